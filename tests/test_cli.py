"""The C++ FishFry driver and CLI (fishfry_amd/fishfry, built from
fishfry_amd/host over libfishfry.so): argument handling as main.cpp:12-23, the
per-run line and phase-table format of FishFry.hpp:198-227, and, on the GPU,
the analytic-phi norms at round-off."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fishfry_amd", "fishfry")

RUN_LINE = re.compile(
    r"^(\d+): Time (\S+) Rate (\S+) L_1 (\S+) L_2 (\S+) L_\\inf (\S+)( \(warmup, ignored\))?$")
PHASES = ["Total", "Reorder 3D block into sub-pencils", "Redistribute into Z pencils", "Real-to-complex FFT in Z",
          "Redistribute for Y pencils", "Forward FFT in Y", "Redistribute for X pencils", "X FFTs and filter",
          "Redistribute for Y pencils", "Backward FFT in Y", "Redistribute in Z pencils", "Complex-to-real FFT in Z",
          "Redistribute for 3D blocks", "Rearrange into 3D blocks and normalize"]


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-C", os.path.join(ROOT, "fishfry_amd", "host")], check=True,
                       stdout=subprocess.DEVNULL)
    return EXE


def run(exe, *args, timeout=120):
    return subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=timeout)


def test_usage_without_enough_arguments(exe):
    r = run(exe, 1, 1, 1, 64, 64)
    assert r.returncode == 1 and "usage" in r.stderr


def test_multi_task_grid_is_refused_before_gpu_work(exe):
    r = run(exe, 1, 2, 1, 64, 64, 64)
    assert r.returncode == 2 and "1x1x1" in r.stderr


def test_invalid_grid_is_refused(exe):
    r = run(exe, 1, 1, 1, 0, 64, 64)
    assert r.returncode == 1


def parse(out):
    lines = out.splitlines()
    runs = [RUN_LINE.match(l) for l in lines]
    runs = [m for m in runs if m]
    k = lines.index("Times Averaged across Iterations")
    table = lines[k + 3: k + 3 + len(PHASES)]
    return lines, runs, table


@pytest.mark.gpu
@pytest.mark.parametrize("args", [(64, 64, 64, 2), (48, 40, 36, 1), (40, 32, 50, 1, "--fp32-refine", 2)])
def test_driver_output_and_norms(exe, args):
    print("cli", args, flush=True)
    r = run(exe, 1, 1, 1, *args)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    lines, runs, table = parse(r.stdout)
    n = args[:3]
    assert lines[0] == f"FishFry<ParisPencil>: {n[0]}x{n[1]}x{n[2]} points over 1x1x1 MPI processes"
    iters = args[3]
    assert [int(m.group(1)) for m in runs] == list(range(iters + 1))
    assert runs[0].group(7) and not any(m.group(7) for m in runs[1:])
    for m in runs:
        t, rate, l1, l2, linf = (float(m.group(i)) for i in range(2, 7))
        assert t > 0 and abs(rate * t / (n[0] * n[1] * n[2]) - 1) < 1e-3
        assert max(l1, l2, linf) < 1e-12
    assert [row.split("| ")[-1] for row in table] == PHASES
    assert "(100.0%) | Total" in table[0]
    assert any(l.startswith("Solve latency p50") for l in lines)


@pytest.mark.gpu
@pytest.mark.parametrize("n,iters,precision,steps", [((32, 32, 32), 2, 0, 0), ((40, 32, 50), 1, 1, 2)])
def test_python_driver_matches_cli_format(n, iters, precision, steps):
    """fishfry_amd.FishFry, the Python mirror of FishFry<ParisPencil>, prints
    the same header, per-run lines and phase table as the C++ driver and the
    reference (FishFry.hpp:198-227), with the analytic norms at round-off."""
    import fishfry_amd as ff

    out = []
    drv = ff.FishFry((1, 1, 1), n, precision=precision, refine_steps=steps, printer=out.append)
    res = drv.run(iters)
    drv.close()
    text = "\n".join(out)
    print(text, flush=True)
    lines, runs, table = parse(text)
    assert lines[0] == f"FishFry<ParisPencil>: {n[0]}x{n[1]}x{n[2]} points over 1x1x1 MPI processes"
    assert [int(m.group(1)) for m in runs] == list(range(iters + 1))
    assert runs[0].group(7) and not any(m.group(7) for m in runs[1:])
    assert len(res) == iters + 1 and res[0]["warmup"] and not any(r["warmup"] for r in res[1:])
    for m, r in zip(runs, res):
        assert abs(float(m.group(3)) / r["rate"] - 1) < 1e-3
        assert max(r["l1"], r["l2"], r["linf"]) < 1e-12
    assert [row.split("| ")[-1] for row in table] == PHASES
    with pytest.raises(ff.FishFryError):
        ff.FishFry((1, 2, 1), n)
