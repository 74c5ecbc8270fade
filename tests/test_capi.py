"""CPU checks of the C-ABI library: it loads, exports every symbol the headers
declare, and its host-only paths (byte formula, argument validation, field
defaults) behave like the reference without touching a GPU."""
import ctypes
import math
import os
import re

import pytest

import fishfry_amd as ff
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ff_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = ff.lib()
    names = declared_symbols()
    assert len(names) >= 12
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_constants_match_the_header():
    """Every FF_* integer constant the Python package defines has the value
    the C header gives it (flags, precisions, status codes)."""
    text = open(os.path.join(ROOT, "include", "fishfry.h")).read()
    defines = {m.group(1): int(m.group(2), 0)
               for m in re.finditer(r"#define\s+(FF_[A-Z0-9_]+)\s+\(?(-?(?:0x[0-9a-fA-F]+|\d+))\)?", text)}
    enums = {m.group(1): int(m.group(2)) for m in re.finditer(r"\b(FF_[A-Z0-9_]+)\s*=\s*(-?\d+)", text)}
    header = {**enums, **defines}
    checked = 0
    for name in dir(ff):
        if name.startswith("FF_") and isinstance(getattr(ff, name), int) and name in header:
            assert getattr(ff, name) == header[name], name
            checked += 1
    assert checked >= 15 and "FF_SERVE_LARGE_ALONE" in header


@pytest.mark.parametrize("n,m", [
    ((64, 64, 64), (1, 1, 1)), ((1024, 1024, 1024), (1, 1, 1)), ((12, 10, 9), (1, 1, 1)),
    ((128, 256, 384), (1, 2, 3)), ((100, 90, 80), (2, 2, 2)), ((33, 17, 65), (3, 1, 4)),
    ((64, 64, 64), (1, 1, 8)), ((7, 5, 11), (1, 1, 1)),
])
def test_reference_bytes_matches_oracle(n, m):
    assert ff.reference_bytes(n, m) == oracle.ref_bytes(n, m)


def test_reference_bytes_known_values():
    assert ff.reference_bytes((64, 64, 64)) == 2162688
    assert ff.reference_bytes((1024, 1024, 1024)) == 8606711808
    assert ff.reference_bytes((0, 4, 4)) == 0


def test_plan_create_rejects_without_gpu_work():
    lo, hi, _ = ff.benchmark_grid((16, 16, 16))
    # more than one task is the reference's multi-rank case: unsupported on one GPU
    with pytest.raises(ff.FishFryError) as e:
        ff.ParisPencil((16, 16, 16), lo, hi, (1, 1, 2), (0, 0, 0))
    assert e.value.status == ff.FF_ERR_UNSUPPORTED
    with pytest.raises(ff.FishFryError) as e:
        ff.ParisPencil((16, 16, 16), lo, hi, (1, 1, 1), (0, 1, 0))
    assert e.value.status == ff.FF_ERR_INVALID
    # a single cell on an axis makes hi == lo: the reference's dd is 0/0 there
    lo1, hi1, _ = ff.benchmark_grid((1, 8, 8))
    with pytest.raises(ff.FishFryError) as e:
        ff.ParisPencil((1, 8, 8), lo1, hi1)
    assert e.value.status == ff.FF_ERR_INVALID
    # option flags: unknown bits, and layouts that exclude each other
    for flags in (128, ff.FF_PENCIL_LAYOUT | ff.FF_PLAIN_LAYOUT, ff.FF_PENCIL_LAYOUT | ff.FF_BLOCKED_LAYOUT,
                  ff.FF_PENCIL_LAYOUT | ff.FF_SLAB_PIPELINE, ff.FF_SYMBOL_3PT | ff.FF_SYMBOL_5PT):
        with pytest.raises(ff.FishFryError) as e:
            ff.ParisPencil((16, 16, 16), lo, hi, flags=flags)
        assert e.value.status == ff.FF_ERR_INVALID, flags
    L = ff.lib()
    h = ctypes.c_void_p()
    rc = L.ff_plan_create(None, None, None, None, None, None, ctypes.byref(h))
    assert rc == ff.FF_ERR_INVALID and "null" in ff.last_error()
    assert L.ff_plan_bytes(None) == 0
    assert L.ff_plan_destroy(None) == ff.FF_OK


def test_solve_rejects_bad_arguments_without_gpu_work():
    L = ff.lib()
    assert L.ff_solve(None, 0, None, None, None, None, None) == ff.FF_ERR_INVALID
    assert L.ff_solve_async(None, 0, None, None, None) == ff.FF_ERR_INVALID
    assert L.ff_solve_pass(None, 0, None, None, None) == ff.FF_ERR_INVALID


def test_field_default_matches_reference_grid():
    n = (64, 48, 1024)
    f = ff.field_default(n)
    for a in range(3):
        assert f.n[a] == n[a]
        assert f.d[a] == 2.0 * math.pi / n[a]
        assert f.lo[a] == -math.pi
        lo, hi, _ = oracle.grid(n)
        assert f.dd[a] == oracle.dd(n[a], lo[a], hi[a])
    assert f.scale == 1.0 and f.nmodes == 0
    assert tuple(f.shift) == (0.0, 0.0, 0.0)


def test_field_validation_without_gpu_work():
    f = ff.field_default((8, 8, 8))
    f.nmodes = 99
    with pytest.raises(ff.FishFryError) as e:
        ff.field_init(f, 0)
    assert e.value.status == ff.FF_ERR_INVALID


def test_persistent_batch_is_rejected_without_gpu_work():
    with pytest.raises(ff.FishFryError) as e:
        ff.Batch([{"n": (8, 8, 8)}], ff.FF_BATCH_PERSISTENT)
    assert e.value.status == ff.FF_ERR_UNSUPPORTED
    assert "persistent" in str(e.value)


def test_batch_flags_are_validated_without_gpu_work():
    with pytest.raises(ff.FishFryError) as e:
        ff.Batch([{"n": (8, 8, 8)}], 32)  # no such flag
    assert e.value.status == ff.FF_ERR_INVALID
    with pytest.raises(ff.FishFryError) as e:
        ff.Batch([{"n": (8, 8, 8)}], ff.FF_BATCH_GROUPED | ff.FF_BATCH_STEPS)  # exclusive modes
    assert e.value.status == ff.FF_ERR_INVALID
    with pytest.raises(ff.FishFryError) as e:
        ff.Batch([{"n": (8, 8, 1)}], ff.FF_BATCH_GROUPED)  # extents below 2 are refused first
    assert e.value.status == ff.FF_ERR_INVALID


def test_tuning_host_logic_without_gpu_work():
    """ff_tune_query / ff_tune_set are host logic: the default configurations
    of the column passes of 1024, 256 and 32, 48, 64, 96, 128, 160, and of the x passes
    of 512, 768 and 1000, run the kernels with the line plan compiled in
    (flags without FF_TUNE_GENERIC -- which also needs the compile-time plan
    to equal the planner's); the y passes of 512, 768 and 1000 and other
    lengths run the generic kernel, and the overrides round-trip.  At 1024 the
    column passes take their twiddles from LDS by default."""
    TWL = ff.FF_TUNE_LDS_TWIDDLES
    ff.tune_reset()
    try:
        for kind in (1, 2):
            # the FP64 x pass at 1024 runs the split exchange by default (three workgroups per CU)
            t = ff.tune_query(kind, 1024, 0, 513)
            assert (t.E, t.T, t.flags) == (16, 4, (ff.FF_TUNE_SPLIT_LDS if kind == 2 else 0) | TWL)
            t = ff.tune_query(kind, 1024, 1, 513)
            assert (t.E, t.T, t.flags) == (16, 8, TWL)
        assert ff.tune_query(1, 512, 0, 257).flags & ff.FF_TUNE_GENERIC
        assert ff.tune_query(2, 600, 0, 301).flags & ff.FF_TUNE_GENERIC
        for kind in (1, 2):
            for n, E, T in ((128, 16, 32), (256, 16, 16), (48, 12, 32)):
                for prec in (0, 1) if n != 256 else (0,):
                    t = ff.tune_query(kind, n, prec, n // 2 + 1)
                    assert (t.E, t.T, t.flags) == (E, T, 0), (kind, n, prec)
        t = ff.tune_query(2, 1000, 0, 501)
        assert (t.E, t.T, t.flags) == (20, 4, 0)
        t = ff.tune_query(2, 768, 0, 385)
        assert (t.E, t.T, t.flags) == (12, 4, 0)
        for n, E in ((1000, 20), (768, 24)):
            t = ff.tune_query(1, n, 0, n // 2 + 1)
            assert (t.E, t.T, t.flags) == (E, 4, ff.FF_TUNE_GENERIC)
        assert ff.tune_query(2, 1000, 1, 501).flags & ff.FF_TUNE_GENERIC  # FP32: generic
        ff.tune_set(1, 1000, 0, 20, 4, 0)  # the x-only kernel is not used for the y pass
        assert ff.tune_query(1, 1000, 0, 501).flags == ff.FF_TUNE_GENERIC
        ff.tune_reset()
        ff.tune_set(2, 1024, 0, 16, 4, ff.FF_TUNE_GENERIC)
        assert ff.tune_query(2, 1024, 0, 513).flags == ff.FF_TUNE_GENERIC
        assert ff.tune_query(1, 1024, 0, 513).flags == TWL
        ff.tune_set(1, 1024, 0, 16, 4, ff.FF_TUNE_SPLIT_LDS)  # the compiled-in split kernel (pencil default)
        assert ff.tune_query(1, 1024, 0, 513).flags == ff.FF_TUNE_SPLIT_LDS
        ff.tune_set(1, 1024, 1, 16, 8, ff.FF_TUNE_SPLIT_LDS)  # no compiled-in FP32 split kernel
        assert ff.tune_query(1, 1024, 1, 513).flags == ff.FF_TUNE_SPLIT_LDS | ff.FF_TUNE_GENERIC
        ff.tune_set(2, 1024, 0, 16, 4, 0)  # the non-split x kernel with twiddles from memory, on request
        assert ff.tune_query(2, 1024, 0, 513).flags == 0
        ff.tune_set(2, 1024, 0, 16, 4, TWL)
        assert ff.tune_query(2, 1024, 0, 513).flags == TWL
        with pytest.raises(ff.FishFryError):
            ff.tune_set(2, 1024, 0, 16, 4, 16)  # unknown flag
        with pytest.raises(ff.FishFryError):
            ff.tune_set(1, 512, 0, 8, 4, TWL)  # LDS twiddles need a kernel with the plan compiled in
        with pytest.raises(ff.FishFryError):
            ff.tune_set(1, 1024, 0, 16, 4, TWL | ff.FF_TUNE_GENERIC)
        with pytest.raises(ff.FishFryError):
            ff.tune_set(0, 512, 0, 8, 2, ff.FF_TUNE_SPLIT_LDS)  # row passes: FF_TUNE_GENERIC or LDS twiddles
        with pytest.raises(ff.FishFryError):
            ff.tune_set(0, 300, 0, 10, 2, TWL)  # LDS twiddles need a row kernel with the plan compiled in
        with pytest.raises(ff.FishFryError):
            ff.tune_set(3, 512, 0, 8, 2, TWL | ff.FF_TUNE_GENERIC)
        ff.tune_set(3, 256, 1, 16, 4, TWL)  # on request at other compiled-in row lengths
        t = ff.tune_query(3, 256, 1, 1)
        assert (t.E, t.T, t.flags) == (16, 4, TWL)
        ff.tune_set(0, 512, 0, 8, 2, 0)  # the row kernel with twiddles from memory, on request
        assert ff.tune_query(0, 512, 0, 1).flags == 0
        ff.tune_reset()
        assert ff.tune_query(2, 1024, 0, 513).flags == ff.FF_TUNE_SPLIT_LDS | TWL
        # row passes: compiled-in plans at the default configurations of even
        # nz = 32 ... 1024 (length nz/2), generic otherwise or on request; at
        # nz = 1024 the twiddles come from LDS (FP64 two rows per workgroup,
        # FP32 four)
        for kind in (0, 3):
            for M, E in ((512, 8), (256, 16), (80, 20), (16, 16)):
                for prec in (0, 1):
                    t = ff.tune_query(kind, M, prec, 1)
                    assert (t.E, t.flags) == (E, TWL if M == 512 else 0), (kind, M, prec)
                    if M == 512:
                        assert t.T == (4 if prec else 2)
            assert ff.tune_query(kind, 450, 0, 1).flags == ff.FF_TUNE_GENERIC
            ff.tune_set(kind, 512, 0, 8, 2, ff.FF_TUNE_GENERIC)
            assert ff.tune_query(kind, 512, 0, 1).flags == ff.FF_TUNE_GENERIC
        ff.tune_reset()
    finally:
        ff.tune_reset()
