"""The single persistent launch over every batch instance (FF_BATCH_PERSISTENT).

Its kernel (an atomic (op, tile) queue in step-major order with per-op
completion counters) did not complete on MI355X: the first run's dependency
polls invalidated the XCD L2 on every poll; after polls were made relaxed and
every wait bounded in wall time, the run still did not finish in 120 s, so the
cause is not established.  The kernel was removed (FF_BATCH_PERSISTENT is
rejected with FF_ERR_UNSUPPORTED); FF_BATCH_STEPS (tests/test_gpu_batch_steps.py)
runs one launch per pass step with no workgroup waiting on another.  These
tests stay skipped.
"""
import numpy as np
import pytest

import fishfry_amd as ff

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="persistent batch kernel did not complete on MI355X (twice); removed")]


def test_persistent_batch_fp64_matches_oracle():
    import oracle
    import torch

    shapes = [(32, 32, 32), (48, 40, 36), (12, 10, 9), (30, 25, 27)]
    rng = np.random.default_rng(1)
    b = ff.Batch([{"n": n} for n in shapes], ff.FF_BATCH_PERSISTENT)
    d = [torch.zeros(b.bytes(i) // 8, dtype=torch.float64, device="cuda") for i in range(len(shapes))]
    p = [torch.zeros(b.bytes(i) // 8, dtype=torch.float64, device="cuda") for i in range(len(shapes))]
    fields = [rng.standard_normal(n) for n in shapes]
    for k, f in enumerate(fields):
        d[k][: f.size].copy_(torch.from_numpy(f.ravel()))
    b.solve(d, p)
    b.wait()
    for k, n in enumerate(shapes):
        got = p[k][: n[0] * n[1] * n[2]].cpu().numpy().reshape(n)
        want = oracle.solve(fields[k])
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13
