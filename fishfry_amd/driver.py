"""Python mirror of the reference's benchmark driver FishFry<P>
(FishFry.hpp:103-246) over the C-ABI: the same header, per-run lines and
phase table as the C++ driver (fishfry_amd/host/FishFry.hpp, binary
``fishfry``), which is the drop-in equivalent of the reference's ``main``.
"""
from __future__ import annotations

import math
from typing import Callable, List, Optional, Sequence

from . import (FF_FP64, Field, FishFryError, ParisPencil, benchmark_grid, field_default, field_diff,
               field_init)


class FishFry:
    """FishFry<ParisPencil>(nTasks, nPoints).run(nIters) at 1x1x1 tasks."""

    def __init__(self, n_tasks: Sequence[int], n_points: Sequence[int], precision: int = FF_FP64,
                 refine_steps: int = 0, field: Optional[Field] = None, printer: Callable[[str], None] = print):
        import torch

        if tuple(n_tasks) != (1, 1, 1):
            raise FishFryError(2, "only 1x1x1 tasks (one GPU) are supported")
        self.n = tuple(int(v) for v in n_points)
        self.total_points = float(self.n[0]) * float(self.n[1]) * float(self.n[2])
        lo, hi, _ = benchmark_grid(self.n)
        self.printer = printer
        self.printer(f"FishFry<ParisPencil>: {self.n[0]}x{self.n[1]}x{self.n[2]} points over 1x1x1 MPI processes")
        self.poisson = ParisPencil(self.n, lo, hi, (1, 1, 1), (0, 0, 0), precision, refine_steps)
        # FishFry.hpp:141-144
        self.bytes = max(self.poisson.bytes(), int(self.total_points) * 8)
        self.f = torch.empty(self.bytes // 8, dtype=torch.float64, device="cuda")
        self.phi = torch.zeros(self.bytes // 8, dtype=torch.float64, device="cuda")
        self.field = field if field is not None else field_default(self.n)
        # FishFry.hpp:146-159: norms of the exact potential
        s = field_diff(self.field, None)
        self.scales = (1.0 / s[0], 1.0 / s[1], 1.0 / s[2])

    def run(self, n_iters: int) -> List[dict]:
        """FishFry.hpp:169-229: n_iters+1 solves, the first is warmup."""
        import torch

        results = []
        total = None
        for i in range(n_iters + 1):
            field_init(self.field, self.f)
            torch.cuda.synchronize()
            stamps: list = []
            self.poisson.solve(self.bytes, self.f, self.phi, stamps)
            t = stamps[-1][1] - stamps[0][1]
            if i == 0:
                total = [(name, 0.0) for name, _ in stamps]
            else:
                total = [(name, acc + s) for (name, acc), (_, s) in zip(total, stamps)]
            d = field_diff(self.field, self.phi)
            l1, l2, linf = d[0] * self.scales[0], math.sqrt(d[1] * self.scales[1]), d[2] * self.scales[2]
            line = "%d: Time %10.4e Rate %10.4e L_1 %10.4e L_2 %10.4e L_\\inf %10.4e" % (
                i, t, self.total_points / t, l1, l2, linf)
            if i == 0:
                line += " (warmup, ignored)"
            self.printer(line)
            results.append({"iter": i, "time": t, "rate": self.total_points / t, "l1": l1, "l2": l2, "linf": linf,
                            "warmup": i == 0})
        self.printer("\nTimes Averaged across Iterations")
        self.printer("Min Rank   | Avg Rank   | Max Rank   (%Total) | Phase")
        self.printer("--------------------------------------------------------")
        per_iter = 1.0 / max(n_iters, 1)
        percent = None
        for k in range(len(total)):
            dt = total[-1][1] - total[0][1] if k == 0 else total[k][1] - total[k - 1][1]
            t = per_iter * dt
            if k == 0:
                percent = 100.0 / t if t else 0.0
            self.printer("%10.4e | %10.4e | %10.4e (%5.1f%%) | %s" % (t, t, t, t * percent, total[k][0]))
        return results

    def close(self) -> None:
        self.poisson.close()
