"""The CPU baseline leg of bench.py (oracle/cpu_step.py) on a small fold, on the
CPU: one whole step with nothing scaled, on 1 and on 4 threads; the threaded PCG
(parallel products, reduction-order dots) solves the same system as the serial
one."""
import types

import numpy as np
import torch

import oracle
from cpu_step import cpu_step_baseline
from untangling26_amd import initializer, scenes


class _FakeSolver:
    def __init__(self, mesh, rest):
        self._setup = initializer.setup_cloth(mesh, x_rest=rest)
        self._x = torch.from_numpy(mesh.x.copy())

    def positions(self):
        return self._x


def test_whole_step_serial_and_threaded():
    m = scenes.folded_cloth(0, n=32)
    rest = scenes.folded_cloth_rest(n=32)
    s = _FakeSolver(m, rest)
    args = types.SimpleNamespace(dirs=10, pcg_iters=400)
    a = cpu_step_baseline(m, rest, s, args, threads=1)
    b = cpu_step_baseline(m, rest, s, args, threads=4)
    assert a["cores"] == 1 and b["cores"] == 4 and a["kind"] == "port"
    assert a["value"] > 0 and b["value"] > 0
    assert set(a["seconds"]) == {"detect", "contours", "prp", "assemble", "pcg", "ccd"}
    assert "nothing sampled or scaled" in b["sample"]


def test_threaded_pcg_matches_serial():
    from autograd_cases import ip_case

    m, st, x_tilde, recs = ip_case(energy=1)
    fr = oracle.friction_init(recs, np.full(m.nv, 0.5, np.float32))
    g, vals, _ = oracle.assemble(st, m.x, x_tilde, 3600.0, fr)
    # the friction blocks make the system stiff: solve far below 1e-6 so that the two
    # summation orders must meet (at 1e-6 the solutions may still differ by 0.5%)
    x1 = oracle.pcg_sell(st.pattern, vals, -g, fr, max_iters=20000, rel_tol=1e-11)
    oracle.set_threads(4)
    try:
        x4 = oracle.pcg_sell(st.pattern, vals, -g, fr, max_iters=20000, rel_tol=1e-11)
    finally:
        oracle.set_threads(1)
    assert x1[3] and x4[3]
    assert np.abs(x1[0] - x4[0]).max() <= 1e-4 * np.abs(x1[0]).max()
