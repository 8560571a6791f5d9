"""Python mirror of lcs::ConjugateGradientSolver (Solver/LinearSolver/precond_cg.h:471-543)."""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import PcgParams, PcgReport, lib, ptr
from .runtime import Context

FP64, MIXED, FP32 = 0, 1, 2
BLOCK_JACOBI, POINT_JACOBI = 0, 1
# Krylov recurrence: the reference's loop (precond_cg.cpp:706-914), or its
# single-reduction (Chronopoulos-Gear) form with two launches per iteration
CLASSIC, SINGLE_REDUCTION = 0, 1


class ConjugateGradientSolver:
    def __init__(self, ctx: Context, pattern, device_pattern=None):
        """pattern: bsr.SellPattern; its arrays are uploaded unless device_pattern
        (slice_off, col, diag_slot) tensors are given."""
        self.ctx = ctx
        self.n = pattern.n
        self.nslots = pattern.nslots
        if device_pattern is None:
            device_pattern = tuple(torch.from_numpy(a.view("int32")).to(ctx.device)
                                   for a in (pattern.slice_off, pattern.col, pattern.diag_slot))
        self.dev_pattern = device_pattern
        so, col, ds = device_pattern
        h = C.c_void_p()
        ctx.check(lib().ut_pcg_create(ctx.handle, self.n, pattern.nslices, self.nslots, ptr(so), ptr(col), ptr(ds),
                                      C.byref(h)))
        self.handle = h
        self.last_solve_report = None

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().ut_pcg_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def device_solve(self, vals, b, x, precision=MIXED, preconditioner=BLOCK_JACOBI, max_iters=100, rel_tol=1e-6,
                     inner_tol=1e-2, max_refine=20, regularization=0.0, check_every=8, vals64=None,
                     recurrence=CLASSIC):
        """Solves A x = b; returns the PcgSolveReport as a dict (also kept in last_solve_report).
        In MIXED mode `vals64` (the FP64 matrix) makes the refinement residual FP64, so the
        result solves the FP64 system; without it the FP32 matrix is refined against."""
        want = torch.float64 if precision == FP64 else torch.float32
        assert vals.dtype == want and vals.is_contiguous(), f"values must be {want}"
        assert b.dtype == torch.float64 and x.dtype == torch.float64
        p = PcgParams(precision, preconditioner, max_iters, rel_tol, inner_tol, max_refine, regularization,
                      check_every, recurrence)
        r = PcgReport()
        if vals64 is not None:
            assert precision == MIXED and vals64.dtype == torch.float64 and vals64.is_contiguous()
            rc = lib().ut_pcg_solve_refined(self.handle, C.byref(p), ptr(vals), ptr(vals64), ptr(b), ptr(x),
                                            C.byref(r))
        else:
            rc = lib().ut_pcg_solve(self.handle, C.byref(p), ptr(vals), ptr(b), ptr(x), C.byref(r))
        self.ctx.check(rc)
        rep = {k: getattr(r, k) for k, _ in PcgReport._fields_}
        rep["converged"] = bool(rep["converged"])
        self.last_solve_report = rep
        return rep

    def set_contacts(self, pairs=None, count=None, cap: int = 0):
        """Contact records (int32 words, 16 per record) whose off-diagonal blocks
        w_a w_b k2 n n^T every product with A applies; None clears them."""
        self._contacts = (pairs, count)  # the library keeps the pointers
        self.ctx.check(lib().ut_pcg_set_contacts(self.handle, ptr(pairs), ptr(count), int(cap) if pairs is not None
                                                 else 0))

    def spmv(self, vals, x, y):
        self.ctx.check(lib().ut_bsr_spmv(self.handle, ptr(vals), int(vals.dtype == torch.float32), ptr(x), ptr(y)))
