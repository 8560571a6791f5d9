"""Affine bodies on the device: the q-space side of a rigid-body scene
(initializer.BodySetup) over ut_abd_*.

q holds 12 FP64 values per body, (p, a1, a2, a3), with the vertex positions
x_v = p + [a1 a2 a3] xbar_v (fn_apply_q_to_x_template, newton_solver.cpp:1882-1907).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import AbdAssembleArgs, AbdDesc, PcgReport, lib, ptr
from .runtime import Context

FP64, MIXED = 0, 1


def _up(a, device, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype == torch.uint32:
        t = t.view(torch.int32)
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


class AffineBodies:
    """Device copy of the bodies of a BodySetup and the q-space operations."""

    def __init__(self, ctx: Context, st):
        self.ctx, self.st = ctx, st
        dv = ctx.device
        self.nb, self.nv = st.nb, st.nv
        self.d = {
            "vert_off": _up(st.vert_off, dv), "body": _up(st.body, dv), "xbar": _up(st.xbar, dv),
            "mass4": _up(st.mass4.reshape(-1), dv, torch.float64), "volume": _up(st.volume, dv, torch.float64),
            "kappa": _up(st.kappa, dv, torch.float64),
        }
        self.desc = AbdDesc()
        self.desc.nb, self.desc.nv = self.nb, self.nv
        for k, t in self.d.items():
            setattr(self.desc, k, ptr(t))
        self.grad = torch.zeros(12 * self.nb, dtype=torch.float64, device=dv)
        self.blocks = torch.zeros(144 * self.nb, dtype=torch.float64, device=dv)
        self.energy = torch.zeros(2, dtype=torch.float64, device=dv)
        self._pcg = None

    def __del__(self):
        try:
            if self._pcg:
                lib().ut_abd_pcg_destroy(self._pcg)
                self._pcg = None
        except Exception:
            pass

    # -- q -> x
    def q_to_x(self, q, fp64=False):
        q = q.detach().to(torch.float64).reshape(-1).contiguous()
        x = torch.empty(3 * self.nv, dtype=torch.float64 if fp64 else torch.float32, device=self.ctx.device)
        self.ctx.check(lib().ut_abd_q_to_x(self.ctx.handle, C.byref(self.desc), ptr(q),
                                           None if fp64 else ptr(x), ptr(x) if fp64 else None))
        return x.view(-1, 3)

    # -- gradient and blocks
    def assemble(self, q, q_tilde=None, inv_h2=3600.0, contacts=None, count=None, cap=0, quasi_static=False,
                 qs_no_inertia_hessian=False, blocks=True):
        q = q.detach().to(torch.float64).reshape(-1).contiguous()
        qt = None if q_tilde is None else q_tilde.detach().to(torch.float64).reshape(-1).contiguous()
        a = AbdAssembleArgs()
        a.q, a.q_tilde, a.inv_h2 = ptr(q), ptr(qt), float(inv_h2)
        a.quasi_static, a.qs_no_inertia_hessian = int(quasi_static), int(qs_no_inertia_hessian)
        if contacts is not None:
            a.contacts, a.contact_count, a.contact_cap = ptr(contacts), ptr(count), int(cap)
        a.grad, a.blocks, a.energy = ptr(self.grad), ptr(self.blocks) if blocks else None, ptr(self.energy)
        self.ctx.check(lib().ut_abd_assemble(self.ctx.handle, C.byref(self.desc), C.byref(a)))
        return self.grad, self.blocks, self.energy

    def apply(self, p, contacts=None, count=None, cap=0, blocks=None):
        p = p.detach().to(torch.float64).reshape(-1).contiguous()
        y = torch.empty_like(p)
        blk = self.blocks if blocks is None else blocks
        self.ctx.check(lib().ut_abd_apply(self.ctx.handle, C.byref(self.desc), ptr(blk), ptr(contacts), ptr(count),
                                          int(cap) if contacts is not None else 0, ptr(p), ptr(y)))
        return y

    def solve(self, b, contacts=None, count=None, cap=0, precision=FP64, max_iters=1000, rel_tol=1e-6, blocks=None):
        """Block-Jacobi PCG of blocks + contacts against b; returns (x, report dict)."""
        if self._pcg is None:
            h = C.c_void_p()
            self.ctx.check(lib().ut_abd_pcg_create(self.ctx.handle, self.nb, C.byref(h)))
            self._pcg = h
        b = b.detach().to(torch.float64).reshape(-1).contiguous()
        x = torch.zeros_like(b)
        r = PcgReport()
        blk = self.blocks if blocks is None else blocks
        rc = lib().ut_abd_pcg_solve(self._pcg, C.byref(self.desc), ptr(blk), ptr(contacts), ptr(count),
                                    int(cap) if contacts is not None else 0, ptr(b), ptr(x), int(precision),
                                    int(max_iters), float(rel_tol), C.byref(r))
        if rc not in (0, 2):
            self.ctx.check(rc)
        rep = {k: getattr(r, k) for k, _ in PcgReport._fields_}
        rep["converged"] = bool(rep["converged"])
        return x, rep
