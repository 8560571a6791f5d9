"""Differentiable incremental potential on the GPU behind torch.autograd.Function.

    E(x) = sum_i m_i / (2 h^2) |x_i - x_tilde_i|^2          inertia (soft_inertia_energy.cpp:25-110)
         + sum_faces  BW98 stretch / shear                  (stretch_energy.h:200-520)
         + sum_hinges dihedral bending                      (bending_energy.h:164-260)
         + sum_pairs  kappa * A * b(d - th, d_hat)          IPC contact barrier (cipc_kernel.hpp:15-36,
                                                            narrow_phase.cpp:1091-1112), or the quadratic form

The reference assembles exactly these terms inside its Newton solver but has no
energy or gradient entry point; this module is the differentiable boundary the
untangling26 Python package gets on top of the same kernels:

* forward: proximity DCD at x (VF + EE pairs within d_hat, FP64 distances),
  then the assembly's inertia / stretch / bending energies plus the contact
  energy, whose closest-point distances are recomputed from x in FP64;
* backward: the hand-written gradient, recomputed (assembly at x: the contact
  gradient is kappa A b'(d) w_k n, exact because the closest point is optimal);
* hvp: the product with the Newton matrix the solver uses (assembled 3x3 blocks
  plus the matrix-free rank-1 contact blocks k2 (W n)(W n)^T, i.e. the contact
  Hessian without the curvature term k1 d''(x), as the reference's solver).

Positions enter the kernels as FP32 (the solver's storage); energies, gradients
and products are FP64.  A batch x[B, V, 3] is evaluated state by state.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import AssembleArgs, lib, ptr
from .lbvh import LBVH, TREE_EDGE, TREE_FACE
from .narrow_phase import BARRIER, QUADRATIC, NarrowPhasesDetector
from .pcg import ConjugateGradientSolver
from .runtime import Context

__all__ = ["IncrementalPotential", "IPFunction", "BARRIER", "QUADRATIC"]


def _up(a, device):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.view(torch.int32) if t.dtype == torch.uint32 else t).to(device)


class IncrementalPotential:
    """The potential of one cloth setup (initializer.ClothSetup).  x_tilde is the
    inertia target (None: no inertia term); quasi_static keeps only its Hessian."""

    def __init__(self, ctx: Context, setup, x_tilde=None, inv_h2=3600.0, kappa=1e8, energy=BARRIER,
                 contact_cap=1 << 16):
        self.ctx, self.st = ctx, setup
        dv = ctx.device
        st = setup
        self.nv, self.nf, self.ne, self.nh = st.nv, len(st.faces), len(st.edges), len(st.hinges)
        names = dict(edges=st.edges, faces=st.faces, mass=st.mass, d_hat=st.d_hat, offset=st.offset,
                     va=st.rest_vert_area, ea=st.rest_edge_area, fa=st.rest_face_area, dm=st.dm,
                     mu_lambda=st.mu_lambda, face_slots=st.face_slots, hinges=st.hinges, rest_angle=st.rest_angle,
                     bend=st.bend_stiff_area, hinge_slots=st.hinge_slots, diag_slot=st.pattern.diag_slot)
        self.d = {k: _up(v, dv) for k, v in names.items()}
        self.x_tilde = None if x_tilde is None else torch.as_tensor(
            np.asarray(x_tilde, np.float64).reshape(-1), device=dv).contiguous()
        self.inv_h2, self.kappa, self.energy_type = float(inv_h2), float(kappa), int(energy)
        self.narrow = NarrowPhasesDetector(ctx, self.nv, self.d["edges"], self.d["faces"], self.d["d_hat"],
                                           self.d["offset"], self.d["va"], self.d["ea"], self.d["fa"], kappa=kappa,
                                           energy=energy)
        self.face_tree = LBVH(ctx, TREE_FACE, self.nf)
        self.edge_tree = LBVH(ctx, TREE_EDGE, self.ne)
        self.pcg = ConjugateGradientSolver(ctx, st.pattern)
        self.nslots = st.pattern.nslots
        self.count = torch.zeros(1, dtype=torch.int32, device=dv)
        self._reserve(contact_cap)
        self.grad = torch.zeros(3 * self.nv, dtype=torch.float64, device=dv)
        self.vals = torch.zeros(9 * self.nslots, dtype=torch.float64, device=dv)
        self.parts = torch.zeros(4, dtype=torch.float64, device=dv)
        self.ce = torch.zeros(257, dtype=torch.float64, device=dv)
        self._x = None  # positions the contact list and trees were built for

    def _reserve(self, cap):
        self.cap = int(cap)
        self.pairs = torch.zeros(16 * self.cap, dtype=torch.int32, device=self.ctx.device)
        self.pair_ids = torch.zeros(2 * self.cap, dtype=torch.int32, device=self.ctx.device)

    @staticmethod
    def _f32(x):
        return x.detach().to(torch.float32).reshape(-1).contiguous()

    # -- proximity pairs at x
    def contacts(self, x):
        """Detects the VF and EE pairs within d_hat at x (FP32 [V, 3] on the device);
        returns the number of records."""
        x32 = self._f32(x)
        d = self.d
        for tree, prims, red, upd in ((self.face_tree, d["faces"], "reduce_face_tree_aabb", "update_face_tree_leave_aabb"),
                                      (self.edge_tree, d["edges"], "reduce_edge_tree_aabb", "update_edge_tree_leave_aabb")):
            getattr(tree, red)(x32, prims)
            tree.construct_tree()
            getattr(tree, upd)(d["offset"], x32, x32, prims)
            tree.refit()
        while True:
            self.count.zero_()
            self.narrow.vf_dcd_query_repulsion(x32, self.count, self.pairs, self.pair_ids, self.face_tree)
            self.narrow.ee_dcd_query_repulsion(x32, self.count, self.pairs, self.pair_ids, self.edge_tree)
            n = int(self.count.item())
            if n <= self.cap:
                break
            self._reserve(2 * n)
        # a copy: x32 may share storage with the caller's tensor, which an in-place
        # update (an optimiser step) would change under the cache
        self._x = x32.clone()
        return n

    def _assemble(self, x32, want_vals):
        d = self.d
        self.grad.zero_()
        self.vals.zero_()
        a = AssembleArgs()
        a.nv, a.nf, a.nh, a.nslots = self.nv, self.nf, self.nh, self.nslots
        a.x = ptr(x32)
        a.x_tilde, a.mass, a.inv_h2 = ptr(self.x_tilde), ptr(d["mass"]), self.inv_h2
        a.faces, a.dm, a.face_area, a.mu_lambda, a.face_slots = (ptr(d[k]) for k in
                                                                 ("faces", "dm", "fa", "mu_lambda", "face_slots"))
        a.hinges, a.rest_angle, a.bend_stiff_area, a.hinge_slots = (ptr(d[k]) for k in
                                                                    ("hinges", "rest_angle", "bend", "hinge_slots"))
        a.contacts, a.contact_count, a.contact_cap = ptr(self.pairs), ptr(self.count), self.cap
        a.diag_slot, a.grad, a.vals, a.vals32 = ptr(d["diag_slot"]), ptr(self.grad), ptr(self.vals), None
        a.energy = ptr(self.parts)
        self.ctx.check(lib().ut_assemble(self.ctx.handle, C.byref(a)))

    def _prepare(self, x):
        x32 = self._f32(x)
        if self._x is None or not torch.equal(self._x, x32):
            self.contacts(x32)
        return x32

    # -- the potential
    def energies(self, x):
        """(inertia, stretch, bending, contact) as FP64 tensors."""
        x32 = self._prepare(x)
        self._assemble(x32, False)
        self.ctx.check(lib().ut_contact_energy(self.ctx.handle, ptr(self.pairs), ptr(self.count), self.cap, ptr(x32),
                                               ptr(self.d["d_hat"]), ptr(self.d["offset"]), self.kappa,
                                               self.energy_type, ptr(self.ce)))
        return torch.stack([self.parts[0], self.parts[1], self.parts[2], self.ce[0]]).clone()

    def energy(self, x):
        return self.energies(x).sum()

    def gradient(self, x):
        """dE/dx, FP64 [V, 3] (hand-written; recomputed at x)."""
        x32 = self._prepare(x)
        self._assemble(x32, False)
        return self.grad.view(-1, 3).clone()

    def hvp(self, x, v):
        """H(x) v with the Newton matrix (assembled blocks + matrix-free contact blocks), FP64 [V, 3]."""
        x32 = self._prepare(x)
        self._assemble(x32, True)
        vin = v.detach().to(torch.float64).reshape(-1).contiguous()
        out = torch.zeros_like(vin)
        self.ctx.check(lib().ut_pcg_set_contacts(self.pcg.handle, ptr(self.pairs), ptr(self.count), self.cap))
        try:
            self.pcg.spmv(self.vals, vin, out)
        finally:
            self.ctx.check(lib().ut_pcg_set_contacts(self.pcg.handle, None, None, 0))
        return out.view(-1, 3)


class IPFunction(torch.autograd.Function):
    """E = IPFunction.apply(x, potential): x [V, 3] or [B, V, 3] on the device;
    E is FP64 (scalar or [B]).  backward multiplies the hand-written gradient."""

    @staticmethod
    def forward(ctx, x, potential: IncrementalPotential):
        ctx.potential = potential
        ctx.save_for_backward(x)
        if x.dim() == 2:
            return potential.energy(x)
        return torch.stack([potential.energy(xb) for xb in x])

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        p = ctx.potential
        if x.dim() == 2:
            gx = p.gradient(x) * g
        else:
            gx = torch.stack([p.gradient(xb) * g[b] for b, xb in enumerate(x)])
        return gx.to(x.dtype), None
