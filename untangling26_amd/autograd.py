"""Differentiable incremental potential on the GPU behind torch.autograd.Function.

    E(x) = sum_i m_i / (2 h^2) |x_i - x_tilde_i|^2          inertia (soft_inertia_energy.cpp:25-110)
         + sum_faces  BW98 stretch / shear                  (stretch_energy.h:200-520)
         + sum_hinges dihedral bending                      (bending_energy.h:164-260)
         + sum_pairs  kappa * A * b(d - th, d_hat)          IPC contact barrier (cipc_kernel.hpp:15-36,
                                                            narrow_phase.cpp:1091-1112), or the quadratic form
         + sum_lagged mu_lambda f0(|P (sum_k w_k (x_k - x_lag_k))|)
                                                            lagged IPC friction (friction_kernel.hpp:137-301,
                                                            narrow_phase.cpp:1455-1521, 3485-3512)

The reference assembles exactly these terms inside its Newton solver but has no
energy or gradient entry point; this module is the differentiable boundary the
untangling26 Python package gets on top of the same kernels:

* forward: proximity DCD at x (VF + EE pairs within d_hat), then the assembly's
  inertia / stretch / bending energies, the contact energy with the closest
  points re-solved at x in FP64, and the friction energy of the lagged pairs;
* backward: the hand-written gradient, recomputed at x (material terms from the
  assembly kernels, the contact term kappa A b'(d) w_k n from distances re-solved
  at x, ut_contact_potential; the friction term from the lagged pairs);
* hvp: the product with the Newton matrix the solver uses (assembled 3x3 blocks
  plus the matrix-free contact blocks w_a w_b K, K = k2 n n^T + the friction
  Hessian, i.e. the contact Hessian without the curvature term, as the
  reference's solver).

Friction is lagged as in the reference's step: set_lag(x_lag) detects the
proximity pairs at x_lag and fixes their normals, weights and multipliers
mu_lambda = -mu k1 (mu per vertex); the friction term then depends on x only
through the tangential motion sum_k w_k (x_k - x_lag_k).

Positions may be FP32 (the solver's storage) or FP64: with FP64 positions every
term is evaluated from the FP64 coordinates (the contact set is detected at the
positions rounded to FP32), so FP64 central differences of the energy check the
gradient.  Energies, gradients and products are FP64.  A batch x[B, V, 3] is
evaluated state by state.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import TERM_BARRIER, TERM_FRICTION, AssembleArgs, PotentialArgs, lib, ptr
from .lbvh import LBVH, TREE_EDGE, TREE_FACE
from .narrow_phase import BARRIER, QUADRATIC, NarrowPhasesDetector
from .pcg import ConjugateGradientSolver
from .runtime import Context

__all__ = ["IncrementalPotential", "IPFunction", "AbdPotential", "AbdFunction", "BARRIER", "QUADRATIC"]


def _up(a, device):
    t = torch.from_numpy(np.ascontiguousarray(a))
    return (t.view(torch.int32) if t.dtype == torch.uint32 else t).to(device)


class _ContactList:
    """A device list of contact records with its count."""

    def __init__(self, device, cap):
        self.device = device
        self.count = torch.zeros(1, dtype=torch.int32, device=device)
        self.reserve(cap)

    def reserve(self, cap):
        self.cap = int(cap)
        self.pairs = torch.zeros(16 * self.cap, dtype=torch.int32, device=self.device)
        self.ids = torch.zeros(2 * self.cap, dtype=torch.int32, device=self.device)


class IncrementalPotential:
    """The potential of one cloth setup (initializer.ClothSetup).  x_tilde is the
    inertia target (None: no inertia term).  friction_mu (scalar or per vertex) with
    x_lag adds the lagged friction of the proximity pairs at x_lag."""

    def __init__(self, ctx: Context, setup, x_tilde=None, inv_h2=3600.0, kappa=1e8, energy=BARRIER,
                 contact_cap=1 << 16, friction_mu=None, x_lag=None):
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
        self.cl = _ContactList(dv, contact_cap)  # proximity pairs at the current x
        self.fl = None  # lagged friction pairs
        self.comb = _ContactList(dv, 2 * contact_cap)  # both, for the Newton matrix
        self.grad = torch.zeros(3 * self.nv, dtype=torch.float64, device=dv)
        self.vals = torch.zeros(9 * self.nslots, dtype=torch.float64, device=dv)
        self.parts = torch.zeros(4, dtype=torch.float64, device=dv)
        self.pot = torch.zeros(2 + 512, dtype=torch.float64, device=dv)
        self._x = None  # FP32 positions the contact list and trees were built for
        self.mu = None
        self.x_lag32 = self.x_lag64 = None
        if friction_mu is not None:
            mu = np.broadcast_to(np.asarray(friction_mu, np.float32), (self.nv,))
            self.mu = torch.as_tensor(np.ascontiguousarray(mu), device=dv)
            self.set_lag(x_lag)

    @staticmethod
    def _f32(x):
        return x.detach().to(torch.float32).reshape(-1).contiguous()

    @staticmethod
    def _f64(x):
        return x.detach().to(torch.float64).reshape(-1).contiguous()

    # -- proximity pairs
    def _detect(self, x32, lst):
        d = self.d
        for tree, prims, red, upd in ((self.face_tree, d["faces"], "reduce_face_tree_aabb", "update_face_tree_leave_aabb"),
                                      (self.edge_tree, d["edges"], "reduce_edge_tree_aabb", "update_edge_tree_leave_aabb")):
            getattr(tree, red)(x32, prims)
            tree.construct_tree()
            getattr(tree, upd)(d["offset"], x32, x32, prims)
            tree.refit()
        while True:
            lst.count.zero_()
            self.narrow.vf_dcd_query_repulsion(x32, lst.count, lst.pairs, lst.ids, self.face_tree)
            self.narrow.ee_dcd_query_repulsion(x32, lst.count, lst.pairs, lst.ids, self.edge_tree)
            n = int(lst.count.item())
            if n <= lst.cap:
                return n
            lst.reserve(2 * n)

    def contacts(self, x):
        """Detects the VF and EE pairs within d_hat at x rounded to FP32; returns their number."""
        x32 = self._f32(x)
        n = self._detect(x32, self.cl)
        # a copy: x32 may share storage with the caller's tensor, which an in-place
        # update (an optimiser step) would change under the cache
        self._x = x32.clone()
        return n

    def set_lag(self, x_lag):
        """Lagged friction about x_lag: the proximity pairs at x_lag with mu_lambda = -mu k1."""
        if self.mu is None:
            raise RuntimeError("set_lag needs friction_mu")
        if x_lag is None:
            raise ValueError("friction needs the lagged positions x_lag")
        xl = torch.as_tensor(x_lag, device=self.ctx.device)
        self.x_lag64 = self._f64(xl)
        self.x_lag32 = self._f32(xl)
        if self.fl is None:
            self.fl = _ContactList(self.ctx.device, self.cl.cap)
        self.nfric = self._detect(self.x_lag32, self.fl)
        self.ctx.check(lib().ut_friction_init(self.ctx.handle, ptr(self.fl.pairs), ptr(self.fl.count), self.fl.cap,
                                              ptr(self.mu)))
        self._x = None  # the trees now hold x_lag

    def _prepare(self, x):
        x32 = self._f32(x)
        if self._x is None or not torch.equal(self._x, x32):
            self.contacts(x32)
        return x32, (self._f64(x) if x.dtype == torch.float64 else None)

    def _assemble(self, x32, x64, contacts=None):
        d = self.d
        self.grad.zero_()
        self.vals.zero_()
        a = AssembleArgs()
        a.nv, a.nf, a.nh, a.nslots = self.nv, self.nf, self.nh, self.nslots
        a.x, a.x64 = ptr(x32), ptr(x64)
        a.x_tilde, a.mass, a.inv_h2 = ptr(self.x_tilde), ptr(d["mass"]), self.inv_h2
        a.faces, a.dm, a.face_area, a.mu_lambda, a.face_slots = (ptr(d[k]) for k in
                                                                 ("faces", "dm", "fa", "mu_lambda", "face_slots"))
        a.hinges, a.rest_angle, a.bend_stiff_area, a.hinge_slots = (ptr(d[k]) for k in
                                                                    ("hinges", "rest_angle", "bend", "hinge_slots"))
        if contacts is not None:
            a.contacts, a.contact_count, a.contact_cap = ptr(contacts.pairs), ptr(contacts.count), contacts.cap
        a.diag_slot, a.grad, a.vals, a.vals32 = ptr(d["diag_slot"]), ptr(self.grad), ptr(self.vals), None
        a.energy = ptr(self.parts)
        self.ctx.check(lib().ut_assemble(self.ctx.handle, C.byref(a)))

    def _potential(self, lst, x32, x64, terms, grad):
        p = PotentialArgs()
        p.pairs, p.count, p.cap = ptr(lst.pairs), ptr(lst.count), lst.cap
        if x64 is not None:
            p.x64, p.x_start64 = ptr(x64), ptr(self.x_lag64)
        else:
            p.x, p.x_start = ptr(x32), ptr(self.x_lag32)
        p.d_hat, p.offset = ptr(self.d["d_hat"]), ptr(self.d["offset"])
        p.kappa, p.energy, p.terms = self.kappa, self.energy_type, terms
        p.grad, p.out = ptr(grad), ptr(self.pot)
        self.ctx.check(lib().ut_contact_potential(self.ctx.handle, C.byref(p)))
        return self.pot[:2].clone()

    # -- the potential
    def energies(self, x):
        """(inertia, stretch, bending, contact, friction) as FP64 tensors."""
        x32, x64 = self._prepare(x)
        self._assemble(x32, x64)
        parts = self.parts[:3].clone()
        contact = self._potential(self.cl, x32, x64, TERM_BARRIER, None)[0]
        friction = (self._potential(self.fl, x32, x64, TERM_FRICTION, None)[1] if self.fl is not None
                    else torch.zeros((), dtype=torch.float64, device=self.ctx.device))
        return torch.stack([parts[0], parts[1], parts[2], contact, friction])

    def energy(self, x):
        return self.energies(x).sum()

    def gradient(self, x):
        """dE/dx, FP64 [V, 3] (hand-written; recomputed at x)."""
        x32, x64 = self._prepare(x)
        self._assemble(x32, x64)
        self._potential(self.cl, x32, x64, TERM_BARRIER, self.grad)
        if self.fl is not None:
            self._potential(self.fl, x32, x64, TERM_FRICTION, self.grad)
        return self.grad.view(-1, 3).clone()

    def _newton_contacts(self, x32, x64):
        """The proximity records at x followed by the lagged friction records with their
        motion at x (repulsion off), in one list for the assembly and the PCG."""
        n = int(self.cl.count.item())
        nf = 0 if self.fl is None else int(self.fl.count.item())
        if n + nf > self.comb.cap:
            self.comb.reserve(2 * (n + nf))
        self.comb.pairs[:16 * n].copy_(self.cl.pairs[:16 * n])
        if nf:
            out = self.comb.pairs[16 * n:]
            self.ctx.check(lib().ut_friction_lagged(self.ctx.handle, ptr(self.fl.pairs), ptr(self.fl.count),
                                                    self.fl.cap, None if x64 is not None else ptr(x32), ptr(x64),
                                                    None if x64 is not None else ptr(self.x_lag32),
                                                    ptr(self.x_lag64) if x64 is not None else None, ptr(out)))
        self.comb.count.fill_(n + nf)
        return self.comb

    def hvp(self, x, v):
        """H(x) v with the Newton matrix (assembled blocks + matrix-free contact blocks), FP64 [V, 3]."""
        x32, x64 = self._prepare(x)
        comb = self._newton_contacts(x32, x64)
        self._assemble(x32, x64, comb)
        vin = v.detach().to(torch.float64).reshape(-1).contiguous()
        out = torch.zeros_like(vin)
        self.ctx.check(lib().ut_pcg_set_contacts(self.pcg.handle, ptr(comb.pairs), ptr(comb.count), comb.cap))
        try:
            self.pcg.spmv(self.vals, vin, out)
        finally:
            self.ctx.check(lib().ut_pcg_set_contacts(self.pcg.handle, None, None, 0))
        return out.view(-1, 3)

    # compatibility: the proximity list of the last detection
    @property
    def pairs(self):
        return self.cl.pairs

    @property
    def count(self):
        return self.cl.count

    @property
    def cap(self):
        return self.cl.cap


class IPFunction(torch.autograd.Function):
    """E = IPFunction.apply(x, potential): x [V, 3] or [B, V, 3] on the device (FP32 or
    FP64); E is FP64 (scalar or [B]).  backward multiplies the hand-written gradient."""

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


class AbdPotential:
    """The incremental potential of affine bodies (initializer.BodySetup) as a function
    of q [nb, 12]:

        E(q) = sum_b (q_b - qt_b)^T (M4_b (x) I3) (q_b - qt_b) / (2 h^2)   (abd_inertia_energy.cpp:25-140)
             + sum_b kappa_b V_b ||A_b^T A_b - I||_F^2                    (abd_ortho_energy.cpp:25-140)
             + contact barrier and lagged friction at the vertex positions x(q) = p + A xbar

    Forward and backward run ut_abd_assemble (inertia and orthogonality), the
    proximity DCD at x(q) and ut_contact_potential (distances re-solved at x); the
    contact gradient reaches q through J^T (ut_abd_gather).  hvp applies the q-space
    Newton matrix (body blocks + the contacts' couplings, ut_abd_apply).  q may be
    FP32 or FP64; with FP64 q, x(q) is evaluated in FP64."""

    def __init__(self, ctx: Context, setup, q_tilde=None, inv_h2=3600.0, kappa=1e8, energy=BARRIER,
                 quasi_static=False, friction_mu=None, q_lag=None, contact_cap=1 << 16):
        from .abd import AffineBodies

        self.ctx, self.st = ctx, setup
        dv = ctx.device
        self.bodies = AffineBodies(ctx, setup)
        self.nb, self.nv = setup.nb, setup.nv
        st = setup
        names = dict(edges=st.edges, faces=st.faces, d_hat=st.d_hat, offset=st.offset, vprop=st.vprop,
                     va=st.rest_vert_area, ea=st.rest_edge_area, fa=st.rest_face_area)
        self.d = {k: _up(v, dv) for k, v in names.items()}
        self.q_tilde = None if q_tilde is None else torch.as_tensor(
            np.asarray(q_tilde, np.float64).reshape(-1), device=dv).contiguous()
        self.inv_h2, self.kappa, self.energy_type = float(inv_h2), float(kappa), int(energy)
        self.quasi_static = bool(quasi_static)
        self.narrow = NarrowPhasesDetector(ctx, self.nv, self.d["edges"], self.d["faces"], self.d["d_hat"],
                                           self.d["offset"], self.d["va"], self.d["ea"], self.d["fa"],
                                           vprop=self.d["vprop"], kappa=kappa, energy=energy)
        self.face_tree = LBVH(ctx, TREE_FACE, len(st.faces))
        self.edge_tree = LBVH(ctx, TREE_EDGE, len(st.edges))
        self.cl = _ContactList(dv, contact_cap)
        self.fl = None
        self.comb = _ContactList(dv, 2 * contact_cap)
        self.gx = torch.zeros(3 * self.nv, dtype=torch.float64, device=dv)
        self.pot = torch.zeros(2 + 512, dtype=torch.float64, device=dv)
        self._x = None
        self.x_lag32 = self.x_lag64 = None
        if friction_mu is not None:
            mu = np.broadcast_to(np.asarray(friction_mu, np.float32), (self.nv,))
            self.mu = torch.as_tensor(np.ascontiguousarray(mu), device=dv)
            if q_lag is None:
                raise ValueError("friction needs the lagged state q_lag")
            ql = torch.as_tensor(np.asarray(q_lag, np.float64).reshape(-1), device=dv)
            self.x_lag64 = self.bodies.q_to_x(ql, fp64=True).reshape(-1).contiguous()
            self.x_lag32 = self.x_lag64.to(torch.float32)
            self.fl = _ContactList(dv, contact_cap)
            self._detect(self.x_lag32, self.fl)
            self.ctx.check(lib().ut_friction_init(self.ctx.handle, ptr(self.fl.pairs), ptr(self.fl.count),
                                                  self.fl.cap, ptr(self.mu)))

    _detect = IncrementalPotential._detect

    def _positions(self, q):
        """(x32, x64 or None) at q."""
        if q.dtype == torch.float64:
            x64 = self.bodies.q_to_x(q, fp64=True).reshape(-1).contiguous()
            return x64.to(torch.float32), x64
        return self.bodies.q_to_x(q).reshape(-1).contiguous(), None

    def _prepare(self, q):
        x32, x64 = self._positions(q)
        if self._x is None or not torch.equal(self._x, x32):
            self._detect(x32, self.cl)
            self._x = x32.clone()
        return x32, x64

    _potential = IncrementalPotential._potential
    _newton_contacts = IncrementalPotential._newton_contacts

    def _bodies(self, q, contacts=None, blocks=False):
        c = contacts
        return self.bodies.assemble(q, self.q_tilde, self.inv_h2, None if c is None else c.pairs,
                                    None if c is None else c.count, 0 if c is None else c.cap,
                                    quasi_static=self.quasi_static, blocks=blocks)

    def energies(self, q):
        """(inertia, orthogonality, contact, friction) as FP64 tensors."""
        x32, x64 = self._prepare(q)
        _, _, e = self._bodies(q)
        e = e.clone()
        contact = self._potential(self.cl, x32, x64, TERM_BARRIER, None)[0]
        friction = (self._potential(self.fl, x32, x64, TERM_FRICTION, None)[1] if self.fl is not None
                    else torch.zeros((), dtype=torch.float64, device=self.ctx.device))
        return torch.stack([e[0], e[1], contact, friction])

    def energy(self, q):
        return self.energies(q).sum()

    def gradient(self, q):
        """dE/dq, FP64 [nb, 12]."""
        x32, x64 = self._prepare(q)
        g, _, _ = self._bodies(q)
        self.gx.zero_()
        self._potential(self.cl, x32, x64, TERM_BARRIER, self.gx)
        if self.fl is not None:
            self._potential(self.fl, x32, x64, TERM_FRICTION, self.gx)
        self.ctx.check(lib().ut_abd_gather(self.ctx.handle, C.byref(self.bodies.desc), ptr(self.gx), ptr(g), 1))
        return g.view(-1, 12).clone()

    def hvp(self, q, v):
        """The q-space Newton matrix times v, FP64 [nb, 12]."""
        x32, x64 = self._prepare(q)
        comb = self._newton_contacts(x32, x64)
        self._bodies(q, comb, blocks=True)
        return self.bodies.apply(v, comb.pairs, comb.count, comb.cap).view(-1, 12)


class AbdFunction(torch.autograd.Function):
    """E = AbdFunction.apply(q, potential): q [nb, 12] or [B, nb, 12]; backward multiplies
    the hand-written gradient dE/dq."""

    @staticmethod
    def forward(ctx, q, potential: AbdPotential):
        ctx.potential = potential
        ctx.save_for_backward(q)
        if q.dim() == 2:
            return potential.energy(q)
        return torch.stack([potential.energy(qb) for qb in q])

    @staticmethod
    def backward(ctx, g):
        (q,) = ctx.saved_tensors
        p = ctx.potential
        if q.dim() == 2:
            gq = p.gradient(q) * g
        else:
            gq = torch.stack([p.gradient(qb) * g[b] for b, qb in enumerate(q)])
        return gq.to(q.dtype), None
