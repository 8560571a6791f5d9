"""GPU parity of the incremental-potential assembly (inertia, BW98 stretch and
shear, dihedral bending, contact gradient and diagonal blocks) against the
FP64 oracle (oracle_energy.c or_assemble, itself pinned against the
per-element functions on the CPU).  Both sum the same FP64 terms; the GPU sums them with atomics in
a nondeterministic order, so gradient and blocks agree within 1e-9 of the
largest magnitude."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from assembly_cases import contact_state
from untangling26_amd._lib import AssembleArgs, lib, ptr

pytestmark = pytest.mark.gpu


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype == torch.uint32:
        t = t.view(torch.int32)
    return t.cuda()


def test_assembly_matches_oracle(ctx):
    # contact records from the oracle's VF DCD on a folded cloth
    m, st, x_tilde, inv_h2, recs = contact_state(n=24)
    assert len(recs) > 0
    g_want, v_want, _ = oracle.assemble(st, m.x, x_tilde, inv_h2, recs)

    pat = st.pattern
    D = {k: dev(v) for k, v in dict(x=m.x, x_tilde=x_tilde, mass=st.mass, faces=st.faces, dm=st.dm,
                                     face_area=st.rest_face_area, mu_lambda=st.mu_lambda, face_slots=st.face_slots,
                                     hinges=st.hinges, rest_angle=st.rest_angle, bend=st.bend_stiff_area,
                                     hinge_slots=st.hinge_slots, diag_slot=pat.diag_slot,
                                     contacts=recs.view(np.int32)).items()}
    cnt = torch.tensor([len(recs)], dtype=torch.int32, device="cuda")
    grad = torch.zeros(3 * m.nv, dtype=torch.float64, device="cuda")
    vals = torch.zeros(9 * pat.nslots, dtype=torch.float64, device="cuda")
    vals32 = torch.zeros(9 * pat.nslots, dtype=torch.float32, device="cuda")
    a = AssembleArgs()
    a.nv, a.nf, a.nh, a.nslots = m.nv, m.nf, len(st.hinges), pat.nslots
    a.x, a.x_tilde, a.mass, a.inv_h2 = ptr(D["x"]), ptr(D["x_tilde"]), ptr(D["mass"]), inv_h2
    a.faces, a.dm, a.face_area, a.mu_lambda, a.face_slots = (ptr(D[k]) for k in
                                                             ("faces", "dm", "face_area", "mu_lambda", "face_slots"))
    a.hinges, a.rest_angle, a.bend_stiff_area, a.hinge_slots = (ptr(D[k]) for k in
                                                                ("hinges", "rest_angle", "bend", "hinge_slots"))
    a.contacts, a.contact_count, a.contact_cap = ptr(D["contacts"]), ptr(cnt), len(recs)
    a.diag_slot, a.grad, a.vals, a.vals32 = ptr(D["diag_slot"]), ptr(grad), ptr(vals), ptr(vals32)
    ctx.check(lib().ut_assemble(ctx.handle, C.byref(a)))
    g = grad.cpu().numpy()
    assert np.abs(g - g_want).max() <= 1e-9 * np.abs(g_want).max()
    V = vals.cpu().numpy().reshape(9, -1)
    assert np.abs(V - v_want).max() <= 1e-9 * np.abs(v_want).max()
    V32 = vals32.cpu().numpy().reshape(9, -1)
    assert np.array_equal(V32, V.astype(np.float32))
