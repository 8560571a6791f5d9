"""GPU incremental potential behind torch.autograd (untangling26_amd.autograd)
against the FP64 oracle on two parallel sheets with 3.5k contacts inside d_hat:
the energies (inertia, stretch, bending from or_assemble; contact energy from the
distances re-solved at x), the hand-written gradient (or_assemble, itself pinned
to central differences on the CPU), the Hessian-vector product with the Newton
matrix (assembled blocks + rank-1 contact blocks), and the autograd plumbing
(single state and a batch)."""
import numpy as np
import pytest
import torch

import oracle
from autograd_cases import ip_case
from untangling26_amd.autograd import BARRIER, QUADRATIC, IncrementalPotential, IPFunction

pytestmark = pytest.mark.gpu


def _dense_hvp(st, vals, recs, v):
    """H v from planar SELL values plus the contact off-diagonal blocks w_a w_b k2 n n^T."""
    pat = st.pattern
    n = pat.n
    rows, cols = pat.keys // n, pat.keys % n
    B = np.asarray(vals)[:, pat.key_slots].T.reshape(-1, 3, 3)
    out = np.zeros((n, 3))
    np.add.at(out, rows, np.einsum("kij,kj->ki", B, v[cols]))
    for rec in recs:
        idx = [int(q & 0x7FFFFFFF) for q in rec[:4]]
        f = rec[4:12].view(np.float32).astype(np.float64)
        nn, k2, b0, b1 = f[0:3], f[5], f[6], f[7]
        w = (1.0, -b0, -b1, b0 + b1 - 1.0) if rec[2] & 0x80000000 else (b0, 1.0 - b0, -b1, b1 - 1.0)
        for a in range(4):
            for b in range(4):
                if a != b:
                    out[idx[a]] += w[a] * w[b] * k2 * nn * (nn @ v[idx[b]])
    return out


@pytest.mark.parametrize("energy", [BARRIER, QUADRATIC])
def test_potential_matches_oracle(ctx, energy):
    m, st, x_tilde, recs = ip_case(energy=energy)
    g_want, v_want, en = oracle.assemble(st, m.x, x_tilde, 3600.0, recs)
    ce_want = oracle.contact_energy(recs, m.x, st.d_hat, st.offset, energy=energy)
    ip = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=energy)
    x = torch.from_numpy(m.x).cuda()
    assert ip.contacts(x) == len(recs)
    e = ip.energies(x).cpu().numpy()
    for k in range(3):
        assert np.isclose(e[k], en[k], rtol=1e-9, atol=1e-18), (k, e[k], en[k])
    assert np.isclose(e[3], ce_want, rtol=1e-9), (e[3], ce_want)
    g = ip.gradient(x).cpu().numpy().ravel()
    assert np.abs(g - g_want).max() <= 1e-9 * np.abs(g_want).max()
    rng = np.random.default_rng(5)
    v = rng.normal(size=(m.nv, 3))
    hv = ip.hvp(x, torch.from_numpy(v).cuda()).cpu().numpy()
    want = _dense_hvp(st, v_want, recs, v)
    assert np.abs(hv - want).max() <= 1e-9 * np.abs(want).max()


def test_autograd_function(ctx):
    m, st, x_tilde, recs = ip_case(energy=BARRIER)
    ip = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=BARRIER)
    x = torch.from_numpy(m.x).cuda().requires_grad_(True)
    E = IPFunction.apply(x, ip)
    assert E.dtype == torch.float64 and E.dim() == 0
    E.backward()
    g = ip.gradient(x.detach())
    assert torch.equal(x.grad, g.to(torch.float32))
    # a batch of three states (the second and third slightly moved): energy per state,
    # gradient of a weighted sum
    rng = np.random.default_rng(9)
    xs = np.stack([m.x, m.x + rng.normal(0, 2e-5, m.x.shape), m.x + rng.normal(0, 2e-5, m.x.shape)]).astype(np.float32)
    xb = torch.from_numpy(xs).cuda().requires_grad_(True)
    Eb = IPFunction.apply(xb, ip)
    assert Eb.shape == (3,)
    for b in range(3):
        assert torch.isclose(Eb[b], ip.energy(xb[b].detach()), rtol=1e-12)
    wts = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64, device="cuda")
    (Eb * wts).sum().backward()
    for b in range(3):
        gb = ip.gradient(xb[b].detach()) * wts[b]
        assert torch.allclose(xb.grad[b].double(), gb, rtol=1e-6, atol=1e-6 * gb.abs().max().item())


def test_in_place_update_refreshes_contacts(ctx):
    """An in-place update of x between two evaluations (as an optimiser step does)
    must re-detect the contacts: the result equals a fresh potential's at the new x."""
    m, st, x_tilde, recs = ip_case(energy=BARRIER)
    ip = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=BARRIER)
    x = torch.from_numpy(m.x.copy()).cuda()
    ip.gradient(x)
    # move the upper sheet (the second half, 0.6 mm above in y) towards the lower one by
    # 15% of d_hat: every contact's distance, normal and k1 change
    upper = torch.arange(m.nv, device=x.device) >= m.nv // 2
    with torch.no_grad():
        x[upper, 1] -= 1.5e-4
    g = ip.gradient(x)
    e = ip.energies(x)
    fresh = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=BARRIER)
    xc = x.clone()
    g_want, e_want = fresh.gradient(xc), fresh.energies(xc)
    # FP64 atomics in the assembly sum in any order: equal to rounding, while a stale
    # contact list (k1 and normals of the old positions) is off by far more
    assert torch.allclose(g, g_want, rtol=1e-10, atol=1e-10 * g_want.abs().max().item())
    assert torch.allclose(e, e_want, rtol=1e-12)
    g_old = ip.gradient(torch.from_numpy(m.x.copy()).cuda())
    assert (g_old - g_want).abs().max() > 1e-3 * g_want.abs().max()
