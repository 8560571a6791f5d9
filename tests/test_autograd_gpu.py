"""GPU incremental potential behind torch.autograd (untangling26_amd.autograd)
against the FP64 oracle on two parallel sheets with 3.5k contacts inside d_hat:
the energies (inertia, stretch, bending from or_assemble; contact and friction
energies from the distances and motions at x, or_contact_potential), the
hand-written gradient (or_assemble's material terms + or_contact_potential, both
pinned to central differences on the CPU), the Hessian-vector product with the
Newton matrix (assembled blocks + the contacts' pair blocks, or_pair_block), and
the autograd function's gradient against the oracle's (single state and a
batch), with FP32 and FP64 positions."""
import numpy as np
import pytest
import torch

import oracle
from autograd_cases import ip_case
from untangling26_amd.autograd import BARRIER, QUADRATIC, IncrementalPotential, IPFunction

pytestmark = pytest.mark.gpu


def _dense_hvp(st, vals, recs, v):
    """H v from planar SELL values plus the contact off-diagonal blocks w_a w_b K."""
    pat = st.pattern
    n = pat.n
    rows, cols = pat.keys // n, pat.keys % n
    B = np.asarray(vals)[:, pat.key_slots].T.reshape(-1, 3, 3)
    out = np.zeros((n, 3))
    np.add.at(out, rows, np.einsum("kij,kj->ki", B, v[cols]))
    for rec in recs:
        idx = [int(q & 0x7FFFFFFF) for q in rec[:4]]
        K, _ = oracle.pair_block(rec)
        w = oracle.record_weights(rec)
        for a in range(4):
            for b in range(4):
                if a != b:
                    out[idx[a]] += w[a] * w[b] * (K @ v[idx[b]])
    return out


def _oracle_gradient(st, x, x_tilde, recs, fric=None, x_lag=None, energy=BARRIER):
    """Material gradient (or_assemble without contacts) + the contact potential's."""
    # or_assemble takes FP32 positions: callers pass states exactly representable in FP32
    g, _, en = oracle.assemble(st, np.asarray(x, np.float32), x_tilde, 3600.0, None)
    eb, _, gb = oracle.contact_potential(recs, x, st.d_hat, st.offset, energy=energy, terms=1)
    ef = 0.0
    gf = np.zeros_like(gb)
    if fric is not None:
        _, ef, gf = oracle.contact_potential(fric, x, st.d_hat, st.offset, energy=energy, terms=2, x_start=x_lag)
    return g.reshape(-1, 3) + gb + gf, en, eb, ef


@pytest.mark.parametrize("energy", [BARRIER, QUADRATIC])
def test_potential_matches_oracle(ctx, energy):
    m, st, x_tilde, recs = ip_case(energy=energy)
    g_want, en, ce_want, _ = _oracle_gradient(st, m.x.astype(np.float64), x_tilde, recs, energy=energy)
    _, v_want, _ = oracle.assemble(st, m.x, x_tilde, 3600.0, recs)
    ip = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=energy)
    x = torch.from_numpy(m.x).cuda()
    assert ip.contacts(x) == len(recs)
    e = ip.energies(x).cpu().numpy()
    for k in range(3):
        assert np.isclose(e[k], en[k], rtol=1e-9, atol=1e-18), (k, e[k], en[k])
    assert np.isclose(e[3], ce_want, rtol=1e-9), (e[3], ce_want)
    assert e[4] == 0.0
    g = ip.gradient(x).cpu().numpy()
    assert np.abs(g - g_want).max() <= 1e-9 * np.abs(g_want).max()
    rng = np.random.default_rng(5)
    v = rng.normal(size=(m.nv, 3))
    hv = ip.hvp(x, torch.from_numpy(v).cuda()).cpu().numpy()
    want = _dense_hvp(st, v_want, recs, v)
    assert np.abs(hv - want).max() <= 1e-9 * np.abs(want).max()


def test_friction_potential_matches_oracle(ctx):
    """Lagged friction (mu = 0.3) about the unperturbed sheets, evaluated at FP64
    positions moved by N(0, 0.05 mm): energies, gradient and the Newton-matrix product."""
    m, st, x_tilde, recs = ip_case(energy=BARRIER)
    x_lag = m.x.astype(np.float64)
    fric = oracle.friction_init(recs, np.full(m.nv, 0.3, np.float32))
    rng = np.random.default_rng(11)
    x = x_lag + rng.normal(0, 5e-5, x_lag.shape)
    x32 = x.astype(np.float32)
    # the proximity records at x (the GPU detects them at x rounded to FP32)
    from autograd_cases import records_at

    recs_x = records_at(st, x32, BARRIER)
    ip = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=BARRIER,
                              friction_mu=0.3, x_lag=torch.from_numpy(x_lag).cuda())
    xt = torch.from_numpy(x).cuda()
    e = ip.energies(xt).cpu().numpy()
    _, _, eb, ef = _oracle_gradient(st, x, x_tilde, recs_x, fric, x_lag)
    assert ef > 0 and np.isclose(e[4], ef, rtol=1e-9), (e[4], ef)
    assert np.isclose(e[3], eb, rtol=1e-9), (e[3], eb)
    # gradient at the FP32 state (exactly representable, so the oracle's FP32 assembly of
    # the material terms sees the same x): material + contact + friction terms; the FP64
    # path of the gradient is pinned by central differences (test_c4_differentiation_gpu)
    gm32 = ip.gradient(torch.from_numpy(x32).cuda()).cpu().numpy()
    gmat_o, _, _ = oracle.assemble(st, x32, x_tilde, 3600.0, None)
    gc32 = oracle.contact_potential(recs_x, x32.astype(np.float64), st.d_hat, st.offset, terms=1)[2] + \
        oracle.contact_potential(fric, x32.astype(np.float64), st.d_hat, st.offset, terms=2, x_start=x_lag)[2]
    want32 = gmat_o.reshape(-1, 3) + gc32
    err = float(np.abs(gm32 - want32).max())
    assert err <= 1e-9 * np.abs(want32).max(), err
    # the Newton matrix: proximity records at x followed by the lagged friction records
    v = rng.normal(size=(m.nv, 3))
    hv = ip.hvp(torch.from_numpy(x32).cuda(), torch.from_numpy(v).cuda()).cpu().numpy()
    lag32 = oracle.friction_lagged(fric, x32.astype(np.float64), x_lag)
    _, v_want32, _ = oracle.assemble(st, x32, x_tilde, 3600.0, np.concatenate([recs_x, lag32]))
    want = _dense_hvp(st, v_want32, np.concatenate([recs_x, lag32]), v)
    err = float(np.abs(hv - want).max())
    assert err <= 1e-8 * np.abs(want).max(), err


def test_autograd_function(ctx):
    m, st, x_tilde, recs = ip_case(energy=BARRIER)
    ip = IncrementalPotential(ctx, st, x_tilde=x_tilde, inv_h2=3600.0, kappa=1e8, energy=BARRIER)
    x = torch.from_numpy(m.x).cuda().requires_grad_(True)
    E = IPFunction.apply(x, ip)
    assert E.dtype == torch.float64 and E.dim() == 0
    E.backward()
    g_want, _, _, _ = _oracle_gradient(st, m.x.astype(np.float64), x_tilde, recs)
    g = x.grad.double().cpu().numpy()
    assert np.abs(g - g_want).max() <= 1e-6 * np.abs(g_want).max()  # x.grad is FP32 like x
    # a batch of three FP64 states (the second and third slightly moved): energy per state,
    # gradient of a weighted sum against the oracle's gradient at each state
    from autograd_cases import records_at

    rng = np.random.default_rng(9)
    xs = np.stack([m.x, m.x + rng.normal(0, 2e-5, m.x.shape), m.x + rng.normal(0, 2e-5, m.x.shape)])
    xs = xs.astype(np.float32).astype(np.float64)  # states exactly representable in FP32
    xb = torch.from_numpy(xs).cuda().requires_grad_(True)
    Eb = IPFunction.apply(xb, ip)
    assert Eb.shape == (3,)
    wts = np.array([1.0, -2.0, 0.5])
    (Eb * torch.from_numpy(wts).cuda()).sum().backward()
    for b in range(3):
        rb = records_at(st, xs[b].astype(np.float32), BARRIER)
        gw, en, eb, _ = _oracle_gradient(st, xs[b], x_tilde, rb)
        assert np.isclose(Eb[b].item(), sum(en) + eb, rtol=1e-9)
        got = xb.grad[b].cpu().numpy()
        assert np.abs(got - wts[b] * gw).max() <= 1e-9 * np.abs(wts[b] * gw).max()


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
