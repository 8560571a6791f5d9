"""The differentiable affine-body potential (AbdPotential / AbdFunction) on the GPU:
ABD inertia + orthogonality + contact barrier + lagged friction (mu = 0.3) as a
function of q, on two stacked boxes and a sphere with every contact inside d_hat:

* per state of a batch of 16 FP64 states, the central differences (h = 1e-7 and
  5e-8, Richardson-combined) of the GPU energy along a random unit direction of q
  equal the backward's gradient along it;
* at two states the gradient equals the oracle's (abd_bodies + J^T of
  or_contact_potential's vertex gradient), and the q-space Newton product equals
  the oracle's (abd_bodies + abd_contacts blocks, abd_apply couplings)."""
import numpy as np
import pytest
import torch

import abd_cases
import oracle
from untangling26_amd.autograd import BARRIER, AbdFunction, AbdPotential

pytestmark = pytest.mark.gpu

MU = 0.3


@pytest.fixture(scope="module")
def case():
    sc = abd_cases.stacked_boxes()
    st = abd_cases.setup(sc, stiffness=1e5)
    rng = np.random.default_rng(2)
    Q = np.empty((16, st.nb, 12))
    for b in range(16):
        d = np.zeros((st.nb, 4, 3))
        d[:, 0] = rng.normal(0, 3e-5, (st.nb, 3))
        d[:, 1:] = rng.normal(0, 1e-4, (st.nb, 3, 3))
        Q[b] = st.q0 + d.reshape(st.nb, 12)
    return st, Q


def _potential(ctx, st):
    return AbdPotential(ctx, st, q_tilde=st.q0, inv_h2=3600.0, kappa=1e8, energy=BARRIER, friction_mu=MU,
                        q_lag=st.q0)


def test_central_differences_over_a_batch(ctx, case):
    st, Q = case
    pot = _potential(ctx, st)
    qb = torch.from_numpy(Q).cuda().requires_grad_(True)
    E = AbdFunction.apply(qb, pot)
    assert E.shape == (16,) and torch.isfinite(E).all()
    E.sum().backward()
    G = qb.grad.cpu().numpy()
    e0 = pot.energies(torch.from_numpy(Q[0]).cuda()).cpu().numpy()
    assert (e0 > 0).all(), e0  # inertia, orthogonality, contact and friction all active
    rng = np.random.default_rng(3)
    h = 1e-7
    worst = 0.0

    def central(q, v, step):
        return (pot.energy(torch.from_numpy(q + step * v).cuda()).item()
                - pot.energy(torch.from_numpy(q - step * v).cuda()).item()) / (2 * step)

    for b in range(16):
        v = rng.normal(size=Q[b].shape)
        v /= np.linalg.norm(v)
        # central differences at h and h / 2, Richardson-combined: the friction term in its
        # sticking range (s < 1e-4) has third derivatives ~ mu_lambda / 1e-8 that leave an
        # O(h^2) error of up to 1e-5 relative in a plain central difference (it falls 100x
        # per 10x smaller h); the combination removes it
        fd = (4.0 * central(Q[b], v, h / 2) - central(Q[b], v, h)) / 3.0
        an = float(np.sum(G[b] * v))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-3 * np.linalg.norm(G[b])))
    assert worst <= 1e-6, worst


def test_gradient_and_hvp_match_oracle(ctx, case):
    st, Q = case
    pot = _potential(ctx, st)
    xlag = oracle.abd_q_to_x(st.body, st.xbar, st.q0)
    fric = oracle.friction_init(abd_cases.records(st, xlag.astype(np.float32)), np.full(st.nv, MU, np.float32))
    for b in (0, 9):
        q = Q[b]
        x = oracle.abd_q_to_x(st.body, st.xbar, q)
        rec = abd_cases.records(st, x.astype(np.float32))
        _, _, gb, Hb = oracle.abd_bodies(q, st.q0, st.mass4, st.volume, st.kappa, 3600.0)
        gx = oracle.contact_potential(rec, x, st.d_hat, st.offset, terms=1)[2] + \
            oracle.contact_potential(fric, x, st.d_hat, st.offset, terms=2, x_start=xlag)[2]
        J = np.zeros((st.nb, 12))
        for v in range(st.nv):
            w = np.concatenate([[1.0], st.xbar[v].astype(np.float64)])
            J[st.body[v]] += np.kron(w, gx[v])
        want = gb + J
        g = pot.gradient(torch.from_numpy(q).cuda()).cpu().numpy()
        err = float(np.abs(g - want).max())
        assert err <= 1e-8 * np.abs(want).max(), (b, err)
        # Newton product: proximity records at x, then the lagged friction records with their motion
        comb = np.concatenate([rec, oracle.friction_lagged(fric, x, xlag)])
        _, Hc = oracle.abd_contacts(comb, st.body, st.xbar, st.nb)
        rng = np.random.default_rng(4 + b)
        p = rng.normal(size=12 * st.nb)
        y = pot.hvp(torch.from_numpy(q).cuda(), torch.from_numpy(p).cuda()).cpu().numpy().ravel()
        yw = oracle.abd_apply(Hb + Hc, comb, st.body, st.xbar, p)
        err = float(np.abs(y - yw).max())
        assert err <= 1e-8 * np.abs(yw).max(), (b, err)
