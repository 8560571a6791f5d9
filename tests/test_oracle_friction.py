"""The oracle's lagged IPC friction (oracle_friction.c), pinned on the CPU by what
the mathematics fixes (friction_kernel.hpp:137-301, narrow_phase.cpp:1455-1521):

* closed forms of the smoothed friction in both regimes: sliding (s >= E):
  D = mu_lambda (s - E/3), gradient mu_lambda u / s, Hessian
  mu_lambda (P - u u^T / s^2) / s; sticking (s < E): D = mu_lambda (s^2/E - s^3/(3E^2));
  below s = 1e-6 energy and gradient vanish and the Hessian is mu_lambda (2/E - s/E^2) P;
* motion along the normal changes nothing;
* FP64 central differences: energy -> gradient -> Hessian, in both regimes;
* the Hessian is positive semidefinite;
* first-iteration values mu_lambda = -mu k1 (mean of the coefficients of record
  vertices 0 and 2), zero motion; later iterations: motion sum w (x - x_start), no
  repulsion;
* the contact potential's gradient (barrier re-solved at x, plus friction) is the
  central difference of its energy on the parallel-sheet proximity records;
* the assembled contact gradient / diagonal blocks and the PCG's off-diagonal
  contact product include the friction block (a wrong sign or a dropped term
  changes the Newton matrix and fails the dense comparison)."""
import numpy as np
import pytest

import oracle
from autograd_cases import ip_case

E = oracle.FRICTION_EPS


def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    t = np.cross(n, rng.normal(size=3))
    t /= np.linalg.norm(t)
    return n, t


@pytest.mark.parametrize("s", [3e-4, 2.5e-3])
def test_sliding_closed_form(s):
    rng = np.random.default_rng(1)
    n, t = _frame(rng)
    ml = 0.7
    rel = s * t + 0.3 * n
    e, g, h = oracle.friction_eval(ml, n, rel)
    assert e == pytest.approx(ml * (s - E / 3), rel=1e-12)
    assert np.allclose(g, ml * t, rtol=1e-12, atol=1e-15)
    P = np.eye(3) - np.outer(n, n)
    assert np.allclose(h, ml * (P - np.outer(t, t)) / s, rtol=1e-10, atol=1e-9)


@pytest.mark.parametrize("s", [2e-6, 4e-5, 9.9e-5])
def test_sticking_closed_form(s):
    rng = np.random.default_rng(2)
    n, t = _frame(rng)
    ml = 1.3
    e, g, h = oracle.friction_eval(ml, n, s * t - 0.01 * n)
    assert e == pytest.approx(ml * (s * s / E - s ** 3 / (3 * E * E)), rel=1e-10)
    assert np.allclose(g, ml * (2 / E - s / E ** 2) * s * t, rtol=1e-10)
    P = np.eye(3) - np.outer(n, n)
    # (f2 - f1/s) t t^T + (f1/s) P with f2 = 2/E - 2s/E^2, f1/s = 2/E - s/E^2
    want = ml * ((-s / E ** 2) * np.outer(t, t) + (2 / E - s / E ** 2) * P)
    assert np.allclose(h, want, rtol=1e-10, atol=1e-6)


def test_below_small_s_only_the_hessian_remains():
    n = np.array([0.0, 0.0, 1.0])
    rel = np.array([3e-7, 0.0, 5.0])
    e, g, h = oracle.friction_eval(2.0, n, rel)
    assert e == 0.0 and not g.any()
    P = np.diag([1.0, 1.0, 0.0])
    assert np.allclose(h, 2.0 * (2 / E - 3e-7 / E ** 2) * P, rtol=1e-12)


def test_normal_motion_changes_nothing():
    rng = np.random.default_rng(3)
    n, t = _frame(rng)
    a = oracle.friction_eval(0.9, n, 7e-5 * t)
    b = oracle.friction_eval(0.9, n, 7e-5 * t + 0.4 * n)
    assert a[0] == pytest.approx(b[0], rel=1e-9)
    assert np.allclose(a[1], b[1], rtol=1e-9, atol=1e-12) and np.allclose(a[2], b[2], rtol=1e-9, atol=1e-6)


@pytest.mark.parametrize("s", [3e-5, 6e-4])
def test_central_differences(s):
    rng = np.random.default_rng(4)
    n, t = _frame(rng)
    ml = 0.8
    rel = s * t + 0.02 * n
    e, g, H = oracle.friction_eval(ml, n, rel)
    h = 1e-6 * s
    for k in range(3):
        d = np.zeros(3)
        d[k] = h
        ep, gp, _ = oracle.friction_eval(ml, n, rel + d)
        em, gm, _ = oracle.friction_eval(ml, n, rel - d)
        assert (ep - em) / (2 * h) == pytest.approx(g[k], rel=1e-6, abs=1e-9)
        assert np.allclose((gp - gm) / (2 * h), H[:, k], rtol=1e-5, atol=1e-3 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()


def test_first_and_later_iteration_values():
    m, st, _, recs = ip_case(energy=1)
    mu = np.linspace(0.1, 0.9, m.nv).astype(np.float32)
    R = oracle.friction_init(recs, mu)
    f = R[:, 4:16].view(np.float32)
    k1 = recs[:, 4:16].view(np.float32)[:, 4]
    i0, i2 = R[:, 0] & 0x7FFFFFFF, R[:, 2] & 0x7FFFFFFF
    assert np.array_equal(f[:, 11], -(np.float32(0.5) * (mu[i0] + mu[i2])) * k1)
    assert not f[:, 8:11].any() and np.array_equal(R[:, :12], recs[:, :12])
    rng = np.random.default_rng(5)
    x0 = m.x.astype(np.float64)
    x = x0 + rng.normal(0, 1e-4, x0.shape)
    L = oracle.friction_lagged(R, x, x0)
    fl = L[:, 4:16].view(np.float32)
    assert not fl[:, 4:6].any() and np.array_equal(fl[:, 11], f[:, 11])
    for i in range(0, len(R), 37):
        w = oracle.record_weights(R[i])
        idx = (R[i, :4] & 0x7FFFFFFF).astype(np.int64)
        want = sum(w[k] * (x[idx[k]] - x0[idx[k]]) for k in range(4))
        assert np.allclose(fl[i, 8:11], want, rtol=1e-6, atol=1e-12)


def test_contact_potential_gradient_is_its_derivative():
    m, st, _, recs = ip_case(energy=1)
    x0 = m.x.astype(np.float64)
    fr = oracle.friction_init(recs, np.full(m.nv, 0.3, np.float32))
    rng = np.random.default_rng(6)
    x = x0 + rng.normal(0, 2e-5, x0.shape)
    allrec = np.concatenate([recs, oracle.friction_lagged(fr, x0, x0)])
    eb, ef, g = oracle.contact_potential(allrec, x, st.d_hat, st.offset, x_start=x0)
    assert eb > 0 and ef > 0
    h = 1e-9
    for _ in range(4):
        v = rng.normal(size=x.shape)
        v /= np.linalg.norm(v)
        ep = sum(oracle.contact_potential(allrec, x + h * v, st.d_hat, st.offset, x_start=x0)[:2])
        em = sum(oracle.contact_potential(allrec, x - h * v, st.d_hat, st.offset, x_start=x0)[:2])
        an = float(np.sum(g * v))
        assert (ep - em) / (2 * h) == pytest.approx(an, rel=1e-6), ((ep - em) / (2 * h), an)
    # the barrier part alone matches the record-based contact energy
    assert eb == pytest.approx(oracle.contact_energy(recs, x, st.d_hat, st.offset), rel=1e-12)


def test_assembly_and_pcg_carry_the_friction_block():
    m, st, x_tilde, recs = ip_case(energy=1)
    x0 = m.x.astype(np.float64)
    fr = oracle.friction_init(recs, np.full(m.nv, 0.5, np.float32))
    rng = np.random.default_rng(7)
    x = x0 + rng.normal(0, 3e-5, x0.shape)
    lag = oracle.friction_lagged(fr, x, x0)
    g, vals, _ = oracle.assemble(st, m.x, x_tilde, 3600.0, lag)
    # dense contact matrix of the lagged records: sum_c (W_c^T K_c W_c)
    n3 = 3 * m.nv
    A = np.zeros((n3, n3))
    gw = np.zeros(n3)
    for r in lag:
        K, gp = oracle.pair_block(r)
        w = oracle.record_weights(r)
        idx = (r[:4] & 0x7FFFFFFF).astype(np.int64)
        for a in range(4):
            gw[3 * idx[a]:3 * idx[a] + 3] += w[a] * gp
            for b in range(4):
                A[3 * idx[a]:3 * idx[a] + 3, 3 * idx[b]:3 * idx[b] + 3] += w[a] * w[b] * K
    g0, vals0, _ = oracle.assemble(st, m.x, x_tilde, 3600.0, None)
    assert np.allclose(g - g0, gw, rtol=1e-9, atol=1e-9 * np.abs(gw).max())
    # friction of the lagged records: the repulsion is off, so the contact gradient is the friction one
    _, ef, gf = oracle.contact_potential(lag, x, st.d_hat, st.offset, x_start=x0, terms=2)
    assert np.allclose(gw.reshape(-1, 3), gf, rtol=1e-5, atol=1e-6 * np.abs(gf).max())
    # the PCG's operator = assembled blocks + matrix-free off-diagonal contact blocks
    pat = st.pattern
    Ad = np.zeros((n3, n3))
    for key, slot in zip(pat.keys, pat.key_slots):
        i, j = divmod(int(key), m.nv)
        Ad[3 * i:3 * i + 3, 3 * j:3 * j + 3] = vals.reshape(9, -1)[:, slot].reshape(3, 3)
    diag = np.zeros((n3, n3))
    for v in range(m.nv):
        diag[3 * v:3 * v + 3, 3 * v:3 * v + 3] = A[3 * v:3 * v + 3, 3 * v:3 * v + 3]
    Afull = Ad + (A - diag)
    b = rng.normal(size=n3)
    xs = oracle.pcg_sell(pat, vals.reshape(9, -1), b, lag, max_iters=4000, rel_tol=1e-12)[0]
    assert np.linalg.norm(Afull @ xs - b) <= 1e-9 * np.linalg.norm(b)
