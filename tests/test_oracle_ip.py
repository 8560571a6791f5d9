"""The contact part of the incremental potential on the CPU: its gradient, as
assembled from the DCD records (sum over pairs of k1 w_k n, k1 = kappa A b'(d)),
is the FP64 central difference of the contact energy recomputed from x
(closest points re-solved at every evaluation), so the autograd backward that
returns the assembled gradient is the derivative of its forward energy."""
import numpy as np

import oracle
from autograd_cases import ip_case


def _record_gradient(recs, nv):
    g = np.zeros((nv, 3))
    for rec in recs:
        idx = [int(v & 0x7FFFFFFF) for v in rec[:4]]
        f = rec[4:12].view(np.float32).astype(np.float64)
        n, k1, b0, b1 = f[0:3], f[4], f[6], f[7]
        w = (1.0, -b0, -b1, b0 + b1 - 1.0) if rec[2] & 0x80000000 else (b0, 1.0 - b0, -b1, b1 - 1.0)
        for k in range(4):
            g[idx[k]] += k1 * w[k] * n
    return g


def test_contact_gradient_is_the_derivative_of_the_contact_energy():
    m, st, _, recs = ip_case(energy=1)
    x = m.x.astype(np.float64)
    g = _record_gradient(recs, m.nv)
    rng = np.random.default_rng(3)
    h = 1e-8
    for _ in range(3):
        v = rng.normal(size=x.shape)
        v /= np.linalg.norm(v)
        fd = (oracle.contact_energy(recs, x + h * v, st.d_hat, st.offset)
              - oracle.contact_energy(recs, x - h * v, st.d_hat, st.offset)) / (2 * h)
        an = float(np.sum(g * v))
        # records hold k1, n and the barycentrics in FP32
        assert abs(fd - an) <= 1e-5 * abs(an) + 1e-6, (fd, an)


def test_quadratic_contact_energy_closed_form():
    m, st, _, recs = ip_case(energy=0)
    # every record: kappa A (d - d_hat)^2 / 2 with d recomputed; k1 = kappa A (d - d_hat)
    e = oracle.contact_energy(recs, m.x, st.d_hat, st.offset, energy=0)
    f = recs[:, 4:12].view(np.float32).astype(np.float64)
    k1, k2 = f[:, 4], f[:, 5]
    want = float(np.sum(0.5 * k1 * k1 / k2))
    assert abs(e - want) <= 1e-5 * want
