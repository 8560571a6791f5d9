"""SPD 3x3-block systems with the cloth sparsity (implicit-Euler-like:
mass / h^2 on the diagonal + edge springs + bending couplings), seeded."""
import numpy as np


def spring_system(mesh, eaf, seed=0, mass=1e-5, h=1.0 / 60.0, k_stretch=1e3, k_bend=1e-2, eps=0.1):
    rng = np.random.default_rng(seed)
    x = mesh.x.astype(np.float64)
    n = mesh.nv
    rows, cols, blocks = [], [], []
    I3 = np.eye(3)
    # mass
    rows.append(np.arange(n))
    cols.append(np.arange(n))
    blocks.append(np.repeat((mass / h**2 * I3)[None], n, 0))

    def couple(a, b, K):
        rows.extend([a, b, a, b])
        cols.extend([a, b, b, a])
        blocks.extend([K, K, -K, -K])

    e = mesh.edges.astype(np.int64)
    d = x[e[:, 1]] - x[e[:, 0]]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ks = k_stretch * rng.uniform(0.5, 1.5, len(e))
    K = ks[:, None, None] * (d[:, :, None] * d[:, None, :] + eps * I3)
    couple(e[:, 0], e[:, 1], K)
    eaf = np.asarray(eaf, np.int64)
    inner = (eaf[:, 0] != 0xFFFFFFFF) & (eaf[:, 1] != 0xFFFFFFFF)
    f = mesh.faces.astype(np.int64)
    o0 = f[eaf[inner, 0]].sum(1) - e[inner].sum(1)
    o1 = f[eaf[inner, 1]].sum(1) - e[inner].sum(1)
    Kb = np.repeat((k_bend * I3)[None], inner.sum(), 0)
    couple(o0, o1, Kb)
    rows = np.concatenate(rows)
    cols = np.concatenate(cols)
    blocks = np.concatenate(blocks)
    b = rng.normal(0, 1, 3 * n) * (mass / h**2)
    return rows, cols, blocks, b
