"""Graph fixtures for the contour tests (shapes of the cases in
UnitTest/test_parallel_contour_construction.cpp:609-667, with our own seeds)."""
import numpy as np

# test_parallel_contour_construction.cpp:613-620
SMALL_FIXED_N = 14
SMALL_FIXED_EDGES = [(0, 1, 1), (1, 2, 0), (3, 4, 1), (4, 5, 1), (5, 3, 0), (7, 8, 0), (8, 9, 1), (9, 10, 1),
                     (11, 12, 1), (12, 13, 1), (13, 11, 1)]
SMALL_FIXED_COMPONENTS = [[0, 1, 2], [3, 4, 5], [6], [7, 8, 9, 10], [11, 12, 13]]


def make_undirected_csr(n, edges):
    """make_undirected_csr (test_parallel_contour_construction.cpp:29-57): both directions, input order."""
    row = np.zeros(n + 1, np.int64)
    for u, v, _ in edges:
        row[u + 1] += 1
        row[v + 1] += 1
    row = np.cumsum(row)
    col = np.zeros(row[-1], np.uint32)
    w = np.zeros(row[-1], np.uint32)
    cur = row[:-1].copy()
    for u, v, f in edges:
        col[cur[u]], w[cur[u]] = v, f & 1
        cur[u] += 1
        col[cur[v]], w[cur[v]] = u, f & 1
        cur[v] += 1
    return row.astype(np.uint32), col, w


def chains(num, length, rng):
    edges = []
    for c in range(num):
        base = c * length
        for i in range(length - 1):
            edges.append((base + i, base + i + 1, int(rng.integers(0, 2))))
    perm = rng.permutation(num * length)
    return num * length, [(int(perm[u]), int(perm[v]), w) for u, v, w in edges]


def grid(nx, ny, rng):
    """Consistent parity field: weight = parity(u) xor parity(v) for a random 0/1 labelling."""
    lab = rng.integers(0, 2, nx * ny)
    edges = []
    for i in range(nx):
        for j in range(ny):
            u = i * ny + j
            if i + 1 < nx:
                edges.append((u, u + ny, int(lab[u] ^ lab[u + ny])))
            if j + 1 < ny:
                edges.append((u, u + 1, int(lab[u] ^ lab[u + 1])))
    return nx * ny, edges


def random_sparse(n, deg, comps, rng):
    edges = []
    comp = rng.integers(0, comps, n)
    members = [np.nonzero(comp == c)[0] for c in range(comps)]
    for m in members:
        if len(m) < 2:
            continue
        for _ in range(len(m) * deg // 2):
            u, v = rng.choice(m, 2, replace=False)
            edges.append((int(u), int(v), int(rng.integers(0, 2))))
    return n, edges


def components_from_min(comp_min):
    groups = {}
    for v, m in enumerate(comp_min.tolist()):
        groups.setdefault(m, []).append(v)
    return sorted(groups.values())
