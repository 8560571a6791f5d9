"""Pins the CPU oracle against the reference's own known answers, closed forms
and brute force (no GPU needed)."""
import numpy as np
import pytest

import oracle


# lbvh.cpp:934-946: LBVH::unit_test expected Morton codes for pos = i / 10.
ANSWER_MORTON32 = [0, 2064888, 16519104, 117698623, 132152839, 939524096, 941588984, 956043200]
ANSWER_MORTON64 = [0, 8868626429902849, 70949011439222786, 505511736569233411, 567592121578553348,
                   4035225266123964421, 4044093892553867270, 4106174277563187207]
# lbvh.cpp:566-597: expected children / parents of the 8-leaf unit test tree.
ANSWER_CHILDREN = [[4, 5], [8, 7], [9, 1], [11, 10], [2, 3], [14, 6], [13, 12]]
ANSWER_PARENT = [0, 2, 4, 4, 0, 0, 5, 1, 1, 2, 3, 3, 6, 6, 5]


def unit_test_points():
    p = (np.arange(8, dtype=np.float32) / np.float32(10.0)).astype(np.float32)
    return np.repeat(p[:, None], 3, axis=1)


def test_morton_known_answers():
    m = oracle.morton30(unit_test_points())
    assert m.tolist() == ANSWER_MORTON32
    keys = (m.astype(np.uint64) << np.uint64(32)) | np.arange(8, dtype=np.uint64)
    assert keys.tolist() == ANSWER_MORTON64


def test_karras_known_answers():
    t = oracle.karras(np.array(ANSWER_MORTON64, dtype=np.uint64))
    assert t["children"][:7].tolist() == ANSWER_CHILDREN
    assert t["parent"][1:].tolist() == ANSWER_PARENT[1:]
    assert t["parent"][0] == 0xFFFFFFFF  # root has no parent (fn_reset_tree, lbvh.cpp:503)
    assert t["object_idx"][7:].tolist() == list(range(8))


def _karras_invariants(t, n):
    par, ch = t["parent"], t["children"]
    ni = n - 1
    seen = np.zeros(2 * n - 1, int)
    for i in range(ni):
        for c in ch[i]:
            assert par[c] == i
            seen[c] += 1
    assert seen[0] == 0 and (seen[1:] == 1).all()

    def leaves(node):
        if node >= ni:
            return [node - ni]
        return leaves(ch[node][0]) + leaves(ch[node][1])

    got = leaves(0)
    assert sorted(got) == list(range(n))
    # every subtree covers a contiguous range of sorted leaves
    for i in range(ni):
        lv = sorted(leaves(i))
        assert lv == list(range(lv[0], lv[-1] + 1))


@pytest.mark.parametrize("n", [2, 3, 5, 17, 64, 257])
def test_karras_invariants_random(n):
    rng = np.random.default_rng(n)
    codes = rng.integers(0, 1 << 30, size=n).astype(np.uint64)
    codes[: n // 3] = codes[0]  # duplicated Morton codes are separated by the index bits
    keys = np.sort((codes << np.uint64(32)) | np.arange(n, dtype=np.uint64))
    _karras_invariants(oracle.karras(keys), n)


def test_lbvh_build_sorted_and_consistent():
    from untangling26_amd import scenes

    m = scenes.two_sheets(3)
    t = oracle.lbvh_build(1, m.x, m.faces)
    ms = t["morton_sorted"]
    assert (np.diff(ms.astype(np.float64)) > 0).all() or (ms[1:] > ms[:-1]).all()
    assert sorted(t["sorted_get_original"].tolist()) == list(range(m.nf))
    _karras_invariants(t, m.nf)
    lo = np.minimum.reduce(m.x[m.faces.ravel()], 0)
    hi = np.maximum.reduce(m.x[m.faces.ravel()], 0)
    assert np.array_equal(t["global_box"], np.concatenate([lo, hi]).astype(np.float32))


@pytest.mark.parametrize("cull", [False, True])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_overlap_grid_equals_brute(seed, cull):
    rng = np.random.default_rng(seed)
    nq, nl = 300, 400
    c = rng.uniform(0, 1, (nq, 3)).astype(np.float32)
    q = np.concatenate([c - rng.uniform(0, 0.05, (nq, 3)), c + rng.uniform(0, 0.05, (nq, 3))], 1).astype(np.float32)
    c = rng.uniform(0, 1, (nl, 3)).astype(np.float32)
    lb = np.concatenate([c - rng.uniform(0, 0.1, (nl, 3)), c + rng.uniform(0, 0.1, (nl, 3))], 1).astype(np.float32)
    # touching boxes: strict comparisons must reject them
    lb[0, :3] = q[0, 3:]
    lb[0, 3:] = q[0, 3:] + 0.1
    lb[1] = q[1]  # identical boxes overlap
    g = oracle.overlap_pairs(q, lb, cull)
    b = oracle.overlap_pairs(q, lb, cull, brute=True)
    assert np.array_equal(g, b)
    assert len(b) > 100
    assert not ((b[:, 0] == 0) & (b[:, 1] == 0)).any()


def test_overlap_grid_equals_brute_on_sheets():
    from untangling26_amd import scenes

    m = scenes.two_sheets(1, n=12)
    d, th = scenes.per_vertex(m)
    q = oracle.query_boxes(2, m.x, m.x, m.edges, d, th)
    lb = oracle.leaf_boxes(1, m.x, m.x, m.faces, th)
    assert np.array_equal(oracle.overlap_pairs(q, lb), oracle.overlap_pairs(q, lb, brute=True))


def test_line_intersection_closed_form():
    A, B, C = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    # vertical segment through (0.25, 0.25) from z=-1 to z=3: t = 0.25, bary = (0.5, 0.25, 0.25)
    hit, bary, t = oracle.line_intersection([0.25, 0.25, -1], [0.25, 0.25, 3], A, B, C)
    assert hit and t == 0.25
    assert np.allclose(bary, [0.5, 0.25, 0.25], atol=1e-7)
    # misses: outside the triangle, parallel, short of the plane, degenerate
    assert not oracle.line_intersection([0.8, 0.8, -1], [0.8, 0.8, 1], A, B, C)[0]
    assert not oracle.line_intersection([0.2, 0.2, 1], [0.3, 0.2, 1], A, B, C)[0]
    assert not oracle.line_intersection([0.2, 0.2, 1], [0.2, 0.2, 2], A, B, C)[0]
    assert not oracle.line_intersection([0.2, 0.2, 0], [0.2, 0.2, 0], A, B, C)[0]
    # reversed segment: t measured from the start
    hit, _, t = oracle.line_intersection([0.25, 0.25, 3], [0.25, 0.25, -1], A, B, C)
    assert hit and t == 0.75


def _one_ef(p0, p1):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], p0, p1], np.float32)
    faces = np.array([[0, 1, 2]], np.uint32)
    edges = np.array([[3, 4]], np.uint32)
    d_hat = np.full(5, 1e-3, np.float32)
    return oracle.ef_dcd(np.array([[0, 0]], np.uint32), x, edges, faces, d_hat)


def test_ef_filters():
    pairs, ids = _one_ef([0.25, 0.25, -1], [0.25, 0.25, 1])
    assert ids.tolist() == [[0, 0]]
    w = pairs[0]
    assert w[:5].tolist() == [3, 4, 0, 1, 2]
    f = w[5:8].view(np.float32)
    assert f[0] == np.float32(0.5) and np.allclose(f[1:], [0.5, 0.25], atol=1e-7)
    # penetration of the shallow end below 1e-4 * sqrt(|N|) = 1e-4 -> rejected (grazing)
    assert len(_one_ef([0.25, 0.25, -5e-5], [0.25, 0.25, 1])[1]) == 0
    assert len(_one_ef([0.25, 0.25, -2e-4], [0.25, 0.25, 1])[1]) == 1
    # face barycentric below 1e-4 -> rejected
    assert len(_one_ef([0.99995e-4 / 2, 0.5, -1], [0.99995e-4 / 2, 0.5, 1])[1]) == 0
    # shared vertex -> rejected before the test
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.25, 0.25, 1]], np.float32)
    p, i = oracle.ef_dcd(np.array([[0, 0]], np.uint32), x, np.array([[0, 3]], np.uint32),
                         np.array([[0, 1, 2]], np.uint32), np.full(4, 1e-3, np.float32))
    assert len(i) == 0


def test_ef_detect_two_sheets_nonempty_and_consistent():
    from untangling26_amd import scenes

    m = scenes.two_sheets(0)
    d, th = scenes.per_vertex(m)
    pairs, ids = oracle.ef_detect(m.x, m.edges, m.faces, d, th)
    assert len(ids) > 10
    # brute force over all edges x faces gives the same set
    ne, nf = m.ne, m.nf
    rng = np.random.default_rng(0)
    sel = np.unique(ids[:, 0])
    broad = np.stack(np.meshgrid(sel, np.arange(nf), indexing="ij"), -1).reshape(-1, 2).astype(np.uint32)
    _, bid = oracle.ef_dcd(broad, m.x, m.edges, m.faces, d)
    assert np.array_equal(bid, ids)
    # edges of one sheet only cut faces of the other sheet
    half_e = m.edges[ids[:, 0], 0] < 32 * 32
    half_f = ids[:, 1] < m.nf // 2
    assert (half_e != half_f).all()
    del rng, ne


@pytest.mark.parametrize("kind,tree_type,cull", [(0, 1, False), (1, 2, True), (2, 1, False)])
def test_cpu_lbvh_traversal_equals_grid(kind, tree_type, cull):
    from untangling26_amd import scenes

    m = scenes.two_sheets(2)
    d, th = scenes.per_vertex(m, thickness=1e-4)
    prims = m.faces if tree_type == 1 else m.edges
    tree = oracle.lbvh_build(tree_type, m.x, prims)
    lb = oracle.leaf_boxes(tree_type, m.x, m.x, prims, th)
    qb = oracle.query_boxes(0 if kind == 0 else 1, m.x, m.x, None if kind == 0 else m.edges, d, th)
    a = oracle.lbvh_query(tree, lb, qb, cull)
    b = oracle.overlap_pairs(qb, lb, cull)
    assert np.array_equal(a, b) and len(a) > 0


def test_graph_bfs_small_fixed_known_answer():
    from contour_cases import (SMALL_FIXED_COMPONENTS, SMALL_FIXED_EDGES, SMALL_FIXED_N, components_from_min,
                               make_undirected_csr)

    row, col, w = make_undirected_csr(SMALL_FIXED_N, SMALL_FIXED_EDGES)
    cm, par, inc, any_inc = oracle.graph_bfs(row, col, w)
    assert components_from_min(cm) == SMALL_FIXED_COMPONENTS
    assert any_inc  # odd triangle 11-12-13
    assert inc[11] == 1 and inc[[0, 3, 6, 7]].sum() == 0
    # chain A: 0 -1- 1 -0- 2 ; cycle B consistent: 3 -1- 4 -1- 5 -0- 3
    assert par[[0, 1, 2]].tolist() == [0, 1, 1]
    assert par[[3, 4, 5]].tolist() == [0, 1, 0]


def _ef_fixture(name):
    from untangling26_amd import scenes

    m = {"two_sheets": lambda: scenes.two_sheets(0), "two_sheets_s3": lambda: scenes.two_sheets(3),
         "folded_small": lambda: scenes.folded_cloth(0, n=64)}[name]()
    d, th = scenes.per_vertex(m)
    pairs, ids = oracle.ef_detect(m.x, m.edges, m.faces, d, th)
    fae, eaf = oracle.mesh_topology(m.edges, m.faces)
    return m, pairs, ids, fae, eaf


@pytest.mark.parametrize("name", ["two_sheets", "two_sheets_s3", "folded_small"])
def test_ef_preprocess_and_dfs_consistent(name):
    m, pairs, ids, fae, eaf = _ef_fixture(name)
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(ids))  # input order must not matter
    oid, opr, adj, adjn = oracle.ef_preprocess(ids[perm], pairs[perm], m.edges, m.faces, fae, eaf)
    P = len(ids)
    assert np.array_equal(oid[:P], ids)  # canonical order = sorted (eid, fid)
    # adjacency is symmetric
    N = len(oid)
    nb = [set(r.tolist()) for r in oracle.adj_rows(adj, adjn)]
    for i in range(N):
        for j in nb[i]:
            assert i in nb[j]
    nc, ci, mi = oracle.contours_dfs(opr, adj, adjn)
    # DFS agrees with the unit test's BFS reference on components and parity
    edges = [(i, j, int(not np.array_equal(opr[i, 2:5], opr[j, 2:5]))) for i in range(N) for j in nb[i] if i < j]
    from contour_cases import make_undirected_csr

    row, col, w = make_undirected_csr(N, edges)
    cm, par, inc, any_inc = oracle.graph_bfs(row, col, w)
    first = {}
    for i in range(N):
        first.setdefault(int(ci[i]), i)
    assert all(cm[i] == first[int(ci[i])] for i in range(N))
    for i in range(N):
        if not inc[cm[i]]:
            assert par[i] == mi[i]
    assert nc >= 1
    if name.startswith("two_sheets"):
        # two flat sheets cross along one open curve: chain interior pairs have two neighbours
        deg = adjn[:N]
        assert (deg <= 2).all() and (deg == 1).sum() % 2 == 0
        assert N == P  # no loop vertices between two separate sheets


def _small_system(seed=0, n_side=8):
    from pcg_cases import spring_system
    from untangling26_amd import scenes, topology

    m = scenes.two_sheets(seed, n=n_side)
    _, eaf = topology.mesh_topology(m.edges, m.faces)
    rows, cols, blocks, b = spring_system(m, eaf, seed=seed, k_stretch=10.0)
    return m.nv, rows, cols, blocks, b


@pytest.mark.parametrize("precond", [0, 1])
def test_pcg_oracle_matches_direct_solve(precond):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    n, rows, cols, blocks, b = _small_system()
    rp, ci, bl = oracle.block_csr(n, rows, cols, blocks)
    A = sp.bsr_matrix((bl.reshape(-1, 3, 3), ci, rp), shape=(3 * n, 3 * n)).tocsc()
    xs = spla.spsolve(A, b)
    x, it, rel, conv, fail = oracle.pcg(n, rp, ci, bl, b, precond=precond, max_iters=5000, rel_tol=1e-10)
    assert conv and fail == 0 and rel <= 1e-10
    assert np.linalg.norm(x - xs) <= 1e-6 * np.linalg.norm(xs)


def test_pcg_oracle_block_diagonal_one_iteration():
    rng = np.random.default_rng(1)
    n = 37
    blocks = []
    for _ in range(n):
        q = rng.normal(size=(3, 3))
        blocks.append(q @ q.T + 3 * np.eye(3))
    rp, ci, bl = oracle.block_csr(n, np.arange(n), np.arange(n), np.array(blocks))
    b = rng.normal(size=3 * n)
    x, it, rel, conv, fail = oracle.pcg(n, rp, ci, bl, b, precond=0, max_iters=10, rel_tol=1e-12)
    assert it == 1 and conv
    want = np.concatenate([np.linalg.solve(blocks[i], b[3 * i:3 * i + 3]) for i in range(n)])
    assert np.allclose(x, want, rtol=1e-12, atol=1e-14)


def test_pcg_oracle_edge_cases():
    n, rows, cols, blocks, b = _small_system(1)
    rp, ci, bl = oracle.block_csr(n, rows, cols, blocks)
    x, it, rel, conv, fail = oracle.pcg(n, rp, ci, bl, np.zeros_like(b))
    assert conv and it == 0 and not x.any()
    x, it, rel, conv, fail = oracle.pcg(n, rp, ci, bl, b, max_iters=0)
    assert not conv and fail == 3 and rel == 1.0
    bad = b.copy()
    bad[5] = np.nan
    x, it, rel, conv, fail = oracle.pcg(n, rp, ci, bl, bad)
    assert not conv and fail == 1


def test_point_triangle_coeff_closed_forms():
    t0, t1, t2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    assert np.allclose(oracle.pt_coeff([0.25, 0.25, 0.7], t0, t1, t2), [0.5, 0.25, 0.25])  # interior
    assert np.allclose(oracle.pt_coeff([0.5, -0.3, 0.2], t0, t1, t2), [0.5, 0.5, 0.0])  # beyond edge t0-t1
    assert np.allclose(oracle.pt_coeff([2.0, -0.5, 0.0], t0, t1, t2), [0.0, 1.0, 0.0])  # beyond vertex t1
    assert np.allclose(oracle.pt_coeff([-0.2, 0.5, 0.1], t0, t1, t2), [0.5, 0.0, 0.5])  # beyond edge t0-t2
    # brute force over a barycentric grid wherever at most one planar barycentric is negative
    # (the classification of distance.hpp:96-160 only looks at the first negative one)
    rng = np.random.default_rng(0)
    checked = 0
    u, v = np.meshgrid(np.linspace(0, 1, 201), np.linspace(0, 1, 201))
    ok = u + v <= 1
    while checked < 50:
        P = rng.normal(size=(4, 3))
        A = np.stack([P[2] - P[1], P[3] - P[1]], 1)
        cc = np.linalg.lstsq(A, P[0] - P[1], rcond=None)[0]
        plane = np.array([1 - cc.sum(), cc[0], cc[1]])
        if (plane < 0).sum() > 1:
            continue
        checked += 1
        c = oracle.pt_coeff(P[0], P[1], P[2], P[3])
        d = np.linalg.norm(P[0] - c @ P[1:])
        pts = (1 - u - v)[ok, None] * P[1] + u[ok, None] * P[2] + v[ok, None] * P[3]
        assert d <= np.linalg.norm(pts - P[0], axis=1).min() + 1e-9
    # two negative barycentrics: the reference keeps the edge opposite the first negative one
    # (here edge t1-t2) although vertex t2 is closer -- reproduced deliberately
    c = oracle.pt_coeff([-1.0, 3.0, 0.0], t0, t1, t2)
    assert np.allclose(c, [0.0, 0.0, 1.0]) or c[0] == 0.0


def test_edge_edge_coeff_closed_forms():
    # skew perpendicular segments one unit apart
    c = oracle.ee_coeff([-1, 0, 0], [1, 0, 0], [0, -1, 1], [0, 1, 1])
    assert np.allclose(c, [0.5, 0.5, 0.5, 0.5])
    # parallel segments: falls back to point-edge candidates, distance = separation
    a0, a1, b0, b1 = map(np.array, ([0, 0, 0], [1, 0, 0], [0.25, 0.5, 0], [2, 0.5, 0]))
    c = oracle.ee_coeff(a0, a1, b0, b1)
    xa = c[0] * a0 + c[1] * a1
    xb = c[2] * b0 + c[3] * b1
    assert np.isclose(np.linalg.norm(xa - xb), 0.5)
    rng = np.random.default_rng(1)
    for _ in range(50):
        P = rng.normal(size=(4, 3))
        c = oracle.ee_coeff(*P)
        d = np.linalg.norm((c[0] * P[0] + c[1] * P[1]) - (c[2] * P[2] + c[3] * P[3]))
        s, t = np.meshgrid(np.linspace(0, 1, 401), np.linspace(0, 1, 401))
        pa = (1 - s)[..., None] * P[0] + s[..., None] * P[1]
        pb = (1 - t)[..., None] * P[2] + t[..., None] * P[3]
        assert d <= np.linalg.norm(pa - pb, axis=-1).min() + 1e-9


@pytest.mark.parametrize("d", [1e-6, 3e-5, 2e-4, 7e-4, 9.9e-4])
def test_barrier_derivatives_central_differences(d):
    dhat = 1e-3
    h = 1e-7 * d
    fd1 = (oracle.barrier(d + h, dhat) - oracle.barrier(d - h, dhat)) / (2 * h)
    fd2 = (oracle.barrier(d + h, dhat, 1) - oracle.barrier(d - h, dhat, 1)) / (2 * h)
    assert np.isclose(oracle.barrier(d, dhat, 1), fd1, rtol=1e-6, atol=1e-12)
    assert np.isclose(oracle.barrier(d, dhat, 2), fd2, rtol=1e-6, atol=1e-9)
    assert oracle.barrier(dhat, dhat) == 0.0 and oracle.barrier(dhat, dhat, 1) == 0.0


def test_accd_point_plane_constant_velocity():
    # point above the triangle plane moving straight down at speed v
    tri = np.array([[-1, -1, 0], [2, -1, 0], [-1, 2, 0]], float)
    for d0, v, off in [(0.1, 0.5, 0.0), (0.1, 0.5, 0.01), (0.05, 1.0, 0.0), (0.3, 0.2, 0.0)]:
        x0 = np.concatenate([[0.2, 0.2, d0], tri.ravel()])
        dx = np.concatenate([[0, 0, -v], np.zeros(9)])
        t = oracle.accd(False, x0, dx, off)
        t_true = (d0 - off) / v
        if t_true >= 1.0:
            assert t == 1.0
        else:
            assert 0 < t <= t_true
            assert d0 - v * t >= off  # never closer than the offset
            assert t >= 0.9 * t_true  # conservative advancement gets within a few percent
    # no relative motion -> 1 (accd.hpp:169-175)
    x0 = np.concatenate([[0.2, 0.2, 0.1], tri.ravel()])
    assert oracle.accd(False, x0, np.tile([0.3, 0.1, -0.2], 4), 0.0) == 1.0
    # post-processing of the reference (narrow_phase.cpp:473-522)
    assert oracle.toi_postprocess(0.001) == 1.0 and oracle.toi_postprocess(1e-3 * 2) == 0.5
    assert oracle.toi_postprocess(-1) == 1.0 and oracle.toi_postprocess(0.4) == 0.4


def _fd_grad(f, x, h):
    g = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        g.flat[i] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_bw98_gradient_and_hessian():
    from untangling26_amd import initializer

    rng = np.random.default_rng(0)
    rest = np.array([[0, 0, 0], [0.01, 0, 0], [0.003, 0.009, 0]], np.float64)
    dm = initializer.dm_inv(rest, np.array([[0, 1, 2]]))[0]
    area = 0.5 * np.linalg.norm(np.cross(rest[1] - rest[0], rest[2] - rest[0]))
    mu, lam = 370.0, 864.0
    e0, g0, _ = oracle.bw98_face(rest, dm, area, mu, lam)
    assert abs(e0) < 1e-12 * mu * area and np.abs(g0).max() < 1e-6 * mu * 0.01  # rest state
    for trial in range(20):
        x = rest + rng.normal(0, 2e-3, rest.shape)
        e, g, H = oracle.bw98_face(x, dm, area, mu, lam)
        fd = _fd_grad(lambda y: oracle.bw98_face(y, dm, area, mu, lam)[0], x, 1e-7 * 0.01)
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(g).max())
        Hm = H.transpose(0, 2, 1, 3).reshape(9, 9)
        assert np.allclose(Hm, Hm.T, atol=1e-9 * np.abs(Hm).max())
        assert np.linalg.eigvalsh(Hm).min() >= -1e-8 * np.abs(Hm).max()
    # pure stretch along the fibres (no shear, |Fu|, |Fv| > 1): the stretch projection is inactive and
    # the shear term vanishes to second order, so the Hessian equals the true one
    x = rest * np.array([1.2, 1.1, 1.0])
    _, _, H = oracle.bw98_face(x, dm, area, mu, 0.0)
    Hm = H.transpose(0, 2, 1, 3).reshape(9, 9)
    fdH = np.stack([_fd_grad(lambda y: oracle.bw98_face(y, dm, area, mu, 0.0)[1].ravel()[i], x, 1e-9).ravel()
                    for i in range(9)])
    assert np.allclose(Hm, fdH, rtol=1e-4, atol=1e-4 * np.abs(Hm).max())


def test_bending_gradient_and_rest_angle():
    from untangling26_amd import initializer

    rng = np.random.default_rng(1)
    # hinge: edge along x, two wings
    e0, e1 = np.array([0.0, 0, 0]), np.array([0.01, 0, 0])
    o0, o1 = np.array([0.004, 0.008, 0.0]), np.array([0.005, -0.008, 0.002])
    x = np.stack([e0, e1, o0, o1])
    th_np = initializer.dihedral_angles(x, np.array([[0, 1, 2, 3]]))[0]
    _, _, _, th = oracle.bending_hinge(x, 0.0, 1.0)
    assert np.isclose(th, th_np, atol=1e-6)
    e, g, H, _ = oracle.bending_hinge(x, th, 1.0)
    assert e == 0.0 and not g.any()
    for _ in range(20):
        y = x + rng.normal(0, 1e-3, x.shape)
        e, g, H, th = oracle.bending_hinge(y, 0.1, 2.0)
        fd = _fd_grad(lambda z: oracle.bending_hinge(z, 0.1, 2.0)[0], y, 1e-9)
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * np.abs(g).max())
        Hm = H.transpose(0, 2, 1, 3).reshape(12, 12)
        assert np.linalg.eigvalsh(Hm).min() >= -1e-9 * np.abs(Hm).max()


def test_assemble_equals_term_sum():
    """The C assembly into SELL slots equals the element-by-element sum of the
    FD-pinned per-element functions (pins the slot scatter and contact weights)."""
    from assembly_cases import contact_state, term_sum

    m, st, x_tilde, inv_h2, recs = contact_state(n=12)
    assert len(recs) > 0
    g, vals, en = oracle.assemble(st, m.x, x_tilde, inv_h2, recs)
    g_want, b_want = term_sum(st, m.x, x_tilde, inv_h2, recs)
    assert np.allclose(g.reshape(-1, 3), g_want, rtol=0, atol=1e-12 * np.abs(g_want).max())
    r, c, B = st.pattern.blocks_of(vals)
    want = np.stack([b_want.get((int(i), int(j)), np.zeros((3, 3))) for i, j in zip(r, c)])
    assert np.allclose(B, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert set(b_want) <= set(zip(r.tolist(), c.tolist()))
    # padding slots stay zero
    mask = np.ones(st.pattern.nslots, bool)
    mask[st.pattern.key_slots] = False
    assert not vals[:, mask].any()
    # inertia energy closed form
    d = m.x.astype(np.float64).ravel() - x_tilde
    assert np.isclose(en[0], 0.5 * inv_h2 * (np.repeat(st.mass.astype(np.float64), 3) * d * d).sum(), rtol=1e-12)


def test_pcg_with_contact_offdiagonals_matches_dense_solve():
    """The matrix-free contact off-diagonal blocks w_a w_b k2 n n^T (a != b),
    added to the assembled matrix, give the dense system a direct solve answers."""
    from assembly_cases import contact_state, contact_weights

    m, st, x_tilde, inv_h2, recs = contact_state(n=12)
    g, vals, _ = oracle.assemble(st, m.x, x_tilde, inv_h2, recs)
    n = st.nv
    A = np.zeros((3 * n, 3 * n))
    r, c, B = st.pattern.blocks_of(vals)
    for i, j, blk in zip(r, c, B):
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] += blk
    for rec in recs:
        idx, w, nn, _, k2 = contact_weights(rec)
        for a in range(4):
            for b in range(4):
                if a != b:
                    A[3 * idx[a]:3 * idx[a] + 3, 3 * idx[b]:3 * idx[b] + 3] += w[a] * w[b] * k2 * np.outer(nn, nn)
    want = np.linalg.solve(A, -g)
    x, it, rel, conv, fail = oracle.pcg_sell(st.pattern, vals, -g, recs, max_iters=20000, rel_tol=1e-12)
    assert conv and fail == 0
    assert np.linalg.norm(x - want) <= 1e-8 * np.linalg.norm(want)
    # without the off-diagonal contact terms the answer is measurably different
    x2, *_ = oracle.pcg_sell(st.pattern, vals, -g, None, max_iters=20000, rel_tol=1e-12)
    assert np.linalg.norm(x2 - want) > 1e-6 * np.linalg.norm(want)
