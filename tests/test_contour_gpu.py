"""GPU parity of the contour stage: union-find components with parity against
the unit test's BFS reference, the EF preprocessing + contours against the
restated host code (exact order, adjacency and labels, with and without the
degenerate filter), and the EF pairs made from zero-distance proximity pairs."""
import numpy as np
import pytest
import torch

import oracle
from contour_cases import (SMALL_FIXED_COMPONENTS, SMALL_FIXED_EDGES, SMALL_FIXED_N, chains, components_from_min,
                           grid, make_undirected_csr, random_sparse)
from untangling26_amd import scenes, topology
from untangling26_amd.contours import (CT_ADJ, CT_CONTOUR_TYPE, CT_MESH_CLOSED, CT_ADJ_N, CT_ADJ_OFF, CT_CONTOUR_INDEX, CT_INCONSISTENT,
                                       CT_MESH_INDEX, CT_PAIRS, CT_PAIR_IDS, ContourBuilder, graph_components)
from untangling26_amd.intersection import EfDetector, add_ef_from_proximity_pairs
from untangling26_amd.lbvh import LBVH, TREE_EDGE, TREE_FACE
from untangling26_amd.narrow_phase import QUADRATIC, NarrowPhasesDetector
from zero_dist_cases import proximity_records, sheet_and_wall

pytestmark = pytest.mark.gpu


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype == torch.uint32:
        t = t.view(torch.int32)
    return t.cuda()


def u32(t):
    return t.cpu().numpy().view(np.uint32)


def _check_graph(ctx, n, edges):
    row, col, w = make_undirected_csr(n, edges)
    cm, par, inc = graph_components(ctx, dev(row), dev(col), dev(w))
    cm, par, inc = u32(cm), u32(par), u32(inc)
    rcm, rpar, rinc, _ = oracle.graph_bfs(row, col, w)
    assert np.array_equal(cm, rcm)
    consistent = rinc[rcm] == 0
    assert np.array_equal(par[consistent], rpar[consistent])
    assert np.array_equal(inc[rcm] != 0, rinc[rcm] != 0)
    return cm


def test_small_fixed_known_answer(ctx):
    cm = _check_graph(ctx, SMALL_FIXED_N, SMALL_FIXED_EDGES)
    assert components_from_min(cm) == SMALL_FIXED_COMPONENTS


@pytest.mark.parametrize("case", ["chains_10x10k", "chains_100x1k", "grid_100", "grid_200x250", "random_50k",
                                  "random_100k", "singletons_pairs"])
def test_graph_components_match_bfs(ctx, case):
    rng = np.random.default_rng(42)
    if case == "chains_10x10k":
        n, e = chains(10000, 10, rng)
    elif case == "chains_100x1k":
        n, e = chains(1000, 100, rng)
    elif case == "grid_100":
        n, e = grid(100, 100, rng)
    elif case == "grid_200x250":
        n, e = grid(200, 250, rng)
    elif case == "random_50k":
        n, e = random_sparse(50000, 4, 100, rng)
    elif case == "random_100k":
        n, e = random_sparse(100000, 6, 500, rng)
    else:
        n = 100000
        e = [(i, i + 1, int(rng.integers(0, 2))) for i in range(50000, n, 2)]
    _check_graph(ctx, n, e)


def _mesh(name):
    return {"two_sheets": lambda: scenes.two_sheets(0), "two_sheets_s3": lambda: scenes.two_sheets(3),
            "folded_cloth": lambda: scenes.folded_cloth(0), "nested_layers": lambda: scenes.nested_layers(0)}[name]()


def _check_build(cb, N, nc, in_ids, in_pairs, m, fae, eaf, filter_degenerate=False):
    """GPU contour build against the oracle on the same input: canonical list,
    packed adjacency, contour index and parity (on consistent contours)."""
    oid, opr, oadj, oadjn, kept = oracle.ef_preprocess(in_ids, in_pairs, m.edges, m.faces, fae, eaf,
                                                       filter_degenerate=filter_degenerate, with_kept=True)
    onc, oci, omi = oracle.contours_dfs(opr, oadj, oadjn)
    assert N == len(oid) and nc == onc
    counts = cb.counts()
    assert counts[0] == len(in_ids) and counts[1] == kept and counts[2] == N and counts[3] == len(oadj)
    assert np.array_equal(u32(cb.get(CT_PAIR_IDS)), oid)
    assert np.array_equal(u32(cb.get(CT_PAIRS)), opr)
    assert np.array_equal(u32(cb.get(CT_ADJ_N)), oadjn)
    assert np.array_equal(u32(cb.get(CT_ADJ)), oadj)
    off = np.concatenate([[0], np.cumsum(oadjn.astype(np.int64))])
    assert np.array_equal(u32(cb.get(CT_ADJ_OFF)), off)
    assert np.array_equal(u32(cb.get(CT_CONTOUR_INDEX)), oci)
    inc = u32(cb.get(CT_INCONSISTENT))
    mi = u32(cb.get(CT_MESH_INDEX))
    first = {}
    for i in range(N):
        first.setdefault(int(oci[i]), i)
    consistent = np.array([inc[first[int(c)]] == 0 for c in oci], bool)
    assert np.array_equal(mi[consistent], omi[consistent])
    # classify_contour_type / mesh_is_closed on the contours whose parity is defined
    ot, ocl = oracle.classify_contours(opr, oracle.pair_is_boundary(oid, eaf), oadjn, oci, omi, onc)
    gt, gcl = u32(cb.get(CT_CONTOUR_TYPE)), u32(cb.get(CT_MESH_CLOSED)).reshape(-1, 2)
    ok = np.array([inc[first[c]] == 0 for c in range(onc)], bool)
    assert gt.shape == (onc,) and np.array_equal(gt[ok], ot[ok])
    assert np.array_equal(gcl[ok], ocl[ok])
    return kept


@pytest.mark.parametrize("name", ["two_sheets", "two_sheets_s3", "folded_cloth", "nested_layers"])
def test_ef_contours_match_oracle(ctx, name):
    m = _mesh(name)
    d_hat, th = scenes.per_vertex(m)
    fae, eaf = topology.mesh_topology(m.edges, m.faces)
    E, F = dev(m.edges), dev(m.faces)
    det = EfDetector(ctx, m.nv, E, F)
    P = det.detect(dev(m.x), dev(d_hat), dev(th))
    cb = ContourBuilder(ctx, m.nv, E, F, dev(fae), dev(eaf))
    for ignore in (True, False):
        N, nc = cb.build(P, det.pairs, det.pair_ids, ignore_near_zero_dist_pairs=ignore)
        in_ids = u32(det.pair_ids[: 2 * P]).reshape(-1, 2)
        in_pairs = u32(det.pairs[: 12 * P]).reshape(-1, 12)
        assert nc > 0
        _check_build(cb, N, nc, in_ids, in_pairs, m, fae, eaf, filter_degenerate=not ignore)


def _rows(ids, pairs, words=12):
    """(id pair, record words) rows in lexicographic order, for multiset comparison."""
    r = np.concatenate([np.asarray(ids, np.uint32).reshape(-1, 2), np.asarray(pairs, np.uint32).reshape(-1, words)],
                       1)
    return r[np.lexsort(r.T[::-1])] if len(r) else r


@pytest.mark.parametrize("mixed", [False, True])
def test_ef_from_proximity_matches_oracle(ctx, mixed):
    """ut_ef_from_proximity on the oracle's proximity records of the wall scene,
    appended after the DCD EF pairs: the same multiset of pairs as the oracle."""
    m, _, _, _ = sheet_and_wall(mixed=mixed)
    st, recs, ids, _ = proximity_records(m)
    voff, vae = topology.vert_adj_edges(m.nv, st.edges)
    d_hat, th = scenes.per_vertex(m)
    E, F = dev(st.edges), dev(st.faces)
    det = EfDetector(ctx, m.nv, E, F)
    P = det.detect(dev(m.x), dev(d_hat), dev(th))
    zp, zi = oracle.ef_from_proximity(recs, ids, m.x, st.edges, st.faces, voff, vae, st.edge_adj_faces)
    assert len(zp) > 0
    cap = P + len(zp) + 64
    pairs = torch.zeros(12 * cap, dtype=torch.int32, device="cuda")
    pids = torch.zeros(2 * cap, dtype=torch.int32, device="cuda")
    pairs[: 12 * P] = det.pairs[: 12 * P]
    pids[: 2 * P] = det.pair_ids[: 2 * P]
    count = torch.tensor([P], dtype=torch.int32, device="cuda")
    rc = torch.tensor([len(recs)], dtype=torch.int32, device="cuda")
    add_ef_from_proximity_pairs(ctx, dev(recs.ravel()), dev(ids.ravel()), rc, dev(m.x), E, F, dev(voff), dev(vae),
                                dev(st.edge_adj_faces), count, pairs, pids)
    n = int(count.item())
    assert n == P + len(zp)
    assert np.array_equal(u32(pids[: 2 * P]), u32(det.pair_ids[: 2 * P]))
    got = _rows(u32(pids[2 * P: 2 * n]), u32(pairs[12 * P: 12 * n]))
    assert np.array_equal(got, _rows(zi, zp))
    # capacity overflow: the count still counts every pair, only cap are written
    count.fill_(0)
    small = 5
    add_ef_from_proximity_pairs(ctx, dev(recs.ravel()), dev(ids.ravel()), rc, dev(m.x), E, F, dev(voff), dev(vae),
                                dev(st.edge_adj_faces), count, pairs[: 12 * small], pids[: 2 * small])
    assert int(count.item()) == len(zp)


@pytest.mark.parametrize("mixed", [False, True])
def test_zero_distance_contours_match_oracle(ctx, mixed):
    """The wall scene through the GPU pipeline as the solver runs it with
    ignore_near_zero_dist_pairs = False: proximity DCD, EF DCD, zero-distance EF
    pairs, degenerate filter, contours; every stage against the oracle."""
    m, _, _, _ = sheet_and_wall(mixed=mixed)
    st, recs, ids, nvf = proximity_records(m)
    voff, vae = topology.vert_adj_edges(m.nv, st.edges)
    d_hat, th = scenes.per_vertex(m)
    x = dev(m.x)
    E, F, EAF = dev(st.edges), dev(st.faces), dev(st.edge_adj_faces)
    # proximity pairs on the GPU: the same records as the oracle's
    narrow = NarrowPhasesDetector(ctx, m.nv, E, F, dev(d_hat), dev(th), dev(st.rest_vert_area),
                                  dev(st.rest_edge_area), dev(st.rest_face_area), kappa=1e8, energy=QUADRATIC)
    ft, et = LBVH(ctx, TREE_FACE, m.nf), LBVH(ctx, TREE_EDGE, st.edges.shape[0])
    for tree, prims, red, upd in ((ft, F, "reduce_face_tree_aabb", "update_face_tree_leave_aabb"),
                                  (et, E, "reduce_edge_tree_aabb", "update_edge_tree_leave_aabb")):
        getattr(tree, red)(x, prims)
        tree.construct_tree()
        getattr(tree, upd)(dev(th), x, x, prims)
        tree.refit()
    cap = 4 * len(recs) + 64
    rc = torch.zeros(1, dtype=torch.int32, device="cuda")
    grec = torch.zeros(16 * cap, dtype=torch.int32, device="cuda")
    gid = torch.zeros(2 * cap, dtype=torch.int32, device="cuda")
    narrow.vf_dcd_query_repulsion(x, rc, grec, gid, ft)
    nvf_gpu = int(rc.item())
    narrow.ee_dcd_query_repulsion(x, rc, grec, gid, et)
    nr = int(rc.item())
    assert nvf_gpu == nvf and nr == len(recs)
    g, w = _rows(u32(gid[: 2 * nr]), u32(grec[: 16 * nr]), 16), _rows(ids, recs, 16)
    exact = [c for c in range(18) if c not in (10, 11)]  # record words 8, 9 (k1, k2) carry log/pow rounding
    assert np.array_equal(g[:, exact], w[:, exact])
    assert np.allclose(g[:, 10:12].view(np.float32), w[:, 10:12].view(np.float32), rtol=1e-5, atol=0)
    # EF pairs: DCD + zero-distance
    det = EfDetector(ctx, m.nv, E, F)
    P = det.detect(x, dev(d_hat), dev(th))
    zp, zi = oracle.ef_from_proximity(recs, ids, m.x, st.edges, st.faces, voff, vae, st.edge_adj_faces)
    ecap = P + len(zp) + 64
    pairs = torch.zeros(12 * ecap, dtype=torch.int32, device="cuda")
    pids = torch.zeros(2 * ecap, dtype=torch.int32, device="cuda")
    pairs[: 12 * P] = det.pairs[: 12 * P]
    pids[: 2 * P] = det.pair_ids[: 2 * P]
    count = torch.tensor([P], dtype=torch.int32, device="cuda")
    add_ef_from_proximity_pairs(ctx, grec, gid, rc, x, E, F, dev(voff), dev(vae), EAF, count, pairs, pids)
    n = int(count.item())
    assert n == P + len(zp)
    cb = ContourBuilder(ctx, m.nv, E, F, dev(st.face_adj_edges), EAF)
    N, nc = cb.build(n, pairs, pids, ignore_near_zero_dist_pairs=False)
    in_ids, in_pairs = u32(pids[: 2 * n]).reshape(-1, 2), u32(pairs[: 12 * n]).reshape(-1, 12)
    kept = _check_build(cb, N, nc, in_ids, in_pairs, st, st.face_adj_edges, st.edge_adj_faces,
                        filter_degenerate=True)
    if mixed:
        assert P < kept < n  # degenerate pairs both kept (next to DCD pairs) and dropped
    else:
        assert P == 0 and kept == n  # all degenerate: all stay
