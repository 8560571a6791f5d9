"""Oracle pins for the zero-distance EF pairs and the degenerate-pair filter:
fn_add_ef_from_proximity_pairs (Solver/CollisionDetector/intersection_resolver.cpp:199-363),
is_degenerate_ef_pair (:561-569) and the filter and canonical order of
host_preprocess_ef_pairs (:695-726, :841-897).  The expected pairs of the
closed-form cases are listed by hand from the meshes."""
import numpy as np
import pytest

import oracle
from untangling26_amd import scenes
from zero_dist_cases import TYPE_BIT, proximity_records, sheet_and_wall


def f32(*v):
    return np.array(v, np.float32).view(np.uint32)


def pair_words(idx, bary):
    w = np.zeros(12, np.uint32)
    w[:5] = idx
    w[5:8] = f32(*bary)
    return w


def vf_rec(v, f3, b0, b1):
    r = np.zeros(16, np.uint32)
    r[:4] = [v, f3[0], f3[1] | TYPE_BIT, f3[2]]
    r[10:12] = f32(b0, b1)
    return r


def ee_rec(e1, e2, b0, b1):
    r = np.zeros(16, np.uint32)
    r[:4] = [e1[0], e1[1], e2[0], e2[1] | TYPE_BIT]
    r[10:12] = f32(b0, b1)
    return r


def test_degenerate_predicate_closed_form():
    assert not oracle.ef_degenerate(pair_words([0] * 5, (0.5, 0.3, 0.3)))
    assert oracle.ef_degenerate(pair_words([0] * 5, (0.99995, 0.3, 0.3)))  # edge bary (0.99995, 5e-5)
    assert oracle.ef_degenerate(pair_words([0] * 5, (0.00005, 0.3, 0.3)))
    assert oracle.ef_degenerate(pair_words([0] * 5, (0.5, 0.6, 0.39995)))  # third face bary 5e-5
    assert oracle.ef_degenerate(pair_words([0] * 5, (0.5, 0.0, 0.5)))
    # the threshold is strict: 1e-4 itself is not degenerate
    one_minus = float(np.float32(1.0) - np.float32(1e-4))
    assert not oracle.ef_degenerate(pair_words([0] * 5, (1e-4, 0.3, 0.3)))
    assert not oracle.ef_degenerate(pair_words([0] * 5, (0.5, 1e-4, one_minus - 2e-4)))


def _tent():
    """Face 0 = (0, 1, 2) in the plane z = 0; a tent of faces (3, 4, 5), (4, 3, 6)
    whose apex vertex 3 lies on face 0 at (0.25, 0.25)."""
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.25, 0.25, 0], [0.25, 0.75, 0.5], [0.7, 0.2, 0.5],
                  [-0.2, 0.2, 0.5]], np.float32)
    F = np.array([[0, 1, 2], [3, 4, 5], [4, 3, 6]], np.uint32)
    E = np.array([[0, 1], [1, 2], [2, 0], [3, 4], [4, 5], [5, 3], [3, 6], [6, 4]], np.uint32)
    fae, eaf = oracle.mesh_topology(E, F)
    return x, E, F, fae, eaf


def test_vf_zero_distance_closed_form():
    x, E, F, _, eaf = _tent()
    off, col = oracle.vert_adj_edges(len(x), E)
    assert list(col[off[3]:off[4]]) == [3, 5, 6]
    # point = 0.5 x0 + 0.25 x1 + 0.25 x2 = (0.25, 0.25, 0) = x3
    recs = vf_rec(3, (0, 1, 2), 0.5, 0.25)[None]
    pr, pid = oracle.ef_from_proximity(recs, np.array([[3, 0]], np.uint32), x, E, F, off, col, eaf)
    want_ids = [[3, 0], [5, 0], [6, 0]]
    want = [pair_words([3, 4, 0, 1, 2], (1.0, 0.5, 0.25)),  # vertex 3 first on edge (3, 4): bary (1, 0)
            pair_words([5, 3, 0, 1, 2], (0.0, 0.5, 0.25)),  # second on edge (5, 3): bary (0, 1)
            pair_words([3, 6, 0, 1, 2], (1.0, 0.5, 0.25))]
    assert pid.tolist() == want_ids
    assert np.array_equal(pr, np.array(want))
    assert all(oracle.ef_degenerate(p) for p in pr)  # on the edge's end: always degenerate


@pytest.mark.parametrize("z,hit", [(9e-6, True), (1.2e-5, False), (-9e-6, True), (0.01, False)])
def test_distance_threshold(z, hit):
    # |d|^2 < 1e-10: 9e-6 (8.1e-11) passes, 1.2e-5 (1.44e-10) does not
    x, E, F, _, eaf = _tent()
    x[3, 2] = z
    off, col = oracle.vert_adj_edges(len(x), E)
    pr, _ = oracle.ef_from_proximity(vf_rec(3, (0, 1, 2), 0.5, 0.25)[None], np.array([[3, 0]], np.uint32), x, E,
                                     F, off, col, eaf)
    assert len(pr) == (3 if hit else 0)


def _crossing():
    """Faces (0, 1, 2), (1, 0, 3) around edge 0 = (0, 1) in z = 0, and faces
    (4, 5, 6), (5, 4, 7) around edge 5 = (4, 5), which crosses edge 0 at
    (0.375, 0, 0): 0.75 x4 + 0.25 x5 on edge 5, 0.625 x0 + 0.375 x1 on edge 0."""
    x = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1, 0], [0.5, -1, 0], [0.375, 0, -1], [0.375, 0, 3],
                  [0.375, 1, 1], [0.375, -1, 1]], np.float32)
    F = np.array([[0, 1, 2], [1, 0, 3], [4, 5, 6], [5, 4, 7]], np.uint32)
    E = np.array([[0, 1], [1, 2], [2, 0], [0, 3], [3, 1], [4, 5], [5, 6], [6, 4], [4, 7], [7, 5]], np.uint32)
    fae, eaf = oracle.mesh_topology(E, F)
    return x, E, F, fae, eaf


def test_ee_zero_distance_closed_form():
    x, E, F, _, eaf = _crossing()
    off, col = oracle.vert_adj_edges(len(x), E)
    recs = ee_rec((4, 5), (0, 1), 0.75, 0.625)[None]  # e1 = edge 5, e2 = edge 0
    pr, pid = oracle.ef_from_proximity(recs, np.array([[5, 0]], np.uint32), x, E, F, off, col, eaf)
    # faces around e2 (0, 1): EF (edge 5, face) with e1's bary (0.75, 0.25); the face
    # corners 0 and 1 get e1's bary (0.75 at vertex 0, 0.25 at vertex 1), the third 0
    # faces around e1 (4, 5): EF (edge 0, face) with e2's bary 0.625; corners 4, 5 get e1's bary
    want_ids = [[5, 0], [5, 1], [0, 2], [0, 3]]
    want = [pair_words([4, 5, 0, 1, 2], (0.75, 0.75, 0.25)),
            pair_words([4, 5, 1, 0, 3], (0.75, 0.25, 0.75)),
            pair_words([0, 1, 4, 5, 6], (0.625, 0.75, 0.25)),
            pair_words([0, 1, 5, 4, 7], (0.625, 0.25, 0.75))]
    assert pid.tolist() == want_ids
    assert np.array_equal(pr, np.array(want))
    # edges sharing vertex 0 touch at distance 0, but every face holds a vertex of the other edge
    r = ee_rec((0, 1), (0, 3), 1.0, 1.0)[None]
    pr, _ = oracle.ef_from_proximity(r, np.array([[0, 3]], np.uint32), x, E, F, off, col, eaf)
    assert len(pr) == 0
    # a record without a VF / EE type bit is ignored
    r = ee_rec((4, 5), (0, 1), 0.75, 0.625)
    r[3] &= ~np.uint32(TYPE_BIT)
    pr, _ = oracle.ef_from_proximity(r[None], np.array([[5, 0]], np.uint32), x, E, F, off, col, eaf)
    assert len(pr) == 0


def test_wall_row_on_sheet():
    """The wall's middle row lies in the sheet's plane, inside sheet faces: one
    zero-distance VF pair per row vertex, each giving one EF pair per wall edge
    around the vertex (valence 6 inside the row, 4 at its two ends)."""
    m, first, nbx, nby = sheet_and_wall()
    st, recs, ids, nvf = proximity_records(m)
    row = first + np.arange(nbx) * nby + 2
    vf_verts = sorted(int(r[0]) for r in recs[:nvf])
    assert vf_verts == row.tolist()
    off, col = oracle.vert_adj_edges(m.nv, st.edges)
    pr, pid = oracle.ef_from_proximity(recs[:nvf], ids[:nvf], m.x, st.edges, st.faces, off, col,
                                       st.edge_adj_faces)
    assert len(pr) == 6 * (nbx - 2) + 2 * 4
    b = pr[:, 5:8].view(np.float32)
    assert set(b[:, 0].tolist()) == {0.0, 1.0}
    # no ordinary EF pair exists (every crossing is on a wall row edge or at a wall vertex), so
    # every pair is degenerate and the filter keeps them all
    dh, th = scenes.per_vertex(m)
    pairs, pids = oracle.ef_detect(m.x, st.edges, st.faces, dh, th)
    assert len(pairs) == 0
    pr, pid = oracle.ef_from_proximity(recs, ids, m.x, st.edges, st.faces, off, col, st.edge_adj_faces)
    oid, opr, adj, adjn, kept = oracle.ef_preprocess(pid, pr, st.edges, st.faces, st.face_adj_edges,
                                                     st.edge_adj_faces, filter_degenerate=True, with_kept=True)
    assert kept == len(pr)


def _two_sheet_pairs():
    m = scenes.two_sheets(0)
    dh, th = scenes.per_vertex(m)
    pairs, ids = oracle.ef_detect(m.x, m.edges, m.faces, dh, th)
    fae, eaf = oracle.mesh_topology(m.edges, m.faces)
    return m, pairs, ids, fae, eaf


def test_degenerate_filter_known_sets():
    m, pairs, ids, fae, eaf = _two_sheet_pairs()
    assert not any(oracle.ef_degenerate(p) for p in pairs)  # the DCD rejects boundary hits
    base = oracle.ef_preprocess(ids, pairs, m.edges, m.faces, fae, eaf)
    nc0, ci0, _ = oracle.contours_dfs(base[1], base[2], base[3])
    # a degenerate twin of a contour pair (same key, edge bary 0.99995): its neighbours are the
    # pair's neighbours, which are not degenerate, so it stays
    k = len(ids) // 2
    twin = pairs[k].copy()
    twin[5] = f32(0.99995)[0]
    # an isolated degenerate pair far from the crossing: an edge of sheet A cutting nothing
    used_e = set(ids[:, 0].tolist()) | set(ids[:, 1].tolist())
    far_e = next(e for e in range(m.ne) if e not in used_e and m.edges[e][0] < 1024 and m.edges[e][1] < 1024
                 and all(e2 not in used_e for e2 in range(max(0, e - 200), e + 200)))
    far_f = next(f for f in range(m.nf) if m.faces[f][0] >= 1024 and not set(m.faces[f].tolist())
                 & set(m.edges[far_e].tolist()))
    iso = pair_words(list(m.edges[far_e]) + list(m.faces[far_f]), (0.5, 0.99999, 0.000005))
    assert oracle.ef_degenerate(twin) and oracle.ef_degenerate(iso)
    P2 = np.concatenate([pairs, twin[None], iso[None]])
    I2 = np.concatenate([ids, ids[k][None], np.array([[far_e, far_f]], np.uint32)])
    oid, opr, adj, adjn, kept = oracle.ef_preprocess(I2, P2, m.edges, m.faces, fae, eaf, filter_degenerate=True,
                                                     with_kept=True)
    assert kept == len(pairs) + 1
    rows = {tuple(r) for r in np.concatenate([oid, opr], 1).tolist()}
    assert tuple(np.concatenate([ids[k], twin]).tolist()) in rows
    assert tuple(np.concatenate([[far_e, far_f], iso]).tolist()) not in rows
    # the twin joins its key twin's contour; the contour count is unchanged
    nc, ci, _ = oracle.contours_dfs(opr, adj, adjn)
    assert nc == nc0
    same = [i for i in range(kept) if tuple(oid[i]) == tuple(ids[k])]
    assert len(same) == 2 and ci[same[0]] == ci[same[1]]
    # the twin's neighbours list it back, next to the original pair
    nb = oracle.adj_rows(adj, adjn)
    for q in nb[same[0]]:
        assert same[0] in nb[q] and same[1] in nb[q]
    # without the filter nothing is dropped
    off_ = oracle.ef_preprocess(I2, P2, m.edges, m.faces, fae, eaf, with_kept=True)
    assert off_[4] == len(pairs) + 2
    # only degenerate pairs: all stay (intersection_resolver.cpp:874-877)
    only = oracle.ef_preprocess(I2[-2:], P2[-2:], m.edges, m.faces, fae, eaf, filter_degenerate=True,
                                with_kept=True)
    assert only[4] == 2


def test_duplicate_keys_canonical_order():
    m, pairs, ids, fae, eaf = _two_sheet_pairs()
    k = 7
    dup = [pairs[k].copy() for _ in range(3)]
    dup[0][5] = f32(0.25)[0]
    dup[1][5] = f32(0.5)[0]
    dup[2][6] = f32(0.125)[0]
    P2 = np.concatenate([pairs, np.array(dup)])
    I2 = np.concatenate([ids, np.repeat(ids[k][None], 3, 0)])
    ref = oracle.ef_preprocess(I2, P2, m.edges, m.faces, fae, eaf)
    rng = np.random.default_rng(3)
    for _ in range(3):
        perm = rng.permutation(len(I2))
        out = oracle.ef_preprocess(I2[perm], P2[perm], m.edges, m.faces, fae, eaf)
        for a, b in zip(ref, out):
            assert np.array_equal(a, b)
    # within the key, memcmp order of the 48 bytes: the first differing byte decides
    run = [i for i in range(len(ref[0])) if tuple(ref[0][i]) == tuple(ids[k])]
    assert len(run) == 4
    blobs = [ref[1][i].tobytes() for i in run]
    assert blobs == sorted(blobs)
