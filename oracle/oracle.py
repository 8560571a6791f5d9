"""ctypes access to liboracle.so -- the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline, never
by the product package.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None


def build() -> pathlib.Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        L = _lib
        p = C.c_void_p
        L.or_morton30.argtypes = [p, C.c_uint32, p]
        L.or_lbvh_build.argtypes = [C.c_int, p, p, C.c_uint32, p, p, p, p, p, p]
        L.or_karras.argtypes = [p, C.c_uint32, p, p, p]
        L.or_leaf_boxes.argtypes = [C.c_int, p, p, p, C.c_uint32, p, p]
        L.or_query_boxes.argtypes = [C.c_int, p, p, p, C.c_uint32, p, p, p]
        L.or_overlap_pairs.argtypes = [p, C.c_uint32, p, C.c_uint32, C.c_int, p, C.c_size_t]
        L.or_overlap_pairs.restype = C.c_size_t
        L.or_overlap_pairs_brute.argtypes = [p, C.c_uint32, p, C.c_uint32, C.c_int, p, C.c_size_t]
        L.or_overlap_pairs_brute.restype = C.c_size_t
        L.or_lbvh_node_boxes.argtypes = [C.c_uint32, p, p, p, p]
        L.or_lbvh_query.argtypes = [C.c_uint32, p, p, p, p, C.c_uint32, C.c_int, p, C.c_size_t]
        L.or_lbvh_query.restype = C.c_size_t
        L.or_mesh_topology.argtypes = [C.c_uint32, C.c_uint32, p, p, p, p]
        L.or_mesh_topology.restype = C.c_int
        L.or_ef_preprocess.argtypes = [C.c_uint32, p, p, p, p, p, p, C.c_int, p, p, C.POINTER(p), p, p]
        L.or_ef_preprocess.restype = C.c_long
        L.or_free.argtypes = [p]
        L.or_ef_degenerate.argtypes = [p]
        L.or_ef_degenerate.restype = C.c_int
        L.or_ef_from_proximity.argtypes = [C.c_size_t, p, p, p, p, p, p, p, p, p, p, C.c_size_t]
        L.or_ef_from_proximity.restype = C.c_size_t
        L.or_contours_dfs.argtypes = [C.c_uint32, p, p, p, p, p]
        L.or_contours_dfs.restype = C.c_uint32
        L.or_graph_bfs.argtypes = [C.c_uint32, p, p, p, p, p, p]
        L.or_graph_bfs.restype = C.c_int
        L.or_pcg.argtypes = [C.c_uint32, p, p, p, p, p, C.c_int, C.c_uint32, C.c_double, p]
        L.or_pcg.restype = C.c_uint32
        L.or_pt_coeff.argtypes = [p, p, p, p, p]
        L.or_ee_coeff.argtypes = [p, p, p, p, p]
        L.or_barrier.argtypes = [C.c_double, C.c_double, C.c_int]
        L.or_barrier.restype = C.c_double
        L.or_set_threads.argtypes = [C.c_int]
        L.or_prp_hit_components.argtypes = [p, p, C.c_uint32, C.c_int, p]
        L.or_friction_eval.argtypes = [C.c_double, p, p, C.c_double, p, p]
        L.or_friction_eval.restype = C.c_double
        L.or_record_weights.argtypes = [p, p]
        L.or_pair_block.argtypes = [p, p, p]
        L.or_friction_init.argtypes = [p, C.c_uint32, p]
        L.or_friction_lagged.argtypes = [p, C.c_uint32, p, p]
        L.or_contact_potential.argtypes = [p, C.c_uint32, p, p, p, p, C.c_double, C.c_int, C.c_int, p, p]
        L.or_accd.argtypes = [C.c_int, p, p, C.c_double]
        L.or_accd.restype = C.c_double
        L.or_toi_postprocess.argtypes = [C.c_double]
        L.or_toi_postprocess.restype = C.c_double
        L.or_dcd_list.argtypes = [C.c_int, p, C.c_size_t, p, p, p, p, p, p, p, p, p, C.c_double, C.c_int, p, p,
                                  C.c_size_t]
        L.or_dcd_list.restype = C.c_size_t
        L.or_ccd_list.argtypes = [C.c_int, p, C.c_size_t, p, p, p, p, p, p]
        L.or_ccd_list.restype = C.c_double
        L.or_bw98_face.argtypes = [p, p, C.c_double, C.c_double, C.c_double, p, p]
        L.or_bw98_face.restype = C.c_double
        L.or_bending_hinge.argtypes = [p, C.c_double, C.c_double, p, p, p]
        L.or_bending_hinge.restype = C.c_double
        L.or_assemble.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, p, p, p, C.c_double, p, p, p, p, p, p, p, p,
                                  p, p, C.c_uint32, p, C.c_uint32, p, p, p]
        L.or_pcg_contacts.argtypes = [C.c_uint32, p, p, p, p, C.c_uint32, p, p, C.c_int, C.c_uint32, C.c_double, p]
        L.or_pcg_contacts.restype = C.c_uint32
        L.or_prp_mesh_create.argtypes = [C.c_uint32] * 3 + [p] * 10
        L.or_prp_mesh_create.restype = C.c_void_p
        L.or_prp_mesh_destroy.argtypes = [C.c_void_p]
        L.or_prp_normals.argtypes = [C.c_void_p, p, p, p]
        L.or_prp_distances.argtypes = [C.c_void_p, C.c_uint32, p, p, p]
        L.or_prp_owns.argtypes = [C.c_float, C.c_float, C.c_float]
        L.or_prp_owns_dual.argtypes = [C.c_float] * 5 + [C.c_int, C.c_int]
        L.or_prp_directions.argtypes = [C.c_uint32, p, p, C.c_uint32, p, C.c_uint32, p, p, p, C.c_uint32, p,
                                        C.c_uint32, p]
        L.or_prp_directions.restype = C.c_uint32
        L.or_prp_prepare.argtypes = [C.c_void_p, C.c_uint32, p, p, p, p, p, C.c_uint32, C.c_uint32]
        L.or_prp_prepare_n.argtypes = [C.c_void_p, C.c_uint32, p, p, p, p, p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_prp_choose.argtypes = [C.c_void_p, C.c_uint32, p]
        L.or_prp_eval_slot.argtypes = [C.c_void_p, C.c_uint32, p, C.c_uint32, C.c_uint32, p, C.c_uint32, p]
        L.or_prp_eval_slot.restype = C.c_uint32
        L.or_prp_get_state.argtypes = [C.c_void_p, C.c_uint32, p, p, p, p]
        L.or_prp_set_dirs.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, p]
        L.or_prp_hop.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.or_prp_hop.restype = C.c_void_p
        L.or_prp_geo.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.or_prp_geo.restype = C.c_void_p
        L.or_prp_select.argtypes = [C.c_uint32, p, p, p, p, p, p, p]
        L.or_prp_drive.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, p, p, p, p, p, p, p]
        L.or_prp_drive.restype = C.c_int
        L.or_prp_suppress.argtypes = [C.c_uint32, p, p, C.c_uint32, p, p]
        L.or_prp_suppress.restype = C.c_uint32
        L.or_prp_penetration.argtypes = [C.c_uint32, p, p, p, p, p, p, p, p, p, p, p, p]
        L.or_classify_contours.argtypes = [C.c_uint32, p, p, p, p, p, C.c_uint32, p, p]
        L.or_prp_hit_attached.argtypes = [p, C.c_uint32, p]
        L.or_prp_hit_attached.restype = C.c_int
        L.or_prp_vertex_flags.argtypes = [p, C.c_uint32, p, C.c_uint32]
        L.or_prp_vertex_flags.restype = C.c_uint32
        L.or_prp_pick_component.argtypes = [C.c_uint32, p, p, p, p]
        L.or_prp_pick_component.restype = C.c_int
        L.or_prp_hits_adjacent.argtypes = [C.c_void_p, p, p]
        L.or_prp_resolve.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_float, p, p, C.c_uint32]
        L.or_prp_resolve.restype = C.c_uint32
        L.or_prp_resolve_meta.argtypes = [C.c_void_p, C.c_uint32, C.c_float, C.c_float, p, p, p, C.c_uint32]
        L.or_prp_resolve_meta.restype = C.c_uint32
        L.or_line_intersection.argtypes = [p, p, p, p, p, p, p]
        L.or_line_intersection.restype = C.c_int
        L.or_ef_dcd.argtypes = [p, C.c_size_t, p, p, p, p, p, p, p, C.c_size_t]
        L.or_ef_dcd.restype = C.c_size_t
        L.or_ef_detect.argtypes = [p, C.c_uint32, p, C.c_uint32, p, C.c_uint32, p, p, p, p, p, C.c_size_t]
        L.or_ef_detect.restype = C.c_size_t
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def morton30(points):
    points = _c(points, np.float32)
    out = np.empty(points.shape[0], np.uint32)
    lib().or_morton30(_p(points), points.shape[0], _p(out))
    return out


def lbvh_build(tree_type, x, prims):
    x = _c(x, np.float32)
    prims = _c(prims, np.uint32)
    n = prims.shape[0]
    nn = 2 * n - 1
    ms = np.empty(n, np.uint64)
    sgo = np.empty(n, np.uint32)
    par = np.empty(nn, np.uint32)
    ch = np.empty((nn, 2), np.uint32)
    obj = np.empty(nn, np.uint32)
    gb = np.empty(6, np.float32)
    lib().or_lbvh_build(tree_type, _p(x), _p(prims), n, _p(ms), _p(sgo), _p(par), _p(ch), _p(obj), _p(gb))
    return dict(morton_sorted=ms, sorted_get_original=sgo, parent=par, children=ch, object_idx=obj, global_box=gb)


def karras(keys):
    keys = _c(keys, np.uint64)
    n = keys.shape[0]
    nn = 2 * n - 1
    par = np.empty(nn, np.uint32)
    ch = np.empty((nn, 2), np.uint32)
    obj = np.empty(nn, np.uint32)
    lib().or_karras(_p(keys), n, _p(par), _p(ch), _p(obj))
    return dict(parent=par, children=ch, object_idx=obj)


def leaf_boxes(tree_type, xs, xe, prims, thickness):
    prims = _c(prims, np.uint32)
    out = np.empty((prims.shape[0], 6), np.float32)
    lib().or_leaf_boxes(tree_type, _p(_c(xs, np.float32)), _p(_c(xe, np.float32)), _p(prims), prims.shape[0],
                        _p(_c(thickness, np.float32)), _p(out))
    return out


def query_boxes(kind, xb, xe, edges, d_hat, thickness):
    xb = _c(xb, np.float32)
    n = xb.shape[0] if kind == 0 else edges.shape[0]
    out = np.empty((n, 6), np.float32)
    e = None if edges is None else _c(edges, np.uint32)
    lib().or_query_boxes(kind, _p(xb), _p(_c(xe, np.float32)), _p(e), n, _p(_c(d_hat, np.float32)),
                         _p(_c(thickness, np.float32)), _p(out))
    return out


def overlap_pairs(qbox, lbox, cull=False, brute=False):
    qbox = _c(qbox, np.float32)
    lbox = _c(lbox, np.float32)
    fn = lib().or_overlap_pairs_brute if brute else lib().or_overlap_pairs
    n = fn(_p(qbox), qbox.shape[0], _p(lbox), lbox.shape[0], int(cull), None, 0)
    out = np.empty((max(n, 1), 2), np.uint32)
    fn(_p(qbox), qbox.shape[0], _p(lbox), lbox.shape[0], int(cull), _p(out), n)
    out = out[:n]
    if brute:
        out = out[np.lexsort((out[:, 1], out[:, 0]))]
    return out


def line_intersection(s, e, a, b, c):
    bary = np.zeros(3, np.float32)
    t = np.zeros(1, np.float32)
    arr = [_c(v, np.float32) for v in (s, e, a, b, c)]
    hit = lib().or_line_intersection(*(_p(v) for v in arr), _p(bary), _p(t))
    return bool(hit), bary, float(t[0])


def ef_dcd(broad, x, edges, faces, d_hat, vprop=None):
    broad = _c(broad, np.uint32).reshape(-1, 2)
    nb = broad.shape[0]
    pairs = np.empty((max(nb, 1), 12), np.uint32)
    ids = np.empty((max(nb, 1), 2), np.uint32)
    n = lib().or_ef_dcd(_p(broad), nb, _p(_c(x, np.float32)), _p(_c(edges, np.uint32)), _p(_c(faces, np.uint32)),
                        _p(None if vprop is None else _c(vprop, np.uint32)), _p(_c(d_hat, np.float32)),
                        _p(pairs), _p(ids), nb)
    return pairs[:n], ids[:n]


def ef_detect(x, edges, faces, d_hat, thickness, vprop=None):
    """Sorted (eid, fid) of all EF intersections and the matching EfPair words."""
    x = _c(x, np.float32)
    edges = _c(edges, np.uint32)
    faces = _c(faces, np.uint32)
    args = (_p(x), x.shape[0], _p(edges), edges.shape[0], _p(faces), faces.shape[0],
            _p(None if vprop is None else _c(vprop, np.uint32)), _p(_c(d_hat, np.float32)),
            _p(_c(thickness, np.float32)))
    n = lib().or_ef_detect(*args, None, None, 0)
    pairs = np.empty((max(n, 1), 12), np.uint32)
    ids = np.empty((max(n, 1), 2), np.uint32)
    lib().or_ef_detect(*args, _p(pairs), _p(ids), n)
    return pairs[:n], ids[:n]


def lbvh_query(tree, leaf_box, qbox, cull=False):
    """CPU LBVH traversal over an or_lbvh_build tree; returns sorted pairs."""
    n = tree["sorted_get_original"].shape[0]
    nb = np.empty((2 * n - 1, 6), np.float32)
    ch = _c(tree["children"], np.uint32)
    obj = _c(tree["object_idx"], np.uint32)
    lib().or_lbvh_node_boxes(n, _p(ch), _p(_c(tree["sorted_get_original"], np.uint32)),
                             _p(_c(leaf_box, np.float32)), _p(nb))
    qbox = _c(qbox, np.float32)
    cnt = lib().or_lbvh_query(n, _p(ch), _p(obj), _p(nb), _p(qbox), qbox.shape[0], int(cull), None, 0)
    out = np.empty((max(cnt, 1), 2), np.uint32)
    lib().or_lbvh_query(n, _p(ch), _p(obj), _p(nb), _p(qbox), qbox.shape[0], int(cull), _p(out), cnt)
    out = out[:cnt]
    return out[np.lexsort((out[:, 1], out[:, 0]))]


def mesh_topology(edges, faces):
    edges = _c(edges, np.uint32)
    faces = _c(faces, np.uint32)
    fae = np.empty((faces.shape[0], 3), np.uint32)
    eaf = np.empty((edges.shape[0], 2), np.uint32)
    ok = lib().or_mesh_topology(edges.shape[0], faces.shape[0], _p(edges), _p(faces), _p(fae), _p(eaf))
    if not ok:
        raise ValueError("non-manifold or inconsistent mesh")
    return fae, eaf


def ef_preprocess(ids, pairs, edges, faces, fae, eaf, filter_degenerate=False, with_kept=False):
    """Canonical EF list with loop pairs appended, and the packed adjacency:
    neighbours of pair i are adj[sum(adjn[:i]) : sum(adjn[:i + 1])].
    filter_degenerate is the reference's not ignore_near_zero_dist_pairs."""
    ids = _c(ids, np.uint32).reshape(-1, 2)
    pairs = _c(pairs, np.uint32).reshape(-1, 12)
    P = ids.shape[0]
    cap = 3 * max(P, 1)
    oid = np.zeros((cap, 2), np.uint32)
    opr = np.zeros((cap, 12), np.uint32)
    adjn = np.zeros(cap, np.uint32)
    kept = C.c_uint32()
    ap = C.c_void_p()
    n = lib().or_ef_preprocess(P, _p(ids), _p(pairs), _p(_c(edges, np.uint32)), _p(_c(faces, np.uint32)),
                               _p(_c(fae, np.uint32)), _p(_c(eaf, np.uint32)), int(bool(filter_degenerate)), _p(oid),
                               _p(opr), C.byref(ap), _p(adjn), C.byref(kept))
    m = int(adjn[:n].sum(dtype=np.int64))
    adj = np.ctypeslib.as_array(C.cast(ap, C.POINTER(C.c_uint32)), shape=(max(m, 1),))[:m].copy()
    lib().or_free(ap)
    out = (oid[:n], opr[:n], adj, adjn[:n])
    return out + (int(kept.value),) if with_kept else out


def adj_rows(adj, adjn):
    """Packed adjacency as a list of rows."""
    off = np.concatenate([[0], np.cumsum(np.asarray(adjn, np.int64))])
    return [adj[off[i]:off[i + 1]] for i in range(len(adjn))]


def ef_degenerate(pair):
    return bool(lib().or_ef_degenerate(_p(_c(pair, np.uint32))))


def vert_adj_edges(nv, edges):
    """Vertex -> edge CSR (offsets [V+1], edges ascending per vertex)."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    v = e.ravel()
    eid = np.repeat(np.arange(len(e)), 2)
    order = np.lexsort((eid, v))
    off = np.zeros(nv + 1, np.uint32)
    off[1:] = np.cumsum(np.bincount(v, minlength=nv))
    return off, eid[order].astype(np.uint32)


def ef_from_proximity(recs, ids, x, edges, faces, vae_off, vae, eaf):
    """EF pairs from zero-distance proximity pairs: (pairs[n, 12], ids[n, 2]) in generation order."""
    recs = _c(recs, np.uint32).reshape(-1, 16)
    ids = _c(ids, np.uint32).reshape(-1, 2)
    n = recs.shape[0]
    cap = 64
    while True:
        pr = np.zeros((cap, 12), np.uint32)
        pid = np.zeros((cap, 2), np.uint32)
        k = lib().or_ef_from_proximity(n, _p(recs), _p(ids), _p(_c(x, np.float32)), _p(_c(edges, np.uint32)),
                                       _p(_c(faces, np.uint32)), _p(_c(vae_off, np.uint32)), _p(_c(vae, np.uint32)),
                                       _p(_c(eaf, np.uint32)), _p(pr), _p(pid), cap)
        if k <= cap:
            return pr[:k], pid[:k]
        cap = int(k)


def contours_dfs(pairs, adj, adjn):
    pairs = _c(pairs, np.uint32)
    N = pairs.shape[0]
    ci = np.zeros(max(N, 1), np.uint32)
    mi = np.zeros(max(N, 1), np.uint32)
    nc = lib().or_contours_dfs(N, _p(pairs), _p(_c(adj, np.uint32)), _p(_c(adjn, np.uint32)), _p(ci), _p(mi))
    return int(nc), ci[:N], mi[:N]


def classify_contours(pairs, is_boundary, adjn, contour_index, mesh_index, nc):
    """(ContourType per contour, mesh_is_closed [nc, 2]) of classify_contour_type."""
    pairs = _c(pairs, np.uint32).reshape(-1, 12)
    N = pairs.shape[0]
    t = np.zeros(max(nc, 1), np.uint32)
    cl = np.zeros((max(nc, 1), 2), np.uint32)
    lib().or_classify_contours(N, _p(pairs if N else np.zeros((1, 12), np.uint32)),
                               _p(_c(is_boundary, np.uint8) if N else np.zeros(1, np.uint8)),
                               _p(_c(adjn, np.uint32) if N else np.zeros(1, np.uint32)),
                               _p(_c(contour_index, np.uint32) if N else np.zeros(1, np.uint32)),
                               _p(_c(mesh_index, np.uint32) if N else np.zeros(1, np.uint32)), nc, _p(t), _p(cl))
    return t[:nc], cl[:nc]


def pair_is_boundary(ids, eaf):
    """1 where the pair's edge (ids[:, 0]) has at most one adjacent face."""
    e = np.asarray(ids, np.uint32).reshape(-1, 2)[:, 0]
    return (np.asarray(eaf, np.uint32).reshape(-1, 2)[e, 1] == 0xFFFFFFFF).astype(np.uint8)


def graph_bfs(row_ptr, col, w):
    row_ptr = _c(row_ptr, np.uint32)
    n = row_ptr.shape[0] - 1
    cm = np.zeros(max(n, 1), np.uint32)
    par = np.zeros(max(n, 1), np.uint32)
    inc = np.zeros(max(n, 1), np.uint32)
    any_inc = lib().or_graph_bfs(n, _p(row_ptr), _p(_c(col, np.uint32)), _p(_c(w, np.uint32)), _p(cm), _p(par),
                                 _p(inc))
    return cm[:n], par[:n], inc[:n], bool(any_inc)


def block_csr(n, rows, cols, blocks):
    """Block-CSR (row_ptr, col, blocks[k,9]) from block coordinates, duplicates summed."""
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    key = rows * n + cols
    uk, inv = np.unique(key, return_inverse=True)
    bl = np.zeros((len(uk), 9))
    np.add.at(bl, inv, np.asarray(blocks, np.float64).reshape(-1, 9))
    r = uk // n
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(np.bincount(r, minlength=n))
    return rp, (uk % n).astype(np.uint32), bl


def pcg(n, rp, ci, bl, b, precond=0, max_iters=1000, rel_tol=1e-6):
    """FP64 PCG restating host_solve_core; returns (x, iterations, rel_residual, converged, failure)."""
    b = _c(b, np.float64).ravel()
    x = np.zeros(3 * n)
    out = np.zeros(3)
    it = lib().or_pcg(n, _p(_c(rp, np.uint32)), _p(_c(ci, np.uint32)), _p(_c(bl, np.float64)), _p(b), _p(x),
                      precond, max_iters, rel_tol, _p(out))
    return x, int(it), float(out[0]), bool(out[1]), int(out[2])


def pt_coeff(p, t0, t1, t2):
    out = np.zeros(3)
    lib().or_pt_coeff(*(_p(_c(v, np.float64)) for v in (p, t0, t1, t2)), _p(out))
    return out


def ee_coeff(a0, a1, b0, b1):
    out = np.zeros(4)
    lib().or_ee_coeff(*(_p(_c(v, np.float64)) for v in (a0, a1, b0, b1)), _p(out))
    return out


def barrier(d, dhat, deriv=0):
    return lib().or_barrier(float(d), float(dhat), int(deriv))


def accd(is_ee, x0, dx, offset):
    return lib().or_accd(int(is_ee), _p(_c(x0, np.float64)), _p(_c(dx, np.float64)), float(offset))


def toi_postprocess(t):
    return lib().or_toi_postprocess(float(t))


def dcd_list(kind, broad, x, edges, faces, d_hat, offset, rva, rea, rfa, kappa=1e8, energy=1, vprop=None):
    """VF (kind 0) or EE (kind 1) DCD over broad pairs: (records uint32[n,16], ids uint32[n,2])."""
    broad = _c(broad, np.uint32).reshape(-1, 2)
    nb = broad.shape[0]
    recs = np.zeros((max(nb, 1), 16), np.uint32)
    ids = np.zeros((max(nb, 1), 2), np.uint32)
    n = lib().or_dcd_list(kind, _p(broad), nb, _p(_c(x, np.float32)), _p(_c(edges, np.uint32)),
                          _p(_c(faces, np.uint32)), _p(None if vprop is None else _c(vprop, np.uint32)),
                          _p(_c(d_hat, np.float32)), _p(_c(offset, np.float32)), _p(_c(rva, np.float32)),
                          _p(_c(rea, np.float32)), _p(_c(rfa, np.float32)), float(kappa), int(energy), _p(recs),
                          _p(ids), nb)
    return recs[:n], ids[:n]


def ccd_list(kind, broad, xs, xe, edges, faces, offset, vprop=None):
    """Global minimum TOI (FP32-rounded) of VF (kind 2) or EE (kind 3) pairs; 1.0 when none."""
    broad = _c(broad, np.uint32).reshape(-1, 2)
    return lib().or_ccd_list(kind, _p(broad), broad.shape[0], _p(_c(xs, np.float32)), _p(_c(xe, np.float32)),
                             _p(_c(edges, np.uint32)), _p(_c(faces, np.uint32)),
                             _p(None if vprop is None else _c(vprop, np.uint32)), _p(_c(offset, np.float32)))


def bw98_face(x3, dm, area, mu, lam):
    """(energy, grad[3,3], hess[3,3,3,3] as block (a,b) entry (i,j)) of one BW98 face."""
    x = _c(np.asarray(x3, np.float64).ravel(), np.float64)
    g = np.zeros(9)
    h = np.zeros(81)
    e = lib().or_bw98_face(_p(x), _p(_c(dm, np.float32)), float(area), float(mu), float(lam), _p(g), _p(h))
    return e, g.reshape(3, 3), h.reshape(3, 3, 3, 3)


def bending_hinge(x4, theta0, k):
    x = _c(np.asarray(x4, np.float64).ravel(), np.float64)
    g = np.zeros(12)
    h = np.zeros(144)
    th = np.zeros(1)
    e = lib().or_bending_hinge(_p(x), float(theta0), float(k), _p(g), _p(h), _p(th))
    return e, g.reshape(4, 3), h.reshape(4, 4, 3, 3), float(th[0])


def assemble(st, x, x_tilde, inv_h2, contacts=None):
    """(grad[3V], planar vals[9, nslots], energies[3]) of the setup st (initializer.ClothSetup)."""
    pat = st.pattern
    grad = np.zeros(3 * st.nv)
    vals = np.zeros(9 * pat.nslots)
    en = np.zeros(3)
    c = None if contacts is None or len(contacts) == 0 else _c(contacts, np.uint32)
    lib().or_assemble(st.nv, len(st.faces), len(st.hinges), _p(_c(x, np.float32)), _p(_c(x_tilde, np.float64)),
                      _p(_c(st.mass, np.float32)), float(inv_h2), _p(_c(st.faces, np.uint32)),
                      _p(_c(st.dm, np.float32)), _p(_c(st.rest_face_area, np.float32)),
                      _p(_c(st.mu_lambda, np.float32)), _p(_c(st.face_slots, np.uint32)),
                      _p(_c(st.hinges, np.uint32)), _p(_c(st.rest_angle, np.float32)),
                      _p(_c(st.bend_stiff_area, np.float32)), _p(_c(st.hinge_slots, np.uint32)), _p(c),
                      0 if c is None else len(c), _p(_c(pat.diag_slot, np.uint32)), pat.nslots, _p(grad), _p(vals),
                      _p(en))
    return grad, vals.reshape(9, -1), en


def pcg_sell(pattern, vals, b, contacts=None, precond=0, max_iters=100, rel_tol=1e-6):
    """FP64 PCG on planar SELL values (+ matrix-free contact off-diagonals)."""
    n = pattern.n
    rows = pattern.keys // n
    rp = np.zeros(n + 1, np.uint32)
    rp[1:] = np.cumsum(np.bincount(rows, minlength=n))
    ci = (pattern.keys % n).astype(np.uint32)
    bl = np.ascontiguousarray(np.asarray(vals)[:, pattern.key_slots].T)
    x = np.zeros(3 * n)
    out = np.zeros(3)
    c = None if contacts is None or len(contacts) == 0 else _c(contacts, np.uint32)
    it = lib().or_pcg_contacts(n, _p(rp), _p(ci), _p(bl), _p(c), 0 if c is None else len(c),
                               _p(_c(b, np.float64)), _p(x), precond, max_iters, rel_tol, _p(out))
    return x, int(it), float(out[0]), bool(out[1]), int(out[2])


PRP_HIT = np.dtype([("type", "u4"), ("src", "u4"), ("dst", "u4"), ("b0", "f4"), ("b1", "f4"), ("dist", "f4"),
                    ("valid", "u4"), ("pad", "u4")])
PRP_METRICS = np.dtype([("raw_hits", "u4"), ("hit_count", "u4"), ("covered", "u4"), ("ef_count", "u4"),
                        ("depth", "f4"), ("volume", "f4"), ("objective", "f4"), ("area", "f4"), ("disp_cost", "f4"),
                        ("max_hop_src", "u4"), ("max_hop_dst", "u4"), ("usable", "u4")])
PRP_CHOICE = np.dtype([("combo", "i4"), ("ks", "u4"), ("kd", "u4"), ("hit_count", "u4"), ("covered", "u4"),
                       ("ef_count", "u4"), ("levels", "u4"), ("suppressed", "u4"), ("dir", "f4", 3),
                       ("depth", "f4"), ("objective", "f4"), ("area", "f4"), ("disp_cost", "f4")])
HOP_INF = 0xFFFFFFFF


def prp_owns(own, opp, tau=0.25):
    return bool(lib().or_prp_owns(own, opp, tau))


def prp_owns_dual(rest_own, rest_opp, cur_own, cur_opp, tau=0.25, use_rest=True, use_cur=True):
    return bool(lib().or_prp_owns_dual(rest_own, rest_opp, cur_own, cur_opp, tau, int(use_rest), int(use_cur)))


def prp_directions(points, weights, b0, b1, x, vert_area, global_c=None, total=10):
    points = _c(points, np.float32).reshape(-1, 3)
    out = np.zeros((64, 3), np.float32)
    b0 = _c(b0, np.uint32)
    b1 = _c(b1, np.uint32)
    gc = None if global_c is None else _c(global_c, np.float64)
    n = lib().or_prp_directions(points.shape[0], _p(points), _p(_c(weights, np.float32)), len(b0), _p(b0), len(b1),
                                _p(b1), _p(_c(x, np.float32)), _p(_c(vert_area, np.float32)),
                                int(np.asarray(x).size // 3), _p(gc), total, _p(out))
    return out[:n]


def prp_select(objective, covered, ef_count, depth, cost, hits, eligible):
    n = len(objective)
    return lib().or_prp_select(n, _p(_c(objective, np.float32)), _p(_c(covered, np.uint32)),
                               _p(_c(ef_count, np.uint32)), _p(_c(depth, np.float32)), _p(_c(cost, np.float32)),
                               _p(_c(hits, np.uint32)), _p(_c(eligible, np.uint32)))


PRP_EVAL_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p)


def prp_drive(nd, ks0, kd0, diam_src, diam_dst, evaluate):
    """The oracle's adaptive k-ring driver with a Python evaluator:
    evaluate(combo, ks, kd) -> dict of PRP_METRICS fields (missing ones zero).
    Returns dict(selected, levels, no_response, metrics[nd], eligible, converged,
    pruned, evaluated, radii[nd, 2])."""
    def cb(_user, combo, ks, kd, out):
        m = np.zeros(1, PRP_METRICS)
        for k, v in evaluate(int(combo), int(ks), int(kd)).items():
            m[k] = v
        C.memmove(out, m.ctypes.data, PRP_METRICS.itemsize)

    fn = PRP_EVAL_FN(cb)
    met = np.zeros(max(nd, 1), PRP_METRICS)
    flags = np.zeros(max(nd, 1), np.uint32)
    radii = np.zeros((max(nd, 1), 2), np.uint32)
    levels = np.zeros(1, np.uint32)
    nr = np.zeros(1, np.uint32)
    sel = lib().or_prp_drive(nd, ks0, kd0, diam_src, diam_dst, C.cast(fn, C.c_void_p), None, _p(met), _p(flags),
                             _p(radii), _p(levels), _p(nr))
    f = flags[:nd]
    return dict(selected=int(sel), levels=int(levels[0]), no_response=bool(nr[0]), metrics=met[:nd],
                eligible=(f & 1) != 0, converged=(f & 2) != 0, pruned=(f & 4) != 0, evaluated=(f & 8) != 0,
                radii=radii[:nd])


def prp_suppress(adjacency, area, active):
    """Contours suppressed by the oracle's suppression pass: adjacency[c] = contours
    that contour c's kept hits reach; area[c]; active contour ids."""
    nc = len(adjacency)
    adjm = np.zeros((max(nc, 1), max(nc, 1)), np.uint8)
    for c, row in enumerate(adjacency):
        for o in row:
            adjm[c, o] = 1
    sup = np.zeros(max(nc, 1), np.uint8)
    act = _c(active, np.uint32)
    lib().or_prp_suppress(nc, _p(adjm), _p(_c(area, np.float32)), len(act), _p(act), _p(sup))
    return [int(c) for c in np.nonzero(sup[:nc])[0]]


def prp_penetration(hits, face_area, edge_area, vert_area, mass, face_normal, edge_normal, vert_normal, edges,
                    faces, direction):
    """The oracle's penetration summary (volume, depth, objective, area, displacement
    cost) of the valid hits (PRP_HIT records) with the given per-primitive data."""
    h = np.ascontiguousarray(np.asarray(hits, PRP_HIT))
    arr = [_c(a if len(a) else np.zeros(1), np.float32) for a in (face_area, edge_area, vert_area, mass)]
    nrm = [_c(np.asarray(a, np.float32).reshape(-1, 3) if len(a) else np.zeros((1, 3)), np.float32)
           for a in (face_normal, edge_normal, vert_normal)]
    top = [_c(a if len(a) else np.zeros((1, 3)), np.uint32) for a in (edges, faces)]
    out = np.zeros(1, PRP_METRICS)
    lib().or_prp_penetration(len(h), _p(h), *[_p(a) for a in arr], *[_p(a) for a in nrm], *[_p(a) for a in top],
                             _p(_c(direction, np.float32)), _p(out))
    return out[0]


def prp_pick_component(attached, cov, cnt, key):
    """The oracle's kept-component rule over per-component data; index or -1."""
    n = len(attached)
    a = _c(attached, np.uint8)
    return lib().or_prp_pick_component(n, _p(a if n else np.zeros(1, np.uint8)), _p(_c(cov, np.uint32)),
                                       _p(_c(cnt, np.uint32)), _p(_c(key, np.uint64)))


class PrpOracle:
    """CPU restatement of the PRP path on one mesh (oracle_prp.c)."""

    def __init__(self, x, x_rest, edges, faces, fae, eaf, vert_area, edge_area, face_area, mass):
        self._keep = [_c(a, dt) for a, dt in ((x, np.float32), (x_rest, np.float32), (edges, np.uint32),
                                              (faces, np.uint32), (fae, np.uint32), (eaf, np.uint32),
                                              (vert_area, np.float32), (edge_area, np.float32),
                                              (face_area, np.float32), (mass, np.float32))]
        k = self._keep
        self.nv, self.ne, self.nf = k[0].size // 3, k[2].size // 2, k[3].size // 3
        self.h = lib().or_prp_mesh_create(self.nv, self.ne, self.nf, *[_p(a) for a in k])

    def __del__(self):
        try:
            lib().or_prp_mesh_destroy(self.h)
        except Exception:
            pass

    def normals(self):
        fn = np.zeros((self.nf, 3), np.float32)
        en = np.zeros((self.ne, 3), np.float32)
        vn = np.zeros((self.nv, 3), np.float32)
        lib().or_prp_normals(self.h, _p(fn), _p(en), _p(vn))
        return fn, en, vn

    def distances(self, bverts):
        b = _c(bverts, np.uint32)
        hop = np.zeros(self.nv, np.uint32)
        geo = np.zeros(self.nv, np.float32)
        lib().or_prp_distances(self.h, len(b), _p(b), _p(hop), _p(geo))
        return hop, geo

    def prepare(self, pairs, adj, adjn, contour_index, mesh_index, nc, total_dirs=10, n_prepare=None):
        self._ct = [_c(pairs, np.uint32), _c(adj, np.uint32), _c(adjn, np.uint32), _c(contour_index, np.uint32),
                    _c(mesh_index, np.uint32)]
        self.nc = int(nc)
        N = self._ct[0].size // 12
        n_prepare = self.nc if n_prepare is None else int(n_prepare)
        lib().or_prp_prepare_n(self.h, N, *[_p(a) for a in self._ct], self.nc, total_dirs, n_prepare)

    def choose(self, c):
        ch = np.zeros(1, PRP_CHOICE)
        lib().or_prp_choose(self.h, c, _p(ch))
        return ch[0]

    def state(self, c):
        dirs = np.zeros((64, 3), np.float32)
        nd = np.zeros(1, np.uint32)
        diam = np.zeros(2, np.uint32)
        nb = np.zeros(2, np.uint32)
        lib().or_prp_get_state(self.h, c, _p(dirs), _p(nd), _p(diam), _p(nb))
        return dirs[:nd[0]].copy(), diam, nb

    def set_dirs(self, c, dirs):
        d = _c(dirs, np.float32).reshape(-1, 3)
        lib().or_prp_set_dirs(self.h, c, d.shape[0], _p(d))

    def hop(self, c, role):
        return np.ctypeslib.as_array(C.cast(lib().or_prp_hop(self.h, c, role), C.POINTER(C.c_uint32)),
                                     (self.nv,)).copy()

    def geo(self, c, role):
        return np.ctypeslib.as_array(C.cast(lib().or_prp_geo(self.h, c, role), C.POINTER(C.c_float)),
                                     (self.nv,)).copy()

    def eval_slot(self, c, direction, k_src, k_dst):
        d = _c(direction, np.float32)
        met = np.zeros(1, PRP_METRICS)
        n = lib().or_prp_eval_slot(self.h, c, _p(d), k_src, k_dst, None, 0, _p(met))
        hits = np.zeros(max(n, 1), PRP_HIT)
        lib().or_prp_eval_slot(self.h, c, _p(d), k_src, k_dst, _p(hits), n, _p(met))
        return hits[:n], met[0]

    def hit_components(self, hits, brute=False):
        """Component root (smallest member index) of every hit of a slot's hit list."""
        h = np.ascontiguousarray(np.asarray(hits, PRP_HIT))
        root = np.zeros(max(len(h), 1), np.uint32)
        lib().or_prp_hit_components(self.h, _p(h), len(h), int(bool(brute)), _p(root))
        return root[:len(h)]

    def hit_attached(self, c, hit):
        h = np.ascontiguousarray(np.asarray([hit], PRP_HIT))
        return bool(lib().or_prp_hit_attached(self.h, c, _p(h)))

    def vertex_flags(self, v):
        out = np.zeros(64, np.uint32)
        n = lib().or_prp_vertex_flags(self.h, v, _p(out), 64)
        return out[:min(n, 64)].tolist()

    def hits_adjacent(self, a, b):
        a = np.ascontiguousarray(np.asarray([a], PRP_HIT))
        b = np.ascontiguousarray(np.asarray([b], PRP_HIT))
        return bool(lib().or_prp_hits_adjacent(self.h, _p(a), _p(b)))

    def resolve(self, max_contours=256, response_depth=1.0, kappa=1e8):
        ch = np.zeros(max(self.nc, 1), PRP_CHOICE)
        n = lib().or_prp_resolve(self.h, max_contours, response_depth, kappa, _p(ch), None, 0)
        recs = np.zeros((max(n, 1), 16), np.uint32)
        lib().or_prp_resolve(self.h, max_contours, response_depth, kappa, _p(ch), _p(recs), n)
        return ch[:self.nc], recs[:n]

    def resolve_meta(self, max_contours=256, response_depth=1.0, kappa=1e8):
        """(choices, records uint32[n, 16], meta uint32[n, 2] = (contour, hit type 0 VF / 1 EE / 2 FV))."""
        ch = np.zeros(max(self.nc, 1), PRP_CHOICE)
        n = lib().or_prp_resolve_meta(self.h, max_contours, response_depth, kappa, _p(ch), None, None, 0)
        recs = np.zeros((max(n, 1), 16), np.uint32)
        meta = np.zeros((max(n, 1), 2), np.uint32)
        lib().or_prp_resolve_meta(self.h, max_contours, response_depth, kappa, _p(ch), _p(recs), _p(meta), n)
        return ch[:self.nc], recs[:n], meta[:n]


# ---------------------------------------------------------------- contact region filter
# IntersectionResolver::filter_invalid_pairs / filter_broad_phase_pairs
# (intersection_resolver.cpp:2233-2535, 2594-2682) with the region sets of
# host_apply_untangling_constraint (intersection_resolver2.cpp:2296-2426).

TYPE_BIT = np.uint32(0x80000000)
INDEX_MASK = np.uint32(0x7FFFFFFF)


def region_sets(nv, ef_pairs, contour_index, mesh_index, responses=None, meta=None):
    """Per-vertex sorted region lists.  EF pair p (edge verts idx[0:2], face verts
    idx[2:5]): edge verts get 2c + m, face verts 2c + m ^ 1 (:2354-2368).  PRP
    response i of contour c: VF -> vertex 2c, face verts 2c ^ 1; FV -> vertex 2c ^ 1,
    face verts 2c; EE -> source edge 2c, target edge 2c ^ 1 (:2339-2345).  Edges and
    faces take the union of their vertices' regions (:2386-2418)."""
    sets = [set() for _ in range(nv)]
    ef = np.asarray(ef_pairs, np.uint32).reshape(-1, 5)
    for p in range(len(ef)):
        er = 2 * int(contour_index[p]) + int(mesh_index[p])
        for k in range(5):
            sets[int(ef[p, k])].add(er if k < 2 else er ^ 1)
    if responses is not None and len(responses):
        R = np.asarray(responses, np.uint32).reshape(-1, 16)
        M = np.asarray(meta, np.uint32).reshape(-1, 2)
        for i in range(len(R)):
            r = 2 * int(M[i, 0])
            t = int(M[i, 1])
            info = {0: (r, r ^ 1, r ^ 1, r ^ 1), 2: (r ^ 1, r, r, r), 1: (r, r, r ^ 1, r ^ 1)}[t]
            for k in range(4):
                sets[int(R[i, k] & INDEX_MASK)].add(info[k])
    return [sorted(s) for s in sets]


def _regions_match(regions, a, b):
    ra = set()
    for v in a:
        ra.update(regions[v])
    for v in b:
        for q in regions[v]:
            if (q ^ 1) in ra:
                return True
    return False


def _record_sides(rec):
    idx = [int(rec[k] & INDEX_MASK) for k in range(4)]
    if rec[2] & TYPE_BIT:
        return idx, [idx[0]], idx[1:]
    return idx, idx[:2], idx[2:]


def filter_proximity(records, x, offset, regions):
    """(records with k1 = k2 = 0 for the disabled pairs, vert_is_invalid uint8[V]).
    A vertex is invalid when it is part of a proximity pair closer than the pair's
    thickness (fn_identify_invalid_verts_from_proximity_pairs, FP32 as the kernel);
    a pair is disabled when it touches an invalid vertex or its two sides lie in
    regions 2c + r and 2c + (r ^ 1) of one contour."""
    R = np.array(records, np.uint32).reshape(-1, 16)
    x = np.asarray(x, np.float32).reshape(-1, 3)
    off = np.asarray(offset, np.float32)
    invalid = np.zeros(len(x), np.uint8)
    for rec in R:
        idx, _, _ = _record_sides(rec)
        b0, b1 = rec[10:12].view(np.float32)
        one = np.float32(1)
        if rec[2] & TYPE_BIT:
            w = (one, -b0, -b1, b0 + b1 - one)
        else:
            w = (b0, one - b0, -b1, b1 - one)
        d = np.zeros(3, np.float32)
        for k in range(4):
            d = d + x[idx[k]] * np.float32(w[k])
        t = off[idx[0]] + off[idx[2]]
        if np.float32(np.dot(d, d)) < t * t + np.float32(1e-10):
            invalid[idx] = 1
    out = R.copy()
    for i, rec in enumerate(R):
        idx, a, b = _record_sides(rec)
        if invalid[idx].any() or _regions_match(regions, a, b):
            out[i, 8] = 0  # k1
            out[i, 9] = 0  # k2
    return out, invalid


def ccd_pair_mask(kind, broad, edges, faces, regions, invalid):
    """True for the swept broad-phase pairs CCD keeps (fn_filter_broadphase_{vf,ee}_pairs):
    kind 2 = (vertex, face), kind 3 = (edge, edge)."""
    B = np.asarray(broad, np.uint32).reshape(-1, 2)
    keep = np.ones(len(B), bool)
    for i, (q, l) in enumerate(B):
        if kind == 2:
            a, b = [int(q)], [int(v) for v in faces[l]]
        else:
            a, b = [int(v) for v in edges[q]], [int(v) for v in edges[l]]
        if invalid[a + b].any() or _regions_match(regions, a, b):
            keep[i] = False
    return keep


# ---------------------------------------------------------------- contact energy
def contact_energy(records, x, d_hat, offset, kappa=1e8, energy=1):
    """Sum over contact records of kappa * area * b(d - th, d_hat) (energy 1, barrier) or
    kappa * area * (d - th - d_hat)^2 / 2 (energy 0), with d the FP64 closest-point
    distance recomputed from x (float64 [V, 3]) and th, d_hat from the two vertices the
    DCD used (VF: vertex and first face vertex; EE: first vertex of each edge); records
    with k2 = 0 carry no energy (narrow_phase.cpp:1091-1112, cipc_kernel.hpp:15-36)."""
    R = np.asarray(records, np.uint32).reshape(-1, 16)
    X = np.asarray(x, np.float64).reshape(-1, 3)
    total = 0.0
    for rec in R:
        i = [int(v & INDEX_MASK) for v in rec[:4]]
        area, k2 = rec[7:8].view(np.float32)[0], rec[9:10].view(np.float32)[0]
        if k2 == 0.0:
            continue
        if rec[2] & TYPE_BIT:
            b = pt_coeff(X[i[0]], X[i[1]], X[i[2]], X[i[3]])
            diff = b[0] * (X[i[0]] - X[i[1]]) + b[1] * (X[i[0]] - X[i[2]]) + b[2] * (X[i[0]] - X[i[3]])
            ka, kb = i[0], i[1]
        else:
            c = ee_coeff(X[i[0]], X[i[1]], X[i[2]], X[i[3]])
            diff = (c[0] * X[i[0]] + c[1] * X[i[1]]) - (c[2] * X[i[2]] + c[3] * X[i[3]])
            ka, kb = i[0], i[2]
        d = float(np.sqrt(diff @ diff))
        dh = (float(d_hat[ka]) + float(d_hat[kb])) * 0.5
        th = float(offset[ka]) + float(offset[kb])
        if energy == 0:
            C_ = (d - th) - dh
            total += 0.5 * float(area) * kappa * C_ * C_
        else:
            total += float(area) * kappa * barrier(d - th, dh)
    return total


# ---- lagged IPC friction and the contact potential (oracle_friction.c) ----------
FRICTION_EPS = 1e-4  # Friction::ando_barrier::friction_eps


def friction_eval(mu_lambda, n, rel, E=FRICTION_EPS):
    """(energy, grad[3], hess[3, 3]) of the lagged friction of one pair with respect to its
    relative displacement rel (friction_kernel.hpp:137-301)."""
    g = np.zeros(3)
    h = np.zeros(9)
    e = lib().or_friction_eval(float(mu_lambda), _p(_c(n, np.float64)), _p(_c(rel, np.float64)), float(E), _p(g),
                               _p(h))
    return e, g, h.reshape(3, 3)


def pair_block(record):
    """(K[3, 3], g[3]): the record's block k2 n n^T + friction Hessian and its point gradient."""
    r = _c(np.asarray(record, np.uint32).reshape(16), np.uint32)
    K = np.zeros(9)
    g = np.zeros(3)
    lib().or_pair_block(_p(r), _p(K), _p(g))
    return K.reshape(3, 3), g


def record_weights(record):
    r = _c(np.asarray(record, np.uint32).reshape(16), np.uint32)
    w = np.zeros(4)
    lib().or_record_weights(_p(r), _p(w))
    return w


def friction_init(records, mu):
    """First-iteration friction values (mu_lambda = -mu k1, dv = 0) of a copy of the records."""
    R = np.array(np.asarray(records, np.uint32).reshape(-1, 16), np.uint32, order="C")
    if len(R):
        lib().or_friction_init(_p(R), len(R), _p(_c(mu, np.float32)))
    return R


def friction_lagged(records, x, x_start):
    """Later-iteration friction records (dv = sum w (x - x_start), repulsion off), a copy."""
    R = np.array(np.asarray(records, np.uint32).reshape(-1, 16), np.uint32, order="C")
    if len(R):
        lib().or_friction_lagged(_p(R), len(R), _p(_c(x, np.float64)), _p(_c(x_start, np.float64)))
    return R


def contact_potential(records, x, d_hat, offset, kappa=1e8, energy=1, terms=3, x_start=None):
    """(barrier energy, friction energy, gradient[V, 3]) at FP64 x."""
    R = _c(np.asarray(records, np.uint32).reshape(-1, 16), np.uint32)
    X = _c(x, np.float64)
    g = np.zeros(X.size)
    out = np.zeros(2)
    xs = None if x_start is None else _c(x_start, np.float64)
    lib().or_contact_potential(_p(R), len(R), _p(X), _p(xs), _p(_c(d_hat, np.float32)),
                               _p(_c(offset, np.float32)), float(kappa), int(energy), int(terms), _p(g), _p(out))
    return out[0], out[1], g.reshape(-1, 3)


def set_threads(n):
    """Threads of the oracle's parallel loops (1: serial, the default the tests use)."""
    lib().or_set_threads(int(n))


# ---- affine bodies (numpy restatement) ------------------------------------------------
# q holds 12 values per body, (p, a1, a2, a3); x_v = p + [a1 a2 a3] xbar_v
# (fn_apply_q_to_x_template, newton_solver.cpp:1882-1907).

def abd_q_to_x(body, xbar, q):
    Q = np.asarray(q, np.float64).reshape(-1, 4, 3)[np.asarray(body, np.int64)]
    xb = np.asarray(xbar, np.float64)
    return Q[:, 0] + xb[:, 0:1] * Q[:, 1] + xb[:, 1:2] * Q[:, 2] + xb[:, 2:3] * Q[:, 3]


def abd_bodies(q, q_tilde, mass4, volume, kappa, inv_h2, grad_scale=1.0, hess_scale=1.0):
    """AbdInertiaEnergy (abd_inertia_energy.cpp:25-140) and AbdOrthoEnergy
    (abd_ortho_energy.cpp:25-140) per body: (E_inertia, E_ortho, grad [nb, 12],
    blocks [nb, 12, 12])."""
    Q = np.asarray(q, np.float64).reshape(-1, 4, 3)
    nb = Q.shape[0]
    M4 = np.asarray(mass4, np.float64).reshape(nb, 4, 4)
    g = np.zeros((nb, 12))
    H = np.zeros((nb, 12, 12))
    e_in = e_or = 0.0
    I3 = np.eye(3)
    for b in range(nb):
        if q_tilde is not None:
            d = Q[b] - np.asarray(q_tilde, np.float64).reshape(-1, 4, 3)[b]
            G = inv_h2 * M4[b]
            e_in += grad_scale * 0.5 * float(np.sum(G * (d @ d.T)))
            g[b] += grad_scale * (G @ d).ravel()
            H[b] += hess_scale * np.kron(G, I3)
        A = Q[b, 1:]  # rows = the columns a_i of A
        s = kappa[b] * volume[b]
        D = A @ A.T
        e_or += s * float(np.sum((D - I3) ** 2))
        g[b, 3:] += 4.0 * s * ((D - I3) @ A).ravel()
        for i in range(3):
            for j in range(3):
                if i == j:
                    h = np.outer(A[i], A[i]) + (D[i, i] - 1.0) * I3 + A.T @ A
                else:
                    h = np.outer(A[j], A[i]) + D[i, j] * I3
                H[b, 3 + 3 * i:6 + 3 * i, 3 + 3 * j:6 + 3 * j] += 4.0 * s * h
    return e_in, e_or, g, H


def abd_record_sides(rec, body, xbar):
    """The two sides of a record (VF: vertex | face, EE: edge | edge) as (body, W) with
    W = sum over the side's vertices of w_k (1, xbar_k) (add_to_rigid_body,
    narrow_phase.cpp:2679-2698)."""
    w = record_weights(rec)
    na = 1 if rec[2] & TYPE_BIT else 2
    sides = []
    for ks in (range(0, na), range(na, 4)):
        W = np.zeros(4)
        b = None
        for k in ks:
            v = int(rec[k] & INDEX_MASK)
            b = int(body[v])
            W += w[k] * np.concatenate([[1.0], np.asarray(xbar[v], np.float64)])
        sides.append((b, W))
    return sides


def abd_contacts(records, body, xbar, nb):
    """Contact gradient [nb, 12] and own-body blocks [nb, 12, 12] of the records."""
    g = np.zeros((nb, 12))
    H = np.zeros((nb, 12, 12))
    for rec in np.asarray(records, np.uint32).reshape(-1, 16):
        (ba, Wa), (bb, Wb) = abd_record_sides(rec, body, xbar)
        K, gp = pair_block(rec)
        for b, W in ([(ba, Wa + Wb)] if ba == bb else [(ba, Wa), (bb, Wb)]):
            g[b] += np.kron(W, gp)
            H[b] += np.kron(np.outer(W, W), K)
    return g, H


def abd_apply(blocks, records, body, xbar, p):
    """y = A p: per-body blocks plus the records' couplings between two bodies."""
    nb = blocks.shape[0]
    P = np.asarray(p, np.float64).reshape(nb, 12)
    y = np.einsum("bij,bj->bi", blocks, P)
    for rec in np.asarray(records, np.uint32).reshape(-1, 16):
        (ba, Wa), (bb, Wb) = abd_record_sides(rec, body, xbar)
        if ba == bb:
            continue
        K, _ = pair_block(rec)
        ua = P[ba].reshape(4, 3).T @ Wa
        ub = P[bb].reshape(4, 3).T @ Wb
        y[ba] += np.kron(Wa, K @ ub)
        y[bb] += np.kron(Wb, K @ ua)
    return y.ravel()
