"""GPU stages on inputs with nothing to find, against the oracle: a flat sheet
(the unfolded 24 x 24 cloth) has no EF intersections (zero pairs, zero contours
with and without the degenerate filter), no proximity pair inside d_hat, no PRP
work (zero contours: no responses), and, lifted rigidly by 1 mm, no time of
impact (the running minimum stays at its reset value, 1, as the oracle's CCD
over the same broad phase says) -- also when the broad-phase list is empty."""
import numpy as np
import pytest
import torch

import oracle
from untangling26_amd import initializer, scenes, topology
from untangling26_amd.contours import ContourBuilder
from untangling26_amd.intersection import EfDetector
from untangling26_amd.lbvh import LBVH, TREE_EDGE, TREE_FACE
from untangling26_amd.narrow_phase import NarrowPhasesDetector
from untangling26_amd.prp import PrpResolver

pytestmark = pytest.mark.gpu


def dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if t.dtype == torch.uint32:
        t = t.view(torch.int32)
    return t.cuda()


def _flat(n=24):
    m = scenes.folded_cloth(0, n=n)
    return scenes.Mesh(x=scenes.folded_cloth_rest(n=n), faces=m.faces, edges=m.edges, name="flat")


def test_flat_sheet_has_no_ef_pairs_contours_or_responses(ctx):
    m = _flat()
    d_hat, th = scenes.per_vertex(m)
    pairs, ids = oracle.ef_detect(m.x, m.edges, m.faces, d_hat, th)
    assert len(ids) == 0
    fae, eaf = topology.mesh_topology(m.edges, m.faces)
    X, E, F = dev(m.x), dev(m.edges), dev(m.faces)
    det = EfDetector(ctx, m.nv, E, F)
    assert det.detect(X, dev(d_hat), dev(th)) == 0
    cb = ContourBuilder(ctx, m.nv, E, F, dev(fae), dev(eaf))
    for ignore in (True, False):
        assert tuple(cb.build(0, det.pairs, det.pair_ids, ignore_near_zero_dist_pairs=ignore)) == (0, 0)
    et = LBVH(ctx, TREE_EDGE, m.ne)
    et.reduce_edge_tree_aabb(X, E)
    et.construct_tree()
    et.update_edge_tree_leave_aabb(dev(th), X, X, E)
    et.refit()
    st = initializer.setup_cloth(m, x_rest=m.x)
    prp = PrpResolver.from_setup(ctx, st, m.x, max_contours=256, response_depth=1.0, kappa=1e8)
    choices, recs, stats = prp.resolve(X, cb, det.face_tree, et)
    assert choices == [] and recs.shape[0] == 0 and stats["responses"] == 0 and stats["contours"] == 0


def test_flat_sheet_has_no_proximity_pairs_or_impacts(ctx):
    m = _flat()
    d_hat, off = scenes.per_vertex(m, d_hat=1e-3, thickness=5e-4)
    _, eaf = topology.mesh_topology(m.edges, m.faces)
    va, ea, fa = topology.rest_areas(m.x, m.edges, m.faces, eaf)
    xe = (m.x + np.array([0.0, 0.0, 1e-3], np.float32)).astype(np.float32)
    X0, X1, E, F, O = dev(m.x), dev(xe), dev(m.edges), dev(m.faces), dev(off)
    npd = NarrowPhasesDetector(ctx, m.nv, E, F, dev(d_hat), O, dev(va), dev(ea), dev(fa))
    ft, et = LBVH(ctx, TREE_FACE, m.nf), LBVH(ctx, TREE_EDGE, m.ne)
    # proximity at x: nothing inside d_hat (the grid spacing is 87 mm)
    for tree, prims, red, upd in ((ft, F, "reduce_face_tree_aabb", "update_face_tree_leave_aabb"),
                                  (et, E, "reduce_edge_tree_aabb", "update_edge_tree_leave_aabb")):
        getattr(tree, red)(X0, prims)
        tree.construct_tree()
        getattr(tree, upd)(O, X0, X0, prims)
        tree.refit()
    cap = 1024
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    recs = torch.zeros(16 * cap, dtype=torch.int32, device="cuda")
    rids = torch.zeros(2 * cap, dtype=torch.int32, device="cuda")
    npd.vf_dcd_query_repulsion(X0, cnt, recs, rids, ft)
    npd.ee_dcd_query_repulsion(X0, cnt, recs, rids, et)
    assert int(cnt.item()) == 0
    # swept from x to x + 1 mm z: broad pairs between neighbours, no impact
    for tree, prims, upd in ((ft, F, "update_face_tree_leave_aabb"), (et, E, "update_edge_tree_leave_aabb")):
        getattr(tree, upd)(O, X0, X1, prims)
        tree.refit()
    toi = torch.zeros(1, dtype=torch.float32, device="cuda")
    npd.reset_toi(toi)
    reset = float(toi.item())
    assert reset == 1.0
    for kind, tree in ((2, ft), (3, et)):
        qb = oracle.query_boxes(0 if kind == 2 else 1, m.x, xe, None if kind == 2 else m.edges, d_hat, off)
        lb = oracle.leaf_boxes(1 if kind == 2 else 2, m.x, xe, m.faces if kind == 2 else m.edges, off)
        broad = oracle.overlap_pairs(qb, lb, cull=(kind == 3))
        assert len(broad) > 0
        assert oracle.ccd_list(kind, broad, m.x, xe, m.edges, m.faces, off) == reset
        query = npd.vf_ccd_query if kind == 2 else npd.ee_ccd_query
        query(X0, X1, toi, tree)
        assert float(toi.item()) == reset, kind
        # an empty broad-phase list
        empty = torch.zeros(2, dtype=torch.int32, device="cuda")
        zero = torch.zeros(1, dtype=torch.int32, device="cuda")
        query(X0, X1, toi, None, empty, zero)
        assert float(toi.item()) == reset, kind
