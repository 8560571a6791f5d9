"""Affine bodies through the lcs_py mirror (one rigid world = one 12-DOF body,
ut_solver_* in q-space): the untangling configuration of the CLI separates two
intersecting boxes and a sphere; every step's EF pairs and contours at the step
start equal the oracle's on the vertex positions x(q) of the oracle's q -> x; and
the 512-body workload (alternating level-3 icospheres and 8x8x8-subdivided
boxes, radii U(0.05, 0.2) m, <= 30% overlap, orthogonality stiffness 1e5) steps
with finite bodies and a falling intersection count."""
import numpy as np
import pytest

import abd_cases
import oracle
from untangling26_amd import lcs_py, scenes

pytestmark = pytest.mark.gpu


def _solver(sc, stiffness=1e5, dirs=16):
    s = lcs_py.NewtonSolver()
    s.init_device("hip")
    for b in range(len(sc.vert_off) - 1):
        lo, hi = int(sc.vert_off[b]), int(sc.vert_off[b + 1])
        f = sc.mesh.faces[(sc.mesh.faces[:, 0] >= lo) & (sc.mesh.faces[:, 0] < hi)] - lo
        w = s.create_world_data_from_array(f"body{b}", sc.mesh.x[lo:hi].astype(np.float64) - sc.q0[b, 0], f)
        w.set_physics_material_rigid(stiffness=stiffness)
        w.set_translation(*sc.q0[b, 0])
        s.register_world_data(w)
    s.init_solver()
    c = s.get_config()
    # untangling26.cli.configure_solver (cli.py:93-118)
    c.use_untangling = True
    c.use_untangling_PRP = True
    c.use_quasi_static_mode = True
    c.gravity = (0.0, 0.0, 0.0)
    c.contact_energy_type = 0
    c.untangling_response_depth = 1.0
    c.nonlinear_iter_count = 1
    c.use_ccd_linesearch = True
    c.ignore_near_zero_dist_pairs = False
    c.pcg_iter_count = 1000
    c.pcg_rel_tol = 1e-6
    c.prp_directions = dirs
    return s


def _oracle_counts(s, x):
    """EF pairs (intersections + pairs from zero-distance proximity records, as the
    solver detects them with ignore_near_zero_dist_pairs = 0) and contours at x."""
    st = s._setup
    x = np.asarray(x, np.float32)
    pairs, ids = oracle.ef_detect(x, st.edges, st.faces, st.d_hat, st.offset, vprop=st.vprop)
    recs, rids = [], []
    for kind, qb, lb, cull in ((0, oracle.query_boxes(0, x, x, None, st.d_hat, st.offset),
                                oracle.leaf_boxes(1, x, x, st.faces, st.offset), False),
                               (1, oracle.query_boxes(1, x, x, st.edges, st.d_hat, st.offset),
                                oracle.leaf_boxes(2, x, x, st.edges, st.offset), True)):
        r, i = oracle.dcd_list(kind, oracle.overlap_pairs(qb, lb, cull=cull), x, st.edges, st.faces, st.d_hat,
                               st.offset, st.rest_vert_area, st.rest_edge_area, st.rest_face_area, 1e8, 0,
                               vprop=st.vprop)
        recs.append(r)
        rids.append(i)
    vae = oracle.vert_adj_edges(st.nv, st.edges)
    zp, zi = oracle.ef_from_proximity(np.concatenate(recs).reshape(-1, 16), np.concatenate(rids).reshape(-1, 2), x,
                                      st.edges, st.faces, *vae, st.edge_adj_faces)
    pairs, ids = np.concatenate([pairs, zp]), np.concatenate([ids, zi])
    _, opr, adj, adjn = oracle.ef_preprocess(ids, pairs, st.edges, st.faces, st.face_adj_edges, st.edge_adj_faces,
                                             filter_degenerate=True)
    nc, _, _ = oracle.contours_dfs(opr, adj, adjn)
    return len(ids), nc


def test_intersecting_bodies_separate(ctx):
    sc = abd_cases.hand_scene([(1, 0.1, (0.0, 0.0, 0.0)), (1, 0.1, (0.17, 0.03, 0.01)),
                               (0, 0.08, (-0.05, 0.15, 0.0))], seed=7)
    s = _solver(sc)
    history = []
    for _ in range(8):
        x = s.positions().cpu().numpy()
        q = s.q().cpu().numpy()
        # the vertices follow q
        assert np.abs(x - oracle.abd_q_to_x(s._setup.body, s._setup.xbar, q)).max() <= 1e-6
        r = s.physics_step_gpu()
        assert r["ef_pairs"] == _oracle_counts(s, x)[0]
        history.append((r["ef_total"], r["contours"]))
        if r["ef_total"] == 0 and r["contours"] == 0:
            break
    assert history[0][0] > 50 and history[-1] == (0, 0) and len(history) <= 5, history
    q = s.q().cpu().numpy().reshape(-1, 4, 3)
    A = q[:, 1:]
    assert np.abs(np.einsum("bik,bjk->bij", A, A) - np.eye(3)).max() < 1e-2
    s.restart_system()
    assert np.allclose(s.q().cpu().numpy().ravel(), s._setup.q0.ravel())


def test_512_body_workload_steps(ctx):
    sc = scenes.rigid_bodies(0)
    assert sc.mesh.nv == 263168 and sc.mesh.nf == 524288
    s = _solver(sc)
    x0 = s.positions().cpu().numpy()
    ef0, nc0 = _oracle_counts(s, x0)
    ef = []
    for i in range(4):
        r = s.physics_step_gpu()
        if i == 0:
            assert (r["ef_pairs"], r["contours"]) == (ef0, nc0)
        ef.append(r["ef_total"])
        assert r["prp_responses"] > 0 and np.isfinite(r["pcg_relative_residual"])
    q = s.q().cpu().numpy()
    assert np.isfinite(q).all()
    assert ef[-1] < ef[0], ef
