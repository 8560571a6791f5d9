"""The lcs_py mirror driven the way the reference's command-line repair drives
lcs_py (untangling26/cli.py:93-150, 168-184), on two crossing 32 x 32 sheets:
world data with auto-scaling off and the default cloth material, init_solver, the
solver configured after set-up (quadratic contacts, PRP, CCD, zero-distance EF
pairs, PCG budget 200), then the repair loop -- step, read the intersection
report, stop at zero EF pairs and zero contours -- and cleanup_device.

Also checked against the oracle: the get_intersection_contour_data dictionary of
the first step (the reference's keys, EF list, adjacency, contours, contour
types, positions), get_response_pair_data, and the .state / OBJ round trips."""
import numpy as np
import pytest
import torch

import oracle
from untangling26_amd import lcs_py, scenes, topology

pytestmark = pytest.mark.gpu


def _configure(config, max_contours=256, pcg_iters=200):
    config.use_untangling = True
    config.use_untangling_PRP = True
    config.use_untangling_ICM = False
    config.use_untangling_GIA = False
    config.use_self_collision = True
    config.use_floor = False
    config.gravity.x = config.gravity.y = config.gravity.z = 0.0
    config.use_quasi_static_mode = True
    config.use_static_mode = False
    config.fix_scene = False
    assert config.num_substep == 1
    config.nonlinear_iter_count = 1
    config.ignore_near_zero_dist_pairs = False
    config.contact_energy_type = 0
    config.use_ccd_linesearch = True
    config.use_energy_linesearch = False
    config.use_gpu_untangling = True
    config.PRP_solid_interior_adjacency = True
    config.PRP_direction_optimization_iterations = 0
    config.untangling_response_depth = 1.0
    config.untangling_process_contours_count = max_contours
    config.pcg_iter_count = pcg_iters
    config.print_collision_info = False
    config.print_pcg_info = False


def _solver(m):
    s = lcs_py.NewtonSolver()
    s.init_device(backend_name="hip")
    w = s.create_world_data_from_array("sheets", m.x, m.faces)
    w.set_auto_scaling(False)
    w.set_physics_material_cloth()
    s.register_world_data(w)
    s.init_solver()
    _configure(s.get_config())
    return s


def test_cli_style_repair_reaches_zero(ctx):
    m = scenes.two_sheets(0)
    s = _solver(m)
    try:
        records, done = [], False
        for it in range(21):
            verts, faces = s.get_sim_result()
            snapshot = [v.copy() for v in verts]
            s.physics_step_cpu()  # the CLI's default --physics cpu: the same GPU step here
            data = s.get_intersection_contour_data()
            ef = int(np.asarray(data["num_pairs"]).reshape(-1)[0])
            nc = int(np.asarray(data["num_contours"]).reshape(-1)[0])
            records.append((ef, nc))
            if ef == 0 and nc == 0:
                done = True
                break
        assert done and records[0][0] > 0 and len(records) <= 4, records
        assert snapshot[0].shape == (m.nv, 3) and np.isfinite(snapshot[0]).all()
        assert np.array_equal(faces[0], m.faces)
    finally:
        s.cleanup_device()
    with pytest.raises(RuntimeError):
        s.get_sim_result()


def test_contour_and_response_dicts_match_oracle(ctx):
    m = scenes.two_sheets(0)
    s = _solver(m)
    s.physics_step_gpu()
    data = s.get_intersection_contour_data()
    # the oracle on the state the step detected at (its iteration start = the registered positions)
    x0 = m.x
    d_hat, th = scenes.per_vertex(m)
    pairs, ids = oracle.ef_detect(x0, m.edges, m.faces, d_hat, th)
    fae, eaf = topology.mesh_topology(m.edges, m.faces)
    # the solver runs with ignore_near_zero_dist_pairs = False: the crossing sheets have no
    # zero-distance proximity pairs, so the degenerate filter decides alone
    oid, opr, adj, adjn = oracle.ef_preprocess(ids, pairs, m.edges, m.faces, fae, eaf, filter_degenerate=True)
    nc, ci, mi = oracle.contours_dfs(opr, adj, adjn)
    n = len(oid)
    assert int(data["num_pairs"][0]) == n and int(data["num_contours"][0]) == nc
    assert np.array_equal(data["narrow_phase_list_ef"], opr[:, :5])
    assert np.array_equal(data["ef_pair_contour_index"], ci)
    assert np.array_equal(data["ef_pair_mesh_index"], mi)
    off = np.concatenate([[0], np.cumsum(adjn.astype(np.int64))]).astype(np.uint32)
    assert np.array_equal(data["ef_pair_adj_pairs_ext_prefix"], off)
    assert np.array_equal(data["ef_pair_adj_pairs_ext_data"], adj)
    pre, dat = data["intersection_contours_prefix"], data["intersection_contours_data"]
    assert pre[0] == 0 and pre[-1] == n and len(pre) == nc + 1
    for c in range(nc):
        assert sorted(dat[pre[c]:pre[c + 1]].tolist()) == np.nonzero(ci == c)[0].tolist()
    ot, _ = oracle.classify_contours(opr, oracle.pair_is_boundary(oid, eaf), adjn, ci, mi, nc)
    assert np.array_equal(data["contour_types"], ot)
    # positions: the edge point at the iteration start, material coordinates (x, z)
    b = opr[:, 5:8].view(np.float32)
    e = opr[:, :5].astype(np.int64)
    want3 = x0[e[:, 0]] * b[:, :1] + x0[e[:, 1]] * (1.0 - b[:, :1])
    assert np.allclose(data["ef_pair_pos_3d"], want3, atol=1e-6)
    mx = x0[:, [0, 2]]
    uv_f = mx[e[:, 2]] * b[:, 1:2] + mx[e[:, 3]] * b[:, 2:3] + mx[e[:, 4]] * (1.0 - b[:, 1:2] - b[:, 2:3])
    assert np.allclose(data["ef_pair_pos_2d"][:, 2:], uv_f, atol=1e-6)
    assert data["prp_debug_uv_side0_faces"].shape == (0, 3)
    # responses: the PRP records the step appended, with their index quadruples
    r = s.get_response_pair_data()
    R = int(r["num_template_pairs"][0])
    assert R == s.get_step_report()["prp_responses"] > 0
    idx, w = r["target_point_template_pair_indices"], r["target_point_template_pair_weights"]
    assert idx.shape == (R, 4) and np.all(idx < m.nv)
    ct = r["target_point_template_pair_collision_type"]
    assert set(np.unique(ct).tolist()) <= {4, 8}
    vf = ct == 4
    # VF weights (1, b0, b1, 1 - b0 - b1) in magnitude: the face weights sum to one
    assert np.allclose(w[vf, 0], 1.0) and np.allclose(w[vf, 1:].sum(1), 1.0, atol=1e-5)
    meta = r["target_point_template_pairs_indices"]
    assert np.all(np.isin(meta[:, 3], [0, 1, 2])) and np.array_equal(meta[:, 3] == 2, ~vf)
    assert np.all(r["target_point_template_pair_stiff"][:, 1] > 0)
    s.cleanup_device()


def test_state_files_round_trip(ctx, tmp_path):
    m = scenes.two_sheets(0)
    s = _solver(m)
    s.physics_step_gpu()
    x1, v1 = s.positions().cpu().numpy(), s.velocities().cpu().numpy()
    p = tmp_path / "out" / "frame.state"
    s.save_current_state(str(p))
    s.restart_system()
    assert np.array_equal(s.positions().cpu().numpy(), m.x)
    s.load_target_state(str(p))
    # the file keeps 6 significant digits (the reference's stream default)
    assert np.allclose(s.positions().cpu().numpy(), x1, rtol=1e-5, atol=1e-6)
    assert np.allclose(s.velocities().cpu().numpy(), v1, rtol=1e-5, atol=1e-6)
    o = tmp_path / "mesh.obj"
    s.save_sim_result(str(o))
    v, f = lcs_py.read_obj(o)
    assert np.array_equal(f, m.faces) and np.allclose(v, x1, rtol=1e-5, atol=1e-6)
    s.restart_system()
    s.load_target_state_from_obj(str(o))
    assert np.allclose(s.positions().cpu().numpy(), x1, rtol=1e-5, atol=1e-6)
    s.cleanup_device()
