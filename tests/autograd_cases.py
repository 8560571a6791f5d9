"""A scene for the incremental-potential tests: two parallel jittered sheets
0.6 mm apart (every vertex has proximity pairs inside d_hat = 1 mm, none closer
than ~0.2 mm, no intersections), with an inertia target a little off x."""
import numpy as np

import oracle
from untangling26_amd import initializer, scenes


def parallel_sheets(n=10, side=0.08, gap=6e-4, jitter=8e-5, seed=0):
    rng = np.random.default_rng(seed)
    a = scenes.grid_sheet(n, n, side, side)
    b = a.copy()
    b[:, 1] += gap
    a[:, 1] += rng.normal(0.0, jitter, len(a))
    b[:, 1] += rng.normal(0.0, jitter, len(b))
    x = np.concatenate([a, b], 0)
    faces = np.concatenate([scenes.grid_faces(n, n, 0), scenes.grid_faces(n, n, n * n)], 0)
    return scenes.Mesh(x, faces, name=f"parallel_sheets_{n}")


def ip_case(energy=1, seed=0):
    m = parallel_sheets(seed=seed)
    st = initializer.setup_cloth(m)
    rng = np.random.default_rng(seed + 1)
    x_tilde = (m.x.astype(np.float64) + rng.normal(0.0, 2e-4, m.x.shape)).ravel()
    x = m.x
    qv = oracle.query_boxes(0, x, x, None, st.d_hat, st.offset)
    vf, _ = oracle.dcd_list(0, oracle.overlap_pairs(qv, oracle.leaf_boxes(1, x, x, st.faces, st.offset)), x,
                            st.edges, st.faces, st.d_hat, st.offset, st.rest_vert_area, st.rest_edge_area,
                            st.rest_face_area, 1e8, energy)
    qe = oracle.query_boxes(1, x, x, st.edges, st.d_hat, st.offset)
    ee, _ = oracle.dcd_list(1, oracle.overlap_pairs(qe, oracle.leaf_boxes(2, x, x, st.edges, st.offset), cull=True),
                            x, st.edges, st.faces, st.d_hat, st.offset, st.rest_vert_area, st.rest_edge_area,
                            st.rest_face_area, 1e8, energy)
    recs = np.concatenate([vf, ee])
    return m, st, x_tilde, recs


def records_at(st, x, energy=1):
    """The oracle's VF + EE proximity records of setup st at FP32 positions x."""
    x = np.asarray(x, np.float32)
    qv = oracle.query_boxes(0, x, x, None, st.d_hat, st.offset)
    vf, _ = oracle.dcd_list(0, oracle.overlap_pairs(qv, oracle.leaf_boxes(1, x, x, st.faces, st.offset)), x,
                            st.edges, st.faces, st.d_hat, st.offset, st.rest_vert_area, st.rest_edge_area,
                            st.rest_face_area, 1e8, energy)
    qe = oracle.query_boxes(1, x, x, st.edges, st.d_hat, st.offset)
    ee, _ = oracle.dcd_list(1, oracle.overlap_pairs(qe, oracle.leaf_boxes(2, x, x, st.edges, st.offset), cull=True),
                            x, st.edges, st.faces, st.d_hat, st.offset, st.rest_vert_area, st.rest_edge_area,
                            st.rest_face_area, 1e8, energy)
    return np.concatenate([vf, ee])
