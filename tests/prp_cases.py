"""Scenes for the PRP tests: a mesh, its cloth setup, the oracle's EF contours
and a prepared PRP oracle."""
import numpy as np

import oracle
from untangling26_amd import initializer, scenes


def prp_fixture(name="two_sheets", total_dirs=10):
    if name == "two_sheets":
        m = scenes.two_sheets(0)
        rest = m.x
    elif name == "two_sheets_s3":
        m = scenes.two_sheets(3)
        rest = m.x
    elif name in ("deep_sheets_8", "deep_sheets_16"):
        # deeply crossing small sheets: the chosen k-rings reach the contour's hop
        # diameter (16 x 16: after growing past the first distance threshold;
        # 8 x 8: within it)
        m = scenes.two_sheets(0, n=int(name.rsplit("_", 1)[1]), depth=0.2)
        rest = m.x
    elif name == "folded_small":
        m = scenes.folded_cloth(0, n=48)
        rest = scenes.folded_cloth_rest(n=48)
    else:
        raise KeyError(name)
    st = initializer.setup_cloth(m, x_rest=rest)
    d, th = scenes.per_vertex(m)
    pairs, ids = oracle.ef_detect(m.x, m.edges, m.faces, d, th)
    oid, opr, adj, adjn = oracle.ef_preprocess(ids, pairs, m.edges, m.faces, st.face_adj_edges, st.edge_adj_faces)
    nc, ci, mi = oracle.contours_dfs(opr, adj, adjn)
    o = oracle.PrpOracle(m.x, rest, st.edges, st.faces, st.face_adj_edges, st.edge_adj_faces, st.rest_vert_area,
                         st.rest_edge_area, st.rest_face_area, st.mass)
    o.prepare(opr, adj, adjn, ci, mi, nc, total_dirs)
    return dict(mesh=m, setup=st, rest=np.asarray(rest, np.float32), ids=oid, pairs=opr, adj=adj, adjn=adjn,
                contour_index=ci, mesh_index=mi, nc=nc, oracle=o)
