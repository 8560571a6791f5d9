"""CPU baseline of one untangling step, timed with the oracle (test and
measurement infrastructure only: the product path never imports this).

The oracle restates the reference's algorithms in plain single-threaded C;
this module times it on a bounded sample of the benchmark's workload and
scales each part to one step:
  detection   VF/EE proximity DCD (grid broad phase + narrow tests) and the EF
              intersection pass on the full mesh, timed once;
  contours    EF preprocessing + contour DFS on the full mesh;
  PRP         distances + directions + adaptive k-ring choice of the first
              contours until a time budget is spent, scaled to the number of
              contours a step processes (at most 256);
  assembly    the incremental potential with the detected contacts;
  PCG         FP64 iterations (with the contact off-diagonal terms) timed on a
              fixed count and scaled to the GPU's mean iteration count;
  CCD         the additive CCD over the swept broad-phase pairs of one step.
"""
from __future__ import annotations

import time

import numpy as np

import oracle


def _t(fn):
    t = time.perf_counter()
    out = fn()
    return time.perf_counter() - t, out


def cpu_step_baseline(mesh, rest, solver, args, prp_budget_s: float = 8.0, pcg_iters: int = 30):
    from untangling26_amd import scenes

    st = solver._setup
    x = np.ascontiguousarray(mesh.x, np.float32)
    edges, faces = st.edges, st.faces
    d_hat, off = scenes.per_vertex(mesh, d_hat=1e-3, thickness=0.0)
    times = {}
    vae = oracle.vert_adj_edges(mesh.nv, edges)

    def detect():
        qb = oracle.query_boxes(0, x, x, None, d_hat, off)
        lb = oracle.leaf_boxes(1, x, x, faces, off)
        vf = oracle.overlap_pairs(qb, lb)
        rv, iv = oracle.dcd_list(0, vf, x, edges, faces, d_hat, off, st.rest_vert_area, st.rest_edge_area,
                                 st.rest_face_area, 1e8, 0)
        qe = oracle.query_boxes(1, x, x, edges, d_hat, off)
        le = oracle.leaf_boxes(2, x, x, edges, off)
        ee = oracle.overlap_pairs(qe, le, cull=True)
        re, ie = oracle.dcd_list(1, ee, x, edges, faces, d_hat, off, st.rest_vert_area, st.rest_edge_area,
                                 st.rest_face_area, 1e8, 0)
        pairs, ids = oracle.ef_detect(x, edges, faces, d_hat, off)
        recs = np.concatenate([rv, re]).reshape(-1, 16)
        # zero-distance proximity pairs -> EF pairs (ignore_near_zero_dist_pairs = False)
        zp, zi = oracle.ef_from_proximity(recs, np.concatenate([iv, ie]).reshape(-1, 2), x, edges, faces, *vae,
                                          st.edge_adj_faces)
        return recs, np.concatenate([pairs, zp]), np.concatenate([ids, zi])

    times["detect"], (recs, pairs, ids) = _t(detect)

    def contours():
        oid, opr, adj, adjn = oracle.ef_preprocess(ids, pairs, edges, faces, st.face_adj_edges, st.edge_adj_faces,
                                                   filter_degenerate=True)
        nc, ci, mi = oracle.contours_dfs(opr, adj, adjn)
        return opr, adj, adjn, ci, mi, nc

    times["contours"], (opr, adj, adjn, ci, mi, nc) = _t(contours)

    # PRP: the first contours until the budget is spent
    o = oracle.PrpOracle(x, rest, edges, faces, st.face_adj_edges, st.edge_adj_faces, st.rest_vert_area,
                         st.rest_edge_area, st.rest_face_area, st.mass)
    active = min(nc, 256)
    done, t_prp = 0, 0.0
    while done < active and t_prp < prp_budget_s:
        n = min(active, max(1, 2 * done) if done else 1)
        tp, _ = _t(lambda: o.prepare(opr, adj, adjn, ci, mi, nc, args.dirs, n_prepare=n))
        tc = 0.0
        for c in range(done, n):
            dt, _ = _t(lambda: o.choose(c))
            tc += dt
        # prepare is recomputed for all first n contours each round: charge its share
        t_prp += tp * (n - done) / n + tc
        done = n
    times["prp"] = t_prp / max(done, 1) * active

    x_tilde = x.astype(np.float64).ravel()
    times["assemble"], (g, vals, _) = _t(lambda: oracle.assemble(st, x, x_tilde, 3600.0, recs))
    tp, _ = _t(lambda: oracle.pcg_sell(st.pattern, vals, -g, recs, max_iters=pcg_iters, rel_tol=1e-300))
    gpu_iters = max(1.0, getattr(solver, "_last_report", {}).get("pcg_iterations", pcg_iters) or pcg_iters)
    times["pcg"] = tp / pcg_iters * gpu_iters

    xe = solver.positions().cpu().numpy().astype(np.float32)

    def ccd():
        qb = oracle.query_boxes(0, x, xe, None, d_hat, off)
        lb = oracle.leaf_boxes(1, x, xe, faces, off)
        vf = oracle.overlap_pairs(qb, lb)
        qe = oracle.query_boxes(1, x, xe, edges, d_hat, off)
        le = oracle.leaf_boxes(2, x, xe, edges, off)
        ee = oracle.overlap_pairs(qe, le, cull=True)
        return min(oracle.ccd_list(2, vf, x, xe, edges, faces, off), oracle.ccd_list(3, ee, x, xe, edges, faces, off))

    times["ccd"], _ = _t(ccd)
    total = sum(times.values())
    return {
        "value": round(1.0 / total, 5),
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle restatement (single-threaded C) of one step on the same {mesh.nf}-triangle state: "
                   f"detection, contours ({nc} contours), assembly, CCD timed in full; PRP timed on the first "
                   f"{done} contours and scaled to {active}; PCG timed for {pcg_iters} iterations and scaled to "
                   f"{gpu_iters:.0f}"),
        "seconds": {k: round(v, 3) for k, v in times.items()},
    }
