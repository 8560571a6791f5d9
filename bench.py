"""Benchmark of the per-step untangling hot path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config folded_cloth]

A step is one pass of the per-step collision pipeline over the folded
256x256-vertex cloth (130,050 triangles), inputs resident in HBM:
  face + edge LBVH rebuild (global box, Morton, radix sort, Karras),
  leaf update + refit of both trees,
  VF and EE broad-phase queries,
  EF intersection detection (edge-vs-face traversal fused with the EF test).
Prints one JSON line (see the "roofline" and "cpu_baseline" keys).
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="folded_cloth")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from untangling26_amd import scenes
    from untangling26_amd.lbvh import LBVH, TREE_EDGE, TREE_FACE
    from untangling26_amd.intersection import ef_detect_fused
    from untangling26_amd.runtime import Context

    ctx = Context(0)
    mesh = scenes.CONFIGS[args.config](0)
    d_hat_np, th_np = scenes.per_vertex(mesh)

    def dev(a):
        t = torch.from_numpy(np.ascontiguousarray(a))
        if t.dtype == torch.uint32:
            t = t.view(torch.int32)
        return t.cuda()

    X, F, E = dev(mesh.x), dev(mesh.faces), dev(mesh.edges)
    D, T = dev(d_hat_np), dev(th_np)
    face_tree = LBVH(ctx, TREE_FACE, mesh.nf)
    edge_tree = LBVH(ctx, TREE_EDGE, mesh.ne)
    counts = torch.zeros(8, dtype=torch.int32, device="cuda")  # vf, ee, ef broad, ef narrow
    caps = {"vf": 1 << 20, "ee": 1 << 21, "ef": 1 << 22, "efn": 1 << 18}
    bufs = {}

    def alloc():
        bufs["vf"] = torch.empty(2 * caps["vf"], dtype=torch.int32, device="cuda")
        bufs["ee"] = torch.empty(2 * caps["ee"], dtype=torch.int32, device="cuda")
        bufs["ef"] = torch.empty(2 * caps["ef"], dtype=torch.int32, device="cuda")
        bufs["efp"] = torch.empty(12 * caps["efn"], dtype=torch.int32, device="cuda")
        bufs["efi"] = torch.empty(2 * caps["efn"], dtype=torch.int32, device="cuda")

    alloc()

    def st_build():
        face_tree.reduce_face_tree_aabb(X, F)
        face_tree.construct_tree()
        edge_tree.reduce_edge_tree_aabb(X, E)
        edge_tree.construct_tree()

    def st_refit():
        face_tree.update_face_tree_leave_aabb(T, X, X, F)
        face_tree.refit()
        edge_tree.update_edge_tree_leave_aabb(T, X, X, E)
        edge_tree.refit()

    def st_query():
        counts[0:2].zero_()
        face_tree.broad_phase_query_from_verts(X, X, counts[0:1], bufs["vf"], D, T)
        edge_tree.broad_phase_query_from_edges(X, X, E, counts[1:2], bufs["ee"], D, T)

    def st_ef():
        counts[3:4].zero_()
        ef_detect_fused(ctx, face_tree, X, E, F, D, T, counts[3:4], bufs["efp"], bufs["efi"])

    def step():
        st_build()
        st_refit()
        st_query()
        st_ef()

    # size the lists once (the reference's resize-and-retry), then run without host syncs
    for _ in range(4):
        step()
        c = counts.cpu().tolist()
        need = {"vf": c[0], "ee": c[1], "efn": c[3]}
        if all(need[k] <= caps[k] for k in need):
            break
        for k in need:
            caps[k] = max(caps[k], 2 * need[k])
        alloc()
    c = counts.cpu().tolist()
    n_vf, n_ee, n_ef = c[0], c[1], c[3]
    # candidate count of the fused EF kernel (for its byte count): one broad query
    counts[2:3].zero_()
    face_tree.broad_phase_query_from_edges_unculled(X, X, E, counts[2:3], bufs["ef"], D, T)
    n_ef_broad = int(counts[2].item())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(args.steps):
        step()
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.steps

    # per-stage average durations (HIP events on the launch stream)
    def timed(fn, reps=20):
        fn()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    stage_ms = {"build": timed(st_build), "refit": timed(st_refit), "query": timed(st_query), "ef_dcd": timed(st_ef)}

    # Roofline of the fused EF kernel (broad traversal + narrow test): per
    # candidate pair 5 x 12 B vertex positions + 8 B edge + 12 B face topology
    # = 80 B, per edge query 8 B edge + 24 B positions + 8 B d_hat/thickness,
    # plus 56 B per emitted pair (48 B EfPair + 8 B ids); node visits (64 B)
    # hit in L2 and are not counted.
    ef_bytes = 80.0 * n_ef_broad + 40.0 * mesh.ne + 56.0 * n_ef
    ef_ms = stage_ms["ef_dcd"]
    achieved = ef_bytes / (ef_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mesh, d_hat_np, th_np)

    out = {
        "metric": "steps/s (per-step collision pipeline: LBVH rebuild+refit, VF/EE queries, fused EF detection)",
        "value": 1000.0 / ms,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded folded cloth)",
        "config": {"workload": f"{mesh.name}: {mesh.nf} tris, {mesh.ne} edges, {mesh.nv} verts",
                   "vf_pairs": n_vf, "ee_pairs": n_ee, "ef_broad_pairs": n_ef_broad, "ef_pairs": n_ef,
                   "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "k_ef_detect"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def cpu_baseline(mesh, d_hat, th):
    """Reference-algorithm CPU restatement of the same step (oracle, 1 thread)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle

    def one():
        ft = oracle.lbvh_build(1, mesh.x, mesh.faces)
        et = oracle.lbvh_build(2, mesh.x, mesh.edges)
        fl = oracle.leaf_boxes(1, mesh.x, mesh.x, mesh.faces, th)
        el = oracle.leaf_boxes(2, mesh.x, mesh.x, mesh.edges, th)
        qv = oracle.query_boxes(0, mesh.x, mesh.x, None, d_hat, th)
        qe = oracle.query_boxes(1, mesh.x, mesh.x, mesh.edges, d_hat, th)
        oracle.lbvh_query(ft, fl, qv)
        oracle.lbvh_query(et, el, qe, cull=True)
        ef = oracle.lbvh_query(ft, fl, qe)
        oracle.ef_dcd(ef, mesh.x, mesh.edges, mesh.faces, d_hat)

    t = time.perf_counter()
    n = 0
    while True:
        one()
        n += 1
        if time.perf_counter() - t > 10.0 or n >= 20:
            break
    dt = (time.perf_counter() - t) / n
    return {"value": 1.0 / dt, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} full steps of the same pipeline on the same mesh (C oracle, -O2, 1 thread)"}


if __name__ == "__main__":
    main()
