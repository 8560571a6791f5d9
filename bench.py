"""Benchmark of the per-step untangling path on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--dirs 64] [--no-cpu]

Workload: the folded self-intersecting cloth, 256 x 256 vertices (130,050
triangles, 2 m side, figure-eight crossing, seed 0), repaired with the
untangling configuration of untangling26.cli.configure_solver (quasi-static,
no gravity, quadratic contacts, PRP responses, CCD line search, one Newton
iteration per step).  A step is one physics_step_gpu call: LBVH rebuild,
proximity + EF detection, contours, PRP (per-contour ray casting of candidate
directions, clustering, J(r), responses), contact filtering, assembly, PCG to
relative residual 1e-6, CCD line search.  Everything is resident in HBM.

The K timed steps start from the same tangled state for both PCG precisions:
mixed (FP32 SpMV / vectors with FP64 refinement, the reported value) and plain
FP64; the FP64 rate and the speedup go into `config`.
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256, help="cloth resolution (vertices per side)")
    ap.add_argument("--dirs", type=int, default=64, help="candidate directions per contour")
    ap.add_argument("--prp-batch", type=int, default=128,
                    help="contours evaluated together by the PRP (PRP_cpu_contour_batch_size, the reference's default 32); "
                         "a throughput setting: the step's results do not depend on it")
    ap.add_argument("--pcg-iters", type=int, default=1000)
    ap.add_argument("--inner-tol", type=float, default=1e-2,
                    help="relative tolerance of each FP32 inner solve of the mixed-precision PCG")
    ap.add_argument("--max-repair", type=int, default=200, help="step cap of the repair episode (cli.py --max-iters)")
    ap.add_argument("--probe-every", type=int, default=3,
                    help="re-solve every n-th repair step's system at both precisions (0: never)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison leg")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run one plain device copy of this many MiB before the legs (counter calibration)")
    return ap.parse_args()


def make_solver(args, precision):
    import numpy as np

    from untangling26_amd import scenes
    from untangling26_amd.lcs_py import NewtonSolver

    m = scenes.folded_cloth(0, n=args.n)
    rest = scenes.folded_cloth_rest(n=args.n)
    s = NewtonSolver()
    s.init_device("hip")
    s.register_world_data(s.create_world_data_from_array("folded_cloth", m.x, m.faces, rest))
    cfg = s.get_config()
    # untangling26/cli.py:93-118
    cfg.use_untangling = True
    cfg.use_quasi_static_mode = True
    cfg.gravity = (0.0, 0.0, 0.0)
    cfg.contact_energy_type = 0
    cfg.use_untangling_PRP = True
    cfg.untangling_response_depth = 1.0
    cfg.nonlinear_iter_count = 1
    cfg.use_ccd_linesearch = True
    cfg.ignore_near_zero_dist_pairs = False
    cfg.pcg_iter_count = args.pcg_iters
    cfg.pcg_rel_tol = 1e-6
    cfg.pcg_inner_tol = args.inner_tol
    cfg.pcg_precision = precision
    cfg.prp_directions = args.dirs
    cfg.PRP_cpu_contour_batch_size = getattr(args, "prp_batch", 32)
    s.init_solver()
    return s, m, np.asarray(rest, np.float32)


def run_leg(args, precision, probe_every=0):
    """Warm-up, the K timed repair steps, then one whole repair episode.

    Timed region: exactly K physics_step_gpu calls from the tangled state.  A step
    whose report shows no EF pair and no contour at its start has found the fold
    intersection-free (the CLI's stopping test, cli.py:134-148): the repair is
    complete, so the state is reset to the tangled one and the next step starts a
    new repair.  Every timed step is therefore a repair step.

    Episode (after the timed region): from the tangled state, step until a step
    reports zero EF pairs and zero contours (at most --max-repair steps), each step
    timed on the host around its synchronising report read-back.  With
    probe_every > 0, every probe_every-th step's system is solved again at FP64 and
    mixed precision (ut_solver_pcg_resolve; outside the step times)."""
    import torch

    from untangling26_amd.lcs_py import PCG_FP64, PCG_MIXED

    s, m, rest = make_solver(args, precision)
    x0 = s.positions().clone()
    v0 = torch.zeros_like(x0)
    name = precision_name(precision)
    warm_ee = 0
    for i in range(args.warmup):
        r = s.physics_step_gpu()
        warm_ee += int(r["prp_ee_launches"])
        log(f"{name} warmup step {i}: {r['ms_total']:.1f} ms, {r['ef_total']} EF pairs, "
            f"{r['contours']} contours, {r['prp_responses']} PRP responses, {r['pcg_iterations']} PCG iterations")
    s.set_state(x0, v0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps, restarts = [], 0
    for i in range(args.steps):
        reps.append(s.physics_step_gpu())
        r = reps[-1]
        # the report read-back synchronizes the step; one line per step keeps a long run visibly alive
        log(f"{name} step {i}: {r['ms_total']:.1f} ms, {r['ef_total']} EF pairs, {r['contours']} contours, "
            f"PCG {r['pcg_iterations']} it, residual {r['pcg_relative_residual']:.2e}")
        if r["ef_total"] == 0 and r["contours"] == 0:
            s.set_state(x0, v0)
            restarts += 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0

    s.set_state(x0, v0)
    torch.cuda.synchronize()
    episode, probes = [], []
    for i in range(args.max_repair):
        ts = time.perf_counter()
        r = s.physics_step_gpu()
        sec = time.perf_counter() - ts
        episode.append((r, sec))
        done = r["ef_total"] == 0 and r["contours"] == 0
        log(f"{name} repair {i}: {1e3 * sec:.1f} ms, {r['ef_total']} EF pairs, {r['contours']} contours, "
            f"PCG {r['pcg_iterations']} it, residual {r['pcg_relative_residual']:.2e}")
        if probe_every and not done and i % probe_every == 0:
            p64, pmx = s.pcg_resolve(PCG_FP64), s.pcg_resolve(PCG_MIXED)
            probes.append({"step": i, "fp64": p64, "mixed": pmx})
            log(f"  PCG probe: fp64 {p64['ms']:.2f} ms {p64['iterations']} it {p64['relative_residual']:.2e} | "
                f"mixed {pmx['ms']:.2f} ms {pmx['iterations']} it {pmx['relative_residual']:.2e}")
        if done:
            break
    # k_prp_ray_edge dispatches of this leg: the warm-up ones, then the timed ones; the
    # timed launches are dispatches [start, end) of the leg's list (a counter pass
    # picks them out of its dispatch list by this window)
    timed_ee = int(sum(r["prp_ee_launches"] for r in reps))
    window = [warm_ee, warm_ee + timed_ee]
    return dict(solver=s, mesh=m, rest=rest, reps=reps, dt=dt, restarts=restarts, episode=episode, probes=probes,
                ee_window=window)


def episode_summary(ep, max_repair):
    """Repair episode: the number of update steps before a step found the fold
    intersection-free (that last call only validates), their time, and the
    per-step record of the trajectory."""
    done = bool(ep) and ep[-1][0]["ef_total"] == 0 and ep[-1][0]["contours"] == 0
    upd = ep[:-1] if done else ep
    secs = sum(t for _, t in upd)
    return {
        "reached": done,
        "steps_to_intersection_free": len(upd) if done else None,
        "seconds_to_intersection_free": round(secs, 4) if done else None,
        "steps_per_s": round(len(upd) / secs, 4) if upd and secs > 0 else None,
        "step_cap": max_repair,
        "ef_pairs": [int(r["ef_total"]) for r, _ in ep],
        "contours": [int(r["contours"]) for r, _ in ep],
        "pcg_iterations": [int(r["pcg_iterations"]) for r, _ in upd],
        "pcg_converged_steps": int(sum(1 for r, _ in upd if r["pcg_converged"])),
        "pcg_relative_residual_max": max((float(r["pcg_relative_residual"]) for r, _ in upd), default=None),
        "ms_per_step": [round(1e3 * t, 2) for _, t in ep],
    }


def probe_summary(probes):
    """Mixed vs FP64 PCG on identical systems, counted where both reached the tolerance."""
    both = [p for p in probes if p["fp64"]["converged"] and p["mixed"]["converged"]]
    if not both:
        return {"systems": len(probes), "both_converged": 0}
    f = sum(p["fp64"]["ms"] for p in both)
    x = sum(p["mixed"]["ms"] for p in both)
    return {
        "systems": len(probes), "both_converged": len(both),
        "fp64_ms": round(f, 3), "mixed_ms": round(x, 3), "speedup": round(f / x, 3),
        "fp64_iterations_mean": round(sum(p["fp64"]["iterations"] for p in both) / len(both), 1),
        "mixed_iterations_mean": round(sum(p["mixed"]["iterations"] for p in both) / len(both), 1),
        "fp64_residual_max": max(p["fp64"]["relative_residual"] for p in both),
        "mixed_residual_max": max(p["mixed"]["relative_residual"] for p in both),
        "steps": [p["step"] for p in probes],
    }


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def precision_name(p):
    from untangling26_amd.lcs_py import PCG_FP64

    return "fp64" if p == PCG_FP64 else "mixed"


def summarize(reps):
    import numpy as np

    keys = ("ms_detect", "ms_contours", "ms_prp", "ms_assemble", "ms_pcg", "ms_ccd", "ms_total")
    out = {k: float(np.mean([r[k] for r in reps])) for k in keys}
    out["ms_contours"] -= out["ms_prp"]  # the PRP time is inside the contour window
    out["pcg_iterations"] = float(np.mean([r["pcg_iterations"] for r in reps]))
    out["pcg_converged_steps"] = int(sum(1 for r in reps if r["pcg_converged"]))
    out["ef_pairs"] = [int(r["ef_total"]) for r in reps]
    out["contours"] = [int(r["contours"]) for r in reps]
    out["prp_responses"] = float(np.mean([r["prp_responses"] for r in reps]))
    out["prp_levels"] = float(np.mean([r["prp_levels"] for r in reps]))
    out["prp_slots"] = float(np.mean([r["prp_slots"] for r in reps]))
    out["prp_rays"] = {k: float(np.mean([r[f"prp_rays_{k}"] for r in reps])) for k in ("vf", "ee", "fv")}
    out["prp_hits"] = float(np.mean([r["prp_hits"] for r in reps]))
    out["ee"] = {k: float(np.sum([r[f"prp_ee_{k}"] for r in reps]))
                 for k in ("launches", "requests", "visits", "leaf_tests", "hits", "cand_tests", "wave_steps")}
    out["ee"]["ms"] = float(np.sum([r["ms_prp_ee"] for r in reps]))
    return out


# Algorithmic bytes of a ray-cast kernel per unit of work (compulsory HBM reads and
# writes, no cache reuse credit): a request 16 B, a BVH node visit 40 B (a box and a
# child reference), a leaf test 48 B (the target's 36 B of positions + 12 B of
# indices), a hit 32 B.  The PRP EE sweep (k_prp_ray_edge) counts these units itself.
EE_BYTES = {"requests": 16, "visits": 40, "leaf_tests": 48, "hits": 32}


def measured_traffic(args):
    """HBM bytes per k_prp_ray_edge launch from the committed counter passes of this
    workload (FETCH_SIZE + WRITE_SIZE, calibrated on a 256 MiB copy), or None."""
    p = ROOT / "profiles" / "k_prp_ray_edge_folded_cloth_256.traffic.json"
    if not p.exists():
        return None
    t = json.loads(p.read_text())
    if t.get("n") != args.n or t.get("dirs") != args.dirs:
        return None
    return t


def ee_roofline(ee):
    launches = max(ee["launches"], 1.0)
    per_launch = sum(EE_BYTES[k] * ee[k] for k in EE_BYTES) / launches
    avg_ms = ee["ms"] / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return per_launch, avg_ms, achieved


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the path has no multi-GPU implementation "
              "(it would shard by Morton ranges, contours and DOF stripes)", file=sys.stderr)
        sys.exit(2)

    import torch

    from untangling26_amd.lcs_py import PCG_FP64, PCG_MIXED

    if args.copy_probe_mb:
        # a copy of known size: its FETCH_SIZE / WRITE_SIZE counters calibrate the kernel's
        n = args.copy_probe_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    leg = run_leg(args, PCG_MIXED, probe_every=args.probe_every)
    s, m, rest, reps, dt = leg["solver"], leg["mesh"], leg["rest"], leg["reps"], leg["dt"]
    mixed = summarize(reps)
    mixed_ep = episode_summary(leg["episode"], args.max_repair)
    probes = probe_summary(leg["probes"])
    value = args.steps / dt
    # roofline of the dominant kernel, the PRP EE sweep: HIP-event time of its launches in
    # the timed steps and the counters of the same launches
    ee_bytes, ee_ms, ee_achieved = ee_roofline(mixed["ee"])
    traffic = measured_traffic(args)
    # the PCG SpMV (the kernel of the PCG iteration loop), measured live on the last matrix
    pat = s._setup.pattern
    nnzb = int(len(pat.keys))
    spmv_ms = s.profile_spmv(100)
    spmv_bytes = nnzb * (9 * 4 + 4) + m.nv * (12 + 12)
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    data = s.get_intersection_contour_data()

    fp64 = fp64_ep = None
    if not args.no_fp64:
        leg64 = run_leg(args, PCG_FP64)
        fp64 = summarize(leg64["reps"])
        fp64["steps_per_s"] = args.steps / leg64["dt"]
        fp64_ep = episode_summary(leg64["episode"], args.max_repair)
        del leg64

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        from cpu_step import cpu_step_baseline

        cpu = cpu_step_baseline(m, rest, s, args)

    if traffic is not None and abs(traffic.get("algorithmic_bytes_per_launch", 0) / max(ee_bytes, 1.0) - 1.0) > 0.15:
        traffic = None  # the committed counter passes were taken on a different kernel version or workload
    line = {
        "metric": "steps/s to intersection-free, 131k-tri fold; mixed-prec vs FP64 PCG speedup",
        "value": round(value, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded folded cloth, figure-eight crossing)",
        "config": {
            "workload": f"folded_cloth {args.n}x{args.n} ({m.nf} tris) untangling repair step, "
                        f"{args.dirs} directions/contour, PCG rel. residual 1e-6",
            "generator": "scenes.folded_cloth(seed 0): 2 m sheet folded at its midline, the folded half "
                         "offset by a quarter cell in-plane (coincident layers make the EF predicates "
                         "degenerate), z += 0.08 sin(2 pi u) sin(pi v) + N(0, 0.005)",
            "pcg": "mixed FP32 + FP64 refinement",
            "prp_contour_batch": args.prp_batch,
            "timed_steps": "K repair steps from the tangled state; a step that finds the fold "
                           "intersection-free restarts the repair from the tangled state",
            "timed_restarts": leg["restarts"],
            "repair_mixed": mixed_ep,
            "repair_fp64": fp64_ep,
            "steps_to_intersection_free": mixed_ep["steps_to_intersection_free"],
            "seconds_to_intersection_free_mixed": mixed_ep["seconds_to_intersection_free"],
            "seconds_to_intersection_free_fp64": None if fp64_ep is None else fp64_ep["seconds_to_intersection_free"],
            "mixed_vs_fp64_speedup_to_intersection_free": (
                round(fp64_ep["seconds_to_intersection_free"] / mixed_ep["seconds_to_intersection_free"], 3)
                if fp64_ep and fp64_ep["reached"] and mixed_ep["reached"] else None),
            "pcg_mixed_vs_fp64_equal_residual": probes,
            "fp64_steps_per_s": None if fp64 is None else round(fp64["steps_per_s"], 4),
            "mixed_vs_fp64_speedup": None if fp64 is None else round(value / fp64["steps_per_s"], 3),
            "pcg_speedup_mixed_vs_fp64": None if fp64 is None else round(fp64["ms_pcg"] / max(mixed["ms_pcg"], 1e-9), 3),
            "ef_pairs_first_last": [mixed["ef_pairs"][0], mixed["ef_pairs"][-1]],
            "contours_first_last": [mixed["contours"][0], mixed["contours"][-1]],
            "final_num_pairs": int(data["num_pairs"][0]),
            "stage_ms_mixed": {k: round(v, 3) for k, v in mixed.items() if k.startswith("ms_")},
            "stage_ms_fp64": None if fp64 is None else {k: round(v, 3) for k, v in fp64.items()
                                                         if k.startswith("ms_")},
            "pcg_iterations_mixed": mixed["pcg_iterations"],
            "pcg_iterations_fp64": None if fp64 is None else fp64["pcg_iterations"],
            "pcg_converged_steps_mixed": mixed["pcg_converged_steps"],
            "prp_responses_per_step": mixed["prp_responses"],
            "prp_levels_per_step": mixed["prp_levels"],
            "prp_slots_per_step": mixed["prp_slots"],
            "prp_rays_per_step": mixed["prp_rays"],
            "prp_raw_hits_per_step": mixed["prp_hits"],
            "prp_ee_kernel": {"launches": mixed["ee"]["launches"], "avg_launch_ms": round(ee_ms, 4),
                              "timed_dispatch_window": leg["ee_window"],
                              "requests_per_launch": mixed["ee"]["requests"] / max(mixed["ee"]["launches"], 1),
                              "visits_per_launch": mixed["ee"]["visits"] / max(mixed["ee"]["launches"], 1),
                              "leaf_tests_per_launch": mixed["ee"]["leaf_tests"] / max(mixed["ee"]["launches"], 1),
                              "hits_per_launch": mixed["ee"]["hits"] / max(mixed["ee"]["launches"], 1),
                              "candidate_leaf_tests_per_launch":
                                  mixed["ee"]["cand_tests"] / max(mixed["ee"]["launches"], 1),
                              # node visits per lane-step of the traversal loop: 1.0 when every lane of
                              # every wave visits a node at every step (no divergence, no idle lanes)
                              "lane_utilisation": round(mixed["ee"]["visits"]
                                                        / max(64.0 * mixed["ee"]["wave_steps"], 1.0), 4),
                              "bytes_per_unit": EE_BYTES},
            "roofline_pcg_spmv": {"kernel": "k_pcg_spmv<float>", "bound": "hbm", "achieved": round(achieved, 2),
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                                  "bytes_per_launch": spmv_bytes, "avg_launch_ms": round(spmv_ms, 5)},
        },
        "roofline": {
            "kernel": "k_prp_ray_edge (PRP EE sweep: LBVH traversal + edge-edge tests)",
            "bound": "hbm",
            "limiter": "latency of dependent BVH node fetches: the counter passes of this workload show "
                       "HBM traffic well below the algorithmic bytes (most node reads hit L2) and the "
                       "VALU busy under half of its issue slots",
            "achieved": round(ee_achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(ee_achieved / HBM_PEAK_GBS, 4),
            "traffic": None if traffic is None else traffic["traffic_bytes_per_launch"],
            "traffic_source": None if traffic is None else traffic["source"],
            "bytes_per_launch": round(ee_bytes),
            "avg_launch_ms": round(ee_ms, 4),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
