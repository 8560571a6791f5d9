"""bench.py's host arithmetic, on the CPU: the roofline of the PRP EE sweep
(algorithmic bytes per launch from the per-unit figures, achieved rate from the
HIP-event time), the guard on the committed counter passes (used only for the
workload they were taken on), and the per-step summary of solver reports."""
import argparse
import math

import bench


def test_ee_roofline_closed_form():
    ee = {"launches": 4.0, "requests": 400.0, "visits": 4000.0, "leaf_tests": 800.0, "cand_tests": 200.0,
          "hits": 40.0, "ms": 8.0}
    per_launch, avg_ms, achieved = bench.ee_roofline(ee)
    # per-unit bytes of a ray cast: 16 B request, 40 B node visit, 48 B leaf test, 32 B hit
    # (candidate tests are leaf tests already counted)
    want = (16 * 400 + 40 * 4000 + 48 * 800 + 32 * 40) / 4
    assert per_launch == want
    assert avg_ms == 2.0
    assert math.isclose(achieved, want / 2e-3 / 1e9)
    # no launch timed: no rate
    _, _, a0 = bench.ee_roofline(dict(ee, ms=0.0))
    assert a0 == 0.0


def test_measured_traffic_only_for_its_workload():
    t = bench.measured_traffic(argparse.Namespace(n=256, dirs=64))
    assert t is not None and t["n"] == 256 and t["dirs"] == 64
    # each figure is rounded to whole bytes after scaling
    assert abs(t["traffic_bytes_per_launch"] - t["fetch_bytes_per_launch"] - t["write_bytes_per_launch"]) <= 1
    assert bench.measured_traffic(argparse.Namespace(n=128, dirs=64)) is None
    assert bench.measured_traffic(argparse.Namespace(n=256, dirs=10)) is None


def _rep(ms, ef, conv):
    r = {k: ms for k in ("ms_detect", "ms_contours", "ms_prp", "ms_assemble", "ms_pcg", "ms_ccd", "ms_total")}
    r.update(ms_contours=3 * ms, ms_prp=2 * ms, pcg_iterations=100, pcg_converged=conv, pcg_relative_residual=1e-7,
             ef_total=ef, contours=2,
             prp_responses=10, prp_levels=3, prp_slots=5, prp_hits=7, ms_prp_ee=1.0)
    for k in ("vf", "ee", "fv"):
        r[f"prp_rays_{k}"] = 11
    for k in ("launches", "requests", "visits", "leaf_tests", "hits", "cand_tests", "wave_steps"):
        r[f"prp_ee_{k}"] = 2
    return r


def test_summarize_means_and_prp_window():
    s = bench.summarize([_rep(1.0, 30, True), _rep(3.0, 20, False)])
    assert s["ms_pcg"] == 2.0 and s["ms_total"] == 2.0
    # the PRP time is inside the contour window: contours alone = 3 ms - 2 ms (mean 6 - 4)
    assert s["ms_prp"] == 4.0 and s["ms_contours"] == 2.0
    assert s["pcg_converged_steps"] == 1
    assert s["ef_pairs"] == [30, 20]
    assert s["ee"]["launches"] == 4.0 and s["ee"]["ms"] == 2.0


def test_episode_summary_counts_update_steps():
    ep = [(_rep(1.0, 30, True), 0.10), (_rep(1.0, 12, True), 0.20), (dict(_rep(1.0, 0, True), contours=0), 0.05)]
    e = bench.episode_summary(ep, 200)
    # the last call found the fold intersection-free: two update steps, 0.3 s
    assert e["reached"] and e["steps_to_intersection_free"] == 2
    assert math.isclose(e["seconds_to_intersection_free"], 0.3) and math.isclose(e["steps_per_s"], 2 / 0.3, rel_tol=1e-3)
    assert e["ef_pairs"] == [30, 12, 0]
    open_ep = bench.episode_summary(ep[:2], 2)
    assert not open_ep["reached"] and open_ep["steps_to_intersection_free"] is None


def test_probe_summary_counts_only_converged_pairs():
    def p(step, f_ms, m_ms, fc=1, mc=1):
        return {"step": step, "fp64": {"ms": f_ms, "iterations": 10, "relative_residual": 1e-7, "converged": fc},
                "mixed": {"ms": m_ms, "iterations": 12, "relative_residual": 9e-7, "converged": mc}}
    s = bench.probe_summary([p(0, 4.0, 2.0), p(3, 6.0, 3.0), p(6, 100.0, 1.0, mc=0)])
    assert s["both_converged"] == 2 and s["speedup"] == 2.0 and s["systems"] == 3
