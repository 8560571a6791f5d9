"""ctypes binding of libuntangle_hip.so (the C ABI declared in include/untangle.h).

The shared library is the only compute path.  If it is missing, importing the
package still works (so CPU-only tooling can inspect it), but every call that
needs the library raises ``RuntimeError`` -- there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

_PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = pathlib.Path(os.environ.get("UT_LIB_PATH", _PKG / "libuntangle_hip.so"))

_lib = None

u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p

# name -> (restype, argtypes); mirrors include/untangle.h
SIGNATURES = {
    "ut_version": (C.c_int, []),
    "ut_ctx_create": (C.c_int, [C.c_int, vp, C.c_int, C.POINTER(vp)]),
    "ut_ctx_destroy": (C.c_int, [vp]),
    "ut_last_error": (C.c_char_p, [vp]),
    "ut_ctx_synchronize": (C.c_int, [vp]),
    "ut_ctx_stream": (vp, [vp]),
    "ut_lbvh_create": (C.c_int, [vp, C.c_int, C.c_uint32, C.POINTER(vp)]),
    "ut_lbvh_destroy": (C.c_int, [vp]),
    "ut_lbvh_reduce_aabb": (C.c_int, [vp, vp, vp]),
    "ut_lbvh_construct_tree": (C.c_int, [vp]),
    "ut_lbvh_build_from_sorted_keys": (C.c_int, [vp, vp]),
    "ut_lbvh_refit": (C.c_int, [vp, vp, vp, vp, vp]),
    "ut_lbvh_query": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, C.c_uint32]),
    "ut_lbvh_tree_arrays": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "ut_lbvh_boxes": (C.c_int, [vp, vp, vp]),
    "ut_lbvh_status": (C.c_int, [vp, u32p, u32p]),
    "ut_morton30": (C.c_int, [vp, vp, C.c_uint32, vp]),
    "ut_ef_dcd": (C.c_int, [vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, C.c_uint32]),
    "ut_ef_detect": (C.c_int, [vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp, vp, vp, C.c_uint32]),
    "ut_contours_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, vp, vp, C.POINTER(vp)]),
    "ut_contours_destroy": (C.c_int, [vp]),
    "ut_ef_from_proximity": (C.c_int, [vp, vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                       C.c_uint32]),
    "ut_contours_build": (C.c_int, [vp, C.c_uint32, vp, vp, C.c_int, u32p, u32p]),
    "ut_contours_counts": (C.c_int, [vp, u32p]),
    "ut_contours_get": (C.c_int, [vp, C.c_int, vp]),
    "ut_graph_components": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp, vp]),
    "ut_solver_pcg_resolve": (C.c_int, [vp, C.c_int, vp, vp]),
    "ut_solver_get_positions": (C.c_int, [vp, C.c_int, vp]),
    "ut_solver_responses": (C.c_int, [vp, vp, vp, C.c_uint32, vp]),
    "ut_graph_components_static": (C.c_int, [vp, C.c_uint32, vp, vp, vp, vp, vp, vp]),
    "ut_pcg_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, vp, C.POINTER(vp)]),
    "ut_pcg_destroy": (C.c_int, [vp]),
    "ut_pcg_solve": (C.c_int, [vp, vp, vp, vp, vp, vp]),
    "ut_pcg_solve_refined": (C.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "ut_pcg_set_contacts": (C.c_int, [vp, vp, vp, C.c_uint32]),
    "ut_contact_energy": (C.c_int, [vp, vp, vp, C.c_uint32, vp, vp, vp, C.c_double, C.c_int, vp]),
    "ut_friction_init": (C.c_int, [vp, vp, vp, C.c_uint32, vp]),
    "ut_friction_lagged": (C.c_int, [vp, vp, vp, C.c_uint32, vp, vp, vp, vp, vp]),
    "ut_contact_potential": (C.c_int, [vp, vp]),
    "ut_abd_q_to_x": (C.c_int, [vp, vp, vp, vp, vp]),
    "ut_abd_assemble": (C.c_int, [vp, vp, vp]),
    "ut_abd_apply": (C.c_int, [vp, vp, vp, vp, vp, C.c_uint32, vp, vp]),
    "ut_abd_gather": (C.c_int, [vp, vp, vp, vp, C.c_int]),
    "ut_abd_pcg_create": (C.c_int, [vp, C.c_uint32, C.POINTER(vp)]),
    "ut_abd_pcg_destroy": (C.c_int, [vp]),
    "ut_abd_pcg_solve": (C.c_int, [vp, vp, vp, vp, vp, C.c_uint32, vp, vp, C.c_int, C.c_uint32, C.c_double, vp]),
    "ut_solver_get_q": (C.c_int, [vp, vp, vp]),
    "ut_solver_set_q": (C.c_int, [vp, vp, vp]),
    "ut_solver_profile": (C.c_int, [vp, C.c_int, C.c_uint32, C.POINTER(C.c_float)]),
    "ut_prp_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, C.c_uint32] + [vp] * 10 + [C.POINTER(vp)]),
    "ut_prp_destroy": (C.c_int, [vp]),
    "ut_prp_set_params": (C.c_int, [vp, vp]),
    "ut_prp_resolve": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_uint32, u32p, vp, C.c_uint32, vp]),
    "ut_prp_prepare": (C.c_int, [vp, vp, vp, C.c_uint32, C.c_uint32]),
    "ut_prp_directions": (C.c_int, [vp, C.c_uint32, vp, u32p]),
    "ut_prp_distances": (C.c_int, [vp, C.c_uint32, C.c_uint32, vp, vp, u32p]),
    "ut_prp_eval": (C.c_int, [vp, vp, vp, vp, C.c_uint32, vp, vp, u32p]),
    "ut_prp_hits": (C.c_int, [vp, vp, C.c_uint32]),
    "ut_prp_normals": (C.c_int, [vp, vp]),
    "ut_bsr_spmv": (C.c_int, [vp, vp, C.c_int, vp, vp]),
    "ut_contact_query": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, C.c_uint32]),
    "ut_toi_reset": (C.c_int, [vp, vp]),
    "ut_assemble": (C.c_int, [vp, vp]),
    "ut_solver_create": (C.c_int, [vp, vp, vp, vp, vp, C.POINTER(vp)]),
    "ut_solver_destroy": (C.c_int, [vp]),
    "ut_solver_step": (C.c_int, [vp, vp]),
    "ut_solver_set_params": (C.c_int, [vp, vp]),
    "ut_solver_get_state": (C.c_int, [vp, vp, vp]),
    "ut_solver_set_state": (C.c_int, [vp, vp, vp]),
    "ut_solver_contour_get": (C.c_int, [vp, C.c_int, vp]),
}


class PcgParams(C.Structure):
    _fields_ = [("precision", C.c_int), ("preconditioner", C.c_int), ("max_iters", C.c_uint32),
                ("rel_tol", C.c_double), ("inner_tol", C.c_double), ("max_refine", C.c_uint32),
                ("regularization", C.c_double), ("check_every", C.c_uint32), ("recurrence", C.c_int)]


class PcgReport(C.Structure):
    _fields_ = [("relative_residual", C.c_double), ("infinity_norm", C.c_double), ("iterations", C.c_uint32),
                ("refinements", C.c_uint32), ("converged", C.c_int), ("failure", C.c_int)]


class ContactArgs(C.Structure):
    _fields_ = [("num_verts", C.c_uint32), ("num_edges", C.c_uint32), ("x", vp), ("x_end", vp), ("edges", vp),
                ("faces", vp), ("vprop", vp), ("d_hat", vp), ("offset", vp), ("rest_vert_area", vp),
                ("rest_edge_area", vp), ("rest_face_area", vp), ("kappa", C.c_double), ("energy", C.c_int),
                ("count", vp), ("pairs", vp), ("pair_ids", vp), ("max_pairs", C.c_uint32), ("toi", vp)]


class AssembleArgs(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("nf", C.c_uint32), ("nh", C.c_uint32), ("nslots", C.c_uint32), ("x", vp),
                ("x_tilde", vp), ("mass", vp), ("inv_h2", C.c_double), ("faces", vp), ("dm", vp), ("face_area", vp),
                ("mu_lambda", vp), ("face_slots", vp), ("hinges", vp), ("rest_angle", vp), ("bend_stiff_area", vp),
                ("hinge_slots", vp), ("contacts", vp), ("contact_count", vp), ("contact_cap", C.c_uint32),
                ("diag_slot", vp), ("grad", vp), ("vals", vp), ("vals32", vp), ("energy", vp),
                ("quasi_static", C.c_int), ("qs_no_inertia_hessian", C.c_int), ("x64", vp)]


class PrpParams(C.Structure):
    _fields_ = [("max_contours", C.c_uint32), ("total_dirs", C.c_uint32), ("batch_contours", C.c_uint32),
                ("response_depth", C.c_float), ("kappa", C.c_float), ("tau", C.c_float)]


class SolverDesc(C.Structure):
    _fields_ = [("nv", C.c_uint32), ("ne", C.c_uint32), ("nf", C.c_uint32), ("nh", C.c_uint32)] + [
        (n, vp) for n in ("edges", "faces", "face_adj_edges", "edge_adj_faces", "mass", "d_hat", "offset", "vprop",
                          "rest_vert_area", "rest_edge_area", "rest_face_area", "dm", "mu_lambda", "hinges",
                          "rest_angle", "bend_stiff_area")] + [
        ("nslices", C.c_uint32), ("nslots", C.c_uint32)] + [
        (n, vp) for n in ("slice_off", "col", "diag_slot", "face_slots", "hinge_slots", "rest_x", "friction_mu",
                          "abd", "q0")]


class PotentialArgs(C.Structure):
    _fields_ = [("pairs", vp), ("count", vp), ("cap", C.c_uint32), ("x", vp), ("x64", vp), ("x_start", vp),
                ("x_start64", vp), ("d_hat", vp), ("offset", vp), ("kappa", C.c_double), ("energy", C.c_int),
                ("terms", C.c_int), ("grad", vp), ("out", vp)]


TERM_BARRIER, TERM_FRICTION = 1, 2


class AbdDesc(C.Structure):
    _fields_ = [("nb", C.c_uint32), ("nv", C.c_uint32), ("vert_off", vp), ("body", vp), ("xbar", vp),
                ("mass4", vp), ("volume", vp), ("kappa", vp)]


class AbdAssembleArgs(C.Structure):
    _fields_ = [("q", vp), ("q_tilde", vp), ("inv_h2", C.c_double), ("quasi_static", C.c_int),
                ("qs_no_inertia_hessian", C.c_int), ("contacts", vp), ("contact_count", vp),
                ("contact_cap", C.c_uint32), ("grad", vp), ("blocks", vp), ("energy", vp)]


class SolverParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3), ("kappa", C.c_double), ("energy", C.c_int),
                ("damping", C.c_double), ("nonlinear_iters", C.c_int), ("use_self_collision", C.c_int),
                ("use_ccd", C.c_int), ("use_untangling", C.c_int), ("pcg", PcgParams), ("use_prp", C.c_int),
                ("quasi_static", C.c_int), ("prp", PrpParams), ("region_filter", C.c_int),
                ("ignore_near_zero_dist_pairs", C.c_int)]


class StepReport(C.Structure):
    _fields_ = [("vf_pairs", C.c_uint32), ("ee_pairs", C.c_uint32), ("ef_pairs", C.c_uint32),
                ("ef_total", C.c_uint32), ("contours", C.c_uint32), ("pcg_iterations", C.c_uint32),
                ("pcg_refinements", C.c_uint32), ("pcg_relative_residual", C.c_double), ("pcg_converged", C.c_int),
                ("toi", C.c_float), ("energy", C.c_double * 3), ("ms_detect", C.c_float), ("ms_contours", C.c_float),
                ("ms_assemble", C.c_float), ("ms_pcg", C.c_float), ("ms_ccd", C.c_float), ("ms_total", C.c_float),
                ("prp_responses", C.c_uint32), ("prp_contours", C.c_uint32), ("prp_levels", C.c_uint32),
                ("prp_hits", C.c_uint32), ("ms_prp", C.c_float), ("prp_slots", C.c_uint32),
                ("prp_rays_vf", C.c_uint64), ("prp_rays_ee", C.c_uint64), ("prp_rays_fv", C.c_uint64),
                ("prp_ee_launches", C.c_uint32), ("ms_prp_ee", C.c_float), ("prp_ee_requests", C.c_uint64),
                ("prp_ee_visits", C.c_uint64), ("prp_ee_leaf_tests", C.c_uint64), ("prp_ee_hits", C.c_uint64),
                ("prp_ee_cand_tests", C.c_uint64), ("prp_ee_wave_steps", C.c_uint64)]


class PrpChoice(C.Structure):
    _fields_ = [("combo", C.c_int32), ("ks", C.c_uint32), ("kd", C.c_uint32), ("hit_count", C.c_uint32),
                ("covered", C.c_uint32), ("ef_count", C.c_uint32), ("levels", C.c_uint32),
                ("suppressed", C.c_uint32), ("dir", C.c_float * 3), ("depth", C.c_float), ("objective", C.c_float),
                ("area", C.c_float), ("disp_cost", C.c_float), ("pad", C.c_uint32)]


class PrpStats(C.Structure):
    _fields_ = [("contours", C.c_uint32), ("levels", C.c_uint32), ("slots", C.c_uint32), ("hits", C.c_uint32),
                ("responses", C.c_uint32), ("ms_prepare", C.c_float), ("ms_levels", C.c_float),
                ("ms_final", C.c_float), ("rays_vf", C.c_uint64), ("rays_ee", C.c_uint64), ("rays_fv", C.c_uint64)]


def lib():
    """Loads the library once; raises RuntimeError if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python -m untangling26_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None passes NULL)."""
    if t is None:
        return None
    return t.data_ptr()


class UntangleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"untangle error {code}: {msg}")
        self.code = code


def check(rc: int, ctx_handle) -> int:
    if rc < 0:
        msg = lib().ut_last_error(ctx_handle)
        raise UntangleError(rc, msg.decode() if msg else "")
    return rc
