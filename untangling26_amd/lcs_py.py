"""Python mirror of the reference's `lcs_py` binding (PythonBindings/src/python_bindings.cpp:1207-1500,
stub PythonBindings/python/lcs_py/__init__.pyi) for the per-step path.

Supported surface: NewtonSolver with create_world_data_from_array,
register_world_data, init_device / set_device, init_solver, physics_step_gpu,
get_sim_result, get_intersection_contour_data, get_config, get_step_report;
SceneParams holds the fields the step reads.  physics_step_cpu is not offered:
this build has no CPU physics path (every call runs the gfx950 kernels).
Errors surface as RuntimeError, as the pybind11 binding raises for
std::runtime_error.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np
import torch

from . import initializer
from ._lib import PcgParams, PrpParams, SolverDesc, SolverParams, StepReport, lib, ptr
from .contours import CT_COMP_MIN, CT_CONTOUR_INDEX, CT_MESH_INDEX, CT_PAIR_IDS, CT_PAIRS
from .runtime import Context
from .scenes import Mesh

PCG_FP64, PCG_MIXED, PCG_FP32 = 0, 1, 2


@dataclass
class SceneParams:
    """SceneParams fields used by the step (Solver/SimulationCore/scene_params.h:24-157)."""
    implicit_dt: float = 1.0 / 60.0
    gravity: tuple = (0.0, -9.8, 0.0)
    stiffness_collision: float = 1e8
    contact_energy_type: int = 1
    damping_rate: float = 2.0
    nonlinear_iter_count: int = 1
    pcg_iter_count: int = 100
    pcg_rel_tol: float = 1e-6
    use_self_collision: bool = True
    use_ccd_linesearch: bool = True
    use_untangling: bool = True
    d_hat: float = 1e-3
    use_untangling_PRP: bool = True
    use_quasi_static_mode: bool = False
    untangling_response_depth: float = 3e-3
    untangling_process_contours_count: int = 256
    PRP_cpu_contour_batch_size: int = 32
    PRP_intrinsic_tau: float = 0.25
    stiffness_untangling: float = 1e8
    # additions of this build
    pcg_precision: int = PCG_MIXED
    pcg_inner_tol: float = 1e-2
    pcg_block_jacobi: bool = True
    prp_directions: int = 10  # 10 = the reference's direction generator; more appends a Fibonacci set
    # filter_invalid_pairs / filter_broad_phase_pairs on proximity and CCD pairs; the reference
    # runs both whenever use_untangling is set (newton_solver.cpp:1308-1311, :2513-2517).
    # Without them the CCD line search blocks the repair motion through the contours.
    use_region_filter: bool = True
    # scene_params.h:55 (the CLI sets it False, cli.py:107): False adds EF pairs from
    # zero-distance proximity pairs and drops isolated degenerate EF pairs
    ignore_near_zero_dist_pairs: bool = True


@dataclass
class WorldData:
    name: str
    vertices: np.ndarray
    triangles: np.ndarray
    rest_vertices: np.ndarray | None = None
    material: initializer.ClothMaterial = field(default_factory=initializer.ClothMaterial)


class NewtonSolver:
    def __init__(self):
        self._worlds: list[WorldData] = []
        self._ctx: Context | None = None
        self._handle = None
        self._config = SceneParams()
        self._setup = None
        self._dev = {}
        self._last_report = None

    # -- scene ------------------------------------------------------------------------
    def create_world_data_from_array(self, name, vertices, triangles, rest_vertices=None):
        return WorldData(name, np.asarray(vertices, np.float32).reshape(-1, 3),
                         np.asarray(triangles, np.uint32).reshape(-1, 3),
                         None if rest_vertices is None else np.asarray(rest_vertices, np.float32).reshape(-1, 3))

    def register_world_data(self, wd: WorldData) -> int:
        self._worlds.append(wd)
        return len(self._worlds) - 1

    def get_config(self) -> SceneParams:
        return self._config

    # -- device -----------------------------------------------------------------------
    def init_device(self, backend_name: str = "hip", binary_path: str = "", device: int = 0):
        if backend_name not in ("hip", "cuda", "default"):
            raise RuntimeError(f"unsupported backend {backend_name!r}: this build runs on HIP only")
        self._ctx = Context(device)

    def set_device(self, ctx: Context):
        self._ctx = ctx

    # -- solver -----------------------------------------------------------------------
    def _merged_mesh(self):
        xs, fs, rs, off = [], [], [], 0
        for w in self._worlds:
            xs.append(w.vertices)
            fs.append(w.triangles.astype(np.int64) + off)
            rs.append(w.vertices if w.rest_vertices is None else w.rest_vertices)
            off += len(w.vertices)
        return Mesh(np.concatenate(xs), np.concatenate(fs).astype(np.uint32)), np.concatenate(rs)

    def _params(self) -> SolverParams:
        c = self._config
        p = SolverParams()
        p.dt = c.implicit_dt
        for k in range(3):
            p.gravity[k] = c.gravity[k]
        p.kappa = c.stiffness_collision
        p.energy = c.contact_energy_type
        p.damping = c.damping_rate
        p.nonlinear_iters = c.nonlinear_iter_count
        p.use_self_collision = int(c.use_self_collision)
        p.use_ccd = int(c.use_ccd_linesearch)
        p.use_untangling = int(c.use_untangling)
        p.pcg = PcgParams(c.pcg_precision, 0 if c.pcg_block_jacobi else 1, c.pcg_iter_count, c.pcg_rel_tol,
                          c.pcg_inner_tol, 20, 0.0, 8)
        p.use_prp = int(c.use_untangling_PRP)
        p.quasi_static = int(c.use_quasi_static_mode)
        p.prp = PrpParams(c.untangling_process_contours_count, c.prp_directions, c.PRP_cpu_contour_batch_size,
                          c.untangling_response_depth, c.stiffness_untangling, c.PRP_intrinsic_tau)
        p.region_filter = int(c.use_region_filter)
        p.ignore_near_zero_dist_pairs = int(c.ignore_near_zero_dist_pairs)
        return p

    def init_solver(self):
        if self._ctx is None:
            self.init_device()
        if not self._worlds:
            raise RuntimeError("no world data registered")
        mesh, rest = self._merged_mesh()
        mat = self._worlds[0].material
        mat.d_hat = self._config.d_hat
        st = initializer.setup_cloth(mesh, mat, x_rest=rest)
        self._setup = st
        dv = self._ctx.device

        def up(a):
            t = torch.from_numpy(np.ascontiguousarray(a))
            if t.dtype == torch.uint32:
                t = t.view(torch.int32)
            return t.to(dv)

        pat = st.pattern
        names = dict(edges=st.edges, faces=st.faces, face_adj_edges=st.face_adj_edges,
                     edge_adj_faces=st.edge_adj_faces, mass=st.mass, d_hat=st.d_hat, offset=st.offset,
                     rest_vert_area=st.rest_vert_area, rest_edge_area=st.rest_edge_area,
                     rest_face_area=st.rest_face_area, dm=st.dm, mu_lambda=st.mu_lambda, hinges=st.hinges,
                     rest_angle=st.rest_angle, bend_stiff_area=st.bend_stiff_area, slice_off=pat.slice_off,
                     col=pat.col, diag_slot=pat.diag_slot, face_slots=st.face_slots, hinge_slots=st.hinge_slots)
        self._dev = {k: up(v) for k, v in names.items()}
        d = SolverDesc()
        d.nv, d.ne, d.nf, d.nh = st.nv, len(st.edges), len(st.faces), len(st.hinges)
        for k, t in self._dev.items():
            setattr(d, k, ptr(t))
        d.vprop = None
        d.nslices, d.nslots = pat.nslices, pat.nslots
        self._dev["rest_x"] = up(np.asarray(rest, np.float32))
        d.rest_x = ptr(self._dev["rest_x"])
        x0 = up(mesh.x)
        h = C.c_void_p()
        self._ctx.check(lib().ut_solver_create(self._ctx.handle, C.byref(d), C.byref(self._params()), ptr(x0),
                                               None, C.byref(h)))
        self._handle = h
        self._desc = d

    def __del__(self):
        try:
            if self._handle:
                lib().ut_solver_destroy(self._handle)
                self._handle = None
        except Exception:
            pass

    def physics_step_gpu(self):
        if self._handle is None:
            raise RuntimeError("init_solver() first")
        self._ctx.check(lib().ut_solver_set_params(self._handle, C.byref(self._params())))
        r = StepReport()
        self._ctx.check(lib().ut_solver_step(self._handle, C.byref(r)))
        self._last_report = {k: (list(getattr(r, k)) if k == "energy" else getattr(r, k)) for k, _ in
                             StepReport._fields_}
        return self._last_report

    def profile_spmv(self, reps=50):
        """Average ms of the PCG SpMV kernel on the last assembled matrix (HIP events)."""
        ms = C.c_float()
        self._ctx.check(lib().ut_solver_profile(self._handle, 0, reps, C.byref(ms)))
        return ms.value

    def pcg_resolve(self, precision):
        """Solve the last step's system again at `precision` (PCG_FP64, or PCG_MIXED when
        the solver runs mixed); the solver state is unchanged.  Returns the solve's report
        (relative_residual, iterations, refinements, converged, failure) and its ms."""
        from ._lib import PcgReport

        r = PcgReport()
        ms = C.c_float()
        self._ctx.check(lib().ut_solver_pcg_resolve(self._handle, int(precision), C.byref(r), C.byref(ms)))
        out = {k: getattr(r, k) for k, _ in PcgReport._fields_}
        out["ms"] = ms.value
        return out

    def physics_step_cpu(self):
        raise RuntimeError("physics_step_cpu is not provided: this build runs the step on the GPU only")

    def get_step_report(self):
        return self._last_report

    # -- results (python_bindings.cpp:485-829) ------------------------------------------
    def positions(self) -> torch.Tensor:
        x = torch.empty((self._setup.nv, 3), dtype=torch.float32, device=self._ctx.device)
        self._ctx.check(lib().ut_solver_get_state(self._handle, ptr(x), None))
        return x

    def set_state(self, x: torch.Tensor, v: torch.Tensor | None = None):
        self._ctx.check(lib().ut_solver_set_state(self._handle, ptr(x.contiguous()),
                                                  None if v is None else ptr(v.contiguous())))

    def set_positions(self, x: torch.Tensor):
        self._ctx.check(lib().ut_solver_set_state(self._handle, ptr(x.contiguous()), None))

    def get_sim_result(self):
        """(list of float32[N,3] per world, list of uint32[M,3] per world) -- fresh host copies."""
        x = self.positions().cpu().numpy()
        verts, tris, off = [], [], 0
        for w in self._worlds:
            n = len(w.vertices)
            verts.append(x[off:off + n].copy())
            tris.append(w.triangles.copy())
            off += n
        return verts, tris

    def get_intersection_contour_data(self):
        """num_pairs / num_contours and per-pair arrays of the last step (python_bindings.cpp:635-829)."""
        r = self._last_report or {}
        n = int(r.get("ef_total", 0))
        out = {"num_pairs": n, "num_contours": int(r.get("contours", 0))}
        if n == 0:
            out.update(ef_pairs=np.zeros((0, 5), np.uint32), ef_indices=np.zeros((0, 2), np.uint32),
                       contour_index=np.zeros(0, np.uint32), mesh_index=np.zeros(0, np.uint32))
            return out

        def get(what, shape):
            t = torch.empty(shape, dtype=torch.int32, device=self._ctx.device)
            self._ctx.check(lib().ut_solver_contour_get(self._handle, what, ptr(t)))
            return t.cpu().numpy().view(np.uint32)

        pairs = get(CT_PAIRS, (n, 12))
        out.update(ef_pairs=pairs[:, :5].copy(), ef_indices=get(CT_PAIR_IDS, (n, 2)),
                   contour_index=get(CT_CONTOUR_INDEX, (n,)), mesh_index=get(CT_MESH_INDEX, (n,)),
                   contour_min_pair=get(CT_COMP_MIN, (n,)))
        return out
