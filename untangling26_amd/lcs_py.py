"""Python mirror of the reference's `lcs_py` binding (PythonBindings/src/python_bindings.cpp:1207-1500,
stub PythonBindings/python/lcs_py/__init__.pyi) for the per-step untangling path.

The module surface a caller of the reference uses (untangling26/cli.py:93-227) is kept:
``MaterialType``, ``QuasiStaticInertiaMode``, ``Float3``, ``SceneParams`` (the reference's
field names, ``gravity`` a ``Float3``), ``WorldData`` (material, auto-scaling, fixed
points, transforms) and ``NewtonSolver`` (scene registration, ``init_device`` /
``set_device`` / ``cleanup_device``, ``init_solver``, ``physics_step_gpu`` /
``physics_step_cpu``, ``get_sim_result``, ``get_intersection_contour_data`` and
``get_response_pair_data`` with the reference's dictionary keys, ``.state`` / OBJ state
save and load, ``get_prp_debug_info``, vertex-id queries).

Every step runs the gfx950 kernels through the C ABI (include/untangle.h); there is no CPU
physics in this build, so ``physics_step_cpu`` runs the same GPU step (the CLI's default
``--physics cpu`` then repairs on the GPU).  Parameters may change between steps, including
after ``init_solver`` (as cli.py does).  Scene features outside this path -- rigid,
tetrahedral and rod materials, the floor, ICM / GIA untangling, PRP direction optimisation,
sub-stepping, energy line search, static mode -- raise ``RuntimeError`` when enabled,
instead of being ignored.  Errors surface as RuntimeError, as the pybind11 binding raises for
std::runtime_error.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
import pathlib

import numpy as np
import torch

from . import initializer
from ._lib import AbdDesc, PcgParams, PrpParams, SolverDesc, SolverParams, StepReport, lib, ptr
from .contours import (CONTOUR_TYPES, CT_ADJ, CT_ADJ_OFF, CT_COMP_MIN, CT_CONTOUR_INDEX, CT_CONTOUR_TYPE,
                       CT_MESH_CLOSED, CT_MESH_INDEX, CT_PAIR_IDS, CT_PAIRS)
from .runtime import Context
from .scenes import Mesh, edges_from_faces

PCG_FP64, PCG_MIXED, PCG_FP32 = 0, 1, 2


class MaterialType(enum.IntEnum):
    Cloth = 0
    Tetrahedral = 1
    Rigid = 2
    Rod = 3


class QuasiStaticInertiaMode(enum.IntEnum):
    Off = 0
    HessianOnly = 1
    Full = 2


class Float3:
    """luisa::float3 as bound by the reference (python_bindings.cpp:1423-1431)."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __getitem__(self, i):
        return (self.x, self.y, self.z)[i]

    def __eq__(self, other):
        return tuple(self) == tuple(other)

    def __repr__(self):
        return f"Float3({self.x}, {self.y}, {self.z})"


def mat_friction_mu(mat):
    """MaterialBase::default_friction_mu (physical_material.h:66) unless the material sets one."""
    return float(getattr(mat, "friction_mu", 0.5))


def _float3(v) -> Float3:
    return v if isinstance(v, Float3) else Float3(*v)


class SceneParams:
    """SceneParams (Solver/SimulationCore/scene_params.h:24-157) with the reference's
    field names and defaults for the fields the step reads, and this build's additions
    (pcg_precision, pcg_rel_tol, pcg_inner_tol, pcg_block_jacobi, pcg_single_reduction, prp_directions,
    use_region_filter).  ``gravity`` and ``floor`` accept a Float3 or any 3-sequence."""

    _DEFAULTS = dict(
        use_gpu=True, fix_scene=False, use_energy_linesearch=False, use_ccd_linesearch=True, use_global_ccd=True,
        use_quasi_static_mode=False, use_static_mode=False, qs_inertia_mode=QuasiStaticInertiaMode.HessianOnly,
        print_system_energy=False, print_pcg_info=False, print_collision_info=False, use_floor=False,
        use_self_collision=True, output_per_frame=False, output_per_iteration=False, scene_id=0,
        load_state_frame=0, num_substep=1, nonlinear_iter_count=1, pcg_iter_count=100, current_frame=0,
        collision_detection_frequece=1, contact_energy_type=1, implicit_dt=1.0 / 60.0, explicit_dt=1.0 / 60.0,
        stiffness_bending_ui=1.0, stiffness_collision=1e8, stiffness_untangling=1e8, stiffness_dirichlet=1e9,
        damping_rate=2.0, d_hat=1e-3, use_untangling=False, use_untangling_ICM=False, use_untangling_GIA=False,
        use_untangling_PRP=False, use_gpu_untangling=False, PRP_use_intrinsic_contour_side_candidates=True,
        PRP_use_rest_geodesic_distance_for_intrinsic_candidates=True, PRP_intrinsic_tau=0.25,
        PRP_use_blended_intrinsic_coordinates=False, PRP_intrinsic_blend_weight=0.5,
        PRP_use_deformed_boundary_distance_for_intrinsic_candidates=False, PRP_direction_optimization_iterations=0,
        accumulate_ICM_correction=False, GIA_use_EF_response=False, ignore_near_zero_dist_pairs=True,
        PRP_cpu_contour_batch_size=32, untangling_process_contours_count=256, untangling_response_depth=3e-3,
        pcg_lm_adaptive=False, pcg_lm_escalation_factor=10.0, consistent_solve=False,
        collect_iteration_debug=False, prp_debug=False, PRP_solid_interior_adjacency=True,
        PRP_solid_interior_adjacency_all_meshes=False, PRP_solid_interior_adjacency_normal_eps=1e-3,
        PRP_solid_interior_adjacency_dist_eps=1e-4,
        # additions of this build
        pcg_precision=PCG_MIXED, pcg_rel_tol=1e-6, pcg_inner_tol=1e-2, pcg_block_jacobi=True,
        # the single-reduction (Chronopoulos-Gear) form of the PCG loop: two launches per
        # iteration instead of three; False runs precond_cg.cpp's recurrence
        pcg_single_reduction=True,
        prp_directions=10,  # 10 = the reference's direction generator; more appends a Fibonacci set
        use_region_filter=True,  # filter_invalid_pairs / filter_broad_phase_pairs under use_untangling
    )

    def __init__(self):
        for k, v in self._DEFAULTS.items():
            setattr(self, k, v)
        self.gravity = Float3(0.0, -9.8, 0.0)
        self.floor = Float3(0.0, 0.0, 0.0)

    def __setattr__(self, k, v):
        if k in ("gravity", "floor"):
            v = _float3(v)
        object.__setattr__(self, k, v)

    def update_dt(self, dt):
        self.implicit_dt = float(dt)

    def get_substep_dt(self):
        return self.implicit_dt / max(int(self.num_substep), 1)

    def get_bending_stiffness_scaling(self):
        return float(self.stiffness_bending_ui)

    def unsupported(self):
        """The enabled features this build does not implement (empty when the step can run)."""
        out = []
        if int(self.num_substep) != 1:
            out.append("num_substep != 1")
        for k in ("use_floor", "use_untangling_ICM", "use_untangling_GIA", "use_energy_linesearch",
                  "use_static_mode"):
            if getattr(self, k):
                out.append(k)
        if int(self.PRP_direction_optimization_iterations) != 0:
            out.append("PRP_direction_optimization_iterations")
        if self.use_quasi_static_mode and int(self.qs_inertia_mode) != QuasiStaticInertiaMode.HessianOnly:
            out.append("qs_inertia_mode other than HessianOnly")
        return out


class WorldData:
    """WorldData (python_bindings.cpp:1234-1296): one registered mesh and its material."""

    def __init__(self, name, vertices, triangles, rest_vertices=None):
        self._name = str(name)
        self.vertices = np.ascontiguousarray(np.asarray(vertices, np.float32).reshape(-1, 3))
        self.triangles = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1, 3))
        self.rest_vertices = None if rest_vertices is None else np.ascontiguousarray(
            np.asarray(rest_vertices, np.float32).reshape(-1, 3))
        self.material = initializer.ClothMaterial()
        self.material_type = MaterialType.Cloth
        self.simulation_type = 0
        self.fixed_points = np.zeros(0, np.uint32)
        self.fixed_stiffness = 1.0
        self.translation, self.rotation, self.scale = Float3(), Float3(), Float3(1.0, 1.0, 1.0)
        self.registration_index = -1
        self._input = self.vertices.copy()
        self._auto_scaling = False

    # -- configuration
    def set_name(self, name):
        self._name = str(name)
        return self

    def set_simulation_type(self, t):
        self.simulation_type = int(t)
        return self

    def set_auto_scaling(self, enabled):
        """WorldData::apply_auto_scaling_if_needed (world_data.cpp:336-383): a mesh whose
        largest bounding-box span exceeds 1.5 m is scaled about the box centre to 0.5 m."""
        self._auto_scaling = bool(enabled)
        self.vertices = self._input.copy()
        if self._auto_scaling and len(self.vertices):
            lo, hi = self.vertices.min(0), self.vertices.max(0)
            span = float((hi - lo).max())
            if span > 1.5:
                c = 0.5 * (lo + hi)
                self.vertices = (c + (self.vertices - c) * np.float32(0.5 / span)).astype(np.float32)
        return self

    def set_physics_material_cloth(self, stretch_model="FEM_BW98", bending_model="DihedralAngle", thickness=1e-3,
                                   youngs_modulus=1e6, poisson_ratio=0.35, area_bending_stiffness=5e-2, d_hat=1e-3,
                                   contact_offset=0.0):
        if stretch_model not in ("FEM_BW98", "BW98"):
            raise RuntimeError(f"stretch model {stretch_model!r} is not part of this build (FEM_BW98 only)")
        if bending_model not in ("DihedralAngle", "Dihedral"):
            raise RuntimeError(f"bending model {bending_model!r} is not part of this build (DihedralAngle only)")
        self.material_type = MaterialType.Cloth
        self.material = initializer.ClothMaterial(thickness=thickness, youngs_modulus=youngs_modulus,
                                                  poisson_ratio=poisson_ratio,
                                                  area_bending_stiffness=area_bending_stiffness, d_hat=d_hat,
                                                  contact_offset=contact_offset)
        return self

    def _unsupported_material(self, kind):
        raise RuntimeError(f"{kind} materials are not part of this build (cloth only)")

    def set_physics_material_rigid(self, model="Orthogonality", thickness=1e-3, stiffness=1e6, density=1e3, mass=0.0,
                                   d_hat=1e-3, contact_offset=0.0):
        """A rigid body as one 12-DOF affine body (python_bindings.cpp:98-118, RigidMaterial
        defaults of physical_material.h:108-120); the orthogonality model only."""
        if model not in ("Orthogonality", "ABD"):
            raise RuntimeError(f"rigid model {model!r} is not part of this build (Orthogonality only)")
        if mass:
            raise RuntimeError("a prescribed rigid body mass is not part of this build (density only)")
        self.material_type = MaterialType.Rigid
        self.material = initializer.RigidMaterial(stiffness=stiffness, density=density, d_hat=d_hat,
                                                  contact_offset=contact_offset, thickness=thickness)
        return self

    def set_physics_material_tet(self, *a, **k):
        self._unsupported_material("tetrahedral")

    def set_physics_material_rod(self, *a, **k):
        self._unsupported_material("rod")

    def add_fixed_point_by_indices(self, indices, stiffness=1.0):
        self.fixed_points = np.union1d(self.fixed_points, np.asarray(indices, np.uint32)).astype(np.uint32)
        self.fixed_stiffness = float(stiffness)
        return self

    def add_fixed_point_by_method(self, *a, **k):
        raise RuntimeError("fixed points by method are not part of this build (use add_fixed_point_by_indices)")

    def set_translation(self, x, y=None, z=None):
        self.translation = _float3((x, y, z) if y is not None else x)
        return self

    def set_rotation(self, x, y=None, z=None):
        self.rotation = _float3((x, y, z) if y is not None else x)
        return self

    def set_scale(self, x, y=None, z=None):
        self.scale = _float3((x, y, z) if y is not None else (x, x, x) if np.isscalar(x) else x)
        return self

    def set_face_orientation(self, along_triangle_winding=True):
        if not along_triangle_winding:
            self.triangles = np.ascontiguousarray(self.triangles[:, [0, 2, 1]])
        return self

    # -- queries
    def get_name(self):
        return self._name

    def get_id(self):
        return self.registration_index

    def get_registration_index(self):
        return self.registration_index

    def get_fixed_point_indices(self):
        return self.fixed_points.copy()

    def get_rest_positions(self):
        return (self.vertices if self.rest_vertices is None else self.rest_vertices).copy()

    def get_rest_translation(self):
        return self.translation

    def get_rest_rotation(self):
        return self.rotation

    def get_rest_scale(self):
        return self.scale

    def world_positions(self):
        """Model positions through make_model_matrix (Core/affine_position.h:40-100):
        translate(t) * rotate(r) * scale(s), rotate(r) = X(r.x) Y(r.y) Z(r.z) with angles
        in radians and the rows those functions give (X and Y rotate by -angle in the
        textbook convention, Z by +angle)."""
        x = self.vertices.astype(np.float64) * np.array(tuple(self.scale))
        x = x @ self.rotation_matrix().T + np.array(tuple(self.translation))
        return x.astype(np.float32)

    def rotation_matrix(self):
        """rotate(r) of make_model_matrix (Core/affine_position.h:40-100)."""
        rx, ry, rz = tuple(self.rotation)
        cx, sx, cy, sy, cz, sz = np.cos(rx), np.sin(rx), np.cos(ry), np.sin(ry), np.cos(rz), np.sin(rz)
        X = np.array([[1, 0, 0], [0, cx, sx], [0, -sx, cx]])
        Y = np.array([[cy, 0, -sy], [0, 1, 0], [sy, 0, cy]])
        Z = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
        return X @ Y @ Z


class PyPRPDebugInfo:
    def __init__(self):
        self.bool_stats, self.uint_stats, self.float_stats = {}, {}, {}


class NewtonSolver:
    def __init__(self):
        self._worlds: list[WorldData] = []
        self._ctx: Context | None = None
        self._own_ctx = False
        self._handle = None
        self._config = SceneParams()
        self._setup = None
        self._dev = {}
        self._last_report = None
        self._x0 = None

    # -- scene ------------------------------------------------------------------------
    def create_world_data_from_array(self, name, vertices, triangles, rest_vertices=None):
        """WorldData of a triangle mesh.  rest_vertices (an addition of this build) gives a
        rest shape other than the registered positions."""
        return WorldData(name, vertices, triangles, rest_vertices)

    def create_world_data_from_file_path(self, name, obj_file_path):
        v, f = read_obj(obj_file_path)
        return WorldData(name, v, f)

    def register_world_data(self, wd: WorldData) -> int:
        if self._handle is not None:
            raise RuntimeError("register world data before init_solver()")
        wd.registration_index = len(self._worlds)
        self._worlds.append(wd)
        return wd.registration_index

    def num_meshes(self):
        return len(self._worlds)

    def get_mesh_names(self):
        return [w.get_name() for w in self._worlds]

    def print_registered_meshes_info(self):
        for w in self._worlds:
            print(f"mesh {w.registration_index} {w.get_name()!r}: {len(w.vertices)} vertices, "
                  f"{len(w.triangles)} triangles, {w.material_type.name}")

    def get_config(self) -> SceneParams:
        return self._config

    # -- device -----------------------------------------------------------------------
    def init_device(self, backend_name: str = "hip", binary_path: str = "", device: int = 0):
        """The reference picks a LuisaCompute backend by name; this build has one, HIP on
        gfx950 ("auto" and "cuda" select it too, other backend names raise)."""
        if backend_name not in ("hip", "cuda", "auto", "default"):
            raise RuntimeError(f"unsupported backend {backend_name!r}: this build runs on HIP (gfx950) only")
        self._ctx = Context(device)
        self._own_ctx = True

    def set_device(self, ctx: Context):
        """Borrow an existing context (the reference's set_device takes raw device / stream
        pointers, solver_interface.h:35-67)."""
        self._ctx = ctx
        self._own_ctx = False

    def cleanup_device(self):
        """Release the solver and, when owned, the context (no-op for a borrowed one)."""
        if self._handle is not None:
            lib().ut_solver_destroy(self._handle)
            self._handle = None
        if self._ctx is not None and self._own_ctx:
            self._ctx.close()
        self._ctx = None

    def get_device_ptr(self):
        return self._ctx.handle.value if self._ctx is not None else 0

    def get_stream_ptr(self):
        if self._ctx is None:
            return 0
        return lib().ut_ctx_stream(self._ctx.handle) or 0

    # -- solver -----------------------------------------------------------------------
    def _merged_mesh(self):
        xs, fs, rs, off = [], [], [], 0
        for w in self._worlds:
            xw = w.world_positions()
            xs.append(xw)
            fs.append(w.triangles.astype(np.int64) + off)
            rs.append(xw if w.rest_vertices is None else w.rest_vertices)
            off += len(xw)
        return Mesh(np.concatenate(xs), np.concatenate(fs).astype(np.uint32)), np.concatenate(rs)

    def _check_supported(self):
        bad = self._config.unsupported()
        if bad:
            raise RuntimeError("not part of this build's untangling path: " + ", ".join(bad))

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
                          c.pcg_inner_tol, 20, 0.0, 8, int(bool(c.pcg_single_reduction)))
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
        self._check_supported()
        rigid = [w.material_type == MaterialType.Rigid for w in self._worlds]
        if any(rigid):
            if not all(rigid):
                raise RuntimeError("cloth and rigid bodies in one scene are not part of this build")
            return self._init_rigid()
        mats = {tuple(vars(w.material).items()) for w in self._worlds}
        if len(mats) > 1:
            raise RuntimeError("one cloth material for all meshes in this build")
        mesh, rest = self._merged_mesh()
        mat = initializer.ClothMaterial(**vars(self._worlds[0].material))
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
        # fixed vertices: vertex property bit 0 (kPropFixed), no prediction and no motion
        fixed = np.zeros(st.nv, np.uint32)
        off = 0
        for w in self._worlds:
            fixed[off + w.fixed_points.astype(np.int64)] = 1
            off += len(w.vertices)
        d = SolverDesc()
        d.nv, d.ne, d.nf, d.nh = st.nv, len(st.edges), len(st.faces), len(st.hinges)
        for k, t in self._dev.items():
            setattr(d, k, ptr(t))
        if fixed.any():
            self._dev["vprop"] = up(fixed)
            d.vprop = ptr(self._dev["vprop"])
        else:
            d.vprop = None
        d.nslices, d.nslots = pat.nslices, pat.nslots
        self._dev["rest_x"] = up(np.asarray(rest, np.float32))
        d.rest_x = ptr(self._dev["rest_x"])
        # sa_contact_active_verts_friction_coeff: the material's friction_mu per vertex
        # (init_collision_data.cpp:188); every contact step carries lagged friction
        self._dev["friction_mu"] = up(np.full(st.nv, mat.friction_mu, np.float32))
        d.friction_mu = ptr(self._dev["friction_mu"])
        x0 = up(mesh.x)
        self._x0 = x0.clone()
        self._material_x = np.concatenate([w.vertices[:, [0, 2]] for w in self._worlds]).astype(np.float32)
        h = C.c_void_p()
        self._ctx.check(lib().ut_solver_create(self._ctx.handle, C.byref(d), C.byref(self._params()), ptr(x0),
                                               None, C.byref(h)))
        self._handle = h
        self._desc = d

    def _init_rigid(self):
        """Every world one affine body: model positions xbar = scale * vertices, q0 = (translation,
        the columns of the world rotation), the solid's mass matrix from its closed surface."""
        from .scenes import RigidScene

        mats = {tuple(vars(w.material).items()) for w in self._worlds}
        if len(mats) > 1:
            raise RuntimeError("one rigid material for all bodies in this build")
        xs, fs, body, xbar, q0, off = [], [], [], [], [], [0]
        for b, w in enumerate(self._worlds):
            model = w.vertices.astype(np.float64) * np.array(tuple(w.scale))
            R = w.rotation_matrix()
            t = np.array(tuple(w.translation), np.float64)
            xs.append(model @ R.T + t)
            fs.append(w.triangles.astype(np.int64) + off[-1])
            body.append(np.full(len(model), b, np.uint32))
            xbar.append(model)
            q0.append(np.stack([t, R[:, 0], R[:, 1], R[:, 2]]))
            off.append(off[-1] + len(model))
        mesh = Mesh(np.concatenate(xs), np.concatenate(fs).astype(np.uint32))
        scene = RigidScene(mesh, np.concatenate(body), np.concatenate(xbar).astype(np.float32), np.asarray(q0),
                           np.zeros(len(q0)), np.zeros(len(q0), np.uint8), np.asarray(off, np.uint32))
        mat = initializer.RigidMaterial(**vars(self._worlds[0].material))
        mat.d_hat = self._config.d_hat
        st = initializer.setup_bodies(scene, mat)
        self._setup = st
        dv = self._ctx.device

        def up(a, dtype=None):
            t = torch.from_numpy(np.ascontiguousarray(a))
            if t.dtype == torch.uint32:
                t = t.view(torch.int32)
            return t.to(dv) if dtype is None else t.to(dtype).to(dv)

        names = dict(edges=st.edges, faces=st.faces, face_adj_edges=st.face_adj_edges,
                     edge_adj_faces=st.edge_adj_faces, mass=st.mass, d_hat=st.d_hat, offset=st.offset,
                     vprop=st.vprop, rest_vert_area=st.rest_vert_area, rest_edge_area=st.rest_edge_area,
                     rest_face_area=st.rest_face_area, rest_x=st.x,
                     friction_mu=np.full(st.nv, mat_friction_mu(mat), np.float32))
        self._dev = {k: up(v) for k, v in names.items()}
        self._dev.update(vert_off=up(st.vert_off), body=up(st.body), xbar=up(st.xbar),
                         mass4=up(st.mass4.reshape(-1), torch.float64), volume=up(st.volume, torch.float64),
                         kappa=up(st.kappa, torch.float64), q0=up(st.q0.reshape(-1), torch.float64))
        a = AbdDesc()
        a.nb, a.nv = st.nb, st.nv
        for k in ("vert_off", "body", "xbar", "mass4", "volume", "kappa"):
            setattr(a, k, ptr(self._dev[k]))
        d = SolverDesc()
        d.nv, d.ne, d.nf, d.nh = st.nv, len(st.edges), len(st.faces), 0
        for k in names:
            setattr(d, k, ptr(self._dev[k]))
        d.abd = C.cast(C.pointer(a), C.c_void_p)
        d.q0 = ptr(self._dev["q0"])
        x0 = up(st.x)
        self._x0 = x0.clone()
        self._q0 = self._dev["q0"].clone()
        self._material_x = np.zeros((st.nv, 2), np.float32)
        h = C.c_void_p()
        self._ctx.check(lib().ut_solver_create(self._ctx.handle, C.byref(d), C.byref(self._params()), ptr(x0),
                                               None, C.byref(h)))
        self._handle = h
        self._desc, self._abd_desc = d, a
        self._rigid = True

    def q(self):
        """Affine bodies: the device q [nb, 12] (p, a1, a2, a3 per body)."""
        self._require()
        out = torch.empty(12 * self._setup.nb, dtype=torch.float64, device=self._ctx.device)
        self._ctx.check(lib().ut_solver_get_q(self._handle, ptr(out), None))
        return out.view(-1, 12)

    def set_q(self, q, vq=None):
        self._require()
        qq = q.detach().to(torch.float64).reshape(-1).contiguous()
        vv = torch.zeros_like(qq) if vq is None else vq.detach().to(torch.float64).reshape(-1).contiguous()
        self._ctx.check(lib().ut_solver_set_q(self._handle, ptr(qq), ptr(vv)))

    def __del__(self):
        try:
            if self._handle:
                lib().ut_solver_destroy(self._handle)
                self._handle = None
        except Exception:
            pass

    def _require(self):
        if self._handle is None:
            raise RuntimeError("Solver not initialized. Call init_solver() first.")

    def physics_step_gpu(self):
        """One physics_step_GPU (newton_solver.cpp:2482-2767); returns this build's step
        report (counts, PCG residual, stage times)."""
        self._require()
        self._check_supported()
        self._ctx.check(lib().ut_solver_set_params(self._handle, C.byref(self._params())))
        r = StepReport()
        self._ctx.check(lib().ut_solver_step(self._handle, C.byref(r)))
        self._last_report = {k: (list(getattr(r, k)) if k == "energy" else getattr(r, k)) for k, _ in
                             StepReport._fields_}
        self._config.current_frame += 1
        return self._last_report

    def physics_step_cpu(self):
        """The reference's CPU physics step; this build has no CPU physics, so it runs the
        GPU step (same results as physics_step_gpu)."""
        return self.physics_step_gpu()

    def restart_system(self):
        """Positions back to the registered ones, velocities zero."""
        self._require()
        if getattr(self, "_rigid", False):
            self.set_q(self._q0)
        else:
            self.set_state(self._x0, torch.zeros_like(self._x0))
        self._last_report = None
        self._config.current_frame = 0

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

    def get_step_report(self):
        return self._last_report

    # -- state on the device ----------------------------------------------------------
    def positions(self, which=0) -> torch.Tensor:
        """Device positions [V, 3]: 0 current, 1 start of the last Newton iteration, 2 start
        of the last step."""
        self._require()
        x = torch.empty((self._setup.nv, 3), dtype=torch.float32, device=self._ctx.device)
        self._ctx.check(lib().ut_solver_get_positions(self._handle, int(which), ptr(x)))
        return x

    def velocities(self) -> torch.Tensor:
        v = torch.empty((self._setup.nv, 3), dtype=torch.float32, device=self._ctx.device)
        self._ctx.check(lib().ut_solver_get_state(self._handle, None, ptr(v)))
        return v

    def set_state(self, x: torch.Tensor, v: torch.Tensor | None = None):
        if getattr(self, "_rigid", False):
            raise RuntimeError("affine bodies: set the body state with set_q (positions follow q)")
        self._ctx.check(lib().ut_solver_set_state(self._handle, ptr(x.contiguous()),
                                                  None if v is None else ptr(v.contiguous())))

    def set_positions(self, x: torch.Tensor):
        self._ctx.check(lib().ut_solver_set_state(self._handle, ptr(x.contiguous()), None))

    # -- results (python_bindings.cpp:485-1206) -----------------------------------------
    def _split(self, x):
        out, off = [], 0
        for w in self._worlds:
            n = len(w.vertices)
            out.append(np.ascontiguousarray(x[off:off + n]))
            off += n
        return out

    def get_sim_result(self):
        """(list of float32[N,3] per world, list of uint32[M,3] per world) -- fresh host copies."""
        self._require()
        return self._split(self.positions().cpu().numpy()), [w.triangles.copy() for w in self._worlds]

    def get_object_sim_result_by_registration_id(self, registration_id):
        v, f = self.get_sim_result()
        return v[registration_id], f[registration_id]

    def get_object_by_registration_id(self, registration_id):
        return self._worlds[registration_id]

    def set_object_vertex_positions(self, registration_id, positions):
        self._require()
        x = self.positions()
        off = sum(len(w.vertices) for w in self._worlds[:registration_id])
        n = len(self._worlds[registration_id].vertices)
        x[off:off + n] = torch.as_tensor(np.asarray(positions, np.float32).reshape(n, 3), device=x.device)
        self.set_positions(x)

    def get_surface_edges(self):
        out = []
        for w in self._worlds:
            out.append(edges_from_faces(w.triangles).astype(np.uint32))
        return out

    def query_local_vid_from_global_vid(self, global_vid):
        return int(global_vid) - self._prefix(self.query_registration_id_from_global_vid(global_vid))

    def query_registration_id_from_global_vid(self, global_vid):
        off = 0
        for w in self._worlds:
            if global_vid < off + len(w.vertices):
                return w.registration_index
            off += len(w.vertices)
        raise RuntimeError(f"global vertex {global_vid} out of range")

    def query_global_vid_from_registration_id_and_local_vid(self, registration_id, local_vid):
        return self._prefix(registration_id) + int(local_vid)

    def _prefix(self, registration_id):
        return sum(len(w.vertices) for w in self._worlds[:registration_id])

    def _contour_get(self, what, shape):
        t = torch.empty(shape, dtype=torch.int32, device=self._ctx.device)
        if t.numel():
            self._ctx.check(lib().ut_solver_contour_get(self._handle, what, ptr(t)))
        return t.cpu().numpy().view(np.uint32)

    def get_intersection_contour_data(self):
        """The EF pairs and contours of the last step's detection (at its Newton iteration
        start), with the keys of python_bindings.cpp:635-829.

        narrow_phase_list_ef uint32[P, 5] (edge vertices, face vertices; loop-vertex pairs
        appended), ef_pair_contour_index / ef_pair_mesh_index uint32[P], ef_pair_pos_3d
        float32[P, 3] (the edge point at the iteration start), ef_pair_pos_2d float32[P, 4]
        (material coordinates (x, z) of the model positions at the edge point and the face
        point, init_mesh_data.cpp:186-187), ef_pair_adj_pairs_ext_prefix / _data (CSR of
        the EF adjacency), intersection_contours_prefix / _data (each contour's pairs,
        ascending), contour_types uint32[C] (ContourType), num_pairs and num_contours
        (uint32[1]), and the disabled projected-winding debug fields as empty arrays.
        Additions: ef_indices uint32[P, 2] (edge id, face id) and contour_mesh_is_closed
        uint32[C, 2]."""
        self._require()
        r = self._last_report or {}
        n = int(r.get("ef_total", 0))
        nc = int(r.get("contours", 0)) if n else 0
        pairs = self._contour_get(CT_PAIRS, (n, 12))
        ids = self._contour_get(CT_PAIR_IDS, (n, 2))
        ci = self._contour_get(CT_CONTOUR_INDEX, (n,))
        mi = self._contour_get(CT_MESH_INDEX, (n,))
        adj_off = self._contour_get(CT_ADJ_OFF, (n + 1,)) if n else np.zeros(1, np.uint32)
        adj = self._contour_get(CT_ADJ, (int(adj_off[-1]),)) if n else np.zeros(0, np.uint32)
        ctype = self._contour_get(CT_CONTOUR_TYPE, (nc,))
        closed = self._contour_get(CT_MESH_CLOSED, (nc, 2))
        ef5 = pairs[:, :5].copy()
        bary = pairs[:, 5:8].view(np.float32)
        wb = np.stack([bary[:, 0], 1.0 - bary[:, 0]], 1)
        wf = np.stack([bary[:, 1], bary[:, 2], 1.0 - bary[:, 1] - bary[:, 2]], 1)
        if n:
            x0 = self.positions(1).cpu().numpy()
            pos3 = (x0[ef5[:, 0]] * wb[:, :1] + x0[ef5[:, 1]] * wb[:, 1:]).astype(np.float32)
            m = self._material_x
            uv_e = m[ef5[:, 0]] * wb[:, :1] + m[ef5[:, 1]] * wb[:, 1:]
            uv_f = m[ef5[:, 2]] * wf[:, :1] + m[ef5[:, 3]] * wf[:, 1:2] + m[ef5[:, 4]] * wf[:, 2:]
            pos2 = np.concatenate([uv_e, uv_f], 1).astype(np.float32)
        else:
            pos3, pos2 = np.zeros((0, 3), np.float32), np.zeros((0, 4), np.float32)
        order = np.argsort(ci, kind="stable")
        cprefix = np.zeros(nc + 1, np.uint32)
        if nc:
            cprefix[1:] = np.cumsum(np.bincount(ci, minlength=nc))
        out = {
            "narrow_phase_list_ef": ef5, "ef_pair_contour_index": ci, "ef_pair_mesh_index": mi,
            "ef_pair_pos_3d": pos3, "ef_pair_pos_2d": pos2,
            "ef_pair_adj_pairs_ext_prefix": adj_off.astype(np.uint32), "ef_pair_adj_pairs_ext_data": adj,
            "intersection_contours_prefix": cprefix, "intersection_contours_data": order.astype(np.uint32),
            "contour_types": ctype, "num_pairs": np.array([n], np.uint32),
            "num_contours": np.array([nc], np.uint32),
            "ef_indices": ids, "contour_mesh_is_closed": closed,
        }
        none = np.array([0xFFFFFFFF], np.uint32)
        out["prp_debug_uv_contour_idx"] = none
        out["prp_debug_uv_combo_idx"] = none.copy()
        for side in ("0", "1"):
            out[f"prp_debug_uv_side{side}_vertex_ids"] = np.zeros(0, np.uint32)
            out[f"prp_debug_uv_side{side}_vertex_uv"] = np.zeros((0, 2), np.float32)
            out[f"prp_debug_uv_side{side}_vertex_has_uv"] = np.zeros(0, np.uint8)
            out[f"prp_debug_uv_side{side}_vertex_is_boundary"] = np.zeros(0, np.uint8)
            out[f"prp_debug_uv_side{side}_faces"] = np.zeros((0, 3), np.uint32)
            out[f"prp_debug_uv_side{side}_segments"] = np.zeros((0, 4), np.float32)
        for k in ("src_uv", "dst_uv"):
            out[f"prp_debug_uv_hit_{k}"] = np.zeros((0, 2), np.float32)
        for k in ("src_has_uv", "dst_has_uv", "is_valid"):
            out[f"prp_debug_uv_hit_{k}"] = np.zeros(0, np.uint8)
        for k in ("src_key", "dst_key", "winding_key"):
            out[f"prp_debug_uv_hit_{k}"] = np.zeros(0, np.uint32)
        return out

    def get_response_pair_data(self):
        """The PRP response pairs of the last Newton iteration with the keys of
        python_bindings.cpp:830-946: indices uint32[R, 4] (type bit masked), |weights|
        float32[R, 4], collision type (4 VF, 8 EE), normal, stiffness (k1, k2), area, the
        points src / dst at the iteration start, the index quadruples and the counts."""
        self._require()
        n = C.c_uint32()
        self._ctx.check(lib().ut_solver_responses(self._handle, None, None, 0, C.byref(n)))
        R = n.value
        rec = torch.empty((max(R, 1), 16), dtype=torch.int32, device=self._ctx.device)
        meta = torch.empty((max(R, 1), 4), dtype=torch.int32, device=self._ctx.device)
        if R:
            self._ctx.check(lib().ut_solver_responses(self._handle, ptr(rec), ptr(meta), R, C.byref(n)))
        rec = rec.cpu().numpy().view(np.uint32)[:R]
        meta = meta.cpu().numpy().view(np.uint32)[:R]
        idx = rec[:, :4] & np.uint32(0x7FFFFFFF)
        f = rec[:, 4:12].view(np.float32)
        vf = (rec[:, 2] & np.uint32(0x80000000)) != 0
        b0, b1 = f[:, 6], f[:, 7]
        w = np.where(vf[:, None], np.stack([np.ones_like(b0), -b0, -b1, b0 + b1 - 1.0], 1),
                     np.stack([b0, 1.0 - b0, -b1, b1 - 1.0], 1)).astype(np.float32)
        aw = np.abs(w)
        x0 = self.positions(1).cpu().numpy() if R else np.zeros((1, 3), np.float32)
        P = x0[idx] if R else np.zeros((0, 4, 3), np.float32)
        src = np.where(vf[:, None], P[:, 0], aw[:, 0:1] * P[:, 0] + aw[:, 1:2] * P[:, 1])
        dst = np.where(vf[:, None], aw[:, 1:2] * P[:, 1] + aw[:, 2:3] * P[:, 2] + aw[:, 3:4] * P[:, 3],
                       aw[:, 2:3] * P[:, 2] + aw[:, 3:4] * P[:, 3])
        return {
            "target_point_template_pair_indices": idx,
            "target_point_template_pair_weights": aw,
            "target_point_template_pair_collision_type": np.where(vf, 4, 8).astype(np.uint32),
            "target_point_template_pair_normal": f[:, 0:3].copy(),
            "target_point_template_pair_stiff": f[:, 4:6].copy(),
            "target_point_template_pair_area": f[:, 3].copy(),
            "target_point_template_pair_src": src.astype(np.float32),
            "target_point_template_pair_dst": dst.astype(np.float32),
            "target_point_template_pairs_indices": meta,
            "num_template_pairs": np.array([R], np.uint32),
            "num_template_pairs_indices": np.array([R], np.uint32),
        }

    def get_prp_debug_info(self):
        """PRP counters of the last step (the reference fills its maps only with prp_debug)."""
        info = PyPRPDebugInfo()
        r = self._last_report or {}
        for k in ("prp_contours", "prp_levels", "prp_slots", "prp_hits", "prp_responses", "prp_rays_vf",
                  "prp_rays_ee", "prp_rays_fv"):
            if k in r:
                info.uint_stats["prp." + k[4:]] = int(r[k])
        if "ms_prp" in r:
            info.float_stats["prp.ms"] = float(r["ms_prp"])
        info.bool_stats["prp.gpu"] = True
        return info

    # -- state files (solver_interface.cpp:772-803, :969-1260) ---------------------------
    def save_sim_result(self, obj_path):
        """save_mesh_to_obj: "o mesh_k" per mesh, its vertices, its faces (1-based, global)."""
        self._require()
        x = self.positions().cpu().numpy()
        _mkdir(obj_path)
        lines, off = ["# Simulated Reulst"], 0
        for k, w in enumerate(self._worlds):
            n = len(w.vertices)
            lines.append(f"o mesh_{k}")
            lines += [f"v {_g(a)} {_g(b)} {_g(c)}" for a, b, c in x[off:off + n]]
            lines += [f"f {a + off + 1} {b + off + 1} {c + off + 1}" for a, b, c in w.triangles.astype(np.int64)]
            off += n
        pathlib.Path(obj_path).write_text("\n".join(lines) + "\n")

    def save_current_state(self, state_path):
        """save_current_frame_state_file: "o State" and one "v x y z" line per degree of
        freedom (cloth: per vertex), then "o Velocity" and the velocities, at the stream's
        default precision (6 significant digits)."""
        self._require()
        _mkdir(state_path)
        pathlib.Path(state_path).write_text(state_text(self.positions().cpu().numpy(),
                                                       self.velocities().cpu().numpy()))

    def load_target_state(self, state_path):
        """load_state_from_state_file: positions and velocities from a .state file (a file
        with more entries than degrees of freedom stops the load, as the reference does)."""
        self._require()
        q, qv = parse_state_text(pathlib.Path(state_path).read_text())
        nv = self._setup.nv
        if len(q) > nv or len(qv) > nv:
            return
        x = self.positions().cpu().numpy()
        v = self.velocities().cpu().numpy()
        if q:
            x[:len(q)] = np.asarray(q, np.float32)
        if qv:
            v[:len(qv)] = np.asarray(qv, np.float32)
        self.set_state(torch.from_numpy(x).to(self._ctx.device), torch.from_numpy(v).to(self._ctx.device))

    def load_target_state_from_obj(self, obj_path):
        """load_state_from_obj_file: positions from an OBJ with one "v" line per vertex."""
        self._require()
        v, _ = read_obj(obj_path)
        if len(v) != self._setup.nv:
            raise RuntimeError(f"Loaded vertex count {len(v)} does not match the number of vertices "
                               f"{self._setup.nv} in the sim data.")
        self.set_positions(torch.from_numpy(np.asarray(v, np.float32)).to(self._ctx.device))


def _g(a):
    """C++ ostream default formatting of a float (%g, 6 significant digits)."""
    return "%g" % float(a)


def state_text(x, v):
    """The .state file of positions x[V, 3] and velocities v[V, 3]."""
    lines = ["o State"] + [f"v {_g(a)} {_g(b)} {_g(c)}" for a, b, c in np.asarray(x).reshape(-1, 3)]
    lines += ["o Velocity"] + [f"v {_g(a)} {_g(b)} {_g(c)}" for a, b, c in np.asarray(v).reshape(-1, 3)]
    return "\n".join(lines) + "\n"


def parse_state_text(text):
    """(positions, velocities) lists of a .state file: "v" lines after "o State" and
    after "o Velocity"; other lines are skipped."""
    q, qv, sec = [], [], None
    for line in text.splitlines():
        if not line:
            continue
        if line.startswith("o State"):
            sec = q
            continue
        if line.startswith("o Velocity"):
            sec = qv
            continue
        if line[0] == "v" and sec is not None:
            sec.append([float(t) for t in line[1:].split()[:3]])
    return q, qv


def _mkdir(path):
    d = os.path.dirname(os.fspath(path))
    if d:
        os.makedirs(d, exist_ok=True)


def read_obj(path):
    """Vertices ("v x y z") and triangles ("f a b c", 1-based, optional /vt/vn parts) of an OBJ."""
    vs, fs = [], []
    for line in pathlib.Path(path).read_text().splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            vs.append([float(a) for a in t[1:4]])
        elif t[0] == "f":
            idx = [int(a.split("/")[0]) - 1 for a in t[1:]]
            for k in range(1, len(idx) - 1):
                fs.append([idx[0], idx[k], idx[k + 1]])
    return np.asarray(vs, np.float32).reshape(-1, 3), np.asarray(fs, np.uint32).reshape(-1, 3)
