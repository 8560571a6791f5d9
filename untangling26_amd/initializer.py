"""Host-side initializer: the static arrays the per-step path reads.

Restates the parts of Solver/Initializer/{init_mesh_data,init_sim_data,
init_collision_data}.cpp that the hot path consumes (the initializer itself is
outside the per-step path): mesh adjacency, rest areas, lumped masses, BW98
inverse rest matrices and stiffness, bending hinges with rest angles, per-vertex
d_hat / contact offset, and the SELL-64 block pattern with per-element slots.
Everything is numpy; `to_device` uploads it for ut_solver_create.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import bsr, topology
from .scenes import Mesh, edges_from_faces

NONE = 0xFFFFFFFF


@dataclass
class ClothMaterial:
    """ClothMaterial defaults of SimulationCore/physical_material.h:83-96."""
    thickness: float = 1e-3
    youngs_modulus: float = 1e6
    poisson_ratio: float = 0.35
    area_bending_stiffness: float = 5e-2
    density: float = 1e3
    d_hat: float = 1e-3
    contact_offset: float = 0.0
    friction_mu: float = 0.5  # MaterialBase::default_friction_mu (physical_material.h:66)


def convert_prop(young: float, poisson: float):
    """StretchEnergy::convert_prop (Energy/stretch_energy.h:417-422): (mu, lambda)."""
    mu = young / (2.0 * (1.0 + poisson))
    lam = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson))
    return mu, lam


def dm_inv(x: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """get_Dm_inv (stretch_energy.h:423-438) per face, as (d0, d1, d2, d3) with Dm = [[d0, d2], [d1, d3]]."""
    x = np.asarray(x, np.float64)
    f = np.asarray(faces, np.int64)
    x0, x1, x2 = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    r1, r2 = x1 - x0, x2 - x0
    cr = np.cross(r1, r2)
    a1 = r1 / np.linalg.norm(r1, axis=1, keepdims=True)
    a2 = np.cross(cr, a1)
    a2 /= np.linalg.norm(a2, axis=1, keepdims=True)
    uv = [np.stack([(a1 * p).sum(1), (a2 * p).sum(1)], 1) for p in (x0, x1, x2)]
    du0, du1 = uv[1] - uv[0], uv[2] - uv[0]
    # duv has columns du0, du1; inverse of [[a, b], [c, d]] = [[d, -b], [-c, a]] / det
    a, b, c, d = du0[:, 0], du1[:, 0], du0[:, 1], du1[:, 1]
    det = a * d - b * c
    # inverse [[i00, i01], [i10, i11]] = [[d, -b], [-c, a]] / det, stored as (i00, i10, i01, i11)
    return np.stack([d / det, -c / det, -b / det, a / det], 1).astype(np.float32)


def dihedral_angles(x: np.ndarray, hinges: np.ndarray) -> np.ndarray:
    """Dihedral angle of hinges (e0, e1, o0, o1), same formula as the GPU kernel."""
    x = np.asarray(x, np.float64)
    h = np.asarray(hinges, np.int64)
    e0, e1, o0, o1 = (x[h[:, k]] for k in range(4))
    n1 = np.cross(e1 - o0, e0 - o0)
    n2 = np.cross(e0 - o1, e1 - o1)
    den = np.maximum(np.sqrt((n1 * n1).sum(1) * (n2 * n2).sum(1)), 1e-12)
    c = np.clip((n1 * n2).sum(1) / den, -1.0, 1.0)
    s = np.sign((np.cross(n2, n1) * (e1 - e0)).sum(1))
    return (s * np.arccos(c)).astype(np.float32)


@dataclass
class ClothSetup:
    nv: int
    x: np.ndarray
    edges: np.ndarray
    faces: np.ndarray
    face_adj_edges: np.ndarray
    edge_adj_faces: np.ndarray
    mass: np.ndarray
    d_hat: np.ndarray
    offset: np.ndarray
    rest_vert_area: np.ndarray
    rest_edge_area: np.ndarray
    rest_face_area: np.ndarray
    dm: np.ndarray
    mu_lambda: np.ndarray
    hinges: np.ndarray
    rest_angle: np.ndarray
    bend_stiff_area: np.ndarray
    pattern: bsr.SellPattern
    face_slots: np.ndarray
    hinge_slots: np.ndarray
    extra: dict = field(default_factory=dict)


def setup_cloth(mesh: Mesh, mat: ClothMaterial | None = None, x_rest: np.ndarray | None = None) -> ClothSetup:
    mat = mat or ClothMaterial()
    faces = np.ascontiguousarray(mesh.faces, np.uint32)
    edges = mesh.edges if mesh.edges is not None else edges_from_faces(faces)
    x = np.ascontiguousarray(mesh.x, np.float32)
    xr = x if x_rest is None else np.asarray(x_rest, np.float32)
    fae, eaf = topology.mesh_topology(edges, faces)
    va, ea, fa = topology.rest_areas(xr, edges, faces, eaf)
    areal_density = mat.density * mat.thickness
    mass = (va * areal_density).astype(np.float32)
    mu, lam = convert_prop(mat.youngs_modulus, mat.poisson_ratio)
    mu_lambda = np.tile(np.array([mu * mat.thickness, lam * mat.thickness], np.float32), (len(faces), 1))
    interior = (eaf[:, 0] != NONE) & (eaf[:, 1] != NONE)
    e = edges[interior].astype(np.int64)
    fi = faces.astype(np.int64)
    o0 = fi[eaf[interior, 0]].sum(1) - e.sum(1)
    o1 = fi[eaf[interior, 1]].sum(1) - e.sum(1)
    hinges = np.stack([e[:, 0], e[:, 1], o0, o1], 1).astype(np.uint32)
    rest_angle = dihedral_angles(xr, hinges)
    bend = (mat.area_bending_stiffness * ea[interior]).astype(np.float32)
    pattern = bsr.cloth_pattern(len(x), faces, edges, eaf)
    f = faces.astype(np.int64)
    face_slots = pattern.slots(np.repeat(f, 3, axis=1).ravel(), np.tile(f, (1, 3)).ravel()).reshape(-1, 9)
    hh = hinges.astype(np.int64)
    hinge_slots = pattern.slots(np.repeat(hh, 4, axis=1).ravel(), np.tile(hh, (1, 4)).ravel()).reshape(-1, 16)
    return ClothSetup(
        nv=len(x), x=x, edges=np.ascontiguousarray(edges, np.uint32), faces=faces, face_adj_edges=fae,
        edge_adj_faces=eaf, mass=mass, d_hat=np.full(len(x), mat.d_hat, np.float32),
        offset=np.full(len(x), 0.5 * mat.contact_offset, np.float32), rest_vert_area=va, rest_edge_area=ea,
        rest_face_area=fa, dm=dm_inv(xr, faces), mu_lambda=mu_lambda, hinges=hinges, rest_angle=rest_angle,
        bend_stiff_area=bend, pattern=pattern, face_slots=face_slots.astype(np.uint32),
        hinge_slots=hinge_slots.astype(np.uint32))


# ---- affine bodies -------------------------------------------------------------------
@dataclass
class RigidMaterial:
    """RigidMaterial fields used by the affine-body path: orthogonality stiffness
    (RigidMaterial::stiffness), density, contact d_hat and offset."""
    stiffness: float = 1e6
    density: float = 1e3
    d_hat: float = 1e-3
    contact_offset: float = 0.0
    thickness: float = 1e-3  # lumped vertex mass of the PRP displacement cost


def dyadic_mass(xbar: np.ndarray, faces: np.ndarray, density: float):
    """Mass, first moment and second moment (dyadic) of the solid bounded by a closed,
    outward-oriented triangle mesh (compute_trimesh_dyadic_mass, init_sim_data.cpp:988-997):
    sums over the tetrahedra (0, a, b, c) with det = a . (b x c):
    M = rho sum det / 6, MI = rho sum det (a + b + c) / 24,
    I = rho sum det (a a^T + b b^T + c c^T + s s^T) / 120, s = a + b + c."""
    x = np.asarray(xbar, np.float64)
    f = np.asarray(faces, np.int64)
    a, b, c = x[f[:, 0]], x[f[:, 1]], x[f[:, 2]]
    det = np.einsum("ij,ij->i", a, np.cross(b, c))
    s = a + b + c
    M = density * det.sum() / 6.0
    MI = density * (det[:, None] * s).sum(0) / 24.0
    outer = lambda u: u[:, :, None] * u[:, None, :]
    I = density * (det[:, None, None] * (outer(a) + outer(b) + outer(c) + outer(s))).sum(0) / 120.0
    return M, MI, I


def mass4(M, MI, I, eps=1.1920929e-7):
    """The compressed 4x4 mass matrix [[M, MI^T], [MI, I]] with its diagonal clamped
    to Epsilon (init_sim_data.cpp:1000-1021)."""
    m = np.zeros((4, 4))
    m[0, 0] = M
    m[0, 1:] = m[1:, 0] = MI
    m[1:, 1:] = I
    d = np.maximum(np.diag(m), eps)
    np.fill_diagonal(m, d)
    return m


@dataclass
class BodySetup:
    nb: int
    nv: int
    vert_off: np.ndarray  # uint32 [nb + 1]
    body: np.ndarray  # uint32 [nv]
    xbar: np.ndarray  # float32 [nv, 3]
    mass4: np.ndarray  # float64 [nb, 4, 4]
    volume: np.ndarray  # float64 [nb]
    kappa: np.ndarray  # float64 [nb]
    q0: np.ndarray  # float64 [nb, 12]
    # collision data of the vertex mesh
    x: np.ndarray
    edges: np.ndarray
    faces: np.ndarray
    face_adj_edges: np.ndarray
    edge_adj_faces: np.ndarray
    mass: np.ndarray
    d_hat: np.ndarray
    offset: np.ndarray
    vprop: np.ndarray
    rest_vert_area: np.ndarray
    rest_edge_area: np.ndarray
    rest_face_area: np.ndarray


def setup_bodies(scene, mat: RigidMaterial | None = None) -> BodySetup:
    """Affine-body data of a scenes.RigidScene: per body the compressed mass matrix of
    its solid (surface integration), its volume and orthogonality stiffness; per vertex
    its body, model position and the rigid property word (kPropRigid | body << 8: the
    DCD skips pairs inside one body)."""
    mat = mat or RigidMaterial()
    m = scene.mesh
    nb = len(scene.vert_off) - 1
    faces = np.ascontiguousarray(m.faces, np.uint32)
    edges = np.ascontiguousarray(m.edges, np.uint32)
    fae, eaf = topology.mesh_topology(edges, faces)
    va, ea, fa = topology.rest_areas(m.x, edges, faces, eaf)
    M4 = np.zeros((nb, 4, 4))
    vol = np.zeros(nb)
    fb = scene.body[faces[:, 0]]
    order = np.argsort(fb, kind="stable")
    foff = np.searchsorted(fb[order], np.arange(nb + 1))
    for b in range(nb):
        fsel = faces[order[foff[b]:foff[b + 1]]]
        Mb, MI, I = dyadic_mass(scene.xbar, fsel, mat.density)
        M4[b] = mass4(Mb, MI, I)
        vol[b] = Mb / mat.density
    q0 = scene.q0.reshape(nb, 12)
    vprop = (np.uint32(2) | (scene.body.astype(np.uint32) << np.uint32(8))).astype(np.uint32)
    return BodySetup(
        nb=nb, nv=m.nv, vert_off=np.asarray(scene.vert_off, np.uint32), body=np.asarray(scene.body, np.uint32),
        xbar=np.ascontiguousarray(scene.xbar, np.float32), mass4=M4, volume=vol, kappa=np.full(nb, mat.stiffness),
        q0=np.ascontiguousarray(q0, np.float64), x=np.ascontiguousarray(m.x, np.float32), edges=edges, faces=faces,
        face_adj_edges=fae, edge_adj_faces=eaf, mass=(va * mat.density * mat.thickness).astype(np.float32),
        d_hat=np.full(m.nv, mat.d_hat, np.float32), offset=np.full(m.nv, 0.5 * mat.contact_offset, np.float32),
        vprop=vprop, rest_vert_area=va, rest_edge_area=ea, rest_face_area=fa)
