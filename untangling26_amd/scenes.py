"""Seeded synthetic scenes for the untangling workloads (numpy, host side).

World is y-up (the reference's gravity is (0, -9.8, 0), scene_params.h:131),
so cloth sheets lie in the x-z plane and "vertical" means y.

  two_sheets     two 32x32-vertex 1 m sheets, B rotated by U[10,30] deg about a
                 random horizontal axis and cutting A with max penetration 0.05 m
  folded_cloth   256x256-vertex 2 m sheet folded about its midline; the folded
                 half is displaced by 0.08 sin(2 pi u) sin(pi v) + N(0, 0.005)
  nested_layers  8 stacked 128x128-vertex sheets, 4 mm apart, vertex jitter
                 N(0, 6 mm) and a per-layer in-plane shift U(-0.05, 0.05) m
  draped_cloth   a 64x64-vertex 1 m sheet draped over a level-5 icosphere of
                 radius 0.3 m, plus seeded states perturbed by N(0, 1 mm) with
                 every cloth-sphere gap kept in (0.3, 0.9) mm
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    x: np.ndarray  # float32 [V, 3]
    faces: np.ndarray  # uint32 [F, 3]
    edges: np.ndarray = field(default=None)  # uint32 [E, 2]
    name: str = ""

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.float32)
        self.faces = np.ascontiguousarray(self.faces, dtype=np.uint32)
        if self.edges is None:
            self.edges = edges_from_faces(self.faces)

    @property
    def nv(self):
        return self.x.shape[0]

    @property
    def nf(self):
        return self.faces.shape[0]

    @property
    def ne(self):
        return self.edges.shape[0]


def grid_faces(nx: int, ny: int, offset: int = 0) -> np.ndarray:
    """Two triangles per quad of an nx-by-ny vertex grid (row-major, index i*ny+j)."""
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    a = (i * ny + j).ravel()
    b = a + ny
    c = a + 1
    d = b + 1
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([c, b, d], 1)], 0)
    # interleave the two triangles of each quad for locality
    f = f.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3)
    return (f + offset).astype(np.uint32)


def edges_from_faces(faces: np.ndarray) -> np.ndarray:
    """Unique undirected edges (min, max), sorted lexicographically."""
    f = faces.astype(np.int64)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], 0)
    e.sort(axis=1)
    m = int(f.max()) + 1
    key = np.unique(e[:, 0] * m + e[:, 1])
    return np.stack([key // m, key % m], 1).astype(np.uint32)


def grid_sheet(nx: int, ny: int, size_x: float, size_y: float) -> np.ndarray:
    u = np.linspace(0.0, 1.0, nx)
    v = np.linspace(0.0, 1.0, ny)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = np.stack([uu * size_x, np.zeros_like(uu), vv * size_y], -1).reshape(-1, 3)
    return x


def _rotation(axis: np.ndarray, angle: float) -> np.ndarray:
    axis = axis / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * k + (1 - np.cos(angle)) * (k @ k)


def two_sheets(seed: int = 0, n: int = 32, side: float = 1.0, depth: float = 0.05) -> Mesh:
    rng = np.random.default_rng(seed)
    a = grid_sheet(n, n, side, side) - np.array([side / 2, 0.0, side / 2])
    theta = np.deg2rad(rng.uniform(10.0, 30.0))
    phi = rng.uniform(0.0, 2 * np.pi)
    axis = np.array([np.cos(phi), 0.0, np.sin(phi)])
    b = a @ _rotation(axis, theta).T
    b[:, 1] += -depth - b[:, 1].min()
    x = np.concatenate([a, b], 0)
    faces = np.concatenate([grid_faces(n, n, 0), grid_faces(n, n, n * n)], 0)
    return Mesh(x, faces, name=f"two_sheets_{n}x{n}")


def folded_cloth(seed: int = 0, n: int = 256, side: float = 2.0, amp: float = 0.08, noise: float = 0.005,
                 shift: float = 0.25) -> Mesh:
    """Sheet folded about its midline; the folded half carries the figure-eight
    height field plus noise.  `shift` (in grid cells) offsets the folded half in
    the plane: without it the mirrored grid projects exactly onto the lower
    grid, every crossing lies on an edge or vertex of the other layer and the
    EF filters (intersection_resolver.cpp:160-181) reject most of them."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n)
    v = np.linspace(0.0, 1.0, n)
    uu, vv = np.meshgrid(u, v, indexing="ij")
    x = np.empty((n, n, 3))
    folded = uu > 0.5
    cell = side / (n - 1)
    x[..., 0] = np.where(folded, (1.0 - uu) * side + shift * cell, uu * side)
    x[..., 2] = vv * side + np.where(folded, shift * cell, 0.0)
    up = 2.0 * uu - 1.0  # coordinate along the folded half, 0 at the fold
    h = amp * np.sin(2 * np.pi * up) * np.sin(np.pi * vv) + rng.normal(0.0, noise, size=uu.shape)
    x[..., 1] = np.where(folded, h, 0.0)
    return Mesh(x.reshape(-1, 3), grid_faces(n, n), name=f"folded_cloth_{n}x{n}")


def nested_layers(seed: int = 0, layers: int = 8, n: int = 128, side: float = 1.0, spacing: float = 0.004,
                  jitter: float = 0.006, shift: float = 0.05) -> Mesh:
    rng = np.random.default_rng(seed)
    xs, fs = [], []
    for k in range(layers):
        x = grid_sheet(n, n, side, side)
        x[:, 1] = k * spacing + rng.normal(0.0, jitter, size=x.shape[0])
        d = rng.uniform(-shift, shift, size=2)
        x[:, 0] += d[0]
        x[:, 2] += d[1]
        xs.append(x)
        fs.append(grid_faces(n, n, k * n * n))
    return Mesh(np.concatenate(xs, 0), np.concatenate(fs, 0), name=f"nested_{layers}x{n}x{n}")


def folded_cloth_rest(n: int = 256, side: float = 2.0) -> np.ndarray:
    """Rest shape of folded_cloth: the flat, unfolded sheet (same vertex order)."""
    return grid_sheet(n, n, side, side).astype(np.float32)


def nested_layers_rest(layers: int = 8, n: int = 128, side: float = 1.0, spacing: float = 0.004) -> np.ndarray:
    xs = []
    for k in range(layers):
        x = grid_sheet(n, n, side, side)
        x[:, 1] = k * spacing
        xs.append(x)
    return np.concatenate(xs, 0).astype(np.float32)


def rest_shape(name: str, mesh: Mesh, **kw) -> np.ndarray:
    """Rest positions of a named scene (two_sheets is at rest as generated)."""
    if name == "folded_cloth":
        return folded_cloth_rest(**kw)
    if name == "nested_layers":
        return nested_layers_rest(**kw)
    return mesh.x.copy()


def per_vertex(mesh: Mesh, d_hat: float = 1e-3, thickness: float = 0.0):
    """Per-vertex d_hat and contact offset arrays (init_collision_data.cpp:119-240)."""
    return (np.full(mesh.nv, d_hat, dtype=np.float32), np.full(mesh.nv, thickness, dtype=np.float32))


CONFIGS = {
    "two_sheets": two_sheets,
    "folded_cloth": folded_cloth,
    "nested_layers": nested_layers,
}


def icosphere(level: int = 5, radius: float = 0.3) -> tuple[np.ndarray, np.ndarray]:
    """Icosahedron subdivided `level` times (each triangle into four, new vertices
    projected onto the sphere): 10 4^level + 2 vertices, 20 4^level faces."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = [tuple(x) for x in f]
    for _ in range(level):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                p = verts[a] + verts[b]
                verts.append(p / np.linalg.norm(p))
                mid[k] = len(verts) - 1
            return mid[k]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return (np.asarray(verts) * radius).astype(np.float64), np.asarray(faces, np.uint32)


def _drape(n: int, side: float, radius: float, cap_angle: float) -> np.ndarray:
    """The n x n grid wrapped over the sphere (centre at the origin, y up) keeping arc
    length from the sheet's centre: within `cap_angle` of the pole the sheet lies on
    the sphere, beyond it continues along the tangent at the cap's rim."""
    g = grid_sheet(n, n, side, side) - np.array([side / 2, 0.0, side / 2])
    rho = np.hypot(g[:, 0], g[:, 2])
    phi = np.arctan2(g[:, 2], g[:, 0])
    d = np.stack([np.cos(phi), np.zeros_like(phi), np.sin(phi)], 1)
    theta = np.minimum(rho / radius, cap_angle)
    on = np.stack([np.sin(theta), np.cos(theta), np.sin(theta)], 1)
    p = radius * np.stack([on[:, 0] * d[:, 0], on[:, 1], on[:, 2] * d[:, 2]], 1)
    extra = np.maximum(rho - radius * cap_angle, 0.0)
    tangent = np.stack([np.cos(cap_angle) * d[:, 0], -np.sin(cap_angle) * np.ones_like(phi), np.cos(cap_angle) * d[:, 2]], 1)
    return p + extra[:, None] * tangent


def _gap_states(base: np.ndarray, radius: float, rng, sigma: float, gap_lo: float, gap_hi: float) -> np.ndarray:
    """base perturbed by N(0, sigma) per coordinate; every vertex that comes within
    1 mm of the sphere is moved radially to a gap drawn from U(gap_lo, gap_hi)."""
    x = base + rng.normal(0.0, sigma, base.shape)
    r = np.linalg.norm(x, axis=1)
    near = r < radius + 1e-3
    gap = rng.uniform(gap_lo, gap_hi, int(near.sum()))
    x[near] *= ((radius + gap) / r[near])[:, None]
    return x


def draped_cloth(seed: int = 0, n: int = 64, side: float = 1.0, radius: float = 0.3, level: int = 5,
                 cap_angle: float = np.pi / 3, states: int = 0, sigma: float = 1e-3, gap=(3e-4, 9e-4)):
    """The differentiation workload: cloth (vertices 0 .. n^2 - 1) over a fixed
    icosphere (the rest).  Returns (mesh, rest, cloth_vertex_count, states[B, V, 3]):
    the base drape has every cloth vertex near the sphere at a gap U(gap) and the
    states perturb the cloth by N(0, sigma) with the same gap rule, so every
    contact distance stays inside (1e-6, d_hat = 1e-3) (a cloth edge of 1.6 cm sags
    at most 0.11 mm below its end points over the sphere and a sphere facet lies
    at most 0.04 mm inside it)."""
    rng = np.random.default_rng(seed)
    sv, sf = icosphere(level, radius)
    cloth = _gap_states(_drape(n, side, radius, cap_angle), radius, rng, 0.0, *gap)
    x = np.concatenate([cloth, sv], 0)
    faces = np.concatenate([grid_faces(n, n, 0), sf + n * n], 0)
    rest = np.concatenate([grid_sheet(n, n, side, side) - np.array([side / 2, 0.0, side / 2]), sv], 0)
    st = np.empty((states, len(x), 3), np.float64)
    for b in range(states):
        st[b, :n * n] = _gap_states(cloth, radius, rng, sigma, *gap)
        st[b, n * n:] = sv
    mesh = Mesh(x, faces, name=f"draped_cloth_{n}x{n}_ico{level}")
    return mesh, rest.astype(np.float32), n * n, st


def subdivided_box(n: int = 8, half: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """Surface of the cube [-half, half]^3 with every face an n x n grid of quads
    (two triangles each, outward normals): 6 n^2 + 2 vertices... precisely
    (n+1)^3 - (n-1)^3 vertices and 12 n^2 triangles."""
    idx = {}
    verts = []

    def vid(p):
        k = tuple(int(c) for c in p)
        if k not in idx:
            idx[k] = len(verts)
            verts.append(k)
        return idx[k]

    faces = []
    # each face: fixed axis a at side s (0 or n), the other two axes (u, v) ordered for an outward normal
    for a in range(3):
        u, v = (a + 1) % 3, (a + 2) % 3
        for s in (0, n):
            for i in range(n):
                for j in range(n):
                    q = []
                    for di, dj in ((0, 0), (1, 0), (1, 1), (0, 1)):
                        p = [0, 0, 0]
                        p[a], p[u], p[v] = s, i + di, j + dj
                        q.append(vid(p))
                    if s == 0:
                        q = q[::-1]
                    faces += [(q[0], q[1], q[2]), (q[0], q[2], q[3])]
    x = (np.asarray(verts, np.float64) / n * 2.0 - 1.0) * half
    return x, np.asarray(faces, np.uint32)


def _random_rotation(rng) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@dataclass
class RigidScene:
    mesh: Mesh  # all bodies' vertices (body by body) at q0
    body: np.ndarray  # uint32 [V]: body of each vertex
    xbar: np.ndarray  # float32 [V, 3]: model position of each vertex (body frame, centred)
    q0: np.ndarray  # float64 [B, 4, 3]: p, then the columns a1, a2, a3 of A (x = p + A xbar)
    radius: np.ndarray  # float64 [B]
    kind: np.ndarray  # uint8 [B]: 0 icosphere, 1 box
    vert_off: np.ndarray  # uint32 [B + 1]


def rigid_bodies(seed: int = 0, n_bodies: int = 512, cube: float = 2.0, r_lo: float = 0.05, r_hi: float = 0.2,
                 max_overlap: float = 0.3, ico_level: int = 3, box_n: int = 8) -> RigidScene:
    """Affine-body workload: bodies alternating icosphere (level 3, 642 vertices) and
    8x8x8-subdivided box (386 vertices, half-extent = the radius), radii U(r_lo, r_hi),
    random orientations, centres drawn uniformly in the cube; a centre is kept when
    it overlaps no placed body by more than max_overlap of the smaller radius
    (centre distance >= r_i + r_j - max_overlap min(r_i, r_j))."""
    rng = np.random.default_rng(seed)
    sv, sf = icosphere(ico_level, 1.0)
    bv, bf = subdivided_box(box_n, 1.0)
    centres, radii = [], []
    attempts = 0
    while len(centres) < n_bodies:
        attempts += 1
        if attempts > 2_000_000:
            raise RuntimeError("rigid_bodies: could not place the bodies")
        r = rng.uniform(r_lo, r_hi)
        c = rng.uniform(r, cube - r, 3)
        if centres:
            C = np.asarray(centres)
            R = np.asarray(radii)
            d = np.linalg.norm(C - c, axis=1)
            if np.any(d < r + R - max_overlap * np.minimum(r, R)):
                continue
        centres.append(c)
        radii.append(r)
    xs, fs, body, xbar, q0, kinds, off = [], [], [], [], [], [], [0]
    for b in range(n_bodies):
        kind = b % 2
        v, f = (sv, sf) if kind == 0 else (bv, bf)
        model = v * radii[b]
        Rm = _random_rotation(rng)
        xs.append(centres[b] + model @ Rm.T)
        fs.append(f + off[-1])
        body.append(np.full(len(v), b, np.uint32))
        xbar.append(model)
        q0.append(np.stack([centres[b], Rm[:, 0], Rm[:, 1], Rm[:, 2]]))
        kinds.append(kind)
        off.append(off[-1] + len(v))
    mesh = Mesh(np.concatenate(xs), np.concatenate(fs), name=f"rigid_bodies_{n_bodies}")
    return RigidScene(mesh, np.concatenate(body), np.concatenate(xbar).astype(np.float32), np.asarray(q0),
                      np.asarray(radii), np.asarray(kinds, np.uint8), np.asarray(off, np.uint32))
