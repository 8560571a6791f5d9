"""Seeded synthetic scenes for the untangling workloads (numpy, host side).

World is y-up (the reference's gravity is (0, -9.8, 0), scene_params.h:131),
so cloth sheets lie in the x-z plane and "vertical" means y.

  two_sheets     two 32x32-vertex 1 m sheets, B rotated by U[10,30] deg about a
                 random horizontal axis and cutting A with max penetration 0.05 m
  folded_cloth   256x256-vertex 2 m sheet folded about its midline; the folded
                 half is displaced by 0.08 sin(2 pi u) sin(pi v) + N(0, 0.005)
  nested_layers  8 stacked 128x128-vertex sheets, 4 mm apart, vertex jitter
                 N(0, 6 mm) and a per-layer in-plane shift U(-0.05, 0.05) m
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
