"""Static mesh adjacency used by the per-step path (host side, computed once).

Restates the parts of Initializer/init_mesh_data.cpp:419-500 the hot path
reads: face_adj_edges (the edge between face[j] and face[(j+1)%3]),
edge_adj_faces (the faces around each edge, ascending, 0xFFFFFFFF if absent)
and the edges around each vertex (sa_vert_adj_edges_csr, ascending).
Meshes with an edge shared by more than two faces are rejected.
"""
from __future__ import annotations

import numpy as np

NONE = np.uint32(0xFFFFFFFF)


def mesh_topology(edges: np.ndarray, faces: np.ndarray):
    edges = np.asarray(edges, np.int64)
    faces = np.asarray(faces, np.int64)
    nv = int(max(edges.max(initial=0), faces.max(initial=0))) + 1
    ekey = np.minimum(edges[:, 0], edges[:, 1]) * nv + np.maximum(edges[:, 0], edges[:, 1])
    order = np.argsort(ekey, kind="stable")
    skey = ekey[order]
    a = faces
    b = np.roll(faces, -1, axis=1)  # face[(j+1)%3]
    fkey = np.minimum(a, b) * nv + np.maximum(a, b)
    pos = np.searchsorted(skey, fkey)
    if (pos >= len(skey)).any() or (skey[np.minimum(pos, len(skey) - 1)] != fkey).any():
        raise ValueError("a face edge is missing from the edge list")
    fae = order[pos].astype(np.uint32)  # [F, 3]
    e_ids = fae.ravel().astype(np.int64)
    f_ids = np.repeat(np.arange(faces.shape[0], dtype=np.int64), 3)
    o = np.lexsort((f_ids, e_ids))
    e_sorted, f_sorted = e_ids[o], f_ids[o]
    counts = np.bincount(e_sorted, minlength=edges.shape[0])
    if (counts > 2).any():
        raise ValueError("non-manifold edge (more than two faces)")
    first = np.searchsorted(e_sorted, np.arange(edges.shape[0]))
    eaf = np.full((edges.shape[0], 2), NONE, np.uint32)
    has1 = counts >= 1
    has2 = counts >= 2
    eaf[has1, 0] = f_sorted[first[has1]]
    eaf[has2, 1] = f_sorted[first[has2] + 1]
    return fae, eaf


def rest_areas(x_rest: np.ndarray, edges: np.ndarray, faces: np.ndarray, edge_adj_faces: np.ndarray,
               min_area: float = 1e-10):
    """Face, vertex and edge rest areas (init_mesh_data.cpp:621-700): face area
    = |cross|/2, vertex and edge areas = sum of adjacent face areas / 3, all
    clamped from below like safe_area."""
    x = np.asarray(x_rest, np.float64)
    f = np.asarray(faces, np.int64)
    fa = 0.5 * np.linalg.norm(np.cross(x[f[:, 1]] - x[f[:, 0]], x[f[:, 2]] - x[f[:, 0]]), axis=1)
    fa = np.maximum(fa, min_area)
    va = np.zeros(x.shape[0])
    np.add.at(va, f.ravel(), np.repeat(fa / 3.0, 3))
    eaf = np.asarray(edge_adj_faces, np.int64)
    ea = np.zeros(eaf.shape[0])
    for k in range(2):
        m = eaf[:, k] != 0xFFFFFFFF
        ea[m] += fa[eaf[m, k]] / 3.0
    return (np.maximum(va, min_area).astype(np.float32), np.maximum(ea, min_area).astype(np.float32),
            fa.astype(np.float32))


def vert_adj_edges(num_verts: int, edges: np.ndarray):
    """Edges around each vertex as CSR: (offsets uint32[V+1], edge ids ascending per vertex)."""
    e = np.asarray(edges, np.int64).reshape(-1, 2)
    v = e.ravel()
    eid = np.repeat(np.arange(len(e), dtype=np.int64), 2)
    order = np.lexsort((eid, v))
    off = np.zeros(num_verts + 1, np.uint32)
    off[1:] = np.cumsum(np.bincount(v, minlength=num_verts))
    return off, eid[order].astype(np.uint32)
