"""Intersection contours on the GPU (ut_contours_* of include/untangle.h).

Mirrors IntersectionResolver::host_preprocess_ef_pairs
(Solver/CollisionDetector/intersection_resolver.cpp:683-1096) and
intersection_contour_construction_from_ext_adjacent
(intersection_resolver2.cpp:254-309): canonical EF order, loop-vertex pairs,
EF adjacency, contour index and mesh parity per pair.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import lib, ptr
from .runtime import Context

(CT_PAIR_IDS, CT_PAIRS, CT_CONTOUR_INDEX, CT_MESH_INDEX, CT_ADJ, CT_ADJ_N, CT_INCONSISTENT, CT_COMP_MIN,
 CT_ADJ_OFF, CT_CONTOUR_TYPE, CT_MESH_CLOSED) = range(11)

# ContourType (intersection_resolver2.h:27-37)
CONTOUR_TYPES = ("Undifined", "Closed", "Eight", "LL", "BLI", "Cross", "BIBI", "BBII")


class ContourBuilder:
    def __init__(self, ctx: Context, num_verts: int, edges, faces, face_adj_edges, edge_adj_faces):
        self.ctx = ctx
        self._keep = (edges, faces, face_adj_edges, edge_adj_faces)  # the library keeps the pointers
        h = C.c_void_p()
        ctx.check(lib().ut_contours_create(ctx.handle, num_verts, edges.shape[0], faces.shape[0], ptr(edges),
                                           ptr(faces), ptr(face_adj_edges), ptr(edge_adj_faces), C.byref(h)))
        self.handle = h
        self.num_total = 0
        self.num_contours = 0

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().ut_contours_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def build(self, num_pairs: int, pairs, pair_ids, ignore_near_zero_dist_pairs: bool = True):
        """ignore_near_zero_dist_pairs=False (the CLI's setting) drops degenerate
        pairs that have no non-degenerate neighbour."""
        nt, nc = C.c_uint32(), C.c_uint32()
        self.ctx.check(lib().ut_contours_build(self.handle, num_pairs, ptr(pairs), ptr(pair_ids),
                                               int(bool(ignore_near_zero_dist_pairs)), C.byref(nt), C.byref(nc)))
        self.num_total, self.num_contours = nt.value, nc.value
        return self.num_total, self.num_contours

    def counts(self):
        """(pairs given, pairs after the degenerate filter, total with loop pairs, adjacency entries)."""
        out = (C.c_uint32 * 4)()
        self.ctx.check(lib().ut_contours_counts(self.handle, out))
        return tuple(out)

    def get(self, what: int):
        n = self.num_total
        shape = {CT_PAIR_IDS: (n, 2), CT_PAIRS: (n, 12), CT_ADJ: (self.counts()[3],),
                 CT_ADJ_OFF: (n + 1 if n else 0,), CT_CONTOUR_TYPE: (self.num_contours,),
                 CT_MESH_CLOSED: (self.num_contours, 2)}.get(what, (n,))
        out = torch.empty(shape, dtype=torch.int32, device=self.ctx.device)
        if out.numel():
            self.ctx.check(lib().ut_contours_get(self.handle, what, ptr(out)))
        return out


def graph_components(ctx: Context, row_ptr, col_idx, flip_w):
    """(comp_min, parity, inconsistent) of a CSR graph with 0/1 flip weights."""
    n = row_ptr.shape[0] - 1
    dev = ctx.device
    scratch = torch.empty(2 * max(n, 1), dtype=torch.int32, device=dev)
    cm = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    par = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    inc = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ctx.check(lib().ut_graph_components(ctx.handle, n, ptr(row_ptr), ptr(col_idx), ptr(flip_w), ptr(scratch),
                                        ptr(cm), ptr(par), ptr(inc)))
    return cm[:n], par[:n], inc[:n]


def graph_components_static(ctx: Context, row_ptr, col_idx, flip_w):
    """graph_components through ut_graph_components_static: the union-find
    parents live in the library's static device array (no workspace, no
    allocation inside the library; n <= 2^20)."""
    n = row_ptr.shape[0] - 1
    dev = ctx.device
    cm = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    par = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    inc = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    ctx.check(lib().ut_graph_components_static(ctx.handle, n, ptr(row_ptr), ptr(col_idx), ptr(flip_w), ptr(cm),
                                               ptr(par), ptr(inc)))
    return cm[:n], par[:n], inc[:n]
