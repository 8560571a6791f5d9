"""Python mirror of lcs::LBVH (Solver/CollisionDetector/lbvh.h:164-230).

Method names and argument meaning follow the reference; buffers are torch
tensors on the context's device.  ``update_*_tree_leave_aabb`` records the
leaf inputs and ``refit`` runs the fused leaf-update + bottom-up refit kernel.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import lib, ptr
from .runtime import Context

TREE_FACE = 1
TREE_EDGE = 2
QUERY_VERTS = 0
QUERY_EDGES = 1
QUERY_EDGES_UNCULLED = 2


def _f32(t):
    assert t.dtype == torch.float32 and t.is_contiguous() and t.is_cuda, "expected contiguous float32 device tensor"
    return t


def _u32(t):
    assert t.dtype in (torch.int32, torch.uint32) and t.is_contiguous() and t.is_cuda, \
        "expected contiguous 32-bit integer device tensor"
    return t


class LBVH:
    def __init__(self, ctx: Context, tree_type: int, num_leaves: int):
        self.ctx = ctx
        self.tree_type = tree_type
        self.num_leaves = int(num_leaves)
        h = C.c_void_p()
        ctx.check(lib().ut_lbvh_create(ctx.handle, tree_type, self.num_leaves, C.byref(h)))
        self.handle = h
        self._leaf_args = None

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().ut_lbvh_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    # -- construct (lbvh.cpp:1003-1047) ----------------------------------
    def reduce_face_tree_aabb(self, x, faces):
        assert self.tree_type == TREE_FACE
        self.ctx.check(lib().ut_lbvh_reduce_aabb(self.handle, ptr(_f32(x)), ptr(_u32(faces))))

    def reduce_edge_tree_aabb(self, x, edges):
        assert self.tree_type == TREE_EDGE
        self.ctx.check(lib().ut_lbvh_reduce_aabb(self.handle, ptr(_f32(x)), ptr(_u32(edges))))

    def construct_tree(self):
        self.ctx.check(lib().ut_lbvh_construct_tree(self.handle))

    def build_from_sorted_keys(self, keys):
        assert keys.dtype == torch.int64 and keys.is_cuda
        self.ctx.check(lib().ut_lbvh_build_from_sorted_keys(self.handle, ptr(keys)))

    # -- refit (lbvh.cpp:1050-1092) --------------------------------------
    def update_face_tree_leave_aabb(self, thickness, start_position, end_position, faces):
        self._leaf_args = (_f32(thickness), _f32(start_position), _f32(end_position), _u32(faces))

    def update_edge_tree_leave_aabb(self, thickness, start_position, end_position, edges):
        self._leaf_args = (_f32(thickness), _f32(start_position), _f32(end_position), _u32(edges))

    def refit(self):
        if self._leaf_args is None:
            raise RuntimeError("refit() needs update_*_tree_leave_aabb() first")
        t, xs, xe, prims = self._leaf_args
        self.ctx.check(lib().ut_lbvh_refit(self.handle, ptr(t), ptr(xs), ptr(xe), ptr(prims)))

    # -- queries (lbvh.cpp:1193-1262) ------------------------------------
    def _query(self, kind, x_begin, x_end, edges, n, count, pair_list, d_hat, thickness):
        cap = pair_list.numel() // 2
        self.ctx.check(lib().ut_lbvh_query(self.handle, kind, ptr(_f32(x_begin)), ptr(_f32(x_end)),
                                           ptr(edges), n, ptr(_f32(d_hat)), ptr(_f32(thickness)),
                                           ptr(_u32(count)), ptr(_u32(pair_list)), cap))

    def broad_phase_query_from_verts(self, x_begin, x_end, count, pair_list, d_hat, thickness):
        self._query(QUERY_VERTS, x_begin, x_end, None, x_begin.shape[0], count, pair_list, d_hat, thickness)

    def broad_phase_query_from_edges(self, x_begin, x_end, edges, count, pair_list, d_hat, thickness):
        self._query(QUERY_EDGES, x_begin, x_end, _u32(edges), edges.shape[0], count, pair_list, d_hat, thickness)

    def broad_phase_query_from_edges_unculled(self, x_begin, x_end, edges, count, pair_list, d_hat, thickness):
        self._query(QUERY_EDGES_UNCULLED, x_begin, x_end, _u32(edges), edges.shape[0], count, pair_list,
                    d_hat, thickness)

    # -- inspection ------------------------------------------------------
    def tree_arrays(self):
        n = self.num_leaves
        nn = 2 * n - 1
        dev = self.ctx.device
        out = dict(
            morton_sorted=torch.empty(n, dtype=torch.int64, device=dev),
            sorted_get_original=torch.empty(n, dtype=torch.int32, device=dev),
            parent=torch.empty(nn, dtype=torch.int32, device=dev),
            children=torch.empty((nn, 2), dtype=torch.int32, device=dev),
            object_idx=torch.empty(nn, dtype=torch.int32, device=dev),
        )
        self.ctx.check(lib().ut_lbvh_tree_arrays(self.handle, *(ptr(out[k]) for k in
                                                                ("morton_sorted", "sorted_get_original", "parent",
                                                                 "children", "object_idx"))))
        return out

    def boxes(self):
        root = torch.empty(6, dtype=torch.float32, device=self.ctx.device)
        glob = torch.empty(6, dtype=torch.float32, device=self.ctx.device)
        self.ctx.check(lib().ut_lbvh_boxes(self.handle, ptr(root), ptr(glob)))
        return root, glob

    def status(self):
        h, t = C.c_uint32(), C.c_uint32()
        self.ctx.check(lib().ut_lbvh_status(self.handle, C.byref(h), C.byref(t)))
        return h.value, t.value


def morton30(ctx: Context, points):
    out = torch.empty(points.shape[0], dtype=torch.int32, device=ctx.device)
    ctx.check(lib().ut_morton30(ctx.handle, ptr(_f32(points)), points.shape[0], ptr(out)))
    return out
