"""EF intersection detection, mirroring IntersectionResolver::ef_dcd_intersection_query
(Solver/CollisionDetector/intersection_resolver.cpp:365-423) and the broad phase
that feeds it (NewtonSolver::device_broadphase_dcd, newton_solver.cpp:1260-1299).

``EfDetector.detect`` runs the whole per-step detection: face/edge LBVH
rebuild, refit, the unculled edge-vs-face query and the narrow EF test, with the
reference's resize-and-retry on capacity overflow (newton_solver.cpp:1402-1438).
"""
from __future__ import annotations

import torch

from ._lib import lib, ptr
from .lbvh import LBVH, TREE_FACE
from .runtime import Context

EF_PAIR_WORDS = 12  # CollisionPair::EfPair, collision_data.h:152-220


def ef_dcd_intersection_query(ctx: Context, broad, broad_count, x, edges, faces, d_hat, count, pairs, pair_ids,
                              vprop=None):
    cap_b = broad.numel() // 2
    cap = pair_ids.numel() // 2
    ctx.check(lib().ut_ef_dcd(ctx.handle, ptr(broad), ptr(broad_count), cap_b, ptr(x), ptr(edges), ptr(faces),
                              ptr(vprop), ptr(d_hat), ptr(count), ptr(pairs), ptr(pair_ids), cap))


def ef_detect_fused(ctx: Context, face_tree: LBVH, x, edges, faces, d_hat, thickness, count, pairs, pair_ids,
                    vprop=None):
    """Broad + narrow EF detection in one kernel (ut_ef_detect); face_tree must be refit."""
    cap = pair_ids.numel() // 2
    ctx.check(lib().ut_ef_detect(face_tree.handle, ptr(x), ptr(edges), edges.shape[0], ptr(faces), ptr(vprop),
                                 ptr(d_hat), ptr(thickness), ptr(count), ptr(pairs), ptr(pair_ids), cap))


def add_ef_from_proximity_pairs(ctx: Context, recs, rec_ids, rec_count, x, edges, faces, vert_adj_off,
                                vert_adj_edges, edge_adj_faces, count, pairs, pair_ids):
    """fn_add_ef_from_proximity_pairs (intersection_resolver.cpp:199-363): appends
    EF pairs made from the zero-distance VF / EE proximity pairs (ut_ef_from_proximity)."""
    rec_cap = rec_ids.numel() // 2
    cap = pair_ids.numel() // 2
    ctx.check(lib().ut_ef_from_proximity(ctx.handle, ptr(recs), ptr(rec_ids), ptr(rec_count), rec_cap, ptr(x),
                                         ptr(edges), ptr(faces), ptr(vert_adj_off), ptr(vert_adj_edges),
                                         ptr(edge_adj_faces), ptr(count), ptr(pairs), ptr(pair_ids), cap))


def decode_ef_pairs(pairs: torch.Tensor, n: int):
    """Splits EfPair words into (indices uint32[n,5], baries float[n,3])."""
    w = pairs[: n * EF_PAIR_WORDS].view(n, EF_PAIR_WORDS)
    return w[:, :5], w[:, 5:8].view(torch.float32)


class EfDetector:
    """Owns the face LBVH and the broad/narrow lists for EF intersection detection."""

    def __init__(self, ctx: Context, num_verts: int, edges: torch.Tensor, faces: torch.Tensor,
                 broad_cap: int | None = None, ef_cap: int | None = None):
        self.ctx = ctx
        self.edges = edges.contiguous()
        self.faces = faces.contiguous()
        self.num_verts = num_verts
        self.ne = edges.shape[0]
        self.nf = faces.shape[0]
        self.face_tree = LBVH(ctx, TREE_FACE, self.nf)
        self.counts = ctx.zeros(4, torch.int32)  # [0] broad EF, [1] narrow EF
        self.broad_cap = broad_cap or max(1024, 4 * self.ne)
        self.ef_cap = ef_cap or max(1024, self.ne)
        self._alloc()

    def _alloc(self):
        self.broad = self.ctx.empty(2 * self.broad_cap, torch.int32)
        self.pairs = self.ctx.empty(EF_PAIR_WORDS * self.ef_cap, torch.int32)
        self.pair_ids = self.ctx.empty(2 * self.ef_cap, torch.int32)

    def build(self, x):
        self.face_tree.reduce_face_tree_aabb(x, self.faces)
        self.face_tree.construct_tree()

    def _broad(self, x, d_hat, thickness):
        self.counts[0].zero_()
        self.face_tree.update_face_tree_leave_aabb(thickness, x, x, self.faces)
        self.face_tree.refit()
        self.face_tree.broad_phase_query_from_edges_unculled(x, x, self.edges, self.counts[0:1], self.broad,
                                                             d_hat, thickness)

    def _narrow(self, x, d_hat, vprop):
        self.counts[1].zero_()
        ef_dcd_intersection_query(self.ctx, self.broad, self.counts[0:1], x, self.edges, self.faces, d_hat,
                                  self.counts[1:2], self.pairs, self.pair_ids, vprop)

    def detect(self, x, d_hat, thickness, vprop=None, rebuild: bool = True, max_attempts: int = 10,
               fused: bool = True):
        """Returns the number of EF pairs; pair_ids[:n] holds (eid, fid).

        fused=True runs ut_ef_detect (no broad list); fused=False runs the
        reference's two steps (broad query, then the EF narrow phase)."""
        if rebuild:
            self.build(x)
        if fused:
            self.face_tree.update_face_tree_leave_aabb(thickness, x, x, self.faces)
            self.face_tree.refit()
            for _ in range(max_attempts):
                self.counts[1].zero_()
                ef_detect_fused(self.ctx, self.face_tree, x, self.edges, self.faces, d_hat, thickness,
                                self.counts[1:2], self.pairs, self.pair_ids, vprop)
                n = int(self.counts[1].item())
                if n <= self.ef_cap:
                    return n
                self.ef_cap = 2 * n
                self._alloc()
            raise RuntimeError("EF count still out of range after resizing")
        for _ in range(max_attempts):
            self._broad(x, d_hat, thickness)
            nb = int(self.counts[0].item())
            if nb <= self.broad_cap:
                break
            self.broad_cap = 2 * nb
            self._alloc()
        else:
            raise RuntimeError("broad phase EF count still out of range after resizing")
        for _ in range(max_attempts):
            self._narrow(x, d_hat, vprop)
            n = int(self.counts[1].item())
            if n <= self.ef_cap:
                return n
            self.ef_cap = 2 * n
            self._alloc()
            self._broad(x, d_hat, thickness)
        raise RuntimeError("narrow phase EF count still out of range after resizing")
