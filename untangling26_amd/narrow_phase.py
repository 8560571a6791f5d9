"""Python mirror of lcs::NarrowPhasesDetector's proximity DCD and CCD queries
(Solver/CollisionDetector/narrow_phase.h:25-346) over ut_contact_query."""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import ContactArgs, lib, ptr
from .lbvh import LBVH
from .runtime import Context

VF_DCD, EE_DCD, VF_CCD, EE_CCD = 0, 1, 2, 3
QUADRATIC, BARRIER = 0, 1
PAIR_WORDS = 16  # CollisionPairTemplate, 64 B


class NarrowPhasesDetector:
    def __init__(self, ctx: Context, num_verts, edges, faces, d_hat, offset, rest_vert_area, rest_edge_area,
                 rest_face_area, vprop=None, kappa=1e8, energy=BARRIER):
        self.ctx = ctx
        self.nv = int(num_verts)
        self.edges, self.faces = edges, faces
        self.d_hat, self.offset = d_hat, offset
        self.areas = (rest_vert_area, rest_edge_area, rest_face_area)
        self.vprop = vprop
        self.kappa = float(kappa)
        self.energy = int(energy)

    def _args(self, x, x_end=None, count=None, pairs=None, pair_ids=None, toi=None):
        a = ContactArgs()
        a.num_verts, a.num_edges = self.nv, self.edges.shape[0]
        a.x, a.x_end = ptr(x), ptr(x_end)
        a.edges, a.faces, a.vprop = ptr(self.edges), ptr(self.faces), ptr(self.vprop)
        a.d_hat, a.offset = ptr(self.d_hat), ptr(self.offset)
        a.rest_vert_area, a.rest_edge_area, a.rest_face_area = (ptr(t) for t in self.areas)
        a.kappa, a.energy = self.kappa, self.energy
        a.count, a.pairs, a.pair_ids = ptr(count), ptr(pairs), ptr(pair_ids)
        a.max_pairs = 0 if pair_ids is None else pair_ids.numel() // 2
        a.toi = ptr(toi)
        return a

    def _run(self, kind, args, tree: LBVH | None, broad=None, broad_count=None):
        cap = 0 if broad is None else broad.numel() // 2
        self.ctx.check(lib().ut_contact_query(self.ctx.handle, kind, C.byref(args),
                                              tree.handle if tree is not None else None, ptr(broad),
                                              ptr(broad_count), cap))

    # DCD: fn_narrow_phase_{vf,ee}_dcd_query
    def vf_dcd_query_repulsion(self, x, count, pairs, pair_ids, face_tree=None, broad=None, broad_count=None):
        self._run(VF_DCD, self._args(x, count=count, pairs=pairs, pair_ids=pair_ids), face_tree, broad, broad_count)

    def ee_dcd_query_repulsion(self, x, count, pairs, pair_ids, edge_tree=None, broad=None, broad_count=None):
        self._run(EE_DCD, self._args(x, count=count, pairs=pairs, pair_ids=pair_ids), edge_tree, broad, broad_count)

    # CCD: fn_narrow_phase_{vf,ee}_ccd_query with the global TOI
    def reset_toi(self, toi):
        self.ctx.check(lib().ut_toi_reset(self.ctx.handle, ptr(toi)))

    def vf_ccd_query(self, x_begin, x_end, toi, face_tree=None, broad=None, broad_count=None):
        self._run(VF_CCD, self._args(x_begin, x_end, toi=toi), face_tree, broad, broad_count)

    def ee_ccd_query(self, x_begin, x_end, toi, edge_tree=None, broad=None, broad_count=None):
        self._run(EE_CCD, self._args(x_begin, x_end, toi=toi), edge_tree, broad, broad_count)
