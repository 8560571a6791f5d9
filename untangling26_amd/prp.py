"""PRP untangling on the GPU (ut_prp_* of include/untangle.h).

Mirrors the PRP branch of IntersectionResolver::host_resolve
(Solver/CollisionDetector/intersection_resolver.cpp:2131-2217) ->
host_resolve_intersections_PRP (intersection_resolver2.cpp:2750-4069): per
contour, candidate separation directions are ray cast from the contour's two
sides, hits are clustered and culled, J(r) picks the direction and the kept
hits become response pairs (CollisionPairTemplate records).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import PrpChoice, PrpParams, PrpStats, lib, ptr
from .runtime import Context

PRP_SLOT = np.dtype([("contour", "u4"), ("combo", "u4"), ("ks", "u4"), ("kd", "u4"), ("dir", "f4", 3),
                     ("ee_pmin", "f4"), ("ee_pmax", "f4"), ("pad", "u4")])
PRP_SUMMARY = np.dtype([("raw_hits", "u4"), ("hit_count", "u4"), ("covered", "u4"), ("ef_count", "u4"),
                        ("depth", "f4"), ("volume", "f4"), ("objective", "f4"), ("area", "f4"),
                        ("disp_cost", "f4"), ("max_hop_src", "u4"), ("max_hop_dst", "u4"), ("usable", "u4"),
                        ("root", "u4"), ("pad", "u4", 3)])
PRP_HIT = np.dtype([("slot", "u4"), ("type", "u4"), ("src", "u4"), ("dst", "u4"), ("b0", "f4"), ("b1", "f4"),
                    ("dist", "f4"), ("root", "u4")])


def _np(a):
    return a.ctypes.data_as(C.c_void_p)


class PrpResolver:
    """GPU PRP resolver of one mesh.  `setup` is an initializer.ClothSetup; the
    device arrays passed here must stay alive while the resolver is used."""

    def __init__(self, ctx: Context, dev: dict, nv: int, ne: int, nf: int, max_contours=256, total_dirs=10,
                 batch_contours=32, response_depth=1.0, kappa=1e8, tau=0.25):
        self.ctx = ctx
        self._dev = dev
        self.nv, self.ne, self.nf = nv, ne, nf
        self.params = PrpParams(max_contours, total_dirs, batch_contours, response_depth, kappa, tau)
        h = C.c_void_p()
        ctx.check(lib().ut_prp_create(ctx.handle, nv, ne, nf, *[ptr(dev[k]) for k in (
            "edges", "faces", "face_adj_edges", "edge_adj_faces", "rest_x", "rest_vert_area", "rest_edge_area",
            "rest_face_area", "mass")], C.byref(self.params), C.byref(h)))
        self.handle = h
        self.batch_n = 0

    @classmethod
    def from_setup(cls, ctx: Context, setup, rest_x, **kw):
        def up(a):
            t = torch.from_numpy(np.ascontiguousarray(a))
            return (t.view(torch.int32) if t.dtype == torch.uint32 else t).to(ctx.device)

        dev = {k: up(getattr(setup, k)) for k in ("edges", "faces", "face_adj_edges", "edge_adj_faces",
                                                   "rest_vert_area", "rest_edge_area", "rest_face_area", "mass")}
        dev["rest_x"] = up(np.asarray(rest_x, np.float32))
        return cls(ctx, dev, setup.nv, len(setup.edges), len(setup.faces), **kw)

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib().ut_prp_destroy(self.handle)
                self.handle = None
        except Exception:
            pass

    def resolve(self, x, contours, face_tree, edge_tree, cap=1 << 16):
        """(choices list of dicts, response records int32[n, 16] on the device, stats dict)."""
        nc = max(contours.num_contours, 1)
        ch = (PrpChoice * nc)()
        st = PrpStats()
        n32 = C.c_uint32()
        while True:
            out = torch.empty((cap, 16), dtype=torch.int32, device=self.ctx.device)
            rc = lib().ut_prp_resolve(self.handle, ptr(x), contours.handle, face_tree.handle, edge_tree.handle,
                                      ptr(out), cap, C.byref(n32), ch, nc, C.byref(st))
            self.ctx.check(rc)
            if n32.value <= cap:
                break
            cap = n32.value
        n = n32.value
        choices = [{k: (list(getattr(c, k)) if k == "dir" else getattr(c, k)) for k, _ in PrpChoice._fields_}
                   for c in ch[:contours.num_contours]]
        return choices, out[:min(n, cap)], {k: getattr(st, k) for k, _ in PrpStats._fields_}

    # ---- parity hooks
    def prepare(self, x, contours, c0, n):
        self.ctx.check(lib().ut_prp_prepare(self.handle, ptr(x), contours.handle, c0, n))
        self.batch_n = n

    def directions(self, c):
        d = np.zeros((64, 3), np.float32)
        n = C.c_uint32()
        self.ctx.check(lib().ut_prp_directions(self.handle, c, _np(d), C.byref(n)))
        return d[:n.value].copy()

    def distances(self, local_c, role):
        hop = np.zeros(self.nv, np.uint32)
        geo = np.zeros(self.nv, np.float32)
        diam = C.c_uint32()
        self.ctx.check(lib().ut_prp_distances(self.handle, local_c, role, _np(hop), _np(geo), C.byref(diam)))
        return hop, geo, diam.value

    def normals(self):
        out = np.zeros(3 * (self.nf + self.ne + self.nv), np.float32)
        self.ctx.check(lib().ut_prp_normals(self.handle, _np(out)))
        fn = out[:3 * self.nf].reshape(-1, 3)
        en = out[3 * self.nf:3 * (self.nf + self.ne)].reshape(-1, 3)
        vn = out[3 * (self.nf + self.ne):].reshape(-1, 3)
        return fn, en, vn

    def eval(self, face_tree, edge_tree, slots, kring):
        """Summaries of the slots and the sorted raw hits (with cluster roots)."""
        slots = np.ascontiguousarray(slots, PRP_SLOT)
        kr = np.ascontiguousarray(kring, np.uint32)
        assert kr.size == 2 * self.batch_n
        out = np.zeros(len(slots), PRP_SUMMARY)
        nh = C.c_uint32()
        self.ctx.check(lib().ut_prp_eval(self.handle, face_tree.handle, edge_tree.handle, _np(slots), len(slots),
                                         _np(kr), _np(out), C.byref(nh)))
        hits = np.zeros(max(nh.value, 1), PRP_HIT)
        self.ctx.check(lib().ut_prp_hits(self.handle, _np(hits), nh.value))
        return out, hits[:nh.value]
