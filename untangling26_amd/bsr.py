"""3x3-block sparse matrices in the SELL-64 layout of the PCG kernels.

Host-side (numpy) construction of the static block pattern: block rows are
grouped in slices of 64 (one wavefront), each slice padded to its longest
row, columns of padding slots point at the row itself with zero values.
Values are planar: vals[c, slot] for entry c = 3*i + j of a block.
"""
from __future__ import annotations

import numpy as np

SLICE = 64


class SellPattern:
    def __init__(self, n: int, rows: np.ndarray, cols: np.ndarray):
        rows = np.asarray(rows, np.int64)
        cols = np.asarray(cols, np.int64)
        key = np.unique(np.concatenate([rows * n + cols, np.arange(n, dtype=np.int64) * (n + 1)]))
        r, c = key // n, key % n
        self.n = n
        counts = np.bincount(r, minlength=n)
        nslices = (n + SLICE - 1) // SLICE
        pad_counts = np.zeros(nslices * SLICE, np.int64)
        pad_counts[:n] = counts
        widths = pad_counts.reshape(nslices, SLICE).max(1)
        slice_off = np.zeros(nslices + 1, np.int64)
        slice_off[1:] = np.cumsum(widths * SLICE)
        self.nslices = nslices
        self.nslots = int(slice_off[-1])
        row_start = np.zeros(n + 1, np.int64)
        row_start[1:] = np.cumsum(counts)
        k = np.arange(len(key)) - row_start[r]  # position of the block within its row
        s = r // SLICE
        slots = slice_off[s] + k * SLICE + (r % SLICE)
        col = np.empty(self.nslots, np.int64)
        # padding: column = own row
        for sl in range(nslices):
            base = slice_off[sl]
            w = widths[sl]
            lanes = np.arange(SLICE)
            rr = np.minimum(sl * SLICE + lanes, n - 1)
            col[base:base + w * SLICE] = np.tile(rr, w)
        col[slots] = c
        self.slice_off = slice_off.astype(np.uint32)
        self.col = col.astype(np.uint32)
        self.keys = key  # sorted row * n + col
        self.key_slots = slots
        self.diag_slot = self.slots(np.arange(n), np.arange(n)).astype(np.uint32)

    def slots(self, r, c) -> np.ndarray:
        q = np.asarray(r, np.int64) * self.n + np.asarray(c, np.int64)
        pos = np.searchsorted(self.keys, q)
        if (pos >= len(self.keys)).any() or (self.keys[np.minimum(pos, len(self.keys) - 1)] != q).any():
            raise KeyError("block not in pattern")
        return self.key_slots[pos]

    def scatter(self, r, c, blocks, dtype=np.float64) -> np.ndarray:
        """Planar values (9, nslots) with blocks[k] (3x3) summed into (r[k], c[k])."""
        vals = np.zeros((9, self.nslots), np.float64)
        sl = self.slots(r, c)
        b = np.asarray(blocks, np.float64).reshape(-1, 9)
        for comp in range(9):
            np.add.at(vals[comp], sl, b[:, comp])
        return vals.astype(dtype)

    def blocks_of(self, vals):
        """(rows, cols, blocks[k,3,3]) of the stored (non-padding) blocks."""
        r = self.keys // self.n
        c = self.keys % self.n
        b = np.asarray(vals)[:, self.key_slots].T.reshape(-1, 3, 3)
        return r, c, b


def cloth_pattern(n_verts: int, faces: np.ndarray, edges: np.ndarray, edge_adj_faces: np.ndarray) -> SellPattern:
    """Blocks coupled by stretch (vertices of a face) and bending (the four
    vertices of the two faces around an interior edge)."""
    f = np.asarray(faces, np.int64)
    rr = [np.repeat(f, 3, axis=1).ravel()]
    cc = [np.tile(f, (1, 3)).ravel()]
    eaf = np.asarray(edge_adj_faces, np.int64)
    interior = (eaf[:, 1] != 0xFFFFFFFF) & (eaf[:, 0] != 0xFFFFFFFF)
    e = np.asarray(edges, np.int64)[interior]
    f0 = f[eaf[interior, 0]]
    f1 = f[eaf[interior, 1]]
    o0 = f0.sum(1) - e.sum(1)
    o1 = f1.sum(1) - e.sum(1)
    quad = np.stack([e[:, 0], e[:, 1], o0, o1], 1)
    rr.append(np.repeat(quad, 4, axis=1).ravel())
    cc.append(np.tile(quad, (1, 4)).ravel())
    return SellPattern(n_verts, np.concatenate(rr), np.concatenate(cc))
