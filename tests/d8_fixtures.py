"""The integer-terrain fixtures of the reference's D8 byte-identity test,
restated, with its FNV-1a pins.

Recipe: tests/cpp/test_flow_d8_identity.cpp:176-294 (LCG, masks, forced
outlets, BFS distance to a cell that may drain out, z = 1000*d + noise);
specs :296-322; hashing :118-134 and :343-427; pins :517-588.
"""
from __future__ import annotations

from collections import deque

import numpy as np

MASK64 = (1 << 64) - 1
D8R = (-1, -1, 0, 1, 1, 1, 0, -1)
D8C = (0, 1, 1, 1, 0, -1, -1, -1)
ABSENT = 0xABAD1DEAABAD1DEA
COMPONENTS = ("receiver", "direction", "donors", "outlets", "topo", "svector",
              "area", "weighted")


class LCG:
    def __init__(self, seed):
        self.s = seed & MASK64

    def __call__(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & MASK64
        return self.s >> 33


# name, rows, cols, cell, bc, mask, nodata_is_outlet, forced, pits, seed
SPECS = [
    ("open_square", 61, 67, 10.0, "oooo", "none", True, 0, False, 11),
    ("open_nonint_cell", 47, 53, 30.137, "oooo", "none", True, 0, False, 12),
    ("holes_nodata_outlet", 59, 71, 10.0, "oooo", "holes", True, 0, False, 13),
    ("holes_nodata_blocked", 59, 71, 10.0, "oooo", "holes", False, 0, False, 14),
    ("nodata_border", 43, 57, 10.0, "oooo", "border", True, 0, False, 15),
    ("noflux_forced", 41, 45, 10.0, "nnnn", "none", True, 6, False, 16),
    ("periodic_forced", 37, 39, 10.0, "pppp", "none", True, 4, False, 17),
    ("periodic_ns_open_ew", 35, 41, 10.0, "popo", "none", True, 0, False, 18),
    ("baselevel_edges", 33, 36, 10.0, "bbbb", "none", True, 0, False, 19),
    ("mixed_onbn", 38, 44, 10.0, "onbn", "holes", True, 0, False, 20),
    ("forced_on_open", 40, 40, 10.0, "oooo", "none", True, 5, False, 21),
    ("pits_as_outlets", 31, 37, 10.0, "nnnn", "none", True, 0, True, 22),
    ("periodic_2_rows", 2, 9, 10.0, "pppp", "none", True, 1, False, 23),
    ("periodic_1_col", 7, 1, 10.0, "pppp", "none", True, 1, False, 24),
    ("tiny_3x3_periodic", 3, 3, 10.0, "pppp", "none", True, 1, False, 25),
    ("single_cell", 1, 1, 10.0, "oooo", "none", True, 0, False, 26),
    ("two_by_n", 2, 11, 10.0, "oooo", "none", True, 0, False, 27),
]

# (topo kept, s-vector built) -- test_flow_d8_identity.cpp:334-339
CONFIGS = {"sv_on": (True, True), "sv_off": (True, False),
           "sv_only": (False, True), "receivers_only": (False, False)}


class Fixture:
    def __init__(self, spec):
        (self.name, self.nr, self.nc, self.dx, self.bc, self.mask_kind,
         self.nodata_is_outlet, self.n_forced, self.pits, self.seed) = spec
        self.nodata = -9999.0
        self._make()

    def _step(self, r, c, d):
        """(kind, nr, nc): kind 0 blocked, 1 inside, 2 off an o/b edge"""
        rr, cc = r + D8R[d], c + D8C[d]
        bc = self.bc
        off = False
        if rr < 0:
            if bc[0] == "p":
                rr = self.nr - 1
            elif bc[0] in "ob":
                off = True
            else:
                return 0, 0, 0
        if rr >= self.nr:
            if bc[2] == "p":
                rr = 0
            elif bc[2] in "ob":
                off = True
            else:
                return 0, 0, 0
        if cc < 0:
            if bc[3] == "p":
                cc = self.nc - 1
            elif bc[3] in "ob":
                off = True
            else:
                return 0, 0, 0
        if cc >= self.nc:
            if bc[1] == "p":
                cc = 0
            elif bc[1] in "ob":
                off = True
            else:
                return 0, 0, 0
        if off:
            return 2, 0, 0
        return 1, rr, cc

    def _make(self):
        R, Cc = self.nr, self.nc
        N = R * Cc
        s = LCG(self.seed)
        m = np.ones(N, np.uint8)
        for r in range(R):
            for c in range(Cc):
                i = r * Cc + c
                if self.mask_kind == "holes" and s() % 100 < 4:
                    m[i] = 0
                if self.mask_kind == "border" and (r < 2 or c < 3 or r >= R - 1):
                    m[i] = 0
        forced = []
        for _ in range(self.n_forced):
            f = s() % N
            while m[f] == 0:
                f = (f + 1) % N
            if f not in forced:
                forced.append(f)
        dist = np.full(N, -1, np.int64)
        q = deque()
        bc = self.bc
        for r in range(R):
            for c in range(Cc):
                i = r * Cc + c
                if not m[i]:
                    continue
                sink = i in forced
                if ((r == 0 and bc[0] == "b") or (c == Cc - 1 and bc[1] == "b") or
                        (r == R - 1 and bc[2] == "b") or (c == 0 and bc[3] == "b")):
                    sink = True
                for d in range(8):
                    if sink:
                        break
                    k, rr, cc = self._step(r, c, d)
                    if k == 0:
                        continue
                    if k == 2:
                        sink = True
                        break
                    if not m[rr * Cc + cc] and self.nodata_is_outlet:
                        sink = True
                if sink:
                    dist[i] = 0
                    q.append(i)
        while q:
            i = q.popleft()
            r, c = divmod(i, Cc)
            for d in range(8):
                k, rr, cc = self._step(r, c, d)
                if k != 1:
                    continue
                j = rr * Cc + cc
                if not m[j] or dist[j] >= 0:
                    continue
                dist[j] = dist[i] + 1
                q.append(j)
        dem = np.full(N, np.float32(self.nodata), np.float32)
        mask = np.zeros(N, np.uint8)
        for i in range(N):
            if not m[i]:
                continue
            noise = (s() % 4) * 250
            if self.pits:
                z = 1000 + noise
            else:
                if dist[i] < 0:
                    raise ValueError(f"fixture {self.name} has an unroutable cell")
                z = 1000 * int(dist[i]) + noise
            dem[i] = np.float32(z)
            mask[i] = 1
        self.dem = dem.reshape(R, Cc)
        self.mask = mask.reshape(R, Cc)
        self.forced = np.array(forced, np.int64)


class FNV:
    """64-bit FNV-1a, test_flow_d8_identity.cpp:118-134"""

    def __init__(self):
        self.h = 1469598103934665603

    def bytes(self, b: bytes):
        from oracle.oracle import fnv1a
        self.h = fnv1a(self.h, b)

    def arr(self, a):
        self.bytes(np.ascontiguousarray(a).tobytes())

    def vec(self, a):
        a = np.ascontiguousarray(a)
        self.bytes(np.uint64(a.size).tobytes())
        self.arr(a)


def hash_flow(fx, cfg, flow, area_data, area_mask, weighted):
    """Hash row in the reference's component order.  `flow` holds int64
    receiver, uint8 dir/donors, int64 outlets/topo/s_vector/s_start and
    uint64 s_size as full arrays (topo/s-vector None when absent)."""
    keep_topo, sv = CONFIGS[cfg]
    out = [ABSENT] * 8
    n = fx.nr * fx.nc
    h = FNV(); h.arr(flow["receiver"].astype(np.int64)); out[0] = h.h
    h = FNV(); h.arr(flow["dir"].astype(np.uint8)); out[1] = h.h
    h = FNV(); h.arr(flow["donors"].astype(np.uint8)); out[2] = h.h
    h = FNV(); h.vec(flow["outlets"].astype(np.int64)); out[3] = h.h
    if keep_topo:
        h = FNV(); h.vec(flow["topo"].astype(np.int64)); out[4] = h.h
    if sv:
        h = FNV()
        ss = flow["s_start"].astype(np.int64)
        sz = flow["s_size"].astype(np.uint64)
        svec = flow["s_vector"].astype(np.int64)
        total = 0
        parts = []
        for o in flow["outlets"]:
            st, size = int(ss[o]), int(sz[o])
            parts.append(svec[st:st + size].tobytes())
            total += size
        for p in parts:
            h.bytes(p)
        h.bytes(np.uint64(total).tobytes())
        inter = np.empty(2 * n, np.int64)
        inter[0::2] = ss
        inter[1::2] = sz.view(np.int64)
        h.bytes(inter.tobytes())
        out[5] = h.h
    if keep_topo or sv:
        h = FNV(); h.vec(area_data.astype(np.float32)); h.vec(area_mask.astype(np.uint8))
        out[6] = h.h
    if keep_topo:
        h = FNV(); h.vec(weighted.astype(np.float64)); out[7] = h.h
    return out


def weights_for(n):
    f = np.arange(n, dtype=np.int64)
    return 0.125 + 0.375 * (f % 7).astype(np.float64)


def load_pins(path):
    pins = {}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            key, *vals = line.split()
            pins[key] = [int(v, 16) for v in vals]
    return pins
