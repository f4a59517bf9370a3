"""Many boxes in one HIP-graph launch, with autotuned tiles.

The batched-realisation workload: 48 independent 2LPT FP32 boxes (16 x 128^3,
16 x 192^3, 16 x 256^3), seed_i = 12345 + i, box size L_i = 100 + 700 U_i with
U_i the i-th draw of std::mt19937_64(20261202) through
std::uniform_real_distribution<double>(0, 1) (libstdc++).  Each box is an
``ICRun`` on the same context; the whole sequence is captured once with
``mfic_graph_begin`` / ``mfic_graph_end`` and replayed with one
``mfic_graph_launch``.  There is no counterpart in monofonIC, which generates
one box per process.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence, Tuple

from . import _lib as L
from .grid import Context
from .lpt import ICConfig, ICRun


class MT19937_64:
    """std::mt19937_64 (C++11 [rand.predef])."""

    N, M = 312, 156
    MATRIX_A = 0xB5026F5AA96619E9
    UPPER, LOWER = 0xFFFFFFFF80000000, 0x7FFFFFFF
    MASK = (1 << 64) - 1

    def __init__(self, seed: int = 5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & self.MASK
        for i in range(1, self.N):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & self.MASK
        self.idx = self.N

    def _twist(self):
        mt = self.mt
        for i in range(self.N):
            x = (mt[i] & self.UPPER) | (mt[(i + 1) % self.N] & self.LOWER)
            xa = x >> 1
            if x & 1:
                xa ^= self.MATRIX_A
            mt[i] = mt[(i + self.M) % self.N] ^ xa
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & self.MASK


def uniform01(g: MT19937_64) -> float:
    """std::uniform_real_distribution<double>(0, 1) on a 64-bit engine with
    libstdc++'s generate_canonical: one draw, x / 2^64, capped below 1."""
    u = float(g()) / 18446744073709551616.0
    return u if u < 1.0 else 1.0 - 2.0 ** -53


def batch_boxes(seed: int = 20261202) -> List[Tuple[int, int, float]]:
    """(n, seed, L) of the 48-box workload"""
    g = MT19937_64(seed)
    out = []
    for i in range(48):
        n = 128 if i < 16 else (192 if i < 32 else 256)
        out.append((n, 12345 + i, 100.0 + 700.0 * uniform01(g)))
    return out


class Graph:
    """mfic_graph: replay of every call captured on a context stream."""

    def __init__(self, h: C.c_void_p):
        self.h = h

    def launch(self, blocking: bool = True) -> None:
        L.call("mfic_graph_launch", self.h, 1 if blocking else 0)

    def close(self) -> None:
        if self.h:
            L.call("mfic_graph_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def capture(ctx: Context, fn) -> Graph:
    """run fn() (library calls on ctx) under stream capture and return the
    instantiated graph; fn must have run once before (lazy allocations)"""
    L.call("mfic_graph_begin", ctx.h)
    try:
        fn()
    except Exception:
        h = C.c_void_p()
        L.load().mfic_graph_end(ctx.h, C.byref(h))
        raise
    h = C.c_void_p()
    L.call("mfic_graph_end", ctx.h, C.byref(h))
    g = Graph(h)
    ctx.adopt(g)
    return g


def autotune(ctx: Context, n: int, prec: int, lpt_order: int) -> float:
    """radix plan and tile autotuner for boxes of size n (sc lattice,
    component fields); returns the tuned time (ms)"""
    ms = C.c_double()
    L.call("mfic_autotune", ctx.h, n, prec, lpt_order, C.byref(ms))
    return ms.value


def autotune_config(ctx: Context, cfg, cosmo: L.CosmoParams) -> float:
    """the autotuner on a given run configuration (lpt.ICConfig) and
    cosmology; returns the tuned time (ms)"""
    ms = C.c_double()
    c = cfg.to_c()
    L.call("mfic_autotune_config", ctx.h, C.byref(c), C.byref(cosmo), C.byref(ms))
    return ms.value


def plan_count(length: int) -> int:
    """number of compiled radix plans of a transform length"""
    return L.load().mfic_fft_plan_count(length)


def plan_info(length: int, variant: int):
    """(elements per thread, radices) of radix plan `variant` of a length"""
    e = C.c_int()
    r = (C.c_int * 4)()
    L.call("mfic_fft_plan_info", length, variant, C.byref(e), r)
    return e.value, tuple(x for x in r if x)


def set_plan(ctx: Context, length: int, pass_kind: int, prec: int, variant: int) -> None:
    """select radix plan `variant` (< 0: the default) for one strided pass kind"""
    L.call("mfic_fft_set_plan", ctx.h, length, pass_kind, prec, variant)


def get_plan(ctx: Context, length: int, pass_kind: int, prec: int) -> int:
    v = C.c_int()
    L.call("mfic_fft_get_plan", ctx.h, length, pass_kind, prec, C.byref(v))
    return v.value


def chosen_plans(ctx: Context, n: int, prec: int) -> dict:
    """{(length, pass kind): variant} of every non-default plan a box of
    size n can use"""
    out = {}
    for length in (n, 3 * n // 2):
        for p in L.PASSES:
            if p == L.PASS_FUSED_ROWS:  # two forms at the padded length
                if length != 3 * n // 2:
                    continue
            elif plan_count(length) < 2:
                continue
            v = get_plan(ctx, length, p, prec)
            if v:
                out[(length, p)] = v
    return out


def chosen_tiles(ctx: Context, n: int, prec: int) -> dict:
    """{(length, kind): tile} of every tile override (autotuned or set) a
    box of size n can use: kinds 0-2 at n and n/2, the padded pass kinds at
    3n/2"""
    out = {}
    keys = [(n, 0), (n // 2, 1), (n, 2), (3 * n // 2, 0), (3 * n // 4, 1)]
    keys += [(3 * n // 2, k) for k in range(L.PASS_PAD_C2R_A, L.PASS_TRUNC_AX1 + 1)]
    for length, kind in keys:
        if not L.load().mfic_fft_size_supported(length):
            continue
        t, pad = C.c_int(), C.c_int()
        L.call("mfic_fft_get_tile", ctx.h, length, kind, prec, C.byref(t), C.byref(pad))
        if t.value:
            out[(length, kind)] = t.value
    return out


class Batch:
    """A list of boxes run back to back on one context, optionally as one graph."""

    def __init__(self, ctx: Context, boxes: Sequence[Tuple[int, int, float]], lpt_order: int = 2, prec: int = L.F32,
                 zstart: float = 24.0, tune: bool = True):
        self.ctx = ctx
        self.boxes = list(boxes)
        self.tuned = {}
        if tune:
            for n in sorted({b[0] for b in self.boxes}):
                self.tuned[n] = autotune(ctx, n, prec, lpt_order)
        self.runs = [ICRun(ctx, ICConfig(n=n, box=box, zstart=zstart, lpt_order=lpt_order, prec=prec, seed=seed))
                     for n, seed, box in self.boxes]
        self.graph = None

    @property
    def cells(self) -> int:
        return sum(b[0] ** 3 for b in self.boxes)

    def execute(self, blocking: bool = True) -> None:
        """the sequential form: every box's pipeline enqueued one after another"""
        for r in self.runs:
            r.execute(blocking=False)
        if blocking:
            self.ctx.sync()

    def capture(self) -> Graph:
        self.execute()  # lazy scratch / twiddle allocations happen here
        self.graph = capture(self.ctx, lambda: self.execute(blocking=False))
        return self.graph

    def launch(self, blocking: bool = True) -> None:
        if self.graph is None:
            self.capture()
        self.graph.launch(blocking)

    def close(self) -> None:
        if self.graph is not None:
            self.graph.close()
            self.graph = None
        for r in self.runs:
            r.close()
        self.runs = []
