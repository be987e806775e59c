"""Thin Python handles over the C-ABI: device buffers, the grouped aggregate,
the filter, the hash join and the generators. Used by the tests and bench.py;
numpy arrays are the host-side representation.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib as L
from ._lib import call

NP_OF = {L.INT32: np.int32, L.INT64: np.int64, L.FLOAT64: np.float64, L.DECIMAL: np.int64, L.IPV4: np.int64,
         L.UINT32: np.uint32, L.UINT16: np.uint16}


class Context:
    """One device + one HIP stream. Closing it first closes every handle and
    buffer still alive on it (the library frees them against this stream)."""

    def __init__(self, device: int = 0):
        import weakref
        h = C.c_void_p()
        call("wg_ctx_create", device, C.byref(h))
        self.h = h
        self._children = weakref.WeakSet()
        cus, hbm, lds = C.c_int32(), C.c_uint64(), C.c_int32()
        call("wg_ctx_info", h, C.byref(cus), C.byref(hbm), C.byref(lds))
        self.cus, self.hbm, self.lds = cus.value, hbm.value, lds.value

    def close(self):
        if self.h:
            for c in list(self._children):
                try:
                    c.close()
                except Exception:
                    pass
            call("wg_ctx_destroy", self.h)
            self.h = None

    def _adopt(self, child):
        self._children.add(child)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        call("wg_ctx_sync", self.h)

    def inject_alloc_failure(self, nth: int, sticky: bool = False):
        """The nth device allocation made through this context from now on
        fails as out of memory (and every later one when sticky); 0 clears."""
        call("wg_ctx_inject_alloc_failure", self.h, nth, 1 if sticky else 0)

    def memset(self, buf: "DeviceBuffer", value: int, nbytes: Optional[int] = None):
        call("wg_memset", self.h, buf.ptr, value, buf.nbytes if nbytes is None else nbytes)

    def timer_start(self):
        call("wg_ctx_timer_start", self.h)

    def timer_stop(self) -> float:
        ms = C.c_float()
        call("wg_ctx_timer_stop", self.h, C.byref(ms))
        return ms.value

    # -- memory
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, max(arr.nbytes, 1))
        if arr.nbytes:
            call("wg_memcpy_h2d", self.h, b.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes)
        return b

    def copy(self, dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int):
        """Device-to-device copy through the library's streaming copy kernel."""
        call("wg_memcpy_d2d", self.h, dst.ptr, src.ptr, nbytes)

    def gen(self, kind: int, dtype, n: int, seed: int, i0: int = 0, i1: int = 0, f0: float = 0.0, f1: float = 0.0,
            row0: int = 0, out: Optional["DeviceBuffer"] = None) -> "DeviceBuffer":
        nbytes = n * np.dtype(dtype).itemsize
        b = out if out is not None else DeviceBuffer(self, max(nbytes, 1))
        call("wg_gen", self.h, kind, b.ptr, row0, n, seed, i0, i1, f0, f1)
        return b


class DeviceBuffer:
    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        p = C.c_void_p()
        call("wg_malloc", ctx.h, nbytes, C.byref(p))
        self.ptr = p
        ctx._adopt(self)

    def close(self):
        self.free()

    def to_host(self, dtype, count: Optional[int] = None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dtype=dt)
        if out.nbytes:
            call("wg_memcpy_d2h", self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes)
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            call("wg_free", self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                L.lib().wg_free(self.ctx.h, self.ptr)
        except Exception:
            pass


class BorrowedBuffer:
    """Device memory the library owns (a join's retained build column):
    readable through the same calls as a DeviceBuffer, never freed here."""

    def __init__(self, ctx: Context, ptr, nbytes: int):
        self.ctx = ctx
        self.ptr = C.c_void_p(ptr)
        self.nbytes = nbytes

    to_host = DeviceBuffer.to_host

    def free(self):
        pass


TYPE_BYTES = {L.INT32: 4, L.INT64: 8, L.FLOAT64: 8, L.DECIMAL: 8, L.DECIMAL128: 16, L.IPV4: 8, L.UINT32: 4,
              L.UINT16: 2}


@dataclass
class Col:
    """A column handed to the library: device buffer or host numpy array."""
    data: object  # DeviceBuffer | np.ndarray
    type: int
    validity: object = None  # DeviceBuffer | np.ndarray(uint64) | None
    scale: int = 0

    def c(self) -> L.Column:
        def ptr(x):
            if x is None:
                return None
            if isinstance(x, (DeviceBuffer, BorrowedBuffer)):
                return x.ptr
            return x.ctypes.data_as(C.c_void_p)
        return L.Column(ptr(self.data), ptr(self.validity), self.type, self.scale)


def columns(cols: Sequence[Col]):
    arr = (L.Column * len(cols))(*[c.c() for c in cols])
    return arr


@dataclass
class AggSpec:
    func: int
    col: int = -1
    col2: int = -1


@dataclass
class Pred:
    col: int
    op: int
    i64: int = 0
    f64: float = 0.0


def bitmap_from_mask(valid: np.ndarray) -> np.ndarray:
    """batch.Bitmap words (bit i%64 of word i/64, 1 = valid) from a bool mask."""
    n = len(valid)
    words = np.zeros((n + 63) // 64, dtype=np.uint64)
    idx = np.nonzero(valid)[0]
    np.bitwise_or.at(words, idx >> 6, (np.uint64(1) << (idx & 63).astype(np.uint64)))
    return words


class HashAggregate:
    """GROUP BY keys with SUM/COUNT/MIN/MAX/AVG on the GPU (wg_agg_*)."""

    def __init__(self, ctx: Context, types: Sequence[int], keys: Sequence[int], aggs: Sequence[AggSpec],
                 scales: Optional[Sequence[int]] = None, nullable: Optional[Sequence[bool]] = None,
                 group_hint: int = 0):
        self.ctx = ctx
        self.types = list(types)
        self.keys = list(keys)
        self.aggs = list(aggs)
        n = len(types)
        t = (C.c_int32 * n)(*types)
        s = (C.c_int32 * n)(*(scales or [0] * n))
        nl = (C.c_uint8 * n)(*[1 if x else 0 for x in (nullable or [False] * n)])
        k = (C.c_int32 * len(keys))(*keys)
        a = (L.AggSpecC * max(len(aggs), 1))(*[L.AggSpecC(x.func, x.col, x.col2, 0) for x in aggs])
        h = C.c_void_p()
        call("wg_agg_create", ctx.h, n, t, s, nl, len(keys), k, len(aggs), a, group_hint, C.byref(h))
        self.h = h
        ctx._adopt(self)

    def close(self):
        if self.h and self.ctx.h:
            call("wg_agg_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                L.lib().wg_agg_destroy(self.h)
        except Exception:
            pass

    def consume(self, cols: Sequence[Col], n_rows: int, sel=None, pred: Optional[Pred] = None, device: bool = True):
        """sel: None, a host uint32 array, a DeviceBuffer, or (DeviceBuffer, count)."""
        arr = columns(cols)
        keep = [arr]
        if sel is None:
            sp, ns = None, 0
        elif isinstance(sel, tuple):
            sp, ns = sel[0].ptr, int(sel[1])
        elif isinstance(sel, DeviceBuffer):
            sp, ns = sel.ptr, sel.nbytes // 4
        else:
            sel = np.ascontiguousarray(sel, dtype=np.uint32)
            keep.append(sel)
            sp, ns = sel.ctypes.data_as(C.c_void_p), len(sel)
        preds = [] if pred is None else (list(pred) if isinstance(pred, (list, tuple)) else [pred])
        pa = (L.PredicateC * max(len(preds), 1))(*[L.PredicateC(p.col, p.op, p.i64, p.f64) for p in preds])
        call("wg_agg_consume", self.h, arr, n_rows, sp, ns, pa, len(preds), L.MEM_DEVICE if device else L.MEM_HOST)

    def merge(self, other: "HashAggregate"):
        call("wg_agg_merge", self.h, other.h)

    def finalize(self):
        call("wg_agg_finalize", self.h)

    def reset(self):
        call("wg_agg_reset", self.h)

    def set_tuning(self, lds_slots: int, block: int):
        call("wg_agg_set_tuning", self.h, lds_slots, block)

    def num_groups(self) -> int:
        n = C.c_uint64()
        call("wg_agg_num_groups", self.h, C.byref(n))
        return n.value

    def footprint(self) -> int:
        n = C.c_uint64()
        call("wg_agg_footprint", self.h, C.byref(n))
        return n.value

    def state_bytes(self) -> int:
        n = C.c_uint64()
        call("wg_agg_state_bytes", self.h, C.byref(n))
        return n.value

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        call("wg_agg_stats", self.h, *[C.byref(x) for x in v])
        return dict(lds_rows=v[0].value, global_rows=v[1].value, flushed=v[2].value, capacity=v[3].value)

    def set_strategy(self, strategy: int, part_lds_slots: int = 2048, part_block: int = 1024,
                     part_chunk_rows: int = 1 << 16):
        call("wg_agg_set_strategy", self.h, strategy, part_lds_slots, part_block, part_chunk_rows)

    def set_heavy_hitters(self, hot_slots: int):
        """Heavy-hitter slots of the partitioned consume (0: none)."""
        call("wg_agg_set_heavy_hitters", self.h, hot_slots)

    def spill_run(self, names=None, type_ids=None) -> bytes:
        """The finalized groups as one partial-state spill run (bytes)."""
        k = len(self.keys)
        nm = None if names is None else (C.c_char_p * k)(*[x.encode() for x in names])
        ti = None if type_ids is None else (C.c_uint16 * k)(*type_ids)
        size = C.c_uint64()
        call("wg_agg_spill_run", self.h, nm, ti, None, 0, C.byref(size))
        buf = C.create_string_buffer(max(size.value, 1))
        call("wg_agg_spill_run", self.h, nm, ti, buf, size.value, C.byref(size))
        return buf.raw[:size.value]

    def merge_run(self, run: bytes):
        """Merge a partial-state spill run into this (not finalized) aggregate."""
        buf = C.create_string_buffer(run, len(run))
        call("wg_agg_merge_run", self.h, buf, len(run))

    def pass_times(self):
        a = (C.c_float * 4)()
        call("wg_agg_pass_times", self.h, a)
        return list(a)

    def kernel_time(self):
        ms, k = C.c_float(), C.c_uint32()
        call("wg_agg_kernel_time", self.h, C.byref(ms), C.byref(k))
        return ms.value, k.value

    def fetch(self) -> dict:
        """All groups to host: keys (list of arrays), key_valid, values, valid."""
        n = self.num_groups()
        keys = [np.empty(n, dtype=NP_OF[self.types[k]]) for k in self.keys]
        vals, valid = [], []
        for a in self.aggs:
            is_f = a.func in (L.SUM, L.AVG) or (a.func in (L.MIN, L.MAX) and self.types[a.col] == L.FLOAT64)
            vals.append(np.empty(n, dtype=np.float64 if is_f else np.int64))
            valid.append(np.empty(n, dtype=np.uint8))
        kp = (C.c_void_p * max(len(keys), 1))(*[k.ctypes.data_as(C.c_void_p) for k in keys])
        vp = (C.c_void_p * max(len(vals), 1))(*[v.ctypes.data_as(C.c_void_p) for v in vals])
        vv = (C.c_void_p * max(len(valid), 1))(*[v.ctypes.data_as(C.c_void_p) for v in valid])
        if n:
            call("wg_agg_fetch", self.h, 0, n, kp, vp, vv, L.MEM_HOST)
        kvalid = np.empty(n, dtype=np.uint8)
        if n:
            call("wg_agg_fetch_key_valid", self.h, 0, n, kvalid.ctypes.data_as(C.c_void_p), L.MEM_HOST)
        return dict(n=n, keys=keys, key_valid=kvalid.astype(bool), values=vals, valid=[v.astype(bool) for v in valid])

    def output_buffers(self, n: int):
        """Device buffers for n groups of every output column: (keys, values, valid)."""
        ks = [self.ctx.alloc(max(n, 1) * np.dtype(NP_OF[self.types[k]]).itemsize) for k in self.keys]
        vs = [self.ctx.alloc(max(n, 1) * 8) for _ in self.aggs]
        vv = [self.ctx.alloc(max(n, 1)) for _ in self.aggs]
        return ks, vs, vv

    def fetch_device(self, keys, vals, valid, offset: int = 0, count: Optional[int] = None):
        """Groups [offset, offset + count) into device buffers (None skips a
        column): keys in their input type, values as float64 / int64, validity
        bytes; one pass over the groups (the Next drain on the device)."""
        count = self.num_groups() - offset if count is None else count
        ptr = (lambda b: None if b is None else b.ptr)
        kp = (C.c_void_p * max(len(keys), 1))(*[ptr(k) for k in keys])
        vp = (C.c_void_p * max(len(vals), 1))(*[ptr(v) for v in vals])
        vv = (C.c_void_p * max(len(valid), 1))(*[ptr(v) for v in valid])
        if count:
            call("wg_agg_fetch", self.h, offset, count, kp, vp, vv, L.MEM_DEVICE)

    def fetch_state(self, agg: int) -> dict:
        n = self.num_groups()
        cnt = np.empty(n, np.int64)
        lo = np.empty(n, np.uint64)
        hi = np.empty(n, np.int64)
        mm = np.empty(n, np.int64)
        if n:
            call("wg_agg_fetch_state", self.h, agg, 0, n, *(x.ctypes.data_as(C.c_void_p) for x in (cnt, lo, hi, mm)),
                 L.MEM_HOST)
        return dict(count=cnt, sum_lo=lo, sum_hi=hi, minmax=mm)


def filter_sel(ctx: Context, col: Col, n_rows: int, op: int, i64: int = 0, f64: float = 0.0,
               sel: Optional[DeviceBuffer] = None, n_sel: int = 0, out: Optional[DeviceBuffer] = None):
    """KernelFilter on the device: returns (device sel buffer, count)."""
    n = n_sel if sel is not None else n_rows
    out = out if out is not None else ctx.alloc(max(n, 1) * 4)
    nout = C.c_uint64()
    cc = col.c()
    call("wg_filter", ctx.h, C.byref(cc), n_rows, sel.ptr if sel is not None else None, n_sel, op, i64, f64,
         out.ptr, C.byref(nout))
    return out, nout.value


class HashJoin:
    """Inner hash join on one int-class key (wg_join_*): build, then probe."""

    def __init__(self, ctx: Context, build_types: Sequence[int], build_key, inline: Sequence[int] = (),
                 scales: Optional[Sequence[int]] = None, build_hint: int = 0):
        """build_key: a column index, or a pair of them (two-column key)."""
        self.ctx = ctx
        self.types = list(build_types)
        n = len(build_types)
        t = (C.c_int32 * n)(*build_types)
        s = (C.c_int32 * n)(*(scales or [0] * n))
        il = (C.c_int32 * max(len(inline), 1))(*inline)
        h = C.c_void_p()
        self.dual = isinstance(build_key, (tuple, list))
        if self.dual:
            call("wg_join_create_dual", ctx.h, n, t, s, build_key[0], build_key[1], len(inline), il, build_hint,
                 C.byref(h))
        else:
            call("wg_join_create", ctx.h, n, t, s, build_key, len(inline), il, build_hint, C.byref(h))
        self.h = h
        ctx._adopt(self)

    def dual_probe_key(self, key0: Col, key1: Col, n_rows: int, sel=None) -> Col:
        """The id column a two-column-key join probes with (wg_join_dual_probe_key)."""
        ids = self.ctx.alloc(max(n_rows, 1) * 8)
        valid = self.ctx.alloc(max((n_rows + 63) // 64, 1) * 8)
        sp, ns = (None, 0) if sel is None else (sel[0].ptr, int(sel[1]))
        k0, k1 = key0.c(), key1.c()
        call("wg_join_dual_probe_key", self.h, C.byref(k0), C.byref(k1), n_rows, sp, ns, ids.ptr, valid.ptr)
        return Col(ids, L.INT64, valid)

    def close(self):
        if self.h and self.ctx.h:
            call("wg_join_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                L.lib().wg_join_destroy(self.h)
        except Exception:
            pass

    def build(self, cols: Sequence[Col], n_rows: int, device: bool = True):
        call("wg_join_build_append", self.h, columns(cols), n_rows, L.MEM_DEVICE if device else L.MEM_HOST)

    def finish(self):
        call("wg_join_build_finish", self.h)

    def rows(self) -> int:
        n = C.c_uint64()
        call("wg_join_build_rows", self.h, C.byref(n))
        return n.value

    def build_column(self, col: int) -> Col:
        """The retained build column as a device column (the base of the
        build side's view columns); owned by the join."""
        c = L.Column()
        call("wg_join_build_column", self.h, col, C.byref(c))
        n = self.rows()
        data = BorrowedBuffer(self.ctx, c.data, n * TYPE_BYTES[c.type])
        valid = BorrowedBuffer(self.ctx, c.validity, ((n + 63) // 64) * 8) if c.validity else None
        return Col(data, c.type, valid, c.scale)

    def probe(self, key: Col, n_rows: int, out_p: DeviceBuffer, out_b: DeviceBuffer, out_cap: int,
              cursor: Optional[L.ProbeCursor] = None, sel=None):
        """One bounded probe call; returns (pairs written, cursor)."""
        cur = cursor if cursor is not None else L.ProbeCursor(0, -1, 0, 0)
        n = C.c_uint64()
        kc = key.c()
        sp, ns = (None, 0) if sel is None else (sel[0].ptr, int(sel[1]))
        call("wg_join_probe", self.h, C.byref(kc), n_rows, sp, ns, out_p.ptr, out_b.ptr, out_cap, C.byref(cur),
             C.byref(n))
        return n.value, cur

    def probe_agg(self, probe_cols: Sequence[Col], probe_key: int, n_rows: int, agg: HashAggregate, sel=None):
        sp, ns = (None, 0) if sel is None else (sel[0].ptr, int(sel[1]))
        call("wg_join_probe_agg", self.h, columns(probe_cols), len(probe_cols), probe_key, n_rows, sp, ns, agg.h)

    def set_tuning(self, lean: bool = True, block: int = 512, replicas: int = 0):
        """Kernel choice for probe_agg (results do not depend on it)."""
        call("wg_join_set_tuning", self.h, 1 if lean else 0, block, replicas)

    def semi(self, key: Col, n_rows: int, out: DeviceBuffer, anti: bool = False, sel=None) -> int:
        """Semi (anti=False) / anti join: positions written to `out`, returns the count."""
        n = C.c_uint64()
        kc = key.c()
        sp, ns = (None, 0) if sel is None else (sel[0].ptr, int(sel[1]))
        call("wg_join_semi", self.h, C.byref(kc), n_rows, sp, ns, 1 if anti else 0, out.ptr, C.byref(n))
        return n.value

    def bloom_build(self):
        call("wg_join_bloom_build", self.h)

    def bloom_filter(self, key: Col, n_rows: int, out: DeviceBuffer, sel=None) -> int:
        """Probe positions whose key may be in the build (device sel in `out`)."""
        n = C.c_uint64()
        kc = key.c()
        sp, ns = (None, 0) if sel is None else (sel[0].ptr, int(sel[1]))
        call("wg_join_bloom_filter", self.h, C.byref(kc), n_rows, sp, ns, out.ptr, C.byref(n))
        return n.value

    def bloom_words(self) -> np.ndarray:
        w = C.c_uint64()
        call("wg_join_bloom_words", self.h, C.byref(w), None, 0)
        out = np.empty(w.value, np.uint64)
        call("wg_join_bloom_words", self.h, C.byref(w), out.ctypes.data_as(C.c_void_p), w.value)
        return out

    def gather_build(self, col: int, rows: DeviceBuffer, n: int, out: DeviceBuffer):
        call("wg_join_gather_build", self.h, col, rows.ptr, n, out.ptr)

    def digest(self, probe_cols: Sequence[Col], prow: DeviceBuffer, brow: DeviceBuffer, n: int,
               probe_sel: Sequence[int], build_sel: Sequence[int]):
        d = (C.c_uint64 * 2)()
        ps = (C.c_int32 * max(len(probe_sel), 1))(*probe_sel)
        bs = (C.c_int32 * max(len(build_sel), 1))(*build_sel)
        call("wg_join_digest", self.h, columns(probe_cols), prow.ptr, brow.ptr, n, ps, len(probe_sel), bs,
             len(build_sel), d)
        return int(d[0]), int(d[1])


def gather(ctx: Context, col: Col, rows: DeviceBuffer, n: int, out: DeviceBuffer):
    cc = col.c()
    call("wg_gather", ctx.h, C.byref(cc), rows.ptr, n, out.ptr)


def view_flatten(ctx: Context, base: Col, base_rows: int, indices: DeviceBuffer, n: int,
                 own_valid: Optional[DeviceBuffer] = None, out: Optional[DeviceBuffer] = None,
                 out_valid: Optional[DeviceBuffer] = None) -> Col:
    """A view column (base, indices, own NULL bitmap) materialized on the
    device (wg_view_flatten): Col of n rows with its validity words."""
    w = TYPE_BYTES[base.type]
    out = out if out is not None else ctx.alloc(max(n, 1) * w)
    out_valid = out_valid if out_valid is not None else ctx.alloc(max((n + 63) // 64, 1) * 8)
    cc = base.c()
    call("wg_view_flatten", ctx.h, C.byref(cc), base_rows, indices.ptr, own_valid.ptr if own_valid is not None else None,
         n, out.ptr, out_valid.ptr)
    return Col(out, base.type, out_valid, base.scale)


def shuffle_partition(ctx: Context, cols: Sequence[Col], key_cols: Sequence[int], num_parts: int, n_rows: int,
                      sel: Optional[DeviceBuffer] = None, n_sel: int = 0, want_parts: bool = True,
                      parts: Optional[DeviceBuffer] = None, rows: Optional[DeviceBuffer] = None):
    """Routing of a partitioned shuffle sink on the device (wg_shuffle_partition):
    returns (parts device buffer or None, rows device buffer, offsets host
    int64 array of num_parts + 1). parts / rows may be passed in (n active
    entries of 4 bytes each)."""
    n = n_sel if sel is not None else n_rows
    if parts is None and want_parts:
        parts = ctx.alloc(max(n, 1) * 4)
    if rows is None:
        rows = ctx.alloc(max(n, 1) * 4)
    offsets = np.zeros(max(num_parts, 1) + 1, np.uint64)
    arr = columns(cols)
    k = (C.c_int32 * max(len(key_cols), 1))(*key_cols)
    call("wg_shuffle_partition", ctx.h, arr, len(cols), k, len(key_cols), n_rows,
         sel.ptr if sel is not None else None, n_sel, num_parts, parts.ptr if parts is not None else None, rows.ptr,
         offsets.ctypes.data_as(C.c_void_p))
    return parts, rows, offsets.astype(np.int64)


class Batch:
    """Many GROUP BY key SUM(val), COUNT(*) queries in one persistent launch."""

    def __init__(self, ctx: Context, max_queries: int, max_groups_total: int):
        self.ctx = ctx
        h = C.c_void_p()
        call("wg_batch_create", ctx.h, max_queries, max_groups_total, C.byref(h))
        self.h = h
        ctx._adopt(self)

    def close(self):
        if self.h and self.ctx.h:
            call("wg_batch_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                L.lib().wg_batch_destroy(self.h)
        except Exception:
            pass

    @staticmethod
    def queries(qs):
        """qs: list of (key Col, val Col, n_rows, est_groups[, extras dict]) where
        extras may hold row_begin, range=(Col, lo, hi), dim=(HashJoin, payload)."""
        arr = (L.BatchQuery * max(len(qs), 1))()
        for i, q in enumerate(qs):
            k, v, n, g = q[:4]
            x = q[4] if len(q) > 4 else {}
            bq = L.BatchQuery()
            bq.key, bq.val, bq.n_rows, bq.est_groups = k.c(), v.c(), n, g
            bq.row_begin = x.get("row_begin", 0)
            if "range" in x:
                rc, lo, hi = x["range"]
                bq.range_col, bq.range_lo, bq.range_hi = rc.c(), lo, hi
            if "dim" in x:
                bq.dim, bq.dim_payload = x["dim"][0].h, x["dim"][1]
            arr[i] = bq
        return arr

    def run(self, qarr, n: int, use_graph: bool = True, plain: bool = False):
        call("wg_batch_run", self.h, qarr, n, 1 if use_graph else 0, 1 if plain else 0)

    def autotune(self, qarr, n: int):
        call("wg_batch_autotune", self.h, qarr, n)

    def tuning(self):
        b, t, r, f, c = C.c_int32(), C.c_int32(), C.c_uint64(), C.c_int32(), C.c_int32()
        call("wg_batch_tuning", self.h, C.byref(b), C.byref(t), C.byref(r), C.byref(f), C.byref(c))
        return dict(block=b.value, lds_threshold=t.value, tile_rows=r.value, max_fanout=f.value, lds_cache=c.value)

    def set_tuning(self, block: int, lds_threshold: int, tile_rows: int, max_fanout: int = 1, lds_cache: int = 1):
        call("wg_batch_set_tuning", self.h, block, lds_threshold, tile_rows, max_fanout, lds_cache)

    def choice(self, q: int):
        """(layout, workgroup size, key partitions per tile) of query q; layout
        1 = LDS table(s), 2 = global table behind an LDS cache, 0 = global."""
        u, b, f = C.c_int32(), C.c_int32(), C.c_int32()
        call("wg_batch_choice", self.h, q, C.byref(u), C.byref(b), C.byref(f))
        return u.value, b.value, f.value

    def result(self, q: int) -> dict:
        n = C.c_uint64()
        call("wg_batch_result", self.h, q, C.byref(n), None, None, None, None, 0)
        k = np.empty(n.value, np.int64)
        lo = np.empty(n.value, np.uint64)
        hi = np.empty(n.value, np.int64)
        c = np.empty(n.value, np.int64)
        call("wg_batch_result", self.h, q, C.byref(n), *(x.ctypes.data_as(C.c_void_p) for x in (k, lo, hi, c)), n.value)
        o = np.argsort(k, kind="stable")
        return dict(n=n.value, keys=k[o], sum_lo=lo[o], sum_hi=hi[o], counts=c[o])


class Server:
    """Continuous-batching request loop over device-resident columns (wg_serve_*)."""

    def __init__(self, ctx: Context, cols: Sequence[Col], n_rows: int, ts_col: int, lanes: int = 3,
                 hbm_budget: int = 4 << 30, max_batch: int = 64, plain: bool = False):
        self.ctx = ctx
        self.cols = list(cols)
        h = C.c_void_p()
        call("wg_serve_create", ctx.h, columns(cols), len(cols), n_rows, ts_col, lanes, hbm_budget, max_batch,
             1 if plain else 0, C.byref(h))
        self.h = h
        ctx._adopt(self)

    def close(self):
        if self.h and self.ctx.h:
            call("wg_serve_destroy", self.h)
        self.h = None

    def __del__(self):
        try:
            if self.h and self.ctx.h:
                L.lib().wg_serve_destroy(self.h)
        except Exception:
            pass

    def add_dim(self, j: "HashJoin") -> int:
        i = C.c_int32()
        call("wg_serve_add_dim", self.h, j.h, C.byref(i))
        return i.value

    @staticmethod
    def query(kind, key_col, val_col, ts_lo, ts_hi, est_groups, dim=0, dim_payload=0):
        return L.ServeQuery(kind, key_col, val_col, dim, dim_payload, 0, ts_lo, ts_hi, est_groups)

    def submit(self, q) -> int:
        i = C.c_uint64()
        call("wg_serve_submit", self.h, C.byref(q), C.byref(i))
        return i.value

    def wait(self, rid: int):
        n, ms = C.c_uint64(), C.c_double()
        call("wg_serve_wait", self.h, rid, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def result(self, rid: int) -> dict:
        n, _ = self.wait(rid)
        k = np.empty(n, np.int64)
        lo = np.empty(n, np.uint64)
        hi = np.empty(n, np.int64)
        c = np.empty(n, np.int64)
        call("wg_serve_result", self.h, rid, *(x.ctypes.data_as(C.c_void_p) for x in (k, lo, hi, c)), n)
        o = np.argsort(k, kind="stable")
        return dict(n=n, keys=k[o], sum_lo=lo[o], sum_hi=hi[o], counts=c[o])

    def replay(self, qs, qps: float, seed: int = 0):
        n = len(qs)
        arr = (L.ServeQuery * n)(*qs)
        lat = np.empty(n, np.float64)
        ids = np.empty(n, np.uint64)
        call("wg_serve_replay", self.h, arr, n, qps, seed, lat.ctypes.data_as(C.c_void_p),
             ids.ctypes.data_as(C.c_void_p))
        return lat, ids

    def stats(self) -> dict:
        st = L.ServeStats()
        call("wg_serve_stats_get", self.h, C.byref(st))
        return {f: getattr(st, f) for f, _ in L.ServeStats._fields_}

    def reset(self):
        call("wg_serve_reset", self.h)
