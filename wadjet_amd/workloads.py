"""Synthetic workloads of the benchmark configurations, generated on the
device from fixed seeds (wadjet_gen.h), with their operator specs."""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib as L
from .gpu import AggSpec, Col, Context, Pred

# --- config 1: k int64 over 1000 values, v int64 in [0, 1e6) -------------------
CFG1_TYPES = [L.INT64, L.INT64]
CFG1_AGGS = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1)]


def gen_config1(ctx: Context, n: int = 1_000_000, seed: int = 1):
    k = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, seed, 0, 999)
    v = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, seed + 1, 0, 999_999)
    return [Col(k, L.INT64), Col(v, L.INT64)]


# --- config 2: composite (a Zipf(1.1) over 1M, b uniform 16) GROUP BY ---------
# SUM(price DECIMAL(15,2)) into Int128, COUNT(*), MIN/MAX/AVG(f), WHERE f > -0.5
CFG2_TYPES = [L.INT32, L.INT32, L.DECIMAL, L.FLOAT64]
CFG2_SCALES = [0, 0, 2, 0]
CFG2_KEYS = [0, 1]
CFG2_AGGS = [AggSpec(L.SUM, 2), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 3), AggSpec(L.MAX, 3), AggSpec(L.AVG, 3)]
CFG2_PRED = Pred(3, L.GT, 0, -0.5)
CFG2_BYTES_PER_ROW = 4 + 4 + 8 + 8
CFG2_BYTES_PER_GROUP = 56


def gen_config2(ctx: Context, n: int = 100_000_000, seed: int = 2, row0: int = 0):
    a = ctx.gen(L.GEN_ZIPF_I32, "int32", n, seed, 1_000_000, 0, 1.1, row0=row0)
    b = ctx.gen(L.GEN_UNIFORM_I32, "int32", n, seed + 1, 0, 15, row0=row0)
    price = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, seed + 2, 100, 10_000_000, row0=row0)
    f = ctx.gen(L.GEN_NORMAL_F64, "float64", n, seed + 3, 0, 0, 0.0, 1.0, row0=row0)
    return [Col(a, L.INT32), Col(b, L.INT32), Col(price, L.DECIMAL, None, 2), Col(f, L.FLOAT64)]


# --- config 3: build 10M unique int64 keys, probe 200M fk (95% matching) -------
@dataclass
class Cfg3:
    n_build: int = 10_000_000
    n_probe: int = 200_000_000
    match_permille: int = 950
    seed: int = 3


CFG3_BUILD_TYPES = [L.INT64, L.INT32, L.DECIMAL]  # key, p1, p2 DECIMAL(12,2)
CFG3_BUILD_SCALES = [0, 0, 2]
CFG3_PROBE_TYPES = [L.INT64, L.INT32]  # fk, qty
CFG3_BYTES_PER_PROBE_ROW = 12
CFG3_BYTES_PER_BUILD_ROW = 20


def gen_config3_build(ctx: Context, cfg: Cfg3):
    key = ctx.gen(L.GEN_PERM_KEY_I64, "int64", cfg.n_build, cfg.seed, cfg.n_build)
    p1 = ctx.gen(L.GEN_UNIFORM_I32, "int32", cfg.n_build, cfg.seed + 1, 0, 24)
    p2 = ctx.gen(L.GEN_UNIFORM_I64, "int64", cfg.n_build, cfg.seed + 2, 0, 10**12 - 1)
    return [Col(key, L.INT64), Col(p1, L.INT32), Col(p2, L.DECIMAL, None, 2)]


def gen_config3_probe(ctx: Context, cfg: Cfg3, n: int = None, row0: int = 0):
    n = cfg.n_probe if n is None else n
    fk = ctx.gen(L.GEN_FK_I64, "int64", n, cfg.seed + 10, cfg.n_build, cfg.match_permille, row0=row0)
    qty = ctx.gen(L.GEN_UNIFORM_I32, "int32", n, cfg.seed + 11, 1, 50, row0=row0)
    return [Col(fk, L.INT64), Col(qty, L.INT32)]
