"""Synthetic workloads of the benchmark configurations, generated on the
device from fixed seeds (wadjet_gen.h), with their operator specs."""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib as L
from .gpu import AggSpec, Col, Context, Pred

# --- k int64 over 1000 values, v int64 in [0, 1e6) -------------------
KV1K_TYPES = [L.INT64, L.INT64]
KV1K_AGGS = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1)]


def gen_kv1k(ctx: Context, n: int = 1_000_000, seed: int = 1):
    k = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, seed, 0, 999)
    v = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, seed + 1, 0, 999_999)
    return [Col(k, L.INT64), Col(v, L.INT64)]


# --- composite (a Zipf(1.1) over 1M, b uniform 16) GROUP BY ---------
# SUM(price DECIMAL(15,2)) into Int128, COUNT(*), MIN/MAX/AVG(f), WHERE f > -0.5
ZIPF_TYPES = [L.INT32, L.INT32, L.DECIMAL, L.FLOAT64]
ZIPF_SCALES = [0, 0, 2, 0]
ZIPF_KEYS = [0, 1]
ZIPF_AGGS = [AggSpec(L.SUM, 2), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 3), AggSpec(L.MAX, 3), AggSpec(L.AVG, 3)]
ZIPF_PRED = Pred(3, L.GT, 0, -0.5)
ZIPF_BYTES_PER_ROW = 4 + 4 + 8 + 8
ZIPF_BYTES_PER_GROUP = 56


def gen_zipf_groupby(ctx: Context, n: int = 100_000_000, seed: int = 2, row0: int = 0):
    a = ctx.gen(L.GEN_ZIPF_I32, "int32", n, seed, 1_000_000, 0, 1.1, row0=row0)
    b = ctx.gen(L.GEN_UNIFORM_I32, "int32", n, seed + 1, 0, 15, row0=row0)
    price = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, seed + 2, 100, 10_000_000, row0=row0)
    f = ctx.gen(L.GEN_NORMAL_F64, "float64", n, seed + 3, 0, 0, 0.0, 1.0, row0=row0)
    return [Col(a, L.INT32), Col(b, L.INT32), Col(price, L.DECIMAL, None, 2), Col(f, L.FLOAT64)]


# --- build 10M unique int64 keys, probe 200M fk (95% matching) -------
@dataclass
class FkJoin:
    n_build: int = 10_000_000
    n_probe: int = 200_000_000
    match_permille: int = 950
    seed: int = 3


FKJ_BUILD_TYPES = [L.INT64, L.INT32, L.DECIMAL]  # key, p1, p2 DECIMAL(12,2)
FKJ_BUILD_SCALES = [0, 0, 2]
FKJ_PROBE_TYPES = [L.INT64, L.INT32]  # fk, qty
FKJ_BYTES_PER_PROBE_ROW = 12
FKJ_BYTES_PER_BUILD_ROW = 20


def gen_fk_join_build(ctx: Context, cfg: FkJoin):
    key = ctx.gen(L.GEN_PERM_KEY_I64, "int64", cfg.n_build, cfg.seed, cfg.n_build)
    p1 = ctx.gen(L.GEN_UNIFORM_I32, "int32", cfg.n_build, cfg.seed + 1, 0, 24)
    p2 = ctx.gen(L.GEN_UNIFORM_I64, "int64", cfg.n_build, cfg.seed + 2, 0, 10**12 - 1)
    return [Col(key, L.INT64), Col(p1, L.INT32), Col(p2, L.DECIMAL, None, 2)]


def gen_fk_join_probe(ctx: Context, cfg: FkJoin, n: int = None, row0: int = 0):
    n = cfg.n_probe if n is None else n
    fk = ctx.gen(L.GEN_FK_I64, "int64", n, cfg.seed + 10, cfg.n_build, cfg.match_permille, row0=row0)
    qty = ctx.gen(L.GEN_UNIFORM_I32, "int32", n, cfg.seed + 11, 1, 50, row0=row0)
    return [Col(fk, L.INT64), Col(qty, L.INT32)]


# --- many concurrent GROUP BY / join queries of uneven shapes ---------
QMIX_DIM_ROWS = 50_000                     # the dimension the join queries probe (unique keys)
QMIX_DIM_GROUPS = (16, 256, 4096, 50_000)  # cardinalities of its four payload columns
QMIX_DIM_MATCH_PERMILLE = 900              # probe keys that find a dimension row


def plan_query_mix(n_queries: int = 256, seed: int = 4, max_rows: int = 4_000_000, min_rows: int = 1000,
                 median_rows: int = 65_536, sigma: float = 1.39, min_groups: int = 10, max_groups: int = 100_000,
                 join_every: int = 4):
    """Per query: rows ~ lognormal clipped to [min_rows, max_rows], groups
    log-uniform in [min_groups, max_groups], key type cycling int64 / int32 /
    uint32 (IPv4-like), value int64 or DECIMAL(18,2). Every join_every-th
    query (0: none) is JOIN dim ON key GROUP BY a dimension payload column
    instead: int64 foreign keys, the payload whose cardinality is nearest the
    drawn group count."""
    import numpy as np
    rng = np.random.default_rng(seed)
    rows = np.clip(np.exp(rng.normal(np.log(median_rows), sigma, n_queries)), min_rows, max_rows).astype(np.int64)
    groups = np.exp(rng.uniform(np.log(min_groups), np.log(max_groups), n_queries)).astype(np.int64)
    key_types = [(L.INT64, L.INT32, L.UINT32)[i % 3] for i in range(n_queries)]
    val_types = [(L.INT64, L.DECIMAL)[(i // 3) % 2] for i in range(n_queries)]
    out = []
    for i, (r, g, k, v) in enumerate(zip(rows, groups, key_types, val_types)):
        q = dict(rows=int(r), groups=int(g), key_type=k, val_type=v, seed=seed * 1000 + i, kind="groupby")
        if join_every and i % join_every == join_every - 1:
            pay = int(np.argmin([abs(np.log(c) - np.log(g)) for c in QMIX_DIM_GROUPS]))
            q.update(kind="join", key_type=L.INT64, payload=pay, groups=QMIX_DIM_GROUPS[pay])
        out.append(q)
    return out


def gen_query_dim(ctx: Context, seed: int = 77):
    """The dimension of the join queries: QMIX_DIM_ROWS unique int64 keys and
    one int32 payload column per QMIX_DIM_GROUPS cardinality, built into a
    HashJoin with every payload inline. Returns (HashJoin, device columns)."""
    from .gpu import HashJoin
    n = QMIX_DIM_ROWS
    cols = [Col(ctx.gen(L.GEN_PERM_KEY_I64, "int64", n, seed, n), L.INT64)]
    for j, g in enumerate(QMIX_DIM_GROUPS):
        cols.append(Col(ctx.gen(L.GEN_UNIFORM_I32, "int32", n, seed + 1 + j, 0, g - 1), L.INT32))
    jn = HashJoin(ctx, [L.INT64] + [L.INT32] * len(QMIX_DIM_GROUPS), 0, inline=list(range(1, 1 + len(QMIX_DIM_GROUPS))),
                  build_hint=n)
    jn.build(cols, n)
    jn.finish()
    return jn, cols


def gen_query_mix(ctx: Context, plan, dim=None):
    """Device columns of every planned query: list of (key Col, val Col, rows,
    est_groups[, extras]) — a join query's extras hold dim=(HashJoin, payload)
    and the dimension columns. `dim` (gen_query_dim's result) is built on
    demand."""
    out = []
    for q in plan:
        n, g, s = q["rows"], q["groups"], q["seed"]
        if q.get("kind") == "join":
            if dim is None:
                dim = gen_query_dim(ctx)
            k = ctx.gen(L.GEN_FK_I64, "int64", n, s, QMIX_DIM_ROWS, QMIX_DIM_MATCH_PERMILLE)
            v = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, s + 7, -10**9, 10**9) if q["val_type"] == L.INT64 else \
                ctx.gen(L.GEN_UNIFORM_I64, "int64", n, s + 7, 0, 10**12)
            vc = Col(v, L.INT64) if q["val_type"] == L.INT64 else Col(v, L.DECIMAL, None, 2)
            out.append((Col(k, L.INT64), vc, n, g, dict(dim=(dim[0], q["payload"]), dim_cols=dim[1],
                                                          payload=q["payload"])))
            continue
        if q["key_type"] == L.INT64:
            k = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, s, 0, g - 1)
        elif q["key_type"] == L.INT32:
            k = ctx.gen(L.GEN_UNIFORM_I32, "int32", n, s, -(g // 2), g - 1 - g // 2)
        else:
            k = ctx.gen(L.GEN_ZIPF_U32, "uint32", n, s, g, 0, 0.0)  # s = 0: uniform over g scrambled addresses
        if q["val_type"] == L.INT64:
            v = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, s + 7, -10**9, 10**9)
            vc = Col(v, L.INT64)
        else:
            v = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, s + 7, 0, 10**12)
            vc = Col(v, L.DECIMAL, None, 2)
        out.append((Col(k, q["key_type"]), vc, n, g))
    return out


# --- resident flow table served under Poisson load ------------------
FLOW_TYPES = [L.UINT32, L.UINT16, L.INT64, L.INT64]  # src_ip, dst_port, bytes, ts
FLOW_TS0 = 1_700_000_000_000_000_000
FLOW_TS_STEP = 1000
FLOW_DIM_TYPES = [L.UINT32, L.INT32]  # ip, asn


def gen_flow_table(ctx: Context, n: int = 500_000_000, seed: int = 5):
    src = ctx.gen(L.GEN_ZIPF_U32, "uint32", n, seed, 1_000_000, 0, 1.2)
    port = ctx.gen(L.GEN_UNIFORM_U16, "uint16", n, seed + 1, 0, 1023)
    nbytes = ctx.gen(L.GEN_LOGNORMAL_I64, "int64", n, seed + 2, 0, 0, 6.0, 1.5)
    ts = ctx.gen(L.GEN_SEQ_I64, "int64", n, seed + 3, FLOW_TS0, FLOW_TS_STEP)
    return [Col(src, L.UINT32), Col(port, L.UINT16), Col(nbytes, L.INT64), Col(ts, L.INT64)]


def gen_flow_dim(ctx: Context, n: int = 50_000, seed: int = 5, n_asn: int = 500):
    """Dimension keyed by the n most frequent source addresses, with an ASN."""
    ip = ctx.gen(L.GEN_RANK_U32, "uint32", n, seed)
    asn = ctx.gen(L.GEN_UNIFORM_I32, "int32", n, seed + 9, 0, n_asn - 1)
    return [Col(ip, L.UINT32), Col(asn, L.INT32)]


def plan_flow_requests(n_table: int, n_queries: int, seed: int = 55, scale: float = 1.0):
    """Mixed requests: GROUP BY dst_port / GROUP BY src_ip / JOIN dim GROUP BY asn,
    each over a random ts window; `scale` shrinks the windows for small tables."""
    import numpy as np
    from .gpu import Server
    rng = np.random.default_rng(seed)
    qs, shapes = [], []
    for i in range(n_queries):
        kind = rng.choice(3, p=[0.5, 0.3, 0.2])
        if kind == 0:
            w = int(rng.integers(1_000_000, 4_000_000) * scale)
            spec = (0, 1, 2, 1024)
        elif kind == 1:
            w = int(rng.integers(250_000, 1_000_000) * scale)
            spec = (0, 0, 2, max(w // 2, 64))
        else:
            w = int(rng.integers(1_000_000, 4_000_000) * scale)
            spec = (1, 0, 2, 512)
        w = max(1000, min(w, n_table - 1))
        start = int(rng.integers(0, n_table - w))
        lo = FLOW_TS0 + start * FLOW_TS_STEP + int(rng.integers(0, FLOW_TS_STEP))
        hi = lo + w * FLOW_TS_STEP
        k, key_col, val_col, est = spec
        qs.append(Server.query(int(k), key_col, val_col, lo, hi, est, dim=0, dim_payload=0))
        shapes.append(dict(kind=int(k), rows=w, start=start))
    return qs, shapes
