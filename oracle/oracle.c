/*
 * oracle.c — CPU restatement of wadjet's grouped aggregation, Int128 decimal
 * arithmetic, typed filter and inner hash join (the semantics the HIP library
 * must reproduce). TEST INFRASTRUCTURE ONLY: linked by tests/, smoke() and the
 * bench's CPU-baseline leg, never by the product.
 *
 * Reference files restated (paths relative to the wadjet repository):
 *   internal/engine/batch/decimal.go:17-187         Int128 From/Add/Neg/Less/ToFloat64/FromFloat64/Format/Parse
 *   internal/engine/exec/agg_scatter.go:178-578     SoA scatter kernels and their dispatch
 *   internal/engine/exec/aggregate.go:5982-6119     count sharing + accumulator field choice (AVG(int64) -> float64)
 *   internal/engine/exec/aggregate_partial_spill.go:1855-1945   writeAccToColumn (finalization)
 *   internal/engine/exec/kernel/types.go:106-175    Accumulator.Merge
 *   internal/engine/exec/kernel/compare.go:47-178   typed filter kernels
 *   internal/engine/exec/join.go:551-560,2670-2836  arenaAppendInt (LIFO chains) and inlineIntProbe
 */
#define _GNU_SOURCE /* qsort_r */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------- Int128 */

/* decimal.go:17-23 */
void or_i128_from_i64(int64_t v, uint64_t* lo, int64_t* hi) {
  *hi = v < 0 ? -1 : 0;
  *lo = (uint64_t)v;
}

/* decimal.go:77-85 */
void or_i128_add(uint64_t alo, int64_t ahi, uint64_t blo, int64_t bhi, uint64_t* lo, int64_t* hi) {
  uint64_t l = alo + blo;
  int64_t h = (int64_t)((uint64_t)ahi + (uint64_t)bhi);
  if (l < alo) h = (int64_t)((uint64_t)h + 1);
  *lo = l;
  *hi = h;
}

/* decimal.go:61-69 */
void or_i128_neg(uint64_t lo, int64_t hi, uint64_t* olo, int64_t* ohi) {
  uint64_t l = ~lo + 1;
  int64_t h = ~hi;
  if (l == 0) h = (int64_t)((uint64_t)h + 1);
  *olo = l;
  *ohi = h;
}

/* decimal.go:93-98 */
int or_i128_less(uint64_t alo, int64_t ahi, uint64_t blo, int64_t bhi) {
  if (ahi != bhi) return ahi < bhi;
  return alo < blo;
}

/* math.Pow10: exact table for 0..31 (pow10tab), products of the 1e32 powers
 * (pow10postab32) above that. */
static const double pow10tab[32] = {1e00, 1e01, 1e02, 1e03, 1e04, 1e05, 1e06, 1e07, 1e08, 1e09, 1e10,
                                    1e11, 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21,
                                    1e22, 1e23, 1e24, 1e25, 1e26, 1e27, 1e28, 1e29, 1e30, 1e31};
static const double pow10postab32[10] = {1e00, 1e32, 1e64, 1e96, 1e128, 1e160, 1e192, 1e224, 1e256, 1e288};
static double go_pow10(int n) {
  if (n >= 0 && n <= 308) return pow10postab32[n / 32] * pow10tab[n % 32];
  if (n < 0 && n >= -323) return 1.0 / go_pow10(-n); /* unused by decimals */
  return n > 0 ? INFINITY : 0.0;
}

/* decimal.go:39-50 */
double or_i128_to_f64(uint64_t lo, int64_t hi, int scale) {
  if ((hi == 0 && lo <= (uint64_t)INT64_MAX) || (hi == -1 && lo > (uint64_t)INT64_MAX))
    return (double)(int64_t)lo / go_pow10(scale);
  double f = (double)hi * 18446744073709551616.0 + (double)lo;
  return f / go_pow10(scale);
}

/* decimal.go:26-36 */
void or_i128_from_f64(double f, int scale, uint64_t* lo, int64_t* hi) {
  double scaled = f * go_pow10(scale);
  if (scaled >= 0) scaled += 0.5;
  else scaled -= 0.5;
  int64_t v = (int64_t)scaled;
  *lo = (uint64_t)v;
  *hi = v < 0 ? -1 : 0;
}

/* decimal.go:107-140 (FormatDecimal) */
int or_i128_format(uint64_t lo, int64_t hi, int scale, char* buf, int cap) {
  if (scale <= 0) return snprintf(buf, cap, "%lld", (long long)(int64_t)lo);
  int neg = hi < 0;
  uint64_t alo = lo;
  int64_t ahi = hi;
  if (neg) or_i128_neg(lo, hi, &alo, &ahi);
  int64_t v = (int64_t)alo;
  int64_t divisor = (int64_t)go_pow10(scale);
  int64_t ip = v / divisor, fp = v % divisor;
  char frac[64];
  snprintf(frac, sizeof frac, "%0*lld", scale, (long long)fp);
  int L = (int)strlen(frac);
  while (L > 0 && frac[L - 1] == '0') frac[--L] = 0;
  if (L == 0) strcpy(frac, "0");
  return snprintf(buf, cap, "%s%lld.%s", neg ? "-" : "", (long long)ip, frac);
}

/* decimal.go:155-187 (ParseDecimalString) */
int64_t or_parse_decimal(const char* s, int scale) {
  while (*s == ' ' || *s == '\t' || *s == '\n') ++s;
  char tmp[128];
  snprintf(tmp, sizeof tmp, "%s", s);
  int L = (int)strlen(tmp);
  while (L > 0 && (tmp[L - 1] == ' ' || tmp[L - 1] == '\t' || tmp[L - 1] == '\n')) tmp[--L] = 0;
  char* p = tmp;
  int neg = 0;
  if (*p == '-') { neg = 1; ++p; }
  else if (*p == '+') ++p;
  char ip[128] = {0}, fp[128] = {0};
  char* dot = strchr(p, '.');
  if (dot) {
    memcpy(ip, p, dot - p);
    snprintf(fp, sizeof fp, "%s", dot + 1);
  } else {
    snprintf(ip, sizeof ip, "%s", p);
  }
  int fl = (int)strlen(fp);
  if (fl < scale) {
    for (int i = fl; i < scale; ++i) fp[i] = '0';
    fp[scale] = 0;
  } else if (fl > scale) {
    fp[scale] = 0;
  }
  char comb[256];
  snprintf(comb, sizeof comb, "%s%s", ip, fp);
  long long v = 0;
  sscanf(comb, "%lld", &v);
  uint64_t lo = (uint64_t)v;
  int64_t hi = v < 0 ? -1 : 0;
  if (neg) or_i128_neg(lo, hi, &lo, &hi);
  return (int64_t)lo;
}

/* ------------------------------------------------------------- columns */

static int valid_at(const or_col* c, uint64_t row) {
  return c->validity == NULL || ((c->validity[row >> 6] >> (row & 63)) & 1ull);
}

/* intKeyFromVector / the consume loops' widening (join.go:504-515, aggregate.go:1910-1913) */
static int64_t int_at(const or_col* c, uint64_t row) {
  switch (c->type) {
    case OR_INT32: return ((const int32_t*)c->data)[row];
    case OR_UINT32: return ((const uint32_t*)c->data)[row];
    case OR_UINT16: return ((const uint16_t*)c->data)[row];
    case OR_DECIMAL128: return (int64_t)((const uint64_t*)c->data)[2 * row + 1];
    default: return ((const int64_t*)c->data)[row];
  }
}
static double f64_at(const or_col* c, uint64_t row) { return ((const double*)c->data)[row]; }
static void dec_at(const or_col* c, uint64_t row, uint64_t* lo, int64_t* hi) {
  if (c->type == OR_DECIMAL128) {
    const uint64_t* d = (const uint64_t*)c->data;
    *hi = (int64_t)d[2 * row];
    *lo = d[2 * row + 1];
  } else {
    int64_t v = ((const int64_t*)c->data)[row];
    *lo = (uint64_t)v;
    *hi = v < 0 ? -1 : 0; /* Int128From, decimal.go:17-23 */
  }
}

/* kernel/compare.go:26-40 and ResolveFilterKernel :145-178 */
static int cmp_i64(int64_t a, int64_t b, int op) {
  switch (op) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}
static int cmp_f64(double a, double b, int op) {
  switch (op) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}
static int pred_pass(const or_col* c, uint64_t row, int op, int64_t ci, double cf) {
  if (!valid_at(c, row)) return 0;
  switch (c->type) {
    case OR_FLOAT64: return cmp_f64(f64_at(c, row), cf, op);
    case OR_INT32: case OR_UINT16: return cmp_i64((int32_t)int_at(c, row), (int32_t)ci, op);
    default: return cmp_i64(int_at(c, row), ci, op);
  }
}

uint64_t or_filter(const or_col* col, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel, int32_t op, int64_t ci,
                   double cf, uint32_t* out) {
  uint64_t n = sel ? n_sel : n_rows, k = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t row = sel ? sel[i] : i;
    if (pred_pass(col, row, op, ci, cf)) out[k++] = (uint32_t)row;
  }
  return k;
}

/* -------------------------------------------------------------- mixing */

uint64_t or_mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

void or_digest_rows(const int64_t* rows, uint64_t n, int32_t width, uint64_t d[2]) {
  uint64_t a = 0, b = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t x = 0x243F6A8885A308D3ull, y = 0x13198A2E03707344ull;
    for (int w = 0; w < width; ++w) {
      uint64_t v = (uint64_t)rows[i * width + w];
      x = or_mix64(x ^ v);
      y = or_mix64(y + v * 0x9E3779B97F4A7C15ull);
    }
    a += x;
    b += y;
  }
  d[0] = a;
  d[1] = b;
}

/* ------------------------------------------------------- group table */

typedef struct gtab {
  int32_t n_keys;
  int64_t cap, n;
  int64_t* slot_gid; /* -1 empty */
  int64_t* keys;     /* n * n_keys */
  uint8_t* nulls;    /* n * n_keys */
  int64_t max_groups;
} gtab;

static uint64_t tuple_hash(const int64_t* k, const uint8_t* nl, int n) {
  uint64_t h = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < n; ++i) h = or_mix64(h ^ (uint64_t)k[i] ^ ((uint64_t)nl[i] << 63 ^ (uint64_t)nl[i]));
  return h;
}

static int gtab_init(gtab* t, int n_keys, int64_t max_groups) {
  t->n_keys = n_keys;
  t->cap = 1024;
  t->n = 0;
  t->max_groups = max_groups;
  t->slot_gid = (int64_t*)malloc(t->cap * 8);
  t->keys = (int64_t*)malloc((size_t)(max_groups > 0 ? max_groups : 1) * n_keys * 8);
  t->nulls = (uint8_t*)malloc((size_t)(max_groups > 0 ? max_groups : 1) * n_keys);
  if (!t->slot_gid || !t->keys || !t->nulls) return -1;
  memset(t->slot_gid, 0xFF, t->cap * 8);
  return 0;
}
static void gtab_free(gtab* t) {
  free(t->slot_gid);
  free(t->keys);
  free(t->nulls);
}
static int gtab_grow(gtab* t) {
  int64_t nc = t->cap * 2;
  int64_t* ns = (int64_t*)malloc(nc * 8);
  if (!ns) return -1;
  memset(ns, 0xFF, nc * 8);
  for (int64_t g = 0; g < t->n; ++g) {
    uint64_t h = tuple_hash(t->keys + g * t->n_keys, t->nulls + g * t->n_keys, t->n_keys);
    int64_t i = (int64_t)(h & (uint64_t)(nc - 1));
    while (ns[i] >= 0) i = (i + 1) & (nc - 1);
    ns[i] = g;
  }
  free(t->slot_gid);
  t->slot_gid = ns;
  t->cap = nc;
  return 0;
}
/* group id of the tuple, inserting in first-seen order; -1 past max_groups */
static int64_t gtab_find(gtab* t, const int64_t* k, const uint8_t* nl, int* is_new) {
  uint64_t h = tuple_hash(k, nl, t->n_keys);
  int64_t i = (int64_t)(h & (uint64_t)(t->cap - 1));
  for (;;) {
    int64_t g = t->slot_gid[i];
    if (g < 0) break;
    if (memcmp(t->keys + g * t->n_keys, k, 8 * t->n_keys) == 0 && memcmp(t->nulls + g * t->n_keys, nl, t->n_keys) == 0) {
      *is_new = 0;
      return g;
    }
    i = (i + 1) & (t->cap - 1);
  }
  if (t->n >= t->max_groups) return -1;
  int64_t g = t->n++;
  memcpy(t->keys + g * t->n_keys, k, 8 * t->n_keys);
  memcpy(t->nulls + g * t->n_keys, nl, t->n_keys);
  t->slot_gid[i] = g;
  *is_new = 1;
  if (t->n * 2 > t->cap && gtab_grow(t) != 0) return -1;
  return g;
}

/* --------------------------------------------------- accumulator update */

/* Accumulator fields chosen by initFlatAccums (aggregate.go:6030-6119) */
static void acc_init(or_acc* a, int func, const or_col* c) {
  memset(a, 0, sizeof *a);
  if (!c) return;
  if (func == OR_SUM || func == OR_AVG) {
    if (c->type == OR_FLOAT64) a->is_float = 1;
    else if (c->type == OR_DECIMAL || c->type == OR_DECIMAL128) a->is_decimal = 1;
    else if (c->type == OR_INT64 && func == OR_AVG) a->is_float = 1;
  } else if (func == OR_MIN || func == OR_MAX) {
    if (c->type == OR_FLOAT64) a->is_float = 1;
    else if (c->type == OR_DECIMAL || c->type == OR_DECIMAL128) a->is_decimal = 1;
  }
}

/* one row into one accumulator: scatterFlatAggUpdate (agg_scatter.go:519-578) */
static void acc_update(or_acc* a, int func, const or_col* c, uint64_t row) {
  if (!c) { /* COUNT(*): scatterCountStar */
    a->count++;
    return;
  }
  if (!valid_at(c, row)) return;
  int t = c->type;
  switch (func) {
    case OR_COUNT:
      a->count++;
      break;
    case OR_SUM:
    case OR_AVG:
      if (t == OR_FLOAT64) {
        a->sum_f64 += f64_at(c, row);
        a->count++;
      } else if (t == OR_DECIMAL || t == OR_DECIMAL128) {
        uint64_t lo;
        int64_t hi;
        dec_at(c, row, &lo, &hi);
        or_i128_add(a->sum_lo, a->sum_hi, lo, hi, &a->sum_lo, &a->sum_hi);
        a->count++;
      } else if (t == OR_INT64 && func == OR_AVG) {
        a->sum_f64 += (double)int_at(c, row);
        a->count++;
      } else if (t == OR_INT64 || t == OR_INT32 || t == OR_UINT16) {
        a->sum_i64 = (int64_t)((uint64_t)a->sum_i64 + (uint64_t)int_at(c, row));
        a->count++;
      }
      /* IPv4 class: no SUM/AVG case -> count stays 0 -> NULL */
      break;
    case OR_MIN:
      if (t == OR_FLOAT64) {
        double v = f64_at(c, row);
        if (!a->has_min || v < a->min_f64) { a->min_f64 = v; a->has_min = 1; }
      } else if (t != OR_DECIMAL && t != OR_DECIMAL128) {
        int64_t v = int_at(c, row);
        if (!a->has_min || v < a->min_i64) { a->min_i64 = v; a->has_min = 1; }
      }
      break;
    case OR_MAX:
      if (t == OR_FLOAT64) {
        double v = f64_at(c, row);
        if (!a->has_max || v > a->max_f64) { a->max_f64 = v; a->has_max = 1; }
      } else if (t != OR_DECIMAL && t != OR_DECIMAL128) {
        int64_t v = int_at(c, row);
        if (!a->has_max || v > a->max_i64) { a->max_i64 = v; a->has_max = 1; }
      }
      break;
  }
}

/* Accumulator.Merge (kernel/types.go:106-175), numeric fields */
static void acc_merge(or_acc* a, const or_acc* o) {
  a->sum_i64 = (int64_t)((uint64_t)a->sum_i64 + (uint64_t)o->sum_i64);
  a->sum_f64 += o->sum_f64;
  a->count += o->count;
  if (o->is_float) a->is_float = 1;
  if (o->is_decimal) {
    a->is_decimal = 1;
    or_i128_add(a->sum_lo, a->sum_hi, o->sum_lo, o->sum_hi, &a->sum_lo, &a->sum_hi);
  }
  if (o->has_min) {
    if (!a->has_min) {
      a->min_i64 = o->min_i64;
      a->min_f64 = o->min_f64;
      a->has_min = 1;
    } else if (o->is_float || a->is_float) {
      if (o->min_f64 < a->min_f64) a->min_f64 = o->min_f64;
    } else if (o->min_i64 < a->min_i64) {
      a->min_i64 = o->min_i64;
    }
  }
  if (o->has_max) {
    if (!a->has_max) {
      a->max_i64 = o->max_i64;
      a->max_f64 = o->max_f64;
      a->has_max = 1;
    } else if (o->is_float || a->is_float) {
      if (o->max_f64 > a->max_f64) a->max_f64 = o->max_f64;
    } else if (o->max_i64 > a->max_i64) {
      a->max_i64 = o->max_i64;
    }
  }
}

int or_finalize(const or_acc* a, int32_t func, int32_t in_type, int32_t scale, double* f, int64_t* i) {
  *f = 0.0;
  *i = 0;
  switch (func) {
    case OR_SUM:
      if (a->count == 0) return 0;
      if (a->is_decimal) *f = or_i128_to_f64(a->sum_lo, a->sum_hi, scale);
      else if (a->is_float) *f = a->sum_f64;
      else *f = (double)a->sum_i64;
      return 1;
    case OR_AVG:
      if (a->count == 0) return 0;
      if (a->is_decimal) *f = or_i128_to_f64(a->sum_lo, a->sum_hi, scale) / (double)a->count;
      else if (a->is_float) *f = a->sum_f64 / (double)a->count;
      else *f = (double)a->sum_i64 / (double)a->count;
      return 1;
    case OR_COUNT:
      *i = a->count;
      return 1;
    case OR_MIN:
      if (!a->has_min) return 0;
      if (in_type == OR_FLOAT64) *f = a->min_f64;
      else *i = a->min_i64;
      return 1;
    case OR_MAX:
      if (!a->has_max) return 0;
      if (in_type == OR_FLOAT64) *f = a->max_f64;
      else *i = a->max_i64;
      return 1;
  }
  return 0;
}

/* ------------------------------------------------------- the aggregate */

static void key_of(const or_col* cols, const int32_t* key_cols, int n_keys, uint64_t row, int64_t* k, uint8_t* nl) {
  for (int i = 0; i < n_keys; ++i) {
    const or_col* c = &cols[key_cols[i]];
    if (!valid_at(c, row)) {
      nl[i] = 1;
      k[i] = 0;
    } else {
      nl[i] = 0;
      k[i] = int_at(c, row);
    }
  }
}

int64_t or_hash_agg(const or_col* cols, int32_t n_cols, uint64_t n_rows, const uint32_t* sel, uint64_t n_sel,
                    int32_t pred_col, int32_t pred_op, int64_t pred_i64, double pred_f64, const int32_t* key_cols,
                    int32_t n_keys, const int32_t* agg_func, const int32_t* agg_col, int32_t n_aggs,
                    int64_t* keys_out, uint8_t* key_null_out, or_acc* acc_out, int64_t max_groups) {
  (void)n_cols;
  gtab t;
  if (gtab_init(&t, n_keys, max_groups) != 0) return -1;
  int64_t k[8];
  uint8_t nl[8];
  uint64_t n = sel ? n_sel : n_rows;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t row = sel ? sel[i] : i;
    if (pred_col >= 0 && !pred_pass(&cols[pred_col], row, pred_op, pred_i64, pred_f64)) continue;
    key_of(cols, key_cols, n_keys, row, k, nl);
    int is_new = 0;
    int64_t g = gtab_find(&t, k, nl, &is_new);
    if (g < 0) {
      gtab_free(&t);
      return -1;
    }
    or_acc* acc = acc_out + g * n_aggs;
    if (is_new)
      for (int a = 0; a < n_aggs; ++a) acc_init(&acc[a], agg_func[a], agg_col[a] >= 0 ? &cols[agg_col[a]] : NULL);
    for (int a = 0; a < n_aggs; ++a) acc_update(&acc[a], agg_func[a], agg_col[a] >= 0 ? &cols[agg_col[a]] : NULL, row);
  }
  memcpy(keys_out, t.keys, (size_t)t.n * n_keys * 8);
  if (key_null_out) memcpy(key_null_out, t.nulls, (size_t)t.n * n_keys);
  int64_t ng = t.n;
  gtab_free(&t);
  return ng;
}

/* --------------------------------------------- parallel CPU baseline */

typedef struct part_job {
  const or_col* cols;
  int32_t n_cols;
  uint64_t beg, end;
  int32_t pred_col, pred_op;
  int64_t pred_i64;
  double pred_f64;
  const int32_t *key_cols, *agg_func, *agg_col;
  int32_t n_keys, n_aggs;
  int64_t max_groups;
  int64_t* keys;
  or_acc* accs;
  int64_t n_groups;
} part_job;

static void* part_run(void* arg) {
  part_job* j = (part_job*)arg;
  /* a chunk is a contiguous row range: express it as offset columns */
  or_col c[32];
  for (int i = 0; i < j->n_cols; ++i) {
    c[i] = j->cols[i];
    int w = c[i].type == OR_INT32 || c[i].type == OR_UINT32 ? 4 : c[i].type == OR_UINT16 ? 2
            : c[i].type == OR_DECIMAL128 ? 16 : 8;
    c[i].data = (const char*)c[i].data + j->beg * w;
    c[i].validity = NULL; /* baseline inputs carry no nulls */
  }
  j->n_groups = or_hash_agg(c, j->n_cols, j->end - j->beg, NULL, 0, j->pred_col, j->pred_op, j->pred_i64, j->pred_f64,
                            j->key_cols, j->n_keys, j->agg_func, j->agg_col, j->n_aggs, j->keys, NULL, j->accs,
                            j->max_groups);
  return NULL;
}

int64_t or_hash_agg_parallel(const or_col* cols, int32_t n_cols, uint64_t n_rows, int32_t pred_col, int32_t pred_op,
                             int64_t pred_i64, double pred_f64, const int32_t* key_cols, int32_t n_keys,
                             const int32_t* agg_func, const int32_t* agg_col, int32_t n_aggs, int32_t threads,
                             int64_t* keys_out, or_acc* acc_out, int64_t max_groups) {
  if (threads < 1) threads = 1;
  if (n_cols > 32) return -1;
  part_job* jobs = (part_job*)calloc(threads, sizeof(part_job));
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  for (int t = 0; t < threads; ++t) {
    part_job* j = &jobs[t];
    j->cols = cols;
    j->n_cols = n_cols;
    j->beg = n_rows * t / threads;
    j->end = n_rows * (t + 1) / threads;
    j->pred_col = pred_col;
    j->pred_op = pred_op;
    j->pred_i64 = pred_i64;
    j->pred_f64 = pred_f64;
    j->key_cols = key_cols;
    j->agg_func = agg_func;
    j->agg_col = agg_col;
    j->n_keys = n_keys;
    j->n_aggs = n_aggs;
    j->max_groups = max_groups;
    j->keys = (int64_t*)malloc((size_t)max_groups * n_keys * 8);
    j->accs = (or_acc*)malloc((size_t)max_groups * n_aggs * sizeof(or_acc));
    pthread_create(&th[t], NULL, part_run, j);
  }
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  /* MergeSink: fold partials in worker order */
  gtab g;
  int64_t ng = -1;
  if (gtab_init(&g, n_keys, max_groups) == 0) {
    ng = 0;
    uint8_t nl[8] = {0};
    for (int t = 0; t < threads && ng >= 0; ++t) {
      part_job* j = &jobs[t];
      if (j->n_groups < 0) { ng = -1; break; }
      for (int64_t q = 0; q < j->n_groups; ++q) {
        int is_new = 0;
        int64_t gi = gtab_find(&g, j->keys + q * n_keys, nl, &is_new);
        if (gi < 0) { ng = -1; break; }
        for (int a = 0; a < n_aggs; ++a) {
          if (is_new) acc_out[gi * n_aggs + a] = j->accs[q * n_aggs + a];
          else acc_merge(&acc_out[gi * n_aggs + a], &j->accs[q * n_aggs + a]);
        }
      }
    }
    if (ng >= 0) {
      ng = g.n;
      memcpy(keys_out, g.keys, (size_t)ng * n_keys * 8);
    }
    gtab_free(&g);
  }
  for (int t = 0; t < threads; ++t) {
    free(jobs[t].keys);
    free(jobs[t].accs);
  }
  free(jobs);
  free(th);
  return ng;
}

/* ---------------------------------------------------------- hash join */

int64_t or_hash_join(const or_col* build_key, uint64_t n_build, const or_col* probe_key, uint64_t n_probe,
                     const uint32_t* probe_sel, uint64_t n_probe_sel, uint32_t* out_probe, uint32_t* out_build,
                     int64_t cap) {
  /* HashJoin.Build: key -> head, arenaNext chains, newest first (join.go:551-560) */
  uint64_t tcap = 16;
  while (tcap < n_build * 2 + 2) tcap <<= 1;
  int64_t* tkey = (int64_t*)malloc(tcap * 8);
  int64_t* thead = (int64_t*)malloc(tcap * 8);
  int64_t* next = (int64_t*)malloc((n_build ? n_build : 1) * 8);
  uint8_t* used = (uint8_t*)calloc(tcap, 1);
  for (uint64_t r = 0; r < n_build; ++r) {
    if (!valid_at(build_key, r)) continue; /* NULL keys never join */
    int64_t k = int_at(build_key, r);
    uint64_t i = or_mix64((uint64_t)k) & (tcap - 1);
    while (used[i] && tkey[i] != k) i = (i + 1) & (tcap - 1);
    if (!used[i]) {
      used[i] = 1;
      tkey[i] = k;
      next[r] = -1;
    } else {
      next[r] = thead[i];
    }
    thead[i] = (int64_t)r;
  }
  /* inlineIntProbe: probe order, then chain order */
  int64_t out = 0;
  uint64_t n = probe_sel ? n_probe_sel : n_probe;
  for (uint64_t p = 0; p < n; ++p) {
    uint64_t row = probe_sel ? probe_sel[p] : p;
    if (!valid_at(probe_key, row)) continue;
    int64_t k = int_at(probe_key, row);
    uint64_t i = or_mix64((uint64_t)k) & (tcap - 1);
    while (used[i] && tkey[i] != k) i = (i + 1) & (tcap - 1);
    if (!used[i]) continue;
    for (int64_t b = thead[i]; b >= 0; b = next[b]) {
      if (out >= cap) { out = -1; goto done; }
      out_probe[out] = (uint32_t)row;
      out_build[out] = (uint32_t)b;
      ++out;
    }
  }
done:
  free(tkey);
  free(thead);
  free(next);
  free(used);
  return out;
}

/* ------------------------------------------------------- semi / anti join */

/* executeSemiAntiJoin without a residual filter (join.go:3245-3310) and its
 * typed loops semiProbeInt64 / antiProbeInt64 (join.go:4486-4650): the probe
 * positions, in input order, whose key is (semi) / is not (anti) among the
 * non-NULL build keys; a NULL probe key never matches (the anti join keeps
 * it, join.go:3290-3296). Returns the number of positions written. */
int64_t or_semi_join(const or_col* build_key, uint64_t n_build, const or_col* probe_key, uint64_t n_probe,
                     const uint32_t* probe_sel, uint64_t n_probe_sel, int anti, uint32_t* out) {
  uint64_t tcap = 16;
  while (tcap < n_build * 2 + 2) tcap <<= 1;
  int64_t* tkey = (int64_t*)malloc(tcap * 8);
  uint8_t* used = (uint8_t*)calloc(tcap, 1);
  for (uint64_t r = 0; r < n_build; ++r) {
    if (!valid_at(build_key, r)) continue;
    int64_t k = int_at(build_key, r);
    uint64_t i = or_mix64((uint64_t)k) & (tcap - 1);
    while (used[i] && tkey[i] != k) i = (i + 1) & (tcap - 1);
    used[i] = 1;
    tkey[i] = k;
  }
  int64_t nout = 0;
  uint64_t n = probe_sel ? n_probe_sel : n_probe;
  for (uint64_t p = 0; p < n; ++p) {
    uint64_t row = probe_sel ? probe_sel[p] : p;
    int found = 0;
    if (valid_at(probe_key, row)) {
      int64_t k = int_at(probe_key, row);
      uint64_t i = or_mix64((uint64_t)k) & (tcap - 1);
      while (used[i] && tkey[i] != k) i = (i + 1) & (tcap - 1);
      found = used[i];
    }
    if (anti ? !found : found) out[nout++] = (uint32_t)row;
  }
  free(tkey);
  free(used);
  return nout;
}

/* ------------------------------------------------ join -> aggregate */

/* Inner join build_key = probe_key feeding GROUP BY (an int build column)
 * SUM(probe int column x build DECIMAL column), COUNT(*): or_hash_join's
 * build and probe (HashJoin.Build, join.go:551-560; inlineIntProbe,
 * :2670-2836) with every matched pair fed to scatterSumDecimal
 * (agg_scatter.go:330-348) and scatterCountStar (:387-406). The product
 * (aggPreProject, plan.go:7900-7990, at the decimal's scale) is formed exactly
 * in 128 bits and added with Int128.Add (decimal.go:77-85). The probe rows
 * split into `threads` contiguous ranges, each aggregated into its own
 * partial groups, folded in range order with Accumulator.Merge's Int128 add
 * (kernel/types.go:106-175). Groups come out ascending by key. Returns the
 * group count, or -1 past max_groups / out of memory. */
typedef struct ja_job {
  const int64_t* tkey;
  const int64_t* thead;
  const int64_t* next;
  const uint8_t* used;
  uint64_t tcap;
  const or_col *bgroup, *bdec, *pkey, *pmul;
  uint64_t beg, end;
  int64_t max_groups;
  int64_t* keys; /* partial groups, unordered */
  int64_t* cnt;
  uint64_t* lo;
  int64_t* hi;
  int64_t n;     /* -1: over max_groups */
} ja_job;

static int64_t ja_slot(ja_job* j, int64_t* gslot, uint64_t gcap, int64_t g) {
  uint64_t i = or_mix64((uint64_t)g) & (gcap - 1);
  while (gslot[i] >= 0 && j->keys[gslot[i]] != g) i = (i + 1) & (gcap - 1);
  if (gslot[i] < 0) {
    if (j->n >= j->max_groups) return -1;
    gslot[i] = j->n;
    j->keys[j->n] = g;
    j->cnt[j->n] = 0;
    j->lo[j->n] = 0;
    j->hi[j->n] = 0;
    ++j->n;
  }
  return gslot[i];
}

static void* ja_run(void* arg) {
  ja_job* j = (ja_job*)arg;
  uint64_t gcap = 16;
  while (gcap < (uint64_t)j->max_groups * 2) gcap <<= 1;
  int64_t* gslot = (int64_t*)malloc(gcap * 8);
  if (!gslot) { j->n = -1; return NULL; }
  memset(gslot, 0xFF, gcap * 8);
  j->n = 0;
  for (uint64_t row = j->beg; row < j->end; ++row) {
    if (!valid_at(j->pkey, row)) continue;
    int64_t k = int_at(j->pkey, row);
    uint64_t i = or_mix64((uint64_t)k) & (j->tcap - 1);
    while (j->used[i] && j->tkey[i] != k) i = (i + 1) & (j->tcap - 1);
    if (!j->used[i]) continue;
    for (int64_t b = j->thead[i]; b >= 0; b = j->next[b]) {
      int64_t gi = ja_slot(j, gslot, gcap, int_at(j->bgroup, (uint64_t)b));
      if (gi < 0) { j->n = -1; free(gslot); return NULL; }
      j->cnt[gi] += 1;
      if (!valid_at(j->pmul, row) || !valid_at(j->bdec, (uint64_t)b)) continue;
      __int128 p = (__int128)int_at(j->pmul, row) * (__int128)int_at(j->bdec, (uint64_t)b);
      or_i128_add(j->lo[gi], j->hi[gi], (uint64_t)p, (int64_t)(p >> 64), &j->lo[gi], &j->hi[gi]);
    }
  }
  free(gslot);
  return NULL;
}

static int cmp_i64_idx(const void* a, const void* b, void* keys) {
  int64_t x = ((const int64_t*)keys)[*(const int64_t*)a], y = ((const int64_t*)keys)[*(const int64_t*)b];
  return x < y ? -1 : x > y;
}

int64_t or_join_agg_parallel(const or_col* build_key, const or_col* build_group, const or_col* build_dec,
                             uint64_t n_build, const or_col* probe_key, const or_col* probe_mul, uint64_t n_probe,
                             int32_t threads, int64_t* keys_out, int64_t* count_out, uint64_t* sum_lo_out,
                             int64_t* sum_hi_out, int64_t max_groups) {
  if (threads < 1) threads = 1;
  uint64_t tcap = 16;
  while (tcap < n_build * 2 + 2) tcap <<= 1;
  int64_t* tkey = (int64_t*)malloc(tcap * 8);
  int64_t* thead = (int64_t*)malloc(tcap * 8);
  int64_t* next = (int64_t*)malloc((n_build ? n_build : 1) * 8);
  uint8_t* used = (uint8_t*)calloc(tcap, 1);
  ja_job* jobs = (ja_job*)calloc(threads, sizeof(ja_job));
  pthread_t* th = (pthread_t*)calloc(threads, sizeof(pthread_t));
  int64_t ng = -1;
  if (!tkey || !thead || !next || !used || !jobs || !th) goto out;
  for (uint64_t r = 0; r < n_build; ++r) {
    if (!valid_at(build_key, r)) continue;
    int64_t k = int_at(build_key, r);
    uint64_t i = or_mix64((uint64_t)k) & (tcap - 1);
    while (used[i] && tkey[i] != k) i = (i + 1) & (tcap - 1);
    if (!used[i]) {
      used[i] = 1;
      tkey[i] = k;
      next[r] = -1;
    } else {
      next[r] = thead[i];
    }
    thead[i] = (int64_t)r;
  }
  for (int t = 0; t < threads; ++t) {
    ja_job* j = &jobs[t];
    *j = (ja_job){tkey, thead, next, used, tcap, build_group, build_dec, probe_key, probe_mul,
                  n_probe * t / threads, n_probe * (t + 1) / threads, max_groups, NULL, NULL, NULL, NULL, 0};
    j->keys = (int64_t*)malloc((size_t)max_groups * 8);
    j->cnt = (int64_t*)malloc((size_t)max_groups * 8);
    j->lo = (uint64_t*)malloc((size_t)max_groups * 8);
    j->hi = (int64_t*)malloc((size_t)max_groups * 8);
    pthread_create(&th[t], NULL, ja_run, j);
  }
  for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  {
    /* fold the partials in range order */
    ja_job all = {0};
    all.max_groups = max_groups;
    all.keys = keys_out;
    all.cnt = count_out;
    all.lo = sum_lo_out;
    all.hi = sum_hi_out;
    uint64_t gcap = 16;
    while (gcap < (uint64_t)max_groups * 2) gcap <<= 1;
    int64_t* gslot = (int64_t*)malloc(gcap * 8);
    if (!gslot) goto out;
    memset(gslot, 0xFF, gcap * 8);
    int ok = 1;
    for (int t = 0; t < threads && ok; ++t) {
      ja_job* j = &jobs[t];
      if (j->n < 0) { ok = 0; break; }
      for (int64_t q = 0; q < j->n; ++q) {
        int64_t gi = ja_slot(&all, gslot, gcap, j->keys[q]);
        if (gi < 0) { ok = 0; break; }
        count_out[gi] += j->cnt[q];
        or_i128_add(sum_lo_out[gi], sum_hi_out[gi], j->lo[q], j->hi[q], &sum_lo_out[gi], &sum_hi_out[gi]);
      }
    }
    free(gslot);
    if (ok) {
      /* ascending by key */
      int64_t n = all.n;
      int64_t* idx = (int64_t*)malloc((size_t)(n ? n : 1) * 8 * 5);
      if (!idx) goto out;
      int64_t *k2 = idx + n, *c2 = idx + 2 * n, *h2 = idx + 3 * n;
      uint64_t* l2 = (uint64_t*)(idx + 4 * n);
      for (int64_t i = 0; i < n; ++i) idx[i] = i;
      qsort_r(idx, (size_t)n, 8, cmp_i64_idx, keys_out);
      for (int64_t i = 0; i < n; ++i) {
        k2[i] = keys_out[idx[i]];
        c2[i] = count_out[idx[i]];
        l2[i] = sum_lo_out[idx[i]];
        h2[i] = sum_hi_out[idx[i]];
      }
      memcpy(keys_out, k2, (size_t)n * 8);
      memcpy(count_out, c2, (size_t)n * 8);
      memcpy(sum_lo_out, l2, (size_t)n * 8);
      memcpy(sum_hi_out, h2, (size_t)n * 8);
      free(idx);
      ng = n;
    }
  }
out:
  if (jobs)
    for (int t = 0; t < threads; ++t) {
      free(jobs[t].keys);
      free(jobs[t].cnt);
      free(jobs[t].lo);
      free(jobs[t].hi);
    }
  free(jobs);
  free(th);
  free(tkey);
  free(thead);
  free(next);
  free(used);
  return ng;
}
