/*
 * wadjet_gen.h — seeded synthetic column generators that write straight into
 * device memory (test and benchmark workloads; not a reference interface).
 *
 * Every value is a pure function of (seed, row): a SplitMix64 stream indexed by
 * row, so any row range can be regenerated anywhere and host-side checks can
 * rebuild the same inputs from the same seed.
 */
#ifndef WADJET_GEN_H
#define WADJET_GEN_H

#include <stdint.h>

#include "wadjet_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum wg_gen_kind {
  WG_GEN_UNIFORM_I64 = 1, /* int64 uniform in [i0, i1] */
  WG_GEN_UNIFORM_I32 = 2, /* int32 uniform in [i0, i1] */
  WG_GEN_UNIFORM_U16 = 3, /* uint16 uniform in [i0, i1] */
  WG_GEN_ZIPF_I32 = 4,    /* int32 rank in [1, i0], P(r) ~ r^-f0 */
  WG_GEN_ZIPF_U32 = 5,    /* uint32 rank in [1, i0], P(r) ~ r^-f0, scrambled by a bijection of [0,2^32) */
  WG_GEN_NORMAL_F64 = 6,  /* float64 N(f0, f1), exact zeros rejected */
  WG_GEN_LOGNORMAL_I64 = 7, /* int64 round(exp(N(f0, f1))) */
  WG_GEN_PERM_KEY_I64 = 8,  /* int64 mix64(pi(row)) for a seeded permutation pi of [0, i0): unique keys */
  WG_GEN_FK_I64 = 9,        /* int64 foreign key: with prob i1/1000 mix64(u % i0) (a build key), else mix64(i0 + u % i0) */
  WG_GEN_SEQ_I64 = 10,      /* int64 i0 + row * i1 (monotonic timestamps) */
  WG_GEN_RANK_U32 = 11      /* uint32: the address WG_GEN_ZIPF_U32 (same seed) gives rank row + 1 */
} wg_gen_kind;

/* Fill `out` (device) with rows [row0, row0 + n) of the generator. */
wg_status wg_gen(wg_ctx* ctx, int32_t kind, void* out, uint64_t row0, uint64_t n, uint64_t seed,
                 int64_t i0, int64_t i1, double f0, double f1);

#ifdef __cplusplus
}
#endif
#endif
