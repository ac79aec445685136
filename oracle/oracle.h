/*
 * oracle.h — CPU oracle for the in-place 213 transpose.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker or as the timed CPU
 * baseline, never as the thing measured on the GPU.
 *
 * Contents (each function cites the reference file:line it restates):
 *   - input generators (iota bit patterns, seeded random bits),
 *   - the closed-form out-of-place permutation seq_tt (rank 3, {1,0,2}) and the
 *     general rank-N seq_tt,
 *   - extended gcd / modular inverse and the reduced-divisor arithmetic,
 *   - every C2R / R2C / skinny pass map (gathers, rotators, scatter permuters),
 *   - an in-place CPU C2R / R2C / 213 driver built from those maps,
 *     parallelised with OpenMP: the CPU baseline path.
 *
 * Parity pinning (what checks each function; see tests/test_oracle.py):
 *   seq_tt              — the reference's own gen_ans built from source
 *                         (oracle/_ref/gen_ans) on iota inputs, and brute force.
 *   extended_gcd        — the reference's gcd.cpp built from source, plus the
 *                         identity (a*mmi) mod b == 1.
 *   find_divisor, rdiv — the reference's reduced_math.cpp built from source,
 *                         plus exact '/' and '%'.
 *   pass maps           — bijection per row/column, composition of the passes
 *                         equal to seq_tt on every shape with d1,d2 <= 40 and on
 *                         random larger shapes, fused == unfused forms, p(pi(i)) = i.
 *   in-place drivers    — seq_tt on seeded random bits.
 */
#ifndef IPT_ORACLE_H
#define IPT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generators -------------------------------------------------------- */
/* data[i] holds the bit pattern of the integer i (int32 for 4-byte elements,
 * int64 for 8-byte ones), as TensorUtil::init_data (src/tensor_util.cpp:42-47)
 * does for the int / long long types the reference test driver exercises. */
void ipt_oracle_fill_iota(void* data, int esz, uint64_t n);
/* splitmix64 stream of random bits, seeded; NaN payloads are included. */
void ipt_oracle_fill_random(void* data, int esz, uint64_t n, uint64_t seed);

/* ---- closed-form oracle ------------------------------------------------ */
/* out[x1 + x0*d2 + x2*d1*d2] = in[x0 + x1*d1 + x2*d1*d2]: TensorUtil::seq_tt
 * (src/tensor_util.cpp:80-92) with permutation {1,0,2}
 * (src/gen_ans.cpp:30, src/test_213inplace.cu:78).  Out-of-place. */
void ipt_oracle_seq_tt_213(const void* in, void* out, int esz,
                           int64_t d1, int64_t d2, int64_t d3);
/* General rank-N seq_tt (src/tensor_util.cpp:80-92): dims fastest-first,
 * perm[j] = source dim placed at output position j.  Returns 0 on success. */
int ipt_oracle_seq_tt(const void* in, void* out, int esz, int rank,
                      const int64_t* dims, const int* perm);

/* ---- number theory ----------------------------------------------------- */
/* src/gcd.cpp:3-30: Euclid with Bezout; mmi = a^-1 mod b when gcd == 1,
 * normalised to [0,b); mmi = 0 when gcd != 1 (and when b == 1). */
void ipt_oracle_extended_gcd(int64_t a, int64_t b, int64_t* gcd, int64_t* mmi);
/* src/reduced_math.cpp:36-50: mul = ceil(2^p/d), p = 31 + ceil(log2 d),
 * shift = p - 32; d == 1 gives mul = 0.  Exact for x < 2^31. */
void ipt_oracle_find_divisor32(uint32_t d, uint32_t* mul, uint32_t* shift);
uint32_t ipt_oracle_rdiv32(uint32_t x, uint32_t d, uint32_t mul, uint32_t shift);
/* src/reduced_math.cpp:53-67: 64-bit analogue, p = 63 + ceil(log2 d).
 * Exact for x < 2^63. */
void ipt_oracle_find_divisor64(uint64_t d, uint64_t* mul, uint32_t* shift);
uint64_t ipt_oracle_rdiv64(uint64_t x, uint64_t d, uint64_t mul, uint32_t shift);

/* ---- plan and pass maps ------------------------------------------------ */
/* View of one slice: R rows x C columns, row-major.  c = gcd(R,C), a = R/c,
 * b = C/c, k = a^-1 mod b (C2R, src/transpose.cu:35-40), q = b^-1 mod a
 * (R2C, src/transpose.cu:77-82). */
typedef struct {
    int64_t R, C, c, a, b, k, q;
} ipt_oracle_plan;
void ipt_oracle_make_plan(int64_t R, int64_t C, ipt_oracle_plan* p);

enum {
    /* column gathers: new[i][j] = old[g(i,j)][j] */
    IPT_OMAP_C2R_PRE = 0,      /* (i + j/b) mod R            include/equations.h:78-94, skinny.h:12-26 */
    IPT_OMAP_C2R_POST = 1,     /* (i*C - i/a + j) mod R      include/skinny.h:69-78 (with c, not d2/c) */
    IPT_OMAP_R2C_PRE = 2,      /* p((i - j) mod R)           include/skinny.h:87-107 */
    IPT_OMAP_R2C_POST = 3,     /* (i + (R - j/b) mod R) mod R include/equations.h:196-222, skinny.h:109-117 */
    /* row gathers: new[i][j] = old[i][s_i(j)] */
    IPT_OMAP_C2R_SHUFFLE = 4,  /* include/equations.h:12-54, skinny.h:28-67 (exact 64-bit) */
    IPT_OMAP_R2C_SHUFFLE = 5,  /* include/equations.h:178-194 */
    IPT_OMAP_COUNT = 6
};
/* Source index of the gather map `which` at (i, j). */
int64_t ipt_oracle_map(int which, const ipt_oracle_plan* p, int64_t i, int64_t j);

enum {
    IPT_OROT_C2R_PRE = 0,   /* j / b               include/equations.h:78-94 */
    IPT_OROT_C2R_POST = 1,  /* j mod R             include/equations.h:97-113 */
    IPT_OROT_R2C_PRE = 2,   /* (R - j mod R) mod R include/equations.h:159-176 */
    IPT_OROT_R2C_POST = 3   /* 0 | R - j/b | 2R - j/b   include/equations.h:196-222 */
};
/* Rotation amount of column j; rotate(fn): new[i][j] = old[(i + fn(j)) mod R][j]
 * (src/rotate.cu:217-235). */
int64_t ipt_oracle_rotator(int which, const ipt_oracle_plan* p, int64_t j);

enum {
    IPT_OPERM_C2R_POST = 0, /* c2r::postpermuter (i*C - i/a) mod R, include/equations.h:115-128 */
    IPT_OPERM_R2C_PRE = 1   /* r2c::prepermuter  a*((c-1)i mod c) + ((ceil-ish)*q mod a), :134-157 */
};
/* Row scatter map; scatter_permute(f): new[f(i)] = old[i] (src/permute.cu:116-137).
 * C2R scatters by IPT_OPERM_R2C_PRE, R2C scatters by IPT_OPERM_C2R_POST
 * (include/equations.h:226-236). */
int64_t ipt_oracle_permuter(int which, const ipt_oracle_plan* p, int64_t i);

/* ---- in-place drivers (the OpenMP CPU path) ---------------------------- */
/* Each call transposes d3 slices in place.  fused != 0 uses the single
 * column gather for the column op, fused == 0 uses rotate + scatter_permute,
 * as the two branches of col_op (src/transpose.cu:18-28, 57-68).
 * threads <= 0 uses the OpenMP default.  Return 0 on success, -1 on bad
 * arguments, -2 when a buffer cannot be allocated. */
/* inplace::c2r::transpose (src/transpose.cu:30-47): slices of d2 rows x d1 cols. */
int ipt_oracle_c2r(void* data, int esz, int64_t d1, int64_t d2, int64_t d3,
                   int fused, int threads);
/* inplace::r2c::transpose (src/transpose.cu:70-89): slices of d1 rows x d2 cols
 * (view R = d2, C = d1). */
int ipt_oracle_r2c(void* data, int esz, int64_t d1, int64_t d2, int64_t d3,
                   int fused, int threads);
/* inplace::transpose (src/transpose.cu:99-124): slices of d2 rows x d1 cols. */
int ipt_oracle_transpose_213(void* data, int esz, int64_t d1, int64_t d2,
                             int64_t d3, int threads);
/* inplace::openmp::transpose(row_major, data, m, n, tmp) (include/openmp.h:7-8),
 * generalised to 4- and 8-byte elements.  Row-major m x n maps to
 * (d1=n, d2=m, d3=1), column-major to (d1=m, d2=n, d3=1).  tmp is unused
 * (per-thread buffers are internal) and may be NULL. */
int ipt_oracle_openmp_transpose(int row_major, void* data, int esz, int64_t m,
                                int64_t n, void* tmp, int threads);

/* Apply one gather pass in place (test hook for per-pass GPU parity). */
int ipt_oracle_apply_col_gather(void* data, int esz, int64_t R, int64_t C,
                                int64_t d3, int which, int threads);
int ipt_oracle_apply_row_gather(void* data, int esz, int64_t R, int64_t C,
                                int64_t d3, int which, int threads);

#ifdef __cplusplus
}
#endif
#endif
