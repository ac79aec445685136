/*
 * oracle.c — CPU oracle for the in-place 213 transpose.
 *
 * TEST INFRASTRUCTURE ONLY: linked by nothing in the product library.  Used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker / CPU baseline.  Declarations, citations and the list of checks that
 * pin each function are in oracle.h.
 *
 * All index arithmetic is exact 64-bit integer arithmetic.  Payloads are
 * copied as raw bytes (memcpy of esz bytes), so every bit pattern, NaN
 * payloads included, is preserved.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- generators */

static inline uint64_t splitmix64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void ipt_oracle_fill_iota(void* data, int esz, uint64_t n) {
    if (esz == 4) {
        uint32_t* d = (uint32_t*)data;
        for (uint64_t i = 0; i < n; i++) d[i] = (uint32_t)(int32_t)i;
    } else {
        uint64_t* d = (uint64_t*)data;
        for (uint64_t i = 0; i < n; i++) d[i] = i;
    }
}

/* Word t of the stream is splitmix64_mix(seed + (t+1)*golden).  8-byte
 * elements take word i; 4-byte element i takes half (i&1) of word i/2. */
void ipt_oracle_fill_random(void* data, int esz, uint64_t n, uint64_t seed) {
    const uint64_t golden = 0x9E3779B97F4A7C15ull;
    if (esz == 8) {
        uint64_t* d = (uint64_t*)data;
        for (uint64_t i = 0; i < n; i++) d[i] = splitmix64_mix(seed + (i + 1) * golden);
    } else {
        uint32_t* d = (uint32_t*)data;
        for (uint64_t i = 0; i < n; i++) {
            uint64_t w = splitmix64_mix(seed + (i / 2 + 1) * golden);
            d[i] = (uint32_t)(w >> (32 * (i & 1)));
        }
    }
}

/* ------------------------------------------------------------ closed form */

void ipt_oracle_seq_tt_213(const void* in, void* out, int esz,
                           int64_t d1, int64_t d2, int64_t d3) {
    const unsigned char* src = (const unsigned char*)in;
    unsigned char* dst = (unsigned char*)out;
    const int64_t d1d2 = d1 * d2;
    for (int64_t x2 = 0; x2 < d3; x2++)
        for (int64_t x1 = 0; x1 < d2; x1++)
            for (int64_t x0 = 0; x0 < d1; x0++) {
                int64_t s = x0 + x1 * d1 + x2 * d1d2;
                int64_t d = x1 + x0 * d2 + x2 * d1d2;
                memcpy(dst + d * esz, src + s * esz, (size_t)esz);
            }
}

int ipt_oracle_seq_tt(const void* in, void* out, int esz, int rank,
                      const int64_t* dims, const int* perm) {
    if (rank <= 0 || rank > 16) return -1;
    int64_t stride[16], stride_d[16];
    uint64_t vol = 1;
    for (int i = 0; i < rank; i++) {
        stride[i] = (int64_t)vol;
        vol *= (uint64_t)dims[i];
    }
    /* stride_d[i]: product of the output dims that precede source dim i in
     * the output order (the inner loop of src/tensor_util.cpp:84-87). */
    for (int i = 0; i < rank; i++) {
        int64_t s = 1;
        int j = 0;
        for (; j < rank && perm[j] != i; j++) s *= dims[perm[j]];
        if (j == rank) return -1;
        stride_d[i] = s;
    }
    const unsigned char* src = (const unsigned char*)in;
    unsigned char* dst = (unsigned char*)out;
    for (uint64_t idx_s = 0; idx_s < vol; idx_s++) {
        uint64_t idx_d = 0;
        for (int i = 0; i < rank; i++)
            idx_d += ((idx_s / (uint64_t)stride[i]) % (uint64_t)dims[i]) * (uint64_t)stride_d[i];
        memcpy(dst + idx_d * esz, src + idx_s * esz, (size_t)esz);
    }
    return 0;
}

/* ----------------------------------------------------------- number theory */

void ipt_oracle_extended_gcd(int64_t a, int64_t b, int64_t* gcd, int64_t* mmi) {
    int64_t x = 0, lastx = 1, y = 1, lasty = 0, origb = b;
    while (b != 0) {
        int64_t quotient = a / b;
        int64_t newb = a % b;
        a = b;
        b = newb;
        int64_t newx = lastx - quotient * x;
        lastx = x;
        x = newx;
        int64_t newy = lasty - quotient * y;
        lasty = y;
        y = newy;
    }
    *gcd = a;
    *mmi = 0;
    if (a == 1) *mmi = (lastx < 0) ? lastx + origb : lastx;
}

static int ceil_log2_u64(uint64_t d) {
    int l = 0;
    while (((uint64_t)1 << l) < d) l++;
    return l;
}

void ipt_oracle_find_divisor32(uint32_t d, uint32_t* mul, uint32_t* shift) {
    if (d <= 1) { *mul = 0; *shift = 0; return; }
    uint32_t p = 31 + (uint32_t)ceil_log2_u64(d);
    uint64_t m = (((uint64_t)1 << p) + d - 1) / d;
    *mul = (uint32_t)m;
    *shift = p - 32;
}

uint32_t ipt_oracle_rdiv32(uint32_t x, uint32_t d, uint32_t mul, uint32_t shift) {
    (void)d;
    if (!mul) return x;
    return (uint32_t)(((uint64_t)x * mul) >> 32) >> shift;
}

void ipt_oracle_find_divisor64(uint64_t d, uint64_t* mul, uint32_t* shift) {
    if (d <= 1) { *mul = 0; *shift = 0; return; }
    uint32_t p = 63 + (uint32_t)ceil_log2_u64(d);
    unsigned __int128 m = ((((unsigned __int128)1) << p) + d - 1) / d;
    *mul = (uint64_t)m;
    *shift = p - 64;
}

uint64_t ipt_oracle_rdiv64(uint64_t x, uint64_t d, uint64_t mul, uint32_t shift) {
    (void)d;
    if (!mul) return x;
    return (uint64_t)(((unsigned __int128)x * mul) >> 64) >> shift;
}

/* ---------------------------------------------------------------- pass maps */

void ipt_oracle_make_plan(int64_t R, int64_t C, ipt_oracle_plan* p) {
    int64_t c, t;
    p->R = R;
    p->C = C;
    ipt_oracle_extended_gcd(R, C, &c, &t);
    p->c = c;
    p->a = R / c;
    p->b = C / c;
    /* C2R: extended_gcd(d2/c, d1/c, t, k) (src/transpose.cu:35-40). */
    ipt_oracle_extended_gcd(p->a, p->b, &t, &p->k);
    /* R2C: extended_gcd(d1/c, d2/c, t, q) with r2c's d1 = C, d2 = R
     * (src/transpose.cu:77-82). */
    ipt_oracle_extended_gcd(p->b, p->a, &t, &p->q);
}

static inline int64_t pmod(int64_t x, int64_t m) {
    int64_t r = x % m;
    return r < 0 ? r + m : r;
}

int64_t ipt_oracle_permuter(int which, const ipt_oracle_plan* p, int64_t i) {
    const int64_t R = p->R, C = p->C, c = p->c, a = p->a, q = p->q;
    if (which == IPT_OPERM_C2R_POST) {
        /* c2r::postpermuter: (i*n - i/a) % m with m = R, n = C. */
        return (i * C - i / a) % R;
    }
    /* r2c::prepermuter: k = ((c-1)*i) % c, l = (c-1+i)/c,
     * r = k*a + (l*q) % a. */
    int64_t kk = ((c - 1) * i) % c;
    int64_t l = (c - 1 + i) / c;
    return kk * a + (l * q) % a;
}

int64_t ipt_oracle_rotator(int which, const ipt_oracle_plan* p, int64_t j) {
    const int64_t R = p->R, b = p->b;
    switch (which) {
    case IPT_OROT_C2R_PRE: return j / b;
    case IPT_OROT_C2R_POST: return j % R;
    case IPT_OROT_R2C_PRE: { int64_t m = j % R; return m == 0 ? 0 : R - m; }
    default: {
        int64_t bd = j / b;
        if (bd == 0) return 0;
        return bd < R ? R - bd : R + R - bd;
    }
    }
}

int64_t ipt_oracle_map(int which, const ipt_oracle_plan* p, int64_t i, int64_t j) {
    const int64_t R = p->R, C = p->C, c = p->c, a = p->a, b = p->b, k = p->k;
    switch (which) {
    case IPT_OMAP_C2R_PRE:
        return (i + j / b) % R;
    case IPT_OMAP_C2R_POST:
        return (i * C - i / a + j) % R;
    case IPT_OMAP_R2C_PRE:
        return ipt_oracle_permuter(IPT_OPERM_R2C_PRE, p, pmod(i - j, R));
    case IPT_OMAP_R2C_POST:
        return (i + ipt_oracle_rotator(IPT_OROT_R2C_POST, p, j)) % R;
    case IPT_OMAP_C2R_SHUFFLE: {
        int64_t f = j + i * (C - 1);
        if (i - (j % c) > R - c) f += R;
        int64_t u = (k * ((f / c) % b)) % b;
        return u + (f % c) * b;
    }
    case IPT_OMAP_R2C_SHUFFLE:
        return ((j / b + i) % R + j * R) % C;
    default:
        return -1;
    }
}

/* ---------------------------------------------------------- in-place passes */

#define IPT_COL_BLOCK 32

static int set_threads(int threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

typedef int64_t (*col_fn)(const ipt_oracle_plan*, int64_t i, int64_t j, int which);

static int64_t col_via_map(const ipt_oracle_plan* p, int64_t i, int64_t j, int which) {
    return ipt_oracle_map(which, p, i, j);
}
static int64_t col_via_rotator(const ipt_oracle_plan* p, int64_t i, int64_t j, int which) {
    return (i + ipt_oracle_rotator(which, p, j)) % p->R;
}

/* new[i][j] = old[g(i,j)][j] for every slice, in place, one column block at a
 * time through a per-thread buffer. */
static int col_gather(unsigned char* data, int esz, const ipt_oracle_plan* p,
                      int64_t d3, col_fn g, int which, int threads) {
    const int64_t R = p->R, C = p->C;
    const int64_t nblk = (C + IPT_COL_BLOCK - 1) / IPT_COL_BLOCK;
    const int64_t work = nblk * d3;
    int fail = 0;
    int nt = set_threads(threads);
#pragma omp parallel num_threads(nt)
    {
        unsigned char* tmp = (unsigned char*)malloc((size_t)R * IPT_COL_BLOCK * esz);
        if (!tmp) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t w = 0; w < work; w++) {
            if (!tmp) continue;
            unsigned char* s = data + (size_t)(w / nblk) * R * C * esz;
            int64_t j0 = (w % nblk) * IPT_COL_BLOCK;
            int64_t nj = C - j0 < IPT_COL_BLOCK ? C - j0 : IPT_COL_BLOCK;
            for (int64_t i = 0; i < R; i++)
                for (int64_t jj = 0; jj < nj; jj++) {
                    int64_t src = g(p, i, j0 + jj, which);
                    memcpy(tmp + (i * IPT_COL_BLOCK + jj) * esz,
                           s + (src * C + j0 + jj) * esz, (size_t)esz);
                }
            for (int64_t i = 0; i < R; i++)
                memcpy(s + (i * C + j0) * esz, tmp + i * IPT_COL_BLOCK * esz, (size_t)(nj * esz));
        }
        free(tmp);
    }
    return fail ? -2 : 0;
}

/* scatter_permute: new[f(i)] = old[i] for whole rows (src/permute.cu:116-137). */
static int row_scatter(unsigned char* data, int esz, const ipt_oracle_plan* p,
                       int64_t d3, int which, int threads) {
    const int64_t R = p->R, C = p->C;
    const int64_t nblk = (C + IPT_COL_BLOCK - 1) / IPT_COL_BLOCK;
    const int64_t work = nblk * d3;
    int64_t* f = (int64_t*)malloc((size_t)R * sizeof(int64_t));
    if (!f) return -2;
    for (int64_t i = 0; i < R; i++) f[i] = ipt_oracle_permuter(which, p, i);
    int fail = 0;
    int nt = set_threads(threads);
#pragma omp parallel num_threads(nt)
    {
        unsigned char* tmp = (unsigned char*)malloc((size_t)R * IPT_COL_BLOCK * esz);
        if (!tmp) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 1)
        for (int64_t w = 0; w < work; w++) {
            if (!tmp) continue;
            unsigned char* s = data + (size_t)(w / nblk) * R * C * esz;
            int64_t j0 = (w % nblk) * IPT_COL_BLOCK;
            int64_t nj = C - j0 < IPT_COL_BLOCK ? C - j0 : IPT_COL_BLOCK;
            for (int64_t i = 0; i < R; i++)
                memcpy(tmp + f[i] * IPT_COL_BLOCK * esz, s + (i * C + j0) * esz, (size_t)(nj * esz));
            for (int64_t i = 0; i < R; i++)
                memcpy(s + (i * C + j0) * esz, tmp + i * IPT_COL_BLOCK * esz, (size_t)(nj * esz));
        }
        free(tmp);
    }
    free(f);
    return fail ? -2 : 0;
}

/* new[i][j] = old[i][s_i(j)] for every row of every slice. */
static int row_gather(unsigned char* data, int esz, const ipt_oracle_plan* p,
                      int64_t d3, int which, int threads) {
    const int64_t R = p->R, C = p->C;
    const int64_t work = R * d3;
    int fail = 0;
    int nt = set_threads(threads);
#pragma omp parallel num_threads(nt)
    {
        unsigned char* tmp = (unsigned char*)malloc((size_t)C * esz);
        if (!tmp) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 4)
        for (int64_t w = 0; w < work; w++) {
            if (!tmp) continue;
            int64_t i = w % R;
            unsigned char* row = data + (size_t)w * C * esz;
            if (esz == 4) {
                const uint32_t* r = (const uint32_t*)row;
                uint32_t* t = (uint32_t*)tmp;
                for (int64_t j = 0; j < C; j++) t[j] = r[ipt_oracle_map(which, p, i, j)];
            } else {
                const uint64_t* r = (const uint64_t*)row;
                uint64_t* t = (uint64_t*)tmp;
                for (int64_t j = 0; j < C; j++) t[j] = r[ipt_oracle_map(which, p, i, j)];
            }
            memcpy(row, tmp, (size_t)(C * esz));
        }
        free(tmp);
    }
    return fail ? -2 : 0;
}

static int valid_args(const void* data, int esz, int64_t d1, int64_t d2, int64_t d3) {
    return data && (esz == 4 || esz == 8) && d1 > 0 && d2 > 0 && d3 > 0;
}

int ipt_oracle_c2r(void* data, int esz, int64_t d1, int64_t d2, int64_t d3,
                   int fused, int threads) {
    if (!valid_args(data, esz, d1, d2, d3)) return -1;
    ipt_oracle_plan p;
    ipt_oracle_make_plan(d2, d1, &p);
    unsigned char* d = (unsigned char*)data;
    int rc = 0;
    if (p.c > 1) rc = col_gather(d, esz, &p, d3, col_via_map, IPT_OMAP_C2R_PRE, threads);
    if (!rc) rc = row_gather(d, esz, &p, d3, IPT_OMAP_C2R_SHUFFLE, threads);
    if (rc) return rc;
    if (fused) return col_gather(d, esz, &p, d3, col_via_map, IPT_OMAP_C2R_POST, threads);
    /* Unfused col_op: rotate(postrotator(R)), then
     * scatter_permute(c2r::scatter_postpermuter) = r2c::prepermuter. */
    rc = col_gather(d, esz, &p, d3, col_via_rotator, IPT_OROT_C2R_POST, threads);
    if (!rc) rc = row_scatter(d, esz, &p, d3, IPT_OPERM_R2C_PRE, threads);
    return rc;
}

int ipt_oracle_r2c(void* data, int esz, int64_t d1, int64_t d2, int64_t d3,
                   int fused, int threads) {
    if (!valid_args(data, esz, d1, d2, d3)) return -1;
    ipt_oracle_plan p;
    ipt_oracle_make_plan(d2, d1, &p);
    unsigned char* d = (unsigned char*)data;
    int rc;
    if (fused) {
        rc = col_gather(d, esz, &p, d3, col_via_map, IPT_OMAP_R2C_PRE, threads);
    } else {
        /* scatter_permute(r2c::scatter_prepermuter = c2r::postpermuter),
         * then rotate(r2c::prerotator(R)). */
        rc = row_scatter(d, esz, &p, d3, IPT_OPERM_C2R_POST, threads);
        if (!rc) rc = col_gather(d, esz, &p, d3, col_via_rotator, IPT_OROT_R2C_PRE, threads);
    }
    if (!rc) rc = row_gather(d, esz, &p, d3, IPT_OMAP_R2C_SHUFFLE, threads);
    if (!rc && p.c > 1) rc = col_gather(d, esz, &p, d3, col_via_map, IPT_OMAP_R2C_POST, threads);
    return rc;
}

int ipt_oracle_transpose_213(void* data, int esz, int64_t d1, int64_t d2,
                             int64_t d3, int threads) {
    if (!valid_args(data, esz, d1, d2, d3)) return -1;
    /* Dispatch of src/transpose.cu:105-123: skinny when either side <= 32,
     * C2R when the kernel view keeps the shorter side as rows. */
    if (d2 <= 32 || d1 <= 32) {
        if (d2 < d1) return ipt_oracle_c2r(data, esz, d1, d2, d3, 1, threads);
        return ipt_oracle_r2c(data, esz, d2, d1, d3, 1, threads);
    }
    if (d1 >= d2) return ipt_oracle_c2r(data, esz, d1, d2, d3, 1, threads);
    return ipt_oracle_r2c(data, esz, d2, d1, d3, 1, threads);
}

int ipt_oracle_openmp_transpose(int row_major, void* data, int esz, int64_t m,
                                int64_t n, void* tmp, int threads) {
    (void)tmp;
    if (row_major) return ipt_oracle_transpose_213(data, esz, n, m, 1, threads);
    return ipt_oracle_transpose_213(data, esz, m, n, 1, threads);
}

int ipt_oracle_apply_col_gather(void* data, int esz, int64_t R, int64_t C,
                                int64_t d3, int which, int threads) {
    if (!valid_args(data, esz, C, R, d3) || which < 0 || which > IPT_OMAP_R2C_POST) return -1;
    ipt_oracle_plan p;
    ipt_oracle_make_plan(R, C, &p);
    return col_gather((unsigned char*)data, esz, &p, d3, col_via_map, which, threads);
}

int ipt_oracle_apply_row_gather(void* data, int esz, int64_t R, int64_t C,
                                int64_t d3, int which, int threads) {
    if (!valid_args(data, esz, C, R, d3) ||
        (which != IPT_OMAP_C2R_SHUFFLE && which != IPT_OMAP_R2C_SHUFFLE)) return -1;
    ipt_oracle_plan p;
    ipt_oracle_make_plan(R, C, &p);
    return row_gather((unsigned char*)data, esz, &p, d3, which, threads);
}
