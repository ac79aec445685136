// ipt_plan.h — decomposition plan and index maps of the in-place transpose.
//
// Shared by host code (planning, ipt_debug_eval_map) and gfx950 kernels.
//
// Kernel view of one slice: R rows x C columns, row-major.  c = gcd(R,C),
// a = R/c, b = C/c.  C2R turns an R x C matrix into its C x R transpose with
// three in-place passes (src/transpose.cu:30-47); R2C is its inverse
// (src/transpose.cu:70-89).  Every pass is a gather: a column gather
// new[i][j] = old[g(i,j)][j] or a row gather new[i][j] = old[i][s_i(j)].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define IPT_HD __host__ __device__ __forceinline__

namespace ipt {

// ---------------------------------------------------------------- arithmetic

// Extended Euclid (src/gcd.cpp:3-30): gcd and mmi = a^-1 mod b when
// gcd == 1 (normalised to [0,b)); mmi = 0 otherwise and when b == 1.
inline void extended_gcd(int64_t a, int64_t b, int64_t& gcd, int64_t& mmi) {
    int64_t x = 0, lastx = 1, origb = b;
    while (b != 0) {
        int64_t q = a / b, r = a % b;
        a = b;
        b = r;
        int64_t nx = lastx - q * x;
        lastx = x;
        x = nx;
    }
    gcd = a;
    mmi = (a == 1) ? (lastx < 0 ? lastx + origb : lastx) : 0;
}

inline int ceil_log2(uint64_t d) {
    int l = 0;
    while ((uint64_t(1) << l) < d) l++;
    return l;
}

// Division by an invariant divisor with the round-up multiply-shift method
// (Granlund & Montgomery 1994; the reference's reduced_divisor,
// include/reduced_math.h:47-77, src/reduced_math.cpp:36-67):
// mul = ceil(2^p / d), p = (W-1) + ceil(log2 d); d == 1 is mul = 0.
// Div32 is exact for numerators < 2^31, Div64 for numerators < 2^63; the
// planner keeps every numerator inside that domain (see fits_index32).
struct Div32 {
    uint32_t d, mul, shift;
    static Div32 make(uint32_t d) {
        Div32 r{d, 0u, 0u};
        if (d > 1) {
            uint32_t p = 31 + (uint32_t)ceil_log2(d);
            r.mul = (uint32_t)(((uint64_t(1) << p) + d - 1) / d);
            r.shift = p - 32;
        }
        return r;
    }
    IPT_HD uint32_t div(uint32_t x) const {
        return mul ? (uint32_t)(((uint64_t)x * mul) >> 32) >> shift : x;
    }
    IPT_HD uint32_t mod(uint32_t x) const { return x - div(x) * d; }
};

struct Div64 {
    uint64_t d, mul;
    uint32_t shift;
    static Div64 make(uint64_t d) {
        Div64 r{d, 0u, 0u};
        if (d > 1) {
            uint32_t p = 63 + (uint32_t)ceil_log2(d);
            r.mul = (uint64_t)(((((unsigned __int128)1) << p) + d - 1) / d);
            r.shift = p - 64;
        }
        return r;
    }
    IPT_HD uint64_t div(uint64_t x) const {
        return mul ? (uint64_t)(((unsigned __int128)x * mul) >> 64) >> shift : x;
    }
    IPT_HD uint64_t mod(uint64_t x) const { return x - div(x) * d; }
};

template <typename Ix> struct DivOf;
template <> struct DivOf<uint32_t> { using type = Div32; };
template <> struct DivOf<uint64_t> { using type = Div64; };

// ---------------------------------------------------------------- plan

// Host-side plan of a view R x C.
struct HostPlan {
    int64_t R, C, c, a, b, k, q;
    static HostPlan make(int64_t R, int64_t C) {
        HostPlan p{};
        int64_t t;
        p.R = R;
        p.C = C;
        extended_gcd(R, C, p.c, t);
        p.a = R / p.c;
        p.b = C / p.c;
        extended_gcd(p.a, p.b, t, p.k);  // C2R: a^-1 mod b (src/transpose.cu:35-40)
        extended_gcd(p.b, p.a, t, p.q);  // R2C: b^-1 mod a (src/transpose.cu:77-82)
        return p;
    }
    // True when every numerator the maps form stays below 2^31, so the
    // 32-bit plan (Div32) is exact: i*C + j and f of the C2R shuffle are
    // <= R*C; j*R + R-1 < R*C + R; k*(x mod b) < b^2; (c-1)*x < c*R;
    // ((c-1+x)/c)*q < (a+2)*a.
    bool fits_index32() const {
        const unsigned __int128 lim = (unsigned __int128)1 << 31;
        auto u = [](int64_t v) { return (unsigned __int128)v; };
        return u(R) * u(C) + u(R) < lim && u(b) * u(b) < lim && u(c) * u(R) + u(R) < lim &&
               (u(a) + 2) * u(a) < lim;
    }
    // The same bounds at 2^63 for the 64-bit plan (Div64).  Slices up to 2^40
    // elements pass every bound but b^2 (k times a residue mod b in the C2R
    // shuffle): a view with b above ~3e9 has no exact plan.
    bool fits_index64() const {
        const unsigned __int128 lim = (unsigned __int128)1 << 63;
        auto u = [](int64_t v) { return (unsigned __int128)v; };
        return u(R) * u(C) + u(R) < lim && u(b) * u(b) < lim && u(c) * u(R) + u(R) < lim &&
               (u(a) + 2) * u(a) < lim;
    }
};

// Device-side plan: the same numbers plus reduced divisors in index type Ix.
template <typename Ix> struct Plan {
    using Div = typename DivOf<Ix>::type;
    Ix R, C, c, a, b, k, q;
    Div dR, dC, dc, da, db;
    static Plan make(const HostPlan& h) {
        Plan p;
        p.R = (Ix)h.R;
        p.C = (Ix)h.C;
        p.c = (Ix)h.c;
        p.a = (Ix)h.a;
        p.b = (Ix)h.b;
        p.k = (Ix)h.k;
        p.q = (Ix)h.q;
        p.dR = Div::make(p.R);
        p.dC = Div::make(p.C);
        p.dc = Div::make(p.c);
        p.da = Div::make(p.a);
        p.db = Div::make(p.b);
        return p;
    }
};

// ---------------------------------------------------------------- maps
//
// Column gathers return the source row g(i,j); row gathers return the source
// column s_i(j).  All results are exact (no reference overflow, App. defects
// of the 32-bit shuffle do not apply: the product k*(x mod b) is reduced in
// the plan's index width, which the planner widens to 64 bits when b^2 does
// not fit).

// c2r prerotator as a gather: (i + j/b) mod R (include/equations.h:78-94,
// include/skinny.h:12-26).  i + j/b < R + c <= 2R.
struct MapC2RPre {
    static constexpr bool kRunConstant = true;  // same source row across runs of b columns
    static constexpr int kDiag = 0;             // 1: C2R post form, 2: R2C pre form
    template <typename Ix> static IPT_HD Ix src(const Plan<Ix>& p, Ix i, Ix j) {
        Ix x = i + p.db.div(j);
        return x >= p.R ? x - p.R : x;
    }
};

// c2r col_op, correct fused form: (i*C - i/a + j) mod R
// (include/skinny.h:69-78 as used at src/skinny.cu:332; the large-path call
// at src/transpose.cu:22 passes d2/c instead of c and is not followed).
struct MapC2RPost {
    static constexpr bool kRunConstant = false;
    static constexpr int kDiag = 1;
    template <typename Ix> static IPT_HD Ix row_term(const Plan<Ix>& p, Ix i) {
        return p.dR.mod(i * p.C - p.da.div(i));
    }
    template <typename Ix> static IPT_HD Ix src(const Plan<Ix>& p, Ix i, Ix j) {
        Ix x = row_term(p, i) + p.dR.mod(j);
        return x >= p.R ? x - p.R : x;
    }
};

// r2c::prepermuter p(x) = a*((c-1)x mod c) + (((c-1+x)/c)*q mod a)
// (include/equations.h:134-157, include/skinny.h:95-101).
template <typename Ix> IPT_HD Ix r2c_prepermute(const Plan<Ix>& p, Ix x) {
    Ix t1 = p.a * p.dc.mod((p.c - 1) * x);
    Ix t2 = p.da.mod(p.dc.div(p.c - 1 + x) * p.q);
    return t1 + t2;
}

// r2c col_op, fused form: p((i - j) mod R) (include/skinny.h:87-107).
struct MapR2CPre {
    static constexpr bool kRunConstant = false;
    static constexpr int kDiag = 2;
    template <typename Ix> static IPT_HD Ix src(const Plan<Ix>& p, Ix i, Ix j) {
        Ix x = i + p.R - p.dR.mod(j);  // in [1, 2R)
        if (x >= p.R) x -= p.R;
        return r2c_prepermute(p, x);
    }
};

// r2c postrotator as a gather: (i + (R - j/b) mod R) mod R
// (include/equations.h:196-222, include/skinny.h:109-117).  j/b < c <= R.
struct MapR2CPost {
    static constexpr bool kRunConstant = true;
    static constexpr int kDiag = 0;
    template <typename Ix> static IPT_HD Ix src(const Plan<Ix>& p, Ix i, Ix j) {
        Ix bd = p.db.div(j);
        Ix x = bd == 0 ? i : i + p.R - bd;
        return x >= p.R ? x - p.R : x;
    }
};

// c2r::shuffle / long_shuffle (include/equations.h:12-54,
// include/skinny.h:28-67): f = j + i(C-1) + (R if i - (j mod c) > R - c),
// s = (k * ((f/c) mod b)) mod b + (f mod c) * b.
struct MapC2RShuffle {
    template <typename Ix> static IPT_HD Ix src(const Plan<Ix>& p, Ix i, Ix j) {
        Ix f = j + i * (p.C - 1);
        if (i + p.c > p.R + p.dc.mod(j)) f += p.R;
        Ix fd = p.dc.div(f);
        Ix fm = f - fd * p.c;
        Ix u = p.db.mod(p.k * p.db.mod(fd));
        return u + fm * p.b;
    }
};

// r2c::shuffle (include/equations.h:178-194): ((j/b + i) mod R + j*R) mod C.
struct MapR2CShuffle {
    template <typename Ix> static IPT_HD Ix src(const Plan<Ix>& p, Ix i, Ix j) {
        Ix x = p.db.div(j) + i;  // < c + R <= 2R
        if (x >= p.R) x -= p.R;
        return p.dC.mod(x + j * p.R);
    }
};

enum MapId : int {
    MAP_C2R_PRE = 0,
    MAP_C2R_POST = 1,
    MAP_R2C_PRE = 2,
    MAP_R2C_POST = 3,
    MAP_C2R_SHUFFLE = 4,
    MAP_R2C_SHUFFLE = 5,
};

}  // namespace ipt
