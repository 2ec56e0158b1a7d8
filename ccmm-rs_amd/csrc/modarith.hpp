// Modular arithmetic on 64-bit residues for gfx950.
//
// CDNA has no 64x64->128 multiply instruction; the compiler lowers the
// 64-bit products below to v_mad_u64_u32 / v_mul_hi_u32 sequences.  The
// same source also compiles for the host (g++) so the kernels' index logic
// and arithmetic can be exercised on the CPU.
//
// Residue ranges: every kernel boundary holds canonical residues in [0, q).
// Inside the NTT, butterflies keep values lazily in [0, 4q), which needs
// q < 2^62 (all plans enforce it).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CK_HD __host__ __device__ __forceinline__
#else
#define CK_HD inline
#endif

namespace ck {

typedef uint64_t u64;
typedef uint32_t u32;

CK_HD u64 mulhi64(u64 a, u64 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (u64)(((unsigned __int128)a * b) >> 64);
#endif
}

// Full 128-bit product as (hi, lo).
CK_HD void mul128(u64 a, u64 b, u64 &hi, u64 &lo) {
    lo = a * b;
    hi = mulhi64(a, b);
}

// Harvey/Shoup multiplication by a fixed w < q with wq = floor(w 2^64 / q).
// For any x < 2^64 the result lies in [0, 2q).
CK_HD u64 shoup_lazy(u64 x, u64 w, u64 wq, u64 q) {
    return x * w - mulhi64(x, wq) * q;
}

CK_HD u64 shoup(u64 x, u64 w, u64 wq, u64 q) {
    u64 r = shoup_lazy(x, w, wq, q);
    return r >= q ? r - q : r;
}

// ---- moduli up to 2^64 (the reference's 64-bit NTT parameter set) ---------
// Canonical residues throughout.  Shoup's remainder x w - floor(x wq / 2^64) q
// lies in [0, 2q) for any q < 2^64 and x < 2^64, which can exceed 2^64 when
// q > 2^63, so its high word (0 or 1) is formed from the two products' high
// words and decides the one correction.
CK_HD u64 shoup_wide(u64 x, u64 w, u64 wq, u64 q) {
    const u64 qh = mulhi64(x, wq);
    const u64 pl = x * w, ql = qh * q;
    const u64 lo = pl - ql;
    const u64 hi = mulhi64(x, w) - mulhi64(qh, q) - (pl < ql ? 1u : 0u);
    return (hi != 0 || lo >= q) ? lo - q : lo;
}

// a + b mod q for a, b < q < 2^64: a carry out of 64 bits means the sum
// exceeds q.
CK_HD u64 add_mod_wide(u64 a, u64 b, u64 q) {
    const u64 s = a + b;
    return (s < a || s >= q) ? s - q : s;
}

CK_HD u64 add_mod(u64 a, u64 b, u64 q) {
    u64 s = a + b;
    return s >= q ? s - q : s;
}

CK_HD u64 sub_mod(u64 a, u64 b, u64 q) {
    return a >= b ? a - b : a + q - b;
}

CK_HD u64 neg_mod(u64 a, u64 q) { return a == 0 ? 0 : q - a; }

// Barrett constant for 128-bit inputs: r = floor(2^128 / q) split in two words.
struct Barrett {
    u64 q;
    u64 r_hi, r_lo;
};

// Reduce a 128-bit value hi 2^64 + lo to [0, q) (q < 2^62) as two Shoup
// multiplications by constants the Barrett pair already holds:
// floor(2^64 / q) = r_hi, and with W = 2^64 mod q = -r_hi q (mod 2^64),
// floor(W 2^64 / q) = floor(2^128 / q) - r_hi 2^64 = r_lo.  So
//   hi W - floor(hi r_lo / 2^64) q  is hi 2^64 mod q, in [0, 2q),
//   lo   - floor(lo r_hi / 2^64) q  is lo mod q,      in [0, 2q),
// for any hi, lo < 2^64; their sum is below 4q < 2^64.
CK_HD u64 barrett_reduce128(u64 hi, u64 lo, const Barrett &b) {
    const u64 w = 0 - b.r_hi * b.q;
    u64 r = (hi * w - mulhi64(hi, b.r_lo) * b.q) + (lo - mulhi64(lo, b.r_hi) * b.q);
    const u64 two_q = b.q << 1;
    r = r >= two_q ? r - two_q : r;
    return r >= b.q ? r - b.q : r;
}

CK_HD u64 mul_mod(u64 a, u64 b, const Barrett &br) {
    u64 hi, lo;
    mul128(a, b, hi, lo);
    return barrett_reduce128(hi, lo, br);
}

// ---- exact FP64 modular arithmetic for primes q < 2^42 --------------------
//
// CDNA4 runs FP64 FMA at full rate but 32-bit integer multiplies at a
// fraction of it, so for the 40-bit primes residues are carried as doubles
// (exact integers) in a signed lazy range.  For integer-valued doubles y, w
// with |y| < 2^45 and 0 <= w < q < 2^42, and wp = fl(w / q):
//   hi = fl(y w), lo = y w - hi exactly (TwoProduct by FMA),
//   k = rint(fl(y wp)) is round(y w / q) or off by one (|error| < 2^-7),
//   r = (hi - k q) + lo is then exact (every intermediate is an integer
//       below 2^53) and r = y w (mod q) with |r| <= 0.51 q.
// No rounding error ever reaches a residue: the result is the exact
// integer, merely not yet canonical.
CK_HD double fp_mulmod(double y, double w, double wp, double q) {
#pragma clang fp contract(off)
    double hi = y * w;
    double lo = __builtin_fma(y, w, -hi);
    double k = __builtin_rint(y * wp);
    double r = __builtin_fma(-k, q, hi);
    return r + lo;
}

// x (integer-valued, |x| < 2^45) -> x mod q in (-0.5q, 0.5q].
CK_HD double fp_reduce(double x, double q, double qinv) {
#pragma clang fp contract(off)
    double k = __builtin_rint(x * qinv);
    return __builtin_fma(-k, q, x);
}

// Canonical residue in [0, q) of an integer-valued double, |x| < 2^45.
CK_HD u64 fp_canonical(double x, double q, double qinv) {
    double r = fp_reduce(x, q, qinv);
    r = r < 0.0 ? r + q : r;
    r = r >= q ? r - q : r;
    return (u64)r;
}

CK_HD double u64_as_f64(u64 v) {
    double d;
    __builtin_memcpy(&d, &v, 8);
    return d;
}

CK_HD u64 f64_as_u64(double d) {
    u64 v;
    __builtin_memcpy(&v, &d, 8);
    return v;
}

// Exact conversions of integers in [0, 2^52) through the 2^52 exponent bias:
// the double 2^52 + v carries v in its mantissa bits.
CK_HD double small_to_f64(u64 v) { return u64_as_f64(v | 0x4330000000000000ull) - 4503599627370496.0; }
CK_HD u64 f64_to_small(double d) { return f64_as_u64(d + 4503599627370496.0) - 0x4330000000000000ull; }

// The base-conversion target sum_m c_m w_m mod q over K <= 8 digits
// c_m < 2^60 given as halves cl_m, ch_m < 2^30 (c = ch 2^30 + cl), and
// weights w_m < q < 2^41 given as w[2m] = wl_m < 2^20, w[2m+1] = wh_m < 2^21
// (w = wh 2^20 + wl).  Each product of a half and a weight piece is below
// 2^51, an exact double, and the partial sums stay exact integers below
// 2^53: sum_m cl wl and sum_m ch wl have at most 8 terms below 2^50, the
// wh sums are cut into groups of at most 4 terms below 2^51.  So
//   sum = S_ll + 2^20 S_lh + 2^30 S_hl + 2^50 S_hh
// takes FMAs, one fp_reduce per partial sum and three fp_mulmod by
// k20 = 2^20, k30 = 2^30, k50 = 2^50 mod q (with fl(k / q)): 28 FP64
// operations for K = 7 where a fp_mulmod per half and weight takes 84.
// fp_reduce of a sum below 2^53 stays exact (the FMA's exact result is the
// remainder); its quotient estimate is off by less than 2^-13 for
// q >= 2^20 (|r| <= 0.51 q, |result| <= 2.1 q) and by less than 1 for any
// q (|r| < 1.5 q), inside fp_mulmod's |y| < 2^45 either way.
template <int K, class Row>
CK_HD double fp_split_target(const double *cl, const double *ch, const Row &w, double q, double qinv, double k20,
                             double k20p, double k30, double k30p, double k50, double k50p) {
#pragma clang fp contract(off)
    double sll = 0.0, shl = 0.0, slh0 = 0.0, slh1 = 0.0, shh0 = 0.0, shh1 = 0.0;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        const double wl = w[2 * m], wh = w[2 * m + 1];
        sll = __builtin_fma(cl[m], wl, sll);
        shl = __builtin_fma(ch[m], wl, shl);
        if (m < 4) {
            slh0 = __builtin_fma(cl[m], wh, slh0);
            shh0 = __builtin_fma(ch[m], wh, shh0);
        } else {
            slh1 = __builtin_fma(cl[m], wh, slh1);
            shh1 = __builtin_fma(ch[m], wh, shh1);
        }
    }
    double rlh = fp_reduce(slh0, q, qinv), rhh = fp_reduce(shh0, q, qinv);
    if (K > 4) {
        rlh += fp_reduce(slh1, q, qinv);
        rhh += fp_reduce(shh1, q, qinv);
    }
    return (fp_reduce(sll, q, qinv) + fp_mulmod(rlh, k20, k20p, q)) +
           (fp_mulmod(fp_reduce(shl, q, qinv), k30, k30p, q) + fp_mulmod(rhh, k50, k50p, q));
}

// Canonical FP64 residue in [0, q) of an integer-valued double, |x| < 2^45.
CK_HD double fp_canon(double x, double q, double qinv) {
    double r = fp_reduce(x, q, qinv);
    r = r < 0.0 ? r + q : r;
    return r >= q ? r - q : r;
}

// fp_canonical for results that stay below 2^52 (every q < 2^41 here).
CK_HD u64 fp_residue(double x, double q, double qinv) {
    double r = fp_reduce(x, q, qinv);
    r = r < 0.0 ? r + q : r;
    r = r >= q ? r - q : r;
    return f64_to_small(r);
}

// Streaming accesses: data a kernel reads or writes exactly once go through
// the caches with the nontemporal hint (the `nt` bit of global loads and
// stores), so they do not evict what the kernel reuses from its XCD's L2:
// the source limbs of a Galois gather, twiddle and constant tables.
CK_HD u64 ld_stream(const u64 *p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

CK_HD void st_stream(u64 *p, u64 v) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// 128-bit accumulator for sums of products of 64-bit words before one
// reduction.  The second factor is cut into 21-bit pieces b = b0 + b1 2^21 +
// b2 2^42 (b < 2^63) and the first into its 32-bit words, so each of the six
// partial products a_i b_j is below 2^53 and goes straight into its own
// 64-bit sum by one v_mad_u64_u32, with no carries: up to 1024 products fit.
// A full 128-bit product per term would recompute its cross terms for the
// low and the high word and track two carries.  fold() recombines the sums
// s_ij 2^(32 i + 21 j) modulo 2^128.  A factor that stays fixed over many
// products (a key entry, a table constant) is cut once with split21.
struct Split21 {
    u64 b0, b1, b2;
};

CK_HD Split21 split21(u64 b) { return Split21{b & 0x1fffffull, (b >> 21) & 0x1fffffull, b >> 42}; }

struct Acc128 {
    u64 s00, s01, s02, s10, s11, s12;  // s_ij = sum a_i b_j
    CK_HD void zero() { s00 = s01 = s02 = s10 = s11 = s12 = 0; }
    CK_HD void mac(u64 a, const Split21 &b) {
        const u64 a0 = (u32)a, a1 = a >> 32;
        s00 += a0 * b.b0;
        s01 += a0 * b.b1;
        s02 += a0 * b.b2;
        s10 += a1 * b.b0;
        s11 += a1 * b.b1;
        s12 += a1 * b.b2;
    }
    CK_HD void mac(u64 a, u64 b) { mac(a, split21(b)); }
    CK_HD void fold(u64 &hi, u64 &lo) const {
        // weights 2^0, 2^21, 2^42, 2^32, 2^53, 2^74
        u64 l = s00, h = 0, t;
        t = s01 << 21;
        l += t;
        h += (s01 >> 43) + (l < t);
        t = s02 << 42;
        l += t;
        h += (s02 >> 22) + (l < t);
        t = s10 << 32;
        l += t;
        h += (s10 >> 32) + (l < t);
        t = s11 << 53;
        l += t;
        h += (s11 >> 11) + (l < t);
        h += s12 << 10;
        hi = h;
        lo = l;
    }
    // sum mod q for a sum below q 2^64 (barrett_reduce128's input range)
    CK_HD u64 reduce(const Barrett &b) const {
        u64 hi, lo;
        fold(hi, lo);
        return barrett_reduce128(hi, lo, b);
    }
};

#if defined(__HIPCC__)
// Read-only kernel tables through the scalar unit: a wave-uniform address in
// the constant address space (4) compiles to s_load, so a table entry costs
// neither a vector-memory round trip nor a VGPR per lane.  Only for memory
// that no kernel writes (the scalar cache is not coherent with vector stores).
#define CK_CONST_AS __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const CK_CONST_AS T *const_table(const T *p) {
    return (const CK_CONST_AS T *)p;
}
#endif

}  // namespace ck
