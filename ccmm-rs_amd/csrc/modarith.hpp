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

// Reduce a 128-bit value (hi, lo) with hi < q (true for any sum of a few
// products of canonical residues when q < 2^62) to [0, q).
// Quotient estimate floor(x * r / 2^128) is short by at most 2.
CK_HD u64 barrett_reduce128(u64 hi, u64 lo, const Barrett &b) {
    // x * r = (hi 2^64 + lo)(r_hi 2^64 + r_lo); need bits >= 128.
    u64 p0h = mulhi64(lo, b.r_lo);
    u64 p1l = lo * b.r_hi, p1h = mulhi64(lo, b.r_hi);
    u64 p2l = hi * b.r_lo, p2h = mulhi64(hi, b.r_lo);
    u64 p3l = hi * b.r_hi;
    // middle column sum with carries
    u64 m = p0h + p1l;
    u64 c = m < p0h;
    u64 m2 = m + p2l;
    c += m2 < m;
    u64 quot = p3l + p1h + p2h + c;
    u64 r = lo - quot * b.q;  // true remainder < 3q fits in 64 bits
    if (r >= b.q) r -= b.q;
    if (r >= b.q) r -= b.q;
    return r;
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

// 128-bit accumulator for sums of products before one reduction.
struct Acc128 {
    u64 hi, lo;
    CK_HD void zero() { hi = lo = 0; }
    CK_HD void mac(u64 a, u64 b) {
        u64 h, l;
        mul128(a, b, h, l);
        lo += l;
        hi += h + (lo < l);
    }
    CK_HD void add(u64 v) {
        lo += v;
        hi += (lo < v);
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
