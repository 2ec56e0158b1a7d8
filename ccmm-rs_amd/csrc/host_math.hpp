// Host-side setup arithmetic: root selection and twiddle tables.
// Runs once per prime when a plan is created; never on the hot path.
#pragma once
#include <stdint.h>

#include <vector>

namespace ck {
namespace host {

typedef unsigned __int128 u128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        b = mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}

inline uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a, q - 2, q); }

inline uint64_t shoup_quotient(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 64) / q); }

inline void barrett(uint64_t q, uint64_t &hi, uint64_t &lo) {
    u128 max = ~(u128)0;
    u128 r = max / q;
    if (max % q == (u128)q - 1) r += 1;
    hi = (uint64_t)(r >> 64);
    lo = (uint64_t)r;
}

// Root selection of ring/ntt_params.rs:16-50: psi = g^((q-1)/2N) for the
// first g >= 2 with psi^N = -1 and psi^2N = 1.  Returns 0 if none exists
// (2N must divide q-1).
inline uint64_t negacyclic_root(uint64_t q, uint64_t degree) {
    uint64_t two_n = 2 * degree;
    if (degree == 0 || (q - 1) % two_n) return 0;
    uint64_t e = (q - 1) / two_n;
    for (uint64_t g = 2; g < q; ++g) {
        uint64_t psi = powmod(g, e, q);
        if (powmod(psi, degree, q) == q - 1 && powmod(psi, two_n, q) == 1) return psi;
    }
    return 0;
}

inline uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

// Twiddle tables T[k] = root^bitrev_n(k) with Shoup quotients, interleaved
// (w, wq) per entry, k in [0, N).  Entry 0 is unused by the kernels.
inline void twiddle_table(uint64_t q, uint64_t root, uint32_t logn, std::vector<uint64_t> &out) {
    uint32_t n = 1u << logn;
    std::vector<uint64_t> pw(n);
    uint64_t v = 1;
    for (uint32_t j = 0; j < n; ++j) {
        pw[j] = v;
        v = mulmod(v, root, q);
    }
    out.assign(2 * (size_t)n, 0);
    for (uint32_t k = 0; k < n; ++k) {
        uint64_t w = pw[bitrev(k, logn)];
        out[2 * k] = w;
        out[2 * k + 1] = shoup_quotient(w, q);
    }
}

// Moduli below this bound run the NTT on the exact FP64 path.
constexpr uint64_t kFpModulusBound = 1ull << 41;

inline uint64_t f64_bits(double d) {
    uint64_t v;
    __builtin_memcpy(&v, &d, 8);
    return v;
}

// FP64-path table: (double w, double fl(w / q)) as raw bits.
inline void twiddle_table_fp(uint64_t q, uint64_t root, uint32_t logn, std::vector<uint64_t> &out) {
    twiddle_table(q, root, logn, out);
    for (size_t k = 0; k < out.size() / 2; ++k) {
        double w = (double)out[2 * k];
        out[2 * k] = f64_bits(w);
        out[2 * k + 1] = f64_bits(w / (double)q);
    }
}

// Row factors of the two-pass NTT (n = n1 + n2, rows u < 2^n1, row stages
// s < n2), FP64 layout like twiddle_table_fp, entry u * n2 + s.  Row u's
// twiddle at stage s and offset j < 2^s is T[(2^n1 + bitrev_n1(u)) 2^s + j]
// with T[k] = root^bitrev_n(k); bit-reversing that index gives the exponent
//   bitrev_s(j) 2^(n-s) + u 2^(n2-s) + 2^(n2-1-s),
// while T[2^s + j] = root^(bitrev_s(j) 2^(n-s) + 2^(n-1-s)), so
//   T[(2^n1 + bitrev_n1(u)) 2^s + j] = T[2^s + j] * root^((2u+1) 2^(n2-1-s) - 2^(n-1-s)):
// the first factor is shared by every row, the second is this table.
inline void row_factor_table_fp(uint64_t q, uint64_t root, uint32_t logn, std::vector<uint64_t> &out) {
    const uint32_t n2 = (logn + 1) / 2, n1 = logn - n2;
    const uint64_t two_n = 2ull << logn;  // root^(2N) = 1
    out.assign(2 * ((size_t)1 << n1) * n2, 0);
    for (uint64_t u = 0; u < (1ull << n1); ++u)
        for (uint32_t s = 0; s < n2; ++s) {
            const uint64_t e = ((2 * u + 1) << (n2 - 1 - s)) % two_n;
            const uint64_t sub = (1ull << (logn - 1 - s)) % two_n;
            const uint64_t w = powmod(root, (e + two_n - sub) % two_n, q);
            const size_t k = (size_t)u * n2 + s;
            out[2 * k] = f64_bits((double)w);
            out[2 * k + 1] = f64_bits((double)w / (double)q);
        }
}

// Deterministic Miller-Rabin for 64-bit values (witness set of
// ring/ntt_prime.rs:8-33).
inline bool is_prime(uint64_t v) {
    static const uint64_t small[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    static const uint64_t wit[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (v < 2) return false;
    for (uint64_t p : small) {
        if (v == p) return true;
        if (v % p == 0) return false;
    }
    uint64_t d = v - 1;
    unsigned s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    for (uint64_t w : wit) {
        uint64_t a = w % v;
        if (!a) continue;
        uint64_t x = powmod(a, d, v);
        if (x == 1 || x == v - 1) continue;
        bool ok = false;
        for (unsigned r = 1; r < s; ++r) {
            x = mulmod(x, x, v);
            if (x == v - 1) {
                ok = true;
                break;
            }
        }
        if (!ok) return false;
    }
    return true;
}

}  // namespace host
}  // namespace ck
