// Exact CRT of RNS residues to a centered multiword integer held in
// registers (ring/wide_crt.rs:11-120 reconstruct_coefficients_big +
// centered_representative_big), shared by the bounded gadget decomposition
// and the decoder.
//
// One thread per coefficient builds X in [0, Q) in W 64-bit words:
//   v_i = (x_i - X mod q_i) * (q_0 .. q_{i-1})^-1 mod q_i,
//   X  += v_i * (q_0 .. q_{i-1})                      (Garner, limb by limb)
// with X mod q_i = sum_k X[k] * (2^(64k) mod q_i); then X > floor(Q/2) is
// replaced by X - Q in two's complement (value > floor(Q/2) -> negative, as
// the reference's centered map).
#pragma once
#include <vector>

#include "internal.hpp"
#include "host_math.hpp"

namespace ck {

struct CrtTable {
    const u64 *pow64;   // [q_limbs][W]: 2^(64k) mod q_i
    const u64 *prefix;  // [q_limbs][W]: q_0 .. q_{i-1} (1 for i = 0)
    const u64 *inv;     // [q_limbs]: (q_0 .. q_{i-1})^-1 mod q_i
    const u64 *half;    // [W]: floor(Q / 2)
    const u64 *qw;      // [W]: Q
};

// x: limb i of the coefficient at x[i * limb_stride]; neg: the gathered
// coefficient is negated (Galois gather).  Returns true for a negative value.
// The prime and CRT tables are wave-uniform and read through the scalar unit
// (no kernel writes them).
template <int W>
__device__ __forceinline__ bool crt_centered(const u64 *x, u64 limb_stride, bool neg, u32 q_limbs,
                                             const DevPrime *primes, const CrtTable &t, u64 (&X)[W]) {
#pragma unroll
    for (int k = 0; k < W; ++k) X[k] = 0;
    const CK_CONST_AS DevPrime *pr = const_table(primes);
    for (u32 i = 0; i < q_limbs; ++i) {
        const Barrett br{pr[i].q, pr[i].br_hi, pr[i].br_lo};
        const u64 q = br.q;
        u64 xi = x[(u64)i * limb_stride];
        if (neg) xi = neg_mod(xi, q);
        const CK_CONST_AS u64 *pw = const_table(t.pow64) + (u64)i * W;
        u64 r = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) r = add_mod(r, mul_mod(X[k], pw[k], br), q);
        const u64 v = mul_mod(sub_mod(xi, r, q), const_table(t.inv)[i], br);
        const CK_CONST_AS u64 *pre = const_table(t.prefix) + (u64)i * W;
        u64 carry = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            u64 hi, lo;
            mul128(v, pre[k], hi, lo);
            const u64 l1 = lo + carry;
            hi += l1 < lo;
            const u64 l2 = l1 + X[k];
            hi += l2 < l1;
            X[k] = l2;
            carry = hi;
        }
    }
    const CK_CONST_AS u64 *half = const_table(t.half);
    const CK_CONST_AS u64 *qw = const_table(t.qw);
    int cmp = 0;
#pragma unroll
    for (int k = W - 1; k >= 0; --k)
        if (cmp == 0 && X[k] != half[k]) cmp = X[k] > half[k] ? 1 : -1;
    const bool negative = cmp > 0;
    if (negative) {
        u64 borrow = 0;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const u64 q = qw[k];
            const u64 d1 = X[k] - q;
            const u64 b1 = X[k] < q;
            const u64 d2 = d1 - borrow;
            const u64 b2 = d1 < borrow;
            X[k] = d2;
            borrow = b1 | b2;
        }
    }
    return negative;
}

// ---- host side ----------------------------------------------------------------

// Instantiated word counts: the smallest holding bits(Q) + 1 bits (the sign
// of X - Q) is used.
static const unsigned kCrtWords[] = {4, 8, 12, 16, 24, 32};

struct CrtHost {
    unsigned words = 0, bits = 0;
    std::vector<u64> blob;  // pow64 | prefix | inv | half | qw
};

inline void words_mul_small(std::vector<u64> &x, u64 m) {
    unsigned __int128 carry = 0;
    for (auto &w : x) {
        unsigned __int128 t = (unsigned __int128)w * m + carry;
        w = (u64)t;
        carry = t >> 64;
    }
}

// Returns false if Q needs more than 2047 bits.
inline bool build_crt(const std::vector<u64> &mods, size_t q_limbs, CrtHost &h) {
    std::vector<u64> q(40, 0);
    q[0] = 1;
    for (size_t i = 0; i < q_limbs; ++i) words_mul_small(q, mods[i]);
    h.bits = 0;
    for (size_t k = q.size(); k-- > 0;)
        if (q[k]) {
            h.bits = (unsigned)(64 * k + 64 - __builtin_clzll(q[k]));
            break;
        }
    h.words = 0;
    for (unsigned w : kCrtWords)
        if (64 * w >= h.bits + 1) {
            h.words = w;
            break;
        }
    if (!h.words) return false;
    const unsigned W = h.words;
    h.blob.assign((size_t)q_limbs * W * 2 + q_limbs + 2 * W, 0);
    u64 *pow64 = h.blob.data(), *prefix = pow64 + q_limbs * W, *inv = prefix + q_limbs * W;
    u64 *half = inv + q_limbs, *qw = half + W;
    std::vector<u64> pre(W, 0);
    pre[0] = 1;
    for (size_t i = 0; i < q_limbs; ++i) {
        const u64 qi = mods[i];
        const u64 two64 = (u64)((((unsigned __int128)1) << 64) % qi);
        u64 p = 1 % qi;
        for (unsigned k = 0; k < W; ++k) {
            pow64[i * W + k] = p;
            p = host::mulmod(p, two64, qi);
        }
        u64 pm = 0;  // prefix mod q_i, Horner over the words
        for (unsigned k = W; k-- > 0;) pm = (host::mulmod(pm, two64, qi) + pre[k] % qi) % qi;
        inv[i] = host::invmod(pm, qi);
        for (unsigned k = 0; k < W; ++k) prefix[i * W + k] = pre[k];
        words_mul_small(pre, qi);
    }
    for (unsigned k = 0; k < W; ++k) {
        qw[k] = q[k];
        half[k] = (q[k] >> 1) | (k + 1 < q.size() ? q[k + 1] << 63 : 0);
    }
    return true;
}

inline CrtTable crt_table(const u64 *d_blob, size_t q_limbs, unsigned W) {
    CrtTable t;
    t.pow64 = d_blob;
    t.prefix = t.pow64 + q_limbs * W;
    t.inv = t.prefix + q_limbs * W;
    t.half = t.inv + q_limbs;
    t.qw = t.half + W;
    return t;
}

}  // namespace ck
