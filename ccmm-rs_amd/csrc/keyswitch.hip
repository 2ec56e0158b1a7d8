// Key switching (gfx950): RNS-gadget (reference family, no special primes)
// and the dnum-block hybrid with special primes P.
//
// Both run the same exact pipeline over the key basis B (B = Q for the RNS
// gadget, B = Q u P for the hybrid):
//   1. ModUp: digit_j = canonical x mod Q_j (grafting/rns_gadget.rs:143-228)
//      lifted exactly to every limb of B by mixed-radix (Garner) digits
//      (ring/basis_transition.rs:72-138).  For limbs inside block j the digit
//      is the input residue itself.
//   2. forward NTT of every digit limb;
//   3. MAC with the key: acc_c[i] = sum_j D_j[i] * key_c[j][i]  (c = b, a),
//      one 128-bit accumulation and one reduction per residue;
//   4. inverse NTT of both accumulators;
//   5. hybrid: centered ModDown by P (ring/auxiliary_modulus.rs:89-171) with
//      an exact P-side Garner and a lexicographic mixed-radix comparison
//      against floor(P/2); then the pass-through component is added
//      (b for key switch / rotation, c0 and c1 for relinearization).
// Adding b after ModDown gives the same residues as ModDown(P b + acc_b)
// (ckks/hybrid_rns_galois_key.rs:633-663) because P b = 0 mod P.
#include <algorithm>
#include <cstring>

#include "baseconv.hpp"
#include "crt.hpp"
#include "host_math.hpp"
#include "internal.hpp"

namespace ck {

constexpr int kMaxAlpha = 12;   // limbs per gadget block handled by ModUp
constexpr int kMaxSpecial = 8;  // special primes handled by ModDown
constexpr int kKsThreads = 256;
constexpr int kFpRow = 4 + 2 * kMaxAlpha;      // doubles per FP64 ModUp row
// doubles per FP64 ModDown row: the fp_mulmod weights of each P digit, then
// the split weights and constants of fp_split_target
constexpr int kMdSplitW = 6 + 4 * kMaxSpecial;
constexpr int kMdSplitK = kMdSplitW + 2 * kMaxSpecial;
constexpr int kMdRow = kMdSplitK + 6;

// Constants of the exact lift of one gadget block.
struct ModUpBlock {
    u32 start, size;
    // Garner: c_k = (x_k - sum_{m<k} c_m Pm[k][m]) * inv[k]  mod q_{start+k}
    u64 pm[kMaxAlpha][kMaxAlpha];
    u64 inv[kMaxAlpha], inv_q[kMaxAlpha];
    // evaluation of the mixed-radix value at every basis limb i:
    // w[i * kMaxAlpha + m] = (q_start ... q_{start+m-1}) mod q_i
    const u64 *w;
    // FP64 modes (fp_mode 1 / 2: every block prime after the first is below
    // 2^41; mode 2 when the first is not, and digit 0 enters split at bit 30,
    // c_0 = ch 2^30 + cl, both below 2^45 as fp_mulmod requires); every
    // other digit is below 2^41.
    // Garner row gf[k] (k >= 1, q_k < 2^41): q_k, 1/q_k, inv[k], fl(inv[k] / q_k),
    // pm[k][0] 2^30 mod q_k, its quotient, then (pm[k][m], fl(pm[k][m] / q_k))
    // for m < k.  Target row tf + i * kFpRow (q_i < 2^41): q_i, 1/q_i,
    // w_i0 2^30 mod q_i, its quotient, then (w_im, fl(w_im / q_i)) for m < size.
    u32 fp_mode;
    double gf[kMaxAlpha][kFpRow];
    const double *tf;
};

struct ModDownTable {
    u32 q_limbs, p_limbs;
    u64 pm[kMaxSpecial][kMaxSpecial];  // Garner over P
    u64 inv[kMaxSpecial], inv_q[kMaxSpecial];
    u64 half[kMaxSpecial];  // mixed-radix digits of floor(P/2)
    // per Q limb i: w[i*kMaxSpecial + m] = (p_0..p_{m-1}) mod q_i,
    // pmod[i] = P mod q_i, pinv[i] = P^-1 mod q_i (Shoup pair)
    const u64 *w;
    const u64 *pmod, *pinv, *pinv_q;
    // FP64 target row tf + i * kMdRow (q_i < 2^41): q_i, 1/q_i, P mod q_i,
    // P^-1 mod q_i, fl(P^-1 / q_i), 0, then per P digit m (split at bit 30):
    // w_im, fl(w_im / q_i), w_im 2^30 mod q_i, fl(that / q_i); from
    // kMdSplitW per digit w_im mod 2^20, w_im >> 20, and from kMdSplitK
    // 2^20, 2^30, 2^50 mod q_i each with fl(that / q_i)
    const double *tf;
};

}  // namespace ck

// permuted key copies kept per bounded key for hoisted rotations (LRU)
constexpr size_t kHoistKeyCacheMax = 8;

struct ckks_ksk {
    ckks_context *ctx = nullptr;
    const ckks_rns_plan *plan = nullptr;  // key basis B
    int hybrid = 0;
    size_t q_limbs = 0, p_limbs = 0, limbs = 0, dnum = 0, degree = 0;
    std::vector<size_t> block_start, block_size;
    std::vector<int> block_fp;  // ModUp of the block runs in FP64 mode
    ck::u64 fp_targets = 0, fp_q_targets = 0;  // FP64 target limbs of B / of Q
    ck::u64 *d_kb = nullptr, *d_ka = nullptr;  // [dnum][B][N], NTT domain
    ck::ModUpBlock *d_modup = nullptr;
    std::vector<ck::ModUpBlock> h_modup;
    ck::u64 *d_consts = nullptr;
    ck::ModDownTable *d_moddown = nullptr;
    // exact int8-MFMA base conversion: dnum ModUp plans then the ModDown plan
    int mfma = 0;
    // bounded (balanced base-2^w) gadget: dnum = digit count, no blocks
    int bounded = 0;
    unsigned base_log = 0, bounded_words = 0;
    ck::u64 *d_bounded = nullptr;  // table blob (BoundedTable points into it)
    int tuned_engine = CKKS_BASECONV_VALU;  // engine of CKKS_BASECONV_AUTO for this key
    int tuned_fusion = CKKS_KS_FUSION_OFF;  // mode of CKKS_KS_FUSION_AUTO for this key
    int tuned_split = CKKS_KS_SPLIT_OFF;    // mode of CKKS_KS_SPLIT_AUTO for this key
    // hoisted rotations: key entries [2][D][B][N] permuted by k^-1, per k,
    // at most kHoistKeyCacheMax of them; hoist_lru lists the exponents most
    // recently used first, and the last one is freed when the cache is full
    mutable std::map<ck::u64, ck::u64 *> hoist_keys;
    mutable std::vector<ck::u64> hoist_lru;
    ck::BcPlan *d_bc = nullptr;
    signed char *d_bt = nullptr;
};

// Digits of a ciphertext batch's a components for hoisted bounded rotations:
// decomposed, lifted to the key basis and transformed once.
struct ckks_hoisted {
    ckks_context *ctx = nullptr;
    const ckks_rns_plan *plan = nullptr;
    size_t q_limbs = 0, digit_count = 0, n_cts = 0;
    unsigned base_log = 0;
    // RNS-gadget / dnum-hybrid digits (ckks_hoist): the block layout they
    // were cut with; empty for bounded digits
    std::vector<size_t> blocks;
    ck::u64 *d_digits = nullptr;  // [ct][digit][B][N], NTT domain
    // (ct, coefficient) whose digit vector holds a -2^(w-1): negating it
    // carries, so a rotation that negates one takes the canonical path
    std::vector<std::pair<ck::u32, ck::u32>> ties;
    // canonical digits: some coefficient has every residue of a gadget block
    // zero, so rotations take the canonical path
    bool zero_block = false;
    int *d_flag = nullptr, *h_flag = nullptr;  // the zero-block check's result (device, pinned host)
    // ckks_hoist_with_b: P NTT(b) of the b components, [ct][q_limbs][N]
    // (P the product of the key's special primes, 1 without them)
    ck::u64 *d_bhat = nullptr;
    ~ckks_hoisted() {
        if (d_digits) hipFree(d_digits);
        if (d_bhat) hipFree(d_bhat);
        if (d_flag) hipFree(d_flag);
        if (h_flag) hipHostFree(h_flag);
    }
};

namespace ck {

// ---- device helpers --------------------------------------------------------

CK_HD u64 reduce128(const Acc128 &a, const DevPrime &P) {
    return a.reduce(Barrett{P.q, P.br_hi, P.br_lo});
}

// ---- 1. ModUp -------------------------------------------------------------------

struct ModUpArgs {
    const u64 *in;  // component being decomposed: limb i of ct c at in + c*in_stride + i*N
    u64 in_stride;
    u64 kinv;  // nonzero: decompose sigma_k(in), k^-1 mod 2N (rotation fused in)
    u64 *digits;  // [ct][dnum][B][N]
    const ModUpBlock *blocks;
    const DevPrime *primes;  // key basis B
    u64 fp_targets;  // bit i: limb i takes the FP64 target path (fp_mode blocks)
    u32 dnum, limbs, logn, chunks;
};

// Coefficient n of sigma_k(p) is +-p[g]: g = n k^-1 mod 2N, negated when
// g >= N (apply_automorphism, ckks/automorphism.rs:8-46, as a gather).
struct Gather {
    u64 idx;
    bool neg;
};

CK_HD Gather galois_gather(u64 n, u64 kinv, u32 logn) {
    if (kinv == 0) return Gather{n, false};
    const u64 N = (u64)1 << logn;
    u64 g = (n * kinv) & (2 * N - 1);
    return g < N ? Gather{g, false} : Gather{g - N, true};
}

__device__ __forceinline__ Barrett barrett_at(const CK_CONST_AS DevPrime *p, u32 i) {
    return Barrett{p[i].q, p[i].br_hi, p[i].br_lo};
}

// The block and prime tables are read through the scalar unit (const_table):
// every address below except the residues' is uniform over the wave.
//
// Integer mode: Garner digits and every target by 128-bit accumulation and
// Barrett reduction.
//
// FP64 modes (MODE = blk->fp_mode = 1 or 2): Garner digits c_k (k >= 1) and
// the targets whose bit is set in a.fp_targets (q_i < 2^41) as sums of
// fp_mulmod terms, each |term| <= 0.51 q, at most kMaxAlpha + 1 of them, so
// every sum stays below 2^45 and is exact; fp_canon / fp_residue make it
// canonical.  Digit 0 is x_0 itself: whole in mode 1 (first block prime below
// 2^41), split at bit 30 in mode 2; the other digits are below 2^41.
template <int A, int MODE>
__device__ void modup_thread(const ModUpArgs &a, const CK_CONST_AS ModUpBlock *blk, u32 ct, u32 j, u64 n) {
    const u64 N = (u64)1 << a.logn;
    const CK_CONST_AS DevPrime *primes = const_table(a.primes);
    const u32 start = blk->start;
    const Gather gi = galois_gather(n, a.kinv, a.logn);
    const u64 *src = a.in + ct * a.in_stride + (u64)start * N + gi.idx;
    u64 *dst = a.digits + (((u64)ct * a.dnum + j) * a.limbs) * N + n;
    u64 x[A];
#pragma unroll
    for (int m = 0; m < A; ++m) {
        u64 v = src[(u64)m * N];
        x[m] = gi.neg ? neg_mod(v, primes[start + m].q) : v;
        // inside the block the canonical digit reduces to the residue
        st_stream(dst + (u64)(start + m) * N, x[m]);
    }
    const CK_CONST_AS u64 *wt = const_table(blk->w);
    if (MODE == 0) {
        u64 c[A];
#pragma unroll
        for (int k = 0; k < A; ++k) {
            if (k == 0) {
                c[0] = x[0];
            } else {
                const Barrett bk = barrett_at(primes, start + k);
                Acc128 t;
                t.zero();
#pragma unroll
                for (int m = 0; m < k; ++m) t.mac(c[m], blk->pm[k][m]);
                u64 tr = t.reduce(bk);
                c[k] = shoup(sub_mod(x[k], tr, bk.q), blk->inv[k], blk->inv_q[k], bk.q);
            }
        }
        for (u32 i = 0; i < a.limbs; ++i) {
            if (i - start < (u32)A) continue;  // block limbs are stored above
            const Barrett bi = barrett_at(primes, i);
            const CK_CONST_AS u64 *w = wt + (u64)i * kMaxAlpha;
            u64 v;
            if (A == 1) {
                v = c[0] >= bi.q ? barrett_reduce128(0, c[0], bi) : c[0];
            } else {
                Acc128 t;
                t.zero();
#pragma unroll
                for (int m = 0; m < A; ++m) t.mac(c[m], w[m]);
                v = t.reduce(bi);
            }
            st_stream(dst + (u64)i * N, v);
        }
        return;
    }
    constexpr bool SPLIT0 = MODE == 2;
    double cl[A];  // cl[0]: c_0 = x_0, or its low 30 bits when split
    const double ch0 = SPLIT0 ? small_to_f64(x[0] >> 30) : 0.0;
    cl[0] = small_to_f64(SPLIT0 ? x[0] & 0x3fffffffull : x[0]);
#pragma unroll
    for (int k = 1; k < A; ++k) {
        const CK_CONST_AS double *g = blk->gf[k];
        const double q = g[0];
        double t = small_to_f64(x[k]) - fp_mulmod(cl[0], g[6], g[7], q);
        if (SPLIT0) t -= fp_mulmod(ch0, g[4], g[5], q);
#pragma unroll
        for (int m = 1; m < k; ++m) t -= fp_mulmod(cl[m], g[6 + 2 * m], g[7 + 2 * m], q);
        cl[k] = fp_canon(fp_mulmod(t, g[2], g[3], q), q, g[1]);
    }
    const CK_CONST_AS double *tf = const_table(blk->tf);
    const u64 fpm = a.fp_targets;
    for (u32 i = 0; i < a.limbs; ++i) {
        if (i - start < (u32)A) continue;  // block limbs are stored above
        u64 v;
        if (i < 64 && ((fpm >> i) & 1)) {  // the mask is empty for bases over 64 limbs
            const CK_CONST_AS double *r = tf + (u64)i * kFpRow;
            const double q = r[0];
            double acc = fp_mulmod(cl[0], r[4], r[5], q);
            if (SPLIT0) acc += fp_mulmod(ch0, r[2], r[3], q);
#pragma unroll
            for (int m = 1; m < A; ++m) acc += fp_mulmod(cl[m], r[4 + 2 * m], r[5 + 2 * m], q);
            v = fp_residue(acc, q, r[1]);
        } else {
            const Barrett bi = barrett_at(primes, i);
            const CK_CONST_AS u64 *w = wt + (u64)i * kMaxAlpha;
            Acc128 t;
            t.zero();
            t.mac(x[0], w[0]);
#pragma unroll
            for (int m = 1; m < A; ++m) t.mac(f64_to_small(cl[m]), w[m]);
            v = t.reduce(bi);
        }
        st_stream(dst + (u64)i * N, v);
    }
}

// One instantiation per block size A (and mode), so each gets its own
// register budget; a launch covers the consecutive digits
// [j_first, j_first + n_j) of one size.
template <int A, int MODE>
__global__ __launch_bounds__(kKsThreads) void k_modup(ModUpArgs a, u32 j_first, u32 n_j) {
    const u32 b = xcd_contiguous_block(blockIdx.x, gridDim.x);
    const u32 chunk = b % a.chunks;
    const u32 cj = b / a.chunks;
    const u32 j = j_first + cj % n_j, ct = cj / n_j;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    modup_thread<A, MODE>(a, const_table(a.blocks) + j, ct, j, n);
}

// sizes / fp: per digit, its block size and its mode (0 integer, 1 / 2 FP64).
static int launch_modup(hipStream_t s, const ModUpArgs &mu, const std::vector<size_t> &sizes,
                        const std::vector<int> &fp, size_t n_cts) {
    for (size_t j0 = 0; j0 < sizes.size();) {
        size_t j1 = j0 + 1;
        while (j1 < sizes.size() && sizes[j1] == sizes[j0] && fp[j1] == fp[j0]) ++j1;
        const dim3 g((u32)(n_cts * (j1 - j0) * mu.chunks));
        const u32 jf = (u32)j0, nj = (u32)(j1 - j0);
        const int mode = fp[j0];
        switch (sizes[j0]) {
#define CK_MU(A)                                                                                  \
    case A:                                                                                       \
        if (mode == 1) hipLaunchKernelGGL((k_modup<A, 1>), g, dim3(kKsThreads), 0, s, mu, jf, nj); \
        else if (mode == 2) hipLaunchKernelGGL((k_modup<A, 2>), g, dim3(kKsThreads), 0, s, mu, jf, nj); \
        else hipLaunchKernelGGL((k_modup<A, 0>), g, dim3(kKsThreads), 0, s, mu, jf, nj);         \
        break;
            CK_MU(1) CK_MU(2) CK_MU(3) CK_MU(4) CK_MU(5) CK_MU(6)
            CK_MU(7) CK_MU(8) CK_MU(9) CK_MU(10) CK_MU(11) CK_MU(12)
#undef CK_MU
            default: return fail(CKKS_ERR_UNSUPPORTED, "gadget blocks of at most 12 limbs are supported");
        }
        j0 = j1;
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- 3. key MAC -------------------------------------------------------------------

// ---- 1b. bounded gadget decomposition ---------------------------------------------
//
// BoundedGadgetLayout::decompose + lift_digit (grafting/bounded_gadget.rs:
// 59-122, 382-402): the centered value of every coefficient over Q, written
// in balanced base-2^w digits d_j in [-2^(w-1), 2^(w-1)), each lifted to
// every limb of the key basis.  One thread per coefficient; the canonical
// value X in [0, Q) is built in W 64-bit words held in registers:
//   v_i = (x_i - X mod q_i) * (q_0 .. q_{i-1})^-1 mod q_i,
//   X  += v_i * (q_0 .. q_{i-1})                        (Garner, limb by limb)
// with X mod q_i = sum_k X[k] * (2^(64k) mod q_i).  X > floor(Q/2) is
// replaced by X - Q in two's complement, and the digits are read off w bits
// at a time with a carry (r + c >= 2^(w-1) -> digit r + c - 2^w, carry 1),
// which is the reference's repeated (value - digit) / 2^w.

using BoundedTable = CrtTable;

struct BoundedArgs {
    const u64 *in;  // component: limb i of ct c at in + c*in_stride + i*N
    u64 in_stride;
    u64 kinv;       // nonzero: decompose sigma_k(in) (rotation fused in)
    u64 *digits;    // [ct][D][B][N] lifted digits, or
    int64_t *raw;   // [ct][D][N] the balanced digits themselves (when non-null)
    unsigned char *ties;  // [ct][N] (when non-null): 1 where a digit is -2^(w-1)
    BoundedTable t;
    const DevPrime *primes;  // key basis B
    u32 q_limbs, limbs, logn, digit_count, base_log;
    u32 small;  // 1: 2^(base_log-1) < every key prime (see lift_digit)
    u64 total;  // n_cts * N
};

// A balanced digit (|d| <= 2^(w-1)) lifted to [0, q) of key prime t: one
// conditional add when 2^(w-1) < q (small, wave-uniform), a Barrett reduction
// otherwise.  The prime table is read through the scalar unit.
__device__ __forceinline__ u64 lift_digit(int64_t d, const CK_CONST_AS DevPrime *pr, u32 t, u32 small) {
    if (small) {
        const u64 q = pr[t].q;
        return d >= 0 ? (u64)d : q - (u64)(-d);
    }
    const Barrett br = barrett_at(pr, t);
    if (d >= 0) return barrett_reduce128(0, (u64)d, br);
    const u64 m = barrett_reduce128(0, (u64)(-(d + 1)) + 1, br);
    return m ? br.q - m : 0;
}

template <int W>
__global__ __launch_bounds__(kKsThreads) void k_bounded_decompose(BoundedArgs a) {
    const u64 gid = (u64)blockIdx.x * kKsThreads + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u64 ct = gid >> a.logn, n = gid & (N - 1);
    const Gather gi = galois_gather(n, a.kinv, a.logn);
    const u64 *x = a.in + ct * a.in_stride + gi.idx;
    u64 X[W];
    const bool negative = crt_centered<W>(x, N, gi.neg, a.q_limbs, a.primes, a.t, X);
    // balanced digits, w <= 63 bits at a time, from a 128-bit bit buffer
    const u32 w = a.base_log;
    const u64 mask = ((u64)1 << w) - 1, half_base = (u64)1 << (w - 1);
    const u64 fill = negative ? ~(u64)0 : 0;
    u64 *out = a.digits + ct * (u64)a.digit_count * a.limbs * N + n;
    u64 lo = 0, hi = 0, carry = 0;
    u32 nb = 0, j = 0;
    bool tie = false;
#pragma unroll
    for (int k = 0; k < W + 2; ++k) {
        const u64 word = k < W ? X[k] : fill;
        // nb < w <= 63 bits are buffered; append 64 more
        lo |= nb ? (word << nb) : word;
        hi = nb ? (word >> (64 - nb)) : 0;
        nb += 64;
        while (nb >= w && j < a.digit_count) {
            const u64 r = lo & mask;
            lo = (lo >> w) | (hi << (64 - w));
            hi >>= w;
            nb -= w;
            const u64 t = r + carry;
            int64_t d;
            if (t >= half_base) {
                d = (int64_t)t - (int64_t)mask - 1;
                carry = 1;
            } else {
                d = (int64_t)t;
                carry = 0;
            }
            tie |= d == -(int64_t)half_base;
            if (a.raw) {
                a.raw[(ct * a.digit_count + j) * N + n] = d;
            } else {
                u64 *o = out + (u64)j * a.limbs * N;
                const CK_CONST_AS DevPrime *pr = const_table(a.primes);
                for (u32 t2 = 0; t2 < a.limbs; ++t2) o[(u64)t2 * N] = lift_digit(d, pr, t2, a.small);
            }
            ++j;
        }
    }
    if (a.ties) a.ties[ct * N + n] = tie ? 1 : 0;
}

// Hoisted digit automorphism (BoundedGadgetDecomposition::automorphism_hoisted,
// grafting/bounded_gadget.rs:273-312) as a gather: output coefficient n takes
// the digits of source g = n k^-1 mod 2N (g - N, negated, when g >= N).  A
// negated digit vector is renormalised into [-2^(w-1), 2^(w-1)) from the
// lowest digit up with a carry, which is what makes the result equal to the
// balanced decomposition of sigma_k(x); the digit count leaves one bit of
// headroom above the centered value, so the final carry is zero.  Writes the
// digits themselves (raw) or lifts them to every limb (lifted).
struct HoistArgs {
    const int64_t *digits;  // [ct][D][N] balanced digits of the unrotated component
    int64_t *raw;           // [ct][D][N] output digits, or
    u64 *lifted;            // [ct][D][B][N] output digits lifted to every limb
    const DevPrime *primes;
    u64 kinv;  // k^-1 mod 2N
    u32 limbs, logn, digit_count, base_log;
    u32 small;  // as BoundedArgs::small
    u64 total;  // n_cts * N
};

__global__ __launch_bounds__(kKsThreads) void k_bounded_hoist(HoistArgs a) {
    const u64 gid = (u64)blockIdx.x * kKsThreads + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u64 ct = gid >> a.logn, n = gid & (N - 1);
    const Gather gi = galois_gather(n, a.kinv, a.logn);
    const int64_t *src = a.digits + ct * a.digit_count * N + gi.idx;
    const u32 w = a.base_log;
    const __int128 half = (__int128)1 << (w - 1);
    const __int128 mask = ((__int128)1 << w) - 1;
    int64_t carry = 0;
    for (u32 j = 0; j < a.digit_count; ++j) {
        int64_t d = src[(u64)j * N];
        if (gi.neg) {
            // value in [-2^(w-1) - 1, 2^(w-1) + 1]; balanced remainder and carry
            const __int128 x = (__int128)carry - d + half;
            d = (int64_t)((x & mask) - half);
            carry = (int64_t)(x >> w);
        }
        if (a.raw) {
            a.raw[(ct * a.digit_count + j) * N + n] = d;
        } else {
            u64 *o = a.lifted + ((ct * a.digit_count + j) * a.limbs) * N + n;
            const CK_CONST_AS DevPrime *pr = const_table(a.primes);
            for (u32 t = 0; t < a.limbs; ++t) o[(u64)t * N] = lift_digit(d, pr, t, a.small);
        }
    }
}

static int launch_hoist(hipStream_t s, const HoistArgs &a) {
    if (a.total == 0) return CKKS_OK;
    hipLaunchKernelGGL(k_bounded_hoist, dim3((u32)((a.total + kKsThreads - 1) / kKsThreads)), dim3(kKsThreads), 0,
                       s, a);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

struct MacArgs {
    const u64 *digits;  // [ct][dnum][B][N] NTT domain
    const u64 *kb, *ka;  // [dnum][B][N]
    u64 *acc;            // [ct][2][B][N]
    const DevPrime *primes;
    u32 n_cts, dnum, limbs, logn, chunks;
    // hoisted canonical rotations: the key-side correction [B][N] of each
    // component, added to every ciphertext's sums (null otherwise)
    const u64 *tb, *ta;
    // hoisted with b: P NTT(b) [ct][q_limbs][N], added to component 0's sums
    // on the Q limbs (null otherwise; fixed-D kernels only)
    const u64 *bhat;
    u32 q_limbs;
};

// NTT-domain Galois permutation (ring/ntt_galois_permutation.rs:5-19) over
// whole limbs: out[n] = in[((k (2n + 1) mod 2N) - 1) / 2], so that
// NTT(sigma_k(p)) = permute(NTT(p), k); permuting by k^-1 inverts it.
struct NttPermArgs {
    const u64 *in;
    u64 *out;
    u64 k;
    u64 total;  // limbs * N
    u32 logn;
};

__global__ __launch_bounds__(kKsThreads) void k_ntt_permute(NttPermArgs a) {
    const u64 gid = (u64)blockIdx.x * kKsThreads + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u64 n = gid & (N - 1);
    a.out[gid] = a.in[(gid - n) + (((a.k * (2 * n + 1)) & (2 * N - 1)) >> 1)];
}

static int launch_ntt_permute(hipStream_t s, const u64 *in, u64 *out, u64 k, u64 total, u32 logn) {
    if (total == 0) return CKKS_OK;
    NttPermArgs a{in, out, k, total, logn};
    hipLaunchKernelGGL(k_ntt_permute, dim3((u32)((total + kKsThreads - 1) / kKsThreads)), dim3(kKsThreads), 0, s, a);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// Each thread owns one residue position (i, n) and keeps the key entries of
// every digit in registers while it walks the batch, so a key is read from
// HBM once per launch, not once per ciphertext.
//
// FP64 limbs (q < 2^41): each product is fp_mulmod(digit, key), |term| <=
// 0.51 q, so D <= 8 terms and a canonical correction sum below 2^44 (exact)
// and fp_residue makes the sum canonical; the key entries enter as doubles
// with fl(w / q) divided once per thread.
template <int D>
__device__ __forceinline__ void ks_mac_fp(const MacArgs &a, const DevPrime &P, u32 i, u64 N, u64 n) {
    double kb[D], kbp[D], ka[D], kap[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        kb[j] = small_to_f64(a.kb[((u64)j * a.limbs + i) * N + n]);
        ka[j] = small_to_f64(a.ka[((u64)j * a.limbs + i) * N + n]);
        kbp[j] = kb[j] / P.qd;
        kap[j] = ka[j] / P.qd;
    }
    // the correction enters as the sums' starting value (canonical, < 2^41)
    const double t0b = a.tb ? small_to_f64(a.tb[(u64)i * N + n]) : 0.0;
    const double t0a = a.ta ? small_to_f64(a.ta[(u64)i * N + n]) : 0.0;
    const u64 *bh = a.bhat && i < a.q_limbs ? a.bhat + (u64)i * N + n : nullptr;
    for (u32 ct = 0; ct < a.n_cts; ++ct) {
        const u64 *d = a.digits + (((u64)ct * D) * a.limbs + i) * N + n;
        double sb = bh ? t0b + small_to_f64(ld_stream(bh + (u64)ct * a.q_limbs * N)) : t0b, sa = t0a;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const double dv = small_to_f64(ld_stream(d + (u64)j * a.limbs * N));
            sb += fp_mulmod(dv, kb[j], kbp[j], P.qd);
            sa += fp_mulmod(dv, ka[j], kap[j], P.qd);
        }
        u64 *o = a.acc + ((u64)ct * 2 * a.limbs + i) * N + n;
        st_stream(o, fp_residue(sb, P.qd, P.qinv));
        st_stream(o + (u64)a.limbs * N, fp_residue(sa, P.qd, P.qinv));
    }
}

template <int D>
__global__ __launch_bounds__(kKsThreads) void k_ks_mac(MacArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 i = b / a.chunks;
    const u64 N = (u64)1 << a.logn;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    const DevPrime P = a.primes[i];
    if (P.fp) {
        ks_mac_fp<D>(a, P, i, N, n);
        return;
    }
    u64 kb[D], ka[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
        kb[j] = a.kb[((u64)j * a.limbs + i) * N + n];
        ka[j] = a.ka[((u64)j * a.limbs + i) * N + n];
    }
    const u64 tb = a.tb ? a.tb[(u64)i * N + n] : 0, ta = a.ta ? a.ta[(u64)i * N + n] : 0;
    const u64 *bh = a.bhat && i < a.q_limbs ? a.bhat + (u64)i * N + n : nullptr;
    for (u32 ct = 0; ct < a.n_cts; ++ct) {
        const u64 *d = a.digits + (((u64)ct * D) * a.limbs + i) * N + n;
        Acc128 sb, sa;
        sb.zero();
        sa.zero();
        if (bh) sb.mac(ld_stream(bh + (u64)ct * a.q_limbs * N), 1);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            u64 dv = ld_stream(d + (u64)j * a.limbs * N);
            sb.mac(dv, kb[j]);
            sa.mac(dv, ka[j]);
        }
        sb.mac(tb, 1);
        sa.mac(ta, 1);
        u64 *o = a.acc + ((u64)ct * 2 * a.limbs + i) * N + n;
        st_stream(o, reduce128(sb, P));
        st_stream(o + (u64)a.limbs * N, reduce128(sa, P));
    }
}

// Hoisted rotations of one batch by up to kMacMaxExp exponents in one MAC:
// every exponent's key entries (permuted by its k^-1) and correction stay
// in registers while the thread walks the batch, and each digit (and the
// hoisted P NTT(b)) is read once for all of them instead of once per
// exponent.  Same sums, products and reductions as k_ks_mac per exponent.
constexpr int kMacMaxExp = 4;

struct MacMultiArgs {
    const u64 *digits;  // [ct][D][B][N] NTT domain
    const u64 *kb[kMacMaxExp], *ka[kMacMaxExp];  // per exponent [D][B][N]
    const u64 *tb[kMacMaxExp], *ta[kMacMaxExp];  // per exponent [B][N], or null
    u64 *acc[kMacMaxExp];                        // per exponent [ct][2][B][N]
    const u64 *bhat;  // P NTT(b) [ct][q_limbs][N], or null
    const DevPrime *primes;
    u32 n_cts, limbs, q_limbs, logn, chunks;
};

template <int D, int E>
__global__ __launch_bounds__(kKsThreads) void k_ks_mac_multi(MacMultiArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 i = b / a.chunks;
    const u64 N = (u64)1 << a.logn;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    const DevPrime P = a.primes[i];
    const u64 *bh = a.bhat && i < a.q_limbs ? a.bhat + (u64)i * N + n : nullptr;
    const u64 kstride = (u64)a.limbs * N;
    if (P.fp) {
        double kb[E][D], kbp[E][D], ka[E][D], kap[E][D], t0b[E], t0a[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
#pragma unroll
            for (int j = 0; j < D; ++j) {
                kb[e][j] = small_to_f64(a.kb[e][(u64)j * kstride + (u64)i * N + n]);
                ka[e][j] = small_to_f64(a.ka[e][(u64)j * kstride + (u64)i * N + n]);
                kbp[e][j] = kb[e][j] / P.qd;
                kap[e][j] = ka[e][j] / P.qd;
            }
            t0b[e] = a.tb[e] ? small_to_f64(a.tb[e][(u64)i * N + n]) : 0.0;
            t0a[e] = a.ta[e] ? small_to_f64(a.ta[e][(u64)i * N + n]) : 0.0;
        }
        for (u32 ct = 0; ct < a.n_cts; ++ct) {
            const u64 *d = a.digits + (((u64)ct * D) * a.limbs + i) * N + n;
            double dv[D];
#pragma unroll
            for (int j = 0; j < D; ++j) dv[j] = small_to_f64(ld_stream(d + (u64)j * kstride));
            const double bv = bh ? small_to_f64(ld_stream(bh + (u64)ct * a.q_limbs * N)) : 0.0;
            const u64 o = ((u64)ct * 2 * a.limbs + i) * N + n;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                double sb = t0b[e] + bv, sa = t0a[e];
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    sb += fp_mulmod(dv[j], kb[e][j], kbp[e][j], P.qd);
                    sa += fp_mulmod(dv[j], ka[e][j], kap[e][j], P.qd);
                }
                st_stream(a.acc[e] + o, fp_residue(sb, P.qd, P.qinv));
                st_stream(a.acc[e] + o + kstride, fp_residue(sa, P.qd, P.qinv));
            }
        }
        return;
    }
    u64 kb[E][D], ka[E][D], tb[E], ta[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            kb[e][j] = a.kb[e][(u64)j * kstride + (u64)i * N + n];
            ka[e][j] = a.ka[e][(u64)j * kstride + (u64)i * N + n];
        }
        tb[e] = a.tb[e] ? a.tb[e][(u64)i * N + n] : 0;
        ta[e] = a.ta[e] ? a.ta[e][(u64)i * N + n] : 0;
    }
    for (u32 ct = 0; ct < a.n_cts; ++ct) {
        const u64 *d = a.digits + (((u64)ct * D) * a.limbs + i) * N + n;
        u64 dv[D];
#pragma unroll
        for (int j = 0; j < D; ++j) dv[j] = ld_stream(d + (u64)j * kstride);
        const u64 bv = bh ? ld_stream(bh + (u64)ct * a.q_limbs * N) : 0;
        const u64 o = ((u64)ct * 2 * a.limbs + i) * N + n;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            Acc128 sb, sa;
            sb.zero();
            sa.zero();
#pragma unroll
            for (int j = 0; j < D; ++j) {
                sb.mac(dv[j], kb[e][j]);
                sa.mac(dv[j], ka[e][j]);
            }
            sb.mac(tb[e], 1);
            sa.mac(ta[e], 1);
            sb.mac(bv, 1);
            st_stream(a.acc[e] + o, reduce128(sb, P));
            st_stream(a.acc[e] + o + kstride, reduce128(sa, P));
        }
    }
}

// More than 8 digits over a batch (the bounded gadget: 23 digits at the
// reference's N=2^16 shape): chunks of C digits, one launch each, with the
// chunk's key entries in registers while the thread walks the batch, so a
// key is read once per launch; the first chunk writes the accumulators and
// later ones add to them.
template <int C>
__global__ __launch_bounds__(kKsThreads) void k_ks_mac_chunk(MacArgs a, u32 j0, u32 first) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 i = b / a.chunks;
    const u64 N = (u64)1 << a.logn;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    const DevPrime P = a.primes[i];
    const u32 cnt = a.dnum - j0 < (u32)C ? a.dnum - j0 : (u32)C;
    u64 kb[C], ka[C];
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const bool on = (u32)j < cnt;
        kb[j] = on ? a.kb[((u64)(j0 + j) * a.limbs + i) * N + n] : 0;
        ka[j] = on ? a.ka[((u64)(j0 + j) * a.limbs + i) * N + n] : 0;
    }
    // the next ciphertext's digits are loaded before this one's products,
    // so a thread keeps 2C loads in flight
    const u64 d_ct = (u64)a.dnum * a.limbs * N, d_j = (u64)a.limbs * N;
    const u64 *d = a.digits + ((u64)j0 * a.limbs + i) * N + n;
    u64 dv[C];
#pragma unroll
    for (int j = 0; j < C; ++j) dv[j] = (u32)j < cnt ? d[j * d_j] : 0;
    const bool corr = first && a.tb;
    const u64 tbv = corr ? a.tb[(u64)i * N + n] : 0, tav = corr ? a.ta[(u64)i * N + n] : 0;
    for (u32 ct = 0; ct < a.n_cts; ++ct) {
        u64 nx[C];
        const bool more = ct + 1 < a.n_cts;
        d += d_ct;
#pragma unroll
        for (int j = 0; j < C; ++j) nx[j] = more && (u32)j < cnt ? d[j * d_j] : 0;
        Acc128 sb, sa;
        sb.zero();
        sa.zero();
#pragma unroll
        for (int j = 0; j < C; ++j) {
            sb.mac(dv[j], kb[j]);
            sa.mac(dv[j], ka[j]);
        }
#pragma unroll
        for (int j = 0; j < C; ++j) dv[j] = nx[j];
        u64 *o = a.acc + ((u64)ct * 2 * a.limbs + i) * N + n;
        if (corr) {
            sb.mac(tbv, 1);
            sa.mac(tav, 1);
        }
        u64 rb = reduce128(sb, P), ra = reduce128(sa, P);
        if (!first) {
            rb = add_mod(rb, o[0], P.q);
            ra = add_mod(ra, o[(u64)a.limbs * N], P.q);
        }
        o[0] = rb;
        o[(u64)a.limbs * N] = ra;
    }
}

// Generic digit count (the RNS gadget with singleton blocks has dnum = L):
// key entries are streamed per ciphertext; partial sums are reduced every 8
// products so the 128-bit accumulator cannot overflow.
__global__ __launch_bounds__(kKsThreads) void k_ks_mac_any(MacArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 rest = b / a.chunks;
    const u32 i = rest % a.limbs, ct = rest / a.limbs;
    const u64 N = (u64)1 << a.logn;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    const DevPrime P = a.primes[i];
    const u64 *d = a.digits + (((u64)ct * a.dnum) * a.limbs + i) * N + n;
    u64 rb = a.tb ? a.tb[(u64)i * N + n] : 0, ra = a.ta ? a.ta[(u64)i * N + n] : 0;
    for (u32 j0 = 0; j0 < a.dnum; j0 += 8) {
        Acc128 sb, sa;
        sb.zero();
        sa.zero();
        u32 j1 = j0 + 8 < a.dnum ? j0 + 8 : a.dnum;
        for (u32 j = j0; j < j1; ++j) {
            u64 dv = d[(u64)j * a.limbs * N];
            u64 off = ((u64)j * a.limbs + i) * N + n;
            sb.mac(dv, a.kb[off]);
            sa.mac(dv, a.ka[off]);
        }
        rb = add_mod(rb, reduce128(sb, P), P.q);
        ra = add_mod(ra, reduce128(sa, P), P.q);
    }
    u64 *o = a.acc + ((u64)ct * 2 * a.limbs + i) * N + n;
    o[0] = rb;
    o[(u64)a.limbs * N] = ra;
}

// ---- 5. ModDown / finalize ---------------------------------------------------------

struct DownArgs {
    const u64 *acc;  // [ct][2][B][N] coefficient domain
    const u64 *add0, *add1;  // pass-through for component 0 / 1 (may be null)
    u64 add_stride;
    u64 kinv;  // nonzero: component 0 pass-through is sigma_k(add0)
    u64 *out;  // [ct][2][q_limbs][N] with out_stride
    u64 out_stride;
    const ModDownTable *md;  // null: no special primes
    const DevPrime *primes;  // key basis
    u64 fp_targets;  // bit i: Q limb i takes the FP64 target path
    u32 limbs, q_limbs, logn, chunks;
    u32 split;  // every special prime below 2^60: FP64 targets by fp_split_target
};

template <int K>
__global__ __launch_bounds__(kKsThreads) void k_moddown(DownArgs a) {
    const u32 b = xcd_contiguous_block(blockIdx.x, gridDim.x);
    const u32 chunk = b % a.chunks;
    const u32 rest = b / a.chunks;
    const u32 comp = rest % 2, ct = rest / 2;
    const u64 N = (u64)1 << a.logn;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    const CK_CONST_AS DevPrime *primes = const_table(a.primes);
    const u64 *x = a.acc + ((u64)ct * 2 + comp) * a.limbs * N + n;
    const u64 *add = comp == 0 ? a.add0 : a.add1;
    const Gather ga = galois_gather(n, comp == 0 ? a.kinv : 0, a.logn);
    if (add) add += ct * a.add_stride + ga.idx;
    u64 *o = a.out + ct * a.out_stride + (u64)comp * a.q_limbs * N + n;
    if (K == 0) {
        for (u32 i = 0; i < a.q_limbs; ++i) {
            u64 v = x[(u64)i * N];
            if (add) {
                const u64 q = primes[i].q;
                u64 ad = add[(u64)i * N];
                v = ga.neg ? sub_mod(v, ad, q) : add_mod(v, ad, q);
            }
            o[(u64)i * N] = v;
        }
        return;
    }
    // tables through the scalar unit (const_table): uniform addresses only
    const CK_CONST_AS ModDownTable *md = const_table(a.md);
    u64 c[K > 0 ? K : 1];
    bool neg = false;
    if (K > 0) {
        u64 xp[K > 0 ? K : 1];
#pragma unroll
        for (int m = 0; m < K; ++m) xp[m] = ld_stream(x + (u64)(a.q_limbs + m) * N);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k == 0) {
                c[0] = xp[0];
            } else {
                const Barrett bk = barrett_at(primes, a.q_limbs + k);
                Acc128 t;
                t.zero();
#pragma unroll
                for (int m = 0; m < k; ++m) t.mac(c[m], md->pm[k][m]);
                c[k] = shoup(sub_mod(xp[k], t.reduce(bk), bk.q), md->inv[k],
                             md->inv_q[k], bk.q);
            }
        }
        // x_P > floor(P/2)  <=>  mixed-radix digits compare greater from the top
        int cmp = 0;
#pragma unroll
        for (int k = K - 1; k >= 0; --k)
            if (cmp == 0 && c[k] != md->half[k]) cmp = c[k] > md->half[k] ? 1 : -1;
        neg = cmp > 0;
    }
    // FP64 targets (bit i of a.fp_targets, q_i < 2^41): every P digit enters
    // split at bit 30; the 2K terms sum to |acc| <= 2K 0.51 q_i, so
    // t = x_i - acc (+ P) stays below 2^45 for the final fp_mulmod by P^-1.
    double cl[K > 0 ? K : 1], ch[K > 0 ? K : 1];
#pragma unroll
    for (int m = 0; m < K; ++m) {
        cl[m] = small_to_f64(c[m] & 0x3fffffffull);
        ch[m] = small_to_f64(c[m] >> 30);
    }
    const CK_CONST_AS double *tf = const_table(md->tf);
    const CK_CONST_AS u64 *wt = const_table(md->w);
    const CK_CONST_AS u64 *pmod = const_table(md->pmod);
    const CK_CONST_AS u64 *pinv = const_table(md->pinv);
    const CK_CONST_AS u64 *pinv_q = const_table(md->pinv_q);
    const u64 fpm = a.fp_targets;
    for (u32 i = 0; i < a.q_limbs; ++i) {
        if (i < 64 && ((fpm >> i) & 1)) {  // the mask is empty for bases over 64 limbs
            const CK_CONST_AS double *r = tf + (u64)i * kMdRow;
            const double q = r[0];
            double acc;
            if (a.split) {  // every P digit below 2^60
                const CK_CONST_AS double *sk = r + kMdSplitK;
                acc = fp_split_target<K>(cl, ch, r + kMdSplitW, q, r[1], sk[0], sk[1], sk[2], sk[3], sk[4], sk[5]);
            } else {
                acc = fp_mulmod(cl[0], r[6], r[7], q) + fp_mulmod(ch[0], r[8], r[9], q);
#pragma unroll
                for (int m = 1; m < K; ++m)
                    acc += fp_mulmod(cl[m], r[6 + 4 * m], r[7 + 4 * m], q) +
                           fp_mulmod(ch[m], r[8 + 4 * m], r[9 + 4 * m], q);
            }
            double t = small_to_f64(ld_stream(x + (u64)i * N)) - acc;  // x_i - (x_P mod q_i)
            if (neg) t += r[2];                             // x_P - P
            double v = fp_mulmod(t, r[3], r[4], q);
            if (add) {
                const double ad = small_to_f64(add[(u64)i * N]);
                v = ga.neg ? v - ad : v + ad;
            }
            st_stream(o + (u64)i * N, fp_residue(v, q, r[1]));
            continue;
        }
        const Barrett bi = barrett_at(primes, i);
        const CK_CONST_AS u64 *w = wt + (u64)i * kMaxSpecial;
        Acc128 t;
        t.zero();
#pragma unroll
        for (int m = 0; m < K; ++m) t.mac(c[m], w[m]);
        u64 r = t.reduce(bi);  // x_P mod q_i
        if (neg) r = sub_mod(r, pmod[i], bi.q);  // (x_P - P) mod q_i
        u64 v = shoup(sub_mod(ld_stream(x + (u64)i * N), r, bi.q), pinv[i], pinv_q[i], bi.q);
        if (add) {
            u64 ad = add[(u64)i * N];
            v = ga.neg ? sub_mod(v, ad, bi.q) : add_mod(v, ad, bi.q);
        }
        st_stream(o + (u64)i * N, v);
    }
}

// ---- exact base conversion on the int8 matrix cores ----------------------------------
//
// ModUp and ModDown as one GEMM each (see baseconv.hpp): the VALU computes
// the mixed-radix digits of every coefficient and writes their 7-bit chunks
// (9 per digit, row k = 9m + a) to LDS; v_mfma_i32_32x32x32_i8 multiplies
// them by the constant chunk matrix; the epilogue recombines
// sum_b E_b 2^(7b) and reduces it once per target.  256 coefficients per
// workgroup, each wave owns 64 of them (two 32-row M-tiles).

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr int bc_krows(int A) { return ((9 * A + 31) / 32) * 32; }

// Chunk row of one coefficient: byte k = 9m + a is bits [7a, 7a+7) of c_m.
template <int A>
__device__ __forceinline__ void bc_store_row(const u64 (&c)[A], signed char *row) {
    constexpr int KR = bc_krows(A);
#pragma unroll
    for (int w = 0; w < KR / 4; ++w) {
        u32 word = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * w + i;
            if (k < 9 * A) {
                const int m = k / 9, sh = 7 * (k % 9);
                word |= (u32)((c[m] >> sh) & 127u) << (8 * i);
            }
        }
        reinterpret_cast<u32 *>(row)[w] = word;
    }
}

// One column tile for the wave's 64 rows: E tile -> per-wave LDS staging.
template <int KR>
__device__ __forceinline__ void bc_gemm_tile(const signed char *lds_a, const signed char *bt, u32 tile,
                                             u32 wave, u32 lane, int *lds_e) {
    v16i acc0 = {}, acc1 = {};
    const u32 r = lane & 31, kh = (lane >> 5) * 16;
#pragma unroll
    for (int ks = 0; ks < KR / 32; ++ks) {
        const u32 k0 = (u32)ks * 32 + kh;
        v4i a0 = *reinterpret_cast<const v4i *>(lds_a + (wave * 64 + r) * KR + k0);
        v4i a1 = *reinterpret_cast<const v4i *>(lds_a + (wave * 64 + 32 + r) * KR + k0);
        v4i b = *reinterpret_cast<const v4i *>(bt + (tile * kBcTileCols + r) * KR + k0);
        acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a0, b, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(a1, b, acc1, 0, 0, 0);
    }
    // accumulator i of lane l holds row (i%4) + 8(i/4) + 4(l/32), column l%32
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const u32 row = (i & 3) + 8 * (i >> 2) + 4 * (lane >> 5);
        lds_e[row * 33 + r] = acc0[i];
        lds_e[(row + 32) * 33 + r] = acc1[i];
    }
}

// sum_b E_b 2^(7b) for the chunk columns [col, col+nch) of one row, as a
// 128-bit (hi, lo) with hi < 2^15.
__device__ __forceinline__ void bc_combine(const int *e_row, u32 col, u32 nch, u64 &hi, u64 &lo) {
    u64 g0 = 0, g1 = 0;
    for (u32 b = 0; b < nch; ++b) {
        u64 e = (u64)(u32)e_row[col + b];
        if (b < 6) g0 += e << (7 * b);
        else g1 += e << (7 * (b - 6));
    }
    const u64 add_lo = g1 << 42;
    lo = g0 + add_lo;
    hi = (g1 >> 22) + (lo < add_lo ? 1u : 0u);
}

struct ModUpMfmaArgs {
    ModUpArgs base;
    const BcPlan *plans;  // one per gadget block
};

template <int A>
__device__ __forceinline__ void modup_mfma_body(const ModUpMfmaArgs &args, signed char *lds_a,
                                                int (*lds_e)[64 * 33]) {
    constexpr int KR = bc_krows(A);
    const ModUpArgs &a = args.base;
    const u32 N = 1u << a.logn;
    const u32 chunks = N / kBcRows;
    const u32 b = blockIdx.x;
    const u32 chunk = b % chunks, cj = b / chunks;
    const u32 j = cj % a.dnum, ct = cj / a.dnum;
    const ModUpBlock &blk = a.blocks[j];
    const BcPlan &bp = args.plans[j];
    const u32 tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const u64 n = (u64)chunk * kBcRows + tid;
    u64 *dst = a.digits + (((u64)ct * a.dnum + j) * a.limbs) * N;
    // ---- digits (VALU): gather sigma_k, Garner, in-block limbs, chunk row
    {
        const Gather gi = galois_gather(n, a.kinv, a.logn);
        const u64 *src = a.in + ct * a.in_stride + (u64)blk.start * N + gi.idx;
        u64 x[A], c[A];
#pragma unroll
        for (int m = 0; m < A; ++m) {
            u64 v = src[(u64)m * N];
            x[m] = gi.neg ? neg_mod(v, a.primes[blk.start + m].q) : v;
            dst[(u64)(blk.start + m) * N + n] = x[m];  // digit_j mod q_i = x_i inside the block
        }
#pragma unroll
        for (int k = 0; k < A; ++k) {
            const DevPrime &Pk = a.primes[blk.start + k];
            if (k == 0) {
                c[0] = x[0];
            } else {
                Acc128 t;
                t.zero();
#pragma unroll
                for (int m = 0; m < k; ++m) t.mac(c[m], blk.pm[k][m]);
                c[k] = shoup(sub_mod(x[k], reduce128(t, Pk), Pk.q), blk.inv[k], blk.inv_q[k], Pk.q);
            }
        }
        bc_store_row<A>(c, lds_a + tid * KR);
    }
    __syncthreads();
    // ---- GEMM per column tile, then the epilogue for its targets
    int *e = lds_e[wave];
    const u64 row_n = (u64)chunk * kBcRows + wave * 64 + lane;
    for (u32 tile = 0; tile < bp.n_tiles; ++tile) {
        bc_gemm_tile<KR>(lds_a, bp.bt, tile, wave, lane, e);
        __syncthreads();
        const BcTile &T = bp.tiles[tile];
        for (u32 t = 0; t < T.n_targets; ++t) {
            const BcTarget tg = T.t[t];
            const DevPrime &Pt = a.primes[tg.limb];
            u64 hi, lo;
            bc_combine(e + lane * 33, tg.col, tg.nch, hi, lo);
            dst[(u64)tg.limb * N + row_n] = barrett_reduce128(hi, lo, Barrett{Pt.q, Pt.br_hi, Pt.br_lo});
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_modup_mfma(ModUpMfmaArgs args) {
    __shared__ __align__(16) signed char lds_a[kBcRows * kBcMaxK];
    __shared__ int lds_e[4][64 * 33];
    const u32 N = 1u << args.base.logn;
    const u32 j = (blockIdx.x / (N / kBcRows)) % args.base.dnum;
    switch (args.base.blocks[j].size) {
#define CK_MM(A) \
    case A: modup_mfma_body<A>(args, lds_a, lds_e); break;
        CK_MM(2) CK_MM(3) CK_MM(4) CK_MM(5) CK_MM(6) CK_MM(7)
        CK_MM(8) CK_MM(9)
#undef CK_MM
        default: break;
    }
}

struct ModDownMfmaArgs {
    DownArgs base;
    const BcPlan *plan;  // P digits -> Q limbs
};

template <int K>
__global__ __launch_bounds__(256) void k_moddown_mfma(ModDownMfmaArgs args) {
    constexpr int KR = bc_krows(K);
    __shared__ __align__(16) signed char lds_a[kBcRows * KR];
    __shared__ int lds_e[4][64 * 33];
    __shared__ unsigned char lds_neg[kBcRows];
    const DownArgs &a = args.base;
    const BcPlan &bp = *args.plan;
    const ModDownTable &md = *a.md;
    const u32 N = 1u << a.logn;
    const u32 chunks = N / kBcRows;
    const u32 b = blockIdx.x;
    const u32 chunk = b % chunks, rest = b / chunks;
    const u32 comp = rest % 2, ct = rest / 2;
    const u32 tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const u64 *x = a.acc + ((u64)ct * 2 + comp) * a.limbs * N;
    // ---- P-side mixed-radix digits and the centering decision (VALU)
    {
        const u64 n = (u64)chunk * kBcRows + tid;
        u64 c[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const DevPrime &Pk = a.primes[a.q_limbs + k];
            const u64 xk = x[(u64)(a.q_limbs + k) * N + n];
            if (k == 0) {
                c[0] = xk;
            } else {
                Acc128 t;
                t.zero();
#pragma unroll
                for (int m = 0; m < k; ++m) t.mac(c[m], md.pm[k][m]);
                c[k] = shoup(sub_mod(xk, reduce128(t, Pk), Pk.q), md.inv[k], md.inv_q[k], Pk.q);
            }
        }
        int cmp = 0;
#pragma unroll
        for (int k = K - 1; k >= 0; --k)
            if (cmp == 0 && c[k] != md.half[k]) cmp = c[k] > md.half[k] ? 1 : -1;
        lds_neg[tid] = cmp > 0 ? 1 : 0;
        bc_store_row<K>(c, lds_a + tid * KR);
    }
    __syncthreads();
    const u64 *add = comp == 0 ? a.add0 : a.add1;
    const u64 row_n = (u64)chunk * kBcRows + wave * 64 + lane;
    const Gather ga = galois_gather(row_n, comp == 0 ? a.kinv : 0, a.logn);
    if (add) add += ct * a.add_stride + ga.idx;
    u64 *o = a.out + ct * a.out_stride + (u64)comp * a.q_limbs * N + row_n;
    const bool neg = lds_neg[wave * 64 + lane] != 0;
    int *e = lds_e[wave];
    for (u32 tile = 0; tile < bp.n_tiles; ++tile) {
        bc_gemm_tile<KR>(lds_a, bp.bt, tile, wave, lane, e);
        __syncthreads();
        const BcTile &T = bp.tiles[tile];
        for (u32 t = 0; t < T.n_targets; ++t) {
            const BcTarget tg = T.t[t];
            const u32 i = tg.limb;
            const DevPrime &Pi = a.primes[i];
            u64 hi, lo;
            bc_combine(e + lane * 33, tg.col, tg.nch, hi, lo);
            u64 r = barrett_reduce128(hi, lo, Barrett{Pi.q, Pi.br_hi, Pi.br_lo});  // x_P mod q_i
            if (neg) r = sub_mod(r, md.pmod[i], Pi.q);
            u64 v = shoup(sub_mod(x[(u64)i * N + row_n], r, Pi.q), md.pinv[i], md.pinv_q[i], Pi.q);
            if (add) {
                u64 ad = add[(u64)i * N];
                v = ga.neg ? sub_mod(v, ad, Pi.q) : add_mod(v, ad, Pi.q);
            }
            o[(u64)i * N] = v;
        }
        __syncthreads();
    }
}

// ---- host-side setup ----------------------------------------------------------------

// Constant side of an exact base conversion from `src` digits to `targets`
// (pairs of destination limb index and modulus): targets packed whole into
// 32-column tiles; B^T[column][k], k = 9m + a, column = tile*32 + col + b,
// entry = chunk_b((2^(7a) M_m) mod p_t), M_m = prod_{l<m} src_l.
static bool build_bc_plan(const std::vector<u64> &src, const std::vector<std::pair<u32, u64>> &targets,
                          BcPlan &plan, std::vector<signed char> &bt) {
    const int A = (int)src.size();
    if (A < 1 || A > 12) return false;
    plan = BcPlan{};
    plan.k_rows = (u32)bc_krows(A);
    u32 tile = 0, col = 0;
    for (auto &t : targets) {
        u32 nch = (u32)chunks7(t.second);
        if (col + nch > (u32)kBcTileCols || plan.tiles[tile].n_targets == (u32)kBcMaxPerTile) {
            ++tile;
            col = 0;
        }
        if (tile >= (u32)kBcMaxTiles) return false;
        BcTile &T = plan.tiles[tile];
        T.t[T.n_targets++] = BcTarget{t.first, col, nch, 0};
        col += nch;
    }
    plan.n_tiles = targets.empty() ? 0 : tile + 1;
    const u32 KR = plan.k_rows;
    bt.assign((size_t)plan.n_tiles * kBcTileCols * KR, 0);
    for (u32 ti = 0; ti < plan.n_tiles; ++ti) {
        const BcTile &T = plan.tiles[ti];
        for (u32 e = 0; e < T.n_targets; ++e) {
            u64 p = 0;
            for (auto &t : targets)
                if (t.first == T.t[e].limb) p = t.second;
            u64 Mm = 1 % p;  // M_m mod p
            for (int m = 0; m < A; ++m) {
                u64 v = Mm;  // (2^(7a) M_m) mod p for a = 0
                for (int a = 0; a < 9; ++a) {
                    for (u32 b = 0; b < T.t[e].nch; ++b)
                        bt[((size_t)ti * kBcTileCols + T.t[e].col + b) * KR + 9 * m + a] =
                            (signed char)((v >> (7 * b)) & 127u);
                    v = host::mulmod(v, 128 % p, p);
                }
                Mm = host::mulmod(Mm, src[m] % p, p);
            }
        }
    }
    return true;
}

static void fill_garner(const std::vector<u64> &mods, u64 (*pm)[kMaxAlpha], u64 *inv, u64 *inv_q,
                        int stride_dummy) {
    (void)stride_dummy;
    size_t A = mods.size();
    for (size_t k = 0; k < A; ++k) {
        u64 qk = mods[k];
        u64 prod = 1 % qk;
        for (size_t m = 0; m < k; ++m) {
            pm[k][m] = prod;  // (q_0 .. q_{m-1}) mod q_k
            prod = host::mulmod(prod, mods[m] % qk, qk);
        }
        if (k > 0) {
            inv[k] = host::invmod(prod, qk);
            inv_q[k] = host::shoup_quotient(inv[k], qk);
        }
    }
}

static u64 double_bits(double d) {
    u64 v;
    std::memcpy(&v, &d, 8);
    return v;
}

static bool fp_prime(u64 q) { return q < host::kFpModulusBound; }

// Appends the FP64 pair of a weight w < q: w and fl(w / q).
static void push_fp_pair(std::vector<u64> &consts, u64 w, u64 q) {
    consts.push_back(double_bits((double)w));
    consts.push_back(double_bits((double)w / (double)q));
}

// Bit i set when limb i (i < count <= 64) takes the FP64 target path.
static u64 fp_target_mask(const std::vector<u64> &mods, size_t count) {
    if (count > 64) return 0;
    u64 mask = 0;
    for (size_t i = 0; i < count; ++i)
        if (fp_prime(mods[i])) mask |= 1ull << i;
    return mask;
}

// ModDown constants of Q u P (the last mods.size() - q_limbs limbs are P):
// Garner over P into md; per Q limb the mixed-radix weights, P mod q_i,
// P^-1 mod q_i and its Shoup quotient appended to consts (the table's
// device pointers are set once consts is uploaded: md.w at the offset
// consts.size() had on entry).
// The ModDown digits are Garner digits below their special primes: with
// every special prime below 2^60 the FP64 targets take fp_split_target.
static u32 moddown_split(const std::vector<u64> &mods, size_t q_limbs) {
    for (size_t m = q_limbs; m < mods.size(); ++m)
        if (mods[m] >= (1ull << 60)) return 0;
    return 1;
}

static void build_moddown(const std::vector<u64> &mods, size_t q_limbs, std::vector<u64> &consts,
                          ModDownTable &md) {
    const size_t K = mods.size() - q_limbs;
    const size_t md_w_off = consts.size();
    md.q_limbs = (u32)q_limbs;
    md.p_limbs = (u32)K;
    std::vector<u64> pmods(mods.begin() + q_limbs, mods.end());
    u64 tmp_pm[kMaxAlpha][kMaxAlpha] = {};
    u64 tmp_inv[kMaxAlpha] = {}, tmp_inv_q[kMaxAlpha] = {};
    fill_garner(pmods, tmp_pm, tmp_inv, tmp_inv_q, 0);
    for (size_t a = 0; a < K; ++a) {
        for (size_t b = 0; b < K; ++b) md.pm[a][b] = tmp_pm[a][b];
        md.inv[a] = tmp_inv[a];
        md.inv_q[a] = tmp_inv_q[a];
    }
    // mixed-radix digits of floor(P/2): P odd, so floor(P/2) = (P-1)/2.
    // Digits of P-1 are (p_0 - 1, p_1 - 1, ...) ... computed by halving
    // the positional representation from the top digit down.
    {
        std::vector<u64> d(K);
        for (size_t m = 0; m < K; ++m) d[m] = pmods[m] - 1;  // P - 1 = sum (p_m - 1) M_m
        // divide by 2 in mixed radix: process from most significant digit
        unsigned __int128 rem = 0;
        for (size_t m = K; m-- > 0;) {
            unsigned __int128 cur = rem * pmods[m] + d[m];
            md.half[m] = (u64)(cur / 2);
            rem = cur % 2;
        }
    }
    for (size_t i = 0; i < q_limbs; ++i) {
        u64 prod = 1 % mods[i];
        for (int m = 0; m < kMaxSpecial; ++m) {
            consts.push_back(m < (int)K ? prod : 0);
            if (m < (int)K) prod = host::mulmod(prod, pmods[m] % mods[i], mods[i]);
        }
    }
    // P mod q_i, P^-1 mod q_i
    for (size_t i = 0; i < q_limbs; ++i) {
        u64 p = 1 % mods[i];
        for (size_t m = 0; m < K; ++m) p = host::mulmod(p, pmods[m] % mods[i], mods[i]);
        consts.push_back(p);
    }
    for (size_t i = 0; i < q_limbs; ++i) {
        u64 p = 1 % mods[i];
        for (size_t m = 0; m < K; ++m) p = host::mulmod(p, pmods[m] % mods[i], mods[i]);
        u64 inv = host::invmod(p, mods[i]);
        consts.push_back(inv);
    }
    for (size_t i = 0; i < q_limbs; ++i)
        consts.push_back(host::shoup_quotient(consts[md_w_off + q_limbs * kMaxSpecial + q_limbs + i],
                                              mods[i]));
    // FP64 target rows (ModDownTable::tf), zero rows for integer targets
    for (size_t i = 0; i < q_limbs; ++i) {
        const u64 q = mods[i];
        const size_t row0 = consts.size();
        if (fp_prime(q)) {
            const u64 pmod = consts[md_w_off + q_limbs * kMaxSpecial + i];
            const u64 pinv = consts[md_w_off + q_limbs * kMaxSpecial + q_limbs + i];
            consts.push_back(double_bits((double)q));
            consts.push_back(double_bits(1.0 / (double)q));
            consts.push_back(double_bits((double)pmod));
            push_fp_pair(consts, pinv, q);
            consts.push_back(0);
            for (size_t m = 0; m < K; ++m) {
                const u64 w = consts[md_w_off + i * kMaxSpecial + m];
                push_fp_pair(consts, w, q);
                push_fp_pair(consts, host::mulmod(w, (1ull << 30) % q, q), q);
            }
            consts.resize(row0 + kMdSplitW, 0);
            for (size_t m = 0; m < K; ++m) {
                const u64 w = consts[md_w_off + i * kMaxSpecial + m];
                consts.push_back(double_bits((double)(w & 0xfffffull)));
                consts.push_back(double_bits((double)(w >> 20)));
            }
            consts.resize(row0 + kMdSplitK, 0);
            for (int e : {20, 30, 50}) push_fp_pair(consts, (u64)(((unsigned __int128)1 << e) % q), q);
        }
        consts.resize(row0 + kMdRow, 0);
    }
}

// Device pointers of a ModDown table whose constants start at w (the
// layout build_moddown appends).
static void patch_moddown(ModDownTable &md, const u64 *w, size_t q_limbs) {
    md.w = w;
    md.pmod = md.w + q_limbs * kMaxSpecial;
    md.pinv = md.pmod + q_limbs;
    md.pinv_q = md.pinv + q_limbs;
    md.tf = (const double *)(md.pinv_q + q_limbs);
}

// ---- key material --------------------------------------------------------------------

// Flags a residue that is not below its limb's prime (key material handed in
// as device memory is checked where it lies).
__global__ __launch_bounds__(256) void k_check_canonical(const u64 *x, u64 elems, u32 logn, u32 limbs,
                                                         const DevPrime *primes, int *bad) {
    for (u64 i = (u64)blockIdx.x * 256 + threadIdx.x; i < elems; i += (u64)gridDim.x * 256)
        if (x[i] >= primes[(u32)((i >> logn) % limbs)].q) *bad = 1;
}

static bool is_device_pointer(const void *p) {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable host memory is not known to HIP
        return false;
    }
    return at.type == hipMemoryTypeDevice;
}

// Copies key entries [entries][limbs][N] (b and a) into new device buffers
// *db, *da after checking every residue is canonical.  The sources may be
// host arrays or device buffers (ckks_keygen writes device buffers); device
// sources are read on the context stream, after the work queued there.
static int upload_key_material(const ckks_rns_plan *plan, size_t kelems, const u64 *key_b, const u64 *key_a,
                               u64 **db, u64 **da) {
    ckks_context *ctx = plan->ctx;
    const size_t N = plan->degree, B = plan->moduli.size();
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "keys cannot be created during graph capture");
    const bool dev_b = is_device_pointer(key_b), dev_a = is_device_pointer(key_a);
    if (dev_b != dev_a) return fail(CKKS_ERR_INVALID, "key_b and key_a must both be host or both device memory");
    if (!dev_b) {
        const auto &mods = plan->moduli;
        for (size_t e = 0; e < kelems; ++e) {
            u64 q = mods[(e / N) % B];
            if (key_b[e] >= q || key_a[e] >= q) return fail(CKKS_ERR_INVALID, "key residues must be canonical");
        }
    }
    CK_HIP(hipMalloc(db, kelems * sizeof(u64)));
    CK_HIP(hipMalloc(da, kelems * sizeof(u64)));
    if (!dev_b) {
        CK_HIP(hipMemcpy(*db, key_b, kelems * sizeof(u64), hipMemcpyHostToDevice));
        CK_HIP(hipMemcpy(*da, key_a, kelems * sizeof(u64), hipMemcpyHostToDevice));
        return CKKS_OK;
    }
    void *w;
    int rc = ctx->workspace(2, sizeof(int), &w);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    CK_HIP(hipMemcpyAsync(*db, key_b, kelems * sizeof(u64), hipMemcpyDeviceToDevice, s));
    CK_HIP(hipMemcpyAsync(*da, key_a, kelems * sizeof(u64), hipMemcpyDeviceToDevice, s));
    CK_HIP(hipMemsetAsync(w, 0, sizeof(int), s));
    const u32 grid = (u32)std::min<u64>((kelems + 255) / 256, 4096);
    for (const u64 *x : {(const u64 *)*db, (const u64 *)*da})
        hipLaunchKernelGGL(k_check_canonical, dim3(grid), dim3(256), 0, s, x, (u64)kelems, plan->logn, (u32)B,
                           plan->d_primes, (int *)w);
    CK_HIP(hipGetLastError());
    int bad = 0;
    CK_HIP(hipMemcpyAsync(&bad, w, sizeof(int), hipMemcpyDeviceToHost, s));
    CK_HIP(hipStreamSynchronize(s));
    if (bad) return fail(CKKS_ERR_INVALID, "key residues must be canonical");
    return CKKS_OK;
}

static int create_key(const ckks_rns_plan *plan, size_t q_limbs, const size_t *block_sizes,
                      size_t n_blocks, const u64 *key_b, const u64 *key_a, bool hybrid,
                      ckks_ksk **out) {
    if (!plan || !block_sizes || !out) return fail(CKKS_ERR_INVALID, "null argument");
    if (n_blocks == 0) return fail(CKKS_ERR_INVALID, "RNS gadget layout must contain at least one block");
    const size_t B = plan->moduli.size();
    if (q_limbs == 0 || q_limbs > B) return fail(CKKS_ERR_INVALID, "Q limb count must be within the key basis");
    const size_t K = B - q_limbs;
    if (!hybrid && K != 0) return fail(CKKS_ERR_INVALID, "RNS gadget keys have no special primes");
    if (hybrid && (K == 0 || K > (size_t)kMaxSpecial))
        return fail(CKKS_ERR_UNSUPPORTED, "hybrid keys need 1..8 special primes");
    size_t total = 0;
    for (size_t j = 0; j < n_blocks; ++j) {
        if (block_sizes[j] == 0) return fail(CKKS_ERR_INVALID, "RNS gadget blocks must be nonempty");
        if (block_sizes[j] > (size_t)kMaxAlpha)
            return fail(CKKS_ERR_UNSUPPORTED, "gadget blocks of at most 12 limbs are supported");
        total += block_sizes[j];
    }
    if (total != q_limbs) return fail(CKKS_ERR_INVALID, "RNS gadget blocks must cover the full basis exactly");
    ckks_context *ctx = plan->ctx;
    CK_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<ckks_ksk> k(new ckks_ksk());
    k->ctx = ctx;
    k->plan = plan;
    k->hybrid = hybrid ? 1 : 0;
    k->q_limbs = q_limbs;
    k->p_limbs = K;
    k->limbs = B;
    k->dnum = n_blocks;
    k->degree = plan->degree;
    const size_t N = plan->degree;
    const auto &mods = plan->moduli;
    // ModUp constants per block
    std::vector<u64> consts;  // w tables of every block, then ModDown tables
    size_t start = 0;
    for (size_t j = 0; j < n_blocks; ++j) {
        ModUpBlock mb{};
        mb.start = (u32)start;
        mb.size = (u32)block_sizes[j];
        std::vector<u64> bm(mods.begin() + start, mods.begin() + start + block_sizes[j]);
        fill_garner(bm, mb.pm, mb.inv, mb.inv_q, 0);
        // FP64 modes: every block prime after the first below 2^41; mode 2
        // when the first one is not (its digit is split at bit 30)
        bool fp_ok = B <= 64;
        for (size_t kk = 1; kk < bm.size(); ++kk) fp_ok = fp_ok && fp_prime(bm[kk]);
        mb.fp_mode = fp_ok ? (fp_prime(bm[0]) ? 1u : 2u) : 0u;
        if (mb.fp_mode) {
            for (size_t kk = 1; kk < bm.size(); ++kk) {
                const u64 qk = bm[kk];
                double *g = mb.gf[kk];
                g[0] = (double)qk;
                g[1] = 1.0 / (double)qk;
                g[2] = (double)mb.inv[kk];
                g[3] = (double)mb.inv[kk] / (double)qk;
                const u64 w2 = host::mulmod(mb.pm[kk][0], (1ull << 30) % qk, qk);
                g[4] = (double)w2;
                g[5] = (double)w2 / (double)qk;
                for (size_t m = 0; m < kk; ++m) {
                    g[6 + 2 * m] = (double)mb.pm[kk][m];
                    g[7 + 2 * m] = (double)mb.pm[kk][m] / (double)qk;
                }
            }
        }
        mb.w = nullptr;  // patched after upload
        k->h_modup.push_back(mb);
        k->block_start.push_back(start);
        k->block_size.push_back(block_sizes[j]);
        k->block_fp.push_back((int)mb.fp_mode);
        for (size_t i = 0; i < B; ++i) {
            u64 prod = 1 % mods[i];
            for (int m = 0; m < kMaxAlpha; ++m) {
                consts.push_back(m < (int)bm.size() ? prod : 0);
                if (m < (int)bm.size()) prod = host::mulmod(prod, bm[m] % mods[i], mods[i]);
            }
        }
        start += block_sizes[j];
    }
    size_t md_w_off = consts.size();
    ModDownTable md{};
    if (hybrid) build_moddown(mods, q_limbs, consts, md);
    // FP64 ModUp target rows of every block (ModUpBlock::tf)
    const size_t tf_off = consts.size();
    for (size_t j = 0; j < n_blocks; ++j) {
        for (size_t i = 0; i < B; ++i) {
            const size_t row0 = consts.size();
            const u64 q = mods[i];
            if (k->block_fp[j] && fp_prime(q)) {
                const u64 w0 = consts[(j * B + i) * kMaxAlpha];
                consts.push_back(double_bits((double)q));
                consts.push_back(double_bits(1.0 / (double)q));
                push_fp_pair(consts, host::mulmod(w0, (1ull << 30) % q, q), q);
                for (size_t m = 0; m < block_sizes[j]; ++m) push_fp_pair(consts, consts[(j * B + i) * kMaxAlpha + m], q);
            }
            consts.resize(row0 + kFpRow, 0);
        }
    }
    k->fp_targets = fp_target_mask(mods, B);
    k->fp_q_targets = fp_target_mask(mods, q_limbs);
    CK_HIP(hipMalloc(&k->d_consts, std::max<size_t>(consts.size(), 1) * sizeof(u64)));
    CK_HIP(hipMemcpy(k->d_consts, consts.data(), consts.size() * sizeof(u64), hipMemcpyHostToDevice));
    for (size_t j = 0; j < n_blocks; ++j) {
        k->h_modup[j].w = k->d_consts + j * B * kMaxAlpha;
        k->h_modup[j].tf = (const double *)(k->d_consts + tf_off) + j * B * kFpRow;
    }
    CK_HIP(hipMalloc(&k->d_modup, n_blocks * sizeof(ModUpBlock)));
    CK_HIP(hipMemcpy(k->d_modup, k->h_modup.data(), n_blocks * sizeof(ModUpBlock), hipMemcpyHostToDevice));
    if (hybrid) {
        patch_moddown(md, k->d_consts + md_w_off, q_limbs);
        CK_HIP(hipMalloc(&k->d_moddown, sizeof(ModDownTable)));
        CK_HIP(hipMemcpy(k->d_moddown, &md, sizeof(ModDownTable), hipMemcpyHostToDevice));
    }
    // exact int8-MFMA base conversion for the hybrid (blocks of >= 2 limbs)
    bool mfma = hybrid;
    // (blocks of 2..9 limbs; larger blocks keep the VALU kernel, whose
    // register footprint does not limit the MFMA kernel's occupancy)
    for (size_t j = 0; j < n_blocks; ++j) mfma &= block_sizes[j] >= 2 && block_sizes[j] <= 9;
    if (mfma) {
        std::vector<BcPlan> plans(n_blocks + 1);
        std::vector<std::vector<signed char>> bts(n_blocks + 1);
        for (size_t j = 0; j <= n_blocks && mfma; ++j) {
            std::vector<u64> src;
            std::vector<std::pair<u32, u64>> tg;
            if (j < n_blocks) {  // ModUp of block j: every other limb of Q u P
                src.assign(mods.begin() + k->block_start[j], mods.begin() + k->block_start[j] + block_sizes[j]);
                for (size_t i = 0; i < B; ++i)
                    if (i < k->block_start[j] || i >= k->block_start[j] + block_sizes[j])
                        tg.emplace_back((u32)i, mods[i]);
            } else {  // ModDown: P digits onto every Q limb
                src.assign(mods.begin() + q_limbs, mods.end());
                for (size_t i = 0; i < q_limbs; ++i) tg.emplace_back((u32)i, mods[i]);
            }
            mfma &= build_bc_plan(src, tg, plans[j], bts[j]);
        }
        if (mfma) {
            size_t total = 0;
            std::vector<size_t> off(n_blocks + 1);
            for (size_t j = 0; j <= n_blocks; ++j) {
                off[j] = total;
                total += (bts[j].size() + 255) / 256 * 256;
            }
            CK_HIP(hipMalloc(&k->d_bt, std::max<size_t>(total, 256)));
            for (size_t j = 0; j <= n_blocks; ++j) {
                CK_HIP(hipMemcpy(k->d_bt + off[j], bts[j].data(), bts[j].size(), hipMemcpyHostToDevice));
                plans[j].bt = k->d_bt + off[j];
            }
            CK_HIP(hipMalloc(&k->d_bc, plans.size() * sizeof(BcPlan)));
            CK_HIP(hipMemcpy(k->d_bc, plans.data(), plans.size() * sizeof(BcPlan), hipMemcpyHostToDevice));
            k->mfma = 1;
        }
    }
    // key material (host or device)
    ctx->keys.insert(k.get());
    if (key_b && key_a) {
        const int rc = upload_key_material(plan, n_blocks * B * N, key_b, key_a, &k->d_kb, &k->d_ka);
        if (rc) {
            ckks_ksk_destroy(k.release());
            return rc;
        }
    }
    *out = k.release();
    return CKKS_OK;
}

// ---- bounded gadget tables ---------------------------------------------------------

struct BoundedHost {
    CrtHost crt;
    unsigned words = 0, digit_count = 0;
    std::vector<u64> &blob = crt.blob;
};

static int build_bounded(const std::vector<u64> &mods, size_t q_limbs, unsigned base_log, BoundedHost &h) {
    if (base_log < 1 || base_log > 63) return fail(CKKS_ERR_INVALID, "bounded gadget base_log must be in 1..=63");
    if (!build_crt(mods, q_limbs, h.crt)) return fail(CKKS_ERR_UNSUPPORTED, "bounded gadget: Q above 2047 bits");
    h.words = h.crt.words;
    h.digit_count = (h.crt.bits + 1 + base_log - 1) / base_log;
    return CKKS_OK;
}

static BoundedTable bounded_table(const u64 *d_blob, size_t q_limbs, unsigned W) {
    return crt_table(d_blob, q_limbs, W);
}

// 1 when every one of the first `limbs` primes of the plan exceeds 2^(w-1)
static u32 digits_below_primes(const ckks_rns_plan *plan, size_t limbs, unsigned w) {
    if (w < 1 || w > 63) return 0;
    for (size_t t = 0; t < limbs && t < plan->moduli.size(); ++t)
        if (plan->moduli[t] <= ((u64)1 << (w - 1))) return 0;
    return 1;
}

static int launch_bounded(hipStream_t s, const BoundedArgs &a, unsigned W) {
    const dim3 g((u32)((a.total + kKsThreads - 1) / kKsThreads));
    switch (W) {
#define CK_BD(WW) \
    case WW: hipLaunchKernelGGL(k_bounded_decompose<WW>, g, dim3(kKsThreads), 0, s, a); break;
        CK_BD(4) CK_BD(8) CK_BD(12) CK_BD(16) CK_BD(24) CK_BD(32)
#undef CK_BD
        default: return fail(CKKS_ERR_UNSUPPORTED, "bounded gadget word count");
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// Key of the bounded family: no special primes (bounded relinearization,
// grafting/bounded_key_switch.rs) or with them (hybrid-QP Galois key,
// ckks/hybrid_rns_galois_key.rs).  Entries [digit][limb] pairs (b, a).
static int create_bounded_key(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                              const u64 *key_b, const u64 *key_a, ckks_ksk **out) {
    if (!plan || !key_b || !key_a || !out) return fail(CKKS_ERR_INVALID, "null argument");
    const size_t B = plan->moduli.size();
    if (q_limbs == 0 || q_limbs > B) return fail(CKKS_ERR_INVALID, "Q limb count must be within the key basis");
    const size_t K = B - q_limbs;
    if (K > (size_t)kMaxSpecial) return fail(CKKS_ERR_UNSUPPORTED, "at most 8 special primes");
    BoundedHost h;
    int rc = build_bounded(plan->moduli, q_limbs, base_log, h);
    if (rc) return rc;
    // the ModDown tables come from the block-key builder (its block layout
    // and placeholder key entries are replaced below; the bounded key never
    // runs the block ModUp)
    ckks_ksk *k = nullptr;
    if (K > 0) {
        std::vector<size_t> blocks;
        for (size_t left = q_limbs; left;) {
            size_t b = std::min<size_t>(left, kMaxAlpha);
            blocks.push_back(b);
            left -= b;
        }
        // the key material of the digit count is uploaded below
        rc = create_key(plan, q_limbs, blocks.data(), blocks.size(), nullptr, nullptr, true, &k);
    } else {
        std::vector<size_t> blocks(q_limbs, 1);
        rc = create_key(plan, q_limbs, blocks.data(), blocks.size(), nullptr, nullptr, false, &k);
    }
    if (rc) return rc;
    std::unique_ptr<ckks_ksk, int (*)(ckks_ksk *)> guard(k, ckks_ksk_destroy);
    hipFree(k->d_bc);
    hipFree(k->d_bt);
    k->d_bc = nullptr;
    k->d_bt = nullptr;
    k->mfma = 0;
    k->bounded = 1;
    k->base_log = base_log;
    k->bounded_words = h.words;
    k->dnum = h.digit_count;
    const size_t N = plan->degree, kelems = (size_t)h.digit_count * B * N;
    if ((rc = upload_key_material(plan, kelems, key_b, key_a, &k->d_kb, &k->d_ka))) return rc;
    CK_HIP(hipMalloc(&k->d_bounded, h.blob.size() * sizeof(u64)));
    CK_HIP(hipMemcpy(k->d_bounded, h.blob.data(), h.blob.size() * sizeof(u64), hipMemcpyHostToDevice));
    *out = guard.release();
    return CKKS_OK;
}

// ---- the pipeline --------------------------------------------------------------------

// Switches the key of component `x` (coefficient domain, q_limbs limbs per
// ct at x + c*x_stride) and writes
//   out[c][0] = add0 + KS_b(x),  out[c][1] = add1 + KS_a(x)
// (add0/add1 may be null) over the key's Q limbs.
// Inputs of a hoisted rotation by k: NTT-domain digits of the unrotated
// component and the key entries permuted by k^-1.
struct HoistedMac {
    const u64 *digits;   // [ct][D][B][N]
    const u64 *kb, *ka;  // [D][B][N]
    u64 k;
    const u64 *tb, *ta;  // canonical-lift correction [B][N] (null for bounded digits)
    // P NTT(b) [ct][q_limbs][N] of a handle hoisted with b (null otherwise):
    // the MAC adds it and ModDown reads no pass-through for component 0
    const u64 *bhat = nullptr;
};

static int key_switch_lane(const ckks_ksk *k, const u64 *x, u64 x_stride, const u64 *add0,
                           const u64 *add1, u64 add_stride, u64 *out, u64 out_stride, size_t n_cts,
                           u64 galois_kinv, const int64_t *hoisted, const HoistedMac *hm);

// Runs the key switch on one lane, or splits the batch over the context
// stream (first half) and the auxiliary stream (second half): fork with an
// event recorded on the context stream, join by making the context stream
// wait for the auxiliary one.  Each lane has its own workspaces.
// hoisted: bounded keys only; the balanced digits [ct][D][N] of the unrotated
// component x (ckks_bounded_digits), which replace its decomposition.
// hm: bounded keys only; NTT-domain digits of the unrotated component
// (ckks_bounded_hoist) with the key permuted by k^-1: the decomposition and
// forward NTT are skipped and the MAC output is permuted by k.
static int key_switch_core(const ckks_ksk *k, const u64 *x, u64 x_stride, const u64 *add0,
                           const u64 *add1, u64 add_stride, u64 *out, u64 out_stride, size_t n_cts,
                           u64 galois_kinv = 0, const int64_t *hoisted = nullptr,
                           const HoistedMac *hm = nullptr) {
    ckks_context *ctx = k->ctx;
    if (ctx->lane == 0) {
        // one key switch per ciphertext; a Galois one is a rotation, or an
        // automorphism when it conjugates (k = 2N - 1 is its own inverse)
        ckks_execution_profile &pc = ctx->counts();
        pc.key_switches += n_cts;
        if (galois_kinv) (galois_kinv == 2 * (u64)k->degree - 1 ? pc.automorphisms : pc.rotations) += n_cts;
    }
    const int split = ctx->ks_split == CKKS_KS_SPLIT_AUTO ? k->tuned_split : ctx->ks_split;
    if (split != CKKS_KS_SPLIT_ON || n_cts < 8 || ctx->capturing || ctx->lane != 0)
        return key_switch_lane(k, x, x_stride, add0, add1, add_stride, out, out_stride, n_cts, galois_kinv,
                               hoisted, hm);
    if (!ctx->aux) {
        CK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
        CK_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        CK_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    }
    const size_t n1 = n_cts / 2, n2 = n_cts - n1;
    CK_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    CK_HIP(hipStreamWaitEvent(ctx->aux, ctx->ev_fork, 0));
    hipStream_t main = ctx->stream;
    ctx->stream = ctx->aux;
    ctx->lane = 1;
    HoistedMac hm2{};
    if (hm) {
        hm2 = *hm;
        hm2.digits += n1 * k->dnum * k->limbs * k->degree;
        if (hm2.bhat) hm2.bhat += n1 * k->q_limbs * k->degree;
    }
    int rc = key_switch_lane(k, x + n1 * x_stride, x_stride, add0 ? add0 + n1 * add_stride : nullptr,
                             add1 ? add1 + n1 * add_stride : nullptr, add_stride, out + n1 * out_stride,
                             out_stride, n2, galois_kinv, hoisted ? hoisted + n1 * k->dnum * k->degree : nullptr,
                             hm ? &hm2 : nullptr);
    ctx->stream = main;
    ctx->lane = 0;
    if (rc) return rc;
    rc = key_switch_lane(k, x, x_stride, add0, add1, add_stride, out, out_stride, n1, galois_kinv, hoisted, hm);
    CK_HIP(hipEventRecord(ctx->ev_join, ctx->aux));
    CK_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
    return rc;
}

// One key switch of a ciphertext batch, split into the phases that a
// multi-key batch (key_switch_multi) interleaves with other entries.
struct KsJob {
    const ckks_ksk *k;
    const u64 *x;
    u64 x_stride;
    const u64 *add0, *add1;
    u64 add_stride;
    u64 *out;
    u64 out_stride;
    size_t n_cts;
    u64 galois_kinv;
    const int64_t *hoisted;
    const HoistedMac *hm;
    u64 *digits, *acc, *mac_out;  // [ct][D][B][N] (or the hoisted digits), [ct][2][B][N], MAC output
    bool mfma;
};

static bool ks_uses_mfma(const ckks_ksk *k) {
    const ckks_context *ctx = k->ctx;
    const int engine = ctx->base_conversion == CKKS_BASECONV_AUTO ? k->tuned_engine : ctx->base_conversion;
    return k->mfma && engine == CKKS_BASECONV_MFMA;
}

// 1. decomposition: ModUp, bounded digits or hoisted-digit automorphism
static int ks_decompose(const KsJob &j, hipStream_t s) {
    const ckks_ksk *k = j.k;
    const ckks_rns_plan *plan = k->plan;
    const u64 N = k->degree, B = k->limbs, D = k->dnum;
    const u32 chunks = (u32)(N / kKsThreads);
    int rc;
    ModUpArgs mu{};
    mu.in = j.x;
    mu.in_stride = j.x_stride;
    mu.kinv = j.galois_kinv;
    mu.digits = j.digits;
    mu.blocks = k->d_modup;
    mu.primes = plan->d_primes;
    mu.fp_targets = k->fp_targets;
    mu.dnum = (u32)D;
    mu.limbs = (u32)B;
    mu.logn = plan->logn;
    mu.chunks = chunks;
    if (j.hm) {
        // decomposed, lifted and transformed once (ckks_bounded_hoist)
    } else if (k->bounded) {
        BoundedArgs ba{};
        ba.in = j.x;
        ba.in_stride = j.x_stride;
        ba.kinv = j.galois_kinv;
        ba.digits = j.digits;
        ba.t = bounded_table(k->d_bounded, k->q_limbs, k->bounded_words);
        ba.primes = plan->d_primes;
        ba.q_limbs = (u32)k->q_limbs;
        ba.limbs = (u32)B;
        ba.logn = plan->logn;
        ba.digit_count = (u32)D;
        ba.base_log = k->base_log;
        ba.small = digits_below_primes(plan, B, k->base_log);
        ba.total = (u64)j.n_cts * N;
        if (j.hoisted) {
            HoistArgs ha{};
            ha.digits = j.hoisted;
            ha.lifted = j.digits;
            ha.primes = plan->d_primes;
            ha.kinv = j.galois_kinv;
            ha.limbs = (u32)B;
            ha.logn = plan->logn;
            ha.digit_count = (u32)D;
            ha.base_log = k->base_log;
            ha.small = ba.small;
            ha.total = (u64)j.n_cts * N;
            if ((rc = launch_hoist(s, ha))) return rc;
        } else if ((rc = launch_bounded(s, ba, k->bounded_words))) {
            return rc;
        }
    } else if (j.mfma) {
        ModUpMfmaArgs mm{mu, k->d_bc};
        hipLaunchKernelGGL(k_modup_mfma, dim3((u32)(j.n_cts * D * (N / kBcRows))), dim3(256), 0, s, mm);
    } else {
        if ((rc = launch_modup(s, mu, k->block_size, k->block_fp, j.n_cts))) return rc;
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// 3. key MAC of the NTT-domain digits into mac_out
// The MAC adds the hoisted P NTT(b) (the fixed-D kernels, dnum <= 8); the
// ModDown of such a job reads no pass-through for component 0.
static bool ks_mac_takes_b(const KsJob &j) { return j.hm && j.hm->bhat && j.k->dnum <= 8; }

static int ks_mac(const KsJob &j, hipStream_t s) {
    const ckks_ksk *k = j.k;
    const u64 N = k->degree, B = k->limbs, D = k->dnum;
    const u32 chunks = (u32)(N / kKsThreads);
    MacArgs ma{};
    ma.digits = j.digits;
    ma.kb = j.hm ? j.hm->kb : k->d_kb;
    ma.ka = j.hm ? j.hm->ka : k->d_ka;
    ma.tb = j.hm ? j.hm->tb : nullptr;
    ma.ta = j.hm ? j.hm->ta : nullptr;
    ma.bhat = ks_mac_takes_b(j) ? j.hm->bhat : nullptr;
    ma.q_limbs = (u32)k->q_limbs;
    ma.acc = j.mac_out;
    ma.primes = k->plan->d_primes;
    ma.n_cts = (u32)j.n_cts;
    ma.dnum = (u32)D;
    ma.limbs = (u32)B;
    ma.logn = k->plan->logn;
    ma.chunks = chunks;
    switch (D) {
#define CK_MAC(DD) \
    case DD: hipLaunchKernelGGL(k_ks_mac<DD>, dim3((u32)(B * chunks)), dim3(kKsThreads), 0, s, ma); break;
        CK_MAC(1) CK_MAC(2) CK_MAC(3) CK_MAC(4) CK_MAC(5) CK_MAC(6) CK_MAC(7) CK_MAC(8)
#undef CK_MAC
        default:
            if (j.n_cts >= 2) {
                for (u32 j0 = 0; j0 < (u32)D; j0 += 8)
                    hipLaunchKernelGGL(k_ks_mac_chunk<8>, dim3((u32)(B * chunks)), dim3(kKsThreads), 0, s, ma, j0,
                                       j0 == 0 ? 1u : 0u);
            } else {
                hipLaunchKernelGGL(k_ks_mac_any, dim3((u32)(j.n_cts * B * chunks)), dim3(kKsThreads), 0, s, ma);
            }
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// The MACs of hoisted jobs jobs[g] (g in grp) that share their digits:
// one k_ks_mac_multi when 2..kMacMaxExp of them have dnum <= 8, else one
// ks_mac each.  (More digits, chunked: a two-exponent chunk kernel took
// 0.41 ms per chunk of 8 against 2 x 0.16 ms for the one-exponent chunks,
// at 132 VGPRs and without their next-ciphertext prefetch.)
static int ks_mac_group(const std::vector<KsJob> &jobs, const std::vector<size_t> &grp, hipStream_t s) {
    const KsJob &j0 = jobs[grp[0]];
    const ckks_ksk *k = j0.k;
    const u64 D = k->dnum;
    if (grp.size() < 2 || D > 8 || !j0.hm) {
        for (size_t g : grp)
            if (int rc = ks_mac(jobs[g], s)) return rc;
        return CKKS_OK;
    }
    const u64 N = k->degree, B = k->limbs;
    MacMultiArgs ma{};
    ma.digits = j0.digits;
    for (size_t e = 0; e < grp.size(); ++e) {
        const KsJob &j = jobs[grp[e]];
        ma.kb[e] = j.hm->kb;
        ma.ka[e] = j.hm->ka;
        ma.tb[e] = j.hm->tb;
        ma.ta[e] = j.hm->ta;
        ma.acc[e] = j.mac_out;
    }
    ma.bhat = ks_mac_takes_b(j0) ? j0.hm->bhat : nullptr;
    ma.primes = k->plan->d_primes;
    ma.n_cts = (u32)j0.n_cts;
    ma.limbs = (u32)B;
    ma.q_limbs = (u32)k->q_limbs;
    ma.logn = k->plan->logn;
    ma.chunks = (u32)(N / kKsThreads);
    const dim3 g((u32)(B * ma.chunks)), t(kKsThreads);
    const int E = (int)grp.size();
    switch (D * 8 + E) {
#define CK_MM(DD, EE) \
    case DD * 8 + EE: hipLaunchKernelGGL((k_ks_mac_multi<DD, EE>), g, t, 0, s, ma); break;
#define CK_MM_D(DD) CK_MM(DD, 2) CK_MM(DD, 3) CK_MM(DD, 4)
        CK_MM_D(1) CK_MM_D(2) CK_MM_D(3) CK_MM_D(4) CK_MM_D(5) CK_MM_D(6) CK_MM_D(7) CK_MM_D(8)
#undef CK_MM_D
#undef CK_MM
        default: return fail(CKKS_ERR_INVALID, "hoisted MAC group: 2..4 exponents, dnum <= 8");
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// 5. ModDown (or plain finalize) of the coefficient-domain accumulators
// with the pass-through adds
static int ks_moddown(const KsJob &j, hipStream_t s) {
    const ckks_ksk *k = j.k;
    const u64 N = k->degree, B = k->limbs, D = k->dnum;
    const u32 chunks = (u32)(N / kKsThreads);
    DownArgs da{};
    da.acc = j.acc;
    da.add0 = ks_mac_takes_b(j) ? nullptr : j.add0;
    da.add1 = j.add1;
    da.add_stride = j.add_stride;
    da.kinv = j.galois_kinv;
    da.out = j.out;
    da.out_stride = j.out_stride;
    da.md = k->d_moddown;
    da.primes = k->plan->d_primes;
    da.fp_targets = k->fp_q_targets;
    da.split = moddown_split(k->plan->moduli, k->q_limbs);
    da.limbs = (u32)B;
    da.q_limbs = (u32)k->q_limbs;
    da.logn = k->plan->logn;
    da.chunks = chunks;
    dim3 g((u32)(j.n_cts * 2 * chunks));
    if (j.mfma) {
        ModDownMfmaArgs dm{da, k->d_bc + D};
        dim3 gm((u32)(j.n_cts * 2 * (N / kBcRows)));
        switch (k->p_limbs) {
#define CK_MDM(KK) \
    case KK: hipLaunchKernelGGL(k_moddown_mfma<KK>, gm, dim3(256), 0, s, dm); break;
            CK_MDM(1) CK_MDM(2) CK_MDM(3) CK_MDM(4) CK_MDM(5) CK_MDM(6) CK_MDM(7) CK_MDM(8)
#undef CK_MDM
            default: return fail(CKKS_ERR_UNSUPPORTED, "too many special primes");
        }
        CK_HIP(hipGetLastError());
        return CKKS_OK;
    }
    switch (k->p_limbs) {
#define CK_MD(KK) \
    case KK: hipLaunchKernelGGL(k_moddown<KK>, g, dim3(kKsThreads), 0, s, da); break;
        CK_MD(0) CK_MD(1) CK_MD(2) CK_MD(3) CK_MD(4) CK_MD(5) CK_MD(6) CK_MD(7) CK_MD(8)
#undef CK_MD
        default: return fail(CKKS_ERR_UNSUPPORTED, "too many special primes");
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// With CKKS_KS_FUSION_INVERSE set on the context, hoisted rotations of
// two-pass degrees with dnum <= 8 compute the key MAC inside the
// Galois-read first inverse NTT pass (the NTT-domain accumulators never
// reach HBM).  Not the default: every output then reads its dnum digit,
// key and correction entries (7 loads at dnum = 3) where the separate MAC
// keeps the key entries in registers across the batch; on MI355X at
// N = 2^16, 64 + 16 ciphertexts, the fused pass took 3.27 ms against
// 1.26 ms (MAC) + 1.30 ms (Galois-read pass).
static bool hoisted_mac_fused(const ckks_context *ctx, u32 logn, size_t dnum) {
    return logn > 12 && logn <= 17 && dnum <= 8 && ctx->ks_fusion == CKKS_KS_FUSION_INVERSE;
}

static int key_switch_lane(const ckks_ksk *k, const u64 *x, u64 x_stride, const u64 *add0,
                           const u64 *add1, u64 add_stride, u64 *out, u64 out_stride, size_t n_cts,
                           u64 galois_kinv, const int64_t *hoisted, const HoistedMac *hm) {
    ckks_context *ctx = k->ctx;
    const ckks_rns_plan *plan = k->plan;
    const u64 N = k->degree, B = k->limbs, D = k->dnum;
    const u32 logn = plan->logn;
    hipStream_t s = ctx->stream;
    int rc;
    void *wd = nullptr, *wa;
    // hoisted rotations of two-pass degrees with dnum <= 8: the MAC runs in
    // the first inverse NTT pass, which reads it through the permutation
    const bool two_pass = logn > 12 && logn <= 17;
    const bool hoist_fused = hm && !hm->bhat && hoisted_mac_fused(ctx, logn, D);
    // hoisted: slot 4 holds the unpermuted MAC output instead of the digits
    const size_t slot4 = hm ? n_cts * 2 * B * N : n_cts * D * B * N;
    if (!hoist_fused && (rc = ctx->workspace(4, slot4 * sizeof(u64), &wd))) return rc;
    if ((rc = ctx->workspace(5, n_cts * 2 * B * N * sizeof(u64), &wa))) return rc;
    KsJob j{k, x, x_stride, add0, add1, add_stride, out, out_stride, n_cts, galois_kinv, hoisted, hm};
    j.digits = hm ? const_cast<u64 *>(hm->digits) : (u64 *)wd;
    j.acc = (u64 *)wa;
    j.mac_out = hm ? (u64 *)wd : j.acc;
    j.mfma = ks_uses_mfma(k);
    // 1. ModUp
    if ((rc = ks_decompose(j, s))) return rc;
    // 2+3. forward NTT of every digit limb, key MAC
    const int fusion = ctx->ks_fusion == CKKS_KS_FUSION_AUTO ? k->tuned_fusion : ctx->ks_fusion;
    // CKKS_KS_FUSION_INVERSE: the MAC is computed inside the first inverse
    // NTT pass (step 4) and the accumulators never reach HBM in the NTT domain
    const bool mac_in_inverse = fusion == CKKS_KS_FUSION_INVERSE && two_pass && D <= 8 && !hm;
    if ((fusion == CKKS_KS_FUSION_XCD || fusion == CKKS_KS_FUSION_LINEAR) && two_pass && !hm) {
        if ((rc = launch_ntt_mac(ctx, plan, j.digits, n_cts, D, B, k->d_kb, k->d_ka, j.acc,
                                 fusion == CKKS_KS_FUSION_XCD)))
            return rc;
    } else {
        NttCall f{plan, j.digits, j.digits, n_cts * D, B, B * N, B * N};
        if (!hm && (rc = launch_ntt(ctx, &f, 1, true))) return rc;
        if (!mac_in_inverse && !hoist_fused && (rc = ks_mac(j, s))) return rc;
    }
    // 4. inverse NTT of both accumulators.  Hoisted: the MAC runs over
    // sigma_k^-1-permuted keys; the inverse NTT reads its output permuted by
    // k, which is sum_j NTT(sigma_k(d_j)) * key_j
    const NttMac mac = hm ? NttMac{hm->digits, hm->kb, hm->ka, (u32)D, hm->tb, hm->ta}
                          : NttMac{j.digits, k->d_kb, k->d_ka, (u32)D};
    NttCall iv{plan, hm ? (hoist_fused ? hm->digits : j.mac_out) : j.acc, j.acc, n_cts * 2, B, B * N, B * N,
               mac_in_inverse || hoist_fused ? &mac : nullptr, hm ? hm->k : 0};
    if ((rc = launch_ntt(ctx, &iv, 1, false))) return rc;
    // 5. ModDown (or plain finalize) with pass-through adds
    return ks_moddown(j, s);
}

// Key switches of several ciphertext batches, each with its own key (keys
// of different ring degrees, levels and kinds may mix): every phase runs for
// all entries before the next, and the forward and inverse NTTs of all
// entries go out as one launch per pass (mixed degrees share the descriptor
// table), so small batches ride in the big batches' launches.  Forward
// row-pass fusion of the key MAC applies per entry (its key's mode); no
// inverse-pass fusion, hoisting or stream split; at most 16 entries.
struct KsMultiEntry {
    const ckks_ksk *k;
    const u64 *x, *add0, *add1;
    u64 x_stride, add_stride;
    u64 *out;
    u64 out_stride;
    size_t n_cts;
    u64 galois_kinv;
};

static int key_switch_multi(ckks_context *ctx, const KsMultiEntry *e, size_t count) {
    if (count > 16) return fail(CKKS_ERR_INVALID, "a key-switch batch holds at most 16 entries");
    size_t dig = 0, acc = 0;
    for (size_t i = 0; i < count; ++i) {
        dig += e[i].n_cts * e[i].k->dnum * e[i].k->limbs * e[i].k->degree;
        acc += e[i].n_cts * 2 * e[i].k->limbs * e[i].k->degree;
        ckks_execution_profile &pc = ctx->counts();
        pc.key_switches += e[i].n_cts;
        if (e[i].galois_kinv)
            (e[i].galois_kinv == 2 * (u64)e[i].k->degree - 1 ? pc.automorphisms : pc.rotations) += e[i].n_cts;
    }
    int rc;
    void *wd, *wa;
    if ((rc = ctx->workspace(4, dig * sizeof(u64), &wd))) return rc;
    if ((rc = ctx->workspace(5, acc * sizeof(u64), &wa))) return rc;
    hipStream_t s = ctx->stream;
    std::vector<KsJob> jobs(count);
    std::vector<NttCall> fwd(count), inv(count);
    u64 *pd = (u64 *)wd, *pa = (u64 *)wa;
    for (size_t i = 0; i < count; ++i) {
        const ckks_ksk *k = e[i].k;
        const size_t B = k->limbs, N = k->degree;
        KsJob &j = jobs[i];
        j = KsJob{k, e[i].x, e[i].x_stride, e[i].add0, e[i].add1, e[i].add_stride, e[i].out, e[i].out_stride,
                  e[i].n_cts, e[i].galois_kinv, nullptr, nullptr};
        j.digits = pd;
        j.acc = j.mac_out = pa;
        j.mfma = ks_uses_mfma(k);
        pd += e[i].n_cts * k->dnum * B * N;
        pa += e[i].n_cts * 2 * B * N;
        fwd[i] = NttCall{k->plan, j.digits, j.digits, e[i].n_cts * k->dnum, B, B * N, B * N};
        inv[i] = NttCall{k->plan, j.acc, j.acc, e[i].n_cts * 2, B, B * N, B * N};
        if ((rc = ks_decompose(j, s))) return rc;
    }
    // entries whose key runs the MAC fused into the digits' forward row pass
    // (tuned, or set on the context) take launch_ntt_mac; the others share
    // one forward launch per pass
    std::vector<char> fused(count, 0);
    std::vector<NttCall> plain;
    for (size_t i = 0; i < count; ++i) {
        const ckks_ksk *k = e[i].k;
        const int f = ctx->ks_fusion == CKKS_KS_FUSION_AUTO ? k->tuned_fusion : ctx->ks_fusion;
        const u32 logn = k->plan->logn;
        fused[i] = (f == CKKS_KS_FUSION_XCD || f == CKKS_KS_FUSION_LINEAR) && logn > 12 && logn <= 17;
        if (!fused[i]) plain.push_back(fwd[i]);
    }
    if ((rc = launch_ntt(ctx, plain.data(), plain.size(), true))) return rc;
    for (size_t i = 0; i < count; ++i) {
        const ckks_ksk *k = e[i].k;
        if (!fused[i]) {
            if ((rc = ks_mac(jobs[i], s))) return rc;
            continue;
        }
        const int f = ctx->ks_fusion == CKKS_KS_FUSION_AUTO ? k->tuned_fusion : ctx->ks_fusion;
        if ((rc = launch_ntt_mac(ctx, k->plan, jobs[i].digits, e[i].n_cts, k->dnum, k->limbs, k->d_kb, k->d_ka,
                                 jobs[i].acc, f == CKKS_KS_FUSION_XCD)))
            return rc;
    }
    if ((rc = launch_ntt(ctx, inv.data(), count, false))) return rc;
    for (size_t i = 0; i < count; ++i)
        if ((rc = ks_moddown(jobs[i], s))) return rc;
    return CKKS_OK;
}

size_t ksk_bytes(const ckks_ksk *k) {
    const size_t entries = 2 * k->dnum * k->limbs * k->degree * sizeof(u64);
    // hoisting tables: a permuted key copy per exponent, plus the correction
    // of canonical-digit keys
    const size_t corr = k->bounded ? 0 : 2 * k->limbs * k->degree * sizeof(u64);
    return entries * (1 + k->hoist_keys.size()) + corr * k->hoist_keys.size();
}

int ksk_info(const ckks_ksk *k, size_t *q_limbs, size_t *degree) {
    if (!k) return CKKS_ERR_INVALID;
    *q_limbs = k->q_limbs;
    *degree = k->degree;
    return CKKS_OK;
}

static int check_ct_args(const ckks_ksk *k, const void *in, const void *out, size_t n_cts, size_t &in_stride,
                         size_t &out_stride, size_t in_polys) {
    if (!k) return fail(CKKS_ERR_INVALID, "null key");
    if (n_cts && (!in || !out)) return fail(CKKS_ERR_INVALID, "null buffer");
    const size_t poly = k->q_limbs * k->degree;
    if (in_stride == 0) in_stride = in_polys * poly;
    if (out_stride == 0) out_stride = 2 * poly;
    if (in_stride < in_polys * poly || out_stride < 2 * poly)
        return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    return CKKS_OK;
}

}  // namespace ck

using namespace ck;

extern "C" {

int ckks_rns_ksk_create(const ckks_rns_plan *plan, const size_t *block_sizes, size_t n_blocks,
                        const uint64_t *key_b, const uint64_t *key_a, ckks_ksk **out) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    if (!key_b || !key_a) return fail(CKKS_ERR_INVALID, "null argument");
    return create_key(plan, plan->moduli.size(), block_sizes, n_blocks, key_b, key_a, false, out);
}

int ckks_hybrid_ksk_create(const ckks_rns_plan *qp_plan, size_t q_limbs, const size_t *block_sizes,
                           size_t n_blocks, const uint64_t *key_b, const uint64_t *key_a, ckks_ksk **out) {
    if (!key_b || !key_a) return fail(CKKS_ERR_INVALID, "null argument");
    return create_key(qp_plan, q_limbs, block_sizes, n_blocks, key_b, key_a, true, out);
}

int ckks_bounded_ksk_create(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                            const uint64_t *key_b, const uint64_t *key_a, ckks_ksk **out) {
    return create_bounded_key(plan, q_limbs, base_log, key_b, key_a, out);
}

int ckks_bounded_digit_count(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                             size_t *count) {
    if (!plan || !count) return fail(CKKS_ERR_INVALID, "null argument");
    if (q_limbs == 0 || q_limbs > plan->moduli.size())
        return fail(CKKS_ERR_INVALID, "Q limb count must be within the plan basis");
    BoundedHost h;
    int rc = build_bounded(plan->moduli, q_limbs, base_log, h);
    if (rc) return rc;
    *count = h.digit_count;
    return CKKS_OK;
}

// Balanced decomposition on the device: lifted digits (lifted_out) or the
// signed digits themselves (raw_out).
static int bounded_decompose_impl(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log, const uint64_t *in,
                                  uint64_t *lifted_out, int64_t *raw_out, size_t n_polys, size_t in_stride) {
    uint64_t *out = lifted_out;
    if (!plan || (n_polys && (!in || (!lifted_out && !raw_out)))) return fail(CKKS_ERR_INVALID, "null argument");
    if (q_limbs == 0 || q_limbs > plan->moduli.size())
        return fail(CKKS_ERR_INVALID, "Q limb count must be within the plan basis");
    ckks_context *ctx = plan->ctx;
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "ckks_bounded_decompose uploads its tables: not capturable");
    BoundedHost h;
    int rc = build_bounded(plan->moduli, q_limbs, base_log, h);
    if (rc || n_polys == 0) return rc;
    const size_t N = plan->degree;
    if (in_stride == 0) in_stride = q_limbs * N;
    if (in_stride < q_limbs * N) return fail(CKKS_ERR_INVALID, "polynomial stride smaller than its limbs");
    void *ws;
    if ((rc = ctx->workspace(3, h.blob.size() * sizeof(u64), &ws))) return rc;
    CK_HIP(hipSetDevice(ctx->device));
    CK_HIP(hipStreamSynchronize(ctx->stream));
    CK_HIP(hipMemcpy(ws, h.blob.data(), h.blob.size() * sizeof(u64), hipMemcpyHostToDevice));
    BoundedArgs ba{};
    ba.in = in;
    ba.in_stride = in_stride;
    ba.kinv = 0;
    ba.digits = out;
    ba.raw = raw_out;
    ba.t = bounded_table((const u64 *)ws, q_limbs, h.words);
    ba.primes = plan->d_primes;
    ba.q_limbs = (u32)q_limbs;
    ba.limbs = (u32)plan->moduli.size();
    ba.logn = plan->logn;
    ba.digit_count = h.digit_count;
    ba.base_log = base_log;
    ba.small = digits_below_primes(plan, plan->moduli.size(), base_log);
    ba.total = (u64)n_polys * N;
    return launch_bounded(ctx->stream, ba, h.words);
}

int ckks_bounded_decompose(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                           const uint64_t *in, uint64_t *out, size_t n_polys, size_t in_stride) {
    if (n_polys && !out) return fail(CKKS_ERR_INVALID, "null argument");
    return bounded_decompose_impl(plan, q_limbs, base_log, in, out, nullptr, n_polys, in_stride);
}

int ckks_bounded_digits(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log, const uint64_t *in,
                        int64_t *out, size_t n_polys, size_t in_stride) {
    if (n_polys && !out) return fail(CKKS_ERR_INVALID, "null argument");
    return bounded_decompose_impl(plan, q_limbs, base_log, in, nullptr, out, n_polys, in_stride);
}

// k^-1 mod 2N for an odd Galois exponent (Newton iteration on a power of two).
static u64 galois_inverse(u64 k, u64 two_n) {
    u64 inv = k;
    for (int it = 0; it < 6; ++it) inv = inv * (2 - k * inv);
    return inv & (two_n - 1);
}

int ckks_bounded_digits_automorphism(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                                     const int64_t *in, int64_t *out, size_t n_polys, uint64_t exponent) {
    if (!plan || (n_polys && (!in || !out))) return fail(CKKS_ERR_INVALID, "null argument");
    if (q_limbs == 0 || q_limbs > plan->moduli.size())
        return fail(CKKS_ERR_INVALID, "Q limb count must be within the plan basis");
    if (in == out && n_polys) return fail(CKKS_ERR_INVALID, "automorphism output must not alias its input");
    const u64 two_n = 2 * (u64)plan->degree;
    const u64 k = exponent % two_n;
    if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
    BoundedHost h;
    int rc = build_bounded(plan->moduli, q_limbs, base_log, h);
    if (rc || n_polys == 0) return rc;
    HoistArgs ha{};
    ha.digits = in;
    ha.raw = out;
    ha.primes = plan->d_primes;
    ha.kinv = galois_inverse(k, two_n);
    ha.limbs = (u32)plan->moduli.size();
    ha.logn = plan->logn;
    ha.digit_count = h.digit_count;
    ha.base_log = base_log;
    ha.small = digits_below_primes(plan, plan->moduli.size(), base_log);
    ha.total = (u64)n_polys * plan->degree;
    return launch_hoist(plan->ctx->stream, ha);
}

int ckks_mod_down(const ckks_rns_plan *qp_plan, size_t q_limbs, const uint64_t *in, uint64_t *out,
                  size_t n_polys) {
    if (!qp_plan || (n_polys && (!in || !out))) return fail(CKKS_ERR_INVALID, "null argument");
    const size_t B = qp_plan->moduli.size();
    if (q_limbs == 0 || q_limbs >= B) return fail(CKKS_ERR_INVALID, "mod-down needs Q limbs and special primes");
    const size_t K = B - q_limbs;
    if (K > (size_t)kMaxSpecial) return fail(CKKS_ERR_UNSUPPORTED, "at most 8 special primes");
    if (n_polys == 0) return CKKS_OK;
    ckks_context *ctx = qp_plan->ctx;
    auto *mp = const_cast<ckks_rns_plan *>(qp_plan);
    auto it = mp->moddown.find(q_limbs);
    if (it == mp->moddown.end()) {
        if (ctx->capturing) return fail(CKKS_ERR_INVALID, "mod-down tables cannot be built during graph capture");
        std::vector<u64> consts;
        ModDownTable md{};
        build_moddown(qp_plan->moduli, q_limbs, consts, md);
        void *dc = nullptr, *dm = nullptr;
        CK_HIP(hipSetDevice(ctx->device));
        CK_HIP(hipMalloc(&dc, consts.size() * sizeof(u64)));
        CK_HIP(hipMemcpy(dc, consts.data(), consts.size() * sizeof(u64), hipMemcpyHostToDevice));
        patch_moddown(md, (const u64 *)dc, q_limbs);
        CK_HIP(hipMalloc(&dm, sizeof(ModDownTable)));
        CK_HIP(hipMemcpy(dm, &md, sizeof(ModDownTable), hipMemcpyHostToDevice));
        it = mp->moddown.emplace(q_limbs, std::make_pair(dc, dm)).first;
    }
    const size_t N = qp_plan->degree;
    DownArgs da{};
    da.acc = in;
    da.out = out;
    da.out_stride = 2 * q_limbs * N;  // polynomials are consecutive [p][q_limbs][N]
    da.md = (const ModDownTable *)it->second.second;
    da.primes = qp_plan->d_primes;
    da.fp_targets = fp_target_mask(qp_plan->moduli, q_limbs);
    da.split = moddown_split(qp_plan->moduli, q_limbs);
    da.limbs = (u32)B;
    da.q_limbs = (u32)q_limbs;
    da.logn = qp_plan->logn;
    da.chunks = (u32)(N / kKsThreads);
    const dim3 g((u32)(n_polys * da.chunks));
    switch (K) {
#define CK_MDS(KK) \
    case KK: hipLaunchKernelGGL(k_moddown<KK>, g, dim3(kKsThreads), 0, ctx->stream, da); break;
        CK_MDS(1) CK_MDS(2) CK_MDS(3) CK_MDS(4) CK_MDS(5) CK_MDS(6) CK_MDS(7) CK_MDS(8)
#undef CK_MDS
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

int ckks_ksk_bytes(const ckks_ksk *k, size_t *bytes) {
    if (!k || !bytes) return fail(CKKS_ERR_INVALID, "null argument");
    *bytes = ksk_bytes(k);
    return CKKS_OK;
}

int ckks_ksk_destroy(ckks_ksk *k) {
    if (!k) return CKKS_OK;
    hipSetDevice(k->ctx->device);
    hipStreamSynchronize(k->ctx->stream);
    k->ctx->keys.erase(k);
    for (auto &e : k->hoist_keys) hipFree(e.second);
    hipFree(k->d_kb);
    hipFree(k->d_ka);
    hipFree(k->d_modup);
    hipFree(k->d_consts);
    hipFree(k->d_moddown);
    hipFree(k->d_bc);
    hipFree(k->d_bt);
    hipFree(k->d_bounded);
    delete k;
    return CKKS_OK;
}

int ckks_autotune_key_switch(ckks_ksk *k, size_t n_cts, int *best_engine, int *best_fusion,
                             float *ms) {
    if (!k || !best_engine || !best_fusion) return fail(CKKS_ERR_INVALID, "null argument");
    ckks_context *ctx = k->ctx;
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "autotuning is not allowed during graph capture");
    ProfileGuard profile_guard(ctx);
    if (n_cts == 0) return fail(CKKS_ERR_INVALID, "autotune needs at least one ciphertext");
    if (ms)
        for (int i = 0; i < 8; ++i) ms[i] = -1.f;
    const size_t poly = k->q_limbs * k->degree, per = 2 * poly;
    CK_HIP(hipSetDevice(ctx->device));
    u64 *buf = nullptr;
    CK_HIP(hipMalloc(&buf, 2 * n_cts * per * sizeof(u64)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int saved_mode = ctx->base_conversion, saved_fusion = ctx->ks_fusion;
    int best_e = CKKS_BASECONV_VALU, best_f = CKKS_KS_FUSION_OFF;
    float best_ms = 0.f;
    const bool can_fuse = k->plan->logn > 12 && k->plan->logn <= 17;
    auto body = [&]() -> int {
        CK_HIP(hipMemsetAsync(buf, 0, 2 * n_cts * per * sizeof(u64), ctx->stream));
        CK_HIP(hipEventCreate(&e0));
        CK_HIP(hipEventCreate(&e1));
        u64 *in = buf, *out = buf + n_cts * per;
        for (int engine = CKKS_BASECONV_VALU; engine <= CKKS_BASECONV_MFMA; ++engine) {
            if (engine == CKKS_BASECONV_MFMA && !k->mfma) continue;
            for (int fusion = CKKS_KS_FUSION_OFF; fusion <= CKKS_KS_FUSION_INVERSE; ++fusion) {
                if (fusion != CKKS_KS_FUSION_OFF && !can_fuse) continue;
                if (fusion == CKKS_KS_FUSION_INVERSE && k->dnum > 8) continue;
                ctx->base_conversion = engine;
                ctx->ks_fusion = fusion;
                int r = key_switch_core(k, in + poly, per, in, nullptr, per, out, per, n_cts);
                if (r) return r;
                const int reps = 2;
                CK_HIP(hipEventRecord(e0, ctx->stream));
                for (int i = 0; i < reps; ++i) {
                    r = key_switch_core(k, in + poly, per, in, nullptr, per, out, per, n_cts);
                    if (r) return r;
                }
                CK_HIP(hipEventRecord(e1, ctx->stream));
                CK_HIP(hipEventSynchronize(e1));
                float t = 0.f;
                CK_HIP(hipEventElapsedTime(&t, e0, e1));
                t /= reps;
                if (ms) ms[(engine - 1) * 4 + fusion] = t;
                if (best_ms == 0.f || t < best_ms) {
                    best_ms = t;
                    best_e = engine;
                    best_f = fusion;
                }
            }
        }
        return CKKS_OK;
    };
    const int saved_split = ctx->ks_split;
    ctx->ks_split = CKKS_KS_SPLIT_OFF;
    int rc = body();
    // then the best pair with the batch split over two streams
    float split_ms = -1.f;
    if (!rc && n_cts >= 8) {
        auto timed_split = [&]() -> int {
            ctx->base_conversion = best_e;
            ctx->ks_fusion = best_f;
            ctx->ks_split = CKKS_KS_SPLIT_ON;
            u64 *in = buf, *out = buf + n_cts * per;
            int r = key_switch_core(k, in + poly, per, in, nullptr, per, out, per, n_cts);
            if (r) return r;
            CK_HIP(hipEventRecord(e0, ctx->stream));
            for (int i = 0; i < 2; ++i)
                if ((r = key_switch_core(k, in + poly, per, in, nullptr, per, out, per, n_cts))) return r;
            CK_HIP(hipEventRecord(e1, ctx->stream));
            CK_HIP(hipEventSynchronize(e1));
            CK_HIP(hipEventElapsedTime(&split_ms, e0, e1));
            split_ms /= 2;
            return CKKS_OK;
        };
        rc = timed_split();
    }
    ctx->ks_split = saved_split;
    ctx->base_conversion = saved_mode;
    ctx->ks_fusion = saved_fusion;
    if (!rc) k->tuned_split = (split_ms > 0 && split_ms < best_ms) ? CKKS_KS_SPLIT_ON : CKKS_KS_SPLIT_OFF;
    hipStreamSynchronize(ctx->stream);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(buf);
    if (rc) return rc;
    k->tuned_engine = best_e;
    k->tuned_fusion = best_f;
    *best_engine = best_e;
    *best_fusion = best_f;
    return CKKS_OK;
}

int ckks_key_switch(const ckks_ksk *key, const uint64_t *in, uint64_t *out, size_t n_cts,
                    size_t in_stride, size_t out_stride) {
    int rc = check_ct_args(key, in, out, n_cts, in_stride, out_stride, 2);
    if (rc || n_cts == 0) return rc;
    const size_t poly = key->q_limbs * key->degree;
    if (in == out) {
        // the a component is consumed by ModUp before any output is written
        // only if ModUp finishes first, which stream order guarantees; b is
        // read by the last kernel while out is written element-by-element at
        // the same positions, which is safe.
    }
    return key_switch_core(key, in + poly, in_stride, in, nullptr, in_stride, out, out_stride, n_cts);
}

int ckks_apply_galois(const ckks_ksk *key, uint64_t exponent, const uint64_t *in, uint64_t *out,
                      size_t n_cts, size_t in_stride, size_t out_stride) {
    int rc = check_ct_args(key, in, out, n_cts, in_stride, out_stride, 2);
    if (rc || n_cts == 0) return rc;
    const u64 two_n = 2 * (u64)key->degree;
    const u64 k = exponent % two_n;
    if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
    if (in == out) return fail(CKKS_ERR_INVALID, "rotation output must not alias its input");
    // sigma_k is fused into the key switch as gathers with k^-1 mod 2N
    const size_t poly = key->q_limbs * key->degree;
    return key_switch_core(key, in + poly, in_stride, in, nullptr, in_stride, out, out_stride, n_cts,
                           galois_inverse(k, two_n));
}

int ckks_apply_galois_batch(ckks_context *ctx, const ckks_galois_batch *entries, size_t count) {
    if (!ctx || (count && !entries)) return fail(CKKS_ERR_INVALID, "null argument");
    std::vector<KsMultiEntry> e;
    for (size_t i = 0; i < count; ++i) {
        const ckks_galois_batch &b = entries[i];
        size_t in_stride = b.in_stride, out_stride = b.out_stride;
        int rc = check_ct_args(b.key, b.in, b.out, b.n_cts, in_stride, out_stride, 2);
        if (rc) return rc;
        if (b.key->ctx != ctx) return fail(CKKS_ERR_INVALID, "batch keys must belong to the context");
        if (b.n_cts == 0) continue;
        const u64 two_n = 2 * (u64)b.key->degree;
        const u64 k = b.exponent % two_n;
        if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
        if (b.in == b.out) return fail(CKKS_ERR_INVALID, "rotation output must not alias its input");
        const size_t poly = b.key->q_limbs * b.key->degree;
        e.push_back(KsMultiEntry{b.key, b.in + poly, b.in, nullptr, in_stride, in_stride, b.out, out_stride, b.n_cts,
                                 galois_inverse(k, two_n)});
    }
    if (e.empty()) return CKKS_OK;
    return key_switch_multi(ctx, e.data(), e.size());
}

int ckks_bounded_apply_galois_hoisted(const ckks_ksk *key, uint64_t exponent, const uint64_t *in,
                                      const int64_t *a_digits, uint64_t *out, size_t n_cts, size_t in_stride,
                                      size_t out_stride) {
    int rc = check_ct_args(key, in, out, n_cts, in_stride, out_stride, 2);
    if (rc || n_cts == 0) return rc;
    if (!key->bounded) return fail(CKKS_ERR_INVALID, "hoisted rotations need a bounded gadget key");
    if (!a_digits) return fail(CKKS_ERR_INVALID, "null argument");
    const u64 two_n = 2 * (u64)key->degree;
    const u64 k = exponent % two_n;
    if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
    if (in == out) return fail(CKKS_ERR_INVALID, "rotation output must not alias its input");
    const size_t poly = key->q_limbs * key->degree;
    return key_switch_core(key, in + poly, in_stride, in, nullptr, in_stride, out, out_stride, n_cts,
                           galois_inverse(k, two_n), a_digits);
}

int ckks_bounded_hoist(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride,
                       ckks_hoisted **out) {
    if (!key || !out || (n_cts && !in)) return fail(CKKS_ERR_INVALID, "null argument");
    if (!key->bounded) return fail(CKKS_ERR_INVALID, "hoisted rotations need a bounded gadget key");
    ckks_context *ctx = key->ctx;
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "hoisting allocates and synchronises: not capturable");
    const size_t N = key->degree, B = key->limbs, D = key->dnum, poly = key->q_limbs * N;
    if (in_stride == 0) in_stride = 2 * poly;
    if (in_stride < 2 * poly) return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    std::unique_ptr<ckks_hoisted> h(new ckks_hoisted());
    h->ctx = ctx;
    h->plan = key->plan;
    h->q_limbs = key->q_limbs;
    h->digit_count = D;
    h->n_cts = n_cts;
    h->base_log = key->base_log;
    if (n_cts) {
        CK_HIP(hipSetDevice(ctx->device));
        CK_HIP(hipMalloc(&h->d_digits, n_cts * D * B * N * sizeof(u64)));
        struct Flags {
            unsigned char *p = nullptr;
            ~Flags() {
                if (p) hipFree(p);
            }
        } flags;
        CK_HIP(hipMalloc(&flags.p, n_cts * N));
        BoundedArgs ba{};
        ba.in = in + poly;
        ba.in_stride = in_stride;
        ba.kinv = 0;
        ba.digits = h->d_digits;
        ba.ties = flags.p;
        ba.t = bounded_table(key->d_bounded, key->q_limbs, key->bounded_words);
        ba.primes = key->plan->d_primes;
        ba.q_limbs = (u32)key->q_limbs;
        ba.limbs = (u32)B;
        ba.logn = key->plan->logn;
        ba.digit_count = (u32)D;
        ba.base_log = key->base_log;
        ba.small = digits_below_primes(key->plan, B, key->base_log);
        ba.total = (u64)n_cts * N;
        int rc = launch_bounded(ctx->stream, ba, key->bounded_words);
        if (rc) return rc;
        NttCall f{key->plan, h->d_digits, h->d_digits, n_cts * D, B, B * N, B * N};
        if ((rc = launch_ntt(ctx, &f, 1, true))) return rc;
        std::vector<unsigned char> host(n_cts * N);
        CK_HIP(hipStreamSynchronize(ctx->stream));
        CK_HIP(hipMemcpy(host.data(), flags.p, host.size(), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < host.size(); ++i)
            if (host[i]) h->ties.emplace_back((u32)(i / N), (u32)(i % N));
    }
    *out = h.release();
    return CKKS_OK;
}

// Return code of the per-exponent tables below when device memory is short.
constexpr int kNoRoomForKeyCopy = -1;

// ---- exact hoisting of canonical (RNS-gadget / dnum) digits -----------------------
//
// The canonical digit of block j is the lift L_j(x) in [0, Q_j) reduced mod
// every key-basis prime p_i.  sigma_k moves coefficient g to n with a sign;
// for a negated position, L_j(-v) = Q_j - L_j(v) when v != 0 mod Q_j, so
//   digit_j(sigma_k(x))_i = sigma_k(digit_j(x)_i) + (Q_j mod p_i) [n negated]
// unless the block value at the source is 0 mod Q_j (every residue of the
// block zero).  In the NTT domain sigma_k is the Galois permutation, so the
// key MAC of the rotated digits is the MAC of the permuted hoisted digits
// plus T_i = NTT_i(neg_k) * sum_j (Q_j mod p_i) key_j,i, a per-key, per-exponent
// constant: added to the sums of every ciphertext, it makes the hoisted
// rotation give the residues of the canonical one.  A batch with a zero
// block value takes the canonical path.

// neg_k over every key-basis limb: 1 where sigma_k negates (n k^-1 mod 2N >= N).
__global__ __launch_bounds__(kKsThreads) void k_neg_pattern(u64 *out, u64 kinv, u32 logn, u64 total) {
    const u64 x = (u64)blockIdx.x * kKsThreads + threadIdx.x;
    if (x >= total) return;
    const u64 N = (u64)1 << logn, n = x & (N - 1);
    out[x] = ((n * kinv) & (2 * N - 1)) >= N ? 1u : 0u;
}

// T[c][i][n] = NEG[i][n] * sum_j qmod[j][i] key_c[j][i][n] mod p_i.
struct CorrArgs {
    const u64 *neg;      // [B][N], NTT domain
    const u64 *kb, *ka;  // [D][B][N]
    const u64 *qmod;     // [D][B]: Q_j mod p_i
    u64 *out;            // [2][B][N]
    const DevPrime *primes;
    u32 dnum, limbs, logn;
    u64 total;  // B * N
};

__global__ __launch_bounds__(kKsThreads) void k_hoist_correction(CorrArgs a) {
    const u64 x = (u64)blockIdx.x * kKsThreads + threadIdx.x;
    if (x >= a.total) return;
    const u32 i = (u32)(x >> a.logn);
    const CK_CONST_AS DevPrime *pr = const_table(a.primes);
    const Barrett br = barrett_at(pr, i);
    Acc128 sb, sa;
    sb.zero();
    sa.zero();
    for (u32 j = 0; j < a.dnum; ++j) {
        const u64 c = a.qmod[(u64)j * a.limbs + i];
        sb.mac(a.kb[(u64)j * a.total + x], c);
        sa.mac(a.ka[(u64)j * a.total + x], c);
    }
    const u64 ng = a.neg[x];
    a.out[x] = mul_mod(ng, sb.reduce(br), br);
    a.out[a.total + x] = mul_mod(ng, sa.reduce(br), br);
}

// Sets *flag when some coefficient of some ciphertext has every residue of a
// gadget block zero (its canonical digit cannot be hoisted exactly).
struct ZeroArgs {
    const u64 *in;  // the a components: limb m of ct c at in + c * stride + m * N
    u64 stride;
    const ModUpBlock *blocks;
    int *flag;
    u32 dnum, logn, chunks;
};

__global__ __launch_bounds__(kKsThreads) void k_zero_blocks(ZeroArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks, ct = b / a.chunks;
    const u64 N = (u64)1 << a.logn;
    const u64 n = (u64)chunk * kKsThreads + threadIdx.x;
    const CK_CONST_AS ModUpBlock *blk = const_table(a.blocks);
    const u64 *x = a.in + ct * a.stride + n;
    for (u32 j = 0; j < a.dnum; ++j) {
        const u32 start = blk[j].start, size = blk[j].size;
        bool zero = true;
        for (u32 m = 0; m < size; ++m) zero = zero && x[(u64)(start + m) * N] == 0;
        if (zero) *a.flag = 1;  // every writer stores the same value
    }
}

// Q_j mod p_i for the key's blocks and basis, [D][B].
static std::vector<u64> block_moduli_mod(const ckks_ksk *key) {
    const std::vector<uint64_t> &mods = key->plan->moduli;
    std::vector<u64> out(key->dnum * key->limbs);
    for (size_t j = 0; j < key->dnum; ++j)
        for (size_t i = 0; i < key->limbs; ++i) {
            unsigned __int128 r = 1 % mods[i];
            for (size_t m = 0; m < key->block_size[j]; ++m) r = r * (mods[key->block_start[j] + m] % mods[i]) % mods[i];
            out[j * key->limbs + i] = (u64)r;
        }
    return out;
}

// The correction of exponent k permuted by k^-1 (it joins the MAC output
// before the inverse NTT reads that permuted by k) into out[2][B][N].
static int hoist_correction(const ckks_ksk *key, u64 k, u64 *out) {
    ckks_context *ctx = key->ctx;
    const u64 two_n = 2 * (u64)key->degree, kinv = galois_inverse(k, two_n);
    const size_t B = key->limbs, N = key->degree, BN = B * N;
    const std::vector<u64> qm = block_moduli_mod(key);
    u64 *tmp = nullptr;  // NEG [B][N], T [2][B][N], qmod [D][B]
    if (hipMalloc(&tmp, (3 * BN + qm.size()) * sizeof(u64)) != hipSuccess) {
        (void)hipGetLastError();
        return kNoRoomForKeyCopy;
    }
    u64 *neg = tmp, *t = tmp + BN, *qmod = tmp + 3 * BN;
    int rc = CKKS_OK;
    hipStream_t s = ctx->stream;
    const dim3 g((u32)((BN + kKsThreads - 1) / kKsThreads));
    if (hipMemcpyAsync(qmod, qm.data(), qm.size() * sizeof(u64), hipMemcpyHostToDevice, s) != hipSuccess)
        rc = fail(CKKS_ERR_HIP, "correction table upload failed");
    if (!rc) {
        hipLaunchKernelGGL(k_neg_pattern, g, dim3(kKsThreads), 0, s, neg, kinv, key->plan->logn, (u64)BN);
        ProfileGuard pg(ctx);  // table setup, not an operation of the caller
        NttCall f{key->plan, neg, neg, 1, B, BN, BN};
        rc = launch_ntt(ctx, &f, 1, true);
    }
    if (!rc) {
        CorrArgs a{neg, key->d_kb, key->d_ka, qmod, t, key->plan->d_primes, (u32)key->dnum, (u32)B,
                   key->plan->logn, (u64)BN};
        hipLaunchKernelGGL(k_hoist_correction, g, dim3(kKsThreads), 0, s, a);
        rc = launch_ntt_permute(s, t, out, kinv, 2 * BN, key->plan->logn);
    }
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = fail(CKKS_ERR_HIP, "correction table build failed");
    hipFree(tmp);
    return rc;
}

// The key entries permuted by k^-1 for hoisted rotations by k (canonical
// digits: and the correction above), built on first use and cached (LRU,
// kHoistKeyCacheMax copies per key).
static int hoisted_key(const ckks_ksk *key, u64 k, const u64 **out) {
    const u64 two_n = 2 * (u64)key->degree;
    const u64 kinv = galois_inverse(k, two_n);
    const size_t DBN = key->dnum * key->limbs * key->degree;
    int rc;
    auto it = key->hoist_keys.find(k);
    auto &lru = key->hoist_lru;
    if (it == key->hoist_keys.end()) {
        ckks_context *ctx = key->ctx;
        if (ctx->capturing) return fail(CKKS_ERR_INVALID, "hoisted key tables cannot be built during graph capture");
        CK_HIP(hipSetDevice(ctx->device));
        if (lru.size() >= kHoistKeyCacheMax) {
            // the evicted copy may still be read by queued work on the stream
            CK_HIP(hipStreamSynchronize(ctx->stream));
            hipFree(key->hoist_keys[lru.back()]);
            key->hoist_keys.erase(lru.back());
            lru.pop_back();
        }
        // canonical-digit keys also keep the correction [2][B][N] after the copy
        const size_t extra = key->bounded ? 0 : 2 * key->limbs * key->degree;
        u64 *pk = nullptr;
        if (hipMalloc(&pk, (2 * DBN + extra) * sizeof(u64)) != hipSuccess) {
            (void)hipGetLastError();
            return kNoRoomForKeyCopy;
        }
        if ((rc = launch_ntt_permute(ctx->stream, key->d_kb, pk, kinv, DBN, key->plan->logn)) ||
            (rc = launch_ntt_permute(ctx->stream, key->d_ka, pk + DBN, kinv, DBN, key->plan->logn)) ||
            (extra && (rc = hoist_correction(key, k, pk + 2 * DBN)))) {
            hipStreamSynchronize(ctx->stream);
            hipFree(pk);
            return rc;
        }
        it = key->hoist_keys.emplace(k, pk).first;
    } else {
        lru.erase(std::find(lru.begin(), lru.end(), k));
    }
    lru.insert(lru.begin(), k);
    *out = it->second;
    return CKKS_OK;
}

// The key switch of a hoisted rotation by k: MAC over the key permuted by
// k^-1 (pk), inverse NTT reading the MAC output permuted by k, ModDown with
// sigma_k(b) passed through.
static int rotate_from_hoisted(const ckks_hoisted *h, const ckks_ksk *key, u64 k, const u64 *pk, const u64 *in,
                               u64 *out, size_t in_stride, size_t out_stride) {
    const size_t DBN = key->dnum * key->limbs * key->degree, BN = key->limbs * key->degree;
    const u64 *tb = key->bounded ? nullptr : pk + 2 * DBN;
    const HoistedMac hm{h->d_digits, pk, pk + DBN, k, tb, tb ? tb + BN : nullptr, h->d_bhat};
    const size_t poly = key->q_limbs * key->degree;
    return key_switch_core(key, in + poly, in_stride, in, nullptr, in_stride, out, out_stride, h->n_cts,
                           galois_inverse(k, 2 * (u64)key->degree), nullptr, &hm);
}

int ckks_bounded_rotate_hoisted(const ckks_hoisted *h, const ckks_ksk *key, uint64_t exponent, const uint64_t *in,
                                uint64_t *out, size_t in_stride, size_t out_stride) {
    if (!h) return fail(CKKS_ERR_INVALID, "null argument");
    const size_t n_cts = h->n_cts;
    int rc = check_ct_args(key, in, out, n_cts, in_stride, out_stride, 2);
    if (rc || n_cts == 0) return rc;
    if (!key->bounded || !h->blocks.empty() || key->plan != h->plan || key->q_limbs != h->q_limbs ||
        key->base_log != h->base_log || key->dnum != h->digit_count)
        return fail(CKKS_ERR_INVALID, "key does not match the hoisted digits (plan, Q limbs, base_log)");
    const u64 two_n = 2 * (u64)key->degree;
    const u64 k = exponent % two_n;
    if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
    if (in == out) return fail(CKKS_ERR_INVALID, "rotation output must not alias its input");
    // sigma_k maps coefficient s to s k mod 2N, negated when that is >= N; a
    // negated tie carries, and the hoisted digits would differ from the
    // decomposition of the rotated a
    for (const auto &t : h->ties)
        if (((u64)t.second * k & (two_n - 1)) >= key->degree)
            return ckks_apply_galois(key, exponent, in, out, n_cts, in_stride, out_stride);
    const u64 *pk = nullptr;
    rc = hoisted_key(key, k, &pk);
    if (rc == kNoRoomForKeyCopy)  // the canonical rotation needs no permuted copy, same residues
        return ckks_apply_galois(key, exponent, in, out, n_cts, in_stride, out_stride);
    if (rc) return rc;
    return rotate_from_hoisted(h, key, k, pk, in, out, in_stride, out_stride);
}

int ckks_bounded_hoist_destroy(ckks_hoisted *h) {
    delete h;
    return CKKS_OK;
}

// The canonical digits of the a components of `in` into h (sized for
// n_cts ciphertexts of key's shape), their forward NTT, and the zero-block
// check (one host synchronisation).
static int hoist_into(ckks_hoisted *h, const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride) {
    ckks_context *ctx = key->ctx;
    const size_t N = key->degree, B = key->limbs, D = key->dnum, poly = key->q_limbs * N;
    if (n_cts == 0) return CKKS_OK;
    CK_HIP(hipSetDevice(ctx->device));
    // digits of the unrotated a: the same decomposition as a key switch
    KsJob j{key, in + poly, in_stride, nullptr, nullptr, 0, nullptr, 0, n_cts, 0, nullptr, nullptr};
    j.digits = h->d_digits;
    j.mfma = ks_uses_mfma(key);
    int rc = ks_decompose(j, ctx->stream);
    if (rc) return rc;
    NttCall f{key->plan, h->d_digits, h->d_digits, n_cts * D, B, B * N, B * N};
    if ((rc = launch_ntt(ctx, &f, 1, true))) return rc;
    if (h->d_bhat) {
        // P NTT(b) on the Q limbs: the reference adds P b before ModDown,
        // which the MAC does here in the NTT domain
        const size_t L = key->q_limbs;
        NttCall fb{key->plan, in, h->d_bhat, n_cts, L, in_stride, L * N};
        if ((rc = launch_ntt(ctx, &fb, 1, true))) return rc;
        if (key->p_limbs) {
            const std::vector<uint64_t> &mods = key->plan->moduli;
            std::vector<uint64_t> pm(L);
            for (size_t i = 0; i < L; ++i) {
                unsigned __int128 r = 1;
                for (size_t m = L; m < key->limbs; ++m) r = r * (mods[m] % mods[i]) % mods[i];
                pm[i] = (uint64_t)r;
            }
            if ((rc = ckks_poly_scalar_mul(key->plan, h->d_bhat, pm.data(), h->d_bhat, n_cts, L, L * N))) return rc;
        }
    }
    // exactness of the hoisted rotations (see hoist_correction)
    ZeroArgs za{in + poly, in_stride, key->d_modup, h->d_flag, (u32)D, key->plan->logn, (u32)(N / kKsThreads)};
    hipError_t e = hipMemsetAsync(h->d_flag, 0, sizeof(int), ctx->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero_blocks, dim3((u32)(n_cts * za.chunks)), dim3(kKsThreads), 0, ctx->stream, za);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h->h_flag, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(CKKS_ERR_HIP, std::string("zero-block check: ") + hipGetErrorString(e));
    h->zero_block = *h->h_flag != 0;
    return CKKS_OK;
}

static int hoist_new(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride, bool with_b,
                     ckks_hoisted **out) {
    if (!key || !out || (n_cts && !in)) return fail(CKKS_ERR_INVALID, "null argument");
    if (key->bounded) {
        if (with_b) return fail(CKKS_ERR_UNSUPPORTED, "hoisting with b serves RNS-gadget and dnum-hybrid keys");
        return ckks_bounded_hoist(key, in, n_cts, in_stride, out);
    }
    ckks_context *ctx = key->ctx;
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "hoisting allocates: not capturable");
    const size_t N = key->degree, B = key->limbs, D = key->dnum, poly = key->q_limbs * N;
    if (in_stride == 0) in_stride = 2 * poly;
    if (in_stride < 2 * poly) return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    std::unique_ptr<ckks_hoisted> h(new ckks_hoisted());
    h->ctx = ctx;
    h->plan = key->plan;
    h->q_limbs = key->q_limbs;
    h->digit_count = D;
    h->n_cts = n_cts;
    h->blocks = key->block_size;
    if (n_cts) {
        CK_HIP(hipSetDevice(ctx->device));
        CK_HIP(hipMalloc(&h->d_digits, n_cts * D * B * N * sizeof(u64)));
        if (with_b) CK_HIP(hipMalloc(&h->d_bhat, n_cts * key->q_limbs * N * sizeof(u64)));
        CK_HIP(hipMalloc(&h->d_flag, sizeof(int)));
        CK_HIP(hipHostMalloc(&h->h_flag, sizeof(int)));
        int rc = hoist_into(h.get(), key, in, n_cts, in_stride);
        if (rc) return rc;
    }
    *out = h.release();
    return CKKS_OK;
}

int ckks_hoist(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride, ckks_hoisted **out) {
    return hoist_new(key, in, n_cts, in_stride, false, out);
}

int ckks_hoist_with_b(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride,
                      ckks_hoisted **out) {
    return hoist_new(key, in, n_cts, in_stride, true, out);
}

int ckks_rehoist(ckks_hoisted *h, const ckks_ksk *key, const uint64_t *in, size_t in_stride) {
    if (!h || !key || (h->n_cts && !in)) return fail(CKKS_ERR_INVALID, "null argument");
    if (key->bounded || h->blocks.empty())
        return fail(CKKS_ERR_UNSUPPORTED, "re-hoisting serves RNS-gadget and dnum-hybrid digits");
    if (key->plan != h->plan || key->q_limbs != h->q_limbs || key->block_size != h->blocks)
        return fail(CKKS_ERR_INVALID, "key does not match the hoisted digits (plan, Q limbs, gadget blocks)");
    if (key->ctx->capturing) return fail(CKKS_ERR_INVALID, "re-hoisting synchronises: not capturable");
    const size_t poly = key->q_limbs * key->degree;
    if (in_stride == 0) in_stride = 2 * poly;
    if (in_stride < 2 * poly) return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    return hoist_into(h, key, in, h->n_cts, in_stride);
}

int ckks_rotate_hoisted(const ckks_hoisted *h, const ckks_ksk *key, uint64_t exponent, const uint64_t *in,
                        uint64_t *out, size_t in_stride, size_t out_stride) {
    if (!h) return fail(CKKS_ERR_INVALID, "null argument");
    if (key && key->bounded) return ckks_bounded_rotate_hoisted(h, key, exponent, in, out, in_stride, out_stride);
    const size_t n_cts = h->n_cts;
    int rc = check_ct_args(key, in, out, n_cts, in_stride, out_stride, 2);
    if (rc || n_cts == 0) return rc;
    if (h->blocks.empty() || key->plan != h->plan || key->q_limbs != h->q_limbs || key->block_size != h->blocks)
        return fail(CKKS_ERR_INVALID, "key does not match the hoisted digits (plan, Q limbs, gadget blocks)");
    const u64 two_n = 2 * (u64)key->degree;
    const u64 k = exponent % two_n;
    if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
    if (in == out) return fail(CKKS_ERR_INVALID, "rotation output must not alias its input");
    // a zero block value, or no room for the per-exponent tables: the
    // canonical key switch, whose residues the hoisted one reproduces
    const u64 kinv = galois_inverse(k, two_n);
    const size_t poly = key->q_limbs * key->degree;
    if (h->zero_block)
        return key_switch_core(key, in + poly, in_stride, in, nullptr, in_stride, out, out_stride, n_cts, kinv);
    const u64 *pk = nullptr;
    rc = hoisted_key(key, k, &pk);
    if (rc == kNoRoomForKeyCopy)
        return key_switch_core(key, in + poly, in_stride, in, nullptr, in_stride, out, out_stride, n_cts, kinv);
    if (rc) return rc;
    return rotate_from_hoisted(h, key, k, pk, in, out, in_stride, out_stride);
}

int ckks_rotate_hoisted_batch(ckks_context *ctx, const ckks_hoisted_rotation *e, size_t count) {
    if (!ctx || (count && !e)) return fail(CKKS_ERR_INVALID, "null argument");
    if (count > 16) return fail(CKKS_ERR_INVALID, "a hoisted rotation batch holds at most 16 entries");
    struct Fast {
        const ckks_hoisted *h;
        const ckks_ksk *key;
        const u64 *in;
        u64 *out;
        size_t in_stride, out_stride;
        u64 k;
        const u64 *pk;
    };
    std::vector<Fast> fast;
    int rc;
    for (size_t i = 0; i < count; ++i) {
        const ckks_hoisted_rotation &r = e[i];
        if (!r.hoisted || !r.key) return fail(CKKS_ERR_INVALID, "null argument");
        const ckks_ksk *key = r.key;
        const ckks_hoisted *h = r.hoisted;
        if (key->ctx != ctx) return fail(CKKS_ERR_INVALID, "hoisted rotation batch entries must share the context");
        size_t in_stride = r.in_stride, out_stride = r.out_stride;
        if ((rc = check_ct_args(key, r.in, r.out, h->n_cts, in_stride, out_stride, 2))) return rc;
        if (h->n_cts == 0) continue;
        if (key->bounded ? (!h->blocks.empty() || key->plan != h->plan || key->q_limbs != h->q_limbs ||
                            key->base_log != h->base_log || key->dnum != h->digit_count)
                         : (h->blocks.empty() || key->plan != h->plan || key->q_limbs != h->q_limbs ||
                            key->block_size != h->blocks))
            return fail(CKKS_ERR_INVALID, "key does not match the hoisted digits (plan, Q limbs, gadget)");
        const u64 two_n = 2 * (u64)key->degree, k = r.exponent % two_n;
        if ((k & 1) == 0) return fail(CKKS_ERR_INVALID, "CKKS automorphism exponent must be odd modulo 2N");
        if (r.in == r.out) return fail(CKKS_ERR_INVALID, "rotation output must not alias its input");
        const u64 kinv = galois_inverse(k, two_n);
        const size_t poly = key->q_limbs * key->degree;
        const u64 *pk = nullptr;
        // the canonical path: a zero block value (canonical digits) or a
        // negated tie (bounded digits, see ckks_bounded_rotate_hoisted)
        bool canonical = h->zero_block;
        for (const auto &t : h->ties)
            canonical = canonical || ((u64)t.second * k & (two_n - 1)) >= key->degree;
        rc = canonical ? kNoRoomForKeyCopy : hoisted_key(key, k, &pk);
        if (rc == kNoRoomForKeyCopy) {  // the canonical key switch, same residues
            if ((rc = key_switch_core(key, r.in + poly, in_stride, r.in, nullptr, in_stride, r.out, out_stride,
                                      h->n_cts, kinv)))
                return rc;
            continue;
        }
        if (rc) return rc;
        fast.push_back(Fast{h, key, r.in, r.out, in_stride, out_stride, k, pk});
    }
    if (fast.empty()) return CKKS_OK;
    // every entry's MAC in the first inverse pass, or none
    bool fused = true;
    for (const Fast &f : fast)
        fused = fused && !f.h->d_bhat && hoisted_mac_fused(ctx, f.key->plan->logn, f.key->dnum);
    // the MAC outputs (slot 4, unfused only) and the accumulators (slot 5)
    size_t total = 0;
    for (const Fast &f : fast) total += f.h->n_cts * 2 * f.key->limbs * f.key->degree;
    void *wd = nullptr, *wa;
    if (!fused && (rc = ctx->workspace(4, total * sizeof(u64), &wd))) return rc;
    if ((rc = ctx->workspace(5, total * sizeof(u64), &wa))) return rc;
    hipStream_t s = ctx->stream;
    std::vector<HoistedMac> hms(fast.size());
    std::vector<KsJob> jobs(fast.size());
    std::vector<NttCall> inv(fast.size());
    std::vector<NttMac> macs(fast.size());
    u64 *pd = (u64 *)wd, *pa = (u64 *)wa;
    for (size_t i = 0; i < fast.size(); ++i) {
        const Fast &f = fast[i];
        const ckks_ksk *key = f.key;
        const size_t B = key->limbs, N = key->degree, DBN = key->dnum * B * N, poly = key->q_limbs * N;
        // bounded keys keep no correction tables after the permuted copy
        const u64 *tb = key->bounded ? nullptr : f.pk + 2 * DBN;
        hms[i] = HoistedMac{f.h->d_digits, f.pk, f.pk + DBN, f.k, tb, tb ? tb + B * N : nullptr, f.h->d_bhat};
        const u64 kinv = galois_inverse(f.k, 2 * (u64)N);
        KsJob &j = jobs[i];
        j = KsJob{key, f.in + poly, f.in_stride, f.in, nullptr, f.in_stride, f.out, f.out_stride, f.h->n_cts, kinv,
                  nullptr, &hms[i]};
        j.digits = const_cast<u64 *>(f.h->d_digits);
        j.mac_out = fused ? nullptr : pd;
        j.acc = pa;
        j.mfma = ks_uses_mfma(key);
        if (!fused) pd += f.h->n_cts * 2 * B * N;
        pa += f.h->n_cts * 2 * B * N;
        const HoistedMac &hm = hms[i];
        macs[i] = NttMac{hm.digits, hm.kb, hm.ka, (u32)key->dnum, hm.tb, hm.ta};
        inv[i] = NttCall{key->plan, fused ? hm.digits : j.mac_out, j.acc, f.h->n_cts * 2, B, B * N, B * N,
                         fused ? &macs[i] : nullptr, f.k};
        ckks_execution_profile &pc = ctx->counts();
        pc.key_switches += f.h->n_cts;
        (f.k == 2 * (u64)N - 1 ? pc.automorphisms : pc.rotations) += f.h->n_cts;
    }
    // the MACs: entries of one hoisted batch share their digit reads, up to
    // kMacMaxExp exponents per launch
    if (!fused) {
        std::vector<bool> done(fast.size(), false);
        for (size_t i = 0; i < fast.size(); ++i) {
            if (done[i]) continue;
            std::vector<size_t> grp{i};
            done[i] = true;
            for (size_t m = i + 1; m < fast.size() && grp.size() < (size_t)kMacMaxExp; ++m)
                if (!done[m] && fast[m].h == fast[i].h && fast[m].key->dnum == fast[i].key->dnum &&
                    fast[m].key->limbs == fast[i].key->limbs) {
                    grp.push_back(m);
                    done[m] = true;
                }
            if ((rc = ks_mac_group(jobs, grp, s))) return rc;
        }
    }
    // one inverse NTT launch per pass for every entry, each reading its MAC
    // (computed in the first pass when fused) through its own Galois
    // permutation
    if ((rc = launch_ntt(ctx, inv.data(), inv.size(), false))) return rc;
    for (KsJob &j : jobs)
        if ((rc = ks_moddown(j, s))) return rc;
    return CKKS_OK;
}

int ckks_relinearize(const ckks_ksk *key, const uint64_t *in3, uint64_t *out, size_t n_cts,
                     size_t in_stride, size_t out_stride) {
    int rc = check_ct_args(key, in3, out, n_cts, in_stride, out_stride, 3);
    if (rc || n_cts == 0) return rc;
    const size_t poly = key->q_limbs * key->degree;
    rc = key_switch_core(key, in3 + 2 * poly, in_stride, in3, in3 + poly, in_stride, out, out_stride, n_cts);
    if (rc == CKKS_OK) key->ctx->counts().relinearizations += n_cts;
    return rc;
}

int ckks_multiply_relinearize_rescale(const ckks_ksk *key, const uint64_t *ct0, const uint64_t *ct1,
                                      uint64_t *out, size_t n_cts, size_t in_stride, size_t out_stride) {
    if (!key) return fail(CKKS_ERR_INVALID, "null key");
    if (n_cts == 0) return CKKS_OK;
    const size_t L = key->q_limbs, N = key->degree, poly = L * N;
    if (L < 2) return fail(CKKS_ERR_INVALID, "RNS CKKS rescale requires at least two modulus limbs");
    if (in_stride == 0) in_stride = 2 * poly;
    if (out_stride == 0) out_stride = 2 * (L - 1) * N;
    ckks_context *ctx = key->ctx;
    int rc;
    void *w6, *w7;
    if ((rc = ctx->workspace(6, n_cts * 3 * poly * sizeof(u64), &w6))) return rc;
    if ((rc = ctx->workspace(7, n_cts * 2 * poly * sizeof(u64), &w7))) return rc;
    u64 *t3 = (u64 *)w6, *t2 = (u64 *)w7;
    if ((rc = ckks_tensor(key->plan, ct0, ct1, t3, n_cts, L, in_stride, 3 * poly))) return rc;
    if ((rc = key_switch_core(key, t3 + 2 * poly, 3 * poly, t3, t3 + poly, 3 * poly, t2, 2 * poly, n_cts)))
        return rc;
    ctx->counts().relinearizations += n_cts;
    return ckks_rescale(key->plan, t2, out, n_cts, L, 2 * poly, out_stride);
}

}  // extern "C"
