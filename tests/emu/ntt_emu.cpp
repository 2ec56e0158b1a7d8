// CPU build of the NTT kernel tiles (the same source the gfx950 kernels are
// compiled from), with every LDS, HBM and twiddle index bounds-checked.  The
// tests use it to check the kernels' addressing and arithmetic on the host
// before any GPU run.  Test infrastructure only.
#define CK_CHECK_BOUNDS 1
#include <cstring>
#include <vector>

#include "../../ccmm-rs_amd/csrc/host_math.hpp"
#include "../../ccmm-rs_amd/csrc/ntt_pass.hpp"

using namespace ck;

template <bool FWD, bool FIRST, bool FP, int TB, bool BLK, bool WIDE = false, bool GU = false>
static void run_pass(const TileJob &base, u32 tiles) {
    static u64 regs[1 << (TB - 4)][16];
    static u64 lds[1 << TB];
    for (u32 tile = 0; tile < tiles; ++tile) {
        TileJob job = base;
        job.tile = tile;
        BlockExec ex{1u << (TB - 4)};
        std::memset(lds, 0xA5, sizeof lds);
        auto regs_of = [&](u32 tid) -> u64(&)[16] { return regs[tid]; };
        ntt_dispatch<FWD, FIRST, FP, TB, BLK, WIDE, GU>(ex, regs_of, job, lds);
    }
}

template <bool FWD, bool FIRST, bool FP, bool WIDE = false, bool GU = false>
static void run_variant(const TileJob &job, u32 logn, int tb, int blk) {
    const bool two = ntt_two_pass(logn);
    if (two && tb == 13) {
        if (blk) run_pass<FWD, FIRST, FP, 13, true, WIDE, GU>(job, ntt_tiles(logn, 13));
        else run_pass<FWD, FIRST, FP, 13, false, WIDE, GU>(job, ntt_tiles(logn, 13));
    } else {
        if (blk && two) run_pass<FWD, FIRST, FP, 12, true, WIDE, GU>(job, ntt_tiles(logn));
        else run_pass<FWD, FIRST, FP, 12, false, WIDE, GU>(job, ntt_tiles(logn));
    }
}

// Transforms one limb of 2^logn residues modulo q (< 2^64) with kernel
// variant (tb = log2 two-pass tile size 12 or 13, blk = blocked inter-pass
// layout).  FP64 primes of two-pass degrees build their row twiddles from
// the row factors, as the kernels do.
extern "C" int emu_ntt_variant(const uint64_t *in, uint64_t *out, uint64_t q, uint32_t logn,
                               int forward, int tb, int blk) {
    if (logn < 10 || logn > 17 || (tb != 12 && tb != 13)) return -1;
    if (tb == 13 && logn < 14) return -1;
    uint64_t psi = host::negacyclic_root(q, 1ull << logn);
    if (!psi) return -2;
    std::vector<uint64_t> tw;
    const bool fp = q < host::kFpModulusBound;
    if (fp)
        host::twiddle_table_fp(q, forward ? psi : host::invmod(psi, q), logn, tw);
    else
        host::twiddle_table(q, forward ? psi : host::invmod(psi, q), logn, tw);
    uint32_t n = 1u << logn;
    std::vector<uint64_t> tmp(n, 0xDEADBEEFull), src(in, in + n), rowf;
    if (fp && ntt_two_pass(logn))
        host::row_factor_table_fp(q, forward ? psi : host::invmod(psi, q), logn, rowf);
    TileJob job{};
    job.tw = reinterpret_cast<const Tw *>(tw.data());
    job.rowf = rowf.empty() ? nullptr : reinterpret_cast<const Tw *>(rowf.data());
    job.q = q;
    job.ninv = host::invmod(n % q, q);
    job.ninv_q = host::shoup_quotient(job.ninv, q);
    job.qd = (double)q;
    job.qinv = 1.0 / (double)q;
    job.ninv_d = (double)job.ninv;
    job.ninv_p = (double)job.ninv / (double)q;
    job.logn = logn;
    bool two = ntt_two_pass(logn);
    job.src = src.data();
    job.dst = two ? tmp.data() : out;
    // moduli from 2^62 up take the canonical-residue (wide) butterflies
    const bool wide = q >= (1ull << 62);
#define CK_RUN(F, S)                                                   \
    do {                                                               \
        if (fp) run_variant<F, S, true>(job, logn, tb, blk);           \
        else if (wide) run_variant<F, S, false, true>(job, logn, tb, blk); \
        else run_variant<F, S, false>(job, logn, tb, blk);             \
    } while (0)
    if (forward) CK_RUN(true, true);
    else CK_RUN(false, true);
    if (two) {
        job.src = tmp.data();
        job.dst = out;
        if (forward) CK_RUN(true, false);
        else CK_RUN(false, false);
    }
    return 0;
}

// Inverse transform of a two-pass degree whose first pass reads its input
// through the NTT-domain Galois map of the odd exponent k < 2N (element n
// of the transformed limb is in[(k n + (k - 1) / 2) mod N]), as the rotation
// kernels run it: whole source segments loaded, permuted in LDS.  The
// inverse of k mod N is computed the way the launch code does.
extern "C" int emu_intt_galois(const uint64_t *in, uint64_t *out, uint64_t q, uint32_t logn, int tb, int blk,
                               uint32_t k) {
    if (logn < 13 || logn > 17 || (tb != 12 && tb != 13) || (tb == 13 && logn < 14)) return -1;
    const uint32_t n = 1u << logn;
    if ((k & 1u) == 0 || k >= 2 * n) return -1;
    uint64_t psi = host::negacyclic_root(q, n);
    if (!psi) return -2;
    const bool fp = q < host::kFpModulusBound;
    if (q >= (1ull << 62)) return -1;
    const uint64_t ipsi = host::invmod(psi, q);
    std::vector<uint64_t> tw, rowf, tmp(n, 0xDEADBEEFull), src(in, in + n);
    if (fp) {
        host::twiddle_table_fp(q, ipsi, logn, tw);
        host::row_factor_table_fp(q, ipsi, logn, rowf);
    } else {
        host::twiddle_table(q, ipsi, logn, tw);
    }
    TileJob job{};
    job.tw = reinterpret_cast<const Tw *>(tw.data());
    job.rowf = rowf.empty() ? nullptr : reinterpret_cast<const Tw *>(rowf.data());
    job.q = q;
    job.ninv = host::invmod(n % q, q);
    job.ninv_q = host::shoup_quotient(job.ninv, q);
    job.qd = (double)q;
    job.qinv = 1.0 / (double)q;
    job.ninv_d = (double)job.ninv;
    job.ninv_p = (double)job.ninv / (double)q;
    job.logn = logn;
    u32 inv = k;
    for (int it = 0; it < 4; ++it) inv *= 2u - k * inv;
    job.gk = k;
    job.gkinv = inv & (n - 1);
    job.src = src.data();
    job.dst = tmp.data();
    if (fp) run_variant<false, true, true, false, true>(job, logn, tb, blk);
    else run_variant<false, true, false, false, true>(job, logn, tb, blk);
    job.gk = job.gkinv = 0;
    job.src = tmp.data();
    job.dst = out;
    if (fp) run_variant<false, false, true>(job, logn, tb, blk);
    else run_variant<false, false, false>(job, logn, tb, blk);
    return 0;
}

extern "C" int emu_ntt(const uint64_t *in, uint64_t *out, uint64_t q, uint32_t logn, int forward) {
    return emu_ntt_variant(in, out, q, logn, forward, 12, 0);
}

// ---- FP64 base-conversion arithmetic ---------------------------------------
// The FP64 target / Garner evaluations of the ModUp and ModDown kernels,
// restated over the same modarith.hpp primitives so the tests can drive
// them with extreme operands and compare with exact integer arithmetic.

static double fp_quot(u64 w, u64 q) { return (double)w / (double)q; }

// ModUp FP64 target: sum_m c_m w_m mod q.  c[0] any residue < 2^62 (split at
// bit 30), c[m > 0] < 2^41; w[m] < q < 2^41; w0s = w[0] 2^30 mod q.
extern "C" uint64_t emu_fp_modup_target(const uint64_t *c, const uint64_t *w, uint64_t w0s, int A, uint64_t q) {
    const double qd = (double)q;
    double acc = fp_mulmod(small_to_f64(c[0] >> 30), (double)w0s, fp_quot(w0s, q), qd);
    for (int m = 0; m < A; ++m) {
        const u64 cm = m == 0 ? (c[0] & 0x3fffffffull) : c[m];
        acc += fp_mulmod(small_to_f64(cm), (double)w[m], fp_quot(w[m], q), qd);
    }
    return fp_residue(acc, qd, 1.0 / qd);
}

// ModUp FP64 Garner step k: (x_k - sum_{m<k} c_m pm[m]) inv mod q, same
// operand ranges; pm0s = pm[0] 2^30 mod q.
extern "C" uint64_t emu_fp_garner_step(uint64_t xk, const uint64_t *c, const uint64_t *pm, uint64_t pm0s, int k,
                                       uint64_t inv, uint64_t q) {
    const double qd = (double)q;
    double t = small_to_f64(xk) - (fp_mulmod(small_to_f64(c[0] & 0x3fffffffull), (double)pm[0], fp_quot(pm[0], q), qd) +
                                   fp_mulmod(small_to_f64(c[0] >> 30), (double)pm0s, fp_quot(pm0s, q), qd));
    for (int m = 1; m < k; ++m) t -= fp_mulmod(small_to_f64(c[m]), (double)pm[m], fp_quot(pm[m], q), qd);
    return f64_to_small(fp_canon(fp_mulmod(t, (double)inv, fp_quot(inv, q), qd), qd, 1.0 / qd));
}

// ModDown FP64 target: ((x - (x_P mod q) + neg P) P^-1 +- add) mod q with
// x_P = sum_m c_m w_m, every digit c_m < 2^62 split at bit 30; w2[m] =
// w[m] 2^30 mod q.
extern "C" uint64_t emu_fp_moddown_target(uint64_t x, const uint64_t *c, const uint64_t *w, const uint64_t *w2, int K,
                                          int neg, uint64_t pmod, uint64_t pinv, uint64_t add, int add_neg,
                                          uint64_t q) {
    const double qd = (double)q;
    double acc = 0.0;
    for (int m = 0; m < K; ++m)
        acc += fp_mulmod(small_to_f64(c[m] & 0x3fffffffull), (double)w[m], fp_quot(w[m], q), qd) +
               fp_mulmod(small_to_f64(c[m] >> 30), (double)w2[m], fp_quot(w2[m], q), qd);
    double t = small_to_f64(x) - acc;
    if (neg) t += (double)pmod;
    double v = fp_mulmod(t, (double)pinv, fp_quot(pinv, q), qd);
    const double ad = small_to_f64(add);
    v = add_neg ? v - ad : v + ad;
    return fp_residue(v, qd, 1.0 / qd);
}

// The same with the split-weight target (fp_split_target), which the ModDown
// kernel uses when every special prime is below 2^60 (digits c_m < 2^60).
template <int K>
static double split_target(const uint64_t *c, const uint64_t *w, uint64_t q) {
    double cl[K], ch[K], wr[2 * K];
    for (int m = 0; m < K; ++m) {
        cl[m] = small_to_f64(c[m] & 0x3fffffffull);
        ch[m] = small_to_f64(c[m] >> 30);
        wr[2 * m] = small_to_f64(w[m] & 0xfffffull);
        wr[2 * m + 1] = small_to_f64(w[m] >> 20);
    }
    const uint64_t k20 = (1ull << 20) % q, k30 = (1ull << 30) % q,
                   k50 = (uint64_t)(((unsigned __int128)1 << 50) % q);
    const double qd = (double)q;
    return fp_split_target<K>(cl, ch, wr, qd, 1.0 / qd, (double)k20, fp_quot(k20, q), (double)k30, fp_quot(k30, q),
                              (double)k50, fp_quot(k50, q));
}

extern "C" uint64_t emu_fp_moddown_target_split(uint64_t x, const uint64_t *c, const uint64_t *w, int K, int neg,
                                                uint64_t pmod, uint64_t pinv, uint64_t add, int add_neg,
                                                uint64_t q) {
    const double qd = (double)q;
    double acc;
    switch (K) {
        case 1: acc = split_target<1>(c, w, q); break;
        case 2: acc = split_target<2>(c, w, q); break;
        case 3: acc = split_target<3>(c, w, q); break;
        case 4: acc = split_target<4>(c, w, q); break;
        case 5: acc = split_target<5>(c, w, q); break;
        case 6: acc = split_target<6>(c, w, q); break;
        case 7: acc = split_target<7>(c, w, q); break;
        case 8: acc = split_target<8>(c, w, q); break;
        default: return ~0ull;
    }
    double t = small_to_f64(x) - acc;
    if (neg) t += (double)pmod;
    double v = fp_mulmod(t, (double)pinv, fp_quot(pinv, q), qd);
    const double ad = small_to_f64(add);
    v = add_neg ? v - ad : v + ad;
    return fp_residue(v, qd, 1.0 / qd);
}

// ---- wave-uniform twiddle stages ------------------------------------------
// Stages that uniform_twiddle_mask flags read their twiddle through the
// scalar unit (readfirstlane of the index).  For every pass geometry the
// kernels instantiate, walk each flagged stage and count waves whose 64
// lanes disagree on the twiddle index.
template <int L, int CL, PassKind KIND, bool FWD>
static void uniform_pass(u32 logn, int &flagged, int &bad) {
    typedef NttPass<L, CL, KIND, FWD, false> P;
    const int T = (1 << (L + CL)) / 16;
    PassAddr pa{};
    pa.logn = logn;
    pa.n2 = (logn + 1) / 2;
    pa.n1 = logn - pa.n2;
    const u32 tiles = KIND == PASS_FULL ? 1 : (1u << logn) >> (L + CL);
    const bool row = KIND == PASS_ROW;
    for (int I = 0; I < n_rounds(L); ++I) {
        const Geom g = P::geom(I);
        const int K = round_k(L, I), s0 = round_start(L, I);
        const u32 um = uniform_twiddle_mask(g, L, CL, K, s0, FWD, row);
        for (int j = 0; j < K; ++j) {
            if (!((um >> j) & 1)) continue;
            ++flagged;
            const int s = s0 + j, rb = FWD ? K - 1 - j : j;
            for (u32 tile = 0; tile < tiles; ++tile) {
                pa.tile = tile;
                for (int i = 0; i < 16; ++i) {
                    if (i & (1 << rb)) continue;
                    for (int w0 = 0; w0 < T; w0 += 64) {
                        long first = -1;
                        for (int tid = w0; tid < w0 + 64 && tid < T; ++tid) {
                            const u32 e = coord(g, tid, i), t = e & ((1u << L) - 1);
                            const u32 base = row ? row_twiddle_base<L, CL>(pa, e) : 1u;
                            const u32 idx = FWD ? (base << s) + (t >> (L - s)) : (base << (L - 1 - s)) + (t >> (s + 1));
                            if (first < 0) first = idx;
                            else if ((long)idx != first) ++bad;
                        }
                    }
                }
            }
        }
    }
}

// Row passes of FP64 primes build twiddles from row factors: the shared
// index (stage flagged in the mask without row bits) and, when the geometry
// says so, the row itself are read through the scalar unit.
template <int L, int CL, PassKind KIND, bool FWD>
static void uniform_row_factors(u32 logn, int &flagged, int &bad) {
    typedef NttPass<L, CL, KIND, FWD, true> P;
    const int T = (1 << (L + CL)) / 16;
    const u32 tiles = (1u << logn) >> (L + CL);
    for (int I = 0; I < n_rounds(L); ++I) {
        const Geom g = P::geom(I);
        const int K = round_k(L, I), s0 = round_start(L, I);
        const u32 sm = uniform_twiddle_mask(g, L, CL, K, s0, FWD, false);
        const bool ru = wave_uniform_bits(g, L, L + CL);
        for (int j = 0; j < K; ++j) {
            const bool s_flag = (sm >> j) & 1;
            if (!s_flag && !ru) continue;
            flagged += (int)s_flag + (int)ru;
            const int s = s0 + j, rb = FWD ? K - 1 - j : j;
            for (u32 tile = 0; tile < tiles; ++tile)
                for (int i = 0; i < 16; ++i) {
                    if (i & (1 << rb)) continue;
                    for (int w0 = 0; w0 < T; w0 += 64) {
                        long first_s = -1, first_r = -1;
                        for (int tid = w0; tid < w0 + 64 && tid < T; ++tid) {
                            const u32 e = coord(g, tid, i), t = e & ((1u << L) - 1);
                            const u32 sidx = FWD ? (1u << s) + (t >> (L - s)) : (1u << (L - 1 - s)) + (t >> (s + 1));
                            const u32 r = tile * (1u << CL) + (e >> L);
                            if (s_flag) {
                                if (first_s < 0) first_s = sidx;
                                else if ((long)sidx != first_s) ++bad;
                            }
                            if (ru) {
                                if (first_r < 0) first_r = r;
                                else if ((long)r != first_r) ++bad;
                            }
                        }
                    }
                }
        }
    }
}

template <int L, int CL, PassKind KIND>
static void uniform_both(u32 logn, int &flagged, int &bad) {
    if (KIND == PASS_ROW) {
        uniform_row_factors<L, CL, KIND, true>(logn, flagged, bad);
        uniform_row_factors<L, CL, KIND, false>(logn, flagged, bad);
    }
    uniform_pass<L, CL, KIND, true>(logn, flagged, bad);
    uniform_pass<L, CL, KIND, false>(logn, flagged, bad);
}

// Returns the number of non-uniform waves over every flagged stage; the
// number of flagged stages goes to *flagged.
extern "C" int emu_twiddle_uniform(int *flagged_out) {
    int flagged = 0, bad = 0;
    uniform_both<10, 0, PASS_FULL>(10, flagged, bad);
    uniform_both<11, 0, PASS_FULL>(11, flagged, bad);
    uniform_both<12, 0, PASS_FULL>(12, flagged, bad);
    // two-pass, 4096-residue tiles: column L = n1, row L = n2 = ceil(logn / 2)
    uniform_both<6, 6, PASS_COL>(13, flagged, bad);
    uniform_both<7, 5, PASS_ROW>(13, flagged, bad);
    uniform_both<7, 5, PASS_COL>(14, flagged, bad);
    uniform_both<7, 5, PASS_ROW>(14, flagged, bad);
    uniform_both<7, 5, PASS_COL>(15, flagged, bad);
    uniform_both<8, 4, PASS_ROW>(15, flagged, bad);
    uniform_both<8, 4, PASS_COL>(16, flagged, bad);
    uniform_both<8, 4, PASS_ROW>(16, flagged, bad);
    uniform_both<8, 4, PASS_COL>(17, flagged, bad);
    uniform_both<9, 3, PASS_ROW>(17, flagged, bad);
    // 8192-residue tiles (N >= 2^14)
    uniform_both<7, 6, PASS_COL>(14, flagged, bad);
    uniform_both<7, 6, PASS_ROW>(14, flagged, bad);
    uniform_both<7, 6, PASS_COL>(15, flagged, bad);
    uniform_both<8, 5, PASS_ROW>(15, flagged, bad);
    uniform_both<8, 5, PASS_COL>(16, flagged, bad);
    uniform_both<8, 5, PASS_ROW>(16, flagged, bad);
    uniform_both<8, 5, PASS_COL>(17, flagged, bad);
    uniform_both<9, 4, PASS_ROW>(17, flagged, bad);
    *flagged_out = flagged;
    return bad;
}

// ---- 128-bit accumulation and reduction ------------------------------------
// Acc128 (six carry-free partial sums) against unsigned __int128 sums of up to
// 16 products of operands below 2^bits (bits <= 62, all-ones operands
// included), and barrett_reduce128 (two Shoup products) against the 128-bit
// remainder for odd q < 2^62 and arbitrary hi, lo.  Returns the mismatches.
extern "C" long emu_wide_arith_check(uint64_t seed, long iters) {
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&]() {  // splitmix64
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    typedef unsigned __int128 u128;
    long bad = 0;
    for (long it = 0; it < iters; ++it) {
        const int n = 1 + (int)(next() % 16), bits = 40 + (int)(next() % 23);
        Acc128 a;
        a.zero();
        u128 ref = 0;
        for (int k = 0; k < n; ++k) {
            u64 x = next() >> (64 - bits), y = next() >> (64 - bits);
            if (it % 7 == 0) x = y = (1ull << bits) - 1;
            a.mac(x, y);
            ref += (u128)x * y;
        }
        u64 hi, lo;
        a.fold(hi, lo);
        bad += hi != (u64)(ref >> 64) || lo != (u64)ref;
        u64 q = (next() >> (2 + next() % 61)) | 1;
        if (q < 3) q = 3;
        if (it % 5 == 0) q = (1ull << 62) - 57;
        u64 bh, bl;
        host::barrett(q, bh, bl);
        const Barrett b{q, bh, bl};
        u64 h = next(), l = next();
        if (it % 3 == 0) h %= q;
        if (it % 11 == 0) h = l = ~0ull;
        const u128 x = ((u128)h << 64) | l;
        bad += barrett_reduce128(h, l, b) != (u64)(x % q);
    }
    return bad;
}

// host::negacyclic_root, the root every plan is built with.
extern "C" uint64_t emu_negacyclic_root(uint64_t q, uint64_t degree) { return host::negacyclic_root(q, degree); }
