// Negacyclic NTT tiles for 64-bit residues, shared by the gfx950 kernels and
// a CPU build of the same code (used by the tests to check every index).
//
// Transform (ring/ntt.rs:100-127): X_k = sum_j a_j psi^((2k+1) j), natural
// order in and out.  Internally it is the Cooley-Tukey negacyclic network
// with twiddle table T[k] = psi^bitrev_n(k) (k in [1, N)), whose output at
// position p is X_{bitrev_n(p)}; the inverse is the Gentleman-Sande network
// with Tinv[k] = psi^-bitrev_n(k) followed by N^-1.  Because psi_{N/2} =
// psi_N^2, the table for N/2 is the first N/2 entries of the table for N,
// so one table per prime serves every supported degree.
//
// Decomposition: N = N1 * N2 (n = n1 + n2).
//   column pass: stages 0..n1-1 act on columns {c + N2 t : t < N1};
//   row pass:    stages n1..n-1 act on rows    {m N2 + r : r < N2}.
// Forward = column pass then row pass; inverse = row pass then column pass.
// A tile is 4096 elements (COLS sub-transforms of R = 2^L elements), held
// as 16 residues per thread in 256 threads; radix-2^k rounds of up to 4
// stages run in registers and exchange through LDS between rounds.  For
// N <= 4096 a single "full" pass holds the whole transform in one tile.
//
// Twiddle index for a butterfly at local stage s of a sub-transform whose
// element index is t (lower element of the pair):
//   forward: (base << s) + (t >> (L - s))
//   inverse: (base << (L - 1 - u)) + (t >> (u + 1))   (u = local GS stage)
// with base = 1 for column / full passes and base = 2^n1 + m for row m.
#pragma once
#include "modarith.hpp"

// Host builds may define CK_CHECK_BOUNDS to trap any out-of-range index of
// the kernel code (the device build compiles the checks out).
#if defined(CK_CHECK_BOUNDS) && !defined(__HIP_DEVICE_COMPILE__)
#include <cstdio>
#include <cstdlib>
#define CK_BOUNDS(x, n)                                                        \
    do {                                                                       \
        if (!((x) < (n))) {                                                    \
            std::fprintf(stderr, "index %u out of range %u at %s:%d\n",        \
                         (unsigned)(x), (unsigned)(n), __FILE__, __LINE__);    \
            std::abort();                                                      \
        }                                                                      \
    } while (0)
#else
#define CK_BOUNDS(x, n) ((void)0)
#endif

namespace ck {

enum PassKind { PASS_COL = 0, PASS_ROW = 1, PASS_FULL = 2 };

struct Tw {
    u64 w, wq;
};

// Per-launch view of one limb transform, resolved by the kernel.
struct TileJob {
    const u64 *src;  // input residues of this limb transform
    u64 *dst;        // output residues
    const Tw *tw;    // forward or inverse twiddle table (index 1..N-1)
    // FP64 primes of two-pass degrees: row factors of tw
    // (row_factor_table_fp); row passes build each twiddle as tw[2^s + j]
    // times its row's factor
    const Tw *rowf;
    u64 q;
    u64 ninv, ninv_q;  // N^-1 and its Shoup quotient (inverse only)
    // FP64 path (q < 2^42): twiddle entries are (double w, double w/q)
    double qd, qinv;    // q and 1/q as doubles
    double ninv_d, ninv_p;  // N^-1 and N^-1/q as doubles
    u32 logn;
    u32 tile;  // tile index within the limb transform
    // Galois-read inverse row pass: input element n is src[(gk n + (gk - 1) / 2) mod N]
    // for the odd exponent gk < 2N; gkinv = gk^-1 mod N
    u32 gk, gkinv;
};

CK_HD u32 bitrev(u32 x, u32 bits) {
#if defined(__HIP_DEVICE_COMPILE__)
    return bits ? __builtin_bitreverse32(x) >> (32 - bits) : 0u;  // v_bfrev_b32
#else
    u32 r = 0;
    for (u32 i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
#endif
}

// ---- compile-time round geometry ------------------------------------------

// Round sizes: split L stages into ceil(L/4) rounds as evenly as possible.
constexpr int n_rounds(int L) { return (L + 3) / 4; }
constexpr int round_k(int L, int i) {
    // first (L % nr) rounds get one extra stage
    return L / n_rounds(L) + (i < L % n_rounds(L) ? 1 : 0);
}
constexpr int round_start(int L, int i) {
    int s = 0;
    for (int j = 0; j < i; ++j) s += round_k(L, j);
    return s;
}

// Bit layout of a tile coordinate e = t | (col << L), TB = L + CL bits
// (TB = 12 for two-pass tiles).  A round keeps 4 "register" bits per thread
// (the k butterfly bits plus 4-k extra bits) and maps the other TB-4 bits to
// the thread id in a preference order that decides coalescing:
//   col-first: lanes walk columns (contiguous in HBM for column passes and
//              for the natural-order side of row passes);
//   t-first:   lanes walk t (contiguous for the internal side of row passes).
struct Geom {
    int reg_pos[4];
    int thr_pos[12];
    int n_thr;
};

constexpr Geom make_geom(int L, int CL, int bfly_lo, int k, bool col_first) {
    Geom g{};
    int TB = L + CL;
    // preference order of free bits
    int order[16] = {};
    int no = 0;
    if (col_first) {
        for (int b = L; b < TB; ++b) order[no++] = b;
        for (int b = 0; b < L; ++b)
            if (b < bfly_lo || b >= bfly_lo + k) order[no++] = b;
    } else {
        for (int b = 0; b < L; ++b)
            if (b < bfly_lo || b >= bfly_lo + k) order[no++] = b;
        for (int b = L; b < TB; ++b) order[no++] = b;
    }
    for (int b = 0; b < k; ++b) g.reg_pos[b] = bfly_lo + b;
    // extra register bits: the least preferred free bits
    int extra = 4 - k;
    for (int i = 0; i < extra; ++i) g.reg_pos[k + i] = order[no - extra + i];
    g.n_thr = no - extra;
    for (int i = 0; i < g.n_thr; ++i) g.thr_pos[i] = order[i];
    return g;
}

template <int L, int CL>
struct TileShape {
    static constexpr int TB = L + CL;
    static constexpr int ELEMS = 1 << TB;
    static constexpr int THREADS = ELEMS / 16;
    static constexpr int COLS = 1 << CL;
    static constexpr int R = 1 << L;
};

// Coordinate of register v for thread tid under geometry g.
CK_HD u32 coord(const Geom &g, u32 tid, int v) {
    u32 e = 0;
    for (int i = 0; i < g.n_thr; ++i) e |= ((tid >> i) & 1u) << g.thr_pos[i];
    for (int b = 0; b < 4; ++b) e |= (u32)((v >> b) & 1) << g.reg_pos[b];
    return e;
}

// LDS placement.  The linear address a = (t << CL) | col is swizzled by
// XORing its bits 0..4 with M * a[5..12] over GF(2), M's columns below.
// Every LDS access of the passes moves 64 lanes whose addresses differ in
// six bit positions (a fixed bit permutation of the lane id); M was picked
// by checking all 22 such position sets of the column, row and full passes
// (N = 2^10 .. 2^17, 4096- and 8192-residue tiles): each set maps onto five
// independent low bits, so a wave's 64 u64 cover each 8-byte bank slot
// exactly twice (and each of 16 slots four times with 32 banks).  The map is
// a bijection and linear: lds_index(a ^ b) = lds_index(a) ^ lds_index(b) for
// disjoint bit sets, which lets the kernels add a per-register constant to a
// per-thread base.
constexpr u32 kLdsSwizzle[8] = {20, 7, 9, 29, 31, 25, 22, 19};

CK_HD u32 lds_swizzle(u32 a) {
    u32 x = a;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if ((a >> (5 + k)) & 1u) x ^= kLdsSwizzle[k];
    return x;
}

template <int L, int CL>
CK_HD u32 lds_index(u32 e) {
    CK_BOUNDS(e, 1u << (L + CL));
    u32 t = e & ((1u << L) - 1), col = e >> L;
    return lds_swizzle((t << CL) | col);
}

// Swizzled LDS row offsets for the Galois-read row pass, whose load places
// each residue at a data-dependent row: entry h is
// lds_swizzle(bitrev_L(h) << CL), so lds_index(bitrev_L(h) | col << L) =
// table[h] ^ lds_swizzle(col) (the swizzle is linear over disjoint bits),
// one cached read instead of a bit reversal and eight conditional XORs.
constexpr u32 lds_swizzle_c(u32 a) {
    u32 x = a;
    for (int k = 0; k < 8; ++k)
        if ((a >> (5 + k)) & 1u) x ^= kLdsSwizzle[k];
    return x;
}
constexpr u32 bitrev_c(u32 x, u32 bits) {
    u32 r = 0;
    for (u32 i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}
template <int L, int CL>
struct GuRowTable {
    unsigned short v[1 << L];
    constexpr GuRowTable() : v() {
        for (u32 h = 0; h < (1u << L); ++h) v[h] = (unsigned short)lds_swizzle_c(bitrev_c(h, L) << CL);
    }
};
#if defined(__HIPCC__)
template <int L, int CL>
__device__ constexpr GuRowTable<L, CL> kGuRow{};
#else
template <int L, int CL>
constexpr GuRowTable<L, CL> kGuRow{};
#endif

// ---- butterflies ------------------------------------------------------------

CK_HD void ct_bfly(u64 &x, u64 &y, const Tw &w, u64 q) {
    u64 two_q = q << 1;
    u64 a = x >= two_q ? x - two_q : x;
    u64 t = shoup_lazy(y, w.w, w.wq, q);
    x = a + t;
    y = a - t + two_q;
}

CK_HD void gs_bfly(u64 &x, u64 &y, const Tw &w, u64 q) {
    u64 two_q = q << 1;
    u64 s = x + y;
    s = s >= two_q ? s - two_q : s;
    u64 d = x - y + two_q;
    x = s;
    y = shoup_lazy(d, w.w, w.wq, q);
}

// Butterflies of moduli in [2^62, 2^64) on canonical residues (shoup_wide);
// sub_mod is exact for any q < 2^64.
CK_HD void ct_bfly_wide(u64 &x, u64 &y, const Tw &w, u64 q) {
    const u64 t = shoup_wide(y, w.w, w.wq, q);
    const u64 a = x;
    x = add_mod_wide(a, t, q);
    y = sub_mod(a, t, q);
}

CK_HD void gs_bfly_wide(u64 &x, u64 &y, const Tw &w, u64 q) {
    const u64 a = x, b = y;
    x = add_mod_wide(a, b, q);
    y = shoup_wide(sub_mod(a, b, q), w.w, w.wq, q);
}

// FP64 butterflies on bit-cast doubles (see fp_mulmod).  Forward: no
// reduction; |x| grows by at most 0.51 q per stage, so after 17 stages it is
// below 10 q < 2^46 (q < 2^42).  Inverse: the product output is within
// 0.51 q; the sum doubles per stage and is reduced (into 0.51 q) only at the
// last stage of each round (REDUCE), i.e. after at most 3 lazy stages: from
// inputs below q, the reducing stage sees |x|, |y| < 8 q, so x - y and x + y
// stay below 16 q < 2^45 (q < 2^41), inside fp_mulmod's and fp_reduce's range.
CK_HD void ct_bfly_fp(u64 &x, u64 &y, const Tw &w, double q) {
    double a = u64_as_f64(x);
    double t = fp_mulmod(u64_as_f64(y), u64_as_f64(w.w), u64_as_f64(w.wq), q);
    x = f64_as_u64(a + t);
    y = f64_as_u64(a - t);
}

template <bool REDUCE>
CK_HD void gs_bfly_fp(u64 &x, u64 &y, const Tw &w, double q, double qinv) {
    double a = u64_as_f64(x), b = u64_as_f64(y);
    x = f64_as_u64(REDUCE ? fp_reduce(a + b, q, qinv) : a + b);
    y = f64_as_u64(fp_mulmod(a - b, u64_as_f64(w.w), u64_as_f64(w.wq), q));
}

// ---- per-pass element addressing ----------------------------------------------

// Global element offsets (within the limb) for tile coordinate e.
// Column pass: t indexes rows of a column, col indexes columns.
// Row pass: col indexes rows u (natural side), t = r.
// Galois-read inverse row passes (GU below) also carry the permutation
// n -> gk n + gc (mod N) of the source and gkinv = gk^-1 mod N.
struct PassAddr {
    u32 logn, n1, n2;
    u32 tile;
    u32 gk, gc, gkinv;
};

// 24-bit product: v_mul_u32_u24 issues at full rate where the 32-bit
// v_mul_lo_u32 is quarter rate.  Both operands below 2^24.
CK_HD u32 mul24(u32 a, u32 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    CK_BOUNDS(a, 1u << 24);
    CK_BOUNDS(b, 1u << 24);
    return a * b;
#endif
}

// Row u of tile column col in a row pass.  Plain: the tile's 2^CL
// consecutive rows.  GU (inverse row pass reading its input through the
// NTT-domain Galois map n -> gk n + gc): the rows whose images have low
// n1 bits v = tile 2^CL + col, so the tile's source is 2^CL whole source
// columns (contiguous segments of every source row, see the staged load);
// u = (v - gc) gk^-1 mod 2^n1.
template <int CL, bool GU>
CK_HD u32 row_u(const PassAddr &pa, u32 col) {
    const u32 v = pa.tile * (1u << CL) + col;
    const u32 m1 = (1u << pa.n1) - 1;
    return GU ? mul24((v - pa.gc) & m1, pa.gkinv) & m1 : v;
}

template <int L, int CL>
CK_HD u32 col_pass_offset(const PassAddr &pa, u32 e) {
    u32 t = e & ((1u << L) - 1), col = e >> L;
    u32 c = pa.tile * (1u << CL) + col;
    return c + (t << pa.n2);
}

// Inter-pass (internal) position of element (m, c): row m of the row pass,
// column c of the column pass.  Plain layout: m * N2 + c.  Blocked layout
// (BLK): the intersection of one column tile (2^CLc columns) and one row
// tile (2^CLr rows m sharing their low lt = log2(tiles) bits) is stored as
// one contiguous chunk, so both passes move whole chunks of
// 2^(CLr + CLc) residues (2 KiB at N = 2^16 with 4096-residue tiles) instead
// of 2^CL-residue segments.  TB = log2 of the tile size.
template <int TB, bool BLK>
CK_HD u32 internal_offset(const PassAddr &pa, u32 m, u32 c) {
    if (!BLK) return (m << pa.n2) + c;
    const u32 lt = pa.logn - TB;
    const u32 clc = TB - pa.n1, clr = TB - pa.n2;
    const u32 m_lo = m & ((1u << lt) - 1), m_hi = m >> lt;
    const u32 c_hi = c >> clc, c_lo = c & ((1u << clc) - 1);
    return ((((m_lo << lt) + c_hi) << (clr + clc)) + (m_hi << clc)) + c_lo;
}

// Row pass, internal side: element (m, r) with m = bitrev_n1(u).
template <int L, int CL, bool BLK, bool GU = false>
CK_HD u32 row_internal_offset(const PassAddr &pa, u32 e) {
    u32 r = e & ((1u << L) - 1), col = e >> L;
    u32 u = row_u<CL, GU>(pa, col);
    u32 m = bitrev(u, pa.n1);
    return internal_offset<L + CL, BLK>(pa, m, r);
}

// Column pass, internal side: element (m = t, c).
template <int L, int CL, bool BLK>
CK_HD u32 col_internal_offset(const PassAddr &pa, u32 e) {
    u32 t = e & ((1u << L) - 1), col = e >> L;
    u32 c = pa.tile * (1u << CL) + col;
    return internal_offset<L + CL, BLK>(pa, t, c);
}

// Row pass, natural side: bitrev_n2(r) * N1 + u.
template <int L, int CL>
CK_HD u32 row_natural_offset(const PassAddr &pa, u32 e) {
    u32 r = e & ((1u << L) - 1), col = e >> L;
    u32 u = pa.tile * (1u << CL) + col;
    return (bitrev(r, pa.n2) << pa.n1) + u;
}

template <int L, int CL, bool GU = false>
CK_HD u32 row_twiddle_base(const PassAddr &pa, u32 e) {
    u32 col = e >> L;
    u32 u = row_u<CL, GU>(pa, col);
    return (1u << pa.n1) + bitrev(u, pa.n1);
}

// ---- one round of butterflies on the 16 registers -------------------------------

// True when bits [lo, hi) of the tile coordinate are the same in every lane
// of a wave64 for each register: register bits, or thread bits mapped to
// thread-id bits >= 6.
constexpr bool wave_uniform_bits(const Geom &g, int lo, int hi) {
    for (int b = lo; b < hi; ++b) {
        bool ok = false;
        for (int r = 0; r < 4; ++r) ok = ok || g.reg_pos[r] == b;
        for (int i = 6; i < g.n_thr; ++i) ok = ok || g.thr_pos[i] == b;
        if (!ok) return false;
    }
    return true;
}

// Bit j set: the twiddle index of stage j of the round is wave-uniform (it
// reads t bits [L - s, L) forward, [u + 1, L) inverse, and the row bits in a
// row pass), so the twiddle is read through the scalar unit.
constexpr u32 uniform_twiddle_mask(const Geom &g, int L, int CL, int K, int s0, bool fwd, bool row) {
    u32 m = 0;
    for (int j = 0; j < K; ++j) {
        const int s = s0 + j;
        const bool t_ok = fwd ? wave_uniform_bits(g, L - s, L) : wave_uniform_bits(g, s + 1, L);
        if (t_ok && (!row || wave_uniform_bits(g, L, L + CL))) m |= 1u << j;
    }
    return m;
}

// Twiddle idx of the table; UNIFORM: idx is the same in every lane.
template <bool UNIFORM>
CK_HD Tw load_twiddle(const Tw *tw, u32 idx) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (UNIFORM) {
        const u32 ui = __builtin_amdgcn_readfirstlane(idx);
        const CK_CONST_AS Tw *ct = const_table(tw);
        return Tw{ct[ui].w, ct[ui].wq};
    }
#endif
    return tw[idx];
}

// Row-pass twiddle of an FP64 prime as the product of the shared entry
// tw[sidx] (sidx = 2^level + j, the column pass's index) and the factor of
// (row, level): fp_mulmod gives it exactly, a signed residue within 0.52 q,
// and wp = fl(w * fl(1/q)) is at most two ulps from w / q, which moves y wp by
// under 2^-7 for |y| < 2^45.  The butterfly products then stay within 0.52 q
// (0.51 q with table twiddles), inside the margins of the lazy ranges above:
// forward |x| < q + 17 * 0.52 q < 10 q, inverse sums below 16 q.
// SU / RU: the shared index / the row is wave-uniform.
template <bool SU, bool RU>
CK_HD Tw row_twiddle_fp(const TileJob &job, u32 sidx, u32 fidx) {
    const Tw ts = load_twiddle<SU>(job.tw, sidx);
    const Tw f = load_twiddle<RU>(job.rowf, fidx);
    const double w = fp_mulmod(u64_as_f64(ts.w), u64_as_f64(f.w), u64_as_f64(f.wq), job.qd);
    return Tw{f64_as_u64(w), f64_as_u64(w * job.qinv)};
}

// Forward round: stages s0..s0+k-1 (butterfly bit of stage s is L-1-s).
// SM: stages whose shared index (row-factored twiddles) is wave-uniform;
// RU: the row of an element is wave-uniform.
template <int L, int CL, int K, bool FP, u32 UM = 0, u32 SM = 0, bool RU = false, bool WIDE = false,
          bool GU = false>
CK_HD void fwd_round(u64 (&v)[16], const Geom &g, u32 tid, int s0, bool row_base,
                     const PassAddr &pa, const TileJob &job) {
    const bool fac = FP && row_base;  // FP64 row passes always read row factors
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int s = s0 + j;
        const int rb = K - 1 - j;  // register bit of this stage
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i & (1 << rb)) continue;
            u32 e = coord(g, tid, i);
            u32 t = e & ((1u << L) - 1);
            Tw w;
            if (fac) {
                const u32 sidx = (1u << s) + (t >> (L - s));
                const u32 fidx = row_u<CL, GU>(pa, e >> L) * L + (u32)s;
                CK_BOUNDS(sidx, 1u << L);
                CK_BOUNDS(fidx, (1u << (pa.logn - L)) * L);
                w = (SM >> j) & 1u ? row_twiddle_fp<true, RU>(job, sidx, fidx) : row_twiddle_fp<false, RU>(job, sidx, fidx);
            } else {
                u32 base = row_base ? row_twiddle_base<L, CL, GU>(pa, e) : 1u;
                u32 idx = (base << s) + (t >> (L - s));
                CK_BOUNDS(idx, 1u << pa.logn);
                w = (UM >> j) & 1u ? load_twiddle<true>(job.tw, idx) : load_twiddle<false>(job.tw, idx);
            }
            if (FP)
                ct_bfly_fp(v[i], v[i | (1 << rb)], w, job.qd);
            else if (WIDE)
                ct_bfly_wide(v[i], v[i | (1 << rb)], w, job.q);
            else
                ct_bfly(v[i], v[i | (1 << rb)], w, job.q);
        }
    }
}

// Inverse round: GS stages u0..u0+k-1 (butterfly bit of stage u is u; its
// twiddle level is L - 1 - u).
template <int L, int CL, int K, bool FP, u32 UM = 0, u32 SM = 0, bool RU = false, bool WIDE = false,
          bool GU = false>
CK_HD void inv_round(u64 (&v)[16], const Geom &g, u32 tid, int u0, bool row_base,
                     const PassAddr &pa, const TileJob &job) {
    const bool fac = FP && row_base;  // FP64 row passes always read row factors
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int u = u0 + j;
        const int rb = j;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            if (i & (1 << rb)) continue;
            u32 e = coord(g, tid, i);
            u32 t = e & ((1u << L) - 1);
            Tw w;
            if (fac) {
                const u32 sidx = (1u << (L - 1 - u)) + (t >> (u + 1));
                const u32 fidx = row_u<CL, GU>(pa, e >> L) * L + (u32)(L - 1 - u);
                CK_BOUNDS(sidx, 1u << L);
                CK_BOUNDS(fidx, (1u << (pa.logn - L)) * L);
                w = (SM >> j) & 1u ? row_twiddle_fp<true, RU>(job, sidx, fidx) : row_twiddle_fp<false, RU>(job, sidx, fidx);
            } else {
                u32 base = row_base ? row_twiddle_base<L, CL, GU>(pa, e) : 1u;
                u32 idx = (base << (L - 1 - u)) + (t >> (u + 1));
                CK_BOUNDS(idx, 1u << pa.logn);
                w = (UM >> j) & 1u ? load_twiddle<true>(job.tw, idx) : load_twiddle<false>(job.tw, idx);
            }
            if (FP && j == K - 1)
                gs_bfly_fp<true>(v[i], v[i | (1 << rb)], w, job.qd, job.qinv);
            else if (FP)
                gs_bfly_fp<false>(v[i], v[i | (1 << rb)], w, job.qd, job.qinv);
            else if (WIDE)
                gs_bfly_wide(v[i], v[i | (1 << rb)], w, job.q);
            else
                gs_bfly(v[i], v[i | (1 << rb)], w, job.q);
        }
    }
}

}  // namespace ck
