// One tile of one NTT pass, written as a sequence of phases separated by
// block barriers.  The GPU kernel runs the phases with __syncthreads()
// between them; the CPU build runs each phase for every thread in turn
// (BlockExec below), which lets the tests check addressing and results of
// the exact kernel code on the host.
#pragma once
#include "ntt_tile.hpp"

namespace ck {

// Executes `phase(tid)` for all threads, then a barrier.
#if defined(__HIP_DEVICE_COMPILE__)
struct BlockExec {
    u32 threads;
    // thread id handed to the phases; a caller that runs several passes in a
    // loop can make it opaque per iteration (see k_ntt_row_mac) so the
    // compiler recomputes the index math instead of keeping it live
    u32 tid = threadIdx.x;
    template <class F>
    __device__ __forceinline__ void run(F &&phase) {
        phase(tid);
    }
    __device__ __forceinline__ void barrier() { __syncthreads(); }
};
#else
struct BlockExec {
    u32 threads;
    u32 tid = 0;  // unused on the host: run() walks every thread
    template <class F>
    void run(F &&phase) {
        for (u32 t = 0; t < threads; ++t) phase(t);
    }
    void barrier() {}
};
#endif

// Input of a pass's first load: the residue at job.src[off], converted for
// the FP64 path when the pass reads canonical residues.  A fused caller
// passes its own functor with the same signature (for instance one that
// computes the value from other operands); it returns the register bits the
// pass works on.
struct PlainSrc {
    CK_HD u64 operator()(const TileJob &job, u32 off, bool fp_initial) const {
        const u64 x = ld_stream(job.src + off);
        return fp_initial ? f64_as_u64((double)x) : x;
    }
};

// BLK: the inter-pass buffer uses the blocked layout (internal_offset).
// WIDE: integer arithmetic on canonical residues for moduli in [2^62, 2^64)
// (the lazy [0, 4q) butterflies need q < 2^62).
// GU: inverse row pass whose input is read through the NTT-domain Galois
// map of job.gk (row_u): the load reads whole source segments and places
// each residue at its permuted position in LDS.
template <int L, int CL, PassKind KIND, bool FWD, bool FP, bool BLK = false, bool WIDE = false, bool GU = false>
struct NttPass {
    static_assert(!GU || (KIND == PASS_ROW && !FWD), "Galois-read rows belong to the inverse row pass");
    typedef TileShape<L, CL> S;
    static constexpr int NR = n_rounds(L);

    // geometry of round i.  Column passes: lanes walk columns (contiguous in
    // HBM, and every column shares its twiddles).  Row passes and the full
    // pass: lanes walk t, so the twiddles a wave loads are contiguous; the
    // natural-order side of a row pass is staged through LDS (see run).
    static constexpr Geom geom(int i) {
        int k = round_k(L, i);
        int lo = FWD ? (L - round_start(L, i) - k) : round_start(L, i);
        return make_geom(L, CL, lo, k, KIND == PASS_COL);
    }

    // Linear staging order for LDS <-> HBM copies: x = i * THREADS + tid.
    // Row passes map it so lanes walk columns (the natural side is
    // contiguous along columns); the full pass maps it to the natural index.
    CK_HD static u32 stage_coord(u32 x) {
        if (KIND == PASS_ROW) return (x >> CL) | ((x & ((1u << CL) - 1)) << L);
        return x;
    }

    template <int I>
    CK_HD static void do_round(u64 (&v)[16], u32 tid, const PassAddr &pa, const TileJob &job) {
        constexpr Geom g = geom(I);
        constexpr int K = round_k(L, I);
        constexpr int s0 = round_start(L, I);
        constexpr bool row = (KIND == PASS_ROW);
        constexpr u32 UM = uniform_twiddle_mask(g, L, CL, K, s0, FWD, row);
        // row-factored twiddles: the shared index reads t bits only
        constexpr u32 SM = uniform_twiddle_mask(g, L, CL, K, s0, FWD, false);
        constexpr bool RU = wave_uniform_bits(g, L, L + CL);
        if (FWD)
            fwd_round<L, CL, K, FP, UM, SM, RU, WIDE, GU>(v, g, tid, s0, row, pa, job);
        else
            inv_round<L, CL, K, FP, UM, SM, RU, WIDE, GU>(v, g, tid, s0, row, pa, job);
    }

    // global offset (within the limb) of the element at tile coordinate e on
    // the read side and the write side of this pass.
    CK_HD static u32 read_off(const PassAddr &pa, u32 e) {
        if (KIND == PASS_COL)
            return FWD ? col_pass_offset<L, CL>(pa, e) : col_internal_offset<L, CL, BLK>(pa, e);
        if (KIND == PASS_ROW)
            return FWD ? row_internal_offset<L, CL, BLK, GU>(pa, e) : row_natural_offset<L, CL>(pa, e);
        return e;  // PASS_FULL reads are staged linearly (see run)
    }
    CK_HD static u32 write_off(const PassAddr &pa, u32 e) {
        if (KIND == PASS_COL)
            return FWD ? col_internal_offset<L, CL, BLK>(pa, e) : col_pass_offset<L, CL>(pa, e);
        if (KIND == PASS_ROW)
            return FWD ? row_natural_offset<L, CL>(pa, e) : row_internal_offset<L, CL, BLK, GU>(pa, e);
        return e;
    }

    // final pass of the whole transform: canonicalise (and scale by N^-1).
    static constexpr bool FINAL = (KIND == PASS_FULL) || (FWD ? KIND == PASS_ROW : KIND == PASS_COL);
    // first pass: reads canonical residues (the FP64 path converts them;
    // between passes the FP64 path keeps its lazy doubles as raw bits).
    static constexpr bool INITIAL = (KIND == PASS_FULL) || (FWD ? KIND == PASS_COL : KIND == PASS_ROW);

    // lds_index is linear: thread part (computed once) ^ register part
    // (a compile-time constant per unrolled i).
    template <int I>
    CK_HD static void exchange_out(u64 (&v)[16], u32 tid, u64 *lds) {
        constexpr Geom g = geom(I);
        const u32 base = lds_index<L, CL>(coord(g, tid, 0));
#pragma unroll
        for (int i = 0; i < 16; ++i) lds[base ^ lds_index<L, CL>(coord(g, 0, i))] = v[i];
    }
    template <int I>
    CK_HD static void exchange_in(u64 (&v)[16], u32 tid, const u64 *lds) {
        constexpr Geom g = geom(I);
        const u32 base = lds_index<L, CL>(coord(g, tid, 0));
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = lds[base ^ lds_index<L, CL>(coord(g, 0, i))];
    }

    CK_HD static u64 finish(u64 x, const TileJob &job) {
        if (FP) {
            double d = u64_as_f64(x);
            if (!FWD) d = fp_mulmod(d, job.ninv_d, job.ninv_p, job.qd);
            return fp_canonical(d, job.qd, job.qinv);
        }
        u64 q = job.q;
        if (WIDE) return FWD ? x : shoup_wide(x, job.ninv, job.ninv_q, q);
        if (FWD) {
            u64 two_q = q << 1;
            x = x >= two_q ? x - two_q : x;
            return x >= q ? x - q : x;
        }
        x = shoup_lazy(x, job.ninv, job.ninv_q, q);
        return x >= q ? x - q : x;
    }

    CK_HD static PassAddr pass_addr(const TileJob &job) {
        PassAddr pa;
        pa.logn = job.logn;
        pa.n2 = (job.logn + 1) / 2;
        pa.n1 = job.logn - pa.n2;
        pa.tile = job.tile;
        pa.gk = job.gk;
        pa.gc = (job.gk - 1) >> 1;
        pa.gkinv = job.gkinv;
        return pa;
    }

    template <class Exec, class Regs, class Src = PlainSrc>
    CK_HD static void run(Exec &ex, Regs &&regs_of, const TileJob &job, u64 *lds, const Src &src = Src{}) {
        run_impl<false>(ex, regs_of, job, lds, src);
    }

    // Forward row pass (and forward full pass) that stops with the tile's
    // lazy results staged in LDS (after a barrier): element x of the linear
    // staging order (x = i * THREADS + tid) sits at lds[lds_index(e)] with
    // e = stage_coord(x) (full pass: e = x), destined for
    // row_natural_offset(pass_addr(job), e) (full pass: x); finish() makes it
    // canonical.  Callers fuse their own epilogue on the NTT output.
    template <class Exec, class Regs>
    CK_HD static void run_to_lds(Exec &ex, Regs &&regs_of, const TileJob &job, u64 *lds) {
        static_assert(FWD && KIND != PASS_COL, "only passes that stage their output");
        run_impl<true>(ex, regs_of, job, lds, PlainSrc{});
    }

    template <bool TO_LDS, class Exec, class Regs, class Src>
    CK_HD static void run_impl(Exec &ex, Regs &&regs_of, const TileJob &job, u64 *lds, const Src &src) {
        const PassAddr pa = pass_addr(job);
        const u32 T = S::THREADS;
        const u32 n = job.logn;

        // ---- load
        // staged: the full inverse pass (natural x -> internal bitrev(x)) and
        // the inverse row pass (natural side, lanes walk columns)
        constexpr bool STAGED_IN = !FWD && (KIND == PASS_FULL || KIND == PASS_ROW);
        constexpr bool STAGED_OUT = FWD && (KIND == PASS_FULL || KIND == PASS_ROW);
        if (STAGED_IN) {
            ex.run([&](u32 tid) {
                if (tid >= T) return;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u32 x = (u32)i * T + tid;
                    u32 e, off, li;
                    if (KIND == PASS_FULL) {
                        e = bitrev(x, n);
                        off = x;
                        li = lds_index<L, CL>(e);
                    } else {
                        e = stage_coord(x);
                        off = row_natural_offset<L, CL>(pa, e);
                        if constexpr (GU) {
                            // off = R 2^n1 + v is source element gk n + gc of
                            // destination n = h 2^n1 + u, u = row_u(col):
                            // gk h + ((gk u + gc) >> n1) = R (mod 2^n2); its
                            // tile coordinate is bitrev_n2(h) | col << L
                            const u32 col = e >> L, m2 = (1u << pa.n2) - 1;
                            const u32 z = mul24(pa.gk, row_u<CL, true>(pa, col)) + pa.gc;
                            const u32 h = mul24(((off >> pa.n1) - (z >> pa.n1)) & m2, pa.gkinv) & m2;
                            CK_BOUNDS(h, 1u << L);
                            li = kGuRow<L, CL>.v[h] ^ lds_swizzle(col);
                        } else {
                            li = lds_index<L, CL>(e);
                        }
                    }
                    CK_BOUNDS(off, 1u << n);
                    CK_BOUNDS(li, 1u << (L + CL));
                    lds[li] = src(job, off, FP && INITIAL);
                }
            });
            ex.barrier();
            ex.run([&](u32 tid) {
                if (tid >= T) return;
                exchange_in<0>(regs_of(tid), tid, lds);
            });
        } else {
            ex.run([&](u32 tid) {
                if (tid >= T) return;
                constexpr Geom g = geom(0);
                u64(&v)[16] = regs_of(tid);
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u32 off = read_off(pa, coord(g, tid, i));
                    CK_BOUNDS(off, 1u << n);
                    v[i] = src(job, off, FP && INITIAL);
                }
            });
        }

        // ---- rounds
        rounds<0>(ex, regs_of, pa, job, lds);

        // ---- store
        if (STAGED_OUT) {
            ex.barrier();
            ex.run([&](u32 tid) {
                if (tid >= T) return;
                constexpr Geom g = geom(NR - 1);
                u64(&v)[16] = regs_of(tid);
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u32 e = coord(g, tid, i);
                    // full pass: internal position p holds X_bitrev(p)
                    if (KIND == PASS_FULL) e = bitrev(e, n);
                    lds[lds_index<L, CL>(e)] = v[i];
                }
            });
            ex.barrier();
            if (TO_LDS) return;
            ex.run([&](u32 tid) {
                if (tid >= T) return;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u32 x = (u32)i * T + tid;
                    u32 e = stage_coord(x);
                    u32 off = KIND == PASS_FULL ? x : row_natural_offset<L, CL>(pa, e);
                    CK_BOUNDS(off, 1u << n);
                    st_stream(job.dst + off, finish(lds[lds_index<L, CL>(e)], job));
                }
            });
        } else {
            ex.run([&](u32 tid) {
                if (tid >= T) return;
                constexpr Geom g = geom(NR - 1);
                u64(&v)[16] = regs_of(tid);
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u64 x = FINAL ? finish(v[i], job) : v[i];
                    u32 off = write_off(pa, coord(g, tid, i));
                    CK_BOUNDS(off, 1u << n);
                    st_stream(job.dst + off, x);
                }
            });
        }
    }

    template <int I, class Exec, class Regs>
    CK_HD static void rounds(Exec &ex, Regs &&regs_of, const PassAddr &pa, const TileJob &job,
                             u64 *lds) {
        if constexpr (I < NR) {
            const u32 T = S::THREADS;
            if constexpr (I > 0) {
                ex.barrier();  // previous readers of lds are done
                ex.run([&](u32 tid) {
                    if (tid >= T) return;
                    exchange_out<I - 1>(regs_of(tid), tid, lds);
                });
                ex.barrier();
                ex.run([&](u32 tid) {
                    if (tid >= T) return;
                    exchange_in<I>(regs_of(tid), tid, lds);
                });
            }
            ex.run([&](u32 tid) {
                if (tid >= T) return;
                do_round<I>(regs_of(tid), tid, pa, job);
            });
            rounds<I + 1>(ex, regs_of, pa, job, lds);
        }
    }
};

// Split of log N into (n1 column stages, n2 row stages); N <= 4096 uses a
// single full pass.
CK_HD u32 ntt_n2(u32 logn) { return (logn + 1) / 2; }
CK_HD u32 ntt_n1(u32 logn) { return logn - ntt_n2(logn); }
CK_HD bool ntt_two_pass(u32 logn) { return logn > 12; }
// tiles per limb transform for two-pass tiles of 2^tb residues
CK_HD u32 ntt_tiles(u32 logn, u32 tb = 12) { return logn > 12 ? (1u << (logn - tb)) : 1u; }

// Dispatch on log N to the instantiated pass.  first = first pass of the
// transform (forward: column/full, inverse: row/full).  TB = log2 of the
// two-pass tile size (12: 256 threads, 13: 512 threads; the full pass for
// N <= 4096 always uses 256 threads); BLK = blocked inter-pass layout.
// GU: the inverse first (row) pass of a two-pass degree reads its input
// through the Galois map of job.gk (NttPass); other passes ignore it.
template <bool FWD, bool FIRST, bool FP, int TB = 12, bool BLK = false, bool WIDE = false, bool GU = false,
          class Exec, class Regs, class Src = PlainSrc>
CK_HD void ntt_dispatch(Exec &ex, Regs &&regs_of, const TileJob &job, u64 *lds, const Src &src = Src{}) {
    switch (job.logn) {
#define CK_FULL(LG)                                                   \
    case LG:                                                          \
        if (FIRST) NttPass<LG, 0, PASS_FULL, FWD, FP, false, WIDE>::run(ex, regs_of, job, lds, src); \
        break;
        CK_FULL(10)
        CK_FULL(11)
        CK_FULL(12)
#undef CK_FULL
#define CK_TWO(LG)                                                                          \
    case LG: {                                                                              \
        constexpr int N1 = LG - (LG + 1) / 2, N2 = (LG + 1) / 2;                            \
        if (FWD == FIRST)                                                                   \
            NttPass<N1, TB - N1, PASS_COL, FWD, FP, BLK, WIDE>::run(ex, regs_of, job, lds, src);  \
        else                                                                                \
            NttPass<N2, TB - N2, PASS_ROW, FWD, FP, BLK, WIDE, GU && !FWD>::run(ex, regs_of, job, lds, src); \
        break;                                                                              \
    }
        CK_TWO(13)
        CK_TWO(14)
        CK_TWO(15)
        CK_TWO(16)
        CK_TWO(17)
#undef CK_TWO
        default:
            break;
    }
}

}  // namespace ck
