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
    template <class F>
    __device__ __forceinline__ void run(F &&phase) {
        phase((u32)threadIdx.x);
    }
    __device__ __forceinline__ void barrier() { __syncthreads(); }
};
#else
struct BlockExec {
    u32 threads;
    template <class F>
    void run(F &&phase) {
        for (u32 t = 0; t < threads; ++t) phase(t);
    }
    void barrier() {}
};
#endif

template <int L, int CL, PassKind KIND, bool FWD>
struct NttPass {
    typedef TileShape<L, CL> S;
    static constexpr int NR = n_rounds(L);

    // geometry of round i; first/last rounds choose the lane walk that makes
    // their HBM side contiguous.
    static constexpr Geom geom(int i) {
        int k = round_k(L, i);
        int lo = FWD ? (L - round_start(L, i) - k) : round_start(L, i);
        bool col_first = false;
        if (KIND == PASS_COL) {
            col_first = true;
        } else if (KIND == PASS_ROW) {
            // forward: internal side (read) is t-contiguous, natural side
            // (write) col-contiguous; inverse the other way round.
            if (FWD) col_first = (i == NR - 1);
            else col_first = (i == 0);
        } else {
            col_first = false;  // PASS_FULL: single column, lanes walk t
        }
        return make_geom(L, CL, lo, k, col_first);
    }

    template <int I>
    CK_HD static void do_round(u64 (&v)[16], u32 tid, const PassAddr &pa, const Tw *tw, u64 q) {
        constexpr Geom g = geom(I);
        constexpr int K = round_k(L, I);
        constexpr int s0 = round_start(L, I);
        constexpr bool row = (KIND == PASS_ROW);
        if (FWD)
            fwd_round<L, CL, K>(v, g, tid, s0, 1u, row, pa, tw, q);
        else
            inv_round<L, CL, K>(v, g, tid, s0, 1u, row, pa, tw, q);
    }

    // global offset (within the limb) of the element at tile coordinate e on
    // the read side and the write side of this pass.
    CK_HD static u32 read_off(const PassAddr &pa, u32 e) {
        if (KIND == PASS_COL) return col_pass_offset<L, CL>(pa, e);
        if (KIND == PASS_ROW)
            return FWD ? row_internal_offset<L, CL>(pa, e) : row_natural_offset<L, CL>(pa, e);
        return e;  // PASS_FULL reads are staged linearly (see run)
    }
    CK_HD static u32 write_off(const PassAddr &pa, u32 e) {
        if (KIND == PASS_COL) return col_pass_offset<L, CL>(pa, e);
        if (KIND == PASS_ROW)
            return FWD ? row_natural_offset<L, CL>(pa, e) : row_internal_offset<L, CL>(pa, e);
        return e;
    }

    // final pass of the whole transform: canonicalise (and scale by N^-1).
    static constexpr bool FINAL = (KIND == PASS_FULL) || (FWD ? KIND == PASS_ROW : KIND == PASS_COL);

    template <int I>
    CK_HD static void exchange_out(u64 (&v)[16], u32 tid, u64 *lds) {
        constexpr Geom g = geom(I);
#pragma unroll
        for (int i = 0; i < 16; ++i) lds[lds_index<L, CL>(coord(g, tid, i))] = v[i];
    }
    template <int I>
    CK_HD static void exchange_in(u64 (&v)[16], u32 tid, const u64 *lds) {
        constexpr Geom g = geom(I);
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = lds[lds_index<L, CL>(coord(g, tid, i))];
    }

    CK_HD static u64 finish(u64 x, const TileJob &job) {
        u64 q = job.q;
        if (FWD) {
            u64 two_q = q << 1;
            x = x >= two_q ? x - two_q : x;
            return x >= q ? x - q : x;
        }
        x = shoup_lazy(x, job.ninv, job.ninv_q, q);
        return x >= q ? x - q : x;
    }

    template <class Exec, class Regs>
    CK_HD static void run(Exec &ex, Regs &&regs_of, const TileJob &job, u64 *lds) {
        PassAddr pa;
        pa.logn = job.logn;
        pa.n2 = (job.logn + 1) / 2;
        pa.n1 = job.logn - pa.n2;
        pa.tile = job.tile;
        const u32 T = S::THREADS;
        const u32 n = job.logn;

        // ---- load
        if (KIND == PASS_FULL && !FWD) {
            // natural input x goes to internal position bitrev(x)
            ex.run([&](u32 tid) {
                if (tid >= T) return;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u32 x = (u32)i * T + tid;
                    CK_BOUNDS(x, 1u << n);
                    lds[lds_index<L, CL>(bitrev(x, n))] = job.src[x];
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
                    v[i] = job.src[off];
                }
            });
        }

        // ---- rounds
        rounds<0>(ex, regs_of, pa, job, lds);

        // ---- store
        if (KIND == PASS_FULL && FWD) {
            // internal position p holds X_bitrev(p)
            ex.barrier();
            ex.run([&](u32 tid) {
                if (tid >= T) return;
                constexpr Geom g = geom(NR - 1);
                u64(&v)[16] = regs_of(tid);
#pragma unroll
                for (int i = 0; i < 16; ++i)
                    lds[lds_index<L, CL>(bitrev(coord(g, tid, i), n))] = v[i];
            });
            ex.barrier();
            ex.run([&](u32 tid) {
                if (tid >= T) return;
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    u32 x = (u32)i * T + tid;
                    CK_BOUNDS(x, 1u << n);
                    job.dst[x] = finish(lds[lds_index<L, CL>(x)], job);
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
                    job.dst[off] = x;
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
                do_round<I>(regs_of(tid), tid, pa, job.tw, job.q);
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
CK_HD u32 ntt_tiles(u32 logn) { return logn > 12 ? (1u << (logn - 12)) : 1u; }

// Dispatch on log N to the instantiated pass.  first = first pass of the
// transform (forward: column/full, inverse: row/full).
template <bool FWD, bool FIRST, class Exec, class Regs>
CK_HD void ntt_dispatch(Exec &ex, Regs &&regs_of, const TileJob &job, u64 *lds) {
    switch (job.logn) {
#define CK_FULL(LG)                                                   \
    case LG:                                                          \
        if (FIRST) NttPass<LG, 0, PASS_FULL, FWD>::run(ex, regs_of, job, lds); \
        break;
        CK_FULL(10)
        CK_FULL(11)
        CK_FULL(12)
#undef CK_FULL
#define CK_TWO(LG)                                                                          \
    case LG: {                                                                              \
        constexpr int N1 = LG - (LG + 1) / 2, N2 = (LG + 1) / 2;                            \
        if (FWD == FIRST)                                                                   \
            NttPass<N1, 12 - N1, PASS_COL, FWD>::run(ex, regs_of, job, lds);                \
        else                                                                                \
            NttPass<N2, 12 - N2, PASS_ROW, FWD>::run(ex, regs_of, job, lds);                \
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
