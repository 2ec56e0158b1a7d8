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

template <bool FWD, bool FIRST, bool FP, int TB, bool BLK>
static void run_pass(const TileJob &base, u32 tiles) {
    static u64 regs[1 << (TB - 4)][16];
    static u64 lds[1 << TB];
    for (u32 tile = 0; tile < tiles; ++tile) {
        TileJob job = base;
        job.tile = tile;
        BlockExec ex{1u << (TB - 4)};
        std::memset(lds, 0xA5, sizeof lds);
        auto regs_of = [&](u32 tid) -> u64(&)[16] { return regs[tid]; };
        ntt_dispatch<FWD, FIRST, FP, TB, BLK>(ex, regs_of, job, lds);
    }
}

template <bool FWD, bool FIRST, bool FP>
static void run_variant(const TileJob &job, u32 logn, int tb, int blk) {
    const bool two = ntt_two_pass(logn);
    if (two && tb == 13) {
        if (blk) run_pass<FWD, FIRST, FP, 13, true>(job, ntt_tiles(logn, 13));
        else run_pass<FWD, FIRST, FP, 13, false>(job, ntt_tiles(logn, 13));
    } else {
        if (blk && two) run_pass<FWD, FIRST, FP, 12, true>(job, ntt_tiles(logn));
        else run_pass<FWD, FIRST, FP, 12, false>(job, ntt_tiles(logn));
    }
}

// Transforms one limb of 2^logn residues modulo q (< 2^62) with kernel
// variant (tb = log2 two-pass tile size 12 or 13, blk = blocked inter-pass
// layout).
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
    std::vector<uint64_t> tmp(n, 0xDEADBEEFull), src(in, in + n);
    TileJob job{};
    job.tw = reinterpret_cast<const Tw *>(tw.data());
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
#define CK_RUN(F, S)                                              \
    do {                                                          \
        if (fp) run_variant<F, S, true>(job, logn, tb, blk);      \
        else run_variant<F, S, false>(job, logn, tb, blk);        \
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

extern "C" int emu_ntt(const uint64_t *in, uint64_t *out, uint64_t q, uint32_t logn, int forward) {
    return emu_ntt_variant(in, out, q, logn, forward, 12, 0);
}
