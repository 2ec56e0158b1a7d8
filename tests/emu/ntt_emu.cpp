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

template <bool FWD, bool FIRST>
static void run_pass(const TileJob &base, u32 tiles) {
    static u64 regs[256][16];
    static u64 lds[4096];
    for (u32 tile = 0; tile < tiles; ++tile) {
        TileJob job = base;
        job.tile = tile;
        BlockExec ex{256};
        std::memset(lds, 0xA5, sizeof lds);
        auto regs_of = [&](u32 tid) -> u64(&)[16] { return regs[tid]; };
        ntt_dispatch<FWD, FIRST>(ex, regs_of, job, lds);
    }
}

// Transforms one limb of 2^logn residues modulo q (< 2^62).
extern "C" int emu_ntt(const uint64_t *in, uint64_t *out, uint64_t q, uint32_t logn, int forward) {
    if (logn < 10 || logn > 17) return -1;
    uint64_t psi = host::negacyclic_root(q, 1ull << logn);
    if (!psi) return -2;
    std::vector<uint64_t> tw;
    host::twiddle_table(q, forward ? psi : host::invmod(psi, q), logn, tw);
    uint32_t n = 1u << logn;
    std::vector<uint64_t> tmp(n, 0xDEADBEEFull), src(in, in + n);
    TileJob job{};
    job.tw = reinterpret_cast<const Tw *>(tw.data());
    job.q = q;
    job.ninv = host::invmod(n % q, q);
    job.ninv_q = host::shoup_quotient(job.ninv, q);
    job.logn = logn;
    bool two = ntt_two_pass(logn);
    job.src = src.data();
    job.dst = two ? tmp.data() : out;
    if (forward) run_pass<true, true>(job, ntt_tiles(logn));
    else run_pass<false, true>(job, ntt_tiles(logn));
    if (two) {
        job.src = tmp.data();
        job.dst = out;
        if (forward) run_pass<true, false>(job, ntt_tiles(logn));
        else run_pass<false, false>(job, ntt_tiles(logn));
    }
    return 0;
}
