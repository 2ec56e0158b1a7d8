// Batched negacyclic NTT / iNTT kernels (gfx950) and their launch.
//
// One launch per pass covers every (polynomial, limb, tile) of up to 16
// batch entries of possibly different degree and level.  Blocks are ordered
// limb-major (limb outer, polynomial inner, tile innermost) so blocks in
// flight share a prime and its twiddle table stays hot in each XCD's L2.
#include "internal.hpp"
#include "ntt_pass.hpp"

namespace ck {

struct DevNttDesc {
    const u64 *src;
    u64 *dst;
    u64 *tmp;
    const DevPrime *primes;
    u64 in_stride, out_stride;
    u32 n_polys, n_limbs, logn, block_start;
};

constexpr int kMaxNttDesc = 16;

struct NttLaunch {
    DevNttDesc d[kMaxNttDesc];
    u32 n;
};

template <bool FWD, bool FIRST>
__global__ __launch_bounds__(256) void k_ntt(NttLaunch launch) {
    __shared__ u64 lds[4096];
    const u32 b = blockIdx.x;
    u32 di = 0;
    while (di + 1 < launch.n && b >= launch.d[di + 1].block_start) ++di;
    const DevNttDesc &D = launch.d[di];
    const u32 tiles = ntt_tiles(D.logn);
    const u32 local = b - D.block_start;
    const u32 tile = local % tiles;
    const u32 lt = local / tiles;
    const u32 limb = lt / D.n_polys, poly = lt % D.n_polys;
    const DevPrime &P = D.primes[limb];
    const size_t N = (size_t)1 << D.logn;

    TileJob job;
    job.tw = FWD ? P.fwd : P.inv;
    job.q = P.q;
    job.ninv = P.ninv;
    job.ninv_q = P.ninv_q;
    job.logn = D.logn;
    job.tile = tile;
    const u64 *src = D.src + poly * D.in_stride + limb * N;
    u64 *dst = D.dst + poly * D.out_stride + limb * N;
    u64 *tmp = D.tmp + ((size_t)poly * D.n_limbs + limb) * N;
    const bool two = ntt_two_pass(D.logn);
    if (FIRST) {
        job.src = src;
        job.dst = two ? tmp : dst;
    } else {
        job.src = tmp;
        job.dst = dst;
    }
    u64 regs[16];
    BlockExec ex{256};
    auto regs_of = [&](u32) -> u64(&)[16] { return regs; };
    ntt_dispatch<FWD, FIRST>(ex, regs_of, job, lds);
}

int launch_ntt(ckks_context *ctx, const NttCall *calls, size_t count, bool forward) {
    if (count == 0) return CKKS_OK;
    if (count > (size_t)kMaxNttDesc) return fail(CKKS_ERR_INVALID, "NTT batch holds at most 16 entries");
    NttLaunch first{}, second{};
    size_t tmp_elems = 0;
    u32 blocks1 = 0, blocks2 = 0;
    for (size_t i = 0; i < count; ++i) {
        const NttCall &c = calls[i];
        int rc = check_plan_limbs(c.plan, c.n_limbs, "NTT");
        if (rc) return rc;
        if (c.plan->ctx != ctx) return fail(CKKS_ERR_INVALID, "NTT batch plans must share the context");
        if (!c.in || !c.out) return fail(CKKS_ERR_INVALID, "null buffer");
        if (c.n_polys == 0) continue;
        if (c.n_polys > 1 && (c.in_stride < c.n_limbs * c.plan->degree ||
                              c.out_stride < c.n_limbs * c.plan->degree))
            return fail(CKKS_ERR_INVALID, "polynomial stride smaller than its limbs");
        if (ntt_two_pass(c.plan->logn)) tmp_elems += c.n_polys * c.n_limbs * c.plan->degree;
    }
    u64 *tmp = nullptr;
    if (tmp_elems) {
        void *ws;
        int rc = ctx->workspace(0, tmp_elems * sizeof(u64), &ws);
        if (rc) return rc;
        tmp = (u64 *)ws;
    }
    size_t tmp_off = 0;
    for (size_t i = 0; i < count; ++i) {
        const NttCall &c = calls[i];
        if (c.n_polys == 0) continue;
        DevNttDesc d{};
        d.src = c.in;
        d.dst = c.out;
        d.primes = c.plan->d_primes;
        d.in_stride = c.n_polys > 1 ? c.in_stride : 0;
        d.out_stride = c.n_polys > 1 ? c.out_stride : 0;
        d.n_polys = (u32)c.n_polys;
        d.n_limbs = (u32)c.n_limbs;
        d.logn = c.plan->logn;
        u32 tiles = ntt_tiles(d.logn);
        u32 nblk = d.n_polys * d.n_limbs * tiles;
        bool two = ntt_two_pass(d.logn);
        d.tmp = two ? tmp + tmp_off : nullptr;
        if (two) tmp_off += c.n_polys * c.n_limbs * c.plan->degree;
        d.block_start = blocks1;
        first.d[first.n++] = d;
        blocks1 += nblk;
        if (two) {
            d.block_start = blocks2;
            second.d[second.n++] = d;
            blocks2 += nblk;
        }
    }
    hipStream_t s = ctx->stream;
    if (forward) {
        if (blocks1) hipLaunchKernelGGL((k_ntt<true, true>), dim3(blocks1), dim3(256), 0, s, first);
        if (blocks2) hipLaunchKernelGGL((k_ntt<true, false>), dim3(blocks2), dim3(256), 0, s, second);
    } else {
        if (blocks1) hipLaunchKernelGGL((k_ntt<false, true>), dim3(blocks1), dim3(256), 0, s, first);
        if (blocks2) hipLaunchKernelGGL((k_ntt<false, false>), dim3(blocks2), dim3(256), 0, s, second);
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

}  // namespace ck

using namespace ck;

extern "C" {

int ckks_ntt_forward(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                     size_t n_polys, size_t n_limbs, size_t poly_stride) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    NttCall c{plan, in, out, n_polys, n_limbs, poly_stride, poly_stride};
    return launch_ntt(plan->ctx, &c, 1, true);
}

int ckks_ntt_inverse(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                     size_t n_polys, size_t n_limbs, size_t poly_stride) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    NttCall c{plan, in, out, n_polys, n_limbs, poly_stride, poly_stride};
    return launch_ntt(plan->ctx, &c, 1, false);
}

static int batch_call(ckks_context *ctx, const ckks_ntt_batch *batch, size_t count, bool fwd) {
    if (!ctx || (!batch && count)) return fail(CKKS_ERR_INVALID, "null argument");
    if (count > (size_t)kMaxNttDesc) return fail(CKKS_ERR_INVALID, "NTT batch holds at most 16 entries");
    NttCall calls[kMaxNttDesc];
    for (size_t i = 0; i < count; ++i)
        calls[i] = NttCall{batch[i].plan, batch[i].in, batch[i].out, batch[i].n_polys,
                           batch[i].n_limbs, batch[i].poly_stride, batch[i].poly_stride};
    return launch_ntt(ctx, calls, count, fwd);
}

int ckks_ntt_forward_batch(ckks_context *ctx, const ckks_ntt_batch *batch, size_t count) {
    return batch_call(ctx, batch, count, true);
}

int ckks_ntt_inverse_batch(ckks_context *ctx, const ckks_ntt_batch *batch, size_t count) {
    return batch_call(ctx, batch, count, false);
}

}  // extern "C"
