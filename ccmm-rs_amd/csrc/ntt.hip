// Batched negacyclic NTT / iNTT kernels (gfx950) and their launch.
//
// A launch covers every (polynomial, limb, tile) of up to 16 batch entries
// of possibly different degree and level.  Blocks are ordered limb-major
// (limb outer, polynomial inner, tile innermost) so the blocks in flight
// share a prime and its twiddle table stays hot in each XCD's L2.
//
// Degrees above 4096 take two passes (column pass, row pass) through an
// inter-pass buffer.  The batch may be cut into chunks whose buffer is
// reused chunk after chunk (ckks_set_ntt_chunk_bytes); by default the whole
// batch is one chunk.
#include <algorithm>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <vector>

#include "host_math.hpp"
#include "internal.hpp"
#include "ntt_pass.hpp"

namespace ck {

struct DevNttDesc {
    const u64 *src;
    u64 *dst;
    u64 *tmp;  // inter-pass buffer of this entry's first limb transform in the chunk
    const DevPrime *primes;
    u64 in_stride, out_stride;
    u32 n_polys, n_limbs, logn, block_start;
    u32 lt_begin;  // first limb transform (limb-major index) covered
    u32 lt_count;  // limb transforms covered (two-pass launches)
    NttMac mac;    // fused key inner product (k_ntt<.., kSrcMac> only)
    u32 galois_k;  // NTT-domain Galois permutation on the load (k_ntt<.., kSrcGalois> only)
    u32 galois_kinv;  // galois_k^-1 mod N
};

constexpr int kMaxNttDesc = 16;

struct NttLaunch {
    DevNttDesc d[kMaxNttDesc];
    u32 n;
};

// First load of the fused inverse NTT: the key inner product of one
// (ciphertext, component, limb) at limb offset off.  FP64 limbs sum D terms
// fp_mulmod(digit, key) (|sum| <= 0.51 D q < 2^44 for D <= 8) and reduce the
// sum into 0.51 q, the input range of the inverse FP64 butterflies; integer
// limbs accumulate 128-bit products and reduce once (D q^2 < 2^64 q for
// q < 2^61).  Digits and key entries are canonical residues.
// A correction table c (hoisted rotations: canonical residues) joins the
// sum as one more term.
struct MacSrc {
    const u64 *d;  // digit 0 of this (ct, limb); digit j at d + j * d_stride
    const u64 *k;  // key entry 0 of this (comp, limb); entry j at k + j * k_stride
    const u64 *c;  // correction of this (comp, limb) or null
    u64 d_stride, k_stride;
    u32 dnum;
    Barrett br;
    double qd, qinv;
    __device__ __forceinline__ u64 operator()(const TileJob &, u32 off, bool fp) const {
        // every load of the element issued before the first use
        u64 dv[8], kv[8];
        const u64 cv = c ? c[off] : 0;
#pragma unroll
        for (u32 j = 0; j < 8; ++j)
            if (j < dnum) {
                dv[j] = d[j * d_stride + off];
                kv[j] = k[j * k_stride + off];
            }
        if (fp) {
            double acc = small_to_f64(cv);
#pragma unroll
            for (u32 j = 0; j < 8; ++j)
                if (j < dnum) {
                    const double w = small_to_f64(kv[j]);
                    acc += fp_mulmod(small_to_f64(dv[j]), w, w * qinv, qd);
                }
            return f64_as_u64(fp_reduce(acc, qd, qinv));
        }
        Acc128 t;
        t.zero();
        t.mac(cv, 1);
#pragma unroll
        for (u32 j = 0; j < 8; ++j)
            if (j < dnum) t.mac(dv[j], kv[j]);
        return t.reduce(br);
    }
};

// Input of the first inverse pass read through the NTT-domain Galois
// permutation (ring/ntt_galois_permutation.rs:5-19): element n of the limb is
// in[((k (2n + 1)) mod 2N - 1) / 2], k odd; 2N divides 2^32, so the 32-bit
// product may wrap.  Single-pass degrees gather with it; two-pass degrees
// read whole source segments and permute in LDS (row_u).
struct GaloisSrc {
    u32 k;
    __device__ __forceinline__ u64 operator()(const TileJob &job, u32 off, bool fp_initial) const {
        const u32 m = (((2u * off + 1u) * k) & ((2u << job.logn) - 1u)) >> 1;
        const u64 x = job.src[m];
        return fp_initial ? f64_as_u64((double)x) : x;
    }
};

// kSrcMacGalois: the key inner product read through the Galois permutation
// (two-pass degrees).  Its blocks run tile-major, limb transform fastest,
// remapped so each XCD runs a contiguous range: the two components of one
// ciphertext (adjacent limb transforms) read the same digit segments, and
// every ciphertext of one limb the same key and correction segments, from
// one L2.
enum { kSrcPlain = 0, kSrcMac = 1, kSrcGalois = 2, kSrcMacGalois = 3 };

// TB = log2 of the two-pass tile (12: 256 threads, 13: 512 threads);
// BLK = blocked inter-pass layout.  Single-pass launches use <.., 12, false>.
// SRC: the first inverse pass may read the descriptor's key inner product
// (kSrcMac) or its input permuted by a Galois exponent (kSrcGalois).
template <bool FWD, bool FIRST, int TB, bool BLK, int SRC = kSrcPlain>
__global__ __launch_bounds__(1 << (TB - 4)) void k_ntt(NttLaunch launch) {
    __shared__ u64 lds[1 << TB];
    const u32 b = SRC == kSrcMacGalois ? xcd_contiguous_block(blockIdx.x, gridDim.x) : blockIdx.x;
    u32 di = 0;
    while (di + 1 < launch.n && b >= launch.d[di + 1].block_start) ++di;
    const DevNttDesc &D = launch.d[di];
    const u32 tiles = ntt_tiles(D.logn, TB);
    const u32 local = b - D.block_start;
    const u32 tile = SRC == kSrcMacGalois ? local / D.lt_count : local % tiles;
    const u32 rel = SRC == kSrcMacGalois ? local % D.lt_count : local / tiles;
    const u32 lt = D.lt_begin + rel;
    const u32 limb = lt / D.n_polys, poly = lt % D.n_polys;
    const DevPrime &P = D.primes[limb];
    const size_t N = (size_t)1 << D.logn;

    TileJob job{};
    job.tw = FWD ? P.fwd : P.inv;
    job.rowf = FWD ? P.rowf : P.rowf_inv;
    job.q = P.q;
    job.ninv = P.ninv;
    job.ninv_q = P.ninv_q;
    job.qd = P.qd;
    job.qinv = P.qinv;
    job.ninv_d = P.ninv_d;
    job.ninv_p = P.ninv_p;
    job.logn = D.logn;
    job.tile = tile;
    const u64 *src = D.src + poly * D.in_stride + limb * N;
    u64 *dst = D.dst + poly * D.out_stride + limb * N;
    u64 *tmp = D.tmp + (size_t)rel * N;
    const bool two = ntt_two_pass(D.logn);
    if (FIRST) {
        job.src = src;
        job.dst = two ? tmp : dst;
    } else {
        job.src = tmp;
        job.dst = dst;
    }
    u64 regs[16];
    BlockExec ex{1u << (TB - 4)};
    auto regs_of = [&](u32) -> u64(&)[16] { return regs; };
    static_assert(SRC == kSrcPlain || (!FWD && FIRST), "load variants feed the first inverse pass");
    if (SRC == kSrcGalois && two) {
        job.gk = D.galois_k;
        job.gkinv = D.galois_kinv;
        if (P.fp)
            ntt_dispatch<FWD, FIRST, true, TB, BLK, false, true>(ex, regs_of, job, lds);
        else
            ntt_dispatch<FWD, FIRST, false, TB, BLK, false, true>(ex, regs_of, job, lds);
        return;
    }
    if (SRC == kSrcGalois) {
        const GaloisSrc gs{D.galois_k};
        if (P.fp)
            ntt_dispatch<FWD, FIRST, true, TB, BLK>(ex, regs_of, job, lds, gs);
        else
            ntt_dispatch<FWD, FIRST, false, TB, BLK>(ex, regs_of, job, lds, gs);
        return;
    }
    if (SRC == kSrcMac || SRC == kSrcMacGalois) {
        const u32 ct = poly >> 1, comp = poly & 1;
        MacSrc ms;
        ms.d_stride = (u64)D.n_limbs * N;
        ms.k_stride = (u64)D.n_limbs * N;
        ms.d = D.mac.digits + (u64)ct * D.mac.dnum * ms.d_stride + (u64)limb * N;
        ms.k = (comp ? D.mac.ka : D.mac.kb) + (u64)limb * N;
        const u64 *corr = comp ? D.mac.ta : D.mac.tb;
        ms.c = corr ? corr + (u64)limb * N : nullptr;
        ms.dnum = D.mac.dnum;
        ms.br = Barrett{P.q, P.br_hi, P.br_lo};
        ms.qd = P.qd;
        ms.qinv = P.qinv;
        if (SRC == kSrcMacGalois) {
            job.gk = D.galois_k;
            job.gkinv = D.galois_kinv;
            if (P.fp)
                ntt_dispatch<FWD, FIRST, true, TB, BLK, false, true>(ex, regs_of, job, lds, ms);
            else
                ntt_dispatch<FWD, FIRST, false, TB, BLK, false, true>(ex, regs_of, job, lds, ms);
            return;
        }
        if (P.fp)
            ntt_dispatch<FWD, FIRST, true, TB, BLK>(ex, regs_of, job, lds, ms);
        else
            ntt_dispatch<FWD, FIRST, false, TB, BLK>(ex, regs_of, job, lds, ms);
        return;
    }
    if (P.fp)
        ntt_dispatch<FWD, FIRST, true, TB, BLK>(ex, regs_of, job, lds);
    else
        ntt_dispatch<FWD, FIRST, false, TB, BLK>(ex, regs_of, job, lds);
}

template <int TB, bool BLK>
static void launch_variant(hipStream_t s, bool forward, bool first, const NttLaunch &l, u32 blocks, int src) {
    const dim3 g(blocks), t(1u << (TB - 4));
    if (forward) {
        if (first) hipLaunchKernelGGL((k_ntt<true, true, TB, BLK>), g, t, 0, s, l);
        else hipLaunchKernelGGL((k_ntt<true, false, TB, BLK>), g, t, 0, s, l);
    } else {
        if (first && src == kSrcMac) hipLaunchKernelGGL((k_ntt<false, true, TB, BLK, kSrcMac>), g, t, 0, s, l);
        else if (first && src == kSrcMacGalois)
            hipLaunchKernelGGL((k_ntt<false, true, TB, BLK, kSrcMacGalois>), g, t, 0, s, l);
        else if (first && src == kSrcGalois)
            hipLaunchKernelGGL((k_ntt<false, true, TB, BLK, kSrcGalois>), g, t, 0, s, l);
        else if (first) hipLaunchKernelGGL((k_ntt<false, true, TB, BLK>), g, t, 0, s, l);
        else hipLaunchKernelGGL((k_ntt<false, false, TB, BLK>), g, t, 0, s, l);
    }
}

static void launch_pass(hipStream_t s, bool forward, bool first, const NttLaunch &l, u32 blocks,
                        int variant = 0, int src = kSrcPlain) {
    if (!blocks) return;
    switch (variant) {
    case CKKS_NTT_VARIANT_BLOCKED: launch_variant<12, true>(s, forward, first, l, blocks, src); break;
    case CKKS_NTT_VARIANT_TILE8K: launch_variant<13, false>(s, forward, first, l, blocks, src); break;
    case CKKS_NTT_VARIANT_TILE8K | CKKS_NTT_VARIANT_BLOCKED:
        launch_variant<13, true>(s, forward, first, l, blocks, src);
        break;
    default: launch_variant<12, false>(s, forward, first, l, blocks, src); break;
    }
}

int ntt_variant_for(const ckks_context *ctx, u32 logn) {
    if (!ntt_two_pass(logn) || logn >= 32) return 0;
    int v = ctx->ntt_variant[logn];
    if ((v & CKKS_NTT_VARIANT_TILE8K) && logn < 14) v &= ~CKKS_NTT_VARIANT_TILE8K;
    return v;
}


// The inter-pass buffer of ckks_ntt_forward_noalloc: a fixed array each
// device holds from the library's load to the process end, shared by every
// call on that device (calls that overlap on one device race).  Host code
// uses the address ntt_scratch_lookup returns, never the symbol itself.
constexpr size_t kNttScratchBytes = 32u << 20;  // 64 limb transforms at N = 2^16
__device__ __align__(256) u64 scratch[kNttScratchBytes / sizeof(u64)];

int ntt_scratch_lookup(int device, u64 **out) {
    static std::mutex mu;
    static std::map<int, u64 *> by_device;
    std::lock_guard<std::mutex> lock(mu);
    auto it = by_device.find(device);
    if (it == by_device.end()) {
        // the caller made `device` current
        void *p = nullptr;
        CK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        it = by_device.emplace(device, (u64 *)p).first;
    }
    *out = it->second;
    return CKKS_OK;
}

int launch_ntt(ckks_context *ctx, const NttCall *calls, size_t count, bool forward, u64 *ext_scratch,
               size_t scratch_elems) {
    if (count == 0) return CKKS_OK;
    if (count > (size_t)kMaxNttDesc) return fail(CKKS_ERR_INVALID, "NTT batch holds at most 16 entries");
    size_t total_two = 0, max_n = 0;  // residues of all two-pass limb transforms
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
        if (c.mac && (forward || (count != 1 && !c.galois_k) || !ntt_two_pass(c.plan->logn) || c.n_polys % 2 ||
                      c.mac->dnum == 0 || c.mac->dnum > 8))
            return fail(CKKS_ERR_INVALID,
                        "fused key inner product: inverse two-pass calls (several only Galois-read), 1..8 digits");
        if ((c.galois_k != 0) != (calls[0].galois_k != 0) || (c.mac != nullptr) != (calls[0].mac != nullptr))
            return fail(CKKS_ERR_INVALID, "Galois-permuted loads and key products: every entry of a launch or none");
        if (c.galois_k && (forward || (c.galois_k & 1) == 0 || c.galois_k >= 2 * (u64)c.plan->degree ||
                           (!c.mac && c.in == c.out)))
            return fail(CKKS_ERR_INVALID, "Galois-permuted load: inverse only, odd exponent below 2N, in != out");
        if (ntt_two_pass(c.plan->logn)) {
            total_two += c.n_polys * c.n_limbs * c.plan->degree;
            max_n = std::max(max_n, c.plan->degree);
        }
    }
    {
        ckks_execution_profile &pc = ctx->counts();
        for (size_t i = 0; i < count; ++i) {
            (forward ? pc.ntt_forwards : pc.ntt_inverses) += calls[i].n_polys;
            pc.ntt_limb_transforms += calls[i].n_polys * calls[i].n_limbs;
        }
    }
    hipStream_t s = ctx->stream;
    auto base_desc = [&](const NttCall &c) {
        DevNttDesc d{};
        d.src = c.in;
        d.dst = c.out;
        d.primes = c.plan->d_primes;
        d.in_stride = c.n_polys > 1 ? c.in_stride : 0;
        d.out_stride = c.n_polys > 1 ? c.out_stride : 0;
        d.n_polys = (u32)c.n_polys;
        d.n_limbs = (u32)c.n_limbs;
        d.logn = c.plan->logn;
        if (c.mac) d.mac = *c.mac;
        d.galois_k = (u32)c.galois_k;
        if (c.galois_k) {
            // inverse modulo 2^32 by Newton steps (each doubles the correct
            // low bits; k k == 1 mod 8 seeds three), reduced mod N
            u32 inv = d.galois_k;
            for (int it = 0; it < 4; ++it) inv *= 2u - d.galois_k * inv;
            d.galois_kinv = inv & (u32)(c.plan->degree - 1);
        }
        return d;
    };
    const int src = calls[0].mac ? (calls[0].galois_k ? kSrcMacGalois : kSrcMac)
                                 : calls[0].galois_k ? kSrcGalois : kSrcPlain;
    auto two_pass = [&](size_t i) { return calls[i].n_polys != 0 && ntt_two_pass(calls[i].plan->logn); };
    // single-pass entries (N <= 4096): one launch
    {
        NttLaunch l{};
        u32 blocks = 0;
        for (size_t i = 0; i < count; ++i) {
            const NttCall &c = calls[i];
            if (c.n_polys == 0 || two_pass(i)) continue;
            DevNttDesc d = base_desc(c);
            d.block_start = blocks;
            l.d[l.n++] = d;
            blocks += d.n_polys * d.n_limbs;
        }
        launch_pass(s, forward, true, l, blocks, 0, src);
    }
    if (total_two) {
        size_t cap = ctx->ntt_chunk_bytes / sizeof(u64);
        if (cap == 0 || cap > total_two) cap = total_two;
        if (cap < max_n) cap = max_n;
        u64 *tmp;
        if (ext_scratch) {
            if (scratch_elems < max_n) return fail(CKKS_ERR_INVALID, "one limb transform exceeds the scratch");
            cap = std::min(cap, scratch_elems);
            tmp = ext_scratch;
        } else {
            void *ws;
            int rc = ctx->workspace(0, cap * sizeof(u64), &ws);
            if (rc) return rc;
            tmp = (u64 *)ws;
        }
        // The variant is chosen per degree; in a batch of several degrees the
        // entry with the most residues lends its variant to every entry it
        // applies to (8192-residue tiles need N >= 2^14), so the batch runs
        // as one launch per pass instead of a small trailing launch per
        // variant.  One sweep per variant left; each sweep walks the limb
        // transforms of its entries in order, closing a chunk when the
        // inter-pass buffer is full.
        int vv[kMaxNttDesc] = {};
        size_t dom = count, dom_res = 0;
        for (size_t i = 0; i < count; ++i) {
            if (!two_pass(i)) continue;
            vv[i] = ntt_variant_for(ctx, calls[i].plan->logn);
            const size_t res = calls[i].n_polys * calls[i].n_limbs * calls[i].plan->degree;
            if (res > dom_res) dom = i, dom_res = res;
        }
        for (size_t i = 0; i < count && dom < count; ++i)
            if (two_pass(i) && (!(vv[dom] & CKKS_NTT_VARIANT_TILE8K) || calls[i].plan->logn >= 14)) vv[i] = vv[dom];
        bool seen[CKKS_NTT_VARIANT_COUNT] = {};
        for (size_t i = 0; i < count; ++i)
            if (two_pass(i)) seen[vv[i]] = true;
        for (int var = 0; var < CKKS_NTT_VARIANT_COUNT; ++var) {
        if (!seen[var]) continue;
        const u32 tb = (var & CKKS_NTT_VARIANT_TILE8K) ? 13 : 12;
        auto in_sweep = [&](size_t i) { return two_pass(i) && vv[i] == var; };
        size_t ci = 0, lt = 0;
        while (ci < count) {
            NttLaunch l{};
            u32 blocks = 0;
            size_t used = 0;
            while (ci < count && l.n < (u32)kMaxNttDesc) {
                if (!in_sweep(ci)) {
                    ++ci;
                    lt = 0;
                    continue;
                }
                const NttCall &c = calls[ci];
                const size_t total_lt = c.n_polys * c.n_limbs, N = c.plan->degree;
                const size_t room = (cap - used) / N;
                if (room == 0) break;
                const size_t take = std::min(room, total_lt - lt);
                DevNttDesc d = base_desc(c);
                d.block_start = blocks;
                d.lt_begin = (u32)lt;
                d.lt_count = (u32)take;
                d.tmp = tmp + used;
                l.d[l.n++] = d;
                blocks += (u32)(take * ntt_tiles(d.logn, tb));
                used += take * N;
                lt += take;
                if (lt >= total_lt) {
                    ++ci;
                    lt = 0;
                }
            }
            launch_pass(s, forward, true, l, blocks, var, src);
            launch_pass(s, forward, false, l, blocks, var);
        }
        }
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- forward NTT fused with the key-switch inner product -----------------------
//
// The digits of a key switch ([ct][dnum][B][N], coefficient domain) go
// through the column pass as usual; the row pass of every digit j of one
// (ciphertext, limb, tile) then runs in the same block, and instead of
// storing the NTT-domain digit it multiplies it with the key entries
// (kb_j, ka_j) at the same positions and accumulates both products in
// registers.  Only the two accumulators reach HBM: the NTT-domain digits
// are never written or read back.
//
// Block order: (limb, tile) outer, ciphertext inner, remapped so that each
// XCD (blocks are dispatched round-robin over the 8 XCDs) runs a contiguous
// range: the blocks sharing one key tile meet in the same L2.

struct RowMacArgs {
    const u64 *tmp;      // column-pass output, limb-major limb transforms
    const u64 *kb, *ka;  // [dnum][B][N] NTT domain
    u64 *acc;            // [ct][2][B][N]
    const DevPrime *primes;
    u32 n_cts, dnum, limbs, logn;
    u32 xcd_order;  // 1: remap blocks so each XCD runs a contiguous range
};

// Reloads the thread id as a value the compiler cannot see through, so index
// math derived from it is recomputed where it is used.
__device__ __forceinline__ void opaque_tid(BlockExec &ex) {
    ex.tid = threadIdx.x;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(ex.tid));
#endif
}

// FP64 limbs multiply the row pass's lazy doubles (|x| < 10 q < 2^45, the
// input range of fp_mulmod) by the key entries without making them
// canonical first; a key entry w enters with wp = fl(fl(w) * fl(1/q)), at
// most two ulps from w / q, which moves y wp by under 2^-6 for |y| < 2^45,
// so each term stays an exact integer within 0.52 q; the sums are reduced
// into half a q after every 8 digits, so they stay below 2^45 (fp_reduce's
// and fp_residue's range) for any digit count.  Integer limbs make the digit
// canonical and multiply by Barrett reduction.
template <int TB, bool BLK, bool FP, int LG>
__device__ void row_mac_body(const RowMacArgs &a, const DevPrime &P, u32 ct, u32 tile, u32 i,
                             u64 *lds) {
    constexpr int N1 = LG - (LG + 1) / 2, N2 = (LG + 1) / 2;
    typedef NttPass<N2, TB - N2, PASS_ROW, true, FP, BLK> Pass;
    constexpr u32 T = 1u << (TB - 4);
    (void)N1;
    const u64 N = (u64)1 << LG;
    TileJob job{};
    job.tw = P.fwd;
    job.rowf = P.rowf;
    job.q = P.q;
    job.ninv = P.ninv;
    job.ninv_q = P.ninv_q;
    job.qd = P.qd;
    job.qinv = P.qinv;
    job.ninv_d = P.ninv_d;
    job.ninv_p = P.ninv_p;
    job.logn = LG;
    job.tile = tile;
    job.dst = nullptr;
    const PassAddr pa = Pass::pass_addr(job);
    const Barrett br{P.q, P.br_hi, P.br_lo};
    u64 regs[16];
    BlockExec ex{T};
    auto off = [&](int r) { return row_natural_offset<N2, TB - N2>(pa, Pass::stage_coord((u32)r * T + ex.tid)); };
    u64 sb[16], sa[16];  // FP: the sums as double bits
#pragma unroll
    for (int r = 0; r < 16; ++r) sb[r] = sa[r] = 0;
    auto regs_of = [&](u32) -> u64(&)[16] { return regs; };
    const u64 polys = (u64)a.n_cts * a.dnum;
#pragma unroll 1
    for (u32 j = 0; j < a.dnum; ++j) {
        job.src = a.tmp + ((u64)i * polys + (u64)ct * a.dnum + j) * N;
        // opaque per digit: stops the index math of 16 elements x every
        // butterfly from being hoisted out of the loop (and kept in VGPRs)
        opaque_tid(ex);
        Pass::run_to_lds(ex, regs_of, job, lds);
        const u64 *kb = a.kb + ((u64)j * a.limbs + i) * N;
        const u64 *ka = a.ka + ((u64)j * a.limbs + i) * N;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const u32 e = Pass::stage_coord((u32)r * T + ex.tid);
            const u64 x = lds[lds_index<N2, TB - N2>(e)];
            const u32 o = off(r);
            if (FP) {
                const double y = u64_as_f64(x);
                const double wb = small_to_f64(kb[o]), wa = small_to_f64(ka[o]);
                sb[r] = f64_as_u64(u64_as_f64(sb[r]) + fp_mulmod(y, wb, wb * P.qinv, P.qd));
                sa[r] = f64_as_u64(u64_as_f64(sa[r]) + fp_mulmod(y, wa, wa * P.qinv, P.qd));
            } else {
                const u64 d = Pass::finish(x, job);
                sb[r] = add_mod(sb[r], mul_mod(d, kb[o], br), P.q);
                sa[r] = add_mod(sa[r], mul_mod(d, ka[o], br), P.q);
            }
            // keep the scheduler from hoisting all 48 loads of the epilogue
            if ((r & 3) == 3) __builtin_amdgcn_sched_barrier(0);
        }
        if (FP && (j & 7) == 7) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                sb[r] = f64_as_u64(fp_reduce(u64_as_f64(sb[r]), P.qd, P.qinv));
                sa[r] = f64_as_u64(fp_reduce(u64_as_f64(sa[r]), P.qd, P.qinv));
            }
        }
        __syncthreads();  // the next digit's pass reuses lds
    }
    opaque_tid(ex);
    u64 *ob = a.acc + ((u64)ct * 2 * a.limbs + i) * N;
    u64 *oa = ob + (u64)a.limbs * N;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const u32 o = off(r);
        if (FP) {
            st_stream(ob + o, fp_residue(u64_as_f64(sb[r]), P.qd, P.qinv));
            st_stream(oa + o, fp_residue(u64_as_f64(sa[r]), P.qd, P.qinv));
        } else {
            st_stream(ob + o, sb[r]);
            st_stream(oa + o, sa[r]);
        }
    }
}

// At most 128 VGPRs (4 waves per SIMD): the unbounded build takes 144
// (occupancy 3) and the bound costs no spills.
template <int TB, bool BLK, int LG>
__global__ __launch_bounds__(1 << (TB - 4), 4) void k_ntt_row_mac(RowMacArgs a) {
    __shared__ u64 lds[1 << TB];
    const u32 tiles = 1u << (LG - TB);
    const u32 lb = a.xcd_order ? xcd_contiguous_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 ct = lb % a.n_cts;
    const u32 rest = lb / a.n_cts;
    const u32 tile = rest % tiles, i = rest / tiles;
    const DevPrime &P = a.primes[i];
    if (P.fp) row_mac_body<TB, BLK, true, LG>(a, P, ct, tile, i, lds);
    else row_mac_body<TB, BLK, false, LG>(a, P, ct, tile, i, lds);
}

template <int TB, bool BLK>
static int launch_row_mac(const RowMacArgs &a, dim3 g, hipStream_t s) {
    const dim3 t(1u << (TB - 4));
    switch (a.logn) {
#define CK_RM(LG) \
    case LG: hipLaunchKernelGGL((k_ntt_row_mac<TB, BLK, LG>), g, t, 0, s, a); return CKKS_OK;
        CK_RM(13) CK_RM(14) CK_RM(15) CK_RM(16) CK_RM(17)
#undef CK_RM
        default: break;
    }
    return fail(CKKS_ERR_UNSUPPORTED, "fused NTT-MAC: unsupported degree");
}

int launch_ntt_mac(ckks_context *ctx, const ckks_rns_plan *plan, const u64 *digits, size_t n_cts,
                   size_t dnum, size_t limbs, const u64 *kb, const u64 *ka, u64 *acc,
                   bool xcd_order) {
    const u32 logn = plan->logn;
    if (!ntt_two_pass(logn) || logn > 17) return fail(CKKS_ERR_UNSUPPORTED, "fused NTT-MAC needs 2^13 <= N <= 2^17");
    // 4096-residue tiles: the 8192-residue fused kernel needs more registers
    // than a CU can give two waves per SIMD without spilling
    const int var = ntt_variant_for(ctx, logn) & CKKS_NTT_VARIANT_BLOCKED;
    const u32 tb = 12;
    const size_t N = plan->degree, polys = n_cts * dnum, lts = polys * limbs;
    const size_t tiles = ntt_tiles(logn, tb);
    if (lts * tiles >= (1ull << 31) || n_cts * limbs * tiles >= (1ull << 31))
        return fail(CKKS_ERR_INVALID, "key-switch batch too large for one launch");
    void *ws;
    int rc = ctx->workspace(0, lts * N * sizeof(u64), &ws);
    if (rc) return rc;
    ctx->counts().ntt_forwards += polys;
    ctx->counts().ntt_limb_transforms += polys * limbs;
    // column pass of every digit limb into the inter-pass buffer
    NttLaunch l{};
    DevNttDesc &d = l.d[0];
    d.src = digits;
    d.dst = nullptr;
    d.tmp = (u64 *)ws;
    d.primes = plan->d_primes;
    d.in_stride = limbs * N;
    d.out_stride = limbs * N;
    d.n_polys = (u32)polys;
    d.n_limbs = (u32)limbs;
    d.logn = logn;
    d.block_start = 0;
    d.lt_begin = 0;
    l.n = 1;
    launch_pass(ctx->stream, true, true, l, (u32)(lts * tiles), var);
    RowMacArgs a{(const u64 *)ws, kb, ka, acc, plan->d_primes, (u32)n_cts, (u32)dnum, (u32)limbs, logn,
                 xcd_order ? 1u : 0u};
    const dim3 g((u32)(n_cts * limbs * tiles));
    rc = var ? launch_row_mac<12, true>(a, g, ctx->stream) : launch_row_mac<12, false>(a, g, ctx->stream);
    if (rc) return rc;
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- single-modulus NTT for primes up to 64 bits ----------------------------------
//
// The tile passes of the RNS kernels with canonical-residue butterflies
// (NttPass<..., WIDE>): one block per (polynomial, tile), 4096-residue tiles,
// the plain inter-pass layout.

struct NttWideArgs {
    const u64 *src;
    u64 *dst;
    u64 *tmp;  // inter-pass buffer, one polynomial after another
    DevPrime p;
    u32 logn;
};

template <bool FWD, bool FIRST>
__global__ __launch_bounds__(256) void k_ntt_wide(NttWideArgs a) {
    __shared__ u64 lds[1 << 12];
    const u32 tiles = ntt_tiles(a.logn);
    const u32 poly = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    const size_t N = (size_t)1 << a.logn;
    TileJob job{};
    job.tw = FWD ? a.p.fwd : a.p.inv;
    job.rowf = nullptr;
    job.q = a.p.q;
    job.ninv = a.p.ninv;
    job.ninv_q = a.p.ninv_q;
    job.qd = job.qinv = job.ninv_d = job.ninv_p = 0.0;
    job.logn = a.logn;
    job.tile = tile;
    const bool two = ntt_two_pass(a.logn);
    job.src = FIRST ? a.src + poly * N : a.tmp + poly * N;
    job.dst = FIRST && two ? a.tmp + poly * N : a.dst + poly * N;
    u64 regs[16];
    BlockExec ex{256};
    auto regs_of = [&](u32) -> u64(&)[16] { return regs; };
    ntt_dispatch<FWD, FIRST, false, 12, false, true>(ex, regs_of, job, lds);
}

static int ntt64_run(const ckks_ntt64_plan *plan, const u64 *in, u64 *out, size_t n_polys, bool fwd) {
    if (!plan || !in || !out) return fail(CKKS_ERR_INVALID, "null argument");
    if (n_polys == 0) return CKKS_OK;
    ckks_context *ctx = plan->ctx;
    const size_t N = plan->degree;
    const u32 tiles = ntt_tiles(plan->logn);
    if (n_polys * tiles >= (1ull << 31)) return fail(CKKS_ERR_INVALID, "NTT batch too large for one launch");
    NttWideArgs a{in, out, nullptr, plan->prime, plan->logn};
    const bool two = ntt_two_pass(plan->logn);
    if (two) {
        void *ws;
        int rc = ctx->workspace(0, n_polys * N * sizeof(u64), &ws);
        if (rc) return rc;
        a.tmp = (u64 *)ws;
    }
    ckks_execution_profile &pc = ctx->counts();
    (fwd ? pc.ntt_forwards : pc.ntt_inverses) += n_polys;
    pc.ntt_limb_transforms += n_polys;
    const dim3 g((u32)(n_polys * tiles)), t(256);
    if (fwd) {
        hipLaunchKernelGGL((k_ntt_wide<true, true>), g, t, 0, ctx->stream, a);
        if (two) hipLaunchKernelGGL((k_ntt_wide<true, false>), g, t, 0, ctx->stream, a);
    } else {
        hipLaunchKernelGGL((k_ntt_wide<false, true>), g, t, 0, ctx->stream, a);
        if (two) hipLaunchKernelGGL((k_ntt_wide<false, false>), g, t, 0, ctx->stream, a);
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

}  // namespace ck

using namespace ck;

extern "C" {

int ckks_ntt64_plan_create(ckks_context *ctx, uint64_t q, size_t degree, ckks_ntt64_plan **out) {
    if (!ctx || !out) return fail(CKKS_ERR_INVALID, "null argument");
    if (degree == 0 || (degree & (degree - 1))) return fail(CKKS_ERR_INVALID, "NTT degree must be a positive power of two");
    uint32_t logn = 0;
    while (((size_t)1 << logn) < degree) ++logn;
    if (logn < 10 || logn > 17) return fail(CKKS_ERR_UNSUPPORTED, "GPU NTT supports degrees 2^10 .. 2^17");
    if (q < 3 || (q - 1) % (2 * (uint64_t)degree) != 0) return fail(CKKS_ERR_INVALID, "2N must divide q - 1");
    if (!host::is_prime(q)) return fail(CKKS_ERR_INVALID, "NTT modulus must be prime");
    const uint64_t psi = host::negacyclic_root(q, degree);
    if (!psi) return fail(CKKS_ERR_INVALID, "no primitive 2N-th root found");
    CK_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<ckks_ntt64_plan> plan(new ckks_ntt64_plan());
    plan->ctx = ctx;
    plan->q = q;
    plan->psi = psi;
    plan->logn = logn;
    plan->degree = degree;
    std::vector<uint64_t> fwd, inv;
    host::twiddle_table(q, psi, logn, fwd);
    host::twiddle_table(q, host::invmod(psi, q), logn, inv);
    const size_t bytes = fwd.size() * sizeof(uint64_t);
    CK_HIP(hipMalloc(&plan->d_fwd, bytes));
    if (hipMalloc(&plan->d_inv, bytes) != hipSuccess) {
        hipFree(plan->d_fwd);
        return fail(CKKS_ERR_OOM, "twiddle table allocation failed");
    }
    CK_HIP(hipMemcpy(plan->d_fwd, fwd.data(), bytes, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(plan->d_inv, inv.data(), bytes, hipMemcpyHostToDevice));
    DevPrime &p = plan->prime;
    p.q = q;
    p.fwd = plan->d_fwd;
    p.inv = plan->d_inv;
    p.ninv = host::invmod(degree % q, q);
    p.ninv_q = host::shoup_quotient(p.ninv, q);
    p.logn = logn;
    ctx->ntt64_plans.insert(plan.get());
    *out = plan.release();
    return CKKS_OK;
}

int ckks_ntt64_plan_destroy(ckks_ntt64_plan *plan) {
    if (!plan) return CKKS_OK;
    hipSetDevice(plan->ctx->device);
    hipStreamSynchronize(plan->ctx->stream);
    plan->ctx->ntt64_plans.erase(plan);
    hipFree(plan->d_fwd);
    hipFree(plan->d_inv);
    delete plan;
    return CKKS_OK;
}

int ckks_ntt64_plan_psi(const ckks_ntt64_plan *plan, uint64_t *psi) {
    if (!plan || !psi) return fail(CKKS_ERR_INVALID, "null argument");
    *psi = plan->psi;
    return CKKS_OK;
}

int ckks_ntt64_forward(const ckks_ntt64_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys) {
    return ntt64_run(plan, in, out, n_polys, true);
}

int ckks_ntt64_inverse(const ckks_ntt64_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys) {
    return ntt64_run(plan, in, out, n_polys, false);
}


int ckks_ntt_forward(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                     size_t n_polys, size_t n_limbs, size_t poly_stride) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    NttCall c{plan, in, out, n_polys, n_limbs, poly_stride, poly_stride};
    return launch_ntt(plan->ctx, &c, 1, true);
}

int ckks_ntt_forward_noalloc(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys,
                             size_t n_limbs, size_t poly_stride) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    const ckks_context *ctx = plan->ctx;
    if (!ctx->ntt_scratch) return fail(CKKS_ERR_INVALID, "no scratch on this device");
    if (plan->degree * sizeof(u64) > kNttScratchBytes)
        return fail(CKKS_ERR_INVALID, "one limb transform exceeds the 32 MiB scratch");
    NttCall c{plan, in, out, n_polys, n_limbs, poly_stride, poly_stride};
    return launch_ntt(plan->ctx, &c, 1, true, ctx->ntt_scratch, kNttScratchBytes / sizeof(u64));
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

extern "C" int ckks_autotune_ntt(const ckks_rns_plan *plan, size_t n_polys, size_t n_limbs,
                                 int *best_variant, float *ms_per_variant) {
    if (!plan || !best_variant) return fail(CKKS_ERR_INVALID, "null argument");
    ckks_context *ctx = plan->ctx;
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "autotuning is not allowed during graph capture");
    ProfileGuard profile_guard(ctx);
    int rc = check_plan_limbs(plan, n_limbs, "NTT autotune");
    if (rc) return rc;
    if (n_polys == 0) return fail(CKKS_ERR_INVALID, "autotune needs at least one polynomial");
    if (ms_per_variant)
        for (int v = 0; v < CKKS_NTT_VARIANT_COUNT; ++v) ms_per_variant[v] = -1.f;
    const u32 logn = plan->logn;
    *best_variant = 0;
    if (!ntt_two_pass(logn)) return CKKS_OK;  // one kernel serves N <= 4096
    const size_t per = n_limbs * plan->degree, elems = n_polys * per;
    CK_HIP(hipSetDevice(ctx->device));
    u64 *buf = nullptr;
    CK_HIP(hipMalloc(&buf, 2 * elems * sizeof(u64)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const int saved = ctx->ntt_variant[logn];
    float best_ms = 0.f;
    int best = saved;
    rc = CKKS_OK;
    auto body = [&]() -> int {
        CK_HIP(hipMemsetAsync(buf, 0, 2 * elems * sizeof(u64), ctx->stream));
        CK_HIP(hipEventCreate(&e0));
        CK_HIP(hipEventCreate(&e1));
        const NttCall fwd{plan, buf, buf + elems, n_polys, n_limbs, per, per};
        const NttCall inv{plan, buf + elems, buf, n_polys, n_limbs, per, per};
        for (int v = 0; v < CKKS_NTT_VARIANT_COUNT; ++v) {
            if ((v & CKKS_NTT_VARIANT_TILE8K) && logn < 14) continue;
            ctx->ntt_variant[logn] = (uint8_t)v;
            int r = launch_ntt(ctx, &fwd, 1, true);
            if (!r) r = launch_ntt(ctx, &inv, 1, false);
            if (r) return r;
            const int reps = 3;
            CK_HIP(hipEventRecord(e0, ctx->stream));
            for (int i = 0; i < reps && !r; ++i) {
                r = launch_ntt(ctx, &fwd, 1, true);
                if (!r) r = launch_ntt(ctx, &inv, 1, false);
            }
            if (r) return r;
            CK_HIP(hipEventRecord(e1, ctx->stream));
            CK_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            CK_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= reps;
            if (ms_per_variant) ms_per_variant[v] = ms;
            if (best_ms == 0.f || ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        return CKKS_OK;
    };
    rc = body();
    hipStreamSynchronize(ctx->stream);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(buf);
    ctx->ntt_variant[logn] = (uint8_t)(rc ? saved : best);
    if (!rc) *best_variant = best;
    return rc;
}
