// CKKS encode / decode on the device: the integer halves of encode_rns
// (application_support/ckks.rs:39-66) and of the reference's decode
// (bin/r30_13b_ckks_primitives.rs:55-80).  The FP64 canonical embedding
// between slots and real coefficients runs as an FFT on the device (the
// Python side); these kernels move between real coefficients and residues.
//   encode: signed = round(coefficient * scale), halfway cases away from
//           zero (f64::round); residue = signed rem_euclid q per limb.
//   decode: exact CRT to the centered integer (crt.hpp), then / scale.
#include <cmath>

#include "crt.hpp"
#include "internal.hpp"

namespace ck {

struct EncodeArgs {
    const double *coeffs;  // [p][N]
    u64 *out;              // [p][limb][N]
    const DevPrime *primes;
    double scale;
    u32 n_limbs, logn;
    u64 total;  // p * N
    int *overflow;
};

__global__ __launch_bounds__(256) void k_encode_round(EncodeArgs a) {
    const u64 gid = (u64)blockIdx.x * 256 + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u64 p = gid >> a.logn, n = gid & (N - 1);
    const double x = a.coeffs[gid] * a.scale;
    double r = trunc(x);
    if (fabs(x - r) >= 0.5) r += x > 0 ? 1.0 : -1.0;
    // |r| < 2^63 is representable as int64
    if (!(fabs(r) < 9223372036854775808.0)) {
        *a.overflow = 1;  // plain vector store
        r = 0;
    }
    const int64_t v = (int64_t)r;
    u64 *o = a.out + p * (u64)a.n_limbs * N + n;
    for (u32 i = 0; i < a.n_limbs; ++i) {
        const u64 q = a.primes[i].q;
        u64 res;
        if (v >= 0) {
            res = (u64)v % q;
        } else {
            const u64 mag = ((u64)(-(v + 1)) + 1) % q;
            res = mag ? q - mag : 0;
        }
        o[(u64)i * N] = res;
    }
}

struct DecodeArgs {
    const u64 *in;  // [p][limb][N]
    double *out;    // [p][N]
    CrtTable t;
    const DevPrime *primes;
    double inv_scale;
    u32 n_limbs, logn;
    u64 total;
};

template <int W>
__global__ __launch_bounds__(256) void k_decode_crt(DecodeArgs a) {
    const u64 gid = (u64)blockIdx.x * 256 + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u64 p = gid >> a.logn, n = gid & (N - 1);
    u64 X[W];
    const bool negative = crt_centered<W>(a.in + p * (u64)a.n_limbs * N + n, N, false, a.n_limbs, a.primes, a.t, X);
    if (negative) {  // magnitude of the two's-complement value
        u64 carry = 1;
#pragma unroll
        for (int k = 0; k < W; ++k) {
            const u64 v = ~X[k] + carry;
            carry = (carry && v == 0) ? 1 : 0;
            X[k] = v;
        }
    }
    // FP64 value from the top three nonzero words (192 significant bits)
    int top = -1;
#pragma unroll
    for (int k = 0; k < W; ++k)
        if (X[k]) top = k;
    double d = 0.0;
    if (top >= 0) {
        const double two64 = 18446744073709551616.0;
        u64 w0 = X[0], w1 = 0, w2 = 0;
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k == top) w2 = X[k];
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k == top - 1) w1 = X[k];
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k == top - 2) w0 = X[k];
        if (top < 2) w0 = top == 1 ? X[0] : 0;
        if (top == 0) w1 = 0;
        d = ((double)w2 * two64 + (double)w1) * two64 + (double)w0;
        const int shift = top >= 2 ? top - 2 : (top == 1 ? -1 : -2);
        d = ldexp(d, 64 * shift);
    }
    a.out[gid] = (negative ? -d : d) * a.inv_scale;
}

static int crt_table_for(const ckks_rns_plan *plan, size_t n_limbs, CrtTable *t, unsigned *words) {
    auto *mp = const_cast<ckks_rns_plan *>(plan);
    auto it = mp->crt.find(n_limbs);
    if (it == mp->crt.end()) {
        if (plan->ctx->capturing) return fail(CKKS_ERR_INVALID, "decode tables cannot be built during graph capture");
        CrtHost h;
        if (!build_crt(plan->moduli, n_limbs, h)) return fail(CKKS_ERR_UNSUPPORTED, "decode: Q above 2047 bits");
        u64 *d = nullptr;
        CK_HIP(hipMalloc(&d, h.blob.size() * sizeof(u64)));
        CK_HIP(hipMemcpy(d, h.blob.data(), h.blob.size() * sizeof(u64), hipMemcpyHostToDevice));
        it = mp->crt.emplace(n_limbs, std::make_pair(d, h.words)).first;
    }
    *t = crt_table(it->second.first, n_limbs, it->second.second);
    *words = it->second.second;
    return CKKS_OK;
}

}  // namespace ck

using namespace ck;

extern "C" {

int ckks_encode_coefficients(const ckks_rns_plan *plan, const double *coeffs, double scale, uint64_t *out,
                             size_t n_polys, size_t n_limbs) {
    int rc = check_plan_limbs(plan, n_limbs, "encode");
    if (rc || n_polys == 0) return rc;
    if (!coeffs || !out) return fail(CKKS_ERR_INVALID, "null buffer");
    if (!(scale > 0.0) || !std::isfinite(scale)) return fail(CKKS_ERR_INVALID, "CKKS scale must be positive and finite");
    ckks_context *ctx = plan->ctx;
    // the overflow flag is read back on the host before returning
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "encode reads its overflow flag back: not capturable");
    void *w;
    if ((rc = ctx->workspace(2, sizeof(int), &w))) return rc;
    CK_HIP(hipMemsetAsync(w, 0, sizeof(int), ctx->stream));
    EncodeArgs a{coeffs, out, plan->d_primes, scale, (u32)n_limbs, plan->logn, (u64)n_polys * plan->degree, (int *)w};
    hipLaunchKernelGGL(k_encode_round, dim3((u32)((a.total + 255) / 256)), dim3(256), 0, ctx->stream, a);
    CK_HIP(hipGetLastError());
    int overflow = 0;
    CK_HIP(hipMemcpyAsync(&overflow, w, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CK_HIP(hipStreamSynchronize(ctx->stream));
    if (overflow) return fail(CKKS_ERR_INVALID, "encoded coefficient magnitude reaches 2^63");
    return CKKS_OK;
}

int ckks_decode_coefficients(const ckks_rns_plan *plan, const uint64_t *in, double scale, double *out,
                             size_t n_polys, size_t n_limbs) {
    int rc = check_plan_limbs(plan, n_limbs, "decode");
    if (rc || n_polys == 0) return rc;
    if (!in || !out) return fail(CKKS_ERR_INVALID, "null buffer");
    if (!(scale > 0.0) || !std::isfinite(scale)) return fail(CKKS_ERR_INVALID, "CKKS scale must be positive and finite");
    CrtTable t;
    unsigned W = 0;
    if ((rc = crt_table_for(plan, n_limbs, &t, &W))) return rc;
    DecodeArgs a{in, out, t, plan->d_primes, 1.0 / scale, (u32)n_limbs, plan->logn, (u64)n_polys * plan->degree};
    const dim3 g((u32)((a.total + 255) / 256));
    hipStream_t s = plan->ctx->stream;
    switch (W) {
#define CK_DC(WW) \
    case WW: hipLaunchKernelGGL(k_decode_crt<WW>, g, dim3(256), 0, s, a); break;
        CK_DC(4) CK_DC(8) CK_DC(12) CK_DC(16) CK_DC(24) CK_DC(32)
#undef CK_DC
        default: return fail(CKKS_ERR_UNSUPPORTED, "decode word count");
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

}  // extern "C"
