// Element-wise ring kernels (gfx950): pointwise arithmetic, Galois
// automorphisms in both domains, exact rescale, mod switch and the tensor.
// All are HBM-bound streams; each thread owns 2 consecutive residues of one
// limb so loads/stores are 16-byte and a block never straddles limbs.
#include <algorithm>

#include "host_math.hpp"
#include "internal.hpp"

namespace ck {

constexpr int kEwThreads = 256;
constexpr int kEwPerThread = 2;
constexpr int kEwChunk = kEwThreads * kEwPerThread;  // residues per block

enum EwOp { EW_MUL = 0, EW_MAC = 1, EW_ADD = 2, EW_SUB = 3, EW_NEG = 4 };

struct EwArgs {
    const u64 *x, *y;
    u64 *out;
    const DevPrime *primes;
    u64 poly_stride;
    u32 n_limbs, logn, chunks;  // chunks per limb
};

template <int OP>
__global__ __launch_bounds__(kEwThreads) void k_elementwise(EwArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const DevPrime P = a.primes[limb];
    Barrett br{P.q, P.br_hi, P.br_lo};
    const size_t base = poly * a.poly_stride + ((size_t)limb << a.logn) +
                        (size_t)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    const u64 q = P.q;
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        u64 x = a.x[base + i];
        u64 r;
        if (OP == EW_NEG) {
            r = neg_mod(x, q);
        } else {
            u64 y = a.y[base + i];
            if (OP == EW_MUL) r = mul_mod(x, y, br);
            else if (OP == EW_ADD) r = add_mod(x, y, q);
            else if (OP == EW_SUB) r = sub_mod(x, y, q);
            else r = add_mod(a.out[base + i], mul_mod(x, y, br), q);  // EW_MAC
        }
        a.out[base + i] = r;
    }
}

static int elementwise(int op, const ckks_rns_plan *plan, const u64 *x, const u64 *y, u64 *out,
                       size_t n_polys, size_t n_limbs, size_t poly_stride) {
    int rc = check_plan_limbs(plan, n_limbs, "pointwise");
    if (rc) return rc;
    if (n_polys == 0) return CKKS_OK;
    if (!x || !out || (op != EW_NEG && !y)) return fail(CKKS_ERR_INVALID, "null buffer");
    if (n_polys > 1 && poly_stride < n_limbs * plan->degree)
        return fail(CKKS_ERR_INVALID, "polynomial stride smaller than its limbs");
    EwArgs a{};
    a.x = x;
    a.y = y;
    a.out = out;
    a.primes = plan->d_primes;
    a.poly_stride = n_polys > 1 ? poly_stride : n_limbs * plan->degree;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(plan->degree / kEwChunk);
    dim3 grid((u32)(n_polys * n_limbs * a.chunks));
    hipStream_t s = plan->ctx->stream;
    switch (op) {
        case EW_MUL: hipLaunchKernelGGL(k_elementwise<EW_MUL>, grid, dim3(kEwThreads), 0, s, a); break;
        case EW_MAC: hipLaunchKernelGGL(k_elementwise<EW_MAC>, grid, dim3(kEwThreads), 0, s, a); break;
        case EW_ADD: hipLaunchKernelGGL(k_elementwise<EW_ADD>, grid, dim3(kEwThreads), 0, s, a); break;
        case EW_SUB: hipLaunchKernelGGL(k_elementwise<EW_SUB>, grid, dim3(kEwThreads), 0, s, a); break;
        default: hipLaunchKernelGGL(k_elementwise<EW_NEG>, grid, dim3(kEwThreads), 0, s, a); break;
    }
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- Galois automorphisms ------------------------------------------------------

struct PermArgs {
    const u64 *in;
    u64 *out;
    const DevPrime *primes;
    u64 poly_stride;
    u64 k;  // coefficient domain: k^-1 mod 2N; NTT domain: k mod 2N
    u32 n_limbs, logn, chunks;
};

// Coefficient domain (ckks/automorphism.rs:8-46) as a gather: output j takes
// source i' = j k^-1 mod 2N; i' < N gives +in[i'], otherwise -in[i' - N].
__global__ __launch_bounds__(kEwThreads) void k_automorphism(PermArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const u64 q = a.primes[limb].q;
    const u64 n = (u64)1 << a.logn, mask2 = 2 * n - 1;
    const size_t base = poly * a.poly_stride + ((size_t)limb << a.logn);
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        u64 j = (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread + i;
        u64 src = (j * a.k) & mask2;
        u64 v = src < n ? a.in[base + src] : neg_mod(a.in[base + src - n], q);
        a.out[base + j] = v;
    }
}

// NTT domain (ring/ntt_galois_permutation.rs:5-19):
// out[j] = in[((k (2j+1) mod 2N) - 1) / 2].
__global__ __launch_bounds__(kEwThreads) void k_ntt_galois(PermArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const u64 n = (u64)1 << a.logn, mask2 = 2 * n - 1;
    const size_t base = poly * a.poly_stride + ((size_t)limb << a.logn);
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        u64 j = (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread + i;
        u64 rot = (a.k * (2 * j + 1)) & mask2;
        a.out[base + j] = a.in[base + ((rot - 1) >> 1)];
    }
}

static int permute(bool ntt_domain, const ckks_rns_plan *plan, const u64 *in, u64 *out,
                   size_t n_polys, size_t n_limbs, size_t poly_stride, u64 exponent) {
    int rc = check_plan_limbs(plan, n_limbs, "automorphism");
    if (rc) return rc;
    if (n_polys == 0) return CKKS_OK;
    if (!in || !out) return fail(CKKS_ERR_INVALID, "null buffer");
    if (in == out) return fail(CKKS_ERR_INVALID, "automorphism output must not alias its input");
    const u64 two_n = 2 * (u64)plan->degree;
    u64 k = exponent % two_n;
    if ((k & 1) == 0)
        return fail(CKKS_ERR_INVALID, ntt_domain ? "Galois exponent must be odd"
                                                 : "CKKS automorphism exponent must be odd modulo 2N");
    if (n_polys > 1 && poly_stride < n_limbs * plan->degree)
        return fail(CKKS_ERR_INVALID, "polynomial stride smaller than its limbs");
    PermArgs a{};
    a.in = in;
    a.out = out;
    a.primes = plan->d_primes;
    a.poly_stride = n_polys > 1 ? poly_stride : n_limbs * plan->degree;
    if (ntt_domain) {
        a.k = k;
    } else {
        // inverse of k modulo 2N (odd k, power-of-two modulus): Newton
        u64 inv = k;
        for (int it = 0; it < 6; ++it) inv = inv * (2 - k * inv);
        a.k = inv & (two_n - 1);
    }
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(plan->degree / kEwChunk);
    dim3 grid((u32)(n_polys * n_limbs * a.chunks));
    if (ntt_domain)
        hipLaunchKernelGGL(k_ntt_galois, grid, dim3(kEwThreads), 0, plan->ctx->stream, a);
    else
        hipLaunchKernelGGL(k_automorphism, grid, dim3(kEwThreads), 0, plan->ctx->stream, a);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- NTT-resident plaintext multiply-accumulate --------------------------------

constexpr u32 kMaxMacTerms = 64;

struct MacListArgs {
    const u64 *ct[kMaxMacTerms];  // NTT-domain ciphertexts [2][limbs][N]
    const u64 *pt[kMaxMacTerms];  // NTT-domain plaintexts [limbs][N]
    u64 *out;                     // [2][limbs][N]
    const DevPrime *primes;
    u32 count, n_limbs, logn, chunks;
};

// out = sum_t ct_t * pt_t, limb by limb and evaluation point by point
// (NttPolynomial::pointwise_mul_add_assign, ring/ntt_polynomial.rs:222-235,
// over the NTT-resident ciphertexts of ckks/rns_plain_multiply.rs:119-191).
__global__ __launch_bounds__(kEwThreads) void k_mac_list(MacListArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;  // component * limbs + limb
    const u32 limb = pl % a.n_limbs;
    const DevPrime P = a.primes[limb];
    const Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 off = ((u64)pl << a.logn) + (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    const u64 poff = ((u64)limb << a.logn) + (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    u64 acc[kEwPerThread] = {};
    for (u32 t = 0; t < a.count; ++t) {
#pragma unroll
        for (int i = 0; i < kEwPerThread; ++i)
            acc[i] = add_mod(acc[i], mul_mod(a.ct[t][off + i], a.pt[t][poff + i], br), P.q);
    }
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) a.out[off + i] = acc[i];
}

// ---- monomials and scalars -----------------------------------------------------

struct MonoArgs {
    const u64 *in;
    u64 *out;
    const DevPrime *primes;
    u64 poly_stride;
    u64 e;       // exponent mod 2N
    u32 negate;  // coefficient domain: the external sign is -1
    u32 n_limbs, logn, chunks;
};

// NTT domain (ring/ntt_polynomial.rs:133-175): evaluation k times
// r_k^e = psi^((2k+1) e mod 2N), read from the forward twiddle table
// T[i] = psi^bitrev(i) (psi^t = T[bitrev(t)] for t < N, psi^(N+t) = -psi^t).
// The reference walks the same powers as a geometric progression.
__global__ __launch_bounds__(kEwThreads) void k_ntt_monomial(MonoArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const DevPrime P = a.primes[limb];
    const Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 n = (u64)1 << a.logn, mask2 = 2 * n - 1;
    const size_t base = poly * a.poly_stride + ((size_t)limb << a.logn);
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        const u64 k = (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread + i;
        const u64 t = ((2 * k + 1) * a.e) & mask2;
        const u64 idx = __brev((u32)(t & (n - 1))) >> (32 - a.logn);
        const u64 raw = P.fwd[idx].w;
        const u64 w = P.fp ? (u64)__longlong_as_double((long long)raw) : raw;
        const u64 r = mul_mod(a.in[base + k], w, br);
        a.out[base + k] = t >= n ? neg_mod(r, P.q) : r;
    }
}

// Coefficient domain (ring/polynomial.rs:164-200) as a gather: output d
// takes source (d - shift) mod N, negated by the exponent's sign (e >= N),
// the external sign and the wrap (d < shift).
__global__ __launch_bounds__(kEwThreads) void k_monomial(MonoArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const u64 q = a.primes[limb].q;
    const u64 n = (u64)1 << a.logn;
    const u64 shift = a.e & (n - 1);
    const u32 base_neg = (a.e >= n ? 1u : 0u) ^ a.negate;
    const size_t base = poly * a.poly_stride + ((size_t)limb << a.logn);
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        const u64 d = (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread + i;
        const u64 src = (d - shift) & (n - 1);
        const u64 v = a.in[base + src];
        a.out[base + d] = (base_neg ^ (d < shift ? 1u : 0u)) ? neg_mod(v, q) : v;
    }
}

constexpr u32 kMaxScalarLimbs = 128;

struct ScalarArgs {
    const u64 *in;
    u64 *out;
    const DevPrime *primes;
    u64 poly_stride;
    u32 n_limbs, logn, chunks;
    u64 w[kMaxScalarLimbs];  // scalar mod q_i, Shoup quotient
    u64 wq[kMaxScalarLimbs];
};

__global__ __launch_bounds__(kEwThreads) void k_scalar_mul(ScalarArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const u64 q = a.primes[limb].q;
    const u64 w = a.w[limb], wq = a.wq[limb];
    const size_t base = poly * a.poly_stride + ((size_t)limb << a.logn) + (size_t)chunk * kEwChunk +
                        threadIdx.x * kEwPerThread;
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) a.out[base + i] = shoup(a.in[base + i], w, wq, q);
}

static int check_poly_call(const ckks_rns_plan *plan, const u64 *in, const u64 *out, size_t n_polys,
                           size_t n_limbs, size_t &poly_stride, const char *what) {
    int rc = check_plan_limbs(plan, n_limbs, what);
    if (rc) return rc;
    if (n_polys && (!in || !out)) return fail(CKKS_ERR_INVALID, "null buffer");
    if (n_polys > 1 && poly_stride < n_limbs * plan->degree)
        return fail(CKKS_ERR_INVALID, "polynomial stride smaller than its limbs");
    if (n_polys <= 1) poly_stride = n_limbs * plan->degree;
    return CKKS_OK;
}

static int monomial(bool ntt_domain, const ckks_rns_plan *plan, const u64 *in, u64 *out, size_t n_polys,
                    size_t n_limbs, size_t poly_stride, u64 e, int sign) {
    int rc = check_poly_call(plan, in, out, n_polys, n_limbs, poly_stride, "monomial");
    if (rc || n_polys == 0) return rc;
    if (!ntt_domain && in == out) return fail(CKKS_ERR_INVALID, "monomial output must not alias its input");
    if (sign != 1 && sign != -1) return fail(CKKS_ERR_INVALID, "monomial sign must be either 1 or -1");
    MonoArgs a{};
    a.in = in;
    a.out = out;
    a.primes = plan->d_primes;
    a.poly_stride = poly_stride;
    a.e = e % (2 * (u64)plan->degree);
    a.negate = sign == -1 ? 1u : 0u;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(plan->degree / kEwChunk);
    const dim3 grid((u32)(n_polys * n_limbs * a.chunks));
    if (ntt_domain)
        hipLaunchKernelGGL(k_ntt_monomial, grid, dim3(kEwThreads), 0, plan->ctx->stream, a);
    else
        hipLaunchKernelGGL(k_monomial, grid, dim3(kEwThreads), 0, plan->ctx->stream, a);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

static int scalar_mul(const ckks_rns_plan *plan, const u64 *in, const u64 *scalars, u64 *out, size_t n_polys,
                      size_t n_limbs, size_t poly_stride) {
    int rc = check_poly_call(plan, in, out, n_polys, n_limbs, poly_stride, "scalar multiply");
    if (rc || n_polys == 0) return rc;
    if (!scalars) return fail(CKKS_ERR_INVALID, "null scalars");
    if (n_limbs > kMaxScalarLimbs) return fail(CKKS_ERR_UNSUPPORTED, "scalar multiply of more than 128 limbs");
    ScalarArgs a{};
    a.in = in;
    a.out = out;
    a.primes = plan->d_primes;
    a.poly_stride = poly_stride;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(plan->degree / kEwChunk);
    for (size_t i = 0; i < n_limbs; ++i) {
        const u64 q = plan->moduli[i];
        a.w[i] = scalars[i] % q;
        a.wq[i] = host::shoup_quotient(a.w[i], q);
    }
    hipLaunchKernelGGL(k_scalar_mul, dim3((u32)(n_polys * n_limbs * a.chunks)), dim3(kEwThreads), 0,
                       plan->ctx->stream, a);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- rescale / mod switch ------------------------------------------------------

struct RescaleArgs {
    const u64 *in;
    u64 *out;
    const DevPrime *primes;
    const RescaleConst *rc;  // row of the dropped limb: rc[i] for target i
    u64 in_stride, out_stride;
    u32 n_limbs, logn;
    u32 rescale;  // 0: mod switch (drop only)
};

// One thread per coefficient of one ciphertext component; loops over the
// kept limbs so the dropped limb is read once.
__global__ __launch_bounds__(kEwThreads) void k_rescale(RescaleArgs a) {
    const u64 n = (u64)1 << a.logn;
    const u64 chunks = n / kEwThreads;
    const u64 b = blockIdx.x;
    const u64 c = (b % chunks) * kEwThreads + threadIdx.x;
    const u64 comp = (b / chunks) % 2, ct = (b / chunks) / 2;
    const u32 L = a.n_limbs, last = L - 1;
    const u64 *src = a.in + ct * a.in_stride + comp * (u64)L * n;
    u64 *dst = a.out + ct * a.out_stride + comp * (u64)last * n;
    if (!a.rescale) {
        for (u32 i = 0; i < last; ++i) dst[(u64)i * n + c] = src[(u64)i * n + c];
        return;
    }
    const u64 q_drop = a.primes[last].q;
    const u64 rem = src[(u64)last * n + c];
    const bool neg = rem > q_drop / 2;
    for (u32 i = 0; i < last; ++i) {
        const DevPrime &P = a.primes[i];
        const RescaleConst k = a.rc[i];
        u64 rq = rem >= P.q ? barrett_reduce128(0, rem, Barrett{P.q, P.br_hi, P.br_lo}) : rem;
        if (neg) rq = sub_mod(rq, k.ql_mod, P.q);  // (rem - q_drop) mod q_i
        u64 diff = sub_mod(src[(u64)i * n + c], rq, P.q);
        dst[(u64)i * n + c] = shoup(diff, k.inv, k.inv_q, P.q);
    }
}

static int rescale_like(bool rescale, const ckks_rns_plan *plan, const u64 *in, u64 *out,
                        size_t n_cts, size_t n_limbs, size_t in_stride, size_t out_stride) {
    int rc = check_plan_limbs(plan, n_limbs, "rescale");
    if (rc) return rc;
    if (n_limbs < 2) return fail(CKKS_ERR_INVALID, "RNS CKKS rescale requires at least two modulus limbs");
    if (n_cts == 0) return CKKS_OK;
    if (!in || !out) return fail(CKKS_ERR_INVALID, "null buffer");
    const size_t n = plan->degree;
    if (in_stride == 0) in_stride = 2 * n_limbs * n;
    if (out_stride == 0) out_stride = 2 * (n_limbs - 1) * n;
    if (in_stride < 2 * n_limbs * n || out_stride < 2 * (n_limbs - 1) * n)
        return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    RescaleArgs a{};
    a.in = in;
    a.out = out;
    a.primes = plan->d_primes;
    a.rc = plan->d_rescale + (n_limbs - 1) * plan->moduli.size();
    a.in_stride = in_stride;
    a.out_stride = out_stride;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.rescale = rescale ? 1u : 0u;
    dim3 grid((u32)(n_cts * 2 * (n / kEwThreads)));
    hipLaunchKernelGGL(k_rescale, grid, dim3(kEwThreads), 0, plan->ctx->stream, a);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// ---- tensor --------------------------------------------------------------------

// NTT-domain tensor of (b0,a0) x (b1,a1): c0 = b0 b1, c1 = b0 a1 + a0 b1,
// c2 = a0 a1, each reduced once from a 128-bit sum.
struct TensorArgs {
    const u64 *x;  // [ct][b0,a0,b1,a1][limb][N]
    u64 *y;        // [ct][c0,c1,c2][limb][N]
    const DevPrime *primes;
    u32 n_limbs, logn, chunks;
};

__global__ __launch_bounds__(kEwThreads) void k_tensor(TensorArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, ct = pl / a.n_limbs;
    const DevPrime P = a.primes[limb];
    Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 n = (u64)1 << a.logn;
    const u64 poly = (u64)a.n_limbs * n;
    const u64 off = (u64)limb * n + (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    const u64 *x = a.x + ct * 4 * poly + off;
    u64 *y = a.y + ct * 3 * poly + off;
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        u64 b0 = x[i], a0 = x[poly + i], b1 = x[2 * poly + i], a1 = x[3 * poly + i];
        y[i] = mul_mod(b0, b1, br);
        Acc128 acc;
        acc.zero();
        acc.mac(b0, a1);
        acc.mac(a0, b1);
        y[poly + i] = acc.reduce(br);
        y[2 * poly + i] = mul_mod(a0, a1, br);
    }
}

static int tensor(const ckks_rns_plan *plan, const u64 *ct0, const u64 *ct1, u64 *out, size_t n_cts,
                  size_t n_limbs, size_t in_stride, size_t out_stride) {
    int rc = check_plan_limbs(plan, n_limbs, "tensor");
    if (rc) return rc;
    if (n_cts == 0) return CKKS_OK;
    if (!ct0 || !ct1 || !out) return fail(CKKS_ERR_INVALID, "null buffer");
    const size_t n = plan->degree, poly = n_limbs * n;
    if (in_stride == 0) in_stride = 2 * poly;
    if (out_stride == 0) out_stride = 3 * poly;
    if (in_stride < 2 * poly || out_stride < 3 * poly)
        return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    ckks_context *ctx = plan->ctx;
    void *w1, *w2;
    if ((rc = ctx->workspace(1, n_cts * 4 * poly * sizeof(u64), &w1))) return rc;
    if ((rc = ctx->workspace(2, n_cts * 3 * poly * sizeof(u64), &w2))) return rc;
    u64 *x = (u64 *)w1, *y = (u64 *)w2;
    // x = [ct][b0, a0, b1, a1][limb][N] in the NTT domain
    NttCall fwd[4] = {
        {plan, ct0, x, n_cts, n_limbs, in_stride, 4 * poly},
        {plan, ct0 + poly, x + poly, n_cts, n_limbs, in_stride, 4 * poly},
        {plan, ct1, x + 2 * poly, n_cts, n_limbs, in_stride, 4 * poly},
        {plan, ct1 + poly, x + 3 * poly, n_cts, n_limbs, in_stride, 4 * poly},
    };
    if ((rc = launch_ntt(ctx, fwd, 4, true))) return rc;
    TensorArgs a{};
    a.x = x;
    a.y = y;
    a.primes = plan->d_primes;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(n / kEwChunk);
    dim3 grid((u32)(n_cts * n_limbs * a.chunks));
    hipLaunchKernelGGL(k_tensor, grid, dim3(kEwThreads), 0, ctx->stream, a);
    CK_HIP(hipGetLastError());
    NttCall inv[3] = {
        {plan, y, out, n_cts, n_limbs, 3 * poly, out_stride},
        {plan, y + poly, out + poly, n_cts, n_limbs, 3 * poly, out_stride},
        {plan, y + 2 * poly, out + 2 * poly, n_cts, n_limbs, 3 * poly, out_stride},
    };
    return launch_ntt(ctx, inv, 3, false);
}

// ---- plaintext multiply ----------------------------------------------------------

// NTT-domain ciphertext components times a plaintext: x[c][comp][i] * pt[c][i].
struct MulPlainArgs {
    const u64 *x;   // [ct][2][limb][N] NTT domain (workspace)
    const u64 *pt;  // plaintext of ct c at pt + c * pt_stride
    u64 *y;         // [ct][2][limb][N]
    const DevPrime *primes;
    u64 pt_stride;
    u32 n_limbs, logn, chunks;
};

__global__ __launch_bounds__(kEwThreads) void k_mul_plain(MulPlainArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, cc = pl / a.n_limbs;  // cc = ct * 2 + comp
    const u64 ct = cc / 2;
    const DevPrime P = a.primes[limb];
    Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 n = (u64)1 << a.logn;
    const u64 off = (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    const u64 *x = a.x + ((u64)cc * a.n_limbs + limb) * n + off;
    const u64 *p = a.pt + ct * a.pt_stride + (u64)limb * n + off;
    u64 *y = a.y + ((u64)cc * a.n_limbs + limb) * n + off;
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) y[i] = mul_mod(x[i], p[i], br);
}

static int multiply_plain(const ckks_rns_plan *plan, const u64 *in, const u64 *pt, u64 *out, size_t n_cts,
                          size_t n_limbs, size_t in_stride, size_t pt_stride, size_t out_stride) {
    int rc = check_plan_limbs(plan, n_limbs, "plaintext multiply");
    if (rc) return rc;
    if (n_cts == 0) return CKKS_OK;
    if (!in || !pt || !out) return fail(CKKS_ERR_INVALID, "null buffer");
    const size_t n = plan->degree, poly = n_limbs * n;
    if (in_stride == 0) in_stride = 2 * poly;
    if (out_stride == 0) out_stride = 2 * poly;
    if (in_stride < 2 * poly || out_stride < 2 * poly || (pt_stride && pt_stride < poly))
        return fail(CKKS_ERR_INVALID, "ciphertext stride smaller than its limbs");
    ckks_context *ctx = plan->ctx;
    void *w;
    if ((rc = ctx->workspace(1, n_cts * 2 * poly * sizeof(u64), &w))) return rc;
    u64 *x = (u64 *)w;
    NttCall fwd[2] = {{plan, in, x, n_cts, n_limbs, in_stride, 2 * poly},
                      {plan, in + poly, x + poly, n_cts, n_limbs, in_stride, 2 * poly}};
    if ((rc = launch_ntt(ctx, fwd, 2, true))) return rc;
    MulPlainArgs a{};
    a.x = x;
    a.pt = pt;
    a.y = x;
    a.primes = plan->d_primes;
    a.pt_stride = pt_stride;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(n / kEwChunk);
    hipLaunchKernelGGL(k_mul_plain, dim3((u32)(n_cts * 2 * n_limbs * a.chunks)), dim3(kEwThreads), 0,
                       ctx->stream, a);
    CK_HIP(hipGetLastError());
    NttCall inv[2] = {{plan, x, out, n_cts, n_limbs, 2 * poly, out_stride},
                      {plan, x + poly, out + poly, n_cts, n_limbs, 2 * poly, out_stride}};
    return launch_ntt(ctx, inv, 2, false);
}

// ---- RLWE raw encryption / decryption ---------------------------------------
//
// encrypt_rns_raw_with_distribution_ntt_rng (grafting/rns_key_switch.rs:
// 245-383) once its randomness is drawn on the host: b = m + e - a*s per
// limb, with the one logical error e and the ternary secret s projected to
// every limb (project_error, rlwe/error.rs:150-166), and a*s through the
// NTT.  decrypt_rns_raw_with_ntt (:906-947): b + a*s.

struct ProjectArgs {
    const int8_t *s;  // N signed secret coefficients
    u64 *out;         // [limb][N]
    const DevPrime *primes;
    u32 n_limbs, logn;
};

__global__ __launch_bounds__(kEwThreads) void k_project_secret(ProjectArgs a) {
    const u64 N = (u64)1 << a.logn;
    const u64 gid = (u64)blockIdx.x * kEwThreads + threadIdx.x;
    if (gid >= (u64)a.n_limbs * N) return;
    const u32 limb = (u32)(gid >> a.logn);
    const u64 n = gid & (N - 1);
    const u64 q = a.primes[limb].q;
    const int v = a.s[n];
    a.out[gid] = v >= 0 ? (u64)v % q : neg_mod((u64)(-v) % q, q);
}

struct RlweArgs {
    const u64 *x;       // [p][limb][N]: a (then a*s in place)
    const u64 *s_ntt;   // [limb][N]
    u64 *y;             // [p][limb][N]
    const u64 *m;       // encrypt: message; decrypt: b
    const int64_t *e;   // encrypt: [p][N] errors; decrypt: null
    u64 *out;
    const DevPrime *primes;
    u64 m_stride, out_stride;  // elements between polynomials of m / out
    u32 n_limbs, logn, chunks, encrypt;
};

__global__ __launch_bounds__(kEwThreads) void k_mul_secret(RlweArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const DevPrime P = a.primes[limb];
    const Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 off = ((u64)limb << a.logn) + (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    const u64 base = (u64)poly * ((u64)a.n_limbs << a.logn) + off;
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) a.y[base + i] = mul_mod(a.x[base + i], a.s_ntt[off + i], br);
}

__global__ __launch_bounds__(kEwThreads) void k_rlwe_finish(RlweArgs a) {
    const u32 b = blockIdx.x;
    const u32 chunk = b % a.chunks;
    const u32 pl = b / a.chunks;
    const u32 limb = pl % a.n_limbs, poly = pl / a.n_limbs;
    const u64 q = a.primes[limb].q;
    const u64 N = (u64)1 << a.logn;
    const u64 n0 = (u64)chunk * kEwChunk + threadIdx.x * kEwPerThread;
    const u64 as_base = (u64)poly * ((u64)a.n_limbs << a.logn) + ((u64)limb << a.logn) + n0;
    const u64 m_base = (u64)poly * a.m_stride + ((u64)limb << a.logn) + n0;
    const u64 out_base = (u64)poly * a.out_stride + ((u64)limb << a.logn) + n0;
#pragma unroll
    for (int i = 0; i < kEwPerThread; ++i) {
        const u64 as = a.y[as_base + i];
        u64 r;
        if (a.encrypt) {
            const int64_t ev = a.e[(u64)poly * N + n0 + i];
            u64 ep;
            if (ev >= 0) {
                ep = (u64)ev % q;
            } else {
                const u64 mag = ((u64)(-(ev + 1)) + 1) % q;  // |ev| mod q, no overflow at INT64_MIN
                ep = mag ? q - mag : 0;
            }
            r = sub_mod(add_mod(a.m[m_base + i], ep, q), as, q);
        } else {
            r = add_mod(a.m[m_base + i], as, q);
        }
        a.out[out_base + i] = r;
    }
}

static int rlwe_raw(const ckks_rns_plan *plan, const u64 *m, size_t m_stride, const int64_t *e,
                    const int8_t *secret, const u64 *a_in, size_t a_stride, u64 *out, size_t out_stride,
                    size_t n_polys, size_t n_limbs, bool encrypt) {
    int rc = check_plan_limbs(plan, n_limbs, encrypt ? "encrypt" : "decrypt");
    if (rc || n_polys == 0) return rc;
    if (!m || !secret || !a_in || !out || (encrypt && !e)) return fail(CKKS_ERR_INVALID, "null buffer");
    const size_t N = plan->degree, poly = n_limbs * N;
    if (m_stride == 0) m_stride = poly;
    if (a_stride == 0) a_stride = poly;
    if (out_stride == 0) out_stride = poly;
    if (m_stride < poly || a_stride < poly || out_stride < poly)
        return fail(CKKS_ERR_INVALID, "polynomial stride smaller than its limbs");
    ckks_context *ctx = plan->ctx;
    void *w;
    if ((rc = ctx->workspace(1, (1 + n_polys) * poly * sizeof(u64), &w))) return rc;
    u64 *s_ntt = (u64 *)w, *as = s_ntt + poly;
    ProjectArgs pa{secret, s_ntt, plan->d_primes, (u32)n_limbs, plan->logn};
    hipLaunchKernelGGL(k_project_secret, dim3((u32)((poly + kEwThreads - 1) / kEwThreads)), dim3(kEwThreads), 0,
                       ctx->stream, pa);
    CK_HIP(hipGetLastError());
    NttCall f[2] = {{plan, s_ntt, s_ntt, 1, n_limbs, poly, poly},
                    {plan, a_in, as, n_polys, n_limbs, a_stride, poly}};
    if ((rc = launch_ntt(ctx, f, 2, true))) return rc;
    RlweArgs ra{};
    ra.x = as;
    ra.s_ntt = s_ntt;
    ra.y = as;
    ra.m = m;
    ra.e = e;
    ra.out = out;
    ra.primes = plan->d_primes;
    ra.m_stride = m_stride;
    ra.out_stride = out_stride;
    ra.n_limbs = (u32)n_limbs;
    ra.logn = plan->logn;
    ra.chunks = (u32)(N / kEwChunk);
    ra.encrypt = encrypt ? 1u : 0u;
    const dim3 grid((u32)(n_polys * n_limbs * ra.chunks));
    hipLaunchKernelGGL(k_mul_secret, grid, dim3(kEwThreads), 0, ctx->stream, ra);
    CK_HIP(hipGetLastError());
    NttCall iv{plan, as, as, n_polys, n_limbs, poly, poly};
    if ((rc = launch_ntt(ctx, &iv, 1, false))) return rc;
    hipLaunchKernelGGL(k_rlwe_finish, grid, dim3(kEwThreads), 0, ctx->stream, ra);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

// Key generation from host-sampled randomness (see ckks_keygen): entry j of
// the key is the encryption under s of g_j * source, g_j the gadget factor of
// entry j per limb, computed in the NTT domain by linearity,
//   NTT(b_j) = NTT(g_j * source + e_j) - NTT(a_j) * NTT(s),
// which equals the forward NTT of encrypt_rns_raw's b = m + e - a*s.
struct KeygenArgs {
    const u64 *factors;   // [entries][limbs] gadget factors, canonical
    const u64 *source;    // [limbs][N], coefficient domain
    const int64_t *e;     // [entries][N]
    const u64 *s_ntt;     // [limbs][N]
    const u64 *ka;        // [entries][limbs][N], NTT(a)
    u64 *kb;              // [entries][limbs][N]
    const DevPrime *primes;
    u32 limbs, logn;
    u64 total;            // entries * limbs * N
};

__global__ __launch_bounds__(kEwThreads) void k_keygen_message(KeygenArgs a) {
    const u64 gid = (u64)blockIdx.x * kEwThreads + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u64 row = gid >> a.logn;  // entry * limbs + limb
    const u32 limb = (u32)(row % a.limbs);
    const u64 j = row / a.limbs;
    const DevPrime P = a.primes[limb];
    const Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 q = P.q;
    const u64 x = gid & (N - 1);
    const u64 m = mul_mod(a.factors[row], a.source[((u64)limb << a.logn) + x], br);
    const int64_t ev = a.e[j * N + x];
    u64 ep;
    if (ev >= 0) {
        ep = (u64)ev % q;
    } else {
        const u64 mag = ((u64)(-(ev + 1)) + 1) % q;
        ep = mag ? q - mag : 0;
    }
    a.kb[gid] = add_mod(m, ep, q);
}

__global__ __launch_bounds__(kEwThreads) void k_keygen_finish(KeygenArgs a) {
    const u64 gid = (u64)blockIdx.x * kEwThreads + threadIdx.x;
    if (gid >= a.total) return;
    const u64 N = (u64)1 << a.logn;
    const u32 limb = (u32)((gid >> a.logn) % a.limbs);
    const DevPrime P = a.primes[limb];
    const Barrett br{P.q, P.br_hi, P.br_lo};
    const u64 as = mul_mod(a.ka[gid], a.s_ntt[((u64)limb << a.logn) + (gid & (N - 1))], br);
    a.kb[gid] = sub_mod(a.kb[gid], as, P.q);
}

// Flags (bit 0) a secret coefficient outside {-1, 0, 1}, (bit 1) a nonzero
// one, (bit 2) a residue of `a` that is not canonical.
__global__ __launch_bounds__(kEwThreads) void k_keygen_check(const int8_t *s, const u64 *a, u64 a_total,
                                                             const DevPrime *primes, u32 limbs, u32 logn,
                                                             int *flags) {
    const u64 N = (u64)1 << logn;
    for (u64 i = (u64)blockIdx.x * kEwThreads + threadIdx.x; i < a_total; i += (u64)gridDim.x * kEwThreads) {
        if (i < N) {
            const int v = s[i];
            if (v < -1 || v > 1) flags[0] = 1;
            if (v != 0) flags[1] = 1;
        }
        if (a[i] >= primes[(u32)((i >> logn) % limbs)].q) flags[2] = 1;
    }
}

static int keygen(const ckks_rns_plan *plan, int kind, size_t q_limbs, const size_t *block_sizes, size_t n_blocks,
                  unsigned base_log, const u64 *source, const int8_t *secret, const u64 *a, const int64_t *e,
                  u64 *key_b, u64 *key_a) {
    if (!plan || !source || !secret || !a || !e || !key_b || !key_a) return fail(CKKS_ERR_INVALID, "null argument");
    const auto &mods = plan->moduli;
    const size_t B = mods.size(), N = plan->degree;
    if (q_limbs == 0 || q_limbs > B) return fail(CKKS_ERR_INVALID, "Q limb count must be within the key basis");
    ckks_context *ctx = plan->ctx;
    if (ctx->capturing) return fail(CKKS_ERR_INVALID, "key generation reads its checks back: not capturable");
    // P mod every limb (1 without special primes)
    std::vector<u64> p_mod(B);
    for (size_t i = 0; i < B; ++i) {
        u64 v = 1 % mods[i];
        for (size_t k = q_limbs; k < B; ++k) v = host::mulmod(v, mods[k] % mods[i], mods[i]);
        p_mod[i] = v;
    }
    std::vector<u64> factors;
    size_t entries = 0;
    if (kind == CKKS_KEYGEN_RNS_GADGET || kind == CKKS_KEYGEN_HYBRID) {
        if (kind == CKKS_KEYGEN_RNS_GADGET && q_limbs != B)
            return fail(CKKS_ERR_INVALID, "RNS gadget keys have no special primes");
        if (kind == CKKS_KEYGEN_HYBRID && q_limbs == B)
            return fail(CKKS_ERR_INVALID, "hybrid keys need special primes");
        if (!block_sizes || n_blocks == 0)
            return fail(CKKS_ERR_INVALID, "RNS gadget layout must contain at least one block");
        size_t start = 0;
        for (size_t j = 0; j < n_blocks; ++j) {
            if (block_sizes[j] == 0) return fail(CKKS_ERR_INVALID, "RNS gadget blocks must be nonempty");
            for (size_t i = 0; i < B; ++i)
                factors.push_back(i >= start && i < start + block_sizes[j] ? p_mod[i] : 0);
            start += block_sizes[j];
        }
        if (start != q_limbs) return fail(CKKS_ERR_INVALID, "RNS gadget blocks must cover the full basis exactly");
        entries = n_blocks;
    } else if (kind == CKKS_KEYGEN_BOUNDED) {
        size_t count = 0;
        int rc = ckks_bounded_digit_count(plan, q_limbs, base_log, &count);
        if (rc) return rc;
        std::vector<u64> g(B);
        for (size_t i = 0; i < B; ++i) g[i] = p_mod[i];
        const u64 two_w = base_log >= 64 ? 0 : (1ull << base_log);
        for (size_t j = 0; j < count; ++j) {
            for (size_t i = 0; i < B; ++i) {
                factors.push_back(g[i]);
                g[i] = host::mulmod(g[i], two_w % mods[i], mods[i]);
            }
        }
        entries = count;
    } else {
        return fail(CKKS_ERR_INVALID, "unknown key kind");
    }
    const size_t poly = B * N, total = entries * poly;
    void *w;
    const size_t f_off = poly, flag_off = poly + factors.size();
    int rc = ctx->workspace(1, (flag_off + 2) * sizeof(u64), &w);
    if (rc) return rc;
    u64 *s_ntt = (u64 *)w, *d_factors = s_ntt + f_off;
    int *flags = (int *)(s_ntt + flag_off);
    hipStream_t st = ctx->stream;
    CK_HIP(hipMemcpyAsync(d_factors, factors.data(), factors.size() * sizeof(u64), hipMemcpyHostToDevice, st));
    CK_HIP(hipMemsetAsync(flags, 0, 3 * sizeof(int), st));
    const u32 check_grid = (u32)std::min<u64>((total + kEwThreads - 1) / kEwThreads, 4096);
    hipLaunchKernelGGL(k_keygen_check, dim3(check_grid), dim3(kEwThreads), 0, st, secret, a, (u64)total,
                       plan->d_primes, (u32)B, plan->logn, flags);
    CK_HIP(hipGetLastError());
    int hflags[3] = {0, 0, 0};
    CK_HIP(hipMemcpyAsync(hflags, flags, 3 * sizeof(int), hipMemcpyDeviceToHost, st));
    CK_HIP(hipStreamSynchronize(st));
    if (hflags[0]) return fail(CKKS_ERR_INVALID, "RNS secret coefficients must be ternary");
    if (!hflags[1]) return fail(CKKS_ERR_INVALID, "RNS target secret must be nonzero");
    if (hflags[2]) return fail(CKKS_ERR_INVALID, "uniform residues must be canonical");
    ProjectArgs pa{secret, s_ntt, plan->d_primes, (u32)B, plan->logn};
    hipLaunchKernelGGL(k_project_secret, dim3((u32)((poly + kEwThreads - 1) / kEwThreads)), dim3(kEwThreads), 0,
                       st, pa);
    CK_HIP(hipGetLastError());
    KeygenArgs ka{d_factors, source, e, s_ntt, key_a, key_b, plan->d_primes, (u32)B, plan->logn, (u64)total};
    const dim3 grid((u32)((total + kEwThreads - 1) / kEwThreads));
    hipLaunchKernelGGL(k_keygen_message, grid, dim3(kEwThreads), 0, st, ka);
    CK_HIP(hipGetLastError());
    NttCall f[3] = {{plan, s_ntt, s_ntt, 1, B, poly, poly},
                    {plan, a, key_a, entries, B, poly, poly},
                    {plan, key_b, key_b, entries, B, poly, poly}};
    if ((rc = launch_ntt(ctx, f, 3, true))) return rc;
    hipLaunchKernelGGL(k_keygen_finish, grid, dim3(kEwThreads), 0, st, ka);
    CK_HIP(hipGetLastError());
    return CKKS_OK;
}

}  // namespace ck

using namespace ck;

extern "C" {

int ckks_encrypt_raw(const ckks_rns_plan *plan, const uint64_t *message, const int64_t *error,
                     const int8_t *secret, const uint64_t *a, uint64_t *b, size_t n_polys, size_t n_limbs,
                     size_t poly_stride) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    return rlwe_raw(plan, message, poly_stride, error, secret, a, poly_stride, b, poly_stride, n_polys,
                    n_limbs, true);
}

int ckks_keygen(const ckks_rns_plan *plan, int kind, size_t q_limbs, const size_t *block_sizes, size_t n_blocks,
                unsigned base_log, const uint64_t *source, const int8_t *secret, const uint64_t *a,
                const int64_t *e, uint64_t *key_b, uint64_t *key_a) {
    return keygen(plan, kind, q_limbs, block_sizes, n_blocks, base_log, source, secret, a, e, key_b, key_a);
}

int ckks_decrypt_raw(const ckks_rns_plan *plan, const uint64_t *b, const uint64_t *a, const int8_t *secret,
                     uint64_t *out, size_t n_polys, size_t n_limbs, size_t poly_stride, size_t out_stride) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    return rlwe_raw(plan, b, poly_stride, nullptr, secret, a, poly_stride, out, out_stride, n_polys,
                    n_limbs, false);
}

// ExecutionProfile count of a successful call (the plan is valid then).
static int counted(int rc, const ckks_rns_plan *plan, uint64_t ckks_execution_profile::*field, size_t n) {
    if (rc == CKKS_OK && n) plan->ctx->counts().*field += n;
    return rc;
}

int ckks_multiply_plain(const ckks_rns_plan *plan, const uint64_t *in, const uint64_t *pt, uint64_t *out,
                        size_t n_cts, size_t n_limbs, size_t in_stride, size_t pt_stride, size_t out_stride) {
    return counted(multiply_plain(plan, in, pt, out, n_cts, n_limbs, in_stride, pt_stride, out_stride), plan,
                   &ckks_execution_profile::ciphertext_plaintext_multiplies, n_cts);
}

int ckks_poly_mul(const ckks_rns_plan *plan, const uint64_t *x, const uint64_t *y, uint64_t *out,
                  size_t n_polys, size_t n_limbs, size_t poly_stride) {
    return elementwise(EW_MUL, plan, x, y, out, n_polys, n_limbs, poly_stride);
}
int ckks_poly_mac(const ckks_rns_plan *plan, uint64_t *acc, const uint64_t *x, const uint64_t *y,
                  size_t n_polys, size_t n_limbs, size_t poly_stride) {
    return elementwise(EW_MAC, plan, x, y, acc, n_polys, n_limbs, poly_stride);
}
int ckks_poly_add(const ckks_rns_plan *plan, const uint64_t *x, const uint64_t *y, uint64_t *out,
                  size_t n_polys, size_t n_limbs, size_t poly_stride) {
    return counted(elementwise(EW_ADD, plan, x, y, out, n_polys, n_limbs, poly_stride), plan,
                   &ckks_execution_profile::additions, n_polys);
}
int ckks_poly_sub(const ckks_rns_plan *plan, const uint64_t *x, const uint64_t *y, uint64_t *out,
                  size_t n_polys, size_t n_limbs, size_t poly_stride) {
    return counted(elementwise(EW_SUB, plan, x, y, out, n_polys, n_limbs, poly_stride), plan,
                   &ckks_execution_profile::additions, n_polys);
}
int ckks_poly_neg(const ckks_rns_plan *plan, const uint64_t *x, uint64_t *out, size_t n_polys,
                  size_t n_limbs, size_t poly_stride) {
    return elementwise(EW_NEG, plan, x, nullptr, out, n_polys, n_limbs, poly_stride);
}

int ckks_multiply_plain_accumulate_ntt(const ckks_rns_plan *plan, const uint64_t *const *cts,
                                       const uint64_t *const *pts, size_t count, uint64_t *out, size_t n_limbs) {
    int rc = check_plan_limbs(plan, n_limbs, "plaintext multiply-accumulate");
    if (rc) return rc;
    if (!cts || !pts || !out || count == 0) return fail(CKKS_ERR_INVALID, "null argument");
    if (count > kMaxMacTerms) return fail(CKKS_ERR_UNSUPPORTED, "plaintext multiply-accumulate of at most 64 terms");
    MacListArgs a{};
    for (size_t t = 0; t < count; ++t) {
        if (!cts[t] || !pts[t]) return fail(CKKS_ERR_INVALID, "null buffer");
        a.ct[t] = cts[t];
        a.pt[t] = pts[t];
    }
    a.out = out;
    a.primes = plan->d_primes;
    a.count = (u32)count;
    a.n_limbs = (u32)n_limbs;
    a.logn = plan->logn;
    a.chunks = (u32)(plan->degree / kEwChunk);
    hipLaunchKernelGGL(k_mac_list, dim3((u32)(2 * n_limbs * a.chunks)), dim3(kEwThreads), 0, plan->ctx->stream, a);
    CK_HIP(hipGetLastError());
    plan->ctx->counts().ciphertext_plaintext_multiplies += count;
    plan->ctx->counts().additions += 2 * (count - 1);
    return CKKS_OK;
}

int ckks_ntt_mul_monomial(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys,
                          size_t n_limbs, size_t poly_stride, int64_t exponent) {
    if (!plan) return fail(CKKS_ERR_INVALID, "null plan");
    const int64_t period = 2 * (int64_t)plan->degree;
    const int64_t e = ((exponent % period) + period) % period;  // rem_euclid
    return monomial(true, plan, in, out, n_polys, n_limbs, poly_stride, (u64)e, 1);
}

int ckks_mul_monomial(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys,
                      size_t n_limbs, size_t poly_stride, uint64_t exponent, int sign) {
    return monomial(false, plan, in, out, n_polys, n_limbs, poly_stride, exponent, sign);
}

int ckks_poly_scalar_mul(const ckks_rns_plan *plan, const uint64_t *in, const uint64_t *scalars, uint64_t *out,
                         size_t n_polys, size_t n_limbs, size_t poly_stride) {
    return scalar_mul(plan, in, scalars, out, n_polys, n_limbs, poly_stride);
}

int ckks_automorphism(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys,
                      size_t n_limbs, size_t poly_stride, uint64_t exponent) {
    return counted(permute(false, plan, in, out, n_polys, n_limbs, poly_stride, exponent), plan,
                   &ckks_execution_profile::automorphisms, n_polys);
}

int ckks_ntt_galois_permute(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                            size_t n_polys, size_t n_limbs, size_t poly_stride, uint64_t exponent) {
    return counted(permute(true, plan, in, out, n_polys, n_limbs, poly_stride, exponent), plan,
                   &ckks_execution_profile::automorphisms, n_polys);
}

int ckks_rescale(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_cts,
                 size_t n_limbs, size_t in_stride, size_t out_stride) {
    return counted(rescale_like(true, plan, in, out, n_cts, n_limbs, in_stride, out_stride), plan,
                   &ckks_execution_profile::rescales, n_cts);
}

int ckks_mod_switch(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_cts,
                    size_t n_limbs, size_t in_stride, size_t out_stride) {
    return counted(rescale_like(false, plan, in, out, n_cts, n_limbs, in_stride, out_stride), plan,
                   &ckks_execution_profile::modulus_switches, n_cts);
}

int ckks_tensor(const ckks_rns_plan *plan, const uint64_t *ct0, const uint64_t *ct1, uint64_t *out,
                size_t n_cts, size_t n_limbs, size_t in_stride, size_t out_stride) {
    return counted(tensor(plan, ct0, ct1, out, n_cts, n_limbs, in_stride, out_stride), plan,
                   &ckks_execution_profile::ciphertext_ciphertext_multiplies, n_cts);
}

}  // extern "C"
