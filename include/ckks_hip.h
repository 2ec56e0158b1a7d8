/*
 * ckks_hip.h -- C ABI of the MI355X (gfx950) backend for the leveled
 * RNS-CKKS evaluation path of ccmm-rs ("FHE-rs").
 *
 * The reference crate is pure Rust with no FFI; every entry point below
 * replaces one Rust function of the hot path (cited as file:line into the
 * reference crate) and computes bit-identical residues.  A Rust binding
 * (`ccmm-rs-hip-sys`, a separate crate because the reference is
 * #![forbid(unsafe_code)]) maps the status codes back to the reference's
 * panic messages, which ckks_last_error() reproduces.
 *
 * Conventions
 *   - residues are canonical uint64_t in [0, q); every prime q < 2^62;
 *   - device buffers are plain device pointers (hipMalloc'd or any other
 *     device allocation); the caller owns host and device memory passed in,
 *     the library owns only its plans, keys and workspace;
 *   - polynomial batches are limb-major: polynomial p, limb i occupies
 *     [p * poly_stride + i * N, ... + N) in elements;
 *   - a ciphertext is (b, a) decrypting as b + a*s, stored as
 *     [b limbs][a limbs]: component c, limb i at (c * n_limbs + i) * N
 *     within the ciphertext, ciphertexts ct_stride elements apart;
 *   - NTT-domain index k holds the evaluation at psi^(2k+1) (natural order,
 *     ring/ntt.rs:100-127);
 *   - calls are asynchronous on the context's stream; ckks_sync() waits;
 *   - a context is single-stream and thread-compatible: use one context per
 *     host thread.  Functions return CKKS_OK or an error code and never
 *     unwind across the ABI.
 */
#ifndef CKKS_HIP_H
#define CKKS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum ckks_status {
    CKKS_OK = 0,
    CKKS_ERR_INVALID = 1,     /* argument violates a reference assertion      */
    CKKS_ERR_HIP = 2,         /* HIP runtime error                            */
    CKKS_ERR_OOM = 3,         /* device allocation failed / budget exceeded   */
    CKKS_ERR_UNSUPPORTED = 4, /* outside the GPU path (e.g. q >= 2^62)        */
    CKKS_ERR_BUSY = 5         /* serving queue full                           */
};

/* Message of the last failed call on this host thread.  Where the reference
 * panics with a fixed message the same text is used. */
const char *ckks_last_error(void);

/* Library and device identification. */
const char *ckks_version(void);

/* ---------------------------------------------------------------- context */
typedef struct ckks_context ckks_context;

/* stream: the hipStream_t to run on; NULL means the device's null
 * (default) stream, the one torch uses by default.  The context owns every
 * plan and key created from it: ckks_context_destroy releases those still
 * alive, after which their handles must not be used. */
int ckks_context_create(int device, void *stream, ckks_context **out);
int ckks_context_destroy(ckks_context *ctx);
int ckks_context_set_stream(ckks_context *ctx, void *stream);
int ckks_sync(ckks_context *ctx);

/* Tuning: bytes of the NTT inter-pass buffer per chunk of a batched
 * transform (default 0 = the whole batch in one chunk).  Results do not
 * depend on it.  No reference counterpart. */
int ckks_set_ntt_chunk_bytes(ckks_context *ctx, size_t bytes);

/* Tuning: kernel variant of the two-pass NTT (N > 4096) for one degree
 * 2^logn, a bit set: CKKS_NTT_VARIANT_BLOCKED stores the inter-pass buffer
 * in chunks shared by one column tile and one row tile;
 * CKKS_NTT_VARIANT_TILE8K uses 8192-residue tiles (512 threads; N >= 2^14,
 * ignored below).  Every variant gives identical residues.  Default 0.
 * No reference counterpart. */
enum {
    CKKS_NTT_VARIANT_BLOCKED = 1,
    CKKS_NTT_VARIANT_TILE8K = 2,
    CKKS_NTT_VARIANT_COUNT = 4
};
int ckks_set_ntt_variant(ckks_context *ctx, uint32_t logn, int variant);
int ckks_get_ntt_variant(const ckks_context *ctx, uint32_t logn, int *variant);

/* Tuning: how a key switch runs the forward NTT of its digits and the key
 * inner product.  CKKS_KS_FUSION_OFF: separate NTT and MAC kernels.
 * CKKS_KS_FUSION_XCD / CKKS_KS_FUSION_LINEAR: the digits' row pass is
 * fused with the MAC (the NTT-domain digits never reach HBM), with the
 * blocks sharing a key tile placed on one XCD or in plain order.
 * CKKS_KS_FUSION_INVERSE: the MAC is computed inside the first inverse NTT
 * pass (the NTT-domain accumulators never reach HBM; dnum <= 8); set on the
 * context, it also fuses the MAC of hoisted rotations of degrees above
 * 2^12 into their Galois-read first inverse pass (off otherwise: slower
 * than the separate MAC, which reads each key entry once per batch).
 * CKKS_KS_FUSION_AUTO (default): the choice ckks_autotune_key_switch made
 * for the key, OFF for an untuned key.  Identical residues in every mode.
 * No reference counterpart. */
enum {
    CKKS_KS_FUSION_AUTO = -1,
    CKKS_KS_FUSION_OFF = 0,
    CKKS_KS_FUSION_XCD = 1,
    CKKS_KS_FUSION_LINEAR = 2,
    CKKS_KS_FUSION_INVERSE = 3
};
int ckks_set_key_switch_fusion(ckks_context *ctx, int mode);

/* Tuning: a key switch of at least 8 ciphertexts may run its two halves on
 * the context stream and on an auxiliary stream of the context (default
 * priority), joined before the call returns, so the VALU-bound base
 * conversions of one half overlap the HBM-bound NTT and MAC of the other.
 * CKKS_KS_SPLIT_AUTO (default): the choice ckks_autotune_key_switch made
 * for the key, OFF for an untuned key.  Never during graph capture.
 * Identical residues either way.  No reference counterpart. */
enum { CKKS_KS_SPLIT_AUTO = -1, CKKS_KS_SPLIT_OFF = 0, CKKS_KS_SPLIT_ON = 1 };
int ckks_set_key_switch_split(ckks_context *ctx, int mode);

/* Tuning: engine of the exact ModUp/ModDown base conversion of hybrid keys.
 * CKKS_BASECONV_MFMA uses the int8 matrix cores (7-bit chunks, int32
 * accumulation) wherever every gadget block has 2..9 limbs;
 * CKKS_BASECONV_VALU uses the vector-ALU kernels; CKKS_BASECONV_AUTO (the
 * default) uses the engine ckks_autotune_key_switch picked for the key, or
 * VALU for an untuned key.  All engines are exact and give identical
 * residues.  No reference counterpart. */
enum { CKKS_BASECONV_AUTO = 0, CKKS_BASECONV_VALU = 1, CKKS_BASECONV_MFMA = 2 };
int ckks_set_base_conversion(ckks_context *ctx, int mode);

/* Device memory helpers for callers without their own allocator. */
int ckks_malloc(ckks_context *ctx, size_t bytes, void **dev);
int ckks_free(ckks_context *ctx, void *dev);
int ckks_upload(ckks_context *ctx, void *dev, const void *host, size_t bytes);
int ckks_download(ckks_context *ctx, void *host, const void *dev, size_t bytes);

/* ---------------------------------------------------------------- ring */
typedef struct ckks_rns_plan ckks_rns_plan;

/* find_ntt_prime_below_bits (ring/ntt_prime.rs:67-94) continued downward:
 * the first `count` primes q = k*2N + 1 below 2^bits (bits <= 62), skipping
 * the n_skip values in skip.  Writes them to out and returns how many were
 * found through *found.  Host-only; needs no context. */
int ckks_find_ntt_primes(unsigned bits, size_t degree, size_t count, const uint64_t *skip,
                         size_t n_skip, uint64_t *out, size_t *found);

/* PreparedRnsNttPlan::new (ring/prepared_rns_ntt.rs:8-37) over
 * RnsNttPlan::new (ring/rns_ntt.rs:4-40): one negacyclic plan per modulus,
 * psi chosen as find_negacyclic_root (ring/ntt_params.rs:16-50).
 * degree: 2^10 .. 2^17; every modulus prime, < 2^62, 2N | q-1. */
int ckks_rns_plan_create(ckks_context *ctx, const uint64_t *moduli, size_t n_limbs,
                         size_t degree, ckks_rns_plan **out);
int ckks_rns_plan_destroy(ckks_rns_plan *plan);
/* psi of every limb (NttPlan::psi, ring/ntt.rs:87-89). */
int ckks_rns_plan_psi(const ckks_rns_plan *plan, uint64_t *psi_out);

/* PreparedRnsNttPlan::forward / inverse (ring/prepared_rns_ntt.rs:39-84,
 * per limb PreparedNttPlan::forward prepared_ntt.rs:200 and ::inverse
 * prepared_ntt.rs:370) over n_polys polynomials of the first n_limbs limbs.
 * in may equal out. */
int ckks_ntt_forward(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                     size_t n_polys, size_t n_limbs, size_t poly_stride);
int ckks_ntt_inverse(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                     size_t n_polys, size_t n_limbs, size_t poly_stride);
/* No counterpart in the reference: ckks_ntt_forward for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and residues as
 * ckks_ntt_forward; it allocates nothing: the inter-pass buffer of degrees
 * above 4096 is a fixed 32 MiB array each device holds for the life of the
 * process, and the batch runs in chunks of at most 32 MiB of limb
 * transforms (64 at N = 2^16).  Every call on a device shares that array, so
 * calls that overlap on one device (other streams, other contexts, other
 * host threads) race; calls on one stream are ordered.  The array needs no
 * clearing: each chunk writes it fully before reading it. */
int ckks_ntt_forward_noalloc(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                             size_t n_polys, size_t n_limbs, size_t poly_stride);

/* PreparedNttPlan::new / forward / inverse (ring/prepared_ntt.rs:34,200,370)
 * for one prime of any width up to 64 bits, the reference's 64-bit
 * parameter set included (ring/ntt_parameter_sets.rs:11-12: q = 2^64 -
 * 114687, psi at N = 4096).  degree 2^10 .. 2^17 with 2N | q-1; psi as
 * find_negacyclic_root.  Residues stay canonical in every butterfly (the RNS
 * plan's lazy butterflies need q < 2^62), so this plan serves moduli the RNS
 * plan refuses.  in may equal out; n_polys polynomials of `degree`
 * residues each, contiguous. */
typedef struct ckks_ntt64_plan ckks_ntt64_plan;
int ckks_ntt64_plan_create(ckks_context *ctx, uint64_t q, size_t degree, ckks_ntt64_plan **out);
int ckks_ntt64_plan_destroy(ckks_ntt64_plan *plan);
int ckks_ntt64_plan_psi(const ckks_ntt64_plan *plan, uint64_t *psi);
int ckks_ntt64_forward(const ckks_ntt64_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys);
int ckks_ntt64_inverse(const ckks_ntt64_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys);

/* Mixed-degree / mixed-level batch in one launch per pass (at most 16
 * entries; every entry's plan must belong to the same context). */
typedef struct ckks_ntt_batch {
    const ckks_rns_plan *plan;
    const uint64_t *in;
    uint64_t *out;
    size_t n_polys;
    size_t n_limbs;
    size_t poly_stride;
} ckks_ntt_batch;
int ckks_ntt_forward_batch(ckks_context *ctx, const ckks_ntt_batch *batch, size_t count);
int ckks_ntt_inverse_batch(ckks_context *ctx, const ckks_ntt_batch *batch, size_t count);

/* Autotuner: times every NTT variant on a forward + inverse transform of
 * n_polys x n_limbs limbs of `plan` (scratch buffers allocated and freed
 * here; synchronises the context stream), keeps the fastest for the plan's
 * degree in the context and returns it.  ms_per_variant (optional, length
 * CKKS_NTT_VARIANT_COUNT) receives the measured time of each variant in
 * milliseconds, -1 for a variant that does not apply.  Not callable during
 * graph capture.  No reference counterpart. */
int ckks_autotune_ntt(const ckks_rns_plan *plan, size_t n_polys, size_t n_limbs,
                      int *best_variant, float *ms_per_variant);

/* NttPolynomial::pointwise_mul / pointwise_mul_add_assign / add / sub
 * (ring/ntt_polynomial.rs:160-270) and Polynomial::{add,sub,neg}
 * (ring/polynomial.rs); the same arithmetic serves both domains.
 * out may alias inputs. */
int ckks_poly_mul(const ckks_rns_plan *plan, const uint64_t *x, const uint64_t *y,
                  uint64_t *out, size_t n_polys, size_t n_limbs, size_t poly_stride);
int ckks_poly_mac(const ckks_rns_plan *plan, uint64_t *acc, const uint64_t *x, const uint64_t *y,
                  size_t n_polys, size_t n_limbs, size_t poly_stride);
int ckks_poly_add(const ckks_rns_plan *plan, const uint64_t *x, const uint64_t *y,
                  uint64_t *out, size_t n_polys, size_t n_limbs, size_t poly_stride);
int ckks_poly_sub(const ckks_rns_plan *plan, const uint64_t *x, const uint64_t *y,
                  uint64_t *out, size_t n_polys, size_t n_limbs, size_t poly_stride);
int ckks_poly_neg(const ckks_rns_plan *plan, const uint64_t *x, uint64_t *out,
                  size_t n_polys, size_t n_limbs, size_t poly_stride);

/* NttPolynomial::mul_monomial_signed (ring/ntt_polynomial.rs:133-175) and
 * RnsNttPolynomial::mul_monomial_signed (ring/rns_ntt.rs:215-235): NTT-domain
 * product by X^exponent, evaluation k scaled by psi^((2k+1) e mod 2N) with
 * e = exponent rem_euclid 2N.  out may alias in. */
int ckks_ntt_mul_monomial(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys,
                          size_t n_limbs, size_t poly_stride, int64_t exponent);
/* Polynomial::mul_monomial_signed (ring/polynomial.rs:164-200) per limb:
 * coefficient-domain sign * X^exponent * x (exponent reduced mod 2N, sign 1
 * or -1).  out must not alias in. */
int ckks_mul_monomial(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_polys,
                      size_t n_limbs, size_t poly_stride, uint64_t exponent, int sign);
/* Polynomial::scalar_mul (ring/polynomial.rs:144-155) per limb: limb i times
 * scalars[i] mod q_i (scalars: host array of n_limbs values, any u64; at
 * most 128 limbs).  Either domain.  out may alias in. */
int ckks_poly_scalar_mul(const ckks_rns_plan *plan, const uint64_t *in, const uint64_t *scalars, uint64_t *out,
                         size_t n_polys, size_t n_limbs, size_t poly_stride);
/* apply_automorphism (ckks/automorphism.rs:8-46) per limb, coefficient
 * domain; exponent odd modulo 2N.  out must not alias in. */
int ckks_automorphism(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                      size_t n_polys, size_t n_limbs, size_t poly_stride, uint64_t exponent);
/* permute_negacyclic_ntt_galois (ring/ntt_galois_permutation.rs:5-19),
 * NTT domain.  out must not alias in. */
int ckks_ntt_galois_permute(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out,
                            size_t n_polys, size_t n_limbs, size_t poly_stride,
                            uint64_t exponent);

/* ------------------------------------------------------------ ciphertexts */

/* rescale_rns_ckks_to_next -> rescale_rns_rlwe -> rescale_polynomial_limb
 * (ckks/rns_rescale.rs:19-167): n_cts ciphertexts over the first n_limbs
 * limbs (coefficient domain) to n_limbs-1 limbs.  in_stride/out_stride are
 * the ciphertext strides in elements.  out must not alias in. */
int ckks_rescale(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_cts,
                 size_t n_limbs, size_t in_stride, size_t out_stride);

/* mod_switch_rns_ckks_to_next (ckks/rns_mod_switch.rs:14-59): drop the
 * trailing limb of both components. */
int ckks_mod_switch(const ckks_rns_plan *plan, const uint64_t *in, uint64_t *out, size_t n_cts,
                    size_t n_limbs, size_t in_stride, size_t out_stride);

/* multiply_plain_rns_ckks_with_prepared_ntt (ckks/rns_plain_multiply.rs:
 * 193-260) over PreparedRnsPlaintextNtt (:76-110): ciphertexts
 * (coefficient domain) times NTT-domain plaintexts of n_limbs limbs;
 * plaintext c is at pt + c * pt_stride (pt_stride 0: one plaintext for the
 * whole batch).  The CKKS scale bookkeeping (scale product) stays with the
 * caller.  out may equal in. */
int ckks_multiply_plain(const ckks_rns_plan *plan, const uint64_t *in, const uint64_t *pt,
                        uint64_t *out, size_t n_cts, size_t n_limbs, size_t in_stride,
                        size_t pt_stride, size_t out_stride);
/* multiply_ntt_resident_rlwe_with_prepared_plaintext(_resident)
 * (ckks/rns_plain_multiply.rs:119-191) summed with
 * pointwise_mul_add_assign (ring/ntt_polynomial.rs:222-235): the
 * plaintext-diagonal inner sum of a baby-step / giant-step schedule in one
 * launch,  out = sum_{t < count} cts[t] * pts[t]  over NTT-domain
 * ciphertexts ([2][n_limbs][N] each) and NTT-domain plaintexts ([n_limbs][N]),
 * limb by limb.  cts and pts are host arrays of device pointers, count in
 * 1..64; out may be one of the cts. */
int ckks_multiply_plain_accumulate_ntt(const ckks_rns_plan *plan, const uint64_t *const *cts,
                                       const uint64_t *const *pts, size_t count, uint64_t *out, size_t n_limbs);

/* encrypt_rns_raw_with_distribution_ntt_rng (grafting/rns_key_switch.rs:
 * 245-383) after its sampling, which stays on the host (ChaCha20, Gaussian
 * table): b = m + e - a*s over n_limbs limbs for n_polys messages.  message,
 * a and b are [p][limb][N] with poly_stride elements between polynomials (0:
 * n_limbs * N); error is [p][N] signed (one logical error per message,
 * projected to every limb as rlwe/error.rs:150-166); secret is N signed
 * coefficients (ternary).  All device pointers. */
int ckks_encrypt_raw(const ckks_rns_plan *plan, const uint64_t *message, const int64_t *error,
                     const int8_t *secret, const uint64_t *a, uint64_t *b, size_t n_polys,
                     size_t n_limbs, size_t poly_stride);

/* decrypt_rns_raw_with_ntt (grafting/rns_key_switch.rs:906-947):
 * out = b + a*s, [p][limb][N]; b and a with poly_stride, out with
 * out_stride (0: n_limbs * N). */
int ckks_decrypt_raw(const ckks_rns_plan *plan, const uint64_t *b, const uint64_t *a,
                     const int8_t *secret, uint64_t *out, size_t n_polys, size_t n_limbs,
                     size_t poly_stride, size_t out_stride);

/* auxiliary_mod_down_rns_candidate (ring/auxiliary_modulus.rs:89-171):
 * n_polys consecutive polynomials [p][qp limbs][N] over qp_plan (q_limbs Q
 * limbs then the special primes P) to [p][q_limbs][N]:
 * round(centered(x) / P) mod Q, centered by x_P > floor(P/2).  Tables are
 * built on the first call per q_limbs (not capturable). */
int ckks_mod_down(const ckks_rns_plan *qp_plan, size_t q_limbs, const uint64_t *in, uint64_t *out,
                  size_t n_polys);

/* encode_rns (application_support/ckks.rs:39-66), integer half: real
 * coefficients [p][N] (device doubles, from the canonical embedding) times
 * scale, rounded half away from zero (f64::round), rem_euclid every limb ->
 * out [p][limb][N].  CKKS_ERR_INVALID if a rounded value reaches 2^63.
 * Synchronises the context stream (overflow check). */
int ckks_encode_coefficients(const ckks_rns_plan *plan, const double *coeffs, double scale,
                             uint64_t *out, size_t n_polys, size_t n_limbs);

/* Decode (bin/r30_13b_ckks_primitives.rs:55-80), integer half: exact CRT of
 * [p][limb][N] residues over the first n_limbs limbs to the centered integer
 * (value > floor(Q/2) -> negative, ring/wide_crt.rs), converted to FP64 and
 * divided by scale -> out [p][N] doubles.  The CRT tables are built on the
 * first call per limb count (not capturable) and kept by the plan. */
int ckks_decode_coefficients(const ckks_rns_plan *plan, const uint64_t *in, double scale,
                             double *out, size_t n_polys, size_t n_limbs);

/* rns_tensor_with_prepared_ntt (grafting/rns_key_switch.rs:1724-1777):
 * (b0,a0) x (b1,a1) -> (c0,c1,c2) coefficient domain, out laid out as
 * [c0 limbs][c1 limbs][c2 limbs] per ciphertext (out_stride elements). */
int ckks_tensor(const ckks_rns_plan *plan, const uint64_t *ct0, const uint64_t *ct1,
                uint64_t *out, size_t n_cts, size_t n_limbs, size_t in_stride,
                size_t out_stride);

/* ------------------------------------------------------------ key switching */
typedef struct ckks_ksk ckks_ksk;

/* PreparedRnsKeySwitchKey::new / PreparedRnsMultiplicationKey
 * (grafting/rns_key_switch.rs:1930-2013, 2347-2430): NTT-domain key
 * entries, block-major [block][limb] pairs (b, a), copied from key_b,
 * key_a of n_blocks * n_limbs * N residues each (host arrays or device
 * buffers; device buffers are read on the context stream and checked there).
 * RnsGadgetLayout::new (grafting/rns_gadget.rs:37-67) block sizes. */
int ckks_rns_ksk_create(const ckks_rns_plan *plan, const size_t *block_sizes, size_t n_blocks,
                        const uint64_t *key_b, const uint64_t *key_a, ckks_ksk **out);

/* dnum-block hybrid key with special primes P.  qp_plan holds Q limbs
 * followed by P limbs; q_limbs of them are Q.  Entry j encrypts
 * P * E_j * s' over QP in the NTT domain, [block][qp limb] pairs. */
int ckks_hybrid_ksk_create(const ckks_rns_plan *qp_plan, size_t q_limbs, const size_t *block_sizes,
                           size_t n_blocks, const uint64_t *key_b, const uint64_t *key_a,
                           ckks_ksk **out);
/* Bounded balanced base-2^w gadget key (BoundedGadgetLayout,
 * grafting/bounded_gadget.rs:18-37; digit_count = ceil((bits(Q)+1)/w),
 * base_log w in 1..63).  plan holds the Q limbs (q_limbs of them) followed
 * by the special primes P, if any.  Without P: BoundedRnsMultiplicationKey
 * (grafting/bounded_key_switch.rs:27-103), entry j encrypting B^j * s^2
 * (or any B^j * t).  With P: PreparedHybridRnsGaloisKey
 * (ckks/hybrid_rns_galois_key.rs:61-170), entry j encrypting
 * P * B^j * sigma(s).  Entries are NTT-domain [digit][limb] pairs (b, a) of
 * digit_count * plan limbs * N residues each.  The key switch decomposes
 * the centered value of its input into the digits, lifts each to every limb
 * (lift_digit, :382-402), and runs NTT, key MAC, iNTT and (with P) the
 * ModDown; the P * b base term of the reference is added after the ModDown,
 * which gives the same residues. */
int ckks_bounded_ksk_create(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                            const uint64_t *key_b, const uint64_t *key_a, ckks_ksk **out);

/* Key generation on the device from host-sampled randomness, for the three
 * key kinds above:
 *   RnsKeySwitchKey::generate_with_distribution_ntt_rng and
 *   RnsMultiplicationKey::generate_with_distribution_ntt_rng
 *     (grafting/rns_key_switch.rs:1155-1258, 1430-1512): CKKS_KEYGEN_RNS_GADGET;
 *   the dnum-block hybrid key of ckks_hybrid_ksk_create: CKKS_KEYGEN_HYBRID;
 *   PreparedHybridRnsGaloisKey::generate_with_distribution_ntt_rng
 *     (ckks/hybrid_rns_galois_key.rs:61-170) and BoundedRnsMultiplicationKey
 *     (grafting/bounded_key_switch.rs:22-118): CKKS_KEYGEN_BOUNDED.
 * Each entry is encrypt_rns_raw_with_distribution_ntt_rng
 * (grafting/rns_key_switch.rs:245-383) of a gadget multiple of `source`
 * under `secret`, in the NTT domain:
 *   key_b[j] = NTT(g_j * source + e_j - a_j * secret),  key_a[j] = NTT(a_j)
 * limb by limb over the plan's basis, where g_j is
 *   RNS gadget: 1 on the limbs of block j, 0 elsewhere (q_limbs = plan limbs);
 *   hybrid:     P on the limbs of block j, 0 elsewhere (P = product of the
 *               plan limbs after the first q_limbs);
 *   bounded:    P * 2^(base_log * j) on every limb (P = 1 without special
 *               primes), j < ckks_bounded_digit_count.
 * The reference samples a_j (uniform per limb) and e_j (one logical error
 * per entry) from its RNG inside; here the caller samples them.  All
 * buffers are device memory: source [limb][N] coefficient domain (sigma_k(s)
 * for a Galois key, s^2 for a relinearization key), secret [N] ternary,
 * a [entries][limb][N] canonical, e [entries][N], outputs key_b and key_a
 * [entries][limb][N], which ckks_*_ksk_create accept as device pointers.
 * Checks the secret (ternary, nonzero) and a (canonical) and synchronises
 * the stream: not capturable. */
enum ckks_keygen_kind { CKKS_KEYGEN_RNS_GADGET = 0, CKKS_KEYGEN_HYBRID = 1, CKKS_KEYGEN_BOUNDED = 2 };
int ckks_keygen(const ckks_rns_plan *plan, int kind, size_t q_limbs, const size_t *block_sizes, size_t n_blocks,
                unsigned base_log, const uint64_t *source, const int8_t *secret, const uint64_t *a,
                const int64_t *e, uint64_t *key_b, uint64_t *key_a);
int ckks_bounded_digit_count(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                             size_t *count);

/* BoundedGadgetLayout::decompose + lift_digit (grafting/bounded_gadget.rs:
 * 59-122, 382-402): n_polys polynomials over the first q_limbs limbs of
 * plan (poly p at in + p * in_stride, in_stride 0 = q_limbs * N) to
 * out[p][digit][limb][N] over every limb of the plan.  Synchronises the
 * context stream to upload its tables; not capturable. */
int ckks_bounded_decompose(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                           const uint64_t *in, uint64_t *out, size_t n_polys, size_t in_stride);

/* BoundedGadgetLayout::decompose (grafting/bounded_gadget.rs:59-118): the
 * balanced digits themselves, out[p][digit][N] signed, in
 * [-2^(base_log-1), 2^(base_log-1)).  Arguments as ckks_bounded_decompose;
 * synchronises the context stream to upload its tables. */
int ckks_bounded_digits(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log, const uint64_t *in,
                        int64_t *out, size_t n_polys, size_t in_stride);

/* BoundedGadgetDecomposition::automorphism_hoisted (grafting/
 * bounded_gadget.rs:273-312): sigma_exponent applied to decomposed digits
 * [p][digit][N] (from ckks_bounded_digits with the same q_limbs and
 * base_log), negated digit vectors renormalised with a carry, which equals
 * the decomposition of the rotated polynomial.  out must not alias in. */
int ckks_bounded_digits_automorphism(const ckks_rns_plan *plan, size_t q_limbs, unsigned base_log,
                                     const int64_t *in, int64_t *out, size_t n_polys, uint64_t exponent);

/* Hoisted rotation with a bounded key: ckks_apply_galois with the
 * decomposition of the a components taken from a_digits [ct][digit][N]
 * (ckks_bounded_digits of the unrotated a, computed once and reused for
 * every exponent) through the hoisted digit automorphism, as
 * PreparedCanonicalGaloisSource (grafting/bounded_gadget.rs:176-249) and
 * automorphism_hoisted do for the reference.  Same residues as
 * ckks_apply_galois; the a limbs of in are not read. */
int ckks_bounded_apply_galois_hoisted(const ckks_ksk *key, uint64_t exponent, const uint64_t *in,
                                      const int64_t *a_digits, uint64_t *out, size_t n_cts, size_t in_stride,
                                      size_t out_stride);

/* Hoisted rotations in the NTT domain (the use PreparedCanonicalGaloisSource,
 * grafting/bounded_gadget.rs:176-249, serves: many rotations of one
 * ciphertext).  ckks_bounded_hoist decomposes the a components of n_cts
 * ciphertexts once, lifts the digits to the key basis and transforms them
 * (device memory owned by the handle; synchronises the context stream).
 * ckks_bounded_rotate_hoisted then rotates the same ciphertexts (in: their
 * b components are read) with any bounded key of the same plan, Q limbs and
 * base_log: sigma_k of a digit is a permutation in the NTT domain, read by
 * the key MAC, so no decomposition or forward NTT runs.  A digit vector
 * holding -2^(base_log-1) carries when negated; when sigma_k negates one,
 * the call takes the ckks_apply_galois path.  Same residues as
 * ckks_apply_galois either way.  The first call with a given key and
 * exponent builds a copy of that key permuted by exponent^-1 (as much device
 * memory as the key; at most 8 such copies per key, least recently used
 * dropped first; refused under graph capture; without device memory for it
 * the call takes the ckks_apply_galois path), so calls with one key from
 * several host threads must be serialised by the caller.  Destroy the handle
 * before the context. */
typedef struct ckks_hoisted ckks_hoisted;
int ckks_bounded_hoist(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride,
                       ckks_hoisted **out);
int ckks_bounded_rotate_hoisted(const ckks_hoisted *hoisted, const ckks_ksk *key, uint64_t exponent,
                                const uint64_t *in, uint64_t *out, size_t in_stride, size_t out_stride);
int ckks_bounded_hoist_destroy(ckks_hoisted *hoisted);

/* Hoisted rotations with RNS-gadget and dnum-hybrid keys, residue for
 * residue those of ckks_apply_galois (rotate_left_rns_ckks_with_prepared_ntt,
 * ckks/rns_galois_key.rs:730; the reference hoists bounded digits only).
 * ckks_hoist cuts the canonical digits of the UNROTATED a components once
 * and transforms them; ckks_rotate_hoisted(k) runs the key MAC on them
 * permuted by sigma_k (read through the key permuted by k^-1), so no
 * decomposition or forward NTT runs per rotation.  The canonical digit of a
 * negated coefficient is Q_j - L, not -L: the difference, (Q_j mod p_i) at
 * every position sigma_k negates, enters as a per-key, per-exponent
 * correction T = NTT(neg_k) * sum_j (Q_j mod p_i) key_j added to every
 * ciphertext's MAC sums.  It holds unless a coefficient has every residue
 * of a gadget block zero; ckks_hoist checks for that (one pass over the a
 * components and a host synchronisation) and such a batch, like a call
 * without device memory for the per-exponent tables, runs the canonical
 * key switch.  For bounded keys both calls are ckks_bounded_hoist /
 * ckks_bounded_rotate_hoisted.  The key must have the plan, Q limbs and
 * gadget blocks the digits were cut with; the permuted key copies and
 * corrections are cached per exponent as for bounded keys.  Destroy the
 * handle with ckks_bounded_hoist_destroy, before the context. */
int ckks_hoist(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride, ckks_hoisted **out);
/* No reference counterpart: ckks_hoist that also keeps P NTT(b) of the b
 * components (Q limbs; P the product of the key's special primes, 1 for an
 * RNS-gadget key).  The reference adds P b to the key-switch sums before
 * ModDown; rotations from this handle add P NTT(b) inside the key MAC,
 * where the Galois-read inverse NTT turns it into P sigma_k(b), so ModDown
 * reads no sigma_k gather of b.  Same residues as ckks_hoist, provided the
 * in argument of each rotation still holds the hoisted ciphertexts (b is
 * taken as it was at the hoist; a batch that takes the canonical path reads
 * it from in).  The MAC adds it for dnum <= 8 (larger dnum: b read from in
 * as for ckks_hoist).  Not for bounded keys; ckks_rehoist refreshes it. */
int ckks_hoist_with_b(const ckks_ksk *key, const uint64_t *in, size_t n_cts, size_t in_stride,
                      ckks_hoisted **out);
/* No reference counterpart: ckks_hoist of a new batch of the same count and
 * shape into an existing handle (RNS-gadget / dnum keys), reusing its
 * device buffers, for loops that hoist batch after batch without
 * allocating; the handle's digits of the previous batch are replaced. */
int ckks_rehoist(ckks_hoisted *hoisted, const ckks_ksk *key, const uint64_t *in, size_t in_stride);
int ckks_rotate_hoisted(const ckks_hoisted *hoisted, const ckks_ksk *key, uint64_t exponent, const uint64_t *in,
                        uint64_t *out, size_t in_stride, size_t out_stride);
/* Several hoisted rotations in one launch sequence (no reference
 * counterpart; the same residues as ckks_rotate_hoisted entry by entry):
 * the key MACs of entries that share a hoisted handle (one batch rotated by
 * several exponents) run up to 4 per launch, reading each digit once; the
 * inverse NTTs of all entries are one launch per pass (mixed degrees, each
 * entry read through its own Galois permutation), then ModDown per entry.
 * Entries of bounded keys, and entries that take the canonical path, run on
 * their own.  At most 16 entries, each with its own out; in and out follow
 * ckks_rotate_hoisted (strides 0 = dense). */
typedef struct ckks_hoisted_rotation {
    const ckks_hoisted *hoisted;
    const ckks_ksk *key;
    uint64_t exponent;
    const uint64_t *in;
    uint64_t *out;
    size_t in_stride, out_stride;
} ckks_hoisted_rotation;
int ckks_rotate_hoisted_batch(ckks_context *ctx, const ckks_hoisted_rotation *entries, size_t count);

int ckks_ksk_destroy(ckks_ksk *key);
/* LogicalPayloadBytes::logical_payload_bytes of a key (ckks/key_accounting.rs:
 * 13-120): the device bytes of its entries (2 * entries * limbs * N * 8) plus
 * the tables it caches for hoisted rotations (a permuted copy per exponent;
 * canonical-digit keys add 2 * limbs * N * 8 of correction per exponent),
 * the figure the serving loop charges while a request reads the key. */
int ckks_ksk_bytes(const ckks_ksk *key, size_t *bytes);

/* rns_key_switch_with_prepared_dynamic_ntt (grafting/rns_key_switch.rs:
 * 2526-2608) for an RNS-gadget key, or the dnum hybrid key switch for a
 * hybrid key (ModUp -> NTT -> key MAC -> iNTT -> ModDown by P).
 * Coefficient-domain ciphertexts over the key's full Q basis in and out;
 * out may equal in. */
int ckks_key_switch(const ckks_ksk *key, const uint64_t *in, uint64_t *out, size_t n_cts,
                    size_t in_stride, size_t out_stride);

/* apply_rns_galois_automorphism_with_prepared_ntt (ckks/rns_galois_key.rs:
 * 439-469) / rotate_left_rns_ckks_with_prepared_ntt (:730-751): sigma_k on
 * both components then key switch with a key for sigma_k(s) -> s. */
int ckks_apply_galois(const ckks_ksk *key, uint64_t exponent, const uint64_t *in, uint64_t *out,
                      size_t n_cts, size_t in_stride, size_t out_stride);

/* Galois key switches of several ciphertext batches, each with its own key,
 * in one launch sequence: the batched form of
 * apply_hybrid_rns_galois_automorphism_* / apply_rns_galois_automorphism_*
 * (ckks/hybrid_rns_galois_key.rs:202-780, ckks/rns_galois_key.rs:439-469),
 * for ciphertexts of uneven ring degree and level.  Every entry's residues
 * equal ckks_apply_galois on it alone.  Decomposition, key MAC and ModDown
 * run per entry; the forward and inverse NTTs of all entries run as one
 * launch per pass.  At most 16 entries, keys of this context. */
typedef struct ckks_galois_batch {
    const ckks_ksk *key;
    uint64_t exponent;  /* odd modulo 2N */
    const uint64_t *in; /* [n_cts] ciphertexts over the key's Q limbs */
    uint64_t *out;      /* must not alias in */
    size_t n_cts, in_stride, out_stride;  /* strides 0: 2 * q_limbs * N */
} ckks_galois_batch;
int ckks_apply_galois_batch(ckks_context *ctx, const ckks_galois_batch *entries, size_t count);

/* rns_relinearize_with_prepared_dynamic_ntt (grafting/rns_key_switch.rs:
 * 2113-2259): (c0,c1,c2) -> (c0,c1) + switch(c2). */
int ckks_relinearize(const ckks_ksk *key, const uint64_t *in3, uint64_t *out, size_t n_cts,
                     size_t in_stride, size_t out_stride);

/* multiply_relinearize_rescale_rns_ckks_with_prepared_ntt
 * (ckks/rns_multiply.rs:161-215): tensor, relinearize, rescale. */
int ckks_multiply_relinearize_rescale(const ckks_ksk *relin_key, const uint64_t *ct0,
                                      const uint64_t *ct1, uint64_t *out, size_t n_cts,
                                      size_t in_stride, size_t out_stride);

/* Autotuner: times one key switch of n_cts ciphertexts with each
 * base-conversion engine the key supports and each fusion mode (scratch
 * ciphertexts allocated and freed here; synchronises the context stream)
 * and records the fastest pair in the key for CKKS_BASECONV_AUTO and
 * CKKS_KS_FUSION_AUTO; then times that pair with the batch split over two
 * streams and records the faster for CKKS_KS_SPLIT_AUTO.  ms (optional, 2 x 4 floats) receives the times as
 * ms[engine - 1][fusion mode], engine VALU = 1 / MFMA = 2, -1 where a pair
 * does not apply.  Not callable during graph capture.  No reference
 * counterpart. */
int ckks_autotune_key_switch(ckks_ksk *key, size_t n_cts, int *best_engine, int *best_fusion,
                             float *ms);

/* ------------------------------------------------------------ serving */
/* Continuous-batching request loop on one GPU.  No reference counterpart
 * (the reference evaluates one job at a time on the CPU).  A job is a
 * straight-line program over ciphertext registers [2][limbs][N] (registers
 * 0.. n_inputs-1 start as copies of the inputs; `output` receives register
 * `out_reg` at the end), validated at submission (register use, levels, key bases).  Each
 * ckks_server_step retires finished requests, admits waiting ones in
 * arrival order while their footprint fits the HBM budget and the active
 * cap -- the registers (n_regs * 2 * n_limbs * N * 8 bytes) plus every key and
 * plaintext the program reads that no admitted request holds yet (a key or
 * plaintext is charged once while any admitted request reads it, and
 * released with its last reader) -- then runs the next op of
 * every active request, batching ops that share (op, key, exponent, plan,
 * level) into one launch.  Evaluation keys and plans stay resident and are
 * owned by the caller; they must outlive the jobs that name them.  All work
 * runs on the context stream. */
enum {
    CKKS_OP_ROTATE = 1,             /* dst = keyswitch(sigma_exponent(a)), Galois key */
    CKKS_OP_MUL_PLAIN = 2,          /* dst = a * plaintext (NTT domain, a's limbs) */
    CKKS_OP_ADD = 3,                /* dst = a + b */
    CKKS_OP_RESCALE = 4,            /* dst = rescale(a) */
    CKKS_OP_MUL_RELIN_RESCALE = 5,  /* dst = rescale(relin(a x b)), relinearization key */
    CKKS_OP_COPY = 6,               /* dst = a */
    CKKS_OP_MOD_SWITCH = 7,         /* dst = a without its last limb */
    CKKS_OP_SUB = 8,                /* dst = a - b */
    CKKS_OP_NTT = 9,                /* dst = NTT(a): the register moves to the NTT domain */
    CKKS_OP_INTT = 10,              /* dst = NTT^-1(a): back to the coefficient domain */
    CKKS_OP_MAC_PLAIN = 11          /* dst = sum_{t < b} reg[a + t] * pt_t (NTT domain); `plaintext`
                                       points to a host array of b device pointers to NTT-domain
                                       plaintexts, copied at submission; 1 <= b <= 64 */
};
/* Registers hold coefficient-domain ciphertexts except after CKKS_OP_NTT or
 * CKKS_OP_MAC_PLAIN; CKKS_OP_INTT and CKKS_OP_MAC_PLAIN read NTT-domain
 * registers, every other op reads coefficient-domain ones (add, sub and copy
 * keep their operands' domain); the output register ends coefficient domain. */
typedef struct ckks_op {
    uint32_t code, dst, a, b;
    const ckks_ksk *key;
    uint64_t exponent;
    const uint64_t *plaintext;
} ckks_op;
typedef struct ckks_job {
    const ckks_rns_plan *plan; /* basis whose first n_limbs limbs hold the input */
    uint32_t n_limbs, n_regs, out_reg;
    uint32_t n_inputs;         /* registers 0.. n_inputs-1 <- input[i] (0 means 1) */
    const uint64_t *input;     /* [n_inputs][2][n_limbs][N] */
    uint64_t *output;
    const ckks_op *ops;
    size_t n_ops;
} ckks_job;
typedef struct ckks_server_stats {
    uint64_t submitted, completed, rejected, waiting, active, peak_active, launches, steps;
    uint64_t peak_bytes;
    double p50_ms, p99_ms, mean_ms, max_ms; /* submit -> completion observed */
    uint64_t cancelled, failed;             /* see ckks_server_cancel / poll state 4 */
    uint64_t resident_bytes;                /* keys and plaintexts charged now */
} ckks_server_statistics;
typedef struct ckks_server ckks_server;
int ckks_server_create(ckks_context *ctx, size_t hbm_budget_bytes, size_t max_active, ckks_server **out);
/* CKKS_ERR_INVALID for a malformed program, CKKS_ERR_OOM for a job whose
 * footprint exceeds the whole budget; the op list is copied. */
int ckks_server_submit(ckks_server *server, const ckks_job *job, uint64_t *ticket);
int ckks_server_step(ckks_server *server, size_t *launched);
/* state: 0 waiting, 1 admitted, 2 done (latency_ms set, else -1),
 * 3 cancelled, 4 failed: its registers could not be allocated while no other
 * request was in flight (device memory exhausted outside the budget). */
int ckks_server_poll(ckks_server *server, uint64_t ticket, int *state, double *latency_ms);
/* Cancels a waiting request (it leaves the queue) or an admitted one (no
 * further op is issued; its registers and resident charges are released at
 * the next step; the output is not written).  CKKS_ERR_INVALID for an unknown
 * ticket or a request that already finished. */
int ckks_server_cancel(ckks_server *server, uint64_t ticket);
int ckks_server_drain(ckks_server *server);
int ckks_server_stats(ckks_server *server, ckks_server_statistics *out);
int ckks_server_destroy(ckks_server *server);

/* ----------------------------------------------------- execution profile */
/* ExecutionProfile (execution/mod.rs:162-230) observed on the device: the
 * context counts the operations its entry points run, per ciphertext (per
 * polynomial for NTTs, additions and Galois permutations), in the
 * reference's ExecutionEventKind vocabulary:
 *   ciphertext_plaintext_multiplies  ckks_multiply_plain (per ciphertext)
 *   ciphertext_ciphertext_multiplies ckks_tensor (also inside
 *                                    ckks_multiply_relinearize_rescale)
 *   additions                        ckks_poly_add / ckks_poly_sub (per polynomial)
 *   relinearizations                 ckks_relinearize, ckks_multiply_relinearize_rescale
 *   rescales / modulus_switches      ckks_rescale / ckks_mod_switch
 *   automorphisms                    ckks_automorphism, ckks_ntt_galois_permute
 *                                    (per polynomial); key-switched conjugations
 *   rotations                        key-switched Galois automorphisms other
 *                                    than conjugation (ckks_apply_galois, hoisted)
 *   ntt_forwards / ntt_inverses      RNS polynomial transforms, including those
 *                                    inside key switching, encryption and keygen
 * and two counts the reference has no field for: key_switches (every key
 * switch, per ciphertext) and ntt_limb_transforms (single-limb transforms
 * behind the NTT counts).  A captured graph adds the counts of the calls it
 * recorded each time it is launched (the capture itself counts nothing). */
typedef struct ckks_execution_profile {
    uint64_t ciphertext_plaintext_multiplies, ciphertext_ciphertext_multiplies, additions, relinearizations;
    uint64_t rescales, modulus_switches, automorphisms, rotations, ntt_forwards, ntt_inverses;
    uint64_t key_switches, ntt_limb_transforms;
} ckks_execution_profile;
int ckks_profile_get(const ckks_context *ctx, ckks_execution_profile *out);
int ckks_profile_reset(ckks_context *ctx);

/* ------------------------------------------------------------ graphs */
/* Rotation schedules (BSGS of CCMM / FFT) captured once as a hipGraph and
 * replayed: ckks_graph_begin starts capturing the context's stream (which
 * must not be the null stream); every library call until ckks_graph_end is
 * recorded, not run.  Workspaces must already be large enough: run the
 * schedule once eagerly first (a call that would have to grow a workspace
 * during capture fails).  A replay reads and writes the buffers the calls
 * named at capture time.  No reference counterpart (the reference runs the
 * same schedule as a plain loop). */
typedef struct ckks_graph ckks_graph;
int ckks_graph_begin(ckks_context *ctx);
int ckks_graph_end(ckks_context *ctx, ckks_graph **out);
int ckks_graph_launch(ckks_graph *graph);
int ckks_graph_destroy(ckks_graph *graph);

#ifdef __cplusplus
}
#endif
#endif
