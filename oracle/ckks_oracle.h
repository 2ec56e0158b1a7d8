/*
 * ckks_oracle.h -- CPU oracle for the leveled RNS-CKKS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * Rust algorithms (Confidential-Intelligence-Lab/ccmm-rs, src/ring, src/grafting,
 * src/ckks).  It is the oracle the HIP product path is compared against; it
 * is never linked into, loaded by, or called from the product library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Conventions (all follow the reference):
 *   - residues are canonical uint64_t in [0, q);
 *   - an RNS polynomial is limb-major: limb i occupies [i*N, (i+1)*N);
 *   - an RLWE ciphertext is (b, a) and decrypts as b + a*s;
 *   - NTT-domain index k holds the evaluation at psi^(2k+1) (natural order).
 */
#ifndef CKKS_ORACLE_H
#define CKKS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- R1: scalar modular arithmetic (ring/modulus.rs) -------------------- */
uint64_t orc_add(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_sub(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_neg(uint64_t a, uint64_t q);
uint64_t orc_mul(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_pow(uint64_t base, uint64_t exponent, uint64_t q);
uint64_t orc_inv(uint64_t a, uint64_t q);
uint64_t orc_barrett_reciprocal_hi(uint64_t q);
uint64_t orc_barrett_reciprocal_lo(uint64_t q);
uint64_t orc_barrett_mul(uint64_t a, uint64_t b, uint64_t q);
uint64_t orc_shoup_quotient(uint64_t w, uint64_t q);
uint64_t orc_shoup_mul(uint64_t x, uint64_t w, uint64_t w_quotient, uint64_t q);

/* ---- R14: NTT-prime search (ring/ntt_prime.rs) --------------------------- */
int orc_is_prime(uint64_t value);
size_t orc_ntt_primes(unsigned bits, size_t degree, size_t count,
                      const uint64_t *skip, size_t n_skip, uint64_t *out);

/* ---- R2: negacyclic root (ring/ntt_params.rs) ---------------------------- */
uint64_t orc_find_negacyclic_root(uint64_t q, size_t degree);

/* ---- R3: O(N^2) definition transforms (ring/ntt.rs) ---------------------- */
void orc_ntt_forward_reference(const uint64_t *in, uint64_t *out, size_t n,
                               uint64_t q, uint64_t psi);
void orc_ntt_inverse_reference(const uint64_t *in, uint64_t *out, size_t n,
                               uint64_t q, uint64_t psi);

/* ---- R5: prepared radix-2 Shoup NTT (ring/prepared_ntt.rs) --------------- */
typedef struct orc_plan orc_plan;
orc_plan *orc_plan_new(uint64_t q, size_t degree, uint64_t psi /* 0: search */);
void orc_plan_free(orc_plan *plan);
uint64_t orc_plan_psi(const orc_plan *plan);
void orc_ntt_forward(const orc_plan *plan, const uint64_t *in, uint64_t *out);
void orc_ntt_inverse(const orc_plan *plan, const uint64_t *in, uint64_t *out);
/* Limb-major RNS transforms: `limbs` consecutive limbs, limb i with plans[i]. */
void orc_rns_ntt_forward(const orc_plan *const *plans, size_t limbs,
                         const uint64_t *in, uint64_t *out);
void orc_rns_ntt_inverse(const orc_plan *const *plans, size_t limbs,
                         const uint64_t *in, uint64_t *out);

/* ---- R7/R9: pointwise and coefficient ops (ring/ntt_polynomial.rs, polynomial.rs) */
void orc_pointwise_mul(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q);
void orc_pointwise_mac(uint64_t *acc, const uint64_t *x, const uint64_t *y, size_t n, uint64_t q);
void orc_poly_add(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q);
void orc_poly_sub(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q);
void orc_negacyclic_mul_naive(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q);
void orc_ntt_mul_monomial(const orc_plan *plan, const uint64_t *in, uint64_t *out, int64_t exponent);
void orc_mul_monomial(const uint64_t *in, uint64_t *out, size_t n, uint64_t q, uint64_t exponent, int sign);
void orc_scalar_mul(const uint64_t *in, uint64_t *out, size_t n, uint64_t q, uint64_t scalar);

/* ---- R8 / C1: Galois automorphisms ------------------------------------- */
void orc_permute_ntt_galois(const uint64_t *in, uint64_t *out, size_t n, uint64_t exponent);
void orc_apply_automorphism(const uint64_t *in, uint64_t *out, size_t n, uint64_t q, uint64_t exponent);
uint64_t orc_rotation_exponent_left(size_t degree, size_t steps);
uint64_t orc_rotation_exponent_right(size_t degree, size_t steps);
uint64_t orc_conjugation_exponent(size_t degree);
uint64_t orc_inverse_automorphism_exponent(size_t degree, uint64_t exponent);

/* ---- C5/C6: exact rescale and mod switch (ckks/rns_rescale.rs) ---------- */
void orc_rescale_limb(const uint64_t *src, const uint64_t *dropped, uint64_t *out,
                      size_t n, uint64_t q, uint64_t q_drop);
/* (b,a) over `limbs` moduli -> (b,a) over limbs-1 moduli, dropping the last. */
void orc_rescale_ct(const uint64_t *moduli, size_t limbs, size_t n,
                    const uint64_t *b, const uint64_t *a, uint64_t *b_out, uint64_t *a_out);

/* ---- R11: exact basis extension by mixed-radix digits (ring/basis_transition.rs) */
void orc_mixed_radix_digits(const uint64_t *residues, const uint64_t *moduli, size_t k,
                            uint64_t *digits);
uint64_t orc_eval_mixed_radix(const uint64_t *digits, const uint64_t *moduli, size_t k,
                              uint64_t target);
/* in: src_limbs x n residues over src_moduli; out: n_targets x n residues. */
void orc_extend_basis(const uint64_t *in, const uint64_t *src_moduli, size_t src_limbs,
                      const uint64_t *targets, size_t n_targets, size_t n, uint64_t *out);

/* ---- R12: exact centered mod-down by P (ring/auxiliary_modulus.rs) ------ */
/* in: (q_limbs + p_limbs) x n over Q then P; out: q_limbs x n. */
void orc_mod_down(const uint64_t *in, const uint64_t *q_moduli, size_t q_limbs,
                  const uint64_t *p_moduli, size_t p_limbs, size_t n, uint64_t *out);

/* ---- G1: RNS gadget decomposition (grafting/rns_gadget.rs) -------------- */
/* Digit j = canonical x mod Q_j lifted to every limb of the full basis.
 * out: n_blocks x limbs x n.  Blocks must satisfy Q_j < 2^128. */
int orc_rns_gadget_decompose(const uint64_t *in, const uint64_t *moduli, size_t limbs,
                             const size_t *block_sizes, size_t n_blocks, size_t n,
                             uint64_t *out);

/* ---- G2: bounded balanced gadget (grafting/bounded_gadget.rs) ----------- */
/* digit_count = ceil((bits(Q) + 1) / base_log), Q = product of moduli
 * (BoundedGadgetLayout::new, :18-37).  base_log in 1..63. */
size_t orc_bounded_digit_count(const uint64_t *moduli, size_t limbs, unsigned base_log);
/* decompose / decompose_canonical (:59-122): balanced base-2^w digits of the
 * centered value of each coefficient (limbs x n input), digit in
 * [-2^(w-1), 2^(w-1)); out: digit_count x n.  Returns 0, or -1 if the digit
 * count is insufficient (the reference's assert). */
int orc_bounded_decompose(const uint64_t *in, const uint64_t *moduli, size_t limbs, size_t n,
                          unsigned base_log, int64_t *digits);
/* lift_digit (:382-402): signed digit rem_euclid q. */
void orc_lift_signed(const int64_t *digits, size_t n, uint64_t q, uint64_t *out);

/* ---- K6 / C3: bounded key switch (grafting/bounded_key_switch.rs:139-219;
 * ckks/hybrid_rns_galois_key.rs:550-780).  plans/qp_moduli: q_limbs Q limbs
 * then p_limbs P limbs (p_limbs = 0: no P, K6 form).  Digits of a over Q;
 * key: digit_count x (q_limbs+p_limbs) x n NTT domain.  K6: b_out = b +
 * sum d_j kb_j, a_out = sum d_j ka_j.  C3: the QP sums start from P*b on the
 * Q limbs (0 on P), then both components are moved down by P. */
int orc_bounded_key_switch(const orc_plan *const *plans, const uint64_t *qp_moduli,
                           size_t q_limbs, size_t p_limbs, unsigned base_log,
                           const uint64_t *b, const uint64_t *a,
                           const uint64_t *key_b, const uint64_t *key_a,
                           uint64_t *b_out, uint64_t *a_out);

/* ---- K2: tensor (grafting/rns_key_switch.rs rns_tensor_with_prepared_ntt) */
void orc_tensor(const orc_plan *const *plans, size_t limbs,
                const uint64_t *b0, const uint64_t *a0, const uint64_t *b1, const uint64_t *a1,
                uint64_t *c0, uint64_t *c1, uint64_t *c2);

/* ---- K3: RNS-gadget key switch without P --------------------------------
 * key_b/key_a: n_blocks x limbs x n, NTT domain.  out = (b + sum d_j kb_j,
 * sum d_j ka_j), coefficient domain. */
int orc_rns_key_switch(const orc_plan *const *plans, const uint64_t *moduli, size_t limbs,
                       const size_t *block_sizes, size_t n_blocks,
                       const uint64_t *b, const uint64_t *a,
                       const uint64_t *key_b, const uint64_t *key_a,
                       uint64_t *b_out, uint64_t *a_out);

/* ---- K4: dnum-block hybrid key switch with special primes P ------------
 * plans: (q_limbs + p_limbs) plans, Q first then P.  key_b/key_a:
 * n_blocks x (q_limbs+p_limbs) x n, NTT domain, encrypting P*E_j*s'.
 * b/a: q_limbs x n coefficient domain. */
int orc_hybrid_key_switch(const orc_plan *const *plans, const uint64_t *qp_moduli,
                          size_t q_limbs, size_t p_limbs,
                          const size_t *block_sizes, size_t n_blocks,
                          const uint64_t *b, const uint64_t *a,
                          const uint64_t *key_b, const uint64_t *key_a,
                          uint64_t *b_out, uint64_t *a_out);

/* ---- E1: raw RLWE encryption/decryption with caller-sampled inputs ------
 * b = m + e - a*s per limb (NTT product); decrypt = b + a*s. */
void orc_encrypt(const orc_plan *const *plans, const uint64_t *moduli, size_t limbs,
                 const uint64_t *message, const int64_t *error, const int8_t *secret,
                 const uint64_t *a, uint64_t *b_out);
void orc_decrypt(const orc_plan *const *plans, const uint64_t *moduli, size_t limbs,
                 const uint64_t *b, const uint64_t *a, const int8_t *secret, uint64_t *out);

/* ---- Host samplers: ChaCha20 stream seeded from u64 --------------------- */
typedef struct orc_rng { uint32_t key[8]; uint64_t counter; uint32_t buf[64]; size_t idx; } orc_rng;
void orc_rng_seed(orc_rng *rng, uint64_t seed);
uint32_t orc_rng_next_u32(orc_rng *rng);
uint64_t orc_rng_next_u64(orc_rng *rng);
void orc_chacha20_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]);
void orc_sample_uniform(orc_rng *rng, uint64_t q, size_t n, uint64_t *out);
void orc_sample_gaussian(orc_rng *rng, double sigma, size_t n, int64_t *out);
void orc_sample_ternary(orc_rng *rng, size_t n, int8_t *out);
uint32_t orc_pcg32_xsh_rr(uint64_t state);
uint64_t orc_pcg32_step(uint64_t state, uint64_t inc);
void orc_rng_set_key(orc_rng *rng, const uint32_t key[8], uint64_t counter);
size_t orc_gaussian_table(double sigma, double *weights, size_t cap, double *total);
double orc_uniform_open_unit(orc_rng *rng);
void orc_ternary_census(uint64_t counts[4]);

#ifdef __cplusplus
}
#endif
#endif
