/*
 * ckks_oracle.c -- CPU oracle for the leveled RNS-CKKS hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see ckks_oracle.h).  Every function restates the
 * reference Rust algorithm named in its comment (paths are relative to the
 * reference crate root, Confidential-Intelligence-Lab/ccmm-rs).  The HIP product
 * never links or calls this file.
 */
#include "ckks_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ======================================================================
 * R1  src/ring/modulus.rs
 * ==================================================================== */

/* Modulus::add_canonical (modulus.rs:158-178), generalised to any inputs
 * by a first reduction as Modulus::add does (modulus.rs:140-143). */
uint64_t orc_add(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)(((u128)(a % q) + (u128)(b % q)) % q);
}

/* Modulus::sub (modulus.rs:146-152). */
uint64_t orc_sub(uint64_t a, uint64_t b, uint64_t q) {
    u128 aa = a % q, bb = b % q;
    return (uint64_t)((aa + q - bb) % q);
}

/* Modulus::neg (modulus.rs:196-204). */
uint64_t orc_neg(uint64_t a, uint64_t q) {
    uint64_t r = a % q;
    return r == 0 ? 0 : q - r;
}

/* Modulus::mul (modulus.rs:207-211). */
uint64_t orc_mul(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)(((u128)a * b) % q);
}

/* Modulus::pow (modulus.rs:229-244). */
uint64_t orc_pow(uint64_t base, uint64_t exponent, uint64_t q) {
    uint64_t b = base % q, r = 1 % q;
    while (exponent > 0) {
        if (exponent & 1) r = orc_mul(r, b, q);
        b = orc_mul(b, b, q);
        exponent >>= 1;
    }
    return r;
}

/* Modulus::inverse_prime (modulus.rs:256-262): Fermat. */
uint64_t orc_inv(uint64_t a, uint64_t q) {
    uint64_t r = a % q;
    if (r == 0) abort(); /* "zero has no multiplicative inverse" */
    return orc_pow(r, q - 2, q);
}

/* PreparedModulus::new (modulus.rs:20-37): floor(2^128 / q). */
static u128 barrett_reciprocal(uint64_t q) {
    u128 max = ~(u128)0;
    u128 rec = max / q;
    if (max % q == (u128)q - 1) rec += 1;
    return rec;
}
uint64_t orc_barrett_reciprocal_hi(uint64_t q) { return (uint64_t)(barrett_reciprocal(q) >> 64); }
uint64_t orc_barrett_reciprocal_lo(uint64_t q) { return (uint64_t)barrett_reciprocal(q); }

/* PreparedModulus::mul_canonical (modulus.rs:53-89). */
uint64_t orc_barrett_mul(uint64_t a, uint64_t b, uint64_t q) {
    u128 rec = barrett_reciprocal(q);
    u128 x = (u128)a * b;
    u128 x_hi = x >> 64, x_lo = (uint64_t)x;
    u128 r_hi = rec >> 64, r_lo = (uint64_t)rec;
    u128 p0 = x_lo * r_lo, p1 = x_lo * r_hi, p2 = x_hi * r_lo, p3 = x_hi * r_hi;
    u128 mask = ((u128)1 << 64) - 1;
    u128 carry = ((p0 >> 64) + (p1 & mask) + (p2 & mask)) >> 64;
    u128 quotient = p3 + (p1 >> 64) + (p2 >> 64) + carry;
    u128 reduced = x - quotient * q;
    while (reduced >= q) reduced -= q;
    return (uint64_t)reduced;
}

/* ShoupMultiplier::new (modulus.rs:99-106). */
uint64_t orc_shoup_quotient(uint64_t w, uint64_t q) {
    return (uint64_t)(((u128)w << 64) / q);
}

/* ShoupMultiplier::mul (modulus.rs:108-124). */
uint64_t orc_shoup_mul(uint64_t x, uint64_t w, uint64_t wq, uint64_t q) {
    uint64_t est = (uint64_t)(((u128)x * wq) >> 64);
    uint64_t r = x * w - est * q;
    return r >= q ? r - q : r;
}

/* ======================================================================
 * R14  src/ring/ntt_prime.rs
 * ==================================================================== */

static int mr_round(uint64_t n, unsigned s, uint64_t d, uint64_t witness) {
    uint64_t a = witness % n;
    if (a == 0) return 1;
    uint64_t x = orc_pow(a, d, n);
    if (x == 1 || x == n - 1) return 1;
    for (unsigned r = 1; r < s; ++r) {
        x = orc_mul(x, x, n);
        if (x == n - 1) return 1;
    }
    return 0;
}

/* is_ntt_prime_candidate for values < 2^64 (ntt_prime.rs:8-33): small-prime
 * screen then deterministic Miller-Rabin with the 7-witness set. */
int orc_is_prime(uint64_t value) {
    static const uint64_t small[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    static const uint64_t wit[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    if (value < 2) return 0;
    for (int i = 0; i < 12; ++i) {
        if (value == small[i]) return 1;
        if (value % small[i] == 0) return 0;
    }
    uint64_t d = value - 1;
    unsigned s = 0;
    while ((d & 1) == 0) { d >>= 1; ++s; }
    for (int i = 0; i < 7; ++i)
        if (!mr_round(value, s, d, wit[i])) return 0;
    return 1;
}

/* find_ntt_prime_below_bits (ntt_prime.rs:67-94), continued downward to
 * collect `count` primes q = k*2N + 1 below 2^bits, skipping `skip`. */
size_t orc_ntt_primes(unsigned bits, size_t degree, size_t count,
                      const uint64_t *skip, size_t n_skip, uint64_t *out) {
    if (bits < 2 || bits > 63 || degree < 2 || (degree & (degree - 1))) return 0;
    uint64_t two_n = 2 * (uint64_t)degree;
    uint64_t cand = ((uint64_t)1 << bits) - 1;
    cand -= (cand - 1) % two_n;
    size_t found = 0;
    while (found < count) {
        if (cand > 2 && (cand - 1) % two_n == 0 && orc_is_prime(cand)) {
            int skipped = 0;
            for (size_t i = 0; i < n_skip; ++i) skipped |= (skip[i] == cand);
            if (!skipped) out[found++] = cand;
        }
        if (cand <= two_n + 1) break;
        cand -= two_n;
    }
    return found;
}

/* ======================================================================
 * R2  src/ring/ntt_params.rs:16-50  find_negacyclic_root
 * ==================================================================== */
uint64_t orc_find_negacyclic_root(uint64_t q, size_t degree) {
    uint64_t two_n = 2 * (uint64_t)degree;
    if ((q - 1) % two_n != 0) abort(); /* "2N must divide q - 1" */
    uint64_t e = (q - 1) / two_n;
    for (uint64_t g = 2; g < q; ++g) {
        uint64_t psi = orc_pow(g, e, q);
        if (orc_pow(psi, degree, q) == q - 1 && orc_pow(psi, two_n, q) == 1) return psi;
    }
    abort(); /* "no primitive 2N-th root found" */
}

/* ======================================================================
 * R3  src/ring/ntt.rs:106-158  O(N^2) reference transforms
 * ==================================================================== */
void orc_ntt_forward_reference(const uint64_t *in, uint64_t *out, size_t n, uint64_t q,
                               uint64_t psi) {
    for (size_t k = 0; k < n; ++k) {
        uint64_t point = orc_pow(psi, 2 * k + 1, q), power = 1, sum = 0;
        for (size_t j = 0; j < n; ++j) {
            sum = orc_add(sum, orc_mul(in[j], power, q), q);
            power = orc_mul(power, point, q);
        }
        out[k] = sum;
    }
}

void orc_ntt_inverse_reference(const uint64_t *in, uint64_t *out, size_t n, uint64_t q,
                               uint64_t psi) {
    uint64_t psi_inv = orc_inv(psi, q), n_inv = orc_inv(n % q, q);
    for (size_t j = 0; j < n; ++j) {
        uint64_t sum = 0;
        for (size_t k = 0; k < n; ++k) {
            uint64_t e = (uint64_t)((2 * k + 1) * j);
            sum = orc_add(sum, orc_mul(in[k], orc_pow(psi_inv, e, q), q), q);
        }
        out[j] = orc_mul(sum, n_inv, q);
    }
}

/* ======================================================================
 * R5  src/ring/prepared_ntt.rs  PreparedNttPlan
 * ==================================================================== */
struct orc_plan {
    uint64_t q, psi, psi_inv, n_inv;
    size_t n, logn;
    int shoup;                       /* q < 2^63 (prepared_ntt.rs:54) */
    uint64_t *fwd_twist, *fwd_twist_sh;   /* psi^j and Shoup quotients       */
    uint64_t *inv_twist;                  /* psi^-j                          */
    uint64_t *fwd_stage, *fwd_stage_sh;   /* stage-major twiddles, N-1 total */
    uint64_t *inv_stage;
    size_t *swap_i, *swap_j;
    size_t n_swaps;
};

/* powers() and stage_twiddles() (prepared_ntt.rs:401-436). */
static void stage_twiddles(uint64_t q, size_t n, uint64_t root, uint64_t *dst) {
    size_t pos = 0;
    for (size_t len = 2; len <= n; len *= 2) {
        size_t half = len / 2;
        uint64_t step = orc_pow(root, n / len, q), tw = 1;
        for (size_t i = 0; i < half; ++i) {
            dst[pos++] = tw;
            tw = orc_mul(tw, step, q);
        }
    }
}

/* PreparedNttPlan::new (prepared_ntt.rs:35-124). */
orc_plan *orc_plan_new(uint64_t q, size_t degree, uint64_t psi) {
    if (degree == 0 || (degree & (degree - 1))) return NULL;
    if ((q - 1) % (2 * (uint64_t)degree) != 0) return NULL;
    orc_plan *p = (orc_plan *)calloc(1, sizeof(orc_plan));
    p->q = q;
    p->n = degree;
    while (((size_t)1 << p->logn) < degree) ++p->logn;
    p->psi = psi ? psi % q : orc_find_negacyclic_root(q, degree);
    if (orc_pow(p->psi, 2 * degree, q) != 1 || orc_pow(p->psi, degree, q) != q - 1) {
        free(p);
        return NULL;
    }
    p->psi_inv = orc_inv(p->psi, q);
    p->n_inv = orc_inv(degree % q, q);
    p->shoup = q < ((uint64_t)1 << 63);
    size_t n = degree;
    p->fwd_twist = malloc(n * 8);
    p->fwd_twist_sh = malloc(n * 8);
    p->inv_twist = malloc(n * 8);
    p->fwd_stage = malloc(n * 8);
    p->fwd_stage_sh = malloc(n * 8);
    p->inv_stage = malloc(n * 8);
    uint64_t v = 1, w = 1;
    for (size_t j = 0; j < n; ++j) {
        p->fwd_twist[j] = v;
        p->inv_twist[j] = w;
        v = orc_mul(v, p->psi, q);
        w = orc_mul(w, p->psi_inv, q);
    }
    uint64_t omega = orc_mul(p->psi, p->psi, q);
    uint64_t omega_inv = orc_mul(p->psi_inv, p->psi_inv, q);
    stage_twiddles(q, n, omega, p->fwd_stage);
    stage_twiddles(q, n, omega_inv, p->inv_stage);
    if (p->shoup) {
        for (size_t j = 0; j < n; ++j) p->fwd_twist_sh[j] = orc_shoup_quotient(p->fwd_twist[j], q);
        for (size_t j = 0; j + 1 < n; ++j) p->fwd_stage_sh[j] = orc_shoup_quotient(p->fwd_stage[j], q);
    }
    /* prepare_bit_reverse_swaps (prepared_ntt.rs:608-631) */
    p->swap_i = malloc(n * sizeof(size_t));
    p->swap_j = malloc(n * sizeof(size_t));
    size_t j = 0;
    for (size_t i = 1; i < n; ++i) {
        size_t bit = n >> 1;
        while (j & bit) { j ^= bit; bit >>= 1; }
        j ^= bit;
        if (i < j) { p->swap_i[p->n_swaps] = i; p->swap_j[p->n_swaps] = j; ++p->n_swaps; }
    }
    return p;
}

void orc_plan_free(orc_plan *p) {
    if (!p) return;
    free(p->fwd_twist); free(p->fwd_twist_sh); free(p->inv_twist);
    free(p->fwd_stage); free(p->fwd_stage_sh); free(p->inv_stage);
    free(p->swap_i); free(p->swap_j);
    free(p);
}

uint64_t orc_plan_psi(const orc_plan *p) { return p->psi; }

/* cyclic_ntt_shoup (prepared_ntt.rs:531-606): bit-reverse, then radix-2
 * DIT stages with Shoup twiddles and canonical add/sub. */
static void cyclic_ntt_shoup(uint64_t *v, const orc_plan *p) {
    uint64_t q = p->q;
    for (size_t s = 0; s < p->n_swaps; ++s) {
        uint64_t t = v[p->swap_i[s]]; v[p->swap_i[s]] = v[p->swap_j[s]]; v[p->swap_j[s]] = t;
    }
    size_t base = 0;
    for (size_t len = 2; len <= p->n; len *= 2) {
        size_t half = len / 2;
        for (size_t start = 0; start < p->n; start += len) {
            for (size_t o = 0; o < half; ++o) {
                uint64_t u = v[start + o];
                uint64_t x = orc_shoup_mul(v[start + o + half], p->fwd_stage[base + o],
                                           p->fwd_stage_sh[base + o], q);
                uint64_t sum = u + x;
                v[start + o] = sum >= q ? sum - q : sum;
                v[start + o + half] = u >= x ? u - x : q - (x - u);
            }
        }
        base += half;
    }
}

/* cyclic_ntt_prepared (prepared_ntt.rs:438-479): u128 multiply path. */
static void cyclic_ntt_prepared(uint64_t *v, const orc_plan *p, const uint64_t *stages) {
    uint64_t q = p->q;
    for (size_t s = 0; s < p->n_swaps; ++s) {
        uint64_t t = v[p->swap_i[s]]; v[p->swap_i[s]] = v[p->swap_j[s]]; v[p->swap_j[s]] = t;
    }
    size_t base = 0;
    for (size_t len = 2; len <= p->n; len *= 2) {
        size_t half = len / 2;
        for (size_t start = 0; start < p->n; start += len) {
            for (size_t o = 0; o < half; ++o) {
                uint64_t u = v[start + o];
                uint64_t x = orc_mul(v[start + o + half], stages[base + o], q);
                v[start + o] = orc_add(u, x, q);
                v[start + o + half] = u >= x ? u - x : q - (x - u);
            }
        }
        base += half;
    }
}

/* PreparedNttPlan::forward (prepared_ntt.rs:200-233). */
void orc_ntt_forward(const orc_plan *p, const uint64_t *in, uint64_t *out) {
    if (out != in) memmove(out, in, p->n * 8);
    if (p->shoup) {
        for (size_t j = 0; j < p->n; ++j)
            out[j] = orc_shoup_mul(out[j], p->fwd_twist[j], p->fwd_twist_sh[j], p->q);
        cyclic_ntt_shoup(out, p);
    } else {
        for (size_t j = 0; j < p->n; ++j) out[j] = orc_mul(out[j], p->fwd_twist[j], p->q);
        cyclic_ntt_prepared(out, p, p->fwd_stage);
    }
}

/* PreparedNttPlan::inverse (prepared_ntt.rs:370-398). */
void orc_ntt_inverse(const orc_plan *p, const uint64_t *in, uint64_t *out) {
    if (out != in) memmove(out, in, p->n * 8);
    cyclic_ntt_prepared(out, p, p->inv_stage);
    for (size_t j = 0; j < p->n; ++j) {
        if (p->shoup) {
            out[j] = orc_barrett_mul(out[j], p->n_inv, p->q);
            out[j] = orc_barrett_mul(out[j], p->inv_twist[j], p->q);
        } else {
            out[j] = orc_mul(out[j], p->n_inv, p->q);
            out[j] = orc_mul(out[j], p->inv_twist[j], p->q);
        }
    }
}

/* PreparedRnsNttPlan::{forward,inverse} (prepared_rns_ntt.rs:39-84): limb loop. */
void orc_rns_ntt_forward(const orc_plan *const *plans, size_t limbs, const uint64_t *in,
                         uint64_t *out) {
    for (size_t i = 0; i < limbs; ++i) {
        size_t n = plans[i]->n;
        orc_ntt_forward(plans[i], in + i * n, out + i * n);
    }
}

void orc_rns_ntt_inverse(const orc_plan *const *plans, size_t limbs, const uint64_t *in,
                         uint64_t *out) {
    for (size_t i = 0; i < limbs; ++i) {
        size_t n = plans[i]->n;
        orc_ntt_inverse(plans[i], in + i * n, out + i * n);
    }
}

/* ======================================================================
 * R7 / R9  ntt_polynomial.rs, polynomial.rs
 * ==================================================================== */

/* NttPolynomial::pointwise_mul (ntt_polynomial.rs; Modulus::mul_canonical). */
void orc_pointwise_mul(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_mul(x[i], y[i], q);
}

/* pointwise_mul_add_assign_prevalidated (ntt_polynomial.rs:222-235). */
void orc_pointwise_mac(uint64_t *acc, const uint64_t *x, const uint64_t *y, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; ++i) acc[i] = orc_add(acc[i], orc_mul(x[i], y[i], q), q);
}

void orc_poly_add(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_add(x[i], y[i], q);
}

void orc_poly_sub(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n, uint64_t q) {
    for (size_t i = 0; i < n; ++i) out[i] = orc_sub(x[i], y[i], q);
}

/* Polynomial::negacyclic_mul (polynomial.rs:210-234): O(N^2) oracle. */
void orc_negacyclic_mul_naive(const uint64_t *x, const uint64_t *y, uint64_t *out, size_t n,
                              uint64_t q) {
    uint64_t *acc = calloc(n, 8);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j) {
            uint64_t prod = orc_mul(x[i], y[j], q);
            size_t d = i + j;
            if (d < n) acc[d] = orc_add(acc[d], prod, q);
            else acc[d - n] = orc_sub(acc[d - n], prod, q);
        }
    memcpy(out, acc, n * 8);
    free(acc);
}

/* NttPolynomial::mul_monomial_signed (ntt_polynomial.rs): multiplying by
 * X^e scales evaluation k by psi^((2k+1) e). */
void orc_ntt_mul_monomial(const orc_plan *p, const uint64_t *in, uint64_t *out, int64_t exponent) {
    int64_t period = 2 * (int64_t)p->n;
    int64_t e = ((exponent % period) + period) % period;
    for (size_t k = 0; k < p->n; ++k) {
        uint64_t f = orc_pow(p->psi, (uint64_t)(((2 * (int64_t)k + 1) * e) % period), p->q);
        out[k] = orc_mul(in[k], f, p->q);
    }
}

/* Polynomial::mul_monomial_signed (ring/polynomial.rs:164-200): sign * X^e * x
 * in Z_q[X]/(X^N + 1) as a signed coefficient permutation; e reduced mod 2N. */
void orc_mul_monomial(const uint64_t *in, uint64_t *out, size_t n, uint64_t q, uint64_t exponent, int sign) {
    uint64_t e = exponent % (2 * (uint64_t)n);
    size_t shift = (size_t)(e % n);
    int exp_neg = e >= n, ext_neg = sign == -1;
    for (size_t src = 0; src < n; ++src) {
        size_t sum = src + shift;
        int wrap = sum >= n;
        size_t dst = wrap ? sum - n : sum;
        out[dst] = (exp_neg ^ ext_neg ^ wrap) ? orc_neg(in[src], q) : in[src];
    }
}

/* Polynomial::scalar_mul (ring/polynomial.rs:144-155): coefficient * scalar
 * mod q (Modulus::mul reduces the full product; the scalar may exceed q). */
void orc_scalar_mul(const uint64_t *in, uint64_t *out, size_t n, uint64_t q, uint64_t scalar) {
    for (size_t i = 0; i < n; ++i) out[i] = (uint64_t)(((unsigned __int128)in[i] * scalar) % q);
}

/* ======================================================================
 * R8 / C1  Galois automorphisms
 * ==================================================================== */

/* permute_negacyclic_ntt_galois (ring/ntt_galois_permutation.rs:5-19). */
void orc_permute_ntt_galois(const uint64_t *in, uint64_t *out, size_t n, uint64_t exponent) {
    uint64_t m = 2 * (uint64_t)n;
    uint64_t k = exponent % m;
    if ((k & 1) == 0) abort(); /* "Galois exponent must be odd" */
    for (size_t j = 0; j < n; ++j) {
        uint64_t rot = (uint64_t)(((u128)k * (2 * j + 1)) % m);
        out[j] = in[(rot - 1) / 2];
    }
}

/* apply_automorphism (ckks/automorphism.rs:8-46). */
void orc_apply_automorphism(const uint64_t *in, uint64_t *out, size_t n, uint64_t q,
                            uint64_t exponent) {
    uint64_t m = 2 * (uint64_t)n;
    uint64_t k = exponent % m;
    if ((k & 1) == 0) abort(); /* "CKKS automorphism exponent must be odd modulo 2N" */
    for (size_t i = 0; i < n; ++i) {
        uint64_t e = (uint64_t)(((u128)i * k) % m);
        if (e < n) out[e] = in[i];
        else out[e - n] = orc_neg(in[i], q);
    }
}

/* rotation_exponent_left / right, conjugation, inverse (automorphism.rs:51-240). */
uint64_t orc_rotation_exponent_left(size_t degree, size_t steps) {
    size_t slots = degree / 2;
    steps %= slots;
    uint64_t m = 2 * (uint64_t)degree, r = 1, b = 5 % m;
    while (steps) {
        if (steps & 1) r = (r * b) % m;
        b = (b * b) % m;
        steps >>= 1;
    }
    return r;
}

uint64_t orc_rotation_exponent_right(size_t degree, size_t steps) {
    size_t slots = degree / 2;
    steps %= slots;
    return orc_rotation_exponent_left(degree, (slots - steps) % slots);
}

uint64_t orc_conjugation_exponent(size_t degree) { return 2 * (uint64_t)degree - 1; }

uint64_t orc_inverse_automorphism_exponent(size_t degree, uint64_t exponent) {
    uint64_t m = 2 * (uint64_t)degree, k = exponent % m;
    for (uint64_t c = 1; c < m; c += 2)
        if ((k * c) % m == 1) return c;
    abort();
}

/* ======================================================================
 * C5  ckks/rns_rescale.rs:107-167  rescale_polynomial_limb
 * ==================================================================== */
void orc_rescale_limb(const uint64_t *src, const uint64_t *dropped, uint64_t *out, size_t n,
                      uint64_t q, uint64_t q_drop) {
    uint64_t inverse = orc_inv(q_drop % q, q);
    for (size_t i = 0; i < n; ++i) {
        uint64_t rem = dropped[i], rq;
        if (rem > q_drop / 2) {
            /* centered = rem - q_drop < 0: magnitude (q_drop - rem) */
            uint64_t mag = (q_drop - rem) % q;
            rq = mag == 0 ? 0 : q - mag;
        } else {
            rq = rem % q;
        }
        out[i] = orc_mul(orc_sub(src[i], rq, q), inverse, q);
    }
}

/* rescale_rns_rlwe (rns_rescale.rs:63-95): every limb but the trailing one. */
void orc_rescale_ct(const uint64_t *moduli, size_t limbs, size_t n, const uint64_t *b,
                    const uint64_t *a, uint64_t *b_out, uint64_t *a_out) {
    size_t last = limbs - 1;
    for (size_t i = 0; i < last; ++i) {
        orc_rescale_limb(b + i * n, b + last * n, b_out + i * n, n, moduli[i], moduli[last]);
        orc_rescale_limb(a + i * n, a + last * n, a_out + i * n, n, moduli[i], moduli[last]);
    }
}

/* ======================================================================
 * R11  ring/basis_transition.rs:72-138  mixed-radix basis extension
 * ==================================================================== */

/* mixed_radix_digits (basis_transition.rs:88-112).  `inv` (k x k, entry
 * [i*k + j] = (q_i mod q_j)^-1 mod q_j) holds the inverses the reference
 * recomputes per coefficient; NULL computes them on the fly. */
static void mixed_radix_digits_inv(const uint64_t *residues, const uint64_t *moduli, size_t k,
                                   const uint64_t *inv_table, uint64_t *digits) {
    for (size_t i = 0; i < k; ++i) digits[i] = residues[i];
    for (size_t i = 0; i < k; ++i)
        for (size_t j = i + 1; j < k; ++j) {
            uint64_t qj = moduli[j];
            uint64_t diff = orc_sub(digits[j], digits[i] % qj, qj);
            uint64_t inv = inv_table ? inv_table[i * k + j] : orc_inv(moduli[i] % qj, qj);
            digits[j] = orc_mul(diff, inv, qj);
        }
}

static uint64_t *mixed_radix_inverses(const uint64_t *moduli, size_t k) {
    uint64_t *t = calloc(k > 0 ? k * k : 1, 8);
    for (size_t i = 0; i < k; ++i)
        for (size_t j = i + 1; j < k; ++j) t[i * k + j] = orc_inv(moduli[i] % moduli[j], moduli[j]);
    return t;
}

void orc_mixed_radix_digits(const uint64_t *residues, const uint64_t *moduli, size_t k,
                            uint64_t *digits) {
    mixed_radix_digits_inv(residues, moduli, k, NULL, digits);
}

/* evaluate_mixed_radix_modulus (basis_transition.rs:114-138). */
uint64_t orc_eval_mixed_radix(const uint64_t *digits, const uint64_t *moduli, size_t k,
                              uint64_t target) {
    uint64_t value = 0, product = 1 % target;
    for (size_t i = 0; i < k; ++i) {
        value = orc_add(value, orc_mul(digits[i] % target, product, target), target);
        product = orc_mul(product, moduli[i] % target, target);
    }
    return value;
}

/* extend_coefficients_to_modulus (basis_transition.rs:72-86), over many targets. */
void orc_extend_basis(const uint64_t *in, const uint64_t *src_moduli, size_t src_limbs,
                      const uint64_t *targets, size_t n_targets, size_t n, uint64_t *out) {
    uint64_t *res = malloc(src_limbs * 8), *dig = malloc(src_limbs * 8);
    uint64_t *invt = mixed_radix_inverses(src_moduli, src_limbs);
    for (size_t c = 0; c < n; ++c) {
        for (size_t i = 0; i < src_limbs; ++i) res[i] = in[i * n + c];
        mixed_radix_digits_inv(res, src_moduli, src_limbs, invt, dig);
        for (size_t t = 0; t < n_targets; ++t)
            out[t * n + c] = orc_eval_mixed_radix(dig, src_moduli, src_limbs, targets[t]);
    }
    free(invt);
    free(res);
    free(dig);
}

/* ======================================================================
 * Minimal multiword unsigned integers (restating the num-bigint 0.4.8 uses
 * of ring/wide_crt.rs and ring/auxiliary_modulus.rs for P-side CRT).
 * ==================================================================== */
#define BIG_WORDS 48
typedef struct { uint64_t w[BIG_WORDS]; } big;

static void big_set_u64(big *x, uint64_t v) { memset(x, 0, sizeof *x); x->w[0] = v; }

static void big_mul_add_small(big *x, uint64_t m, uint64_t a) {
    u128 carry = a;
    for (int i = 0; i < BIG_WORDS; ++i) {
        u128 t = (u128)x->w[i] * m + carry;
        x->w[i] = (uint64_t)t;
        carry = t >> 64;
    }
}

static int big_cmp(const big *x, const big *y) {
    for (int i = BIG_WORDS - 1; i >= 0; --i) {
        if (x->w[i] != y->w[i]) return x->w[i] < y->w[i] ? -1 : 1;
    }
    return 0;
}

static void big_shr1(big *x) {
    for (int i = 0; i < BIG_WORDS; ++i)
        x->w[i] = (x->w[i] >> 1) | (i + 1 < BIG_WORDS ? x->w[i + 1] << 63 : 0);
}

/* value from mixed-radix digits: x = d0 + d1 m0 + d2 m0 m1 + ... (Horner). */
static void big_from_mixed_radix(big *x, const uint64_t *digits, const uint64_t *moduli, size_t k) {
    big_set_u64(x, 0);
    for (size_t i = k; i-- > 0;) big_mul_add_small(x, moduli[i], digits[i]);
}

static void big_product(big *x, const uint64_t *moduli, size_t k) {
    big_set_u64(x, 1);
    for (size_t i = 0; i < k; ++i) big_mul_add_small(x, moduli[i], 0);
}

/* ======================================================================
 * R12  ring/auxiliary_modulus.rs:89-171  auxiliary_mod_down_rns_candidate
 * r = centered(x mod P) (value > floor(P/2) -> negative, wide_crt.rs);
 * out_i = (x_i - r mod q_i) * P^-1 mod q_i.
 * ==================================================================== */
void orc_mod_down(const uint64_t *in, const uint64_t *q_moduli, size_t q_limbs,
                  const uint64_t *p_moduli, size_t p_limbs, size_t n, uint64_t *out) {
    big half_p, xp;
    big_product(&half_p, p_moduli, p_limbs);
    big_shr1(&half_p);
    uint64_t *p_inv = malloc(q_limbs * 8), *p_mod = malloc(q_limbs * 8);
    for (size_t i = 0; i < q_limbs; ++i) {
        uint64_t pm = 1;
        for (size_t k = 0; k < p_limbs; ++k) pm = orc_mul(pm, p_moduli[k] % q_moduli[i], q_moduli[i]);
        p_mod[i] = pm;
        p_inv[i] = orc_inv(pm, q_moduli[i]);
    }
    uint64_t *res = malloc(p_limbs * 8), *dig = malloc(p_limbs * 8);
    uint64_t *invt = mixed_radix_inverses(p_moduli, p_limbs);
    for (size_t c = 0; c < n; ++c) {
        for (size_t k = 0; k < p_limbs; ++k) res[k] = in[(q_limbs + k) * n + c];
        mixed_radix_digits_inv(res, p_moduli, p_limbs, invt, dig);
        big_from_mixed_radix(&xp, dig, p_moduli, p_limbs);
        int negative = big_cmp(&xp, &half_p) > 0;
        for (size_t i = 0; i < q_limbs; ++i) {
            uint64_t qi = q_moduli[i];
            uint64_t r = orc_eval_mixed_radix(dig, p_moduli, p_limbs, qi); /* x_P mod q_i */
            if (negative) r = orc_sub(r, p_mod[i], qi);                  /* x_P - P     */
            uint64_t diff = orc_sub(in[i * n + c], r, qi);
            out[i * n + c] = orc_mul(diff, p_inv[i], qi);
        }
    }
    free(invt); free(res); free(dig); free(p_inv); free(p_mod);
}

/* ======================================================================
 * G1  grafting/rns_gadget.rs:143-228  decompose_block_local
 * ==================================================================== */
int orc_rns_gadget_decompose(const uint64_t *in, const uint64_t *moduli, size_t limbs,
                             const size_t *block_sizes, size_t n_blocks, size_t n,
                             uint64_t *out) {
    size_t start = 0;
    for (size_t blk = 0; blk < n_blocks; ++blk) {
        size_t end = start + block_sizes[blk];
        if (end > limbs) return -1;
        /* prefix products and inverses within the block (u128) */
        u128 prefix[64];
        uint64_t inverses[64];
        u128 pre = moduli[start];
        for (size_t l = start + 1; l < end; ++l) {
            prefix[l - start - 1] = pre;
            inverses[l - start - 1] = orc_inv((uint64_t)(pre % moduli[l]), moduli[l]);
            u128 next = pre * moduli[l];
            if (next / moduli[l] != pre) return -2; /* "gadget block composite modulus exceeds u128" */
            pre = next;
        }
        for (size_t c = 0; c < n; ++c) {
            u128 value = in[start * n + c];
            for (size_t l = start + 1; l < end; ++l) {
                uint64_t ql = moduli[l];
                uint64_t vm = (uint64_t)(value % ql);
                uint64_t diff = orc_sub(in[l * n + c], vm, ql);
                uint64_t corr = orc_mul(diff, inverses[l - start - 1], ql);
                value += prefix[l - start - 1] * corr;
            }
            for (size_t i = 0; i < limbs; ++i)
                out[(blk * limbs + i) * n + c] = (uint64_t)(value % moduli[i]);
        }
        start = end;
    }
    return start == limbs ? 0 : -1;
}

/* ======================================================================
 * G2  grafting/bounded_gadget.rs:18-37 (layout), :75-122
 *     (decompose_canonical), :382-402 (lift_digit)
 * Signed multiword values are kept in two's complement over BIG_WORDS words.
 * ==================================================================== */
static unsigned big_bits(const big *x) {
    for (int i = BIG_WORDS - 1; i >= 0; --i)
        if (x->w[i]) return (unsigned)(64 * i + 64 - __builtin_clzll(x->w[i]));
    return 0;
}

static void big_sub(big *x, const big *y) { /* x -= y mod 2^(64 BIG_WORDS) */
    uint64_t borrow = 0;
    for (int i = 0; i < BIG_WORDS; ++i) {
        uint64_t a = x->w[i], b = y->w[i];
        uint64_t d = a - b - borrow;
        borrow = (a < b) || (a - b < borrow);
        x->w[i] = d;
    }
}

static void big_add_signed_small(big *x, int64_t v) { /* x += v, two's complement */
    uint64_t add = (uint64_t)v, ext = v < 0 ? ~0ull : 0;
    u128 carry = 0;
    for (int i = 0; i < BIG_WORDS; ++i) {
        u128 t = (u128)x->w[i] + (i == 0 ? add : ext) + carry;
        x->w[i] = (uint64_t)t;
        carry = t >> 64;
    }
}

static void big_sar(big *x, unsigned s) { /* arithmetic shift right, 0 < s < 64 */
    uint64_t sign = (x->w[BIG_WORDS - 1] >> 63) ? ~0ull : 0;
    for (int i = 0; i < BIG_WORDS; ++i) {
        uint64_t next = i + 1 < BIG_WORDS ? x->w[i + 1] : sign;
        x->w[i] = (x->w[i] >> s) | (next << (64 - s));
    }
}

static int big_is_zero(const big *x) {
    for (int i = 0; i < BIG_WORDS; ++i)
        if (x->w[i]) return 0;
    return 1;
}

size_t orc_bounded_digit_count(const uint64_t *moduli, size_t limbs, unsigned base_log) {
    big q;
    big_product(&q, moduli, limbs);
    unsigned bits = big_bits(&q);
    return (bits + 1 + base_log - 1) / base_log;
}

int orc_bounded_decompose(const uint64_t *in, const uint64_t *moduli, size_t limbs, size_t n,
                          unsigned base_log, int64_t *digits) {
    const size_t count = orc_bounded_digit_count(moduli, limbs, base_log);
    big q, half, x;
    big_product(&q, moduli, limbs);
    half = q;
    big_shr1(&half);
    uint64_t *invt = mixed_radix_inverses(moduli, limbs);
    uint64_t *res = malloc(limbs * 8), *dig = malloc(limbs * 8);
    const uint64_t mask = (base_log == 64) ? ~0ull : ((1ull << base_log) - 1);
    const uint64_t half_base = 1ull << (base_log - 1);
    int rc = 0;
    for (size_t c = 0; c < n && !rc; ++c) {
        for (size_t i = 0; i < limbs; ++i) res[i] = in[i * n + c];
        mixed_radix_digits_inv(res, moduli, limbs, invt, dig);
        big_from_mixed_radix(&x, dig, moduli, limbs);
        /* centered_representative_big: value > floor(Q/2) -> value - Q */
        if (big_cmp(&x, &half) > 0) big_sub(&x, &q);
        for (size_t j = 0; j < count; ++j) {
            /* residue = value mod base (Euclidean): the low bits in two's complement */
            uint64_t r = x.w[0] & mask;
            int64_t d = r >= half_base ? (int64_t)r - (int64_t)(mask) - 1 : (int64_t)r;
            digits[j * n + c] = d;
            big_add_signed_small(&x, -d); /* value - digit, divisible by base */
            big_sar(&x, base_log);
        }
        if (!big_is_zero(&x)) rc = -1; /* "bounded gadget digit count is insufficient" */
    }
    free(invt); free(res); free(dig);
    return rc;
}

void orc_lift_signed(const int64_t *digits, size_t n, uint64_t q, uint64_t *out) {
    for (size_t c = 0; c < n; ++c) {
        int64_t v = digits[c];
        if (v >= 0) out[c] = (uint64_t)v % q;
        else {
            uint64_t m = ((uint64_t)(-(v + 1)) + 1) % q;
            out[c] = m == 0 ? 0 : q - m;
        }
    }
}

/* K6 / C3, following bounded_key_switch.rs:139-219 and
 * hybrid_rns_galois_key.rs:627-780 term by term. */
int orc_bounded_key_switch(const orc_plan *const *plans, const uint64_t *qp_moduli,
                           size_t q_limbs, size_t p_limbs, unsigned base_log,
                           const uint64_t *b, const uint64_t *a,
                           const uint64_t *key_b, const uint64_t *key_a,
                           uint64_t *b_out, uint64_t *a_out) {
    const size_t n = plans[0]->n, qp = q_limbs + p_limbs;
    const size_t count = orc_bounded_digit_count(qp_moduli, q_limbs, base_log);
    int64_t *digits = malloc(count * n * 8);
    int rc = orc_bounded_decompose(a, qp_moduli, q_limbs, n, base_log, digits);
    if (rc) { free(digits); return rc; }
    uint64_t *acc_b = malloc(qp * n * 8), *acc_a = calloc(qp * n, 8);
    uint64_t *t = malloc(n * 8), *dn = malloc(n * 8);
    for (size_t i = 0; i < qp; ++i) {
        const uint64_t qi = qp_moduli[i];
        /* base term: b (K6) or P*b on Q limbs and 0 on P limbs (C3, :627-663) */
        if (i >= q_limbs) memset(t, 0, n * 8);
        else if (p_limbs == 0) memcpy(t, b + i * n, n * 8);
        else {
            uint64_t pf = 1;
            for (size_t k = 0; k < p_limbs; ++k) pf = orc_mul(pf, qp_moduli[q_limbs + k] % qi, qi);
            for (size_t c = 0; c < n; ++c) t[c] = orc_mul(b[i * n + c], pf, qi);
        }
        orc_ntt_forward(plans[i], t, acc_b + i * n);
        for (size_t j = 0; j < count; ++j) {
            orc_lift_signed(digits + j * n, n, qi, t);
            orc_ntt_forward(plans[i], t, dn);
            orc_pointwise_mac(acc_b + i * n, dn, key_b + (j * qp + i) * n, n, qi);
            orc_pointwise_mac(acc_a + i * n, dn, key_a + (j * qp + i) * n, n, qi);
        }
        orc_ntt_inverse(plans[i], acc_b + i * n, acc_b + i * n);
        orc_ntt_inverse(plans[i], acc_a + i * n, acc_a + i * n);
    }
    if (p_limbs == 0) {
        memcpy(b_out, acc_b, q_limbs * n * 8);
        memcpy(a_out, acc_a, q_limbs * n * 8);
    } else {
        orc_mod_down(acc_b, qp_moduli, q_limbs, qp_moduli + q_limbs, p_limbs, n, b_out);
        orc_mod_down(acc_a, qp_moduli, q_limbs, qp_moduli + q_limbs, p_limbs, n, a_out);
    }
    free(digits); free(acc_b); free(acc_a); free(t); free(dn);
    return 0;
}

/* ======================================================================
 * K2  grafting/rns_key_switch.rs:1724-1777  rns_tensor_with_prepared_ntt
 * ==================================================================== */
void orc_tensor(const orc_plan *const *plans, size_t limbs, const uint64_t *b0,
                const uint64_t *a0, const uint64_t *b1, const uint64_t *a1, uint64_t *c0,
                uint64_t *c1, uint64_t *c2) {
    size_t n = plans[0]->n;
    uint64_t *t0 = malloc(n * 8), *t1 = malloc(n * 8), *t2 = malloc(n * 8), *t3 = malloc(n * 8);
    uint64_t *u = malloc(n * 8);
    for (size_t i = 0; i < limbs; ++i) {
        const orc_plan *p = plans[i];
        uint64_t q = p->q;
        orc_ntt_forward(p, b0 + i * n, t0);
        orc_ntt_forward(p, a0 + i * n, t1);
        orc_ntt_forward(p, b1 + i * n, t2);
        orc_ntt_forward(p, a1 + i * n, t3);
        orc_pointwise_mul(t0, t2, u, n, q);
        orc_ntt_inverse(p, u, c0 + i * n);
        orc_pointwise_mul(t0, t3, u, n, q);
        orc_pointwise_mac(u, t1, t2, n, q);
        orc_ntt_inverse(p, u, c1 + i * n);
        orc_pointwise_mul(t1, t3, u, n, q);
        orc_ntt_inverse(p, u, c2 + i * n);
    }
    free(t0); free(t1); free(t2); free(t3); free(u);
}

/* ======================================================================
 * K3  grafting/rns_key_switch.rs:2526-2608
 *     rns_key_switch_with_prepared_dynamic_ntt
 * ==================================================================== */
int orc_rns_key_switch(const orc_plan *const *plans, const uint64_t *moduli, size_t limbs,
                       const size_t *block_sizes, size_t n_blocks, const uint64_t *b,
                       const uint64_t *a, const uint64_t *key_b, const uint64_t *key_a,
                       uint64_t *b_out, uint64_t *a_out) {
    size_t n = plans[0]->n;
    uint64_t *digits = malloc(n_blocks * limbs * n * 8);
    int rc = orc_rns_gadget_decompose(a, moduli, limbs, block_sizes, n_blocks, n, digits);
    if (rc) { free(digits); return rc; }
    uint64_t *bn = malloc(n * 8), *an = malloc(n * 8), *dn = malloc(n * 8);
    for (size_t i = 0; i < limbs; ++i) {
        const orc_plan *p = plans[i];
        uint64_t q = p->q;
        orc_ntt_forward(p, b + i * n, bn);
        memset(an, 0, n * 8);
        for (size_t j = 0; j < n_blocks; ++j) {
            orc_ntt_forward(p, digits + (j * limbs + i) * n, dn);
            orc_pointwise_mac(bn, dn, key_b + (j * limbs + i) * n, n, q);
            orc_pointwise_mac(an, dn, key_a + (j * limbs + i) * n, n, q);
        }
        orc_ntt_inverse(p, bn, b_out + i * n);
        orc_ntt_inverse(p, an, a_out + i * n);
    }
    free(bn); free(an); free(dn); free(digits);
    return 0;
}

/* ======================================================================
 * K4  dnum-block hybrid key switch with special primes P.
 * Composed from reference pieces:
 *   digit_j = canonical x mod Q_j          (rns_gadget.rs:143-228, G1)
 *   exact lift of digit_j to Q u P         (basis_transition.rs:72-138, R11)
 *   NTT-domain MAC with key[j] over Q u P  (rns_key_switch.rs:2580-2598, K3)
 *   centered ModDown by P                  (auxiliary_modulus.rs:89-171, R12)
 *   b_out = b + ModDown(acc_b): the same residues as ModDown(P*b + acc_b)
 *   used by hybrid_rns_galois_key.rs:633-663, since (P*b) mod P = 0.
 * ==================================================================== */
int orc_hybrid_key_switch(const orc_plan *const *plans, const uint64_t *qp_moduli,
                          size_t q_limbs, size_t p_limbs, const size_t *block_sizes,
                          size_t n_blocks, const uint64_t *b, const uint64_t *a,
                          const uint64_t *key_b, const uint64_t *key_a, uint64_t *b_out,
                          uint64_t *a_out) {
    size_t n = plans[0]->n, qp = q_limbs + p_limbs;
    uint64_t *acc_b = calloc(qp * n, 8), *acc_a = calloc(qp * n, 8);
    uint64_t *lift = malloc(qp * n * 8), *dn = malloc(n * 8);
    uint64_t *targets = malloc(qp * 8);
    size_t start = 0;
    for (size_t j = 0; j < n_blocks; ++j) {
        size_t end = start + block_sizes[j];
        if (end > q_limbs) { free(acc_b); free(acc_a); free(lift); free(dn); free(targets); return -1; }
        size_t nt = 0;
        for (size_t i = 0; i < qp; ++i)
            if (i < start || i >= end) targets[nt++] = qp_moduli[i];
        uint64_t *ext = malloc((nt ? nt : 1) * n * 8);
        orc_extend_basis(a + start * n, qp_moduli + start, end - start, targets, nt, n, ext);
        size_t t = 0;
        for (size_t i = 0; i < qp; ++i) {
            const uint64_t *src = (i >= start && i < end) ? a + i * n : ext + (t++) * n;
            memcpy(lift + i * n, src, n * 8);
        }
        free(ext);
        for (size_t i = 0; i < qp; ++i) {
            orc_ntt_forward(plans[i], lift + i * n, dn);
            orc_pointwise_mac(acc_b + i * n, dn, key_b + (j * qp + i) * n, n, qp_moduli[i]);
            orc_pointwise_mac(acc_a + i * n, dn, key_a + (j * qp + i) * n, n, qp_moduli[i]);
        }
        start = end;
    }
    if (start != q_limbs) { free(acc_b); free(acc_a); free(lift); free(dn); free(targets); return -1; }
    orc_rns_ntt_inverse(plans, qp, acc_b, acc_b);
    orc_rns_ntt_inverse(plans, qp, acc_a, acc_a);
    orc_mod_down(acc_b, qp_moduli, q_limbs, qp_moduli + q_limbs, p_limbs, n, b_out);
    orc_mod_down(acc_a, qp_moduli, q_limbs, qp_moduli + q_limbs, p_limbs, n, a_out);
    for (size_t i = 0; i < q_limbs; ++i)
        orc_poly_add(b_out + i * n, b + i * n, b_out + i * n, n, qp_moduli[i]);
    free(acc_b); free(acc_a); free(lift); free(dn); free(targets);
    return 0;
}

/* ======================================================================
 * E1  grafting/rns_key_switch.rs:245-383 (encrypt), :906-947 (decrypt)
 * ==================================================================== */
static void project_signed(const int64_t *src, uint64_t *dst, size_t n, uint64_t q) {
    /* project_error / signed_to_mod (rlwe/error.rs:150-166) */
    for (size_t i = 0; i < n; ++i) {
        int64_t v = src[i];
        if (v >= 0) dst[i] = (uint64_t)v % q;
        else {
            uint64_t m = (uint64_t)(-(v + 1)) + 1;
            m %= q;
            dst[i] = m == 0 ? 0 : q - m;
        }
    }
}

void orc_encrypt(const orc_plan *const *plans, const uint64_t *moduli, size_t limbs,
                 const uint64_t *message, const int64_t *error, const int8_t *secret,
                 const uint64_t *a, uint64_t *b_out) {
    size_t n = plans[0]->n;
    int64_t *s64 = malloc(n * 8);
    for (size_t i = 0; i < n; ++i) s64[i] = secret[i];
    uint64_t *s = malloc(n * 8), *e = malloc(n * 8), *an = malloc(n * 8), *sn = malloc(n * 8);
    for (size_t i = 0; i < limbs; ++i) {
        uint64_t q = moduli[i];
        project_signed(s64, s, n, q);
        project_signed(error, e, n, q);
        orc_ntt_forward(plans[i], a + i * n, an);
        orc_ntt_forward(plans[i], s, sn);
        orc_pointwise_mul(an, sn, an, n, q);
        orc_ntt_inverse(plans[i], an, an); /* a*s */
        for (size_t c = 0; c < n; ++c)
            b_out[i * n + c] = orc_sub(orc_add(message[i * n + c], e[c], q), an[c], q);
    }
    free(s64); free(s); free(e); free(an); free(sn);
}

void orc_decrypt(const orc_plan *const *plans, const uint64_t *moduli, size_t limbs,
                 const uint64_t *b, const uint64_t *a, const int8_t *secret, uint64_t *out) {
    size_t n = plans[0]->n;
    int64_t *s64 = malloc(n * 8);
    for (size_t i = 0; i < n; ++i) s64[i] = secret[i];
    uint64_t *s = malloc(n * 8), *an = malloc(n * 8), *sn = malloc(n * 8);
    for (size_t i = 0; i < limbs; ++i) {
        uint64_t q = moduli[i];
        project_signed(s64, s, n, q);
        orc_ntt_forward(plans[i], a + i * n, an);
        orc_ntt_forward(plans[i], s, sn);
        orc_pointwise_mul(an, sn, an, n, q);
        orc_ntt_inverse(plans[i], an, an);
        orc_poly_add(b + i * n, an, out + i * n, n, q);
    }
    free(s64); free(s); free(an); free(sn);
}

/* ======================================================================
 * Host samplers.  ChaCha20 block function (RFC 8439 quarter round, 20
 * rounds, 64-bit block counter in words 12-13, zero stream id in 14-15, as
 * rand_chacha 0.3.1 uses it), seeded by rand_core 0.6.4's seed_from_u64
 * (PCG32 fill of the 32-byte key), consumed through a 4-block buffer the way
 * rand_core's BlockRng does.  The samplers follow rlwe/uniform.rs:12-30,
 * rlwe/error.rs:33-140 and rlwe/secret_key.rs:24-38.
 * ==================================================================== */
#define ROTL32(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d)                                                 \
    a += b; d ^= a; d = ROTL32(d, 16); c += d; b ^= c; b = ROTL32(b, 12); \
    a += b; d ^= a; d = ROTL32(d, 8);  c += d; b ^= c; b = ROTL32(b, 7);

void orc_chacha20_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                      key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                      (uint32_t)counter, (uint32_t)(counter >> 32), 0, 0};
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    for (int i = 0; i < 10; ++i) {
        QR(x[0], x[4], x[8], x[12]); QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]); QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; ++i) out[i] = x[i] + s[i];
}

static void rng_refill(orc_rng *rng) {
    for (int blk = 0; blk < 4; ++blk) orc_chacha20_block(rng->key, rng->counter++, rng->buf + 16 * blk);
    rng->idx = 0;
}

/* PCG32's XSH-RR output permutation of a 64-bit LCG state (O'Neill, "PCG:
 * A Family of Simple Fast Space-Efficient Statistically Good Algorithms",
 * pcg32_random_r). */
uint32_t orc_pcg32_xsh_rr(uint64_t state) {
    uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
    uint32_t rot = (uint32_t)(state >> 59);
    return (xs >> rot) | (xs << ((32 - rot) & 31));
}

/* The LCG step state * 6364136223846793005 + inc (PCG32's multiplier). */
uint64_t orc_pcg32_step(uint64_t state, uint64_t inc) { return state * 6364136223846793005ull + inc; }

/* rand_core 0.6.4 SeedableRng::seed_from_u64: PCG32 with the increment
 * 11634580027462260723, output taken from the advanced state, filling the
 * 32-byte ChaCha key word by word (little endian). */
void orc_rng_seed(orc_rng *rng, uint64_t state) {
    const uint64_t inc = 11634580027462260723ull;
    for (int i = 0; i < 8; ++i) {
        state = orc_pcg32_step(state, inc);
        rng->key[i] = orc_pcg32_xsh_rr(state);
    }
    rng->counter = 0;
    rng->idx = 64; /* empty: refill on first use */
}

/* A stream with the given key at the given block counter (test access to
 * the ChaCha stream layout). */
void orc_rng_set_key(orc_rng *rng, const uint32_t key[8], uint64_t counter) {
    memcpy(rng->key, key, sizeof rng->key);
    rng->counter = counter;
    rng->idx = 64;
}

uint32_t orc_rng_next_u32(orc_rng *rng) {
    if (rng->idx >= 64) rng_refill(rng);
    return rng->buf[rng->idx++];
}

uint64_t orc_rng_next_u64(orc_rng *rng) {
    if (rng->idx < 63) {
        uint64_t lo = rng->buf[rng->idx], hi = rng->buf[rng->idx + 1];
        rng->idx += 2;
        return (hi << 32) | lo;
    }
    if (rng->idx >= 64) {
        rng_refill(rng);
        rng->idx = 2;
        return ((uint64_t)rng->buf[1] << 32) | rng->buf[0];
    }
    uint64_t lo = rng->buf[63];
    rng_refill(rng);
    rng->idx = 1;
    return ((uint64_t)rng->buf[0] << 32) | lo;
}

/* sample_uniform_residues (rlwe/uniform.rs:12-30): rejection below
 * threshold = (2^64 - q) mod q, then candidate mod q. */
void orc_sample_uniform(orc_rng *rng, uint64_t q, size_t n, uint64_t *out) {
    uint64_t threshold = (0 - q) % q;
    size_t c = 0;
    while (c < n) {
        uint64_t cand = orc_rng_next_u64(rng);
        if (cand >= threshold) out[c++] = cand % q;
    }
}

/* DiscreteGaussianTable::new (rlwe/error.rs:73-88): weights
 * exp(-x^2 / (2 sigma^2)) for x in [-tail, tail], tail = ceil(16 sigma),
 * summed left to right.  Returns the entry count (cap: room in weights). */
size_t orc_gaussian_table(double sigma, double *weights, size_t cap, double *total) {
    int64_t tail = (int64_t)ceil(16.0 * sigma);
    size_t cnt = (size_t)(2 * tail + 1);
    double t = 0.0;
    for (int64_t x = -tail; x <= tail; ++x) {
        double xf = (double)x;
        double w = exp(-(xf * xf) / (2.0 * sigma * sigma));
        if ((size_t)(x + tail) < cap) weights[x + tail] = w;
        t += w;
    }
    *total = t;
    return cnt;
}

/* uniform_open_unit (rlwe/error.rs:131-140): the top 53 bits of next_u64,
 * centred in their interval, in (0, 1). */
double orc_uniform_open_unit(orc_rng *rng) {
    uint64_t v = orc_rng_next_u64(rng) >> 11;
    return ((double)v + 0.5) * (1.0 / (double)((uint64_t)1 << 53));
}

/* Census of rand 0.8's single-sample u8 range 0..3 over every u32 draw
 * (widening multiply by 3, accept while the low half is within the zone):
 * counts[0..2] the accepted draws giving 0, 1, 2; counts[3] the rejected. */
void orc_ternary_census(uint64_t counts[4]) {
    const uint32_t range = 3;
    const uint32_t zone = 0xffffffffu - ((0xffffffffu - range + 1) % range);
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    uint32_t v = 0;
    do {
        uint64_t m = (uint64_t)v * range;
        if ((uint32_t)m <= zone) ++counts[m >> 32];
        else ++counts[3];
    } while (++v != 0);
}

/* DiscreteGaussianTable::sample (rlwe/error.rs:90-108). */
void orc_sample_gaussian(orc_rng *rng, double sigma, size_t n, int64_t *out) {
    int64_t tail = (int64_t)ceil(16.0 * sigma);
    size_t cnt = (size_t)(2 * tail + 1);
    double *wt = malloc(cnt * sizeof(double)), total = 0.0;
    orc_gaussian_table(sigma, wt, cnt, &total);
    for (size_t i = 0; i < n; ++i) {
        double target = orc_uniform_open_unit(rng) * total;
        double cum = 0.0;
        int64_t res = tail;
        for (size_t k = 0; k < cnt; ++k) {
            cum += wt[k];
            if (target < cum) { res = (int64_t)k - tail; break; }
        }
        out[i] = res;
    }
    free(wt);
}

/* SecretKey::generate_with_rng (rlwe/secret_key.rs:24-38) with rand 0.8's
 * single-sample u8 range 0..3 (widening multiply over u32 with a zone). */
void orc_sample_ternary(orc_rng *rng, size_t n, int8_t *out) {
    for (;;) {
        int any = 0;
        const uint32_t range = 3;
        const uint32_t zone = 0xffffffffu - ((0xffffffffu - range + 1) % range);
        for (size_t i = 0; i < n; ++i) {
            uint32_t hi;
            for (;;) {
                uint32_t v = orc_rng_next_u32(rng);
                uint64_t m = (uint64_t)v * range;
                hi = (uint32_t)(m >> 32);
                if ((uint32_t)m <= zone) break;
            }
            out[i] = hi == 0 ? 0 : (hi == 1 ? 1 : -1);
            any |= out[i] != 0;
        }
        if (any) return;
    }
}
