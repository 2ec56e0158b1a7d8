// Host-sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer) of the
// CPU code: the oracle (oracle/ckks_oracle.c), the serving scheduler and the
// CPU build of the NTT tiles, driven through representative calls at small
// sizes.  Test infrastructure only; exits non-zero on a failed check (the
// sanitizers abort on their own findings).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ntt_emu.cpp"
#include "serve_emu.cpp"

extern "C" {
#include "../../oracle/ckks_oracle.h"
}

static int failures = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

static std::vector<uint64_t> rand_limbs(orc_rng *rng, const std::vector<uint64_t> &mods, size_t n) {
    std::vector<uint64_t> out(mods.size() * n);
    for (size_t i = 0; i < mods.size(); ++i) orc_sample_uniform(rng, mods[i], n, out.data() + i * n);
    return out;
}

int main() {
    // serving scheduler scenarios and a mixed run
    long out[16];
    for (int w = 0; w < 4; ++w) CHECK(serve_emu_scenario(w, out) == 0);
    int kinds[6] = {0, 1, 0, 0, 1, 0}, arrive[6] = {0, 0, 1, 2, 3, 3};
    uint64_t order[6];
    CHECK(serve_emu_run(6, kinds, arrive, 3, 2, 1, 0, out, order) == 0 && out[0] == 6);
    char msg[256];
    for (int w = 0; w <= 6; ++w) serve_emu_validate(w, msg, sizeof msg);

    // NTT tiles (CPU build, bounds-checked) against the oracle's prepared NTT
    const size_t sizes[3] = {10, 13, 16};
    uint64_t q40 = 0, q60 = 0;
    size_t found = 0;
    uint64_t primes[2];
    for (uint64_t c = (1ull << 40) + 1; found < 1 && c > (1ull << 39); c -= 2ull << 16)
        if (orc_is_prime(c)) primes[found++] = c;
    q40 = primes[0];
    found = 0;
    for (uint64_t c = (1ull << 60) + 1; found < 1 && c > (1ull << 59); c -= 2ull << 16)
        if (orc_is_prime(c)) primes[found++] = c;
    q60 = primes[0];
    orc_rng rng;
    orc_rng_seed(&rng, 1);
    for (size_t s : sizes) {
        for (uint64_t q : {q40, q60}) {
            const size_t n = (size_t)1 << s;
            std::vector<uint64_t> x(n), a(n), b(n), c(n);
            orc_sample_uniform(&rng, q, n, x.data());
            orc_plan *p = orc_plan_new(q, n, 0);
            orc_ntt_forward(p, x.data(), a.data());
            CHECK(emu_ntt(x.data(), b.data(), q, (uint32_t)s, 1) == 0);
            CHECK(std::memcmp(a.data(), b.data(), n * 8) == 0);
            CHECK(emu_ntt(b.data(), c.data(), q, (uint32_t)s, 0) == 0);
            CHECK(std::memcmp(c.data(), x.data(), n * 8) == 0);
            orc_plan_free(p);
        }
    }
    int flagged = 0;
    emu_twiddle_uniform(&flagged);

    // oracle: hybrid and RNS-gadget key switches, bounded key switch, tensor,
    // rescale, mod-down, encryption round trip, samplers, monomials
    const size_t n = 1024;
    std::vector<uint64_t> qm, pm;
    for (uint64_t cand = (1ull << 40) + 1; qm.size() < 4; cand -= 2 * n)
        if (orc_is_prime(cand)) qm.push_back(cand);
    for (uint64_t cand = (1ull << 60) + 1; pm.size() < 2; cand -= 2 * n)
        if (orc_is_prime(cand)) pm.push_back(cand);
    std::vector<uint64_t> qp(qm);
    qp.insert(qp.end(), pm.begin(), pm.end());
    std::vector<orc_plan *> plans;
    for (uint64_t q : qp) plans.push_back(orc_plan_new(q, n, 0));
    const size_t L = qm.size(), K = pm.size(), B = L + K;
    std::vector<uint64_t> b = rand_limbs(&rng, qm, n), a = rand_limbs(&rng, qm, n);
    const size_t blocks[2] = {2, 2};
    std::vector<uint64_t> kb = rand_limbs(&rng, qp, n), ka = rand_limbs(&rng, qp, n);
    kb.resize(2 * B * n);
    ka.resize(2 * B * n);
    for (size_t i = B * n; i < 2 * B * n; ++i) kb[i] = ka[i] = kb[i - B * n] % qp[(i / n) % B];
    std::vector<uint64_t> bo(L * n), ao(L * n);
    CHECK(orc_hybrid_key_switch(plans.data(), qp.data(), L, K, blocks, 2, b.data(), a.data(), kb.data(),
                                ka.data(), bo.data(), ao.data()) == 0);
    std::vector<uint64_t> rkb(L * L * n), rka(L * L * n);
    for (size_t i = 0; i < L * L * n; ++i) rkb[i] = rka[i] = (i * 2654435761u) % qm[(i / n) % L];
    const size_t singles[4] = {1, 1, 1, 1};
    CHECK(orc_rns_key_switch(plans.data(), qm.data(), L, singles, 4, b.data(), a.data(), rkb.data(), rka.data(),
                             bo.data(), ao.data()) == 0);
    const unsigned w = 32;
    const size_t cnt = orc_bounded_digit_count(qm.data(), L, w);
    std::vector<uint64_t> bkb(cnt * B * n), bka(cnt * B * n);
    for (size_t i = 0; i < bkb.size(); ++i) bkb[i] = bka[i] = (i * 40503u) % qp[(i / n) % B];
    CHECK(orc_bounded_key_switch(plans.data(), qp.data(), L, K, w, b.data(), a.data(), bkb.data(), bka.data(),
                                 bo.data(), ao.data()) == 0);
    std::vector<int64_t> digits(cnt * n);
    CHECK(orc_bounded_decompose(a.data(), qm.data(), L, n, w, digits.data()) == 0);
    std::vector<uint64_t> c0(L * n), c1(L * n), c2(L * n);
    orc_tensor(plans.data(), L, b.data(), a.data(), b.data(), a.data(), c0.data(), c1.data(), c2.data());
    std::vector<uint64_t> rb((L - 1) * n), ra((L - 1) * n);
    orc_rescale_ct(qm.data(), L, n, b.data(), a.data(), rb.data(), ra.data());
    std::vector<uint64_t> x = rand_limbs(&rng, qp, n), md(L * n);
    orc_mod_down(x.data(), qm.data(), L, pm.data(), K, n, md.data());
    std::vector<int8_t> s(n);
    orc_sample_ternary(&rng, n, s.data());
    std::vector<int64_t> e(n);
    orc_sample_gaussian(&rng, 3.2, n, e.data());
    std::vector<uint64_t> m = rand_limbs(&rng, qm, n), ct_b(L * n), dec(L * n);
    orc_encrypt(plans.data(), qm.data(), L, m.data(), e.data(), s.data(), a.data(), ct_b.data());
    orc_decrypt(plans.data(), qm.data(), L, ct_b.data(), a.data(), s.data(), dec.data());
    for (size_t i = 0; i < L; ++i)
        for (size_t j = 0; j < n; ++j) {
            const uint64_t q = qm[i], got = dec[i * n + j], want = m[i * n + j];
            const uint64_t diff = got >= want ? got - want : q - (want - got);
            const int64_t centered = diff > q / 2 ? -(int64_t)(q - diff) : (int64_t)diff;
            CHECK(centered == e[j]);
        }
    std::vector<uint64_t> mono(n);
    orc_mul_monomial(b.data(), mono.data(), n, qm[0], 3 * n / 2, -1);
    orc_ntt_mul_monomial(plans[0], b.data(), mono.data(), -7);
    std::vector<uint64_t> ext(2 * n);
    orc_extend_basis(b.data(), qm.data(), 2, pm.data(), 2, n, ext.data());
    for (orc_plan *p : plans) orc_plan_free(p);
    uint64_t census_dummy = orc_pcg32_xsh_rr(12345);
    (void)census_dummy;
    if (failures) std::fprintf(stderr, "%d checks failed\n", failures);
    else std::printf("sanitized checks passed\n");
    return failures ? 1 : 0;
}
