"""Pins the CPU oracle (a C restatement of the reference algorithms) before
it is trusted as the oracle of the GPU path.

Every oracle function is checked against something independent of its own
code: the reference's known-answer vectors and fixtures, closed forms
evaluated with Python big integers, O(N^2) definitions, or invariants that a
dropped term, wrong sign or transposed index would break."""
from __future__ import annotations

import math
import random

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues


def py_psi(q, n):
    """closed form: psi = (smallest quadratic non-residue)^((q-1)/2N)."""
    g = 2
    while pow(g, (q - 1) // 2, q) != q - 1:
        g += 1
    return pow(g, (q - 1) // (2 * n), q)


def crt(residues, moduli):
    """Python big-integer CRT to the canonical value in [0, prod)."""
    M = math.prod(moduli)
    x = 0
    for r, q in zip(residues, moduli):
        Mi = M // q
        x += int(r) * Mi * pow(Mi, -1, q)
    return x % M


# ------------------------------------------------------------------ R1
EDGE_MODULI = [2, 3, 97, 12289, Q40[0], Q0[0], (1 << 62) - 57, (1 << 63) + 29, (1 << 64) - 59]


def xorshift(seed):
    # xorshift64 generator as defined inside ring/modulus.rs tests
    x = seed
    while True:
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        yield x


@pytest.mark.parametrize("q", EDGE_MODULI)
def test_scalar_modular_arithmetic_matches_big_integers(q):
    vals = sorted({0, 1, q // 2, max(q - 2, 0), q - 1})
    gen = xorshift(0x9E3779B97F4A7C15 ^ q)
    vals += [next(gen) % q for _ in range(64)]
    for a in vals:
        for b in vals[:12]:
            assert O.add(a, b, q) == (a + b) % q
            assert O.sub(a, b, q) == (a - b) % q
            assert O.mul(a, b, q) == (a * b) % q
            if q > 2:
                assert O.barrett_mul(a, b, q) == (a * b) % q
            if 2 <= q < (1 << 63) and b < q:
                wq = O.shoup_quotient(b, q)
                assert wq == (b << 64) // q
                assert O.shoup_mul(a, b, wq, q) == (a * b) % q
    assert O.pow_(3, q - 1, q) == (1 if q > 3 and O.is_prime(q) else pow(3, q - 1, q))


def test_modular_inverse_and_pow():
    for q in [97, 12289, Q40[3], Q0[0]]:
        for a in [1, 2, 5, q - 1, q // 3]:
            assert O.inv(a, q) * a % q == 1
            assert O.pow_(a, 12345, q) == pow(a, 12345, q)


# ------------------------------------------------------------------ R14
# RESEARCH_65536_MODULUS_POOL (ckks/parameters.rs:397-426): the first 28 NTT
# primes below 2^55 for N = 65536, in search order.
RESEARCH_65536_POOL = [
    0x7fffffffba0001, 0x7fffffffaa0001, 0x7fffffff7e0001, 0x7fffffff380001,
    0x7ffffffef00001, 0x7ffffffeba0001, 0x7ffffffeac0001, 0x7ffffffe700001,
    0x7ffffffe600001, 0x7ffffffe4c0001, 0x7ffffffe220001, 0x7ffffffe160001,
    0x7ffffffdec0001, 0x7ffffffdd00001, 0x7ffffffdce0001, 0x7ffffffdbc0001,
    0x7ffffffd5c0001, 0x7ffffffd440001, 0x7ffffffd340001, 0x7ffffffd280001,
    0x7ffffffd160001, 0x7ffffffcf60001, 0x7ffffffcea0001, 0x7ffffffcda0001,
    0x7ffffffccc0001, 0x7ffffffcc20001, 0x7ffffffc800001, 0x7ffffffc080001,
]


def test_prime_search_reproduces_reference_pool():
    assert O.ntt_primes(55, 65536, 28) == RESEARCH_65536_POOL


def test_prime_search_against_trial_division():
    # brute force on small widths: every q = 1 mod 2N below 2^bits, descending
    for bits, n in [(16, 8), (20, 64), (24, 256)]:
        expect = []
        q = (1 << bits) - 1
        q -= (q - 1) % (2 * n)
        while len(expect) < 6:
            if q > 2 and all(q % d for d in range(2, int(q ** 0.5) + 1)):
                expect.append(q)
            q -= 2 * n
        assert O.ntt_primes(bits, n, 6) == expect


def test_primality_deterministic_witnesses():
    assert O.is_prime((1 << 61) - 1)
    assert not O.is_prime(3215031751)           # strong pseudoprime to 2,3,5,7
    assert not O.is_prime(3825123056546413051)  # spsp to the first 9 primes
    for q in Q40 + Q0 + P60:
        assert (q - 1) % (1 << 17) == 0


# ------------------------------------------------------------------ R2
def test_negacyclic_root_known_answers():
    # ring/ntt_parameter_sets.rs:8-12 and results/r3.3d/canonical-ntt-parameters.txt
    assert O.find_negacyclic_root(4294828033, 4096) == 1953722822
    assert O.find_negacyclic_root(18446744073709436929, 4096) == 5975861664659593359
    # ring/ntt.rs tests use psi = 8 for N = 8, q = 97
    assert O.find_negacyclic_root(97, 8) == 8


def test_negacyclic_root_closed_form():
    for q, n in [(12289, 16), (12289, 256), (40961, 512), (65537, 1024), (786433, 2048)]:
        assert O.find_negacyclic_root(q, n) == py_psi(q, n)
    for q in Q40[:4] + Q0:
        for n in [1024, 4096, 65536]:
            psi = O.find_negacyclic_root(q, n)
            assert psi == py_psi(q, n)
            assert pow(psi, n, q) == q - 1
        # psi_{N/2} = psi_N^2 (shared twiddle-table prefixes)
        assert O.find_negacyclic_root(q, 32768) == pow(O.find_negacyclic_root(q, 65536), 2, q)


# ------------------------------------------------------------------ R3-R5
def test_ntt_small_known_answers():
    plan = O.Plan(97, 8, 8)
    # delta -> all ones (ring/ntt.rs reference_roundtrip_recovers_constant_one)
    assert plan.forward([1, 0, 0, 0, 0, 0, 0, 0]).tolist() == [1] * 8
    # reference vectors: NTT product == naive negacyclic product (ring/ntt.rs:261-305)
    pairs = [([1, 8, 23, 42, 7, 11, 19, 31], [5, 7, 11, 13, 17, 29, 37, 41]),
             ([1, 2, 3, 4, 5, 6, 7, 8], [8, 7, 6, 5, 4, 3, 2, 1]),
             ([0, 1, 0, 1, 0, 1, 0, 1], [1, 0, 1, 0, 1, 0, 1, 0]),
             ([96, 95, 94, 93, 92, 91, 90, 89], [9, 10, 11, 12, 13, 14, 15, 16])]
    for lhs, rhs in pairs:
        prod = plan.inverse(O.pointwise_mul(plan.forward(lhs), plan.forward(rhs), 97))
        naive = [0] * 8
        for i in range(8):
            for j in range(8):
                d = i + j
                s = 1 if d < 8 else -1
                naive[d % 8] = (naive[d % 8] + s * lhs[i] * rhs[j]) % 97
        assert prod.tolist() == naive
        assert O.negacyclic_mul_naive(lhs, rhs, 97).tolist() == naive


def definition_ntt(a, q, psi):
    n = len(a)
    return [sum(int(a[j]) * pow(psi, (2 * k + 1) * j, q) for j in range(n)) % q for k in range(n)]


@pytest.mark.parametrize("q,n", [(12289, 16), (12289, 64), (12289, 256), (Q40[0], 32), (Q0[0], 64),
                                 (18446744073709436929, 16)])
def test_prepared_ntt_matches_definition(q, n):
    rng = random.Random(q ^ n)
    a = [rng.randrange(q) for _ in range(n)]
    plan = O.Plan(q, n)
    fwd = plan.forward(a)
    assert fwd.tolist() == definition_ntt(a, q, plan.psi)
    assert O.ntt_forward_reference(a, q, plan.psi).tolist() == fwd.tolist()
    assert plan.inverse(fwd).tolist() == a
    assert O.ntt_inverse_reference(fwd, q, plan.psi).tolist() == a


@pytest.mark.parametrize("n", [1024, 4096])
def test_prepared_ntt_matches_o_n2_reference_at_config_primes(n):
    rng = np.random.default_rng(n)
    for q in [Q0[0], Q40[0], Q40[23], P60[6]]:
        a = rng.integers(0, q, n, dtype=np.uint64)
        plan = O.Plan(q, n)
        fwd = plan.forward(a)
        assert (fwd == O.ntt_forward_reference(a, q, plan.psi)).all()
        assert (plan.inverse(fwd) == a).all()


def test_ntt_product_equals_naive_product_at_config_primes():
    rng = np.random.default_rng(7)
    n = 256
    for q in [Q0[0], Q40[5]]:
        plan = O.Plan(q, n)
        x = rng.integers(0, q, n, dtype=np.uint64)
        y = rng.integers(0, q, n, dtype=np.uint64)
        prod = plan.inverse(O.pointwise_mul(plan.forward(x), plan.forward(y), q))
        assert (prod == O.negacyclic_mul_naive(x, y, q)).all()


def test_monomial_multiplication_in_ntt_domain():
    # ring/ntt_polynomial.rs:339-377: X^e in the NTT domain equals the signed
    # coefficient shift
    n, q = 64, 12289
    plan = O.Plan(q, n)
    rng = np.random.default_rng(3)
    a = rng.integers(0, q, n, dtype=np.uint64)
    for e in [-129, -65, -1, 0, 1, 17, 63, 64, 65, 127, 257]:
        got = plan.inverse(O.ntt_mul_monomial(plan, plan.forward(a), e))
        mono = np.zeros(n, dtype=np.uint64)
        ee = e % (2 * n)
        mono[ee % n] = 1 if ee < n else q - 1
        assert (got == O.negacyclic_mul_naive(a, mono, q)).all()


def test_coefficient_monomial_is_a_negacyclic_product():
    # ring/polynomial.rs:164-200 (and its tests): sign * X^e * a equals the
    # naive negacyclic product with the monomial; X^(2N) = 1, X^N = -1
    n, q = 64, 12289
    rng = np.random.default_rng(5)
    a = rng.integers(0, q, n, dtype=np.uint64)
    for e in [0, 1, 5, 63, 64, 65, 127, 128, 129, 1000]:
        for sign in (1, -1):
            mono = np.zeros(n, dtype=np.uint64)
            ee = e % (2 * n)
            neg = (ee >= n) ^ (sign == -1)
            mono[ee % n] = q - 1 if neg else 1
            assert (O.mul_monomial(a, q, e, sign) == O.negacyclic_mul_naive(a, mono, q)).all(), (e, sign)
    # the NTT-domain form of the same product (ring/ntt_polynomial.rs:133)
    plan = O.Plan(q, n)
    for e in [3, 64, 100]:
        assert (plan.inverse(O.ntt_mul_monomial(plan, plan.forward(a), e)) == O.mul_monomial(a, q, e)).all()


def test_scalar_mul_known_values():
    # Modulus::mul reduces the full 128-bit product: scalars above q, 0, q-1
    q = 1152921504606584833
    x = np.array([0, 1, q - 1, 123456789123456789, q // 2], dtype=np.uint64)
    for scalar in [0, 1, q - 1, q, q + 5, (1 << 64) - 1, 0x9E3779B97F4A7C15]:
        want = [(int(v) * scalar) % q for v in x]
        assert O.scalar_mul(x, q, scalar).tolist() == want
    # linearity through the NTT: NTT(c a) = c NTT(a)
    n, q2 = 64, 12289
    plan = O.Plan(q2, n)
    a = np.random.default_rng(6).integers(0, q2, n, dtype=np.uint64)
    assert (plan.forward(O.scalar_mul(a, q2, 777)) == O.scalar_mul(plan.forward(a), q2, 777)).all()


# ------------------------------------------------------------------ R8 / C1
def test_automorphism_known_answer():
    # ckks/automorphism.rs automorphism_matches_manual_negacyclic_mapping
    a = [1, 2, 3, 4, 5, 6, 7, 8]
    q = 97
    assert O.apply_automorphism(a, q, 3).tolist() == [1, q - 4, 7, 2, q - 5, 8, 3, q - 6]
    assert O.apply_automorphism(a, q, 1).tolist() == a
    assert (O.apply_automorphism(a, q, 3) == O.apply_automorphism(a, q, 19)).all()


def test_galois_permutation_matches_coefficient_automorphism():
    # ring/ntt_galois_permutation.rs:39-66 (psi = 11^((q-1)/2N) there)
    for n in [8, 16, 32, 64]:
        q = 12289
        psi = pow(11, (q - 1) // (2 * n), q)
        plan = O.Plan(q, n, psi)
        a = [(i * i + 17 * i + 3) % q for i in range(n)]
        fa = plan.forward(a)
        for k in range(1, 2 * n, 2):
            assert (O.permute_ntt_galois(fa, k) == plan.forward(O.apply_automorphism(a, q, k))).all()


def test_rotation_and_conjugation_exponents():
    for n in [8, 4096, 65536]:
        for r in [0, 1, 2, 5, n // 2 - 1, n // 2, 3 * n]:
            assert O.rotation_exponent_left(n, r) == pow(5, r % (n // 2), 2 * n)
            left = O.rotation_exponent_left(n, r)
            right = O.rotation_exponent_right(n, r)
            assert left * right % (2 * n) == 1
        assert O.conjugation_exponent(n) == 2 * n - 1
        for k in [1, 3, 5, 2 * n - 1]:
            assert k * O.inverse_automorphism_exponent(n, k) % (2 * n) == 1


# ------------------------------------------------------------------ C5
def test_rescale_matches_integer_closed_form():
    # ckks/rns_rescale.rs:178-215: round(centered(x) / q_last) mod q_i
    rng = np.random.default_rng(11)
    moduli = [Q0[0]] + Q40[:3]
    n = 64
    L = len(moduli)
    Q = math.prod(moduli)
    coeffs = [int(v) for v in rng.integers(0, 1 << 62, n, dtype=np.uint64)]
    coeffs = [(c * 0x9E3779B97F4A7C15 + c * c) % Q for c in coeffs]
    coeffs[:6] = [0, 1, Q - 1, Q // 2, Q // 2 + 1, moduli[-1] // 2]
    b = np.array([[c % q for c in coeffs] for q in moduli], dtype=np.uint64)
    a = np.array([[(c * 3) % Q % q for c in coeffs] for q in moduli], dtype=np.uint64)
    bo, ao = O.rescale_ct(moduli, b, a)
    qd = moduli[-1]
    for comp, src in ((bo, coeffs), (ao, [(c * 3) % Q for c in coeffs])):
        for j, x in enumerate(src):
            rem = x % qd
            cr = rem - qd if rem > qd // 2 else rem
            r = (x - cr) // qd
            for i in range(L - 1):
                assert int(comp[i, j]) == r % moduli[i]


# ------------------------------------------------------------------ R11 / R12
def test_extend_basis_matches_crt():
    rng = np.random.default_rng(5)
    src = [Q0[0]] + Q40[:4]
    tgt = Q40[4:9] + P60[:3]
    n = 32
    x = rand_residues(rng, src, n)[0]
    out = O.extend_basis(x, src, tgt)
    for c in range(n):
        v = crt(x[:, c], src)
        for t, q in enumerate(tgt):
            assert int(out[t, c]) == v % q


def round_div_away(x, d):
    # ring/auxiliary_modulus.rs:14-25, ties away from zero
    r = (abs(x) + d // 2) // d
    return -r if x < 0 else r


def test_mod_down_signed_boundaries_reference_bases():
    # ring/auxiliary_modulus.rs:210-242 with Q = {97, 193}, P = {257}
    qm, pm = [97, 193], [257]
    QP, Qv, Pv = 97 * 193 * 257, 97 * 193, 257
    vals = [-1000, -386, -257, -129, -128, -1, 0, 1, 128, 129, 257, 386, 1000]
    x = np.array([[v % m for v in vals] for m in qm + pm], dtype=np.uint64)
    out = O.mod_down(x, qm, pm)
    for j, v in enumerate(vals):
        expect = round_div_away(v, Pv) % Qv
        assert crt(out[:, j], qm) == expect
    # signed rounding table of the same file
    for v, e in [(0, 0), (4, 0), (5, 1), (6, 1), (14, 1), (15, 2), (-4, 0), (-5, -1), (-6, -1),
                 (-14, -1), (-15, -2)]:
        assert round_div_away(v, 10) == e


def test_mod_down_matches_big_integer_rounding_at_config_sizes():
    rng = np.random.default_rng(9)
    qm = [Q0[0]] + Q40[:8]
    pm = P60[:4]
    n = 48
    x = rand_residues(rng, qm + pm, n)[0]
    # plant values close to 0, +-P/2 and QP
    QP = math.prod(qm + pm)
    Pv = math.prod(pm)
    planted = [0, 1, QP - 1, Pv // 2, Pv // 2 + 1, QP - Pv // 2, QP - Pv // 2 - 1, Pv, QP - Pv]
    for j, v in enumerate(planted):
        for i, m in enumerate(qm + pm):
            x[i, j] = v % m
    out = O.mod_down(x, qm, pm)
    Qv = math.prod(qm)
    for c in range(n):
        v = crt(x[:, c], qm + pm)
        cen = v - QP if v > QP // 2 else v
        assert crt(out[:, c], qm) == round_div_away(cen, Pv) % Qv


# ------------------------------------------------------------------ G1
def test_rns_gadget_digits_reconstruct():
    rng = np.random.default_rng(13)
    moduli = [Q0[0]] + Q40[:5]
    n = 16
    x = rand_residues(rng, moduli, n)[0]
    Q = math.prod(moduli)
    # a block whose composite modulus exceeds 128 bits is rejected, as the
    # reference panics "gadget block composite modulus exceeds u128"
    with pytest.raises(ValueError):
        O.rns_gadget_decompose(x, moduli, [3, 3])
    for blocks in ([1] * 6, [2, 2, 2], [2, 1, 3], [1, 2, 3]):
        d = O.rns_gadget_decompose(x, moduli, blocks)
        start = 0
        rec = [0] * n
        for j, sz in enumerate(blocks):
            Qj = math.prod(moduli[start:start + sz])
            E = (Q // Qj) * pow(Q // Qj, -1, Qj) % Q
            for c in range(n):
                v = crt(x[:, c], moduli)
                digit = v % Qj  # canonical digit in [0, Q_j)
                for i, q in enumerate(moduli):
                    assert int(d[j, i, c]) == digit % q
                rec[c] = (rec[c] + digit * E) % Q
            start += sz
        for c in range(n):
            assert rec[c] == crt(x[:, c], moduli)


# ------------------------------------------------------------------ K2/K3/K4/E1
def _keygen_rns(plan, blocks, s_src, s_dst, rng, sigma0=True):
    """RNS-gadget key: entry j encrypts E_j * s_src under s_dst, NTT domain."""
    moduli = plan.moduli
    n = plan.n
    L = len(moduli)
    kb = np.zeros((len(blocks), L, n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    start = 0
    for j, sz in enumerate(blocks):
        a = rand_residues(rng, moduli, n)[0]
        # E_j is 1 on the block's limbs and 0 elsewhere (CRT idempotent)
        msg = np.zeros((L, n), dtype=np.uint64)
        for i in range(start, start + sz):
            msg[i] = np.array([v % moduli[i] for v in s_src.tolist()], dtype=np.uint64)
        err = np.zeros(n, dtype=np.int64)
        b = O.encrypt(plan, msg, err, s_dst, a)
        kb[j] = plan.forward(b)
        ka[j] = plan.forward(a)
        start += sz
    return kb, ka


def _centered(v, q):
    v = int(v)
    return v - q if v > q // 2 else v


def test_encrypt_decrypt_roundtrip_is_message_plus_error():
    rng = np.random.default_rng(21)
    moduli = [Q0[0]] + Q40[:3]
    n = 1024
    plan = O.RnsPlan(moduli, n)
    s = rng.integers(-1, 2, n).astype(np.int8)
    m = rand_residues(rng, moduli, n)[0]
    e = rng.integers(-20, 21, n).astype(np.int64)
    a = rand_residues(rng, moduli, n)[0]
    b = O.encrypt(plan, m, e, s, a)
    dec = O.decrypt(plan, b, a, s)
    for i, q in enumerate(moduli):
        assert (dec[i] == (m[i].astype(object) + e) % q).all()


def test_rns_key_switch_preserves_decryption_exactly_with_zero_noise():
    # grafting/rns_key_switch.rs:3735-3895
    rng = np.random.default_rng(17)
    moduli = [Q0[0]] + Q40[:3]
    n = 1024
    plan = O.RnsPlan(moduli, n)
    s_src = rng.integers(-1, 2, n).astype(np.int8)
    s_dst = rng.integers(-1, 2, n).astype(np.int8)
    for blocks in ([1, 1, 1, 1], [2, 2]):
        kb, ka = _keygen_rns(plan, blocks, s_src, s_dst, rng)
        b = rand_residues(rng, moduli, n)[0]
        a = rand_residues(rng, moduli, n)[0]
        bo, ao = O.rns_key_switch(plan, blocks, b, a, kb, ka)
        assert (O.decrypt(plan, bo, ao, s_dst) == O.decrypt(plan, b, a, s_src)).all()


def test_rns_key_switch_is_linear_in_the_key():
    # digits are canonical residues, so the switch is linear in the key
    # entries (not in the ciphertext: digit lifts wrap modulo q_j)
    rng = np.random.default_rng(19)
    moduli = Q40[:3]
    n = 1024
    plan = O.RnsPlan(moduli, n)
    k1b, k1a, k2b, k2a = (rand_residues(rng, moduli, n, 3) for _ in range(4))
    add = lambda x, y: np.array([(x[..., i, :].astype(object) + y[..., i, :]) % q
                                 for i, q in enumerate(moduli)], dtype=np.uint64).swapaxes(0, -2)
    zero = np.zeros((3, n), dtype=np.uint64)
    b, a = rand_residues(rng, moduli, n)[0], rand_residues(rng, moduli, n)[0]
    o1 = O.rns_key_switch(plan, [1, 1, 1], zero, a, k1b, k1a)
    o2 = O.rns_key_switch(plan, [1, 1, 1], zero, a, k2b, k2a)
    o12 = O.rns_key_switch(plan, [1, 1, 1], zero, a, np.ascontiguousarray(add(k1b, k2b)),
                           np.ascontiguousarray(add(k1a, k2a)))
    assert (o12[0] == add(o1[0], o2[0])).all() and (o12[1] == add(o1[1], o2[1])).all()
    # a = 0 leaves b and gives a zero a-component
    ob, oa = O.rns_key_switch(plan, [1, 1, 1], b, zero, k1b, k1a)
    assert (ob == b).all() and (oa == 0).all()


def _keygen_hybrid(qp_plan, q_limbs, blocks, s_prime, s, rng, err=None):
    """dnum key over QP: entry j encrypts P * E_j * s' under s."""
    moduli = qp_plan.moduli
    n = qp_plan.n
    QP = len(moduli)
    pm = moduli[q_limbs:]
    Pv = math.prod(pm)
    kb = np.zeros((len(blocks), QP, n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    start = 0
    for j, sz in enumerate(blocks):
        a = rand_residues(rng, moduli, n)[0]
        msg = np.zeros((QP, n), dtype=np.uint64)
        for i in range(start, start + sz):
            q = moduli[i]
            msg[i] = np.array([(int(v) * Pv) % q for v in s_prime.tolist()], dtype=np.uint64)
        e = np.zeros(n, dtype=np.int64) if err is None else err(n)
        b = O.encrypt(qp_plan, msg, e, s, a)
        kb[j] = qp_plan.forward(b)
        ka[j] = qp_plan.forward(a)
        start += sz
    return kb, ka


@pytest.mark.parametrize("blocks,K", [([2, 2], 2), ([3, 2], 3), ([1, 2, 2], 2)])
def test_hybrid_key_switch_decrypts_within_mod_down_rounding(blocks, K):
    rng = np.random.default_rng(23)
    q_moduli = [Q0[0]] + Q40[: sum(blocks) - 1]
    p_moduli = P60[:K]
    n = 1024
    qp = O.RnsPlan(q_moduli + p_moduli, n)
    q_plan = qp.prefix(len(q_moduli))
    s = rng.integers(-1, 2, n).astype(np.int8)
    s_prime = rng.integers(-1, 2, n).astype(np.int8)
    kb, ka = _keygen_hybrid(qp, len(q_moduli), blocks, s_prime, s, rng)
    b = rand_residues(rng, q_moduli, n)[0]
    a = rand_residues(rng, q_moduli, n)[0]
    bo, ao = O.hybrid_key_switch(qp, len(q_moduli), blocks, b, a, kb, ka)
    got = O.decrypt(q_plan, bo, ao, s)
    want = O.decrypt(q_plan, b, a, s_prime)
    # zero-noise key: acc_b + acc_a s = P a s' exactly mod QP, so the only
    # error is the mod-down rounding: |e| <= (1 + ||s||_1) / 2
    bound = (1 + int(np.abs(s).sum())) // 2 + 1
    for i, q in enumerate(q_moduli):
        diff = (got[i].astype(object) - want[i]) % q
        worst = max(abs(_centered(v, q)) for v in diff)
        assert worst <= bound
    # the error polynomial is the same integer in every limb (exact CRT)
    d0 = [(_centered((int(g) - int(w)) % q_moduli[0], q_moduli[0])) for g, w in zip(got[0], want[0])]
    for i, q in enumerate(q_moduli[1:], 1):
        di = [(_centered((int(g) - int(w)) % q, q)) for g, w in zip(got[i], want[i])]
        assert di == d0


def test_hybrid_key_switch_matches_explicit_composition():
    # b + ModDown(sum_j NTT^-1(NTT(lift(d_j)) kb_j)) built from the separately
    # pinned oracle pieces (gadget digits, big-integer mod-down)
    rng = np.random.default_rng(29)
    q_moduli = [Q0[0]] + Q40[:3]
    p_moduli = P60[:2]
    blocks = [2, 2]
    n = 1024
    qp = O.RnsPlan(q_moduli + p_moduli, n)
    kb = rand_residues(rng, qp.moduli, n, 2)
    ka = rand_residues(rng, qp.moduli, n, 2)
    b = rand_residues(rng, q_moduli, n)[0]
    a = rand_residues(rng, q_moduli, n)[0]
    bo, ao = O.hybrid_key_switch(qp, 4, blocks, b, a, kb, ka)
    Q = math.prod(q_moduli)
    acc_b = np.zeros((6, n), dtype=object)
    acc_a = np.zeros((6, n), dtype=object)
    start = 0
    for j, sz in enumerate(blocks):
        Qj = math.prod(q_moduli[start:start + sz])
        digit = np.array([[crt(a[:, c], q_moduli) % Qj % q for c in range(n)] for q in qp.moduli],
                         dtype=np.uint64)
        dn = qp.forward(digit)
        for i, q in enumerate(qp.moduli):
            acc_b[i] = (acc_b[i] + dn[i].astype(object) * kb[j, i].astype(object)) % q
            acc_a[i] = (acc_a[i] + dn[i].astype(object) * ka[j, i].astype(object)) % q
        start += sz
    ib = qp.inverse(acc_b.astype(np.uint64))
    ia = qp.inverse(acc_a.astype(np.uint64))
    Pv = math.prod(p_moduli)
    QPv = Q * Pv
    for c in range(0, n, 37):
        for comp, acc, out, add in ((0, ib, bo, b), (1, ia, ao, None)):
            v = crt(acc[:, c], qp.moduli)
            cen = v - QPv if v > QPv // 2 else v
            r = round_div_away(cen, Pv)
            for i, q in enumerate(q_moduli):
                e = (r + (int(add[i, c]) if add is not None else 0)) % q
                assert int(out[i, c]) == e


def test_tensor_matches_naive_products():
    rng = np.random.default_rng(31)
    moduli = Q40[:2]
    n = 1024
    plan = O.RnsPlan(moduli, n)
    b0, a0, b1, a1 = (rand_residues(rng, moduli, n)[0] for _ in range(4))
    c0, c1, c2 = O.tensor(plan, b0, a0, b1, a1)
    for i, q in enumerate(moduli):
        assert (c0[i] == O.negacyclic_mul_naive(b0[i], b1[i], q)).all()
        t = O.negacyclic_mul_naive(b0[i], a1[i], q).astype(object) + \
            O.negacyclic_mul_naive(a0[i], b1[i], q)
        assert (c1[i] == (t % q).astype(np.uint64)).all()
        assert (c2[i] == O.negacyclic_mul_naive(a0[i], a1[i], q)).all()


# ------------------------------------------------------------------ samplers
def test_chacha20_block_known_answer():
    # ChaCha20 with all-zero key and nonce, block 0 (the published keystream
    # 76 b8 e0 ad a0 f1 3d 90 ...), as rand_chacha's own tests use it
    blk = O.chacha20_block([0] * 8, 0)
    assert blk[:4].tolist() == [0xade0b876, 0x903df1a0, 0xe56a5d40, 0x28bd8653]


def test_samplers_ranges_and_determinism():
    r1, r2 = O.Rng(0), O.Rng(0)
    q = Q40[0]
    u1, u2 = r1.uniform(q, 4096), r2.uniform(q, 4096)
    assert (u1 == u2).all() and int(u1.max()) < q
    g = O.Rng(1).gaussian(3.2, 20000)
    assert abs(g.mean()) < 0.1 and abs(g.std() - 3.2) < 0.1 and np.abs(g).max() <= 52
    t = O.Rng(2).ternary(3000)
    assert set(np.unique(t).tolist()) <= {-1, 0, 1}
    assert all(abs((t == v).mean() - 1 / 3) < 0.04 for v in (-1, 0, 1))


@pytest.mark.parametrize("blocks,K", [([2, 2], 2), ([1, 1, 1, 1], 0), ([3, 1], 1)])
def test_hoisted_canonical_digit_identity(blocks, K):
    """The identity exact hoisting of canonical digits rests on, in the
    oracle: NTT of the canonical digits of sigma_k(x) = the NTT-domain Galois
    permutation of the NTT of the digits of x, plus (Q_j mod p_i) times
    NTT(neg_k) (neg_k = 1 where sigma_k negates), unless a block value is
    zero; with a zero block value at a negated source the identity fails
    there, which is why such batches take the canonical path."""
    n = 64
    L = sum(blocks)
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    orc = O.RnsPlan(basis, n)
    rng = np.random.default_rng(sum(blocks) * 10 + K)
    x = np.stack([rng.integers(0, qi, n, dtype=np.uint64) for qi in q])

    def digits(v):
        out, start = [], 0
        for sz in blocks:
            blk = list(range(start, start + sz))
            others = [i for i in range(len(basis)) if i not in blk]
            d = np.empty((len(basis), n), dtype=np.uint64)
            d[blk] = v[start:start + sz]
            d[others] = O.extend_basis(v[start:start + sz], [basis[i] for i in blk], [basis[i] for i in others])
            out.append(d)
            start += sz
        return out

    for r in (1, 3, 7):
        k = O.rotation_exponent_left(n, r)
        kinv = pow(k, -1, 2 * n)
        neg = np.array([1 if (t * kinv) % (2 * n) >= n else 0 for t in range(n)], dtype=np.uint64)
        rot = np.stack([O.apply_automorphism(x[i], qi, k) for i, qi in enumerate(q)])
        want = [orc.forward(d) for d in digits(rot)]
        base = [orc.forward(d) for d in digits(x)]
        start = 0
        for j, sz in enumerate(blocks):
            Qj = int(np.prod([int(v) for v in q[start:start + sz]], dtype=object))
            for i, p in enumerate(basis):
                c = Qj % p
                got = (O.permute_ntt_galois(base[j][i], k).astype(object) +
                       c * O.Plan(p, n).forward(neg).astype(object)) % p
                assert (got.astype(np.uint64) == want[j][i]).all(), (r, j, i)
            start += sz
    # a zero block value at a negated source breaks it for that block
    x[0:blocks[0], 5] = 0
    k = O.rotation_exponent_left(n, 1)
    kinv = pow(k, -1, 2 * n)
    neg = np.array([1 if (t * kinv) % (2 * n) >= n else 0 for t in range(n)], dtype=np.uint64)
    negated_sources = {(t * kinv) % (2 * n) - n for t in range(n) if (t * kinv) % (2 * n) >= n}
    rot = np.stack([O.apply_automorphism(x[i], qi, k) for i, qi in enumerate(q)])
    want = orc.forward(digits(rot)[0])
    base = orc.forward(digits(x)[0])
    Q0b = int(np.prod([int(v) for v in q[:blocks[0]]], dtype=object))
    i = len(basis) - 1 if K else blocks[0]  # a limb outside block 0
    p = basis[i]
    got = (O.permute_ntt_galois(base[i], k).astype(object) +
           (Q0b % p) * O.Plan(p, n).forward(neg).astype(object)) % p
    assert (got.astype(np.uint64) == want[i]).all() == (5 not in negated_sources)
