"""Oracle pinning for the bounded (balanced base-2^w) gadget family:
BoundedGadgetLayout (grafting/bounded_gadget.rs), bounded relinearization
(grafting/bounded_key_switch.rs) and the hybrid-QP Galois key switch
(ckks/hybrid_rns_galois_key.rs).  Balanced digits in [-B/2, B/2) whose
weighted sum is the centered coefficient are unique, so reconstruction plus
the digit bound pin the decomposition; the key switches are pinned by exact
decryption identities with zero-noise keys and by a big-integer composition."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

# bounded_gadget.rs:419-425 test basis
SMALL = [12_289, 40_961, 65_537]
# parameters.rs:397-426 (research 65536 pool) and :265-279 (research 8192)
POOL65536 = [0x7fffffffba0001, 0x7fffffffaa0001, 0x7fffffff7e0001, 0x7fffffff380001,
             0x7ffffffef00001, 0x7ffffffeba0001, 0x7ffffffeac0001, 0x7ffffffe700001,
             0x7ffffffe600001, 0x7ffffffe4c0001, 0x7ffffffe220001, 0x7ffffffe160001,
             0x7ffffffdec0001]
R8192 = [0xfffffdc001, 0xfffff4c001, 0xfffff3c001, 0xffffe80001, 0xffffe74001]


def _rns(coeffs, moduli):
    return np.array([[int(c) % m for c in coeffs] for m in moduli], dtype=np.uint64)


def _reconstruct(d, w):
    return [sum(int(d[j, c]) << (w * j) for j in range(d.shape[0])) for c in range(d.shape[1])]


def _centered(v, q):
    v = int(v) % q
    return v - q if v > q // 2 else v


def test_digit_count_known_answers():
    # hybrid_rns_galois_key.rs:1182-1183 (Q = 13 limbs of the 65536 pool, w = 32)
    assert O.bounded_digit_count(POOL65536, 32) == 23
    # bounded_gadget.rs:755-756 (research 8192, w = 20)
    assert O.bounded_digit_count(R8192, 20) == 11
    # closed form ceil((bits(Q) + 1) / w)
    for moduli in (SMALL, R8192, [Q0[0]] + Q40[:3]):
        Q = math.prod(moduli)
        for w in (2, 7, 20, 32, 61, 63):
            assert O.bounded_digit_count(moduli, w) == -(-(Q.bit_length() + 1) // w)


@pytest.mark.parametrize("w", [2, 4, 8, 12, 16, 20, 24, 32])
def test_decomposition_reconstructs_reference_vectors(w):
    # bounded_gadget.rs:587-604 and the deterministic campaign :637-660
    Q = math.prod(SMALL)
    cases = [[0, 1, 17, 42, 1_001, Q // 3, Q // 2, Q - 42, Q - 1]]
    cases += [[(97 * seed + 31 * i + 7 * i * i + 11) % Q for i in range(64)] for seed in (0, 1, 17, 63)]
    cases += [[(97 * i + 31 * i * i + 11) % Q for i in range(128)]]  # :606-623
    for coeffs in cases:
        d = O.bounded_decompose(_rns(coeffs, SMALL), SMALL, w)
        assert d.shape[0] == O.bounded_digit_count(SMALL, w)
        assert d.min() >= -(1 << (w - 1)) and d.max() < (1 << (w - 1))
        rec = _reconstruct(d, w)
        assert rec == [_centered(c, Q) for c in coeffs]
        assert [r % Q for r in rec] == coeffs


def test_negative_centered_values_reconstruct():
    # bounded_gadget.rs:662-675
    Q = math.prod(SMALL)
    coeffs = [Q - 1, Q - 2, Q - 17, Q - 1_001]
    d = O.bounded_decompose(_rns(coeffs, SMALL), SMALL, 8)
    assert _reconstruct(d, 8) == [-1, -2, -17, -1_001]


def test_research_8192_roundtrip_above_u128():
    # bounded_gadget.rs:737-775
    Q = math.prod(R8192)
    coeffs = [0, 1, (1 << 150) + 0x1234_5678, Q - 1, Q - 17]
    d = O.bounded_decompose(_rns(coeffs, R8192), R8192, 20)
    assert d.shape[0] == 11
    assert [r % Q for r in _reconstruct(d, 20)] == coeffs
    assert np.abs(d).max() <= 1 << 19


@pytest.mark.parametrize("w", [13, 32, 61, 63])
def test_decomposition_at_config_sizes(w):
    moduli = [Q0[0]] + Q40[:8] + P60[:2]
    Q = math.prod(moduli)
    rng = np.random.default_rng(w)
    x = rand_residues(rng, moduli, 256)[0]
    x[:, 0] = 0
    x[:, 1] = np.array([(Q // 2) % m for m in moduli], dtype=np.uint64)
    x[:, 2] = np.array([(Q // 2 + 1) % m for m in moduli], dtype=np.uint64)
    d = O.bounded_decompose(x, moduli, w)
    vals = [sum(int(x[i, c]) * (Q // m) * pow(Q // m, -1, m) for i, m in enumerate(moduli)) % Q
            for c in range(256)]
    assert _reconstruct(d, w) == [_centered(v, Q) for v in vals]
    assert d.min() >= -(1 << (w - 1)) and d.max() < (1 << (w - 1))


def test_lift_signed_is_rem_euclid():
    q = Q40[0]
    d = np.array([0, 1, -1, q - 1, -(q - 1), (1 << 62) - 1, -(1 << 62)], dtype=np.int64)
    assert O.lift_signed(d, q).tolist() == [int(v) % q for v in d.tolist()]


def _square(plan, s):
    sn = plan.forward(np.stack([np.array([v % q for v in s.tolist()], dtype=np.uint64)
                                for q in plan.moduli]))
    return plan.inverse(np.stack([(sn[i].astype(object) ** 2 % q).astype(np.uint64)
                                  for i, q in enumerate(plan.moduli)]))


def _keygen_bounded(plan, q_limbs, w, target, s, rng):
    """Entry j encrypts P * B^j * target under s (P = 1 without special
    primes), NTT domain: BoundedRnsMultiplicationKey (bounded_key_switch.rs:
    27-103) and PreparedHybridRnsGaloisKey (hybrid_rns_galois_key.rs:61-170)."""
    moduli = plan.moduli
    n = plan.n
    Pv = math.prod(moduli[q_limbs:])
    count = O.bounded_digit_count(moduli[:q_limbs], w)
    kb = np.zeros((count, len(moduli), n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    for j in range(count):
        a = rand_residues(rng, moduli, n)[0]
        msg = np.stack([(target[i].astype(object) * (Pv * (1 << (w * j)) % q) % q).astype(np.uint64)
                        for i, q in enumerate(moduli)])
        b = O.encrypt(plan, msg, np.zeros(n, dtype=np.int64), s, a)
        kb[j] = plan.forward(b)
        ka[j] = plan.forward(a)
    return kb, ka


@pytest.mark.parametrize("w", [20, 61])
def test_bounded_relinearization_is_exact_with_zero_noise(w):
    rng = np.random.default_rng(31 + w)
    moduli = [Q0[0]] + Q40[:3]
    n = 256
    plan = O.RnsPlan(moduli, n)
    s = rng.integers(-1, 2, n).astype(np.int8)
    s2 = _square(plan, s)
    kb, ka = _keygen_bounded(plan, len(moduli), w, s2, s, rng)
    c0 = rand_residues(rng, moduli, n)[0]
    c2 = rand_residues(rng, moduli, n)[0]
    bo, ao = O.bounded_key_switch(plan, len(moduli), w, c0, c2, kb, ka)
    # out_b + out_a s = c0 + sum_j d_j B^j s^2 = c0 + c2 s^2 exactly
    got = O.decrypt(plan, bo, ao, s)
    f2, fs = plan.forward(c2), plan.forward(s2)
    c2s2 = plan.inverse(np.stack([(f2[i].astype(object) * fs[i].astype(object) % q).astype(np.uint64)
                                  for i, q in enumerate(moduli)]))
    for i, q in enumerate(moduli):
        want = (c0[i].astype(object) + c2s2[i].astype(object)) % q
        assert (got[i] == want).all()


@pytest.mark.parametrize("w,K", [(61, 1), (32, 2), (20, 3)])
def test_hybrid_qp_key_switch_decrypts_within_mod_down_rounding(w, K):
    rng = np.random.default_rng(41 + w)
    q_moduli = [Q0[0]] + Q40[:3]
    qp = O.RnsPlan(q_moduli + P60[:K], 256)
    q_plan = qp.prefix(len(q_moduli))
    n = qp.n
    s = rng.integers(-1, 2, n).astype(np.int8)
    s_prime = rng.integers(-1, 2, n).astype(np.int8)
    target = np.stack([np.array([v % q for v in s_prime.tolist()], dtype=np.uint64) for q in qp.moduli])
    kb, ka = _keygen_bounded(qp, len(q_moduli), w, target, s, rng)
    b = rand_residues(rng, q_moduli, n)[0]
    a = rand_residues(rng, q_moduli, n)[0]
    bo, ao = O.bounded_key_switch(qp, len(q_moduli), w, b, a, kb, ka)
    got = O.decrypt(q_plan, bo, ao, s)
    want = O.decrypt(q_plan, b, a, s_prime)
    # acc_b + acc_a s = P (b + a s') exactly mod QP; mod-down rounds each
    # component once: |error| <= (1 + ||s||_1) / 2
    bound = (1 + int(np.abs(s).sum())) // 2 + 1
    d0 = None
    for i, q in enumerate(q_moduli):
        di = [_centered(int(g) - int(v), q) for g, v in zip(got[i], want[i])]
        assert max(abs(v) for v in di) <= bound
        d0 = di if d0 is None else d0
        assert di == d0


def test_hybrid_qp_key_switch_matches_big_integer_composition():
    """hybrid_rns_galois_key.rs:627-780 term by term with Python integers:
    QP sums from P*b, digit MACs, then round(centered(x) / P) per component."""
    rng = np.random.default_rng(53)
    q_moduli = [Q0[0]] + Q40[:2]
    p_moduli = P60[:2]
    w = 40
    qp = O.RnsPlan(q_moduli + p_moduli, 64)
    n = qp.n
    count = O.bounded_digit_count(q_moduli, w)
    kb = rand_residues(rng, qp.moduli, n, count)
    ka = rand_residues(rng, qp.moduli, n, count)
    b = rand_residues(rng, q_moduli, n)[0]
    a = rand_residues(rng, q_moduli, n)[0]
    bo, ao = O.bounded_key_switch(qp, 3, w, b, a, kb, ka)
    Q = math.prod(q_moduli)
    Pv = math.prod(p_moduli)
    QP = Q * Pv
    a_int = [sum(int(a[i, c]) * (Q // m) * pow(Q // m, -1, m) for i, m in enumerate(q_moduli)) % Q
             for c in range(n)]
    a_c = [_centered(v, Q) for v in a_int]
    digits = []
    for c in range(n):
        v, row = a_c[c], []
        for _ in range(count):
            r = v % (1 << w)
            dgt = r - (1 << w) if r >= 1 << (w - 1) else r
            row.append(dgt)
            v = (v - dgt) >> w
        assert v == 0
        digits.append(row)
    # negacyclic products over the integers mod each QP prime (schoolbook)
    key_b = [[qp.inverse(kb[j])[i] for i in range(len(qp.moduli))] for j in range(count)]
    key_a = [[qp.inverse(ka[j])[i] for i in range(len(qp.moduli))] for j in range(count)]

    def negacyclic(x, y, q):
        out = [0] * n
        for i in range(n):
            if x[i] == 0:
                continue
            for k in range(n):
                t = x[i] * int(y[k])
                if i + k < n:
                    out[i + k] += t
                else:
                    out[i + k - n] -= t
        return [v % q for v in out]

    for comp, key in ((0, key_b), (1, key_a)):
        limbs = []
        for i, q in enumerate(qp.moduli):
            acc = [(Pv * int(b[i, c])) % q if (comp == 0 and i < 3) else 0 for c in range(n)]
            for j in range(count):
                dj = [digits[c][j] % q for c in range(n)]
                prod = negacyclic(dj, key[j][i], q)
                acc = [(u + v) % q for u, v in zip(acc, prod)]
            limbs.append(acc)
        vals = [sum(limbs[i][c] * (QP // m) * pow(QP // m, -1, m) for i, m in enumerate(qp.moduli)) % QP
                for c in range(n)]
        cen = [_centered(v, QP) for v in vals]
        rounded = [(2 * v + Pv) // (2 * Pv) for v in cen]  # P odd: no ties
        got = bo if comp == 0 else ao
        for i, q in enumerate(q_moduli):
            assert got[i].tolist() == [r % q for r in rounded], f"component {comp} limb {i}"


@pytest.mark.parametrize("w", [7, 20, 61, 63])
def test_hoisted_digit_automorphism_equals_canonical(w):
    """automorphism_hoisted(decompose(x)) == decompose(sigma(x)), the
    reference's own check (grafting/bounded_gadget.rs:428-488), including
    ties: coefficients whose digits sit at -2^(w-1), where negation carries."""
    n = 64
    moduli = [Q0[0]] + Q40[:3]
    Q = math.prod(moduli)
    rng = np.random.default_rng(w)
    x = np.stack([rng.integers(0, q, n, dtype=np.uint64) for q in moduli])
    half = 1 << (w - 1)
    ties = [-half, -half - (half << w), -(half << w), half - 1 + (-half << w)]
    for c, v in enumerate(ties):
        x[:, 2 * c + 1] = [v % q for q in moduli]
    d = O.bounded_decompose(x, moduli, w)
    assert (d == -half).any()
    for e in (3, 5, 2 * n - 1, 5 ** 7 % (2 * n), 13):
        sx = np.stack([O.apply_automorphism(x[i], q, e) for i, q in enumerate(moduli)])
        want = O.bounded_decompose(sx, moduli, w)
        assert (O.bounded_automorphism_hoisted(d, e, w) == want).all(), f"exponent {e}"
