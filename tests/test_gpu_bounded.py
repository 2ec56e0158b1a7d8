"""GPU parity of the bounded (balanced base-2^w) gadget family against the
oracle, residue for residue: the decomposition with its lifted digits
(grafting/bounded_gadget.rs), bounded relinearization without special primes
(grafting/bounded_key_switch.rs) and the hybrid-QP Galois key switch
(ckks/hybrid_rns_galois_key.rs)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu

POOL65536 = [0x7fffffffba0001, 0x7fffffffaa0001, 0x7fffffff7e0001, 0x7fffffff380001,
             0x7ffffffef00001, 0x7ffffffeba0001, 0x7ffffffeac0001, 0x7ffffffe700001,
             0x7ffffffe600001, 0x7ffffffe4c0001, 0x7ffffffe220001, 0x7ffffffe160001,
             0x7ffffffdec0001, 0x7ffffffdd00001, 0x7ffffffdce0001]


@pytest.mark.parametrize("logn,L,K,w", [(12, 4, 1, 61), (12, 4, 0, 20), (13, 9, 2, 32),
                                         (15, 13, 4, 45), (16, 25, 7, 63), (14, 25, 0, 7)])
def test_bounded_decomposition_bit_exact(pkg, ctx, logn, L, K, w):
    n = 1 << logn
    moduli = [Q0[0]] + Q40[:L - 1] + P60[:K]
    rng = np.random.default_rng(logn * 100 + w)
    x = rand_residues(rng, moduli[:L], n, polys=2)
    Q = math.prod(moduli[:L])
    # exact boundaries: 0, floor(Q/2), floor(Q/2)+1, Q-1
    for c, v in enumerate((0, Q // 2, Q // 2 + 1, Q - 1, 1)):
        x[0, :, c] = [v % q for q in moduli[:L]]
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    count = pkg.bounded_digit_count(plan, L, w)
    assert count == O.bounded_digit_count(moduli[:L], w)
    got = pkg.to_host(pkg.bounded_decompose(plan, L, w, pkg.to_device(x)))
    assert got.shape == (2, count, len(moduli), n)
    for p in range(2):
        d = O.bounded_decompose(x[p], moduli[:L], w)
        for t, q in enumerate(moduli):
            assert (got[p, :, t] == O.lift_signed(d, q)).all(), f"poly {p} limb {t}"


def _key(rng, plan_moduli, n, count):
    return rand_residues(rng, plan_moduli, n, count), rand_residues(rng, plan_moduli, n, count)


# digit counts 10, 5, 10, 21 and 31 over two ciphertexts: one key MAC launch
# of up to 8, 16 or 24 digits, and two launches of 24 + 7
@pytest.mark.parametrize("logn,L,w", [(12, 4, 20), (13, 6, 61), (16, 9, 40), (12, 6, 13), (12, 8, 11)])
def test_bounded_relinearize_bit_exact(pkg, ctx, logn, L, w):
    n = 1 << logn
    moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(7 * logn + w)
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    count = O.bounded_digit_count(moduli, w)
    kb, ka = _key(rng, moduli, n, count)
    key = pkg.BoundedKeySwitchKey(plan, L, w, kb, ka)
    ct3 = rand_residues(rng, moduli, n, polys=6).reshape(2, 3, L, n)
    got = pkg.to_host(key.relinearize(pkg.to_device(ct3)))
    orc = O.RnsPlan(moduli, n)
    for c in range(2):
        bo, ao = O.bounded_key_switch(orc, L, w, ct3[c, 0], ct3[c, 2], kb, ka)
        ao = np.stack([(ao[i].astype(object) + ct3[c, 1, i]) % q for i, q in enumerate(moduli)]).astype(np.uint64)
        assert (got[c, 0] == bo).all() and (got[c, 1] == ao).all(), f"ct {c}"


@pytest.mark.parametrize("logn,L,K,w,steps", [(12, 4, 1, 61, 1), (13, 9, 2, 32, 5), (15, 13, 4, 45, -3),
                                               (12, 6, 2, 13, 3)])
def test_hybrid_qp_rotation_bit_exact(pkg, ctx, logn, L, K, w, steps):
    n = 1 << logn
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(logn + w)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, w)
    kb, ka = _key(rng, qp, n, count)
    key = pkg.BoundedKeySwitchKey(plan, L, w, kb, ka)
    k = O.rotation_exponent_left(n, steps) if steps > 0 else O.rotation_exponent_right(n, -steps)
    ct = rand_residues(rng, q_moduli, n, polys=4).reshape(2, 2, L, n)
    got = pkg.to_host(key.apply_galois(pkg.to_device(ct), k))
    orc = O.RnsPlan(qp, n)
    for c in range(2):
        sb = np.stack([O.apply_automorphism(ct[c, 0, i], q, k) for i, q in enumerate(q_moduli)])
        sa = np.stack([O.apply_automorphism(ct[c, 1, i], q, k) for i, q in enumerate(q_moduli)])
        bo, ao = O.bounded_key_switch(orc, L, w, sb, sa, kb, ka)
        assert (got[c, 0] == bo).all() and (got[c, 1] == ao).all(), f"ct {c}"


def test_reference_rotation_benchmark_shape(pkg, ctx):
    """The reference's UHD rotation configuration (hybrid_rns_galois_key.rs:
    1152-1200): N = 2^16, Q = 13 limbs of the 55-bit pool, P = 2 limbs,
    base_log 32 -> 23 digits; one ciphertext against the oracle."""
    n = 1 << 16
    q_moduli, p_moduli = POOL65536[:13], POOL65536[13:15]
    qp = q_moduli + p_moduli
    rng = np.random.default_rng(65536)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, 32)
    assert count == 23
    kb, ka = _key(rng, qp, n, count)
    key = pkg.BoundedKeySwitchKey(plan, 13, 32, kb, ka)
    ct = rand_residues(rng, q_moduli, n, polys=2).reshape(1, 2, 13, n)
    got = pkg.to_host(key.apply_galois(pkg.to_device(ct), 3))
    sb = np.stack([O.apply_automorphism(ct[0, 0, i], q, 3) for i, q in enumerate(q_moduli)])
    sa = np.stack([O.apply_automorphism(ct[0, 1, i], q, 3) for i, q in enumerate(q_moduli)])
    bo, ao = O.bounded_key_switch(O.RnsPlan(qp, n), 13, 32, sb, sa, kb, ka)
    assert (got[0, 0] == bo).all() and (got[0, 1] == ao).all()


def _with_ties(x, moduli, w):
    """Puts coefficients whose balanced digits sit at -2^(w-1) (where a
    negated digit vector carries) at odd positions 1, 3, 5, 7."""
    half = 1 << (w - 1)
    for c, v in enumerate((-half, -half - (half << w), -(half << w), half - 1 + (-half << w))):
        x[..., :, 2 * c + 1] = [v % q for q in moduli]
    return x


@pytest.mark.parametrize("logn,L,w", [(12, 4, 20), (13, 9, 61), (14, 6, 63)])
def test_bounded_digits_and_hoisted_automorphism(pkg, ctx, logn, L, w):
    """ckks_bounded_digits equals the oracle's balanced digits, and
    ckks_bounded_digits_automorphism equals both the oracle's
    automorphism_hoisted and the digits of the rotated polynomial."""
    n = 1 << logn
    moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(3 * logn + w)
    x = _with_ties(rand_residues(rng, moduli, n, polys=2), moduli, w)
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    dx = pkg.to_device(x)
    digits = pkg.bounded_digits(plan, L, w, dx)
    got = digits.cpu().numpy()
    for p in range(2):
        assert (got[p] == O.bounded_decompose(x[p], moduli, w)).all(), f"poly {p}"
    for e in (2 * n - 1, O.rotation_exponent_left(n, 3)):
        hd = pkg.bounded_digits_automorphism(plan, L, w, digits, e).cpu().numpy()
        rot = pkg.bounded_digits(plan, L, w, plan.apply_automorphism(dx, e, n_limbs=L)).cpu().numpy()
        assert (hd == rot).all(), f"exponent {e}: hoisted != decomposition of the rotation"
        if logn == 12:
            for p in range(2):
                assert (hd[p] == O.bounded_automorphism_hoisted(got[p], e, w)).all()


@pytest.mark.parametrize("logn,L,K,w", [(12, 4, 1, 61), (13, 9, 2, 32), (13, 6, 0, 20)])
def test_hoisted_rotation_matches_apply_galois(pkg, ctx, logn, L, K, w):
    """Rotations from digits decomposed once give the residues of
    ckks_apply_galois for every exponent, with and without special primes."""
    n = 1 << logn
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(logn * 7 + w + K)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, w)
    kb, ka = _key(rng, qp, n, count)
    key = pkg.BoundedKeySwitchKey(plan, L, w, kb, ka)
    ct = rand_residues(rng, q_moduli, n, polys=6).reshape(3, 2, L, n)
    ct[:, 1] = _with_ties(ct[:, 1], q_moduli, w)
    dct = pkg.to_device(ct)
    digits = key.digits(dct)
    for steps in (1, 5, -2):
        k = O.rotation_exponent_left(n, steps) if steps > 0 else O.rotation_exponent_right(n, -steps)
        want = pkg.to_host(key.apply_galois(dct, k))
        got = pkg.to_host(key.apply_galois_hoisted(dct, k, digits))
        assert (got == want).all(), f"steps {steps}"


@pytest.mark.parametrize("logn,L,K,w", [(12, 4, 1, 61), (13, 9, 2, 32), (14, 6, 0, 20)])
def test_ntt_domain_hoisted_rotations(pkg, ctx, logn, L, K, w):
    """Digits decomposed and transformed once, rotated by permutation in the
    NTT domain: the residues of apply_galois for every exponent; exponents
    that negate a tie coefficient (2N - 1 here) take the canonical path."""
    n = 1 << logn
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(logn * 11 + w + K)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, w)
    kb, ka = _key(rng, qp, n, count)
    key = pkg.BoundedKeySwitchKey(plan, L, w, kb, ka)
    other = pkg.BoundedKeySwitchKey(plan, L, w, *_key(rng, qp, n, count))
    ct = rand_residues(rng, q_moduli, n, polys=6).reshape(3, 2, L, n)
    ct[1, 1] = _with_ties(ct[1, 1], q_moduli, w)  # ties at coefficients 1, 3, 5, 7 of ct 1
    dct = pkg.to_device(ct)
    hoisted = key.hoist(dct)
    try:
        for k in (O.rotation_exponent_left(n, 1), O.rotation_exponent_left(n, 7), 2 * n - 1, 5):
            for kk in (key, other):
                want = pkg.to_host(kk.apply_galois(dct, k))
                got = pkg.to_host(kk.rotate_hoisted(hoisted, dct, k))
                assert (got == want).all(), f"exponent {k}"
    finally:
        hoisted.close()


def test_hoisted_rotations_split_over_two_streams(pkg, ctx):
    """Hoisted rotations of a batch split over two streams (lane offsets of
    the hoisted digits, per-lane MAC workspaces) give the single-stream
    residues of apply_galois."""
    n, L, K, w = 1 << 13, 6, 2, 32
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(99)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, w)
    key = pkg.BoundedKeySwitchKey(plan, L, w, *_key(rng, qp, n, count))
    dct = pkg.to_device(rand_residues(rng, q_moduli, n, polys=18).reshape(9, 2, L, n))
    k = O.rotation_exponent_left(n, 3)
    ctx.set_key_switch_split(ctx.KS_SPLIT_OFF)
    want = pkg.to_host(key.apply_galois(dct, k))
    hoisted = key.hoist(dct)
    try:
        ctx.set_key_switch_split(ctx.KS_SPLIT_ON)
        got = pkg.to_host(key.rotate_hoisted(hoisted, dct, k))
        got_canonical = pkg.to_host(key.apply_galois(dct, k))
    finally:
        ctx.set_key_switch_split(ctx.KS_SPLIT_AUTO)
        hoisted.close()
    assert (got == want).all() and (got_canonical == want).all()


@pytest.mark.parametrize("logn,variant", [(13, 0), (13, 1), (14, 2), (14, 3)])
def test_hoisted_rotations_in_every_ntt_variant(pkg, ctx, logn, variant):
    """The hoisted rotation's inverse NTT reads the key MAC output through the
    Galois permutation in its first pass; every two-pass kernel variant
    (4096/8192-residue tiles, plain/blocked inter-pass layout) gives the
    residues of the oracle's rotation."""
    n, L, K, w = 1 << logn, 4, 1, 32
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(logn * 13 + variant)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, w)
    kb, ka = _key(rng, qp, n, count)
    key = pkg.BoundedKeySwitchKey(plan, L, w, kb, ka)
    ct = rand_residues(rng, q_moduli, n, polys=4).reshape(2, 2, L, n)
    k = O.rotation_exponent_left(n, 3)
    saved = ctx.ntt_variant(logn)
    ctx.set_ntt_variant(logn, variant)
    try:
        hoisted = key.hoist(pkg.to_device(ct))
        try:
            got = pkg.to_host(key.rotate_hoisted(hoisted, pkg.to_device(ct), k))
        finally:
            hoisted.close()
    finally:
        ctx.set_ntt_variant(logn, saved)
    orc = O.RnsPlan(qp, n)
    for c in range(2):
        sb = np.stack([O.apply_automorphism(ct[c, 0, i], q, k) for i, q in enumerate(q_moduli)])
        sa = np.stack([O.apply_automorphism(ct[c, 1, i], q, k) for i, q in enumerate(q_moduli)])
        bo, ao = O.bounded_key_switch(orc, L, w, sb, sa, kb, ka)
        assert (got[c, 0] == bo).all() and (got[c, 1] == ao).all(), f"ct {c}"


def test_hoisted_key_cache_evicts_and_rebuilds(pkg, ctx):
    """A key keeps at most 8 permuted copies for hoisted rotations; rotating by
    12 exponents, then by the first ones again (evicted and rebuilt), gives the
    residues of apply_galois every time."""
    n, L, K, w = 1 << 12, 4, 1, 61
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(4242)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    count = O.bounded_digit_count(q_moduli, w)
    key = pkg.BoundedKeySwitchKey(plan, L, w, *_key(rng, qp, n, count))
    dct = pkg.to_device(rand_residues(rng, q_moduli, n, polys=4).reshape(2, 2, L, n))
    hoisted = key.hoist(dct)
    try:
        steps = list(range(1, 13)) + [1, 2, 3, 12]
        for s in steps:
            k = O.rotation_exponent_left(n, s)
            want = pkg.to_host(key.apply_galois(dct, k))
            got = pkg.to_host(key.rotate_hoisted(hoisted, dct, k))
            assert (got == want).all(), f"rotation by {s}"
    finally:
        hoisted.close()
