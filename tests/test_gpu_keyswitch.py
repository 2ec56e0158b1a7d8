"""GPU parity of key switching, rotation, relinearization and the
multiply-relinearize-rescale pipeline against the oracle (bit-exact), plus
decryption semantics with noisy keys."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def _keys(rng, moduli, n, dnum):
    return rand_residues(rng, moduli, n, dnum), rand_residues(rng, moduli, n, dnum)


def _ct(rng, moduli, n, n_cts):
    return rand_residues(rng, moduli, n, 2 * n_cts).reshape(n_cts, 2, len(moduli), n)


@pytest.mark.parametrize("logn,L,blocks", [(12, 4, [1, 1, 1, 1]), (12, 4, [2, 2]), (13, 5, [2, 1, 2]),
                                           (16, 6, [1] * 6)])
def test_rns_gadget_key_switch_bit_exact(pkg, ctx, logn, L, blocks):
    n = 1 << logn
    moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(logn * 10 + L)
    kb, ka = _keys(rng, moduli, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    key = pkg.RnsKeySwitchKey(plan, blocks, kb, ka)
    ct = _ct(rng, moduli, n, 3)
    out = pkg.to_host(key.switch(pkg.to_device(ct)))
    orc = O.RnsPlan(moduli, n)
    for c in range(3):
        bo, ao = O.rns_key_switch(orc, blocks, ct[c, 0], ct[c, 1], kb, ka)
        assert (out[c, 0] == bo).all() and (out[c, 1] == ao).all(), f"ct {c}"


@pytest.mark.parametrize("engine", ["mfma", "valu"])
@pytest.mark.parametrize("logn,L,K,blocks", [(12, 4, 1, [2, 2]), (12, 4, 2, [1, 3]), (13, 9, 4, [5, 4]),
                                             (15, 13, 4, [5, 4, 4]), (16, 25, 7, [9, 8, 8]),
                                             (16, 17, 4, [5, 4, 4, 4]), (12, 12, 3, [3, 9])])
def test_hybrid_key_switch_bit_exact(pkg, ctx, logn, L, K, blocks, engine):
    """Both base-conversion engines (int8 MFMA and VALU) against the oracle."""
    n = 1 << logn
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(logn * 100 + L)
    kb, ka = _keys(rng, qp, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    n_cts = 2
    ct = _ct(rng, q_moduli, n, n_cts)
    # residues at the extremes exercise the largest chunk sums
    ct[0, 1, :, :8] = np.array(q_moduli, dtype=np.uint64)[:, None] - 1
    ct[0, 1, :, 8:16] = 0
    ctx.set_base_conversion(ctx.BASECONV_VALU if engine == "valu" else ctx.BASECONV_MFMA)
    try:
        out = pkg.to_host(key.switch(pkg.to_device(ct)))
    finally:
        ctx.set_base_conversion(ctx.BASECONV_AUTO)
    orc = O.RnsPlan(qp, n)
    for c in range(n_cts):
        bo, ao = O.hybrid_key_switch(orc, L, blocks, ct[c, 0], ct[c, 1], kb, ka)
        assert (out[c, 0] == bo).all() and (out[c, 1] == ao).all(), f"ct {c}"


def test_hybrid_mod_down_centering_boundaries(pkg, ctx):
    """Accumulators that land exactly on +-P/2 after the key MAC: zero key
    b-entries except a constant chosen so acc_P = floor(P/2), floor(P/2)+1."""
    n = 1 << 12
    q_moduli = [Q0[0]] + Q40[:3]
    pm = P60[:2]
    qp = q_moduli + pm
    import math
    Pv = math.prod(pm)
    rng = np.random.default_rng(5)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    orc = O.RnsPlan(qp, n)
    # a = constant 1 polynomial -> every digit is the constant 1; the key
    # b-entry of block 0 encodes a constant c, so acc_b = c exactly
    for cval in [Pv // 2, Pv // 2 + 1, Pv // 2 - 1, Pv - 1, 1, 0]:
        kb = np.zeros((2, len(qp), n), dtype=np.uint64)
        ka = np.zeros_like(kb)
        const = np.zeros((len(qp), n), dtype=np.uint64)
        for i, q in enumerate(qp):
            const[i, 0] = cval % q
        kb[0] = orc.forward(const)
        key = pkg.HybridKeySwitchKey(plan, 4, [2, 2], kb, ka)
        ct = np.zeros((1, 2, 4, n), dtype=np.uint64)
        ct[0, 1, :, 0] = 1  # a = 1 -> digit_0 = 1, digit_1 = 1
        out = pkg.to_host(key.switch(pkg.to_device(ct)))
        bo, ao = O.hybrid_key_switch(orc, 4, [2, 2], ct[0, 0], ct[0, 1], kb, ka)
        assert (out[0, 0] == bo).all() and (out[0, 1] == ao).all()


@pytest.mark.parametrize("hybrid", [False, True])
def test_rotation_bit_exact(pkg, ctx, hybrid):
    n = 1 << 13
    L = 5
    q_moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(77)
    if hybrid:
        basis, blocks = q_moduli + P60[:2], [3, 2]
    else:
        basis, blocks = q_moduli, [1] * L
    kb, ka = _keys(rng, basis, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = (pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka) if hybrid
           else pkg.RnsKeySwitchKey(plan, blocks, kb, ka))
    ct = _ct(rng, q_moduli, n, 2)
    orc = O.RnsPlan(basis, n)
    for steps in [1, 7, 2 ** 11]:
        k = O.rotation_exponent_left(n, steps)
        out = pkg.to_host(key.apply_galois(pkg.to_device(ct), k))
        for c in range(2):
            rb = np.stack([O.apply_automorphism(ct[c, 0, i], q, k) for i, q in enumerate(q_moduli)])
            ra = np.stack([O.apply_automorphism(ct[c, 1, i], q, k) for i, q in enumerate(q_moduli)])
            if hybrid:
                bo, ao = O.hybrid_key_switch(orc, L, blocks, rb, ra, kb, ka)
            else:
                bo, ao = O.rns_key_switch(orc, blocks, rb, ra, kb, ka)
            assert (out[c, 0] == bo).all() and (out[c, 1] == ao).all()


@pytest.mark.parametrize("hybrid", [False, True])
def test_multiply_relinearize_rescale_bit_exact(pkg, ctx, hybrid):
    n = 1 << 12
    L = 4
    q_moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(91)
    basis, blocks = (q_moduli + P60[:1], [2, 2]) if hybrid else (q_moduli, [1] * L)
    kb, ka = _keys(rng, basis, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = (pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka) if hybrid
           else pkg.RnsKeySwitchKey(plan, blocks, kb, ka))
    ct0, ct1 = _ct(rng, q_moduli, n, 2), _ct(rng, q_moduli, n, 2)
    d0, d1 = pkg.to_device(ct0), pkg.to_device(ct1)
    rel = pkg.to_host(key.relinearize(pkg.to_device(np.concatenate([
        np.stack(O.tensor(O.RnsPlan(q_moduli, n), ct0[c, 0], ct0[c, 1], ct1[c, 0], ct1[c, 1]))[None]
        for c in range(2)]))))
    out = pkg.to_host(key.multiply_relinearize_rescale(d0, d1))
    orc_b = O.RnsPlan(basis, n)
    orc_q = O.RnsPlan(q_moduli, n)
    for c in range(2):
        c0, c1, c2 = O.tensor(orc_q, ct0[c, 0], ct0[c, 1], ct1[c, 0], ct1[c, 1])
        if hybrid:
            sb, sa = O.hybrid_key_switch(orc_b, L, blocks, np.zeros_like(c0), c2, kb, ka)
        else:
            sb, sa = O.rns_key_switch(orc_b, blocks, np.zeros_like(c0), c2, kb, ka)
        rb = np.array([(c0[i].astype(object) + sb[i]) % q for i, q in enumerate(q_moduli)], dtype=np.uint64)
        ra = np.array([(c1[i].astype(object) + sa[i]) % q for i, q in enumerate(q_moduli)], dtype=np.uint64)
        assert (rel[c, 0] == rb).all() and (rel[c, 1] == ra).all()
        bo, ao = O.rescale_ct(q_moduli, rb, ra)
        assert (out[c, 0] == bo).all() and (out[c, 1] == ao).all()


def test_rotation_decrypts_with_noisy_hybrid_key(pkg, ctx):
    """Semantics: a Gaussian-noise Galois key for sigma_k(s) -> s; the
    rotated ciphertext decrypts under s to sigma_k(m) plus small noise.
    P (two 60-bit primes) exceeds every block modulus Q_j (~2^100), which
    keeps the key-switch noise sum_j d_j e_j / P small."""
    import math
    n = 1 << 12
    L, K = 4, 2
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    blocks = [2, 2]
    rng = O.Rng(0)
    s = rng.ternary(n)
    k = O.rotation_exponent_left(n, 1)
    # sigma_k(s) as a ternary vector
    s_rot3 = O.apply_automorphism(np.array([v % 3 for v in s.tolist()], dtype=np.uint64), 3, k)
    s_rot = np.array([{0: 0, 1: 1, 2: -1}[int(v)] for v in s_rot3], dtype=np.int8)
    orc_qp = O.RnsPlan(qp, n)
    orc_q = O.RnsPlan(q_moduli, n)
    Pv = math.prod(P60[:K])
    kb = np.zeros((2, len(qp), n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    start = 0
    for j, sz in enumerate(blocks):
        a = np.stack([rng.uniform(q, n) for q in qp])
        msg = np.zeros((len(qp), n), dtype=np.uint64)
        for i in range(start, start + sz):
            msg[i] = np.array([(int(v) * Pv) % qp[i] for v in s_rot.tolist()], dtype=np.uint64)
        e = rng.gaussian(3.2, n)
        b = O.encrypt(orc_qp, msg, e, s, a)
        kb[j], ka[j] = orc_qp.forward(b), orc_qp.forward(a)
        start += sz
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    mi = np.random.default_rng(3).integers(-(1 << 30), 1 << 30, n)
    m = np.stack([np.array([int(v) % q for v in mi.tolist()], dtype=np.uint64) for q in q_moduli])
    a = np.stack([rng.uniform(q, n) for q in q_moduli])
    b = O.encrypt(orc_q, m, rng.gaussian(3.2, n), s, a)
    out = pkg.to_host(key.apply_galois(pkg.to_device(np.stack([b, a])[None]), k))
    dec = O.decrypt(orc_q, out[0, 0], out[0, 1], s)
    want = np.stack([O.apply_automorphism(m[i], q, k) for i, q in enumerate(q_moduli)])
    for i, q in enumerate(q_moduli):
        diff = (dec[i].astype(object) - want[i]) % q
        worst = max(min(int(v), q - int(v)) for v in diff)
        assert worst < 2 ** 20, worst


@pytest.mark.parametrize("case", ["wide_in_block", "basis_over_64"])
def test_hybrid_key_switch_integer_mode_bit_exact(pkg, ctx, case):
    """ModUp / ModDown keep their integer path where the FP64 one does not
    apply: a gadget block with a 60-bit prime after its first limb, and a
    key basis of more than 64 limbs (no FP64 target mask)."""
    if case == "wide_in_block":
        logn, blocks, K = 12, [2, 2, 1], 3
        q_moduli = [Q40[0], P60[3], Q40[1], P60[4], Q0[0]]
        p_moduli = P60[:3]
    else:
        logn, blocks, K = 12, [12] * 5 + [6], 4
        q_moduli = [Q0[0]] + O.ntt_primes(40, 1 << logn, 65)[:65]
        p_moduli = O.ntt_primes(61, 1 << logn, K)
    n = 1 << logn
    qp = q_moduli + p_moduli
    L = len(q_moduli)
    assert sum(blocks) == L and len(set(qp)) == len(qp)
    rng = np.random.default_rng(len(qp))
    kb, ka = _keys(rng, qp, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    ct = _ct(rng, q_moduli, n, 1)
    ct[0, 1, :, :8] = np.array(q_moduli, dtype=np.uint64)[:, None] - 1
    ctx.set_base_conversion(ctx.BASECONV_VALU)
    try:
        out = pkg.to_host(key.switch(pkg.to_device(ct)))
    finally:
        ctx.set_base_conversion(ctx.BASECONV_AUTO)
    bo, ao = O.hybrid_key_switch(O.RnsPlan(qp, n), L, blocks, ct[0, 0], ct[0, 1], kb, ka)
    assert (out[0, 0] == bo).all() and (out[0, 1] == ao).all()
