"""Autotuner: every NTT kernel variant and every key-switch engine it can
pick gives the oracle's residues bit-exactly, and a tuned context keeps
giving them."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("logn", [13, 14, 15, 16, 17])
@pytest.mark.parametrize("variant", [1, 2, 3])
def test_ntt_variant_bit_exact(pkg, ctx, logn, variant):
    n = 1 << logn
    moduli = [Q0[0], Q40[0], Q40[1], P60[0]] if logn < 17 else O.ntt_primes(60, 1 << 17, 2)
    rng = np.random.default_rng(10 * logn + variant)
    x = rand_residues(rng, moduli, n, polys=3)
    x[1, :, :4] = np.array(moduli, dtype=np.uint64)[:, None] - 1
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    orc = O.RnsPlan(moduli, n)
    saved = ctx.ntt_variant(logn)
    ctx.set_ntt_variant(logn, variant)
    try:
        out = pkg.empty(x.shape)
        plan.forward(pkg.to_device(x), out)
        got = pkg.to_host(out)
        back = pkg.to_host(plan.inverse(pkg.to_device(x)))
    finally:
        ctx.set_ntt_variant(logn, saved)
    for p in range(3):
        assert (got[p] == orc.forward(x[p])).all(), f"forward poly {p}"
        assert (back[p] == orc.inverse(x[p])).all(), f"inverse poly {p}"


def test_mixed_degree_batch_across_variants(pkg, ctx):
    """One batched call whose entries run different kernel variants."""
    rng = np.random.default_rng(5)
    moduli = [Q0[0], Q40[0], Q40[1]]
    entries, want = [], []
    for logn, var in ((13, 1), (14, 2), (15, 0), (16, 3)):
        ctx.set_ntt_variant(logn, var)
        n = 1 << logn
        plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
        x = rand_residues(rng, moduli, n, polys=2)
        orc = O.RnsPlan(moduli, n)
        entries.append((plan, pkg.to_device(x), pkg.empty(x.shape), None))
        want.append(np.stack([orc.forward(x[p]) for p in range(2)]))
    try:
        pkg.ntt_batch(ctx, entries, forward=True)
        for e, w in zip(entries, want):
            assert (pkg.to_host(e[2]) == w).all()
    finally:
        for logn in (13, 14, 15, 16):
            ctx.set_ntt_variant(logn, 0)


def test_autotune_ntt_then_bit_exact(pkg, ctx):
    n = 1 << 16
    moduli = [Q0[0]] + Q40[:5] + P60[:2]
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    best, ms = plan.autotune_ntt(4)
    assert 0 <= best < 4 and ms[best] > 0
    assert all(m > 0 for m in ms)  # all four variants apply at 2^16
    assert ctx.ntt_variant(16) == best
    rng = np.random.default_rng(6)
    x = rand_residues(rng, moduli, n, polys=2)
    got = pkg.to_host(plan.forward(pkg.to_device(x)))
    orc = O.RnsPlan(moduli, n)
    for p in range(2):
        assert (got[p] == orc.forward(x[p])).all()
    ctx.set_ntt_variant(16, 0)


def test_autotune_key_switch_engine(pkg, ctx):
    n, L, K, blocks = 1 << 13, 6, 3, [3, 3]
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(8)
    kb = rand_residues(rng, qp, n, len(blocks))
    ka = rand_residues(rng, qp, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    ct = rand_residues(rng, q_moduli, n, polys=4).reshape(2, 2, L, n)
    ctx.set_base_conversion(ctx.BASECONV_VALU)
    want = pkg.to_host(key.switch(pkg.to_device(ct)))
    ctx.set_base_conversion(ctx.BASECONV_AUTO)
    engine, fusion, times = key.autotune(2)
    assert engine in (ctx.BASECONV_VALU, ctx.BASECONV_MFMA)
    assert fusion in (ctx.KS_FUSION_OFF, ctx.KS_FUSION_XCD, ctx.KS_FUSION_LINEAR, ctx.KS_FUSION_INVERSE)
    assert len(times) == 8 and all(t > 0 for t in times.values())
    assert times[(engine, fusion)] == min(times.values())
    got = pkg.to_host(key.switch(pkg.to_device(ct)))
    assert (got == want).all()


@pytest.mark.parametrize("logn,variant", [(13, 0), (14, 1), (15, 2), (16, 3)])
def test_fused_ntt_mac_matches_unfused(pkg, ctx, logn, variant):
    """The key switch with the key MAC fused into the digits' forward row pass,
    or into the accumulators' first inverse pass, gives the residues of the
    separate NTT + MAC kernels, for every NTT variant."""
    n, L, K, blocks = 1 << logn, 5, 2, [3, 2]
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(logn)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, rand_residues(rng, qp, n, 2), rand_residues(rng, qp, n, 2))
    ct = pkg.to_device(rand_residues(rng, q_moduli, n, polys=6).reshape(3, 2, L, n))
    k = O.rotation_exponent_left(n, 3)
    ctx.set_ntt_variant(logn, variant)
    try:
        ctx.set_key_switch_fusion(ctx.KS_FUSION_OFF)
        want = pkg.to_host(key.apply_galois(ct, k))
        got = {}
        for mode in (ctx.KS_FUSION_XCD, ctx.KS_FUSION_LINEAR, ctx.KS_FUSION_INVERSE):
            ctx.set_key_switch_fusion(mode)
            got[mode] = pkg.to_host(key.apply_galois(ct, k))
    finally:
        ctx.set_key_switch_fusion(ctx.KS_FUSION_AUTO)
        ctx.set_ntt_variant(logn, 0)
    for mode, g in got.items():
        assert (g == want).all(), f"fusion mode {mode}"


def test_fused_ntt_mac_with_many_digits(pkg, ctx):
    """A bounded key with more than 8 digits (its FP64 sums are reduced
    every 8 digits) gives the unfused residues with the MAC fused into the
    forward row pass, and the unfused path matches the oracle."""
    n, L, K, base_log = 1 << 13, 5, 2, 16
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    cnt = O.bounded_digit_count(q_moduli, base_log)
    assert cnt > 8
    rng = np.random.default_rng(77)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    kb, ka = rand_residues(rng, qp, n, cnt), rand_residues(rng, qp, n, cnt)
    key = pkg.BoundedKeySwitchKey(plan, L, base_log, kb, ka)
    ct = pkg.to_device(rand_residues(rng, q_moduli, n, polys=4).reshape(2, 2, L, n))
    k = O.rotation_exponent_left(n, 5)
    try:
        ctx.set_key_switch_fusion(ctx.KS_FUSION_OFF)
        want = pkg.to_host(key.apply_galois(ct, k))
        for mode in (ctx.KS_FUSION_XCD, ctx.KS_FUSION_LINEAR):
            ctx.set_key_switch_fusion(mode)
            assert (pkg.to_host(key.apply_galois(ct, k)) == want).all(), f"fusion mode {mode}"
    finally:
        ctx.set_key_switch_fusion(ctx.KS_FUSION_AUTO)


@pytest.mark.parametrize("n_cts", [8, 9, 16])
def test_split_key_switch_matches_single_stream(pkg, ctx, n_cts):
    """Halves of the batch on two streams give the single-stream residues,
    for rotations and for relinearization (pass-through terms offset too)."""
    n, L, K, blocks = 1 << 13, 5, 2, [3, 2]
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(n_cts)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, rand_residues(rng, qp, n, 2), rand_residues(rng, qp, n, 2))
    ct = pkg.to_device(rand_residues(rng, q_moduli, n, polys=2 * n_cts).reshape(n_cts, 2, L, n))
    ct3 = pkg.to_device(rand_residues(rng, q_moduli, n, polys=3 * n_cts).reshape(n_cts, 3, L, n))
    k = O.rotation_exponent_left(n, 7)
    try:
        ctx.set_key_switch_split(ctx.KS_SPLIT_OFF)
        want = pkg.to_host(key.apply_galois(ct, k)), pkg.to_host(key.relinearize(ct3))
        ctx.set_key_switch_split(ctx.KS_SPLIT_ON)
        got = pkg.to_host(key.apply_galois(ct, k)), pkg.to_host(key.relinearize(ct3))
    finally:
        ctx.set_key_switch_split(ctx.KS_SPLIT_AUTO)
    assert (got[0] == want[0]).all() and (got[1] == want[1]).all()
