"""Hoisted rotations with dnum-hybrid and RNS-gadget keys (ckks_hoist /
ckks_hoist_with_b / ckks_rotate_hoisted): the canonical digits of the
unrotated a are cut and transformed once; each rotation is a key MAC over
the key permuted by k^-1
plus the per-exponent correction of the canonical lift of negated
coefficients, so every residue equals the canonical rotation
(ckks_apply_galois, itself checked against the oracle here): at toy sizes,
for every key kind and for conjugation, at the N = 2^16, 25-limb,
[9, 8, 8], K = 7 shape, and for batches holding a zero block value (which
take the canonical path)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def canonical_rotation_oracle(qp_moduli, q_limbs, blocks, ct, k, kb, ka):
    """rotate_left via the canonical key switch: sigma_k of both components,
    canonical digits of sigma_k(a), MAC, inverse NTT, ModDown by P."""
    n = ct.shape[-1]
    q = qp_moduli[:q_limbs]
    rb = np.stack([O.apply_automorphism(ct[0, i], qi, k) for i, qi in enumerate(q)])
    ra = np.stack([O.apply_automorphism(ct[1, i], qi, k) for i, qi in enumerate(q)])
    if len(qp_moduli) > q_limbs:
        return O.hybrid_key_switch(O.RnsPlan(qp_moduli, n), q_limbs, blocks, rb, ra, kb, ka)
    return O.rns_key_switch(O.RnsPlan(qp_moduli, n), blocks, rb, ra, kb, ka)


@pytest.mark.parametrize("with_b", [False, True])
@pytest.mark.parametrize("kind,logn,L,K,blocks", [("hybrid", 12, 4, 2, [2, 2]), ("hybrid", 13, 9, 4, [5, 4]),
                                                  ("rns", 12, 4, 0, [1, 1, 1, 1]), ("rns", 12, 4, 0, [2, 2])])
def test_hoisted_rotation_equals_canonical(pkg, ctx, kind, logn, L, K, blocks, with_b):
    n = 1 << logn
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    rng = np.random.default_rng(logn * 7 + L + K)
    kb, ka = rand_residues(rng, basis, n, len(blocks)), rand_residues(rng, basis, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = (pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka) if kind == "hybrid"
           else pkg.RnsKeySwitchKey(plan, blocks, kb, ka))
    ct = rand_residues(rng, q, n, polys=6).reshape(3, 2, L, n)
    ct[0, 1, :, :8] = np.array(q, dtype=np.uint64)[:, None] - 1
    ct[1, 0, :, :8] = np.array(q, dtype=np.uint64)[:, None] - 1
    dct = pkg.to_device(ct)
    h = key.hoist(dct, with_b=with_b)
    try:
        for k in (O.rotation_exponent_left(n, 1), O.rotation_exponent_left(n, 5), 2 * n - 1,
                  O.rotation_exponent_left(n, n // 4 - 1)):
            ctx.reset_execution_profile()
            got = key.rotate_hoisted(h, dct, k)
            # the hoisted path itself: no decomposition, no forward NTT
            assert ctx.execution_profile().ntt_forwards == 0
            assert torch.equal(got, key.apply_galois(dct, k)), k
        k = O.rotation_exponent_left(n, 3)
        got = pkg.to_host(key.rotate_hoisted(h, dct, k))
        bo, ao = canonical_rotation_oracle(basis, L, blocks, ct[2], k, kb, ka)
        assert (got[2, 0] == bo).all() and (got[2, 1] == ao).all()
    finally:
        h.close()


@pytest.mark.parametrize("kind", ["hybrid", "rns"])
def test_zero_block_values_take_the_canonical_path(pkg, ctx, kind):
    """A coefficient whose every residue in a gadget block is zero has a
    canonical digit that does not follow the negation; such a batch is
    rotated canonically, so the residues still equal ckks_apply_galois for
    every exponent (including ones that negate that coefficient)."""
    n, L = 1 << 12, 4
    blocks, K = ([2, 2], 2) if kind == "hybrid" else ([1, 1, 1, 1], 0)
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    rng = np.random.default_rng(91)
    kb, ka = rand_residues(rng, basis, n, len(blocks)), rand_residues(rng, basis, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = (pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka) if kind == "hybrid"
           else pkg.RnsKeySwitchKey(plan, blocks, kb, ka))
    ct = rand_residues(rng, q, n, polys=4).reshape(2, 2, L, n)
    ct[1, 1, 0:2, 17] = 0      # block 0 of a at coefficient 17 is zero
    ct[1, 1, 0:2, n - 3] = 0   # and at a coefficient near the top
    dct = pkg.to_device(ct)
    h = key.hoist(dct)
    try:
        for r in (1, 2, 7, 100):
            k = O.rotation_exponent_left(n, r)
            ctx.reset_execution_profile()
            got = key.rotate_hoisted(h, dct, k)
            assert ctx.execution_profile().ntt_forwards == 2 * len(blocks)  # the canonical path ran
            assert torch.equal(got, key.apply_galois(dct, k)), r
    finally:
        h.close()


@pytest.mark.parametrize("with_b", [False, True])
def test_hoisted_rotation_at_the_benchmark_shape(pkg, ctx, with_b):
    n, L, K, blocks = 1 << 16, 25, 7, [9, 8, 8]
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    rng = np.random.default_rng(1616)
    kb, ka = rand_residues(rng, basis, n, 3), rand_residues(rng, basis, n, 3)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    ct = rand_residues(rng, q, n, polys=4).reshape(2, 2, L, n)
    dct = pkg.to_device(ct)
    h = key.hoist(dct, with_b=with_b)
    try:
        for i in (0, 7, 14):
            k = O.rotation_exponent_left(n, 1 << i)
            assert torch.equal(key.rotate_hoisted(h, dct, k), key.apply_galois(dct, k)), i
        k = O.rotation_exponent_left(n, 1 << 7)
        got = pkg.to_host(key.rotate_hoisted(h, dct, k))
    finally:
        h.close()
    bo, ao = canonical_rotation_oracle(basis, L, blocks, ct[1], k, kb, ka)
    assert (got[1, 0] == bo).all() and (got[1, 1] == ao).all()


def test_hoisted_rotation_decrypts_to_the_rotated_slots(pkg, ctx):
    """With keys generated on the device, the hoisted rotation is the
    canonical one, residue for residue, and decrypts to the rotated slots
    (FP64, 1e-6 of the scale)."""
    from workloads import Workload
    wl = Workload(pkg, ctx, logn=13, L=6, K=3, dnum=3, seed=77)
    L = len(wl.Q)
    rng = np.random.default_rng(5)
    z = rng.uniform(-1, 1, wl.slots) + 1j * rng.uniform(-1, 1, wl.slots)
    ct = wl.encrypt(z)[None].contiguous()
    for steps in (1, 3, 100):
        key, k = wl.key_for(L, steps)
        h = key.hoist(ct)
        hoisted = key.rotate_hoisted(h, ct, k)
        canonical = key.apply_galois(ct, k)
        h.close()
        assert torch.equal(hoisted, canonical), steps
        err = np.max(np.abs(wl.decrypt(hoisted[0], L, wl.scale) - np.roll(z, -steps)))
        assert err < 1e-6, (steps, err)


@pytest.mark.parametrize("with_b", [False, True])
def test_hoisted_rotation_batch_of_uneven_degrees(pkg, ctx, with_b):
    """ckks_rotate_hoisted_batch: the benchmark's two shapes (N = 2^16, 25
    limbs, [9, 8, 8], K = 7 and N = 2^15, 13 limbs, [5, 4, 4], K = 4), an
    RNS-gadget key and a bounded key in one call; every entry equals
    apply_galois, and its inverse NTTs run as one launch per pass."""
    rng = np.random.default_rng(2024)
    specs = [("hybrid", 16, 25, 7, [9, 8, 8], 3), ("hybrid", 15, 13, 4, [5, 4, 4], 2),
             ("rns", 12, 4, 0, [2, 1, 1], 2), ("bounded", 13, 6, 2, 32, 2)]
    entries, hoisted = [], []
    for idx, (kind, logn, L, K, blocks, n_cts) in enumerate(specs):
        n = 1 << logn
        q = [Q0[0]] + Q40[:L - 1]
        basis = q + P60[:K]
        plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
        if kind == "hybrid":
            key = pkg.HybridKeySwitchKey(plan, L, blocks, rand_residues(rng, basis, n, len(blocks)),
                                         rand_residues(rng, basis, n, len(blocks)))
        elif kind == "rns":
            key = pkg.RnsKeySwitchKey(plan, blocks, rand_residues(rng, basis, n, len(blocks)),
                                      rand_residues(rng, basis, n, len(blocks)))
        else:
            cnt = O.bounded_digit_count(q, blocks)
            key = pkg.BoundedKeySwitchKey(plan, L, blocks, rand_residues(rng, basis, n, cnt),
                                          rand_residues(rng, basis, n, cnt))
        ct = pkg.to_device(rand_residues(rng, q, n, polys=2 * n_cts).reshape(n_cts, 2, L, n))
        h = key.hoist(ct, with_b=True) if with_b and kind != "bounded" else key.hoist(ct)
        hoisted.append(h)
        entries.append((h, key, O.rotation_exponent_left(n, 3 + idx), ct, torch.empty_like(ct)))
    try:
        ctx.reset_execution_profile()
        pkg.rotate_hoisted_batch(ctx, entries)
        prof = ctx.execution_profile()
        assert prof.ntt_forwards == 0  # no decomposition of canonical digits ran again
        for h, key, k, ct, out in entries:
            assert torch.equal(out, key.apply_galois(ct, k)), k
    finally:
        for h in hoisted:
            h.close()


@pytest.mark.parametrize("with_b", [False, True])
def test_hoisted_batch_of_exponents_shares_the_digits(pkg, ctx, with_b):
    """Several exponents of one hoisted batch in one ckks_rotate_hoisted_batch
    call (their MACs read each digit once, up to 4 exponents per launch):
    5 exponents of an N = 2^15 batch, interleaved with 2 of an N = 2^13
    batch and 3 of an RNS-gadget batch (no special primes: P = 1), each
    equal to apply_galois."""
    rng = np.random.default_rng(515)
    shapes = [("hybrid", 15, 13, 4, [5, 4, 4], 2, 5), ("hybrid", 13, 9, 4, [5, 4], 3, 2), ("rns", 12, 4, 0, [2, 2], 2, 3)]
    entries, hoisted = [], []
    per = []
    for kind, logn, L, K, blocks, n_cts, n_exp in shapes:
        n = 1 << logn
        q = [Q0[0]] + Q40[:L - 1]
        basis = q + P60[:K]
        plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
        make = pkg.HybridKeySwitchKey if kind == "hybrid" else (lambda p, L_, b, kb, ka: pkg.RnsKeySwitchKey(p, b, kb, ka))
        ct = pkg.to_device(rand_residues(rng, q, n, polys=2 * n_cts).reshape(n_cts, 2, L, n))
        keys = [make(plan, L, blocks, rand_residues(rng, basis, n, len(blocks)),
                     rand_residues(rng, basis, n, len(blocks))) for _ in range(n_exp)]
        h = keys[0].hoist(ct, with_b=with_b)
        hoisted.append(h)
        exps = [O.rotation_exponent_left(n, 1 << e) for e in range(n_exp - 1)] + [2 * n - 1]
        per.append([(h, key, k, ct, torch.empty_like(ct)) for key, k in zip(keys, exps)])
    # interleave the entries of the three batches
    for i in range(max(len(p) for p in per)):
        for p in per:
            if i < len(p):
                entries.append(p[i])
    try:
        ctx.reset_execution_profile()
        pkg.rotate_hoisted_batch(ctx, entries)
        assert ctx.execution_profile().ntt_forwards == 0
        for h, key, k, ct, out in entries:
            assert torch.equal(out, key.apply_galois(ct, k)), k
    finally:
        for h in hoisted:
            h.close()


def test_bounded_hoisted_batch_shares_the_digits(pkg, ctx):
    """Bounded (base 2^w) digits of one hoisted batch rotated by five
    exponents in one ckks_rotate_hoisted_batch call (more than 8 digits: one
    chunked MAC per exponent, one inverse NTT launch for all); a batch
    holding a -2^(w-1) tie takes the canonical path for the exponents that
    negate it.  Every entry equals apply_galois."""
    rng = np.random.default_rng(2323)
    n, L, K, w = 1 << 13, 6, 2, 32
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    cnt = O.bounded_digit_count(q, w)
    keys = [pkg.BoundedKeySwitchKey(plan, L, w, rand_residues(rng, basis, n, cnt), rand_residues(rng, basis, n, cnt))
            for _ in range(5)]
    ct = rand_residues(rng, q, n, polys=6).reshape(3, 2, L, n)
    batches = [ct]
    tie = ct.copy()
    # coefficient 7 of ciphertext 1: a = -2^(w-1), whose balanced digit vector holds -2^(w-1)
    tie[1, 1, :, 7] = [(-(1 << (w - 1))) % qi for qi in q]
    batches.append(tie)
    for b in batches:
        d = pkg.to_device(b)
        h = keys[0].hoist(d)
        try:
            exps = [O.rotation_exponent_left(n, 1 << e) for e in range(4)] + [2 * n - 1]
            entries = [(h, key, k, d, torch.empty_like(d)) for key, k in zip(keys, exps)]
            pkg.rotate_hoisted_batch(ctx, entries)
            for _, key, k, _, out in entries:
                assert torch.equal(out, key.apply_galois(d, k)), k
        finally:
            h.close()


def test_hoisted_mac_inside_the_inverse_pass(pkg, ctx):
    """With KS_FUSION_INVERSE, hoisted rotations of two-pass degrees compute
    the key MAC (and the canonical-lift correction) inside the Galois-read
    first inverse pass: single calls and a batch of the benchmark's two
    shapes give the residues of apply_galois."""
    rng = np.random.default_rng(99)
    entries, hoisted = [], []
    for logn, L, K, blocks, n_cts in ((16, 25, 7, [9, 8, 8], 2), (15, 13, 4, [5, 4, 4], 3), (13, 9, 4, [5, 4], 2)):
        n = 1 << logn
        q = [Q0[0]] + Q40[:L - 1]
        basis = q + P60[:K]
        plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
        key = pkg.HybridKeySwitchKey(plan, L, blocks, rand_residues(rng, basis, n, len(blocks)),
                                     rand_residues(rng, basis, n, len(blocks)))
        ct = pkg.to_device(rand_residues(rng, q, n, polys=2 * n_cts).reshape(n_cts, 2, L, n))
        hoisted.append(key.hoist(ct))
        entries.append((hoisted[-1], key, O.rotation_exponent_left(n, 1 << (logn - 4)), ct, torch.empty_like(ct)))
    try:
        want = [key.apply_galois(ct, k) for h, key, k, ct, out in entries]
        ctx.set_key_switch_fusion(ctx.KS_FUSION_INVERSE)
        for (h, key, k, ct, out), w in zip(entries, want):
            assert torch.equal(key.rotate_hoisted(h, ct, k), w), k
        pkg.rotate_hoisted_batch(ctx, entries[:2])
        for (h, key, k, ct, out), w in zip(entries[:2], want):
            assert torch.equal(out, w), k
    finally:
        ctx.set_key_switch_fusion(ctx.KS_FUSION_AUTO)
        for h in hoisted:
            h.close()


@pytest.mark.parametrize("with_b", [False, True])
def test_rehoist_replaces_the_digits_in_place(pkg, ctx, with_b):
    """ckks_rehoist: a second batch hoisted into the first one's buffers
    rotates to apply_galois of the second batch; a batch with a zero block
    value re-hoisted after a clean one takes the canonical path, and a clean
    one after it the hoisted path again."""
    n, L, K, blocks = 1 << 13, 9, 4, [5, 4]
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    rng = np.random.default_rng(4242)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, rand_residues(rng, basis, n, 2), rand_residues(rng, basis, n, 2))
    batches = [rand_residues(rng, q, n, polys=6).reshape(3, 2, L, n) for _ in range(3)]
    batches[1][2, 1, 5:9, 100] = 0  # block 1 of a zero at coefficient 100
    k = O.rotation_exponent_left(n, 9)
    d0 = pkg.to_device(batches[0])
    h = key.hoist(d0, with_b=with_b)
    try:
        for i, b in enumerate(batches):
            d = pkg.to_device(b)
            if i:
                h.rehoist(key, d)
            ctx.reset_execution_profile()
            got = key.rotate_hoisted(h, d, k)
            assert (ctx.execution_profile().ntt_forwards > 0) == (i == 1), i
            assert torch.equal(got, key.apply_galois(d, k)), i
        with pytest.raises(ValueError, match="as many"):
            h.rehoist(key, pkg.to_device(batches[0][:2]))
    finally:
        h.close()
