"""ckks_apply_galois_batch: rotations of ciphertext batches of uneven ring
degree, level and key kind in one launch sequence (one forward and one
inverse NTT launch for all of them) give, entry by entry, the residues of
ckks_apply_galois on that entry alone; one entry is recomputed by the
oracle."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def _dev_batch(moduli, n, n_cts, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = torch.empty((n_cts, 2, len(moduli), n), dtype=torch.int64, device="cuda")
    for i, q in enumerate(moduli):
        out[:, :, i, :] = torch.randint(0, int(q), (n_cts, 2, n), generator=g, device="cuda", dtype=torch.int64)
    return out


@pytest.mark.parametrize("fusion", ["off", "xcd", "linear"])
def test_mixed_degree_batch_matches_single_calls(pkg, ctx, fusion):
    """Also with the key MAC fused into the forward row pass of the entries
    whose degree takes two passes (the bounded and RNS-gadget keys stay
    unfused at their degrees or with more than 8 digits' registers)."""
    rng = np.random.default_rng(404)
    entries, keys, mats = [], [], []
    # the benchmark's pair: 64 ct at N = 2^16 (25 limbs, [9, 8, 8], K = 7) and
    # 16 ct at N = 2^15 (13 limbs, [5, 4, 4], K = 4); plus an RNS-gadget key
    # and a bounded hybrid-QP key at small degrees
    specs = [("hybrid", 16, 25, 7, [9, 8, 8], 64), ("hybrid", 15, 13, 4, [5, 4, 4], 16),
             ("rns", 12, 4, 0, [1, 1, 1, 1], 3), ("bounded", 13, 6, 2, 32, 2)]
    for idx, (kind, logn, L, K, blocks, n_cts) in enumerate(specs):
        n = 1 << logn
        q = [Q0[0]] + Q40[:L - 1]
        basis = q + P60[:K]
        plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
        if kind == "hybrid":
            kb, ka = rand_residues(rng, basis, n, len(blocks)), rand_residues(rng, basis, n, len(blocks))
            key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
            mats.append((kb, ka))
        elif kind == "rns":
            key = pkg.RnsKeySwitchKey(plan, blocks, rand_residues(rng, basis, n, len(blocks)),
                                      rand_residues(rng, basis, n, len(blocks)))
        else:
            cnt = O.bounded_digit_count(q, blocks)
            key = pkg.BoundedKeySwitchKey(plan, L, blocks, rand_residues(rng, basis, n, cnt),
                                          rand_residues(rng, basis, n, cnt))
        ct = _dev_batch(q, n, n_cts, seed=idx)
        k = O.rotation_exponent_left(n, 1 << idx)
        keys.append((key, k, ct, q, basis, blocks, kind))
        entries.append((key, k, ct, torch.empty_like(ct)))
    mode = {"off": ctx.KS_FUSION_OFF, "xcd": ctx.KS_FUSION_XCD, "linear": ctx.KS_FUSION_LINEAR}[fusion]
    try:
        ctx.set_key_switch_fusion(mode)
        pkg.apply_galois_batch(ctx, entries)
        torch.cuda.synchronize()
        ctx.set_key_switch_fusion(ctx.KS_FUSION_OFF)
        for (key, k, ct, out) in entries:
            assert torch.equal(out, key.apply_galois(ct, k)), "batched entry != single call"
    finally:
        ctx.set_key_switch_fusion(ctx.KS_FUSION_AUTO)
    # one ciphertext of the N = 2^15 entry against the oracle
    key, k, ct, q, basis, blocks, _ = keys[1]
    kb, ka = mats[1]
    c = pkg.to_host(ct[5])
    rb = np.stack([O.apply_automorphism(c[0, i], qi, k) for i, qi in enumerate(q)])
    ra = np.stack([O.apply_automorphism(c[1, i], qi, k) for i, qi in enumerate(q)])
    bo, ao = O.hybrid_key_switch(O.RnsPlan(basis, 1 << 15), len(q), blocks, rb, ra, kb, ka)
    got = pkg.to_host(entries[1][3][5])
    assert (got[0] == bo).all() and (got[1] == ao).all()


def test_batch_rejections(pkg, ctx):
    n = 1 << 12
    q = [Q0[0]] + Q40[:3]
    basis = q + P60[:1]
    rng = np.random.default_rng(1)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = pkg.HybridKeySwitchKey(plan, 4, [2, 2], rand_residues(rng, basis, n, 2), rand_residues(rng, basis, n, 2))
    ct = pkg.to_device(rand_residues(rng, q, n, 2).reshape(1, 2, 4, n))
    with pytest.raises(pkg.CkksError, match="odd"):
        pkg.apply_galois_batch(ctx, [(key, 4, ct, torch.empty_like(ct))])
    with pytest.raises(pkg.CkksError, match="alias"):
        pkg.apply_galois_batch(ctx, [(key, 5, ct, ct)])
    with pytest.raises(pkg.CkksError, match="16 entries"):
        pkg.apply_galois_batch(ctx, [(key, 5, ct, torch.empty_like(ct)) for _ in range(17)])
