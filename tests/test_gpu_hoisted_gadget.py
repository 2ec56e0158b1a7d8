"""Hoisted rotations with dnum-hybrid and RNS-gadget keys (ckks_hoist /
ckks_rotate_hoisted): digits of the unrotated a cut and transformed once,
each rotation a key MAC over the key permuted by k^-1.  Bit-exact against
the oracle's composition of the same steps -- canonical digits of a
(basis_transition.rs lift per block), sigma_k of every digit limb, NTT, key
MAC, inverse NTT, ModDown by P, sigma_k(b) added -- at toy sizes and at the
N = 2^16, 25-limb, [9, 8, 8], K = 7 shape; and, with device-generated keys,
a hoisted rotation decrypts to the rotated slots like the canonical one."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def hoisted_switch_oracle(qp_moduli, q_limbs, blocks, b, a, k, kb, ka):
    n = a.shape[-1]
    q = qp_moduli[:q_limbs]
    p = qp_moduli[q_limbs:]
    orc = O.RnsPlan(qp_moduli, n)
    acc_b = [np.zeros(n, dtype=np.uint64) for _ in qp_moduli]
    acc_a = [np.zeros(n, dtype=np.uint64) for _ in qp_moduli]
    start = 0
    for j, sz in enumerate(blocks):
        blk = list(range(start, start + sz))
        others = [i for i in range(len(qp_moduli)) if i not in blk]
        ext = O.extend_basis(a[start:start + sz], [qp_moduli[i] for i in blk], [qp_moduli[i] for i in others])
        digit = np.empty((len(qp_moduli), n), dtype=np.uint64)
        digit[blk] = a[start:start + sz]
        digit[others] = ext
        rot = np.stack([O.apply_automorphism(digit[i], qi, k) for i, qi in enumerate(qp_moduli)])
        f = orc.forward(rot)
        for i, qi in enumerate(qp_moduli):
            acc_b[i] = O.pointwise_mac(acc_b[i], f[i], kb[j, i], qi)
            acc_a[i] = O.pointwise_mac(acc_a[i], f[i], ka[j, i], qi)
        start += sz
    cb, ca = orc.inverse(np.stack(acc_b)), orc.inverse(np.stack(acc_a))
    if p:
        cb, ca = O.mod_down(cb, q, p), O.mod_down(ca, q, p)
    rb = np.stack([O.apply_automorphism(b[i], qi, k) for i, qi in enumerate(q)])
    bo = np.array([(cb[i].astype(object) + rb[i]) % qi for i, qi in enumerate(q)], dtype=np.uint64)
    return bo, ca


@pytest.mark.parametrize("kind,logn,L,K,blocks", [("hybrid", 12, 4, 2, [2, 2]), ("hybrid", 13, 9, 4, [5, 4]),
                                                  ("rns", 12, 4, 0, [1, 1, 1, 1]), ("rns", 12, 4, 0, [2, 2])])
def test_hoisted_rotation_matches_oracle_composition(pkg, ctx, kind, logn, L, K, blocks):
    n = 1 << logn
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    rng = np.random.default_rng(logn * 7 + L + K)
    kb, ka = rand_residues(rng, basis, n, len(blocks)), rand_residues(rng, basis, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = (pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka) if kind == "hybrid"
           else pkg.RnsKeySwitchKey(plan, blocks, kb, ka))
    ct = rand_residues(rng, q, n, polys=4).reshape(2, 2, L, n)
    ct[0, 1, :, :8] = np.array(q, dtype=np.uint64)[:, None] - 1
    dct = pkg.to_device(ct)
    h = key.hoist(dct)
    try:
        for k in (O.rotation_exponent_left(n, 1), O.rotation_exponent_left(n, 5), 2 * n - 1):
            got = pkg.to_host(key.rotate_hoisted(h, dct, k))
            for c in range(2):
                bo, ao = hoisted_switch_oracle(basis, L, blocks, ct[c, 0], ct[c, 1], k, kb, ka)
                assert (got[c, 0] == bo).all() and (got[c, 1] == ao).all(), (k, c)
    finally:
        h.close()


def test_hoisted_rotation_at_the_benchmark_shape(pkg, ctx):
    n, L, K, blocks = 1 << 16, 25, 7, [9, 8, 8]
    q = [Q0[0]] + Q40[:L - 1]
    basis = q + P60[:K]
    rng = np.random.default_rng(1616)
    kb, ka = rand_residues(rng, basis, n, 3), rand_residues(rng, basis, n, 3)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    ct = rand_residues(rng, q, n, polys=4).reshape(2, 2, L, n)
    dct = pkg.to_device(ct)
    h = key.hoist(dct)
    k = O.rotation_exponent_left(n, 1 << 7)
    got = pkg.to_host(key.rotate_hoisted(h, dct, k))
    h.close()
    bo, ao = hoisted_switch_oracle(basis, L, blocks, ct[1, 0], ct[1, 1], k, kb, ka)
    assert (got[1, 0] == bo).all() and (got[1, 1] == ao).all()


def test_hoisted_rotation_decrypts_like_the_canonical_one(pkg, ctx):
    """With keys generated on the device, the hoisted rotation and the
    canonical one both decrypt to the rotated slots (FP64, 1e-6 of the
    scale); the hoisted variant is a key switch of sigma_k(a) of the same
    noise size."""
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
        want = np.roll(z, -steps)
        for out in (hoisted, canonical):
            err = np.max(np.abs(wl.decrypt(out[0], L, wl.scale) - want))
            assert err < 1e-6, (steps, err)
        assert not torch.equal(hoisted, canonical)  # different digits, different residues
