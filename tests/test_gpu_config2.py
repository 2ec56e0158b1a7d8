"""Bit-exact parity at the batched-core-primitives shape itself: N = 2^16,
Q = one 60-bit base prime + 24 40-bit primes (25 limbs), dnum = 3 hybrid
key blocks [9, 8, 8] with 7 60-bit special primes, a 64-ciphertext batch;
and the 16-ciphertext batch at N = 2^15 (13 limbs, blocks [5, 4, 4], 4
special primes) that runs beside it.

Each GPU call runs on the whole batch; the oracle recomputes a few of the
ciphertexts (first, middle, last) from their own residues:
- rotations by 2^i slots, k = 5^(2^i) mod 2N (ckks/rns_galois_key.rs:730-751;
  the dnum-block hybrid switch composes rns_gadget.rs, basis_transition.rs
  and auxiliary_modulus.rs, as ckks/hybrid_rns_galois_key.rs:550-780 does for
  bounded digits);
- multiply + relinearize + rescale (ckks/rns_multiply.rs:161-215)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu

SHAPES = {
    16: dict(L=25, K=7, blocks=[9, 8, 8], n_cts=64, check=(0, 31, 63)),
    15: dict(L=13, K=4, blocks=[5, 4, 4], n_cts=16, check=(0, 15)),
}


def _device_batch(torch, moduli, n, n_cts, seed):
    """[n_cts][2][L][n] canonical residues drawn on the device (the 64-ct
    batch is 1.7 GB); extreme residues q-1 and 0 in ciphertext 0."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    out = torch.empty((n_cts, 2, len(moduli), n), dtype=torch.int64, device="cuda")
    for i, q in enumerate(moduli):
        out[:, :, i, :] = torch.randint(0, int(q), (n_cts, 2, n), generator=g, device="cuda", dtype=torch.int64)
        out[0, :, i, :8] = int(q) - 1
        out[0, :, i, 8:16] = 0
    return out


def _setup(pkg, ctx, logn, seed):
    sh = SHAPES[logn]
    n, L, K, blocks = 1 << logn, sh["L"], sh["K"], sh["blocks"]
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(seed)
    kb, ka = rand_residues(rng, qp, n, len(blocks)), rand_residues(rng, qp, n, len(blocks))
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    key = pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)
    return sh, n, L, blocks, q_moduli, qp, kb, ka, key


def _rotate_oracle(orc, L, blocks, q_moduli, ct, k, kb, ka):
    rb = np.stack([O.apply_automorphism(ct[0, i], q, k) for i, q in enumerate(q_moduli)])
    ra = np.stack([O.apply_automorphism(ct[1, i], q, k) for i, q in enumerate(q_moduli)])
    return O.hybrid_key_switch(orc, L, blocks, rb, ra, kb, ka)


@pytest.mark.parametrize("logn,i_list", [(16, (0, 7, 14)), (15, (0, 13))])
def test_batched_rotations_by_powers_of_two(pkg, ctx, logn, i_list):
    import torch

    sh, n, L, blocks, q_moduli, qp, kb, ka, key = _setup(pkg, ctx, logn, 1000 + logn)
    dct = _device_batch(torch, q_moduli, n, sh["n_cts"], seed=logn)
    orc = O.RnsPlan(qp, n)
    out = torch.empty_like(dct)
    for j, i in enumerate(i_list):
        k = O.rotation_exponent_left(n, 1 << i)
        assert k == pow(5, 1 << i, 2 * n)
        key.apply_galois(dct, k, out=out)
        torch.cuda.synchronize()
        # every ciphertext of the batch for the first exponent, the last
        # ciphertext for the others
        for c in (sh["check"] if j == 0 else sh["check"][-1:]):
            ct = pkg.to_host(dct[c])
            got = pkg.to_host(out[c])
            bo, ao = _rotate_oracle(orc, L, blocks, q_moduli, ct, k, kb, ka)
            assert (got[0] == bo).all() and (got[1] == ao).all(), f"rotation 2^{i}, ct {c}"


def test_batched_multiply_relinearize_rescale_n65536(pkg, ctx):
    import torch

    sh, n, L, blocks, q_moduli, qp, kb, ka, key = _setup(pkg, ctx, 16, 2016)
    n_cts = 4
    d0 = _device_batch(torch, q_moduli, n, n_cts, seed=1)
    d1 = _device_batch(torch, q_moduli, n, n_cts, seed=2)
    out = pkg.to_host(key.multiply_relinearize_rescale(d0, d1))
    assert out.shape == (n_cts, 2, L - 1, n)
    orc_qp, orc_q = O.RnsPlan(qp, n), O.RnsPlan(q_moduli, n)
    for c in (0, n_cts - 1):
        c0_, c1_ = pkg.to_host(d0[c]), pkg.to_host(d1[c])
        t0, t1, t2 = O.tensor(orc_q, c0_[0], c0_[1], c1_[0], c1_[1])
        sb, sa = O.hybrid_key_switch(orc_qp, L, blocks, np.zeros_like(t0), t2, kb, ka)
        rb = np.array([(t0[i].astype(object) + sb[i]) % q for i, q in enumerate(q_moduli)], dtype=np.uint64)
        ra = np.array([(t1[i].astype(object) + sa[i]) % q for i, q in enumerate(q_moduli)], dtype=np.uint64)
        bo, ao = O.rescale_ct(q_moduli, rb, ra)
        assert (out[c, 0] == bo).all() and (out[c, 1] == ao).all(), f"ct {c}"
