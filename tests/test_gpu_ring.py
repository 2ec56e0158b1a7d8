"""GPU parity of the ring kernels against the oracle: every residue must be
bit-exact.  Calls go through the C ABI (libckks_hip.so)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu

MODULI = [Q0[0]] + Q40 + P60  # 25 Q limbs + 7 special primes


def _plan(pkg, ctx, n, moduli=MODULI):
    return pkg.PreparedRnsNttPlan(ctx, moduli, n)


@pytest.mark.parametrize("logn", [10, 11, 12, 13, 14, 15, 16])
def test_ntt_forward_inverse_bit_exact(pkg, ctx, logn):
    n = 1 << logn
    moduli = MODULI[:6] + MODULI[-2:]
    rng = np.random.default_rng(100 + logn)
    x = rand_residues(rng, moduli, n, polys=3)
    x[0, :, 0] = 0
    x[1, :, :4] = np.array(moduli, dtype=np.uint64)[:, None] - 1
    plan = _plan(pkg, ctx, n, moduli)
    orc = O.RnsPlan(moduli, n)
    d = pkg.to_device(x)
    out = pkg.empty(x.shape)
    plan.forward(d, out)
    got = pkg.to_host(out)
    for p in range(3):
        assert (got[p] == orc.forward(x[p])).all(), f"forward mismatch poly {p}"
    back = pkg.empty(x.shape)
    plan.inverse(out, back)
    assert (pkg.to_host(back) == x).all()
    # inverse of arbitrary NTT-domain input
    y = rand_residues(rng, moduli, n, polys=2)
    dy = pkg.to_device(y)
    plan.inverse(dy)  # in place
    gy = pkg.to_host(dy)
    for p in range(2):
        assert (gy[p] == orc.inverse(y[p])).all()


def test_ntt_full_chain_at_2_16_and_levels(pkg, ctx):
    n = 1 << 16
    rng = np.random.default_rng(7)
    plan = _plan(pkg, ctx, n)
    assert plan.psi() == [O.find_negacyclic_root(q, n) for q in MODULI]
    x = rand_residues(rng, MODULI, n, polys=2)
    d = pkg.to_device(x)
    plan.forward(d)
    got = pkg.to_host(d)
    orc = O.RnsPlan(MODULI, n)
    for p in range(2):
        assert (got[p] == orc.forward(x[p])).all()
    # a lower level: only the first 13 limbs of each polynomial
    lvl = 13
    x2 = rand_residues(rng, MODULI[:lvl], n, polys=2)
    d2 = pkg.to_device(x2)
    plan.forward(d2, n_limbs=lvl)
    got2 = pkg.to_host(d2)
    orc2 = orc.prefix(lvl)
    for p in range(2):
        assert (got2[p] == orc2.forward(x2[p])).all()


def test_mixed_degree_batch_single_launch(pkg, ctx):
    """Config-2 shape in one launch per pass: N=2^16 at 25 limbs plus
    N=2^15 at 13 limbs (and a 2^12 entry)."""
    rng = np.random.default_rng(8)
    q = MODULI[:25]
    p16, p15, p12 = (_plan(pkg, ctx, 1 << k, q) for k in (16, 15, 12))
    x16 = rand_residues(rng, q, 1 << 16, polys=2)
    x15 = rand_residues(rng, q[:13], 1 << 15, polys=3)
    x12 = rand_residues(rng, q[:4], 1 << 12, polys=2)
    d16, d15, d12 = pkg.to_device(x16), pkg.to_device(x15), pkg.to_device(x12)
    o16, o15, o12 = pkg.empty(x16.shape), pkg.empty(x15.shape), pkg.empty(x12.shape)
    pkg.ntt_batch(ctx, [(p16, d16, o16, 25), (p15, d15, o15, 13), (p12, d12, o12, 4)], forward=True)
    for x, o, nn, L in ((x16, o16, 1 << 16, 25), (x15, o15, 1 << 15, 13), (x12, o12, 1 << 12, 4)):
        orc = O.RnsPlan(q[:L], nn)
        g = pkg.to_host(o)
        for p in range(x.shape[0]):
            assert (g[p] == orc.forward(x[p])).all()
    pkg.ntt_batch(ctx, [(p16, o16, None, 25), (p15, o15, None, 13), (p12, o12, None, 4)], forward=False)
    assert (pkg.to_host(o16) == x16).all()
    assert (pkg.to_host(o15) == x15).all()
    assert (pkg.to_host(o12) == x12).all()


@pytest.mark.parametrize("chunk_limbs", [1, 3, 7])
def test_chunked_two_pass_batches(pkg, ctx, chunk_limbs):
    """Inter-pass chunks that cut entries and limbs at arbitrary places
    (tiny chunks force every boundary case); results must not change."""
    rng = np.random.default_rng(20 + chunk_limbs)
    q = MODULI[:6]
    p16, p15, p13 = (_plan(pkg, ctx, 1 << k, q) for k in (16, 15, 13))
    xs = [rand_residues(rng, q[:5], 1 << 16, polys=2), rand_residues(rng, q[:3], 1 << 15, polys=3),
          rand_residues(rng, q[:6], 1 << 13, polys=2)]
    plans = [(p16, 5, 1 << 16), (p15, 3, 1 << 15), (p13, 6, 1 << 13)]
    ds = [pkg.to_device(x) for x in xs]
    outs = [pkg.empty(x.shape) for x in xs]
    try:
        ctx.set_ntt_chunk_bytes(chunk_limbs * (1 << 16) * 8)
        pkg.ntt_batch(ctx, [(p, d, o, L) for (p, L, _), d, o in zip(plans, ds, outs)], forward=True)
        got = [pkg.to_host(o) for o in outs]
        pkg.ntt_batch(ctx, [(p, o, None, L) for (p, L, _), o in zip(plans, outs)], forward=False)
        back = [pkg.to_host(o) for o in outs]
    finally:
        ctx.set_ntt_chunk_bytes(0)
    for x, g, b, (_, L, n) in zip(xs, got, back, plans):
        orc = O.RnsPlan(q[:L], n)
        for p in range(x.shape[0]):
            assert (g[p] == orc.forward(x[p])).all()
        assert (b == x).all()


@pytest.mark.parametrize("logn", [12, 16])
def test_pointwise_ops(pkg, ctx, logn):
    n = 1 << logn
    moduli = MODULI[:4] + MODULI[-1:]
    rng = np.random.default_rng(9)
    x = rand_residues(rng, moduli, n, polys=2)
    y = rand_residues(rng, moduli, n, polys=2)
    x[0, :, :3] = np.array(moduli, dtype=np.uint64)[:, None] - 1
    y[0, :, :3] = np.array(moduli, dtype=np.uint64)[:, None] - 1
    plan = _plan(pkg, ctx, n, moduli)
    dx, dy = pkg.to_device(x), pkg.to_device(y)
    mul = pkg.to_host(plan.pointwise_mul(dx, dy))
    add = pkg.to_host(plan.add(dx, dy))
    sub = pkg.to_host(plan.sub(dx, dy))
    neg = pkg.to_host(plan.neg(dx))
    acc = pkg.to_device(x)
    plan.pointwise_mul_add_assign(acc, dx, dy)
    mac = pkg.to_host(acc)
    for p in range(2):
        for i, q in enumerate(moduli):
            assert (mul[p, i] == O.pointwise_mul(x[p, i], y[p, i], q)).all()
            assert (mac[p, i] == O.pointwise_mac(x[p, i], x[p, i], y[p, i], q)).all()
            xo, yo = x[p, i].astype(object), y[p, i].astype(object)
            assert (add[p, i] == ((xo + yo) % q).astype(np.uint64)).all()
            assert (sub[p, i] == ((xo - yo) % q).astype(np.uint64)).all()
            assert (neg[p, i] == ((-xo) % q).astype(np.uint64)).all()


@pytest.mark.parametrize("logn", [12, 16])
def test_automorphisms_both_domains(pkg, ctx, logn):
    n = 1 << logn
    moduli = MODULI[:3]
    rng = np.random.default_rng(10)
    x = rand_residues(rng, moduli, n, polys=2)
    plan = _plan(pkg, ctx, n, moduli)
    dx = pkg.to_device(x)
    for k in [1, 3, 5, O.rotation_exponent_left(n, 1), O.rotation_exponent_left(n, 1 << 14 % (n // 2)),
              O.rotation_exponent_right(n, 7), O.conjugation_exponent(n), 2 * n + 5]:
        got = pkg.to_host(plan.apply_automorphism(dx, k))
        gnt = pkg.to_host(plan.permute_negacyclic_ntt_galois(dx, k))
        for p in range(2):
            for i, q in enumerate(moduli):
                assert (got[p, i] == O.apply_automorphism(x[p, i], q, k)).all()
                assert (gnt[p, i] == O.permute_ntt_galois(x[p, i], k)).all()
    with pytest.raises(pkg.CkksError, match="odd"):
        plan.apply_automorphism(dx, 4)


@pytest.mark.parametrize("logn,limbs", [(12, 4), (16, 25)])
def test_rescale_and_mod_switch(pkg, ctx, logn, limbs):
    n = 1 << logn
    moduli = MODULI[:limbs]
    rng = np.random.default_rng(11)
    ct = rand_residues(rng, moduli, n, polys=2 * 3).reshape(3, 2, limbs, n)
    # boundary residues of the dropped limb: 0, q/2, q/2+1, q-1
    ql = moduli[-1]
    ct[0, :, -1, :4] = [0, ql // 2, ql // 2 + 1, ql - 1]
    plan = _plan(pkg, ctx, n, moduli)
    d = pkg.to_device(ct)
    out = pkg.to_host(plan.rescale_to_next(d, limbs))
    for c in range(3):
        bo, ao = O.rescale_ct(moduli, ct[c, 0], ct[c, 1])
        assert (out[c, 0] == bo).all() and (out[c, 1] == ao).all()
    ms = pkg.to_host(plan.mod_switch_to_next(d, limbs))
    assert (ms == ct[:, :, :-1, :]).all()
    with pytest.raises(pkg.CkksError, match="at least two"):
        plan.rescale_to_next(d, 1)


@pytest.mark.parametrize("logn,limbs", [(12, 4), (16, 5)])
def test_tensor(pkg, ctx, logn, limbs):
    n = 1 << logn
    moduli = MODULI[:limbs]
    rng = np.random.default_rng(12)
    ct0 = rand_residues(rng, moduli, n, polys=4).reshape(2, 2, limbs, n)
    ct1 = rand_residues(rng, moduli, n, polys=4).reshape(2, 2, limbs, n)
    plan = _plan(pkg, ctx, n, moduli)
    out = pkg.to_host(plan.tensor(pkg.to_device(ct0), pkg.to_device(ct1), limbs))
    orc = O.RnsPlan(moduli, n)
    for c in range(2):
        c0, c1, c2 = O.tensor(orc, ct0[c, 0], ct0[c, 1], ct1[c, 0], ct1[c, 1])
        assert (out[c, 0] == c0).all() and (out[c, 1] == c1).all() and (out[c, 2] == c2).all()


@pytest.mark.parametrize("logn", [10, 12, 16])
def test_monomials_and_scalars_bit_exact(pkg, ctx, logn):
    """RnsNttPolynomial::mul_monomial_signed (ring/rns_ntt.rs:215), the
    coefficient-domain Polynomial::mul_monomial_signed and scalar_mul
    (ring/polynomial.rs:144, 164) against the oracle, on 40-bit (FP64 table)
    and 60-bit primes."""
    n = 1 << logn
    moduli = MODULI[:3] + MODULI[-2:]
    rng = np.random.default_rng(300 + logn)
    x = rand_residues(rng, moduli, n, polys=2)
    x[0, :, :4] = np.array(moduli, dtype=np.uint64)[:, None] - 1
    plan = _plan(pkg, ctx, n, moduli)
    orc = O.RnsPlan(moduli, n)
    d = pkg.to_device(x)
    for e in [0, 1, -1, n - 1, n, n + 1, 2 * n - 1, 2 * n + 3, 12345, -54321]:
        got = pkg.to_host(plan.ntt_mul_monomial_signed(d, e))
        for p in range(2):
            want = np.stack([O.ntt_mul_monomial(orc.plans[i], x[p, i], e) for i in range(len(moduli))])
            assert (got[p] == want).all(), f"NTT monomial e={e}"
        if e >= 0:
            for sign in (1, -1):
                got = pkg.to_host(plan.mul_monomial_signed(d, e, sign))
                for p in range(2):
                    want = np.stack([O.mul_monomial(x[p, i], q, e, sign) for i, q in enumerate(moduli)])
                    assert (got[p] == want).all(), f"monomial e={e} sign={sign}"
    # in place in the NTT domain
    dd = d.clone()
    plan.ntt_mul_monomial_signed(dd, 7, out=dd)
    assert (pkg.to_host(dd) == pkg.to_host(plan.ntt_mul_monomial_signed(d, 7))).all()
    scalars = [0, moduli[1] + 5, (1 << 64) - 1, 0x9E3779B97F4A7C15, moduli[4] - 1]
    got = pkg.to_host(plan.scalar_mul(d, scalars))
    for p in range(2):
        want = np.stack([O.scalar_mul(x[p, i], q, scalars[i]) for i, q in enumerate(moduli)])
        assert (got[p] == want).all()
    with pytest.raises(pkg.CkksError, match="sign"):
        plan.mul_monomial_signed(d, 3, 2)
    with pytest.raises(pkg.CkksError, match="alias"):
        plan.mul_monomial_signed(d, 3, 1, out=d)
