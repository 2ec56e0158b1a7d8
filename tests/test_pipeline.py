"""End-to-end CKKS pipeline of the small config: encrypt two slot vectors,
HMult + relinearize + rescale, rotate left by one slot (both key switches
hybrid with the one 60-bit special prime), decrypt, decode; the decrypted slots
must match the FP64 cleartext (z0 * z1 rotated) within the reference's own
multiply+rescale tolerance (1e-3 max-abs at N=4096, ckks/rns_multiply.rs:526).
The CPU test pins the oracle pipeline; the GPU test requires every residue
of the GPU pipeline to equal the oracle's and the same decrypted error."""
from __future__ import annotations

import numpy as np
import pytest

import ckks_pipeline as CP
from conftest import O

TOL = 1e-3


def oracle_pipeline(d):
    plan = d["plan"]
    ct0, ct1 = d["ct0"], d["ct1"]
    c0, c1, c2 = O.tensor(plan, ct0[0], ct0[1], ct1[0], ct1[1])
    kb, ka = d["rk"]
    L = len(plan.moduli)
    sb, sa = O.hybrid_key_switch(d["qp4"], L, [1] * L, np.zeros_like(c0), c2, kb, ka)
    rb = np.array([(c0[i].astype(object) + sb[i]) % q for i, q in enumerate(plan.moduli)], dtype=np.uint64)
    ra = np.array([(c1[i].astype(object) + sa[i]) % q for i, q in enumerate(plan.moduli)], dtype=np.uint64)
    mb, ma = O.rescale_ct(plan.moduli, rb, ra)
    q3 = plan.moduli[:3]
    k = d["k"]
    ab = np.stack([O.apply_automorphism(mb[i], q, k) for i, q in enumerate(q3)])
    aa = np.stack([O.apply_automorphism(ma[i], q, k) for i, q in enumerate(q3)])
    gb, ga = d["gk"]
    ob, oa = O.hybrid_key_switch(d["qp_plan"], 3, [1, 1, 1], ab, aa, gb, ga)
    return np.stack([mb, ma]), np.stack([ob, oa])


def oracle_pipeline_bounded(d):
    """The same pipeline with the reference's bounded algorithms:
    multiply_relinearize_rescale_rns_ckks_bounded_with_ntt
    (ckks/rns_multiply.rs, bounded relinearization of
    grafting/bounded_key_switch.rs:126-219) and the hybrid-QP rotation
    (ckks/hybrid_rns_galois_key.rs:550-780)."""
    plan = d["plan"]
    ct0, ct1 = d["ct0"], d["ct1"]
    c0, c1, c2 = O.tensor(plan, ct0[0], ct0[1], ct1[0], ct1[1])
    kb, ka = d["brk"]
    L = len(plan.moduli)
    sb, sa = O.bounded_key_switch(plan, L, CP.BOUNDED_RELIN_LOG, c0, c2, kb, ka)
    ra = np.array([(c1[i].astype(object) + sa[i]) % q for i, q in enumerate(plan.moduli)], dtype=np.uint64)
    mb, ma = O.rescale_ct(plan.moduli, sb, ra)
    q3 = plan.moduli[:3]
    k = d["k"]
    ab = np.stack([O.apply_automorphism(mb[i], q, k) for i, q in enumerate(q3)])
    aa = np.stack([O.apply_automorphism(ma[i], q, k) for i, q in enumerate(q3)])
    gb, ga = d["bgk"]
    ob, oa = O.bounded_key_switch(d["qp_plan"], 3, CP.BOUNDED_ROT_LOG, ab, aa, gb, ga)
    return np.stack([mb, ma]), np.stack([ob, oa])


@pytest.fixture(scope="module")
def data():
    return CP.setup(0)


def test_encode_decode_roundtrip(data):
    emb, plan = data["emb"], data["plan"]
    dec = CP.decrypt_slots(plan, data["ct0"], data["s"], CP.SCALE, emb)
    assert np.max(np.abs(dec - data["z0"])) < 1e-7


def test_embedding_matches_o_n2_definition():
    # ckks/embedding.rs:401-444: FFT embedding == direct evaluation
    import embedding as E
    n = 64
    emb = E.Embedding(n)
    c = np.random.default_rng(1).standard_normal(n)
    roots = np.exp(1j * np.pi * (2 * np.arange(n) + 1) / n)
    direct = np.array([np.polyval(c[::-1], r) for r in roots])
    assert np.allclose(emb.evaluate_all(c), direct, atol=1e-9)
    z = emb.coefficients_to_slots(c)
    assert np.allclose(emb.slots_to_coefficients(z), c, atol=1e-9)
    # sigma_5 rotates logical slots left by one (automorphism.rs:179-207)
    rot = O.apply_automorphism(np.arange(n, dtype=np.uint64), 2 ** 61 - 1, 5)


def test_oracle_pipeline_matches_fp64(data):
    mid, out = oracle_pipeline(data)
    plan = data["plan"]
    q3 = O.RnsPlan(plan.moduli[:3], CP.N)
    got = CP.decrypt_slots(q3, out, data["s"], CP.scale_after_rescale(), data["emb"])
    err = np.max(np.abs(got - CP.expected_slots(data)))
    assert err < TOL, err


@pytest.mark.gpu
def test_gpu_pipeline_bit_exact_and_within_tolerance(pkg, ctx, data):
    plan = data["plan"]
    L = len(plan.moduli)
    qp4 = pkg.PreparedRnsNttPlan(ctx, data["qp4"].moduli, CP.N)
    rkey = pkg.HybridKeySwitchKey(qp4, L, [1] * L, *data["rk"])
    qp = pkg.PreparedRnsNttPlan(ctx, data["qp_plan"].moduli, CP.N)
    gkey = pkg.HybridKeySwitchKey(qp, 3, [1, 1, 1], *data["gk"])
    d0 = pkg.to_device(data["ct0"][None])
    d1 = pkg.to_device(data["ct1"][None])
    mid = rkey.multiply_relinearize_rescale(d0, d1)
    out = pkg.to_host(gkey.apply_galois(mid, data["k"]))
    want_mid, want_out = oracle_pipeline(data)
    assert (pkg.to_host(mid)[0] == want_mid).all()
    assert (out[0] == want_out).all()
    q3 = O.RnsPlan(plan.moduli[:3], CP.N)
    got = CP.decrypt_slots(q3, out[0], data["s"], CP.scale_after_rescale(), data["emb"])
    err = np.max(np.abs(got - CP.expected_slots(data)))
    assert err < TOL, err


def test_oracle_bounded_pipeline_matches_fp64(data):
    _, out = oracle_pipeline_bounded(data)
    q3 = O.RnsPlan(data["plan"].moduli[:3], CP.N)
    got = CP.decrypt_slots(q3, out, data["s"], CP.scale_after_rescale(), data["emb"])
    err = np.max(np.abs(got - CP.expected_slots(data)))
    assert err < TOL, err


@pytest.mark.gpu
def test_gpu_bounded_pipeline_bit_exact_and_within_tolerance(pkg, ctx, data):
    plan = data["plan"]
    L = len(plan.moduli)
    qplan = pkg.PreparedRnsNttPlan(ctx, plan.moduli, CP.N)
    rkey = pkg.BoundedKeySwitchKey(qplan, L, CP.BOUNDED_RELIN_LOG, *data["brk"])
    qp = pkg.PreparedRnsNttPlan(ctx, data["qp_plan"].moduli, CP.N)
    gkey = pkg.BoundedKeySwitchKey(qp, 3, CP.BOUNDED_ROT_LOG, *data["bgk"])
    d0 = pkg.to_device(data["ct0"][None])
    d1 = pkg.to_device(data["ct1"][None])
    mid = rkey.multiply_relinearize_rescale(d0, d1)
    out = pkg.to_host(gkey.apply_galois(mid, data["k"]))
    want_mid, want_out = oracle_pipeline_bounded(data)
    assert (pkg.to_host(mid)[0] == want_mid).all()
    assert (out[0] == want_out).all()
    q3 = O.RnsPlan(plan.moduli[:3], CP.N)
    got = CP.decrypt_slots(q3, out[0], data["s"], CP.scale_after_rescale(), data["emb"])
    err = np.max(np.abs(got - CP.expected_slots(data)))
    assert err < TOL, err
