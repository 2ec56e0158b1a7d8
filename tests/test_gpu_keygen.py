"""Key generation on the device (ckks_keygen) from host-sampled randomness,
bit-exact against the oracle's encryption of the same gadget multiples with
the same a and e, for every key kind:
- RNS gadget (RnsKeySwitchKey / RnsMultiplicationKey::generate_with_
  distribution_ntt_rng, grafting/rns_key_switch.rs:1155-1258, 1430-1512);
- the dnum-block hybrid key (entry j encrypts P * E_j * source);
- bounded with and without P (PreparedHybridRnsGaloisKey::generate_with_
  distribution_ntt_rng, ckks/hybrid_rns_galois_key.rs:61-170;
  BoundedRnsMultiplicationKey, grafting/bounded_key_switch.rs:22-118);
each entry is encrypt_rns_raw (grafting/rns_key_switch.rs:245-383) then the
forward NTT, as the reference's prepared keys hold them."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import ckks_pipeline as CP
from conftest import O, P60, Q0, Q40

pytestmark = pytest.mark.gpu

N = 1 << 12
Q = [Q0[0]] + Q40[:3]

CASES = {
    "rns": dict(basis=Q, q_limbs=4, blocks=[1, 1, 1, 1], w=0),
    "rns_blocks": dict(basis=Q, q_limbs=4, blocks=[2, 2], w=0),
    "hybrid": dict(basis=Q + P60[:2], q_limbs=4, blocks=[2, 2], w=0),
    "bounded_qp": dict(basis=Q + P60[:1], q_limbs=4, blocks=None, w=61),
    "bounded_relin": dict(basis=Q, q_limbs=4, blocks=None, w=20),
}


def _factors(case):
    basis, ql = case["basis"], case["q_limbs"]
    Pv = math.prod(basis[ql:])
    if case["blocks"] is not None:
        out, start = [], 0
        for sz in case["blocks"]:
            out.append([Pv % q if start <= i < start + sz else 0 for i, q in enumerate(basis)])
            start += sz
        return out
    count = O.bounded_digit_count(basis[:ql], case["w"])
    return [[Pv * (1 << (case["w"] * j)) % q for q in basis] for j in range(count)]


@pytest.mark.parametrize("source_kind", ["galois", "square"])
@pytest.mark.parametrize("name", list(CASES))
def test_keygen_matches_oracle(pkg, ctx, name, source_kind):
    case = CASES[name]
    basis, ql = case["basis"], case["q_limbs"]
    rng = O.Rng(31 + len(name))
    s = rng.ternary(N)
    orc = O.RnsPlan(basis, N)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, N)
    fac = _factors(case)
    entries = len(fac)
    a = np.stack([np.stack([rng.uniform(q, N) for q in basis]) for _ in range(entries)])
    e = np.stack([rng.gaussian(3.2, N) for _ in range(entries)])
    k = O.rotation_exponent_left(N, 3)
    if source_kind == "galois":
        src = np.stack([CP.ternary_to_mod(CP.sigma_ternary(s, k, N), q) for q in basis])
    else:
        src = CP.secret_square(orc, s)
    want_b = np.empty((entries, len(basis), N), dtype=np.uint64)
    want_a = np.empty_like(want_b)
    for j in range(entries):
        msg = np.stack([(src[i].astype(object) * fac[j][i] % q).astype(np.uint64) for i, q in enumerate(basis)])
        b = O.encrypt(orc, msg, e[j], s, a[j])
        want_b[j], want_a[j] = orc.forward(b), orc.forward(a[j])
    ds = torch.from_numpy(s.astype(np.int8)).cuda()
    dsrc = pkg.galois_source(plan, ds, k) if source_kind == "galois" else pkg.square_source(plan, ds)
    assert (pkg.to_host(dsrc) == src).all(), "device source != oracle source"
    kind = (pkg.KEYGEN_BOUNDED if case["blocks"] is None else
            pkg.KEYGEN_HYBRID if ql < len(basis) else pkg.KEYGEN_RNS_GADGET)
    kb, ka = pkg.keygen(plan, kind, ql, dsrc, ds, pkg.to_device(a), torch.from_numpy(e).cuda(),
                        block_sizes=case["blocks"] or (), base_log=case["w"])
    assert (pkg.to_host(ka) == want_a).all()
    assert (pkg.to_host(kb) == want_b).all()


def test_generated_key_equals_host_key(pkg, ctx):
    """A key built from device-generated entries switches exactly as one built
    from the same entries copied through host memory."""
    case = CASES["hybrid"]
    basis, ql, blocks = case["basis"], case["q_limbs"], case["blocks"]
    rng = O.Rng(77)
    s = rng.ternary(N)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, N)
    ds = torch.from_numpy(s.astype(np.int8)).cuda()
    k = O.rotation_exponent_left(N, 1)
    a = pkg.to_device(np.stack([np.stack([rng.uniform(q, N) for q in basis]) for _ in blocks]))
    e = torch.from_numpy(np.stack([rng.gaussian(3.2, N) for _ in blocks])).cuda()
    key = pkg.HybridKeySwitchKey.generate(plan, ql, blocks, pkg.galois_source(plan, ds, k), ds, a, e)
    kb, ka = pkg.keygen(plan, pkg.KEYGEN_HYBRID, ql, pkg.galois_source(plan, ds, k), ds, a, e, block_sizes=blocks)
    host_key = pkg.HybridKeySwitchKey(plan, ql, blocks, pkg.to_host(kb), pkg.to_host(ka))
    ct = pkg.to_device(np.stack([np.stack([rng.uniform(q, N) for q in Q]) for _ in range(2)])[None])
    assert (pkg.to_host(key.apply_galois(ct, k)) == pkg.to_host(host_key.apply_galois(ct, k))).all()


def test_keygen_rejections(pkg, ctx):
    """The reference's keygen assertions: a ternary, nonzero secret; canonical
    uniform residues; blocks covering the basis."""
    basis = Q + P60[:1]
    plan = pkg.PreparedRnsNttPlan(ctx, basis, N)
    rng = O.Rng(5)
    s = torch.from_numpy(rng.ternary(N).astype(np.int8)).cuda()
    src = pkg.project_secret(plan, s)
    a = pkg.to_device(np.stack([np.stack([rng.uniform(q, N) for q in basis]) for _ in range(2)]))
    e = torch.zeros((2, N), dtype=torch.int64, device="cuda")
    bad = s.clone()
    bad[7] = 2
    with pytest.raises(pkg.CkksError, match="ternary"):
        pkg.keygen(plan, pkg.KEYGEN_HYBRID, 4, src, bad, a, e, block_sizes=[2, 2])
    with pytest.raises(pkg.CkksError, match="nonzero"):
        pkg.keygen(plan, pkg.KEYGEN_HYBRID, 4, src, torch.zeros_like(s), a, e, block_sizes=[2, 2])
    a_bad = a.clone()
    a_bad[1, 2, 5] = int(basis[2])
    with pytest.raises(pkg.CkksError, match="canonical"):
        pkg.keygen(plan, pkg.KEYGEN_HYBRID, 4, src, s, a_bad, e, block_sizes=[2, 2])
    with pytest.raises(pkg.CkksError, match="cover"):
        pkg.keygen(plan, pkg.KEYGEN_HYBRID, 4, src, s, a, e, block_sizes=[2, 1])
    with pytest.raises(pkg.CkksError, match="special primes"):
        pkg.keygen(plan, pkg.KEYGEN_RNS_GADGET, 4, src, s, a, e, block_sizes=[2, 2])
    # the context still works afterwards
    kb, ka = pkg.keygen(plan, pkg.KEYGEN_HYBRID, 4, src, s, a, e, block_sizes=[2, 2])
    torch.cuda.synchronize()


def test_key_bytes_and_storage_report(pkg, ctx):
    """ckks_ksk_bytes: 2 x entries x limbs x N x 8 device bytes, growing by a
    permuted copy per exponent cached for hoisted rotations; the level-map
    report adds them up."""
    CK = pkg.ckks
    case = CASES["bounded_qp"]
    basis, ql, w = case["basis"], case["q_limbs"], case["w"]
    rng = np.random.default_rng(3)
    plan = pkg.PreparedRnsNttPlan(ctx, basis, N)
    count = pkg.bounded_digit_count(plan, ql, w)
    key = pkg.BoundedKeySwitchKey(plan, ql, w, *[np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64)
                                                                        for q in basis]) for _ in range(count)])
                                                 for _ in range(2)])
    base = 2 * count * len(basis) * N * 8
    assert key.nbytes == base
    ct = pkg.to_device(np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in Q]) for _ in range(2)])[None])
    h = key.hoist(ct)
    key.rotate_hoisted(h, ct, O.rotation_exponent_left(N, 1))
    key.rotate_hoisted(h, ct, O.rotation_exponent_left(N, 2))
    h.close()
    assert key.nbytes == 3 * base
    keys = CK.RnsCkksEvaluationKeys()
    lvl = CK.RnsCkksLevelKeys(0, Q)
    lvl.insert_galois_key(5, key)
    keys.insert_level(lvl)
    rep = CK.evaluation_key_storage(keys)
    assert rep.galois_bytes == rep.total_resident_bytes == 3 * base
