"""Level-aware evaluation layer: chain states (ckks/chain.rs tests), the
evaluation-key map (ckks/evaluation_keys.rs) and evaluator dispatch
(ckks/evaluator.rs), with the reference's panic messages.  The CPU tests
cover the bookkeeping without touching a GPU; the GPU test runs an
evaluator sequence and checks every residue against the oracle."""
from __future__ import annotations

import types

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, load_pkg, rand_residues

CK = load_pkg().ckks
SMALL = [12_289, 40_961, 65_537]


def test_chain_levels_and_states():
    # chain.rs:130-190
    chain = CK.ModulusChain.from_top_basis(SMALL)
    assert chain.max_level() == 2 and chain.level(1) == (12_289, 40_961)
    st = CK.CkksChainState.top(chain, 65_537.0)
    assert st.level == 0 and st.basis == chain.level(0)
    assert st.rescale_divisor(chain) == 65_537
    nxt = st.after_rescale(chain)
    assert (nxt.level, nxt.basis, nxt.scale) == (1, chain.level(1), 1.0)
    prod = CK.CkksChainState.new(chain, 1, 40_961.0).after_multiply(CK.CkksChainState.new(chain, 1, 2.0), chain)
    assert prod.scale == 81_922.0
    last = CK.CkksChainState.new(chain, 2, 1.0)
    assert not last.can_rescale(chain)
    with pytest.raises(CK.CkksAssertion, match="cannot rescale the final CKKS chain level"):
        last.after_rescale(chain)
    with pytest.raises(CK.CkksAssertion, match="CKKS level must exist in the modulus chain"):
        CK.CkksChainState.new(chain, 3, 1.0)
    with pytest.raises(CK.CkksAssertion, match="CKKS scale must be positive and finite"):
        CK.CkksChainState.new(chain, 0, 0.0)
    with pytest.raises(CK.CkksAssertion, match="CKKS multiplication requires matching levels"):
        st.after_multiply(nxt, chain)


def test_ciphertext_must_match_its_state():
    import torch

    chain = CK.ModulusChain(SMALL)
    st = CK.CkksChainState.new(chain, 1, 1.0)
    ok = CK.RnsCkksCiphertext(torch.zeros(2 * 2 * 8, dtype=torch.int64), st, chain, 8)
    assert ok.level == 1 and ok.basis == chain.level(1)
    with pytest.raises(CK.CkksAssertion, match="basis must match CKKS chain state"):
        CK.RnsCkksCiphertext(torch.zeros(2 * 3 * 8, dtype=torch.int64), st, chain, 8)


def _fake_key(moduli, q_limbs):
    return types.SimpleNamespace(plan=types.SimpleNamespace(moduli=list(moduli)), q_limbs=q_limbs)


def test_evaluation_key_map_lookup_and_errors():
    chain = CK.ModulusChain(SMALL)
    keys = CK.RnsCkksEvaluationKeys()
    lvl0 = CK.RnsCkksLevelKeys(0, chain.level(0))
    mk = _fake_key(SMALL + [P60[0]], 3)
    lvl0.set_multiplication_key(mk)
    lvl0.insert_galois_key(5, _fake_key(SMALL, 3))
    with pytest.raises(CK.CkksAssertion, match="RNS Galois key exponent already exists at this level"):
        lvl0.insert_galois_key(5, _fake_key(SMALL, 3))
    with pytest.raises(CK.CkksAssertion, match="RNS multiplication key basis must match level-key basis"):
        lvl0.set_multiplication_key(_fake_key(SMALL[:2], 2))
    keys.insert_level(lvl0)
    st0 = CK.CkksChainState.top(chain, 1.0)
    assert keys.multiplication_for(st0) is mk
    with pytest.raises(CK.CkksAssertion, match="Galois key is missing for requested exponent at active CKKS level"):
        keys.galois_for(st0, 25)
    with pytest.raises(CK.CkksAssertion, match="evaluation keys are missing for active CKKS level"):
        keys.for_state(CK.CkksChainState.new(chain, 1, 1.0))
    lvl1 = CK.RnsCkksLevelKeys(1, chain.level(1))
    keys.insert_level(lvl1)
    with pytest.raises(CK.CkksAssertion, match="multiplication key is missing for active CKKS level"):
        keys.multiplication_for(CK.CkksChainState.new(chain, 1, 1.0))


def test_exponent_helpers_match_oracle():
    for n in (16, 4096, 1 << 16):
        for r in (1, 2, 7, 100):
            assert CK.rotation_exponent_left(n, r) == O.rotation_exponent_left(n, r)
            assert CK.rotation_exponent_right(n, r) == O.rotation_exponent_right(n, r)
        assert CK.conjugation_exponent(n) == O.conjugation_exponent(n)


def test_trace_counts():
    tr = CK.ExecutionTrace()
    tr.record(CK.ExecutionEvent(CK.ExecutionEventKind.ROTATE, ring_degree=8))
    tr.record(CK.ExecutionEvent(CK.ExecutionEventKind.RESCALE))
    tr.record(CK.ExecutionEvent(CK.ExecutionEventKind.ROTATE))
    assert len(tr) == 3 and tr.count(CK.ExecutionEventKind.ROTATE) == 2 and not tr.is_empty()


def test_execution_profile_counts_and_total():
    """ExecutionProfile::{count, total_operations} (execution/mod.rs:176-230):
    eBLAS / CCMM kinds count 0, the total sums the ten counted kinds (not the
    two device-only fields), and a trace's events give the same counts."""
    K = CK.ExecutionEventKind
    p = CK.ExecutionProfile(ciphertext_plaintext_multiplies=2, rotations=3, ntt_forwards=5, rescales=1,
                            key_switches=3, ntt_limb_transforms=40)
    assert p.count(K.ROTATE) == 3 and p.count(K.NTT_FORWARD) == 5 and p.count(K.CIPHERTEXT_PLAINTEXT_MULTIPLY) == 2
    for kind in (K.EBLAS_DOT, K.EBLAS_GEMV, K.EBLAS_GEMM, K.CPMM, K.CCMM_SCALAR, K.CCMM_STRUCTURED):
        assert p.count(kind) == 0
    assert p.total_operations() == 11
    tr = CK.ExecutionTrace()
    for kind in (K.ROTATE, K.ROTATE, K.RESCALE, K.EBLAS_GEMM, K.NTT_INVERSE):
        tr.record(CK.ExecutionEvent(kind))
    q = CK.ExecutionProfile.from_trace(tr)
    assert q.rotations == 2 and q.rescales == 1 and q.ntt_inverses == 1 and q.total_operations() == 4
    assert (p - CK.ExecutionProfile(rotations=1)).rotations == 2


@pytest.mark.gpu
def test_gpu_evaluator_sequence_matches_oracle(pkg, ctx):
    """add -> multiply (relin + rescale) -> rotate_left(3) -> conjugate ->
    mod_switch, every key from the level map, every residue against the
    oracle's composition of the same steps."""
    n, L, K = 1 << 12, 4, 1
    q = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(77)
    chain = CK.ModulusChain(q)
    plan = pkg.PreparedRnsNttPlan(ctx, q, n)
    trace = CK.ExecutionTrace()
    keys = CK.RnsCkksEvaluationKeys()
    # level 0: multiplication key (dnum hybrid with P); level 1: Galois keys
    qp0 = pkg.PreparedRnsNttPlan(ctx, q + P60[:K], n)
    mkb, mka = rand_residues(rng, q + P60[:K], n, L), rand_residues(rng, q + P60[:K], n, L)
    lvl0 = CK.RnsCkksLevelKeys(0, q)
    lvl0.set_multiplication_key(pkg.HybridKeySwitchKey(qp0, L, [1] * L, mkb, mka))
    keys.insert_level(lvl0)
    q1 = q[:L - 1]
    qp1 = pkg.PreparedRnsNttPlan(ctx, q1 + P60[:K], n)
    lvl1 = CK.RnsCkksLevelKeys(1, q1)
    gk = {}
    for e in (CK.rotation_exponent_left(n, 3), CK.conjugation_exponent(n)):
        kb, ka = rand_residues(rng, q1 + P60[:K], n, 3), rand_residues(rng, q1 + P60[:K], n, 3)
        gk[e] = (kb, ka)
        lvl1.insert_galois_key(e, pkg.HybridKeySwitchKey(qp1, L - 1, [1] * (L - 1), kb, ka))
    keys.insert_level(lvl1)
    ev = CK.RnsCkksEvaluator(chain, keys, plan, trace)
    a = rand_residues(rng, q, n, polys=2)
    b = rand_residues(rng, q, n, polys=2)
    x = ev.encrypt_state(pkg.to_device(a), 2.0 ** 40)
    y = ev.encrypt_state(pkg.to_device(b), 2.0 ** 40)
    s = ev.add(x, y)
    m = ev.multiply(s, y)
    assert m.level == 1 and m.scale == 2.0 ** 80 / q[-1]
    r = ev.rotate_left(m, 3)
    c = ev.conjugate(r)
    z = ev.mod_switch(c)
    assert z.level == 2 and z.scale == m.scale
    with pytest.raises(CK.CkksAssertion, match="RNS CKKS addition requires matching levels"):
        ev.add(x, m)
    # oracle composition
    orc = O.RnsPlan(q, n)
    sa = np.stack([np.stack([(a[p, i].astype(object) + b[p, i]) % qi for i, qi in enumerate(q)]) for p in range(2)]).astype(np.uint64)
    c0, c1, c2 = O.tensor(orc, sa[0], sa[1], b[0], b[1])
    kb_, ka_ = O.hybrid_key_switch(O.RnsPlan(q + P60[:K], n), L, [1] * L, np.zeros_like(c0), c2, mkb, mka)
    rb = np.stack([(c0[i].astype(object) + kb_[i]) % qi for i, qi in enumerate(q)]).astype(np.uint64)
    ra = np.stack([(c1[i].astype(object) + ka_[i]) % qi for i, qi in enumerate(q)]).astype(np.uint64)
    mb, ma = O.rescale_ct(q, rb, ra)
    assert (pkg.to_host(m.data) == np.stack([mb, ma])).all()
    cur = np.stack([mb, ma])
    for e in (CK.rotation_exponent_left(n, 3), CK.conjugation_exponent(n)):
        gb = np.stack([O.apply_automorphism(cur[0, i], qi, e) for i, qi in enumerate(q1)])
        ga = np.stack([O.apply_automorphism(cur[1, i], qi, e) for i, qi in enumerate(q1)])
        ob, oa = O.hybrid_key_switch(O.RnsPlan(q1 + P60[:K], n), L - 1, [1] * (L - 1), gb, ga, *gk[e])
        cur = np.stack([ob, oa])
    assert (pkg.to_host(c.data) == cur).all()
    assert (pkg.to_host(z.data) == cur[:, :L - 2]).all()
    kinds = [ev_.kind for ev_ in trace.events()]
    assert kinds.count(CK.ExecutionEventKind.ROTATE) == 1 and kinds.count(CK.ExecutionEventKind.MOD_SWITCH) == 1


def test_evaluation_key_storage_report():
    """evaluation_key_storage (ckks/key_accounting.rs:276-344) over the level
    map: per-level multiplication and Galois bytes, totals, and the report
    text layout."""
    class FakeKey:
        def __init__(self, moduli, q_limbs, nbytes):
            self.plan = type("P", (), {"moduli": moduli})()
            self.q_limbs = q_limbs
            self.nbytes = nbytes

    q = [11, 13, 17]
    keys = CK.RnsCkksEvaluationKeys()
    l0 = CK.RnsCkksLevelKeys(0, q)
    l0.set_multiplication_key(FakeKey(q + [19], 3, 1000))
    l0.insert_galois_key(5, FakeKey(q + [19], 3, 300))
    l0.insert_galois_key(25, FakeKey(q + [19], 3, 300))
    l1 = CK.RnsCkksLevelKeys(1, q[:2])
    l1.insert_galois_key(5, FakeKey(q[:2] + [19], 2, 200))
    keys.insert_level(l1)
    keys.insert_level(l0)
    rep = CK.evaluation_key_storage(keys)
    assert [lv.level for lv in rep.levels] == [0, 1]
    assert (rep.levels[0].multiplication_bytes, rep.levels[0].galois_key_count, rep.levels[0].galois_bytes,
            rep.levels[0].total_resident_bytes) == (1000, 2, 600, 1600)
    assert (rep.multiplication_bytes, rep.galois_key_count, rep.galois_bytes, rep.total_resident_bytes) == \
        (1000, 3, 800, 1800)
    text = rep.to_report_string()
    assert text.startswith("CKKS_EVALUATION_KEY_STORAGE\nlevel=0 multiplication_bytes=1000 galois_key_count=2 ")
    assert text.endswith("total_resident_bytes=1800\n")
