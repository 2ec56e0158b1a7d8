"""Committed golden fixtures (tests/golden, written by make_golden.py in the
versioned wire format): the oracle must keep reproducing them (CPU), and the
GPU path must reproduce them residue for residue."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

from conftest import O, load_pkg

W = load_pkg().wire
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MAN = json.load(open(os.path.join(G, "golden.json")))


def blob(name):
    return W.load(os.path.join(G, name + ".ckrns"))


def closed_form_input(q, n):
    j = np.arange(n, dtype=object)
    return np.array([(v * v + 7 * v + 11) % q for v in j], dtype=np.uint64)


# ---------------------------------------------------------------- wire format

def test_wire_round_trip_and_rejections():
    rng = np.random.default_rng(1)
    q = [97, 193]
    data = np.stack([rng.integers(0, m, (3, 16), dtype=np.uint64) for m in q], axis=1)
    b = W.RnsBlob(q, data, W.KIND_CIPHERTEXT, W.DOMAIN_NTT, 2.0 ** 40)
    raw = W.to_bytes(b)
    back = W.from_bytes(raw)
    assert back.moduli == q and back.kind == W.KIND_CIPHERTEXT and back.domain == W.DOMAIN_NTT
    assert back.scale == 2.0 ** 40 and np.array_equal(back.data, data)
    with pytest.raises(W.WireError, match="not an RNS blob"):
        W.from_bytes(b"XXXXXXXX" + raw[8:])
    with pytest.raises(W.WireError, match="unsupported format version"):
        W.from_bytes(raw[:7] + b"\x02" + raw[8:])
    with pytest.raises(W.WireError, match="does not match the header"):
        W.from_bytes(raw[:-8])
    bad = data.copy()
    bad[0, 0, 0] = 97
    with pytest.raises(W.WireError, match="canonical"):
        W.to_bytes(W.RnsBlob(q, bad))


# ---------------------------------------------------------------- oracle vs fixtures

def test_oracle_reproduces_ntt_fixtures():
    for c in MAN["ntt"]:
        x = blob(c["in"]).data[0, 0]
        assert np.array_equal(O.Plan(c["q"], c["n"]).forward(x), blob(c["out"]).data[0, 0])
    for c in MAN["ntt_prefix"]:
        y = O.Plan(c["q"], c["n"]).forward(closed_form_input(c["q"], c["n"]))
        assert np.array_equal(y[:c["outputs"]], blob(c["file"]).data[0, 0])


def _oracle_keyswitch(c):
    kb, ka = blob(c["prefix"] + "_key_b").data, blob(c["prefix"] + "_key_a").data
    ct = blob(c["prefix"] + "_ct").data
    moduli, L, n = c["moduli"], c["q_limbs"], c["n"]
    plan = O.RnsPlan(moduli, n)
    b, a = ct[0], ct[1]
    if c["name"] == "rns":
        return O.rns_key_switch(plan, c["blocks"], b, a, kb, ka)
    if c["name"] == "dnum":
        return O.hybrid_key_switch(plan, L, c["blocks"], b, a, kb, ka)
    k = c.get("exponent")
    if k:
        b = np.stack([O.apply_automorphism(b[i], q, k) for i, q in enumerate(moduli[:L])])
        a = np.stack([O.apply_automorphism(a[i], q, k) for i, q in enumerate(moduli[:L])])
    return O.bounded_key_switch(plan, L, c["base_log"], b, a, kb, ka)


def test_oracle_reproduces_keyswitch_rescale_moddown_fixtures():
    for c in MAN["keyswitch"]:
        ob, oa = _oracle_keyswitch(c)
        want = blob(c["prefix"] + "_out").data
        assert np.array_equal(ob, want[0]) and np.array_equal(oa, want[1]), c["name"]
    c = MAN["rescale"][0]
    ct = blob(c["in"]).data
    mb, ma = O.rescale_ct(c["moduli"], ct[0], ct[1])
    assert np.array_equal(np.stack([mb, ma]), blob(c["out"]).data)
    c = MAN["moddown"][0]
    assert np.array_equal(O.mod_down(blob(c["in"]).data[0], c["q"], c["p"]), blob(c["out"]).data[0])


# ---------------------------------------------------------------- GPU vs fixtures

@pytest.mark.gpu
def test_gpu_reproduces_ntt_fixtures(pkg, ctx):
    for c in MAN["ntt"]:
        plan = pkg.PreparedRnsNttPlan(ctx, [c["q"]], c["n"])
        got = pkg.to_host(plan.forward(pkg.to_device(blob(c["in"]).data[0])))
        assert np.array_equal(got[0], blob(c["out"]).data[0, 0]), (c["q"], c["n"])
    for c in MAN["ntt_prefix"]:
        plan = pkg.PreparedRnsNttPlan(ctx, [c["q"]], c["n"])
        got = pkg.to_host(plan.forward(pkg.to_device(closed_form_input(c["q"], c["n"])[None])))
        assert np.array_equal(got[0, :c["outputs"]], blob(c["file"]).data[0, 0])


@pytest.mark.gpu
def test_gpu_reproduces_keyswitch_fixtures(pkg, ctx):
    for c in MAN["keyswitch"]:
        moduli, L, n = c["moduli"], c["q_limbs"], c["n"]
        plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
        kb, ka = blob(c["prefix"] + "_key_b").data, blob(c["prefix"] + "_key_a").data
        if c["name"] == "rns":
            key = pkg.RnsKeySwitchKey(plan, c["blocks"], kb, ka)
        elif c["name"] == "dnum":
            key = pkg.HybridKeySwitchKey(plan, L, c["blocks"], kb, ka)
        else:
            key = pkg.BoundedKeySwitchKey(plan, L, c["base_log"], kb, ka)
        ct = pkg.to_device(blob(c["prefix"] + "_ct").data[None])
        out = key.apply_galois(ct, c["exponent"]) if c.get("exponent") else key.switch(ct)
        assert np.array_equal(pkg.to_host(out)[0], blob(c["prefix"] + "_out").data), c["name"]


@pytest.mark.gpu
def test_gpu_reproduces_rescale_and_moddown_fixtures(pkg, ctx):
    c = MAN["rescale"][0]
    plan = pkg.PreparedRnsNttPlan(ctx, c["moduli"], c["n"])
    got = pkg.to_host(plan.rescale_to_next(pkg.to_device(blob(c["in"]).data[None]), len(c["moduli"])))
    assert np.array_equal(got[0], blob(c["out"]).data)
    c = MAN["moddown"][0]
    q, p = c["q"], c["p"]
    qp = pkg.PreparedRnsNttPlan(ctx, q + p, c["n"])
    x = blob(c["in"]).data
    got = pkg.to_host(qp.mod_down(pkg.to_device(np.concatenate([x, x, x])), len(q)))
    for i in range(3):  # an odd count of polynomials
        assert np.array_equal(got[i], blob(c["out"]).data[0])
