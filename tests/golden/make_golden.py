"""Writes the golden fixtures of this directory (run from the repository
root: python tests/golden/make_golden.py).  Every output comes from the
oracle, whose functions are pinned against the reference's known answers in
tests/test_oracle*.py; the fixtures freeze them in the versioned wire format
(ccmm-rs_amd/wire.py) so both the oracle and the GPU path are checked
against the same committed bytes.  Cases:
  ntt_*      forward NTT vectors per config prime at N = 2^12 and 2^13, and
             the first 256 outputs at N = 2^16 of a closed-form input
  ks_*       one toy key switch per family at N = 2^10 (RNS gadget, dnum
             hybrid with P, bounded relinearization, hybrid-QP rotation)
  rescale_*  rescale at the centering boundaries of the dropped limb
  moddown_*  mod-down of signed boundary values (auxiliary_modulus.rs tests)
"""
from __future__ import annotations

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from conftest import O, P60, Q0, Q40, load_pkg, rand_residues  # noqa: E402

W = load_pkg().wire


def save(name, moduli, data, kind=0, domain=0):
    W.save(os.path.join(HERE, name + ".ckrns"), W.RnsBlob(list(moduli), np.asarray(data, dtype=np.uint64),
                                                          kind, domain))
    return name


def closed_form_input(q, n):
    j = np.arange(n, dtype=object)
    return np.array([(v * v + 7 * v + 11) % q for v in j], dtype=np.uint64)


def main():
    manifest = {"ntt": [], "ntt_prefix": [], "keyswitch": [], "rescale": [], "moddown": []}
    primes = [Q0[0], Q40[0], Q40[1], P60[0], P60[1]]
    rng = np.random.default_rng(20260101)
    for n in (1 << 12, 1 << 13):
        for q in primes:
            x = rng.integers(0, q, n, dtype=np.uint64)
            y = O.Plan(q, n).forward(x)
            tag = f"ntt_{q}_{n}"
            save(tag + "_in", [q], x[None, None])
            save(tag + "_out", [q], y[None, None], domain=1)
            manifest["ntt"].append({"q": q, "n": n, "in": tag + "_in", "out": tag + "_out"})
    for q in (Q0[0], Q40[0]):
        n = 1 << 16
        y = O.Plan(q, n).forward(closed_form_input(q, n))
        tag = f"nttprefix_{q}_{n}"
        save(tag, [q], y[:256].reshape(1, 1, 256), domain=1)
        manifest["ntt_prefix"].append({"q": q, "n": n, "outputs": 256, "file": tag,
                                       "input": "x_j = (j^2 + 7 j + 11) mod q"})
    # key switches, N = 1024
    n = 1024
    q = [Q0[0]] + Q40[:3]
    cases = [
        ("rns", q, 0, {"blocks": [1, 1, 1, 1]}),
        ("dnum", q + P60[:2], 4, {"blocks": [2, 2]}),
        ("bounded_relin", q, 0, {"base_log": 20}),
        ("bounded_qp", q + P60[:1], 4, {"base_log": 61, "exponent": 5}),
    ]
    for name, moduli, q_limbs, params in cases:
        L = q_limbs or len(moduli)
        plan = O.RnsPlan(moduli, n)
        if "blocks" in params:
            digits = len(params["blocks"])
        else:
            digits = O.bounded_digit_count(moduli[:L], params["base_log"])
        kb = rand_residues(rng, moduli, n, digits)
        ka = rand_residues(rng, moduli, n, digits)
        ct = rand_residues(rng, moduli[:L], n, polys=2)
        b, a = ct[0], ct[1]
        if name == "rns":
            ob, oa = O.rns_key_switch(plan, params["blocks"], b, a, kb, ka)
        elif name == "dnum":
            ob, oa = O.hybrid_key_switch(plan, L, params["blocks"], b, a, kb, ka)
        else:
            k = params.get("exponent")
            if k:
                b = np.stack([O.apply_automorphism(b[i], qi, k) for i, qi in enumerate(moduli[:L])])
                a = np.stack([O.apply_automorphism(a[i], qi, k) for i, qi in enumerate(moduli[:L])])
            ob, oa = O.bounded_key_switch(plan, L, params["base_log"], b, a, kb, ka)
        tag = f"ks_{name}"
        save(tag + "_key_b", moduli, kb, kind=2, domain=1)
        save(tag + "_key_a", moduli, ka, kind=2, domain=1)
        save(tag + "_ct", moduli[:L], ct.reshape(2, L, n), kind=1)
        save(tag + "_out", moduli[:L], np.stack([ob, oa]), kind=1)
        manifest["keyswitch"].append(dict(name=name, n=n, q_limbs=L, moduli=list(moduli), prefix=tag, **params))
    # rescale boundaries of the dropped limb (rns_rescale.rs:133-150)
    ql = q[-1]
    ct = rand_residues(rng, q, n, polys=2)
    bvals = [0, 1, ql // 2 - 1, ql // 2, ql // 2 + 1, ql - 1]
    for c, v in enumerate(bvals):
        ct[:, -1, c] = v
    mb, ma = O.rescale_ct(q, ct[0], ct[1])
    save("rescale_in", q, ct, kind=1)
    save("rescale_out", q[:-1], np.stack([mb, ma]), kind=1)
    manifest["rescale"].append({"n": n, "moduli": q, "in": "rescale_in", "out": "rescale_out",
                                "boundary_values_of_last_limb": bvals})
    # mod-down of signed boundaries (auxiliary_modulus.rs:210-242)
    qp = q + P60[:2]
    Pv = math.prod(P60[:2])
    QP = math.prod(qp)
    signed = []
    for t in (1, 128, 129, 257, 386, 1000):
        for s in (1, -1):
            signed += [s * (t * Pv // 256), s * (Pv // 2 + t), s * (Pv // 2 - t)]
    x = rand_residues(rng, qp, n)[0]
    for c, v in enumerate(signed):
        x[:, c] = [v % m for m in qp]
    out = O.mod_down(x, q, P60[:2]).reshape(len(q), n)
    save("moddown_in", qp, x[None])
    save("moddown_out", q, out[None])
    manifest["moddown"].append({"n": n, "q": q, "p": P60[:2], "in": "moddown_in", "out": "moddown_out",
                                "signed_values": [str(v) for v in signed]})
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
