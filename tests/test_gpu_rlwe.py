"""GPU raw RLWE encryption / decryption (grafting/rns_key_switch.rs:245-383,
906-947) from host-sampled randomness against the oracle, residue for
residue, and an encrypt -> decrypt round trip through the device only."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("logn,L", [(12, 4), (13, 9), (16, 25)])
def test_encrypt_decrypt_bit_exact(pkg, ctx, logn, L):
    n = 1 << logn
    moduli = [Q0[0]] + Q40[:L - 1]
    rng = O.Rng(logn)
    s = rng.ternary(n)
    m = rand_residues(np.random.default_rng(logn), moduli, n, polys=3)
    e = np.stack([rng.gaussian(3.2, n) for _ in range(3)])
    e[0, :4] = [np.iinfo(np.int64).min + 1, np.iinfo(np.int64).max, -1, 0]
    a = np.stack([np.stack([rng.uniform(q, n) for q in moduli]) for _ in range(3)])
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    ds = torch.from_numpy(s.astype(np.int8)).cuda()
    de = torch.from_numpy(e.astype(np.int64)).cuda()
    da = pkg.to_device(a)
    b = plan.encrypt_raw(pkg.to_device(m), de, ds, da)
    got = pkg.to_host(b)
    orc = O.RnsPlan(moduli, n)
    for p in range(3):
        assert (got[p] == O.encrypt(orc, m[p], e[p], s, a[p])).all(), f"poly {p}"
    dec = pkg.to_host(plan.decrypt_raw(b, da, ds))
    for p in range(3):
        assert (dec[p] == O.decrypt(orc, got[p], a[p], s)).all()
        want = np.stack([(m[p, i].astype(object) + np.array([int(v) % q for v in e[p]], dtype=object)) % q
                         for i, q in enumerate(moduli)]).astype(np.uint64)
        assert (dec[p] == want).all()
