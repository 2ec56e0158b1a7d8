"""CKKS encode / decode on the device (ckks/embedding.rs,
application_support/ckks.rs:39-66, bin/r30_13b_ckks_primitives.rs:55-80)
against the oracle's FP64 embedding and exact big-integer CRT.  The FP64
DFTs of the two sides round differently in the last bits, so encoded
integers may differ by one where coefficient * scale sits at a rounding
boundary; decoded values agree to FP64 precision."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import embedding as E
from conftest import O, P60, Q0, Q40, load_pkg, rand_residues

ENC = load_pkg().encoding


@pytest.mark.parametrize("n", [16, 1024, 1 << 13])
def test_embedding_matches_oracle_on_cpu(n):
    rng = np.random.default_rng(n)
    z = rng.uniform(-1, 1, n // 2) + 1j * rng.uniform(-1, 1, n // 2)
    ours = ENC.CanonicalEmbedding(n, device="cpu")
    ref = E.Embedding(n)
    c = ours.slots_to_coefficients(torch.from_numpy(z)).numpy()
    assert np.allclose(c, ref.slots_to_coefficients(z), atol=1e-12)
    back = ours.coefficients_to_slots(torch.from_numpy(c)).numpy()
    assert np.allclose(back, z, atol=1e-12)


def test_round_half_away_reference_semantics():
    # f64::round, as encode_rns uses it
    assert [E.round_half_away(v) for v in (0.5, -0.5, 1.5, -1.5, 2.4999999999999996)] == [1, -1, 2, -2, 2]


@pytest.mark.gpu
@pytest.mark.parametrize("logn,L", [(12, 4), (13, 9), (16, 25)])
def test_gpu_encode_decode_against_oracle(pkg, ctx, logn, L):
    n = 1 << logn
    moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(logn)
    z = rng.uniform(-1, 1, (2, n // 2)) + 1j * rng.uniform(-1, 1, (2, n // 2))
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    emb = ENC.CanonicalEmbedding(n)
    scale = 2.0 ** 40
    got = pkg.to_host(ENC.encode(plan, emb, torch.from_numpy(z), scale))
    ref_emb = E.Embedding(n)
    for p in range(2):
        want = E.encode_rns(z[p], ref_emb, moduli, scale)
        # centered integers differ by at most one (FP64 DFT rounding)
        g0 = np.array([E.crt_centered(got[p][:, c:c + 1], moduli)[0] for c in range(0, n, max(1, n // 512))])
        w0 = np.array([E.crt_centered(want[:, c:c + 1], moduli)[0] for c in range(0, n, max(1, n // 512))])
        assert np.abs(g0 - w0).max() <= 1
        assert (got[p] == want).mean() > 0.99
        # decode: exact CRT on the device vs the oracle's big integers
        dec = ENC.decode(plan, emb, pkg.to_device(want), scale).cpu().numpy()
        ref = E.decode(want, moduli, scale, ref_emb)
        assert np.allclose(dec, ref, rtol=0, atol=1e-9)
        assert np.max(np.abs(dec - z[p])) < 1e-6


@pytest.mark.gpu
def test_gpu_decode_exact_boundaries(pkg, ctx):
    """Centered map at 0, floor(Q/2), floor(Q/2) + 1 and Q - 1, and values
    beyond 2^64 in magnitude."""
    n = 4096
    moduli = [Q0[0]] + Q40[:3] + P60[:2]
    Q = math.prod(moduli)
    vals = [0, 1, Q // 2, Q // 2 + 1, Q - 1, (1 << 100) + 12345, Q - (1 << 90), 3 << 63]
    x = np.zeros((len(moduli), n), dtype=np.uint64)
    for c, v in enumerate(vals):
        x[:, c] = [v % q for q in moduli]
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    got = ENC.decode_coefficients(plan, pkg.to_device(x), 1.0).cpu().numpy()[0]
    for c, v in enumerate(vals):
        cen = v - Q if v > Q // 2 else v
        assert got[c] == pytest.approx(float(cen), rel=1e-15, abs=0), (c, v)
