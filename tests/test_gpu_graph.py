"""CP multiply parity, and hipGraph-captured rotation schedules: a captured
BSGS-style sequence (rotations + plaintext-diagonal multiplies) replayed
several times, with eager work and synchronisations between replays, must
match the eager results every time."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def test_multiply_plain_bit_exact(pkg, ctx):
    n, L = 1 << 13, 5
    moduli = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(3)
    ct = rand_residues(rng, moduli, n, polys=6).reshape(3, 2, L, n)
    pt = rand_residues(rng, moduli, n)[0]
    pts = rand_residues(rng, moduli, n, polys=3)
    plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
    orc = O.RnsPlan(moduli, n)
    got1 = pkg.to_host(plan.multiply_plain(pkg.to_device(ct), pkg.to_device(pt), L))
    got3 = pkg.to_host(plan.multiply_plain(pkg.to_device(ct), pkg.to_device(pts), L))
    for c in range(3):
        for comp in range(2):
            for p, got in ((pt, got1), (pts[c], got3)):
                want = orc.inverse(np.stack([O.pointwise_mul(f, p[i], q) for i, (f, q) in
                                             enumerate(zip(orc.forward(ct[c, comp]), moduli))]))
                assert (got[c, comp] == want).all()


def test_graph_captured_rotation_schedule(pkg):
    import torch

    n, L, K, blocks = 1 << 13, 5, 2, [3, 2]
    q_moduli = [Q0[0]] + Q40[:L - 1]
    qp = q_moduli + P60[:K]
    rng = np.random.default_rng(4)
    stream = torch.cuda.Stream()  # capture needs a non-null stream
    with torch.cuda.stream(stream):
        ctx = pkg.Context(0, stream=stream)
        plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
        qplan = pkg.PreparedRnsNttPlan(ctx, q_moduli, n)
        steps = [1, 2, 4]
        keys = []
        for r in steps:
            kb = rand_residues(rng, qp, n, len(blocks))
            ka = rand_residues(rng, qp, n, len(blocks))
            keys.append((O.rotation_exponent_left(n, r), pkg.HybridKeySwitchKey(plan, L, blocks, kb, ka)))
        diag = pkg.to_device(rand_residues(rng, q_moduli, n)[0])
        x = pkg.to_device(rand_residues(rng, q_moduli, n, polys=4).reshape(2, 2, L, n))
        bufs = [pkg.empty((2, 2, L, n)) for _ in steps]

        def schedule():
            src = x
            for (k, key), out in zip(keys, bufs):
                key.apply_galois(src, k, out)
                qplan.multiply_plain(out, diag, L, out=out)
                src = out

        schedule()  # eager run sizes every workspace
        stream.synchronize()
        eager = pkg.to_host(bufs[-1])
        ctx.graph_begin()
        schedule()
        graph = ctx.graph_end()
        for rep in range(3):
            # new input between replays, eager reference recomputed
            newx = rand_residues(np.random.default_rng(100 + rep), q_moduli, n, polys=4).reshape(2, 2, L, n)
            x.copy_(pkg.to_device(newx))
            schedule()
            stream.synchronize()
            want = pkg.to_host(bufs[-1]).copy()
            for b in bufs:
                b.zero_()
            graph.launch()
            stream.synchronize()
            got = pkg.to_host(bufs[-1])
            assert (got == want).all(), f"replay {rep}"
            assert not (want == eager).all()
        graph.close()
        ctx.close()


def test_encode_refuses_capture_and_context_stays_usable(pkg):
    """Encoding reads an overflow flag back to the host, which a captured
    graph cannot do: inside graph_begin/graph_end it fails with a status
    code, the capture still closes, and the context encodes afterwards."""
    import torch

    n, L = 1 << 12, 3
    moduli = [Q0[0]] + Q40[:L - 1]
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx = pkg.Context(0, stream=stream)
        plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
        emb = pkg.encoding.CanonicalEmbedding(n)
        slots = torch.zeros(n // 2, dtype=torch.complex128)
        slots[:4] = torch.tensor([0.5, -0.25, 1.0, 0.0], dtype=torch.complex128)
        want = pkg.to_host(pkg.encoding.encode(plan, emb, slots, 2.0 ** 30))
        # the embedding's FFT runs before the capture opens
        coeffs = emb.slots_to_coefficients(slots.cuda()[None]).contiguous()
        out = pkg.empty((1, L, n))
        ctx.graph_begin()
        with pytest.raises(pkg.CkksError, match="capturable"):
            pkg._check(pkg.lib().ckks_encode_coefficients(plan.handle, pkg._ptr(coeffs, dtype="float64"),
                                                          2.0 ** 30, pkg._ptr(out), 1, L))
        g = ctx.graph_end()
        g.close()
        again = pkg.to_host(pkg.encoding.encode(plan, emb, slots, 2.0 ** 30))
        stream.synchronize()
        assert (again == want).all()
        plan.close()
        ctx.close()
