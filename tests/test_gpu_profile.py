"""The device execution profile (ckks_profile_get, the reference's
ExecutionProfile, execution/mod.rs:162-230) counts what the entry points ran:
NTTs per polynomial and per limb, rotations, conjugations, multiplies,
relinearizations, rescales; autotuning trials count nothing; a captured
graph counts its recorded calls once per launch; the evaluator can emit the
device's NTT events into its trace."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import O, P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def test_profile_counts_device_operations(pkg):
    n, L, K = 1 << 12, 4, 1
    q = [Q0[0]] + Q40[:L - 1]
    qp = q + P60[:K]
    rng = np.random.default_rng(8)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx = pkg.Context(0, stream=stream)
        plan = pkg.PreparedRnsNttPlan(ctx, q, n)
        qp_plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
        key = pkg.HybridKeySwitchKey(qp_plan, L, [2, 2], rand_residues(rng, qp, n, 2), rand_residues(rng, qp, n, 2))
        ctx.reset_execution_profile()
        x = pkg.to_device(rand_residues(rng, q, n, polys=3))
        plan.forward(x)
        p = ctx.execution_profile()
        assert p.ntt_forwards == 3 and p.ntt_limb_transforms == 3 * L and p.total_operations() == 3
        ct = pkg.to_device(rand_residues(rng, q, n, polys=4).reshape(2, 2, L, n))
        ctx.reset_execution_profile()
        key.apply_galois(ct, O.rotation_exponent_left(n, 1))
        key.apply_galois(ct, O.conjugation_exponent(n))
        p = ctx.execution_profile()
        assert p.rotations == 2 and p.automorphisms == 2 and p.key_switches == 4
        # each key switch transforms its 2 digits forward and 2 accumulators back
        assert p.ntt_forwards == 4 * 2 and p.ntt_inverses == 4 * 2
        ctx.reset_execution_profile()
        key.multiply_relinearize_rescale(ct, ct)
        plan.multiply_plain(ct, pkg.to_device(rand_residues(rng, q, n)[0]), L)
        plan.add(ct, ct)
        p = ctx.execution_profile()
        assert (p.ciphertext_ciphertext_multiplies, p.relinearizations, p.rescales) == (2, 2, 2)
        assert p.ciphertext_plaintext_multiplies == 2 and p.additions == 4  # 4 polynomials
        # autotuning runs trials but leaves the profile as it was
        before = ctx.execution_profile()
        key.autotune(2)
        plan.autotune_ntt(2)
        assert ctx.execution_profile() == before
        # a captured rotation counts once per launch, not at capture
        out = torch.empty_like(ct)
        key.apply_galois(ct, 5, out=out)  # sizes the workspaces eagerly
        ctx.reset_execution_profile()
        ctx.graph_begin()
        key.apply_galois(ct, 5, out=out)
        g = ctx.graph_end()
        assert ctx.execution_profile().rotations == 0
        g.launch()
        g.launch()
        stream.synchronize()
        assert ctx.execution_profile().rotations == 4
        g.close()
        ctx.close()


def test_evaluator_trace_carries_device_ntt_events(pkg, ctx):
    CK = pkg.ckks
    n, L, K = 1 << 12, 3, 1
    q = [Q0[0]] + Q40[:L - 1]
    rng = np.random.default_rng(9)
    chain = CK.ModulusChain(q)
    plan = pkg.PreparedRnsNttPlan(ctx, q, n)
    q1 = q
    qp = pkg.PreparedRnsNttPlan(ctx, q1 + P60[:K], n)
    lvl = CK.RnsCkksLevelKeys(0, q1)
    e = CK.rotation_exponent_left(n, 1)
    lvl.insert_galois_key(e, pkg.HybridKeySwitchKey(qp, L, [2, 1], rand_residues(rng, q1 + P60[:K], n, 2),
                                                   rand_residues(rng, q1 + P60[:K], n, 2)))
    keys = CK.RnsCkksEvaluationKeys()
    keys.insert_level(lvl)
    trace = CK.ExecutionTrace()
    ev = CK.RnsCkksEvaluator(chain, keys, plan, trace, device_events=True)
    x = ev.encrypt_state(pkg.to_device(rand_residues(rng, q, n, polys=2)), 2.0 ** 40)
    ev.rotate_left(x, 1)
    kinds = [ev_.kind for ev_ in trace.events()]
    assert kinds[0] == CK.ExecutionEventKind.ROTATE
    assert trace.count(CK.ExecutionEventKind.NTT_FORWARD) == 2 and trace.count(CK.ExecutionEventKind.NTT_INVERSE) == 2
