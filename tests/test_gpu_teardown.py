"""Teardown order: a context closed before the plans, keys, hoisted digits,
servers and graphs made on it, whose finalizers then run (explicitly, and at
interpreter exit), ends the process cleanly.  Runs in a child process so that
its exit status is the check."""
from __future__ import annotations

import os
import subprocess
import sys
import textwrap

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = textwrap.dedent("""
    import gc, sys
    import numpy as np
    sys.path.insert(0, {tests!r})
    from conftest import load_pkg, Q0, Q40, P60, rand_residues
    pkg = load_pkg()
    n, L, K, w = 1 << 12, 4, 1, 61
    q = [Q0[0]] + Q40[:L - 1]
    qp = q + P60[:K]
    rng = np.random.default_rng(5)
    import torch
    stream = torch.cuda.Stream()  # graph capture needs a non-null stream
    torch.cuda.set_stream(stream)
    ctx = pkg.Context(0, stream=stream)
    plan = pkg.PreparedRnsNttPlan(ctx, qp, n)
    qplan = pkg.PreparedRnsNttPlan(ctx, q, n)
    cnt = pkg.bounded_digit_count(plan, L, w)
    bkey = pkg.BoundedKeySwitchKey(plan, L, w, rand_residues(rng, qp, n, cnt), rand_residues(rng, qp, n, cnt))
    hkey = pkg.HybridKeySwitchKey(plan, L, [2, 2], rand_residues(rng, qp, n, 2), rand_residues(rng, qp, n, 2))
    rkey = pkg.RnsKeySwitchKey(qplan, [1] * L, rand_residues(rng, q, n, L), rand_residues(rng, q, n, L))
    ct = pkg.to_device(rand_residues(rng, q, n, 2).reshape(1, 2, L, n))
    out = bkey.apply_galois(ct, 5)
    hoisted = bkey.hoist(ct)
    bkey.rotate_hoisted(hoisted, ct, 25)
    hkey.apply_galois(ct, 5)
    rkey.switch(ct)
    server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 28)
    ctx.graph_begin()
    hkey.apply_galois(ct, 5, out=out)
    graph = ctx.graph_end()
    graph.launch()
    ctx.sync()
    ctx.close()           # the context first
    del hoisted, server   # then some finalizers explicitly
    gc.collect()
    rkey.close()
    print("teardown-ok", flush=True)
    # graph, keys and plans are left to interpreter exit
""")


def test_context_closed_before_its_objects():
    script = SCRIPT.format(tests=os.path.join(ROOT, "tests"))
    r = subprocess.run([sys.executable, "-c", script], cwd=ROOT, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, f"exit {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    assert "teardown-ok" in r.stdout
