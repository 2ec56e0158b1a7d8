"""ckks_ntt_forward_noalloc, the NTT entry point that allocates nothing:
residues identical to ckks_ntt_forward on two calls in a row, the second
with a smaller input (a 75-limb-transform batch at N = 2^16 that runs in two
chunks of the 32 MiB scratch, then 10 limbs at N = 2^13), in place as well;
the device's free memory does not move across the calls."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import P60, Q0, Q40, rand_residues

pytestmark = pytest.mark.gpu


def test_noalloc_forward_matches_forward(pkg, ctx):
    cases = [(16, [Q0[0]] + Q40[:24], 3), (13, [Q0[0]] + Q40[:3] + P60[:1], 2)]
    rng = np.random.default_rng(2027)
    data = []
    for logn, moduli, polys in cases:
        n = 1 << logn
        plan = pkg.PreparedRnsNttPlan(ctx, moduli, n)
        x = pkg.to_device(rand_residues(rng, moduli, n, polys=polys))
        want = plan.forward(x, out=torch.empty_like(x))
        data.append((plan, x, want, torch.empty_like(x), x.clone()))
    torch.cuda.synchronize()
    free_before = torch.cuda.mem_get_info()[0]
    for plan, x, want, out, inplace in data:
        plan.forward_noalloc(x, out)
        plan.forward_noalloc(inplace, inplace)
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free_before
    for plan, x, want, out, inplace in data:
        assert torch.equal(out, want)
        assert torch.equal(inplace, want)
