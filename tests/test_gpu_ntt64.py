"""The single-modulus NTT for primes up to 64 bits (ckks_ntt64_*), the
reference's 64-bit parameter set (ring/ntt_parameter_sets.rs:11-12)
included: psi, the transform of X (which is psi^(2k+1)), and random
polynomials bit-exact against the oracle, forward and inverse."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import O, Q40

pytestmark = pytest.mark.gpu

NTT64_MODULUS = 18446744073709436929  # 2^64 - 114687, 2^14 | q - 1
NTT64_PSI = 5975861664659593359


def test_64_bit_parameter_set(pkg, ctx):
    n = 4096
    plan = pkg.PreparedNtt64(ctx, NTT64_MODULUS, n)
    assert plan.psi == NTT64_PSI
    x = np.zeros(n, dtype=np.uint64)
    x[1] = 1
    out = pkg.to_host(plan.forward(pkg.to_device(x)))
    q = NTT64_MODULUS
    want = [pow(NTT64_PSI, 2 * k + 1, q) for k in range(n)]
    assert [int(v) for v in out] == want


@pytest.mark.parametrize("case", ["q64_4096", "q64_8192", "q63_65536", "q63_131072", "q40_16384"])
def test_ntt64_bit_exact(pkg, ctx, case):
    q, n = {"q64_4096": (NTT64_MODULUS, 4096), "q64_8192": (NTT64_MODULUS, 8192),
            "q63_65536": (O.ntt_primes(63, 1 << 16, 1)[0], 1 << 16),
            "q63_131072": (O.ntt_primes(63, 1 << 17, 1)[0], 1 << 17),
            "q40_16384": (Q40[0], 1 << 14)}[case]
    rng = np.random.default_rng(n % 1000 + 7)
    x = rng.integers(0, q, (3, n), dtype=np.uint64)
    x[1] = q - 1
    x[2, :5] = 0
    plan = pkg.PreparedNtt64(ctx, q, n)
    orc = O.Plan(q, n)
    assert plan.psi == orc.psi
    d = pkg.to_device(x)
    f = pkg.to_host(plan.forward(d, out=pkg.empty(d.shape)))
    for i in range(3):
        assert (f[i] == orc.forward(x[i])).all(), f"forward poly {i}"
    back = pkg.to_host(plan.inverse(pkg.to_device(f)))
    assert (back == x).all()
    y = rng.integers(0, q, (2, n), dtype=np.uint64)
    dy = pkg.to_device(y)
    plan.inverse(dy)  # in place
    got = pkg.to_host(dy)
    for i in range(2):
        assert (got[i] == orc.inverse(y[i])).all(), f"inverse poly {i}"


def test_ntt64_rejections(pkg, ctx):
    with pytest.raises(pkg.CkksError, match="2N must divide"):
        pkg.PreparedNtt64(ctx, NTT64_MODULUS, 1 << 15)  # 2^16 does not divide q - 1
    with pytest.raises(pkg.CkksError, match="prime"):
        pkg.PreparedNtt64(ctx, 16385, 4096)  # 2 * 8192 + 1 = 5 * 29 * 113
    with pytest.raises(pkg.CkksError, match="degrees"):
        pkg.PreparedNtt64(ctx, NTT64_MODULUS, 256)
