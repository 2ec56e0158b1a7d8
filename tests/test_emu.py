"""The NTT kernel source compiled for the CPU (tests/emu), with every LDS,
HBM and twiddle index bounds-checked, against the oracle: checks the
kernels' addressing and arithmetic before they run on a GPU."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import O, P60, Q0, Q40

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(os.path.join(HERE, "libntt_emu.so"))
    lib.emu_ntt.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int]
    return lib


@pytest.mark.parametrize("logn", list(range(10, 18)))
def test_emulated_kernel_ntt_matches_oracle(emu, logn):
    n = 1 << logn
    rng = np.random.default_rng(logn)
    primes = [Q0[0], Q40[0], P60[0]]
    if logn == 17:
        primes = [O.ntt_primes(60, 1 << 17, 1)[0]]
    for q in primes:
        x = rng.integers(0, q, n, dtype=np.uint64)
        x[:3] = [0, q - 1, q - 1]
        out = np.zeros(n, dtype=np.uint64)
        assert emu.emu_ntt(x.ctypes.data, out.ctypes.data, q, logn, 1) == 0
        plan = O.Plan(q, n)
        assert (out == plan.forward(x)).all()
        back = np.zeros(n, dtype=np.uint64)
        assert emu.emu_ntt(out.ctypes.data, back.ctypes.data, q, logn, 0) == 0
        assert (back == x).all()
        # inverse of arbitrary NTT-domain input equals the oracle inverse
        y = rng.integers(0, q, n, dtype=np.uint64)
        assert emu.emu_ntt(y.ctypes.data, back.ctypes.data, q, logn, 0) == 0
        assert (back == plan.inverse(y)).all()


@pytest.mark.parametrize("logn", [13, 14, 15, 16, 17])
@pytest.mark.parametrize("tb,blk", [(12, 1), (13, 0), (13, 1)])
def test_emulated_kernel_variants_match_oracle(emu, logn, tb, blk):
    """Every NTT kernel variant the autotuner may pick (tile of 2^tb
    residues, blocked inter-pass layout) gives the oracle's residues."""
    if tb == 13 and logn < 14:
        pytest.skip("8192-residue tiles serve N >= 2^14")
    emu.emu_ntt_variant.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                    C.c_int, C.c_int]
    n = 1 << logn
    rng = np.random.default_rng(100 + logn)
    primes = [Q40[1], P60[1]] if logn < 17 else [O.ntt_primes(60, 1 << 17, 1)[0]]
    for q in primes:
        plan = O.Plan(q, n)
        x = rng.integers(0, q, n, dtype=np.uint64)
        out = np.zeros(n, dtype=np.uint64)
        assert emu.emu_ntt_variant(x.ctypes.data, out.ctypes.data, q, logn, 1, tb, blk) == 0
        assert (out == plan.forward(x)).all()
        back = np.zeros(n, dtype=np.uint64)
        assert emu.emu_ntt_variant(x.ctypes.data, back.ctypes.data, q, logn, 0, tb, blk) == 0
        assert (back == plan.inverse(x)).all()
