"""Rotation-schedule programs for the linear-algebra workloads (BSGS GEMV,
radix-4 FFT, Jiang CCMM): the FP64 cleartext execution of every program
(rotations as cyclic shifts, plaintext products elementwise) must equal
numpy's matrix product / FFT.  This pins the schedules themselves; the GPU
tests run the same programs encrypted."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_pkg

PKG = load_pkg()
LIN = PKG.linear


def _builder(slots):
    n2 = 2 * (2 * slots)  # 2N with N = 2 * slots
    return LIN.Builder(slots, lambda limbs, steps: (None, pow(5, steps, n2)), lambda values, limbs: None,
                       moduli=[0] * 16)


@pytest.mark.parametrize("d,slots", [(16, 64), (64, 256), (100, 200)])
def test_gemv_bsgs_schedule(d, slots):
    rng = np.random.default_rng(d)
    M = rng.uniform(-1, 1, (d, d))
    x = rng.uniform(-1, 1, d)
    prog = LIN.gemv_program(_builder(slots), M, limbs=3)
    got = prog.simulate([LIN.tile(x, d, slots)])
    assert np.allclose(got, LIN.tile(M @ x, d, slots), atol=1e-12)
    # BSGS: about 2 sqrt(d) rotations instead of d - 1
    assert prog.counts()["rotate"] <= 2 * int(np.ceil(np.sqrt(d)))


@pytest.mark.parametrize("d,slots", [(4, 8), (64, 128), (128, 128), (256, 1024), (4096, 4096)])
def test_radix4_fft_schedule(d, slots):
    rng = np.random.default_rng(d)
    x = rng.uniform(-1, 1, d) + 1j * rng.uniform(-1, 1, d)
    prog = LIN.fft_program(_builder(slots), d, limbs=16)
    perm = LIN.bitrev_permutation(d)
    got = prog.simulate([LIN.tile(x[perm], d, slots)])
    assert np.allclose(got, LIN.tile(np.fft.fft(x), d, slots), atol=1e-9 * d)
    stages = LIN.fft_stages(d)
    assert len(stages) == (d.bit_length() - 1 + 1) // 2
    assert all(len(s) <= 7 for s in stages)


@pytest.mark.parametrize("d,slots", [(4, 16), (8, 64), (16, 512)])
def test_jiang_ccmm_schedule(d, slots):
    rng = np.random.default_rng(d)
    A = rng.uniform(-1, 1, (d, d))
    B = rng.uniform(-1, 1, (d, d))
    prog = LIN.ccmm_program(_builder(slots), d, limbs=5, relin_key=None)
    got = prog.simulate([LIN.tile(A.ravel(), d * d, slots), LIN.tile(B.ravel(), d * d, slots)])
    assert np.allclose(got, LIN.tile((A @ B).ravel(), d * d, slots), atol=1e-12)
    assert prog.counts()["mul_relin_rescale"] == d


def test_diagonal_algebra():
    rng = np.random.default_rng(3)
    d = 12
    A, B = rng.standard_normal((d, d)), rng.standard_normal((d, d))
    comp = LIN.compose_diagonals(LIN.matrix_diagonals(A), LIN.matrix_diagonals(B), d)
    ref = LIN.matrix_diagonals(A @ B)
    for i, v in ref.items():
        assert np.allclose(comp[i], v)
    cols = rng.permutation(d)
    P = np.zeros((d, d))
    P[np.arange(d), cols] = 1
    pd = LIN.permutation_diagonals(cols)
    md = LIN.matrix_diagonals(P)
    assert sorted(pd) == sorted(md) and all(np.allclose(pd[i], md[i]) for i in md)
