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


@pytest.mark.parametrize("rows,cols", [(16, 16), (8, 32), (32, 8), (256, 256)])
def test_fft2_schedule(rows, cols):
    """2-D DFT of a rows x cols image held as two half-image ciphertexts:
    the even / odd programs' FP64 execution equals numpy's fft2 at the
    spectrum rows 2 bitrev(j) and 2 bitrev(j) + 1, with every rotation
    confined to its row / column by the signed-offset diagonals."""
    slots = rows // 2 * cols
    rng = np.random.default_rng(rows * cols)
    X = rng.uniform(-1, 1, (rows, cols))
    cp = LIN.bitrev_permutation(cols)
    top, bot = X[: rows // 2][:, cp].ravel(), X[rows // 2:][:, cp].ravel()
    pe, po = LIN.fft2_programs(_builder(slots), rows, cols, limbs=20)
    Y = np.fft.fft2(X)
    rp = LIN.bitrev_permutation(rows // 2)
    assert np.allclose(pe.simulate([top, bot]).reshape(rows // 2, cols), Y[2 * rp], atol=1e-9)
    assert np.allclose(po.simulate([top, bot]).reshape(rows // 2, cols), Y[2 * rp + 1], atol=1e-9)
    levels = LIN.fft2_levels(rows, cols)
    assert pe.counts()["rescale"] == po.counts()["rescale"] == levels
    # rotations per stage: at most 6 signed offsets of a radix-4 stage
    assert pe.counts()["rotate"] <= 6 * levels


def test_signed_diagonals_do_not_wrap():
    """A signed-offset diagonal never pairs an element with one of another
    row: D_o[j] = 0 wherever j + o leaves [0, d)."""
    d = 16
    M = LIN.dit_stage_matrix(d, 4) @ LIN.dit_stage_matrix(d, 2)
    sd = LIN.signed_diagonals(M)
    rebuilt = np.zeros_like(M)
    for o, v in sd.items():
        for j in range(d):
            if v[j] != 0:
                assert 0 <= j + o < d
                rebuilt[j, j + o] = v[j]
    assert np.allclose(rebuilt, M)
