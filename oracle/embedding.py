"""FP64 cleartext side of the oracle: canonical embedding, encode, decode.

TEST INFRASTRUCTURE ONLY.  numpy restatement of
  CkksCanonicalEmbedding (ckks/embedding.rs:40-347): roots xi_j =
    exp(i pi (2j+1)/N); logical slot j is the root with exponent 5^j mod 2N;
    evaluation = phase twist zeta^k then a positive-sign DFT; interpolation =
    negative-sign DFT, 1/N and the inverse twist;
  encode_rns (application_support/ckks.rs:39-66): round(coef * scale), half
    away from zero as f64::round, then rem_euclid per limb;
  decode (bin/r30_13b_ckks_primitives.rs:55-80): exact CRT, centered,
    divided by the scale, then coefficients_to_slots.
FP64 results agree with the reference's radix-2 FFT to rounding (not bit for
bit); the tests compare slots against FP64 products with stated tolerances.
"""
from __future__ import annotations

import math

import numpy as np


class Embedding:
    def __init__(self, degree: int):
        self.n = degree
        two_n = 2 * degree
        idx, e = [], 1
        for _ in range(degree // 2):
            idx.append((e - 1) // 2)
            e = (e * 5) % two_n
        self.slot_root_indices = np.array(idx)
        k = np.arange(degree)
        self.twist = np.exp(1j * np.pi * k / degree)

    def evaluate_all(self, coeffs) -> np.ndarray:
        c = np.asarray(coeffs, dtype=np.float64) * self.twist
        return np.fft.ifft(c) * self.n  # positive-sign DFT, no normalisation

    def coefficients_to_slots(self, coeffs) -> np.ndarray:
        return self.evaluate_all(coeffs)[self.slot_root_indices]

    def slots_to_coefficients(self, slots) -> np.ndarray:
        ev = np.zeros(self.n, dtype=np.complex128)
        ev[self.slot_root_indices] = slots
        ev[self.n - 1 - self.slot_root_indices] = np.conj(slots)
        coef = np.fft.fft(ev) / self.n * np.conj(self.twist)
        assert np.all(np.abs(coef.imag) <= 1e-10 * (1 + np.abs(coef.real)))
        return coef.real


def round_half_away(x: float) -> int:
    """f64::round: halfway cases away from zero (x - trunc(x) is exact)."""
    t = math.trunc(x)
    if abs(x - t) >= 0.5:
        t += 1 if x > 0 else -1
    return int(t)


def encode_rns(slots, emb: Embedding, moduli, scale: float) -> np.ndarray:
    coeffs = emb.slots_to_coefficients(slots)
    signed = [round_half_away(v * scale) for v in coeffs]
    return np.array([[s % q for s in signed] for q in moduli], dtype=np.uint64)


def crt_centered(residues: np.ndarray, moduli) -> list[int]:
    """Exact CRT of every coefficient, centered to (-Q/2, Q/2]."""
    Q = math.prod(moduli)
    x = [0] * residues.shape[1]
    for r, q in zip(residues, moduli):
        Mi = Q // q
        f = Mi * pow(Mi, -1, q)
        for c, v in enumerate(r.tolist()):
            x[c] += int(v) * f
    out = []
    for v in x:
        v %= Q
        out.append(v - Q if v > Q // 2 else v)
    return out


def decode(plain_residues: np.ndarray, moduli, scale: float, emb: Embedding) -> np.ndarray:
    coeffs = [v / scale for v in crt_centered(plain_residues, moduli)]
    return emb.coefficients_to_slots(coeffs)
