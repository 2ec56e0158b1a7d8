"""CKKS encode / decode on the device.

CanonicalEmbedding restates CkksCanonicalEmbedding (ckks/embedding.rs:
40-347): roots xi_k = exp(i pi (2k+1)/N); logical slot j is the root with
exponent 5^j mod 2N; evaluation is the phase twist exp(i pi k/N) followed by
a positive-sign DFT, interpolation the negative-sign DFT with 1/N and the
inverse twist.  The DFTs run in FP64 on the GPU (torch.fft); the integer
halves (rounding to residues, exact centered CRT) are HIP kernels
(ckks_encode_coefficients / ckks_decode_coefficients).
"""
from __future__ import annotations

import math

import torch

from . import PreparedRnsNttPlan, _check, _ptr, empty, lib


class CanonicalEmbedding:
    def __init__(self, degree: int, device="cuda"):
        self.n = degree
        two_n = 2 * degree
        idx, e = [], 1
        for _ in range(degree // 2):
            idx.append((e - 1) // 2)
            e = (e * 5) % two_n
        self.slot_root_indices = torch.tensor(idx, dtype=torch.long, device=device)
        self.conj_indices = degree - 1 - self.slot_root_indices
        k = torch.arange(degree, dtype=torch.float64, device=device)
        self.twist = torch.polar(torch.ones_like(k), math.pi * k / degree)

    def slots_to_coefficients(self, slots: torch.Tensor) -> torch.Tensor:
        """slots [..., N/2] complex -> real coefficients [..., N] (float64)."""
        slots = slots.to(torch.complex128)
        ev = torch.zeros(slots.shape[:-1] + (self.n,), dtype=torch.complex128, device=slots.device)
        ev[..., self.slot_root_indices] = slots
        ev[..., self.conj_indices] = slots.conj()
        coef = torch.fft.fft(ev) / self.n * self.twist.conj()
        return coef.real.contiguous()

    def coefficients_to_slots(self, coeffs: torch.Tensor) -> torch.Tensor:
        c = coeffs.to(torch.float64) * self.twist
        ev = torch.fft.ifft(c) * self.n
        return ev[..., self.slot_root_indices]


def encode(plan: PreparedRnsNttPlan, emb: CanonicalEmbedding, slots, scale: float, n_limbs: int | None = None):
    """encode_rns (application_support/ckks.rs:39-66) on the device: slots
    [p][N/2] (or [N/2]) -> coefficient-domain residues [p][limbs][N]."""
    n_limbs = len(plan.moduli) if n_limbs is None else n_limbs
    if not isinstance(slots, torch.Tensor):
        slots = torch.as_tensor(slots)
    slots = slots.to("cuda")
    single = slots.dim() == 1
    if single:
        slots = slots[None]
    coeffs = emb.slots_to_coefficients(slots).contiguous()
    p = coeffs.shape[0]
    out = empty((p, n_limbs, plan.degree))
    _check(lib().ckks_encode_coefficients(plan.handle, _ptr(coeffs, p * plan.degree, dtype="float64", name="coeffs"),
                                          float(scale), _ptr(out), p, n_limbs))
    return out[0] if single else out


def decode_coefficients(plan: PreparedRnsNttPlan, residues, scale: float, n_limbs: int | None = None):
    """Exact centered CRT / scale on the device: [p][limbs][N] -> [p][N] float64."""
    n_limbs = len(plan.moduli) if n_limbs is None else n_limbs
    p = plan._count(residues, n_limbs)
    out = torch.empty((p, plan.degree), dtype=torch.float64, device="cuda")
    _check(lib().ckks_decode_coefficients(plan.handle, _ptr(residues, p * n_limbs * plan.degree, name="residues"),
                                          float(scale), _ptr(out, dtype="float64"), p, n_limbs))
    return out


def decode(plan: PreparedRnsNttPlan, emb: CanonicalEmbedding, residues, scale: float, n_limbs: int | None = None):
    """Plaintext residues -> slots [p][N/2] (bin/r30_13b_ckks_primitives.rs:55-80)."""
    return emb.coefficients_to_slots(decode_coefficients(plan, residues, scale, n_limbs))
