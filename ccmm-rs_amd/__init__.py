"""MI355X (gfx950) backend for the leveled RNS-CKKS hot path of ccmm-rs.

Python binding of ``libckks_hip.so`` (the C ABI declared in
``include/ckks_hip.h``).  Class and method names mirror the reference's Rust
API (``PreparedRnsNttPlan::forward``, ``rescale_rns_ckks_to_next``,
``rns_key_switch_with_prepared_dynamic_ntt`` ...) so callers and tests read
like the reference's own.  Device memory is held in torch tensors of dtype
int64 whose bits are the uint64 residues; torch is only the allocator and
stream provider.  There is no CPU fallback: every operation runs the HIP
kernels, and a missing library raises.

The directory name contains a hyphen, so load it with :func:`load` from a
path (``importlib``) rather than ``import``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libckks_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "ckks_hip.h")

_lib = None

_sz = C.c_size_t
_u64 = C.c_uint64
_p = C.c_void_p


class CkksError(RuntimeError):
    """A library call returned a non-zero status; message from ckks_last_error()."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def build(quiet: bool = True) -> str:
    """Compile libckks_hip.so in-tree for gfx950 (hipcc cross-compiles)."""
    import subprocess

    cmd = ["make", "-j8", "-C", _HERE]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL if quiet else None)
    return LIB_PATH


_SIGNATURES = {
    "ckks_last_error": (C.c_char_p, []),
    "ckks_version": (C.c_char_p, []),
    "ckks_context_create": (C.c_int, [C.c_int, _p, C.POINTER(_p)]),
    "ckks_context_destroy": (C.c_int, [_p]),
    "ckks_context_set_stream": (C.c_int, [_p, _p]),
    "ckks_sync": (C.c_int, [_p]),
    "ckks_set_ntt_chunk_bytes": (C.c_int, [_p, _sz]),
    "ckks_set_base_conversion": (C.c_int, [_p, C.c_int]),
    "ckks_set_ntt_variant": (C.c_int, [_p, C.c_uint32, C.c_int]),
    "ckks_set_key_switch_fusion": (C.c_int, [_p, C.c_int]),
    "ckks_set_key_switch_split": (C.c_int, [_p, C.c_int]),
    "ckks_get_ntt_variant": (C.c_int, [_p, C.c_uint32, C.POINTER(C.c_int)]),
    "ckks_autotune_ntt": (C.c_int, [_p, _sz, _sz, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "ckks_autotune_key_switch": (C.c_int, [_p, _sz, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_float)]),
    "ckks_malloc": (C.c_int, [_p, _sz, C.POINTER(_p)]),
    "ckks_free": (C.c_int, [_p, _p]),
    "ckks_upload": (C.c_int, [_p, _p, _p, _sz]),
    "ckks_download": (C.c_int, [_p, _p, _p, _sz]),
    "ckks_find_ntt_primes": (C.c_int, [C.c_uint, _sz, _sz, _p, _sz, _p, C.POINTER(_sz)]),
    "ckks_rns_plan_create": (C.c_int, [_p, _p, _sz, _sz, C.POINTER(_p)]),
    "ckks_rns_plan_destroy": (C.c_int, [_p]),
    "ckks_rns_plan_psi": (C.c_int, [_p, _p]),
    "ckks_rotate_hoisted_batch": (C.c_int, [_p, _p, _sz]),
    "ckks_rehoist": (C.c_int, [_p, _p, _p, _sz]),
    "ckks_ntt64_plan_create": (C.c_int, [_p, C.c_uint64, _sz, C.POINTER(_p)]),
    "ckks_ntt64_plan_destroy": (C.c_int, [_p]),
    "ckks_ntt64_plan_psi": (C.c_int, [_p, C.POINTER(C.c_uint64)]),
    "ckks_ntt64_forward": (C.c_int, [_p, _p, _p, _sz]),
    "ckks_ntt64_inverse": (C.c_int, [_p, _p, _p, _sz]),
    "ckks_ntt_forward": (C.c_int, [_p, _p, _p, _sz, _sz, _sz]),
    "ckks_ntt_inverse": (C.c_int, [_p, _p, _p, _sz, _sz, _sz]),
    "ckks_ntt_forward_noalloc": (C.c_int, [_p, _p, _p, _sz, _sz, _sz]),
    "ckks_ntt_forward_batch": (C.c_int, [_p, _p, _sz]),
    "ckks_ntt_inverse_batch": (C.c_int, [_p, _p, _sz]),
    "ckks_poly_mul": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_poly_mac": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_poly_add": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_poly_sub": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_poly_neg": (C.c_int, [_p, _p, _p, _sz, _sz, _sz]),
    "ckks_automorphism": (C.c_int, [_p, _p, _p, _sz, _sz, _sz, _u64]),
    "ckks_ntt_mul_monomial": (C.c_int, [_p, _p, _p, _sz, _sz, _sz, C.c_int64]),
    "ckks_mul_monomial": (C.c_int, [_p, _p, _p, _sz, _sz, _sz, _u64, C.c_int]),
    "ckks_poly_scalar_mul": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_ntt_galois_permute": (C.c_int, [_p, _p, _p, _sz, _sz, _sz, _u64]),
    "ckks_rescale": (C.c_int, [_p, _p, _p, _sz, _sz, _sz, _sz]),
    "ckks_mod_switch": (C.c_int, [_p, _p, _p, _sz, _sz, _sz, _sz]),
    "ckks_tensor": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz, _sz]),
    "ckks_multiply_plain": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz, _sz, _sz]),
    "ckks_multiply_plain_accumulate_ntt": (C.c_int, [_p, _p, _p, _sz, _p, _sz]),
    "ckks_graph_begin": (C.c_int, [_p]),
    "ckks_graph_end": (C.c_int, [_p, C.POINTER(_p)]),
    "ckks_graph_launch": (C.c_int, [_p]),
    "ckks_graph_destroy": (C.c_int, [_p]),
    "ckks_rns_ksk_create": (C.c_int, [_p, _p, _sz, _p, _p, C.POINTER(_p)]),
    "ckks_hybrid_ksk_create": (C.c_int, [_p, _sz, _p, _sz, _p, _p, C.POINTER(_p)]),
    "ckks_bounded_ksk_create": (C.c_int, [_p, _sz, C.c_uint, _p, _p, C.POINTER(_p)]),
    "ckks_keygen": (C.c_int, [_p, C.c_int, _sz, _p, _sz, C.c_uint, _p, _p, _p, _p, _p, _p]),
    "ckks_encrypt_raw": (C.c_int, [_p, _p, _p, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_encode_coefficients": (C.c_int, [_p, _p, C.c_double, _p, _sz, _sz]),
    "ckks_mod_down": (C.c_int, [_p, _sz, _p, _p, _sz]),
    "ckks_decode_coefficients": (C.c_int, [_p, _p, C.c_double, _p, _sz, _sz]),
    "ckks_server_create": (C.c_int, [_p, _sz, _sz, C.POINTER(_p)]),
    "ckks_server_submit": (C.c_int, [_p, _p, C.POINTER(C.c_uint64)]),
    "ckks_server_step": (C.c_int, [_p, C.POINTER(_sz)]),
    "ckks_server_poll": (C.c_int, [_p, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "ckks_server_drain": (C.c_int, [_p]),
    "ckks_profile_get": (C.c_int, [_p, _p]),
    "ckks_profile_reset": (C.c_int, [_p]),
    "ckks_server_cancel": (C.c_int, [_p, C.c_uint64]),
    "ckks_server_stats": (C.c_int, [_p, _p]),
    "ckks_server_destroy": (C.c_int, [_p]),
    "ckks_decrypt_raw": (C.c_int, [_p, _p, _p, _p, _p, _sz, _sz, _sz, _sz]),
    "ckks_bounded_digit_count": (C.c_int, [_p, _sz, C.c_uint, C.POINTER(_sz)]),
    "ckks_bounded_decompose": (C.c_int, [_p, _sz, C.c_uint, _p, _p, _sz, _sz]),
    "ckks_bounded_digits": (C.c_int, [_p, _sz, C.c_uint, _p, _p, _sz, _sz]),
    "ckks_bounded_digits_automorphism": (C.c_int, [_p, _sz, C.c_uint, _p, _p, _sz, _u64]),
    "ckks_bounded_apply_galois_hoisted": (C.c_int, [_p, _u64, _p, _p, _p, _sz, _sz, _sz]),
    "ckks_bounded_hoist": (C.c_int, [_p, _p, _sz, _sz, C.POINTER(_p)]),
    "ckks_bounded_rotate_hoisted": (C.c_int, [_p, _p, _u64, _p, _p, _sz, _sz]),
    "ckks_bounded_hoist_destroy": (C.c_int, [_p]),
    "ckks_hoist": (C.c_int, [_p, _p, _sz, _sz, C.POINTER(_p)]),
    "ckks_hoist_with_b": (C.c_int, [_p, _p, _sz, _sz, C.POINTER(_p)]),
    "ckks_rotate_hoisted": (C.c_int, [_p, _p, _u64, _p, _p, _sz, _sz]),
    "ckks_ksk_destroy": (C.c_int, [_p]),
    "ckks_ksk_bytes": (C.c_int, [_p, C.POINTER(_sz)]),
    "ckks_key_switch": (C.c_int, [_p, _p, _p, _sz, _sz, _sz]),
    "ckks_apply_galois": (C.c_int, [_p, _u64, _p, _p, _sz, _sz, _sz]),
    "ckks_apply_galois_batch": (C.c_int, [_p, _p, _sz]),
    "ckks_relinearize": (C.c_int, [_p, _p, _p, _sz, _sz, _sz]),
    "ckks_multiply_relinearize_rescale": (C.c_int, [_p, _p, _p, _p, _sz, _sz, _sz]),
}


def header_symbols() -> list[str]:
    """Every function name declared in include/ckks_hip.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ckks_\w+)\s*\(", text, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CkksError(-1, f"{LIB_PATH} is missing: build it with make -C {_HERE}")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        raise CkksError(rc, lib().ckks_last_error().decode())


# --------------------------------------------------------------- tensors
def _torch():
    import torch

    return torch


def to_device(x: np.ndarray, device="cuda"):
    """uint64 numpy array -> int64 torch tensor on the GPU (same bits)."""
    torch = _torch()
    a = np.ascontiguousarray(x, dtype=np.uint64)
    return torch.from_numpy(a.view(np.int64).copy()).to(device)


def to_host(t) -> np.ndarray:
    return t.detach().cpu().numpy().view(np.uint64)


def empty(shape, device="cuda"):
    torch = _torch()
    return torch.empty(shape, dtype=torch.int64, device=device)


def _ptr(t, numel: int | None = None, dtype: str = "int64", name: str = "tensor") -> int:
    """Device address of a contiguous GPU tensor of the given dtype holding at
    least `numel` elements; anything else raises ValueError before the library
    sees the pointer (a host pointer or a short buffer would fault the GPU)."""
    if t is None:
        return None
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != getattr(torch, dtype):
        raise ValueError(f"{name} must be {dtype}, not {str(t.dtype).replace('torch.', '')}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() < numel:
        raise ValueError(f"{name} holds {t.numel()} elements, {numel} needed")
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    return t.data_ptr()


# --------------------------------------------------------------- parameters
def find_ntt_primes(bits: int, degree: int, count: int, skip=()) -> list[int]:
    """find_ntt_prime_below_bits (ring/ntt_prime.rs:67) continued downward."""
    out = np.zeros(max(count, 1), dtype=np.uint64)
    sk = np.ascontiguousarray(list(skip) or [0], dtype=np.uint64)
    found = _sz(0)
    _check(lib().ckks_find_ntt_primes(bits, degree, count, sk.ctypes.data, len(skip),
                                      out.ctypes.data, C.byref(found)))
    return [int(v) for v in out[: found.value]]


def chain_moduli(levels: int, special: int, degree: int = 65536):
    """Moduli of the benchmark chains: one 60-bit base prime, `levels`
    40-bit primes and `special` 60-bit special primes, all q = 1 mod 2N
    (found for N = 2^16, so they also serve every smaller degree).
    Returns (Q moduli, P moduli)."""
    q0 = find_ntt_primes(60, degree, 1)
    qs = find_ntt_primes(40, degree, levels)
    ps = find_ntt_primes(60, degree, special, skip=q0) if special else []
    return q0 + qs, ps


# --------------------------------------------------------------- context
class Context:
    """ckks_context on one GPU; runs on torch's current stream by default so
    library calls and torch ops are ordered."""

    def __init__(self, device: int = 0, stream=None):
        torch = _torch()
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self._stream = stream
        h = _p()
        # handle 0 is the null stream, which the library then uses too
        _check(lib().ckks_context_create(device, _p(stream.cuda_stream or None), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def stream(self):
        return self._stream

    def set_stream(self, stream):
        """ckks_context_set_stream: finish the work queued on the current
        stream, then run on `stream` (a torch stream)."""
        _check(lib().ckks_context_set_stream(self._h, _p(stream.cuda_stream or None)))
        self._stream = stream

    def sync(self):
        _check(lib().ckks_sync(self._h))

    def set_ntt_chunk_bytes(self, nbytes: int):
        """NTT inter-pass buffer per chunk (tuning only; results unchanged)."""
        _check(lib().ckks_set_ntt_chunk_bytes(self._h, int(nbytes)))

    def graph_begin(self):
        """Start capturing library calls on this context's stream."""
        _check(lib().ckks_graph_begin(self._h))

    def graph_end(self) -> "Graph":
        h = _p()
        _check(lib().ckks_graph_end(self._h, C.byref(h)))
        return Graph(self, h)

    BASECONV_AUTO, BASECONV_VALU, BASECONV_MFMA = 0, 1, 2
    NTT_VARIANT_BLOCKED, NTT_VARIANT_TILE8K, NTT_VARIANT_COUNT = 1, 2, 4

    def set_base_conversion(self, mode: int):
        """Hybrid ModUp/ModDown engine: AUTO (the key's tuned engine, VALU
        when untuned), VALU, or MFMA (int8 matrix cores where the key allows)."""
        _check(lib().ckks_set_base_conversion(self._h, int(mode)))

    def set_ntt_variant(self, logn: int, variant: int):
        """Two-pass NTT kernel variant for degree 2^logn (bit set of
        NTT_VARIANT_BLOCKED / NTT_VARIANT_TILE8K); results unchanged."""
        _check(lib().ckks_set_ntt_variant(self._h, int(logn), int(variant)))

    KS_FUSION_AUTO, KS_FUSION_OFF, KS_FUSION_XCD, KS_FUSION_LINEAR, KS_FUSION_INVERSE = -1, 0, 1, 2, 3

    def set_key_switch_fusion(self, mode: int):
        """Key switch: digits' forward row pass fused with the key MAC
        (XCD or LINEAR block order), the MAC inside the first inverse pass
        (INVERSE; also for hoisted rotations), separate kernels (OFF), or
        the key's tuned choice (AUTO, the default)."""
        _check(lib().ckks_set_key_switch_fusion(self._h, int(mode)))

    KS_SPLIT_AUTO, KS_SPLIT_OFF, KS_SPLIT_ON = -1, 0, 1

    def set_key_switch_split(self, mode: int):
        """Key switches of >= 8 ciphertexts split over two streams (ON), not
        (OFF), or the key's tuned choice (AUTO, the default)."""
        _check(lib().ckks_set_key_switch_split(self._h, int(mode)))

    def execution_profile(self):
        """ckks_profile_get: the operations run on this context so far, as a
        ckks.ExecutionProfile (execution/mod.rs:162-230)."""
        from .ckks import ExecutionProfile

        raw = (C.c_uint64 * 12)()
        _check(lib().ckks_profile_get(self._h, C.cast(raw, _p)))
        return ExecutionProfile(*[int(v) for v in raw])

    def reset_execution_profile(self):
        _check(lib().ckks_profile_reset(self._h))

    def ntt_variant(self, logn: int) -> int:
        v = C.c_int()
        _check(lib().ckks_get_ntt_variant(self._h, int(logn), C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "_h", None):
            lib().ckks_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured schedule of library calls (hipGraph), replayed by launch()."""

    def __init__(self, ctx: "Context", h):
        self.ctx = ctx
        self._h = h

    def launch(self):
        _check(lib().ckks_graph_launch(self._h))

    def close(self):
        if getattr(self, "_h", None) and self.ctx.handle:
            lib().ckks_graph_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PreparedRnsNttPlan:
    """PreparedRnsNttPlan (ring/prepared_rns_ntt.rs) on the GPU: one prepared
    negacyclic plan per modulus of the basis."""

    def __init__(self, ctx: Context, moduli, degree: int):
        self.ctx = ctx
        self.moduli = [int(q) for q in moduli]
        self.degree = int(degree)
        arr = np.ascontiguousarray(self.moduli, dtype=np.uint64)
        h = _p()
        _check(lib().ckks_rns_plan_create(ctx.handle, arr.ctypes.data, len(arr), self.degree,
                                          C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def __len__(self):
        return len(self.moduli)

    def psi(self) -> list[int]:
        out = np.zeros(len(self.moduli), dtype=np.uint64)
        _check(lib().ckks_rns_plan_psi(self._h, out.ctypes.data))
        return [int(v) for v in out]

    def _shape(self, t, n_limbs):
        n_limbs = n_limbs or len(self.moduli)
        per = n_limbs * self.degree
        if t.numel() % per != 0 or t.numel() == 0:
            raise ValueError("tensor size must be a positive multiple of limbs x N")
        return t.numel() // per, n_limbs, per

    def _count(self, t, limbs_per_item: int) -> int:
        """Number of items of limbs_per_item x N residues in t (ValueError
        unless t holds a positive whole number of them)."""
        per = int(limbs_per_item) * self.degree
        if per <= 0 or t.numel() == 0 or t.numel() % per != 0:
            raise ValueError(f"tensor size must be a positive multiple of {limbs_per_item} x N")
        return t.numel() // per

    def forward(self, x, out=None, n_limbs=None):
        """Forward negacyclic NTT of every limb of every polynomial in x
        (limb-major [poly][limb][N]); out may be x."""
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        out = x if out is None else out
        _check(lib().ckks_ntt_forward(self._h, _ptr(x, n_polys * per, name="x"),
                                      _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per))
        return out

    def forward_noalloc(self, x, out, n_limbs=None):
        """ckks_ntt_forward_noalloc: forward() without any device allocation
        (fixed per-device scratch; out must be given, and may be x)."""
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        _check(lib().ckks_ntt_forward_noalloc(self._h, _ptr(x, n_polys * per, name="x"),
                                              _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per))
        return out

    def inverse(self, x, out=None, n_limbs=None):
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        out = x if out is None else out
        _check(lib().ckks_ntt_inverse(self._h, _ptr(x, n_polys * per, name="x"),
                                      _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per))
        return out

    # --- pointwise (NttPolynomial / Polynomial arithmetic) ---
    def _ew(self, fn, x, y, out, n_limbs):
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        m = n_polys * per
        _check(fn(self._h, _ptr(x, m, name="x"), _ptr(y, m, name="y"), _ptr(out, m, name="out"),
                  n_polys, n_limbs, per))
        return out

    def pointwise_mul(self, x, y, out=None, n_limbs=None):
        out = empty(x.shape) if out is None else out
        return self._ew(lib().ckks_poly_mul, x, y, out, n_limbs)

    def pointwise_mul_add_assign(self, acc, x, y, n_limbs=None):
        n_polys, n_limbs, per = self._shape(acc, n_limbs)
        m = n_polys * per
        _check(lib().ckks_poly_mac(self._h, _ptr(acc, m, name="acc"), _ptr(x, m, name="x"),
                                   _ptr(y, m, name="y"), n_polys, n_limbs, per))
        return acc

    def add(self, x, y, out=None, n_limbs=None):
        out = empty(x.shape) if out is None else out
        return self._ew(lib().ckks_poly_add, x, y, out, n_limbs)

    def sub(self, x, y, out=None, n_limbs=None):
        out = empty(x.shape) if out is None else out
        return self._ew(lib().ckks_poly_sub, x, y, out, n_limbs)

    def neg(self, x, out=None, n_limbs=None):
        out = empty(x.shape) if out is None else out
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        _check(lib().ckks_poly_neg(self._h, _ptr(x, n_polys * per, name="x"),
                                   _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per))
        return out

    def ntt_mul_monomial_signed(self, x, exponent: int, out=None, n_limbs=None):
        """RnsNttPolynomial::mul_monomial_signed (ring/rns_ntt.rs:215): x * X^e
        in the NTT domain (e signed, reduced rem_euclid 2N)."""
        out = empty(x.shape) if out is None else out
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        _check(lib().ckks_ntt_mul_monomial(self._h, _ptr(x, n_polys * per, name="x"),
                                           _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per,
                                           int(exponent)))
        return out

    def mul_monomial_signed(self, x, exponent: int, sign: int = 1, out=None, n_limbs=None):
        """Polynomial::mul_monomial_signed (ring/polynomial.rs:164) per limb:
        sign * X^exponent * x, coefficient domain."""
        if exponent < 0:
            raise ValueError("the coefficient-domain monomial exponent is unsigned")
        out = empty(x.shape) if out is None else out
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        _check(lib().ckks_mul_monomial(self._h, _ptr(x, n_polys * per, name="x"),
                                       _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per,
                                       int(exponent), int(sign)))
        return out

    def scalar_mul(self, x, scalars, out=None, n_limbs=None):
        """Polynomial::scalar_mul (ring/polynomial.rs:144) per limb: limb i
        times scalars[i] mod q_i (one int for every limb, or one per limb)."""
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        if isinstance(scalars, int):
            scalars = [scalars] * n_limbs
        sc = np.ascontiguousarray([int(v) % (1 << 64) for v in scalars], dtype=np.uint64)
        if len(sc) != n_limbs:
            raise ValueError("one scalar per limb")
        out = empty(x.shape) if out is None else out
        _check(lib().ckks_poly_scalar_mul(self._h, _ptr(x, n_polys * per, name="x"), sc.ctypes.data,
                                          _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per))
        return out

    def apply_automorphism(self, x, exponent: int, out=None, n_limbs=None):
        """apply_automorphism (ckks/automorphism.rs:8) per limb, coefficient domain."""
        out = empty(x.shape) if out is None else out
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        _check(lib().ckks_automorphism(self._h, _ptr(x, n_polys * per, name="x"),
                                       _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per, exponent))
        return out

    def permute_negacyclic_ntt_galois(self, x, exponent: int, out=None, n_limbs=None):
        """permute_negacyclic_ntt_galois (ring/ntt_galois_permutation.rs:5)."""
        out = empty(x.shape) if out is None else out
        n_polys, n_limbs, per = self._shape(x, n_limbs)
        _check(lib().ckks_ntt_galois_permute(self._h, _ptr(x, n_polys * per, name="x"),
                                             _ptr(out, n_polys * per, name="out"), n_polys, n_limbs, per,
                                             exponent))
        return out

    # --- ciphertext level ---
    def rescale_to_next(self, ct, n_limbs: int, out=None):
        """rescale_rns_ckks_to_next (ckks/rns_rescale.rs:19) on a batch of
        ciphertexts [ct][2][n_limbs][N] -> [ct][2][n_limbs-1][N]."""
        n = self.degree
        n_cts = self._count(ct, 2 * n_limbs)
        out = empty((n_cts, 2, n_limbs - 1, n)) if out is None else out
        _check(lib().ckks_rescale(self._h, _ptr(ct, n_cts * 2 * n_limbs * n, name="ct"),
                                  _ptr(out, n_cts * 2 * (n_limbs - 1) * n, name="out"), n_cts, n_limbs,
                                  2 * n_limbs * n, 2 * (n_limbs - 1) * n))
        return out

    def mod_switch_to_next(self, ct, n_limbs: int, out=None):
        n = self.degree
        n_cts = self._count(ct, 2 * n_limbs)
        out = empty((n_cts, 2, n_limbs - 1, n)) if out is None else out
        _check(lib().ckks_mod_switch(self._h, _ptr(ct, n_cts * 2 * n_limbs * n, name="ct"),
                                     _ptr(out, n_cts * 2 * (n_limbs - 1) * n, name="out"), n_cts, n_limbs,
                                     2 * n_limbs * n, 2 * (n_limbs - 1) * n))
        return out

    def multiply_plain(self, ct, pt_ntt, n_limbs: int, out=None):
        """multiply_plain_rns_ckks_with_prepared_ntt (ckks/rns_plain_multiply.rs:193):
        ct [n][2][n_limbs][N] (coefficient domain) times NTT-domain plaintexts
        pt_ntt ([n_limbs][N] shared, or [n][n_limbs][N] one per ciphertext)."""
        n = self.degree
        n_cts = self._count(ct, 2 * n_limbs)
        per_ct = pt_ntt.numel() == n_cts * n_limbs * n and n_cts > 1
        out = empty((n_cts, 2, n_limbs, n)) if out is None else out
        _check(lib().ckks_multiply_plain(self._h, _ptr(ct, n_cts * 2 * n_limbs * n, name="ct"),
                                         _ptr(pt_ntt, (n_cts if per_ct else 1) * n_limbs * n, name="pt_ntt"),
                                         _ptr(out, n_cts * 2 * n_limbs * n, name="out"), n_cts, n_limbs,
                                         2 * n_limbs * n, n_limbs * n if per_ct else 0,
                                         2 * n_limbs * n))
        return out

    def multiply_plain_accumulate_ntt(self, cts, pts, n_limbs: int, out=None):
        """ckks_multiply_plain_accumulate_ntt: sum_t cts[t] * pts[t] over
        NTT-domain ciphertexts [2][n_limbs][N] and plaintexts [n_limbs][N]."""
        n = self.degree
        if not cts or len(cts) != len(pts):
            raise ValueError("one plaintext per ciphertext, at least one term")
        out = empty((2, n_limbs, n)) if out is None else out
        ca = (C.c_uint64 * len(cts))(*[_ptr(c, 2 * n_limbs * n, name="ct") for c in cts])
        pa = (C.c_uint64 * len(pts))(*[_ptr(p, n_limbs * n, name="pt") for p in pts])
        _check(lib().ckks_multiply_plain_accumulate_ntt(self._h, C.cast(ca, _p), C.cast(pa, _p), len(cts),
                                                        _ptr(out, 2 * n_limbs * n, name="out"), int(n_limbs)))
        return out

    def tensor(self, ct0, ct1, n_limbs: int, out=None):
        """rns_tensor_with_prepared_ntt (grafting/rns_key_switch.rs:1724)."""
        n = self.degree
        n_cts = self._count(ct0, 2 * n_limbs)
        out = empty((n_cts, 3, n_limbs, n)) if out is None else out
        m = n_cts * 2 * n_limbs * n
        _check(lib().ckks_tensor(self._h, _ptr(ct0, m, name="ct0"), _ptr(ct1, m, name="ct1"),
                                 _ptr(out, n_cts * 3 * n_limbs * n, name="out"), n_cts, n_limbs,
                                 2 * n_limbs * n, 3 * n_limbs * n))
        return out

    def mod_down(self, x, q_limbs: int, out=None):
        """auxiliary_mod_down_rns_candidate (ring/auxiliary_modulus.rs:89): x
        [p][plan limbs][N] (Q then P) -> [p][q_limbs][N]."""
        n = self.degree
        polys = self._count(x, len(self.moduli))
        out = empty((polys, q_limbs, n)) if out is None else out
        _check(lib().ckks_mod_down(self._h, int(q_limbs), _ptr(x, polys * len(self.moduli) * n, name="x"),
                                   _ptr(out, polys * int(q_limbs) * n, name="out"), polys))
        return out

    def encrypt_raw(self, message, error, secret, a, out=None, n_limbs=None):
        """b = m + e - a*s (grafting/rns_key_switch.rs:245-383) for messages
        [p][limbs][N], errors [p][N] (int64), secret [N] (int8), uniform a
        [p][limbs][N]; returns b."""
        n_polys, n_limbs, per = self._shape(message, n_limbs)
        out = empty(message.shape) if out is None else out
        m, n = n_polys * per, self.degree
        _check(lib().ckks_encrypt_raw(self._h, _ptr(message, m, name="message"),
                                      _ptr(error, n_polys * n, name="error"),
                                      _ptr(secret, n, dtype="int8", name="secret"), _ptr(a, m, name="a"),
                                      _ptr(out, m, name="out"),
                                      n_polys, n_limbs, per))
        return out

    def decrypt_raw(self, b, a, secret, out=None, n_limbs=None):
        """b + a*s (grafting/rns_key_switch.rs:906-947)."""
        n_polys, n_limbs, per = self._shape(b, n_limbs)
        out = empty(b.shape) if out is None else out
        m = n_polys * per
        _check(lib().ckks_decrypt_raw(self._h, _ptr(b, m, name="b"), _ptr(a, m, name="a"),
                                      _ptr(secret, self.degree, dtype="int8", name="secret"),
                                      _ptr(out, m, name="out"), n_polys, n_limbs,
                                      per, per))
        return out

    def autotune_ntt(self, n_polys: int, n_limbs: int | None = None):
        """Times every NTT kernel variant on n_polys x n_limbs limbs, keeps
        the fastest for this degree in the context; returns (best, ms list)."""
        n_limbs = len(self.moduli) if n_limbs is None else n_limbs
        best = C.c_int()
        ms = (C.c_float * 4)()
        _check(lib().ckks_autotune_ntt(self._h, int(n_polys), int(n_limbs), C.byref(best), ms))
        return best.value, list(ms)

    def close(self):
        # the context releases its plans itself when it is closed first
        if getattr(self, "_h", None) and self.ctx.handle:
            lib().ckks_rns_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PreparedNtt64:
    """PreparedNttPlan (ring/prepared_ntt.rs:34, forward :200, inverse :370)
    for one prime of any width up to 64 bits, the reference's 64-bit
    parameter set included (ring/ntt_parameter_sets.rs:11-12).  Residues are
    u64 values carried in int64 tensors ([polys][N], values >= 2^63 as their
    two's-complement bit patterns)."""

    def __init__(self, ctx: Context, q: int, degree: int):
        self.ctx = ctx
        self.q, self.degree = int(q), int(degree)
        h = _p()
        _check(lib().ckks_ntt64_plan_create(ctx.handle, self.q, self.degree, C.byref(h)))
        self._h = h

    @property
    def psi(self) -> int:
        v = C.c_uint64()
        _check(lib().ckks_ntt64_plan_psi(self._h, C.byref(v)))
        return int(v.value)

    def _run(self, fn, x, out):
        if x.numel() == 0 or x.numel() % self.degree:
            raise ValueError("tensor size must be a positive multiple of N")
        n = x.numel() // self.degree
        out = x if out is None else out
        _check(fn(self._h, _ptr(x, x.numel(), name="x"), _ptr(out, x.numel(), name="out"), n))
        return out

    def forward(self, x, out=None):
        """Forward negacyclic NTT of every polynomial in x; out may be x."""
        return self._run(lib().ckks_ntt64_forward, x, out)

    def inverse(self, x, out=None):
        return self._run(lib().ckks_ntt64_inverse, x, out)

    def close(self):
        # the context releases its plans itself when it is closed first
        if getattr(self, "_h", None) and self.ctx.handle:
            lib().ckks_ntt64_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _KeySwitchKey:
    """Device-resident key-switching key; `switch`, `apply_galois`,
    `relinearize` and `multiply_relinearize_rescale` take ciphertext batches
    [ct][2][L][N] (relinearize: [ct][3][L][N]) over the key's Q basis."""

    q_limbs: int

    def _finish_init(self, plan, h):
        self.plan = plan
        self.ctx = plan.ctx
        self._h = h

    @property
    def handle(self):
        return self._h

    def _n_cts(self, t, polys):
        per = polys * self.q_limbs * self.plan.degree
        if t.numel() % per != 0 or t.numel() == 0:
            raise ValueError("ciphertext batch size mismatch")
        return t.numel() // per

    def _ct_ptrs(self, n, ins, out, out_limbs=None):
        """Checked pointers of the input batches [(tensor, polys)] and of out
        ([n][2][out_limbs][N])."""
        N = self.plan.degree
        ptrs = [_ptr(t, n * polys * self.q_limbs * N, name=nm) for nm, t, polys in ins]
        ol = self.q_limbs if out_limbs is None else out_limbs
        return ptrs + [_ptr(out, n * 2 * ol * N, name="out")]

    def switch(self, ct, out=None):
        """rns_key_switch_with_prepared_dynamic_ntt (grafting/rns_key_switch.rs:2526)
        or the dnum hybrid switch."""
        n = self._n_cts(ct, 2)
        out = empty((n, 2, self.q_limbs, self.plan.degree)) if out is None else out
        _check(lib().ckks_key_switch(self._h, *self._ct_ptrs(n, [("ct", ct, 2)], out), n, 0, 0))
        return out

    def apply_galois(self, ct, exponent: int, out=None):
        """apply_rns_galois_automorphism_with_prepared_ntt (ckks/rns_galois_key.rs:439)."""
        n = self._n_cts(ct, 2)
        out = empty((n, 2, self.q_limbs, self.plan.degree)) if out is None else out
        _check(lib().ckks_apply_galois(self._h, exponent, *self._ct_ptrs(n, [("ct", ct, 2)], out), n, 0, 0))
        return out

    def relinearize(self, ct3, out=None):
        """rns_relinearize_with_prepared_dynamic_ntt (grafting/rns_key_switch.rs:2113)."""
        n = self._n_cts(ct3, 3)
        out = empty((n, 2, self.q_limbs, self.plan.degree)) if out is None else out
        _check(lib().ckks_relinearize(self._h, *self._ct_ptrs(n, [("ct3", ct3, 3)], out), n, 0, 0))
        return out

    def multiply_relinearize_rescale(self, ct0, ct1, out=None):
        """multiply_relinearize_rescale_rns_ckks_with_prepared_ntt (ckks/rns_multiply.rs:161)."""
        n = self._n_cts(ct0, 2)
        out = empty((n, 2, self.q_limbs - 1, self.plan.degree)) if out is None else out
        ptrs = self._ct_ptrs(n, [("ct0", ct0, 2), ("ct1", ct1, 2)], out, self.q_limbs - 1)
        _check(lib().ckks_multiply_relinearize_rescale(self._h, *ptrs, n,
                                                       0, 0))
        return out

    def hoist(self, ct, with_b: bool = False) -> "HoistedDigits":
        """ckks_hoist: digits of the a components of ct cut and transformed
        once, for rotate_hoisted, whose residues are those of apply_galois.
        with_b (ckks_hoist_with_b): P NTT(b) is kept too and rotations add
        it inside the key MAC instead of reading sigma_k(b) from their ct
        argument, which must then still hold these ciphertexts."""
        return HoistedDigits(self, ct, with_b)

    def rotate_hoisted(self, hoisted: "HoistedDigits", ct, exponent: int, out=None):
        """ckks_rotate_hoisted: apply_galois(ct, exponent) from the hoisted
        digits, residue for residue (ct: the ciphertexts hoisted, whose b
        components are read; the first call per exponent builds the key's
        permuted copy and correction, kept with the key)."""
        if self._n_cts(ct, 2) != hoisted.n_cts:
            raise ValueError("ct must hold the ciphertexts the digits were hoisted from")
        out = empty((hoisted.n_cts, 2, self.q_limbs, self.plan.degree)) if out is None else out
        _check(lib().ckks_rotate_hoisted(hoisted.handle, self._h, int(exponent),
                                         *self._ct_ptrs(hoisted.n_cts, [("ct", ct, 2)], out), 0, 0))
        return out

    @property
    def nbytes(self) -> int:
        """Device bytes of the key (ckks_ksk_bytes; logical_payload_bytes)."""
        b = _sz()
        _check(lib().ckks_ksk_bytes(self._h, C.byref(b)))
        return b.value

    def autotune(self, n_cts: int):
        """Times the key switch of n_cts ciphertexts for every (base-conversion
        engine, fusion mode) pair and keeps the fastest for the AUTO modes;
        returns (engine, fusion, {(engine, fusion): ms}) over the pairs that apply."""
        be, bf = C.c_int(), C.c_int()
        ms = (C.c_float * 8)()
        _check(lib().ckks_autotune_key_switch(self._h, int(n_cts), C.byref(be), C.byref(bf), ms))
        times = {(e, f): ms[(e - 1) * 4 + f] for e in (1, 2) for f in range(4) if ms[(e - 1) * 4 + f] >= 0}
        return be.value, bf.value, times

    def close(self):
        if getattr(self, "_h", None) and self.ctx.handle:
            lib().ckks_ksk_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sizes(seq):
    return np.ascontiguousarray(list(seq), dtype=np.uintp)


def _key_material(plan, key_b, key_a, entries=None):
    """Pointers to key entries [entries][plan limbs][N]: numpy arrays (copied
    from host memory) or int64 GPU tensors such as keygen() returns (copied
    on the device).  Returns (ptr_b, ptr_a, keep-alive objects, entries)."""
    torch = _torch()
    per = len(plan.moduli) * plan.degree
    if isinstance(key_b, torch.Tensor) or isinstance(key_a, torch.Tensor):
        n = key_b.numel() // per if entries is None else entries
        if n == 0 or key_b.numel() != n * per or key_a.numel() != n * per:
            raise ValueError("key entries must have shape [entries][plan limbs][N]")
        return _ptr(key_b, n * per, name="key_b"), _ptr(key_a, n * per, name="key_a"), (key_b, key_a), n
    kb = np.ascontiguousarray(key_b, dtype=np.uint64)
    ka = np.ascontiguousarray(key_a, dtype=np.uint64)
    if kb.size != ka.size or kb.size % per or kb.size == 0 or (entries is not None and kb.size != entries * per):
        raise ValueError("key entries must have shape [entries][plan limbs][N]")
    return kb.ctypes.data, ka.ctypes.data, (kb, ka), kb.size // per


KEYGEN_RNS_GADGET, KEYGEN_HYBRID, KEYGEN_BOUNDED = 0, 1, 2


def keygen(plan, kind: int, q_limbs: int, source, secret, a, e, block_sizes=(), base_log: int = 0):
    """ckks_keygen: NTT-domain key entries (key_b, key_a), each
    [entries][plan limbs][N], from host-sampled randomness a
    [entries][limbs][N] (uniform canonical) and e [entries][N] (int64), for
    the gadget multiple of `source` ([limbs][N] coefficient domain) under the
    ternary `secret` ([N] int8).  Entry counts: blocks (RNS gadget, hybrid)
    or the bounded digit count."""
    N, B = plan.degree, len(plan.moduli)
    if kind == KEYGEN_BOUNDED:
        entries = bounded_digit_count(plan, q_limbs, base_log)
    else:
        entries = len(block_sizes)
    bs = _sizes(block_sizes or [0])
    kb, ka = empty((entries, B, N)), empty((entries, B, N))
    _check(lib().ckks_keygen(plan.handle, int(kind), int(q_limbs), bs.ctypes.data, len(block_sizes),
                             int(base_log), _ptr(source, B * N, name="source"),
                             _ptr(secret, N, dtype="int8", name="secret"), _ptr(a, entries * B * N, name="a"),
                             _ptr(e, entries * N, name="e"), _ptr(kb), _ptr(ka)))
    return kb, ka


def project_secret(plan, secret):
    """The ternary secret [N] (int8) as residues [plan limbs][N] (project_secret)."""
    torch = _torch()
    s = secret.to(torch.int64)
    q = torch.tensor([int(v) for v in plan.moduli], dtype=torch.int64, device=s.device)[:, None]
    return torch.where(s[None, :] < 0, q + s[None, :], s[None, :].expand(len(plan.moduli), -1)).contiguous()


def galois_source(plan, secret, exponent: int):
    """sigma_k(s) over the plan's limbs: the source of a Galois key."""
    return plan.apply_automorphism(project_secret(plan, secret), int(exponent))


def square_source(plan, secret):
    """s^2 over the plan's limbs: the source of a relinearization key."""
    f = plan.forward(project_secret(plan, secret))
    return plan.inverse(plan.pointwise_mul(f, f))


class RnsKeySwitchKey(_KeySwitchKey):
    """PreparedRnsKeySwitchKey / PreparedRnsMultiplicationKey
    (grafting/rns_key_switch.rs:1930-2430): NTT-domain entries
    key_b/key_a of shape [blocks][L][N] over the plan's full basis."""

    def __init__(self, plan: PreparedRnsNttPlan, block_sizes, key_b, key_a):
        self.q_limbs = len(plan.moduli)
        bs = _sizes(block_sizes)
        pb, pa, _keep, _ = _key_material(plan, key_b, key_a, len(bs))
        h = _p()
        _check(lib().ckks_rns_ksk_create(plan.handle, bs.ctypes.data, len(bs), pb, pa, C.byref(h)))
        self._finish_init(plan, h)

    @classmethod
    def generate(cls, plan: PreparedRnsNttPlan, block_sizes, source, secret, a, e):
        """RnsKeySwitchKey::generate_with_distribution_ntt_rng
        (grafting/rns_key_switch.rs:1155-1258) on the device from host-sampled
        a, e: entry j encrypts E_j * source under secret."""
        kb, ka = keygen(plan, KEYGEN_RNS_GADGET, len(plan.moduli), source, secret, a, e, block_sizes=block_sizes)
        return cls(plan, block_sizes, kb, ka)


class HybridKeySwitchKey(_KeySwitchKey):
    """dnum-block hybrid key over Q u P: entry j encrypts P*E_j*s' (NTT
    domain, [blocks][Q+P limbs][N]); `plan` holds the Q limbs then P."""

    def __init__(self, plan: PreparedRnsNttPlan, q_limbs: int, block_sizes, key_b, key_a):
        self.q_limbs = int(q_limbs)
        bs = _sizes(block_sizes)
        pb, pa, _keep, _ = _key_material(plan, key_b, key_a, len(bs))
        h = _p()
        _check(lib().ckks_hybrid_ksk_create(plan.handle, self.q_limbs, bs.ctypes.data, len(bs), pb, pa,
                                            C.byref(h)))
        self._finish_init(plan, h)

    @classmethod
    def generate(cls, plan: PreparedRnsNttPlan, q_limbs: int, block_sizes, source, secret, a, e):
        """dnum hybrid key on the device from host-sampled a, e: entry j
        encrypts P * E_j * source under secret over Q u P."""
        kb, ka = keygen(plan, KEYGEN_HYBRID, q_limbs, source, secret, a, e, block_sizes=block_sizes)
        return cls(plan, q_limbs, block_sizes, kb, ka)


class BoundedKeySwitchKey(_KeySwitchKey):
    """Bounded balanced base-2^w gadget key (grafting/bounded_gadget.rs):
    without special primes (q_limbs = plan limbs) the bounded multiplication
    key of grafting/bounded_key_switch.rs, with them the hybrid-QP Galois key
    of ckks/hybrid_rns_galois_key.rs.  Entries [digit_count][plan limbs][N],
    NTT domain."""

    def __init__(self, plan: PreparedRnsNttPlan, q_limbs: int, base_log: int, key_b, key_a):
        self.q_limbs = int(q_limbs)
        self.base_log = int(base_log)
        count = bounded_digit_count(plan, self.q_limbs, self.base_log)
        pb, pa, _keep, _ = _key_material(plan, key_b, key_a, count)
        h = _p()
        _check(lib().ckks_bounded_ksk_create(plan.handle, self.q_limbs, self.base_log, pb, pa, C.byref(h)))
        self.digit_count = count
        self._finish_init(plan, h)

    @classmethod
    def generate(cls, plan: PreparedRnsNttPlan, q_limbs: int, base_log: int, source, secret, a, e):
        """PreparedHybridRnsGaloisKey::generate_with_distribution_ntt_rng
        (ckks/hybrid_rns_galois_key.rs:61-170) / BoundedRnsMultiplicationKey
        (grafting/bounded_key_switch.rs:22-118) on the device from host-sampled
        a, e: entry j encrypts P * 2^(w j) * source under secret."""
        kb, ka = keygen(plan, KEYGEN_BOUNDED, q_limbs, source, secret, a, e, base_log=base_log)
        return cls(plan, q_limbs, base_log, kb, ka)

    def digits(self, ct):
        """Balanced digits [ct][digit_count][N] (int64) of the a components of
        a ciphertext batch [ct][2][q_limbs][N], for apply_galois_hoisted."""
        torch = _torch()
        n = self._n_cts(ct, 2)
        poly = self.q_limbs * self.plan.degree
        _ptr(ct, n * 2 * poly, name="ct")
        out = torch.empty((n, self.digit_count, self.plan.degree), dtype=torch.int64, device=ct.device)
        # a component of ciphertext c at element (2 c + 1) * poly
        _check(lib().ckks_bounded_digits(self.plan.handle, self.q_limbs, self.base_log, ct.data_ptr() + 8 * poly,
                                         _ptr(out), n, 2 * poly))
        return out

    def hoist(self, ct) -> HoistedDigits:
        """Decompose, lift and transform the a components of ct once."""
        return HoistedDigits(self, ct)

    def rotate_hoisted(self, hoisted: HoistedDigits, ct, exponent: int, out=None):
        """ckks_bounded_rotate_hoisted: same residues as apply_galois(ct, exponent).
        ct must hold the ciphertexts the digits were hoisted from (their b
        components are read); the first call per exponent builds a permuted
        copy of this key, kept with the key."""
        if self._n_cts(ct, 2) != hoisted.n_cts:
            raise ValueError("ct must hold the ciphertexts the digits were hoisted from")
        if out is not None and self._n_cts(out, 2) != hoisted.n_cts:
            raise ValueError("out must hold as many ciphertexts as were hoisted")
        out = empty((hoisted.n_cts, 2, self.q_limbs, self.plan.degree)) if out is None else out
        _check(lib().ckks_bounded_rotate_hoisted(hoisted.handle, self._h, int(exponent),
                                                 *self._ct_ptrs(hoisted.n_cts, [("ct", ct, 2)], out), 0, 0))
        return out

    def apply_galois_hoisted(self, ct, exponent: int, a_digits, out=None):
        """Rotation from hoisted digits (grafting/bounded_gadget.rs:176-312):
        same residues as apply_galois; a_digits = self.digits(ct)."""
        n = self._n_cts(ct, 2)
        if a_digits.numel() != n * self.digit_count * self.plan.degree:
            raise ValueError("a_digits must hold digit_count x N digits per ciphertext")
        out = empty((n, 2, self.q_limbs, self.plan.degree)) if out is None else out
        p_ct, p_out = self._ct_ptrs(n, [("ct", ct, 2)], out)
        _check(lib().ckks_bounded_apply_galois_hoisted(self._h, int(exponent), p_ct,
                                                       _ptr(a_digits, name="a_digits"), p_out, n, 0, 0))
        return out


class HoistedDigits:
    """ckks_bounded_hoist: the NTT-domain digits of a ciphertext batch's a
    components, decomposed once for many rotations (rotate_hoisted)."""

    def __init__(self, key: "BoundedKeySwitchKey", ct, with_b: bool = False):
        self.ctx = key.ctx
        self.n_cts = key._n_cts(ct, 2)
        h = _p()
        p_ct = _ptr(ct, self.n_cts * 2 * key.q_limbs * key.plan.degree, name="ct")
        fn = lib().ckks_hoist_with_b if with_b else lib().ckks_hoist
        _check(fn(key.handle, p_ct, self.n_cts, 0, C.byref(h)))
        self._h = h

    def rehoist(self, key, ct):
        """ckks_rehoist: the digits of a new batch of as many ciphertexts,
        in place (RNS-gadget / dnum keys); no device allocation."""
        if key._n_cts(ct, 2) != self.n_cts:
            raise ValueError("ct must hold as many ciphertexts as were hoisted")
        p_ct = _ptr(ct, self.n_cts * 2 * key.q_limbs * key.plan.degree, name="ct")
        _check(lib().ckks_rehoist(self._h, key.handle, p_ct, 0))
        return self

    @property
    def handle(self):
        return self._h

    def close(self):
        # after the context is gone (interpreter exit) the handle is left alone
        if getattr(self, "_h", None) and self.ctx.handle:
            lib().ckks_bounded_hoist_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bounded_digit_count(plan: PreparedRnsNttPlan, q_limbs: int, base_log: int) -> int:
    """BoundedGadgetLayout::new digit count over the first q_limbs limbs."""
    c = _sz()
    _check(lib().ckks_bounded_digit_count(plan.handle, int(q_limbs), int(base_log), C.byref(c)))
    return c.value


def bounded_decompose(plan: PreparedRnsNttPlan, q_limbs: int, base_log: int, x, out=None):
    """BoundedGadgetLayout::decompose + lift_digit on the device: x
    [polys][q_limbs][N] -> [polys][digit_count][plan limbs][N]."""
    n = plan.degree
    polys = plan._count(x, q_limbs)
    count = bounded_digit_count(plan, q_limbs, base_log)
    out = empty((polys, count, len(plan.moduli), n)) if out is None else out
    _check(lib().ckks_bounded_decompose(plan.handle, int(q_limbs), int(base_log), _ptr(x, polys * q_limbs * n, name="x"),
                                        _ptr(out, polys * count * len(plan.moduli) * n, name="out"), polys, 0))
    return out


def bounded_digits(plan: PreparedRnsNttPlan, q_limbs: int, base_log: int, x):
    """BoundedGadgetLayout::decompose on the device: x [polys][q_limbs][N]
    -> int64 digits [polys][digit_count][N]."""
    torch = _torch()
    n = plan.degree
    polys = plan._count(x, q_limbs)
    count = bounded_digit_count(plan, q_limbs, base_log)
    out = torch.empty((polys, count, n), dtype=torch.int64, device=x.device)
    _check(lib().ckks_bounded_digits(plan.handle, int(q_limbs), int(base_log), _ptr(x, polys * q_limbs * n, name="x"),
                                     _ptr(out), polys, 0))
    return out


def bounded_digits_automorphism(plan: PreparedRnsNttPlan, q_limbs: int, base_log: int, digits, exponent: int):
    """BoundedGadgetDecomposition::automorphism_hoisted on the device."""
    n = plan.degree
    count = bounded_digit_count(plan, q_limbs, base_log)
    polys = plan._count(digits, count)
    out = _torch().empty_like(digits)
    _check(lib().ckks_bounded_digits_automorphism(plan.handle, int(q_limbs), int(base_log),
                                                  _ptr(digits, polys * count * n, name="digits"),
                                                  _ptr(out), polys, int(exponent)))
    return out


class GaloisBatchEntry(C.Structure):
    _fields_ = [("key", _p), ("exponent", _u64), ("input", _p), ("output", _p), ("n_cts", _sz),
                ("in_stride", _sz), ("out_stride", _sz)]


def apply_galois_batch(ctx: Context, entries):
    """ckks_apply_galois_batch over [(key, exponent, ct, out)]: rotations of
    batches of uneven degree and level in one launch sequence; each out
    equals key.apply_galois(ct, exponent)."""
    arr = (GaloisBatchEntry * max(len(entries), 1))()
    for i, (key, exponent, ct, out) in enumerate(entries):
        n = key._n_cts(ct, 2)
        p_ct, p_out = key._ct_ptrs(n, [("ct", ct, 2)], out)
        arr[i] = GaloisBatchEntry(key.handle, int(exponent), p_ct, p_out, n, 0, 0)
    _check(lib().ckks_apply_galois_batch(ctx.handle, C.cast(arr, _p), len(entries)))


class HoistedRotationEntry(C.Structure):
    _fields_ = [("hoisted", _p), ("key", _p), ("exponent", _u64), ("input", _p), ("output", _p),
                ("in_stride", _sz), ("out_stride", _sz)]


def rotate_hoisted_batch(ctx: Context, entries):
    """ckks_rotate_hoisted_batch over [(hoisted, key, exponent, ct, out)]:
    hoisted rotations of uneven degree in one launch sequence; each out
    equals key.rotate_hoisted(hoisted, ct, exponent), i.e.
    key.apply_galois(ct, exponent)."""
    arr = (HoistedRotationEntry * max(len(entries), 1))()
    for i, (hoisted, key, exponent, ct, out) in enumerate(entries):
        if key._n_cts(ct, 2) != hoisted.n_cts:
            raise ValueError("ct must hold the ciphertexts the digits were hoisted from")
        p_ct, p_out = key._ct_ptrs(hoisted.n_cts, [("ct", ct, 2)], out)
        arr[i] = HoistedRotationEntry(hoisted.handle, key.handle, int(exponent), p_ct, p_out, 0, 0)
    _check(lib().ckks_rotate_hoisted_batch(ctx.handle, C.cast(arr, _p), len(entries)))


class NttBatchEntry(C.Structure):
    _fields_ = [("plan", _p), ("input", _p), ("output", _p), ("n_polys", _sz),
                ("n_limbs", _sz), ("poly_stride", _sz)]


def ntt_batch(ctx: Context, entries, forward: bool = True):
    """One launch per pass over entries [(plan, x, out, n_limbs)] of mixed
    degree and level."""
    arr = (NttBatchEntry * len(entries))()
    for i, (plan, x, out, n_limbs) in enumerate(entries):
        n_polys, n_limbs, per = plan._shape(x, n_limbs)
        arr[i] = NttBatchEntry(plan.handle, _ptr(x, n_polys * per, name="x"),
                               _ptr(out if out is not None else x, n_polys * per, name="out"), n_polys,
                               n_limbs, per)
    fn = lib().ckks_ntt_forward_batch if forward else lib().ckks_ntt_inverse_batch
    _check(fn(ctx.handle, C.cast(arr, _p), len(entries)))


def load():
    """Return this module (for callers that locate the package by path)."""
    import sys

    return sys.modules[__name__]


from . import ckks  # noqa: E402  (level-aware evaluator over the classes above)
from . import serving  # noqa: E402  (continuous-batching serving loop)
from . import encoding  # noqa: E402  (CKKS encode / decode on the device)
from . import linear  # noqa: E402  (rotation-schedule programs of the linear-algebra workloads)
from . import wire  # noqa: E402  (versioned binary layout of RNS data)
