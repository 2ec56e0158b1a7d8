"""ctypes wrapper around the CPU oracle (libckks_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, as the oracle.  The HIP product never
imports this module.

All arrays are numpy ``uint64`` (residues), limb-major ``[limb][N]``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libckks_oracle.so")
_lib = None

u64 = C.c_uint64
sz = C.c_size_t
P = C.c_void_p


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        sig = {
            "orc_add": (u64, [u64, u64, u64]),
            "orc_sub": (u64, [u64, u64, u64]),
            "orc_neg": (u64, [u64, u64]),
            "orc_mul": (u64, [u64, u64, u64]),
            "orc_pow": (u64, [u64, u64, u64]),
            "orc_inv": (u64, [u64, u64]),
            "orc_barrett_mul": (u64, [u64, u64, u64]),
            "orc_barrett_reciprocal_hi": (u64, [u64]),
            "orc_barrett_reciprocal_lo": (u64, [u64]),
            "orc_shoup_quotient": (u64, [u64, u64]),
            "orc_shoup_mul": (u64, [u64, u64, u64, u64]),
            "orc_is_prime": (C.c_int, [u64]),
            "orc_ntt_primes": (sz, [C.c_uint, sz, sz, P, sz, P]),
            "orc_find_negacyclic_root": (u64, [u64, sz]),
            "orc_ntt_forward_reference": (None, [P, P, sz, u64, u64]),
            "orc_ntt_inverse_reference": (None, [P, P, sz, u64, u64]),
            "orc_plan_new": (P, [u64, sz, u64]),
            "orc_plan_free": (None, [P]),
            "orc_plan_psi": (u64, [P]),
            "orc_ntt_forward": (None, [P, P, P]),
            "orc_ntt_inverse": (None, [P, P, P]),
            "orc_rns_ntt_forward": (None, [P, sz, P, P]),
            "orc_rns_ntt_inverse": (None, [P, sz, P, P]),
            "orc_pointwise_mul": (None, [P, P, P, sz, u64]),
            "orc_pointwise_mac": (None, [P, P, P, sz, u64]),
            "orc_poly_add": (None, [P, P, P, sz, u64]),
            "orc_poly_sub": (None, [P, P, P, sz, u64]),
            "orc_negacyclic_mul_naive": (None, [P, P, P, sz, u64]),
            "orc_ntt_mul_monomial": (None, [P, P, P, C.c_int64]),
            "orc_mul_monomial": (None, [P, P, C.c_size_t, C.c_uint64, C.c_uint64, C.c_int]),
            "orc_scalar_mul": (None, [P, P, C.c_size_t, C.c_uint64, C.c_uint64]),
            "orc_permute_ntt_galois": (None, [P, P, sz, u64]),
            "orc_apply_automorphism": (None, [P, P, sz, u64, u64]),
            "orc_rotation_exponent_left": (u64, [sz, sz]),
            "orc_rotation_exponent_right": (u64, [sz, sz]),
            "orc_conjugation_exponent": (u64, [sz]),
            "orc_inverse_automorphism_exponent": (u64, [sz, u64]),
            "orc_rescale_limb": (None, [P, P, P, sz, u64, u64]),
            "orc_rescale_ct": (None, [P, sz, sz, P, P, P, P]),
            "orc_mixed_radix_digits": (None, [P, P, sz, P]),
            "orc_eval_mixed_radix": (u64, [P, P, sz, u64]),
            "orc_extend_basis": (None, [P, P, sz, P, sz, sz, P]),
            "orc_mod_down": (None, [P, P, sz, P, sz, sz, P]),
            "orc_rns_gadget_decompose": (C.c_int, [P, P, sz, P, sz, sz, P]),
            "orc_tensor": (None, [P, sz, P, P, P, P, P, P, P]),
            "orc_rns_key_switch": (C.c_int, [P, P, sz, P, sz, P, P, P, P, P, P]),
            "orc_hybrid_key_switch": (C.c_int, [P, P, sz, sz, P, sz, P, P, P, P, P, P]),
            "orc_bounded_digit_count": (sz, [P, sz, C.c_uint]),
            "orc_bounded_decompose": (C.c_int, [P, P, sz, sz, C.c_uint, P]),
            "orc_lift_signed": (None, [P, sz, C.c_uint64, P]),
            "orc_bounded_key_switch": (C.c_int, [P, P, sz, sz, C.c_uint, P, P, P, P, P, P]),
            "orc_encrypt": (None, [P, P, sz, P, P, P, P, P]),
            "orc_decrypt": (None, [P, P, sz, P, P, P, P]),
            "orc_rng_seed": (None, [P, u64]),
            "orc_rng_next_u32": (C.c_uint32, [P]),
            "orc_rng_next_u64": (u64, [P]),
            "orc_chacha20_block": (None, [P, u64, P]),
            "orc_sample_uniform": (None, [P, u64, sz, P]),
            "orc_sample_gaussian": (None, [P, C.c_double, sz, P]),
            "orc_sample_ternary": (None, [P, sz, P]),
            "orc_pcg32_xsh_rr": (C.c_uint32, [u64]),
            "orc_pcg32_step": (u64, [u64, u64]),
            "orc_rng_set_key": (None, [P, P, u64]),
            "orc_gaussian_table": (sz, [C.c_double, P, sz, C.POINTER(C.c_double)]),
            "orc_uniform_open_unit": (C.c_double, [P]),
            "orc_ternary_census": (None, [P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _u64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64)


def _sizes(seq) -> np.ndarray:
    return np.ascontiguousarray(seq, dtype=np.uintp)


# ---------------------------------------------------------------- scalars
def add(a, b, q): return lib().orc_add(a, b, q)
def sub(a, b, q): return lib().orc_sub(a, b, q)
def mul(a, b, q): return lib().orc_mul(a, b, q)
def pow_(a, e, q): return lib().orc_pow(a, e, q)
def inv(a, q): return lib().orc_inv(a, q)
def barrett_mul(a, b, q): return lib().orc_barrett_mul(a, b, q)
def shoup_quotient(w, q): return lib().orc_shoup_quotient(w, q)
def shoup_mul(x, w, wq, q): return lib().orc_shoup_mul(x, w, wq, q)
def is_prime(v): return bool(lib().orc_is_prime(v))
def find_negacyclic_root(q, n): return lib().orc_find_negacyclic_root(q, n)
def rotation_exponent_left(n, r): return lib().orc_rotation_exponent_left(n, r)
def rotation_exponent_right(n, r): return lib().orc_rotation_exponent_right(n, r)
def conjugation_exponent(n): return lib().orc_conjugation_exponent(n)
def inverse_automorphism_exponent(n, k): return lib().orc_inverse_automorphism_exponent(n, k)


def ntt_primes(bits: int, degree: int, count: int, skip=()) -> list[int]:
    out = np.zeros(count, dtype=np.uint64)
    sk = _u64(list(skip) or [0])
    got = lib().orc_ntt_primes(bits, degree, count, _p(sk), len(skip), _p(out))
    return [int(x) for x in out[:got]]


# ---------------------------------------------------------------- plans
class Plan:
    """PreparedNttPlan for one prime (prepared_ntt.rs)."""

    def __init__(self, q: int, n: int, psi: int = 0):
        self.q, self.n = int(q), int(n)
        self._h = lib().orc_plan_new(self.q, self.n, psi)
        if not self._h:
            raise ValueError(f"invalid NTT plan q={q} n={n}")
        self.psi = lib().orc_plan_psi(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_plan_free(h)
            self._h = None

    def forward(self, x) -> np.ndarray:
        x = _u64(x)
        out = np.empty_like(x)
        lib().orc_ntt_forward(self._h, _p(x), _p(out))
        return out

    def inverse(self, x) -> np.ndarray:
        x = _u64(x)
        out = np.empty_like(x)
        lib().orc_ntt_inverse(self._h, _p(x), _p(out))
        return out


class RnsPlan:
    """PreparedRnsNttPlan: one Plan per limb (prepared_rns_ntt.rs)."""

    def __init__(self, moduli, n: int):
        self.moduli = [int(q) for q in moduli]
        self.n = int(n)
        self.plans = [Plan(q, n) for q in self.moduli]
        self._arr = (C.c_void_p * len(self.plans))(*[p._h for p in self.plans])

    def sub(self, idx) -> "RnsPlan":
        r = RnsPlan.__new__(RnsPlan)
        r.moduli = [self.moduli[i] for i in idx]
        r.n = self.n
        r.plans = [self.plans[i] for i in idx]
        r._arr = (C.c_void_p * len(r.plans))(*[p._h for p in r.plans])
        return r

    def prefix(self, k: int) -> "RnsPlan":
        return self.sub(range(k))

    @property
    def handle(self):
        return C.cast(self._arr, C.c_void_p)

    def mod_arr(self) -> np.ndarray:
        return _u64(self.moduli)

    def forward(self, x) -> np.ndarray:
        x = _u64(x).reshape(-1, self.n)
        assert x.shape[0] == len(self.plans)
        out = np.empty_like(x)
        lib().orc_rns_ntt_forward(self.handle, len(self.plans), _p(x), _p(out))
        return out

    def inverse(self, x) -> np.ndarray:
        x = _u64(x).reshape(-1, self.n)
        assert x.shape[0] == len(self.plans)
        out = np.empty_like(x)
        lib().orc_rns_ntt_inverse(self.handle, len(self.plans), _p(x), _p(out))
        return out


# ---------------------------------------------------------------- limb ops
def ntt_forward_reference(x, q, psi):
    x = _u64(x); out = np.empty_like(x)
    lib().orc_ntt_forward_reference(_p(x), _p(out), x.size, q, psi)
    return out


def ntt_inverse_reference(x, q, psi):
    x = _u64(x); out = np.empty_like(x)
    lib().orc_ntt_inverse_reference(_p(x), _p(out), x.size, q, psi)
    return out


def negacyclic_mul_naive(x, y, q):
    x = _u64(x); y = _u64(y); out = np.empty_like(x)
    lib().orc_negacyclic_mul_naive(_p(x), _p(y), _p(out), x.size, q)
    return out


def pointwise_mul(x, y, q):
    x = _u64(x); y = _u64(y); out = np.empty_like(x)
    lib().orc_pointwise_mul(_p(x), _p(y), _p(out), x.size, q)
    return out


def pointwise_mac(acc, x, y, q):
    acc = _u64(acc).copy(); x = _u64(x); y = _u64(y)
    lib().orc_pointwise_mac(_p(acc), _p(x), _p(y), x.size, q)
    return acc


def permute_ntt_galois(x, k):
    x = _u64(x); out = np.empty_like(x)
    lib().orc_permute_ntt_galois(_p(x), _p(out), x.size, k)
    return out


def apply_automorphism(x, q, k):
    x = _u64(x); out = np.empty_like(x)
    lib().orc_apply_automorphism(_p(x), _p(out), x.size, q, k)
    return out


def mul_monomial(x, q, exponent: int, sign: int = 1):
    """Polynomial::mul_monomial_signed (ring/polynomial.rs:164-200)."""
    x = _u64(x)
    out = np.empty_like(x)
    lib().orc_mul_monomial(_p(x), _p(out), len(x), q, exponent, sign)
    return out


def scalar_mul(x, q, scalar: int):
    """Polynomial::scalar_mul (ring/polynomial.rs:144-155)."""
    x = _u64(x)
    out = np.empty_like(x)
    lib().orc_scalar_mul(_p(x), _p(out), len(x), q, scalar)
    return out


def ntt_mul_monomial(plan: Plan, x, e: int):
    x = _u64(x); out = np.empty_like(x)
    lib().orc_ntt_mul_monomial(plan._h, _p(x), _p(out), e)
    return out


# ---------------------------------------------------------------- RNS / ct ops
def rescale_ct(moduli, b, a):
    m = _u64(moduli); L = len(moduli)
    b = _u64(b); a = _u64(a)
    n = b.size // L
    bo = np.empty((L - 1) * n, dtype=np.uint64); ao = np.empty_like(bo)
    lib().orc_rescale_ct(_p(m), L, n, _p(b), _p(a), _p(bo), _p(ao))
    return bo.reshape(L - 1, n), ao.reshape(L - 1, n)


def extend_basis(x, src_moduli, targets):
    x = _u64(x); sm = _u64(src_moduli); tg = _u64(targets)
    n = x.size // len(src_moduli)
    out = np.empty(len(targets) * n, dtype=np.uint64)
    lib().orc_extend_basis(_p(x), _p(sm), len(src_moduli), _p(tg), len(targets), n, _p(out))
    return out.reshape(len(targets), n)


def mod_down(x, q_moduli, p_moduli):
    x = _u64(x); qm = _u64(q_moduli); pm = _u64(p_moduli)
    n = x.size // (len(q_moduli) + len(p_moduli))
    out = np.empty(len(q_moduli) * n, dtype=np.uint64)
    lib().orc_mod_down(_p(x), _p(qm), len(q_moduli), _p(pm), len(p_moduli), n, _p(out))
    return out.reshape(len(q_moduli), n)


def rns_gadget_decompose(x, moduli, block_sizes):
    x = _u64(x); m = _u64(moduli); bs = _sizes(block_sizes)
    L = len(moduli); n = x.size // L
    out = np.empty(len(block_sizes) * L * n, dtype=np.uint64)
    rc = lib().orc_rns_gadget_decompose(_p(x), _p(m), L, _p(bs), len(block_sizes), n, _p(out))
    if rc:
        raise ValueError(f"rns gadget decompose failed rc={rc}")
    return out.reshape(len(block_sizes), L, n)


def tensor(plan: RnsPlan, b0, a0, b1, a1):
    L = len(plan.plans); n = plan.n
    ins = [_u64(v).reshape(L, n) for v in (b0, a0, b1, a1)]
    outs = [np.empty((L, n), dtype=np.uint64) for _ in range(3)]
    lib().orc_tensor(plan.handle, L, *[_p(v) for v in ins], *[_p(v) for v in outs])
    return tuple(outs)


def rns_key_switch(plan: RnsPlan, block_sizes, b, a, key_b, key_a):
    L = len(plan.plans); n = plan.n
    b = _u64(b); a = _u64(a); kb = _u64(key_b); ka = _u64(key_a)
    bs = _sizes(block_sizes)
    bo = np.empty((L, n), dtype=np.uint64); ao = np.empty_like(bo)
    rc = lib().orc_rns_key_switch(plan.handle, _p(plan.mod_arr()), L, _p(bs), len(block_sizes),
                                  _p(b), _p(a), _p(kb), _p(ka), _p(bo), _p(ao))
    if rc:
        raise ValueError(f"rns key switch failed rc={rc}")
    return bo, ao


def hybrid_key_switch(qp_plan: RnsPlan, q_limbs, block_sizes, b, a, key_b, key_a):
    QP = len(qp_plan.plans); n = qp_plan.n
    b = _u64(b); a = _u64(a); kb = _u64(key_b); ka = _u64(key_a)
    bs = _sizes(block_sizes)
    bo = np.empty((q_limbs, n), dtype=np.uint64); ao = np.empty_like(bo)
    rc = lib().orc_hybrid_key_switch(qp_plan.handle, _p(qp_plan.mod_arr()), q_limbs, QP - q_limbs,
                                     _p(bs), len(block_sizes), _p(b), _p(a), _p(kb), _p(ka),
                                     _p(bo), _p(ao))
    if rc:
        raise ValueError(f"hybrid key switch failed rc={rc}")
    return bo, ao


def bounded_digit_count(moduli, base_log):
    """BoundedGadgetLayout::new digit count (grafting/bounded_gadget.rs:18-37)."""
    m = np.array(moduli, dtype=np.uint64)
    return int(lib().orc_bounded_digit_count(_p(m), len(moduli), base_log))


def bounded_decompose(x, moduli, base_log):
    """Balanced base-2^w digits of the centered coefficients of x (limbs x n):
    BoundedGadgetLayout::decompose (grafting/bounded_gadget.rs:59-122)."""
    x = _u64(x); m = np.array(moduli, dtype=np.uint64)
    n = x.shape[-1]
    out = np.empty((bounded_digit_count(moduli, base_log), n), dtype=np.int64)
    if lib().orc_bounded_decompose(_p(x), _p(m), len(moduli), n, base_log, _p(out)):
        raise ValueError("bounded gadget digit count is insufficient for centered coefficient")
    return out


def bounded_automorphism_hoisted(digits, exponent, base_log):
    """BoundedGadgetDecomposition::automorphism_hoisted
    (grafting/bounded_gadget.rs:273-312): sigma_exponent on decomposed
    balanced digits [D][N]; a negated source's digit vector is renormalised
    from the lowest digit up with a carry.  Python integers, small N only."""
    d = np.asarray(digits, dtype=np.int64)
    count, n = d.shape
    two_n, base = 2 * n, 1 << base_log
    half = base // 2
    e = exponent % two_n
    assert e % 2 == 1
    out = np.zeros_like(d)
    for src in range(n):
        mapped = src * e % two_n
        target, negate = mapped % n, mapped >= n
        if not negate:
            out[:, target] = d[:, src]
            continue
        carry = 0
        for j in range(count):
            value = -int(d[j, src]) + carry
            normalized = (value + half) % base - half  # Python % is rem_euclid for base > 0
            out[j, target] = normalized
            carry = (value - normalized) // base
        assert carry == 0, "hoisted digit negation exceeded gadget capacity"
    return out


def lift_signed(d, q):
    """lift_digit (grafting/bounded_gadget.rs:382-402): rem_euclid q."""
    d = np.ascontiguousarray(d, dtype=np.int64)
    out = np.empty(d.shape, dtype=np.uint64)
    lib().orc_lift_signed(_p(d), d.size, q, _p(out))
    return out


def bounded_key_switch(qp_plan: RnsPlan, q_limbs, base_log, b, a, key_b, key_a):
    """Bounded-digit key switch: no P (bounded_key_switch.rs:139-219) when the
    plan has only Q limbs, else the hybrid-QP form of
    hybrid_rns_galois_key.rs:550-780."""
    QP = len(qp_plan.plans); n = qp_plan.n
    b = _u64(b); a = _u64(a); kb = _u64(key_b); ka = _u64(key_a)
    bo = np.empty((q_limbs, n), dtype=np.uint64); ao = np.empty_like(bo)
    rc = lib().orc_bounded_key_switch(qp_plan.handle, _p(qp_plan.mod_arr()), q_limbs, QP - q_limbs,
                                      base_log, _p(b), _p(a), _p(kb), _p(ka), _p(bo), _p(ao))
    if rc:
        raise ValueError(f"bounded key switch failed rc={rc}")
    return bo, ao


def encrypt(plan: RnsPlan, message, error, secret, a):
    L = len(plan.plans); n = plan.n
    m = _u64(message); e = np.ascontiguousarray(error, dtype=np.int64)
    s = np.ascontiguousarray(secret, dtype=np.int8); a = _u64(a)
    b = np.empty((L, n), dtype=np.uint64)
    lib().orc_encrypt(plan.handle, _p(plan.mod_arr()), L, _p(m), _p(e), _p(s), _p(a), _p(b))
    return b


def decrypt(plan: RnsPlan, b, a, secret):
    L = len(plan.plans); n = plan.n
    b = _u64(b); a = _u64(a); s = np.ascontiguousarray(secret, dtype=np.int8)
    out = np.empty((L, n), dtype=np.uint64)
    lib().orc_decrypt(plan.handle, _p(plan.mod_arr()), L, _p(b), _p(a), _p(s), _p(out))
    return out


# ---------------------------------------------------------------- sampling
class Rng:
    """ChaCha20 stream seeded from a u64 (host sampler)."""

    def __init__(self, seed: int):
        self._buf = (C.c_uint8 * 512)()
        lib().orc_rng_seed(C.cast(self._buf, C.c_void_p), seed)

    @property
    def h(self):
        return C.cast(self._buf, C.c_void_p)

    def next_u32(self): return lib().orc_rng_next_u32(self.h)
    def next_u64(self): return lib().orc_rng_next_u64(self.h)

    def uniform(self, q, n):
        out = np.empty(n, dtype=np.uint64)
        lib().orc_sample_uniform(self.h, q, n, _p(out))
        return out

    def gaussian(self, sigma, n):
        out = np.empty(n, dtype=np.int64)
        lib().orc_sample_gaussian(self.h, sigma, n, _p(out))
        return out

    def ternary(self, n):
        out = np.empty(n, dtype=np.int8)
        lib().orc_sample_ternary(self.h, n, _p(out))
        return out

    def uniform_open_unit(self):
        return lib().orc_uniform_open_unit(self.h)

    @classmethod
    def with_key(cls, key_words, counter=0):
        """A ChaCha20 stream with the given 8-word key from block `counter`."""
        r = cls(0)
        k = np.ascontiguousarray(key_words, dtype=np.uint32)
        lib().orc_rng_set_key(r.h, _p(k), counter)
        return r


def pcg32_xsh_rr(state: int) -> int:
    return lib().orc_pcg32_xsh_rr(state)


def pcg32_step(state: int, inc: int) -> int:
    return lib().orc_pcg32_step(state, inc)


def gaussian_table(sigma: float):
    """(weights, total) of DiscreteGaussianTable::new (rlwe/error.rs:73-88)."""
    total = C.c_double()
    cnt = lib().orc_gaussian_table(sigma, None, 0, C.byref(total))
    w = np.empty(cnt, dtype=np.float64)
    lib().orc_gaussian_table(sigma, _p(w), cnt, C.byref(total))
    return w, total.value


def ternary_census():
    out = np.zeros(4, dtype=np.uint64)
    lib().orc_ternary_census(_p(out))
    return [int(v) for v in out]


def chacha20_block(key_words, counter):
    k = np.ascontiguousarray(key_words, dtype=np.uint32)
    out = np.empty(16, dtype=np.uint32)
    lib().orc_chacha20_block(_p(k), counter, _p(out))
    return out
