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


@pytest.mark.parametrize("logn", [13, 14, 15, 16, 17])
@pytest.mark.parametrize("tb,blk", [(12, 0), (12, 1), (13, 0), (13, 1)])
def test_emulated_galois_read_inverse_matches_oracle(emu, logn, tb, blk):
    """The rotation's first inverse pass, which loads whole source segments
    and places each residue at its Galois-permuted position in LDS, gives
    the oracle's inverse NTT of the oracle's NTT-domain permutation, for
    the identity, the rotation generators, conjugation and odd exponents."""
    if tb == 13 and logn < 14:
        pytest.skip("8192-residue tiles serve N >= 2^14")
    emu.emu_intt_galois.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_int,
                                    C.c_uint32]
    n = 1 << logn
    rng = np.random.default_rng(300 + logn)
    primes = [Q40[2], P60[2]] if logn < 17 else [O.ntt_primes(60, 1 << 17, 1)[0]]
    exps = [1, 5, pow(5, 7, 2 * n), pow(5, n // 4 - 1, 2 * n), 2 * n - 1, 2 * n - 5,
            int(rng.integers(0, n)) * 2 + 1]
    for q in primes:
        plan = O.Plan(q, n)
        x = rng.integers(0, q, n, dtype=np.uint64)
        for k in exps:
            out = np.zeros(n, dtype=np.uint64)
            assert emu.emu_intt_galois(x.ctypes.data, out.ctypes.data, q, logn, tb, blk, k) == 0
            assert (out == plan.inverse(O.permute_ntt_galois(x, k))).all(), (q, k)


def _is_prime(n):
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _prime_below(bound):
    n = bound - 1
    while not _is_prime(n):
        n -= 2 if n % 2 else 1
    return n


# the largest FP64-path prime (< 2^41) and a chain prime; 60-bit source primes
_FP_TARGETS = [_prime_below(1 << 41), Q40[0], Q40[-1]]
_WIDE = [Q0[0], _prime_below(1 << 62)] + list(P60[:2])


def _u64(vals):
    return np.array([int(v) for v in vals], dtype=np.uint64)


def _operand_sets(rng, q, n, lo_bound, hi_bound, trials):
    """Random operands plus the extremes (0, 1, bound - 1) in every slot."""
    for extreme in (0, 1, None):
        if extreme is None:
            yield [hi_bound - 1] * n, q - 1
            yield [lo_bound - 1] + [hi_bound - 1] * (n - 1), q - 1
        else:
            yield [extreme] * n, extreme % q
    for _ in range(trials):
        yield [int(rng.integers(0, hi_bound)) for _ in range(n)], int(rng.integers(0, q))


@pytest.mark.parametrize("q", _FP_TARGETS)
def test_fp64_modup_target_and_garner_exact(emu, q):
    """The FP64 target sum and Garner step of the ModUp kernel equal exact
    integer arithmetic at the operand extremes the kernel admits: digit 0 up
    to 2^62 (split at bit 30), other digits and every weight up to q - 1."""
    emu.emu_fp_modup_target.restype = C.c_uint64
    emu.emu_fp_modup_target.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_uint64]
    emu.emu_fp_garner_step.restype = C.c_uint64
    emu.emu_fp_garner_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_uint64,
                                       C.c_uint64]
    rng = np.random.default_rng(q % 1000)
    for A in (1, 8, 9, 12):
        for wide0 in _WIDE:
            for digits, x in _operand_sets(rng, q, A, wide0, min(q, 1 << 41), 60):
                digits[0] = digits[0] % wide0
                for w in ([q - 1] * A, [int(v) for v in rng.integers(0, q, A)]):
                    w0s = w[0] * (1 << 30) % q
                    cd, wd = _u64(digits), _u64(w)  # kept alive across the calls
                    got = emu.emu_fp_modup_target(cd.ctypes.data, wd.ctypes.data, w0s, A, q)
                    assert got == sum(c * ww for c, ww in zip(digits, w)) % q
                    if A > 1:
                        k = A - 1
                        inv = int(rng.integers(1, q))
                        got = emu.emu_fp_garner_step(x, cd.ctypes.data, wd.ctypes.data, w0s, k, inv, q)
                        want = (x - sum(c * ww for c, ww in zip(digits[:k], w[:k]))) * inv % q
                        assert got == want


@pytest.mark.parametrize("q", _FP_TARGETS)
def test_fp64_moddown_target_exact(emu, q):
    """The FP64 target evaluation of the ModDown kernel equals exact integer
    arithmetic: 7 and 8 special-prime digits up to 2^62, x and the added b
    component up to q - 1, both centering branches and both signs of b."""
    f = emu.emu_fp_moddown_target
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_uint64,
                  C.c_uint64, C.c_int, C.c_uint64]
    rng = np.random.default_rng(q % 997)
    for K in (1, 7, 8):
        for digits, x in _operand_sets(rng, q, K, 1 << 62, 1 << 62, 60):
            for w in ([q - 1] * K, [int(v) for v in rng.integers(0, q, K)]):
                w2 = [ww * (1 << 30) % q for ww in w]
                pmod, pinv = q - 1, int(rng.integers(1, q))
                cd, wd, w2d = _u64(digits), _u64(w), _u64(w2)  # kept alive across the calls
                for neg in (0, 1):
                    for add, add_neg in ((0, 0), (q - 1, 0), (q - 1, 1), (int(rng.integers(0, q)), 1)):
                        got = f(x, cd.ctypes.data, wd.ctypes.data, w2d.ctypes.data, K, neg, pmod, pinv, add,
                                add_neg, q)
                        xp = sum(c * ww for c, ww in zip(digits, w))
                        want = ((x - xp + neg * pmod) * pinv + (-add if add_neg else add)) % q
                        assert got == want


@pytest.mark.parametrize("q", _FP_TARGETS + [_prime_below(1 << 21)])
def test_fp64_moddown_split_target_exact(emu, q):
    """The split-weight FP64 ModDown target (digits below 2^60, used when
    every special prime is) equals exact integer arithmetic: 1 to 8 digits
    at 2^60 - 1 and 0, weights at q - 1 and at the split boundaries, x and
    the added b component up to q - 1, both centering branches, both signs."""
    f = emu.emu_fp_moddown_target_split
    f.restype = C.c_uint64
    f.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_uint64, C.c_uint64,
                  C.c_int, C.c_uint64]
    rng = np.random.default_rng(q % 991)
    edge_w = [w for w in (q - 1, (1 << 20) - 1, 1 << 20, ((1 << 21) - 1) << 20) if w < q]
    for K in range(1, 9):
        for digits, x in _operand_sets(rng, q, K, 1 << 60, 1 << 60, 12):
            for w in [[q - 1] * K, [int(v) for v in rng.integers(0, q, K)],
                      [edge_w[i % len(edge_w)] for i in range(K)]]:
                pmod, pinv = q - 1, int(rng.integers(1, q))
                cd, wd = _u64(digits), _u64(w)  # kept alive across the calls
                for neg in (0, 1):
                    for add, add_neg in ((0, 0), (q - 1, 0), (q - 1, 1)):
                        got = f(x, cd.ctypes.data, wd.ctypes.data, K, neg, pmod, pinv, add, add_neg, q)
                        xp = sum(c * ww for c, ww in zip(digits, w))
                        want = ((x - xp + neg * pmod) * pinv + (-add if add_neg else add)) % q
                        assert got == want, (K, digits, w, x, neg, add, add_neg)


@pytest.mark.parametrize("logn", [12, 16, 17])
def test_emulated_fp64_inverse_extremes(emu, logn):
    """The inverse FP64 butterflies reduce the sum only once per round: pinned
    with the largest NTT prime below 2^41 (the FP64 bound) and inputs at q - 1
    and 0, whose sums grow fastest, against the oracle."""
    emu.emu_ntt_variant.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                    C.c_int, C.c_int]
    n = 1 << logn
    q = O.ntt_primes(41, n, 1)[0]
    assert (1 << 40) < q < (1 << 41)
    plan = O.Plan(q, n)
    rng = np.random.default_rng(logn + 7)
    cases = [np.full(n, q - 1, dtype=np.uint64),
             np.where(np.arange(n) % 2 == 0, q - 1, 0).astype(np.uint64),
             rng.integers(q - 64, q, n, dtype=np.uint64)]
    for y in cases:
        back = np.zeros(n, dtype=np.uint64)
        assert emu.emu_ntt(y.ctypes.data, back.ctypes.data, q, logn, 0) == 0
        assert (back == plan.inverse(y)).all()
        assert emu.emu_ntt_variant(y.ctypes.data, back.ctypes.data, q, logn, 0, 13 if logn >= 14 else 12, 1) == 0
        assert (back == plan.inverse(y)).all()


@pytest.mark.parametrize("logn", [13, 14, 15, 16, 17])
@pytest.mark.parametrize("tb,blk", [(12, 0), (13, 1)])
def test_row_factor_twiddles_at_the_fp64_bound(emu, logn, tb, blk):
    """FP64 row passes build each twiddle from the shared column-pass entry
    and a per-(row, stage) factor; with the largest NTT prime below the FP64
    bound and inputs that grow the lazy ranges fastest, both directions give
    the oracle's residues."""
    if tb == 13 and logn < 14:
        pytest.skip("8192-residue tiles serve N >= 2^14")
    emu.emu_ntt_variant.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                    C.c_int, C.c_int]
    n = 1 << logn
    q = O.ntt_primes(41, n, 1)[0]
    plan = O.Plan(q, n)
    rng = np.random.default_rng(300 + logn)
    for x in (rng.integers(0, q, n, dtype=np.uint64), np.full(n, q - 1, dtype=np.uint64),
              np.where(np.arange(n) % 3 == 0, q - 1, 0).astype(np.uint64)):
        out = np.zeros(n, dtype=np.uint64)
        assert emu.emu_ntt_variant(x.ctypes.data, out.ctypes.data, q, logn, 1, tb, blk) == 0
        assert (out == plan.forward(x)).all()
        assert emu.emu_ntt_variant(x.ctypes.data, out.ctypes.data, q, logn, 0, tb, blk) == 0
        assert (out == plan.inverse(x)).all()


def test_scalar_twiddle_stages_are_wave_uniform(emu):
    """The NTT stages whose twiddle the kernels read through the scalar unit
    (readfirstlane of the index) have one twiddle index per wave64 in every
    pass geometry the kernels instantiate."""
    flagged = C.c_int(0)
    bad = emu.emu_twiddle_uniform(C.byref(flagged))
    assert flagged.value > 0
    assert bad == 0


def test_wide_accumulation_and_reduction(emu):
    """The carry-free 128-bit accumulator and the two-Shoup wide reduction
    that every integer-limb kernel uses agree with exact 128-bit arithmetic
    on random and all-ones operands, for odd moduli up to 2^62."""
    emu.emu_wide_arith_check.restype = C.c_long
    emu.emu_wide_arith_check.argtypes = [C.c_uint64, C.c_long]
    assert emu.emu_wide_arith_check(1, 400000) == 0


@pytest.mark.parametrize("logn", [10, 12, 13, 14, 16, 17])
def test_emulated_wide_modulus_ntt_matches_oracle(emu, logn):
    """Moduli from 2^62 to 2^64 run canonical-residue butterflies (the
    64-bit parameter set's prime where 2N | q - 1, else a 63-bit NTT prime),
    in every kernel variant, against the oracle; inputs at q - 1 included."""
    emu.emu_ntt_variant.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_int,
                                    C.c_int, C.c_int]
    n = 1 << logn
    qs = [O.ntt_primes(63, n, 1)[0]]
    if logn <= 13:
        qs.append(18446744073709436929)
    rng = np.random.default_rng(500 + logn)
    for q in qs:
        assert q >= 1 << 62
        plan = O.Plan(q, n)
        variants = [(12, 0)] if logn <= 12 else [(12, 0), (12, 1)] + ([(13, 1)] if logn >= 14 else [])
        for x in (rng.integers(0, q, n, dtype=np.uint64), np.full(n, q - 1, dtype=np.uint64)):
            for tb, blk in variants:
                out = np.zeros(n, dtype=np.uint64)
                assert emu.emu_ntt_variant(x.ctypes.data, out.ctypes.data, q, logn, 1, tb, blk) == 0
                assert (out == plan.forward(x)).all()
                assert emu.emu_ntt_variant(x.ctypes.data, out.ctypes.data, q, logn, 0, tb, blk) == 0
                assert (out == plan.inverse(x)).all()


def test_root_selection_reproduces_the_64_bit_parameter_set(emu):
    """The plans' psi search gives the reference's 64-bit parameter (and the
    32-bit one): ring/ntt_parameter_sets.rs:8-12."""
    emu.emu_negacyclic_root.restype = C.c_uint64
    emu.emu_negacyclic_root.argtypes = [C.c_uint64, C.c_uint64]
    assert emu.emu_negacyclic_root(18446744073709436929, 4096) == 5975861664659593359
    assert emu.emu_negacyclic_root(4294828033, 4096) == 1953722822
