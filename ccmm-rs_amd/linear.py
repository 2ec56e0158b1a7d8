"""Plaintext-diagonal x rotation programs: the linear-algebra callers of the
rotation hot path (the BSGS / radix-4 FFT and GEMV / CCMM schedules of the
reference's eblas layer, eblas/prepared_radix4_fft256.rs, eblas/fft.rs,
and the structured CCMM of bin/structured_ccmm_4096.rs).

Slot vectors have S = N/2 slots and hold data of period d (d | S), repeated
S/d times.  A matrix M acting on the period is applied by its diagonals
diag_i[j] = M[j][(j + i) mod d]:  M x = sum_i diag_i * rot_i(x), rot_i the
left rotation by i slots (Galois exponent 5^i).  Dense matrices use
baby-step / giant-step:  i = g k + b,
    M x = sum_k rot_{g k}( sum_b rot_{-g k}(diag_{g k + b}) * rot_b(x) ),
which takes g - 1 + K - 1 rotations instead of d - 1.  Sparse matrices (FFT
stages, the CCMM masks) use their few diagonals directly.

Builders emit serving.Program ops.  `key_for(limbs, steps)` returns (key,
Galois exponent) for a left rotation by `steps` at a level of `limbs` limbs
(None keys build cleartext-only programs for serving.Program.simulate), and
`encode(values, limbs)` returns the NTT-domain device plaintext of a slot
vector at that level (None likewise).  Every plaintext multiply is followed
by one rescale of the sum, so a transform uses one level.
"""
from __future__ import annotations

import math
from typing import Callable, Optional

import numpy as np

from .serving import Program


def tile(values, period: int, slots: int) -> np.ndarray:
    v = np.asarray(values, dtype=np.complex128)
    assert v.shape == (period,) and slots % period == 0
    return np.tile(v, slots // period)


def matrix_diagonals(M: np.ndarray) -> dict:
    """diag_i[j] = M[j][(j + i) mod d] for the nonzero diagonals of M."""
    d = M.shape[0]
    j = np.arange(d)
    out = {}
    for i in range(d):
        v = M[j, (j + i) % d]
        if np.any(v != 0):
            out[i] = v.astype(np.complex128)
    return out


def permutation_diagonals(cols: np.ndarray) -> dict:
    """Diagonals of the permutation y[j] = x[cols[j]] on a period of len(cols)."""
    d = len(cols)
    off = (np.asarray(cols) - np.arange(d)) % d
    out = {}
    for i in np.unique(off):
        v = np.zeros(d, dtype=np.complex128)
        v[off == i] = 1.0
        out[int(i)] = v
    return out


def compose_diagonals(A: dict, B: dict, d: int) -> dict:
    """Diagonals of A @ B: (AB)_i[j] = sum_a A_a[j] B_{i-a}[(j + a) mod d]."""
    out: dict = {}
    for a, va in A.items():
        for b, vb in B.items():
            i = (a + b) % d
            term = va * np.roll(vb, -a)
            out[i] = out.get(i, 0) + term
    return {i: v for i, v in out.items() if np.any(np.abs(v) > 1e-300)}


class Builder:
    """ntt_resident (default): every plaintext-diagonal inner sum runs in the
    NTT domain -- the rotated copies are transformed once each, one
    multiply-accumulate launch forms sum_i diag_i * rot_i(x), and one
    inverse transform returns it -- instead of a forward transform, a
    product and an inverse transform per diagonal.  Both forms give the same
    residues (exact modular arithmetic, linear transforms)."""

    def __init__(self, slots: int, key_for: Callable, encode: Callable, moduli, ntt_resident: bool = True):
        self.S = slots
        self.key_for = key_for
        self.encode = encode
        self.moduli = list(moduli)
        self.ntt_resident = ntt_resident
        self._zeros = {}

    def _zero_pt(self, limbs):
        if limbs not in self._zeros:
            self._zeros[limbs] = self._pt(np.zeros(self.S, dtype=np.complex128), limbs)
        return self._zeros[limbs]

    def _mac_terms(self, prog: Program, x, y, terms: list, limbs: int):
        """y = sum over (offset, slot values) terms of values * rot_offset(x),
        NTT-resident: each rotated copy is transformed in a scratch register,
        one multiply-accumulate, one inverse transform, then one rescale."""
        regs = prog.scratch(len(terms))
        pts, vals = [], []
        for r, (o, v) in zip(regs, terms):
            if o % self.S == 0:
                prog.ntt(r, x)
            else:
                self._rot(prog, r, x, o, limbs)
                prog.ntt(r, r)
            pts.append(self._pt(v, limbs))
            vals.append(v)
        prog.mac_plain(y, regs[0], pts, values=vals)
        prog.intt(y, y)
        prog.rescale(y, y)

    def _rot(self, prog: Program, dst, src, steps, limbs):
        steps %= self.S
        if steps == 0:
            prog.copy(dst, src)
            return
        key, exponent = self.key_for(limbs, steps)
        prog.rotate(dst, src, key, exponent, steps=steps)

    def _pt(self, values, limbs):
        return self.encode(values, limbs)

    def transform_direct(self, prog: Program, x, y, tmp, diags: dict, period: int, limbs: int):
        """y = M x from M's (few) diagonals, then one rescale (in place on y)."""
        if self.ntt_resident:
            return self._mac_terms(prog, x, y, [(i, tile(v, period, self.S)) for i, v in sorted(diags.items())],
                                   limbs)
        first = True
        for i, v in sorted(diags.items()):
            vals = tile(v, period, self.S)
            if i % period == 0:
                src = x
            else:
                self._rot(prog, tmp, x, i, limbs)
                src = tmp
            if first:
                prog.mul_plain(y, src, self._pt(vals, limbs), values=vals)
                first = False
            else:
                prog.mul_plain(tmp, src, self._pt(vals, limbs), values=vals)
                prog.add(y, y, tmp)
        prog.rescale(y, y)

    def transform_signed(self, prog: Program, x, y, tmp, sdiags: dict, limbs: int):
        """y = M x for an operator given by signed-offset diagonals over the
        whole slot vector, y[s] = sum_o D_o[s] x[s + o], where D_o[s] is zero
        unless s + o stays inside s's own row / block: no cyclic rotation
        couples independent rows (the masks of the reference's packed FFT2
        stages, eblas/fft.rs:2700-2800).  One rescale at the end."""
        if self.ntt_resident:
            return self._mac_terms(prog, x, y, [(o, np.asarray(v, dtype=np.complex128))
                                                for o, v in sorted(sdiags.items())], limbs)
        first = True
        for o, vals in sorted(sdiags.items()):
            vals = np.asarray(vals, dtype=np.complex128)
            assert vals.shape == (self.S,)
            if o == 0:
                src = x
            else:
                self._rot(prog, tmp, x, o, limbs)
                src = tmp
            if first:
                prog.mul_plain(y, src, self._pt(vals, limbs), values=vals)
                first = False
            else:
                prog.mul_plain(tmp, src, self._pt(vals, limbs), values=vals)
                prog.add(y, y, tmp)
        prog.rescale(y, y)

    def transform_bsgs(self, prog: Program, x, y, regs: list, diags: dict, period: int, limbs: int,
                       unit: int = 1, g: Optional[int] = None):
        """y = M x by baby-step / giant-step over diagonal offsets that are
        multiples of `unit`; regs: scratch registers (g + 2 of them)."""
        offs = sorted(diags)
        assert all(i % unit == 0 for i in offs)
        span = max(offs) // unit + 1
        g = g or max(1, int(math.ceil(math.sqrt(span))))
        assert len(regs) >= g + 2
        baby, acc, tmp = regs[:g], regs[g], regs[g + 1]
        used_b = sorted({(i // unit) % g for i in offs})
        if self.ntt_resident:
            assert list(baby) == list(range(baby[0], baby[0] + g)), "baby-step registers must be consecutive"
            used_b = list(range(used_b[0], used_b[-1] + 1))  # a multiply-accumulate reads a register range
        for b in used_b:
            self._rot(prog, baby[b], x, b * unit, limbs)
            if self.ntt_resident:
                prog.ntt(baby[b], baby[b])
        first_y = True
        for k in range((span + g - 1) // g):
            terms = [(b, diags[(g * k + b) * unit]) for b in range(g) if (g * k + b) * unit in diags]
            if not terms:
                continue
            if self.ntt_resident:
                lo, hi = terms[0][0], terms[-1][0]
                have = dict(terms)
                pts, vals = [], []
                for b in range(lo, hi + 1):
                    if b in have:
                        v = np.roll(tile(have[b], period, self.S), g * k * unit)
                        pts.append(self._pt(v, limbs))
                    else:
                        v = np.zeros(self.S, dtype=np.complex128)
                        pts.append(self._zero_pt(limbs))
                    vals.append(v)
                prog.mac_plain(acc, baby[lo], pts, values=vals)
                prog.intt(acc, acc)
                terms = []
            for t, (b, v) in enumerate(terms):
                # rot_{-gk}(diag) so the giant rotation lands it on diag
                vals = np.roll(tile(v, period, self.S), g * k * unit)
                if t == 0:
                    prog.mul_plain(acc, baby[b], self._pt(vals, limbs), values=vals)
                else:
                    prog.mul_plain(tmp, baby[b], self._pt(vals, limbs), values=vals)
                    prog.add(acc, acc, tmp)
            if k:
                self._rot(prog, tmp, acc, g * k * unit, limbs)
                src = tmp
            else:
                src = acc
            if first_y:
                prog.copy(y, src)
                first_y = False
            else:
                prog.add(y, y, src)
        prog.rescale(y, y)


# ---------------------------------------------------------------- workloads

def gemv_program(b: Builder, M: np.ndarray, limbs: int) -> Program:
    """Public d x d matrix times an encrypted d-periodic vector (Halevi-Shoup
    diagonals, BSGS)."""
    d = M.shape[0]
    diags = matrix_diagonals(M)
    span = max(diags) + 1
    g = max(1, int(math.ceil(math.sqrt(span))))
    prog = Program(n_regs=g + 4, out_reg=1)
    b.transform_bsgs(prog, 0, 1, list(range(2, g + 4)), diags, d, limbs, g=g)
    return prog


def bitrev_permutation(d: int) -> np.ndarray:
    bits = d.bit_length() - 1
    return np.array([int(format(i, f"0{bits}b")[::-1], 2) if bits else 0 for i in range(d)])


def dft_stage_diagonals(m: int, d: int) -> dict:
    """Radix-2 decimation-in-time stage of span m on a period of d:
    out[s+p] = in[s+p] + w^p in[s+p+m], out[s+p+m] = in[s+p] - w^p in[s+p+m],
    w = exp(-2 pi i / 2m) (numpy's forward sign)."""
    j = np.arange(d)
    p = j % (2 * m)
    low = p < m
    w = np.exp(-2j * np.pi * (np.where(low, p, p - m)) / (2 * m))
    d0 = np.where(low, 1.0, -w)
    dm = np.where(low, w, 0.0)
    dneg = np.where(low, 0.0, 1.0)
    out = {0: d0.astype(np.complex128)}
    if m % d:
        out[m % d] = dm.astype(np.complex128)
        out[(d - m) % d] = out.get((d - m) % d, 0) + dneg.astype(np.complex128)
    return out


def fft_stages(d: int, radix: int = 4) -> list:
    """Diagonal sets of the DFT of length d as a product of radix-4 stages
    (pairs of radix-2 stages; a final radix-2 stage when log2 d is odd).
    Input in bit-reversed order, output in natural order."""
    spans = [1 << s for s in range(d.bit_length() - 1)]
    stages = []
    i = 0
    while i < len(spans):
        st = dft_stage_diagonals(spans[i], d)
        if radix == 4 and i + 1 < len(spans):
            st = compose_diagonals(dft_stage_diagonals(spans[i + 1], d), st, d)
            i += 2
        else:
            i += 1
        stages.append(st)
    return stages


def fft_program(b: Builder, d: int, limbs: int) -> Program:
    """Encrypted length-d DFT (input slots in bit-reversed order, as the
    caller encodes them): one level per radix-4 stage."""
    stages = fft_stages(d)
    prog = Program(n_regs=3, out_reg=0)
    for t, st in enumerate(stages):
        b.transform_direct(prog, 0, 1, 2, st, d, limbs - t)
        prog.copy(0, 1)
    return prog


def dif_stage_matrix(d: int, m: int, sign: float = -1.0) -> np.ndarray:
    """Radix-2 decimation-in-frequency stage of half-span m on length d
    (PackedFft2DifStageDiagonals, eblas/fft.rs:2728-2800): in each block of
    2m, out[p] = in[p] + in[p+m], out[p+m] = (in[p] - in[p+m]) w^p,
    w = exp(sign 2 pi i / 2m)."""
    M = np.zeros((d, d), dtype=np.complex128)
    for j in range(d):
        p = j % (2 * m)
        if p < m:
            M[j, j] = 1.0
            M[j, j + m] = 1.0
        else:
            w = np.exp(sign * 2j * np.pi * (p - m) / (2 * m))
            M[j, j - m] = w
            M[j, j] = -w
    return M


def dit_stage_matrix(d: int, m: int) -> np.ndarray:
    """Radix-2 decimation-in-time stage of span m on length d (dense form of
    dft_stage_diagonals, forward sign)."""
    M = np.zeros((d, d), dtype=np.complex128)
    for j in range(d):
        p = j % (2 * m)
        if p < m:
            w = np.exp(-2j * np.pi * p / (2 * m))
            M[j, j] = 1.0
            M[j, j + m] = w
        else:
            w = np.exp(-2j * np.pi * (p - m) / (2 * m))
            M[j, j - m] = 1.0
            M[j, j] = -w
    return M


def radix4_matrices(stages: list) -> list:
    """Pairs of consecutive radix-2 stage matrices multiplied into radix-4
    stages (a final radix-2 stage when the count is odd)."""
    out = []
    for i in range(0, len(stages), 2):
        out.append(stages[i + 1] @ stages[i] if i + 1 < len(stages) else stages[i])
    return out


def signed_diagonals(M: np.ndarray) -> dict:
    """{o: D_o} with D_o[j] = M[j][j + o] (zero where j + o leaves [0, d))."""
    d = M.shape[0]
    out = {}
    for o in range(-(d - 1), d):
        j = np.arange(max(0, -o), min(d, d - o))
        v = M[j, j + o]
        if np.any(v != 0):
            full = np.zeros(d, dtype=np.complex128)
            full[j] = v
            out[o] = full
    return out


def fft2_programs(b: Builder, rows: int, cols: int, limbs: int):
    """Encrypted 2-D DFT of a rows x cols real or complex image held as two
    ciphertexts of rows/2 x cols slots each (top and bottom halves,
    row-major, columns in bit-reversed order; rows/2 * cols = the slot count).
    The column DFT starts with the decimation-in-frequency butterfly across
    the halves (u = top + bottom, v = (top - bottom) w^r, w = exp(-2 pi i /
    rows)), which needs no rotation; the twiddle w^r of v is folded into the
    first row stage's diagonals (it is constant along a row and the row
    stages never mix rows).  Then each half runs the row DFT (radix-4
    decimation-in-time stages, natural-order output) and the column DFT of
    length rows/2 (radix-4 decimation-in-frequency stages, bit-reversed
    output), one level per stage.  Returns (even, odd) programs with inputs
    (top, bottom): their outputs hold the spectrum rows 2 bitrev(j) and
    2 bitrev(j) + 1 at slot row j."""
    half = rows // 2
    S = b.S
    assert half * cols == S and rows & (rows - 1) == 0 and cols & (cols - 1) == 0
    row_stages = radix4_matrices([dit_stage_matrix(cols, 1 << t) for t in range(cols.bit_length() - 1)])
    col_stages = radix4_matrices([dif_stage_matrix(half, half >> (t + 1)) for t in range(half.bit_length() - 1)])
    slot = np.arange(S)
    r_of, c_of = slot // cols, slot % cols

    def row_sdiags(M, twist=None):
        out = {}
        for o, v in signed_diagonals(M).items():
            full = v[c_of] * (twist if twist is not None else 1.0)
            if np.any(full != 0):
                out[o] = full
        return out

    def col_sdiags(M):
        return {o * cols: v[r_of] for o, v in signed_diagonals(M).items()}

    twist = np.exp(-2j * np.pi * r_of / rows)
    progs = []
    for odd in (False, True):
        prog = Program(n_regs=5, out_reg=2, n_inputs=2)
        (prog.sub if odd else prog.add)(2, 0, 1)
        lv = limbs
        for t, M in enumerate(row_stages):
            b.transform_signed(prog, 2, 3, 4, row_sdiags(M, twist if (odd and t == 0) else None), lv)
            prog.copy(2, 3)
            lv -= 1
        for M in col_stages:
            b.transform_signed(prog, 2, 3, 4, col_sdiags(M), lv)
            prog.copy(2, 3)
            lv -= 1
        progs.append(prog)
    return progs[0], progs[1]


def fft2_levels(rows: int, cols: int) -> int:
    """Levels fft2_programs uses: one per radix-4 (or final radix-2) stage."""
    lr, lc = cols.bit_length() - 1, (rows // 2).bit_length() - 1
    return (lr + 1) // 2 + (lc + 1) // 2


def ccmm_permutations(d: int):
    r, c = np.divmod(np.arange(d * d), d)
    sigma = r * d + (r + c) % d
    tau = ((r + c) % d) * d + c
    phi = [r * d + (c + k) % d for k in range(d)]
    return sigma, tau, phi


def ccmm_program(b: Builder, d: int, limbs: int, relin_key) -> Program:
    """Encrypted d x d matrix product A B (Jiang-Kim-Lauter-Song): A, B
    row-major in registers 0 and 1;
      A B = sum_k phi^k(sigma(A)) * psi^k(tau(B)),
    sigma, tau by BSGS (one level), phi^k by its two diagonals (one level),
    psi^k a rotation by d k then a mod-switch, the products relinearized and
    rescaled (one level): three levels in all."""
    dd = d * d
    sigma, tau, phi = ccmm_permutations(d)
    sd = permutation_diagonals(sigma)
    td = permutation_diagonals(tau)
    g_s = max(1, int(math.ceil(math.sqrt(2 * d - 1))))
    g_t = max(1, int(math.ceil(math.sqrt(d))))
    g = max(g_s, g_t)
    # registers: 0 A, 1 B, 2 sA, 3 tB, 4 Ak, 5 Bk, 6 prod, 7 C, 8 tmp, 9.. scratch
    prog = Program(n_regs=10 + g + 2, out_reg=7, n_inputs=2)
    scratch = list(range(10, 10 + g + 2))
    # sigma's offsets are in (-d, d): rotate A right by d-1 first so they are 0..2d-2
    shifted = {(i + d - 1) % dd: v for i, v in sd.items()}
    b._rot(prog, 8, 0, -(d - 1), limbs)
    b.transform_bsgs(prog, 8, 2, scratch, shifted, dd, limbs, g=g_s)
    b.transform_bsgs(prog, 1, 3, scratch, td, dd, limbs, unit=d, g=g_t)
    lv = limbs - 1
    for k in range(d):
        b.transform_direct(prog, 2, 4, 8, permutation_diagonals(phi[k]), dd, lv)
        b._rot(prog, 5, 3, d * k, lv)
        prog.mod_switch(5, 5)
        if k == 0:
            prog.mul_relin_rescale(7, 4, 5, relin_key)
        else:
            prog.mul_relin_rescale(6, 4, 5, relin_key)
            prog.add(7, 7, 6)
    return prog


def rotation_steps(prog: Program) -> set:
    return {int(m) for (code, *_), m in zip(prog.ops, prog.meta) if code == 1}
