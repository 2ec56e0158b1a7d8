"""Continuous-batching serving loop (libckks_hip.so ckks_server_*) and the
program format its jobs use.

A program is a list of ops over ciphertext registers [2][limbs][N]
(coefficient domain).  Registers 0 .. n_inputs-1 start as the job's inputs;
`out_reg` is copied to the job's output at the end.  The server validates a
program at submission, admits jobs by register footprint against an HBM
budget, and batches the same op of different jobs into one launch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

from . import PreparedRnsNttPlan, _check, _p, _ptr, _sz, lib

OP_ROTATE, OP_MUL_PLAIN, OP_ADD, OP_RESCALE, OP_MUL_RELIN_RESCALE, OP_COPY, OP_MOD_SWITCH, OP_SUB = 1, 2, 3, 4, 5, 6, 7, 8
OP_NTT, OP_INTT, OP_MAC_PLAIN = 9, 10, 11


class _Op(C.Structure):
    _fields_ = [("code", C.c_uint32), ("dst", C.c_uint32), ("a", C.c_uint32), ("b", C.c_uint32),
                ("key", _p), ("exponent", C.c_uint64), ("plaintext", _p)]


class _Job(C.Structure):
    _fields_ = [("plan", _p), ("n_limbs", C.c_uint32), ("n_regs", C.c_uint32), ("out_reg", C.c_uint32),
                ("n_inputs", C.c_uint32), ("input", _p), ("output", _p), ("ops", _p), ("n_ops", _sz)]


class _Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "rejected", "waiting", "active",
                                         "peak_active", "launches", "steps", "peak_bytes")] + \
               [(n, C.c_double) for n in ("p50_ms", "p99_ms", "mean_ms", "max_ms")] + \
               [(n, C.c_uint64) for n in ("cancelled", "failed", "resident_bytes")]


def _fn(name):
    return getattr(lib(), name)


@dataclass
class Program:
    """Straight-line op list; `keep` holds device plaintexts alive."""
    n_regs: int
    out_reg: int
    n_inputs: int = 1
    ops: list = field(default_factory=list)
    keep: list = field(default_factory=list)

    meta: list = field(default_factory=list)  # per op: slot rotation or plaintext slot values
    _scratch_base: int = -1
    _scratch_n: int = 0

    def scratch(self, k: int) -> list:
        """k scratch registers after the program's own (one shared area,
        grown as needed; consecutive indices)."""
        if self._scratch_base < 0:
            self._scratch_base = self.n_regs
        if k > self._scratch_n:
            self._scratch_n = k
            self.n_regs = self._scratch_base + k
        return list(range(self._scratch_base, self._scratch_base + k))

    def rotate(self, dst, a, key, exponent, steps=None):
        self.ops.append((OP_ROTATE, dst, a, 0, key, int(exponent), None))
        self.meta.append(steps)

    def mul_plain(self, dst, a, pt, values=None):
        if pt is not None:
            self.keep.append(pt)
        self.ops.append((OP_MUL_PLAIN, dst, a, 0, None, 0, pt))
        self.meta.append(values)

    def add(self, dst, a, b):
        self.ops.append((OP_ADD, dst, a, b, None, 0, None))
        self.meta.append(None)

    def ntt(self, dst, a):
        """dst = NTT(a): the register moves to the NTT domain."""
        self.ops.append((OP_NTT, dst, a, 0, None, 0, None))
        self.meta.append(None)

    def intt(self, dst, a):
        self.ops.append((OP_INTT, dst, a, 0, None, 0, None))
        self.meta.append(None)

    def mac_plain(self, dst, first, pts, values=None):
        """dst = sum_t reg[first + t] * pts[t] over NTT-domain registers and
        NTT-domain plaintexts (ckks_multiply_plain_accumulate_ntt): one
        launch for a whole baby-step inner sum."""
        pts = list(pts)
        self.keep.extend(p for p in pts if p is not None)
        self.ops.append((OP_MAC_PLAIN, dst, first, len(pts), None, 0, pts))
        self.meta.append(values)

    def sub(self, dst, a, b):
        self.ops.append((OP_SUB, dst, a, b, None, 0, None))
        self.meta.append(None)

    def rescale(self, dst, a):
        self.ops.append((OP_RESCALE, dst, a, 0, None, 0, None))
        self.meta.append(None)

    def mul_relin_rescale(self, dst, a, b, key):
        self.ops.append((OP_MUL_RELIN_RESCALE, dst, a, b, key, 0, None))
        self.meta.append(None)

    def copy(self, dst, a):
        self.ops.append((OP_COPY, dst, a, 0, None, 0, None))
        self.meta.append(None)

    def mod_switch(self, dst, a):
        self.ops.append((OP_MOD_SWITCH, dst, a, 0, None, 0, None))
        self.meta.append(None)

    def simulate(self, inputs):
        """Runs the program on cleartext slot vectors (FP64): rotations are
        cyclic shifts, plaintext products elementwise, rescale and
        mod-switch identities.  The FP64 reference of a program's result."""
        import numpy as np

        regs = [None] * self.n_regs
        for i, v in enumerate(inputs):
            regs[i] = np.asarray(v, dtype=np.complex128)
        for (code, dst, a, b, _k, _e, _pt), meta in zip(self.ops, self.meta):
            if code == OP_ROTATE:
                regs[dst] = np.roll(regs[a], -int(meta))
            elif code == OP_MUL_PLAIN:
                regs[dst] = regs[a] * meta
            elif code == OP_ADD:
                regs[dst] = regs[a] + regs[b]
            elif code == OP_SUB:
                regs[dst] = regs[a] - regs[b]
            elif code == OP_MAC_PLAIN:
                regs[dst] = sum(regs[a + t] * meta[t] for t in range(b))
            elif code == OP_MUL_RELIN_RESCALE:
                regs[dst] = regs[a] * regs[b]
            else:
                regs[dst] = regs[a].copy()
        return regs[self.out_reg]

    def register_limbs(self, n_limbs: int):
        """Per op, (limbs of its source register a, limbs of the output
        register after the program) as levels move: rescale, mod switch and
        multiply+relinearize+rescale drop one limb; everything else keeps a's.
        The last entry's second field is the output's limb count (the server
        itself rejects programs that read empty registers)."""
        regs = [int(n_limbs) if i < self.n_inputs else 0 for i in range(self.n_regs)]
        out = []
        for code, dst, a, _b, _k, _e, _pt in self.ops:
            la = regs[a] if 0 <= a < self.n_regs else 0
            if 0 <= dst < self.n_regs:
                regs[dst] = la - 1 if code in (OP_RESCALE, OP_MOD_SWITCH, OP_MUL_RELIN_RESCALE) else la
            out.append((la, 0))
        final = regs[self.out_reg] if 0 <= self.out_reg < self.n_regs else 0
        return [(la, final) for la, _ in out] or [(0, final)]

    def footprint(self, n_limbs: int, degree: int) -> int:
        """Bytes the server charges to admit this program alone: its registers
        (n_regs x 2 x n_limbs x N x 8) plus each distinct key and plaintext it
        reads (ckks_server_submit's admission model)."""
        regs = self.n_regs * 2 * int(n_limbs) * int(degree) * 8
        seen, res = set(), 0
        for (code, _d, _a, _b, key, _e, pt), (la, _) in zip(self.ops, self.register_limbs(n_limbs)):
            if key is not None and code in (OP_ROTATE, OP_MUL_RELIN_RESCALE) and id(key) not in seen:
                seen.add(id(key))
                res += key.nbytes
            for p in (pt if code == OP_MAC_PLAIN else [pt] if code == OP_MUL_PLAIN else []):
                if p is not None and id(p) not in seen:
                    seen.add(id(p))
                    res += la * int(degree) * 8
        return regs + res

    def counts(self) -> dict:
        names = {OP_ROTATE: "rotate", OP_MUL_PLAIN: "mul_plain", OP_ADD: "add", OP_SUB: "sub", OP_RESCALE: "rescale",
                 OP_NTT: "ntt", OP_INTT: "intt", OP_MAC_PLAIN: "mac_plain",
                 OP_MUL_RELIN_RESCALE: "mul_relin_rescale", OP_COPY: "copy", OP_MOD_SWITCH: "mod_switch"}
        out: dict = {}
        for op in self.ops:
            out[names[op[0]]] = out.get(names[op[0]], 0) + 1
        return out


class Server:
    WAITING, ADMITTED, DONE, CANCELLED, FAILED = 0, 1, 2, 3, 4

    def __init__(self, ctx, hbm_budget_bytes: int, max_active: int = 0):
        self.ctx = ctx
        h = _p()
        _check(_fn("ckks_server_create")(ctx.handle, int(hbm_budget_bytes), int(max_active), C.byref(h)))
        self._h = h
        self._jobs = {}  # ticket -> (job struct, op array, program, tensors) kept alive until done

    def submit(self, plan: PreparedRnsNttPlan, n_limbs: int, inputs, output, program: Program) -> int:
        ops = (_Op * max(len(program.ops), 1))()
        n = plan.degree
        limbs = program.register_limbs(int(n_limbs))
        tables = []
        for i, (code, dst, a, b, key, exp, pt) in enumerate(program.ops):
            if code == OP_MAC_PLAIN:
                # a host array of the terms' device pointers (copied at submission)
                tab = (C.c_uint64 * len(pt))(*[_ptr(p, limbs[i][0] * n, name="plaintext") for p in pt])
                tables.append(tab)
                ptr = C.cast(tab, _p)
            else:
                ptr = _ptr(pt, limbs[i][0] * n, name="plaintext") if pt is not None else None
            ops[i] = _Op(code, dst, a, b, key.handle if key is not None else None, exp, ptr)
        job = _Job(plan.handle, int(n_limbs), program.n_regs, program.out_reg, program.n_inputs,
                   _ptr(inputs, program.n_inputs * 2 * int(n_limbs) * n, name="inputs"),
                   _ptr(output, 2 * limbs[-1][1] * n, name="output"), C.cast(ops, _p), len(program.ops))
        t = C.c_uint64()
        _check(_fn("ckks_server_submit")(self._h, C.addressof(job), C.byref(t)))
        self._jobs[t.value] = (job, ops, program, inputs, output, tables)
        return t.value

    def step(self) -> int:
        n = _sz()
        _check(_fn("ckks_server_step")(self._h, C.byref(n)))
        return n.value

    def poll(self, ticket: int):
        st, lat = C.c_int(), C.c_double()
        _check(_fn("ckks_server_poll")(self._h, int(ticket), C.byref(st), C.byref(lat)))
        if st.value in (self.DONE, self.CANCELLED, self.FAILED):
            self._jobs.pop(ticket, None)
        return st.value, (lat.value if st.value == self.DONE else None)

    def cancel(self, ticket: int):
        """Drop a waiting or admitted request (ckks_server_cancel)."""
        _check(_fn("ckks_server_cancel")(self._h, int(ticket)))
        self._jobs.pop(ticket, None) if self.poll(ticket)[0] == self.CANCELLED else None

    def drain(self):
        _check(_fn("ckks_server_drain")(self._h))

    def stats(self) -> dict:
        s = _Stats()
        _check(_fn("ckks_server_stats")(self._h, C.addressof(s)))
        return {n: getattr(s, n) for n, _ in _Stats._fields_}

    def close(self):
        # the context destroys its servers itself when it is closed first
        if getattr(self, "_h", None) and self.ctx.handle:
            _fn("ckks_server_destroy")(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ProgramRunner:
    """Runs one Program directly -- one library call per op on registers it
    owns -- outside the serving loop, so that a whole rotation schedule can
    be captured once as a hipGraph (capture()) and replayed (replay())
    without per-op host work.  Registers hold compact [2][limbs][N]
    ciphertexts at their current level (as the server's).  The program's
    keys must belong to the plan's context; each run switches that context
    to the runner's own (non-null, default-priority) stream and back, which
    waits for the run.  The first run() sizes every library workspace, so
    run() must precede capture(); nothing is allocated while capturing."""

    def __init__(self, plan: PreparedRnsNttPlan, n_limbs: int, program: Program, stream=None):
        import torch

        self.ctx, self.plan, self.prog = plan.ctx, plan, program
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self.L = int(n_limbs)
        n = plan.degree
        self.per = 2 * self.L * n
        # two extra registers: the output of ops that cannot run in place
        self.regs = torch.empty((program.n_regs + 2, self.per), dtype=torch.int64, device="cuda")
        self.tmp = program.n_regs
        self.limbs = program.register_limbs(self.L)
        self.graph = None

    def _reg(self, i, limbs):
        return self.regs[i, : 2 * limbs * self.plan.degree]

    def _ct(self, i, limbs):
        return self._reg(i, limbs).view(1, 2, limbs, self.plan.degree)

    def _run_ops(self):
        plan, n = self.plan, self.plan.degree
        regs = [self.L if i < self.prog.n_inputs else 0 for i in range(self.prog.n_regs)]
        for (code, dst, a, b, key, exp, pt) in self.prog.ops:
            la = regs[a]
            out_l = la - 1 if code in (OP_RESCALE, OP_MOD_SWITCH, OP_MUL_RELIN_RESCALE) else la
            # ops whose output may not alias an input write the spare register
            spare = code in (OP_ROTATE, OP_RESCALE, OP_MOD_SWITCH, OP_MUL_RELIN_RESCALE) and \
                (dst == a or (code == OP_MUL_RELIN_RESCALE and dst == b))
            d = self.tmp if spare else dst
            if code == OP_ROTATE:
                key.apply_galois(self._ct(a, la), exp, out=self._ct(d, la))
            elif code == OP_MUL_PLAIN:
                plan.multiply_plain(self._ct(a, la), pt, la, out=self._ct(d, la))
            elif code in (OP_ADD, OP_SUB):
                (plan.add if code == OP_ADD else plan.sub)(self._reg(a, la), self._reg(b, la),
                                                           out=self._reg(d, la), n_limbs=la)
            elif code == OP_RESCALE:
                plan.rescale_to_next(self._ct(a, la), la, out=self._ct(d, out_l))
            elif code == OP_MOD_SWITCH:
                plan.mod_switch_to_next(self._ct(a, la), la, out=self._ct(d, out_l))
            elif code == OP_MUL_RELIN_RESCALE:
                key.multiply_relinearize_rescale(self._ct(a, la), self._ct(b, la), out=self._ct(d, out_l))
            elif code == OP_COPY:
                if d != a:
                    self._reg(d, la).copy_(self._reg(a, la))
            elif code == OP_NTT:
                plan.forward(self._reg(a, la), out=self._reg(d, la), n_limbs=la)
            elif code == OP_INTT:
                plan.inverse(self._reg(a, la), out=self._reg(d, la), n_limbs=la)
            elif code == OP_MAC_PLAIN:
                plan.multiply_plain_accumulate_ntt([self._reg(a + t, la) for t in range(b)], pt, la,
                                                   out=self._reg(d, la))
            else:
                raise ValueError(f"unknown op code {code}")
            if spare:
                self._reg(dst, out_l).copy_(self._reg(self.tmp, out_l))
            regs[dst] = out_l

    def _on_stream(self):
        import contextlib

        import torch

        @contextlib.contextmanager
        def cm():
            prev = self.ctx.stream
            torch.cuda.current_stream().synchronize()
            self.ctx.set_stream(self.stream)
            try:
                with torch.cuda.stream(self.stream):
                    yield
            finally:
                self.ctx.set_stream(prev)  # waits for the runner's stream
        return cm()

    def _load(self, inputs):
        self.regs[: self.prog.n_inputs].copy_(inputs.reshape(self.prog.n_inputs, self.per))

    def output(self):
        lo = self.limbs[-1][1]
        return self._ct(self.prog.out_reg, lo)[0]

    def run(self, inputs=None):
        """Run the program on `inputs` ([n_inputs][2][n_limbs][N]; None: the
        registers as they are); returns the output register (complete)."""
        with self._on_stream():
            if inputs is not None:
                self._load(inputs)
            self._run_ops()
        return self.output()

    def capture(self):
        """Record one run of the ops (not the input load) as a graph."""
        with self._on_stream():
            self.ctx.graph_begin()
            try:
                self._run_ops()
            finally:
                self.graph = self.ctx.graph_end()
        return self.graph

    def replay(self, inputs=None):
        with self._on_stream():
            if inputs is not None:
                self._load(inputs)
            self.graph.launch()
        return self.output()
