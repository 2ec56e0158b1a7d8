"""Serving loop on the GPU: encrypted GEMV, radix-4 FFT and CCMM jobs with
real keys run through ckks_server_*; each decrypted, decoded result matches
its FP64 product / FFT, every result is bit-identical whether the job ran
alone or in a mix of arrivals, and admission respects the HBM budget and the
active cap."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import ckks_pipeline as CP
from conftest import O, P60, Q0, Q40

pytestmark = pytest.mark.gpu

N = 4096
SLOTS = N // 2
Q = [Q0[0]] + Q40[:4]  # 5 limbs: three levels of transforms + margin
P = P60[:1]
SCALE = 2.0 ** 40


class Fixture:
    def __init__(self, pkg, ctx):
        self.pkg, self.ctx = pkg, ctx
        self.rng = O.Rng(2024)
        self.s = self.rng.ternary(N)
        self.plan = pkg.PreparedRnsNttPlan(ctx, Q, N)
        self.emb = pkg.encoding.CanonicalEmbedding(N)
        self.keys = {}
        self.qp_plans = {}
        self.o_plans = {}
        self.relin = {}

    def _qp(self, limbs):
        if limbs not in self.qp_plans:
            self.qp_plans[limbs] = self.pkg.PreparedRnsNttPlan(self.ctx, Q[:limbs] + P, N)
            self.o_plans[limbs] = O.RnsPlan(Q[:limbs] + P, N)
        return self.qp_plans[limbs], self.o_plans[limbs]

    def key_for(self, limbs, steps):
        k = pow(5, steps, 2 * N)
        if (limbs, k) not in self.keys:
            dev, orc = self._qp(limbs)
            kb, ka = CP.keygen_galois_hybrid(orc, limbs, [1] * limbs, self.s, k, self.rng)
            self.keys[(limbs, k)] = self.pkg.HybridKeySwitchKey(dev, limbs, [1] * limbs, kb, ka)
        return self.keys[(limbs, k)], k

    def relin_key(self, limbs):
        if limbs not in self.relin:
            dev, orc = self._qp(limbs)
            kb, ka = CP.keygen_relin_hybrid(orc, limbs, [1] * limbs, self.s, self.rng)
            self.relin[limbs] = self.pkg.HybridKeySwitchKey(dev, limbs, [1] * limbs, kb, ka)
        return self.relin[limbs]

    def encode_pt(self, values, limbs):
        enc = self.pkg.encoding
        pt = enc.encode(self.plan, self.emb, torch.from_numpy(np.asarray(values)), float(Q[limbs - 1]), limbs)
        return self.plan.forward(pt, n_limbs=limbs)

    def builder(self):
        return self.pkg.linear.Builder(SLOTS, self.key_for, self.encode_pt, Q)

    def encrypt(self, slots):
        enc = self.pkg.encoding
        m = enc.encode(self.plan, self.emb, torch.from_numpy(slots), SCALE)
        e = torch.from_numpy(self.rng.gaussian(3.2, N).astype(np.int64)).cuda()
        a = self.pkg.to_device(np.stack([self.rng.uniform(q, N) for q in Q]))
        s = torch.from_numpy(self.s.astype(np.int8)).cuda()
        b = self.plan.encrypt_raw(m, e, s, a)
        return torch.stack([b, a])

    def decrypt(self, ct, limbs, scale):
        s = torch.from_numpy(self.s.astype(np.int8)).cuda()
        ct = ct.view(2, limbs, N)
        m = self.plan.decrypt_raw(ct[0].contiguous(), ct[1].contiguous(), s, n_limbs=limbs)
        return self.pkg.encoding.decode(self.plan, self.emb, m, scale, limbs)[0].cpu().numpy()


@pytest.fixture(scope="module")
def fx(pkg, ctx):
    return Fixture(pkg, ctx)


def _jobs(fx):
    """(name, program, inputs, out_limbs, out_scale, expected slots)."""
    lin = fx.pkg.linear
    b = fx.builder()
    rng = np.random.default_rng(9)
    L = len(Q)
    jobs = []
    d = 32
    M = rng.uniform(-1, 1, (d, d))
    x = rng.uniform(-1, 1, d)
    jobs.append(("gemv", lin.gemv_program(b, M, L), [lin.tile(x, d, SLOTS)], L - 1, SCALE,
                 lin.tile(M @ x, d, SLOTS)))
    d = 64
    z = rng.uniform(-1, 1, d) + 1j * rng.uniform(-1, 1, d)
    perm = lin.bitrev_permutation(d)
    jobs.append(("fft", lin.fft_program(b, d, L), [lin.tile(z[perm], d, SLOTS)], L - 3, SCALE,
                 lin.tile(np.fft.fft(z), d, SLOTS)))
    d = 16
    A = rng.uniform(-1, 1, (d, d))
    B = rng.uniform(-1, 1, (d, d))
    relin = fx.relin_key(L - 2)
    jobs.append(("ccmm", lin.ccmm_program(b, d, L, relin), [lin.tile(A.ravel(), d * d, SLOTS),
                                                         lin.tile(B.ravel(), d * d, SLOTS)],
                 L - 3, SCALE * SCALE / Q[L - 3], lin.tile((A @ B).ravel(), d * d, SLOTS)))
    return jobs


def _submit(fx, server, job):
    name, prog, inputs, out_limbs, _, _ = job
    cts = torch.stack([fx.encrypt(v) for v in inputs]).contiguous()
    out = torch.zeros((2, out_limbs, N), dtype=torch.int64, device="cuda")
    t = server.submit(fx.plan, len(Q), cts, out, prog)
    return t, cts, out


def test_workloads_decrypt_to_fp64(pkg, ctx, fx):
    server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 30)
    jobs = _jobs(fx)
    subs = [_submit(fx, server, j) for j in jobs]
    server.drain()
    for job, (t, _, out) in zip(jobs, subs):
        name, prog, _, out_limbs, scale, want = job
        assert server.poll(t)[0] == server.DONE
        got = fx.decrypt(out, out_limbs, scale)
        err = np.max(np.abs(got - want))
        tol = {"gemv": 1e-4, "fft": 1e-3, "ccmm": 1e-3}[name]
        assert err < tol, (name, err)
        # the FP64 program itself equals the FP64 product
        assert np.allclose(prog.simulate(job[2]), want, atol=1e-9)
    st = server.stats()
    assert st["completed"] == 3 and st["rejected"] == 0
    server.close()


def test_alone_and_in_mix_bit_identical(pkg, ctx, fx):
    jobs = _jobs(fx)
    torch.manual_seed(0)
    inputs = {}
    alone = {}
    for idx, job in enumerate(jobs):
        server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 30)
        t, cts, out = _submit(fx, server, job)
        inputs[idx] = cts
        server.drain()
        alone[idx] = out.clone()
        server.close()
    # mix: arrivals interleaved with steps, duplicates of every job, budget
    # for about three jobs at once
    # budget for about three jobs at once: registers plus the keys and
    # diagonals each reads (shared ones are charged once)
    foot = max(j[1].footprint(len(Q), N) for j in jobs)
    server = pkg.serving.Server(ctx, hbm_budget_bytes=3 * foot, max_active=3)
    order = np.random.default_rng(5).permutation([0, 1, 2] * 3)
    pending = list(order)
    outs = []
    while pending or server.stats()["active"] or server.stats()["waiting"]:
        if pending:
            idx = int(pending.pop(0))
            prog = jobs[idx][1]
            out = torch.zeros_like(alone[idx])
            t = server.submit(fx.plan, len(Q), inputs[idx], out, prog)
            outs.append((idx, t, out))
        server.step()
        server.step()
    server.drain()
    st = server.stats()
    assert st["completed"] == len(order)
    assert st["peak_active"] <= 3 and st["peak_bytes"] <= 3 * foot
    assert st["launches"] < sum(len(jobs[i][1].ops) for i in order)  # ops were batched across jobs
    for idx, t, out in outs:
        assert server.poll(t)[0] == server.DONE
        assert torch.equal(out, alone[idx]), f"job {idx} differs in the mix"
    assert st["p99_ms"] >= st["p50_ms"] > 0
    server.close()


def test_rejections(pkg, ctx, fx):
    server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 16)
    prog = pkg.serving.Program(n_regs=2, out_reg=1)
    prog.copy(1, 0)
    cts = torch.zeros((2, len(Q), N), dtype=torch.int64, device="cuda")
    out = torch.zeros_like(cts)
    with pytest.raises(pkg.CkksError, match="footprint exceeds"):
        server.submit(fx.plan, len(Q), cts, out, prog)
    bad = pkg.serving.Program(n_regs=2, out_reg=1)
    bad.add(1, 0, 1)
    server.close()
    server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 30)
    with pytest.raises(pkg.CkksError, match="addition needs two ciphertexts at one level"):
        server.submit(fx.plan, len(Q), cts, out, bad)
    assert server.stats()["rejected"] == 1
    server.close()


def test_cancel_waiting_and_active_requests(pkg, ctx, fx):
    """A cancelled waiting request never runs; a cancelled active one stops
    and gives its registers and resident key charges back; the request that
    remains produces the residues it produces alone."""
    jobs = _jobs(fx)
    gemv = jobs[0]
    server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 30)
    t_alone, cts, out_alone = _submit(fx, server, gemv)
    server.drain()
    ref = out_alone.clone()
    server.close()
    foot = gemv[1].n_regs * 2 * len(Q) * N * 8
    server = pkg.serving.Server(ctx, hbm_budget_bytes=1 << 30, max_active=1)
    outs = []
    for _ in range(3):
        out = torch.zeros_like(ref)
        outs.append((server.submit(fx.plan, len(Q), cts, out, gemv[1]), out))
    server.step()
    (t0, o0), (t1, o1), (t2, o2) = outs
    assert server.poll(t0)[0] == server.ADMITTED and server.poll(t2)[0] == server.WAITING
    server.cancel(t2)
    assert server.poll(t2)[0] == server.CANCELLED
    server.cancel(t0)
    server.drain()
    assert server.poll(t0)[0] == server.CANCELLED and server.poll(t1)[0] == server.DONE
    assert torch.equal(o1, ref)
    assert not torch.any(o2)  # never ran
    st = server.stats()
    assert st["cancelled"] == 2 and st["completed"] == 1 and st["resident_bytes"] == 0
    assert st["peak_bytes"] >= foot  # registers plus the keys and diagonals the job reads
    with pytest.raises(pkg.CkksError, match="finished"):
        server.cancel(t1)
    server.close()


def test_mac_accumulate_ntt_bit_exact(pkg, ctx):
    """ckks_multiply_plain_accumulate_ntt: sum_t ct_t * pt_t in the NTT domain,
    against the oracle's pointwise products and sums."""
    from conftest import rand_residues as rr
    rng = np.random.default_rng(12)
    L, n = 4, N
    q = Q[:L]
    plan = pkg.PreparedRnsNttPlan(ctx, q, n)
    cts = [rr(rng, q, n, polys=2) for _ in range(5)]
    pts = [rr(rng, q, n)[0] for _ in range(5)]
    out = pkg.to_host(plan.multiply_plain_accumulate_ntt([pkg.to_device(c) for c in cts],
                                                         [pkg.to_device(p) for p in pts], L))
    for comp in range(2):
        for i, qi in enumerate(q):
            want = np.zeros(n, dtype=np.uint64)
            for c, p in zip(cts, pts):
                want = O.pointwise_mac(want, c[comp, i], p[i], qi)
            acc = np.zeros(n, dtype=object)
            for c, p in zip(cts, pts):
                acc = (acc + c[comp, i].astype(object) * p[i].astype(object)) % qi
            assert (out[comp, i] == acc.astype(np.uint64)).all()
            assert (want == out[comp, i]).all()


def test_ntt_resident_programs_match_classic_ones(pkg, ctx, fx):
    """The NTT-resident form of every program (one multiply-accumulate per
    inner sum) gives the residues of the classic form (a transform pair per
    diagonal) through the serving loop, with fewer ops."""
    lin = pkg.linear
    rng = np.random.default_rng(21)
    L = len(Q)
    d = 32
    M = rng.uniform(-1, 1, (d, d))
    x = rng.uniform(-1, 1, d)
    progs = {}
    for nr in (False, True):
        b = lin.Builder(SLOTS, fx.key_for, fx.encode_pt, Q, ntt_resident=nr)
        progs[nr] = [("gemv", lin.gemv_program(b, M, L), [lin.tile(x, d, SLOTS)], L - 1),
                     ("fft", lin.fft_program(b, 64, L), [lin.tile(rng.uniform(-1, 1, 64), 64, SLOTS)], L - 3),
                     ("ccmm", lin.ccmm_program(b, 16, L, fx.relin_key(L - 2)),
                      [lin.tile(rng.uniform(-1, 1, 256), 256, SLOTS)] * 2, L - 3)]
    inputs = [torch.stack([fx.encrypt(v) for v in job[2]]).contiguous() for job in progs[False]]
    outs = {}
    for nr in (False, True):
        server = pkg.serving.Server(ctx, hbm_budget_bytes=8 << 30)
        res = []
        for (name, prog, _, ol), cts in zip(progs[nr], inputs):
            out = torch.zeros((2, ol, N), dtype=torch.int64, device="cuda")
            res.append((server.submit(fx.plan, L, cts, out, prog), out))
        server.drain()
        outs[nr] = [o for _, o in res]
        server.close()
    for (name, p0, _, _), (_, p1, _, _), o0, o1 in zip(progs[False], progs[True], outs[False], outs[True]):
        assert torch.equal(o0, o1), name
        assert len(p1.ops) <= len(p0.ops), name
    assert len(progs[True][0][1].ops) * 2 < len(progs[False][0][1].ops)  # GEMV: far fewer ops


def test_program_runner_graph_replays_match_eager(pkg, ctx, fx):
    """A GEMV program run directly (one library call per op) equals its
    serving-loop result; captured once as a hipGraph and replayed with new
    inputs and eager work between the replays, every replay equals the eager
    run on the same inputs."""
    lin = pkg.linear
    rng = np.random.default_rng(33)
    L = len(Q)
    d = 32
    M = rng.uniform(-1, 1, (d, d))
    prog = lin.gemv_program(fx.builder(), M, L)
    runner = pkg.serving.ProgramRunner(fx.plan, L, prog)
    xs = [torch.stack([fx.encrypt(lin.tile(rng.uniform(-1, 1, d), d, SLOTS))]).contiguous() for _ in range(4)]
    eager = [runner.run(x).clone() for x in xs]
    server = pkg.serving.Server(fx.ctx, hbm_budget_bytes=8 << 30)
    out = torch.zeros((2, L - 1, N), dtype=torch.int64, device="cuda")
    server.submit(fx.plan, L, xs[0], out, prog)
    server.drain()
    assert torch.equal(out, eager[0])
    server.close()
    runner.capture()
    for i in (1, 3, 0, 2):
        got = runner.replay(xs[i]).clone()
        assert torch.equal(got, eager[i]), f"replay {i}"
        assert torch.equal(runner.run(xs[(i + 1) % 4]), eager[(i + 1) % 4])  # eager work between replays
    runner.graph.close()
