"""Benchmark of the leveled RNS-CKKS rotation hot path on one MI355X.

Workload (batched core primitives, N = 2^16):
  Q = one 60-bit base prime + 24 40-bit primes (L = 24, 25 limbs), hybrid
  key switching with dnum = 3 blocks [9, 8, 8] and 7 60-bit special primes;
  a batch of 64 ciphertexts at N = 2^16 plus 16 ciphertexts at N = 2^15
  (13 limbs, blocks [5, 4, 4], 4 special primes); one step rotates every
  ciphertext by each of 2^0 .. 2^14 slots (15 Galois keys per degree).
  Inputs, keys and outputs stay resident in HBM; residues are synthetic
  (uniform, canonical) and keys are synthetic NTT-domain material: the work
  done per rotation does not depend on the values.

  The step (--rotation hoisted, the default) cuts and transforms each
  batch's canonical digits and P NTT(b) once (ckks_hoist_with_b, then
  ckks_rehoist into the same buffers on later steps) and then runs one
  ckks_rotate_hoisted_batch per four exponents (their MACs read each digit
  once); the hoisted MAC plus a per-exponent correction of the negated
  coefficients' lift gives the residues of the per-rotation key switch,
  which --rotation canonical times (ckks_apply_galois_batch per exponent)
  and which is reported beside it.

Reported (one JSON line):
  value           batched hybrid-key-switched rotations per second of the step
  batched_n65536_rotations_per_s
                  one step over the N = 2^16 batch alone
  hoisted / canonical
                  both steps: rates, N = 2^16 rates, call latencies
  plain_*         the plain per-op GPU path: the reference's own rotation
                  algorithm (RNS-gadget key switch with singleton blocks,
                  rotate_left_rns_ckks_with_prepared_ntt) one ciphertext per
                  call, timed on a sample of rotations; speedup_vs_plain is the
                  N = 2^16 batched rate over it
  hybrid_one_ct_per_call_*
                  the dnum = 3 hybrid key one ciphertext per call (batching
                  alone, same algorithm)
  bounded_hybrid_qp
                  the reference's bounded-digit rotation shape, canonical
                  and hoisted
  ntt_*           forward NTT over the batch's 96 digit limbs per ciphertext
  p50/p99_ms      latency of one batched rotation call of the step (64
                  ciphertexts at N = 2^16 and the 16 at N = 2^15 in one
                  launch sequence; hoisted: by --exponents-per-call
                  exponents, after the step's hoist)
  rotation_roofline
                  the step against HBM peak: per exponent every ciphertext
                  in and out plus each key once
  roofline        the forward NTT (dominant kernel pair) against HBM peak
  cpu_baseline    the CPU oracle's hybrid rotation on all the host cores the
                  job may use (one ciphertext per thread), bounded sample;
                  cpu_baseline_one_core the same on one core
"""
from __future__ import annotations

import argparse
import gc
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBPS = 8000.0


def _load_pkg():
    name = "ccmm_rs_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "ccmm-rs_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def rotation_exponent_left(n, steps):
    return pow(5, steps % (n // 2), 2 * n)


def rand_residues(torch, shape, device):
    # < 2^39 is canonical for every 40- and 60-bit prime of the chain
    return torch.randint(0, 1 << 39, shape, dtype=torch.int64, device=device)


def synthetic_key(torch, blocks_total, limbs, n):
    g = torch.Generator().manual_seed(limbs * 1000 + blocks_total)
    kb = torch.randint(0, 1 << 39, (blocks_total, limbs, n), dtype=torch.int64, generator=g).numpy()
    ka = torch.randint(0, 1 << 39, (blocks_total, limbs, n), dtype=torch.int64, generator=g).numpy()
    return kb.view(np.uint64), ka.view(np.uint64)


class Timer:
    """HIP events on torch's current stream, which the library runs on."""

    def __init__(self, torch):
        self.torch = torch

    def __enter__(self):
        self.e0 = self.torch.cuda.Event(enable_timing=True)
        self.e1 = self.torch.cuda.Event(enable_timing=True)
        self.e0.record()
        return self

    def __exit__(self, *a):
        self.e1.record()
        self.e1.synchronize()
        self.ms = self.e0.elapsed_time(self.e1)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(q_moduli, p_moduli, n, blocks, seconds=10.0, threads=1):
    """The oracle's rotation (automorphism + hybrid key switch) on `threads`
    host threads, each rotating its own ciphertext (independent work items;
    the oracle's C calls release the GIL), for about `seconds`."""
    import concurrent.futures as cf

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the oracle, timed here as the CPU baseline

    O.build()
    rng = np.random.default_rng(0)
    qp = q_moduli + p_moduli
    plan = O.RnsPlan(qp, n)
    L = len(q_moduli)
    kb = rng.integers(0, 1 << 39, (len(blocks), len(qp), n), dtype=np.uint64)
    ka = rng.integers(0, 1 << 39, (len(blocks), len(qp), n), dtype=np.uint64)
    k = rotation_exponent_left(n, 1)

    def worker(seed, t0):
        r = np.random.default_rng(seed)
        b = r.integers(0, 1 << 39, (L, n), dtype=np.uint64)
        a = r.integers(0, 1 << 39, (L, n), dtype=np.uint64)
        done = 0
        while time.perf_counter() - t0 < seconds:
            rb = np.stack([O.apply_automorphism(b[i], q, k) for i, q in enumerate(q_moduli)])
            ra = np.stack([O.apply_automorphism(a[i], q, k) for i, q in enumerate(q_moduli)])
            O.hybrid_key_switch(plan, L, blocks, rb, ra, kb, ka)
            done += 1
        return done

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        done = sum(ex.map(lambda s: worker(s, t0), range(threads)))
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "rotations/s", "cores": threads, "kind": "port",
            "sample": f"{done} hybrid rotations at N=2^16, L=24, dnum=3, K=7, one ciphertext per thread "
                      f"({el:.1f} s, oracle in C, {threads} thread{'s' if threads > 1 else ''}; "
                      f"host: {os.cpu_count()} logical CPUs, {_cpu_model()})"}


def measure_ntt(torch, plan, polys, limbs, n, reps=5):
    """Forward and inverse NTT of `polys` x `limbs` limbs (the digit limbs of
    a key-switch batch), timed live on the library's stream (the null
    stream, which torch's events also use)."""
    digits = rand_residues(torch, (polys, limbs, n), "cuda")
    plan.forward(digits)
    plan.inverse(digits)
    torch.cuda.synchronize()
    with Timer(torch) as tn:
        for _ in range(reps):
            plan.forward(digits)
    ntt_ms = tn.ms / reps
    limb_transforms = polys * limbs
    ntt_bytes = 2 * n * 8 * limb_transforms
    ntt_gbps = ntt_bytes / (ntt_ms / 1e3) / 1e9
    with Timer(torch) as ti:
        for _ in range(reps):
            plan.inverse(digits)
    intt_gbps = ntt_bytes / (ti.ms / reps / 1e3) / 1e9
    del digits
    print(f"# NTT fwd {ntt_ms:.2f} ms over {limb_transforms} limbs: {ntt_gbps:.0f} GB/s; "
          f"inv {intt_gbps:.0f} GB/s", file=sys.stderr, flush=True)
    return limb_transforms, ntt_ms, ntt_gbps, intt_gbps


# The 55-bit NTT prime pool of the reference's N=2^16 rotation benchmark
# (ckks/hybrid_rns_galois_key.rs:1152-1200): Q = the first 13, P = the next 2.
UHD_POOL = [0x7fffffffba0001, 0x7fffffffaa0001, 0x7fffffff7e0001, 0x7fffffff380001, 0x7ffffffef00001,
            0x7ffffffeba0001, 0x7ffffffeac0001, 0x7ffffffe700001, 0x7ffffffe600001, 0x7ffffffe4c0001,
            0x7ffffffe220001, 0x7ffffffe160001, 0x7ffffffdec0001, 0x7ffffffdd00001, 0x7ffffffdce0001]


def bounded_rotation_benchmark(torch, pkg, ctx, batch=8, rotations=8):
    """The reference's own rotation type: hybrid-QP Galois keys with bounded
    base-2^32 digits (23 digits) at N=2^16, Q = 13 x 55-bit, P = 2 x 55-bit;
    `batch` ciphertexts rotated by 2^0..2^(rotations-1) slots, each key
    switch decomposing the rotated a (canonical) or hoisting the digits once
    (ckks_bounded_hoist, then one ckks_rotate_hoisted_batch over every
    exponent, whose MACs read each digit once per two exponents; same
    residues)."""
    N = 1 << 16
    plan = pkg.PreparedRnsNttPlan(ctx, UHD_POOL, N)
    D = pkg.bounded_digit_count(plan, 13, 32)
    exps = [rotation_exponent_left(N, 1 << i) for i in range(rotations)]
    keys = []
    for i in range(rotations):
        kb, ka = synthetic_key(torch, D, 15, N)
        keys.append(pkg.BoundedKeySwitchKey(plan, 13, 32, kb, ka))
    cts = rand_residues(torch, (batch, 2, 13, N), "cuda")
    out = torch.empty_like(cts)
    outs = [torch.empty_like(cts) for _ in exps]
    # warm: workspaces and the permuted key tables of every exponent
    h = keys[0].hoist(cts)
    for e, key in zip(exps, keys):
        key.apply_galois(cts, e, out)
    pkg.rotate_hoisted_batch(ctx, [(h, key, e, cts, o) for e, key, o in zip(exps, keys, outs)])
    h.close()
    torch.cuda.synchronize()
    with Timer(torch) as tc:
        for e, key in zip(exps, keys):
            key.apply_galois(cts, e, out)
    with Timer(torch) as th:
        h = keys[0].hoist(cts)
        pkg.rotate_hoisted_batch(ctx, [(h, key, e, cts, o) for e, key, o in zip(exps, keys, outs)])
    h.close()
    n_rot = batch * rotations
    return {"workload": f"{batch} ct x {rotations} rotations at N=2^16, Q 13 x 55-bit, P 2 x 55-bit, "
                        f"bounded digits base 2^32 ({D} digits)",
            "canonical_rotations_per_s": round(n_rot / (tc.ms / 1e3), 1),
            "hoisted_rotations_per_s": round(n_rot / (th.ms / 1e3), 1),
            "hoist_speedup": round(tc.ms / th.ms, 2)}


def serving_benchmark(torch, pkg, ctx, n_requests=40, rate=4.0, seed=0):
    """Continuous-batching serving of a Poisson stream (seeded) of encrypted
    jobs on one GPU: 40% public 1024x1024 matrix x encrypted vector (N=2^14,
    L=8), 40% 64x64 encrypted CCMM (N=2^13, L=9), 20% length-4096 encrypted
    FFT1 (N=2^16, L=8).  Evaluation keys, matrix diagonals and DFT stage
    diagonals stay resident; admission by HBM footprint.  Synthetic
    ciphertexts and keys (the work does not depend on their values)."""
    lin, srv, enc = pkg.linear, pkg.serving, pkg.encoding
    q_all, p_all = pkg.chain_moduli(24, 7)
    key_cache = {}

    def blocks_for(limbs):
        return [limbs - limbs // 2, limbs // 2] if limbs > 1 else [1]

    def setup(n, L):
        q = q_all[:L]
        plan = pkg.PreparedRnsNttPlan(ctx, q, n)
        emb = enc.CanonicalEmbedding(n)
        qp_plans = {}
        keys = {}

        def key_for(limbs, steps):
            k = pow(5, steps, 2 * n)
            if (limbs, k) not in keys:
                if limbs not in qp_plans:
                    qp_plans[limbs] = pkg.PreparedRnsNttPlan(ctx, q[:limbs] + p_all[:4], n)
                blocks = blocks_for(limbs)
                shape = (len(blocks), limbs + 4, n)
                if shape not in key_cache:
                    key_cache[shape] = synthetic_key(torch, len(blocks), limbs + 4, n)
                keys[(limbs, k)] = pkg.HybridKeySwitchKey(qp_plans[limbs], limbs, blocks, *key_cache[shape])
            return keys[(limbs, k)], k

        def encode(values, limbs):
            pt = enc.encode(plan, emb, torch.from_numpy(np.asarray(values)), float(q[limbs - 1]), limbs)
            return plan.forward(pt, n_limbs=limbs)

        return plan, lin.Builder(n // 2, key_for, encode, q), key_for, qp_plans, keys

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    kinds = {}
    # GEMV
    plan, b, key_for, qp, keys = setup(1 << 14, 8)
    M = rng.uniform(-1, 1, (1024, 1024))
    kinds["gemv"] = (plan, 8, 1, lin.gemv_program(b, M, 8), 7)
    # CCMM
    plan, b, key_for, qp, keys = setup(1 << 13, 9)
    key_for(7, 1)
    relin = pkg.HybridKeySwitchKey(qp[7], 7, blocks_for(7), *key_cache[(2, 11, 1 << 13)])
    kinds["ccmm"] = (plan, 9, 2, lin.ccmm_program(b, 64, 9, relin), 6)
    kinds["ccmm"][3].keep.append(relin)
    # FFT1
    plan, b, key_for, qp, keys = setup(1 << 16, 8)
    kinds["fft"] = (plan, 8, 1, lin.fft_program(b, 4096, 8), 2)
    setup_s = time.perf_counter() - t0
    inputs = {name: rand_residues(torch, (n_in, 2, L, pl.degree), "cuda")
              for name, (pl, L, n_in, _, _) in kinds.items()}
    mix = rng.choice(["gemv", "ccmm", "fft"], size=n_requests, p=[0.4, 0.4, 0.2])
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    poisson = _serve_run(torch, srv, ctx, kinds, inputs, mix, arrivals)
    # capacity: the same mix all queued at once (the server runs 16 at a time)
    closed = _serve_run(torch, srv, ctx, kinds, inputs, mix, np.zeros(n_requests))
    # one GEMV request run directly, eagerly and as a replayed hipGraph
    direct = _direct_gemv(torch, srv, ctx, kinds["gemv"], inputs["gemv"])
    return {
        "requests": poisson["completed"],
        "requests_per_s": round(poisson["completed"] / poisson["makespan"], 2),
        "offered_rate_per_s": rate,
        "p50_ms": round(poisson["p50_ms"], 2),
        "p99_ms": round(poisson["p99_ms"], 2),
        "per_kind_p50_ms": poisson["per_kind_p50_ms"],
        "peak_active": poisson["peak_active"],
        "launches_per_request": round(poisson["launches"] / max(poisson["completed"], 1), 1),
        "capacity_requests_per_s": round(closed["completed"] / closed["makespan"], 2),
        "capacity_peak_active": closed["peak_active"],
        "capacity_p50_ms": round(closed["p50_ms"], 2),
        "capacity_p99_ms": round(closed["p99_ms"], 2),
        "gemv_direct": direct,
        "peak_admitted_bytes": poisson["peak_bytes"],
        "resident_key_and_diagonal_bytes": poisson["resident"],
        "program_ops": {k: v[3].counts() for k, v in kinds.items()},
        "setup_s": round(setup_s, 1),
        "workload": "Poisson (seeded) mix: 40% GEMV 1024x1024 N=2^14 L=8, 40% CCMM 64x64 N=2^13 L=9, "
                    "20% FFT1 4096 N=2^16 L=8; synthetic ciphertexts and keys; plaintext-diagonal inner "
                    "sums NTT-resident (one multiply-accumulate launch each)",
    }


def _serve_run(torch, srv, ctx, kinds, inputs, mix, arrivals):
    n_requests = len(mix)
    # budget: registers for 16 requests at once plus every kind's resident
    # keys and diagonals (charged once while any admitted request reads them)
    regs = max(pr.n_regs * 2 * L * pl.degree * 8 for pl, L, _, pr, _ in kinds.values())
    resident = sum(pr.footprint(L, pl.degree) - pr.n_regs * 2 * L * pl.degree * 8
                   for pl, L, _, pr, _ in kinds.values())
    server = srv.Server(ctx, hbm_budget_bytes=16 * regs + resident, max_active=16)
    arrivals = np.asarray(arrivals, dtype=np.float64)
    torch.cuda.synchronize()
    outs, tickets = [], []
    start = time.perf_counter()
    i = 0
    while i < n_requests or server.stats()["active"] or server.stats()["waiting"]:
        now = time.perf_counter() - start
        while i < n_requests and arrivals[i] <= now:
            name = str(mix[i])
            pl, L, n_in, prog, out_l = kinds[name]
            out = torch.empty((2, out_l, pl.degree), dtype=torch.int64, device="cuda")
            tickets.append((name, server.submit(pl, L, inputs[name], out, prog)))
            outs.append(out)
            i += 1
        if server.stats()["active"] or server.stats()["waiting"]:
            server.step()
        elif i < n_requests:
            time.sleep(max(0.0, min(0.01, arrivals[i] - (time.perf_counter() - start))))
    server.drain()
    makespan = time.perf_counter() - start
    st = server.stats()
    lat = {}
    for name, t in tickets:
        state, ms = server.poll(t)
        lat.setdefault(name, []).append(ms)
    server.close()
    return {"completed": int(st["completed"]), "makespan": makespan, "p50_ms": st["p50_ms"],
            "p99_ms": st["p99_ms"], "peak_active": int(st["peak_active"]), "launches": int(st["launches"]),
            "peak_bytes": int(st["peak_bytes"]), "resident": int(resident),
            "per_kind_p50_ms": {k: round(float(np.percentile(v, 50)), 2) for k, v in lat.items()}}


def _direct_gemv(torch, srv, ctx, kind, inputs, reps=10):
    """One GEMV request (1024 x 1024 at N=2^14) run outside the serving loop:
    eagerly (one library call per op) and as a captured hipGraph replayed;
    latency p50 of each.  The runner needs its own non-null stream."""
    plan, L, _, prog, _ = kind
    runner = srv.ProgramRunner(plan, L, prog)
    runner.run(inputs)

    def timed(fn):
        out = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()  # returns when the run is complete
            out.append((time.perf_counter() - t0) * 1e3)
        return float(np.percentile(out, 50))

    eager = timed(lambda: runner.run())
    runner.capture()
    graph = timed(lambda: runner.replay())
    runner.graph.close()
    return {"ops": len(prog.ops), "eager_p50_ms": round(eager, 2), "graph_p50_ms": round(graph, 2)}


def copy_calibration(torch, nbytes=1 << 30, reps=5):
    """A plain device copy of known size: the reference point for reading
    the HBM counters (FETCH_SIZE / WRITE_SIZE) of the NTT kernels."""
    src = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda")
    src.fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    with Timer(torch) as t:
        for _ in range(reps):
            dst.copy_(src)
    gbps = 2 * nbytes / (t.ms / reps / 1e3) / 1e9
    print(f"# copy {nbytes >> 20} MiB: {gbps:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    return gbps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--mixed", type=int, default=16, help="ciphertexts at N=2^15 in each step")
    ap.add_argument("--mixed-launch", choices=["batch", "separate"], default="batch",
                    help="the N=2^15 ciphertexts ride in the N=2^16 call (ckks_apply_galois_batch) "
                         "or run as calls of their own")
    ap.add_argument("--rotation", choices=["hoisted", "canonical"], default="hoisted",
                    help="timed step: each batch's digits hoisted once per step and rotated 15 times "
                         "(ckks_hoist + ckks_rotate_hoisted_batch, the canonical residues), or every "
                         "rotation decomposing its own digits (ckks_apply_galois_batch)")
    ap.add_argument("--hoist-b", action=argparse.BooleanOptionalAction, default=True,
                    help="hoist with b (ckks_hoist_with_b: P NTT(b) added inside the key MAC, no sigma_k "
                         "gather of b in ModDown; same residues)")
    ap.add_argument("--exponents-per-call", type=int, default=4,
                    help="hoisted step: exponents per ckks_rotate_hoisted_batch call (1-4; the MACs of one "
                         "batch's exponents share their digit reads)")
    ap.add_argument("--plain-sample", type=int, default=24, help="plain-path rotations timed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=16, help="threads of the all-core CPU baseline")
    ap.add_argument("--fusion", choices=["auto", "off", "xcd", "linear", "inverse"], default="auto",
                    help="digits' forward row pass fused with the key MAC: autotuned per key, "
                         "off, or fused with XCD-contiguous / linear block order")
    ap.add_argument("--only", choices=["all", "ntt", "serve", "bounded"], default="all",
                    help="ntt: only the NTT measurement and a copy of known size (for counter runs); "
                         "serve: only the serving loop; bounded: only the bounded-digit rotations")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-bounded", action="store_true",
                    help="skip the bounded-digit (canonical vs hoisted) rotation measurement")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--split", choices=["auto", "off", "on"], default="auto",
                    help="key switches split over two streams: autotuned per key, off, on")
    ap.add_argument("--ntt-variant", type=int, choices=[0, 1, 2, 3], default=None,
                    help="two-pass NTT kernel variant for both degrees (bit set of NTT_VARIANT_BLOCKED / "
                         "NTT_VARIANT_TILE8K), set after the autotuner")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default NTT variant and base-conversion engine")
    ap.add_argument("--base-conversion", choices=["auto", "valu", "mfma"], default="auto",
                    help="ModUp/ModDown engine: autotuned per key (auto), VALU, or int8 MFMA")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; sharding across GPUs is by ciphertext and is not "
              "run by this benchmark", file=sys.stderr)
        sys.exit(2)

    import torch

    pkg = _load_pkg()
    dev = "cuda"
    torch.cuda.init()
    ctx = pkg.Context(0)
    ctx.set_key_switch_split({"auto": ctx.KS_SPLIT_AUTO, "off": ctx.KS_SPLIT_OFF,
                              "on": ctx.KS_SPLIT_ON}[args.split])
    ctx.set_key_switch_fusion({"auto": ctx.KS_FUSION_AUTO, "off": ctx.KS_FUSION_OFF,
                               "xcd": ctx.KS_FUSION_XCD, "linear": ctx.KS_FUSION_LINEAR,
                               "inverse": ctx.KS_FUSION_INVERSE}[args.fusion])
    ctx.set_base_conversion({"auto": ctx.BASECONV_AUTO, "valu": ctx.BASECONV_VALU,
                             "mfma": ctx.BASECONV_MFMA}[args.base_conversion])

    N = 1 << 16
    N2 = 1 << 15
    q_moduli, p_moduli = pkg.chain_moduli(24, 7)
    L, K = 25, 7
    blocks = [9, 8, 8]
    L2, K2, blocks2 = 13, 4, [5, 4, 4]
    rot_steps = [1 << i for i in range(15)]

    qp_plan = pkg.PreparedRnsNttPlan(ctx, q_moduli + p_moduli, N)
    qp2 = q_moduli[:L2] + p_moduli[:K2]
    qp_plan2 = pkg.PreparedRnsNttPlan(ctx, qp2, N2)

    if args.only == "bounded":
        print(json.dumps(bounded_rotation_benchmark(torch, pkg, ctx)), flush=True)
        return

    if args.only == "serve":
        print(json.dumps(serving_benchmark(torch, pkg, ctx, args.serve_requests, args.serve_rate)), flush=True)
        return

    if args.only == "ntt":
        if not args.no_autotune:
            v16, _ = qp_plan.autotune_ntt(16)
            print(f"# autotune NTT: 2^16 variant {v16}", file=sys.stderr, flush=True)
        lt, ntt_ms, fwd, inv = measure_ntt(torch, qp_plan, args.batch * len(blocks), L + K, N)
        cp = copy_calibration(torch)
        print(json.dumps({"ntt_limb_transforms": lt, "ntt_forward_ms": round(ntt_ms, 3),
                          "ntt_forward_gbps": round(fwd, 1), "ntt_inverse_gbps": round(inv, 1),
                          "copy_gbps": round(cp, 1)}), flush=True)
        return

    print(f"# building {len(rot_steps)} Galois keys per degree", file=sys.stderr, flush=True)
    keys, keys2 = [], []
    for r in rot_steps:
        kb, ka = synthetic_key(torch, len(blocks), L + K, N)
        keys.append((rotation_exponent_left(N, r), pkg.HybridKeySwitchKey(qp_plan, L, blocks, kb, ka)))
        if args.mixed:
            kb2, ka2 = synthetic_key(torch, len(blocks2), L2 + K2, N2)
            keys2.append((rotation_exponent_left(N2, r),
                          pkg.HybridKeySwitchKey(qp_plan2, L2, blocks2, kb2, ka2)))

    # ---- autotuner: NTT kernel variant per degree, base-conversion engine per key
    tuned = {}
    if not args.no_autotune:
        v16, ms16 = qp_plan.autotune_ntt(16)
        v15, ms15 = qp_plan2.autotune_ntt(16)
        tuned["ntt_variant"] = {"2^16": v16, "2^15": v15}
        print(f"# autotune NTT: 2^16 variant {v16} {['%.2f' % m for m in ms16]} ms; "
              f"2^15 variant {v15}", file=sys.stderr, flush=True)
        if args.base_conversion == "auto" and args.fusion == "auto":
            picks = {}
            for deg, kl, n in (("2^16", keys, args.batch), ("2^15", keys2, max(args.mixed, 1))):
                for _, key in kl:
                    eng, fus, times = key.autotune(n)
                    name = ("mfma" if eng == ctx.BASECONV_MFMA else "valu") + "/" + \
                        ["off", "xcd", "linear", "inverse"][fus]
                    picks.setdefault(deg, {}).setdefault(name, 0)
                    picks[deg][name] += 1
                if kl:
                    print(f"# autotune key switch {deg} (last key, ms per (engine, fusion)): "
                          f"{ {k: round(v, 2) for k, v in times.items()} }", file=sys.stderr, flush=True)
            tuned["key_switch"] = picks

    if args.ntt_variant is not None:
        ctx.set_ntt_variant(16, args.ntt_variant)
        ctx.set_ntt_variant(15, args.ntt_variant)
        tuned["ntt_variant"] = {"2^16": args.ntt_variant, "2^15": args.ntt_variant, "set": True}

    B = args.batch
    cts = rand_residues(torch, (B, 2, L, N), dev)
    out = torch.empty_like(cts)
    cts2 = rand_residues(torch, (args.mixed, 2, L2, N2), dev) if args.mixed else None
    out2 = torch.empty_like(cts2) if args.mixed else None

    batched = args.mixed and args.mixed_launch == "batch"

    def step_canonical(lat=None, only16=False):
        """15 calls, each rotating the 64 + 16 ciphertexts (one launch
        sequence) and decomposing their rotated digits itself."""
        for idx, (k, key) in enumerate(keys):
            if lat is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            if batched and not only16:
                # both degrees in one launch sequence (one NTT launch per pass)
                k2, key2 = keys2[idx]
                pkg.apply_galois_batch(ctx, [(key, k, cts, out), (key2, k2, cts2, out2)])
            else:
                key.apply_galois(cts, k, out)
            if lat is not None:
                e1.record()
                lat.append((e0, e1))
            if args.mixed and not batched and not only16:
                k2, key2 = keys2[idx]
                key2.apply_galois(cts2, k2, out2)

    held = {}  # hoisted-digit handles, created by the first (warmup) step and re-hoisted after

    def hoist(name, key, batch):
        if name in held:
            return held[name].rehoist(key, batch)
        held[name] = key.hoist(batch, with_b=args.hoist_b)
        return held[name]

    host_call_ms = []  # host time of each ckks_rotate_hoisted_batch call
    # exponents per hoisted call: their MACs share the digit reads, and each
    # needs its own output
    E = max(1, min(args.exponents_per_call, 4))
    outs = [out] + [torch.empty_like(out) for _ in range(E - 1)]
    outs2 = [out2] + [torch.empty_like(out2) for _ in range(E - 1)] if args.mixed else [None] * E

    def step_hoisted(lat=None, only16=False, hoist_lat=None):
        """The canonical digits of each batch cut and transformed once per
        step (ckks_hoist the first time, ckks_rehoist into the same buffers
        after), then 15 calls each rotating the 64 + 16 ciphertexts from
        them (ckks_rotate_hoisted_batch: the inverse NTTs of both degrees in
        one launch per pass), E exponents per call (their MACs read each
        digit once); the residues are those of step_canonical."""
        if hoist_lat is not None:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        h16 = hoist("16", keys[0][1], cts)
        h15 = hoist("15", keys2[0][1], cts2) if args.mixed and not only16 else None
        if hoist_lat is not None:
            e1.record()
            hoist_lat.append((e0, e1))
        for c0 in range(0, len(keys), E):
            chunk = range(c0, min(c0 + E, len(keys)))
            if lat is not None:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
            entries = []
            for e, idx in enumerate(chunk):
                k, key = keys[idx]
                entries.append((h16, key, k, cts, outs[e]))
                if h15 is not None and batched:
                    k2, key2 = keys2[idx]
                    entries.append((h15, key2, k2, cts2, outs2[e]))
            th = time.perf_counter()
            pkg.rotate_hoisted_batch(ctx, entries)
            host_call_ms.append((time.perf_counter() - th) * 1e3)
            if lat is not None:
                e1.record()
                lat.append((e0, e1))
            if h15 is not None and not batched:
                for e, idx in enumerate(chunk):
                    k2, key2 = keys2[idx]
                    key2.rotate_hoisted(h15, cts2, k2, outs2[e])

    def timed(step_fn, label):
        print(f"# warmup ({label})", file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            step_fn()
        torch.cuda.synchronize()
        lat_events, hoist_events = [], []
        # The host enqueues far ahead of the GPU except right after the
        # per-step host synchronisation of the hoisted step's zero-block
        # check; a full cyclic garbage collection there (tens of ms over
        # torch's objects) left the GPU idle.  Objects that exist before the
        # timed steps are moved out of the collector's reach.
        host_call_ms.clear()
        gc.collect()
        gc.freeze()
        pauses = []
        def on_gc(phase, info, _t=[0.0]):
            if phase == "start":
                _t[0] = time.perf_counter()
            else:
                pauses.append((info["generation"], (time.perf_counter() - _t[0]) * 1e3))
        gc.callbacks.append(on_gc)
        t0 = time.perf_counter()
        with Timer(torch) as tm:
            for _ in range(args.steps):
                if step_fn is step_hoisted:
                    step_fn(lat_events, hoist_lat=hoist_events)
                else:
                    step_fn(lat_events)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gc.callbacks.remove(on_gc)
        gc.unfreeze()
        print(f"# {label}: {len(pauses)} garbage collections in the timed steps, longest "
              f"{max((p[1] for p in pauses), default=0.0):.2f} ms", file=sys.stderr, flush=True)
        if step_fn is step_hoisted and host_call_ms:
            hc = sorted(host_call_ms)
            print(f"# {label}: host time per rotation call p50 {hc[len(hc) // 2]:.3f} ms, max {hc[-1]:.3f} ms",
                  file=sys.stderr, flush=True)
        ms = tm.ms / args.steps
        lats = sorted(e0.elapsed_time(e1) for e0, e1 in lat_events)
        hoists = [e0.elapsed_time(e1) for e0, e1 in hoist_events]
        # the N=2^16 batch alone (one step of it), the like-for-like figure
        # for the comparison with the plain path, which runs at N=2^16
        t16 = time.perf_counter()
        with Timer(torch) as t16ev:
            step_fn(only16=True)
        torch.cuda.synchronize()
        n16 = B * len(keys) / (t16ev.ms / 1e3)
        print(f"# {label}: {ms:.1f} ms/step (wall {(wall / args.steps) * 1e3:.1f}); N=2^16 batch alone "
              f"{n16:.0f} rot/s ({(time.perf_counter() - t16) * 1e3:.0f} ms wall)", file=sys.stderr, flush=True)
        return ms, lats, hoists, n16

    rot_per_step = len(rot_steps) * (B + args.mixed)
    can_ms, can_lats, _, can_n16 = timed(step_canonical, "canonical")
    can_value = rot_per_step / (can_ms / 1e3)
    hoi_ms, hoi_lats, hoi_hoists, hoi_n16 = timed(step_hoisted, "hoisted")
    hoi_value = rot_per_step / (hoi_ms / 1e3)
    for h in held.values():
        h.close()
    if args.rotation == "hoisted":
        ms_step, lats, n16_value, value = hoi_ms, hoi_lats, hoi_n16, hoi_value
    else:
        ms_step, lats, n16_value, value = can_ms, can_lats, can_n16, can_value
    p50 = float(np.percentile(lats, 50))
    p99 = float(np.percentile(lats, 99))
    print(f"# step: canonical {can_value:.0f} rot/s, hoisted {hoi_value:.0f} rot/s", file=sys.stderr, flush=True)

    # ---- plain per-op path: reference algorithm (RNS gadget, singleton
    # blocks), one ciphertext per call
    q_plan = pkg.PreparedRnsNttPlan(ctx, q_moduli, N)
    kbp, kap = synthetic_key(torch, L, L, N)
    plain_key = pkg.RnsKeySwitchKey(q_plan, [1] * L, kbp, kap)
    k1 = rotation_exponent_left(N, 1)
    one = cts[0:1].contiguous()
    one_out = torch.empty_like(one)
    plain_key.apply_galois(one, k1, one_out)
    torch.cuda.synchronize()
    with Timer(torch) as tp:
        for i in range(args.plain_sample):
            plain_key.apply_galois(cts[i % B:i % B + 1], k1, one_out)
    plain_value = args.plain_sample / (tp.ms / 1e3)
    del plain_key
    print(f"# plain: {tp.ms / args.plain_sample:.2f} ms/rotation, {plain_value:.0f} rot/s", file=sys.stderr,
          flush=True)
    # the same dnum=3 hybrid key, one ciphertext per call (no batching)
    k0, key0 = keys[0]
    key0.apply_galois(one, k0, one_out)
    torch.cuda.synchronize()
    with Timer(torch) as th1:
        for i in range(args.plain_sample):
            key0.apply_galois(cts[i % B:i % B + 1], k0, one_out)
    single_hybrid_value = args.plain_sample / (th1.ms / 1e3)
    print(f"# hybrid, one ciphertext per call: {single_hybrid_value:.0f} rot/s", file=sys.stderr, flush=True)

    limb_transforms, ntt_ms, ntt_gbps, intt_gbps = measure_ntt(torch, qp_plan, B * len(blocks), L + K, N)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_forward_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    result = {
        "metric": "key-switched rotations/s at N=2^16 L=24 batched vs plain GPU; NTT GB/s; p99 lat",
        "value": round(value, 1),
        "unit": "rotations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": f"rotate {B} ct at N=2^16 L=24 (dnum=3 [9,8,8], K=7) + {args.mixed} ct at "
                        f"N=2^15 L=12 (dnum=3 [5,4,4], K=4) by 2^0..2^14 slots",
            "ring_degree": N, "limbs": L, "dnum": 3, "special_primes": K,
            "batch": B, "mixed_batch": args.mixed, "rotations_per_step": rot_per_step,
        },
        "rotation": args.rotation,
        "batched_n65536_rotations_per_s": round(n16_value, 1),
        "plain_rotations_per_s": round(plain_value, 1),
        "speedup_vs_plain": round(n16_value / plain_value, 2),
        "hybrid_one_ct_per_call_rotations_per_s": round(single_hybrid_value, 1),
        "speedup_vs_hybrid_one_ct_per_call": round(n16_value / single_hybrid_value, 2),
        "hoisted": {
            "rotations_per_s": round(hoi_value, 1), "n65536_rotations_per_s": round(hoi_n16, 1),
            "hoist_ms_p50": round(float(np.percentile(hoi_hoists, 50)), 3) if hoi_hoists else None,
            "p50_ms": round(float(np.percentile(hoi_lats, 50)), 3),
            "p99_ms": round(float(np.percentile(hoi_lats, 99)), 3),
            "with_b": args.hoist_b,
            "exponents_per_call": E,
            "note": "per step: each batch's digits re-hoisted (ckks_rehoist into the buffers of the "
                    "warmup step's ckks_hoist" + ("_with_b, which also transforms P b" if args.hoist_b else "") +
                    "; the zero-block check and its host synchronisation included), then one "
                    f"ckks_rotate_hoisted_batch per {E} exponent{'s' if E > 1 else ''} (their MACs share "
                    "the digit reads); residues equal the canonical rotation's (tests compare them)"},
        "canonical": {
            "rotations_per_s": round(can_value, 1), "n65536_rotations_per_s": round(can_n16, 1),
            "speedup_vs_plain": round(can_n16 / plain_value, 2),
            "speedup_vs_hybrid_one_ct_per_call": round(can_n16 / single_hybrid_value, 2),
            "p50_ms": round(float(np.percentile(can_lats, 50)), 3),
            "p99_ms": round(float(np.percentile(can_lats, 99)), 3),
            "note": "every rotation decomposes its own rotated digits (ckks_apply_galois_batch per exponent)"},
        "ntt_forward_gbps": round(ntt_gbps, 1),
        "ntt_inverse_gbps": round(intt_gbps, 1),
        "autotune": tuned,
        "p50_ms": round(p50, 3),
        "p99_ms": round(p99, 3),
        "latency_call": (f"one {args.rotation} rotation call of {B} ct at N=2^16" +
                         (f" and {args.mixed} ct at N=2^15 (one launch sequence)" if batched else "") +
                         (f" by {E} exponent{'s' if E > 1 else ''}, from digits hoisted once per step"
                          if args.rotation == "hoisted" else " by one exponent")),
        "roofline": {
            "bound": "hbm",
            "kernel": "forward NTT (column + row pass), 2*N*8 bytes per limb transform",
            "achieved": round(ntt_gbps, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(ntt_gbps / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
        },
    }
    # rotation-level roofline: per ciphertext 2 l N 8 bytes in and out, and
    # each key's dnum 2 (l + K) N 8 bytes once per call (amortised over its
    # batch), against the step's time
    call_bytes = B * 4 * L * N * 8 + len(blocks) * 2 * (L + K) * N * 8
    if args.mixed:
        call_bytes += args.mixed * 4 * L2 * N2 * 8 + len(blocks2) * 2 * (L2 + K2) * N2 * 8
    rot_gbps = call_bytes * len(rot_steps) / (ms_step / 1e3) / 1e9
    result["rotation_roofline"] = {
        "bound": "hbm", "achieved": round(rot_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(rot_gbps / HBM_PEAK_GBPS, 4), "algorithmic_bytes_per_call": call_bytes,
        "note": "whole timed step (hoist included): per exponent, every ciphertext in and out plus each "
                "degree's Galois key once, over ms_per_step"}
    if not args.no_bounded:
        print("# bounded-digit rotations", file=sys.stderr, flush=True)
        result["bounded_hybrid_qp"] = bounded_rotation_benchmark(torch, pkg, ctx)
        print(f"# bounded: {result['bounded_hybrid_qp']}", file=sys.stderr, flush=True)
    if not args.no_serving:
        print("# serving loop", file=sys.stderr, flush=True)
        result["serving"] = serving_benchmark(torch, pkg, ctx, args.serve_requests, args.serve_rate)
        print(f"# serving: {result['serving']['requests_per_s']} req/s, p50 {result['serving']['p50_ms']} ms, "
              f"p99 {result['serving']['p99_ms']} ms", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        print("# cpu baseline", file=sys.stderr, flush=True)
        # all the host cores this job may use (the GPU box grants 16 per GPU),
        # then one core as the reference's own single-threaded path runs
        threads = max(1, min(args.cpu_threads, len(os.sched_getaffinity(0))))
        result["cpu_baseline"] = cpu_baseline(q_moduli, p_moduli, N, blocks, args.cpu_seconds, threads)
        result["cpu_baseline_one_core"] = cpu_baseline(q_moduli, p_moduli, N, blocks, args.cpu_seconds / 2, 1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
