"""The encrypted linear-algebra workloads at their stated sizes, on keys
generated on the device, decrypted and checked against FP64:
- structured CCMM of 64 x 64 matrices at N = 2^13, Q one 60-bit + 8 40-bit
  primes, dnum 2, 4 special primes (several of the 256 pairs, batched by the
  serving loop), against the FP64 matmul;
- FFT2 of 256 x 256 synthetic images (8 Gaussian blobs plus uniform noise)
  at N = 2^16, Q one 60-bit + 16 40-bit primes, dnum 4, against numpy's fft2
  (radix-4 stages, plaintext diagonals and rotations);
- a serving mix of uneven ring degrees (1024 x 1024 GEMV at N = 2^14, 64 x 64
  CCMM at N = 2^13, length-4096 FFT1 at N = 2^16, 8 levels): every result is
  bit-identical alone and in the mix, and decrypts to its FP64 value.
Tolerances follow the reference's own: 2e-3 max-abs for CP products
(bin/r30_13b_ckks_primitives.rs:250), 1e-3 for multiply + rescale
(ckks/rns_multiply.rs:526), relative L2 below 1e-2 at N = 65536
(tests/r30_9a_13_hybrid_65536.rs:259-262)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from workloads import Workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ccmm_wl(pkg, ctx):
    return Workload(pkg, ctx, logn=13, L=8, K=4, dnum=2, seed=3)


def _ccmm_job(pkg, wl, rng, d=64):
    lin = pkg.linear
    L = len(wl.Q)
    A = rng.uniform(-1, 1, (d, d))
    B = rng.uniform(-1, 1, (d, d))
    prog = lin.ccmm_program(wl.builder(), d, L, wl.relin_key(L - 2))
    inputs = [lin.tile(A.ravel(), d * d, wl.slots), lin.tile(B.ravel(), d * d, wl.slots)]
    return prog, inputs, L - 3, wl.scale * wl.scale / wl.Q[L - 3], lin.tile((A @ B).ravel(), d * d, wl.slots)


def test_ccmm_64_at_n8192(pkg, ctx, ccmm_wl):
    wl = ccmm_wl
    rng = np.random.default_rng(11)
    jobs = [_ccmm_job(pkg, wl, rng) for _ in range(4)]
    server = pkg.serving.Server(ctx, hbm_budget_bytes=8 << 30)
    subs = [wl.submit(server, prog, inputs, out_limbs) for prog, inputs, out_limbs, _, _ in jobs]
    server.drain()
    for (prog, inputs, out_limbs, scale, want), (t, _, out) in zip(jobs, subs):
        assert server.poll(t)[0] == server.DONE
        got = wl.decrypt(out, out_limbs, scale)
        err = np.max(np.abs(got - want))
        assert err < 1e-3, err
    st = server.stats()
    assert st["completed"] == 4
    # the four jobs' identical op streams were batched into shared launches
    assert st["launches"] < sum(len(j[0].ops) for j in jobs)
    server.close()


def _blob_image(rng, size=256, blobs=8):
    y, x = np.mgrid[0:size, 0:size]
    img = np.zeros((size, size))
    for _ in range(blobs):
        cy, cx = rng.uniform(0, size, 2)
        sig = rng.uniform(4, 32)
        amp = rng.uniform(0, 1)
        img += amp * np.exp(-((y - cy) ** 2 + (x - cx) ** 2) / (2 * sig * sig))
    return img + rng.uniform(-0.01, 0.01, (size, size))


def test_fft2_256x256_at_n65536(pkg, ctx):
    lin = pkg.linear
    wl = Workload(pkg, ctx, logn=16, L=16, K=4, dnum=4, seed=4)
    rows = cols = 256
    L = len(wl.Q)
    even, odd = lin.fft2_programs(wl.builder(), rows, cols, L)
    out_limbs = L - lin.fft2_levels(rows, cols)
    rng = np.random.default_rng(16)
    images = [_blob_image(rng) for _ in range(2)]
    cp = lin.bitrev_permutation(cols)
    rp = lin.bitrev_permutation(rows // 2)
    server = pkg.serving.Server(ctx, hbm_budget_bytes=16 << 30)
    subs = []
    for img in images:
        halves = [img[: rows // 2][:, cp].ravel(), img[rows // 2:][:, cp].ravel()]
        cts = torch.stack([wl.encrypt(h) for h in halves]).contiguous()
        for prog in (even, odd):
            out = torch.zeros((2, out_limbs, wl.n), dtype=torch.int64, device="cuda")
            subs.append((server.submit(wl.plan, L, cts, out, prog), out))
    server.drain()
    for i, img in enumerate(images):
        Y = np.fft.fft2(img)
        for par in (0, 1):
            t, out = subs[2 * i + par]
            assert server.poll(t)[0] == server.DONE
            got = wl.decrypt(out, out_limbs, wl.scale).reshape(rows // 2, cols)
            want = Y[2 * rp + par]
            rel = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert rel < 1e-2, (i, par, rel)
            assert rel < 1e-6, (i, par, rel)  # far inside the reference's bound
    server.close()


def test_serving_mix_uneven_degrees(pkg, ctx, ccmm_wl):
    """GEMV (N = 2^14), CCMM (N = 2^13) and FFT1 (N = 2^16) requests, first
    each alone, then all arriving interleaved with server steps: the outputs
    are bit-identical and decrypt to the FP64 results."""
    lin = pkg.linear
    rng = np.random.default_rng(55)
    gemv_wl = Workload(pkg, ctx, logn=14, L=8, K=3, dnum=3, seed=5)
    fft_wl = Workload(pkg, ctx, logn=16, L=8, K=3, dnum=3, seed=6)
    jobs = []  # (workload, program, inputs, out_limbs, scale, want, tol)
    d = 1024
    M = rng.uniform(-1, 1, (d, d))
    x = rng.uniform(-1, 1, d)
    L = len(gemv_wl.Q)
    jobs.append((gemv_wl, lin.gemv_program(gemv_wl.builder(), M, L), [lin.tile(x, d, gemv_wl.slots)], L - 1,
                 gemv_wl.scale, lin.tile(M @ x, d, gemv_wl.slots), 2e-3))
    prog, inputs, ol, sc, want = _ccmm_job(pkg, ccmm_wl, rng)
    jobs.append((ccmm_wl, prog, inputs, ol, sc, want, 1e-3))
    d = 4096
    z = rng.uniform(-1, 1, d) + 1j * rng.uniform(-1, 1, d)
    L = len(fft_wl.Q)
    perm = lin.bitrev_permutation(d)
    stages = len(lin.fft_stages(d))
    jobs.append((fft_wl, lin.fft_program(fft_wl.builder(), d, L), [lin.tile(z[perm], d, fft_wl.slots)],
                 L - stages, fft_wl.scale, lin.tile(np.fft.fft(z), d, fft_wl.slots), 2e-3))
    inputs, alone = {}, {}
    for idx, (wl, prog, ins, ol, _, _, _) in enumerate(jobs):
        server = pkg.serving.Server(ctx, hbm_budget_bytes=8 << 30)
        t, cts, out = wl.submit(server, prog, ins, ol)
        server.drain()
        assert server.poll(t)[0] == server.DONE
        inputs[idx], alone[idx] = cts, out.clone()
        server.close()
    # the mix: a seeded arrival order with repeats, two server steps between
    # arrivals, so requests join while others are in flight
    order = [int(i) for i in np.random.default_rng(7).permutation([0, 0, 1, 1, 2])]
    server = pkg.serving.Server(ctx, hbm_budget_bytes=8 << 30)
    outs = []
    for idx in order:
        wl, prog, _, ol, _, _, _ = jobs[idx]
        out = torch.zeros_like(alone[idx])
        outs.append((idx, server.submit(wl.plan, len(wl.Q), inputs[idx], out, prog), out))
        server.step()
        server.step()
    server.drain()
    st = server.stats()
    assert st["completed"] == len(order) and st["peak_active"] >= 2
    for idx, t, out in outs:
        assert server.poll(t)[0] == server.DONE
        assert torch.equal(out, alone[idx]), f"job {idx} differs in the mix"
        wl, _, _, ol, scale, want, tol = jobs[idx]
        err = np.max(np.abs(wl.decrypt(out, ol, scale) - want))
        assert err < tol, (idx, err)
    server.close()
