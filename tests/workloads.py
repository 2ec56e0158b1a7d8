"""Device-side harness for the encrypted linear-algebra workloads at their
stated sizes: one parameter set (ring degree N, Q = one 60-bit base prime +
L 40-bit primes, K 60-bit special primes, dnum gadget blocks), a ternary
secret, keys generated on the device (ckks_keygen) from randomness sampled
on the host with numpy, CKKS encode/encrypt/decrypt/decode through the
product, and the programs of ccmm-rs_amd/linear.py run by the serving loop."""
from __future__ import annotations

import numpy as np
import torch

from conftest import O, P60, Q0, Q40

SIGMA = 3.2


def blocks_for(limbs: int, dnum: int) -> list:
    """dnum contiguous gadget blocks over `limbs` limbs, sizes differing by at
    most one, larger blocks first ([9, 8, 8] for 25 limbs, dnum 3)."""
    base, extra = divmod(limbs, dnum)
    return [base + (1 if j < extra else 0) for j in range(dnum) if base + (1 if j < extra else 0)]


class Workload:
    def __init__(self, pkg, ctx, logn: int, L: int, K: int, dnum: int, seed: int = 0, scale_bits: int = 40):
        self.pkg, self.ctx = pkg, ctx
        self.n = 1 << logn
        self.slots = self.n // 2
        self.Q = [Q0[0]] + Q40[:L]
        self.P = P60[:K]
        self.dnum = dnum
        self.scale = float(2 ** scale_bits)
        self.s = O.Rng(seed).ternary(self.n)
        self.ds = torch.from_numpy(self.s.astype(np.int8)).cuda()
        self.rng = np.random.default_rng(seed + 1)
        self.plan = pkg.PreparedRnsNttPlan(ctx, self.Q, self.n)
        self.emb = pkg.encoding.CanonicalEmbedding(self.n)
        self.qp_plans, self.keys, self.relin = {}, {}, {}
        self.key_bytes = 0

    # ------------------------------------------------------------- keys
    def _qp(self, limbs):
        if limbs not in self.qp_plans:
            self.qp_plans[limbs] = self.pkg.PreparedRnsNttPlan(self.ctx, self.Q[:limbs] + self.P, self.n)
        return self.qp_plans[limbs]

    def _sample(self, moduli, entries):
        a = np.empty((entries, len(moduli), self.n), dtype=np.uint64)
        for i, q in enumerate(moduli):
            a[:, i, :] = self.rng.integers(0, q, size=(entries, self.n), dtype=np.uint64)
        e = np.rint(self.rng.normal(0.0, SIGMA, size=(entries, self.n))).astype(np.int64)
        return self.pkg.to_device(a), torch.from_numpy(e).cuda()

    def _hybrid(self, limbs, source_fn):
        qp = self._qp(limbs)
        blocks = blocks_for(limbs, self.dnum)
        a, e = self._sample(qp.moduli, len(blocks))
        key = self.pkg.HybridKeySwitchKey.generate(qp, limbs, blocks, source_fn(qp), self.ds, a, e)
        self.key_bytes += 2 * len(blocks) * len(qp.moduli) * self.n * 8
        return key

    def key_for(self, limbs, steps):
        k = pow(5, steps, 2 * self.n)
        if (limbs, k) not in self.keys:
            self.keys[(limbs, k)] = self._hybrid(limbs, lambda qp: self.pkg.galois_source(qp, self.ds, k))
        return self.keys[(limbs, k)], k

    def relin_key(self, limbs):
        if limbs not in self.relin:
            self.relin[limbs] = self._hybrid(limbs, lambda qp: self.pkg.square_source(qp, self.ds))
        return self.relin[limbs]

    # ------------------------------------------------------- plaintexts
    def encode_pt(self, values, limbs):
        """NTT-domain plaintext of a slot vector at `limbs` limbs, scaled by
        the limb that the following rescale drops (the scale stays put)."""
        enc = self.pkg.encoding
        pt = enc.encode(self.plan, self.emb, torch.from_numpy(np.asarray(values, dtype=np.complex128)),
                        float(self.Q[limbs - 1]), limbs)
        return self.plan.forward(pt, n_limbs=limbs)

    def builder(self):
        return self.pkg.linear.Builder(self.slots, self.key_for, self.encode_pt, self.Q)

    def encrypt(self, slots):
        enc = self.pkg.encoding
        L = len(self.Q)
        m = enc.encode(self.plan, self.emb, torch.from_numpy(np.asarray(slots, dtype=np.complex128)), self.scale)
        e = torch.from_numpy(np.rint(self.rng.normal(0.0, SIGMA, self.n)).astype(np.int64)).cuda()
        a = np.empty((L, self.n), dtype=np.uint64)
        for i, q in enumerate(self.Q):
            a[i] = self.rng.integers(0, q, size=self.n, dtype=np.uint64)
        a = self.pkg.to_device(a)
        b = self.plan.encrypt_raw(m, e, self.ds, a)
        return torch.stack([b, a])

    def decrypt(self, ct, limbs, scale):
        ct = ct.view(2, limbs, self.n)
        m = self.plan.decrypt_raw(ct[0].contiguous(), ct[1].contiguous(), self.ds, n_limbs=limbs)
        return self.pkg.encoding.decode(self.plan, self.emb, m, scale, limbs)[0].cpu().numpy()

    # ---------------------------------------------------------- running
    def submit(self, server, prog, inputs, out_limbs):
        cts = torch.stack([self.encrypt(v) for v in inputs]).contiguous()
        out = torch.zeros((2, out_limbs, self.n), dtype=torch.int64, device="cuda")
        t = server.submit(self.plan, len(self.Q), cts, out, prog)
        return t, cts, out
