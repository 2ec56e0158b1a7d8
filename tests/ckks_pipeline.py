"""Host-side helpers for end-to-end CKKS checks: parameters of the small
config (N=4096, one 60-bit base prime + 3 40-bit primes, one 60-bit special
prime, scale 2^40, ternary secret, Gaussian sigma 3.2, seed 0), key
generation from host-sampled randomness, encryption and decryption.  Built
on the oracle; the GPU pipeline is compared with it."""
from __future__ import annotations

import math

import numpy as np

from conftest import O, P60, Q0, Q40

import embedding as E  # oracle/embedding.py

N = 4096
SCALE = float(2 ** 40)
SIGMA = 3.2
Q = [Q0[0]] + Q40[:3]
P = P60[:1]


def ternary_to_mod(s, q):
    return np.array([int(v) % q for v in s.tolist()], dtype=np.uint64)


def sigma_ternary(s, k, n):
    """sigma_k(s) for a ternary secret, as a ternary vector."""
    r = O.apply_automorphism(np.array([int(v) % 3 for v in s.tolist()], dtype=np.uint64), 3, k)
    return np.array([{0: 0, 1: 1, 2: -1}[int(v)] for v in r], dtype=np.int8)


def secret_square(plan: O.RnsPlan, s):
    """s^2 as an RNS polynomial (NTT product per limb)."""
    out = []
    for i, q in enumerate(plan.moduli):
        sq = ternary_to_mod(s, q)
        f = plan.plans[i].forward(sq)
        out.append(plan.plans[i].inverse(O.pointwise_mul(f, f, q)))
    return np.stack(out)


def keygen_relin_rns(plan: O.RnsPlan, s, rng: O.Rng):
    """RnsMultiplicationKey with singleton blocks (grafting/rns_key_switch.rs:
    1430-1512): entry j encrypts E_j * s^2 (E_j one-hot on limb j), NTT domain."""
    L, n = len(plan.moduli), plan.n
    s2 = secret_square(plan, s)
    kb = np.zeros((L, L, n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    for j in range(L):
        msg = np.zeros((L, n), dtype=np.uint64)
        msg[j] = s2[j]
        e = rng.gaussian(SIGMA, n)
        a = np.stack([rng.uniform(q, n) for q in plan.moduli])
        b = O.encrypt(plan, msg, e, s, a)
        kb[j], ka[j] = plan.forward(b), plan.forward(a)
    return kb, ka


def keygen_relin_hybrid(qp_plan: O.RnsPlan, q_limbs: int, blocks, s, rng: O.Rng):
    """dnum hybrid relinearization key: entry j encrypts P * E_j * s^2 under
    s over Q u P, NTT domain."""
    s2 = secret_square(qp_plan, s)
    return _keygen_hybrid_target(qp_plan, q_limbs, blocks, s, rng,
                                 lambda i, q: s2[i])


def keygen_galois_hybrid(qp_plan: O.RnsPlan, q_limbs: int, blocks, s, k, rng: O.Rng):
    """dnum hybrid Galois key: entry j encrypts P * E_j * sigma_k(s) under s
    over Q u P, NTT domain."""
    s_rot = sigma_ternary(s, k, qp_plan.n)
    return _keygen_hybrid_target(qp_plan, q_limbs, blocks, s, rng,
                                 lambda i, q: ternary_to_mod(s_rot, q))


def _keygen_hybrid_target(qp_plan: O.RnsPlan, q_limbs: int, blocks, s, rng: O.Rng, target):
    n = qp_plan.n
    moduli = qp_plan.moduli
    Pv = math.prod(moduli[q_limbs:])
    kb = np.zeros((len(blocks), len(moduli), n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    start = 0
    for j, sz in enumerate(blocks):
        msg = np.zeros((len(moduli), n), dtype=np.uint64)
        for i in range(start, start + sz):
            q = moduli[i]
            msg[i] = np.array([(int(v) * Pv) % q for v in target(i, q).tolist()], dtype=np.uint64)
        e = rng.gaussian(SIGMA, n)
        a = np.stack([rng.uniform(q, n) for q in moduli])
        b = O.encrypt(qp_plan, msg, e, s, a)
        kb[j], ka[j] = qp_plan.forward(b), qp_plan.forward(a)
        start += sz
    return kb, ka


def _keygen_bounded_target(plan: O.RnsPlan, q_limbs: int, w: int, s, rng: O.Rng, target):
    """Entry j encrypts P * B^j * target under s (P = 1 without special
    primes), NTT domain: BoundedRnsMultiplicationKey (grafting/
    bounded_key_switch.rs:27-103) / PreparedHybridRnsGaloisKey
    (ckks/hybrid_rns_galois_key.rs:61-170)."""
    n, moduli = plan.n, plan.moduli
    Pv = math.prod(moduli[q_limbs:])
    count = O.bounded_digit_count(moduli[:q_limbs], w)
    kb = np.zeros((count, len(moduli), n), dtype=np.uint64)
    ka = np.zeros_like(kb)
    for j in range(count):
        msg = np.stack([(target(i, q).astype(object) * (Pv * (1 << (w * j)) % q) % q).astype(np.uint64)
                        for i, q in enumerate(moduli)])
        e = rng.gaussian(SIGMA, n)
        a = np.stack([rng.uniform(q, n) for q in moduli])
        b = O.encrypt(plan, msg, e, s, a)
        kb[j], ka[j] = plan.forward(b), plan.forward(a)
    return kb, ka


def keygen_relin_bounded(plan: O.RnsPlan, w: int, s, rng: O.Rng):
    s2 = secret_square(plan, s)
    return _keygen_bounded_target(plan, len(plan.moduli), w, s, rng, lambda i, q: s2[i])


def keygen_galois_bounded(qp_plan: O.RnsPlan, q_limbs: int, w: int, s, k, rng: O.Rng):
    s_rot = sigma_ternary(s, k, qp_plan.n)
    return _keygen_bounded_target(qp_plan, q_limbs, w, s, rng, lambda i, q: ternary_to_mod(s_rot, q))


def encrypt_slots(plan: O.RnsPlan, slots, s, rng: O.Rng, emb: E.Embedding):
    m = E.encode_rns(slots, emb, plan.moduli, SCALE)
    e = rng.gaussian(SIGMA, plan.n)
    a = np.stack([rng.uniform(q, plan.n) for q in plan.moduli])
    b = O.encrypt(plan, m, e, s, a)
    return np.stack([b, a])


def decrypt_slots(plan: O.RnsPlan, ct, s, scale, emb: E.Embedding):
    dec = O.decrypt(plan, ct[0], ct[1], s)
    return E.decode(dec, plan.moduli, scale, emb)


def setup(seed: int = 0):
    """Secret, two encrypted slot vectors, relin and rotation-by-1 keys."""
    rng = O.Rng(seed)
    emb = E.Embedding(N)
    plan = O.RnsPlan(Q, N)
    s = rng.ternary(N)
    srng = np.random.default_rng(seed)
    z0 = srng.uniform(-1, 1, N // 2) + 1j * srng.uniform(-1, 1, N // 2)
    z1 = srng.uniform(-1, 1, N // 2) + 1j * srng.uniform(-1, 1, N // 2)
    ct0 = encrypt_slots(plan, z0, s, rng, emb)
    ct1 = encrypt_slots(plan, z1, s, rng, emb)
    # relinearization with P (singleton blocks, each Q_j <= 2^60 ~ P): the
    # RNS-gadget key without P would leave sum_j d_j e_j ~ 2^67 of noise on
    # a 2^80 scale with the 60-bit base prime
    qp4 = O.RnsPlan(Q + P, N)
    rk = keygen_relin_hybrid(qp4, len(Q), [1] * len(Q), s, rng)
    k = O.rotation_exponent_left(N, 1)
    # after the rescale the ciphertext has 3 limbs; hybrid key over them + P
    qp_plan = O.RnsPlan(Q[:3] + P, N)
    gk = keygen_galois_hybrid(qp_plan, 3, [1, 1, 1], s, k, rng)
    # the reference's own algorithms: bounded relinearization (base 2^20,
    # as bin/bounded_ckks_4096.rs) and the hybrid-QP rotation key with base
    # 2^61 (3 digits over the 3 limbs left after the rescale)
    brk = keygen_relin_bounded(plan, BOUNDED_RELIN_LOG, s, rng)
    bgk = keygen_galois_bounded(qp_plan, 3, BOUNDED_ROT_LOG, s, k, rng)
    return dict(rng=rng, emb=emb, plan=plan, s=s, z0=z0, z1=z1, ct0=ct0, ct1=ct1, rk=rk, gk=gk, k=k,
                qp_plan=qp_plan, qp4=qp4, brk=brk, bgk=bgk)


BOUNDED_RELIN_LOG = 20
BOUNDED_ROT_LOG = 61


def expected_slots(d):
    """FP64 cleartext result: (z0 * z1) rotated left by one slot."""
    return np.roll(d["z0"] * d["z1"], -1)


def scale_after_rescale():
    return SCALE * SCALE / Q[-1]
