"""Pins the oracle's host samplers to the published algorithms the
reference's samplers are built from (rand_core 0.6.4, rand_chacha 0.3.1,
rand 0.8.8 -- pinned in the reference's Cargo.lock, absent here), without
running the reference:
- PCG32 (rand_core's seed_from_u64 expansion): the XSH-RR output and the LCG
  multiplier reproduce PCG's published demo stream (pcg32_srandom_r(42, 54));
- ChaCha20 stream layout: with a zero key the word stream is RFC 8439 A.1's
  keystream test vectors #1 and #2 (blocks 0 and 1, 64-bit counter from 0);
  next_u64 pairs consecutive words across the 64-word buffer refill the way
  rand_core's BlockRng does (low word first, no word skipped);
- gen_range(0_u8..3) (SecretKey::generate_with_rng, rlwe/secret_key.rs:24-38):
  exhaustive census of all 2^32 draws: one draw rejected, the three values
  exactly equally likely;
- sample_uniform_residues (rlwe/uniform.rs:12-30): the rejection threshold
  leaves a multiple of q accepted draws (exactly uniform);
- DiscreteGaussianTable (rlwe/error.rs:58-108): every weight is exp of the
  (IEEE-exact) argument to within one ulp of a 60-digit evaluation, and the
  first 200 000 draws of a seed lie at least 1e-9 away from every decision
  boundary, so a last-ulp difference of f64::exp between libm versions
  (about 1e-14 on the cumulative sums) cannot change any of those samples."""
from __future__ import annotations

import decimal
import math

import numpy as np

from conftest import O

PCG_DEMO = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]

# RFC 8439 appendix A.1, test vectors #1 (block counter 0) and #2 (counter
# 1): zero key, zero nonce; keystream bytes as little-endian words
RFC_TV1 = bytes.fromhex("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
                        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
RFC_TV2 = bytes.fromhex("9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
                        "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")


def test_pcg32_matches_the_published_demo_stream():
    mask = (1 << 64) - 1
    inc = (54 << 1) | 1
    state = O.pcg32_step(0, inc)
    state = O.pcg32_step((state + 42) & mask, inc)
    out = []
    for _ in range(6):
        out.append(O.pcg32_xsh_rr(state))  # pcg32_random_r outputs the old state
        state = O.pcg32_step(state, inc)
    assert out == PCG_DEMO


def test_chacha_stream_is_the_rfc8439_keystream():
    r = O.Rng.with_key([0] * 8)
    words = [r.next_u32() for _ in range(32)]
    want = list(np.frombuffer(RFC_TV1 + RFC_TV2, dtype="<u4"))
    assert words == [int(v) for v in want]
    # the block function alone, at counter 1
    assert O.chacha20_block([0] * 8, 1).tolist() == [int(v) for v in np.frombuffer(RFC_TV2, dtype="<u4")]


def test_next_u64_pairs_words_across_the_refill():
    # a flat u32 stream against u64 reads at every buffer offset, including
    # the draw that spans word 63 and the next buffer's word 0
    for skip in (0, 1, 61, 62, 63):
        a, b = O.Rng(99), O.Rng(99)
        flat = [a.next_u32() for _ in range(skip + 2 * 80)]
        for _ in range(skip):
            b.next_u32()
        pairs = [b.next_u64() for _ in range(80)]
        want = [flat[skip + 2 * i] | (flat[skip + 2 * i + 1] << 32) for i in range(80)]
        assert pairs == want, skip


def test_seed_expansion_uses_pcg32_outputs_as_key():
    # seed_from_u64: the key words are the PCG32 outputs of the advanced
    # states; the stream is then block 0 of that key
    inc = 11634580027462260723
    state, key = 7, []
    for _ in range(8):
        state = O.pcg32_step(state, inc)
        key.append(O.pcg32_xsh_rr(state))
    r = O.Rng(7)
    assert [r.next_u32() for _ in range(16)] == O.chacha20_block(key, 0).tolist()


def test_ternary_range_zone_exhaustive():
    zero, one, two, rejected = O.ternary_census()
    assert rejected == (2 ** 32 - 3) % 3 == 1
    assert zero == one == two == (2 ** 32 - 1) // 3


def test_uniform_residue_threshold_is_exact():
    for q in [3, 97, 12289, (1 << 40) - 87, 1152921504606584833, (1 << 62) - 57, (1 << 64) - 59]:
        threshold = (2 ** 64 - q) % q
        assert (2 ** 64 - threshold) % q == 0  # accepted draws: whole copies of Z_q
    r = O.Rng(5)
    q = 1152921504606584833
    v = r.uniform(q, 2000)
    assert int(v.max()) < q


def test_gaussian_table_weights_and_decision_margins():
    decimal.getcontext().prec = 60
    sigma = 3.2
    w, total = O.gaussian_table(sigma)
    tail = math.ceil(16 * sigma)
    assert len(w) == 2 * tail + 1
    for i, x in enumerate(range(-tail, tail + 1)):
        arg = -(float(x) * float(x)) / (2.0 * sigma * sigma)  # IEEE double, as the reference
        exact = decimal.Decimal(arg).exp()
        ulp = math.ulp(float(w[i]))
        assert abs(decimal.Decimal(float(w[i])) - exact) <= decimal.Decimal(ulp), x
    cum = np.cumsum(w)  # left to right, as the table walk
    assert abs(cum[-1] - total) <= 1e-12
    r = O.Rng(0)
    targets = np.array([r.uniform_open_unit() for _ in range(200_000)]) * total
    idx = np.searchsorted(cum, targets, side="right")
    lo = np.abs(targets - cum[np.clip(idx - 1, 0, len(cum) - 1)])
    hi = np.abs(cum[np.clip(idx, 0, len(cum) - 1)] - targets)
    assert min(lo.min(), hi.min()) > 1e-9
    # the oracle's samples are those decisions
    got = O.Rng(0).gaussian(sigma, 200_000)
    assert (got == idx - tail).all()
