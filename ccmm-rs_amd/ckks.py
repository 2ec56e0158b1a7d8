"""Level-aware CKKS evaluation on the device: the host-side mirror of the
reference's state and dispatch layer.

  ModulusChain           ring/modulus_chain.rs:19-106 (level i = prefix of
                         length L - i of the top basis)
  CkksChainState         ckks/chain.rs:9-121 (level, basis, scale)
  RnsCkksCiphertext      ckks/rns_ciphertext.rs:124-175 (device data
                         [2][limbs][N], coefficient domain, plus its state)
  RnsCkksLevelKeys,
  RnsCkksEvaluationKeys  ckks/evaluation_keys.rs:13-151 (keys by level and
                         Galois exponent)
  RnsCkksEvaluator       ckks/evaluator.rs:17-216 (add, sub, multiply,
                         rotate_left/right, conjugate) plus rescale,
                         mod_switch and multiply_plain
  ExecutionTrace         execution/mod.rs:13-266 (operation accounting)

Every arithmetic step runs in libckks_hip.so; this module only checks the
reference's preconditions (raising CkksAssertion with the reference's panic
text) and carries the scale and level bookkeeping.
"""
from __future__ import annotations

import enum
from dataclasses import dataclass
from typing import Optional

from . import (PreparedRnsNttPlan, _KeySwitchKey, empty)


class CkksAssertion(AssertionError):
    """A precondition of the reference failed; the message is the
    reference's own panic text."""


def _require(cond: bool, msg: str):
    if not cond:
        raise CkksAssertion(msg)


# ---------------------------------------------------------------- tracing

class ExecutionDevice(enum.Enum):
    CPU = "Cpu"
    GPU = "Gpu"
    SIMULATOR = "Simulator"


class ExecutionEventKind(enum.Enum):
    EBLAS_DOT = "EblasDot"
    EBLAS_GEMV = "EblasGemv"
    EBLAS_GEMM = "EblasGemm"
    CPMM = "Cpmm"
    CCMM_SCALAR = "CcmmScalar"
    CCMM_STRUCTURED = "CcmmStructured"
    CIPHERTEXT_PLAINTEXT_MULTIPLY = "CiphertextPlaintextMultiply"
    CIPHERTEXT_CIPHERTEXT_MULTIPLY = "CiphertextCiphertextMultiply"
    ADD = "Add"
    RELINEARIZE = "Relinearize"
    RESCALE = "Rescale"
    MOD_SWITCH = "ModSwitch"
    AUTOMORPHISM = "Automorphism"
    ROTATE = "Rotate"
    NTT_FORWARD = "NttForward"
    NTT_INVERSE = "NttInverse"


@dataclass(frozen=True)
class ExecutionEvent:
    kind: ExecutionEventKind
    ring_degree: Optional[int] = None
    level: Optional[int] = None
    level_after: Optional[int] = None
    rns_limbs: Optional[int] = None
    rns_limbs_after: Optional[int] = None


class ExecutionTrace:
    """Ordered record of the operations an evaluator ran (execution/mod.rs:
    ExecutionTrace::{new, record, events, len, count})."""

    def __init__(self, device: ExecutionDevice = ExecutionDevice.GPU):
        self.device = device
        self._events: list[ExecutionEvent] = []

    def record(self, event: ExecutionEvent):
        self._events.append(event)

    def events(self) -> list[ExecutionEvent]:
        return list(self._events)

    def __len__(self):
        return len(self._events)

    def is_empty(self) -> bool:
        return not self._events

    def count(self, kind: ExecutionEventKind) -> int:
        return sum(1 for e in self._events if e.kind == kind)


_PROFILE_FIELDS = {
    ExecutionEventKind.CIPHERTEXT_PLAINTEXT_MULTIPLY: "ciphertext_plaintext_multiplies",
    ExecutionEventKind.CIPHERTEXT_CIPHERTEXT_MULTIPLY: "ciphertext_ciphertext_multiplies",
    ExecutionEventKind.ADD: "additions",
    ExecutionEventKind.RELINEARIZE: "relinearizations",
    ExecutionEventKind.RESCALE: "rescales",
    ExecutionEventKind.MOD_SWITCH: "modulus_switches",
    ExecutionEventKind.AUTOMORPHISM: "automorphisms",
    ExecutionEventKind.ROTATE: "rotations",
    ExecutionEventKind.NTT_FORWARD: "ntt_forwards",
    ExecutionEventKind.NTT_INVERSE: "ntt_inverses",
}


@dataclass
class ExecutionProfile:
    """ExecutionProfile (execution/mod.rs:162-230): O(1) operation counts.
    The device fills one per context (Context.execution_profile); the two
    last fields have no reference counterpart: every key switch, and the
    single-limb transforms behind the NTT counts."""
    ciphertext_plaintext_multiplies: int = 0
    ciphertext_ciphertext_multiplies: int = 0
    additions: int = 0
    relinearizations: int = 0
    rescales: int = 0
    modulus_switches: int = 0
    automorphisms: int = 0
    rotations: int = 0
    ntt_forwards: int = 0
    ntt_inverses: int = 0
    key_switches: int = 0
    ntt_limb_transforms: int = 0

    def count(self, kind: ExecutionEventKind) -> int:
        """The count of an event kind; the eBLAS / CCMM kinds have none (0)."""
        f = _PROFILE_FIELDS.get(kind)
        return getattr(self, f) if f else 0

    def total_operations(self) -> int:
        """Sum of the reference's ten counted kinds."""
        return sum(getattr(self, f) for f in _PROFILE_FIELDS.values())

    def __sub__(self, other: "ExecutionProfile") -> "ExecutionProfile":
        return ExecutionProfile(**{k: getattr(self, k) - getattr(other, k) for k in self.__dataclass_fields__})

    @classmethod
    def from_trace(cls, trace: ExecutionTrace) -> "ExecutionProfile":
        """The counts of a trace's events."""
        return cls(**{f: trace.count(k) for k, f in _PROFILE_FIELDS.items()})


# ---------------------------------------------------------------- chain

class ModulusChain:
    def __init__(self, top_moduli):
        top = [int(q) for q in top_moduli]
        _require(len(top) > 0, "modulus chain requires a nonempty top basis")
        self.levels = [tuple(top[:length]) for length in range(len(top), 0, -1)]

    @classmethod
    def from_top_basis(cls, top_moduli) -> "ModulusChain":
        return cls(top_moduli)

    def level(self, index: int) -> tuple:
        return self.levels[index]

    def top(self) -> tuple:
        return self.levels[0]

    def max_level(self) -> int:
        return len(self.levels) - 1

    def has_next_level(self, level: int) -> bool:
        return level + 1 < len(self.levels)

    def dropped_modulus(self, level: int) -> Optional[int]:
        if level + 1 >= len(self.levels):
            return None
        return self.levels[level][-1]


@dataclass(frozen=True)
class CkksChainState:
    level: int
    basis: tuple
    scale: float

    @classmethod
    def new(cls, chain: ModulusChain, level: int, scale: float) -> "CkksChainState":
        _require(level <= chain.max_level(), "CKKS level must exist in the modulus chain")
        _require(scale == scale and scale not in (float("inf"),) and scale > 0.0,
                 "CKKS scale must be positive and finite")
        return cls(level, chain.level(level), float(scale))

    @classmethod
    def top(cls, chain: ModulusChain, scale: float) -> "CkksChainState":
        return cls.new(chain, 0, scale)

    def can_rescale(self, chain: ModulusChain) -> bool:
        return chain.has_next_level(self.level)

    def rescale_divisor(self, chain: ModulusChain) -> Optional[int]:
        self.assert_matches_chain(chain)
        return chain.dropped_modulus(self.level)

    def after_rescale(self, chain: ModulusChain) -> "CkksChainState":
        self.assert_matches_chain(chain)
        divisor = self.rescale_divisor(chain)
        _require(divisor is not None, "cannot rescale the final CKKS chain level")
        return CkksChainState.new(chain, self.level + 1, self.scale / float(divisor))

    def after_multiply(self, rhs: "CkksChainState", chain: ModulusChain) -> "CkksChainState":
        self.assert_matches_chain(chain)
        rhs.assert_matches_chain(chain)
        _require(self.level == rhs.level, "CKKS multiplication requires matching levels")
        _require(self.basis == rhs.basis, "CKKS multiplication requires matching modulus bases")
        return CkksChainState.new(chain, self.level, self.scale * rhs.scale)

    def assert_matches_chain(self, chain: ModulusChain):
        _require(self.level <= chain.max_level(), "CKKS level must exist in the modulus chain")
        _require(self.basis == chain.level(self.level), "CKKS basis must match its modulus-chain level")


class RnsCkksCiphertext:
    """Device ciphertext [2][limbs][N] (coefficient domain) and its state."""

    def __init__(self, data, state: CkksChainState, chain: ModulusChain, degree: int):
        state.assert_matches_chain(chain)
        limbs = data.numel() // (2 * degree)
        _require(limbs == len(state.basis), "RNS CKKS ciphertext basis must match CKKS chain state")
        self.data = data.view(2, limbs, degree)
        self.state = state
        self.degree = degree

    @property
    def level(self) -> int:
        return self.state.level

    @property
    def basis(self) -> tuple:
        return self.state.basis

    @property
    def scale(self) -> float:
        return self.state.scale

    def assert_matches_chain(self, chain: ModulusChain):
        self.state.assert_matches_chain(chain)


# ---------------------------------------------------------------- keys

class RnsCkksLevelKeys:
    def __init__(self, level: int, basis):
        self.level = level
        self.basis = tuple(int(q) for q in basis)
        self.multiplication_key: Optional[_KeySwitchKey] = None
        self.galois_keys: dict[int, _KeySwitchKey] = {}

    @staticmethod
    def _key_basis(key: _KeySwitchKey) -> tuple:
        return tuple(key.plan.moduli[:key.q_limbs])

    def set_multiplication_key(self, key: _KeySwitchKey):
        _require(self._key_basis(key) == self.basis,
                 "RNS multiplication key basis must match level-key basis")
        self.multiplication_key = key

    def insert_galois_key(self, exponent: int, key: _KeySwitchKey):
        _require(self._key_basis(key) == self.basis, "RNS Galois key basis must match level-key basis")
        _require(exponent not in self.galois_keys, "RNS Galois key exponent already exists at this level")
        self.galois_keys[exponent] = key

    def galois_key(self, exponent: int) -> Optional[_KeySwitchKey]:
        return self.galois_keys.get(exponent)

    def assert_matches_state(self, state: CkksChainState):
        _require(self.level == state.level, "evaluation-key level must match CKKS ciphertext level")
        _require(self.basis == state.basis, "evaluation-key basis must match CKKS ciphertext basis")


class RnsCkksEvaluationKeys:
    def __init__(self):
        self.levels: dict[int, RnsCkksLevelKeys] = {}

    def insert_level(self, level_keys: RnsCkksLevelKeys):
        _require(level_keys.level not in self.levels, "evaluation keys already exist for CKKS level")
        self.levels[level_keys.level] = level_keys

    def level(self, level: int) -> Optional[RnsCkksLevelKeys]:
        return self.levels.get(level)

    def for_state(self, state: CkksChainState) -> RnsCkksLevelKeys:
        keys = self.levels.get(state.level)
        _require(keys is not None, "evaluation keys are missing for active CKKS level")
        keys.assert_matches_state(state)
        return keys

    def multiplication_for(self, state: CkksChainState) -> _KeySwitchKey:
        key = self.for_state(state).multiplication_key
        _require(key is not None, "multiplication key is missing for active CKKS level")
        return key

    def galois_for(self, state: CkksChainState, exponent: int) -> _KeySwitchKey:
        key = self.for_state(state).galois_key(exponent)
        _require(key is not None, "Galois key is missing for requested exponent at active CKKS level")
        return key

    def level_count(self) -> int:
        return len(self.levels)

    def is_empty(self) -> bool:
        return not self.levels


@dataclass
class RnsCkksLevelKeyStorage:
    level: int = 0
    multiplication_bytes: int = 0
    galois_key_count: int = 0
    galois_bytes: int = 0
    total_resident_bytes: int = 0


@dataclass
class RnsCkksEvaluationKeyStorage:
    levels: list = None
    multiplication_bytes: int = 0
    galois_key_count: int = 0
    galois_bytes: int = 0
    total_resident_bytes: int = 0

    def to_report_string(self) -> str:
        """Deterministic text report (RnsCkksEvaluationKeyStorage::
        to_report_string, ckks/key_accounting.rs:346-395; the hybrid-policy
        columns of the reference are zero here: its Grafting hybrid policy is
        not on this path)."""
        out = ["CKKS_EVALUATION_KEY_STORAGE"]
        for lv in self.levels or []:
            out.append(f"level={lv.level} multiplication_bytes={lv.multiplication_bytes} "
                       f"galois_key_count={lv.galois_key_count} galois_bytes={lv.galois_bytes} "
                       f"hybrid_base_bytes=0 prepared_cache_bytes=0 hybrid_prepared_total_bytes=0 "
                       f"total_resident_bytes={lv.total_resident_bytes}")
        out.append(f"TOTAL multiplication_bytes={self.multiplication_bytes} "
                   f"galois_key_count={self.galois_key_count} galois_bytes={self.galois_bytes} "
                   f"hybrid_base_bytes=0 prepared_cache_bytes=0 hybrid_prepared_total_bytes=0 "
                   f"total_resident_bytes={self.total_resident_bytes}")
        return "\n".join(out) + "\n"


def evaluation_key_storage(keys: RnsCkksEvaluationKeys, key_bytes=None) -> RnsCkksEvaluationKeyStorage:
    """evaluation_key_storage (ckks/key_accounting.rs:276-344): resident key
    bytes per level and in total, in level order.  key_bytes(key) gives a
    key's bytes (default: the device figure, key.nbytes)."""
    key_bytes = key_bytes or (lambda k: k.nbytes)
    rep = RnsCkksEvaluationKeyStorage(levels=[])
    for level in sorted(keys.levels):
        lk = keys.levels[level]
        mb = key_bytes(lk.multiplication_key) if lk.multiplication_key is not None else 0
        gb = sum(key_bytes(k) for k in lk.galois_keys.values())
        lv = RnsCkksLevelKeyStorage(level, mb, len(lk.galois_keys), gb, mb + gb)
        rep.levels.append(lv)
        rep.multiplication_bytes += mb
        rep.galois_key_count += lv.galois_key_count
        rep.galois_bytes += gb
        rep.total_resident_bytes += lv.total_resident_bytes
    return rep


# ---------------------------------------------------------------- evaluator

def rotation_exponent_left(degree: int, steps: int) -> int:
    """ckks/rns_automorphism.rs:179-195: 5^steps mod 2N."""
    return pow(5, steps, 2 * degree)


def rotation_exponent_right(degree: int, steps: int) -> int:
    """ckks/rns_automorphism.rs:197-207: inverse of the left exponent."""
    return pow(rotation_exponent_left(degree, steps), -1, 2 * degree)


def conjugation_exponent(degree: int) -> int:
    """ckks/rns_automorphism.rs:233-240."""
    return 2 * degree - 1


class RnsCkksEvaluator:
    """RnsCkksEvaluator (ckks/evaluator.rs:17-216) running on the device.
    `plan` is a PreparedRnsNttPlan over the chain's top basis (every level is
    a prefix of it)."""

    def __init__(self, chain: ModulusChain, keys: RnsCkksEvaluationKeys, plan: PreparedRnsNttPlan,
                 trace: Optional[ExecutionTrace] = None, device_events: bool = False):
        _require(tuple(plan.moduli) == chain.top(), "prepared RNS NTT plan basis must match the chain top basis")
        self.chain, self.keys, self.plan, self.trace = chain, keys, plan, trace
        # device_events: after each operation's event, the NTT transforms the
        # device ran for it (from the context's execution profile)
        self.device_events = device_events
        self._seen = plan.ctx.execution_profile() if (trace is not None and device_events) else None

    def _record(self, kind, ct: RnsCkksCiphertext, after: Optional[RnsCkksCiphertext] = None):
        if self.trace is None:
            return
        after = after or ct
        self.trace.record(ExecutionEvent(kind, ring_degree=ct.degree, level=ct.level,
                                         level_after=after.level, rns_limbs=len(ct.basis),
                                         rns_limbs_after=len(after.basis)))
        if self._seen is not None:
            now = self.plan.ctx.execution_profile()
            d = now - self._seen
            self._seen = now
            for ek, n in ((ExecutionEventKind.NTT_FORWARD, d.ntt_forwards),
                          (ExecutionEventKind.NTT_INVERSE, d.ntt_inverses)):
                for _ in range(n):
                    self.trace.record(ExecutionEvent(ek, ring_degree=ct.degree, rns_limbs=len(ct.basis)))

    def _wrap(self, data, state) -> RnsCkksCiphertext:
        return RnsCkksCiphertext(data, state, self.chain, self.plan.degree)

    def encrypt_state(self, data, scale: float, level: int = 0) -> RnsCkksCiphertext:
        return self._wrap(data, CkksChainState.new(self.chain, level, scale))

    def _binary(self, lhs, rhs, what):
        lhs.assert_matches_chain(self.chain)
        rhs.assert_matches_chain(self.chain)
        _require(lhs.level == rhs.level, f"RNS CKKS {what} requires matching levels")
        _require(lhs.basis == rhs.basis, f"RNS CKKS {what} requires matching bases")
        _require(lhs.scale == rhs.scale, f"RNS CKKS {what} requires matching scales")

    def add(self, lhs: RnsCkksCiphertext, rhs: RnsCkksCiphertext) -> RnsCkksCiphertext:
        self._binary(lhs, rhs, "addition")
        out = self.plan.add(lhs.data, rhs.data, n_limbs=len(lhs.basis))
        res = self._wrap(out, lhs.state)
        self._record(ExecutionEventKind.ADD, lhs, res)
        return res

    def sub(self, lhs: RnsCkksCiphertext, rhs: RnsCkksCiphertext) -> RnsCkksCiphertext:
        self._binary(lhs, rhs, "subtraction")
        out = self.plan.sub(lhs.data, rhs.data, n_limbs=len(lhs.basis))
        return self._wrap(out, lhs.state)

    def multiply(self, lhs: RnsCkksCiphertext, rhs: RnsCkksCiphertext) -> RnsCkksCiphertext:
        """multiply_with_evaluation_keys (ckks/evaluation_keys.rs:158-183):
        tensor, relinearize with the level's key, rescale."""
        lhs.assert_matches_chain(self.chain)
        rhs.assert_matches_chain(self.chain)
        _require(lhs.level == rhs.level, "automatic CKKS multiplication requires matching levels")
        _require(lhs.basis == rhs.basis, "automatic CKKS multiplication requires matching bases")
        key = self.keys.multiplication_for(lhs.state)
        state = lhs.state.after_multiply(rhs.state, self.chain).after_rescale(self.chain)
        out = key.multiply_relinearize_rescale(lhs.data, rhs.data)
        res = self._wrap(out, state)
        self._record(ExecutionEventKind.CIPHERTEXT_CIPHERTEXT_MULTIPLY, lhs, res)
        self._record(ExecutionEventKind.RELINEARIZE, lhs)
        self._record(ExecutionEventKind.RESCALE, lhs, res)
        return res

    def _galois(self, ct: RnsCkksCiphertext, exponent: int, kind) -> RnsCkksCiphertext:
        ct.assert_matches_chain(self.chain)
        key = self.keys.galois_for(ct.state, exponent)
        out = key.apply_galois(ct.data, exponent)
        res = self._wrap(out, ct.state)
        self._record(kind, ct, res)
        return res

    def rotate_left(self, ct: RnsCkksCiphertext, steps: int) -> RnsCkksCiphertext:
        return self._galois(ct, rotation_exponent_left(ct.degree, steps), ExecutionEventKind.ROTATE)

    def rotate_right(self, ct: RnsCkksCiphertext, steps: int) -> RnsCkksCiphertext:
        return self._galois(ct, rotation_exponent_right(ct.degree, steps), ExecutionEventKind.ROTATE)

    def conjugate(self, ct: RnsCkksCiphertext) -> RnsCkksCiphertext:
        return self._galois(ct, conjugation_exponent(ct.degree), ExecutionEventKind.AUTOMORPHISM)

    def rescale(self, ct: RnsCkksCiphertext) -> RnsCkksCiphertext:
        """rescale_rns_ckks_to_next (ckks/rns_rescale.rs:19-60)."""
        state = ct.state.after_rescale(self.chain)
        out = self.plan.rescale_to_next(ct.data, len(ct.basis))
        res = self._wrap(out, state)
        self._record(ExecutionEventKind.RESCALE, ct, res)
        return res

    def mod_switch(self, ct: RnsCkksCiphertext) -> RnsCkksCiphertext:
        """mod_switch_rns_ckks_to_next (ckks/rns_mod_switch.rs:14-59): scale kept."""
        ct.assert_matches_chain(self.chain)
        _require(self.chain.has_next_level(ct.level), "cannot mod-switch the final CKKS chain level")
        state = CkksChainState.new(self.chain, ct.level + 1, ct.scale)
        out = self.plan.mod_switch_to_next(ct.data, len(ct.basis))
        res = self._wrap(out, state)
        self._record(ExecutionEventKind.MOD_SWITCH, ct, res)
        return res

    def multiply_plain(self, ct: RnsCkksCiphertext, pt_ntt, pt_scale: float) -> RnsCkksCiphertext:
        """multiply_plain_rns_ckks_with_prepared_ntt (ckks/rns_plain_multiply.rs:
        193-260): the plaintext is NTT-domain over the ciphertext's basis."""
        ct.assert_matches_chain(self.chain)
        limbs = len(ct.basis)
        _require(pt_ntt.numel() == limbs * ct.degree,
                 "prepared RNS plaintext basis must match CKKS ciphertext basis")
        state = CkksChainState.new(self.chain, ct.level, ct.scale * pt_scale)
        out = self.plan.multiply_plain(ct.data, pt_ntt, limbs)
        res = self._wrap(out, state)
        self._record(ExecutionEventKind.CIPHERTEXT_PLAINTEXT_MULTIPLY, ct, res)
        return res
