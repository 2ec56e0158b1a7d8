"""Versioned limb-major binary layout for RNS data: golden vectors, serving
I/O and anything that crosses a process boundary.  The reference has no
wire format (its types are only Clone/Debug); this one carries exactly the
metadata its RnsPolynomial / RnsNttPolynomial / ciphertext types hold.

Little-endian:
    8 bytes  magic  b"CKRNS\\x00\\x00\\x01"  (format version 1 in the last byte)
    u32      kind    0 polynomial batch, 1 ciphertext batch ([b][a] pairs),
                     2 key entries ([digit] pairs, NTT domain)
    u32      domain  0 coefficient, 1 NTT (index k <-> psi^(2k+1))
    u32      degree  N (power of two)
    u32      limbs   number of moduli
    u64      count   polynomials ([limb][N] blocks) in the payload
    f64      scale   CKKS scale, 0 when not meaningful
    u64[limbs]       moduli
    u64[count][limbs][N]  canonical residues
"""
from __future__ import annotations

import io
import struct
from dataclasses import dataclass

import numpy as np

MAGIC = b"CKRNS\x00\x00\x01"
KIND_POLY, KIND_CIPHERTEXT, KIND_KEY = 0, 1, 2
DOMAIN_COEFF, DOMAIN_NTT = 0, 1
_HEADER = struct.Struct("<8sIIIIQd")


class WireError(ValueError):
    pass


@dataclass
class RnsBlob:
    moduli: list
    data: np.ndarray  # uint64 [count][limbs][N]
    kind: int = KIND_POLY
    domain: int = DOMAIN_COEFF
    scale: float = 0.0

    @property
    def degree(self) -> int:
        return int(self.data.shape[-1])


def to_bytes(blob: RnsBlob) -> bytes:
    data = np.ascontiguousarray(blob.data, dtype=np.uint64)
    if data.ndim == 2:
        data = data[None]
    count, limbs, n = data.shape
    if limbs != len(blob.moduli):
        raise WireError("payload limbs must match the moduli")
    if n & (n - 1) or n < 2:
        raise WireError("degree must be a power of two")
    q = np.array(blob.moduli, dtype=np.uint64)
    if np.any(data >= q[None, :, None]):
        raise WireError("residues must be canonical")
    out = io.BytesIO()
    out.write(_HEADER.pack(MAGIC, blob.kind, blob.domain, n, limbs, count, float(blob.scale)))
    out.write(q.astype("<u8").tobytes())
    out.write(data.astype("<u8").tobytes())
    return out.getvalue()


def from_bytes(buf: bytes) -> RnsBlob:
    if len(buf) < _HEADER.size:
        raise WireError("truncated header")
    magic, kind, domain, n, limbs, count, scale = _HEADER.unpack_from(buf, 0)
    if magic[:7] != MAGIC[:7]:
        raise WireError("not an RNS blob")
    if magic[7] != MAGIC[7]:
        raise WireError(f"unsupported format version {magic[7]}")
    if kind not in (KIND_POLY, KIND_CIPHERTEXT, KIND_KEY) or domain not in (DOMAIN_COEFF, DOMAIN_NTT):
        raise WireError("unknown kind or domain")
    off = _HEADER.size
    need = off + 8 * limbs + 8 * count * limbs * n
    if len(buf) != need:
        raise WireError(f"payload size {len(buf)} does not match the header ({need})")
    q = np.frombuffer(buf, dtype="<u8", count=limbs, offset=off).astype(np.uint64)
    data = np.frombuffer(buf, dtype="<u8", count=count * limbs * n, offset=off + 8 * limbs)
    data = data.astype(np.uint64).reshape(count, limbs, n)
    if np.any(data >= q[None, :, None]):
        raise WireError("residues must be canonical")
    return RnsBlob([int(v) for v in q], data, kind, domain, scale)


def save(path, blob: RnsBlob):
    with open(path, "wb") as f:
        f.write(to_bytes(blob))


def load(path) -> RnsBlob:
    with open(path, "rb") as f:
        return from_bytes(f.read())
