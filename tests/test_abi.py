"""The C-ABI library loads on a host without a GPU and exports every entry
point that include/ckks_hip.h declares (no compute calls are made)."""
from __future__ import annotations

import ctypes as C
import os

import pytest

from conftest import load_pkg


def test_library_exports_every_header_symbol():
    pkg = load_pkg()
    if not os.path.exists(pkg.LIB_PATH):
        pytest.fail("libckks_hip.so is not built")
    lib = C.CDLL(pkg.LIB_PATH)
    names = pkg.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every header symbol has a ctypes signature in the binding
    assert set(names) <= set(pkg._SIGNATURES), set(names) - set(pkg._SIGNATURES)


def test_host_only_entry_points_work_without_gpu():
    pkg = load_pkg()
    assert pkg.lib().ckks_version().startswith(b"ckks_hip")
    # host prime search matches the oracle's restatement
    from conftest import O
    assert pkg.find_ntt_primes(55, 65536, 5) == O.ntt_primes(55, 65536, 5)
    q, p = pkg.chain_moduli(24, 7)
    assert len(q) == 25 and len(p) == 7 and q[0] not in p
    # errors are reported through status codes and ckks_last_error
    with pytest.raises(pkg.CkksError, match="bits"):
        pkg.find_ntt_primes(70, 16, 1)
