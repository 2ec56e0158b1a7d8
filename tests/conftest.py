"""Shared fixtures: the product package (loaded by path), the CPU oracle and
the moduli of the benchmark chains."""
from __future__ import annotations

import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402  (CPU oracle: test infrastructure)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def load_pkg():
    name = "ccmm_rs_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "ccmm-rs_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def oracle():
    return O


# Moduli of the benchmark chains (one 60-bit base prime, 40-bit primes,
# 60-bit special primes), all = 1 mod 2^17 so they serve N <= 2^16.
Q0 = O.ntt_primes(60, 65536, 1)
Q40 = O.ntt_primes(40, 65536, 24)
P60 = O.ntt_primes(60, 65536, 7, skip=Q0)


@pytest.fixture(scope="session")
def primes():
    return {"q0": Q0[0], "q40": Q40, "p60": P60}


@pytest.fixture(scope="session")
def ctx(pkg):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = pkg.Context(0)
    yield c
    c.close()


def rand_residues(rng, moduli, n, polys=1):
    """[polys][limbs][n] uniform canonical residues."""
    out = np.empty((polys, len(moduli), n), dtype=np.uint64)
    for i, q in enumerate(moduli):
        out[:, i, :] = rng.integers(0, q, size=(polys, n), dtype=np.uint64)
    return out
