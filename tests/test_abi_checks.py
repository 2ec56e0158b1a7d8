"""The Python binding refuses tensors the kernels cannot take (host memory,
a wrong dtype, non-contiguous views, buffers too short for the call) with
ValueError, before any pointer reaches the library.  These run on the CPU:
the objects are built without a device and no compute call is made."""
from __future__ import annotations

import pytest
import torch

from conftest import load_pkg

N = 1024
MODULI = [1152921504606584833, 1099511922689, 1099512004609]


class _Dev(torch.Tensor):
    """A host tensor that reports itself as device memory, so the checks
    after the device check can be reached; no test lets such a tensor get
    past every check (the call would then reach the library)."""

    @property
    def is_cuda(self):
        return True


def dev(*shape, dtype=torch.int64):
    return torch.zeros(*shape, dtype=dtype).as_subclass(_Dev)


def _plan(pkg):
    plan = object.__new__(pkg.PreparedRnsNttPlan)
    plan.ctx = None
    plan.moduli = list(MODULI)
    plan.degree = N
    plan._h = None
    return plan


def _key(pkg, q_limbs=3):
    key = object.__new__(pkg.HybridKeySwitchKey)
    key.q_limbs = q_limbs
    key.plan = _plan(pkg)
    key.ctx = None
    key._h = None
    return key


def test_ptr_rejects_bad_tensors():
    pkg = load_pkg()
    good = torch.zeros(8, dtype=torch.int64)
    with pytest.raises(ValueError, match="GPU"):
        pkg._ptr(good)
    with pytest.raises(ValueError, match="int64"):
        pkg._ptr(torch.zeros(8, dtype=torch.int32))
    with pytest.raises(ValueError, match="contiguous"):
        pkg._ptr(torch.zeros(4, 4, dtype=torch.int64).t())
    with pytest.raises(ValueError, match="9 needed"):
        pkg._ptr(good, 9)
    with pytest.raises(ValueError, match="int8"):
        pkg._ptr(good, dtype="int8", name="secret")
    with pytest.raises(ValueError, match="torch tensor"):
        pkg._ptr([1, 2, 3])
    assert pkg._ptr(None) is None


def test_plan_wrappers_reject_before_the_library():
    pkg = load_pkg()
    plan = _plan(pkg)
    x = torch.zeros(2 * 3 * N, dtype=torch.int64)
    # host tensors of the right size and dtype: the device check fires
    with pytest.raises(ValueError, match="GPU"):
        plan.forward(x)
    # a short output buffer
    xd = dev(2 * 3 * N)
    with pytest.raises(ValueError, match="out holds"):
        plan.add(xd, xd, out=dev(3 * N))
    with pytest.raises(ValueError, match="y holds"):
        plan.pointwise_mul(xd, dev(3 * N), out=dev(2 * 3 * N))
    # a batch that is not a whole number of polynomials
    with pytest.raises(ValueError, match="multiple"):
        plan.forward(torch.zeros(3 * N + 1, dtype=torch.int64))
    with pytest.raises(ValueError, match="int64"):
        plan.neg(x.to(torch.int32), out=dev(2 * 3 * N))
    # rescale output one limb short per ciphertext
    with pytest.raises(ValueError, match="out holds"):
        plan.rescale_to_next(dev(2 * 3 * N), 3, out=dev(2 * 1 * N))
    with pytest.raises(ValueError, match="multiple"):
        plan.rescale_to_next(torch.zeros(5 * N, dtype=torch.int64), 3)
    # the secret of encrypt/decrypt is int8
    with pytest.raises(ValueError, match="int8"):
        plan.decrypt_raw(dev(3 * N), dev(3 * N), dev(N), out=dev(3 * N))


def test_key_switch_wrappers_check_out_and_second_operand():
    pkg = load_pkg()
    key = _key(pkg)
    ct = dev(2 * 2 * 3 * N)  # two ciphertexts
    short = dev(2 * 3 * N)   # room for one
    for call in (lambda: key.switch(ct, out=short),
                 lambda: key.apply_galois(ct, 5, out=short)):
        with pytest.raises(ValueError, match="out holds"):
            call()
    with pytest.raises(ValueError, match="ct1 holds"):
        key.multiply_relinearize_rescale(ct, short, out=dev(2 * 2 * 2 * N))
    with pytest.raises(ValueError, match="out holds"):
        key.multiply_relinearize_rescale(ct, ct, out=dev(2 * 2 * 1 * N))
    with pytest.raises(ValueError, match="mismatch"):
        key.relinearize(dev(4 * 3 * N))  # not a whole number of [3][L][N] ciphertexts
    with pytest.raises(ValueError, match="GPU"):
        key.switch(torch.zeros(2 * 2 * 3 * N, dtype=torch.int64), out=dev(2 * 2 * 3 * N))


def test_serving_submit_checks_buffer_sizes():
    pkg = load_pkg()
    from importlib import import_module
    serving = import_module(pkg.__name__ + ".serving")
    prog = serving.Program(n_regs=2, out_reg=1)
    prog.rescale(1, 0)
    # the output register drops one limb: 2 x 2 x N residues
    assert prog.register_limbs(3)[-1] == (3, 2)
    srv = object.__new__(serving.Server)
    srv._h = None
    srv._jobs = {}
    with pytest.raises(ValueError, match="output holds"):
        srv.submit(_plan(pkg), 3, dev(2 * 3 * N), dev(2 * 1 * N), prog)
    with pytest.raises(ValueError, match="inputs holds"):
        srv.submit(_plan(pkg), 3, dev(2 * 2 * N), dev(2 * 2 * N), prog)
