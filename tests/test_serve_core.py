"""The serving scheduler (serve_core.hpp) compiled for the CPU with a
recording executor: admission against the byte budget and the active cap,
arrival-order fairness, batching of equal ops across requests, program
validation messages and error propagation, without a GPU."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emu")


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    lib = C.CDLL(os.path.join(HERE, "libserve_emu.so"))
    lib.serve_emu_run.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_void_p, C.c_void_p]
    lib.serve_emu_validate.argtypes = [C.c_int, C.c_char_p, C.c_int]
    lib.serve_emu_scenario.argtypes = [C.c_int, C.c_void_p]
    return lib


def scenario(emu, which):
    out = np.zeros(16, dtype=np.int64)
    assert emu.serve_emu_scenario(which, out.ctypes.data) == 0
    return out.tolist()


WAITING, ACTIVE, FINISHING, DONE, CANCELLED, FAILED = range(6)


def test_cancel_waiting_and_active_requests(emu):
    o = scenario(emu, 0)
    assert o[0] == ACTIVE
    assert o[1] == 1 and o[2] == CANCELLED      # waiting: out of the queue at once
    assert o[3] == 1 and o[4] == CANCELLED      # active: dropped at the next step
    assert o[5] == ACTIVE                       # its room admits the next request
    assert o[6] == 0 and o[7] == DONE
    assert o[8] == 1 and o[9] == 2 and o[10] == 0  # everything released
    assert o[11] == 0                           # a finished request cannot be cancelled


def test_allocation_failure_with_nothing_in_flight_fails_the_request(emu):
    o = scenario(emu, 1)
    assert o[0] == 0                            # the loop ends (no endless retry)
    assert o[1] == FAILED and o[2] == DONE
    assert o[3] == 1 and o[4] == 1 and o[5] == 0


def test_allocation_failure_while_others_run_waits(emu):
    o = scenario(emu, 2)
    assert o[0] == ACTIVE and o[1] == WAITING
    assert o[2] == 0 and o[3] == DONE and o[4] == 0 and o[5] == 2


def test_shared_key_is_charged_once(emu):
    o = scenario(emu, 3)
    foot = o[8]
    plaintext = 3 * 8 * 8
    assert o[0] == 2                            # both admitted under 2 x regs + one key
    assert o[1] == 1000 + plaintext
    assert o[2] == 2 * foot + 1000 + plaintext
    assert o[3] == 0 and o[4] == 0 and o[5] == 0
    assert o[6] == 1 and o[7] == 0              # a budget without room for the key rejects


def run(emu, kinds, arrive, budget, cap=0, lag=1, fail=0):
    kinds = np.array(kinds, dtype=np.int32)
    arrive = np.array(arrive, dtype=np.int32)
    out = np.zeros(8, dtype=np.int64)
    order = np.zeros(len(kinds), dtype=np.uint64)
    assert emu.serve_emu_run(len(kinds), kinds.ctypes.data, arrive.ctypes.data, budget, cap, lag, fail,
                             out.ctypes.data, order.ctypes.data) == 0
    return dict(completed=out[0], peak_active=out[1], launches=out[2], members=out[3], steps=out[4],
                peak_jobs=out[5], rc=out[6], order=order.tolist())


def test_same_program_batches_into_one_launch_per_op(emu):
    r = run(emu, [0, 0, 0, 0], [0, 0, 0, 0], budget=8)
    assert r["completed"] == 4 and r["peak_active"] == 4
    assert r["launches"] == 4 and r["members"] == 16  # 4 ops, each one launch for all 4 jobs


def test_budget_and_cap_bound_admission_and_keep_arrival_order(emu):
    r = run(emu, [0] * 6, [0] * 6, budget=2)
    assert r["completed"] == 6 and r["peak_active"] <= 2 and r["peak_jobs"] <= 2
    assert r["order"] == sorted(r["order"])  # first come, first served
    r = run(emu, [0] * 6, [0] * 6, budget=6, cap=3)
    assert r["peak_active"] == 3 and r["completed"] == 6


def test_late_arrivals_join_requests_in_flight(emu):
    # job 2 (add, copy) arrives at step 2 while job 0 is at its add: the two
    # adds (same op, same level) go out as one launch
    r = run(emu, [0, 0, 1], [0, 1, 2], budget=8)
    assert r["completed"] == 3
    assert r["members"] == 4 + 4 + 2
    assert r["launches"] == r["members"] - 1


def test_executor_error_stops_the_step(emu):
    r = run(emu, [0, 0], [0, 0], budget=4, fail=2)  # plaintext multiply fails
    assert r["rc"] == 7 and r["completed"] == 0


@pytest.mark.parametrize("which,msg", [
    (0, "rotation needs a key and an odd exponent"),
    (1, "rotation needs a key and an odd exponent"),
    (2, "addition needs two ciphertexts at one level"),
    (3, "output register outside the register file"),
    (4, "cannot leave the final level"),
    (5, "key basis must match the ciphertext level"),
    (6, ""),
    (7, "op needs an NTT-domain ciphertext"),
    (8, "op needs a coefficient-domain ciphertext"),
    (9, "output register must be coefficient domain"),
    (10, "op needs an NTT-domain ciphertext"),
])
def test_validation_messages(emu, which, msg):
    buf = C.create_string_buffer(256)
    emu.serve_emu_validate(which, buf, 256)
    assert buf.value.decode().startswith(msg)
