"""CPU checks of the C++ driver layer's own control flow (libsbp_host.so, no device work): its
early stop and its search over the number of blocks, on the cases worked by hand from the
reference (finetune.cpp:144-162; sbp.cpp:61-112, blockmodel_triplet.cpp:3-100) that also pin
the oracle, so the product and the oracle are held to the same answers."""
import ctypes as C

import numpy as np

from pkg import sbpx


def _host():
    H = sbpx().load_host()
    f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    H.sbp_early_stop.argtypes = [C.c_int, C.c_int, C.c_double, f64p, C.c_int]
    H.sbp_golden_trace.argtypes = [C.c_int, f64p, i32p, C.c_int]
    return H


def _early_stop(iteration, bracket, initial, deltas):
    return _host().sbp_early_stop(iteration, int(bracket), initial, np.asarray(deltas, np.float64), len(deltas))


def _golden_trace(n, score):
    out = np.zeros(256, np.int32)
    k = _host().sbp_golden_trace(n, np.asarray(score, np.float64), out, len(out))
    assert k > 0
    return list(out[:k - 1]), int(out[k - 1])


def test_driver_early_stop_hand_worked_cases():
    assert _early_stop(1, False, 1000.0, [-5.0, 0.0]) == 1
    assert _early_stop(2, False, 1000.0, [-1.0, -1.0, -1e-9]) == 0
    assert _early_stop(3, False, 1000.0, [-1.0, -0.4, -0.05]) == 1
    assert _early_stop(3, True, 1000.0, [-1.0, -0.4, -0.05]) == 0
    assert _early_stop(3, False, 1000.0, [-0.75, -0.5, -0.25]) == 0
    assert _early_stop(4, False, 1000.0, [-100.0, -0.1, -0.1, -0.1]) == 1
    assert _early_stop(3, True, 1000.0, [0.3, 0.3, 0.3]) == 1


def test_driver_search_over_block_counts_hand_worked_traces():
    score = [(b - 10) ** 2 + 100.0 for b in range(65)]
    assert _golden_trace(64, score) == ([32, 16, 8, 4, 12, 14, 10, 11, 9], 10)
    assert _golden_trace(8, [float(b) for b in range(9)]) == ([4, 2, 1], 1)
    assert _golden_trace(40, [abs(b - 20.0) for b in range(41)]) == ([20, 10, 16, 18, 19], 20)
