"""CPU checks of the C ABI library: it loads, exports every entry point include/sbpx.h declares,
and its host-only generator makes valid graphs of the configured shapes. No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from pkg import sbpx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sbpx.h")).read()
    return sorted(set(re.findall(r"\b(sbpx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    m = sbpx()
    L = m.load()
    decl = declared_symbols()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(L, name), name
    assert sorted(m.EXPORTED) == decl


def test_host_library_loads():
    H = sbpx().load_host()
    assert hasattr(H, "sbp_run")


def test_create_without_gpu_reports_an_error_not_a_crash():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = sbpx().load()
    h = C.c_void_p()
    assert L.sbpx_create(C.byref(h), 0) != 0


def _check_graph(N, out_ptr, out_idx, in_ptr, in_idx):
    E = out_ptr[-1]
    assert in_ptr[-1] == E and len(out_idx) == E and len(in_idx) == E
    src = np.repeat(np.arange(N), np.diff(out_ptr))
    assert (src != out_idx).all()  # no self-loops
    key = src.astype(np.int64) * N + out_idx
    assert len(np.unique(key)) == E  # no duplicates
    dst = np.repeat(np.arange(N), np.diff(in_ptr))
    key2 = in_idx.astype(np.int64) * N + dst
    assert (np.sort(key) == np.sort(key2)).all()  # CSC is the transpose of CSR
    for v in range(0, N, max(1, N // 50)):
        assert (np.diff(out_idx[out_ptr[v]:out_ptr[v + 1]]) > 0).all()


def test_generator_small_config_shape_and_determinism():
    m = sbpx()
    cfg = m.CONFIGS["C1"]
    g1 = m.generate_dcsbm(**cfg, seed=0)
    g2 = m.generate_dcsbm(**cfg, seed=0)
    for a, b in zip(g1, g2):
        assert (a == b).all()
    out_ptr, out_idx, in_ptr, in_idx, truth = g1
    N = cfg["num_vertices"]
    _check_graph(N, out_ptr, out_idx, in_ptr, in_idx)
    E = out_ptr[-1]
    assert 15000 < E < 22000  # about 18k edges
    deg_out = np.diff(out_ptr)
    assert deg_out.max() <= 100
    assert set(np.unique(truth)) == set(range(8))
    # p_in = 0.8: most edges stay inside their block
    src = np.repeat(np.arange(N), deg_out)
    inside = (truth[src] == truth[out_idx]).mean()
    assert 0.74 < inside < 0.86
    g3 = m.generate_dcsbm(**cfg, seed=1)
    assert g3[1].shape != g1[1].shape or (g3[1] != g1[1]).any()


def test_generator_medium_config():
    m = sbpx()
    cfg = m.CONFIGS["C2"]
    out_ptr, out_idx, in_ptr, in_idx, truth = m.generate_dcsbm(**cfg, seed=0)
    _check_graph(cfg["num_vertices"], out_ptr, out_idx, in_ptr, in_idx)
    assert 0.85e6 < out_ptr[-1] < 1.3e6  # about 1M edges
    sizes = np.bincount(truth, minlength=44)
    assert sizes.min() >= 1 and sizes.max() > 5 * np.median(sizes)  # Dirichlet(1): uneven sizes


def test_oracle_runs_on_generated_graph():
    m = sbpx()
    out_ptr, out_idx, in_ptr, in_idx, truth = m.generate_dcsbm(300, 4, 0.8, 20.0, 5, 30, 2.1, seed=2)
    g = O.Graph(300, out_ptr, out_idx, in_ptr, in_idx)
    bm = O.Blockmodel(g, 4, truth)
    assert bm.mdl() < O.Blockmodel(g, 300).mdl()
