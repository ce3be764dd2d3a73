"""ctypes wrapper over the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: a CPU restatement of the reference's hot path used only to
check the MI355X library. Nothing in the product imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


class OracleStats(C.Structure):
    _fields_ = [("attempts", C.c_long), ("evals", C.c_long), ("moves", C.c_long), ("dS_sum", C.c_double)]


class OracleRun(C.Structure):
    _fields_ = [("B", C.c_int), ("outer_iterations", C.c_int), ("sweeps", C.c_long),
                ("attempts", C.c_long), ("evals", C.c_long), ("moves", C.c_long),
                ("seconds", C.c_double), ("mcmc_seconds", C.c_double),
                ("merge_seconds", C.c_double), ("mdl", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
    L = C.CDLL(ORACLE_SO)
    L.oracle_last_error.restype = C.c_char_p
    L.oracle_early_stop.argtypes = [C.c_int, C.c_int, C.c_double, f64p, C.c_int]
    L.oracle_golden_trace.argtypes = [C.c_int, f64p, i32p, C.c_int]
    L.oracle_philox.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p]
    L.oracle_logf.argtypes = [C.c_float]
    L.oracle_logf.restype = C.c_float
    L.oracle_logf_check.argtypes = [C.c_uint32, C.c_uint32, f32p]
    L.oracle_logf_check.restype = C.c_long
    L.oracle_graph_new.argtypes = [C.c_int, i32p, i32p, i32p, i32p]
    L.oracle_graph_new.restype = C.c_void_p
    L.oracle_graph_free.argtypes = [C.c_void_p]
    L.oracle_graph_E.argtypes = [C.c_void_p]
    L.oracle_graph_E.restype = C.c_long
    L.oracle_graph_nhigh.argtypes = [C.c_void_p]
    L.oracle_graph_split.argtypes = [C.c_void_p, i32p, i32p]
    L.oracle_bm_new.argtypes = [C.c_void_p, C.c_int, i32p]
    L.oracle_bm_new.restype = C.c_void_p
    L.oracle_bm_clone.argtypes = [C.c_void_p]
    L.oracle_bm_clone.restype = C.c_void_p
    L.oracle_bm_free.argtypes = [C.c_void_p]
    for f in ("oracle_bm_B", "oracle_bm_num_to_merge"):
        getattr(L, f).argtypes = [C.c_void_p]
    L.oracle_bm_set_num_to_merge.argtypes = [C.c_void_p, C.c_int]
    L.oracle_bm_state.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.oracle_bm_nnz.argtypes = [C.c_void_p]
    L.oracle_bm_nnz.restype = C.c_long
    L.oracle_bm_dense.argtypes = [C.c_void_p, i32p]
    L.oracle_bm_coo.argtypes = [C.c_void_p, i32p, i32p, i32p]
    L.oracle_mdl.argtypes = [C.c_void_p, C.c_long, C.c_long, C.POINTER(C.c_double)]
    L.oracle_log_posterior.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.oracle_null_mdl_v1.argtypes = [C.c_long]
    L.oracle_null_mdl_v1.restype = C.c_double
    L.oracle_null_mdl_v2.argtypes = [C.c_long, C.c_long]
    L.oracle_null_mdl_v2.restype = C.c_double
    L.oracle_vertex_delta.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, i32p, i32p,
                                      i32p, C.POINTER(C.c_int32)]
    L.oracle_merge_delta.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, i32p, i32p, i32p]
    L.oracle_merge_delta_mdl.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.oracle_eval_moves.argtypes = [C.c_void_p, C.c_void_p, C.c_int, i32p, i32p, f64p, f64p, f64p, u8p]
    L.oracle_apply_moves.argtypes = [C.c_void_p, C.c_void_p, C.c_int, i32p, i32p]
    L.oracle_propose.argtypes = [C.c_void_p, C.c_void_p, C.c_int, i32p, C.c_uint64, C.c_uint32, C.c_int,
                                 i32p, f64p]
    L.oracle_async_sweep.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_uint32, C.c_int,
                                     C.POINTER(OracleStats)]
    L.oracle_mh_pass.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_uint64, C.c_uint32,
                                 C.POINTER(OracleStats)]
    L.oracle_finetune.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64,
                                  C.POINTER(C.c_uint32), C.POINTER(C.c_int), C.POINTER(OracleStats)]
    L.oracle_bm_overall.argtypes = [C.c_void_p]
    L.oracle_bm_overall.restype = C.c_float
    L.oracle_merge_proposals.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, f64p, i32p]
    L.oracle_carry_out.argtypes = [C.c_void_p, C.c_void_p, f64p, i32p, C.c_int]
    L.oracle_merge_blocks.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_uint64, C.c_uint32]
    L.oracle_sbp.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, i32p, C.POINTER(OracleRun)]
    L.oracle_propose_merge_draw.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_uint64, C.c_uint32,
                                            C.POINTER(C.c_int32)]
    L.oracle_learned_features.argtypes = [C.c_void_p, C.c_void_p, C.c_int, i32p, i32p, f32p]
    L.oracle_mixture_hastings.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, f32p, i32p, C.c_double,
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.oracle_hastings_terms.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
    L.oracle_plain_probabilities.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                             C.POINTER(C.c_double)]
    L.oracle_mdl_cells.argtypes = [C.c_long, i32p, i32p, i32p, i32p, i32p, C.c_int, C.c_long, C.c_long,
                                   C.POINTER(C.c_double)]
    L.oracle_sbp_traced.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64, i32p, C.POINTER(OracleRun),
                                    i32p, f32p, C.c_int]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().oracle_last_error().decode())


def csr_from_edges(n, edges):
    """CSR (out) and CSC (in) lists in first-insertion order, duplicates dropped, as
    utils::insert_nodup builds them (utils.cpp:52-60)."""
    out = [[] for _ in range(n)]
    inn = [[] for _ in range(n)]
    seen = set()
    for u, v in edges:
        if (u, v) in seen:
            continue
        seen.add((u, v))
        out[u].append(v)
        inn[v].append(u)
    out_ptr = np.zeros(n + 1, np.int32)
    in_ptr = np.zeros(n + 1, np.int32)
    out_ptr[1:] = np.cumsum([len(x) for x in out])
    in_ptr[1:] = np.cumsum([len(x) for x in inn])
    out_idx = np.array([w for x in out for w in x], np.int32)
    in_idx = np.array([w for x in inn for w in x], np.int32)
    return out_ptr, out_idx, in_ptr, in_idx


class Graph:
    def __init__(self, n, out_ptr, out_idx, in_ptr, in_idx):
        self.n = int(n)
        self.out_ptr = np.ascontiguousarray(out_ptr, np.int32)
        self.out_idx = np.ascontiguousarray(out_idx, np.int32)
        self.in_ptr = np.ascontiguousarray(in_ptr, np.int32)
        self.in_idx = np.ascontiguousarray(in_idx, np.int32)
        self.h = lib().oracle_graph_new(self.n, self.out_ptr, self.out_idx, self.in_ptr, self.in_idx)
        if not self.h:
            raise RuntimeError("oracle: " + lib().oracle_last_error().decode())

    @classmethod
    def from_edges(cls, n, edges):
        return cls(n, *csr_from_edges(n, edges))

    @property
    def E(self):
        return lib().oracle_graph_E(self.h)

    def split(self):
        nh = lib().oracle_graph_nhigh(self.h)
        high = np.zeros(nh, np.int32)
        low = np.zeros(self.n - nh, np.int32)
        lib().oracle_graph_split(self.h, high, low)
        return high, low

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_graph_free(self.h)
            self.h = None


class Blockmodel:
    def __init__(self, graph, B, assignment=None, _handle=None):
        self.g = graph
        if _handle is not None:
            self.h = _handle
            return
        a = np.ascontiguousarray(np.arange(graph.n) if assignment is None else assignment, np.int32)
        self.h = lib().oracle_bm_new(graph.h, int(B), a)
        if not self.h:
            raise RuntimeError("oracle: " + lib().oracle_last_error().decode())

    def clone(self):
        return Blockmodel(self.g, 0, _handle=lib().oracle_bm_clone(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_bm_free(self.h)
            self.h = None

    @property
    def B(self):
        return lib().oracle_bm_B(self.h)

    @property
    def num_to_merge(self):
        return lib().oracle_bm_num_to_merge(self.h)

    @num_to_merge.setter
    def num_to_merge(self, n):
        lib().oracle_bm_set_num_to_merge(self.h, int(n))

    @property
    def overall_entropy(self):
        return lib().oracle_bm_overall(self.h)

    def state(self):
        B = self.B
        a = np.zeros(self.g.n, np.int32)
        dout = np.zeros(B, np.int32)
        din = np.zeros(B, np.int32)
        deg = np.zeros(B, np.int32)
        lib().oracle_bm_state(self.h, a.ctypes.data, dout.ctypes.data, din.ctypes.data, deg.ctypes.data)
        return a, dout, din, deg

    def dense(self):
        B = self.B
        M = np.zeros(B * B, np.int32)
        lib().oracle_bm_dense(self.h, M)
        return M.reshape(B, B)

    def coo(self):
        n = lib().oracle_bm_nnz(self.h)
        r = np.zeros(n, np.int32)
        c = np.zeros(n, np.int32)
        v = np.zeros(n, np.int32)
        _check(lib().oracle_bm_coo(self.h, r, c, v))
        return r, c, v

    def mdl(self):
        out = C.c_double()
        _check(lib().oracle_mdl(self.h, self.g.n, self.g.E, C.byref(out)))
        return out.value

    def log_posterior(self):
        out = C.c_double()
        _check(lib().oracle_log_posterior(self.h, C.byref(out)))
        return out.value

    def vertex_delta(self, v, s):
        cap = 4 * (self.g.n + 4)
        r = np.zeros(cap, np.int32)
        c = np.zeros(cap, np.int32)
        d = np.zeros(cap, np.int32)
        sw = C.c_int32()
        n = lib().oracle_vertex_delta(self.g.h, self.h, v, s, cap, r, c, d, C.byref(sw))
        if n < 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return {(int(r[i]), int(c[i])): int(d[i]) for i in range(n)}, sw.value

    def merge_delta(self, r_, s):
        cap = 4 * (self.B + 4)
        r = np.zeros(cap, np.int32)
        c = np.zeros(cap, np.int32)
        d = np.zeros(cap, np.int32)
        n = lib().oracle_merge_delta(self.h, r_, s, cap, r, c, d)
        if n < 0:
            raise RuntimeError(lib().oracle_last_error().decode())
        return {(int(r[i]), int(c[i])): int(d[i]) for i in range(n)}

    def merge_delta_mdl(self, r, s):
        out = C.c_double()
        _check(lib().oracle_merge_delta_mdl(self.h, r, s, C.byref(out)))
        return out.value

    def eval_moves(self, v, s, u):
        v = np.ascontiguousarray(v, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        u = np.ascontiguousarray(u, np.float64)
        n = len(v)
        dS = np.zeros(n)
        H = np.zeros(n)
        acc = np.zeros(n, np.uint8)
        _check(lib().oracle_eval_moves(self.g.h, self.h, n, v, s, u, dS, H, acc))
        return dS, H, acc

    def apply_moves(self, v, s):
        v = np.ascontiguousarray(v, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        _check(lib().oracle_apply_moves(self.g.h, self.h, len(v), v, s))

    def propose(self, v, seed, call, tag=0):
        v = np.ascontiguousarray(v, np.int32)
        s = np.zeros(len(v), np.int32)
        u = np.zeros(len(v))
        _check(lib().oracle_propose(self.g.h, self.h, len(v), v, seed, call, tag, s, u))
        return s, u

    def async_sweep(self, batches, seed, call, hybrid_low=False):
        st = OracleStats()
        _check(lib().oracle_async_sweep(self.g.h, self.h, batches, seed, call, int(hybrid_low), C.byref(st)))
        return st

    def mh_pass(self, seed, call, order=None):
        st = OracleStats()
        if order is None:
            _check(lib().oracle_mh_pass(self.g.h, self.h, 0, None, seed, call, C.byref(st)))
        else:
            o = np.ascontiguousarray(order, np.int32)
            _check(lib().oracle_mh_pass(self.g.h, self.h, len(o), o.ctypes.data, seed, call, C.byref(st)))
        return st

    def finetune(self, algo, batches, bracket, seed, call):
        c = C.c_uint32(call)
        sweeps = C.c_int()
        st = OracleStats()
        _check(lib().oracle_finetune(self.g.h, self.h, algo, batches, int(bracket), seed, C.byref(c),
                                     C.byref(sweeps), C.byref(st)))
        return c.value, sweeps.value, st

    def merge_proposals(self, seed, call):
        B = self.B
        dS = np.zeros(B)
        best = np.zeros(B, np.int32)
        _check(lib().oracle_merge_proposals(self.h, seed, call, dS, best))
        return dS, best

    def propose_merge_draw(self, r, i, seed, call):
        out = C.c_int32()
        _check(lib().oracle_propose_merge_draw(self.h, r, i, seed, call, C.byref(out)))
        return out.value

    def learned_features(self, v):
        """Learned proposal inputs of vertex v: candidates[32], counts[32], features[64]."""
        cand = np.zeros(32, np.int32)
        cnt = np.zeros(32, np.int32)
        x = np.zeros(64, np.float32)
        _check(lib().oracle_learned_features(self.g.h, self.h, int(v), cand, cnt, x))
        return cand, cnt, x

    def mixture_hastings(self, v, s, z, cand, w=0.2):
        """(H, q_net(s), q_net(r)) of the learned mixture for v -> s with scores z of candidates cand."""
        H, qf, qb = C.c_double(), C.c_double(), C.c_double()
        _check(lib().oracle_mixture_hastings(self.g.h, self.h, int(v), int(s), np.ascontiguousarray(z, np.float32),
                                             np.ascontiguousarray(cand, np.int32), float(w), C.byref(H),
                                             C.byref(qf), C.byref(qb)))
        return H.value, qf.value, qb.value

    def hastings_terms(self, v, s):
        f, b = C.c_double(), C.c_double()
        _check(lib().oracle_hastings_terms(self.g.h, self.h, int(v), int(s), C.byref(f), C.byref(b)))
        return f.value, b.value

    def plain_probabilities(self, v, s):
        """(P(s | now), P(r | after v moves to s)) of the plain proposal sampler."""
        f, b = C.c_double(), C.c_double()
        _check(lib().oracle_plain_probabilities(self.g.h, self.h, int(v), int(s), C.byref(f), C.byref(b)))
        return f.value, b.value

    def carry_out(self, dS, best, advanced=False):
        _check(lib().oracle_carry_out(self.g.h, self.h, np.ascontiguousarray(dS, np.float64),
                                      np.ascontiguousarray(best, np.int32), int(advanced)))

    def merge_blocks(self, approximate, seed, call):
        _check(lib().oracle_merge_blocks(self.g.h, self.h, int(approximate), seed, call))


def sbp(graph, algo, batches, approximate, seed):
    a = np.zeros(graph.n, np.int32)
    run = OracleRun()
    _check(lib().oracle_sbp(graph.h, algo, batches, int(approximate), seed, a, C.byref(run)))
    return a, run


def sbp_traced(graph, algo, batches, approximate, seed, cap=256):
    """As sbp(), plus the per-outer-iteration trace: blocks after each merge phase and the
    float MDL after its sweeps."""
    a = np.zeros(graph.n, np.int32)
    run = OracleRun()
    tb = np.zeros(cap, np.int32)
    tm = np.zeros(cap, np.float32)
    _check(lib().oracle_sbp_traced(graph.h, algo, batches, int(approximate), seed, a, C.byref(run), tb, tm, cap))
    k = min(run.outer_iterations, cap)
    return a, run, tb[:k].copy(), tm[:k].copy()


def mdl_cells(rows, cols, vals, dout, din, B, N, E):
    """entropy::mdl of a block matrix given as nonzero cells and block degrees (FP32 logf rule)."""
    out = C.c_double()
    a = [np.ascontiguousarray(x, np.int32) for x in (rows, cols, vals, dout, din)]
    _check(lib().oracle_mdl_cells(len(a[0]), *a, int(B), int(N), int(E), C.byref(out)))
    return out.value


def null_mdl_v1(E):
    return lib().oracle_null_mdl_v1(int(E))


def null_mdl_v2(N, E):
    return lib().oracle_null_mdl_v2(int(N), int(E))


def philox(seed, c0, c1, c2, c3):
    out = np.zeros(4, np.uint32)
    lib().oracle_philox(seed, c0, c1, c2, c3, out)
    return out


def logf(x):
    return lib().oracle_logf(float(x))
