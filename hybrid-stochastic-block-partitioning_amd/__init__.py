"""Python access to the MI355X stochastic-block-partitioning library (libsbpx.so, C ABI in
include/sbpx.h) and to the C++ driver (libsbp_host.so). The compute path is the HIP library;
this module only moves host arrays in and out. It raises if the libraries are not built."""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbpx.so")
HOST_LIB_PATH = os.path.join(HERE, "libsbp_host.so")

ALGO_MH, ALGO_ASYNC_GIBBS, ALGO_HYBRID = 0, 1, 2
MERGE_APPROXIMATE, MERGE_ADVANCED = 0, 1
LEARNED_PARAMS = 90656  # W1[256][64], b1[256], W2[256][256], b2[256], W3[32][256], b3[32]
OPT_COUNT_BYTES, OPT_MH_GROUP, OPT_EVAL_VARIANT, OPT_PROPOSAL, OPT_INCREMENTAL, OPT_MH_WIDE = 1, 2, 3, 4, 5, 6
OPT_MH_HEADROOM = 7
OPT_LEARN_MAXDEG = 8
OPT_BATCH_GRAPH = 9
OPT_MH_TEAM = 10
ALGORITHMS = {"metropolis_hastings": ALGO_MH, "async_gibbs": ALGO_ASYNC_GIBBS, "hybrid_mcmc": ALGO_HYBRID}

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

# Synthetic DC-SBM shapes (num_vertices, num_blocks, p_in, dirichlet, deg_min, deg_max, exponent).
CONFIGS = {
    "C1": dict(num_vertices=1000, num_blocks=8, p_in=0.8, dirichlet=20.0, deg_min=10, deg_max=100, exponent=2.1),
    "C2": dict(num_vertices=50000, num_blocks=44, p_in=0.6, dirichlet=1.0, deg_min=10, deg_max=100, exponent=2.1),
    "C4": dict(num_vertices=1000000, num_blocks=200, p_in=0.6, dirichlet=1.0, deg_min=10, deg_max=200, exponent=2.1),
    "C5": dict(num_vertices=5000000, num_blocks=400, p_in=0.55, dirichlet=1.0, deg_min=5, deg_max=500, exponent=2.1),
}


def c3_graph_params(count=512, seed=0):
    """Generator parameters of the batched workload: `count` DC-SBM graphs with N log-uniform in
    [1000, 20000], K = round(1.5 N^0.4), p_in uniform in [0.55, 0.85], Dirichlet concentration 1
    or 20, out-degrees on [5, 100] with exponent 2.1; graph g is generated with seed
    seed * 1000003 + g."""
    rng = np.random.default_rng(seed)
    out = []
    for g in range(count):
        n = int(round(np.exp(rng.uniform(np.log(1000.0), np.log(20000.0)))))
        out.append(dict(num_vertices=n, num_blocks=int(round(1.5 * n ** 0.4)), p_in=float(rng.uniform(0.55, 0.85)),
                        dirichlet=float(rng.choice([1.0, 20.0])), deg_min=5, deg_max=100, exponent=2.1,
                        seed=seed * 1000003 + g))
    return out


class SweepStats(C.Structure):
    _fields_ = [("attempts", C.c_int64), ("evals", C.c_int64), ("moves", C.c_int64), ("dS_accepted", C.c_double),
                ("eval_ms", C.c_double * 4), ("eval_bytes", C.c_double * 4), ("eval_launches", C.c_int32 * 4)]


class DcsbmParams(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("num_blocks", C.c_int32), ("p_in", C.c_double),
                ("dirichlet", C.c_double), ("deg_min", C.c_int32), ("deg_max", C.c_int32),
                ("exponent", C.c_double), ("seed", C.c_uint64)]


class RunStats(C.Structure):
    _fields_ = [("num_blocks", C.c_int32), ("outer_iterations", C.c_int32), ("sweeps", C.c_int64),
                ("attempts", C.c_int64), ("evals", C.c_int64), ("moves", C.c_int64), ("seconds", C.c_double),
                ("mcmc_seconds", C.c_double), ("eval_ms", C.c_double * 4), ("eval_bytes", C.c_double * 4),
                ("eval_launches", C.c_int64 * 4), ("mdl", C.c_double), ("stopped_at_limit", C.c_int32)]


_lib = None
_host = None


def load():
    """The HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with build.py (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.sbpx_create.argtypes = [C.POINTER(vp), C.c_int]
    L.sbpx_destroy.argtypes = [vp]
    L.sbpx_destroy.restype = None
    L.sbpx_last_error.argtypes = [vp]
    L.sbpx_last_error.restype = C.c_char_p
    L.sbpx_set_option.argtypes = [vp, C.c_int, C.c_int64]
    L.sbpx_graph_upload.argtypes = [vp, C.c_int32, C.c_int64, i32p, i32p, i32p, i32p]
    L.sbpx_graph_split.argtypes = [vp, vp, C.POINTER(C.c_int32), vp]
    L.sbpx_blockmodel_build.argtypes = [vp, C.c_int32, i32p]
    L.sbpx_blockmodel_info.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
    L.sbpx_blockmodel_download.argtypes = [vp, vp, vp, vp, vp]
    L.sbpx_blockmodel_coo.argtypes = [vp, C.c_int, i32p, i32p, i32p]
    L.sbpx_mdl.argtypes = [vp, C.POINTER(C.c_double)]
    L.sbpx_log_posterior.argtypes = [vp, C.POINTER(C.c_double)]
    L.sbpx_eval_moves.argtypes = [vp, C.c_int32, i32p, i32p, f64p, f64p, f64p, u8p]
    L.sbpx_propose.argtypes = [vp, C.c_int32, i32p, C.c_uint64, C.c_uint32, C.c_int, i32p, f64p]
    L.sbpx_apply_moves.argtypes = [vp, C.c_int32, i32p, i32p]
    L.sbpx_sweep.argtypes = [vp, C.c_int, C.c_int32, C.c_uint64, C.c_uint32, C.POINTER(SweepStats)]
    L.sbpx_merge_proposals.argtypes = [vp, C.c_uint64, C.c_uint32, f64p, i32p]
    L.sbpx_carry_out_merges.argtypes = [vp, C.c_int32, C.c_int, f64p, i32p]
    L.sbpx_merge_blocks.argtypes = [vp, C.c_int32, C.c_int, C.c_uint64, C.c_uint32]
    L.sbpx_selftest_logf.argtypes = [vp, C.c_uint32, C.c_uint32, np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")]
    L.sbpx_generate_dcsbm.argtypes = [C.POINTER(DcsbmParams), C.POINTER(C.c_int64), vp, vp, vp, vp, vp]
    f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.sbpx_learned_init.argtypes = [vp, C.c_uint64]
    L.sbpx_learned_params.argtypes = [vp, vp, vp]
    L.sbpx_learned_log.argtypes = [vp, C.c_int32]
    L.sbpx_learned_log_count.argtypes = [vp, C.POINTER(C.c_int32)]
    L.sbpx_learned_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, C.POINTER(C.c_float)]
    L.sbpx_learned_forward.argtypes = [vp, C.c_int32, f32p, f32p]
    L.sbpx_learned_eval_moves.argtypes = [vp, C.c_int32, i32p, i32p, f64p, f64p, f64p, u8p, f32p, f32p]
    L.sbpx_learned_scores.argtypes = [vp, C.c_int32, i32p, C.c_int, i32p, f32p]
    L.sbpx_log_posterior_noalloc.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp]
    L.sbpx_learned_gradients.argtypes = [vp, C.c_int32, f32p, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"),
                                         i32p, f32p, f32p]
    L.sbpx_batch_define.argtypes = [vp, C.c_int32, i32p]
    L.sbpx_batch_info.argtypes = [vp, C.POINTER(C.c_int32), vp]
    L.sbpx_batch_build.argtypes = [vp, C.c_int32, i32p, i32p, i32p]
    L.sbpx_batch_download.argtypes = [vp, C.c_int32, i32p, i32p]
    L.sbpx_batch_log_posterior.argtypes = [vp, f64p]
    L.sbpx_batch_sweep.argtypes = [vp, C.c_int, C.c_int32, C.c_int32, i32p,
                                   np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS"),
                                   np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS"), C.POINTER(SweepStats)]
    L.sbpx_batch_replays.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    L.sbpx_batch_merge.argtypes = [vp, C.c_int32, i32p, i32p, C.c_int,
                                   np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS"),
                                   np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]
    _lib = L
    return L


def load_host():
    """The C++ driver library (reference API on top of the C ABI)."""
    global _host
    if _host is not None:
        return _host
    load()
    if not os.path.exists(HOST_LIB_PATH):
        raise RuntimeError(f"{HOST_LIB_PATH} is missing: build it with build.py")
    H = C.CDLL(HOST_LIB_PATH)
    H.sbp_run.argtypes = [C.c_int32, i32p, i32p, i32p, i32p, C.c_int, C.c_int32, C.c_int, C.c_uint64, C.c_int,
                          C.c_int, i32p, C.POINTER(RunStats), C.c_char_p, C.c_int]
    H.sbp_graph_new.argtypes = [C.c_int32, i32p, i32p, i32p, i32p, C.c_int, C.c_int64, C.c_char_p, C.c_int]
    H.sbp_set_time_limit.argtypes = [C.c_double]
    H.sbp_file_load.argtypes = [C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    H.sbp_file_load.restype = C.c_void_p
    H.sbp_file_info.argtypes = [C.c_void_p] + [C.c_void_p] * 5
    H.sbp_file_info.restype = None
    H.sbp_file_arrays.argtypes = [C.c_void_p] + [C.c_void_p] * 5
    H.sbp_file_arrays.restype = None
    H.sbp_file_free.argtypes = [C.c_void_p]
    H.sbp_file_free.restype = None
    H.sbp_graph_new.restype = C.c_void_p
    H.sbp_graph_free.argtypes = [C.c_void_p]
    H.sbp_graph_free.restype = None
    H.sbp_last_trajectory.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    H.sbp_last_intermediates.argtypes = [C.c_void_p, C.c_int]
    H.sbp_partition_checkpointed.argtypes = [C.c_void_p, C.c_int, C.c_int32, C.c_int, C.c_uint64, C.c_int, C.c_char_p,
                                             C.c_int, C.c_int32, i32p, C.POINTER(RunStats), C.c_char_p, C.c_int]
    H.sbp_graph_option.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_char_p, C.c_int]
    H.sbp_graph_learned.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS"),
                                    C.c_char_p, C.c_int]
    H.sbp_partition.argtypes = [C.c_void_p, C.c_int, C.c_int32, C.c_int, C.c_uint64, C.c_int, i32p,
                                C.POINTER(RunStats), C.c_char_p, C.c_int]
    H.sbp_partition_batch.argtypes = [C.c_int32, i32p, i32p, i32p, i32p, i32p, C.c_void_p, C.c_int32,
                                      np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS"), C.c_int, C.c_int,
                                      C.c_int, C.c_int, i32p, i32p, f64p, i32p, f64p, C.c_char_p, C.c_int]
    _host = H
    return H


INTERMEDIATE_FIELDS = ("iteration", "mdl", "normalized_mdl_v1", "normalized_mdl_v2", "modularity", "interblock_edges",
                       "block_size_variation", "num_blocks", "mcmc_iterations", "mcmc_time", "seconds")

EXPORTED = ["sbpx_create", "sbpx_destroy", "sbpx_last_error", "sbpx_set_option", "sbpx_graph_upload",
            "sbpx_graph_split", "sbpx_blockmodel_build", "sbpx_blockmodel_info", "sbpx_blockmodel_download",
            "sbpx_blockmodel_coo", "sbpx_mdl", "sbpx_log_posterior", "sbpx_eval_moves", "sbpx_propose",
            "sbpx_apply_moves", "sbpx_sweep", "sbpx_merge_proposals", "sbpx_carry_out_merges",
            "sbpx_merge_blocks", "sbpx_selftest_logf", "sbpx_generate_dcsbm", "sbpx_learned_init",
            "sbpx_learned_params", "sbpx_learned_log", "sbpx_learned_log_count", "sbpx_learned_train",
            "sbpx_learned_forward", "sbpx_learned_gradients", "sbpx_batch_define", "sbpx_batch_info",
            "sbpx_batch_build", "sbpx_batch_download", "sbpx_batch_log_posterior", "sbpx_batch_sweep",
            "sbpx_batch_merge", "sbpx_blockmodel_trace", "sbpx_batch_trace", "sbpx_learned_eval_moves",
            "sbpx_learned_scores", "sbpx_log_posterior_noalloc", "sbpx_delta_mdl", "sbpx_merge_eval",
            "sbpx_blockmodel_cells", "sbpx_blockmodel_line", "sbpx_batch_replays"]


class SbpxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"sbpx status {status}: {msg}")
        self.status = status


def generate_dcsbm(num_vertices, num_blocks, p_in, dirichlet, deg_min, deg_max, exponent, seed=0):
    """Synthetic directed DC-SBM graph: (out_ptr, out_idx, in_ptr, in_idx, truth)."""
    L = load()
    p = DcsbmParams(num_vertices, num_blocks, p_in, dirichlet, deg_min, deg_max, exponent, seed)
    E = C.c_int64()
    truth = np.zeros(num_vertices, np.int32)
    st = L.sbpx_generate_dcsbm(C.byref(p), C.byref(E), None, None, None, None, truth.ctypes.data)
    if st != 0:
        raise SbpxError(st, "bad generator parameters")
    out_ptr = np.zeros(num_vertices + 1, np.int32)
    in_ptr = np.zeros(num_vertices + 1, np.int32)
    out_idx = np.zeros(E.value, np.int32)
    in_idx = np.zeros(E.value, np.int32)
    st = L.sbpx_generate_dcsbm(C.byref(p), C.byref(E), out_ptr.ctypes.data, out_idx.ctypes.data, in_ptr.ctypes.data,
                               in_idx.ctypes.data, truth.ctypes.data)
    if st != 0:
        raise SbpxError(st, "generator failed")
    return out_ptr, out_idx, in_ptr, in_idx, truth


class Context:
    """One GPU context holding a graph and its current Blockmodel."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        st = self.L.sbpx_create(C.byref(h), device)
        if st != 0:
            raise SbpxError(st, "sbpx_create failed (no GPU?)")
        self.h = h
        self.N = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.sbpx_destroy(self.h)
            self.h = None

    def __del__(self):
        # no device calls while the interpreter shuts down: a profiler or the HIP runtime may
        # already have finalised; close() contexts explicitly instead
        if not sys.is_finalizing():
            self.close()

    def _chk(self, st):
        if st != 0:
            raise SbpxError(st, self.L.sbpx_last_error(self.h).decode())

    def set_option(self, option, value):
        self._chk(self.L.sbpx_set_option(self.h, option, value))

    def upload_graph(self, out_ptr, out_idx, in_ptr, in_idx):
        out_ptr, out_idx, in_ptr, in_idx = (np.ascontiguousarray(x, np.int32) for x in (out_ptr, out_idx, in_ptr, in_idx))
        self.N = len(out_ptr) - 1
        self._chk(self.L.sbpx_graph_upload(self.h, self.N, len(out_idx), out_ptr, out_idx, in_ptr, in_idx))

    def split(self):
        nh = C.c_int32()
        self._chk(self.L.sbpx_graph_split(self.h, None, C.byref(nh), None))
        high = np.zeros(nh.value, np.int32)
        low = np.zeros(self.N - nh.value, np.int32)
        self._chk(self.L.sbpx_graph_split(self.h, high.ctypes.data, C.byref(nh), low.ctypes.data))
        return high, low

    def build(self, B, assignment):
        self._chk(self.L.sbpx_blockmodel_build(self.h, int(B), np.ascontiguousarray(assignment, np.int32)))

    def info(self):
        B = C.c_int32()
        nnz = C.c_int64()
        self._chk(self.L.sbpx_blockmodel_info(self.h, C.byref(B), C.byref(nnz)))
        return B.value, nnz.value

    def state(self):
        B, _ = self.info()
        a = np.zeros(self.N, np.int32)
        dout = np.zeros(B, np.int32)
        din = np.zeros(B, np.int32)
        deg = np.zeros(B, np.int32)
        self._chk(self.L.sbpx_blockmodel_download(self.h, a.ctypes.data, dout.ctypes.data, din.ctypes.data,
                                                  deg.ctypes.data))
        return a, dout, din, deg

    def coo(self, by_column=False):
        _, nnz = self.info()
        r = np.zeros(nnz, np.int32)
        c = np.zeros(nnz, np.int32)
        v = np.zeros(nnz, np.int32)
        self._chk(self.L.sbpx_blockmodel_coo(self.h, int(by_column), r, c, v))
        return r, c, v

    def mdl(self):
        out = C.c_double()
        self._chk(self.L.sbpx_mdl(self.h, C.byref(out)))
        return out.value

    def log_posterior(self):
        out = C.c_double()
        self._chk(self.L.sbpx_log_posterior(self.h, C.byref(out)))
        return out.value

    def eval_moves(self, v, s, u):
        v = np.ascontiguousarray(v, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        u = np.ascontiguousarray(u, np.float64)
        n = len(v)
        dS = np.zeros(n)
        H = np.zeros(n)
        acc = np.zeros(n, np.uint8)
        self._chk(self.L.sbpx_eval_moves(self.h, n, v, s, u, dS, H, acc))
        return dS, H, acc

    def propose(self, v, seed, call, tag=0):
        v = np.ascontiguousarray(v, np.int32)
        s = np.zeros(len(v), np.int32)
        u = np.zeros(len(v))
        self._chk(self.L.sbpx_propose(self.h, len(v), v, seed, call, tag, s, u))
        return s, u

    def apply_moves(self, v, s):
        self._chk(self.L.sbpx_apply_moves(self.h, len(v), np.ascontiguousarray(v, np.int32),
                                          np.ascontiguousarray(s, np.int32)))

    def sweep(self, algo, batches, seed, call):
        st = SweepStats()
        self._chk(self.L.sbpx_sweep(self.h, algo, batches, seed, call, C.byref(st)))
        return st

    def merge_proposals(self, seed, call):
        B, _ = self.info()
        dS = np.zeros(B)
        best = np.zeros(B, np.int32)
        self._chk(self.L.sbpx_merge_proposals(self.h, seed, call, dS, best))
        return dS, best

    def carry_out_merges(self, num_to_merge, dS, best, mode=MERGE_APPROXIMATE):
        self._chk(self.L.sbpx_carry_out_merges(self.h, int(num_to_merge), mode, np.ascontiguousarray(dS, np.float64),
                                               np.ascontiguousarray(best, np.int32)))

    def selftest_logf(self, lo, n):
        out = np.zeros(n, np.float32)
        self._chk(self.L.sbpx_selftest_logf(self.h, lo, n, out))
        return out

    def merge_blocks(self, num_to_merge, seed, call, mode=MERGE_APPROXIMATE):
        self._chk(self.L.sbpx_merge_blocks(self.h, int(num_to_merge), mode, seed, call))

    # a batch of graphs in this context (the uploaded graph is their union)
    def batch_define(self, vertex_offsets):
        offs = np.ascontiguousarray(vertex_offsets, np.int32)
        self._chk(self.L.sbpx_batch_define(self.h, len(offs) - 1, offs))
        self._batch_sizes = np.diff(offs)

    def batch_info(self):
        G = C.c_int32()
        self._chk(self.L.sbpx_batch_info(self.h, C.byref(G), None))
        nb = np.zeros(G.value, np.int32)
        self._chk(self.L.sbpx_batch_info(self.h, None, nb.ctypes.data))
        return nb

    def batch_build(self, graphs, num_blocks, assignment):
        g = np.ascontiguousarray(graphs, np.int32)
        self._chk(self.L.sbpx_batch_build(self.h, len(g), g, np.ascontiguousarray(num_blocks, np.int32),
                                          np.ascontiguousarray(assignment, np.int32)))

    def batch_download(self, graphs=None):
        """Graph-local assignments of the listed graphs (default all), one after the other."""
        g = np.arange(len(self.batch_info()), dtype=np.int32) if graphs is None else np.ascontiguousarray(graphs, np.int32)
        n = int(sum(self._batch_sizes[x] for x in g))
        a = np.zeros(n, np.int32)
        self._chk(self.L.sbpx_batch_download(self.h, len(g), g, a))
        return a

    def batch_log_posterior(self):
        out = np.zeros(len(self.batch_info()))
        self._chk(self.L.sbpx_batch_log_posterior(self.h, out))
        return out

    def batch_sweep(self, graphs, seeds, calls, batches=1, algo=ALGO_ASYNC_GIBBS):
        g = np.ascontiguousarray(graphs, np.int32)
        st = SweepStats()
        self._chk(self.L.sbpx_batch_sweep(self.h, algo, batches, len(g), g, np.ascontiguousarray(seeds, np.uint64),
                                          np.ascontiguousarray(calls, np.uint32), C.byref(st)))
        return st

    def batch_replays(self):
        """(rounds replayed from the captured graph, graph captures) since batch_define."""
        r, k = C.c_int64(), C.c_int64()
        self._chk(self.L.sbpx_batch_replays(self.h, C.byref(r), C.byref(k)))
        return r.value, k.value

    def batch_merge(self, graphs, num_to_merge, seeds, calls, mode=MERGE_APPROXIMATE):
        g = np.ascontiguousarray(graphs, np.int32)
        self._chk(self.L.sbpx_batch_merge(self.h, len(g), g, np.ascontiguousarray(num_to_merge, np.int32), mode,
                                          np.ascontiguousarray(seeds, np.uint64), np.ascontiguousarray(calls, np.uint32)))

    # proposal network
    def learned_init(self, seed=0):
        self._chk(self.L.sbpx_learned_init(self.h, seed))

    def learned_params(self, new=None):
        """Returns the flat parameters; with `new`, installs them afterwards."""
        out = np.zeros(LEARNED_PARAMS, np.float32)
        src = None if new is None else np.ascontiguousarray(new, np.float32)
        if src is not None and src.size != LEARNED_PARAMS:
            raise ValueError("expected %d parameters" % LEARNED_PARAMS)
        self._chk(self.L.sbpx_learned_params(self.h, out.ctypes.data, None if src is None else src.ctypes.data))
        return out

    def learned_log(self, capacity):
        self._chk(self.L.sbpx_learned_log(self.h, int(capacity)))

    def learned_log_count(self):
        n = C.c_int32()
        self._chk(self.L.sbpx_learned_log_count(self.h, C.byref(n)))
        return n.value

    def learned_train(self, steps, batch=4096, lr=1e-3, seed=0):
        loss = C.c_float()
        self._chk(self.L.sbpx_learned_train(self.h, int(steps), int(batch), float(lr), seed, C.byref(loss)))
        return loss.value

    def learned_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        z = np.zeros((x.shape[0], 32), np.float32)
        self._chk(self.L.sbpx_learned_forward(self.h, x.shape[0], x, z))
        return z

    def learned_eval_moves(self, v, s, u):
        """Fixed proposals evaluated as the learned mixture's: (dS, H, accept, q_net(s), q_net(r))."""
        v = np.ascontiguousarray(v, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        u = np.ascontiguousarray(u, np.float64)
        n = len(v)
        dS, H = np.zeros(n), np.zeros(n)
        acc = np.zeros(n, np.uint8)
        qf, qb = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._chk(self.L.sbpx_learned_eval_moves(self.h, n, v, s, u, dS, H, acc, qf, qb))
        return dS, H, acc, qf, qb

    def learned_scores(self, v, scorer=0):
        """Candidates [n][32] and scores [n][32] of the listed vertices: scorer 0 the batched path,
        1 the sequential pass's in-wavefront scorer."""
        v = np.ascontiguousarray(v, np.int32)
        cand = np.zeros((len(v), 32), np.int32)
        z = np.zeros((len(v), 32), np.float32)
        self._chk(self.L.sbpx_learned_scores(self.h, len(v), v, int(scorer), cand, z))
        return cand, z

    def learned_gradients(self, x, mask, y):
        x = np.ascontiguousarray(x, np.float32)
        loss = np.zeros(x.shape[0], np.float32)
        grads = np.zeros(LEARNED_PARAMS, np.float32)
        self._chk(self.L.sbpx_learned_gradients(self.h, x.shape[0], x, np.ascontiguousarray(mask, np.uint32),
                                                np.ascontiguousarray(y, np.int32), loss, grads))
        return loss, grads


def run_sbp(out_ptr, out_idx, in_ptr, in_idx, algorithm="async_gibbs", batches=1, approximate=True, seed=0,
            device=0, count_bytes=False):
    """sbp::stochastic_block_partition through the C++ driver; returns (assignment, RunStats)."""
    H = load_host()
    N = len(out_ptr) - 1
    a = np.zeros(N, np.int32)
    st = RunStats()
    err = C.create_string_buffer(512)
    algo = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
    rc = H.sbp_run(N, np.ascontiguousarray(out_ptr, np.int32), np.ascontiguousarray(out_idx, np.int32),
                   np.ascontiguousarray(in_ptr, np.int32), np.ascontiguousarray(in_idx, np.int32), algo, batches,
                   int(approximate), seed, device, int(count_bytes), a, C.byref(st), err, 512)
    if rc != 0:
        raise RuntimeError("sbp_run: " + err.value.decode())
    return a, st


def load_graph(base_path, undirected=False):
    """Reads base_path.tsv (else .mtx) and base_path_truePartition.tsv as the reference's
    Graph::load does (graph.cpp:24-62). Returns dict(out_ptr, out_idx, in_ptr, in_idx, truth,
    num_edges, undirected); num_edges is the count the reference's MDL uses."""
    H = load_host()
    err = C.create_string_buffer(512)
    h = H.sbp_file_load(str(base_path).encode(), int(undirected), err, 512)
    if not h:
        raise RuntimeError("sbp_file_load: " + err.value.decode())
    try:
        n, stored, E, und, tl = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        H.sbp_file_info(h, C.byref(n), C.byref(stored), C.byref(E), C.byref(und), C.byref(tl))
        op = np.zeros(n.value + 1, np.int32)
        ip = np.zeros(n.value + 1, np.int32)
        oi = np.zeros(stored.value, np.int32)
        ii = np.zeros(stored.value, np.int32)
        truth = np.zeros(tl.value, np.int32)
        H.sbp_file_arrays(h, op.ctypes.data, oi.ctypes.data, ip.ctypes.data, ii.ctypes.data, truth.ctypes.data)
    finally:
        H.sbp_file_free(h)
    return dict(out_ptr=op, out_idx=oi, in_ptr=ip, in_idx=ii, truth=truth, num_edges=E.value,
                undirected=bool(und.value))


def union_graphs(graphs):
    """Joins graphs given as (out_ptr, out_idx, in_ptr, in_idx) into one graph without edges between
    them: returns (vertex_offsets, out_ptr, out_idx, in_ptr, in_idx)."""
    offs, ops, ois, ips, iis = [0], [np.zeros(1, np.int64)], [], [np.zeros(1, np.int64)], []
    eo = 0
    for op, oi, ip, ii in graphs:
        n, v0 = len(op) - 1, offs[-1]
        ops.append(np.asarray(op[1:], np.int64) + eo)
        ips.append(np.asarray(ip[1:], np.int64) + eo)
        ois.append(np.asarray(oi, np.int64) + v0)
        iis.append(np.asarray(ii, np.int64) + v0)
        eo += int(op[-1])
        offs.append(v0 + n)
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs), np.int32)
    return np.array(offs, np.int32), cat(ops), cat(ois), cat(ips), cat(iis)


def partition_batch(graphs, batches=1, seeds=None, num_edges=None, device=0, eval_variant=0, count_bytes=False,
                    graph_replay=False):
    """sbp::partition_batch: each graph of the list (out_ptr, out_idx, in_ptr, in_idx) partitioned by
    async_gibbs with approximate merges, all on one GPU at once; graph g runs as it would alone with
    seed seeds[g] (eval_variant as OPT_EVAL_VARIANT; count_bytes counts the evaluation kernels'
    algorithmic bytes; graph_replay replays sweep rounds from a captured HIP graph once the kernel
    variants are timed). Returns (list of (assignment, num_blocks, mdl, mcmc_iterations), totals) with
    totals = dict(seconds, rounds, sweeps, evals, eval_ms[4], eval_bytes[4], eval_launches[4], replayed,
    captures)."""
    H = load_host()
    G = len(graphs)
    offs, op, oi, ip, ii = union_graphs(graphs)
    seeds = np.ascontiguousarray(np.zeros(G) if seeds is None else seeds, np.uint64)
    ne = None if num_edges is None else np.ascontiguousarray(num_edges, np.int64)
    a = np.zeros(int(offs[-1]), np.int32)
    B = np.zeros(G, np.int32)
    mdl = np.zeros(G)
    its = np.zeros(G, np.int32)
    tot = np.zeros(18)
    err = C.create_string_buffer(512)
    rc = H.sbp_partition_batch(G, offs, op, oi, ip, ii, None if ne is None else ne.ctypes.data, batches, seeds, device,
                               int(eval_variant), int(count_bytes), int(graph_replay), a, B, mdl, its, tot, err, 512)
    if rc != 0:
        raise RuntimeError("sbp_partition_batch: " + err.value.decode())
    out = [(a[offs[g]:offs[g + 1]].copy(), int(B[g]), float(mdl[g]), int(its[g])) for g in range(G)]
    return out, dict(seconds=tot[0], rounds=int(tot[1]), sweeps=int(tot[2]), evals=int(tot[3]),
                     eval_ms=tot[4:8].copy(), eval_bytes=tot[8:12].copy(), eval_launches=tot[12:16].copy(),
                     replayed=int(tot[16]), captures=int(tot[17]))


class ResidentGraph:
    """A graph held on one GPU by the C++ driver, for repeated partitioning runs."""

    def __init__(self, out_ptr, out_idx, in_ptr, in_idx, device=0, num_edges=-1):
        self.H = load_host()
        self.N = len(out_ptr) - 1
        err = C.create_string_buffer(512)
        self.h = self.H.sbp_graph_new(self.N, np.ascontiguousarray(out_ptr, np.int32),
                                      np.ascontiguousarray(out_idx, np.int32), np.ascontiguousarray(in_ptr, np.int32),
                                      np.ascontiguousarray(in_idx, np.int32), device, int(num_edges), err, 512)
        if not self.h:
            raise RuntimeError("sbp_graph_new: " + err.value.decode())

    def set_option(self, option, value):
        err = C.create_string_buffer(512)
        if self.H.sbp_graph_option(self.h, option, int(value), err, 512) != 0:
            raise RuntimeError("sbp_graph_option: " + err.value.decode())

    def load_learned(self, params):
        """Installs proposal-network parameters (flat, LEARNED_PARAMS floats) for learned runs."""
        err = C.create_string_buffer(512)
        p = np.ascontiguousarray(params, np.float32)
        if p.size != LEARNED_PARAMS:
            raise ValueError("expected %d parameters" % LEARNED_PARAMS)
        if self.H.sbp_graph_learned(self.h, p, err, 512) != 0:
            raise RuntimeError("sbp_graph_learned: " + err.value.decode())

    def partition(self, algorithm="async_gibbs", batches=1, approximate=True, seed=0, count_bytes=False,
                  learned=False, checkpoint=None, resume=False, max_iterations=0, max_seconds=0.0):
        """sbp::stochastic_block_partition; with `checkpoint`, the driver state is written after every
        outer iteration, `resume` continues from it and `max_iterations` > 0 stops early;
        `max_seconds` > 0 stops the run after the sweep in progress once it has run that long
        (stats.stopped_at_limit is then 1)."""
        self.H.sbp_set_time_limit(float(max_seconds))
        # True / 1: batched sweeps; 2: also the sequential pass (one wavefront per vertex); 3: also
        # its wavefront groups
        self.set_option(OPT_PROPOSAL, int(learned))
        a = np.zeros(self.N, np.int32)
        st = RunStats()
        err = C.create_string_buffer(512)
        algo = ALGORITHMS[algorithm] if isinstance(algorithm, str) else int(algorithm)
        ck = None if checkpoint is None else str(checkpoint).encode()
        rc = self.H.sbp_partition_checkpointed(self.h, algo, batches, int(approximate), seed, int(count_bytes), ck,
                                               int(resume), int(max_iterations), a, C.byref(st), err, 512)
        if rc != 0:
            raise RuntimeError("sbp_partition: " + err.value.decode())
        return a, st

    def trajectory(self):
        """Outer iterations of the last run: (seconds since its start, MDL, number of blocks)."""
        n = self.H.sbp_last_trajectory(None, None, None, 0)
        t, mdl, B = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
        self.H.sbp_last_trajectory(t.ctypes.data, mdl.ctypes.data, B.ctypes.data, n)
        return t, mdl, B

    def intermediates(self):
        """sbp::get_intermediates of the last run: one dict per outer iteration (iteration, mdl,
        normalized_mdl_v1, normalized_mdl_v2, modularity, interblock_edges, block_size_variation,
        num_blocks, mcmc_iterations, mcmc_time, seconds)."""
        n = self.H.sbp_last_intermediates(None, 0)
        rows = np.zeros((n, 11))
        self.H.sbp_last_intermediates(rows.ctypes.data, n)
        return [dict(zip(INTERMEDIATE_FIELDS, r)) for r in rows]

    def close(self):
        if getattr(self, "h", None):
            self.H.sbp_graph_free(self.h)
            self.h = None

    def __del__(self):
        if not sys.is_finalizing():
            self.close()
