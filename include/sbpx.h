/*
 * sbpx — C ABI of the MI355X stochastic-block-partitioning hot path.
 *
 * One context owns one GPU's copy of a graph (int32 CSR + CSC) and of a Blockmodel (block
 * assignment, block matrix as CSR + CSC, block degrees). Host buffers are copied in and out;
 * the library never keeps caller pointers. Calls on one context are serialised on the
 * context's stream; distinct contexts are independent. Every entry point returns SBPX_OK or an
 * error code, with a message in sbpx_last_error().
 *
 * The reference exposes this path as C++ free functions and classes, not as a C interface.
 * Each entry point below names the reference function whose work it replaces (file:line in
 * vtsynergy/Hybrid-Stochastic-Block-Partitioning). The C++ layer in
 * hybrid-stochastic-block-partitioning_amd/host restores the reference's signatures on top.
 *
 * Randomness: proposals and accept uniforms come from Philox4x32-10 keyed by `seed`, with the
 * counter (vertex or block, `call`, proposal index, tag). A call id names one sweep or one
 * merge phase; the caller advances it.
 */
#ifndef SBPX_H
#define SBPX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbpx_ctx sbpx_ctx;
typedef int sbpx_status;

#define SBPX_OK 0
#define SBPX_ERR_ARG 1         /* bad argument: block id >= B, edge endpoint >= N, ... */
#define SBPX_ERR_HIP 2         /* HIP runtime failure or out of device memory */
#define SBPX_ERR_STATE 3       /* call out of order (no graph, no blockmodel) */
#define SBPX_ERR_NUMERIC 4     /* zero block degree under a nonzero cell, NaN delta MDL */
#define SBPX_ERR_UNSUPPORTED 5 /* configuration outside what this build handles */

#define SBPX_ALGO_METROPOLIS_HASTINGS 0
#define SBPX_ALGO_ASYNC_GIBBS 1
#define SBPX_ALGO_HYBRID 2

#define SBPX_MERGE_APPROXIMATE 0 /* Blockmodel::carry_out_best_merges */
#define SBPX_MERGE_ADVANCED 1    /* block_merge::carry_out_best_merges_advanced */

typedef struct sbpx_sweep_stats {
    int64_t attempts;     /* proposals drawn */
    int64_t evals;        /* proposals with s != r: delta MDL, Hastings and accept computed */
    int64_t moves;        /* accepted moves */
    double dS_accepted;   /* sum of delta MDL over accepted moves */
    /* Per vertex-move evaluation kernel: 0-2 the batched kernels by degree bin (0: degree <= 127,
     * 1: <= 2047, 2: larger), 3 the sequential Metropolis-Hastings kernel. Device time,
     * algorithmic bytes of the evaluations that count (when byte counting is on), launches. */
    double eval_ms[4];
    double eval_bytes[4];
    int32_t eval_launches[4];
} sbpx_sweep_stats;

/* Context. Replaces the process-global state of the reference (args, static RNG engines). */
sbpx_status sbpx_create(sbpx_ctx** ctx, int device);
void sbpx_destroy(sbpx_ctx* ctx);
const char* sbpx_last_error(const sbpx_ctx* ctx);
/* option 1 = count algorithmic bytes in the evaluation kernels (0/1).
 * option 2 = wavefronts that share one vertex's evaluation in the sequential Metropolis-Hastings
 *            kernel: 1, 2, 4, 8 or 16, or 0 to choose by the longest block-matrix row (default).
 * option 3 = kernel variants of the batched evaluation per degree bin: 1 the first (default
 *            unroll, one wavefront per vertex), 2 the second (four row slots per lane per trip
 *            for degrees <= 126; 4 wavefronts sharing a vertex for degrees 127-510), 3 the third
 *            where a bin has one (degrees <= 30: 16 lanes per vertex, four vertices per
 *            wavefront; other bins take the first), or 0 to time them and keep the fastest per
 *            bin after every change of the number of blocks (default). Variants that split a
 *            vertex's sums over other lanes add them in another order, so delta MDL may differ
 *            in its last bits between them.
 * option 4 = proposal: 0 the reference's, 1 the learned mixture in the batched evaluation
 *            (see sbpx_learned_init), 2 also in the sequential Metropolis-Hastings pass, where
 *            each window's entries are scored on the current state (needs option 6 != 1).
 * option 5 = block matrix after a batch of moves: 1 (default) updated from the moved vertices'
 *            edges when they are at most a quarter of the edges, else rebuilt by a full sort;
 *            0 always rebuilt. Both give the same arrays.
 * option 6 = workgroups of the sequential Metropolis-Hastings pass when one wavefront evaluates
 *            each vertex: 1 (one workgroup, 16 speculative evaluations per window), a power of two
 *            up to 512 (8 per workgroup in one cooperative launch; capped at the workgroups the
 *            device holds at once), or 0 to follow the window yield (default). Every choice
 *            applies the same moves in the same order.
 * option 7 = spare slots of the Metropolis-Hastings / heap-merge table pool when it is built:
 *            0 (default) as many as the initial layout holds (at least 2^20); a positive value
 *            limits the pool to exactly that many (and restarts the doubling that follows each
 *            time the pool runs out), so that the rebuild-and-resume path runs; for tests.
 * option 8 = degree bound (out + in) of the learned mixture: 0 to 510, default 126. The
 *            sequential pass (proposal 2) covers every vertex within it; the batched
 *            sweeps cover degree <= 126 and, when the bound reaches 254 or 510, the vertices of
 *            degree 127-254 or 127-510 too. Other vertices keep the plain proposal.
 * option 9 = batched sweeps (sbpx_batch_sweep) replayed from a captured HIP graph: 1 once the
 *            kernel-variant timing of option 3 has settled, 0 never (default). A round's
 *            evaluation, application and block-matrix update are captured once per plan shape
 *            (number of batches, key width, kernel variants) with every size in device memory,
 *            and replayed for any subset of the graphs; the evaluation launches are persistent
 *            workgroups that take vertices from per-bin work counters. The states are the same
 *            either way. Rounds replayed from the graph report no per-kernel times.
 * option 10 = workgroups sharing one entry of the sequential Metropolis-Hastings pass's windows
 *            (a team of 8-wavefront workgroups, each scanning a slice of the rows): 0 to follow
 *            the row slots of the largest evaluation per window (teams from 8192 slots, about
 *            4096 per member, up to 32, as long as the entries the windows advance still fit;
 *            default), 1 never, or a power of two up to 32 (then 8 wavefronts per member).
 *            Every choice applies the same moves in the same order. */
sbpx_status sbpx_set_option(sbpx_ctx* ctx, int option, int64_t value);

/* Graph: deduplicated directed adjacency as int32 CSR (out) and CSC (in), neighbour order as
 * given. Replaces Graph (graph.hpp:15-113) and its degree split (graph.cpp:15-22, 215-228). */
sbpx_status sbpx_graph_upload(sbpx_ctx* ctx, int32_t num_vertices, int64_t num_edges,
                              const int32_t* out_ptr, const int32_t* out_idx,
                              const int32_t* in_ptr, const int32_t* in_idx);
/* High-degree (top 15%) and low-degree vertex lists, as Graph::sort_vertices orders them. */
sbpx_status sbpx_graph_split(sbpx_ctx* ctx, int32_t* high, int32_t* num_high, int32_t* low);

/* Blockmodel(num_blocks, graph, rate, assignment) + initialize_edge_counts
 * (blockmodel.hpp:63-73, blockmodel.cpp:187-255). Resets the total-degree drift. */
sbpx_status sbpx_blockmodel_build(sbpx_ctx* ctx, int32_t num_blocks, const int32_t* assignment);
sbpx_status sbpx_blockmodel_info(sbpx_ctx* ctx, int32_t* num_blocks, int64_t* nnz);
/* Any pointer may be NULL. deg includes the drift of Blockmodel::move_vertex
 * (blockmodel.cpp:344-360). */
sbpx_status sbpx_blockmodel_download(sbpx_ctx* ctx, int32_t* assignment, int32_t* dout,
                                     int32_t* din, int32_t* deg);
/* Nonzero cells, row-major sorted (from the CSR) or column-major sorted (from the CSC). */
sbpx_status sbpx_blockmodel_coo(sbpx_ctx* ctx, int by_column, int32_t* rows, int32_t* cols,
                                int32_t* vals);

/* Trace of the block matrix (sum of M[b, b]: edges inside a block) and the number of stored
 * edges (sum of the in-degrees), the two terms of Blockmodel::interblock_edges
 * (blockmodel.cpp:257-261). */
sbpx_status sbpx_blockmodel_trace(sbpx_ctx* ctx, int64_t* trace, int64_t* edges);

/* entropy::mdl (entropy.cpp:584-589) and Blockmodel::log_posterior_probability
 * (blockmodel.cpp:267-285). */
sbpx_status sbpx_mdl(sbpx_ctx* ctx, double* mdl);
sbpx_status sbpx_log_posterior(sbpx_ctx* ctx, double* log_posterior);

/* entropy::delta_mdl(blockmodel, delta, proposal) (entropy.cpp:296-401) of a given Delta of a
 * vertex move r -> s: its n stored cells (rows[i], cols[i]) with changes deltas[i] (a cell stored
 * with change 0 counts as the reference's Delta counts it), the proposal's k_out and k_in and the
 * Delta's self-edge weight. Each cell must lie in row or column r or s and appear once. */
sbpx_status sbpx_delta_mdl(sbpx_ctx* ctx, int32_t r, int32_t s, int32_t n, const int32_t* rows, const int32_t* cols,
                           const int32_t* deltas, int32_t k_out, int32_t k_in, int32_t self_weight, double* dS);
/* entropy::block_merge_delta_mdl (entropy.cpp:127-180) of merging block r[i] into s[i], each on the
 * current state (block_merge::blockmodel_delta, block_merge.cpp:13-37, gives the Delta). */
sbpx_status sbpx_merge_eval(sbpx_ctx* ctx, int32_t n, const int32_t* r, const int32_t* s, double* dS);
/* Block-matrix cells M[rows[i], cols[i]] (0 where absent): blockmatrix()->get(row, col)
 * (dict_transpose_matrix.cpp:74-84). */
sbpx_status sbpx_blockmodel_cells(sbpx_ctx* ctx, int32_t n, const int32_t* rows, const int32_t* cols, int32_t* vals);
/* The nonzeros of row x (by_column = 0) or column x (1) in ascending index order: getrow_sparse /
 * getcol_sparse (dict_transpose_matrix.cpp:86-122). count receives their number; at most cap are
 * written. */
sbpx_status sbpx_blockmodel_line(sbpx_ctx* ctx, int32_t x, int by_column, int32_t cap, int32_t* idx, int32_t* vals,
                                 int32_t* count);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid allocation). The log posterior of
 * sbpx_log_posterior, of a block matrix the caller holds in device memory: nnz nonzero cells in
 * row-major order (coo_row[i], col[i], val[i]), block degrees dout/din; the result, one double, is
 * written to the device address `out` on `stream` (a hipStream_t; NULL: the context's stream) and
 * the call returns without waiting. All pointers are device addresses the caller owns. No device
 * memory is allocated: partial sums live in a fixed array of the library on the context's device,
 * sized for nnz <= SBPX_NOALLOC_MAX_CELLS (larger: SBPX_ERR_ARG), held until the process ends.
 * Calls that overlap on one device race on that array. A nonzero cell whose row's dout or column's
 * din is zero makes the result infinite (no status: the call does not read back). */
#define SBPX_NOALLOC_MAX_CELLS (1LL << 28)
sbpx_status sbpx_log_posterior_noalloc(sbpx_ctx* ctx, int64_t nnz, const int32_t* coo_row, const int32_t* col,
                                       const int32_t* val, const int32_t* dout, const int32_t* din, double* out,
                                       void* stream);

/* Evaluate fixed proposals (v[i] -> s[i], uniform u[i]) against the current state without
 * applying them: entropy::delta_mdl (entropy.cpp:296-401), entropy::hastings_correction
 * (entropy.cpp:495-557), finetune::accept (finetune.cpp:19-29). */
sbpx_status sbpx_eval_moves(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s,
                            const double* u, double* dS, double* H, uint8_t* accept);
/* The proposal each vertex draws on the current state (common::propose_new_block,
 * common.cpp:131-192); tag 0 = asynchronous Gibbs stream, 1 = Metropolis-Hastings stream. */
sbpx_status sbpx_propose(sbpx_ctx* ctx, int32_t n, const int32_t* v, uint64_t seed, uint32_t call,
                         int tag, int32_t* s, double* u);
/* Apply moves as one asynchronous batch does (finetune.cpp:58-71): all moves take effect, the
 * block matrix and dout/din equal a rebuild, and deg carries the drift of applying them one by
 * one in ascending vertex id. Each vertex may appear at most once. */
sbpx_status sbpx_apply_moves(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s);
/* One MCMC sweep (one iteration of the loops in finetune.cpp:41-84, 356-405, 420-439). */
sbpx_status sbpx_sweep(sbpx_ctx* ctx, int algo, int32_t batches, uint64_t seed, uint32_t call,
                       sbpx_sweep_stats* stats);

/* block_merge::merge_blocks proposal phase (block_merge.cpp:278-301, 335-368): best merge per
 * block and its delta MDL (DBL_MAX marks a block whose proposals all repeated). */
sbpx_status sbpx_merge_proposals(sbpx_ctx* ctx, uint64_t seed, uint32_t call, double* dS,
                                 int32_t* best);
/* Carry out num_to_merge merges from (dS, best) and rebuild (blockmodel.cpp:50-78,
 * block_merge.cpp:123-217, :309). */
sbpx_status sbpx_carry_out_merges(sbpx_ctx* ctx, int32_t num_to_merge, int mode, const double* dS,
                                  const int32_t* best);
/* Both steps: block_merge::merge_blocks (block_merge.cpp:278-311). */
sbpx_status sbpx_merge_blocks(sbpx_ctx* ctx, int32_t num_to_merge, int mode, uint64_t seed,
                              uint32_t call);

/* ---------------------------------------------------------------------------------------
 * Learned proposal. No counterpart in the reference: a proposal network mixed with the
 * reference's proposal and corrected in the Hastings ratio, so the chain's target and MDL stay
 * the reference's. Network: 64 features -> 256 -> 256 -> 32 scores (ReLU), BF16 weights on the
 * matrix cores, FP32 accumulation; it scores the 32 blocks a vertex's neighbours fall in most
 * often. The features leave out the vertex's own contribution to every block degree, so they are
 * the same before and after the vertex moves, and one set of scores gives the network's forward
 * and reverse probabilities. Selected per context with option 4 (0 plain proposal, 1 mixture
 * 0.2 * plain + 0.8 * network in the batched sweeps, 2 also in the sequential pass) for vertices
 * within the degree bound of option 8.
 * Parameters, flat, in this order: W1[256][64], b1[256], W2[256][256], b2[256], W3[32][256],
 * b3[32] (SBPX_LEARNED_PARAMS floats; weights output-major). */
#define SBPX_LEARNED_PARAMS 90656
/* (Re)initialise the network: He-normal weights from the seed, zero biases, fresh Adam state. */
sbpx_status sbpx_learned_init(sbpx_ctx* ctx, uint64_t seed);
/* Copy the parameters out (out != NULL), then in (in != NULL). */
sbpx_status sbpx_learned_params(sbpx_ctx* ctx, float* out, const float* in);
/* Log training samples from the following plain sweeps: for every accepted move of a vertex of
 * degree <= 126 whose target is one of its candidates, its features and the candidate index.
 * capacity > 0 starts a fresh log of up to that many rows; 0 stops logging (rows are kept). */
sbpx_status sbpx_learned_log(sbpx_ctx* ctx, int32_t capacity);
sbpx_status sbpx_learned_log_count(sbpx_ctx* ctx, int32_t* count);
/* Adam steps on the logged rows: `steps` minibatches of `batch` (<= 8192) rows drawn uniformly
 * with replacement from the seed; softmax cross-entropy over each row's candidates. Reports the
 * mean loss of the last minibatch. */
sbpx_status sbpx_learned_train(sbpx_ctx* ctx, int32_t steps, int32_t batch, float lr, uint64_t seed,
                               float* mean_loss);
/* Evaluate fixed proposals (v[i] -> s[i], uniform u[i], each vertex at most once) as the learned
 * mixture proposal would have drawn them, on the current state, without applying them: delta MDL
 * (entropy.cpp:296-401), the mixture's Hastings ratio H = [w q_plain(r) + (1 - w) q_net(r)] /
 * [w q_plain(s) + (1 - w) q_net(s)] with q_plain the reference's closed form
 * (entropy.cpp:495-557) divided by the vertex degree, and the accept decision
 * (finetune.cpp:19-29). qf / qb (may be NULL) receive q_net(s) and q_net(r), -1 where the vertex
 * has no learned part (degree beyond option 8's bound, or no candidate outside its block); there H
 * is the plain ratio. The network must be initialised (sbpx_learned_init). */
sbpx_status sbpx_learned_eval_moves(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s, const double* u,
                                    double* dS, double* H, uint8_t* accept, float* qf, float* qb);
/* The candidates (cand[n][32], -1 past the last) and scores (z[n][32]) of the listed vertices
 * (degree <= 510) on the current state: scorer 0 = the batched sweeps' path (features kernel and
 * matrix-core products), 1 = the sequential pass's scorer (each vertex's features formed by one
 * wavefront, a workgroup's vertices scored together with the same matrix-core products). Both
 * give the same bits. */
sbpx_status sbpx_learned_scores(sbpx_ctx* ctx, int32_t n, const int32_t* v, int scorer, int32_t* cand, float* z);
/* The network on n (<= 8192) rows of 64 features (rounded to BF16): z[n][32] scores. */
sbpx_status sbpx_learned_forward(sbpx_ctx* ctx, int32_t n, const float* x, float* z);
/* Mean softmax cross-entropy of n rows (candidate bit masks, labels in their masks): per-row
 * losses and the gradient of every parameter, in the flat order above. */
sbpx_status sbpx_learned_gradients(sbpx_ctx* ctx, int32_t n, const float* x, const uint32_t* mask,
                                   const int32_t* y, float* loss, float* grads);

/* ---- A batch of independent graphs in one context ------------------------------------
 * The reference partitions one graph per run; these calls hold many graphs as one disjoint
 * union on the device and run each step of the driver for any subset of them in one set of
 * launches. Every graph follows exactly the run it would have alone with the same seed and
 * calls: randomness is keyed by graph-local ids, block ids keep the graph's own table slots,
 * and the log posterior is summed per graph. Graph g is the vertex range
 * [vertex_offsets[g], vertex_offsets[g + 1]) of the uploaded graph; block ids in these calls
 * are graph-local. While a batch is defined, the one-graph blockmodel, sweep and merge calls
 * return SBPX_ERR_STATE. */

/* Defines the batch over the uploaded graph (no edge may join two graphs) and builds every
 * graph's blockmodel with one block per vertex (Blockmodel(N, graph, rate), blockmodel.hpp:63-66). */
sbpx_status sbpx_batch_define(sbpx_ctx* ctx, int32_t num_graphs, const int32_t* vertex_offsets);

/* Number of graphs and the blocks in use per graph (num_blocks[num_graphs]); either may be NULL. */
sbpx_status sbpx_batch_info(sbpx_ctx* ctx, int32_t* num_graphs, int32_t* num_blocks);

/* Rebuilds the blockmodels of the n listed graphs from their assignments (Blockmodel(num_blocks,
 * graph, rate, assignment), blockmodel.hpp:67-73); `assignment` holds the graphs' vertex
 * assignments one after the other, in the order of `graphs`. */
sbpx_status sbpx_batch_build(sbpx_ctx* ctx, int32_t n, const int32_t* graphs, const int32_t* num_blocks,
                             const int32_t* assignment);

/* The graph-local vertex assignments of the n listed graphs, one after the other. */
sbpx_status sbpx_batch_download(sbpx_ctx* ctx, int32_t n, const int32_t* graphs, int32_t* assignment);

/* Trace of every graph's block matrix (trace[num_graphs]), as sbpx_blockmodel_trace. */
sbpx_status sbpx_batch_trace(sbpx_ctx* ctx, int64_t* trace);

/* Log posterior of every graph (out[num_graphs]; Blockmodel::log_posterior_p, blockmodel.cpp:267-285). */
sbpx_status sbpx_batch_log_posterior(sbpx_ctx* ctx, double* out);

/* One asynchronous Gibbs sweep (finetune.cpp:31-89, one iteration) of each listed graph, with
 * its own seed and call; stats are totals over the graphs. */
sbpx_status sbpx_batch_sweep(sbpx_ctx* ctx, int algo, int32_t batches, int32_t n, const int32_t* graphs,
                             const uint64_t* seeds, const uint32_t* calls, sbpx_sweep_stats* stats);

/* Rounds of sbpx_batch_sweep replayed from the captured graph (option 9) since the batch was
 * defined, and the graph captures made for them. No counterpart in the reference (which runs one
 * graph per process); for callers that check how a batch ran. */
sbpx_status sbpx_batch_replays(sbpx_ctx* ctx, int64_t* rounds, int64_t* captures);

/* One merge phase (block_merge::merge_blocks, block_merge.cpp:278-311) of each listed graph,
 * merging num_to_merge[i] blocks of graphs[i]; SBPX_MERGE_APPROXIMATE only. */
sbpx_status sbpx_batch_merge(sbpx_ctx* ctx, int32_t n, const int32_t* graphs, const int32_t* num_to_merge, int mode,
                             const uint64_t* seeds, const uint32_t* calls);

/* Self-test of the device's FP32 log (the function the objective uses, utils.hpp:145-151):
 * out[i] = log of the float whose bit pattern is lo + i, for i < n. */
sbpx_status sbpx_selftest_logf(sbpx_ctx* ctx, uint32_t lo, uint32_t n, float* out);

/* Synthetic degree-corrected SBM graph (host only; no GPU needed; no counterpart in the
 * reference, whose graphs come from graph-tool files). Returns CSR/CSC sized by the caller after
 * a first call with NULL outputs reports N and E. */
typedef struct sbpx_dcsbm_params {
    int32_t num_vertices;
    int32_t num_blocks;
    double p_in;          /* probability that an edge stays inside its source block */
    double dirichlet;     /* concentration of the block-size Dirichlet */
    int32_t deg_min, deg_max;
    double exponent;      /* truncated power law P(k) ~ k^-exponent on [deg_min, deg_max] */
    uint64_t seed;
} sbpx_dcsbm_params;
sbpx_status sbpx_generate_dcsbm(const sbpx_dcsbm_params* p, int64_t* num_edges, int32_t* out_ptr,
                                int32_t* out_idx, int32_t* in_ptr, int32_t* in_idx,
                                int32_t* truth);

#ifdef __cplusplus
}
#endif

#endif /* SBPX_H */
