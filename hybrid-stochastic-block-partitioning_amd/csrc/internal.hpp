// Internal state of an sbpx context: the device-resident graph and blockmodel, scratch
// buffers, and the launchers each translation unit provides.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "mlp.hpp"
#include "../../include/sbpx.h"

namespace sbpx {

#define SBPX_HIP(call)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) throw HipError(std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct ArgError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct NumericError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// An int kernel argument that the host knows (p null: the value v) or that a kernel wrote to
// device memory earlier on the stream (*p; v is then an upper bound the host sizes launches by).
// Launches over device-held counts can be captured in a graph once and replayed with whatever
// sizes the device computes in each replay.
struct Dyn {
    int v = 0;
    const int* p = nullptr;
    Dyn() = default;
    Dyn(int value) : v(value) {}
    Dyn(int bound, const int* at) : v(bound), p(at) {}
    __device__ __forceinline__ int get() const { return p ? *p : v; }
};

// A vertex list in device memory: base + *off (off null: base) and its length.
struct VList {
    const int* base = nullptr;
    const int* off = nullptr;
    Dyn n;
    VList() = default;
    VList(const int* list, int count) : base(list), n(count) {}
    VList(const int* b, const int* o, Dyn count) : base(b), off(o), n(count) {}
    __device__ __forceinline__ const int* ptr() const { return off ? base + *off : base; }
};

// Grid-stride loop over [0, count): one trip per thread when the launch covers the count, more
// when a fixed grid covers a device-held count.
#define GRID_LOOP(i, count)                                                                       \
    for (int i = (int)(blockIdx.x * blockDim.x + threadIdx.x), i##_end = (count); i < i##_end; \
         i += (int)(gridDim.x * blockDim.x))

// Workgroups of a launch over n items of `per` each: all of them for a host-known count, at most
// DYN_GRID (the kernels loop) for a device-held one.
inline int grid_for(const Dyn& n, int per);

// Device words of a batched sweep round replayed from a graph (Ctx::words): the sizes and
// decisions of the block-matrix update, which kernels write and later kernels read.
enum DynWord : int {
    W_NNZ = 0,     // cells of the block matrix (kept equal to bm.nnz)
    W_TOTAL,       // change entries of the batch's movers
    W_INC_TOTAL,   // the same when the incremental update was chosen, else 0
    W_EMIT_N,      // list length when incremental, else 0
    W_FULL_E,      // edge count when the full rebuild was chosen, else 0
    W_NOLD,        // previous cells when incremental, else 0
    W_NRUNS, W_ND, W_NNEW, W_NM, W_NNZ_INC, W_NNZ_FULL,
    NUM_WORDS = 16
};

// Device error word bits raised by kernels.
enum : int {
    DEV_ERR_ZERO_DEGREE = 1,   // nonzero cell with a zero block degree (blockmodel.cpp:276-281)
    DEV_ERR_NAN = 2,           // NaN/Inf delta MDL (entropy.cpp:323-340)
    DEV_ERR_TABLE = 4,         // neighbour-block table overflow
    DEV_ERR_BLOCK_RANGE = 8,   // block id out of range
    DEV_ERR_NO_PROPOSAL = 16,  // a merge candidate without a proposed block
    DEV_ERR_BARRIER = 32,      // a grid barrier of a multi-workgroup kernel was not reached in time
};

// Degree bins for the vertex-move kernels: neighbour-block table capacity per bin.
// Degree bins of the batched evaluation: bin b < NUM_LDS_BINS takes degrees up to
// BIN_MAXDEG[b], one wavefront per vertex with a BIN_CAP[b]-slot neighbour-block table in LDS
// (smaller tables leave room for more wavefronts); bin NUM_LDS_BINS (degree <= GROUP_MAXDEG)
// gives each vertex 4 wavefronts and a 4096-slot LDS table; the last bin 16 wavefronts and a
// table in global memory.
constexpr int NUM_LDS_BINS = 5;
constexpr int BIN_MAXDEG[NUM_LDS_BINS] = {30, 62, 126, 254, 510};
constexpr int BIN_CAP[NUM_LDS_BINS] = {64, 128, 256, 512, 1024};
constexpr int GROUP_MAXDEG = 2047;
// The learned proposal covers the vertices up to a degree bound chosen per context (126 by
// default, at most LEARN_MH_MAXDEG): in the sequential pass every vertex within it, in the
// batched sweeps the first three bins (degree <= 126) and the bins 3-4 whose largest degree is
// within it. Training rows come from the first three bins. The mixture keeps the plain proposal
// with this weight.
constexpr int LEARN_MAXDEG = 126;
constexpr int LEARN_MH_MAXDEG = 510;
constexpr float LEARN_W_PLAIN = 0.2f;
constexpr int NUM_BINS = NUM_LDS_BINS + 2;
// timing / byte classes reported per sweep: 0 = LDS bins with tables up to 256 slots,
// 1 = LDS bins with 512-1024 slots, 2 = group bins (degree > 510), 3 = sequential MH kernel
inline int bin_class(int b) { return b < 3 ? 0 : (b < NUM_LDS_BINS ? 1 : 2); }
constexpr int KERNEL_MH = 3;            // timing / byte class of the sequential Metropolis-Hastings kernel
constexpr int NUM_KERNEL_CLASSES = 4;

struct DevGraph {
    int N = 0;
    int E = 0;
    int max_deg = 0;
    int *out_ptr = nullptr, *out_idx = nullptr, *in_ptr = nullptr, *in_idx = nullptr;
    int* edge_src = nullptr;  // source vertex of each CSR edge
    uint8_t* self_flag = nullptr;  // [N] 1 where the vertex has a self-loop
    std::vector<int> h_out_ptr, h_out_idx, h_in_ptr, h_in_idx;
    std::vector<int> degree;             // |out| + |in| - self (graph.cpp:15-22)
    std::vector<char> self;
    std::vector<int> high, low;          // degree split (graph.cpp:215-228)
};

struct DevBlockmodel {
    int B = 0;
    int nnz = 0;
    bool valid = false;
    int* assign = nullptr;  // [N]
    // CSR of the block matrix (rows sorted by column) and CSC (columns sorted by row);
    // *_cum are inclusive prefix sums of the values inside each row / column.
    int *row_ptr = nullptr, *row_col = nullptr, *row_val = nullptr, *row_cum = nullptr, *coo_row = nullptr;
    int *col_ptr = nullptr, *col_row = nullptr, *col_val = nullptr, *col_cum = nullptr;
    float *row_h = nullptr, *col_h = nullptr;  // cell_entropy(M, din[col], dout[row]) per nonzero
    int *dout = nullptr, *din = nullptr, *deg = nullptr, *diag = nullptr, *drift = nullptr;
};

// A vertex list split by degree bin, for one batch.
struct BinnedList {
    int offset[NUM_BINS] = {};  // into the device list buffer
    int count[NUM_BINS] = {};
    int total = 0;
    int huge_base = 0;  // index of this batch's first huge-table offset
};

struct SweepPlan {
    int algo = -1, batches = 0;
    int* d_lists = nullptr;  // concatenated binned vertex lists
    std::vector<BinnedList> batches_async;
    int huge_pool_slots = 0;
    int* d_huge_offsets = nullptr;
    int* d_high = nullptr;   // high-degree vertices in sorted order (Metropolis-Hastings queue)
    int n_high = 0;
};

// Mutable block matrix of a Metropolis-Hastings pass: per-row open-addressing tables.
struct MhState {
    int *rbase = nullptr, *cbase = nullptr;  // [N + 1] first slot of each row / column
    int *rcap = nullptr, *ccap = nullptr;    // [N] slots of each row / column (powers of two)
    int *rkeys = nullptr, *rvals = nullptr, *ckeys = nullptr, *cvals = nullptr;
    size_t slots = 0;      // pool size of each of the four arrays
    size_t headroom = 0;   // slots beyond the initial layout, for rows that outgrow their segment
    size_t pool_limit = 0;  // slots of each pool array the kernels may use
    int *rocc = nullptr, *cocc = nullptr;    // [N] keys held by each row / column
    int* tables = nullptr;  // neighbour-block tables, one per wavefront
    size_t table_slots = 0;
    int* status = nullptr;  // [12]: next queue position, stop reason, row pool top, column pool top,
                            // commit serial, longest row, pending merge (2), heap length
    int* touch = nullptr;   // [N] serial of the last commit that moved a neighbour of each vertex
    int max_cap = 0;        // slots of the longest row or column when the tables were built
    double mean_scan = -1.0;  // mean row/column slots scanned per evaluation in the last chunk
    double mean_prop = 0.0;   // mean slots of the proposal's row per evaluation (window kernel)
    double mean_wmax = 0.0;   // mean over windows of the most row slots one evaluation visited
    int last_B = -1;          // number of blocks when mean_scan was measured
    int g_yield = 1;  // wavefronts per vertex the last chunk's window yield calls for
    double y_est = 16.0;  // queue entries a window advanced in the last chunk
    int k_max = 1;    // the most workgroups that fit on the device at once (a power of two)
    int k_max_l = 1;  // the same for the learned-mixture instantiation
    int coop_refused = 0;  // cooperative launches the runtime refused (then k_max = 1)
    unsigned long long* wstamp = nullptr;  // [N] accepted-entry stamps of the multi-workgroup pass
    void* wide = nullptr;                  // its per-entry records
    int* wst = nullptr;                    // their evaluation status, dense
    int* ctrl = nullptr;                   // its barrier counter and window control words
    int* team_hist = nullptr;     // [k_max * 512] radix histograms of workgroup teams ([2][team][256] per entry)
    double* team_part = nullptr;  // [k_max * 2] their partial delta-MDL sums and scanned slots
    int* team_arrivals = nullptr;  // [k_max] their barrier arrival counters
    int last_team = 1;            // workgroups per entry in the last chunk
    int* assign0 = nullptr;  // [N] the assignment the block-matrix snapshot was built from
    int* d_all = nullptr;   // 0..N-1
    int* d_high = nullptr;  // high-degree vertices in degree order
    int max_deg_high = 0;
};

// Several independent graphs held as one disjoint union. Graph g owns the vertices
// [vo[g], vo[g + 1]) and the block ids [bo[g], bo[g + 1]), of which the first nb[g] are in use.
// Every bo[g] is a multiple of `align`, the largest neighbour-block table any vertex uses, so the
// table slot of a block (hash_block(b) & (cap - 1)) is the same as for the graph alone; with the
// randomness keyed by graph-local ids and the MDL summed per graph, each graph follows the run it
// would have on its own.
struct DevBatch {
    int G = 0;  // 0: the context holds one graph
    int align = 0;
    int range = 0;  // block ids in use by the union: bo[G]
    std::vector<int> vo, bo, nb;
    int *d_vgraph = nullptr;  // [N] graph of each vertex
    int *d_bgraph = nullptr;  // [range] graph of each block id
    int *d_vo = nullptr, *d_bo = nullptr, *d_nb = nullptr;
    uint64_t* d_seed = nullptr;
    uint32_t* d_call = nullptr;
    uint8_t* d_active = nullptr;  // [G] graphs taking part in the current operation
    int* d_req = nullptr;         // [2 G + 2] request staging
    int* d_dbo = nullptr;         // [G] first dense id of each graph (prefix of nb)
    int* d_dense = nullptr;       // [range] dense number of each block id in use, -1 otherwise
    int* d_inv = nullptr;         // [range] block id of each dense number
    int used = 0;                 // blocks in use over all graphs (sum of nb)
    std::vector<uint8_t> h_active;
    std::vector<uint64_t> h_seed;
    std::vector<uint32_t> h_call;
    std::vector<int> h_dbo;
    // the reference's batches of every graph, split by degree bin: lists[g][k][bin] (union ids)
    std::vector<std::vector<std::array<std::vector<int>, NUM_BINS>>> lists;
    int lists_batches = 0;
    // the union plan of the last sweep and the graphs it covered
    std::vector<char> plan_active;
    int plan_batches = 0;
    std::vector<BinnedList> plan;
    int* d_plan = nullptr;
    int* d_plan_huge = nullptr;
    size_t plan_slots = 0, plan_huge_slots = 0;
    // pinned host staging of a merge round over block ids [lo, hi): scores, proposals, block
    // sizes and the new labels (each `range` long, allocated on the first merge)
    double* h_dS = nullptr;
    int *h_best = nullptr, *h_count = nullptr, *h_map = nullptr;
    // Graph mode (option 9): upper bounds over all graphs of batch k's bin counts (bound_bin[k *
    // NUM_BINS + bin]), list lengths and huge-table slots, fixed per number of batches; the plan
    // buffers sized for them, so that their addresses do not change between rounds.
    std::vector<int> bound_bin, bound_total;
    size_t bound_pool = 0;
    // the captured round: its shape (number of plan batches, key width, kernel variants, buffer
    // addresses), the executable graph, the per-round words it reads (ROUND_WORDS per plan
    // batch: list offset and count per bin, huge-table base, first list entry, list length,
    // epoch), their pinned staging, and the work counters of its persistent launches
    std::vector<long long> gkey;
    hipGraphExec_t gexec = nullptr;
    int* d_round = nullptr;
    int* h_round = nullptr;
    int* d_queue = nullptr;
    size_t round_slots = 0;
    int h_nnz = 0;
    long long graph_rounds = 0;  // rounds replayed from the graph
    long long graph_captures = 0;
};
constexpr int ROUND_OFF = 0, ROUND_CNT = NUM_BINS, ROUND_HUGE = 2 * NUM_BINS, ROUND_FIRST = 2 * NUM_BINS + 1,
              ROUND_TOTAL = 2 * NUM_BINS + 2, ROUND_EPOCH = 2 * NUM_BINS + 3, ROUND_WORDS = 2 * NUM_BINS + 4;

// Buffers of the incremental block-matrix update (allocated on first use).
struct IncBuffers {
    void *okey = nullptr, *mkey = nullptr, *dkey = nullptr;
    int *ocount = nullptr, *mcount = nullptr, *dsum = nullptr, *lb = nullptr, *rank = nullptr, *flag = nullptr,
        *tmp = nullptr;
    size_t cap = 0, tmp_ints = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    void* noalloc_scratch = nullptr;  // this device's address of the allocation-free path's array
    std::string err;
    DevGraph g;
    DevBlockmodel bm;
    SweepPlan plan;
    MhState mh;
    MlpState mlp;
    DevBatch batch;
    IncBuffers inc;
    int incremental = 1;  // 1: update the block matrix from the moved vertices' edges when few changed
    int proposal = 0;  // 0 plain, 1 learned mixture for the batched low-degree vertices
    int mh_group = 0;  // wavefronts per vertex in the Metropolis-Hastings kernel; 0 = by row length
    int mh_wide = 0;   // workgroups of the Metropolis-Hastings pass; 0 = by window yield
    int mh_team = 0;   // workgroups per window entry; 0 = by scanned row length, 1 = one
    int64_t mh_headroom = 0;  // initial table pool headroom in slots; 0 = the initial layout (>= 1M)
    int learn_maxdeg = LEARN_MAXDEG;  // learned mixture's degree bound (sequential pass; whole batched bins)
    int batch_graph = 0;  // batched sweeps replayed from a captured graph once the variant tuner settled

    // scratch
    void* keys0 = nullptr;
    void* keys1 = nullptr;
    int* vals0 = nullptr;
    int* vals1 = nullptr;
    int* flags = nullptr;   // E + 1
    int* pos = nullptr;     // E + 1
    int* hist = nullptr;
    int* scan_tmp = nullptr;
    int* tmp_b = nullptr;   // N + 1
    double* red = nullptr;  // reduction partials
    double* lp_part = nullptr;  // log posterior chunk partials and per-graph results
    size_t lp_part_slots = 0;
    int* lp_seg = nullptr;      // log posterior segments: [3 G + 1]
    int lp_seg_graphs = 0;
    int* dev_err = nullptr;
    long long* counters = nullptr;  // [8]: attempts, evals, moves
    int* words = nullptr;           // [NUM_WORDS] device words of graph-mode rounds (DynWord)

    // per-vertex move buffers
    int* ev_s = nullptr;
    uint8_t* ev_acc = nullptr;
    double* ev_dS = nullptr;
    double* ev_H = nullptr;
    double* ev_u = nullptr;
    int* list_buf = nullptr;      // [N + 1] staging for caller-given vertex lists
    int* aux_buf = nullptr;       // [2 (N + 1)]
    int* moved_epoch = nullptr;   // [N]
    int* new_block = nullptr;     // [N]
    unsigned long long* last_touch = nullptr;  // [N]
    int epoch = 0;
    int* table_pool = nullptr;    // neighbour-block tables of huge-degree vertices
    size_t table_pool_slots = 0;

    // merge buffers
    double* m_dS = nullptr;
    int* m_best = nullptr;
    int* m_map = nullptr;
    int* m_count = nullptr;

    // host mirrors
    std::vector<double> h_dS;
    std::vector<int> h_best;

    // timing of the last sweep's eval kernels (ms) and algorithmic bytes
    std::vector<int> timed;  // kernel class of each timed launch in this sweep (KERNEL_* below)
    struct TimedBin {
        int bin, variant, n;  // batched-evaluation launch: degree bin, kernel variant, vertices
    };
    std::vector<TimedBin> timed_bins;  // parallel to `timed`; bin -1 for other launches
    std::vector<hipEvent_t> events;  // two per timed launch, reused across sweeps
    // Brackets one launch on the stream with a pair of events, recorded under kernel class k.
    void timer_begin() {
        const size_t need = timed.size() * 2 + 2;
        while (events.size() < need) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) throw std::runtime_error("hipEventCreate failed");
            events.push_back(e);
        }
        (void)hipEventRecord(events[timed.size() * 2], stream);
    }
    void timer_end(int k, int bin = -1, int variant = 0, int n = 0) {
        (void)hipEventRecord(events[timed.size() * 2 + 1], stream);
        timed.push_back(k);
        timed_bins.push_back(TimedBin{bin, variant, n});
    }
    // Online choice of the kernel variant per degree bin (bins 0-2: two or four row slots per lane
    // per trip of the scans; bins 3-4: one wavefront per vertex or 4 sharing each vertex, which
    // shortens the slowest vertex when the bin holds few): while a bin's variants have fewer than
    // TUNE_TRIALS timed launches each they take turns; then the one with the lower time per
    // vertex is kept. The choice is made again after every change of the number of blocks.
    // Bin 0 (degree <= 30) has a third variant: 16 lanes per vertex, four vertices per wavefront.
    static constexpr int NUM_VARIANTS = 3;
    static constexpr int TUNE_TRIALS = 2;
    static int variants_of(int bin) { return bin == 0 ? 3 : (bin < NUM_LDS_BINS ? 2 : 1); }
    struct BinTuner {
        double ms[NUM_VARIANTS] = {0, 0, 0};
        double vertices[NUM_VARIANTS] = {0, 0, 0};
        int trials[NUM_VARIANTS] = {0, 0, 0};
        int next = 0;
    };
    BinTuner tuner[8];
    long tuner_key = -1;
    int eval_variant = 0;  // 0 = tuned, k = variant k - 1 of every bin that has it (else the first)
    // The choice is made again when the number of blocks changes (for a batch of graphs: when the
    // blocks in use cross a power of two).
    long tuning_key() const {
        if (batch.G <= 0) return bm.B;
        long total = 0;
        for (int x : batch.nb) total += x;
        int bits = 0;
        while ((1L << (bits + 1)) <= total) ++bits;
        return bits;
    }
    int pick_variant(int bin) {
        const int nv = variants_of(bin);
        if (nv == 1) return 0;
        if (eval_variant) return eval_variant - 1 < nv ? eval_variant - 1 : 0;
        const long key = tuning_key();
        if (tuner_key != key) {
            for (auto& t : tuner) t = BinTuner();
            tuner_key = key;
        }
        BinTuner& t = tuner[bin];
        for (int v = 0; v < nv; ++v)
            if (t.trials[v] < TUNE_TRIALS) {
                int w = (t.next + v) % nv;
                if (t.trials[w] < TUNE_TRIALS) {
                    t.next = (w + 1) % nv;
                    return w;
                }
            }
        int best = 0;
        for (int v = 1; v < nv; ++v)
            if (t.ms[v] * t.vertices[best] < t.ms[best] * t.vertices[v]) best = v;
        return best;
    }
    // The variant pick_variant keeps for a bin once its timing is done (without starting a trial);
    // whether every bin with vertices in `counts` is done.
    int settled_variant(int bin) const {
        const int nv = variants_of(bin);
        if (nv == 1) return 0;
        if (eval_variant) return eval_variant - 1 < nv ? eval_variant - 1 : 0;
        const BinTuner& t = tuner[bin];
        for (int v = 0; v < nv; ++v)
            if (tuner_key != tuning_key() || t.trials[v] < TUNE_TRIALS) return 0;
        int best = 0;
        for (int v = 1; v < nv; ++v)
            if (t.ms[v] * t.vertices[best] < t.ms[best] * t.vertices[v]) best = v;
        return best;
    }
    bool variants_settled(const int* counts) const {
        if (eval_variant) return true;
        for (int b = 0; b < NUM_LDS_BINS; ++b) {
            if (counts[b] == 0 || variants_of(b) == 1) continue;
            if (tuner_key != tuning_key()) return false;
            for (int v = 0; v < variants_of(b); ++v)
                if (tuner[b].trials[v] < TUNE_TRIALS) return false;
        }
        return true;
    }
    void record_variant(int bin, int variant, int n, double ms) {
        if (bin < 0 || bin >= NUM_LDS_BINS || n <= 0) return;
        BinTuner& t = tuner[bin];
        t.ms[variant] += ms;
        t.vertices[variant] += n;
        t.trials[variant] += 1;
    }

    void sync() { SBPX_HIP(hipStreamSynchronize(stream)); }
};

// sort_scan.hip
size_t scan_tmp_ints(long n);
size_t radix_hist_ints(long n);
void scan_exclusive(const int* in, int* out, int n, int* tmp, hipStream_t st);
template <typename K>
int radix_sort(K* keys, K* keys_alt, int* vals, int* vals_alt, long n, const int* shifts, int npass,
               int* hist, int* scan_tmp, hipStream_t st);
// The same over a count held in device memory (Dyn with p set; v is its upper bound), in launches
// of a fixed grid, so that a captured graph replays them with whatever count the device wrote.
// The scan writes the total to out[n] and, when total_at is set, to *total_at.
void scan_exclusive(const int* in, int* out, Dyn n, int* tmp, hipStream_t st, int* total_at = nullptr);
template <typename K>
int radix_sort(K* keys, K* keys_alt, int* vals, int* vals_alt, Dyn n, const int* shifts, int npass,
               int* hist, int* scan_tmp, hipStream_t st);
constexpr int DYN_SCAN_BLOCKS = 1024;   // workgroups of a scan over a device-held count
constexpr int DYN_SORT_BLOCKS = 2048;   // workgroups of a radix pass over a device-held count
constexpr int DYN_GRID = 2048;          // workgroups of an elementwise launch over a device-held count
inline size_t dyn_hist_ints() { return 256 * (size_t)DYN_SORT_BLOCKS + 1; }
inline int grid_for(const Dyn& n, int per) {
    const long g = ((long)n.v + per - 1) / per;
    return (int)std::max<long>(1, n.p ? std::min<long>(g, DYN_GRID) : g);
}

// blockmodel.hip
void build_blockmodel(Ctx& c);               // from c.bm.assign, resets drift
void rebuild_keep_drift(Ctx& c);             // from c.bm.assign, keeps drift
// Incremental update after a batch: prepare before the moves are committed (returns the change
// count, or -1 for a full rebuild), finish after. The moved vertices of `list` are those with
// ev_acc == 2 (new block in new_block, moved_epoch == epoch); old_assign is the assignment the
// block matrix was built from (c.bm.assign when null).
int prepare_update(Ctx& c, const int* list, int n, const int* old_assign = nullptr);
void finish_update(Ctx& c, const int* list, int n, int total);
// The same with device-held sizes (graph mode): prepare before the commit, finish after.
void prepare_update_dyn(Ctx& c, const VList& L, Dyn epoch);
void finish_update_dyn(Ctx& c);
void ensure_update_buffers(Ctx& c);
int sort_key_bits(const Ctx& c);  // bits per block id in the block matrix's sort keys
double log_posterior(Ctx& c);                // sum of M * logf(M / (dout din)) in FP64
// Per-graph log posteriors of a union (out[G]); a one-graph context gives out[0]. Each graph's
// cells are summed in fixed chunks from its first cell, so the value does not depend on the
// other graphs.
void log_posterior_segments(Ctx& c, double* out);
void check_device_error(Ctx& c);
void launch_logf_table(Ctx& c, uint32_t lo, uint32_t n, float* out);

// sweep.hip
struct EvalLaunch {
    const int* list;   // vertex ids
    int n;
    int bin;
    int learned;            // 1: proposals from the learned mixture (bins 0-2 only)
    int variant;            // kernel variant of the bin: bins 0-2: 1 = four row slots per lane per trip;
                            // bins 3-4: 1 = 4 wavefronts share each vertex
    const int* given_s;     // optional fixed proposals
    const double* given_u;  // optional fixed uniforms
    uint64_t seed;
    uint32_t call;
    uint32_t tag;
    int out_by_vertex;      // 1: outputs at index v; 0: at out_offset + position
    int out_offset;
    int propose_only;       // 1: stop after drawing the proposal
    double* out_u;          // optional: the accept uniform used
    const int* huge_offsets;
    long long* bytes;       // optional algorithmic byte counter
    const double* learned_u;  // learned bins: accept uniforms by vertex (null: the propose pass's)
    // graph mode: the count in device memory (n its upper bound), the list at list + *list_off,
    // the huge-table offsets at huge_offsets + *huge_off_at, and the launch's work counter
    const int* n_at;
    const int* list_off;
    const int* huge_off_at;
    int* queue;
};
void launch_eval(Ctx& c, const EvalLaunch& L);
void apply_accepted(Ctx& c, const int* list, int n);
void apply_accepted_dyn(Ctx& c, const VList& L, Dyn epoch);
void launch_round_zero(Ctx& c, int* queue, int nq);
void finish_degrees(Ctx& c);
// learned.hip / mlp.hip
// max_deg: the largest degree in the list (a larger neighbour-block table above LEARN_MAXDEG);
// vertices with more than skip_deg out + in list entries get an empty row (no learned part).
void learned_features(Ctx& c, const int* list, int n, int max_deg = LEARN_MAXDEG, int skip_deg = INT_MAX);
void learned_log(Ctx& c, const int* list, int n);
// The mixture's proposals (drawn, or s_given[i] for list entry i) and the net's forward and
// reverse probabilities, for the listed vertices whose features learned_features wrote.
void learned_mixture(Ctx& c, const int* list, int n, uint64_t seed, uint32_t call, const int* s_given = nullptr);
void ensure_mlp_rows(Ctx& c);
void mlp_forward(Ctx& c, const __bf16* X, int V, __bf16* H1, __bf16* H2, float* Z);
void mlp_backward(Ctx& c, const __bf16* X, const uint32_t* cmask, const int* y, int V, float* loss,
                  float* grads[3], float* gbias[3]);
void mlp_adam_step(Ctx& c, float lr);
void mlp_gather(Ctx& c, const int* idx, int n);
void mlp_sync_bf16(Ctx& c);
// cells.hip: one Delta's delta MDL, fixed merges' delta MDL, cell lookups
void launch_delta_mdl(Ctx& c, int r, int s, int n, const int* er, const int* ec, const int* ed, int k_out, int k_in,
                      int self_w, double* d_out);
void launch_merge_eval(Ctx& c, int n, const int* r, const int* s, double* d_out);
void launch_cells(Ctx& c, int n, const int* rows, const int* cols, int* d_out);
// blockmodel.hip: the fixed scratch array of the allocation-free log posterior
void* noalloc_scratch_address();
void log_posterior_noalloc(void* base, long long nnz, const int* coo_row, const int* col, const int* val,
                           const int* dout, const int* din, double* out, hipStream_t st);
// The sequential pass's workgroup scorer on the listed vertices (degree <= max_deg): candidates
// and scores [n][32].
void learned_scores_probe(Ctx& c, const int* d_list, int n, int max_deg, int* d_cand, float* d_z);
void mlp_convert_to_bf16(Ctx& c, const float* x, long n, __bf16* y);
void mlp_convert_to_f32(Ctx& c, const __bf16* x, long n, float* y);

// Heap-ordered merge selection on the device; parent: the union-find of merged blocks (B).
void merge_heap_select(Ctx& c, int num_to_merge, const double* dS, const int32_t* best, std::vector<int>& parent);
void mh_sequential(Ctx& c, const int* d_list, int n, int max_deg, uint64_t seed, uint32_t call, uint32_t tag,
                   long long* bytes, bool timed);

// merge.hip
void launch_merge_proposals(Ctx& c, uint64_t seed, uint32_t call);  // union: per-graph seeds and calls
void launch_batch_relabel(Ctx& c, int lo, int hi);  // assign[v] = map[assign[v] - lo] where in [lo, hi)
void launch_zero_drift_active(Ctx& c);              // drift = 0 on the blocks of active graphs
void launch_batch_scatter(Ctx& c, int n, int total);  // assignments of n requests into the union
void launch_batch_dense(Ctx& c);                       // dense/inv from nb (d_nb, d_dbo uploaded)
void launch_batch_trace(Ctx& c, long long* d_trace);   // per-graph sum of the block-matrix diagonal
void launch_block_counts(Ctx& c);
void launch_relabel(Ctx& c);

}  // namespace sbpx
