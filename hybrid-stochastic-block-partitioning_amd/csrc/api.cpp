// C ABI (include/sbpx.h) over the device kernels.
#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <exception>
#include <mutex>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <thread>

#include "internal.hpp"

using namespace sbpx;

struct sbpx_ctx {
    Ctx c;
    bool count_bytes = false;
};

namespace {

template <typename F>
sbpx_status guarded(sbpx_ctx* ctx, F&& f) {
    if (!ctx) return SBPX_ERR_ARG;
    try {
        f();
        return SBPX_OK;
    } catch (const ArgError& e) {
        ctx->c.err = e.what();
        return SBPX_ERR_ARG;
    } catch (const NumericError& e) {
        ctx->c.err = e.what();
        return SBPX_ERR_NUMERIC;
    } catch (const StateError& e) {
        ctx->c.err = e.what();
        return SBPX_ERR_STATE;
    } catch (const HipError& e) {
        ctx->c.err = e.what();
        return SBPX_ERR_HIP;
    } catch (const std::exception& e) {
        ctx->c.err = e.what();
        return SBPX_ERR_UNSUPPORTED;
    }
}

template <typename T>
void dalloc(T** p, size_t n) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    SBPX_HIP(hipMalloc((void**)p, std::max<size_t>(1, n) * sizeof(T)));
}
template <typename T>
void dfree(T** p) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
}

void free_batch(Ctx& c);

void free_graph_buffers(Ctx& c) {
    free_batch(c);
    dfree(&c.g.out_ptr);
    dfree(&c.g.out_idx);
    dfree(&c.g.in_ptr);
    dfree(&c.g.in_idx);
    dfree(&c.g.edge_src);
    dfree(&c.g.self_flag);
    DevBlockmodel& b = c.bm;
    dfree(&b.assign);
    dfree(&b.row_ptr);
    dfree(&b.row_col);
    dfree(&b.row_val);
    dfree(&b.row_cum);
    dfree(&b.coo_row);
    dfree(&b.col_ptr);
    dfree(&b.col_row);
    dfree(&b.col_val);
    dfree(&b.col_cum);
    dfree(&b.row_h);
    dfree(&b.col_h);
    dfree(&b.dout);
    dfree(&b.din);
    dfree(&b.deg);
    dfree(&b.diag);
    dfree(&b.drift);
    dfree(&c.keys0);
    dfree(&c.keys1);
    dfree(&c.vals0);
    dfree(&c.vals1);
    dfree(&c.flags);
    dfree(&c.pos);
    dfree(&c.hist);
    dfree(&c.scan_tmp);
    dfree(&c.tmp_b);
    dfree(&c.ev_s);
    dfree(&c.ev_acc);
    dfree(&c.ev_dS);
    dfree(&c.ev_H);
    dfree(&c.ev_u);
    dfree(&c.moved_epoch);
    dfree(&c.new_block);
    dfree(&c.last_touch);
    dfree(&c.m_dS);
    dfree(&c.m_best);
    dfree(&c.m_map);
    dfree(&c.m_count);
    dfree(&c.list_buf);
    dfree(&c.aux_buf);
    dfree(&c.plan.d_lists);
    dfree(&c.plan.d_huge_offsets);
    dfree(&c.table_pool);
    c.table_pool_slots = 0;
    MhState& h = c.mh;
    for (int** p : {&h.rbase, &h.cbase, &h.rkeys, &h.rvals, &h.ckeys, &h.cvals, &h.rocc, &h.cocc, &h.rcap,
                    &h.ccap, &h.tables, &h.status, &h.touch, &h.d_all, &h.d_high, &h.ctrl, &h.assign0})
        dfree(p);
    dfree(&h.wstamp);
    dfree(&h.team_hist);
    dfree(&h.team_part);
    dfree(&h.team_arrivals);
    if (h.wide) (void)hipFree(h.wide);
    if (h.wst) (void)hipFree(h.wst);
    h = MhState();
    c.plan = SweepPlan();
    c.bm = DevBlockmodel();
}

// ------------------------------------------------------------------------------------------
// Proposal network buffers.
void free_mlp(Ctx& c) {
    MlpState& m = c.mlp;
    for (int l = 0; l < 3; ++l) {
        for (float** p : {&m.w[l], &m.b[l], &m.gw[l], &m.gb[l], &m.m1w[l], &m.m2w[l], &m.m1b[l], &m.m2b[l]}) dfree(p);
        dfree(&m.w16[l]);
    }
    for (__bf16** p : {&m.h1, &m.h2, &m.dz, &m.dh1, &m.dh2, &m.ones, &m.ds_x, &m.x, &m.sh1, &m.sh2, &m.tx})
        dfree(p);
    for (float** p : {&m.z, &m.loss, &m.sz, &m.qf, &m.qb}) dfree(p);
    dfree(&m.ds_mask);
    dfree(&m.tmask);
    for (int** p : {&m.ds_y, &m.ds_count, &m.cand, &m.cnt, &m.ps, &m.ty, &m.tidx, &m.log_flag, &m.log_rank}) dfree(p);
    dfree(&m.pu);
    m = MlpState();
}

// Allocates the weights, Adam state and training buffers; initialises the weights (He normal
// for the hidden layers, scaled down for the scores; zero biases) from a seeded host generator.
void init_mlp(Ctx& c, uint64_t seed) {
    MlpState& m = c.mlp;
    if (!m.ready) {
        for (int l = 0; l < 3; ++l) {
            const long n = mlp_weight_count(l), nb = mlp_bias_count(l);
            for (float** p : {&m.w[l], &m.gw[l], &m.m1w[l], &m.m2w[l]}) dalloc(p, n);
            for (float** p : {&m.b[l], &m.gb[l], &m.m1b[l], &m.m2b[l]}) dalloc(p, nb);
            dalloc(&m.w16[l], n);
        }
        for (__bf16** p : {&m.h1, &m.h2, &m.dh1, &m.dh2}) dalloc(p, (size_t)MLP_BATCH * MLP_H);
        dalloc(&m.dz, (size_t)MLP_BATCH * MLP_OUT);
        dalloc(&m.ones, (size_t)MLP_BATCH);
        dalloc(&m.z, (size_t)MLP_BATCH * MLP_OUT);
        dalloc(&m.loss, (size_t)MLP_BATCH);
        dalloc(&m.ds_count, 1);
        dalloc(&m.tx, (size_t)MLP_BATCH * MLP_IN);
        dalloc(&m.tmask, (size_t)MLP_BATCH);
        dalloc(&m.ty, (size_t)MLP_BATCH);
        dalloc(&m.tidx, (size_t)MLP_BATCH);
        SBPX_HIP(hipMemset(m.ds_count, 0, sizeof(int)));
        std::vector<float> one(MLP_BATCH, 1.0f);
        float* tmp = nullptr;
        dalloc(&tmp, MLP_BATCH);
        SBPX_HIP(hipMemcpy(tmp, one.data(), sizeof(float) * MLP_BATCH, hipMemcpyHostToDevice));
        mlp_convert_to_bf16(c, tmp, MLP_BATCH, m.ones);
        c.sync();
        dfree(&tmp);
        m.ready = true;
    }
    // xoshiro256** stream from the seed, Box-Muller normals
    uint64_t st[4];
    uint64_t z = seed + 0x9E3779B97F4A7C15ull;
    for (auto& x : st) {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t y = z;
        y = (y ^ (y >> 30)) * 0xBF58476D1CE4E5B9ull;
        y = (y ^ (y >> 27)) * 0x94D049BB133111EBull;
        x = y ^ (y >> 31);
    }
    auto next = [&st]() {
        auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
        const uint64_t result = rotl(st[1] * 5, 7) * 9;
        const uint64_t t = st[1] << 17;
        st[2] ^= st[0];
        st[3] ^= st[1];
        st[1] ^= st[2];
        st[0] ^= st[3];
        st[2] ^= t;
        st[3] = rotl(st[3], 45);
        return result;
    };
    auto unit = [&next]() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); };
    for (int l = 0; l < 3; ++l) {
        const long n = mlp_weight_count(l), nb = mlp_bias_count(l);
        const int fan_in = l == 0 ? MLP_IN : MLP_H;
        const double sd = (l == 2 ? 0.5 : 1.0) * std::sqrt(2.0 / fan_in);
        std::vector<float> w(n);
        for (long i = 0; i < n; ++i) w[i] = (float)(sd * std::sqrt(-2.0 * std::log(unit())) * std::cos(6.283185307179586 * unit()));
        std::vector<float> zero(nb, 0.0f), zw(n, 0.0f);
        SBPX_HIP(hipMemcpy(m.w[l], w.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(m.b[l], zero.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(m.m1w[l], zw.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(m.m2w[l], zw.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(m.m1b[l], zero.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(m.m2b[l], zero.data(), sizeof(float) * nb, hipMemcpyHostToDevice));
    }
    m.adam_t = 0;
    mlp_sync_bf16(c);
    c.sync();
}

}  // namespace

// Per-vertex buffers of the learned sweeps ([N] rows).
void sbpx::ensure_mlp_rows(Ctx& c) {
    MlpState& m = c.mlp;
    const size_t N = (size_t)std::max(c.g.N, 1);
    if (m.rows >= (int)N) return;
    dalloc(&m.x, N * MLP_IN);
    dalloc(&m.sh1, N * MLP_H);
    dalloc(&m.sh2, N * MLP_H);
    dalloc(&m.sz, N * MLP_OUT);
    dalloc(&m.cand, N * MLP_CAND);
    dalloc(&m.cnt, N * MLP_CAND);
    dalloc(&m.qf, N);
    dalloc(&m.qb, N);
    dalloc(&m.ps, N);
    dalloc(&m.pu, N);
    m.rows = (int)N;
}

namespace {

int bin_of(int degree) {
    for (int b = 0; b < NUM_LDS_BINS; ++b)
        if (degree <= BIN_MAXDEG[b]) return b;
    return degree <= GROUP_MAXDEG ? NUM_LDS_BINS : NUM_LDS_BINS + 1;
}
int table_cap(int degree) {
    int cap = 16;
    while (cap < 2 * (degree + 2)) cap <<= 1;
    return cap;
}

void ensure_pool(Ctx& c, size_t slots) {
    if (slots <= c.table_pool_slots) return;
    dalloc(&c.table_pool, slots);
    c.table_pool_slots = slots;
}

// Orders `vertices` by degree bin (ascending); returns counts and, for the last bin,
// table offsets (7 * cap ints each) starting at 0.
void bin_vertices(const Ctx& c, const int* vertices, int n, std::vector<int>& out, int counts[NUM_BINS],
                  std::vector<int>& huge_off, std::vector<int>* perm = nullptr) {
    out.clear();
    out.reserve(n);
    if (perm) perm->clear();
    huge_off.assign(1, 0);
    for (int b = 0; b < NUM_BINS; ++b) {
        counts[b] = 0;
        for (int i = 0; i < n; ++i) {
            int v = vertices[i];
            if (bin_of(c.g.degree[v]) != b) continue;
            out.push_back(v);
            if (perm) perm->push_back(i);
            counts[b]++;
            if (b == NUM_BINS - 1) huge_off.push_back(huge_off.back() + 7 * table_cap(c.g.degree[v] + 1));
        }
    }
}

void require_graph(Ctx& c) {
    if (c.g.N <= 0) throw StateError("no graph uploaded");
}
void require_bm(Ctx& c) {
    require_graph(c);
    if (c.batch.G > 0) throw StateError("the context holds a batch of graphs: use the sbpx_batch_* calls");
    if (!c.bm.valid) throw StateError("no blockmodel built");
}
void require_batch(Ctx& c) {
    require_graph(c);
    if (c.batch.G <= 0) throw StateError("no batch of graphs defined");
}

void free_batch(Ctx& c) {
    DevBatch& b = c.batch;
    if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
    b.gexec = nullptr;
    dfree(&b.d_round);
    dfree(&b.d_queue);
    if (b.h_round) (void)hipHostFree(b.h_round);
    for (int** p : {&b.d_vgraph, &b.d_bgraph, &b.d_vo, &b.d_bo, &b.d_nb, &b.d_req, &b.d_plan, &b.d_plan_huge,
                    &b.d_dbo, &b.d_dense, &b.d_inv})
        dfree(p);
    dfree(&b.d_seed);
    dfree(&b.d_call);
    dfree(&b.d_active);
    if (b.h_dS) (void)hipHostFree(b.h_dS);
    for (int* p : {b.h_best, b.h_count, b.h_map})
        if (p) (void)hipHostFree(p);
    b = DevBatch();
}

double model_term(long N, long E, int B) {
    // entropy.cpp:584-589 without the log posterior
    double x = pow(B, 2) / E;
    double h = ((1 + x) * log(1 + x)) - (x * log(x));
    return (E * h) + (N * log(B));
}

// Evaluates a binned vertex list (each wave one vertex) with outputs at the vertex index. Each
// bin's launch is bracketed by events when timed.
void eval_vertex_list(sbpx_ctx* ctx, const int* d_list, const BinnedList& bl, const int* d_huge_off,
                      uint64_t seed, uint32_t call, uint32_t tag, long long* bytes, bool timed, bool learned = false) {
    Ctx& c = ctx->c;
    // the vertices the learned proposal covers: bins 0-2, and the bins 3-4 whose degrees stay
    // within the context's learned degree bound; contiguous in the list
    int nb = 3;
    while (nb < NUM_LDS_BINS && BIN_MAXDEG[nb] <= c.learn_maxdeg) ++nb;
    const int* low = d_list + bl.offset[0];
    int nlow = 0;
    for (int b = 0; b < nb; ++b) {
        if (bl.count[b] > 0 && bl.offset[b] != bl.offset[0] + nlow) throw StateError("degree bins not contiguous");
        nlow += bl.count[b];
    }
    const int nlog = bl.count[0] + bl.count[1] + bl.count[2];  // training rows: degree <= 126
    if (learned && nlow > 0) {
        // plain proposals and accept uniforms (vertex-indexed), then the mixture's draws
        for (int b = 0; b < nb; ++b) {
            if (bl.count[b] == 0) continue;
            EvalLaunch L{};
            L.list = d_list + bl.offset[b];
            L.n = bl.count[b];
            L.bin = b;
            L.seed = seed;
            L.call = call;
            L.tag = tag;
            L.out_by_vertex = 1;
            L.propose_only = 1;
            L.out_u = c.ev_u;
            launch_eval(c, L);
        }
        learned_features(c, low, nlow, BIN_MAXDEG[nb - 1]);
        learned_mixture(c, low, nlow, seed, call);
    } else if (c.mlp.logging && nlog > 0) {
        learned_features(c, low, nlog);  // same snapshot as the evaluation
    }
    for (int b = 0; b < NUM_BINS; ++b) {
        if (bl.count[b] == 0) continue;
        if (timed) c.timer_begin();
        EvalLaunch L{};
        L.list = d_list + bl.offset[b];
        L.n = bl.count[b];
        L.bin = b;
        L.variant = c.pick_variant(b);
        L.learned = learned && b < nb;
        L.seed = seed;
        L.call = call;
        L.tag = tag;
        L.out_by_vertex = 1;
        L.huge_offsets = d_huge_off;
        L.bytes = bytes ? bytes + bin_class(b) : nullptr;
        launch_eval(c, L);
        if (timed) c.timer_end(bin_class(b), b, L.variant, L.n);
    }
}

// Reads back a sweep's counters and event timings into stats and feeds the variant tuner.
void sweep_stats(sbpx_ctx* ctx, sbpx_sweep_stats* stats) {
    Ctx& c = ctx->c;
    long long cnt[8];
    double dsum = 0.0;
    SBPX_HIP(hipMemcpyAsync(cnt, c.counters, sizeof(cnt), hipMemcpyDeviceToHost, c.stream));
    SBPX_HIP(hipMemcpyAsync(&dsum, c.red + 2048, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    c.sync();
    check_device_error(c);
    std::vector<float> launch_ms(c.timed.size(), 0.0f);
    for (size_t i = 0; i < c.timed.size(); ++i) {
        SBPX_HIP(hipEventElapsedTime(&launch_ms[i], c.events[2 * i], c.events[2 * i + 1]));
        const Ctx::TimedBin& tb = c.timed_bins[i];
        c.record_variant(tb.bin, tb.variant, tb.n, launch_ms[i]);
    }
    if (stats) {
        *stats = sbpx_sweep_stats{};
        for (size_t i = 0; i < c.timed.size(); ++i) {
            stats->eval_ms[c.timed[i]] += launch_ms[i];
            stats->eval_launches[c.timed[i]] += 1;
        }
        stats->attempts = cnt[0];
        stats->evals = cnt[1];
        stats->moves = cnt[2];
        stats->dS_accepted = dsum;
        for (int b = 0; b < NUM_KERNEL_CLASSES; ++b)
            stats->eval_bytes[b] = ctx->count_bytes ? (double)cnt[4 + b] : 0.0;
    }
}

void build_plan(Ctx& c, int algo, int batches) {
    if (c.plan.algo == algo && c.plan.batches == batches) return;
    SweepPlan& p = c.plan;
    dfree(&p.d_lists);
    dfree(&p.d_huge_offsets);
    p.batches_async.clear();
    p.algo = algo;
    p.batches = batches;
    const int N = c.g.N;
    std::vector<int> base;
    if (algo == SBPX_ALGO_HYBRID) {
        base = c.g.low;
    } else {
        base.resize(N);
        std::iota(base.begin(), base.end(), 0);
    }
    std::vector<int> lists;
    std::vector<int> huge_all;
    size_t pool_need = 0;
    const int n = (int)base.size();
    if (n > 0 && algo != SBPX_ALGO_METROPOLIS_HASTINGS) {
        // finetune.cpp:43-47, 352-366: batch_size = ceil(n / batches), n / batch_size batches
        double num_batches = batches;
        int batch_size = int(ceil(n / num_batches));
        int nb = n / batch_size;
        for (int b = 0; b < nb; ++b) {
            int start = b * batch_size;
            int end = std::min(n, (b + 1) * batch_size);
            BinnedList bl;
            std::vector<int> ordered, hoff;
            bin_vertices(c, base.data() + start, end - start, ordered, bl.count, hoff);
            int off = (int)lists.size();  // the bins follow one another
            for (int k = 0; k < NUM_BINS; ++k) {
                bl.offset[k] = off;
                off += bl.count[k];
            }
            bl.total = end - start;
            bl.huge_base = (int)huge_all.size();
            lists.insert(lists.end(), ordered.begin(), ordered.end());
            huge_all.insert(huge_all.end(), hoff.begin(), hoff.end());
            pool_need = std::max(pool_need, (size_t)hoff.back());
            p.batches_async.push_back(bl);
        }
    }
    dalloc(&p.d_lists, lists.size());
    if (!lists.empty())
        SBPX_HIP(hipMemcpy(p.d_lists, lists.data(), sizeof(int) * lists.size(), hipMemcpyHostToDevice));
    dalloc(&p.d_huge_offsets, huge_all.size());
    if (!huge_all.empty())
        SBPX_HIP(hipMemcpy(p.d_huge_offsets, huge_all.data(), sizeof(int) * huge_all.size(),
                           hipMemcpyHostToDevice));
    ensure_pool(c, pool_need);
}

}  // namespace

// Device copies of small host arrays for the one-move entry points below, freed on scope exit.
namespace {
template <typename T>
struct DevArray {
    T* p = nullptr;
    explicit DevArray(size_t n) { dalloc(&p, std::max<size_t>(n, 1)); }
    DevArray(const T* src, size_t n, hipStream_t st) : DevArray(n) {
        if (n) SBPX_HIP(hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
    }
    ~DevArray() { dfree(&p); }
};
}  // namespace

extern "C" {

sbpx_status sbpx_create(sbpx_ctx** out, int device) {
    if (!out) return SBPX_ERR_ARG;
    *out = nullptr;
    sbpx_ctx* ctx = new sbpx_ctx();
    sbpx_status st = guarded(ctx, [&] {
        int n = 0;
        SBPX_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw ArgError("no such device");
        ctx->c.device = device;
        SBPX_HIP(hipSetDevice(device));
        SBPX_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
        ctx->c.noalloc_scratch = noalloc_scratch_address();  // with the context's device current
        dalloc(&ctx->c.dev_err, 4);
        SBPX_HIP(hipMemset(ctx->c.dev_err, 0, 4 * sizeof(int)));
        dalloc(&ctx->c.counters, 8);
        dalloc(&ctx->c.red, 4096);
        dalloc(&ctx->c.words, NUM_WORDS);
        SBPX_HIP(hipMemset(ctx->c.words, 0, NUM_WORDS * sizeof(int)));
    });
    if (st != SBPX_OK) {
        static thread_local std::string last;
        last = ctx->c.err;
        delete ctx;
        return st;
    }
    *out = ctx;
    return SBPX_OK;
}

void sbpx_destroy(sbpx_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    free_graph_buffers(ctx->c);
    free_mlp(ctx->c);
    dfree(&ctx->c.dev_err);
    dfree(&ctx->c.counters);
    dfree(&ctx->c.red);
    dfree(&ctx->c.words);
    dfree(&ctx->c.lp_part);
    dfree(&ctx->c.lp_seg);
    for (hipEvent_t e : ctx->c.events) (void)hipEventDestroy(e);
    if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
}

const char* sbpx_last_error(const sbpx_ctx* ctx) { return ctx ? ctx->c.err.c_str() : "null context"; }

sbpx_status sbpx_set_option(sbpx_ctx* ctx, int option, int64_t value) {
    return guarded(ctx, [&] {
        if (option == 1) {
            ctx->count_bytes = value != 0;
        } else if (option == 4) {
            if (value < 0 || value > 2)
                throw ArgError("proposal must be 0 (plain), 1 (learned mixture in the batched sweeps) or 2 (also in "
                               "the sequential Metropolis-Hastings pass)");
            ctx->c.proposal = (int)value;
        } else if (option == 3) {
            if (value < 0 || value > 3) throw ArgError("evaluation kernel variant must be 0 (tuned), 1, 2 or 3");
            ctx->c.eval_variant = (int)value;
        } else if (option == 5) {
            if (value != 0 && value != 1) throw ArgError("incremental update must be 0 (off) or 1 (on)");
            ctx->c.incremental = (int)value;
        } else if (option == 2) {
            if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8 && value != 16)
                throw ArgError("wavefronts per vertex must be 0 (by row length), 1, 2, 4, 8 or 16");
            ctx->c.mh_group = (int)value;
        } else if (option == 6) {
            if (value < 0 || value > 512 || (value & (value - 1)) != 0)
                throw ArgError("Metropolis-Hastings workgroups must be 0 (by window yield) or a power of two up to 512");
            ctx->c.mh_wide = (int)value;
        } else if (option == 7) {
            if (value < 0 || value > (int64_t)1 << 30) throw ArgError("table headroom must be in [0, 2^30] slots");
            ctx->c.mh_headroom = value;
            ctx->c.mh.headroom = (size_t)value;
        } else if (option == 8) {
            if (value < 0 || value > LEARN_MH_MAXDEG)
                throw ArgError("learned degree bound must be in [0, 510]");
            ctx->c.learn_maxdeg = (int)value;
        } else if (option == 9) {
            if (value != 0 && value != 1) throw ArgError("graph replay of batched sweeps must be 0 (off) or 1 (on)");
            ctx->c.batch_graph = (int)value;
        } else if (option == 10) {
            if (value < 0 || value > 32 || (value & (value - 1)) != 0)
                throw ArgError("workgroups per entry must be 0 (by row length) or a power of two up to 32");
            ctx->c.mh_team = (int)value;
        } else {
            throw ArgError("unknown option");
        }
    });
}

sbpx_status sbpx_graph_upload(sbpx_ctx* ctx, int32_t N, int64_t E, const int32_t* out_ptr, const int32_t* out_idx,
                              const int32_t* in_ptr, const int32_t* in_idx) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        if (N <= 0 || E < 0 || E >= INT32_MAX / 2) throw ArgError("bad graph size");
        if (!out_ptr || !in_ptr || (E > 0 && (!out_idx || !in_idx))) throw ArgError("null graph array");
        if (out_ptr[0] != 0 || in_ptr[0] != 0 || out_ptr[N] != E || in_ptr[N] != E)
            throw ArgError("CSR/CSC pointers do not span num_edges");
        SBPX_HIP(hipSetDevice(c.device));
        free_graph_buffers(c);
        DevGraph& g = c.g;
        g.N = N;
        g.E = (int)E;
        g.h_out_ptr.assign(out_ptr, out_ptr + N + 1);
        g.h_in_ptr.assign(in_ptr, in_ptr + N + 1);
        g.h_out_idx.assign(out_idx, out_idx + E);
        g.h_in_idx.assign(in_idx, in_idx + E);
        g.self.assign(N, 0);
        std::vector<int> edge_src(E);
        std::vector<int> in_self(N, 0);
        for (int v = 0; v < N; ++v) {
            if (out_ptr[v + 1] < out_ptr[v] || in_ptr[v + 1] < in_ptr[v]) throw ArgError("pointers not monotone");
            for (int j = out_ptr[v]; j < out_ptr[v + 1]; ++j) {
                int w = out_idx[j];
                if (w < 0 || w >= N) throw ArgError("edge endpoint out of range");
                if (w == v) g.self[v] = 1;
                edge_src[j] = v;
            }
            for (int j = in_ptr[v]; j < in_ptr[v + 1]; ++j) {
                int w = in_idx[j];
                if (w < 0 || w >= N) throw ArgError("edge endpoint out of range");
                if (w == v) in_self[v] = 1;
            }
            if (in_self[v] != g.self[v]) throw ArgError("self-loop present in only one of CSR/CSC");
        }
        // degree = |out| + |in| - self (graph.cpp:15-22); top 15% after std::sort with '>'
        // are the high-degree vertices (graph.cpp:215-228)
        g.degree.resize(N);
        g.max_deg = 0;
        for (int v = 0; v < N; ++v) {
            g.degree[v] = (out_ptr[v + 1] - out_ptr[v]) + (in_ptr[v + 1] - in_ptr[v]) - g.self[v];
            g.max_deg = std::max(g.max_deg, g.degree[v]);
        }
        std::vector<int> idx(N);
        std::iota(idx.begin(), idx.end(), 0);
        const std::vector<int>& vd = g.degree;
        std::sort(idx.data(), idx.data() + idx.size(), [&vd](size_t i1, size_t i2) { return vd[i1] > vd[i2]; });
        g.high.clear();
        g.low.clear();
        for (int index = 0; index < N; ++index) {
            if (index < 0.15 * N) g.high.push_back(idx[index]);
            else g.low.push_back(idx[index]);
        }
        // device buffers
        dalloc(&g.out_ptr, N + 1);
        dalloc(&g.in_ptr, N + 1);
        dalloc(&g.out_idx, E);
        dalloc(&g.in_idx, E);
        dalloc(&g.edge_src, E);
        dalloc(&g.self_flag, N);
        SBPX_HIP(hipMemcpy(g.self_flag, g.self.data(), N, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(g.out_ptr, out_ptr, sizeof(int) * (N + 1), hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(g.in_ptr, in_ptr, sizeof(int) * (N + 1), hipMemcpyHostToDevice));
        if (E > 0) {
            SBPX_HIP(hipMemcpy(g.out_idx, out_idx, sizeof(int) * E, hipMemcpyHostToDevice));
            SBPX_HIP(hipMemcpy(g.in_idx, in_idx, sizeof(int) * E, hipMemcpyHostToDevice));
            SBPX_HIP(hipMemcpy(g.edge_src, edge_src.data(), sizeof(int) * E, hipMemcpyHostToDevice));
        }
        DevBlockmodel& b = c.bm;
        const size_t NB = (size_t)N + 1, EE = (size_t)E + 1;
        dalloc(&b.assign, N);
        dalloc(&b.row_ptr, NB);
        dalloc(&b.col_ptr, NB);
        dalloc(&b.row_col, EE);
        dalloc(&b.row_val, EE);
        dalloc(&b.row_cum, EE);
        dalloc(&b.coo_row, EE);
        dalloc(&b.col_row, EE);
        dalloc(&b.col_val, EE);
        dalloc(&b.col_cum, EE);
        dalloc(&b.row_h, EE);
        dalloc(&b.col_h, EE);
        dalloc(&b.dout, NB);
        dalloc(&b.din, NB);
        dalloc(&b.deg, NB);
        dalloc(&b.diag, NB);
        dalloc(&b.drift, NB);
        SBPX_HIP(hipMemset(b.drift, 0, sizeof(int) * NB));
        dalloc((uint64_t**)&c.keys0, EE);
        dalloc((uint64_t**)&c.keys1, EE);
        dalloc(&c.vals0, EE);
        dalloc(&c.vals1, EE);
        dalloc(&c.flags, EE);
        dalloc(&c.pos, EE);
        dalloc(&c.hist, radix_hist_ints(E));
        dalloc(&c.scan_tmp, scan_tmp_ints(std::max<long>(E, 256L * ((E + 4095) / 4096) + 1) + NB));
        dalloc(&c.tmp_b, NB);
        dalloc(&c.ev_s, NB);
        dalloc(&c.ev_acc, NB);
        dalloc(&c.ev_dS, NB);
        dalloc(&c.ev_H, NB);
        dalloc(&c.ev_u, NB);
        dalloc(&c.moved_epoch, NB);
        dalloc(&c.new_block, NB);
        dalloc(&c.last_touch, NB);
        SBPX_HIP(hipMemset(c.moved_epoch, 0, sizeof(int) * NB));
        SBPX_HIP(hipMemset(c.last_touch, 0, sizeof(unsigned long long) * NB));
        c.epoch = 0;
        dalloc(&c.m_dS, NB);
        dalloc(&c.m_best, NB);
        dalloc(&c.m_map, NB);
        dalloc(&c.m_count, NB);
        dalloc(&c.list_buf, NB);
        dalloc(&c.aux_buf, 2 * NB);
        c.plan = SweepPlan();
        MhState& h = c.mh;
        dalloc(&h.rbase, NB + 1);
        dalloc(&h.cbase, NB + 1);
        dalloc(&h.rocc, NB);
        dalloc(&h.cocc, NB);
        dalloc(&h.rcap, NB);
        dalloc(&h.ccap, NB);
        dalloc(&h.status, 12);
        dalloc(&h.touch, N);
        SBPX_HIP(hipMemset(h.status, 0, sizeof(int) * 12));
        SBPX_HIP(hipMemset(h.touch, 0, sizeof(int) * std::max(N, 1)));
        dalloc(&h.d_all, NB);
        dalloc(&h.d_high, NB);
        {
            std::vector<int> all(N);
            std::iota(all.begin(), all.end(), 0);
            SBPX_HIP(hipMemcpy(h.d_all, all.data(), sizeof(int) * N, hipMemcpyHostToDevice));
            if (!g.high.empty())
                SBPX_HIP(hipMemcpy(h.d_high, g.high.data(), sizeof(int) * g.high.size(), hipMemcpyHostToDevice));
            h.max_deg_high = 0;
            for (int v : g.high) h.max_deg_high = std::max(h.max_deg_high, g.degree[v]);
        }
        SBPX_HIP(hipDeviceSynchronize());
    });
}

sbpx_status sbpx_graph_split(sbpx_ctx* ctx, int32_t* high, int32_t* num_high, int32_t* low) {
    return guarded(ctx, [&] {
        require_graph(ctx->c);
        const DevGraph& g = ctx->c.g;
        if (num_high) *num_high = (int32_t)g.high.size();
        if (high) std::copy(g.high.begin(), g.high.end(), high);
        if (low) std::copy(g.low.begin(), g.low.end(), low);
    });
}

sbpx_status sbpx_blockmodel_build(sbpx_ctx* ctx, int32_t B, const int32_t* assignment) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_graph(c);
        if (c.batch.G > 0) throw StateError("the context holds a batch of graphs: use the sbpx_batch_* calls");
        if (B < 1 || B > c.g.N) throw ArgError("num_blocks must be in [1, N]");
        if (!assignment) throw ArgError("null assignment");
        for (int v = 0; v < c.g.N; ++v)
            if (assignment[v] < 0 || assignment[v] >= B) throw ArgError("block id out of range [0, B)");
        SBPX_HIP(hipSetDevice(c.device));
        SBPX_HIP(hipMemcpyAsync(c.bm.assign, assignment, sizeof(int) * c.g.N, hipMemcpyHostToDevice, c.stream));
        c.bm.B = B;
        build_blockmodel(c);
        check_device_error(c);
    });
}

sbpx_status sbpx_blockmodel_info(sbpx_ctx* ctx, int32_t* B, int64_t* nnz) {
    return guarded(ctx, [&] {
        require_bm(ctx->c);
        if (B) *B = ctx->c.bm.B;
        if (nnz) *nnz = ctx->c.bm.nnz;
    });
}

sbpx_status sbpx_blockmodel_download(sbpx_ctx* ctx, int32_t* assignment, int32_t* dout, int32_t* din,
                                     int32_t* deg) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        const int B = c.bm.B;
        if (assignment)
            SBPX_HIP(hipMemcpyAsync(assignment, c.bm.assign, sizeof(int) * c.g.N, hipMemcpyDeviceToHost, c.stream));
        if (dout) SBPX_HIP(hipMemcpyAsync(dout, c.bm.dout, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
        if (din) SBPX_HIP(hipMemcpyAsync(din, c.bm.din, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
        if (deg) SBPX_HIP(hipMemcpyAsync(deg, c.bm.deg, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

sbpx_status sbpx_blockmodel_coo(sbpx_ctx* ctx, int by_column, int32_t* rows, int32_t* cols, int32_t* vals) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        const int nnz = c.bm.nnz, B = c.bm.B;
        std::vector<int> ptr(B + 1), idx(nnz), val(nnz);
        int* dptr = by_column ? c.bm.col_ptr : c.bm.row_ptr;
        int* didx = by_column ? c.bm.col_row : c.bm.row_col;
        int* dval = by_column ? c.bm.col_val : c.bm.row_val;
        SBPX_HIP(hipMemcpyAsync(ptr.data(), dptr, sizeof(int) * (B + 1), hipMemcpyDeviceToHost, c.stream));
        if (nnz) {
            SBPX_HIP(hipMemcpyAsync(idx.data(), didx, sizeof(int) * nnz, hipMemcpyDeviceToHost, c.stream));
            SBPX_HIP(hipMemcpyAsync(val.data(), dval, sizeof(int) * nnz, hipMemcpyDeviceToHost, c.stream));
        }
        c.sync();
        for (int x = 0; x < B; ++x) {
            for (int j = ptr[x]; j < ptr[x + 1]; ++j) {
                if (by_column) {
                    rows[j] = idx[j];
                    cols[j] = x;
                } else {
                    rows[j] = x;
                    cols[j] = idx[j];
                }
                vals[j] = val[j];
            }
        }
    });
}

sbpx_status sbpx_blockmodel_trace(sbpx_ctx* ctx, int64_t* trace, int64_t* edges) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        std::vector<int> diag(c.bm.B);
        SBPX_HIP(hipMemcpyAsync(diag.data(), c.bm.diag, sizeof(int) * c.bm.B, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (trace) *trace = std::accumulate(diag.begin(), diag.end(), (int64_t)0);
        if (edges) *edges = c.g.E;
    });
}

sbpx_status sbpx_log_posterior(sbpx_ctx* ctx, double* out) {
    return guarded(ctx, [&] {
        require_bm(ctx->c);
        *out = log_posterior(ctx->c);
    });
}

sbpx_status sbpx_delta_mdl(sbpx_ctx* ctx, int32_t r, int32_t s, int32_t n, const int32_t* rows, const int32_t* cols,
                           const int32_t* deltas, int32_t k_out, int32_t k_in, int32_t self_weight, double* dS) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        const int B = c.bm.B;
        if (r < 0 || r >= B || s < 0 || s >= B || r == s) throw ArgError("blocks must be distinct and in [0, B)");
        if (n < 0 || (n > 0 && (!rows || !cols || !deltas)) || !dS) throw ArgError("bad Delta arrays");
        std::vector<std::array<int, 3>> e(n);
        for (int i = 0; i < n; ++i) {
            if (rows[i] < 0 || rows[i] >= B || cols[i] < 0 || cols[i] >= B) throw ArgError("Delta cell out of range");
            if (rows[i] != r && rows[i] != s && cols[i] != r && cols[i] != s)
                throw ArgError("Delta cell in neither row/column r nor s");
            e[i] = {rows[i], cols[i], deltas[i]};
        }
        std::sort(e.begin(), e.end());
        for (int i = 1; i < n; ++i)
            if (e[i][0] == e[i - 1][0] && e[i][1] == e[i - 1][1]) throw ArgError("Delta cell listed twice");
        std::vector<int> er(n), ec(n), ed(n);
        for (int i = 0; i < n; ++i) {
            er[i] = e[i][0];
            ec[i] = e[i][1];
            ed[i] = e[i][2];
        }
        DevArray<int> der(er.data(), n, c.stream), dec(ec.data(), n, c.stream), ded(ed.data(), n, c.stream);
        DevArray<double> dout(1);
        launch_delta_mdl(c, r, s, n, der.p, dec.p, ded.p, k_out, k_in, self_weight, dout.p);
        SBPX_HIP(hipMemcpyAsync(dS, dout.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (std::isnan(*dS) || std::isinf(*dS)) throw NumericError("NaN or Inf in a delta MDL");
    });
}

sbpx_status sbpx_merge_eval(sbpx_ctx* ctx, int32_t n, const int32_t* r, const int32_t* s, double* dS) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        if (n < 0 || (n > 0 && (!r || !s || !dS))) throw ArgError("bad merge arrays");
        for (int i = 0; i < n; ++i)
            if (r[i] < 0 || r[i] >= c.bm.B || s[i] < 0 || s[i] >= c.bm.B || r[i] == s[i])
                throw ArgError("merge blocks must be distinct and in [0, B)");
        if (n == 0) return;
        DevArray<int> dr(r, n, c.stream), ds(s, n, c.stream);
        DevArray<double> out(n);
        launch_merge_eval(c, n, dr.p, ds.p, out.p);
        SBPX_HIP(hipMemcpyAsync(dS, out.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

sbpx_status sbpx_blockmodel_cells(sbpx_ctx* ctx, int32_t n, const int32_t* rows, const int32_t* cols, int32_t* vals) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        if (n < 0 || (n > 0 && (!rows || !cols || !vals))) throw ArgError("bad cell arrays");
        for (int i = 0; i < n; ++i)
            if (rows[i] < 0 || rows[i] >= c.bm.B || cols[i] < 0 || cols[i] >= c.bm.B)
                throw ArgError("block id out of range [0, B)");
        if (n == 0) return;
        DevArray<int> dr(rows, n, c.stream), dc(cols, n, c.stream), out(n);
        launch_cells(c, n, dr.p, dc.p, out.p);
        SBPX_HIP(hipMemcpyAsync(vals, out.p, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

sbpx_status sbpx_blockmodel_line(sbpx_ctx* ctx, int32_t x, int by_column, int32_t cap, int32_t* idx, int32_t* vals,
                                 int32_t* count) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        if (x < 0 || x >= c.bm.B) throw ArgError("block id out of range [0, B)");
        if (!count || cap < 0 || (cap > 0 && (!idx || !vals))) throw ArgError("bad output arrays");
        const int* ptr = by_column ? c.bm.col_ptr : c.bm.row_ptr;
        int lohi[2];
        SBPX_HIP(hipMemcpyAsync(lohi, ptr + x, sizeof(lohi), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        const int len = lohi[1] - lohi[0];
        *count = len;
        const int k = std::min(len, (int)cap);
        if (k > 0) {
            SBPX_HIP(hipMemcpyAsync(idx, (by_column ? c.bm.col_row : c.bm.row_col) + lohi[0], sizeof(int) * k,
                                    hipMemcpyDeviceToHost, c.stream));
            SBPX_HIP(hipMemcpyAsync(vals, (by_column ? c.bm.col_val : c.bm.row_val) + lohi[0], sizeof(int) * k,
                                    hipMemcpyDeviceToHost, c.stream));
            c.sync();
        }
    });
}

sbpx_status sbpx_log_posterior_noalloc(sbpx_ctx* ctx, int64_t nnz, const int32_t* coo_row, const int32_t* col,
                                       const int32_t* val, const int32_t* dout, const int32_t* din, double* out,
                                       void* stream) {
    return guarded(ctx, [&] {
        if (nnz < 0 || nnz > SBPX_NOALLOC_MAX_CELLS) throw ArgError("cell count must be in [0, SBPX_NOALLOC_MAX_CELLS]");
        if (nnz > 0 && (!coo_row || !col || !val || !dout || !din)) throw ArgError("null cell or degree array");
        if (!out) throw ArgError("null output");
        if (!ctx->c.noalloc_scratch) throw StateError("no scratch array on this device");
        SBPX_HIP(hipSetDevice(ctx->c.device));
        log_posterior_noalloc(ctx->c.noalloc_scratch, nnz, coo_row, col, val, dout, din, out,
                              stream ? (hipStream_t)stream : ctx->c.stream);
    });
}

sbpx_status sbpx_mdl(sbpx_ctx* ctx, double* out) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        double lp = log_posterior(c);
        *out = model_term(c.g.N, c.g.E, c.bm.B) - lp;
    });
}

// Accept uniforms of a fixed evaluation, by vertex, for the learned bins.
__global__ void scatter_u_kernel(const int* __restrict__ list, int n, const double* __restrict__ u,
                                 double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[list[i]] = u[i];
}

static void eval_fixed_chunk(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s, const double* u,
                             double* dS, double* H, uint8_t* acc, uint64_t seed, uint32_t call, int tag,
                             int32_t* s_out, double* u_out, bool learned, float* qf, float* qb);

// Requests longer than the per-vertex buffers (N entries) run in chunks. learned: the proposals
// are the learned mixture's (Hastings ratio of the mixture; qf / qb receive the network's forward
// and reverse probabilities, -1 where the vertex has no learned part).
static void eval_fixed(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s, const double* u, double* dS,
                       double* H, uint8_t* acc, uint64_t seed, uint32_t call, int tag, int32_t* s_out,
                       double* u_out, bool learned = false, float* qf = nullptr, float* qb = nullptr) {
    Ctx& c = ctx->c;
    require_bm(c);
    if (n < 0) throw ArgError("negative count");
    const int chunk = c.g.N;
    for (int off = 0; off < n; off += chunk) {
        int m = std::min(chunk, n - off);
        eval_fixed_chunk(ctx, m, v + off, s ? s + off : nullptr, u ? u + off : nullptr, dS ? dS + off : nullptr,
                         H ? H + off : nullptr, acc ? acc + off : nullptr, seed, call, tag,
                         s_out ? s_out + off : nullptr, u_out ? u_out + off : nullptr, learned,
                         qf ? qf + off : nullptr, qb ? qb + off : nullptr);
    }
}

static void eval_fixed_chunk(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s, const double* u,
                             double* dS, double* H, uint8_t* acc, uint64_t seed, uint32_t call, int tag,
                             int32_t* s_out, double* u_out, bool learned, float* qf, float* qb) {
    Ctx& c = ctx->c;
    if (n == 0) return;
    for (int i = 0; i < n; ++i)
        if (v[i] < 0 || v[i] >= c.g.N) throw ArgError("vertex out of range");
    if (s)
        for (int i = 0; i < n; ++i)
            if (s[i] < 0 || s[i] >= c.bm.B) throw ArgError("block id out of range [0, B)");
    std::vector<int> ordered, hoff, perm;
    int counts[NUM_BINS];
    bin_vertices(c, v, n, ordered, counts, hoff, &perm);
    std::vector<int> sv(n);
    std::vector<double> uv(n);
    if (s) {
        for (int i = 0; i < n; ++i) {
            sv[i] = s[perm[i]];
            uv[i] = u[perm[i]];
        }
    }
    int* d_list = c.list_buf;
    int* d_s = c.aux_buf;
    double* d_u = (double*)(c.m_dS);
    SBPX_HIP(hipMemcpyAsync(d_list, ordered.data(), sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
    if (s) {
        SBPX_HIP(hipMemcpyAsync(d_s, sv.data(), sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
        SBPX_HIP(hipMemcpyAsync(d_u, uv.data(), sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    }
    ensure_pool(c, (size_t)hoff.back());
    int* d_hoff = c.aux_buf + (c.g.N + 1);
    SBPX_HIP(hipMemcpyAsync(d_hoff, hoff.data(), sizeof(int) * hoff.size(), hipMemcpyHostToDevice, c.stream));
    // the learned bins (as a sweep covers them): contiguous at the start of the ordered list
    int nb = 0, nlow = 0;
    if (learned) {
        nb = 3;
        while (nb < NUM_LDS_BINS && BIN_MAXDEG[nb] <= c.learn_maxdeg) ++nb;
        for (int b = 0; b < nb; ++b) nlow += counts[b];
        if (nlow > 0) {
            learned_features(c, d_list, nlow, BIN_MAXDEG[nb - 1]);
            learned_mixture(c, d_list, nlow, 0, 0, d_s);
            scatter_u_kernel<<<(nlow + 255) / 256, 256, 0, c.stream>>>(d_list, nlow, d_u, c.mlp.pu);
            SBPX_HIP(hipGetLastError());
        }
    }
    int off = 0;
    for (int b = 0; b < NUM_BINS; ++b) {
        if (counts[b]) {
            EvalLaunch L{};
            L.list = d_list + off;
            L.n = counts[b];
            L.bin = b;
            L.learned = learned && b < nb;
            L.learned_u = c.mlp.pu;
            L.variant = c.eval_variant ? c.pick_variant(b) : 0;  // a fixed variant applies here too
            L.given_s = s ? d_s + off : nullptr;
            L.given_u = s ? d_u + off : nullptr;
            L.seed = seed;
            L.call = call;
            L.tag = tag ? TAG_MH : TAG_ASYNC;
            L.out_by_vertex = 0;
            L.out_offset = off;
            L.propose_only = s ? 0 : 1;
            L.out_u = c.ev_u;
            L.huge_offsets = d_hoff;
            launch_eval(c, L);
        }
        off += counts[b];
    }
    std::vector<int> rs(n);
    std::vector<uint8_t> racc(n);
    std::vector<double> rdS(n), rH(n), ru(n);
    std::vector<float> rqf, rqb;
    if (learned && (qf || qb)) {
        rqf.assign(c.g.N, -1.0f);
        rqb.assign(c.g.N, -1.0f);
        SBPX_HIP(hipMemcpyAsync(rqf.data(), c.mlp.qf, sizeof(float) * c.g.N, hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipMemcpyAsync(rqb.data(), c.mlp.qb, sizeof(float) * c.g.N, hipMemcpyDeviceToHost, c.stream));
    }
    SBPX_HIP(hipMemcpyAsync(rs.data(), c.ev_s, sizeof(int) * n, hipMemcpyDeviceToHost, c.stream));
    SBPX_HIP(hipMemcpyAsync(racc.data(), c.ev_acc, n, hipMemcpyDeviceToHost, c.stream));
    SBPX_HIP(hipMemcpyAsync(rdS.data(), c.ev_dS, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SBPX_HIP(hipMemcpyAsync(rH.data(), c.ev_H, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    SBPX_HIP(hipMemcpyAsync(ru.data(), c.ev_u, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    check_device_error(c);
    for (int i = 0; i < n; ++i) {
        int o = perm[i];
        if (dS) dS[o] = rdS[i];
        if (H) H[o] = rH[i];
        if (acc) acc[o] = racc[i] == 2 ? 1 : 0;
        if (s_out) s_out[o] = rs[i];
        if (u_out) u_out[o] = ru[i];
        if (qf) qf[o] = i < nlow ? rqf[ordered[i]] : -1.0f;
        if (qb) qb[o] = i < nlow ? rqb[ordered[i]] : -1.0f;
    }
}

sbpx_status sbpx_eval_moves(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s, const double* u,
                            double* dS, double* H, uint8_t* accept) {
    return guarded(ctx, [&] {
        if (n > 0 && (!v || !s || !u)) throw ArgError("null proposal array");
        eval_fixed(ctx, n, v, s, u, dS, H, accept, 0, 0, 0, nullptr, nullptr);
    });
}

sbpx_status sbpx_learned_eval_moves(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s, const double* u,
                                    double* dS, double* H, uint8_t* accept, float* qf, float* qb) {
    return guarded(ctx, [&] {
        if (n > 0 && (!v || !s || !u)) throw ArgError("null proposal array");
        Ctx& c = ctx->c;
        if (!c.mlp.ready) throw StateError("no proposal network initialised (sbpx_learned_init)");
        require_bm(c);
        std::vector<char> seen(c.g.N, 0);
        for (int i = 0; i < n; ++i) {
            if (v[i] < 0 || v[i] >= c.g.N) throw ArgError("vertex out of range");
            if (seen[v[i]]) throw ArgError("vertex listed twice");
            seen[v[i]] = 1;
        }
        ensure_mlp_rows(c);
        eval_fixed(ctx, n, v, s, u, dS, H, accept, 0, 0, 0, nullptr, nullptr, true, qf, qb);
    });
}

sbpx_status sbpx_learned_scores(sbpx_ctx* ctx, int32_t n, const int32_t* v, int scorer, int32_t* cand, float* z) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        if (!c.mlp.ready) throw StateError("no proposal network initialised (sbpx_learned_init)");
        require_bm(c);
        if (n < 0 || n > c.g.N || n > MLP_BATCH) throw ArgError("count must be in [0, min(N, 8192)]");
        if (scorer != 0 && scorer != 1) throw ArgError("scorer must be 0 (batched) or 1 (sequential pass)");
        if (n > 0 && (!v || !cand || !z)) throw ArgError("null array");
        int max_deg = 0;
        for (int i = 0; i < n; ++i) {
            if (v[i] < 0 || v[i] >= c.g.N) throw ArgError("vertex out of range");
            max_deg = std::max(max_deg, c.g.degree[v[i]]);
        }
        if (max_deg > LEARN_MH_MAXDEG) throw ArgError("vertex degree above 510");
        if (n == 0) return;
        ensure_mlp_rows(c);
        int* d_list = c.list_buf;
        SBPX_HIP(hipMemcpyAsync(d_list, v, sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
        int* d_cand = c.mlp.cand;
        float* d_z = c.mlp.sz;
        if (scorer == 0) {
            learned_features(c, d_list, n, max_deg);
            mlp_forward(c, c.mlp.x, n, c.mlp.sh1, c.mlp.sh2, d_z);
        } else {
            learned_scores_probe(c, d_list, n, max_deg, d_cand, d_z);
        }
        SBPX_HIP(hipMemcpyAsync(cand, d_cand, sizeof(int) * n * MLP_CAND, hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipMemcpyAsync(z, d_z, sizeof(float) * n * MLP_OUT, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        check_device_error(c);
    });
}

sbpx_status sbpx_propose(sbpx_ctx* ctx, int32_t n, const int32_t* v, uint64_t seed, uint32_t call, int tag,
                         int32_t* s, double* u) {
    return guarded(ctx, [&] {
        if (n > 0 && !v) throw ArgError("null vertex array");
        eval_fixed(ctx, n, v, nullptr, nullptr, nullptr, nullptr, nullptr, seed, call, tag, s, u);
    });
}

sbpx_status sbpx_apply_moves(sbpx_ctx* ctx, int32_t n, const int32_t* v, const int32_t* s) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        if (n < 0) throw ArgError("negative count");
        if (n == 0) return;
        std::vector<char> seen(c.g.N, 0);
        for (int i = 0; i < n; ++i) {
            if (v[i] < 0 || v[i] >= c.g.N) throw ArgError("vertex out of range");
            if (s[i] < 0 || s[i] >= c.bm.B) throw ArgError("block id out of range [0, B)");
            if (seen[v[i]]) throw ArgError("vertex listed twice");
            seen[v[i]] = 1;
        }
        std::vector<int> a(c.g.N);
        SBPX_HIP(hipMemcpyAsync(a.data(), c.bm.assign, sizeof(int) * c.g.N, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        std::vector<uint8_t> stat(n);
        for (int i = 0; i < n; ++i) stat[i] = (a[v[i]] != s[i]) ? 2 : 0;
        // vertex-indexed outputs: scatter s and status through a small kernel-free path
        std::vector<int> sv(c.g.N, 0);
        std::vector<uint8_t> st(c.g.N, 0);
        for (int i = 0; i < n; ++i) {
            sv[v[i]] = s[i];
            st[v[i]] = stat[i];
        }
        SBPX_HIP(hipMemcpyAsync(c.ev_s, sv.data(), sizeof(int) * c.g.N, hipMemcpyHostToDevice, c.stream));
        SBPX_HIP(hipMemcpyAsync(c.ev_acc, st.data(), c.g.N, hipMemcpyHostToDevice, c.stream));
        SBPX_HIP(hipMemsetAsync(c.ev_dS, 0, sizeof(double) * c.g.N, c.stream));
        SBPX_HIP(hipMemcpyAsync(c.list_buf, v, sizeof(int) * n, hipMemcpyHostToDevice, c.stream));
        apply_accepted(c, c.list_buf, n);
        c.sync();
        check_device_error(c);
    });
}

sbpx_status sbpx_sweep(sbpx_ctx* ctx, int algo, int32_t batches, uint64_t seed, uint32_t call,
                       sbpx_sweep_stats* stats) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        if (batches < 1) throw ArgError("batches must be >= 1");
        if (algo != SBPX_ALGO_ASYNC_GIBBS && algo != SBPX_ALGO_HYBRID && algo != SBPX_ALGO_METROPOLIS_HASTINGS)
            throw ArgError("unknown algorithm");
        build_plan(c, algo, batches);
        SBPX_HIP(hipMemsetAsync(c.counters, 0, sizeof(long long) * 8, c.stream));
        SBPX_HIP(hipMemsetAsync(c.red + 2048, 0, sizeof(double), c.stream));
        c.timed.clear();
        c.timed_bins.clear();
        long long* bytes = ctx->count_bytes ? c.counters + 4 : nullptr;
        long long* mh_bytes = ctx->count_bytes ? c.counters + 4 + KERNEL_MH : nullptr;
        if (algo == SBPX_ALGO_METROPOLIS_HASTINGS)
            mh_sequential(c, c.mh.d_all, c.g.N, c.g.max_deg, seed, call, TAG_MH, mh_bytes, true);
        else if (algo == SBPX_ALGO_HYBRID)
            mh_sequential(c, c.mh.d_high, (int)c.g.high.size(), c.mh.max_deg_high, seed, call, TAG_MH, mh_bytes,
                          true);
        const bool learned = c.proposal >= 1;
        if (learned && !c.mlp.ready) throw StateError("learned proposal selected but no network initialised");
        if ((learned || c.mlp.logging) && !c.mlp.rows) ensure_mlp_rows(c);
        for (const BinnedList& bl : c.plan.batches_async) {
            const int* d_list = c.plan.d_lists + bl.offset[0];
            eval_vertex_list(ctx, c.plan.d_lists, bl, c.plan.d_huge_offsets + bl.huge_base, seed, call, TAG_ASYNC,
                             bytes, true, learned);
            if (c.mlp.logging && !learned) learned_log(c, d_list, bl.count[0] + bl.count[1] + bl.count[2]);
            apply_accepted(c, d_list, bl.total);
        }
        sweep_stats(ctx, stats);
    });
}

sbpx_status sbpx_merge_proposals(sbpx_ctx* ctx, uint64_t seed, uint32_t call, double* dS, int32_t* best) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        launch_merge_proposals(c, seed, call);
        if (dS) SBPX_HIP(hipMemcpyAsync(dS, c.m_dS, sizeof(double) * c.bm.B, hipMemcpyDeviceToHost, c.stream));
        if (best) SBPX_HIP(hipMemcpyAsync(best, c.m_best, sizeof(int) * c.bm.B, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        check_device_error(c);
    });
}

static int uf_root(std::vector<int>& parent, int x) {
    int root = x;
    while (parent[root] != root) root = parent[root];
    while (parent[x] != root) {
        int nx = parent[x];
        parent[x] = root;
        x = nx;
    }
    return root;
}

// Blockmodel::carry_out_best_merges (blockmodel.cpp:50-78) over blocks [0, B) of one graph:
// candidates by ascending dS (utils::sort_indices, utils.hpp:94-101: std::sort with the same
// comparator, so ties fall in the same order as the reference's). block_map always names the
// block a block has been merged into; a union-find with the merge target as root keeps the same
// mapping in near-linear time.
constexpr int MERGE_SELECT_THREADS = 8;  // host threads sharing a batched merge round's selections

static void select_merges_sorted(int B, int num_to_merge, const double* dS, const int32_t* best,
                                 std::vector<int>& parent) {
    std::vector<int> order(B);
    std::iota(order.begin(), order.end(), 0);
    std::vector<double> d(dS, dS + B);
    std::sort(order.data(), order.data() + order.size(), [&d](size_t i1, size_t i2) { return d[i1] < d[i2]; });
    int num_merged = 0;
    size_t counter = 0;
    while (num_merged < num_to_merge) {
        if (counter >= order.size()) throw StateError("ran out of merge candidates");
        int from = order[counter];
        int b = best[from];
        if (b < 0 || b >= B) throw StateError("block without a merge proposal");
        int to = uf_root(parent, b);
        counter++;
        if (to != from) {
            parent[from] = to;
            num_merged++;
        }
    }
}

// build_mapping (blockmodel.cpp:29-41): surviving labels of non-empty blocks, ascending, from
// `base` on; map[x] for the blocks [0, B) of one graph.
static void merge_mapping(int B, int num_to_merge, const int* count, std::vector<int>& parent, int base, int* map) {
    std::vector<char> used(B, 0);
    for (int x = 0; x < B; ++x)
        if (count[x] > 0) used[uf_root(parent, x)] = 1;
    std::vector<int> rank(B, -1);
    int next = 0;
    for (int x = 0; x < B; ++x)
        if (used[x]) rank[x] = next++;
    if (next > B - num_to_merge) throw StateError("more surviving labels than blocks after the merge");
    for (int x = 0; x < B; ++x) map[x] = base + rank[uf_root(parent, x)];
}

static void carry_out(Ctx& c, int num_to_merge, int mode, const double* dS, const int32_t* best) {
    const int B = c.bm.B;
    if (num_to_merge < 0 || num_to_merge >= B) throw ArgError("num_to_merge must be in [0, B)");
    if (mode != SBPX_MERGE_APPROXIMATE && mode != SBPX_MERGE_ADVANCED) throw ArgError("unknown merge mode");
    std::vector<int> parent(B);
    std::iota(parent.begin(), parent.end(), 0);
    if (mode == SBPX_MERGE_ADVANCED) {
        for (int x = 0; x < B; ++x)
            if (best[x] >= B) throw ArgError("proposed block out of range [0, B)");
        merge_heap_select(c, num_to_merge, dS, best, parent);
    } else {
        select_merges_sorted(B, num_to_merge, dS, best, parent);
    }
    launch_block_counts(c);
    std::vector<int> count(B);
    SBPX_HIP(hipMemcpyAsync(count.data(), c.m_count, sizeof(int) * B, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    std::vector<int> map(B);
    merge_mapping(B, num_to_merge, count.data(), parent, 0, map.data());
    SBPX_HIP(hipMemcpyAsync(c.m_map, map.data(), sizeof(int) * B, hipMemcpyHostToDevice, c.stream));
    launch_relabel(c);
    c.bm.B = B - num_to_merge;
    build_blockmodel(c);
    check_device_error(c);
}

sbpx_status sbpx_carry_out_merges(sbpx_ctx* ctx, int32_t num_to_merge, int mode, const double* dS,
                                  const int32_t* best) {
    return guarded(ctx, [&] {
        require_bm(ctx->c);
        if (!dS || !best) throw ArgError("null merge arrays");
        carry_out(ctx->c, num_to_merge, mode, dS, best);
    });
}

sbpx_status sbpx_merge_blocks(sbpx_ctx* ctx, int32_t num_to_merge, int mode, uint64_t seed, uint32_t call) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_bm(c);
        launch_merge_proposals(c, seed, call);
        c.h_dS.resize(c.bm.B);
        c.h_best.resize(c.bm.B);
        SBPX_HIP(hipMemcpyAsync(c.h_dS.data(), c.m_dS, sizeof(double) * c.bm.B, hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipMemcpyAsync(c.h_best.data(), c.m_best, sizeof(int) * c.bm.B, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        check_device_error(c);
        carry_out(c, num_to_merge, mode, c.h_dS.data(), c.h_best.data());
    });
}

sbpx_status sbpx_selftest_logf(sbpx_ctx* ctx, uint32_t lo, uint32_t n, float* out) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        if (!out) throw ArgError("null output");
        if (n == 0) return;
        float* d = nullptr;
        SBPX_HIP(hipMalloc((void**)&d, sizeof(float) * n));
        try {
            launch_logf_table(c, lo, n, d);
            SBPX_HIP(hipMemcpyAsync(out, d, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
            c.sync();
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        SBPX_HIP(hipFree(d));
    });
}

// ------------------------------------------------------------------------------------------
// Proposal network (no counterpart in the reference; see include/sbpx.h).
sbpx_status sbpx_learned_init(sbpx_ctx* ctx, uint64_t seed) {
    return guarded(ctx, [&] { init_mlp(ctx->c, seed); });
}

sbpx_status sbpx_learned_params(sbpx_ctx* ctx, float* out, const float* in) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        if (!c.mlp.ready) throw StateError("no network initialised");
        long off = 0;
        for (int l = 0; l < 3; ++l) {
            const long n = mlp_weight_count(l), nb = mlp_bias_count(l);
            if (out) {
                SBPX_HIP(hipMemcpy(out + off, c.mlp.w[l], sizeof(float) * n, hipMemcpyDeviceToHost));
                SBPX_HIP(hipMemcpy(out + off + n, c.mlp.b[l], sizeof(float) * nb, hipMemcpyDeviceToHost));
            }
            if (in) {
                SBPX_HIP(hipMemcpy(c.mlp.w[l], in + off, sizeof(float) * n, hipMemcpyHostToDevice));
                SBPX_HIP(hipMemcpy(c.mlp.b[l], in + off + n, sizeof(float) * nb, hipMemcpyHostToDevice));
            }
            off += n + nb;
        }
        if (in) {
            mlp_sync_bf16(c);
            c.sync();
        }
    });
}

sbpx_status sbpx_learned_log(sbpx_ctx* ctx, int32_t capacity) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        MlpState& m = c.mlp;
        if (!m.ready) throw StateError("no network initialised");
        if (capacity < 0) throw ArgError("capacity must be >= 0");
        if (capacity == 0) {
            m.logging = false;
            return;
        }
        if (capacity != m.ds_cap) {
            dalloc(&m.ds_x, (size_t)capacity * MLP_IN);
            dalloc(&m.ds_mask, (size_t)capacity);
            dalloc(&m.ds_y, (size_t)capacity);
            m.ds_cap = capacity;
        }
        SBPX_HIP(hipMemset(m.ds_count, 0, sizeof(int)));
        m.logging = true;
    });
}

sbpx_status sbpx_learned_log_count(sbpx_ctx* ctx, int32_t* count) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        if (!count) throw ArgError("null output");
        int n = 0;
        if (c.mlp.ready) {
            SBPX_HIP(hipMemcpyAsync(&n, c.mlp.ds_count, sizeof(int), hipMemcpyDeviceToHost, c.stream));
            c.sync();
        }
        *count = std::min(n, c.mlp.ds_cap);
    });
}

sbpx_status sbpx_learned_train(sbpx_ctx* ctx, int32_t steps, int32_t batch, float lr, uint64_t seed,
                               float* mean_loss) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        MlpState& m = c.mlp;
        if (!m.ready) throw StateError("no network initialised");
        if (batch <= 0 || batch > MLP_BATCH) throw ArgError("batch must be in [1, 8192]");
        int n = 0;
        SBPX_HIP(hipMemcpyAsync(&n, m.ds_count, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        n = std::min(n, m.ds_cap);
        if (n <= 0) throw StateError("no logged samples to train on");
        uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
        std::vector<int> idx(batch);
        float* grads[3] = {m.gw[0], m.gw[1], m.gw[2]};
        float* gbias[3] = {m.gb[0], m.gb[1], m.gb[2]};
        for (int t = 0; t < steps; ++t) {
            for (int i = 0; i < batch; ++i) {  // splitmix64 draws, uniform rows with replacement
                x += 0x9E3779B97F4A7C15ull;
                uint64_t z = x;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                idx[i] = (int)((unsigned __int128)z * (uint64_t)n >> 64);
            }
            SBPX_HIP(hipMemcpyAsync(m.tidx, idx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, c.stream));
            mlp_gather(c, m.tidx, batch);
            mlp_backward(c, m.tx, m.tmask, m.ty, batch, m.loss, grads, gbias);
            mlp_adam_step(c, lr);
            c.sync();  // idx is reused by the next step's copy
        }
        if (mean_loss) {
            std::vector<float> l(batch);
            SBPX_HIP(hipMemcpyAsync(l.data(), m.loss, sizeof(float) * batch, hipMemcpyDeviceToHost, c.stream));
            c.sync();
            double sum = 0.0;
            for (float v : l) sum += v;
            *mean_loss = (float)(sum / batch);
        }
    });
}

sbpx_status sbpx_learned_forward(sbpx_ctx* ctx, int32_t n, const float* x, float* z) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        MlpState& m = c.mlp;
        if (!m.ready) throw StateError("no network initialised");
        if (n <= 0 || n > MLP_BATCH) throw ArgError("n must be in [1, 8192]");
        if (!x || !z) throw ArgError("null array");
        float* tmp = nullptr;
        dalloc(&tmp, (size_t)n * MLP_IN);
        try {
            SBPX_HIP(hipMemcpy(tmp, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice));
            mlp_convert_to_bf16(c, tmp, (long)n * MLP_IN, m.tx);
            mlp_forward(c, m.tx, n, m.h1, m.h2, m.z);
            SBPX_HIP(hipMemcpyAsync(z, m.z, sizeof(float) * n * MLP_OUT, hipMemcpyDeviceToHost, c.stream));
            c.sync();
        } catch (...) {
            dfree(&tmp);
            throw;
        }
        dfree(&tmp);
    });
}

sbpx_status sbpx_learned_gradients(sbpx_ctx* ctx, int32_t n, const float* x, const uint32_t* mask, const int32_t* y,
                                   float* loss, float* grads) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        MlpState& m = c.mlp;
        if (!m.ready) throw StateError("no network initialised");
        if (n <= 0 || n > MLP_BATCH) throw ArgError("n must be in [1, 8192]");
        if (!x || !mask || !y || !grads) throw ArgError("null array");
        for (int i = 0; i < n; ++i)
            if (y[i] < 0 || y[i] >= MLP_CAND || !(mask[i] >> y[i] & 1u)) throw ArgError("label outside its mask");
        float* tmp = nullptr;
        dalloc(&tmp, (size_t)n * MLP_IN);
        try {
            SBPX_HIP(hipMemcpy(tmp, x, sizeof(float) * n * MLP_IN, hipMemcpyHostToDevice));
            mlp_convert_to_bf16(c, tmp, (long)n * MLP_IN, m.tx);
            SBPX_HIP(hipMemcpy(m.tmask, mask, sizeof(uint32_t) * n, hipMemcpyHostToDevice));
            SBPX_HIP(hipMemcpy(m.ty, y, sizeof(int) * n, hipMemcpyHostToDevice));
            float* gw[3] = {m.gw[0], m.gw[1], m.gw[2]};
            float* gb[3] = {m.gb[0], m.gb[1], m.gb[2]};
            mlp_backward(c, m.tx, m.tmask, m.ty, n, m.loss, gw, gb);
            long off = 0;
            for (int l = 0; l < 3; ++l) {
                const long nw = mlp_weight_count(l), nb = mlp_bias_count(l);
                SBPX_HIP(hipMemcpyAsync(grads + off, m.gw[l], sizeof(float) * nw, hipMemcpyDeviceToHost, c.stream));
                SBPX_HIP(hipMemcpyAsync(grads + off + nw, m.gb[l], sizeof(float) * nb, hipMemcpyDeviceToHost, c.stream));
                off += nw + nb;
            }
            if (loss) SBPX_HIP(hipMemcpyAsync(loss, m.loss, sizeof(float) * n, hipMemcpyDeviceToHost, c.stream));
            c.sync();
        } catch (...) {
            dfree(&tmp);
            throw;
        }
        dfree(&tmp);
    });
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// A batch of independent graphs held as one disjoint union on the device (DevBatch): every
// operation takes the graphs it applies to, runs over the union in one set of launches, and
// leaves the other graphs' state as it was.
namespace {

// Neighbour-block table size the batched evaluation gives a vertex of this degree.
int eval_table_cap(int degree) {
    const int b = bin_of(degree);
    if (b < NUM_LDS_BINS) return BIN_CAP[b];
    if (b == NUM_LDS_BINS) return 4096;
    return table_cap(degree + 1);
}

// Marks the listed graphs active (each at most once) and uploads their seeds and calls.
void batch_select(Ctx& c, int n, const int32_t* graphs, const uint64_t* seeds, const uint32_t* calls,
                  std::vector<char>& active) {
    DevBatch& b = c.batch;
    if (n < 0 || n > b.G) throw ArgError("number of graphs out of range");
    if (n > 0 && !graphs) throw ArgError("null graph list");
    active.assign(b.G, 0);
    for (int i = 0; i < n; ++i) {
        if (graphs[i] < 0 || graphs[i] >= b.G) throw ArgError("graph index out of range");
        if (active[graphs[i]]) throw ArgError("graph listed twice");
        active[graphs[i]] = 1;
    }
    b.h_active.assign(active.begin(), active.end());
    SBPX_HIP(hipMemcpyAsync(b.d_active, b.h_active.data(), b.G, hipMemcpyHostToDevice, c.stream));
    if (seeds && calls) {
        for (int i = 0; i < n; ++i) {
            b.h_seed[graphs[i]] = seeds[i];
            b.h_call[graphs[i]] = calls[i];
        }
        SBPX_HIP(hipMemcpyAsync(b.d_seed, b.h_seed.data(), sizeof(uint64_t) * b.G, hipMemcpyHostToDevice, c.stream));
        SBPX_HIP(hipMemcpyAsync(b.d_call, b.h_call.data(), sizeof(uint32_t) * b.G, hipMemcpyHostToDevice, c.stream));
    }
}

// Uploads the blocks in use per graph and renumbers the blocks in use densely (the rebuild's
// sort keys).
void batch_upload_nb(Ctx& c) {
    DevBatch& b = c.batch;
    b.h_dbo.resize(b.G);
    int used = 0;
    for (int g = 0; g < b.G; ++g) {
        b.h_dbo[g] = used;
        used += b.nb[g];
    }
    b.used = used;
    SBPX_HIP(hipMemcpyAsync(b.d_nb, b.nb.data(), sizeof(int) * b.G, hipMemcpyHostToDevice, c.stream));
    SBPX_HIP(hipMemcpyAsync(b.d_dbo, b.h_dbo.data(), sizeof(int) * b.G, hipMemcpyHostToDevice, c.stream));
    launch_batch_dense(c);
}

// The reference's batches of every graph (finetune.cpp:43-47, 352-366, as in build_plan), split
// by degree bin.
void batch_lists(Ctx& c, int batches) {
    DevBatch& b = c.batch;
    if (b.lists_batches == batches) return;
    b.lists.assign(b.G, {});
    for (int g = 0; g < b.G; ++g) {
        const int n = b.vo[g + 1] - b.vo[g];
        double num_batches = batches;
        int batch_size = int(ceil(n / num_batches));
        int nbat = n / batch_size;
        for (int k = 0; k < nbat; ++k) {
            std::array<std::vector<int>, NUM_BINS> bins;
            for (int v = k * batch_size; v < std::min(n, (k + 1) * batch_size); ++v) {
                const int u = b.vo[g] + v;
                bins[bin_of(c.g.degree[u])].push_back(u);
            }
            b.lists[g].push_back(std::move(bins));
        }
    }
    b.lists_batches = batches;
    b.plan_active.clear();
    // bounds over all graphs, and plan buffers that hold any subset of them
    size_t K = 0;
    for (int g = 0; g < b.G; ++g) K = std::max(K, b.lists[g].size());
    b.bound_bin.assign(K * NUM_BINS, 0);
    b.bound_total.assign(K, 0);
    b.bound_pool = 0;
    size_t slots = 0, huge_slots = 0;
    for (size_t k = 0; k < K; ++k) {
        size_t pool = 0;
        for (int g = 0; g < b.G; ++g) {
            if (k >= b.lists[g].size()) continue;
            for (int bin = 0; bin < NUM_BINS; ++bin) {
                b.bound_bin[k * NUM_BINS + bin] += (int)b.lists[g][k][bin].size();
                b.bound_total[k] += (int)b.lists[g][k][bin].size();
            }
            for (int u : b.lists[g][k][NUM_BINS - 1]) pool += 7 * (size_t)table_cap(c.g.degree[u] + 1);
        }
        b.bound_pool = std::max(b.bound_pool, pool);
        slots += b.bound_total[k];
        huge_slots += b.bound_bin[k * NUM_BINS + NUM_BINS - 1] + 1;
    }
    if (slots > b.plan_slots) {
        dalloc(&b.d_plan, slots);
        b.plan_slots = slots;
    }
    if (huge_slots > b.plan_huge_slots) {
        dalloc(&b.d_plan_huge, huge_slots);
        b.plan_huge_slots = huge_slots;
    }
    ensure_pool(c, b.bound_pool);
}

// Union batch k: batch k of every active graph that has one, binned; kept while the same graphs
// sweep with the same number of batches.
void batch_plan(Ctx& c, const std::vector<char>& active, int batches) {
    DevBatch& b = c.batch;
    if (b.plan_active == active && b.plan_batches == batches) return;
    size_t K = 0;
    for (int g = 0; g < b.G; ++g)
        if (active[g]) K = std::max(K, b.lists[g].size());
    std::vector<int> all, huge_all;
    size_t pool_need = 0;
    b.plan.clear();
    for (size_t k = 0; k < K; ++k) {
        BinnedList bl;
        const int first = (int)all.size();
        std::vector<int> hoff(1, 0);
        for (int bin = 0; bin < NUM_BINS; ++bin) {
            bl.offset[bin] = (int)all.size();
            for (int g = 0; g < b.G; ++g) {
                if (!active[g] || k >= b.lists[g].size()) continue;
                for (int u : b.lists[g][k][bin]) {
                    all.push_back(u);
                    if (bin == NUM_BINS - 1) hoff.push_back(hoff.back() + 7 * table_cap(c.g.degree[u] + 1));
                }
            }
            bl.count[bin] = (int)all.size() - bl.offset[bin];
        }
        bl.total = (int)all.size() - first;
        bl.huge_base = (int)huge_all.size();
        huge_all.insert(huge_all.end(), hoff.begin(), hoff.end());
        pool_need = std::max(pool_need, (size_t)hoff.back());
        b.plan.push_back(bl);
    }
    if (all.size() > b.plan_slots) {
        dalloc(&b.d_plan, all.size());
        b.plan_slots = all.size();
    }
    if (huge_all.size() > b.plan_huge_slots) {
        dalloc(&b.d_plan_huge, huge_all.size());
        b.plan_huge_slots = huge_all.size();
    }
    if (!all.empty())
        SBPX_HIP(hipMemcpy(b.d_plan, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice));
    if (!huge_all.empty())
        SBPX_HIP(hipMemcpy(b.d_plan_huge, huge_all.data(), sizeof(int) * huge_all.size(), hipMemcpyHostToDevice));
    ensure_pool(c, pool_need);
    b.plan_active = active;
    b.plan_batches = batches;
}

// The shape a captured round is valid for: plan batches, number of batches, key width, kernel
// variants, the addresses of the buffers its launches name, and the update rule.
std::vector<long long> round_key(const Ctx& c, int K, int batches, const int* variants) {
    const DevBatch& b = c.batch;
    std::vector<long long> key = {K, batches, sort_key_bits(c),
                                  (long long)(intptr_t)c.table_pool, (long long)(intptr_t)b.d_plan,
                                  (long long)(intptr_t)b.d_plan_huge, (long long)(intptr_t)b.d_round,
                                  c.incremental, c.g.E, c.bm.B};
    for (int bin = 0; bin < NUM_BINS; ++bin) key.push_back(variants[bin]);
    return key;
}

// Captures one round of the plan's K batches: per batch, the evaluation of every bin with vertices
// in some graph (persistent launches over the bin's list, counts in the round words), then the
// application and the block-matrix update with device-held sizes.
void capture_round(Ctx& c, int K, const int* variants) {
    DevBatch& b = c.batch;
    if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
    b.gexec = nullptr;
    b.gkey.clear();
    ensure_update_buffers(c);
    const size_t need = (size_t)K * ROUND_WORDS;
    if (b.round_slots < need) {
        dfree(&b.d_round);
        dfree(&b.d_queue);
        if (b.h_round) (void)hipHostFree(b.h_round);
        b.h_round = nullptr;
        b.round_slots = 0;
        dalloc(&b.d_round, need);
        dalloc(&b.d_queue, (size_t)K * NUM_BINS);
        SBPX_HIP(hipHostMalloc((void**)&b.h_round, sizeof(int) * need));
        b.round_slots = need;
    }
    hipGraph_t graph = nullptr;
    SBPX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
        launch_round_zero(c, b.d_queue, K * NUM_BINS);
        for (int k = 0; k < K; ++k) {
            const int* w = b.d_round + (size_t)k * ROUND_WORDS;
            for (int bin = 0; bin < NUM_BINS; ++bin) {
                const int bound = b.bound_bin[(size_t)k * NUM_BINS + bin];
                if (bound == 0) continue;
                EvalLaunch L{};
                L.list = b.d_plan;
                L.list_off = w + ROUND_OFF + bin;
                L.n = bound;
                L.n_at = w + ROUND_CNT + bin;
                L.bin = bin;
                L.variant = variants[bin];
                L.tag = TAG_ASYNC;
                L.out_by_vertex = 1;
                L.huge_offsets = b.d_plan_huge;
                L.huge_off_at = w + ROUND_HUGE;
                L.queue = b.d_queue + (size_t)k * NUM_BINS + bin;
                launch_eval(c, L);
            }
            apply_accepted_dyn(c, VList(b.d_plan, w + ROUND_FIRST, Dyn(b.bound_total[k], w + ROUND_TOTAL)),
                               Dyn(0, w + ROUND_EPOCH));
        }
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    SBPX_HIP(hipStreamEndCapture(c.stream, &graph));
    const hipError_t e = hipGraphInstantiate(&b.gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SBPX_HIP(e);
}

// One round replayed from the captured graph: the round words of the plan's batches (list
// offsets and counts per bin, huge-table bases, epochs) staged and copied in, the graph launched,
// the cell count read back with the counters.
void replay_round(sbpx_ctx* ctx, sbpx_sweep_stats* stats) {
    Ctx& c = ctx->c;
    DevBatch& b = c.batch;
    const int K = (int)b.plan.size();
    for (int k = 0; k < K; ++k) {
        const BinnedList& bl = b.plan[k];
        int* w = b.h_round + (size_t)k * ROUND_WORDS;
        for (int bin = 0; bin < NUM_BINS; ++bin) {
            if (bl.count[bin] > b.bound_bin[(size_t)k * NUM_BINS + bin]) throw StateError("plan exceeds its bound");
            w[ROUND_OFF + bin] = bl.offset[bin];
            w[ROUND_CNT + bin] = bl.count[bin];
        }
        w[ROUND_HUGE] = bl.huge_base;
        w[ROUND_FIRST] = bl.offset[0];
        w[ROUND_TOTAL] = bl.total;
        w[ROUND_EPOCH] = c.epoch + 1 + k;
    }
    SBPX_HIP(hipMemcpyAsync(b.d_round, b.h_round, sizeof(int) * (size_t)K * ROUND_WORDS, hipMemcpyHostToDevice,
                            c.stream));
    SBPX_HIP(hipGraphLaunch(b.gexec, c.stream));
    c.epoch += K;
    SBPX_HIP(hipMemcpyAsync(&b.h_nnz, c.words + W_NNZ, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    c.timed.clear();
    c.timed_bins.clear();
    sweep_stats(ctx, stats);  // waits for the stream
    c.bm.nnz = b.h_nnz;
    b.graph_rounds++;
}

}  // namespace

extern "C" {

sbpx_status sbpx_batch_define(sbpx_ctx* ctx, int32_t num_graphs, const int32_t* vertex_offsets) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_graph(c);
        const int N = c.g.N, G = num_graphs;
        if (G < 1 || G > N) throw ArgError("number of graphs must be in [1, N]");
        if (!vertex_offsets) throw ArgError("null vertex offsets");
        if (vertex_offsets[0] != 0 || vertex_offsets[G] != N) throw ArgError("vertex offsets must run from 0 to N");
        for (int g = 0; g < G; ++g)
            if (vertex_offsets[g + 1] <= vertex_offsets[g]) throw ArgError("every graph needs a vertex");
        std::vector<int> vgraph(N);
        for (int g = 0; g < G; ++g)
            for (int v = vertex_offsets[g]; v < vertex_offsets[g + 1]; ++v) vgraph[v] = g;
        long align = 64;
        for (int v = 0; v < N; ++v) {
            for (int j = c.g.h_out_ptr[v]; j < c.g.h_out_ptr[v + 1]; ++j)
                if (vgraph[c.g.h_out_idx[j]] != vgraph[v]) throw ArgError("an edge joins two graphs of the batch");
            align = std::max<long>(align, eval_table_cap(c.g.degree[v]));
        }
        std::vector<int> bo(G + 1, 0);
        long top = 0;
        for (int g = 0; g < G; ++g) {
            bo[g] = (int)top;
            top = (top + (vertex_offsets[g + 1] - vertex_offsets[g]) + align - 1) / align * align;
            if (top > (1L << 30)) throw ArgError("the batch needs more than 2^30 block ids");
        }
        bo[G] = (int)top;
        SBPX_HIP(hipSetDevice(c.device));
        SBPX_HIP(hipStreamSynchronize(c.stream));
        free_batch(c);
        DevBatch& b = c.batch;
        b.G = G;
        b.align = (int)align;
        b.range = bo[G];
        b.vo.assign(vertex_offsets, vertex_offsets + G + 1);
        b.bo = bo;
        b.nb.resize(G);
        for (int g = 0; g < G; ++g) b.nb[g] = b.vo[g + 1] - b.vo[g];
        b.h_seed.assign(G, 0);
        b.h_call.assign(G, 0);
        // block-indexed buffers over the union's block ids
        const size_t NBK = (size_t)std::max(N, b.range) + 1;
        DevBlockmodel& m = c.bm;
        for (int** p : {&m.row_ptr, &m.col_ptr, &m.dout, &m.din, &m.deg, &m.diag, &m.drift, &c.m_best, &c.m_map,
                        &c.m_count, &c.tmp_b})
            dalloc(p, NBK);
        dalloc(&c.m_dS, NBK);
        dalloc(&c.last_touch, NBK);
        const long E = c.g.E;
        dalloc(&c.scan_tmp, scan_tmp_ints(std::max<long>(E, 256L * ((E + 4095) / 4096) + 1) + NBK));
        SBPX_HIP(hipMemset(m.drift, 0, sizeof(int) * NBK));
        SBPX_HIP(hipMemset(c.last_touch, 0, sizeof(unsigned long long) * NBK));
        dalloc(&b.d_vgraph, N);
        dalloc(&b.d_bgraph, b.range);
        dalloc(&b.d_vo, G + 1);
        dalloc(&b.d_bo, G + 1);
        dalloc(&b.d_nb, G);
        dalloc(&b.d_seed, G);
        dalloc(&b.d_call, G);
        dalloc(&b.d_active, G);
        dalloc(&b.d_req, 2 * (size_t)G + 2);
        dalloc(&b.d_dbo, G);
        dalloc(&b.d_dense, b.range);
        dalloc(&b.d_inv, b.range);
        std::vector<int> bgraph(b.range), assign(N);
        for (int g = 0; g < G; ++g) {
            for (int x = bo[g]; x < bo[g + 1]; ++x) bgraph[x] = g;
            for (int v = b.vo[g]; v < b.vo[g + 1]; ++v) assign[v] = bo[g] + (v - b.vo[g]);
        }
        SBPX_HIP(hipMemcpy(b.d_vgraph, vgraph.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(b.d_bgraph, bgraph.data(), sizeof(int) * b.range, hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(b.d_vo, b.vo.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice));
        SBPX_HIP(hipMemcpy(b.d_bo, b.bo.data(), sizeof(int) * (G + 1), hipMemcpyHostToDevice));
        batch_upload_nb(c);
        SBPX_HIP(hipMemset(b.d_seed, 0, sizeof(uint64_t) * G));
        SBPX_HIP(hipMemset(b.d_call, 0, sizeof(uint32_t) * G));
        SBPX_HIP(hipMemset(b.d_active, 0, G));
        SBPX_HIP(hipMemcpy(m.assign, assign.data(), sizeof(int) * N, hipMemcpyHostToDevice));
        c.plan = SweepPlan();
        m.B = b.range;
        build_blockmodel(c);
        check_device_error(c);
    });
}

sbpx_status sbpx_batch_info(sbpx_ctx* ctx, int32_t* num_graphs, int32_t* num_blocks) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        if (num_graphs) *num_graphs = c.batch.G;
        if (num_blocks) std::copy(c.batch.nb.begin(), c.batch.nb.end(), num_blocks);
    });
}

sbpx_status sbpx_batch_build(sbpx_ctx* ctx, int32_t n, const int32_t* graphs, const int32_t* num_blocks,
                             const int32_t* assignment) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        DevBatch& b = c.batch;
        std::vector<char> active;
        batch_select(c, n, graphs, nullptr, nullptr, active);
        if (n == 0) return;
        if (!num_blocks || !assignment) throw ArgError("null block counts or assignment");
        std::vector<int> req(2 * (size_t)n + 1);
        req[0] = 0;
        for (int i = 0; i < n; ++i) {
            const int g = graphs[i], Ng = b.vo[g + 1] - b.vo[g], B = num_blocks[i];
            if (B < 1 || B > Ng) throw ArgError("num_blocks must be in [1, N] of its graph");
            const int32_t* a = assignment + req[i];
            for (int v = 0; v < Ng; ++v)
                if (a[v] < 0 || a[v] >= B) throw ArgError("block id out of range [0, B)");
            req[i + 1] = req[i] + Ng;
            req[n + 1 + i] = g;
        }
        const int total = req[n];
        SBPX_HIP(hipMemcpyAsync(b.d_req, req.data(), sizeof(int) * req.size(), hipMemcpyHostToDevice, c.stream));
        SBPX_HIP(hipMemcpyAsync(c.list_buf, assignment, sizeof(int) * total, hipMemcpyHostToDevice, c.stream));
        launch_batch_scatter(c, n, total);
        for (int i = 0; i < n; ++i) b.nb[graphs[i]] = num_blocks[i];
        batch_upload_nb(c);
        launch_zero_drift_active(c);
        rebuild_keep_drift(c);
        check_device_error(c);
    });
}

sbpx_status sbpx_batch_download(sbpx_ctx* ctx, int32_t n, const int32_t* graphs, int32_t* assignment) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        const DevBatch& b = c.batch;
        if (n < 0 || n > b.G) throw ArgError("number of graphs out of range");
        if (n > 0 && (!graphs || !assignment)) throw ArgError("null graph list or assignment");
        std::vector<size_t> at(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            if (graphs[i] < 0 || graphs[i] >= b.G) throw ArgError("graph index out of range");
            at[i + 1] = at[i] + (b.vo[graphs[i] + 1] - b.vo[graphs[i]]);
        }
        for (int i = 0; i < n; ++i) {
            const int g = graphs[i];
            SBPX_HIP(hipMemcpyAsync(assignment + at[i], c.bm.assign + b.vo[g], sizeof(int) * (at[i + 1] - at[i]),
                                    hipMemcpyDeviceToHost, c.stream));
        }
        c.sync();
        for (int i = 0; i < n; ++i)
            for (size_t j = at[i]; j < at[i + 1]; ++j) assignment[j] -= b.bo[graphs[i]];
    });
}

sbpx_status sbpx_batch_trace(sbpx_ctx* ctx, int64_t* trace) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        if (!trace) throw ArgError("null output");
        const DevBatch& b = c.batch;
        long long* d_trace = (long long*)c.m_dS;  // merge scratch, free between merge phases
        launch_batch_trace(c, d_trace);
        SBPX_HIP(hipMemcpyAsync(trace, d_trace, sizeof(int64_t) * b.G, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

sbpx_status sbpx_batch_log_posterior(sbpx_ctx* ctx, double* out) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        if (!out) throw ArgError("null output");
        log_posterior_segments(c, out);
    });
}

sbpx_status sbpx_batch_sweep(sbpx_ctx* ctx, int algo, int32_t batches, int32_t n, const int32_t* graphs,
                             const uint64_t* seeds, const uint32_t* calls, sbpx_sweep_stats* stats) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        if (algo != SBPX_ALGO_ASYNC_GIBBS) throw ArgError("a batch of graphs runs asynchronous Gibbs sweeps");
        if (batches < 1) throw ArgError("batches must be >= 1");
        if (c.proposal != 0 || c.mlp.logging) throw StateError("the learned proposal runs on one graph at a time");
        if (n > 0 && (!seeds || !calls)) throw ArgError("null seeds or calls");
        std::vector<char> active;
        batch_select(c, n, graphs, seeds, calls, active);
        batch_lists(c, batches);
        batch_plan(c, active, batches);
        DevBatch& bw = c.batch;
        const int K = (int)bw.plan.size();
        int present[NUM_BINS] = {};
        for (const BinnedList& bl : bw.plan)
            for (int bin = 0; bin < NUM_BINS; ++bin) present[bin] += bl.count[bin];
        if (c.batch_graph && !ctx->count_bytes && K > 0 && c.g.E > 0 && c.bm.valid && c.variants_settled(present)) {
            int variants[NUM_BINS];
            for (int bin = 0; bin < NUM_BINS; ++bin) variants[bin] = c.settled_variant(bin);
            if (!bw.gexec || bw.gkey != round_key(c, K, batches, variants) || bw.round_slots < (size_t)K * ROUND_WORDS) {
                capture_round(c, K, variants);
                bw.gkey = round_key(c, K, batches, variants);
                bw.graph_captures++;
            }
            replay_round(ctx, stats);
            return;
        }
        SBPX_HIP(hipMemsetAsync(c.counters, 0, sizeof(long long) * 8, c.stream));
        SBPX_HIP(hipMemsetAsync(c.red + 2048, 0, sizeof(double), c.stream));
        c.timed.clear();
        c.timed_bins.clear();
        long long* bytes = ctx->count_bytes ? c.counters + 4 : nullptr;
        const DevBatch& b = c.batch;
        for (const BinnedList& bl : b.plan) {
            eval_vertex_list(ctx, b.d_plan, bl, b.d_plan_huge + bl.huge_base, 0, 0, TAG_ASYNC, bytes, true, false);
            apply_accepted(c, b.d_plan + bl.offset[0], bl.total);
        }
        sweep_stats(ctx, stats);
    });
}

sbpx_status sbpx_batch_replays(sbpx_ctx* ctx, int64_t* rounds, int64_t* captures) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        if (rounds) *rounds = c.batch.graph_rounds;
        if (captures) *captures = c.batch.graph_captures;
    });
}

sbpx_status sbpx_batch_merge(sbpx_ctx* ctx, int32_t n, const int32_t* graphs, const int32_t* num_to_merge, int mode,
                             const uint64_t* seeds, const uint32_t* calls) {
    return guarded(ctx, [&] {
        Ctx& c = ctx->c;
        require_batch(c);
        if (mode != SBPX_MERGE_APPROXIMATE) throw ArgError("a batch of graphs merges in the approximate mode");
        if (n > 0 && (!num_to_merge || !seeds || !calls)) throw ArgError("null merge counts, seeds or calls");
        DevBatch& b = c.batch;
        std::vector<char> active;
        batch_select(c, n, graphs, seeds, calls, active);
        if (n == 0) return;
        int lo = INT32_MAX, hi = 0;
        for (int i = 0; i < n; ++i) {
            const int g = graphs[i];
            if (num_to_merge[i] < 0 || num_to_merge[i] >= b.nb[g]) throw ArgError("num_to_merge must be in [0, B)");
            lo = std::min(lo, b.bo[g]);
            hi = std::max(hi, b.bo[g] + b.nb[g]);
        }
        launch_merge_proposals(c, 0, 0);
        launch_block_counts(c);
        const int span = hi - lo;
        if (!b.h_dS) {
            SBPX_HIP(hipHostMalloc((void**)&b.h_dS, sizeof(double) * b.range));
            for (int** p : {&b.h_best, &b.h_count, &b.h_map}) SBPX_HIP(hipHostMalloc((void**)p, sizeof(int) * b.range));
        }
        double* dS = b.h_dS;
        int *best = b.h_best, *count = b.h_count, *map = b.h_map;
        SBPX_HIP(hipMemcpyAsync(dS, c.m_dS + lo, sizeof(double) * span, hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipMemcpyAsync(best, c.m_best + lo, sizeof(int) * span, hipMemcpyDeviceToHost, c.stream));
        SBPX_HIP(hipMemcpyAsync(count, c.m_count + lo, sizeof(int) * span, hipMemcpyDeviceToHost, c.stream));
        std::iota(map, map + span, lo);  // blocks of graphs not listed keep their ids
        c.sync();
        check_device_error(c);
        // the graphs' selections are independent (each reads and writes its own slice): a few
        // host threads share them
        auto select = [&](int i) {
            const int g = graphs[i], B = b.nb[g], o = b.bo[g] - lo;
            std::vector<int> best_local(B), parent(B);
            for (int x = 0; x < B; ++x) best_local[x] = best[o + x] >= 0 ? best[o + x] - b.bo[g] : -1;
            std::iota(parent.begin(), parent.end(), 0);
            select_merges_sorted(B, num_to_merge[i], dS + o, best_local.data(), parent);
            merge_mapping(B, num_to_merge[i], count + o, parent, b.bo[g], map + o);
        };
        const int nt = std::min(n, MERGE_SELECT_THREADS);
        if (nt <= 1) {
            for (int i = 0; i < n; ++i) select(i);
        } else {
            std::atomic<int> next(0);
            std::mutex err_mu;
            std::exception_ptr err;
            std::vector<std::thread> pool;
            for (int t = 0; t < nt; ++t)
                pool.emplace_back([&] {
                    for (int i = next++; i < n; i = next++) {
                        try {
                            select(i);
                        } catch (...) {
                            std::lock_guard<std::mutex> g(err_mu);
                            if (!err) err = std::current_exception();
                        }
                    }
                });
            for (std::thread& t : pool) t.join();
            if (err) std::rethrow_exception(err);
        }
        SBPX_HIP(hipMemcpyAsync(c.m_map, map, sizeof(int) * span, hipMemcpyHostToDevice, c.stream));
        launch_batch_relabel(c, lo, hi);
        for (int i = 0; i < n; ++i) b.nb[graphs[i]] -= num_to_merge[i];
        batch_upload_nb(c);
        launch_zero_drift_active(c);
        rebuild_keep_drift(c);
        check_device_error(c);
    });
}

}  // extern "C"
