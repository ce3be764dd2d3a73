// Many graphs partitioned side by side on one GPU. Each graph runs the reference's driver
// unchanged (sbp::stochastic_block_partition) in a thread of its own; its device steps go to a
// coordinator instead of a context of its own. When every running driver waits on a step, the
// last to arrive executes the round: one batched call (sbpx_batch_*) per kind of step over the
// shared context, which holds all graphs as one disjoint union. Each graph therefore follows the
// same sequence of steps, with the same seeds and calls, as it would alone, while a round costs
// one set of launches for all graphs.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>

#include "device.hpp"
#include "sbp_host.hpp"

namespace sbpx_host {

namespace {

struct Request {
    enum Kind { BUILD, DOWNLOAD, LOGP, SWEEP, MERGE, TRACE } kind;
    int slot = -1;
    int num_blocks = 0;
    const int32_t* assignment = nullptr;
    int32_t* out_assignment = nullptr;
    int algo = 0, batches = 0, num_to_merge = 0, mode = 0;
    uint64_t seed = 0;
    uint32_t call = 0;
    sbpx_sweep_stats* stats = nullptr;
    double value = 0.0;
    int blocks_left = 0;
    std::string error;
    // set when the round that ran the request is over; each driver waits on its own request, so a
    // round wakes its drivers one by one instead of all of them on one condition variable
    bool done = false;
    std::mutex wake_mu;
    std::condition_variable wake;
};

}  // namespace

struct Batch {
    sbpx_ctx* ctx = nullptr;
    std::vector<int> vertex_offsets;
    std::mutex mu;
    int live = 0;  // drivers still running
    std::vector<Request*> pending;
    long rounds = 0;
    sbpx_sweep_stats totals{};  // sweep statistics over all graphs
    long long sweeps = 0;
    // log posterior of every graph after the last round that changed a state; a graph's entry
    // stays valid until that graph asks for a step that changes its state (no other graph's step
    // touches it), so its MDL queries need no round of their own
    std::vector<double> lp;
    std::vector<char> lp_valid;
    std::vector<int32_t> assign;
    // wall time of the rounds by request kind (the log posterior refresh under LOGP) and between
    // rounds, reported on stderr when SBP_TRACE is set
    double kind_s[6] = {0, 0, 0, 0, 0, 0};
    double between_s = 0.0;
    std::chrono::steady_clock::time_point last_end{};
};

namespace {

void fail(sbpx_ctx* ctx, const std::vector<Request*>& rs, const char* what) {
    const std::string msg = std::string(what) + ": " + sbpx_last_error(ctx);
    for (Request* r : rs) r->error = msg;
}

// Executes one round: every pending request, one batched call per kind (and per parameters).
// Each graph has at most one request in a round and every call changes only the state of the
// graphs it lists, so the order of the calls does not matter.
void execute(Batch& b, std::vector<Request*>& reqs) {
    std::map<int, std::vector<Request*>> by_kind;
    for (Request* r : reqs) by_kind[r->kind].push_back(r);
    sbpx_ctx* ctx = b.ctx;
    // steps that change states first, then the log posteriors of all graphs, then the reads
    for (int kind : {Request::BUILD, Request::MERGE, Request::SWEEP, Request::LOGP, Request::DOWNLOAD,
                     Request::TRACE}) {
        const auto t_kind = std::chrono::steady_clock::now();
        struct KindTimer {
            double& acc;
            std::chrono::steady_clock::time_point t;
            ~KindTimer() { acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count(); }
        } kind_timer{b.kind_s[kind], t_kind};
        if (kind == Request::LOGP) {
            const bool changed = by_kind.count(Request::BUILD) || by_kind.count(Request::MERGE) ||
                                 by_kind.count(Request::SWEEP);
            if (changed || by_kind.count(Request::LOGP)) {
                b.lp.resize(b.vertex_offsets.size() - 1);
                const bool ok = sbpx_batch_log_posterior(ctx, b.lp.data()) == SBPX_OK;
                b.lp_valid.assign(b.lp.size(), ok ? 1 : 0);
                if (!ok && by_kind.count(Request::LOGP)) fail(ctx, by_kind[Request::LOGP], "sbpx_batch_log_posterior");
            }
        }
        if (!by_kind.count(kind)) continue;
        std::vector<Request*>& rs = by_kind[kind];
        const int n = (int)rs.size();
        std::vector<int32_t> graphs(n);
        for (int i = 0; i < n; ++i) graphs[i] = rs[i]->slot;
        switch (kind) {
            case Request::BUILD: {
                std::vector<int32_t> blocks(n), concat;
                for (int i = 0; i < n; ++i) {
                    const int g = rs[i]->slot, Ng = b.vertex_offsets[g + 1] - b.vertex_offsets[g];
                    blocks[i] = rs[i]->num_blocks;
                    concat.insert(concat.end(), rs[i]->assignment, rs[i]->assignment + Ng);
                }
                if (sbpx_batch_build(ctx, n, graphs.data(), blocks.data(), concat.data()) != SBPX_OK)
                    fail(ctx, rs, "sbpx_batch_build");
                break;
            }
            case Request::DOWNLOAD: {
                size_t total = 0;
                for (Request* r : rs) total += b.vertex_offsets[r->slot + 1] - b.vertex_offsets[r->slot];
                b.assign.resize(total);
                if (sbpx_batch_download(ctx, n, graphs.data(), b.assign.data()) != SBPX_OK) {
                    fail(ctx, rs, "sbpx_batch_download");
                    break;
                }
                size_t at = 0;
                for (Request* r : rs) {
                    const size_t len = b.vertex_offsets[r->slot + 1] - b.vertex_offsets[r->slot];
                    std::copy(b.assign.begin() + at, b.assign.begin() + at + len, r->out_assignment);
                    at += len;
                }
                break;
            }
            case Request::LOGP: {
                for (Request* r : rs)
                    if (r->error.empty()) r->value = b.lp[r->slot];
                break;
            }
            case Request::TRACE: {
                std::vector<int64_t> tr(b.vertex_offsets.size() - 1);
                if (sbpx_batch_trace(ctx, tr.data()) != SBPX_OK) {
                    fail(ctx, rs, "sbpx_batch_trace");
                    break;
                }
                for (Request* r : rs) r->value = (double)tr[r->slot];
                break;
            }
            case Request::SWEEP: {
                std::map<std::pair<int, int>, std::vector<Request*>> groups;
                for (Request* r : rs) groups[{r->algo, r->batches}].push_back(r);
                for (auto& gv : groups) {
                    std::vector<Request*>& q = gv.second;
                    const int m = (int)q.size();
                    std::vector<int32_t> gs(m);
                    std::vector<uint64_t> seeds(m);
                    std::vector<uint32_t> calls(m);
                    for (int i = 0; i < m; ++i) {
                        gs[i] = q[i]->slot;
                        seeds[i] = q[i]->seed;
                        calls[i] = q[i]->call;
                    }
                    sbpx_sweep_stats st{};
                    if (sbpx_batch_sweep(ctx, gv.first.first, gv.first.second, m, gs.data(), seeds.data(), calls.data(),
                                         &st) != SBPX_OK) {
                        fail(ctx, q, "sbpx_batch_sweep");
                        continue;
                    }
                    b.sweeps += 1;
                    b.totals.attempts += st.attempts;
                    b.totals.evals += st.evals;
                    b.totals.moves += st.moves;
                    for (int k = 0; k < 4; ++k) {
                        b.totals.eval_ms[k] += st.eval_ms[k];
                        b.totals.eval_bytes[k] += st.eval_bytes[k];
                        b.totals.eval_launches[k] += st.eval_launches[k];
                    }
                    for (Request* r : q)
                        if (r->stats) *r->stats = sbpx_sweep_stats{};  // the batch keeps the totals
                }
                break;
            }
            case Request::MERGE: {
                std::map<int, std::vector<Request*>> groups;
                for (Request* r : rs) groups[r->mode].push_back(r);
                for (auto& gv : groups) {
                    std::vector<Request*>& q = gv.second;
                    const int m = (int)q.size();
                    std::vector<int32_t> gs(m), k(m);
                    std::vector<uint64_t> seeds(m);
                    std::vector<uint32_t> calls(m);
                    for (int i = 0; i < m; ++i) {
                        gs[i] = q[i]->slot;
                        k[i] = q[i]->num_to_merge;
                        seeds[i] = q[i]->seed;
                        calls[i] = q[i]->call;
                    }
                    if (sbpx_batch_merge(ctx, m, gs.data(), k.data(), gv.first, seeds.data(), calls.data()) != SBPX_OK) {
                        fail(ctx, q, "sbpx_batch_merge");
                        continue;
                    }
                    std::vector<int32_t> nb(b.vertex_offsets.size() - 1);
                    if (sbpx_batch_info(ctx, nullptr, nb.data()) != SBPX_OK) {
                        fail(ctx, q, "sbpx_batch_info");
                        continue;
                    }
                    for (Request* r : q) r->blocks_left = nb[r->slot];
                }
                break;
            }
        }
    }
}

// Called with the lock held when every running driver waits: runs the round without the lock
// (no driver can act meanwhile) and wakes them. Returns with the lock released.
void run_round(Batch& b, std::unique_lock<std::mutex>& lk) {
    std::vector<Request*> reqs;
    reqs.swap(b.pending);
    const auto t_start = std::chrono::steady_clock::now();
    if (b.rounds > 0) b.between_s += std::chrono::duration<double>(t_start - b.last_end).count();
    b.rounds++;
    lk.unlock();
    try {
        execute(b, reqs);
    } catch (const std::exception& e) {
        for (Request* r : reqs)
            if (r->error.empty()) r->error = e.what();
    }
    b.last_end = std::chrono::steady_clock::now();
    // notified under the request's own lock: its driver may return (and free the request) as
    // soon as it sees `done`, which it cannot before this lock is released
    for (Request* r : reqs) {
        std::lock_guard<std::mutex> g(r->wake_mu);
        r->done = true;
        r->wake.notify_one();
    }
}

void submit(Batch& b, Request& r) {
    {
        std::unique_lock<std::mutex> lk(b.mu);
        if (r.kind == Request::LOGP && r.slot < (int)b.lp_valid.size() && b.lp_valid[r.slot]) {
            r.value = b.lp[r.slot];
            return;
        }
        if ((r.kind == Request::BUILD || r.kind == Request::MERGE || r.kind == Request::SWEEP) &&
            r.slot < (int)b.lp_valid.size())
            b.lp_valid[r.slot] = 0;
        b.pending.push_back(&r);
        if ((int)b.pending.size() == b.live) {
            run_round(b, lk);
        } else {
            lk.unlock();
            std::unique_lock<std::mutex> w(r.wake_mu);
            r.wake.wait(w, [&r] { return r.done; });
        }
    }
    if (!r.error.empty()) throw std::runtime_error(r.error);
}

void leave(Batch& b) {
    std::unique_lock<std::mutex> lk(b.mu);
    b.live--;
    if (b.live > 0 && (int)b.pending.size() == b.live) run_round(b, lk);
}

}  // namespace

void batch_build(Batch& b, int slot, int num_blocks, const int32_t* assignment) {
    Request r{Request::BUILD};
    r.slot = slot;
    r.num_blocks = num_blocks;
    r.assignment = assignment;
    submit(b, r);
}

void batch_download(Batch& b, int slot, int32_t* assignment) {
    Request r{Request::DOWNLOAD};
    r.slot = slot;
    r.out_assignment = assignment;
    submit(b, r);
}

double batch_log_posterior(Batch& b, int slot) {
    Request r{Request::LOGP};
    r.slot = slot;
    submit(b, r);
    return r.value;
}

void batch_sweep(Batch& b, int slot, int algo, int batches, uint64_t seed, uint32_t call, sbpx_sweep_stats* stats) {
    Request r{Request::SWEEP};
    r.slot = slot;
    r.algo = algo;
    r.batches = batches;
    r.seed = seed;
    r.call = call;
    r.stats = stats;
    submit(b, r);
}

int batch_merge(Batch& b, int slot, int num_to_merge, int mode, uint64_t seed, uint32_t call) {
    Request r{Request::MERGE};
    r.slot = slot;
    r.num_to_merge = num_to_merge;
    r.mode = mode;
    r.seed = seed;
    r.call = call;
    submit(b, r);
    return r.blocks_left;
}

long batch_trace(Batch& b, int slot) {
    Request r{Request::TRACE};
    r.slot = slot;
    submit(b, r);
    return (long)r.value;
}

}  // namespace sbpx_host

namespace sbp {

BatchRun partition_batch(const std::vector<int32_t>& vertex_offsets, const std::vector<int32_t>& out_ptr,
                         const std::vector<int32_t>& out_idx, const std::vector<int32_t>& in_ptr,
                         const std::vector<int32_t>& in_idx, const std::vector<long>& num_edges,
                         const std::vector<Args>& graph_args, int device, int eval_variant, bool count_bytes,
                         bool graph_replay) {
    using sbpx_host::Batch;
    using sbpx_host::Device;
    const int G = (int)vertex_offsets.size() - 1;
    if (G < 1) throw std::runtime_error("partition_batch: no graphs");
    if ((int)graph_args.size() != G || (int)num_edges.size() != G)
        throw std::runtime_error("partition_batch: one Args and one edge count per graph");
    for (const Args& a : graph_args) {
        if (a.algorithm != "async_gibbs")
            throw std::runtime_error("partition_batch: a batch of graphs runs asynchronous Gibbs");
        if (!a.approximate) throw std::runtime_error("partition_batch: a batch of graphs merges in the approximate mode");
        if (!a.checkpoint.empty()) throw std::runtime_error("partition_batch: no checkpoints for a batch of graphs");
    }
    const int N = vertex_offsets.back();
    Batch batch;
    batch.vertex_offsets = vertex_offsets;
    auto check = [&](sbpx_status st, const char* what) {
        if (st != SBPX_OK) throw std::runtime_error(std::string(what) + ": " + sbpx_last_error(batch.ctx));
    };
    if (sbpx_create(&batch.ctx, device) != SBPX_OK) throw std::runtime_error("sbpx_create failed");
    struct CtxGuard {
        sbpx_ctx*& c;
        ~CtxGuard() {
            if (c) sbpx_destroy(c);
        }
    } guard{batch.ctx};
    check(sbpx_graph_upload(batch.ctx, N, (int64_t)out_idx.size(), out_ptr.data(), out_idx.data(), in_ptr.data(),
                            in_idx.data()),
          "sbpx_graph_upload");
    check(sbpx_batch_define(batch.ctx, G, vertex_offsets.data()), "sbpx_batch_define");
    check(sbpx_set_option(batch.ctx, 3, eval_variant), "sbpx_set_option");
    check(sbpx_set_option(batch.ctx, 1, count_bytes ? 1 : 0), "sbpx_set_option");
    check(sbpx_set_option(batch.ctx, 9, graph_replay ? 1 : 0), "sbpx_set_option");

    BatchRun run;
    run.graphs.resize(G);
    std::vector<std::string> errors(G);
    batch.live = G;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> threads;
    threads.reserve(G);
    for (int g = 0; g < G; ++g) {
        threads.emplace_back([&, g] {
            try {
                auto dev = std::make_shared<Device>();
                dev->batch = &batch;
                dev->slot = g;
                dev->N = vertex_offsets[g + 1] - vertex_offsets[g];
                long E = num_edges[g] >= 0 ? num_edges[g]
                                           : (long)(out_ptr[vertex_offsets[g + 1]] - out_ptr[vertex_offsets[g]]);
                dev->E = (int)E;
                dev->stored_edges = (long)(out_ptr[vertex_offsets[g + 1]] - out_ptr[vertex_offsets[g]]);
                Graph graph(dev->N, E, dev);
                Args a = graph_args[g];
                Blockmodel bm = stochastic_block_partition(graph, a);
                BatchRun::Graph& out = run.graphs[g];
                out.assignment = bm.block_assignment();
                out.num_blocks = bm.getNum_blocks();
                out.mdl = bm.getOverall_entropy();
                out.intermediates = get_intermediates();
                out.mcmc_iterations = finetune::MCMC_iterations;
            } catch (const std::exception& e) {
                errors[g] = e.what();
            }
            sbpx_host::leave(batch);
        });
    }
    for (std::thread& t : threads) t.join();
    run.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int g = 0; g < G; ++g)
        if (!errors[g].empty()) throw std::runtime_error("graph " + std::to_string(g) + ": " + errors[g]);
    run.rounds = batch.rounds;
    run.sweeps = batch.sweeps;
    run.totals = batch.totals;
    int64_t replayed = 0, captures = 0;
    check(sbpx_batch_replays(batch.ctx, &replayed, &captures), "sbpx_batch_replays");
    run.replayed = replayed;
    run.captures = captures;
    if (std::getenv("SBP_TRACE"))
        std::fprintf(stderr,
                     "batch rounds %ld: build %.3f s, merge %.3f s, sweep %.3f s, log posterior %.3f s, "
                     "download %.3f s, trace %.3f s; between rounds %.3f s\n",
                     batch.rounds, batch.kind_s[sbpx_host::Request::BUILD], batch.kind_s[sbpx_host::Request::MERGE],
                     batch.kind_s[sbpx_host::Request::SWEEP], batch.kind_s[sbpx_host::Request::LOGP], batch.kind_s[sbpx_host::Request::DOWNLOAD],
                     batch.kind_s[sbpx_host::Request::TRACE], batch.between_s);
    return run;
}

}  // namespace sbp
