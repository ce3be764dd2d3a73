// A caller of the C++ layer written the way the reference's own code and tests call it
// (src/main.cpp:78-82; test/entropy_test.cpp, test/finetune_test.cpp, test/block_merge_test.cpp,
// test/blockmodel_test.cpp): Graph, Blockmodel with an assignment, entropy::mdl, the Delta
// builders, entropy::delta_mdl / hastings_correction / block_merge_delta_mdl,
// Blockmodel::move_vertex, blockmatrix()->get, and sbp::stochastic_block_partition.
//
// Usage: toy_caller EDGES_FILE   (first line: N E; then E lines "from to", 0-based)
// Prints "key value..." lines for the test to check against the reference's known answers.
#include <cstdio>
#include <fstream>
#include <tuple>
#include <vector>

#include "sbp_host.hpp"

static Graph read_graph(const char* path) {
    std::ifstream in(path);
    int n = 0, m = 0;
    in >> n >> m;
    std::vector<std::vector<int>> out(n), inn(n);
    for (int i = 0; i < m; ++i) {
        int u, w;
        in >> u >> w;
        out[u].push_back(w);
        inn[w].push_back(u);
    }
    std::vector<int32_t> op(n + 1, 0), oi, ip(n + 1, 0), ii;
    for (int v = 0; v < n; ++v) {
        op[v + 1] = op[v] + (int)out[v].size();
        ip[v + 1] = ip[v] + (int)inn[v].size();
        oi.insert(oi.end(), out[v].begin(), out[v].end());
        ii.insert(ii.end(), inn[v].begin(), inn[v].end());
    }
    return Graph(n, op, oi, ip, ii);
}

static void print_vec(const char* key, const std::vector<int>& v) {
    std::printf("%s", key);
    for (int x : v) std::printf(" %d", x);
    std::printf("\n");
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    Graph graph = read_graph(argv[1]);
    std::vector<int> assignment = {0, 0, 0, 0, 1, 1, 1, 2, 2, 2, 2};
    Blockmodel B(3, graph, 0.5, assignment);
    std::printf("mdl %.9f\n", entropy::mdl(B, graph.num_vertices(), graph.num_edges()));
    print_vec("degrees_out", B.degrees_out());
    print_vec("degrees_in", B.degrees_in());
    print_vec("degrees", B.degrees());
    std::printf("get_0_0 %d\nget_2_1 %d\n", B.blockmatrix()->get(0, 0), B.blockmatrix()->get(2, 1));

    // the vertex-move Delta of vertex 7, block 2 -> 0 (test/finetune_test.cpp:87-102)
    const int vertex = 7;
    const int current_block = B.block_assignment(vertex);
    EdgeWeights out_edges = finetune::edge_weights(graph.out_neighbors(), vertex, false);
    EdgeWeights in_edges = finetune::edge_weights(graph.in_neighbors(), vertex, false);
    Delta delta = finetune::blockmodel_delta(vertex, current_block, 0, out_edges, in_edges, B);
    std::printf("delta_size %zu\n", delta.entries().size());
    for (const auto& e : delta.entries())
        std::printf("delta %d %d %d\n", std::get<0>(e), std::get<1>(e), std::get<2>(e));
    utils::ProposalAndEdgeCounts proposal{0, 2, 3, 5};
    std::printf("delta_mdl %.9f\n", entropy::delta_mdl(B, delta, proposal));
    std::printf("hastings %.9f\n", entropy::hastings_correction(vertex, graph, B, delta, current_block, proposal));

    // merging block 0 into block 1 (test/block_merge_test.cpp:11-21)
    Delta merge = block_merge::blockmodel_delta(0, 1, B);
    for (const auto& e : merge.entries())
        std::printf("merge_delta %d %d %d\n", std::get<0>(e), std::get<1>(e), std::get<2>(e));
    utils::ProposalAndEdgeCounts merge_proposal{1, 8, 9, 17};
    std::printf("merge_delta_mdl %.9f\n", entropy::block_merge_delta_mdl(0, merge_proposal, B, merge));

    // move_vertex on the six-block assignment: vertex 6, block 3 -> 0 (toy_example.hpp:174-176)
    std::vector<int> assignment6 = {0, 0, 0, 1, 2, 3, 3, 4, 5, 1, 5};
    Blockmodel B6(6, graph, 0.5, assignment6);
    const int v6 = 6, r6 = B6.block_assignment(v6);
    EdgeWeights o6 = finetune::edge_weights(graph.out_neighbors(), v6, false);
    EdgeWeights i6 = finetune::edge_weights(graph.in_neighbors(), v6, true);
    Delta d6 = finetune::blockmodel_delta(v6, r6, 0, o6, i6, B6);
    utils::ProposalAndEdgeCounts p6{0, (int)o6.indices.size(), (int)i6.indices.size(),
                                    (int)(o6.indices.size() + i6.indices.size())};
    B6.move_vertex(v6, d6, p6);
    print_vec("moved_degrees_out", B6.degrees_out());
    print_vec("moved_degrees_in", B6.degrees_in());
    print_vec("moved_degrees", B6.degrees());

    // the driver as src/main.cpp:79 calls it
    Args a;
    a.algorithm = "hybrid_mcmc";
    a.batches = 1;
    a.approximate = true;
    a.seed = 1;
    Blockmodel result = sbp::stochastic_block_partition(graph, a);
    std::printf("partition_blocks %d\npartition_mdl %.6f\n", result.getNum_blocks(),
                (double)result.getOverall_entropy());
    return 0;
}
