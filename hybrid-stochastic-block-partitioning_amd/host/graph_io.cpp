// Graph input with the reference's semantics (Graph::load, graph.cpp:24-62; load_matrix_market
// :64-115; load_text :118-140; parse_directed / parse_undirected :165-210; utils::read_csv,
// utils.cpp:19-42): whitespace-separated fields, 1-based vertex ids, neighbour lists in order of
// first appearance without duplicates, self-loops flagged, vertex count = the largest id seen.
//   .tsv (one "from to" pair per line; further fields ignored): the edge count is the number of
//     distinct out-neighbour entries, halved for an undirected graph (whose reverse edges are
//     stored as well).
//   .mtx (MatrixMarket coordinate): the edge count starts from the size line and grows by one
//     for each reverse edge a symmetric matrix adds; "symmetric" makes the graph undirected.
// The edge count is the one the reference's MDL uses (entropy::mdl takes it as an argument); it
// can differ from the number of stored edges.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbp_host.hpp"

namespace graph_io {

namespace {

std::vector<std::vector<std::string>> read_fields(const std::string& path, bool& found) {
    std::vector<std::vector<std::string>> rows;
    std::ifstream f(path);
    found = (bool)f;
    if (!f) return rows;
    std::string line;
    while (std::getline(f, line)) {
        std::vector<std::string> row;
        std::stringstream ss(line);
        std::string v;
        while (ss >> v) row.push_back(v);
        rows.push_back(row);
    }
    return rows;
}

int to_int(const std::string& s, const std::string& path) {
    size_t pos = 0;
    int v = 0;
    try {
        v = std::stoi(s, &pos);
    } catch (const std::exception&) {
        throw std::runtime_error(path + ": not an integer: '" + s + "'");
    }
    return v;
}

// Neighbour lists that keep the order of first insertion (utils::insert_nodup, utils.cpp:52-60).
struct Lists {
    std::vector<std::vector<int>> out, in;
    void grow(size_t n) {
        if (out.size() < n) out.resize(n);
        if (in.size() < n) in.resize(n);
    }
    static void insert_nodup(std::vector<std::vector<int>>& l, int from, int to) {
        if ((size_t)from >= l.size()) l.resize(from + 1);
        for (int x : l[from])
            if (x == to) return;
        l[from].push_back(to);
    }
};

LoadedGraph finish(Lists& L, int num_vertices, long num_edges, bool undirected) {
    L.grow(num_vertices);
    LoadedGraph g;
    g.num_vertices = num_vertices;
    g.num_edges = num_edges;
    g.undirected = undirected;
    g.out_ptr.assign(1, 0);
    g.in_ptr.assign(1, 0);
    for (int v = 0; v < num_vertices; ++v) {
        g.out_idx.insert(g.out_idx.end(), L.out[v].begin(), L.out[v].end());
        g.out_ptr.push_back((int32_t)g.out_idx.size());
        g.in_idx.insert(g.in_idx.end(), L.in[v].begin(), L.in[v].end());
        g.in_ptr.push_back((int32_t)g.in_idx.size());
    }
    return g;
}

LoadedGraph load_matrix_market(const std::vector<std::vector<std::string>>& rows, const std::string& path) {
    const auto& head = rows[0];
    if (head.size() < 5 || head[2] != "coordinate") throw std::runtime_error(path + ": dense matrices are not supported");
    const bool undirected = head[4] == "symmetric";
    size_t index = rows.size();
    int num_vertices = 0;
    long num_edges = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
        if (rows[i].empty() || rows[i][0][0] == '%') continue;
        if (rows[i].size() < 3) throw std::runtime_error(path + ": malformed size line");
        num_vertices = to_int(rows[i][0], path);
        if (num_vertices != to_int(rows[i][1], path)) throw std::runtime_error(path + ": rectangular matrices are not supported");
        num_edges = to_int(rows[i][2], path);
        index = i + 1;
        break;
    }
    Lists L;
    for (size_t i = index; i < rows.size(); ++i) {
        if (rows[i].size() < 2) continue;
        const int from = to_int(rows[i][0], path) - 1, to = to_int(rows[i][1], path) - 1;
        if (from < 0 || to < 0) throw std::runtime_error(path + ": vertex ids start at 1");
        num_vertices = std::max(num_vertices, std::max(from, to) + 1);
        Lists::insert_nodup(L.out, from, to);
        Lists::insert_nodup(L.in, to, from);
        if (undirected && from != to) {
            Lists::insert_nodup(L.out, to, from);
            Lists::insert_nodup(L.in, from, to);
            num_edges++;
        }
    }
    return finish(L, num_vertices, num_edges, undirected);
}

LoadedGraph load_text(const std::vector<std::vector<std::string>>& rows, bool undirected, const std::string& path) {
    Lists L;
    int num_vertices = 0;
    for (const auto& r : rows) {
        if (r.size() < 2) continue;
        const int from = to_int(r[0], path) - 1, to = to_int(r[1], path) - 1;
        if (from < 0 || to < 0) throw std::runtime_error(path + ": vertex ids start at 1");
        num_vertices = std::max(num_vertices, std::max(from, to) + 1);
        Lists::insert_nodup(L.out, from, to);
        Lists::insert_nodup(L.in, to, from);
        if (undirected && from != to) {  // parse_undirected: both directions stored
            Lists::insert_nodup(L.out, to, from);
            Lists::insert_nodup(L.in, from, to);
        }
    }
    L.grow(num_vertices);
    long num_edges = 0;
    for (const auto& l : L.out) num_edges += (long)l.size();
    if (undirected) num_edges /= 2;
    return finish(L, num_vertices, num_edges, undirected);
}

}  // namespace

LoadedGraph load(const std::string& base_path, bool undirected) {
    bool found = false;
    std::string path = base_path + ".tsv";
    std::vector<std::vector<std::string>> rows = read_fields(path, found);
    if (!found || rows.empty()) {
        path = base_path + ".mtx";
        rows = read_fields(path, found);
    }
    if (!found || rows.empty()) throw std::runtime_error("no graph at " + base_path + ".tsv or .mtx");
    LoadedGraph g = (!rows[0].empty() && rows[0][0] == "%%MatrixMarket") ? load_matrix_market(rows, path)
                                                                         : load_text(rows, undirected, path);
    // truth: "vertex block" lines, both 1-based (graph.cpp:44-61); no file gives all zeros
    std::vector<std::vector<std::string>> truth = read_fields(base_path + "_truePartition.tsv", found);
    g.truth.clear();
    if (found && !truth.empty()) {
        for (const auto& r : truth) {
            if (r.size() < 2) continue;
            const int v = to_int(r[0], path) - 1, b = to_int(r[1], path) - 1;
            if (v < 0) throw std::runtime_error(base_path + "_truePartition.tsv: vertex ids start at 1");
            if ((size_t)v >= g.truth.size()) g.truth.resize(v + 1, -1);
            g.truth[v] = b;
        }
    } else {
        g.truth.assign(g.num_vertices, 0);
    }
    return g;
}

}  // namespace graph_io

Graph Graph::load(const std::string& base_path, bool undirected, int device) {
    graph_io::LoadedGraph g = graph_io::load(base_path, undirected);
    return Graph(g.num_vertices, g.out_ptr, g.out_idx, g.in_ptr, g.in_idx, device, g.num_edges);
}
