"""Partition quality against a planted partition (host side, numpy).

nmi: evaluate::calculate_nmi (evaluate.cpp:61-96) on the contingency table of
evaluate::hungarian (evaluate.cpp:114-206): H terms use the FP32 log the reference applies through
utils::nat_log, the mutual information uses the FP64 log. Empty blocks contribute 0 log 0 = 0
(the reference's table keeps empty algorithm blocks as zero rows and its H term then turns NaN).

f1: evaluate::calculate_f1_score (evaluate.cpp:7-59), pair counting.

modularity: Graph::modularity (graph.cpp:142-163) sums A_ij - dout_i din_j / E over all ordered
vertex pairs in the same block, O(N^2 deg). Grouping the pairs by block gives the same value in
O(E + N): Q = (1/E) sum_b [e_bb - Dout_b Din_b / E], with e_bb the edges inside block b
(self-loops included) and Dout_b, Din_b the summed out/in degrees of its vertices.
"""
import numpy as np


def contingency(truth, assignment):
    truth = np.asarray(truth, np.int64)
    assignment = np.asarray(assignment, np.int64)
    _, t = np.unique(truth, return_inverse=True)
    B = int(assignment.max()) + 1
    K = int(t.max()) + 1
    table = np.zeros((B, K), np.int64)
    np.add.at(table, (assignment, t), 1)
    return table


def nmi(truth, assignment):
    table = contingency(truth, assignment)
    n = table.sum()
    joint = table / float(n)
    p1 = joint.sum(1)
    p2 = joint.sum(0)

    def h(p):
        p = p[p > 0]
        return -float(np.sum(p * np.log(p.astype(np.float32)).astype(np.float64)))

    nz = joint > 0
    outer = np.outer(p1, p2)
    mi = float(np.sum(joint[nz] * np.log(joint[nz] / outer[nz])))
    return mi / np.sqrt(h(p1) * h(p2))


def f1_score(num_vertices, table):
    table = np.asarray(table, np.float64)
    num_pairs = (num_vertices * (num_vertices - 1.0)) / 2.0
    rowsums = table.sum(1)
    colsums = table.sum(0)
    cell_pairs = float(np.sum(table * (table - 1))) / 2.0
    row_pairs = float(np.sum(rowsums * (rowsums - 1))) / 2.0
    col_pairs = float(np.sum(colsums * (colsums - 1))) / 2.0
    recall = cell_pairs / row_pairs
    precision = cell_pairs / col_pairs
    return (2.0 * precision * recall) / (precision + recall)


def modularity(out_ptr, out_idx, in_ptr, assignment):
    out_ptr = np.asarray(out_ptr, np.int64)
    in_ptr = np.asarray(in_ptr, np.int64)
    a = np.asarray(assignment, np.int64)
    N = len(out_ptr) - 1
    E = float(out_ptr[-1])
    src = np.repeat(np.arange(N), np.diff(out_ptr))
    dst = np.asarray(out_idx, np.int64)
    B = int(a.max()) + 1
    inside = np.bincount(a[src][a[src] == a[dst]], minlength=B).astype(np.float64)
    dout = np.bincount(a, weights=np.diff(out_ptr), minlength=B)
    din = np.bincount(a, weights=np.diff(in_ptr), minlength=B)
    return float(np.sum(inside - dout * din / E) / E)
