"""The C++ layer called as the reference's code and tests call it (tests/cpp/toy_caller.cpp, built
by build.py): on the reference's toy graph, the known answers of test/entropy_test.cpp,
test/finetune_test.cpp, test/block_merge_test.cpp and test/blockmodel_test.cpp, independent
cross-check values for delta MDL and the Hastings correction, move_vertex's degrees with the
total-degree drift, and a whole sbp::stochastic_block_partition against the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "cpp", "toy_caller")


def _run(toy, tmp_path):
    if not os.path.exists(CALLER):
        pytest.fail("tests/cpp/toy_caller is not built (build.py builds it)")
    path = tmp_path / "toy_edges.txt"
    edges = sorted(set(map(tuple, toy["edges"])))
    path.write_text(f"{toy['num_vertices']} {len(edges)}\n" + "".join(f"{u} {w}\n" for u, w in edges))
    out = subprocess.run([CALLER, str(path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = {}
    for line in out.stdout.splitlines():
        key, *vals = line.split()
        rows.setdefault(key, []).append(vals)
    return rows, edges


def test_reference_style_caller_reproduces_known_answers(toy, crosschecks, tmp_path):
    rows, edges = _run(toy, tmp_path)
    one = lambda k: rows[k][0]  # noqa: E731
    assert float(one("mdl")[0]) == pytest.approx(toy["entropy"], rel=1e-7)
    assert [int(x) for x in one("degrees_out")] == toy["block_degrees_out"]
    assert [int(x) for x in one("degrees_in")] == toy["block_degrees_in"]
    assert [int(x) for x in one("degrees")] == toy["block_degrees"]
    g = O.Graph.from_edges(toy["num_vertices"], edges)
    dense = O.Blockmodel(g, 3, toy["assignment"]).dense()
    assert int(one("get_0_0")[0]) == dense[0, 0] and int(one("get_2_1")[0]) == dense[2, 1]
    # finetune_test.cpp:87-102: six entries with these values; (0, 1) and (2, 1) read 0
    delta = sorted(tuple(int(x) for x in r) for r in rows["delta"])
    assert int(one("delta_size")[0]) == 6
    assert delta == sorted(tuple(e) for e in toy["move"]["delta"])
    move = next(mv for mv in crosschecks["moves"] if mv["B"] == 3 and mv["v"] == 7)
    assert float(one("delta_mdl")[0]) == pytest.approx(move["dS"], rel=1e-6)
    assert float(one("hastings")[0]) == pytest.approx(move["H"], rel=1e-6)
    # block_merge_test.cpp:11-21
    merge = sorted(tuple(int(x) for x in r) for r in rows["merge_delta"])
    assert merge == sorted(tuple(e) for e in toy["merge"]["delta"])
    assert float(one("merge_delta_mdl")[0]) == pytest.approx(crosschecks["merge_0_to_1_dS"], rel=1e-6)
    # toy_example.hpp:174-176, with the reference's drift of the current block's total degree
    assert [int(x) for x in one("moved_degrees_out")] == toy["complex"]["after_out"]
    assert [int(x) for x in one("moved_degrees_in")] == toy["complex"]["after_in"]
    want = list(toy["complex"]["after_total"])
    want[3] = crosschecks["q9_drift"]["reference_deg3"]
    assert [int(x) for x in one("moved_degrees")] == want
    # the driver as main.cpp calls it, against the oracle's run with the same seed and flags
    _, run = O.sbp(g, 2, 1, True, 1)
    assert int(one("partition_blocks")[0]) == run.B
    assert float(one("partition_mdl")[0]) == pytest.approx(run.mdl, rel=5e-3)
