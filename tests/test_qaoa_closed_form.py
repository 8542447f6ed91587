"""The p=1 QAOA MaxCut closed form (tests/helpers.qaoa_p1_edge_expectation)
against brute force on the oracle: 3-regular graphs with and without
triangles, the committed reference graphs, and random angles.  The GPU batch
tests use the closed form at sizes the oracle cannot reach."""

import numpy as np
import pytest

from aleph_amd import workloads as W
from tests.helpers import GOLDEN, qaoa_p1_cut_expectation


def read_edges(name):
    return [tuple(int(t) for t in line.split()) for line in (GOLDEN / "qaoa" / f"{name}.edges").read_text().split("\n")
            if line.strip() and not line.startswith("#")]


def brute(orc, n, edges, g, b):
    st = orc.OracleSv().run(W.qaoa_maxcut(n, edges, [g], [b]))
    p = np.abs(st.amplitudes()) ** 2
    idx = np.arange(1 << n)
    return float(sum(np.sum(p * (((idx >> u) ^ (idx >> v)) & 1)) for u, v in edges))


@pytest.mark.parametrize("seed", range(6))
def test_closed_form_matches_brute_force_on_3_regular(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([6, 8, 10, 12]))
    edges = W.random_3_regular(n, 100 + seed)
    g, b = float(rng.uniform(0, np.pi)), float(rng.uniform(0, np.pi))
    assert abs(qaoa_p1_cut_expectation(n, edges, g, b) - brute(oracle_lib, n, edges, g, b)) < 1e-12


def test_closed_form_with_triangles(oracle_lib):
    # K4 plus a pendant path: edges with 0, 1 and 2 common neighbours.
    n, edges = 6, [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3), (3, 4), (4, 5)]
    for g, b in ((0.3, 2.1), (1.9, 0.4)):
        assert abs(qaoa_p1_cut_expectation(n, edges, g, b) - brute(oracle_lib, n, edges, g, b)) < 1e-12


@pytest.mark.parametrize("name", ["qaoa_n6", "qaoa_n10", "qaoa_n14"])
def test_closed_form_on_reference_graphs(oracle_lib, name):
    edges = read_edges(name)
    n = 1 + max(max(e) for e in edges)
    maxcut = float((GOLDEN / "qaoa" / f"{name}.maxcut").read_text().split()[0])
    for g, b in ((0.6, 0.35), (2.3, 1.7)):
        want = brute(oracle_lib, n, edges, g, b)
        assert abs(qaoa_p1_cut_expectation(n, edges, g, b) - want) < 1e-12
        assert want <= maxcut + 1e-12
