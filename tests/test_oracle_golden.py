"""The C oracle against the golden vectors written by the reference's own
Fortran (tests/golden/make_golden.py): lists bit-exact, blocks to rounding."""
import numpy as np
import pytest

import golden_cases
import oracle
from golden.make_golden import tables_digest  # noqa: F401  (same digest as the generator)

LIST_KEYS = ("neighn", "neigh_j", "neigh_b", "neigh_back", "neigh_self", "ncommon", "com_iatom_j",
             "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh")
MATS = ("S", "T", "DIP", "VNA", "VNAO")


@pytest.fixture(scope="module")
def oracle_runs(golden):
    out = {}
    for name in golden_cases.NAMES + golden_cases.NAMES_D:
        g = golden(name)
        ts, st = golden_cases.inputs(g)
        out[name] = (g, ts, oracle.System(oracle.Tables(ts), st).run_all())
    return out


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_tables_regenerate(oracle_runs, name):
    g, ts, _ = oracle_runs[name]
    assert str(g["tables_digest"]) == tables_digest(ts)


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_lists_bit_exact(oracle_runs, name):
    g, _, o = oracle_runs[name]
    for k in LIST_KEYS:
        assert np.array_equal(np.asarray(o[k]), g[k]), k


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_blocks_match_reference(oracle_runs, name):
    g, _, o = oracle_runs[name]
    n = len(g["S"])
    for m in MATS:
        np.testing.assert_allclose(o[m][:n], g[m], rtol=0, atol=1e-13, err_msg=m)
        assert abs(np.sum(o[m]) - g[m + "_sum"]) <= 1e-12 * max(1.0, g[m + "_sumabs"]), m
        assert abs(np.sum(o[m] ** 2) - g[m + "_sumsq"]) <= 1e-12 * max(1.0, g[m + "_sumsq"]), m


def test_fallback_branches_exercised(oracle_runs):
    # the chain cluster is built so that collinear frames occur
    _, _, o = oracle_runs["chain_cluster"]
    assert o["frame_fallbacks"] > 0
