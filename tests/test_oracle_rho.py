"""The oracle's McWEDA density assemblers (j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90):
bit-exact against the reference's own module on the golden cases, plus the
properties the mathematics fixes (pair-reversal symmetry of the two-centre
densities, linearity in the shell charges, rotation covariance of the orbital
blocks and invariance of the shell blocks)."""
import copy

import numpy as np
import pytest

import golden_cases
import oracle
from fireball_hip import synth
from golden.make_golden import RKEYS, tables_digest


def rho_inputs(g, gr):
    ts, st = golden_cases.inputs(g)
    tsr = copy.copy(ts)
    tsr.cells2c = dict(ts.cells2c)
    tsr.cells3c = dict(ts.cells3c)
    synth.add_rho_tables(tsr, seed=int(gr["rho_table_seed"]))
    return tsr, st


@pytest.fixture(scope="module")
def rho_runs(golden):
    out = {}
    for name in golden_cases.NAMES + golden_cases.NAMES_D:
        g, gr = golden(name), golden(name + "_rho")
        ts, st = rho_inputs(g, gr)
        o = oracle.System(oracle.Tables(ts), st)
        o.build_lists()
        out[name] = (gr, ts, o.run_rho(gr["qin"]))
    return out


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_rho_tables_regenerate(rho_runs, name):
    gr, ts, _ = rho_runs[name]
    assert tables_digest(ts) == str(gr["rho_tables_digest"])


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_rho_bit_exact_against_reference(rho_runs, name):
    gr, _, r = rho_runs[name]
    for k in RKEYS:
        n = gr[k].size
        assert np.array_equal(r[k][:n], gr[k]), k
        assert np.sum(r[k]) == gr[k + "_sum"], k
        assert np.sum(r[k] ** 2) == gr[k + "_sumsq"], k


def _pairs(L):
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    pair_atom = np.repeat(np.arange(len(L["neighn"])), L["neighn"])
    back = off[L["neigh_j"] - 1] + L["neigh_back"] - 1
    return off, pair_atom, back


def _split(flat, L, imass, size):
    off, pair_atom, _ = _pairs(L)
    out, pos = [], 0
    for p in range(len(L["neigh_j"])):
        a, b = size[imass[pair_atom[p]] - 1], size[imass[L["neigh_j"][p] - 1] - 1]
        out.append(flat[pos:pos + a * b].reshape(b, a).T)
        pos += a * b
    return out


def _system(seed=3):
    sp = [synth.C_SP, synth.H_SP, synth.O_SP]
    ts = synth.add_rho_tables(synth.make_tables(sp, seed=seed), seed=seed + 1)
    st = synth.random_box({6: 6, 1: 8, 8: 4}, 1.0, 0.9, 4, sp)
    return sp, ts, st


def test_rho_2c_pair_reversal_symmetry():
    # rho_bond holds only two-centre terms: with tables mirrored across the two
    # centres, rho_bond(j, i) = rho_bond(i, j)^T, and the shell blocks likewise
    sp, ts, st = _system()
    o = oracle.System(oracle.Tables(ts), st)
    L = o.build_lists()
    r = o.run_rho(synth.shell_qin(st, sp, seed=9))
    norb = [s.norb for s in sp]
    nssh = [s.nssh for s in sp]
    _, _, back = _pairs(L)
    for key, size in (("RHO_BOND", norb), ("RHOW_BOND", nssh), ("RHOW_IN", None)):
        if size is None:
            continue
        B = _split(r[key], L, st.imass, size)
        for p in range(len(B)):
            np.testing.assert_allclose(B[p], B[back[p]].T, atol=1e-12, err_msg=key)
    # the 2c part of rho_in equals rho_bond off the self blocks
    o2 = oracle.System(oracle.Tables(ts), st)
    o2.build_lists()
    o2.set_qin(synth.shell_qin(st, sp, seed=9))
    o2.assemble_rho(("RHO2C", "RHOW2C"))
    off, _, _ = _pairs(L)
    selfs = {off[i] + L["neigh_self"][i] - 1 for i in range(st.natoms)}
    Bi = _split(o2.rho("RHO_IN"), L, st.imass, norb)
    Bb = _split(o2.rho("RHO_BOND"), L, st.imass, norb)
    for p in range(len(Bi)):
        if p not in selfs:
            assert np.array_equal(Bi[p], Bb[p])


def test_rho_linear_in_shell_charges():
    sp, ts, st = _system(seed=6)
    tab = oracle.Tables(ts)
    q1 = synth.shell_qin(st, sp, seed=1)
    q2 = synth.shell_qin(st, sp, seed=2) * 0.3
    res = []
    for q in (q1, q2, q1 + q2):
        o = oracle.System(tab, st)
        o.build_lists()
        res.append(o.run_rho(q))
    for k in RKEYS:
        scale = np.abs(res[2][k]).max()
        np.testing.assert_allclose(res[0][k] + res[1][k], res[2][k], atol=1e-13 * scale * 50, err_msg=k)


def test_rho_rotation_covariance():
    # rotating a cluster rotates the p part of every non-self orbital block
    # (2c ontop and 3c terms) and leaves every shell block unchanged
    sp = [synth.C_SP, synth.H_SP]
    ts = synth.add_rho_tables(synth.make_tables(sp, seed=8), seed=9)
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 4.5, size=(8, 3)) + 10.0
    imass = np.array([1, 2, 1, 1, 2, 1, 2, 1], np.int32)
    st = synth.Structure(pos, imass, np.eye(3) * 40, 1)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    st2 = synth.Structure(pos @ q.T, imass, st.lattice, 1)
    qin = synth.shell_qin(st, sp, seed=4)
    tab = oracle.Tables(ts)
    res, lists = [], []
    for s in (st, st2):
        o = oracle.System(tab, s)
        L = o.build_lists()
        assert L["ntriplets"] > 20
        res.append(o.run_rho(qin))
        lists.append(L)
    L = lists[0]
    for k in ("neigh_j", "neigh_b", "com_mneigh"):
        assert np.array_equal(lists[0][k], lists[1][k])
    off, pair_atom, _ = _pairs(L)
    selfs = {off[i] + L["neigh_self"][i] - 1 for i in range(st.natoms)}
    norb = [s.norb for s in sp]
    perm = [1, 2, 0]                   # (y, z, x) orbital order
    R = q[np.ix_(perm, perm)]
    A = _split(res[0]["RHO_IN"], L, imass, norb)
    B = _split(res[1]["RHO_IN"], L, imass, norb)
    checked = 0
    for p in range(len(A)):
        if p in selfs:
            continue
        a, b = A[p], B[p]
        scale = max(1.0, np.abs(a).max())
        np.testing.assert_allclose(b[0, 0], a[0, 0], atol=1e-12 * scale)
        if a.shape[1] == 4:
            np.testing.assert_allclose(b[0, 1:], R @ a[0, 1:], atol=1e-12 * scale)
        if a.shape[0] == 4:
            np.testing.assert_allclose(b[1:, 0], R @ a[1:, 0], atol=1e-12 * scale)
        if a.shape == (4, 4):
            np.testing.assert_allclose(b[1:, 1:], R @ a[1:, 1:] @ R.T, atol=1e-12 * scale)
        checked += 1
    assert checked > 20
    for k in ("RHOW_IN", "RHOW_BOND"):
        np.testing.assert_allclose(res[1][k], res[0][k], atol=1e-12 * np.abs(res[0][k]).max())


def test_rho_order_errors():
    sp, ts, st = _system()
    o = oracle.System(oracle.Tables(ts), st)
    o.build_lists()
    with pytest.raises(oracle.OracleError):
        o.assemble_rho(("RHO2C",))          # no shell charges yet
    o.set_qin(synth.shell_qin(st, sp))
    with pytest.raises(oracle.OracleError):
        o.assemble_rho(("RHO3C",))          # before rho_2c
    with pytest.raises(oracle.OracleError):
        o.assemble_rho(("RHOW3C",))         # before rho_weighted_2c
    o.assemble_rho()
