"""The oracle's pseudopotential path (c.SYSTEM/M_neighbors_PP.f90,
j.ASSEMBLERS/M_assemble_PP_2c.f90, M_assemble_PP_3c.f90): bit-exact against
the reference's own modules on the golden cases, and pinned independently of
the reference by what the definitions fix:

* range duality of the PP and PPx lists, and the maps and points pointing at
  the entries they name;
* the assembled vnl is the Kleinman-Bylander operator of every PP atom,
  V = sum_alpha P_alpha diag(cl_alpha) P_alpha^T with P_alpha the sVNL blocks
  <phi_i | Psi_alpha> of all orbitals near alpha: dense linear algebra on the
  sVNL blocks must reproduce the list bookkeeping of the atom, ontop and
  three-center terms (cluster, and Gamma-folded on a periodic cell);
* rotating a cluster rotates the p blocks of vnl."""
import numpy as np
import pytest

import golden_cases
import oracle
from fireball_hip import synth
from golden.make_golden import PP_LIST_KEYS, pp_tables, tables_digest


@pytest.fixture(scope="module")
def pp_runs(golden):
    out = {}
    for name in golden_cases.NAMES + golden_cases.NAMES_D:
        g, gp = golden(name), golden(name + "_pp")
        ts, st = golden_cases.inputs(g)
        tsp = pp_tables(ts, int(gp["pp_table_seed"]))
        o = oracle.System(oracle.Tables(tsp), st)
        o.build_lists()
        out[name] = (gp, tsp, st, o.run_pp())
    return out


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_pp_tables_regenerate(pp_runs, name):
    gp, tsp, _, _ = pp_runs[name]
    assert tables_digest(tsp) == str(gp["pp_tables_digest"])


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_pp_bit_exact_against_reference(pp_runs, name):
    gp, _, _, r = pp_runs[name]
    for k in PP_LIST_KEYS:
        assert np.array_equal(r[k], gp[k]), k
    for k in ("SVNL", "VNL"):
        assert np.array_equal(r[k][:gp[k].size], gp[k]), k
        assert np.sum(r[k]) == gp[k + "_sum"], k
        assert np.sum(r[k] ** 2) == gp[k + "_sumsq"], k


def _csr(r, name):
    off = np.concatenate([[0], np.cumsum(r[name + "_n"])]).astype(np.int64)
    return off


def _image_index(lattice):
    xl = oracle.make_cells(lattice, 0)
    return {tuple(np.round(np.linalg.solve(lattice.T, xl[b])).astype(int)): b for b in range(len(xl))}, xl


@pytest.mark.parametrize("name", ["si8_cell", "cho40_dq", "chain_cluster"])
def test_pp_list_duality_and_maps(pp_runs, name):
    _, _, st, r = pp_runs[name]
    N = st.natoms
    opp, opx, opq = _csr(r, "pp"), _csr(r, "ppx"), _csr(r, "ppp")
    if st.icluster == 1:
        neg = {0: 0}
    else:
        idx, xl = _image_index(st.lattice)
        neg = {b: idx[tuple(-np.array(k))] for k, b in idx.items()}
    pp = {(i, int(r["pp_j"][p]) - 1, int(r["pp_b"][p])) for i in range(N) for p in range(opp[i], opp[i + 1])}
    px = {(i, int(r["ppx_j"][p]) - 1, int(r["ppx_b"][p])) for i in range(N) for p in range(opx[i], opx[i + 1])}
    # range duality: j at image b is a PP neighbour of i  <=>  i at image -b is a PPx neighbour of j
    assert {(j, i, neg[b]) for (i, j, b) in pp} == px
    for i in range(N):
        q = {(int(r["ppp_j"][p]), int(r["ppp_b"][p])): p - opq[i] + 1 for p in range(opq[i], opq[i + 1])}
        assert len(q) == opq[i + 1] - opq[i]                     # PPp holds no duplicates
        for p in range(opp[i], opp[i + 1]):
            assert r["pp_map"][p] == q[(int(r["pp_j"][p]), int(r["pp_b"][p]))]
        for p in range(opx[i], opx[i + 1]):
            assert r["ppx_map"][p] == q[(int(r["ppx_j"][p]), int(r["ppx_b"][p]))]
            j, b = int(r["ppx_j"][p]) - 1, int(r["ppx_b"][p])
            k = opp[j] + r["ppx_point"][p] - 1                   # i among j's PP neighbours at image -b
            assert r["pp_j"][k] - 1 == i and r["pp_b"][k] == neg[b]
        for name_, off in (("pp", opp), ("ppx", opx), ("ppp", opq)):
            s = off[i] + r[name_ + "_self"][i] - 1
            assert r[name_ + "_j"][s] - 1 == i and r[name_ + "_b"][s] == 0


def _kb_operator_check(tsp, st, r):
    """Gamma-folded vnl against sum_alpha P_alpha C_alpha P_alpha^T."""
    sp = tsp.species
    norb = np.array([s.norb for s in sp])
    norb_pp = np.array([s.norb_pp for s in sp])
    N = st.natoms
    orb0 = np.concatenate([[0], np.cumsum(norb[st.imass - 1])])
    ntot = orb0[-1]
    opp, opq = _csr(r, "pp"), _csr(r, "ppp")
    # dense P_alpha (ntot x norb_pp(alpha)), summing the sVNL blocks over images
    P = [np.zeros((ntot, norb_pp[st.imass[a] - 1])) for a in range(N)]
    pos = 0
    for i in range(N):
        for p in range(opp[i], opp[i + 1]):
            a = int(r["pp_j"][p]) - 1
            nm, nc = norb[st.imass[i] - 1], norb_pp[st.imass[a] - 1]
            P[a][orb0[i]:orb0[i + 1], :] += r["SVNL"][pos:pos + nm * nc].reshape(nc, nm).T
            pos += nm * nc
    V = np.zeros((ntot, ntot))
    for a in range(N):
        s = sp[st.imass[a] - 1]
        cl = np.repeat(s.cl_pp, [2 * l + 1 for l in s.lssh_pp])
        V += P[a] @ np.diag(cl) @ P[a].T
    F = np.zeros((ntot, ntot))
    pos = 0
    for i in range(N):
        for p in range(opq[i], opq[i + 1]):
            j = int(r["ppp_j"][p]) - 1
            nm, nn = norb[st.imass[i] - 1], norb[st.imass[j] - 1]
            F[orb0[i]:orb0[i + 1], orb0[j]:orb0[j + 1]] += r["VNL"][pos:pos + nm * nn].reshape(nn, nm).T
            pos += nm * nn
    np.testing.assert_allclose(F, V, atol=1e-12 * np.abs(V).max())
    np.testing.assert_allclose(F, F.T, atol=1e-12 * np.abs(V).max())


def test_vnl_is_the_kb_operator_cluster(pp_runs):
    _, tsp, st, r = pp_runs["chain_cluster"]
    _kb_operator_check(tsp, st, r)
    sp = synth.with_pp([synth.C_SP, synth.H_SP])
    ts = synth.add_pp_tables(synth.make_tables(sp, seed=12, with_3c=False), seed=13)
    rng = np.random.default_rng(4)
    pos = rng.uniform(0, 5.0, size=(10, 3)) + 10.0
    st2 = synth.Structure(pos, np.array([1, 2, 1, 1, 2, 1, 2, 1, 1, 2], np.int32), np.eye(3) * 40, 1)
    o = oracle.System(oracle.Tables(ts), st2)
    o.build_lists()
    r2 = o.run_pp()
    assert len(r2["ppc_m"]) > 50
    _kb_operator_check(ts, st2, r2)


@pytest.mark.parametrize("name", ["si8_cell", "cho40_dq"])
def test_vnl_is_the_kb_operator_periodic(pp_runs, name):
    _, tsp, st, r = pp_runs[name]
    _kb_operator_check(tsp, st, r)


def _isotropic_at_origin(ts):
    """Make the <phi|Psi_PP> tables physical at zero separation: the s-p
    elements vanish there (odd parity) and the p-p pi elements meet the sigma
    one, so the self block is rotation invariant and its default frame at z = 0
    does not matter."""
    for (a, b, it, isub), c in ts.cells2c.items():
        if it != synth.P_VNL:
            continue
        r = np.arange(c.nx) * (c.xmax / (c.nx - 1))
        w = r / c.xmax
        _, _, _, shells = synth.munu_pp(ts.species[a - 1], ts.species[b - 1])
        li = ts.species[a - 1].lssh
        lj = ts.species[b - 1].lssh_pp
        sig = {}
        for k, (i, j, m) in enumerate(shells):
            if li[i] == 1 and lj[j] == 1 and m == 0:
                sig[(i, j)] = c.data[:, k].copy()
        for k, (i, j, m) in enumerate(shells):
            if li[i] + lj[j] == 1:
                c.data[:, k] *= w
            elif li[i] == 1 and lj[j] == 1 and m != 0:
                c.data[:, k] = sig[(i, j)] + (c.data[:, k] - sig[(i, j)]) * w * w
    return ts


def test_vnl_rotation_covariance():
    sp = synth.with_pp([synth.C_SP])
    ts = _isotropic_at_origin(synth.add_pp_tables(synth.make_tables(sp, seed=8, with_3c=False), seed=9))
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 4.0, size=(7, 3)) + 10.0
    st = synth.Structure(pos, np.ones(7, np.int32), np.eye(3) * 40, 1)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    st2 = synth.Structure(pos @ q.T, st.imass, st.lattice, 1)
    res = []
    for s in (st, st2):
        o = oracle.System(oracle.Tables(ts), s)
        o.build_lists()
        res.append(o.run_pp())
    for k in ("ppp_j", "ppp_b", "ppc_m"):
        assert np.array_equal(res[0][k], res[1][k])
    perm = [1, 2, 0]
    R = q[np.ix_(perm, perm)]
    D = np.eye(4)
    D[1:, 1:] = R
    A, B = res[0]["VNL"].reshape(-1, 4, 4), res[1]["VNL"].reshape(-1, 4, 4)
    scale = np.abs(A).max()
    assert len(res[0]["ppc_m"]) > 20
    for p in range(len(A)):
        a, b = A[p].T, B[p].T      # column-major blocks
        np.testing.assert_allclose(b, D @ a @ D.T, atol=1e-12 * scale)
