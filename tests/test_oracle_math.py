"""Pins of the oracle that do not depend on the golden vectors: closed forms,
textbook Slater-Koster rotations, polynomial exactness of the interpolators,
brute-force lists and physical invariants.  Each is chosen so that a plausible
slip (transposed block, wrong sign, dropped term, off-by-one window) fails."""
import math

import numpy as np
import pytest

import oracle
from fireball_hip import synth
from fireball_hip.synth import Cell2c, Cell3c, Species, TableSet

RNG = np.random.default_rng(1234)


def _one_species_tables(lssh, rc=5.0):
    sp = Species(14, list(lssh), [rc] * len(lssh))
    return sp, TableSet(species=[sp])


# --------------------------------------------------------------- smoother
def test_smoother_closed_form():
    assert oracle.smoother(0.5, 2.0, 0.8) == 1.0
    assert oracle.smoother(2.5, 2.0, 0.8) == 0.0
    # midpoint of [1.6, 2.0]: ratio = 1/2 -> (1 - 1/4)^2
    assert oracle.smoother(1.8, 2.0, 0.8) == pytest.approx(0.5625, abs=1e-15)


# ---------------------------------------------------------------- epsilon
@pytest.mark.parametrize("seed", range(5))
def test_epsilon_generic_frame(seed):
    rng = np.random.default_rng(seed)
    r1, r2 = rng.normal(size=3), rng.normal(size=3)
    e = oracle.epsilon(r1, r2)
    np.testing.assert_allclose(e.T @ e, np.eye(3), atol=1e-14)
    np.testing.assert_allclose(e[:, 2], r2 / np.linalg.norm(r2), atol=1e-15)
    assert abs(e[:, 1] @ r1) < 1e-13           # y' = z' x r1 is perpendicular to r1
    assert np.linalg.det(e) == pytest.approx(1.0, abs=1e-14)


@pytest.mark.parametrize("r2", [[0, 0, 2.0], [3.0, 0, 0], [0, -1.0, 0], [0, 0, -1.0]])
def test_epsilon_collinear_fallbacks(r2):
    r2 = np.array(r2, float)
    for r1 in (2.5 * r2, np.zeros(3)):
        e = oracle.epsilon(r1, r2)
        np.testing.assert_allclose(e.T @ e, np.eye(3), atol=1e-14)
        np.testing.assert_allclose(e[:, 2], r2 / np.linalg.norm(r2), atol=1e-15)


def test_epsilon_zero_r2_is_identity():
    np.testing.assert_array_equal(oracle.epsilon(np.ones(3), np.zeros(3)), np.eye(3))


# --------------------------------------------- rotation: textbook Slater-Koster
def _direction_frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    return n, oracle.epsilon(rng.normal(size=3), n)


@pytest.mark.parametrize("seed", range(4))
def test_rotate_sp_slater_koster(seed):
    rng = np.random.default_rng(seed)
    sp, ts = _one_species_tables([0, 1])
    tab = oracle.Tables(ts)
    n, eps = _direction_frame(rng)
    ss, spS, psS, ppS, ppP = rng.normal(size=5)
    # molecular frame, orbital order s, py, pz, px; bond along z
    M = np.zeros((4, 4))
    M[0, 0] = ss
    M[0, 2] = spS
    M[2, 0] = psS
    M[1, 1] = M[3, 3] = ppP
    M[2, 2] = ppS
    X = tab.rotate(1, 1, eps, M)
    # Slater-Koster: direction cosines in (y, z, x) orbital order
    c = np.array([n[1], n[2], n[0]])
    ref = np.zeros((4, 4))
    ref[0, 0] = ss
    ref[0, 1:] = spS * c
    ref[1:, 0] = psS * c
    ref[1:, 1:] = (ppS - ppP) * np.outer(c, c) + ppP * np.eye(3)
    np.testing.assert_allclose(X, ref, atol=1e-14)


@pytest.mark.parametrize("seed", range(4))
def test_rotate_sd_slater_koster(seed):
    rng = np.random.default_rng(seed)
    sp, ts = _one_species_tables([0, 2])
    tab = oracle.Tables(ts)
    n, eps = _direction_frame(rng)
    sdS = rng.normal()
    M = np.zeros((6, 6))
    M[0, 3] = sdS          # s - d(z^2) in the molecular frame (d order xy, yz, z2, xz, x2-y2)
    X = tab.rotate(1, 1, eps, M)
    l, m, k = n
    ref = sdS * np.array([math.sqrt(3) * l * m, math.sqrt(3) * m * k, k * k - (l * l + m * m) / 2,
                          math.sqrt(3) * l * k, math.sqrt(3) / 2 * (l * l - m * m)])
    np.testing.assert_allclose(X[0, 1:], ref, atol=1e-14)
    dmat, pmat = oracle.twister(eps)
    np.testing.assert_allclose(dmat.T @ dmat, np.eye(5), atol=1e-13)
    np.testing.assert_allclose(pmat.T @ pmat, np.eye(3), atol=1e-14)


# -------------------------------------------------- 2c Lagrange interpolation
def _cell2c_poly(coefs, nx=64, xmax=5.0):
    x = np.arange(nx) * (xmax / (nx - 1))
    data = np.stack([np.polyval(c, x) for c in coefs], axis=1)
    nME = len(coefs)
    idx = np.arange(1, nME + 1, dtype=np.int32)
    return Cell2c(nME, nx, xmax, idx, idx, np.zeros(nME, np.int32), data)


def test_lagrange_exact_for_degree6_and_edges():
    sp = Species(1, [0, 0, 0], [5.0] * 3)        # three s shells: ME k sits at (k, k)
    ts = TableSet(species=[sp])
    c6 = RNG.normal(size=7)
    c3 = RNG.normal(size=4)
    c7 = RNG.normal(size=8)
    ts.cells2c[(1, 1, 1, 0)] = _cell2c_poly([c6, c3, c7])
    tab = oracle.Tables(ts)
    worst7 = 0.0
    for x in [0.0, 1e-6, 0.03, 0.2, 1.234, 2.5, 4.9, 4.99, 5.0, 5.0 * (1 + 5e-6)]:
        h = tab.getMEs_2c(1, 1, 1, 0, x, 3, 3)
        assert h[0, 0] == pytest.approx(np.polyval(c6, x), abs=1e-10)
        assert h[1, 1] == pytest.approx(np.polyval(c3, x), abs=1e-11)
        worst7 = max(worst7, abs(h[2, 2] - np.polyval(c7, x)))
        assert h[0, 1] == 0.0
    assert worst7 > 1e-12          # degree 7 is not reproduced: a 7-point window is used


def test_lagrange_range_errors():
    sp = Species(1, [0], [5.0])
    ts = TableSet(species=[sp])
    ts.cells2c[(1, 1, 1, 0)] = _cell2c_poly([[1.0, 0.0]])
    tab = oracle.Tables(ts)
    with pytest.raises(oracle.OracleError):
        tab.getMEs_2c(1, 1, 1, 0, -1e-9, 1, 1)
    with pytest.raises(oracle.OracleError):
        tab.getMEs_2c(1, 1, 1, 0, 5.0 * (1 + 2e-5), 1, 1)


# -------------------------------------------------- 3c bivariate interpolation
def _cell3c(fun_per_theta, n=20, xmax=5.0, ymax=4.0, mvalue=(0,)):
    xg = np.arange(n) * (xmax / (n - 1))
    yg = np.arange(n) * (ymax / (n - 1))
    X, Y = np.meshgrid(xg, yg)                 # [iy, ix]
    cells = []
    nME = len(mvalue)
    for f in fun_per_theta:
        data = np.stack([f(X, Y, k) for k in range(nME)], axis=-1)
        idx = np.arange(1, nME + 1, dtype=np.int32)
        cells.append(Cell3c(nME, n, n, xmax, ymax, idx, idx, np.array(mvalue, np.int32), data))
    return cells


def _tables3c(cells):
    sp = Species(1, [0] * len(cells[0].mu), [5.0] * len(cells[0].mu))
    ts = TableSet(species=[sp])
    for ith, c in enumerate(cells, start=1):
        ts.cells3c[(1, 1, 1, 3, 0, ith)] = c
    return oracle.Tables(ts)


def test_bicubic_exact_and_legendre_sum():
    a = RNG.normal(size=(5, 4, 4))
    def bicubic(t):
        return lambda X, Y, k: sum(a[t, i, j] * X ** i * Y ** j for i in range(4) for j in range(4))
    tab = _tables3c(_cell3c([bicubic(t) for t in range(5)]))
    for (x, y) in [(0.01, 0.02), (2.3, 1.7), (4.99, 3.99), (5.0, 4.0), (0.3, 3.9)]:
        for cost in (-1.0, -0.3, 0.0, 0.71, 1.0):
            h = tab.getMEs_3c(1, 1, 1, 3, 0, x, y, 1, 1, cost)
            vals = [sum(a[t, i, j] * x ** i * y ** j for i in range(4) for j in range(4)) for t in range(5)]
            expect = np.polynomial.legendre.legval(cost, vals)
            assert h[0, 0] == pytest.approx(expect, rel=1e-10, abs=1e-10)


def test_method_chosen_from_first_me_first_theta():
    # ME 1 flat (tiny gradient) -> linear 3-point rule for every ME; a linear ME is
    # still exact, a quadratic one is not.  With ME 1 steep, the quadratic is exact.
    def cells(steep):
        def f(X, Y, k):
            if k == 0:
                return (5.0 * X if steep else 1e-5 * X) + 0.0 * Y
            if k == 1:
                return 0.3 + 0.2 * X - 0.1 * Y
            return 0.05 * X * X + 0.02 * Y * Y
        return _cell3c([f] + [lambda X, Y, k: 0.0 * X] * 4, mvalue=(0, 0, 0))
    for steep in (False, True):
        tab = _tables3c(cells(steep))
        x, y = 1.37, 2.11
        h = tab.getMEs_3c(1, 1, 1, 3, 0, x, y, 3, 3, 1.0)
        assert h[1, 1] == pytest.approx(0.3 + 0.2 * x - 0.1 * y, abs=1e-12)
        quad = 0.05 * x * x + 0.02 * y * y
        if steep:
            assert h[2, 2] == pytest.approx(quad, abs=1e-12)
        else:
            assert abs(h[2, 2] - quad) > 1e-4


def test_sint_on_mvalue_one():
    f = lambda X, Y, k: 1.0 + 0.0 * X         # noqa: E731
    z = lambda X, Y, k: 0.0 * X               # noqa: E731
    tab = _tables3c(_cell3c([f, z, z, z, z], mvalue=(0, 1)))
    for cost in (0.0, 0.6, 0.999, 1.0):
        h = tab.getMEs_3c(1, 1, 1, 3, 0, 1.0, 1.0, 2, 2, cost)
        assert h[0, 0] == pytest.approx(1.0, abs=1e-14)
        assert h[1, 1] == pytest.approx(math.sqrt(max(1e-5, 1 - cost * cost)), abs=1e-14)


# ------------------------------------------------------------------ lists
def _brute_lists(st, species):
    """Neighbour pairs by a vectorised all-images search, sorted by (image, atom)."""
    xl = oracle.make_cells(st.lattice, st.icluster)
    rc = np.array([max(species[i - 1].rcutoffA) for i in st.imass])
    out = []
    for ia in range(st.natoms):
        r2 = st.ratom[None, :, :] + xl[:, None, :]              # [img, atom, 3]
        d = np.sqrt(np.sum((r2 - st.ratom[ia]) ** 2, axis=-1))
        ok = d < (rc[ia] + rc[None, :] - 0.01)
        b, j = np.nonzero(ok)
        out.append(sorted(zip(b.tolist(), (j + 1).tolist())))
    return out


def test_make_cells_closed_form():
    lat = np.array([[2.0, 0.1, 0], [0, 3.0, 0.2], [0.3, 0, 5.0]])
    xl = oracle.make_cells(lat)
    assert xl.shape == (729, 3)
    assert np.all(xl[0] == 0)
    ijk = np.rint(np.linalg.solve(lat.T, xl.T).T).astype(int)
    assert len({tuple(v) for v in ijk}) == 729
    assert tuple(ijk[1]) == (-4, -4, -4) and tuple(ijk[2]) == (-3, -4, -4)   # x runs fastest
    assert oracle.make_cells(lat, icluster=1).shape == (1, 3)


@pytest.mark.parametrize("cells", [(1, 1, 1), (1, 1, 2)])
def test_lists_brute_force_and_invariants(cells):
    sp = [synth.SI]
    st = synth.diamond_si(cells, sigma=0.1, seed=3)
    ts = synth.make_tables(sp, with_3c=False)
    o = oracle.System(oracle.Tables(ts), st)
    L = o.build_lists()
    ref = _brute_lists(st, sp)
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    xl = oracle.make_cells(st.lattice)
    for ia in range(st.natoms):
        got = list(zip(L["neigh_b"][off[ia]:off[ia + 1]].tolist(), L["neigh_j"][off[ia]:off[ia + 1]].tolist()))
        assert got == ref[ia]
        for k in range(off[ia], off[ia + 1]):
            ja, b, back = L["neigh_j"][k], L["neigh_b"][k], L["neigh_back"][k]
            kb = off[ja - 1] + back - 1
            assert L["neigh_j"][kb] == ia + 1
            np.testing.assert_allclose(xl[L["neigh_b"][kb]], -xl[b], atol=1e-9)
        assert L["neigh_j"][off[ia] + L["neigh_self"][ia] - 1] == ia + 1
        assert L["neigh_b"][off[ia] + L["neigh_self"][ia] - 1] == 0
    # triplets: mneigh points at (jatom, jbeta - ibeta) in iatom's list; both orders present
    trip = set()
    for t in range(L["ntriplets"]):
        ia, ib, ja, jb, m = (L[k][t] for k in ("com_iatom_j", "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh"))
        k = off[ia - 1] + m - 1
        assert L["neigh_j"][k] == ja
        np.testing.assert_allclose(xl[L["neigh_b"][k]], xl[jb] - xl[ib], atol=1e-9)
        trip.add((ia, ib, ja, jb))
    for (ia, ib, ja, jb) in list(trip)[:200]:
        assert (ja, jb, ia, ib) in trip


# ----------------------------------------------------------- assembled blocks
def _blocks_by_pair(o, L, species, st, mat):
    norb = [sp.norb for sp in species]
    blk = o.blocks(mat)
    out, p, pos = {}, 0, 0
    for ia in range(st.natoms):
        for _ in range(L["neighn"][ia]):
            ja = L["neigh_j"][p]
            nmu, nnu = norb[st.imass[ia] - 1], norb[st.imass[ja - 1] - 1]
            out[p] = blk[pos:pos + nmu * nnu].reshape(nnu, nmu).T
            pos += nmu * nnu
            p += 1
    return out


def test_S_T_symmetry_and_translation_invariance():
    sp = [synth.C_SP, synth.H_SP, synth.O_SP]
    ts = synth.make_tables(sp, seed=5, with_3c=False)
    st = synth.random_box({6: 6, 1: 8, 8: 4}, 1.0, 0.9, 4, sp)
    tab = oracle.Tables(ts)
    o = oracle.System(tab, st)
    L = o.build_lists()
    o.assemble(("S", "T", "DIP", "VNA2C"))
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    for mat in ("S", "T"):
        B = _blocks_by_pair(o, L, sp, st, mat)
        for ia in range(st.natoms):
            for p in range(off[ia], off[ia + 1]):
                if p == off[ia] + L["neigh_self"][ia] - 1:
                    continue
                q = off[L["neigh_j"][p] - 1] + L["neigh_back"][p] - 1
                np.testing.assert_allclose(B[p], B[q].T, atol=1e-13)
    # whole-structure translation: the 2c frame uses absolute positions, the
    # result must not
    st2 = synth.Structure(st.ratom + np.array([0.37, -1.21, 2.05]), st.imass, st.lattice, 0)
    o2 = oracle.System(tab, st2)
    o2.build_lists()
    o2.assemble(("S", "T", "DIP", "VNA2C"))
    for mat in ("S", "T", "VNA", "VNAO"):
        np.testing.assert_allclose(o.blocks(mat), o2.blocks(mat), atol=2e-12)


def test_rotation_covariance_of_p_blocks():
    # rotating a cluster rotates the p part of every 2c block
    sp = [synth.C_SP]
    ts = synth.make_tables(sp, seed=8, with_3c=False)
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 4, size=(6, 3)) + 10.0
    st = synth.Structure(pos, np.ones(6, np.int32), np.eye(3) * 40, 1)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    st2 = synth.Structure(pos @ q.T, st.imass, st.lattice, 1)
    tab = oracle.Tables(ts)
    res = []
    for s in (st, st2):
        o = oracle.System(tab, s)
        L = o.build_lists()
        o.assemble(("S",))
        res.append(_blocks_by_pair(o, L, sp, s, "S"))
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    selfs = {off[i] + L["neigh_self"][i] - 1 for i in range(6)}
    perm = [1, 2, 0]                   # (y, z, x) orbital order
    R = q[np.ix_(perm, perm)]
    for p in res[0]:
        if p == off[0] + L["neigh_self"][0] - 1 or p in selfs:
            continue                   # z = 0: the frame defaults to the crystal z axis
        A, B = res[0][p], res[1][p]
        np.testing.assert_allclose(B[1:, 1:], R @ A[1:, 1:] @ R.T, atol=1e-12)
        np.testing.assert_allclose(B[0, 1:], R @ A[0, 1:], atol=1e-12)
        assert B[0, 0] == pytest.approx(A[0, 0], abs=1e-13)
