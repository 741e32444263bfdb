"""The force restatement of the C oracle (Dassemble_S/T/dipole_z/vna_2c/vna_3c
of k.DASSEMBLERS) against the reference's own Fortran (golden fixtures made by
tests/golden/make_golden.py), and against the mathematics: every derivative
primitive against central differences of the value primitive it differentiates,
the Dblocks against differences of the pair blocks, and the assembled forces
against -dE/dR of E = sum_pairs rho . (Vna + Vna_ontop) on a cluster."""
import numpy as np
import pytest

import golden_cases
import oracle
from fireball_hip import synth

DKEYS = ("DS", "DT", "DDIP")
FKEYS = ("vna_atom", "vna_ontop", "f3naa", "f3nab", "f3nac")


@pytest.fixture(scope="module")
def force_runs(golden):
    out = {}
    for name in golden_cases.NAMES + golden_cases.NAMES_D:
        g = golden(name)
        gf = golden(name + "_forces")
        ts, st = golden_cases.inputs(g)
        o = oracle.System(oracle.Tables(ts), st)
        lists = o.run_all()
        rho = synth.denmat(ts.species, st.imass, lists["neighn"], lists["neigh_j"], lists["neigh_back"],
                           seed=int(gf["denmat_seed"]))
        out[name] = (gf, o.run_forces(rho))
    return out


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_forces_match_reference(force_runs, name):
    gf, o = force_runs[name]
    for k in DKEYS:
        n = len(gf[k])
        np.testing.assert_allclose(o[k][:n], gf[k], rtol=0, atol=1e-13, err_msg=k)
        assert abs(np.sum(o[k]) - gf[k + "_sum"]) <= 1e-12 * max(1.0, gf[k + "_sumabs"]), k
        assert abs(np.sum(o[k] ** 2) - gf[k + "_sumsq"]) <= 1e-12 * max(1.0, gf[k + "_sumsq"]), k
    for k in FKEYS:
        np.testing.assert_allclose(o[k], gf[k], rtol=0, atol=1e-12, err_msg=k)


def test_charged_fixtures_differ_from_neutral(force_runs):
    # si64 and si64_dq share geometry and density: the dQ terms must show
    a, b = force_runs["si64"][1], force_runs["si64_dq"][1]
    assert np.max(np.abs(a["vna_atom"] - b["vna_atom"])) > 1e-4
    assert np.max(np.abs(a["f3naa"] - b["f3naa"])) > 1e-5
    np.testing.assert_array_equal(a["DS"], b["DS"])


# ---------------------------------------------------------------- primitives

def test_Dsmoother_is_derivative():
    rend, xs = 5.0, 0.8
    for r in (0.5, 4.1, 4.5, 4.99, 5.5):
        h = 1e-6
        fd = (oracle.smoother(r + h, rend, xs) - oracle.smoother(r - h, rend, xs)) / (2 * h)
        assert abs(oracle.Dsmoother(r, rend, xs) - fd) < 1e-7
    assert oracle.Dsmoother(3.0, rend, xs) == 0.0 and oracle.Dsmoother(6.0, rend, xs) == 0.0


def _sighat(r1, r2):
    z = np.linalg.norm(r2 - r1)
    return (r2 - r1) / z, z


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_Depsilon_2c_is_derivative(seed):
    # d eps(r2, sighat(r1, r2)) / d r1 with r2 held (b.FUNCTIONS/Depsilon_2c.f90)
    rng = np.random.default_rng(seed)
    r1, r2 = rng.normal(size=3) * 2, rng.normal(size=3) * 2
    s, z = _sighat(r1, r2)
    deps, warned = oracle.Depsilon_2c(r1, r2, z, oracle.epsilon(r2, s))
    assert not warned
    h = 1e-6
    for ix in range(3):
        e = np.zeros(3)
        e[ix] = h
        fd = (oracle.epsilon(r2, _sighat(r1 + e, r2)[0]) - oracle.epsilon(r2, _sighat(r1 - e, r2)[0])) / (2 * h)
        np.testing.assert_allclose(deps[ix], fd, atol=2e-9)


def test_Depsilon_2c_guards():
    r1 = np.array([0.0, 0.0, 1.0])
    deps, warned = oracle.Depsilon_2c(r1, r1, 0.0, np.eye(3))
    assert not warned and not deps.any()                       # the atom case, z < 1e-5
    r2 = np.array([0.0, 0.0, 2.5])                             # r1 x r2 = 0: colinear with the origin
    deps, warned = oracle.Depsilon_2c(r1, r2, 1.5, oracle.epsilon(r2, _sighat(r1, r2)[0]))
    assert warned and not deps.any()


def _frame3(ratm, r1, r2):
    r21 = r2 - r1
    z = np.linalg.norm(r21)
    rnabc = ratm - (r1 + r21 / 2.0)
    x = np.linalg.norm(rnabc)
    return oracle.epsilon(rnabc / x, r21 / z), r21, z, rnabc


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_Depsilon_3c_is_derivative(seed):
    rng = np.random.default_rng(seed)
    ratm, r1, r2 = (rng.normal(size=3) * 2 for _ in range(3))
    eps, r21, z, rnabc = _frame3(ratm, r1, r2)
    dA, dB, warned = oracle.Depsilon_3c(r1, r2, r21, z, ratm, rnabc, eps)
    assert not warned
    h = 1e-6
    for ix in range(3):
        e = np.zeros(3)
        e[ix] = h
        fa = (_frame3(ratm + e, r1, r2)[0] - _frame3(ratm - e, r1, r2)[0]) / (2 * h)
        fb = (_frame3(ratm, r1 + e, r2)[0] - _frame3(ratm, r1 - e, r2)[0]) / (2 * h)
        np.testing.assert_allclose(dA[ix], fa, atol=2e-9)
        np.testing.assert_allclose(dB[ix], fb, atol=2e-9)
    assert not dA[:, :, 2].any()       # sighat does not depend on the third atom


def test_Depsilon_3c_colinear_guard():
    r1, r2 = np.zeros(3), np.array([0.0, 0.0, 2.0])
    ratm = np.array([0.0, 0.0, 3.5])   # on the bond axis: r21 x rnabc = 0
    r21 = r2 - r1
    rnabc = ratm - (r1 + r21 / 2)
    dA, dB, warned = oracle.Depsilon_3c(r1, r2, r21, 2.0, ratm, rnabc, np.eye(3))
    assert warned and not dA.any() and not dB.any()


@pytest.fixture(scope="module")
def sp_tables():
    sp = [synth.C_SP, synth.H_SP]
    ts = synth.make_tables(sp, seed=5)
    return ts, oracle.Tables(ts)


def test_getDMEs_2c_derivative(sp_tables):
    ts, T = sp_tables
    h = 1e-6
    for (a, b, it, isub), c in list(ts.cells2c.items())[:12]:
        nmu, nnu = ts.species[a - 1].norb, ts.species[b - 1].norb
        for x in (0.3, 1.37, 2.91, 4.2):
            hm, dh = T.getDMEs_2c(a, b, it, isub, x, nmu, nnu)
            np.testing.assert_array_equal(hm, T.getMEs_2c(a, b, it, isub, x, nmu, nnu))
            fd = (T.getMEs_2c(a, b, it, isub, x + h, nmu, nnu) - T.getMEs_2c(a, b, it, isub, x - h, nmu, nnu)) / (2 * h)
            np.testing.assert_allclose(dh, fd, atol=1e-7 * max(1.0, np.max(np.abs(fd))))


def test_getDMEs_2c_at_zero_distance(sp_tables):
    # x < 1e-5 returns grid point 1 with a zero derivative (M_Fdata_2c.f90:508-517)
    ts, T = sp_tables
    (a, b, it, isub), c = next(iter(ts.cells2c.items()))
    nmu, nnu = ts.species[a - 1].norb, ts.species[b - 1].norb
    hm, dh = T.getDMEs_2c(a, b, it, isub, 0.0, nmu, nnu)
    assert not dh.any()
    vals = sorted(hm[hm != 0.0].tolist())
    assert vals == sorted(np.asarray(c.data)[0][np.asarray(c.data)[0] != 0.0].tolist())


def _ramp_first_cell(ts):
    """Give ME 1 of theta cell 1 a slope so that every 3c cell takes the
    bicubic branch (e.FDATA/M_Fdata_3c.f90:636-643 switches on its gradient)."""
    for key, c in ts.cells3c.items():
        if key[5] == 1:
            iy, ix = np.meshgrid(np.arange(c.ny), np.arange(c.nx), indexing="ij")
            c.data = np.array(c.data, copy=True)
            c.data[:, :, 0] += 0.002 * (ix + iy)
    return ts


def test_getDMEs_3c_bicubic_derivatives():
    sp = [synth.C_SP, synth.H_SP]
    ts = _ramp_first_cell(synth.make_tables(sp, seed=5))
    T = oracle.Tables(ts)
    h = 1e-6
    for (a, b, k) in ((1, 1, 1), (1, 2, 1), (2, 1, 2)):
        nmu, nnu = ts.species[a - 1].norb, ts.species[b - 1].norb
        for x, y, ct in ((0.7, 1.1, 0.3), (2.3, 1.9, -0.8), (3.9, 4.4, 0.95)):
            hm, dp, dx, dy = T.getDMEs_3c(a, b, k, 3, 0, x, y, nmu, nnu, ct)
            f = lambda X, Y, C: T.getMEs_3c(a, b, k, 3, 0, X, Y, nmu, nnu, C)  # noqa: E731
            np.testing.assert_allclose(hm, f(x, y, ct), rtol=0, atol=1e-14)
            np.testing.assert_allclose(dx, (f(x + h, y, ct) - f(x - h, y, ct)) / (2 * h), atol=1e-7)
            np.testing.assert_allclose(dy, (f(x, y + h, ct) - f(x, y - h, ct)) / (2 * h), atol=1e-7)
            np.testing.assert_allclose(dp, (f(x, y, ct + h) - f(x, y, ct - h)) / (2 * h), atol=1e-7)


def test_getDMEs_3c_three_point_quirk(sp_tables):
    # In the 3-point branch the x and y derivatives of every matrix element are
    # the slopes of ME 1 of theta cell 1 times the Legendre sum (:641-642).
    ts, T = sp_tables
    c1 = ts.cells3c[(1, 1, 1, 3, 0, 1)]
    dxg, dyg = c1.xmax / (c1.nx - 1), c1.ymax / (c1.ny - 1)
    found = False
    for iy in range(2, c1.ny - 2):
        for ix in range(2, c1.nx - 2):
            gx = (c1.data[iy - 1, ix, 0] - c1.data[iy - 1, ix - 1, 0]) / dxg
            gy = (c1.data[iy, ix - 1, 0] - c1.data[iy - 1, ix - 1, 0]) / dyg
            if gx * gx + gy * gy < 1e-6 and abs(gx) > 1e-9:
                found = True
                break
        if found:
            break
    assert found
    x, y, ct = (ix - 1 + 0.4) * dxg, (iy - 1 + 0.3) * dyg, 0.25
    hm, dp, dx, dy = T.getDMEs_3c(1, 1, 1, 3, 0, x, y, 4, 4, ct)
    p = np.array([1.0, ct, 0.5 * (3 * ct * ct - 1), 0.5 * ct * (5 * ct * ct - 3),
                  0.125 * (3 - ct * ct * (30 - 35 * ct * ct))])
    sint = np.sqrt(max(1e-5, 1 - ct * ct))
    mv = {(int(m), int(n)): int(v) for m, n, v in zip(c1.mu, c1.nu, c1.mvalue)}
    for (m, n), v in mv.items():
        scale = sint if v == 1 else 1.0
        assert dx[m - 1, n - 1] == pytest.approx(gx * p.sum() * scale, rel=1e-12)
        assert dy[m - 1, n - 1] == pytest.approx(gy * p.sum() * scale, rel=1e-12)


def test_Drotate_product_rule(sp_tables):
    # d/dr1 of rotate(eps(r1), M(r1)) with M linear in r1
    ts, T = sp_tables
    rng = np.random.default_rng(9)
    r1, r2 = rng.normal(size=3), rng.normal(size=3) + 2.0
    M0, M1 = rng.normal(size=(4, 4)), rng.normal(size=(4, 4))
    a = rng.normal(size=3)

    def frame(r):
        s, z = _sighat(r, r2)
        return oracle.epsilon(r2, s), z

    eps, z = frame(r1)
    deps, _ = oracle.Depsilon_2c(r1, r2, z, eps)
    DM = np.einsum("k,ij->kij", a, M1)
    got = T.Drotate(1, 1, eps, deps, M0 + (a @ r1) * M1, DM)
    h = 1e-6
    for ix in range(3):
        e = np.zeros(3)
        e[ix] = h
        Xp = T.rotate(1, 1, frame(r1 + e)[0], M0 + (a @ (r1 + e)) * M1)
        Xm = T.rotate(1, 1, frame(r1 - e)[0], M0 + (a @ (r1 - e)) * M1)
        np.testing.assert_allclose(got[ix], (Xp - Xm) / (2 * h), atol=2e-8)


# ---------------------------------------------------------------- assembled

def _cluster(charged):
    sp3 = [synth.C_SP, synth.H_SP, synth.O_SP]
    ts = _ramp_first_cell(synth.make_tables(sp3, seed=99))
    rng = np.random.default_rng(5)
    pos = []
    while len(pos) < 7:
        p = rng.uniform(0, 3.2, size=3)
        if all(np.linalg.norm(p - q) > 1.1 for q in pos):
            pos.append(p)
    pos = np.array(pos) + 10.0
    imass = np.array([1, 2, 3, 1, 2, 1, 3], np.int32)
    dq = None
    if charged:
        dq = np.zeros((7, 4))
        r2 = np.random.default_rng(3)
        for ia in range(7):
            dq[ia, :sp3[imass[ia] - 1].nssh] = r2.uniform(-0.1, 0.1, size=sp3[imass[ia] - 1].nssh)
    return ts, pos, imass, dq


@pytest.mark.parametrize("charged", [False, True])
def test_forces_are_minus_energy_gradient(charged):
    """Pieces of the force (vna_atom, vna_ontop, f3na*) against central
    differences of the matching pieces of E = sum_pairs rho . H, with the pair
    forces distributed as build_forces does (c.SYSTEM/M_build_forces.f90:326-343)."""
    ts, pos, imass, dq = _cluster(charged)
    T = oracle.Tables(ts)
    lat = np.diag([40.0] * 3)
    o = oracle.System(T, synth.Structure(pos, imass, lat, icluster=1, dQ=dq))
    r = o.run_all()
    n = len(pos)
    rho = synth.denmat(ts.species, imass, r["neighn"], r["neigh_j"], r["neigh_back"])
    f = o.run_forces(rho)
    norb = np.array([sp.norb for sp in ts.species])
    pa = np.repeat(np.arange(n), r["neighn"])
    boff = np.concatenate([[0], np.cumsum(norb[imass[pa] - 1] * norb[imass[r["neigh_j"] - 1] - 1])])
    noff = np.concatenate([[0], np.cumsum(r["neighn"])])
    selfel = np.zeros(len(rho), bool)
    for ia in range(n):
        p = noff[ia] + r["neigh_self"][ia] - 1
        selfel[boff[p]:boff[p + 1]] = True

    def parts(P):
        rr = oracle.System(T, synth.Structure(P, imass, lat, icluster=1, dQ=dq)).run_all()
        assert np.array_equal(rr["com_mneigh"], r["com_mneigh"])
        return np.array([rho[selfel] @ rr["VNA"][selfel], rho[~selfel] @ rr["VNA"][~selfel], rho @ rr["VNAO"]])

    h = 1e-5
    G = np.zeros((3, n, 3))
    for ia in range(n):
        for k in range(3):
            P = pos.copy()
            P[ia, k] += h
            ep = parts(P)
            P[ia, k] -= 2 * h
            G[:, ia, k] = -(ep - parts(P)) / (2 * h)
    Fa, Fo = np.zeros((n, 3)), np.zeros((n, 3))
    for p in range(r["npairs"]):
        Fa[pa[p]] += f["vna_atom"][p]
        Fa[r["neigh_j"][p] - 1] -= f["vna_atom"][p]
        Fo[pa[p]] += f["vna_ontop"][p]
        Fo[r["neigh_j"][p] - 1] -= f["vna_ontop"][p]
    F3 = f["f3naa"] + f["f3nab"] + f["f3nac"]
    for got, ref in ((Fa, G[0]), (F3, G[1]), (Fo, G[2])):
        assert np.max(np.abs(got)) > 1e-2
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-7 * max(1.0, np.max(np.abs(ref))))


def test_dblocks_are_pair_block_derivatives(golden):
    # Dblock(ix) of S, T and dipole_z = d block / d r1 (r2 held) for sampled pairs of si64
    g = golden("si64")
    ts, st = golden_cases.inputs(g)
    T = oracle.Tables(ts)
    o = oracle.System(T, st)
    r = o.run_all()
    rho = synth.denmat(ts.species, st.imass, r["neighn"], r["neigh_j"], r["neigh_back"])
    f = o.run_forces(rho)
    xl = oracle.make_cells(st.lattice, st.icluster)
    noff = np.concatenate([[0], np.cumsum(r["neighn"])])
    h = 1e-6
    for p in (1, 17, 40, 333, 1200, 2239):
        ia = int(np.searchsorted(noff, p, side="right") - 1)
        ja, mb = int(r["neigh_j"][p]) - 1, int(r["neigh_b"][p])
        r1, r2 = st.ratom[ia], st.ratom[ja] + xl[mb]
        for key, it in (("DS", 1), ("DT", 2), ("DDIP", 10)):
            D = f[key][48 * p:48 * (p + 1)].reshape(4, 4, 3).transpose(2, 1, 0)

            def block(rr):
                s, z = _sighat(rr, r2)
                return T.rotate(1, 1, oracle.epsilon(r2, s), T.getMEs_2c(1, 1, it, 0, z, 4, 4))

            for ix in range(3):
                e = np.zeros(3)
                e[ix] = h
                np.testing.assert_allclose(D[ix], (block(r1 + e) - block(r1 - e)) / (2 * h), atol=1e-8,
                                           err_msg=f"{key} pair {p}")
