"""Synthetic species, Fdata tables and structures for the 2c/3c assembly workloads.

Everything here is plain numpy and deterministic for a given seed.  The tables
follow the layout the reference reads from disk:

* 2c cells: ``data[ix, iME]`` on a uniform grid ``x = ix*dx``, ``dx = xmax/(nx-1)``
  (e.FDATA/M_Fdata_2c.f90:212-232), matrix-element order from ``make_munu`` /
  ``make_munu_atom`` (c.SYSTEM/M_integrals_2c.f90:232-340, 379-488).
* 3c cells: ``data[iy, ix, iME]`` (the reference stores ``Fdata_3c(ix, iy, iME)``
  with ix fastest, e.FDATA/M_Fdata_3c.f90:227-231); x is the distance of the
  third atom from the bond centre, y the bond length.  Matrix-element order from
  ``make_munu_3c`` (c.SYSTEM/M_integrals_3c.f90:230-587).

Interaction ids and subtypes (include/interactions_2c.h, interactions_3c.h):
overlap 1, kinetic 2, vna_ontopL 3 (isorp 0..nssh(left)), vna_ontopR 4 (isorp
0..nssh(right)), vna_atom 5 (isorp 0..nssh(right)), dipole_z 10; 3c bcna 3 with
isorp 0..nssh(third atom) and five Legendre theta cells.

The 2c tables are built so that the physical symmetries hold: swapping the two
centres multiplies a two-centre function by (-1)^(l1+l2) (inversion through the
bond centre), so S and T come out symmetric and ``vna_ontopL(A,B)`` is the
mirror of ``vna_ontopR(B,A)``.  These give the invariants the parity tests use.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

P_ABOHR = 0.529177          # include/constants.h:2
P_EQ2 = 14.39975            # include/constants.h:4
P_MAXTHETA = 5              # include/interactions_3c.h:4

# interaction ids (include/interactions_2c.h:5-25, interactions_3c.h:6-8)
P_OVERLAP, P_KINETIC, P_VNA_ONTOPL, P_VNA_ONTOPR, P_VNA_ATOM, P_DIPOLE_Z = 1, 2, 3, 4, 5, 10
P_VNL = 6
P_RHO_ONTOPL, P_RHO_ONTOPR, P_RHO_ATOM = 15, 16, 17
P_RHOS_ONTOPL, P_RHOS_ONTOPR, P_RHOS_ATOM = 19, 20, 21
P_RHO_3C, P_RHOS_3C, P_BCNA = 1, 2, 3

MASS = {1: 1.008, 3: 6.94, 6: 12.011, 7: 14.007, 8: 15.999, 14: 28.0855, 26: 55.845}
SYMBOL = {1: "H", 6: "C", 7: "N", 8: "O", 14: "Si", 26: "Fe"}
AMU_G = 1.66053906660e-24


@dataclass
class Species:
    nZ: int
    lssh: list            # angular momentum per shell
    rcutoff: list         # shell cutoffs in bohr (as read from info.dat)
    # pseudopotential shells (info.dat: nssh_PP, lssh_PP, rcutoff_PP) and their
    # Kleinman-Bylander coefficients (clPP.ZZ.dat); empty: no PP
    lssh_pp: list = field(default_factory=list)
    rcutoff_pp: float = 0.0
    cl_pp: list = field(default_factory=list)

    @property
    def nssh_pp(self):
        return len(self.lssh_pp)

    @property
    def norb_pp(self):
        return sum(2 * l + 1 for l in self.lssh_pp)

    @property
    def nssh(self):
        return len(self.lssh)

    @property
    def norb(self):
        return sum(2 * l + 1 for l in self.lssh)

    @property
    def rcutoffA(self):
        # c.SYSTEM/M_species.f90:1177-1179: rcutoffA = rcutoff*P_abohr (FP64)
        return [float(np.float64(r) * np.float64(P_ABOHR)) for r in self.rcutoff]

    @property
    def symbol(self):
        return SYMBOL.get(self.nZ, "X")


@dataclass
class Cell2c:
    nME: int
    nx: int
    xmax: float
    mu: np.ndarray
    nu: np.ndarray
    mvalue: np.ndarray
    data: np.ndarray       # (nx, nME) float64


@dataclass
class Cell3c:
    nME: int
    nx: int
    ny: int
    xmax: float
    ymax: float
    mu: np.ndarray
    nu: np.ndarray
    mvalue: np.ndarray
    data: np.ndarray       # (ny, nx, nME) float64


@dataclass
class TableSet:
    species: list                                   # list[Species], index = ispecies-1
    cells2c: dict = field(default_factory=dict)     # (is, js, itype, isub) -> Cell2c (1-based species)
    cells3c: dict = field(default_factory=dict)     # (is, js, ks, itype, isub, itheta) -> Cell3c


@dataclass
class Structure:
    ratom: np.ndarray      # (N, 3) float64, Angstrom, as handed to the assembler
    imass: np.ndarray      # (N,) int32, 1-based species index
    lattice: np.ndarray    # (3, 3) float64, rows a1, a2, a3
    icluster: int = 0
    dQ: np.ndarray | None = None   # (N, MAXSH) float64 shell charges, None = neutral
    name: str = ""

    @property
    def natoms(self):
        return int(self.ratom.shape[0])


# ---------------------------------------------------------------------------
# matrix-element layouts
# ---------------------------------------------------------------------------
def munu_2c(sp_i: Species, sp_j: Species):
    """(mu, nu, mvalue) of the non-zero 2c MEs, c.SYSTEM/M_integrals_2c.f90:297-335."""
    mu, nu, mv, shells = [], [], [], []
    n1 = 0
    for issh, l1 in enumerate(sp_i.lssh):
        n1 += l1 + 1
        n2 = 0
        for jssh, l2 in enumerate(sp_j.lssh):
            n2 += l2 + 1
            for m in range(-min(l1, l2), min(l1, l2) + 1):
                mu.append(n1 + m)
                nu.append(n2 + m)
                mv.append(0)
                shells.append((issh, jssh, m))
            n2 += l2
        n1 += l1
    return (np.array(mu, np.int32), np.array(nu, np.int32), np.array(mv, np.int32), shells)


def munu_3c(sp_i: Species, sp_j: Species):
    """(mu, nu, mvalue) of the non-zero 3c MEs, c.SYSTEM/M_integrals_3c.f90:299-583."""
    mu, nu, mv = [], [], []
    n1 = 0
    for l1 in sp_i.lssh:                       # delta m = 0
        n1 += l1 + 1
        n2 = 0
        for l2 in sp_j.lssh:
            n2 += l2 + 1
            for m in range(-min(l1, l2), min(l1, l2) + 1):
                mu.append(n1 + m); nu.append(n2 + m); mv.append(0)
            n2 += l2
        n1 += l1
    n1 = 0
    for l1 in sp_i.lssh:                       # delta m = 1
        n1 += l1 + 1
        n2 = 0
        for l2 in sp_j.lssh:
            n2 += l2 + 1
            if l1 == 0 and l2 != 0:
                mu.append(n1); nu.append(n2 + 1); mv.append(1)
            if l1 == 1:
                mu.append(n1 + 1); nu.append(n2); mv.append(1)
                if l2 != 0:
                    mu.append(n1); nu.append(n2 + 1); mv.append(1)
                if l2 == 2:
                    mu.append(n1 + 1); nu.append(n2 + 2); mv.append(1)
                    mu.append(n1 - 1); nu.append(n2 - 2); mv.append(1)
            if l1 == 2:
                mu.append(n1 + 1); nu.append(n2); mv.append(1)
                if l2 != 0:
                    mu.append(n1); nu.append(n2 + 1); mv.append(1)
                    mu.append(n1 - 2); nu.append(n2 - 1); mv.append(1)
                    mu.append(n1 + 2); nu.append(n2 + 1); mv.append(1)
                if l2 == 2:
                    mu.append(n1 + 1); nu.append(n2 + 2); mv.append(1)
                    mu.append(n1 - 1); nu.append(n2 - 2); mv.append(1)
            n2 += l2
        n1 += l1
    n1 = 0
    for l1 in sp_i.lssh:                       # delta m = 2
        n1 += l1 + 1
        n2 = 0
        for l2 in sp_j.lssh:
            n2 += l2 + 1
            if l1 == 2:
                mu.append(n1 + 2); nu.append(n2); mv.append(2)
            if l2 == 2:
                mu.append(n1); nu.append(n2 + 2); mv.append(2)
            n2 += l2
        n1 += l1
    return np.array(mu, np.int32), np.array(nu, np.int32), np.array(mv, np.int32)


# ---------------------------------------------------------------------------
# synthetic tables
# ---------------------------------------------------------------------------
XMAX_BOHR = 10.8


def _grid_bohr(n, xmax_bohr=XMAX_BOHR):
    return np.arange(n, dtype=np.float64) * (xmax_bohr / (n - 1))


def make_tables(species: list, seed: int = 2024, nx2c: int = 256, n3c: int = 32,
                with_3c: bool = True) -> TableSet:
    """Synthetic 2c and 3c Fdata for ``species`` (list of Species, 1-based ids).

    2c: every ME is ``A*exp(-a r)*(1-(r/rc)^2)^3`` on 0..10.8 bohr with
    A ~ U(-2, 2) eV, a ~ U(0.5, 1.5)/bohr, rc = 10.8 bohr; 3c: every Legendre
    coefficient is ``A*exp(-a x)*exp(-b y)`` with A ~ U(-0.5, 0.5) eV,
    a, b ~ U(0.5, 1.5)/bohr on a 32x32 grid over 0..10.8 bohr.  Grids are
    stored in Angstrom (xmax = 10.8*P_abohr).
    """
    rng = np.random.default_rng(seed)
    ts = TableSet(species=list(species))
    ns = len(species)
    xmaxA = float(np.float64(XMAX_BOHR) * np.float64(P_ABOHR))
    r = _grid_bohr(nx2c)
    env = (1.0 - (r / XMAX_BOHR) ** 2) ** 3
    params = {}

    def radial(key):
        if key not in params:
            A = rng.uniform(-2.0, 2.0)
            a = rng.uniform(0.5, 1.5)
            params[key] = A * np.exp(-a * r) * env
        return params[key]

    def parity(l1, l2):
        return -1.0 if (l1 + l2) % 2 else 1.0

    for a_ in range(1, ns + 1):
        for b_ in range(1, ns + 1):
            spA, spB = species[a_ - 1], species[b_ - 1]
            mu, nu, mv, shells = munu_2c(spA, spB)
            nME = len(mu)

            def two_center(kind, extra=()):
                data = np.empty((nx2c, nME))
                for k, (i, j, m) in enumerate(shells):
                    li, lj = spA.lssh[i], spB.lssh[j]
                    if (a_, i) <= (b_, j):
                        data[:, k] = radial((kind, a_, i, b_, j, abs(m)) + extra)
                    else:
                        sgn = parity(li, lj)
                        if kind == "dip":
                            sgn = -sgn
                        data[:, k] = sgn * radial((kind, b_, j, a_, i, abs(m)) + extra)
                return data

            for itype, kind in ((P_OVERLAP, "S"), (P_KINETIC, "T"), (P_DIPOLE_Z, "dip")):
                ts.cells2c[(a_, b_, itype, 0)] = Cell2c(nME, nx2c, xmaxA, mu, nu, mv,
                                                        two_center(kind))
            # ontop: potential on the left atom (A, shell k) or on the right atom (B, shell k).
            for isorp in range(0, spA.nssh + 1):
                data = np.empty((nx2c, nME))
                for k, (i, j, m) in enumerate(shells):
                    data[:, k] = radial(("onP", a_, isorp, a_, i, b_, j, abs(m)))
                ts.cells2c[(a_, b_, P_VNA_ONTOPL, isorp)] = Cell2c(nME, nx2c, xmaxA, mu, nu, mv, data)
            for isorp in range(0, spB.nssh + 1):
                data = np.empty((nx2c, nME))
                for k, (i, j, m) in enumerate(shells):
                    li, lj = spA.lssh[i], spB.lssh[j]
                    # <A i|V_B|B j> mirrors <B j|V_B|A i> with the inversion parity
                    data[:, k] = parity(li, lj) * radial(("onP", b_, isorp, b_, j, a_, i, abs(m)))
                ts.cells2c[(a_, b_, P_VNA_ONTOPR, isorp)] = Cell2c(nME, nx2c, xmaxA, mu, nu, mv, data)
            # atom: both orbitals on A, potential of B (shell isorp); symmetric in (i, j)
            mua, nua, mva, shells_a = munu_2c(spA, spA)
            for isorp in range(0, spB.nssh + 1):
                data = np.empty((nx2c, len(mua)))
                for k, (i, j, m) in enumerate(shells_a):
                    data[:, k] = radial(("atom", a_, b_, isorp, min(i, j), max(i, j), abs(m)))
                ts.cells2c[(a_, b_, P_VNA_ATOM, isorp)] = Cell2c(len(mua), nx2c, xmaxA, mua, nua, mva, data)

    if with_3c:
        xg = _grid_bohr(n3c)
        for a_ in range(1, ns + 1):
            for b_ in range(1, ns + 1):
                mu, nu, mv = munu_3c(species[a_ - 1], species[b_ - 1])
                nME = len(mu)
                for c_ in range(1, ns + 1):
                    for isorp in range(0, species[c_ - 1].nssh + 1):
                        for itheta in range(1, P_MAXTHETA + 1):
                            A = rng.uniform(-0.5, 0.5, size=nME)
                            ax = rng.uniform(0.5, 1.5, size=nME)
                            by = rng.uniform(0.5, 1.5, size=nME)
                            data = (A[None, None, :] * np.exp(-ax[None, None, :] * xg[None, :, None])
                                    * np.exp(-by[None, None, :] * xg[:, None, None]))
                            ts.cells3c[(a_, b_, c_, P_BCNA, isorp, itheta)] = Cell3c(
                                nME, n3c, n3c, xmaxA, xmaxA, mu, nu, mv, np.ascontiguousarray(data))
    return ts


def munu_S(sp_i: Species, sp_j: Species):
    """(mu, nu, mvalue) of the shell-averaged ("spherical") tables: one element
    per shell pair, issh major, mu = issh, nu = jssh, mvalue 0
    (make_munuS / make_munuS_atom, c.SYSTEM/M_integrals_2c.f90:521-755, and
    make_munuS_3c, c.SYSTEM/M_integrals_3c.f90:623-753)."""
    mu, nu = [], []
    for i in range(sp_i.nssh):
        for j in range(sp_j.nssh):
            mu.append(i + 1)
            nu.append(j + 1)
    n = len(mu)
    return np.array(mu, np.int32), np.array(nu, np.int32), np.zeros(n, np.int32)


def add_rho_tables(ts: TableSet, seed: int = 4048, with_3c: bool = True) -> TableSet:
    """Add the McWEDA density tables to ``ts`` (own random stream, so the cells
    already in ``ts`` do not change):

    * 2c rho_ontopL 15 (isorp 1..nssh(left)), rho_ontopR 16 (isorp
      1..nssh(right)), rho_atom 17 (isorp 1..nssh(right), both orbitals on the
      left atom), in the orbital layout of make_munu / make_munu_atom;
    * 2c rhoS_ontopL 19, rhoS_ontopR 20, rhoS_atom 21, one element per shell
      pair (make_munuS, make_munuS_atom);
    * 3c rho_3c 1 (make_munu_3c layout) and rhoS_3c 2 (make_munuS_3c), isorp
      1..nssh(third atom), five Legendre theta cells.

    Radial shapes as in make_tables; the ontop pairs are mirrored across the
    two centres the way make_tables mirrors the Vna ontop tables, so the 2c
    density blocks satisfy rho(j, i) = rho(i, j)^T."""
    rng = np.random.default_rng(seed)
    species = ts.species
    ns = len(species)
    any2 = next(iter(ts.cells2c.values()))
    nx2c = any2.nx
    xmaxA = any2.xmax
    r = _grid_bohr(nx2c)
    env = (1.0 - (r / XMAX_BOHR) ** 2) ** 3
    params = {}

    def radial(key):
        if key not in params:
            A = rng.uniform(-2.0, 2.0)
            a = rng.uniform(0.5, 1.5)
            params[key] = A * np.exp(-a * r) * env
        return params[key]

    def parity(l1, l2):
        return -1.0 if (l1 + l2) % 2 else 1.0

    for a_ in range(1, ns + 1):
        for b_ in range(1, ns + 1):
            spA, spB = species[a_ - 1], species[b_ - 1]
            mu, nu, mv, shells = munu_2c(spA, spB)
            nME = len(mu)
            for isorp in range(1, spA.nssh + 1):
                data = np.empty((nx2c, nME))
                for k, (i, j, m) in enumerate(shells):
                    data[:, k] = radial(("rhoP", a_, isorp, a_, i, b_, j, abs(m)))
                ts.cells2c[(a_, b_, P_RHO_ONTOPL, isorp)] = Cell2c(nME, nx2c, xmaxA, mu, nu, mv, data)
            for isorp in range(1, spB.nssh + 1):
                data = np.empty((nx2c, nME))
                for k, (i, j, m) in enumerate(shells):
                    li, lj = spA.lssh[i], spB.lssh[j]
                    data[:, k] = parity(li, lj) * radial(("rhoP", b_, isorp, b_, j, a_, i, abs(m)))
                ts.cells2c[(a_, b_, P_RHO_ONTOPR, isorp)] = Cell2c(nME, nx2c, xmaxA, mu, nu, mv, data)
            mua, nua, mva, shells_a = munu_2c(spA, spA)
            for isorp in range(1, spB.nssh + 1):
                data = np.empty((nx2c, len(mua)))
                for k, (i, j, m) in enumerate(shells_a):
                    data[:, k] = radial(("rhoA", a_, b_, isorp, min(i, j), max(i, j), abs(m)))
                ts.cells2c[(a_, b_, P_RHO_ATOM, isorp)] = Cell2c(len(mua), nx2c, xmaxA, mua, nua, mva, data)
            # shell-averaged tables: functions of the shell pair only
            smu, snu, smv = munu_S(spA, spB)
            for isorp in range(1, spA.nssh + 1):
                data = np.empty((nx2c, len(smu)))
                for k in range(len(smu)):
                    data[:, k] = radial(("rhoSP", a_, isorp, a_, smu[k] - 1, b_, snu[k] - 1))
                ts.cells2c[(a_, b_, P_RHOS_ONTOPL, isorp)] = Cell2c(len(smu), nx2c, xmaxA, smu, snu, smv, data)
            for isorp in range(1, spB.nssh + 1):
                data = np.empty((nx2c, len(smu)))
                for k in range(len(smu)):
                    data[:, k] = radial(("rhoSP", b_, isorp, b_, snu[k] - 1, a_, smu[k] - 1))
                ts.cells2c[(a_, b_, P_RHOS_ONTOPR, isorp)] = Cell2c(len(smu), nx2c, xmaxA, smu, snu, smv, data)
            amu, anu, amv = munu_S(spA, spA)
            for isorp in range(1, spB.nssh + 1):
                data = np.empty((nx2c, len(amu)))
                for k in range(len(amu)):
                    i, j = amu[k] - 1, anu[k] - 1
                    data[:, k] = radial(("rhoSA", a_, b_, isorp, min(i, j), max(i, j)))
                ts.cells2c[(a_, b_, P_RHOS_ATOM, isorp)] = Cell2c(len(amu), nx2c, xmaxA, amu, anu, amv, data)

    if with_3c:
        any3 = next(iter(ts.cells3c.values()))
        n3c = any3.nx
        xg = _grid_bohr(n3c)
        for a_ in range(1, ns + 1):
            for b_ in range(1, ns + 1):
                lay = {P_RHO_3C: munu_3c(species[a_ - 1], species[b_ - 1]),
                       P_RHOS_3C: munu_S(species[a_ - 1], species[b_ - 1])}
                for c_ in range(1, ns + 1):
                    for itype in (P_RHO_3C, P_RHOS_3C):
                        mu, nu, mv = lay[itype]
                        nME = len(mu)
                        for isorp in range(1, species[c_ - 1].nssh + 1):
                            for itheta in range(1, P_MAXTHETA + 1):
                                A = rng.uniform(-0.5, 0.5, size=nME)
                                ax = rng.uniform(0.5, 1.5, size=nME)
                                by = rng.uniform(0.5, 1.5, size=nME)
                                data = (A[None, None, :] * np.exp(-ax[None, None, :] * xg[None, :, None])
                                        * np.exp(-by[None, None, :] * xg[:, None, None]))
                                ts.cells3c[(a_, b_, c_, itype, isorp, itheta)] = Cell3c(
                                    nME, n3c, n3c, any3.xmax, any3.ymax, mu, nu, mv, np.ascontiguousarray(data))
    return ts


# pseudopotential shells of the synthetic species: (lssh_PP, rcutoff_PP as the
# neighbour code uses it (Angstrom), Kleinman-Bylander cl per PP shell)
PP_SHELLS = {1: ([0], 1.05, [-2.1]), 6: ([0, 1], 1.40, [-3.2, 1.7]), 7: ([0, 1], 1.30, [-3.6, 1.9]),
             8: ([0, 1], 1.25, [-4.1, 2.3]), 14: ([0, 1], 1.90, [-2.6, 1.3]), 26: ([0, 1], 1.70, [-3.4, 1.6])}


def with_pp(species):
    """Copies of ``species`` carrying their pseudopotential shells (PP_SHELLS)."""
    out = []
    for sp in species:
        l, rc, cl = PP_SHELLS[sp.nZ]
        out.append(Species(sp.nZ, list(sp.lssh), list(sp.rcutoff), list(l), rc, list(cl)))
    return out


def munu_pp(sp_i: Species, sp_j: Species):
    """(mu, nu, mvalue) of the <phi_i | Psi_PP(j)> tables: the make_munu order with
    the PP shells of sp_j on the right."""
    ghost = Species(sp_j.nZ, list(sp_j.lssh_pp), [0.0] * sp_j.nssh_pp)
    return munu_2c(sp_i, ghost)


def add_pp_tables(ts: TableSet, seed: int = 6066) -> TableSet:
    """Add the non-local pseudopotential tables P_vnl (itype 6, isubtype 0) for
    every ordered species pair (own random stream): <phi_mu(A) | Psi_ncc(B)>
    with the orbitals of A on the left and the PP projectors of B on the right,
    radial shapes as in make_tables.  The species of ``ts`` must carry PP shells
    (with_pp)."""
    rng = np.random.default_rng(seed)
    species = ts.species
    ns = len(species)
    any2 = next(iter(ts.cells2c.values()))
    nx2c, xmaxA = any2.nx, any2.xmax
    r = _grid_bohr(nx2c)
    env = (1.0 - (r / XMAX_BOHR) ** 2) ** 3
    for a_ in range(1, ns + 1):
        for b_ in range(1, ns + 1):
            spA, spB = species[a_ - 1], species[b_ - 1]
            if spB.nssh_pp == 0:
                continue
            mu, nu, mv, shells = munu_pp(spA, spB)
            data = np.empty((nx2c, len(mu)))
            for k in range(len(mu)):
                A = rng.uniform(-2.0, 2.0)
                a = rng.uniform(0.8, 2.0)
                data[:, k] = A * np.exp(-a * r) * env
            ts.cells2c[(a_, b_, P_VNL, 0)] = Cell2c(len(mu), nx2c, xmaxA, mu, nu, mv, data)
    return ts


# neutral shell occupations of the synthetic species (info.dat Qneutral)
QNEUTRAL = {1: [1.0], 6: [2.0, 2.0], 7: [2.0, 3.0], 8: [2.0, 4.0], 14: [2.0, 2.0]}


def shell_qin(s: Structure, species, seed=5, scale=0.05):
    """Input shell charges Qin [natoms][4] = Qneutral + dQ (the DOGS charges of
    the structure, or a fresh random dQ with ``seed`` when it has none), as
    c.SYSTEM/M_configuraciones.f90:872-908 sets s%atom(:)%shell(:)%Qin."""
    rng = np.random.default_rng(seed)
    q = np.zeros((s.natoms, 4))
    for ia in range(s.natoms):
        sp_ = species[s.imass[ia] - 1]
        qn = np.array(QNEUTRAL.get(sp_.nZ, [1.0] * sp_.nssh)[:sp_.nssh], float)
        dq = s.dQ[ia, :sp_.nssh] if s.dQ is not None else rng.uniform(-scale, scale, size=sp_.nssh)
        q[ia, :sp_.nssh] = qn + dq
    return q


# ---------------------------------------------------------------------------
# structures
# ---------------------------------------------------------------------------
SI = Species(14, [0, 1], [5.4, 5.4])
H_SP = Species(1, [0], [4.5])
C_SP = Species(6, [0, 1], [5.0, 5.0])
N_SP = Species(7, [0, 1], [5.0, 5.0])
O_SP = Species(8, [0, 1], [5.0, 5.0])
# a transition-metal-like species with s, p and d shells (9 orbitals)
FE_SPD = Species(26, [0, 1, 2], [5.0, 5.0, 4.5])

DIAMOND_BASIS = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                          [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


def origin_shift(ratom: np.ndarray) -> np.ndarray:
    """Shift every atom by (pi, 1/e, sqrt 2) if any atom is within 1e-4 A of the
    origin (c.SYSTEM/M_configuraciones.f90:757-775)."""
    mags = np.sqrt(np.einsum("ij,ij->i", ratom, ratom))
    if np.any(mags < 1.0e-4):
        shifter = np.array([4.0 * math.atan(1.0), 1.0 / math.exp(1.0), math.sqrt(2.0)])
        ratom = ratom + shifter
    return ratom


def diamond_si(ncell=(2, 2, 2), a=5.43, sigma=0.05, seed=0) -> Structure:
    n1, n2, n3 = ncell
    cells = np.array([[i, j, k] for i in range(n1) for j in range(n2) for k in range(n3)], float)
    frac = (cells[:, None, :] + DIAMOND_BASIS[None, :, :]).reshape(-1, 3)
    ratom = frac * a
    rng = np.random.default_rng(seed)
    ratom = ratom + rng.normal(0.0, sigma, size=ratom.shape)
    lattice = np.diag([n1 * a, n2 * a, n3 * a]).astype(np.float64)
    return Structure(origin_shift(ratom), np.ones(len(ratom), np.int32), lattice, 0,
                     name=f"Si{len(ratom)}")


def _rejection_box(n, L, rmin, rng):
    """Uniform positions in a periodic cube with minimum-image separation >= rmin."""
    pos = np.empty((n, 3))
    ncell = max(1, int(L // rmin))
    cs = L / ncell
    grid = {}
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > 200 * n:
            raise RuntimeError("rejection sampling did not converge")
        p = rng.uniform(0.0, L, size=3)
        c = tuple((p // cs).astype(int) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    cc = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                    for q in grid.get(cc, ()):
                        d = p - pos[q]
                        d -= L * np.round(d / L)
                        if d @ d < rmin * rmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[k] = p
            grid.setdefault(c, []).append(k)
            k += 1
    return pos


def random_box(counts: dict, density: float, rmin: float, seed: int, species_order: list) -> Structure:
    """Random periodic cube.  ``counts`` maps nZ -> number of atoms; ``density`` in g/cm^3."""
    rng = np.random.default_rng(seed)
    labels = np.concatenate([np.full(c, z) for z, c in counts.items()])
    labels = rng.permutation(labels)
    mass = sum(MASS[z] * c for z, c in counts.items()) * AMU_G
    L = (mass / density) ** (1.0 / 3.0) * 1.0e8
    pos = _rejection_box(len(labels), L, rmin, rng)
    zs = [sp.nZ for sp in species_order]
    imass = np.array([zs.index(z) + 1 for z in labels], np.int32)
    lattice = np.diag([L, L, L]).astype(np.float64)
    return Structure(origin_shift(pos), imass, lattice, 0, name=f"box{len(labels)}")


def chno_counts(n):
    """C:H:N:O = 30:50:10:10 for n atoms (largest-remainder rounding)."""
    frac = {6: 0.3, 1: 0.5, 7: 0.1, 8: 0.1}
    base = {z: int(math.floor(f * n)) for z, f in frac.items()}
    rem = n - sum(base.values())
    order = sorted(frac, key=lambda z: -(frac[z] * n - base[z]))
    for z in order[:rem]:
        base[z] += 1
    return base


CHNO_SPECIES = [C_SP, H_SP, N_SP, O_SP]


def config(idx: int):
    """(species, structures) of the five benchmark workloads."""
    if idx == 0:
        return [SI], [diamond_si((2, 2, 2), sigma=0.05, seed=0)]
    if idx == 1:
        return [SI], [diamond_si((8, 8, 8), sigma=0.08, seed=0)]
    if idx == 2:
        return CHNO_SPECIES, [random_box(chno_counts(3000), 1.0, 0.9, 1, CHNO_SPECIES)]
    if idx == 3:
        return batch_species(), batch_structures()
    if idx == 4:
        return [C_SP], [random_box({6: 8192}, 2.0, 1.2, 3, [C_SP])]
    raise ValueError(idx)


def batch_species():
    # one species list for the whole batch: Si plus C/H/N/O
    return [SI, C_SP, H_SP, N_SP, O_SP]


def batch_structures(nstruct=512, seed=2):
    """Half Si diamond supercells (sigma 0.08 A), half C/H/N/O boxes at 1 g/cm^3;
    atom counts log-uniform in [16, 1024]."""
    rng = np.random.default_rng(seed)
    sp = batch_species()
    out = []
    for k in range(nstruct):
        n = int(round(math.exp(rng.uniform(math.log(16), math.log(1024)))))
        sub_seed = int(rng.integers(0, 2**31 - 1))
        if k % 2 == 0:
            # diamond supercell n1 x n2 x n3 with 8*n1*n2*n3 closest to n (>= 16 atoms)
            best = None
            for n1 in range(1, 9):
                for n2 in range(n1, 9):
                    for n3 in range(n2, 9):
                        m = 8 * n1 * n2 * n3
                        if m < 16:
                            continue
                        score = (abs(m - n), n3 - n1)
                        if best is None or score < best[0]:
                            best = (score, (n1, n2, n3))
            s = diamond_si(best[1], sigma=0.08, seed=sub_seed)
            s.imass[:] = 1
        else:
            s = random_box(chno_counts(n), 1.0, 0.9, sub_seed, sp)
        s.name = f"b{k}"
        out.append(s)
    return out


def random_dQ(structs, species, seed=7, scale=0.05):
    """Shell charges for exercising the charged (DOGS) terms."""
    rng = np.random.default_rng(seed)
    for s in structs:
        dq = np.zeros((s.natoms, 4))
        for ia in range(s.natoms):
            sp_ = species[s.imass[ia] - 1]
            dq[ia, :sp_.nssh] = rng.uniform(-scale, scale, size=sp_.nssh)
        s.dQ = dq
    return structs


def denmat(species, imass, neighn, neigh_j, neigh_back, seed=11, scale=0.5):
    """Synthetic density matrix for the force assemblers: one block per pair in
    the pair order of the neighbour lists (iatom major), column-major
    norb_mu x norb_nu, symmetric under the pair reversal of neigh_back
    (rho(j, i) = rho(i, j)^T), entries uniform in [-scale, scale] scaled by
    exp(-|pair index distance| / 8) so that the blocks are not all alike."""
    norb = np.array([sp.norb for sp in species], np.int64)
    imass = np.asarray(imass)
    neighn = np.asarray(neighn, np.int64)
    neigh_j = np.asarray(neigh_j, np.int64)
    neigh_back = np.asarray(neigh_back, np.int64)
    noff = np.concatenate([[0], np.cumsum(neighn)])
    pair_atom = np.repeat(np.arange(len(neighn)), neighn)
    nmu = norb[imass[pair_atom] - 1]
    nnu = norb[imass[neigh_j - 1] - 1]
    boff = np.concatenate([[0], np.cumsum(nmu * nnu)])
    back = noff[neigh_j - 1] + neigh_back - 1
    rng = np.random.default_rng(seed)
    rho = np.zeros(int(boff[-1]))
    for p in range(len(neigh_j)):
        q = int(back[p])
        if q < p:
            continue
        a, b = int(nmu[p]), int(nnu[p])
        blk = rng.uniform(-scale, scale, size=(a, b)) * np.exp(-abs(int(neigh_j[p]) - 1 - int(pair_atom[p])) / 8.0)
        if q == p:
            blk = 0.5 * (blk + blk.T)
        rho[boff[p]:boff[p + 1]] = blk.ravel(order="F")
        if q != p:
            rho[boff[q]:boff[q + 1]] = blk.T.ravel(order="F")
    return rho
