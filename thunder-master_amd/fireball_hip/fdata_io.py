"""Write synthetic tables and structures in the reference's on-disk formats.

Fdata (read by e.FDATA/M_Fdata_2c.f90:149-272 and e.FDATA/M_Fdata_3c.f90:143-264):

* ``2c.Z1.Z2.dat``   number of 2c cells of the ordered species pair
* ``2c.Z1.Z2.dir``   one record per cell: ``iindex itype isubtype fname nME nx xmax``
* ``<fname>``        nx lines of nME values (grid point major)
* ``TT_munu_2c.Z1.Z2.dat``  mu / nu / mvalue lines, one triple per cell of
  interaction TT in directory order (the reader keeps the unit connected across
  cells of one interaction, so the triples are consumed in sequence)
* ``clPP.ZZ.dat``    Kleinman-Bylander coefficients, one line per PP shell (empty: no PP shells)
* ``3c.Z1.Z2.Z3.dat`` / ``.dir`` with ``iindex itype isubtype itheta fname nME nx xmax ny ymax``;
  data files hold ny*nx lines (iy outer, ix inner) of nME values.

Values are written with 17 significant digits so that they read back to the
same binary64 numbers.

The structure file read by the oracle's Fortran harness is this project's own
format (``write_harness_input``).
"""
from __future__ import annotations

import os

import numpy as np

from .synth import Cell2c, Cell3c, Structure, TableSet


def _fmt(v):
    return repr(float(v))


def _write_rows(path, rows):
    with open(path, "w") as f:
        for row in rows:
            f.write(" ".join(_fmt(v) for v in row))
            f.write("\n")


def write_fdata(ts: TableSet, outdir: str):
    os.makedirs(outdir, exist_ok=True)
    sp = ts.species
    ns = len(sp)
    for a in range(1, ns + 1):
        za = sp[a - 1].nZ
        with open(os.path.join(outdir, f"clPP.{za:02d}.dat"), "w") as f:
            for c in getattr(sp[a - 1], "cl_pp", []):
                f.write(_fmt(c) + "\n")
        for b in range(1, ns + 1):
            zb = sp[b - 1].nZ
            keys = sorted(k for k in ts.cells2c if k[0] == a and k[1] == b)
            keys.sort(key=lambda k: (k[2], k[3]))
            with open(os.path.join(outdir, f"2c.{za:02d}.{zb:02d}.dat"), "w") as f:
                f.write(f"{len(keys)}\n")
            munu = {}
            with open(os.path.join(outdir, f"2c.{za:02d}.{zb:02d}.dir"), "w") as fd:
                for idx, k in enumerate(keys, start=1):
                    c = ts.cells2c[k]
                    fname = f"t2_{k[2]:02d}_{k[3]}_{za:02d}_{zb:02d}.dat"
                    fd.write(f"{idx} {k[2]} {k[3]} {fname} {c.nME} {c.nx} {_fmt(c.xmax)}\n")
                    _write_rows(os.path.join(outdir, fname), c.data)
                    munu.setdefault(k[2], []).append(c)
            for itype, cells in munu.items():
                with open(os.path.join(outdir, f"{itype:02d}_munu_2c.{za:02d}.{zb:02d}.dat"), "w") as f:
                    for c in cells:
                        for arr in (c.mu, c.nu, c.mvalue):
                            f.write(" ".join(str(int(v)) for v in arr) + "\n")
            for cc in range(1, ns + 1):
                zc = sp[cc - 1].nZ
                keys3 = sorted(k for k in ts.cells3c if k[:3] == (a, b, cc))
                keys3.sort(key=lambda k: (k[3], k[4], k[5]))
                if not keys3:
                    continue
                with open(os.path.join(outdir, f"3c.{za:02d}.{zb:02d}.{zc:02d}.dat"), "w") as f:
                    f.write(f"{len(keys3)}\n")
                munu3 = {}
                with open(os.path.join(outdir, f"3c.{za:02d}.{zb:02d}.{zc:02d}.dir"), "w") as fd:
                    for idx, k in enumerate(keys3, start=1):
                        c = ts.cells3c[k]
                        fname = f"t3_{k[3]:02d}_{k[4]}_{k[5]}_{za:02d}_{zb:02d}_{zc:02d}.dat"
                        fd.write(f"{idx} {k[3]} {k[4]} {k[5]} {fname} {c.nME} {c.nx} "
                                 f"{_fmt(c.xmax)} {c.ny} {_fmt(c.ymax)}\n")
                        _write_rows(os.path.join(outdir, fname), c.data.reshape(c.ny * c.nx, c.nME))
                        munu3.setdefault(k[3], []).append(c)
                for itype, cells in munu3.items():
                    with open(os.path.join(outdir, f"{itype:02d}_munu_3c.{za:02d}.{zb:02d}.{zc:02d}.dat"), "w") as f:
                        for c in cells:
                            for arr in (c.mu, c.nu, c.mvalue):
                                f.write(" ".join(str(int(v)) for v in arr) + "\n")


def write_pp_input(path: str, ts: TableSet):
    """Pseudopotential shells of every species for the PP harness: nssh_PP,
    then lssh_PP, then rcutoff_PP (the cl values travel in clPP.ZZ.dat)."""
    with open(path, "w") as f:
        for sp in ts.species:
            f.write(f"{sp.nssh_pp}\n")
            f.write(" ".join(str(l) for l in sp.lssh_pp) + "\n")
            f.write(_fmt(sp.rcutoff_pp) + "\n")


def write_harness_input(path: str, fdata_dir: str, ts: TableSet, s: Structure):
    """Input for the oracle's Fortran harness: species, lattice, atoms, shell charges."""
    with open(path, "w") as f:
        f.write(f"'{fdata_dir}'\n")
        f.write(f"{len(ts.species)}\n")
        for sp in ts.species:
            f.write(f"{sp.nZ} {sp.nssh}\n")
            f.write(" ".join(str(l) for l in sp.lssh) + "\n")
            f.write(" ".join(_fmt(r) for r in sp.rcutoff) + "\n")
        f.write(f"{s.natoms} {s.icluster}\n")
        for i in range(3):
            f.write(" ".join(_fmt(v) for v in s.lattice[i]) + "\n")
        for ia in range(s.natoms):
            f.write(f"{int(s.imass[ia])} " + " ".join(_fmt(v) for v in s.ratom[ia]) + "\n")
        for ia in range(s.natoms):
            nssh = ts.species[s.imass[ia] - 1].nssh
            dq = np.zeros(nssh) if s.dQ is None else s.dQ[ia, :nssh]
            f.write(" ".join(_fmt(v) for v in dq) + "\n")


def write_basis_file(path: str, species, s: Structure, z=None):
    """A structure file in the layout read_positions reads
    (c.SYSTEM/M_configuraciones.f90:651-681): "natoms icluster", the three
    lattice vectors, one "Z x y z" record per atom (positions as given, before
    any origin shift).  z overrides the atomic numbers per atom."""
    with open(path, "w") as f:
        f.write(f"{s.natoms} {s.icluster}\n")
        for i in range(3):
            f.write(" ".join(_fmt(v) for v in s.lattice[i]) + "\n")
        for ia in range(s.natoms):
            zi = species[s.imass[ia] - 1].nZ if z is None else int(z[ia])
            f.write(f"{zi} " + " ".join(_fmt(v) for v in s.ratom[ia]) + "\n")


# ---------------------------------------------------------------- reader
class _ListReader:
    """Fortran list-directed input over one text file: every read starts at a
    new record and takes values across records; blanks and commas separate,
    d/D exponents and r*c repeats are accepted, '/' ends a read."""

    def __init__(self, path):
        self.path = path
        with open(path) as f:
            self.lines = f.read().splitlines()
        self.pos = 0

    def values(self, n):
        out = []
        while len(out) < n:
            if self.pos >= len(self.lines):
                raise ValueError(f"short record in {self.path}")
            line = self.lines[self.pos]
            self.pos += 1
            stop = False
            for tok in line.replace(",", " ").split():
                if tok.startswith("/"):
                    stop = True
                    break
                if "*" in tok and tok.split("*", 1)[0].isdigit():
                    r, v = tok.split("*", 1)
                    out.extend([v] * int(r))
                else:
                    out.append(tok.strip("'\""))
                if len(out) >= n:
                    break
            if stop:
                break
        if len(out) < n:
            raise ValueError(f"short record in {self.path}")
        return out[:n]


def _f(tok):
    return float(tok.replace("d", "e").replace("D", "e"))


def read_fdata(fdir, species, types2c=(1, 2, 3, 4, 5, 10), types3c=(3,)):
    """TableSet of the given species (their nZ) from an Fdata directory, with
    the reference reader's semantics (e.FDATA/M_Fdata_2c.f90:149-272,
    M_Fdata_3c.f90:143-264): (itype, isubtype[, itheta]) names the cell at
    directory position iindex, and the mu/nu/mvalue unit stays open across
    cells, restarting a file whenever the interaction changes (create.x
    directories keep each interaction's cells together)."""
    ts = TableSet(species=list(species))
    z = [sp.nZ for sp in species]
    ns = len(species)
    for a in range(1, ns + 1):
        for b in range(1, ns + 1):
            za, zb = z[a - 1], z[b - 1]
            n = int(_ListReader(os.path.join(fdir, f"2c.{za:02d}.{zb:02d}.dat")).values(1)[0])
            d = _ListReader(os.path.join(fdir, f"2c.{za:02d}.{zb:02d}.dir"))
            unit = None
            cells, index = [], {}
            for _ in range(n):
                t = d.values(7)
                iindex, itype, isub, fname = int(t[0]), int(t[1]), int(t[2]), t[3]
                nme, nx, xmax = int(t[4]), int(t[5]), _f(t[6])
                data = None
                if itype in types2c:
                    r = _ListReader(os.path.join(fdir, fname))
                    data = np.array([[_f(v) for v in r.values(nme)] for _ in range(nx)])
                mpath = os.path.join(fdir, f"{itype:02d}_munu_2c.{za:02d}.{zb:02d}.dat")
                if unit is None or unit.path != mpath:
                    unit = _ListReader(mpath)
                mu, nu, mv = (np.array([int(v) for v in unit.values(nme)], np.int32) for _ in range(3))
                cells.append((itype, Cell2c(nme, nx, xmax, mu, nu, mv, data)))
                index[(itype, isub)] = iindex
            for (itype, isub), iindex in index.items():
                if itype in types2c:
                    ctype, cell = cells[iindex - 1]
                    if ctype not in types2c:
                        raise ValueError("2c index points at a cell of another interaction")
                    ts.cells2c[(a, b, itype, isub)] = cell
            for k in range(1, ns + 1):
                zk = z[k - 1]
                n3 = int(_ListReader(os.path.join(fdir, f"3c.{za:02d}.{zb:02d}.{zk:02d}.dat")).values(1)[0])
                d3 = _ListReader(os.path.join(fdir, f"3c.{za:02d}.{zb:02d}.{zk:02d}.dir"))
                unit3 = None
                cells3, index3 = [], {}
                for _ in range(n3):
                    t = d3.values(10)
                    iindex, itype, isub, ith, fname = int(t[0]), int(t[1]), int(t[2]), int(t[3]), t[4]
                    nme, nx, xmax, ny, ymax = int(t[5]), int(t[6]), _f(t[7]), int(t[8]), _f(t[9])
                    data = None
                    if itype in types3c:
                        r = _ListReader(os.path.join(fdir, fname))
                        data = np.array([[_f(v) for v in r.values(nme)] for _ in range(nx * ny)]).reshape(ny, nx, nme)
                    mpath = os.path.join(fdir, f"{itype:02d}_munu_3c.{za:02d}.{zb:02d}.{zk:02d}.dat")
                    if unit3 is None or unit3.path != mpath:
                        unit3 = _ListReader(mpath)
                    mu, nu, mv = (np.array([int(v) for v in unit3.values(nme)], np.int32) for _ in range(3))
                    cells3.append((itype, Cell3c(nme, nx, ny, xmax, ymax, mu, nu, mv, data)))
                    index3[(itype, isub, ith)] = iindex
                for (itype, isub, ith), iindex in index3.items():
                    if itype in types3c:
                        ctype, cell = cells3[iindex - 1]
                        if ctype not in types3c:
                            raise ValueError("3c index points at a cell of another interaction")
                        ts.cells3c[(a, b, k, itype, isub, ith)] = cell
    return ts
