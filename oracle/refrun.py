"""Run the reference harness (oracle/_ref/fb_harness_ref) and read its output.

Test infrastructure only: used by tests/, tests/golden/make_golden.py and the
cpu_baseline leg of bench.py.  The harness binary is compiled from the
reference's Fortran sources by ``make -C oracle ref`` (see oracle/Makefile).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "thunder-master_amd"))

from fireball_hip import fdata_io  # noqa: E402

HARNESS_REF = os.path.join(HERE, "_ref", "fb_harness_ref")
# the same harness with the OpenMP loop of M_assemble_3c_DOGS.f90 (make -C oracle ompref)
HARNESS_OMP = os.path.join(HERE, "_ref", "fb_harness_omp")


def available(binary=HARNESS_REF):
    return os.path.isfile(binary) and os.access(binary, os.X_OK)


class _Reader:
    def __init__(self, buf):
        self.buf = buf
        self.pos = 0

    def i32(self, n):
        a = np.frombuffer(self.buf, np.int32, n, self.pos)
        self.pos += 4 * n
        return a.copy()

    def f64(self, n):
        a = np.frombuffer(self.buf, np.float64, n, self.pos)
        self.pos += 8 * n
        return a.copy()


def parse_output(path, species, imass):
    """Return a dict with flat CSR lists (0-based atom indices kept 1-based as in
    the reference) and per-matrix flat block arrays in pair order."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    natoms, mbeta_max, npairs, ncommon_tot, nrep = r.i32(5)
    times = r.f64(6)
    neighn = r.i32(natoms)
    neigh_j, neigh_b, neigh_back = [], [], []
    for ia in range(natoms):
        n = int(neighn[ia])
        neigh_j.append(r.i32(n)); neigh_b.append(r.i32(n)); neigh_back.append(r.i32(n))
    neigh_self = r.i32(natoms)
    ncommon = r.i32(natoms)
    com = {k: [] for k in ("iatom_j", "iatom_b", "jatom_j", "jatom_b", "mneigh")}
    for ia in range(natoms):
        n = int(ncommon[ia])
        for k in ("iatom_j", "iatom_b", "jatom_j", "jatom_b", "mneigh"):
            com[k].append(r.i32(n))
    cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, np.int32)  # noqa: E731
    neigh_j, neigh_b, neigh_back = cat(neigh_j), cat(neigh_b), cat(neigh_back)
    norb = np.array([sp.norb for sp in species], np.int64)
    mats = {k: [] for k in ("S", "T", "DIP", "VNA", "VNAO")}
    p = 0
    for ia in range(natoms):
        nmu = norb[imass[ia] - 1]
        for _ in range(int(neighn[ia])):
            nnu = norb[imass[neigh_j[p] - 1] - 1]
            for k in ("S", "T", "DIP", "VNA", "VNAO"):
                mats[k].append(r.f64(int(nmu * nnu)))
            p += 1
    assert r.pos == len(r.buf), "trailing bytes in harness output"
    out = dict(natoms=int(natoms), mbeta_max=int(mbeta_max), npairs=int(npairs),
               ntriplets=int(ncommon_tot), nrep=int(nrep),
               times=dict(zip(("lists", "S", "T", "DIP", "VNA2C", "VNA3C"), times)),
               neighn=neighn, neigh_j=neigh_j, neigh_b=neigh_b, neigh_back=neigh_back,
               neigh_self=neigh_self, ncommon=ncommon,
               **{"com_" + k: cat(v) for k, v in com.items()})
    for k, v in mats.items():
        out[k] = np.concatenate(v) if v else np.zeros(0)
    return out


def run(ts, structure, binary=HARNESS_REF, nrep=1, workdir=None, timeout=3600, vna_bump=None, threads=1):
    """Write tables + structure, run the harness, return the parsed output.
    vna_bump: add it to every Vna block element between assemble_vna_2c and
    assemble_vna_3c.  threads: OMP_NUM_THREADS (the OpenMP harness's 3c loop).
    res["rep_times"] holds each repetition's (S, T, dipole_z, vna_2c, vna_3c)
    seconds."""
    own = workdir is None
    tmp = tempfile.mkdtemp(prefix="fbref_") if own else workdir
    fdir = os.path.join(tmp, "Fdata")
    if not os.path.isdir(fdir):
        fdata_io.write_fdata(ts, fdir)
    inp = os.path.join(tmp, "harness.in")
    fdata_io.write_harness_input(inp, fdir, ts, structure)
    outp = os.path.join(tmp, "harness.out")
    env = dict(os.environ, OMP_NUM_THREADS=str(int(threads)))
    if threads > 1:
        env.setdefault("OMP_PROC_BIND", "close")
    args = [binary, inp, outp, str(nrep)] + ([repr(float(vna_bump))] if vna_bump is not None else [])
    proc = subprocess.run(args, cwd=tmp, check=True, env=env,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout)
    res = parse_output(outp, ts.species, structure.imass)
    res["rep_times"] = [tuple(float(x) for x in line.split()[2:7])
                        for line in proc.stdout.decode().splitlines() if line.split()[:1] == ["rep"]]
    if own:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    return res


HARNESS_DREF = os.path.join(HERE, "_ref", "fb_dharness_ref")


def parse_forces(path, species, imass, neighn, neigh_j):
    """Dblocks (S, T, DIP) in pair order as flat [3*block] arrays, vna_atom and
    vna_ontop as [npairs, 3], f3naa/f3nab/f3nac as [natoms, 3]."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    natoms, npairs, ntrip, nrep = r.i32(4)
    times = r.f64(5)
    norb = np.array([sp.norb for sp in species], np.int64)
    D = {k: [] for k in ("DS", "DT", "DDIP")}
    p = 0
    for ia in range(natoms):
        nmu = norb[imass[ia] - 1]
        for _ in range(int(neighn[ia])):
            nnu = norb[imass[neigh_j[p] - 1] - 1]
            for k in ("DS", "DT", "DDIP"):
                D[k].append(r.f64(int(3 * nmu * nnu)))
            p += 1
    fa, fo = [], []
    for ia in range(natoms):
        n = int(neighn[ia])
        fa.append(r.f64(3 * n).reshape(n, 3))
        fo.append(r.f64(3 * n).reshape(n, 3))
    f3 = r.f64(9 * natoms).reshape(natoms, 3, 3)
    assert r.pos == len(r.buf), "trailing bytes in force harness output"
    out = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in D.items()}
    out["vna_atom"] = np.concatenate(fa) if fa else np.zeros((0, 3))
    out["vna_ontop"] = np.concatenate(fo) if fo else np.zeros((0, 3))
    out["f3naa"], out["f3nab"], out["f3nac"] = f3[:, 0].copy(), f3[:, 1].copy(), f3[:, 2].copy()
    out["times"] = dict(zip(("DS", "DT", "DDIP", "DVNA2C", "DVNA3C"), times))
    out["nrep"] = int(nrep)
    out["ntriplets"] = int(ntrip)
    return out


def run_forces(ts, structure, rho, lists, binary=HARNESS_DREF, nrep=1, workdir=None, timeout=3600,
               dq_scale=None):
    """Run the force harness on (tables, structure) with density blocks rho
    (pair order of `lists`, the reference's neighbour lists).  dq_scale:
    multiply every shell charge by it between the H+S assembly and the force
    assemblers (what a charge mixer does in an SCF loop)."""
    own = workdir is None
    tmp = tempfile.mkdtemp(prefix="fbdref_") if own else workdir
    fdir = os.path.join(tmp, "Fdata")
    if not os.path.isdir(fdir):
        fdata_io.write_fdata(ts, fdir)
    inp = os.path.join(tmp, "harness.in")
    fdata_io.write_harness_input(inp, fdir, ts, structure)
    rhof = os.path.join(tmp, "denmat.bin")
    np.ascontiguousarray(rho, np.float64).tofile(rhof)
    outp = os.path.join(tmp, "dharness.out")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    args = [binary, inp, rhof, outp, str(nrep)] + ([repr(float(dq_scale))] if dq_scale is not None else [])
    subprocess.run(args, cwd=tmp, check=True, env=env,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout)
    res = parse_forces(outp, ts.species, structure.imass, lists["neighn"], lists["neigh_j"])
    if own:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    return res


HARNESS_RREF = os.path.join(HERE, "_ref", "fb_rharness_ref")


def parse_rho(path, species, imass, neighn, neigh_j):
    """rho_in, rho_bond (orbital blocks) and rho_in_weighted, rho_bond_weighted
    (shell blocks nssh_mu x nssh_nu), each flat in pair order."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    natoms, npairs, ntrip, nrep = r.i32(4)
    times = r.f64(4)
    norb = np.array([sp.norb for sp in species], np.int64)
    nssh = np.array([sp.nssh for sp in species], np.int64)
    out = {k: [] for k in ("RHO_IN", "RHO_BOND", "RHOW_IN", "RHOW_BOND")}
    for size, keys in ((norb, ("RHO_IN", "RHO_BOND")), (nssh, ("RHOW_IN", "RHOW_BOND"))):
        p = 0
        for ia in range(natoms):
            a = size[imass[ia] - 1]
            for _ in range(int(neighn[ia])):
                b = size[imass[neigh_j[p] - 1] - 1]
                for k in keys:
                    out[k].append(r.f64(int(a * b)))
                p += 1
    assert r.pos == len(r.buf), "trailing bytes in density harness output"
    res = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in out.items()}
    res["times"] = dict(zip(("RHO2C", "RHO3C", "RHOW2C", "RHOW3C"), times))
    res["nrep"] = int(nrep)
    res["ntriplets"] = int(ntrip)
    return res


def write_qin(path, qin, species, imass):
    """Stream of binary64 shell charges, nssh values per atom in atom order."""
    vals = [np.asarray(qin[ia, :species[imass[ia] - 1].nssh], np.float64) for ia in range(len(imass))]
    np.concatenate(vals).tofile(path)


def run_rho(ts, structure, qin, lists, binary=HARNESS_RREF, nrep=1, workdir=None, timeout=3600):
    """Run the McWEDA density harness on (tables, structure) with shell charges
    qin [natoms][4]; `lists` are the reference's neighbour lists (pair order)."""
    own = workdir is None
    tmp = tempfile.mkdtemp(prefix="fbrref_") if own else workdir
    fdir = os.path.join(tmp, "Fdata")
    if not os.path.isdir(fdir):
        fdata_io.write_fdata(ts, fdir)
    inp = os.path.join(tmp, "harness.in")
    fdata_io.write_harness_input(inp, fdir, ts, structure)
    qf = os.path.join(tmp, "qin.bin")
    write_qin(qf, qin, ts.species, structure.imass)
    outp = os.path.join(tmp, "rharness.out")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([binary, inp, qf, outp, str(nrep)], cwd=tmp, check=True, env=env,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout)
    res = parse_rho(outp, ts.species, structure.imass, lists["neighn"], lists["neigh_j"])
    if own:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    return res


HARNESS_PREF = os.path.join(HERE, "_ref", "fb_pharness_ref")


def parse_pp(path, species, imass):
    """PP lists (per list: neighn and the flat per-atom arrays, 1-based as in the
    reference), sVNL blocks (per PP pair, norb_mu x norb_PP_nu) and vnl blocks
    (per PPp pair, norb_mu x norb_nu), flat in pair order."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    natoms, nrep = r.i32(2)
    times = r.f64(3)
    out = {"nrep": int(nrep), "times": dict(zip(("SVNL", "VNL2C", "VNL3C"), times))}

    def block_list(prefix, fields):
        n = r.i32(natoms)
        cols = {k: [] for k in fields}
        for ia in range(natoms):
            for k in fields:
                cols[k].append(r.i32(int(n[ia])))
        out[prefix + "_n"] = n
        for k in fields:
            out[prefix + "_" + k] = np.concatenate(cols[k]) if cols[k] else np.zeros(0, np.int32)

    block_list("pp", ("j", "b", "map"))
    out["pp_self"] = r.i32(natoms)
    block_list("ppx", ("j", "b", "map", "point"))
    out["ppx_self"] = r.i32(natoms)
    block_list("ppp", ("j", "b"))
    out["ppp_self"] = r.i32(natoms)
    block_list("ppc", ("ij", "ib", "jj", "jb", "m"))
    norb = np.array([sp.norb for sp in species], np.int64)
    norb_pp = np.array([sp.norb_pp for sp in species], np.int64)
    sv, vl = [], []
    p = 0
    for ia in range(natoms):
        for _ in range(int(out["pp_n"][ia])):
            sv.append(r.f64(int(norb[imass[ia] - 1] * norb_pp[imass[out["pp_j"][p] - 1] - 1])))
            p += 1
    p = 0
    for ia in range(natoms):
        for _ in range(int(out["ppp_n"][ia])):
            vl.append(r.f64(int(norb[imass[ia] - 1] * norb[imass[out["ppp_j"][p] - 1] - 1])))
            p += 1
    assert r.pos == len(r.buf), "trailing bytes in PP harness output"
    out["SVNL"] = np.concatenate(sv) if sv else np.zeros(0)
    out["VNL"] = np.concatenate(vl) if vl else np.zeros(0)
    return out


def run_pp(ts, structure, binary=HARNESS_PREF, nrep=1, workdir=None, timeout=3600):
    """Run the pseudopotential harness on (tables, structure); the species of
    ts carry their PP shells (synth.with_pp) and ts the P_vnl tables."""
    own = workdir is None
    tmp = tempfile.mkdtemp(prefix="fbpref_") if own else workdir
    fdir = os.path.join(tmp, "Fdata")
    if not os.path.isdir(fdir):
        fdata_io.write_fdata(ts, fdir)
    inp = os.path.join(tmp, "harness.in")
    fdata_io.write_harness_input(inp, fdir, ts, structure)
    ppin = os.path.join(tmp, "pp.in")
    fdata_io.write_pp_input(ppin, ts)
    outp = os.path.join(tmp, "pharness.out")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([binary, inp, ppin, outp, str(nrep)], cwd=tmp, check=True, env=env,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout)
    res = parse_pp(outp, ts.species, structure.imass)
    if own:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    return res


HARNESS_POSREF = os.path.join(HERE, "_ref", "fb_posharness_ref")


def run_positions(species, basis_file, binary=HARNESS_POSREF, timeout=600):
    """The reference's read_positions on basis_file (oracle/ref/fb_posharness.f90):
    natoms, icluster, shifted (ishiftO), norbitals, lattice [3][3], imass,
    iblock_slot and ratom [natoms][3] as read_positions leaves them."""
    tmp = tempfile.mkdtemp(prefix="fbpos_")
    spf = os.path.join(tmp, "species.in")
    with open(spf, "w") as f:
        f.write(f"{len(species)}\n")
        for sp in species:
            f.write(f"{sp.nZ} {float(getattr(sp, 'xmass', 2.0 * sp.nZ))!r} {sp.norb}\n")
    outp = os.path.join(tmp, "pos.out")
    # the reference keeps the file name in 25 characters (T_structure%basisfile)
    import shutil
    shutil.copyfile(basis_file, os.path.join(tmp, "s.inp"))
    subprocess.run([binary, spf, "s.inp", outp], cwd=tmp, check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=timeout)
    with open(outp, "rb") as f:
        r = _Reader(f.read())
    natoms, icluster, shifted, norbitals = (int(v) for v in r.i32(4))
    lattice = r.f64(9).reshape(3, 3)
    imass = np.zeros(natoms, np.int32)
    slot = np.zeros(natoms, np.int32)
    ratom = np.zeros((natoms, 3))
    for ia in range(natoms):
        imass[ia], slot[ia] = r.i32(2)
        ratom[ia] = r.f64(3)
    import shutil
    shutil.rmtree(tmp, ignore_errors=True)
    return {"natoms": natoms, "icluster": icluster, "shifted": shifted, "norbitals": norbitals,
            "lattice": lattice, "imass": imass, "iblock_slot": slot, "ratom": ratom}
