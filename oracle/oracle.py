"""ctypes binding of the C oracle (oracle/libfboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this
module; it is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfboracle.so")
MAXSH = 4
MATS = {"S": 0, "T": 1, "DIP": 2, "VNA": 3, "VNAO": 4}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, i64 = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int64
        L.fbo_tables_new.restype = vp
        L.fbo_tables_new.argtypes = [C.c_int, ip, ip, dp]
        L.fbo_tables_free.argtypes = [vp]
        L.fbo_add_cell2c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, ip, ip, ip, dp]
        L.fbo_add_cell3c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_double, C.c_double, ip, ip, ip, dp]
        L.fbo_distance.restype = C.c_double
        L.fbo_distance.argtypes = [dp, dp]
        L.fbo_smoother.restype = C.c_double
        L.fbo_smoother.argtypes = [C.c_double, C.c_double, C.c_double]
        L.fbo_epsilon.argtypes = [dp, dp, dp]
        L.fbo_twister.argtypes = [dp, C.c_int, dp, dp]
        L.fbo_rotate.argtypes = [vp, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp]
        L.fbo_getMEs_2c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, dp]
        L.fbo_getMEs_3c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_int, C.c_int, C.c_double, dp]
        L.fbo_make_cells.argtypes = [dp, C.c_int, dp]
        L.fbo_system_new.restype = vp
        L.fbo_system_new.argtypes = [vp, C.c_int, dp, ip, dp, C.c_int, dp]
        L.fbo_system_free.argtypes = [vp]
        L.fbo_build_lists.argtypes = [vp]
        for f in ("fbo_npairs", "fbo_ntriplets", "fbo_block_len"):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [vp]
        L.fbo_get_lists.argtypes = [vp] + [ip] * 11
        for f in ("fbo_assemble_S", "fbo_assemble_T", "fbo_assemble_dipole_z", "fbo_assemble_vna_2c",
                  "fbo_assemble_vna_3c", "fbo_frame_fallbacks"):
            getattr(L, f).argtypes = [vp]
        L.fbo_get_blocks.argtypes = [vp, C.c_int, dp]
        L.fbo_Dsmoother.restype = C.c_double
        L.fbo_Dsmoother.argtypes = [C.c_double, C.c_double, C.c_double]
        L.fbo_Depsilon_2c.argtypes = [dp, dp, C.c_double, dp, dp]
        L.fbo_Depsilon_3c.argtypes = [dp, dp, dp, C.c_double, dp, dp, dp, dp, dp]
        L.fbo_Dtwister.argtypes = [dp, dp, dp, dp]
        L.fbo_Drotate.argtypes = [vp, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, dp, dp, dp]
        L.fbo_getDMEs_2c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, dp, dp]
        L.fbo_getDMEs_3c.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_double, dp, dp, dp, dp]
        L.fbo_set_denmat.argtypes = [vp, dp]
        for f in ("fbo_Dassemble_S", "fbo_Dassemble_T", "fbo_Dassemble_dipole_z", "fbo_Dassemble_vna_2c",
                  "fbo_Dassemble_vna_3c", "fbo_deps_warnings"):
            getattr(L, f).argtypes = [vp]
        L.fbo_get_dblocks.argtypes = [vp, C.c_int, dp]
        L.fbo_get_forces.argtypes = [vp, dp, dp, dp, dp, dp]
        L.fbo_set_qin.argtypes = [vp, dp]
        for f in ("fbo_assemble_rho_2c", "fbo_assemble_rho_3c", "fbo_assemble_rho_weighted_2c",
                  "fbo_assemble_rho_weighted_3c"):
            getattr(L, f).argtypes = [vp]
        L.fbo_shell_block_len.restype = i64
        L.fbo_shell_block_len.argtypes = [vp]
        L.fbo_get_rho.argtypes = [vp, C.c_int, dp]
        L.fbo_tables_set_pp.argtypes = [vp, ip, ip, dp, dp]
        L.fbo_rotate_PP.argtypes = [vp, C.c_int, C.c_int, dp, C.c_int, C.c_int, dp, dp]
        for f in ("fbo_build_pp_lists", "fbo_assemble_svnl", "fbo_assemble_vnl_2c", "fbo_assemble_vnl_3c"):
            getattr(L, f).argtypes = [vp]
        L.fbo_pp_sizes.argtypes = [vp, C.POINTER(C.c_int64)]
        L.fbo_get_pp_list.argtypes = [vp, C.c_int, ip, ip, ip, ip, ip, ip, ip]
        L.fbo_get_pp_blocks.argtypes = [vp, C.c_int, dp]
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with status {rc}")


class Tables:
    def __init__(self, ts):
        L = lib()
        self.ts = ts
        ns = len(ts.species)
        nssh = np.array([sp.nssh for sp in ts.species], np.int32)
        lssh = np.zeros((ns, MAXSH), np.int32)
        rc = np.zeros((ns, MAXSH), np.float64)
        for i, sp in enumerate(ts.species):
            lssh[i, :sp.nssh] = sp.lssh
            rc[i, :sp.nssh] = sp.rcutoffA
        self.h = L.fbo_tables_new(ns, _ip(nssh), _ip(lssh), _dp(rc))
        keep = []
        for (a, b, it, isub), c in ts.cells2c.items():
            arrs = [np.ascontiguousarray(x, np.int32) for x in (c.mu, c.nu, c.mvalue)]
            d = np.ascontiguousarray(c.data, np.float64)
            keep.append((arrs, d))
            _check(L.fbo_add_cell2c(self.h, a, b, it, isub, c.nME, c.nx, c.xmax, *(_ip(x) for x in arrs), _dp(d)),
                   "add_cell2c")
        for (a, b, cc, it, isub, ith), c in ts.cells3c.items():
            arrs = [np.ascontiguousarray(x, np.int32) for x in (c.mu, c.nu, c.mvalue)]
            d = np.ascontiguousarray(c.data, np.float64)
            _check(L.fbo_add_cell3c(self.h, a, b, cc, it, isub, ith, c.nME, c.nx, c.ny, c.xmax, c.ymax,
                                    *(_ip(x) for x in arrs), _dp(d)), "add_cell3c")
        if any(getattr(sp, "nssh_pp", 0) for sp in ts.species):
            npp = np.array([sp.nssh_pp for sp in ts.species], np.int32)
            lpp = np.zeros((ns, MAXSH), np.int32)
            clv = np.zeros((ns, MAXSH))
            rpp = np.array([sp.rcutoff_pp for sp in ts.species], np.float64)
            for i, sp in enumerate(ts.species):
                lpp[i, :sp.nssh_pp] = sp.lssh_pp
                clv[i, :sp.nssh_pp] = sp.cl_pp
            _check(L.fbo_tables_set_pp(self.h, _ip(npp), _ip(lpp), _dp(rpp), _dp(clv)), "set_pp")

    def __del__(self):
        if getattr(self, "h", None):
            lib().fbo_tables_free(self.h)
            self.h = None

    def getMEs_2c(self, is_, js, iint, isub, x, nmu, nnu):
        out = np.zeros(nmu * nnu)
        _check(lib().fbo_getMEs_2c(self.h, is_, js, iint, isub, x, nmu, nnu, _dp(out)), "getMEs_2c")
        return out.reshape(nnu, nmu).T

    def getMEs_3c(self, is_, js, ks, iint, isub, x, y, nmu, nnu, cost):
        out = np.zeros(nmu * nnu)
        _check(lib().fbo_getMEs_3c(self.h, is_, js, ks, iint, isub, x, y, nmu, nnu, cost, _dp(out)), "getMEs_3c")
        return out.reshape(nnu, nmu).T

    def getDMEs_2c(self, is_, js, iint, isub, x, nmu, nnu):
        """(value, d/dx) blocks of getDMEs_Fdata_2c."""
        h, d = np.zeros(nmu * nnu), np.zeros(nmu * nnu)
        _check(lib().fbo_getDMEs_2c(self.h, is_, js, iint, isub, x, nmu, nnu, _dp(h), _dp(d)), "getDMEs_2c")
        return h.reshape(nnu, nmu).T, d.reshape(nnu, nmu).T

    def getDMEs_3c(self, is_, js, ks, iint, isub, x, y, nmu, nnu, cost):
        """(value, d/dcos, d/dx, d/dy) blocks of getDMEs_Fdata_3c."""
        out = [np.zeros(nmu * nnu) for _ in range(4)]
        _check(lib().fbo_getDMEs_3c(self.h, is_, js, ks, iint, isub, x, y, nmu, nnu, cost, *(_dp(o) for o in out)),
               "getDMEs_3c")
        return tuple(o.reshape(nnu, nmu).T for o in out)

    def Drotate(self, in1, in2, eps, deps, M, DM):
        """Drotate: M (nmu, nnu), DM and deps as (3, ...) arrays indexed [ix, ...];
        returns Dmatx[ix, imu, inu]."""
        nmu, nnu = M.shape
        out = np.zeros(3 * nmu * nnu)
        e = np.asfortranarray(eps, np.float64).ravel(order="F")
        de = np.asfortranarray(deps, np.float64).ravel(order="F")
        Mf = np.asfortranarray(M, np.float64).ravel(order="F")
        DMf = np.asfortranarray(DM, np.float64).ravel(order="F")
        lib().fbo_Drotate(self.h, in1, in2, _dp(e), _dp(de), nmu, nnu, _dp(Mf), _dp(DMf), _dp(out))
        return out.reshape(nnu, nmu, 3).transpose(2, 1, 0)

    def rotate(self, in1, in3, eps, M):
        nmu, nnu = M.shape
        Mf = np.asfortranarray(M, np.float64)
        epsf = np.asfortranarray(eps, np.float64)
        out = np.zeros(nmu * nnu)
        lib().fbo_rotate(self.h, in1, in3, _dp(epsf.ravel(order="F")), nmu, nnu, _dp(Mf.ravel(order="F")), _dp(out))
        return out.reshape(nnu, nmu).T


def epsilon(r1, r2):
    spe = np.zeros(9)
    lib().fbo_epsilon(_dp(np.ascontiguousarray(r1, np.float64)), _dp(np.ascontiguousarray(r2, np.float64)), _dp(spe))
    return spe.reshape(3, 3).T


def twister(eps, have_d=True):
    dmat = np.zeros(25)
    pmat = np.zeros(9)
    e = np.asfortranarray(eps, np.float64).ravel(order="F")
    lib().fbo_twister(_dp(e), int(have_d), _dp(dmat), _dp(pmat))
    return dmat.reshape(5, 5).T, pmat.reshape(3, 3).T


def smoother(r, rend, x):
    return lib().fbo_smoother(r, rend, x)


def Dsmoother(r, rend, x):
    return lib().fbo_Dsmoother(r, rend, x)


def Depsilon_2c(r1, r2, z, eps):
    """deps[ix, imu, k] = d eps(imu, k) / d r1(ix); second value: colinear guard taken."""
    d = np.zeros(27)
    e = np.asfortranarray(eps, np.float64).ravel(order="F")
    w = lib().fbo_Depsilon_2c(_dp(np.ascontiguousarray(r1, np.float64)), _dp(np.ascontiguousarray(r2, np.float64)),
                              z, _dp(e), _dp(d))
    return d.reshape(3, 3, 3).transpose(2, 1, 0), bool(w)


def Depsilon_3c(r1, r2, r21, z, ratm, rnabc, eps):
    """(d eps/d ratm, d eps/d r1) as [ix, imu, k] arrays, and the colinear flag."""
    a, b = np.zeros(27), np.zeros(27)
    c = lambda v: _dp(np.ascontiguousarray(v, np.float64))  # noqa: E731
    e = np.asfortranarray(eps, np.float64).ravel(order="F")
    w = lib().fbo_Depsilon_3c(c(r1), c(r2), c(r21), z, c(ratm), c(rnabc), _dp(e), _dp(a), _dp(b))
    return a.reshape(3, 3, 3).transpose(2, 1, 0), b.reshape(3, 3, 3).transpose(2, 1, 0), bool(w)


def make_cells(lattice, icluster=0):
    xl = np.zeros((729, 3))
    lat = np.ascontiguousarray(lattice, np.float64)
    n = lib().fbo_make_cells(_dp(lat), icluster, _dp(xl))
    return xl[: n + 1].copy()


class System:
    """Oracle run of the full path on one structure."""

    def __init__(self, tables: Tables, st):
        L = lib()
        self.tables = tables
        self.st = st
        ratom = np.ascontiguousarray(st.ratom, np.float64)
        imass = np.ascontiguousarray(st.imass, np.int32)
        lat = np.ascontiguousarray(st.lattice, np.float64)
        dq = None if st.dQ is None else np.ascontiguousarray(st.dQ[:, :MAXSH], np.float64)
        self.h = L.fbo_system_new(tables.h, st.natoms, _dp(ratom), _ip(imass), _dp(lat), st.icluster,
                                  _dp(dq) if dq is not None else None)
        if not self.h:
            raise OracleError("fbo_system_new failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().fbo_system_free(self.h)
            self.h = None

    def build_lists(self):
        _check(lib().fbo_build_lists(self.h), "build_lists")
        L = lib()
        n = self.st.natoms
        npairs, ntrip = L.fbo_npairs(self.h), L.fbo_ntriplets(self.h)
        out = {k: np.zeros(n, np.int32) for k in ("neighn", "neigh_self", "ncommon")}
        for k in ("neigh_j", "neigh_b", "neigh_back"):
            out[k] = np.zeros(max(npairs, 1), np.int32)
        for k in ("com_iatom_j", "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh"):
            out[k] = np.zeros(max(ntrip, 1), np.int32)
        _check(L.fbo_get_lists(self.h, _ip(out["neighn"]), _ip(out["neigh_j"]), _ip(out["neigh_b"]),
                               _ip(out["neigh_back"]), _ip(out["neigh_self"]), _ip(out["ncommon"]),
                               _ip(out["com_iatom_j"]), _ip(out["com_iatom_b"]), _ip(out["com_jatom_j"]),
                               _ip(out["com_jatom_b"]), _ip(out["com_mneigh"])), "get_lists")
        for k in ("neigh_j", "neigh_b", "neigh_back"):
            out[k] = out[k][:npairs]
        for k in ("com_iatom_j", "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh"):
            out[k] = out[k][:ntrip]
        out["npairs"], out["ntriplets"] = int(npairs), int(ntrip)
        self.lists = out
        return out

    def assemble(self, which=("S", "T", "DIP", "VNA2C", "VNA3C")):
        L = lib()
        fn = {"S": L.fbo_assemble_S, "T": L.fbo_assemble_T, "DIP": L.fbo_assemble_dipole_z,
              "VNA2C": L.fbo_assemble_vna_2c, "VNA3C": L.fbo_assemble_vna_3c}
        for w in which:
            _check(fn[w](self.h), "assemble_" + w)

    def blocks(self, mat):
        n = lib().fbo_block_len(self.h)
        out = np.zeros(max(n, 1))
        _check(lib().fbo_get_blocks(self.h, MATS[mat], _dp(out)), "get_blocks")
        return out[:n]

    def set_denmat(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        assert rho.size == lib().fbo_block_len(self.h)
        self._rho = rho
        _check(lib().fbo_set_denmat(self.h, _dp(rho)), "set_denmat")

    def dassemble(self, which=("S", "T", "DIP", "VNA2C", "VNA3C")):
        L = lib()
        fn = {"S": L.fbo_Dassemble_S, "T": L.fbo_Dassemble_T, "DIP": L.fbo_Dassemble_dipole_z,
              "VNA2C": L.fbo_Dassemble_vna_2c, "VNA3C": L.fbo_Dassemble_vna_3c}
        for w in which:
            _check(fn[w](self.h), "Dassemble_" + w)

    def dblocks(self, mat):
        n = 3 * lib().fbo_block_len(self.h)
        out = np.zeros(max(n, 1))
        _check(lib().fbo_get_dblocks(self.h, {"S": 0, "T": 1, "DIP": 2}[mat], _dp(out)), "get_dblocks")
        return out[:n]

    def forces(self):
        L = lib()
        npairs, n = L.fbo_npairs(self.h), self.st.natoms
        fa, fo = np.zeros((max(npairs, 1), 3)), np.zeros((max(npairs, 1), 3))
        f3 = [np.zeros((n, 3)) for _ in range(3)]
        _check(L.fbo_get_forces(self.h, _dp(fa), _dp(fo), *(_dp(x) for x in f3)), "get_forces")
        return {"vna_atom": fa[:npairs], "vna_ontop": fo[:npairs], "f3naa": f3[0], "f3nab": f3[1], "f3nac": f3[2]}

    def run_forces(self, rho):
        """Dassemble_S/T/dipole_z/vna_2c/vna_3c after run_all(); Dblocks and forces."""
        self.set_denmat(rho)
        self.dassemble()
        res = {"D" + m: self.dblocks(m) for m in ("S", "T", "DIP")}
        res.update(self.forces())
        res["deps_warnings"] = lib().fbo_deps_warnings(self.h)
        return res

    RHO = {"RHO_IN": 0, "RHO_BOND": 1, "RHOW_IN": 2, "RHOW_BOND": 3}

    def set_qin(self, qin):
        qin = np.ascontiguousarray(qin[:, :MAXSH], np.float64)
        assert qin.shape == (self.st.natoms, MAXSH)
        self._qin = qin
        _check(lib().fbo_set_qin(self.h, _dp(qin)), "set_qin")

    def assemble_rho(self, which=("RHO2C", "RHO3C", "RHOW2C", "RHOW3C")):
        """assemble_rho_2c, _3c, assemble_rho_weighted_2c, _3c, in the order of
        j.ASSEMBLERS/M_assemble_vxc_McWEDA_DOGS.f90:148-154."""
        L = lib()
        fn = {"RHO2C": L.fbo_assemble_rho_2c, "RHO3C": L.fbo_assemble_rho_3c,
              "RHOW2C": L.fbo_assemble_rho_weighted_2c, "RHOW3C": L.fbo_assemble_rho_weighted_3c}
        for w in which:
            _check(fn[w](self.h), "assemble_" + w)

    def rho(self, which):
        L = lib()
        k = self.RHO[which]
        n = L.fbo_block_len(self.h) if k < 2 else L.fbo_shell_block_len(self.h)
        out = np.zeros(max(n, 1))
        _check(L.fbo_get_rho(self.h, k, _dp(out)), "get_rho")
        return out[:n]

    def run_rho(self, qin):
        """The four McWEDA density assemblers after build_lists(); all four arrays."""
        self.set_qin(qin)
        self.assemble_rho()
        return {k: self.rho(k) for k in self.RHO}

    PP_LISTS = {"pp": (0, ("j", "b", "map")), "ppx": (1, ("j", "b", "map", "point")), "ppp": (2, ("j", "b")),
                "ppc": (3, ("ij", "ib", "jj", "jb", "m"))}

    def run_pp(self):
        """driver_neighbors_PP, assemble_svnl, assemble_vnl_2c, assemble_vnl_3c
        (after build_lists); the PP lists and the sVNL and vnl blocks."""
        L = lib()
        for f, w in ((L.fbo_build_pp_lists, "build_pp_lists"), (L.fbo_assemble_svnl, "assemble_svnl"),
                     (L.fbo_assemble_vnl_2c, "assemble_vnl_2c"), (L.fbo_assemble_vnl_3c, "assemble_vnl_3c")):
            _check(f(self.h), w)
        sz = (C.c_int64 * 6)()
        L.fbo_pp_sizes(self.h, sz)
        n = self.st.natoms
        out = {}
        for name, (which, fields) in self.PP_LISTS.items():
            tot = max(int(sz[which]), 1)
            cnt = np.zeros(n, np.int32)
            arrs = [np.zeros(tot, np.int32) for _ in range(5)]
            self_ = np.zeros(n, np.int32)
            _check(L.fbo_get_pp_list(self.h, which, _ip(cnt), *(_ip(a) for a in arrs), _ip(self_)), "get_pp_list")
            out[name + "_n"] = cnt
            for k, f in enumerate(fields):
                out[name + "_" + f] = arrs[k][:int(sz[which])]
            if which < 3:
                out[name + "_self"] = self_
        for which, key, k in ((0, "SVNL", 4), (1, "VNL", 5)):
            buf = np.zeros(max(int(sz[k]), 1))
            _check(L.fbo_get_pp_blocks(self.h, which, _dp(buf)), "get_pp_blocks")
            out[key] = buf[:int(sz[k])]
        return out

    def run_all(self):
        lists = self.build_lists()
        self.assemble()
        res = dict(lists)
        for m in MATS:
            res[m] = self.blocks(m)
        res["frame_fallbacks"] = lib().fbo_frame_fallbacks(self.h)
        return res
