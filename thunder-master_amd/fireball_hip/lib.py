"""ctypes binding of libfbhip.so: the host-side mirror of the reference's
assembler interface (assemble_S / assemble_T / assemble_dipole_z /
assemble_vna_2c / assemble_vna_3c on a structure, after driver_neighbors).

There is no CPU fallback: if the shared library is missing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfbhip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "fireball_hip.h")
MAXSH = 4

HS_MATS = ("S", "T", "DIP", "VNA", "VNAO")        # the H+S build's block arrays
FB_MAT = {"S": 0, "T": 1, "DIP": 2, "VNA": 3, "VNAO": 4,
          "RHO_IN": 5, "RHO_BOND": 6, "RHOW_IN": 7, "RHOW_BOND": 8}
PREC = {"fp64": 0, "mixed": 1}
METHOD3C = {"table": 0, "mlp": 1}
STATUS = {0: "FB_OK", -1: "FB_E_ARG", -2: "FB_E_X_NEGATIVE", -3: "FB_E_X_OUT_OF_RANGE",
          -4: "FB_E_ATOMS_TOO_CLOSE", -5: "FB_E_NEIGH_SELF", -6: "FB_E_NEIGH_BACK", -7: "FB_E_NO_TABLE",
          -8: "FB_E_ORDER", -9: "FB_E_HIP", -10: "FB_E_CAPACITY", -11: "FB_E_UNSUPPORTED"}

_lib = None


class FireballError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def header_functions(path=HEADER):
    """Names of the functions declared in include/fireball_hip.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fb_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise FireballError(-9, f"{LIB_PATH} is missing: build it with thunder-master_amd/build.py")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, i64p = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_int64)
        fp = C.POINTER(C.c_float)
        sig = {
            "fb_create": [C.POINTER(vp), C.c_int], "fb_destroy": [vp], "fb_set_stream": [vp, vp],
            "fb_synchronize": [vp], "fb_set_precision": [vp, C.c_int], "fb_set_3c_method": [vp, C.c_int],
            "fb_set_species": [vp, C.c_int, ip, ip, ip, dp],
            "fb_set_table_2c": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, ip, ip, ip, dp],
            "fb_set_table_3c": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_double, C.c_int, C.c_double, ip, ip, ip, dp],
            "fb_commit_tables": [vp], "fb_clear_structures": [vp],
            "fb_add_structure": [vp, C.c_int, dp, ip, dp, C.c_int, dp, ip],
            "fb_set_positions": [vp, C.c_int, dp], "fb_set_charges": [vp, C.c_int, dp],
            "fb_build_lists": [vp, i64p, i64p],
            "fb_get_list_sizes": [vp, C.c_int, i64p, i64p, ip],
            "fb_get_neighbors": [vp, C.c_int, ip, ip, ip, ip, ip],
            "fb_get_common": [vp, C.c_int, ip, ip, ip, ip, ip, ip],
            "fb_assemble_S": [vp], "fb_assemble_T": [vp], "fb_assemble_dipole_z": [vp],
            "fb_assemble_vna_2c": [vp], "fb_assemble_vna_3c": [vp], "fb_assemble_all": [vp],
            "fb_get_block_len": [vp, C.c_int, i64p], "fb_get_blocks": [vp, C.c_int, C.c_int, dp],
            "fb_get_device_blocks": [vp, C.c_int, C.POINTER(vp), i64p],
            "fb_enable_timing": [vp, C.c_int],
            "fb_set_tuning": [vp, C.c_int, C.c_int], "fb_get_tuning": [vp, C.c_int, ip],
            "fb_autotune": [vp, C.c_int, ip],
            "fb_get_timings": [vp, C.c_int, C.c_char_p, C.POINTER(C.c_float), ip],
            "fb_gamma_eigenvalues": [vp, C.c_int, C.c_int, dp, C.c_int64, i64p],
            "fb_read_fdata": [vp, C.c_char_p],
            "fb_mlp_train": [vp, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_uint64],
            "fb_mlp_groups": [vp, ip],
            "fb_mlp_samples": [vp, C.c_int, C.c_int64, C.c_uint64, fp, fp],
            "fb_mlp_set_tolerance": [vp, C.c_double],
            "fb_mlp_info": [vp, C.c_int, dp, C.c_int],
            "fb_mlp_get_params": [vp, C.c_int, fp, fp],
            "fb_mlp_stats": [vp, i64p, i64p],
            "fb_mlp_gradients": [vp, C.c_int, fp, C.c_int, fp, fp, fp, dp],
            "fb_mlp_forward": [vp, C.c_int, fp, fp, C.c_int, fp, fp],
            "fb_mlp_gradients_noalloc": [vp, C.c_int, vp, C.c_int, vp, vp, vp, vp],
            "fb_read_positions": [C.c_char_p, C.c_int, ip, ip, C.c_int, ip, ip, dp, dp, ip, ip, ip, ip],
            "fb_set_denmat": [vp, C.c_int, dp],
            "fb_dassemble_S": [vp], "fb_dassemble_T": [vp], "fb_dassemble_dipole_z": [vp],
            "fb_dassemble_vna_2c": [vp], "fb_dassemble_vna_3c": [vp], "fb_dassemble_all": [vp],
            "fb_get_dblocks": [vp, C.c_int, C.c_int, dp],
            "fb_get_forces": [vp, C.c_int, dp, dp, dp, dp, dp],
            "fb_set_qin": [vp, C.c_int, dp],
            "fb_assemble_rho_2c": [vp], "fb_assemble_rho_3c": [vp], "fb_assemble_rho_weighted_2c": [vp],
            "fb_assemble_rho_weighted_3c": [vp], "fb_assemble_rho_all": [vp],
            "fb_get_shell_block_len": [vp, C.c_int, i64p],
            "fb_set_species_pp": [vp, ip, ip, dp, dp],
            "fb_build_lists_pp": [vp, i64p, i64p, i64p],
            "fb_get_pp_list_sizes": [vp, C.c_int, i64p],
            "fb_get_pp_lists": [vp, C.c_int, C.c_int, ip, ip, ip, ip, ip, ip, ip],
            "fb_assemble_svnl": [vp], "fb_assemble_vnl_2c": [vp], "fb_assemble_vnl_3c": [vp],
            "fb_get_pp_blocks": [vp, C.c_int, C.c_int, dp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        L.fb_last_error.argtypes = [vp]
        L.fb_last_error.restype = C.c_char_p
        for name in header_functions():
            if name in ("fb_last_error",) or name in sig:
                continue
            f = getattr(L, name)
            f.restype = C.c_int
        _lib = L
    return _lib


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_positions(path, species, max_atoms=None):
    """fb_read_positions: the in-scope steps of the reference's read_positions
    on a structure file (species matched by atomic number, origin shift,
    norbitals, iblock_slot).  No GPU is needed."""
    L = lib()
    if max_atoms is None:
        with open(path) as f:
            max_atoms = int(f.readline().replace(",", " ").split()[0])
    ns = len(species)
    nZ = np.array([sp.nZ for sp in species], np.int32)
    norb = np.array([sp.norb for sp in species], np.int32)
    n, ic, no, sh = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    lat = np.zeros(9)
    m = max(int(max_atoms), 1)
    ratom = np.zeros(3 * m)
    imass = np.zeros(m, np.int32)
    slot = np.zeros(m, np.int32)
    rc = L.fb_read_positions(str(path).encode(), ns, _ip(nZ), _ip(norb), int(max_atoms), C.byref(n), C.byref(ic),
                             _dp(lat), _dp(ratom), _ip(imass), C.byref(no), _ip(slot), C.byref(sh))
    if rc != 0:
        msg = L.fb_last_error(None)
        raise FireballError(rc, msg.decode() if msg else "")
    k = n.value
    return {"natoms": k, "icluster": ic.value, "shifted": sh.value, "norbitals": no.value,
            "lattice": lat.reshape(3, 3), "imass": imass[:k].copy(), "iblock_slot": slot[:k].copy(),
            "ratom": ratom[:3 * k].reshape(k, 3).copy()}


class Assembler:
    """One context of the GPU assembler: species + tables + a batch of structures."""

    def __init__(self, tables, device=0, precision="fp64", method3c="table", fdata_dir=None):
        """tables: a TableSet whose species and cells are uploaded; with
        fdata_dir, only its species are used and the cells are read from that
        directory in the reference's Fdata layout (fb_read_fdata)."""
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.fb_create(C.byref(self.h), device)
        if rc != 0:
            raise FireballError(rc, "fb_create failed (no usable GPU?)")
        self.tables = tables
        self.structs = []
        self._set_species(tables)
        if fdata_dir is None:
            self._set_tables(tables)
        else:
            self._check(self.L.fb_read_fdata(self.h, os.fsencode(fdata_dir)))
            self._check(self.L.fb_commit_tables(self.h))
        self.set_precision(precision)
        self.set_3c_method(method3c)

    def _check(self, rc):
        if rc != 0:
            msg = self.L.fb_last_error(self.h)
            raise FireballError(rc, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.fb_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _set_species(self, ts):
        ns = len(ts.species)
        nZ = np.array([sp.nZ for sp in ts.species], np.int32)
        nssh = np.array([sp.nssh for sp in ts.species], np.int32)
        lssh = np.zeros((ns, MAXSH), np.int32)
        rc = np.zeros((ns, MAXSH))
        for i, sp in enumerate(ts.species):
            lssh[i, :sp.nssh] = sp.lssh
            rc[i, :sp.nssh] = sp.rcutoffA
        self._check(self.L.fb_set_species(self.h, ns, _ip(nZ), _ip(nssh), _ip(lssh), _dp(rc)))
        if any(getattr(sp, "nssh_pp", 0) for sp in ts.species):
            npp = np.array([sp.nssh_pp for sp in ts.species], np.int32)
            lpp = np.zeros((ns, MAXSH), np.int32)
            clv = np.zeros((ns, MAXSH))
            rpp = np.array([sp.rcutoff_pp for sp in ts.species], np.float64)
            for i, sp in enumerate(ts.species):
                lpp[i, :sp.nssh_pp] = sp.lssh_pp
                clv[i, :sp.nssh_pp] = sp.cl_pp
            self._check(self.L.fb_set_species_pp(self.h, _ip(npp), _ip(lpp), _dp(rpp), _dp(clv)))

    def _set_tables(self, ts):
        for (a, b, it, isub), c in ts.cells2c.items():
            mu, nu, mv = (np.ascontiguousarray(x, np.int32) for x in (c.mu, c.nu, c.mvalue))
            fdata = np.ascontiguousarray(c.data.T, np.float64)      # Fdata_2c(ix, iME), ix fastest
            self._check(self.L.fb_set_table_2c(self.h, a, b, it, isub, c.nME, c.nx, c.xmax, _ip(mu), _ip(nu),
                                               _ip(mv), _dp(fdata)))
        for (a, b, k, it, isub, ith), c in ts.cells3c.items():
            mu, nu, mv = (np.ascontiguousarray(x, np.int32) for x in (c.mu, c.nu, c.mvalue))
            fdata = np.ascontiguousarray(np.transpose(c.data, (2, 0, 1)), np.float64)  # (iME, iy, ix)
            self._check(self.L.fb_set_table_3c(self.h, a, b, k, it, isub, ith, c.nME, c.nx, c.xmax, c.ny, c.ymax,
                                               _ip(mu), _ip(nu), _ip(mv), _dp(fdata)))
        self._check(self.L.fb_commit_tables(self.h))

    def set_precision(self, precision):
        self._check(self.L.fb_set_precision(self.h, PREC[precision]))
        self.precision = precision

    def set_3c_method(self, method):
        self._check(self.L.fb_set_3c_method(self.h, METHOD3C[method]))

    def clear_structures(self):
        self._check(self.L.fb_clear_structures(self.h))
        self.structs = []

    def add_structure(self, st):
        ratom = np.ascontiguousarray(st.ratom, np.float64)
        imass = np.ascontiguousarray(st.imass, np.int32)
        lat = np.ascontiguousarray(st.lattice, np.float64)
        dq = None if st.dQ is None else np.ascontiguousarray(st.dQ[:, :MAXSH], np.float64)
        idx = C.c_int()
        self._check(self.L.fb_add_structure(self.h, st.natoms, _dp(ratom), _ip(imass), _dp(lat), int(st.icluster),
                                            _dp(dq) if dq is not None else None, C.byref(idx)))
        self.structs.append(st)
        return idx.value

    def set_positions(self, istruct, ratom):
        r = np.ascontiguousarray(ratom, np.float64)
        self._check(self.L.fb_set_positions(self.h, istruct, _dp(r)))

    def build_lists(self):
        npairs, ntrip = C.c_int64(), C.c_int64()
        self._check(self.L.fb_build_lists(self.h, C.byref(npairs), C.byref(ntrip)))
        return npairs.value, ntrip.value

    def list_sizes(self, istruct=0):
        p, t, mb = C.c_int64(), C.c_int64(), C.c_int()
        self._check(self.L.fb_get_list_sizes(self.h, istruct, C.byref(p), C.byref(t), C.byref(mb)))
        return p.value, t.value, mb.value

    def lists(self, istruct=0):
        st = self.structs[istruct]
        npairs, ntrip, mbeta_max = self.list_sizes(istruct)
        n = st.natoms
        out = {k: np.zeros(n, np.int32) for k in ("neighn", "neigh_self", "ncommon")}
        for k in ("neigh_j", "neigh_b", "neigh_back"):
            out[k] = np.zeros(max(npairs, 1), np.int32)
        for k in ("com_iatom_j", "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh"):
            out[k] = np.zeros(max(ntrip, 1), np.int32)
        self._check(self.L.fb_get_neighbors(self.h, istruct, _ip(out["neighn"]), _ip(out["neigh_j"]),
                                            _ip(out["neigh_b"]), _ip(out["neigh_back"]), _ip(out["neigh_self"])))
        self._check(self.L.fb_get_common(self.h, istruct, _ip(out["ncommon"]), _ip(out["com_iatom_j"]),
                                         _ip(out["com_iatom_b"]), _ip(out["com_jatom_j"]), _ip(out["com_jatom_b"]),
                                         _ip(out["com_mneigh"])))
        for k in ("neigh_j", "neigh_b", "neigh_back"):
            out[k] = out[k][:npairs]
        for k in ("com_iatom_j", "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh"):
            out[k] = out[k][:ntrip]
        out.update(npairs=npairs, ntriplets=ntrip, mbeta_max=mbeta_max)
        return out

    def assemble_S(self):
        self._check(self.L.fb_assemble_S(self.h))

    def assemble_T(self):
        self._check(self.L.fb_assemble_T(self.h))

    def assemble_dipole_z(self):
        self._check(self.L.fb_assemble_dipole_z(self.h))

    def assemble_vna_2c(self):
        self._check(self.L.fb_assemble_vna_2c(self.h))

    def assemble_vna_3c(self):
        self._check(self.L.fb_assemble_vna_3c(self.h))

    def assemble_all(self, sync=True):
        """One H+S build; with sync=False it only enqueues (errors surface at
        the next synchronize / blocks / assemble call)."""
        self._check(self.L.fb_assemble_all(self.h))
        if sync:
            self.synchronize()

    def synchronize(self):
        self._check(self.L.fb_synchronize(self.h))

    def blocks(self, mat, istruct=0):
        n = C.c_int64()
        self._check(self.L.fb_get_block_len(self.h, istruct, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._check(self.L.fb_get_blocks(self.h, istruct, FB_MAT[mat], _dp(out)))
        return out[:n.value]

    # ---------------------------------------------------------------- forces
    def set_denmat(self, rho, istruct=0):
        """Density blocks of one structure (s%denmat), block layout of blocks()."""
        n = C.c_int64()
        self._check(self.L.fb_get_block_len(self.h, istruct, C.byref(n)))
        rho = np.ascontiguousarray(rho, np.float64)
        if rho.size != n.value:
            raise FireballError(-1, f"density matrix has {rho.size} values, the blocks {n.value}")
        self._check(self.L.fb_set_denmat(self.h, istruct, _dp(rho)))

    def dassemble(self, which=("S", "T", "DIP", "VNA2C", "VNA3C")):
        fn = {"S": self.L.fb_dassemble_S, "T": self.L.fb_dassemble_T, "DIP": self.L.fb_dassemble_dipole_z,
              "VNA2C": self.L.fb_dassemble_vna_2c, "VNA3C": self.L.fb_dassemble_vna_3c}
        for w in which:
            self._check(fn[w](self.h))

    def dassemble_all(self):
        self._check(self.L.fb_dassemble_all(self.h))

    def dblocks(self, mat, istruct=0):
        n = C.c_int64()
        self._check(self.L.fb_get_block_len(self.h, istruct, C.byref(n)))
        out = np.zeros(max(3 * n.value, 1))
        self._check(self.L.fb_get_dblocks(self.h, istruct, FB_MAT[mat], _dp(out)))
        return out[:3 * n.value]

    def forces(self, istruct=0):
        npairs = self.list_sizes(istruct)[0]
        n = self.structs[istruct].natoms
        fa, fo = np.zeros((max(npairs, 1), 3)), np.zeros((max(npairs, 1), 3))
        f3 = [np.zeros((n, 3)) for _ in range(3)]
        self._check(self.L.fb_get_forces(self.h, istruct, _dp(fa), _dp(fo), *(_dp(x) for x in f3)))
        return {"vna_atom": fa[:npairs], "vna_ontop": fo[:npairs], "f3naa": f3[0], "f3nab": f3[1], "f3nac": f3[2]}

    # ------------------------------------------------------ McWEDA densities
    def set_qin(self, qin, istruct=0):
        """Input shell charges s%atom(:)%shell(:)%Qin of one structure, [natoms][4]."""
        qin = np.ascontiguousarray(np.asarray(qin, np.float64)[:, :MAXSH])
        if qin.shape != (self.structs[istruct].natoms, MAXSH):
            raise FireballError(-1, f"Qin must be [natoms][{MAXSH}]")
        self._check(self.L.fb_set_qin(self.h, istruct, _dp(qin)))

    def assemble_rho(self, which=("RHO2C", "RHO3C", "RHOW2C", "RHOW3C")):
        """assemble_rho_2c, _3c, assemble_rho_weighted_2c, _3c (in this order in
        j.ASSEMBLERS/M_assemble_vxc_McWEDA_DOGS.f90:148-154)."""
        fn = {"RHO2C": self.L.fb_assemble_rho_2c, "RHO3C": self.L.fb_assemble_rho_3c,
              "RHOW2C": self.L.fb_assemble_rho_weighted_2c, "RHOW3C": self.L.fb_assemble_rho_weighted_3c,
              "ALL": self.L.fb_assemble_rho_all}
        for w in which:
            self._check(fn[w](self.h))

    def rho(self, which, istruct=0):
        """RHO_IN / RHO_BOND (orbital blocks, block layout) or RHOW_IN / RHOW_BOND
        (shell blocks nssh_mu x nssh_nu per pair, column-major, pair order)."""
        n = C.c_int64()
        if which.startswith("RHOW"):
            self._check(self.L.fb_get_shell_block_len(self.h, istruct, C.byref(n)))
        else:
            self._check(self.L.fb_get_block_len(self.h, istruct, C.byref(n)))
        out = np.zeros(max(n.value, 1))
        self._check(self.L.fb_get_blocks(self.h, istruct, FB_MAT[which], _dp(out)))
        return out[:n.value]

    # ------------------------------------------------------- pseudopotential
    PP_LISTS = {"pp": (0, ("j", "b", "map")), "ppx": (1, ("j", "b", "map", "point")), "ppp": (2, ("j", "b")),
                "ppc": (3, ("ij", "ib", "jj", "jb", "m"))}

    def build_lists_pp(self):
        """driver_neighbors_PP for every structure; returns (PP pairs, PPp pairs, PP triplets)."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.L.fb_build_lists_pp(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def pp_sizes(self, istruct=0):
        sz = (C.c_int64 * 6)()
        self._check(self.L.fb_get_pp_list_sizes(self.h, istruct, sz))
        return [int(x) for x in sz]

    def pp_lists(self, istruct=0):
        """The PP lists of one structure, keyed like the oracle's run_pp()."""
        sz = self.pp_sizes(istruct)
        n = self.structs[istruct].natoms
        out = {}
        for name, (which, fields) in self.PP_LISTS.items():
            tot = max(sz[which], 1)
            cnt = np.zeros(n, np.int32)
            arrs = [np.zeros(tot, np.int32) for _ in range(5)]
            self_ = np.zeros(n, np.int32)
            ptr = [_ip(a) if k < len(fields) else None for k, a in enumerate(arrs)]
            self._check(self.L.fb_get_pp_lists(self.h, istruct, which, _ip(cnt), *ptr,
                                               _ip(self_) if which < 3 else None))
            out[name + "_n"] = cnt
            for k, f in enumerate(fields):
                out[name + "_" + f] = arrs[k][:sz[which]]
            if which < 3:
                out[name + "_self"] = self_
        return out

    def assemble_pp(self, which=("SVNL", "VNL2C", "VNL3C")):
        """assemble_svnl, assemble_vnl_2c, assemble_vnl_3c (in this order)."""
        fn = {"SVNL": self.L.fb_assemble_svnl, "VNL2C": self.L.fb_assemble_vnl_2c, "VNL3C": self.L.fb_assemble_vnl_3c}
        for w in which:
            self._check(fn[w](self.h))

    def pp_blocks(self, which, istruct=0):
        """SVNL (per PP pair, norb_mu x norb_PP_nu) or VNL (per PPp pair), flat, column-major."""
        sz = self.pp_sizes(istruct)
        n = sz[4] if which == "SVNL" else sz[5]
        out = np.zeros(max(n, 1))
        self._check(self.L.fb_get_pp_blocks(self.h, istruct, 0 if which == "SVNL" else 1, _dp(out)))
        return out[:n]

    def gamma_eigenvalues(self, which="H", istruct=0):
        """Eigenvalues (ascending) of the Gamma-point folded S or
        H = T + Vna block + Vna blocko of one structure, on the GPU."""
        st = self.structs[istruct]
        norb = np.array([sp.norb for sp in self.tables.species])
        n = C.c_int64(int(norb[np.asarray(st.imass) - 1].sum()))
        w = np.zeros(max(n.value, 1))
        self._check(self.L.fb_gamma_eigenvalues(self.h, istruct, {"S": 0, "H": 1}[which], _dp(w), n.value,
                                                C.byref(n)))
        return w[:n.value]

    # ---------------------------------------------------- learned 3c surrogate
    MLP_NP = 8576
    MLP_INFO = ("is", "js", "ks", "isub", "trained", "nsamples", "nsteps", "loss_before", "loss_after",
                "max_cell_err_eV", "train_ms", "calib_samples")

    def train_mlp(self, nsamples=1 << 21, nsteps=40000, batch=8192, lr=3e-3, seed=7):
        """Train one network per neutral 3c group on samples of the table path
        (GPU forward/backward/Adam kernels), then calibrate its error."""
        self._check(self.L.fb_mlp_train(self.h, int(nsamples), int(nsteps), int(batch), float(lr), int(seed)))

    def mlp_groups(self):
        n = C.c_int()
        self._check(self.L.fb_mlp_groups(self.h, C.byref(n)))
        return n.value

    def mlp_info(self, group):
        out = np.zeros(12)
        self._check(self.L.fb_mlp_info(self.h, int(group), _dp(out), 12))
        return dict(zip(self.MLP_INFO, out.tolist()))

    def mlp_params(self, group):
        p = np.zeros(self.MLP_NP, np.float32)
        osc = np.zeros(64, np.float32)
        self._check(self.L.fb_mlp_get_params(self.h, int(group), _fp(p), _fp(osc)))
        return p, osc

    def mlp_samples(self, group, n, seed=1):
        X = np.zeros((n, 3), np.float32)
        T = np.zeros((n, 64), np.float32)
        self._check(self.L.fb_mlp_samples(self.h, int(group), int(n), int(seed), _fp(X), _fp(T)))
        return X, T

    def set_mlp_tolerance(self, tau_ev):
        self._check(self.L.fb_mlp_set_tolerance(self.h, float(tau_ev)))

    def mlp_stats(self):
        a, b = C.c_int64(), C.c_int64()
        self._check(self.L.fb_mlp_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def mlp_gradients(self, nme, params, X, T):
        p = np.ascontiguousarray(params, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.MLP_NP, np.float32)
        loss = C.c_double()
        self._check(self.L.fb_mlp_gradients(self.h, int(nme), _fp(p), X.shape[0], _fp(X), _fp(T), _fp(g),
                                            C.byref(loss)))
        return loss.value, g

    def mlp_gradients_noalloc(self, nme, params_ptr, n, X_ptr, T_ptr, grads_ptr, loss_ptr):
        """fb_mlp_gradients_noalloc on device arrays (integer device addresses:
        params [8576], X [n][3], T [n][64], grads [8576], loss [1], float32).
        Enqueued on the context's stream; call synchronize() before reading."""
        self._check(self.L.fb_mlp_gradients_noalloc(self.h, int(nme), C.c_void_p(params_ptr), int(n),
                                                    C.c_void_p(X_ptr), C.c_void_p(T_ptr),
                                                    C.c_void_p(grads_ptr), C.c_void_p(loss_ptr)))

    def mlp_forward(self, nme, params, X, oscale=None):
        p = np.ascontiguousarray(params, np.float32)
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], 64), np.float32)
        osc = None if oscale is None else np.ascontiguousarray(oscale, np.float32)
        self._check(self.L.fb_mlp_forward(self.h, int(nme), _fp(p), _fp(osc) if osc is not None else None,
                                          X.shape[0], _fp(X), _fp(Y)))
        return Y

    def tuning(self, param):
        v = C.c_int()
        self._check(self.L.fb_get_tuning(self.h, int(param), C.byref(v)))
        return v.value

    def set_tuning(self, param, value):
        self._check(self.L.fb_set_tuning(self.h, int(param), int(value)))

    def set_chunk3c(self, chunk):
        self._check(self.L.fb_set_tuning(self.h, 0, int(chunk)))

    def chunk3c(self):
        v = C.c_int()
        self._check(self.L.fb_get_tuning(self.h, 0, C.byref(v)))
        return v.value

    def autotune(self, nrep=5):
        """Time every workgroup chunk of the mixed 3c kernel on the current
        lists and keep the fastest; returns the chosen chunk."""
        v = C.c_int()
        self._check(self.L.fb_autotune(self.h, int(nrep), C.byref(v)))
        return v.value

    def enable_timing(self, on=True):
        self._check(self.L.fb_enable_timing(self.h, int(on)))

    def timings(self, maxn=64):
        names = C.create_string_buffer(32 * maxn)
        ms = (C.c_float * maxn)()
        cnt = C.c_int()
        self._check(self.L.fb_get_timings(self.h, maxn, names, ms, C.byref(cnt)))
        raw = names.raw
        return [(raw[32 * i:32 * i + 32].split(b"\0")[0].decode(), float(ms[i])) for i in range(cnt.value)]

    def run_all(self, istruct=0):
        """Lists + every matrix for one structure, as a dict shaped like the oracle's."""
        res = self.lists(istruct)
        for m in HS_MATS:
            res[m] = self.blocks(m, istruct)
        return res
