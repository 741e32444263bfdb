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

FB_MAT = {"S": 0, "T": 1, "DIP": 2, "VNA": 3, "VNAO": 4}
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


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Assembler:
    """One context of the GPU assembler: species + tables + a batch of structures."""

    def __init__(self, tables, device=0, precision="fp64", method3c="table"):
        self.L = lib()
        self.h = C.c_void_p()
        rc = self.L.fb_create(C.byref(self.h), device)
        if rc != 0:
            raise FireballError(rc, "fb_create failed (no usable GPU?)")
        self.tables = tables
        self.structs = []
        self._set_species(tables)
        self._set_tables(tables)
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
        for m in FB_MAT:
            res[m] = self.blocks(m, istruct)
        return res
