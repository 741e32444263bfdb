"""Benchmark of the FIREBALL 2c/3c H+S assembly on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 1] [--precision mixed|fp64]

A step is one H+S build of the workload -- S, T, dipole_z, Vna-2c (ontop +
atom) and Vna-3c blocks for every pair of the structure -- with positions,
tables and lists already resident in HBM, replayed from the library's
captured HIP graph (fb_assemble_all, the default path).  Per-kernel times come
from a separate pass with HIP events around every launch (launches then run
one after another).  The default workload is the
4096-atom periodic Si supercell (8x8x8 diamond cells, displacements
N(0, 0.08 A), sp3 basis, rc = 5.4 bohr, synthetic smooth 2c/3c Fdata).
``value`` is 3c triplets per second of whole H+S builds; ms_per_step is the H+S
build time.  The plain binary64 kernels run beside the selected precision and
the per-matrix max-abs difference is reported.

Prints one JSON line.  Only one GPU is supported: --gpus N > 1 exits.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "thunder-master_amd"))

from fireball_hip import synth  # noqa: E402
from fireball_hip.lib import Assembler  # noqa: E402

METRIC = "3c triplets/s + H+S build ms at <=1e-6 eV err; FP32/MLP vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and
# scaled by the same counters over a copy of known size; one line per kernel:
#   traffic <launch name>:<3c chunk> <workload config> <bytes per launch>
TRAFFIC_FILE = os.path.join(REPO, "profiles", "hbm_traffic.txt")
WORKLOADS = {
    0: "Si64 periodic diamond 2x2x2 (a=5.43 A, sigma 0.05 A, seed 0), sp3, rc 5.4 bohr",
    1: "Si4096 periodic diamond 8x8x8 (a=5.43 A, sigma 0.08 A, seed 0), sp3, rc 5.4 bohr",
    2: "CHNO3000 random box 1.0 g/cm3 (C:H:N:O 30:50:10:10, seed 1), rc 4.5/5.0 bohr",
    3: "batch of 512 structures (16..1024 atoms, half Si, half CHNO, seed 2)",
    4: "aC8192 random box 2.0 g/cm3 (seed 3), sp3, rc 5.0 bohr",
}
MATS = ("S", "T", "DIP", "VNA", "VNAO")
# launch shape of a neutral s/p mixed build (tuning knob 4)
LAUNCH = {0: "2c and 3c launches on two streams", 3: "one launch, 3 waves/SIMD", 4: "one launch, 4 waves/SIMD"}


def launch_name(asm):
    """The H+S launch shape the autotuner chose (FB_TUNE_FUSED, FB_TUNE_PERSIST)."""
    fused = asm.tuning(4)
    return LAUNCH[fused] + (", 3c waves on a tile queue" if fused and asm.tuning(5) else "")


def algorithmic_bytes_3c(asm, nstruct):
    """Algorithmic HBM bytes of one 3c launch: 24 B per triplet (six int32:
    ialpha, iatom, jatom, the two cells and the target pair) plus one binary64
    read-modify-write (16 B) per element of every targeted Vna block, the
    accumulate-into-existing-blocks semantics of assemble_vna_3c."""
    ntrip = 0
    tgt_bytes = 0
    for k in range(nstruct):
        L = asm.lists(k)
        ntrip += L["ntriplets"]
        if L["ntriplets"] == 0:
            continue
        st = asm.structs[k]
        norb = np.array([sp.norb for sp in asm.tables.species])
        off = np.concatenate([[0], np.cumsum(L["neighn"])])
        atom_of_pair = np.repeat(np.arange(st.natoms), L["neighn"])
        tgt = np.unique(off[L["com_iatom_j"] - 1] + L["com_mneigh"] - 1)
        nmu = norb[st.imass[atom_of_pair[tgt]] - 1]
        nnu = norb[st.imass[L["neigh_j"][tgt] - 1] - 1]
        tgt_bytes += int(np.sum(nmu * nnu * 16))
    return 24 * ntrip + tgt_bytes, ntrip


def time_builds(asm, steps, warmup):
    """K back-to-back H+S builds (enqueued asynchronously) with every kernel
    bracketed by HIP events on the launch stream; returns wall ms per build and
    the mean duration of each kernel (launches serialised by the events)."""
    asm.enable_timing(True)
    for _ in range(warmup):
        asm.assemble_all(sync=False)
    asm.synchronize()
    asm.timings(4096)                     # drop the warm-up events
    kern = {}
    t0 = time.perf_counter()
    for _ in range(steps):
        asm.assemble_all(sync=False)
    asm.synchronize()
    t1 = time.perf_counter()
    for name, ms in asm.timings(4096):
        kern.setdefault(name, []).append(ms)
    asm.enable_timing(False)
    return (t1 - t0) * 1e3 / steps, {k: float(np.mean(v)) for k, v in kern.items() if k != "clear_status"}


def time_graph(asm, steps, warmup):
    """K H+S builds replayed from the captured HIP graph, no per-kernel events."""
    for _ in range(warmup):
        asm.assemble_all(sync=False)
    asm.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        asm.assemble_all(sync=False)
    asm.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def batch_leg(steps, warmup, config=3):
    """The 512-structure batch (config 3) as one H+S build: every structure's
    2c and 3c blocks in the same launches (the reference assembles the
    structures one after another, M_configuraciones.f90:375-381), autotuned,
    replayed from the captured graph; structures/s and triplets/s."""
    species, structs = synth.config(config)
    asm = Assembler(synth.make_tables(species), precision="mixed")
    try:
        for st in structs:
            asm.add_structure(st)
        npairs, ntrip = asm.build_lists()
        asm.assemble_all()
        chunk = asm.autotune(20)
        ms = time_graph(asm, steps, warmup)
        return {"workload": WORKLOADS[config], "nstructures": len(structs),
                "natoms": int(sum(s.natoms for s in structs)), "npairs": npairs, "ntriplets": ntrip,
                "ms_per_step": ms, "structures_per_s": len(structs) / (ms * 1e-3),
                "triplets_per_s": ntrip / (ms * 1e-3), "chunk3c": chunk,
                "kernel_2c": ("LDS-staged", "register")[asm.tuning(1)], "launch": launch_name(asm)}
    finally:
        asm.close()


def config_leg(steps, warmup, config=2):
    """Another single-structure workload (config 2, CHNO3000) beside the main
    one: the autotuned mixed-precision H+S build and the binary64 build, both
    replayed from their captured graphs, and the largest block difference
    between the two."""
    species, structs = synth.config(config)
    asm = Assembler(synth.make_tables(species), precision="mixed")
    try:
        for st in structs:
            asm.add_structure(st)
        npairs, ntrip = asm.build_lists()
        asm.assemble_all()
        chunk = asm.autotune(20)
        ms = time_graph(asm, steps, warmup)
        mixed = {m: asm.blocks(m) for m in MATS}
        launch = launch_name(asm)
        asm.set_precision("fp64")
        ms64 = time_graph(asm, steps, warmup)
        err = {m: float(np.max(np.abs(mixed[m] - asm.blocks(m)))) for m in MATS}
        return {"workload": WORKLOADS[config], "natoms": int(sum(s.natoms for s in structs)),
                "npairs": npairs, "ntriplets": ntrip, "ms_per_step": ms, "triplets_per_s": ntrip / (ms * 1e-3),
                "chunk3c": chunk, "launch": launch, "fp64_ms_per_step": ms64, "speedup_vs_fp64": ms64 / ms,
                "max_abs_err_vs_fp64_eV": err}
    finally:
        asm.close()


def gamma_eigen_shift(asm, precision_a, precision_b, max_norb=20000):
    """Eigenvalue shift of the Gamma-point folded H = T + Vna (block + blocko)
    of structure 0 between two precisions, from the library's own fold and
    rocSOLVER dsyevd on the upper triangle; returns max |lambda_a - lambda_b|
    in eV and the orbital count (None above max_norb orbitals, where the dense
    solve alone takes tens of seconds).  S is binary64 in both paths."""
    st = asm.structs[0]
    norb = np.array([sp.norb for sp in asm.tables.species])
    if int(norb[np.asarray(st.imass) - 1].sum()) > max_norb:
        return None
    lam = {}
    for prec in (precision_a, precision_b):
        asm.set_precision(prec)
        asm.assemble_all()
        lam[prec] = asm.gamma_eigenvalues("H", 0)
    asm.set_precision(precision_a)
    return {"H": float(np.max(np.abs(lam[precision_a] - lam[precision_b]))), "norb": int(lam[precision_a].size),
            "lambda_range_eV": [float(lam[precision_b][0]), float(lam[precision_b][-1])]}


# MFMA work of one network evaluation as the kernel executes it: layer 1 on a
# 16x16x16 tile (K = 16 slots for the split inputs), layer 2 (64 x 64), layer 3
# three times (high x high, low x high, high x low parts)
MLP_MFMA_FLOPS_PER_EVAL = 2 * (64 * 16 + 64 * 64 + 3 * 64 * 64)
MFMA_BF16_PEAK_TFLOPS = 2500.0
MLP_NSTEPS = 40000


def mlp_flops_algorithmic(nme):
    """Flops of one evaluation by the workload definition: 2 (3*64 + 64*64 + 64*5*nME)."""
    return 2 * (3 * 64 + 64 * 64 + 64 * 5 * nme)


def _kernel_ms(asm, name, n):
    asm.assemble_all()
    asm.enable_timing(True)
    asm.timings(4096)
    for _ in range(n):
        asm.assemble_all()
    t = [ms for k, ms in asm.timings(4096) if k == name]
    asm.enable_timing(False)
    return float(np.mean(t)) if t else float("nan")


def mlp_leg(steps, warmup, config=4):
    """The learned 3c surrogate on its own workload (aC8192, config 4): the
    binary64 build as the reference, the mixed table build, then the network
    trained on the workload's tables (2M samples, 40000 steps) and the same H+S
    build with it (table stencil wherever the calibrated bound exceeds the
    tolerance).  Reports the 3c kernel times, the fraction of triplets on the
    network, the error against binary64 and the MFMA rate."""
    species, structs = synth.config(config)
    ts = synth.make_tables(species)
    ref = Assembler(ts, precision="fp64")
    for st in structs:
        ref.add_structure(st)
    ref.build_lists()
    ref.assemble_all()
    vna64 = np.concatenate([ref.blocks("VNA", k) for k in range(len(structs))])
    ref.close()
    asm = Assembler(ts, precision="mixed")
    for st in structs:
        asm.add_structure(st)
    npairs, ntrip = asm.build_lists()
    asm.assemble_all()
    asm.autotune(20)
    table_ms = time_graph(asm, steps, warmup)
    table_kernel_ms = _kernel_ms(asm, "assemble_3c_mixed", steps)
    t0 = time.perf_counter()
    asm.set_3c_method("mlp")
    asm.train_mlp(nsamples=1 << 21, nsteps=MLP_NSTEPS, batch=8192, lr=3e-3, seed=7)
    train_s = time.perf_counter() - t0
    asm.assemble_all()
    asm.mlp_stats()
    asm.assemble_all()
    n_net, n_seen = asm.mlp_stats()
    err = float(np.max(np.abs(np.concatenate([asm.blocks("VNA", k) for k in range(len(structs))]) - vna64)))
    kms = _kernel_ms(asm, "assemble_3c_mlp", steps)
    ms = time_graph(asm, steps, warmup)
    groups = [asm.mlp_info(gid) for gid in range(asm.mlp_groups())]
    trained = [g for g in groups if g["trained"] == 1]
    nme = 10
    executed = n_net * MLP_MFMA_FLOPS_PER_EVAL / (kms * 1e-3) / 1e12
    algorithmic = n_net * mlp_flops_algorithmic(nme) / (kms * 1e-3) / 1e12
    asm.close()
    return {"workload": WORKLOADS[config], "npairs": int(npairs), "ntriplets": int(ntrip),
            "train_s": train_s, "train_steps": MLP_NSTEPS, "networks": len(trained),
            "loss_after_max": max((g["loss_after"] for g in trained), default=None),
            "max_subcell_err_eV": max((g["max_cell_err_eV"] for g in trained), default=None),
            "tolerance_eV": 2.5e-5, "fraction_on_network": n_net / max(n_seen, 1),
            "ms_per_step": ms, "table_ms_per_step": table_ms, "kernel_ms": kms, "table_kernel_ms": table_kernel_ms,
            "speedup_vs_table_kernel": table_kernel_ms / kms, "speedup_vs_table_build": table_ms / ms,
            "max_abs_err_vs_fp64_eV": err,
            "mfma": {"executed_tflops": executed, "algorithmic_tflops": algorithmic,
                     "peak_tflops": MFMA_BF16_PEAK_TFLOPS, "utilisation": executed / MFMA_BF16_PEAK_TFLOPS,
                     "mfma_flops_per_eval": MLP_MFMA_FLOPS_PER_EVAL,
                     "algorithmic_flops_per_eval": mlp_flops_algorithmic(nme)}}


FP64_PEAK_TFLOPS = 78.6
# FP64 operations per triplet of the neutral 3c force kernel (FMA counted as 2),
# from rocprofv3 --pmc SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F64 over one launch on
# the 4096-atom Si workload; committed with the counter summary
FORCE3C_FLOPS_FILE = os.path.join(REPO, "profiles", "forces_fp64_ops.txt")


def force3c_flops_per_triplet():
    try:
        with open(FORCE3C_FLOPS_FILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 2 and w[0] == "fp64_ops_per_triplet":
                    return float(w[1])
    except OSError:
        pass
    return None


def _force_kernel_ms(asm, n):
    asm.enable_timing(True)
    asm.timings(4096)
    for _ in range(n):
        asm.dassemble_all()
    asm.synchronize()
    kern = {}
    for name, t in asm.timings(4096):
        kern.setdefault(name, []).append(t)
    asm.enable_timing(False)
    return {k: float(np.mean(v)) for k, v in kern.items() if k != "clear_status"}


def forces_leg(asm, structs, steps, warmup):
    """One force build = fb_dassemble_all (Dblocks of S, T and dipole_z, the
    vna_atom / vna_ontop pair forces and the per-atom f3naa/f3nab/f3nac of
    Dassemble_vna_3c, binary64) on the current H+S blocks, with a synthetic
    symmetric density matrix; wall time over K builds and per-kernel times from
    an event-bracketed pass."""
    for k, st in enumerate(structs):
        L = asm.lists(k)
        asm.set_denmat(synth.denmat(asm.tables.species, st.imass, L["neighn"], L["neigh_j"], L["neigh_back"]), k)
    for _ in range(warmup):
        asm.dassemble_all()
    asm.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        asm.dassemble_all()
    asm.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    kern = _force_kernel_ms(asm, max(1, min(steps, 5)))
    kern64 = kern
    ntrip = sum(asm.list_sizes(k)[1] for k in range(len(structs)))
    prec = asm.precision
    out = {"ms_per_step": ms, "triplets_per_s": ntrip / (ms * 1e-3), "kernels_ms": kern,
           "dtype": "f64" if prec == "fp64" else "f32 interpolation of the 3c forces, f64 otherwise"}
    if prec != "fp64":
        # the binary64 force build beside it, and the error of the mixed one
        got = [asm.forces(k) for k in range(len(structs))]
        asm.set_precision("fp64")
        for _ in range(warmup):
            asm.dassemble_all()
        asm.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            asm.dassemble_all()
        asm.synchronize()
        ms64 = (time.perf_counter() - t0) * 1e3 / steps
        kern64 = _force_kernel_ms(asm, max(1, min(steps, 5)))
        ref = [asm.forces(k) for k in range(len(structs))]
        asm.set_precision(prec)
        err = {key: float(max(np.max(np.abs(g[key] - r[key])) for g, r in zip(got, ref)))
               for key in ("vna_atom", "vna_ontop", "f3naa", "f3nab", "f3nac")}
        out.update(fp64_ms_per_step=ms64, fp64_kernels_ms=kern64, speedup_vs_fp64=ms64 / ms,
                   max_abs_err_vs_fp64_eV_per_A=err)
    fl = force3c_flops_per_triplet()
    k3 = kern64.get("Dassemble_vna_3c")
    if fl and k3:
        # the binary64 kernel, whose FP64 operation count the counter pass measured
        tf = fl * ntrip / (k3 * 1e-3) / 1e12
        out["roofline_3c_fp64"] = {"bound": "fp64", "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": tf / FP64_PEAK_TFLOPS, "ops_per_triplet": fl, "kernel_ms": k3}
    return out


def rho_leg(species, structs, steps, warmup):
    """One density build = fb_assemble_rho_all (assemble_rho_2c, _3c,
    assemble_rho_weighted_2c, _3c of the McWEDA module, binary64) on the same
    structures with the synthetic density tables and shell charges Qin =
    Qneutral + dQ; its own context, so the H+S measurement above is untouched."""
    ts = synth.add_rho_tables(synth.make_tables(species), seed=4048)
    a = Assembler(ts, precision="fp64")
    for st in structs:
        a.add_structure(st)
    a.build_lists()
    for k, st in enumerate(structs):
        a.set_qin(synth.shell_qin(st, species, seed=5), k)
    for _ in range(warmup):
        a.assemble_rho(("ALL",))
    a.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a.assemble_rho(("ALL",))
    a.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    a.enable_timing(True)
    a.timings(4096)
    for _ in range(max(1, min(steps, 5))):
        a.assemble_rho(("ALL",))
    a.synchronize()
    kern = {}
    for name, t in a.timings(4096):
        kern.setdefault(name, []).append(t)
    a.enable_timing(False)
    kern = {k: float(np.mean(v)) for k, v in kern.items() if k != "clear_status"}
    ntrip = sum(a.list_sizes(k)[1] for k in range(len(structs)))
    a.close()
    return {"ms_per_step": ms, "triplets_per_s": ntrip / (ms * 1e-3), "kernels_ms": kern, "dtype": "f64"}


def pp_leg(species, structs, steps, warmup):
    """One pseudopotential build = assemble_svnl + assemble_vnl_2c +
    assemble_vnl_3c (binary64) on the same structures with the synthetic PP
    shells and P_vnl tables, after driver_neighbors_PP (timed once, like the
    H+S lists); a context of its own."""
    sp = synth.with_pp(species)
    ts = synth.add_pp_tables(synth.make_tables(sp), seed=6066)
    a = Assembler(ts, precision="fp64")
    for st in structs:
        a.add_structure(st)
    a.build_lists()
    t0 = time.perf_counter()
    npp, nppp, ntrip = a.build_lists_pp()
    lists_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    a.build_lists_pp()
    a.synchronize()
    lists_rebuild_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(warmup):
        a.assemble_pp()
    a.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a.assemble_pp()
    a.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    a.enable_timing(True)
    a.timings(4096)
    for _ in range(max(1, min(steps, 5))):
        a.assemble_pp()
    a.synchronize()
    kern = {}
    for name, t in a.timings(4096):
        kern.setdefault(name, []).append(t)
    a.enable_timing(False)
    kern = {k: float(np.mean(v)) for k, v in kern.items() if k != "clear_status"}
    a.close()
    return {"ms_per_step": ms, "pp_triplets_per_s": ntrip / (ms * 1e-3), "kernels_ms": kern, "dtype": "f64",
            "pp_pairs": npp, "ppp_pairs": nppp, "pp_triplets": ntrip, "lists_pp_ms": lists_ms,
            "lists_pp_rebuild_ms": lists_rebuild_ms}


def pp_cpu_baseline():
    """The reference's PP assemblers (serial flang build) on the Si512 sample."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import refrun
    if not refrun.available(refrun.HARNESS_PREF):
        return None
    st = synth.diamond_si((4, 4, 4), sigma=0.08, seed=0)
    sp = synth.with_pp([synth.SI])
    ts = synth.add_pp_tables(synth.make_tables(sp), seed=6066)
    nrep = 3
    r = refrun.run_pp(ts, st, nrep=nrep)
    t = sum(r["times"].values()) / nrep
    ntrip = int(len(r["ppc_m"]))
    return {"value": ntrip / t, "unit": "PP triplets/s", "cores": 1, "kind": "reference",
            "sample": f"Si512 4x4x4 diamond (sigma 0.08 A): {ntrip} PP triplets, PP build {t * 1e3:.1f} ms "
                      f"(assemble_svnl, assemble_vnl_2c, assemble_vnl_3c, serial flang -O2, mean of {nrep})",
            "pp_build_ms": t * 1e3}


def rho_cpu_baseline():
    """The reference's McWEDA density assemblers (serial flang build) on the
    bounded Si512 sample of the other baselines."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import refrun
    if not refrun.available(refrun.HARNESS_RREF):
        return None
    st = synth.diamond_si((4, 4, 4), sigma=0.08, seed=0)
    ts = synth.add_rho_tables(synth.make_tables([synth.SI]), seed=4048)
    lists = refrun.run(ts, st)
    nrep = 3
    r = refrun.run_rho(ts, st, synth.shell_qin(st, [synth.SI], seed=5), lists, nrep=nrep)
    t = sum(r["times"].values()) / nrep
    return {"value": r["ntriplets"] / t, "unit": "triplets/s", "cores": 1, "kind": "reference",
            "sample": f"Si512 4x4x4 diamond (sigma 0.08 A): {r['ntriplets']} triplets, density build "
                      f"{t * 1e3:.1f} ms (assemble_rho_2c/3c, assemble_rho_weighted_2c/3c, serial flang -O2, "
                      f"mean of {nrep})",
            "rho_build_ms": t * 1e3}


def force_cpu_baseline():
    """The reference's force assemblers (serial flang build) on the same
    bounded sample as the H+S baseline: Dassemble_S/T/dipole_z/vna_2c/vna_3c."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import refrun
    if not refrun.available(refrun.HARNESS_DREF):
        return None
    st = synth.diamond_si((4, 4, 4), sigma=0.08, seed=0)
    ts = synth.make_tables([synth.SI])
    lists = refrun.run(ts, st)
    rho = synth.denmat(ts.species, st.imass, lists["neighn"], lists["neigh_j"], lists["neigh_back"])
    nrep = 3
    r = refrun.run_forces(ts, st, rho, lists, nrep=nrep)
    t = sum(r["times"].values()) / nrep
    return {"value": r["ntriplets"] / t, "unit": "triplets/s", "cores": 1, "kind": "reference",
            "sample": f"Si512 4x4x4 diamond (sigma 0.08 A): {r['ntriplets']} triplets, force build "
                      f"{t * 1e3:.1f} ms (Dassemble_S/T/dipole_z/vna_2c/vna_3c, serial flang -O2, mean of {nrep})",
            "force_build_ms": t * 1e3}


def measured_traffic(kernel, config, chunk):
    """HBM bytes per launch of `kernel` (at 3c chunk `chunk`) on workload
    `config` from the committed counter summary, or None when it holds no
    such line."""
    key = f"{kernel}:{chunk}" if chunk else kernel
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == "traffic" and w[1] == key and w[2] == str(config):
                    return float(w[3])
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """One device-to-device elementwise copy of `mib` MiB (reads and writes
    mib MiB each): the reference launch for scaling HBM byte counters."""
    import torch
    n = (mib << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(f"# copy calibration: elementwise mul of {mib} MiB float32 (reads {mib} MiB, writes {mib} MiB)",
          file=sys.stderr, flush=True)
    del a, b


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline_openmp(steps=5):
    """The reference's H+S assemblers with the OpenMP loop of
    M_assemble_3c_DOGS.f90 on every host thread (oracle/_ref/fb_harness_nbomp:
    the 2c assemblers have no OpenMP and run serially), on the Si4096 workload
    itself; its neighbour lists come from the drop-in driver_neighbors (GPU),
    outside the timed assemblers.  Median of `steps` repetitions."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import refrun
    binary = os.path.join(os.path.dirname(refrun.HARNESS_REF), "fb_harness_nbomp")
    if not refrun.available(binary):
        return {"error": "make -C oracle nbhip not built"}
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    species, structs = synth.config(1)
    st = structs[0]
    r = refrun.run(synth.make_tables(species), st, binary=binary, nrep=steps, threads=threads)
    reps = np.array(r["rep_times"])
    build = np.median(reps.sum(axis=1))
    t3 = np.median(reps[:, 4])
    return {"value": r["ntriplets"] / build, "unit": "triplets/s", "cores": threads, "kind": "reference",
            "cpu_model": _cpu_model(),
            "sample": f"Si4096 ({r['ntriplets']} triplets): H+S build {build * 1e3:.1f} ms, median of {steps} "
                      f"(assemble_S/T/dipole_z/vna_2c serial, vna_3c OpenMP, flang -O2 -fopenmp, "
                      f"OMP_PROC_BIND=close); vna_3c alone {t3 * 1e3:.1f} ms",
            "h_s_build_ms": build * 1e3, "vna_3c_ms": t3 * 1e3,
            "rep_ms": [[round(x * 1e3, 3) for x in rep] for rep in r["rep_times"]]}


def cpu_baseline():
    """The reference's own Fortran (serial flang build, oracle/_ref) on a
    bounded sample: a 512-atom Si cell (4x4x4 diamond, sigma 0.08 A)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import refrun
    st = synth.diamond_si((4, 4, 4), sigma=0.08, seed=0)
    ts = synth.make_tables([synth.SI])
    if refrun.available():
        nrep = 3
        r = refrun.run(ts, st, nrep=nrep)
        t = sum(r["times"][k] for k in ("S", "T", "DIP", "VNA2C", "VNA3C")) / nrep
        return {"value": r["ntriplets"] / t, "unit": "triplets/s", "cores": 1, "kind": "reference",
                "cpu_model": _cpu_model(),
                "sample": f"Si512 4x4x4 diamond (sigma 0.08 A): {r['ntriplets']} triplets, H+S build "
                          f"{t * 1e3:.1f} ms (assemble_S/T/dipole_z/vna_2c/vna_3c, serial flang -O2, "
                          f"mean of {nrep}); neighbour lists {r['times']['lists']:.2f} s not included",
                "h_s_build_ms": t * 1e3}
    import oracle
    o = oracle.System(oracle.Tables(ts), st)
    o.build_lists()
    t0 = time.perf_counter()
    o.assemble()
    t = time.perf_counter() - t0
    return {"value": o.lists["ntriplets"] / t, "unit": "triplets/s", "cores": 1, "kind": "port",
            "sample": f"Si512 4x4x4 diamond: C oracle H+S build {t * 1e3:.1f} ms", "h_s_build_ms": t * 1e3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--precision", default="mixed", choices=("mixed", "fp64"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-compare", action="store_true")
    ap.add_argument("--no-eig", action="store_true", help="skip the Gamma-point eigenvalue shift")
    ap.add_argument("--no-mlp", action="store_true", help="skip the learned 3c surrogate leg")
    ap.add_argument("--no-forces", action="store_true", help="skip the force-assembler leg")
    ap.add_argument("--no-rho", action="store_true", help="skip the McWEDA density leg")
    ap.add_argument("--no-pp", action="store_true", help="skip the pseudopotential leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-structure batch and CHNO3000 legs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one elementwise copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
    species, structs = synth.config(args.config)
    ts = synth.make_tables(species)
    asm = Assembler(ts, precision=args.precision)
    for st in structs:
        asm.add_structure(st)
    t0 = time.perf_counter()
    npairs, ntrip = asm.build_lists()
    lists_ms = (time.perf_counter() - t0) * 1e3
    # a rebuild on the same structures (what an MD step pays: buffers already allocated)
    t0 = time.perf_counter()
    asm.build_lists()
    asm.synchronize()
    lists_rebuild_ms = (time.perf_counter() - t0) * 1e3
    print(f"# lists: {npairs} pairs, {ntrip} triplets in {lists_ms:.1f} ms (rebuild {lists_rebuild_ms:.1f} ms)",
          file=sys.stderr, flush=True)

    chunk = None
    if args.precision == "mixed":
        asm.assemble_all()
        chunk = asm.autotune(20)
        print(f"# autotuned 3c chunk: {chunk} triplets per workgroup; 2c kernel "
              f"{('LDS-staged', 'register')[asm.tuning(1)]}; {launch_name(asm)}", file=sys.stderr, flush=True)
    # per-kernel times from separate 2c and 3c launches (the one-launch form
    # has no 3c kernel of its own to put on the roofline); the build time is
    # that of the tuned launch shape
    shape = asm.tuning(4)
    asm.set_tuning(4, 0)
    ms_events, kern = time_builds(asm, args.steps, args.warmup)
    asm.set_tuning(4, shape)
    ms = time_graph(asm, args.steps, args.warmup)
    print(f"# {args.precision}: {ms:.3f} ms per H+S build (graph replay; {ms_events:.3f} ms with per-kernel "
          f"events); kernels {kern}", file=sys.stderr, flush=True)
    res_main = {m: np.concatenate([asm.blocks(m, k) for k in range(len(structs))]) for m in MATS}

    extra = {}
    if not args.no_fp64_compare and args.precision != "fp64":
        asm.set_precision("fp64")
        ms64_events, kern64 = time_builds(asm, args.steps, args.warmup)
        ms64 = time_graph(asm, args.steps, args.warmup)
        res64 = {m: np.concatenate([asm.blocks(m, k) for k in range(len(structs))]) for m in MATS}
        err = {m: float(np.max(np.abs(res_main[m] - res64[m]))) for m in MATS}
        frob = {m: float(np.linalg.norm(res_main[m] - res64[m]) / max(np.linalg.norm(res64[m]), 1e-300))
                for m in MATS}
        t0 = time.perf_counter()
        eig = gamma_eigen_shift(asm, args.precision, "fp64") if not args.no_eig else None
        print(f"# eigenvalue shift {eig} ({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        extra.update(rel_frobenius_vs_fp64=frob, gamma_eigen_shift_eV=eig)
        extra.update(fp64_ms_per_step=ms64, fp64_ms_with_kernel_events=ms64_events, speedup_vs_fp64=ms64 / ms,
                     kernel_speedup_3c_vs_fp64=(kern64.get("assemble_3c_fp64", float("nan")) /
                                                max(kern.get("assemble_3c_mixed", float("nan")), 1e-12)),
                     max_abs_err_vs_fp64_eV=err,
                     fp64_kernels_ms=kern64)
        asm.set_precision(args.precision)
        print(f"# fp64: {ms64:.3f} ms; err {err}", file=sys.stderr, flush=True)

    mlp = None
    if not args.no_mlp and args.gpus == 1:
        try:
            mlp = mlp_leg(args.steps, args.warmup)
            print(f"# mlp (aC8192): {mlp['ms_per_step']:.3f} ms per build (table {mlp['table_ms_per_step']:.3f}), "
                  f"3c kernel {mlp['kernel_ms']:.3f} ms (table {mlp['table_kernel_ms']:.3f}, "
                  f"{mlp['speedup_vs_table_kernel']:.2f}x), fraction {mlp['fraction_on_network']:.3f}, "
                  f"err {mlp['max_abs_err_vs_fp64_eV']:.2e} eV, MFMA {mlp['mfma']['utilisation']:.3f}",
                  file=sys.stderr, flush=True)
        except Exception as e:  # noqa: BLE001 -- reported, not fatal to the main metric
            mlp = {"error": str(e)}

    chno = None
    if not args.no_batch and args.config != 2:
        try:
            chno = config_leg(args.steps, args.warmup)
            print(f"# CHNO3000: {chno['ms_per_step']:.3f} ms per build ({chno['launch']}, chunk {chno['chunk3c']}), "
                  f"binary64 {chno['fp64_ms_per_step']:.3f} ms, VNA err {chno['max_abs_err_vs_fp64_eV']['VNA']:.2e} eV",
                  file=sys.stderr, flush=True)
        except Exception as e:  # noqa: BLE001 -- reported, not fatal to the main metric
            chno = {"error": str(e)}

    batch = None
    if not args.no_batch and args.config != 3:
        try:
            batch = batch_leg(args.steps, args.warmup)
            print(f"# batch (512 structures): {batch['ms_per_step']:.3f} ms per build, "
                  f"{batch['structures_per_s']:.0f} structures/s ({batch['launch']}, chunk {batch['chunk3c']})",
                  file=sys.stderr, flush=True)
        except Exception as e:  # noqa: BLE001 -- reported, not fatal to the main metric
            batch = {"error": str(e)}

    forces = None
    if not args.no_forces:
        try:
            forces = forces_leg(asm, structs, args.steps, args.warmup)
            if not args.no_cpu_baseline:
                forces["cpu_baseline"] = force_cpu_baseline()
            print(f"# forces: {forces['ms_per_step']:.3f} ms per force build; kernels {forces['kernels_ms']}",
                  file=sys.stderr, flush=True)
        except Exception as e:  # reported, never required
            forces = {"error": str(e)}

    rho = None
    if not args.no_rho:
        try:
            rho = rho_leg(species, structs, args.steps, args.warmup)
            if not args.no_cpu_baseline:
                rho["cpu_baseline"] = rho_cpu_baseline()
            print(f"# densities: {rho['ms_per_step']:.3f} ms per McWEDA density build; kernels {rho['kernels_ms']}",
                  file=sys.stderr, flush=True)
        except Exception as e:  # reported, never required
            rho = {"error": str(e)}

    pp = None
    if not args.no_pp:
        try:
            pp = pp_leg(species, structs, args.steps, args.warmup)
            if not args.no_cpu_baseline:
                pp["cpu_baseline"] = pp_cpu_baseline()
            print(f"# pseudopotential: {pp['ms_per_step']:.3f} ms per PP build ({pp['pp_triplets']} PP triplets, "
                  f"lists {pp['lists_pp_ms']:.1f} ms); kernels {pp['kernels_ms']}", file=sys.stderr, flush=True)
        except Exception as e:  # reported, never required
            pp = {"error": str(e)}

    # dominant kernel: the 3c launch
    k3 = [k for k in kern if k.startswith("assemble_3c")]
    dom = max(kern, key=kern.get) if not k3 else k3[0]
    bytes3c, _ = algorithmic_bytes_3c(asm, len(structs))
    achieved = bytes3c / (kern[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, args.config, chunk), "kernel": dom,
            "kernel_ms": kern[dom], "algorithmic_bytes": bytes3c}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline()
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": str(e)}
        try:
            cpu["openmp"] = cpu_baseline_openmp()
            print(f"# cpu baseline (OpenMP, {cpu['openmp'].get('cores')} threads): {cpu['openmp'].get('sample')}",
                  file=sys.stderr, flush=True)
        except Exception as e:  # noqa: BLE001
            cpu["openmp"] = {"error": str(e)}

    out = {
        "metric": METRIC, "value": ntrip / (ms * 1e-3), "unit": "triplets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.config], "natoms": int(sum(s.natoms for s in structs)),
                   "nstructures": len(structs), "npairs": npairs, "ntriplets": ntrip,
                   "precision": args.precision, "h_s_build_ms": ms, "lists_ms": lists_ms,
                   "lists_rebuild_ms": lists_rebuild_ms,
                   "chunk3c": chunk, "launch": launch_name(asm) if chunk else None,
                   "ms_per_step_with_kernel_events": ms_events,
                   "kernels_ms": kern, **extra, "mlp_3c": mlp, "chno3000": chno, "batch": batch, "forces": forces, "rho_mcweda": rho, "pseudopotential": pp,
                   **({"structures_per_s": len(structs) / (ms * 1e-3)} if len(structs) > 1 else {})},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
