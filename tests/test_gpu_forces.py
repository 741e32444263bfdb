"""GPU force assemblers (fb_dassemble_*) through the C ABI against the
reference's own Fortran (golden fixtures) and the C oracle.

Dblocks and the 2c force pieces are per-pair values computed in the
reference's order: they agree to rounding (1e-12 of the largest value).  The
3c forces are per-atom sums over hundreds of triplets whose partial sums the
GPU adds in another order: 1e-11 of the largest per-atom value."""
import numpy as np
import pytest

import golden_cases
import oracle
from fireball_hip import synth

pytestmark = pytest.mark.gpu

DKEYS = {"DS": "S", "DT": "T", "DDIP": "DIP"}
FKEYS = ("vna_atom", "vna_ontop", "f3naa", "f3nab", "f3nac")


def _gpu_forces(ts, structs, rhos, dassemble_all=True):
    from fireball_hip.lib import Assembler
    asm = Assembler(ts, precision="fp64")
    for st in structs:
        asm.add_structure(st)
    asm.build_lists()
    asm.assemble_all()
    for k, rho in enumerate(rhos):
        asm.set_denmat(rho, k)
    if dassemble_all:
        asm.dassemble_all()
    else:
        asm.dassemble()
    return asm


def _rho(ts, st, lists, seed=11):
    return synth.denmat(ts.species, st.imass, lists["neighn"], lists["neigh_j"], lists["neigh_back"], seed=seed)


def _close(got, ref, rel, what):
    scale = max(1.0, float(np.max(np.abs(ref)))) if ref.size else 1.0
    np.testing.assert_allclose(got, ref, rtol=0, atol=rel * scale, err_msg=what)


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_forces_golden(golden, name):
    print(f"case {name}", flush=True)
    g, gf = golden(name), golden(name + "_forces")
    ts, st = golden_cases.inputs(g)
    lists = {k: g[k] for k in ("neighn", "neigh_j", "neigh_back")}
    asm = _gpu_forces(ts, [st], [_rho(ts, st, lists, int(gf["denmat_seed"]))])
    for k, m in DKEYS.items():
        d = asm.dblocks(m)
        n = len(gf[k])
        _close(d[:n], gf[k], 1e-12, k)
        assert abs(np.sum(d) - gf[k + "_sum"]) <= 1e-11 * max(1.0, gf[k + "_sumabs"]), k
    f = asm.forces()
    for k in ("vna_atom", "vna_ontop"):
        _close(f[k], gf[k], 1e-12, k)
    for k in ("f3naa", "f3nab", "f3nac"):
        _close(f[k], gf[k], 1e-11, k)


def test_forces_batch_matches_oracle_per_structure(golden):
    # si64 neutral and charged in one batch: both structures from one launch each
    g = golden("si64")
    ts, st0 = golden_cases.inputs(g)
    _, st1 = golden_cases.inputs(golden("si64_dq"))
    st2 = synth.diamond_si((1, 1, 1), sigma=0.05, seed=5)
    structs = [st0, st1, st2]
    refs, rhos = [], []
    for k, st in enumerate(structs):
        o = oracle.System(oracle.Tables(ts), st)
        lists = o.run_all()
        rho = _rho(ts, st, lists, seed=20 + k)
        rhos.append(rho)
        refs.append(o.run_forces(rho))
    asm = _gpu_forces(ts, structs, rhos)
    for k, ref in enumerate(refs):
        print(f"structure {k}", flush=True)
        for key, m in DKEYS.items():
            _close(asm.dblocks(m, k), ref[key], 1e-12, key)
        f = asm.forces(k)
        for key in ("vna_atom", "vna_ontop"):
            _close(f[key], ref[key], 1e-12, key)
        for key in ("f3naa", "f3nab", "f3nac"):
            _close(f[key], ref[key], 1e-11, key)


def test_stepwise_entry_points_equal_dassemble_all(golden):
    g = golden("cho40_dq")
    ts, st = golden_cases.inputs(g)
    rho = _rho(ts, st, {k: g[k] for k in ("neighn", "neigh_j", "neigh_back")})
    a = _gpu_forces(ts, [st], [rho], dassemble_all=True)
    b = _gpu_forces(ts, [st], [rho], dassemble_all=False)
    for m in DKEYS.values():
        np.testing.assert_array_equal(a.dblocks(m), b.dblocks(m))
    fa, fb = a.forces(), b.forces()
    for k in FKEYS:
        np.testing.assert_array_equal(fa[k], fb[k])


def test_order_errors():
    from fireball_hip.lib import Assembler, FireballError
    ts = synth.make_tables([synth.SI], seed=2024)
    st = synth.diamond_si((1, 1, 1), sigma=0.05, seed=5)
    asm = Assembler(ts, precision="fp64")
    asm.add_structure(st)
    asm.build_lists()
    with pytest.raises(FireballError):
        asm.dassemble(("VNA2C",))          # no density matrix yet
    asm.assemble_all()
    asm.set_denmat(np.zeros(len(asm.blocks("S"))))
    with pytest.raises(FireballError):
        asm.dassemble(("VNA3C",))          # Dassemble_S / _dipole_z first
    with pytest.raises(FireballError):
        asm.set_denmat(np.zeros(3))        # wrong size


@pytest.mark.parametrize("cfg", [1, 2])
def test_full_size_force_invariants(cfg):
    """Si4096 and CHNO3000 at full size: Dblock(S) and Dblock(T) of the reversed
    pair are minus the transpose (translation invariance of the 2c blocks),
    the per-triplet 3c forces sum to zero over all atoms (Newton's third law of
    f3naXc = -f3naXa - f3naXb), and a second run repeats every value bit for bit."""
    species, structs = synth.config(cfg)
    ts = synth.make_tables(species)
    st = structs[0]
    print(f"config {cfg}", flush=True)
    from fireball_hip.lib import Assembler
    asm = Assembler(ts, precision="fp64")
    asm.add_structure(st)
    asm.build_lists()
    asm.assemble_all()
    L = asm.lists()
    rho = _rho(ts, st, L)
    asm.set_denmat(rho)
    asm.dassemble_all()
    f1 = asm.forces()
    d1 = {m: asm.dblocks(m) for m in ("S", "T")}
    asm.dassemble_all()
    f2 = asm.forces()
    for k in FKEYS:
        np.testing.assert_array_equal(f1[k], f2[k])
    norb = np.array([sp.norb for sp in ts.species])
    pa = np.repeat(np.arange(st.natoms), L["neighn"])
    nmu = norb[st.imass[pa] - 1]
    nnu = norb[st.imass[L["neigh_j"] - 1] - 1]
    boff = np.concatenate([[0], np.cumsum(nmu * nnu)])
    noff = np.concatenate([[0], np.cumsum(L["neighn"])])
    back = noff[L["neigh_j"] - 1] + L["neigh_back"] - 1
    rng = np.random.default_rng(0)
    for p in rng.choice(len(back), size=2000, replace=False):
        q = back[p]
        for m in ("S", "T"):
            Dp = d1[m][3 * boff[p]:3 * boff[p + 1]].reshape(nnu[p], nmu[p], 3)
            Dq = d1[m][3 * boff[q]:3 * boff[q + 1]].reshape(nnu[q], nmu[q], 3)
            np.testing.assert_allclose(Dq, -np.transpose(Dp, (1, 0, 2)), rtol=0, atol=1e-12)
    tot = f1["f3naa"] + f1["f3nab"] + f1["f3nac"]
    scale = np.abs(f1["f3naa"]).sum() + np.abs(f1["f3nab"]).sum() + np.abs(f1["f3nac"]).sum()
    assert np.max(np.abs(tot.sum(axis=0))) <= 1e-12 * scale


@pytest.mark.parametrize("name", ["cho40_dq", "si64_dq", "chain_cluster"])
def test_sp_kernels_equal_generic_kernels(golden, name):
    """The s/p force kernels (compile-time block shapes) against the generic
    shell-loop kernels: the same operations in the same order."""
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    rho = _rho(ts, st, {k: g[k] for k in ("neighn", "neigh_j", "neigh_back")})
    a = _gpu_forces(ts, [st], [rho])
    from fireball_hip.lib import Assembler
    b = Assembler(ts, precision="fp64")
    b.add_structure(st)
    b.build_lists()
    b.assemble_all()
    b.set_tuning(3, 0)
    b.set_denmat(rho)
    b.dassemble_all()
    for m in DKEYS.values():
        np.testing.assert_array_equal(a.dblocks(m), b.dblocks(m))
    fa, fb = a.forces(), b.forces()
    for k in ("vna_atom", "vna_ontop"):
        np.testing.assert_array_equal(fa[k], fb[k])
    for k in ("f3naa", "f3nab", "f3nac"):
        np.testing.assert_allclose(fa[k], fb[k], rtol=0, atol=1e-13 * max(1.0, np.max(np.abs(fb[k]))))


TOL_FORCE_MIXED = 1e-6   # eV/A: mixed-precision 3c forces against binary64, per component


@pytest.mark.parametrize("cfg", [0, 1, 2])
def test_mixed_forces_within_tolerance(cfg):
    """Mixed precision (FP32 interpolation of the 3c force kernel, binary64 frames,
    frame derivatives, rotation and sums): Dblocks and pair forces are binary64 in
    both modes and equal; the per-atom 3c forces are within TOL_FORCE_MIXED of
    binary64."""
    from fireball_hip.lib import Assembler
    species, structs = synth.config(cfg)
    ts = synth.make_tables(species)
    st = structs[0]
    print(f"config {cfg}", flush=True)
    asm = Assembler(ts, precision="fp64")
    asm.add_structure(st)
    asm.build_lists()
    asm.assemble_all()
    L = asm.lists()
    asm.set_denmat(_rho(ts, st, L))
    asm.dassemble_all()
    ref = asm.forces()
    dref = {m: asm.dblocks(m) for m in DKEYS.values()}
    asm.set_precision("mixed")
    asm.assemble_all()
    asm.dassemble_all()
    got = asm.forces()
    for m in DKEYS.values():
        np.testing.assert_array_equal(asm.dblocks(m), dref[m])
    for k in ("vna_atom", "vna_ontop"):
        np.testing.assert_array_equal(got[k], ref[k])
    for k in ("f3naa", "f3nab", "f3nac"):
        err = float(np.max(np.abs(got[k] - ref[k])))
        print(f"config {cfg} {k} mixed vs fp64 max-abs {err:.3e} (max |F| {np.max(np.abs(ref[k])):.3e})", flush=True)
        assert err <= TOL_FORCE_MIXED, k
