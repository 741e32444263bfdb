"""GPU pseudopotential path (fb_build_lists_pp, fb_assemble_svnl / _vnl_2c /
_vnl_3c) through the C ABI against the reference's own modules (golden
fixtures) and the C oracle: the lists bit for bit, the sVNL and vnl blocks
value for value (the kernels keep the reference's operations and order).  At
the full benchmark size the dense Kleinman-Bylander identity of the oracle
tests stands in for the oracle, whose O(N^2) list build takes minutes."""
import numpy as np
import pytest

import golden_cases
import oracle
from fireball_hip import synth
from golden.make_golden import PP_LIST_KEYS, pp_tables

pytestmark = pytest.mark.gpu


def _gpu_pp(ts, structs, which=("SVNL", "VNL2C", "VNL3C")):
    from fireball_hip.lib import Assembler
    asm = Assembler(ts, precision="fp64")
    for st in structs:
        asm.add_structure(st)
    asm.build_lists()
    asm.build_lists_pp()
    asm.assemble_pp(which)
    return asm


def _same(got, ref, what):
    assert got.shape == ref.shape, f"{what}: shape {got.shape} vs {ref.shape}"
    assert not np.any(np.isnan(got)), what
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{what}: {bad.size} values differ, first {bad[:5]}"


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_pp_golden(golden, name):
    print(f"case {name}", flush=True)
    g, gp = golden(name), golden(name + "_pp")
    ts, st = golden_cases.inputs(g)
    tsp = pp_tables(ts, int(gp["pp_table_seed"]))
    asm = _gpu_pp(tsp, [st])
    L = asm.pp_lists()
    for k in PP_LIST_KEYS:
        _same(L[k], gp[k], k)
    for k in ("SVNL", "VNL"):
        b = asm.pp_blocks(k)
        _same(b[:gp[k].size], gp[k], k)
        assert np.sum(b) == gp[k + "_sum"], k


def test_pp_batch_matches_oracle_per_structure():
    sp = synth.with_pp(synth.batch_species())
    ts = synth.add_pp_tables(synth.make_tables(sp, seed=31), seed=33)
    structs = [s for s in synth.batch_structures(nstruct=12, seed=4) if s.natoms <= 160][:4]
    tab = oracle.Tables(ts)
    refs = []
    for st in structs:
        print(f"oracle {st.name} natoms={st.natoms}", flush=True)
        o = oracle.System(tab, st)
        o.build_lists()
        refs.append(o.run_pp())
    asm = _gpu_pp(ts, structs)
    for k, ref in enumerate(refs):
        L = asm.pp_lists(k)
        for key in PP_LIST_KEYS:
            _same(L[key], ref[key], f"structure {k} {key}")
        for key in ("SVNL", "VNL"):
            _same(asm.pp_blocks(key, k), ref[key], f"structure {k} {key}")


def _triclinic(n=32, seed=11):
    """Si atoms at random in a skewed periodic cell, at least 2.0 A apart
    across the cell's images."""
    rng = np.random.default_rng(seed)
    lat = np.array([[8.2, 0.0, 0.0], [2.6, 7.6, 0.0], [1.4, 2.3, 7.1]])
    shifts = np.array([i * lat[0] + j * lat[1] + k * lat[2] for i in (-1, 0, 1) for j in (-1, 0, 1)
                       for k in (-1, 0, 1)])
    pos = np.zeros((0, 3))
    while len(pos) < n:
        p = rng.uniform(0.0, 1.0, 3) @ lat
        if len(pos) == 0 or np.min(np.linalg.norm(p - pos[:, None, :] - shifts[None, :, :], axis=2)) >= 2.0:
            pos = np.vstack([pos, p])
    return synth.Structure(pos, np.ones(n, np.int32), lat, 0, name="tri32")


def test_lists_triclinic_cell_against_oracle():
    """A skewed periodic cell (the cell-list neighbour search over images whose
    boxes are not axis-aligned with the lattice): the H lists, the PP, PPx and
    PPp lists with their maps and triplets, and the sVNL / vnl blocks against
    the oracle, bit for bit."""
    from golden.make_golden import LIST_KEYS
    sp = synth.with_pp([synth.SI])
    ts = synth.add_pp_tables(synth.make_tables(sp, seed=41), seed=43)
    st = _triclinic()
    o = oracle.System(oracle.Tables(ts), st)
    ref = o.run_all()
    ref_pp = o.run_pp()
    asm = _gpu_pp(ts, [st])
    got = asm.lists(0)
    for k in LIST_KEYS:
        _same(got[k], ref[k], k)
    L = asm.pp_lists()
    for k in PP_LIST_KEYS:
        _same(L[k], ref_pp[k], k)
    for k in ("SVNL", "VNL"):
        _same(asm.pp_blocks(k), ref_pp[k], k)


@pytest.mark.parametrize("name", ["si64", "cho40_dq", "chain_cluster"])
def test_pp_sp_kernels_equal_generic_kernels(golden, name):
    """The s/p PP kernels (compile-time shapes, register rotations, unrolled
    contractions, batched atom-case loads, one-pass 3c tiles) against the
    generic shell-loop kernels."""
    from fireball_hip.lib import Assembler
    g, gp = golden(name), golden(name + "_pp")
    ts, st = golden_cases.inputs(g)
    tsp = pp_tables(ts, int(gp["pp_table_seed"]))
    a = _gpu_pp(tsp, [st])
    b = Assembler(tsp, precision="fp64")
    b.add_structure(st)
    b.build_lists()
    b.set_tuning(3, 0)
    b.build_lists_pp()
    b.assemble_pp()
    for k in ("SVNL", "VNL"):
        _same(a.pp_blocks(k), b.pp_blocks(k), k)


def test_pp_order_errors(golden):
    from fireball_hip.lib import Assembler, FireballError
    g, gp = golden("si8_cell"), golden("si8_cell_pp")
    ts, st = golden_cases.inputs(g)
    tsp = pp_tables(ts, int(gp["pp_table_seed"]))
    asm = Assembler(tsp, precision="fp64")
    asm.add_structure(st)
    asm.build_lists()
    with pytest.raises(FireballError) as e:
        asm.assemble_pp(("SVNL",))                 # lists missing
    assert e.value.code == -8
    asm.build_lists_pp()
    for step in ("VNL2C", "VNL3C"):
        with pytest.raises(FireballError) as e:
            asm.assemble_pp((step,))
        assert e.value.code == -8
    asm.assemble_pp()
    _same(asm.pp_blocks("VNL")[:gp["VNL"].size], gp["VNL"], "VNL after the error paths")


@pytest.mark.parametrize("cfg", [1, 2])
def test_pp_full_size_kb_operator(cfg):
    """Si4096 / CHNO3000: the assembled vnl blocks against the dense sum over PP
    atoms of P_alpha diag(cl) P_alpha^T formed from the GPU's own sVNL blocks,
    on a sample of atoms (Gamma-folded rows), and repeat runs bit-identical."""
    sp, structs = synth.config(cfg)
    sp = synth.with_pp(sp)
    st = structs[0]
    ts = synth.add_pp_tables(synth.make_tables(sp), seed=6066)
    print(f"config {cfg}: natoms={st.natoms}", flush=True)
    asm = _gpu_pp(ts, [st])
    L = asm.pp_lists()
    sv, vl = asm.pp_blocks("SVNL"), asm.pp_blocks("VNL")
    asm.assemble_pp()
    _same(asm.pp_blocks("VNL"), vl, "repeat")
    norb = np.array([s.norb for s in sp])
    norb_pp = np.array([s.norb_pp for s in sp])
    N = st.natoms
    opp = np.concatenate([[0], np.cumsum(L["pp_n"])])
    opq = np.concatenate([[0], np.cumsum(L["ppp_n"])])
    sv_sz = norb[st.imass[np.repeat(np.arange(N), L["pp_n"])] - 1] * norb_pp[st.imass[L["pp_j"] - 1] - 1]
    sv_off = np.concatenate([[0], np.cumsum(sv_sz)])
    vl_sz = norb[st.imass[np.repeat(np.arange(N), L["ppp_n"])] - 1] * norb[st.imass[L["ppp_j"] - 1] - 1]
    vl_off = np.concatenate([[0], np.cumsum(vl_sz)])
    # P_alpha rows of atom i: sum over the images of the PP pairs (i, alpha)
    P = {}
    for i in range(N):
        for p in range(opp[i], opp[i + 1]):
            a = int(L["pp_j"][p]) - 1
            nm, nc = norb[st.imass[i] - 1], norb_pp[st.imass[a] - 1]
            blk = sv[sv_off[p]:sv_off[p + 1]].reshape(nc, nm).T
            P.setdefault(a, {}).setdefault(i, np.zeros((nm, nc)))
            P[a][i] = P[a][i] + blk
    rng = np.random.default_rng(cfg)
    sample = rng.choice(N, size=24, replace=False)
    scale = np.abs(vl).max()
    for i in sample:
        row = {}
        for p in range(opq[i], opq[i + 1]):
            j = int(L["ppp_j"][p]) - 1
            nm, nn = norb[st.imass[i] - 1], norb[st.imass[j] - 1]
            row[j] = row.get(j, np.zeros((nm, nn))) + vl[vl_off[p]:vl_off[p + 1]].reshape(nn, nm).T
        ref = {}
        for a, rows in P.items():
            if i not in rows:
                continue
            s_ = sp[st.imass[a] - 1]
            cl = np.repeat(s_.cl_pp, [2 * l + 1 for l in s_.lssh_pp])
            for j, Pj in rows.items():
                ref[j] = ref.get(j, 0.0) + rows[i] @ np.diag(cl) @ Pj.T
        assert set(ref) == set(row), f"atom {i}: block pattern"
        for j in ref:
            np.testing.assert_allclose(row[j], ref[j], atol=1e-12 * scale, err_msg=f"atom {i} block {j}")


def test_pp_atoms_and_structures_without_triplets():
    """PP triplets are tiled per atom like the 3c ones.  A lone dimer (first
    in the batch, so every per-structure offset starts from it), and a cluster
    with an isolated atom and a distant dimer, give atoms that own no PP
    triplet; lists and blocks must equal the oracle's bit for bit."""
    sp = synth.with_pp([synth.SI])
    ts = synth.add_pp_tables(synth.make_tables(sp, seed=5), seed=6)
    box = np.eye(3) * 60.0
    dimer = np.array([[10.0, 10.0, 10.0], [12.3, 10.0, 10.0]])
    si8 = synth.diamond_si((1, 1, 1), sigma=0.05, seed=3).ratom + 10.0
    pos1 = np.vstack([si8, [[45.0, 45.0, 45.0]], dimer + 25.0])
    structs = [synth.Structure(dimer.copy(), np.ones(2, np.int32), box, 1),
               synth.Structure(pos1, np.ones(len(pos1), np.int32), box, 1),
               synth.diamond_si((1, 1, 1), sigma=0.06, seed=5)]
    tab = oracle.Tables(ts)
    refs = []
    for st in structs:
        o = oracle.System(tab, st)
        o.build_lists()
        refs.append(o.run_pp())
    asm = _gpu_pp(ts, structs)
    for k, ref in enumerate(refs):
        print(f"structure {k} natoms={structs[k].natoms}", flush=True)
        L = asm.pp_lists(k)
        for key in PP_LIST_KEYS:
            _same(L[key], ref[key], f"structure {k} {key}")
        for key in ("SVNL", "VNL"):
            _same(asm.pp_blocks(key, k), ref[key], f"structure {k} {key}")
