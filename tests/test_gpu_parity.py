"""GPU parity: the HIP assembler through its C ABI against the golden vectors
of the reference Fortran and against the oracle.

Lists and binary64 blocks must be bit-identical: the binary64 kernels keep
the reference's operations and their order (no contraction), so any
difference is a defect.  The mixed-precision path is held to 1e-6 eV."""
import numpy as np
import pytest

import golden_cases
import oracle

pytestmark = pytest.mark.gpu

LIST_KEYS = ("neighn", "neigh_j", "neigh_b", "neigh_back", "neigh_self", "ncommon", "com_iatom_j",
             "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh")
MATS = ("S", "T", "DIP", "VNA", "VNAO")
def assert_bits(a, b, what):
    """Bit-identical binary64 arrays (NaN nowhere); the message gives the
    largest difference when they are not."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        i = int(np.argmax(d))
        raise AssertionError(f"{what}: {int(np.count_nonzero(a != b))} of {a.size} values differ, "
                             f"max |diff| {d[i]:.3e} at {i} ({a[i]!r} vs {b[i]!r})")


def _gpu(ts, structs, precision="fp64"):
    from fireball_hip.lib import Assembler
    asm = Assembler(ts, precision=precision)
    for st in structs:
        asm.add_structure(st)
    asm.build_lists()
    asm.assemble_all()
    return asm


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_golden_fp64(golden, name):
    print(f"case {name}", flush=True)
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    asm = _gpu(ts, [st])
    r = asm.run_all(0)
    for k in LIST_KEYS:
        assert np.array_equal(r[k], g[k]), k
    n = len(g["S"])
    o = oracle.System(oracle.Tables(ts), st).run_all()
    for m in MATS:
        assert_bits(r[m][:n], g[m], m + " vs golden")
        assert_bits(r[m], o[m], m + " vs oracle")


def test_separate_entry_points_match_fused(golden):
    g = golden("cho40_dq")
    ts, st = golden_cases.inputs(g)
    from fireball_hip.lib import Assembler
    asm = Assembler(ts)
    asm.add_structure(st)
    asm.build_lists()
    asm.assemble_S()
    asm.assemble_T()
    asm.assemble_dipole_z()
    asm.assemble_vna_2c()
    asm.assemble_vna_3c()
    sep = {m: asm.blocks(m) for m in MATS}
    asm.assemble_all()
    for m in MATS:
        assert_bits(asm.blocks(m), sep[m], m)


def test_batch_of_structures_matches_one_by_one():
    from fireball_hip import synth
    sp = [synth.SI]
    ts = synth.make_tables(sp)
    structs = [synth.diamond_si(c, sigma=0.07, seed=k) for k, c in enumerate([(1, 1, 1), (2, 1, 1), (2, 2, 1), (1, 1, 3)])]
    structs[1].icluster = 1
    asm = _gpu(ts, structs)
    tab = oracle.Tables(ts)
    for k, st in enumerate(structs):
        print(f"batch member {k}", flush=True)
        r = asm.run_all(k)
        o = oracle.System(tab, st).run_all()
        for key in LIST_KEYS:
            assert np.array_equal(r[key], o[key]), (k, key)
        for m in MATS:
            assert_bits(r[m], o[m], f"{k} {m}")


def test_si216_against_oracle():
    from fireball_hip import synth
    sp = [synth.SI]
    ts = synth.make_tables(sp)
    st = synth.diamond_si((3, 3, 3), sigma=0.08, seed=4)
    synth.random_dQ([st], sp, seed=9)
    asm = _gpu(ts, [st])
    r = asm.run_all(0)
    o = oracle.System(oracle.Tables(ts), st).run_all()
    for key in LIST_KEYS:
        assert np.array_equal(r[key], o[key]), key
    for m in MATS:
        assert_bits(r[m], o[m], m)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("charged", [False, True])
def test_autotune_leaves_complete_blocks(precision, charged):
    """fb_autotune times 2c and 3c launches that overwrite parts of the blocks;
    afterwards every matrix must again hold a whole H+S build (Vna with its
    three-centre terms), equal to the oracle, and a later build must give the
    same values."""
    from fireball_hip import synth
    print(f"autotune {precision} charged={charged}", flush=True)
    sp = [synth.SI]
    ts = synth.make_tables(sp)
    st = synth.diamond_si((2, 2, 2), sigma=0.06, seed=12)
    if charged:
        synth.random_dQ([st], sp, seed=13)
    asm = _gpu(ts, [st], precision)
    chunk = asm.autotune(nrep=2)
    assert chunk in (64, 128, 256)
    o = oracle.System(oracle.Tables(ts), st).run_all()
    tuned = {m: asm.blocks(m) for m in MATS}
    for m in MATS:
        if precision == "fp64":
            assert_bits(tuned[m], o[m], m)
        else:
            np.testing.assert_allclose(tuned[m], o[m], rtol=0, atol=1e-6, err_msg=m)
    asm.assemble_all()
    for m in MATS:
        assert_bits(asm.blocks(m), tuned[m], m + " rebuilt")


@pytest.mark.parametrize("idx", [1, 2])
def test_full_size_fp64_digest(idx):
    """Si4096 and CHNO3000 at their benchmark sizes: the binary64 GPU lists
    hash to the oracle's (bit-identical) and the blocks' value hashes equal
    the oracle's (every element equal; tests/golden/make_digests.py ran the C
    oracle on the same inputs)."""
    import json
    import os
    from fireball_hip import synth
    sys_path = os.path.join(os.path.dirname(__file__), "golden")
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_digests", os.path.join(sys_path, "make_digests.py"))
    md = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(md)
    with open(md.path_for(idx)) as f:
        want = json.load(f)
    species, structs = synth.config(idx)
    print(f"digest workload {idx}: {structs[0].natoms} atoms", flush=True)
    asm = _gpu(synth.make_tables(species), structs, "fp64")
    got = md.summarize(asm.run_all(0))
    for k, rec in want["lists"].items():
        assert got["lists"][k] == rec, k
    for m, rec in want["blocks"].items():
        g = got["blocks"][m]
        assert g["n"] == rec["n"], m
        if g["sha256_values"] != rec["sha256_values"]:
            d = np.max(np.abs(np.array(g["sample"]) - np.array(rec["sample"])))
            raise AssertionError(f"{m}: values differ (sum {g['sum']!r} vs {rec['sum']!r}, "
                                 f"sampled max |diff| {d:.3e})")
        # equal values; the bytes may differ only in the sign of zeros
        print(f"{m}: values identical, bytes {'identical' if g['sha256'] == rec['sha256'] else 'differ in zero signs'}")


@pytest.mark.parametrize("persist", [0, 1])
@pytest.mark.parametrize("waves", [3, 4])
def test_fused_build_matches_separate_launches(waves, persist):
    """One launch for the whole neutral mixed-precision H+S build (2c atom
    groups, then 3c tiles: four per workgroup, or taken one per wavefront
    from per-XCD queues by persistent workgroups) gives the same bits as the
    separate 2c and 3c launches, on a batch with a structure without
    triplets, and from a replayed graph as well as eagerly."""
    from fireball_hip import synth
    box = np.eye(3) * 60.0
    dimer = synth.Structure(np.array([[10.0, 10.0, 10.0], [12.3, 10.0, 10.0]]), np.ones(2, np.int32), box, 1)
    structs = [dimer, synth.diamond_si((3, 3, 3), sigma=0.08, seed=4), synth.diamond_si((1, 2, 1), sigma=0.06, seed=5)]
    ts = synth.make_tables([synth.SI])
    asm = _gpu(ts, structs, precision="mixed")
    ref = [{m: asm.blocks(m, k) for m in MATS} for k in range(len(structs))]
    asm.set_tuning(4, waves)
    asm.set_tuning(5, persist)
    for graph in (1, 0):
        asm.set_tuning(2, graph)
        print(f"fused waves={waves} persist={persist} graph={graph}", flush=True)
        for _ in range(2):
            asm.assemble_all()
            for k in range(len(structs)):
                for m in MATS:
                    assert_bits(asm.blocks(m, k), ref[k][m], f"{k} {m}")


def test_error_atoms_too_close():
    from fireball_hip import synth
    from fireball_hip.lib import Assembler, FireballError
    ts = synth.make_tables([synth.SI], with_3c=False)
    pos = np.array([[1.0, 1.0, 1.0], [1.3, 1.0, 1.0], [3.0, 3.0, 3.0]])
    st = synth.Structure(pos, np.ones(3, np.int32), np.eye(3) * 20, 1)
    asm = Assembler(ts)
    asm.add_structure(st)
    with pytest.raises(FireballError) as e:
        asm.build_lists()
    assert e.value.code == -4


def test_error_x_out_of_range():
    from fireball_hip import synth
    from fireball_hip.lib import Assembler, FireballError
    ts = synth.make_tables([synth.SI], with_3c=False)
    for c in ts.cells2c.values():
        c.xmax = 2.0                       # shorter than the neighbour range
    pos = np.array([[1.0, 1.0, 1.0], [3.5, 1.0, 1.0]])
    st = synth.Structure(pos, np.ones(2, np.int32), np.eye(3) * 20, 1)
    asm = Assembler(ts)
    asm.add_structure(st)
    asm.build_lists()
    with pytest.raises(FireballError) as e:
        asm.assemble_S()
    assert e.value.code == -3


def test_order_errors():
    from fireball_hip import synth
    from fireball_hip.lib import Assembler, FireballError
    ts = synth.make_tables([synth.SI])
    asm = Assembler(ts)
    asm.add_structure(synth.diamond_si((1, 1, 1), seed=1))
    with pytest.raises(FireballError) as e:
        asm.assemble_S()
    assert e.value.code == -8
    asm.build_lists()
    with pytest.raises(FireballError) as e:
        asm.assemble_vna_3c()
    assert e.value.code == -8


# ----------------------------------------------------------- mixed precision
TOL_MIXED = 1.0e-6      # eV, max-abs per block element (the bar of the FP32 path)


@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_golden_mixed(golden, name):
    print(f"mixed case {name}", flush=True)
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    asm = _gpu(ts, [st], precision="mixed")
    r = asm.run_all(0)
    for k in LIST_KEYS:
        assert np.array_equal(r[k], g[k]), k
    o = oracle.System(oracle.Tables(ts), st).run_all()
    for m in MATS:
        err = float(np.max(np.abs(r[m] - o[m])))
        assert err <= TOL_MIXED, (m, err)
    # the 2c blocks stay binary64 in the mixed path
    for m in ("S", "T", "DIP", "VNAO"):
        assert_bits(r[m], o[m], m)


def test_si4096_mixed_vs_fp64():
    """Config-1 scale: the mixed path against the binary64 GPU path (itself
    bit-identical to the oracle at the smaller sizes above)."""
    from fireball_hip import synth
    species, structs = synth.config(1)
    ts = synth.make_tables(species)
    asm = _gpu(ts, structs, precision="mixed")
    mixed = asm.blocks("VNA")
    asm.set_precision("fp64")
    asm.assemble_all()
    ref = asm.blocks("VNA")
    err = float(np.max(np.abs(mixed - ref)))
    print(f"Si4096 mixed vs fp64 VNA max-abs {err:.3e} eV", flush=True)
    assert err <= TOL_MIXED
    # size-independent invariants of the lists at full size
    L = asm.lists(0)
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    back = off[L["neigh_j"] - 1] + L["neigh_back"] - 1
    atom_of_pair = np.repeat(np.arange(1, structs[0].natoms + 1), L["neighn"])
    assert np.array_equal(L["neigh_j"][back], atom_of_pair)
    assert np.array_equal(back[back], np.arange(len(back)))


def test_mixed_segments_longer_than_chunk():
    """A dense Si cluster in which every pair is a neighbour pair and every
    other atom a common neighbour: each Vna block takes N-2 = 268 triplets,
    more than any workgroup chunk (64, 128, 256), so every workgroup sums its
    segment over several chunks.  The three chunk sizes split the same
    binary64 sums differently and must agree bit for bit; against binary64 the
    bound is the FP32 rounding of 268 contributions per element (ten times
    the per-element bar of the benchmark workloads, whose blocks take ~14),
    far below one lost or repeated contribution."""
    from fireball_hip import synth
    st = _dense_si_cluster()       # ball of 2.8 A: every distance below 2 rc = 5.7 A
    ts = synth.make_tables([synth.SI])
    asm = _gpu(ts, [st], precision="fp64")
    L = asm.lists(0)
    n = st.natoms
    assert np.all(L["neighn"] == n) and np.all(L["ncommon"] == (n - 1) * (n - 2))
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    seg = np.bincount(off[L["com_iatom_j"] - 1] + L["com_mneigh"] - 1)
    assert seg[seg > 0].min() == n - 2 > 256
    ref = asm.blocks("VNA")
    asm.set_precision("mixed")
    got = {}
    for chunk in (64, 128, 256):
        asm.set_chunk3c(chunk)
        asm.assemble_all()
        got[chunk] = asm.blocks("VNA")
        err = float(np.max(np.abs(got[chunk] - ref)))
        print(f"dense cluster chunk {chunk}: mixed vs fp64 VNA max-abs {err:.3e} eV "
              f"(max |VNA| {float(np.max(np.abs(ref))):.3e})", flush=True)
        assert err <= 10 * TOL_MIXED, chunk
    assert np.array_equal(got[64], got[128]) and np.array_equal(got[64], got[256])


def _dense_si_cluster(n=270, radius=2.8, seed=7):
    """n Si atoms in a ball of the given radius (closest approach 0.55 A):
    every pair is a neighbour pair, so every atom has n neighbours."""
    from fireball_hip import synth
    rng = np.random.default_rng(seed)
    pos = np.zeros((0, 3))
    while len(pos) < n:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) <= radius and (len(pos) == 0 or np.min(np.linalg.norm(pos - p, axis=1)) >= 0.55):
            pos = np.vstack([pos, p])
    return synth.Structure(pos + 12.0, np.ones(len(pos), np.int32), np.eye(3) * 40.0, 1)


def test_2c_passes_and_partial_wavefronts_same_bits():
    """Atoms with more neighbours (270) than a 2c workgroup pass (256 pairs):
    each atom's workgroup runs a full pass of four wavefronts and a pass whose
    only wavefront holds 14 pairs, so the wavefront-coalesced block stores see
    full and partial wavefronts and a second pass of the self-block sum.  The
    register kernel (coalesced stores), the LDS-staged kernel (workgroup
    stores) and the one-launch build must give the same bits on every
    matrix."""
    from fireball_hip import synth
    st = _dense_si_cluster()
    ts = synth.make_tables([synth.SI])
    asm = _gpu(ts, [st], precision="fp64")
    L = asm.lists(0)
    assert np.all(L["neighn"] == 270)
    got = {}
    for var in (0, 1):
        asm.set_tuning(1, var)
        asm.assemble_all()
        got[var] = {m: asm.blocks(m) for m in MATS}
    for m in MATS:
        assert np.array_equal(got[0][m], got[1][m]), m
    # mixed precision: two launches against the one-launch build at 3 and 4
    # waves per SIMD (the 2c matrices are binary64 in every shape)
    asm.set_precision("mixed")
    asm.set_tuning(1, 1)
    shapes = {}
    for shape in (0, 3, 4):
        asm.set_tuning(4, shape)
        asm.assemble_all()
        shapes[shape] = {m: asm.blocks(m) for m in MATS}
    for m in MATS:
        assert np.array_equal(shapes[0][m], shapes[3][m]), m
        assert np.array_equal(shapes[0][m], shapes[4][m]), m
        if m != "VNA":
            assert np.array_equal(shapes[0][m], got[1][m]), m


@pytest.mark.parametrize("name", ["cho40_dq", "si64"])
def test_2c_kernel_variants_same_bits(golden, name):
    """The two s/p 2c kernels (table cells staged in LDS, register-resident
    with per-lane gathers) give the same bits on every matrix, neutral and
    charged."""
    print(f"2c variants {name}", flush=True)
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    asm = _gpu(ts, [st])
    got = {}
    for var in (0, 1):
        asm.set_tuning(1, var)
        asm.assemble_all()
        got[var] = {m: asm.blocks(m) for m in MATS}
    for m in MATS:
        assert np.array_equal(got[0][m], got[1][m]), m
    from fireball_hip.lib import FireballError
    with pytest.raises(FireballError):
        asm.set_tuning(1, 2)


def test_generic_shells_against_oracle():
    """Species outside the s / sp fast path (two s shells) run the generic kernels."""
    from fireball_hip import synth
    sp = [synth.Species(3, [0, 0], [5.0, 4.6]), synth.C_SP]
    ts = synth.make_tables(sp, seed=17)
    st = synth.random_box({3: 6, 6: 10}, 0.8, 1.0, 21, sp)
    synth.random_dQ([st], sp, seed=3)
    asm = _gpu(ts, [st])
    r = asm.run_all(0)
    o = oracle.System(oracle.Tables(ts), st).run_all()
    for key in LIST_KEYS:
        assert np.array_equal(r[key], o[key]), key
    for m in MATS:
        assert_bits(r[m], o[m], m)


def test_tables_committed_after_lists():
    """Lists built under a table set that lacks the 3c groups with a carbon
    third atom, then the full set committed: the per-triplet group index
    written at list build must follow the new tables (their group numbering
    shifts), in binary64 and mixed precision, without rebuilding the lists."""
    import dataclasses
    from fireball_hip import synth
    from fireball_hip.lib import Assembler
    sp = [synth.SI, synth.C_SP]
    ts = synth.make_tables(sp, seed=5)
    part = dataclasses.replace(ts, cells3c={k: c for k, c in ts.cells3c.items() if k[2] != 2})
    assert len(part.cells3c) < len(ts.cells3c)
    st = synth.random_box({14: 8, 6: 8}, 0.8, 1.0, 23, sp)
    o = oracle.System(oracle.Tables(ts), st).run_all()
    for precision in ("fp64", "mixed"):
        asm = Assembler(part, precision=precision)
        asm.add_structure(st)
        asm.build_lists()
        asm._set_tables(ts)        # the missing groups are added and the table set committed again
        asm.assemble_all()
        for m in MATS:
            got = asm.blocks(m)
            if precision == "fp64":
                assert_bits(got, o[m], m)
            else:
                np.testing.assert_allclose(got, o[m], rtol=0, atol=TOL_MIXED, err_msg=m)


def test_graph_replays_match_eager():
    """assemble_all replays a captured HIP graph; every replay, with eager
    builds and host synchronisations in between, must equal the eager result."""
    from fireball_hip import synth
    from fireball_hip.lib import Assembler
    sp = [synth.SI]
    ts = synth.make_tables(sp)
    st = synth.diamond_si((2, 2, 2), sigma=0.06, seed=11)
    for precision in ("fp64", "mixed"):
        asm = Assembler(ts, precision=precision)
        asm.add_structure(st)
        asm.build_lists()
        asm.enable_timing(True)            # timing on: eager launches
        asm.assemble_all()
        asm.timings(64)
        eager = {m: asm.blocks(m) for m in MATS}
        asm.enable_timing(False)           # timing off: graph replays
        for rep in range(4):
            asm.assemble_all(sync=(rep % 2 == 0))
            asm.synchronize()
            for m in MATS:
                assert np.array_equal(asm.blocks(m), eager[m]), (precision, rep, m)
            asm.assemble_S()               # eager work between replays
            asm.synchronize()


def test_async_error_surfaces_at_synchronize():
    from fireball_hip import synth
    from fireball_hip.lib import Assembler, FireballError
    ts = synth.make_tables([synth.SI])
    for c in ts.cells2c.values():
        c.xmax = 2.0
    pos = np.array([[1.0, 1.0, 1.0], [3.5, 1.0, 1.0], [1.0, 3.4, 1.0]])
    st = synth.Structure(pos, np.ones(3, np.int32), np.eye(3) * 20, 1)
    asm = Assembler(ts)
    asm.add_structure(st)
    asm.build_lists()
    asm.assemble_all(sync=False)
    with pytest.raises(FireballError) as e:
        asm.synchronize()
    assert e.value.code == -3


def _fold_upper(blocks, neighn, neigh_j, imass, norb_sp):
    """Gamma fold of pair blocks into the upper triangle, symmetric completion."""
    no = norb_sp[np.asarray(imass) - 1]
    off = np.concatenate([[0], np.cumsum(no)])
    n = int(off[-1])
    A = np.zeros((n, n))
    ia = np.repeat(np.arange(len(neighn)), neighn)
    b = 0
    for p in range(len(ia)):
        i, j = ia[p], neigh_j[p] - 1
        nm, nn = no[i], no[j]
        A[off[i]:off[i] + nm, off[j]:off[j] + nn] += blocks[b:b + nm * nn].reshape(nn, nm).T
        b += nm * nn
    U = np.triu(A)
    return U + np.triu(A, 1).T


@pytest.mark.parametrize("name", ["si64", "cho40_dq"])
def test_gamma_eigenvalues_match_numpy(golden, name):
    """Gamma fold + rocSOLVER dsyevd (upper triangle) against numpy eigvalsh of
    the same fold of the oracle's blocks."""
    print(f"case {name}", flush=True)
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    asm = _gpu(ts, [st])
    o = oracle.System(oracle.Tables(ts), st).run_all()
    norb_sp = np.array([sp.norb for sp in ts.species])
    for which, mats in (("S", ("S",)), ("H", ("T", "VNA", "VNAO"))):
        ref_blocks = o[mats[0]].copy()
        for m in mats[1:]:
            ref_blocks = ref_blocks + o[m]
        ref = np.linalg.eigvalsh(_fold_upper(ref_blocks, o["neighn"], o["neigh_j"], st.imass, norb_sp))
        w = asm.gamma_eigenvalues(which)
        assert w.shape == ref.shape
        np.testing.assert_allclose(w, ref, rtol=0, atol=1e-9 * max(1.0, float(np.max(np.abs(ref)))), err_msg=which)


def _list_invariants(asm, istruct, st):
    """Size-independent properties of the lists of one structure: neigh_back
    is an involution onto the mirrored pair, and every triplet's mneigh names
    iatom's neighbour (jatom, jbeta - ibeta)."""
    from fireball_hip import synth  # noqa: F401
    import oracle as orc
    L = asm.lists(istruct)
    off = np.concatenate([[0], np.cumsum(L["neighn"])])
    back = off[L["neigh_j"] - 1] + L["neigh_back"] - 1
    atom_of_pair = np.repeat(np.arange(1, st.natoms + 1), L["neighn"])
    assert np.array_equal(L["neigh_j"][back], atom_of_pair)
    assert np.array_equal(back[back], np.arange(len(back)))
    xl = orc.make_cells(np.asarray(st.lattice), st.icluster)
    assert np.allclose(xl[L["neigh_b"][back]], -xl[L["neigh_b"]], atol=1e-9)
    tgt = off[L["com_iatom_j"] - 1] + L["com_mneigh"] - 1
    assert np.array_equal(L["neigh_j"][tgt], L["com_jatom_j"])
    d = xl[L["neigh_b"][tgt]] - (xl[L["com_jatom_b"]] - xl[L["com_iatom_b"]])
    assert np.max(np.abs(d)) < 1e-4


@pytest.mark.parametrize("cfg", [2, 4, 3])
def test_full_size_workloads_mixed_vs_fp64(cfg):
    """CHNO3000, aC8192 and the 512-structure batch at their real sizes: the
    mixed path within 1e-6 eV of binary64 on every matrix, and the list
    invariants on (a sample of) the structures."""
    from fireball_hip import synth
    print(f"config {cfg}", flush=True)
    species, structs = synth.config(cfg)
    ts = synth.make_tables(species)
    asm = _gpu(ts, structs, precision="mixed")
    ks = range(len(structs)) if len(structs) < 8 else [0, 1, 2, len(structs) // 2, len(structs) - 1]
    mixed = {m: [asm.blocks(m, k) for k in ks] for m in MATS}
    asm.set_precision("fp64")
    asm.assemble_all()
    for m in MATS:
        err = max(float(np.max(np.abs(a - asm.blocks(m, k)))) for a, k in zip(mixed[m], ks))
        print(f"config {cfg} {m} mixed vs fp64 max-abs {err:.3e} eV", flush=True)
        assert err <= TOL_MIXED, m
    for k in ks:
        _list_invariants(asm, k, structs[k])


def test_atoms_and_structures_without_triplets():
    """Tiles are packed per atom from each atom's range of target segments.
    An isolated atom and the two atoms of a dimer own no segment (no two
    distinct non-self neighbours), and a lone dimer is a whole structure with
    no triplet at all; it comes first in the batch so that the per-structure
    tile and block offsets start from an empty structure.  Lists must equal
    the oracle's bit for bit and the blocks its values, in binary64 and in the
    mixed path at every workgroup chunk."""
    from fireball_hip import synth
    box = np.eye(3) * 60.0
    dimer = np.array([[10.0, 10.0, 10.0], [12.3, 10.0, 10.0]])
    st0 = synth.Structure(dimer.copy(), np.ones(2, np.int32), box, 1)
    si8 = synth.diamond_si((1, 1, 1), sigma=0.05, seed=3).ratom + 10.0
    pos1 = np.vstack([si8, [[45.0, 45.0, 45.0]], dimer + 25.0])
    st1 = synth.Structure(pos1, np.ones(len(pos1), np.int32), box, 1)
    st2 = synth.diamond_si((1, 1, 1), sigma=0.06, seed=5)
    structs = [st0, st1, st2]
    ts = synth.make_tables([synth.SI])
    asm = _gpu(ts, structs, precision="fp64")
    tab = oracle.Tables(ts)
    ref = [oracle.System(tab, st).run_all() for st in structs]
    assert np.all(ref[0]["ncommon"] == 0)
    assert np.all(ref[1]["ncommon"][8:] == 0) and np.all(ref[1]["ncommon"][:8] > 0)
    fp64 = []
    for k in range(len(structs)):
        r = asm.run_all(k)
        for key in LIST_KEYS:
            assert np.array_equal(r[key], ref[k][key]), (k, key)
        for m in MATS:
            assert_bits(r[m], ref[k][m], f"{k} {m}")
        fp64.append(r)
    asm.set_precision("mixed")
    got = {}
    for chunk in (64, 128, 256):
        print(f"structures without triplets: mixed, chunk {chunk}", flush=True)
        asm.set_chunk3c(chunk)
        asm.assemble_all()
        got[chunk] = [asm.run_all(k)["VNA"] for k in range(len(structs))]
        for k in range(len(structs)):
            err = float(np.max(np.abs(got[chunk][k] - ref[k]["VNA"])))
            assert err <= TOL_MIXED, (chunk, k, err)
    for k in range(len(structs)):
        assert np.array_equal(got[64][k], got[128][k]) and np.array_equal(got[64][k], got[256][k]), k


@pytest.mark.parametrize("name", golden_cases.NAMES_D)
def test_d_shells_mixed_context_takes_binary64_path(golden, name):
    """Species with d shells go through the generic binary64 kernels in either
    precision mode: the mixed-precision context gives the reference's bits."""
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    r = _gpu(ts, [st], precision="mixed").run_all(0)
    n = len(g["S"])
    for m in MATS:
        assert_bits(r[m][:n], g[m], m + " vs golden")


def test_pp_shells_beyond_p_rejected(golden):
    """The pseudopotential kernels take s and p projector shells (and orbital
    shells up to d): a d projector shell stops with FB_E_UNSUPPORTED."""
    import dataclasses
    from fireball_hip import synth
    from fireball_hip.lib import Assembler, FireballError
    g = golden("spd_cluster_dq")
    ts, _ = golden_cases.inputs(g)
    species = [synth.Species(sp.nZ, list(sp.lssh), list(sp.rcutoff), [0, 2], 1.6, [-3.0, 1.5]) for sp in ts.species]
    with pytest.raises(FireballError) as e:
        Assembler(dataclasses.replace(ts, species=species))
    assert e.value.code == -11
