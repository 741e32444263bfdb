"""GPU McWEDA density assemblers (fb_assemble_rho_*) through the C ABI against
the reference's own M_assemble_rho_McWEDA (golden fixtures) and the C oracle.

The kernels add every term in the reference's order, in binary64 with
contraction off: the four density arrays are compared bit for bit.  At the
full benchmark sizes, where the oracle's O(N^2) list build takes minutes,
exact properties stand in: doubling every shell charge doubles every block
exactly (binary scaling commutes with the sums), the two-centre densities are
symmetric under pair reversal, and repeated runs give the same bits."""
import copy

import numpy as np
import pytest

import golden_cases
import oracle
from fireball_hip import synth

pytestmark = pytest.mark.gpu

RKEYS = ("RHO_IN", "RHO_BOND", "RHOW_IN", "RHOW_BOND")


def _with_rho_tables(ts, seed):
    tsr = copy.copy(ts)
    tsr.cells2c = dict(ts.cells2c)
    tsr.cells3c = dict(ts.cells3c)
    return synth.add_rho_tables(tsr, seed=seed)


def _gpu_rho(ts, structs, qins, which=("ALL",)):
    from fireball_hip.lib import Assembler
    asm = Assembler(ts, precision="fp64")
    for st in structs:
        asm.add_structure(st)
    asm.build_lists()
    for k, q in enumerate(qins):
        asm.set_qin(q, k)
    asm.assemble_rho(which)
    return asm


def _same(got, ref, what):
    """Equal values (an exact zero may carry either sign)."""
    assert got.shape == ref.shape, what
    assert not np.any(np.isnan(got)), what
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"{what}: {bad.size} values differ, first {bad[:5]}, max abs {np.abs(got - ref).max():.3e}"


@pytest.mark.parametrize("name", golden_cases.NAMES + golden_cases.NAMES_D)
def test_rho_golden(golden, name):
    print(f"case {name}", flush=True)
    g, gr = golden(name), golden(name + "_rho")
    ts, st = golden_cases.inputs(g)
    ts = _with_rho_tables(ts, int(gr["rho_table_seed"]))
    asm = _gpu_rho(ts, [st], [gr["qin"]])
    for k in RKEYS:
        r = asm.rho(k)
        _same(r[:gr[k].size], gr[k], k)
        assert np.sum(r) == gr[k + "_sum"], k


def test_rho_batch_matches_oracle_per_structure():
    sp = synth.batch_species()
    ts = synth.add_rho_tables(synth.make_tables(sp, seed=31), seed=32)
    structs = [s for s in synth.batch_structures(nstruct=12, seed=4) if s.natoms <= 160][:5]
    assert len(structs) >= 3
    qins = [synth.shell_qin(s, sp, seed=40 + k) for k, s in enumerate(structs)]
    tab = oracle.Tables(ts)
    refs = []
    for st, q in zip(structs, qins):
        print(f"oracle {st.name} natoms={st.natoms}", flush=True)
        o = oracle.System(tab, st)
        o.build_lists()
        refs.append(o.run_rho(q))
    asm = _gpu_rho(ts, structs, qins)
    for k, ref in enumerate(refs):
        for key in RKEYS:
            _same(asm.rho(key, k), ref[key], f"structure {k} {key}")


def test_rho_stepwise_entry_points_equal_rho_all(golden):
    g, gr = golden("cho40_dq"), golden("cho40_dq_rho")
    ts, st = golden_cases.inputs(g)
    ts = _with_rho_tables(ts, int(gr["rho_table_seed"]))
    a = _gpu_rho(ts, [st], [gr["qin"]], which=("ALL",))
    b = _gpu_rho(ts, [st], [gr["qin"]], which=("RHO2C", "RHO3C", "RHOW2C", "RHOW3C"))
    for k in RKEYS:
        _same(a.rho(k), b.rho(k), k)


@pytest.mark.parametrize("name", ["si64_dq", "cho40_dq", "chain_cluster"])
def test_rho_sp_kernels_equal_generic_kernels(golden, name):
    """The s/p density kernels (compile-time block shapes; the 3c one in four
    phases, its elements spread over lanes) against the generic shell-loop
    kernels: same operations, same order, so the same bits."""
    g, gr = golden(name), golden(name + "_rho")
    ts, st = golden_cases.inputs(g)
    ts = _with_rho_tables(ts, int(gr["rho_table_seed"]))
    a = _gpu_rho(ts, [st], [gr["qin"]])
    from fireball_hip.lib import Assembler
    b = Assembler(ts, precision="fp64")
    b.add_structure(st)
    b.build_lists()
    b.set_tuning(3, 0)
    b.set_qin(gr["qin"])
    b.assemble_rho(("ALL",))
    for k in RKEYS:
        _same(a.rho(k), b.rho(k), k)


def test_rho_order_errors(golden):
    from fireball_hip.lib import Assembler, FireballError
    g, gr = golden("si8_cell"), golden("si8_cell_rho")
    ts, st = golden_cases.inputs(g)
    ts = _with_rho_tables(ts, int(gr["rho_table_seed"]))
    asm = Assembler(ts, precision="fp64")
    asm.add_structure(st)
    asm.build_lists()
    with pytest.raises(FireballError) as e:
        asm.assemble_rho(("RHO2C",))            # shell charges missing
    assert e.value.code == -8
    asm.set_qin(gr["qin"])
    for step in ("RHO3C", "RHOW3C"):            # 3c before its 2c blocks
        with pytest.raises(FireballError) as e:
            asm.assemble_rho((step,))
        assert e.value.code == -8
    with pytest.raises(FireballError):
        asm.rho("RHO_IN")                       # nothing assembled yet
    asm.assemble_rho(("RHO2C", "RHO3C"))
    _same(asm.rho("RHO_IN"), gr["RHO_IN"], "RHO_IN after the error paths")


def _pairs(L):
    off = np.concatenate([[0], np.cumsum(L["neighn"])]).astype(np.int64)
    pair_atom = np.repeat(np.arange(len(L["neighn"])), L["neighn"])
    back = off[L["neigh_j"] - 1] + L["neigh_back"] - 1
    return off, pair_atom, back


@pytest.mark.parametrize("cfg", [1, 2])
def test_rho_full_size_properties(cfg):
    sp, structs = synth.config(cfg)
    st = structs[0]
    ts = synth.add_rho_tables(synth.make_tables(sp), seed=4048)
    q = synth.shell_qin(st, sp, seed=5)
    print(f"config {cfg}: natoms={st.natoms}", flush=True)
    a = _gpu_rho(ts, [st], [q])
    r1 = {k: a.rho(k) for k in RKEYS}
    a.assemble_rho(("ALL",))
    for k in RKEYS:
        _same(a.rho(k), r1[k], f"repeat {k}")
    a.set_qin(2.0 * q)
    a.assemble_rho(("ALL",))
    for k in RKEYS:
        _same(a.rho(k), 2.0 * r1[k], f"doubled charges {k}")
    # pair-reversal symmetry of the two-centre densities
    L = a.lists()
    off, pair_atom, back = _pairs(L)
    for key, size in (("RHO_BOND", [s.norb for s in sp]), ("RHOW_BOND", [s.nssh for s in sp])):
        dims_a = np.array(size)[st.imass[pair_atom] - 1]
        dims_b = np.array(size)[st.imass[L["neigh_j"] - 1] - 1]
        boff = np.concatenate([[0], np.cumsum(dims_a * dims_b)])
        vals = r1[key]
        scale = np.abs(vals).max()
        rng = np.random.default_rng(cfg)
        for p in rng.choice(len(dims_a), size=4000, replace=False):
            q_ = back[p]
            A = vals[boff[p]:boff[p + 1]].reshape(dims_b[p], dims_a[p]).T
            B = vals[boff[q_]:boff[q_ + 1]].reshape(dims_b[q_], dims_a[q_]).T
            np.testing.assert_allclose(A, B.T, atol=1e-12 * scale, err_msg=f"{key} pair {p}")
    # rho_in differs from rho_bond on the 3c targets and the self blocks only
    tgt = np.zeros(len(pair_atom), bool)
    m = L["com_mneigh"] > 0
    tgt[off[L["com_iatom_j"][m] - 1] + L["com_mneigh"][m] - 1] = True
    selfs = off[:-1] + L["neigh_self"] - 1
    tgt[selfs] = True
    norb = np.array([s.norb for s in sp])
    nb = norb[st.imass[pair_atom] - 1] * norb[st.imass[L["neigh_j"] - 1] - 1]
    mask = np.repeat(~tgt, nb)
    _same(r1["RHO_IN"][mask], r1["RHO_BOND"][mask], "rho_in = rho_bond off the 3c targets")
    assert np.any(r1["RHO_IN"][~mask] != r1["RHO_BOND"][~mask])
