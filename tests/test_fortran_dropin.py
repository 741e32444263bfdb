"""The Fortran drop-in (thunder-master_amd/fortran: modules M_assemble_2c,
M_assemble_3c, M_Dassemble_2c, M_Dassemble_3c and M_assemble_rho_McWEDA over
the C ABI) linked into the same harness programs as the reference's own
modules.

The harness (oracle/ref/fb_harness.f90) runs the reference's read_Fdata_2c /
read_Fdata_3c, make_cells and driver_neighbors, then calls assemble_S, T,
dipole_z, vna_2c and vna_3c through whichever M_assemble_2c / M_assemble_3c it
was linked with, and dumps lists and blocks.  oracle/_ref/fb_harness_ref uses
the reference's modules, oracle/_ref/fb_harness_hip the drop-in ones.  Both
are built by `make -C oracle ref hip`.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import golden_cases
import refrun

HIP_HARNESS = os.path.join(os.path.dirname(refrun.HARNESS_REF), "fb_harness_hip")
LIST_KEYS = ("neighn", "neigh_j", "neigh_b", "neigh_back", "neigh_self", "ncommon", "com_iatom_j",
             "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh")
MATS = ("S", "T", "DIP", "VNA", "VNAO")
TOL_MIXED = 1e-6

need_hip = pytest.mark.skipif(not refrun.available(HIP_HARNESS), reason="make -C oracle hip not built")
# the reference's assemblers with the drop-in driver_neighbors (M_neighbors_hip.f90)
NB_HARNESS = os.path.join(os.path.dirname(refrun.HARNESS_REF), "fb_harness_nbhip")
need_nb = pytest.mark.skipif(not refrun.available(NB_HARNESS), reason="make -C oracle nbhip not built")


def _run_hip(ts, st, nrep=1, env=None, binary=HIP_HARNESS):
    print(f"{os.path.basename(binary)} natoms={st.natoms} nrep={nrep} env={env}", flush=True)
    full = dict(os.environ, **(env or {}))
    with tempfile.TemporaryDirectory(prefix="fbhip_") as d:
        fdir = os.path.join(d, "Fdata")
        from fireball_hip import fdata_io
        fdata_io.write_fdata(ts, fdir)
        inp = os.path.join(d, "harness.in")
        fdata_io.write_harness_input(inp, fdir, ts, st)
        outp = os.path.join(d, "harness.out")
        r = subprocess.run([binary, inp, outp, str(nrep)], cwd=d, env=full, capture_output=True,
                           text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(f"hip harness exit {r.returncode}: {r.stdout[-2000:]} {r.stderr[-2000:]}")
        return refrun.parse_output(outp, ts.species, st.imass)


@need_hip
def test_hip_harness_fails_loudly_without_gpu():
    """With no GPU visible the drop-in stops with the library's message and a
    nonzero status; it never falls back to CPU assembly."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    from fireball_hip import synth
    ts = synth.make_tables([synth.SI])
    st = synth.diamond_si((1, 1, 1), sigma=0.05, seed=0)
    with pytest.raises(RuntimeError, match="fb_create"):
        _run_hip(ts, st)


@need_hip
@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_dropin_matches_reference_golden(golden, name):
    print(f"case {name}", flush=True)
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    r = _run_hip(ts, st)
    for k in LIST_KEYS:
        assert np.array_equal(r[k], g[k]), k
    n = len(g["S"])
    for m in MATS:
        assert np.array_equal(r[m][:n], g[m]), (m, float(np.max(np.abs(r[m][:n] - g[m]))))
        # whole-matrix checksums of the golden vectors cover the blocks beyond n
        assert abs(float(np.sum(r[m])) - float(g[m + "_sum"])) <= 1e-9 * max(1.0, float(g[m + "_sumabs"])), m


@need_hip
@pytest.mark.gpu
def test_dropin_matches_reference_live_repeated():
    """Si216 with shell charges, assembled twice in one process (the second
    pass exercises destroy_assemble_2c / destroy_assemble_vna and the cached
    lists), against the reference modules run on the same input."""
    if not refrun.available():
        pytest.skip("reference harness not built")
    from fireball_hip import synth
    ts = synth.make_tables([synth.SI])
    st = synth.diamond_si((3, 3, 3), sigma=0.08, seed=5)
    synth.random_dQ([st], [synth.SI], seed=5)
    ref = refrun.run(ts, st)
    hip = _run_hip(ts, st, nrep=2)
    for k in LIST_KEYS:
        assert np.array_equal(hip[k], ref[k]), k
    for m in MATS:
        assert np.array_equal(hip[m], ref[m]), (m, float(np.max(np.abs(hip[m] - ref[m]))))


@need_hip
@pytest.mark.gpu
def test_dropin_vna_3c_adds_into_caller_blocks():
    """assemble_vna_3c adds the three-centre terms into the caller's
    s%vna%block (M_assemble_3c_DOGS.f90:301): a caller that changed the
    blocks after assemble_vna_2c keeps its change.  The harness adds 0.25 to
    every Vna block element between the two calls.  The library adds a
    block's three-centre terms as one sum (the reference adds them one by one
    onto the edited block), so the edited blocks agree to rounding; every
    other matrix is bit for bit the reference's."""
    if not refrun.available():
        pytest.skip("reference harness not built")
    from fireball_hip import synth
    ts = synth.make_tables([synth.SI])
    st = synth.diamond_si((2, 2, 2), sigma=0.08, seed=9)
    synth.random_dQ([st], [synth.SI], seed=10)
    ref = refrun.run(ts, st, vna_bump=0.25)
    hip = refrun.run(ts, st, binary=HIP_HARNESS, vna_bump=0.25)
    plain = refrun.run(ts, st)
    assert not np.array_equal(ref["VNA"], plain["VNA"])
    for m in MATS:
        if m == "VNA":
            np.testing.assert_allclose(hip[m], ref[m], rtol=0, atol=1e-13 * max(1.0, float(np.max(np.abs(ref[m])))))
        else:
            assert np.array_equal(hip[m], ref[m]), (m, float(np.max(np.abs(hip[m] - ref[m]))))


@need_hip
@pytest.mark.gpu
def test_dropin_mixed_precision_within_tolerance(golden):
    g = golden("cho40_dq")
    ts, st = golden_cases.inputs(g)
    r = _run_hip(ts, st, env={"FIREBALL_HIP_PRECISION": "mixed"})
    n = len(g["S"])
    for m in MATS:
        np.testing.assert_allclose(r[m][:n], g[m], rtol=0, atol=TOL_MIXED, err_msg=m)


# ---------------------------------------------------------------- forces
DHIP_HARNESS = os.path.join(os.path.dirname(refrun.HARNESS_REF), "fb_dharness_hip")
need_dhip = pytest.mark.skipif(not refrun.available(DHIP_HARNESS), reason="make -C oracle hip not built")
DKEYS = ("DS", "DT", "DDIP")
F2KEYS = ("vna_atom", "vna_ontop")
F3KEYS = ("f3naa", "f3nab", "f3nac")


def _scale(a):
    return max(1.0, float(np.max(np.abs(a)))) if np.size(a) else 1.0


@need_dhip
@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_force_dropin_matches_reference_golden(golden, name):
    """The force harness (oracle/ref/fb_dharness.f90) linked with the drop-in
    M_Dassemble_2c / M_Dassemble_3c: Dblocks and force pieces against the
    reference's own Fortran (golden fixtures)."""
    from fireball_hip import synth
    print(f"force case {name}", flush=True)
    g, gf = golden(name), golden(name + "_forces")
    ts, st = golden_cases.inputs(g)
    lists = {k: g[k] for k in ("neighn", "neigh_j", "neigh_back")}
    rho = synth.denmat(ts.species, st.imass, lists["neighn"], lists["neigh_j"], lists["neigh_back"],
                       seed=int(gf["denmat_seed"]))
    r = refrun.run_forces(ts, st, rho, lists, binary=DHIP_HARNESS)
    for k in DKEYS:
        n = len(gf[k])
        np.testing.assert_allclose(r[k][:n], gf[k], rtol=0, atol=1e-12 * _scale(gf[k]), err_msg=k)
    for k in F2KEYS:
        np.testing.assert_allclose(r[k], gf[k], rtol=0, atol=1e-12 * _scale(gf[k]), err_msg=k)
    for k in F3KEYS:
        np.testing.assert_allclose(r[k], gf[k], rtol=0, atol=1e-11 * _scale(gf[k]), err_msg=k)


@need_dhip
@pytest.mark.gpu
def test_force_dropin_matches_reference_live_repeated():
    """Si216 with shell charges, force assemblers run twice in one process (the
    second pass exercises destroy_Dassemble_2c and the re-uploaded density
    matrix), against the reference's force modules on the same input."""
    if not refrun.available(refrun.HARNESS_DREF):
        pytest.skip("reference force harness not built")
    from fireball_hip import synth
    ts = synth.make_tables([synth.SI])
    st = synth.diamond_si((3, 3, 3), sigma=0.08, seed=5)
    synth.random_dQ([st], [synth.SI], seed=5)
    lists = refrun.run(ts, st)
    rho = synth.denmat(ts.species, st.imass, lists["neighn"], lists["neigh_j"], lists["neigh_back"], seed=3)
    ref = refrun.run_forces(ts, st, rho, lists)
    hip = refrun.run_forces(ts, st, rho, lists, binary=DHIP_HARNESS, nrep=2)
    for k in DKEYS + F2KEYS:
        np.testing.assert_allclose(hip[k], ref[k], rtol=0, atol=1e-12 * _scale(ref[k]), err_msg=k)
    for k in F3KEYS:
        np.testing.assert_allclose(hip[k], ref[k], rtol=0, atol=1e-11 * _scale(ref[k]), err_msg=k)


@need_dhip
@pytest.mark.gpu
def test_force_dropin_reads_charges_at_call_time():
    """A charge mixer rewrites s%atom(:)%shell(:)%dQ after the last
    assemble_vna and before the force assemblers; the reference's
    Dassemble_vna_2c / _3c read the new charges.  The harness scales every dQ
    by 0.5 between the two (fb_dharness.f90, fifth argument): the drop-in must
    give the reference's forces for the new charges."""
    if not refrun.available(refrun.HARNESS_DREF):
        pytest.skip("reference force harness not built")
    from fireball_hip import synth
    ts = synth.make_tables([synth.SI])
    st = synth.diamond_si((2, 2, 2), sigma=0.08, seed=6)
    synth.random_dQ([st], [synth.SI], seed=8, scale=0.2)
    lists = refrun.run(ts, st)
    rho = synth.denmat(ts.species, st.imass, lists["neighn"], lists["neigh_j"], lists["neigh_back"], seed=4)
    ref = refrun.run_forces(ts, st, rho, lists, dq_scale=0.5)
    same = refrun.run_forces(ts, st, rho, lists)
    hip = refrun.run_forces(ts, st, rho, lists, binary=DHIP_HARNESS, dq_scale=0.5)
    # the charged terms do change with dQ, so the check is not vacuous
    assert np.max(np.abs(ref["f3naa"] - same["f3naa"])) > 1e-6
    for k in DKEYS + F2KEYS:
        np.testing.assert_allclose(hip[k], ref[k], rtol=0, atol=1e-12 * _scale(ref[k]), err_msg=k)
    for k in F3KEYS:
        np.testing.assert_allclose(hip[k], ref[k], rtol=0, atol=1e-11 * _scale(ref[k]), err_msg=k)


# ---------------------------------------------------------- McWEDA densities
RHIP_HARNESS = os.path.join(os.path.dirname(refrun.HARNESS_REF), "fb_rharness_hip")
need_rhip = pytest.mark.skipif(not refrun.available(RHIP_HARNESS), reason="make -C oracle hip not built")
RKEYS = ("RHO_IN", "RHO_BOND", "RHOW_IN", "RHOW_BOND")


def _rho_tables(ts, seed):
    import copy
    from fireball_hip import synth
    tsr = copy.copy(ts)
    tsr.cells2c = dict(ts.cells2c)
    tsr.cells3c = dict(ts.cells3c)
    return synth.add_rho_tables(tsr, seed=seed)


@need_rhip
@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_rho_dropin_matches_reference_golden(golden, name):
    """The density harness (oracle/ref/fb_rharness.f90) linked with the drop-in
    M_assemble_rho_McWEDA: the four density arrays equal the reference's own
    (golden fixtures) value for value."""
    print(f"density case {name}", flush=True)
    g, gr = golden(name), golden(name + "_rho")
    ts, st = golden_cases.inputs(g)
    ts = _rho_tables(ts, int(gr["rho_table_seed"]))
    lists = {k: g[k] for k in ("neighn", "neigh_j")}
    r = refrun.run_rho(ts, st, gr["qin"], lists, binary=RHIP_HARNESS)
    for k in RKEYS:
        assert np.array_equal(r[k][:gr[k].size], gr[k]), k
        assert float(np.sum(r[k])) == float(gr[k + "_sum"]), k


@need_rhip
@pytest.mark.gpu
def test_rho_dropin_matches_reference_live_repeated():
    """Si216 with shell charges, the four density assemblers run twice in one
    process (the second pass exercises destroy_assemble_rho and the cached
    lists), against the reference's module on the same input."""
    if not refrun.available(refrun.HARNESS_RREF):
        pytest.skip("reference density harness not built")
    from fireball_hip import synth
    ts = _rho_tables(synth.make_tables([synth.SI]), 77)
    st = synth.diamond_si((3, 3, 3), sigma=0.08, seed=5)
    synth.random_dQ([st], [synth.SI], seed=5)
    qin = synth.shell_qin(st, [synth.SI])
    lists = refrun.run(ts, st)
    ref = refrun.run_rho(ts, st, qin, lists)
    hip = refrun.run_rho(ts, st, qin, lists, binary=RHIP_HARNESS, nrep=2)
    for k in RKEYS:
        assert np.array_equal(hip[k], ref[k]), k


# ------------------------------------------------------------ pseudopotential
PHIP_HARNESS = os.path.join(os.path.dirname(refrun.HARNESS_REF), "fb_pharness_hip")
need_phip = pytest.mark.skipif(not refrun.available(PHIP_HARNESS), reason="make -C oracle hip not built")


@need_phip
@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_pp_dropin_matches_reference_golden(golden, name):
    """The PP harness (oracle/ref/fb_pharness.f90: the caller's
    driver_neighbors_PP, then assemble_svnl, assemble_vnl_2c, assemble_vnl_3c)
    linked with the drop-in M_assemble_PP_2c / M_assemble_PP_3c: lists, sVNL
    and vnl equal the reference's own (golden fixtures)."""
    from golden.make_golden import PP_LIST_KEYS, pp_tables
    print(f"pp case {name}", flush=True)
    g, gp = golden(name), golden(name + "_pp")
    ts, st = golden_cases.inputs(g)
    r = refrun.run_pp(pp_tables(ts, int(gp["pp_table_seed"])), st, binary=PHIP_HARNESS)
    for k in PP_LIST_KEYS:
        assert np.array_equal(r[k], gp[k]), k
    for k in ("SVNL", "VNL"):
        assert np.array_equal(r[k][:gp[k].size], gp[k]), k
        assert float(np.sum(r[k])) == float(gp[k + "_sum"]), k


@need_phip
@pytest.mark.gpu
def test_pp_dropin_matches_reference_live_repeated():
    """Si216, the PP assemblers run twice in one process (destroy_assemble_PP_2c
    and the cached PP lists), against the reference's modules on the same input."""
    if not refrun.available(refrun.HARNESS_PREF):
        pytest.skip("reference PP harness not built")
    from fireball_hip import synth
    sp = synth.with_pp([synth.SI])
    ts = synth.add_pp_tables(synth.make_tables(sp), seed=91)
    st = synth.diamond_si((3, 3, 3), sigma=0.08, seed=5)
    ref = refrun.run_pp(ts, st)
    hip = refrun.run_pp(ts, st, binary=PHIP_HARNESS, nrep=2)
    for k in ref:
        if k in ("times", "nrep"):
            continue
        assert np.array_equal(hip[k], ref[k]), k


@need_nb
@pytest.mark.gpu
@pytest.mark.parametrize("name", golden_cases.NAMES)
def test_driver_neighbors_dropin_golden(golden, name):
    """driver_neighbors from the GPU lists under the reference's own
    assemblers: every list array and every block equal to the reference's run
    (the blocks can only match if the lists match entry for entry, in order)."""
    print(f"case {name}", flush=True)
    g = golden(name)
    ts, st = golden_cases.inputs(g)
    r = _run_hip(ts, st, binary=NB_HARNESS)
    for k in LIST_KEYS:
        assert np.array_equal(r[k], g[k]), k
    n = len(g["S"])
    for m in MATS:
        assert np.array_equal(r[m][:n], g[m]), (m, float(np.max(np.abs(r[m][:n] - g[m]))))


@need_nb
@pytest.mark.gpu
def test_driver_neighbors_dropin_chno_box():
    """A 300-atom C/H/N/O box (four species, mixed s and sp orbitals): lists and
    blocks of the drop-in run equal the reference harness's, run here on the
    same input."""
    from fireball_hip import synth
    sp = synth.CHNO_SPECIES
    ts = synth.make_tables(sp)
    st = synth.random_box(synth.chno_counts(300), 1.0, 0.9, 7, sp)
    ref = refrun.run(ts, st)
    r = _run_hip(ts, st, binary=NB_HARNESS)
    for k in LIST_KEYS:
        assert np.array_equal(r[k], ref[k]), k
    for m in MATS:
        assert np.array_equal(r[m], ref[m]), m
