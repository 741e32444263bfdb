"""fb_read_positions (the in-scope steps of read_positions,
c.SYSTEM/M_configuraciones.f90:604-797) against the reference's own
read_positions, run on the same structure files by oracle/ref/fb_posharness.f90
(built from the reference's sources by `make -C oracle posref`): atomic-number
matching, the origin shift by (pi, 1/e, sqrt 2) and its 1e-4 threshold,
norbitals and iblock_slot, all bit for bit.  The library function needs no
GPU; the reference harness runs on the host."""
import math
import os
import tempfile

import numpy as np
import pytest

import refrun

need_ref = pytest.mark.skipif(not refrun.available(refrun.HARNESS_POSREF), reason="make -C oracle posref not built")


def _both(species, st, z=None):
    from fireball_hip import fdata_io
    from fireball_hip.lib import read_positions
    with tempfile.TemporaryDirectory(prefix="fbpos_") as d:
        path = os.path.join(d, "structure.inp")
        fdata_io.write_basis_file(path, species, st, z)
        return read_positions(path, species), refrun.run_positions(species, path)


def _same(got, ref):
    for k in ("natoms", "icluster", "shifted", "norbitals"):
        assert got[k] == ref[k], k
    for k in ("lattice", "imass", "iblock_slot", "ratom"):
        assert np.array_equal(got[k], ref[k]), (k, got[k], ref[k])


@need_ref
def test_atom_at_origin_shifts_everything():
    """A diamond cell has an atom exactly at the origin: every atom moves by
    (pi, 1/e, sqrt 2)."""
    from fireball_hip import synth
    st = synth.Structure(synth.DIAMOND_BASIS * 5.43, np.ones(8, np.int32), np.eye(3) * 5.43, 0)
    assert np.min(np.linalg.norm(st.ratom, axis=1)) == 0.0
    got, ref = _both([synth.SI], st)
    assert ref["shifted"] == 1
    _same(got, ref)
    shifter = np.array([4.0 * math.atan(1.0), 1.0 / math.exp(1.0), math.sqrt(2.0)])
    assert np.array_equal(got["ratom"], st.ratom + shifter)


@need_ref
def test_no_atom_near_origin_no_shift():
    from fireball_hip import synth
    st = synth.diamond_si((2, 2, 2), sigma=0.05, seed=0)
    st.ratom = st.ratom + 0.3
    got, ref = _both([synth.SI], st)
    assert ref["shifted"] == 0
    _same(got, ref)
    assert np.array_equal(got["ratom"], st.ratom)


@pytest.mark.parametrize("r0, shifted", [(0.9e-4, 1), (1.1e-4, 0)])
@need_ref
def test_shift_threshold(r0, shifted):
    """The shift applies when some atom is closer than 1e-4 A to the origin."""
    from fireball_hip import synth
    pos = np.array([[r0 / math.sqrt(3.0)] * 3, [2.0, 1.0, 0.5], [4.1, 3.3, 2.2]])
    st = synth.Structure(pos, np.ones(3, np.int32), np.eye(3) * 20.0, 1)
    got, ref = _both([synth.SI], st)
    assert ref["shifted"] == shifted
    _same(got, ref)


@need_ref
def test_species_by_atomic_number_and_orbital_slots():
    """Mixed C/H/N/O box: species matched by Z in the species order, norbitals
    and the orbital offsets follow the per-species orbital counts (H: 1)."""
    from fireball_hip import synth
    sp = synth.CHNO_SPECIES
    st = synth.random_box(synth.chno_counts(40), 1.0, 0.9, 4, sp)
    got, ref = _both(sp, st)
    _same(got, ref)
    assert np.array_equal(got["imass"], st.imass)
    norb = np.array([s.norb for s in sp])[st.imass - 1]
    assert got["norbitals"] == int(norb.sum())
    assert np.array_equal(got["iblock_slot"], np.concatenate([[0], np.cumsum(norb)[:-1]]))


def test_unknown_atomic_number_and_capacity_fail():
    from fireball_hip import fdata_io, synth
    from fireball_hip.lib import FireballError, read_positions
    st = synth.diamond_si((1, 1, 1), sigma=0.0, seed=0)
    with tempfile.TemporaryDirectory(prefix="fbpos_") as d:
        path = os.path.join(d, "structure.inp")
        fdata_io.write_basis_file(path, [synth.SI], st, z=[14] * 7 + [79])
        with pytest.raises(FireballError) as e:
            read_positions(path, [synth.SI])
        assert e.value.code == -1
        fdata_io.write_basis_file(path, [synth.SI], st)
        with pytest.raises(FireballError) as e:
            read_positions(path, [synth.SI], max_atoms=4)
        assert e.value.code == -10
