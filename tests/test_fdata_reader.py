"""The Fdata readers: fdata_io.read_fdata (Python, used with the oracle) and
fb_read_fdata (the library's native reader).  Both are pinned against the
reference's own read_Fdata_2c / read_Fdata_3c, run through the reference
harness (oracle/_ref/fb_harness_ref) on the same directory, including a
directory written with the format's corner cases (tests/fdata_quirks.py)."""
import os
import tempfile

import numpy as np
import pytest

import fdata_quirks
import oracle
import refrun
from fireball_hip import fdata_io, synth

MATS = ("S", "T", "DIP", "VNA", "VNAO")
need_ref = pytest.mark.skipif(not refrun.available(), reason="make -C oracle ref not built")


def _case():
    ts = synth.make_tables([synth.SI])
    st = synth.diamond_si((2, 2, 2), sigma=0.05, seed=4)
    synth.random_dQ([st], [synth.SI], seed=4)
    return ts, st


def test_read_fdata_round_trip():
    ts = synth.make_tables([synth.SI, synth.H_SP])
    with tempfile.TemporaryDirectory() as d:
        fdata_io.write_fdata(ts, d)
        rt = fdata_io.read_fdata(d, ts.species)
    assert set(rt.cells2c) == {k for k in ts.cells2c if k[2] in (1, 2, 3, 4, 5, 10)}
    assert set(rt.cells3c) == set(ts.cells3c)
    for k, c in rt.cells2c.items():
        o = ts.cells2c[k]
        assert c.xmax == o.xmax and np.array_equal(c.data, o.data)
        assert np.array_equal(c.mu, o.mu) and np.array_equal(c.nu, o.nu) and np.array_equal(c.mvalue, o.mvalue)
    for k, c in rt.cells3c.items():
        o = ts.cells3c[k]
        assert (c.xmax, c.ymax) == (o.xmax, o.ymax) and np.array_equal(c.data, o.data)


@need_ref
def test_python_reader_matches_reference_reader_on_quirky_directory():
    ts, st = _case()
    with tempfile.TemporaryDirectory() as work:
        fdata_quirks.write_quirky_fdata(ts, os.path.join(work, "Fdata"))
        ref = refrun.run(ts, st, workdir=work)
        rt = fdata_io.read_fdata(os.path.join(work, "Fdata"), ts.species)
    o = oracle.System(oracle.Tables(rt), st).run_all()
    for m in MATS:
        np.testing.assert_allclose(o[m], ref[m], rtol=0, atol=1e-12, err_msg=m)
    # the exchanged directory entries are visible: the plain tables give other blocks
    plain = oracle.System(oracle.Tables(ts), st).run_all()
    assert np.max(np.abs(plain["VNAO"] - ref["VNAO"])) > 1e-6


@pytest.mark.gpu
@need_ref
def test_native_reader_matches_reference_reader():
    from fireball_hip.lib import Assembler
    ts, st = _case()
    with tempfile.TemporaryDirectory() as work:
        fdir = os.path.join(work, "Fdata")
        fdata_quirks.write_quirky_fdata(ts, fdir)
        ref = refrun.run(ts, st, workdir=work)
        asm = Assembler(ts, fdata_dir=fdir)
    asm.add_structure(st)
    asm.build_lists()
    asm.assemble_all()
    for m in MATS:
        np.testing.assert_allclose(asm.blocks(m), ref[m], rtol=0, atol=1e-12, err_msg=m)
