"""An Fdata directory that exercises the corners of the reference's reader
(read_Fdata_2c / read_Fdata_3c): values written with d exponents and split
over two records, mvalue records with r*c repeats, and two directory records
whose iindex values are exchanged, so that each of their (itype, isubtype)
keys names the other record's cell.  (The cells of one interaction stay
contiguous in the directory: what a compiler does when a connected unit is
reopened on another file and then on the first one again is not something
the reference fixes, so no test depends on it.)"""
import os

import numpy as np

from fireball_hip import fdata_io


def _dfmt(v):
    return repr(float(v)).replace("e", "d") if "e" in repr(float(v)) else repr(float(v)) + "d0"


def _split_records(path, nme):
    rows = np.loadtxt(path, ndmin=2)
    with open(path, "w") as f:
        for row in rows:
            h = max(1, nme // 2)
            f.write(" ".join(_dfmt(v) for v in row[:h]) + "\n")
            if nme > h:
                f.write(", ".join(_dfmt(v) for v in row[h:]) + "\n")


def _repeat_mvalues(path):
    with open(path) as f:
        lines = f.read().splitlines()
    out = []
    for i, line in enumerate(lines):
        if i % 3 == 2:   # mvalue record of each triple
            vals = line.split()
            if vals and all(v == vals[0] for v in vals):
                line = f"{len(vals)}*{vals[0]}"
        out.append(line)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_quirky_fdata(ts, fdir, swap_pair=(1, 1)):
    fdata_io.write_fdata(ts, fdir)
    for name in os.listdir(fdir):
        p = os.path.join(fdir, name)
        if name.startswith("t2_") or name.startswith("t3_"):
            rows = np.loadtxt(p, ndmin=2)
            _split_records(p, rows.shape[1])
        elif "_munu_" in name:
            _repeat_mvalues(p)
    z = [sp.nZ for sp in ts.species]
    za, zb = z[swap_pair[0] - 1], z[swap_pair[1] - 1]
    dpath = os.path.join(fdir, f"2c.{za:02d}.{zb:02d}.dir")
    with open(dpath) as f:
        recs = [line.split() for line in f.read().splitlines() if line.strip()]
    # exchange the iindex of the first two ontop-L records (interaction 3,
    # subtypes 0 and 1): subtype 0 then reads subtype 1's table and back
    onl = [i for i, r in enumerate(recs) if r[1] == "3"]
    recs[onl[0]][0], recs[onl[1]][0] = recs[onl[1]][0], recs[onl[0]][0]
    with open(dpath, "w") as f:
        for r in recs:
            f.write(" ".join(r) + "\n")
