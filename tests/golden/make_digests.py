"""Digests of the oracle's binary64 lists and blocks at full benchmark size.

The C oracle (oracle/fb_oracle.c) needs minutes for a 3000- or 4096-atom
structure (its neighbour search is the reference's O(729 N^2) scan), which is
too long for a GPU test.  This script runs it once, in the build container,
and writes a small JSON fixture per workload: the SHA-256 of every list array
(int32) and of every block array, their lengths, sums, and every 7976th block
value as a diagnostic sample.  Each block array has two hashes: of its bytes
as they are ("sha256") and of its values with -0.0 written as +0.0
("sha256_values"), so that equal value hashes mean equal values element for
element (what numpy's array_equal checks) and equal byte hashes mean equal
bits, signs of zeros included.  The GPU test hashes its own arrays the same
way.

    python tests/golden/make_digests.py            # Si4096 and CHNO3000
    python tests/golden/make_digests.py 1          # one workload (synth.config index)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.join(REPO, "thunder-master_amd"))

from fireball_hip import synth  # noqa: E402

LIST_KEYS = ("neighn", "neigh_j", "neigh_b", "neigh_back", "neigh_self", "ncommon", "com_iatom_j",
             "com_iatom_b", "com_jatom_j", "com_jatom_b", "com_mneigh")
MATS = ("S", "T", "DIP", "VNA", "VNAO")
SAMPLE_STRIDE = 7976
NAMES = {1: "si4096", 2: "chno3000"}


def digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a, dtype).tobytes()).hexdigest()


def summarize(r):
    """The digest record of one run_all result (oracle or GPU)."""
    out = {"lists": {}, "blocks": {}}
    for k in LIST_KEYS:
        a = np.asarray(r[k])
        out["lists"][k] = {"n": int(a.size), "sha256": digest(a, np.int32)}
    for m in MATS:
        a = np.asarray(r[m], np.float64)
        out["blocks"][m] = {"n": int(a.size), "sha256": digest(a, np.float64),
                            "sha256_values": digest(a + 0.0, np.float64), "sum": float(np.sum(a)),
                            "sumabs": float(np.sum(np.abs(a))),
                            "sample": a[::SAMPLE_STRIDE].tolist()}
    return out


def path_for(idx):
    return os.path.join(HERE, f"digest_{NAMES[idx]}.json")


def make(idx):
    import oracle
    species, structs = synth.config(idx)
    ts = synth.make_tables(species)
    t0 = time.time()
    r = oracle.System(oracle.Tables(ts), structs[0]).run_all()
    rec = summarize(r)
    rec["workload"] = f"synth.config({idx}), synth.make_tables default seed"
    rec["oracle_s"] = round(time.time() - t0, 1)
    with open(path_for(idx), "w") as f:
        json.dump(rec, f)
    print(f"{path_for(idx)}: {rec['oracle_s']} s", flush=True)


if __name__ == "__main__":
    which = [int(a) for a in sys.argv[1:]] or [1, 2]
    for i in which:
        make(i)
