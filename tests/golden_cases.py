"""Rebuild the inputs of a golden fixture (tables from their seed, structure
from the stored positions)."""
import numpy as np

from fireball_hip import synth

SPECIES_BY_Z = {sp.nZ: sp for sp in (synth.SI, synth.C_SP, synth.H_SP, synth.N_SP, synth.O_SP, synth.FE_SPD)}
NAMES = ("si64", "si64_dq", "si8_cell", "cho40_dq", "chain_cluster")
# species with d shells: the H+S assemblers only (generic binary64 kernels)
NAMES_D = ("spd_cluster_dq", "spd_cell")


def inputs(g):
    species = [SPECIES_BY_Z[int(z)] for z in g["species_z"]]
    ts = synth.make_tables(species, seed=int(g["table_seed"]))
    dq = g["dQ"] if np.any(g["dQ"] != 0) else None
    st = synth.Structure(np.array(g["ratom"]), np.array(g["imass"], np.int32), np.array(g["lattice"]),
                         int(g["icluster"]), dQ=dq)
    return ts, st

