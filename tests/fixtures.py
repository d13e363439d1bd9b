"""Loader for the golden fixtures in tests/golden/ (written by
tests/golden/make_golden.py from the reference)."""
import os

import numpy as np

from hbtplus_amd.abi import PARAMS_DTYPE, PROPS_DTYPE, SUB_DTYPE
from hbtplus_amd.synth import Catalogue

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UNBIND_CASES = ["c1_small", "c1_small_bmp05", "c2_small", "c3_small", "c4_small", "hubble", "edge",
                "floor"]
# sampling on (MaxSampleSizeOfPotentialEstimate > 0), the reference driven level by level
SAMPLING_CASES = ["samp_c1", "samp_c2", "samp_c4", "samp_edge"]
TREE_CASES = ["tree_c1", "tree_straddle", "tree_floor"]
PROPS_CASES = ["c2_small", "c3_small", "c4_small", "hubble", "edge"]
MASK_CASES = ["mask_a", "mask_b"]
# masking with the member table in the reference's Mbound order (visit ranks)
MASK_MASS = "mask_mass"
MERGE_CASES = ["merge_a", "merge_b"]
NEST_CASES = ["nest_a", "nest_b"]


def load_unbind(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    cols = {k[4:]: z[k] for k in z.files if k.startswith("col_")}
    params = z["params"].view(PARAMS_DTYPE).copy()
    subs = z["subs_in"].view(SUB_DTYPE).copy()
    cat = Catalogue(cols, subs, params, str(z["name"]))
    ref = (z["ref_subs"].view(SUB_DTYPE).copy(), z["ref_index"].copy(), z["ref_energy"].copy())
    return cat, ref


def load_tree(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    params = z["params"].view(PARAMS_DTYPE).copy()
    return params, (z["x"], z["y"], z["z"], z["mass"]), z["ref_phi"]


def load_props(name):
    """(catalogue, reference unbinding outputs, properties in, reference properties)"""
    cat, ref = load_unbind(name)
    z = np.load(os.path.join(GOLDEN, "props_" + name + ".npz"))
    return cat, ref, z["props_in"].view(PROPS_DTYPE).copy(), z["ref_props"].view(PROPS_DTYPE).copy()


def load_mask(name):
    """(columns with the ID column, records, host groups, reference out_list, reference counts)"""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    n = len(z["id"])
    cols = dict(id=z["id"], x=np.zeros(n, np.float32), y=np.zeros(n, np.float32), z=np.zeros(n, np.float32),
                vx=np.zeros(n, np.float32), vy=np.zeros(n, np.float32), vz=np.zeros(n, np.float32),
                mass=np.ones(n, np.float32), type=np.ones(n, np.int32))
    return cols, z["subs"].view(SUB_DTYPE).copy(), z["host_group"], z["ref_out"], z["ref_count"]


class MergeFixture:
    """A merge fixture: params, columns, the unbinding's state (list_index,
    records, host groups) and the reference's MergeSubhalos outputs with
    MergeTrappedSubhalos on (out_*: records, lists, energies) and off; pipe:
    merging on the unbinding's state without the hand-set fields."""

    def __init__(self, name):
        z = np.load(os.path.join(GOLDEN, name + ".npz"))
        self.name = name
        self.params = z["params"].view(PARAMS_DTYPE).copy()
        self.cols = {k: z[k] for k in ("id", "x", "y", "z", "vx", "vy", "vz", "mass", "type")}
        self.list_index = z["list_index"].copy()
        self.subs = z["subs"].view(SUB_DTYPE).copy()
        self.host_group = z["host_group"].copy()
        self.out = (z["out_subs"].view(SUB_DTYPE).copy(), z["out_list"].copy(), z["out_energy"].copy())
        self.off = (z["off_subs"].view(SUB_DTYPE).copy(), z["off_list"].copy())
        self.pipe = (z["pipe_subs"].view(SUB_DTYPE).copy(), z["pipe_list"].copy(), z["pipe_energy"].copy())
        self.parent = z["parent"].copy()
        self.cores = z["cores"].copy()


def load_merge(name):
    return MergeFixture(name)


def load_nest(name):
    """(host_group, is_central, parent, rank, reference parent, reference rank)"""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return tuple(z[k] for k in ("host_group", "is_central", "parent", "rank", "ref_parent", "ref_rank"))


def load_merge_mass():
    """Merging with the member table in Mbound order: (params, cols,
    list_index, records, host groups, reference (records, lists, energies))."""
    z = np.load(os.path.join(GOLDEN, "merge_mass.npz"))
    cols = {k: z[k] for k in ("id", "x", "y", "z", "vx", "vy", "vz", "mass", "type")}
    return (z["params"].view(PARAMS_DTYPE).copy(), cols, z["list_index"].copy(), z["subs"].view(SUB_DTYPE).copy(),
            z["host_group"].copy(),
            (z["out_subs"].view(SUB_DTYPE).copy(), z["out_list"].copy(), z["out_energy"].copy()))
