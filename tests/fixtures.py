"""Loader for the golden fixtures in tests/golden/ (written by
tests/golden/make_golden.py from the reference)."""
import os

import numpy as np

from hbtplus_amd.abi import PARAMS_DTYPE, PROPS_DTYPE, SUB_DTYPE
from hbtplus_amd.synth import Catalogue

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
UNBIND_CASES = ["c1_small", "c1_small_bmp05", "c2_small", "c3_small", "c4_small", "hubble", "edge",
                "floor"]
TREE_CASES = ["tree_c1", "tree_straddle", "tree_floor"]
PROPS_CASES = ["c2_small", "c3_small", "c4_small", "hubble", "edge"]


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
