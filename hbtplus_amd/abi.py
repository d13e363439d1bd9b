"""Python view of the C-ABI in include/hbt_unbind.h.

numpy dtypes that match the C records byte for byte, a ctypes mirror of
``hbt_particles``, and the loader for the HIP library ``libhbt_unbind.so``.
The loader raises if the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HBT_OK = 0
HBT_ERR_INVALID_ARG = -1
HBT_ERR_OUT_OF_MEMORY = -2
HBT_ERR_DEVICE = -3
HBT_ERR_UNSUPPORTED = -4
HBT_ERR_CAPACITY = -5
HBT_PREC_FAST = 0
HBT_PREC_REPLAY = 1
HBT_PREC_FP32 = 2
HBT_DIRECT_FP32 = 0
HBT_DIRECT_FP32_KAHAN = 1
HBT_DIRECT_FP64 = 2
HBT_STATUS_CONVERGED = 1
HBT_STATUS_DISRUPTED = 2
HBT_STATUS_EARLY = 4
TYPE_MAX = 6

PARAMS_DTYPE = np.dtype([
    ("G", "<f4"), ("softening", "<f4"), ("open_angle", "<f4"), ("box_size", "<f4"),
    ("periodic", "<i4"), ("bound_mass_precision", "<f4"), ("min_num_part_of_sub", "<i4"),
    ("source_sub_relax_factor", "<f4"), ("max_sample_size", "<i4"), ("scale_factor", "<f4"),
    ("hz", "<f4"), ("snapshot_index", "<i4"), ("precision", "<i4"),
    ("refine_mostbound_particle", "<i4"), ("reserved", "<i4", (2,)),
], align=True)

SUB_DTYPE = np.dtype([
    ("offset", "<i8"), ("count", "<i8"), ("parent", "<i8"), ("nbound_prev", "<i8"),
    ("sink_track_id", "<i8"), ("snapshot_index_of_death", "<i4"), ("snapshot_index_of_sink", "<i4"),
    ("ref_pos", "<f4", (3,)), ("ref_vel", "<f4", (3,)), ("mostbound_pos", "<f4", (3,)),
    ("mostbound_vel", "<f4", (3,)),
    ("nbound", "<i8"), ("list_len", "<i8"), ("nsource", "<i8"), ("nbound_type", "<i8", (TYPE_MAX,)),
    ("mbound_type", "<f4", (TYPE_MAX,)), ("mbound", "<f4"), ("specific_self_potential_energy", "<f4"),
    ("specific_self_kinetic_energy", "<f4"), ("specific_angular_momentum", "<f4", (3,)),
    ("n_iter", "<i4"), ("status", "<i4"), ("n_evals", "<i8"), ("out_offset", "<i8"),
], align=True)

assert PARAMS_DTYPE.itemsize == 64, PARAMS_DTYPE.itemsize
assert SUB_DTYPE.itemsize == 240, SUB_DTYPE.itemsize

# hbt_properties (80 bytes): CalculateProfileProperties + CalculateShape outputs
PROPS_DTYPE = np.dtype([("rmax_comoving", "<f4"), ("vmax_physical", "<f4"), ("rhalf_comoving", "<f4"),
                        ("r2sigma_comoving", "<f4"), ("bound_r200crit_comoving", "<f4"),
                        ("bound_m200crit", "<f4"), ("last_max_vmax_physical", "<f4"),
                        ("snapshot_index_of_last_max_vmax", "<i4"), ("inertial_tensor", "<f4", (6,)),
                        ("inertial_tensor_weighted", "<f4", (6,))])
assert PROPS_DTYPE.itemsize == 80
PROPS_FIELDS = PROPS_DTYPE.names

# hbt_tuning (32 bytes): walk group sizes per size bin, bin edges, graph settings, XCD run length
TUNING_DTYPE = np.dtype([("walk_group", "<i4", (3,)), ("bin_edge", "<i4", (2,)),
                         ("graph_passes", "<i4"), ("use_graph", "<i4"), ("xcd_run", "<i4")])
assert TUNING_DTYPE.itemsize == 32

OUTPUT_FIELDS = ("nbound", "list_len", "nsource", "nbound_type", "mbound_type", "mbound",
                 "specific_self_potential_energy", "specific_self_kinetic_energy",
                 "specific_angular_momentum", "ref_pos", "ref_vel", "mostbound_pos",
                 "mostbound_vel", "snapshot_index_of_death", "snapshot_index_of_sink",
                 "sink_track_id")


class Particles(ctypes.Structure):
    """ctypes mirror of ``hbt_particles``."""
    _fields_ = [("n", ctypes.c_int64), ("id", ctypes.c_void_p),
                ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("z", ctypes.c_void_p),
                ("vx", ctypes.c_void_p), ("vy", ctypes.c_void_p), ("vz", ctypes.c_void_p),
                ("mass", ctypes.c_void_p), ("type", ctypes.c_void_p)]


PARTICLE_COLUMNS = ("id", "x", "y", "z", "vx", "vy", "vz", "mass", "type")


def make_params(**kw) -> np.ndarray:
    """An ``hbt_params`` record; defaults are HBT's (src/config_parser.h:68-110)
    in kpc/h, km/s, 1e10 Msun/h units, with sampling off (the reference's
    default MaxSampleSizeOfPotentialEstimate is 1000; RefineMostboundParticle
    keeps its default, on)."""
    p = np.zeros(1, PARAMS_DTYPE)
    defaults = dict(G=43007.1, softening=0.5, open_angle=0.45, box_size=1.0e5, periodic=0,
                    bound_mass_precision=0.995, min_num_part_of_sub=20,
                    source_sub_relax_factor=3.0, max_sample_size=0, scale_factor=1.0, hz=0.0,
                    snapshot_index=0, precision=HBT_PREC_FAST, refine_mostbound_particle=1)
    defaults.update(kw)
    for k, v in defaults.items():
        p[k] = v
    return p


def particles_struct(cols: dict, ptr=None) -> Particles:
    """Fill an ``hbt_particles`` from a dict of columns.  ``ptr(a)`` maps a
    column to its address (numpy: a.ctypes.data; torch: t.data_ptr())."""
    if ptr is None:
        ptr = lambda a: a.ctypes.data  # noqa: E731
    s = Particles()
    s.n = len(cols["x"])
    for k in PARTICLE_COLUMNS:
        setattr(s, k, ptr(cols[k]) if cols.get(k) is not None else None)
    return s


_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhbt_unbind.so")
_lib = None


def load_library() -> ctypes.CDLL:
    """Load libhbt_unbind.so (built in-tree by __graft_entry__.build()).
    Raises if it is missing -- the product has no CPU fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import "
                           f"__graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    sig = {
        "hbt_last_error": ([], ctypes.c_char_p),
        "hbt_version": ([vp, vp], i32),
        "hbt_ctx_create": ([vp, i32, vp], i32),
        "hbt_ctx_destroy": ([vp], i32),
        "hbt_ctx_set_params": ([vp, vp], i32),
        "hbt_ctx_stream": ([vp, vp], i32),
        "hbt_plan_batch": ([vp, i64, vp], i32),
        "hbt_unbind_batch": ([vp, vp, vp, vp, i64, vp, vp, i32], i32),
        "hbt_unbind_batch_indexed": ([vp, vp, vp, i64, vp, vp, i64, vp, vp, i32], i32),
        "hbt_unbind_one": ([vp, vp, vp, vp, vp], i32),
        "hbt_tree_potential": ([vp, i64, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp], i32),
        "hbt_direct_potential": ([vp, i64, vp, vp, vp, vp, vp, i32, vp], i32),
        "hbt_direct_potential_noalloc": ([vp, i64, vp, vp, vp, vp, vp, i32, vp], i32),
        "hbt_id_to_index": ([vp, i64, vp, i64, vp, vp], i32),
        "hbt_last_batch_stats": ([vp, vp, vp, vp], i32),
        "hbt_profile": ([vp, i32], i32),
        "hbt_profile_read": ([vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
        "hbt_profile_sorts": ([vp, vp, vp, vp, vp], i32),
        "hbt_set_rand_state": ([vp, ctypes.c_uint32, i64], i32),
        "hbt_get_rand_state": ([vp, vp, vp], i32),
        "hbt_subhalo_properties": ([vp, vp, vp, i64, vp, vp], i32),
        "hbt_mask_subhalos": ([vp, vp, vp, vp, vp, i64, vp, vp], i32),
        "hbt_merge_subhalos": ([vp, vp, vp, i64, vp, vp, vp, i64, i32, vp, i64, vp], i32),
        "hbt_nest_subhalos": ([vp, i64, vp, vp, vp, vp], i32),
        "hbt_get_tuning": ([vp, vp], i32),
        "hbt_set_tuning": ([vp, vp], i32),
        "hbt_autotune": ([vp, vp, vp, vp, i64, vp, vp, i32, vp], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != HBT_OK:
        msg = _lib.hbt_last_error().decode() if _lib is not None else ""
        raise RuntimeError(f"{what} failed with status {rc}: {msg}")


def plan_batch(subs: np.ndarray) -> int:
    """Host-side planning identical to hbt_plan_batch (no library needed):
    sets subs['out_offset'] and returns the total output length."""
    n = len(subs)
    cap = subs["count"].astype(np.int64).copy()
    parent = subs["parent"].astype(np.int64)
    depth = np.zeros(n, np.int64)
    for s in range(n):
        d, p = 0, parent[s]
        while p >= 0:
            d += 1
            p = parent[p]
            if d > n:
                raise ValueError("cycle in parent links")
        depth[s] = d
    for s in np.argsort(-depth, kind="stable"):
        if parent[s] >= 0:
            cap[parent[s]] += cap[s]
    off = np.zeros(n, np.int64)
    if n:
        off[1:] = np.cumsum(cap)[:-1]
    subs["out_offset"] = off
    return int(cap.sum())
