"""The C-ABI library without a GPU: it loads, exports every entry point the
header declares, plans batches like the Python mirror, rejects bad input, and
its records match the header layout."""
import ctypes
import os
import re

import numpy as np

import oracle
from hbtplus_amd import abi, synth

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "hbt_unbind.h")


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = set(re.findall(r"\b(hbt_[a-z_0-9]+)\s*\(", open(HEADER).read()))
    assert len(names) >= 13
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing


def test_record_layout_matches_header():
    sp, ss, sparts = oracle.abi_sizes()
    assert sp == abi.PARAMS_DTYPE.itemsize == 64
    assert ss == abi.SUB_DTYPE.itemsize == 240
    assert sparts == ctypes.sizeof(abi.Particles) == 80


def test_plan_batch_matches_python():
    lib = abi.load_library()
    cat = synth.config4(n_host=3000, n_sat=7, sat_lo=100, sat_hi=400)
    a = cat.subs.copy()
    t = ctypes.c_int64()
    assert lib.hbt_plan_batch(a.ctypes.data, len(a), ctypes.byref(t)) == 0
    b = cat.subs.copy()
    assert t.value == abi.plan_batch(b)
    assert np.array_equal(a["out_offset"], b["out_offset"])


def test_plan_batch_large_nested_batch():
    """70,000 records (the planner gathers them with several threads above
    65,536): random forests of nests up to 6 deep, children listed after and
    before their hosts; capacities and offsets equal the plain restatement's,
    and a bad parent index anywhere is an error"""
    lib = abi.load_library()
    rng = np.random.default_rng(8)
    n = 70000
    s = np.zeros(n, abi.SUB_DTYPE)
    s["count"] = rng.integers(0, 500, n)
    parent = np.full(n, -1, np.int64)
    depth = np.zeros(n, np.int64)
    for i in rng.permutation(n)[: n // 2]:
        p = int(rng.integers(0, n))
        # keep the forest acyclic and shallow: no attachment under i itself or 6 deep
        q, d = p, 0
        while q >= 0 and q != i and d < 6:
            q, d = parent[q], d + 1
        if q == i or d >= 6 or p == i:
            continue
        parent[i] = p
    s["parent"] = parent
    a, b = s.copy(), s.copy()
    t = ctypes.c_int64()
    assert lib.hbt_plan_batch(a.ctypes.data, n, ctypes.byref(t)) == 0
    assert t.value == abi.plan_batch(b)
    assert np.array_equal(a["out_offset"], b["out_offset"])
    for bad_parent in (n + 3, n - 7):  # past the batch; the record itself
        c = s.copy()
        c["parent"][n - 7] = bad_parent
        assert lib.hbt_plan_batch(c.ctypes.data, n, ctypes.byref(t)) == -1


def test_plan_batch_rejects_cycles():
    lib = abi.load_library()
    s = np.zeros(2, abi.SUB_DTYPE)
    s["count"] = 3
    s["parent"] = [1, 0]
    t = ctypes.c_int64()
    assert lib.hbt_plan_batch(s.ctypes.data, 2, ctypes.byref(t)) == -1
    assert b"cycle" in lib.hbt_last_error()


def test_null_arguments_are_errors_not_crashes():
    lib = abi.load_library()
    assert lib.hbt_unbind_batch(None, None, None, None, 1, None, None, 0) == -1
    assert lib.hbt_tree_potential(None, 0, None, None, None, None, 0, None, None, None, None, None) == -1
    assert lib.hbt_ctx_create(None, 0, None) == -1
    ctx = ctypes.c_void_p()
    # parameters the unbinding cannot run with are rejected before any device call
    for bad in (dict(bound_mass_precision=1.5), dict(bound_mass_precision=0.0), dict(max_sample_size=-1),
                dict(softening=0.0), dict(precision=7)):
        p = abi.make_params(**bad)
        assert lib.hbt_ctx_create(p.ctypes.data, 0, ctypes.byref(ctx)) == -1, bad
