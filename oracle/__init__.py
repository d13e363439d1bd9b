"""ctypes wrapper of the CPU restatement oracle (liboracle.so) and a runner
for the reference harness (oracle/_ref/hbt_ref_harness).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

from hbtplus_amd.abi import PARAMS_DTYPE, PROPS_DTYPE, SUB_DTYPE, particles_struct, plan_batch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "hbt_ref_harness")
REF_TRACKING = os.path.join(os.path.dirname(REF_HARNESS), "hbt_ref_tracking")
_lib = None


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.oracle_unbind_batch.argtypes = [vp, vp, vp, i64, vp, vp, i32]
        L.oracle_unbind_batch_rs.argtypes = [vp, vp, vp, i64, vp, vp, i32, ctypes.c_uint32, vp]
        L.oracle_glibc_rand.argtypes = [ctypes.c_uint32, i64, i64, vp]
        L.oracle_random_shuffle.argtypes = [ctypes.c_uint32, i64, i64, vp, vp]
        L.oracle_tree_potential.argtypes = [vp, i64, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp]
        L.oracle_tree_ncell.argtypes = [vp, i64, vp, vp, vp, vp]
        L.oracle_tree_ncell.restype = i64
        L.oracle_direct_potential.argtypes = [vp, i64, vp, vp, vp, vp, vp, i32, vp, i32]
        L.oracle_id_to_index.argtypes = [i64, vp, i64, vp, vp]
        L.oracle_subhalo_properties.argtypes = [vp, vp, vp, i64, vp, vp, i32]
        L.oracle_mask_subhalos.argtypes = [vp, vp, vp, vp, i64, vp, vp]
        L.oracle_merge_subhalos.argtypes = [vp, vp, vp, i64, vp, vp, vp, i64, i32, vp, vp, i64, vp]
        L.oracle_nest_subhalos.argtypes = [i64, vp, vp, vp, vp]
        L.oracle_merge_cores.argtypes = [vp, vp, vp, vp, i64, vp]
        L.oracle_abi_sizes.argtypes = [vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def unbind_batch(cat, nthreads: int = 1, params=None, rand_state=None):
    """Run the restatement on a catalogue.  Returns (subs_out, out_index, out_energy).
    rand_state: {"seed": s, "draws": d}, the rand() stream the sampling
    shuffles draw from (srand(s) then d calls; default srand(1), no calls);
    "draws" is updated in place."""
    params = cat.params if params is None else params
    cols = {k: np.ascontiguousarray(v) for k, v in cat.cols.items()}
    subs = cat.subs.copy()
    total = plan_batch(subs)
    out_index = np.full(total, -1, np.int64)
    out_energy = np.full(total, np.nan, np.float32)
    ps = particles_struct(cols)
    rs = rand_state if rand_state is not None else {"seed": 1, "draws": 0}
    draws = ctypes.c_int64(int(rs["draws"]))
    rc = lib().oracle_unbind_batch_rs(_p(params), ctypes.addressof(ps), _p(subs), len(subs),
                                      _p(out_index), _p(out_energy), nthreads,
                                      ctypes.c_uint32(int(rs["seed"])), ctypes.byref(draws))
    if rc != 0:
        raise RuntimeError(f"oracle_unbind_batch failed: {rc}")
    rs["draws"] = draws.value
    return subs, out_index, out_energy


def glibc_rand(n, seed=1, skip=0):
    """n values of the restated glibc rand() after srand(seed) and skip calls."""
    out = np.zeros(max(n, 1), np.int32)
    if lib().oracle_glibc_rand(seed, skip, n, _p(out)) != 0:
        raise RuntimeError("oracle_glibc_rand failed")
    return out[:n]


def random_shuffle(lengths, seed=1, skip=0):
    """The restated std::random_shuffle of each range [0, len) in turn
    (concatenated), drawing from srand(seed) after skip calls."""
    ln = np.ascontiguousarray(lengths, np.int64)
    out = np.zeros(max(int(ln.sum()), 1), np.int64)
    if lib().oracle_random_shuffle(seed, skip, len(ln), _p(ln), _p(out)) != 0:
        raise RuntimeError("oracle_random_shuffle failed")
    return out[:int(ln.sum())]


def tree_potential(params, src, tgt, tmass=None):
    sx, sy, sz, sm = (_f32(a) for a in src)
    tx, ty, tz = (_f32(a) for a in tgt)
    tm = _f32(tmass) if tmass is not None else np.zeros(len(tx), np.float32)
    out = np.zeros(len(tx), np.float64)
    rc = lib().oracle_tree_potential(_p(params), len(sx), _p(sx), _p(sy), _p(sz), _p(sm), len(tx),
                                     _p(tx), _p(ty), _p(tz), _p(tm), _p(out))
    if rc != 0:
        raise RuntimeError(f"oracle_tree_potential failed: {rc}")
    return out


def tree_ncell(params, src):
    sx, sy, sz, sm = (_f32(a) for a in src)
    return lib().oracle_tree_ncell(_p(params), len(sx), _p(sx), _p(sy), _p(sz), _p(sm))


def direct_potential(params, seg_offset, x, y, z, m, mode, nthreads=1):
    seg = np.ascontiguousarray(seg_offset, np.int64)
    x, y, z, m = (_f32(a) for a in (x, y, z, m))
    out = np.zeros(len(x), np.float64)
    rc = lib().oracle_direct_potential(_p(params), len(seg) - 1, _p(seg), _p(x), _p(y), _p(z),
                                       _p(m), mode, _p(out), nthreads)
    if rc != 0:
        raise RuntimeError("oracle_direct_potential failed")
    return out


def subhalo_properties(params, cols, subs, out_index, props_in=None, nthreads=1):
    """CalculateProfileProperties + CalculateShape restated, from unbinding
    outputs (records and lists); props_in holds the in/out fields."""
    c = {k: np.ascontiguousarray(v) for k, v in cols.items()}
    ps = particles_struct(c)
    subs = np.ascontiguousarray(subs, SUB_DTYPE)
    oi = np.ascontiguousarray(out_index, np.int64)
    props = (np.zeros(len(subs), PROPS_DTYPE) if props_in is None
             else np.ascontiguousarray(props_in, PROPS_DTYPE).copy())
    rc = lib().oracle_subhalo_properties(_p(params), ctypes.byref(ps), _p(subs), len(subs), _p(oi),
                                         _p(props), nthreads)
    if rc != 0:
        raise RuntimeError("oracle_subhalo_properties failed")
    return props


def _rank(visit_rank, n):
    return None if visit_rank is None else np.ascontiguousarray(visit_rank, np.int64)


def mass_visit_rank(subs, host_group):
    """The reference's visit order as ranks: a group's members by Mbound,
    most massive first (MemberShipTable_t, src/subhalo_tracking.cpp:96-104;
    the records' mbound must not tie within a group), nested subhalos in
    index order (NestedSubhalos as the catalogue lists them)."""
    grp = np.asarray(host_group, np.int64)
    order = np.lexsort((-np.asarray(subs["mbound"], np.float64), grp))
    rank = np.empty(len(subs), np.int64)
    start = np.searchsorted(grp[order], grp[order], side="left")
    rank[order] = np.arange(len(order)) - start
    nested = np.asarray(subs["parent"]) >= 0
    rank[nested] = np.nonzero(nested)[0]
    return rank


def mask_subhalos(cols, subs, host_group, visit_rank=None):
    """MaskSubhalos restated: (out_list laid out like the columns, new counts).
    visit_rank: sibling order (None: index order)."""
    c = {k: np.ascontiguousarray(v) for k, v in cols.items()}
    ps = particles_struct(c)
    subs = np.ascontiguousarray(subs, SUB_DTYPE)
    grp = np.ascontiguousarray(host_group, np.int64)
    out = np.full(len(c["id"]), -1, np.int64)
    cnt = np.zeros(len(subs), np.int64)
    vr = _rank(visit_rank, len(subs))
    rc = lib().oracle_mask_subhalos(ctypes.byref(ps), _p(subs), _p(grp), None if vr is None else _p(vr),
                                    len(subs), _p(out), _p(cnt))
    if rc != 0:
        raise RuntimeError("oracle_mask_subhalos failed")
    return out, cnt


def merge_cores(params, cols, list_index, subs):
    """The merging cores restated: float[nsub, 8] (position, velocity,
    sigma_R, sigma_V)."""
    c = {k: np.ascontiguousarray(v) for k, v in cols.items()}
    ps = particles_struct(c)
    li = np.ascontiguousarray(list_index, np.int64)
    subs = np.ascontiguousarray(subs, SUB_DTYPE)
    out = np.zeros((len(subs), 8), np.float32)
    rc = lib().oracle_merge_cores(_p(params), ctypes.byref(ps), _p(li), _p(subs), len(subs), _p(out))
    if rc != 0:
        raise RuntimeError("oracle_merge_cores failed")
    return out


def ref_merge_cores(params, cols, list_index, subs):
    """The reference's own SubHelper_t::BuildPosition / BuildVelocity
    (oracle/_ref/hbt_ref_tracking mode cores): float[nsub, 8]."""
    n = len(cols["x"])
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.ascontiguousarray(params, PARAMS_DTYPE).tobytes())
            f.write(np.array([len(subs), n, len(list_index)], np.int64).tobytes())
            for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):
                f.write(_f32(cols[k]).tobytes())
            f.write(np.ascontiguousarray(list_index, np.int64).tobytes())
            f.write(np.ascontiguousarray(subs, SUB_DTYPE).tobytes())
        subprocess.run([REF_TRACKING, "cores", fin, fout], check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        return np.fromfile(fout, np.float32).reshape(len(subs), 8)


def nest_subhalos(host_group, is_central, parent, rank):
    """LevelUpDetachedSubhalos restated: (parent, rank) afterwards."""
    g = np.ascontiguousarray(host_group, np.int64)
    c = np.ascontiguousarray(is_central, np.int32)
    p = np.ascontiguousarray(parent, np.int64).copy()
    r = np.ascontiguousarray(rank, np.int32).copy()
    rc = lib().oracle_nest_subhalos(len(g), _p(g), _p(c), _p(p), _p(r))
    if rc != 0:
        raise RuntimeError(f"oracle_nest_subhalos failed: {rc}")
    return p, r


def ref_nest_subhalos(host_group, is_central, parent, rank):
    """The reference's own LevelUpDetachedSubhalos
    (oracle/_ref/hbt_ref_tracking mode nest): (parent, rank) afterwards."""
    n = len(host_group)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([n], np.int64).tobytes())
            f.write(np.ascontiguousarray(host_group, np.int64).tobytes())
            f.write(np.ascontiguousarray(is_central, np.int32).tobytes())
            f.write(np.ascontiguousarray(parent, np.int64).tobytes())
            f.write(np.ascontiguousarray(rank, np.int32).tobytes())
        subprocess.run([REF_TRACKING, "nest", fin, fout], check=True, env=dict(os.environ, OMP_NUM_THREADS="1"))
        raw = open(fout, "rb").read()
    return np.frombuffer(raw, np.int64, n, 0).copy(), np.frombuffer(raw, np.int32, n, 8 * n).copy()


def merge_subhalos(params, cols, list_index, subs, host_group, merge_trapped=True, visit_rank=None):
    """MergeSubhalos restated (one thread): (records after merging, their
    lists as column positions, packed in subhalo order, and the binding
    energies along the lists of the subhalos unbound again, NaN elsewhere)."""
    c = {k: np.ascontiguousarray(v) for k, v in cols.items()}
    ps = particles_struct(c)
    li = np.ascontiguousarray(list_index, np.int64)
    subs = np.ascontiguousarray(subs, SUB_DTYPE).copy()
    grp = np.ascontiguousarray(host_group, np.int64)
    cap = int(subs["count"].sum()) + len(subs)
    out = np.full(max(cap, 1), -1, np.int64)
    en = np.full(max(cap, 1), np.nan, np.float32)
    used = ctypes.c_int64()
    vr = _rank(visit_rank, len(subs))
    rc = lib().oracle_merge_subhalos(_p(params), ctypes.byref(ps), _p(li), len(li), _p(subs), _p(grp),
                                     None if vr is None else _p(vr), len(subs), 1 if merge_trapped else 0,
                                     _p(out), _p(en), cap, ctypes.byref(used))
    if rc != 0:
        raise RuntimeError(f"oracle_merge_subhalos failed: {rc}")
    return subs, out[:used.value].copy(), en[:used.value].copy()


def id_to_index(table_ids, query_ids):
    t = np.ascontiguousarray(table_ids, np.int64)
    q = np.ascontiguousarray(query_ids, np.int64)
    out = np.zeros(len(q), np.int64)
    lib().oracle_id_to_index(len(t), _p(t), len(q), _p(q), _p(out))
    return out


def abi_sizes():
    a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    lib().oracle_abi_sizes(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return a.value, b.value, c.value


# ---- reference harness (development container only) ----------------------

def have_reference() -> bool:
    return os.path.exists(REF_HARNESS)


def _write_catalogue(path, cat, params):
    cols = cat.cols
    n = len(cols["x"])
    with open(path, "wb") as f:
        f.write(np.ascontiguousarray(params, PARAMS_DTYPE).tobytes())
        f.write(np.array([len(cat.subs), n], np.int64).tobytes())
        f.write(np.ascontiguousarray(cols["id"], np.int64).tobytes())
        for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):
            f.write(_f32(cols[k]).tobytes())
        f.write(np.ascontiguousarray(cols["type"], np.int32).tobytes())
        f.write(np.ascontiguousarray(cat.subs, SUB_DTYPE).tobytes())


def ref_time_unbind(cat, nthreads, params=None):
    """Wall seconds of the reference's RecursiveUnbind over the catalogue's
    nests, OpenMP over nests with schedule(dynamic,1) as RefineParticles runs
    them (timed inside oracle/_ref/hbt_ref_harness, input/output excluded)."""
    params = cat.params if params is None else params
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        _write_catalogue(fin, cat, params)
        env = dict(os.environ, OMP_NUM_THREADS=str(int(nthreads)))
        subprocess.run([REF_HARNESS, "time", fin, fout], check=True, env=env)
        return float(np.fromfile(fout, np.float64, 1)[0])


def kick_null_particles(cat):
    """Subhalo_t::KickNullParticles (src/subhalo.cpp:446-477) on every list of
    a catalogue: entries whose ID is the null ID (-1) are dropped, the rest
    keep their order, and the Nbound the orphan rule reads (nbound_prev) loses
    the null entries among its first nbound_prev.  Returns (a catalogue with
    the lists compacted, rows: for each kept entry its row in cat.cols)."""
    from hbtplus_amd.synth import Catalogue
    subs = cat.subs.copy()
    cnt = subs["count"].astype(np.int64)
    seg = np.repeat(np.arange(len(subs)), cnt)
    pos = np.arange(len(seg)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    rows = np.repeat(subs["offset"], cnt) + pos
    keep = cat.cols["id"][rows] != -1
    front = pos < np.repeat(np.minimum(subs["nbound_prev"], cnt), cnt)
    subs["nbound_prev"] = np.bincount(seg[keep & front], minlength=len(subs))
    subs["count"] = np.bincount(seg[keep], minlength=len(subs))
    subs["offset"] = np.concatenate([[0], np.cumsum(subs["count"])[:-1]]) if len(subs) else subs["offset"]
    kept = rows[keep]
    cols = {k: v[kept] for k, v in cat.cols.items()}
    return Catalogue(cols, subs, cat.params.copy(), cat.name + " (null particles kicked)"), kept


def ref_unbind_batch(cat, params=None, props_in=None, levels=False, kick=False):
    """Run the reference's own RecursiveUnbind/Unbind/TruncateSource on the
    catalogue through oracle/_ref/hbt_ref_harness (single thread).  With
    props_in (an array of PROPS_DTYPE): also CalculateProfileProperties and
    CalculateShape on every subhalo after unbinding; returns the properties
    as a fourth element."""
    params = cat.params if params is None else params
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        _write_catalogue(fin, cat, params)
        if props_in is not None:
            with open(fin, "ab") as f:
                f.write(np.ascontiguousarray(props_in, PROPS_DTYPE).tobytes())
        env = dict(os.environ, OMP_NUM_THREADS="1")
        mode = "kick" if kick else ("levels" if levels else ("unbind" if props_in is None else "props"))
        subprocess.run([REF_HARNESS, mode, fin, fout], check=True, env=env)
        raw = open(fout, "rb").read()
    nsub = int(np.frombuffer(raw, np.int64, 1)[0])
    pos = 16
    subs = np.zeros(nsub, SUB_DTYPE)
    plan = cat.subs.copy()
    total = plan_batch(plan)
    out_index = np.full(total, -1, np.int64)
    out_energy = np.full(total, np.nan, np.float32)
    for s in range(nsub):
        r = np.frombuffer(raw, SUB_DTYPE, 1, pos)
        pos += SUB_DTYPE.itemsize
        subs[s] = r[0]
        L = int(r["list_len"][0])
        o = int(plan["out_offset"][s])
        subs["out_offset"][s] = o
        out_index[o:o + L] = np.frombuffer(raw, np.int64, L, pos)
        pos += 8 * L
        out_energy[o:o + L] = np.frombuffer(raw, np.float32, L, pos)
        pos += 4 * L
    if props_in is not None:
        props = np.frombuffer(raw, PROPS_DTYPE, nsub, pos).copy()
        return subs, out_index, out_energy, props
    return subs, out_index, out_energy


def ref_mask_subhalos(cols, subs, host_group, by_mass=False):
    """The reference's own SubhaloSnapshot_t::MaskSubhalos on the catalogue
    (oracle/_ref/hbt_ref_tracking): (out_list laid out like the columns, new
    counts), as oracle.mask_subhalos returns them.  by_mass: the member table
    ordered by the records' mbound (the reference's own order); otherwise by
    index."""
    ids = np.ascontiguousarray(cols["id"], np.int64)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.array([len(subs), len(ids)], np.int64).tobytes())
            f.write(ids.tobytes())
            f.write(np.ascontiguousarray(subs, SUB_DTYPE).tobytes())
            f.write(np.ascontiguousarray(host_group, np.int64).tobytes())
        subprocess.run([REF_TRACKING, "mask_mass" if by_mass else "mask", fin, fout], check=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        raw = np.fromfile(fout, np.int64)
    return raw[:len(ids)].copy(), raw[len(ids):].copy()


def merge_input(cat, subs_out, out_index):
    """The state MergeSubhalos starts from, built from an unbinding's outputs
    (records and output lists): each subhalo's list truncated to nsource
    (TruncateSource), the lists concatenated into list_index and the records'
    offset/count pointed at them.  Returns (list_index, records)."""
    subs = np.ascontiguousarray(subs_out, SUB_DTYPE).copy()
    parts, off = [], 0
    for s in range(len(subs)):
        o, n = int(subs["out_offset"][s]), int(subs["nsource"][s])
        parts.append(np.asarray(out_index[o:o + n], np.int64))
        subs["offset"][s], subs["count"][s] = off, n
        off += n
    subs["parent"] = cat.subs["parent"]
    lst = np.concatenate(parts) if parts else np.zeros(0, np.int64)
    return lst, subs


def ref_merge_subhalos(params, cols, list_index, subs, host_group, merge_trapped=True, time_threads=0,
                       by_mass=False):
    """The reference's own SubhaloSnapshot_t::MergeSubhalos
    (oracle/_ref/hbt_ref_tracking, single thread) on the state given by the
    columns, the lists (column positions) and the records.  Returns (records
    after merging, their lists as column positions, per-entry energies: the
    Energies of the bound block of every subhalo unbound again, NaN
    elsewhere).  time_threads > 0: returns only the wall seconds the
    reference spends in MergeSubhalos on that many OpenMP threads."""
    n = len(cols["x"])
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.ascontiguousarray(params, PARAMS_DTYPE).tobytes())
            f.write(np.array([1 if merge_trapped else 0, len(subs), n, len(list_index)], np.int64).tobytes())
            f.write(np.ascontiguousarray(cols["id"], np.int64).tobytes())
            for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):
                f.write(_f32(cols[k]).tobytes())
            f.write(np.ascontiguousarray(cols["type"], np.int32).tobytes())
            f.write(np.ascontiguousarray(list_index, np.int64).tobytes())
            f.write(np.ascontiguousarray(subs, SUB_DTYPE).tobytes())
            f.write(np.ascontiguousarray(host_group, np.int64).tobytes())
        nt = str(time_threads) if time_threads > 0 else "1"
        mode = "merge_time" if time_threads > 0 else ("merge_mass" if by_mass else "merge")
        subprocess.run([REF_TRACKING, mode, fin, fout], check=True,
                       env=dict(os.environ, OMP_NUM_THREADS=nt))
        raw = open(fout, "rb").read()
    if time_threads > 0:
        return float(np.frombuffer(raw, np.float64, 1)[0])
    ns = len(subs)
    out = np.frombuffer(raw, SUB_DTYPE, ns, 0).copy()
    pos = ns * SUB_DTYPE.itemsize
    total = int(np.frombuffer(raw, np.int64, 1, pos)[0])
    pos += 8
    lst = np.frombuffer(raw, np.int64, total, pos).copy()
    pos += 8 * total
    en = np.frombuffer(raw, np.float32, total, pos).copy()
    return out, lst, en


def ref_tree_potential(params, src, tgt, tmass=None):
    sx, sy, sz, sm = (_f32(a) for a in src)
    tx, ty, tz = (_f32(a) for a in tgt)
    tm = _f32(tmass) if tmass is not None else np.zeros(len(tx), np.float32)
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.ascontiguousarray(params, PARAMS_DTYPE).tobytes())
            f.write(np.array([len(sx), len(tx)], np.int64).tobytes())
            for a in (sx, sy, sz, sm, tx, ty, tz, tm):
                f.write(a.tobytes())
        subprocess.run([REF_HARNESS, "tree", fin, fout], check=True)
        return np.fromfile(fout, np.float64)
