"""Comparison helpers shared by the parity tests."""
from __future__ import annotations

import numpy as np

from hbtplus_amd.abi import HBT_STATUS_CONVERGED, OUTPUT_FIELDS


def as_ids(ids, idx):
    """Particle IDs of output positions; slots outside every list (the
    unused capacity of a nested host's output region holds no position) map
    to -1."""
    ok = (idx >= 0) & (idx < len(ids))
    return np.where(ok, ids[np.where(ok, idx, 0)].astype(np.int64), -1)


def unused_slots(subs, total):
    """Mask of the output slots outside every subhalo's list."""
    L = subs["list_len"].astype(np.int64)
    pos = np.repeat(subs["out_offset"], L) + np.arange(L.sum()) - np.repeat(np.cumsum(L) - L, L)
    free = np.ones(total, bool)
    free[pos] = False
    return free


def list_of(subs, out, s):
    o, L = int(subs["out_offset"][s]), int(subs["list_len"][s])
    return out[o:o + L]


def canonical_order(idx, en):
    """Index list with runs of exactly equal energy put in index order
    (the reference's std::sort leaves the order of equal keys unspecified)."""
    idx = idx.copy()
    if len(idx) < 2:
        return idx
    e = np.nan_to_num(en, nan=np.inf)
    start = 0
    for i in range(1, len(idx) + 1):
        if i == len(idx) or e[i] != e[start]:
            if i - start > 1:
                idx[start:i] = np.sort(idx[start:i])
            start = i
    return idx


def assert_exact(ref, got, fields=OUTPUT_FIELDS):
    """Bit-identical record fields, and identical particle lists up to the
    order of exactly tied energies."""
    rs, ri, re_ = ref
    gs, gi, ge = got
    for f in fields:
        a, b = rs[f], gs[f]
        if not np.array_equal(a, b):
            bad = np.nonzero(~np.all((a == b).reshape(len(a), -1), axis=1))[0]
            raise AssertionError(f"field {f} differs for subhalos {bad[:10]}: "
                                 f"ref {a[bad[:3]]} got {b[bad[:3]]}")
    for s in range(len(rs)):
        a_i, b_i = list_of(rs, ri, s), list_of(gs, gi, s)
        a_e, b_e = list_of(rs, re_, s), list_of(gs, ge, s)
        # an orphan (nbound_prev <= 1) is unbound on a copy of its own list and
        # keeps its extended list in the original order, so the reference's
        # Energies of that copy do not describe its list (the library and the
        # oracle report NaN)
        if gs["status"][s] & HBT_STATUS_CONVERGED and gs["nbound_prev"][s] > 1:
            # energies per particle must agree bit for bit; the reference side
            # may store only the bound block (SAVE_BINDING_ENERGY)
            nb = int(gs["nbound"][s])
            e_of = dict(zip(b_i.tolist(), b_e.tolist()))
            a_e2 = np.array([e_of.get(p, np.nan) for p in a_i.tolist()], np.float32)
            known = np.isfinite(a_e)
            known[nb:] &= np.isfinite(a_e2[nb:])
            assert np.array_equal(a_e[:nb], a_e2[:nb]), f"bound energies differ in subhalo {s}"
            assert np.array_equal(a_e[known], a_e2[known]), f"energies differ in subhalo {s}"
            a_i, b_i = canonical_order(a_i, a_e2), canonical_order(b_i, b_e)
        assert np.array_equal(a_i, b_i), f"particle list differs in subhalo {s}"


def assert_tolerant(ref, got, band_rel=1e-5, refined=0):
    """FAST-mode parity: bound sets identical except particles whose oracle
    binding energy lies within band_rel*|E_min| of zero; bound counts differ
    by at most the number of such particles; most-bound IDs agree unless the
    two most bound oracle energies are within the band; the bound block is in
    oracle-energy order up to the band.  refined > 0: a bound block longer
    than that (RefineBindingEnergyOrder with MaxSampleSize = refined) has its
    first `refined` entries re-ordered by another energy: they must be the
    same set, the most-bound IDs must agree, and the order is checked after
    them."""
    rs, ri, re_ = ref
    gs, gi, ge = got
    stats = dict(n=0, band=0, set_diff=0, mb_mismatch=0)
    for s in range(len(rs)):
        a_i, b_i = list_of(rs, ri, s), list_of(gs, gi, s)
        a_e = list_of(rs, re_, s)
        na, nb = int(rs["nbound"][s]), int(gs["nbound"][s])
        stats["n"] += 1
        if not (gs["status"][s] & HBT_STATUS_CONVERGED):
            # disrupted or early exit: one (or no) particle, the list keeps its order
            assert na == nb and np.array_equal(a_i, b_i), f"subhalo {s}: non-converged outcome differs"
            continue
        if na <= 1:
            raise AssertionError(f"subhalo {s}: converged on the GPU, disrupted in the reference")
        emin = np.nanmin(a_e)
        band = band_rel * abs(emin)
        e_of = dict(zip(a_i.tolist(), a_e.tolist()))
        near = {p for p, e in e_of.items() if abs(e) < band}
        A, B = set(a_i[:na].tolist()), set(b_i[:nb].tolist())
        diff = A ^ B
        stats["band"] += len(near)
        stats["set_diff"] += len(diff)
        assert diff <= near, f"subhalo {s}: {len(diff - near)} bound-set differences outside the band"
        assert abs(na - nb) <= len(near), f"subhalo {s}: Nbound {na} vs {nb}"
        skip = refined if (refined and na > refined) else 0
        if skip:
            assert set(a_i[:skip].tolist()) ^ set(b_i[:skip].tolist()) <= near, \
                f"subhalo {s}: refined prefix differs"
            assert a_i[0] == b_i[0], f"subhalo {s}: most-bound ID after the refinement differs"
        elif a_i[0] != b_i[0]:
            top = np.sort(a_e[:na])[:2]
            assert len(top) > 1 and top[1] - top[0] <= band, f"subhalo {s}: most-bound ID differs"
            stats["mb_mismatch"] += 1
        eg = np.array([e_of.get(p, np.nan) for p in b_i[skip:nb].tolist()])
        eg = eg[~np.isnan(eg)]
        if len(eg):
            run = np.maximum.accumulate(eg)
            assert np.all(run - eg <= band), f"subhalo {s}: bound order differs beyond the band"
    return stats


MERGE_FIELDS = ["count", "nbound", "mbound", "nbound_type", "mbound_type", "sink_track_id",
                "snapshot_index_of_death", "snapshot_index_of_sink", "ref_pos", "ref_vel", "mostbound_pos",
                "mostbound_vel", "specific_self_potential_energy", "specific_self_kinetic_energy",
                "specific_angular_momentum"]


def merge_lists(subs, lst):
    return [lst[int(o):int(o) + int(c)] for o, c in zip(subs["offset"], subs["count"])]


def assert_merge(ref, got, ties=False):
    """MergeSubhalos outputs: record fields bit for bit and every list entry
    for entry.  ref = (records, lists, energies or None), got = (records,
    lists); with ties, runs of exactly equal reference energy may come in any
    order (the reference's std::sort leaves them unspecified)."""
    rs, rl = ref[0], ref[1]
    re_ = ref[2] if len(ref) > 2 else None
    gs, gl = got
    for f in MERGE_FIELDS:
        a = np.ascontiguousarray(rs[f]).view(np.uint8).reshape(len(rs), -1)
        b = np.ascontiguousarray(gs[f]).view(np.uint8).reshape(len(gs), -1)
        bad = np.nonzero(~np.all(a == b, axis=1))[0]
        assert len(bad) == 0, f"field {f} differs for subhalos {bad[:10]}: ref {rs[f][bad[:3]]} got {gs[f][bad[:3]]}"
    ra, ga = merge_lists(rs, rl), merge_lists(gs, gl)
    ea = merge_lists(rs, re_) if re_ is not None else None
    for s in range(len(rs)):
        a, b = ra[s], ga[s]
        if ties and ea is not None and np.isfinite(ea[s]).any():
            e_of = dict(zip(a.tolist(), ea[s].tolist()))
            eb = np.array([e_of.get(p, np.nan) for p in b.tolist()], np.float32)
            a, b = canonical_order(a, ea[s]), canonical_order(b, eb)
        assert np.array_equal(a, b), f"list of subhalo {s} differs"


def check_unbind_invariants(cat, out, conv=None, refined=0):
    """Properties of an unbinding result that hold at any size, checked with
    array operations (no per-subhalo loop): every output list is a
    permutation of its subhalo's source particles plus the unbound tails of
    its children's output lists (RecursiveUnbind hands a child's unbound
    particles to its host); nbound <= list_len; a converged subhalo's bound
    block has negative, non-decreasing energies (most bound first) and its
    first particle is the most-bound particle of the record; the per-type
    bound counts add up to the bound particles of each type; mbound is their
    mass sum.  conv: which subhalos converged (default: the records'
    status).  refined > 0: the first `refined` entries of a bound block
    longer than that are re-ordered by RefineBindingEnergyOrder; the energy
    order is checked from there on.  Returns their number."""
    subs, idx, en = out
    src = cat.subs
    n = len(subs)
    L = subs["list_len"].astype(np.int64)
    nb = subs["nbound"].astype(np.int64)
    assert np.all((nb >= 0) & (nb <= L)), "nbound outside [0, list_len]"
    seg = np.repeat(np.arange(n), L)
    pos = np.arange(len(seg)) - np.repeat(np.cumsum(L) - L, L)
    rows = idx[np.repeat(subs["out_offset"], L) + pos]
    # expected multiset per subhalo: own rows, then each child's tail
    cnt = src["count"].astype(np.int64)
    own_seg = np.repeat(np.arange(n), cnt)
    own_rows = np.repeat(src["offset"], cnt) + np.arange(len(own_seg)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    kids = np.nonzero(src["parent"] >= 0)[0]
    tl = (L - nb)[kids]
    t_seg = np.repeat(src["parent"][kids], tl)
    t_pos = np.arange(len(t_seg)) - np.repeat(np.cumsum(tl) - tl, tl)
    t_rows = idx[np.repeat(subs["out_offset"][kids] + nb[kids], tl) + t_pos]
    e_seg = np.concatenate([own_seg, t_seg])
    e_rows = np.concatenate([own_rows, t_rows])
    assert len(e_seg) == len(seg), "list lengths differ from the source counts plus the children's tails"
    a, b = np.lexsort((rows, seg)), np.lexsort((e_rows, e_seg))
    assert np.array_equal(seg[a], e_seg[b]) and np.array_equal(rows[a], e_rows[b]), "list is not a permutation"
    conv = (subs["status"] & HBT_STATUS_CONVERGED) != 0 if conv is None else np.asarray(conv, bool)
    bound = conv[seg] & (pos < nb[seg])
    e = en[np.repeat(subs["out_offset"], L) + pos]
    assert np.all(e[bound] < 0), "a bound particle has E >= 0"
    step = bound[1:] & bound[:-1] & (seg[1:] == seg[:-1])
    if refined:
        step &= ~((nb[seg[1:]] > refined) & (pos[1:] < refined))
    assert np.all(e[1:][step] >= e[:-1][step]), "bound block is not in ascending energy"
    first = subs["out_offset"][conv]
    for k, c in enumerate("xyz"):
        assert np.array_equal(subs["mostbound_pos"][conv, k], cat.cols[c][idx[first]]), "most-bound position"
    typ = cat.cols["type"][rows].astype(np.int64)
    tmax = subs["nbound_type"].shape[1]
    cnt = np.bincount(seg[bound] * tmax + typ[bound], minlength=n * tmax).reshape(n, tmax)
    assert np.array_equal(cnt[conv], subs["nbound_type"][conv]), "per-type bound counts"
    assert np.array_equal(cnt[conv].sum(1), nb[conv]), "per-type counts do not add up to nbound"
    mass = np.bincount(seg[bound], weights=cat.cols["mass"][rows][bound].astype(np.float64), minlength=n)
    assert np.allclose(subs["mbound"][conv], mass[conv], rtol=1e-5, atol=0), "mbound differs from the mass sum"
    return int(conv.sum())
