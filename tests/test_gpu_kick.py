"""Null particles in the source lists (IDs the snapshot does not hold, which
the ID resolution marks -1): hbt_unbind_batch_indexed drops them as
Subhalo_t::KickNullParticles does (src/subhalo.cpp:446-477, called after the
particle query, src/particle_exchanger.h:239-240) -- each list keeps its other
entries in order and the Nbound the orphan rule reads counts the kept entries
among its first Nbound -- and then unbinds.

Against the reference's own KickNullParticles + unbinding (fixture kick_c2:
random nulls, a list whose first entries are null so that it becomes an
orphan, an all-null list, a nest): REPLAY bit for bit, FAST by the band rule.
On 5,000 subhalos of the one-rank share law with 3% unknown IDs resolved on the device:
REPLAY equals the oracle's KickNullParticles restatement followed by its
unbinding, and the output slots the shorter lists leave free hold -1."""
import numpy as np
import pytest

import oracle
from fixtures import load_unbind
from hbtplus_amd import abi, synth
from helpers import assert_exact, assert_tolerant, unused_slots

pytestmark = pytest.mark.gpu


def _with(params, **kw):
    p = params.copy()
    for k, v in kw.items():
        p[k] = v
    return p


def _fixture_positions(cat):
    """the fixture's lists as positions in its columns, -1 for null rows"""
    cnt = cat.subs["count"].astype(np.int64)
    rows = np.repeat(cat.subs["offset"], cnt) + np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    return np.where(cat.cols["id"][rows] == -1, -1, rows), rows


def _as_positions(cat):
    """records whose offset/count index the concatenated position list"""
    subs = cat.subs.copy()
    subs["offset"] = np.concatenate([[0], np.cumsum(subs["count"])[:-1]])
    return subs


@pytest.mark.parametrize("prec", [abi.HBT_PREC_REPLAY, abi.HBT_PREC_FAST])
def test_kick_matches_reference(gpu, prec):
    cat, ref = load_unbind("kick_c2")
    pos, _ = _fixture_positions(cat)
    assert (pos < 0).sum() > 0
    U = gpu(_with(cat.params, precision=prec))
    gs, gi, ge = U.unbind_positions(cat.cols, pos, _as_positions(cat))
    # the records keep their inputs; the orphan rule read the kicked Nbound
    kicked, _ = oracle.kick_null_particles(cat)
    gs["nbound_prev"] = kicked.subs["nbound_prev"]
    rs = ref[0].copy()
    rs["out_offset"] = gs["out_offset"]
    ro = np.full(len(gi), -1, np.int64)
    reo = np.full(len(ge), np.nan, np.float32)
    for s in range(len(rs)):
        a, L = int(ref[0]["out_offset"][s]), int(ref[0]["list_len"][s])
        b = int(gs["out_offset"][s])
        ro[b:b + L] = ref[1][a:a + L]
        reo[b:b + L] = ref[2][a:a + L]
    if prec == abi.HBT_PREC_REPLAY:
        assert_exact((rs, ro, reo), (gs, gi, ge))
    else:
        print("fast:", assert_tolerant((rs, ro, reo), (gs, gi, ge)))
    assert np.all(gi[unused_slots(gs, len(gi))] == -1)


def test_kick_unknown_ids_in_a_share(gpu):
    snap, ids, cat = synth.config5(seed=1016, nsub=5000)
    rng = np.random.default_rng(16)
    unknown = rng.random(len(ids)) < 0.03
    q = ids.copy()
    q[unknown] = np.int64(2) ** 62 + np.arange(int(unknown.sum()))  # not in the snapshot
    U = gpu(_with(cat.params, precision=abi.HBT_PREC_REPLAY))
    h = U.upload_snapshot(snap, q, cat.subs)
    idx = U.resolve_ids(h)
    got_pos = idx.cpu().numpy()
    assert np.array_equal(got_pos < 0, unknown)
    s1, i1, e1 = U.unbind_indexed(h, idx)
    gs = s1.cpu().numpy().view(abi.SUB_DTYPE).copy()
    total = h["total"]
    gi, ge = i1.cpu().numpy()[:total], e1.cpu().numpy()[:total]
    # the oracle: the catalogue's rows with the unknown entries null, kicked, unbound
    ck = synth.Catalogue({k: v.copy() for k, v in cat.cols.items()}, cat.subs.copy(), cat.params.copy())
    ck.cols["id"][unknown] = -1
    kicked, kept = oracle.kick_null_particles(ck)
    rs, ri, re_ = oracle.unbind_batch(kicked, nthreads=16)
    where = np.argsort(snap["id"])
    row_pos = where[np.searchsorted(snap["id"][where], ids)]  # snapshot position of each catalogue row
    gs["nbound_prev"] = kicked.subs["nbound_prev"]
    rs["out_offset"] = gs["out_offset"]
    ro = np.full(total, -1, np.int64)
    reo = np.full(total, np.nan, np.float32)
    plan = kicked.subs.copy()
    abi.plan_batch(plan)
    for s in np.nonzero(rs["list_len"] > 0)[0]:
        a, L, b = int(plan["out_offset"][s]), int(rs["list_len"][s]), int(gs["out_offset"][s])
        ro[b:b + L] = row_pos[kept[ri[a:a + L]]]
        reo[b:b + L] = re_[a:a + L]
    assert_exact((rs, ro, reo), (gs, gi, ge))
    assert np.all(gi[unused_slots(gs, total)] == -1)
    print(f"{int(unknown.sum())} unknown IDs, {int((kicked.subs['nbound_prev'] <= 1).sum())} orphans after the kick")
