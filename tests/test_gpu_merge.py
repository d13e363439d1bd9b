"""GPU parity of hbt_merge_subhalos (SubhaloSnapshot_t::MergeSubhalos,
src/subhalo_merge.cpp:136-262) through the C-ABI.

Against the reference's own MergeSubhalos on the golden fixtures (REPLAY
re-unbinding, whose lists and records are the reference's bit for bit), on
the hand-derived configuration, and against the oracle on the state the
GPU's own unbinding leaves for larger catalogues."""
import ctypes

import numpy as np
import pytest

import oracle
from fixtures import MERGE_CASES, load_merge, load_merge_mass
from hbtplus_amd import abi, synth
from hbtplus_amd.abi import HBT_ERR_CAPACITY, SUB_DTYPE, particles_struct
from helpers import assert_merge
from merge_cases import check_hand_merge, hand_merge_case

pytestmark = pytest.mark.gpu


def _with(params, **kw):
    p = params.copy()
    for k, v in kw.items():
        p[k] = v
    return p


@pytest.mark.parametrize("name", MERGE_CASES)
@pytest.mark.parametrize("merge_trapped", [True, False])
def test_merge_matches_reference(gpu, name, merge_trapped):
    F = load_merge(name)
    U = gpu(_with(F.params, precision=abi.HBT_PREC_REPLAY))
    got = U.merge_subhalos(F.cols, F.list_index, F.subs, F.host_group, merge_trapped)
    assert_merge(F.out if merge_trapped else F.off, got, ties=True)


def test_merge_in_mass_order_matches_reference(gpu):
    """the member table in the reference's Mbound order, given as visit ranks
    (REPLAY re-unbinding)"""
    params, cols, lst, subs, grp, ref = load_merge_mass()
    U = gpu(_with(params, precision=abi.HBT_PREC_REPLAY))
    got = U.merge_subhalos(cols, lst, subs, grp, True, visit_rank=oracle.mass_visit_rank(subs, grp))
    assert_merge(ref, got, ties=True)


@pytest.mark.parametrize("precision", [abi.HBT_PREC_REPLAY, abi.HBT_PREC_FAST])
@pytest.mark.parametrize("merge_trapped", [True, False])
def test_merge_hand_derived(gpu, precision, merge_trapped):
    params, cols, lst, subs, grp = hand_merge_case()
    U = gpu(_with(params, precision=precision))
    out, ol = U.merge_subhalos(cols, lst, subs, grp, merge_trapped)
    check_hand_merge(subs, out, ol, merge_trapped)


@pytest.mark.parametrize("seed,ngroups,scale", [(2001, 12, 1.0), (2002, 40, 2.0)])
def test_merge_after_gpu_unbinding_matches_oracle(gpu, seed, ngroups, scale):
    """The GPU's FAST unbinding produces the state; merging it on the GPU
    (REPLAY re-unbinding) and in the oracle gives the same records and lists"""
    cat, grp = synth.merging_catalogue(seed=seed, ngroups=ngroups, scale=scale)
    U = gpu(cat.params)
    subs, idx, _ = U.unbind_batch(cat)
    lst, sin = oracle.merge_input(cat, subs, idx)
    U.set_params(_with(cat.params, precision=abi.HBT_PREC_REPLAY))
    got = U.merge_subhalos(cat.cols, lst, sin, grp)
    ref = oracle.merge_subhalos(_with(cat.params, precision=abi.HBT_PREC_REPLAY), cat.cols, lst, sin, grp)
    assert_merge(ref, got, ties=True)
    merged = int(((got[0]["count"] == 1) & (sin["count"] > 1)).sum())
    assert merged >= ngroups, merged
    print(f"seed {seed}: {len(sin)} subhalos, {cat.n_particles} particles, {merged} merged away")


def test_merge_fast_keeps_decisions(gpu):
    """FAST re-unbinding: trap decisions, deaths and the merged-away lists are
    exactly the oracle's (they do not depend on the precision); the
    re-unbound hosts keep bound counts within 0.5% of the oracle's"""
    F = load_merge("merge_a")
    U = gpu(_with(F.params, precision=abi.HBT_PREC_FAST))
    gs, gl = U.merge_subhalos(F.cols, F.list_index, F.subs, F.host_group)
    rs, rl = F.out[0], F.out[1]
    for f in ("sink_track_id", "snapshot_index_of_sink", "snapshot_index_of_death"):
        assert np.array_equal(gs[f], rs[f]), f
    away = (rs["count"] == 1) & (F.subs["count"] > 1)
    assert np.array_equal(gl[gs["offset"][away]], rl[rs["offset"][away]])
    big = rs["nbound"] > 1
    assert np.all(np.abs(gs["nbound"][big] - rs["nbound"][big]) <= np.maximum(2, 0.005 * rs["nbound"][big]))


def _merge_raw(U, cols, lst, subs, grp, cap):
    t = U.torch
    c = {k: U._dev(np.ascontiguousarray(cols[k])) for k in abi.PARTICLE_COLUMNS}
    ps = particles_struct(c, ptr=lambda x: x.data_ptr())
    li = U._dev(np.ascontiguousarray(lst, np.int64))
    subs = np.ascontiguousarray(subs, SUB_DTYPE).copy()
    grp = np.ascontiguousarray(grp, np.int64)
    out = t.full((max(cap, 1),), -1, dtype=t.int64, device="cuda:0")
    used = ctypes.c_int64()
    t.cuda.synchronize()
    rc = U.lib.hbt_merge_subhalos(U.ctx, ctypes.byref(ps), li.data_ptr(), len(lst), subs.ctypes.data,
                                  grp.ctypes.data, None, len(subs), 1, out.data_ptr(), cap, ctypes.byref(used))
    return rc, used.value


def test_merge_capacity_and_rejections(gpu):
    params, cols, lst, subs, grp = hand_merge_case()
    U = gpu(params)
    rc, need = _merge_raw(U, cols, lst, subs, grp, 3)
    assert rc == HBT_ERR_CAPACITY and need > 3
    rc, used = _merge_raw(U, cols, lst, subs, grp, need)
    assert rc == 0 and used == need
    bad = subs.copy()  # trapped, alive and bound: unbinding never leaves one
    bad["sink_track_id"][1] = 0
    with pytest.raises(RuntimeError, match="trapped"):
        U.merge_subhalos(cols, lst, bad, grp)
    g2 = grp.copy()
    g2[1] = 5  # nested in a subhalo of another group
    with pytest.raises(RuntimeError, match="host group"):
        U.merge_subhalos(cols, lst, subs, g2)
    cyc = subs.copy()
    cyc["parent"][0] = 2  # H -> B -> H
    with pytest.raises(RuntimeError, match="cycle|host group"):
        U.merge_subhalos(cols, lst, cyc, grp)
