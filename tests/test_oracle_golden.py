"""The CPU restatement against the reference's own outputs (golden fixtures
written by tests/golden/make_golden.py from the reference compiled from its
sources).  Bit-identical: records, particle lists (up to the order of exactly
tied energies, which std::sort leaves unspecified), energies, potentials."""
import numpy as np
import pytest

import oracle
from fixtures import (MASK_CASES, MASK_MASS, MERGE_CASES, NEST_CASES, PROPS_CASES, SAMPLING_CASES, TREE_CASES,
                      UNBIND_CASES, load_mask, load_merge, load_merge_mass, load_nest, load_props, load_tree,
                      load_unbind)
from hbtplus_amd.abi import HBT_STATUS_CONVERGED, PROPS_FIELDS
from helpers import assert_exact, assert_merge, check_unbind_invariants


@pytest.mark.parametrize("name", UNBIND_CASES)
def test_oracle_matches_reference_unbinding(name):
    cat, ref = load_unbind(name)
    got = oracle.unbind_batch(cat)
    assert_exact(ref, got)


@pytest.mark.parametrize("name", SAMPLING_CASES)
def test_oracle_matches_reference_sampling(name):
    """Sampling on: the shuffle's rand() draws, the sampled trees with their
    scaled masses, the correction switch's sample bound and
    RefineBindingEnergyOrder, against the reference's own Unbind driven level
    by level from a fresh rand() stream; the same with four threads (each
    level's draws are taken before it is unbound)."""
    cat, ref = load_unbind(name)
    assert cat.params["max_sample_size"][0] > 0
    rs = {"seed": 1, "draws": 0}
    got = oracle.unbind_batch(cat, rand_state=rs)
    assert_exact(ref, got)
    # one draw per element after the first of every unbound list longer than
    # the sample (an orphan unbinds its own list, the others their extended one)
    ms = int(cat.params["max_sample_size"][0])
    rsub = ref[0]
    ulen = np.where(cat.subs["nbound_prev"] <= 1, cat.subs["count"], rsub["list_len"])
    assert rs["draws"] == int(np.sum(np.where(ulen > ms, ulen - 1, 0))) > 0
    assert_exact(got, oracle.unbind_batch(cat, nthreads=4))


@pytest.mark.parametrize("name", UNBIND_CASES)
def test_size_free_invariants_hold_for_reference_outputs(name):
    """The properties the full-size GPU tests check hold for the reference's
    own outputs (converged = nbound > 1: its records carry no status) and for
    the oracle's"""
    cat, ref = load_unbind(name)
    assert check_unbind_invariants(cat, ref, conv=ref[0]["nbound"] > 1) >= 1
    assert check_unbind_invariants(cat, oracle.unbind_batch(cat)) >= 1


def test_size_free_invariants_catch_planted_mistakes():
    cat, _ = load_unbind("c2_small")
    good = oracle.unbind_batch(cat)
    check_unbind_invariants(cat, good)
    s0 = int(np.argmax(good[0]["nbound"]))
    o = int(good[0]["out_offset"][s0])

    def planted(fn):
        subs, idx, en = good[0].copy(), good[1].copy(), good[2].copy()
        fn(subs, idx, en)
        with pytest.raises(AssertionError):
            check_unbind_invariants(cat, (subs, idx, en))

    def swap(subs, idx, en):  # two bound particles out of energy order
        idx[[o + 1, o + 5]] = idx[[o + 5, o + 1]]
        en[[o + 1, o + 5]] = en[[o + 5, o + 1]]

    def foreign(subs, idx, en):  # a particle of another subhalo
        idx[o + 2] = good[1][int(good[0]["out_offset"][(s0 + 1) % len(subs)])]

    def type_count(subs, idx, en):
        subs["nbound_type"][s0, 1] += 1

    def mass(subs, idx, en):
        subs["mbound"][s0] *= 1.001

    def most_bound(subs, idx, en):
        subs["mostbound_pos"][s0, 0] += 1.0

    def unbound_in_block(subs, idx, en):
        en[o + 3] = abs(en[o + 3])

    def too_many(subs, idx, en):
        subs["nbound"][s0] = subs["list_len"][s0] + 1

    for fn in (swap, foreign, type_count, mass, most_bound, unbound_in_block, too_many):
        planted(fn)

    # nested: a host list missing one particle of a child's unbound tail
    cat, _ = load_unbind("c4_small")
    subs, idx, en = oracle.unbind_batch(cat)
    check_unbind_invariants(cat, (subs, idx, en))
    k = int(np.nonzero(cat.subs["parent"] >= 0)[0][0])
    h = int(cat.subs["parent"][k])
    tail = idx[int(subs["out_offset"][k] + subs["nbound"][k])]
    o, L = int(subs["out_offset"][h]), int(subs["list_len"][h])
    j = o + int(np.nonzero(idx[o:o + L] == tail)[0][0])
    idx2 = idx.copy()
    idx2[j] = idx[int(subs["out_offset"][k])]  # a bound particle of the child instead
    with pytest.raises(AssertionError):
        check_unbind_invariants(cat, (subs, idx2, en))


@pytest.mark.parametrize("name", TREE_CASES)
def test_oracle_matches_reference_tree(name):
    params, (x, y, z, m), ref = load_tree(name)
    got = oracle.tree_potential(params, (x, y, z, m), (x, y, z), m)
    assert np.array_equal(got, ref)


def test_fixture_outcomes_are_nontrivial():
    """The fixtures exercise convergence, disruption, early exits, nesting,
    orphans and correction passes."""
    seen = set()
    for name in UNBIND_CASES:
        cat, (rs, ri, re_) = load_unbind(name)
        got = oracle.unbind_batch(cat)[0]
        seen |= {int(s) for s in got["status"]}
        if (got["n_iter"] > 2).any():
            seen.add("multi-pass")
        if (cat.subs["parent"] >= 0).any():
            seen.add("nested")
        if ((cat.subs["nbound_prev"] <= 1) & (rs["list_len"] > cat.subs["count"])).any():
            seen.add("orphan-with-tail")
    assert {1, 2, 4, "multi-pass", "nested", "orphan-with-tail"} <= seen


@pytest.mark.parametrize("name", PROPS_CASES)
def test_oracle_properties_match_reference(name):
    """CalculateProfileProperties + CalculateShape restated: bit-identical to
    the reference's on its own unbinding outputs"""
    cat, (subs, idx, _), pin, ref = load_props(name)
    got = oracle.subhalo_properties(cat.params, cat.cols, subs, idx, pin)
    for f in PROPS_FIELDS:
        assert np.array_equal(np.ascontiguousarray(got[f]).view(np.uint8),
                              np.ascontiguousarray(ref[f]).view(np.uint8)), f


@pytest.mark.parametrize("name", MASK_CASES)
def test_oracle_mask_matches_reference(name):
    """MaskSubhalos restated: the kept lists of the reference's own
    SubhaloSnapshot_t::MaskSubhalos, entry for entry"""
    cols, subs, grp, ref_out, ref_cnt = load_mask(name)
    out, cnt = oracle.mask_subhalos(cols, subs, grp)
    assert np.array_equal(cnt, ref_cnt)
    for o, c in zip(subs["offset"], cnt):
        assert np.array_equal(out[o:o + c], ref_out[o:o + c])


def _same_kept(out, cnt, ref_out, ref_cnt, subs):
    return np.array_equal(cnt, ref_cnt) and all(
        np.array_equal(out[o:o + c], ref_out[o:o + c]) for o, c in zip(subs["offset"], cnt))


def test_oracle_mask_in_mass_order():
    """The reference's member table ordered by the subhalos' own Mbound (the
    central a head other than the group's first in index order): the
    restatement given that order as visit ranks keeps what the reference
    keeps, and in index order it would not"""
    cols, subs, grp, ref_out, ref_cnt = load_mask(MASK_MASS)
    rank = oracle.mass_visit_rank(subs, grp)
    assert _same_kept(*oracle.mask_subhalos(cols, subs, grp, visit_rank=rank), ref_out, ref_cnt, subs)
    assert not _same_kept(*oracle.mask_subhalos(cols, subs, grp), ref_out, ref_cnt, subs)


def test_oracle_merge_in_mass_order():
    params, cols, lst, subs, grp, ref = load_merge_mass()
    rank = oracle.mass_visit_rank(subs, grp)
    got = oracle.merge_subhalos(params, cols, lst, subs, grp, True, visit_rank=rank)
    assert_merge(ref, (got[0], got[1]), ties=True)
    plain = oracle.merge_subhalos(params, cols, lst, subs, grp, True)
    with pytest.raises(AssertionError):
        assert_merge(ref, (plain[0], plain[1]), ties=True)


@pytest.mark.parametrize("name", MERGE_CASES)
@pytest.mark.parametrize("merge_trapped", [True, False])
def test_oracle_merge_matches_reference(name, merge_trapped):
    """MergeSubhalos restated (cores, trap detection, MergeRecursive/MergeTo,
    re-unbinding, TruncateSource): records and lists of the reference's own
    SubhaloSnapshot_t::MergeSubhalos, bit for bit and entry for entry"""
    F = load_merge(name)
    got = oracle.merge_subhalos(F.params, F.cols, F.list_index, F.subs, F.host_group, merge_trapped)
    ref = F.out if merge_trapped else F.off
    assert_merge(ref, got[:2])
    if merge_trapped:  # binding energies of the bound blocks of the re-unbound subhalos
        rs = ref[0]
        for s in np.nonzero(got[0]["status"] & HBT_STATUS_CONVERGED)[0]:
            o, nb = int(rs["offset"][s]), int(rs["nbound"][s])
            assert np.array_equal(ref[2][o:o + nb], got[2][o:o + nb]), s


@pytest.mark.parametrize("name", NEST_CASES)
def test_oracle_nest_matches_reference(name):
    """NestSubhalos' level-up restated: the hosts and Ranks the reference's
    own LevelUpDetachedSubhalos leaves"""
    g, c, p, r, rp, rr = load_nest(name)
    op, orr = oracle.nest_subhalos(g, c, p, r)
    assert np.array_equal(op, rp) and np.array_equal(orr, rr)


@pytest.mark.parametrize("name", MERGE_CASES)
def test_oracle_merge_cores_match_reference(name):
    """The merging cores restated (mass-weighted position about the first
    particle with the periodic origin in double, velocity, dispersions): the
    reference's own SubHelper_t::BuildPosition / BuildVelocity bit for bit"""
    F = load_merge(name)
    got = oracle.merge_cores(F.params, F.cols, F.list_index, F.subs)
    ok = F.subs["nbound"] > 0  # the reference leaves the mean unset for empty lists
    assert np.array_equal(got[ok].view(np.uint32), F.cores[ok].view(np.uint32))
    assert np.array_equal(got[~ok][:, 6:], F.cores[~ok][:, 6:])
