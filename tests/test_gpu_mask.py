"""GPU parity of hbt_mask_subhalos (MaskSubhalos, src/subhalo_tracking.cpp:
793-841) through the C-ABI against the oracle: identical kept lists (as
particle column positions) and counts."""
import numpy as np
import pytest

import oracle
from fixtures import MASK_CASES, MASK_MASS, load_mask
from hbtplus_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,ngroups,big", [(11, 30, 0), (12, 200, 0), (13, 10, 6000), (14, 3, 150000)])
def test_mask_matches_oracle(gpu, seed, ngroups, big):
    cols, subs, grp = synth.masking_catalogue(seed=seed, ngroups=ngroups, big_group=big)
    ref_out, ref_cnt = oracle.mask_subhalos(cols, subs, grp)
    out, cnt = _run(gpu, cols, subs, grp)
    assert np.array_equal(cnt, ref_cnt)
    for s in range(len(subs)):
        o = subs["offset"][s]
        assert np.array_equal(out[o:o + cnt[s]], ref_out[o:o + cnt[s]]), s
    print(f"seed {seed}: {len(subs)} subhalos, {len(cols['id'])} entries, kept {cnt.sum()}")


def _run(gpu, cols, subs, grp):
    from hbtplus_amd.abi import make_params
    return gpu(make_params()).mask_subhalos(cols, subs, grp)


def test_mask_rejects_cross_group_nesting(gpu):
    cols, subs, grp = synth.masking_catalogue(seed=15, ngroups=4)
    child = int(np.nonzero(subs["parent"] >= 0)[0][0])
    grp = grp.copy()
    grp[child] = grp[subs["parent"][child]] + 1000
    with pytest.raises(RuntimeError, match="host group"):
        _run(gpu, cols, subs, grp)


@pytest.mark.parametrize("name", MASK_CASES)
def test_mask_matches_reference(gpu, name):
    """against the reference's own MaskSubhalos (golden fixtures)"""
    cols, subs, grp, ref_out, ref_cnt = load_mask(name)
    out, cnt = _run(gpu, cols, subs, grp)
    assert np.array_equal(cnt, ref_cnt)
    for o, c in zip(subs["offset"], cnt):
        assert np.array_equal(out[o:o + c], ref_out[o:o + c])


def test_mask_in_mass_order_matches_reference(gpu):
    """the member table in the reference's Mbound order, given as visit ranks"""
    cols, subs, grp, ref_out, ref_cnt = load_mask(MASK_MASS)
    from hbtplus_amd.abi import make_params
    out, cnt = gpu(make_params()).mask_subhalos(cols, subs, grp, visit_rank=oracle.mass_visit_rank(subs, grp))
    assert np.array_equal(cnt, ref_cnt)
    for o, c in zip(subs["offset"], cnt):
        assert np.array_equal(out[o:o + c], ref_out[o:o + c])
