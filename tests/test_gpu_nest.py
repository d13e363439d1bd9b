"""GPU parity of hbt_nest_subhalos (NestSubhalos' LevelUpDetachedSubhalos,
src/subhalo_tracking.cpp:669-688, 735-786) through the C-ABI: hosts and Ranks
identical to the reference's on the golden fixtures and to the oracle on
larger random nests."""
import numpy as np
import pytest

import oracle
from fixtures import NEST_CASES, load_nest
from hbtplus_amd import abi, synth
from test_oracle_math import NEST_HAND

pytestmark = pytest.mark.gpu


def _gpu(gpu):
    return gpu(abi.make_params())


@pytest.mark.parametrize("name", NEST_CASES)
def test_nest_matches_reference(gpu, name):
    g, c, p, r, rp, rr = load_nest(name)
    op, orr = _gpu(gpu).nest_subhalos(g, c, p, r)
    assert np.array_equal(op, rp) and np.array_equal(orr, rr)


def test_nest_hand_derived(gpu):
    a = np.array(NEST_HAND, np.int64)
    op, orr = _gpu(gpu).nest_subhalos(a[:, 0], a[:, 1].astype(np.int32), a[:, 2], a[:, 3].astype(np.int32))
    assert op.tolist() == a[:, 4].tolist() and orr.tolist() == a[:, 5].tolist()


@pytest.mark.parametrize("seed,ngroups,nsub", [(21, 500, 20000), (22, 20000, 200000)])
def test_nest_matches_oracle(gpu, seed, ngroups, nsub):
    g, c, p, r = synth.nesting_catalogue(seed=seed, ngroups=ngroups, nsub=nsub, nfield=nsub // 20)
    rp, rr = oracle.nest_subhalos(g, c, p, r)
    op, orr = _gpu(gpu).nest_subhalos(g, c, p, r)
    assert np.array_equal(op, rp) and np.array_equal(orr, rr)


def test_nest_rejects_cycles_and_bad_hosts(gpu):
    g, c, p, r = synth.nesting_catalogue(seed=23, ngroups=10, nsub=100, nfield=4)
    U = _gpu(gpu)
    bad = p.copy()
    child = int(np.nonzero(p >= 0)[0][0])
    bad[p[child]] = child  # host <-> member: a cycle
    with pytest.raises(RuntimeError, match="chain"):
        U.nest_subhalos(g, c, bad, r)
    bad = p.copy()
    bad[child] = len(p) + 5
    with pytest.raises(RuntimeError, match="host outside"):
        U.nest_subhalos(g, c, bad, r)
