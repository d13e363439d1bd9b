"""Generators are deterministic, give unique IDs and the configured shapes."""
import numpy as np

from hbtplus_amd import synth


def test_deterministic_and_unique():
    a = synth.config2(seed=5, nsub=50, hi=2000)
    b = synth.config2(seed=5, nsub=50, hi=2000)
    for k in a.cols:
        assert np.array_equal(a.cols[k], b.cols[k])
    assert len(np.unique(a.cols["id"])) == a.n_particles
    assert (a.cols["id"] >= 0).all()


def test_config_shapes():
    c1 = synth.config1()
    assert c1.n_particles == 24000 and len(c1.subs) == 1
    c2 = synth.config2(nsub=200)
    assert len(c2.subs) == 200
    n = c2.subs["count"]
    assert n.min() >= 20 and (np.abs(c2.cols["x"]) <= 1e5).all()
    c4 = synth.config4(n_host=2000, n_sat=4, sat_lo=100, sat_hi=300)
    assert c4.subs["parent"][0] == -1 and (c4.subs["parent"][1:] == 0).all()


def test_power_law_mean():
    rng = np.random.Generator(np.random.Philox(key=1))
    n = synth.power_law_counts(rng, 200000, 20, 2e5)
    assert 230 < n.mean() < 320  # analytic mean ~272
