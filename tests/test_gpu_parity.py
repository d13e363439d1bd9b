"""GPU parity: the HIP library through its C-ABI against the CPU oracle
(which is itself pinned bit-for-bit to the reference, see test_oracle_golden).

REPLAY precision must be bit-identical to the oracle (particle lists up to the
order of exactly tied energies); FAST precision must meet the tolerant
criteria (bound sets identical outside the 1e-5*|E_min| band around E=0,
potentials within 1e-4 relative RMS of the reference tree)."""
import numpy as np
import pytest

import oracle
from hbtplus_amd import abi, synth
from helpers import assert_exact, assert_tolerant

pytestmark = pytest.mark.gpu


def _params(cat, **kw):
    p = cat.params.copy()
    for k, v in kw.items():
        p[k] = v
    return p


def _cols(cat):
    c = cat.cols
    return (c["x"], c["y"], c["z"], c["mass"]), (c["x"], c["y"], c["z"])


def test_tree_potential_replay_bit_exact(gpu):
    cat = synth.config1(n_nfw=3000, n_cont=600)
    p = _params(cat, precision=abi.HBT_PREC_REPLAY)
    src, tgt = _cols(cat)
    ref = oracle.tree_potential(p, src, tgt, cat.cols["mass"])
    got = gpu(p).tree_potential(src, tgt, cat.cols["mass"])
    nbad = int(np.sum(ref != got))
    print(f"replay tree potential: {nbad} of {len(ref)} differ, "
          f"max rel {np.max(np.abs(got / ref - 1)):.3e}")
    assert nbad == 0


def test_tree_potential_fast_rms(gpu):
    cat = synth.config1(n_nfw=3000, n_cont=600)
    p = _params(cat, precision=abi.HBT_PREC_FAST)
    src, tgt = _cols(cat)
    ref = oracle.tree_potential(p, src, tgt, cat.cols["mass"])
    got = gpu(p).tree_potential(src, tgt, cat.cols["mass"])
    rms = np.sqrt(np.mean(((got - ref) / ref) ** 2))
    print(f"fast tree potential rel RMS {rms:.3e}")
    assert rms < 1e-4


def test_tree_potential_periodic_straddle(gpu):
    # a subhalo across the periodic boundary: the tree is built on raw
    # coordinates and every separation is wrapped (NEAREST)
    cat = synth.config2(seed=11, nsub=1, hi=2000)
    c = cat.cols
    x = np.mod(c["x"] - c["x"].mean() + 5.0, 1e5).astype(np.float32)  # centre near x=5
    src = (x, c["y"], c["z"], c["mass"])
    tgt = (x, c["y"], c["z"])
    for prec in (abi.HBT_PREC_REPLAY, abi.HBT_PREC_FAST):
        p = _params(cat, precision=prec)
        ref = oracle.tree_potential(p, src, tgt, c["mass"])
        got = gpu(p).tree_potential(src, tgt, c["mass"])
        rms = np.sqrt(np.mean(((got - ref) / ref) ** 2))
        print(f"straddle prec={prec}: rel RMS {rms:.3e}, differ {int(np.sum(got != ref))}")
        if prec == abi.HBT_PREC_REPLAY:
            assert np.array_equal(ref, got)
        else:
            assert rms < 1e-4


def test_unbind_replay_small_batch(gpu):
    cat = synth.config2(seed=21, nsub=60, hi=3000)
    p = _params(cat, precision=abi.HBT_PREC_REPLAY)
    ref = oracle.unbind_batch(cat, params=p)
    U = gpu(p)
    got = U.unbind_batch(cat)
    assert_exact(ref, got)


def test_unbind_replay_c1(gpu):
    for bmp in (0.5, 0.995):
        cat = synth.config1(bound_mass_precision=bmp)
        p = _params(cat, precision=abi.HBT_PREC_REPLAY)
        ref = oracle.unbind_batch(cat, params=p)
        got = gpu(p).unbind_batch(cat)
        print(f"C1 bmp={bmp}: nbound ref {ref[0]['nbound']} got {got[0]['nbound']}")
        assert_exact(ref, got)


def test_unbind_fast_small_batch(gpu):
    cat = synth.config2(seed=21, nsub=60, hi=3000)
    p = _params(cat, precision=abi.HBT_PREC_FAST)
    ref = oracle.unbind_batch(cat, params=p)
    got = gpu(p).unbind_batch(cat)
    st = assert_tolerant(ref, got)
    print("fast vs oracle:", st)


def test_plain_equals_batched(gpu):
    cat = synth.config2(seed=23, nsub=30, hi=3000)
    for prec in (abi.HBT_PREC_FAST, abi.HBT_PREC_REPLAY):
        U = gpu(_params(cat, precision=prec))
        a = U.unbind_batch(cat, plain=False)
        b = U.unbind_batch(cat, plain=True)
        for x, y in zip(a, b):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_nested_replay(gpu):
    cat = synth.config4(n_host=20000, n_sat=12, sat_lo=200, sat_hi=4000)
    p = _params(cat, precision=abi.HBT_PREC_REPLAY)
    ref = oracle.unbind_batch(cat, params=p)
    got = gpu(p).unbind_batch(cat)
    assert_exact(ref, got)
