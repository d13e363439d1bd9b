"""Writes tests/golden/c4_full_oracle.npz: the oracle's unbinding of the whole
C4 catalogue (synth.config4(): one 4,000,000-particle host holding 200
satellites, 5.33 M particles), reduced to what a FAST-precision result can be
checked against without shipping 4 M particle IDs:

* per subhalo: nbound, mbound, the most-bound particle's ID, whether the two
  most bound energies lie within the band (then the most-bound ID may differ);
* per subhalo: the IDs of its listed particles whose binding energy lies within
  the band |E| < 1e-5 |E_min| (the only particles whose bound/unbound outcome
  FAST may flip), and the SHA-1 of the sorted IDs of its bound particles
  outside the band;
* the SHA-1 of the catalogue's position columns, so that a change of the
  generator is caught instead of read as a parity failure.

The host's oracle run is serial (about 15 minutes on one core of a 2.7 GHz
Xeon):

    python tests/golden/make_c4_full.py
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from hbtplus_amd import synth  # noqa: E402

BAND_REL = 1e-5
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c4_full_oracle.npz")


def catalogue_digest(cat) -> bytes:
    h = hashlib.sha1()
    for k in ("x", "y", "z", "vx", "vy", "vz", "mass"):
        h.update(np.ascontiguousarray(cat.cols[k]).tobytes())
    h.update(np.ascontiguousarray(cat.subs).tobytes())
    return h.digest()


def reduce_result(cat, out, band_rel=BAND_REL):
    """Per-subhalo summary of an unbinding result (see the module docstring);
    the band is taken from the energies in `out` (the oracle's)."""
    subs, idx, en = out
    ids = cat.cols["id"]
    n = len(subs)
    mbid = np.zeros(n, np.uint64)
    tie = np.zeros(n, bool)
    digest = np.zeros((n, 20), np.uint8)
    band_ids, band_off = [], np.zeros(n + 1, np.int64)
    for s in range(n):
        o, L, nb = int(subs["out_offset"][s]), int(subs["list_len"][s]), int(subs["nbound"][s])
        li, le = idx[o:o + L], en[o:o + L]
        emin = np.nanmin(le) if L else 0.0
        band = band_rel * abs(emin)
        near = np.abs(le) < band
        bid = np.sort(ids[li[:nb][~near[:nb]]].astype(np.uint64))
        digest[s] = np.frombuffer(hashlib.sha1(bid.tobytes()).digest(), np.uint8)
        band_ids.append(ids[li[near]].astype(np.uint64))
        band_off[s + 1] = band_off[s] + int(near.sum())
        if L:
            mbid[s] = ids[li[0]]
            top = np.sort(le[:nb])[:2]
            tie[s] = len(top) > 1 and top[1] - top[0] <= band
    return dict(nbound=subs["nbound"].astype(np.int64), mbound=subs["mbound"].astype(np.float32),
                mostbound_id=mbid, top_tie=tie, bound_digest=digest,
                band_ids=np.concatenate(band_ids) if band_ids else np.zeros(0, np.uint64), band_off=band_off)


def main():
    t = time.time()
    cat = synth.config4()
    print(f"catalogue: {cat.n_particles} particles, {len(cat.subs)} subhalos", flush=True)
    out = oracle.unbind_batch(cat, nthreads=os.cpu_count() or 1)
    print(f"oracle: {time.time() - t:.0f} s", flush=True)
    red = reduce_result(cat, out)
    np.savez_compressed(OUT, catalogue_sha1=np.frombuffer(catalogue_digest(cat), np.uint8), **red)
    print(f"wrote {OUT}: host nbound {red['nbound'][0]}, {len(red['band_ids'])} band particles", flush=True)


if __name__ == "__main__":
    main()
