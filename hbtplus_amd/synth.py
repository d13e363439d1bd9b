"""Synthetic subhalo catalogues for the five unbinding workloads.

Units: kpc/h, km/s, 1e10 Msun/h, so G = 43007.1.  All randomness comes from
numpy's counter-based Philox generator with a fixed seed per workload, so the
same inputs are produced on every machine.

Each generator returns a ``Catalogue``: particle columns (id int64, x/y/z,
vx/vy/vz, mass float32, type int32), an ``hbt_subhalo`` record array with the
inputs filled (offset, count, parent, nbound_prev, initial frame, previous
most-bound coordinates, death/sink markers) and the ``hbt_params`` record.

Physics of the haloes (used only to make realistic inputs):
  NFW density truncated at r200; radii by inverse CDF of
  mu(x) = ln(1+x) - x/(1+x); isotropic Gaussian velocities whose 1-D
  dispersion solves the spherical Jeans equation with sigma(r200) = 0,
      sigma^2(x) = V^2 x (1+x)^2 * int_x^c mu(t) / (t^3 (1+t)^2) dt,
      V^2 = G M200 / (r_s mu(c));
  escape speed of the truncated profile
      v_esc^2(x) = 2 V^2 [ln(1+x)/x - 1/(1+c)].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .abi import SUB_DTYPE, make_params

G_DEFAULT = 43007.1
PARTICLE_MASS = 1e-4


@dataclass
class Catalogue:
    cols: dict
    subs: np.ndarray
    params: np.ndarray
    name: str = ""

    @property
    def n_particles(self) -> int:
        return len(self.cols["x"])

    def subset(self, sel) -> "Catalogue":
        """A catalogue with only the subhalos `sel` (top-level only; parents
        are not followed), particles re-packed."""
        sel = np.asarray(sel)
        subs = self.subs[sel].copy()
        idx = np.concatenate([np.arange(s["offset"], s["offset"] + s["count"]) for s in subs]) \
            if len(subs) else np.zeros(0, np.int64)
        cols = {k: v[idx].copy() for k, v in self.cols.items()}
        off = np.zeros(len(subs), np.int64)
        if len(subs):
            off[1:] = np.cumsum(subs["count"])[:-1]
        subs["offset"] = off
        subs["parent"] = -1
        return Catalogue(cols, subs, self.params.copy(), self.name + "[subset]")


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=seed))


def _mu(x):
    return np.log1p(x) - x / (1.0 + x)


def nfw_unit_radii(rng, c: np.ndarray) -> np.ndarray:
    """x = r/r_s with M(<x) uniform: solve mu(x) = u mu(c) by inverting a table
    of mu on a log grid (per 1% bin of c), then three Newton steps."""
    c = np.asarray(c, np.float64)
    target = rng.random(c.shape) * _mu(c)
    if c.size == 0:
        return c.copy()
    lc = np.log(c)
    lc0 = lc.min()
    bins = np.floor((lc - lc0) / 0.01).astype(np.int64)
    x = np.empty_like(c)
    ly = np.linspace(np.log(1e-7), 0.0, 2048)
    order = np.argsort(bins, kind="stable")
    edges = np.searchsorted(bins[order], np.arange(bins.max() + 2))
    for b in range(len(edges) - 1):
        sel = order[edges[b]:edges[b + 1]]
        if sel.size == 0:
            continue
        cb = float(np.exp(lc0 + 0.01 * (b + 0.5)))
        ms = _mu(cb * np.exp(ly)) / _mu(cb)
        x[sel] = c[sel] * np.exp(np.interp(target[sel] / _mu(c[sel]), ms, ly))
    for _ in range(3):  # Newton on mu(x) - target, mu'(x) = x / (1+x)^2
        f = _mu(x) - target
        x = np.clip(x - f * (1.0 + x) ** 2 / np.maximum(x, 1e-12), 1e-9, c)
    return x


class _JeansTable:
    """sigma^2(x; c) / V^2 tabulated on a log grid of c and of x/c."""

    def __init__(self, cmin: float, cmax: float, nc: int = 48, nx: int = 1024):
        self.lc = np.linspace(np.log(cmin * 0.99), np.log(cmax * 1.01), nc)
        self.ly = np.linspace(np.log(1e-6), 0.0, nx)  # y = x/c
        tab = np.zeros((nc, nx))
        for k, lc in enumerate(self.lc):
            c = np.exp(lc)
            x = c * np.exp(self.ly)
            f = _mu(x) / (x ** 3 * (1 + x) ** 2) * x  # integrand * dt/dln t
            # cumulative integral from x to c (trapezoid in ln x)
            seg = 0.5 * (f[1:] + f[:-1]) * np.diff(np.log(x))
            I = np.concatenate([np.cumsum(seg[::-1])[::-1], [0.0]])
            tab[k] = x * (1 + x) ** 2 * I
        self.tab = tab

    def __call__(self, x: np.ndarray, c: np.ndarray) -> np.ndarray:
        kc = np.clip(np.rint((np.log(c) - self.lc[0]) / (self.lc[1] - self.lc[0])), 0,
                     len(self.lc) - 1).astype(np.int64)
        ly = np.log(np.clip(x / c, 1e-6, 1.0))
        fy = (ly - self.ly[0]) / (self.ly[1] - self.ly[0])
        jy = np.clip(np.floor(fy).astype(np.int64), 0, len(self.ly) - 2)
        w = fy - jy
        return self.tab[kc, jy] * (1 - w) + self.tab[kc, jy + 1] * w


def _isotropic(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _uniform_ball(rng, n, r):
    return _isotropic(rng, n) * (r * rng.random(n) ** (1.0 / 3.0))[:, None]


def power_law_counts(rng, n, lo, hi, alpha=1.9):
    """dN/dN_p ~ N_p^-alpha on [lo, hi] by inverse CDF, rounded."""
    a, b = lo ** (1 - alpha), hi ** (1 - alpha)
    u = rng.random(n)
    return np.rint((a + u * (b - a)) ** (1.0 / (1 - alpha))).astype(np.int64)


def unique_ids(rng, n) -> np.ndarray:
    """n distinct int64 IDs in [0, 2^62), scattered and shuffled by seed."""
    perm = rng.permutation(n).astype(np.uint64)
    return ((perm * np.uint64(2654435761) + np.uint64(1000003)) & np.uint64((1 << 62) - 1)).astype(np.int64)


def _nfw_members(rng, owner, c, r200, jeans, G, m):
    """NFW positions (relative) and Jeans velocities (relative) for particles
    whose subhalo parameters are c[owner], r200[owner]."""
    cc = c[owner]
    rs = r200[owner] / cc
    x = nfw_unit_radii(rng, cc)
    pos = _isotropic(rng, len(owner)) * (x * rs)[:, None]
    n_of = np.bincount(owner, minlength=len(c)).astype(np.float64)
    M200 = n_of * m
    V2 = G * M200[owner] / (rs * _mu(cc))
    sig = np.sqrt(np.maximum(V2 * jeans(x, cc), 0.0))
    vel = rng.normal(size=(len(owner), 3)) * sig[:, None]
    return pos, vel


def _escape_speed(r, c, r200, M200, G):
    rs = r200 / c
    x = np.maximum(r / rs, 1e-9)
    V2 = G * M200 / (rs * _mu(c))
    return np.sqrt(2.0 * V2 * np.maximum(np.log1p(x) / x - 1.0 / (1.0 + c), 0.0))


def _assemble(rng, groups, box, periodic, params, name, parent=None, shuffle_lists=True):
    """groups: list over subhalos of (centre[3], bulk[3], rel_pos[n,3], rel_vel[n,3]).
    Builds columns with lists contiguous per subhalo, positions wrapped into the
    box when periodic, IDs unique, and the initial frame = mass-weighted mean."""
    counts = np.array([len(g[2]) for g in groups], np.int64)
    n = int(counts.sum())
    off = np.zeros(len(groups), np.int64)
    if len(groups):
        off[1:] = np.cumsum(counts)[:-1]
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    frame_p = np.empty((len(groups), 3))
    frame_v = np.empty((len(groups), 3))
    for s, (cen, bulk, rp, rv) in enumerate(groups):
        k = len(rp)
        order = rng.permutation(k) if shuffle_lists else np.arange(k)
        rp, rv = rp[order], rv[order]
        pos[off[s]:off[s] + k] = cen + rp
        vel[off[s]:off[s] + k] = bulk + rv
        frame_p[s] = cen + (rp.mean(axis=0) if k else 0.0)
        frame_v[s] = bulk + (rv.mean(axis=0) if k else 0.0)
    if periodic:
        pos = np.mod(pos, box)
        frame_p = np.mod(frame_p, box)
    cols = {
        "id": unique_ids(rng, n),
        "x": pos[:, 0].astype(np.float32), "y": pos[:, 1].astype(np.float32),
        "z": pos[:, 2].astype(np.float32),
        "vx": vel[:, 0].astype(np.float32), "vy": vel[:, 1].astype(np.float32),
        "vz": vel[:, 2].astype(np.float32),
        "mass": np.full(n, PARTICLE_MASS, np.float32),
        "type": np.ones(n, np.int32),
    }
    subs = np.zeros(len(groups), SUB_DTYPE)
    subs["offset"] = off
    subs["count"] = counts
    subs["parent"] = -1 if parent is None else parent
    subs["nbound_prev"] = counts
    subs["sink_track_id"] = -1
    subs["snapshot_index_of_death"] = -1
    subs["snapshot_index_of_sink"] = -1
    subs["ref_pos"] = frame_p.astype(np.float32)
    subs["ref_vel"] = frame_v.astype(np.float32)
    subs["mostbound_pos"] = subs["ref_pos"]
    subs["mostbound_vel"] = subs["ref_vel"]
    return Catalogue(cols, subs, params, name)


def _power_law_population(rng, nsub, lo, hi, cont_frac, cont_mode, c_median, c_sigma_ln, G,
                          box, centre_fn, log_uniform=False):
    """Arrays of a power-law subhalo population, grouped by subhalo with a
    random order inside each group: (owner, relative positions, relative
    velocities, centres, bulk velocities)."""
    if log_uniform:
        npart = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), nsub))).astype(np.int64)
    else:
        npart = power_law_counts(rng, nsub, lo, hi)
    if c_sigma_ln > 0:
        c = c_median * np.exp(c_sigma_ln * rng.normal(size=nsub))
    else:
        c = np.full(nsub, float(c_median))
    r200 = 200.0 * (npart / 2.0e4) ** (1.0 / 3.0)
    jeans = _JeansTable(c.min(), c.max())
    owner = np.repeat(np.arange(nsub), npart)
    rel_p, rel_v = _nfw_members(rng, owner, c, r200, jeans, G, PARTICLE_MASS)
    ncont = np.rint(cont_frac * npart).astype(np.int64)
    cowner = np.repeat(np.arange(nsub), ncont)
    cp = _isotropic(rng, len(cowner)) * (r200[cowner] * rng.random(len(cowner)) ** (1 / 3.0))[:, None]
    if cont_mode == "escape":
        vesc = _escape_speed(np.linalg.norm(cp, axis=1), c[cowner], r200[cowner],
                             npart[cowner] * PARTICLE_MASS, G)
        cv = _isotropic(rng, len(cowner)) * (3.0 * vesc)[:, None]
    else:  # Gaussian sigma=1000 km/s plus a 300 km/s bulk offset along +x
        cv = rng.normal(size=(len(cowner), 3)) * 1000.0
        cv[:, 0] += 300.0
    centres = centre_fn(nsub)
    bulk = rng.normal(size=(nsub, 3)) * 300.0
    own = np.concatenate([owner, cowner])
    order = np.argsort(own + rng.random(len(own)))  # group by subhalo, random order inside
    return (own[order], np.concatenate([rel_p, cp])[order], np.concatenate([rel_v, cv])[order],
            centres, bulk)


def _assemble_arrays(rng, own, rel_p, rel_v, centres, bulk, box, periodic, params, name):
    """Vectorised form of _assemble for particles already grouped by owner."""
    nsub = len(centres)
    counts = np.bincount(own, minlength=nsub).astype(np.int64)
    off = np.zeros(nsub, np.int64)
    off[1:] = np.cumsum(counts)[:-1]
    pos = centres[own] + rel_p
    vel = bulk[own] + rel_v
    safe = np.maximum(counts, 1)[:, None]
    frame_p = centres + np.stack([np.bincount(own, rel_p[:, j], nsub) for j in range(3)], 1) / safe
    frame_v = bulk + np.stack([np.bincount(own, rel_v[:, j], nsub) for j in range(3)], 1) / safe
    if periodic:
        pos = np.mod(pos, box)
        frame_p = np.mod(frame_p, box)
    n = len(own)
    cols = {
        "id": unique_ids(rng, n),
        "x": pos[:, 0].astype(np.float32), "y": pos[:, 1].astype(np.float32),
        "z": pos[:, 2].astype(np.float32),
        "vx": vel[:, 0].astype(np.float32), "vy": vel[:, 1].astype(np.float32),
        "vz": vel[:, 2].astype(np.float32),
        "mass": np.full(n, PARTICLE_MASS, np.float32),
        "type": np.ones(n, np.int32),
    }
    subs = np.zeros(nsub, SUB_DTYPE)
    subs["offset"] = off
    subs["count"] = counts
    subs["parent"] = -1
    subs["nbound_prev"] = counts
    subs["sink_track_id"] = -1
    subs["snapshot_index_of_death"] = -1
    subs["snapshot_index_of_sink"] = -1
    subs["ref_pos"] = frame_p.astype(np.float32)
    subs["ref_vel"] = frame_v.astype(np.float32)
    subs["mostbound_pos"] = subs["ref_pos"]
    subs["mostbound_vel"] = subs["ref_vel"]
    return Catalogue(cols, subs, params, name)


def config1(seed: int = 1001, bound_mass_precision: float = 0.5, n_nfw=20000, n_cont=4000) -> Catalogue:
    """Single isolated subhalo: 20,000 NFW particles (r_s=20, c=10, r200=200)
    plus 4,000 uniform contaminants with Gaussian sigma=1000 km/s and a
    300 km/s bulk offset; eps=0.5, theta=0.45."""
    rng = _rng(seed)
    c = np.array([10.0])
    r200 = np.array([200.0])
    jeans = _JeansTable(9.9, 10.1)
    owner = np.zeros(n_nfw, np.int64)
    rp, rv = _nfw_members(rng, owner, c, r200, jeans, G_DEFAULT, PARTICLE_MASS)
    cp = _uniform_ball(rng, n_cont, 200.0)
    cv = rng.normal(size=(n_cont, 3)) * 1000.0
    cv[:, 0] += 300.0
    groups = [(np.zeros(3), np.zeros(3), np.concatenate([rp, cp]), np.concatenate([rv, cv]))]
    params = make_params(softening=0.5, open_angle=0.45, periodic=0,
                         bound_mass_precision=bound_mass_precision)
    return _assemble(rng, groups, 1e5, False, params, f"C1 single 24k NFW (BoundMassPrecision={bound_mass_precision})")


def config2(seed: int = 1002, nsub: int = 5000, hi: float = 2.0e5, at_edges: bool = False) -> Catalogue:
    """5,000 subhalos in a periodic 100 Mpc/h box, N_p ~ N^-1.9 on [20, 2e5],
    log-normal c (median 10, sigma_ln 0.3), r200 = 200 (N_p/2e4)^(1/3),
    20% uniform contaminants at 3x the local escape speed.  at_edges: each
    centre coordinate is put on the box edge (0) with probability 0.6, so
    most subhalos wrap around one to three faces of the periodic box."""
    rng = _rng(seed)
    box = 1.0e5

    def centres(n):
        c = rng.random((n, 3)) * box
        if at_edges:
            c[rng.random((n, 3)) < 0.6] = 0.0
        return c

    pop = _power_law_population(rng, nsub, 20, hi, 0.2, "escape", 10.0, 0.3, G_DEFAULT, box, centres)
    params = make_params(softening=0.5, open_angle=0.45, periodic=1, box_size=box)
    name = f"C2 {nsub} power-law subhalos" + (" at the box edges" if at_edges else "")
    return _assemble_arrays(rng, *pop, box, True, params, name)


def config3(seed: int = 1003, nsub: int = 50000, contaminants: str = "gauss") -> Catalogue:
    """Direct-sum precision study: N_p log-uniform on [20, 4096], c=10,
    15% contaminants as in config 1 (contaminants="gauss"), or at 3x the
    escape speed as in config 2 ("escape": most subhalos then converge)."""
    rng = _rng(seed)
    box = 1.0e5
    pop = _power_law_population(rng, nsub, 20, 4096, 0.15, contaminants, 10.0, 0.0, G_DEFAULT, box,
                                lambda n: rng.random((n, 3)) * box, log_uniform=True)
    params = make_params(softening=0.5, open_angle=0.45, periodic=1, box_size=box)
    return _assemble_arrays(rng, *pop, box, True, params, f"C3 {nsub} log-uniform subhalos")


def config4(seed: int = 1004, n_host: int = 4_000_000, n_sat: int = 200,
            sat_lo: float = 1e3, sat_hi: float = 1e5) -> Catalogue:
    """Nested host: an NFW host (c=5, r200=1000) with n_sat NFW satellites
    (N_p ~ N^-1.9 on [1e3, 1e5], c=10) at positions uniform in r<r200; the
    satellites are nested in the host (parent = 0), lists disjoint as after
    exclusive masking.  eps=1, theta=0.45."""
    rng = _rng(seed)
    host_c, host_r200 = np.array([5.0]), np.array([1000.0])
    jeans_h = _JeansTable(4.9, 5.1)
    hp, hv = _nfw_members(rng, np.zeros(n_host, np.int64), host_c, host_r200, jeans_h, G_DEFAULT,
                          PARTICLE_MASS)
    npart = power_law_counts(rng, n_sat, sat_lo, sat_hi)
    c = np.full(n_sat, 10.0)
    r200 = 200.0 * (npart / 2.0e4) ** (1.0 / 3.0)
    jeans = _JeansTable(9.9, 10.1)
    owner = np.repeat(np.arange(n_sat), npart)
    sp, sv = _nfw_members(rng, owner, c, r200, jeans, G_DEFAULT, PARTICLE_MASS)
    centres = _uniform_ball(rng, n_sat, 1000.0)
    # satellite bulk velocity ~ host dispersion at its radius
    xs = np.linalg.norm(centres, axis=1) / (1000.0 / 5.0)
    V2 = G_DEFAULT * n_host * PARTICLE_MASS / ((1000.0 / 5.0) * _mu(5.0))
    sig = np.sqrt(np.maximum(V2 * jeans_h(xs, np.full(n_sat, 5.0)), 0.0))
    bulk = rng.normal(size=(n_sat, 3)) * sig[:, None]
    start = np.concatenate([[0], np.cumsum(npart)])
    groups = [(np.zeros(3), np.zeros(3), hp, hv)]
    for s in range(n_sat):
        groups.append((centres[s], bulk[s], sp[start[s]:start[s + 1]], sv[start[s]:start[s + 1]]))
    parent = np.concatenate([[-1], np.zeros(n_sat, np.int64)])
    params = make_params(softening=1.0, open_angle=0.45, periodic=0)
    return _assemble(rng, groups, 1e5, False, params, f"C4 {n_host} host + {n_sat} satellites",
                     parent=parent)


def config5(seed: int = 1005, nsub: int = 200_000, background_frac: float = 0.1):
    """One rank's share: nsub subhalos from the config-2 law in a periodic
    100 Mpc/h box inside a snapshot that also holds a uniform background
    (background_frac of the listed particles).  The snapshot is shuffled and
    every subhalo is given as a list of particle IDs to be resolved on the
    device.  Returns (snapshot columns, id lists, catalogue-of-lists)."""
    rng = _rng(seed)
    box = 1.0e5
    pop = _power_law_population(rng, nsub, 20, 2.0e5, 0.2, "escape", 10.0, 0.3, G_DEFAULT, box,
                                lambda n: rng.random((n, 3)) * box)
    params = make_params(softening=0.5, open_angle=0.45, periodic=1, box_size=box)
    cat = _assemble_arrays(rng, *pop, box, True, params, f"C5 {nsub} subhalos")
    nlist = cat.n_particles
    nbg = int(background_frac * nlist)
    bg = {
        "x": (rng.random(nbg) * box).astype(np.float32), "y": (rng.random(nbg) * box).astype(np.float32),
        "z": (rng.random(nbg) * box).astype(np.float32),
        "vx": (rng.normal(size=nbg) * 300).astype(np.float32),
        "vy": (rng.normal(size=nbg) * 300).astype(np.float32),
        "vz": (rng.normal(size=nbg) * 300).astype(np.float32),
        "mass": np.full(nbg, PARTICLE_MASS, np.float32), "type": np.ones(nbg, np.int32),
    }
    all_ids = unique_ids(rng, nlist + nbg)
    cat.cols["id"] = all_ids[:nlist]
    bg["id"] = all_ids[nlist:]
    snap = {k: np.concatenate([cat.cols[k], bg[k]]) for k in cat.cols}
    perm = rng.permutation(len(snap["x"]))
    snap = {k: v[perm] for k, v in snap.items()}
    return snap, cat.cols["id"].copy(), cat


CONFIGS = {"c1": config1, "c2": config2, "c3": config3, "c4": config4, "c5": config5}


def masking_catalogue(seed=1006, ngroups=40, max_subs=12, pool=4000, big_group=0):
    """Host groups whose subhalo source lists overlap, as they do before
    MaskSubhalos: each group draws its lists from a shared ID pool (the same
    ID appears as several rows of the columns), with nesting inside the group,
    orphans (nbound_prev <= 1), repeated IDs inside a list, null IDs (-1) and
    field subhalos (host group -1).  big_group > 0 adds one group whose lists
    hold about that many entries in total.  Returns (cols, subs, host_group)."""
    from .abi import SUB_DTYPE
    rng = np.random.Generator(np.random.Philox(key=seed))
    ids, subs, groups = [], [], []
    next_id = 1
    off = 0

    def add(g, parent, nprev, lst):
        nonlocal off
        r = np.zeros(1, SUB_DTYPE)[0]
        r["offset"], r["count"], r["parent"], r["nbound_prev"] = off, len(lst), parent, nprev
        subs.append(r)
        groups.append(g)
        ids.append(np.asarray(lst, np.int64))
        off += len(lst)
        return len(subs) - 1

    sizes = [(g, int(rng.integers(1, max_subs + 1)), pool) for g in range(ngroups)]
    if big_group:
        sizes.append((ngroups, 6, big_group))
    for g, ns, psize in sizes:
        pool_ids = np.arange(next_id, next_id + psize, dtype=np.int64)
        next_id += psize
        members = []
        for k in range(ns):
            n = int(rng.integers(0, max(2, psize // max(ns, 1)) * 2))
            lst = rng.choice(pool_ids, size=n, replace=True) if n else np.zeros(0, np.int64)
            if n and rng.random() < 0.2:
                lst[rng.integers(0, n, max(1, n // 50))] = -1
            parent = -1 if (k == 0 or rng.random() < 0.3) else int(members[int(rng.integers(0, len(members)))])
            nprev = 1 if rng.random() < 0.15 else int(max(2, n))
            members.append(add(g, parent, nprev, lst))
    for k in range(5):  # field subhalos
        lst = rng.integers(1, next_id, 50)
        add(-1, -1, 50, lst)
    ids = np.concatenate(ids) if ids else np.zeros(0, np.int64)
    n = len(ids)
    cols = dict(id=ids, x=rng.random(n, dtype=np.float32), y=rng.random(n, dtype=np.float32),
                z=rng.random(n, dtype=np.float32), vx=np.zeros(n, np.float32), vy=np.zeros(n, np.float32),
                vz=np.zeros(n, np.float32), mass=np.ones(n, np.float32), type=np.ones(n, np.int32))
    return cols, np.array(subs, SUB_DTYPE), np.array(groups, np.int64)


def merging_catalogue(seed=1007, ngroups=3, periodic=True, box=1000.0, hz=0.0, scale_factor=1.0,
                      snapshot_index=5, scale=1.0):
    """Host groups laid out so that merging has work to do.  Per group: a
    central NFW halo; nested in it a satellite sinking onto its core (with a
    subhalo nested in the satellite and sinking onto the satellite's core), a
    satellite orbiting far out (with one nested subhalo sinking onto that
    satellite and one lying at the central's core; in the first group also
    an empty one), a hot subhalo that disrupts and a one-particle orphan at
    the centre; a second head of the
    group sinking onto the central and a head far out.  Plus a field halo with
    a nested subhalo sinking onto it.  The first group sits across the
    periodic boundary.  `scale` multiplies every particle count.  Returns
    (Catalogue, host_group)."""
    rng = _rng(seed)
    jeans = _JeansTable(5.0, 20.0)
    G = G_DEFAULT

    def halo(n, c):
        n = max(1, int(round(n * scale))) if n > 1 else 1
        r200 = 200.0 * (n / 2.0e4) ** (1.0 / 3.0)
        rp, rv = _nfw_members(rng, np.zeros(n, np.int64), np.array([c]), np.array([r200]), jeans, G,
                              PARTICLE_MASS)
        return rp, rv, r200, np.sqrt(G * n * PARTICLE_MASS / r200)

    entries, parent, group, nprev = [], [], [], []

    def add(cen, bulk, rp, rv, par, grp, orphan=False):
        entries.append((np.asarray(cen, np.float64), np.asarray(bulk, np.float64), rp, rv))
        parent.append(par)
        group.append(grp)
        nprev.append(1 if orphan else len(rp))
        return len(entries) - 1

    def near(c, r):
        return np.asarray(c) + _isotropic(rng, 1)[0] * r

    for g in range(ngroups):
        C = np.array([box - 3.0, 2.0, 0.5 * box]) if (g == 0 and periodic) else rng.uniform(0.2, 0.8, 3) * box
        B = rng.normal(size=3) * 50.0
        rp, rv, r200, v200 = halo(rng.integers(1200, 1800), 8.0)
        cen = add(C, B, rp, rv, -1, g)
        c1, b1 = near(C, 0.3), B + rng.normal(size=3) * 0.5
        rp, rv, _, _ = halo(rng.integers(150, 300), 15.0)
        s1 = add(c1, b1, rp, rv, cen, g)
        rp, rv, _, _ = halo(60, 15.0)
        add(near(c1, 0.1), b1 + rng.normal(size=3) * 0.3, rp, rv, s1, g)
        c2, b2 = near(C, 0.5 * r200), B + _isotropic(rng, 1)[0] * v200
        rp, rv, _, _ = halo(rng.integers(300, 500), 12.0)
        s2 = add(c2, b2, rp, rv, cen, g)
        rp, rv, _, _ = halo(rng.integers(50, 80), 15.0)
        add(near(c2, 0.1), b2 + rng.normal(size=3) * 0.2, rp, rv, s2, g)
        rp, rv, _, _ = halo(rng.integers(50, 80), 15.0)
        add(near(C, 0.2), B + rng.normal(size=3) * 0.3, rp, rv, s2, g)
        if g == 0:  # an empty list nested in the far satellite
            add(C, B, np.zeros((0, 3)), np.zeros((0, 3)), s2, g)
        rp, rv, _, _ = halo(40, 10.0)
        add(near(C, 5.0), B, rp, rv * 20.0, cen, g)  # hot: disrupts
        add(C, B, np.zeros((1, 3)), np.zeros((1, 3)), cen, g, orphan=True)
        rp, rv, _, _ = halo(rng.integers(100, 200), 12.0)
        add(near(C, 0.3), B + rng.normal(size=3) * 0.5, rp, rv, -1, g)
        rp, rv, _, _ = halo(100, 12.0)
        add(near(C, 0.8 * r200), B + _isotropic(rng, 1)[0] * 2.0 * v200, rp, rv, -1, g)
    F = rng.uniform(0.2, 0.8, 3) * box
    BF = rng.normal(size=3) * 50.0
    rp, rv, _, _ = halo(600, 8.0)
    f = add(F, BF, rp, rv, -1, -1)
    rp, rv, _, _ = halo(60, 15.0)
    add(near(F, 0.2), BF + rng.normal(size=3) * 0.3, rp, rv, f, -1)
    params = make_params(softening=0.5, open_angle=0.45, periodic=1 if periodic else 0, box_size=box,
                         hz=hz, scale_factor=scale_factor, snapshot_index=snapshot_index)
    cat = _assemble(rng, entries, box, periodic, params, f"merging {ngroups} groups",
                    parent=np.array(parent, np.int64))
    cat.subs["nbound_prev"] = np.array(nprev, np.int64)
    return cat, np.array(group, np.int64)


def nesting_catalogue(seed=1011, ngroups=60, nsub=600, nfield=40):
    """Inputs of the nest bookkeeping: host groups with one flagged central
    each, field subhalos, and nests that cross groups (a subhalo nested in a
    host of another halo), with Ranks as a previous snapshot left them: roots
    mostly heads (Rank 0) but some not, nested subhalos mostly satellites
    (Rank >= 1) but some Rank 0.  Indices are shuffled; nests are acyclic.
    Returns (host_group, is_central, parent, rank)."""
    rng = np.random.Generator(np.random.Philox(key=seed))
    n = nsub + nfield
    group = np.concatenate([rng.integers(0, ngroups, nsub), np.full(nfield, -1)]).astype(np.int64)
    parent = np.full(n, -1, np.int64)
    rank = np.zeros(n, np.int32)
    u = rng.random(n)
    pick = rng.random(n)
    members = {}
    for s in range(n):
        same = members.setdefault(int(group[s]), [])
        if s > 0 and u[s] >= 0.25:  # else a root
            if u[s] < 0.85 and same:
                parent[s] = same[int(pick[s] * len(same))]
            else:
                parent[s] = int(pick[s] * s)
        same.append(s)
    roots = parent < 0
    rank[~roots] = rng.integers(1, 9, int((~roots).sum()))
    rank[~roots & (rng.random(n) < 0.08)] = 0
    rank[roots & (rng.random(n) < 0.1)] = rng.integers(1, 5)
    central = np.zeros(n, np.int32)
    for g in range(ngroups):
        m = np.nonzero(group == g)[0]
        if len(m):
            central[m[rng.integers(0, len(m))]] = 1
    perm = rng.permutation(n)      # new index of old subhalo i: inv[i]
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    parent_new = np.where(parent[perm] >= 0, inv[np.maximum(parent[perm], 0)], -1)
    return group[perm], central[perm], parent_new.astype(np.int64), rank[perm]
