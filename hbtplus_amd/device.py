"""Host-side driver of libhbt_unbind.so through its C-ABI.

torch is used only to hold device memory; every computation is a call into
the HIP library.  Mirrors the reference's entry points:
  GpuUnbinder.unbind_batch   SubhaloSnapshot_t::RefineParticles (batched or plain)
  GpuUnbinder.unbind_one     Subhalo_t::Unbind (host buffers)
  GpuUnbinder.tree_potential GravityTree_t::Build + EvaluatePotential
  GpuUnbinder.direct_potential, GpuUnbinder.id_to_index
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import abi
from .abi import (PARAMS_DTYPE, PARTICLE_COLUMNS, SUB_DTYPE, check, load_library,
                  particles_struct, plan_batch)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the HIP library needs a device")
    return torch


class GpuUnbinder:
    def __init__(self, params, device: int = 0):
        self.torch = _torch()
        self.lib = load_library()
        self.device = device
        self.params = np.ascontiguousarray(params, PARAMS_DTYPE).copy()
        self.ctx = ctypes.c_void_p()
        check(self.lib.hbt_ctx_create(self.params.ctypes.data, device, ctypes.byref(self.ctx)),
              "hbt_ctx_create")

    # context management -------------------------------------------------
    def close(self):
        if self.ctx:
            self.lib.hbt_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        self.params = np.ascontiguousarray(params, PARAMS_DTYPE).copy()
        check(self.lib.hbt_ctx_set_params(self.ctx, self.params.ctypes.data), "hbt_ctx_set_params")

    def set_rand_state(self, seed=1, draws=0):
        """The rand() stream the sampling shuffles draw from: srand(seed) then
        `draws` calls (a new context: seed 1, no calls)."""
        check(self.lib.hbt_set_rand_state(self.ctx, ctypes.c_uint32(seed), ctypes.c_int64(draws)),
              "hbt_set_rand_state")

    def rand_state(self):
        seed, draws = ctypes.c_uint32(), ctypes.c_int64()
        check(self.lib.hbt_get_rand_state(self.ctx, ctypes.byref(seed), ctypes.byref(draws)),
              "hbt_get_rand_state")
        return {"seed": seed.value, "draws": draws.value}

    def _dev(self, a):
        return self.torch.from_numpy(np.ascontiguousarray(a)).to(f"cuda:{self.device}")

    # device-resident inputs ----------------------------------------------
    def upload(self, cat):
        """Copy a catalogue to the device once; returns a handle for unbind_device."""
        cols = {k: self._dev(cat.cols[k]) for k in PARTICLE_COLUMNS}
        subs = cat.subs.copy()
        total = plan_batch(subs)
        return dict(cols=cols, subs_host=subs, total=total,
                    subs0=self._dev(subs.view(np.uint8)),
                    ps=particles_struct(cols, ptr=lambda t: t.data_ptr()))

    def unbind_device(self, h, plain=False):
        """Run on device-resident inputs (inputs are not modified; the record
        array is re-initialised from a device copy).  Returns device tensors."""
        t = self.torch
        subs_dev = h["subs0"].clone()
        out_i = t.empty(max(h["total"], 1), dtype=t.int64, device=f"cuda:{self.device}")
        out_e = t.empty(max(h["total"], 1), dtype=t.float32, device=f"cuda:{self.device}")
        t.cuda.synchronize(self.device)
        rc = self.lib.hbt_unbind_batch(self.ctx, ctypes.byref(h["ps"]), subs_dev.data_ptr(),
                                       h["subs_host"].ctypes.data, len(h["subs_host"]),
                                       out_i.data_ptr(), out_e.data_ptr(), 1 if plain else 0)
        check(rc, "hbt_unbind_batch")
        return subs_dev, out_i, out_e

    def unbind_batch(self, cat, plain=False):
        """RefineParticles on a catalogue.  Returns (subs, out_index, out_energy)
        as numpy arrays in the layout of oracle.unbind_batch."""
        h = self.upload(cat)
        subs_dev, out_i, out_e = self.unbind_device(h, plain=plain)
        subs = subs_dev.cpu().numpy().view(SUB_DTYPE).copy()
        total = h["total"]
        return subs, out_i.cpu().numpy()[:total].copy(), out_e.cpu().numpy()[:total].copy()

    # post-unbind properties -------------------------------------------------
    def properties_device(self, h, out, props_in=None):
        """CalculateProfileProperties + CalculateShape on the outputs of
        unbind_device (out = (records, out_index, out_energy), device); props_in:
        numpy array of PROPS_DTYPE (in/out fields) or None.  Returns numpy."""
        t = self.torch
        subs_dev, out_i, _ = out
        n = len(h["subs_host"])
        pin = (np.zeros(n, abi.PROPS_DTYPE) if props_in is None
               else np.ascontiguousarray(props_in, abi.PROPS_DTYPE))
        pd = self._dev(pin.view(np.uint8))
        t.cuda.synchronize(self.device)
        rc = self.lib.hbt_subhalo_properties(self.ctx, ctypes.byref(h["ps"]), subs_dev.data_ptr(), n,
                                             out_i.data_ptr(), pd.data_ptr())
        check(rc, "hbt_subhalo_properties")
        return pd.cpu().numpy().view(abi.PROPS_DTYPE).copy()

    def unbind_properties(self, cat, props_in=None, plain=False):
        """unbind_batch followed by the properties: (subs, out_index, out_energy, props)."""
        h = self.upload(cat)
        out = self.unbind_device(h, plain=plain)
        props = self.properties_device(h, out, props_in)
        subs = out[0].cpu().numpy().view(SUB_DTYPE).copy()
        total = h["total"]
        return subs, out[1].cpu().numpy()[:total].copy(), out[2].cpu().numpy()[:total].copy(), props

    # exclusive ownership inside host groups ----------------------------------
    def mask_subhalos(self, cols, subs, host_group, visit_rank=None):
        """MaskSubhalos on host columns/records: returns (out_list laid out
        like the columns, new counts), numpy.  visit_rank: sibling order
        (None: index order)."""
        t = self.torch
        c = {k: self._dev(np.ascontiguousarray(cols[k])) for k in PARTICLE_COLUMNS}
        ps = particles_struct(c, ptr=lambda x: x.data_ptr())
        subs = np.ascontiguousarray(subs, SUB_DTYPE)
        grp = np.ascontiguousarray(host_group, np.int64)
        n = len(cols["id"])
        out = t.full((max(n, 1),), -1, dtype=t.int64, device=f"cuda:{self.device}")
        cnt = np.zeros(len(subs), np.int64)
        t.cuda.synchronize(self.device)
        vr = None if visit_rank is None else np.ascontiguousarray(visit_rank, np.int64)
        rc = self.lib.hbt_mask_subhalos(self.ctx, ctypes.byref(ps), subs.ctypes.data, grp.ctypes.data,
                                        None if vr is None else vr.ctypes.data, len(subs), out.data_ptr(),
                                        cnt.ctypes.data)
        check(rc, "hbt_mask_subhalos")
        return out.cpu().numpy()[:n].copy(), cnt

    def nest_subhalos(self, host_group, is_central, parent, rank):
        """NestSubhalos' LevelUpDetachedSubhalos: returns (parent, rank) after
        the detached subhalos were levelled up, numpy."""
        t = self.torch
        g = self._dev(np.ascontiguousarray(host_group, np.int64))
        c = self._dev(np.ascontiguousarray(is_central, np.int32))
        p = self._dev(np.ascontiguousarray(parent, np.int64))
        r = self._dev(np.ascontiguousarray(rank, np.int32))
        t.cuda.synchronize(self.device)
        check(self.lib.hbt_nest_subhalos(self.ctx, len(host_group), g.data_ptr(), c.data_ptr(), p.data_ptr(),
                                         r.data_ptr()), "hbt_nest_subhalos")
        return p.cpu().numpy(), r.cpu().numpy()

    def upload_merge(self, cols, list_index):
        """Particle columns and the lists (column positions) to the device."""
        c = {k: self._dev(np.ascontiguousarray(cols[k])) for k in PARTICLE_COLUMNS}
        return dict(cols=c, ps=particles_struct(c, ptr=lambda x: x.data_ptr()),
                    li=self._dev(np.ascontiguousarray(list_index, np.int64)), n_list=len(list_index))

    def merge_device(self, hm, subs, host_group, merge_trapped=True, out=None, visit_rank=None):
        """hbt_merge_subhalos on uploaded columns and lists.  Returns (records
        after merging, the device output tensor, entries used)."""
        t = self.torch
        subs = np.ascontiguousarray(subs, SUB_DTYPE).copy()
        grp = np.ascontiguousarray(host_group, np.int64)
        cap = int(subs["count"].sum()) + len(subs)
        if out is None or out.numel() < cap:
            out = t.full((max(cap, 1),), -1, dtype=t.int64, device=f"cuda:{self.device}")
        used = ctypes.c_int64()
        t.cuda.synchronize(self.device)
        vr = None if visit_rank is None else np.ascontiguousarray(visit_rank, np.int64)
        rc = self.lib.hbt_merge_subhalos(self.ctx, ctypes.byref(hm["ps"]), hm["li"].data_ptr(), hm["n_list"],
                                         subs.ctypes.data, grp.ctypes.data, None if vr is None else vr.ctypes.data,
                                         len(subs),
                                         1 if merge_trapped else 0, out.data_ptr(), out.numel(),
                                         ctypes.byref(used))
        check(rc, "hbt_merge_subhalos")
        return subs, out, used.value

    def merge_subhalos(self, cols, list_index, subs, host_group, merge_trapped=True, visit_rank=None):
        """MergeSubhalos after unbinding.  cols: host particle columns;
        list_index: the lists as column positions (subs offset/count index
        it); subs: post-unbinding records.  Returns (records after merging,
        their lists as column positions laid out by the records), numpy."""
        hm = self.upload_merge(cols, list_index)
        subs, out, used = self.merge_device(hm, subs, host_group, merge_trapped, visit_rank=visit_rank)
        return subs, out.cpu().numpy()[:used].copy()

    # execution tuning (results do not depend on it) -----------------------
    def get_tuning(self):
        t = np.zeros(1, abi.TUNING_DTYPE)
        check(self.lib.hbt_get_tuning(self.ctx, t.ctypes.data), "hbt_get_tuning")
        return t[0]

    def set_tuning(self, t):
        t = np.asarray(t, abi.TUNING_DTYPE).reshape(1)
        check(self.lib.hbt_set_tuning(self.ctx, t.ctypes.data), "hbt_set_tuning")

    def autotune(self, h, reps=2):
        """Time the batch on device-resident inputs under candidate tunings;
        the fastest is set on the context and returned."""
        t = self.torch
        subs_dev = h["subs0"].clone()
        out_i = t.empty(max(h["total"], 1), dtype=t.int64, device=f"cuda:{self.device}")
        out_e = t.empty(max(h["total"], 1), dtype=t.float32, device=f"cuda:{self.device}")
        t.cuda.synchronize(self.device)
        chosen = np.zeros(1, abi.TUNING_DTYPE)
        rc = self.lib.hbt_autotune(self.ctx, ctypes.byref(h["ps"]), subs_dev.data_ptr(),
                                   h["subs_host"].ctypes.data, len(h["subs_host"]), out_i.data_ptr(),
                                   out_e.data_ptr(), int(reps), chosen.ctypes.data)
        check(rc, "hbt_autotune")
        return chosen[0]

    # snapshot + ID lists (lists resolved on the device) ---------------------
    def upload_snapshot(self, snap_cols, list_ids, subs):
        """Snapshot columns, the concatenated source-ID lists and the records
        (offset/count index the ID list) to the device."""
        cols = {k: self._dev(snap_cols[k]) for k in PARTICLE_COLUMNS}
        subs = subs.copy()
        total = plan_batch(subs)
        return dict(cols=cols, ids=self._dev(np.asarray(list_ids, np.int64)), subs_host=subs,
                    total=total, subs0=self._dev(subs.view(np.uint8)),
                    ps=particles_struct(cols, ptr=lambda t: t.data_ptr()))

    def subset_handle(self, h, sel):
        """A handle over the records sel of h (same device columns and lists;
        the output layout is planned for the subset alone)."""
        subs = h["subs_host"][np.asarray(sel)].copy()
        subs["parent"] = -1
        total = plan_batch(subs)
        out = dict(h)
        out.update(subs_host=subs, total=total, subs0=self._dev(subs.view(np.uint8)))
        return out

    def resolve_ids(self, h):
        """ID -> snapshot position for every list entry (hbt_id_to_index)."""
        t = self.torch
        n = len(h["ids"])
        out = t.empty(max(n, 1), dtype=t.int64, device=f"cuda:{self.device}")
        t.cuda.synchronize(self.device)
        check(self.lib.hbt_id_to_index(self.ctx, len(h["cols"]["id"]), h["cols"]["id"].data_ptr(), n,
                                       h["ids"].data_ptr(), out.data_ptr()), "hbt_id_to_index")
        return out

    def unbind_indexed(self, h, list_index, plain=False):
        t = self.torch
        subs_dev = h["subs0"].clone()
        out_i = t.empty(max(h["total"], 1), dtype=t.int64, device=f"cuda:{self.device}")
        out_e = t.empty(max(h["total"], 1), dtype=t.float32, device=f"cuda:{self.device}")
        t.cuda.synchronize(self.device)
        check(self.lib.hbt_unbind_batch_indexed(self.ctx, ctypes.byref(h["ps"]), list_index.data_ptr(),
                                                len(h["ids"]), subs_dev.data_ptr(),
                                                h["subs_host"].ctypes.data, len(h["subs_host"]),
                                                out_i.data_ptr(), out_e.data_ptr(), 1 if plain else 0),
              "hbt_unbind_batch_indexed")
        return subs_dev, out_i, out_e

    def unbind_positions(self, cols, positions, subs, plain=False):
        """hbt_unbind_batch_indexed on host inputs: snapshot columns, the lists
        as snapshot positions (negative: null particles, which the call drops
        as KickNullParticles does) and records whose offset/count index them.
        Returns (records, out_index, out_energy) as numpy arrays."""
        t = self.torch
        c = {k: self._dev(np.ascontiguousarray(cols[k])) for k in PARTICLE_COLUMNS}
        ps = particles_struct(c, ptr=lambda x: x.data_ptr())
        li = self._dev(np.ascontiguousarray(positions, np.int64))
        subs = np.ascontiguousarray(subs, SUB_DTYPE).copy()
        total = plan_batch(subs)
        sd = self._dev(subs.view(np.uint8))
        out_i = t.empty(max(total, 1), dtype=t.int64, device=f"cuda:{self.device}")
        out_e = t.empty(max(total, 1), dtype=t.float32, device=f"cuda:{self.device}")
        t.cuda.synchronize(self.device)
        check(self.lib.hbt_unbind_batch_indexed(self.ctx, ctypes.byref(ps), li.data_ptr(), len(positions),
                                                sd.data_ptr(), subs.ctypes.data, len(subs), out_i.data_ptr(),
                                                out_e.data_ptr(), 1 if plain else 0), "hbt_unbind_batch_indexed")
        return (sd.cpu().numpy().view(SUB_DTYPE).copy(), out_i.cpu().numpy()[:total].copy(),
                out_e.cpu().numpy()[:total].copy())

    def last_stats(self):
        ms = ctypes.c_double()
        ev = ctypes.c_int64()
        ps = ctypes.c_int64()
        check(self.lib.hbt_last_batch_stats(self.ctx, ctypes.byref(ms), ctypes.byref(ev),
                                            ctypes.byref(ps)), "hbt_last_batch_stats")
        return ms.value, ev.value, ps.value

    def profile(self, mode):
        check(self.lib.hbt_profile(self.ctx, mode), "hbt_profile")

    def profile_read(self):
        ms = ctypes.c_double()
        vals = [ctypes.c_int64() for _ in range(7)]
        check(self.lib.hbt_profile_read(self.ctx, ctypes.byref(ms), *[ctypes.byref(v) for v in vals]),
              "hbt_profile_read")
        return dict(walk_ms=ms.value, walk_launches=vals[0].value, interactions=vals[1].value,
                    opened=vals[2].value, targets=vals[3].value, lane_slots=vals[4].value,
                    tree_items=vals[5].value, wave_steps=vals[6].value)

    def profile_sorts(self):
        ms = ctypes.c_double()
        vals = [ctypes.c_int64() for _ in range(3)]
        check(self.lib.hbt_profile_sorts(self.ctx, ctypes.byref(ms), *[ctypes.byref(v) for v in vals]),
              "hbt_profile_sorts")
        return dict(sort_ms=ms.value, sort_calls=vals[0].value, elements=vals[1].value,
                    element_levels=vals[2].value)

    # per-subhalo form with host buffers --------------------------------------
    def unbind_one(self, cols: dict, sub: np.ndarray):
        cols = {k: np.ascontiguousarray(v) for k, v in cols.items()}
        rec = np.ascontiguousarray(sub, SUB_DTYPE).copy()
        n = int(rec["count"][0])
        out_i = np.zeros(max(n, 1), np.int64)
        out_e = np.zeros(max(n, 1), np.float32)
        ps = particles_struct(cols)
        check(self.lib.hbt_unbind_one(self.ctx, ctypes.byref(ps), rec.ctypes.data,
                                      out_i.ctypes.data, out_e.ctypes.data), "hbt_unbind_one")
        return rec, out_i[:n], out_e[:n]

    # tree / direct / hash ---------------------------------------------------
    def tree_potential(self, src, tgt, tmass=None):
        f = lambda a: self._dev(np.asarray(a, np.float32))  # noqa: E731
        sx, sy, sz, sm = (f(a) for a in src)
        tx, ty, tz = (f(a) for a in tgt)
        tm = f(tmass) if tmass is not None else None
        out = self.torch.empty(len(tgt[0]), dtype=self.torch.float64, device=f"cuda:{self.device}")
        self.torch.cuda.synchronize(self.device)
        check(self.lib.hbt_tree_potential(self.ctx, len(src[0]), sx.data_ptr(), sy.data_ptr(),
                                          sz.data_ptr(), sm.data_ptr(), len(tgt[0]), tx.data_ptr(),
                                          ty.data_ptr(), tz.data_ptr(),
                                          tm.data_ptr() if tm is not None else None,
                                          out.data_ptr()), "hbt_tree_potential")
        return out.cpu().numpy()

    def direct_potential(self, seg_offset, x, y, z, m, mode):
        f = lambda a: self._dev(np.asarray(a, np.float32))  # noqa: E731
        seg = self._dev(np.asarray(seg_offset, np.int64))
        xs, ys, zs, ms = f(x), f(y), f(z), f(m)
        out = self.torch.empty(max(len(x), 1), dtype=self.torch.float64, device=f"cuda:{self.device}")
        self.torch.cuda.synchronize(self.device)
        check(self.lib.hbt_direct_potential(self.ctx, len(seg_offset) - 1, seg.data_ptr(),
                                            xs.data_ptr(), ys.data_ptr(), zs.data_ptr(),
                                            ms.data_ptr(), mode, out.data_ptr()),
              "hbt_direct_potential")
        return out.cpu().numpy()[:len(x)]

    def id_to_index(self, table_ids, query_ids):
        t = self._dev(np.asarray(table_ids, np.int64))
        q = self._dev(np.asarray(query_ids, np.int64))
        out = self.torch.empty(max(len(query_ids), 1), dtype=self.torch.int64,
                               device=f"cuda:{self.device}")
        self.torch.cuda.synchronize(self.device)
        check(self.lib.hbt_id_to_index(self.ctx, len(table_ids), t.data_ptr(), len(query_ids),
                                       q.data_ptr(), out.data_ptr()), "hbt_id_to_index")
        return out.cpu().numpy()[:len(query_ids)]
