// hbt_host.hpp -- C++ host mirror of the reference's unbinding interface on
// top of the C-ABI in include/hbt_unbind.h.
//
// The types and member functions keep the reference's names, argument
// meaning and behaviour so that a caller written against HBT+ reads the same:
//   Particle_t                      src/snapshot.h:51-72 (40-byte AoS record)
//   Subhalo_t::Unbind               src/subhalo.h:111, src/subhalo_unbind.cpp:263-431
//   Subhalo_t::RecursiveUnbind      src/subhalo.h:112, src/subhalo_unbind.cpp:432-447
//   Subhalo_t::TruncateSource       src/subhalo.h:114, src/subhalo_unbind.cpp:449-458
//   RefineParticles                 src/subhalo.h:245, src/subhalo_unbind.cpp:460-516
//   SubhaloSnapshot_t::MergeSubhalos
//                                   src/subhalo.h:247, src/subhalo_merge.cpp:136-262
//   GravityTree_t::Build / EvaluatePotential / BindingEnergy
//                                   src/oct_tree.h:62-63, src/gravity_tree.h:15-16
// Particle lists stay AoS on the host (as in the reference) and are converted
// to device columns at the boundary.  Errors from the library are thrown as
// hbt::Error (the reference has no error path; callers may catch it).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hbt_unbind.h"

namespace hbt {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &w) : std::runtime_error(w), code(c) {}
};

inline void check(int rc, const char *what) {
  if (rc != HBT_OK) throw Error(rc, std::string(what) + ": " + hbt_last_error());
}

typedef std::array<float, 3> HBTxyz;

struct Particle_t {
  int64_t Id;
  HBTxyz ComovingPosition;
  HBTxyz PhysicalVelocity;
  float Mass;
  int Type;
};
static_assert(sizeof(Particle_t) == 40, "Particle_t must keep the reference's 40-byte layout");

struct Cosmology_t {
  float ScaleFactor = 1.f;
  float Hz = 0.f;
};

// HBTConfig fields read by the unbinding, with the reference's defaults
// (src/config_parser.h:68-110).  With MaxSampleSizeOfPotentialEstimate > 0
// lists longer than the sample are shuffled with the rand() stream the
// context keeps (srand(1) at creation, as a process that never seeds;
// Unbinder::SetRandState sets it), full-pass trees use the sample, and
// RefineMostboundParticle re-orders the most bound entries.
struct UnbindConfig {
  float G = 43007.1f;
  float SofteningHalo = 0.5f;
  float TreeNodeOpenAngle = 0.45f;
  float BoxSize = 1e5f;
  bool PeriodicBoundaryOn = true;
  float BoundMassPrecision = 0.995f;
  int MinNumPartOfSub = 20;
  float SourceSubRelaxFactor = 3.f;
  int MaxSampleSizeOfPotentialEstimate = 1000;
  bool RefineMostboundParticle = true;
  int precision = HBT_PREC_FAST;
  hbt_params params(const Cosmology_t &c, int snapshot_index) const {
    hbt_params p{};
    p.G = G;
    p.softening = SofteningHalo;
    p.open_angle = TreeNodeOpenAngle;
    p.box_size = BoxSize;
    p.periodic = PeriodicBoundaryOn ? 1 : 0;
    p.bound_mass_precision = BoundMassPrecision;
    p.min_num_part_of_sub = MinNumPartOfSub;
    p.source_sub_relax_factor = SourceSubRelaxFactor;
    p.max_sample_size = MaxSampleSizeOfPotentialEstimate;
    p.refine_mostbound_particle = RefineMostboundParticle ? 1 : 0;
    p.scale_factor = c.ScaleFactor;
    p.hz = c.Hz;
    p.snapshot_index = snapshot_index;
    p.precision = precision;
    return p;
  }
};

struct Subhalo_t {
  int64_t TrackId = -1;
  int64_t Nbound = 0;
  float Mbound = 0.f;
  int64_t NboundType[HBT_TYPE_MAX] = {0};
  float MboundType[HBT_TYPE_MAX] = {0};
  int SnapshotIndexOfDeath = -1;
  int SnapshotIndexOfSink = -1;
  int64_t SinkTrackId = -1;
  float SpecificSelfPotentialEnergy = 0.f;
  float SpecificSelfKineticEnergy = 0.f;
  float SpecificAngularMomentum[3] = {0, 0, 0};
  HBTxyz ComovingAveragePosition{{0, 0, 0}};
  HBTxyz PhysicalAverageVelocity{{0, 0, 0}};
  HBTxyz ComovingMostBoundPosition{{0, 0, 0}};
  HBTxyz PhysicalMostBoundVelocity{{0, 0, 0}};
  std::vector<Particle_t> Particles;
  std::vector<float> Energies;           // SAVE_BINDING_ENERGY: E of the bound particles
  std::vector<int64_t> NestedSubhalos;   // indices of nested subhalos

  bool IsAlive() const { return SnapshotIndexOfDeath == -1; }
  bool IsTrapped() const { return SinkTrackId != -1; }
  void TruncateSource(const UnbindConfig &cfg) {
    int64_t n = Nbound <= 1 ? Nbound : (int64_t)((float)Nbound * cfg.SourceSubRelaxFactor);
    if (n > (int64_t)Particles.size()) n = (int64_t)Particles.size();
    Particles.resize(n);
  }
};

// A context (HIP stream + device buffers) and the calls that replace the
// reference's unbinding loop.
class Unbinder {
 public:
  Unbinder(const UnbindConfig &cfg, const Cosmology_t &cosmo, int snapshot_index, int device = 0)
      : cfg_(cfg), cosmo_(cosmo), snap_(snapshot_index) {
    hbt_params p = cfg_.params(cosmo_, snap_);
    check(hbt_ctx_create(&p, device, &ctx_), "hbt_ctx_create");
  }
  ~Unbinder() { hbt_ctx_destroy(ctx_); }
  Unbinder(const Unbinder &) = delete;
  Unbinder &operator=(const Unbinder &) = delete;

  // The rand() stream the sampling shuffles draw from: srand(seed) followed
  // by `draws` calls (a caller that calls rand() elsewhere keeps it in step).
  void SetRandState(uint32_t seed, int64_t draws) { check(hbt_set_rand_state(ctx_, seed, draws), "hbt_set_rand_state"); }

  // Subhalo_t::Unbind: one subhalo, host buffers, list permuted in place.
  void Unbind(Subhalo_t &sub) {
    std::vector<Subhalo_t *> one{&sub};
    std::vector<int64_t> parent{-1};
    run(one, parent, /*plain=*/true, /*truncate=*/false, /*recursive=*/false);
  }

  // SubhaloSnapshot_t::RefineParticles: every subhalo whose index is not in
  // another's NestedSubhalos is unbound with RecursiveUnbind semantics (nested
  // subhalos first, their unbound tails appended to the host), then
  // TruncateSource on all.  plain=true drives the GPU one subhalo at a time.
  void RefineParticles(std::vector<Subhalo_t> &subs, bool plain = false) {
    std::vector<Subhalo_t *> ptr(subs.size());
    std::vector<int64_t> parent(subs.size(), -1);
    for (size_t i = 0; i < subs.size(); i++) {
      ptr[i] = &subs[i];
      for (int64_t c : subs[i].NestedSubhalos) {
        if (c < 0 || c >= (int64_t)subs.size()) throw Error(HBT_ERR_INVALID_ARG, "bad nested index");
        parent[c] = (int64_t)i;
      }
    }
    run(ptr, parent, plain, /*truncate=*/true, /*recursive=*/true);
  }

  // SubhaloSnapshot_t::MergeSubhalos after RefineParticles: trap detection
  // along the (glued) nests, MergeTo in post-order from every group's central,
  // re-unbinding and TruncateSource of the hosts that received subhalos.
  // HostHaloId: the subhalos' host halos (-1 for field subhalos); a nested
  // subhalo shares its host's halo.  GroupRank: each subhalo's place in
  // MemberTable.SubGroups of its halo (most massive first, the central at 0);
  // empty: index order.  Nested subhalos go in NestedSubhalos order.
  void MergeSubhalos(std::vector<Subhalo_t> &subs, const std::vector<int64_t> &HostHaloId,
                     bool MergeTrappedSubhalos = true, const std::vector<int64_t> &GroupRank = {}) {
    const int64_t nsub = (int64_t)subs.size();
    if ((int64_t)HostHaloId.size() != nsub) throw Error(HBT_ERR_INVALID_ARG, "one host halo per subhalo");
    if (!GroupRank.empty() && (int64_t)GroupRank.size() != nsub)
      throw Error(HBT_ERR_INVALID_ARG, "one group rank per subhalo");
    std::vector<int64_t> parent(nsub, -1), VisitRank(nsub);
    for (int64_t i = 0; i < nsub; i++) VisitRank[i] = GroupRank.empty() ? i : GroupRank[i];
    for (int64_t i = 0; i < nsub; i++)
      for (size_t k = 0; k < subs[i].NestedSubhalos.size(); k++) {
        const int64_t c = subs[i].NestedSubhalos[k];
        if (c >= 0 && c < nsub) VisitRank[c] = (int64_t)k;
      }
    for (int64_t i = 0; i < nsub; i++)
      for (int64_t c : subs[i].NestedSubhalos) {
        if (c < 0 || c >= nsub) throw Error(HBT_ERR_INVALID_ARG, "bad nested index");
        parent[c] = i;
      }
    int64_t n = 0;
    for (auto &S : subs) n += (int64_t)S.Particles.size();
    std::vector<int64_t> id(n), lst(n);
    std::vector<float> col[7];
    for (auto &c : col) c.resize(n);
    std::vector<int32_t> type(n);
    std::vector<const Particle_t *> src(n);
    std::vector<hbt_subhalo> rec(nsub);
    int64_t off = 0;
    for (int64_t s = 0; s < nsub; s++) {
      const Subhalo_t &S = subs[s];
      hbt_subhalo &r = rec[s];
      std::memset(&r, 0, sizeof(r));
      r.offset = off;
      r.count = (int64_t)S.Particles.size();
      r.parent = parent[s];
      r.nbound_prev = S.Nbound;
      r.nbound = S.Nbound;
      r.mbound = S.Mbound;
      for (int t = 0; t < HBT_TYPE_MAX; t++) { r.nbound_type[t] = S.NboundType[t]; r.mbound_type[t] = S.MboundType[t]; }
      r.sink_track_id = S.SinkTrackId;
      r.snapshot_index_of_death = S.SnapshotIndexOfDeath;
      r.snapshot_index_of_sink = S.SnapshotIndexOfSink;
      for (int j = 0; j < 3; j++) {
        r.ref_pos[j] = S.ComovingAveragePosition[j];
        r.ref_vel[j] = S.PhysicalAverageVelocity[j];
        r.mostbound_pos[j] = S.ComovingMostBoundPosition[j];
        r.mostbound_vel[j] = S.PhysicalMostBoundVelocity[j];
        r.specific_angular_momentum[j] = S.SpecificAngularMomentum[j];
      }
      r.specific_self_potential_energy = S.SpecificSelfPotentialEnergy;
      r.specific_self_kinetic_energy = S.SpecificSelfKineticEnergy;
      for (const Particle_t &p : S.Particles) {
        id[off] = p.Id;
        for (int j = 0; j < 3; j++) { col[j][off] = p.ComovingPosition[j]; col[3 + j][off] = p.PhysicalVelocity[j]; }
        col[6][off] = p.Mass;
        type[off] = p.Type;
        src[off] = &p;
        lst[off] = off;
        off++;
      }
    }
    const int64_t cap = n + nsub;
    DevArray<int64_t> d_id(n), d_lst(n), d_out(cap);
    DevArray<float> d_col[7] = {DevArray<float>(n), DevArray<float>(n), DevArray<float>(n), DevArray<float>(n),
                                DevArray<float>(n), DevArray<float>(n), DevArray<float>(n)};
    DevArray<int32_t> d_type(n);
    d_id.put(id);
    d_lst.put(lst);
    for (int i = 0; i < 7; i++) d_col[i].put(col[i]);
    d_type.put(type);
    hbt_particles P{n, d_id.p, d_col[0].p, d_col[1].p, d_col[2].p, d_col[3].p, d_col[4].p, d_col[5].p,
                    d_col[6].p, d_type.p};
    int64_t used = 0;
    check(hbt_merge_subhalos(ctx_, &P, d_lst.p, n, rec.data(), HostHaloId.data(),
                             VisitRank.empty() ? nullptr : VisitRank.data(), nsub, MergeTrappedSubhalos ? 1 : 0,
                             d_out.p, cap, &used),
          "hbt_merge_subhalos");
    std::vector<int64_t> out(used);
    d_out.get(out);
    std::vector<std::vector<Particle_t>> lists(nsub);
    for (int64_t s = 0; s < nsub; s++) {
      lists[s].resize(rec[s].count);
      for (int64_t i = 0; i < rec[s].count; i++) lists[s][i] = *src[out[rec[s].offset + i]];
    }
    for (int64_t s = 0; s < nsub; s++) {
      Subhalo_t &S = subs[s];
      const hbt_subhalo &r = rec[s];
      S.Particles.swap(lists[s]);
      S.Nbound = r.nbound;
      S.Mbound = r.mbound;
      for (int t = 0; t < HBT_TYPE_MAX; t++) { S.NboundType[t] = r.nbound_type[t]; S.MboundType[t] = r.mbound_type[t]; }
      S.SnapshotIndexOfDeath = r.snapshot_index_of_death;
      S.SnapshotIndexOfSink = r.snapshot_index_of_sink;
      S.SinkTrackId = r.sink_track_id;
      S.SpecificSelfPotentialEnergy = r.specific_self_potential_energy;
      S.SpecificSelfKineticEnergy = r.specific_self_kinetic_energy;
      for (int j = 0; j < 3; j++) {
        S.SpecificAngularMomentum[j] = r.specific_angular_momentum[j];
        S.ComovingAveragePosition[j] = r.ref_pos[j];
        S.PhysicalAverageVelocity[j] = r.ref_vel[j];
        S.ComovingMostBoundPosition[j] = r.mostbound_pos[j];
        S.PhysicalMostBoundVelocity[j] = r.mostbound_vel[j];
      }
    }
  }

  hbt_ctx *context() { return ctx_; }

 private:
  template <class T>
  struct DevArray {
    T *p = nullptr;
    explicit DevArray(size_t n) {
      if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
        throw Error(HBT_ERR_OUT_OF_MEMORY, "hipMalloc");
    }
    ~DevArray() { (void)hipFree(p); }
    void put(const std::vector<T> &v) {
      if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
        throw Error(HBT_ERR_DEVICE, "hipMemcpy");
    }
    void get(std::vector<T> &v) {
      if (!v.empty() && hipMemcpy(v.data(), p, v.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess)
        throw Error(HBT_ERR_DEVICE, "hipMemcpy");
    }
  };

  void run(std::vector<Subhalo_t *> &subs, const std::vector<int64_t> &parent, bool plain, bool truncate,
           bool recursive) {
    const int64_t nsub = (int64_t)subs.size();
    // particle columns: every subhalo's own list, in order
    int64_t n = 0;
    for (auto *s : subs) n += (int64_t)s->Particles.size();
    std::vector<int64_t> id(n);
    std::vector<float> col[7];
    for (auto &c : col) c.resize(n);
    std::vector<int32_t> type(n);
    std::vector<hbt_subhalo> rec(nsub);
    std::vector<const Particle_t *> src(n);
    int64_t off = 0;
    for (int64_t s = 0; s < nsub; s++) {
      Subhalo_t &S = *subs[s];
      hbt_subhalo &r = rec[s];
      std::memset(&r, 0, sizeof(r));
      r.offset = off;
      r.count = (int64_t)S.Particles.size();
      r.parent = parent[s];
      // the orphan rule (Nbound <= 1) is part of RecursiveUnbind only
      r.nbound_prev = recursive ? S.Nbound : INT64_MAX;
      r.sink_track_id = S.SinkTrackId;
      r.snapshot_index_of_death = S.SnapshotIndexOfDeath;
      r.snapshot_index_of_sink = S.SnapshotIndexOfSink;
      for (int j = 0; j < 3; j++) {
        r.ref_pos[j] = S.ComovingAveragePosition[j];
        r.ref_vel[j] = S.PhysicalAverageVelocity[j];
        r.mostbound_pos[j] = S.ComovingMostBoundPosition[j];
        r.mostbound_vel[j] = S.PhysicalMostBoundVelocity[j];
        r.specific_angular_momentum[j] = S.SpecificAngularMomentum[j];
      }
      r.specific_self_potential_energy = S.SpecificSelfPotentialEnergy;
      r.specific_self_kinetic_energy = S.SpecificSelfKineticEnergy;
      for (const Particle_t &p : S.Particles) {
        id[off] = p.Id;
        for (int j = 0; j < 3; j++) { col[j][off] = p.ComovingPosition[j]; col[3 + j][off] = p.PhysicalVelocity[j]; }
        col[6][off] = p.Mass;
        type[off] = p.Type;
        src[off] = &p;
        off++;
      }
    }
    int64_t total = 0;
    check(hbt_plan_batch(rec.data(), nsub, &total), "hbt_plan_batch");
    DevArray<int64_t> d_id(n);
    DevArray<float> d_col[7] = {DevArray<float>(n), DevArray<float>(n), DevArray<float>(n), DevArray<float>(n),
                                DevArray<float>(n), DevArray<float>(n), DevArray<float>(n)};
    DevArray<int32_t> d_type(n);
    DevArray<hbt_subhalo> d_rec(nsub);
    DevArray<int64_t> d_out(total);
    DevArray<float> d_en(total);
    d_id.put(id);
    for (int i = 0; i < 7; i++) d_col[i].put(col[i]);
    d_type.put(type);
    d_rec.put(rec);
    hbt_particles P{n, d_id.p, d_col[0].p, d_col[1].p, d_col[2].p, d_col[3].p, d_col[4].p, d_col[5].p,
                    d_col[6].p, d_type.p};
    check(hbt_unbind_batch(ctx_, &P, d_rec.p, rec.data(), nsub, d_out.p, d_en.p, plain ? 1 : 0),
          "hbt_unbind_batch");
    std::vector<hbt_subhalo> out(nsub);
    std::vector<int64_t> idx(total);
    std::vector<float> en(total);
    d_rec.get(out);
    d_out.get(idx);
    d_en.get(en);
    // write back: the reference's Subhalo_t fields and permuted particle lists
    std::vector<std::vector<Particle_t>> lists(nsub);
    for (int64_t s = 0; s < nsub; s++) {
      const hbt_subhalo &r = out[s];
      lists[s].resize(r.list_len);
      for (int64_t i = 0; i < r.list_len; i++) lists[s][i] = *src[idx[r.out_offset + i]];
    }
    for (int64_t s = 0; s < nsub; s++) {
      Subhalo_t &S = *subs[s];
      const hbt_subhalo &r = out[s];
      S.Particles.swap(lists[s]);
      S.Nbound = r.nbound;
      S.Mbound = r.mbound;
      for (int t = 0; t < HBT_TYPE_MAX; t++) { S.NboundType[t] = r.nbound_type[t]; S.MboundType[t] = r.mbound_type[t]; }
      S.SnapshotIndexOfDeath = r.snapshot_index_of_death;
      S.SnapshotIndexOfSink = r.snapshot_index_of_sink;
      S.SinkTrackId = r.sink_track_id;
      S.SpecificSelfPotentialEnergy = r.specific_self_potential_energy;
      S.SpecificSelfKineticEnergy = r.specific_self_kinetic_energy;
      for (int j = 0; j < 3; j++) {
        S.SpecificAngularMomentum[j] = r.specific_angular_momentum[j];
        S.ComovingAveragePosition[j] = r.ref_pos[j];
        S.PhysicalAverageVelocity[j] = r.ref_vel[j];
        S.ComovingMostBoundPosition[j] = r.mostbound_pos[j];
        S.PhysicalMostBoundVelocity[j] = r.mostbound_vel[j];
      }
      S.Energies.clear();
      if (r.status & HBT_STATUS_CONVERGED)
        S.Energies.assign(en.begin() + r.out_offset, en.begin() + r.out_offset + r.nbound);
      if (truncate) S.TruncateSource(cfg_);
    }
  }

  UnbindConfig cfg_;
  Cosmology_t cosmo_;
  int snap_;
  hbt_ctx *ctx_ = nullptr;
};

// GravityTree_t: Build over a particle list, then EvaluatePotential /
// BindingEnergy at arbitrary targets (batched on the device).
class GravityTree_t {
 public:
  GravityTree_t(Unbinder &u, const UnbindConfig &cfg, const Cosmology_t &cosmo)
      : ctx_(u.context()), cfg_(cfg), cosmo_(cosmo) {}
  void Build(const std::vector<Particle_t> &parts) { src_ = parts; }
  // potentials of many targets at once; target_mass != 0 removes the self term
  std::vector<double> EvaluatePotential(const std::vector<HBTxyz> &targets,
                                        const std::vector<float> &target_mass = {}) {
    const int64_t ns = (int64_t)src_.size(), nt = (int64_t)targets.size();
    std::vector<float> s[4], t[4];
    for (auto &v : s) v.resize(ns);
    for (auto &v : t) v.resize(nt);
    for (int64_t i = 0; i < ns; i++) {
      for (int j = 0; j < 3; j++) s[j][i] = src_[i].ComovingPosition[j];
      s[3][i] = src_[i].Mass;
    }
    for (int64_t i = 0; i < nt; i++) {
      for (int j = 0; j < 3; j++) t[j][i] = targets[i][j];
      t[3][i] = target_mass.empty() ? 0.f : target_mass[i];
    }
    std::vector<float *> d(8, nullptr);
    double *dphi = nullptr;
    auto cleanup = [&]() { for (float *p : d) (void)hipFree(p); (void)hipFree(dphi); };
    for (int k = 0; k < 4; k++) {
      if (hipMalloc(&d[k], std::max<int64_t>(ns, 1) * 4) != hipSuccess ||
          hipMalloc(&d[4 + k], std::max<int64_t>(nt, 1) * 4) != hipSuccess) { cleanup(); throw Error(HBT_ERR_OUT_OF_MEMORY, "hipMalloc"); }
      (void)hipMemcpy(d[k], s[k].data(), ns * 4, hipMemcpyHostToDevice);
      (void)hipMemcpy(d[4 + k], t[k].data(), nt * 4, hipMemcpyHostToDevice);
    }
    if (hipMalloc(&dphi, std::max<int64_t>(nt, 1) * 8) != hipSuccess) { cleanup(); throw Error(HBT_ERR_OUT_OF_MEMORY, "hipMalloc"); }
    hbt_params p = cfg_.params(cosmo_, 0);
    hbt_ctx_set_params(ctx_, &p);
    int rc = hbt_tree_potential(ctx_, ns, d[0], d[1], d[2], d[3], nt, d[4], d[5], d[6], d[7], dphi);
    std::vector<double> phi(nt);
    if (rc == HBT_OK) (void)hipMemcpy(phi.data(), dphi, nt * 8, hipMemcpyDeviceToHost);
    cleanup();
    check(rc, "hbt_tree_potential");
    return phi;
  }
  // BindingEnergy = 0.5 |v - v_ref + Hz a (x - x_ref)|^2 + phi (src/gravity_tree.cpp:166-175)
  std::vector<double> BindingEnergy(const std::vector<Particle_t> &targets, const HBTxyz &refPos,
                                    const HBTxyz &refVel, bool targets_in_tree) {
    std::vector<HBTxyz> pos(targets.size());
    std::vector<float> m(targets.size(), 0.f);
    for (size_t i = 0; i < targets.size(); i++) {
      pos[i] = targets[i].ComovingPosition;
      if (targets_in_tree) m[i] = targets[i].Mass;
    }
    std::vector<double> e = EvaluatePotential(pos, m);
    for (size_t i = 0; i < targets.size(); i++) {
      float dv[3];
      for (int j = 0; j < 3; j++) {
        float dx = targets[i].ComovingPosition[j] - refPos[j];
        if (cfg_.PeriodicBoundaryOn) {
          const float h = (float)((double)cfg_.BoxSize / 2.);
          dx = dx > h ? dx - cfg_.BoxSize : (dx < -h ? dx + cfg_.BoxSize : dx);
        }
        dv[j] = targets[i].PhysicalVelocity[j] - refVel[j];
        dv[j] += cosmo_.Hz * cosmo_.ScaleFactor * dx;
      }
      e[i] += (double)(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]) * 0.5;
    }
    return e;
  }

 private:
  hbt_ctx *ctx_;
  UnbindConfig cfg_;
  Cosmology_t cosmo_;
  std::vector<Particle_t> src_;
};

}  // namespace hbt
