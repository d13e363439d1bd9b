// ref_harness.cpp -- drives the HBT+ reference's own unbinding code.
//
// TEST INFRASTRUCTURE ONLY, built in the development container by
// oracle/Makefile (target `ref`) against the reference sources where they lie
// (subhalo_unbind.cpp, gravity_tree.cpp, config_parser.cpp, mymath.cpp,
// subhalo.cpp, snapshot.cpp, mpi_wrapper.cpp); output goes to oracle/_ref/.
// It pins the CPU restatement (oracle/hbt_oracle.cpp) and writes the golden
// fixtures under tests/golden/.  No MPI call is made on this path.
//
// Usage:  hbt_ref_harness unbind|time|props|levels|kick|tree <in.bin> <out.bin>
// Binary formats are produced/consumed by oracle/refrun.py.
//
// `unbind` mirrors SubhaloSnapshot_t::RefineParticles (src/subhalo_unbind.cpp:478-515):
// every top-level subhalo in index order through Subhalo_t::RecursiveUnbind,
// then TruncateSource on every subhalo.  Run with OMP_NUM_THREADS=1.
// `levels` unbinds the same subhalos level by level (by nest height, index
// order within a level), each with the steps RecursiveUnbind takes for it
// (src/subhalo_unbind.cpp:432-447: its nested subhalos' unbound tails
// appended, orphans unbound on their own list): the order in which the
// library and the oracle consume the process's rand() stream when sampling
// (MaxSampleSizeOfPotentialEstimate > 0, src/subhalo_unbind.cpp:302).
// `kick` first drops null particles (Id -1, an ID the snapshot did not
// resolve) from every list with Subhalo_t::KickNullParticles (src/subhalo.cpp:446-477,
// as ParticleExchanger_t::RestoreParticles does, src/particle_exchanger.h:239-240),
// the record's nbound_prev as the Nbound it adjusts, then runs `unbind`.
// `tree` builds GravityTree_t over the sources and calls EvaluatePotential.

#include <omp.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "subhalo.h"
#include "gravity_tree.h"
#include "config_parser.h"
#include "../include/hbt_unbind.h"

namespace {

struct Epoch : public Snapshot_t {
  HBTxyz dummy{{0.f, 0.f, 0.f}};
  HBTInt size() const { return 0; }
  const HBTxyz &GetComovingPosition(const HBTInt) const { return dummy; }
  const HBTxyz &GetPhysicalVelocity(const HBTInt) const { return dummy; }
  HBTReal GetMass(const HBTInt) const { return 0.f; }
};

struct Cols : public Snapshot_t {
  std::vector<HBTxyz> pos;
  std::vector<HBTReal> mass;
  HBTInt size() const { return (HBTInt)pos.size(); }
  const HBTxyz &GetComovingPosition(const HBTInt i) const { return pos[i]; }
  const HBTxyz &GetPhysicalVelocity(const HBTInt i) const { return pos[i]; }
  HBTReal GetMass(const HBTInt i) const { return mass[i]; }
};

void configure(const hbt_params &p, Snapshot_t &epoch) {
  HBTConfig.SofteningHalo = p.softening;
  HBTConfig.BoxSize = p.box_size;
  HBTConfig.PeriodicBoundaryOn = p.periodic != 0;
  HBTConfig.MinNumPartOfSub = p.min_num_part_of_sub;
  HBTConfig.BoundMassPrecision = p.bound_mass_precision;
  HBTConfig.SourceSubRelaxFactor = p.source_sub_relax_factor;
  HBTConfig.TreeNodeOpenAngle = p.open_angle;
  HBTConfig.MaxSampleSizeOfPotentialEstimate = p.max_sample_size;
  HBTConfig.RefineMostboundParticle = p.refine_mostbound_particle != 0;
  // derived parameters exactly as Parameter_t::ParseConfigFile computes them
  HBTConfig.BoxHalf = HBTConfig.BoxSize / 2.;
  HBTConfig.TreeNodeResolution = HBTConfig.SofteningHalo * 0.1;
  HBTConfig.TreeNodeResolutionHalf = HBTConfig.TreeNodeResolution / 2.;
  HBTConfig.TreeNodeOpenAngleSquare = HBTConfig.TreeNodeOpenAngle * HBTConfig.TreeNodeOpenAngle;
  PhysicalConst::G = p.G;
  PhysicalConst::H0 = 0.1f;
  HBTConfig.MinSnapshotIndex = 0;
  HBTConfig.MaxSnapshotIndex = 1 << 20;
  epoch.SetSnapshotIndex(p.snapshot_index);
  epoch.Cosmology.ScaleFactor = p.scale_factor;
  epoch.Cosmology.Hz = p.hz;
  epoch.Cosmology.OmegaM0 = 0.3f;
  epoch.Cosmology.OmegaLambda0 = 0.7f;
  epoch.Cosmology.OmegaZ = 0.3f;
}

template <class T>
void rd(FILE *f, T *p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}
template <class T>
void wr(FILE *f, const T *p, size_t n) {
  if (n && fwrite(p, sizeof(T), n, f) != n) { fprintf(stderr, "short write\n"); exit(2); }
}

// mode 0: unbind and write the outputs; 1: time the nest-parallel unbinding;
// 2: as 0, then CalculateProfileProperties + CalculateShape of every subhalo
// (src/subhalo_tracking.cpp:904-905), with hbt_properties in and out;
// 3: as 0 with the subhalos unbound level by level; 4: as 0 after
// KickNullParticles on every subhalo.
int run_unbind(FILE *in, FILE *out, int mode) {
  const bool timed = mode == 1;
  hbt_params p;
  int64_t nsub, npart;
  rd(in, &p, 1); rd(in, &nsub, 1); rd(in, &npart, 1);
  std::vector<int64_t> id(npart);
  std::vector<float> x(npart), y(npart), z(npart), vx(npart), vy(npart), vz(npart), m(npart);
  std::vector<int32_t> type(npart);
  rd(in, id.data(), npart);
  rd(in, x.data(), npart); rd(in, y.data(), npart); rd(in, z.data(), npart);
  rd(in, vx.data(), npart); rd(in, vy.data(), npart); rd(in, vz.data(), npart);
  rd(in, m.data(), npart); rd(in, type.data(), npart);
  std::vector<hbt_subhalo> rec(nsub);
  rd(in, rec.data(), nsub);
  std::vector<hbt_properties> props(mode == 2 ? nsub : 0);
  if (mode == 2) rd(in, props.data(), nsub);

  Epoch epoch;
  configure(p, epoch);
  std::unordered_map<int64_t, int64_t> where;
  where.reserve(npart * 2);
  for (int64_t i = 0; i < npart; i++) where[id[i]] = i;

  SubhaloList_t subs(nsub);
  std::vector<int64_t> roots;
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t &S = subs[s];
    const hbt_subhalo &r = rec[s];
    S.Particles.resize(r.count);
    for (int64_t i = 0; i < r.count; i++) {
      int64_t g = r.offset + i;
      Particle_t &q = S.Particles[i];
      q.Id = id[g];
      q.ComovingPosition = {{x[g], y[g], z[g]}};
      q.PhysicalVelocity = {{vx[g], vy[g], vz[g]}};
      q.Mass = m[g];
      q.Type = (ParticleType_t)type[g];
    }
    S.Nbound = r.nbound_prev;
    S.SinkTrackId = r.sink_track_id;
    S.SnapshotIndexOfDeath = r.snapshot_index_of_death;
    S.SnapshotIndexOfSink = r.snapshot_index_of_sink;
    for (int j = 0; j < 3; j++) {
      S.ComovingAveragePosition[j] = r.ref_pos[j];
      S.PhysicalAverageVelocity[j] = r.ref_vel[j];
      S.ComovingMostBoundPosition[j] = r.mostbound_pos[j];
      S.PhysicalMostBoundVelocity[j] = r.mostbound_vel[j];
    }
    S.SpecificSelfPotentialEnergy = r.specific_self_potential_energy;
    S.SpecificSelfKineticEnergy = r.specific_self_kinetic_energy;
    for (int j = 0; j < 3; j++) S.SpecificAngularMomentum[j] = r.specific_angular_momentum[j];
    if (r.parent >= 0) subs[r.parent].NestedSubhalos.push_back(s);
    else roots.push_back(s);
  }
  if (timed) {
    // RefineParticles' halo-parallel layout (src/subhalo_unbind.cpp:479-513):
    // the nests are independent, one per OpenMP iteration, dynamic schedule
    const int64_t nroots = (int64_t)roots.size();
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < nroots; i++) subs[roots[i]].RecursiveUnbind(subs, epoch);
    const double secs = omp_get_wtime() - t0;
    wr(out, &secs, 1);
    return 0;
  }
  if (mode == 4)
    for (int64_t s = 0; s < nsub; s++) subs[s].KickNullParticles();
  if (mode == 3) {
    std::vector<int> height(nsub, -1);
    std::vector<int64_t> stack;
    int hmax = 0;
    for (int64_t r : roots) {  // heights by a post-order walk
      stack.push_back(r);
      while (!stack.empty()) {
        const int64_t s = stack.back();
        bool ready = true;
        for (HBTInt c : subs[s].NestedSubhalos)
          if (height[c] < 0) { stack.push_back(c); ready = false; }
        if (!ready) continue;
        int h = 0;
        for (HBTInt c : subs[s].NestedSubhalos) h = std::max(h, height[c] + 1);
        height[s] = h;
        hmax = std::max(hmax, h);
        stack.pop_back();
      }
    }
    for (int h = 0; h <= hmax; h++)
      for (int64_t s = 0; s < nsub; s++) {
        if (height[s] != h) continue;
        Subhalo_t &S = subs[s];
        const bool orphan = S.Nbound <= 1;
        Subhalo_t::ParticleList_t own;
        if (orphan) own = S.Particles;
        for (HBTInt c : S.NestedSubhalos) {
          const Subhalo_t &K = subs[c];
          S.Particles.insert(S.Particles.end(), K.Particles.begin() + K.Nbound, K.Particles.end());
        }
        if (orphan) S.Particles.swap(own);
        S.Unbind(epoch);
        if (orphan) S.Particles.swap(own);
      }
  } else {
    for (int64_t s : roots) subs[s].RecursiveUnbind(subs, epoch);
  }
  for (size_t s = 0; s < props.size(); s++) {
    Subhalo_t &S = subs[s];
    hbt_properties &q = props[s];
    S.BoundR200CritComoving = q.bound_r200crit_comoving;
    S.BoundM200Crit = q.bound_m200crit;
    S.LastMaxVmaxPhysical = q.last_max_vmax_physical;
    S.SnapshotIndexOfLastMaxVmax = q.snapshot_index_of_last_max_vmax;
    S.CalculateProfileProperties(epoch);
    S.CalculateShape();
    q.rmax_comoving = S.RmaxComoving;
    q.vmax_physical = S.VmaxPhysical;
    q.rhalf_comoving = S.RHalfComoving;
    q.r2sigma_comoving = S.R2SigmaComoving;
    q.bound_r200crit_comoving = S.BoundR200CritComoving;
    q.bound_m200crit = S.BoundM200Crit;
    q.last_max_vmax_physical = S.LastMaxVmaxPhysical;
    q.snapshot_index_of_last_max_vmax = S.SnapshotIndexOfLastMaxVmax;
    for (int j = 0; j < 6; j++) {
      q.inertial_tensor[j] = S.InertialTensor[j];
      q.inertial_tensor_weighted[j] = S.InertialTensorWeighted[j];
    }
  }

  std::vector<int64_t> list_len(nsub);
  int64_t total = 0;
  for (int64_t s = 0; s < nsub; s++) {
    list_len[s] = (int64_t)subs[s].Particles.size();
    total += list_len[s];
  }
  wr(out, &nsub, 1);
  wr(out, &total, 1);
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t &S = subs[s];
    hbt_subhalo r = rec[s];
    r.nbound = S.Nbound;
    r.list_len = list_len[s];
    r.mbound = S.Mbound;
    for (int t = 0; t < HBT_TYPE_MAX; t++) { r.nbound_type[t] = S.NboundType[t]; r.mbound_type[t] = S.MboundType[t]; }
    for (int j = 0; j < 3; j++) {
      r.ref_pos[j] = S.ComovingAveragePosition[j];
      r.ref_vel[j] = S.PhysicalAverageVelocity[j];
      r.mostbound_pos[j] = S.ComovingMostBoundPosition[j];
      r.mostbound_vel[j] = S.PhysicalMostBoundVelocity[j];
      r.specific_angular_momentum[j] = S.SpecificAngularMomentum[j];
    }
    r.specific_self_potential_energy = S.SpecificSelfPotentialEnergy;
    r.specific_self_kinetic_energy = S.SpecificSelfKineticEnergy;
    r.snapshot_index_of_death = S.SnapshotIndexOfDeath;
    r.snapshot_index_of_sink = S.SnapshotIndexOfSink;
    r.sink_track_id = S.SinkTrackId;
    std::vector<int64_t> idx(list_len[s]);
    for (int64_t i = 0; i < list_len[s]; i++) idx[i] = where[S.Particles[i].Id];
    std::vector<float> en(list_len[s], __builtin_nanf(""));
    for (size_t i = 0; i < S.Energies.size() && i < (size_t)list_len[s]; i++) en[i] = S.Energies[i];
    S.TruncateSource();
    r.nsource = (int64_t)S.Particles.size();
    wr(out, &r, 1);
    wr(out, idx.data(), idx.size());
    wr(out, en.data(), en.size());
  }
  if (!props.empty()) wr(out, props.data(), nsub);
  return 0;
}

int run_tree(FILE *in, FILE *out) {
  hbt_params p;
  int64_t ns, nt;
  rd(in, &p, 1); rd(in, &ns, 1); rd(in, &nt, 1);
  std::vector<float> sx(ns), sy(ns), sz(ns), sm(ns), tx(nt), ty(nt), tz(nt), tm(nt);
  rd(in, sx.data(), ns); rd(in, sy.data(), ns); rd(in, sz.data(), ns); rd(in, sm.data(), ns);
  rd(in, tx.data(), nt); rd(in, ty.data(), nt); rd(in, tz.data(), nt); rd(in, tm.data(), nt);
  Epoch epoch;
  configure(p, epoch);
  Cols src;
  src.Cosmology = epoch.Cosmology;
  src.pos.resize(ns);
  src.mass.assign(sm.begin(), sm.end());
  for (int64_t i = 0; i < ns; i++) src.pos[i] = {{sx[i], sy[i], sz[i]}};
  GravityTree_t tree;
  tree.Reserve(ns);
  tree.Build(src);
  std::vector<double> phi(nt);
  for (int64_t i = 0; i < nt; i++) {
    HBTxyz t = {{tx[i], ty[i], tz[i]}};
    phi[i] = tree.EvaluatePotential(t, tm[i]);
  }
  wr(out, phi.data(), nt);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: %s unbind|time|props|levels|kick|tree in.bin out.bin\n", argv[0]); return 1; }
  FILE *in = fopen(argv[2], "rb");
  FILE *out = fopen(argv[3], "wb");
  if (!in || !out) { fprintf(stderr, "cannot open files\n"); return 1; }
  int rc = 1;
  if (!strcmp(argv[1], "unbind")) rc = run_unbind(in, out, 0);
  else if (!strcmp(argv[1], "time")) rc = run_unbind(in, out, 1);
  else if (!strcmp(argv[1], "props")) rc = run_unbind(in, out, 2);
  else if (!strcmp(argv[1], "levels")) rc = run_unbind(in, out, 3);
  else if (!strcmp(argv[1], "kick")) rc = run_unbind(in, out, 4);
  else if (!strcmp(argv[1], "tree")) rc = run_tree(in, out);
  fclose(in);
  fclose(out);
  return rc;
}
