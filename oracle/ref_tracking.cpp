// ref_tracking.cpp -- drives the reference's own SubhaloSnapshot_t::MaskSubhalos
// (src/subhalo_tracking.cpp:826-841) and SubhaloSnapshot_t::MergeSubhalos
// (src/subhalo_merge.cpp:187-220) on catalogues in the layouts of
// hbt_mask_subhalos and hbt_merge_subhalos, so that the oracle restatement and
// the library can be checked against them.  TEST INFRASTRUCTURE ONLY: linked
// with every reference source except the HBT.cpp driver (oracle/Makefile
// target `ref`); a singleton MPI process, since SubhaloSnapshot_t builds its
// MPI datatypes on construction.  Run with OMP_NUM_THREADS=1.
//
//   ref_tracking mask|mask_mass in.bin out.bin
//     in : int64 nsub, int64 ncol; int64 id[ncol];
//          hbt_subhalo subs[nsub] (offset, count, parent, nbound_prev, mbound read);
//          int64 host_group[nsub]
//     out: int64 out_list[ncol] (column positions, laid out like the columns;
//          -1 where unused); int64 new_count[nsub]
//
//   ref_tracking merge|merge_mass in.bin out.bin
//     in : hbt_params; int64 merge_trapped; int64 nsub, ncol, nlist;
//          columns id[ncol] (int64), x y z vx vy vz mass [ncol] (float),
//          type[ncol] (int32); int64 list_index[nlist];
//          hbt_subhalo subs[nsub] (offset/count index list_index; parent,
//          nbound, mbound and its per-type parts, trap and death fields,
//          frames, most bound coordinates, kinematics);
//          int64 host_group[nsub]
//     out: hbt_subhalo subs[nsub] (merged state; offset/count index the list
//          below); int64 total; int64 list[total] (column positions);
//          float energy[total] (Energies of the bound block where the
//          subhalo was unbound again, NaN elsewhere)
//   ref_tracking nest in.bin out.bin
//     in : int64 nsub; int64 host_group[nsub]; int32 is_central[nsub];
//          int64 parent[nsub]; int32 rank[nsub]
//     out: int64 parent[nsub]; int32 rank[nsub] after LevelUpDetachedSubhalos
//
//   ref_tracking cores in.bin out.bin
//     in : hbt_params; int64 nsub, ncol, nlist; x y z vx vy vz mass [ncol]
//          (float); int64 list_index[nlist]; hbt_subhalo subs[nsub]
//          (offset/count index list_index; nbound read)
//     out: float[8 * nsub]: each subhalo's merging core (position[3],
//          velocity[3], sigma_R, sigma_V; zeros where nbound = 0)
//
//   ref_tracking merge_time in.bin out.bin
//     as merge; out: double seconds spent in MergeSubhalos (OMP_NUM_THREADS
//     threads, the reference's own parallel loops)
//
// The member table is built by the reference (MemberShipTable_t::Build, which
// orders each group by Mbound).  mask / merge: every subhalo gets an Mbound
// that decreases with its index and the group's first head the largest, so the
// group's central and the order of its heads are index order (the library's
// order without visit ranks).  mask_mass / merge_mass: the subhalos' own
// Mbound (the records' mbound) orders each group, the heads are collected in
// that order as NestSubhalos does, and the library is given that order as
// visit ranks.
#include <mpi.h>
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "config_parser.h"
#include "snapshot.h"
#include "halo.h"
// MaskSubhalos is a private member reached in the reference only through
// AssignHosts (which needs a halo catalogue); this translation unit alone sees
// the class with its members public (access control is not part of the
// member's symbol, so the call binds to the reference's compiled function).
#define private public
#include "subhalo.h"
#undef private
#include "../include/hbt_unbind.h"

// The merging helper of src/subhalo_merge.cpp:14-27, declared here with the
// same members so that the harness can call the reference's compiled
// SubHelper_t::BuildPosition / BuildVelocity.
struct SubHelper_t {
  HBTInt HostTrackId;
  bool IsMerged;
  HBTxyz ComovingPosition;
  HBTxyz PhysicalVelocity;
  float ComovingSigmaR;
  float PhysicalSigmaV;
  void BuildPosition(const Subhalo_t &sub);
  void BuildVelocity(const Subhalo_t &sub);
  SubHelper_t() : HostTrackId(-1), IsMerged(false) {}
};

namespace {

template <class T>
void rd(FILE *f, T *p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}
template <class T>
void wr(FILE *f, const T *p, size_t n) {
  if (n) fwrite(p, sizeof(T), n, f);
}

// The member table is built by MemberShipTable_t::Build, which orders each
// group by Mbound: every subhalo is given an Mbound that decreases with its
// index and the group's first head the largest while the table is built, so
// the group's central and the order of its heads are those of the library.
void build_members(SubhaloSnapshot_t &snap, const std::vector<int64_t> &grp,
                   const std::vector<std::vector<int64_t>> &heads) {
  const int64_t nsub = (int64_t)snap.Subhalos.size();
  std::vector<float> keep(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    keep[s] = snap.Subhalos[s].Mbound;
    snap.Subhalos[s].HostHaloId = grp[s];
    snap.Subhalos[s].Mbound = (float)(nsub - s);
  }
  for (const auto &h : heads)
    if (!h.empty()) snap.Subhalos[h[0]].Mbound = (float)(2 * nsub + 1);
  snap.MemberTable.Build((HBTInt)heads.size(), snap.Subhalos, true);
  snap.MemberTable.SubGroupsOfHeads.assign(heads.begin(), heads.end());
  for (int64_t s = 0; s < nsub; s++) snap.Subhalos[s].Mbound = keep[s];
}

// The member table as the reference has it: MemberShipTable_t::Build orders
// each group by the subhalos' own Mbound (CompareMass_t, src/subhalo_tracking.cpp:96-104)
// and the heads are collected from that order (the Rank 0 members,
// :676-686; Rank 0 here: no nesting host), so the central is the most massive
// member.  The library and the oracle are given the same order as visit ranks.
void build_members_by_mass(SubhaloSnapshot_t &snap, const std::vector<int64_t> &grp,
                           const std::vector<int64_t> &parent, int64_t nhalos) {
  const int64_t nsub = (int64_t)snap.Subhalos.size();
  for (int64_t s = 0; s < nsub; s++) snap.Subhalos[s].HostHaloId = grp[s];
  snap.MemberTable.Build((HBTInt)nhalos, snap.Subhalos, true);
  snap.MemberTable.SubGroupsOfHeads.assign(nhalos, std::vector<HBTInt>());
  for (int64_t g = 0; g < nhalos; g++)
    for (HBTInt s : snap.MemberTable.SubGroups[g])
      if (parent[s] < 0) snap.MemberTable.SubGroupsOfHeads[g].push_back(s);
}

int run_mask(FILE *in, FILE *out, bool by_mass) {
  int64_t nsub, ncol;
  rd(in, &nsub, 1); rd(in, &ncol, 1);
  std::vector<int64_t> id(ncol), grp(nsub);
  std::vector<hbt_subhalo> rec(nsub);
  rd(in, id.data(), ncol);
  rd(in, rec.data(), nsub);
  rd(in, grp.data(), nsub);
  int64_t nhalos = 0;
  for (int64_t g : grp) nhalos = std::max(nhalos, g + 1);
  SubhaloSnapshot_t snap;
  snap.Subhalos.resize(nsub);
  std::vector<std::vector<int64_t>> heads(nhalos);
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t &S = snap.Subhalos[s];
    S.Particles.resize(rec[s].count);
    for (int64_t i = 0; i < rec[s].count; i++) S.Particles[i].Id = id[rec[s].offset + i];
    S.Nbound = rec[s].nbound_prev;
    S.Mbound = rec[s].mbound;
    if (rec[s].parent >= 0) snap.Subhalos[rec[s].parent].NestedSubhalos.push_back(s);
    else if (grp[s] >= 0) heads[grp[s]].push_back(s);
  }
  if (by_mass) {
    std::vector<int64_t> parent(nsub);
    for (int64_t s = 0; s < nsub; s++) parent[s] = rec[s].parent;
    build_members_by_mass(snap, grp, parent, nhalos);
  } else {
    build_members(snap, grp, heads);
  }
  snap.MaskSubhalos();
  std::vector<int64_t> out_list(ncol, -1), cnt(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    // the kept particles are a subsequence of the list; each kept ID is its
    // first entry in the list
    const auto &P = snap.Subhalos[s].Particles;
    int64_t j = 0;
    for (size_t k = 0; k < P.size(); k++) {
      while (j < rec[s].count && id[rec[s].offset + j] != P[k].Id) j++;
      if (j == rec[s].count) { fprintf(stderr, "kept list is not a subsequence\n"); return 3; }
      out_list[rec[s].offset + k] = rec[s].offset + j;
      j++;
    }
    cnt[s] = (int64_t)P.size();
  }
  wr(out, out_list.data(), ncol);
  wr(out, cnt.data(), nsub);
  return 0;
}

void configure(const hbt_params &p, Snapshot_t &epoch) {
  HBTConfig.SofteningHalo = p.softening;
  HBTConfig.BoxSize = p.box_size;
  HBTConfig.PeriodicBoundaryOn = p.periodic != 0;
  HBTConfig.MinNumPartOfSub = p.min_num_part_of_sub;
  HBTConfig.BoundMassPrecision = p.bound_mass_precision;
  HBTConfig.SourceSubRelaxFactor = p.source_sub_relax_factor;
  HBTConfig.TreeNodeOpenAngle = p.open_angle;
  HBTConfig.MaxSampleSizeOfPotentialEstimate = p.max_sample_size;
  HBTConfig.RefineMostboundParticle = true;
  // derived parameters as Parameter_t::ParseConfigFile computes them
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

// The cores of the merging (SubHelper_t::BuildPosition / BuildVelocity) of
// every subhalo: position[3], velocity[3], sigma_R, sigma_V as floats.
int run_cores(FILE *in, FILE *out) {
  hbt_params p;
  int64_t nsub, ncol, nlist;
  rd(in, &p, 1); rd(in, &nsub, 1); rd(in, &ncol, 1); rd(in, &nlist, 1);
  std::vector<float> c[7];
  for (auto &v : c) v.resize(ncol);
  for (auto &v : c) rd(in, v.data(), ncol);
  std::vector<int64_t> list(nlist);
  rd(in, list.data(), nlist);
  std::vector<hbt_subhalo> rec(nsub);
  rd(in, rec.data(), nsub);
  SubhaloSnapshot_t snap;
  configure(p, snap);
  std::vector<float> o(8 * nsub, 0.f);
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t S;
    S.Particles.resize(rec[s].count);
    for (int64_t i = 0; i < rec[s].count; i++) {
      const int64_t g = list[rec[s].offset + i];
      S.Particles[i].ComovingPosition = {{c[0][g], c[1][g], c[2][g]}};
      S.Particles[i].PhysicalVelocity = {{c[3][g], c[4][g], c[5][g]}};
      S.Particles[i].Mass = c[6][g];
    }
    S.Nbound = rec[s].nbound;
    SubHelper_t h;
    h.ComovingPosition = {{0.f, 0.f, 0.f}};
    h.PhysicalVelocity = {{0.f, 0.f, 0.f}};
    h.BuildPosition(S);
    h.BuildVelocity(S);
    for (int j = 0; j < 3; j++) { o[8 * s + j] = h.ComovingPosition[j]; o[8 * s + 3 + j] = h.PhysicalVelocity[j]; }
    o[8 * s + 6] = h.ComovingSigmaR;
    o[8 * s + 7] = h.PhysicalSigmaV;
  }
  wr(out, o.data(), o.size());
  return 0;
}

// timed: only the wall seconds of MergeSubhalos (OpenMP as the reference runs
// it) are written
// LevelUpDetachedSubhalos (src/subhalo_tracking.cpp:735-769) on a snapshot
// whose member table has each group's flagged central first.
int run_nest(FILE *in, FILE *out) {
  int64_t nsub;
  rd(in, &nsub, 1);
  std::vector<int64_t> grp(nsub), parent(nsub);
  std::vector<int32_t> central(nsub), rank(nsub);
  rd(in, grp.data(), nsub);
  rd(in, central.data(), nsub);
  rd(in, parent.data(), nsub);
  rd(in, rank.data(), nsub);
  int64_t nhalos = 0;
  for (int64_t g : grp) nhalos = std::max(nhalos, g + 1);
  SubhaloSnapshot_t snap;
  snap.Subhalos.resize(nsub);
  std::vector<std::vector<int64_t>> heads(nhalos);
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t &S = snap.Subhalos[s];
    S.TrackId = s;
    S.Rank = rank[s];
    if (parent[s] >= 0) snap.Subhalos[parent[s]].NestedSubhalos.push_back(s);
  }
  // the flagged central of a group as SubGroups[g][0]
  for (int64_t s = 0; s < nsub; s++)
    if (central[s] && grp[s] >= 0) heads[grp[s]].insert(heads[grp[s]].begin(), s);
  build_members(snap, grp, heads);
  snap.LevelUpDetachedSubhalos();
  std::vector<int64_t> np(nsub, -1);
  std::vector<int32_t> nr(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    for (auto c : snap.Subhalos[s].NestedSubhalos) np[c] = s;
    nr[s] = snap.Subhalos[s].Rank;
  }
  wr(out, np.data(), nsub);
  wr(out, nr.data(), nsub);
  return 0;
}

int run_merge(FILE *in, FILE *out, bool timed, bool by_mass) {
  hbt_params p;
  int64_t flag, nsub, ncol, nlist;
  rd(in, &p, 1); rd(in, &flag, 1); rd(in, &nsub, 1); rd(in, &ncol, 1); rd(in, &nlist, 1);
  std::vector<int64_t> id(ncol);
  std::vector<float> c[7];
  for (auto &v : c) v.resize(ncol);
  std::vector<int32_t> type(ncol);
  rd(in, id.data(), ncol);
  for (auto &v : c) rd(in, v.data(), ncol);
  rd(in, type.data(), ncol);
  std::vector<int64_t> list(nlist), grp(nsub);
  rd(in, list.data(), nlist);
  std::vector<hbt_subhalo> rec(nsub);
  rd(in, rec.data(), nsub);
  rd(in, grp.data(), nsub);
  SubhaloSnapshot_t snap;
  configure(p, snap);
  HBTConfig.MergeTrappedSubhalos = flag != 0;
  std::unordered_map<int64_t, int64_t> where;
  where.reserve(ncol * 2);
  for (int64_t i = 0; i < ncol; i++) where[id[i]] = i;
  int64_t nhalos = 0;
  for (int64_t g : grp) nhalos = std::max(nhalos, g + 1);
  std::vector<std::vector<int64_t>> heads(nhalos);
  snap.Subhalos.resize(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t &S = snap.Subhalos[s];
    const hbt_subhalo &r = rec[s];
    S.TrackId = s;
    S.Particles.resize(r.count);
    for (int64_t i = 0; i < r.count; i++) {
      const int64_t g = list[r.offset + i];
      Particle_t &q = S.Particles[i];
      q.Id = id[g];
      q.ComovingPosition = {{c[0][g], c[1][g], c[2][g]}};
      q.PhysicalVelocity = {{c[3][g], c[4][g], c[5][g]}};
      q.Mass = c[6][g];
      q.Type = (ParticleType_t)type[g];
    }
    S.Nbound = r.nbound;
    S.Mbound = r.mbound;
    for (int t = 0; t < HBT_TYPE_MAX; t++) { S.NboundType[t] = r.nbound_type[t]; S.MboundType[t] = r.mbound_type[t]; }
    S.SinkTrackId = r.sink_track_id;
    S.SnapshotIndexOfDeath = r.snapshot_index_of_death;
    S.SnapshotIndexOfSink = r.snapshot_index_of_sink;
    for (int j = 0; j < 3; j++) {
      S.ComovingAveragePosition[j] = r.ref_pos[j];
      S.PhysicalAverageVelocity[j] = r.ref_vel[j];
      S.ComovingMostBoundPosition[j] = r.mostbound_pos[j];
      S.PhysicalMostBoundVelocity[j] = r.mostbound_vel[j];
      S.SpecificAngularMomentum[j] = r.specific_angular_momentum[j];
    }
    S.SpecificSelfPotentialEnergy = r.specific_self_potential_energy;
    S.SpecificSelfKineticEnergy = r.specific_self_kinetic_energy;
    if (r.parent >= 0) snap.Subhalos[r.parent].NestedSubhalos.push_back(s);
    else if (grp[s] >= 0) heads[grp[s]].push_back(s);
  }
  if (by_mass) {
    std::vector<int64_t> parent(nsub);
    for (int64_t s = 0; s < nsub; s++) parent[s] = rec[s].parent;
    build_members_by_mass(snap, grp, parent, nhalos);
  } else {
    build_members(snap, grp, heads);
  }
  const double t0 = omp_get_wtime();
  snap.MergeSubhalos();
  const double secs = omp_get_wtime() - t0;
  if (timed) {
    wr(out, &secs, 1);
    return 0;
  }
  int64_t total = 0;
  for (auto &S : snap.Subhalos) total += (int64_t)S.Particles.size();
  std::vector<int64_t> olist;
  std::vector<float> oen;
  olist.reserve(total);
  oen.reserve(total);
  for (int64_t s = 0; s < nsub; s++) {
    Subhalo_t &S = snap.Subhalos[s];
    hbt_subhalo &r = rec[s];
    r.offset = (int64_t)olist.size();
    r.count = (int64_t)S.Particles.size();
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
    for (size_t i = 0; i < S.Particles.size(); i++) {
      olist.push_back(where.at(S.Particles[i].Id));
      oen.push_back(i < S.Energies.size() ? S.Energies[i] : __builtin_nanf(""));
    }
  }
  wr(out, rec.data(), nsub);
  wr(out, &total, 1);
  wr(out, olist.data(), total);
  wr(out, oen.data(), total);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: %s mask|mask_mass|merge|merge_mass|merge_time|nest|cores in.bin out.bin\n", argv[0]); return 1; }
  MPI_Init(&argc, &argv);
  FILE *in = fopen(argv[2], "rb");
  FILE *out = fopen(argv[3], "wb");
  int rc = 1;
  if (!in || !out) fprintf(stderr, "cannot open files\n");
  else if (!strcmp(argv[1], "mask")) rc = run_mask(in, out, false);
  else if (!strcmp(argv[1], "mask_mass")) rc = run_mask(in, out, true);
  else if (!strcmp(argv[1], "merge")) rc = run_merge(in, out, false, false);
  else if (!strcmp(argv[1], "merge_mass")) rc = run_merge(in, out, false, true);
  else if (!strcmp(argv[1], "merge_time")) rc = run_merge(in, out, true, false);
  else if (!strcmp(argv[1], "nest")) rc = run_nest(in, out);
  else if (!strcmp(argv[1], "cores")) rc = run_cores(in, out);
  if (in) fclose(in);
  if (out) fclose(out);
  MPI_Finalize();
  return rc;
}
