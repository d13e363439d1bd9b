// Drives the C++ host mirror (hbtplus_amd/host/hbt_host.hpp) the way a
// reference caller would: fills Subhalo_t objects with AoS Particle_t lists,
// calls RefineParticles (or Unbind per subhalo), writes the results.
//
// usage: hbt_host_test <in.bin> <out.bin> <mode: batch|plain|unbind|merge>
// in.bin : hbt_params, int64 nsub, int64 npart, columns (id, x..vz, mass, type),
//          hbt_subhalo records (offset/count/parent/nbound_prev/frames/markers)
//          merge: then int64 HostHaloId[nsub]; RefineParticles followed by
//          MergeSubhalos
// out.bin: per subhalo: int64 nbound, int64 nparticles, float mbound,
//          float ref_pos[3], float ref_vel[3], int64 SinkTrackId,
//          int32 SnapshotIndexOfDeath, int32 SnapshotIndexOfSink, then the
//          particle IDs in order
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../hbtplus_amd/host/hbt_host.hpp"

template <class T>
static void rd(FILE *f, T *p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: %s in out batch|plain|unbind\n", argv[0]); return 1; }
  FILE *in = fopen(argv[1], "rb");
  if (!in) return 1;
  hbt_params p;
  int64_t nsub, n;
  rd(in, &p, 1); rd(in, &nsub, 1); rd(in, &n, 1);
  std::vector<int64_t> id(n);
  std::vector<float> c[7];
  for (auto &v : c) v.resize(n);
  std::vector<int32_t> type(n);
  rd(in, id.data(), n);
  for (auto &v : c) rd(in, v.data(), n);
  rd(in, type.data(), n);
  std::vector<hbt_subhalo> rec(nsub);
  rd(in, rec.data(), nsub);
  const bool merge = !strcmp(argv[3], "merge");
  std::vector<int64_t> host(merge ? nsub : 0);
  if (merge) rd(in, host.data(), nsub);
  fclose(in);

  hbt::UnbindConfig cfg;
  cfg.G = p.G; cfg.SofteningHalo = p.softening; cfg.TreeNodeOpenAngle = p.open_angle;
  cfg.BoxSize = p.box_size; cfg.PeriodicBoundaryOn = p.periodic != 0;
  cfg.BoundMassPrecision = p.bound_mass_precision; cfg.MinNumPartOfSub = p.min_num_part_of_sub;
  cfg.SourceSubRelaxFactor = p.source_sub_relax_factor; cfg.precision = p.precision;
  cfg.MaxSampleSizeOfPotentialEstimate = p.max_sample_size;
  cfg.RefineMostboundParticle = p.refine_mostbound_particle != 0;
  hbt::Cosmology_t cosmo;
  cosmo.ScaleFactor = p.scale_factor;
  cosmo.Hz = p.hz;
  std::vector<hbt::Subhalo_t> subs(nsub);
  for (int64_t s = 0; s < nsub; s++) {
    hbt::Subhalo_t &S = subs[s];
    const hbt_subhalo &r = rec[s];
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
    for (int64_t i = 0; i < r.count; i++) {
      const int64_t g = r.offset + i;
      hbt::Particle_t q;
      q.Id = id[g];
      q.ComovingPosition = {{c[0][g], c[1][g], c[2][g]}};
      q.PhysicalVelocity = {{c[3][g], c[4][g], c[5][g]}};
      q.Mass = c[6][g];
      q.Type = type[g];
      S.Particles.push_back(q);
    }
    if (r.parent >= 0) subs[r.parent].NestedSubhalos.push_back(s);
  }
  try {
    hbt::Unbinder U(cfg, cosmo, p.snapshot_index);
    if (!strcmp(argv[3], "unbind")) {
      for (auto &S : subs) U.Unbind(S);
    } else {
      U.RefineParticles(subs, !strcmp(argv[3], "plain"));
      if (merge) U.MergeSubhalos(subs, host, true);
    }
  } catch (const hbt::Error &e) {
    fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 3;
  }
  FILE *out = fopen(argv[2], "wb");
  for (auto &S : subs) {
    int64_t np = (int64_t)S.Particles.size();
    fwrite(&S.Nbound, 8, 1, out);
    fwrite(&np, 8, 1, out);
    fwrite(&S.Mbound, 4, 1, out);
    fwrite(S.ComovingAveragePosition.data(), 4, 3, out);
    fwrite(S.PhysicalAverageVelocity.data(), 4, 3, out);
    fwrite(&S.SinkTrackId, 8, 1, out);
    fwrite(&S.SnapshotIndexOfDeath, 4, 1, out);
    fwrite(&S.SnapshotIndexOfSink, 4, 1, out);
    for (auto &q : S.Particles) fwrite(&q.Id, 8, 1, out);
  }
  fclose(out);
  return 0;
}
