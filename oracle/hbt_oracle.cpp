// hbt_oracle.cpp -- CPU restatement of HBT+ subhalo unbinding.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
// MI355X library: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load liboracle.so, and only as the checker or as the
// timed CPU baseline.  The product path (hbtplus_amd/, libhbt_unbind.so)
// never links, loads or calls it.
//
// It restates the reference algorithm step by step, including the
// float/double rounding of every expression, so that with one thread it is
// bit-identical to the reference built with g++ on x86-64 (no FMA):
//   serial-insertion octree          src/oct_tree.tpp:17-144
//   node moments and threading       src/gravity_tree.cpp:18-77
//   softened potential walk          src/gravity_tree.cpp:79-164
//   binding energy                   src/gravity_tree.cpp:166-175
//   relative velocity                src/snapshot.h:100-111
//   two-pointer partition            src/subhalo_unbind.cpp:21-58
//   frame averages                   src/subhalo_unbind.cpp:108-188
//   kinematics                       src/subhalo_unbind.cpp:189-232
//   Unbind iteration                 src/subhalo_unbind.cpp:263-431
//   RecursiveUnbind / TruncateSource src/subhalo_unbind.cpp:432-458
//   CountParticleTypes/CountParticles src/subhalo.cpp:400-444,479-499
//   derived parameters               src/config_parser.cpp:95-103
//   sampling: random_shuffle, SetMassUnit, RefineBindingEnergyOrder
//                                    src/subhalo_unbind.cpp:74-79,234-262,299-302,
//                                    335-339,385-388,406-407
// Third-party behaviour restated: POSIX drand48 (48-bit LCG, a=0x5DEECE66D,
// c=0xB; glibc starts an unseeded generator at state 0) used at
// src/oct_tree.tpp:118; glibc rand() (random(3)'s TYPE_3 additive feedback
// generator, an unseeded generator being srand(1)) and libstdc++
// std::random_shuffle (GCC 11 bits/stl_algo.h: for i = 1..n-1, swap a[i] with
// a[rand() % (i+1)]) used at src/subhalo_unbind.cpp:302; libstdc++ std::sort
// (called directly, same library as the reference build).
//
// Order of the rand() draws in a batch with sampling on: subhalos are unbound
// level by level (a level: the subhalos whose nested subhalos are all done,
// i.e. by nest height), in index order within a level.  Without nesting this
// is the index order of the reference's single-threaded RefineParticles loop;
// the reference harness (oracle/ref_harness.cpp, mode "levels") drives the
// reference's own Unbind in this order for the sampling fixtures.
//
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off -fopenmp, no -march).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>
#include <unordered_set>
#include <functional>
#include <map>
#include <omp.h>

#include "../include/hbt_unbind.h"

namespace {

typedef int64_t Int;

// ---- parameters with the reference's derived values ------------------------
struct Params {
  float G, eps, theta2, res, res_half, box, box_half, bmp, relax, a, hz;
  bool periodic;
  int min_part, snap;
  Int max_sample;      // MaxSampleSizeOfPotentialEstimate
  bool refine;         // RefineMostboundParticle, only with sampling (:264-265)
  explicit Params(const hbt_params &p) {
    G = p.G;
    eps = p.softening;
    theta2 = p.open_angle * p.open_angle;                  // float*float, config_parser.cpp:103
    res = (float)((double)p.softening * 0.1);             // :101
    res_half = (float)((double)res / 2.);                 // :102
    box = p.box_size;
    box_half = (float)((double)p.box_size / 2.);          // :100
    bmp = p.bound_mass_precision;
    relax = p.source_sub_relax_factor;
    a = p.scale_factor;
    hz = p.hz;
    periodic = p.periodic != 0;
    min_part = p.min_num_part_of_sub;
    snap = p.snapshot_index;
    max_sample = p.max_sample_size;
    refine = max_sample > 0 && p.refine_mostbound_particle != 0;
  }
  // NEAREST macro, config_parser.h:142, in float and in double
  inline float nearest(float x) const {
    return (x > box_half) ? (x - box) : ((x < -box_half) ? (x + box) : x);
  }
  inline double nearest(double x) const {
    return (x > (double)box_half) ? (x - (double)box)
                                  : ((x < -(double)box_half) ? (x + (double)box) : x);
  }
};

// ---- POSIX drand48 ------------------------------------------------------
// glibc leaves an unseeded generator at state 0 (the reference never seeds
// it); X <- (0x5DEECE66D X + 0xB) mod 2^48, result X / 2^48.
struct Drand48 {
  uint64_t x = 0;
  double next() {
    x = (0x5DEECE66DULL * x + 0xBULL) & ((1ULL << 48) - 1);
    return (double)x / 281474976710656.0;
  }
};

// ---- glibc rand() ----------------------------------------------------------
// random(3)'s TYPE_3 generator: srand(seed) sets r[0] = seed (1 when 0) and
// r[i] = 16807 r[i-1] mod (2^31 - 1) for i = 1..30 (Schrage's method in
// 32-bit words), r[i] = r[i-31] for i = 31..33; then r[i] = r[i-31] + r[i-3]
// mod 2^32.  The values r[34..343] are discarded and the k-th rand() call
// returns r[344+k] >> 1.  A process that never calls srand behaves as srand(1).
struct GlibcRand {
  uint32_t w[31];  // w[i % 31] = r[i] for the last 31 indices
  int64_t m = 0;   // next index of r
  int64_t draws = 0;
  explicit GlibcRand(uint32_t seed = 1) {
    int32_t r[34];
    int32_t word = (int32_t)(seed == 0 ? 1 : seed);
    r[0] = word;
    for (int i = 1; i < 31; i++) {
      const long hi = word / 127773, lo = word % 127773;
      word = (int32_t)(16807 * lo - 2836 * hi);
      if (word < 0) word += 2147483647;
      r[i] = word;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 3; i < 34; i++) w[i % 31] = (uint32_t)r[i];
    m = 34;
    for (int i = 0; i < 310; i++) step();
  }
  uint32_t step() {
    const int u = (int)(m % 31);
    w[u] += w[(u + 28) % 31];
    m++;
    return w[u];
  }
  int32_t next() {
    draws++;
    return (int32_t)(step() >> 1);
  }
  void skip(int64_t n) {
    for (int64_t i = 0; i < n; i++) next();
  }
};

// std::random_shuffle(first, last) of libstdc++ (GCC 11, bits/stl_algo.h):
// draws[i-1] is the rand() value of step i.
template <class T>
void random_shuffle_with(std::vector<T> &a, const int32_t *draws) {
  const Int n = (Int)a.size();
  for (Int i = 1; i < n; i++) {
    const Int j = (Int)draws[i - 1] % (i + 1);
    if (i != j) std::swap(a[i], a[j]);
  }
}

struct Vec3f { float v[3]; };

// ---- the reference octree, restated ----------------------------------------
class Tree {
 public:
  Int np = 0;                          // particles are nodes [0,np), cells [np, np+ncell)
  std::vector<Int> sons;               // 8 per cell, build phase
  std::vector<float> cs, clen, cmass;  // 3 per cell, 1, 1
  std::vector<Int> csib, cnext, pnext;
  std::vector<float> px, py, pz, pm;   // gathered sources in insertion order

  void build(const Params &P, Drand48 &rng) {
    Int n = np;
    sons.clear(); cs.clear(); clen.clear(); cmass.clear(); csib.clear(); cnext.clear();
    pnext.assign(n, -1);
    double xmin[3], xmax[3], Center[3], Len;
    const float *pos[3] = {px.data(), py.data(), pz.data()};
    for (int j = 0; j < 3; j++) xmin[j] = xmax[j] = pos[j][0];
    for (Int i = 1; i < n; i++)
      for (int j = 0; j < 3; j++) {
        if (pos[j][i] > xmax[j]) xmax[j] = pos[j][i];
        else if (pos[j][i] < xmin[j]) xmin[j] = pos[j][i];
      }
    Len = xmax[0] - xmin[0];
    for (int j = 1; j < 3; j++)
      if ((xmax[j] - xmin[j]) > Len) Len = xmax[j] - xmin[j];
    for (int j = 0; j < 3; j++) Center[j] = 0.5 * (xmax[j] + xmin[j]);
    double Lenhalf = 0.5 * Len;
    append_cell();
    const Int root = n;
    for (Int i = 0; i < n; i++) {
      Int node = root;
      double lenhalf = Lenhalf, c[3] = {Center[0], Center[1], Center[2]};
      while (true) {
        lenhalf *= 0.5;
        int sub = 0;
        for (int j = 0; j < 3; j++) {
          if (pos[j][i] > c[j]) { c[j] += lenhalf; sub += 1 << j; }
          else c[j] -= lenhalf;
        }
        Int subid = son(node, sub);
        if (subid < 0) { son(node, sub) = i; break; }
        if (subid < n) {
          Int nid = n + ncell();
          append_cell();
          son(node, sub) = nid;
          node = nid;
          if (lenhalf < (double)P.res_half) {
            sub = (int)(Int)(8.0 * rng.next());
            if (sub >= 8) sub = 7;
          } else {
            sub = 0;
            for (int j = 0; j < 3; j++)
              if (pos[j][subid] > c[j]) sub += 1 << j;
          }
          son(node, sub) = subid;
        } else {
          node = subid;
        }
      }
    }
    update_internal(P, root, -1, Len);
  }

  double potential(const Params &P, double tx, double ty, double tz, float target_mass) const {
    const double h = 2.8 * (double)P.eps, h_inv = 1.0 / h;
    double pot = (double)(float)(target_mass / P.eps);
    Int no = np;
    while (no >= 0) {
      double dx, dy, dz, mass, r2;
      if (no < np) {
        dx = (double)px[no] - tx; dy = (double)py[no] - ty; dz = (double)pz[no] - tz;
        if (P.periodic) { dx = P.nearest(dx); dy = P.nearest(dy); dz = P.nearest(dz); }
        mass = pm[no];
        no = pnext[no];
        r2 = dx * dx + dy * dy + dz * dz;
      } else {
        Int c = no - np;
        dx = (double)cs[3 * c] - tx; dy = (double)cs[3 * c + 1] - ty; dz = (double)cs[3 * c + 2] - tz;
        if (P.periodic) { dx = P.nearest(dx); dy = P.nearest(dy); dz = P.nearest(dz); }
        mass = cmass[c];
        r2 = dx * dx + dy * dy + dz * dz;
        float l2 = clen[c] * clen[c];
        if ((double)l2 > r2 * (double)P.theta2) { no = cnext[c]; continue; }
        no = csib[c];
      }
      double r = std::sqrt(r2);
      if (r >= h) {
        pot -= mass / r;
      } else {
        double u = r * h_inv, wp;
        if (u < 0.5)
          wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
        else
          wp = -3.2 + 0.066666666667 / u +
               u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
        pot += mass * h_inv * wp;
      }
    }
    return pot * (double)P.G / (double)P.a;
  }

  Int ncell() const { return (Int)clen.size(); }

 private:
  Int &son(Int node, int k) { return sons[8 * (node - np) + k]; }
  void append_cell() {
    for (int k = 0; k < 8; k++) sons.push_back(-1);
    cs.push_back(0); cs.push_back(0); cs.push_back(0);
    clen.push_back(0); cmass.push_back(0); csib.push_back(-1); cnext.push_back(-1);
  }
  void process(const Params &P, Int id, Int next, double &mass, double com[3], double len) {
    if (id < np) {
      double m = pm[id];
      mass += m;
      com[0] += px[id] * m; com[1] += py[id] * m; com[2] += pz[id] * m;
      pnext[id] = next;
    } else {
      if (len >= (double)P.res) update_internal(P, id, next, len / 2.);
      else update_internal(P, id, next, len);
      Int c = id - np;
      double m = cmass[c];
      mass += m;
      com[0] += cs[3 * c] * m; com[1] += cs[3 * c + 1] * m; com[2] += cs[3 * c + 2] * m;
    }
  }
  void update_internal(const Params &P, Int no, Int sib, double len) {
    Int s[8];
    Int c = no - np;
    for (int k = 0; k < 8; k++) s[k] = sons[8 * c + k];
    clen[c] = (float)len;
    csib[c] = sib;
    int i = 0;
    while (s[i] < 0) i++;
    Int pp = s[i];
    cnext[c] = pp;
    double mass = 0., com[3] = {0., 0., 0.};
    for (i++; i < 8; i++) {
      if (s[i] >= 0) {
        Int p = pp;
        pp = s[i];
        process(P, p, pp, mass, com, len);
      }
    }
    process(P, pp, sib, mass, com, len);
    cmass[c] = (float)mass;
    for (int j = 0; j < 3; j++) cs[3 * c + j] = (float)(com[j] / mass);
  }
};

// ---- the Unbind iteration --------------------------------------------------
struct PE { Int pid; float E; };
inline bool comp_energy(const PE &a, const PE &b) { return a.E < b.E; }

// two-pointer partition, src/subhalo_unbind.cpp:21-58
Int partition_energy(std::vector<PE> &El, Int len) {
  if (len == 0) return 0;
  if (len == 1) return El[0].E < 0;
  PE tmp = El[0];
  Int f = 0, b = len;
  while (true) {
    while (true) {
      b--;
      if (b == f) { El[f] = tmp; if (tmp.E < 0) b++; return b; }
      if (El[b].E < 0) break;
    }
    El[f] = El[b];
    while (true) {
      f++;
      if (f == b) { El[b] = tmp; if (tmp.E < 0) b++; return b; }
      if (El[f].E > 0) break;
    }
    El[b] = El[f];
  }
}

struct PartCols {  // host particle columns
  const hbt_particles *p;
  float x(Int g, int j) const { return j == 0 ? p->x[g] : (j == 1 ? p->y[g] : p->z[g]); }
  float v(Int g, int j) const { return j == 0 ? p->vx[g] : (j == 1 ? p->vy[g] : p->vz[g]); }
  float m(Int g) const { return p->mass[g]; }
  int t(Int g) const { return p->type ? p->type[g] : 1; }
};

// RelativeVelocity, src/snapshot.h:100-111 (all float)
inline void rel_vel(const Params &P, const float tp[3], const float tv[3], const float rp[3],
                    const float rv[3], float dv[3]) {
  for (int j = 0; j < 3; j++) {
    float dx = tp[j] - rp[j];
    if (P.periodic) dx = P.nearest(dx);
    dv[j] = tv[j] - rv[j];
    dv[j] += P.hz * P.a * dx;
  }
}
inline float vnorm(const float v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }
inline float vdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

struct Sub {
  hbt_subhalo *rec;
  std::vector<Int> list;        // Particles: global particle indices
  std::vector<float> energy;    // Elist E aligned with `list` after convergence
  Int nbound;
};

void count_types(const PartCols &C, Sub &S, bool skip_null) {
  hbt_subhalo &r = *S.rec;
  Int nb[HBT_TYPE_MAX] = {0};
  float mb[HBT_TYPE_MAX] = {0};
  for (Int i = 0; i < S.nbound; i++) {
    Int g = S.list[i];
    if (skip_null && C.p->id[g] == -1) continue;
    int t = C.t(g);
    nb[t]++;
    mb[t] += C.m(g);
  }
  for (int t = 0; t < HBT_TYPE_MAX; t++) { r.nbound_type[t] = nb[t]; r.mbound_type[t] = mb[t]; }
}

// draws: the rand() values of the list's shuffle (read only when the list is
// longer than MaxSampleSizeOfPotentialEstimate > 0: one per element after the
// first).
void unbind(const Params &P, const PartCols &C, Sub &S, Drand48 &rng, Tree &tree,
            const int32_t *draws = nullptr) {
  hbt_subhalo &r = *S.rec;
  const Int N = (Int)S.list.size();
  auto alive = [&]() { return r.snapshot_index_of_death == -1; };
  S.energy.assign(N, std::numeric_limits<float>::quiet_NaN());
  r.n_iter = 0; r.n_evals = 0; r.status = 0;
  if (N < P.min_part) {
    if (alive()) r.snapshot_index_of_death = P.snap;
  }
  if (N == 0 || N == 1) {  // early exits, CountParticles()
    S.nbound = N;
    count_types(C, S, false);
    float m = 0.f;
    for (int t = 0; t < HBT_TYPE_MAX; t++) m += r.mbound_type[t];
    r.mbound = m;
    r.status = HBT_STATUS_EARLY;
    return;
  }
  float *RefPos = r.ref_pos, *RefVel = r.ref_vel;
  float OldRefPos[3] = {0, 0, 0}, OldRefVel[3] = {0, 0, 0};
  const Int old_mb_id = C.p->id[S.list[0]];
  Int Nbound = N, Nlast = 0;
  const Int MS = P.max_sample;
  // shuffle for the sampled trees (:302)
  if (MS > 0 && N > MS) random_shuffle_with(S.list, draws);
  std::vector<PE> El(N);
  for (Int i = 0; i < N; i++) { El[i].pid = i; El[i].E = 0.f; }
  auto gpos = [&](Int i, float o[3]) { Int g = S.list[El[i].pid]; for (int j = 0; j < 3; j++) o[j] = C.x(g, j); };
  auto gvel = [&](Int i, float o[3]) { Int g = S.list[El[i].pid]; for (int j = 0; j < 3; j++) o[j] = C.v(g, j); };
  auto gmass = [&](Int i) { return C.m(S.list[El[i].pid]); };
  // tree over Elist[b, e) with the masses times mf (EnergySnapshot_t::GetMass
  // with MassFactor, :74-79)
  auto load_tree = [&](Int b, Int e, float mf) {
    Int n = e - b;
    tree.np = n;
    tree.px.resize(n); tree.py.resize(n); tree.pz.resize(n); tree.pm.resize(n);
    for (Int i = 0; i < n; i++) {
      float x[3]; gpos(b + i, x);
      tree.px[i] = x[0]; tree.py[i] = x[1]; tree.pz[i] = x[2]; tree.pm[i] = gmass(b + i) * mf;
    }
    tree.build(P, rng);
  };
  // BindingEnergy (src/gravity_tree.cpp:166-175) of Elist entry i in the frame
  auto binding_energy = [&](Int i, float tmass) {
    float x[3], v[3], dv[3];
    gpos(i, x); gvel(i, v);
    double pot = tree.potential(P, x[0], x[1], x[2], tmass);
    rel_vel(P, x, v, RefPos, RefVel, dv);
    return (float)((double)vnorm(dv) * 0.5 + pot);
  };
  bool correction = false;
  while (true) {
    r.n_iter++;
    if (correction) {
      float RefVelDiff[3];
      rel_vel(P, OldRefPos, OldRefVel, RefPos, RefVel, RefVelDiff);
      float dK = (float)(0.5 * (double)vnorm(RefVelDiff));
      load_tree(Nbound, Nlast, 1.f);
      for (Int i = 0; i < Nbound; i++) {
        float x[3], v[3], ov[3];
        gpos(i, x); gvel(i, v);
        rel_vel(P, x, v, OldRefPos, OldRefVel, ov);
        float kin = vdot(ov, RefVelDiff) + dK;
        double rhs = (double)kin - tree.potential(P, x[0], x[1], x[2], 0.f);
        El[i].E = (float)((double)El[i].E + rhs);
      }
      r.n_evals += Nbound;
      Nlast = Nbound;
    } else {
      Nlast = Nbound;
      // sampled tree over the first MaxSampleSize entries with their masses
      // scaled by Nlast / MaxSampleSize (:332-340); targets outside the
      // sample carry no self term
      Int np_tree = Nlast;
      float mf = 1.f;
      if (MS > 0 && Nlast > MS) {
        np_tree = MS;
        mf = (float)Nlast / (float)MS;
      }
      load_tree(0, np_tree, mf);
      for (Int i = 0; i < Nlast; i++) El[i].E = binding_energy(i, i < np_tree ? gmass(i) * mf : 0.f);
      r.n_evals += Nlast;
    }
    Nbound = partition_energy(El, Nlast);
    if (Nbound < P.min_part) {  // disruption
      Nbound = 1;
      Nlast = 1;
      if (alive()) r.snapshot_index_of_death = P.snap;
      for (Int i = 0; i < N; i++)
        if (C.p->id[S.list[i]] == old_mb_id) { std::swap(S.list[i], S.list[0]); break; }
      for (int j = 0; j < 3; j++) { RefPos[j] = r.mostbound_pos[j]; RefVel[j] = r.mostbound_vel[j]; }
      r.mbound = C.m(S.list[0]);
      r.status |= HBT_STATUS_DISRUPTED;
      break;
    }
    std::sort(El.begin() + Nbound, El.begin() + Nlast, comp_energy);
    Int Ndiff = Nlast - Nbound;
    if (Ndiff < Nbound && (MS <= 0 || Ndiff < MS)) {  // :383-389
      correction = true;
      for (int j = 0; j < 3; j++) { OldRefPos[j] = RefPos[j]; OldRefVel[j] = RefVel[j]; }
    }
    // AverageVelocity, src/subhalo_unbind.cpp:108-138
    double msum;
    if (Nbound == 1) {
      gvel(0, RefVel);
      msum = gmass(0);
    } else {
      double sv[3] = {0, 0, 0};
      msum = 0.;
      for (Int i = 0; i < Nbound; i++) {
        float m = gmass(i), v[3];
        gvel(i, v);
        msum += m;
        for (int j = 0; j < 3; j++) sv[j] += v[j] * m;  // float product, double sum
      }
      for (int j = 0; j < 3; j++) RefVel[j] = (float)(sv[j] / msum);
    }
    r.mbound = (float)msum;
    // AveragePosition, src/subhalo_unbind.cpp:139-188
    if (Nbound == 1) {
      gpos(0, RefPos);
    } else {
      double origin[3] = {0, 0, 0}, sx[3] = {0, 0, 0}, ms = 0.;
      if (P.periodic) { float o[3]; gpos(0, o); for (int j = 0; j < 3; j++) origin[j] = o[j]; }
      for (Int i = 0; i < Nbound; i++) {
        float m = gmass(i), x[3];
        gpos(i, x);
        ms += m;
        if (P.periodic) {
          for (int j = 0; j < 3; j++) sx[j] += P.nearest((double)x[j] - origin[j]) * (double)m;
        } else {
          for (int j = 0; j < 3; j++) sx[j] += x[j] * m;  // float product
        }
      }
      for (int j = 0; j < 3; j++) {
        sx[j] /= ms;
        if (P.periodic) sx[j] += origin[j];
        RefPos[j] = (float)sx[j];
      }
    }
    if ((float)Nbound >= (float)Nlast * P.bmp) {  // converged
      if (!alive()) r.snapshot_index_of_death = -1;
      if (r.sink_track_id != -1) { r.snapshot_index_of_sink = -1; r.sink_track_id = -1; }
      std::sort(El.begin(), El.begin() + Nbound, comp_energy);
      if (P.refine && Nbound > MS) {
        // RefineBindingEnergyOrder (:234-262): the first MaxSampleSize entries
        // re-ordered by their energy in a tree over themselves (full masses),
        // keeping their unbinding energies
        load_tree(0, MS, 1.f);
        std::vector<PE> Ein(MS);
        for (Int i = 0; i < MS; i++) { Ein[i].pid = i; Ein[i].E = binding_energy(i, gmass(i)); }
        std::sort(Ein.begin(), Ein.end(), comp_energy);
        for (Int i = 0; i < MS; i++) Ein[i] = El[Ein[i].pid];
        for (Int i = 0; i < MS; i++) El[i] = Ein[i];
      }
      std::vector<Int> nl(N);
      for (Int i = 0; i < N; i++) { nl[i] = S.list[El[i].pid]; S.energy[i] = El[i].E; El[i].pid = i; }
      S.list.swap(nl);
      for (int j = 0; j < 3; j++) {
        r.mostbound_pos[j] = C.x(S.list[0], j);
        r.mostbound_vel[j] = C.v(S.list[0], j);
      }
      r.status |= HBT_STATUS_CONVERGED;
      break;
    }
  }
  S.nbound = Nbound;
  // AverageKinematics, src/subhalo_unbind.cpp:189-232 (Elist order == list order here
  // when converged; NumPart<=1 gives zeros)
  if (Nbound <= 1) {
    r.specific_self_potential_energy = 0.f;
    r.specific_self_kinetic_energy = 0.f;
    r.specific_angular_momentum[0] = r.specific_angular_momentum[1] = r.specific_angular_momentum[2] = 0.f;
  } else {
    double E = 0, K = 0, AM[3] = {0, 0, 0}, M = 0;
    for (Int i = 0; i < Nbound; i++) {
      float m = gmass(i), x[3], v[3];
      gpos(i, x); gvel(i, v);
      E += El[i].E * m;  // float product
      double dx[3], dv[3];
      for (int j = 0; j < 3; j++) {
        dx[j] = x[j] - RefPos[j];  // float difference
        if (P.periodic) dx[j] = P.nearest(dx[j]);
        dx[j] *= P.a;
        dv[j] = (double)(v[j] - RefVel[j]) + P.hz * dx[j];
        K += dv[j] * dv[j] * m;
      }
      AM[0] += (dx[1] * dv[2] - dx[2] * dv[1]) * m;
      AM[1] += (dx[2] * dv[0] - dx[0] * dv[2]) * m;
      AM[2] += (dx[0] * dv[1] - dx[1] * dv[0]) * m;
      M += m;
    }
    E /= M;
    K *= 0.5 / M;
    r.specific_self_potential_energy = (float)(E - K);
    r.specific_self_kinetic_energy = (float)K;
    for (int j = 0; j < 3; j++) r.specific_angular_momentum[j] = (float)(AM[j] / M);
  }
  // CountParticleTypes: the Nbound>100 branch does not skip null IDs
  count_types(C, S, Nbound <= 100);
  if (!(r.status & HBT_STATUS_CONVERGED))
    std::fill(S.energy.begin(), S.energy.end(), std::numeric_limits<float>::quiet_NaN());
}

struct Batch {
  const Params &P;
  const PartCols &C;
  std::vector<Sub> &subs;
  const std::vector<std::vector<Int>> &children;
};

void recursive_unbind(Batch &B, Int s, Drand48 &rng, Tree &tree) {
  Sub &S = B.subs[s];
  bool orphan = S.rec->nbound_prev <= 1;
  std::vector<Int> backup;
  if (orphan) backup = S.list;
  for (Int c : B.children[s]) {
    recursive_unbind(B, c, rng, tree);
    Sub &K = B.subs[c];
    S.list.insert(S.list.end(), K.list.begin() + K.nbound, K.list.end());
  }
  if (orphan) S.list.swap(backup);
  unbind(B.P, B.C, S, rng, tree);
  if (orphan) {
    S.list.swap(backup);
    S.energy.assign(S.list.size(), std::numeric_limits<float>::quiet_NaN());
  }
}

// ---- post-unbind properties (src/subhalo.cpp:242-398) ------------------------
// PeriodicDistance, src/config_parser.h:143-156: float differences, NEAREST,
// float sqrt of the float sum of squares.
inline float periodic_distance(const Params &P, const float c[3], float x, float y, float z) {
  float d[3] = {c[0] - x, c[1] - y, c[2] - z};
  if (P.periodic)
    for (int j = 0; j < 3; j++) d[j] = P.nearest(d[j]);
  return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

struct RadMassVel { float r, m, v; };

// Subhalo_t::CalculateProfileProperties (src/subhalo.cpp:242-333) with the
// spherical-overdensity size of Snapshot_t (src/snapshot.cpp:259-280) at the
// c200 factor (src/snapshot.cpp:328-336).
void profile_properties(const Params &P, const PartCols &C, const hbt_subhalo &s, const Int *list,
                        hbt_properties &o) {
  const Int nb = s.nbound;
  if (nb <= 1) {
    o.rmax_comoving = 0.f; o.vmax_physical = 0.f; o.r2sigma_comoving = 0.f; o.rhalf_comoving = 0.f;
    o.bound_r200crit_comoving = 0.f; o.bound_m200crit = 0.f;
    return;
  }
  const float vel_unit = P.G / P.a;
  std::vector<RadMassVel> prof(nb);
  for (Int i = 0; i < nb; i++) {
    const Int g = list[i];
    prof[i].r = periodic_distance(P, s.mostbound_pos, C.p->x[g], C.p->y[g], C.p->z[g]);
    prof[i].m = C.p->mass[g];
  }
  std::sort(prof.begin(), prof.end(), [](const RadMassVel &a, const RadMassVel &b) { return a.r < b.r; });
  double m_cum = 0.;
  for (auto &p : prof) p.m = (float)(m_cum += p.m);
  for (auto &p : prof) {
    if (p.r < P.eps) p.r = P.eps;  // SofteningHalo
    p.v = p.m / p.r;
  }
  auto maxprof = std::max_element(prof.begin(), prof.end(),
                                  [](const RadMassVel &a, const RadMassVel &b) { return a.v < b.v; });
  o.rmax_comoving = maxprof->r;
  o.vmax_physical = std::sqrt(maxprof->v * vel_unit);
  o.rhalf_comoving = prof[nb / 2].r;
  o.r2sigma_comoving = prof[(Int)(nb * 0.955)].r;
  const float virial_c200 = 200.f;
  const float rho = (float)((double)(virial_c200 * P.hz * P.hz) / 2.0 / (double)P.G * (double)P.a *
                            (double)P.a * (double)P.a);
  for (auto p = prof.rbegin(); p != prof.rend(); ++p) {
    if (p->m > rho * p->r * p->r * p->r) {
      o.bound_m200crit = p->m;
      o.bound_r200crit_comoving = (float)std::pow((double)(p->m / rho), 1.0 / 3);
      break;
    }
  }
  if (o.vmax_physical >= o.last_max_vmax_physical) {
    o.snapshot_index_of_last_max_vmax = P.snap;
    o.last_max_vmax_physical = o.vmax_physical;
  }
}

// Subhalo_t::CalculateShape (src/subhalo.cpp:337-398), serial order (the
// reference's OpenMP reduction with one thread), without the GSL eigenvectors.
void shape(const Params &P, const PartCols &C, const hbt_subhalo &s, const Int *list, hbt_properties &o) {
  const Int nb = s.nbound;
  if (nb <= 1) {
    for (int j = 0; j < 6; j++) { o.inertial_tensor[j] = 0.f; o.inertial_tensor_weighted[j] = 0.f; }
    return;
  }
  const float *c = s.mostbound_pos;
  double Ixx = 0, Iyy = 0, Izz = 0, Ixy = 0, Ixz = 0, Iyz = 0;
  double Ixxw = 0, Iyyw = 0, Izzw = 0, Ixyw = 0, Ixzw = 0, Iyzw = 0;
  for (Int i = 1; i < nb; i++) {
    const Int g = list[i];
    const float m = C.p->mass[g];
    float dx = C.p->x[g] - c[0], dy = C.p->y[g] - c[1], dz = C.p->z[g] - c[2];
    if (P.periodic) { dx = P.nearest(dx); dy = P.nearest(dy); dz = P.nearest(dz); }
    const float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    Ixx += dx2 * m; Iyy += dy2 * m; Izz += dz2 * m;
    Ixy += dx * dy * m; Ixz += dx * dz * m; Iyz += dy * dz * m;
    float dr2 = dx2 + dy2 + dz2;
    dr2 /= m;
    Ixxw += dx2 / dr2; Iyyw += dy2 / dr2; Izzw += dz2 / dr2;
    Ixyw += dx * dy / dr2; Ixzw += dx * dz / dr2; Iyzw += dy * dz / dr2;
  }
  const double t[6] = {Ixx, Ixy, Ixz, Iyy, Iyz, Izz}, w[6] = {Ixxw, Ixyw, Ixzw, Iyyw, Iyzw, Izzw};
  for (int j = 0; j < 6; j++) {
    o.inertial_tensor[j] = (float)t[j];
    o.inertial_tensor[j] /= s.mbound;
    o.inertial_tensor_weighted[j] = (float)w[j];
    o.inertial_tensor_weighted[j] /= s.mbound;
  }
}

// SubHelper_t::BuildPosition / BuildVelocity (src/subhalo_merge.cpp:29-127):
// the core of the first min(nbound, 20) list entries.  The periodic origin is
// a double, so differences, NEAREST and the final `+= origin` are in double.
struct MergeCore { float pos[3], vel[3], sig_r, sig_v; bool merged; };

void merge_core(const Params &P, const PartCols &C, const Int *list, Int nb, MergeCore &h) {
  h.merged = false;
  h.sig_r = h.sig_v = 0.f;
  for (int j = 0; j < 3; j++) h.pos[j] = h.vel[j] = 0.f;
  if (nb <= 0) return;
  const Int g0 = list[0];
  if (nb == 1) {
    for (int j = 0; j < 3; j++) { h.pos[j] = C.x(g0, j); h.vel[j] = C.v(g0, j); }
    return;
  }
  const Int n = nb > 20 ? 20 : nb;
  double sx[3] = {0, 0, 0}, sx2[3] = {0, 0, 0}, msum = 0, origin[3];
  for (int j = 0; j < 3; j++) origin[j] = C.x(g0, j);
  for (Int i = 0; i < n; i++) {
    const Int g = list[i];
    const float m = C.m(g);
    msum += m;
    for (int j = 0; j < 3; j++) {
      double dx;
      if (P.periodic) dx = P.nearest((double)C.x(g, j) - origin[j]);  // float - double
      else dx = C.x(g, j);
      sx[j] += dx * m;
      sx2[j] += dx * dx * m;
    }
  }
  for (int j = 0; j < 3; j++) {
    sx[j] /= msum;
    sx2[j] /= msum;
    h.pos[j] = (float)sx[j];
    if (P.periodic) h.pos[j] = (float)((double)h.pos[j] + origin[j]);  // float += double
    sx2[j] -= sx[j] * sx[j];
  }
  h.sig_r = (float)std::sqrt(sx2[0] + sx2[1] + sx2[2]);
  double sv[3] = {0, 0, 0}, sv2[3] = {0, 0, 0};
  msum = 0;
  for (Int i = 0; i < n; i++) {
    const Int g = list[i];
    const float m = C.m(g);
    msum += m;
    for (int j = 0; j < 3; j++) {
      const double dv = C.v(g, j);
      sv[j] += dv * m;
      sv2[j] += dv * dv * m;
    }
  }
  for (int j = 0; j < 3; j++) {
    sv[j] /= msum;
    sv2[j] /= msum;
    h.vel[j] = (float)sv[j];
    sv2[j] -= sv[j] * sv[j];
  }
  h.sig_v = (float)std::sqrt(sv2[0] + sv2[1] + sv2[2]);
}

// With sampling on (MaxSampleSizeOfPotentialEstimate > 0): the subhalos of
// each level in index order, each with RecursiveUnbind's handling of its
// nested subhalos' tails (:432-447) and its shuffle's draws taken from the
// rand() stream in that order.  The draws of a level are generated before the
// level is unbound, so several threads give the same result as one.
void unbind_levels(Batch &B, const std::vector<std::vector<Int>> &levels, int nthreads, GlibcRand &rnd) {
  const Params &P = B.P;
  Drand48 drand;
  Tree tree;
  for (const auto &lev : levels) {
    const Int nl = (Int)lev.size();
    std::vector<std::vector<Int>> backup(nl);
    std::vector<Int> off(nl + 1, 0);
    for (Int q = 0; q < nl; q++) {
      Sub &S = B.subs[lev[q]];
      const bool orphan = S.rec->nbound_prev <= 1;
      if (orphan) backup[q] = S.list;
      for (Int c : B.children[lev[q]]) {
        Sub &K = B.subs[c];
        S.list.insert(S.list.end(), K.list.begin() + K.nbound, K.list.end());
      }
      if (orphan) S.list.swap(backup[q]);
      const Int ulen = (Int)S.list.size();
      off[q + 1] = off[q] + ((P.max_sample > 0 && ulen > P.max_sample) ? ulen - 1 : 0);
    }
    std::vector<int32_t> dr(std::max<Int>(off[nl], 1));
    for (Int i = 0; i < off[nl]; i++) dr[i] = rnd.next();
    auto one = [&](Int q, Drand48 &rg, Tree &tr) {
      Sub &S = B.subs[lev[q]];
      unbind(P, B.C, S, rg, tr, dr.data() + off[q]);
      if (S.rec->nbound_prev <= 1) {
        S.list.swap(backup[q]);
        S.energy.assign(S.list.size(), std::numeric_limits<float>::quiet_NaN());
      }
    };
    if (nthreads <= 1) {
      for (Int q = 0; q < nl; q++) one(q, drand, tree);
    } else {
#pragma omp parallel num_threads(nthreads)
      {
        Drand48 rg;
        Tree tr;
#pragma omp for schedule(dynamic, 1)
        for (Int q = 0; q < nl; q++) one(q, rg, tr);
      }
    }
  }
}

// The visit order of a group's members: siblings (a subhalo's nested
// subhalos; a group's heads) in ascending visit_rank, index order for ties or
// without ranks.
void sort_siblings(const int64_t *visit_rank, std::vector<std::vector<Int>> &nested,
                   std::map<Int, std::vector<Int>> &heads) {
  if (!visit_rank) return;
  auto by_rank = [&](Int a, Int b) { return visit_rank[a] != visit_rank[b] ? visit_rank[a] < visit_rank[b] : a < b; };
  for (auto &v : nested) std::sort(v.begin(), v.end(), by_rank);
  for (auto &g : heads) std::sort(g.second.begin(), g.second.end(), by_rank);
}

}  // namespace

extern "C" {

int oracle_subhalo_properties(const hbt_params *hp, const hbt_particles *parts, const hbt_subhalo *subs,
                              int64_t nsub, const int64_t *out_index, hbt_properties *props,
                              int nthreads) {
  if (!hp || !parts || (nsub > 0 && (!subs || !out_index || !props))) return HBT_ERR_INVALID_ARG;
  Params P(*hp);
  PartCols C{parts};
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (Int i = 0; i < nsub; i++) {
    const Int *list = out_index + subs[i].out_offset;
    profile_properties(P, C, subs[i], list, props[i]);
    shape(P, C, subs[i], list, props[i]);
  }
  return HBT_OK;
}

// MaskSubhalos (src/subhalo_tracking.cpp:793-841): per host group a masker
// with an exclusion set, Mask(central) recursing into NestedSubhalos first
// (the central's list temporarily extended by the group's other heads,
// :828-834), orphans skipped (:808), first claim of an ID kept (:811-819).
// subs: offset/count/parent/nbound_prev; out_list laid out like the columns.
// visit_rank (or null: index order): siblings in ascending rank -- a group's
// heads in MemberTable.SubGroups order (the central first, :826-834, heads
// collected in that order at :676-686), nested subhalos in NestedSubhalos order.
int oracle_mask_subhalos(const hbt_particles *parts, const hbt_subhalo *subs, const int64_t *group,
                         const int64_t *visit_rank, int64_t nsub, int64_t *out_list, int64_t *new_count) {
  if (!parts || (nsub > 0 && (!subs || !group || !out_list || !new_count))) return HBT_ERR_INVALID_ARG;
  std::vector<std::vector<Int>> nested(nsub);
  std::map<Int, std::vector<Int>> heads;
  for (Int s = 0; s < nsub; s++) {
    const Int p = subs[s].parent;
    if (p >= 0) {
      if (p >= nsub || group[p] != group[s]) return HBT_ERR_INVALID_ARG;
      nested[p].push_back(s);
    } else if (group[s] >= 0) {
      heads[group[s]].push_back(s);
    }
  }
  sort_siblings(visit_rank, nested, heads);
  for (Int s = 0; s < nsub; s++) {
    new_count[s] = subs[s].count;
    for (Int i = 0; i < subs[s].count; i++) out_list[subs[s].offset + i] = subs[s].offset + i;
  }
  for (auto &g : heads) {
    const std::vector<Int> &hd = g.second;
    const Int central = hd[0];
    std::vector<Int> cnest = nested[central];
    cnest.insert(cnest.end(), hd.begin() + 1, hd.end());
    std::unordered_set<int64_t> excl;
    std::function<void(Int)> mask = [&](Int s) {
      for (Int c : (s == central ? cnest : nested[s])) mask(c);
      if (subs[s].nbound_prev <= 1) return;
      Int keep = 0;
      for (Int i = 0; i < subs[s].count; i++) {
        const Int col = subs[s].offset + i;
        if (excl.insert(parts->id[col]).second) out_list[subs[s].offset + keep++] = col;
      }
      new_count[s] = keep;
    };
    mask(central);
  }
  return HBT_OK;
}

// LevelUpDetachedSubhalos (src/subhalo_tracking.cpp:735-769) with
// Subhalo_t::LevelUpDetachedMembers (:771-786), the nest step of
// NestSubhalos (:669-688) on local indices: heads recorded first (Rank 0),
// centrals and field subhalos promoted to Rank 0, then each head's nest walked
// recursively, a member leaving the nest when its host halo differs from its
// host's or its Rank is 0.  Every head on entry walks its nest, also one that
// is itself nested.  Children are the subhalos naming a host, in index order.
// parent and rank in/out.
int oracle_nest_subhalos(int64_t nsub, const int64_t *group, const int32_t *is_central, int64_t *parent,
                         int32_t *rank) {
  if (nsub < 0 || (nsub > 0 && (!group || !is_central || !parent || !rank))) return HBT_ERR_INVALID_ARG;
  std::vector<std::vector<Int>> nested(nsub);
  for (Int s = 0; s < nsub; s++) {
    if (parent[s] < -1 || parent[s] >= nsub || parent[s] == s) return HBT_ERR_INVALID_ARG;
    if (parent[s] >= 0) nested[parent[s]].push_back(s);
  }
  for (Int s = 0; s < nsub; s++) {  // acyclic: every chain of hosts ends
    Int h = s, d = 0;
    while (parent[h] >= 0 && d <= nsub) { h = parent[h]; d++; }
    if (d > nsub) return HBT_ERR_INVALID_ARG;
  }
  std::vector<char> head(nsub);
  for (Int s = 0; s < nsub; s++) head[s] = rank[s] == 0;
  for (Int s = 0; s < nsub; s++)
    if (is_central[s] || group[s] < 0) rank[s] = 0;
  // every head on entry, nested or not, walks its nest (a second visit
  // through a host's nest changes nothing)
  std::function<void(Int)> level_up = [&](Int h) {
    std::vector<Int> keep;
    for (Int c : nested[h]) {
      if (group[c] != group[h] || rank[c] == 0) {
        rank[c] = 0;
        parent[c] = -1;
      } else {
        keep.push_back(c);
      }
      level_up(c);
    }
    nested[h].swap(keep);
  };
  for (Int s = 0; s < nsub; s++)
    if (head[s]) level_up(s);
  return HBT_OK;
}

// MergeSubhalos (src/subhalo_merge.cpp:187-220): GlueHeadNests (the group's
// other heads appended to the central's nest, src/subhalo_tracking.cpp:843-861),
// FillHelpers (HostTrackId from the nests, :163-186; cores from the first
// min(Nbound, 20) particles, BuildPosition/BuildVelocity :29-127),
// DetectTraps (:136-161 with SinkDistance :129-134), then with merge_trapped
// MergeRecursive from every group's central (:222-232) with MergeTo
// (:244-262), Unbind of every IsMerged subhalo in index order and their
// TruncateSource (:207-214).  list_index: the lists as column positions;
// subs in/out (offset/count then index out_list, packed in subhalo order);
// out_energy (optional): the re-unbound subhalos' binding energies along
// their lists, NaN elsewhere.
int oracle_merge_subhalos(const hbt_params *hp, const hbt_particles *parts, const int64_t *list_index,
                          int64_t n_list, hbt_subhalo *subs, const int64_t *group, const int64_t *visit_rank,
                          int64_t nsub, int merge_trapped, int64_t *out_list, float *out_energy,
                          int64_t out_capacity, int64_t *out_total) {
  if (!hp || !parts || (nsub > 0 && (!subs || !group || !list_index || !out_list))) return HBT_ERR_INVALID_ARG;
  Params P(*hp);
  PartCols C{parts};
  std::vector<std::vector<Int>> nested(nsub);
  std::map<Int, std::vector<Int>> heads;
  for (Int s = 0; s < nsub; s++) {
    if (subs[s].offset < 0 || subs[s].count < 0 || subs[s].offset + subs[s].count > n_list) return HBT_ERR_INVALID_ARG;
    const Int p = subs[s].parent;
    if (p >= 0) {
      if (p >= nsub || group[p] != group[s]) return HBT_ERR_INVALID_ARG;
      nested[p].push_back(s);
    } else if (group[s] >= 0) {
      heads[group[s]].push_back(s);
    }
  }
  sort_siblings(visit_rank, nested, heads);
  for (auto &g : heads)  // GlueHeadNests
    nested[g.second[0]].insert(nested[g.second[0]].end(), g.second.begin() + 1, g.second.end());
  std::vector<Int> host(nsub, -1);  // FillHostTrackIds
  for (Int s = 0; s < nsub; s++)
    for (Int c : nested[s]) host[c] = s;
  std::vector<Sub> S(nsub);
  for (Int s = 0; s < nsub; s++) {
    S[s].rec = &subs[s];
    S[s].list.assign(list_index + subs[s].offset, list_index + subs[s].offset + subs[s].count);
    S[s].nbound = subs[s].nbound;
    S[s].energy.assign(S[s].list.size(), std::numeric_limits<float>::quiet_NaN());
  }
  // FillCores: SubHelper_t::BuildPosition / BuildVelocity
  std::vector<MergeCore> H(nsub);
  for (Int s = 0; s < nsub; s++) merge_core(P, C, S[s].list.data(), S[s].nbound, H[s]);
  // DetectTraps
  for (Int i = 0; i < nsub; i++) {
    hbt_subhalo &r = subs[i];
    if (r.sink_track_id != -1) continue;
    for (Int h = host[i]; h >= 0; h = host[h]) {
      if (S[h].nbound <= 1) continue;
      const float d = periodic_distance(P, H[h].pos, r.mostbound_pos[0], r.mostbound_pos[1], r.mostbound_pos[2]);
      float e[3];
      for (int j = 0; j < 3; j++) e[j] = H[h].vel[j] - r.mostbound_vel[j];
      const float v = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
      const float delta = d / H[h].sig_r + v / H[h].sig_v;
      if (delta < 2.f) {
        r.sink_track_id = h;
        r.snapshot_index_of_sink = P.snap;
        if (S[i].nbound > 1) H[h].merged = true;
        break;
      }
    }
  }
  if (merge_trapped) {
    std::function<void(Int)> merge_rec = [&](Int s) {
      for (Int c : nested[s]) merge_rec(c);
      hbt_subhalo &r = subs[s];
      if (r.sink_track_id == -1 || r.snapshot_index_of_death != -1) return;
      if (S[s].nbound > 1) {  // MergeTo
        if (r.sink_track_id < 0 || r.sink_track_id >= nsub) return;
        Sub &T = S[r.sink_track_id];
        std::unordered_set<int64_t> ids;
        for (Int g : T.list) ids.insert(parts->id[g]);
        for (Int g : S[s].list)
          if (ids.insert(parts->id[g]).second) T.list.push_back(g);
        T.nbound += S[s].nbound;
        S[s].list.resize(1);
        S[s].energy.resize(1);
        S[s].nbound = 1;
        count_types(C, S[s], false);
        float m = 0.f;
        for (int t = 0; t < HBT_TYPE_MAX; t++) m += r.mbound_type[t];
        r.mbound = m;
      }
      r.snapshot_index_of_death = P.snap;
    };
    for (auto &g : heads) merge_rec(g.second[0]);
    Drand48 rng;
    Tree tree;
    // the merged hosts are unbound in index order (src/subhalo_merge.cpp:207-210);
    // with sampling their shuffles draw from a fresh rand() stream in that order
    GlibcRand rnd(1);
    std::vector<int32_t> dr;
    for (Int s = 0; s < nsub; s++) {
      if (!H[s].merged) continue;
      subs[s].nbound_prev = INT64_MAX;
      const Int n = (Int)S[s].list.size();
      dr.assign(std::max<Int>(n - 1, 1), 0);
      if (P.max_sample > 0 && n > P.max_sample)
        for (Int i = 0; i + 1 < n; i++) dr[i] = rnd.next();
      unbind(P, C, S[s], rng, tree, dr.data());
      Int ns = S[s].nbound <= 1 ? S[s].nbound : (Int)((float)S[s].nbound * P.relax);
      if (ns > (Int)S[s].list.size()) ns = (Int)S[s].list.size();
      S[s].list.resize(ns);
      S[s].energy.resize(ns);
    }
  }
  Int total = 0;
  for (Int s = 0; s < nsub; s++) total += (Int)S[s].list.size();
  if (out_total) *out_total = total;
  if (total > out_capacity) return HBT_ERR_CAPACITY;
  Int off = 0;
  for (Int s = 0; s < nsub; s++) {
    hbt_subhalo &r = subs[s];
    r.nbound = S[s].nbound;
    r.offset = off;
    r.count = (Int)S[s].list.size();
    r.list_len = r.nsource = r.count;
    r.out_offset = off;
    for (size_t i = 0; i < S[s].list.size(); i++) {
      if (out_energy) out_energy[off] = S[s].energy[i];
      out_list[off++] = S[s].list[i];
    }
  }
  return HBT_OK;
}

// The merging cores alone (BuildPosition / BuildVelocity of every subhalo):
// out[8*i ..] = position[3], velocity[3], sigma_R, sigma_V.
int oracle_merge_cores(const hbt_params *hp, const hbt_particles *parts, const int64_t *list_index,
                       const hbt_subhalo *subs, int64_t nsub, float *out) {
  if (!hp || !parts || (nsub > 0 && (!subs || !list_index || !out))) return HBT_ERR_INVALID_ARG;
  Params P(*hp);
  PartCols C{parts};
  for (Int s = 0; s < nsub; s++) {
    MergeCore h;
    merge_core(P, C, list_index + subs[s].offset, subs[s].nbound, h);
    float *o = out + 8 * s;
    for (int j = 0; j < 3; j++) { o[j] = h.pos[j]; o[3 + j] = h.vel[j]; }
    o[6] = h.sig_r;
    o[7] = h.sig_v;
  }
  return HBT_OK;
}

int oracle_plan_batch(hbt_subhalo *subs, int64_t nsub, int64_t *total) {
  if (nsub < 0 || (nsub > 0 && !subs) || !total) return HBT_ERR_INVALID_ARG;
  std::vector<Int> cap(nsub, 0);
  for (Int s = 0; s < nsub; s++) {
    if (subs[s].count < 0 || subs[s].parent >= nsub || subs[s].parent == s) return HBT_ERR_INVALID_ARG;
    cap[s] = subs[s].count;
  }
  // a parent's list grows by its descendants' lists; parents must precede nothing in particular,
  // so propagate with a depth-first accumulation (children may have any index)
  std::vector<Int> depth(nsub, 0);
  for (Int s = 0; s < nsub; s++) {
    Int d = 0, p = subs[s].parent;
    while (p >= 0) { d++; p = subs[p].parent; if (d > nsub) return HBT_ERR_INVALID_ARG; }
    depth[s] = d;
  }
  std::vector<Int> ord(nsub);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](Int a, Int b) { return depth[a] > depth[b]; });
  for (Int s : ord)
    if (subs[s].parent >= 0) cap[subs[s].parent] += cap[s];
  Int off = 0;
  for (Int s = 0; s < nsub; s++) { subs[s].out_offset = off; off += cap[s]; }
  *total = off;
  return HBT_OK;
}

// Unbinds a batch of subhalos the way RefineParticles does: every top-level
// subhalo (parent < 0) in index order through RecursiveUnbind, then
// TruncateSource on all.  nthreads>1 runs top-level nests in parallel
// (schedule(dynamic,1), src/subhalo_unbind.cpp:479) with one drand48 stream
// per thread; nthreads==1 matches a single-threaded reference run.  With
// sampling on, subhalos go level by level (unbind_levels).
// rand_seed / rand_draws: the rand() stream the shuffles draw from: srand(seed)
// followed by *rand_draws calls; on return *rand_draws counts the calls made.
int oracle_unbind_batch_rs(const hbt_params *hp, const hbt_particles *parts, hbt_subhalo *subs,
                           int64_t nsub, int64_t *out_index, float *out_energy, int nthreads,
                           uint32_t rand_seed, int64_t *rand_draws) {
  if (!hp || !parts || (nsub > 0 && !subs) || !rand_draws || *rand_draws < 0) return HBT_ERR_INVALID_ARG;
  Params P(*hp);
  PartCols C{parts};
  std::vector<Sub> S(nsub);
  std::vector<std::vector<Int>> children(nsub);
  std::vector<Int> roots;
  for (Int s = 0; s < nsub; s++) {
    S[s].rec = &subs[s];
    S[s].list.resize(subs[s].count);
    for (Int i = 0; i < subs[s].count; i++) S[s].list[i] = subs[s].offset + i;
    S[s].nbound = subs[s].nbound_prev;
    if (subs[s].parent >= 0) children[subs[s].parent].push_back(s);
    else roots.push_back(s);
  }
  Batch B{P, C, S, children};
  if (P.max_sample > 0) {
    // levels: nest heights, subhalos in index order within each
    std::vector<int> height(nsub, 0);
    std::function<int(Int)> hgt = [&](Int s) {
      int h = 0;
      for (Int c : children[s]) h = std::max(h, hgt(c) + 1);
      return height[s] = h;
    };
    int hmax = 0;
    for (Int s : roots) hmax = std::max(hmax, hgt(s));
    std::vector<std::vector<Int>> levels(hmax + 1);
    for (Int s = 0; s < nsub; s++) levels[height[s]].push_back(s);
    GlibcRand rnd(rand_seed);
    rnd.skip(*rand_draws);
    unbind_levels(B, levels, nthreads, rnd);
    *rand_draws = rnd.draws;
  } else if (nthreads <= 1) {
    Drand48 rng;
    Tree tree;
    for (Int s : roots) recursive_unbind(B, s, rng, tree);
  } else {
    // largest nests first keeps the dynamic schedule balanced
    std::vector<Int> cost(nsub, 0);
    for (Int s = 0; s < nsub; s++) { Int p = s; while (subs[p].parent >= 0) p = subs[p].parent; cost[p] += subs[s].count; }
    std::vector<Int> order = roots;
    std::stable_sort(order.begin(), order.end(), [&](Int a, Int b) { return cost[a] > cost[b]; });
#pragma omp parallel num_threads(nthreads)
    {
      Drand48 rng;
      Tree tree;
#pragma omp for schedule(dynamic, 1)
      for (Int k = 0; k < (Int)order.size(); k++) recursive_unbind(B, order[k], rng, tree);
    }
  }
  for (Int s = 0; s < nsub; s++) {
    hbt_subhalo &r = subs[s];
    Sub &X = S[s];
    r.nbound = X.nbound;
    r.list_len = (Int)X.list.size();
    Int ns = X.nbound <= 1 ? X.nbound : (Int)((float)X.nbound * P.relax);
    if (ns > r.list_len) ns = r.list_len;
    r.nsource = ns;
    if (out_index)
      for (Int i = 0; i < r.list_len; i++) out_index[r.out_offset + i] = X.list[i];
    if (out_energy)
      for (Int i = 0; i < r.list_len; i++)
        out_energy[r.out_offset + i] = i < (Int)X.energy.size() ? X.energy[i] : std::numeric_limits<float>::quiet_NaN();
  }
  return HBT_OK;
}

int oracle_unbind_batch(const hbt_params *hp, const hbt_particles *parts, hbt_subhalo *subs,
                        int64_t nsub, int64_t *out_index, float *out_energy, int nthreads) {
  int64_t draws = 0;
  return oracle_unbind_batch_rs(hp, parts, subs, nsub, out_index, out_energy, nthreads, 1, &draws);
}

// n rand() values after srand(seed) and `skip` calls.
int oracle_glibc_rand(uint32_t seed, int64_t skip, int64_t n, int32_t *out) {
  if (skip < 0 || n < 0 || (n > 0 && !out)) return HBT_ERR_INVALID_ARG;
  GlibcRand r(seed);
  r.skip(skip);
  for (int64_t i = 0; i < n; i++) out[i] = r.next();
  return HBT_OK;
}

// std::random_shuffle of each index range [0, len[k]) in turn, drawing from
// srand(seed) after `skip` calls: perm holds the ranges one after another.
int oracle_random_shuffle(uint32_t seed, int64_t skip, int64_t nseg, const int64_t *len, int64_t *perm) {
  if (skip < 0 || nseg < 0) return HBT_ERR_INVALID_ARG;
  GlibcRand r(seed);
  r.skip(skip);
  int64_t o = 0;
  for (int64_t k = 0; k < nseg; k++) {
    std::vector<int64_t> a(len[k]);
    std::iota(a.begin(), a.end(), 0);
    std::vector<int32_t> d(std::max<int64_t>(len[k] - 1, 1));
    for (int64_t i = 0; i + 1 < len[k]; i++) d[i] = r.next();
    random_shuffle_with(a, d.data());
    std::copy(a.begin(), a.end(), perm + o);
    o += len[k];
  }
  return HBT_OK;
}

// GravityTree_t::Build over sources in the given order + EvaluatePotential.
int oracle_tree_potential(const hbt_params *hp, int64_t n_src, const float *sx, const float *sy,
                          const float *sz, const float *sm, int64_t n_tgt, const float *tx,
                          const float *ty, const float *tz, const float *tmass, double *out_phi) {
  if (!hp || n_src < 1 || n_tgt < 0) return HBT_ERR_INVALID_ARG;
  Params P(*hp);
  Tree tree;
  Drand48 rng;
  tree.np = n_src;
  tree.px.assign(sx, sx + n_src); tree.py.assign(sy, sy + n_src);
  tree.pz.assign(sz, sz + n_src); tree.pm.assign(sm, sm + n_src);
  tree.build(P, rng);
  for (Int i = 0; i < n_tgt; i++)
    out_phi[i] = tree.potential(P, tx[i], ty[i], tz[i], tmass ? tmass[i] : 0.f);
  return HBT_OK;
}

// Number of cells of the reference tree over the given sources (test aid).
int64_t oracle_tree_ncell(const hbt_params *hp, int64_t n_src, const float *sx, const float *sy,
                          const float *sz, const float *sm) {
  Params P(*hp);
  Tree tree;
  Drand48 rng;
  tree.np = n_src;
  tree.px.assign(sx, sx + n_src); tree.py.assign(sy, sy + n_src);
  tree.pz.assign(sz, sz + n_src); tree.pm.assign(sm, sm + n_src);
  tree.build(P, rng);
  return tree.ncell();
}

// Exact pair sum with the walk's softened kernel (src/gravity_tree.cpp:144-160),
// self term excluded; mode HBT_DIRECT_FP64 / FP32 / FP32_KAHAN.
int oracle_direct_potential(const hbt_params *hp, int64_t nseg, const int64_t *seg, const float *x,
                            const float *y, const float *z, const float *m, int mode,
                            double *out_phi, int nthreads) {
  if (!hp || nseg < 0) return HBT_ERR_INVALID_ARG;
  Params P(*hp);
  const double h = 2.8 * (double)P.eps, hinv = 1.0 / h;
  const float hf = 2.8f * P.eps, hinvf = 1.0f / hf;
  const double scale = (double)P.G / (double)P.a;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (Int s = 0; s < nseg; s++) {
    for (Int i = seg[s]; i < seg[s + 1]; i++) {
      if (mode == HBT_DIRECT_FP64) {
        double pot = 0;
        for (Int j = seg[s]; j < seg[s + 1]; j++) {
          if (j == i) continue;
          double dx = (double)x[j] - x[i], dy = (double)y[j] - y[i], dz = (double)z[j] - z[i];
          if (P.periodic) { dx = P.nearest(dx); dy = P.nearest(dy); dz = P.nearest(dz); }
          double r = std::sqrt(dx * dx + dy * dy + dz * dz);
          if (r >= h) pot -= m[j] / r;
          else {
            double u = r * hinv, wp;
            if (u < 0.5) wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
            else wp = -3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
            pot += m[j] * hinv * wp;
          }
        }
        out_phi[i] = pot * scale;
      } else {
        float pot = 0.f, comp = 0.f;
        for (Int j = seg[s]; j < seg[s + 1]; j++) {
          if (j == i) continue;
          float dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
          if (P.periodic) { dx = P.nearest(dx); dy = P.nearest(dy); dz = P.nearest(dz); }
          float r = std::sqrt(dx * dx + dy * dy + dz * dz), term;
          if (r >= hf) term = -m[j] / r;
          else {
            float u = r * hinvf, wp;
            if (u < 0.5f) wp = -2.8f + u * u * (5.333333333333f + u * u * (6.4f * u - 9.6f));
            else wp = -3.2f + 0.066666666667f / u + u * u * (10.666666666667f + u * (-16.0f + u * (9.6f - 2.133333333333f * u)));
            term = m[j] * hinvf * wp;
          }
          if (mode == HBT_DIRECT_FP32_KAHAN) {
            float yk = term - comp, t = pot + yk;
            comp = (t - pot) - yk;
            pot = t;
          } else {
            pot += term;
          }
        }
        out_phi[i] = (double)pot * scale;
      }
    }
  }
  return HBT_OK;
}

// ID -> index over a table (MappedIndexTable_t semantics, src/hash.tpp:18-56):
// sorted (id, index) pairs and a binary search; unknown or negative IDs -> -1.
int oracle_id_to_index(int64_t n_table, const int64_t *ids, int64_t n_query, const int64_t *q,
                       int64_t *out) {
  std::vector<std::pair<Int, Int>> t(n_table);
  for (Int i = 0; i < n_table; i++) t[i] = {ids[i], i};
  std::sort(t.begin(), t.end());
  for (Int k = 0; k < n_query; k++) {
    Int key = q[k];
    if (key < 0) { out[k] = -1; continue; }
    auto it = std::lower_bound(t.begin(), t.end(), std::make_pair(key, (Int)-1));
    out[k] = (it != t.end() && it->first == key) ? it->second : -1;
  }
  return HBT_OK;
}

int oracle_abi_sizes(int64_t *sz_params, int64_t *sz_sub, int64_t *sz_parts) {
  *sz_params = sizeof(hbt_params);
  *sz_sub = sizeof(hbt_subhalo);
  *sz_parts = sizeof(hbt_particles);
  return 0;
}

}  // extern "C"
