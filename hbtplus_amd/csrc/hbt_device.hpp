// hbt_device.hpp -- device-side data structures and arithmetic helpers shared
// by the unbinding kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hbt_prims.hpp"

namespace hbt {

constexpr int MAX_DEPTH = 21;   // octant digits per 64-bit key (3 bits each)
constexpr int NLEVEL = MAX_DEPTH + 1;

// segment flags
enum : int {
  F_ACTIVE = 1,       // in the pass loop
  F_CORR = 2,         // next pass is a correction pass
  F_CONV = 4,         // converged
  F_DISR = 8,         // disrupted
  F_EARLY = 16,       // 0 or 1 particles: no iteration
  F_ORPHAN = 32,      // orphan: unbinds its own list only, output keeps the extended list
  F_FINAL = 64,       // converged or disrupted in this pass, outputs pending
  F_DEEP = 128,       // tree keys truncated at MAX_DEPTH above the resolution floor
  F_SAMP = 256,       // this full pass's tree holds only the first MaxSampleSize entries
  F_SHUF = 512,       // the unbound list was shuffled (longer than MaxSampleSize)
  F_REFPEND = 1024,   // converged; RefineBindingEnergyOrder and the outputs are pending
  F_REFINE = 2048,    // the refinement pass: energies go to Er, not E
};

// Unbinding parameters with the reference's derived values
// (src/config_parser.cpp:95-103).
struct DevParams {
  float G, eps, theta2, res, box, box_half, bmp, relax, a, hz;
  int periodic, min_part, snap, replay, wide_sum;
  int max_sample;     // MaxSampleSizeOfPotentialEstimate (0: no sampling)
  int refine;         // RefineMostboundParticle, with sampling only
};

// What the host fixes per segment of a level (the rest of Seg starts at zero
// and is set on the device).
struct SegIn {
  int64_t ibase, out_off;
  int64_t list_off;  // first entry of the own list in the list positions
  int64_t nbprev;    // Nbound before the pass (RecursiveUnbind's orphan test)
  int sub, cap, count, wgroup;
  int draw_off;      // first rand() value of the list's shuffle in the level's draws, -1: none
  int shuf_len;      // the list length the host counted those draws for
};

// Per-subhalo state for one level of the batch.
struct Seg {
  int64_t ibase;       // first working slot (aligned to WAVE)
  int64_t out_off;     // caller's out_offset
  int sub;             // index into the caller's hbt_subhalo array
  int64_t list_off;    // first entry of the own list in the list positions
  int64_t nbprev;      // Nbound before the pass (orphan test)
  int cap;             // working capacity (own list + nested tails)
  int count;           // own source list length
  int len;             // working list length (own + appended tails)
  int ulen;            // length that is unbound (len, or count for orphans)
  int nbound, nlast;
  int flags;
  int n_src, src_lo, n_tgt;   // current pass
  int n_items;
  int nP, nG, K, hole;        // partition
  int ncopy;                  // Elist entries the pass's partition moved (its nlast), copied back with the bound flags
  int nsort, sort_lo;         // energy sort: Elist[sort_lo, sort_lo + nsort)
  int iter;
  long long evals;
  int D, dlen, straddle;
  int Dfull;                  // D of the latest full pass: bounds D of every later pass
  int sfull;                  // straddle of the latest full pass, whose sources hold
                              // every later pass's sources and targets
  int wgroup;                 // lanes per traversal group in the walk
  int draw_off;               // shuffle: first draw in the level's draws (-1: not shuffled)
  int shuf_len;               // the unbound list length the host counted the draws for
  int sall;                   // straddle over the whole unbound list (shuffled lists)
  int swapq;                  // disruption of a shuffled list: first position of the old most-bound ID
  float mf;                   // mass factor of the sampled tree (Nlast / MaxSampleSize)
  long long old_mb_id;        // ID of the list's first particle before the shuffle
  float ref_pos[3], ref_vel[3], old_pos[3], old_vel[3];
  float rvd[3], dK;
  float mbound;
  double center[3], Len;
  unsigned int bmin[3], bmax[3];
};

// ---- float helpers -------------------------------------------------------
__device__ __forceinline__ unsigned int f2ord(float f) {
  unsigned int u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned int u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// NEAREST macro (src/config_parser.h:142) in float and double
__device__ __forceinline__ float nearest_f(float x, const DevParams &P) {
  return (x > P.box_half) ? (x - P.box) : ((x < -P.box_half) ? (x + P.box) : x);
}
__device__ __forceinline__ double nearest_d(double x, const DevParams &P) {
  return (x > (double)P.box_half) ? (x - (double)P.box)
                                  : ((x < -(double)P.box_half) ? (x + (double)P.box) : x);
}

// Snapshot_t::RelativeVelocity (src/snapshot.h:100-111), all float
__device__ __forceinline__ void rel_vel(const DevParams &P, float tx, float ty, float tz, float tvx,
                                        float tvy, float tvz, const float rp[3], const float rv[3],
                                        float dv[3]) {
#pragma clang fp contract(off)
  float t[3] = {tx, ty, tz}, v[3] = {tvx, tvy, tvz};
  for (int j = 0; j < 3; j++) {
    float dx = t[j] - rp[j];
    if (P.periodic) dx = nearest_f(dx, P);
    dv[j] = v[j] - rv[j];
    dv[j] += P.hz * P.a * dx;
  }
}

__device__ __forceinline__ float vnorm_f(const float v[3]) {
#pragma clang fp contract(off)
  return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}
__device__ __forceinline__ float vdot_f(const float a[3], const float b[3]) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Softened kernel of the potential walk (src/gravity_tree.cpp:144-160):
// contribution of mass m at distance r (r2 = r*r) to the sum pot.
__device__ __forceinline__ double kernel_term_d(double m, double r, double h, double hinv) {
#pragma clang fp contract(off)
  if (r >= h) return -(m / r);
  double u = r * hinv, wp;
  if (u < 0.5)
    wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
  else
    wp = -3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
  return m * hinv * wp;
}

// FP32 form with explicit fused multiply-adds (the walk variants share this
// code, so their results agree bit for bit): rsq for 1/r, r = r2 * (1/r);
// r2 == 0 is the self term (u = 0).
__device__ __forceinline__ float pair_r2(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ float kernel_term_f(float m, float r2, float h2, float hinv) {
#pragma clang fp contract(off)
  const float rinv = __builtin_amdgcn_rsqf(r2);
  if (r2 >= h2) return -m * rinv;
  const float r = r2 > 0.f ? r2 * rinv : 0.f;
  const float u = r * hinv, u2 = u * u;
  float wp;
  if (u < 0.5f) {
    wp = __builtin_fmaf(u2, __builtin_fmaf(u2, __builtin_fmaf(6.4f, u, -9.6f), 5.333333333333f), -2.8f);
  } else {
    float t = __builtin_fmaf(-2.133333333333f, u, 9.6f);
    t = __builtin_fmaf(t, u, -16.0f);
    t = __builtin_fmaf(t, u, 10.666666666667f);
    wp = __builtin_fmaf(u2, t, __builtin_fmaf(0.066666666667f, __builtin_amdgcn_rcpf(u), -3.2f));
  }
  return m * hinv * wp;
}

template <bool COMPENSATED = true>
struct Kahan {
  float s, c;
  __device__ __forceinline__ void init(float x) { s = x; c = 0.f; }
  __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
    if (!COMPENSATED) { s += x; return; }
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

}  // namespace hbt
