// hbt_kernels.hpp -- the unbinding pass kernels (gfx950).
//
// One pass of the reference's Unbind loop (src/subhalo_unbind.cpp:310-422)
// runs for every active segment (subhalo) of a level at once:
//   pass_setup        full pass over Elist[0,Nlast) or correction pass over the
//                     freshly removed block Elist[Nbound,Nlast) (:312-356)
//   bbox / tree_params / keys
//                     root cube and per-particle octant digits computed with the
//                     reference's double-precision centre iteration
//                     (src/oct_tree.tpp:118-224)
//   (segmented sort of the keys)
//   delta / emit / items / moments
//                     compressed octree in pre-order with skip pointers; node
//                     moments accumulated in double over float children in son
//                     order (src/gravity_tree.cpp:18-77)
//   walk              stackless walk per target (src/gravity_tree.cpp:79-175)
//   partition_*       parallel replica of PartitionBindingEnergy (:21-58)
//   frames            FP64 AverageVelocity/AveragePosition, convergence (:357-421)
//   finalize          permuted list, kinematics, particle-type counts (:405-431)
#pragma once
#include "hbt_device.hpp"
#include "../../include/hbt_unbind.h"

namespace hbt {

// One tree item (leaf or node) in DFS pre-order, one 32-byte record: the
// walk reads both halves of an item at once.
// Rows of Work::ncount, each written by the kernel that decides the count:
// sources of the tree (pass setup), the partition scan (pass setup), the bound
// target list (partition), the energy sort (frames).
enum { NC_SRC = 0, NC_PART = 1, NC_BOUND = 2, NC_SORT = 3, NC_CROSS = 4, NC_ROWS = 5 };

struct __align__(32) Item {
  float4 pm;   // x, y, z, mass (a node: centre of mass and total mass)
  int2 ax;     // float bits of len^2 (0 for a leaf), skip (first item after the subtree)
  int2 sb;     // skip as a byte offset from the segment's first item (32 * skip), unused
};

struct Work {
  // per level
  Seg *seg;
  int nseg;
  int64_t nslots;
  // the slots a pass visits: all (act_chunk null, nact = nslots) or the
  // wave-sized chunks of the segments still active (nact = 64 * chunks)
  const int *act_chunk;
  int64_t nact;
  int *segact;               // [nseg] 1: the segment stays active after the pass
  const int *seg_of_slot;
  const int64_t *ibase;
  int *ncount;               // [NC_ROWS][nseg] element counts of the pass's scans and sorts
  // particle columns of the caller (device)
  const int64_t *g_id;
  const float *g_x, *g_y, *g_z, *g_vx, *g_vy, *g_vz, *g_m;
  const int32_t *g_type;
  const int64_t *g_index;    // optional: list entry -> particle column position
  const float4 *g_packed;    // optional: the columns interleaved, 2 float4 per particle
  // working arrays [nslots]
  int *gidx;
  float4 *pos;               // x, y, z, mass: one record, so a gather reads one line
  float4 *vel;               // vx, vy, vz, type (int bits)
  float *E;
  float *Er;                 // energies of the refinement pass (RefineBindingEnergyOrder)
  int *el, *el2, *sp;
  u64 *k0, *k1;
  u32 *v0, *v1;
  signed char *dl;
  u32 *bits;
  int *pre, *nlo, *nhi, *npos;
  int *cross;                // per super-tile: pairs of the nodes whose leaves span tiles in it
  int *cross2;               // per segment: pairs of the nodes whose leaves span super-tiles
  int *tcount;               // per tile of the pass's table: list length of the super-tile it starts
  Item *items;               // [2*nslots] tree items in pre-order
  u64 *scan64;
  int *pk, *gk;
  unsigned char *isb;
  int *active;               // [4]: segments active after the pass, the largest Nlast among them
  unsigned long long *counters;  // [4]: interactions, opened nodes, targets (instrumented passes)
  // caller outputs
  hbt_subhalo *subs;
  int64_t *out_index;
  float *out_energy;
};

// counts list entries that are not positions in [0, n)
__global__ void check_index_kernel(const int64_t *idx, int64_t n_list, int64_t n, int *bad) {
  int cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_list; i += (int64_t)gridDim.x * blockDim.x)
    cnt += (idx[i] < 0 || idx[i] >= n) ? 1 : 0;
  if (cnt) atomicAdd(bad, cnt);
}

// The slots a wave visits form one wave-aligned chunk, and segments start at
// multiples of WAVE, so a wave's slots share one segment: the lookups read
// one entry for the wave (scalar loads of a wave-uniform address).
__device__ __forceinline__ int seg_lane_local(const Work &W, int64_t s, int &k) {
  const int64_t s0 = (int64_t)__builtin_amdgcn_readfirstlane((int)(s >> 6)) << 6;
  k = W.seg_of_slot[s0];
  if (k < 0) return -1;
  return (int)(s - W.ibase[k]);
}

// i-th visited slot (the active lists hold wave-sized chunks, nact a multiple of WAVE)
__device__ __forceinline__ int64_t slot_at(const Work &W, int64_t i) {
  if (W.act_chunk && i < W.nact) {
    const int c = __builtin_amdgcn_readfirstlane((int)(i >> 6));
    return (int64_t)W.act_chunk[c] * WAVE + (i & (WAVE - 1));
  }
  return i;
}

#define SLOT_LOOP(s, W)                                                                        \
  for (int64_t s##_i = (int64_t)blockIdx.x * BLOCK + threadIdx.x, s = slot_at((W), s##_i);    \
       s##_i < (W).nact; s##_i += (int64_t)gridDim.x * BLOCK, s = slot_at((W), s##_i))

#define SEG_LOOP(k, W) \
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < (W).nseg; k += gridDim.x * blockDim.x)

// One wave per segment (k is wave-uniform).
#define WAVE_SEG_LOOP(k, W)                                                     \
  for (int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) / WAVE); k < (W).nseg; \
       k += (int)(gridDim.x * blockDim.x / WAVE))

// Sums of C values over the tiles t0 .. t0+nt-1 of part: the wave's lanes take
// strided tiles, then wave_total; the sums are valid in every lane.
template <int C>
__device__ __forceinline__ void wave_sum_tiles(const double *__restrict__ part, int t0, int nt, double v[C]) {
#pragma clang fp contract(off)
  for (int c = 0; c < C; c++) v[c] = 0;
  for (int t = t0 + (int)(threadIdx.x & (WAVE - 1)); t < t0 + nt; t += WAVE)
    for (int c = 0; c < C; c++) v[c] += part[(int64_t)t * C + c];
  for (int c = 0; c < C; c++) v[c] = wave_total(v[c]);
}

// ---------------------------------------------------------------------------
// Level initialisation

// Fills seg_of_slot for the level.
__global__ void fill_seg_of_slot_kernel(int nseg, const int64_t *ibase, const int *cap_aligned,
                                        int *seg_of_slot) {
  for (int k = blockIdx.x; k < nseg; k += gridDim.x) {
    const int64_t b = ibase[k];
    const int c = cap_aligned[k];
    for (int i = threadIdx.x; i < c; i += blockDim.x) seg_of_slot[b + i] = k;
  }
}

// The same, slot-parallel, for a level whose segments tile the slots in order
// (ibase ascending, each segment's slots followed by the next one's): a slot's
// segment is the last one starting at or before it.  A giant segment is then
// not left to one workgroup.
__global__ void __launch_bounds__(BLOCK) seg_of_slot_kernel(int nseg, const int64_t *__restrict__ ibase,
                                                            int64_t nslots, int *seg_of_slot) {
  for (int64_t s = (int64_t)blockIdx.x * BLOCK + threadIdx.x; s < nslots; s += (int64_t)gridDim.x * BLOCK) {
    int lo = 0, hi = nseg - 1;  // ibase[lo] <= s
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ibase[mid] <= s) lo = mid; else hi = mid - 1;
    }
    seg_of_slot[s] = lo;
  }
}

// The same for a level whose segments start at multiples of WAVE slots: one
// search per WAVE slots.  Each wave takes WAVE such chunks, lane i searching
// chunk i, then writes the chunks' slots one chunk per store across the lanes.
__global__ void __launch_bounds__(BLOCK) seg_of_chunk_kernel(int nseg, const int64_t *__restrict__ ibase,
                                                             int64_t nslots, int *seg_of_slot) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nchunk = (nslots + WAVE - 1) / WAVE;
  const int64_t waves = (int64_t)gridDim.x * (BLOCK / WAVE);
  for (int64_t g = (int64_t)blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE; g * WAVE < nchunk; g += waves) {
    const int64_t c = g * WAVE + lane;
    int lo = 0;
    if (c < nchunk) {
      const int64_t s = c * WAVE;
      int hi = nseg - 1;  // ibase[lo] <= s
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ibase[mid] <= s) lo = mid; else hi = mid - 1;
      }
    }
    for (int j = 0; j < WAVE; j++) {
      const int k = __shfl(lo, j, WAVE);
      const int64_t s = (g * WAVE + j) * WAVE + lane;
      if (s < nslots) seg_of_slot[s] = k;
    }
  }
}

// The level's segment records: zero, then the host's fixed fields.
__global__ void __launch_bounds__(BLOCK) seg_expand_kernel(Seg *seg, const SegIn *in, int n) {
  for (int k = blockIdx.x * BLOCK + threadIdx.x; k < n; k += gridDim.x * BLOCK) {
    const SegIn a = in[k];
    Seg S{};
    S.ibase = a.ibase;
    S.out_off = a.out_off;
    S.sub = a.sub;
    S.cap = a.cap;
    S.count = a.count;
    S.list_off = a.list_off;
    S.nbprev = a.nbprev;
    S.wgroup = a.wgroup;
    S.draw_off = a.draw_off;
    S.shuf_len = a.shuf_len;
    S.mf = 1.f;
    S.swapq = INT_MAX;
    seg[k] = S;
  }
}

// Builds the working list of every segment: own particles in source order
// (their positions are written by gather_kernel), then the unbound tails of
// nested subhalos (read from their output lists, src/subhalo_unbind.cpp:437-443).
// child_ptr/child_list: CSR list of the subhalos nested in each segment.
__global__ void build_lists_kernel(Work W, const int *child_ptr, const int *child_list) {
  for (int k = blockIdx.x; k < W.nseg; k += gridDim.x) {
    Seg &S = W.seg[k];
    const int64_t b = S.ibase;
    int pos = S.count;
    for (int c = child_ptr[k]; c < child_ptr[k + 1]; c++) {
      const hbt_subhalo &q = W.subs[child_list[c]];
      const int tail = (int)(q.list_len - q.nbound);
      for (int i = threadIdx.x; i < tail; i += blockDim.x)
        W.gidx[b + pos + i] = (int)W.out_index[q.out_offset + q.nbound + i];
      pos += tail;
    }
    if (threadIdx.x == 0) S.len = pos;
    __syncthreads();
  }
}

// The snapshot's columns interleaved per particle (x, y, z, mass | vx, vy, vz,
// type): a gather in list order, which is random in the snapshot, then reads
// one 32-byte record per particle instead of eight scattered words.
__global__ void __launch_bounds__(BLOCK) pack_particles_kernel(hbt_particles P, float4 *out) {
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * BLOCK) {
    out[2 * i] = make_float4(P.x[i], P.y[i], P.z[i], P.mass[i]);
    out[2 * i + 1] = make_float4(P.vx[i], P.vy[i], P.vz[i], __int_as_float(P.type ? P.type[i] : 1));
  }
}

// Column position of list entry l of a segment's own list.
__device__ __forceinline__ int own_entry(const Work &W, const Seg &S, int l) {
  const int64_t o = S.list_off + l;
  return W.g_index ? (int)W.g_index[o] : (int)o;
}

// Column positions of the segments' own particles (phase bit 1), and local
// copies of the listed particles' columns (phase bit 2), slot-parallel.  A
// shuffled list is permuted between the two phases.
__global__ void __launch_bounds__(BLOCK) gather_kernel(Work W, int phase) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0 || l >= W.seg[k].len) continue;
    int g;
    if ((phase & 1) && l < W.seg[k].count) {
      g = own_entry(W, W.seg[k], l);
      W.gidx[s] = g;
    } else {
      g = W.gidx[s];
    }
    if (!(phase & 2)) continue;
    if (W.g_packed) {
      W.pos[s] = W.g_packed[2 * (int64_t)g];
      W.vel[s] = W.g_packed[2 * (int64_t)g + 1];
    } else {
      W.pos[s] = make_float4(W.g_x[g], W.g_y[g], W.g_z[g], W.g_m[g]);
      W.vel[s] = make_float4(W.g_vx[g], W.g_vy[g], W.g_vz[g], __int_as_float(W.g_type ? W.g_type[g] : 1));
    }
  }
}

// Unbind entry (src/subhalo_unbind.cpp:263-309): early exits, initial Elist;
// the shuffle's sort count (steps 1..ulen-1) of a shuffled list.
__global__ void seg_init_kernel(Work W, DevParams P) {
  if (blockIdx.x == 0 && threadIdx.x == 0) W.active[3] = 0;  // refinements pending in the level
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    hbt_subhalo &r = W.subs[S.sub];
    const bool orphan = S.nbprev <= 1;
    S.ulen = orphan ? S.count : S.len;
    S.flags = orphan ? F_ORPHAN : 0;
    S.iter = 0;
    S.evals = 0;
    for (int j = 0; j < 3; j++) { S.ref_pos[j] = r.ref_pos[j]; S.ref_vel[j] = r.ref_vel[j]; }
    if (S.ulen < P.min_part && r.snapshot_index_of_death == -1) r.snapshot_index_of_death = P.snap;
    if (S.ulen <= 1) {
      S.flags |= F_EARLY;
      S.nbound = S.ulen;
      S.nlast = S.ulen;
    } else {
      S.flags |= F_ACTIVE;
      S.nbound = S.ulen;
      S.nlast = S.ulen;
    }
    // the host shuffles exactly the lists longer than the sample (:302), with
    // draws counted for the same length (a list of another length would read
    // other segments' draws: it is not shuffled, and parity tests show it)
    if (S.draw_off >= 0 && S.ulen > P.max_sample && P.max_sample > 0 && S.ulen == S.shuf_len) S.flags |= F_SHUF;
    else S.draw_off = -1;
    W.ncount[NC_SRC * W.nseg + k] = S.draw_off >= 0 ? S.ulen - 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// Shuffle (std::random_shuffle at src/subhalo_unbind.cpp:302; libstdc++'s
// form: step i = 1..n-1 swaps positions i and j_i = rand() % (i+1)), replayed
// in parallel.  Before step i position i still holds element i, so the last
// step that moves an element into position p from above (the largest i > p
// with j_i = p) leaves element i there for good.  A position that no later
// step reaches keeps what step p left: element p if j_p = p, else what
// position q = j_p held just before step p -- the element of the latest step
// i in (q, p) with j_i = q, or, if none, what step q left there, and so on
// down the chain p -> j_p -> ...  Sorting the pairs (j_i, i) groups each
// position's movers in step order; pre = j_i, nlo = the previous mover into
// j_i (-1: none), nhi = the last mover into a position (-1: none), npos =
// the list before the shuffle.

// Keys (j_i << 32 | i) of steps i = 1..ulen-1 at sort positions i-1.
__global__ void __launch_bounds__(BLOCK) shuffle_keys_kernel(Work W, const int *__restrict__ draws) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    Seg &S = W.seg[k];
    if (S.draw_off < 0 || l >= S.ulen) continue;
    const int64_t b = S.ibase;
    W.npos[s] = W.gidx[s];
    W.nlo[s] = -1;
    W.nhi[s] = -1;
    if (l == 0) {
      W.pre[s] = 0;
      S.old_mb_id = W.g_id[W.gidx[s]];  // OldMostboundParticle is taken before the shuffle (:297)
      continue;
    }
    const int j = (int)((unsigned)draws[S.draw_off + l - 1] % (unsigned)(l + 1));
    W.pre[s] = j;
    W.k0[b + l - 1] = ((u64)(unsigned)j << 32) | (u64)(unsigned)l;
    W.v0[b + l - 1] = 0u;
  }
}

// Previous mover into the same position, and the last mover into each position.
__global__ void __launch_bounds__(BLOCK) shuffle_links_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int t = seg_lane_local(W, s, k);
    if (t < 0) continue;
    const Seg &S = W.seg[k];
    const int n = S.ulen - 1;  // sorted steps
    if (S.draw_off < 0 || t >= n) continue;
    const int64_t b = S.ibase;
    const u64 e = W.k0[s];
    const int q = (int)(e >> 32), i = (int)(e & 0xffffffffu);
    if (i == q) continue;  // a step that swaps a position with itself moves nothing
    int prev = -1;
    if (t > 0) {
      const u64 f = W.k0[s - 1];
      if ((int)(f >> 32) == q && (int)(f & 0xffffffffu) > q) prev = (int)(f & 0xffffffffu);
    }
    W.nlo[b + i] = prev;
    if (t + 1 == n || (int)(W.k0[s + 1] >> 32) != q) W.nhi[b + q] = i;
  }
}

// The shuffled list: position p takes the element the chain above names.
__global__ void __launch_bounds__(BLOCK) shuffle_apply_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int p = seg_lane_local(W, s, k);
    if (p < 0) continue;
    const Seg &S = W.seg[k];
    if (S.draw_off < 0 || p >= S.ulen) continue;
    const int64_t b = S.ibase;
    int src = W.nhi[s];
    if (src < 0) {
      int cur = p;  // strictly decreases along the chain
      while (true) {
        if (cur == 0) { src = 0; break; }
        const int q = W.pre[b + cur];
        if (q == cur) { src = cur; break; }
        const int w = W.nlo[b + cur];
        if (w >= 0) { src = w; break; }
        cur = q;
      }
    }
    W.gidx[s] = W.npos[b + src];
  }
}

// Spatial target order of a shuffled list (its sampled trees hold only the
// first MaxSampleSize entries): a sort of every particle's octant digits in
// the cube of the whole list.  Setup: the whole unbound list as the sources.
__global__ void prelude_setup_kernel(Work W) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    const bool on = S.draw_off >= 0 && (S.flags & F_ACTIVE);
    S.n_src = on ? S.ulen : 0;
    S.src_lo = 0;
    S.n_tgt = 0;
    for (int j = 0; j < 3; j++) { S.bmin[j] = 0xffffffffu; S.bmax[j] = 0u; }
    W.ncount[NC_SRC * W.nseg + k] = S.n_src;
  }
}

// After the digit sort: the order becomes the target list, and the whole
// list's periodic straddle bounds every later pass of the segment.
__global__ void __launch_bounds__(BLOCK) prelude_sp_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    Seg &S = W.seg[k];
    if (S.draw_off < 0 || !(S.flags & F_ACTIVE) || l >= S.ulen) continue;
    W.sp[s] = (int)W.v0[s];
    if (l == 0) S.sall = S.straddle;
  }
}

__global__ void el_init_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    if (l < W.seg[k].ulen) W.el[s] = l;
  }
}

// ---------------------------------------------------------------------------
// Pass: setup

__global__ void pass_setup_kernel(Work W, DevParams P) {
  if (blockIdx.x == 0 && threadIdx.x == 0) { W.active[0] = 0; W.active[1] = 0; W.active[2] = 0; }
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    W.segact[k] = 0;
    S.flags &= ~(F_FINAL | F_SAMP);
    S.mf = 1.f;
    S.n_src = 0; S.n_tgt = 0; S.src_lo = 0; S.nsort = 0; S.sort_lo = 0; S.n_items = 0;
    S.K = 0; S.nP = 0; S.nG = 0;
    if (!(S.flags & F_ACTIVE)) continue;
    S.iter++;
    if (S.flags & F_CORR) {
      // RelativeVelocity(OldRefPos, OldRefVel, RefPos, RefVel), dK (:314-316)
      rel_vel(P, S.old_pos[0], S.old_pos[1], S.old_pos[2], S.old_vel[0], S.old_vel[1],
              S.old_vel[2], S.ref_pos, S.ref_vel, S.rvd);
      S.dK = (float)(0.5 * (double)vnorm_f(S.rvd));
      S.src_lo = S.nbound;
      S.n_src = S.nlast - S.nbound;
      S.n_tgt = S.nbound;
      S.nlast = S.nbound;  // Nlast=Nbound after the correction loop (:329)
    } else {
      S.nlast = S.nbound;
      S.src_lo = 0;
      S.n_src = S.nlast;
      S.n_tgt = S.nlast;
      if (P.max_sample > 0 && S.nlast > P.max_sample) {
        // sampled tree: the first MaxSampleSize entries with their masses
        // scaled by Nlast / MaxSampleSize (float / long -> float, :334-339)
        S.n_src = P.max_sample;
        S.mf = (float)S.nlast / (float)P.max_sample;
        S.flags |= F_SAMP;
      }
    }
    S.evals += S.n_tgt;
    S.nsort = S.n_src;
    for (int j = 0; j < 3; j++) { S.bmin[j] = 0xffffffffu; S.bmax[j] = 0u; }
  }
  SEG_LOOP(k, W) {
    const Seg &S = W.seg[k];
    const bool act = (S.flags & F_ACTIVE) != 0;
    W.ncount[NC_SRC * W.nseg + k] = S.n_src;
    W.ncount[NC_PART * W.nseg + k] = act ? S.nlast : 0;
    W.ncount[NC_BOUND * W.nseg + k] = 0;
    W.ncount[NC_SORT * W.nseg + k] = 0;
  }
}

// ---------------------------------------------------------------------------
// Tree: bounding box, root cube, keys

// Bounding box of each segment's sources: a block reduces one tile, then one
// set of atomics per tile (ordered-integer encoding of the floats).
__global__ void __launch_bounds__(BLOCK) bbox_kernel(Work W, const Tile *tiles, int ntiles) {
  __shared__ unsigned int sh[6][BLOCK / WAVE];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    Seg &S = W.seg[tl.seg];
    if (threadIdx.x == 0) W.tcount[t] = 0;  // node lists of the pass's moments
    const int cnt = min(TILE, S.n_src - tl.start);
    if (cnt <= 0) continue;  // uniform
    const int64_t b = S.ibase;
    unsigned int mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (int i = threadIdx.x; i < cnt; i += BLOCK) {
      const int li = W.el[b + S.src_lo + tl.start + i];
      const float4 q = W.pos[b + li];
      const unsigned int e[3] = {f2ord(q.x), f2ord(q.y), f2ord(q.z)};
      for (int j = 0; j < 3; j++) { mn[j] = min(mn[j], e[j]); mx[j] = max(mx[j], e[j]); }
    }
    for (int d = WAVE / 2; d > 0; d >>= 1)
      for (int j = 0; j < 3; j++) {
        mn[j] = min(mn[j], (unsigned int)__shfl_down((int)mn[j], d, WAVE));
        mx[j] = max(mx[j], (unsigned int)__shfl_down((int)mx[j], d, WAVE));
      }
    if ((threadIdx.x & (WAVE - 1)) == 0)
      for (int j = 0; j < 3; j++) { sh[j][threadIdx.x / WAVE] = mn[j]; sh[3 + j][threadIdx.x / WAVE] = mx[j]; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int j = 0; j < 3; j++) {
        unsigned int a = sh[j][0], c = sh[3 + j][0];
        for (int w = 1; w < BLOCK / WAVE; w++) { a = min(a, sh[j][w]); c = max(c, sh[3 + j][w]); }
        atomicMin(&S.bmin[j], a);
        atomicMax(&S.bmax[j], c);
      }
    }
    __syncthreads();
  }
}

// One-wave tiles (up to R * WAVE sources, lane l taking sources l + WAVE r):
// a segment's only tile, so its box is stored directly.
template <int R>
__device__ __forceinline__ void bbox_wave_tile(const Work &W, const Tile &tl, int t) {
  const int lane = threadIdx.x & (WAVE - 1);
  {
    Seg &S = W.seg[tl.seg];
    if (lane == 0) W.tcount[t] = 0;
    const int cnt = min(R * WAVE, S.n_src);
    if (cnt <= 0) return;  // uniform across the wave
    const int64_t b = S.ibase;
    unsigned int mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int e = lane + WAVE * r;
      if (e < cnt) {
        const float4 q = W.pos[b + W.el[b + S.src_lo + e]];
        const unsigned int o[3] = {f2ord(q.x), f2ord(q.y), f2ord(q.z)};
        for (int j = 0; j < 3; j++) { mn[j] = min(mn[j], o[j]); mx[j] = max(mx[j], o[j]); }
      }
    }
    for (int d = WAVE / 2; d > 0; d >>= 1)
      for (int j = 0; j < 3; j++) {
        mn[j] = min(mn[j], (unsigned int)__shfl_xor((int)mn[j], d, WAVE));
        mx[j] = max(mx[j], (unsigned int)__shfl_xor((int)mx[j], d, WAVE));
      }
    if (lane == 0)
      for (int j = 0; j < 3; j++) { S.bmin[j] = mn[j]; S.bmax[j] = mx[j]; }
  }
}

// Tiles t0 <= t < tmid take four sources per lane, tmid <= t < ntiles one.
__global__ void __launch_bounds__(BLOCK) bbox_wave_kernel(Work W, const Tile *tiles, int t0, int tmid, int ntiles) {
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    if (t < tmid) bbox_wave_tile<4>(W, tl, t);
    else bbox_wave_tile<1>(W, tl, t);
  }
}

// Root cube exactly as OctTree_t::Build (src/oct_tree.tpp:118-139) and the
// depth schedule: node length halves while >= TreeNodeResolution
// (src/gravity_tree.cpp:37-40); a node created at depth >= max(1, dlen) is
// below the resolution floor (src/oct_tree.tpp:200).
__global__ void tree_params_kernel(Work W, DevParams P) {
  SEG_LOOP(k, W) {
#pragma clang fp contract(off)
    Seg &S = W.seg[k];
    W.ncount[NC_CROSS * W.nseg + k] = 0;
    if (S.n_src <= 0) continue;
    double xmin[3], xmax[3];
    for (int j = 0; j < 3; j++) { xmin[j] = ord2f(S.bmin[j]); xmax[j] = ord2f(S.bmax[j]); }
    double Len = xmax[0] - xmin[0];
    for (int j = 1; j < 3; j++)
      if ((xmax[j] - xmin[j]) > Len) Len = xmax[j] - xmin[j];
    for (int j = 0; j < 3; j++) S.center[j] = 0.5 * (xmax[j] + xmin[j]);
    S.Len = Len;
    int dlen = 0;
    while (dlen < 1100 && ldexp(Len, -dlen) >= (double)P.res) dlen++;
    S.dlen = dlen;
    int D = dlen < 1 ? 1 : dlen;
    if (D > MAX_DEPTH) { D = MAX_DEPTH; S.flags |= F_DEEP; }
    S.D = D;
    // a correction pass builds its tree over the block just removed, which
    // lies inside the latest full pass's sources but not inside the previous
    // correction pass's: the latest full pass bounds the depth of later passes
    if (!(S.flags & F_CORR)) S.Dfull = D;
    // the FP32 walk drops the periodic wrap unless some target-source pair can
    // be more than half a box apart: a correction pass's sources (the block
    // just removed) may lie on one side of the box edge while its targets lie
    // on both, so its flag also takes the latest full pass's, whose sources
    // hold both sets
    const int st = (P.periodic && Len > 0.5 * (double)P.box_half) ? 1 : 0;
    if (!(S.flags & F_CORR)) S.sfull = st;
    // a shuffled list's sampled sources may span less than its targets: the
    // straddle of the whole list (prelude) bounds every pass
    S.straddle = st | S.sfull | S.sall;
  }
}

// Octant digits of each source for depths 1..D: the reference's insertion
// descends with lenhalf *= 0.5 and centre +/- lenhalf in double and compares
// the float coordinate with strict '>' (src/oct_tree.tpp:151-184).  Digits
// are packed most significant first in bits 62..0.
__global__ void __launch_bounds__(BLOCK) keys_kernel(Work W) {
  SLOT_LOOP(s, W) {
#pragma clang fp contract(off)
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (l >= S.n_src) continue;
    const int64_t b = S.ibase;
    const int li = W.el[b + S.src_lo + l];
    const float4 q = W.pos[b + li];
    const float x = q.x, y = q.y, z = q.z;
    double c0 = S.center[0], c1 = S.center[1], c2 = S.center[2];
    double lh = 0.5 * S.Len;
    u64 key = 0;
    for (int d = 0; d < S.D; d++) {
      lh *= 0.5;
      int sub = 0;
      if (x > c0) { c0 += lh; sub |= 1; } else c0 -= lh;
      if (y > c1) { c1 += lh; sub |= 2; } else c1 -= lh;
      if (z > c2) { c2 += lh; sub |= 4; } else c2 -= lh;
      key = (key << 3) | (u64)sub;
    }
    key <<= 3 * (MAX_DEPTH - S.D);
    W.k0[s] = key;
    W.v0[s] = (u32)li;
    W.bits[s] = 0u;  // node-start masks, set by emit_kernel
  }
}

// ---------------------------------------------------------------------------
// Tree: structure from sorted keys.  Pair l joins sorted sources l and l+1;
// delta = number of leading digits they share (D when the keys are equal: the
// two are in one cell at the depth of the resolution floor).  A node of the
// compressed octree is a maximal run of sources whose pairs all have
// delta >= v, split by pairs with delta == v; it sits at depth v.

__device__ __forceinline__ int key_delta(u64 a, u64 b, int D) {
  if (a == b) return D;
  int d = (__clzll((long long)(a ^ b)) - 1) / 3;
  return d < D ? d : D;
}

// First v octant digits of a key (digit 0 in bits 62..60).
__device__ __forceinline__ u64 key_prefix(u64 k, int v) { return v == 0 ? 0ull : k >> (63 - 3 * v); }

// First index in [lo, hi+1] whose v-digit prefix is >= p (keys ascending).
__device__ __forceinline__ int prefix_lower(const u64 *key, int lo, int hi, int v, u64 p) {
  while (lo <= hi) {
    const int m = (lo + hi) >> 1;
    if (key_prefix(key[m], v) < p) lo = m + 1;
    else hi = m - 1;
  }
  return lo;
}

// Per pair l of sorted sources: its delta, and the node of which it is the
// leftmost split.  The neighbouring pairs' deltas come from the keys.  In a
// full pass the sorted order also becomes the spatial target order.
//
// The keys of a tile (and one on either side) are staged in LDS, so the
// binary searches for a node's first and last sources run there; a search
// goes to device memory only for the part of its range outside the tile (the
// nodes that extend past it).  KeyWin is that window: sources w_lo..w_hi.
struct KeyWin {
  const u64 *sk;   // LDS: keys of sources w_lo..w_hi
  const u64 *key;  // the segment's keys in device memory
  int w_lo, w_hi;
  __device__ __forceinline__ u64 at(int i) const { return (i >= w_lo && i <= w_hi) ? sk[i - w_lo] : key[i]; }
  // first index in [a, h + 1] whose v-digit prefix is >= p (keys ascending)
  __device__ __forceinline__ int lower(int a, int h, int v, u64 p) const {
    const int wa = a > w_lo ? a : w_lo, wh = h < w_hi ? h : w_hi;
    if (wa > wh) return prefix_lower(key, a, h, v, p);
    if (key_prefix(sk[wa - w_lo], v) >= p) return wa == a ? a : prefix_lower(key, a, wa - 1, v, p);
    if (key_prefix(sk[wh - w_lo], v) < p) return wh == h ? h + 1 : prefix_lower(key, wh + 1, h, v, p);
    int lo = wa + 1, hi = wh - 1;  // the answer is in [wa + 1, wh]
    while (lo <= hi) {
      const int m = (lo + hi) >> 1;
      if (key_prefix(sk[m - w_lo], v) < p) lo = m + 1;
      else hi = m - 1;
    }
    return lo;
  }
};

// One source l (slot s) of segment S: the work of emit for it.
__device__ __forceinline__ void emit_one(const Work &W, const Seg &S, const KeyWin &K, int l, int full_copy_sp) {
  const int n = S.n_src;
  const int64_t s = S.ibase + l;
  if (full_copy_sp && !(S.flags & (F_CORR | F_SAMP | F_REFINE))) W.sp[s] = (int)W.v0[s];
  W.nlo[s] = -1;
  const int v = (l + 1 < n) ? key_delta(K.at(l), K.at(l + 1), S.D) : -1;
  W.dl[s] = (signed char)v;
  if (v < 0) return;
  const int vl = l > 0 ? key_delta(K.at(l - 1), K.at(l), S.D) : -1;
  if (vl == v) return;  // source l-1 is in the node, in another child
  const int vr = l + 2 < n ? key_delta(K.at(l + 1), K.at(l + 2), S.D) : -1;
  const u64 p = key_prefix(K.at(l), v);
  int lo = l, hi = l + 1;
  if (vl > v) {
    lo = K.lower(0, l - 1, v, p);
    // leftmost split iff sources lo and l are in the same child
    if (key_delta(K.at(lo), K.at(l), S.D) <= v) return;
  }
  if (vr >= v) hi = K.lower(l + 2, n - 1, v, p + 1) - 1;
  W.nlo[s] = lo;
  W.nhi[s] = hi;
  atomicOr(&W.bits[S.ibase + lo], 1u << v);
}

// Workgroup form: one tile of TILE sources, window t0 - 1 .. t0 + TILE + 1.
__global__ void __launch_bounds__(BLOCK) emit_kernel(Work W, const Tile *__restrict__ tiles, int ntiles,
                                                     int full_copy_sp) {
  __shared__ u64 sk[TILE + 3];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const int n = S.n_src;
    const int cnt = min(TILE, n - tl.start);
    if (cnt <= 0) continue;  // uniform across the block
    KeyWin K{sk, W.k0 + S.ibase, tl.start > 0 ? tl.start - 1 : 0, min(n - 1, tl.start + TILE + 1)};
    for (int i = K.w_lo + (int)threadIdx.x; i <= K.w_hi; i += BLOCK) sk[i - K.w_lo] = K.key[i];
    __syncthreads();
    for (int i = threadIdx.x; i < cnt; i += BLOCK) emit_one(W, S, K, tl.start + i, full_copy_sp);
    __syncthreads();  // the next tile restages sk
  }
}

// One-wave form: a segment of at most SMALL_CAP sources, all of its keys in
// the wave's LDS.
__global__ void __launch_bounds__(BLOCK) emit_wave_kernel(Work W, const Tile *__restrict__ tiles, int t0,
                                                          int ntiles, int full_copy_sp) {
  __shared__ u64 sk[BLOCK / WAVE][SMALL_CAP];
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const int n = min(SMALL_CAP, S.n_src);
    if (n <= 0) continue;  // uniform across the wave
    KeyWin K{sk[wv], W.k0 + S.ibase, 0, n - 1};
    for (int i = lane; i < n; i += WAVE) sk[wv][i] = K.key[i];
    wave_sync();
    for (int i = lane; i < n; i += WAVE) emit_one(W, S, K, i, full_copy_sp);
    wave_sync();  // the next tile restages the wave's keys
  }
}

// Node length at depth v: Len halved while >= resolution, stored as float.
__device__ __forceinline__ float node_len(const Seg &S, int v) {
  const int e = v < S.dlen ? v : S.dlen;
  return (float)ldexp(S.Len, -e);
}

// Writes leaves and node headers in pre-order.  pre = exclusive prefix of
// popcount(bits) over the segment's sorted sources.
__global__ void __launch_bounds__(BLOCK) items_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l >= 0) {
      const Seg &S = W.seg[k];
      if (l < S.n_src) {
        const int64_t b = S.ibase, ib = 2 * b;
        const u32 bl = W.bits[s];
        const int pos = l + W.pre[s] + __popc(bl);
        const int li = (int)W.v0[s];
        // a sampled tree's leaves carry mass * MassFactor (EnergySnapshot_t::GetMass, :79)
        const float4 q = W.pos[b + li];
        const float m = (S.flags & F_SAMP) ? q.w * S.mf : q.w;
        W.items[ib + pos] = Item{make_float4(q.x, q.y, q.z, m),
                                 make_int2(0, pos + 1), make_int2(32 * (pos + 1), 0)};
        if (l + 1 == S.n_src) {
          W.seg[k].n_items = pos + 1;
          // sentinel after the last item: massless, never opened, its skip is
          // itself, so a group walk that has finished reads it and stays put
          W.items[ib + pos + 1] = Item{make_float4(0.f, 0.f, 0.f, 0.f), make_int2(0, pos + 1),
                                       make_int2(32 * (pos + 1), 0)};
        }
        const int lo = W.nlo[s];
        if (lo >= 0) {
          const int v = W.dl[s], hi = W.nhi[s];
          const u32 blo = W.bits[b + lo];
          const int npos = lo + W.pre[b + lo] + __popc(blo & ((1u << v) - 1u));
          const int hpos = hi + W.pre[b + hi] + __popc(W.bits[b + hi]);
          const float len = node_len(S, v);
          const float lsq = len * len;
          W.items[ib + npos] = Item{make_float4(0.f, 0.f, 0.f, 0.f), make_int2(__float_as_int(lsq), hpos + 1),
                                    make_int2(32 * (hpos + 1), 0)};
          W.npos[s] = (int)(ib + npos);
        }
      }
    }
  }
}

// Node moments: mass and centre of mass accumulated in double over the float
// values of the children in son order, stored as float
// (UpdateInternalNodes/ProcessNode/FillNodeCenter, src/gravity_tree.cpp:18-77).
// pm(i) and skip(i) read item i; g is the node, end its skip.
template <class Pm, class Skip>
__device__ __forceinline__ float4 sum_children(Pm pm, Skip skip, int g, int end) {
#pragma clang fp contract(off)
  double mass = 0., cx = 0., cy = 0., cz = 0.;
  int c = g + 1;
  while (c < end) {
    const float4 q = pm(c);
    const double m = q.w;
    mass += m;
    cx += (double)q.x * m;
    cy += (double)q.y * m;
    cz += (double)q.z * m;
    const int nx = skip(c);
    c = nx > c ? nx : end;  // skips always advance
  }
  return make_float4((float)(cx / mass), (float)(cy / mass), (float)(cz / mass), (float)mass);
}

// The same over the items in device memory.  Skips are segment-local item
// positions; ib is the segment's first item.
__device__ __forceinline__ void node_moments(const Work &W, int64_t ib, int g) {
  const Item *it = W.items;
  W.items[g].pm = sum_children([&](int c) { return it[c].pm; }, [&](int c) { return (int)(ib + it[c].ax.y); },
                               g, (int)(ib + W.items[g].ax.y));
}

// Node moments in two or three launches.  moments_tile_kernel: one workgroup
// per tile stages the tile's items in LDS (its leaves, the headers of the
// nodes whose leftmost leaf it holds: fewer than 2 TILE + MAX_DEPTH), computes
// the nodes whose leaves all lie in the tile deepest first (their children do
// too), depth by depth with a workgroup barrier between depths, and lists the
// pairs of the nodes whose leaves cross a tile edge (at most one per depth per
// edge): in W.cross, by super-tile (SUPER_TILES tiles), when the node lies in
// one, else in W.cross2 by segment.  moments_super_kernel: one workgroup per
// super-tile computes its listed nodes, deepest first; moments_cross_kernel
// (only for segments of more than one super-tile): one workgroup per segment
// computes the rest, deepest first.  A node's children are inside a tile (done
// by the first launch) or listed at the same or a lower level and deeper.
// Each node sums its children in son order (sum_children), so the results do
// not depend on which launch computes a node.
constexpr int TILE_ITEMS = 2 * TILE + MAX_DEPTH + 2;
constexpr int SUPER_TILES = 64;
constexpr int SUPER = SUPER_TILES * TILE;

// Lists the pair l (in tile t) of a node with leaves lo..hi that crosses a tile edge.
__device__ __forceinline__ void list_crossing(const Work &W, int t, const Tile &tl, int64_t b, int l, int lo,
                                              int hi, int *ncross) {
  if (lo / SUPER == hi / SUPER) {
    const int st = tl.start / SUPER * SUPER;  // first leaf of the super-tile
    const int ts = t - (tl.start - st) / TILE;  // its first tile in the table
    W.cross[b + st + atomicAdd(&W.tcount[ts], 1)] = l;
  } else {
    W.cross2[b + atomicAdd(ncross, 1)] = l;
  }
}

__global__ void __launch_bounds__(BLOCK) moments_tile_kernel(Work W, const Tile *__restrict__ tiles, int ntiles) {
  __shared__ float4 spm[TILE_ITEMS];
  __shared__ int sskip[TILE_ITEMS];
  __shared__ int snode[TILE];            // local nodes (LDS item index), grouped by depth
  __shared__ int sdepth[MAX_DEPTH + 2];  // nodes per depth, then the next free place
  __shared__ int sfirst[MAX_DEPTH + 2];  // first place of each depth
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const int n = S.n_src;
    const int cnt = min(TILE, n - 1 - tl.start);  // pairs l with l + 1 < n
    if (cnt <= 0) continue;  // uniform across the block
    const int64_t b = S.ibase, ib = 2 * b;
    const int t_lo = tl.start, t_hi = tl.start + TILE - 1;
    const int nl = min(TILE, n - tl.start);
    const int lastl = tl.start + nl - 1;
    const int first = tl.start + W.pre[b + tl.start];
    const int last = lastl + W.pre[b + lastl] + __popc(W.bits[b + lastl]);
    const int nit = last - first + 1;
    int *ncross = W.ncount + NC_CROSS * W.nseg + tl.seg;
    if (nit > TILE_ITEMS) {
      // not expected (leaves plus the headers of the nodes starting in the
      // tile); the same depth order over device memory
      for (int v = min(max(S.D, S.Dfull), MAX_DEPTH); v >= 0; v--) {
        for (int i = threadIdx.x; i < cnt; i += BLOCK) {
          const int64_t s = b + tl.start + i;
          if (W.dl[s] != v || W.nlo[s] < 0) continue;
          if (W.nlo[s] >= t_lo && W.nhi[s] <= t_hi) node_moments(W, ib, W.npos[s]);
          else list_crossing(W, t, tl, b, tl.start + i, W.nlo[s], W.nhi[s], ncross);
        }
        __syncthreads();
      }
      continue;  // uniform across the block
    }
    // one batch of loads per thread: its PER pairs (4-aligned slots; the
    // group stays inside the segment's wave-aligned slots) and its items
    constexpr int PER = TILE / BLOCK, IPER = (TILE_ITEMS + BLOCK - 1) / BLOCK;
    static_assert(PER == 4, "pairs are loaded four at a time");
    const int p0 = PER * (int)threadIdx.x;
    int plo[PER], phi[PER], ppos[PER], pdl[PER];
    if (p0 < cnt) {
      const int64_t s = b + tl.start + p0;
      const int4 a = *reinterpret_cast<const int4 *>(W.nlo + s);
      const int4 h = *reinterpret_cast<const int4 *>(W.nhi + s);
      const int4 g = *reinterpret_cast<const int4 *>(W.npos + s);
      const char4 d = *reinterpret_cast<const char4 *>(W.dl + s);
      plo[0] = a.x; plo[1] = a.y; plo[2] = a.z; plo[3] = a.w;
      phi[0] = h.x; phi[1] = h.y; phi[2] = h.z; phi[3] = h.w;
      ppos[0] = g.x; ppos[1] = g.y; ppos[2] = g.z; ppos[3] = g.w;
      pdl[0] = d.x; pdl[1] = d.y; pdl[2] = d.z; pdl[3] = d.w;
    }
#pragma unroll
    for (int j = 0; j < PER; j++)
      if (p0 + j >= cnt) plo[j] = -1;
    float4 qpm[IPER];
    int qsk[IPER];
#pragma unroll
    for (int k = 0; k < IPER; k++) {
      // past the tile's items the last one is read again (and not stored)
      const int i = min((int)threadIdx.x + k * BLOCK, nit - 1);
      qpm[k] = W.items[ib + first + i].pm;
      qsk[k] = W.items[ib + first + i].ax.y - first;
    }
#pragma unroll
    for (int k = 0; k < IPER; k++) {
      const int i = (int)threadIdx.x + k * BLOCK;
      if (i < nit) {
        spm[i] = qpm[k];
        sskip[i] = qsk[k];
      }
    }
    for (int v = threadIdx.x; v <= MAX_DEPTH + 1; v += BLOCK) sdepth[v] = 0;
    __syncthreads();
    int rank[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
      if (plo[j] < 0) continue;
      if (plo[j] >= t_lo && phi[j] <= t_hi) rank[j] = atomicAdd(&sdepth[pdl[j]], 1);
      else list_crossing(W, t, tl, b, tl.start + p0 + j, plo[j], phi[j], ncross);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int o = 0;
      for (int v = MAX_DEPTH; v >= 0; v--) { sfirst[v] = o; o += sdepth[v]; }
      sfirst[MAX_DEPTH + 1] = o;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++)
      if (plo[j] >= t_lo && phi[j] <= t_hi) snode[sfirst[pdl[j]] + rank[j]] = (int)(ppos[j] - ib) - first;
    __syncthreads();
    // depths deepest first: places sfirst[v] .. sfirst[v] + count, in order
    for (int v = MAX_DEPTH; v >= 0; v--) {
      const int p0 = sfirst[v], p1 = (v > 0 ? sfirst[v - 1] : sfirst[MAX_DEPTH + 1]);
      if (p0 == p1) continue;  // uniform
      for (int p = p0 + (int)threadIdx.x; p < p1; p += BLOCK) {
        const int g = snode[p];
        spm[g] = sum_children([&](int c) { return spm[c]; }, [&](int c) { return sskip[c]; }, g, sskip[g]);
      }
      __syncthreads();
    }
    for (int p = threadIdx.x; p < sfirst[MAX_DEPTH + 1]; p += BLOCK) {
      const int g = snode[p];
      W.items[ib + first + g].pm = spm[g];
    }
    __syncthreads();
  }
}

// The nodes of a list (pairs at list[0..ng)) of segment S, deepest first.
__device__ __forceinline__ void moments_of_list(const Work &W, const Seg &S, const int *list, int ng) {
  const int64_t b = S.ibase, ib = 2 * b;
  for (int v = min(max(S.D, S.Dfull), MAX_DEPTH); v >= 0; v--) {
    for (int i = threadIdx.x; i < ng; i += BLOCK) {
      const int64_t s = b + list[i];
      if (W.dl[s] == v) node_moments(W, ib, W.npos[s]);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BLOCK) moments_super_kernel(Work W, const Tile *__restrict__ tiles, int ntiles) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    if (tl.start % SUPER != 0) continue;  // a super-tile's first tile stands for it
    const int ng = W.tcount[t];
    if (ng == 0) continue;  // uniform
    const Seg &S = W.seg[tl.seg];
    moments_of_list(W, S, W.cross + S.ibase + tl.start, ng);
  }
}

__global__ void __launch_bounds__(BLOCK) moments_cross_kernel(Work W, const Tile *__restrict__ tiles, int ntiles) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    if (tl.start != 0) continue;  // the segment's first tile stands for it
    const int ng = W.ncount[NC_CROSS * W.nseg + tl.seg];
    if (ng == 0) continue;  // uniform
    const Seg &S = W.seg[tl.seg];
    moments_of_list(W, S, W.cross2 + S.ibase, ng);
  }
}

// One-wave tiles: the whole tree of a segment of at most R * WAVE sources
// (fewer than 2 R WAVE items) in LDS; lane l holds pairs l + WAVE r, and the
// depths that hold nodes are taken deepest first with a wave barrier between
// them.
template <int R>
__global__ void __launch_bounds__(BLOCK) moments_wave_kernel(Work W, const Tile *__restrict__ tiles, int t0,
                                                             int ntiles) {
  constexpr int NI = 2 * R * WAVE;
  __shared__ float4 spm[BLOCK / WAVE][NI];
  __shared__ int sskip[BLOCK / WAVE][NI];
  const int lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const int cnt = min(R * WAVE, S.n_src) - 1;  // pairs
    if (cnt <= 0) continue;  // uniform across the wave
    const int64_t b = S.ibase, ib = 2 * b;
    const int nit = S.n_items;
    int lo[R], pos[R], v[R];
    unsigned int dm = 0;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int e = lane + WAVE * r;
      lo[r] = -1;
      pos[r] = 0;
      v[r] = 0;
      if (e < cnt) {
        lo[r] = W.nlo[b + e];
        pos[r] = W.npos[b + e];
        v[r] = W.dl[b + e];
      }
      if (lo[r] >= 0) dm |= 1u << v[r];
    }
    for (int d = WAVE / 2; d > 0; d >>= 1) dm |= (unsigned int)__shfl_xor((int)dm, d, WAVE);
    if (nit > NI) {
      // not expected (a compressed octree has fewer nodes than leaves): the
      // same depth order over device memory
      while (dm) {
        const int d = 31 - __clz(dm);
        dm &= ~(1u << d);
#pragma unroll
        for (int r = 0; r < R; r++)
          if (lo[r] >= 0 && v[r] == d) node_moments(W, ib, pos[r]);
        __threadfence();
        wave_sync();
      }
      continue;
    }
    for (int i = lane; i < nit; i += WAVE) {
      const Item q = W.items[ib + i];
      spm[wv][i] = q.pm;
      sskip[wv][i] = q.ax.y;
    }
    wave_sync();
    while (dm) {  // uniform across the wave
      const int d = 31 - __clz(dm);
      dm &= ~(1u << d);
#pragma unroll
      for (int r = 0; r < R; r++)
        if (lo[r] >= 0 && v[r] == d) {
          const int g = pos[r] - (int)ib;
          spm[wv][g] = sum_children([&](int c) { return spm[wv][c]; }, [&](int c) { return sskip[wv][c]; }, g,
                                    sskip[wv][g]);
        }
      wave_sync();
    }
#pragma unroll
    for (int r = 0; r < R; r++)
      if (lo[r] >= 0) W.items[pos[r]].pm = spm[wv][pos[r] - (int)ib];
    wave_sync();  // the next tile reuses the wave's LDS
  }
}

// ---------------------------------------------------------------------------
// Walk.  Pre-order items: an internal node is opened when
// len^2 > r^2 * theta^2 (then the next item is its first child), otherwise its
// monopole is used and the walk continues at its skip pointer; a leaf has
// len^2 = 0 and skip = next item.

// Reference arithmetic (src/gravity_tree.cpp:79-164): FP64 walk over float data.
template <bool COUNT>
__device__ double walk_replay(const Work &W, const DevParams &P, int64_t ib, int n_items, float tx,
                              float ty, float tz, float tmass, unsigned &ni, unsigned &no) {
#pragma clang fp contract(off)
  const double h = 2.8 * (double)P.eps, hinv = 1.0 / h;
  double pot = (double)(float)(tmass / P.eps);
  const double x = tx, y = ty, z = tz;
  const int end = (int)(ib + n_items);
  int j = (int)ib;
  while (j < end) {
    const float4 q = W.items[j].pm;
    const int2 a = W.items[j].ax;
    double dx = (double)q.x - x, dy = (double)q.y - y, dz = (double)q.z - z;
    if (P.periodic) { dx = nearest_d(dx, P); dy = nearest_d(dy, P); dz = nearest_d(dz, P); }
    const double r2 = dx * dx + dy * dy + dz * dz;
    const float lsq = __int_as_float(a.x);
    if ((double)lsq > r2 * (double)P.theta2) { j++; if (COUNT) no++; continue; }
    const int nj = (int)(ib + a.y);
    j = nj > j ? nj : end;  // skips always advance
    if (COUNT) ni++;
    pot += kernel_term_d((double)q.w, sqrt(r2), h, hinv);
  }
  return pot * (double)P.G / (double)P.a;
}

// FP32 walk, the terms summed in FP64 (or in FP32 when !P.wide_sum), as the
// group walk's FastTerm sums them.  Straddling segments (a periodic subhalo
// wider than a quarter box) form the separation in double.
template <bool STRADDLE, bool COUNT>
__device__ double walk_fast(const Work &W, const DevParams &P, int64_t ib, int n_items, float tx,
                            float ty, float tz, float tmass, unsigned &ni, unsigned &no) {
#pragma clang fp contract(off)
  const float h = 2.8f * P.eps, h2 = h * h, hinv = 1.0f / h;
  const float self = tmass / P.eps;
  double accd = self;
  float accf = self;
  const int end = (int)(ib + n_items);
  int j = (int)ib;
  while (j < end) {
    const float4 q = W.items[j].pm;
    const int2 a = W.items[j].ax;
    float dx, dy, dz;
    if (STRADDLE) {
      dx = (float)nearest_d((double)q.x - (double)tx, P);
      dy = (float)nearest_d((double)q.y - (double)ty, P);
      dz = (float)nearest_d((double)q.z - (double)tz, P);
    } else {
      dx = q.x - tx; dy = q.y - ty; dz = q.z - tz;
    }
    const float r2 = pair_r2(dx, dy, dz);
    if (__int_as_float(a.x) > r2 * P.theta2) { j++; if (COUNT) no++; continue; }
    const int nj = (int)(ib + a.y);
    j = nj > j ? nj : end;  // skips always advance
    if (COUNT) ni++;
    const float t = kernel_term_f(q.w, r2, h2, hinv);
    if (P.wide_sum) accd += t;
    else accf += t;
  }
  return (P.wide_sum ? accd : (double)accf) * (double)P.G / (double)P.a;
}

template <bool COUNT>
__device__ __forceinline__ double walk_dispatch(const Work &W, const DevParams &P, const Seg &S,
                                                float tx, float ty, float tz, float tmass,
                                                unsigned &ni, unsigned &no) {
  const int64_t ib = 2 * S.ibase;
  if (P.replay) return walk_replay<COUNT>(W, P, ib, S.n_items, tx, ty, tz, tmass, ni, no);
  if (S.straddle) return walk_fast<true, COUNT>(W, P, ib, S.n_items, tx, ty, tz, tmass, ni, no);
  return walk_fast<false, COUNT>(W, P, ib, S.n_items, tx, ty, tz, tmass, ni, no);
}

__device__ __forceinline__ void flush_counts(const Work &W, unsigned ni, unsigned no, unsigned nt,
                                             unsigned ns = 0, unsigned nw = 0) {
  unsigned long long a = ni, b = no, c = nt, d = ns, e = nw;
  a = wave_reduce_sum(a); b = wave_reduce_sum(b); c = wave_reduce_sum(c); d = wave_reduce_sum(d);
  e = wave_reduce_sum(e);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    atomicAdd(&W.counters[0], a); atomicAdd(&W.counters[1], b); atomicAdd(&W.counters[2], c);
    atomicAdd(&W.counters[3], d); atomicAdd(&W.counters[6], e);
  }
}

// Counts of an instrumented group walk, per lane: summed sources, opened
// nodes, targets, lane slots spent by the groups (leader lanes), wave steps
// (lane 0 of each wave).
struct WalkCount {
  unsigned ni = 0, no = 0, nt = 0, ns = 0, nw = 0;
  __device__ void flush(const Work &W) const { flush_counts(W, ni, no, nt, ns, nw); }
};

// Group-shared traversal.  The wave's 64 targets (consecutive in spatial
// order, one segment) are split into groups of GROUP lanes; each group walks
// the pre-order items together: the group stands on item j, every lane whose
// own path reaches j decides for itself (open, or use the monopole and skip
// the subtree), and the group moves to j+1 if any of its lanes opened j, else
// to skip(j) -- lanes still skipping an accepted ancestor A have
// skip(A) >= skip(j), so no lane's next item is passed over.  Each lane sees
// exactly the items, decisions and summation order of its own walk, whatever
// GROUP is.  GROUP = 64: the item index is wave-uniform and items are read
// with scalar loads; smaller groups visit fewer items that none of their lanes
// needs, at the price of vector loads and of the wave running until its
// slowest group is done.
template <int GROUP>
__device__ __forceinline__ u64 group_lanes() {
  if (GROUP == WAVE) return ~0ull;
  const int base = (int)(threadIdx.x & (WAVE - 1)) & ~(GROUP - 1);
  return (GROUP >= WAVE ? ~0ull : ((1ull << (GROUP & 63)) - 1ull)) << base;
}

template <int GROUP>
__device__ __forceinline__ bool group_any(bool p, u64 lanes) {
  if (GROUP == WAVE) return __any(p);
  return (__ballot(p) & lanes) != 0ull;
}

// Lane K of the caller's aligned group of GROUP (8, 16 or 32) lanes, read by
// every lane of the group (ds_swizzle, bit mode: source lane = (lane & AND) |
// K within each half-wave).  Every lane of the wave must be active.
template <int GROUP, int K>
__device__ __forceinline__ int group_field(int v) {
  static_assert(GROUP >= 8 && GROUP <= 32 && K < 8, "group_field: 8 to 32 lanes");
  constexpr int AND = 0x1f & ~(GROUP - 1);
  return __builtin_amdgcn_ds_swizzle(v, AND | (K << 5));
}

// Per-target arithmetic of one item in FP32: sep2 = squared separation,
// opens = the opening test, add = the softened monopole term, the terms summed
// in FP64 (WIDE: the FAST precision; one conversion and one FP64 add per term,
// exact for a few thousand terms of one sign) or in FP32.  Straddling segments
// (a periodic subhalo wider than a quarter box) form the separation in double.
template <bool STRADDLE, bool WIDE>
struct FastTerm {
  using real = float;
  float h2, hinv;
  typename std::conditional<WIDE, double, float>::type acc;
  __device__ __forceinline__ void init(const DevParams &P, bool live, float tmass) {
#pragma clang fp contract(off)
    const float h = 2.8f * P.eps;
    h2 = h * h;
    hinv = 1.0f / h;
    acc = live ? tmass / P.eps : 0.f;
  }
  __device__ __forceinline__ float sep2(const DevParams &P, const float4 &q, float tx, float ty,
                                        float tz) const {
#pragma clang fp contract(off)
    float dx, dy, dz;
    if (STRADDLE) {
      dx = (float)nearest_d((double)q.x - (double)tx, P);
      dy = (float)nearest_d((double)q.y - (double)ty, P);
      dz = (float)nearest_d((double)q.z - (double)tz, P);
    } else {
      dx = q.x - tx; dy = q.y - ty; dz = q.z - tz;
    }
    return pair_r2(dx, dy, dz);
  }
  __device__ __forceinline__ bool opens(const DevParams &P, float lsq, float r2) const {
#pragma clang fp contract(off)
    return lsq > r2 * P.theta2;
  }
  // the same test as a wave lane mask (every lane active)
  __device__ __forceinline__ u64 opens_mask(const DevParams &P, float lsq, float r2) const {
#pragma clang fp contract(off)
    return __builtin_amdgcn_fcmpf(lsq, r2 * P.theta2, 2 /* ordered > */);
  }
  __device__ __forceinline__ void add(float m, float r2) {
#pragma clang fp contract(off)
    acc += kernel_term_f(m, r2, h2, hinv);
  }
  // add() for the lanes in `accept` without branching on it: kernel_term_f's
  // Newtonian form is computed for every lane, the spline only for accepting
  // lanes within the softening length (rare), and the others add zero
  __device__ __forceinline__ void add_masked(bool accept, u64, float m, float r2) {
#pragma clang fp contract(off)
    float term = -m * __builtin_amdgcn_rsqf(r2);
    if (accept && !(r2 >= h2)) term = kernel_term_f(m, r2, h2, hinv);
    const float added = accept ? term : 0.f;
    acc += added;
  }
  __device__ __forceinline__ double result() const { return (double)acc; }
};

// Reference arithmetic (src/gravity_tree.cpp:79-164): FP64 over float data.
struct ReplayTerm {
  using real = double;
  double h, hinv, pot;
  __device__ __forceinline__ void init(const DevParams &P, bool live, float tmass) {
#pragma clang fp contract(off)
    h = 2.8 * (double)P.eps;
    hinv = 1.0 / h;
    pot = live ? (double)(float)(tmass / P.eps) : 0.0;
  }
  __device__ __forceinline__ double sep2(const DevParams &P, const float4 &q, float tx, float ty,
                                         float tz) const {
#pragma clang fp contract(off)
    double dx = (double)q.x - (double)tx, dy = (double)q.y - (double)ty, dz = (double)q.z - (double)tz;
    if (P.periodic) { dx = nearest_d(dx, P); dy = nearest_d(dy, P); dz = nearest_d(dz, P); }
    return dx * dx + dy * dy + dz * dz;
  }
  __device__ __forceinline__ bool opens(const DevParams &P, float lsq, double r2) const {
#pragma clang fp contract(off)
    return (double)lsq > r2 * (double)P.theta2;
  }
  __device__ __forceinline__ u64 opens_mask(const DevParams &P, float lsq, double r2) const {
#pragma clang fp contract(off)
    return __builtin_amdgcn_fcmp((double)lsq, r2 * (double)P.theta2, 2 /* ordered > */);
  }
  __device__ __forceinline__ void add(float m, double r2) {
#pragma clang fp contract(off)
    pot += kernel_term_d((double)m, sqrt(r2), h, hinv);
  }
  __device__ __forceinline__ void add_masked(bool accept, u64, float m, double r2) {
    if (accept) add(m, r2);
  }
  __device__ __forceinline__ double result() const { return pot; }
};

// The items come in as a restrict-qualified kernel argument pointing at the
// segment's first item (wave-uniform; item positions and skips are
// segment-local): nothing in the walk kernel writes them, so reads at a
// group-uniform index compile to scalar loads when GROUP = 64, and both
// possible successors of item j (j+1 and skip(j)) are fetched while j is
// evaluated.  Smaller groups read with vector loads at a 32-bit byte offset
// from the scalar base, one dword of the item per lane, and spread the six
// fields over the group with crossbar swizzles; every lane of the wave takes
// part in every step (lanes off their path discard the result), the dwords of
// both possible successors are fetched as soon as skip(j) is known, the
// group's vote takes the group's field of the opened-lane mask, and the
// accepting lanes add their terms without a branch; a finished group sits on
// the sentinel item at `end`, whose skip is itself.  The 64-lane form
// clamps its reads to the segment's last item, a leaf whose skip is `end`, and
// its skips always advance (max with j+1).
template <int GROUP, bool COUNT, class Term>
__device__ double walk_group(const Item *__restrict__ it, const DevParams &P, int ib, int end,
                             bool live, float tx, float ty, float tz, float tmass, WalkCount &wc) {
  Term t;
  t.init(P, live, tmass);
  const u64 lanes = group_lanes<GROUP>();
  const bool leader = (threadIdx.x & (GROUP - 1)) == 0;
  int until = live ? ib : INT_MAX;  // lane is on its path at j iff until <= j
  int j = group_any<GROUP>(live, lanes) ? ib : end;
  const int last = end - 1 > ib ? end - 1 : ib;
  if constexpr (GROUP == WAVE) {
    j = __builtin_amdgcn_readfirstlane(j);
    float4 q = it[(unsigned)min(j, last)].pm;
    int2 a = it[(unsigned)min(j, last)].ax;
    while (j < end) {
      const int j1 = j + 1;
      const int js = max(a.y, j1);
      const unsigned c1 = (unsigned)min(j1, last), cs = (unsigned)min(js, last);
      const float4 q1 = it[c1].pm, qs = it[cs].pm;
      const int2 a1 = it[c1].ax, as = it[cs].ax;
      bool opened = false;
      if (COUNT && leader) wc.ns += GROUP;
      if (COUNT && (threadIdx.x & (WAVE - 1)) == 0) wc.nw++;
      if (until <= j) {
        const typename Term::real r2 = t.sep2(P, q, tx, ty, tz);
        opened = t.opens(P, __int_as_float(a.x), r2);
        if (opened) {
          if (COUNT) wc.no++;
        } else {
          t.add(q.w, r2);
          until = a.y;
          if (COUNT) wc.ni++;
        }
      }
      const bool go = __any(opened);
      j = go ? j1 : js;
      q = go ? q1 : qs;
      a = go ? a1 : as;
    }
  } else {
    // lane r of every 8 fetches dword r of its group's item (r = 5, 7 are
    // not used) and the group takes the six fields it needs from those lanes
    // through the LDS crossbar: 4 bytes per lane through the vector memory
    // pipe.  Positions are byte offsets from the segment's first item: J =
    // 32 j + the lane's dword offset (< 32), so J < 32 end iff j < end, and a
    // skip kept as 32 skip (the item's sb field) is <= J iff skip <= j.
    const char *base = (const char *)it;
    const unsigned dw = 4u * (threadIdx.x & 7u);
    // the group's lanes in a wave mask: the vote takes the group's field of a
    // lane mask with one 64-bit shift per lane (vector unit) instead of
    // folding the mask on the scalar unit
    const unsigned gbase = threadIdx.x & (WAVE - GROUP);
    constexpr u64 GFIELD = GROUP >= 64 ? ~0ull : ((1ull << (GROUP & 63)) - 1ull);
    const unsigned endb = 32u * (unsigned)end;
    unsigned J = 32u * (unsigned)j + dw;
    unsigned untilb = live ? 32u * (unsigned)ib : UINT_MAX;
    int v = *(const int *)(base + J);
    while (__any(J < endb)) {
      const bool on = J < endb;
      const float4 q = make_float4(__int_as_float(group_field<GROUP, 0>(v)), __int_as_float(group_field<GROUP, 1>(v)),
                                   __int_as_float(group_field<GROUP, 2>(v)), __int_as_float(group_field<GROUP, 3>(v)));
      const float lsq = __int_as_float(group_field<GROUP, 4>(v));
      const unsigned sb = (unsigned)group_field<GROUP, 6>(v);
      // J <= 32 end + dw throughout: every item's skip lies in (j, end], the sentinel's is end
      const unsigned J1 = J + 32u, JS = sb + dw;
      // both possible successors are fetched now, under the work on item j
      // (j + 1 = end + 1 reads the next region or the buffer's two spare items)
      const int v1 = *(const int *)(base + J1), vs = *(const int *)(base + JS);
      // lane masks straight from the compares (every lane is active here):
      // the group's vote runs on them without a round trip through VGPRs
      const u64 pmask = __builtin_amdgcn_uicmp(J, endb, 36 /* unsigned < */) &
                        __builtin_amdgcn_uicmp(untilb, J, 37 /* unsigned <= */);
      const typename Term::real r2 = t.sep2(P, q, tx, ty, tz);
      const u64 omask = pmask & t.opens_mask(P, lsq, r2);
      const u64 amask = pmask & ~omask;
      const bool accept = __builtin_amdgcn_inverse_ballot_w64(amask);
      if (COUNT) {
        if (on && leader) wc.ns += GROUP;
        if ((threadIdx.x & (WAVE - 1)) == 0) wc.nw++;
        wc.no += __builtin_amdgcn_inverse_ballot_w64(omask) ? 1u : 0u;
        wc.ni += accept ? 1u : 0u;
      }
      const bool go = ((omask >> gbase) & GFIELD) != 0ull;
      J = go ? J1 : JS;
      v = go ? v1 : vs;
      t.add_masked(accept, amask, q.w, r2);
      untilb = accept ? sb : untilb;
    }
  }
  return t.result();
}

template <int GROUP, bool COUNT, bool REPLAY>
__device__ __forceinline__ double walk_group_dispatch(const Item *__restrict__ it,
                                                      const DevParams &P, const Seg &S, int ib,
                                                      int end, bool live, float x, float y, float z,
                                                      float tm, WalkCount &wc) {
  if (REPLAY)
    return walk_group<GROUP, COUNT, ReplayTerm>(it, P, ib, end, live, x, y, z, tm, wc);
  if (!P.wide_sum) {
    if (S.straddle)
      return walk_group<GROUP, COUNT, FastTerm<true, false>>(it, P, ib, end, live, x, y, z, tm, wc);
    return walk_group<GROUP, COUNT, FastTerm<false, false>>(it, P, ib, end, live, x, y, z, tm, wc);
  }
  if (S.straddle)
    return walk_group<GROUP, COUNT, FastTerm<true, true>>(it, P, ib, end, live, x, y, z, tm, wc);
  return walk_group<GROUP, COUNT, FastTerm<false, true>>(it, P, ib, end, live, x, y, z, tm, wc);
}

// Lanes per group chosen per segment (S.wgroup, set from its size bin).  The
// straddling and uncompensated forms (rare; precision study) keep 64.
template <bool COUNT, bool REPLAY>
__device__ __forceinline__ double walk_binned(const Item *__restrict__ it, const DevParams &P,
                                              const Seg &S, int ib, int end, bool live, float x,
                                              float y, float z, float tm, WalkCount &wc) {
  if (REPLAY) {
    switch (S.wgroup) {
      case 8: return walk_group<8, COUNT, ReplayTerm>(it, P, ib, end, live, x, y, z, tm, wc);
      case 16: return walk_group<16, COUNT, ReplayTerm>(it, P, ib, end, live, x, y, z, tm, wc);
      case 32: return walk_group<32, COUNT, ReplayTerm>(it, P, ib, end, live, x, y, z, tm, wc);
      default: return walk_group<64, COUNT, ReplayTerm>(it, P, ib, end, live, x, y, z, tm, wc);
    }
  }
  if (!P.wide_sum || S.straddle)
    return walk_group_dispatch<64, COUNT, false>(it, P, S, ib, end, live, x, y, z, tm, wc);
  using T = FastTerm<false, true>;
  switch (S.wgroup) {
    case 8: return walk_group<8, COUNT, T>(it, P, ib, end, live, x, y, z, tm, wc);
    case 16: return walk_group<16, COUNT, T>(it, P, ib, end, live, x, y, z, tm, wc);
    case 32: return walk_group<32, COUNT, T>(it, P, ib, end, live, x, y, z, tm, wc);
    default: return walk_group<64, COUNT, T>(it, P, ib, end, live, x, y, z, tm, wc);
  }
}

// The target mass that removes a target's own softened self term: none in a
// correction pass (its tree holds the removed block only); in a sampled pass
// the scaled mass of the sampled entries and none for the rest (:344-349);
// otherwise the particle's mass m.  isb flags the sampled entries.
__device__ __forceinline__ float target_mass(const Work &W, const Seg &S, int64_t b, int li, float m) {
  if (S.flags & F_CORR) return 0.f;
  if (S.flags & F_SAMP) return W.isb[b + li] ? m * S.mf : 0.f;
  return m;
}

// A target's binding energy from its potential.
// Full pass: E = 0.5|dv|^2 + phi (BindingEnergy, src/subhalo_unbind.cpp:341-354).
// Correction pass: E += OldVel.RefVelDiff + dK - phi_removed (:319-328).
template <bool REPLAY>
__device__ __forceinline__ void finish_target(const Work &W, const DevParams &P, const Seg &S, int64_t b,
                                              int li, float x, float y, float z, double pot) {
#pragma clang fp contract(off)
  pot = pot * (double)P.G / (double)P.a;
  const float4 v = W.vel[b + li];
  const float vx = v.x, vy = v.y, vz = v.z;
  if (!(S.flags & F_CORR)) {
    float dv[3];
    rel_vel(P, x, y, z, vx, vy, vz, S.ref_pos, S.ref_vel, dv);
    // the refinement pass keeps the unbinding energies (:252-259)
    float *E = (S.flags & F_REFINE) ? W.Er : W.E;
    if (REPLAY) {
      E[b + li] = (float)((double)vnorm_f(dv) * 0.5 + pot);
    } else {
      const double kin = 0.5 * ((double)dv[0] * dv[0] + (double)dv[1] * dv[1] + (double)dv[2] * dv[2]);
      E[b + li] = (float)(kin + pot);
    }
  } else {
    float ov[3];
    rel_vel(P, x, y, z, vx, vy, vz, S.old_pos, S.old_vel, ov);
    const float kin = vdot_f(ov, S.rvd) + S.dK;
    W.E[b + li] = (float)((double)W.E[b + li] + ((double)kin - pot));
  }
}

// The targets of one wave-aligned chunk of slots, one per lane.
template <int GROUP, bool COUNT, bool REPLAY>
__device__ __forceinline__ void walk_chunk(const Work &W, const DevParams &P, const Item *__restrict__ it,
                                           int64_t s, WalkCount &wc) {
  // segments are wave-aligned: the whole wave belongs to one segment
  const int k = __builtin_amdgcn_readfirstlane(W.seg_of_slot[s]);
  const Seg &S = W.seg[k];
  const int64_t b = S.ibase;
  const int l = (int)(s - b);
  const bool live = l < S.n_tgt;
  if (!__any(live)) return;
  if (COUNT && live) wc.nt++;
  const int li = live ? W.sp[s] : 0;
  const float4 q = W.pos[b + li];
  const float x = q.x, y = q.y, z = q.z;
  const float tm = target_mass(W, S, b, li, q.w);
  const Item *seg_it = it + 2 * b;  // item positions and skips are segment-local
  const int end = S.n_items;
  double pot;
  if (GROUP == 0)
    pot = walk_binned<COUNT, REPLAY>(seg_it, P, S, 0, end, live, x, y, z, tm, wc);
  else
    pot = walk_group_dispatch<GROUP == 0 ? 64 : GROUP, COUNT, REPLAY>(seg_it, P, S, 0, end, live, x, y, z,
                                                                     tm, wc);
  if (live) finish_target<REPLAY>(W, P, S, b, li, x, y, z, pot);
}

// Binding energies with the group-shared walk (same results as walk_kernel
// for every group size).  GROUP = 0: lanes per group from each segment's bin.
template <int GROUP, bool COUNT, bool REPLAY>
__global__ void __launch_bounds__(BLOCK) walk_shared_kernel(Work W, DevParams P,
                                                            const Item *__restrict__ it, int xcd_run) {
  WalkCount wc;
  const int vb = xcd_block(xcd_run);
  for (int64_t i = (int64_t)vb * BLOCK + threadIdx.x; i < W.nact; i += (int64_t)gridDim.x * BLOCK)
    walk_chunk<GROUP, COUNT, REPLAY>(W, P, it, slot_at(W, i), wc);
  if (COUNT) wc.flush(W);
}

// Instrumented passes: the elements a segmented sort ordered (counters[8])
// and the elements times the merge levels each needed (counters[9]).
__global__ void sort_count_kernel(const int *n, int nseg, unsigned long long *counters) {
  unsigned long long e = 0, el = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nseg; k += gridDim.x * blockDim.x) {
    const int c = n[k] > 0 ? n[k] : 0;
    e += (unsigned long long)c;
    el += (unsigned long long)c * (unsigned long long)merge_levels(c);
  }
  e = wave_reduce_sum(e);
  el = wave_reduce_sum(el);
  if ((threadIdx.x & (WAVE - 1)) == 0 && (e | el)) {
    atomicAdd(&counters[8], e);
    atomicAdd(&counters[9], el);
  }
}

// Instrumented passes: the items of the trees the walk read (counters[5]).
__global__ void tree_items_kernel(Work W) {
  SEG_LOOP(k, W) {
    const Seg &S = W.seg[k];
    if (S.n_tgt > 0 && S.n_items > 0) atomicAdd(&W.counters[5], (unsigned long long)S.n_items);
  }
}

// Binding energies of the pass's targets, in spatial order.
// Full pass: E = 0.5|dv|^2 + phi (BindingEnergy, :341-354).
// Correction pass: E += OldVel.RefVelDiff + dK - phi_removed (:319-328).
template <bool COUNT>
__global__ void __launch_bounds__(BLOCK) walk_kernel(Work W, DevParams P) {
  unsigned ni = 0, no = 0, nt = 0;
  SLOT_LOOP(s, W) {
#pragma clang fp contract(off)
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (l >= S.n_tgt) continue;
    if (COUNT) nt++;
    const int64_t b = S.ibase;
    const int li = W.sp[s];
    const float4 q = W.pos[b + li], w = W.vel[b + li];
    const float x = q.x, y = q.y, z = q.z;
    const float vx = w.x, vy = w.y, vz = w.z;
    if (!(S.flags & F_CORR)) {
      const double pot = walk_dispatch<COUNT>(W, P, S, x, y, z, target_mass(W, S, b, li, q.w), ni, no);
      float dv[3];
      rel_vel(P, x, y, z, vx, vy, vz, S.ref_pos, S.ref_vel, dv);
      float *E = (S.flags & F_REFINE) ? W.Er : W.E;
      if (P.replay) {
        E[b + li] = (float)((double)vnorm_f(dv) * 0.5 + pot);
      } else {
        const double kin = 0.5 * ((double)dv[0] * dv[0] + (double)dv[1] * dv[1] + (double)dv[2] * dv[2]);
        E[b + li] = (float)(kin + pot);
      }
    } else {
      const double pot = walk_dispatch<COUNT>(W, P, S, x, y, z, 0.f, ni, no);
      float ov[3];
      rel_vel(P, x, y, z, vx, vy, vz, S.old_pos, S.old_vel, ov);
      const float kin = vdot_f(ov, S.rvd) + S.dK;
      W.E[b + li] = (float)((double)W.E[b + li] + ((double)kin - pot));
    }
  }
  if (COUNT) flush_counts(W, ni, no, nt);
}

// ---------------------------------------------------------------------------
// Partition (PartitionBindingEnergy, src/subhalo_unbind.cpp:21-58), in
// parallel.  Over positions 1..len-1, p_k is the k-th position from the front
// holding E>0 and g_k the k-th from the back holding E<0.  The two-pointer
// loop moves g_k to p_{k-1} (p_0 = 0) and p_k to g_k for k = 1..K, where K is
// the largest k with p_k < g_k; then g_{K+1} moves to p_K if g_{K+1} > p_K.
// Element 0 fills the last hole h, and Nbound = h + (E_0 < 0).

__global__ void __launch_bounds__(BLOCK) partition_rank_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    // only segments partitioned in this pass: the scan covers active segments only
    if (!(S.flags & F_ACTIVE) || l < 1 || l >= S.nlast) continue;
    const int64_t b = S.ibase;
    const float e = W.E[b + W.el[s]];
    const u64 x = W.scan64[s];
    if (e > 0.f) W.pk[b + (int)(x & 0xffffffffu) + 1] = l;
    if (e < 0.f) W.gk[b + S.nG - (int)(x >> 32)] = l;
  }
}

__global__ void __launch_bounds__(BLOCK) partition_k_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    Seg &S = W.seg[k];
    const int m = min(S.nP, S.nG);
    if (!(S.flags & F_ACTIVE) || l < 1 || l > m) continue;
    const int64_t b = S.ibase;
    if (W.pk[b + l] < W.gk[b + l] && (l == m || W.pk[b + l + 1] >= W.gk[b + l + 1])) S.K = l;
  }
}

__global__ void partition_hole_kernel(Work W) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE)) continue;
    const int64_t b = S.ibase;
    const int K = S.K;
    const int pK = K ? W.pk[b + K] : 0;
    const bool gextra = (K + 1 <= S.nG) && (W.gk[b + K + 1] > pK);
    S.hole = gextra ? W.gk[b + K + 1] : pK;
    S.nP = gextra ? K + 1 : K;  // number of g moves (reuse nP)
  }
}

__global__ void __launch_bounds__(BLOCK) partition_move_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE) || l >= S.nlast) continue;
    const int64_t b = S.ibase;
    int dest = l;
    if (l == 0) {
      dest = S.hole;
    } else {
      const float e = W.E[b + W.el[s]];
      const u64 x = W.scan64[s];
      if (e > 0.f) {
        const int r = (int)(x & 0xffffffffu) + 1;
        if (r <= S.K) dest = W.gk[b + r];
      } else if (e < 0.f) {
        const int r = S.nG - (int)(x >> 32);
        if (r <= S.nP) dest = r == 1 ? 0 : W.pk[b + r - 1];
      }
    }
    W.el2[b + dest] = W.el[s];
  }
}

// Nbound after the partition; disruption test (:357-379).
__global__ void partition_done_kernel(Work W, DevParams P) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    S.ncopy = (S.flags & F_ACTIVE) ? S.nlast : 0;  // bound_flags_kernel copies the new Elist back
    if (!(S.flags & F_ACTIVE)) continue;
    const int64_t b = S.ibase;
    // the original element 0 ends in the last hole (Etmp, :37-38 and :50-51);
    // the partitioned Elist is still in el2
    int nb = S.hole + (W.E[b + W.el2[b + S.hole]] < 0.f ? 1 : 0);
    if (S.nlast == 0) nb = 0;
    S.nbound = nb;
    if (nb < P.min_part) {
      S.flags = (S.flags & ~F_ACTIVE) | F_DISR | F_FINAL;
      S.nbound = 1;
      S.nlast = 1;
    }
    S.nsort = 0;  // the energy sort follows the frames
    W.ncount[NC_BOUND * W.nseg + k] = (S.flags & F_ACTIVE) ? S.nlast : 0;
  }
}

// A shuffled list that is disrupted puts its old most-bound particle (the
// first list entry before the shuffle) back in front: the first particle
// with its ID swaps with position 0 (:363-370).  Find, then swap.
__global__ void __launch_bounds__(BLOCK) disrupt_find_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    Seg &S = W.seg[k];
    if (!(S.flags & F_SHUF) || (S.flags & (F_DISR | F_FINAL)) != (F_DISR | F_FINAL) || l >= S.ulen) continue;
    if (W.g_id[W.gidx[s]] == S.old_mb_id) atomicMin(&S.swapq, l);
  }
}

__global__ void disrupt_swap_kernel(Work W) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_SHUF) || (S.flags & (F_DISR | F_FINAL)) != (F_DISR | F_FINAL)) continue;
    const int q = S.swapq;
    S.swapq = INT_MAX;
    if (q == INT_MAX || q == 0) continue;
    const int64_t a = S.ibase, c = S.ibase + q;
    auto sw = [](auto *v, int64_t i, int64_t j) { const auto t = v[i]; v[i] = v[j]; v[j] = t; };
    sw(W.gidx, a, c); sw(W.pos, a, c); sw(W.vel, a, c);
  }
}

// Sampled pass: isb flags the sampled entries Elist[0, MaxSampleSize) (read
// by the walk for the self term; the partition then reuses isb).
__global__ void __launch_bounds__(BLOCK) sample_flags_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (!(S.flags & F_SAMP) || l >= S.nlast) continue;
    W.isb[S.ibase + W.el[s]] = l < S.n_src ? 1 : 0;
  }
}

// The partitioned Elist back from el2 (for every segment partitioned in the
// pass, including those it disrupted), and for the segments still active the
// spatial list's flags: isb[l] = particle l is still bound.
__global__ void __launch_bounds__(BLOCK) bound_flags_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (l >= S.ncopy) continue;
    const int li = W.el2[s];
    W.el[s] = li;
    if ((S.flags & F_ACTIVE) && l < S.nlast) W.isb[S.ibase + li] = l < S.nbound ? 1 : 0;
  }
}

__global__ void __launch_bounds__(BLOCK) spatial_compact_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE) || l >= S.nlast) continue;
    const int li = W.sp[s];
    if (W.isb[S.ibase + li]) W.el2[S.ibase + W.pre[s]] = li;
  }
}

__global__ void __launch_bounds__(BLOCK) spatial_copy_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE) || l >= S.nbound) continue;
    W.sp[s] = W.el2[s];
  }
}

// Sort keys for Elist[sort_lo, sort_lo + nsort) by energy: (block, ordered E,
// position) -> local index.  The block bit keeps the bound block Elist[0, nbound)
// and the removed block Elist[nbound, nlast) apart, so one segmented sort
// orders each block by energy in place; position breaks ties (a stable order).
__global__ void __launch_bounds__(BLOCK) energy_keys_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (l >= S.nsort) continue;
    const int64_t b = S.ibase;
    const int at = S.sort_lo + l;
    const int li = W.el[b + at];
    const u64 block = at >= S.nbound ? 1ull : 0ull;
    W.k0[s] = (block << 63) | ((u64)f2ord(W.E[b + li]) << 31) | (u64)(unsigned)at;
    W.v0[s] = (u32)li;
  }
}

__global__ void __launch_bounds__(BLOCK) energy_writeback_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (l >= S.nsort) continue;
    W.el[S.ibase + S.sort_lo + l] = (int)W.v0[s];
  }
}

// ---------------------------------------------------------------------------
// FAST frame sums, tile-parallel: every tile of an active segment sums its part
// of the bound block in FP64 (mass, m*v, m*x relative to the periodic origin);
// frames_combine_kernel adds the tiles of each segment in tile order, so the
// result does not depend on scheduling.
constexpr int FRAME_SUMS = 7;

// The frame sums of bound entry li: mass, momentum, mass-weighted position
// (relative to the most-bound entry o when periodic).
__device__ __forceinline__ void frame_accum(const Work &W, const DevParams &P, int64_t b, int li,
                                            const double o[3], double v[FRAME_SUMS]) {
#pragma clang fp contract(off)
  const float4 q = W.pos[b + li], w = W.vel[b + li];
  const float m = q.w;
  v[0] += m;
  v[1] += w.x * m; v[2] += w.y * m; v[3] += w.z * m;
  if (P.periodic) {
    v[4] += nearest_d((double)q.x - o[0], P) * (double)m;
    v[5] += nearest_d((double)q.y - o[1], P) * (double)m;
    v[6] += nearest_d((double)q.z - o[2], P) * (double)m;
  } else {
    v[4] += q.x * m; v[5] += q.y * m; v[6] += q.z * m;
  }
}

__global__ void __launch_bounds__(BLOCK)
frames_partial_kernel(Work W, DevParams P, const Tile *tiles, int ntiles, double *part) {
#pragma clang fp contract(off)
  __shared__ double sh[FRAME_SUMS][BLOCK / WAVE];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const int cnt = (S.flags & F_ACTIVE) ? min(TILE, S.nbound - tl.start) : 0;
    if (cnt <= 0) continue;  // uniform
    const int64_t b = S.ibase;
    const float4 p0 = W.pos[b + W.el[b]];
    const double o[3] = {p0.x, p0.y, p0.z};
    double v[FRAME_SUMS] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < cnt; i += BLOCK) frame_accum(W, P, b, W.el[b + tl.start + i], o, v);
    for (int c = 0; c < FRAME_SUMS; c++) {
      const double r = wave_total(v[c]);
      if ((threadIdx.x & (WAVE - 1)) == 0) sh[c][threadIdx.x / WAVE] = r;
    }
    __syncthreads();
    if (threadIdx.x < FRAME_SUMS) {
      double r = 0;
      for (int w = 0; w < BLOCK / WAVE; w++) r += sh[threadIdx.x][w];
      part[(int64_t)t * FRAME_SUMS + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

// One-wave tiles of up to R * WAVE entries: pass r of the wave takes entries
// WAVE r .. WAVE r + WAVE - 1, lane l entry WAVE r + l, as wave r of the
// workgroup form does, and the totals of the passes are added to zero in
// order, as the workgroup form adds its waves' totals, so both give the same
// bits.
template <int R>
__device__ __forceinline__ void frames_wave_tile(const Work &W, const DevParams &P, const Tile &tl, int t,
                                                 double *part) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (WAVE - 1);
  {
    const Seg &S = W.seg[tl.seg];
    const int cnt = (S.flags & F_ACTIVE) ? min(R * WAVE, S.nbound) : 0;
    if (cnt <= 0) return;  // uniform across the wave
    const int64_t b = S.ibase;
    const float4 p0 = W.pos[b + W.el[b]];
    const double o[3] = {p0.x, p0.y, p0.z};
    double acc[FRAME_SUMS] = {0, 0, 0, 0, 0, 0, 0};
    for (int r = 0; r < R; r++) {
      double v[FRAME_SUMS] = {0, 0, 0, 0, 0, 0, 0};
      const int e = lane + WAVE * r;
      if (e < cnt) frame_accum(W, P, b, W.el[b + e], o, v);
      for (int c = 0; c < FRAME_SUMS; c++) acc[c] += wave_total(v[c]);
    }
    if (lane == 0)
      for (int c = 0; c < FRAME_SUMS; c++) part[(int64_t)t * FRAME_SUMS + c] = acc[c];
  }
}

// Tiles t0 <= t < tmid take four entries per lane, tmid <= t < ntiles one.
__global__ void __launch_bounds__(BLOCK)
frames_partial_wave_kernel(Work W, DevParams P, const Tile *tiles, int t0, int tmid, int ntiles, double *part) {
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    if (t < tmid) frames_wave_tile<4>(W, P, tl, t, part);
    else frames_wave_tile<1>(W, P, tl, t, part);
  }
}

// Correction switch, new frame, convergence (:380-421) for one segment, given
// the FP64 sums over its bound block.
__device__ void frame_update(Work &W, const DevParams &P, Seg &S, double msum, double s1, double s2,
                             double s3, double q1, double q2, double q3) {
#pragma clang fp contract(off)
  const int64_t b = S.ibase;
  const int nb = S.nbound;
  const int l0 = W.el[b];
  const float4 p0 = W.pos[b + l0], v0 = W.vel[b + l0];
  const double o[3] = {p0.x, p0.y, p0.z};
  // correction switch (uses the frame of the pass just done); with sampling
  // only when fewer than MaxSampleSize were removed (:383-389)
  if (S.nlast - nb < nb && (P.max_sample <= 0 || S.nlast - nb < P.max_sample)) {
    S.flags |= F_CORR;
    for (int j = 0; j < 3; j++) { S.old_pos[j] = S.ref_pos[j]; S.old_vel[j] = S.ref_vel[j]; }
  }
  if (nb == 0) {
    msum = 0.;
  } else if (nb == 1) {
    S.ref_vel[0] = v0.x; S.ref_vel[1] = v0.y; S.ref_vel[2] = v0.z;
    msum = p0.w;
    S.ref_pos[0] = p0.x; S.ref_pos[1] = p0.y; S.ref_pos[2] = p0.z;
  } else {
    S.ref_vel[0] = (float)(s1 / msum); S.ref_vel[1] = (float)(s2 / msum); S.ref_vel[2] = (float)(s3 / msum);
    double q[3] = {q1 / msum, q2 / msum, q3 / msum};
    for (int j = 0; j < 3; j++) {
      if (P.periodic) q[j] += o[j];
      S.ref_pos[j] = (float)q[j];
    }
  }
  S.mbound = (float)msum;
  // energy sort of the removed block Elist[nb, nlast) (:359-360) and, on
  // convergence, of the bound block Elist[0, nb) too (:401)
  if ((float)nb >= (float)S.nlast * P.bmp) {
    // RefineBindingEnergyOrder (:406-407) runs after the level's passes; the
    // outputs wait for it
    if (P.refine && nb > P.max_sample) {
      S.flags = (S.flags & ~F_ACTIVE) | F_CONV | F_REFPEND;
      atomicAdd(&W.active[3], 1);
    } else {
      S.flags = (S.flags & ~F_ACTIVE) | F_CONV | F_FINAL;
    }
    S.sort_lo = 0;
    S.nsort = S.nlast;
  } else {
    S.sort_lo = nb;
    S.nsort = S.nlast - nb;
    atomicAdd(&W.active[0], 1);
    atomicMax(&W.active[1], S.nlast);  // bounds every count of the next pass
    W.segact[&S - W.seg] = 1;
  }
  W.ncount[NC_SORT * W.nseg + (int)(&S - W.seg)] = S.nsort;
}

// v[c] += part[(t0 + i * step) * C + c] for i = 0 .. cnt-1, in that order, with
// the values of the next few tiles loaded ahead of their use.
template <int C>
__device__ __forceinline__ void sum_tiles_in_order(const double *__restrict__ part, int t0, int cnt, int step,
                                                   double v[C]) {
#pragma clang fp contract(off)
  constexpr int PREF = 4;
  int i = 0;
  for (; i + PREF <= cnt; i += PREF) {
    double x[PREF][C];
#pragma unroll
    for (int u = 0; u < PREF; u++)
#pragma unroll
      for (int c = 0; c < C; c++) x[u][c] = part[(int64_t)(t0 + (i + u) * step) * C + c];
#pragma unroll
    for (int u = 0; u < PREF; u++)
#pragma unroll
      for (int c = 0; c < C; c++) v[c] += x[u][c];
  }
  for (; i < cnt; i++)
#pragma unroll
    for (int c = 0; c < C; c++) v[c] += part[(int64_t)(t0 + i * step) * C + c];
}

// Segments whose bound block spans more than SUPER_TILES tiles: one thread per
// run of SUPER_TILES tiles sums the run's partials in tile order into gpart
// (at the run's first tile), so that frames_combine_kernel adds runs.
__global__ void __launch_bounds__(BLOCK) frames_runs_kernel(Work W, const Tile *__restrict__ tiles, int ntiles,
                                                            const double *part, double *gpart) {
#pragma clang fp contract(off)
  for (int t = blockIdx.x * BLOCK + threadIdx.x; t < ntiles; t += gridDim.x * BLOCK) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const int nt = (S.flags & F_ACTIVE) ? (S.nbound + TILE - 1) / TILE : 0;
    const int r = tl.start / TILE;  // tile number within the segment
    if (nt <= SUPER_TILES || r >= nt || r % SUPER_TILES != 0) continue;
    double v[FRAME_SUMS] = {0, 0, 0, 0, 0, 0, 0};
    sum_tiles_in_order<FRAME_SUMS>(part, t, min(SUPER_TILES, nt - r), 1, v);
    for (int c = 0; c < FRAME_SUMS; c++) gpart[(int64_t)t * FRAME_SUMS + c] = v[c];
  }
}

// One thread per segment: the wave's updates of the active count and of the
// largest Nlast (frame_update) then combine into one atomic each.  (One wave
// per segment, as output_combine_kernel does, measured slower here.)  A bound
// block of more than SUPER_TILES tiles is summed over the runs of
// frames_runs_kernel.
__global__ void frames_combine_kernel(Work W, DevParams P, const int *first_tile, const double *part,
                                      const double *gpart) {
#pragma clang fp contract(off)
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE)) continue;
    double v[FRAME_SUMS] = {0, 0, 0, 0, 0, 0, 0};
    const int nt = (S.nbound + TILE - 1) / TILE;
    if (nt <= SUPER_TILES)
      sum_tiles_in_order<FRAME_SUMS>(part, first_tile[k], nt, 1, v);
    else
      sum_tiles_in_order<FRAME_SUMS>(gpart, first_tile[k], (nt + SUPER_TILES - 1) / SUPER_TILES, SUPER_TILES, v);
    frame_update(W, P, S, v[0], v[1], v[2], v[3], v[4], v[5], v[6]);
  }
}

// ---------------------------------------------------------------------------
// Frame update (:380-421) with the reference's summation order (REPLAY): one
// thread per segment sums the bound block serially in Elist order.
__global__ void frames_serial_kernel(Work W, DevParams P) {
#pragma clang fp contract(off)
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE)) continue;
    const int64_t b = S.ibase;
    const int nb = S.nbound;
    const float4 p0 = W.pos[b + W.el[b]];
    const double o[3] = {p0.x, p0.y, p0.z};
    double msum = 0, s1 = 0, s2 = 0, s3 = 0, q1 = 0, q2 = 0, q3 = 0;
    if (nb > 1) {
      for (int i = 0; i < nb; i++) {
        const int li = W.el[b + i];
        const float4 q = W.pos[b + li], w = W.vel[b + li];
        const float m = q.w;
        msum += m;
        s1 += w.x * m; s2 += w.y * m; s3 += w.z * m;
        if (P.periodic) {
          q1 += nearest_d((double)q.x - o[0], P) * (double)m;
          q2 += nearest_d((double)q.y - o[1], P) * (double)m;
          q3 += nearest_d((double)q.z - o[2], P) * (double)m;
        } else {
          q1 += q.x * m; q2 += q.y * m; q3 += q.z * m;
        }
      }
    }
    frame_update(W, P, S, msum, s1, s2, s3, q1, q2, q3);
  }
}

// ---------------------------------------------------------------------------
// Outputs for segments that finished in this pass, early exits and orphans.

// per-slot: output list and energies
__global__ void __launch_bounds__(BLOCK) output_list_kernel(Work W, int early) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    const bool fin = early ? (S.flags & F_EARLY) != 0 : (S.flags & F_FINAL) != 0;
    if (!fin || l >= S.len) continue;
    const int64_t o = S.out_off + l;
    if ((S.flags & F_CONV) && !(S.flags & F_ORPHAN)) {
      const int li = l < S.ulen ? W.el[s] : l;
      W.out_index[o] = W.gidx[S.ibase + li];
      W.out_energy[o] = l < S.ulen ? W.E[S.ibase + li] : __builtin_nanf("");
    } else {
      // disrupted, early, orphan: the list keeps its order -- an orphan's
      // own list as it came in, even when a shuffled copy of it was unbound
      // (:434-446)
      const bool orig = (S.flags & F_ORPHAN) && (S.flags & F_SHUF) && l < S.count;
      W.out_index[o] = orig ? own_entry(W, S, l) : W.gidx[s];
      W.out_energy[o] = __builtin_nanf("");
    }
  }
}

// Writes the record of a segment that finished, from its sums over the bound
// block: kinematics (AverageKinematics, :189-232), particle types
// (CountParticleTypes / CountParticles, src/subhalo.cpp:400-499) and the
// state markers of Unbind (:269-421).
__device__ void write_record(Work &W, const DevParams &P, Seg &S, double E, double K, double A0,
                             double A1, double A2, double M, const int nt[HBT_TYPE_MAX],
                             const float mts[HBT_TYPE_MAX]) {
#pragma clang fp contract(off)
  hbt_subhalo &r = W.subs[S.sub];
  const int64_t b = S.ibase;
  const bool conv = (S.flags & F_CONV) != 0;
  const int nb = S.nbound;
  if (S.flags & F_EARLY) {
    float mb = 0.f;  // CountParticles: Mbound = float sum of MboundType
    for (int t = 0; t < HBT_TYPE_MAX; t++) mb += mts[t];
    r.mbound = mb;
    r.status = HBT_STATUS_EARLY;
  } else {
    if (nb <= 1) {
      r.specific_self_potential_energy = 0.f;
      r.specific_self_kinetic_energy = 0.f;
      r.specific_angular_momentum[0] = r.specific_angular_momentum[1] = r.specific_angular_momentum[2] = 0.f;
    } else {
      E /= M;
      K *= 0.5 / M;
      r.specific_self_potential_energy = (float)(E - K);
      r.specific_self_kinetic_energy = (float)K;
      r.specific_angular_momentum[0] = (float)(A0 / M);
      r.specific_angular_momentum[1] = (float)(A1 / M);
      r.specific_angular_momentum[2] = (float)(A2 / M);
    }
    if (conv) {
      r.status = HBT_STATUS_CONVERGED;
      if (r.snapshot_index_of_death != -1) r.snapshot_index_of_death = -1;
      if (r.sink_track_id != -1) { r.snapshot_index_of_sink = -1; r.sink_track_id = -1; }
      for (int j = 0; j < 3; j++) { r.ref_pos[j] = S.ref_pos[j]; r.ref_vel[j] = S.ref_vel[j]; }
      const int l0 = W.el[b];
      const float4 q = W.pos[b + l0], w = W.vel[b + l0];
      r.mostbound_pos[0] = q.x; r.mostbound_pos[1] = q.y; r.mostbound_pos[2] = q.z;
      r.mostbound_vel[0] = w.x; r.mostbound_vel[1] = w.y; r.mostbound_vel[2] = w.z;
      r.mbound = S.mbound;
    } else {
      // disruption (:361-379): frame <- previous most-bound coordinates
      r.status = HBT_STATUS_DISRUPTED;
      if (r.snapshot_index_of_death == -1) r.snapshot_index_of_death = P.snap;
      for (int j = 0; j < 3; j++) { r.ref_pos[j] = r.mostbound_pos[j]; r.ref_vel[j] = r.mostbound_vel[j]; }
      r.mbound = W.pos[b].w;
    }
  }
  for (int t = 0; t < HBT_TYPE_MAX; t++) { r.nbound_type[t] = nt[t]; r.mbound_type[t] = mts[t]; }
  r.nbound = nb;
  r.list_len = S.len;
  long long ns = nb <= 1 ? nb : (long long)((float)nb * P.relax);
  if (ns > S.len) ns = S.len;
  r.nsource = ns;
  r.n_iter = S.iter;
  r.n_evals = S.evals;
  S.flags &= ~F_FINAL;
}

// Kinematics of one bound particle (AverageKinematics, :205-224).
__device__ __forceinline__ void kin_accum(const Work &W, const DevParams &P, const Seg &S, int li,
                                          double v[6]) {
#pragma clang fp contract(off)
  const int64_t b = S.ibase;
  const float4 q = W.pos[b + li], w = W.vel[b + li];
  const float m = q.w;
  v[0] += W.E[b + li] * m;
  const float x[3] = {q.x, q.y, q.z};
  const float vel[3] = {w.x, w.y, w.z};
  double dx[3], dv[3];
  for (int j = 0; j < 3; j++) {
    dx[j] = x[j] - S.ref_pos[j];
    if (P.periodic) dx[j] = nearest_d(dx[j], P);
    dx[j] *= P.a;
    dv[j] = (double)(vel[j] - S.ref_vel[j]) + P.hz * dx[j];
    v[1] += dv[j] * dv[j] * m;
  }
  v[2] += (dx[1] * dv[2] - dx[2] * dv[1]) * m;
  v[3] += (dx[2] * dv[0] - dx[0] * dv[2]) * m;
  v[4] += (dx[0] * dv[1] - dx[1] * dv[0]) * m;
  v[5] += m;
}

// Serial form (REPLAY and early exits): one thread per finished segment,
// sums in list order as the reference does with one thread.
__global__ void output_record_kernel(Work W, DevParams P, int early) {
#pragma clang fp contract(off)
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    const bool fin = early ? (S.flags & F_EARLY) != 0 : (S.flags & F_FINAL) != 0;
    if (!fin) continue;
    const int64_t b = S.ibase;
    const bool conv = (S.flags & F_CONV) != 0;
    const int nb = S.nbound;
    const bool skip_null = !(S.flags & F_EARLY) && nb <= 100;
    double v[6] = {0, 0, 0, 0, 0, 0};
    int nt[HBT_TYPE_MAX] = {0, 0, 0, 0, 0, 0};
    float mts[HBT_TYPE_MAX] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < nb; i++) {
      const int li = conv ? W.el[b + i] : i;
      if (nb > 1) kin_accum(W, P, S, li, v);
    }
    for (int i = 0; i < nb; i++) {
      const int li = conv ? W.el[b + i] : i;
      if (skip_null && W.g_id[W.gidx[b + li]] == -1) continue;
      const int t = __float_as_int(W.vel[b + li].w);
      nt[t]++;
      mts[t] += W.pos[b + li].w;
    }
    write_record(W, P, S, v[0], v[1], v[2], v[3], v[4], v[5], nt, mts);
  }
}

// FAST form, tile-parallel: partial sums per tile (6 kinematic sums, 6 type
// counts, 6 type masses in FP64), combined per segment in tile order.  A
// converged segment's bound set is read in slot order -- the slots of the
// list it unbound, [0, ulen), flagged in isb (every pass's partition flags the
// slots it keeps, so the flags hold the final bound set) -- so the reads
// stream through the columns instead of following the energy order; other
// finished segments sum their first nbound slots.
constexpr int REC_SUMS = 18;

__device__ __forceinline__ int record_span(const Seg &S) { return (S.flags & F_CONV) ? S.ulen : S.nbound; }

// The record sums of slot li: kinematics of a bound entry, and its type
// count and mass unless it is a null particle of a small subhalo.
__device__ __forceinline__ void record_accum(const Work &W, const DevParams &P, const Seg &S, int64_t b, int li,
                                             bool conv, int nb, bool skip_null, double (&v)[REC_SUMS]) {
#pragma clang fp contract(off)
  if (conv && !W.isb[b + li]) return;
  if (nb > 1) kin_accum(W, P, S, li, v);
  if (skip_null && W.g_id[W.gidx[b + li]] == -1) return;
  const int ty = __float_as_int(W.vel[b + li].w);
  const float m = W.pos[b + li].w;
  for (int q = 0; q < HBT_TYPE_MAX; q++) {
    const bool hit = q == ty;
    v[6 + q] += hit ? 1.0 : 0.0;
    v[12 + q] += hit ? (double)m : 0.0;
  }
}

__global__ void __launch_bounds__(BLOCK)
output_partial_kernel(Work W, DevParams P, const Tile *tiles, int ntiles, double *part) {
#pragma clang fp contract(off)
  __shared__ double sh[REC_SUMS][BLOCK / WAVE];
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const Tile tl = tiles[t];
    const Seg &S = W.seg[tl.seg];
    const bool fin = (S.flags & F_FINAL) != 0;
    const int cnt = fin ? min(TILE, record_span(S) - tl.start) : 0;
    if (cnt <= 0) continue;  // uniform
    const int64_t b = S.ibase;
    const bool conv = (S.flags & F_CONV) != 0;
    const int nb = S.nbound;
    const bool skip_null = nb <= 100;
    double v[REC_SUMS];
    for (int c = 0; c < REC_SUMS; c++) v[c] = 0;
    for (int i = threadIdx.x; i < cnt; i += BLOCK) record_accum(W, P, S, b, tl.start + i, conv, nb, skip_null, v);
    for (int c = 0; c < REC_SUMS; c++) {
      const double r = wave_total(v[c]);
      if ((threadIdx.x & (WAVE - 1)) == 0) sh[c][threadIdx.x / WAVE] = r;
    }
    __syncthreads();
    if (threadIdx.x < REC_SUMS) {
      double r = 0;
      for (int w = 0; w < BLOCK / WAVE; w++) r += sh[threadIdx.x][w];
      part[(int64_t)t * REC_SUMS + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

// One-wave tiles of up to R * WAVE slots, the same bits as the workgroup
// form (see frames_partial_wave_kernel).
template <int R>
__device__ __forceinline__ void output_wave_tile(const Work &W, const DevParams &P, const Tile &tl, int t,
                                                 double *part) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & (WAVE - 1);
  {
    const Seg &S = W.seg[tl.seg];
    const bool fin = (S.flags & F_FINAL) != 0;
    const int cnt = fin ? min(R * WAVE, record_span(S)) : 0;
    if (cnt <= 0) return;  // uniform across the wave
    const int64_t b = S.ibase;
    const int nb = S.nbound;
    double acc[REC_SUMS];
    for (int c = 0; c < REC_SUMS; c++) acc[c] = 0;
    for (int r = 0; r < R; r++) {
      double v[REC_SUMS];
      for (int c = 0; c < REC_SUMS; c++) v[c] = 0;
      const int e = lane + WAVE * r;
      if (e < cnt) record_accum(W, P, S, b, e, (S.flags & F_CONV) != 0, nb, nb <= 100, v);
      for (int c = 0; c < REC_SUMS; c++) acc[c] += wave_total(v[c]);
    }
    if (lane == 0)
      for (int c = 0; c < REC_SUMS; c++) part[(int64_t)t * REC_SUMS + c] = acc[c];
  }
}

// Tiles t0 <= t < tmid take four slots per lane, tmid <= t < ntiles one.
__global__ void __launch_bounds__(BLOCK)
output_partial_wave_kernel(Work W, DevParams P, const Tile *tiles, int t0, int tmid, int ntiles, double *part) {
  WAVE_TILE_LOOP(t, t0, ntiles) {
    const Tile tl = tiles[t];
    if (t < tmid) output_wave_tile<4>(W, P, tl, t, part);
    else output_wave_tile<1>(W, P, tl, t, part);
  }
}

__global__ void __launch_bounds__(BLOCK) output_combine_kernel(Work W, DevParams P, const int *first_tile,
                                                               const double *part) {
#pragma clang fp contract(off)
  WAVE_SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_FINAL)) continue;  // uniform across the wave
    double v[REC_SUMS];
    wave_sum_tiles<REC_SUMS>(part, first_tile[k], (record_span(S) + TILE - 1) / TILE, v);
    if ((threadIdx.x & (WAVE - 1)) != 0) continue;
    int nt[HBT_TYPE_MAX];
    float mts[HBT_TYPE_MAX];
    for (int q = 0; q < HBT_TYPE_MAX; q++) { nt[q] = (int)v[6 + q]; mts[q] = (float)v[12 + q]; }
    write_record(W, P, S, v[0], v[1], v[2], v[3], v[4], v[5], nt, mts);
  }
}

// ---------------------------------------------------------------------------
// RefineBindingEnergyOrder (src/subhalo_unbind.cpp:234-262), for the segments
// that converged with more bound particles than MaxSampleSize: a tree over the
// first MaxSampleSize entries of the energy-sorted list (full masses), each of
// them a target with its own mass (self term removed) in the final frame,
// then those entries re-ordered by that energy; they keep their unbinding
// energies.  Runs once per level after its passes.

__global__ void refine_setup_kernel(Work W, DevParams P) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    const bool on = (S.flags & F_REFPEND) != 0;
    S.flags &= ~(F_CORR | F_SAMP | F_FINAL);
    if (on) S.flags |= F_REFINE;
    S.mf = 1.f;
    S.src_lo = 0;
    S.n_src = on ? P.max_sample : 0;
    S.n_tgt = S.n_src;
    S.nsort = S.n_src;
    S.sort_lo = 0;
    S.n_items = 0;
    for (int j = 0; j < 3; j++) { S.bmin[j] = 0xffffffffu; S.bmax[j] = 0u; }
    W.ncount[NC_SRC * W.nseg + k] = S.n_src;
    W.ncount[NC_SORT * W.nseg + k] = S.n_src;
  }
}

// targets: the entries themselves (their order does not change any result)
__global__ void __launch_bounds__(BLOCK) refine_targets_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (!(S.flags & F_REFINE) || l >= S.n_tgt) continue;
    W.sp[s] = W.el[s];
  }
}

// sort keys (refinement energy, position): ties keep the energy order
__global__ void __launch_bounds__(BLOCK) refine_keys_kernel(Work W) {
  SLOT_LOOP(s, W) {
    int k;
    int l = seg_lane_local(W, s, k);
    if (l < 0) continue;
    const Seg &S = W.seg[k];
    if (!(S.flags & F_REFINE) || l >= S.nsort) continue;
    const int li = W.el[s];
    W.k0[s] = ((u64)f2ord(W.Er[S.ibase + li]) << 31) | (u64)(unsigned)l;
    W.v0[s] = (u32)li;
  }
}

__global__ void refine_done_kernel(Work W) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_REFINE)) continue;
    S.flags = (S.flags & ~(F_REFINE | F_REFPEND)) | F_FINAL;
    S.n_src = S.n_tgt = 0;
  }
}

}  // namespace hbt
