// hbt_unbind.hip -- host orchestration and C-ABI of libhbt_unbind.so (gfx950).
//
// A batch of subhalos is processed level by level: a level is a set of
// subhalos whose nested children are all done (RecursiveUnbind order,
// src/subhalo_unbind.cpp:432-447).  Batched form: one level per nest height,
// all its subhalos advance through the unbinding passes together.  Plain form:
// one subhalo per level, in the reference's depth-first order -- the same
// kernels driven one subhalo at a time, as the reference's RefineParticles loop
// calls Subhalo_t::Unbind per subhalo.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <initializer_list>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/hbt_unbind.h"
#include "hbt_kernels.hpp"

using namespace hbt;

#include "hbt_rand.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      return fail(HBT_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define RETCHK(x)              \
  do {                         \
    int r_ = (x);              \
    if (r_ != HBT_OK) return r_; \
  } while (0)

DevParams make_dev_params(const hbt_params &p) {
  DevParams d;
  d.G = p.G;
  d.eps = p.softening;
  d.theta2 = p.open_angle * p.open_angle;
  d.res = (float)((double)p.softening * 0.1);
  d.box = p.box_size;
  d.box_half = (float)((double)p.box_size / 2.);
  d.bmp = p.bound_mass_precision;
  d.relax = p.source_sub_relax_factor;
  d.a = p.scale_factor;
  d.hz = p.hz;
  d.periodic = p.periodic != 0;
  d.min_part = p.min_num_part_of_sub;
  d.snap = p.snapshot_index;
  d.replay = p.precision == HBT_PREC_REPLAY;
  d.wide_sum = p.precision != HBT_PREC_FP32;
  d.max_sample = p.max_sample_size > 0 ? p.max_sample_size : 0;
  d.refine = d.max_sample > 0 && p.refine_mostbound_particle != 0;
  return d;
}

// Grow-only device allocation.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return HBT_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (want == 0) return HBT_OK;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) return fail(HBT_ERR_OUT_OF_MEMORY, "hipMalloc failed");
    bytes = want;
    return HBT_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return (T *)p; }
};

// Grow-only pinned host allocation; ensure(want, keep) keeps the first
// `keep` bytes when it grows.
struct HBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want, size_t keep = 0) {
    if (want <= bytes) return HBT_OK;
    void *q = nullptr;
    hipError_t e = hipHostMalloc(&q, want, hipHostMallocDefault);
    if (e != hipSuccess) return fail(HBT_ERR_OUT_OF_MEMORY, "hipHostMalloc failed");
    if (p && keep) memcpy(q, p, std::min(keep, bytes));
    if (p) (void)hipHostFree(p);
    p = q;
    bytes = want;
    return HBT_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// A typed view into a device buffer owned elsewhere.
struct DView {
  void *p = nullptr;
  template <class T> T *as() const { return (T *)p; }
};

}  // namespace

struct hbt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hbt_params params{};
  int cus = 256;
  // per-slot work arrays
  DBuf gidx, pos, vel, E, el, el2, sp, k0, k1, v0, v1, dl, bits, pre, nlo, nhi,
      npos, items, scan64, pk, gk, isb, seg_of_slot, cross, cross2, tcount;
  // per-level tables: staged through one pinned host buffer and one copy
  DBuf lvlbuf;
  void *h_stage = nullptr;
  size_t h_stage_bytes = 0;
  DView segin, ibase, tiles, first_tile, ntile, child_ptr, child_list, capal;
  DBuf seg;             // per-segment state, expanded on the device from segin
  DBuf ncount, tsum, tpart;
  DBuf segact, actbuf;  // segment activity after a pass; lists of the active segments
  HBuf h_list;          // pinned: activity flags read back, lists of the active segments
  DBuf gaps;            // (subhalo, capacity, first flat index) of the hosts whose output regions have unused slots
  DBuf active, counters;
  // instrumentation: 1 = time walk launches with HIP events, 2 = count interactions
  int prof_mode = 0;
  int walk_mode = 0;  // 0: group-shared traversal, 1: independent per-lane walks
  // lanes per traversal group by size bin, graph capture of the pass sequence
  // (extra passes of a graph launch on a converged level are no-ops)
  hbt_tuning tune = {{8, 8, 8}, {0, 1 << 30}, 2, 0, 16};
  hipEvent_t evw0 = nullptr, evw1 = nullptr;
  double walk_ms = 0;
  int64_t walk_launches = 0;
  hipEvent_t evs0 = nullptr, evs1 = nullptr;  // the segmented sorts of the passes
  double sort_ms = 0;
  int64_t sort_calls = 0;
  // staging for hbt_unbind_one / hbt_tree_potential
  DBuf st_cols, st_subs, st_out_i, st_out_e, st_phi;
  int *h_active = nullptr;  // pinned
  const int64_t *list_index = nullptr;  // set for the duration of hbt_unbind_batch_indexed
  const float4 *packed = nullptr;       // likewise, when the snapshot is gathered from packed records
  DBuf pack;                            // the packed snapshot
  DBuf idtab;                           // hbt_id_to_index's table
  DBuf idpart, idcnt;                   // its partitioned (key, position) pairs and bucket counts
  // the rand() stream of the sampling shuffles: srand(rand_seed), rand_pos calls made
  uint32_t rand_seed = 1;
  int64_t rand_pos = 0;
  DBuf Er, draws, randtab;              // refinement energies; a level's draws; chunk jump matrices
  DBuf kick, kickin, kickcnt;           // lists compacted by KickNullParticles, their ranges and counts
  DView rwin;                           // the level's window of the stream (staged)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  int64_t last_evals = 0, last_passes = 0;
  ~hbt_ctx() {
    for (DBuf *b : {&gidx, &pos, &vel, &E, &el, &el2, &sp, &k0, &k1, &v0, &v1, &dl,
                    &bits, &pre, &nlo, &nhi, &npos, &items, &scan64, &pk, &gk,
                    &isb, &seg_of_slot, &cross, &cross2, &tcount, &seg, &lvlbuf, &ncount, &tsum, &tpart, &active, &counters,
                    &st_cols, &st_subs, &st_out_i, &st_out_e, &st_phi, &segact, &actbuf, &gaps, &pack, &idtab, &idpart, &idcnt,
                    &Er, &draws, &randtab, &kick, &kickin, &kickcnt})
      b->release();
    if (h_active) (void)hipHostFree(h_active);
    if (h_stage) (void)hipHostFree(h_stage);
    h_list.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (evw0) (void)hipEventDestroy(evw0);
    if (evw1) (void)hipEventDestroy(evw1);
    if (evs0) (void)hipEventDestroy(evs0);
    if (evs1) (void)hipEventDestroy(evs1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Later passes visit only the active segments' slots and tiles
// (HBT_COMPACT=0 keeps every pass on the whole level).
bool compact_passes() {
  static const bool v = getenv("HBT_COMPACT") ? atoi(getenv("HBT_COMPACT")) != 0 : true;
  return v;
}

// Level description built on the host.
struct Level {
  std::vector<int> subs;        // subhalo indices
  std::vector<SegIn> seg;      // what the host fixes per segment; the rest starts at zero
  std::vector<int64_t> ibase;
  std::vector<int> capal, first_tile, ntile, child_ptr, child_list;
  std::vector<Tile> tiles;
  int64_t nslots = 0;
  int maxcap = 0;
  int64_t ndraws = 0;                 // rand() values the level's shuffles take
  uint32_t win[31] = {};              // the stream's window before the first of them
};

int walk_group_for(const hbt_tuning &t, int64_t len) {
  return len < t.bin_edge[0] ? t.walk_group[0] : (len < t.bin_edge[1] ? t.walk_group[1] : t.walk_group[2]);
}

// The batch's nest structure on the host, with the record fields the
// planning reads gathered once (a record is 240 bytes, the planning needs 40
// of them): capacities (own list plus the nested subhalos' capacities),
// heights, children in index order (CSR) and the post-order (RecursiveUnbind
// order).
struct SubPlan {
  int64_t offset, count, nbp, out_off, cap;  // cap: own list plus the nested subhalos' capacities
};
struct Nest {
  std::vector<SubPlan> f;
  std::vector<int> height, ch_ptr, ch_idx, order;
  int64_t n = 0;
  int nchild(int s) const { return ch_ptr[s + 1] - ch_ptr[s]; }
  const int *child(int s) const { return ch_idx.data() + ch_ptr[s]; }
};

// f(r) for r in [0, parts), each on a thread of its own (r = 0 on the caller's).
// A thread_local the caller fills must reach f by pointer: named inside f, it
// is the worker thread's own instance.
template <class F>
static void parallel_parts(int parts, const F &f) {
  std::vector<std::thread> th;
  for (int r = 1; r < parts; r++) th.emplace_back(f, r);
  f(0);
  for (auto &x : th) x.join();
}

// f(lo, hi) over [0, n), on four threads for large n.
template <class F>
static void parallel_ranges(int64_t n, const F &f) {
  const int nth = n >= 65536 ? 4 : 1;
  parallel_parts(nth, [&](int r) { f(n * r / nth, n * (r + 1) / nth); });
}

// Members of a level sorted by capacity, largest first, ties in member order;
// caps receives the sorted capacities.  Most lists are short, so one stable
// counting pass on the capacity clamped to CLAMP - 1 orders all but the
// members at or above the clamp, which it gathers in front (in member order)
// and a stable comparison sort then orders; that bucket holds few members.
// Large levels count and place on four threads, each over its own quarter of
// the members, a quarter's members of a bucket following the earlier
// quarters' members of that bucket.
void sort_largest_first(const Nest &N, const std::vector<int> &members, std::vector<int> &out,
                        std::vector<int> &caps) {
  const size_t n = members.size();
  out.resize(n);
  caps.resize(n);
  if (n < 64) {  // one-subhalo levels (the plain form, single calls) skip the histograms
    std::copy(members.begin(), members.end(), out.begin());
    std::stable_sort(out.begin(), out.end(), [&](int a, int b) { return N.f[a].cap > N.f[b].cap; });
    for (size_t i = 0; i < n; i++) caps[i] = (int)N.f[out[i]].cap;
    return;
  }
  constexpr int CLAMP = 4096, NT = 4;
  static thread_local std::vector<int> key;
  static thread_local std::vector<uint32_t> cnt;
  key.resize(n);
  cnt.assign((size_t)NT * CLAMP, 0u);
  const int nth = n >= 65536 ? NT : 1;
  int *kp = key.data(), *op = out.data(), *capp = caps.data();
  uint32_t *cp = cnt.data();  // one histogram per quarter
  const int *mp = members.data();
  const Nest *np = &N;
  parallel_parts(nth, [=](int r) {
    uint32_t *h = cp + (size_t)r * CLAMP;
    for (size_t i = n * r / nth; i < n * (r + 1) / nth; i++) {
      const int c = (int)np->f[mp[i]].cap;
      kp[i] = c;
      h[CLAMP - 1 - std::min(c, CLAMP - 1)]++;  // bucket 0: the largest
    }
  });
  uint32_t o = 0, nbig = 0;
  for (int b = 0; b < CLAMP; b++) {
    for (int r = 0; r < nth; r++) {
      const uint32_t x = cp[(size_t)r * CLAMP + b];
      cp[(size_t)r * CLAMP + b] = o;
      o += x;
    }
    if (b == 0) nbig = o;
  }
  parallel_parts(nth, [=](int r) {
    uint32_t *h = cp + (size_t)r * CLAMP;
    for (size_t i = n * r / nth; i < n * (r + 1) / nth; i++) {
      const int c = kp[i];
      const uint32_t p = h[CLAMP - 1 - std::min(c, CLAMP - 1)]++;
      op[p] = mp[i];
      capp[p] = c;
    }
  });
  if (nbig > 1) {
    std::vector<std::pair<int, int>> big(nbig);  // (capacity, member), in member order
    for (uint32_t i = 0; i < nbig; i++) big[i] = {caps[i], out[i]};
    std::stable_sort(big.begin(), big.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
      return a.first > b.first;
    });
    for (uint32_t i = 0; i < nbig; i++) {
      caps[i] = big[i].first;
      out[i] = big[i].second;
    }
  }
}

// draw_off: per subhalo, the first of its shuffle's draws in the level's
// draws (-1: not shuffled) and the list length they were counted for
// (draw_off[nsub + s]); null when nothing is shuffled.
void build_level(const Nest &N, const std::vector<int> &members, const hbt_tuning &tune,
                 const int64_t *draw_off, int64_t nsub, Level &L) {
  const int n = (int)members.size();
  // largest first: the merge levels of a sort then cover a prefix
  static thread_local std::vector<int> caps;
  sort_largest_first(N, members, L.subs, caps);
  L.seg.resize(n);
  L.ibase.resize(n);
  L.capal.resize(n);
  L.first_tile.resize(n);
  L.ntile.resize(n);
  // slot and tile offsets in level order
  int64_t off = 0, t0 = 0;
  L.maxcap = 0;
  for (int k = 0; k < n; k++) {
    const int c = caps[k];
    const int ca = std::max(WAVE, (c + WAVE - 1) / WAVE * WAVE);
    const int nt = (c + TILE - 1) / TILE;
    L.ibase[k] = off;
    L.capal[k] = ca;
    L.first_tile[k] = (int)t0;
    L.ntile[k] = nt;
    off += ca;
    t0 += nt;
    L.maxcap = std::max(L.maxcap, c);
  }
  L.tiles.resize(t0);
  L.nslots = off;
  // nest children in level order (none in a level of top-level subhalos)
  L.child_ptr.assign(n + 1, 0);
  if (N.ch_ptr[N.n] > 0)
    for (int k = 0; k < n; k++) L.child_ptr[k + 1] = L.child_ptr[k] + N.nchild(L.subs[k]);
  L.child_list.resize(L.child_ptr[n]);
  // caps is this thread's: the workers take its data by pointer
  const int *cp = caps.data();
  parallel_ranges(n, [&, cp](int64_t k0, int64_t k1) {
    for (int64_t k = k0; k < k1; k++) {
      const int s = L.subs[k];
      const SubPlan &F = N.f[s];
      const int c = cp[k];
      L.seg[k] = SegIn{L.ibase[k], F.out_off, F.offset, F.nbp, s, c, (int)F.count, walk_group_for(tune, c),
                       draw_off ? (int)draw_off[s] : -1, draw_off ? (int)draw_off[nsub + s] : 0};
      for (int t = 0; t < L.ntile[k]; t++) L.tiles[L.first_tile[k] + t] = Tile{(int)k, t * TILE};
      for (int i = 0; i < N.nchild(s); i++) L.child_list[L.child_ptr[k] + i] = N.child(s)[i];
    }
  });
}

// Copies host tables to the device: packed (256-byte aligned) into the pinned
// staging buffer, one asynchronous copy into lvlbuf; dev[i] receives the
// device address of blob i.  The stream is drained first, so no earlier copy
// still reads the staging buffer (normally it is idle already: every level
// and every entry point ends with a stream synchronisation).
struct Blob {
  const void *p;
  size_t bytes;
};

// Copies blob i to h + at[i]: the blobs' bytes taken end to end and split
// into four equal ranges for large levels.
void pack_blobs(char *h, const std::vector<size_t> &at, std::initializer_list<Blob> blobs) {
  std::vector<const Blob *> bl;
  std::vector<size_t> end;  // running total of the blobs' bytes
  size_t total = 0;
  for (const Blob &b : blobs) {
    bl.push_back(&b);
    total += b.bytes;
    end.push_back(total);
  }
  const int64_t nq = (int64_t)(total >> 4);  // 16-byte units
  parallel_ranges(nq, [&](int64_t q0, int64_t q1) {
    const size_t lo = (size_t)q0 << 4, hi = q1 == nq ? total : (size_t)q1 << 4;
    size_t j = std::upper_bound(end.begin(), end.end(), lo) - end.begin(), p = lo;
    for (; p < hi && j < bl.size(); j++) {
      const size_t b0 = end[j] - bl[j]->bytes, q = std::min(hi, end[j]);
      if (q > p) memcpy(h + at[j] + (p - b0), (const char *)bl[j]->p + (p - b0), q - p);
      p = q;
    }
  });
}

int stage_blobs(hbt_ctx *c, std::initializer_list<Blob> blobs, void **dev) {
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<size_t> at;
  size_t off = 0;
  for (const Blob &b : blobs) {
    at.push_back(off);
    off += (std::max<size_t>(b.bytes, 1) + 255) & ~(size_t)255;
  }
  if (off > c->h_stage_bytes) {
    if (c->h_stage) HIPCHK(hipHostFree(c->h_stage));
    c->h_stage = nullptr;
    c->h_stage_bytes = 0;
    HIPCHK(hipHostMalloc(&c->h_stage, off, hipHostMallocDefault));
    c->h_stage_bytes = off;
  }
  RETCHK(c->lvlbuf.ensure(off));
  char *h = (char *)c->h_stage, *d = (char *)c->lvlbuf.p;
  for (size_t i = 0; i < blobs.size(); i++) dev[i] = d + at[i];
  pack_blobs(h, at, blobs);
  HIPCHK(hipMemcpyAsync(c->lvlbuf.p, c->h_stage, off, hipMemcpyHostToDevice, c->stream));
  return HBT_OK;
}

// The level's tables: segment records, slot bases, aligned capacities, tile
// table, first tile and tile count per segment, nest children (CSR).
int stage_level(hbt_ctx *c, const Level &L) {
  const size_t n = L.seg.size();
  void *d[9];
  RETCHK(c->seg.ensure(std::max<size_t>(n, 1) * sizeof(Seg)));
  RETCHK(stage_blobs(c, {{L.seg.data(), n * sizeof(SegIn)},
                         {L.ibase.data(), n * 8},
                         {L.capal.data(), n * 4},
                         {L.tiles.data(), L.tiles.size() * sizeof(Tile)},
                         {L.first_tile.data(), n * 4},
                         {L.ntile.data(), n * 4},
                         {L.child_ptr.data(), L.child_ptr.size() * 4},
                         {L.child_list.data(), L.child_list.size() * 4},
                         {L.win, sizeof L.win}},
                     d));
  c->segin.p = d[0]; c->ibase.p = d[1]; c->capal.p = d[2]; c->tiles.p = d[3];
  c->first_tile.p = d[4]; c->ntile.p = d[5]; c->child_ptr.p = d[6]; c->child_list.p = d[7];
  c->rwin.p = d[8];
  return HBT_OK;
}

template <class T>
int upload(DBuf &b, const std::vector<T> &v, hipStream_t st) {
  RETCHK(b.ensure(std::max<size_t>(v.size() * sizeof(T), 16)));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return HBT_OK;
}

// The walk's grid covers every slot block once: its blocks finish at very
// different times and the hardware hands the remaining ones to the CUs that
// come free; the count is rounded up to whole XCD runs so that the XCD-aware
// block order applies (the extra blocks find no slots).
int walk_grid(int64_t nslots, int xcd_run) {
  int64_t g = (nslots + BLOCK - 1) / BLOCK;
  const int64_t m = (int64_t)NXCD * std::max(1, xcd_run);
  g = (g + m - 1) / m * m;
  return (int)std::max<int64_t>(g, 1);
}

int grid_for(int64_t n, int cus) {
  int64_t g = (n + BLOCK - 1) / BLOCK;
  int64_t cap = (int64_t)cus * 16;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (int)g;
}

// ---- functors for the segmented scans ---------------------------------------
struct PopLoad {
  const int64_t *ibase; const u32 *bits;
  __device__ u64 operator()(int k, int l) const { return (u64)__popc(bits[ibase[k] + l]); }
};
struct PreStore {
  const int64_t *ibase; int *pre;
  __device__ void operator()(int k, int l, u64 x) const { pre[ibase[k] + l] = (int)x; }
};
struct PreAdd {
  const int64_t *ibase; int *pre;
  __device__ void operator()(int k, int l, u64 x) const { pre[ibase[k] + l] += (int)x; }
};
struct PartLoad {
  const int64_t *ibase; const int *el; const float *E;
  __device__ u64 operator()(int k, int l) const {
    if (l < 1) return 0;
    const float e = E[ibase[k] + el[ibase[k] + l]];
    return (u64)(e > 0.f ? 1u : 0u) | ((u64)(e < 0.f ? 1u : 0u) << 32);
  }
};
struct S64Store {
  const int64_t *ibase; u64 *out;
  __device__ void operator()(int k, int l, u64 x) const { out[ibase[k] + l] = x; }
};
struct S64Add {
  const int64_t *ibase; u64 *out;
  __device__ void operator()(int k, int l, u64 x) const { out[ibase[k] + l] += x; }
};
struct BoundLoad {
  const int64_t *ibase; const int *sp; const unsigned char *isb;
  __device__ u64 operator()(int k, int l) const { return (u64)isb[ibase[k] + sp[ibase[k] + l]]; }
};


// Partition totals of each active segment: the sum of its tile totals.
__global__ void partition_counts_kernel(Work W, const int *first_tile, const u64 *tile_sum) {
  SEG_LOOP(k, W) {
    Seg &S = W.seg[k];
    if (!(S.flags & F_ACTIVE)) { S.nP = S.nG = 0; continue; }
    const int nt = (W.ncount[NC_PART * W.nseg + k] + TILE - 1) / TILE;
    u64 t = 0;
    for (int u = first_tile[k]; u < first_tile[k] + nt; u++) t += tile_sum[u];
    S.nP = (int)(t & 0xffffffffu);
    S.nG = (int)(t >> 32);
    S.K = 0;
  }
}

// A host's output region holds its own list and every nested particle (the
// planned capacity); only its subhalos' unbound tails arrive, so the slots
// past its final list are set to -1 and NaN.  gap: (subhalo, capacity, first
// flat index) triples; the flat index runs over the hosts' regions in turn, so
// the grid spreads over them whatever their sizes.
__global__ void __launch_bounds__(BLOCK) fill_gaps_kernel(const hbt_subhalo *subs, const int64_t *gap, int ngap,
                                                          int64_t total, int64_t *out_index, float *out_energy) {
  for (int64_t f = (int64_t)blockIdx.x * BLOCK + threadIdx.x; f < total; f += (int64_t)gridDim.x * BLOCK) {
    int lo = 0, hi = ngap - 1;  // last host whose first flat index is <= f
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (gap[3 * mid + 2] <= f) lo = mid; else hi = mid - 1;
    }
    const int64_t s = gap[3 * lo], i = f - gap[3 * lo + 2];
    if (i < subs[s].list_len) continue;
    const int64_t o = subs[s].out_offset + i;
    out_index[o] = -1;
    out_energy[o] = __builtin_nanf("");
  }
}

// Sum of n_evals over the records (one block).
__global__ void __launch_bounds__(BLOCK) eval_sum_kernel(const hbt_subhalo *subs, int64_t nsub,
                                                         unsigned long long *out) {
  unsigned long long v = 0;
  for (int64_t i = threadIdx.x; i < nsub; i += BLOCK) v += (unsigned long long)subs[i].n_evals;
  unsigned long long tot;
  (void)block_exclusive_scan(v, &tot);
  if (threadIdx.x == 0) *out = tot;
}

// Particle types outside [0, HBT_TYPE_MAX) (the per-type counts index by type).
__global__ void __launch_bounds__(BLOCK) type_check_kernel(const int32_t *type, int64_t n, int *bad) {
  int cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
    cnt += (type[i] < 0 || type[i] >= HBT_TYPE_MAX) ? 1 : 0;
  cnt = (int)wave_reduce_sum((unsigned)cnt);
  if ((threadIdx.x & (WAVE - 1)) == 0 && cnt) atomicAdd(bad, cnt);
}

// List entries that are not positions in [0, n): cnt[0] counts those >= n
// (rejected), cnt[1] the negative ones (null particles: IDs the snapshot did
// not resolve, dropped by KickNullParticles).
__global__ void __launch_bounds__(BLOCK) list_check_kernel(const int64_t *idx, int64_t n_list, int64_t n, int *cnt) {
  int big = 0, null = 0;
  for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n_list; i += (int64_t)gridDim.x * BLOCK) {
    big += idx[i] >= n ? 1 : 0;
    null += idx[i] < 0 ? 1 : 0;
  }
  big = (int)wave_reduce_sum((unsigned)big);
  null = (int)wave_reduce_sum((unsigned)null);
  if ((threadIdx.x & (WAVE - 1)) == 0) {
    if (big) atomicAdd(&cnt[0], big);
    if (null) atomicAdd(&cnt[1], null);
  }
}

// Subhalo_t::KickNullParticles (src/subhalo.cpp:446-477) on the device: a
// list keeps its entries that are not null, in order, and the Nbound it
// carries (nbound_prev, RecursiveUnbind's orphan test) counts the kept ones
// among its first Nbound entries.  One workgroup per subhalo.
struct KickIn {
  int64_t off, cnt, front, dst;  // list range, entries counted by Nbound, first kept entry's place
};

__global__ void __launch_bounds__(BLOCK) kick_count_kernel(const int64_t *__restrict__ idx, const KickIn *in,
                                                           int64_t nsub, int64_t *keep) {
  for (int64_t s = blockIdx.x; s < nsub; s += gridDim.x) {
    const KickIn k = in[s];
    unsigned long long a = 0, f = 0;
    for (int64_t i = threadIdx.x; i < k.cnt; i += BLOCK) {
      const bool live = idx[k.off + i] >= 0;
      a += live ? 1 : 0;
      f += (live && i < k.front) ? 1 : 0;
    }
    unsigned long long ta, tf;
    (void)block_exclusive_scan(a, &ta);
    (void)block_exclusive_scan(f, &tf);
    if (threadIdx.x == 0) {
      keep[2 * s] = (int64_t)ta;
      keep[2 * s + 1] = (int64_t)tf;
    }
  }
}

__global__ void __launch_bounds__(BLOCK) kick_compact_kernel(const int64_t *__restrict__ idx, const KickIn *in,
                                                             int64_t nsub, int64_t *out) {
  for (int64_t s = blockIdx.x; s < nsub; s += gridDim.x) {
    const KickIn k = in[s];
    int64_t run = 0;
    for (int64_t i0 = 0; i0 < k.cnt; i0 += BLOCK) {  // uniform across the block
      const int64_t i = i0 + threadIdx.x;
      const int64_t v = i < k.cnt ? idx[k.off + i] : -1;
      unsigned long long tot;
      const unsigned long long ex = block_exclusive_scan((unsigned long long)(v >= 0 ? 1 : 0), &tot);
      if (v >= 0) out[k.dst + run + (int64_t)ex] = v;
      run += (int64_t)tot;
    }
  }
}

struct Runner {
  hbt_ctx *c;
  Work W;
  DevParams P;
  Level *L;
  int gs, gseg, gtile, gblk;  // slot, segment-thread, tile and segment-block grids
  bool swap_sp = true;         // swap the spatial list with its compacted copy (direct launches)
  int nt_big = 0;              // tiles of segments of more than SMALL_CAP slots (a prefix)
  int nt_tiny = 0;             // then those of more than TINY_CAP; the rest are of at most TINY_CAP
  int gsmall = 1, gtiny = 1, gwave = 1;  // one-wave tile grids (four tiles per workgroup)
  int gwseg = 1;               // segment-wave grid
  int gw;                      // walk grid: every slot block once
  // the tile table a pass uses: the level's, or that of its active segments
  // (then W.act_chunk lists their slots); caps and first tiles of the listed
  // segments in order, for the merge-level prefixes
  const Tile *tl_cur = nullptr;
  const int *ft_cur = nullptr;
  int nt_cur = 0;
  std::vector<int> cur_cap, cur_first;
  int nlisted = 0;             // segments in the current lists

  void list_all() {
    tl_cur = c->tiles.as<Tile>();
    ft_cur = c->first_tile.as<int>();
    nt_cur = (int)L->tiles.size();
    cur_cap.resize(L->seg.size());
    for (size_t k = 0; k < L->seg.size(); k++) cur_cap[k] = L->seg[k].cap;
    cur_first = L->first_tile;
    nlisted = (int)L->seg.size();
    W.act_chunk = nullptr;
    W.nact = L->nslots;
    set_grids();
  }

  void set_grids() {
    gs = grid_for(W.nact, c->cus);
    gw = walk_grid(W.nact, c->tune.xcd_run);
    nt_big = tiles_above(SMALL_CAP);
    nt_tiny = tiles_above(TINY_CAP);
    gtile = (int)std::max<int64_t>(1, std::min<int64_t>(nt_big, (int64_t)c->cus * 8));
    gsmall = wave_grid(nt_tiny - nt_big);
    gtiny = wave_grid(nt_cur - nt_tiny);
    gwave = wave_grid(nt_cur - nt_big);
  }

  // tile kernels: the workgroup form over the large segments' tiles, the
  // one-wave form over the rest
  int wave_grid(int ntiles) const {
    return (int)std::max<int64_t>(1, std::min<int64_t>((ntiles + BLOCK / WAVE - 1) / (BLOCK / WAVE),
                                                       (int64_t)c->cus * 16));
  }
  bool any_big() const { return nt_big > 0; }
  bool any_small() const { return nt_tiny > nt_big; }
  bool any_tiny() const { return nt_cur > nt_tiny; }
  bool any_wave() const { return nt_cur > nt_big; }

  // Lists the segments still active after a pass: their wave-sized slot
  // chunks, their tiles (a table of its own, with a first-tile index per
  // segment), caps and first tiles in level order.
  int list_active() {
    const int nseg = (int)L->seg.size();
    // one pinned host buffer: the activity flags come down into it and the
    // lists are written in place and go up with one copy
    const size_t b_act = ((size_t)nseg * 4 + 255) & ~(size_t)255;
    RETCHK(c->h_list.ensure(b_act));
    int *act = (int *)c->h_list.p;
    HIPCHK(hipMemcpyAsync(act, W.segact, nseg * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // listed segments, chunks and tiles per quarter of the segments (on four
    // threads for large levels), then each quarter's lists at its offsets
    const int nth = nseg >= 65536 ? 4 : 1;
    int64_t qn[4] = {}, qch[4] = {}, qtl[4] = {};
    const Level *lv = L;
    auto quarter = [nseg, nth](int r, int &k0, int &k1) {
      k0 = (int)((int64_t)nseg * r / nth);
      k1 = (int)((int64_t)nseg * (r + 1) / nth);
    };
    parallel_parts(nth, [&, act, lv](int r) {
      int k0, k1;
      quarter(r, k0, k1);
      int64_t a = 0, t = 0, ch = 0;
      for (int k = k0; k < k1; k++) {
        if (!act[k]) continue;
        a++;
        t += lv->ntile[k];
        ch += lv->capal[k] / WAVE;
      }
      qn[r] = a;
      qtl[r] = t;
      qch[r] = ch;
    });
    int64_t nch = 0, ntl = 0, nl = 0, on[4], och[4], otl[4];
    for (int r = 0; r < nth; r++) {
      on[r] = nl;
      och[r] = nch;
      otl[r] = ntl;
      nl += qn[r];
      nch += qch[r];
      ntl += qtl[r];
    }
    nlisted = (int)nl;
    const size_t b_ch = ((size_t)nch * 4 + 255) & ~(size_t)255, b_tl = ((size_t)ntl * sizeof(Tile) + 255) & ~(size_t)255;
    const size_t b_ft = (size_t)nseg * 4;
    // the flags stay at the front while the lists are written behind them
    RETCHK(c->h_list.ensure(b_act + b_ch + b_tl + b_ft, b_act));
    act = (int *)c->h_list.p;
    char *h = (char *)c->h_list.p + b_act;
    int *chunks = (int *)h;
    Tile *tl = (Tile *)(h + b_ch);
    int *ft = (int *)(h + b_ch + b_tl);
    cur_cap.resize(nl);
    cur_first.resize(nl);
    int *ccap = cur_cap.data(), *cfirst = cur_first.data();
    parallel_parts(nth, [&, act, lv, chunks, tl, ft, ccap, cfirst](int r) {
      int k0, k1;
      quarter(r, k0, k1);
      int64_t qc = och[r], qt = otl[r], ql = on[r];
      for (int k = k0; k < k1; k++) {
        if (!act[k]) {
          ft[k] = 0;
          continue;
        }
        ft[k] = (int)qt;
        ccap[ql] = lv->seg[k].cap;
        cfirst[ql] = (int)qt;
        ql++;
        memcpy(tl + qt, lv->tiles.data() + lv->first_tile[k], (size_t)lv->ntile[k] * sizeof(Tile));
        qt += lv->ntile[k];
        const int64_t c0 = lv->ibase[k] / WAVE, c1 = (lv->ibase[k] + lv->capal[k]) / WAVE;
        for (int64_t q = c0; q < c1; q++) chunks[qc++] = (int)q;
      }
    });
    RETCHK(c->actbuf.ensure(b_ch + b_tl + b_ft + 256));
    char *d = (char *)c->actbuf.p;
    // ordered on the context's stream, which the next pass's launches follow;
    // the pinned buffer is written again only after the next read-back has
    // synchronised the stream
    HIPCHK(hipMemcpyAsync(d, h, b_ch + b_tl + b_ft, hipMemcpyHostToDevice, c->stream));
    const int64_t nchunks = nch;
    const int ntiles_listed = (int)ntl;
    W.act_chunk = (const int *)d;
    W.nact = nchunks * WAVE;
    tl_cur = (const Tile *)(d + b_ch);
    nt_cur = ntiles_listed;
    ft_cur = (const int *)(d + b_ch + b_tl);
    set_grids();
    return HBT_OK;
  }

  // Bound for the pass from the previous pass of the level: merge levels for
  // segments of up to max_n elements (no count of a pass exceeds the Nlast of
  // the pass before).
  int max_n = 0;

  // tiles of the segments with capacity above w (a prefix of the tile table)
  int tiles_above(int w) const {
    int lo = 0, hi = (int)cur_cap.size();  // first listed segment with cap <= w
    while (lo < hi) {
      const int mid = (lo + hi) / 2;
      if (cur_cap[mid] > w) lo = mid + 1; else hi = mid;
    }
    return lo < (int)cur_cap.size() ? cur_first[lo] : nt_cur;
  }

  int sort_pairs(const int *n) {  // sorts (k0,v0) by the counts n[seg], result in (k0,v0)
    const Tile *tl = tl_cur;
    const int nt = nt_cur;
    if (c->prof_mode == 1) HIPCHK(hipEventRecord(c->evs0, c->stream));
    if (any_big()) tile_sort_kernel<<<gtile, BLOCK, 0, c->stream>>>(tl, nt_big, W.ibase, n, W.k0, W.v0, W.k1, W.v1);
    if (any_wave()) sort_wave_kernel<<<gwave, BLOCK, 0, c->stream>>>(tl, nt_big, nt_tiny, nt, W.ibase, n, W.k0, W.v0);
    // one launch per merge level of the largest segment of the level, over
    // the tiles of the segments whose capacity exceeds the level's width
    // (segments are ordered largest first, so those tiles are a prefix)
    int level = 0;
    for (int w = TILE; w < max_n; w *= 2, level++) {
      const int ntw = tiles_above(w);
      if (ntw == 0) break;
      const int gmerge = (int)std::max<int64_t>(1, std::min<int64_t>(ntw, (int64_t)c->cus * 16));
      merge_pass_kernel<<<gmerge, BLOCK, 0, c->stream>>>(tl, ntw, W.ibase, n, level, W.k0, W.v0, W.k1, W.v1);
    }
    // instrumentation: the sort's device time, or the elements it sorted and
    // their merge levels
    if (c->prof_mode == 1) {
      HIPCHK(hipEventRecord(c->evs1, c->stream));
      HIPCHK(hipEventSynchronize(c->evs1));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->evs0, c->evs1));
      c->sort_ms += ms;
      c->sort_calls++;
    } else if (c->prof_mode == 2) {
      sort_count_kernel<<<gseg, BLOCK, 0, c->stream>>>(n, W.nseg, W.counters);
    }
    return HBT_OK;
  }

  // segmented exclusive scan over the counts n[seg]; the tile totals stay in tsum
  template <class Load, class Store, class Add>
  int seg_scan(Load ld, Store st, Add ad, const int *n) {
    u64 *ts = c->tsum.as<u64>();
    if (any_big()) scan_tiles_kernel<u64><<<gtile, BLOCK, 0, c->stream>>>(tl_cur, nt_big, n, ts, ld, st);
    if (any_wave()) scan_wave_kernel<u64><<<gwave, BLOCK, 0, c->stream>>>(tl_cur, nt_big, nt_tiny, nt_cur, n, ts, ld, st);
    // only segments of more than one tile have tiles after their first
    const int nmulti = tiles_above(TILE);
    if (nmulti > 0) scan_add_kernel<u64><<<gtile, BLOCK, 0, c->stream>>>(tl_cur, nmulti, n, ts, ad);
    return HBT_OK;
  }
  const int *count_row(int row) const { return W.ncount + (size_t)row * W.nseg; }

  // tree over the pass's sources: keys, sort, structure, moments
  void bbox() {
    if (any_big()) bbox_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big);
    if (any_wave()) bbox_wave_kernel<<<gwave, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big, nt_tiny, nt_cur);
  }

  int build_tree(bool copy_sp) {
    bbox();
    tree_params_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P);
    keys_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(sort_pairs(count_row(NC_SRC)));
    if (any_big()) emit_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big, copy_sp ? 1 : 0);
    if (any_wave()) emit_wave_kernel<<<gwave, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big, nt_cur, copy_sp ? 1 : 0);
    RETCHK(seg_scan(PopLoad{W.ibase, W.bits}, PreStore{W.ibase, W.pre}, PreAdd{W.ibase, W.pre},
                    count_row(NC_SRC)));
    items_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    if (any_big()) {
      moments_tile_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big);
      moments_super_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big);
      if (max_n > SUPER) moments_cross_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big);
    }
    if (any_small()) moments_wave_kernel<4><<<gsmall, BLOCK, 0, c->stream>>>(W, tl_cur, nt_big, nt_tiny);
    if (any_tiny()) moments_wave_kernel<1><<<gtiny, BLOCK, 0, c->stream>>>(W, tl_cur, nt_tiny, nt_cur);
    return HBT_OK;
  }

  // walk variants: wave-shared traversal (default) or one independent walk per lane
  template <int GROUP>
  void launch_shared_walk(bool count) {
    if (P.replay) {
      if (count) walk_shared_kernel<GROUP, true, true><<<gw, BLOCK, 0, c->stream>>>(W, P, W.items, c->tune.xcd_run);
      else walk_shared_kernel<GROUP, false, true><<<gw, BLOCK, 0, c->stream>>>(W, P, W.items, c->tune.xcd_run);
    } else {
      if (count) walk_shared_kernel<GROUP, true, false><<<gw, BLOCK, 0, c->stream>>>(W, P, W.items, c->tune.xcd_run);
      else walk_shared_kernel<GROUP, false, false><<<gw, BLOCK, 0, c->stream>>>(W, P, W.items, c->tune.xcd_run);
    }
  }

  void launch_walk(bool count) {
    if (c->walk_mode == 0) {
      const int *g = c->tune.walk_group;
      const int uniform = (g[0] == g[1] && g[1] == g[2]) ? g[0] : 0;
      switch (uniform) {
        case 0: launch_shared_walk<0>(count); break;  // per-segment bins
        case 8: launch_shared_walk<8>(count); break;
        case 16: launch_shared_walk<16>(count); break;
        case 32: launch_shared_walk<32>(count); break;
        default: launch_shared_walk<64>(count); break;
      }
    } else {
      if (count) walk_kernel<true><<<gs, BLOCK, 0, c->stream>>>(W, P);
      else walk_kernel<false><<<gs, BLOCK, 0, c->stream>>>(W, P);
    }
  }

  int pass() {
    pass_setup_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P);
    if (P.max_sample > 0) sample_flags_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(build_tree(true));
    if (c->prof_mode == 1) HIPCHK(hipEventRecord(c->evw0, c->stream));
    launch_walk(c->prof_mode == 2);
    if (c->prof_mode == 2) tree_items_kernel<<<gseg, BLOCK, 0, c->stream>>>(W);
    if (c->prof_mode == 1) HIPCHK(hipEventRecord(c->evw1, c->stream));
    // partition
    RETCHK(seg_scan(PartLoad{W.ibase, W.el, W.E}, S64Store{W.ibase, W.scan64},
                    S64Add{W.ibase, W.scan64}, count_row(NC_PART)));
    partition_counts_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, ft_cur, c->tsum.as<u64>());
    partition_rank_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    partition_k_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    partition_hole_kernel<<<gseg, BLOCK, 0, c->stream>>>(W);
    partition_move_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    partition_done_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P);
    if (P.max_sample > 0) {
      disrupt_find_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
      disrupt_swap_kernel<<<gseg, BLOCK, 0, c->stream>>>(W);
    }
    // spatial target list keeps the bound particles
    bound_flags_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(seg_scan(BoundLoad{W.ibase, W.sp, W.isb}, PreStore{W.ibase, W.pre},
                    PreAdd{W.ibase, W.pre}, count_row(NC_BOUND)));
    spatial_compact_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    // the compacted list (in el2) becomes the spatial list and the old one
    // the next pass's scratch; only the active segments read either, and
    // they are rewritten every pass.  A captured pass sequence replays fixed
    // pointers, so it copies the list back instead.
    if (swap_sp) std::swap(W.sp, W.el2);
    else spatial_copy_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    // frames, convergence (the frame sums read the bound block only, in Elist order)
    if (P.replay) {
      frames_serial_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P);
    } else {
      if (any_big()) frames_partial_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, P, tl_cur, nt_big, c->tpart.as<double>());
      if (any_wave())
        frames_partial_wave_kernel<<<gwave, BLOCK, 0, c->stream>>>(W, P, tl_cur, nt_big, nt_tiny, nt_cur,
                                                                   c->tpart.as<double>());
      // the run sums go after the tile partials in the same buffer
      double *gpart = c->tpart.as<double>() + (int64_t)(nt_cur + 1) * FRAME_SUMS;
      if (max_n > SUPER)
        frames_runs_kernel<<<grid_for(nt_cur, c->cus), BLOCK, 0, c->stream>>>(W, tl_cur, nt_cur, c->tpart.as<double>(),
                                                                             gpart);
      frames_combine_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P, ft_cur, c->tpart.as<double>(), gpart);
    }
    // energy sort of the removed block, and of the bound block on convergence
    energy_keys_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(sort_pairs(count_row(NC_SORT)));
    energy_writeback_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    outputs();
    HIPCHK(hipGetLastError());
    return HBT_OK;
  }

  // The outputs of the record and its list for the segments flagged F_FINAL.
  void outputs() {
    output_list_kernel<<<gs, BLOCK, 0, c->stream>>>(W, 0);
    if (P.replay) {
      output_record_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P, 0);
    } else {
      if (any_big()) output_partial_kernel<<<gtile, BLOCK, 0, c->stream>>>(W, P, tl_cur, nt_big, c->tpart.as<double>());
      if (any_wave())
        output_partial_wave_kernel<<<gwave, BLOCK, 0, c->stream>>>(W, P, tl_cur, nt_big, nt_tiny, nt_cur,
                                                                   c->tpart.as<double>());
      output_combine_kernel<<<gwseg, BLOCK, 0, c->stream>>>(W, P, ft_cur, c->tpart.as<double>());
    }
  }

  // RefineBindingEnergyOrder for the segments that converged with more than
  // MaxSampleSize bound particles, then their outputs (once per level, after
  // its passes; those segments finished in earlier passes, so every slot and
  // tile is listed again).
  int refine() {
    list_all();
    max_n = std::min(L->maxcap, P.max_sample);
    refine_setup_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P);
    refine_targets_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(build_tree(false));
    launch_walk(false);
    refine_keys_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(sort_pairs(count_row(NC_SORT)));
    energy_writeback_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    refine_done_kernel<<<gseg, BLOCK, 0, c->stream>>>(W);
    outputs();
    HIPCHK(hipGetLastError());
    return HBT_OK;
  }

  // The level's shuffles (std::random_shuffle, src/subhalo_unbind.cpp:302):
  // the level's rand() values, then the parallel replay of each list's swaps
  // on its column positions.  The columns are gathered afterwards.
  int shuffle() {
    if (L->ndraws <= 0) return HBT_OK;
    RETCHK(c->draws.ensure((size_t)L->ndraws * 4));
    const int64_t nchunk = (L->ndraws + RCHUNK - 1) / RCHUNK;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>((nchunk * WAVE + BLOCK - 1) / BLOCK, (int64_t)c->cus * 16));
    rand_gen_kernel<<<g, BLOCK, 0, c->stream>>>(c->randtab.as<uint32_t>(), c->rwin.as<uint32_t>(), L->ndraws,
                                                  c->draws.as<int>());
    shuffle_keys_kernel<<<gs, BLOCK, 0, c->stream>>>(W, c->draws.as<int>());
    RETCHK(sort_pairs(count_row(NC_SRC)));
    shuffle_links_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    shuffle_apply_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    HIPCHK(hipGetLastError());
    return HBT_OK;
  }

  // Spatial target order of the shuffled lists: their octant digits sorted.
  int prelude() {
    if (L->ndraws <= 0) return HBT_OK;
    prelude_setup_kernel<<<gseg, BLOCK, 0, c->stream>>>(W);
    bbox();
    tree_params_kernel<<<gseg, BLOCK, 0, c->stream>>>(W, P);
    keys_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    RETCHK(sort_pairs(count_row(NC_SRC)));
    prelude_sp_kernel<<<gs, BLOCK, 0, c->stream>>>(W);
    HIPCHK(hipGetLastError());
    return HBT_OK;
  }
};

int ensure_slots(hbt_ctx *c, int64_t S, int64_t nseg, int64_t ntiles) {
  const size_t s = (size_t)S;
  RETCHK(c->gidx.ensure(s * 4));
  RETCHK(c->pos.ensure(s * 16));
  RETCHK(c->vel.ensure(s * 16));
  RETCHK(c->E.ensure(s * 4));
  RETCHK(c->Er.ensure(s * 4));
  RETCHK(c->el.ensure(s * 4));
  RETCHK(c->el2.ensure(s * 4));
  RETCHK(c->sp.ensure(s * 4));
  RETCHK(c->k0.ensure(s * 8));
  RETCHK(c->k1.ensure(s * 8));
  RETCHK(c->v0.ensure(s * 4));
  RETCHK(c->v1.ensure(s * 4));
  RETCHK(c->dl.ensure(s));
  RETCHK(c->bits.ensure(s * 4));
  RETCHK(c->pre.ensure(s * 4));
  RETCHK(c->nlo.ensure(s * 4));
  RETCHK(c->nhi.ensure(s * 4));
  RETCHK(c->npos.ensure(s * 4));
  // two items per slot (leaves and nodes, plus each segment's sentinel) and
  // two spare items that the group walk may read past the last segment's end
  RETCHK(c->items.ensure((s * 2 + 2) * sizeof(Item)));
  RETCHK(c->scan64.ensure(s * 8));
  RETCHK(c->cross.ensure(s * 4));
  RETCHK(c->cross2.ensure(s * 4));
  RETCHK(c->tcount.ensure((size_t)(ntiles + 1) * 4));
  RETCHK(c->pk.ensure((s + 1) * 4));
  RETCHK(c->gk.ensure((s + 1) * 4));
  RETCHK(c->isb.ensure(s));
  RETCHK(c->seg_of_slot.ensure(s * 4));
  RETCHK(c->ncount.ensure((size_t)(NC_ROWS * nseg + 1) * 4));
  RETCHK(c->segact.ensure((size_t)(nseg + 1) * 4));
  RETCHK(c->tsum.ensure((size_t)(ntiles + 1) * 8));
  RETCHK(c->tpart.ensure((size_t)(ntiles + 1) * 8 * REC_SUMS));
  RETCHK(c->active.ensure(16));
  RETCHK(c->counters.ensure(128));
  return HBT_OK;
}

Work make_work(hbt_ctx *c, const Level &L, const hbt_particles *parts, hbt_subhalo *subs,
               int64_t *out_index, float *out_energy) {
  Work W{};
  W.seg = c->seg.as<Seg>();
  W.nseg = (int)L.seg.size();
  W.nslots = L.nslots;
  W.act_chunk = nullptr;
  W.nact = L.nslots;
  W.segact = c->segact.as<int>();
  W.seg_of_slot = c->seg_of_slot.as<int>();
  W.ibase = c->ibase.as<int64_t>();
  W.ncount = c->ncount.as<int>();
  W.g_index = c->list_index;
  W.g_packed = c->packed;
  if (parts) {
    W.g_id = parts->id; W.g_x = parts->x; W.g_y = parts->y; W.g_z = parts->z;
    W.g_vx = parts->vx; W.g_vy = parts->vy; W.g_vz = parts->vz; W.g_m = parts->mass;
    W.g_type = parts->type;
  }
  const size_t S = (size_t)L.nslots;
  W.gidx = c->gidx.as<int>();
  W.pos = c->pos.as<float4>();
  W.vel = c->vel.as<float4>();
  W.E = c->E.as<float>();
  W.Er = c->Er.as<float>();
  W.el = c->el.as<int>(); W.el2 = c->el2.as<int>(); W.sp = c->sp.as<int>();
  W.k0 = c->k0.as<u64>(); W.k1 = c->k1.as<u64>(); W.v0 = c->v0.as<u32>(); W.v1 = c->v1.as<u32>();
  W.dl = c->dl.as<signed char>();
  W.bits = c->bits.as<u32>();
  W.pre = c->pre.as<int>(); W.nlo = c->nlo.as<int>(); W.nhi = c->nhi.as<int>();
  W.npos = c->npos.as<int>();
  W.items = c->items.as<Item>();
  W.cross = c->cross.as<int>();
  W.cross2 = c->cross2.as<int>();
  W.tcount = c->tcount.as<int>();
  W.scan64 = c->scan64.as<u64>();
  W.pk = c->pk.as<int>(); W.gk = c->gk.as<int>();
  W.isb = c->isb.as<unsigned char>();
  W.active = c->active.as<int>();
  W.counters = c->counters.as<unsigned long long>();
  W.subs = subs;
  W.out_index = out_index;
  W.out_energy = out_energy;
  return W;
}

// After a level's passes: the refinements that wait for them.
int finish_level(hbt_ctx *c, Runner &R) {
  if (!R.P.refine) return HBT_OK;
  HIPCHK(hipMemcpyAsync(c->h_active + 3, R.W.active + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->h_active[3] == 0) return HBT_OK;
  RETCHK(R.refine());
  HIPCHK(hipStreamSynchronize(c->stream));
  return HBT_OK;
}

// Runs one level: builds the working lists, iterates passes until every
// segment has converged or been disrupted, writes outputs.
int run_level(hbt_ctx *c, Level &L, const hbt_particles *parts, hbt_subhalo *subs,
              int64_t *out_index, float *out_energy, int64_t *passes) {
  if (L.seg.empty()) return HBT_OK;
  const int64_t nseg = (int64_t)L.seg.size();
  const auto ts0 = std::chrono::steady_clock::now();
  RETCHK(ensure_slots(c, L.nslots, nseg, (int64_t)L.tiles.size()));
  RETCHK(stage_level(c, L));
  Runner R;
  R.c = c;
  R.L = &L;
  R.P = make_dev_params(c->params);
  R.W = make_work(c, L, parts, subs, out_index, out_energy);
  R.max_n = L.maxcap;
  R.list_all();
  R.gseg = grid_for(nseg, c->cus);
  R.gwseg = (int)std::min<int64_t>(((int64_t)nseg * WAVE + BLOCK - 1) / BLOCK, (int64_t)c->cus * 8);
  const int gsb = (int)std::min<int64_t>(nseg, (int64_t)c->cus * 8);
  R.gblk = gsb;
  seg_expand_kernel<<<R.gseg, BLOCK, 0, c->stream>>>(R.W.seg, c->segin.as<SegIn>(), (int)nseg);
  // segment bases are multiples of WAVE (capacities are rounded up to it)
  seg_of_chunk_kernel<<<R.gs, BLOCK, 0, c->stream>>>((int)nseg, R.W.ibase, L.nslots, c->seg_of_slot.as<int>());
  build_lists_kernel<<<gsb, BLOCK, 0, c->stream>>>(R.W, c->child_ptr.as<int>(), c->child_list.as<int>());
  seg_init_kernel<<<R.gseg, BLOCK, 0, c->stream>>>(R.W, R.P);
  if (L.ndraws > 0) {
    // shuffled lists: positions, the shuffle, then the columns in shuffled order
    gather_kernel<<<R.gs, BLOCK, 0, c->stream>>>(R.W, 1);
    RETCHK(R.shuffle());
    gather_kernel<<<R.gs, BLOCK, 0, c->stream>>>(R.W, 2);
  } else {
    gather_kernel<<<R.gs, BLOCK, 0, c->stream>>>(R.W, 3);
  }
  el_init_kernel<<<R.gs, BLOCK, 0, c->stream>>>(R.W);
  RETCHK(R.prelude());
  output_list_kernel<<<R.gs, BLOCK, 0, c->stream>>>(R.W, 1);
  output_record_kernel<<<R.gseg, BLOCK, 0, c->stream>>>(R.W, R.P, 1);
  HIPCHK(hipGetLastError());
  if (getenv("HBT_PASS_TIMES") && atoi(getenv("HBT_PASS_TIMES"))) {
    HIPCHK(hipStreamSynchronize(c->stream));
    fprintf(stderr, "[hbt] level set up (staging, lists, gather) in %.3f ms\n",
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts0).count());
  }
  // every pass that does not converge removes at least one particle
  static const bool debug = getenv("HBT_DEBUG") != nullptr;
  // The pass sequence has static launch shapes and arguments for the whole
  // level, so it is captured once as a hipGraph (graph_passes passes plus the
  // read-back of the active count) and replayed until no segment is active.
  // the plain form (one subhalo per level) launches directly, as a per-subhalo
  // Unbind loop does; graphs serve the batched form
  const bool graph = c->tune.use_graph && c->prof_mode == 0 && !debug && L.seg.size() > 1;
  const int graph_passes = c->tune.graph_passes;
  if (graph) {
    hipGraph_t g = nullptr;
    hipGraphExec_t gx = nullptr;
    auto release = [&]() {
      if (gx) (void)hipGraphExecDestroy(gx);
      if (g) (void)hipGraphDestroy(g);
    };
    R.swap_sp = false;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = HBT_OK;
    for (int p = 0; p < graph_passes && rc == HBT_OK; p++) rc = R.pass();
    hipError_t e1 = hipMemcpyAsync(c->h_active, R.W.active, sizeof(int), hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (rc != HBT_OK || e1 != hipSuccess || e2 != hipSuccess) {
      release();
      return rc != HBT_OK ? rc : fail(HBT_ERR_DEVICE, "pass graph capture failed");
    }
    hipError_t e3 = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
      release();
      return fail(HBT_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    // every pass that does not converge removes at least one particle
    const int max_launch = (L.maxcap + 2 + graph_passes - 1) / graph_passes;
    for (int it = 0; it < max_launch; it++) {
      hipError_t e = hipGraphLaunch(gx, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        release();
        return fail(HBT_ERR_DEVICE, std::string("pass graph: ") + hipGetErrorString(e));
      }
      *passes += graph_passes;
      if (*c->h_active == 0) break;
    }
    release();
    if (*c->h_active != 0) return fail(HBT_ERR_DEVICE, "unbinding passes did not converge within the pass cap");
    return finish_level(c, R);
  }
  // HBT_PASS_TIMES=1: per pass, the GPU time between its first and last
  // launch and the host time, printed to stderr (diagnostics)
  static const bool pass_times = getenv("HBT_PASS_TIMES") && atoi(getenv("HBT_PASS_TIMES"));
  hipEvent_t pe0 = nullptr, pe1 = nullptr;
  if (pass_times) {
    HIPCHK(hipEventCreate(&pe0));
    HIPCHK(hipEventCreate(&pe1));
  }
  // every pass that does not converge removes at least one particle
  for (int it = 0; it < L.maxcap + 2; it++) {
    const auto h0 = std::chrono::steady_clock::now();
    if (pass_times) HIPCHK(hipEventRecord(pe0, c->stream));
    RETCHK(R.pass());
    if (pass_times) HIPCHK(hipEventRecord(pe1, c->stream));
    (*passes)++;
    if (debug) {
      std::vector<Seg> hs(nseg);
      HIPCHK(hipMemcpyAsync(hs.data(), R.W.seg, nseg * sizeof(Seg), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      const Seg &S = hs[0];
      std::vector<float> e(std::min(8, S.len));
      std::vector<int> el(std::min(8, S.len));
      if (!e.empty()) {
        HIPCHK(hipMemcpyAsync(e.data(), R.W.E + S.ibase, e.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipMemcpyAsync(el.data(), R.W.el + S.ibase, el.size() * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      }
      fprintf(stderr,
              "[hbt] pass %d seg0: flags %d len %d nbound %d nlast %d n_src %d n_tgt %d items %d "
              "nP %d nG %d K %d hole %d D %d dlen %d Len %g iter %d\n",
              it, S.flags, S.len, S.nbound, S.nlast, S.n_src, S.n_tgt, S.n_items, S.nP, S.nG, S.K,
              S.hole, S.D, S.dlen, S.Len, S.iter);
      fprintf(stderr, "[hbt]   E:");
      for (size_t i = 0; i < e.size(); i++) fprintf(stderr, " %g", e[i]);
      fprintf(stderr, "  el:");
      for (size_t i = 0; i < el.size(); i++) fprintf(stderr, " %d", el[i]);
      fprintf(stderr, "\n");
    }
    HIPCHK(hipMemcpyAsync(c->h_active, R.W.active, 4 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (pass_times) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, pe0, pe1));
      const double hms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
      fprintf(stderr, "[hbt] level of %zu: pass %d gpu %.3f ms host %.3f ms, active after %d, listed %d\n",
              L.seg.size(), it + 1, ms, hms, c->h_active[0], R.nlisted);
    }
    // the next pass is bounded by this one
    R.max_n = c->h_active[1];
    // once a quarter of the listed segments have finished, the next passes
    // visit only the slots and tiles of the active ones
    if (c->h_active[0] > 0 && 4 * c->h_active[0] <= 3 * R.nlisted && compact_passes()) RETCHK(R.list_active());
    if (c->prof_mode == 1) {
      float wms = 0;
      HIPCHK(hipEventElapsedTime(&wms, c->evw0, c->evw1));
      c->walk_ms += wms;
      c->walk_launches++;
    }
    if (*c->h_active == 0) break;
  }
  if (pe0) (void)hipEventDestroy(pe0);
  if (pe1) (void)hipEventDestroy(pe1);
  // every pass that does not converge removes a particle, so with a valid
  // BoundMassPrecision the loop ends before the cap
  if (*c->h_active != 0) return fail(HBT_ERR_DEVICE, "unbinding passes did not converge within the pass cap");
  return finish_level(c, R);
}

// Gathers the records' planning fields, checks them and builds the nests:
// capacities, heights, children (CSR) and the post-order; detects cycles.
// n_list >= 0: each record's list must also lie inside [0, n_list), checked
// during the same gather.
int plan_nests(const hbt_subhalo *subs, int64_t nsub, Nest &N, int64_t n_list = -1) {
  N.n = nsub;
  N.f.resize(nsub);
  N.height.assign(nsub, 0);
  N.ch_ptr.assign(nsub + 1, 0);
  static thread_local std::vector<int64_t> parent;
  parent.resize(nsub);
  // the records are read once, by a few threads for large batches (the
  // fields sit in two cache lines of each 240-byte record)
  const int nth = nsub >= 65536 ? 4 : 1;
  std::atomic<int> bad{0}, outside{0};
  // (the worker threads get plain pointers: `parent` is thread-local)
  SubPlan *fp = N.f.data();
  int64_t *pp = parent.data();
  auto gather = [subs, nsub, n_list, fp, pp, &bad, &outside](int64_t lo, int64_t hi) {
    for (int64_t s = lo; s < hi; s++) {
      const hbt_subhalo &r = subs[s];
      if (r.count < 0 || r.parent >= nsub || r.parent == s) bad.store(1, std::memory_order_relaxed);
      if (n_list >= 0 && (r.offset < 0 || r.count < 0 || r.offset + r.count > n_list))
        outside.store(1, std::memory_order_relaxed);
      fp[s] = SubPlan{r.offset, r.count, r.nbound_prev, r.out_offset, r.count};
      pp[s] = r.parent;
    }
  };
  if (nth == 1) {
    gather(0, nsub);
  } else {
    std::vector<std::thread> th;
    for (int t = 1; t < nth; t++) th.emplace_back(gather, nsub * t / nth, nsub * (t + 1) / nth);
    gather(0, nsub / nth);
    for (auto &x : th) x.join();
  }
  if (outside.load()) return fail(HBT_ERR_INVALID_ARG, "subhalo list outside the list positions");
  if (bad.load()) return fail(HBT_ERR_INVALID_ARG, "bad subhalo record (count/parent)");
  bool nested = false;
  for (int64_t s = 0; s < nsub; s++)
    if (parent[s] >= 0) {
      N.ch_ptr[parent[s] + 1]++;
      nested = true;
    }
  for (int64_t s = 0; s < nsub; s++) N.ch_ptr[s + 1] += N.ch_ptr[s];
  N.ch_idx.resize(N.ch_ptr[nsub]);
  N.order.clear();
  N.order.reserve(nsub);
  if (!nested) {
    for (int64_t s = 0; s < nsub; s++) N.order.push_back((int)s);
    return HBT_OK;
  }
  {
    static thread_local std::vector<int> fill;
    fill.assign(N.ch_ptr.begin(), N.ch_ptr.end() - 1);
    for (int64_t s = 0; s < nsub; s++)
      if (parent[s] >= 0) N.ch_idx[fill[parent[s]]++] = (int)s;
  }
  // post-order from the roots (index order) gives heights and capacities
  std::vector<std::pair<int, int>> st;
  for (int64_t r = 0; r < nsub; r++) {
    if (parent[r] >= 0) continue;
    st.push_back({(int)r, 0});
    while (!st.empty()) {
      auto &t = st.back();
      if (t.second < N.nchild(t.first)) {
        const int ch = N.child(t.first)[t.second++];
        st.push_back({ch, 0});
        if ((int64_t)st.size() > nsub + 1) return fail(HBT_ERR_INVALID_ARG, "nest too deep");
      } else {
        const int s = t.first;
        N.order.push_back(s);
        int h = 0;
        for (int i = 0; i < N.nchild(s); i++) {
          const int ch = N.child(s)[i];
          h = std::max(h, N.height[ch] + 1);
          N.f[s].cap += N.f[ch].cap;
        }
        N.height[s] = h;
        st.pop_back();
      }
    }
  }
  if ((int64_t)N.order.size() != nsub) return fail(HBT_ERR_INVALID_ARG, "cycle in parent links");
  return HBT_OK;
}

}  // namespace

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

const char *hbt_last_error(void) { return g_err.c_str(); }

int hbt_version(int *major, int *minor) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  return HBT_OK;
}

namespace {
// the scratch address of hbt_direct_potential_noalloc on a device (hbt_direct.hpp)
int noalloc_scratch_setup(int device);
}  // namespace

// Parameters the unbinding cannot run with: a BoundMassPrecision outside
// (0, 1] never converges (or converges at once), a negative sample size.
static int check_params(const hbt_params &p) {
  if (!(p.bound_mass_precision > 0.f && p.bound_mass_precision <= 1.f))
    return fail(HBT_ERR_INVALID_ARG, "bound_mass_precision must lie in (0, 1]");
  if (p.max_sample_size < 0) return fail(HBT_ERR_INVALID_ARG, "max_sample_size must be >= 0");
  if (!(p.softening > 0.f)) return fail(HBT_ERR_INVALID_ARG, "softening must be > 0");
  if (p.precision != HBT_PREC_FAST && p.precision != HBT_PREC_REPLAY && p.precision != HBT_PREC_FP32)
    return fail(HBT_ERR_INVALID_ARG, "unknown precision");
  return HBT_OK;
}

int hbt_ctx_create(const hbt_params *params, int device, hbt_ctx **out) {
  if (!params || !out) return fail(HBT_ERR_INVALID_ARG, "null argument");
  RETCHK(check_params(*params));
  hbt_ctx *c = new (std::nothrow) hbt_ctx();
  if (!c) return fail(HBT_ERR_OUT_OF_MEMORY, "context allocation");
  c->device = device;
  c->params = *params;
  if (const char *wm = getenv("HBT_WALK_MODE")) c->walk_mode = atoi(wm) ? 1 : 0;
  if (const char *wg = getenv("HBT_WALK_GROUP")) {
    const int g = atoi(wg);
    for (int b = 0; b < 3; b++) c->tune.walk_group[b] = (g == 8 || g == 16 || g == 32) ? g : 64;
  }
  if (const char *g = getenv("HBT_GRAPH")) c->tune.use_graph = atoi(g) ? 1 : 0;
  if (const char *g = getenv("HBT_GRAPH_PASSES")) c->tune.graph_passes = std::max(1, atoi(g));
  if (const char *g = getenv("HBT_XCD_RUN")) {
    const int r = atoi(g);
    c->tune.xcd_run = (r >= 1 && r <= 64 && (r & (r - 1)) == 0) ? r : 1;
  }
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void **)&c->h_active, sizeof(int) * 8, hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipEventCreate(&c->evw0);
  if (e == hipSuccess) e = hipEventCreate(&c->evw1);
  if (e == hipSuccess) e = hipEventCreate(&c->evs0);
  if (e == hipSuccess) e = hipEventCreate(&c->evs1);
  if (e == hipSuccess) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cus = prop.multiProcessorCount;
  }
  if (e != hipSuccess) {
    delete c;
    return fail(HBT_ERR_DEVICE, std::string("context setup: ") + hipGetErrorString(e));
  }
  // the device's scratch address for the entry point that allocates nothing
  if (const int rc = noalloc_scratch_setup(device); rc != HBT_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return HBT_OK;
}

int hbt_ctx_destroy(hbt_ctx *ctx) {
  if (!ctx) return HBT_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return HBT_OK;
}

int hbt_ctx_set_params(hbt_ctx *ctx, const hbt_params *params) {
  if (!ctx || !params) return fail(HBT_ERR_INVALID_ARG, "null argument");
  RETCHK(check_params(*params));
  ctx->params = *params;
  return HBT_OK;
}

int hbt_set_rand_state(hbt_ctx *ctx, uint32_t seed, int64_t draws) {
  if (!ctx || draws < 0) return fail(HBT_ERR_INVALID_ARG, "bad argument");
  ctx->rand_seed = seed;
  ctx->rand_pos = draws;
  return HBT_OK;
}

int hbt_get_rand_state(hbt_ctx *ctx, uint32_t *seed, int64_t *draws) {
  if (!ctx) return fail(HBT_ERR_INVALID_ARG, "null argument");
  if (seed) *seed = ctx->rand_seed;
  if (draws) *draws = ctx->rand_pos;
  return HBT_OK;
}

int hbt_ctx_stream(hbt_ctx *ctx, void **stream) {
  if (!ctx || !stream) return fail(HBT_ERR_INVALID_ARG, "null argument");
  *stream = (void *)ctx->stream;
  return HBT_OK;
}

int hbt_plan_batch(hbt_subhalo *subs, int64_t nsub, int64_t *total) {
  if (nsub < 0 || (nsub > 0 && !subs) || !total) return fail(HBT_ERR_INVALID_ARG, "null argument");
  Nest N;
  RETCHK(plan_nests(subs, nsub, N));
  int64_t off = 0;
  for (int64_t s = 0; s < nsub; s++) {
    subs[s].out_offset = off;
    off += N.f[s].cap;
  }
  *total = off;
  return HBT_OK;
}

static int unbind_batch_impl(hbt_ctx *ctx, const hbt_particles *particles, int64_t n_list,
                             hbt_subhalo *subs, const hbt_subhalo *subs_host, int64_t nsub,
                             int64_t *out_index, float *out_energy, int plain, const int64_t *out_cap,
                             bool planned);

// The host plan of the last batch (memory reused from call to call, so a
// repeated batch pays no page faults for it).
static thread_local Nest t_nest;

int hbt_unbind_batch(hbt_ctx *ctx, const hbt_particles *particles, hbt_subhalo *subs,
                     const hbt_subhalo *subs_host, int64_t nsub, int64_t *out_index,
                     float *out_energy, int plain) {
  if (!ctx || !particles) return fail(HBT_ERR_INVALID_ARG, "null argument");
  ctx->list_index = nullptr;
  return unbind_batch_impl(ctx, particles, particles->n, subs, subs_host, nsub, out_index, out_energy,
                           plain, nullptr, false);
}

static bool tuning_valid(const hbt_tuning &t) {
  for (int b = 0; b < 3; b++) {
    const int g = t.walk_group[b];
    if (g != 8 && g != 16 && g != 32 && g != 64) return false;
  }
  return t.bin_edge[0] >= 0 && t.bin_edge[1] >= t.bin_edge[0] && t.graph_passes >= 1 &&
         t.graph_passes <= 16 && (t.use_graph == 0 || t.use_graph == 1) && t.xcd_run >= 1 &&
         t.xcd_run <= 64 && (t.xcd_run & (t.xcd_run - 1)) == 0;
}

int hbt_get_tuning(hbt_ctx *ctx, hbt_tuning *out) {
  if (!ctx || !out) return fail(HBT_ERR_INVALID_ARG, "null argument");
  *out = ctx->tune;
  return HBT_OK;
}

int hbt_set_tuning(hbt_ctx *ctx, const hbt_tuning *tuning) {
  if (!ctx || !tuning) return fail(HBT_ERR_INVALID_ARG, "null argument");
  if (!tuning_valid(*tuning))
    return fail(HBT_ERR_INVALID_ARG, "tuning: walk groups must be 8/16/32/64, bin edges ascending, "
                                     "graph_passes in 1..16, use_graph 0 or 1, xcd_run a power "
                                     "of two in 1..64");
  ctx->tune = *tuning;
  return HBT_OK;
}

int hbt_autotune(hbt_ctx *ctx, const hbt_particles *particles, hbt_subhalo *subs,
                 const hbt_subhalo *subs_host, int64_t nsub, int64_t *out_index, float *out_energy,
                 int reps, hbt_tuning *chosen) {
  if (!ctx || !particles || !chosen || reps < 1) return fail(HBT_ERR_INVALID_ARG, "bad argument");
  const hbt_tuning start = ctx->tune;
  hbt_tuning best = start;
  double best_ms = 0;
  // every run draws its shuffles from the same point of the rand() stream;
  // the stream is left as after one run
  const int64_t pos0 = ctx->rand_pos;
  // device time of the batch under t, best of reps; t replaces best when it is
  // faster by more than 1% (earlier, simpler settings win ties)
  auto trial = [&](const hbt_tuning &t, bool first) -> int {
    ctx->tune = t;
    double ms = 1e300;
    for (int r = 0; r < reps; r++) {
      ctx->rand_pos = pos0;
      const int rc = hbt_unbind_batch(ctx, particles, subs, subs_host, nsub, out_index, out_energy, 0);
      if (rc != HBT_OK) return rc;
      ms = std::min(ms, ctx->last_ms);
    }
    if (first || ms < 0.99 * best_ms) {
      best = t;
      best_ms = ms;
    }
    return HBT_OK;
  };
  // 0. the current setting is the one to beat
  int rc = trial(start, true);
  // 1. one group size for every bin
  const int groups[4] = {64, 32, 16, 8};
  for (int i = 0; i < 4 && rc == HBT_OK; i++) {
    hbt_tuning t = best;
    for (int b = 0; b < 3; b++) t.walk_group[b] = groups[i];
    t.bin_edge[0] = 0;
    t.bin_edge[1] = 1 << 30;
    if (!memcmp(&t, &best, sizeof t)) continue;
    rc = trial(t, false);
  }
  // 2. a separate group size for small subhalos
  const int edges[2] = {256, 2048};
  for (int e = 0; e < 2 && rc == HBT_OK; e++) {
    for (int i = 0; i < 2 && rc == HBT_OK; i++) {
      const int g0 = groups[i];
      if (g0 == best.walk_group[2]) continue;
      hbt_tuning t = best;
      t.walk_group[0] = g0;
      t.walk_group[1] = t.walk_group[2];
      t.bin_edge[0] = edges[e];
      t.bin_edge[1] = 1 << 30;
      rc = trial(t, false);
    }
  }
  // 3. XCD runs of the walk's blocks
  const int runs[3] = {16, 4, 1};
  for (int i = 0; i < 3 && rc == HBT_OK; i++) {
    if (runs[i] == best.xcd_run) continue;
    hbt_tuning t = best;
    t.xcd_run = runs[i];
    rc = trial(t, false);
  }
  // 4. direct launches (each pass bounded by the previous one) or a HIP graph
  //    of 1-4 passes per launch
  if (rc == HBT_OK) {
    hbt_tuning t = best;
    t.use_graph = 1 - best.use_graph;
    rc = trial(t, false);
  }
  for (int gp = 1; gp <= 4 && rc == HBT_OK && best.use_graph; gp++) {
    if (gp == best.graph_passes) continue;
    hbt_tuning t = best;
    t.graph_passes = gp;
    rc = trial(t, false);
  }
  if (rc != HBT_OK) {
    ctx->tune = start;
    return rc;
  }
  ctx->tune = best;
  *chosen = best;
  return HBT_OK;
}

// Lists that cover a good part of the snapshot are gathered from packed
// records (one streaming pass over the snapshot, then one 32-byte read per
// listed particle); sparse lists read the columns directly.
static int pack_snapshot(hbt_ctx *ctx, const hbt_particles *snapshot, int64_t n_list) {
  ctx->packed = nullptr;
  if (n_list > 0 && snapshot->n <= 4 * n_list) {
    RETCHK(ctx->pack.ensure((size_t)snapshot->n * 2 * sizeof(float4)));
    pack_particles_kernel<<<grid_for(snapshot->n, ctx->cus), BLOCK, 0, ctx->stream>>>(*snapshot,
                                                                                      ctx->pack.as<float4>());
    HIPCHK(hipGetLastError());
    ctx->packed = ctx->pack.as<float4>();
  }
  return HBT_OK;
}

int hbt_unbind_batch_indexed(hbt_ctx *ctx, const hbt_particles *snapshot, const int64_t *list_index,
                             int64_t n_list, hbt_subhalo *subs, const hbt_subhalo *subs_host,
                             int64_t nsub, int64_t *out_index, float *out_energy, int plain) {
  if (!ctx || !snapshot || (n_list > 0 && !list_index) || n_list < 0)
    return fail(HBT_ERR_INVALID_ARG, "null argument");
  if (n_list >= (int64_t)INT32_MAX) return fail(HBT_ERR_UNSUPPORTED, "more than 2^31 list entries");
  HIPCHK(hipSetDevice(ctx->device));
  if (nsub > 0 && !subs_host) return fail(HBT_ERR_INVALID_ARG, "null argument");
  // the packed snapshot and the list positions belong to this call: cleared on every return
  struct Reset {
    hbt_ctx *c;
    ~Reset() { c->packed = nullptr; c->list_index = nullptr; }
  } reset{ctx};
  int cnt[2] = {0, 0};
  bool planned = false;
  if (n_list > 0) {
    // the check, and for dense lists the packing of the snapshot, run on the
    // device while the host plans the nests
    RETCHK(ctx->active.ensure(32));
    int *dcnt = ctx->active.as<int>() + 4;
    HIPCHK(hipMemsetAsync(dcnt, 0, 2 * sizeof(int), ctx->stream));
    list_check_kernel<<<grid_for(n_list, ctx->cus), BLOCK, 0, ctx->stream>>>(list_index, n_list, snapshot->n,
                                                                              dcnt);
    HIPCHK(hipMemcpyAsync(ctx->h_active + 4, dcnt, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    RETCHK(pack_snapshot(ctx, snapshot, n_list));
    if (nsub > 0) {
      // (the lists' range is checked while the records are gathered)
      RETCHK(plan_nests(subs_host, nsub, t_nest, n_list));
      planned = true;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    cnt[0] = ctx->h_active[4];
    cnt[1] = ctx->h_active[5];
    if (cnt[0]) return fail(HBT_ERR_INVALID_ARG, std::to_string(cnt[0]) + " list entries lie past the snapshot");
  }
  // null entries: the lists are compacted first (KickNullParticles); the
  // records keep their input fields, the unbinding sees the compacted lists
  // and the adjusted Nbound, and output slots the shorter lists leave free
  // are set to -1 / NaN
  std::vector<hbt_subhalo> kicked;
  std::vector<int64_t> out_cap;
  if (cnt[1] > 0) {
    // the caller planned its outputs for the uncompacted lists
    out_cap.resize(nsub);
    for (int64_t s = 0; s < nsub; s++) out_cap[s] = t_nest.f[s].cap;
    planned = false;
    std::vector<KickIn> kin(nsub);
    for (int64_t s = 0; s < nsub; s++)
      kin[s] = KickIn{subs_host[s].offset, subs_host[s].count,
                      std::max<int64_t>(0, std::min(subs_host[s].nbound_prev, subs_host[s].count)), 0};
    RETCHK(ctx->kickin.ensure(nsub * sizeof(KickIn)));
    RETCHK(ctx->kickcnt.ensure(nsub * 2 * sizeof(int64_t)));
    HIPCHK(hipMemcpyAsync(ctx->kickin.p, kin.data(), nsub * sizeof(KickIn), hipMemcpyHostToDevice, ctx->stream));
    const int gk = (int)std::max<int64_t>(1, std::min<int64_t>(nsub, (int64_t)ctx->cus * 16));
    kick_count_kernel<<<gk, BLOCK, 0, ctx->stream>>>(list_index, ctx->kickin.as<KickIn>(), nsub,
                                                     ctx->kickcnt.as<int64_t>());
    std::vector<int64_t> keep(2 * nsub);
    HIPCHK(hipMemcpyAsync(keep.data(), ctx->kickcnt.p, keep.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    kicked.assign(subs_host, subs_host + nsub);
    int64_t dst = 0;
    for (int64_t s = 0; s < nsub; s++) {
      kin[s].dst = dst;
      kicked[s].offset = dst;
      kicked[s].count = keep[2 * s];
      kicked[s].nbound_prev = keep[2 * s + 1];
      dst += keep[2 * s];
    }
    RETCHK(ctx->kick.ensure((size_t)std::max<int64_t>(dst, 1) * 8));
    HIPCHK(hipMemcpyAsync(ctx->kickin.p, kin.data(), nsub * sizeof(KickIn), hipMemcpyHostToDevice, ctx->stream));
    kick_compact_kernel<<<gk, BLOCK, 0, ctx->stream>>>(list_index, ctx->kickin.as<KickIn>(), nsub,
                                                       ctx->kick.as<int64_t>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    list_index = ctx->kick.as<int64_t>();
    n_list = dst;
    subs_host = kicked.data();
  }
  ctx->list_index = list_index;
  return unbind_batch_impl(ctx, snapshot, n_list, subs, subs_host, nsub, out_index, out_energy, plain,
                           out_cap.empty() ? nullptr : out_cap.data(), planned);
}

// out_cap: the output capacity the caller planned per subhalo when it exceeds
// what the lists now need (lists compacted by KickNullParticles), or null.
static int unbind_batch_impl(hbt_ctx *ctx, const hbt_particles *particles, int64_t n_list,
                             hbt_subhalo *subs, const hbt_subhalo *subs_host, int64_t nsub,
                             int64_t *out_index, float *out_energy, int plain, const int64_t *out_cap,
                             bool planned) {
  if (nsub > 0 && (!subs || !subs_host || !out_index || !out_energy))
    return fail(HBT_ERR_INVALID_ARG, "null argument");
  if (nsub == 0) return HBT_OK;
  if (particles->n >= (int64_t)INT32_MAX) return fail(HBT_ERR_UNSUPPORTED, "more than 2^31 particles");
  HIPCHK(hipSetDevice(ctx->device));
  // the type check runs on the device while the host plans the nests
  const bool check_types = particles->type && particles->n > 0;
  if (check_types) {
    RETCHK(ctx->active.ensure(32));
    int *dbad = ctx->active.as<int>() + 6;
    HIPCHK(hipMemsetAsync(dbad, 0, sizeof(int), ctx->stream));
    type_check_kernel<<<grid_for(particles->n, ctx->cus), BLOCK, 0, ctx->stream>>>(particles->type, particles->n,
                                                                                  dbad);
    HIPCHK(hipMemcpyAsync(ctx->h_active + 6, dbad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  }
  Nest &N = t_nest;
  static thread_local Level L;
  if (!planned) RETCHK(plan_nests(subs_host, nsub, N));
  for (int64_t s = 0; s < nsub; s++) {
    if (N.f[s].offset < 0 || N.f[s].offset + N.f[s].count > n_list)
      return fail(HBT_ERR_INVALID_ARG, "subhalo list outside the particle columns");
    if (N.f[s].cap >= (int64_t)INT32_MAX / 2) return fail(HBT_ERR_UNSUPPORTED, "subhalo list too long");
  }
  hbt_particles P = *particles;
  int64_t passes = 0;
  // the type check's verdict is taken before the first level runs (the
  // record kernels index per-type sums by type), after the host has laid the
  // level out, so that the check and the packing overlap that work too
  bool types_checked = !check_types;
  auto types_ok = [&]() -> int {
    if (types_checked) return HBT_OK;
    types_checked = true;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int bad = ctx->h_active[6];
    if (bad) return fail(HBT_ERR_INVALID_ARG, std::to_string(bad) + " particle types outside [0, 6)");
    return HBT_OK;
  };
  const int64_t MS = ctx->params.max_sample_size > 0 ? ctx->params.max_sample_size : 0;
  if (MS > 0 && !particles->id) return fail(HBT_ERR_INVALID_ARG, "sampling needs the particle ID column");
  if (MS > 0) {
    if (ctx->rand_pos > ((int64_t)1 << 46)) return fail(HBT_ERR_UNSUPPORTED, "rand() stream position above 2^46");
    if (!ctx->randtab.p) {
      const RandJumps &J = rand_jumps();
      RETCHK(ctx->randtab.ensure(J.chunk.size() * 4));
      HIPCHK(hipMemcpyAsync(ctx->randtab.p, J.chunk.data(), J.chunk.size() * 4, hipMemcpyHostToDevice,
                            ctx->stream));
    }
  }
  // levels: nest heights; within a level the subhalos in index order (the
  // order in which sampled lists draw their shuffles from rand()); the
  // batched form runs a level at once, the plain form one subhalo at a time
  int hmax = 0;
  for (int64_t s = 0; s < nsub; s++) hmax = std::max(hmax, N.height[s]);
  static thread_local std::vector<std::vector<int>> by_h;
  by_h.resize(hmax + 1);
  for (auto &v : by_h) v.clear();
  for (int64_t s = 0; s < nsub; s++) by_h[N.height[s]].push_back((int)s);
  static thread_local std::vector<int64_t> draw_off;
  static thread_local std::vector<hbt_subhalo> recs;
  if (MS > 0) draw_off.assign(2 * nsub, -1);
  // The shuffles of a level: a list longer than MaxSampleSize takes one
  // rand() value per element after its first, in level order.  A nested
  // host's list holds its children's unbound tails, known once their level ran.
  auto plan_draws = [&](const std::vector<int> &members, Level &Lv) -> int {
    Lv.ndraws = 0;
    if (MS <= 0) return HBT_OK;
    bool nested = false;
    for (int s : members) nested = nested || N.nchild(s) > 0;
    if (nested) {
      recs.resize(nsub);
      HIPCHK(hipMemcpyAsync(recs.data(), subs, nsub * sizeof(hbt_subhalo), hipMemcpyDeviceToHost, ctx->stream));
      HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    for (int s : members) {
      int64_t len = N.f[s].count;
      if (N.f[s].nbp > 1)
        for (int i = 0; i < N.nchild(s); i++) len += recs[N.child(s)[i]].list_len - recs[N.child(s)[i]].nbound;
      draw_off[s] = -1;
      draw_off[nsub + s] = len;
      if (len > MS) {
        draw_off[s] = Lv.ndraws;
        Lv.ndraws += len - 1;
      }
    }
    if (Lv.ndraws >= (int64_t)RCHUNK << RJUMPS) return fail(HBT_ERR_UNSUPPORTED, "too many shuffle draws in one level");
    rand_window(ctx->rand_seed, ctx->rand_pos, Lv.win);
    ctx->rand_pos += Lv.ndraws;
    return HBT_OK;
  };
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  static const bool pass_times = getenv("HBT_PASS_TIMES") && atoi(getenv("HBT_PASS_TIMES"));
  for (int h = 0; h <= hmax; h++) {
    if (plain) {
      for (int s : by_h[h]) {
        const std::vector<int> one{s};
        RETCHK(plan_draws(one, L));
        const int64_t nd = L.ndraws;
        uint32_t win[31];
        memcpy(win, L.win, sizeof win);
        build_level(N, one, ctx->tune, MS > 0 ? draw_off.data() : nullptr, nsub, L);
        L.ndraws = nd;
        memcpy(L.win, win, sizeof win);
        RETCHK(types_ok());
        RETCHK(run_level(ctx, L, &P, subs, out_index, out_energy, &passes));
      }
      continue;
    }
    const auto t0 = std::chrono::steady_clock::now();
    RETCHK(plan_draws(by_h[h], L));
    const int64_t nd = L.ndraws;
    uint32_t win[31];
    memcpy(win, L.win, sizeof win);
    // build_level lays each level out largest first
    build_level(N, by_h[h], ctx->tune, MS > 0 ? draw_off.data() : nullptr, nsub, L);
    L.ndraws = nd;
    memcpy(L.win, win, sizeof win);
    if (pass_times)
      fprintf(stderr, "[hbt] level %d of %zu segments built on the host in %.3f ms\n", h, L.seg.size(),
              std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    RETCHK(types_ok());
    RETCHK(run_level(ctx, L, &P, subs, out_index, out_energy, &passes));
  }
  static thread_local std::vector<int64_t> gap;
  gap.clear();
  int64_t gap_total = 0;
  for (int64_t s = 0; s < nsub; s++) {
    const int64_t room = out_cap ? out_cap[s] : N.f[s].cap;
    if (N.nchild(s) > 0 || room > N.f[s].cap) {
      gap.push_back(s);
      gap.push_back(room);
      gap.push_back(gap_total);
      gap_total += room;
    }
  }
  if (!gap.empty()) {
    RETCHK(ctx->gaps.ensure(gap.size() * sizeof(int64_t)));
    HIPCHK(hipMemcpyAsync(ctx->gaps.p, gap.data(), gap.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                          ctx->stream));
    const int ngap = (int)(gap.size() / 3);
    fill_gaps_kernel<<<grid_for(gap_total, ctx->cus), BLOCK, 0, ctx->stream>>>(
        subs, ctx->gaps.as<int64_t>(), ngap, gap_total, out_index, out_energy);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  // evaluation count, summed on the device after the timed work
  long long *h_ev = (long long *)(ctx->h_active + 4);
  eval_sum_kernel<<<1, BLOCK, 0, ctx->stream>>>(subs, nsub, ctx->counters.as<unsigned long long>() + 4);
  HIPCHK(hipMemcpyAsync(h_ev, ctx->counters.as<unsigned long long>() + 4, 8, hipMemcpyDeviceToHost,
                        ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  ctx->last_passes = passes;
  ctx->last_evals = *h_ev;
  return HBT_OK;
}

int hbt_last_batch_stats(hbt_ctx *ctx, double *device_ms, int64_t *n_evals, int64_t *n_passes) {
  if (!ctx) return fail(HBT_ERR_INVALID_ARG, "null argument");
  if (device_ms) *device_ms = ctx->last_ms;
  if (n_evals) *n_evals = ctx->last_evals;
  if (n_passes) *n_passes = ctx->last_passes;
  return HBT_OK;
}

int hbt_profile(hbt_ctx *ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return fail(HBT_ERR_INVALID_ARG, "bad argument");
  HIPCHK(hipSetDevice(ctx->device));
  ctx->prof_mode = mode;
  ctx->walk_ms = 0;
  ctx->walk_launches = 0;
  ctx->sort_ms = 0;
  ctx->sort_calls = 0;
  RETCHK(ctx->counters.ensure(128));
  HIPCHK(hipMemsetAsync(ctx->counters.p, 0, 128, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return HBT_OK;
}

int hbt_profile_read(hbt_ctx *ctx, double *walk_ms, int64_t *walk_launches, int64_t *interactions,
                     int64_t *opened, int64_t *targets, int64_t *lane_slots, int64_t *tree_items,
                     int64_t *wave_steps) {
  if (!ctx) return fail(HBT_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipSetDevice(ctx->device));
  unsigned long long h[7] = {0, 0, 0, 0, 0, 0, 0};
  if (ctx->counters.p) {
    HIPCHK(hipMemcpyAsync(h, ctx->counters.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (walk_ms) *walk_ms = ctx->walk_ms;
  if (walk_launches) *walk_launches = ctx->walk_launches;
  if (interactions) *interactions = (int64_t)h[0];
  if (opened) *opened = (int64_t)h[1];
  if (targets) *targets = (int64_t)h[2];
  if (lane_slots) *lane_slots = (int64_t)h[3];
  if (tree_items) *tree_items = (int64_t)h[5];
  if (wave_steps) *wave_steps = (int64_t)h[6];
  return HBT_OK;
}

int hbt_profile_sorts(hbt_ctx *ctx, double *sort_ms, int64_t *sort_calls, int64_t *elements,
                      int64_t *element_levels) {
  if (!ctx) return fail(HBT_ERR_INVALID_ARG, "null argument");
  HIPCHK(hipSetDevice(ctx->device));
  unsigned long long h[2] = {0, 0};
  if (ctx->counters.p) {
    HIPCHK(hipMemcpyAsync(h, ctx->counters.as<unsigned long long>() + 8, sizeof(h), hipMemcpyDeviceToHost,
                          ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  if (sort_ms) *sort_ms = ctx->sort_ms;
  if (sort_calls) *sort_calls = ctx->sort_calls;
  if (elements) *elements = (int64_t)h[0];
  if (element_levels) *element_levels = (int64_t)h[1];
  return HBT_OK;
}

int hbt_unbind_one(hbt_ctx *ctx, const hbt_particles *particles, hbt_subhalo *sub,
                   int64_t *out_index, float *out_energy) {
  if (!ctx || !particles || !sub || !out_index || !out_energy)
    return fail(HBT_ERR_INVALID_ARG, "null argument");
  const int64_t n = sub->count;
  if (n < 0 || n > particles->n) return fail(HBT_ERR_INVALID_ARG, "bad count");
  HIPCHK(hipSetDevice(ctx->device));
  const size_t nn = (size_t)std::max<int64_t>(n, 1);
  RETCHK(ctx->st_cols.ensure(nn * (8 + 7 * 4 + 4)));
  RETCHK(ctx->st_subs.ensure(sizeof(hbt_subhalo)));
  RETCHK(ctx->st_out_i.ensure(nn * 8));
  RETCHK(ctx->st_out_e.ensure(nn * 4));
  char *base = ctx->st_cols.as<char>();
  hbt_particles D;
  D.n = n;
  int64_t *did = (int64_t *)base;
  float *f = (float *)(base + nn * 8);
  D.id = did;
  D.x = f; D.y = f + nn; D.z = f + 2 * nn; D.vx = f + 3 * nn; D.vy = f + 4 * nn; D.vz = f + 5 * nn;
  D.mass = f + 6 * nn;
  D.type = (int32_t *)(f + 7 * nn);
  hipStream_t st = ctx->stream;
  if (n > 0) {
    HIPCHK(hipMemcpyAsync((void *)D.id, particles->id, n * 8, hipMemcpyHostToDevice, st));
    const float *src[7] = {particles->x, particles->y, particles->z, particles->vx, particles->vy,
                           particles->vz, particles->mass};
    const float *dst[7] = {D.x, D.y, D.z, D.vx, D.vy, D.vz, D.mass};
    for (int i = 0; i < 7; i++)
      HIPCHK(hipMemcpyAsync((void *)dst[i], src[i], n * 4, hipMemcpyHostToDevice, st));
    if (particles->type)
      HIPCHK(hipMemcpyAsync((void *)D.type, particles->type, n * 4, hipMemcpyHostToDevice, st));
    else
      D.type = nullptr;
  }
  const hbt_subhalo keep = *sub;
  hbt_subhalo h = *sub;
  h.offset = 0;
  h.parent = -1;
  h.out_offset = 0;
  // Subhalo_t::Unbind has no orphan rule (that belongs to RecursiveUnbind)
  h.nbound_prev = INT64_MAX;
  HIPCHK(hipMemcpyAsync(ctx->st_subs.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
  RETCHK(hbt_unbind_batch(ctx, &D, ctx->st_subs.as<hbt_subhalo>(), &h, 1, ctx->st_out_i.as<int64_t>(),
                          ctx->st_out_e.as<float>(), 1));
  HIPCHK(hipMemcpyAsync(sub, ctx->st_subs.p, sizeof(hbt_subhalo), hipMemcpyDeviceToHost, st));
  if (n > 0) {
    HIPCHK(hipMemcpyAsync(out_index, ctx->st_out_i.p, n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(out_energy, ctx->st_out_e.p, n * 4, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  sub->offset = keep.offset;
  sub->parent = keep.parent;
  sub->out_offset = keep.out_offset;
  sub->nbound_prev = keep.nbound_prev;
  return HBT_OK;
}

}  // extern "C"

#include "hbt_tree_api.hpp"
#include "hbt_props.hpp"
#include "hbt_mask.hpp"
#include "hbt_merge.hpp"
#include "hbt_nest.hpp"
#include "hbt_direct.hpp"
#include "hbt_idhash.hpp"
