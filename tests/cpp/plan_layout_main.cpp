// Host-side checks of the unbinding planner, run without a GPU (no HIP calls):
//  - the capacity order of a level (largest first, ties in member order)
//    against std::stable_sort, across the counting pass's clamp and its
//    four-thread split;
//  - the level layout build_level writes (wave-aligned slot bases, tile table,
//    segment records, nest children) against its definition, on planned nests;
//  - the records' checks while the planner gathers them (lists inside the
//    list positions, counts and parents), from every gather thread's range;
//  - the rand() stream the sampled shuffles replay (the host's window for a
//    call number and the device's chunk jumps) against glibc's own rand();
//  - the staging copy of a level's tables (pack_blobs) against plain memcpy,
//    padding untouched, across its four-range split.
// Prints "ok" and exits 0, or names the first failed check and exits 1.
#include <cstdio>
#include <random>
#include <string>

#include "../../hbtplus_amd/csrc/hbt_unbind.hip"

#define CHECK(cond, ...)                 \
  do {                                   \
    if (!(cond)) {                       \
      fprintf(stderr, "FAIL: ");         \
      fprintf(stderr, __VA_ARGS__);      \
      fprintf(stderr, "\n");             \
      return 1;                          \
    }                                    \
  } while (0)

static std::vector<int> stable_order(const Nest &N, const std::vector<int> &members) {
  std::vector<int> r = members;
  std::stable_sort(r.begin(), r.end(), [&](int a, int b) { return N.f[a].cap > N.f[b].cap; });
  return r;
}

// capacities drawn from a mix: short lists, the counting pass's clamp and its
// neighbours, long lists, or one value for all
static int64_t draw_cap(std::mt19937 &g, int mode) {
  switch (mode) {
    case 0: return g() % 50;
    case 1: return 4093 + g() % 6;
    case 2: return g() % 4 ? (int64_t)(g() % 300) : (int64_t)(g() % 10000000);
    default: return 777;
  }
}

static int check_sort() {
  std::mt19937 g(11);
  const int sizes[] = {0, 1, 2, 5, 63, 64, 3000, 65535, 65536, 70001, 150000};
  std::vector<int> out, caps;
  for (int n : sizes)
    for (int mode = 0; mode < 4; mode++) {
      Nest M;
      M.f.resize(n + 10);
      for (auto &f : M.f) f.cap = draw_cap(g, mode);
      // members: a shuffled subset (a nested level), or all in order (the top level)
      std::vector<int> mem;
      if (mode & 1) {
        for (int i = 0; i < n; i++) mem.push_back(i);
      } else {
        for (int i = 0; i < n + 10 && (int)mem.size() < n; i++)
          if (g() % 8) mem.push_back(i);
        std::shuffle(mem.begin(), mem.end(), g);
      }
      sort_largest_first(M, mem, out, caps);
      CHECK(out == stable_order(M, mem), "capacity order, n=%d mode=%d", n, mode);
      CHECK(caps.size() == out.size(), "caps size, n=%d", n);
      for (size_t i = 0; i < out.size(); i++)
        CHECK(caps[i] == (int)M.f[out[i]].cap, "caps[%zu], n=%d mode=%d", i, n, mode);
    }
  return 0;
}

// random forests of nests (children listed before and after their hosts)
static std::vector<hbt_subhalo> forest(std::mt19937 &g, int n, int64_t &total) {
  std::vector<hbt_subhalo> s(n);
  std::memset(s.data(), 0, n * sizeof(hbt_subhalo));
  total = 0;
  for (int i = 0; i < n; i++) {
    s[i].count = g() % 16 == 0 ? 2000 + g() % 30000 : g() % 400;
    s[i].offset = total;
    s[i].nbound_prev = s[i].count / 2;
    s[i].parent = -1;
    total += s[i].count;
  }
  // a third attach under a random subhalo that is not their own descendant
  // and sits less than five deep
  for (int i = 0; i < n; i++) {
    if (g() % 3) continue;
    const int p = (int)(g() % n);
    int q = p, d = 0;
    while (q >= 0 && q != i) q = (int)s[q].parent, d++;
    if (q != i && d < 5) s[i].parent = p;
  }
  int64_t o = 0;
  for (int i = 0; i < n; i++) {
    s[i].out_offset = o;
    o += s[i].count + 7;
  }
  return s;
}

static int check_layout() {
  std::mt19937 g(12);
  const hbt_tuning tune = {{8, 16, 32}, {200, 3000}, 2, 0, 16};
  for (int n : {1, 7, 900, 70000}) {
    int64_t total = 0;
    std::vector<hbt_subhalo> s = forest(g, n, total);
    static Nest N;
    CHECK(plan_nests(s.data(), n, N) == HBT_OK, "plan_nests, n=%d: %s", n, hbt_last_error());
    // a level: every subhalo (children and all), and the heads alone
    for (int which = 0; which < 2; which++) {
      std::vector<int> members;
      for (int i = 0; i < n; i++)
        if (which == 0 || s[i].parent < 0) members.push_back(i);
      static Level L;
      build_level(N, members, tune, nullptr, n, L);
      const int m = (int)members.size();
      CHECK(L.subs == stable_order(N, members), "level order, n=%d", n);
      int64_t off = 0, t0 = 0;
      int maxcap = 0;
      for (int k = 0; k < m; k++) {
        const int sub = L.subs[k];
        const SubPlan &F = N.f[sub];
        const int c = (int)F.cap;
        const int ca = std::max(WAVE, (c + WAVE - 1) / WAVE * WAVE);
        const int nt = (c + TILE - 1) / TILE;
        CHECK(L.ibase[k] == off && off % WAVE == 0, "ibase[%d], n=%d", k, n);
        CHECK(L.capal[k] == ca && L.first_tile[k] == t0 && L.ntile[k] == nt, "tiles of %d, n=%d", k, n);
        for (int t = 0; t < nt; t++)
          CHECK(L.tiles[t0 + t].seg == k && L.tiles[t0 + t].start == t * TILE, "tile %d of %d", t, k);
        const SegIn &S = L.seg[k];
        CHECK(S.ibase == off && S.out_off == F.out_off && S.list_off == F.offset && S.nbprev == F.nbp &&
                  S.sub == sub && S.cap == c && S.count == (int)F.count && S.wgroup == walk_group_for(tune, c) &&
                  S.draw_off == -1 && S.shuf_len == 0,
              "segment record %d, n=%d", k, n);
        const int nc = N.nchild(sub);
        CHECK(L.child_ptr[k + 1] - L.child_ptr[k] == nc, "children of %d", k);
        for (int i = 0; i < nc; i++) CHECK(L.child_list[L.child_ptr[k] + i] == N.child(sub)[i], "child %d of %d", i, k);
        off += ca;
        t0 += nt;
        maxcap = std::max(maxcap, c);
      }
      CHECK(L.nslots == off && (int64_t)L.tiles.size() == t0 && L.maxcap == maxcap, "level totals, n=%d", n);
      // a subhalo's capacity is its own list plus its nested subhalos' capacities
      for (int i = 0; i < n; i++) {
        int64_t c = s[i].count;
        for (int j = 0; j < N.nchild(i); j++) c += N.f[N.child(i)[j]].cap;
        CHECK(N.f[i].cap == c, "capacity of %d, n=%d", i, n);
      }
    }
  }
  return 0;
}

// the records' checks: a list outside [0, n_list) or a bad count or parent is
// an error naming which, wherever the record sits (each gather thread's range)
static int check_record_errors() {
  std::mt19937 g(14);
  for (int n : {5, 70000}) {
    int64_t total = 0;
    const std::vector<hbt_subhalo> base = forest(g, n, total);
    static Nest N;
    CHECK(plan_nests(base.data(), n, N, total) == HBT_OK, "valid records rejected, n=%d: %s", n, hbt_last_error());
    CHECK(plan_nests(base.data(), n, N) == HBT_OK, "valid records rejected without n_list, n=%d", n);
    const std::string outside = "subhalo list outside the list positions", badrec = "bad subhalo record";
    for (int at : {0, n / 2, n - 1}) {
      struct Case {
        const char *what;
        void (*edit)(hbt_subhalo &, int64_t, int);
        bool with_n_list;
        const std::string *msg;
      };
      const Case cases[] = {
          {"list past the end", [](hbt_subhalo &r, int64_t t, int) { r.offset = t - r.count + 1; }, true, &outside},
          {"negative offset", [](hbt_subhalo &r, int64_t, int) { r.offset = -1; }, true, &outside},
          {"negative count", [](hbt_subhalo &r, int64_t, int) { r.count = -3; }, true, &outside},
          {"negative count, no n_list", [](hbt_subhalo &r, int64_t, int) { r.count = -3; }, false, &badrec},
          {"parent past the end", [](hbt_subhalo &r, int64_t, int m) { r.parent = m; }, true, &badrec},
          {"both", [](hbt_subhalo &r, int64_t t, int m) { r.parent = m; r.offset = t; r.count = 1; }, true, &outside},
      };
      for (const Case &c : cases) {
        std::vector<hbt_subhalo> s = base;
        c.edit(s[at], total, n);
        const int rc = c.with_n_list ? plan_nests(s.data(), n, N, total) : plan_nests(s.data(), n, N);
        CHECK(rc == HBT_ERR_INVALID_ARG && std::string(hbt_last_error()).rfind(*c.msg, 0) == 0,
              "%s at %d of %d: rc %d, '%s'", c.what, at, n, rc, hbt_last_error());
      }
    }
    // a list that ends exactly at n_list is inside
    std::vector<hbt_subhalo> s = base;
    s[n - 1].offset = total - s[n - 1].count;
    CHECK(plan_nests(s.data(), n, N, total) == HBT_OK, "list ending at n_list rejected, n=%d", n);
  }
  return 0;
}

// The rand() stream the sampled shuffles replay: the window the host computes
// for call number pos (jumping by the matrices M^(2^k)), stepped by the
// recurrence, gives glibc's own rand() values from that call on; and the
// device's chunk jumps M^(RCHUNK 2^k) move a window exactly RCHUNK 2^k calls.
static int check_rand_stream() {
  for (uint32_t seed : {1u, 0u, 12345u})
    for (int64_t pos : {(int64_t)0, (int64_t)1, (int64_t)30, (int64_t)31, (int64_t)344, (int64_t)1000,
                        (int64_t)123457, (int64_t)RCHUNK * 3 + 5, (int64_t)5000011}) {
      srand(seed);
      for (int64_t i = 0; i < pos; i++) (void)rand();
      uint32_t w[RWIN];
      rand_window(seed, pos, w);
      for (int t = 0; t < 200; t++) {
        const uint32_t next = w[0] + w[28];
        const int want = rand();
        CHECK((int)(next >> 1) == want, "rand() call %lld of seed %u: %d, want %d", (long long)(pos + t), seed,
              (int)(next >> 1), want);
        memmove(w, w + 1, (RWIN - 1) * sizeof(uint32_t));
        w[RWIN - 1] = next;
      }
    }
  const RandJumps &J = rand_jumps();
  for (int k = 0; k < 8; k++) {
    const int64_t pos = 777, step = (int64_t)RCHUNK << k;
    uint32_t a[RWIN], b[RWIN];
    rand_window(1, pos, a);
    rmat_apply(&J.chunk[(size_t)k * RWIN * RWIN], a, a);
    rand_window(1, pos + step, b);
    CHECK(memcmp(a, b, sizeof a) == 0, "chunk jump %d (%lld calls)", k, (long long)step);
  }
  return 0;
}

static int check_staging() {
  std::mt19937 g(13);
  const size_t pick[] = {0, 1, 7, 15, 16, 17, 255, 256, 1000, 65536 * 4 + 3, 1048573, 3000001};
  for (int rep = 0; rep < 40; rep++) {
    const int nb = 1 + g() % 9;
    std::vector<std::vector<unsigned char>> src(nb);
    for (auto &v : src) {
      v.resize(pick[g() % (sizeof pick / sizeof pick[0])]);
      for (auto &x : v) x = (unsigned char)g();
    }
    std::vector<size_t> at;
    size_t off = 0;
    for (auto &v : src) {
      at.push_back(off);
      off += (std::max<size_t>(v.size(), 1) + 255) & ~(size_t)255;
    }
    std::vector<unsigned char> h(off + 256, 0xab);
    const auto B = [&](int i) { return Blob{src[i].empty() ? nullptr : src[i].data(), src[i].size()}; };
    // the entry points pass fixed lists of up to nine blobs
    const Blob z{nullptr, 0};
    std::vector<Blob> b(9, z);
    for (int i = 0; i < nb; i++) b[i] = B(i);
    std::vector<size_t> at9 = at;
    for (int i = nb; i < 9; i++) at9.push_back(off);
    pack_blobs((char *)h.data(), at9, {b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8]});
    size_t p = 0;
    for (int i = 0; i < nb; i++) {
      for (; p < at[i]; p++) CHECK(h[p] == 0xab, "padding byte %zu overwritten (rep %d)", p, rep);
      CHECK(src[i].empty() || std::memcmp(h.data() + at[i], src[i].data(), src[i].size()) == 0,
            "blob %d of %zu bytes (rep %d)", i, src[i].size(), rep);
      p = at[i] + src[i].size();
    }
    for (; p < h.size(); p++) CHECK(h[p] == 0xab, "trailing byte %zu overwritten (rep %d)", p, rep);
  }
  return 0;
}

int main() {
  if (check_sort() || check_layout() || check_record_errors() || check_rand_stream() || check_staging())
    return 1;
  printf("ok\n");
  return 0;
}
