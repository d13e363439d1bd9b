// ref_tracking.cpp -- drives the reference's own SubhaloSnapshot_t::MaskSubhalos
// (src/subhalo_tracking.cpp:826-841) on a catalogue in the layout of
// hbt_mask_subhalos, so that the oracle restatement can be checked against
// it.  Linked with every reference source except the HBT.cpp driver
// (oracle/Makefile target `ref`); a singleton MPI process, since
// SubhaloSnapshot_t builds its MPI datatypes on construction.
//
//   ref_tracking mask in.bin out.bin
//     in : int64 nsub, int64 ncol; int64 id[ncol];
//          hbt_subhalo subs[nsub] (offset, count, parent, nbound_prev read);
//          int64 host_group[nsub]
//     out: int64 out_list[ncol] (column positions, laid out like the columns;
//          -1 where unused); int64 new_count[nsub]
//
// The member table is built by the reference (MemberShipTable_t::Build, which
// orders each group by Mbound): every subhalo gets an Mbound that decreases
// with its index and the group's first head the largest, so the group's
// central and the order of its heads are those of hbt_mask_subhalos.
#include <mpi.h>

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

namespace {

template <class T>
void rd(FILE *f, T *p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}
template <class T>
void wr(FILE *f, const T *p, size_t n) {
  if (n) fwrite(p, sizeof(T), n, f);
}

int run_mask(FILE *in, FILE *out) {
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
    S.HostHaloId = grp[s];
    S.Mbound = (float)(nsub - s);
    if (rec[s].parent >= 0) snap.Subhalos[rec[s].parent].NestedSubhalos.push_back(s);
    else if (grp[s] >= 0) heads[grp[s]].push_back(s);
  }
  for (int64_t g = 0; g < nhalos; g++)
    if (!heads[g].empty()) snap.Subhalos[heads[g][0]].Mbound = (float)(2 * nsub + 1);
  snap.MemberTable.Build(nhalos, snap.Subhalos, true);
  snap.MemberTable.SubGroupsOfHeads.assign(heads.begin(), heads.end());
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

}  // namespace

int main(int argc, char **argv) {
  if (argc != 4) { fprintf(stderr, "usage: %s mask in.bin out.bin\n", argv[0]); return 1; }
  MPI_Init(&argc, &argv);
  FILE *in = fopen(argv[2], "rb");
  FILE *out = fopen(argv[3], "wb");
  int rc = 1;
  if (!in || !out) fprintf(stderr, "cannot open files\n");
  else if (!strcmp(argv[1], "mask")) rc = run_mask(in, out);
  if (in) fclose(in);
  if (out) fclose(out);
  MPI_Finalize();
  return rc;
}
