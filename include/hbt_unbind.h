/*
 * hbt_unbind.h -- C-ABI of the MI355X (gfx950) subhalo-unbinding library
 * (libhbt_unbind.so).
 *
 * This is the drop-in boundary for HBT+'s subhalo unbinding hot path.  Every
 * entry point below replaces one reference interface (cited file:line, paths
 * relative to the HBT+ source tree):
 *
 *   hbt_unbind_one      Subhalo_t::Unbind                src/subhalo.h:111,
 *                                                        src/subhalo_unbind.cpp:263-431
 *   hbt_unbind_batch    SubhaloSnapshot_t::RefineParticles loop body
 *                       (RecursiveUnbind + Unbind + TruncateSource)
 *                                                        src/subhalo.h:112,114,245,
 *                                                        src/subhalo_unbind.cpp:432-516
 *   hbt_tree_potential  GravityTree_t::Build + EvaluatePotential
 *                                                        src/oct_tree.h:62-63,
 *                                                        src/gravity_tree.h:15
 *   hbt_direct_potential  exact pair sum with the same softened kernel as
 *                       GravityTree_t::EvaluatePotential (src/gravity_tree.cpp:144-160);
 *                       the FP64 check of the tree potentials and the precision study
 *                       (the unbinding itself always walks the tree, as the reference does)
 *   hbt_id_to_index     ParticleSnapshot_t::GetIndices / MappedIndexTable_t::GetIndices
 *                                                        src/hash_remote.tpp:58-109,
 *                                                        src/hash.tpp:18-56
 *   hbt_unbind_batch_indexed  the RefineParticles loop body on ID-resolved lists
 *   hbt_subhalo_properties    CalculateProfileProperties + CalculateShape
 *                                                        src/subhalo.cpp:242-398
 *   hbt_nest_subhalos   NestSubhalos / LevelUpDetachedSubhalos
 *                                                        src/subhalo_tracking.cpp:669-688,735-786
 *   hbt_mask_subhalos   MaskSubhalos / SubhaloMasker_t   src/subhalo_tracking.cpp:793-841
 *   hbt_merge_subhalos  MergeSubhalos (DetectTraps, MergeRecursive, MergeTo)
 *                                                        src/subhalo_merge.cpp:136-262
 *
 * Conventions
 *   - Plain pointers and sizes only.  Every function returns an int status:
 *     HBT_OK (0) on success, a negative HBT_ERR_* code otherwise.  Nothing
 *     throws or aborts across this boundary; hbt_last_error() gives a message.
 *   - Particle arrays are structure-of-arrays (x[], y[], z[], ...), float32
 *     positions/velocities/masses, int64 IDs, int32 types -- the fields of the
 *     reference Particle_t (src/snapshot.h:51-72), split into columns.
 *   - Pointers passed to hbt_unbind_batch / hbt_tree_potential /
 *     hbt_direct_potential / hbt_id_to_index are DEVICE pointers (HBM); the
 *     work is enqueued on the context's stream and the call returns after the
 *     stream has finished, unless stated otherwise.
 *   - hbt_unbind_one takes HOST pointers (it is the per-subhalo form a
 *     reference caller uses in place of Subhalo_t::Unbind) and copies in/out.
 *   - One context = one HIP stream + device work buffers.  Contexts are
 *     independent, so distinct host threads may drive distinct contexts.
 */
#ifndef HBT_UNBIND_H
#define HBT_UNBIND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HBT_OK                  0
#define HBT_ERR_INVALID_ARG    -1
#define HBT_ERR_OUT_OF_MEMORY  -2
#define HBT_ERR_DEVICE         -3   /* a HIP runtime call failed */
#define HBT_ERR_UNSUPPORTED    -4   /* a configuration the library does not implement */
#define HBT_ERR_CAPACITY       -5   /* an output buffer is too small */

/* ---- precision policies -------------------------------------------------- */
/* FAST: pair/tree arithmetic in FP32, each target's potential terms summed
 *       in FP64; frame averages, kinetic energy and the E<0 cut in FP64.
 * REPLAY: the reference's own arithmetic (FP64 walk over float node data,
 *       float-rounded node moments, serial-order FP64 frame sums, no FMA
 *       contraction) for bit-level certification. */
#define HBT_PREC_FAST    0
#define HBT_PREC_REPLAY  1
#define HBT_PREC_FP32    2   /* FAST with the potential terms summed in FP32 (precision study) */

/* direct-sum accumulation modes (hbt_direct_potential) */
#define HBT_DIRECT_FP32        0   /* plain FP32 sum */
#define HBT_DIRECT_FP32_KAHAN  1   /* FP32 terms, Kahan-compensated sum */
#define HBT_DIRECT_FP64        2   /* FP64 throughout */

/* status bits reported in hbt_subhalo.status */
#define HBT_STATUS_CONVERGED   1   /* left the loop through the BoundMassPrecision test */
#define HBT_STATUS_DISRUPTED   2   /* Nbound fell below MinNumPartOfSub */
#define HBT_STATUS_EARLY       4   /* 0 or 1 source particles: no iteration */

#define HBT_TYPE_MAX 6              /* ParticleType_t::TypeMax, src/datatypes.h:179-188 */

/* Unbinding parameters: the HBTConfig fields Unbind reads
 * (src/config_parser.h:22-110, derived values src/config_parser.cpp:95-103),
 * PhysicalConst::G and the epoch's Cosmology_t (src/snapshot.h:17-39). */
typedef struct hbt_params {
  float   G;                        /* PhysicalConst::G, 43007.1 for kpc/h, km/s, 1e10 Msun/h */
  float   softening;                /* SofteningHalo */
  float   open_angle;               /* TreeNodeOpenAngle (theta) */
  float   box_size;                 /* BoxSize */
  int32_t periodic;                 /* PeriodicBoundaryOn */
  float   bound_mass_precision;     /* BoundMassPrecision */
  int32_t min_num_part_of_sub;      /* MinNumPartOfSub */
  float   source_sub_relax_factor;  /* SourceSubRelaxFactor */
  int32_t max_sample_size;          /* MaxSampleSizeOfPotentialEstimate (0: no sampling) */
  float   scale_factor;             /* Cosmology.ScaleFactor */
  float   hz;                       /* Cosmology.Hz */
  int32_t snapshot_index;           /* epoch.GetSnapshotIndex() */
  int32_t precision;                /* HBT_PREC_FAST or HBT_PREC_REPLAY */
  int32_t refine_mostbound_particle;/* RefineMostboundParticle (read when max_sample_size > 0) */
  int32_t reserved[2];
} hbt_params;

/* Particle columns (device or host pointers, see each entry point). */
typedef struct hbt_particles {
  int64_t        n;
  const int64_t *id;                /* Particle_t::Id */
  const float   *x, *y, *z;         /* ComovingPosition */
  const float   *vx, *vy, *vz;      /* PhysicalVelocity */
  const float   *mass;              /* Mass */
  const int32_t *type;              /* Type (0..HBT_TYPE_MAX-1) */
} hbt_particles;

/* One subhalo record: the Subhalo_t fields Unbind reads and writes
 * (src/subhalo.h:24-110), plus the layout of its particle list. */
typedef struct hbt_subhalo {
  /* -- inputs -- */
  int64_t offset;                   /* first particle of the source list in the particle columns */
  int64_t count;                    /* source list length (Particles.size()) */
  int64_t parent;                   /* subhalo whose NestedSubhalos holds this one, -1 if none */
  int64_t nbound_prev;              /* Nbound before this pass: RecursiveUnbind's orphan test
                                       (Nbound <= 1); pass a value >= 2 for plain Unbind semantics */
  int64_t sink_track_id;            /* SinkTrackId (in/out) */
  int32_t snapshot_index_of_death;  /* SnapshotIndexOfDeath (in/out) */
  int32_t snapshot_index_of_sink;   /* SnapshotIndexOfSink (in/out) */
  float   ref_pos[3];               /* ComovingAveragePosition: initial frame in, final frame out */
  float   ref_vel[3];               /* PhysicalAverageVelocity */
  float   mostbound_pos[3];         /* ComovingMostBoundPosition (in: previous, out: new) */
  float   mostbound_vel[3];         /* PhysicalMostBoundVelocity */
  /* -- outputs -- */
  int64_t nbound;                   /* Nbound */
  int64_t list_len;                 /* Particles.size() after unbinding (own list + appended tails) */
  int64_t nsource;                  /* Particles.size() after TruncateSource */
  int64_t nbound_type[HBT_TYPE_MAX];/* NboundType */
  float   mbound_type[HBT_TYPE_MAX];/* MboundType */
  float   mbound;                   /* Mbound */
  float   specific_self_potential_energy;
  float   specific_self_kinetic_energy;
  float   specific_angular_momentum[3];
  int32_t n_iter;                   /* unbinding passes run */
  int32_t status;                   /* HBT_STATUS_* bits */
  int64_t n_evals;                  /* binding-energy evaluations (full pass: Nlast, correction: Nbound) */
  int64_t out_offset;               /* start of this subhalo's output list in out_index/out_energy */
} hbt_subhalo;

typedef struct hbt_ctx hbt_ctx;

/* Library and context management. */
const char *hbt_last_error(void);
int hbt_version(int *major, int *minor);
int hbt_ctx_create(const hbt_params *params, int device, hbt_ctx **out);
int hbt_ctx_destroy(hbt_ctx *ctx);
int hbt_ctx_set_params(hbt_ctx *ctx, const hbt_params *params);
/* The HIP stream (hipStream_t) the context enqueues on. */
int hbt_ctx_stream(hbt_ctx *ctx, void **stream);

/* The rand() stream the sampling shuffles draw from (MaxSampleSizeOfPotentialEstimate
 * > 0: std::random_shuffle at src/subhalo_unbind.cpp:302 on glibc's rand()):
 * the stream of srand(seed) after `draws` calls.  A new context holds seed 1,
 * no calls (a process that never calls srand); every unbinding call advances
 * it by the values its shuffles take.  Within a call the subhalos draw level
 * by level (a level: the subhalos whose nested subhalos are all done), in
 * index order within a level -- the order of the reference's single-threaded
 * RefineParticles loop for a batch without nesting.  No counterpart in the
 * reference's interface (its stream is the process's global rand() state). */
int hbt_set_rand_state(hbt_ctx *ctx, uint32_t seed, int64_t draws);
int hbt_get_rand_state(hbt_ctx *ctx, uint32_t *seed, int64_t *draws);

/* Output capacity needed by hbt_unbind_batch for a given nest layout:
 * for every subhalo, its own count plus the counts of everything nested
 * below it (unbound tails are appended to the parent, src/subhalo_unbind.cpp:442).
 * subs: HOST array of nsub records with offset/count/parent filled.
 * Writes subs[i].out_offset and returns the total in *total. */
int hbt_plan_batch(hbt_subhalo *subs, int64_t nsub, int64_t *total);

/* Batched unbinding of nsub subhalos (RefineParticles loop body).
 *   particles : device columns holding every source list
 *   subs      : DEVICE array of nsub records (inputs filled, out_offset set by
 *               hbt_plan_batch); outputs are written in place
 *   out_index : DEVICE int64[total]; for each subhalo the final particle list
 *               as indices into the particle columns: bound particles sorted by
 *               binding energy, then the removed blocks (last removed first),
 *               then appended tails, exactly the reference's Particles order;
 *               a host's region keeps room for every nested particle, and the
 *               slots past its final list_len are set to -1
 *   out_energy: DEVICE float[total]; the binding energy at each list position
 *               (the reference's Elist E, SAVE_BINDING_ENERGY), NaN where none
 *               and in the unused slots
 *   subs_host : HOST copy of subs (same layout) used for planning; not modified
 *   plain     : 0 = batched form; 1 = plain per-subhalo form (one subhalo at a
 *               time through the same kernels, the loop RefineParticles runs) */
int hbt_unbind_batch(hbt_ctx *ctx, const hbt_particles *particles,
                     hbt_subhalo *subs, const hbt_subhalo *subs_host, int64_t nsub,
                     int64_t *out_index, float *out_energy, int plain);

/* Batched unbinding of subhalos whose source lists are positions in a
 * snapshot's particle columns -- what RefineParticles sees after the source
 * IDs were resolved (ParticleExchanger_t::QueryParticles,
 * src/particle_exchanger.h:193-208; on the device: hbt_id_to_index).
 *   snapshot  : device particle columns of the snapshot
 *   list_index: DEVICE int64[n_list]; subs[].offset/count are ranges of it and
 *               its entries are snapshot positions.  A negative entry is a null
 *               particle (an ID the snapshot does not hold, -1 from
 *               hbt_id_to_index): before unbinding, every list drops its null
 *               entries keeping the others in order, and the Nbound the orphan
 *               test reads becomes the number of kept entries among the first
 *               nbound_prev (Subhalo_t::KickNullParticles, src/subhalo.cpp:446-477,
 *               after the particle query, src/particle_exchanger.h:239-240).  The
 *               records' input fields are not changed; the output slots that
 *               the shorter lists leave free are set to -1 and NaN.  Positions
 *               at or past snapshot->n are rejected with HBT_ERR_INVALID_ARG
 *   out_index : snapshot positions, layout as in hbt_unbind_batch
 * Other arguments as in hbt_unbind_batch.  When n_list is at least a quarter
 * of snapshot->n the context first interleaves the snapshot's columns into
 * 32-byte records (a device buffer of 32 * snapshot->n bytes that the context
 * keeps for later calls), so the gather in list order reads one record per
 * particle. */
int hbt_unbind_batch_indexed(hbt_ctx *ctx, const hbt_particles *snapshot,
                             const int64_t *list_index, int64_t n_list, hbt_subhalo *subs,
                             const hbt_subhalo *subs_host, int64_t nsub, int64_t *out_index,
                             float *out_energy, int plain);

/* Per-subhalo form with HOST buffers (Subhalo_t::Unbind).  `particles` are
 * host columns of this subhalo's list (n = sub->count, sub->offset ignored);
 * `sub` is a host record updated in place; out_index/out_energy are host
 * arrays of length sub->count. */
int hbt_unbind_one(hbt_ctx *ctx, const hbt_particles *particles, hbt_subhalo *sub,
                   int64_t *out_index, float *out_energy);

/* Tree potential (GravityTree_t::Build over the n_src sources in the given
 * order, then EvaluatePotential at each target).  Returns the specific
 * physical potential G*phi/a.  target_mass[i] != 0 removes the target's own
 * softened self term (pass the target's mass when it is one of the sources,
 * src/gravity_tree.cpp:79-98); target_mass may be NULL (all 0).
 * Device pointers; out_phi is double[n_tgt]. */
int hbt_tree_potential(hbt_ctx *ctx, int64_t n_src, const float *sx, const float *sy,
                       const float *sz, const float *smass, int64_t n_tgt,
                       const float *tx, const float *ty, const float *tz,
                       const float *target_mass, double *out_phi);

/* Direct-sum potential of each of n particles due to all others (self term
 * excluded), softened with the same spline kernel.  Segmented: nseg groups
 * given by seg_offset[nseg+1] (device int64); mode = HBT_DIRECT_*.
 * out_phi is double[n] (device). */
int hbt_direct_potential(hbt_ctx *ctx, int64_t nseg, const int64_t *seg_offset,
                         const float *x, const float *y, const float *z,
                         const float *mass, int mode, double *out_phi);

/* hbt_direct_potential without allocating anything (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.  Same
 * arguments and results as hbt_direct_potential; at most 65536 groups (larger
 * nseg: HBT_ERR_CAPACITY).  Its tile table lives in a file-scope device array
 * of the library (about 512 KiB, held by each GPU the library runs on until
 * the process ends) that every call on a GPU shares: calls that overlap on one
 * GPU race.  Offsets that are not ascending give HBT_ERR_INVALID_ARG after the
 * call (such groups get no targets). */
int hbt_direct_potential_noalloc(hbt_ctx *ctx, int64_t nseg, const int64_t *seg_offset,
                                 const float *x, const float *y, const float *z,
                                 const float *mass, int mode, double *out_phi);

/* Post-unbind properties of one subhalo: the Subhalo_t fields set by
 * CalculateProfileProperties and CalculateShape (src/subhalo.h:48-77,
 * src/subhalo.cpp:242-398).  Fields marked in/out keep their input value
 * where the reference leaves them unchanged. */
typedef struct hbt_properties {
  float   rmax_comoving;            /* RmaxComoving */
  float   vmax_physical;            /* VmaxPhysical */
  float   rhalf_comoving;           /* RHalfComoving */
  float   r2sigma_comoving;         /* R2SigmaComoving (95.5% containment) */
  float   bound_r200crit_comoving;  /* BoundR200CritComoving (in/out: kept when no radius qualifies) */
  float   bound_m200crit;           /* BoundM200Crit (in/out, as above) */
  float   last_max_vmax_physical;   /* LastMaxVmaxPhysical (in/out) */
  int32_t snapshot_index_of_last_max_vmax;  /* SnapshotIndexOfLastMaxVmax (in/out) */
  float   inertial_tensor[6];       /* InertialTensor {xx, xy, xz, yy, yz, zz} / Mbound */
  float   inertial_tensor_weighted[6];  /* InertialTensorWeighted */
} hbt_properties;

/* CalculateProfileProperties + CalculateShape (src/subhalo.cpp:242-398) for
 * nsub subhalos after hbt_unbind_batch: subhalo i's bound particles are
 * out_index[subs[i].out_offset .. + subs[i].nbound) (energy order, most bound
 * first), indices into the particle columns; the centre is mostbound_pos.
 * particles, subs, out_index, props: device pointers; props is in/out. */
int hbt_subhalo_properties(hbt_ctx *ctx, const hbt_particles *particles, const hbt_subhalo *subs,
                           int64_t nsub, const int64_t *out_index, hbt_properties *props);

/* Exclusive particle ownership inside each host halo (SubhaloMasker_t and
 * SubhaloSnapshot_t::MaskSubhalos, src/subhalo_tracking.cpp:793-841).
 * A host group's subhalos are visited from its central -- the first of the
 * group's heads (parent < 0) -- in post-order: a subhalo's nested subhalos
 * before it, the group's other heads counting as nested in the central after
 * its own nested subhalos.  Siblings are taken in ascending visit_rank (HOST
 * int64[nsub], ties and NULL: index order): the reference's order is a group's
 * members as MemberTable.SubGroups[g] holds them (most massive first, the
 * central at [0]; SubGroupsOfHeads keeps that order) and a subhalo's nested
 * subhalos in NestedSubhalos order (src/subhalo_tracking.cpp:96-104,676-686),
 * so visit_rank = a head's place in SubGroups[g] and a nested subhalo's place
 * in its parent's NestedSubhalos.  A visited subhalo with
 * nbound_prev > 1 keeps the list entries whose particle ID no earlier entry of
 * the group has kept (list order); orphans (nbound_prev <= 1) and field
 * subhalos (host_group < 0) keep their lists and claim nothing.
 *   particles  : device columns (id read)
 *   subs_host  : HOST records; offset/count/parent/nbound_prev read
 *   host_group : HOST int64[nsub]; a nested subhalo must share its parent's group
 *   out_list   : DEVICE int64[sum of counts]; subhalo i's retained entries, as
 *                particle column positions, at [offset_i, offset_i + new_count[i])
 *                (so out_list is laid out like the input columns)
 *   new_count  : HOST int64[nsub], output */
int hbt_mask_subhalos(hbt_ctx *ctx, const hbt_particles *particles, const hbt_subhalo *subs_host,
                      const int64_t *host_group, const int64_t *visit_rank, int64_t nsub,
                      int64_t *out_list, int64_t *new_count);

/* Nest bookkeeping before masking (SubhaloSnapshot_t::NestSubhalos ->
 * LevelUpDetachedSubhalos, Subhalo_t::LevelUpDetachedMembers,
 * src/subhalo_tracking.cpp:669-688, 735-786), for one rank's subhalos with
 * local nest indices.  Every group's central and every field subhalo gets
 * Rank 0; then, in the nests below the subhalos that were heads (Rank 0) on
 * entry, a nested subhalo leaves its host's nest (parent -1, Rank 0) when its
 * host halo differs from its host's or its Rank is 0.
 *   host_group : DEVICE int64[nsub], HostHaloId (-1: field)
 *   is_central : DEVICE int32[nsub], 1 for each group's central (its most
 *                massive member, MemberTable.SubGroups[g][0])
 *   parent     : DEVICE int64[nsub], in/out: the subhalo whose NestedSubhalos
 *                holds this one, -1 if none
 *   rank       : DEVICE int32[nsub], in/out: Rank
 * Hosts outside [0, nsub) and nest chains deeper than 4096 are rejected with
 * HBT_ERR_INVALID_ARG (parent and rank then keep their input). */
int hbt_nest_subhalos(hbt_ctx *ctx, int64_t nsub, const int64_t *host_group,
                      const int32_t *is_central, int64_t *parent, int32_t *rank);

/* Merging of trapped subhalos after the unbinding of a snapshot
 * (SubhaloSnapshot_t::MergeSubhalos with DetectTraps, MergeRecursive and
 * Subhalo_t::MergeTo, src/subhalo_merge.cpp:136-262).
 *   1. DetectTraps: every subhalo not yet trapped looks up its chain of hosts
 *      (its parent; a group's other heads count as nested in the group's
 *      central -- its first head in visit_rank order, as in hbt_mask_subhalos) for
 *      the first host with nbound > 1 whose core (mass-weighted mean position
 *      and velocity of its min(nbound, 20) most bound particles) lies within
 *      phase-space distance 2 of the subhalo's most bound particle, in units
 *      of the core's dispersions; that host becomes its sink
 *      (sink_track_id, snapshot_index_of_sink = the context's snapshot index).
 *   2. With merge_trapped (HBTConfig.MergeTrappedSubhalos), each group is
 *      walked in post-order from its central; a trapped subhalo that is alive
 *      is marked dead at this snapshot and, when nbound > 1, its list is
 *      appended to its sink's without the IDs the sink already holds and its
 *      own list shrinks to its first particle (nbound 1).  Every host that
 *      received a subhalo with nbound > 1 is then unbound again (plain Unbind)
 *      and its list truncated (TruncateSource).
 *   snapshot   : device particle columns (every column read)
 *   list_index : DEVICE int64[n_list]; subhalo i's particle list is
 *                list_index[subs_host[i].offset .. + count) as snapshot
 *                positions, its first nbound entries bound and most bound first
 *                (e.g. the out_index of hbt_unbind_batch_indexed with
 *                offset = out_offset, count = nsource)
 *   subs_host  : HOST records, in/out.  Read: offset, count, parent, nbound,
 *                sink_track_id, snapshot indices, most bound position and
 *                velocity, and the Unbind inputs of the re-unbound hosts.
 *                Written: the merged state; offset/count then give each list
 *                in out_list (list_len = nsource = count, out_offset = offset).
 *                A failed call leaves them unchanged.
 *                A trapped, alive subhalo with nbound > 1 on entry is rejected
 *                (unbinding never leaves one: it clears the trap of a subhalo
 *                that stays bound and marks the others dead).
 *   host_group : HOST int64[nsub], -1 for field subhalos; a nested subhalo
 *                shares its parent's group
 *   visit_rank : HOST int64[nsub] or NULL, the sibling order of
 *                hbt_mask_subhalos (MergeRecursive visits in post-order)
 *   out_list   : DEVICE int64[out_capacity]; the lists after merging,
 *                compacted in subhalo order.  Sum of counts + nsub always
 *                suffices; *out_total (may be NULL) receives the length used
 *                (HBT_ERR_CAPACITY, with *out_total set, when it does not fit). */
int hbt_merge_subhalos(hbt_ctx *ctx, const hbt_particles *snapshot, const int64_t *list_index,
                       int64_t n_list, hbt_subhalo *subs_host, const int64_t *host_group,
                       const int64_t *visit_rank, int64_t nsub, int merge_trapped, int64_t *out_list,
                       int64_t out_capacity, int64_t *out_total);

/* ID -> index resolution: builds a device hash table of the n_table IDs
 * (index = position in `table_ids`; an ID given twice keeps its smallest
 * position), then looks up n_query IDs.  Unknown IDs and negative IDs give -1
 * (src/hash.tpp:52).  Device pointers.  The table is 2^pb partitions of 4,096
 * 16-byte slots, pb chosen so that a partition holds at most 2,560 IDs on
 * average (for example 2.1 GB for 70.7 M IDs), plus 24 bytes per ID of
 * scatter buffers; when one partition would hold more than 3,584 IDs (many
 * repeats of one ID) a single table of the next power of two at or above
 * 2 * n_table slots is used instead.  Both stay with the context for later
 * calls. */
int hbt_id_to_index(hbt_ctx *ctx, int64_t n_table, const int64_t *table_ids,
                    int64_t n_query, const int64_t *query_ids, int64_t *out_index);

/* Timing of the last hbt_unbind_batch call (device milliseconds from HIP
 * events on the context stream) and the number of energy evaluations it did. */
int hbt_last_batch_stats(hbt_ctx *ctx, double *device_ms, int64_t *n_evals,
                         int64_t *n_passes);

/* Execution tuning (no counterpart in the reference).  Results are identical
 * for every setting; only the speed changes.
 *   walk_group[b]: lanes that walk the tree together (8, 16, 32 or 64) for
 *                  subhalos of size bin b: list length < bin_edge[0] (b = 0),
 *                  < bin_edge[1] (b = 1), otherwise b = 2
 *   graph_passes : unbinding passes per HIP-graph launch (>= 1)
 *   use_graph    : 1 = replay each level's pass sequence as a captured HIP graph
 *                  (batches of more than one subhalo), 0 = direct launches, each
 *                  pass sized by the previous one (merge levels for its largest
 *                  active subhalo, node depths of its deepest tree)
 *   xcd_run      : the walk's workgroups reach each of the 8 XCDs in runs of
 *                  this many consecutive slot blocks (1, 2, 4, ... 64; 1 = the
 *                  hardware's round-robin order) */
typedef struct hbt_tuning {
  int32_t walk_group[3];
  int32_t bin_edge[2];
  int32_t graph_passes;
  int32_t use_graph;
  int32_t xcd_run;
} hbt_tuning;

int hbt_get_tuning(hbt_ctx *ctx, hbt_tuning *out);
/* Validates and sets; HBT_ERR_INVALID_ARG for an unsupported value. */
int hbt_set_tuning(hbt_ctx *ctx, const hbt_tuning *tuning);

/* Autotuner: times hbt_unbind_batch (same arguments) under the context's
 * current tuning and then candidate tunings -- one group size for all bins, then a separate group for the small-subhalo
 * bin at two bin edges, then the XCD run length, then direct launches or a
 * graph and the passes per graph launch -- keeping at each
 * step the setting with the smallest device time over `reps` runs, sets the
 * winner on the context and returns it in *chosen.  The outputs of the last
 * run are left in subs/out_index/out_energy (identical for every tuning). */
int hbt_autotune(hbt_ctx *ctx, const hbt_particles *particles, hbt_subhalo *subs,
                 const hbt_subhalo *subs_host, int64_t nsub, int64_t *out_index,
                 float *out_energy, int reps, hbt_tuning *chosen);

/* Instrumentation of the potential-walk kernel (no counterpart in the
 * reference; used by the benchmark for the roofline).  mode 0: off;
 * 1: time every walk launch with HIP events on the context stream;
 * 2: count, in every walk launch, the summed sources (leaves and accepted
 * nodes), the opened nodes, the targets, the lane slots spent by the lane
 * groups (items a group visited times its lanes, whether or not a lane
 * needed the item), the items of the trees walked and the wave steps of the
 * group walks (a wave's steps last until its slowest group is done).
 * Resets the accumulators. */
int hbt_profile(hbt_ctx *ctx, int mode);
int hbt_profile_read(hbt_ctx *ctx, double *walk_ms, int64_t *walk_launches,
                     int64_t *interactions, int64_t *opened, int64_t *targets,
                     int64_t *lane_slots, int64_t *tree_items, int64_t *wave_steps);
/* The segmented sorts of the unbinding passes (the octant keys of each tree,
 * the energy order): mode 1 times every sort (tile sort and its merge levels)
 * with HIP events, mode 2 counts the elements sorted and the elements times
 * the merge levels each needed (every level reads and writes each element's
 * 12-byte key and value once, the tile sort once more). */
int hbt_profile_sorts(hbt_ctx *ctx, double *sort_ms, int64_t *sort_calls, int64_t *elements,
                      int64_t *element_levels);

#ifdef __cplusplus
}
#endif
#endif /* HBT_UNBIND_H */
