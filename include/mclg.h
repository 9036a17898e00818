/* mclg.h — C ABI of the MI355X-native collision pipeline.
 *
 * This is the boundary under the `mcl::BVHTree<double,DIM>` facade
 * (include/MCL/BVHTree.hpp): plain pointers and sizes, status codes, no C++ or
 * torch types.  Each entry point names the reference interface it replaces
 * (paths relative to the mattoverby/mclccd repository).
 *
 * Semantics shared by all entry points:
 *   - Every function returns MCLG_OK (0) or a negative MCLG_ERR_* code; the
 *     message of the last error is available from mclg_last_error().
 *   - Host-pointer entry points copy their inputs to the GPU; *_device entry
 *     points take device pointers (HBM-resident inputs) and copy nothing in.
 *   - update/traverse are synchronous: when they return, results are on the
 *     device and their counts are known.  Results stay valid until the next
 *     traverse on the same context.
 *   - A context is bound to the HIP device current at mclg_create and is not
 *     thread-safe: calls on one context must not overlap.
 */
#ifndef MCLG_H
#define MCLG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MCLG_OK = 0,
  MCLG_ERR_ARG = -1,         /* invalid argument (sizes, out-of-range indices, dim) */
  MCLG_ERR_HIP = -2,         /* a HIP runtime call failed */
  MCLG_ERR_NOMEM = -3,       /* device allocation failed */
  MCLG_ERR_STATE = -4,       /* call order (e.g. traverse before update) */
  MCLG_ERR_UNSUPPORTED = -5, /* request with no GPU path (e.g. the TOI adjoint of CTCD hits) */
  MCLG_ERR_CAPACITY = -6     /* caller buffer smaller than the result */
};

typedef struct mclg_ctx mclg_ctx;

/* mcl::BVHTree<T,DIM>::Options — include/MCL/BVHTree.hpp:42-64 (same fields,
 * same defaults; booleans as int). */
typedef struct mclg_options {
  double box_eta;    /* padding of tree boxes (default 1e-6) */
  double vf_ccd_eta; /* VF gap; update() clamps it to box_eta */
  double ee_ccd_eta; /* EE gap; update() clamps it to box_eta */
  int reptri;        /* representative triangles (default 1) */
  int threaded;      /* 0: the reference's serial traversal order for replay (MCLG_ORIENT_KDBVH implied) */
  int discrete;      /* discrete tri-tri (3D) / edge-edge (2D) at V1 (default 1) */
  int continuous;    /* ACCD from V0 to V1 (default 1) */
  int vf_one_sided;  /* VF stencils through CTCD with the wrong-side test (NarrowPhase.cpp:283-419)
                      * instead of ACCD; EE stays ACCD (BVHTree.cpp:443-455) */
  int ee_robust;     /* unused on this path (as in the reference) */
} mclg_options;

/* Fills the reference defaults (BVHTree.hpp:53-63). */
void mclg_default_options(mclg_options* o);

/* mcl::BVHTree<T,DIM>::BVHTree() — include/MCL/BVHTree.cpp:20-24.  dim is 2 or 3. */
int mclg_create(int dim, mclg_ctx** out);
void mclg_destroy(mclg_ctx* ctx);

/* The public `options` member — BVHTree.hpp:64. */
int mclg_set_options(mclg_ctx* ctx, const mclg_options* o);
int mclg_get_options(const mclg_ctx* ctx, mclg_options* o);

/* ctx == NULL: why the calling thread's last mclg_create failed. */
const char* mclg_last_error(const mclg_ctx* ctx);

/* Runs the context's work on a caller-provided hipStream_t (NULL: the context's
 * own stream).  No reference counterpart; used by the PyTorch wrapper. */
int mclg_set_stream(mclg_ctx* ctx, void* hip_stream);

/* BVHTree<T,DIM>::update(const T* V0, const T* V1, const int* P, int np, int pdim,
 * const int* active) — include/MCL/BVHTree.hpp:81-82, BVHTree.cpp:27-126.
 * V0, V1: nv x dim row-major; P: np x pdim row-major; active: nv ints or NULL.
 * nv is explicit here (the reference implies it from P).  np == 0 clears the
 * tree.  Indices outside [0, nv) -> MCLG_ERR_ARG. */
int mclg_update(mclg_ctx* ctx, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                const int* active);
int mclg_update_device(mclg_ctx* ctx, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                       const int* active);

/* Refit-only update.  No reference counterpart (the reference rebuilds its
 * tree on every update, BVHTree.cpp:124).  Same arguments and results as
 * mclg_update for the same mesh: recomputes the leaf boxes (and active flags)
 * from V0/V1 and the node volumes bottom-up, but keeps the previous full
 * update's leaf order and hierarchy, skipping the Morton codes and the sort.
 * Leaf pairs, candidates, hits and discrete pairs equal a rebuild's (the sets
 * do not depend on the hierarchy); pairs are oriented by the kept leaf order.
 * Traversal slows as the kept hierarchy ages under large motion: call
 * mclg_update again then.  P must be the last update's; nv, np and pdim must
 * match it, else MCLG_ERR_STATE.  Representative-triangle flags are kept, as
 * the reference keeps them when the leaf count is unchanged (BVHTree.cpp:44-48). */
int mclg_update_refit(mclg_ctx* ctx, const double* V0, const double* V1, int nv, const int* P, int np, int pdim,
                      const int* active);
int mclg_update_refit_device(mclg_ctx* ctx, const double* V0, const double* V1, int nv, const int* P, int np,
                             int pdim, const int* active);

/* BVHTree<T,DIM>::traverse(const T* V0, const T* V1, const int* P) const —
 * include/MCL/BVHTree.hpp:86, BVHTree.cpp:129-159 (+ collide, get_candidates,
 * default_narrow_phase, default_discrete_test: :230-500).  Uses nv from the last
 * update; P must have pdim == dim.  Instead of invoking callbacks it leaves the
 * continuous hits (stencil, is_vf, toi), the discrete hits (p0, p1) and, on
 * request, the candidate stencils and leaf pairs on the device. */
#define MCLG_KEEP_CANDIDATES 1u /* keep every candidate stencil (for filter_pair / narrow_phase replay) */
#define MCLG_KEEP_LEAF_PAIRS 2u /* keep the broadphase leaf pairs */
/* also run the discrete test with the two primitives of each leaf pair swapped
 * and list the pairs whose result changes (tt_isect.hpp's predicate is not
 * symmetric for degenerate triangles, nor the 2D segment test for collinear
 * segments): exactly the pairs where the orientation of the reference's KdBVH
 * decides the discrete result (mclg_copy_order_dependent) */
#define MCLG_CHECK_ORDER 4u
int mclg_traverse(mclg_ctx* ctx, const double* V0, const double* V1, const int* P, unsigned flags);
int mclg_traverse_device(mclg_ctx* ctx, const double* V0, const double* V1, const int* P, unsigned flags);

/* Sizes of the last traverse.  candidates and dist_evals are always counted. */
typedef struct mclg_counts {
  int64_t leaf_pairs;
  int64_t candidates;
  int64_t hits;
  int64_t discrete;
  int64_t dist_evals; /* ACCD pair_distance evaluations performed */
  int64_t vf_candidates;  /* of candidates: vertex-face stencils (the rest are edge-edge) */
  int64_t ee_dist_evals;  /* of dist_evals: those of edge-edge stencils */
  int64_t order_dependent; /* leaf pairs whose discrete result depends on the argument order (MCLG_CHECK_ORDER) */
  int64_t traverse_steps;  /* traversal wave-steps: child-pair records read, summed over the 64-leaf packets */
  int64_t suspended;       /* ACCD stencils suspended after three iterations and resumed later (same results;
                              0 on frames whose hit list may still outgrow its capacity) */
} mclg_counts;
int mclg_get_counts(const mclg_ctx* ctx, mclg_counts* out);

/* Copies of the last traverse's results to host memory.  Each writes
 * min(count, cap) records and fails with MCLG_ERR_CAPACITY if cap < count.
 *   hits:       sten4 = int[4*cap], is_vf = uint8[cap], toi = double[cap]
 *               (the arguments of append_pair, BVHTree.hpp:102)
 *   discrete:   int[2*cap] (the arguments of append_discrete, BVHTree.hpp:108)
 *   candidates: sten4 / is_vf as above plus the ACCD toi of each candidate
 *               (-1 when it does not hit), every candidate
 *               (MCLG_KEEP_CANDIDATES)
 * Hits are the VF hits then the EE hits, each in the order the GPU finished
 * them; candidates follow the leaf-pair order.  Both vary from run to run (as
 * does the reference's concurrent callback order), so compare them as sets.
 *   leaf pairs: int[2*cap] (only when the traverse had MCLG_KEEP_LEAF_PAIRS;
 *               MCLG_ERR_STATE otherwise)
 * Stencils follow BVHTree.cpp:336-395: VF (v, f0, f1, f2) with -1 padding in
 * 2D, EE (a0, a1, b0, b1).  Pair orientation: the lower primitive index
 * first, or the lower rank when a rank is set (mclg_set_rank). */
int mclg_copy_hits(mclg_ctx* ctx, int* sten4, unsigned char* is_vf, double* toi, int64_t cap);
int mclg_copy_discrete(mclg_ctx* ctx, int* pairs2, int64_t cap);
int mclg_copy_candidates(mclg_ctx* ctx, int* sten4, unsigned char* is_vf, double* toi, int64_t cap);
int mclg_copy_leaf_pairs(mclg_ctx* ctx, int* pairs2, int64_t cap);
/* The order-dependent leaf pairs of a traverse with MCLG_CHECK_ORDER, as
 * int[2*cap] primitive pairs oriented like the discrete hits; a pair is in the
 * discrete list iff the discrete test holds in that orientation.  No reference
 * counterpart (diagnostic of the orientation; MCLG_ERR_STATE without the flag). */
int mclg_copy_order_dependent(mclg_ctx* ctx, int* pairs2, int64_t cap);

/* The leaves of the last update (the reference's public `leaves` member,
 * BVHTree.hpp:39, BVHLeaf.hpp:12-52): per primitive the swept, t0 and t1 boxes
 * as [sw_min, sw_max, t0_min, t0_max, t1_min, t1_max] (6*dim doubles, padded
 * by box_eta), the active flag and the representative-triangle flags
 * (bits 0..2 = v[0..2], bits 3..5 = e[0..2]).  Any output may be NULL. */
int mclg_copy_leaves(mclg_ctx* ctx, double* boxes, unsigned char* active, unsigned char* flags, int64_t cap);

/* The tree of the last update, for visitor traversals on the host (replaces
 * the KdBVH the reference walks in traverse(BVHTraverse*), BVHTree.cpp:162-166).
 * Node ids: internal node k (0 <= k < np-1) is k, the leaf at sorted position
 * s is np-1+s.  lchild / rchild: np-1 entries each (the internal nodes'
 * children); order[s]: the primitive at sorted position s (np entries); *root:
 * the root node id (0 for a single leaf, -1 for an empty tree).  A leaf
 * primitive's boxes are those of mclg_copy_leaves; an internal node's volume
 * is the merge of its children's (BVHLeaf.cpp:9-23).  Any output may be NULL;
 * cap must be at least np. */
int mclg_copy_tree(mclg_ctx* ctx, int* lchild, int* rchild, int* order, int* root, int64_t cap);

/* Device pointers to the last traverse's results (valid until the next traverse).
 * Candidates live in two queues, VF then EE; their per-candidate toi arrays are
 * set only when the traverse kept candidates, leaf_pairs2 only when it kept
 * leaf pairs (NULL otherwise). */
typedef struct mclg_device_results {
  const int* hit_sten4;
  const unsigned char* hit_is_vf;
  const double* hit_toi;
  const int* discrete2;
  const int* vf_sten4;
  const double* vf_toi;
  int64_t n_vf;
  const int* ee_sten4;
  const double* ee_toi;
  int64_t n_ee;
  const int* leaf_pairs2;
} mclg_device_results;
int mclg_get_device_results(const mclg_ctx* ctx, mclg_device_results* out);

/* Orientation of leaf pairs.  The reference orients each pair by the in-order
 * position of its leaves in its KdBVH (BVHTree.cpp:288-301, KdBVH.hpp:202,212);
 * that order decides the EE stencil order and the tri-tri argument order.
 * rank[prim] (np ints) makes the GPU orient pairs the same way (lower rank
 * first); NULL restores the default (lower primitive index first).  Cleared
 * when the primitive count changes. */
int mclg_set_rank(mclg_ctx* ctx, const int* rank, int np);

/* Orientation of leaf pairs (no reference counterpart; the reference always
 * has its KdBVH):
 *   MCLG_ORIENT_INDEX  lower primitive index first (or the rank of
 *                      mclg_set_rank) — the default;
 *   MCLG_ORIENT_KDBVH  the reference's own orientation: every update and refit
 *                      rebuilds, on the host, the order of the reference's
 *                      median-split KdBVH over the swept-box centres
 *                      (KdBVH.hpp:103-134,192-222: std::nth_element, the same
 *                      comparator and input order) and orients each pair by it
 *                      (BVHTree.cpp:288-301).  EE stencils, EE TOIs and discrete
 *                      pairs then equal the reference's record for record; the
 *                      host build costs O(np log np) per update.
 * An update with options.threaded == 0 (the reference's serial traversal,
 * BVHTree.cpp:152-157) uses the KdBVH orientation whatever the mode, so that
 * mclg_copy_events reproduces the reference's serial callback sequence. */
#define MCLG_ORIENT_INDEX 0
#define MCLG_ORIENT_KDBVH 1
int mclg_set_orientation(mclg_ctx* ctx, int mode);
int mclg_get_orientation(const mclg_ctx* ctx, int* mode);

/* The scalar type T of mcl::BVHTree<T,DIM> (BVHTree.cpp:504-507):
 *   MCLG_PRECISION_F64  T = double (the default);
 *   MCLG_PRECISION_F32  T = float: vertex coordinates and the three etas are
 *                       rounded to float at each update / traverse, leaf boxes
 *                       are padded in float (BVHLeaf bounds), the KdBVH order
 *                       uses float centres, and ACCD runs in float arithmetic as
 *                       NarrowPhaseACCD<float,DIM> does (NarrowPhase.cpp:634-711,
 *                       757-770); CTCD and the discrete tests widen to double as
 *                       the reference's float specialisations do
 *                       (NarrowPhase.cpp:208-219,259-269,425-444) and CTCD TOIs
 *                       are rounded to float.  Inputs stay double arrays (a
 *                       float caller widens exactly); outputs are float values
 *                       held in doubles.  The TOI adjoint is FP64 only
 *                       (MCLG_ERR_UNSUPPORTED after a float traverse); the
 *                       batched path is FP64 only.
 * Set it before the first update (MCLG_ERR_STATE afterwards, as T is fixed for
 * the life of a reference tree). */
#define MCLG_PRECISION_F64 0
#define MCLG_PRECISION_F32 1
int mclg_set_precision(mclg_ctx* ctx, int precision);
int mclg_get_precision(const mclg_ctx* ctx, int* precision);

/* The reference KdBVH's leaf order for given swept-box centres, on the host
 * (no GPU): rank[i] = in-order position of leaf i (KdBVH.hpp:192-222, objects
 * reordered as in KdBVH.hpp:130-133); gapnode[g] (n-1 entries, may be NULL) =
 * the index of the internal node that splits positions g and g+1, internal
 * nodes numbered in the order the reference pushes their volumes, which is
 * the order of the traversal's frontlist (BVHTree.cpp:168-185).  centres:
 * n x dim, (min + max) / 2 of each swept box. */
int mclg_kdbvh_order(int dim, const double* centres, int n, int* rank, int* gapnode);

/* The callbacks of the last traverse in the order the reference's serial
 * traversal calls them (BVHTree.cpp:152-157,230-302), for host replay: leaf
 * pairs in visiting order — the reference's frontlist order under
 * MCLG_ORIENT_KDBVH or threaded == 0 (by the node that first separates the two
 * leaves, then by their positions), else by (rank of the first leaf, rank of
 * the second) — and per leaf pair its CCD records in get_candidates order
 * (VF first leaf -> second, VF second -> first, EE; BVHTree.cpp:327-395)
 * followed by its discrete hit.  events[k] >= 0 indexes the hit list
 * (mclg_copy_hits) or, with over_candidates, the candidate list
 * (mclg_copy_candidates); events[k] < 0 is discrete hit -1-events[k]
 * (mclg_copy_discrete).  group[k] numbers the leaf pair of event k (0, 1, ...
 * over the pairs that have events).  Either output may be NULL; *n_events
 * receives the count (hits or candidates, plus discrete hits).  Needs a
 * traverse with MCLG_KEEP_LEAF_PAIRS (and MCLG_KEEP_CANDIDATES for
 * over_candidates).  Host work: O(records log records). */
int mclg_copy_events(mclg_ctx* ctx, int over_candidates, int64_t* events, int64_t* group, int64_t cap,
                     int64_t* n_events);

/* Device milliseconds of the phases of the last update / traverse, measured
 * with HIP events on the context's stream:
 *   0 validate+pack  1 reptri  2 leaf boxes  3 morton  4 sort  5 lbvh (update)
 *   6 validate+pack  7 traversal kernel  8 candidate + discrete kernel
 *   9 ACCD kernels (traverse; 7-9 bracket the final launches only)
 *  10 TOI adjoint (last mclg_toi_grad: sort, adjoint kernels)
 *  11 VF ACCD kernel  12 EE ACCD kernel (both inside 9, which adds the hit
 *     compaction on a frame whose hit list outgrew its capacity)
 * Writes min(n, MCLG_NUM_TIMINGS) values. */
#define MCLG_NUM_TIMINGS 13
int mclg_get_timings(const mclg_ctx* ctx, double* ms, int n);

/* TOI adjoint.  No reference counterpart.  For the CCD hits of the last
 * traverse (in the order mclg_copy_hits returns them) and per-hit weights w,
 * writes dL/dV0 and dL/dV1 (nv x dim each) of L = sum_i w_i toi_i, where toi_i
 * is the ACCD time of impact as computed by NarrowPhase.cpp:634-711 (its
 * derivative wherever it is smooth; hits already within the gap have toi 0 and
 * contribute nothing).  Each distance is differentiated at the closest points
 * the forward pass computed (envelope theorem), except in the reference's
 * point_on_triangle quirk branch (Projection.hpp:131), which is differentiated
 * through.  Host pointers, or device pointers when device_ptrs.  Allocates
 * scratch on the first call (about 16 bytes per hit for the visiting order).
 * MCLG_ERR_UNSUPPORTED when the last traverse ran with vf_one_sided (CTCD
 * times have no adjoint here); MCLG_ERR_STATE when no traverse completed
 * since the last update or refit (the hits would be paired with vertex
 * positions they were not computed from). */
int mclg_toi_grad(mclg_ctx* ctx, const double* w, double* dV0, double* dV1, int device_ptrs);

/* The TOI adjoint without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * gradient as mclg_toi_grad(ctx, w, dV0, dV1, 1) (device pointers only) up to
 * the order of the floating-point sums; w, dV0 and dV1 are the caller's.  Its
 * scratch (the hits' visiting order and the queue of multi-step hits) is one
 * static array of about 25 MB that each GPU the library runs on holds until
 * the process ends, sized for at most MCLG_NOALLOC_MAX_HITS hits: a larger
 * hit list fails with MCLG_ERR_CAPACITY.  Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (from several contexts or
 * threads) race.  The context's device must be current. */
#define MCLG_NOALLOC_MAX_HITS (1 << 20)
int mclg_toi_grad_noalloc(mclg_ctx* ctx, const double* w, double* dV0, double* dV1);

/* ------------------------------------------------------------------------
 * Batched scenes.  No reference counterpart: the reference runs one
 * mcl::BVHTree per scene, one after the other.  Here N independent scenes
 * (each exactly the problem one BVHTree::update + traverse solves, with its
 * own gaps) run as one pass over concatenated arrays: one LBVH keyed by
 * (scene, Morton code), every kernel's queue length read on the device, the
 * whole frame captured once as a hipGraph and replayed.
 * ------------------------------------------------------------------------ */
typedef struct mclg_batch mclg_batch;
int mclg_batch_create(int dim, mclg_batch** out);
void mclg_batch_destroy(mclg_batch* b);
const char* mclg_batch_last_error(const mclg_batch* b);

/* Scene layout: scene s owns vertices [voff[s], voff[s+1]) and primitives
 * [poff[s], poff[s+1]) of the concatenated arrays; P (poff[n] x dim) holds each
 * scene's own vertex indices.  opts: one entry for all scenes or one per scene
 * (gaps per scene; reptri/discrete/continuous must agree).  Computes the
 * representative-triangle flags once for this topology.  Host pointers. */
int mclg_batch_set_scenes(mclg_batch* b, int nscenes, const int* voff, const int* poff, const int* P,
                          const mclg_options* opts, int nopts);

/* The active argument of BVHTree::update (BVHTree.cpp:27,88,114-115) for every
 * scene at once: one int per concatenated vertex (voff[n] of them), a leaf
 * active when any of its vertices is; a scene without a mask of its own takes
 * all ones.  Used by every later frame until changed; NULL: no mask (every
 * leaf active).  Host pointer, or a device pointer when device_ptrs != 0;
 * the values are copied (a later change needs another call). */
int mclg_batch_set_active(mclg_batch* b, const int* active, int device_ptrs);

/* One frame (update + traverse of every scene).  V0, V1: voff[n] x dim,
 * concatenated; host pointers, or device pointers when device_ptrs != 0. */
int mclg_batch_run(mclg_batch* b, const double* V0, const double* V1, int device_ptrs);
/* Host milliseconds of the last frame (graph launch to counters on the host). */
double mclg_batch_last_ms(const mclg_batch* b);

int mclg_batch_get_counts(const mclg_batch* b, mclg_counts* out);
int mclg_batch_scene_counts(mclg_batch* b, int64_t* hits, int64_t* discrete);
/* Results with their scene and scene-local indices. */
int mclg_batch_copy_hits(mclg_batch* b, int* scene, int* sten4, unsigned char* is_vf, double* toi, int64_t cap);
int mclg_batch_copy_discrete(mclg_batch* b, int* scene, int* pairs2, int64_t cap);

/* Execution parameters (negative / zero: keep): hipGraph replay on/off, waves
 * per traversal block (1-4; larger values are clamped), resident blocks per CU of the queue kernels,
 * 64-item chunks a wave takes per step.  out4 receives the current values. */
int mclg_batch_set_params(mclg_batch* b, int use_graph, int tr_waves, int blocks_per_cu, int chunk);
int mclg_batch_get_params(const mclg_batch* b, int* out4);
/* Times frames on the last inputs over a small grid of the parameters above and
 * keeps the fastest; results are cached per (GPU architecture, dim, scene
 * count, primitive count) in cache_path (NULL: no cache). */
int mclg_batch_autotune(mclg_batch* b, const char* cache_path, int reps);

#ifdef __cplusplus
}
#endif

#endif /* MCLG_H */
