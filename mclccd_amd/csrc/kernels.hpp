// Host-side launchers of the gfx950 kernels (defined in build.hip, traverse.hip,
// batch.hip and adjoint.hip).  Plain pointers only.
#pragma once
#include "common.hpp"

namespace mclg {

// Several independent scenes solved together (all pointers null for one scene).
// Primitives and vertices of a scene are contiguous; the tree is keyed by
// (scene, Morton) so each scene is one subtree.
struct SceneArgs {
  int n = 0;                           // number of scenes
  const int* prim_scene = nullptr;     // scene of each primitive
  const int* vert_scene = nullptr;     // scene of each vertex
  const int* prim_off = nullptr;       // n+1 primitive offsets
  const double* box_eta = nullptr;     // per scene
  const double* vf_eta = nullptr;      // per scene
  const double* ee_eta = nullptr;      // per scene
  unsigned long long* bounds = nullptr;  // per scene centre bounds, ordered encoding [lo(dim), hi(dim)]
};

// rec: per-vertex records (both time steps); v1 (may be null): V1 per vertex, v1_stride(dim) doubles each.
// f32: round the coordinates to float (float-precision tree)
void launch_pack(int dim, const double* V0, const double* V1, int nv, double* rec, hipStream_t s, double* v1 = nullptr,
                 bool f32 = false);
void launch_minmax_i32(const int* a, long long n, int* out2, hipStream_t s);
void launch_fill_u8(uint8_t* p, long long n, uint8_t v, hipStream_t s);
void launch_fill_i32(int* p, long long n, int v, hipStream_t s);
void launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t s);
void launch_reptri(const int* P, int np, int pdim, int nv, int* owner, unsigned long long* keys, int* vals,
                   uint32_t table_size, uint8_t* flags, hipStream_t s);
int leaf_box_blocks(int np);
void launch_leaf_boxes(int dim, const double* rec, const int* P, int np, int pdim, const int* active, double eta,
                       void* boxes, uint8_t* act, double* partial, unsigned long long* cbounds, const SceneArgs& S,
                       hipStream_t s, bool f32 = false);  // f32: padded bounds rounded to float
void launch_bounds_init(unsigned long long* b, int nscenes, int dim, unsigned long long lo0, unsigned long long hi0,
                        hipStream_t s);
void launch_morton(int dim, const void* boxes, int np, const unsigned long long* cbounds, const SceneArgs& S,
                   uint32_t* keys, uint32_t* vals, uint32_t* mort, hipStream_t s);
void launch_scene_keys(const uint32_t* order, int n, const int* prim_scene, uint32_t* hi, hipStream_t s);
void launch_gather_u32(const uint32_t* order, int n, const uint32_t* src, uint32_t* dst, hipStream_t s);
int radix_tiles(int n);
size_t radix_hist_words(int n);
// In-place exclusive scan of n 64-bit values; part needs scan_part_words(n) entries.
long long scan_part_words(long long n);
long long expand_chunks(unsigned long long pair_cap);  // 64-pair chunks of the candidate kernels
void launch_scan_u64(unsigned long long* data, long long n, unsigned long long* part, hipStream_t s);  // uint32 words of the hist buffer launch_radix_sort needs
void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* ktmp, uint32_t* vtmp, uint32_t* hist, int n, int nbits,
                       hipStream_t s);
int lbvh_frontier_cap(int n);
inline size_t sorted_leaf_bytes(int dim) { return dim == 3 ? sizeof(SortedLeaf<3>) : sizeof(SortedLeaf<2>); }
// Also writes sboxes[s] = boxes[order[s]] with the leaf's flags (SortedLeaf<dim>,
// sorted_leaf_bytes(dim) each)
// and, when cp is given and n > 1, the child-pair records (ChildPair<dim> per
// internal node) and *root.
void launch_lbvh(int dim, const uint32_t* keys, const uint32_t* hi, const uint32_t* order, int n, const void* boxes,
                 const uint8_t* act, void* nodes, void* sboxes, int* lchild, int* rchild, int* rng_l, int* rng_r,
                 int* visits, int4* frontier, unsigned long long* nfrontier, const SceneArgs& S, hipStream_t s,
                 void* cp, int* root);
// Leaf pairs as sorted positions (s, k), s < k.  Several scenes: scene_of_sorted
// gives each sorted leaf's scene and scene_off the scenes' sorted-leaf offsets
// (n+1 entries); pairs never cross scenes.
// waves per traversal block: 1..kTraverseMaxWaves (larger values are clamped)
constexpr int kTraverseMaxWaves = 4;
void launch_traverse(int dim, const void* nodes, const void* cp, const int* root, const void* sboxes, int n,
                     int2* pairs, unsigned long long cap, unsigned long long* counters,
                     const uint32_t* scene_of_sorted, const int* scene_off, int waves, hipStream_t s);
// The leaf arrays of faces (head, rank, prim) for sorted positions (prim = order[s]); rank may be null.
void launch_face_records(int dim, const uint32_t* order, int n, const int* P, const uint8_t* flags, const int* rank,
                         const FaceArrays& faces, hipStream_t s);
// pairs[i] (sorted positions) -> primitive pairs, lower rank first.
void launch_pairs_to_prims(int2* pairs, long long n, const FaceArrays& faces, hipStream_t s);

// Kernels that follow a variable-size queue read its length from the device
// counter (so a captured graph needs no host round trip); the *_hint sizes
// only shape the grid (-1: the capacity).
struct ExpandLaunch {
  const int2* pairs;
  const unsigned long long* npairs_dev;
  unsigned long long pair_cap;
  long long npairs_hint;
  FaceArrays faces;  // leaf arrays in sorted order (pairs hold sorted positions) and V1 per vertex
  // scratch of expand_chunks(pair_cap) + 1 entries each, and scan_part_words of that
  unsigned long long* off_vf;
  unsigned long long* off_ee;
  unsigned long long* scan_part;
  int discrete, continuous;
  int4* vfq;
  unsigned long long vf_cap;
  int4* eeq;
  unsigned long long ee_cap;
  int2* disc;
  unsigned long long disc_cap;
  int2* odep = nullptr;  // pairs whose discrete result depends on the argument order (null: not checked)
  unsigned long long odep_cap = 0;
  unsigned long long* counters;
  int num_cus;
  int blocks_per_cu;
  int chunk;
};
void launch_expand(int dim, const ExpandLaunch& L, hipStream_t s);

struct AccdLaunch {
  const int4* q;
  const unsigned long long* n_dev;
  unsigned long long cap;
  long long n_hint;
  const double* rec;
  double eta;
  const int* vert_scene;
  const double* scene_eta;
  double* cand_toi;                  // per stencil: toi, -1 = no hit (null: not written; needs hit_sten)
  unsigned long long* chunk_hits;    // hits per 64-stencil chunk (accd_chunks(cap) entries, zeroed by the
                                     // caller; null: not written; needs hit_sten)
  unsigned long long* counters;
  int num_cus;
  int blocks_per_cu;
  int one_sided = 0;                 // VF: CTCD with the wrong-side test instead of ACCD
  int f32 = 0;                       // float-precision tree: float ACCD; CTCD TOIs rounded to float
  // direct hit list: hits appended through counters[C_HITS] (zeroed by the
  // caller), records past hit_cap dropped.  Null, or a queue of 2^32 or
  // more stencils: only cand_toi / chunk_hits, and launch_hits builds the list.
  int4* hit_sten = nullptr;
  uint8_t* hit_vf = nullptr;
  double* hit_toi = nullptr;
  unsigned long long hit_cap = 0;
  // ACCD with the direct hit list and no chunk_hits suspends the stencils still
  // running after three iterations and resumes them in dense chunks (same
  // results); no_suspend runs every stencil in place.
  int no_suspend = 0;
  // 3D CTCD (one_sided) with the direct hit list and no chunk_hits: cap entries
  // for the queue indices of the stencils that pass the prefilter, listed by a
  // selection pass (counters[C_CTCD_SEL], cleared by the caller); null: one pass.
  uint32_t* ctcd_sel = nullptr;
};
// Whether launch_accd appends hits directly for a queue of this capacity.
inline bool accd_direct_hits(unsigned long long cap) { return cap < (1ull << 32); }
void launch_accd(int dim, bool vf, const AccdLaunch& L, hipStream_t s);
long long accd_chunks(unsigned long long cap);

// Hit list from the per-stencil TOIs: VF hits, then EE hits, each in queue
// order.  off holds the hits per 64-stencil chunk written by the ACCD
// kernels (VF chunks, then EE chunks, accd_chunks(vf_cap) + accd_chunks(ee_cap)
// entries, plus one); with scan it is scanned here and the total is written
// to counters[C_HITS].
struct HitsLaunch {
  const int4* vfq;
  const unsigned long long* n_vf_dev;
  unsigned long long vf_cap;
  const double* vf_toi;
  const int4* eeq;
  const unsigned long long* n_ee_dev;
  unsigned long long ee_cap;
  const double* ee_toi;
  unsigned long long* off;
  unsigned long long* scan_part;  // scan_part_words(entries of off)
  int4* hit_sten;
  uint8_t* hit_vf;
  double* hit_toi;
  unsigned long long hit_cap;
  unsigned long long* counters;
  int num_cus;
  int blocks_per_cu;
};
// scan=true: scan off and set counters[C_HITS] first; false: only the
// compaction (after a hit-capacity regrowth).
void launch_hits(int dim, const HitsLaunch& L, bool scan, hipStream_t s);

}  // namespace mclg
