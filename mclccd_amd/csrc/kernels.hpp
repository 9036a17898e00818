// Host-side launchers of the gfx950 kernels (defined in build.hip, traverse.hip,
// batch.hip and adjoint.hip).  Plain pointers only.
#pragma once
#include "common.hpp"

namespace mclg {

void launch_pack(int dim, const double* V0, const double* V1, int nv, double* rec, hipStream_t s);
void launch_minmax_i32(const int* a, long long n, int* out2, hipStream_t s);
void launch_fill_u8(uint8_t* p, long long n, uint8_t v, hipStream_t s);
void launch_fill_i32(int* p, long long n, int v, hipStream_t s);
void launch_fill_u64(unsigned long long* p, long long n, unsigned long long v, hipStream_t s);
void launch_reptri(const int* P, int np, int pdim, int nv, int* owner, unsigned long long* keys, int* vals,
                   uint32_t table_size, uint8_t* flags, hipStream_t s);
int leaf_box_blocks(int np);
void launch_leaf_boxes(int dim, const double* rec, const int* P, int np, int pdim, const int* active, double eta,
                       void* boxes, uint8_t* act, double* partial, unsigned long long* cbounds, hipStream_t s);
void launch_morton(int dim, const void* boxes, int np, const unsigned long long* cbounds, uint32_t* keys,
                   uint32_t* vals, hipStream_t s);
int radix_tiles(int n);
void launch_radix_sort(uint32_t* keys, uint32_t* vals, uint32_t* ktmp, uint32_t* vtmp, uint32_t* hist, int n, int nbits,
                       hipStream_t s);
void launch_lbvh(int dim, const uint32_t* keys, const uint32_t* order, int n, const void* boxes, const uint8_t* act,
                 void* nodes, int* lchild, int* rchild, int* rng_l, int* rng_r, int* visits, hipStream_t s);
void launch_traverse(int dim, const void* nodes, const void* boxes, const uint8_t* act, const uint32_t* order, int n,
                     const int* rank, int2* pairs, unsigned long long cap, unsigned long long* counters, hipStream_t s);

struct ExpandLaunch {
  const int2* pairs;
  long long npairs;
  const int* P;
  const uint8_t* flags;
  const double* rec;
  int discrete, continuous;
  int4* vfq;
  unsigned long long vf_cap;
  int4* eeq;
  unsigned long long ee_cap;
  int2* disc;
  unsigned long long disc_cap;
  unsigned long long* counters;
  int num_cus;
};
void launch_expand(int dim, const ExpandLaunch& L, hipStream_t s);

struct AccdLaunch {
  const int4* q;
  long long n;
  const double* rec;
  double eta;
  double* cand_toi;
  int4* hit_sten;
  uint8_t* hit_vf;
  double* hit_toi;
  unsigned long long hit_cap;
  unsigned long long* counters;
  int num_cus;
};
void launch_accd(int dim, bool vf, const AccdLaunch& L, hipStream_t s);

}  // namespace mclg
