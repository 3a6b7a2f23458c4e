// Internal state of an aggregation handle (the reference's Aggregation).
#pragma once
#include "common.hpp"

// Basis storage (FP64 master copy, FP32 copy refreshed on demand):
//   [coarse site xc][block site xb][spin-colour e 12][basis i nb]
// i.e. for each block a (|B|*12) x nb row-major panel, so that a wavefront
// reading one row's nb entries and the next rows is one contiguous stream
// (project, promote, Gram-Schmidt and the coarse-operator build all walk
// these panels).  Chirality of row e is e / 6 (spin / 2).
struct mgc_aggregation {
  mgc::Lat4 fine, coarse, blk;
  int Vf = 0, Vc = 0, Vb = 0, nb = 0;
  mgc::cplx<double>* basis64 = nullptr;
  mgc::cplx<float>* basis32 = nullptr;
  bool basis_set = false;
  bool basis32_valid = false;
  int* blk_fine = nullptr;  // [Vc][Vb] fine site of (block, block site)
  int* fine_row = nullptr;  // [Vf] panel row (xc*Vb + xb) of a fine site
};

namespace mgc {
// FP32 copy of the basis, refreshed when stale
const cplx<float>* agg_basis32(mgc_aggregation* a, hipStream_t s);
__host__ __device__ inline size_t agg_row(int nb, int Vb, int xc, int xb, int e) {
  return (((size_t)xc * Vb + xb) * 12 + e) * nb;
}
}  // namespace mgc
