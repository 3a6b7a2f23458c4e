// Internal state of an aggregation handle (the reference's Aggregation).
#pragma once
#include "common.hpp"

// Basis storage (FP64 master copy, FP32 copy refreshed on demand):
//   [coarse site xc][block site xb][site element e Ef][basis i nb]
// i.e. for each block a (|B|*Ef) x nb row-major panel, so that a wavefront
// reading one row's nb entries and the next rows is one contiguous stream
// (project, promote, Gram-Schmidt and the coarse-operator build all walk
// these panels).  Ef = 12 (spin-colour) on a fine Dirac field, Ef = N on a
// coarse field; the chirality of element e is e / (Ef/2) (spin / 2 for a
// Dirac field, the coarse chirality n / nb for a coarse one).
struct mgc_aggregation {
  mgc::Lat4 fine, coarse, blk;
  int Vf = 0, Vc = 0, Vb = 0, nb = 0;
  int Ef = 12;
  mgc::cplx<double>* basis64 = nullptr;
  mgc::cplx<float>* basis32 = nullptr;
  bool basis_set = false;
  bool basis32_valid = false;
  int* blk_fine = nullptr;  // [Vc][Vb] fine site of (block, block site)
  int* fine_row = nullptr;  // [Vf] panel row (xc*Vb + xb) of a fine site
  void* part = nullptr;     // partial sums of split projections (grown on demand)
  size_t part_bytes = 0;
  // Cholesky-QR orthonormalisation: per (block, chirality) panel the nb x nb
  // transform T and a breakdown flag (allocated on the first call)
  mgc::cplx<double>* cq_T = nullptr;
  int* cq_flag = nullptr;
};

namespace mgc {
// FP32 copy of the basis, refreshed when stale
const cplx<float>* agg_basis32(mgc_aggregation* a, hipStream_t s);
// project (mask kind 0 all / 1 inner of point / 2 outer of point) or promote
void agg_transfer(mgc_aggregation* a, const void* in, void* out, int prec, int nrhs,
                  bool project, int kind, int point, hipStream_t s);
// fine field [Vf][Ef] holding the chirality-l part of basis vector i (l = -1: all)
void agg_extract_basis(mgc_aggregation* a, int i, int l, void* out, int prec, hipStream_t s);
// FP64 fine fields [Vf][nr][Ef] of the doubled basis columns n0 .. n0+nr-1
// (column n = chirality-(n / nb) part of basis vector n % nb)
void agg_extract_doubled(mgc_aggregation* a, int n0, int nr, void* out, hipStream_t s);
__host__ __device__ inline size_t agg_row(int nb, int Vb, int Ef, int xc, int xb, int e) {
  return (((size_t)xc * Vb + xb) * Ef + e) * nb;
}
}  // namespace mgc
