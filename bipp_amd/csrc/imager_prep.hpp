// Parameter blocks and launchers of the eigenpair imager's preparation kernels (imager_prep.hip).
#pragma once

#include "common.hpp"

namespace bippamd {

template <typename R>
struct EigPairPrep {
  int n_ant, n_beam, n_img;
  const cplx<float>* v_in;  // [nS][n_ant x n_beam] complex64, column-major
  const float* d_in;        // [nS][n_img][n_beam] scaled eigenvalues of each image
  const float* scale;       // [nS]
  const float* wl;          // [nS]
  cplx<R>* v_out;           // [nS][n_ant x n_beam]
  R* w_out;                 // [nS][n_img][n_beam]
  double* uvw_scale;        // [nS] 2 pi / wl
};

template <typename R>
void launch_eigpair_prep(const EigPairPrep<R>& p, int n_samples, hipStream_t st);
template <typename R>
void launch_scale_uvw(long long n, const float* uvw, long long lduvw, double scale, R* u, R* v,
                      R* w, hipStream_t st);

}  // namespace bippamd
