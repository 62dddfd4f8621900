// Stage-B preparation kernels of the eigenpair imager: eigenvectors and scaled eigenvalues as
// a dataset stores them (complex64 / float32, src/dataset_file.cpp:204-249) brought into the
// compute precision, and the per-sample uvw scale 2 pi / wl (src/nufft_synthesis.cpp:149-155).
#include "common.hpp"
#include "imager_prep.hpp"

namespace bippamd {

namespace {

// One workgroup per sample: V (nAnt x nBeam, complex64, ld n_ant) -> complex<R>;
// w[img][e] = scale * dScaled[img][e] (virtual_vis scales the selected eigenvalues by the
// sample's scale, src/host/virtual_vis.cpp:52-57); uvw scale 2 pi / wl.
template <typename R>
__global__ __launch_bounds__(256) void eigpair_prep_kernel(EigPairPrep<R> p) {
  const int s = blockIdx.x;
  const long long nv = (long long)p.n_ant * p.n_beam;
  const cplx<float>* vi = p.v_in + (size_t)s * nv;
  cplx<R>* vo = p.v_out + (size_t)s * nv;
  for (long long i = threadIdx.x; i < nv; i += blockDim.x) vo[i] = cvt<R>(vi[i]);
  const R sc = static_cast<R>(p.scale[s]);
  const float* di = p.d_in + (size_t)s * p.n_img * p.n_beam;
  R* wo = p.w_out + (size_t)s * p.n_img * p.n_beam;
  for (int i = threadIdx.x; i < p.n_img * p.n_beam; i += blockDim.x) wo[i] = sc * static_cast<R>(di[i]);
  if (threadIdx.x == 0)
    p.uvw_scale[s] = 2.0 * 3.14159265358979323846 / static_cast<double>(p.wl[s]);
}

// u, v, w (R) = uvw (float, metres) * 2 pi / wl of one sample, for the direct sum.
template <typename R>
__global__ void scale_uvw_kernel(long long n, const float* uvw, long long lduvw, double scale,
                                 R* u, R* v, R* w) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  u[i] = static_cast<R>(scale * static_cast<double>(uvw[i]));
  v[i] = static_cast<R>(scale * static_cast<double>(uvw[i + lduvw]));
  w[i] = static_cast<R>(scale * static_cast<double>(uvw[i + 2 * lduvw]));
}

}  // namespace

template <typename R>
void launch_eigpair_prep(const EigPairPrep<R>& p, int n_samples, hipStream_t st) {
  if (n_samples <= 0) return;
  eigpair_prep_kernel<R><<<n_samples, 256, 0, st>>>(p);
  check_launch("eigpair_prep_kernel");
}
template void launch_eigpair_prep<float>(const EigPairPrep<float>&, int, hipStream_t);
template void launch_eigpair_prep<double>(const EigPairPrep<double>&, int, hipStream_t);

template <typename R>
void launch_scale_uvw(long long n, const float* uvw, long long lduvw, double scale, R* u, R* v,
                      R* w, hipStream_t st) {
  if (n <= 0) return;
  scale_uvw_kernel<R><<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(n, uvw, lduvw, scale,
                                                                             u, v, w);
  check_launch("scale_uvw_kernel");
}
template void launch_scale_uvw<float>(long long, const float*, long long, double, float*, float*,
                                      float*, hipStream_t);
template void launch_scale_uvw<double>(long long, const float*, long long, double, double*,
                                       double*, double*, hipStream_t);

}  // namespace bippamd
