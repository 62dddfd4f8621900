// Virtual visibilities, the direct NUFT sum, and the small batched complex GEMMs used by
// the general-beamforming (W != I) path.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace bippamd {

namespace {

// --------------------------------------------------------------------------------------
// VV = V diag(dw) V^H per sample, column-major n_ant x n_ant, entry (i,j) at i + n_ant j
// (src/host/virtual_vis.cpp:19-73).  16x16 output tiles, eigenvector index staged in LDS.
// --------------------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void virtual_vis_kernel(int n_ant, int n_cols, const cplx<R>* V,
                                                           long long ldv, long long stride_v,
                                                           const R* dw, long long stride_dw,
                                                           cplx<R>* vv, long long stride_vv) {
  __shared__ cplx<R> ti[16][17];
  __shared__ cplx<R> tj[16][17];
  const int s = blockIdx.z;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const cplx<R>* Vs = V + (size_t)s * stride_v;
  const R* d = dw + (size_t)s * stride_dw;
  cplx<R> acc = mk<R>(0, 0);
  for (int e0 = 0; e0 < n_cols; e0 += 16) {
    const int e = e0 + ty;
    const int i = i0 + tx, j = j0 + tx;
    const R de = e < n_cols ? d[e] : R(0);
    ti[ty][tx] = (e < n_cols && i < n_ant) ? Vs[i + (size_t)e * ldv] * de : mk<R>(0, 0);
    tj[ty][tx] = (e < n_cols && j < n_ant) ? Vs[j + (size_t)e * ldv] : mk<R>(0, 0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = acc + mulc(ti[k][tx], tj[k][ty]);
    __syncthreads();
  }
  const int i = i0 + tx, j = j0 + ty;
  if (i < n_ant && j < n_ant) vv[(size_t)s * stride_vv + i + (size_t)j * n_ant] = acc;
}

// --------------------------------------------------------------------------------------
// Direct NUFT sum (src/host/kernels/nuft_sum.cpp:9-31):
//   out[q] += sum_j Re(c_j exp(i alpha (x_q u_j + y_q v_j + z_q w_j)))
// One thread per output point, input points staged through LDS.
// --------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void sincos_r(R x, R* s, R* c);
template <>
__device__ __forceinline__ void sincos_r<double>(double x, double* s, double* c) {
  sincos(x, s, c);
}
template <>
__device__ __forceinline__ void sincos_r<float>(float x, float* s, float* c) {
  sincosf(x, s, c);
}

constexpr int NUFT_CHUNK = 256;

template <typename R>
__global__ __launch_bounds__(256) void nuft_sum_kernel(R alpha, int n_in, const cplx<R>* in,
                                                        const R* u, const R* v, const R* w,
                                                        int n_out, const R* x, const R* y,
                                                        const R* z, R* out) {
  __shared__ R su[NUFT_CHUNK], sv[NUFT_CHUNK], sw[NUFT_CHUNK];
  __shared__ cplx<R> sc[NUFT_CHUNK];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const R xq = q < n_out ? x[q] : R(0), yq = q < n_out ? y[q] : R(0), zq = q < n_out ? z[q] : R(0);
  R sum = 0;
  for (int j0 = 0; j0 < n_in; j0 += NUFT_CHUNK) {
    const int j = j0 + threadIdx.x;
    if (j < n_in) {
      su[threadIdx.x] = u[j];
      sv[threadIdx.x] = v[j];
      sw[threadIdx.x] = w[j];
      sc[threadIdx.x] = in[j];
    }
    __syncthreads();
    const int m = min(NUFT_CHUNK, n_in - j0);
    for (int k = 0; k < m; ++k) {
      const R ph = alpha * (xq * su[k] + yq * sv[k] + zq * sw[k]);
      R s, c;
      sincos_r<R>(ph, &s, &c);
      sum += sc[k].re * c - sc[k].im * s;
    }
    __syncthreads();
  }
  if (q < n_out) out[q] += sum;
}

// Mixed form: phase alpha*(x.u) formed in float64 and reduced to one turn, float32 sin/cos
// and accumulation.
__global__ __launch_bounds__(256) void nuft_sum_mixed_kernel(double alpha_turns, int n_in,
                                                              const cplx<float>* in,
                                                              const double* u, const double* v,
                                                              const double* w, int n_out,
                                                              const double* x, const double* y,
                                                              const double* z, float* out) {
  __shared__ double su[NUFT_CHUNK], sv[NUFT_CHUNK], sw[NUFT_CHUNK];
  __shared__ cplx<float> sc[NUFT_CHUNK];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const double xq = q < n_out ? x[q] * alpha_turns : 0.0;
  const double yq = q < n_out ? y[q] * alpha_turns : 0.0;
  const double zq = q < n_out ? z[q] * alpha_turns : 0.0;
  float sum = 0;
  for (int j0 = 0; j0 < n_in; j0 += NUFT_CHUNK) {
    const int j = j0 + threadIdx.x;
    if (j < n_in) {
      su[threadIdx.x] = u[j];
      sv[threadIdx.x] = v[j];
      sw[threadIdx.x] = w[j];
      sc[threadIdx.x] = in[j];
    }
    __syncthreads();
    const int m = min(NUFT_CHUNK, n_in - j0);
    for (int k = 0; k < m; ++k) {
      const double t = fma(xq, su[k], fma(yq, sv[k], zq * sw[k]));
      const float f = static_cast<float>(t - rint(t));
      float s, c;
      sincospif(2.0f * f, &s, &c);
      sum += sc[k].re * c - sc[k].im * s;
    }
    __syncthreads();
  }
  if (q < n_out) out[q] += sum;
}

// --------------------------------------------------------------------------------------
// Batched ragged complex GEMM, 16x16 tiles (small sizes; W != I path only).
// --------------------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void gemm_batch_kernel(GemmBatch<R> g) {
  __shared__ cplx<R> ta[16][17];
  __shared__ cplx<R> tb[16][17];
  const int bidx = blockIdx.z;
  const int m = g.m ? g.m[bidx] : g.m0;
  const int n = g.n ? g.n[bidx] : g.n0;
  const int k = g.k ? g.k[bidx] : g.k0;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  if (i0 >= m || j0 >= n) return;  // uniform per workgroup
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const cplx<R>* A = g.a + (size_t)bidx * g.stride_a;
  const cplx<R>* B = g.b + (size_t)bidx * g.stride_b;
  cplx<R>* C = g.c + (size_t)bidx * g.stride_c;
  cplx<R> acc = mk<R>(0, 0);
  for (int l0 = 0; l0 < k; l0 += 16) {
    // ta[l][i] = op(A)[i0+i, l0+l], tb[l][j] = op(B)[l0+l, j0+j]
    {
      const int i = i0 + tx, l = l0 + ty;
      cplx<R> a = mk<R>(0, 0);
      if (i < m && l < k) a = g.op_a ? conj(A[l + (size_t)i * g.lda]) : A[i + (size_t)l * g.lda];
      ta[ty][tx] = a;
      const int j = j0 + tx;
      cplx<R> b = mk<R>(0, 0);
      if (j < n && l < k) b = g.op_b ? conj(B[j + (size_t)l * g.ldb]) : B[l + (size_t)j * g.ldb];
      tb[ty][tx] = b;
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 16; ++l) acc = acc + ta[l][tx] * tb[l][ty];
    __syncthreads();
  }
  const int i = i0 + tx, j = j0 + ty;
  if (i < m && j < n) {
    cplx<R>& c = C[i + (size_t)j * g.ldc];
    const cplx<R> old = (g.beta.re == R(0) && g.beta.im == R(0)) ? mk<R>(0, 0) : c * g.beta;
    c = old + g.alpha * acc;
  }
}

template <typename R>
__global__ void sinc_matrix_kernel(int n_ant, const double* xyz, long long ldxyz,
                                   long long stride_xyz, const double* wl, double pref,
                                   cplx<R>* out, long long ld, long long stride) {
  const int s = blockIdx.z;
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= n_ant || j >= n_ant) return;
  const double* p = xyz + (size_t)s * stride_xyz;
  R val;
  if (i == j) {
    val = static_cast<R>(pref);
  } else {
    const R dx = static_cast<R>(p[i]) - static_cast<R>(p[j]);
    const R dy = static_cast<R>(p[i + ldxyz]) - static_cast<R>(p[j + ldxyz]);
    const R dz = static_cast<R>(p[i + 2 * ldxyz]) - static_cast<R>(p[j + 2 * ldxyz]);
    const R x = R(2) * sqrt(dx * dx + dy * dy + dz * dz) / static_cast<R>(wl[s]);
    R sc;
    if constexpr (sizeof(R) == 8)
      sc = x != R(0) ? sinpi(x) / (R(3.14159265358979323846) * x) : R(1);
    else
      sc = x != R(0) ? sinpif(x) / (R(3.14159265358979323846) * x) : R(1);
    val = static_cast<R>(pref) * sc;
  }
  out[(size_t)s * stride + i + (size_t)j * ld] = mk<R>(val, R(0));
}

template <typename R, typename TIn>
__global__ void gather_w_kernel(int n_ant, int n_beam, const cplx<TIn>* w, long long ldw,
                                long long stride_w, const int* n_red, const int* idx,
                                long long idx_stride, cplx<R>* out, long long ld, long long stride) {
  const int s = blockIdx.y;
  const int c = blockIdx.x;
  const int nr = n_red[s];
  cplx<R>* o = out + (size_t)s * stride + (size_t)c * ld;
  if (c >= nr) {
    for (int i = threadIdx.x; i < n_ant; i += blockDim.x) o[i] = mk<R>(0, 0);
    return;
  }
  const int b = idx[(size_t)s * idx_stride + c];
  const cplx<TIn>* src = w + (size_t)s * stride_w + (size_t)b * ldw;
  for (int i = threadIdx.x; i < n_ant; i += blockDim.x) o[i] = cvt<R>(src[i]);
}

// uvw[i + N j] = (2 pi / wl) (p_i - p_j) for one sample (python/bipp/instrument.py:246-251,
// src/nufft_synthesis.cpp:149-155), as three planes of N^2 entries.
template <typename R>
__global__ void baselines_kernel(int n_ant, const double* xyz, long long ldxyz, double scale,
                                 R* u, R* v, R* w) {
  const long long nn = (long long)n_ant * n_ant;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < nn;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = static_cast<int>(t % n_ant), j = static_cast<int>(t / n_ant);
    u[t] = static_cast<R>(scale * (xyz[i] - xyz[j]));
    v[t] = static_cast<R>(scale * (xyz[i + ldxyz] - xyz[j + ldxyz]));
    w[t] = static_cast<R>(scale * (xyz[i + 2 * ldxyz] - xyz[j + 2 * ldxyz]));
  }
}

template <typename TS, typename TD>
__global__ void copy_cast_kernel(int rows, int cols, const cplx<TS>* src, long long lds,
                                 long long stride_s, cplx<TD>* dst, long long ldd,
                                 long long stride_d) {
  const int s = blockIdx.y;
  const long long n = (long long)rows * cols;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = static_cast<int>(t % rows), j = static_cast<int>(t / rows);
    dst[(size_t)s * stride_d + i + (size_t)j * ldd] =
        cvt<TD>(src[(size_t)s * stride_s + i + (size_t)j * lds]);
  }
}

}  // namespace

template <typename R>
void launch_virtual_vis(int n_samples, int n_ant, int n_eig_cols, const cplx<R>* v, long long ldv,
                        long long stride_v, const R* dw, long long stride_dw, cplx<R>* vv,
                        long long stride_vv, hipStream_t stream) {
  if (n_samples <= 0 || n_ant <= 0) return;
  dim3 grid(ceil_div(n_ant, 16), ceil_div(n_ant, 16), n_samples);
  virtual_vis_kernel<R><<<grid, 256, 0, stream>>>(n_ant, n_eig_cols, v, ldv, stride_v, dw,
                                                  stride_dw, vv, stride_vv);
  check_launch("virtual_vis_kernel");
}

template <typename R>
void launch_nuft_sum(R alpha, int n_in, const cplx<R>* input, const R* u, const R* v, const R* w,
                     int n_out, const R* x, const R* y, const R* z, R* out, hipStream_t stream) {
  if (n_out <= 0) return;
  nuft_sum_kernel<R><<<ceil_div(n_out, 256), 256, 0, stream>>>(alpha, n_in, input, u, v, w, n_out,
                                                               x, y, z, out);
  check_launch("nuft_sum_kernel");
}

void launch_nuft_sum_mixed(double alpha, int n_in, const cplx<float>* input, const double* u,
                           const double* v, const double* w, int n_out, const double* x,
                           const double* y, const double* z, float* out, hipStream_t stream) {
  if (n_out <= 0) return;
  const double alpha_turns = alpha / (2.0 * 3.14159265358979323846);
  nuft_sum_mixed_kernel<<<ceil_div(n_out, 256), 256, 0, stream>>>(alpha_turns, n_in, input, u, v, w,
                                                                  n_out, x, y, z, out);
  check_launch("nuft_sum_mixed_kernel");
}

template <typename R>
void launch_gemm_batch(const GemmBatch<R>& g, int max_m, int max_n, hipStream_t stream) {
  if (g.n_batch <= 0 || max_m <= 0 || max_n <= 0) return;
  dim3 grid(ceil_div(max_m, 16), ceil_div(max_n, 16), g.n_batch);
  gemm_batch_kernel<R><<<grid, 256, 0, stream>>>(g);
  check_launch("gemm_batch_kernel");
}

template <typename R>
void launch_sinc_matrix(int n_samples, int n_ant, const double* xyz, long long ldxyz,
                        long long stride_xyz, const double* wl, double prefactor, cplx<R>* out,
                        long long ld, long long stride, hipStream_t stream) {
  if (n_samples <= 0 || n_ant <= 0) return;
  dim3 grid(ceil_div(n_ant, 16), ceil_div(n_ant, 16), n_samples);
  sinc_matrix_kernel<R><<<grid, 256, 0, stream>>>(n_ant, xyz, ldxyz, stride_xyz, wl, prefactor,
                                                  out, ld, stride);
  check_launch("sinc_matrix_kernel");
}

template <typename R, typename TIn>
void launch_gather_w(int n_samples, int n_ant, int n_beam, const void* w, long long ldw,
                     long long stride_w, const int* n_red, const int* idx, long long idx_stride,
                     cplx<R>* out, long long ld, long long stride, hipStream_t stream) {
  if (n_samples <= 0 || n_beam <= 0) return;
  dim3 grid(n_beam, n_samples);
  gather_w_kernel<R, TIn><<<grid, 256, 0, stream>>>(n_ant, n_beam,
                                                    static_cast<const cplx<TIn>*>(w), ldw, stride_w,
                                                    n_red, idx, idx_stride, out, ld, stride);
  check_launch("gather_w_kernel");
}

template <typename R>
void launch_baselines(int n_ant, const double* xyz, long long ldxyz, double scale, R* u, R* v, R* w,
                      hipStream_t stream) {
  const long long nn = (long long)n_ant * n_ant;
  if (nn <= 0) return;
  baselines_kernel<R><<<static_cast<int>(std::min<long long>((nn + 255) / 256, 65535)), 256, 0,
                        stream>>>(n_ant, xyz, ldxyz, scale, u, v, w);
  check_launch("baselines_kernel");
}
template void launch_baselines<float>(int, const double*, long long, double, float*, float*,
                                      float*, hipStream_t);
template void launch_baselines<double>(int, const double*, long long, double, double*, double*,
                                       double*, hipStream_t);

template <typename TS, typename TD>
void launch_copy_cast(int n_batch, int rows, int cols, const void* src, long long lds,
                      long long stride_s, void* dst, long long ldd, long long stride_d,
                      hipStream_t stream) {
  const long long n = (long long)rows * cols;
  if (n_batch <= 0 || n <= 0) return;
  dim3 grid(static_cast<unsigned>(std::min<long long>((n + 255) / 256, 4096)), n_batch);
  copy_cast_kernel<TS, TD><<<grid, 256, 0, stream>>>(rows, cols, static_cast<const cplx<TS>*>(src),
                                                     lds, stride_s, static_cast<cplx<TD>*>(dst),
                                                     ldd, stride_d);
  check_launch("copy_cast_kernel");
}
template void launch_copy_cast<float, float>(int, int, int, const void*, long long, long long,
                                             void*, long long, long long, hipStream_t);
template void launch_copy_cast<float, double>(int, int, int, const void*, long long, long long,
                                              void*, long long, long long, hipStream_t);
template void launch_copy_cast<double, float>(int, int, int, const void*, long long, long long,
                                              void*, long long, long long, hipStream_t);
template void launch_copy_cast<double, double>(int, int, int, const void*, long long, long long,
                                               void*, long long, long long, hipStream_t);

template void launch_virtual_vis<float>(int, int, int, const cplx<float>*, long long, long long,
                                        const float*, long long, cplx<float>*, long long,
                                        hipStream_t);
template void launch_virtual_vis<double>(int, int, int, const cplx<double>*, long long, long long,
                                         const double*, long long, cplx<double>*, long long,
                                         hipStream_t);
template void launch_nuft_sum<float>(float, int, const cplx<float>*, const float*, const float*,
                                     const float*, int, const float*, const float*, const float*,
                                     float*, hipStream_t);
template void launch_nuft_sum<double>(double, int, const cplx<double>*, const double*,
                                      const double*, const double*, int, const double*,
                                      const double*, const double*, double*, hipStream_t);
template void launch_gemm_batch<float>(const GemmBatch<float>&, int, int, hipStream_t);
template void launch_gemm_batch<double>(const GemmBatch<double>&, int, int, hipStream_t);
template void launch_sinc_matrix<float>(int, int, const double*, long long, long long,
                                        const double*, double, cplx<float>*, long long, long long,
                                        hipStream_t);
template void launch_sinc_matrix<double>(int, int, const double*, long long, long long,
                                         const double*, double, cplx<double>*, long long,
                                         long long, hipStream_t);
template void launch_gather_w<float, float>(int, int, int, const void*, long long, long long,
                                            const int*, const int*, long long, cplx<float>*,
                                            long long, long long, hipStream_t);
template void launch_gather_w<float, double>(int, int, int, const void*, long long, long long,
                                             const int*, const int*, long long, cplx<float>*,
                                             long long, long long, hipStream_t);
template void launch_gather_w<double, float>(int, int, int, const void*, long long, long long,
                                             const int*, const int*, long long, cplx<double>*,
                                             long long, long long, hipStream_t);
template void launch_gather_w<double, double>(int, int, int, const void*, long long, long long,
                                              const int*, const int*, long long, cplx<double>*,
                                              long long, long long, hipStream_t);

}  // namespace bippamd
