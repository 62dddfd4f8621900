// Type-3 NUFFT (3-D, non-uniform baselines -> non-uniform pixels) for the Bluebild
// image sum  f_l(x_q) = sum_j c_lj exp(+i u_j . x_q),  image_l += Re f_l.
//
// Replaces the neonufft PlanT3 calls of src/host/nufft.hpp:196-216 and
// src/gpu/nufft.hpp:198-216 (sign +1, real part accumulated into a float image).  The
// algorithm is the published two-stage type-3 scheme (Lee & Greengard 2005; Barnett,
// Magland & af Klinteberg 2019) with the "exponential of semicircle" kernel
// phi(z) = exp(beta (sqrt(1 - z^2) - 1)), restated per dimension as:
//   centre   u = C + u',  x = D + x';   c'_j = c_j exp(i u'_j . D)
//   rescale  t_j = u'_j / gamma,  s_q = gamma x'_q,  gamma = nf1 / (2 sigma1 S)
//   spread   b_l = sum_j c'_j psi(l h1 - t_j)                 (grid of nf1, h1 = 2 pi / nf1)
//   type 2   F(s) = sum_l b_l exp(i l h1 s)                     (deconvolve by phi2^, FFT on
//            a grid of nf2 = sigma2 nf1, interpolate at h1 s)
//   correct  f(x_q) = exp(i C . x_q) h1 F(s_q) / psi^(s_q)       (per dimension)
// The same points serve every energy level: one spread pass accumulates all levels, the
// kernel weights of a point are computed once.
//
// Data layout in HBM: grid[level][z][y][x] complex (x fastest), nf2 per axis.  Axes longer
// than 128 that are not contiguous are transformed in two passes (n = n1 n2) and keep
// their frequencies in the permuted order  position = (k mod n1) n2 + k / n1,  which the
// interpolation reads directly (no transpose passes).
#include "common.hpp"
#include "kernels.hpp"
#include "nufft3.hpp"

#include <algorithm>
#include <cmath>

namespace bippamd {

namespace {

constexpr int SP_NT = 256;  // spread workgroup

template <typename R>
__device__ __forceinline__ R es_kernel(R z, R beta) {
  const R t = R(1) - z * z;
  return t > R(0) ? exp(beta * (sqrt(t) - R(1))) : R(0);
}

// 2 int_0^1 phi(z) cos(xi z) dz by Gauss-Legendre (nodes/weights on [0, 1])
__device__ double es_ft(double xi, double beta, const double* gl_x, const double* gl_w, int nq) {
  double s = 0;
  for (int i = 0; i < nq; ++i) {
    const double z = gl_x[i];
    s += gl_w[i] * exp(beta * (sqrt(1.0 - z * z) - 1.0)) * cos(xi * z);
  }
  return 2.0 * s;
}

// --------------------------------------------------------------------------------------
// Point preparation: one thread per (sample, antenna pair a >= b).
// --------------------------------------------------------------------------------------
template <typename R>
__global__ void n3_points_kernel(N3PointsParams p) {
  const int s = blockIdx.y;
  const long long npair = (long long)p.n_ant_max * (p.n_ant_max + 1) / 2;
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= npair) return;
  const int n = p.n_rows ? p.n_rows[s] : p.n_rows_const;
  const int a = static_cast<int>(floor((sqrt(8.0 * t + 1.0) - 1.0) * 0.5));
  int a2 = a;
  // guard rounding of the square root
  if ((long long)(a2 + 1) * (a2 + 2) / 2 <= t) ++a2;
  if ((long long)a2 * (a2 + 1) / 2 > t) --a2;
  const int b = static_cast<int>(t - (long long)a2 * (a2 + 1) / 2);
  const long long pid = (long long)s * npair + t;
  if (a2 >= n) {
    p.valid[pid] = 0;
    return;
  }
  const double* pos = p.pos + (size_t)s * p.pos_stride;  // [3][apad] positions / wavelength
  double u[3], cell[3];
  int l0[3], bin[3];
  float fr[3];
  double ph = 0;
  for (int d = 0; d < 3; ++d) {
    u[d] = 6.283185307179586 * (pos[d * p.apad + a2] - pos[d * p.apad + b]);
    const double up = u[d] - p.C[d];
    ph += up * p.D[d];
    cell[d] = up * p.inv_gh[d];  // t / h1 in grid units
    if (p.phase_fp32) cell[d] = static_cast<double>(static_cast<float>(up) * static_cast<float>(p.inv_gh[d]));
    l0[d] = static_cast<int>(ceil(cell[d] - 0.5 * p.w));
    fr[d] = static_cast<float>(cell[d] - l0[d]);
    bin[d] = (l0[d] + p.off[d]) / p.bin_size[d];
  }
  double ps, pc;
  if (p.phase_fp32) {
    float fs, fc;
    const float fph = static_cast<float>(u[0] - p.C[0]) * static_cast<float>(p.D[0]) +
                      static_cast<float>(u[1] - p.C[1]) * static_cast<float>(p.D[1]) +
                      static_cast<float>(u[2] - p.C[2]) * static_cast<float>(p.D[2]);
    sincosf(fph, &fs, &fc);
    ps = fs;
    pc = fc;
  } else {
    const double turns = ph * 0.15915494309189535;
    const double f = turns - rint(turns);
    sincospi(2.0 * f, &ps, &pc);
  }
  const int binid = (bin[2] * p.nbins[1] + bin[1]) * p.nbins[0] + bin[0];
  p.valid[pid] = 1;
  p.bin_of[pid] = binid;
  p.l0x[pid] = l0[0];
  p.l0y[pid] = l0[1];
  p.l0z[pid] = l0[2];
  p.frx[pid] = fr[0];
  p.fry[pid] = fr[1];
  p.frz[pid] = fr[2];
  p.phr[pid] = static_cast<float>(pc);
  p.phi[pid] = static_cast<float>(ps);
  p.src[pid] = a2 + p.ld_vv * b;  // VV(a, b) inside the sample's matrix
  p.mult[pid] = a2 == b ? 1.f : 2.f;  // (a,b) and (b,a) are conjugate: 2 Re(.)
  atomicAdd(&p.bin_count[binid], 1u);
}

// exclusive scan helpers (counts -> starts)
__global__ void scan_block_kernel(const unsigned* in, unsigned* out, unsigned* block_sums, int n) {
  __shared__ unsigned sh[1024];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const unsigned v = i < n ? in[i] : 0u;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == 1023) block_sums[blockIdx.x] = sh[1023];
}

__global__ void scan_add_kernel(unsigned* out, const unsigned* block_offsets, int n) {
  const int i = blockIdx.x * 1024 + threadIdx.x;
  if (i < n) out[i] += block_offsets[blockIdx.x];
}

__global__ void n3_scatter_kernel(long long n_pts, const unsigned char* valid, const int* bin_of,
                                  const unsigned* bin_start, unsigned* bin_fill, unsigned* sorted) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n_pts || !valid[i]) return;
  const int b = bin_of[i];
  const unsigned slot = bin_start[b] + atomicAdd(&bin_fill[b], 1u);
  sorted[slot] = static_cast<unsigned>(i);
}

// --------------------------------------------------------------------------------------
// Spreading: one workgroup per bin, LDS tile (bin + kernel width) for every level.
// Threads own kernel offsets (ox, oy, oz); points are taken in batches whose kernel
// weights are computed once into LDS.
// --------------------------------------------------------------------------------------
template <typename R, typename TV>
__global__ __launch_bounds__(SP_NT) void n3_spread_kernel(N3SpreadParams<R> p) {
  extern __shared__ unsigned char sp_sh[];
  const int bin = blockIdx.x;
  const unsigned beg = p.bin_start[bin];
  const unsigned cnt = p.bin_count[bin];
  if (cnt == 0) return;
  const int w = p.w;
  const int TX = p.bin_size[0] + w, TY = p.bin_size[1] + w, TZ = p.bin_size[2] + w;
  const int tcells = TX * TY * TZ;
  const int L = p.n_levels;
  R* tile = reinterpret_cast<R*>(sp_sh);                     // [L][tcells][2]
  constexpr int PB = 32;                                     // points per batch
  R* wts = tile + (size_t)2 * L * tcells;                    // [PB][3][w]
  R* val = wts + (size_t)PB * 3 * w;                         // [PB][L][2]
  int* org = reinterpret_cast<int*>(val + (size_t)PB * L * 2);  // [PB][3] tile-local origin
  const int bx = bin % p.nbins[0];
  const int by = (bin / p.nbins[0]) % p.nbins[1];
  const int bz = bin / (p.nbins[0] * p.nbins[1]);
  const int ox0 = bx * p.bin_size[0], oy0 = by * p.bin_size[1], oz0 = bz * p.bin_size[2];
  for (int i = threadIdx.x; i < 2 * L * tcells; i += SP_NT) tile[i] = R(0);
  const int w3 = w * w * w;
  const R two_over_w = R(2) / R(w);
  __syncthreads();
  for (unsigned b0 = 0; b0 < cnt; b0 += PB) {
    const int nb = min<unsigned>(PB, cnt - b0);
    // kernel weights and values of this batch
    for (int t = threadIdx.x; t < nb * 3 * w; t += SP_NT) {
      const int pi = t / (3 * w), rem = t % (3 * w), d = rem / w, o = rem % w;
      const unsigned id = p.sorted[beg + b0 + pi];
      const float fr = d == 0 ? p.frx[id] : (d == 1 ? p.fry[id] : p.frz[id]);
      wts[(pi * 3 + d) * w + o] = es_kernel<R>((R(o) - R(fr)) * two_over_w, R(p.beta));
    }
    for (int t = threadIdx.x; t < nb * L; t += SP_NT) {
      const int pi = t / L, lv = t % L;
      const unsigned id = p.sorted[beg + b0 + pi];
      const long long s = id / p.npair;
      const TV* vv = p.vv + ((size_t)s * L + lv) * p.vv_stride;
      const TV c = vv[p.src[id]];
      const R m = R(p.mult[id]);
      const R pr = R(p.phr[id]), pim = R(p.phi[id]);
      val[(pi * L + lv) * 2 + 0] = m * (R(c.re) * pr - R(c.im) * pim);
      val[(pi * L + lv) * 2 + 1] = m * (R(c.re) * pim + R(c.im) * pr);
    }
    for (int t = threadIdx.x; t < nb; t += SP_NT) {
      const unsigned id = p.sorted[beg + b0 + t];
      org[t * 3 + 0] = p.l0x[id] + p.off[0] - ox0;
      org[t * 3 + 1] = p.l0y[id] + p.off[1] - oy0;
      org[t * 3 + 2] = p.l0z[id] + p.off[2] - oz0;
    }
    __syncthreads();
    // work item = (point, kernel offset); consecutive lanes take consecutive offsets of one
    // point, i.e. distinct tile cells
    for (int it = threadIdx.x; it < nb * w3; it += SP_NT) {
      const int pi = it / w3, off = it % w3;
      const int oxo = off % w, oyo = (off / w) % w, ozo = off / (w * w);
      const R wt = wts[(pi * 3 + 0) * w + oxo] * wts[(pi * 3 + 1) * w + oyo] *
                   wts[(pi * 3 + 2) * w + ozo];
      const int cx = org[pi * 3 + 0] + oxo, cy = org[pi * 3 + 1] + oyo, cz = org[pi * 3 + 2] + ozo;
      const int cell = (cz * TY + cy) * TX + cx;
      for (int lv = 0; lv < L; ++lv) {
        atomicAdd(&tile[((size_t)lv * tcells + cell) * 2 + 0], wt * val[(pi * L + lv) * 2 + 0]);
        atomicAdd(&tile[((size_t)lv * tcells + cell) * 2 + 1], wt * val[(pi * L + lv) * 2 + 1]);
      }
    }
    __syncthreads();
  }
  // flush the tile into the nf2 grid (spread-grid index l -> l mod nf2)
  for (int i = threadIdx.x; i < L * tcells; i += SP_NT) {
    const int lv = i / tcells, cell = i % tcells;
    const R re = tile[(size_t)i * 2], im = tile[(size_t)i * 2 + 1];
    if (re == R(0) && im == R(0)) continue;
    const int cx = cell % TX, cy = (cell / TX) % TY, cz = cell / (TX * TY);
    int gx = ox0 + cx - p.off[0], gy = oy0 + cy - p.off[1], gz = oz0 + cz - p.off[2];
    gx = ((gx % p.nf2[0]) + p.nf2[0]) % p.nf2[0];
    gy = ((gy % p.nf2[1]) + p.nf2[1]) % p.nf2[1];
    gz = ((gz % p.nf2[2]) + p.nf2[2]) % p.nf2[2];
    R* g = p.grid + ((size_t)lv * p.grid_stride + ((size_t)gz * p.nf2[1] + gy) * p.nf2[0] + gx) * 2;
    atomicAdd(g, re);
    atomicAdd(g + 1, im);
  }
}

// --------------------------------------------------------------------------------------
// Deconvolution of the spread modes l in [-nf1/2, nf1/2) by 1/phi2^(l).
// --------------------------------------------------------------------------------------
template <typename R>
__global__ void n3_deconv_kernel(N3DeconvParams<R> p) {
  const long long n = (long long)p.nf1[0] * p.nf1[1] * p.nf1[2];
  const int lv = blockIdx.y;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int ix = static_cast<int>(t % p.nf1[0]);
    const int iy = static_cast<int>((t / p.nf1[0]) % p.nf1[1]);
    const int iz = static_cast<int>(t / ((long long)p.nf1[0] * p.nf1[1]));
    const int lx = ix - p.nf1[0] / 2, ly = iy - p.nf1[1] / 2, lz = iz - p.nf1[2] / 2;
    const R f = p.fx[ix] * p.fy[iy] * p.fz[iz];
    const int gx = (lx + p.nf2[0]) % p.nf2[0], gy = (ly + p.nf2[1]) % p.nf2[1],
              gz = (lz + p.nf2[2]) % p.nf2[2];
    R* g = p.grid + ((size_t)lv * p.grid_stride + ((size_t)gz * p.nf2[1] + gy) * p.nf2[0] + gx) * 2;
    g[0] *= f;
    g[1] *= f;
  }
}

// --------------------------------------------------------------------------------------
// Stockham FFT in LDS (sign +): `nl` lines of length n, line stride ld, radices r[].
// Returns the buffer holding the result.
// --------------------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ void dft_small(cplx<R>* v, int rad) {
  if (rad == 2) {
    const cplx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if (rad == 4) {
    const cplx<R> a = v[0], b = v[1], c = v[2], d = v[3];
    const cplx<R> s0 = a + c, s1 = a - c, s2 = b + d, s3 = b - d;
    const cplx<R> is3 = mk<R>(-s3.im, s3.re);  // +i * s3
    v[0] = s0 + s2;
    v[1] = s1 + is3;
    v[2] = s0 - s2;
    v[3] = s1 - is3;
  } else {
    cplx<R> o[5];
    for (int k = 0; k < rad; ++k) {
      cplx<R> acc = mk<R>(0, 0);
      for (int r = 0; r < rad; ++r) {
        R sn, cs;
        if constexpr (sizeof(R) == 8)
          sincospi(2.0 * ((r * k) % rad) / rad, &sn, &cs);
        else
          sincospif(2.0f * ((r * k) % rad) / rad, &sn, &cs);
        acc = acc + v[r] * mk<R>(cs, sn);
      }
      o[k] = acc;
    }
    for (int k = 0; k < rad; ++k) v[k] = o[k];
  }
}

template <typename R>
__device__ cplx<R>* lds_fft(cplx<R>* a, cplx<R>* b, int n, int nl, int ld, const N3Radix& rad,
                            int tid, int nt) {
  int ns = 1;
  cplx<R>* in = a;
  cplx<R>* out = b;
  for (int st = 0; st < rad.count; ++st) {
    const int R_ = rad.r[st];
    const int nb = n / R_;
    for (int t = tid; t < nl * nb; t += nt) {
      const int line = t / nb, j = t % nb;
      const cplx<R>* src = in + (size_t)line * ld;
      cplx<R>* dst = out + (size_t)line * ld;
      const int k = j % ns;
      cplx<R> v[5];
      R sn, cs;
      if constexpr (sizeof(R) == 8)
        sincospi(2.0 * k / (double)(ns * R_), &sn, &cs);
      else
        sincospif(2.0f * k / (float)(ns * R_), &sn, &cs);
      const cplx<R> w1 = mk<R>(cs, sn);
      cplx<R> tw = mk<R>(1, 0);
      for (int r = 0; r < R_; ++r) {
        v[r] = src[j + r * nb] * tw;
        tw = tw * w1;
      }
      dft_small<R>(v, R_);
      const int idx = (j / ns) * ns * R_ + k;
      for (int r = 0; r < R_; ++r) dst[idx + r * ns] = v[r];
    }
    __syncthreads();
    ns *= R_;
    cplx<R>* tmp = in;
    in = out;
    out = tmp;
  }
  return in;
}

// Contiguous axis (x): one line of length n per workgroup.
template <typename R>
__global__ __launch_bounds__(256) void fft_x_kernel(cplx<R>* grid, int n, long long n_lines,
                                                    N3Radix rad) {
  extern __shared__ unsigned char fsh[];
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);
  cplx<R>* b = a + n;
  for (long long line = blockIdx.x; line < n_lines; line += gridDim.x) {
    cplx<R>* g = grid + line * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = g[i];
    __syncthreads();
    cplx<R>* res = lds_fft<R>(a, b, n, 1, n, rad, threadIdx.x, blockDim.x);
    for (int i = threadIdx.x; i < n; i += blockDim.x) g[i] = res[i];
    __syncthreads();
  }
}

// Strided axis: transform of length m over elements axis_index = base + i * estride
// (i < m), for TW consecutive x positions at once.  Optionally multiplies element k1 of
// sub-transform j2 by exp(+2 pi i j2 k1 / ntot) (first pass of a two-pass transform).
template <typename R>
__global__ __launch_bounds__(256) void fft_strided_kernel(N3StridedParams<R> p) {
  extern __shared__ unsigned char fsh[];
  constexpr int TW = 16;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);
  cplx<R>* b = a + TW * p.m;
  // work item: (x tile, outer index over the other axis, sub-transform index)
  const long long n_xt = (p.nx + TW - 1) / TW;
  for (long long item = blockIdx.x; item < n_xt * p.n_outer * p.n_sub; item += gridDim.x) {
    const long long xt = item % n_xt;
    const long long rest = item / n_xt;
    const int sub = static_cast<int>(rest % p.n_sub);
    const long long outer = rest / p.n_sub;
    const long long x0 = xt * TW;
    // element i of the transform lives at axis index sub * sub_stride + i * estride
    for (int t = threadIdx.x; t < TW * p.m; t += blockDim.x) {
      const int xi = t % TW, i = t / TW;
      const long long x = x0 + xi;
      const long long ai = (long long)sub * p.sub_stride + (long long)i * p.estride;
      cplx<R> v = mk<R>(0, 0);
      if (x < p.nx) v = p.grid[outer * p.outer_stride + ai * p.axis_stride + x];
      a[xi * p.m + i] = v;
    }
    __syncthreads();
    cplx<R>* res = lds_fft<R>(a, b, p.m, TW, p.m, p.rad, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < TW * p.m; t += blockDim.x) {
      const int xi = t % TW, k = t / TW;
      const long long x = x0 + xi;
      if (x >= p.nx) continue;
      cplx<R> v = res[xi * p.m + k];
      if (p.twiddle) {
        R sn, cs;
        const long long e = ((long long)sub * k) % p.ntot;
        if constexpr (sizeof(R) == 8)
          sincospi(2.0 * (double)e / (double)p.ntot, &sn, &cs);
        else
          sincospif(2.0f * (float)e / (float)p.ntot, &sn, &cs);
        v = v * mk<R>(cs, sn);
      }
      const long long ai = (long long)sub * p.sub_stride + (long long)k * p.estride;
      p.grid[outer * p.outer_stride + ai * p.axis_stride + x] = v;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------
// Interpolation at the pixels, correction, real part accumulated into the images.
// --------------------------------------------------------------------------------------
template <typename R, typename TI>
__global__ __launch_bounds__(256) void n3_interp_kernel(N3InterpParams<R> p) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.n_pix) return;
  const int w = p.w2;
  R wx[16], wy[16], wz[16];
  int kx[16], ky[16], kz[16];
  double corr = 1.0, cphase = 0.0;
  const double xq[3] = {p.lmn[q], p.lmn[p.n_pix + q], p.lmn[2 * (size_t)p.n_pix + q]};
  for (int d = 0; d < 3; ++d) {
    const double s = p.gam[d] * (xq[d] - p.D[d]);
    const double a2 = p.h1[d] * s / p.h2[d];
    const int k0 = static_cast<int>(ceil(a2 - 0.5 * w));
    R* wd = d == 0 ? wx : (d == 1 ? wy : wz);
    int* kd = d == 0 ? kx : (d == 1 ? ky : kz);
    for (int o = 0; o < w; ++o) {
      wd[o] = es_kernel<R>(static_cast<R>((k0 + o - a2) * 2.0 / w), R(p.beta2));
      int k = (k0 + o) % p.nf2[d];
      if (k < 0) k += p.nf2[d];
      // two-pass axes store frequency k at (k mod n1) n2 + k / n1
      kd[o] = p.n2[d] > 1 ? (k % p.n1[d]) * p.n2[d] + k / p.n1[d] : k;
    }
    const double alpha1 = 0.5 * p.w1 * p.h1[d];
    corr *= p.h1[d] * p.h2[d] / (alpha1 * es_ft(s * alpha1, p.beta1, p.gl_x, p.gl_w, p.nq));
    cphase += p.C[d] * xq[d];
  }
  const double turns = cphase * 0.15915494309189535;
  double cs, cc;
  sincospi(2.0 * (turns - rint(turns)), &cs, &cc);
  for (int lv = 0; lv < p.n_levels; ++lv) {
    const cplx<R>* g = p.grid + (size_t)lv * p.grid_stride;
    R sr = 0, si = 0;
    for (int oz = 0; oz < w; ++oz) {
      for (int oy = 0; oy < w; ++oy) {
        const cplx<R>* row = g + ((size_t)kz[oz] * p.nf2[1] + ky[oy]) * p.nf2[0];
        const R wyz = wy[oy] * wz[oz];
        R rr = 0, ri = 0;
        for (int ox = 0; ox < w; ++ox) {
          const cplx<R> v = row[kx[ox]];
          rr += wx[ox] * v.re;
          ri += wx[ox] * v.im;
        }
        sr += wyz * rr;
        si += wyz * ri;
      }
    }
    // Re( exp(i C.x) * corr * (sr + i si) )
    const double re = corr * (cc * static_cast<double>(sr) - cs * static_cast<double>(si));
    TI* img = static_cast<TI*>(p.img) + (size_t)lv * p.img_ld;
    img[q] += static_cast<TI>(re * p.scale);
  }
}

template <typename R>
__global__ void zero_kernel(R* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = R(0);
}

// Per-sample min/max of the positions of the rows used (positions / wavelength).
__global__ void pos_range_kernel(const double* pos, long long pos_stride, int apad,
                                 const int* n_rows, int n_rows_const, double* out) {
  const int s = blockIdx.x;
  const int n = n_rows ? n_rows[s] : n_rows_const;
  __shared__ double mn[3][256], mx[3][256];
  for (int d = 0; d < 3; ++d) {
    double lo = 1e300, hi = -1e300;
    for (int a = threadIdx.x; a < n; a += blockDim.x) {
      const double v = pos[(size_t)s * pos_stride + d * apad + a];
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
    mn[d][threadIdx.x] = lo;
    mx[d][threadIdx.x] = hi;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < (int)blockDim.x; ++i) {
      lo = fmin(lo, mn[d][i]);
      hi = fmax(hi, mx[d][i]);
    }
    out[s * 6 + d] = lo;
    out[s * 6 + 3 + d] = hi;
  }
}

// VV_l = Z diag(w_l) Z^H for every (sample, level); Z rows = antennas used (ld ldz).
template <typename R, typename TO>
__global__ __launch_bounds__(256) void vv_levels_kernel(N3VVParams<R> p) {
  __shared__ cplx<R> ti[16][17];
  __shared__ cplx<R> tj[16][17];
  const int sl = blockIdx.z;
  const int s = sl / p.n_levels, lv = sl % p.n_levels;
  const int n = p.n_rows ? p.n_rows[s] : p.n_rows_const;
  const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
  if (i0 >= n || j0 > i0 + 15) return;  // only blocks touching the lower triangle
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const cplx<R>* Z = p.z + (size_t)s * p.stride_z;
  const R* dw = p.wd + ((size_t)s * p.n_levels + lv) * p.n_cols;
  cplx<R> acc = mk<R>(0, 0);
  for (int e0 = 0; e0 < p.n_cols; e0 += 16) {
    const int e = e0 + ty;
    const R de = e < p.n_cols ? dw[e] : R(0);
    const int i = i0 + tx, j = j0 + tx;
    ti[ty][tx] = (e < p.n_cols && i < n) ? Z[i + (size_t)e * p.ldz] * de : mk<R>(0, 0);
    tj[ty][tx] = (e < p.n_cols && j < n) ? Z[j + (size_t)e * p.ldz] : mk<R>(0, 0);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) acc = acc + mulc(ti[k][tx], tj[k][ty]);
    __syncthreads();
  }
  const int i = i0 + tx, j = j0 + ty;
  if (i < n && j < n && i >= j) {
    cplx<TO>* out = reinterpret_cast<cplx<TO>*>(p.vv) + (size_t)sl * p.vv_stride;
    out[i + (size_t)p.ld_vv * j] = cvt<TO>(acc);
  }
}

}  // namespace

// --------------------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------------------
void launch_pos_range(int n_samples, const double* pos, long long pos_stride, int apad,
                      const int* n_rows, int n_rows_const, double* out, hipStream_t st) {
  if (n_samples <= 0) return;
  pos_range_kernel<<<n_samples, 256, 0, st>>>(pos, pos_stride, apad, n_rows, n_rows_const, out);
  check_launch("pos_range_kernel");
}

template <typename R>
void launch_zero(R* p, long long n, hipStream_t st) {
  if (n <= 0) return;
  const long long blocks = std::min<long long>((n + 255) / 256, 1 << 16);
  zero_kernel<R><<<static_cast<unsigned>(blocks), 256, 0, st>>>(p, n);
  check_launch("zero_kernel");
}
template void launch_zero<float>(float*, long long, hipStream_t);
template void launch_zero<double>(double*, long long, hipStream_t);
template void launch_zero<unsigned>(unsigned*, long long, hipStream_t);

template <typename R>
void launch_n3_points(const N3PointsParams& p, int n_samples, hipStream_t st) {
  const long long npair = (long long)p.n_ant_max * (p.n_ant_max + 1) / 2;
  if (n_samples <= 0 || npair <= 0) return;
  dim3 grid(static_cast<unsigned>((npair + 255) / 256), n_samples);
  n3_points_kernel<R><<<grid, 256, 0, st>>>(p);
  check_launch("n3_points_kernel");
}
template void launch_n3_points<float>(const N3PointsParams&, int, hipStream_t);
template void launch_n3_points<double>(const N3PointsParams&, int, hipStream_t);

void launch_scan(const unsigned* in, unsigned* out, unsigned* tmp, int n, hipStream_t st) {
  if (n <= 0) return;
  const int nb = (n + 1023) / 1024;
  if (nb > 1024 * 1024) fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT bins");
  unsigned* sums = tmp;
  unsigned* offs = tmp + nb;
  scan_block_kernel<<<nb, 1024, 0, st>>>(in, out, sums, n);
  check_launch("scan_block_kernel");
  if (nb > 1) {
    unsigned* sums2 = offs + nb;
    launch_scan(sums, offs, sums2, nb, st);
    scan_add_kernel<<<nb, 1024, 0, st>>>(out, offs, n);
    check_launch("scan_add_kernel");
  }
}

void launch_n3_scatter(long long n_pts, const unsigned char* valid, const int* bin_of,
                       const unsigned* bin_start, unsigned* bin_fill, unsigned* sorted,
                       hipStream_t st) {
  if (n_pts <= 0) return;
  n3_scatter_kernel<<<static_cast<unsigned>((n_pts + 255) / 256), 256, 0, st>>>(
      n_pts, valid, bin_of, bin_start, bin_fill, sorted);
  check_launch("n3_scatter_kernel");
}

template <typename R, typename TV>
size_t n3_spread_shmem(const N3SpreadParams<R>& p) {
  const int tcells = (p.bin_size[0] + p.w) * (p.bin_size[1] + p.w) * (p.bin_size[2] + p.w);
  return sizeof(R) * ((size_t)2 * p.n_levels * tcells + 32 * 3 * p.w + 32 * p.n_levels * 2) +
         sizeof(int) * 32 * 3;
}

template <typename R, typename TV>
void launch_n3_spread(const N3SpreadParams<R>& p, int n_bins, hipStream_t st) {
  if (n_bins <= 0) return;
  const size_t sh = n3_spread_shmem<R, TV>(p);
  if (sh > 160 * 1024) fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT spread tile exceeds LDS");
  static bool attr_set = false;
  if (!attr_set) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&n3_spread_kernel<R, TV>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  n3_spread_kernel<R, TV><<<n_bins, SP_NT, sh, st>>>(p);
  check_launch("n3_spread_kernel");
}
template void launch_n3_spread<float, cplx<float>>(const N3SpreadParams<float>&, int, hipStream_t);
template void launch_n3_spread<double, cplx<double>>(const N3SpreadParams<double>&, int,
                                                     hipStream_t);

template <typename R>
void launch_n3_deconv(const N3DeconvParams<R>& p, int n_levels, hipStream_t st) {
  const long long n = (long long)p.nf1[0] * p.nf1[1] * p.nf1[2];
  if (n <= 0 || n_levels <= 0) return;
  dim3 grid(static_cast<unsigned>(std::min<long long>((n + 255) / 256, 16384)), n_levels);
  n3_deconv_kernel<R><<<grid, 256, 0, st>>>(p);
  check_launch("n3_deconv_kernel");
}
template void launch_n3_deconv<float>(const N3DeconvParams<float>&, int, hipStream_t);
template void launch_n3_deconv<double>(const N3DeconvParams<double>&, int, hipStream_t);

template <typename R>
void launch_fft_x(cplx<R>* grid, int n, long long n_lines, const N3Radix& rad, hipStream_t st) {
  if (n_lines <= 0) return;
  const size_t sh = 2 * sizeof(cplx<R>) * (size_t)n;
  static bool attr_set = false;
  if (!attr_set) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_x_kernel<R>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(n_lines, 8192));
  fft_x_kernel<R><<<blocks, 256, sh, st>>>(grid, n, n_lines, rad);
  check_launch("fft_x_kernel");
}
template void launch_fft_x<float>(cplx<float>*, int, long long, const N3Radix&, hipStream_t);
template void launch_fft_x<double>(cplx<double>*, int, long long, const N3Radix&, hipStream_t);

template <typename R>
void launch_fft_strided(const N3StridedParams<R>& p, hipStream_t st) {
  const long long items = ((p.nx + 15) / 16) * p.n_outer * p.n_sub;
  if (items <= 0) return;
  const size_t sh = 2 * 16 * sizeof(cplx<R>) * (size_t)p.m;
  static bool attr_set = false;
  if (!attr_set) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_strided_kernel<R>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(items, 16384));
  fft_strided_kernel<R><<<blocks, 256, sh, st>>>(p);
  check_launch("fft_strided_kernel");
}
template void launch_fft_strided<float>(const N3StridedParams<float>&, hipStream_t);
template void launch_fft_strided<double>(const N3StridedParams<double>&, hipStream_t);

template <typename R, typename TI>
void launch_n3_interp(const N3InterpParams<R>& p, hipStream_t st) {
  if (p.n_pix <= 0) return;
  if (p.w2 > 16) fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT kernel width above 16");
  n3_interp_kernel<R, TI><<<ceil_div(p.n_pix, 256), 256, 0, st>>>(p);
  check_launch("n3_interp_kernel");
}
template void launch_n3_interp<float, float>(const N3InterpParams<float>&, hipStream_t);
template void launch_n3_interp<double, double>(const N3InterpParams<double>&, hipStream_t);

template <typename R, typename TO>
void launch_vv_levels(const N3VVParams<R>& p, int n_samples, int n_rows_max, hipStream_t st) {
  if (n_samples <= 0 || n_rows_max <= 0) return;
  dim3 grid(ceil_div(n_rows_max, 16), ceil_div(n_rows_max, 16), n_samples * p.n_levels);
  vv_levels_kernel<R, TO><<<grid, 256, 0, st>>>(p);
  check_launch("vv_levels_kernel");
}
template void launch_vv_levels<double, float>(const N3VVParams<double>&, int, int, hipStream_t);
template void launch_vv_levels<double, double>(const N3VVParams<double>&, int, int, hipStream_t);
template void launch_vv_levels<float, float>(const N3VVParams<float>&, int, int, hipStream_t);
template void launch_vv_levels<float, double>(const N3VVParams<float>&, int, int, hipStream_t);

}  // namespace bippamd
