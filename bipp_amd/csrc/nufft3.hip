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
  N3Point pt;
  int bin[3];
  double ph = 0;
  for (int d = 0; d < 3; ++d) {
    u[d] = 6.283185307179586 * (pos[d * p.apad + a2] - pos[d * p.apad + b]);
    const double up = u[d] - p.C[d];
    ph += up * p.D[d];
    cell[d] = up * p.inv_gh[d];  // t / h1 in grid units
    if (p.phase_fp32) cell[d] = static_cast<double>(static_cast<float>(up) * static_cast<float>(p.inv_gh[d]));
    const int l0 = static_cast<int>(ceil(cell[d] - 0.5 * p.w));
    pt.fr[d] = static_cast<float>(cell[d] - l0);
    pt.l0[d] = l0 + p.off[d];
    // the plan keeps every support inside [-nf1/2, nf1/2); the clamp only keeps the bin
    // index in range (cells outside a block are never written by it)
    bin[d] = min(max(pt.l0[d], 0) / p.bin_size[d], p.nbins[d] - 1);
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
  pt.ph[0] = static_cast<float>(pc);
  pt.ph[1] = static_cast<float>(ps);
  pt.src = static_cast<int>(a2 + p.ld_vv * b);  // VV(a, b) inside the sample's matrix
  pt.mult = a2 == b ? 1.f : 2.f;                 // (a,b) and (b,a) are conjugate: 2 Re(.)
  pt.sample = s;
  pt.pad = 0;
  p.valid[pid] = 1;
  p.bin_of[pid] = binid;
  p.pts[pid] = pt;
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

// counting sort: point records copied to their bin's segment
__global__ void n3_scatter_kernel(long long n_pts, const unsigned char* valid, const int* bin_of,
                                  const N3Point* pts, const unsigned* bin_start,
                                  unsigned* bin_fill, N3Point* sorted) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n_pts || !valid[i]) return;
  const int b = bin_of[i];
  const unsigned slot = bin_start[b] + atomicAdd(&bin_fill[b], 1u);
  sorted[slot] = pts[i];
}

// --------------------------------------------------------------------------------------
// Spreading with output ownership.  The workgroup of block (bx, by, bz) owns the cells
// l + nf1/2 in [b B, b B + B) per axis.  It walks the bins that can reach the block (a point
// of bin c covers cells [l0, l0 + w), l0 in bin c), clips each point's support to the block,
// accumulates kernel weight x value for every level in an LDS tile, and finally writes the
// tile times the deconvolution factors with plain stores.  Every cell of [-nf1/2, nf1/2)^3
// is written by exactly one workgroup; no global atomics, no zero pass over the grid.
// --------------------------------------------------------------------------------------
constexpr int SP_PB = 64;  // points per batch (one wavefront scans the item counts)

template <typename R, typename TV>
__global__ __launch_bounds__(SP_NT) void n3_spread_kernel(N3SpreadParams<R> p) {
  extern __shared__ __align__(16) unsigned char sp_sh[];
  const int w = p.w;
  const int BX = p.bin_size[0], BY = p.bin_size[1], BZ = p.bin_size[2];
  const int tcells = BX * BY * BZ;
  const int L = p.n_levels;
  R* tile = reinterpret_cast<R*>(sp_sh);                          // [L][tcells][2]
  R* wts = tile + (size_t)2 * L * tcells;                         // [PB][3][w]
  R* val = wts + (size_t)SP_PB * 3 * w;                           // [PB][L][2]
  int* info = reinterpret_cast<int*>(val + (size_t)SP_PB * L * 2);  // [PB][8]
  int* pref = info + SP_PB * 8;                                   // [PB + 1]
  const int bin = blockIdx.x;
  const int bx = bin % p.nbins[0];
  const int by = (bin / p.nbins[0]) % p.nbins[1];
  const int bz = bin / (p.nbins[0] * p.nbins[1]);
  const int org[3] = {bx * BX, by * BY, bz * BZ};
  const int bsz[3] = {BX, BY, BZ};
  for (int i = threadIdx.x; i < 2 * L * tcells; i += SP_NT) tile[i] = R(0);
  const R two_over_w = R(2) / R(w);
  const R beta = R(p.beta);
  for (int cz = max(0, bz - p.nback[2]); cz <= bz; ++cz)
    for (int cy = max(0, by - p.nback[1]); cy <= by; ++cy)
      for (int cx = max(0, bx - p.nback[0]); cx <= bx; ++cx) {
        const int nbid = (cz * p.nbins[1] + cy) * p.nbins[0] + cx;
        const unsigned beg = p.bin_start[nbid];
        const unsigned cnt = p.bin_count[nbid];
        for (unsigned b0 = 0; b0 < cnt; b0 += SP_PB) {
          const int nb = min<unsigned>(SP_PB, cnt - b0);
          __syncthreads();  // previous batch fully consumed
          // support of each point clipped to the block
          if (threadIdx.x < SP_PB) {
            int items = 0;
            if ((int)threadIdx.x < nb) {
              const N3Point& pt = p.pts[beg + b0 + threadIdx.x];
              items = 1;
              for (int d = 0; d < 3; ++d) {
                const int lo = max(pt.l0[d], org[d]);
                const int hi = min(pt.l0[d] + w, org[d] + bsz[d]);
                const int c = max(0, hi - lo);
                info[threadIdx.x * 8 + d] = lo - pt.l0[d];      // first kernel offset
                info[threadIdx.x * 8 + 3 + d] = c;              // offsets inside the block
                items *= c;
              }
              info[threadIdx.x * 8 + 6] = pt.l0[0] - org[0];
              info[threadIdx.x * 8 + 7] = (pt.l0[1] - org[1]) * BX + (pt.l0[2] - org[2]) * BX * BY;
            }
            // inclusive scan over the wavefront
            int v = items;
            for (int o = 1; o < SP_PB; o <<= 1) {
              const int u = __shfl_up(v, o, SP_PB);
              if ((int)threadIdx.x >= o) v += u;
            }
            pref[threadIdx.x + 1] = v;
            if (threadIdx.x == 0) pref[0] = 0;
          }
          __syncthreads();
          const int total = pref[nb];
          if (total == 0) continue;
          // kernel weights of the clipped offsets, values of every level
          for (int t = threadIdx.x; t < nb * 3 * w; t += SP_NT) {
            const int pi = t / (3 * w), rem = t % (3 * w), d = rem / w, j = rem % w;
            if (j >= info[pi * 8 + 3 + d]) continue;
            const N3Point& pt = p.pts[beg + b0 + pi];
            const int o = info[pi * 8 + d] + j;
            wts[(pi * 3 + d) * w + j] = es_kernel<R>((R(o) - R(pt.fr[d])) * two_over_w, beta);
          }
          for (int t = threadIdx.x; t < nb * L; t += SP_NT) {
            const int pi = t / L, lv = t % L;
            if (pref[pi + 1] == pref[pi]) continue;
            const N3Point& pt = p.pts[beg + b0 + pi];
            const TV* vv =
                p.vv + ((size_t)pt.sample * p.n_levels_total + p.lv0 + lv) * p.vv_stride;
            const TV c = vv[pt.src];
            const R m = R(pt.mult);
            const R pr = R(pt.ph[0]), pim = R(pt.ph[1]);
            val[(pi * L + lv) * 2 + 0] = m * (R(c.re) * pr - R(c.im) * pim);
            val[(pi * L + lv) * 2 + 1] = m * (R(c.re) * pim + R(c.im) * pr);
          }
          __syncthreads();
          // items (point, clipped offset), x offsets fastest: consecutive lanes take
          // consecutive cells
          for (int it = threadIdx.x; it < total; it += SP_NT) {
            int lo = 0, hi = nb;  // pref[lo] <= it < pref[hi]
            while (hi - lo > 1) {
              const int mid = (lo + hi) >> 1;
              if (pref[mid] <= it) lo = mid;
              else hi = mid;
            }
            const int pi = lo;
            const int r = it - pref[pi];
            const int c0 = info[pi * 8 + 3], c1 = info[pi * 8 + 4];
            const int jx = r % c0, rq = r / c0, jy = rq % c1, jz = rq / c1;
            const R wt = wts[(pi * 3 + 0) * w + jx] * wts[(pi * 3 + 1) * w + jy] *
                         wts[(pi * 3 + 2) * w + jz];
            // tile cell of kernel offset (info0 + jx, info1 + jy, info2 + jz)
            const int cell = info[pi * 8 + 6] + info[pi * 8 + 0] + jx +
                             (info[pi * 8 + 1] + jy) * BX + (info[pi * 8 + 2] + jz) * BX * BY +
                             info[pi * 8 + 7];
            for (int lv = 0; lv < L; ++lv) {
              atomicAdd(&tile[((size_t)lv * tcells + cell) * 2 + 0], wt * val[(pi * L + lv) * 2 + 0]);
              atomicAdd(&tile[((size_t)lv * tcells + cell) * 2 + 1], wt * val[(pi * L + lv) * 2 + 1]);
            }
          }
        }
      }
  __syncthreads();
  // deconvolve and store the owned cells inside [-nf1/2, nf1/2)^3 (grid index l mod nf2)
  for (int i = threadIdx.x; i < L * tcells; i += SP_NT) {
    const int lv = i / tcells, cell = i % tcells;
    const int ix = org[0] + cell % BX, iy = org[1] + (cell / BX) % BY, iz = org[2] + cell / (BX * BY);
    if (ix >= p.nf1[0] || iy >= p.nf1[1] || iz >= p.nf1[2]) continue;
    const R f = p.fx[ix] * p.fy[iy] * p.fz[iz];
    int gx = ix - p.nf1[0] / 2, gy = iy - p.nf1[1] / 2, gz = iz - p.nf1[2] / 2;
    gx += gx < 0 ? p.nf2[0] : 0;
    gy += gy < 0 ? p.nf2[1] : 0;
    gz += gz < 0 ? p.nf2[2] : 0;
    p.grid[(size_t)(p.lv0 + lv) * p.grid_stride + ((size_t)gz * p.nf2[1] + gy) * p.nf2[0] + gx] =
        mk<R>(tile[(size_t)i * 2] * f, tile[(size_t)i * 2 + 1] * f);
  }
}

// --------------------------------------------------------------------------------------
// Stockham FFT in LDS (sign +): `nl` lines of length n, line stride ld, radices r[].
// Returns the buffer holding the result.
// --------------------------------------------------------------------------------------
// Small DFTs with sign +, constant roots.
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
  } else if (rad == 3) {
    // w = exp(+2 pi i / 3): y1,2 = a - (b + c)/2 +- i sqrt(3)/2 (b - c)
    const R h = R(0.86602540378443864676);
    const cplx<R> a = v[0], t1 = v[1] + v[2], t2 = v[1] - v[2];
    const cplx<R> m = mk<R>(a.re - R(0.5) * t1.re, a.im - R(0.5) * t1.im);
    const cplx<R> q = mk<R>(-h * t2.im, h * t2.re);
    v[0] = a + t1;
    v[1] = m + q;
    v[2] = m - q;
  } else {
    // radix 5: roots exp(+2 pi i k / 5)
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    const R cr[5] = {R(1), c1, c2, c2, c1};
    const R sr[5] = {R(0), s1, s2, -s2, -s1};
    cplx<R> o[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      cplx<R> acc = v[0];
#pragma unroll
      for (int r = 1; r < 5; ++r) {
        const int e = (r * k) % 5;
        acc = acc + v[r] * mk<R>(cr[e], sr[e]);
      }
      o[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) v[k] = o[k];
  }
}

// Stockham stages in LDS; `tw` holds exp(+2 pi i t / n) for t < n.
template <typename R>
__device__ cplx<R>* lds_fft(cplx<R>* a, cplx<R>* b, int n, int nl, int ld, const N3Radix& rad,
                            const cplx<R>* __restrict__ tw, int tid, int nt) {
  int ns = 1;
  cplx<R>* in = a;
  cplx<R>* out = b;
  for (int st = 0; st < rad.count; ++st) {
    const int R_ = rad.r[st];
    const int nb = n / R_;
    const int tstride = n / (ns * R_);  // twiddle k r / (ns R) of the full length
    for (int t = tid; t < nl * nb; t += nt) {
      const int line = t / nb, j = t % nb;
      const cplx<R>* src = in + (size_t)line * ld;
      cplx<R>* dst = out + (size_t)line * ld;
      const int k = j % ns;
      cplx<R> v[5];
      v[0] = src[j];
      for (int r = 1; r < R_; ++r) v[r] = src[j + r * nb] * tw[k * r * tstride];
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

// position of index a inside window [start, start + count) of an axis of length n
__device__ __forceinline__ bool in_window(int a, const N3Window& wd, int n) {
  int d = a - wd.start;
  d += d < 0 ? n : 0;
  return d < wd.count;
}

// Contiguous axis (x): `nl` lines per workgroup (lines (level, z, y) of the windows).
template <typename R>
__global__ __launch_bounds__(256) void fft_lines_kernel(N3LineParams<R> p, int nl) {
  extern __shared__ __align__(16) unsigned char fsh[];
  __shared__ long long lbase[16];
  const int n = p.n;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);
  cplx<R>* b = a + (size_t)nl * n;
  const long long n_lines = (long long)p.n_levels * p.zw.count * p.yw.count;
  const long long n_groups = (n_lines + nl - 1) / nl;
  for (long long grp = blockIdx.x; grp < n_groups; grp += gridDim.x) {
    const long long line0 = grp * nl;
    const int nlg = static_cast<int>(min<long long>(nl, n_lines - line0));
    if ((int)threadIdx.x < nlg) {
      const long long line = line0 + threadIdx.x;
      const int yo = static_cast<int>(line % p.yw.count);
      const long long r = line / p.yw.count;
      const int zo = static_cast<int>(r % p.zw.count);
      const int lv = static_cast<int>(r / p.zw.count);
      int y = p.yw.start + yo, z = p.zw.start + zo;
      y -= y >= p.ny ? p.ny : 0;
      z -= z >= p.nz ? p.nz : 0;
      lbase[threadIdx.x] = (long long)lv * p.grid_stride + ((long long)z * p.ny + y) * n;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nlg * n; i += blockDim.x) {
      const int li = i / n, x = i - li * n;
      a[i] = in_window(x, p.in, n) ? p.grid[lbase[li] + x] : mk<R>(0, 0);
    }
    __syncthreads();
    cplx<R>* res = lds_fft<R>(a, b, n, nlg, n, p.rad, p.tw, threadIdx.x, blockDim.x);
    for (int i = threadIdx.x; i < nlg * n; i += blockDim.x) {
      const int li = i / n, k = i - li * n;
      if (in_window(k, p.out, n)) p.grid[lbase[li] + k] = res[i];
    }
    __syncthreads();
  }
}

// Strided axis: transforms of length m for TW consecutive x positions at once; element i of
// sub-transform `sub` at axis index sub * sub_stride + i * estride.  The LDS line stride is
// odd so that the TW lines of a load fall into different banks.
template <typename R>
__global__ __launch_bounds__(256) void fft_pass_kernel(N3PassParams<R> p) {
  extern __shared__ __align__(16) unsigned char fsh[];
  constexpr int TW = 16;
  const int ld = p.m | 1;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);
  cplx<R>* b = a + TW * ld;
  const int n_xt = (p.xw.count + TW - 1) / TW;
  const long long n_items = (long long)n_xt * p.n_sub * p.other.count * p.n_levels;
  for (long long item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int xt = static_cast<int>(item % n_xt);
    long long rest = item / n_xt;
    const int sub = static_cast<int>(rest % p.n_sub);
    rest /= p.n_sub;
    const int o = static_cast<int>(rest % p.other.count);
    const int lv = static_cast<int>(rest / p.other.count);
    int oo = p.other.start + o;
    oo -= oo >= p.no ? p.no : 0;
    if (p.operm_n2 > 1) oo = (oo % p.operm_n1) * p.operm_n2 + oo / p.operm_n1;
    cplx<R>* base = p.grid + (size_t)lv * p.grid_stride + (size_t)oo * p.outer_stride;
    for (int t = threadIdx.x; t < TW * p.m; t += blockDim.x) {
      const int xi = t % TW, i = t / TW;
      const int j = xt * TW + xi;
      cplx<R> v = mk<R>(0, 0);
      if (j < p.xw.count) {
        int x = p.xw.start + j;
        x -= x >= p.nx ? p.nx : 0;
        const long long ai = (long long)sub * p.sub_stride + (long long)i * p.estride;
        if (!p.in_check || in_window(static_cast<int>(ai), p.in, p.nax))
          v = base[ai * p.axis_stride + x];
      }
      a[xi * ld + i] = v;
    }
    __syncthreads();
    cplx<R>* res = lds_fft<R>(a, b, p.m, TW, ld, p.rad, p.tw, threadIdx.x, blockDim.x);
    for (int t = threadIdx.x; t < TW * p.m; t += blockDim.x) {
      const int xi = t % TW, k = t / TW;
      const int j = xt * TW + xi;
      if (j >= p.xw.count) continue;
      if (p.out_check && !in_window(sub * p.fa + k * p.fb, p.out, p.nax)) continue;
      int x = p.xw.start + j;
      x -= x >= p.nx ? p.nx : 0;
      cplx<R> v = res[xi * ld + k];
      if (p.twiddle) v = v * p.tw2[((long long)sub * k) % p.ntot];
      const long long ai = (long long)sub * p.sub_stride + (long long)k * p.estride;
      base[ai * p.axis_stride + x] = v;
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
                       const N3Point* pts, const unsigned* bin_start, unsigned* bin_fill,
                       N3Point* sorted, hipStream_t st) {
  if (n_pts <= 0) return;
  n3_scatter_kernel<<<static_cast<unsigned>((n_pts + 255) / 256), 256, 0, st>>>(
      n_pts, valid, bin_of, pts, bin_start, bin_fill, sorted);
  check_launch("n3_scatter_kernel");
}

template <typename R>
size_t n3_spread_shmem(const N3SpreadParams<R>& p) {
  const size_t tcells = (size_t)p.bin_size[0] * p.bin_size[1] * p.bin_size[2];
  return sizeof(R) * (2 * p.n_levels * tcells + SP_PB * 3 * p.w + SP_PB * p.n_levels * 2) +
         sizeof(int) * (SP_PB * 8 + SP_PB + 1);
}

template <typename R, typename TV>
void launch_n3_spread(const N3SpreadParams<R>& p_all, hipStream_t st) {
  const long long n_blocks = (long long)p_all.nbins[0] * p_all.nbins[1] * p_all.nbins[2];
  if (n_blocks <= 0) return;
  if (n_blocks > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT spread blocks");
  for (int d = 0; d < 3; ++d)
    if ((long long)p_all.nbins[d] * p_all.bin_size[d] < p_all.nf1[d] || p_all.nback[d] < 0)
      fail(BIPP_INTERNAL_ERROR, "NUFFT spread blocks do not cover the grid");
  static bool attr_set = false;
  if (!attr_set) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&n3_spread_kernel<R, TV>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  // as many levels per launch as the LDS tile allows
  const int L = p_all.n_levels_total;
  int group = L;
  for (; group > 1; --group) {
    N3SpreadParams<R> t = p_all;
    t.n_levels = group;
    if (n3_spread_shmem<R>(t) <= 64 * 1024) break;
  }
  for (int lv0 = 0; lv0 < L; lv0 += group) {
    N3SpreadParams<R> p = p_all;
    p.lv0 = lv0;
    p.n_levels = std::min(group, L - lv0);
    const size_t sh = n3_spread_shmem<R>(p);
    if (sh > 160 * 1024) fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT spread tile exceeds LDS");
    n3_spread_kernel<R, TV><<<static_cast<unsigned>(n_blocks), SP_NT, sh, st>>>(p);
    check_launch("n3_spread_kernel");
  }
}
template void launch_n3_spread<float, cplx<float>>(const N3SpreadParams<float>&, hipStream_t);
template void launch_n3_spread<double, cplx<double>>(const N3SpreadParams<double>&, hipStream_t);

namespace {
bool window_ok(const N3Window& w, int n) {
  return n > 0 && w.start >= 0 && w.start < n && w.count >= 0 && w.count <= n;
}
}  // namespace

template <typename R>
void launch_fft_lines(const N3LineParams<R>& p, hipStream_t st) {
  const long long n_lines = (long long)p.n_levels * p.zw.count * p.yw.count;
  if (n_lines <= 0) return;
  if (!window_ok(p.yw, p.ny) || !window_ok(p.zw, p.nz) || !window_ok(p.in, p.n) ||
      !window_ok(p.out, p.n))
    fail(BIPP_INTERNAL_ERROR, "FFT line windows out of range");
  // several short lines per workgroup so that every stage keeps most lanes busy
  const int nl = std::max(1, std::min(16, 2048 / p.n));
  const size_t sh = 2 * sizeof(cplx<R>) * (size_t)p.n * nl;
  // 1 KiB of the 160 KiB stays for the kernel's static line bases
  if (sh > 159 * 1024) fail(BIPP_INTERNAL_ERROR, "FFT line exceeds LDS");
  static bool attr_set = false;
  if (!attr_set) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_lines_kernel<R>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
    attr_set = true;
  }
  const long long groups = (n_lines + nl - 1) / nl;
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(groups, 32768));
  fft_lines_kernel<R><<<blocks, 256, sh, st>>>(p, nl);
  check_launch("fft_lines_kernel");
}
template void launch_fft_lines<float>(const N3LineParams<float>&, hipStream_t);
template void launch_fft_lines<double>(const N3LineParams<double>&, hipStream_t);

template <typename R>
void launch_fft_pass(const N3PassParams<R>& p, hipStream_t st) {
  const long long items =
      (long long)((p.xw.count + 15) / 16) * p.n_sub * p.other.count * p.n_levels;
  if (items <= 0) return;
  if (!window_ok(p.xw, p.nx) || !window_ok(p.other, p.no) || !window_ok(p.in, p.nax) ||
      !window_ok(p.out, p.nax) || (long long)p.m * p.n_sub != p.nax)
    fail(BIPP_INTERNAL_ERROR, "FFT pass windows out of range");
  const size_t sh = 2 * 16 * sizeof(cplx<R>) * (size_t)(p.m | 1);
  if (sh > 160 * 1024) fail(BIPP_INTERNAL_ERROR, "FFT pass exceeds LDS");
  static bool attr_set = false;
  if (!attr_set) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_pass_kernel<R>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(items, 32768));
  fft_pass_kernel<R><<<blocks, 256, sh, st>>>(p);
  check_launch("fft_pass_kernel");
}
template void launch_fft_pass<float>(const N3PassParams<float>&, hipStream_t);
template void launch_fft_pass<double>(const N3PassParams<double>&, hipStream_t);

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
