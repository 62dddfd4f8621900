// Steering-matrix contraction: images straight from eigenvectors.
//
// For one sample with eigenvectors V (antennas x K), per-level weights w_l[e] =
// scale * f(d_e [lo_l <= d_e <= hi_l]) and antenna positions p_i, the reference forms
// virtual visibilities VV_l = V diag(w_l) V^H (src/host/virtual_vis.cpp:19-73) and sums
// Re sum_ij VV_l[i,j] exp(+i 2pi/lambda (p_i - p_j).x) over pixels x
// (src/host/kernels/nuft_sum.cpp:9-31).  The same image is
//
//     I_l(x) = sum_e w_l[e] |B_e(x)|^2 ,   B_e(x) = sum_i V_ie exp(+i 2pi p_i.x / lambda),
//
// i.e. a complex GEMM B = V^T E with E generated on the fly, followed by a weighted
// |.|^2 reduction over e.  The GEMM is written as a real GEMM with interleaved rows
// (Re B_e, Im B_e) and interleaved contraction index (antenna i, re/im of E):
//     Re-row coefficients (Re V, -Im V), Im-row coefficients (Im V, Re V).
//
// Modes:
//   BF16X3: phase p.x/lambda in float64, reduced to one turn, sin/cos in float32, both
//           operands split into bf16 hi + lo, products hi*hi + hi*lo + lo*hi on
//           v_mfma_f32_16x16x32_bf16, float32 accumulation.
//   FP32:   everything float32 (phase included) on v_mfma_f32_16x16x4_f32.
//   FP64:   everything float64 on v_mfma_f64_16x16x4_f64 (the GPU oracle form).
//
// Workgroup = 8 waves (4 for FP64) over a tile of 128 (64) pixels; each wave owns 4 row
// tiles (64 rows) and all pixel columns; the workgroup loops over every sample of the
// batch, so each pixel tile is written once per launch (no atomics in HBM).  The E tile of
// a 16-antenna chunk is generated once per workgroup into LDS (double-buffered) in the
// exact MFMA B-operand layout.
#include "common.hpp"

#include <algorithm>
#include "kernels.hpp"

#include <type_traits>

namespace bippamd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int LMAX = 8;  // energy levels per launch

// fragment element counts per (row tile, k-chunk)
//   BF16X3: [hi|lo][64 lanes][8 bf16]   -> 1024 bf16
//   FP32:   [64 lanes][8 floats]        -> 512 floats
//   FP64:   [64 lanes][8 doubles]       -> 512 doubles
__host__ __device__ constexpr long long frag_elems(int mode) { return mode == STEER_BF16X3 ? 1024 : 512; }

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  __bf16 h = (__bf16)x;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short b) {
  return __builtin_bit_cast(float, (unsigned)b << 16);
}

template <typename R>
__device__ __forceinline__ R apply_filter(int f, R v) {
  switch (f) {
    case BIPP_FILTER_STD:
      return v > R(0) ? R(1) : (v < R(0) ? R(-1) : R(0));
    case BIPP_FILTER_SQRT:
      return v > R(0) ? sqrt(v) : (v < R(0) ? -sqrt(-v) : R(0));
    case BIPP_FILTER_INV:
      return v != R(0) ? R(1) / v : R(0);
    case BIPP_FILTER_INV_SQRT:
      return v > R(0) ? R(1) / sqrt(v) : (v < R(0) ? R(-1) / sqrt(-v) : R(0));
    default:
      return v;
  }
}

// --------------------------------------------------------------------------------------
// Prep: level weights, eigenvector compaction, positions in wavelengths, A fragments.
// One workgroup (256 threads) per sample.
// --------------------------------------------------------------------------------------
template <typename R>
__global__ __launch_bounds__(256) void steer_prep_kernel(SteerPrepParams<R> p) {
  extern __shared__ unsigned char prep_sh[];
  const int sid = blockIdx.x;
  const int tid = threadIdx.x;
  const int ncol = p.n_cols;
  const int L = p.n_levels;
  int* elist = reinterpret_cast<int*>(prep_sh);          // [ncol]
  int* act = elist + ncol;                               // [ncol]
  __shared__ int ne_sh;
  const int n = p.n_red[sid];
  const int nrows = p.n_rows_const > 0 ? p.n_rows_const : n;
  const R* d = p.d + (size_t)sid * p.stride_d;
  const unsigned long long nv = p.n_vis[sid];
  const double scale = nv ? 1.0 / static_cast<double>(nv) : 0.0;

  for (int e = tid; e < ncol; e += blockDim.x) {
    bool any = false;
    if (e < n) {
      const R dv = d[e];
      const float df = static_cast<float>(dv);
      for (int lv = 0; lv < L; ++lv) {
        const float lo = static_cast<float>(p.intervals[2 * lv]);
        const float hi = static_cast<float>(p.intervals[2 * lv + 1]);
        const bool m = df >= lo && df <= hi;
        const R wv = static_cast<R>(scale) * apply_filter<R>(p.filter, m ? dv : R(0));
        any |= (wv != R(0));
      }
    }
    act[e] = any;
  }
  __syncthreads();
  if (tid == 0) {
    int ne = 0;
    for (int e = 0; e < ncol; ++e)
      if (act[e]) elist[ne++] = e;
    ne_sh = ne;
    p.n_eig[sid] = ne;
    p.n_ant[sid] = nrows;
  }
  __syncthreads();
  const int ne = ne_sh;

  // weights [lv][epad]
  for (int i = tid; i < L * p.epad; i += blockDim.x) {
    const int lv = i / p.epad, ec = i % p.epad;
    R wv = 0;
    if (ec < ne) {
      const R dv = d[elist[ec]];
      const float df = static_cast<float>(dv);
      const float lo = static_cast<float>(p.intervals[2 * lv]);
      const float hi = static_cast<float>(p.intervals[2 * lv + 1]);
      const bool m = df >= lo && df <= hi;
      wv = static_cast<R>(scale) * apply_filter<R>(p.filter, m ? dv : R(0));
    }
    if (p.mode == STEER_FP64)
      static_cast<double*>(p.wts)[(size_t)sid * p.w_stride + i] = static_cast<double>(wv);
    else
      static_cast<float*>(p.wts)[(size_t)sid * p.w_stride + i] = static_cast<float>(wv);
  }

  // positions in wavelengths
  {
    const double* xyz = p.xyz + (size_t)sid * p.stride_xyz;
    const double inv_wl = 1.0 / p.wl[sid];
    double* pos = p.pos + (size_t)sid * p.pos_stride;
    const int* idx = p.idx ? p.idx + (size_t)sid * p.idx_stride : nullptr;
    for (int a = tid; a < p.apad; a += blockDim.x) {
      for (int c = 0; c < 3; ++c) {
        double v = 0;
        if (a < nrows) {
          const int ant = idx ? idx[a] : a;
          v = xyz[ant + (size_t)c * p.ldxyz] * inv_wl;
        }
        pos[(size_t)c * p.apad + a] = v;
      }
    }
  }

  // A fragments
  const cplx<R>* Z = p.z + (size_t)sid * p.stride_z;
  const long long items = (long long)p.rtiles * p.kchunks * 64;
  for (long long it = tid; it < items; it += blockDim.x) {
    const int lane = static_cast<int>(it & 63);
    const long long tk = it >> 6;
    const int kc = static_cast<int>(tk % p.kchunks);
    const int rt = static_cast<int>(tk / p.kchunks);
    // Output rows of a 16-row tile hold (Re, Im) of 8 eigenvectors.  The pairing follows the
    // accumulator layout of the MFMA used, so that each lane holds both parts of an
    // eigenvector: f32 16x16 tiles put rows 4g..4g+3 in lane group g (pairs (2k, 2k+1));
    // v_mfma_f64_16x16x4f64 puts rows g, g+4, g+8, g+12 in lane group g (pairs (r, r+4)).
    const int rho = lane & 15;
    int ec, comp;
    if (p.mode == STEER_FP64) {
      ec = rt * 8 + (rho & 3) + 4 * (rho >> 3);
      comp = (rho >> 2) & 1;
    } else {
      ec = rt * 8 + (rho >> 1);
      comp = rho & 1;
    }
    const bool rvalid = ec < ne;
    const int e = rvalid ? elist[ec] : 0;
    float vals32[8];
    double vals64[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      int kp;
      if (p.mode == STEER_BF16X3)
        kp = kc * 32 + 8 * (lane >> 4) + j;
      else
        kp = kc * 32 + 4 * j + (lane >> 4);
      const int a = kp >> 1, c = kp & 1;
      double v = 0;
      if (rvalid && a < nrows) {
        const cplx<R> z = Z[a + (size_t)e * p.ldz];
        if (comp == 0)
          v = c == 0 ? static_cast<double>(z.re) : -static_cast<double>(z.im);
        else
          v = c == 0 ? static_cast<double>(z.im) : static_cast<double>(z.re);
      }
      vals64[j] = v;
      vals32[j] = static_cast<float>(v);
    }
    const long long base = (long long)sid * p.afrag_stride + ((long long)rt * p.kchunks + kc) * frag_elems(p.mode);
    if (p.mode == STEER_BF16X3) {
      unsigned short* f = static_cast<unsigned short*>(p.afrag) + base;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const unsigned short h = bf16_bits(vals32[j]);
        const float res = static_cast<float>(vals64[j] - static_cast<double>(bf16_to_f32(h)));
        f[lane * 8 + j] = h;
        f[512 + lane * 8 + j] = bf16_bits(res);
      }
    } else if (p.mode == STEER_FP32) {
      float* f = static_cast<float*>(p.afrag) + base;
#pragma unroll
      for (int j = 0; j < 8; ++j) f[lane * 8 + j] = vals32[j];
    } else {
      double* f = static_cast<double*>(p.afrag) + base;
#pragma unroll
      for (int j = 0; j < 8; ++j) f[lane * 8 + j] = vals64[j];
    }
  }
}

// Dense per-level weights w[s][lv][e] = scale f(d_e [lo <= float(d_e) <= hi]) for the
// virtual-visibility path (python/bipp/filter.py:48-51, src/host/virtual_vis.cpp:34-60).
template <typename R>
__global__ void level_weights_kernel(int n_levels, int filter, const double* intervals,
                                     const int* n_red, const R* d, long long stride_d, int n_cols,
                                     const unsigned long long* n_vis, R* out, long long stride_out) {
  const int sid = blockIdx.x;
  const int n = n_red[sid];
  const unsigned long long nv = n_vis[sid];
  const double scale = nv ? 1.0 / static_cast<double>(nv) : 0.0;
  for (int i = threadIdx.x; i < n_levels * n_cols; i += blockDim.x) {
    const int lv = i / n_cols, e = i % n_cols;
    R wv = 0;
    if (e < n) {
      const R dv = d[(size_t)sid * stride_d + e];
      const float df = static_cast<float>(dv);
      const bool m = df >= static_cast<float>(intervals[2 * lv]) &&
                     df <= static_cast<float>(intervals[2 * lv + 1]);
      wv = static_cast<R>(scale) * apply_filter<R>(filter, m ? dv : R(0));
    }
    out[(size_t)sid * stride_out + i] = wv;
  }
}

// --------------------------------------------------------------------------------------
// Main contraction kernel
// --------------------------------------------------------------------------------------
template <int MODE>
struct SteerCfg;
template <>
struct SteerCfg<STEER_BF16X3> {
  static constexpr int NT = 512, TP = 128;
  using Acc = f32x4;
  using Img = float;
};
template <>
struct SteerCfg<STEER_FP32> {
  static constexpr int NT = 512, TP = 128;
  using Acc = f32x4;
  using Img = float;
};
template <>
struct SteerCfg<STEER_FP64> {
  static constexpr int NT = 256, TP = 64;
  using Acc = f64x4;
  using Img = double;
};

template <int MODE>
struct SteerSmem;
template <>
struct SteerSmem<STEER_BF16X3> {
  uint4 eb[2][SteerCfg<STEER_BF16X3>::TP / 16][2][64];
};
template <>
struct SteerSmem<STEER_FP32> {
  float eb[2][32][SteerCfg<STEER_FP32>::TP + 16];
};
template <>
struct SteerSmem<STEER_FP64> {
  double eb[2][32][SteerCfg<STEER_FP64>::TP + 8];
};

template <int MODE>
__global__ __launch_bounds__(SteerCfg<MODE>::NT) void steer_kernel(SteerParams p) {
  using Cfg = SteerCfg<MODE>;
  constexpr int NT = Cfg::NT, TP = Cfg::TP, CT = TP / 16, NWV = NT / 64, RTW = 4;
  constexpr int RT_PASS = NWV * RTW;
  using Acc = typename Cfg::Acc;
  using Img = typename Cfg::Img;

  __shared__ SteerSmem<MODE> sm;
  __shared__ Img img_acc[LMAX][TP];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q0 = blockIdx.x * TP;
  const int L = p.n_levels;

  // generator role: one pixel and one quad of antennas of each 16-antenna chunk
  const int gpx = tid % TP, gquad = tid / TP;  // gquad in [0, 4)
  const int gq = q0 + gpx;
  double lx = 0, ly = 0, lz = 0;
  if (gq < p.n_pix) {
    lx = p.lmn[gq];
    ly = p.lmn[p.n_pix + gq];
    lz = p.lmn[2 * (size_t)p.n_pix + gq];
  }
  const float lxf = static_cast<float>(lx), lyf = static_cast<float>(ly), lzf = static_cast<float>(lz);

  for (int i = tid; i < LMAX * TP; i += NT) (&img_acc[0][0])[i] = 0;

  auto gen = [&](const double* pos, int apad, int kc, int buf) {
    const int a0 = kc * 16 + gquad * 4;
    float cs[4], sn[4];
    double cs64[4], sn64[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int a = a0 + j;
      const double px = pos[a], py = pos[apad + a], pz = pos[2 * apad + a];
      if (MODE == STEER_FP64) {
        const double t = fma(px, lx, fma(py, ly, pz * lz));
        const double f = t - rint(t);
        sincospi(2.0 * f, &sn64[j], &cs64[j]);
      } else if (p.phase_fp32) {
        const float t = fmaf(static_cast<float>(px), lxf,
                             fmaf(static_cast<float>(py), lyf, static_cast<float>(pz) * lzf));
        sincospif(2.0f * t, &sn[j], &cs[j]);
      } else {
        const double t = fma(px, lx, fma(py, ly, pz * lz));
        const float f = static_cast<float>(t - rint(t));
        sincospif(2.0f * f, &sn[j], &cs[j]);
      }
    }
    if constexpr (MODE == STEER_BF16X3) {
      unsigned short hi[8], lo[8];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        hi[2 * j] = bf16_bits(cs[j]);
        hi[2 * j + 1] = bf16_bits(sn[j]);
        lo[2 * j] = bf16_bits(cs[j] - bf16_to_f32(hi[2 * j]));
        lo[2 * j + 1] = bf16_bits(sn[j] - bf16_to_f32(hi[2 * j + 1]));
      }
      uint4 h, l;
      h.x = hi[0] | ((unsigned)hi[1] << 16);
      h.y = hi[2] | ((unsigned)hi[3] << 16);
      h.z = hi[4] | ((unsigned)hi[5] << 16);
      h.w = hi[6] | ((unsigned)hi[7] << 16);
      l.x = lo[0] | ((unsigned)lo[1] << 16);
      l.y = lo[2] | ((unsigned)lo[3] << 16);
      l.z = lo[4] | ((unsigned)lo[5] << 16);
      l.w = lo[6] | ((unsigned)lo[7] << 16);
      const int ct = gpx >> 4, fl = (gpx & 15) + 16 * gquad;
      sm.eb[buf][ct][0][fl] = h;
      sm.eb[buf][ct][1][fl] = l;
    } else if constexpr (MODE == STEER_FP32) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        sm.eb[buf][2 * (gquad * 4 + j)][gpx] = cs[j];
        sm.eb[buf][2 * (gquad * 4 + j) + 1][gpx] = sn[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        sm.eb[buf][2 * (gquad * 4 + j)][gpx] = cs64[j];
        sm.eb[buf][2 * (gquad * 4 + j) + 1][gpx] = sn64[j];
      }
    }
  };

  for (int s = 0; s < p.n_samples; ++s) {
    const int ne = p.n_eig[s];
    const int na = p.n_ant[s];
    if (ne <= 0 || na <= 0) continue;  // uniform over the workgroup
    const int nkc = (na + 15) >> 4;
    const int nrt = (2 * ne + 15) >> 4;
    const double* pos = p.pos + (size_t)s * p.pos_stride;
    const long long fbase = (long long)s * p.afrag_stride;
    for (int rb = 0; rb < nrt; rb += RT_PASS) {
      const int my_rt0 = rb + wave * RTW;
      const int nmy = min(max(nrt - my_rt0, 0), RTW);
      Acc acc[RTW][CT];
#pragma unroll
      for (int r = 0; r < RTW; ++r)
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i) acc[r][c][i] = 0;

      gen(pos, p.apad, 0, 0);
      __syncthreads();
      for (int kc = 0; kc < nkc; ++kc) {
        const int buf = kc & 1;
        if (kc + 1 < nkc) gen(pos, p.apad, kc + 1, buf ^ 1);
        if (nmy > 0) {
          if constexpr (MODE == STEER_BF16X3) {
            const unsigned short* fr = static_cast<const unsigned short*>(p.afrag) + fbase;
            bf16x8 ahi[RTW], alo[RTW];
#pragma unroll
            for (int r = 0; r < RTW; ++r) {
              if (r < nmy) {
                const long long off = ((long long)(my_rt0 + r) * p.kchunks + kc) * 1024;
                ahi[r] = *reinterpret_cast<const bf16x8*>(fr + off + lane * 8);
                alo[r] = *reinterpret_cast<const bf16x8*>(fr + off + 512 + lane * 8);
              }
            }
#pragma unroll
            for (int c = 0; c < CT; ++c) {
              const uint4 bh = sm.eb[buf][c][0][lane];
              const uint4 bl = sm.eb[buf][c][1][lane];
              const bf16x8 bhi = __builtin_bit_cast(bf16x8, bh);
              const bf16x8 blo = __builtin_bit_cast(bf16x8, bl);
#pragma unroll
              for (int r = 0; r < RTW; ++r) {
                if (r < nmy) {
                  acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[r], bhi, acc[r][c], 0, 0, 0);
                  acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r], blo, acc[r][c], 0, 0, 0);
                  acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[r], bhi, acc[r][c], 0, 0, 0);
                }
              }
            }
          } else {
            using T = typename std::conditional<MODE == STEER_FP64, double, float>::type;
            const T* fr = static_cast<const T*>(p.afrag) + fbase;
            T af[RTW][8];
#pragma unroll
            for (int r = 0; r < RTW; ++r) {
              if (r < nmy) {
                const long long off = ((long long)(my_rt0 + r) * p.kchunks + kc) * 512 + lane * 8;
#pragma unroll
                for (int j = 0; j < 8; ++j) af[r][j] = fr[off + j];
              }
            }
#pragma unroll
            for (int c = 0; c < CT; ++c) {
#pragma unroll
              for (int kk = 0; kk < 8; ++kk) {
                const T b = sm.eb[buf][4 * kk + (lane >> 4)][c * 16 + (lane & 15)];
#pragma unroll
                for (int r = 0; r < RTW; ++r) {
                  if (r < nmy) {
                    if constexpr (MODE == STEER_FP64)
                      acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[r][kk], b, acc[r][c], 0, 0, 0);
                    else
                      acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[r][kk], b, acc[r][c], 0, 0, 0);
                  }
                }
              }
            }
          }
        }
        __syncthreads();
      }

      // epilogue: I_l(x) += sum_e w_l[e] (Re^2 + Im^2)
      if (nmy > 0) {
        Img pw0[RTW][CT], pw1[RTW][CT];
#pragma unroll
        for (int r = 0; r < RTW; ++r)
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            pw0[r][c] = static_cast<Img>(acc[r][c][0]) * static_cast<Img>(acc[r][c][0]) +
                        static_cast<Img>(acc[r][c][1]) * static_cast<Img>(acc[r][c][1]);
            pw1[r][c] = static_cast<Img>(acc[r][c][2]) * static_cast<Img>(acc[r][c][2]) +
                        static_cast<Img>(acc[r][c][3]) * static_cast<Img>(acc[r][c][3]);
          }
        const Img* wts = static_cast<const Img*>(p.wts) + (size_t)s * p.w_stride;
        for (int lv = 0; lv < L; ++lv) {
          Img w0[RTW], w1[RTW];
#pragma unroll
          for (int r = 0; r < RTW; ++r) {
            // accumulator registers (0,1) and (2,3) of this lane hold eigenvectors e0, e1
            const int g = lane >> 4;
            const int e0 = (my_rt0 + r) * 8 + (MODE == STEER_FP64 ? g : 2 * g);
            const int e1 = e0 + (MODE == STEER_FP64 ? 4 : 1);
            w0[r] = r < nmy ? wts[(size_t)lv * p.epad + e0] : Img(0);
            w1[r] = r < nmy ? wts[(size_t)lv * p.epad + e1] : Img(0);
          }
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            Img v = 0;
#pragma unroll
            for (int r = 0; r < RTW; ++r) v += w0[r] * pw0[r][c] + w1[r] * pw1[r][c];
            v += __shfl_xor(v, 16, 64);
            v += __shfl_xor(v, 32, 64);
            if (lane < 16) atomicAdd(&img_acc[lv][c * 16 + lane], v);
          }
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  Img* img = static_cast<Img*>(p.img);
  for (int i = tid; i < L * TP; i += NT) {
    const int lv = i / TP, px = i % TP;
    const int q = q0 + px;
    if (q < p.n_pix) img[(size_t)lv * p.img_ld + q] += img_acc[lv][px];
  }
}

}  // namespace

size_t steer_frag_elems_per_sample(int mode, int rtiles, int kchunks) {
  return static_cast<size_t>(rtiles) * kchunks * frag_elems(mode);
}

template <typename R>
void launch_steer_prep(const SteerPrepParams<R>& p, hipStream_t stream) {
  if (p.n_samples <= 0) return;
  if (p.n_levels > LMAX) fail(BIPP_INVALID_PARAMETER_ERROR, "at most 8 energy levels per image batch");
  const size_t shmem = 2 * sizeof(int) * (size_t)p.n_cols;
  steer_prep_kernel<R><<<p.n_samples, 256, shmem, stream>>>(p);
  check_launch("steer_prep_kernel");
}

void launch_steer(const SteerParams& p, int mode, hipStream_t stream) {
  if (p.n_samples <= 0 || p.n_pix <= 0) return;
  if (p.n_levels > LMAX) fail(BIPP_INVALID_PARAMETER_ERROR, "at most 8 energy levels per image batch");
  if (mode == STEER_BF16X3) {
    constexpr int TP = SteerCfg<STEER_BF16X3>::TP;
    steer_kernel<STEER_BF16X3><<<ceil_div(p.n_pix, TP), SteerCfg<STEER_BF16X3>::NT, 0, stream>>>(p);
  } else if (mode == STEER_FP32) {
    constexpr int TP = SteerCfg<STEER_FP32>::TP;
    steer_kernel<STEER_FP32><<<ceil_div(p.n_pix, TP), SteerCfg<STEER_FP32>::NT, 0, stream>>>(p);
  } else {
    constexpr int TP = SteerCfg<STEER_FP64>::TP;
    steer_kernel<STEER_FP64><<<ceil_div(p.n_pix, TP), SteerCfg<STEER_FP64>::NT, 0, stream>>>(p);
  }
  check_launch("steer_kernel");
}

namespace {
// flag[0] = 1 if any sample's solve failed (info != 0), else 0
__global__ void any_failure_kernel(const int* info, int n, int* flag) {
  int bad = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) bad |= info[i] != 0;
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) flag[0] = bad;
}
// img += tmp unless flag[0] is set
template <typename T>
__global__ void commit_image_kernel(const T* tmp, T* img, long long n, const int* flag) {
  if (flag[0]) return;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    img[i] += tmp[i];
}
template <typename T>
__global__ void zero_image_kernel(T* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = T(0);
}
}  // namespace

void launch_commit_image(const int* info, int n_samples, const void* tmp, void* img, long long n,
                         bool dbl, int* flag, hipStream_t stream) {
  any_failure_kernel<<<1, 256, 0, stream>>>(info, n_samples, flag);
  check_launch("any_failure_kernel");
  const unsigned blocks = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 4096));
  if (dbl)
    commit_image_kernel<double><<<blocks, 256, 0, stream>>>(static_cast<const double*>(tmp),
                                                           static_cast<double*>(img), n, flag);
  else
    commit_image_kernel<float><<<blocks, 256, 0, stream>>>(static_cast<const float*>(tmp),
                                                          static_cast<float*>(img), n, flag);
  check_launch("commit_image_kernel");
}

void launch_zero_image(void* p, long long n, bool dbl, hipStream_t stream) {
  const unsigned blocks = static_cast<unsigned>(std::min<long long>((n + 255) / 256, 4096));
  if (dbl)
    zero_image_kernel<double><<<blocks, 256, 0, stream>>>(static_cast<double*>(p), n);
  else
    zero_image_kernel<float><<<blocks, 256, 0, stream>>>(static_cast<float*>(p), n);
  check_launch("zero_image_kernel");
}

template <typename R>
void launch_level_weights(int n_samples, int n_levels, int filter, const double* intervals,
                          const int* n_red, const R* d, long long stride_d, int n_cols,
                          const unsigned long long* n_vis, R* out, long long stride_out,
                          hipStream_t stream) {
  if (n_samples <= 0) return;
  level_weights_kernel<R><<<n_samples, 256, 0, stream>>>(n_levels, filter, intervals, n_red, d,
                                                         stride_d, n_cols, n_vis, out, stride_out);
  check_launch("level_weights_kernel");
}
template void launch_level_weights<float>(int, int, int, const double*, const int*, const float*,
                                          long long, int, const unsigned long long*, float*,
                                          long long, hipStream_t);
template void launch_level_weights<double>(int, int, int, const double*, const int*,
                                           const double*, long long, int,
                                           const unsigned long long*, double*, long long,
                                           hipStream_t);

template void launch_steer_prep<float>(const SteerPrepParams<float>&, hipStream_t);
template void launch_steer_prep<double>(const SteerPrepParams<double>&, hipStream_t);

}  // namespace bippamd
