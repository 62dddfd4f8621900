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
#include <type_traits>

namespace bippamd {

namespace {

constexpr int SP_BIN = 8;    // spread block edge: one wavefront per 8^3 block
constexpr int SP_LG = 4;     // levels per spread launch (sums kept in registers)
constexpr int SP_WMAX = 16;  // largest kernel width
constexpr int SP_MAXR = 9;   // neighbour bin ranges per block: (nback + 1)^2 with nback <= 2
constexpr int FFT_NT = 512;  // FFT workgroup: ~2 radix-4 butterflies per thread per stage

template <typename R>
__device__ __forceinline__ R es_kernel(R z, R beta) {
  const R t = R(1) - z * z;
  return t > R(0) ? exp(beta * (sqrt(t) - R(1))) : R(0);
}
// float: the exponent beta (sqrt(t) - 1) lies in [-beta, 0] (beta <= 2.38 * 16), so the
// hardware base-2 exponential and square root (about 1 ulp each, normal arguments only) need
// no range reduction or denormal scaling; their error is below the ~1e-6 relative error that
// rounding the exponent itself already gives.
template <>
__device__ __forceinline__ float es_kernel<float>(float z, float beta) {
  const float t = 1.0f - z * z;
  const float e = __builtin_amdgcn_exp2f(beta * 1.44269504088896341f *
                                         (__builtin_amdgcn_sqrtf(fmaxf(t, 0.0f)) - 1.0f));
  return t > 0.0f ? e : 0.0f;
}

// phi(xi) (the Fourier transform of the kernel) from its table by the cubic through the four
// nearest entries; phi is even and the table starts one step below zero.
__device__ __forceinline__ double ft_lookup(const double* tab, int n, double inv_delta,
                                            double xi) {
  const double u = fabs(xi) * inv_delta + 1.0;
  int k = static_cast<int>(u);
  k = min(max(k, 1), n - 3);
  const double t = u - k;
  const double tm1 = t - 1.0, tm2 = t - 2.0, tp1 = t + 1.0;
  return tab[k - 1] * (-t * tm1 * tm2 / 6.0) + tab[k] * (tp1 * tm1 * tm2 / 2.0) +
         tab[k + 1] * (-tp1 * t * tm2 / 2.0) + tab[k + 2] * (tp1 * t * tm1 / 6.0);
}

// --------------------------------------------------------------------------------------
// Points of a slot (sample s, slot t): whether the slot carries a point, its u (radians per
// unit direction cosine), the offset of its value in the sample's VV matrix (lower
// triangle), whether that value is conjugated, and its multiplicity.
// --------------------------------------------------------------------------------------
struct SlotPoint {
  double u[3];
  int src;
  int conj;
  float mult;
};

__device__ __forceinline__ bool slot_point(const N3Source& p, int s, long long t, SlotPoint& sp) {
  const int n = p.n_rows ? p.n_rows[s] : p.n_rows_const;
  if (p.mode == N3_SRC_PAIRS) {
    // column-major lower triangle: t = T(b) + (a - b), T(b) = b (2 N - b + 1) / 2
    const long long N = p.n_ant_max;
    const double q = 2.0 * N + 1.0;
    int b = static_cast<int>(floor(0.5 * (q - sqrt(q * q - 8.0 * static_cast<double>(t)))));
    auto T = [&](long long c) { return c * (2 * N - c + 1) / 2; };
    if (b < 0) b = 0;
    if (b > N - 1) b = static_cast<int>(N - 1);
    while (b > 0 && T(b) > t) --b;          // guard rounding of the square root
    while (b < N - 1 && T(b + 1) <= t) ++b;
    const int a = static_cast<int>(t - T(b)) + b;
    if (a >= n || (p.merge_auto && a == b && a > 0)) return false;
    const double* pos = p.pos + (size_t)s * p.pos_stride;  // [3][apad] positions / wavelength
#pragma unroll
    for (int d = 0; d < 3; ++d)
      sp.u[d] = 6.283185307179586 * (pos[d * p.apad + a] - pos[d * p.apad + b]);
    // the merged a == b point reads the trace slot (0, 1)
    sp.src = (p.merge_auto && a == b) ? static_cast<int>(p.ld_vv) : static_cast<int>(a + p.ld_vv * b);
    sp.conj = 0;
    sp.mult = a == b ? 1.f : 2.f;  // (a,b) and (b,a) are conjugate: 2 Re(.)
    return true;
  }
  const int N = p.n_ant_max;
  const int j = static_cast<int>(t / N), i = static_cast<int>(t - (long long)j * N);
  if (i >= n || j >= n) return false;
  const bool anti = p.antisym[s] != 0;
  if (anti && i < j) return false;                           // stands with its mirror (j, i)
  if (anti && p.merge_auto && i == j && i > 0) return false;  // diagonal merged at (0, 0)
  const float* uv = p.uvw + (size_t)s * p.uvw_stride + (i + (long long)N * j);
  const double sc = p.uvw_scale[s];
#pragma unroll
  for (int d = 0; d < 3; ++d) sp.u[d] = sc * static_cast<double>(uv[d * p.lduvw]);
  if (anti && p.merge_auto && i == j) {
    sp.src = static_cast<int>(p.ld_vv);
    sp.conj = 0;
  } else if (i >= j) {
    sp.src = static_cast<int>(i + p.ld_vv * j);
    sp.conj = 0;
  } else {
    sp.src = static_cast<int>(j + p.ld_vv * i);  // VV(i, j) = conj(VV(j, i))
    sp.conj = 1;
  }
  sp.mult = (anti && i != j) ? 2.f : 1.f;
  return true;
}

// Point records: one thread per listed slot (or per slot of every sample); counts per bin
// and each point's rank in its bin (fill = 0), or the records and values written into their
// bins' sorted ranges at those ranks (fill = 1).  Both launches compute a point's bin from
// the same operations.
__global__ void n3_points_kernel(N3PointsParams p) {
  const long long sps = p.src.slots_per_sample();
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= p.n_list) return;
  const unsigned long long slot = p.slot_list ? p.slot_list[i] : static_cast<unsigned long long>(i);
  const int s = static_cast<int>(slot / sps);
  const long long t = static_cast<long long>(slot - (unsigned long long)s * sps);
  SlotPoint sp;
  if (!slot_point(p.src, s, t, sp)) return;
  double cell[3];
  N3Point pt;
  int bin[3];
  double ph = 0;
  for (int d = 0; d < 3; ++d) {
    const double up = sp.u[d] - p.C[d];
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
  const int binid = (bin[2] * p.nbins[1] + bin[1]) * p.nbins[0] + bin[0];
  if (!p.fill) {
    p.rank[i] = atomicAdd(&p.bin_count[binid], 1u);
    return;
  }
  const unsigned dst = p.bin_start[binid] + p.rank[i];
  double ps, pc;
  if (p.phase_fp32) {
    float fs, fc;
    const float fph = static_cast<float>(sp.u[0] - p.C[0]) * static_cast<float>(p.D[0]) +
                      static_cast<float>(sp.u[1] - p.C[1]) * static_cast<float>(p.D[1]) +
                      static_cast<float>(sp.u[2] - p.C[2]) * static_cast<float>(p.D[2]);
    sincosf(fph, &fs, &fc);
    ps = fs;
    pc = fc;
  } else {
    const double turns = ph * 0.15915494309189535;
    const double f = turns - rint(turns);
    sincospi(2.0 * f, &ps, &pc);
  }
  // the multiplicity (1 or 2) scales the pre-phase exactly
  pt.ph[0] = static_cast<float>(pc) * sp.mult;
  pt.ph[1] = static_cast<float>(ps) * sp.mult;
  p.pts[dst] = pt;
  // the point's values of every level, conjugated for an upper-triangle entry
  const float sgn = sp.conj ? -1.f : 1.f;
  if (p.vv_double) {
    const cplx<double>* v = static_cast<const cplx<double>*>(p.vv) + (size_t)s * p.n_levels * p.vv_stride + sp.src;
    cplx<double>* o = static_cast<cplx<double>*>(p.vals) + (size_t)dst * p.n_levels;
    for (int lv = 0; lv < p.n_levels; ++lv) {
      const cplx<double> c = v[(size_t)lv * p.vv_stride];
      o[lv] = mk<double>(c.re, sgn * c.im);
    }
  } else {
    const cplx<float>* v = static_cast<const cplx<float>*>(p.vv) + (size_t)s * p.n_levels * p.vv_stride + sp.src;
    cplx<float>* o = static_cast<cplx<float>*>(p.vals) + (size_t)dst * p.n_levels;
    for (int lv = 0; lv < p.n_levels; ++lv) {
      const cplx<float> c = v[(size_t)lv * p.vv_stride];
      o[lv] = mk<float>(c.re, sgn * c.im);
    }
  }
}

// One workgroup per sample: is its uvw antisymmetric with a zero diagonal?  (Exact float
// comparisons: baselines formed as p_i - p_j satisfy it exactly.)
__global__ __launch_bounds__(256) void uvw_antisym_kernel(N3Source p, unsigned char* out) {
  const int s = blockIdx.x;
  const int N = p.n_ant_max;
  const float* uv = p.uvw + (size_t)s * p.uvw_stride;
  int bad = 0;
  for (long long t = threadIdx.x; t < (long long)N * N && !bad; t += blockDim.x) {
    const int j = static_cast<int>(t / N), i = static_cast<int>(t - (long long)j * N);
    if (i < j) continue;
    for (int d = 0; d < 3; ++d) {
      const float a = uv[i + (long long)N * j + d * p.lduvw];
      const float b = uv[j + (long long)N * i + d * p.lduvw];
      if (i == j ? a != 0.0f : a != -b) bad = 1;
    }
  }
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) out[s] = bad ? 0 : 1;
}

// min / max of u over the points of item k (a sample, or a listed range of slots).
__global__ __launch_bounds__(256) void slot_range_kernel(N3Source p,
                                                         const unsigned long long* list,
                                                         const unsigned* begin,
                                                         const unsigned* count, double* out) {
  __shared__ double red[2][3][256];
  const int k = blockIdx.x;
  const long long sps = p.slots_per_sample();
  const long long n = list ? static_cast<long long>(count[k]) : sps;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const unsigned long long slot =
        list ? list[static_cast<long long>(begin[k]) + i] : (unsigned long long)k * sps + i;
    const int s = static_cast<int>(slot / sps);
    SlotPoint sp;
    if (!slot_point(p, s, static_cast<long long>(slot - (unsigned long long)s * sps), sp)) continue;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], sp.u[d]);
      hi[d] = fmax(hi[d], sp.u[d]);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    red[0][d][threadIdx.x] = lo[d];
    red[1][d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        red[0][d][threadIdx.x] = fmin(red[0][d][threadIdx.x], red[0][d][threadIdx.x + off]);
        red[1][d][threadIdx.x] = fmax(red[1][d][threadIdx.x], red[1][d][threadIdx.x + off]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    out[6 * (size_t)k + threadIdx.x] = red[0][threadIdx.x][0];
    out[6 * (size_t)k + 3 + threadIdx.x] = red[1][threadIdx.x][0];
  }
}

// Box of every slot's point (-1 without a point) and the point count of every box.
__global__ void n3_partition_kernel(N3PartParams p) {
  const long long sps = p.src.slots_per_sample();
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= sps * p.n_samples) return;
  const int s = static_cast<int>(i / sps);
  SlotPoint sp;
  if (!slot_point(p.src, s, i - (long long)s * sps, sp)) {
    p.group_of[i] = -1;
    return;
  }
  int g = 0;
  for (int d = 2; d >= 0; --d) {
    const double x = p.inv[d] * (sp.u[d] - p.lo[d]);
    const int k = x > 0.0 ? static_cast<int>(fmin(x, static_cast<double>(p.dims[d] - 1))) : 0;
    g = g * p.dims[d] + k;
  }
  p.group_of[i] = g;
  atomicAdd(&p.group_count[g], 1u);
}

// slot ids in box order (counting sort)
__global__ void n3_group_scatter_kernel(long long n_slots, const int* group_of,
                                        const unsigned* group_start, unsigned* group_fill,
                                        unsigned long long* list) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n_slots) return;
  const int g = group_of[i];
  if (g < 0) return;
  list[group_start[g] + atomicAdd(&group_fill[g], 1u)] = static_cast<unsigned long long>(i);
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


// --------------------------------------------------------------------------------------
// Spreading with output ownership.  One workgroup owns an 8^3 block (cells l + nf1/2 in
// [8 b, 8 b + 8) per axis); in each of its wavefronts lane t owns the x-row
// (y = t mod 8, z = t / 8) and keeps its 8 cells for SP_LG levels in registers.  The points of
// the bins that can reach the block (a point of bin c covers cells [l0, l0 + w), l0 in bin c)
// are taken 64 at a time, batch k by wavefront k mod (wavefronts): each point that meets the block
// is compacted into the wavefront's LDS buffers with its kernel weights (the x weights
// expanded to the block's 8 columns, zero outside the support) and its values; then every
// lane whose row lies inside a point's y/z support adds it to its 8 cells.  The wavefronts'
// sums are reduced through LDS and the block is written once, times the deconvolution
// factors, with plain stores: no atomics on the grid, and every cell of [-nf1/2, nf1/2)^3
// written by exactly one workgroup (no zero pass over the grid).  The sums follow the order
// of the points inside their bins, which the counting pass ranks in arrival order (the
// values an atomic counter returns): float images agree between runs to rounding, not bit
// for bit.
// --------------------------------------------------------------------------------------
template <typename R>
struct SpreadBatch {
  R pwx[64][SP_BIN];      // kernel weight of each block column (x), row (y) and slab (z),
  R pwy[64][SP_BIN];      // zero outside the point's support
  R pwz[64][SP_BIN];
  R pv[64][SP_LG][2];     // values of the levels of this launch
};
static_assert(SP_BIN == 8 && SP_LG * 2 == 8, "batch rows are moved as 8 values");

// Rows of 8 values moved with 16-byte LDS accesses: a wavefront's row stores (rows 32 or 64
// bytes apart) conflict at most 2-way instead of 8-way for 4-byte stores, and a row read
// is two or four instructions.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_row8(float* dst, const float (&v)[8]) {
  reinterpret_cast<f32x4*>(dst)[0] = f32x4{v[0], v[1], v[2], v[3]};
  reinterpret_cast<f32x4*>(dst)[1] = f32x4{v[4], v[5], v[6], v[7]};
}
__device__ __forceinline__ void st_row8(double* dst, const double (&v)[8]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) reinterpret_cast<f64x2*>(dst)[k] = f64x2{v[2 * k], v[2 * k + 1]};
}
__device__ __forceinline__ void ld_row8(const float* src, float (&v)[8]) {
  const f32x4 a = reinterpret_cast<const f32x4*>(src)[0];
  const f32x4 b = reinterpret_cast<const f32x4*>(src)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
  v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void ld_row8(const double* src, double (&v)[8]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const f64x2 a = reinterpret_cast<const f64x2*>(src)[k];
    v[2 * k] = a.x;
    v[2 * k + 1] = a.y;
  }
}

// wavefronts per spread block (batches split between them); static LDS stays under 64 KiB
template <typename R>
struct SpreadWaves {
  static constexpr int value = sizeof(R) == 8 ? 2 : 4;
};

template <typename R, typename TV>
__global__ __launch_bounds__(256) void n3_spread_kernel(N3SpreadParams<R> p) {
  constexpr int B = SP_BIN, PB = 64, NWV = SpreadWaves<R>::value;
  // per-wavefront batch buffers; reused at the end for the reduction of the row sums
  __shared__ __align__(16) union {
    SpreadBatch<R> batch[NWV];
    R red[NWV - 1][64][(B / 2) * SP_LG * 2];  // half of a lane's cells at a time
  } sh;
  const int w = p.w;
  const int L = p.n_levels;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  SpreadBatch<R>& sb = sh.batch[wave];
  // Workgroups go to the 8 XCDs in turn (workgroup i to XCD i mod 8).  With xcd_chunk > 0,
  // runs of xcd_chunk consecutive blocks go to one XCD (neighbouring blocks read overlapping
  // point ranges and then share an L2), the runs dealt to the XCDs in turn.
  int bin = static_cast<int>(blockIdx.x);
  if (p.xcd_chunk > 0) {
    const int x = bin % 8, k = bin / 8;
    bin = ((k / p.xcd_chunk) * 8 + x) * p.xcd_chunk + k % p.xcd_chunk;
  }
  if (bin >= p.nbins[0] * p.nbins[1] * p.nbins[2]) return;
  const int bx = bin % p.nbins[0];
  const int by = (bin / p.nbins[0]) % p.nbins[1];
  const int bz = bin / (p.nbins[0] * p.nbins[1]);
  const int org[3] = {bx * B, by * B, bz * B};
  const int cy = lane & 7, cz = lane >> 3;
  R acc[B][SP_LG][2];
#pragma unroll
  for (int c = 0; c < B; ++c)
#pragma unroll
    for (int lv = 0; lv < SP_LG; ++lv) acc[c][lv][0] = acc[c][lv][1] = R(0);
  const unsigned long long t_start = p.prof ? wall_clock64() : 0;
  unsigned long long n_hit = 0;
  const R two_over_w = R(2) / R(w);
  const R beta = R(p.beta);
  // The bins that can reach the block: for each (cy, cz) of the neighbourhood the bins
  // cx = bx - nback .. bx are consecutive in the sorted point array, so the neighbourhood is
  // at most (nback + 1)^2 contiguous ranges; batches run over their concatenation.
  unsigned rbeg[SP_MAXR], rlen[SP_MAXR];
  int nr = 0;
  unsigned total_pts = 0;
  for (int cz_ = max(0, bz - p.nback[2]); cz_ <= bz; ++cz_)
    for (int cy_ = max(0, by - p.nback[1]); cy_ <= by; ++cy_) {
      const int row = (cz_ * p.nbins[1] + cy_) * p.nbins[0];
      const int lo = row + max(0, bx - p.nback[0]), hi = row + bx;
      rbeg[nr] = p.bin_start[lo];
      rlen[nr] = p.bin_start[hi] + p.bin_count[hi] - rbeg[nr];
      total_pts += rlen[nr];
      ++nr;
    }
  const unsigned n_batches = (total_pts + PB - 1) / PB;
  // wavefronts that own at least one batch (the others reach the barriers below idle)
  const int n_active = static_cast<int>(min<unsigned>(n_batches, NWV));
  for (unsigned bi = wave; bi < n_batches; bi += NWV) {
    const unsigned b0 = bi * PB;
    const int nb = min<unsigned>(PB, total_pts - b0);
    bool hit = false;
    N3Point pt;
    int rel[3] = {0, 0, 0};
    TV vals[SP_LG];
    if (lane < nb) {
      unsigned v = b0 + lane;
      int r = 0;
      while (v >= rlen[r]) v -= rlen[r++];
      pt = p.pts[rbeg[r] + v];
      hit = true;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        rel[d] = pt.l0[d] - org[d];
        hit = hit && rel[d] < B && rel[d] + w > 0;
      }
      if (hit) {
        // values first: their latency overlaps the compaction and the kernel weights
        const TV* pv = p.vals + (size_t)(rbeg[r] + v) * p.n_levels_total + p.lv0;
#pragma unroll
        for (int lv = 0; lv < SP_LG; ++lv)
          if (lv < L) vals[lv] = pv[lv];
      }
    }
    const unsigned long long mask = __ballot(hit);
    const int total = __popcll(mask);
    const int slot = __popcll(mask & ((1ull << lane) - 1ull));
    wave_sync();  // the previous batch of this wavefront fully consumed
    if (hit) {
      R row[3][B];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        // column c is support offset o = c - rel of the point, at (o - fr) 2 / w: the
        // argument stepped from c = 0, the support test 0 <= o < w as one unsigned compare
        const R x0 = (R(-rel[d]) - R(pt.fr[d])) * two_over_w;
#pragma unroll
        for (int c = 0; c < B; ++c)
          row[d][c] = static_cast<unsigned>(c - rel[d]) < static_cast<unsigned>(w)
                          ? es_kernel<R>(x0 + R(c) * two_over_w, beta)
                          : R(0);
      }
      st_row8(sb.pwx[slot], row[0]);
      st_row8(sb.pwy[slot], row[1]);
      st_row8(sb.pwz[slot], row[2]);
      const R pr = R(pt.ph[0]), pim = R(pt.ph[1]);  // multiplicity included
      R pvr[SP_LG * 2];
#pragma unroll
      for (int lv = 0; lv < SP_LG; ++lv) {
        if (lv < L) {
          const TV c = vals[lv];
          pvr[2 * lv] = R(c.re) * pr - R(c.im) * pim;
          pvr[2 * lv + 1] = R(c.re) * pim + R(c.im) * pr;
        } else {
          pvr[2 * lv] = pvr[2 * lv + 1] = R(0);
        }
      }
      st_row8(&sb.pv[slot][0][0], pvr);
    }
    wave_sync();
    n_hit += total;
    // Branch-free: rows outside a point's support take weight 0 (the expanded y and z
    // weights are zero there); two points per iteration keep two load chains in flight.
    auto row_weight = [&](int q) { return sb.pwy[q][cy] * sb.pwz[q][cz]; };
    auto add_point = [&](int q, R wyz) {
      R vr[SP_LG], vi[SP_LG], pvq[SP_LG * 2], wxq[B];
      ld_row8(&sb.pv[q][0][0], pvq);
      ld_row8(sb.pwx[q], wxq);
#pragma unroll
      for (int lv = 0; lv < SP_LG; ++lv) {
        vr[lv] = pvq[2 * lv] * wyz;
        vi[lv] = pvq[2 * lv + 1] * wyz;
      }
#pragma unroll
      for (int cx = 0; cx < B; ++cx) {
        const R wx = wxq[cx];
#pragma unroll
        for (int lv = 0; lv < SP_LG; ++lv) {
          acc[cx][lv][0] += wx * vr[lv];
          acc[cx][lv][1] += wx * vi[lv];
        }
      }
    };
    int q = 0;
    for (; q + 1 < total; q += 2) {
      const R w0 = row_weight(q), w1 = row_weight(q + 1);
      add_point(q, w0);
      add_point(q + 1, w1);
    }
    if (q < total) add_point(q, row_weight(q));
  }
  // ---- reduce the wavefronts' row sums into wavefront 0, half a row at a time ----------
  if (n_active > 1) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // every wavefront is done with its batch buffers (h = 0) or wavefront 0 with the
      // previous half (h = 1)
      __syncthreads();
      if (wave > 0 && wave < n_active) {
        R* dst = sh.red[wave - 1][lane];
#pragma unroll
        for (int c = 0; c < B / 2; ++c)
#pragma unroll
          for (int lv = 0; lv < SP_LG; ++lv) {
            dst[(c * SP_LG + lv) * 2 + 0] = acc[h * (B / 2) + c][lv][0];
            dst[(c * SP_LG + lv) * 2 + 1] = acc[h * (B / 2) + c][lv][1];
          }
      }
      __syncthreads();
      for (int k = 0; k < n_active - 1 && wave == 0; ++k) {
        const R* src = sh.red[k][lane];
#pragma unroll
        for (int c = 0; c < B / 2; ++c)
#pragma unroll
          for (int lv = 0; lv < SP_LG; ++lv) {
            acc[h * (B / 2) + c][lv][0] += src[(c * SP_LG + lv) * 2 + 0];
            acc[h * (B / 2) + c][lv][1] += src[(c * SP_LG + lv) * 2 + 1];
          }
      }
    }
  }
  if (wave > 0) return;  // no barrier below: wavefront 0 writes the block
  if (p.prof && lane == 0) {
    unsigned long long* pr = p.prof + (size_t)bin * 4;
    pr[0] = t_start;
    pr[1] = wall_clock64();
    pr[2] = total_pts;
    pr[3] = n_hit;  // hits of wavefront 0
  }
  // ---- deconvolve and store the owned cells inside [-nf1/2, nf1/2)^3 ------------------
  const int iy = org[1] + cy, iz = org[2] + cz;
  if (iy >= p.nf1[1] || iz >= p.nf1[2]) return;
  int gy = iy - p.nf1[1] / 2, gz = iz - p.nf1[2] / 2;
  gy += gy < 0 ? p.nf2[1] : 0;
  gz += gz < 0 ? p.nf2[2] : 0;
  const R fyz = p.fy[iy] * p.fz[iz];
#pragma unroll
  for (int cx = 0; cx < B; ++cx) {
    const int ix = org[0] + cx;
    if (ix >= p.nf1[0]) continue;
    const R f = p.fx[ix] * fyz;
    int gx = ix - p.nf1[0] / 2;
    gx += gx < 0 ? p.nf2[0] : 0;
    cplx<R>* g = p.grid + ((size_t)gz * p.nf2[1] + gy) * p.nf2[0] + gx;
#pragma unroll
    for (int lv = 0; lv < SP_LG; ++lv)
      if (lv < L)
        g[(size_t)(p.lv0 - p.lv_begin + lv) * p.grid_stride] =
            mk<R>(acc[cx][lv][0] * f, acc[cx][lv][1] * f);
  }
}

// --------------------------------------------------------------------------------------
// Stockham FFT in LDS (sign +): `nl` lines of length n, line stride ld, radices r[].
// Returns the buffer holding the result.
// --------------------------------------------------------------------------------------
// exp(+2 pi i m / N) for the constant twiddles of the composite radices (m is a
// compile-time constant after unrolling, so the selection folds away).
template <typename R, int N>
__device__ __forceinline__ cplx<R> root_const(int m) {
  constexpr double C8 = 0.923879532511286756128, S8 = 0.382683432365089771728,
                   H2 = 0.707106781186547524401, H3 = 0.866025403784438646764;
  double c = 1, s = 0;
  if constexpr (N == 16) {
    constexpr double tc[10] = {1, C8, H2, S8, 0, -S8, -H2, -C8, -1, -C8};
    constexpr double ts[10] = {0, S8, H2, C8, 1, C8, H2, S8, 0, -S8};
    c = tc[m];
    s = ts[m];
  } else if constexpr (N == 9) {
    constexpr double tc[5] = {1, 0.766044443118978035202, 0.173648177666930348852, -0.5,
                              -0.939692620785908384054};
    constexpr double ts[5] = {0, 0.642787609686539326323, 0.984807753012208059367, H3,
                              0.342020143325668733044};
    c = tc[m];
    s = ts[m];
  } else if constexpr (N == 8) {
    constexpr double tc[4] = {1, H2, 0, -H2};
    constexpr double ts[4] = {0, H2, 1, H2};
    c = tc[m];
    s = ts[m];
  } else {
    static_assert(N == 6, "roots of the composite radices");
    constexpr double tc[3] = {1, 0.5, -0.5};
    constexpr double ts[3] = {0, H3, H3};
    c = tc[m];
    s = ts[m];
  }
  return mk<R>(static_cast<R>(c), static_cast<R>(s));
}

template <typename R, int RAD>
__device__ __forceinline__ void dft_small(cplx<R>* v);

// DFT of length N = P Q (sign +) in registers, Cooley-Tukey: input j = Q j1 + j2, output
// k = k1 + P k2, X[k] = sum_j2 w_N^(j2 k1) w_Q^(j2 k2) sum_j1 x[Q j1 + j2] w_P^(j1 k1).
template <typename R, int P, int Q>
__device__ __forceinline__ void dft_ct(cplx<R>* v) {
  constexpr int N = P * Q;
  cplx<R> f[N];
#pragma unroll
  for (int j2 = 0; j2 < Q; ++j2) {
    cplx<R> t[P];
#pragma unroll
    for (int j1 = 0; j1 < P; ++j1) t[j1] = v[Q * j1 + j2];
    dft_small<R, P>(t);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      const int m = (j2 * k1) % N;
      f[j2 * P + k1] = m == 0 ? t[k1] : t[k1] * root_const<R, N>(m);
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) {
    cplx<R> t[Q];
#pragma unroll
    for (int j2 = 0; j2 < Q; ++j2) t[j2] = f[j2 * P + k1];
    dft_small<R, Q>(t);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = t[k2];
  }
}

// Small DFTs with sign +, constant roots.
template <typename R, int RAD>
__device__ __forceinline__ void dft_small(cplx<R>* v) {
  if constexpr (RAD == 1) {
  } else if constexpr (RAD == 2) {
    const cplx<R> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (RAD == 4) {
    const cplx<R> a = v[0], b = v[1], c = v[2], d = v[3];
    const cplx<R> s0 = a + c, s1 = a - c, s2 = b + d, s3 = b - d;
    const cplx<R> is3 = mk<R>(-s3.im, s3.re);  // +i * s3
    v[0] = s0 + s2;
    v[1] = s1 + is3;
    v[2] = s0 - s2;
    v[3] = s1 - is3;
  } else if constexpr (RAD == 3) {
    // w = exp(+2 pi i / 3): y1,2 = a - (b + c)/2 +- i sqrt(3)/2 (b - c)
    const R h = R(0.86602540378443864676);
    const cplx<R> a = v[0], t1 = v[1] + v[2], t2 = v[1] - v[2];
    const cplx<R> m = mk<R>(a.re - R(0.5) * t1.re, a.im - R(0.5) * t1.im);
    const cplx<R> q = mk<R>(-h * t2.im, h * t2.re);
    v[0] = a + t1;
    v[1] = m + q;
    v[2] = m - q;
  } else if constexpr (RAD == 6) {
    dft_ct<R, 2, 3>(v);
  } else if constexpr (RAD == 8) {
    dft_ct<R, 2, 4>(v);
  } else if constexpr (RAD == 9) {
    dft_ct<R, 3, 3>(v);
  } else if constexpr (RAD == 16) {
    dft_ct<R, 4, 4>(v);
  } else {
    static_assert(RAD == 5, "radix 2, 3, 4, 5, 6, 8, 9 or 16");
    // radix 5, roots w^k = exp(+2 pi i k / 5) = c_k + i s_k with w^4 = conj(w), w^3 =
    // conj(w^2): with t1 = x1 + x4, t2 = x2 + x3, t3 = x1 - x4, t4 = x2 - x3,
    //   y1, y4 = x0 + c1 t1 + c2 t2 +- i (s1 t3 + s2 t4)
    //   y2, y3 = x0 + c2 t1 + c1 t2 +- i (s2 t3 - s1 t4)
    // (36 real operations instead of the 80 of the direct sums)
    const R c1 = R(0.30901699437494742410), c2 = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    const cplx<R> x0 = v[0];
    const cplx<R> t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
    const cplx<R> a1 = mk<R>(x0.re + c1 * t1.re + c2 * t2.re, x0.im + c1 * t1.im + c2 * t2.im);
    const cplx<R> a2 = mk<R>(x0.re + c2 * t1.re + c1 * t2.re, x0.im + c2 * t1.im + c1 * t2.im);
    const cplx<R> b1 = mk<R>(s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im);
    const cplx<R> b2 = mk<R>(s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im);
    v[0] = x0 + t1 + t2;
    v[1] = mk<R>(a1.re - b1.im, a1.im + b1.re);
    v[4] = mk<R>(a1.re + b1.im, a1.im - b1.re);
    v[2] = mk<R>(a2.re - b2.im, a2.im + b2.re);
    v[3] = mk<R>(a2.re + b2.im, a2.im - b2.re);
  }
}

// exp(+2 pi i t / n) from two LDS tables: t1[t mod 64] * t2[t / 64]
template <typename R>
struct Twid {
  const cplx<R>* t1;
  const cplx<R>* t2;
  __device__ __forceinline__ cplx<R> operator()(int t) const { return t1[t & 63] * t2[t >> 6]; }
};

// table layout in global memory: [64] fine entries, then [ceil(n / 64)] coarse entries
__host__ __device__ __forceinline__ int twid_len(int n) { return 64 + (n + 63) / 64; }

template <typename R>
__device__ __forceinline__ Twid<R> load_twid(const cplx<R>* __restrict__ g, int n, cplx<R>* lds,
                                             int tid, int nt) {
  const int len = twid_len(n);
  for (int i = tid; i < len; i += nt) lds[i] = g[i];
  return Twid<R>{lds, lds + 64};
}

// x / d for 0 <= x < 2^31 from the multiplier and shift of d (N3Radix)
__device__ __forceinline__ int fast_div(int x, unsigned m, int s) {
  return static_cast<int>((__umulhi(static_cast<unsigned>(x), m) + static_cast<unsigned>(x)) >> s);
}

// One Stockham stage of radix RAD over nl lines (stride ld) of length n; (nb_m, nb_s) and
// (ns_m, ns_s) divide by n / RAD and by ns.  FIRST: the first stage (ns = 1), whose
// twiddles are all 1 (k = 0), so none are formed or applied.  TAB: the twiddles w^(k r) of
// this stage are read from its table (entry (r - 1) ns + k, build_stage_tab) instead of
// being formed from the two-level table.  ONE: a single line (nl = 1), no line index.
template <typename R, int RAD, bool FIRST, bool TAB, bool ONE>
__device__ __forceinline__ void fft_stage(const cplx<R>* in, cplx<R>* out, int n, int nl, int ld,
                                          int ns, const Twid<R>& tw, const cplx<R>* stab,
                                          int tid, int nt, unsigned nb_m, int nb_s,
                                          unsigned ns_m, int ns_s) {
  const int nb = n / RAD;
  for (int t = tid; t < (ONE ? nb : nl * nb); t += nt) {
    const int line = ONE ? 0 : fast_div(t, nb_m, nb_s), j = ONE ? t : t - line * nb;
    const cplx<R>* src = ONE ? in + j : in + (size_t)line * ld + j;
    cplx<R>* dst = ONE ? out : out + (size_t)line * ld;
    cplx<R> v[RAD];
    if constexpr (FIRST) {
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[r] = src[r * nb];
      dft_small<R, RAD>(v);
      dst += j * RAD;
#pragma unroll
      for (int r = 0; r < RAD; ++r) dst[r] = v[r];
      continue;
    }
    const int jq = fast_div(j, ns_m, ns_s);
    const int k = j - jq * ns;
    v[0] = src[0];
    if constexpr (TAB) {
      const cplx<R>* tk = stab + k;  // consecutive lanes read consecutive entries
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[r] = src[r * nb] * tk[(r - 1) * ns];
    } else {
      // twiddles w^r, w = exp(2 pi i k tstride / n): w, w^4 and w^8 from the table, the
      // others as products at most two deep (k r tstride < n for r < RAD)
      const int tstride = n / (ns * RAD);  // twiddle k r / (ns RAD) of the full length
      const int t1 = k * tstride;
      cplx<R> wp[RAD];
      if constexpr (RAD > 1) wp[1] = tw(t1);
      if constexpr (RAD > 2) wp[2] = wp[1] * wp[1];
      if constexpr (RAD > 3) wp[3] = wp[2] * wp[1];
      if constexpr (RAD > 4) {
        wp[4] = tw(4 * t1);
#pragma unroll
        for (int r = 5; r < RAD && r < 8; ++r) wp[r] = wp[4] * wp[r - 4];
      }
      if constexpr (RAD > 8) {
        wp[8] = tw(8 * t1);
#pragma unroll
        for (int r = 9; r < RAD; ++r) wp[r] = wp[8] * wp[r - 8];
      }
#pragma unroll
      for (int r = 1; r < RAD; ++r) v[r] = src[r * nb] * wp[r];
    }
    dft_small<R, RAD>(v);
    dst += jq * ns * RAD + k;
#pragma unroll
    for (int r = 0; r < RAD; ++r) dst[r * ns] = v[r];
  }
}

// Twiddle tables of the stages after the first, one after another: stage st (span ns, radix
// RAD) holds RAD - 1 rows of ns entries, entry (r - 1) ns + k = exp(2 pi i k r tstride / n)
// with tstride = n / (ns RAD) (k r tstride < n); n - r[0] entries in all.
__host__ __device__ __forceinline__ int stage_tab_len(int n, const N3Radix& rad) {
  return rad.count > 1 ? n - rad.r[0] : 0;
}
template <typename R>
__device__ void build_stage_tab(cplx<R>* tab, int n, const N3Radix& rad, const Twid<R>& tw,
                                int tid, int nt) {
  int ns = rad.r[0], off = 0;
  for (int st = 1; st < rad.count; ++st) {
    const int RAD = rad.r[st], rm = RAD - 1, cnt = ns * rm;
    const int tstride = n / (ns * RAD);
    for (int e = tid; e < cnt; e += nt) {
      const int r = e / ns + 1, k = e - (r - 1) * ns;
      tab[off + e] = tw(k * r * tstride);
    }
    off += cnt;
    ns *= RAD;
  }
}

template <typename R, int MAXR, bool FIRST, bool TAB, bool ONE>
__device__ __forceinline__ void fft_stage_radix(int R_, const cplx<R>* in, cplx<R>* out, int n,
                                                int nl, int ld, int ns, const N3Radix& rad, int st,
                                                const Twid<R>& tw, const cplx<R>* stab, int tid,
                                                int nt) {
#define STAGE(RR) fft_stage<R, RR, FIRST, TAB, ONE>(in, out, n, nl, ld, ns, tw, stab, tid, nt, \
                                               rad.nb_m[st], rad.nb_s[st], rad.ns_m[st], rad.ns_s[st])
  switch (R_) {
    case 2: STAGE(2); break;
    case 3: STAGE(3); break;
    case 4: STAGE(4); break;
    case 6:
      if constexpr (MAXR >= 6) STAGE(6);
      break;
    case 8:
      if constexpr (MAXR >= 8) STAGE(8);
      break;
    case 9:
      if constexpr (MAXR >= 9) STAGE(9);
      break;
    case 16:
      if constexpr (MAXR >= 16) STAGE(16);
      break;
    default: STAGE(5); break;
  }
#undef STAGE
}

// Stockham stages in LDS; returns the buffer holding the result.  MAXR: the largest radix
// compiled in (5: radices 2-5 only, which keeps the kernel's register and scalar-register
// footprint small; 16: also the composite radices 6, 8, 9, 16); the launchers pick it from
// the plan.
template <typename R, int MAXR, bool TAB, bool ONE>
__device__ cplx<R>* lds_fft(cplx<R>* a, cplx<R>* b, int n, int nl, int ld, const N3Radix& rad,
                            const Twid<R>& tw, const cplx<R>* stab, int tid, int nt) {
  int ns = 1;
  cplx<R>* in = a;
  cplx<R>* out = b;
  for (int st = 0; st < rad.count; ++st) {
    const int R_ = rad.r[st];
    if (st == 0) {
      fft_stage_radix<R, MAXR, true, false, ONE>(R_, in, out, n, nl, ld, ns, rad, st, tw, stab, tid, nt);
    } else {
      fft_stage_radix<R, MAXR, false, TAB, ONE>(R_, in, out, n, nl, ld, ns, rad, st, tw, stab, tid, nt);
      if (TAB) stab += ns * (R_ - 1);
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
__device__ __forceinline__ bool in_window(int a, N3Window wd, int n) {
  int d = a - wd.start;
  d += d < 0 ? n : 0;
  return d < wd.count;
}

// Elements per thread of one FFT item: items hold at most 4096 elements (the launchers
// check), so 8 per thread of an FFT_NT = 512 workgroup.
constexpr int FFT_PF = 8;

// Contiguous axis (x): `nl` lines per workgroup (lines (level, z, y) of the windows).  The
// next group's inputs are fetched into registers while the current group is transformed.
// TAB: the stages' twiddle tables are built in LDS once per workgroup (build_stage_tab).
// ONE: one line per group (nl = 1): no line index per element, and the stores run over the
// output window only.
template <typename R, int MAXR, bool TAB, bool ONE>
__global__ __launch_bounds__(FFT_NT) void fft_lines_kernel(N3LineParams<R> p, int nl) {
  extern __shared__ __align__(16) unsigned char fsh[];
  __shared__ long long lbase[2][16];
  const int n = p.n;
  const int tid = threadIdx.x;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);
  cplx<R>* b = a + (size_t)nl * n;
  const Twid<R> tw = load_twid<R>(p.tw, n, b + (size_t)nl * n, tid, FFT_NT);
  cplx<R>* stab = b + (size_t)nl * n + twid_len(n);
  if (TAB) {
    __syncthreads();  // the two-level table is in LDS
    build_stage_tab<R>(stab, n, p.rad, tw, tid, FFT_NT);
  }
  const long long n_lines = (long long)p.n_levels * p.zw.count * p.yw.count;
  const long long n_groups = (n_lines + nl - 1) / nl;
  auto set_bases = [&](long long grp, int buf) {
    const long long line0 = grp * nl;
    const int nlg = static_cast<int>(min<long long>(nl, n_lines - line0));
    if (tid < nlg) {
      const long long line = line0 + tid;
      const int yo = static_cast<int>(line % p.yw.count);
      const long long r = line / p.yw.count;
      const int zo = static_cast<int>(r % p.zw.count);
      const int lv = static_cast<int>(r / p.zw.count);
      int y = p.yw.start + yo, z = p.zw.start + zo;
      y -= y >= p.ny ? p.ny : 0;
      z -= z >= p.nz ? p.nz : 0;
      lbase[buf][tid] = (long long)lv * p.grid_stride + ((long long)z * p.ny + y) * n;
    }
  };
  auto fetch = [&](long long grp, int buf, cplx<R>* regs) {
    const int nlg = static_cast<int>(min<long long>(nl, n_lines - grp * nl));
    if (ONE) {
      const cplx<R>* g = p.grid + lbase[buf][0];
#pragma unroll
      for (int k = 0; k < FFT_PF; ++k) {
        const int i = tid + k * FFT_NT;
        if (i < n) regs[k] = in_window(i, p.in, n) ? g[i] : mk<R>(0, 0);
      }
      return;
    }
#pragma unroll
    for (int k = 0; k < FFT_PF; ++k) {
      const int i = tid + k * FFT_NT;
      if (i < nlg * n) {
        const int li = fast_div(i, p.rad.n_m, p.rad.n_s), x = i - li * n;
        regs[k] = in_window(x, p.in, n) ? p.grid[lbase[buf][li] + x] : mk<R>(0, 0);
      }
    }
  };
  long long grp = blockIdx.x;
  if (grp >= n_groups) return;
  cplx<R> cur[FFT_PF], nxt[FFT_PF];
  int cb = 0;
  set_bases(grp, cb);
  __syncthreads();
  fetch(grp, cb, cur);
  for (; grp < n_groups; grp += gridDim.x) {
    const int nlg = static_cast<int>(min<long long>(nl, n_lines - grp * nl));
    const long long gn = grp + gridDim.x;
#pragma unroll
    for (int k = 0; k < FFT_PF; ++k) {
      const int i = tid + k * FFT_NT;
      if (i < nlg * n) a[i] = cur[k];
    }
    if (gn < n_groups) set_bases(gn, cb ^ 1);
    __syncthreads();
    if (gn < n_groups) fetch(gn, cb ^ 1, nxt);  // in flight during the transform
    cplx<R>* res = lds_fft<R, MAXR, TAB, ONE>(a, b, n, nlg, n, p.rad, tw, stab, tid, FFT_NT);
    if (ONE) {
      cplx<R>* g = p.grid + lbase[cb][0];
      for (int e = tid; e < p.out.count; e += FFT_NT) {
        int k = p.out.start + e;  // start < n, count <= n
        k -= k >= n ? n : 0;
        g[k] = res[k];
      }
    } else {
      for (int i = tid; i < nlg * n; i += FFT_NT) {
        const int li = fast_div(i, p.rad.n_m, p.rad.n_s), k = i - li * n;
        if (in_window(k, p.out, n)) p.grid[lbase[cb][li] + k] = res[i];
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < FFT_PF; ++k) cur[k] = nxt[k];
    cb ^= 1;
  }
}

// Strided axis: transforms of length m for TW consecutive x positions at once; element i of
// sub-transform `sub` at axis index sub * sub_stride + i * estride.  The LDS line stride is
// odd so that the TW lines of a load fall into different banks.  (Latency is hidden by
// occupancy: a register prefetch of the next item costs more waves than it saves.)
template <typename R, int MAXR>
__global__ __launch_bounds__(FFT_NT) void fft_pass_kernel(N3PassParams<R> p, int tw_shift) {
  extern __shared__ __align__(16) unsigned char fsh[];
  const int TW = 1 << tw_shift;  // x positions per item (power of two)
  const int ld = p.m | 1;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);
  cplx<R>* b = a + TW * ld;
  cplx<R>* tws = b + TW * ld;
  const Twid<R> tw = load_twid<R>(p.tw, p.m, tws, threadIdx.x, blockDim.x);
  // second table (two-pass twiddles) right after the first; unused when p.twiddle == 0
  cplx<R>* tws2 = tws + twid_len(p.m);
  if (p.twiddle) load_twid<R>(p.tw2, static_cast<int>(p.ntot), tws2, threadIdx.x, blockDim.x);
  const Twid<R> tw2{tws2, tws2 + 64};
  const int n_xt = (p.xw.count + TW - 1) >> tw_shift;
  // the launcher checks that the item count fits in 31 bits
  const int n_items = n_xt * p.n_sub * p.other.count * p.n_levels;
  for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int q1 = fast_div(item, p.xt_m, p.xt_s);
    const int xt = item - q1 * n_xt;
    const int q2 = fast_div(q1, p.sub_m, p.sub_s);
    const int sub = q1 - q2 * p.n_sub;
    const int lv = fast_div(q2, p.oc_m, p.oc_s);
    const int o = q2 - lv * p.other.count;
    int oo = p.other.start + o;
    oo -= oo >= p.no ? p.no : 0;
    if (p.operm_n2 > 1) {
      const int q = fast_div(oo, p.op_m, p.op_s);
      oo = (oo - q * p.operm_n1) * p.operm_n2 + q;
    }
    cplx<R>* base = p.grid + (size_t)lv * p.grid_stride + (size_t)oo * p.outer_stride;
    for (int t = threadIdx.x; t < (p.m << tw_shift); t += blockDim.x) {
      const int xi = t & (TW - 1), i = t >> tw_shift;
      const int j = (xt << tw_shift) + xi;
      cplx<R> v = mk<R>(0, 0);
      if (j < p.xw.count) {
        int x = p.xw.start + j;
        x -= x >= p.nx ? p.nx : 0;
        const int ai = static_cast<int>(sub * p.sub_stride + i * p.estride);
        if (!p.in_check || in_window(ai, p.in, p.nax))
          v = base[(long long)ai * p.axis_stride + x];
      }
      a[xi * ld + i] = v;
    }
    __syncthreads();
    cplx<R>* res = lds_fft<R, MAXR, false, false>(a, b, p.m, TW, ld, p.rad, tw, nullptr, threadIdx.x,
                                           blockDim.x);
    for (int t = threadIdx.x; t < (p.m << tw_shift); t += blockDim.x) {
      const int xi = t & (TW - 1), k = t >> tw_shift;
      const int j = (xt << tw_shift) + xi;
      if (j >= p.xw.count) continue;
      if (p.out_check && !in_window(sub * p.fa + k * p.fb, p.out, p.nax)) continue;
      int x = p.xw.start + j;
      x -= x >= p.nx ? p.nx : 0;
      cplx<R> v = res[xi * ld + k];
      if (p.twiddle) v = v * tw2(sub * k);  // sub < n2, k < n1: sub k < ntot
      const int ai = static_cast<int>(sub * p.sub_stride + k * p.estride);
      base[(long long)ai * p.axis_stride + x] = v;
    }
    __syncthreads();
  }
}

// Strided axis, m = M1 M2 with both factors radices of dft_small: each transform in two
// register stages with one LDS exchange (X[k1 + M1 k2] = sum_c2 w_M2^(c2 k2) w_m^(c2 k1)
// sum_j1 x[c2 + M2 j1] w_M1^(j1 k1)).  Stage 1: thread (x position, c2) loads x[c2 + M2 j1],
// j1 < M1, straight from global memory, transforms them, applies w_m^(c2 k1) and stores row
// k1 M2 + c2 of its line in LDS; stage 2: thread (x position, k1) reads the M2 values of row
// k1, transforms them and stores outputs k1 + M1 k2.  Lanes run over consecutive x
// positions, so global accesses are contiguous and the odd LDS line stride spreads a
// wavefront's accesses over the banks.  Same items, windows, pass twiddle and output
// permutation as fft_pass_kernel.
template <typename R, int M1, int M2>
__global__ __launch_bounds__(FFT_NT) void fft_pass_rr_kernel(N3PassParams<R> p, int tw_shift) {
  extern __shared__ __align__(16) unsigned char fsh[];
  constexpr int M = M1 * M2;
  constexpr int LD = M | 1;
  const int TW = 1 << tw_shift;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(fsh);  // [TW][LD]
  cplx<R>* twm = a + TW * LD;                     // exp(+2 pi i t / M), t < M
  cplx<R>* tws2 = twm + M;                        // two-level table of ntot (pass twiddle)
  const int tid = threadIdx.x;
  for (int t = tid; t < M; t += FFT_NT) twm[t] = p.tw[t & 63] * p.tw[64 + (t >> 6)];
  if (p.twiddle) load_twid<R>(p.tw2, static_cast<int>(p.ntot), tws2, tid, FFT_NT);
  const Twid<R> tw2{tws2, tws2 + 64};
  __syncthreads();
  const int n_xt = (p.xw.count + TW - 1) >> tw_shift;
  const int n_items = n_xt * p.n_sub * p.other.count * p.n_levels;
  for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
    const int q1 = fast_div(item, p.xt_m, p.xt_s);
    const int xt = item - q1 * n_xt;
    const int q2 = fast_div(q1, p.sub_m, p.sub_s);
    const int sub = q1 - q2 * p.n_sub;
    const int lv = fast_div(q2, p.oc_m, p.oc_s);
    const int o = q2 - lv * p.other.count;
    int oo = p.other.start + o;
    oo -= oo >= p.no ? p.no : 0;
    if (p.operm_n2 > 1) {
      const int q = fast_div(oo, p.op_m, p.op_s);
      oo = (oo - q * p.operm_n1) * p.operm_n2 + q;
    }
    cplx<R>* base = p.grid + (size_t)lv * p.grid_stride + (size_t)oo * p.outer_stride;
    const long long sub_off = (long long)sub * p.sub_stride;
    for (int u = tid; u < (M2 << tw_shift); u += FFT_NT) {
      const int xi = u & (TW - 1), c2 = u >> tw_shift;
      const int j = (xt << tw_shift) + xi;
      cplx<R> v[M1];
      if (j < p.xw.count) {
        int x = p.xw.start + j;
        x -= x >= p.nx ? p.nx : 0;
#pragma unroll
        for (int j1 = 0; j1 < M1; ++j1) {
          const int ai = static_cast<int>(sub_off + (long long)(c2 + M2 * j1) * p.estride);
          v[j1] = (!p.in_check || in_window(ai, p.in, p.nax))
                      ? base[(long long)ai * p.axis_stride + x]
                      : mk<R>(0, 0);
        }
      } else {
#pragma unroll
        for (int j1 = 0; j1 < M1; ++j1) v[j1] = mk<R>(0, 0);
      }
      dft_small<R, M1>(v);
      cplx<R>* row = a + xi * LD + c2;
      row[0] = v[0];
#pragma unroll
      for (int k1 = 1; k1 < M1; ++k1) row[k1 * M2] = c2 == 0 ? v[k1] : v[k1] * twm[c2 * k1];
    }
    __syncthreads();
    for (int u = tid; u < (M1 << tw_shift); u += FFT_NT) {
      const int xi = u & (TW - 1), k1 = u >> tw_shift;
      const int j = (xt << tw_shift) + xi;
      if (j >= p.xw.count) continue;
      int x = p.xw.start + j;
      x -= x >= p.nx ? p.nx : 0;
      cplx<R> v[M2];
      const cplx<R>* row = a + xi * LD + k1 * M2;
#pragma unroll
      for (int c2 = 0; c2 < M2; ++c2) v[c2] = row[c2];
      dft_small<R, M2>(v);
#pragma unroll
      for (int k2 = 0; k2 < M2; ++k2) {
        const int k = k1 + M1 * k2;
        if (p.out_check && !in_window(sub * p.fa + k * p.fb, p.out, p.nax)) continue;
        cplx<R> w = v[k2];
        if (p.twiddle) w = w * tw2(sub * k);  // sub < n2, k < n1: sub k < ntot
        const int ai = static_cast<int>(sub_off + (long long)k * p.estride);
        base[(long long)ai * p.axis_stride + x] = w;
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------
// Interpolation at the pixels, correction, real part accumulated into the images.
// --------------------------------------------------------------------------------------
// The kernel width is a template argument so that the per-axis weights and indices stay in
// registers (a run-time width puts them in scratch memory, read in the innermost loop).
template <typename R, typename TI, int W>
__global__ __launch_bounds__(256) void n3_interp_kernel(N3InterpParams<R> p) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= p.n_pix) return;
  constexpr int w = W;
  R wx[W], wy[W], wz[W];
  int kx[W], ky[W], kz[W];
  double corr = 1.0, cphase = 0.0;
  const double xq[3] = {p.lmn[q], p.lmn[p.n_pix + q], p.lmn[2 * (size_t)p.n_pix + q]};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double s = p.gam[d] * (xq[d] - p.D[d]);
    const double a2 = p.h1[d] * s / p.h2[d];
    const int k0 = static_cast<int>(ceil(a2 - 0.5 * w));
    R* wd = d == 0 ? wx : (d == 1 ? wy : wz);
    int* kd = d == 0 ? kx : (d == 1 ? ky : kz);
    // frequency k = (k0 + o) mod nf2 and, for two-pass axes (stored at (k mod n1) n2 +
    // k / n1), its digits (k mod n1, k / n1): one division each for o = 0, then stepped
    // with the wrap-around at nf2 = n1 n2
    const int nf = p.nf2[d], n1 = p.n1[d], n2 = p.n2[d];
    int k = k0 % nf;
    if (k < 0) k += nf;
    int kq = k / n1, kr = k - kq * n1;
#pragma unroll
    for (int o = 0; o < w; ++o) {
      wd[o] = es_kernel<R>(static_cast<R>((k0 + o - a2) * 2.0 / w), R(p.beta2));
      kd[o] = n2 > 1 ? kr * n2 + kq : k;
      ++k;
      ++kr;
      if (kr == n1) {
        kr = 0;
        ++kq;
      }
      if (k == nf) k = kr = kq = 0;
    }
    const double alpha1 = 0.5 * p.w1 * p.h1[d];
    corr *= p.h1[d] * p.h2[d] / (alpha1 * ft_lookup(p.ft, p.n_ft, p.ft_inv_delta, s * alpha1));
    cphase += p.C[d] * xq[d];
  }
  const double turns = cphase * 0.15915494309189535;
  double cs, cc;
  sincospi(2.0 * (turns - rint(turns)), &cs, &cc);
  for (int lv = 0; lv < p.n_levels; ++lv) {
    const cplx<R>* g = p.grid + (size_t)lv * p.grid_stride;
    R sr = 0, si = 0;
#pragma unroll 1
    for (int oz = 0; oz < w; ++oz) {
#pragma unroll 1
      for (int oy = 0; oy < w; ++oy) {
        const cplx<R>* row = g + ((size_t)kz[oz] * p.nf2[1] + ky[oy]) * p.nf2[0];
        const R wyz = wy[oy] * wz[oz];
        R rr = 0, ri = 0;
#pragma unroll
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

// VV_l = Z diag(w_l) Z^H for every (sample, level); Z rows = antennas used (ld ldz).  The
// vector-ALU form, used for float32 eigenvectors (float64 ones take vv_levels_mfma_kernel).
template <typename R, typename TO>
__global__ __launch_bounds__(256) void vv_levels_kernel(N3VVParams<R> p) {
  // 64 x 64 output tile per workgroup (tiles on or below the diagonal), eigenvectors taken
  // VK at a time through LDS; thread (tx, ty) accumulates rows i0 + tx + 16 u and columns
  // j0 + ty + 16 v (u, v < 4), so LDS reads are consecutive over tx and broadcast over ty
  constexpr int VT = 64, VK = 16;
  __shared__ cplx<R> ti[VK][VT + 1];  // Z[i0 + r, e] d_e
  __shared__ cplx<R> tj[VK][VT + 1];  // Z[j0 + c, e]
  const int sl = blockIdx.z;
  const int s = sl / p.n_levels, lv = sl % p.n_levels;
  const int n = p.n_rows ? p.n_rows[s] : p.n_rows_const;
  const int i0 = blockIdx.x * VT, j0 = blockIdx.y * VT;
  if (i0 >= n || j0 > i0) return;  // only tiles touching the lower triangle
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const cplx<R>* Z = p.z + (size_t)s * p.stride_z;
  const R* dw = p.wd + ((size_t)s * p.n_levels + lv) * p.n_cols;
  cplx<R> acc[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[u][v] = mk<R>(0, 0);
  for (int e0 = 0; e0 < p.n_cols; e0 += VK) {
    // eigenvectors outside the level carry weight 0: skip chunks without any (uniform)
    const int ew = e0 + (tid & (VK - 1));
    const R dt = ew < p.n_cols ? dw[ew] : R(0);
    if (!__syncthreads_or(dt != R(0))) continue;
#pragma unroll
    for (int q = 0; q < VK * VT / 256; ++q) {
      const int idx = tid + 256 * q;
      const int k = idx / VT, r = idx % VT, e = e0 + k;
      const bool ok = e < p.n_cols;
      const R de = ok ? dw[e] : R(0);
      ti[k][r] = (ok && i0 + r < n) ? Z[(i0 + r) + (size_t)e * p.ldz] * de : mk<R>(0, 0);
      tj[k][r] = (ok && j0 + r < n) ? Z[(j0 + r) + (size_t)e * p.ldz] : mk<R>(0, 0);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < VK; ++k) {
      cplx<R> a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = ti[k][tx + 16 * u];
#pragma unroll
      for (int v = 0; v < 4; ++v) b[v] = tj[k][ty + 16 * v];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = acc[u][v] + mulc(a[u], b[v]);
    }
    __syncthreads();
  }
  cplx<TO>* out = reinterpret_cast<cplx<TO>*>(p.vv) + (size_t)sl * p.vv_stride;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + tx + 16 * u, j = j0 + ty + 16 * v;
      if (i < n && j < n && i >= j) out[i + (size_t)p.ld_vv * j] = cvt<TO>(acc[u][v]);
    }
}

// vv_levels_kernel in float64 on v_mfma_f64_16x16x4f64 with the operands loaded straight from
// global memory (no LDS, no barriers).  Workgroup = (sample, level, 64 x 64 tile on or below
// the diagonal); wavefront = 32 x 32 quadrant as 2 x 2 blocks of 16 x 16.  The products form
// D[j][i] = sum_e conj(Z[j, e]) (w_e Z[i, e]) = VV[i][j] (A operand: lane l holds row j =
// l % 16, eigenvector l / 16; B operand: eigenvector l / 16, row i = l % 16; result register r:
// j = l / 16 + 4 r, i = l % 16), so one store writes 16 consecutive rows i of a column j.  Only
// the eigenvectors from the level's first to its last nonzero weight are visited, four per
// step, with the next step's operands loaded before the current step's products.  Workgroups
// are numbered so that the tiles and levels of one sample run on one XCD (workgroup x runs on
// XCD x mod 8), which keeps the sample's Z in that XCD's L2.  Held to 128 registers (four
// wavefronts per SIMD; the 2 x 2 blocks' sums stay in vector registers, no spills).
template <typename TO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void
vv_levels_mfma_kernel(N3VVParams<double> p, int n_tiles, int per_xcd, int n_units) {
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  const int unit = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (unit >= n_units) return;
  const int t = unit % n_tiles, sl = unit / n_tiles;
  const int s = sl / p.n_levels, lv = sl % p.n_levels;
  int ti = 0;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int n = p.n_rows ? p.n_rows[s] : p.n_rows_const;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ib = 64 * ti + 32 * (wave & 1), jb = 64 * tj + 32 * (wave >> 1);
  // quadrant outside the matrix or strictly above the diagonal (every i < every j)
  if (ib >= n || jb >= n || jb > ib + 31) return;
  const double* dw = p.wd + ((size_t)s * p.n_levels + lv) * p.n_cols;
  int lo = p.n_cols, hi = -1;
  for (int e = lane; e < p.n_cols; e += 64)
    if (dw[e] != 0.0) {
      lo = e < lo ? e : lo;
      hi = e;
    }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const int ol = __shfl_xor(lo, m), oh = __shfl_xor(hi, m);
    lo = ol < lo ? ol : lo;
    hi = oh > hi ? oh : hi;
  }
  const cplx<double>* Z = p.z + (size_t)s * p.stride_z;
  const int lr = lane & 15, lk = lane >> 4;
  const int jr[2] = {jb + lr, jb + 16 + lr}, ir[2] = {ib + lr, ib + 16 + lr};
  f64x4 dr[2][2], di[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      dr[a][b] = f64x4{0, 0, 0, 0};
      di[a][b] = f64x4{0, 0, 0, 0};
    }
  // P = Z[j, e] (rows j of the two column blocks), Q = w_e Z[i, e] (rows i of the two row
  // blocks) for the eigenvector e = k + l / 16 of this lane; zero outside [lo, hi] and n.
  auto load = [&](int k, cplx<double>* P, cplx<double>* Q) {
    const int e = k + lk;
    const bool ok = e <= hi;
    const double w = ok ? dw[e] : 0.0;
    const cplx<double>* col = Z + (size_t)(ok ? e : 0) * p.ldz;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      P[b] = (ok && jr[b] < n) ? col[jr[b]] : mk<double>(0, 0);
      const cplx<double> z = (ok && ir[b] < n) ? col[ir[b]] : mk<double>(0, 0);
      Q[b] = mk<double>(z.re * w, z.im * w);
    }
  };
  if (lo <= hi) {
    cplx<double> P[2], Q[2];
    load(lo & ~3, P, Q);
    for (int k = lo & ~3; k <= hi; k += 4) {
      cplx<double> Pn[2], Qn[2];
      load(k + 4, Pn, Qn);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          // re: Pr Qr + Pi Qi; im: Pr Qi - Pi Qr
          dr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(P[a].re, Q[b].re, dr[a][b], 0, 0, 0);
          dr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(P[a].im, Q[b].im, dr[a][b], 0, 0, 0);
          di[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(P[a].re, Q[b].im, di[a][b], 0, 0, 0);
          di[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-P[a].im, Q[b].re, di[a][b], 0, 0, 0);
        }
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        P[b] = Pn[b];
        Q[b] = Qn[b];
      }
    }
  }
  cplx<TO>* out = reinterpret_cast<cplx<TO>*>(p.vv) + (size_t)sl * p.vv_stride;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = jb + 16 * a + lk + 4 * r, i = ib + 16 * b + lr;
        if (i < n && j < n && i >= j)
          out[i + (size_t)p.ld_vv * j] = mk<TO>(static_cast<TO>(dr[a][b][r]), static_cast<TO>(di[a][b][r]));
      }
}

// Trace of VV_l for every (sample, level) into the unused upper-triangle slot (0, 1).
template <typename TO>
__global__ __launch_bounds__(256) void vv_trace_kernel(int n_levels, const int* n_rows,
                                                       int n_rows_const, cplx<TO>* vv,
                                                       long long vv_stride, long long ld_vv) {
  __shared__ double red[4];
  const int sl = blockIdx.x;
  const int s = sl / n_levels;
  const int n = n_rows ? n_rows[s] : n_rows_const;
  cplx<TO>* m = vv + (size_t)sl * vv_stride;
  double re = 0, im = 0;
  for (int a = threadIdx.x; a < n; a += 256) {
    const cplx<TO> c = m[a + ld_vv * a];
    re += static_cast<double>(c.re);
    im += static_cast<double>(c.im);
  }
  re = block_sum<256>(re, red);
  im = block_sum<256>(im, red);
  if (threadIdx.x == 0) m[ld_vv] = mk<TO>(static_cast<TO>(re), static_cast<TO>(im));
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

void launch_n3_points(const N3PointsParams& p, int n_samples, hipStream_t st) {
  if (n_samples <= 0 || p.n_list <= 0) return;
  if (!p.slot_list && p.n_list != p.src.slots_per_sample() * n_samples)
    fail(BIPP_INTERNAL_ERROR, "NUFFT point slots do not match the batch");
  n3_points_kernel<<<static_cast<unsigned>((p.n_list + 255) / 256), 256, 0, st>>>(p);
  check_launch("n3_points_kernel");
}

void launch_uvw_antisym(const N3Source& src, int n_samples, unsigned char* out, hipStream_t st) {
  if (n_samples <= 0) return;
  uvw_antisym_kernel<<<n_samples, 256, 0, st>>>(src, out);
  check_launch("uvw_antisym_kernel");
}

void launch_slot_range(const N3Source& src, int n_items, const unsigned long long* list,
                       const unsigned* begin, const unsigned* count, double* out, hipStream_t st) {
  if (n_items <= 0) return;
  slot_range_kernel<<<n_items, 256, 0, st>>>(src, list, begin, count, out);
  check_launch("slot_range_kernel");
}

void launch_n3_partition(const N3PartParams& p, hipStream_t st) {
  const long long n = p.src.slots_per_sample() * p.n_samples;
  if (n <= 0) return;
  n3_partition_kernel<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(p);
  check_launch("n3_partition_kernel");
}

void launch_n3_group_scatter(long long n_slots, const int* group_of, const unsigned* group_start,
                             unsigned* group_fill, unsigned long long* list, hipStream_t st) {
  if (n_slots <= 0) return;
  n3_group_scatter_kernel<<<static_cast<unsigned>((n_slots + 255) / 256), 256, 0, st>>>(
      n_slots, group_of, group_start, group_fill, list);
  check_launch("n3_group_scatter_kernel");
}

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


template <typename R, typename TV>
void launch_n3_spread(const N3SpreadParams<R>& p_all, hipStream_t st) {
  const long long n_blocks = (long long)p_all.nbins[0] * p_all.nbins[1] * p_all.nbins[2];
  if (n_blocks <= 0) return;
  if (n_blocks > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT spread blocks");
  if (p_all.w < 1 || p_all.w > SP_WMAX) fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT kernel width above 16");
  for (int d = 0; d < 3; ++d)
    if (p_all.bin_size[d] != SP_BIN || (long long)p_all.nbins[d] * SP_BIN < p_all.nf1[d] ||
        p_all.nback[d] < 0 || p_all.nback[d] > 2)
      fail(BIPP_INTERNAL_ERROR, "NUFFT spread blocks do not cover the grid");
  // SP_LG levels per launch (the sums of a lane's 8 cells live in registers)
  if (p_all.lv_begin < 0 || p_all.lv_end > p_all.n_levels_total || p_all.lv_begin >= p_all.lv_end)
    fail(BIPP_INTERNAL_ERROR, "NUFFT spread level range");
  for (int lv0 = p_all.lv_begin; lv0 < p_all.lv_end; lv0 += SP_LG) {
    N3SpreadParams<R> p = p_all;
    p.lv0 = lv0;
    p.n_levels = std::min(SP_LG, p_all.lv_end - lv0);
    // with the XCD order (see the kernel), a multiple of 8 chunks of workgroups; the extra
    // ones exit
    const long long unit = p.xcd_chunk > 0 ? 8LL * p.xcd_chunk : 1;
    const long long grid = (n_blocks + unit - 1) / unit * unit;
    n3_spread_kernel<R, TV><<<static_cast<unsigned>(grid), 64 * SpreadWaves<R>::value, 0, st>>>(p);
    check_launch("n3_spread_kernel");
  }
}
template void launch_n3_spread<float, cplx<float>>(const N3SpreadParams<float>&, hipStream_t);
template void launch_n3_spread<double, cplx<double>>(const N3SpreadParams<double>&, hipStream_t);

namespace {
bool window_ok(const N3Window& w, int n) {
  return n > 0 && w.start >= 0 && w.start < n && w.count >= 0 && w.count <= n;
}
inline int max_radix(const N3Radix& r) {
  int m = 0;
  for (int i = 0; i < r.count; ++i) m = std::max(m, r.r[i]);
  return m;
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
  if ((long long)nl * p.n > FFT_NT * FFT_PF) fail(BIPP_INTERNAL_ERROR, "FFT line group too long");
  const size_t sh0 = sizeof(cplx<R>) * (2 * (size_t)p.n * nl + twid_len(p.n));
  // 1 KiB of the 160 KiB stays for the kernel's static line bases
  if (sh0 > 159 * 1024) fail(BIPP_INTERNAL_ERROR, "FFT line exceeds LDS");
  // the stages' twiddle tables where two workgroups still fit a CU
  const size_t sh_tab = sh0 + sizeof(cplx<R>) * stage_tab_len(p.n, p.rad);
  const bool tab = sh_tab <= 79 * 1024;
  const size_t sh = tab ? sh_tab : sh0;
  const long long groups = (n_lines + nl - 1) / nl;
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(groups, 32768));
  auto go = [&](auto kernel, bool& attr_set) {
    if (!attr_set) {
      BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 159 * 1024));
      attr_set = true;
    }
    kernel<<<blocks, FFT_NT, sh, st>>>(p, nl);
  };
  // [max radix > 5][stage tables][one line per group]
  static bool attr[2][2][2] = {};
  const bool r16 = max_radix(p.rad) > 5, one = nl == 1;
  if (r16 && tab && one) go(&fft_lines_kernel<R, 16, true, true>, attr[1][1][1]);
  else if (r16 && tab) go(&fft_lines_kernel<R, 16, true, false>, attr[1][1][0]);
  else if (r16 && one) go(&fft_lines_kernel<R, 16, false, true>, attr[1][0][1]);
  else if (r16) go(&fft_lines_kernel<R, 16, false, false>, attr[1][0][0]);
  else if (tab && one) go(&fft_lines_kernel<R, 5, true, true>, attr[0][1][1]);
  else if (tab) go(&fft_lines_kernel<R, 5, true, false>, attr[0][1][0]);
  else if (one) go(&fft_lines_kernel<R, 5, false, true>, attr[0][0][1]);
  else go(&fft_lines_kernel<R, 5, false, false>, attr[0][0][0]);
  check_launch("fft_lines_kernel");
}
template void launch_fft_lines<float>(const N3LineParams<float>&, hipStream_t);
template void launch_fft_lines<double>(const N3LineParams<double>&, hipStream_t);

// The two-stage register pass for m = M1 M2: x positions per item a power of two in
// [16, 2^tw_shift_max] with the line block at most 80 KiB of LDS (two workgroups per CU).
template <typename R, int M1, int M2>
void go_pass_rr(const N3PassParams<R>& p, hipStream_t st) {
  constexpr int M = M1 * M2, LD = M | 1;
  const int tw_cap = p.tw_shift_max >= 4 && p.tw_shift_max <= 7 ? p.tw_shift_max : 6;
  int tw_shift = 4;
  while (tw_shift < tw_cap && (size_t)LD * (2u << tw_shift) * sizeof(cplx<R>) <= 80 * 1024)
    ++tw_shift;
  const int TW = 1 << tw_shift;
  const long long items =
      (long long)((p.xw.count + TW - 1) / TW) * p.n_sub * p.other.count * p.n_levels;
  if (items <= 0) return;
  if (items > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT grid too large for one pass");
  N3PassParams<R> q = p;
  n3_div_magic((p.xw.count + TW - 1) / TW, q.xt_m, q.xt_s);
  n3_div_magic(p.n_sub, q.sub_m, q.sub_s);
  n3_div_magic(p.other.count, q.oc_m, q.oc_s);
  n3_div_magic(p.operm_n2 > 1 ? p.operm_n1 : 1, q.op_m, q.op_s);
  const size_t sh = sizeof(cplx<R>) * ((size_t)TW * LD + M +
                                       (p.twiddle ? twid_len(static_cast<int>(p.ntot)) : 0));
  if (sh > 96 * 1024) fail(BIPP_INTERNAL_ERROR, "FFT pass exceeds LDS");
  static bool attr = false;
  if (!attr) {
    BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&fft_pass_rr_kernel<R, M1, M2>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
    attr = true;
  }
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(items, 32768));
  fft_pass_rr_kernel<R, M1, M2><<<blocks, FFT_NT, sh, st>>>(q, tw_shift);
}

// factor pairs compiled for the register pass (M1 >= M2, M1 M2 <= 128)
#define BIPP_RR_PAIRS(X)                                                                     \
  X(2, 1) X(2, 2) X(3, 1) X(3, 2) X(3, 3) X(4, 1) X(4, 2) X(4, 3) X(4, 4) X(5, 1) X(5, 2)   \
  X(5, 3) X(5, 4) X(5, 5) X(6, 1) X(6, 2) X(6, 3) X(6, 4) X(6, 5) X(6, 6) X(8, 1) X(8, 2)   \
  X(8, 3) X(8, 4) X(8, 5) X(8, 6) X(8, 8) X(9, 1) X(9, 2) X(9, 3) X(9, 4) X(9, 5) X(9, 6)   \
  X(9, 8) X(9, 9) X(16, 1) X(16, 2) X(16, 3) X(16, 4) X(16, 5) X(16, 6) X(16, 8)

template <typename R>
bool launch_fft_pass_rr(const N3PassParams<R>& p, hipStream_t st) {
  const int m1 = p.rr_m1, m2 = p.m / std::max(1, p.rr_m1);
  if (m1 * m2 != p.m) return false;
#define BIPP_RR_CASE(A, B)              \
  if (m1 == A && m2 == B) {             \
    go_pass_rr<R, A, B>(p, st);         \
    return true;                        \
  }
  BIPP_RR_PAIRS(BIPP_RR_CASE)
#undef BIPP_RR_CASE
  return false;
}

bool fft_rr_factors(int m, int& m1, int& m2) {
#define BIPP_RR_FIND(A, B)   \
  if (A * B == m) {          \
    m1 = A;                  \
    m2 = B;                  \
    found = true;            \
  }
  bool found = false;
  BIPP_RR_PAIRS(BIPP_RR_FIND)  // the last match: the largest first factor
#undef BIPP_RR_FIND
  return found;
}

template <typename R>
void launch_fft_pass(const N3PassParams<R>& p_in, hipStream_t st) {
  N3PassParams<R> p = p_in;
  if (!window_ok(p.xw, p.nx) || !window_ok(p.other, p.no) || !window_ok(p.in, p.nax) ||
      !window_ok(p.out, p.nax) || (long long)p.m * p.n_sub != p.nax ||
      (p.twiddle && p.ntot != p.nax))
    fail(BIPP_INTERNAL_ERROR, "FFT pass windows out of range");
  // x positions per item: a power of two in [16, 64] with about 4096 elements per item
  const int ld = p.m | 1;
  const int tw_cap = p.tw_shift_max >= 4 && p.tw_shift_max <= 6 ? p.tw_shift_max : 6;
  int tw_shift = 4;
  while (tw_shift < tw_cap && (ld << (tw_shift + 1)) <= 4096) ++tw_shift;
  const int TW = 1 << tw_shift;
  if ((long long)p.m * TW > FFT_NT * FFT_PF) fail(BIPP_INTERNAL_ERROR, "FFT pass item too large");
  const long long items =
      (long long)((p.xw.count + TW - 1) / TW) * p.n_sub * p.other.count * p.n_levels;
  if (items <= 0) return;
  if (items > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT grid too large for one pass");
  if (p.operm_n2 > 1 && p.operm_n1 < 1) fail(BIPP_INTERNAL_ERROR, "FFT pass permutation");
  n3_div_magic((p.xw.count + TW - 1) / TW, p.xt_m, p.xt_s);
  n3_div_magic(p.n_sub, p.sub_m, p.sub_s);
  n3_div_magic(p.other.count, p.oc_m, p.oc_s);
  n3_div_magic(p.operm_n2 > 1 ? p.operm_n1 : 1, p.op_m, p.op_s);
  const size_t sh = sizeof(cplx<R>) * (2 * (size_t)TW * ld + twid_len(p.m) +
                                       (p.twiddle ? twid_len(static_cast<int>(p.ntot)) : 0));
  if (sh > 160 * 1024) fail(BIPP_INTERNAL_ERROR, "FFT pass exceeds LDS");
  const unsigned blocks = static_cast<unsigned>(std::min<long long>(items, 32768));
  auto go = [&](auto kernel, bool& attr_set) {
    if (!attr_set) {
      BIPP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      attr_set = true;
    }
    kernel<<<blocks, FFT_NT, sh, st>>>(p, tw_shift);
  };
  if (p.rr_m1 > 0) {
    if (!launch_fft_pass_rr(p, st)) fail(BIPP_INTERNAL_ERROR, "FFT pass factors not compiled");
    check_launch("fft_pass_rr_kernel");
    return;
  }
  static bool attr5 = false, attr16 = false;
  if (max_radix(p.rad) > 5)
    go(&fft_pass_kernel<R, 16>, attr16);
  else
    go(&fft_pass_kernel<R, 5>, attr5);
  check_launch("fft_pass_kernel");
}
template void launch_fft_pass<float>(const N3PassParams<float>&, hipStream_t);
template void launch_fft_pass<double>(const N3PassParams<double>&, hipStream_t);

template <typename R, typename TI>
void launch_n3_interp(const N3InterpParams<R>& p, hipStream_t st) {
  if (p.n_pix <= 0) return;
  const dim3 grid(ceil_div(p.n_pix, 256)), block(256);
  switch (p.w2) {
#define BIPP_INTERP_W(W) \
  case W: n3_interp_kernel<R, TI, W><<<grid, block, 0, st>>>(p); break;
    BIPP_INTERP_W(2) BIPP_INTERP_W(3) BIPP_INTERP_W(4) BIPP_INTERP_W(5) BIPP_INTERP_W(6)
    BIPP_INTERP_W(7) BIPP_INTERP_W(8) BIPP_INTERP_W(9) BIPP_INTERP_W(10) BIPP_INTERP_W(11)
    BIPP_INTERP_W(12) BIPP_INTERP_W(13) BIPP_INTERP_W(14) BIPP_INTERP_W(15) BIPP_INTERP_W(16)
#undef BIPP_INTERP_W
    default: fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT kernel width outside [2, 16]");
  }
  check_launch("n3_interp_kernel");
}
template void launch_n3_interp<float, float>(const N3InterpParams<float>&, hipStream_t);
template void launch_n3_interp<double, double>(const N3InterpParams<double>&, hipStream_t);

template <typename TO>
void launch_vv_trace(int n_samples, int n_levels, const int* n_rows, int n_rows_const, void* vv,
                     long long vv_stride, long long ld_vv, hipStream_t st) {
  if (n_samples <= 0 || n_levels <= 0) return;
  if (ld_vv < 2) fail(BIPP_INTERNAL_ERROR, "trace slot needs at least two rows");
  vv_trace_kernel<TO><<<n_samples * n_levels, 256, 0, st>>>(
      n_levels, n_rows, n_rows_const, static_cast<cplx<TO>*>(vv), vv_stride, ld_vv);
  check_launch("vv_trace_kernel");
}
template void launch_vv_trace<float>(int, int, const int*, int, void*, long long, long long,
                                     hipStream_t);
template void launch_vv_trace<double>(int, int, const int*, int, void*, long long, long long,
                                      hipStream_t);

template <typename R, typename TO>
void launch_vv_levels(const N3VVParams<R>& p, int n_samples, int n_rows_max, hipStream_t st) {
  if (n_samples <= 0 || n_rows_max <= 0) return;
  if constexpr (std::is_same<R, double>::value) {
    const long long nt = ceil_div(n_rows_max, 64), n_tiles = nt * (nt + 1) / 2;
    const long long n_units = n_tiles * n_samples * p.n_levels;
    const long long per_xcd = (n_units + 7) / 8;
    if (8 * per_xcd > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "too many VV tiles for one launch");
    vv_levels_mfma_kernel<TO><<<static_cast<unsigned>(8 * per_xcd), 256, 0, st>>>(
        p, static_cast<int>(n_tiles), static_cast<int>(per_xcd), static_cast<int>(n_units));
    check_launch("vv_levels_mfma_kernel");
  } else {
    dim3 grid(ceil_div(n_rows_max, 64), ceil_div(n_rows_max, 64), n_samples * p.n_levels);
    vv_levels_kernel<R, TO><<<grid, 256, 0, st>>>(p);
    check_launch("vv_levels_kernel");
  }
}
template void launch_vv_levels<double, float>(const N3VVParams<double>&, int, int, hipStream_t);
template void launch_vv_levels<double, double>(const N3VVParams<double>&, int, int, hipStream_t);
template void launch_vv_levels<float, float>(const N3VVParams<float>&, int, int, hipStream_t);
template void launch_vv_levels<float, double>(const N3VVParams<float>&, int, int, hipStream_t);

}  // namespace bippamd
