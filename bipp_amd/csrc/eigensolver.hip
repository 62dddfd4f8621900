// Batched generalized Hermitian eigensolver for the Bluebild pencil (S, G) on gfx950.
//
// Restates, per sample, src/eigensolver.cpp:129-225 (flag -> reduce -> Gram -> ?hegv)
// with the LAPACK steps of ?hegv(itype=1,'V','L') (src/host/lapack_api.hpp:112-158):
//   Cholesky G = L L^H and W = L^-1 (blocked, chol_inv.hpp), C = W S W^H (two LDS-tiled GEMMs),
//   Householder tridiagonalisation C = Q T Q^H on the lower triangle, eigenpairs of T by
//   divide and conquer (tridiag_dc.hpp, eigenvalues ascending, eigenvectors into Z), X = Q Z
//   (eigenvector columns in registers), eigenvectors W^H X (GEMM).  ?hegv's QL iteration on T
//   gives the same eigenpairs to rounding (eigenvectors of repeated eigenvalues up to a
//   rotation within their eigenspace).
//
// Launch shape: a persistent grid (512-thread workgroups for problems above 128, 256
// below); each workgroup pulls sample indices from a device counter (largest problems
// first) and solves one problem at a time, keeping its matrices in a per-workgroup slot of
// global memory and its vectors, tiles and per-merge arrays in LDS.  The column steps are
// workgroup-wide parallel updates separated by barriers, with lanes mapped to matrix rows
// so that global accesses are coalesced.
#include "common.hpp"
#include "kernels.hpp"

#define BIPP_HD __device__ __forceinline__
#define BIPP_GLOBAL __attribute__((address_space(1)))
#include "chol_inv.hpp"
#include "tridiag_dc.hpp"

#include <limits>
#include <type_traits>

namespace bippamd {

namespace {

// Threads per workgroup for the small kernels of this file.
constexpr int NT = 256;
constexpr int NW = NT / 64;

// Eigensolver launch shape per size bucket: wider workgroups for the larger problems, so
// that more columns are in flight per step; the per-sample matrices live in global memory,
// so most steps are bound by dependent round trips.  (Two 256-thread workgroups per CU for
// NMAX = 256, which overlaps one sample's QL chain with the other's work, measured slower.)
template <int NMAX>
struct EigCfg {
  static constexpr int NT = NMAX >= 256 ? 512 : 256;
  static constexpr int NW = NT / 64;
  static constexpr int RM = NMAX / 64;          // rows per lane for a full column
  static constexpr int UC = RM >= 8 ? 2 : 4;    // columns per wave loaded together
};

template <typename R>
__device__ __forceinline__ R sinpi_r(R x);
template <>
__device__ __forceinline__ double sinpi_r<double>(double x) { return sinpi(x); }
template <>
__device__ __forceinline__ float sinpi_r<float>(float x) { return sinpif(x); }

template <typename R>
__device__ __forceinline__ R eps_r();
template <>
__device__ __forceinline__ double eps_r<double>() { return 2.220446049250313e-16; }
template <>
__device__ __forceinline__ float eps_r<float>() { return 1.1920928955078125e-07f; }

// complex reciprocal 1/(a)
template <typename R>
__device__ __forceinline__ cplx<R> crecip(cplx<R> a) {
  // Smith's algorithm
  if (fabs(a.re) >= fabs(a.im)) {
    R r = a.im / a.re, den = a.re + a.im * r;
    return mk<R>(R(1) / den, -r / den);
  } else {
    R r = a.re / a.im, den = a.im + a.re * r;
    return mk<R>(r / den, R(-1) / den);
  }
}

// In-workgroup GEMM tiling: GT x GT output tiles, k taken GK at a time through LDS.
constexpr int GT = 64;
constexpr int GK = 16;

template <typename R>
struct GemmTiles {
  cplx<R> a[GK][GT + 1];  // a[k][r] = op(A)[i0 + r, k0 + k]
  cplx<R> b[GK][GT + 1];  // b[k][c] = op(B)[k0 + k, j0 + c]
};

// Back-transform layout: LPC lanes hold one eigenvector column (RR rows each), reflectors
// staged through LDS RB at a time.
template <int NMAX>
struct BtCfg {
  static constexpr int LPC = NMAX >= 512 ? 32 : 16;
  static constexpr int RR = NMAX / LPC;
  static constexpr int CPWV = 64 / LPC;
  static constexpr int RB = NMAX >= 512 ? 4 : 8;
  static constexpr int CPG = NMAX >= 512 ? 1 : 2;  // columns per lane group
};

// The float64 back-transform on matrix instructions (n <= NMAX <= 256): one block of 16
// reflectors over all rows, the wavefronts' partial products, the block's Gram matrix, T
// factor and tau.
constexpr bool bt_mfma(int nmax) { return nmax <= 256; }
template <typename R, int NMAX>
struct BtBlock {
  cplx<R> v[16][NMAX + 1];            // v[q][i]: reflector q of the block at row i
  R red[EigCfg<NMAX>::NW][64][8];     // per wavefront: a 16 x 16 complex result, lane layout
  cplx<R> g[16][16];                  // V^H V
  cplx<R> t[16][16];                  // T (upper triangular)
  cplx<R> tau[16];
};

template <typename R, int NMAX>
struct BackTile {
  cplx<R> v[BtCfg<NMAX>::RB][NMAX];  // reflector vectors (zero above the unit entry)
  cplx<R> tau[BtCfg<NMAX>::RB];
};

// Two reals accessed together (one 2 x sizeof(R) LDS access).
template <typename R>
struct alignas(2 * sizeof(R)) RPair {
  R x, y;
};

template <typename R, int NMAX>
struct EigShared {
  static constexpr int NW_ = EigCfg<NMAX>::NW;
  union {
    struct {
      cplx<R> v[NMAX];
      cplx<R> w[NMAX];
      cplx<R> vn[NMAX];
      cplx<R> xc[NMAX];  // column (conjugate-transpose) parts of the Hermitian products
    };
    R dct[DcCfg<NMAX>::TILE_ELEMS];  // product tiles of the tridiagonal divide and conquer
  };
  union {
    cplx<R> xw[NW_][NMAX];  // per-wave partial sums of the matrix-vector passes
    GemmTiles<R> gt;        // operand tiles of the GEMM passes
    BackTile<R, NMAX> bt;   // reflectors of the back-transform
    typename std::conditional<std::is_same<R, double>::value && bt_mfma(NMAX), BtBlock<R, NMAX>,
                              char>::type bb;  // blocks of the back-transform on MFMA
    DcShared<R, NMAX> dc;   // per-merge arrays of the tridiagonal divide and conquer
    CiShared<R> ci;         // diagonal blocks and product tiles of the Cholesky inverse
  };
  RPair<R> de[NMAX];      // tridiagonal: (diagonal d_i, subdiagonal e_i)
  int perm[NMAX];
  int idx[NMAX];
  R red[2][2 * NW_];  // two slots: back-to-back reductions need no barrier between
  R scal[4];
  int iscal[4];
};
static_assert(DcCfg<256>::TILE_ELEMS <= 8 * 256 && DcCfg<64>::TILE_ELEMS <= 8 * 64,
              "divide-and-conquer tiles fit in the four vectors");

// The eigensolver's LDS, at namespace scope: the phases kept out of line (own register
// allocation) name it directly, so that their accesses compile to LDS instructions rather
// than generic (flat) ones through a reference parameter.
template <typename R, int NMAX>
__shared__ EigShared<R, NMAX> eig_sh;

// --------------------------------------------------------------------------------------
// C = op(A) op(B) for n x n matrices, GT x GT output tiles (with LOWER_ONLY the tiles on
// and below the diagonal).  fa(i, k) and fb(k, j) return operand elements, zero outside
// their nonzero pattern; krange(i0, j0, klo, khi) bounds the k that can contribute to a
// tile.  A_KFAST / B_KFAST: consecutive k are contiguous in memory (else consecutive i / j),
// which sets how the threads split a tile load.  The loads of the next k chunk are in flight
// while the current one is used.  Every thread executes the same sequence of barriers.
// float64: the tile is 4 x 4 blocks of 16 x 16, 16 / NW blocks per wavefront, each complex
// product as four v_mfma_f64_16x16x4f64 (re re, -im im into the real part, re im, im re into
// the imaginary part); operand layouts: A[row l % 16][k l / 16], B[k l / 16][col l % 16],
// result register r at row l / 16 + 4 r, column l % 16 (l: lane).  float32: lanes own tile
// rows, wavefronts GT / NW tile columns, on the vector ALU.
// --------------------------------------------------------------------------------------
template <typename R, int NMAX, bool A_KFAST, bool B_KFAST, bool LOWER_ONLY, typename FA,
          typename FB, typename KR>
__device__ __forceinline__ void gemm_tiles(int n, FA fa, FB fb, KR krange,
                                           cplx<R>* __restrict__ C, int ldc,
                                           EigShared<R, NMAX>& sh) {
  constexpr int NT_ = EigCfg<NMAX>::NT, NW_ = EigCfg<NMAX>::NW;
  constexpr int CPW = GT / NW_;
  constexpr int PER = GT * GK / NT_;
  static_assert(GT * GK % NT_ == 0, "tile loads split evenly over the workgroup");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  GemmTiles<R>& t = sh.gt;
  const int nt = (n + GT - 1) / GT;
  constexpr bool MF = std::is_same<R, double>::value;
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  constexpr int BPW = 16 / NW_;             // 16 x 16 blocks per wavefront (float64)
  const int br = wave & 3, bc0 = (wave >> 2) * BPW;
  for (int tj = 0; tj < nt; ++tj) {
    for (int ti = LOWER_ONLY ? tj : 0; ti < nt; ++ti) {
      const int i0 = ti * GT, j0 = tj * GT;
      int klo = 0, khi = 0;
      krange(i0, j0, klo, khi);
      cplx<R> acc[CPW];
      f64x4 cr[BPW], ci[BPW];
#pragma unroll
      for (int c = 0; c < CPW; ++c) acc[c] = mk<R>(0, 0);
#pragma unroll
      for (int q = 0; q < BPW; ++q) {
        cr[q] = f64x4{0, 0, 0, 0};
        ci[q] = f64x4{0, 0, 0, 0};
      }
      cplx<R> ra[PER], rb[PER];
      auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT_;
          const int ka = A_KFAST ? e % GK : e / GT;
          const int ia = i0 + (A_KFAST ? e / GK : e % GT);
          ra[q] = (ia < n && k0 + ka < khi) ? fa(ia, k0 + ka) : mk<R>(0, 0);
          const int kb = B_KFAST ? e % GK : e / GT;
          const int jb = j0 + (B_KFAST ? e / GK : e % GT);
          rb[q] = (jb < n && k0 + kb < khi) ? fb(k0 + kb, jb) : mk<R>(0, 0);
        }
      };
      if (klo < khi) fetch(klo);
      for (int k0 = klo; k0 < khi; k0 += GK) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT_;
          t.a[A_KFAST ? e % GK : e / GT][A_KFAST ? e / GK : e % GT] = ra[q];
          t.b[B_KFAST ? e % GK : e / GT][B_KFAST ? e / GK : e % GT] = rb[q];
        }
        __syncthreads();
        if (k0 + GK < khi) fetch(k0 + GK);
        if constexpr (MF) {
#pragma unroll
          for (int ks = 0; ks < GK / 4; ++ks) {
            const int kr = 4 * ks + (lane >> 4);
            const cplx<R> av = t.a[kr][br * 16 + (lane & 15)];
#pragma unroll
            for (int q = 0; q < BPW; ++q) {
              const cplx<R> bv = t.b[kr][(bc0 + q) * 16 + (lane & 15)];
              cr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.re, bv.re, cr[q], 0, 0, 0);
              cr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.im, bv.im, cr[q], 0, 0, 0);
              ci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.re, bv.im, ci[q], 0, 0, 0);
              ci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.im, bv.re, ci[q], 0, 0, 0);
            }
          }
        } else {
#pragma unroll 4
          for (int k = 0; k < GK; ++k) {
            const cplx<R> a = t.a[k][lane];
#pragma unroll
            for (int c = 0; c < CPW; ++c) acc[c] = acc[c] + a * t.b[k][wave * CPW + c];
          }
        }
        __syncthreads();
      }
      if constexpr (MF) {
#pragma unroll
        for (int q = 0; q < BPW; ++q) {
          const int j = j0 + (bc0 + q) * 16 + (lane & 15);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = i0 + br * 16 + (lane >> 4) + 4 * r;
            if (i < n && j < n) C[i + (size_t)j * ldc] = mk<R>(cr[q][r], ci[q][r]);
          }
        }
      } else {
        const int i = i0 + lane;
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
          const int j = j0 + wave * CPW + c;
          if (i < n && j < n) C[i + (size_t)j * ldc] = acc[c];
        }
      }
    }
  }
}

// Sum over the 64 lanes of a wavefront (every lane gets it; DPP within rows of 16, then
// two lane swaps).
template <typename R>
__device__ __forceinline__ R wave_total(R v) {
  v = sum16(v);
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// Sums of a and b over the workgroup (every thread gets them), with one barrier: the
// caller guarantees that another barrier separates this call from the next write of
// sh.red[slot] (it holds 2 values per wavefront).
template <typename R, int NMAX>
__device__ __forceinline__ void block_sum2(R& a, R& b, EigShared<R, NMAX>& sh, int slot = 0) {
  constexpr int NW_ = EigCfg<NMAX>::NW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  R* red = sh.red[slot];
  a = wave_total(a);
  b = wave_total(b);
  if (lane == 0) {
    red[wave] = a;
    red[NW_ + wave] = b;
  }
  __syncthreads();
  a = 0;
  b = 0;
#pragma unroll
  for (int w = 0; w < NW_; ++w) {
    a += red[w];
    b += red[NW_ + w];
  }
}

// Column-major matrix of cplx<R> (leading dimension ld) accessed by raw buffer
// instructions.  They count only against the vector-memory counter (through the generic
// pointer of an out-of-line function they would be flat accesses, which every LDS wait also
// waits for), and an offset past the buffer's end (OOB) reads zero and drops the store, so a
// masked lane needs no branch around its load or store.
template <typename R>
struct BufMat {
  static constexpr unsigned OOB = 0x80000000u;
  __amdgpu_buffer_rsrc_t rs;
  int ld;
  __device__ BufMat(cplx<R>* A, int ld_, int n)
      : rs(__builtin_amdgcn_make_buffer_rsrc(
            A, 0, (int)(sizeof(cplx<R>) * ((size_t)ld_ * (n - 1) + n)), 0x00020000)),
        ld(ld_) {}
  __device__ unsigned off(int i, int j) const {
    return (unsigned)(sizeof(cplx<R>) * (i + (size_t)j * ld));
  }
  __device__ cplx<R> load(unsigned o) const;
  __device__ void store(unsigned o, cplx<R> x) const;
};
template <>
__device__ __forceinline__ cplx<double> BufMat<double>::load(unsigned o) const {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 v = __builtin_amdgcn_raw_buffer_load_b128(rs, o, 0, 0);
  return mk<double>(__builtin_bit_cast(double, ((unsigned long long)v.y << 32) | v.x),
                    __builtin_bit_cast(double, ((unsigned long long)v.w << 32) | v.z));
}
template <>
__device__ __forceinline__ void BufMat<double>::store(unsigned o, cplx<double> x) const {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const unsigned long long a = __builtin_bit_cast(unsigned long long, x.re);
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x.im);
  const u4 v = {(unsigned)a, (unsigned)(a >> 32), (unsigned)b, (unsigned)(b >> 32)};
  __builtin_amdgcn_raw_buffer_store_b128(v, rs, o, 0, 0);
}
// two 32-bit loads: this compiler lowers the 64-bit buffer load builtin to a single dword
template <>
__device__ __forceinline__ cplx<float> BufMat<float>::load(unsigned o) const {
  return mk<float>(__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o, 0, 0)),
                   __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o + 4, 0, 0)));
}
template <>
__device__ __forceinline__ void BufMat<float>::store(unsigned o, cplx<float> x) const {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  const u2 v = {__builtin_bit_cast(unsigned, x.re), __builtin_bit_cast(unsigned, x.im)};
  __builtin_amdgcn_raw_buffer_store_b64(v, rs, o, 0, 0);
}

// Householder reflector (zlarfg semantics) annihilating A[c+2:n, c]; writes v (v[c+1]=1)
// into vec and into A[c+2:n, c], beta into A[c+1, c] and ee[c]; returns tau.  Thread t holds
// row c + 2 + t.
template <typename R, int NMAX>
__device__ __forceinline__ cplx<R> reflector(int n, int c, const BufMat<R>& M, cplx<R>* vec,
                                             EigShared<R, NMAX>& sh) {
  static_assert(EigCfg<NMAX>::NT >= NMAX, "one row per thread");
  const int tid = threadIdx.x;
  const cplx<R> alpha = M.load(M.off(c + 1, c));
  const int i = c + 2 + tid;
  const unsigned oi = i < n ? M.off(i, c) : BufMat<R>::OOB;
  const cplx<R> xi = M.load(oi);
  R part = abs2(xi), unused = 0;
  block_sum2(part, unused, sh);  // sh.red is next written after this function's barrier
  const R xnorm2 = part;
  cplx<R> tau, scal;
  R beta;
  if (xnorm2 == R(0) && alpha.im == R(0)) {
    tau = mk<R>(0, 0);
    beta = alpha.re;
    scal = mk<R>(0, 0);
  } else {
    const R nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
    beta = alpha.re >= R(0) ? -nrm : nrm;
    tau = mk<R>((beta - alpha.re) / beta, -alpha.im / beta);
    scal = crecip(mk<R>(alpha.re - beta, alpha.im));
  }
  const cplx<R> vi = xi * scal;
  M.store(oi, vi);
  if (i < n) vec[i] = vi;
  if (tid == 0) {
    vec[c + 1] = mk<R>(1, 0);
    M.store(M.off(c + 1, c), mk<R>(beta, 0));
    sh.de[c].y = beta;
  }
  __syncthreads();
  return tau;
}

// One pass over the lower triangle of the trailing square [i0,n)^2 of the Hermitian A
// (the strict upper triangle is neither read nor written):
//   optionally A -= v w^H + w v^H (pending rank-2 update), then x = A * vn, where each
//   stored A[i, j] (i >= j) adds A[i, j] vn_j to x_i and, below the diagonal,
//   conj(A[i, j]) vn_i to x_j (a per-column reduction over the wavefront's lanes).
// Result x[i0:n) is written to sh.w (v and w are no longer needed afterwards).
template <typename R, int NMAX, bool APPLY>
__device__ __forceinline__ void trail_pass(int n, int i0, BIPP_GLOBAL cplx<R>* __restrict__ A, int ld,
                                           const cplx<R>* v, const cplx<R>* vn,
                                           EigShared<R, NMAX>& sh) {
  constexpr int NT_ = EigCfg<NMAX>::NT, NW_ = EigCfg<NMAX>::NW, RM = EigCfg<NMAX>::RM,
                UC = EigCfg<NMAX>::UC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cplx<R> acc[RM], vi[RM], wi[RM], vni[RM];
#pragma unroll
  for (int r = 0; r < RM; ++r) {
    acc[r] = mk<R>(0, 0);
    const int i = i0 + lane + 64 * r;
    vi[r] = (APPLY && i < n) ? v[i] : mk<R>(0, 0);
    wi[r] = (APPLY && i < n) ? sh.w[i] : mk<R>(0, 0);
    vni[r] = i < n ? vn[i] : mk<R>(0, 0);
  }
  for (int jb = i0 + wave; jb < n; jb += NW_ * UC) {
    cplx<R> a[UC][RM];
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const int j = jb + u * NW_;
      const BIPP_GLOBAL cplx<R>* col = A + (size_t)j * ld;
#pragma unroll
      for (int r = 0; r < RM; ++r) {
        const int i = i0 + lane + 64 * r;
        if (j < n && i < n && i >= j) a[u][r] = ci_get(col[i]);
      }
    }
    R pre[UC], pim[UC];
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      const int j = jb + u * NW_;
      pre[u] = 0;
      pim[u] = 0;
      if (j >= n) continue;
      cplx<R> cvj = mk<R>(0, 0), cwj = mk<R>(0, 0);
      if (APPLY) {
        cvj = conj(v[j]);
        cwj = conj(sh.w[j]);
      }
      const cplx<R> vnj = vn[j];
      BIPP_GLOBAL cplx<R>* col = A + (size_t)j * ld;
#pragma unroll
      for (int r = 0; r < RM; ++r) {
        const int i = i0 + lane + 64 * r;
        if (i < n && i >= j) {
          cplx<R> x = a[u][r];
          if (APPLY) {
            x = x - (vi[r] * cwj + wi[r] * cvj);
            ci_put(col[i], x);
          }
          acc[r] = acc[r] + x * vnj;
          if (i > j) {
            const cplx<R> t = cmul_conj_a(x, vni[r]);
            pre[u] += t.re;
            pim[u] += t.im;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < UC; ++u) {
      pre[u] = wave_total(pre[u]);
      pim[u] = wave_total(pim[u]);
    }
    if (lane == 0) {
#pragma unroll
      for (int u = 0; u < UC; ++u) {
        const int j = jb + u * NW_;
        if (j < n) sh.xc[j] = mk<R>(pre[u], pim[u]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RM; ++r) {
    const int i = i0 + lane + 64 * r;
    if (i < n) sh.xw[wave][i] = acc[r];
  }
  __syncthreads();
  for (int i = i0 + tid; i < n; i += NT_) {
    cplx<R> s = sh.xc[i];
#pragma unroll
    for (int w = 0; w < NW_; ++w) s = s + sh.xw[w][i];
    sh.w[i] = s;
  }
  __syncthreads();
}

// w = tau*x - 1/2 tau (tau x)^H v v, for the reflector stored in vec over [c+1, n).
// On entry sh.w holds x (un-scaled); on exit sh.w holds w.
template <typename R, int NMAX>
__device__ __forceinline__ void make_w(int n, int c, cplx<R> tau, const cplx<R>* vec,
                                       EigShared<R, NMAX>& sh) {
  constexpr int NT_ = EigCfg<NMAX>::NT;
  const int tid = threadIdx.x;
  R pr = 0, pi = 0;
  for (int i = c + 1 + tid; i < n; i += NT_) {
    cplx<R> x = tau * sh.w[i];
    sh.w[i] = x;
    cplx<R> t = cmul_conj_a(x, vec[i]);  // conj(x) v
    pr += t.re;
    pi += t.im;
  }
  block_sum2(pr, pi, sh);  // sh.red is next written after this function's barrier
  const cplx<R> alpha2 = tau * mk<R>(R(-0.5) * pr, R(-0.5) * pi);
  for (int i = c + 1 + tid; i < n; i += NT_) sh.w[i] = sh.w[i] + alpha2 * vec[i];
  __syncthreads();
}

// Arguments of an out-of-line function arrive in vector registers.  The phases' sizes and
// pointers are the same in every lane: read back from the first lane they are known to be
// uniform, so the loops over them branch on scalars and buffer resources built from them stay
// in scalar registers.
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename T>
__device__ __forceinline__ T* uniform(T* p) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(a)));
  const unsigned hi =
      static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(a >> 32)));
  return reinterpret_cast<T*>((static_cast<unsigned long long>(hi) << 32) | lo);
}
// A pointer into global memory as such (a generic pointer gives flat instructions, which
// count against the LDS counter as well, so every LDS wait would also wait for them).
template <typename T>
__device__ __forceinline__ BIPP_GLOBAL T* global_ptr(T* p) {
  return (BIPP_GLOBAL T*)uniform(p);
}

// ci_gemm's tile arithmetic in float64 on v_mfma_f64_16x16x4f64: the 64 x 64 tile as 4 x 4
// blocks of 16 x 16, 16 / NW blocks per wavefront, each complex product as four MFMAs (the
// operand and result layouts of gemm_tiles).
template <int NT>
struct CiMfmaTile {
  static constexpr int NW = NT / 64, BPW = 16 / NW;
  static_assert(CI_T == 64 && CI_K % 4 == 0 && BPW * NW == 16, "4 x 4 blocks of 16");
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  f64x4 cr[BPW], ci[BPW];
  __device__ void init() {
#pragma unroll
    for (int q = 0; q < BPW; ++q) {
      cr[q] = f64x4{0, 0, 0, 0};
      ci[q] = f64x4{0, 0, 0, 0};
    }
  }
  __device__ void step(const CiShared<double>& sh, int tid) {
    const int lane = tid & 63, wave = tid >> 6, br = wave & 3, bc0 = (wave >> 2) * BPW;
#pragma unroll
    for (int ks = 0; ks < CI_K / 4; ++ks) {
      const int kr = 4 * ks + (lane >> 4);
      const cplx<double> av = sh.gt.a[kr][br * 16 + (lane & 15)];
#pragma unroll
      for (int q = 0; q < BPW; ++q) {
        const cplx<double> bv = sh.gt.b[kr][(bc0 + q) * 16 + (lane & 15)];
        cr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.re, bv.re, cr[q], 0, 0, 0);
        cr[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-av.im, bv.im, cr[q], 0, 0, 0);
        ci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.re, bv.im, ci[q], 0, 0, 0);
        ci[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av.im, bv.re, ci[q], 0, 0, 0);
      }
    }
  }
  template <class ST>
  __device__ void store(int i0, int j0, int m, int nn, ST st, int tid) const {
    const int lane = tid & 63, wave = tid >> 6, br = wave & 3, bc0 = (wave >> 2) * BPW;
#pragma unroll
    for (int q = 0; q < BPW; ++q) {
      const int j = j0 + (bc0 + q) * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + br * 16 + (lane >> 4) + 4 * r;
        if (i < m && j < nn) st(i, j, mk<double>(cr[q][r], ci[q][r]));
      }
    }
  }
};

// dc_product_tiles' tile arithmetic in float64 on v_mfma_f64_16x16x4f64 (64 x 64 tiles,
// 512 threads): 4 x 4 blocks of 16 x 16, two per wavefront (the layouts of gemm_tiles).
template <int TK>
struct DcMfmaTile {
  static constexpr int TR = 64;
  static_assert(TK % 4 == 0, "k chunks of whole MFMA steps");
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  f64x4 acc[2];
  __device__ void init() {
    acc[0] = f64x4{0, 0, 0, 0};
    acc[1] = f64x4{0, 0, 0, 0};
  }
  __device__ void step(const double* tA, const double* tB, int tid) {
    const int lane = tid & 63, wave = tid >> 6, br = wave & 3, bc0 = (wave >> 2) * 2;
    const int lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int s4 = 0; s4 < TK / 4; ++s4) {
      const double a = tA[(4 * s4 + lk) * TR + br * 16 + lr];
#pragma unroll
      for (int q = 0; q < 2; ++q)
        acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, tB[(4 * s4 + lk) * TR + (bc0 + q) * 16 + lr],
                                                      acc[q], 0, 0, 0);
    }
  }
  template <class ST>
  __device__ void store(ST st, int tid) const {
    const int lane = tid & 63, wave = tid >> 6, br = wave & 3, bc0 = (wave >> 2) * 2;
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) st(br * 16 + (lane >> 4) + 4 * r, (bc0 + q) * 16 + (lane & 15), acc[q][r]);
  }
};

// The workgroup as the team of the team-parallel routines (chol_inv.hpp, tridiag_dc.hpp).
struct WgTeam {
  int tid;
  unsigned long long* prof = nullptr;  // step k's wall-clock time added to prof[k] (thread 0)
  unsigned long long t = 0;
  template <typename R, int NT>
  using CiTile =
      typename std::conditional<std::is_same<R, double>::value, CiMfmaTile<NT>, CiValuTile<R, NT>>::type;
  template <typename R, int TR, int TK, int NT>
  using DcTile = typename std::conditional<std::is_same<R, double>::value && NT == 512 && TR == 64,
                                           DcMfmaTile<TK>, DcValuTile<R, TR, TK, NT>>::type;
  __device__ void sync() const { __syncthreads(); }
  // total over the aligned group of P lanes (a + b == b + a: the same bits in each lane)
  template <int P, typename R>
  __device__ R lane_sum(R v) const {
#pragma unroll
    for (int m = 1; m < P; m <<= 1) v += __shfl_xor(v, m);
    return v;
  }
  __device__ void mark(int k) {
    if (prof && tid == 0) {
      const unsigned long long now = wall_clock64();
      if (k >= 0) ((BIPP_GLOBAL unsigned long long*)prof)[k] += now - t;
      t = now;
    }
  }
};

// After the pass for column k (sh.w holds y = A v_k): w_k = tau_k y - 1/2 tau_k (tau_k y)^H
// v_k v_k into sh.w, the pending update (v_k, w_k) applied to column k + 1 and the reflector
// of column k + 1 formed (v_{k+1} into vnext and below the subdiagonal of A, beta into
// A[k + 2, k + 1] and sh.de[k + 1].y, the updated diagonal into sh.de[k + 1].x; returns
// tau_{k+1}).  Thread t holds row k + 1 + t throughout, so the updated column stays in
// registers (the reflector overwrites it in A); w_k at row k + 1 is formed by every thread
// from its own copy of y_{k+1}; alpha of the reflector goes through LDS before the norm's
// barrier.  Three barriers instead of the five of the separate steps.
template <typename R, int NMAX>
__device__ __forceinline__ cplx<R> post_column(int n, int k, cplx<R> tau, const BufMat<R>& M,
                                               const cplx<R>* vcur, cplx<R>* vnext,
                                               EigShared<R, NMAX>& sh) {
  static_assert(EigCfg<NMAX>::NT >= NMAX, "one row per thread");
  const int tid = threadIdx.x;
  const int i = k + 1 + tid;
  const bool live = i < n;
  const unsigned oi = live ? M.off(i, k + 1) : BufMat<R>::OOB;
  const cplx<R> a = M.load(oi);  // column k + 1, row i (in flight during the dot)
  const cplx<R> yi = live ? sh.w[i] : mk<R>(0, 0);
  const cplx<R> vi = live ? vcur[i] : mk<R>(0, 0);
  const cplx<R> y1 = sh.w[k + 1], v1 = vcur[k + 1];  // row k + 1, for every thread
  const cplx<R> xi = tau * yi;
  const cplx<R> dt = cmul_conj_a(xi, vi);
  R pr = dt.re, pi = dt.im;
  block_sum2(pr, pi, sh, 0);
  const cplx<R> alpha2 = tau * mk<R>(R(-0.5) * pr, R(-0.5) * pi);
  const cplx<R> wi = xi + alpha2 * vi;
  const cplx<R> w1 = tau * y1 + alpha2 * v1;
  if (live) sh.w[i] = wi;
  // column k + 1 with the update (v_k, w_k)
  const cplx<R> c = a - (vi * conj(w1) + wi * conj(v1));
  if (tid == 0) sh.de[k + 1].x = c.re;
  // reflector of column k + 1: alpha = row k + 2 (thread 1), the rest below it
  if (tid == 1) sh.scal[0] = c.re, sh.scal[1] = c.im;
  R part = (tid >= 2 && live) ? abs2(c) : R(0), unused = 0;
  block_sum2(part, unused, sh, 1);
  const cplx<R> alpha = n > k + 2 ? mk<R>(sh.scal[0], sh.scal[1]) : mk<R>(0, 0);
  const R xnorm2 = part;
  cplx<R> tn, scal;
  R beta;
  if (xnorm2 == R(0) && alpha.im == R(0)) {
    tn = mk<R>(0, 0);
    beta = alpha.re;
    scal = mk<R>(0, 0);
  } else {
    const R nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xnorm2);
    beta = alpha.re >= R(0) ? -nrm : nrm;
    tn = mk<R>((beta - alpha.re) / beta, -alpha.im / beta);
    scal = crecip(mk<R>(alpha.re - beta, alpha.im));
  }
  if (tid >= 2 && live) {
    const cplx<R> v = c * scal;
    M.store(oi, v);
    vnext[i] = v;
  }
  if (tid == 1 && live) {
    vnext[i] = mk<R>(1, 0);
    M.store(oi, mk<R>(beta, 0));
    sh.de[k + 1].y = beta;
  }
  if (tid == 0) M.store(M.off(k + 1, k + 1), tn);  // the diagonal slot keeps tau
  __syncthreads();
  return tn;
}

// Householder tridiagonalisation (zhetd2 order, lower triangle, n >= 2): d and e into
// sh.de, the reflectors below the subdiagonal of A with tau in the diagonal slots (0 for the
// last column).  Each column's pass applies the pending rank-2 update and forms the
// Hermitian product from the stored lower triangle.  Out of line: its registers are
// allocated apart from the other phases.
template <typename R, int NMAX>
__device__ __noinline__ void tridiag_phase(int n, cplx<R>* __restrict__ A, int ld,
                                           unsigned long long* prof) {
  n = uniform(n);
  ld = uniform(ld);
  A = uniform(A);
  prof = uniform(prof);
  EigShared<R, NMAX>& sh = eig_sh<R, NMAX>;
  const int tid = threadIdx.x;
  cplx<R>* vcur = sh.v;
  cplx<R>* vnext = sh.vn;
  const BufMat<R> M(A, ld, n);
  WgTeam tm{tid, prof};  // step timing: 0 w, column update and reflector, 1 pass
  tm.mark(-1);
  const R d0 = M.load(M.off(0, 0)).re;
  cplx<R> tau = reflector<R, NMAX>(n, 0, M, vcur, sh);
  if (tid == 0) {
    sh.de[0].x = d0;
    M.store(M.off(0, 0), tau);  // diagonal slot keeps tau_0 for the back-transform
  }
  trail_pass<R, NMAX, false>(n, 1, (BIPP_GLOBAL cplx<R>*)A, ld, vcur, vcur, sh);
  for (int k = 0; k + 2 < n; ++k) {
    // w_k, column k + 1 updated, reflector k + 1
    tm.mark(1);
    const cplx<R> tn = post_column<R, NMAX>(n, k, tau, M, vcur, vnext, sh);
    tm.mark(0);
    trail_pass<R, NMAX, true>(n, k + 2, (BIPP_GLOBAL cplx<R>*)A, ld, vcur, vnext, sh);
    tau = tn;
    cplx<R>* t = vcur;
    vcur = vnext;
    vnext = t;
  }
  make_w<R, NMAX>(n, n - 2, tau, vcur, sh);  // w of the last reflector
  tm.mark(1);
  if (tid == 0) {
    cplx<R> a = M.load(M.off(n - 1, n - 1));
    const cplx<R> vv = vcur[n - 1], ww = sh.w[n - 1];
    a = a - (vv * conj(ww) + ww * conj(vv));
    sh.de[n - 1].x = a.re;
    M.store(M.off(n - 1, n - 1), mk<R>(0, 0));  // no reflector for the last column
    sh.de[n - 1].y = 0;
  }
  __syncthreads();
}

// W = L^-1 of the Cholesky factor of G (chol_inv.hpp), out of line; -1 or the failing column.
template <typename R, int NMAX>
__device__ __noinline__ int chol_phase(int n, cplx<R>* L, int ld, cplx<R>* T, int ldt,
                                       unsigned long long* prof) {
  WgTeam tm{static_cast<int>(threadIdx.x), uniform(prof)};
  return chol_inverse<R, EigCfg<NMAX>::NT>(tm, uniform(n), global_ptr(L), uniform(ld),
                                           global_ptr(T), uniform(ldt), eig_sh<R, NMAX>.ci);
}

// eigenpairs of the tridiagonal in sh.de by divide and conquer (tridiag_dc.hpp), eigenvectors
// into Z; work: 2 n^2 reals of scratch; out of line
template <typename R, int NMAX>
__device__ __noinline__ void dc_phase(int n, R* Z, int ld, R* work, unsigned long long* prof) {
  EigShared<R, NMAX>& sh = eig_sh<R, NMAX>;
  n = uniform(n);
  WgTeam tm{static_cast<int>(threadIdx.x), uniform(prof)};
  BIPP_GLOBAL R* w = global_ptr(work);
  dc_tridiag<R, NMAX, EigCfg<NMAX>::NT>(tm, n, &sh.de[0].x, &sh.de[0].y, 2, global_ptr(Z),
                                        uniform(ld), w, w + (size_t)n * n, sh.dc, sh.dct);
}

// records the wall clock at phase boundaries when the caller asked for a profile
template <typename R>
__device__ __forceinline__ void eig_mark(const EigParams<R>& p, int sid, int k) {
  if (p.prof && threadIdx.x == 0) p.prof[(size_t)sid * EIG_PROF_SLOTS + k] = wall_clock64();
}

// X = Q Z[:, perm] for Q = H_0 H_1 ... H_{n-2} (H_c = I - tau_c v_c v_c^H, v_c below the
// subdiagonal of column c of A, tau_c on its diagonal), float64 on v_mfma_f64_16x16x4f64.
// The reflectors go in blocks of 16 from the last, H_cs ... H_c1 = I - V T V^H (compact WY,
// T upper triangular: T_ii = tau_i, T[0:i, i] = -tau_i T[0:i, 0:i] (V^H V)[0:i, i]).  The T
// of every block is formed first (V^H V on MFMA, the recurrence on 16 lanes) and kept in the
// strict upper triangle of A, which nothing reads from the tridiagonalisation until the
// un-whitening overwrites A.  Then panels of Z[:, perm] stay in registers while every block
// is applied: X -= V (T (V^H X)).  A panel holds one 16-column block per wavefront pair;
// the two wavefronts of a pair take alternate 16-row blocks, so V^H X needs one sum of the
// pair's two partial products through LDS.  MFMA layouts (l: lane): A[l % 16][l / 16],
// B[l / 16][l % 16], result register r at row l / 16 + 4 r, column l % 16; a result
// block's register s is therefore the B operand for its rows 4 s .. 4 s + 3, and every
// product chains in registers.  Row blocks above a block's first unit entry (where V is
// zero) are skipped.
template <int NMAX>
__device__ __noinline__ void back_transform_mfma(int n, cplx<double>* A_, int ld,
                                                 const double* Z_, cplx<double>* X_, int ldx) {
  EigShared<double, NMAX>& sh = eig_sh<double, NMAX>;
  n = uniform(n);
  ld = uniform(ld);
  ldx = uniform(ldx);
  BIPP_GLOBAL cplx<double>* A = global_ptr(A_);
  const BIPP_GLOBAL double* Z = global_ptr(Z_);
  BIPP_GLOBAL cplx<double>* X = global_ptr(X_);
  constexpr int NT_ = EigCfg<NMAX>::NT, NW_ = EigCfg<NMAX>::NW;
  constexpr int NRB = NMAX / 16, RBW = NRB / 2, PW = NW_ / 2 * 16;
  static_assert(NRB % 2 == 0 && NW_ % 2 == 0, "row blocks and wavefronts in pairs");
  typedef double f64x4 __attribute__((ext_vector_type(4)));
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int half = wave & 1, cg = wave >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  BtBlock<double, NMAX>& bb = sh.bb;
  const int nblk = (n - 1 + 15) / 16;  // reflectors 0 .. n - 2
  const f64x4 z4 = {0, 0, 0, 0};

  // stage block b: bb.v[q] = v_{cs + q} (zero for cs + q < 0), bb.tau
  auto stage = [&](int b) {
    const int cs = n - 2 - 16 * b - 15;
    for (int e = tid; e < 16 * NMAX; e += NT_) {
      const int q = e / NMAX, i = e % NMAX, c = cs + q;
      cplx<double> val = mk<double>(0, 0);
      if (c >= 0 && i < n) {
        if (i == c + 1)
          val = mk<double>(1, 0);
        else if (i > c + 1)
          val = ci_get(A[i + (size_t)c * ld]);
      }
      bb.v[q][i] = val;
    }
    if (tid < 16)
      bb.tau[tid] =
          cs + tid >= 0 ? ci_get(A[(cs + tid) + (size_t)(cs + tid) * ld]) : mk<double>(0, 0);
  };
  // first row block holding a nonzero of block b's V
  auto rb_first = [&](int b) {
    const int cs = n - 2 - 16 * b - 15;
    return ((cs < 0 ? 0 : cs) + 1) / 16;
  };

  // ---- T of every block, into the strict upper triangle of A --------------------------
  for (int b = 0; b < nblk; ++b) {
    const int cs = n - 2 - 16 * b - 15, rb0 = rb_first(b);
    __syncthreads();  // the previous block's LDS is no longer read
    stage(b);
    __syncthreads();
    f64x4 gr = z4, gi = z4;
    for (int rb = wave; rb < NRB; rb += NW_) {
      if (rb < rb0) continue;  // uniform over the wavefront
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const cplx<double> v = bb.v[lr][rb * 16 + 4 * s4 + lk];
        gr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, v.re, gr, 0, 0, 0);
        gr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, v.im, gr, 0, 0, 0);
        gi = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, v.im, gi, 0, 0, 0);
        gi = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.im, v.re, gi, 0, 0, 0);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bb.red[wave][lane][r] = gr[r];
      bb.red[wave][lane][4 + r] = gi[r];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double sr = 0, si = 0;
        for (int w = 0; w < NW_; ++w) {
          sr += bb.red[w][lane][r];
          si += bb.red[w][lane][4 + r];
        }
        bb.g[lk + 4 * r][lr] = mk<double>(sr, si);
      }
      wave_sync();
      // lanes p < 16: column i of T from the columns before it
      for (int i = 0; i < 16; ++i) {
        if (lane < 16) {
          const int pr = lane;
          cplx<double> val = mk<double>(0, 0);
          if (pr == i) {
            val = bb.tau[i];
          } else if (pr < i) {
            cplx<double> acc = mk<double>(0, 0);
            for (int q = pr; q < i; ++q) acc = acc + bb.t[pr][q] * bb.g[q][i];
            val = mk<double>(0, 0) - bb.tau[i] * acc;
          }
          bb.t[pr][i] = val;
        }
        wave_sync();
      }
      // store T[p][q], p <= q, at A[cs + p, cs + q + 1] (strict upper triangle)
      for (int e = lane; e < 256; e += 64) {
        const int pr = e / 16, q = e % 16;
        if (pr <= q && cs + pr >= 0) ci_put(A[(cs + pr) + (size_t)(cs + q + 1) * ld], bb.t[pr][q]);
      }
    }
  }
  __syncthreads();

  // ---- panels ------------------------------------------------------------------------
  for (int j0 = 0; j0 < n; j0 += PW) {
    const int col = j0 + cg * 16 + lr;
    const int pc = col < n ? sh.perm[col] : 0;
    f64x4 xr[RBW], xi[RBW];
#pragma unroll
    for (int t = 0; t < RBW; ++t) {
      const int rb = half + 2 * t;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + lk + 4 * r;
        xr[t][r] = (row < n && col < n) ? Z[row + (size_t)pc * ld] : 0.0;
      }
      xi[t] = z4;
    }
    for (int b = 0; b < nblk; ++b) {
      const int cs = n - 2 - 16 * b - 15, rb0 = rb_first(b);
      __syncthreads();  // the previous block's LDS is no longer read
      stage(b);
      for (int e = tid; e < 256; e += NT_) {
        const int pr = e / 16, q = e % 16;
        bb.t[pr][q] = (pr <= q && cs + pr >= 0) ? ci_get(A[(cs + pr) + (size_t)(cs + q + 1) * ld])
                                                : mk<double>(0, 0);
      }
      __syncthreads();
      // W1 = V^H X over this wavefront's row blocks
      f64x4 wr = z4, wi = z4;
#pragma unroll
      for (int t = 0; t < RBW; ++t) {
        const int rb = half + 2 * t;
        if (rb < rb0) continue;  // uniform over the wavefront
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const cplx<double> v = bb.v[lr][rb * 16 + 4 * s4 + lk];
          wr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, xr[t][s4], wr, 0, 0, 0);
          wr = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, xi[t][s4], wr, 0, 0, 0);
          wi = __builtin_amdgcn_mfma_f64_16x16x4f64(v.re, xi[t][s4], wi, 0, 0, 0);
          wi = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.im, xr[t][s4], wi, 0, 0, 0);
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        bb.red[wave][lane][r] = wr[r];
        bb.red[wave][lane][4 + r] = wi[r];
      }
      __syncthreads();
      // the pair's sum, in the same order in both wavefronts
      f64x4 w1r, w1i;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        w1r[r] = bb.red[wave & ~1][lane][r] + bb.red[wave | 1][lane][r];
        w1i[r] = bb.red[wave & ~1][lane][4 + r] + bb.red[wave | 1][lane][4 + r];
      }
      // W2 = T W1
      f64x4 w2r = z4, w2i = z4;
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        const cplx<double> tv = bb.t[lr][4 * s4 + lk];
        w2r = __builtin_amdgcn_mfma_f64_16x16x4f64(tv.re, w1r[s4], w2r, 0, 0, 0);
        w2r = __builtin_amdgcn_mfma_f64_16x16x4f64(-tv.im, w1i[s4], w2r, 0, 0, 0);
        w2i = __builtin_amdgcn_mfma_f64_16x16x4f64(tv.re, w1i[s4], w2i, 0, 0, 0);
        w2i = __builtin_amdgcn_mfma_f64_16x16x4f64(tv.im, w1r[s4], w2i, 0, 0, 0);
      }
      // X -= V W2
#pragma unroll
      for (int t = 0; t < RBW; ++t) {
        const int rb = half + 2 * t;
        if (rb < rb0) continue;  // uniform over the wavefront
#pragma unroll
        for (int s4 = 0; s4 < 4; ++s4) {
          const cplx<double> v = bb.v[4 * s4 + lk][rb * 16 + lr];
          xr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.re, w2r[s4], xr[t], 0, 0, 0);
          xr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(v.im, w2i[s4], xr[t], 0, 0, 0);
          xi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.re, w2i[s4], xi[t], 0, 0, 0);
          xi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-v.im, w2r[s4], xi[t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < RBW; ++t) {
      const int rb = half + 2 * t;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb * 16 + lk + 4 * r;
        if (row < n && col < n) ci_put(X[row + (size_t)col * ldx], mk<double>(xr[t][r], xi[t][r]));
      }
    }
  }
  __syncthreads();
}

template <typename R, typename TIn, int NMAX>
__device__ void solve_one(const EigParams<R>& p, int sid, EigShared<R, NMAX>& sh) {
  constexpr int NT_ = EigCfg<NMAX>::NT, NW_ = EigCfg<NMAX>::NW, RM = EigCfg<NMAX>::RM,
                UC = EigCfg<NMAX>::UC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = p.n_red[sid];
  const int nmax = p.nmax;
  R* dout = p.d_out + (size_t)sid * p.stride_d;
  if (n <= 0) {
    for (int j = tid; j < nmax; j += NT_) dout[j] = 0;
    if (tid == 0) p.info[sid] = 0;
    return;
  }
  const int ld = p.ws_ld;
  const size_t slot = blockIdx.x;
  cplx<R>* A = p.ws_a + slot * p.ws_stride;
  cplx<R>* L = p.ws_l + slot * p.ws_stride;
  R* Z = p.ws_z + slot * p.ws_stride;
  cplx<R>* X = p.z_out + (size_t)sid * p.stride_z;
  const int ldx = p.ldz;

  for (int i = tid; i < n; i += NT_) sh.idx[i] = p.idx[(size_t)sid * p.idx_stride + i];
  __syncthreads();
  eig_mark(p, sid, 0);

  // ---- load the reduced pencil -------------------------------------------------------
  {
    const cplx<TIn>* S = reinterpret_cast<const cplx<TIn>*>(p.s) + (size_t)sid * p.stride_s;
    const double* xyz = p.xyz + (size_t)sid * p.stride_xyz;
    const R wl = static_cast<R>(p.wl[sid]);
    const R pref = static_cast<R>(p.gram_prefactor);
    const cplx<R>* Gin = p.g_in ? p.g_in + (size_t)sid * p.stride_g : nullptr;
    for (int j = wave; j < n; j += NW_) {
      const int bj = sh.idx[j];
      for (int i = lane; i < n; i += 64) {
        const int bi = sh.idx[i];
        cplx<TIn> sv = (bi >= bj) ? S[bi + (size_t)bj * p.lds] : conj(S[bj + (size_t)bi * p.lds]);
        cplx<R> a = cvt<R>(sv);
        if (i == j) a.im = 0;
        A[i + (size_t)j * ld] = a;
        if (!p.standard) {
          cplx<R> g;
          if (Gin) {
            g = Gin[i + (size_t)j * p.ld_g];
          } else if (i == j) {
            g = mk<R>(pref, 0);
          } else {
            const R dx = static_cast<R>(xyz[bi]) - static_cast<R>(xyz[bj]);
            const R dy = static_cast<R>(xyz[bi + p.ldxyz]) - static_cast<R>(xyz[bj + p.ldxyz]);
            const R dz =
                static_cast<R>(xyz[bi + 2 * p.ldxyz]) - static_cast<R>(xyz[bj + 2 * p.ldxyz]);
            const R x = R(2) * sqrt(dx * dx + dy * dy + dz * dz) / wl;
            const R sc = x != R(0) ? sinpi_r<R>(x) / (R(3.14159265358979323846) * x) : R(1);
            g = mk<R>(pref * sc, 0);
          }
          L[i + (size_t)j * ld] = g;
        }
      }
    }
    __syncthreads();
  }
  eig_mark(p, sid, 1);

  if (!p.standard) {
    // ---- W = L^-1 of the Cholesky factor G = L L^H (blocked; X is its scratch) ----------
    {
      const int bad = chol_phase<R, NMAX>(n, L, ld, X, ldx,
                                          p.prof ? p.prof + (size_t)sid * EIG_PROF_SLOTS + 18 : nullptr);
      __syncthreads();
      if (bad >= 0) {
        // leading minor bad + 1 of G is not positive definite: ?hegv reports info = n + bad + 1
        for (int jj = tid; jj < nmax; jj += NT_) dout[jj] = 0;
        if (tid == 0) p.info[sid] = n + bad + 1;
        __syncthreads();
        return;
      }
    }
    eig_mark(p, sid, 2);
    // ---- C = W S W^H -------------------------------------------------------------------
    const cplx<R>* W = L;
    // X = W S (W lower triangular: k <= i)
    gemm_tiles<R, NMAX, false, true, false>(
        n, [=](int i, int k) { return k <= i ? W[i + (size_t)k * ld] : mk<R>(0, 0); },
        [=](int k, int j) { return A[k + (size_t)j * ld]; },
        [=](int i0, int, int& klo, int& khi) {
          klo = 0;
          khi = min(n, i0 + GT);
        },
        X, ldx, sh);
    __syncthreads();
    // A = X W^H on and below the diagonal ((W^H)[k, j] = conj(W[j, k]), k <= j; the
    // tridiagonalisation reads the lower triangle only), then a real diagonal
    gemm_tiles<R, NMAX, false, false, true>(
        n, [=](int i, int k) { return X[i + (size_t)k * ldx]; },
        [=](int k, int j) { return k <= j ? conj(W[j + (size_t)k * ld]) : mk<R>(0, 0); },
        [=](int, int j0, int& klo, int& khi) {
          klo = 0;
          khi = min(n, j0 + GT);
        },
        A, ld, sh);
    __syncthreads();
    for (int i = tid; i < n; i += NT_) A[i + (size_t)i * ld].im = 0;
    __syncthreads();
  }
  eig_mark(p, sid, 3);

  // ---- Householder tridiagonalisation (zhetd2, lower, fused her2 + hemv) --------------
  if (n == 1) {
    if (tid == 0) {
      sh.de[0].x = A[0].re;
      sh.de[0].y = 0;
    }
    __syncthreads();
  } else {
    tridiag_phase<R, NMAX>(n, A, ld, p.prof ? p.prof + (size_t)sid * EIG_PROF_SLOTS + 16 : nullptr);
  }
  eig_mark(p, sid, 4);

  // ---- eigenpairs of the real tridiagonal by divide and conquer, eigenvectors into Z ----
  // (X, free until the back-transform, holds the two n x n real work arrays)
  dc_phase<R, NMAX>(n, Z, ld, reinterpret_cast<R*>(X),
                    p.prof ? p.prof + (size_t)sid * EIG_PROF_SLOTS + 8 : nullptr);
  eig_mark(p, sid, 5);

  // ---- ascending order: rank sort --------------------------------------------------
  // (NaN ordered as +inf, ties by index: the ranks always form a permutation of [0, n))
  for (int i = tid; i < n; i += NT_) {
    const R di0 = sh.de[i].x;
    const R di = di0 == di0 ? di0 : std::numeric_limits<R>::infinity();
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const R dj0 = sh.de[j].x;
      const R dj = dj0 == dj0 ? dj0 : std::numeric_limits<R>::infinity();
      rank += (dj < di) || (dj == di && j < i);
    }
    sh.perm[rank] = i;
  }
  __syncthreads();
  for (int j = tid; j < nmax; j += NT_) dout[j] = j < n ? sh.de[sh.perm[j]].x : R(0);

  // ---- back-transform X = Q Z[:, perm] ------------------------------------------------
  if constexpr (std::is_same<R, double>::value && bt_mfma(NMAX)) {
    back_transform_mfma<NMAX>(n, A, ld, Z, X, ldx);
  } else
  // Each group of LPC lanes keeps CPG columns of Z[:, perm] in registers (rows g + LPC r)
  // while H_{n-2}, ..., H_0 are applied to them (the columns' dependent chains of dot
  // product, reduction and update interleave), the reflectors staged through LDS RB at a
  // time; row blocks that lie above a reflector's unit entry are skipped.
  {
    constexpr int LPC = BtCfg<NMAX>::LPC, RR = BtCfg<NMAX>::RR, CPWV = BtCfg<NMAX>::CPWV,
                  BT_RB = BtCfg<NMAX>::RB, RG = RR >= 4 ? 4 : RR, CPG = BtCfg<NMAX>::CPG;
    static_assert(LPC == 16 || LPC == 32, "column groups of 16 or 32 lanes");
    BackTile<R, NMAX>& bt = sh.bt;
    const int g = lane % LPC;
    for (int jb = 0; jb < n; jb += NW_ * CPWV * CPG) {
      int jc[CPG];
      cplx<R> x[CPG][RR];
#pragma unroll
      for (int u = 0; u < CPG; ++u) {
        jc[u] = jb + (wave * CPWV + lane / LPC) * CPG + u;
        const bool live = jc[u] < n;
        const int pj = live ? sh.perm[jc[u]] : 0;
#pragma unroll
        for (int r = 0; r < RR; ++r) {
          const int i = g + LPC * r;
          x[u][r] = mk<R>(live && i < n ? Z[i + (size_t)pj * ld] : R(0), R(0));
        }
      }
      // the chunk of reflectors cb, cb - 1, ... is fetched into registers one chunk ahead
      constexpr int PER = BT_RB * NMAX / NT_;
      static_assert(BT_RB * NMAX % NT_ == 0, "reflector chunk splits evenly");
      cplx<R> rv[PER], rt = mk<R>(0, 0);
      auto fetch = [&](int cb) {
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int e = tid + u * NT_;
          const int q = e / NMAX, i = e % NMAX, c = cb - q;
          cplx<R> val = mk<R>(0, 0);
          if (c >= 0 && i < n) {
            if (i == c + 1)
              val = mk<R>(1, 0);
            else if (i > c + 1)
              val = A[i + (size_t)c * ld];
          }
          rv[u] = val;
        }
        if (tid < BT_RB)
          rt = cb - tid >= 0 ? A[(cb - tid) + (size_t)(cb - tid) * ld] : mk<R>(0, 0);
      };
      if (n >= 2) fetch(n - 2);
      for (int cb = n - 2; cb >= 0; cb -= BT_RB) {
        __syncthreads();  // the previous chunk is no longer read
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int e = tid + u * NT_;
          bt.v[e / NMAX][e % NMAX] = rv[u];
        }
        if (tid < BT_RB) bt.tau[tid] = rt;
        __syncthreads();
        if (cb - BT_RB >= 0) fetch(cb - BT_RB);
        for (int q = 0; q < BT_RB; ++q) {
          const int c = cb - q;
          if (c < 0) break;
          const cplx<R> tau = bt.tau[q];
          if (tau.re == R(0) && tau.im == R(0)) continue;  // uniform over the workgroup
          // row blocks r < RG g0 hold rows <= c only, where v = 0: skipped RG blocks at a
          // time (a uniform branch per group, the group's LDS loads issued together)
          const int g0 = (c + 1) / (LPC * RG);
          cplx<R> dt[CPG];
#pragma unroll
          for (int u = 0; u < CPG; ++u) dt[u] = mk<R>(0, 0);
#pragma unroll
          for (int gq = 0; gq < RR / RG; ++gq) {
            if (gq >= g0) {
              cplx<R> vv[RG];
#pragma unroll
              for (int w = 0; w < RG; ++w) vv[w] = bt.v[q][g + LPC * (gq * RG + w)];
#pragma unroll
              for (int w = 0; w < RG; ++w)
#pragma unroll
                for (int u = 0; u < CPG; ++u)
                  dt[u] = dt[u] + cmul_conj_a(vv[w], x[u][gq * RG + w]);
            }
          }
          cplx<R> t[CPG];
#pragma unroll
          for (int u = 0; u < CPG; ++u) {
            dt[u].re = sum16(dt[u].re);
            dt[u].im = sum16(dt[u].im);
            if (LPC == 32) {
              dt[u].re += __shfl_xor(dt[u].re, 16, 64);
              dt[u].im += __shfl_xor(dt[u].im, 16, 64);
            }
            t[u] = tau * dt[u];
          }
#pragma unroll
          for (int gq = 0; gq < RR / RG; ++gq) {
            if (gq >= g0) {
              cplx<R> vv[RG];
#pragma unroll
              for (int w = 0; w < RG; ++w) vv[w] = bt.v[q][g + LPC * (gq * RG + w)];
#pragma unroll
              for (int w = 0; w < RG; ++w)
#pragma unroll
                for (int u = 0; u < CPG; ++u)
                  x[u][gq * RG + w] = x[u][gq * RG + w] - vv[w] * t[u];
            }
          }
        }
      }
#pragma unroll
      for (int u = 0; u < CPG; ++u) {
        if (jc[u] < n) {
#pragma unroll
          for (int r = 0; r < RR; ++r) {
            const int i = g + LPC * r;
            if (i < n) X[i + (size_t)jc[u] * ldx] = x[u][r];
          }
        }
      }
    }
    __syncthreads();
  }

  eig_mark(p, sid, 6);
  // ---- un-whiten: eigenvectors of the pencil = W^H X ----------------------------------
  if (!p.standard) {
    // X is copied into A (free after the back-transform), then X = W^H A with
    // (W^H)[i, k] = conj(W[k, i]) for k >= i
    for (int j = wave; j < n; j += NW_)
      for (int i = lane; i < n; i += 64) A[i + (size_t)j * ld] = X[i + (size_t)j * ldx];
    __syncthreads();
    const cplx<R>* W = L;
    gemm_tiles<R, NMAX, true, true, false>(
        n, [=](int i, int k) { return k >= i ? conj(W[k + (size_t)i * ld]) : mk<R>(0, 0); },
        [=](int k, int j) { return A[k + (size_t)j * ld]; },
        [=](int i0, int, int& klo, int& khi) {
          klo = i0;
          khi = n;
        },
        X, ldx, sh);
    __syncthreads();
  }
  eig_mark(p, sid, 7);
  if (tid == 0) p.info[sid] = 0;
  __syncthreads();
}

template <typename R, typename TIn, int NMAX>
__global__ __launch_bounds__(EigCfg<NMAX>::NT) void eig_kernel(EigParams<R> p) {
  EigShared<R, NMAX>& sh = eig_sh<R, NMAX>;
  __shared__ int job_sh;
  for (;;) {
    if (threadIdx.x == 0) job_sh = static_cast<int>(atomicAdd(p.counter, 1u));
    __syncthreads();
    const int job = job_sh;
    __syncthreads();
    if (job >= p.n_jobs) break;
    const int sid = p.order ? p.order[job] : job;
    solve_one<R, TIn, NMAX>(p, sid, sh);
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------
// Flag live entries and compact kept beams (src/eigensolver.cpp:145-168).
// One workgroup per sample.
// --------------------------------------------------------------------------------------
// |v|^2 as std::norm forms it: both squares and the sum rounded separately (a fused
// multiply-add would change which entries sit above eps at the threshold)
template <typename T>
__device__ inline T norm_unfused(T re, T im) {
#pragma clang fp contract(off)
  const T a = re * re;
  const T b = im * im;
  return a + b;
}

template <typename TIn>
__global__ __launch_bounds__(NT) void flag_kernel(FlagParams p) {
  extern __shared__ int flag_sh[];
  __shared__ unsigned long long cnt_sh[NW];
  const int sid = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = p.n_beam;
  const cplx<TIn>* S = reinterpret_cast<const cplx<TIn>*>(p.s) + (size_t)sid * p.stride_s;
  const TIn eps = sizeof(TIn) == 4 ? TIn(1.1920928955078125e-07) : TIn(2.220446049250313e-16);
  for (int i = tid; i < nb; i += NT) flag_sh[i] = 0;
  __syncthreads();
  unsigned long long cnt = 0;
  for (int c = wave; c < nb; c += NW) {
    bool any = false;
    for (int r = c + lane; r < nb; r += 64) {
      const cplx<TIn> v = S[r + (size_t)c * p.lds];
      const bool live = p.mode == 0 ? (norm_unfused(v.re, v.im) > eps)
                                    : (fabs(v.re) >= eps || fabs(v.im) >= eps);
      if (live) {
        flag_sh[r] = 1;
        cnt += 1 + (r != c);
        any = true;
      }
    }
    if (__any(any) && lane == 0) flag_sh[c] = 1;
  }
  cnt = wave_sum(cnt);
  if (lane == 0) cnt_sh[wave] = cnt;
  __syncthreads();
  if (tid == 0) {
    unsigned long long total = 0;
    for (int w = 0; w < NW; ++w) total += cnt_sh[w];
    p.n_vis[sid] = total;
    int nr = 0;
    int* idx = p.idx + (size_t)sid * p.idx_stride;
    for (int i = 0; i < nb; ++i)
      if (flag_sh[i]) idx[nr++] = i;
    p.n_red[sid] = nr;
  }
}

// Scatter reduced eigenvectors to (nAntenna x nBeam) with W = I: V[idx[i], e] = Z[i, e].
template <typename R, typename TOut>
__global__ void unbeam_identity_kernel(UnbeamParams<R> p) {
  const int sid = blockIdx.y;
  const int e = blockIdx.x;
  const int n = p.n_red[sid];
  cplx<TOut>* V = reinterpret_cast<cplx<TOut>*>(p.v) + (size_t)sid * p.stride_v + (size_t)e * p.ldv;
  const int* idx = p.idx + (size_t)sid * p.idx_stride;
  for (int i = threadIdx.x; i < p.n_ant; i += blockDim.x) V[i] = mk<TOut>(0, 0);
  __syncthreads();
  if (e >= n) return;
  const cplx<R>* Zc = p.z + (size_t)sid * p.stride_z + (size_t)e * p.ldz;
  for (int i = threadIdx.x; i < n; i += blockDim.x) V[idx[i]] = cvt<TOut>(Zc[i]);
}

template <typename R, typename TOut>
__global__ void cast_d_kernel(int n_samples, int nmax, const R* d, long long stride_d, TOut* out,
                              long long stride_out, const unsigned long long* n_vis, TOut* scale) {
  const int sid = blockIdx.x;
  for (int j = threadIdx.x; j < nmax; j += blockDim.x)
    out[(size_t)sid * stride_out + j] = static_cast<TOut>(d[(size_t)sid * stride_d + j]);
  if (threadIdx.x == 0 && scale) {
    const unsigned long long nv = n_vis[sid];
    scale[sid] = nv ? TOut(1) / static_cast<TOut>(nv) : TOut(0);
  }
}

__global__ void fill_u32_kernel(unsigned* p, unsigned v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

void launch_fill_u32(unsigned* p, unsigned v, int n, hipStream_t stream) {
  fill_u32_kernel<<<ceil_div(n, 256), 256, 0, stream>>>(p, v, n);
  check_launch("fill_u32_kernel");
}

void launch_flag(const FlagParams& p, int n_samples, bool in_double, hipStream_t stream) {
  if (n_samples <= 0) return;
  const size_t shmem = sizeof(int) * (size_t)p.n_beam;
  if (in_double)
    flag_kernel<double><<<n_samples, NT, shmem, stream>>>(p);
  else
    flag_kernel<float><<<n_samples, NT, shmem, stream>>>(p);
  check_launch("flag_kernel");
}

template <typename R>
int eig_nmax_bucket(int nmax) {
  if (nmax <= 64) return 64;
  if (nmax <= 128) return 128;
  if (nmax <= 256) return 256;
  if (nmax <= 512) return 512;
  return -1;
}

template <typename R>
void launch_eig(const EigParams<R>& p, int n_slots, bool in_double, hipStream_t stream) {
  if (p.n_jobs <= 0) return;
  const int b = eig_nmax_bucket<R>(p.nmax);
  if (b < 0) fail(BIPP_INVALID_PARAMETER_ERROR, "eigensolver supports at most 512 beams");
#define BIPP_EIG_CASE(NM)                                                               \
  if (b == NM) {                                                                        \
    if (in_double)                                                                      \
      eig_kernel<R, double, NM><<<n_slots, EigCfg<NM>::NT, 0, stream>>>(p);             \
    else                                                                                \
      eig_kernel<R, float, NM><<<n_slots, EigCfg<NM>::NT, 0, stream>>>(p);              \
  }
  BIPP_EIG_CASE(64)
  BIPP_EIG_CASE(128)
  BIPP_EIG_CASE(256)
  BIPP_EIG_CASE(512)
#undef BIPP_EIG_CASE
  check_launch("eig_kernel");
}

template <typename R>
void launch_unbeam_identity(const UnbeamParams<R>& p, int n_samples, int n_beam, bool out_double,
                            hipStream_t stream) {
  if (n_samples <= 0 || n_beam <= 0) return;
  dim3 grid(n_beam, n_samples);
  if (out_double)
    unbeam_identity_kernel<R, double><<<grid, 256, 0, stream>>>(p);
  else
    unbeam_identity_kernel<R, float><<<grid, 256, 0, stream>>>(p);
  check_launch("unbeam_identity_kernel");
}

template <typename R>
void launch_cast_d(int n_samples, int nmax, const R* d, long long stride_d, void* out,
                   long long stride_out, bool out_double, const unsigned long long* n_vis,
                   void* scale, hipStream_t stream) {
  if (n_samples <= 0) return;
  if (out_double)
    cast_d_kernel<R, double><<<n_samples, 256, 0, stream>>>(
        n_samples, nmax, d, stride_d, static_cast<double*>(out), stride_out, n_vis,
        static_cast<double*>(scale));
  else
    cast_d_kernel<R, float><<<n_samples, 256, 0, stream>>>(
        n_samples, nmax, d, stride_d, static_cast<float*>(out), stride_out, n_vis,
        static_cast<float*>(scale));
  check_launch("cast_d_kernel");
}

template <typename R>
int eig_blocks_per_cu(int nmax, bool in_double) {
  const int b = eig_nmax_bucket<R>(nmax);
  int nb = 1;
  const void* fn = nullptr;
  int nt = NT;
#define BIPP_EIG_FN(NM)                                                                      \
  if (b == NM) {                                                                             \
    fn = in_double ? reinterpret_cast<const void*>(&eig_kernel<R, double, NM>)               \
                   : reinterpret_cast<const void*>(&eig_kernel<R, float, NM>);               \
    nt = EigCfg<NM>::NT;                                                                     \
  }
  BIPP_EIG_FN(64)
  BIPP_EIG_FN(128)
  BIPP_EIG_FN(256)
  BIPP_EIG_FN(512)
#undef BIPP_EIG_FN
  if (!fn) return 1;
  BIPP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, nt, 0));
  return nb > 0 ? nb : 1;
}
template int eig_blocks_per_cu<float>(int, bool);
template int eig_blocks_per_cu<double>(int, bool);

template void launch_eig<float>(const EigParams<float>&, int, bool, hipStream_t);
template void launch_eig<double>(const EigParams<double>&, int, bool, hipStream_t);
template void launch_unbeam_identity<float>(const UnbeamParams<float>&, int, int, bool,
                                            hipStream_t);
template void launch_unbeam_identity<double>(const UnbeamParams<double>&, int, int, bool,
                                             hipStream_t);
template void launch_cast_d<float>(int, int, const float*, long long, void*, long long, bool,
                                   const unsigned long long*, void*, hipStream_t);
template void launch_cast_d<double>(int, int, const double*, long long, void*, long long, bool,
                                    const unsigned long long*, void*, hipStream_t);

}  // namespace bippamd
