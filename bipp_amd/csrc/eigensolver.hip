// Batched generalized Hermitian eigensolver for the Bluebild pencil (S, G) on gfx950.
//
// Restates, per sample, src/eigensolver.cpp:129-225 (flag -> reduce -> Gram -> ?hegv)
// with the LAPACK steps of ?hegv(itype=1,'V','L') (src/host/lapack_api.hpp:112-158):
//   Cholesky G = L L^H, C = L^-1 S L^-H, Householder tridiagonalisation C = Q T Q^H,
//   implicit QL on T (eigenvalues + rotations), X = Q Z, eigenvectors L^-H X,
//   eigenvalues ascending.
//
// Launch shape: a persistent grid of 256-thread workgroups; each workgroup pulls sample
// indices from a device counter (largest problems first) and solves one problem at a
// time, keeping its matrices in a per-workgroup slot of global memory (L2/MALL resident
// at these sizes) and its vectors in LDS.  Every step is a workgroup-wide parallel
// update separated by barriers, with lanes mapped to matrix rows so that global
// accesses are coalesced.
#include "common.hpp"
#include "kernels.hpp"

#include <limits>

namespace bippamd {

namespace {

constexpr int NT = 256;
constexpr int NW = NT / 64;

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

template <typename R, int NMAX>
struct EigShared {
  cplx<R> v[NMAX];
  cplx<R> w[NMAX];
  cplx<R> vn[NMAX];
  cplx<R> xw[NW][NMAX];
  R dd[NMAX];
  R ee[NMAX];
  R cs[NMAX];
  R sn[NMAX];
  int perm[NMAX];
  int idx[NMAX];
  R red[2 * NW];
  R scal[4];
  int iscal[4];
  cplx<R> ctau;
};

// --------------------------------------------------------------------------------------
// Forward substitution B <- L^-1 B on all n columns of B (column-major, ld).
// --------------------------------------------------------------------------------------
template <typename R, int NMAX>
__device__ void fwd_solve(int n, const cplx<R>* __restrict__ L, int ldl, cplx<R>* __restrict__ B,
                          int ldb, EigShared<R, NMAX>& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int k = 0; k < n; ++k) {
    const R linv = R(1) / L[k + (size_t)k * ldl].re;
    for (int j = tid; j < n; j += NT) {
      cplx<R> r = B[k + (size_t)j * ldb] * linv;
      B[k + (size_t)j * ldb] = r;
      sh.w[j] = r;
    }
    for (int i = k + 1 + tid; i < n; i += NT) sh.v[i] = L[i + (size_t)k * ldl];
    __syncthreads();
    for (int j = wave; j < n; j += NW) {
      const cplx<R> rj = sh.w[j];
      cplx<R>* col = B + (size_t)j * ldb;
      for (int i = k + 1 + lane; i < n; i += 64) col[i] = col[i] - sh.v[i] * rj;
    }
    __syncthreads();
  }
}

// Backward substitution B <- L^-H B.
template <typename R, int NMAX>
__device__ void bwd_solve_h(int n, const cplx<R>* __restrict__ L, int ldl, cplx<R>* __restrict__ B,
                            int ldb, EigShared<R, NMAX>& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int k = n - 1; k >= 0; --k) {
    const R linv = R(1) / L[k + (size_t)k * ldl].re;
    for (int j = tid; j < n; j += NT) {
      cplx<R> r = B[k + (size_t)j * ldb] * linv;
      B[k + (size_t)j * ldb] = r;
      sh.w[j] = r;
    }
    for (int i = tid; i < k; i += NT) sh.v[i] = conj(L[k + (size_t)i * ldl]);
    __syncthreads();
    for (int j = wave; j < n; j += NW) {
      const cplx<R> rj = sh.w[j];
      cplx<R>* col = B + (size_t)j * ldb;
      for (int i = lane; i < k; i += 64) col[i] = col[i] - sh.v[i] * rj;
    }
    __syncthreads();
  }
}

// Householder reflector (zlarfg semantics) annihilating A[c+2:n, c]; writes v (v[c+1]=1)
// into vec and into A[c+2:n, c], beta into A[c+1, c] and ee[c]; returns tau.
template <typename R, int NMAX>
__device__ cplx<R> reflector(int n, int c, cplx<R>* __restrict__ A, int ld, cplx<R>* vec,
                             EigShared<R, NMAX>& sh) {
  const int tid = threadIdx.x;
  cplx<R>* col = A + (size_t)c * ld;
  const cplx<R> alpha = col[c + 1];
  R part = 0;
  for (int i = c + 2 + tid; i < n; i += NT) part += abs2(col[i]);
  const R xnorm2 = block_sum<NT>(part, sh.red);
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
  for (int i = c + 2 + tid; i < n; i += NT) {
    cplx<R> vi = col[i] * scal;
    col[i] = vi;
    vec[i] = vi;
  }
  if (tid == 0) {
    vec[c + 1] = mk<R>(1, 0);
    col[c + 1] = mk<R>(beta, 0);
    sh.ee[c] = beta;
  }
  __syncthreads();
  return tau;
}

// One pass over the trailing square [i0,n)^2 of the full Hermitian A:
//   optionally A -= v w^H + w v^H (pending rank-2 update), then x = A * vn.
// Result x[i0:n) is written to sh.w (v and w are no longer needed afterwards).
template <typename R, int NMAX, bool APPLY>
__device__ void trail_pass(int n, int i0, cplx<R>* __restrict__ A, int ld, const cplx<R>* v,
                           const cplx<R>* vn, EigShared<R, NMAX>& sh) {
  constexpr int RMAX = NMAX / 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  cplx<R> acc[RMAX], vi[RMAX], wi[RMAX];
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    acc[r] = mk<R>(0, 0);
    const int i = i0 + lane + 64 * r;
    vi[r] = (APPLY && i < n) ? v[i] : mk<R>(0, 0);
    wi[r] = (APPLY && i < n) ? sh.w[i] : mk<R>(0, 0);
  }
  for (int j = i0 + wave; j < n; j += NW) {
    cplx<R> cvj = mk<R>(0, 0), cwj = mk<R>(0, 0);
    if (APPLY) {
      cvj = conj(v[j]);
      cwj = conj(sh.w[j]);
    }
    const cplx<R> vnj = vn[j];
    cplx<R>* col = A + (size_t)j * ld;
#pragma unroll
    for (int r = 0; r < RMAX; ++r) {
      const int i = i0 + lane + 64 * r;
      if (i < n) {
        cplx<R> a = col[i];
        if (APPLY) {
          a = a - (vi[r] * cwj + wi[r] * cvj);
          col[i] = a;
        }
        acc[r] = acc[r] + a * vnj;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RMAX; ++r) {
    const int i = i0 + lane + 64 * r;
    if (i < n) sh.xw[wave][i] = acc[r];
  }
  __syncthreads();
  for (int i = i0 + tid; i < n; i += NT) {
    cplx<R> s = sh.xw[0][i];
#pragma unroll
    for (int w = 1; w < NW; ++w) s = s + sh.xw[w][i];
    sh.w[i] = s;
  }
  __syncthreads();
}

// w = tau*x - 1/2 tau (tau x)^H v v, for the reflector stored in vec over [c+1, n).
// On entry sh.w holds x (un-scaled); on exit sh.w holds w.
template <typename R, int NMAX>
__device__ void make_w(int n, int c, cplx<R> tau, const cplx<R>* vec, EigShared<R, NMAX>& sh) {
  const int tid = threadIdx.x;
  R pr = 0, pi = 0;
  for (int i = c + 1 + tid; i < n; i += NT) {
    cplx<R> x = tau * sh.w[i];
    sh.w[i] = x;
    cplx<R> t = cmul_conj_a(x, vec[i]);  // conj(x) v
    pr += t.re;
    pi += t.im;
  }
  const R dr = block_sum<NT>(pr, sh.red);
  const R di = block_sum<NT>(pi, sh.red);
  const cplx<R> alpha2 = tau * mk<R>(R(-0.5) * dr, R(-0.5) * di);
  for (int i = c + 1 + tid; i < n; i += NT) sh.w[i] = sh.w[i] + alpha2 * vec[i];
  __syncthreads();
}

template <typename R, typename TIn, int NMAX>
__device__ void solve_one(const EigParams<R>& p, int sid, EigShared<R, NMAX>& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int n = p.n_red[sid];
  const int nmax = p.nmax;
  R* dout = p.d_out + (size_t)sid * p.stride_d;
  if (n <= 0) {
    for (int j = tid; j < nmax; j += NT) dout[j] = 0;
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

  for (int i = tid; i < n; i += NT) sh.idx[i] = p.idx[(size_t)sid * p.idx_stride + i];
  __syncthreads();

  // ---- load the reduced pencil -------------------------------------------------------
  {
    const cplx<TIn>* S = reinterpret_cast<const cplx<TIn>*>(p.s) + (size_t)sid * p.stride_s;
    const double* xyz = p.xyz + (size_t)sid * p.stride_xyz;
    const R wl = static_cast<R>(p.wl[sid]);
    const R pref = static_cast<R>(p.gram_prefactor);
    const cplx<R>* Gin = p.g_in ? p.g_in + (size_t)sid * p.stride_g : nullptr;
    for (int j = wave; j < n; j += NW) {
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

  if (!p.standard) {
    // ---- Cholesky G = L L^H (lower, right-looking) -------------------------------------
    for (int j = 0; j < n; ++j) {
      if (tid == 0) {
        const R djj = L[j + (size_t)j * ld].re;
        const bool ok = djj > R(0) && djj == djj;
        const R sj = ok ? sqrt(djj) : R(1);
        L[j + (size_t)j * ld] = mk<R>(sj, 0);
        sh.scal[0] = R(1) / sj;
        sh.iscal[0] = ok;
      }
      __syncthreads();
      if (!sh.iscal[0]) {
        // leading minor j+1 of G is not positive definite: ?hegv reports info = n + j + 1
        for (int jj = tid; jj < nmax; jj += NT) dout[jj] = 0;
        if (tid == 0) p.info[sid] = n + j + 1;
        __syncthreads();
        return;
      }
      const R inv = sh.scal[0];
      for (int i = j + 1 + tid; i < n; i += NT) {
        cplx<R> u = L[i + (size_t)j * ld] * inv;
        L[i + (size_t)j * ld] = u;
        sh.v[i] = u;
      }
      __syncthreads();
      for (int k = j + 1 + wave; k < n; k += NW) {
        const cplx<R> ck = conj(sh.v[k]);
        cplx<R>* col = L + (size_t)k * ld;
        for (int i = k + lane; i < n; i += 64) col[i] = col[i] - sh.v[i] * ck;
      }
      __syncthreads();
    }
    // ---- C = L^-1 S L^-H ---------------------------------------------------------------
    fwd_solve<R, NMAX>(n, L, ld, A, ld, sh);  // A = L^-1 S
    for (int j = wave; j < n; j += NW)
      for (int i = lane; i < n; i += 64) X[i + (size_t)j * ldx] = conj(A[j + (size_t)i * ld]);
    __syncthreads();
    fwd_solve<R, NMAX>(n, L, ld, X, ldx, sh);  // X = L^-1 (L^-1 S)^H = C
    for (int j = wave; j < n; j += NW)
      for (int i = lane; i < n; i += 64) {
        cplx<R> c = (i >= j) ? X[i + (size_t)j * ldx] : conj(X[j + (size_t)i * ldx]);
        if (i == j) c.im = 0;
        A[i + (size_t)j * ld] = c;
      }
    __syncthreads();
  }

  // ---- Householder tridiagonalisation (zhetd2, lower, fused her2 + hemv) --------------
  if (n == 1) {
    if (tid == 0) {
      sh.dd[0] = A[0].re;
      sh.ee[0] = 0;
    }
    __syncthreads();
  } else {
    cplx<R>* vcur = sh.v;
    cplx<R>* vnext = sh.vn;
    cplx<R> tau = reflector<R, NMAX>(n, 0, A, ld, vcur, sh);
    if (tid == 0) {
      sh.dd[0] = A[0].re;
      A[0] = tau;  // diagonal slot keeps tau_0 for the back-transform
    }
    trail_pass<R, NMAX, false>(n, 1, A, ld, vcur, vcur, sh);
    make_w<R, NMAX>(n, 0, tau, vcur, sh);
    for (int k = 0; k + 2 < n; ++k) {
      // pending update (v_k, w_k) applied to column k+1, rows [k+1, n)
      {
        const cplx<R> cw = conj(sh.w[k + 1]);
        const cplx<R> cv = conj(vcur[k + 1]);
        cplx<R>* col = A + (size_t)(k + 1) * ld;
        for (int i = k + 1 + tid; i < n; i += NT) col[i] = col[i] - (vcur[i] * cw + sh.w[i] * cv);
      }
      __syncthreads();
      if (tid == 0) sh.dd[k + 1] = A[(k + 1) + (size_t)(k + 1) * ld].re;
      tau = reflector<R, NMAX>(n, k + 1, A, ld, vnext, sh);
      if (tid == 0) A[(k + 1) + (size_t)(k + 1) * ld] = tau;
      trail_pass<R, NMAX, true>(n, k + 2, A, ld, vcur, vnext, sh);
      make_w<R, NMAX>(n, k + 1, tau, vnext, sh);
      cplx<R>* t = vcur;
      vcur = vnext;
      vnext = t;
    }
    if (tid == 0) {
      cplx<R> a = A[(n - 1) + (size_t)(n - 1) * ld];
      const cplx<R> vv = vcur[n - 1], ww = sh.w[n - 1];
      a = a - (vv * conj(ww) + ww * conj(vv));
      sh.dd[n - 1] = a.re;
      A[(n - 1) + (size_t)(n - 1) * ld] = mk<R>(0, 0);  // no reflector for the last column
      sh.ee[n - 1] = 0;
    }
    __syncthreads();
  }

  // ---- implicit QL on the real tridiagonal (tql2), rotations into Z -------------------
  for (int j = 0; j < n; ++j)
    for (int i = tid; i < n; i += NT) Z[i + (size_t)j * ld] = (i == j) ? R(1) : R(0);
  if (tid == 0) {
    sh.scal[1] = 0;  // f
    sh.scal[2] = 0;  // tst1
    sh.iscal[2] = 0; // failure flag
  }
  __syncthreads();
  const R eps = eps_r<R>();
  for (int l = 0; l < n; ++l) {
    if (tid == 0) {
      R tst1 = fmax(sh.scal[2], fabs(sh.dd[l]) + fabs(sh.ee[l]));
      sh.scal[2] = tst1;
      int m = l;
      while (m < n - 1 && fabs(sh.ee[m]) > eps * tst1) ++m;
      sh.iscal[1] = m;
    }
    __syncthreads();
    const int m = sh.iscal[1];
    if (m > l) {
      int iter = 0;
      for (;;) {
        if (tid == 0) {
          R* d = sh.dd;
          R* e = sh.ee;
          R g = d[l];
          R pp = (d[l + 1] - g) / (R(2) * e[l]);
          R r = sqrt(pp * pp + R(1));
          if (pp < 0) r = -r;
          d[l] = e[l] / (pp + r);
          d[l + 1] = e[l] * (pp + r);
          const R dl1 = d[l + 1];
          R h = g - d[l];
          for (int i = l + 2; i < n; ++i) d[i] -= h;
          sh.scal[1] += h;
          pp = d[m];
          R c = 1, c2 = 1, c3 = 1, s = 0, s2 = 0;
          const R el1 = e[l + 1];
          for (int i = m - 1; i >= l; --i) {
            c3 = c2;
            c2 = c;
            s2 = s;
            g = c * e[i];
            h = c * pp;
            // Givens rotation zeroing e[i]: one square root and one reciprocal
            const R ei = e[i];
            r = sqrt(pp * pp + ei * ei);
            e[i + 1] = s * r;
            if (r != R(0)) {
              const R rinv = R(1) / r;
              s = ei * rinv;
              c = pp * rinv;
            } else {
              s = R(0);
              c = R(1);
            }
            pp = c * d[i] - s * g;
            d[i + 1] = h + s * (c * g + s * d[i]);
            sh.cs[i] = c;
            sh.sn[i] = s;
          }
          pp = -s * s2 * c3 * el1 * e[l] / dl1;
          e[l] = s * pp;
          d[l] = c * pp;
          sh.iscal[3] = !(fabs(e[l]) > eps * sh.scal[2]);
        }
        __syncthreads();
        // apply rotations i = m-1 .. l to the rows of Z owned by this thread; columns are
        // fetched 16 at a time so that their loads are in flight together
        constexpr int CH = 16;
        for (int row = tid; row < n; row += NT) {
          R carry = Z[row + (size_t)m * ld];
          for (int i0 = m - 1; i0 >= l; i0 -= CH) {
            const int cnt = min(CH, i0 - l + 1);
            R zc[CH];
#pragma unroll
            for (int t = 0; t < CH; ++t)
              if (t < cnt) zc[t] = Z[row + (size_t)(i0 - t) * ld];
#pragma unroll
            for (int t = 0; t < CH; ++t) {
              if (t < cnt) {
                const int i = i0 - t;
                const R c = sh.cs[i], s = sh.sn[i];
                Z[row + (size_t)(i + 1) * ld] = s * zc[t] + c * carry;
                carry = c * zc[t] - s * carry;
              }
            }
          }
          Z[row + (size_t)l * ld] = carry;
        }
        const bool done = sh.iscal[3];
        __syncthreads();
        if (done) break;
        if (++iter > 30 * n) {
          if (tid == 0) sh.iscal[2] = 1;
          break;
        }
      }
    }
    if (tid == 0) {
      sh.dd[l] += sh.scal[1];
      sh.ee[l] = 0;
    }
    __syncthreads();
  }
  if (sh.iscal[2]) {
    for (int jj = tid; jj < nmax; jj += NT) dout[jj] = 0;
    if (tid == 0) p.info[sid] = -1;  // QL failed to converge
    __syncthreads();
    return;
  }

  // ---- ascending order: rank sort --------------------------------------------------
  for (int i = tid; i < n; i += NT) {
    const R di = sh.dd[i];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const R dj = sh.dd[j];
      rank += (dj < di) || (dj == di && j < i);
    }
    sh.perm[rank] = i;
  }
  __syncthreads();
  for (int j = tid; j < nmax; j += NT) dout[j] = j < n ? sh.dd[sh.perm[j]] : R(0);

  // ---- back-transform X = Q Z[:, perm] ------------------------------------------------
  for (int j = wave; j < n; j += NW) {
    const int pj = sh.perm[j];
    for (int i = lane; i < n; i += 64) X[i + (size_t)j * ldx] = mk<R>(Z[i + (size_t)pj * ld], 0);
  }
  __syncthreads();
  for (int c = n - 2; c >= 0; --c) {
    const cplx<R> tau = A[c + (size_t)c * ld];
    if (tau.re == R(0) && tau.im == R(0)) continue;  // uniform: every thread reads the same value
    for (int i = c + 1 + tid; i < n; i += NT)
      sh.v[i] = (i == c + 1) ? mk<R>(1, 0) : A[i + (size_t)c * ld];
    __syncthreads();
    for (int j = wave; j < n; j += NW) {
      cplx<R>* col = X + (size_t)j * ldx;
      R pr = 0, pi = 0;
      for (int i = c + 1 + lane; i < n; i += 64) {
        cplx<R> t = cmul_conj_a(sh.v[i], col[i]);
        pr += t.re;
        pi += t.im;
      }
      pr = wave_sum(pr);
      pi = wave_sum(pi);
      const cplx<R> t = tau * mk<R>(pr, pi);
      for (int i = c + 1 + lane; i < n; i += 64) col[i] = col[i] - sh.v[i] * t;
    }
    __syncthreads();
  }

  // ---- un-whiten: eigenvectors of the pencil = L^-H X ---------------------------------
  if (!p.standard) bwd_solve_h<R, NMAX>(n, L, ld, X, ldx, sh);
  if (tid == 0) p.info[sid] = 0;
  __syncthreads();
}

template <typename R, typename TIn, int NMAX>
__global__ __launch_bounds__(NT) void eig_kernel(EigParams<R> p) {
  __shared__ EigShared<R, NMAX> sh;
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
      const bool live = p.mode == 0 ? (v.re * v.re + v.im * v.im > eps)
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
      eig_kernel<R, double, NM><<<n_slots, NT, 0, stream>>>(p);                         \
    else                                                                                \
      eig_kernel<R, float, NM><<<n_slots, NT, 0, stream>>>(p);                          \
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
#define BIPP_EIG_FN(NM)                                                                      \
  if (b == NM)                                                                               \
    fn = in_double ? reinterpret_cast<const void*>(&eig_kernel<R, double, NM>)               \
                   : reinterpret_cast<const void*>(&eig_kernel<R, float, NM>);
  BIPP_EIG_FN(64)
  BIPP_EIG_FN(128)
  BIPP_EIG_FN(256)
  BIPP_EIG_FN(512)
#undef BIPP_EIG_FN
  if (!fn) return 1;
  BIPP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, NT, 0));
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
