// Cholesky factor G = L L^H of a Hermitian positive definite matrix and its inverse
// W = L^-1, computed by one workgroup: the whitening of the generalized problem in ?hegv
// (?potrf on G, then the transform by L^-1; src/host/lapack_api.hpp:112-158).
//
// Blocked so that the O(n^3) work runs as LDS-tiled products instead of column sweeps over
// the whole trailing matrix:
//   1. block columns of CI_CB: the diagonal block of the updated G is factored in LDS (one
//      barrier per column) and inverted there (one column per thread), its inverse W_JJ
//      stored in place of the block; the panel below becomes L_{>J,J} = G_{>J,J} W_JJ^H and
//      the trailing matrix G_{>J,>J} -= L_{>J,J} L_{>J,J}^H (lower tiles only);
//   2. the off-diagonal blocks of W bottom-up over pairs of blocks of size s = CI_CB, 2 CI_CB,
//      ...: inverse of [L11 0; L21 L22] is [W11 0; -W22 L21 W11 W22], with T = L21 W11 in
//      the scratch array and W21 = -W22 T written over L21.
// Only the lower triangle of L is read; tiles on the diagonal leave the strict upper
// triangle untouched.  Team: tid, sync() over NT threads, mark(k), which may time step k
// (0 diagonal factor, 1 diagonal inverse, 2 panel and trailing products, 3 off-diagonal
// blocks of W).
#pragma once

#ifndef BIPP_HD
#define BIPP_HD __host__ __device__
#endif
#ifndef BIPP_GLOBAL
#define BIPP_GLOBAL  // pointee qualifier of the matrices (the device build: global memory)
#endif

#include "cplx.hpp"

namespace bippamd {

// read and write a matrix element member by member (a struct in another address space has
// no copy constructor or assignment of its own)
template <typename R>
BIPP_HD cplx<R> ci_get(const BIPP_GLOBAL cplx<R>& x) {
  return mk<R>(x.re, x.im);
}
template <typename R>
BIPP_HD void ci_put(BIPP_GLOBAL cplx<R>& x, cplx<R> v) {
  x.re = v.re;
  x.im = v.im;
}

constexpr int CI_CB = 32;  // diagonal block edge
constexpr int CI_T = 64;   // product tile edge
constexpr int CI_K = 16;   // product k chunk

template <typename R>
struct CiShared {
  union {
    struct {
      cplx<R> d[CI_CB][CI_CB + 1];  // diagonal block, then its Cholesky factor
      cplx<R> w[CI_CB][CI_CB + 1];  // inverse of the factor
    } blk;
    struct {
      cplx<R> a[CI_K][CI_T + 1];  // a[k][r] = A[i0 + r, k0 + k]
      cplx<R> b[CI_K][CI_T + 1];  // b[k][c] = B[k0 + k, j0 + c]
    } gt;
  };
};

// Tile arithmetic of ci_gemm on the vector units: thread t accumulates tile row t mod CI_T
// and CPW consecutive columns of group t / CI_T.  A team may bring its own (Team::CiTile,
// the same members), e.g. one on matrix instructions.
template <typename R, int NT>
struct CiValuTile {
  static constexpr int CPW = CI_T * CI_T / NT;
  static_assert(CPW * (NT / CI_T) == CI_T, "tile splits over NT");
  cplx<R> acc[CPW];
  BIPP_HD void init() {
    for (int c = 0; c < CPW; ++c) acc[c] = mk<R>(0, 0);
  }
  // acc += the staged chunk's product (sh.gt.a: [k][row], sh.gt.b: [k][column])
  BIPP_HD void step(const CiShared<R>& sh, int tid) {
    const int row = tid % CI_T, cg = tid / CI_T;
    for (int k = 0; k < CI_K; ++k) {
      const cplx<R> a = sh.gt.a[k][row];
      for (int c = 0; c < CPW; ++c) acc[c] = acc[c] + a * sh.gt.b[k][cg * CPW + c];
    }
  }
  template <class ST>
  BIPP_HD void store(int i0, int j0, int m, int nn, ST st, int tid) const {
    const int i = i0 + tid % CI_T;
    for (int c = 0; c < CPW; ++c) {
      const int j = j0 + (tid / CI_T) * CPW + c;
      if (i < m && j < nn) st(i, j, acc[c]);
    }
  }
};

// C = A B over rows [0, m) and columns [0, nn) in CI_T x CI_T tiles (LOWER: only tiles with
// ti >= tj).  fa(i, k) and fb(k, j) give the operands (zero outside their pattern);
// krange(i0, j0, klo, khi) bounds the k of a tile; store(i, j, acc) writes a result.
// B_KFAST: consecutive k of B are contiguous in memory (else consecutive j), which sets how
// the threads split a tile load.  The arithmetic is the team's tile type (Team::CiTile);
// the next k chunk is loaded into registers while the current one is used.  Ends with a
// barrier (the stores are visible to the next product).
template <typename R, int NT, bool LOWER, bool B_KFAST, class Team, class FA, class FB, class KR,
          class ST>
BIPP_HD void ci_gemm(Team& tm, int m, int nn, FA fa, FB fb, KR krange, ST store,
                     CiShared<R>& sh) {
  constexpr int PER = CI_T * CI_K / NT;
  static_assert(PER * NT == CI_T * CI_K, "tile loads split over NT");
  const int tid = tm.tid;
  const int ntr = (m + CI_T - 1) / CI_T, ntc = (nn + CI_T - 1) / CI_T;
  for (int tj = 0; tj < ntc; ++tj) {
    for (int ti = LOWER ? tj : 0; ti < ntr; ++ti) {
      const int i0 = ti * CI_T, j0 = tj * CI_T;
      int klo = 0, khi = 0;
      krange(i0, j0, klo, khi);
      typename Team::template CiTile<R, NT> acc;
      acc.init();
      cplx<R> ra[PER], rb[PER];
      auto fetch = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT;
          const int ka = e / CI_T, ia = i0 + e % CI_T;
          ra[q] = (ia < m && k0 + ka < khi) ? fa(ia, k0 + ka) : mk<R>(0, 0);
          const int kb = B_KFAST ? e % CI_K : e / CI_T;
          const int jb = j0 + (B_KFAST ? e / CI_K : e % CI_T);
          rb[q] = (jb < nn && k0 + kb < khi) ? fb(k0 + kb, jb) : mk<R>(0, 0);
        }
      };
      if (klo < khi) fetch(klo);
      for (int k0 = klo; k0 < khi; k0 += CI_K) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
          const int e = tid + q * NT;
          sh.gt.a[e / CI_T][e % CI_T] = ra[q];
          if (B_KFAST)
            sh.gt.b[e % CI_K][e / CI_K] = rb[q];
          else
            sh.gt.b[e / CI_T][e % CI_T] = rb[q];
        }
        tm.sync();
        if (k0 + CI_K < khi) fetch(k0 + CI_K);
        acc.step(sh, tid);
        tm.sync();
      }
      acc.store(i0, j0, m, nn, store, tid);
    }
  }
  tm.sync();
}

// G (lower triangle of L, leading dimension ld) -> W = L^-1 in place.  T: scratch of at
// least ldt x n complex (ldt >= n).  Returns -1, or the 0-based column j whose pivot is
// not positive (the leading minor of order j + 1 is not positive definite; L is then left
// partly transformed).
template <typename R, int NT, class Team>
BIPP_HD int chol_inverse(Team& tm, int n, BIPP_GLOBAL cplx<R>* L, int ld, BIPP_GLOBAL cplx<R>* T, int ldt,
                         CiShared<R>& sh) {
  const int tid = tm.tid;
  auto at = [=](int i, int j) -> BIPP_GLOBAL cplx<R>& { return L[i + (long long)j * ld]; };

  // ---- 1. blocked right-looking factorisation, diagonal blocks inverted -----------------
  tm.mark(-1);
  for (int j0 = 0; j0 < n; j0 += CI_CB) {
    const int jb = n - j0 < CI_CB ? n - j0 : CI_CB;
    for (int e = tid; e < CI_CB * CI_CB; e += NT) {
      const int i = e % CI_CB, j = e / CI_CB;
      sh.blk.d[i][j] = (i < jb && i >= j) ? ci_get(at(j0 + i, j0 + j)) : mk<R>(0, 0);
    }
    tm.sync();
    // step k: pivot d_kk (final after step k - 1), trailing columns updated with the
    // unscaled column k, column k - 1 scaled (nothing reads it in this step)
    R p_prev = R(1);
    int fail = -1;
    for (int k = 0; k < jb; ++k) {
      const R p = sh.blk.d[k][k].re;
      if (!(p > R(0) && p == p)) {
        fail = k;
        break;
      }
      const R ip = R(1) / p;
      for (int e = tid; e < CI_CB * CI_CB; e += NT) {
        const int i = e % CI_CB, j = e / CI_CB;
        if (j > k && i >= j && i < jb)
          sh.blk.d[i][j] = sh.blk.d[i][j] - sh.blk.d[i][k] * conj(sh.blk.d[j][k]) * ip;
      }
      if (k > 0) {
        const R s = sqrt(p_prev), is = R(1) / s;
        for (int i = k - 1 + tid; i < jb; i += NT)
          sh.blk.d[i][k - 1] = i == k - 1 ? mk<R>(s, R(0)) : sh.blk.d[i][k - 1] * is;
      }
      p_prev = p;
      tm.sync();
    }
    if (fail >= 0) return j0 + fail;
    {
      const int k = jb - 1;
      const R s = sqrt(p_prev), is = R(1) / s;
      for (int i = k + tid; i < jb; i += NT)
        sh.blk.d[i][k] = i == k ? mk<R>(s, R(0)) : sh.blk.d[i][k] * is;
    }
    tm.sync();
    tm.mark(0);
    // inverse W = L^-1, right-looking over the whole team: a[i][j] = delta_ij - sum_{m < i}
    // L_im W_mj with W_mj = a[m][j] / L_mm, accumulated one pivot m per step (the elements
    // of a step spread over the team); then W_ij = a[i][j] / L_ii
    for (int e = tid; e < CI_CB * CI_CB; e += NT) {
      const int i = e % CI_CB, j = e / CI_CB;
      if (i < jb && i >= j) sh.blk.w[i][j] = mk<R>(i == j ? R(1) : R(0), R(0));
    }
    tm.sync();
    for (int m = 0; m + 1 < jb; ++m) {
      const R im = R(1) / sh.blk.d[m][m].re;
      for (int e = tid; e < CI_CB * CI_CB; e += NT) {
        const int i = e % CI_CB, j = e / CI_CB;
        if (i > m && i < jb && j <= m)
          sh.blk.w[i][j] = sh.blk.w[i][j] - sh.blk.d[i][m] * (sh.blk.w[m][j] * im);
      }
      tm.sync();
    }
    for (int e = tid; e < CI_CB * CI_CB; e += NT) {
      const int i = e % CI_CB, j = e / CI_CB;
      if (i < jb && i >= j) sh.blk.w[i][j] = sh.blk.w[i][j] * (R(1) / sh.blk.d[i][i].re);
    }
    tm.sync();
    for (int e = tid; e < CI_CB * CI_CB; e += NT) {
      const int i = e % CI_CB, j = e / CI_CB;
      if (i < jb && i >= j) ci_put(at(j0 + i, j0 + j), sh.blk.w[i][j]);
    }
    tm.sync();
    tm.mark(1);
    const int mp = n - j0 - jb;
    if (mp > 0) {
      const int r0 = j0 + jb;
      // L_{>J,J} = G_{>J,J} W_JJ^H in place (a tile reads its rows' k < jb before it stores)
      ci_gemm<R, NT, false, false>(
          tm, mp, jb, [=](int i, int k) { return ci_get(L[(r0 + i) + (long long)(j0 + k) * ld]); },
          [=](int k, int j) {
            return k <= j ? conj(ci_get(L[(j0 + j) + (long long)(j0 + k) * ld])) : mk<R>(0, 0);
          },
          [=](int, int, int& klo, int& khi) {
            klo = 0;
            khi = jb;
          },
          [=](int i, int j, cplx<R> v) { ci_put(L[(r0 + i) + (long long)(j0 + j) * ld], v); }, sh);
      // G_{>J,>J} -= L_{>J,J} L_{>J,J}^H on and below the diagonal
      ci_gemm<R, NT, true, false>(
          tm, mp, mp, [=](int i, int k) { return ci_get(L[(r0 + i) + (long long)(j0 + k) * ld]); },
          [=](int k, int j) { return conj(ci_get(L[(r0 + j) + (long long)(j0 + k) * ld])); },
          [=](int, int, int& klo, int& khi) {
            klo = 0;
            khi = jb;
          },
          [=](int i, int j, cplx<R> v) {
            if (i >= j) {
              BIPP_GLOBAL cplx<R>& c = L[(r0 + i) + (long long)(r0 + j) * ld];
              ci_put(c, ci_get(c) - v);
            }
          },
          sh);
    }
    tm.mark(2);
  }

  // ---- 2. off-diagonal blocks of W, bottom-up ------------------------------------------
  for (int s = CI_CB; s < n; s *= 2) {
    for (int b0 = 0; b0 + s < n; b0 += 2 * s) {
      const int s2 = n - b0 - s < s ? n - b0 - s : s;
      const int b1 = b0 + s;
      // T = L21 W11 (W11 lower: k >= j)
      ci_gemm<R, NT, false, true>(
          tm, s2, s, [=](int i, int k) { return ci_get(L[(b1 + i) + (long long)(b0 + k) * ld]); },
          [=](int k, int j) {
            return k >= j ? ci_get(L[(b0 + k) + (long long)(b0 + j) * ld]) : mk<R>(0, 0);
          },
          [=](int, int jt, int& klo, int& khi) {
            klo = jt;
            khi = s;
          },
          [=](int i, int j, cplx<R> v) { ci_put(T[i + (long long)j * ldt], v); }, sh);
      // W21 = -W22 T (W22 lower: k <= i), over L21
      ci_gemm<R, NT, false, true>(
          tm, s2, s,
          [=](int i, int k) {
            return k <= i ? ci_get(L[(b1 + i) + (long long)(b1 + k) * ld]) : mk<R>(0, 0);
          },
          [=](int k, int j) { return ci_get(T[k + (long long)j * ldt]); },
          [=](int it, int, int& klo, int& khi) {
            klo = 0;
            khi = it + CI_T < s2 ? it + CI_T : s2;
          },
          [=](int i, int j, cplx<R> v) {
            ci_put(L[(b1 + i) + (long long)(b0 + j) * ld], mk<R>(-v.re, -v.im));
          },
          sh);
    }
  }
  tm.mark(3);
  return -1;
}

}  // namespace bippamd
