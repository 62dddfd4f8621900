// Eigenvalues and eigenvectors of a real symmetric tridiagonal matrix by divide and conquer,
// computed by one workgroup (the eigensolver's tridiagonal phase; LAPACK's ?stedc is the
// method ?heevd uses for the same step, src/host/lapack_api.hpp:112-158 calls ?hegv, whose
// QL iteration gives the same eigenpairs to rounding).
//
// Bottom-up over levels l = 1 .. L (L = ceil(log2 n)): level l merges the blocks
// [m 2^l, m 2^l + 2^(l-1)) and [m 2^l + 2^(l-1), min((m + 1) 2^l, n)) of every block m; the
// leaves are the 1 x 1 blocks.  Splitting T at off-diagonal beta (rows k, k+1) leaves
// diag(T1, T2) with beta taken off d_k and d_{k+1}, plus rho v v^T, rho = |beta|,
// v = e_k + sign(beta) e_{k+1}; so with T1 = Q1 D1 Q1^T and T2 = Q2 D2 Q2^T,
//   T = diag(Q1, Q2) (D + rho z z^T) diag(Q1, Q2)^T,  z = [last row of Q1, sign(beta) first
//   row of Q2].
// Per merge (Cuppen; deflation and the eigenvector formula of Gu and Eisenstat, the steps
// LAPACK's dlaed2 / dlaed3 / dlaed4 take):
//   1. z scaled to unit norm (rho doubled), the 2 x sz poles merged into ascending order;
//   2. deflation in that order: a pole with rho |z_i| <= tol keeps (d_i, e_i); two poles
//      with |(d_j - d_i) c s| <= tol are rotated so that one z vanishes;
//      tol = 8 eps max(max |d|, max |z|);
//   3. the K remaining poles: root k of 1/rho + sum z_i^2 / (d_i - lambda) = 0 per lane;
//   4. z recomputed from the roots, z_i^2 = -(d_i - lambda_i) prod_{j != i} (d_i - lambda_j)
//      / (d_i - d_j), so the vectors u_k = (z_i / (d_i - lambda_k))_i / norm are orthogonal
//      to working precision however close the roots are;
//   5. the sz eigenvalues of the merge sorted (ties by slot); M (sz x sz, rows = sorted
//      slots, columns = output order) holds the vectors of the rank-one problem in slot
//      space, the deflation rotations applied to its rows in reverse order;
//   6. Q_new = diag(Q1, Q2) M (rows of M gathered back to the children's column order),
//      two half-size products.
// The matrix is scaled to max |entry| = 1 first and the eigenvalues scaled back at the end.
//
// The code runs on a "team" (tid, sync(), mark(k), which may time step k of a level, and
// lane_sum<P>(v), the total of v over the aligned group of P consecutive threads, the same
// bits in each): the eigensolver's workgroup on the GPU.  Every loop over a team's share is
// followed by a sync before another thread reads its results.
#pragma once

#ifndef BIPP_HD
#define BIPP_HD __host__ __device__
#endif
#ifndef BIPP_GLOBAL
#define BIPP_GLOBAL  // pointee qualifier of the work arrays (the device build: global memory)
#endif

#include <limits>

namespace bippamd {

constexpr int DC_TK = 16;  // k chunk of the merge products

template <int NMAX>
struct DcCfg {
  static constexpr int TR = NMAX / 4 < 64 ? NMAX / 4 : 64;  // product tile edge
  static constexpr int TK = DC_TK;                          // product k chunk
  static constexpr int TILE_ELEMS = 2 * TR * TK;            // LDS reals of the tiles
  static constexpr int MAXIT = 100;                         // secular iteration cap
};

template <typename R, int NMAX>
struct DcShared {
  R ds[NMAX];    // per sorted slot: pole (updated by deflation rotations), then item value
  R zs[NMAX];    // per sorted slot: z
  R dk[NMAX];    // per merge at r0 + k: the K non-deflated poles, ascending
  R zk[NMAX];    //                      and their z
  R tau[NMAX];   // root k at r0 + k: lambda_k = dk[org_k] + tau_k
  R zh[NMAX];    // recomputed z
  R nrm[NMAX];   // 1 / |u_k|
  R rc[NMAX];    // deflation rotations at r0 + t: (c, s) on local slots (rp, rn)
  R rs[NMAX];
  int kind[NMAX];     // per slot: -1 deflated, else its compact index k
  int slot_of[NMAX];  // per position (child column order): its sorted slot
  int org[NMAX];      // per root: origin pole (compact index)
  int rp[NMAX];
  int rn[NMAX];
  int itemcol[NMAX];  // per output column r0 + c: the local slot of the item placed there
  int mK[NMAX / 2 + 1];  // per merge: non-deflated count, rotation count
  int mR[NMAX / 2 + 1];
  R scale;
};

template <typename R>
BIPP_HD R dc_abs(R x) {
  return x < R(0) ? -x : x;
}

template <typename R>
BIPP_HD R dc_key(R x) {  // NaN ordered as +inf (the sort keys stay a strict order)
  return x == x ? x : std::numeric_limits<R>::infinity();
}

// d_i - lambda_j from the poles and the root's (origin, tau): the same rounding wherever used
template <typename R>
BIPP_HD R dc_delta(const R* dk, const int* org, const R* tau, int i, int j) {
  return (dk[i] - dk[org[j]]) - tau[j];
}

// Roots eta of c eta^2 - a eta + b = 0 (each in the form without cancellation): tn = x + eta
// for the root with x + eta inside (lo, hi) nearest x; false if neither is.
template <typename R>
BIPP_HD bool dc_quad(R c, R a, R b, R x, R lo, R hi, R& tn) {
  bool have = false;
  auto take = [&](R eta) {
    const R t = x + eta;
    if (t > lo && t < hi && (!have || dc_abs(t - x) < dc_abs(tn - x))) {
      tn = t;
      have = true;
    }
  };
  if (c == R(0)) {
    if (a != R(0)) take(b / a);
  } else {
    const R disc = a * a - R(4) * b * c;
    if (disc >= R(0)) {
      const R sq = sqrt(disc);
      const R r1 = a <= R(0) ? (a - sq) / (R(2) * c) : R(2) * b / (a + sq);
      take(r1);
      if (r1 != R(0)) take(b / (c * r1));
    }
  }
  return have;
}

// Root k of 1/rho + sum_i zk_i^2 / (dk_i - lambda) = 0, dk strictly ascending, zk nonzero,
// rho > 0: lambda_k lies in (dk_k, dk_{k+1}), or in (dk_{K-1}, dk_{K-1} + rho |zk|^2] for
// the last one.  Returns the nearer pole o and tau = lambda - dk_o.  The function increases
// with lambda between poles; its sign at the interval's midpoint picks the half (and o), and
// tau is kept in a bracket narrowed by the sign at every iterate.  Start: the root of the
// two-pole model (poles k and k+1, or K-2 and K-1 for the last root) whose other terms are
// taken at the midpoint (R.-C. Li's initial guess, the one LAPACK's dlaed4 makes).  Step:
// the root of the two-pole model with each pole's partial sum fitted to its value and slope
// and a constant fitted to the function (Li's "middle way"), else a Newton step, else
// bisection, whichever first stays inside the bracket.
// The P lanes of a root (part = 0 .. P - 1) each sum the terms i = part, part + P, ...;
// sum(v) totals a value over the P lanes with the same bits in each, so every lane takes the
// same steps and returns the same root.
template <typename R, int MAXIT, int P, class SUM>
BIPP_HD void dc_secular(int K, int k, int part, const R* dk, const R* zk, R rho, SUM sum,
                        int& org_out, R& tau_out) {
  const R eps = std::numeric_limits<R>::epsilon();
  if (K == 1) {
    org_out = 0;
    tau_out = rho * zk[0] * zk[0];
    return;
  }
  const R rinv = R(1) / rho;
  const int lower = k < K - 1 ? k : K - 2;  // the model's poles: lower, lower + 1
  R base, mid, lo, hi;                      // the midpoint: dk_base + mid
  if (k < K - 1) {
    base = dk[k];
    mid = R(0.5) * (dk[k + 1] - dk[k]);
  } else {
    R s = 0;
    for (int i = part; i < K; i += P) s += zk[i] * zk[i];
    hi = rho * sum(s) * (R(1) + R(4) * eps);
    base = dk[K - 1];
    mid = R(0.5) * hi;
  }
  R f = 0;
  for (int i = part; i < K; i += P) f += zk[i] * zk[i] / ((dk[i] - base) - mid);
  f = rinv + sum(f);
  int o;
  if (k < K - 1) {
    if (f >= R(0)) {
      o = k;
      lo = R(0);
      hi = mid;
    } else {
      o = k + 1;
      lo = -mid;
      hi = R(0);
    }
  } else {
    o = K - 1;
    if (f >= R(0)) {
      lo = R(0);
      hi = mid;
    } else {
      lo = mid;
    }
  }
  const R dorg = dk[o];
  R tau;
  {
    const R z1 = zk[lower] * zk[lower], z2 = zk[lower + 1] * zk[lower + 1];
    const R p1 = dk[lower] - dorg, p2 = dk[lower + 1] - dorg;
    // f minus the two poles' terms, as a constant: C + z1/(p1 - t) + z2/(p2 - t) = 0
    const R cm = f - z1 / ((dk[lower] - base) - mid) - z2 / ((dk[lower + 1] - base) - mid);
    if (!dc_quad(cm, cm * (p1 + p2) + z1 + z2, cm * p1 * p2 + z1 * p2 + z2 * p1, R(0), lo, hi,
                 tau))
      tau = R(0.5) * (lo + hi);
  }
  for (int it = 0; it < MAXIT; ++it) {
    // psi: the poles up to lower (terms <= 0), phi: the rest (terms >= 0)
    R psi = 0, dpsi = 0, phi = 0, dphi = 0;
    int i = part;
#pragma unroll 4
    for (; i <= lower; i += P) {
      const R rdel = R(1) / ((dk[i] - dorg) - tau);  // one division per term
      const R t = zk[i] * zk[i] * rdel;
      psi += t;
      dpsi += t * rdel;
    }
#pragma unroll 4
    for (; i < K; i += P) {
      const R rdel = R(1) / ((dk[i] - dorg) - tau);
      const R t = zk[i] * zk[i] * rdel;
      phi += t;
      dphi += t * rdel;
    }
    psi = sum(psi);
    dpsi = sum(dpsi);
    phi = sum(phi);
    dphi = sum(dphi);
    const R w = rinv + psi + phi;
    const R dw = dpsi + dphi;
    if (w < R(0))
      lo = tau;
    else
      hi = tau;
    // |w| within the rounding of its evaluation (phi - psi: the sum of |terms|): done
    if (dc_abs(w) <= R(8) * eps * (rinv + phi - psi + dc_abs(tau) * dw)) break;
    const R dl = (dk[lower] - dorg) - tau;
    const R du = (dk[lower + 1] - dorg) - tau;
    // model c eta^2 - a eta + b = 0 (eta: step of tau)
    R tn;
    bool have = dc_quad(w - dl * dpsi - du * dphi, (dl + du) * w - dl * du * dw, dl * du * w,
                        tau, lo, hi, tn);
    if (!have && dw > R(0)) {
      const R t = tau - w / dw;
      if (t > lo && t < hi) {
        tn = t;
        have = true;
      }
    }
    if (!have) tn = R(0.5) * (lo + hi);
    if (!(tn > lo && tn < hi) || tn == tau) break;  // bracket at its resolution
    tau = tn;
  }
  org_out = o;
  tau_out = tau;
}

// Number of keys in sorted d[b .. e) below x (strict) or not above x (!strict).
template <typename R>
BIPP_HD int dc_count(const R* d, int st, int b, int e, R x, bool strict) {
  int lo = b, hi = e;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const R v = dc_key(d[(long long)mid * st]);
    if (strict ? (v < x) : (v <= x))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo - b;
}

// Tile arithmetic of dc_product_tiles on the vector units: each thread accumulates an
// MR x MC block (row group tid mod RG, column group tid / RG).  A team may bring its own
// (Team::DcTile, the same members), e.g. one on matrix instructions.
template <typename R, int TR, int TK, int NT>
struct DcValuTile {
  static constexpr int MR = TR * TR / NT < 4 ? TR * TR / NT : 4;
  static constexpr int MC = TR * TR / (NT * MR);
  static constexpr int RG = TR / MR;  // row groups (fastest over threads)
  static_assert(MR >= 1 && MC >= 1 && RG * (TR / MC) == NT, "tile splits over the team");
  R acc[MR][MC];
  BIPP_HD void init() {
    for (int i = 0; i < MR; ++i)
      for (int j = 0; j < MC; ++j) acc[i][j] = R(0);
  }
  // acc += the staged chunk's product (tA, tB: [TK][TR], k-major)
  BIPP_HD void step(const R* tA, const R* tB, int tid) {
    const int rg = tid % RG, cg = tid / RG;
    for (int kk = 0; kk < TK; ++kk) {
      R a[MR], b[MC];
      for (int i = 0; i < MR; ++i) a[i] = tA[kk * TR + rg * MR + i];
      for (int j = 0; j < MC; ++j) b[j] = tB[kk * TR + cg * MC + j];
      for (int i = 0; i < MR; ++i)
        for (int j = 0; j < MC; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
  }
  // st(row, col, value) for the tile's entries (rows and columns relative to the tile)
  template <class ST>
  BIPP_HD void store(ST st, int tid) const {
    const int rg = tid % RG, cg = tid / RG;
    for (int j = 0; j < MC; ++j)
      for (int i = 0; i < MR; ++i) st(rg * MR + i, cg * MC + j, acc[i][j]);
  }
};

// Q_new = diag(Q1, Q2) M for one merge, in TR x TR tiles (rows of a tile lie in one half:
// n1 is a multiple of TR), k in chunks of TK staged through LDS (2 TR TK reals of tile), the
// next chunk's loads in flight while the current one is used; the arithmetic is the team's
// tile type (Team::DcTile).
template <typename R, int TR, int NT, class Team>
BIPP_HD void dc_product_tiles(Team& tm, int r0, int sz, int n1, const BIPP_GLOBAL R* Bp, int ldp,
                              BIPP_GLOBAL R* Bn, int ldn, const BIPP_GLOBAL R* Mb,
                              const int* slot_of, R* tile) {
  constexpr int TK = DC_TK;
  constexpr int LPT = (TK * TR + NT - 1) / NT;   // staged elements per thread and operand
  R* tA = tile;            // [TK][TR]: A[ra + r, kc + kk]
  R* tB = tile + TK * TR;  // [TK][TR]: M'[kc + kk, ca + c]
  const int tid = tm.tid;
  const int ntr = (sz + TR - 1) / TR;
  for (int t = 0; t < ntr * ntr; ++t) {
    const int ra = (t / ntr) * TR, ca = (t % ntr) * TR;
    const int k0 = ra < n1 ? 0 : n1, k1 = ra < n1 ? n1 : sz;
    R pa[LPT], pb[LPT];
    auto load = [&](int kc) {
      for (int u = 0; u < LPT; ++u) {
        const int e = tid + u * NT;
        const int kk = e / TR, rr = e % TR;
        const int k = kc + kk;
        const bool in = e < TK * TR && k < k1;
        pa[u] = (in && ra + rr < sz) ? Bp[(r0 + ra + rr) + (long long)(r0 + k) * ldp] : R(0);
        pb[u] = (in && ca + rr < sz)
                    ? Mb[(long long)(slot_of[r0 + k] - r0) * sz + ca + rr]
                    : R(0);
      }
    };
    typename Team::template DcTile<R, TR, TK, NT> acc;
    acc.init();
    load(k0);
    for (int kc = k0; kc < k1; kc += TK) {
      tm.sync();  // the previous chunk is no longer read
      for (int u = 0; u < LPT; ++u) {
        const int e = tid + u * NT;
        if (e < TK * TR) {
          tA[e] = pa[u];
          tB[e] = pb[u];
        }
      }
      tm.sync();
      if (kc + TK < k1) load(kc + TK);
      acc.step(tA, tB, tid);
    }
    acc.store(
        [&](int i, int j, R v) {
          const int row = ra + i, col = ca + j;
          if (row < sz && col < sz) Bn[(r0 + row) + (long long)(r0 + col) * ldn] = v;
        },
        tid);
  }
  tm.sync();
}

// Eigen-decomposition of the symmetric tridiagonal (d_i = d[i st], e_i = e[i st], i < n - 1):
// on return d holds the eigenvalues ascending and column j of Z (leading dimension ldz) the
// eigenvector of d_j.  W (n x n, leading dimension n) and M (n^2 reals) are scratch memory;
// tile holds DcCfg<NMAX>::TILE_ELEMS reals of fast memory.  Requires 1 <= n <= NMAX and a
// team of NT threads.
template <typename R, int NMAX, int NT, class Team>
BIPP_HD void dc_tridiag(Team& tm, int n, R* d, R* e, int st, BIPP_GLOBAL R* Z, int ldz,
                        BIPP_GLOBAL R* W, BIPP_GLOBAL R* M, DcShared<R, NMAX>& sh, R* tile) {
  constexpr int TR = DcCfg<NMAX>::TR;
  constexpr int MAXIT = DcCfg<NMAX>::MAXIT;
  // lanes per secular root: a level has at most n <= NMAX roots
  constexpr int RL = NT >= 4 * NMAX ? 4 : NT >= 2 * NMAX ? 2 : 1;
  const int tid = tm.tid;
  const R eps = std::numeric_limits<R>::epsilon();
  const R rsqrt2 = R(0.70710678118654752440);
  int L = 0;
  while ((1 << L) < n) ++L;
  // level l's blocks live in Z when L - l is even, else in W
  auto buf = [&](int l) { return ((L - l) & 1) == 0 ? Z : W; };
  auto bld = [&](int l) { return ((L - l) & 1) == 0 ? ldz : n; };

  // ---- scale to max |entry| = 1, modified diagonal, leaves -----------------------------
  if (tid == 0) {
    R s = 0;
    for (int i = 0; i < n; ++i) {
      const R a = dc_abs(d[(long long)i * st]);
      s = a > s ? a : s;
      if (i + 1 < n) {
        const R b = dc_abs(e[(long long)i * st]);
        s = b > s ? b : s;
      }
    }
    sh.scale = (s > R(0) && s <= std::numeric_limits<R>::max()) ? s : R(1);
  }
  tm.sync();
  const R scale = sh.scale;
  const R inv_scale = R(1) / scale;
  for (int i = tid; i < n; i += NT) {
    d[(long long)i * st] *= inv_scale;
    if (i + 1 < n) e[(long long)i * st] *= inv_scale;
  }
  tm.sync();
  {
    BIPP_GLOBAL R* B0 = buf(0);
    const int ld0 = bld(0);
    for (int i = tid; i < n; i += NT) {
      R v = d[(long long)i * st];
      if (i > 0) v -= dc_abs(e[(long long)(i - 1) * st]);
      if (i + 1 < n) v -= dc_abs(e[(long long)i * st]);
      sh.ds[i] = v;
      B0[i + (long long)i * ld0] = R(1);
    }
    tm.sync();
    for (int i = tid; i < n; i += NT) d[(long long)i * st] = sh.ds[i];
    tm.sync();
  }

  tm.mark(-1);
  for (int l = 1; l <= L; ++l) {
    const int h = 1 << (l - 1), S = 1 << l;
    const int nm = (n + S - 1) >> l;
    const BIPP_GLOBAL R* Bp = buf(l - 1);
    const int ldp = bld(l - 1);
    BIPP_GLOBAL R* Bn = buf(l);
    const int ldn = bld(l);

    // ---- 1. z and the merged order of the poles ----------------------------------------
    for (int i = tid; i < n; i += NT) {
      const int r0 = (i >> l) << l;
      if (r0 + h >= n) continue;  // a block without a sibling at this level
      const int sz = n - r0 < S ? n - r0 : S;
      const int q = i - r0;
      const R beta = e[(long long)(r0 + h - 1) * st];
      R zq = q < h ? Bp[(r0 + h - 1) + (long long)i * ldp] : Bp[(r0 + h) + (long long)i * ldp];
      if (q >= h && beta < R(0)) zq = -zq;
      const R di = d[(long long)i * st];
      const R key = dc_key(di);
      const int rank = q < h ? q + dc_count(d, st, r0 + h, r0 + sz, key, true)
                             : (q - h) + dc_count(d, st, r0, r0 + h, key, false);
      const int slot = r0 + rank;
      sh.ds[slot] = di;
      sh.zs[slot] = zq * rsqrt2;
      sh.slot_of[i] = slot;
    }
    tm.sync();
    tm.mark(0);

    // ---- 2. deflation, one thread per merge --------------------------------------------
    for (int m = tid; m < nm; m += NT) {
      const int r0 = m << l;
      if (r0 + h >= n) continue;
      const int sz = n - r0 < S ? n - r0 : S;
      const R rho = R(2) * dc_abs(e[(long long)(r0 + h - 1) * st]);
      R dmax = 0, zmax = 0;
      for (int p = r0; p < r0 + sz; ++p) {
        const R a = dc_abs(sh.ds[p]), b = dc_abs(sh.zs[p]);
        dmax = a > dmax ? a : dmax;
        zmax = b > zmax ? b : zmax;
      }
      const R tol = R(8) * eps * (dmax > zmax ? dmax : zmax);
      // the pending pole pj's (d, z) stay in registers; the rotation test |t c s| <= tol
      // is taken as |t zp zn| <= tol (zp^2 + zn^2) (c s = -zn zp / th^2), so the square root
      // and the divisions are only evaluated for a rotation
      int K = 0, nr = 0, pj = -1;
      R dj = 0, zj = 0;
      for (int p = r0; p < r0 + sz; ++p) {
        const R zn = sh.zs[p], dn = sh.ds[p];
        if (rho * dc_abs(zn) <= tol) {
          sh.kind[p] = -1;
          continue;
        }
        if (pj < 0) {
          pj = p;
          dj = dn;
          zj = zn;
          continue;
        }
        const R t = dn - dj;
        const R th2 = zj * zj + zn * zn;
        if (dc_abs(t * zj * zn) <= tol * th2) {
          // rotate slot pj's z into slot p: pj keeps an eigenvalue of its own
          const R th = sqrt(th2);
          const R c = zn / th, s = -zj / th;
          sh.zs[pj] = R(0);
          sh.rp[r0 + nr] = pj - r0;
          sh.rn[r0 + nr] = p - r0;
          sh.rc[r0 + nr] = c;
          sh.rs[r0 + nr] = s;
          ++nr;
          sh.ds[pj] = dj * c * c + dn * s * s;
          sh.kind[pj] = -1;
          dj = dj * s * s + dn * c * c;  // slot p's pole and z after the rotation
          zj = th;
          sh.ds[p] = dj;
          sh.zs[p] = zj;
        } else {
          sh.kind[pj] = K;
          sh.dk[r0 + K] = dj;
          sh.zk[r0 + K] = zj;
          ++K;
          dj = dn;
          zj = zn;
        }
        pj = p;
      }
      if (pj >= 0) {
        sh.kind[pj] = K;
        sh.dk[r0 + K] = dj;
        sh.zk[r0 + K] = zj;
        ++K;
      }
      sh.mK[m] = K;
      sh.mR[m] = nr;
    }
    tm.sync();
    tm.mark(1);

    // ---- 3. roots, RL lanes each -------------------------------------------------------
    for (int j = tid / RL; j < n; j += NT / RL) {
      const int m = j >> l, r0 = m << l;
      if (r0 + h >= n) continue;
      const int K = sh.mK[m], k = j - r0;
      if (k >= K) continue;
      const R rho = R(2) * dc_abs(e[(long long)(r0 + h - 1) * st]);
      int o;
      R t;
      dc_secular<R, MAXIT, RL>(K, k, tid % RL, sh.dk + r0, sh.zk + r0, rho,
                               [&](R v) { return tm.template lane_sum<RL>(v); }, o, t);
      if (tid % RL == 0) {
        sh.org[j] = o;
        sh.tau[j] = t;
      }
    }
    tm.sync();
    tm.mark(2);

    // ---- 4. recomputed z ------------------------------------------------------------------
    for (int i = tid; i < n; i += NT) {
      const int m = i >> l, r0 = m << l;
      if (r0 + h >= n) continue;
      const int K = sh.mK[m], k = i - r0;
      if (k >= K) continue;
      const R* dk = sh.dk + r0;
      const int* org = sh.org + r0;
      const R* tau = sh.tau + r0;
      R prod = dc_delta(dk, org, tau, k, k);
      for (int j = 0; j < K; ++j)
        if (j != k) prod *= dc_delta(dk, org, tau, k, j) / (dk[k] - dk[j]);
      sh.zh[i] = copysign(sqrt(dc_abs(prod)), sh.zk[i]);
    }
    tm.sync();
    tm.mark(3);

    // ---- 5a. vector norms; item values per slot ----------------------------------------
    for (int j = tid; j < n; j += NT) {
      const int m = j >> l, r0 = m << l;
      if (r0 + h >= n) continue;
      const int K = sh.mK[m], k = j - r0;
      if (k >= K) continue;
      const R* dk = sh.dk + r0;
      const int* org = sh.org + r0;
      const R* tau = sh.tau + r0;
      R s = 0;
      for (int i = 0; i < K; ++i) {
        const R u = sh.zh[r0 + i] / dc_delta(dk, org, tau, i, k);
        s += u * u;
      }
      sh.nrm[j] = R(1) / sqrt(s);
    }
    for (int p = tid; p < n; p += NT) {
      const int r0 = (p >> l) << l;
      if (r0 + h >= n) continue;
      const int k = sh.kind[p];
      // zs is free after the deflation: it now holds the value of the slot's item
      sh.zs[p] = k < 0 ? sh.ds[p] : sh.dk[r0 + sh.org[r0 + k]] + sh.tau[r0 + k];
    }
    tm.sync();
    tm.mark(4);

    // ---- 5b. output order (rank sort of the items, ties by slot) -----------------------
    for (int p = tid; p < n; p += NT) {
      const int r0 = (p >> l) << l;
      if (r0 + h >= n) continue;
      const int sz = n - r0 < S ? n - r0 : S;
      const R v = dc_key(sh.zs[p]);
      int rank = 0;
      for (int q = r0; q < r0 + sz; ++q) {
        const R u = dc_key(sh.zs[q]);
        rank += (u < v) || (u == v && q < p);
      }
      sh.itemcol[r0 + rank] = p - r0;
      d[(long long)(r0 + rank) * st] = sh.zs[p];
    }
    tm.sync();
    tm.mark(5);

    // ---- 5c. M in slot space, row-major [slot][column] at M + r0 S ----------------------
    for (long long g = tid; g < (long long)n * S; g += NT) {
      const int p = static_cast<int>(g % n);  // global slot
      const int c = static_cast<int>(g / n);  // output column
      const int r0 = (p >> l) << l;
      if (r0 + h >= n) continue;
      const int sz = n - r0 < S ? n - r0 : S;
      if (c >= sz) continue;
      const int pl = p - r0;
      const int it = sh.itemcol[r0 + c];
      const int kit = sh.kind[r0 + it];
      R v;
      if (kit < 0) {
        v = pl == it ? R(1) : R(0);
      } else {
        const int k2 = sh.kind[p];
        v = k2 < 0 ? R(0)
                   : sh.zh[r0 + k2] /
                         dc_delta(sh.dk + r0, sh.org + r0, sh.tau + r0, k2, kit) *
                         sh.nrm[r0 + kit];
      }
      M[(long long)r0 * S + (long long)pl * sz + c] = v;
    }
    tm.sync();

    // ---- 5d. deflation rotations on the rows of M, in reverse order ---------------------
    for (int g = tid; g < n; g += NT) {
      const int m = g >> l, r0 = m << l;
      if (r0 + h >= n) continue;
      const int sz = n - r0 < S ? n - r0 : S;
      const int c = g - r0;  // one column per thread
      BIPP_GLOBAL R* Mb = M + (long long)r0 * S;
      for (int t = sh.mR[m] - 1; t >= 0; --t) {
        const int a = sh.rp[r0 + t], b = sh.rn[r0 + t];
        const R cc = sh.rc[r0 + t], ss = sh.rs[r0 + t];
        const R ya = Mb[(long long)a * sz + c], yb = Mb[(long long)b * sz + c];
        Mb[(long long)a * sz + c] = cc * ya - ss * yb;
        Mb[(long long)b * sz + c] = ss * ya + cc * yb;
      }
    }
    tm.sync();
    tm.mark(6);

    // ---- 6. Q_new = diag(Q1, Q2) M ---------------------------------------------------------
    // blocks without a sibling are copied
    for (long long g = tid; g < (long long)n * S; g += NT) {
      const int row = static_cast<int>(g % n);
      const int c = static_cast<int>(g / n);
      const int r0 = (row >> l) << l;
      if (r0 + h < n) continue;
      const int sz = n - r0;
      if (c >= sz) continue;
      Bn[row + (long long)(r0 + c) * ldn] = Bp[row + (long long)(r0 + c) * ldp];
    }
    if (h >= TR) {
      for (int m = 0; m < nm; ++m) {
        const int r0 = m << l;
        if (r0 + h >= n) continue;
        const int sz = n - r0 < S ? n - r0 : S;
        dc_product_tiles<R, TR, NT>(tm, r0, sz, h, Bp, ldp, Bn, ldn, M + (long long)r0 * S,
                                    sh.slot_of, tile);
      }
    } else if (TR == 64 && h == 32) {
      for (int m = 0; m < nm; ++m) {
        const int r0 = m << l;
        if (r0 + h >= n) continue;
        const int sz = n - r0 < S ? n - r0 : S;
        dc_product_tiles<R, 32, NT>(tm, r0, sz, h, Bp, ldp, Bn, ldn, M + (long long)r0 * S,
                                    sh.slot_of, tile);
      }
    } else {
      // small blocks: one output per thread, consecutive threads down a column
      for (long long g = tid; g < (long long)n * S; g += NT) {
        const int row = static_cast<int>(g % n);
        const int c = static_cast<int>(g / n);
        const int r0 = (row >> l) << l;
        if (r0 + h >= n) continue;
        const int sz = n - r0 < S ? n - r0 : S;
        if (c >= sz) continue;
        const int k0 = row - r0 < h ? 0 : h, k1 = row - r0 < h ? h : sz;
        const BIPP_GLOBAL R* Mb = M + (long long)r0 * S;
        R acc = 0;
        for (int k = k0; k < k1; ++k)
          acc = fma(Bp[row + (long long)(r0 + k) * ldp],
                    Mb[(long long)(sh.slot_of[r0 + k] - r0) * sz + c], acc);
        Bn[row + (long long)(r0 + c) * ldn] = acc;
      }
    }
    tm.sync();
    tm.mark(7);
  }

  for (int i = tid; i < n; i += NT) d[(long long)i * st] *= scale;
  tm.sync();
}

}  // namespace bippamd
