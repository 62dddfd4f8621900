// Host side of the type-3 NUFFT: plan (kernel, grid sizes, FFT factorisation, correction
// tables) and the runner that launches points, sort, spread, FFT and interpolation.
#pragma once

#include <array>
#include <map>
#include <mutex>

#include "host_runtime.hpp"
#include "nufft3.hpp"

// ---------------------------------------------------------------------------------------
// Type-3 NUFFT plan (host): kernel, grid sizes, FFT factorisation, correction tables.
// Kernel width and shape follow the published exponential-of-semicircle rule for
// upsampling factor 2 (Barnett, Magland & af Klinteberg 2019): w = ceil(log10(1/tol)) + 1,
// beta = 2.30 w (2.20, 2.26, 2.38 for w = 2, 3, 4); grid size per dimension
// nf1 = smooth(2 sigma / pi * X * S + w + 1) and an inner type-2 grid nf2 = smooth(2 nf1).
// ---------------------------------------------------------------------------------------
namespace bippamd {
namespace {

int next_smooth(double x) {
  long long n = static_cast<long long>(std::ceil(x));
  if (n < 2) n = 2;
  n += n & 1;
  for (;; n += 2) {
    long long m = n;
    for (int p : {2, 3, 5})
      while (m % p == 0) m /= p;
    if (m == 1) return static_cast<int>(n);
  }
}

// Stockham stages for length m, radices at most `max_radix` (16, 9, 8 and 6 are
// in-register composite DFTs; 4, 2, 3, 5 otherwise): large radices first take the factors
// 2 and 3, then the rest; the stages run in ascending radix order.
int radix_cap(const char* name, int dflt) {
  const char* e = std::getenv(name);
  const int x = e ? std::atoi(e) : dflt;
  return x >= 16 ? 16 : (x >= 9 ? 9 : (x >= 8 ? 8 : (x >= 6 ? 6 : 4)));
}
// contiguous-axis lines (BIPP_FFT_RADIX_MAX, default 9) and strided passes
// (BIPP_FFT_PASS_RADIX_MAX, default 4); the defaults were the fastest at configuration 4
// (x lines of 3456 as radices 6, 8, 8, 9: 20.9 ms for the FFT against 22.4 ms with radices
// up to 4; composite radices in the strided passes cost occupancy and were slower)
int fft_radix_max() {
  static const int v = radix_cap("BIPP_FFT_RADIX_MAX", 9);
  return v;
}
int fft_pass_radix_max() {
  static const int v = radix_cap("BIPP_FFT_PASS_RADIX_MAX", 4);
  return v;
}
// 1: strided passes whose length factors into two dft radices run as two register stages
int fft_pass_rr_default() {
  static const int v = [] {
    const char* e = std::getenv("BIPP_FFT_PASS_RR");
    return e ? (std::atoi(e) != 0 ? 1 : 0) : 1;
  }();
  return v;
}
int spread_chunk_default() {
  static const int v = [] {
    const char* e = std::getenv("BIPP_SPREAD_XCD_CHUNK");
    return e ? std::max(0, std::atoi(e)) : 32;
  }();
  return v;
}

// multiplier and shift for exact division by d (see N3Radix)
void div_magic(int d, unsigned& m, int& s) {
  if (d < 1) fail(BIPP_INTERNAL_ERROR, "FFT divisor below 1");
  n3_div_magic(d, m, s);
}

N3Radix radices_of(int m, int max_radix) {
  N3Radix r;
  const int n = m;
  int twos = 0, threes = 0;
  while (m % 2 == 0) {
    ++twos;
    m /= 2;
  }
  while (m % 3 == 0) {
    ++threes;
    m /= 3;
  }
  auto push = [&](int p) {
    if (r.count >= 20) fail(BIPP_INTERNAL_ERROR, "FFT length has too many factors");
    r.r[r.count++] = p;
  };
  if (max_radix >= 16)
    for (; twos >= 4; twos -= 4) push(16);
  if (max_radix >= 8)
    for (; twos >= 3; twos -= 3) push(8);
  if (max_radix >= 9)
    for (; threes >= 2; threes -= 2) push(9);
  if (max_radix >= 6)
    for (; twos >= 1 && threes >= 1; --twos, --threes) push(6);
  for (; twos >= 2; twos -= 2) push(4);
  if (twos == 1) push(2);
  for (; threes >= 1; --threes) push(3);
  while (m % 5 == 0) {
    push(5);
    m /= 5;
  }
  if (m != 1) fail(BIPP_INTERNAL_ERROR, "FFT length is not 2^a 3^b 5^c");
  std::sort(r.r, r.r + r.count);
  div_magic(n, r.n_m, r.n_s);
  int ns = 1;
  for (int st = 0; st < r.count; ++st) {
    div_magic(n / r.r[st], r.nb_m[st], r.nb_s[st]);
    div_magic(ns, r.ns_m[st], r.ns_s[st]);
    ns *= r.r[st];
  }
  return r;
}

// Limits of one plan: x lines are transformed whole in LDS, the strided axes (y, z) in at
// most two passes of length <= 128 each.
constexpr int N3_MAX_X = 4096;
constexpr int N3_MAX_SUB = 128;

// n = n1 * n2 with n1, n2 <= 128 (n2 == 1 when n <= 128 itself); false if there is none
bool split_axis(int n, int& n1, int& n2) {
  if (n <= N3_MAX_SUB) {
    n1 = n;
    n2 = 1;
    return true;
  }
  for (int a = N3_MAX_SUB; a >= 2; --a)
    if (n % a == 0 && n / a <= N3_MAX_SUB) {
      n1 = a;
      n2 = n / a;
      return true;
    }
  return false;
}

// smallest 2^a 3^b 5^c >= x that a strided pass can transform (n1 n2 with n1, n2 <= 128),
// or -1 above 128^2
int next_smooth_strided(double x) {
  for (int n = next_smooth(x); n <= N3_MAX_SUB * N3_MAX_SUB; n = next_smooth(n + 1.0)) {
    int a, b;
    if (split_axis(n, a, b)) return n;
  }
  return -1;
}

// Gauss-Legendre nodes/weights mapped to [0, 1] (Newton iteration on P_n)
void gauss_legendre01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5));
    double dp = 1;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1 - z);
    w[i] = 1.0 / ((1 - z * z) * dp * dp);  // 2/((1-z^2) P'^2) scaled by 1/2 for [0, 1]
  }
}

// Quadrature weights times the kernel at the nodes: ew[i] = w[i] exp(beta (sqrt(1 - x_i^2) - 1)),
// so that phi(xi) = 2 sum_i ew[i] cos(xi x_i).
std::vector<double> es_ft_weights(double beta, const std::vector<double>& x,
                                  const std::vector<double>& w) {
  std::vector<double> ew(x.size());
  for (size_t i = 0; i < x.size(); ++i) ew[i] = w[i] * std::exp(beta * (std::sqrt(1 - x[i] * x[i]) - 1));
  return ew;
}
double es_ft_host(double xi, const std::vector<double>& x, const std::vector<double>& ew) {
  double s = 0;
  for (size_t i = 0; i < x.size(); ++i) s += ew[i] * std::cos(xi * x[i]);
  return 2 * s;
}

}  // namespace

// Grid sizes of a plan for points of half-extent X around C and pixels of half-extent S
// around D, per axis: spread grid nf1 = smooth(2 sigma / pi X S + w + 1) (at least 2 w) and
// FFT grid nf2 = smooth(2 nf1), strided axes restricted to lengths two passes can handle.
// fits is false when an axis exceeds what one plan can transform.
struct N3Sizes {
  int w = 2;
  int nf1[3] = {0, 0, 0}, nf2[3] = {0, 0, 0};
  bool fits = true;
  double grid_cells() const { return double(nf2[0]) * nf2[1] * nf2[2]; }
};

inline int n3_width(double tol) {
  const int w = static_cast<int>(std::ceil(std::log10(1.0 / tol))) + 1;
  return std::max(2, std::min(w, 16));
}

inline N3Sizes n3_sizes(double tol, const double X[3], const double S[3]) {
  N3Sizes z;
  z.w = n3_width(tol);
  const double sigma = 2.0;
  for (int d = 0; d < 3; ++d) {
    const double Sd = std::max(S[d], 1e-9), Xd = std::max(X[d], 1e-9);
    const double want = 2 * sigma / 3.14159265358979323846 * Xd * Sd + z.w + 1;
    if (want > 1e9) {
      z.fits = false;
      z.nf1[d] = z.nf2[d] = 1 << 30;
      continue;
    }
    z.nf1[d] = std::max(next_smooth(want), 2 * z.w);
    if (d == 0) {
      z.nf2[d] = next_smooth(2.0 * z.nf1[d]);
      z.fits = z.fits && z.nf2[d] <= N3_MAX_X;
    } else {
      const int n = next_smooth_strided(2.0 * z.nf1[d]);
      z.fits = z.fits && n > 0;
      z.nf2[d] = n > 0 ? n : next_smooth(2.0 * z.nf1[d]);
    }
  }
  return z;
}

struct N3Plan {
  int w = 6;
  double beta = 13.8;
  int nf1[3], nf2[3], n1[3], n2[3];
  double h1[3], h2[3], gam[3], C[3], D[3];
  int bin[3] = {8, 8, 8}, nbins[3], nback[3], off[3];
  // FFT windows: spread input l in [-nf1/2, nf1/2) and the cells the pixels interpolate from
  N3Window in_w[3], out_w[3];
  long long grid_cells = 0;
  std::vector<double> gl_x, gl_w, gl_ew;
  std::vector<double> deconv[3];
  // kernel (w, beta) and grid sizes the tables were last built for (they are reused)
  int tab_w = -1, dec_nf1[3] = {-1, -1, -1}, dec_nf2[3] = {-1, -1, -1};
  // phi(xi) = int_{-1}^{1} exp(beta (sqrt(1 - z^2) - 1)) cos(xi z) dz at xi = (j - 1) / ft_inv_delta
  std::vector<double> ft_tab;
  double ft_inv_delta = 0;

  void make(double tol, const double X_in[3], const double S_in[3], const double C_in[3],
            const double D_in[3]) {
    const N3Sizes z = n3_sizes(tol, X_in, S_in);
    if (!z.fits)
      fail(BIPP_INVALID_PARAMETER_ERROR,
           "NUFFT grid of " + std::to_string(z.nf2[0]) + " x " + std::to_string(z.nf2[1]) + " x " +
               std::to_string(z.nf2[2]) + " points exceeds one plan (x at most " +
               std::to_string(N3_MAX_X) + ", y and z at most " +
               std::to_string(N3_MAX_SUB * N3_MAX_SUB) +
               "); partition the uvw domain (Partition auto or grid)");
    w = z.w;
    beta = w == 2 ? 2.20 * w : (w == 3 ? 2.26 * w : (w == 4 ? 2.38 * w : 2.30 * w));
    const double sigma = 2.0;
    const bool new_kernel = tab_w != w;
    if (new_kernel) {
      gauss_legendre01(64, gl_x, gl_w);
      gl_ew = es_ft_weights(beta, gl_x, gl_w);
    }
    grid_cells = 1;
    for (int d = 0; d < 3; ++d) {
      const double S = std::max(S_in[d], 1e-9);
      nf1[d] = z.nf1[d];
      nf2[d] = z.nf2[d];
      h1[d] = 2 * 3.14159265358979323846 / nf1[d];
      h2[d] = 2 * 3.14159265358979323846 / nf2[d];
      gam[d] = nf1[d] / (2 * sigma * S);
      C[d] = C_in[d];
      D[d] = D_in[d];
      if (d == 0) {
        n1[0] = nf2[0];
        n2[0] = 1;
      } else if (!split_axis(nf2[d], n1[d], n2[d])) {
        fail(BIPP_INTERNAL_ERROR, "NUFFT strided axis length without a two-pass split");
      }
      // spread blocks over l + nf1/2 in [0, nf1); a point's support [l0, l0 + w) starts in
      // its bin and reaches at most nback blocks further
      off[d] = nf1[d] / 2;
      nbins[d] = (nf1[d] + bin[d] - 1) / bin[d];
      nback[d] = 1 + (w - 2) / bin[d];
      grid_cells *= nf2[d];
      in_w[d] = N3Window{nf2[d] - nf1[d] / 2, nf1[d]};
      // pixels sit at |a| <= h1 gam S / h2 grid cells and read cells ceil(a - w/2) + [0, w);
      // one cell of margin on each side
      const double amax = h1[d] * gam[d] * S / h2[d];
      const long long klo = static_cast<long long>(std::floor(-amax - 0.5 * w)) - 1;
      const long long khi = static_cast<long long>(std::ceil(amax - 0.5 * w)) + w;
      const long long cnt = khi - klo + 1;
      if (cnt >= nf2[d])
        out_w[d] = N3Window{0, nf2[d]};
      else
        out_w[d] = N3Window{static_cast<int>(((klo % nf2[d]) + nf2[d]) % nf2[d]),
                            static_cast<int>(cnt)};
      if (new_kernel || dec_nf1[d] != nf1[d] || dec_nf2[d] != nf2[d]) {
        deconv[d].resize(nf1[d]);
        const double a2 = 0.5 * w * h2[d];
        for (int i = 0; i < nf1[d]; ++i) {
          const int l = i - nf1[d] / 2;
          deconv[d][i] = 1.0 / (a2 * es_ft_host(l * a2, gl_x, gl_ew));
        }
        dec_nf1[d] = nf1[d];
        dec_nf2[d] = nf2[d];
      }
    }
    // the pixel-side correction evaluates phi at |xi| = |gam (x - D)| w h1 / 2, at most
    // pi w / (2 sigma) for pixels inside the extent: tabulated with margin, interpolated
    // by cubics on the device (phi is even; entry 0 is phi(-delta) = phi(delta))
    if (new_kernel) {
      constexpr int NFT = 2048;
      const double xi_max = 1.05 * 3.14159265358979323846 * w / (2 * sigma) + 1e-3;
      const double delta = xi_max / NFT;
      ft_tab.resize(NFT + 3);
      for (int j = 0; j < NFT + 3; ++j) ft_tab[j] = es_ft_host(std::fabs((j - 1) * delta), gl_x, gl_ew);
      ft_inv_delta = 1.0 / delta;
      tab_w = w;
    }
  }
  long long n_bins() const { return (long long)nbins[0] * nbins[1] * nbins[2]; }
};

// Run-time shapes of the spread and the FFT, picked per grid shape by the autotuner:
// largest radix of the x lines and of the strided passes (radices_of), the largest item width
// of the strided passes (2^pass_tw_shift x positions, at most about 4096 elements per item)
// and the spread's blocks per XCD run (0: launch order).  The defaults are the measured best
// at configuration 4 (x lines radix 9: FFT 18.3 ms against 18.5-19.7 ms with 8 or 16; passes
// radix 4; spread runs of 32: 9.4 ms against 11.1 ms in launch order); BIPP_FFT_RADIX_MAX,
// BIPP_FFT_PASS_RADIX_MAX and BIPP_SPREAD_XCD_CHUNK override them.
struct N3Tuning {
  int lines_radix = fft_radix_max();
  int pass_radix = fft_pass_radix_max();
  int pass_tw_shift = 6;
  int spread_chunk = spread_chunk_default();
  int pass_rr = fft_pass_rr_default();
};

inline std::vector<N3Tuning> n3_tune_candidates() {
  std::vector<N3Tuning> out;
  for (int lr : {9, 16, 8})
    for (int pr : {4, 8})
      for (int tw : {6, 5})
        for (int ch : {32, 0}) {
          N3Tuning t;
          t.lines_radix = lr;
          t.pass_radix = pr;
          t.pass_tw_shift = tw;
          t.spread_chunk = ch;
          out.push_back(t);
        }
  return out;
}

inline std::string n3_tune_key(const N3Plan& pl, int L, size_t tg) {
  return std::to_string(pl.nf2[0]) + "x" + std::to_string(pl.nf2[1]) + "x" +
         std::to_string(pl.nf2[2]) + "/w" + std::to_string(pl.w) + "/L" + std::to_string(L) +
         "/" + std::to_string(tg);
}

// process-wide table of tuned shapes
inline std::mutex& n3_tune_mu() {
  static std::mutex m;
  return m;
}
inline std::map<std::string, N3Tuning>& n3_tune_table() {
  static std::map<std::string, N3Tuning> t;
  return t;
}
inline bool n3_tune_lookup(const std::string& key, N3Tuning& out) {
  std::lock_guard<std::mutex> lock(n3_tune_mu());
  auto it = n3_tune_table().find(key);
  if (it == n3_tune_table().end()) return false;
  out = it->second;
  return true;
}
inline void n3_tune_store(const std::string& key, const N3Tuning& t) {
  std::lock_guard<std::mutex> lock(n3_tune_mu());
  n3_tune_table()[key] = t;
}
inline void n3_tune_clear() {
  std::lock_guard<std::mutex> lock(n3_tune_mu());
  n3_tune_table().clear();
}

// The C-ABI view of a tuning: shape[4] = (lines radix, pass radix, pass item shift, spread
// chunk); candidates: 5 doubles per timed candidate (its shape and ms), capacity candidates.
inline void export_tuning(const N3Tuning& t, const std::vector<std::pair<N3Tuning, float>>& log,
                          int32_t* shape, size_t capacity, double* cand, size_t* n_cand) {
  if (shape) {
    shape[0] = t.lines_radix;
    shape[1] = t.pass_radix;
    shape[2] = t.pass_tw_shift;
    shape[3] = t.spread_chunk;
  }
  if (n_cand) *n_cand = log.size();
  if (cand)
    for (size_t k = 0; k < log.size() && k < capacity; ++k) {
      const N3Tuning& c = log[k].first;
      double* r = cand + 5 * k;
      r[0] = c.lines_radix;
      r[1] = c.pass_radix;
      r[2] = c.pass_tw_shift;
      r[3] = c.spread_chunk;
      r[4] = log[k].second;
    }
}

// How the points of a batch are split into uvw boxes, each transformed by its own plan
// (Partition of include/bipp/image_synthesis.hpp:23-53).
enum { N3_PART_AUTO = 0, N3_PART_NONE = 1, N3_PART_GRID = 2 };

struct N3Options {
  double tol = 1e-5;
  int part = N3_PART_AUTO;
  int dims[3] = {1, 1, 1};
  // grid memory of one box's plan; 0: two thirds of the device memory
  double max_grid_bytes = 0;
};

// Partition option from the C-ABI (method BIPP_PARTITION_*, dims[3] for a grid).
inline void set_partition(N3Options& o, int method, const size_t* dims) {
  require(method >= N3_PART_AUTO && method <= N3_PART_GRID, BIPP_INVALID_PARAMETER_ERROR,
          "unknown partition method");
  o.part = method;
  if (method == N3_PART_GRID) {
    require(dims != nullptr, BIPP_INVALID_POINTER_ERROR, "null partition dimensions");
    for (int d = 0; d < 3; ++d) {
      require(dims[d] >= 1 && dims[d] <= 4096, BIPP_INVALID_PARAMETER_ERROR,
              "partition dimensions must be in [1, 4096]");
      o.dims[d] = static_cast<int>(dims[d]);
    }
  }
}

// Boxes per axis for points in [lo, hi] and pixels of half-extent S (the auto partition of
// src/gpu/nufft.hpp:110-125 with the sizing loop of src/nufft_util.hpp:12-50): while one
// box's plan exceeds what a plan can transform, the axis furthest over its limit gets one
// more box; then, while one level's grid of a box exceeds max_bytes, the axis with the
// largest product of box extent and pixel extent does.
inline std::array<int, 3> n3_auto_partition(double tol, const double lo[3], const double hi[3],
                                             const double S[3], size_t elem, double max_bytes) {
  std::array<int, 3> g = {1, 1, 1};
  for (int iter = 0; iter < 100000; ++iter) {
    double X[3], ext[3];
    for (int d = 0; d < 3; ++d) {
      ext[d] = (hi[d] - lo[d]) / g[d];
      X[d] = 0.5 * ext[d];
    }
    const N3Sizes z = n3_sizes(tol, X, S);
    if (!z.fits) {
      int worst = 0;
      double over = -1;
      for (int d = 0; d < 3; ++d) {
        const double lim = d == 0 ? N3_MAX_X : double(N3_MAX_SUB) * N3_MAX_SUB;
        const double r = z.nf2[d] / lim;
        if (r > over) {
          over = r;
          worst = d;
        }
      }
      ++g[worst];
      continue;
    }
    if (z.grid_cells() * double(elem) <= max_bytes) return g;
    int best = 0;
    for (int d = 1; d < 3; ++d)
      if (ext[d] * S[d] > ext[best] * S[best]) best = d;
    ++g[best];
  }
  fail(BIPP_INVALID_PARAMETER_ERROR, "no uvw partition keeps the NUFFT grid within memory");
}

// Summary of a profiled spread launch: block durations (100 MHz wall clock), visited points
// and items, and the slowest blocks with their bin coordinates.
inline void report_spread_profile(const unsigned long long* dprof, long long nb, const N3Plan& pl,
                                  hipStream_t st) {
  std::vector<unsigned long long> h((size_t)nb * 4);
  BIPP_HIP(hipMemcpyAsync(h.data(), dprof, sizeof(unsigned long long) * h.size(),
                          hipMemcpyDeviceToHost, st));
  BIPP_HIP(hipStreamSynchronize(st));
  unsigned long long t0 = ~0ull, t1 = 0;
  double sum_dur = 0, sum_pts = 0, sum_items = 0;
  std::vector<std::pair<double, long long>> dur;
  dur.reserve(nb);
  for (long long b = 0; b < nb; ++b) {
    const unsigned long long* r = &h[(size_t)b * 4];
    if (!r[0]) continue;
    t0 = std::min(t0, r[0]);
    t1 = std::max(t1, r[1]);
    const double d = double(r[1] - r[0]) / 100.0;
    sum_dur += d;
    sum_pts += double(r[2]);
    sum_items += double(r[3]);
    dur.emplace_back(d, b);
  }
  std::sort(dur.begin(), dur.end(), [](auto& a, auto& b) { return a.first > b.first; });
  std::fprintf(stderr,
               "[bipp spread profile] blocks %lld span %.1f us, sum of block us %.0f, visits %.3g, "
               "items %.3g, median block %.1f us\n",
               (long long)dur.size(), double(t1 - t0) / 100.0, sum_dur, sum_pts, sum_items,
               dur.empty() ? 0.0 : dur[dur.size() / 2].first);
  for (size_t i = 0; i < std::min<size_t>(5, dur.size()); ++i) {
    const long long b = dur[i].second;
    const unsigned long long* r = &h[(size_t)b * 4];
    std::fprintf(stderr, "  block (%lld,%lld,%lld) %.1f us, visits %llu, items %llu\n",
                 b % pl.nbins[0], (b / pl.nbins[0]) % pl.nbins[1], b / ((long long)pl.nbins[0] * pl.nbins[1]),
                 dur[i].first, r[2], r[3]);
  }
}

// NUFFT image stage of one batch.  values(): VV_l = Z diag(w_l) Z^H of every (sample, level)
// (lower triangle; the trace in slot (0, 1)).  image(): the points of the source split into
// uvw boxes, each box spread, transformed and interpolated with a plan of its own extent, every
// level accumulated into `img` (TG per pixel).
template <typename TG>
struct N3Runner {
  DevBuf grid, sorted, sorted_vals;
  DevBuf bin_count, bin_start, bin_rank, scan_tmp, vv, deconv, gl, twid, prof_buf;
  DevBuf group_of, group_count, group_start, group_fill, group_tmp, slot_list, ranges;
  std::vector<cplx<TG>> htw;
  std::vector<TG> hdec;
  N3Plan plan;
  long long ld_vv = 0, vv_stride = 0;
  // last batch: FFT grid of its (first) box, boxes per axis, boxes with points
  int last_nf2[3] = {0, 0, 0};
  int last_dims[3] = {1, 1, 1};
  int last_boxes = 0;
  // event pairs per transform pass (a box, or one level group of a box):
  // [pass][sort, spread, FFT, interpolation][begin, end]
  std::vector<hipEvent_t> evs;
  int n_passes = 0;
  hipEvent_t ev_vv[2] = {nullptr, nullptr};
  size_t tw_off[3][3] = {};
  bool autotune = false;
  N3Tuning last_tuning;
  std::vector<std::pair<N3Tuning, float>> tune_log;  // candidates and ms of the last tuning
  double* h_buf = nullptr;  // pinned: ranges, counts and starts read back by the host
  double max_bytes_ = 0;    // grid memory bound of one box (bytes)
  size_t h_cap = 0;

  N3Runner() = default;
  N3Runner(const N3Runner&) = delete;
  ~N3Runner() {
    for (auto e : evs) (void)hipEventDestroy(e);
    for (auto e : ev_vv)
      if (e) (void)hipEventDestroy(e);
    if (h_buf) (void)hipHostFree(h_buf);
  }

  double* host_buf(size_t n) {
    if (h_cap < n) {
      if (h_buf) BIPP_HIP(hipHostFree(h_buf));
      h_buf = nullptr;
      BIPP_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_buf), sizeof(double) * n));
      h_cap = n;
    }
    return h_buf;
  }
  hipEvent_t pass_event(int pass, int stage, int edge) {
    const size_t i = ((size_t)pass * 4 + stage) * 2 + edge;
    while (evs.size() <= i) {
      hipEvent_t e;
      BIPP_HIP(hipEventCreate(&e));
      evs.push_back(e);
    }
    return evs[i];
  }

  template <typename RE>
  void values(ContextImpl& c, int nS, int L, int n_cols, int rows_max, const int* n_rows,
              const cplx<RE>* z, long long stride_z, int ldz, const RE* wd) {
    hipStream_t st = c.stream;
    if (!ev_vv[0]) {
      BIPP_HIP(hipEventCreate(&ev_vv[0]));
      BIPP_HIP(hipEventCreate(&ev_vv[1]));
    }
    BIPP_HIP(hipEventRecord(ev_vv[0], st));
    ld_vv = rows_max;
    vv_stride = ld_vv * ld_vv;
    cplx<TG>* vvp = vv.as<cplx<TG>>((size_t)nS * L * vv_stride);
    N3VVParams<RE> vp;
    vp.n_levels = L;
    vp.n_cols = n_cols;
    vp.n_rows = n_rows;
    vp.n_rows_const = rows_max;
    vp.z = z;
    vp.stride_z = stride_z;
    vp.ldz = ldz;
    vp.wd = wd;
    vp.vv = vvp;
    vp.vv_stride = vv_stride;
    vp.ld_vv = ld_vv;
    launch_vv_levels<RE, TG>(vp, nS, rows_max, st);
    if (merge_auto(rows_max)) launch_vv_trace<TG>(nS, L, n_rows, rows_max, vvp, vv_stride, ld_vv, st);
    BIPP_HIP(hipEventRecord(ev_vv[1], st));
  }
  static bool merge_auto(int rows_max) { return rows_max >= 2; }

  // levels transformed together by the current plan: as many as the grid memory bound of
  // one box holds, at least one
  int lg_levels(int L) const {
    const double lvl_bytes = double(plan.grid_cells) * sizeof(cplx<TG>);
    return static_cast<int>(std::max(1.0, std::min<double>(L, max_bytes_ / lvl_bytes)));
  }

  // Stage times of the last batch (ms): virtual visibilities, point binning and sort,
  // spread, FFT, interpolation.  Waits for the stream.
  void stage_ms(ContextImpl& c, double* ms) {
    c.sync();
    float t = 0;
    BIPP_HIP(hipEventElapsedTime(&t, ev_vv[0], ev_vv[1]));
    ms[0] += t;
    for (int b = 0; b < n_passes; ++b)
      for (int k = 0; k < 4; ++k) {
        BIPP_HIP(hipEventElapsedTime(&t, pass_event(b, k, 0), pass_event(b, k, 1)));
        ms[1 + k] += t;
      }
  }

  void image(ContextImpl& c, N3Source src, int nS, int L, const N3Options& opt,
             const double pix_lo[3], const double pix_hi[3], const double* lmn64, int n_pix,
             void* img, bool phase_fp32) {
    hipStream_t st = c.stream;
    n_passes = 0;
    last_boxes = 0;
    src.merge_auto = merge_auto(src.n_ant_max) ? 1 : 0;
    src.ld_vv = ld_vv;
    const long long sps = src.slots_per_sample();
    const long long n_slots = sps * nS;
    if (n_slots <= 0) return;
    if (n_slots > 0x7fffffffLL)
      fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT points in one batch (reduce the sample batch size)");
    // extent of every sample's points
    double* rg = ranges.as<double>(6 * (size_t)nS);
    launch_slot_range(src, nS, nullptr, nullptr, nullptr, rg, st);
    double* h = host_buf(6 * (size_t)nS);
    BIPP_HIP(hipMemcpyAsync(h, rg, sizeof(double) * 6 * nS, hipMemcpyDeviceToHost, st));
    c.sync();
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int k = 0; k < nS; ++k)
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], h[6 * k + d]);
        hi[d] = std::max(hi[d], h[6 * k + 3 + d]);
      }
    if (lo[0] > hi[0]) return;  // no points
    double S[3], D[3];
    for (int d = 0; d < 3; ++d) {
      S[d] = 0.5 * (pix_hi[d] - pix_lo[d]);
      D[d] = 0.5 * (pix_hi[d] + pix_lo[d]);
    }
    // grid memory of one box: the option, else BIPP_NUFFT_MAX_GRID_BYTES, else two thirds of
    // the device memory (src/gpu/nufft.hpp:112); levels that do not fit together are
    // transformed one group after another
    double maxb = opt.max_grid_bytes;
    if (maxb <= 0) {
      const char* e = std::getenv("BIPP_NUFFT_MAX_GRID_BYTES");
      maxb = e ? std::atof(e) : 0.0;
    }
    if (maxb <= 0) {
      size_t fr = 0, tot = 0;
      BIPP_HIP(hipMemGetInfo(&fr, &tot));
      maxb = double(tot) / 3 * 2;
    }
    max_bytes_ = maxb;
    std::array<int, 3> dims = {1, 1, 1};
    if (opt.part == N3_PART_GRID) {
      for (int d = 0; d < 3; ++d) dims[d] = std::max(1, opt.dims[d]);
    } else if (opt.part == N3_PART_AUTO) {
      dims = n3_auto_partition(opt.tol, lo, hi, S, sizeof(cplx<TG>), maxb);
    }
    for (int d = 0; d < 3; ++d) last_dims[d] = dims[d];
    const long long n_groups = (long long)dims[0] * dims[1] * dims[2];
    if (n_groups > (1 << 20)) fail(BIPP_INVALID_PARAMETER_ERROR, "too many uvw partition boxes");
    if (n_groups == 1) {
      double X[3], C[3];
      for (int d = 0; d < 3; ++d) {
        X[d] = 0.5 * (hi[d] - lo[d]);
        C[d] = 0.5 * (hi[d] + lo[d]);
      }
      plan.make(opt.tol, X, S, C, D);
      transform(c, src, nullptr, n_slots, nS, L, lmn64, n_pix, img, phase_fp32, 0);
      last_boxes = 1;
      return;
    }
    // box of every point, counting sort of the slots by box
    N3PartParams pp;
    pp.src = src;
    pp.n_samples = nS;
    for (int d = 0; d < 3; ++d) {
      pp.dims[d] = dims[d];
      pp.lo[d] = lo[d];
      pp.inv[d] = hi[d] > lo[d] ? dims[d] / (hi[d] - lo[d]) : 0.0;
    }
    pp.group_of = group_of.as<int>(n_slots);
    pp.group_count = group_count.as<unsigned>(n_groups);
    launch_zero<unsigned>(pp.group_count, n_groups, st);
    launch_n3_partition(pp, st);
    unsigned* gstart = group_start.as<unsigned>(n_groups);
    unsigned* gtmp = group_tmp.as<unsigned>(3 * ((size_t)n_groups / 1024 + 2) + 4096);
    launch_scan(pp.group_count, gstart, gtmp, (int)n_groups, st);
    unsigned* gfill = group_fill.as<unsigned>(n_groups);
    launch_zero<unsigned>(gfill, n_groups, st);
    unsigned long long* list = slot_list.as<unsigned long long>(n_slots);
    launch_n3_group_scatter(n_slots, pp.group_of, gstart, gfill, list, st);
    // extent of every box's points
    double* grg = ranges.as<double>(6 * (size_t)n_groups);
    launch_slot_range(src, (int)n_groups, list, gstart, pp.group_count, grg, st);
    double* hb = host_buf(8 * (size_t)n_groups);
    BIPP_HIP(hipMemcpyAsync(hb, grg, sizeof(double) * 6 * n_groups, hipMemcpyDeviceToHost, st));
    // counts and starts as unsigned pairs after the ranges
    unsigned* hc = reinterpret_cast<unsigned*>(hb + 6 * n_groups);
    BIPP_HIP(hipMemcpyAsync(hc, pp.group_count, sizeof(unsigned) * n_groups, hipMemcpyDeviceToHost, st));
    BIPP_HIP(hipMemcpyAsync(hc + n_groups, gstart, sizeof(unsigned) * n_groups,
                            hipMemcpyDeviceToHost, st));
    c.sync();
    std::vector<double> gr(hb, hb + 6 * n_groups);
    std::vector<unsigned> gc(hc, hc + n_groups), gs(hc + n_groups, hc + 2 * n_groups);
    int box = 0;
    for (long long g = 0; g < n_groups; ++g) {
      if (!gc[g]) continue;
      double X[3], C[3];
      for (int d = 0; d < 3; ++d) {
        X[d] = 0.5 * (gr[6 * g + 3 + d] - gr[6 * g + d]);
        C[d] = 0.5 * (gr[6 * g + 3 + d] + gr[6 * g + d]);
      }
      plan.make(opt.tol, X, S, C, D);
      transform(c, src, list + gs[g], gc[g], nS, L, lmn64, n_pix, img, phase_fp32, box);
      ++box;
    }
    last_boxes = box;
  }

 private:
  // Points, sort, spread, FFT and interpolation of the listed slots (all slots when list is
  // null) with the current plan.
  void transform(ContextImpl& c, const N3Source& src, const unsigned long long* list,
                 long long n_list, int nS, int L, const double* lmn64, int n_pix, void* img,
                 bool phase_fp32, int box) {
    hipStream_t st = c.stream;
    const N3Plan& pl = plan;
    for (int d = 0; d < 3; ++d)
      if (box == 0) last_nf2[d] = pl.nf2[d];
    const int first = n_passes;  // the sort is timed with the box's first pass
    auto mark0 = [&](int edge) { BIPP_HIP(hipEventRecord(pass_event(first, 0, edge), st)); };
    mark0(0);
    N3PointsParams pp;
    pp.src = src;
    pp.slot_list = list;
    pp.n_list = n_list;
    for (int d = 0; d < 3; ++d) {
      pp.C[d] = pl.C[d];
      pp.D[d] = pl.D[d];
      pp.inv_gh[d] = 1.0 / (pl.gam[d] * pl.h1[d]);
      pp.off[d] = pl.off[d];
      pp.bin_size[d] = pl.bin[d];
      pp.nbins[d] = pl.nbins[d];
    }
    pp.w = pl.w;
    pp.phase_fp32 = phase_fp32 ? 1 : 0;
    pp.vv = vv.p;
    pp.vv_stride = vv_stride;
    pp.n_levels = L;
    pp.vv_double = sizeof(TG) == 8 ? 1 : 0;
    const long long nb = pl.n_bins();
    if (nb > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT bins");
    // counting sort without an unsorted copy: count per bin, scan, then every point written
    // into its bin's range
    pp.bin_count = bin_count.as<unsigned>(nb);
    pp.rank = bin_rank.as<unsigned>(n_list);
    launch_zero<unsigned>(pp.bin_count, nb, st);
    pp.fill = 0;
    launch_n3_points(pp, nS, st);
    unsigned* bstart = bin_start.as<unsigned>(nb);
    unsigned* btmp = scan_tmp.as<unsigned>(3 * ((size_t)nb / 1024 + 2) + 4096);
    launch_scan(pp.bin_count, bstart, btmp, (int)nb, st);
    N3Point* srt = sorted.as<N3Point>(n_list);
    cplx<TG>* svals = sorted_vals.as<cplx<TG>>((size_t)n_list * L);
    pp.fill = 1;
    pp.bin_start = bstart;
    pp.pts = srt;
    pp.vals = svals;
    launch_n3_points(pp, nS, st);
    mark0(1);
    // levels in groups whose grids fit the memory bound of one box (at least one level)
    const long long G = pl.grid_cells;
    const int lg_max = lg_levels(L);
    cplx<TG>* cg = grid.as<cplx<TG>>((size_t)std::min(L, lg_max) * G);
    upload_tables(pl, st);
    // the run-time shapes of the spread and the FFT: the tuned ones of this grid shape, or
    // the defaults; with autotuning on, every candidate is timed on the first level group of
    // the first batch of a shape (the spread rewrites its whole window and the FFT reads only
    // what the spread and its own earlier passes wrote, so the runs repeat exactly)
    N3Tuning cfg = tuning_for(pl, L, c, bstart, pp.bin_count, srt, svals, cg, std::min(L, lg_max));
    last_tuning = cfg;
    for (int g0 = 0; g0 < L; g0 += lg_max) {
      const int lg = std::min(lg_max, L - g0);
      const int pass = n_passes++;
      auto mark = [&](int stage, int edge) {
        BIPP_HIP(hipEventRecord(pass_event(pass, stage, edge), st));
      };
      if (pass != first) {
        mark(0, 0);  // the points are sorted once per box
        mark(0, 1);
      }
      mark(1, 0);
      spread_group(pl, L, g0, lg, bstart, pp.bin_count, srt, svals, cg, cfg, st);
      mark(1, 1);
      mark(2, 0);
      fft_group(pl, lg, cg, cfg, st);
      mark(2, 1);
      mark(3, 0);
      interp_group(pl, g0, lg, cg, lmn64, n_pix, img, st);
      mark(3, 1);
    }
  }

  // host tables of the plan on the device: deconvolution factors, twiddles, kernel transform
  void upload_tables(const N3Plan& pl, hipStream_t st) {
    const int nd = pl.nf1[0] + pl.nf1[1] + pl.nf1[2];
    hdec.resize(nd);
    for (int i = 0; i < pl.nf1[0]; ++i) hdec[i] = static_cast<TG>(pl.deconv[0][i]);
    for (int i = 0; i < pl.nf1[1]; ++i) hdec[pl.nf1[0] + i] = static_cast<TG>(pl.deconv[1][i]);
    for (int i = 0; i < pl.nf1[2]; ++i)
      hdec[pl.nf1[0] + pl.nf1[1] + i] = static_cast<TG>(pl.deconv[2][i]);
    TG* ddec = deconv.as<TG>(nd);
    BIPP_HIP(hipMemcpyAsync(ddec, hdec.data(), sizeof(TG) * nd, hipMemcpyHostToDevice, st));
    // twiddle tables exp(+2 pi i t / n) computed in FP64: x, then per strided axis n1, n2 and
    // n1 * n2; two-level tables: exp(+2 pi i t / n) = fine[t mod 64] * coarse[t / 64]
    htw.clear();
    auto add_table = [&](long long n) {
      const size_t o = htw.size();
      auto root = [&](long long t) {
        const double ang = 2.0 * 3.14159265358979323846 * static_cast<double>(t % n) / n;
        return mk<TG>(static_cast<TG>(std::cos(ang)), static_cast<TG>(std::sin(ang)));
      };
      for (long long t = 0; t < 64; ++t) htw.push_back(root(t));
      for (long long s = 0; s < (n + 63) / 64; ++s) htw.push_back(root(64 * s));
      return o;
    };
    tw_off[0][0] = add_table(pl.nf2[0]);
    for (int d = 1; d < 3; ++d) {
      tw_off[d][0] = add_table(pl.n1[d]);
      tw_off[d][1] = add_table(pl.n2[d]);
      tw_off[d][2] = pl.n2[d] > 1 ? add_table((long long)pl.n1[d] * pl.n2[d]) : 0;
    }
    cplx<TG>* dtw = twid.as<cplx<TG>>(htw.size());
    BIPP_HIP(hipMemcpyAsync(dtw, htw.data(), sizeof(cplx<TG>) * htw.size(), hipMemcpyHostToDevice, st));
    double* dft = gl.as<double>(pl.ft_tab.size());
    BIPP_HIP(hipMemcpyAsync(dft, pl.ft_tab.data(), sizeof(double) * pl.ft_tab.size(),
                            hipMemcpyHostToDevice, st));
  }

  // spread + deconvolution of levels [g0, g0 + lg) into [-nf1/2, nf1/2)^3 of the nf2 grid
  void spread_group(const N3Plan& pl, int L, int g0, int lg, const unsigned* bstart,
                    const unsigned* bcount, const N3Point* srt, const cplx<TG>* svals,
                    cplx<TG>* cg, const N3Tuning& cfg, hipStream_t st) {
    const TG* ddec = static_cast<const TG*>(deconv.p);
    N3SpreadParams<TG> sp;
    sp.n_levels = lg;
    sp.lv0 = g0;
    sp.n_levels_total = L;
    sp.lv_begin = g0;
    sp.lv_end = g0 + lg;
    sp.w = pl.w;
    sp.beta = pl.beta;
    for (int d = 0; d < 3; ++d) {
      sp.bin_size[d] = pl.bin[d];
      sp.nbins[d] = pl.nbins[d];
      sp.nback[d] = pl.nback[d];
      sp.nf1[d] = pl.nf1[d];
      sp.nf2[d] = pl.nf2[d];
    }
    sp.bin_start = bstart;
    sp.bin_count = bcount;
    sp.pts = srt;
    sp.vals = svals;
    sp.xcd_chunk = cfg.spread_chunk;
    sp.fx = ddec;
    sp.fy = ddec + pl.nf1[0];
    sp.fz = ddec + pl.nf1[0] + pl.nf1[1];
    sp.grid = cg;
    sp.grid_stride = pl.grid_cells;
    // BIPP_N3_PROFILE=1: per-block duration / work summary of the spread on stderr
    const long long nb = pl.n_bins();
    const bool prof = std::getenv("BIPP_N3_PROFILE") != nullptr;
    sp.prof = prof ? prof_buf.as<unsigned long long>((size_t)nb * 4) : nullptr;
    if (prof) BIPP_HIP(hipMemsetAsync(sp.prof, 0, sizeof(unsigned long long) * nb * 4, st));
    launch_n3_spread<TG, cplx<TG>>(sp, st);
    if (prof) report_spread_profile(sp.prof, nb, pl, st);
  }

  // pruned FFT (sign +) along x, y, z of lg level grids
  void fft_group(const N3Plan& pl, int lg, cplx<TG>* cg, const N3Tuning& cfg, hipStream_t st) {
    const long long G = pl.grid_cells;
    const cplx<TG>* dtw = static_cast<const cplx<TG>*>(twid.p);
    const int nx = pl.nf2[0], ny = pl.nf2[1], nz = pl.nf2[2];
    // x: lines (y, z) of the input window; read x of the input window, write the output window
    N3LineParams<TG> lp;
    lp.grid = cg;
    lp.grid_stride = G;
    lp.n_levels = lg;
    lp.n = nx;
    lp.ny = ny;
    lp.nz = nz;
    lp.yw = pl.in_w[1];
    lp.zw = pl.in_w[2];
    lp.in = pl.in_w[0];
    lp.out = pl.out_w[0];
    lp.rad = radices_of(nx, cfg.lines_radix);
    lp.tw = dtw + tw_off[0][0];
    launch_fft_lines<TG>(lp, st);
    // y over (x of the output window, z of the input window); z over (x, y output windows)
    for (int d = 1; d < 3; ++d) {
      N3PassParams<TG> fp{};
      fp.grid = cg;
      fp.grid_stride = G;
      fp.n_levels = lg;
      fp.nx = nx;
      fp.xw = pl.out_w[0];
      fp.tw_shift_max = cfg.pass_tw_shift;
      auto rr = [&](int m) {
        int a = 0, b = 0;
        return cfg.pass_rr && fft_rr_factors(m, a, b) ? a : 0;
      };
      if (d == 1) {
        fp.no = nz;
        fp.other = pl.in_w[2];
        fp.operm_n1 = 1;
        fp.operm_n2 = 1;
        fp.outer_stride = (long long)nx * ny;
        fp.axis_stride = nx;
      } else {
        fp.no = ny;
        fp.other = pl.out_w[1];
        fp.operm_n1 = pl.n1[1];
        fp.operm_n2 = pl.n2[1];
        fp.outer_stride = nx;
        fp.axis_stride = (long long)nx * ny;
      }
      fp.nax = pl.nf2[d];
      fp.in = pl.in_w[d];
      fp.out = pl.out_w[d];
      const int n1 = pl.n1[d], n2 = pl.n2[d];
      if (n2 == 1) {
        fp.m = n1;
        fp.n_sub = 1;
        fp.sub_stride = 0;
        fp.estride = 1;
        fp.in_check = 1;
        fp.out_check = 1;
        fp.fa = 0;
        fp.fb = 1;
        fp.twiddle = 0;
        fp.ntot = n1;
        fp.rad = radices_of(n1, cfg.pass_radix);
        fp.rr_m1 = rr(n1);
        fp.tw = dtw + tw_off[d][0];
        fp.tw2 = nullptr;
        launch_fft_pass<TG>(fp, st);
      } else {
        // n = n1 n2: transforms of length n1 over stride n2 (inputs pruned), twiddle, then
        // transforms of length n2 over contiguous runs (outputs pruned); frequency k is
        // stored at (k mod n1) n2 + k / n1
        fp.m = n1;
        fp.n_sub = n2;
        fp.sub_stride = 1;
        fp.estride = n2;
        fp.in_check = 1;
        fp.out_check = 0;
        fp.twiddle = 1;
        fp.ntot = (long long)n1 * n2;
        fp.rad = radices_of(n1, cfg.pass_radix);
        fp.rr_m1 = rr(n1);
        fp.tw = dtw + tw_off[d][0];
        fp.tw2 = dtw + tw_off[d][2];
        launch_fft_pass<TG>(fp, st);
        fp.m = n2;
        fp.n_sub = n1;
        fp.sub_stride = n2;
        fp.estride = 1;
        fp.in_check = 0;
        fp.out_check = 1;
        fp.fa = 1;
        fp.fb = n1;
        fp.twiddle = 0;
        fp.rad = radices_of(n2, cfg.pass_radix);
        fp.rr_m1 = rr(n2);
        fp.tw = dtw + tw_off[d][1];
        fp.tw2 = nullptr;
        launch_fft_pass<TG>(fp, st);
      }
    }
  }

  // interpolation, correction, real part into the images of levels [g0, g0 + lg)
  void interp_group(const N3Plan& pl, int g0, int lg, const cplx<TG>* cg, const double* lmn64,
                    int n_pix, void* img, hipStream_t st) {
    N3InterpParams<TG> ip;
    ip.n_pix = n_pix;
    ip.n_levels = lg;
    ip.w1 = pl.w;
    ip.w2 = pl.w;
    ip.beta1 = pl.beta;
    ip.beta2 = pl.beta;
    for (int d = 0; d < 3; ++d) {
      ip.gam[d] = pl.gam[d];
      ip.D[d] = pl.D[d];
      ip.C[d] = pl.C[d];
      ip.h1[d] = pl.h1[d];
      ip.h2[d] = pl.h2[d];
      ip.nf2[d] = pl.nf2[d];
      ip.n1[d] = pl.n1[d];
      ip.n2[d] = pl.n2[d];
    }
    ip.lmn = lmn64;
    ip.ft = static_cast<const double*>(gl.p);
    ip.n_ft = static_cast<int>(pl.ft_tab.size());
    ip.ft_inv_delta = pl.ft_inv_delta;
    ip.grid = cg;
    ip.grid_stride = pl.grid_cells;
    ip.img = static_cast<TG*>(img) + (size_t)g0 * n_pix;
    ip.scale = 1.0;
    ip.img_ld = n_pix;
    launch_n3_interp<TG, TG>(ip, st);
  }

  // Run-time shapes for this grid shape: from the process-wide table of tuned shapes, or,
  // with autotuning on and no entry yet, by timing every candidate (spread + FFT of one level
  // group with HIP events) and keeping the fastest.
  N3Tuning tuning_for(const N3Plan& pl, int L, ContextImpl& c, const unsigned* bstart,
                      const unsigned* bcount, const N3Point* srt, const cplx<TG>* svals,
                      cplx<TG>* cg, int lg) {
    const std::string key = n3_tune_key(pl, L, sizeof(TG));
    N3Tuning found;
    if (n3_tune_lookup(key, found)) return found;
    if (!autotune) return N3Tuning{};
    hipStream_t st = c.stream;
    const std::vector<N3Tuning> cands = n3_tune_candidates();
    std::vector<hipEvent_t> ev(2 * cands.size());
    for (auto& e : ev) BIPP_HIP(hipEventCreate(&e));
    for (size_t k = 0; k < cands.size(); ++k) {
      BIPP_HIP(hipEventRecord(ev[2 * k], st));
      spread_group(pl, L, 0, lg, bstart, bcount, srt, svals, cg, cands[k], st);
      fft_group(pl, lg, cg, cands[k], st);
      BIPP_HIP(hipEventRecord(ev[2 * k + 1], st));
    }
    c.sync();
    size_t best = 0;
    float best_ms = 0;
    tune_log.clear();
    for (size_t k = 0; k < cands.size(); ++k) {
      float t = 0;
      BIPP_HIP(hipEventElapsedTime(&t, ev[2 * k], ev[2 * k + 1]));
      tune_log.push_back({cands[k], t});
      if (k == 0 || t < best_ms) {
        best = k;
        best_ms = t;
      }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    n3_tune_store(key, cands[best]);
    return cands[best];
  }
};

}  // namespace bippamd
