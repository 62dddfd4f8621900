// Host driver and C-ABI of libbipp_amd: contexts, the batched eigensolve and the fused
// synthesis pipeline (stage A -> level weights -> image accumulation) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "bipp_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "nufft3.hpp"

namespace bippamd {

namespace {
thread_local std::string g_last_error;

template <typename F>
BippError guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return BIPP_SUCCESS;
  } catch (const BippException& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return BIPP_GENERIC_ERROR;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return BIPP_GENERIC_ERROR;
  } catch (...) {
    g_last_error = "unknown error";
    return BIPP_GENERIC_ERROR;
  }
}

void require(bool cond, BippError code, const char* msg) {
  if (!cond) fail(code, msg);
}
void require(bool cond, BippError code, const std::string& msg) {
  if (!cond) fail(code, msg);
}

// Growable device allocation.  Reallocation frees the old block; hipFree synchronises the
// device, so work still in flight on it has finished first.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void* get(size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > cap) {
      if (p) BIPP_HIP(hipFree(p));
      p = nullptr;
      cap = 0;
      BIPP_HIP(hipMalloc(&p, bytes));
      cap = bytes;
    }
    return p;
  }
  template <typename T>
  T* as(size_t n) {
    return static_cast<T*>(get(n * sizeof(T)));
  }
};

struct PinnedBuf {
  void* p = nullptr;
  size_t cap = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void* get(size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    if (bytes > cap) {
      if (p) BIPP_HIP(hipHostFree(p));
      p = nullptr;
      cap = 0;
      BIPP_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
      cap = bytes;
    }
    return p;
  }
};

}  // namespace

// ---------------------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------------------
struct ContextImpl {
  BippProcessingUnit pu = BIPP_PU_GPU;
  int device = 0;
  int n_cu = 1;
  hipStream_t stream = nullptr;

  explicit ContextImpl(BippProcessingUnit p) {
    require(p != BIPP_PU_CPU, BIPP_INVALID_PARAMETER_ERROR,
            "this library computes on the GPU only (processing unit CPU requested)");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
      fail(BIPP_GPU_SUPPORT_ERROR, "no GPU device available");
    BIPP_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    BIPP_HIP(hipGetDeviceProperties(&prop, device));
    n_cu = prop.multiProcessorCount;
    BIPP_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    BIPP_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    pu = BIPP_PU_GPU;
  }
  ~ContextImpl() {
    if (ready) (void)hipEventDestroy(ready);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // Waits for the stream by polling an event: a blocking wait that spans a long kernel wakes
  // up about a millisecond late, which leaves the device idle between batches.
  void sync() {
    BIPP_HIP(hipEventRecord(ready, stream));
    for (;;) {
      const hipError_t q = hipEventQuery(ready);
      if (q == hipSuccess) break;
      if (q != hipErrorNotReady) BIPP_HIP(q);
    }
  }

 private:
  hipEvent_t ready = nullptr;
};

// ---------------------------------------------------------------------------------------
// Stage A on a batch: flag -> [Gram/W congruence] -> generalized eigensolve.
// Results stay on the device: d (R, stride nmax), z (reduced eigenvectors, ld nmax) and,
// for a general W, v_full = W_I Z (nAntenna x nmax).
// ---------------------------------------------------------------------------------------
template <typename R>
struct EigRunner {
  DevBuf n_red, idx, n_vis, info, counter, order, d, z, ws_a, ws_l, ws_z, prof_buf;
  DevBuf sig, w_i, tmp, g, v_full;
  int n_samples = 0, n_ant = 0, nmax = 0;
  bool general_w = false;
  std::vector<int> h_info;

  void run(ContextImpl& ctx, int nS, int nAnt, int nBeam, const double* wl, const void* s,
           bool s_double, long long lds, long long stride_s, const void* w, bool w_double,
           long long ldw, long long stride_w, const double* xyz, long long ldxyz,
           long long stride_xyz, bool standard, double prefactor) {
    require(nBeam >= 1 && nBeam <= 512, BIPP_INVALID_PARAMETER_ERROR,
            "number of beams must be in [1, 512]");
    require(nAnt >= 1, BIPP_INVALID_PARAMETER_ERROR, "number of antennas must be positive");
    require(w != nullptr || nAnt == nBeam, BIPP_INVALID_PARAMETER_ERROR,
            "identity beamforming needs nAntenna == nBeam");
    require(lds >= nBeam, BIPP_INVALID_PARAMETER_ERROR, "lds < nBeam");
    n_samples = nS;
    n_ant = nAnt;
    nmax = nBeam;
    general_w = w != nullptr;
    hipStream_t st = ctx.stream;
    const long long nn = (long long)nmax * nmax;

    FlagParams fp;
    fp.n_beam = nBeam;
    fp.mode = standard ? 1 : 0;
    fp.s = s;
    fp.lds = lds;
    fp.stride_s = stride_s;
    fp.n_red = n_red.as<int>(nS);
    fp.idx = idx.as<int>((size_t)nS * nmax);
    fp.idx_stride = nmax;
    fp.n_vis = n_vis.as<unsigned long long>(nS);
    launch_flag(fp, nS, s_double, st);

    const cplx<R>* g_in = nullptr;
    cplx<R>* wi = nullptr;
    if (general_w) {
      // W_I = W[:, kept beams] (src/eigensolver.cpp:201-203)
      wi = w_i.as<cplx<R>>((size_t)nS * nAnt * nmax);
      if (w_double)
        launch_gather_w<R, double>(nS, nAnt, nBeam, w, ldw, stride_w, fp.n_red, fp.idx, nmax, wi,
                                   nAnt, (long long)nAnt * nmax, st);
      else
        launch_gather_w<R, float>(nS, nAnt, nBeam, w, ldw, stride_w, fp.n_red, fp.idx, nmax, wi,
                                  nAnt, (long long)nAnt * nmax, st);
    }
    if (general_w && !standard) {
      // G_r = W_I^H Sigma W_I (src/host/gram_matrix.cpp:58-61)
      cplx<R>* sg = sig.as<cplx<R>>((size_t)nS * nAnt * nAnt);
      launch_sinc_matrix<R>(nS, nAnt, xyz, ldxyz, stride_xyz, wl, prefactor, sg, nAnt,
                            (long long)nAnt * nAnt, st);
      cplx<R>* t = tmp.as<cplx<R>>((size_t)nS * nAnt * nmax);
      GemmBatch<R> g1{};
      g1.n_batch = nS;
      g1.m = nullptr;
      g1.m0 = nAnt;
      g1.n = fp.n_red;
      g1.k = nullptr;
      g1.k0 = nAnt;
      g1.op_a = 0;
      g1.op_b = 0;
      g1.alpha = mk<R>(1, 0);
      g1.beta = mk<R>(0, 0);
      g1.a = sg;
      g1.lda = nAnt;
      g1.stride_a = (long long)nAnt * nAnt;
      g1.b = wi;
      g1.ldb = nAnt;
      g1.stride_b = (long long)nAnt * nmax;
      g1.c = t;
      g1.ldc = nAnt;
      g1.stride_c = (long long)nAnt * nmax;
      launch_gemm_batch<R>(g1, nAnt, nmax, st);
      cplx<R>* gr = g.as<cplx<R>>((size_t)nS * nn);
      GemmBatch<R> g2 = g1;
      g2.m = fp.n_red;
      g2.n = fp.n_red;
      g2.k = nullptr;
      g2.k0 = nAnt;
      g2.op_a = 1;
      g2.a = wi;
      g2.lda = nAnt;
      g2.stride_a = (long long)nAnt * nmax;
      g2.b = t;
      g2.ldb = nAnt;
      g2.stride_b = (long long)nAnt * nmax;
      g2.c = gr;
      g2.ldc = nmax;
      g2.stride_c = nn;
      launch_gemm_batch<R>(g2, nmax, nmax, st);
      g_in = gr;
    }

    // persistent launch, largest problems first
    std::vector<int> h_nred(nS);
    BIPP_HIP(hipMemcpyAsync(h_nred.data(), fp.n_red, sizeof(int) * nS, hipMemcpyDeviceToHost, st));
    ctx.sync();
    std::vector<int> h_order(nS);
    std::iota(h_order.begin(), h_order.end(), 0);
    std::stable_sort(h_order.begin(), h_order.end(),
                     [&](int a, int b) { return h_nred[a] > h_nred[b]; });
    int* d_order = order.as<int>(nS);
    BIPP_HIP(hipMemcpyAsync(d_order, h_order.data(), sizeof(int) * nS, hipMemcpyHostToDevice, st));

    const int per_cu = eig_blocks_per_cu<R>(nmax, s_double);
    const int n_slots = std::max(1, std::min(nS, per_cu * ctx.n_cu));
    EigParams<R> ep;
    ep.n_jobs = nS;
    ep.nmax = nmax;
    ep.order = d_order;
    ep.counter = counter.as<unsigned>(1);
    launch_fill_u32(ep.counter, 0u, 1, st);
    ep.n_red = fp.n_red;
    ep.idx = fp.idx;
    ep.idx_stride = nmax;
    ep.s = s;
    ep.lds = lds;
    ep.stride_s = stride_s;
    ep.xyz = xyz;
    ep.ldxyz = ldxyz;
    ep.stride_xyz = stride_xyz;
    ep.wl = wl;
    ep.g_in = g_in;
    ep.stride_g = nn;
    ep.ld_g = nmax;
    ep.gram_prefactor = prefactor;
    ep.standard = standard ? 1 : 0;
    ep.d_out = d.as<R>((size_t)nS * nmax);
    ep.stride_d = nmax;
    ep.z_out = z.as<cplx<R>>((size_t)nS * nn);
    ep.stride_z = nn;
    ep.ldz = nmax;
    ep.ws_ld = nmax;
    ep.ws_stride = nn;
    ep.ws_a = ws_a.as<cplx<R>>((size_t)n_slots * nn);
    ep.ws_l = ws_l.as<cplx<R>>((size_t)n_slots * nn);
    ep.ws_z = ws_z.as<R>((size_t)n_slots * nn);
    ep.info = info.as<int>(nS);
    // BIPP_EIG_PROFILE=1: per-phase wall-clock averages of this launch on stderr
    const bool prof = std::getenv("BIPP_EIG_PROFILE") != nullptr;
    if (prof) {
      ep.prof = prof_buf.as<unsigned long long>((size_t)nS * 16);
      BIPP_HIP(hipMemsetAsync(ep.prof, 0, sizeof(unsigned long long) * nS * 16, st));
    }
    launch_eig<R>(ep, n_slots, s_double, st);
    if (prof) {
      std::vector<unsigned long long> h((size_t)nS * 16);
      BIPP_HIP(hipMemcpyAsync(h.data(), ep.prof, sizeof(unsigned long long) * h.size(),
                              hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
      static const char* names[7] = {"load", "cholesky", "whiten", "tridiag", "ql",
                                     "backtransform", "unwhiten"};
      double acc[7] = {0, 0, 0, 0, 0, 0, 0}, gen = 0, sweeps = 0, rots = 0;
      int cnt = 0;
      for (int k = 0; k < nS; ++k) {
        const unsigned long long* r = &h[(size_t)k * 16];
        if (!r[0] || !r[7]) continue;
        ++cnt;
        for (int q = 0; q < 7; ++q) acc[q] += r[q + 1] >= r[q] && r[q] ? double(r[q + 1] - r[q]) : 0.0;
        gen += double(r[8]);
        sweeps += double(r[9]);
        rots += double(r[10]);
      }
      std::fprintf(stderr, "[bipp eig profile] %d samples, n_slots %d, us/sample (100 MHz clock):", cnt,
                   n_slots);
      for (int q = 0; q < 7; ++q) std::fprintf(stderr, " %s %.1f", names[q], cnt ? acc[q] / cnt / 100.0 : 0.0);
      // time inside the rotation generator, and QL sweeps and rotations per sample
      if (cnt)
        std::fprintf(stderr, " | ql generator %.1f, sweeps %.0f, rotations %.0f", gen / cnt / 100.0,
                     sweeps / cnt, rots / cnt);
      std::fprintf(stderr, "\n");
    }

    if (general_w) {
      // V = W_I Z (src/eigensolver.cpp:214-216)
      cplx<R>* vf = v_full.as<cplx<R>>((size_t)nS * nAnt * nmax);
      BIPP_HIP(hipMemsetAsync(vf, 0, sizeof(cplx<R>) * (size_t)nS * nAnt * nmax, st));
      GemmBatch<R> g3{};
      g3.n_batch = nS;
      g3.m = nullptr;
      g3.m0 = nAnt;
      g3.n = fp.n_red;
      g3.k = fp.n_red;
      g3.op_a = 0;
      g3.op_b = 0;
      g3.alpha = mk<R>(1, 0);
      g3.beta = mk<R>(0, 0);
      g3.a = static_cast<const cplx<R>*>(w_i.p);
      g3.lda = nAnt;
      g3.stride_a = (long long)nAnt * nmax;
      g3.b = ep.z_out;
      g3.ldb = nmax;
      g3.stride_b = nn;
      g3.c = vf;
      g3.ldc = nAnt;
      g3.stride_c = (long long)nAnt * nmax;
      launch_gemm_batch<R>(g3, nAnt, nmax, st);
    }
  }

  // Throws EigensolverError if any solve failed (LAPACK info != 0 analogue).
  void check(ContextImpl& ctx) {
    h_info.resize(n_samples);
    BIPP_HIP(hipMemcpyAsync(h_info.data(), info.p, sizeof(int) * n_samples, hipMemcpyDeviceToHost,
                            ctx.stream));
    ctx.sync();
    for (int k = 0; k < n_samples; ++k)
      if (h_info[k] != 0)
        fail(BIPP_EIGENSOLVER_ERROR, "generalized eigensolve failed for sample " +
                                         std::to_string(k) + " (info " +
                                         std::to_string(h_info[k]) + ")");
  }

  const int* d_n_red() const { return static_cast<const int*>(n_red.p); }
  const int* d_idx() const { return static_cast<const int*>(idx.p); }
  const unsigned long long* d_n_vis() const {
    return static_cast<const unsigned long long*>(n_vis.p);
  }
  const R* d_d() const { return static_cast<const R*>(d.p); }
  const cplx<R>* d_z() const { return static_cast<const cplx<R>*>(z.p); }
  const cplx<R>* d_vfull() const { return static_cast<const cplx<R>*>(v_full.p); }
};

// Writes d, V (W = I scatter or W_I Z), scale and kept counts in the caller's layout.
template <typename R>
void export_eig(ContextImpl& ctx, EigRunner<R>& er, int nS, int nAnt, int nBeam, bool out_double,
                void* d, long long stride_d, void* v, long long ldv, long long stride_v,
                int32_t* n_red_out, void* scale) {
  hipStream_t st = ctx.stream;
  launch_cast_d<R>(nS, nBeam, er.d_d(), nBeam, d, stride_d, out_double, er.d_n_vis(), scale, st);
  if (v) {
    if (!er.general_w) {
      UnbeamParams<R> up;
      up.n_ant = nAnt;
      up.n_red = er.d_n_red();
      up.idx = er.d_idx();
      up.idx_stride = nBeam;
      up.z = er.d_z();
      up.stride_z = (long long)nBeam * nBeam;
      up.ldz = nBeam;
      up.v = v;
      up.stride_v = stride_v;
      up.ldv = ldv;
      launch_unbeam_identity<R>(up, nS, nBeam, out_double, st);
    } else if (out_double) {
      launch_copy_cast<R, double>(nS, nAnt, nBeam, er.d_vfull(), nAnt, (long long)nAnt * nBeam, v,
                                  ldv, stride_v, st);
    } else {
      launch_copy_cast<R, float>(nS, nAnt, nBeam, er.d_vfull(), nAnt, (long long)nAnt * nBeam, v,
                                 ldv, stride_v, st);
    }
  }
  if (n_red_out)
    BIPP_HIP(hipMemcpyAsync(n_red_out, er.d_n_red(), sizeof(int) * nS, hipMemcpyDeviceToDevice, st));
}

}  // namespace bippamd

using namespace bippamd;

// A context is shared by the synthesis handles created from it: each handle holds a
// reference, so destroying the context first leaves it alive until its last handle goes.
struct BippContextImpl {
  std::unique_ptr<ContextImpl> impl;
  EigRunner<float> eig32;
  EigRunner<double> eig64;
  std::atomic<int> refs{1};

  void retain() { refs.fetch_add(1); }
  void release() {
    if (refs.fetch_sub(1) == 1) {
      (void)hipStreamSynchronize(impl->stream);
      delete this;
    }
  }
};

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

// n = n1 * n2 with n1, n2 <= 128 (n2 == 1 when n <= 128 itself)
void split_axis(int n, int& n1, int& n2) {
  if (n <= 128) {
    n1 = n;
    n2 = 1;
    return;
  }
  for (int a = 128; a >= 2; --a)
    if (n % a == 0 && n / a <= 128) {
      n1 = a;
      n2 = n / a;
      return;
    }
  fail(BIPP_INVALID_PARAMETER_ERROR,
       "NUFFT grid axis of " + std::to_string(n) + " points exceeds the single-plan limit (16384)");
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
            const double D_in[3], int max_x) {
    w = static_cast<int>(std::ceil(std::log10(1.0 / tol))) + 1;
    w = std::max(2, std::min(w, 16));
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
      const double X = std::max(X_in[d], 1e-9);
      nf1[d] = std::max(next_smooth(2 * sigma / 3.14159265358979323846 * X * S + w + 1), 2 * w);
      nf2[d] = next_smooth(2.0 * nf1[d]);
      h1[d] = 2 * 3.14159265358979323846 / nf1[d];
      h2[d] = 2 * 3.14159265358979323846 / nf2[d];
      gam[d] = nf1[d] / (2 * sigma * S);
      C[d] = C_in[d];
      D[d] = D_in[d];
      if (d == 0) {
        if (nf2[0] > max_x)
          fail(BIPP_INVALID_PARAMETER_ERROR, "NUFFT grid x axis of " + std::to_string(nf2[0]) +
                                                 " points exceeds the single-plan limit");
        n1[0] = nf2[0];
        n2[0] = 1;
      } else {
        split_axis(nf2[d], n1[d], n2[d]);
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

// Runs the NUFFT image stage for one batch: points from positions, values from
// VV_l = Z diag(w_l) Z^H, all levels accumulated into `img` (TI per pixel).
template <typename TG>
struct N3Runner {
  DevBuf grid, valid, bin_of, pts, sorted;
  DevBuf bin_count, bin_start, bin_fill, scan_tmp, vv, deconv, gl, twid, prof_buf;
  std::vector<cplx<TG>> htw;
  std::vector<TG> hdec;

  template <typename RE>
  void run(ContextImpl& c, const N3Plan& pl, int nS, int L, int n_cols, int rows_max,
           const int* n_rows, const cplx<RE>* z, long long stride_z, int ldz, const RE* wd,
           const double* pos, long long pos_stride, int apad, const double* lmn64, int n_pix,
           void* img, bool phase_fp32, hipEvent_t* evs) {
    hipStream_t st = c.stream;
    auto mark = [&](int i) {
      if (evs) BIPP_HIP(hipEventRecord(evs[i], st));
    };
    mark(0);
    // virtual visibilities of every (sample, level), lower triangle
    const long long ldv = rows_max;
    const long long vv_stride = ldv * ldv;
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
    vp.ld_vv = ldv;
    launch_vv_levels<RE, TG>(vp, nS, rows_max, st);
    const bool merge_auto = rows_max >= 2;
    if (merge_auto) launch_vv_trace<TG>(nS, L, n_rows, rows_max, vvp, vv_stride, ldv, st);
    mark(1);

    // points, bins, sort
    const long long npair = (long long)rows_max * (rows_max + 1) / 2;
    const long long npts = npair * nS;
    if (npts > 0xffffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT points in one batch");
    N3PointsParams pp;
    pp.n_ant_max = rows_max;
    pp.n_rows = n_rows;
    pp.n_rows_const = rows_max;
    pp.pos = pos;
    pp.pos_stride = pos_stride;
    pp.apad = apad;
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
    pp.ld_vv = ldv;
    pp.merge_auto = merge_auto ? 1 : 0;
    pp.valid = valid.as<unsigned char>(npts);
    pp.bin_of = bin_of.as<int>(npts);
    pp.pts = pts.as<N3Point>(npts);
    const long long nb = pl.n_bins();
    if (nb > 0x7fffffffLL) fail(BIPP_INVALID_PARAMETER_ERROR, "too many NUFFT bins");
    pp.bin_count = bin_count.as<unsigned>(nb);
    launch_zero<unsigned>(pp.bin_count, nb, st);
    launch_n3_points<double>(pp, nS, st);
    unsigned* bstart = bin_start.as<unsigned>(nb);
    unsigned* btmp = scan_tmp.as<unsigned>(3 * ((size_t)nb / 1024 + 2) + 4096);
    launch_scan(pp.bin_count, bstart, btmp, (int)nb, st);
    unsigned* bfill = bin_fill.as<unsigned>(nb);
    launch_zero<unsigned>(bfill, nb, st);
    N3Point* srt = sorted.as<N3Point>(npts);
    launch_n3_scatter(npts, pp.valid, pp.bin_of, pp.pts, bstart, bfill, srt, st);
    mark(2);

    // spread + deconvolution into [-nf1/2, nf1/2)^3 of the nf2 grid
    const long long G = pl.grid_cells;
    cplx<TG>* cg = grid.as<cplx<TG>>((size_t)L * G);
    const int nd = pl.nf1[0] + pl.nf1[1] + pl.nf1[2];
    hdec.resize(nd);
    for (int i = 0; i < pl.nf1[0]; ++i) hdec[i] = static_cast<TG>(pl.deconv[0][i]);
    for (int i = 0; i < pl.nf1[1]; ++i) hdec[pl.nf1[0] + i] = static_cast<TG>(pl.deconv[1][i]);
    for (int i = 0; i < pl.nf1[2]; ++i)
      hdec[pl.nf1[0] + pl.nf1[1] + i] = static_cast<TG>(pl.deconv[2][i]);
    TG* ddec = deconv.as<TG>(nd);
    BIPP_HIP(hipMemcpyAsync(ddec, hdec.data(), sizeof(TG) * nd, hipMemcpyHostToDevice, st));
    N3SpreadParams<TG> sp;
    sp.n_levels = L;
    sp.lv0 = 0;
    sp.n_levels_total = L;
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
    sp.bin_count = pp.bin_count;
    sp.pts = srt;
    sp.vv = vvp;
    sp.vv_stride = vv_stride;
    sp.fx = ddec;
    sp.fy = ddec + pl.nf1[0];
    sp.fz = ddec + pl.nf1[0] + pl.nf1[1];
    sp.grid = cg;
    sp.grid_stride = G;
    // BIPP_N3_PROFILE=1: per-block duration / work summary of the spread on stderr
    const bool prof = std::getenv("BIPP_N3_PROFILE") != nullptr;
    sp.prof = prof ? prof_buf.as<unsigned long long>((size_t)nb * 4) : nullptr;
    if (prof) BIPP_HIP(hipMemsetAsync(sp.prof, 0, sizeof(unsigned long long) * nb * 4, st));
    launch_n3_spread<TG, cplx<TG>>(sp, st);
    if (prof) report_spread_profile(sp.prof, nb, pl, st);
    mark(3);

    // pruned FFT (sign +) along x, y, z; twiddle tables exp(+2 pi i t / n) computed in FP64
    const int nx = pl.nf2[0], ny = pl.nf2[1], nz = pl.nf2[2];
    // table segments: x, then per strided axis n1, n2 and n1 * n2
    size_t tw_off[3][3] = {};
    htw.clear();
    // two-level tables: exp(+2 pi i t / n) = fine[t mod 64] * coarse[t / 64]
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
    tw_off[0][0] = add_table(nx);
    for (int d = 1; d < 3; ++d) {
      tw_off[d][0] = add_table(pl.n1[d]);
      tw_off[d][1] = add_table(pl.n2[d]);
      tw_off[d][2] = pl.n2[d] > 1 ? add_table((long long)pl.n1[d] * pl.n2[d]) : 0;
    }
    cplx<TG>* dtw = twid.as<cplx<TG>>(htw.size());
    BIPP_HIP(hipMemcpyAsync(dtw, htw.data(), sizeof(cplx<TG>) * htw.size(),
                            hipMemcpyHostToDevice, st));
    // x: lines (y, z) of the input window; read x of the input window, write the output window
    N3LineParams<TG> lp;
    lp.grid = cg;
    lp.grid_stride = G;
    lp.n_levels = L;
    lp.n = nx;
    lp.ny = ny;
    lp.nz = nz;
    lp.yw = pl.in_w[1];
    lp.zw = pl.in_w[2];
    lp.in = pl.in_w[0];
    lp.out = pl.out_w[0];
    lp.rad = radices_of(nx, fft_radix_max());
    lp.tw = dtw + tw_off[0][0];
    launch_fft_lines<TG>(lp, st);
    // y over (x of the output window, z of the input window); z over (x, y output windows)
    for (int d = 1; d < 3; ++d) {
      N3PassParams<TG> fp{};
      fp.grid = cg;
      fp.grid_stride = G;
      fp.n_levels = L;
      fp.nx = nx;
      fp.xw = pl.out_w[0];
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
        fp.rad = radices_of(n1, fft_pass_radix_max());
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
        fp.rad = radices_of(n1, fft_pass_radix_max());
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
        fp.rad = radices_of(n2, fft_pass_radix_max());
        fp.tw = dtw + tw_off[d][1];
        fp.tw2 = nullptr;
        launch_fft_pass<TG>(fp, st);
      }
    }

    mark(4);
    // interpolation, correction, real part into the images
    double* dft = gl.as<double>(pl.ft_tab.size());
    BIPP_HIP(hipMemcpyAsync(dft, pl.ft_tab.data(), sizeof(double) * pl.ft_tab.size(),
                            hipMemcpyHostToDevice, st));
    N3InterpParams<TG> ip;
    ip.n_pix = n_pix;
    ip.n_levels = L;
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
    ip.ft = dft;
    ip.n_ft = static_cast<int>(pl.ft_tab.size());
    ip.ft_inv_delta = pl.ft_inv_delta;
    ip.grid = cg;
    ip.grid_stride = G;
    ip.img = img;
    ip.scale = 1.0;
    ip.img_ld = n_pix;
    launch_n3_interp<TG, TG>(ip, st);
    mark(5);
  }
};

}  // namespace bippamd

// ---------------------------------------------------------------------------------------
// Fused synthesis pipeline
// ---------------------------------------------------------------------------------------
struct BippSynthesisImpl {
  BippContextImpl* ctx = nullptr;
  BippCompute compute = BIPP_COMPUTE_MIXED;
  int imager = 0;
  int n_ant = 0, n_beam = 0, n_levels = 0, n_pix = 0;
  BippFilter filter = BIPP_FILTER_LSQ;
  bool batched = true;
  size_t n_collected = 0;
  DevBuf intervals, lmn64, lmn32, img;
  // steering buffers
  DevBuf n_eig, n_ant_used, pos, afrag, wts;
  // direct-sum buffers
  DevBuf wdense, vfull, vv, vv32, uvw;
  // type-3 NUFFT state
  double tol = 1e-5;
  double pix_lo[3] = {0, 0, 0}, pix_hi[3] = {0, 0, 0};
  N3Plan n3plan;
  N3Runner<float> n3f;
  N3Runner<double> n3d;
  DevBuf ranges, dummy;
  int last_nf2[3] = {0, 0, 0};
  // staging for host collects
  DevBuf st_s, st_w, st_xyz, st_wl;
  std::vector<double> last_d;
  std::vector<int32_t> last_nred;
  int last_count = 0;
  // stage timings on the context stream: 0 eigensolve, 1 prep/weights, 2 image, 3 launches
  // NUFFT sub-stages: 4 virtual vis, 5 point sort, 6 spread, 7 deconvolve + FFT, 8 interpolate
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t ev_n3[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double* h_rg = nullptr;         // pinned per-sample position ranges
  size_t h_rg_cap = 0;
  static constexpr int N_TIMES = 9;
  double stage_ms[N_TIMES] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool n3_timed = false;

  void init_events() {
    for (auto& e : ev) BIPP_HIP(hipEventCreate(&e));
    for (auto& e : ev_n3) BIPP_HIP(hipEventCreate(&e));
  }
  ~BippSynthesisImpl() {
    if (h_rg) (void)hipHostFree(h_rg);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : ev_n3)
      if (e) (void)hipEventDestroy(e);
  }
  void accumulate_times() {
    float t = 0;
    BIPP_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
    stage_ms[0] += t;
    BIPP_HIP(hipEventElapsedTime(&t, ev[1], ev[2]));
    stage_ms[1] += t;
    BIPP_HIP(hipEventElapsedTime(&t, ev[2], ev[3]));
    stage_ms[2] += t;
    stage_ms[3] += 1;
    if (n3_timed) {
      for (int i = 0; i < 5; ++i) {
        BIPP_HIP(hipEventElapsedTime(&t, ev_n3[i], ev_n3[i + 1]));
        stage_ms[4 + i] += t;
      }
      n3_timed = false;
    }
  }

  bool img_double() const { return compute == BIPP_COMPUTE_FP64; }
  int steer_mode() const {
    return compute == BIPP_COMPUTE_FP64 ? STEER_FP64
                                        : (compute == BIPP_COMPUTE_FP32 ? STEER_FP32 : STEER_BF16X3);
  }

  template <typename R>
  void collect_impl(EigRunner<R>& er, int nS, const double* wl, bool s_double, const void* s,
                    long long lds, long long stride_s, const void* w, long long ldw,
                    long long stride_w, const double* xyz, long long ldxyz, long long stride_xyz) {
    ContextImpl& c = *ctx->impl;
    hipStream_t st = c.stream;
    BIPP_HIP(hipEventRecord(ev[0], st));
    er.run(c, nS, n_ant, n_beam, wl, s, s_double, lds, stride_s, w, s_double, ldw, stride_w, xyz,
           ldxyz, stride_xyz, false, 1.0);
    BIPP_HIP(hipEventRecord(ev[1], st));
    // a failed solve raises before anything of this batch reaches the images
    er.check(c);
    const double* d_int = static_cast<const double*>(intervals.p);
    if (imager == 0) {
      const int mode = steer_mode();
      const int apad = ((n_ant + 15) / 16) * 16;
      const int kchunks = apad / 16;
      const int rtiles = (2 * n_beam + 15) / 16;
      const int epad = rtiles * 8;
      const size_t fe = steer_frag_elems_per_sample(mode, rtiles, kchunks);
      const size_t esz = mode == STEER_BF16X3 ? 2 : (mode == STEER_FP32 ? 4 : 8);
      SteerPrepParams<R> pp;
      pp.n_samples = nS;
      pp.n_levels = n_levels;
      pp.filter = filter;
      pp.intervals = d_int;
      pp.n_red = er.d_n_red();
      pp.idx = er.general_w ? nullptr : er.d_idx();
      pp.idx_stride = n_beam;
      pp.d = er.d_d();
      pp.stride_d = n_beam;
      pp.n_cols = n_beam;
      pp.n_vis = er.d_n_vis();
      if (er.general_w) {
        pp.z = er.d_vfull();
        pp.ldz = n_ant;
        pp.stride_z = (long long)n_ant * n_beam;
      } else {
        pp.z = er.d_z();
        pp.ldz = n_beam;
        pp.stride_z = (long long)n_beam * n_beam;
      }
      pp.xyz = xyz;
      pp.ldxyz = ldxyz;
      pp.stride_xyz = stride_xyz;
      pp.wl = wl;
      pp.n_eig = n_eig.as<int>(nS);
      pp.n_ant = n_ant_used.as<int>(nS);
      pp.pos_stride = 3LL * apad;
      pp.pos = pos.as<double>((size_t)nS * pp.pos_stride);
      pp.apad = apad;
      pp.afrag_stride = static_cast<long long>(fe);
      pp.afrag = afrag.get(nS * fe * esz);
      pp.kchunks = kchunks;
      pp.rtiles = rtiles;
      pp.w_stride = (long long)n_levels * epad;
      pp.wts = wts.get((size_t)nS * pp.w_stride * (mode == STEER_FP64 ? 8 : 4));
      pp.epad = epad;
      pp.mode = mode;
      // rows of V_full are all antennas; n_red then counts eigenvectors only
      pp.n_rows_const = er.general_w ? n_ant : 0;
      launch_steer_prep<R>(pp, st);
      BIPP_HIP(hipEventRecord(ev[2], st));
      SteerParams sp;
      sp.n_samples = nS;
      sp.n_levels = n_levels;
      sp.n_pix = n_pix;
      sp.n_eig = pp.n_eig;
      sp.n_ant = pp.n_ant;
      sp.pos = pp.pos;
      sp.pos_stride = pp.pos_stride;
      sp.apad = apad;
      sp.afrag = pp.afrag;
      sp.afrag_stride = pp.afrag_stride;
      sp.kchunks = kchunks;
      sp.rtiles = rtiles;
      sp.wts = pp.wts;
      sp.w_stride = pp.w_stride;
      sp.epad = epad;
      sp.lmn = static_cast<const double*>(lmn64.p);
      sp.phase_fp32 = compute == BIPP_COMPUTE_FP32 ? 1 : 0;
      sp.img = img.p;
      sp.img_ld = n_pix;
      launch_steer(sp, mode, st);
      BIPP_HIP(hipEventRecord(ev[3], st));
    } else if (imager == 2) {
      // virtual visibilities + type-3 NUFFT over all samples and levels of the batch
      R* wd = wdense.as<R>((size_t)nS * n_levels * n_beam);
      launch_level_weights<R>(nS, n_levels, filter, d_int, er.d_n_red(), er.d_d(), n_beam, n_beam,
                              er.d_n_vis(), wd, (long long)n_levels * n_beam, st);
      // positions / wavelength of the rows used (operand fragments not needed)
      const int apad = ((n_ant + 15) / 16) * 16;
      SteerPrepParams<R> pp{};
      pp.n_samples = nS;
      pp.n_levels = n_levels;
      pp.filter = filter;
      pp.intervals = d_int;
      pp.n_red = er.d_n_red();
      pp.idx = er.general_w ? nullptr : er.d_idx();
      pp.idx_stride = n_beam;
      pp.d = er.d_d();
      pp.stride_d = n_beam;
      pp.n_cols = n_beam;
      pp.n_vis = er.d_n_vis();
      pp.z = er.general_w ? er.d_vfull() : er.d_z();
      pp.ldz = er.general_w ? n_ant : n_beam;
      pp.stride_z = er.general_w ? (long long)n_ant * n_beam : (long long)n_beam * n_beam;
      pp.xyz = xyz;
      pp.ldxyz = ldxyz;
      pp.stride_xyz = stride_xyz;
      pp.wl = wl;
      pp.n_eig = n_eig.as<int>(nS);
      pp.n_ant = n_ant_used.as<int>(nS);
      pp.pos_stride = 3LL * apad;
      pp.pos = pos.as<double>((size_t)nS * pp.pos_stride);
      pp.apad = apad;
      pp.afrag = dummy.get(256);
      pp.afrag_stride = 0;
      pp.kchunks = 1;
      pp.rtiles = 0;
      pp.wts = dummy.get(256);
      pp.w_stride = 0;
      pp.epad = 0;
      pp.mode = steer_mode();
      pp.n_rows_const = er.general_w ? n_ant : 0;
      launch_steer_prep<R>(pp, st);
      // input-point extent from the per-sample position ranges
      double* rg = ranges.as<double>(6 * (size_t)nS);
      launch_pos_range(nS, pp.pos, pp.pos_stride, apad, pp.n_ant, 0, rg, st);
      // pinned host copy: the copy is asynchronous and the host only polls the event below
      const size_t n_rg = 6 * (size_t)nS;
      if (h_rg_cap < n_rg) {
        if (h_rg) BIPP_HIP(hipHostFree(h_rg));
        h_rg = nullptr;
        BIPP_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_rg), sizeof(double) * n_rg));
        h_rg_cap = n_rg;
      }
      BIPP_HIP(hipMemcpyAsync(h_rg, rg, sizeof(double) * n_rg, hipMemcpyDeviceToHost, st));
      c.sync();  // the ranges arrive after the eigensolve
      double X[3] = {0, 0, 0}, Cc[3] = {0, 0, 0}, S[3], D[3];
      for (int k = 0; k < nS; ++k)
        for (int d = 0; d < 3; ++d)
          X[d] = std::max(X[d], 2 * 3.14159265358979323846 * (h_rg[6 * k + 3 + d] - h_rg[6 * k + d]));
      for (int d = 0; d < 3; ++d) {
        S[d] = 0.5 * (pix_hi[d] - pix_lo[d]);
        D[d] = 0.5 * (pix_hi[d] + pix_lo[d]);
      }
      n3plan.make(tol, X, S, Cc, D, compute == BIPP_COMPUTE_FP64 ? 4096 : 4096);
      for (int d = 0; d < 3; ++d) last_nf2[d] = n3plan.nf2[d];
      BIPP_HIP(hipEventRecord(ev[2], st));
      const int rows_max = er.general_w ? n_ant : n_beam;
      if (compute == BIPP_COMPUTE_FP64)
        n3d.run<R>(c, n3plan, nS, n_levels, n_beam, rows_max, pp.n_ant, pp.z, pp.stride_z, pp.ldz,
                   wd, pp.pos, pp.pos_stride, apad, static_cast<const double*>(lmn64.p), n_pix,
                   img.p, false, ev_n3);
      else
        n3f.run<R>(c, n3plan, nS, n_levels, n_beam, rows_max, pp.n_ant, pp.z, pp.stride_z, pp.ldz,
                   wd, pp.pos, pp.pos_stride, apad, static_cast<const double*>(lmn64.p), n_pix,
                   img.p, compute == BIPP_COMPUTE_FP32, ev_n3);
      n3_timed = true;
      BIPP_HIP(hipEventRecord(ev[3], st));
    } else {
      // virtual visibilities + direct NUFT sum, per sample and level
      R* wd = wdense.as<R>((size_t)nS * n_levels * n_beam);
      launch_level_weights<R>(nS, n_levels, filter, d_int, er.d_n_red(), er.d_d(), n_beam, n_beam,
                              er.d_n_vis(), wd, (long long)n_levels * n_beam, st);
      BIPP_HIP(hipEventRecord(ev[2], st));
      cplx<R>* vf;
      if (er.general_w) {
        vf = const_cast<cplx<R>*>(er.d_vfull());
      } else {
        vf = vfull.as<cplx<R>>((size_t)nS * n_ant * n_beam);
        UnbeamParams<R> up;
        up.n_ant = n_ant;
        up.n_red = er.d_n_red();
        up.idx = er.d_idx();
        up.idx_stride = n_beam;
        up.z = er.d_z();
        up.stride_z = (long long)n_beam * n_beam;
        up.ldz = n_beam;
        up.v = vf;
        up.stride_v = (long long)n_ant * n_beam;
        up.ldv = n_ant;
        launch_unbeam_identity<R>(up, nS, n_beam, sizeof(R) == 8, st);
      }
      const long long nn = (long long)n_ant * n_ant;
      cplx<R>* vvb = vv.as<cplx<R>>((size_t)nn);
      std::vector<double> h_wl(nS);
      BIPP_HIP(hipMemcpyAsync(h_wl.data(), wl, sizeof(double) * nS, hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
      const bool mixed = compute == BIPP_COMPUTE_MIXED;
      for (int k = 0; k < nS; ++k) {
        const double uvw_scale = 2.0 * 3.14159265358979323846 / h_wl[k];
        const double* xk = xyz + (size_t)k * stride_xyz;
        for (int lv = 0; lv < n_levels; ++lv) {
          launch_virtual_vis<R>(1, n_ant, n_beam, vf + (size_t)k * n_ant * n_beam, n_ant, 0,
                                wd + ((size_t)k * n_levels + lv) * n_beam, 0, vvb, 0, st);
          if (mixed) {
            double* u = uvw.as<double>(3 * (size_t)nn);
            launch_baselines<double>(n_ant, xk, ldxyz, uvw_scale, u, u + nn, u + 2 * nn, st);
            cplx<float>* v32 = vv32.as<cplx<float>>((size_t)nn);
            launch_copy_cast<R, float>(1, (int)nn, 1, vvb, nn, 0, v32, nn, 0, st);
            const double* l = static_cast<const double*>(lmn64.p);
            launch_nuft_sum_mixed(1.0, (int)nn, v32, u, u + nn, u + 2 * nn, n_pix, l, l + n_pix,
                                  l + 2 * (size_t)n_pix, static_cast<float*>(img.p) + (size_t)lv * n_pix,
                                  st);
          } else {
            R* u = uvw.as<R>(3 * (size_t)nn);
            launch_baselines<R>(n_ant, xk, ldxyz, uvw_scale, u, u + nn, u + 2 * nn, st);
            const R* l = sizeof(R) == 8 ? static_cast<const R*>(lmn64.p) : static_cast<const R*>(lmn32.p);
            launch_nuft_sum<R>(R(1), (int)nn, vvb, u, u + nn, u + 2 * nn, n_pix, l, l + n_pix,
                               l + 2 * (size_t)n_pix, static_cast<R*>(img.p) + (size_t)lv * n_pix, st);
          }
        }
      }
      BIPP_HIP(hipEventRecord(ev[3], st));
    }
    accumulate_times();
    // keep the eigenvalues of this chunk (appended to those of the collect call's earlier chunks)
    const size_t o = (size_t)last_count;
    last_count += nS;
    last_d.resize((size_t)last_count * n_beam);
    last_nred.resize(last_count);
    std::vector<R> tmp((size_t)nS * n_beam);
    BIPP_HIP(hipMemcpyAsync(tmp.data(), er.d_d(), sizeof(R) * tmp.size(), hipMemcpyDeviceToHost, st));
    BIPP_HIP(hipMemcpyAsync(last_nred.data() + o, er.d_n_red(), sizeof(int) * nS,
                            hipMemcpyDeviceToHost, st));
    c.sync();
    for (size_t i = 0; i < tmp.size(); ++i) last_d[o * n_beam + i] = static_cast<double>(tmp[i]);
    n_collected += nS;
  }

  void collect(int nS, const double* wl, bool s_double, const void* s, long long lds,
               long long stride_s, const void* w, long long ldw, long long stride_w,
               const double* xyz, long long ldxyz, long long stride_xyz) {
    const size_t es = s_double ? 16 : 8;
    const int chunk = batched ? 4096 : 1;
    last_count = 0;
    for (int k0 = 0; k0 < nS; k0 += chunk) {
      const int m = std::min(chunk, nS - k0);
      const void* sk = static_cast<const char*>(s) + es * (size_t)k0 * stride_s;
      const void* wk = w ? static_cast<const char*>(w) + es * (size_t)k0 * stride_w : nullptr;
      const double* xk = xyz + (size_t)k0 * stride_xyz;
      if (compute == BIPP_COMPUTE_FP32)
        collect_impl<float>(ctx->eig32, m, wl + k0, s_double, sk, lds, stride_s, wk, ldw, stride_w,
                            xk, ldxyz, stride_xyz);
      else
        collect_impl<double>(ctx->eig64, m, wl + k0, s_double, sk, lds, stride_s, wk, ldw,
                             stride_w, xk, ldxyz, stride_xyz);
    }
  }
};

// ---------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------
extern "C" {

const char* bipp_last_error(void) { return g_last_error.c_str(); }

const char* bipp_version(void) { return "1.0.0-mi355x"; }

BippError bipp_device_count(int* count) {
  return guard([&] {
    require(count != nullptr, BIPP_INVALID_POINTER_ERROR, "null pointer");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
  });
}

BippError bipp_ctx_create(BippProcessingUnit pu, BippContext* ctx) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_POINTER_ERROR, "null pointer");
    auto c = std::make_unique<BippContextImpl>();
    c->impl = std::make_unique<ContextImpl>(pu);
    *ctx = c.release();
  });
}

BippError bipp_ctx_destroy(BippContext ctx) {
  return guard([&] {
    if (ctx) ctx->release();
  });
}

BippError bipp_ctx_processing_unit(BippContext ctx, BippProcessingUnit* pu) {
  return guard([&] {
    require(ctx && pu, BIPP_INVALID_HANDLE_ERROR, "null handle");
    *pu = ctx->impl->pu;
  });
}

BippError bipp_ctx_stream(BippContext ctx, void** stream) {
  return guard([&] {
    require(ctx && stream, BIPP_INVALID_HANDLE_ERROR, "null handle");
    *stream = ctx->impl->stream;
  });
}

BippError bipp_ctx_synchronize(BippContext ctx) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    ctx->impl->sync();
  });
}

BippError bipp_eigh_gram_batched(BippContext ctx, size_t nSamples, size_t nAntenna, size_t nBeam,
                                 BippPrecision sPrec, BippPrecision compute, BippPrecision outPrec,
                                 const double* wl, const void* s, size_t lds, size_t strideS,
                                 const void* w, size_t ldw, size_t strideW, const double* xyz,
                                 size_t ldxyz, size_t strideXyz, void* d, size_t strideD, void* v,
                                 size_t ldv, size_t strideV, int32_t* nBeamReduced, void* scale) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null context");
    require(s && xyz && wl && d, BIPP_INVALID_POINTER_ERROR, "null array");
    if (nSamples == 0) return;
    ContextImpl& c = *ctx->impl;
    const bool sd = sPrec == BIPP_PRECISION_DOUBLE, od = outPrec == BIPP_PRECISION_DOUBLE;
    if (compute == BIPP_PRECISION_DOUBLE) {
      ctx->eig64.run(c, (int)nSamples, (int)nAntenna, (int)nBeam, wl, s, sd, lds, strideS, w, sd,
                     ldw, strideW, xyz, ldxyz, strideXyz, false, 1.0);
      export_eig<double>(c, ctx->eig64, (int)nSamples, (int)nAntenna, (int)nBeam, od, d, strideD, v,
                         ldv, strideV, nBeamReduced, scale);
      ctx->eig64.check(c);
    } else {
      ctx->eig32.run(c, (int)nSamples, (int)nAntenna, (int)nBeam, wl, s, sd, lds, strideS, w, sd,
                     ldw, strideW, xyz, ldxyz, strideXyz, false, 1.0);
      export_eig<float>(c, ctx->eig32, (int)nSamples, (int)nAntenna, (int)nBeam, od, d, strideD, v,
                        ldv, strideV, nBeamReduced, scale);
      ctx->eig32.check(c);
    }
  });
}

}  // extern "C"

namespace {

// Shared body of the single-sample host entry points (bipp::eigh_gram / bipp::eigh).
template <typename T>
void eigh_host(bool gram, T wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
               const void* w, size_t ldw, const T* xyz, size_t ldxyz, T* d, void* v, size_t ldv,
               size_t* nBeamReduced, T* scale) {
  require(s && d, BIPP_INVALID_POINTER_ERROR, "null array");
  require(!gram || xyz, BIPP_INVALID_POINTER_ERROR, "null xyz");
  require(nBeam > 0 && nAntenna > 0, BIPP_INVALID_PARAMETER_ERROR, "empty problem");
  require(lds >= nBeam && (!w || ldw >= nAntenna) && (!v || ldv >= nAntenna) &&
              (!gram || ldxyz >= nAntenna),
          BIPP_INVALID_PARAMETER_ERROR, "leading dimension too small");
  // The reference's eigh_gram takes no Context; a process-wide one serves these calls.  It
  // is never destroyed, so no HIP call runs during static destruction.
  static BippContextImpl* dflt = nullptr;
  if (!dflt) {
    auto c = std::make_unique<BippContextImpl>();
    c->impl = std::make_unique<ContextImpl>(BIPP_PU_AUTO);
    dflt = c.release();
  }
  ContextImpl& c = *dflt->impl;
  hipStream_t st = c.stream;
  constexpr bool dbl = sizeof(T) == 8;
  const size_t ce = 2 * sizeof(T);
  // the identity W of the reference's callers is passed explicitly; keep it general
  std::vector<unsigned char> hs(ce * nBeam * nBeam);
  for (size_t j = 0; j < nBeam; ++j)
    std::memcpy(hs.data() + ce * j * nBeam, static_cast<const unsigned char*>(s) + ce * j * lds,
                ce * nBeam);
  DevBuf ds, dw, dxyz, dwl, dd, dv, dnr, dsc;
  void* s_dev = ds.get(hs.size());
  BIPP_HIP(hipMemcpyAsync(s_dev, hs.data(), hs.size(), hipMemcpyHostToDevice, st));
  const void* w_dev = nullptr;
  std::vector<unsigned char> hw;
  if (w) {
    hw.resize(ce * nAntenna * nBeam);
    for (size_t j = 0; j < nBeam; ++j)
      std::memcpy(hw.data() + ce * j * nAntenna, static_cast<const unsigned char*>(w) + ce * j * ldw,
                  ce * nAntenna);
    void* p = dw.get(hw.size());
    BIPP_HIP(hipMemcpyAsync(p, hw.data(), hw.size(), hipMemcpyHostToDevice, st));
    w_dev = p;
  } else {
    require(nAntenna == nBeam, BIPP_INVALID_PARAMETER_ERROR, "w == NULL needs nAntenna == nBeam");
  }
  std::vector<double> hx(3 * nAntenna, 0.0);
  if (gram)
    for (size_t c3 = 0; c3 < 3; ++c3)
      for (size_t i = 0; i < nAntenna; ++i) hx[i + c3 * nAntenna] = static_cast<double>(xyz[i + c3 * ldxyz]);
  double* xdev = dxyz.as<double>(3 * nAntenna);
  BIPP_HIP(hipMemcpyAsync(xdev, hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice, st));
  double hwl = static_cast<double>(wl);
  double* wldev = dwl.as<double>(1);
  BIPP_HIP(hipMemcpyAsync(wldev, &hwl, sizeof(double), hipMemcpyHostToDevice, st));
  T* d_dev = static_cast<T*>(dd.get(sizeof(T) * nBeam));
  void* v_dev = dv.get(ce * nAntenna * nBeam);
  int32_t* nr_dev = dnr.as<int32_t>(1);
  T* sc_dev = static_cast<T*>(dsc.get(sizeof(T)));
  if constexpr (dbl) {
    dflt->eig64.run(c, 1, (int)nAntenna, (int)nBeam, wldev, s_dev, true, nBeam, 0, w_dev, true,
                    nAntenna, 0, xdev, nAntenna, 0, !gram, 1.0);
    export_eig<double>(c, dflt->eig64, 1, (int)nAntenna, (int)nBeam, true, d_dev, 0, v_dev,
                       nAntenna, 0, nr_dev, sc_dev);
    dflt->eig64.check(c);
  } else {
    dflt->eig32.run(c, 1, (int)nAntenna, (int)nBeam, wldev, s_dev, false, nBeam, 0, w_dev, false,
                    nAntenna, 0, xdev, nAntenna, 0, !gram, 1.0);
    export_eig<float>(c, dflt->eig32, 1, (int)nAntenna, (int)nBeam, false, d_dev, 0, v_dev,
                      nAntenna, 0, nr_dev, sc_dev);
    dflt->eig32.check(c);
  }
  std::vector<unsigned char> hv(ce * nAntenna * nBeam);
  int32_t nr = 0;
  T sc = 0;
  BIPP_HIP(hipMemcpyAsync(d, d_dev, sizeof(T) * nBeam, hipMemcpyDeviceToHost, st));
  BIPP_HIP(hipMemcpyAsync(hv.data(), v_dev, hv.size(), hipMemcpyDeviceToHost, st));
  BIPP_HIP(hipMemcpyAsync(&nr, nr_dev, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  BIPP_HIP(hipMemcpyAsync(&sc, sc_dev, sizeof(T), hipMemcpyDeviceToHost, st));
  BIPP_HIP(hipStreamSynchronize(st));
  if (v)
    for (size_t j = 0; j < nBeam; ++j)
      std::memcpy(static_cast<unsigned char*>(v) + ce * j * ldv, hv.data() + ce * j * nAntenna,
                  ce * nAntenna);
  if (nBeamReduced) *nBeamReduced = static_cast<size_t>(nr);
  if (scale) *scale = sc;
}

}  // namespace

extern "C" {

BippError bipp_eigh_gram_s(float wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                           const void* w, size_t ldw, const float* xyz, size_t ldxyz, float* d,
                           void* v, size_t ldv, size_t* nBeamReduced, float* scale) {
  return guard([&] {
    eigh_host<float>(true, wl, nAntenna, nBeam, s, lds, w, ldw, xyz, ldxyz, d, v, ldv, nBeamReduced,
                     scale);
  });
}
BippError bipp_eigh_gram_d(double wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                           const void* w, size_t ldw, const double* xyz, size_t ldxyz, double* d,
                           void* v, size_t ldv, size_t* nBeamReduced, double* scale) {
  return guard([&] {
    eigh_host<double>(true, wl, nAntenna, nBeam, s, lds, w, ldw, xyz, ldxyz, d, v, ldv,
                      nBeamReduced, scale);
  });
}
BippError bipp_eigh_s(float wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                      const void* w, size_t ldw, float* d, void* v, size_t ldv,
                      size_t* nBeamReduced, float* scale) {
  return guard([&] {
    eigh_host<float>(false, wl, nAntenna, nBeam, s, lds, w, ldw, nullptr, nAntenna, d, v, ldv,
                     nBeamReduced, scale);
  });
}
BippError bipp_eigh_d(double wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                      const void* w, size_t ldw, double* d, void* v, size_t ldv,
                      size_t* nBeamReduced, double* scale) {
  return guard([&] {
    eigh_host<double>(false, wl, nAntenna, nBeam, s, lds, w, ldw, nullptr, nAntenna, d, v, ldv,
                      nBeamReduced, scale);
  });
}

BippError bipp_synthesis_create(BippContext ctx, BippCompute compute, int imager, size_t nAntenna,
                                size_t nBeam, size_t nLevels, const double* intervals,
                                BippFilter filter, size_t nPixel, const float* lmn, size_t ldlmn,
                                int batched, BippSynthesis* out) {
  return guard([&] {
    require(ctx && out, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(intervals && lmn, BIPP_INVALID_POINTER_ERROR, "null array");
    require(nLevels >= 1 && nLevels <= 8, BIPP_INVALID_PARAMETER_ERROR, "1..8 levels supported");
    require(nBeam >= 1 && nBeam <= 512, BIPP_INVALID_PARAMETER_ERROR, "1..512 beams supported");
    require(imager >= 0 && imager <= 2, BIPP_INVALID_PARAMETER_ERROR, "unknown imager");
    require(ldlmn >= nPixel, BIPP_INVALID_PARAMETER_ERROR, "ldlmn < nPixel");
    require(compute == BIPP_COMPUTE_FP64 || compute == BIPP_COMPUTE_FP32 ||
                compute == BIPP_COMPUTE_MIXED,
            BIPP_INVALID_PARAMETER_ERROR, "unknown compute mode");
    auto h = std::make_unique<BippSynthesisImpl>();
    h->ctx = ctx;
    h->compute = compute;
    h->imager = imager;
    h->n_ant = (int)nAntenna;
    h->n_beam = (int)nBeam;
    h->n_levels = (int)nLevels;
    h->n_pix = (int)nPixel;
    h->filter = filter;
    h->batched = batched != 0;
    h->init_events();
    hipStream_t st = ctx->impl->stream;
    double* di = h->intervals.as<double>(2 * nLevels);
    BIPP_HIP(hipMemcpyAsync(di, intervals, sizeof(double) * 2 * nLevels, hipMemcpyHostToDevice, st));
    std::vector<double> l64(3 * nPixel);
    std::vector<float> l32(3 * nPixel);
    for (size_t c3 = 0; c3 < 3; ++c3)
      for (size_t q = 0; q < nPixel; ++q) {
        l32[c3 * nPixel + q] = lmn[q + c3 * ldlmn];
        l64[c3 * nPixel + q] = static_cast<double>(lmn[q + c3 * ldlmn]);
      }
    for (int c3 = 0; c3 < 3; ++c3) {
      const auto mm = std::minmax_element(l64.begin() + c3 * nPixel, l64.begin() + (c3 + 1) * nPixel);
      h->pix_lo[c3] = nPixel ? *mm.first : 0.0;
      h->pix_hi[c3] = nPixel ? *mm.second : 0.0;
    }
    double* d64 = h->lmn64.as<double>(l64.size());
    float* d32 = h->lmn32.as<float>(l32.size());
    BIPP_HIP(hipMemcpyAsync(d64, l64.data(), sizeof(double) * l64.size(), hipMemcpyHostToDevice, st));
    BIPP_HIP(hipMemcpyAsync(d32, l32.data(), sizeof(float) * l32.size(), hipMemcpyHostToDevice, st));
    const size_t isz = h->img_double() ? 8 : 4;
    void* im = h->img.get(isz * nLevels * nPixel);
    BIPP_HIP(hipMemsetAsync(im, 0, isz * nLevels * nPixel, st));
    BIPP_HIP(hipStreamSynchronize(st));
    ctx->retain();
    *out = h.release();
  });
}

BippError bipp_synthesis_destroy(BippSynthesis h) {
  return guard([&] {
    if (!h) return;
    BippContextImpl* c = h->ctx;
    (void)hipStreamSynchronize(c->impl->stream);
    delete h;
    c->release();
  });
}

BippError bipp_synthesis_collect(BippSynthesis h, size_t nSamples, const double* wl,
                                 BippPrecision sPrec, const void* s, size_t lds, size_t strideS,
                                 const void* w, size_t ldw, size_t strideW, const double* xyz,
                                 size_t ldxyz, size_t strideXyz) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(s && xyz && wl, BIPP_INVALID_POINTER_ERROR, "null array");
    if (nSamples == 0) return;
    h->collect((int)nSamples, wl, sPrec == BIPP_PRECISION_DOUBLE, s, lds, strideS, w, ldw, strideW,
               xyz, ldxyz, strideXyz);
  });
}

BippError bipp_synthesis_collect_host(BippSynthesis h, size_t nSamples, const double* wl,
                                      BippPrecision sPrec, const void* s, size_t lds,
                                      size_t strideS, const void* w, size_t ldw, size_t strideW,
                                      const double* xyz, size_t ldxyz, size_t strideXyz) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(s && xyz && wl, BIPP_INVALID_POINTER_ERROR, "null array");
    if (nSamples == 0) return;
    hipStream_t st = h->ctx->impl->stream;
    const size_t es = sPrec == BIPP_PRECISION_DOUBLE ? 16 : 8;
    const size_t sb = es * strideS * nSamples;
    void* sd = h->st_s.get(sb);
    BIPP_HIP(hipMemcpyAsync(sd, s, sb, hipMemcpyHostToDevice, st));
    const void* wd = nullptr;
    if (w) {
      const size_t wb = es * strideW * nSamples;
      void* p = h->st_w.get(wb);
      BIPP_HIP(hipMemcpyAsync(p, w, wb, hipMemcpyHostToDevice, st));
      wd = p;
    }
    const size_t xb = sizeof(double) * strideXyz * nSamples;
    void* xd = h->st_xyz.get(xb);
    BIPP_HIP(hipMemcpyAsync(xd, xyz, xb, hipMemcpyHostToDevice, st));
    double* wld = h->st_wl.as<double>(nSamples);
    BIPP_HIP(hipMemcpyAsync(wld, wl, sizeof(double) * nSamples, hipMemcpyHostToDevice, st));
    h->collect((int)nSamples, wld, sPrec == BIPP_PRECISION_DOUBLE, sd, lds, strideS, wd, ldw,
               strideW, static_cast<const double*>(xd), ldxyz, strideXyz);
  });
}

BippError bipp_synthesis_get(BippSynthesis h, size_t level, int normalize, float* image) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(image != nullptr, BIPP_INVALID_POINTER_ERROR, "null image");
    require(level < (size_t)h->n_levels, BIPP_INVALID_PARAMETER_ERROR, "level out of range");
    hipStream_t st = h->ctx->impl->stream;
    const double f = (normalize && h->n_collected) ? 1.0 / static_cast<double>(h->n_collected) : 1.0;
    if (h->img_double()) {
      std::vector<double> t(h->n_pix);
      BIPP_HIP(hipMemcpyAsync(t.data(), static_cast<double*>(h->img.p) + level * h->n_pix,
                              sizeof(double) * h->n_pix, hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
      for (int q = 0; q < h->n_pix; ++q) image[q] = static_cast<float>(t[q] * f);
    } else {
      BIPP_HIP(hipMemcpyAsync(image, static_cast<float*>(h->img.p) + level * h->n_pix,
                              sizeof(float) * h->n_pix, hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
      const float ff = static_cast<float>(f);
      if (normalize)
        for (int q = 0; q < h->n_pix; ++q) image[q] *= ff;
    }
  });
}

BippError bipp_synthesis_image_device(BippSynthesis h, void** ptr, BippPrecision* type) {
  return guard([&] {
    require(h && ptr && type, BIPP_INVALID_HANDLE_ERROR, "null handle");
    *ptr = h->img.p;
    *type = h->img_double() ? BIPP_PRECISION_DOUBLE : BIPP_PRECISION_SINGLE;
  });
}

BippError bipp_synthesis_num_collected(BippSynthesis h, size_t* n) {
  return guard([&] {
    require(h && n, BIPP_INVALID_HANDLE_ERROR, "null handle");
    *n = h->n_collected;
  });
}

BippError bipp_synthesis_reset(BippSynthesis h) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    hipStream_t st = h->ctx->impl->stream;
    const size_t isz = h->img_double() ? 8 : 4;
    BIPP_HIP(hipMemsetAsync(h->img.p, 0, isz * h->n_levels * h->n_pix, st));
    BIPP_HIP(hipStreamSynchronize(st));
    h->n_collected = 0;
  });
}

BippError bipp_synthesis_set_tolerance(BippSynthesis h, double tol) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(tol > 0 && tol < 1, BIPP_INVALID_PARAMETER_ERROR, "tolerance must be in (0, 1)");
    h->tol = tol;
  });
}

BippError bipp_synthesis_nufft_grid(BippSynthesis h, int32_t* nf) {
  return guard([&] {
    require(h && nf, BIPP_INVALID_HANDLE_ERROR, "null handle");
    for (int d = 0; d < 3; ++d) nf[d] = h->last_nf2[d];
  });
}

BippError bipp_synthesis_timings(BippSynthesis h, double* ms, int reset) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    if (ms) std::copy(h->stage_ms, h->stage_ms + BippSynthesisImpl::N_TIMES, ms);
    if (reset) std::fill(h->stage_ms, h->stage_ms + BippSynthesisImpl::N_TIMES, 0.0);
  });
}

BippError bipp_synthesis_last_eigenvalues(BippSynthesis h, size_t capacity, double* d,
                                          int32_t* nBeamReduced, size_t* count) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    const size_t n = static_cast<size_t>(h->last_count);
    if (count) *count = n;
    if (!d && !nBeamReduced) return;
    require(capacity >= n, BIPP_INVALID_PARAMETER_ERROR,
            "capacity " + std::to_string(capacity) + " below the " + std::to_string(n) +
                " samples of the last collect call");
    if (d) std::copy(h->last_d.begin(), h->last_d.begin() + n * h->n_beam, d);
    if (nBeamReduced) std::copy(h->last_nred.begin(), h->last_nred.begin() + n, nBeamReduced);
  });
}

BippError bipp_virtual_vis(BippContext ctx, BippPrecision prec, size_t nSamples, size_t nAntenna,
                           size_t nEig, const void* v, size_t ldv, size_t strideV, const void* dw,
                           size_t strideDw, void* vv, size_t strideVv) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null context");
    require(v && dw && vv, BIPP_INVALID_POINTER_ERROR, "null array");
    hipStream_t st = ctx->impl->stream;
    if (prec == BIPP_PRECISION_DOUBLE)
      launch_virtual_vis<double>((int)nSamples, (int)nAntenna, (int)nEig,
                                 static_cast<const cplx<double>*>(v), ldv, strideV,
                                 static_cast<const double*>(dw), strideDw,
                                 static_cast<cplx<double>*>(vv), strideVv, st);
    else
      launch_virtual_vis<float>((int)nSamples, (int)nAntenna, (int)nEig,
                                static_cast<const cplx<float>*>(v), ldv, strideV,
                                static_cast<const float*>(dw), strideDw,
                                static_cast<cplx<float>*>(vv), strideVv, st);
  });
}

BippError bipp_nuft_sum(BippContext ctx, BippPrecision prec, double alpha, size_t nIn,
                        const void* input, const void* u, const void* v, const void* w,
                        size_t nOut, const void* x, const void* y, const void* z, void* out) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null context");
    require(input && u && v && w && x && y && z && out, BIPP_INVALID_POINTER_ERROR, "null array");
    hipStream_t st = ctx->impl->stream;
    if (prec == BIPP_PRECISION_DOUBLE)
      launch_nuft_sum<double>(alpha, (int)nIn, static_cast<const cplx<double>*>(input),
                              static_cast<const double*>(u), static_cast<const double*>(v),
                              static_cast<const double*>(w), (int)nOut,
                              static_cast<const double*>(x), static_cast<const double*>(y),
                              static_cast<const double*>(z), static_cast<double*>(out), st);
    else
      launch_nuft_sum<float>(static_cast<float>(alpha), (int)nIn,
                             static_cast<const cplx<float>*>(input), static_cast<const float*>(u),
                             static_cast<const float*>(v), static_cast<const float*>(w), (int)nOut,
                             static_cast<const float*>(x), static_cast<const float*>(y),
                             static_cast<const float*>(z), static_cast<float*>(out), st);
  });
}

BippError bipp_nuft_sum_mixed(BippContext ctx, double alpha, size_t nIn, const void* input,
                              const double* u, const double* v, const double* w, size_t nOut,
                              const double* x, const double* y, const double* z, float* out) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null context");
    require(input && u && v && w && x && y && z && out, BIPP_INVALID_POINTER_ERROR, "null array");
    launch_nuft_sum_mixed(alpha, (int)nIn, static_cast<const cplx<float>*>(input), u, v, w,
                          (int)nOut, x, y, z, out, ctx->impl->stream);
  });
}

BippError bipp_gram(BippContext ctx, BippPrecision prec, size_t nSamples, size_t nAntenna,
                    const double* wl, const double* xyz, size_t ldxyz, size_t strideXyz,
                    double prefactor, void* g, size_t ldg, size_t strideG) {
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null context");
    require(wl && xyz && g, BIPP_INVALID_POINTER_ERROR, "null array");
    hipStream_t st = ctx->impl->stream;
    if (prec == BIPP_PRECISION_DOUBLE)
      launch_sinc_matrix<double>((int)nSamples, (int)nAntenna, xyz, ldxyz, strideXyz, wl, prefactor,
                                 static_cast<cplx<double>*>(g), ldg, strideG, st);
    else
      launch_sinc_matrix<float>((int)nSamples, (int)nAntenna, xyz, ldxyz, strideXyz, wl, prefactor,
                                static_cast<cplx<float>*>(g), ldg, strideG, st);
  });
}

}  // extern "C"
