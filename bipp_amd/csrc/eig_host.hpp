// Host side of the batched generalized eigensolve (flag -> Gram -> eig_kernel -> un-beam)
// and the library context object that owns its workspaces.
#pragma once

#include "host_runtime.hpp"
#include "kernels.hpp"

namespace bippamd {

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
    int n_slots = std::max(1, std::min(nS, per_cu * ctx.n_cu));
    // BIPP_EIG_MAX_SLOTS caps the persistent workgroups (their matrices then share the
    // last-level cache among fewer slots)
    if (const char* e = std::getenv("BIPP_EIG_MAX_SLOTS"))
      if (std::atoi(e) > 0) n_slots = std::min(n_slots, std::atoi(e));
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
      ep.prof = prof_buf.as<unsigned long long>((size_t)nS * EIG_PROF_SLOTS);
      BIPP_HIP(hipMemsetAsync(ep.prof, 0, sizeof(unsigned long long) * nS * EIG_PROF_SLOTS, st));
    }
    launch_eig<R>(ep, n_slots, s_double, st);
    if (prof) {
      std::vector<unsigned long long> h((size_t)nS * EIG_PROF_SLOTS);
      BIPP_HIP(hipMemcpyAsync(h.data(), ep.prof, sizeof(unsigned long long) * h.size(),
                              hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
      static const char* names[7] = {"load", "cholesky", "whiten", "tridiag", "tridiag_eig",
                                     "backtransform", "unwhiten"};
      double acc[7] = {0, 0, 0, 0, 0, 0, 0}, sub[8] = {0, 0, 0, 0, 0, 0, 0, 0}, tsub[3] = {0, 0, 0},
             csub[4] = {0, 0, 0, 0};
      int cnt = 0;
      for (int k = 0; k < nS; ++k) {
        const unsigned long long* r = &h[(size_t)k * EIG_PROF_SLOTS];
        if (!r[0] || !r[7]) continue;
        ++cnt;
        for (int q = 0; q < 7; ++q) acc[q] += r[q + 1] >= r[q] && r[q] ? double(r[q + 1] - r[q]) : 0.0;
        for (int q = 0; q < 8; ++q) sub[q] += double(r[8 + q]);
        for (int q = 0; q < 2; ++q) tsub[q] += double(r[16 + q]);
        for (int q = 0; q < 4; ++q) csub[q] += double(r[18 + q]);
      }
      std::fprintf(stderr, "[bipp eig profile] %d samples, n_slots %d, us/sample (100 MHz clock):", cnt,
                   n_slots);
      for (int q = 0; q < 7; ++q) std::fprintf(stderr, " %s %.1f", names[q], cnt ? acc[q] / cnt / 100.0 : 0.0);
      // steps of the divide and conquer, summed over its levels
      static const char* subn[8] = {"merge order", "deflation", "roots", "z", "norms",
                                    "rank sort", "M", "products"};
      std::fprintf(stderr, " | tridiag_eig:");
      for (int q = 0; q < 8; ++q) std::fprintf(stderr, " %s %.1f", subn[q], cnt ? sub[q] / cnt / 100.0 : 0.0);
      // steps of the tridiagonalisation, summed over its columns
      static const char* tsn[2] = {"w, column update and reflector", "pass"};
      std::fprintf(stderr, " | tridiag:");
      for (int q = 0; q < 2; ++q) std::fprintf(stderr, " %s %.1f", tsn[q], cnt ? tsub[q] / cnt / 100.0 : 0.0);
      // steps of the Cholesky and inverse, summed over its blocks
      static const char* csn[4] = {"diagonal factor", "diagonal inverse", "panel and trailing",
                                   "off-diagonal W"};
      std::fprintf(stderr, " | cholesky:");
      for (int q = 0; q < 4; ++q) std::fprintf(stderr, " %s %.1f", csn[q], cnt ? csub[q] / cnt / 100.0 : 0.0);
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

// A context is shared by the synthesis handles created from it: each handle holds a
// reference, so destroying the context first leaves it alive until its last handle goes.
struct BippContextImpl {
  std::unique_ptr<bippamd::ContextImpl> impl;
  bippamd::EigRunner<float> eig32;
  bippamd::EigRunner<double> eig64;
  std::atomic<int> refs{1};

  void retain() { refs.fetch_add(1); }
  void release() {
    if (refs.fetch_sub(1) == 1) {
      (void)hipStreamSynchronize(impl->stream);
      delete this;
    }
  }
};
