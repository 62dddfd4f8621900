// C-ABI of libbipp_amd: contexts, the batched eigensolve and the fused synthesis pipeline
// (stage A -> level weights -> image accumulation) on one HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "bipp_hip.h"
#include "eig_host.hpp"
#include "host_runtime.hpp"
#include "kernels.hpp"
#include "nufft_host.hpp"

using namespace bippamd;

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
  N3Options n3opt;
  double pix_lo[3] = {0, 0, 0}, pix_hi[3] = {0, 0, 0};
  N3Runner<float> n3f;
  N3Runner<double> n3d;
  DevBuf dummy;
  size_t sample_batch = 0;  // samples per batch; 0: automatic
  // staging for host collects
  DevBuf st_s, st_w, st_xyz, st_wl;
  std::vector<double> last_d;
  std::vector<int32_t> last_nred;
  int last_count = 0;
  // stage timings on the context stream: 0 eigensolve, 1 prep/weights, 2 image, 3 launches
  // NUFFT sub-stages: 4 virtual vis, 5 point sort, 6 spread, 7 deconvolve + FFT, 8 interpolate
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  static constexpr int N_TIMES = 9;
  double stage_ms[N_TIMES] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool n3_timed = false;

  void init_events() {
    for (auto& e : ev) BIPP_HIP(hipEventCreate(&e));
  }
  ~BippSynthesisImpl() {
    for (auto& e : ev)
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
      if (compute == BIPP_COMPUTE_FP64)
        n3d.stage_ms(*ctx->impl, stage_ms + 4);
      else
        n3f.stage_ms(*ctx->impl, stage_ms + 4);
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
      BIPP_HIP(hipEventRecord(ev[2], st));
      const int rows_max = er.general_w ? n_ant : n_beam;
      N3Source src;
      src.mode = N3_SRC_PAIRS;
      src.n_ant_max = rows_max;
      src.n_rows = pp.n_ant;
      src.n_rows_const = rows_max;
      src.pos = pp.pos;
      src.pos_stride = pp.pos_stride;
      src.apad = apad;
      const double* l64 = static_cast<const double*>(lmn64.p);
      if (compute == BIPP_COMPUTE_FP64) {
        n3d.values<R>(c, nS, n_levels, n_beam, rows_max, pp.n_ant, pp.z, pp.stride_z, pp.ldz, wd);
        n3d.image(c, src, nS, n_levels, n3opt, pix_lo, pix_hi, l64, n_pix, img.p, false);
      } else {
        n3f.values<R>(c, nS, n_levels, n_beam, rows_max, pp.n_ant, pp.z, pp.stride_z, pp.ldz, wd);
        n3f.image(c, src, nS, n_levels, n3opt, pix_lo, pix_hi, l64, n_pix, img.p,
                  compute == BIPP_COMPUTE_FP32);
      }
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
    accumulate_times();  // every event of the chunk has completed
    for (size_t i = 0; i < tmp.size(); ++i) last_d[o * n_beam + i] = static_cast<double>(tmp[i]);
    n_collected += nS;
  }

  // ---- per-timestep chain captured in a HIP graph (steering imager, identity beamforming) ----
  // flag -> eigensolve (samples in index order) -> level weights and operands -> steering
  // image into a scratch image -> added to the images unless a solve failed.  No host
  // round trip inside the chain, so it is captured once per argument set and replayed.
  struct GraphState {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nS = 0;
    const double* wl = nullptr;
    bool s_double = false;
    const void* s = nullptr;
    long long lds = 0, stride_s = 0;
    const double* xyz = nullptr;
    long long ldxyz = 0, stride_xyz = 0;
    DevBuf n_red, idx, n_vis, info, counter, d, z, ws_a, ws_l, ws_z, n_eig, n_ant_used, pos,
        afrag, wts, img_tmp, flag;
    long long captures = 0, launches = 0;
    void drop() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
    }
    ~GraphState() { drop(); }
  } gs;

  template <typename R>
  void record_chain(int nS, const double* wl, bool s_double, const void* s, long long lds,
                    long long stride_s, const double* xyz, long long ldxyz, long long stride_xyz,
                    int n_slots, hipStream_t st) {
    const int nmax = n_beam;
    const long long nn = (long long)nmax * nmax;
    FlagParams fp;
    fp.n_beam = n_beam;
    fp.mode = 0;
    fp.s = s;
    fp.lds = lds;
    fp.stride_s = stride_s;
    fp.n_red = static_cast<int*>(gs.n_red.p);
    fp.idx = static_cast<int*>(gs.idx.p);
    fp.idx_stride = nmax;
    fp.n_vis = static_cast<unsigned long long*>(gs.n_vis.p);
    launch_flag(fp, nS, s_double, st);
    EigParams<R> ep;
    ep.n_jobs = nS;
    ep.nmax = nmax;
    ep.order = nullptr;
    ep.counter = static_cast<unsigned*>(gs.counter.p);
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
    ep.stride_g = nn;
    ep.ld_g = nmax;
    ep.gram_prefactor = 1.0;
    ep.standard = 0;
    ep.d_out = static_cast<R*>(gs.d.p);
    ep.stride_d = nmax;
    ep.z_out = static_cast<cplx<R>*>(gs.z.p);
    ep.stride_z = nn;
    ep.ldz = nmax;
    ep.ws_ld = nmax;
    ep.ws_stride = nn;
    ep.ws_a = static_cast<cplx<R>*>(gs.ws_a.p);
    ep.ws_l = static_cast<cplx<R>*>(gs.ws_l.p);
    ep.ws_z = static_cast<R*>(gs.ws_z.p);
    ep.info = static_cast<int*>(gs.info.p);
    launch_eig<R>(ep, n_slots, s_double, st);
    const int mode = steer_mode();
    const int apad = ((n_ant + 15) / 16) * 16;
    const int kchunks = apad / 16;
    const int rtiles = (2 * n_beam + 15) / 16;
    const int epad = rtiles * 8;
    const size_t fe = steer_frag_elems_per_sample(mode, rtiles, kchunks);
    SteerPrepParams<R> pp;
    pp.n_samples = nS;
    pp.n_levels = n_levels;
    pp.filter = filter;
    pp.intervals = static_cast<const double*>(intervals.p);
    pp.n_red = fp.n_red;
    pp.idx = fp.idx;
    pp.idx_stride = n_beam;
    pp.d = ep.d_out;
    pp.stride_d = n_beam;
    pp.n_cols = n_beam;
    pp.n_vis = fp.n_vis;
    pp.z = ep.z_out;
    pp.ldz = n_beam;
    pp.stride_z = nn;
    pp.xyz = xyz;
    pp.ldxyz = ldxyz;
    pp.stride_xyz = stride_xyz;
    pp.wl = wl;
    pp.n_eig = static_cast<int*>(gs.n_eig.p);
    pp.n_ant = static_cast<int*>(gs.n_ant_used.p);
    pp.pos_stride = 3LL * apad;
    pp.pos = static_cast<double*>(gs.pos.p);
    pp.apad = apad;
    pp.afrag_stride = static_cast<long long>(fe);
    pp.afrag = gs.afrag.p;
    pp.kchunks = kchunks;
    pp.rtiles = rtiles;
    pp.w_stride = (long long)n_levels * epad;
    pp.wts = gs.wts.p;
    pp.epad = epad;
    pp.mode = mode;
    pp.n_rows_const = 0;
    launch_steer_prep<R>(pp, st);
    const long long n_img = (long long)n_levels * n_pix;
    launch_zero_image(gs.img_tmp.p, n_img, img_double(), st);
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
    sp.img = gs.img_tmp.p;
    sp.img_ld = n_pix;
    launch_steer(sp, mode, st);
    launch_commit_image(ep.info, nS, gs.img_tmp.p, img.p, n_img, img_double(),
                        static_cast<int*>(gs.flag.p), st);
  }

  template <typename R>
  void collect_graph_impl(int nS, const double* wl, bool s_double, const void* s, long long lds,
                          long long stride_s, const double* xyz, long long ldxyz,
                          long long stride_xyz) {
    ContextImpl& c = *ctx->impl;
    hipStream_t st = c.stream;
    const bool same = gs.exec && gs.nS == nS && gs.wl == wl && gs.s_double == s_double &&
                      gs.s == s && gs.lds == lds && gs.stride_s == stride_s && gs.xyz == xyz &&
                      gs.ldxyz == ldxyz && gs.stride_xyz == stride_xyz;
    if (!same) {
      gs.drop();
      const int nmax = n_beam;
      const long long nn = (long long)nmax * nmax;
      const int per_cu = eig_blocks_per_cu<R>(nmax, s_double);
      const int n_slots = std::max(1, std::min(nS, per_cu * c.n_cu));
      const int mode = steer_mode();
      const int apad = ((n_ant + 15) / 16) * 16;
      const int rtiles = (2 * n_beam + 15) / 16;
      const size_t fe = steer_frag_elems_per_sample(mode, rtiles, apad / 16);
      const size_t esz = mode == STEER_BF16X3 ? 2 : (mode == STEER_FP32 ? 4 : 8);
      // every buffer the graph touches is allocated before the capture and owned by it
      gs.n_red.as<int>(nS);
      gs.idx.as<int>((size_t)nS * nmax);
      gs.n_vis.as<unsigned long long>(nS);
      gs.info.as<int>(nS);
      gs.counter.as<unsigned>(1);
      gs.d.as<R>((size_t)nS * nmax);
      gs.z.as<cplx<R>>((size_t)nS * nn);
      gs.ws_a.as<cplx<R>>((size_t)n_slots * nn);
      gs.ws_l.as<cplx<R>>((size_t)n_slots * nn);
      gs.ws_z.as<R>((size_t)n_slots * nn);
      gs.n_eig.as<int>(nS);
      gs.n_ant_used.as<int>(nS);
      gs.pos.as<double>((size_t)nS * 3 * apad);
      gs.afrag.get((size_t)nS * fe * esz);
      gs.wts.get((size_t)nS * n_levels * rtiles * 8 * (mode == STEER_FP64 ? 8 : 4));
      gs.img_tmp.get((img_double() ? 8 : 4) * (size_t)n_levels * n_pix);
      gs.flag.as<int>(1);
      c.sync();
      BIPP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        record_chain<R>(nS, wl, s_double, s, lds, stride_s, xyz, ldxyz, stride_xyz, n_slots, st);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      BIPP_HIP(hipStreamEndCapture(st, &gs.graph));
      BIPP_HIP(hipGraphInstantiate(&gs.exec, gs.graph, nullptr, nullptr, 0));
      gs.nS = nS;
      gs.wl = wl;
      gs.s_double = s_double;
      gs.s = s;
      gs.lds = lds;
      gs.stride_s = stride_s;
      gs.xyz = xyz;
      gs.ldxyz = ldxyz;
      gs.stride_xyz = stride_xyz;
      ++gs.captures;
    }
    BIPP_HIP(hipGraphLaunch(gs.exec, st));
    ++gs.launches;
    int flag = 0;
    BIPP_HIP(hipMemcpyAsync(&flag, gs.flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    c.sync();
    if (flag)
      fail(BIPP_EIGENSOLVER_ERROR,
           "generalized eigensolve failed for a sample of the batch; the batch was not added");
    n_collected += nS;
  }

  void collect(int nS, const double* wl, bool s_double, const void* s, long long lds,
               long long stride_s, const void* w, long long ldw, long long stride_w,
               const double* xyz, long long ldxyz, long long stride_xyz) {
    const size_t es = s_double ? 16 : 8;
    // samples per batch: the caller's choice, else 4096 (the NUFFT imager: the reference's
    // automatic size, about 2 GB of points at 20 bytes per baseline, src/nufft_synthesis.cpp:97-103)
    int chunk = 4096;
    if (imager == 2)
      chunk = static_cast<int>(std::max<long long>(
          1, std::min<long long>(4096, 2000000000LL / (20LL * n_ant * n_ant))));
    if (sample_batch > 0) chunk = static_cast<int>(std::min<size_t>(sample_batch, 1 << 20));
    if (!batched) chunk = 1;
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

BippError bipp_synthesis_collect_graph(BippSynthesis h, size_t nSamples, const double* wl,
                                       BippPrecision sPrec, const void* s, size_t lds,
                                       size_t strideS, const double* xyz, size_t ldxyz,
                                       size_t strideXyz) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(s && xyz && wl, BIPP_INVALID_POINTER_ERROR, "null array");
    require(h->imager == 0, BIPP_INVALID_PARAMETER_ERROR,
            "graph-captured collects run the steering imager (the NUFFT plan depends on the data)");
    require(h->n_ant == h->n_beam, BIPP_INVALID_PARAMETER_ERROR,
            "graph-captured collects take identity beamforming (nAntenna == nBeam)");
    require(lds >= (size_t)h->n_beam && ldxyz >= (size_t)h->n_ant, BIPP_INVALID_PARAMETER_ERROR,
            "leading dimension too small");
    require(nSamples <= 4096, BIPP_INVALID_PARAMETER_ERROR,
            "a graph-captured collect takes at most 4096 samples");
    if (nSamples == 0) return;
    const bool sd = sPrec == BIPP_PRECISION_DOUBLE;
    if (h->compute == BIPP_COMPUTE_FP32)
      h->collect_graph_impl<float>((int)nSamples, wl, sd, s, lds, strideS, xyz, ldxyz, strideXyz);
    else
      h->collect_graph_impl<double>((int)nSamples, wl, sd, s, lds, strideS, xyz, ldxyz, strideXyz);
  });
}

BippError bipp_synthesis_graph_info(BippSynthesis h, long long* captures, long long* launches) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    if (captures) *captures = h->gs.captures;
    if (launches) *launches = h->gs.launches;
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
    h->n3opt.tol = tol;
  });
}

BippError bipp_synthesis_set_partition(BippSynthesis h, int method, const size_t* dims) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    set_partition(h->n3opt, method, dims);
  });
}

BippError bipp_synthesis_set_sample_batch_size(BippSynthesis h, size_t n) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    h->sample_batch = n;
  });
}

BippError bipp_synthesis_nufft_grid(BippSynthesis h, int32_t* nf) {
  return guard([&] {
    require(h && nf, BIPP_INVALID_HANDLE_ERROR, "null handle");
    const int* src = h->compute == BIPP_COMPUTE_FP64 ? h->n3d.last_nf2 : h->n3f.last_nf2;
    for (int d = 0; d < 3; ++d) nf[d] = src[d];
  });
}

BippError bipp_synthesis_set_autotune(BippSynthesis h, int enable) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    h->n3f.autotune = h->n3d.autotune = enable != 0;
  });
}

BippError bipp_synthesis_nufft_tuning(BippSynthesis h, int32_t* shape, size_t capacity,
                                      double* candidates, size_t* n_candidates) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    const bool d64 = h->compute == BIPP_COMPUTE_FP64;
    const N3Tuning& t = d64 ? h->n3d.last_tuning : h->n3f.last_tuning;
    const auto& log = d64 ? h->n3d.tune_log : h->n3f.tune_log;
    export_tuning(t, log, shape, capacity, candidates, n_candidates);
  });
}

BippError bipp_nufft_tuning_reset(void) {
  return guard([&] { n3_tune_clear(); });
}

BippError bipp_synthesis_nufft_partition(BippSynthesis h, int32_t* dims, int32_t* boxes) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    const bool d64 = h->compute == BIPP_COMPUTE_FP64;
    if (dims)
      for (int d = 0; d < 3; ++d) dims[d] = d64 ? h->n3d.last_dims[d] : h->n3f.last_dims[d];
    if (boxes) *boxes = d64 ? h->n3d.last_boxes : h->n3f.last_boxes;
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
