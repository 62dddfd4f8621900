// Eigenpair imager: stage B of image synthesis on eigenpairs a dataset already holds.
//
// Replaces the per-sample loop of nufft_synthesis (src/nufft_synthesis.cpp:74-170) and the
// NUFFTInterface::add / get_image pair behind it (src/nufft_interface.hpp:13-21,
// src/gpu/nufft.hpp:28-235).  Samples are staged in pinned host memory until a batch is full
// (the reference's sampleBatchSize, src/nufft_synthesis.cpp:97-103), then one launch turns
// every (sample, image) into virtual visibilities VV = V diag(scale dScaled) V^H
// (src/host/virtual_vis.cpp:19-73) and one partitioned type-3 NUFFT, or the direct sum,
// adds Re sum_ij VV_ij exp(+i uvw_ij . x) to every image.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "bipp_hip.h"
#include "eig_host.hpp"
#include "host_runtime.hpp"
#include "imager_prep.hpp"
#include "kernels.hpp"
#include "nufft_host.hpp"

using namespace bippamd;

struct BippImagerImpl {
  BippContextImpl* ctx = nullptr;
  BippCompute compute = BIPP_COMPUTE_MIXED;
  int imager = 2;
  int n_ant = 0, n_beam = 0, n_img = 0, n_pix = 0;
  N3Options opt;
  size_t batch_req = 0;  // requested samples per batch, 0: automatic
  size_t batch = 0;      // samples per batch of the current staging buffers
  size_t count = 0;      // staged samples
  size_t n_added = 0;
  double pix_lo[3] = {0, 0, 0}, pix_hi[3] = {0, 0, 0};
  PinnedBuf h_uvw, h_v, h_d, h_sc, h_wl;
  DevBuf d_uvw, d_v, d_d, d_sc, d_wl, v_r, w_r, uvw_scale, antisym, lmn64, lmn32, img;
  DevBuf vv, vv32, uvw_r;
  N3Runner<float> n3f;
  N3Runner<double> n3d;

  bool img_double() const { return compute == BIPP_COMPUTE_FP64; }
  size_t nn() const { return (size_t)n_ant * n_ant; }
  size_t sample_bytes() const {
    return sizeof(float) * 3 * nn() + 8 * (size_t)n_ant * n_beam +
           sizeof(float) * ((size_t)n_img * n_beam + 2);
  }

  void ensure_staging() {
    if (batch) return;
    size_t b = batch_req;
    if (b == 0) {
      // the reference's automatic batch: about 2 GB of points at 20 bytes per baseline
      b = std::max<size_t>(1, 2000000000ULL / (20ULL * nn()));
      b = std::min<size_t>(b, 4096);
    }
    // staging stays within 1 GiB of pinned memory
    b = std::max<size_t>(1, std::min<size_t>(b, (1ULL << 30) / sample_bytes()));
    batch = b;
    h_uvw.get(sizeof(float) * 3 * nn() * b);
    h_v.get(8 * (size_t)n_ant * n_beam * b);
    h_d.get(sizeof(float) * (size_t)n_img * n_beam * b);
    h_sc.get(sizeof(float) * b);
    h_wl.get(sizeof(float) * b);
  }

  void add(float wl, float scale, const float* uvw, size_t lduvw, const void* v, size_t ldv,
           const float* dsc, size_t ldd) {
    ensure_staging();
    const size_t k = count;
    float* u = static_cast<float*>(h_uvw.p) + k * 3 * nn();
    for (int c3 = 0; c3 < 3; ++c3) std::memcpy(u + c3 * nn(), uvw + c3 * lduvw, sizeof(float) * nn());
    unsigned char* vd = static_cast<unsigned char*>(h_v.p) + k * 8 * (size_t)n_ant * n_beam;
    for (int e = 0; e < n_beam; ++e)
      std::memcpy(vd + 8 * (size_t)e * n_ant, static_cast<const unsigned char*>(v) + 8 * (size_t)e * ldv,
                  8 * (size_t)n_ant);
    float* dd = static_cast<float*>(h_d.p) + k * (size_t)n_img * n_beam;
    for (int l = 0; l < n_img; ++l)
      std::memcpy(dd + (size_t)l * n_beam, dsc + (size_t)l * ldd, sizeof(float) * n_beam);
    static_cast<float*>(h_sc.p)[k] = scale;
    static_cast<float*>(h_wl.p)[k] = wl;
    ++count;
    ++n_added;
    if (count == batch) flush();
  }

  void flush() {
    if (!count) return;
    const int nS = static_cast<int>(count);
    count = 0;  // a failure below drops the staged samples
    ContextImpl& c = *ctx->impl;
    hipStream_t st = c.stream;
    float* du = d_uvw.as<float>(3 * nn() * nS);
    BIPP_HIP(hipMemcpyAsync(du, h_uvw.p, sizeof(float) * 3 * nn() * nS, hipMemcpyHostToDevice, st));
    void* dv = d_v.get(8 * (size_t)n_ant * n_beam * nS);
    BIPP_HIP(hipMemcpyAsync(dv, h_v.p, 8 * (size_t)n_ant * n_beam * nS, hipMemcpyHostToDevice, st));
    float* dd = d_d.as<float>((size_t)n_img * n_beam * nS);
    BIPP_HIP(hipMemcpyAsync(dd, h_d.p, sizeof(float) * n_img * n_beam * nS, hipMemcpyHostToDevice, st));
    float* ds = d_sc.as<float>(nS);
    BIPP_HIP(hipMemcpyAsync(ds, h_sc.p, sizeof(float) * nS, hipMemcpyHostToDevice, st));
    float* dw = d_wl.as<float>(nS);
    BIPP_HIP(hipMemcpyAsync(dw, h_wl.p, sizeof(float) * nS, hipMemcpyHostToDevice, st));
    std::vector<float> hwl(static_cast<float*>(h_wl.p), static_cast<float*>(h_wl.p) + nS);
    if (compute == BIPP_COMPUTE_FP32)
      run<float, float>(nS, du, dv, dd, ds, dw, hwl, n3f);
    else if (compute == BIPP_COMPUTE_FP64)
      run<double, double>(nS, du, dv, dd, ds, dw, hwl, n3d);
    else
      run<double, float>(nS, du, dv, dd, ds, dw, hwl, n3f);
    c.sync();  // the staging buffers are refilled by the next add
  }

  // R: virtual-visibility precision; TG: NUFFT grid and image precision
  template <typename R, typename TG>
  void run(int nS, const float* du, const void* dv, const float* dd, const float* ds,
           const float* dw, const std::vector<float>& hwl, N3Runner<TG>& n3) {
    ContextImpl& c = *ctx->impl;
    hipStream_t st = c.stream;
    EigPairPrep<R> pp;
    pp.n_ant = n_ant;
    pp.n_beam = n_beam;
    pp.n_img = n_img;
    pp.v_in = static_cast<const cplx<float>*>(dv);
    pp.d_in = dd;
    pp.scale = ds;
    pp.wl = dw;
    pp.v_out = v_r.as<cplx<R>>((size_t)n_ant * n_beam * nS);
    pp.w_out = w_r.as<R>((size_t)n_img * n_beam * nS);
    pp.uvw_scale = uvw_scale.as<double>(nS);
    launch_eigpair_prep<R>(pp, nS, st);
    const double* l64 = static_cast<const double*>(lmn64.p);
    if (imager == 2) {
      n3.template values<R>(c, nS, n_img, n_beam, n_ant, nullptr, pp.v_out,
                            (long long)n_ant * n_beam, n_ant, pp.w_out);
      N3Source src;
      src.mode = N3_SRC_UVW;
      src.n_ant_max = n_ant;
      src.n_rows_const = n_ant;
      src.uvw = du;
      src.uvw_stride = 3LL * nn();
      src.lduvw = (long long)nn();
      src.uvw_scale = pp.uvw_scale;
      unsigned char* anti = antisym.as<unsigned char>(nS);
      src.antisym = anti;
      launch_uvw_antisym(src, nS, anti, st);
      n3.image(c, src, nS, n_img, opt, pix_lo, pix_hi, l64, n_pix, img.p,
               compute == BIPP_COMPUTE_FP32);
      return;
    }
    // direct sum per (sample, image) of the reference's arithmetic (src/host/kernels/nuft_sum.cpp)
    const long long n2 = (long long)nn();
    cplx<R>* vvb = vv.as<cplx<R>>(n2);
    for (int k = 0; k < nS; ++k) {
      const double sc = 2.0 * 3.14159265358979323846 / static_cast<double>(hwl[k]);
      const float* uk = du + (size_t)k * 3 * n2;
      for (int l = 0; l < n_img; ++l) {
        launch_virtual_vis<R>(1, n_ant, n_beam, pp.v_out + (size_t)k * n_ant * n_beam, n_ant, 0,
                              pp.w_out + ((size_t)k * n_img + l) * n_beam, 0, vvb, 0, st);
        if (compute == BIPP_COMPUTE_MIXED) {
          double* u = uvw_r.as<double>(3 * n2);
          launch_scale_uvw<double>(n2, uk, n2, sc, u, u + n2, u + 2 * n2, st);
          cplx<float>* v32 = vv32.as<cplx<float>>(n2);
          launch_copy_cast<R, float>(1, (int)n2, 1, vvb, n2, 0, v32, n2, 0, st);
          launch_nuft_sum_mixed(1.0, (int)n2, v32, u, u + n2, u + 2 * n2, n_pix, l64, l64 + n_pix,
                                l64 + 2 * (size_t)n_pix,
                                static_cast<float*>(img.p) + (size_t)l * n_pix, st);
        } else {
          R* u = uvw_r.as<R>(3 * n2);
          launch_scale_uvw<R>(n2, uk, n2, sc, u, u + n2, u + 2 * n2, st);
          const R* lm = sizeof(R) == 8 ? reinterpret_cast<const R*>(lmn64.p)
                                       : reinterpret_cast<const R*>(lmn32.p);
          launch_nuft_sum<R>(R(1), (int)n2, vvb, u, u + n2, u + 2 * n2, n_pix, lm, lm + n_pix,
                             lm + 2 * (size_t)n_pix, static_cast<R*>(img.p) + (size_t)l * n_pix, st);
        }
      }
    }
  }
};

extern "C" {

BippError bipp_imager_create(BippContext ctx, BippCompute compute, int imager, size_t nAntenna,
                             size_t nBeam, size_t nImages, size_t nPixel, const float* lmn,
                             size_t ldlmn, BippImager* out) {
  return guard([&] {
    require(ctx && out, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(lmn != nullptr, BIPP_INVALID_POINTER_ERROR, "null pixel array");
    require(imager == 1 || imager == 2, BIPP_INVALID_PARAMETER_ERROR,
            "imager must be 1 (direct sum) or 2 (type-3 NUFFT)");
    require(nAntenna >= 1 && nAntenna <= 4096, BIPP_INVALID_PARAMETER_ERROR,
            "number of antennas must be in [1, 4096]");
    require(nBeam >= 1 && nBeam <= 4096, BIPP_INVALID_PARAMETER_ERROR,
            "number of beams must be in [1, 4096]");
    require(nImages >= 1 && nImages <= 64, BIPP_INVALID_PARAMETER_ERROR,
            "number of images must be in [1, 64]");
    require(nPixel >= 1 && nPixel < (1ULL << 31), BIPP_INVALID_PARAMETER_ERROR,
            "number of pixels out of range");
    require(ldlmn >= nPixel, BIPP_INVALID_PARAMETER_ERROR, "ldlmn < nPixel");
    require(compute == BIPP_COMPUTE_FP64 || compute == BIPP_COMPUTE_FP32 ||
                compute == BIPP_COMPUTE_MIXED,
            BIPP_INVALID_PARAMETER_ERROR, "unknown compute mode");
    auto h = std::make_unique<BippImagerImpl>();
    h->ctx = ctx;
    h->compute = compute;
    h->imager = imager;
    h->n_ant = (int)nAntenna;
    h->n_beam = (int)nBeam;
    h->n_img = (int)nImages;
    h->n_pix = (int)nPixel;
    hipStream_t st = ctx->impl->stream;
    std::vector<double> l64(3 * nPixel);
    std::vector<float> l32(3 * nPixel);
    for (size_t c3 = 0; c3 < 3; ++c3)
      for (size_t q = 0; q < nPixel; ++q) {
        l32[c3 * nPixel + q] = lmn[q + c3 * ldlmn];
        l64[c3 * nPixel + q] = static_cast<double>(lmn[q + c3 * ldlmn]);
      }
    for (int c3 = 0; c3 < 3; ++c3) {
      const auto mm = std::minmax_element(l64.begin() + c3 * nPixel, l64.begin() + (c3 + 1) * nPixel);
      h->pix_lo[c3] = *mm.first;
      h->pix_hi[c3] = *mm.second;
    }
    double* d64 = h->lmn64.as<double>(l64.size());
    float* d32 = h->lmn32.as<float>(l32.size());
    BIPP_HIP(hipMemcpyAsync(d64, l64.data(), sizeof(double) * l64.size(), hipMemcpyHostToDevice, st));
    BIPP_HIP(hipMemcpyAsync(d32, l32.data(), sizeof(float) * l32.size(), hipMemcpyHostToDevice, st));
    const size_t isz = h->img_double() ? 8 : 4;
    void* im = h->img.get(isz * nImages * nPixel);
    BIPP_HIP(hipMemsetAsync(im, 0, isz * nImages * nPixel, st));
    BIPP_HIP(hipStreamSynchronize(st));
    ctx->retain();
    *out = h.release();
  });
}

BippError bipp_imager_destroy(BippImager h) {
  return guard([&] {
    if (!h) return;
    BippContextImpl* c = h->ctx;
    (void)hipStreamSynchronize(c->impl->stream);
    delete h;
    c->release();
  });
}

BippError bipp_imager_set_tolerance(BippImager h, double tol) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(tol > 0 && tol < 1, BIPP_INVALID_PARAMETER_ERROR, "tolerance must be in (0, 1)");
    h->opt.tol = tol;
  });
}

BippError bipp_imager_set_partition(BippImager h, int method, const size_t* dims) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    set_partition(h->opt, method, dims);
  });
}

BippError bipp_imager_set_sample_batch_size(BippImager h, size_t n) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(h->count == 0, BIPP_INVALID_PARAMETER_ERROR,
            "the sample batch size cannot change while samples are staged");
    h->batch_req = n;
    h->batch = 0;
  });
}

BippError bipp_imager_add(BippImager h, float wl, float scale, const float* uvw, size_t lduvw,
                          const void* v, size_t ldv, const float* dScaled, size_t lddScaled) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(uvw && v && dScaled, BIPP_INVALID_POINTER_ERROR, "null array");
    require(lduvw >= h->nn() && ldv >= (size_t)h->n_ant && lddScaled >= (size_t)h->n_beam,
            BIPP_INVALID_PARAMETER_ERROR, "leading dimension too small");
    require(wl > 0, BIPP_INVALID_PARAMETER_ERROR, "wavelength must be positive");
    h->add(wl, scale, uvw, lduvw, v, ldv, dScaled, lddScaled);
  });
}

BippError bipp_imager_get(BippImager h, size_t imgIdx, float* image) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    require(image != nullptr, BIPP_INVALID_POINTER_ERROR, "null image");
    require(imgIdx < (size_t)h->n_img, BIPP_INVALID_PARAMETER_ERROR, "image index out of range");
    h->flush();
    hipStream_t st = h->ctx->impl->stream;
    if (h->img_double()) {
      std::vector<double> t(h->n_pix);
      BIPP_HIP(hipMemcpyAsync(t.data(), static_cast<double*>(h->img.p) + imgIdx * h->n_pix,
                              sizeof(double) * h->n_pix, hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
      for (int q = 0; q < h->n_pix; ++q) image[q] = static_cast<float>(t[q]);
    } else {
      BIPP_HIP(hipMemcpyAsync(image, static_cast<float*>(h->img.p) + imgIdx * h->n_pix,
                              sizeof(float) * h->n_pix, hipMemcpyDeviceToHost, st));
      BIPP_HIP(hipStreamSynchronize(st));
    }
  });
}

BippError bipp_imager_num_added(BippImager h, size_t* n) {
  return guard([&] {
    require(h && n, BIPP_INVALID_HANDLE_ERROR, "null handle");
    *n = h->n_added;
  });
}

BippError bipp_imager_nufft_partition(BippImager h, int32_t* dims, int32_t* boxes) {
  return guard([&] {
    require(h != nullptr, BIPP_INVALID_HANDLE_ERROR, "null handle");
    const bool d64 = h->compute == BIPP_COMPUTE_FP64;
    if (dims)
      for (int d = 0; d < 3; ++d) dims[d] = d64 ? h->n3d.last_dims[d] : h->n3f.last_dims[d];
    if (boxes) *boxes = d64 ? h->n3d.last_boxes : h->n3f.last_boxes;
  });
}

}  // extern "C"
