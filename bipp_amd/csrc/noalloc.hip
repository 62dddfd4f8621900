// Batched generalized eigensolve without device allocation.
//
// The same kernels as bipp_eigh_gram_batched (flag, eig_kernel, export), with every scratch
// buffer -- per-sample kept-beam lists, counts, eigenvalues and reduced eigenvectors, the
// per-workgroup matrix slots of the persistent eigensolver and its job counter -- carved from
// one file-scope device array of a fixed size instead of hipMalloc'd workspaces.  The
// persistent launch takes as many workgroup slots as the array leaves room for.
//
// The array is one per device and shared by every call on that device: calls that overlap
// on one GPU race (each call clears what must start at zero, nothing else survives a call).
#include <hip/hip_runtime.h>

#include <mutex>

#include "bipp_hip.h"
#include "eig_host.hpp"
#include "host_runtime.hpp"
#include "kernels.hpp"

namespace bippamd {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device

// the no-allocation eigensolve's workspace (one per device, held until the process ends)
__device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;

// Device address of `scratch` on each device (the host-side name is only a handle).
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_on(int device) {
  require(device >= 0 && device < kMaxDevices, BIPP_INVALID_PARAMETER_ERROR,
          "device index beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    BIPP_HIP(hipGetDevice(&cur));
    if (cur != device) BIPP_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) BIPP_HIP(hipSetDevice(cur));
    BIPP_HIP(e);
    require(p != nullptr, BIPP_GPU_ERROR, "no device address for the scratch array");
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

// out[0] = first sample with info != 0 (or -1), out[1] = its info
__global__ void first_failure_kernel(const int* info, int n, int* out) {
  __shared__ int first;
  if (threadIdx.x == 0) first = n;
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    if (info[i] != 0) atomicMin(&first, i);
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = first < n ? first : -1;
    out[1] = first < n ? info[first] : 0;
  }
}

// bump allocator over the scratch array, 256-byte aligned pieces
struct Carve {
  unsigned char* base;
  size_t used = 0, cap;
  template <typename T>
  T* take(size_t n) {
    const size_t off = (used + 255) & ~size_t(255);
    used = off + sizeof(T) * n;
    return reinterpret_cast<T*>(base + off);
  }
};

template <typename R>
void eig_noalloc(ContextImpl& c, int nS, int nAnt, int nBeam, bool s_double, bool out_double,
                 const double* wl, const void* s, long long lds, long long stride_s,
                 const double* xyz, long long ldxyz, long long stride_xyz, void* d,
                 long long stride_d, void* v, long long ldv, long long stride_v,
                 int32_t* n_red_out, void* scale) {
  const int nmax = nBeam;
  const long long nn = (long long)nmax * nmax;
  // bytes of the per-sample part and of one workgroup slot
  const size_t per_sample = sizeof(int) * (2 + nmax) + sizeof(unsigned long long) +
                            sizeof(R) * nmax + sizeof(cplx<R>) * nn + 1024;
  const size_t per_slot = (2 * sizeof(cplx<R>) + sizeof(R)) * nn + 768;
  const size_t fixed = 4096;
  const size_t need = fixed + per_sample * nS + per_slot;
  if (need > kScratchBytes)
    fail(BIPP_INVALID_PARAMETER_ERROR,
         std::to_string(nS) + " samples of " + std::to_string(nBeam) +
             " beams need " + std::to_string(need >> 20) +
             " MiB of scratch, more than the 64 MiB of the no-allocation eigensolver; "
             "use bipp_eigh_gram_batched or fewer samples per call");
  hipStream_t st = c.stream;
  Carve cv{static_cast<unsigned char*>(scratch_on(c.device)), 0, kScratchBytes};
  unsigned* counter = cv.take<unsigned>(1);
  int* fail_out = cv.take<int>(2);
  int* info = cv.take<int>(nS);
  int* n_red = cv.take<int>(nS);
  int* idx = cv.take<int>((size_t)nS * nmax);
  unsigned long long* n_vis = cv.take<unsigned long long>(nS);
  R* dd = cv.take<R>((size_t)nS * nmax);
  cplx<R>* z = cv.take<cplx<R>>((size_t)nS * nn);
  const size_t left = kScratchBytes - ((cv.used + 255) & ~size_t(255));
  const int room = static_cast<int>(std::min<size_t>(left / per_slot, 1 << 20));
  const int per_cu = eig_blocks_per_cu<R>(nmax, s_double);
  const int n_slots = std::max(1, std::min({nS, per_cu * c.n_cu, room}));
  cplx<R>* ws_a = cv.take<cplx<R>>((size_t)n_slots * nn);
  cplx<R>* ws_l = cv.take<cplx<R>>((size_t)n_slots * nn);
  R* ws_z = cv.take<R>((size_t)n_slots * nn);
  if (cv.used > kScratchBytes) fail(BIPP_INTERNAL_ERROR, "no-allocation scratch layout overflow");

  FlagParams fp;
  fp.n_beam = nBeam;
  fp.mode = 0;
  fp.s = s;
  fp.lds = lds;
  fp.stride_s = stride_s;
  fp.n_red = n_red;
  fp.idx = idx;
  fp.idx_stride = nmax;
  fp.n_vis = n_vis;
  launch_flag(fp, nS, s_double, st);

  EigParams<R> ep;
  ep.n_jobs = nS;
  ep.nmax = nmax;
  ep.order = nullptr;  // samples in index order (no host round trip to sort them)
  ep.counter = counter;
  launch_fill_u32(counter, 0u, 1, st);
  ep.n_red = n_red;
  ep.idx = idx;
  ep.idx_stride = nmax;
  ep.s = s;
  ep.lds = lds;
  ep.stride_s = stride_s;
  ep.xyz = xyz;
  ep.ldxyz = ldxyz;
  ep.stride_xyz = stride_xyz;
  ep.wl = wl;
  ep.g_in = nullptr;
  ep.stride_g = nn;
  ep.ld_g = nmax;
  ep.gram_prefactor = 1.0;
  ep.standard = 0;
  ep.d_out = dd;
  ep.stride_d = nmax;
  ep.z_out = z;
  ep.stride_z = nn;
  ep.ldz = nmax;
  ep.ws_ld = nmax;
  ep.ws_stride = nn;
  ep.ws_a = ws_a;
  ep.ws_l = ws_l;
  ep.ws_z = ws_z;
  ep.info = info;
  launch_eig<R>(ep, n_slots, s_double, st);

  // results into the caller's arrays (src/eigensolver.cpp:173-174,214-222)
  launch_cast_d<R>(nS, nBeam, dd, nmax, d, stride_d, out_double, n_vis, scale, st);
  if (v) {
    UnbeamParams<R> up;
    up.n_ant = nAnt;
    up.n_red = n_red;
    up.idx = idx;
    up.idx_stride = nmax;
    up.z = z;
    up.stride_z = nn;
    up.ldz = nmax;
    up.v = v;
    up.stride_v = stride_v;
    up.ldv = ldv;
    launch_unbeam_identity<R>(up, nS, nBeam, out_double, st);
  }
  if (n_red_out)
    BIPP_HIP(hipMemcpyAsync(n_red_out, n_red, sizeof(int) * nS, hipMemcpyDeviceToDevice, st));
  first_failure_kernel<<<1, 256, 0, st>>>(info, nS, fail_out);
  check_launch("first_failure_kernel");
  int res[2] = {-1, 0};
  BIPP_HIP(hipMemcpyAsync(res, fail_out, sizeof(res), hipMemcpyDeviceToHost, st));
  c.sync();
  if (res[0] >= 0)
    fail(BIPP_EIGENSOLVER_ERROR, "generalized eigensolve failed for sample " +
                                     std::to_string(res[0]) + " (info " + std::to_string(res[1]) + ")");
}

}  // namespace
}  // namespace bippamd

using namespace bippamd;

extern "C" BippError bipp_eigh_gram_batched_noalloc(
    BippContext ctx, size_t nSamples, size_t nAntenna, size_t nBeam, BippPrecision sPrec,
    BippPrecision compute, BippPrecision outPrec, const double* wl, const void* s, size_t lds,
    size_t strideS, const void* w, size_t ldw, size_t strideW, const double* xyz, size_t ldxyz,
    size_t strideXyz, void* d, size_t strideD, void* v, size_t ldv, size_t strideV,
    int32_t* nBeamReduced, void* scale) {
  (void)ldw;
  (void)strideW;
  return guard([&] {
    require(ctx != nullptr, BIPP_INVALID_HANDLE_ERROR, "null context");
    require(s && xyz && wl && d, BIPP_INVALID_POINTER_ERROR, "null array");
    require(w == nullptr, BIPP_INVALID_PARAMETER_ERROR,
            "the no-allocation eigensolve takes identity beamforming only (w = NULL); "
            "use bipp_eigh_gram_batched for a beamforming matrix");
    require(nBeam >= 1 && nBeam <= 512, BIPP_INVALID_PARAMETER_ERROR,
            "number of beams must be in [1, 512]");
    require(nAntenna == nBeam, BIPP_INVALID_PARAMETER_ERROR,
            "identity beamforming needs nAntenna == nBeam");
    require(lds >= nBeam && ldxyz >= nAntenna && (!v || ldv >= nAntenna),
            BIPP_INVALID_PARAMETER_ERROR, "leading dimension too small");
    require(nSamples < (1ULL << 24), BIPP_INVALID_PARAMETER_ERROR, "too many samples");
    if (nSamples == 0) return;
    ContextImpl& c = *ctx->impl;
    const bool sd = sPrec == BIPP_PRECISION_DOUBLE, od = outPrec == BIPP_PRECISION_DOUBLE;
    if (compute == BIPP_PRECISION_DOUBLE)
      eig_noalloc<double>(c, (int)nSamples, (int)nAntenna, (int)nBeam, sd, od, wl, s, lds, strideS,
                          xyz, ldxyz, strideXyz, d, strideD, v, ldv, strideV, nBeamReduced, scale);
    else
      eig_noalloc<float>(c, (int)nSamples, (int)nAntenna, (int)nBeam, sd, od, wl, s, lds, strideS,
                         xyz, ldxyz, strideXyz, d, strideD, v, ldv, strideV, nBeamReduced, scale);
  });
}
