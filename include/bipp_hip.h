/*
 * bipp_hip.h -- C-ABI of the MI355X (gfx950) Bluebild synthesis library libbipp_amd.so.
 *
 * Plain pointers, sizes and enums only.  Every entry point returns a BippError code
 * (include/bipp/errors.h); on failure, bipp_last_error() returns a message for the
 * calling thread.  Matrices are column-major with explicit leading dimensions, as in
 * the reference (src/eigensolver.cpp:136-141).  Complex arrays are interleaved
 * (re, im) pairs of float (complex64) or double (complex128).
 *
 * Each declaration names the reference interface it replaces.  The reference's own FFI
 * is the pybind11 module python/pybipp.cpp; its functions map onto these symbols one to
 * one (see the ctypes binding in bipp_amd/_capi.py).
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "bipp/enums.h"
#include "bipp/errors.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this thread ("" if none). */
const char* bipp_last_error(void);

/* Number of visible GPUs (0 on a host without one). */
BippError bipp_device_count(int* count);

/* Library version string. */
const char* bipp_version(void);

/* ------------------------------------------------------------------------------------ *
 * Context -- replaces bipp::Context(BippProcessingUnit) (include/bipp/context.hpp:25-36,
 * src/context_internal.cpp:27-73).  A context owns one HIP stream on the current device.
 * BIPP_PU_AUTO and BIPP_PU_GPU select the GPU; BIPP_PU_CPU is refused with
 * BIPP_INVALID_PARAMETER_ERROR, because this library computes on the GPU only.
 * Creating a context on a host without a GPU fails with BIPP_GPU_SUPPORT_ERROR.
 * ------------------------------------------------------------------------------------ */
typedef struct BippContextImpl* BippContext;

BippError bipp_ctx_create(BippProcessingUnit pu, BippContext* ctx);
BippError bipp_ctx_destroy(BippContext ctx);
BippError bipp_ctx_processing_unit(BippContext ctx, BippProcessingUnit* pu);
/* The context's hipStream_t, for callers that order their own work with it. */
BippError bipp_ctx_stream(BippContext ctx, void** stream);
BippError bipp_ctx_synchronize(BippContext ctx);

/* ------------------------------------------------------------------------------------ *
 * Eigendecomposition of one sample, host arrays -- replaces bipp::eigh_gram<T> and
 * bipp::eigh<T> (include/bipp/bipp.hpp:47-75, src/eigensolver.cpp:40-225) and the pybind
 * functions eigh_gram / eigh (python/pybipp.cpp:234-275, 462-493).
 *   s: nBeam x nBeam (lds), lower triangle read; w: nAntenna x nBeam (ldw), may be NULL
 *   for the identity; xyz: nAntenna x 3 (ldxyz).  d (nBeam) ascending, zero padded;
 *   v: nAntenna x nBeam (ldv), zero padded.  Returns the kept-beam count and 1/nVis.
 * The _s variants take complex64/float32, the _d variants complex128/float64; the
 * eigensolve runs on the GPU in the same precision.
 * ------------------------------------------------------------------------------------ */
BippError bipp_eigh_gram_s(float wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                           const void* w, size_t ldw, const float* xyz, size_t ldxyz, float* d,
                           void* v, size_t ldv, size_t* nBeamReduced, float* scale);
BippError bipp_eigh_gram_d(double wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                           const void* w, size_t ldw, const double* xyz, size_t ldxyz, double* d,
                           void* v, size_t ldv, size_t* nBeamReduced, double* scale);
BippError bipp_eigh_s(float wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                      const void* w, size_t ldw, float* d, void* v, size_t ldv,
                      size_t* nBeamReduced, float* scale);
BippError bipp_eigh_d(double wl, size_t nAntenna, size_t nBeam, const void* s, size_t lds,
                      const void* w, size_t ldw, double* d, void* v, size_t ldv,
                      size_t* nBeamReduced, double* scale);

/* ------------------------------------------------------------------------------------ *
 * Batched eigendecomposition, device arrays -- the batched form of bipp::eigh_gram
 * (src/eigensolver.cpp:129-225) for nSamples samples of one array geometry: all solves
 * run in one persistent launch with a device work queue (largest problems first).
 *   sPrec: element type of s and w (SINGLE: complex64, DOUBLE: complex128);
 *   compute: precision of the solve (SINGLE or DOUBLE); d and v are written in
 *   outPrec.  wl (double, nSamples), xyz (double, per sample nAntenna x 3).
 *   nBeamReduced (int32, nSamples) and scale (outPrec, nSamples) are device arrays.
 *   Sample k's arrays start at base + k * stride (strides in elements).
 * ------------------------------------------------------------------------------------ */
BippError bipp_eigh_gram_batched(BippContext ctx, size_t nSamples, size_t nAntenna, size_t nBeam,
                                 BippPrecision sPrec, BippPrecision compute, BippPrecision outPrec,
                                 const double* wl, const void* s, size_t lds, size_t strideS,
                                 const void* w, size_t ldw, size_t strideW, const double* xyz,
                                 size_t ldxyz, size_t strideXyz, void* d, size_t strideD, void* v,
                                 size_t ldv, size_t strideV, int32_t* nBeamReduced, void* scale);

/* Same as bipp_eigh_gram_batched with identity beamforming (w must be NULL), for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops): no
 * counterpart in the reference.  Every scratch buffer comes from one 64 MiB device array of
 * the library (one per GPU, held until the process ends), so a call needs
 *   nSamples * (c nBeam^2 + (4 + r) nBeam + 1 KiB) + (2 c + r) nBeam^2 + 4 KiB <= 64 MiB
 * (c: 16 bytes for compute DOUBLE, 8 for SINGLE; r: 8 or 4); larger requests fail with
 * BIPP_INVALID_PARAMETER_ERROR.  Samples are solved in index order on as many workgroup
 * slots as the array leaves room for.  Calls that overlap on one GPU race (they share the
 * array); no hipMalloc, hipFree or allocating library routine is called. */
BippError bipp_eigh_gram_batched_noalloc(BippContext ctx, size_t nSamples, size_t nAntenna,
                                         size_t nBeam, BippPrecision sPrec, BippPrecision compute,
                                         BippPrecision outPrec, const double* wl, const void* s,
                                         size_t lds, size_t strideS, const void* w, size_t ldw,
                                         size_t strideW, const double* xyz, size_t ldxyz,
                                         size_t strideXyz, void* d, size_t strideD, void* v,
                                         size_t ldv, size_t strideV, int32_t* nBeamReduced,
                                         void* scale);

/* ------------------------------------------------------------------------------------ *
 * Fused synthesis -- the per-sample chain of src/nufft_synthesis.cpp:74-170 with
 * stage A (src/eigensolver.cpp:129-225) fused in front, behind a collect/get interface
 * shaped on NUFFTInterface<T>::add / get_image (src/nufft_interface.hpp:13-21).
 * The reference has no such object; the collect/get names follow the NufftSynthesis
 * facade of earlier bipp releases.
 *   compute: FP64 / FP32 / MIXED (include/bipp/enums.h).
 *   imager:  0 = steering contraction (MFMA), 1 = virtual visibilities + direct NUFT sum
 *            (src/host/kernels/nuft_sum.cpp:9-31), 2 = virtual visibilities + type-3 NUFFT
 *            over all samples of a collect (replaces neonufft PlanT3,
 *            src/host/nufft.hpp:196-216, src/gpu/nufft.hpp:198-216).
 *   intervals: nLevels x 2 (lo, hi) host doubles, closed intervals on the eigenvalues
 *            rounded to float32 (python/bipp/filter.py:48-51); filter: BippFilter.
 *   lmn: nPixel x 3 host float32 (ldlmn), the ImageProp::pixel_lmn layout
 *            (include/bipp/image_prop.hpp:20).
 *   batched: 1 -> every stage runs once per collect over all its samples;
 *            0 -> one launch sequence per sample (the plain per-sample form).
 * ------------------------------------------------------------------------------------ */
typedef struct BippSynthesisImpl* BippSynthesis;

BippError bipp_synthesis_create(BippContext ctx, BippCompute compute, int imager, size_t nAntenna,
                                size_t nBeam, size_t nLevels, const double* intervals,
                                BippFilter filter, size_t nPixel, const float* lmn, size_t ldlmn,
                                int batched, BippSynthesis* out);
BippError bipp_synthesis_destroy(BippSynthesis h);

/* Adds nSamples samples.  Device arrays (same layout as bipp_eigh_gram_batched); w may be
 * NULL (identity beamforming).  wl is a device array of doubles. */
BippError bipp_synthesis_collect(BippSynthesis h, size_t nSamples, const double* wl,
                                 BippPrecision sPrec, const void* s, size_t lds, size_t strideS,
                                 const void* w, size_t ldw, size_t strideW, const double* xyz,
                                 size_t ldxyz, size_t strideXyz);
/* Per-timestep chain captured in a HIP graph: the same as bipp_synthesis_collect for the
 * steering imager with identity beamforming (w = NULL), with no host round trip inside the
 * chain (samples solved in index order; a failed solve keeps the whole batch out of the
 * images and returns BIPP_EIGENSOLVER_ERROR).  The first call for an argument set captures
 * flag -> eigensolve -> level weights -> steering image into a graph; later calls with the
 * same pointers, sizes and strides replay it, so a caller refills the same device buffers
 * for every timestep.  No counterpart in the reference. */
BippError bipp_synthesis_collect_graph(BippSynthesis h, size_t nSamples, const double* wl,
                                       BippPrecision sPrec, const void* s, size_t lds,
                                       size_t strideS, const double* xyz, size_t ldxyz,
                                       size_t strideXyz);
/* Graph captures and launches of bipp_synthesis_collect_graph on this handle. */
BippError bipp_synthesis_graph_info(BippSynthesis h, long long* captures, long long* launches);
/* Same as bipp_synthesis_collect with host arrays (copied to the device first). */
BippError bipp_synthesis_collect_host(BippSynthesis h, size_t nSamples, const double* wl,
                                      BippPrecision sPrec, const void* s, size_t lds,
                                      size_t strideS, const void* w, size_t ldw, size_t strideW,
                                      const double* xyz, size_t ldxyz, size_t strideXyz);
/* Image of level `level` (nPixel floats, host), divided by the number of collected samples
 * when normalize != 0 (src/image_synthesis.cpp:152-161). */
BippError bipp_synthesis_get(BippSynthesis h, size_t level, int normalize, float* image);
/* Device pointer and element type (SINGLE: float, DOUBLE: double) of the image
 * accumulator (nLevels x nPixel, un-normalised). */
BippError bipp_synthesis_image_device(BippSynthesis h, void** ptr, BippPrecision* type);
BippError bipp_synthesis_num_collected(BippSynthesis h, size_t* n);
BippError bipp_synthesis_reset(BippSynthesis h);
/* Eigenvalues (float64, nBeam per sample, ascending, zero padded) and kept-beam counts of
 * all samples of the last collect call, host arrays with room for `capacity` samples
 * (BIPP_INVALID_PARAMETER_ERROR if the call had more).  *count (may be NULL) receives the
 * number of samples; d and nBeamReduced may both be NULL to query it. */
BippError bipp_synthesis_last_eigenvalues(BippSynthesis h, size_t capacity, double* d,
                                          int32_t* nBeamReduced, size_t* count);
/* GPU time per stage accumulated over collect calls, from HIP events on the context stream
 * (ms must hold 9 doubles): ms[0] eigensolve (flag, Gram, solve), ms[1] level weights /
 * operand preparation, ms[2] imaging kernels, ms[3] number of collect batches timed; for the
 * NUFFT imager ms[4] virtual visibilities, ms[5] point binning and sort, ms[6] spreading,
 * ms[7] deconvolution + FFT, ms[8] interpolation.  The reference's analogue is its
 * wall-clock timer tree (src/rt_graph.hpp:94-161).  reset != 0 clears the counters. */
BippError bipp_synthesis_timings(BippSynthesis h, double* ms, int reset);
/* NUFFT tolerance of imager 2 (NufftSynthesisOptions::tolerance,
 * include/bipp/image_synthesis.hpp:61-65); default 1e-5. */
BippError bipp_synthesis_set_tolerance(BippSynthesis h, double tol);
/* Inner FFT grid (nf2 per axis) of the last NUFFT collect (its first uvw box), for
 * inspection. */
BippError bipp_synthesis_nufft_grid(BippSynthesis h, int32_t* nf);

/* uvw-domain partition of the type-3 NUFFT -- replaces NufftSynthesisOptions::
 * localUVWPartition (include/bipp/image_synthesis.hpp:23-53,55-148; applied at
 * src/gpu/nufft.hpp:110-125).  The points of a batch are split into a grid of boxes in uvw,
 * each transformed by a plan of its own extent:
 *   BIPP_PARTITION_AUTO: the fewest boxes (grown per axis as src/nufft_util.hpp:12-50 does)
 *     for which one box's plan fits the FFT limits and two thirds of the device memory;
 *   BIPP_PARTITION_NONE: one plan (an error if it exceeds the limits);
 *   BIPP_PARTITION_GRID: dims[0] x dims[1] x dims[2] boxes. */
enum { BIPP_PARTITION_AUTO = 0, BIPP_PARTITION_NONE = 1, BIPP_PARTITION_GRID = 2 };
BippError bipp_synthesis_set_partition(BippSynthesis h, int method, const size_t* dims);
/* Boxes per axis and boxes holding points of the last NUFFT collect. */
BippError bipp_synthesis_nufft_partition(BippSynthesis h, int32_t* dims, int32_t* boxes);
/* Autotuning of the type-3 NUFFT's run-time shapes (no counterpart in the reference, whose
 * NUFFT comes from neonufft): with enable != 0, the first batch of every new grid shape times
 * each candidate shape -- largest FFT radix of the x lines and of the strided passes, strided
 * pass item width, spread blocks per XCD run -- on its own spread and FFT, and keeps the
 * fastest for that shape in a process-wide table used by every later batch (and by other
 * handles) whether or not they tune.  Off by default (measured defaults). */
BippError bipp_synthesis_set_autotune(BippSynthesis h, int enable);
/* Shape used by the last NUFFT batch, shape[4] = (lines radix, pass radix, log2 pass item
 * width, spread chunk), and the candidates timed by the last tuning of this handle:
 * 5 doubles each (the 4 shape values and milliseconds), up to capacity of them;
 * *n_candidates receives their number (0 if this handle has not tuned). */
BippError bipp_synthesis_nufft_tuning(BippSynthesis h, int32_t* shape, size_t capacity,
                                      double* candidates, size_t* n_candidates);
/* Forgets every tuned shape of the process. */
BippError bipp_nufft_tuning_reset(void);
/* Samples per batch (NufftSynthesisOptions::sampleBatchSize, src/nufft_synthesis.cpp:97-103);
 * 0: automatic (4096; the NUFFT imager min(4096, 2e9 / (20 nAntenna^2))). */
BippError bipp_synthesis_set_sample_batch_size(BippSynthesis h, size_t n);

/* ------------------------------------------------------------------------------------ *
 * Eigenpair imager -- stage B of bipp::image_synthesis on eigenpairs a dataset holds
 * (include/bipp/image_synthesis.hpp:160-163, src/image_synthesis.cpp:28-182): the per-sample
 * loop of nufft_synthesis (src/nufft_synthesis.cpp:74-170) behind NUFFTInterface::add /
 * get_image (src/nufft_interface.hpp:13-21).  The reference's add() takes virtual
 * visibilities formed on the host; this one takes the sample's eigenpairs as Dataset hands
 * them out (include/bipp/dataset.hpp:14-41) and forms them on the GPU.
 *   imager: 1 = direct sum (src/host/kernels/nuft_sum.cpp:9-31), 2 = type-3 NUFFT.
 *   compute: FP64 (double everywhere), MIXED (double virtual visibilities and point
 *            coordinates, float grid), FP32.
 *   lmn: nPixel x 3 host float32 (ldlmn), ImageProp::pixel_lmn (include/bipp/image_prop.hpp:20).
 * Samples are staged in pinned memory and imaged a batch at a time (sample batch size as in
 * bipp_synthesis_set_sample_batch_size); get() images what is staged first.
 * ------------------------------------------------------------------------------------ */
typedef struct BippImagerImpl* BippImager;

BippError bipp_imager_create(BippContext ctx, BippCompute compute, int imager, size_t nAntenna,
                             size_t nBeam, size_t nImages, size_t nPixel, const float* lmn,
                             size_t ldlmn, BippImager* out);
BippError bipp_imager_destroy(BippImager h);
/* NufftSynthesisOptions::tolerance of the type-3 NUFFT; default 1e-5. */
BippError bipp_imager_set_tolerance(BippImager h, double tol);
/* As bipp_synthesis_set_partition. */
BippError bipp_imager_set_partition(BippImager h, int method, const size_t* dims);
/* As bipp_synthesis_set_sample_batch_size; only while no sample is staged. */
BippError bipp_imager_set_sample_batch_size(BippImager h, size_t n);
/* One sample, host arrays: uvw nAntenna^2 x 3 float (lduvw) in metres, unscaled (entry
 * i + nAntenna j is the baseline of antennas i, j; the imager scales by 2 pi / wl,
 * src/nufft_synthesis.cpp:149-155); v nAntenna x nBeam complex64 (ldv); dScaled: for each
 * of the nImages images nBeam floats (lddScaled apart), the selected, filtered eigenvalues
 * (zero: not used, src/host/virtual_vis.cpp:36-42); scale: the sample's 1/nVis. */
BippError bipp_imager_add(BippImager h, float wl, float scale, const float* uvw, size_t lduvw,
                          const void* v, size_t ldv, const float* dScaled, size_t lddScaled);
/* Image imgIdx (nPixel floats, host) summed over every sample added, not normalised. */
BippError bipp_imager_get(BippImager h, size_t imgIdx, float* image);
BippError bipp_imager_num_added(BippImager h, size_t* n);
/* Boxes per axis and boxes holding points of the last NUFFT batch. */
BippError bipp_imager_nufft_partition(BippImager h, int32_t* dims, int32_t* boxes);

/* ------------------------------------------------------------------------------------ *
 * Kernel-level entry points on device arrays (parity tests and custom pipelines).
 * ------------------------------------------------------------------------------------ */
/* VV = V diag(dw) V^H per sample: src/host/virtual_vis.cpp:19-73.  prec selects
 * complex64/float or complex128/double for v, dw and vv; vv is nAntenna^2 per sample. */
BippError bipp_virtual_vis(BippContext ctx, BippPrecision prec, size_t nSamples, size_t nAntenna,
                           size_t nEig, const void* v, size_t ldv, size_t strideV, const void* dw,
                           size_t strideDw, void* vv, size_t strideVv);

/* out[q] += sum_j Re(c_j exp(i alpha (x_q u_j + y_q v_j + z_q w_j))):
 * src/host/kernels/nuft_sum.cpp:9-31 and src/gpu/kernels/nuft_sum.cu:24-97.
 * prec SINGLE/DOUBLE: everything in that type.  bipp_nuft_sum_mixed: coordinates in
 * double, phases reduced to one turn in double, float32 sin/cos and sums. */
BippError bipp_nuft_sum(BippContext ctx, BippPrecision prec, double alpha, size_t nIn,
                        const void* input, const void* u, const void* v, const void* w,
                        size_t nOut, const void* x, const void* y, const void* z, void* out);
BippError bipp_nuft_sum_mixed(BippContext ctx, double alpha, size_t nIn, const void* input,
                              const double* u, const double* v, const double* w, size_t nOut,
                              const double* x, const double* y, const double* z, float* out);

/* Sinc Gram matrix of each sample's array, G_ij = prefactor sinc(2|p_i - p_j|/wl)
 * (src/host/gram_matrix.cpp:23-56; the C++ reference uses prefactor 1,
 * python/bipp/gram.py:110 uses 4 pi).  xyz double nAntenna x 3 per sample. */
BippError bipp_gram(BippContext ctx, BippPrecision prec, size_t nSamples, size_t nAntenna,
                    const double* wl, const double* xyz, size_t ldxyz, size_t strideXyz,
                    double prefactor, void* g, size_t ldg, size_t strideG);

#ifdef __cplusplus
}
#endif
