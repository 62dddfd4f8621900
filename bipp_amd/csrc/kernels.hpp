// Kernel parameter blocks and launchers shared between the HIP kernels and the host
// driver.  Everything here works on device pointers and a caller-supplied stream.
#pragma once

#include "common.hpp"

namespace bippamd {

// ---------------------------------------------------------------------------------------
// Stage A
// ---------------------------------------------------------------------------------------
struct FlagParams {
  int n_beam;
  int mode;  // 0: |s|^2 > eps (eigh_gram), 1: |Re|>=eps || |Im|>=eps (eigh)
  const void* s;
  long long lds, stride_s;
  int* n_red;
  int* idx;
  long long idx_stride;
  unsigned long long* n_vis;
};

template <typename R>
struct EigParams {
  int n_jobs = 0;
  int nmax = 0;
  const int* order = nullptr;
  unsigned* counter = nullptr;
  const int* n_red = nullptr;
  const int* idx = nullptr;
  long long idx_stride = 0;
  const void* s = nullptr;
  long long lds = 0, stride_s = 0;
  const double* xyz = nullptr;
  long long ldxyz = 0, stride_xyz = 0;
  const double* wl = nullptr;
  const cplx<R>* g_in = nullptr;
  long long stride_g = 0;
  int ld_g = 0;
  double gram_prefactor = 1.0;
  int standard = 0;
  R* d_out = nullptr;
  long long stride_d = 0;
  cplx<R>* z_out = nullptr;
  long long stride_z = 0;
  int ldz = 0;
  cplx<R>* ws_a = nullptr;
  cplx<R>* ws_l = nullptr;
  R* ws_z = nullptr;
  long long ws_stride = 0;
  int ws_ld = 0;
  int* info = nullptr;
  // optional profile (wall clock ticks), EIG_PROF_SLOTS per sample: [0, 8) phase-boundary
  // timestamps, [8, 16) divide-and-conquer step sums, [16, 18) tridiagonalisation step sums,
  // [18, 22) Cholesky step sums
  unsigned long long* prof = nullptr;
};
constexpr int EIG_PROF_SLOTS = 24;

template <typename R>
struct UnbeamParams {
  int n_ant;
  const int* n_red;
  const int* idx;
  long long idx_stride;
  const cplx<R>* z;
  long long stride_z;
  int ldz;
  void* v;
  long long stride_v, ldv;
};

void launch_fill_u32(unsigned* p, unsigned v, int n, hipStream_t stream);
void launch_flag(const FlagParams& p, int n_samples, bool in_double, hipStream_t stream);
template <typename R>
void launch_eig(const EigParams<R>& p, int n_slots, bool in_double, hipStream_t stream);
template <typename R>
void launch_unbeam_identity(const UnbeamParams<R>& p, int n_samples, int n_beam, bool out_double,
                            hipStream_t stream);
template <typename R>
void launch_cast_d(int n_samples, int nmax, const R* d, long long stride_d, void* out,
                   long long stride_out, bool out_double, const unsigned long long* n_vis,
                   void* scale, hipStream_t stream);

// General beamforming W: batched complex GEMM C = alpha op(A) op(B) + beta C over ragged
// sizes (float64 or float32), and the sinc matrix of the full array.
template <typename R>
struct GemmBatch {
  int n_batch;
  const int* m;  // per batch, or nullptr -> m0
  const int* n;
  const int* k;
  int m0, n0, k0;
  int op_a, op_b;  // 0: N, 1: C (conjugate transpose)
  cplx<R> alpha, beta;
  const cplx<R>* a;
  long long lda, stride_a;
  const cplx<R>* b;
  long long ldb, stride_b;
  cplx<R>* c;
  long long ldc, stride_c;
};
template <typename R>
void launch_gemm_batch(const GemmBatch<R>& g, int max_m, int max_n, hipStream_t stream);

template <typename R>
void launch_sinc_matrix(int n_samples, int n_ant, const double* xyz, long long ldxyz,
                        long long stride_xyz, const double* wl, double prefactor, cplx<R>* out,
                        long long ld, long long stride, hipStream_t stream);
// W_I = W[:, idx[:n_red]] (zero beyond), per sample
template <typename R, typename TIn>
void launch_gather_w(int n_samples, int n_ant, int n_beam, const void* w, long long ldw,
                     long long stride_w, const int* n_red, const int* idx, long long idx_stride,
                     cplx<R>* out, long long ld, long long stride, hipStream_t stream);

// ---------------------------------------------------------------------------------------
// Stage B: steering contraction
// ---------------------------------------------------------------------------------------
enum SteerMode { STEER_BF16X3 = 0, STEER_FP32 = 1, STEER_FP64 = 2 };

// Row tiles are 16 output rows (8 eigenvectors x Re/Im); a k-chunk is 16 antennas.
constexpr int STEER_TP = 128;      // pixels per workgroup
constexpr int STEER_NT = 512;      // threads per workgroup (8 waves)
constexpr int STEER_RT_PASS = 32;  // row tiles per pass (256 eigenvectors)

template <typename R>
struct SteerPrepParams {
  int n_samples;
  int n_levels;
  int filter;
  const double* intervals;  // [n_levels][2]
  // eigen-decomposition of each sample (reduced basis)
  const int* n_red;         // rows of Z used / antennas used
  const int* idx;           // kept antenna of each reduced row (W = I), or nullptr
  long long idx_stride;
  const R* d;               // eigenvalues [nS][stride_d]
  long long stride_d;
  const cplx<R>* z;         // eigenvectors, rows = antennas used (ld = ldz)
  long long stride_z;
  int ldz;
  int n_cols;               // eigenvector columns per sample (nmax)
  const unsigned long long* n_vis;  // scale = 1/n_vis
  const double* xyz;        // [nS] column-major N x 3
  long long ldxyz, stride_xyz;
  const double* wl;
  // outputs
  int* n_eig;               // eigenvectors with a nonzero weight
  int* n_ant;               // antennas (rows) used
  double* pos;              // [nS][3][apad] positions / wavelength
  long long pos_stride;
  int apad;                 // antennas padded to 16
  void* afrag;              // operand fragments
  long long afrag_stride;   // per sample, in fragment elements (see steering.hip)
  int kchunks;              // apad / 16
  int rtiles;               // row tiles allocated per sample
  void* wts;                // weights [nS][n_levels][epad] (float, or double for FP64)
  long long w_stride;
  int epad;
  int mode;
  int n_rows_const = 0;     // > 0: rows of z are all n_rows_const antennas (general W)
};

struct SteerParams {
  int n_samples;
  int n_levels;
  int n_pix;
  const int* n_eig;
  const int* n_ant;
  const double* pos;
  long long pos_stride;
  int apad;
  const void* afrag;
  long long afrag_stride;
  int kchunks;
  int rtiles;
  const void* wts;
  long long w_stride;
  int epad;
  const double* lmn;  // [3][n_pix]
  int phase_fp32;     // plain form: phases formed in float32
  void* img;          // [n_levels][img_ld] float (double for FP64 mode), accumulated
  long long img_ld;
};

template <typename R>
void launch_steer_prep(const SteerPrepParams<R>& p, hipStream_t stream);
void launch_steer(const SteerParams& p, int mode, hipStream_t stream);
// flag[0] = any info != 0; img += tmp (n elements, float or double) unless flagged
void launch_commit_image(const int* info, int n_samples, const void* tmp, void* img, long long n,
                         bool dbl, int* flag, hipStream_t stream);
void launch_zero_image(void* p, long long n, bool dbl, hipStream_t stream);
size_t steer_frag_elems_per_sample(int mode, int rtiles, int kchunks);

template <typename R>
void launch_level_weights(int n_samples, int n_levels, int filter, const double* intervals,
                          const int* n_red, const R* d, long long stride_d, int n_cols,
                          const unsigned long long* n_vis, R* out, long long stride_out,
                          hipStream_t stream);
template <typename R>
int eig_blocks_per_cu(int nmax, bool in_double);

// ---------------------------------------------------------------------------------------
// Virtual visibilities and the direct NUFT sum
// ---------------------------------------------------------------------------------------
template <typename R>
void launch_virtual_vis(int n_samples, int n_ant, int n_eig_cols, const cplx<R>* v, long long ldv,
                        long long stride_v, const R* dw, long long stride_dw, cplx<R>* vv,
                        long long stride_vv, hipStream_t stream);

template <typename R>
void launch_nuft_sum(R alpha, int n_in, const cplx<R>* input, const R* u, const R* v, const R* w,
                     int n_out, const R* x, const R* y, const R* z, R* out, hipStream_t stream);
// mixed: float64 phase reduced to one turn, float32 sin/cos and accumulation
void launch_nuft_sum_mixed(double alpha, int n_in, const cplx<float>* input, const double* u,
                           const double* v, const double* w, int n_out, const double* x,
                           const double* y, const double* z, float* out, hipStream_t stream);

template <typename R>
void launch_baselines(int n_ant, const double* xyz, long long ldxyz, double scale, R* u, R* v, R* w,
                      hipStream_t stream);
template <typename TS, typename TD>
void launch_copy_cast(int n_batch, int rows, int cols, const void* src, long long lds,
                      long long stride_s, void* dst, long long ldd, long long stride_d,
                      hipStream_t stream);

}  // namespace bippamd
