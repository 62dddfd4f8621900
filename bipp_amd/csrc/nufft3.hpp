// Parameter blocks and launchers of the type-3 NUFFT (nufft3.hip) and its host plan.
#pragma once

#include "common.hpp"

namespace bippamd {

struct N3Radix {
  int count = 0;
  int r[20] = {};
};

struct N3PointsParams {
  int n_ant_max;                 // rows per sample (pairs a >= b < n_ant_max)
  const int* n_rows;             // active rows per sample, or nullptr -> n_rows_const
  int n_rows_const;
  const double* pos;             // [nS][3][apad] positions / wavelength
  long long pos_stride;
  int apad;
  double C[3], D[3], inv_gh[3];  // centres and 1/(gamma h1)
  int w;
  int off[3], bin_size[3], nbins[3];
  int phase_fp32;
  long long ld_vv;
  // outputs, [nS * npair]
  unsigned char* valid;
  int* bin_of;
  int *l0x, *l0y, *l0z;
  float *frx, *fry, *frz, *phr, *phi, *mult;
  int* src;
  unsigned* bin_count;
};

template <typename R>
struct N3SpreadParams {
  int n_levels, w;
  double beta;
  int bin_size[3], nbins[3], off[3], nf2[3];
  const unsigned *bin_start, *bin_count, *sorted;
  const int *l0x, *l0y, *l0z, *src;
  const float *frx, *fry, *frz, *phr, *phi, *mult;
  long long npair;
  const cplx<R>* vv;             // [nS][L][vv_stride]
  long long vv_stride;
  R* grid;                       // [L][grid_stride][2]
  long long grid_stride;
};

template <typename R>
struct N3DeconvParams {
  int nf1[3], nf2[3];
  const R *fx, *fy, *fz;         // 1/phi2^(l) for l = -nf1/2 .. nf1/2-1
  R* grid;
  long long grid_stride;
};

template <typename R>
struct N3StridedParams {
  cplx<R>* grid;
  long long nx;                  // contiguous extent tiled by 16
  long long axis_stride;         // elements between consecutive axis indices
  long long n_outer, outer_stride;
  int m;                         // transform length
  long long estride, sub_stride;
  int n_sub;
  int twiddle;
  long long ntot;
  N3Radix rad;
};

template <typename R>
struct N3InterpParams {
  int n_pix, n_levels, w1, w2;
  double beta1, beta2;
  double gam[3], D[3], C[3], h1[3], h2[3];
  int nf2[3], n1[3], n2[3];
  const double* lmn;             // [3][n_pix]
  const double *gl_x, *gl_w;
  int nq;
  const cplx<R>* grid;
  long long grid_stride;
  void* img;                     // [L][img_ld], float or double
  double scale;
  long long img_ld;
};

template <typename R>
struct N3VVParams {
  int n_levels, n_cols;
  const int* n_rows;
  int n_rows_const;
  const cplx<R>* z;
  long long stride_z;
  int ldz;
  const R* wd;                   // [nS][L][n_cols]
  void* vv;
  long long vv_stride;
  long long ld_vv;
};

void launch_pos_range(int n_samples, const double* pos, long long pos_stride, int apad,
                      const int* n_rows, int n_rows_const, double* out, hipStream_t st);
template <typename R>
void launch_zero(R* p, long long n, hipStream_t st);
template <typename R>
void launch_n3_points(const N3PointsParams& p, int n_samples, hipStream_t st);
void launch_scan(const unsigned* in, unsigned* out, unsigned* tmp, int n, hipStream_t st);
void launch_n3_scatter(long long n_pts, const unsigned char* valid, const int* bin_of,
                       const unsigned* bin_start, unsigned* bin_fill, unsigned* sorted,
                       hipStream_t st);
template <typename R, typename TV>
void launch_n3_spread(const N3SpreadParams<R>& p, int n_bins, hipStream_t st);
template <typename R>
void launch_n3_deconv(const N3DeconvParams<R>& p, int n_levels, hipStream_t st);
template <typename R>
void launch_fft_x(cplx<R>* grid, int n, long long n_lines, const N3Radix& rad, hipStream_t st);
template <typename R>
void launch_fft_strided(const N3StridedParams<R>& p, hipStream_t st);
template <typename R, typename TI>
void launch_n3_interp(const N3InterpParams<R>& p, hipStream_t st);
template <typename R, typename TO>
void launch_vv_levels(const N3VVParams<R>& p, int n_samples, int n_rows_max, hipStream_t st);

}  // namespace bippamd
