// Parameter blocks and launchers of the type-3 NUFFT (nufft3.hip) and its host plan.
#pragma once

#include "common.hpp"

namespace bippamd {

// Multiplier m and shift s with x / d = (umulhi(x, m) + x) >> s for 0 <= x < 2^31, d >= 1:
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
inline void n3_div_magic(int d, unsigned& m, int& s) {
  s = 0;
  while ((1LL << s) < d) ++s;
  m = static_cast<unsigned>((((1ULL << 32) * ((1ULL << s) - static_cast<unsigned long long>(d))) /
                             static_cast<unsigned long long>(d)) + 1ULL);
}

struct N3Radix {
  int count = 0;
  int r[20] = {};
  // Division by run-time constants as q = (umulhi(x, m) + x) >> s, exact for 0 <= x < 2^31
  // (s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1): the length n, and per stage the
  // butterfly count n / r and the span ns = r[0] ... r[st-1].
  unsigned n_m = 1;
  int n_s = 0;
  unsigned nb_m[20] = {}, ns_m[20] = {};
  int nb_s[20] = {}, ns_s[20] = {};
};

// One non-uniform point of the spread (32 bytes; its level values are stored beside it).
struct alignas(16) N3Point {
  int l0[3];    // first cell of the kernel support, spread index + nf1/2
  float fr[3];  // cell - (l0 - nf1/2) per axis, in [w/2, w/2 + 1)
  float ph[2];  // multiplicity times the pre-phase exp(i u'.D) (2 for a pair whose mirror
                // point is its conjugate, else 1)
};
static_assert(sizeof(N3Point) == 32, "point records are 32 bytes");

// Where the input points of a batch come from.
//   N3_SRC_PAIRS: antenna pairs a >= b of each sample, u = 2 pi (p_a - p_b) from positions
//     in wavelengths; slot t of a sample runs down the columns of the lower triangle,
//     t = b (2 n - b + 1) / 2 + (a - b) for n = n_ant_max rows (consecutive slots read
//     consecutive entries of the sample's VV matrices).
//   N3_SRC_UVW:   every entry (i, j) of the sample's nAnt^2 x 3 uvw array (metres, scaled by
//     2 pi / wl), slot t = i + nAnt j, valued VV(i, j).  Where the sample's uvw is
//     antisymmetric (uvw(j, i) = -uvw(i, j), zero diagonal: baselines p_i - p_j), the pairs
//     i > j stand for both entries (multiplicity 2) and the diagonal is one point at 0.
enum { N3_SRC_PAIRS = 0, N3_SRC_UVW = 1 };

struct N3Source {
  int mode = N3_SRC_PAIRS;
  int n_ant_max = 0;                  // rows per sample
  const int* n_rows = nullptr;        // active rows per sample, or nullptr -> n_rows_const
  int n_rows_const = 0;
  const double* pos = nullptr;        // pairs: [nS][3][apad] positions / wavelength
  long long pos_stride = 0;
  int apad = 0;
  const float* uvw = nullptr;         // uvw: [nS] nAnt^2 x 3 column-major (ld lduvw), metres
  long long uvw_stride = 0, lduvw = 0;
  const double* uvw_scale = nullptr;  // [nS] 2 pi / wl
  const unsigned char* antisym = nullptr;  // [nS] uvw antisymmetric with zero diagonal
  // all diagonal pairs sit at u = 0: one point per sample valued by the trace of VV_l,
  // stored in the unused slot (0, 1) of each matrix
  int merge_auto = 0;
  long long ld_vv = 0;
  __host__ __device__ long long slots_per_sample() const {
    return mode == N3_SRC_PAIRS ? (long long)n_ant_max * (n_ant_max + 1) / 2
                                : (long long)n_ant_max * n_ant_max;
  }
};

struct N3PointsParams {
  N3Source src;
  // slots to process: slot_list[i] (= sample * slots_per_sample + t) for i < n_list, or all
  // slots of all samples when slot_list is null
  const unsigned long long* slot_list;
  long long n_list;
  double C[3], D[3], inv_gh[3];  // centres and 1/(gamma h1)
  int w;
  int off[3], bin_size[3], nbins[3];
  int phase_fp32;
  // values: VV_l(src) of every level (conjugated for upper-triangle entries) copied next to
  // the record, so that the spread streams them with the sorted points
  const void* vv;                // [nS][n_levels][vv_stride] complex (vv_double: double)
  long long vv_stride;
  int n_levels, vv_double;
  // Two launches over the same points: the first (fill = 0) counts the points per bin and
  // keeps each point's arrival rank in its bin; the second (fill = 1), after the scan of the
  // counts into bin_start, writes each point's record and values straight into its slot
  // bin_start + rank of the sorted arrays (no atomics), so the unsorted records are never
  // stored.
  int fill;
  unsigned* bin_count;           // [n_bins] zeroed (fill = 0)
  unsigned* rank;                // [n_list] written (fill = 0), read (fill = 1)
  const unsigned* bin_start;     // [n_bins] (fill = 1)
  N3Point* pts;                  // [points] sorted by bin (fill = 1)
  void* vals;                    // [points][n_levels] complex, same order (fill = 1)
};

// Domain partition of the points over a grid of boxes in u (the uvw-space partition of
// src/host/domain_partition.hpp:35-101): box index per axis = floor(inv (u - lo)) clamped to
// [0, dims - 1], boxes numbered x fastest.
struct N3PartParams {
  N3Source src;
  int n_samples;
  int dims[3];
  double lo[3], inv[3];
  int* group_of;                  // [nS * sps]: box of each slot, -1 for no point
  unsigned* group_count;          // [n_groups]
};

// Spread with output ownership: the workgroup of block b accumulates, in LDS, every point
// of bins b - nback .. b whose kernel support meets the block, then writes its cells once,
// deconvolved, with plain stores.
template <typename R>
struct N3SpreadParams {
  int n_levels, w;               // levels spread by this launch
  int lv0, n_levels_total;       // first level of the launch, levels stored per sample
  int lv_begin, lv_end;          // levels [lv_begin, lv_end) spread by the launcher; grid slot
                                 // 0 holds level lv_begin
  double beta;
  int bin_size[3], nbins[3], nback[3];
  int nf1[3], nf2[3];
  const unsigned *bin_start, *bin_count;
  const N3Point* pts;            // sorted by bin
  const cplx<R>* vals;           // [sorted point][n_levels_total] values
  const R *fx, *fy, *fz;         // 1/phi2^(l) for l = -nf1/2 .. nf1/2-1
  cplx<R>* grid;                 // [L][grid_stride]
  long long grid_stride;
  // optional per-block profile [n_blocks][4]: start, end (wall clock), points, items
  unsigned long long* prof;
  int xcd_chunk;                 // blocks per XCD run (0: blocks in launch order)
};

// Window of an axis of length n: indices (start + i) mod n for i < count.
struct N3Window {
  int start, count;
};

// Contiguous-axis (x) transforms of the pruned FFT: lines (level, z in zw, y in yw); inputs
// outside `in` are zero and are not read, outputs outside `out` are not written.
template <typename R>
struct N3LineParams {
  cplx<R>* grid;
  long long grid_stride;
  int n_levels;
  int n, ny, nz;
  N3Window yw, zw, in, out;
  N3Radix rad;
  const cplx<R>* tw;             // exp(+2 pi i t / n), t < n
};

// Strided-axis transforms of the pruned FFT for x positions (x0 + j) mod nx, j < nxw, and
// entries o < other.count of the remaining axis (storage index perm(start + o mod no)).
// Element i of sub-transform `sub` sits at axis index sub * sub_stride + i * estride.
template <typename R>
struct N3PassParams {
  cplx<R>* grid;
  long long grid_stride;
  int n_levels;
  int nx;
  N3Window xw;
  int no;
  N3Window other;
  int operm_n1, operm_n2;        // operm_n2 > 1: storage (k mod n1) n2 + k / n1
  long long outer_stride, axis_stride;
  int m, n_sub;
  long long sub_stride, estride;
  int nax;                       // axis length
  int in_check;                  // load only axis indices inside `in`
  N3Window in;
  int out_check, fa, fb;         // store output k only if frequency sub fa + k fb is in `out`
  N3Window out;
  int twiddle;                   // multiply output k of sub-transform sub by tw2[sub k mod ntot]
  long long ntot;
  N3Radix rad;
  const cplx<R>* tw;             // exp(+2 pi i t / m), t < m
  const cplx<R>* tw2;            // exp(+2 pi i t / ntot), t < ntot (when twiddle)
  int tw_shift_max;               // items hold at most 2^tw_shift_max x positions (4..6)
  int rr_m1;                      // > 0: two-stage register pass, m = rr_m1 (m / rr_m1)
  // set by the launcher: division multipliers and shifts (see N3Radix) for the item
  // decomposition (x tiles, n_sub, other.count) and for operm_n1
  unsigned xt_m, sub_m, oc_m, op_m;
  int xt_s, sub_s, oc_s, op_s;
};

template <typename R>
struct N3InterpParams {
  int n_pix, n_levels, w1, w2;
  double beta1, beta2;
  double gam[3], D[3], C[3], h1[3], h2[3];
  int nf2[3], n1[3], n2[3];
  const double* lmn;             // [3][n_pix]
  const double* ft;              // phi(xi) at xi = (j - 1) / ft_inv_delta, j < n_ft
  int n_ft;
  double ft_inv_delta;
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
// 1 per sample whose nAnt^2 x 3 uvw is antisymmetric with a zero diagonal, else 0
void launch_uvw_antisym(const N3Source& src, int n_samples, unsigned char* out, hipStream_t st);
// min and max of u over the points of each item: item k covers the slots
// list[begin[k] .. begin[k] + count[k]) (list == nullptr: slots of sample k); out [k][6]
void launch_slot_range(const N3Source& src, int n_items, const unsigned long long* list,
                       const unsigned* begin, const unsigned* count, double* out, hipStream_t st);
void launch_n3_partition(const N3PartParams& p, hipStream_t st);
void launch_n3_group_scatter(long long n_slots, const int* group_of, const unsigned* group_start,
                             unsigned* group_fill, unsigned long long* list, hipStream_t st);
template <typename R>
void launch_zero(R* p, long long n, hipStream_t st);
void launch_n3_points(const N3PointsParams& p, int n_samples, hipStream_t st);
void launch_scan(const unsigned* in, unsigned* out, unsigned* tmp, int n, hipStream_t st);
template <typename R, typename TV>
void launch_n3_spread(const N3SpreadParams<R>& p, hipStream_t st);
template <typename R>
void launch_fft_lines(const N3LineParams<R>& p, hipStream_t st);
template <typename R>
void launch_fft_pass(const N3PassParams<R>& p, hipStream_t st);
// m = m1 m2 with (m1, m2) a factor pair of the two-stage register pass (m1 >= m2)
bool fft_rr_factors(int m, int& m1, int& m2);
template <typename R, typename TI>
void launch_n3_interp(const N3InterpParams<R>& p, hipStream_t st);
template <typename TO>
void launch_vv_trace(int n_samples, int n_levels, const int* n_rows, int n_rows_const, void* vv,
                     long long vv_stride, long long ld_vv, hipStream_t st);
template <typename R, typename TO>
void launch_vv_levels(const N3VVParams<R>& p, int n_samples, int n_rows_max, hipStream_t st);

}  // namespace bippamd
