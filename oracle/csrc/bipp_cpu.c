/*
 * CPU restatement of the reference's per-sample path -- TEST INFRASTRUCTURE ONLY.
 *
 * Used as the checker at sizes numpy finds slow and as the CPU baseline timed by
 * bench.py.  Never linked into, or called by, the product library.
 *
 *   bippc_eigh_gram_c : src/eigensolver.cpp:129-225 in float (the reference's default
 *                       precision): flag (|s|^2 > FLT_EPSILON, :148-159), Gram
 *                       sinc(2|p_i-p_j|/wl) (src/host/gram_matrix.cpp:23-56), LAPACK chegv
 *                       itype=1 'V' 'L' (src/host/lapack_api.hpp:112-134), un-beam with W=I.
 *   bippc_eigh_gram_batch: the same over independent samples, OpenMP over samples with
 *                       single-threaded LAPACK per thread.
 *   bippc_steering_image: I_l(x) += sum_e w_l[e] |sum_i V_ie exp(+i 2pi p_i.x/wl)|^2, the
 *                       image of src/host/virtual_vis.cpp:19-73 + src/host/kernels/nuft_sum.cpp
 *                       in closed form; OpenMP over pixels.
 *   bippc_nuft_sum_d  : src/host/kernels/nuft_sum.cpp:9-31 in double, OpenMP over outputs.
 *
 * LAPACK comes from the OpenBLAS that scipy bundles, opened with dlopen (path passed in).
 */
#include <complex.h>
#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef void (*chegv_t)(const int*, const char*, const char*, const int*, float complex*,
                        const int*, float complex*, const int*, float*, float complex*,
                        const int*, float*, int*, size_t, size_t);
typedef void (*set_threads_t)(int);

static chegv_t p_chegv = 0;
static set_threads_t p_set_threads = 0;

int bippc_init_lapack(const char* path) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  p_chegv = (chegv_t)dlsym(h, "scipy_chegv_");
  if (!p_chegv) p_chegv = (chegv_t)dlsym(h, "chegv_");
  p_set_threads = (set_threads_t)dlsym(h, "scipy_openblas_set_num_threads");
  if (!p_set_threads) p_set_threads = (set_threads_t)dlsym(h, "openblas_set_num_threads");
  return p_chegv ? 0 : -2;
}

void bippc_lapack_threads(int n) {
  if (p_set_threads) p_set_threads(n);
}

int bippc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static float sincf_pi(float x) {
  const float pi = 3.14159265358979323846f;
  return x != 0.f ? sinf(pi * x) / (pi * x) : 1.f;
}

/* s: n x n column-major complex64 (lower used); xyz: n x 3 column-major float.
 * Outputs d[n] (ascending, zero padded), v n x n (zero padded), *nr, *scale.
 * Returns LAPACK info (0 on success). */
int bippc_eigh_gram_c(float wl, int n, const float complex* s, const float* xyz, float* d,
                      float complex* v, int* nr_out, float* scale_out) {
  int* flag = (int*)calloc(n, sizeof(int));
  int* idx = (int*)malloc(sizeof(int) * n);
  size_t nvis = 0;
  for (int c = 0; c < n; ++c)
    for (int r = c; r < n; ++r) {
      const float complex x = s[r + (size_t)c * n];
      const float nrm = crealf(x) * crealf(x) + cimagf(x) * cimagf(x);
      if (nrm > FLT_EPSILON) {
        flag[r] = flag[c] = 1;
        nvis += 1 + (r != c);
      }
    }
  int nr = 0;
  for (int i = 0; i < n; ++i)
    if (flag[i]) idx[nr++] = i;
  memset(d, 0, sizeof(float) * n);
  memset(v, 0, sizeof(float complex) * (size_t)n * n);
  int info = 0;
  if (nr > 0) {
    float complex* a = (float complex*)malloc(sizeof(float complex) * (size_t)nr * nr);
    float complex* g = (float complex*)malloc(sizeof(float complex) * (size_t)nr * nr);
    for (int j = 0; j < nr; ++j)
      for (int i = j; i < nr; ++i) {
        a[i + (size_t)j * nr] = s[idx[i] + (size_t)idx[j] * n];
        const float dx = xyz[idx[i]] - xyz[idx[j]];
        const float dy = xyz[idx[i] + n] - xyz[idx[j] + n];
        const float dz = xyz[idx[i] + 2 * n] - xyz[idx[j] + 2 * n];
        g[i + (size_t)j * nr] = i == j ? 1.f : sincf_pi(2.f * sqrtf(dx * dx + dy * dy + dz * dz) / wl);
      }
    const int itype = 1;
    int lwork = -1;
    float complex wq;
    float* rwork = (float*)malloc(sizeof(float) * (3 * nr > 1 ? 3 * nr : 1));
    p_chegv(&itype, "V", "L", &nr, a, &nr, g, &nr, d, &wq, &lwork, rwork, &info, 1, 1);
    lwork = (int)crealf(wq);
    if (lwork < 2 * nr) lwork = 2 * nr;
    float complex* work = (float complex*)malloc(sizeof(float complex) * lwork);
    p_chegv(&itype, "V", "L", &nr, a, &nr, g, &nr, d, work, &lwork, rwork, &info, 1, 1);
    for (int j = 0; j < nr; ++j)
      for (int i = 0; i < nr; ++i) v[idx[i] + (size_t)j * n] = a[i + (size_t)j * nr];
    free(work);
    free(rwork);
    free(a);
    free(g);
  }
  *nr_out = nr;
  *scale_out = nvis ? 1.f / (float)nvis : 0.f;
  free(flag);
  free(idx);
  return info;
}

/* n_samples independent eigh_gram problems of size n (same layout as bippc_eigh_gram_c,
 * sample k at s + k n^2, xyz + 3 k n, d + k n, v + k n^2), one single-threaded LAPACK
 * call per OpenMP thread (the "independent problems x cores" form of the CPU baseline).
 * Returns the first nonzero LAPACK info, 0 if none. */
int bippc_eigh_gram_batch(int n_samples, const float* wl, int n, const float complex* s,
                          const float* xyz, float* d, float complex* v, int* nr, float* scale) {
  int bad = 0;
  if (p_set_threads) p_set_threads(1);
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < n_samples; ++k) {
    const int info = bippc_eigh_gram_c(wl[k], n, s + (size_t)k * n * n, xyz + (size_t)k * 3 * n,
                                       d + (size_t)k * n, v + (size_t)k * n * n, nr + k, scale + k);
    if (info) {
#pragma omp critical
      if (!bad) bad = info;
    }
  }
  return bad;
}

/* v: n_ant x n_eig column-major complex64; w: n_levels x n_eig (row-major) float;
 * pos: n_ant x 3 column-major double (positions / wavelength); lmn: P x 3 column-major
 * float; img: n_levels x P float, accumulated. */
void bippc_steering_image(int n_ant, int n_eig, const float complex* v, const float* w,
                          int n_levels, const double* pos, int n_pix, const float* lmn,
                          float* img) {
#pragma omp parallel
  {
    float* er = (float*)malloc(sizeof(float) * n_ant);
    float* ei = (float*)malloc(sizeof(float) * n_ant);
#pragma omp for schedule(static)
    for (int q = 0; q < n_pix; ++q) {
      const double l = lmn[q], m = lmn[q + n_pix], nn = lmn[q + 2 * (size_t)n_pix];
      for (int i = 0; i < n_ant; ++i) {
        const double t = pos[i] * l + pos[i + n_ant] * m + pos[i + 2 * n_ant] * nn;
        const float f = (float)(t - rint(t));
        er[i] = cosf(6.28318530717958647692f * f);
        ei[i] = sinf(6.28318530717958647692f * f);
      }
      for (int e = 0; e < n_eig; ++e) {
        const float complex* ve = v + (size_t)e * n_ant;
        float br = 0.f, bi = 0.f;
        for (int i = 0; i < n_ant; ++i) {
          const float vr = crealf(ve[i]), vi = cimagf(ve[i]);
          br += vr * er[i] - vi * ei[i];
          bi += vr * ei[i] + vi * er[i];
        }
        const float p = br * br + bi * bi;
        for (int lv = 0; lv < n_levels; ++lv) img[(size_t)lv * n_pix + q] += w[(size_t)lv * n_eig + e] * p;
      }
    }
    free(er);
    free(ei);
  }
}

void bippc_nuft_sum_d(double alpha, int n_in, const double complex* in, const double* u,
                      const double* v, const double* w, int n_out, const double* x,
                      const double* y, const double* z, double* out) {
#pragma omp parallel for schedule(static)
  for (int q = 0; q < n_out; ++q) {
    double sum = 0.0;
    for (int j = 0; j < n_in; ++j) {
      const double p = alpha * (x[q] * u[j] + y[q] * v[j] + z[q] * w[j]);
      sum += creal(in[j]) * cos(p) - cimag(in[j]) * sin(p);
    }
    out[q] += sum;
  }
}
