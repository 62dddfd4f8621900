// CPU run of the blocked Cholesky + inverse (bipp_amd/csrc/chol_inv.hpp) with one
// std::thread per workgroup thread and std::barrier for the workgroup barrier, under the
// host's sanitizers.
//   ci_harness IN OUT: IN = int32 {nt, n, single} + complex128 G[n * n] (column-major);
//                      OUT = int32 info + complex128 W[n * n] (lower triangle = L^-1).
#include <math.h>

#include <barrier>
#include <cmath>
#include <complex>
#include <cstdio>
#include <memory>
#include <thread>
#include <vector>

#define BIPP_HD
#define BIPP_CPLX_FN inline
#include "../../bipp_amd/csrc/chol_inv.hpp"

namespace {

struct CpuTeam {
  int tid;
  std::barrier<>* bar;
  void sync() { bar->arrive_and_wait(); }
  void mark(int) {}
  template <typename R, int NT>
  using CiTile = bippamd::CiValuTile<R, NT>;
};

template <typename R, int NT>
int solve(int n, std::vector<std::complex<double>>& g) {
  using C = bippamd::cplx<R>;
  std::vector<C> L((size_t)n * n), T((size_t)n * n);
  for (size_t i = 0; i < L.size(); ++i) {
    L[i].re = static_cast<R>(g[i].real());
    L[i].im = static_cast<R>(g[i].imag());
    T[i].re = T[i].im = R(NAN);  // scratch starts as garbage
  }
  // the strict upper triangle must be neither read nor written: garbage in it
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < j; ++i) L[i + (size_t)j * n].re = L[i + (size_t)j * n].im = R(NAN);
  auto sh = std::make_unique<bippamd::CiShared<R>>();
  std::barrier<> bar(NT);
  std::vector<int> ret(NT);
  std::vector<std::thread> th;
  for (int t = 0; t < NT; ++t)
    th.emplace_back([&, t] {
      CpuTeam tm{t, &bar};
      ret[t] = bippamd::chol_inverse<R, NT>(tm, n, L.data(), n, T.data(), n, *sh);
    });
  for (auto& x : th) x.join();
  for (int t = 1; t < NT; ++t)
    if (ret[t] != ret[0]) return -1000;  // every thread must see the same outcome
  for (size_t i = 0; i < L.size(); ++i) g[i] = std::complex<double>(L[i].re, L[i].im);
  return ret[0];
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 3;
  int hdr[3];
  if (fread(hdr, sizeof(int), 3, f) != 3) return 4;
  const int nt = hdr[0], n = hdr[1], single = hdr[2];
  if (n < 1) return 5;
  std::vector<std::complex<double>> g((size_t)n * n);
  if (fread(g.data(), sizeof(std::complex<double>), g.size(), f) != g.size()) return 6;
  fclose(f);
  int info;
  if (nt == 256)
    info = single ? solve<float, 256>(n, g) : solve<double, 256>(n, g);
  else if (nt == 512)
    info = single ? solve<float, 512>(n, g) : solve<double, 512>(n, g);
  else
    return 7;
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 8;
  fwrite(&info, sizeof(int), 1, o);
  fwrite(g.data(), sizeof(std::complex<double>), g.size(), o);
  fclose(o);
  return 0;
}
