// CPU run of the divide-and-conquer tridiagonal eigensolver (bipp_amd/csrc/tridiag_dc.hpp)
// with one std::thread per workgroup thread and std::barrier for the workgroup barrier, so
// that the same index arithmetic and synchronisation as on the GPU run under the host's
// sanitizers.
//   dc_harness IN OUT: IN = int32 {nmax, n, single} + float64 d[n] + float64 e[n - 1];
//                      OUT = float64 d[n] + float64 Z[n * n] (column-major, ld n).
#include <math.h>

#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

#define BIPP_HD
#include "../../bipp_amd/csrc/tridiag_dc.hpp"

namespace {

struct CpuTeam {
  int tid;
  std::barrier<>* bar;
  int gsize;                                             // lanes per lane_sum group
  std::vector<std::unique_ptr<std::barrier<>>>* groups;  // one barrier per group
  double* slot;                                          // one value per thread
  void sync() { bar->arrive_and_wait(); }
  void mark(int) {}
  template <typename R, int TR, int TK, int NT>
  using DcTile = bippamd::DcValuTile<R, TR, TK, NT>;
  // the GPU's butterfly over the group: v += v of lane tid ^ m for m = 1, 2, ...
  template <int P, typename R>
  R lane_sum(R v) {
    if (P == 1) return v;
    if (P != gsize) std::abort();
    std::barrier<>& gb = *(*groups)[tid / P];
    for (int m = 1; m < P; m <<= 1) {
      slot[tid] = static_cast<double>(v);
      gb.arrive_and_wait();
      const R o = static_cast<R>(slot[tid ^ m]);
      gb.arrive_and_wait();
      v += o;
    }
    return v;
  }
};

template <typename R, int NMAX>
void solve(int n, std::vector<double>& d64, std::vector<double>& e64, std::vector<double>& z64) {
  constexpr int NT = NMAX >= 256 ? 512 : 256;
  std::vector<R> d(d64.begin(), d64.end()), e(n, R(0));
  for (int i = 0; i + 1 < n; ++i) e[i] = static_cast<R>(e64[i]);
  // the work arrays hold garbage where the solver must not read (only what it wrote)
  std::vector<R> Z((size_t)n * n, R(NAN)), W((size_t)n * n, R(NAN)), M((size_t)n * n, R(NAN));
  auto sh = std::make_unique<bippamd::DcShared<R, NMAX>>();
  std::vector<R> tile(bippamd::DcCfg<NMAX>::TILE_ELEMS);
  std::barrier<> bar(NT);
  constexpr int G = NT >= 4 * NMAX ? 4 : NT >= 2 * NMAX ? 2 : 1;  // lanes per secular root
  std::vector<std::unique_ptr<std::barrier<>>> groups;
  for (int g = 0; g < NT / G; ++g) groups.push_back(std::make_unique<std::barrier<>>(G));
  std::vector<double> slot(NT);
  std::vector<std::thread> th;
  for (int t = 0; t < NT; ++t)
    th.emplace_back([&, t] {
      CpuTeam tm{t, &bar, G, &groups, slot.data()};
      bippamd::dc_tridiag<R, NMAX, NT>(tm, n, d.data(), e.data(), 1, Z.data(), n, W.data(),
                                       M.data(), *sh, tile.data());
    });
  for (auto& x : th) x.join();
  d64.assign(d.begin(), d.end());
  z64.assign(Z.begin(), Z.end());
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 3;
  int hdr[3];
  if (fread(hdr, sizeof(int), 3, f) != 3) return 4;
  const int nmax = hdr[0], n = hdr[1], single = hdr[2];
  if (n < 1 || n > nmax) return 5;
  std::vector<double> d(n), e(n > 1 ? n - 1 : 0), z;
  if (fread(d.data(), sizeof(double), n, f) != (size_t)n) return 6;
  if (n > 1 && fread(e.data(), sizeof(double), n - 1, f) != (size_t)(n - 1)) return 7;
  fclose(f);
#define RUN(NM)                                         \
  if (nmax == NM) {                                     \
    if (single)                                         \
      solve<float, NM>(n, d, e, z);                     \
    else                                                \
      solve<double, NM>(n, d, e, z);                    \
  }
  RUN(64) else RUN(128) else RUN(256) else RUN(512) else return 8;
  FILE* g = fopen(argv[2], "wb");
  if (!g) return 9;
  fwrite(d.data(), sizeof(double), n, g);
  fwrite(z.data(), sizeof(double), (size_t)n * n, g);
  fclose(g);
  return 0;
}
