// Complex arithmetic on plain structs (layout-compatible with std::complex<R>), for device
// code and, with BIPP_CPLX_FN defined as `inline` beforehand, for host builds of the
// team-parallel routines.
#pragma once

#ifndef BIPP_CPLX_FN
#define BIPP_CPLX_FN __host__ __device__ __forceinline__
#endif

namespace bippamd {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};

template <typename R>
BIPP_CPLX_FN cplx<R> mk(R re, R im) {
  cplx<R> c;
  c.re = re;
  c.im = im;
  return c;
}
template <typename R>
BIPP_CPLX_FN cplx<R> operator+(cplx<R> a, cplx<R> b) {
  return mk(a.re + b.re, a.im + b.im);
}
template <typename R>
BIPP_CPLX_FN cplx<R> operator-(cplx<R> a, cplx<R> b) {
  return mk(a.re - b.re, a.im - b.im);
}
template <typename R>
BIPP_CPLX_FN cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
BIPP_CPLX_FN cplx<R> operator*(cplx<R> a, R s) {
  return mk(a.re * s, a.im * s);
}
template <typename R>
BIPP_CPLX_FN cplx<R> conj(cplx<R> a) {
  return mk(a.re, -a.im);
}
// a * conj(b)
template <typename R>
BIPP_CPLX_FN cplx<R> mulc(cplx<R> a, cplx<R> b) {
  return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
// conj(a) * b
template <typename R>
BIPP_CPLX_FN cplx<R> cmul_conj_a(cplx<R> a, cplx<R> b) {
  return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
template <typename R>
BIPP_CPLX_FN R abs2(cplx<R> a) {
  return a.re * a.re + a.im * a.im;
}
template <typename R, typename S>
BIPP_CPLX_FN cplx<R> cvt(cplx<S> a) {
  return mk<R>(static_cast<R>(a.re), static_cast<R>(a.im));
}

}  // namespace bippamd
