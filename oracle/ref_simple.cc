// ORACLE helper -- test infrastructure only, never part of the product.
//
// Second, independent reference statement of the same operator: the scalar
// (Nsimd = 1, no permute) loop kernel of dslash_simple.h:19-150, compiled from
// the reference sources where they lie (oracle/Makefile, target `ref`).  It
// needs a complex type with + - * += -=; the reference's bench supplies one
// inside bench_simple.cc, so an equivalent minimal type is written here.
#include <cstdint>

#define RESTRICT

template <class T> struct SimpleComplex {
  T re, im;
  SimpleComplex() {}
  template <class F> SimpleComplex(F r, F i) : re(r), im(i) {}
  SimpleComplex &operator+=(const SimpleComplex &r) { re += r.re; im += r.im; return *this; }
  SimpleComplex &operator-=(const SimpleComplex &r) { re -= r.re; im -= r.im; return *this; }
};
template <class T> SimpleComplex<T> operator*(const SimpleComplex<T> &l, const SimpleComplex<T> &r)
{
  return SimpleComplex<T>(l.re * r.re - l.im * r.im, l.re * r.im + l.im * r.re);
}
template <class T> SimpleComplex<T> operator+(const SimpleComplex<T> &l, const SimpleComplex<T> &r)
{
  return SimpleComplex<T>(l.re + r.re, l.im + r.im);
}
template <class T> SimpleComplex<T> operator-(const SimpleComplex<T> &l, const SimpleComplex<T> &r)
{
  return SimpleComplex<T>(l.re - r.re, l.im - r.im);
}

#include "dslash_simple.h"

extern "C" void gbxref_simple_d(double *U, double *out, double *in, uint64_t *nbr, uint64_t nsite, uint64_t Ls,
                                uint8_t *prm)
{
  typedef SimpleComplex<double> C;
  dslash_kernel<C>((C *)U, (C *)out, (C *)in, nbr, nsite, Ls, prm);
}
