// static_data_check -- drop a static_data blob onto libgbx.
//
// Reads the arrays Grid's TableGenerate writes (one static_data.cc, or a
// split pair, plus static_data.h for nsite and Ls), infers the SIMD width
// from the gauge array's length, runs the hopping term on the GPU through
// the reference entry point dslash_kernel<Simd>(...) on U_static,
// Phi_static and the blob's own tables, and prints the reference's check
// (bench.cc:180-199): the squared norm of Psi_cpp - Psi, the two norms, and
// the first 64 values.  Exit status 0 when |Psi_cpp - Psi| / |Psi_cpp| is
// within tol (default 1e-7: the blob holds 9 significant digits).
//
// usage: static_data_check static_data.h file.cc [file2.cc] [--f32] [--tol X] [--nrep N]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gbx.h"
#include "gbx_dslash.hpp"

template <class T, int N>
static int run(int nrep, uint64_t nsite, uint64_t Ls, const std::vector<double> &U, const std::vector<double> &Phi,
               std::vector<double> &Psi, std::vector<uint64_t> &nbr, std::vector<uint8_t> &prm, bool dbl)
{
  typedef gbx::vcomplex<T, N> Simd;
  std::vector<T> u(U.begin(), U.end()), in(Phi.begin(), Phi.end()), out(Phi.size());
  const double us = dslash_kernel<Simd>(nrep, (Simd *)u.data(), (Simd *)out.data(), (Simd *)in.data(), nbr.data(), nsite,
                                        Ls, prm.data());
  if (us < 0) return 1;
  const double flops = 1320.0 * (double)nsite * (double)Ls * N;
  std::printf("\t%g Gflop/s in %s precision; kernel call %g microseconds\n", (nrep - 1) * flops / us / 1000.0,
              dbl ? "double" : "single", us / (nrep - 1));
  for (size_t i = 0; i < out.size(); i++) Psi[i] = (double)out[i];
  return 0;
}

int main(int argc, char **argv)
{
  const char *hdr = nullptr;
  std::vector<const char *> files;
  bool f32 = false;
  double tol = 1e-7;
  int nrep = 10;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--f32")) f32 = true;
    else if (!std::strcmp(argv[i], "--tol") && i + 1 < argc) tol = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "--nrep") && i + 1 < argc) nrep = std::atoi(argv[++i]);
    else if (!hdr) hdr = argv[i];
    else files.push_back(argv[i]);
  }
  if (!hdr || files.empty() || nrep < 2) {
    std::fprintf(stderr, "usage: static_data_check static_data.h file.cc [file2.cc] [--f32] [--tol X] [--nrep N>=2]\n");
    return 2;
  }
  uint64_t nsite;
  int Ls;
  uint64_t sizes[5];
  if (gbx_static_data_read_header(hdr, &nsite, &Ls) || gbx_static_data_sizes(files.data(), (int)files.size(), sizes)) {
    std::fprintf(stderr, "static_data: %s\n", gbx_last_error());
    return 2;
  }
  const uint64_t nsimd = sizes[0] / (144 * nsite);
  if (sizes[0] != nsimd * 144 * nsite || sizes[1] != nsite * 24 * Ls * nsimd || sizes[2] != sizes[1] ||
      sizes[3] != nsite * Ls * 8 || sizes[4] != sizes[3]) {
    std::fprintf(stderr, "static_data: array sizes do not fit nsite=%llu Ls=%d\n", (unsigned long long)nsite, Ls);
    return 2;
  }
  std::vector<double> U(sizes[0]), Phi(sizes[1]), Psi_cpp(sizes[2]), Psi(sizes[2]);
  std::vector<uint64_t> nbr(sizes[3]);
  std::vector<uint8_t> prm(sizes[4]);
  if (gbx_static_data_read(files.data(), (int)files.size(), U.data(), Phi.data(), Psi_cpp.data(), nbr.data(), prm.data(),
                           sizes)) {
    std::fprintf(stderr, "static_data: %s\n", gbx_last_error());
    return 2;
  }
  std::printf("Nsimd %llu nsite %llu Ls %d\n", (unsigned long long)nsimd, (unsigned long long)nsite, Ls);
  int rc = 2;
#define GBX_RUN(N)                                                                                   \
  case N:                                                                                            \
    rc = f32 ? run<float, N>(nrep, nsite, Ls, U, Phi, Psi, nbr, prm, false)                          \
             : run<double, N>(nrep, nsite, Ls, U, Phi, Psi, nbr, prm, true);                          \
    break;
  switch (nsimd) {
    GBX_RUN(1)
    GBX_RUN(2)
    GBX_RUN(4)
    GBX_RUN(8)
    GBX_RUN(16)
  default: std::fprintf(stderr, "static_data: unsupported SIMD width %llu\n", (unsigned long long)nsimd); return 2;
  }
#undef GBX_RUN
  if (rc) {
    std::fprintf(stderr, "dslash_kernel: %s\n", gbx_last_error());
    return rc;
  }
  double err = 0, nref = 0, nres = 0;
  for (size_t i = 0; i < Psi.size(); i++) {
    err += (Psi_cpp[i] - Psi[i]) * (Psi_cpp[i] - Psi[i]);
    nres += Psi[i] * Psi[i];
    nref += Psi_cpp[i] * Psi_cpp[i];
  }
  std::printf("normdiff %g ref %g result %g\n", err, nref, nres);
  for (int i = 0; i < 64 && i < (int)Psi.size(); i++) std::printf("%d ref %g result %g\n", i, Psi_cpp[i], Psi[i]);
  const double rel = std::sqrt(err / nref);
  std::printf("relative difference %.3e (tolerance %.1e)\n", rel, tol);
  return rel <= tol ? 0 : 1;
}
