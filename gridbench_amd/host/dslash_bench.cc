// dslash_bench -- the reference benchmark's flow (bench.cc:59-202) on libgbx.
//
// Builds seeded inputs in the reference host layout for Nsimd lanes, the
// nbr/prm tables in the reference's form, calls
//   dslash_kernel<Simd>(nrep, U, out, in, nbr, nsite, Ls, prm)
// through gbx_dslash.hpp exactly as bench.cc:148-164 calls the reference,
// prints the reference's line (bench.cc:174, but with nrep-1 applications in
// the rate, the ones the clock covers) and checks the result against the
// library's device-resident path (gbx_dhop) on the same inputs.
//
// usage: dslash_bench [L=8] [Ls=8] [nrep=1000] [precision d|f] [nsimd=1]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gbx.h"
#include "gbx_dslash.hpp"

template <class T, int N>
static int run(int L, int Ls, int nrep)
{
  typedef gbx::vcomplex<T, N> Simd;
  const int dims[4] = {L, L, L, L};
  const uint64_t V = (uint64_t)L * L * L * L, nsite = V / N;
  std::vector<double> links(V * 72), dbl(V * 144), psi(V * Ls * 24);
  if (gbx_gen_su3_links(dims, 101, links.data()) || gbx_double_links(dims, links.data(), dbl.data()) ||
      gbx_gen_spinor(dims, Ls, 1, 102, psi.data())) {
    std::fprintf(stderr, "generator: %s\n", gbx_last_error());
    return 1;
  }
  std::vector<uint64_t> nbr(nsite * Ls * 8);
  std::vector<uint8_t> prm(nsite * Ls * 8);
  if (gbx_table_generate(dims, N, Ls, nbr.data(), prm.data())) {
    std::fprintf(stderr, "tables: %s\n", gbx_last_error());
    return 1;
  }
  // Reference layout: pack through the device path (upload lexicographic,
  // download in the Nsimd virtual-node layout).
  gbx_geom g;
  gbx_gauge U64;
  gbx_spinor x, y;
  if (gbx_geom_create(dims, nullptr, &g) || gbx_gauge_create(g, GBX_F64, 0, &U64) || gbx_spinor_create(g, Ls, 1, GBX_F64, &x) ||
      gbx_spinor_create(g, Ls, 1, GBX_F64, &y) || gbx_gauge_upload(U64, dbl.data()) || gbx_spinor_upload(x, psi.data())) {
    std::fprintf(stderr, "device setup: %s\n", gbx_last_error());
    return 1;
  }
  std::vector<double> inref(V * Ls * 24), ref(V * Ls * 24);
  gbx_spinor_download_ref_f64(x, inref.data(), N);
  gbx_dhop(U64, y, x, 0);
  gbx_spinor_download_ref_f64(y, ref.data(), N);
  // Gauge field in the reference layout [nsite][8][3][3][N][2]
  std::vector<T> Uh(V * 144), in(V * Ls * 24), out(V * Ls * 24);
  {
    int simd[4] = {1, 1, 1, 1}, d = 3;
    for (int n = N; n > 1; n >>= 1, d = (d + 3) % 4) simd[d] *= 2;
    int rd[4];
    for (int e = 0; e < 4; e++) rd[e] = L / simd[e];
    for (uint64_t site = 0; site < V; site++) {
      uint64_t r = site;
      int xg[4];
      for (int e = 0; e < 4; e++) {
        xg[e] = (int)(r % L);
        r /= L;
      }
      uint64_t o = 0, lane = 0;
      for (int e = 3; e >= 0; e--) {
        o = o * rd[e] + xg[e] % rd[e];
        lane = lane * simd[e] + xg[e] / rd[e];
      }
      for (int k = 0; k < 72; k++)
        for (int ri = 0; ri < 2; ri++) Uh[((o * 72 + k) * N + lane) * 2 + ri] = (T)dbl[(site * 72 + k) * 2 + ri];
    }
  }
  for (uint64_t i = 0; i < in.size(); i++) in[i] = (T)inref[i];

  const double flops = 1320.0 * (double)V * Ls;
  const double usec = dslash_kernel<Simd>(nrep, (Simd *)Uh.data(), (Simd *)out.data(), (Simd *)in.data(), nbr.data(), nsite,
                                          Ls, prm.data());
  if (usec < 0) return 1;
  const int timed = nrep > 1 ? nrep - 1 : 1;
  std::printf("\t%g Gflop/s in %s precision; kernel call %g microseconds (Nsimd %d, %d^4 x Ls%d, path %d)\n",
              timed * flops / usec / 1000.0, sizeof(T) == 8 ? "double" : "single", usec / timed, N, L, Ls,
              gbx_dslash_last_path());
  double err = 0, nref = 0, nres = 0;
  for (uint64_t i = 0; i < out.size(); i++) {
    err += ((double)out[i] - ref[i]) * ((double)out[i] - ref[i]);
    nres += (double)out[i] * out[i];
    nref += ref[i] * ref[i];
  }
  std::printf("normdiff %g ref %g result %g (relative %g)\n", err, nref, nres, std::sqrt(err / nref));
  gbx_spinor_destroy(x);
  gbx_spinor_destroy(y);
  gbx_gauge_destroy(U64);
  gbx_geom_destroy(g);
  const double tol = sizeof(T) == 8 ? 1e-14 : 1e-6;
  return std::sqrt(err / nref) <= tol ? 0 : 2;
}

int main(int argc, char **argv)
{
  const int L = argc > 1 ? std::atoi(argv[1]) : 8;
  const int Ls = argc > 2 ? std::atoi(argv[2]) : 8;
  const int nrep = argc > 3 ? std::atoi(argv[3]) : 1000;
  const char prec = argc > 4 ? argv[4][0] : 'd';
  const int nsimd = argc > 5 ? std::atoi(argv[5]) : 1;
  if (gbx_device_count() < 1) {
    std::fprintf(stderr, "dslash_bench: no GPU\n");
    return 3;
  }
#define GBX_CASE(T, N) \
  if (nsimd == N) return run<T, N>(L, Ls, nrep);
  if (prec == 'd') {
    GBX_CASE(double, 1)
    GBX_CASE(double, 2)
    GBX_CASE(double, 4)
    GBX_CASE(double, 8)
    GBX_CASE(double, 16)
  } else {
    GBX_CASE(float, 1)
    GBX_CASE(float, 2)
    GBX_CASE(float, 4)
    GBX_CASE(float, 8)
    GBX_CASE(float, 16)
  }
  std::fprintf(stderr, "dslash_bench: nsimd %d not in {1,2,4,8,16}\n", nsimd);
  return 1;
}
