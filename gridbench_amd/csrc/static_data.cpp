// Reader and writer for the reference's static_data blobs.
//
// The reference benchmark links five arrays that Grid's TableGenerate writes
// out as C source (TableGenerate.cc:93-138): the doubled gauge field
// U_static, the source Phi_static, Grid's Dhop of it Psi_cpp_static (all in
// the reference host layout for its SIMD width, bench.cc:105-138), and the
// stencil tables nbr_static / prm_static, plus a header with nsite and Ls.
// The blobs themselves are not in the repository (.MISSING_LARGE_BLOBS).
// The writer reproduces that text format byte for byte (same printf formats,
// same array names, the trailing 0 element); the reader parses any such file
// (or a split pair such as static_data_gauge.cc / static_data_fermion.cc),
// so a real Grid blob can be checked against this library at Grid's 9-digit
// print precision.
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "gbx_internal.h"

namespace {

const char *kNames[5] = {"U_static", "Phi_static", "Psi_cpp_static", "nbr_static", "prm_static"};

struct File {
  FILE *f = nullptr;
  ~File()
  {
    if (f) fclose(f);
  }
};

int read_all(const char *path, std::string &text)
{
  File fh;
  fh.f = fopen(path, "rb");
  GBX_CHECK(fh.f, GBX_EINVAL, "static_data: cannot open %s: %s", path, strerror(errno));
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, fh.f)) > 0) text.append(buf, n);
  GBX_CHECK(!ferror(fh.f), GBX_EINVAL, "static_data: read error on %s", path);
  return GBX_OK;
}

// Position of the first value of array `name` (just after its '{'), or npos.
size_t array_start(const std::string &t, const char *name)
{
  const std::string key = std::string(name) + "[]";
  size_t p = 0;
  while ((p = t.find(key, p)) != std::string::npos) {
    // a definition "<type> name[] = {", not the extern declaration
    const size_t eq = t.find_first_not_of(" \t", p + key.size());
    if (eq != std::string::npos && t[eq] == '=') {
      const size_t br = t.find('{', eq);
      return br == std::string::npos ? std::string::npos : br + 1;
    }
    p += key.size();
  }
  return std::string::npos;
}

// Parse the comma-separated values of an array up to its '}'; the last
// element is the writer's trailing 0, which is not data.
template <typename T>
int parse_array(const std::string &t, size_t p, bool integer, std::vector<T> &out, const char *name)
{
  out.clear();
  const char *s = t.c_str() + p;
  const char *end = t.c_str() + t.size();
  for (;;) {
    while (s < end && (*s == ' ' || *s == '\n' || *s == '\t' || *s == '\r' || *s == ',')) s++;
    GBX_CHECK(s < end, GBX_EINVAL, "static_data: %s is not terminated", name);
    if (*s == '}') break;
    char *e = nullptr;
    if (integer) {
      errno = 0;
      const unsigned long long v = strtoull(s, &e, 0);
      GBX_CHECK(e != s && errno == 0 && *s != '-' && v <= (unsigned long long)std::numeric_limits<T>::max(), GBX_EINVAL,
                "static_data: bad integer in %s at element %zu", name, out.size());
      out.push_back((T)v);
    } else {
      const double v = strtod(s, &e);
      GBX_CHECK(e != s, GBX_EINVAL, "static_data: bad number in %s at element %zu", name, out.size());
      out.push_back((T)v);
    }
    s = e;
  }
  GBX_CHECK(!out.empty(), GBX_EINVAL, "static_data: %s is empty", name);
  out.pop_back();  // the trailing 0
  return GBX_OK;
}

struct Blob {
  std::vector<double> U, Phi, Psi;
  std::vector<uint64_t> nbr;
  std::vector<uint8_t> prm;
};

int read_blob(const char *const *paths, int npaths, Blob &b)
{
  GBX_CHECK(paths && npaths >= 1, GBX_ENULL, "static_data: no files");
  bool found[5] = {false, false, false, false, false};
  for (int i = 0; i < npaths; i++) {
    GBX_CHECK(paths[i], GBX_ENULL, "static_data: NULL path");
    std::string t;
    GBX_TRY(read_all(paths[i], t));
    for (int a = 0; a < 5; a++) {
      const size_t p = array_start(t, kNames[a]);
      if (p == std::string::npos) continue;
      GBX_CHECK(!found[a], GBX_EINVAL, "static_data: %s defined twice", kNames[a]);
      found[a] = true;
      if (a == 0) GBX_TRY(parse_array(t, p, false, b.U, kNames[a]));
      if (a == 1) GBX_TRY(parse_array(t, p, false, b.Phi, kNames[a]));
      if (a == 2) GBX_TRY(parse_array(t, p, false, b.Psi, kNames[a]));
      if (a == 3) GBX_TRY(parse_array(t, p, true, b.nbr, kNames[a]));
      if (a == 4) {
        std::vector<uint64_t> tmp;
        GBX_TRY(parse_array(t, p, true, tmp, kNames[a]));
        b.prm.resize(tmp.size());
        for (size_t k = 0; k < tmp.size(); k++) {
          GBX_CHECK(tmp[k] <= 1, GBX_ETABLE, "static_data: prm_static[%zu] = %llu", k, (unsigned long long)tmp[k]);
          b.prm[k] = (uint8_t)tmp[k];
        }
      }
    }
  }
  for (int a = 0; a < 5; a++) GBX_CHECK(found[a], GBX_EINVAL, "static_data: %s not found", kNames[a]);
  return GBX_OK;
}

}  // namespace

extern "C" int gbx_static_data_write(const char *cc_path, const char *h_path, uint64_t nsite, int Ls, int nsimd,
                                     const double *U, const double *Phi, const double *Psi_cpp, const uint64_t *nbr,
                                     const uint8_t *prm)
{
  GBX_CHECK(cc_path && h_path && U && Phi && Psi_cpp && nbr && prm, GBX_ENULL, "gbx_static_data_write: NULL argument");
  GBX_CHECK(nsite >= 1 && Ls >= 1 && nsimd >= 1, GBX_EINVAL, "gbx_static_data_write: nsite=%llu Ls=%d nsimd=%d",
            (unsigned long long)nsite, Ls, nsimd);
  const uint64_t umax = nsite * 18 * 8 * (uint64_t)nsimd, fmax = nsite * 24 * (uint64_t)Ls * nsimd,
                 nbrmax = nsite * (uint64_t)Ls * 8;
  {
    File fh;
    fh.f = fopen(cc_path, "w");
    GBX_CHECK(fh.f, GBX_EINVAL, "gbx_static_data_write: cannot create %s: %s", cc_path, strerror(errno));
    FILE *fp = fh.f;
    fprintf(fp, "#include <stdint.h>\n");
    const double *arr[3] = {U, Phi, Psi_cpp};
    const uint64_t len[3] = {umax, fmax, fmax};
    for (int a = 0; a < 3; a++) {
      fprintf(fp, "double %s[] = { \n ", kNames[a]);
      for (uint64_t n = 0; n < len[a]; n++) fprintf(fp, "    %16.8le, \n", arr[a][n]);
      fprintf(fp, "    0}; \n ");
    }
    fprintf(fp, "uint64_t nbr_static[] = { \n ");
    for (uint64_t n = 0; n < nbrmax; n++) fprintf(fp, "    0x%llx, \n", (unsigned long long)nbr[n]);
    fprintf(fp, "    0}; \n ");
    fprintf(fp, "uint8_t prm_static[] = { \n ");
    for (uint64_t n = 0; n < nbrmax; n++) fprintf(fp, "    0x%x, \n", (unsigned)prm[n]);
    fprintf(fp, "    0}; \n ");
    GBX_CHECK(!ferror(fp), GBX_EINVAL, "gbx_static_data_write: write error on %s", cc_path);
  }
  File fh;
  fh.f = fopen(h_path, "w");
  GBX_CHECK(fh.f, GBX_EINVAL, "gbx_static_data_write: cannot create %s: %s", h_path, strerror(errno));
  FILE *fp = fh.f;
  fprintf(fp, "#include <stdint.h>\n");
  fprintf(fp, "const uint64_t nsite = %llu ; \n", (unsigned long long)nsite);
  fprintf(fp, "const int Ls = %d ; \n", Ls);
  fprintf(fp, "extern double   U_static[] ; \n");
  fprintf(fp, "extern double   Phi_static[] ; \n");
  fprintf(fp, "extern double   Psi_cpp_static[] ; \n");
  fprintf(fp, "extern uint64_t nbr_static[] ; \n");
  fprintf(fp, "extern uint8_t  prm_static[] ; \n");
  GBX_CHECK(!ferror(fp), GBX_EINVAL, "gbx_static_data_write: write error on %s", h_path);
  return GBX_OK;
}

extern "C" int gbx_static_data_read_header(const char *h_path, uint64_t *nsite, int *Ls)
{
  GBX_CHECK(h_path && nsite && Ls, GBX_ENULL, "gbx_static_data_read_header: NULL argument");
  std::string t;
  GBX_TRY(read_all(h_path, t));
  const size_t a = t.find("nsite"), b = t.find("Ls");
  GBX_CHECK(a != std::string::npos && b != std::string::npos, GBX_EINVAL, "static_data: %s has no nsite / Ls", h_path);
  const size_t ea = t.find('=', a), eb = t.find('=', b);
  GBX_CHECK(ea != std::string::npos && eb != std::string::npos, GBX_EINVAL, "static_data: no value for nsite / Ls in %s",
            h_path);
  unsigned long long ns = 0;
  int ls = 0;
  GBX_CHECK(sscanf(t.c_str() + ea, "= %llu", &ns) == 1 && sscanf(t.c_str() + eb, "= %d", &ls) == 1, GBX_EINVAL,
            "static_data: cannot parse nsite / Ls in %s", h_path);
  GBX_CHECK(ns >= 1 && ls >= 1, GBX_EINVAL, "static_data: nsite=%llu Ls=%d", ns, ls);
  *nsite = ns;
  *Ls = ls;
  return GBX_OK;
}

extern "C" int gbx_static_data_sizes(const char *const *paths, int npaths, uint64_t sizes[5])
{
  GBX_CHECK(sizes, GBX_ENULL, "gbx_static_data_sizes: NULL sizes");
  Blob b;
  GBX_TRY(read_blob(paths, npaths, b));
  sizes[0] = b.U.size();
  sizes[1] = b.Phi.size();
  sizes[2] = b.Psi.size();
  sizes[3] = b.nbr.size();
  sizes[4] = b.prm.size();
  return GBX_OK;
}

extern "C" int gbx_static_data_read(const char *const *paths, int npaths, double *U, double *Phi, double *Psi_cpp,
                                    uint64_t *nbr, uint8_t *prm, const uint64_t sizes[5])
{
  GBX_CHECK(U && Phi && Psi_cpp && nbr && prm && sizes, GBX_ENULL, "gbx_static_data_read: NULL argument");
  Blob b;
  GBX_TRY(read_blob(paths, npaths, b));
  const uint64_t have[5] = {b.U.size(), b.Phi.size(), b.Psi.size(), b.nbr.size(), b.prm.size()};
  for (int a = 0; a < 5; a++)
    GBX_CHECK(have[a] == sizes[a], GBX_EMISMATCH, "gbx_static_data_read: %s has %llu elements, caller expects %llu",
              kNames[a], (unsigned long long)have[a], (unsigned long long)sizes[a]);
  memcpy(U, b.U.data(), have[0] * sizeof(double));
  memcpy(Phi, b.Phi.data(), have[1] * sizeof(double));
  memcpy(Psi_cpp, b.Psi.data(), have[2] * sizeof(double));
  memcpy(nbr, b.nbr.data(), have[3] * sizeof(uint64_t));
  memcpy(prm, b.prm.data(), have[4]);
  return GBX_OK;
}
