// MatrixMarket text I/O for the factor exchange format: the files the
// reference's sv%dump writes (mld_d_base_ainv_solver_dmp.f90:77-106, through
// PSBLAS's CSR print and psb_geprt) and the matrices its sample drivers read
// (tests/fileread/df_sample.f90:156-175).  Host code only.
#include <algorithm>
#include <cctype>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ainv_internal.h"

namespace ainv {
namespace {

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

File open_or_throw(const std::string &path, const char *mode) {
  FILE *f = std::fopen(path.c_str(), mode);
  if (!f) throw Error(AINV_ERR_ARG, "cannot open " + path);
  return File(f);
}

std::string lower(std::string s) {
  for (char &c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Reads the banner and skips comments; returns the size line.
std::string read_header(FILE *f, const std::string &path,
                        std::vector<std::string> &banner) {
  char line[4096];
  if (!std::fgets(line, sizeof line, f))
    throw Error(AINV_ERR_ARG, path + ": empty file");
  char w[5][64] = {};
  if (std::sscanf(line, "%63s %63s %63s %63s %63s", w[0], w[1], w[2], w[3],
                  w[4]) != 5 ||
      lower(w[0]) != "%%matrixmarket" || lower(w[1]) != "matrix")
    throw Error(AINV_ERR_ARG, path + ": not a MatrixMarket matrix file");
  banner.assign({lower(w[2]), lower(w[3]), lower(w[4])});
  while (std::fgets(line, sizeof line, f)) {
    const char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '%' || *p == '\n' || *p == '\r' || *p == 0) continue;
    return std::string(line);
  }
  throw Error(AINV_ERR_ARG, path + ": missing size line");
}

}  // namespace

void mm_write_csr(const std::string &path, const HostCsr &a, const char *head) {
  File f = open_or_throw(path, "w");
  std::fprintf(f.get(), "%%%%MatrixMarket matrix coordinate real general\n");
  if (head && *head) std::fprintf(f.get(), "%% %s\n", head);
  std::fprintf(f.get(), "%d %d %" PRId64 "\n", a.n, a.n, a.nnz());
  for (int32_t i = 0; i < a.n; ++i)
    for (int64_t p = a.ptr[i]; p < a.ptr[i + 1]; ++p)
      std::fprintf(f.get(), "%d %d %.17g\n", i + 1, a.idx[p] + 1, a.val[p]);
  if (std::ferror(f.get())) throw Error(AINV_ERR_INTERNAL, "write failed: " + path);
}

void mm_write_vector(const std::string &path, const std::vector<double> &v,
                     const char *head) {
  File f = open_or_throw(path, "w");
  std::fprintf(f.get(), "%%%%MatrixMarket matrix array real general\n");
  if (head && *head) std::fprintf(f.get(), "%% %s\n", head);
  std::fprintf(f.get(), "%zu 1\n", v.size());
  for (double x : v) std::fprintf(f.get(), "%.17g\n", x);
  if (std::ferror(f.get())) throw Error(AINV_ERR_INTERNAL, "write failed: " + path);
}

// Coordinate files (real / integer / pattern; general / symmetric /
// skew-symmetric) to a square CSR with sorted columns and duplicate
// entries summed.
HostCsr mm_read_csr(const std::string &path) {
  File f = open_or_throw(path, "r");
  std::vector<std::string> b;
  const std::string size = read_header(f.get(), path, b);
  if (b[0] != "coordinate")
    throw Error(AINV_ERR_ARG, path + ": sparse (coordinate) format expected");
  const bool pattern = b[1] == "pattern";
  if (!pattern && b[1] != "real" && b[1] != "integer" && b[1] != "double")
    throw Error(AINV_ERR_ARG, path + ": field " + b[1] + " not supported");
  const bool sym = b[2] == "symmetric", skew = b[2] == "skew-symmetric";
  if (!sym && !skew && b[2] != "general")
    throw Error(AINV_ERR_ARG, path + ": symmetry " + b[2] + " not supported");
  long long m = 0, n = 0, nz = 0;
  if (std::sscanf(size.c_str(), "%lld %lld %lld", &m, &n, &nz) != 3 || m < 1 ||
      n < 1 || nz < 0)
    throw Error(AINV_ERR_ARG, path + ": bad size line");
  if (m != n) throw Error(AINV_ERR_ARG, path + ": matrix is not square");
  if (n > INT32_MAX) throw Error(AINV_ERR_ARG, path + ": order exceeds int32");
  struct E {
    int32_t i, j;
    double v;
  };
  std::vector<E> e;
  e.reserve((size_t)nz * (sym || skew ? 2 : 1));
  char line[512];
  for (long long t = 0; t < nz; ++t) {
    if (!std::fgets(line, sizeof line, f.get()))
      throw Error(AINV_ERR_ARG, path + ": fewer entries than announced");
    long long i = 0, j = 0;
    double v = 1.0;
    const int got = pattern ? std::sscanf(line, "%lld %lld", &i, &j)
                            : std::sscanf(line, "%lld %lld %lf", &i, &j, &v);
    if (got != (pattern ? 2 : 3) || i < 1 || j < 1 || i > n || j > n)
      throw Error(AINV_ERR_ARG, path + ": bad entry line " + std::to_string(t + 1));
    e.push_back({(int32_t)(i - 1), (int32_t)(j - 1), v});
    if ((sym || skew) && i != j)
      e.push_back({(int32_t)(j - 1), (int32_t)(i - 1), skew ? -v : v});
  }
  std::sort(e.begin(), e.end(), [](const E &x, const E &y) {
    return x.i != y.i ? x.i < y.i : x.j < y.j;
  });
  HostCsr a;
  a.n = (int32_t)n;
  a.ptr.assign((size_t)n + 1, 0);
  for (size_t p = 0; p < e.size(); ++p) {
    if (p > 0 && e[p].i == e[p - 1].i && e[p].j == e[p - 1].j) {
      a.val.back() += e[p].v;
      continue;
    }
    a.idx.push_back(e[p].j);
    a.val.push_back(e[p].v);
    a.ptr[(size_t)e[p].i + 1]++;
  }
  for (int32_t i = 0; i < a.n; ++i) a.ptr[(size_t)i + 1] += a.ptr[i];
  return a;
}

std::vector<double> mm_read_vector(const std::string &path) {
  File f = open_or_throw(path, "r");
  std::vector<std::string> b;
  const std::string size = read_header(f.get(), path, b);
  if (b[0] != "array" || (b[1] != "real" && b[1] != "integer"))
    throw Error(AINV_ERR_ARG, path + ": dense real array expected");
  long long m = 0, n = 0;
  if (std::sscanf(size.c_str(), "%lld %lld", &m, &n) != 2 || m < 0 || n != 1)
    throw Error(AINV_ERR_ARG, path + ": a single column expected");
  std::vector<double> v((size_t)m);
  for (long long t = 0; t < m; ++t)
    if (std::fscanf(f.get(), "%lf", &v[(size_t)t]) != 1)
      throw Error(AINV_ERR_ARG, path + ": fewer values than announced");
  return v;
}

}  // namespace ainv

using namespace ainv;

int ainv_mm_read(const char *path, int32_t *n, int64_t *nnz, int64_t *irp,
                 int32_t *ja, double *val) {
  return guard([&] {
    if (!path) throw Error(AINV_ERR_ARG, "NULL path");
    HostCsr a = mm_read_csr(path);
    if (n) *n = a.n;
    if (nnz) *nnz = a.nnz();
    if (irp) std::copy(a.ptr.begin(), a.ptr.end(), irp);
    if (ja) std::copy(a.idx.begin(), a.idx.end(), ja);
    if (val) std::copy(a.val.begin(), a.val.end(), val);
  });
}

int ainv_mm_read_vector(const char *path, int64_t *n, double *val) {
  return guard([&] {
    if (!path) throw Error(AINV_ERR_ARG, "NULL path");
    std::vector<double> v = mm_read_vector(path);
    if (n) *n = (int64_t)v.size();
    if (val) std::copy(v.begin(), v.end(), val);
  });
}
