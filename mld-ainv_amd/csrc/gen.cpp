// Synthetic matrices of the benchmark configurations, generated on the
// host with a seeded xoshiro256** stream so that every path (GPU library,
// CPU baseline, tests) sees byte-identical inputs.
//   pde3d:    tests/pdegen/genpde_mod.f90:316-389 with the coefficient
//             functions of tests/pdegen/pde3d_base_mod.f90:12-64
//             (a = 1, b = 1/sqrt(3), c = 0)
//   aniso2d:  the 5-point pattern of genpde_mod.f90:735-775 with
//             a1 = eps, a2 = 1, b = c = 0
//   random:   lognormal row lengths, banded + long-range columns,
//             diagonally dominant
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ainv_internal.h"

namespace ainv {
namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {  // SplitMix64
      z += 0x9e3779b97f4a7c15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
      s[i] = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double normal() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  uint64_t below(uint64_t m) { return m ? next() % m : 0; }
};

// fill the caller's arrays from an assembled host CSR (or only sizes)
int emit(const HostCsr &a, int32_t *n, int64_t *nnz, int64_t *irp,
         int32_t *ja, double *val) {
  if (n) *n = a.n;
  if (nnz) *nnz = a.nnz();
  if (irp) std::memcpy(irp, a.ptr.data(), a.ptr.size() * sizeof(int64_t));
  if (ja) std::memcpy(ja, a.idx.data(), a.idx.size() * sizeof(int32_t));
  if (val) std::memcpy(val, a.val.data(), a.val.size() * sizeof(double));
  return AINV_SUCCESS;
}

}  // namespace

void gen_pde3d(int32_t N, int hconv, HostCsr &a) {
  if (N < 1 || (int64_t)N * N * N > INT32_MAX)
    throw Error(AINV_ERR_ARG, "pde3d: bad grid size");
  const double deltah = 1.0 / (double)(N + (hconv ? 2 : 1));
  const double sqdeltah = deltah * deltah;
  const double deltah2 = 2.0 * deltah;
  const double a1 = 1.0, a2 = 1.0, a3 = 1.0;
  const double b1 = 1.0 / std::sqrt(3.0), b2 = b1, b3 = b1;
  const double c = 0.0;
  const int64_t n = (int64_t)N * N * N;
  a = HostCsr();
  a.n = (int32_t)n;
  a.ptr.resize(n + 1);
  a.idx.reserve(7 * n);
  a.val.reserve(7 * n);
  a.ptr[0] = 0;
  const int64_t NN = (int64_t)N * N;
  for (int32_t ix = 0; ix < N; ++ix)
    for (int32_t iy = 0; iy < N; ++iy)
      for (int32_t iz = 0; iz < N; ++iz) {
        const int64_t i = (int64_t)ix * NN + (int64_t)iy * N + iz;
        auto put = [&](int64_t j, double v) {
          a.idx.push_back((int32_t)j);
          a.val.push_back(v);
        };
        if (ix > 0) put(i - NN, -a1 / sqdeltah - b1 / deltah2);
        if (iy > 0) put(i - N, -a2 / sqdeltah - b2 / deltah2);
        if (iz > 0) put(i - 1, -a3 / sqdeltah - b3 / deltah2);
        put(i, (2 * 1.0) * (a1 + a2 + a3) / sqdeltah + c);
        if (iz < N - 1) put(i + 1, -a3 / sqdeltah + b3 / deltah2);
        if (iy < N - 1) put(i + N, -a2 / sqdeltah + b2 / deltah2);
        if (ix < N - 1) put(i + NN, -a1 / sqdeltah + b1 / deltah2);
        a.ptr[i + 1] = (int64_t)a.idx.size();
      }
}

void gen_aniso2d(int32_t N, double eps, HostCsr &a) {
  if (N < 1 || (int64_t)N * N > INT32_MAX)
    throw Error(AINV_ERR_ARG, "aniso2d: bad grid size");
  const double h = 1.0 / (double)(N + 1);
  const double h2 = h * h;
  const double a1 = eps, a2 = 1.0;
  const int64_t n = (int64_t)N * N;
  a = HostCsr();
  a.n = (int32_t)n;
  a.ptr.resize(n + 1);
  a.idx.reserve(5 * n);
  a.val.reserve(5 * n);
  a.ptr[0] = 0;
  for (int32_t ix = 0; ix < N; ++ix)
    for (int32_t iy = 0; iy < N; ++iy) {
      const int64_t i = (int64_t)ix * N + iy;
      auto put = [&](int64_t j, double v) {
        a.idx.push_back((int32_t)j);
        a.val.push_back(v);
      };
      if (ix > 0) put(i - N, -a1 / h2);
      if (iy > 0) put(i - 1, -a2 / h2);
      put(i, 2.0 * (a1 + a2) / h2);
      if (iy < N - 1) put(i + 1, -a2 / h2);
      if (ix < N - 1) put(i + N, -a1 / h2);
      a.ptr[i + 1] = (int64_t)a.idx.size();
    }
}

void gen_random(int32_t n, uint64_t seed, double median, double sigma,
                int32_t cap, int32_t band, int32_t nlong, double dom,
                HostCsr &a) {
  if (n < 1) throw Error(AINV_ERR_ARG, "random: n must be positive");
  Rng rng(seed);
  a = HostCsr();
  a.n = n;
  a.ptr.assign((size_t)n + 1, 0);
  std::vector<int32_t> cols;
  std::vector<double> vals;
  const double mu = std::log(median);
  for (int32_t i = 0; i < n; ++i) {
    const int64_t lo = std::max<int64_t>(0, (int64_t)i - band);
    const int64_t hi = std::min<int64_t>((int64_t)n - 1, (int64_t)i + band);
    const int64_t avail = hi - lo;  // band columns other than i
    int64_t len = std::llround(std::exp(mu + sigma * rng.normal()));
    len = std::max<int64_t>(0, std::min<int64_t>(len, cap));
    len = std::min<int64_t>(len, avail);
    cols.clear();
    auto taken = [&](int32_t j) {
      return j == i || std::find(cols.begin(), cols.end(), j) != cols.end();
    };
    while ((int64_t)cols.size() < len) {
      int32_t j = (int32_t)(lo + (int64_t)rng.below((uint64_t)(hi - lo + 1)));
      if (!taken(j)) cols.push_back(j);
    }
    const int64_t want_long = std::min<int64_t>(nlong, (int64_t)n - 1 - len);
    for (int64_t t = 0; t < want_long;) {
      int32_t j = (int32_t)rng.below((uint64_t)n);
      if (!taken(j)) {
        cols.push_back(j);
        ++t;
      }
    }
    vals.resize(cols.size());
    double asum = 0.0;
    for (size_t t = 0; t < cols.size(); ++t) {
      vals[t] = rng.normal();
      asum += std::fabs(vals[t]);
    }
    cols.push_back(i);
    vals.push_back(dom * asum > 0.0 ? dom * asum : 1.0);
    std::vector<size_t> ord(cols.size());
    for (size_t t = 0; t < ord.size(); ++t) ord[t] = t;
    std::sort(ord.begin(), ord.end(),
              [&](size_t p, size_t q) { return cols[p] < cols[q]; });
    for (size_t t : ord) {
      a.idx.push_back(cols[t]);
      a.val.push_back(vals[t]);
    }
    a.ptr[i + 1] = (int64_t)a.idx.size();
  }
}

void gen_normal(uint64_t seed, int64_t count, double *out) {
  Rng rng(seed);
  for (int64_t t = 0; t < count; ++t) out[t] = rng.normal();
}

}  // namespace ainv

// ---------------------------------------------------------------- C ABI
extern "C" {

int ainv_gen_pde3d(int32_t N, int hconv, int32_t *n, int64_t *nnz,
                   int64_t *irp, int32_t *ja, double *val) {
  return ainv::guard([&] {
    ainv::HostCsr a;
    ainv::gen_pde3d(N, hconv, a);
    ainv::emit(a, n, nnz, irp, ja, val);
  });
}

int ainv_gen_aniso2d(int32_t N, double eps, int32_t *n, int64_t *nnz,
                     int64_t *irp, int32_t *ja, double *val) {
  return ainv::guard([&] {
    ainv::HostCsr a;
    ainv::gen_aniso2d(N, eps, a);
    ainv::emit(a, n, nnz, irp, ja, val);
  });
}

int ainv_gen_random(int32_t nrows, uint64_t seed, double median, double sigma,
                    int32_t cap, int32_t band, int32_t nlong, double dom,
                    int32_t *n, int64_t *nnz, int64_t *irp, int32_t *ja,
                    double *val) {
  return ainv::guard([&] {
    ainv::HostCsr a;
    ainv::gen_random(nrows, seed, median, sigma, cap, band, nlong, dom, a);
    ainv::emit(a, n, nnz, irp, ja, val);
  });
}

int ainv_gen_normal(uint64_t seed, int64_t count, double *out) {
  return ainv::guard([&] {
    if (count < 0 || (count > 0 && !out))
      throw ainv::Error(AINV_ERR_ARG, "gen_normal: bad count or output");
    ainv::gen_normal(seed, count, out);
  });
}

}  // extern "C"
