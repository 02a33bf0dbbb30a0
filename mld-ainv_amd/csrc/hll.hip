// Device-resident sparse storage: owned buffers, CSR upload/download, a
// deterministic CSR transpose (radix sort on (column,row) keys) and the
// CSR -> sliced-ELL (HLL) conversion that lays factors out for the apply
// kernels.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdio>
#include <numeric>
#include <vector>

#include "ainv_internal.h"

namespace ainv {

void DevBuf::alloc(size_t b) {
  release();
  if (b == 0) return;
  AINV_HIP_CHECK(hipMalloc(&p, b));
  bytes = b;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

int value_bytes(int prec) {
  switch (prec) {
    case AINV_PREC_F64: return 8;
    case AINV_PREC_F32: return 4;
    case AINV_PREC_BF16: return 2;
  }
  throw Error(AINV_ERR_ARG, "unknown factor precision " + std::to_string(prec));
}

void csr_upload(const HostCsr &h, DevCsr &d, hipStream_t s) {
  d.n = h.n;
  d.nnz = h.nnz();
  d.ptr.alloc((size_t)(h.n + 1) * sizeof(int64_t));
  d.idx.alloc((size_t)d.nnz * sizeof(int32_t));
  d.val.alloc((size_t)d.nnz * sizeof(double));
  AINV_HIP_CHECK(hipMemcpyAsync(d.ptr.p, h.ptr.data(), d.ptr.bytes,
                                hipMemcpyHostToDevice, s));
  if (d.nnz) {
    AINV_HIP_CHECK(hipMemcpyAsync(d.idx.p, h.idx.data(), d.idx.bytes,
                                  hipMemcpyHostToDevice, s));
    AINV_HIP_CHECK(hipMemcpyAsync(d.val.p, h.val.data(), d.val.bytes,
                                  hipMemcpyHostToDevice, s));
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

void csr_download(const DevCsr &d, HostCsr &h, hipStream_t s) {
  h.n = d.n;
  h.ptr.resize(d.n + 1);
  h.idx.resize(d.nnz);
  h.val.resize(d.nnz);
  AINV_HIP_CHECK(hipMemcpyAsync(h.ptr.data(), d.ptr.p, d.ptr.bytes,
                                hipMemcpyDeviceToHost, s));
  if (d.nnz) {
    AINV_HIP_CHECK(hipMemcpyAsync(h.idx.data(), d.idx.p, d.idx.bytes,
                                  hipMemcpyDeviceToHost, s));
    AINV_HIP_CHECK(hipMemcpyAsync(h.val.data(), d.val.p, d.val.bytes,
                                  hipMemcpyDeviceToHost, s));
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

namespace {

constexpr int kThreads = 256;

inline unsigned grid_for(int64_t work, int per_block = kThreads) {
  int64_t g = (work + per_block - 1) / per_block;
  return (unsigned)(g < 1 ? 1 : g);
}

// key = column << 32 | row for every stored entry (one thread per row)
__global__ void k_transpose_keys(int32_t n, const int64_t *__restrict__ ptr,
                                 const int32_t *__restrict__ idx,
                                 uint64_t *__restrict__ keys) {
  int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int64_t e = ptr[r]; e < ptr[r + 1]; ++e)
    keys[e] = ((uint64_t)(uint32_t)idx[e] << 32) | (uint32_t)r;
}

// row pointer of the transpose: first position whose column >= c
__global__ void k_transpose_ptr(int32_t n, int64_t nnz,
                                const uint64_t *__restrict__ keys,
                                int64_t *__restrict__ tptr,
                                int32_t *__restrict__ tidx) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nnz) tidx[t] = (int32_t)(uint32_t)(keys[t] & 0xffffffffu);
  if (t <= n) {
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if ((int64_t)(keys[mid] >> 32) < t)
        lo = mid + 1;
      else
        hi = mid;
    }
    tptr[t] = lo;
  }
}

int bits_for(int64_t v) {
  int b = 1;
  while ((int64_t(1) << b) <= v && b < 62) ++b;
  return b;
}

}  // namespace

void csr_transpose(const DevCsr &a, DevCsr &t, hipStream_t s) {
  const int32_t n = a.n;
  const int64_t nnz = a.nnz;
  t.n = n;
  t.nnz = nnz;
  t.ptr.alloc((size_t)(n + 1) * sizeof(int64_t));
  t.idx.alloc((size_t)nnz * sizeof(int32_t));
  t.val.alloc((size_t)nnz * sizeof(double));
  if (nnz == 0) {
    AINV_HIP_CHECK(hipMemsetAsync(t.ptr.p, 0, t.ptr.bytes, s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
    return;
  }
  DevBuf kin, kout;
  kin.alloc((size_t)nnz * sizeof(uint64_t));
  kout.alloc((size_t)nnz * sizeof(uint64_t));
  hipLaunchKernelGGL(k_transpose_keys, dim3(grid_for(n)), dim3(kThreads), 0, s,
                     n, a.ptr.as<int64_t>(), a.idx.as<int32_t>(),
                     kin.as<uint64_t>());
  AINV_HIP_CHECK(hipGetLastError());
  const int end_bit = 32 + bits_for(n);
  size_t tmp_bytes = 0;
  AINV_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
      nullptr, tmp_bytes, kin.as<uint64_t>(), kout.as<uint64_t>(),
      a.val.as<double>(), t.val.as<double>(), (int)nnz, 0, end_bit, s));
  DevBuf tmp;
  tmp.alloc(tmp_bytes ? tmp_bytes : 16);
  AINV_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(
      tmp.p, tmp_bytes, kin.as<uint64_t>(), kout.as<uint64_t>(),
      a.val.as<double>(), t.val.as<double>(), (int)nnz, 0, end_bit, s));
  hipLaunchKernelGGL(k_transpose_ptr, dim3(grid_for(std::max<int64_t>(nnz, n + 1))),
                     dim3(kThreads), 0, s, n, nnz, kout.as<uint64_t>(),
                     t.ptr.as<int64_t>(), t.idx.as<int32_t>());
  AINV_HIP_CHECK(hipGetLastError());
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

namespace {

// per slice: width = max row length (one thread per slice); rlen per
// stored row (perm: the row each stored row holds, null: in order)
__global__ void k_hll_width(int32_t n, int32_t nslices,
                            const int64_t *__restrict__ ptr,
                            const int32_t *__restrict__ perm,
                            int32_t *__restrict__ rlen,
                            int64_t *__restrict__ w64) {
  int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nslices) return;
  if (s == nslices) {
    w64[s] = 0;
    return;
  }
  int32_t w = 0;
  for (int l = 0; l < kSlice; ++l) {
    int32_t r = s * kSlice + l;
    if (r >= n) break;
    const int32_t pr = perm ? perm[r] : r;
    int32_t len = (int32_t)(ptr[pr + 1] - ptr[pr]);
    rlen[r] = len;
    w = len > w ? len : w;
  }
  w64[s] = (int64_t)w * kSlice;
}

__device__ inline void store_val(double *p, double v) { *p = v; }
__device__ inline void store_val(float *p, double v) { *p = (float)v; }
// BF16: round-to-nearest-even of the FP32 rounding of v
__device__ inline void store_val(uint16_t *p, double v) {
  float f = (float)v;
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) {  // inf / nan
    *p = (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    return;
  }
  u += 0x7fffu + ((u >> 16) & 1u);
  *p = (uint16_t)(u >> 16);
}

// one thread per stored row of the padded row space: writes the row it
// holds (perm, or itself) column-major
template <typename VT>
__global__ void k_hll_fill(int32_t n, int32_t nslices,
                           const int64_t *__restrict__ ptr,
                           const int32_t *__restrict__ idx,
                           const double *__restrict__ val,
                           const int64_t *__restrict__ sptr,
                           const int32_t *__restrict__ perm,
                           int32_t *__restrict__ col, VT *__restrict__ hv,
                           int32_t ncols) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int32_t s = (int32_t)(r / kSlice);
  if (s >= nslices) return;
  int lane = (int)(r % kSlice);
  int64_t base = sptr[s];
  int32_t width = (int32_t)((sptr[s + 1] - base) / kSlice);
  int64_t beg = 0;
  int32_t len = 0;
  const int32_t pr = r < n ? (perm ? perm[r] : (int32_t)r) : n;
  int32_t self = pr < n ? pr : n - 1;
  if (self >= ncols) self = ncols - 1;  // rectangular: stay inside x
  if (pr < n) {
    beg = ptr[pr];
    len = (int32_t)(ptr[pr + 1] - beg);
  }
  for (int32_t j = 0; j < width; ++j) {
    int64_t q = base + (int64_t)j * kSlice + lane;
    if (j < len) {
      col[q] = idx[beg + j];
      store_val(hv + q, val[beg + j]);
    } else {
      col[q] = self;
      store_val(hv + q, 0.0);
    }
  }
}

template <typename VT>
__global__ void k_convert(int64_t n, const double *__restrict__ src,
                          VT *__restrict__ dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) store_val(dst + i, src[i]);
}

}  // namespace

void hll_attach_csr(const HostCsr &a, Hll &h, hipStream_t s) {
  if (a.n != h.n) throw Error(AINV_ERR_INTERNAL, "CSR copy of another matrix");
  if (h.prec == AINV_PREC_BF16)
    throw Error(AINV_ERR_INTERNAL, "CSR copy of a BF16 matrix");
  const int64_t nnz = a.nnz();
  h.cptr.alloc((size_t)(a.n + 1) * sizeof(int64_t));
  h.cidx.alloc((size_t)std::max<int64_t>(nnz, 1) * sizeof(int32_t));
  const int vb = value_bytes(h.prec);
  h.cval.alloc((size_t)std::max<int64_t>(nnz, 1) * vb);
  AINV_HIP_CHECK(hipMemcpyAsync(h.cptr.p, a.ptr.data(), h.cptr.bytes,
                                hipMemcpyHostToDevice, s));
  if (nnz) {
    AINV_HIP_CHECK(hipMemcpyAsync(h.cidx.p, a.idx.data(),
                                  (size_t)nnz * sizeof(int32_t),
                                  hipMemcpyHostToDevice, s));
    DevBuf v64;
    v64.alloc((size_t)nnz * sizeof(double));
    AINV_HIP_CHECK(hipMemcpyAsync(v64.p, a.val.data(), v64.bytes,
                                  hipMemcpyHostToDevice, s));
    DevBuf conv;
    convert_vector(v64.as<double>(), nnz, h.prec, conv, s);
    AINV_HIP_CHECK(hipMemcpyAsync(h.cval.p, conv.p, (size_t)nnz * vb,
                                  hipMemcpyDeviceToDevice, s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
  }
}

int64_t hll_sorted_padding(const std::vector<int32_t> &len, int32_t sigma,
                           std::vector<int32_t> *perm) {
  const int32_t n = (int32_t)len.size();
  const int32_t ns = (n + kSlice - 1) / kSlice;
  std::vector<int32_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  if (sigma > 0)
    for (int64_t b = 0; b < n; b += sigma) {
      const int64_t e = std::min<int64_t>(b + sigma, n);
      std::stable_sort(ord.begin() + b, ord.begin() + e,
                       [&](int32_t x, int32_t y) { return len[x] > len[y]; });
    }
  std::vector<int32_t> sw(ns);  // slice widths
  int64_t padded = 0;
  for (int32_t t = 0; t < ns; ++t) {
    int32_t w = 0;
    for (int32_t i = t * kSlice; i < std::min(n, (t + 1) * kSlice); ++i)
      w = std::max(w, len[ord[i]]);
    sw[t] = w;
    padded += (int64_t)w * kSlice;
  }
  if (!perm) return padded;
  perm->assign((size_t)ns * kSlice, n);
  if (sigma <= 0) {
    std::copy(ord.begin(), ord.end(), perm->begin());
    return padded;
  }
  // Sorting packs each window's long rows into its first slices, so the
  // heavy slices recur with the window's period, which can alias with the
  // stride of the persistent kernels' block loops (one block taking every
  // heavy group).  The groups of 4 slices (a workgroup's) of each of the 8
  // contiguous ranges the XCDs take are therefore stored heaviest first:
  // a launch starts its longest groups first and a persistent block's
  // stride takes one group of each band of weights.  The last group, which
  // holds the rows past n, stays last.
  constexpr int32_t G = kSliceGroup;
  const int32_t ng = (ns + G - 1) / G;
  const int32_t nfull = n / (G * kSlice);  // groups of 256 rows
  std::vector<int64_t> gw(ng, 0);
  for (int32_t t = 0; t < ns; ++t) gw[t / G] += sw[t];
  std::vector<int32_t> go(ng);
  std::iota(go.begin(), go.end(), 0);
  for (int32_t x = 0; x < kXcdCount; ++x) {
    const int32_t b = (int32_t)((int64_t)ng * x / kXcdCount);
    const int32_t e = std::min<int32_t>(
        (int32_t)((int64_t)ng * (x + 1) / kXcdCount), nfull);
    if (e > b)
      std::stable_sort(go.begin() + b, go.begin() + e,
                       [&](int32_t p, int32_t q) { return gw[p] > gw[q]; });
  }
  for (int32_t g = 0, i = 0; g < ng; ++g)
    for (int32_t r = go[g] * G * kSlice;
         r < std::min<int64_t>((int64_t)(go[g] + 1) * G * kSlice, n); ++r)
      (*perm)[i++] = ord[r];
  return padded;
}

std::string hll_layout_json(const Hll &h) {
  char b[96];
  std::snprintf(b, sizeof b, "%.3f", h.width_mean);
  return "{\"nnz\": " + std::to_string(h.nnz) +
         ", \"padded\": " + std::to_string(h.padded) +
         ", \"unsorted_padded\": " + std::to_string(h.unsorted_padded) +
         ", \"sort_window\": " + std::to_string(h.sigma) +
         ", \"slices\": " + std::to_string(h.nslices) +
         ", \"max_width\": " + std::to_string(h.max_width) +
         ", \"width_mean\": " + b +
         ", \"width_p50\": " + std::to_string(h.width_p50) +
         ", \"width_p99\": " + std::to_string(h.width_p99) + "}";
}

int32_t hll_pick_sigma(const std::vector<int32_t> &len) {
  int64_t nnz = 0;
  for (const int32_t l : len) nnz += l;
  const int64_t p0 = hll_sorted_padding(len, 0);
  if (nnz == 0 || p0 * 4 < nnz * 5) return 0;  // pads < 25 %
  // windows beyond 4096 rows pad a little less but spread a slice's rows,
  // and so its gathers, further (C4: 16384 rows padded 1.04x against 1.19x
  // and served 2% fewer solves/s, the single-column path 8% fewer)
  constexpr int32_t win[] = {256, 1024, 4096};
  constexpr int nwin = 3;
  int64_t pw[nwin], best = p0;
  for (int i = 0; i < nwin; ++i) {
    pw[i] = hll_sorted_padding(len, win[i]);
    best = std::min(best, pw[i]);
  }
  if (best * 5 > p0 * 4) return 0;  // saves < 20 %
  for (int i = 0; i < nwin; ++i)
    if (pw[i] * 100 <= best * 105) return win[i];
  return win[nwin - 1];
}

void hll_from_csr(const DevCsr &a, int prec, Hll &h, hipStream_t s,
                  int32_t ncols, int32_t sigma) {
  if (ncols < 0) ncols = a.n;
  if (ncols < 1) ncols = 1;
  if (sigma != 0 && sigma != kSortAuto && (sigma < 0 || sigma % kSlice))
    throw Error(AINV_ERR_ARG, "row-sort window must be a multiple of 64");
  const int vb = value_bytes(prec);
  h = Hll();
  h.n = a.n;
  h.ncols = ncols;
  h.prec = prec;
  h.nnz = a.nnz;
  h.nslices = (a.n + kSlice - 1) / kSlice;
  if (sigma != 0 && a.n > 0) {
    // row lengths on the host: the sort and its window are host work
    std::vector<int64_t> ptr(a.n + 1);
    AINV_HIP_CHECK(hipMemcpyAsync(ptr.data(), a.ptr.p,
                                  ptr.size() * sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
    AINV_HIP_CHECK(hipStreamSynchronize(s));
    std::vector<int32_t> len(a.n);
    for (int32_t i = 0; i < a.n; ++i) len[i] = (int32_t)(ptr[i + 1] - ptr[i]);
    h.unsorted_padded = hll_sorted_padding(len, 0);
    if (sigma == kSortAuto) sigma = hll_pick_sigma(len);
    if (sigma > 0) {
      std::vector<int32_t> perm;
      hll_sorted_padding(len, sigma, &perm);
      h.perm.alloc(perm.size() * sizeof(int32_t));
      AINV_HIP_CHECK(hipMemcpyAsync(h.perm.p, perm.data(), h.perm.bytes,
                                    hipMemcpyHostToDevice, s));
      AINV_HIP_CHECK(hipStreamSynchronize(s));
      h.sigma = sigma;
    }
  }
  h.sptr.alloc((size_t)(h.nslices + 1) * sizeof(int64_t));
  h.rlen.alloc((size_t)std::max(a.n, 1) * sizeof(int32_t));
  DevBuf w64;
  w64.alloc((size_t)(h.nslices + 1) * sizeof(int64_t));
  hipLaunchKernelGGL(k_hll_width, dim3(grid_for(h.nslices + 1)),
                     dim3(kThreads), 0, s, a.n, h.nslices, a.ptr.as<int64_t>(),
                     h.perm.as<int32_t>(), h.rlen.as<int32_t>(),
                     w64.as<int64_t>());
  AINV_HIP_CHECK(hipGetLastError());
  size_t tmp_bytes = 0;
  AINV_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(
      nullptr, tmp_bytes, w64.as<int64_t>(), h.sptr.as<int64_t>(),
      h.nslices + 1, s));
  DevBuf tmp;
  tmp.alloc(tmp_bytes ? tmp_bytes : 16);
  AINV_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(
      tmp.p, tmp_bytes, w64.as<int64_t>(), h.sptr.as<int64_t>(),
      h.nslices + 1, s));
  AINV_HIP_CHECK(hipMemcpyAsync(&h.padded, h.sptr.as<int64_t>() + h.nslices,
                                sizeof(int64_t), hipMemcpyDeviceToHost, s));
  // widest slice, for reporting
  std::vector<int64_t> hs(h.nslices + 1);
  AINV_HIP_CHECK(hipMemcpyAsync(hs.data(), h.sptr.p, hs.size() * sizeof(int64_t),
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  if (!h.perm.p) h.unsorted_padded = h.padded;
  if (h.nslices > 0) {
    std::vector<int32_t> wd(h.nslices);
    for (int32_t i = 0; i < h.nslices; ++i)
      wd[i] = (int32_t)((hs[i + 1] - hs[i]) / kSlice);
    h.width_mean = (double)h.padded / kSlice / h.nslices;
    std::sort(wd.begin(), wd.end());
    h.max_width = wd.back();
    h.width_p50 = wd[wd.size() / 2];
    h.width_p99 = wd[std::min(wd.size() - 1, wd.size() * 99 / 100)];
  }
  h.col.alloc((size_t)h.padded * sizeof(int32_t));
  h.val.alloc((size_t)h.padded * vb);
  if (h.padded > 0) {
    dim3 g(grid_for((int64_t)h.nslices * kSlice));
    switch (prec) {
      case AINV_PREC_F64:
        hipLaunchKernelGGL(k_hll_fill<double>, g, dim3(kThreads), 0, s, a.n,
                           h.nslices, a.ptr.as<int64_t>(), a.idx.as<int32_t>(),
                           a.val.as<double>(), h.sptr.as<int64_t>(),
                           h.perm.as<int32_t>(), h.col.as<int32_t>(),
                           h.val.as<double>(), ncols);
        break;
      case AINV_PREC_F32:
        hipLaunchKernelGGL(k_hll_fill<float>, g, dim3(kThreads), 0, s, a.n,
                           h.nslices, a.ptr.as<int64_t>(), a.idx.as<int32_t>(),
                           a.val.as<double>(), h.sptr.as<int64_t>(),
                           h.perm.as<int32_t>(), h.col.as<int32_t>(),
                           h.val.as<float>(), ncols);
        break;
      default:
        hipLaunchKernelGGL(k_hll_fill<uint16_t>, g, dim3(kThreads), 0, s, a.n,
                           h.nslices, a.ptr.as<int64_t>(), a.idx.as<int32_t>(),
                           a.val.as<double>(), h.sptr.as<int64_t>(),
                           h.perm.as<int32_t>(), h.col.as<int32_t>(),
                           h.val.as<uint16_t>(), ncols);
    }
    AINV_HIP_CHECK(hipGetLastError());
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

namespace {
// reach (rows) of each slice's middle row -- its last row for a short last
// slice -- to its first and its last column (entries sorted by column)
__global__ void k_slice_reach(int32_t n, int32_t nslices,
                              const int64_t *__restrict__ sptr,
                              const int32_t *__restrict__ perm,
                              const int32_t *__restrict__ rlen,
                              const int32_t *__restrict__ col,
                              int32_t *__restrict__ out) {
  const int32_t s = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= nslices) return;
  const int32_t r0 = s * kSlice;
  const int32_t lane = min(kSlice / 2, n - 1 - r0);
  const int32_t len = rlen[r0 + lane];
  const int32_t r = perm ? perm[r0 + lane] : r0 + lane;
  int32_t reach = 0;
  if (len > 0) {
    const int32_t first = col[sptr[s] + lane];
    const int32_t last = col[sptr[s] + (int64_t)(len - 1) * kSlice + lane];
    reach = max(max(r - first, last - r), 0);
  }
  out[s] = reach;
}
}  // namespace

int32_t hll_reach_groups(const Hll &h, hipStream_t s) {
  if (h.nslices == 0 || h.padded == 0) return 0;
  DevBuf out;
  out.alloc((size_t)h.nslices * sizeof(int32_t));
  hipLaunchKernelGGL(k_slice_reach, dim3(grid_for(h.nslices)), dim3(kThreads),
                     0, s, h.n, h.nslices, h.sptr.as<int64_t>(),
                     h.perm.as<int32_t>(), h.rlen.as<int32_t>(),
                     h.col.as<int32_t>(),
                     out.as<int32_t>());
  AINV_HIP_CHECK(hipGetLastError());
  std::vector<int32_t> r(h.nslices);
  AINV_HIP_CHECK(hipMemcpyAsync(r.data(), out.p, out.bytes,
                                hipMemcpyDeviceToHost, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  std::nth_element(r.begin(), r.begin() + r.size() / 2, r.end());
  return r[r.size() / 2] / (kSliceGroup * kSlice);
}

void hll_set_order(Hll &h, int32_t period, int32_t band, hipStream_t s) {
  const int32_t ngrp = (h.nslices + kSliceGroup - 1) / kSliceGroup;
  const bool strips = period > 0 && band > 0 && ngrp >= kXcdCount;
  // once allocated the table stays (captured graphs hold its address);
  // linear order is then the identity
  if (!strips && !h.gorder.p) {
    h.order_period = h.order_band = 0;
    return;
  }
  if (!h.gorder.p) h.gorder.alloc((size_t)ngrp * sizeof(int32_t));
  std::vector<int32_t> ord(ngrp);
  if (!strips) {
    for (int32_t g = 0; g < ngrp; ++g) ord[g] = g;
  } else {
    band = std::min(band, period);
    // the XCD ranges of a launch of ngrp workgroups (the apply kernels'
    // renumbering: XCD x runs a contiguous range, the first ngrp % 8 ranges
    // one group longer)
    const int32_t per = ngrp / kXcdCount, rem = ngrp % kXcdCount;
    int32_t pos = 0;
    for (int32_t x = 0; x < kXcdCount; ++x) {
      const int32_t a = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
      const int32_t len = per + (x < rem ? 1 : 0);
      const int32_t planes = (len + period - 1) / period;
      for (int32_t b0 = 0; b0 < period; b0 += band)
        for (int32_t p = 0; p < planes; ++p)
          for (int32_t o = b0; o < std::min(b0 + band, period); ++o) {
            const int32_t g = p * period + o;
            if (g < len) ord[pos++] = a + g;
          }
    }
    if (pos != ngrp)
      throw Error(AINV_ERR_INTERNAL, "slice-group order is not a permutation");
  }
  AINV_HIP_CHECK(hipMemcpyAsync(h.gorder.p, ord.data(),
                                (size_t)ngrp * sizeof(int32_t),
                                hipMemcpyHostToDevice, s));
  AINV_HIP_CHECK(hipStreamSynchronize(s));
  h.order_period = strips ? period : 0;
  h.order_band = strips ? band : 0;
}

namespace {
__global__ __launch_bounds__(256) void k_stream_copy(
    int64_t n16, const uint4 *__restrict__ src, uint4 *__restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

void stream_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
  const int64_t n16 = (int64_t)(bytes / 16);
  int64_t g = (n16 + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)g), dim3(256), 0, s, n16,
                     static_cast<const uint4 *>(src), static_cast<uint4 *>(dst));
  AINV_HIP_CHECK(hipGetLastError());
}

namespace {
// sv%update_a pivots (mld_d_base_ainv_update_a.f90:53-79): one thread per
// row merges row i of the stored W with row i of Z^T (= column i of Z),
// both sorted, summing W(i,k) * (x_k * Z(k,i)) in ascending k; then
// d_i = 1 / (1/d_i + ee).  Compiled without contraction: same roundings as
// the sequential restatement.
__global__ void k_update_pivots(int32_t n, const int64_t *__restrict__ wp,
                                const int32_t *__restrict__ wi,
                                const double *__restrict__ wv,
                                const int64_t *__restrict__ zp,
                                const int32_t *__restrict__ zi,
                                const double *__restrict__ zv,
                                const double *__restrict__ x,
                                double *__restrict__ d) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  double ee = 0.0;
  int64_t p = wp[i], q = zp[i];
  const int64_t pe = wp[i + 1], qe = zp[i + 1];
  while (p < pe && q < qe) {
    const int32_t a = wi[p], b = zi[q];
    if (a == b) {
      const double sz = x[b] * zv[q];
      ee = ee + wv[p] * sz;
      ++p;
      ++q;
    } else if (a < b) {
      ++p;
    } else {
      ++q;
    }
  }
  double dd = 1.0 / d[i];
  dd = dd + ee;
  d[i] = 1.0 / dd;
}
}  // namespace

void update_pivots(const DevCsr &w, const DevCsr &zt, const double *x,
                   double *d, hipStream_t s) {
  if (w.n == 0) return;
  hipLaunchKernelGGL(k_update_pivots, dim3((unsigned)((w.n + 255) / 256)),
                     dim3(256), 0, s, w.n, w.ptr.as<int64_t>(),
                     w.idx.as<int32_t>(), w.val.as<double>(),
                     zt.ptr.as<int64_t>(), zt.idx.as<int32_t>(),
                     zt.val.as<double>(), x, d);
  AINV_HIP_CHECK(hipGetLastError());
}

void convert_vector(const double *src, int64_t n, int prec, DevBuf &dst,
                    hipStream_t s) {
  // d is kept FP64 for F64 factors and FP32 otherwise
  if (prec == AINV_PREC_F64) {
    dst.alloc((size_t)n * sizeof(double));
    AINV_HIP_CHECK(hipMemcpyAsync(dst.p, src, dst.bytes,
                                  hipMemcpyDeviceToDevice, s));
  } else {
    dst.alloc((size_t)n * sizeof(float));
    hipLaunchKernelGGL(k_convert<float>, dim3(grid_for(n)), dim3(kThreads), 0,
                       s, n, src, dst.as<float>());
    AINV_HIP_CHECK(hipGetLastError());
  }
  AINV_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace ainv
