// The preconditioner apply without any allocation (ainv_apply_noalloc): the
// same two kernel passes as ainv_apply, with the intermediate
// t = d o (W x) kept in a file-scope device array instead of the handle's
// buffer, so that callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc)
// can apply a built preconditioner.
//
// The array is one per GPU the library runs on and lives until the process
// ends.  Its device address is taken once per device with
// hipGetSymbolAddress (at ainv_init, or at the first call on a device),
// with that device current, and only that address is handed to kernels.
// Every call on a GPU shares that GPU's array, so calls that overlap on one
// GPU (other threads, other streams) race on it: serialize them.
#include <mutex>

#include "objects.h"

namespace ainv {
namespace {

// 64 MiB: t of n rows and K columns (K = k rounded up to a power of two
// when ldx = ldy = K, else k) takes n K 4 bytes with FP32 / BF16 factors
// and n K 8 with FP64 factors; larger calls are refused
constexpr size_t kScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace

// The device address of this GPU's scratch array; the caller has made
// `device` current.
void *noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw Error(AINV_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void *p = nullptr;
    AINV_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(AINV_ERR_HIP, "no address for the scratch array");
    g_scratch[device] = p;
  }
  return g_scratch[device];
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

}  // namespace ainv

using namespace ainv;

extern "C" int ainv_apply_noalloc(ainv_h h, double alpha, const double *x,
                                  double beta, double *y, char trans, int k,
                                  int64_t ldx, int64_t ldy, void *stream) {
  return guard([&] {
    if (!h) throw Error(AINV_ERR_ARG, "NULL handle");
    if (!x || !y) throw Error(AINV_ERR_ARG, "NULL vector");
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->built) throw Error(AINV_ERR_STATE, "apply before build");
    if (h->kind == AINV_KIND_ML)
      throw Error(AINV_ERR_ARG,
                  "the multilevel cycle's workspace exceeds the scratch "
                  "array: use ainv_apply");
    const char t = (char)std::toupper((unsigned char)trans);
    if (t != 'N' && t != 'T' && t != 'C')
      throw Error(AINV_ERR_ARG, std::string("invalid trans '") + trans + "'");
    if (k < 1 || k > 64) throw Error(AINV_ERR_ARG, "k must be in 1..64");
    if (ldx < k || ldy < k)
      throw Error(AINV_ERR_ARG, "ldx and ldy must be >= k");
    if (t != 'N' && !h->have_t)
      throw Error(AINV_ERR_STATE,
                  "the transposed layout is built by the first 'T' apply of "
                  "ainv_apply (it allocates): call that once first");
    AINV_HIP_CHECK(hipSetDevice(h->device));
    void *tb = noalloc_setup(h->device);
    int K = 1;
    while (K < k) K <<= 1;
    const bool tuned = (t == 'N' && ldx == K && ldy == K);
    const int map = tuned ? tuned_apply_map(h, K) : 0;
    const int64_t ldt = tuned ? K : k;
    const bool f32t = h->prec != AINV_PREC_F64;
    const size_t need =
        (size_t)h->n * ldt * (f32t ? sizeof(float) : sizeof(double));
    if (need > noalloc_scratch_bytes())
      throw Error(AINV_ERR_CAPACITY,
                  "n x " + std::to_string(ldt) + " intermediate of " +
                      std::to_string(need) + " bytes exceeds the " +
                      std::to_string(noalloc_scratch_bytes()) +
                      "-byte scratch array: use ainv_apply");
    hipStream_t s = stream ? (hipStream_t)stream : h->stream;
    const Hll &first = t == 'N' ? h->hw : h->hzt;
    const Hll &second = t == 'N' ? h->hz : h->hwt;
    hll_spmm_scale(first, h->dv.p, x, false, ldx, tb, ldt, k, s, map_w(map));
    hll_spmm_axpby(second, alpha, tb, f32t, ldt, beta, y, ldy, k, s,
                   map_z(map));
  });
}
