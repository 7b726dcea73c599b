// slab_getri_noalloc: the batched inverse of slab_getri with its workspace
// in a file-scope device array instead of a caller- or library-allocated
// buffer, for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops).  No counterpart in the reference.
#include "slab_internal.h"
#include "../../include/slab_hip.h"

#include <mutex>

// The workspace of every call on a device: W and X (padded copies of the
// triangles and their inverses), the pair products and the inverse
// permutations of getri.  Each device holds its own copy for the life of the
// process.  The getri kernels write every word they read (the padded copies
// in full, the upper strips of the triangular inverses as zeros), so nothing
// in it needs clearing between calls.
static __device__ __align__(256) unsigned char scratch[SLAB_GETRI_NOALLOC_MAX_BYTES];

namespace {

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; host code uses only this
// address, never the symbol's name.
hipError_t scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (p == nullptr) return hipErrorInvalidSymbol;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int slab_getri_noalloc(int dtype, int batch, int n, const void* LU, int64_t lda, int64_t lu_bs,
                                  const int* ipiv, int64_t p_bs, void* X, int64_t ldx, int64_t x_bs, void* stream) {
  if (batch < 0 || n < 0 || (dtype != SLAB_F64 && dtype != SLAB_F32)) return -1;
  if (batch == 0 || n == 0) return 0;
  if (slab_getri_ws(dtype, batch, n) > (size_t)SLAB_GETRI_NOALLOC_MAX_BYTES) return -2;
  void* ws = nullptr;
  const hipError_t e = scratch_address(&ws);
  if (e != hipSuccess) return -(1000 + (int)e);
  return slab_getri(dtype, batch, n, LU, lda, lu_bs, ipiv, p_bs, X, ldx, x_bs, ws, stream);
}
