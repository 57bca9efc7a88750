// Allocation-free design change: uft_mma_design_change_noalloc.
//
// The same two kernels as uft_mma_design_change (k_design_change writes one
// max partial per block of 256 elements and sets xold := x; k_final_max
// reduces the partials in a fixed order), but on caller-owned device arrays
// and with the block partials in a file-scope device array instead of a
// context's hipMalloc'd buffer.  The call itself allocates nothing (no
// hipMalloc, hipFree, stream or event creation), so it can run where
// allocation is forbidden, e.g. under stream capture.
//
// The scratch array exists once per GPU for the whole process (512 KiB).  Its
// device address is looked up once per device with hipGetSymbolAddress on the
// first call that names that device; the array's name is never handed to a
// kernel or a copy from host code.  Every call on one GPU shares that GPU's
// array: calls that overlap on one GPU (different streams or host threads
// without ordering between them) race on it.  Nothing in it needs clearing:
// each call writes every partial it later reads.
#include <hip/hip_runtime.h>

#include <mutex>

#include "uft_reduce.h"

#include "uft_internal.h"

namespace {

constexpr int kBlock = uftr::RTPB;                // threads per block of k_design_change
constexpr int64_t kMaxElems = int64_t(1) << 24;   // largest n accepted
constexpr int kMaxBlocks = (int)(kMaxElems / kBlock);
constexpr int kMaxDevices = 64;

double *g_scratch_ptr[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];

}  // namespace

static __device__ __align__(256) double scratch[kMaxBlocks];

namespace {

// Device address of this GPU's scratch array, looked up with that device
// current (each GPU holds its own copy at its own address); the previously
// current device is restored.
void lookup_scratch(int device) {
  int prev = -1;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess) e = hipSetDevice(device);
  void *p = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? static_cast<double *>(p) : nullptr;
}

}  // namespace

extern "C" int uft_mma_design_change_noalloc(int device, const double *x, double *xold, int64_t n, double *ch,
                                             void *stream) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return uft_set_error(nullptr, UFT_ERR_HIP, "uft_mma_design_change_noalloc: no HIP device");
  if (device < 0 || device >= ndev || device >= kMaxDevices)
    return uft_set_error(nullptr, UFT_ERR_ARG, "uft_mma_design_change_noalloc: device %d out of range", device);
  if (n < 0 || n > kMaxElems)
    return uft_set_error(nullptr, UFT_ERR_ARG,
                         "uft_mma_design_change_noalloc: n = %lld outside [0, %lld] (scratch holds %d block "
                         "partials)", (long long)n, (long long)kMaxElems, kMaxBlocks);
  if (!ch || (n > 0 && (!x || !xold)))
    return uft_set_error(nullptr, UFT_ERR_ARG, "uft_mma_design_change_noalloc: null device pointer");
  std::call_once(g_scratch_once[device], lookup_scratch, device);
  if (g_scratch_err[device] != hipSuccess)
    return uft_set_error(nullptr, UFT_ERR_HIP, "uft_mma_design_change_noalloc: hipGetSymbolAddress: %s",
                         hipGetErrorString(g_scratch_err[device]));
  double *part = g_scratch_ptr[device];
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess || (prev != device && hipSetDevice(device) != hipSuccess))
    return uft_set_error(nullptr, UFT_ERR_HIP, "uft_mma_design_change_noalloc: cannot select device %d", device);
  hipStream_t s = static_cast<hipStream_t>(stream);
  (void)hipGetLastError();
  const int nb = (int)((n + kBlock - 1) / kBlock);
  if (nb > 0) uftk::design_change(s, n, x, xold, part);
  uftk::final_max(s, part, nb, ch);
  const hipError_t e = hipGetLastError();
  if (prev != device) (void)hipSetDevice(prev);
  if (e != hipSuccess)
    return uft_set_error(nullptr, UFT_ERR_HIP, "uft_mma_design_change_noalloc: launch failed: %s",
                         hipGetErrorString(e));
  return UFT_OK;
}
