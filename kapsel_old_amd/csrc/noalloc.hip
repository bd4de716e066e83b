// noalloc.hip -- spm_fft3d_noalloc: the context-free transform of
// spm_fft3d for callers that may not allocate device memory (code under
// HIP graph capture, tight serving loops).  Its root tables go to a
// file-scope device array instead of a workspace.
//
// Each GPU holds one copy of the array until the process ends; every call
// on that GPU rewrites it, so calls that overlap on one GPU (two streams,
// two threads) race on it.  The host never names the array in a kernel
// argument: its address on each device comes from hipGetSymbolAddress,
// taken once per device with that device current.
#include <mutex>

#include "spm_internal.h"

namespace {

// The root tables of one transform of up to 1024 points per axis need
// 49,408 bytes (spm_fft_workspace_size(1024, 1024, 1024)).
constexpr size_t kScratchBytes = 65536;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
void *g_scratch_dev[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) char scratch[kScratchBytes];

extern "C" int spm_fft3d_noalloc(int nx, int ny, int nz, int direction, double *field, void *stream) {
  spm::global_error().clear();
  const size_t need = spm::fft3d_workspace_bytes(nx, ny, nz);
  if (need == 0) {
    spm::global_error() = "transform sizes must be powers of two in [8, 1024] or 48, 80, 96, 160, 192";
    return SPM_ERR_INVALID;
  }
  if (need > kScratchBytes) {
    spm::global_error() = "transform needs more than the 65536-byte static scratch";
    return SPM_ERR_INVALID;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    spm::global_error() = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return SPM_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    spm::global_error() = "device ordinal beyond the 64 this library tracks";
    return SPM_ERR_INVALID;
  }
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_dev[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        spm::global_error() = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return SPM_ERR_HIP;
      }
      g_scratch_dev[dev] = p;
    }
    ptr = g_scratch_dev[dev];
  }
  return spm::fft3d_raw(nx, ny, nz, direction, field, ptr, static_cast<hipStream_t>(stream));
}
