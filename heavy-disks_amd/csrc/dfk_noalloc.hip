// Scratch of the no-allocation entry point (dfk_potential_noalloc): one
// disk's half spectrum and a zeroed control block, in a file-scope device
// array instead of memory the library allocates.  Each device holds its own
// copy of the array, at its own address, until the process ends.
//
// In host code `scratch` names a host-side stand-in, not the device memory:
// its device address is looked up once per device, with that device current,
// by dfk_noalloc_init (called from dfk_create), and only that address is
// ever handed to kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>

#include "dfk_internal.h"
#include "dfk_launch.h"

namespace {

// [0, 256): the DiskCtl the transform kernels read (all zero: "not done",
// never written by them); [256, ...): the half spectrum, DFK_NOALLOC_MAX_SPECTRUM
// double2 entries.
constexpr size_t kCtlBytes = 256;
static_assert(sizeof(DiskCtl) <= kCtlBytes, "control block larger than its scratch slot");
constexpr size_t kScratchBytes = kCtlBytes + (size_t)DFK_NOALLOC_MAX_SPECTRUM * sizeof(double2);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

int dfk_noalloc_init(int device, char *err, size_t errlen) {
  if (device < 0 || device >= kMaxDevices) {
    snprintf(err, errlen, "no-allocation scratch: device %d beyond %d", device, kMaxDevices);
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device]) return 0;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    snprintf(err, errlen, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
             hipGetErrorString(e));
    return 1;
  }
  // the control block must read "not done"; device globals start zeroed,
  // cleared here once more through the looked-up address
  e = hipMemset(p, 0, kCtlBytes);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    snprintf(err, errlen, "clearing the no-allocation scratch on device %d failed: %s", device,
             hipGetErrorString(e));
    return 1;
  }
  g_ptr[device] = p;
  return 0;
}

void *dfk_noalloc_ctl(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  return g_ptr[device];
}

void *dfk_noalloc_spectrum(int device) {
  void *p = dfk_noalloc_ctl(device);
  return p ? (void *)((unsigned char *)p + kCtlBytes) : nullptr;
}
