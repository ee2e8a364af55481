// bbk_noalloc.hip -- bbk_dsytrf_noalloc: the binary64 factorization with
// no device allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The factorization's scratch (exchange records, panel state, interchange
// list, flushed W rows) lives in the file-scope device array `scratch`,
// sized for orders up to kNoAllocMaxN.  Each GPU holds its own copy until
// the process ends.  Host code never passes the array's name to a kernel:
// the first call on a device (with that device current) resolves the
// array's device address with hipGetSymbolAddress and keeps it per device;
// every kernel argument and fill uses that address.
//
// Every call on one GPU shares that GPU's array and its tag counter, so
// calls that overlap on one GPU (two streams, two threads) race and must be
// ordered by the caller.  The array keeps the last call's contents; the
// exchange records are cleared on a device's first call and whenever the
// tag counter wraps, and every launch uses tags no earlier launch used.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/bbk_ldlt.h"
#include "bbk_handle.hpp"

namespace bbk {

constexpr int kNoAllocMaxN = 16384;
constexpr int kNoAllocMaxDevices = 64;

// factor_scratch_bytes<double>(kNoAllocMaxN) by the same carve: records for
// 2 (64 + 16) exchange slots, the panel state, 2 (128 + 1) interchange
// entries and (128 + 1) W rows of 16384 binary64 values, each region
// rounded up to 256 bytes.
constexpr size_t kRegion(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNoAllocScratchBytes =
    kRegion(sizeof(XRec) * 2 * ((kNoAllocMaxN + kPanelThreads - 1) / kPanelThreads + 16)) +
    kRegion(sizeof(int) * PanelState::kInts) + kRegion(sizeof(int) * 2 * (kPanelMaxWidth + 1)) +
    kRegion(sizeof(double) * (kPanelMaxWidth + 1) * static_cast<size_t>(kNoAllocMaxN));

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocDevice {
  void* ptr = nullptr;  // device address of `scratch` on this device
  int epoch = 1024;     // next exchange-tag range
  bool cleared = false;
  long long clock_khz = 0;  // device wall-clock rate
};

std::mutex g_noalloc_mu;
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

}  // namespace

}  // namespace bbk

extern "C" int bbk_dsytrf_noalloc(int n, int nb, double* d_A, int64_t lda, int* d_perm,
                                  int* d_piv, void* stream) {
  using namespace bbk;
  if (n <= 0) return -1;
  if (nb < 0) return -2;
  if (!d_A) return -3;
  if (lda < n) return -4;
  if (!d_perm) return -5;
  if (!d_piv) return -6;
  if (!stream) return -7;  // cooperative launches are refused on the null stream
  if (n > kNoAllocMaxN) return BBK_ERR_UNSUPPORTED;
  const int b = nb == 0 ? default_panel_width(n, sizeof(double)) : nb;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return BBK_ERR_HIP;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return BBK_ERR_UNSUPPORTED;
  // n <= 16384: the 256-column panel variant, which the scratch is sized for
  int nt = 0, max_groups = 0;
  const int pc = panel_plan<double>(n, kPanelThreads, &nt, &max_groups, &e);
  if (pc != BBK_OK) return pc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);

  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoAllocDevice& d = g_noalloc[dev];
  if (!d.ptr) {
    // once per device, with the device current: the array's address here
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return BBK_ERR_HIP;
    d.ptr = p;
  }
  if (factor_scratch_bytes<double>(kNoAllocMaxN, nullptr, nullptr) > kNoAllocScratchBytes)
    return BBK_ERR_UNSUPPORTED;  // carve and array size disagree (build error)
  FactorScratch fs;
  factor_scratch_bytes<double>(kNoAllocMaxN, &fs, reinterpret_cast<char*>(d.ptr));
  if (!d.cleared || d.epoch > (1 << 30)) {
    launch_zero_u32(reinterpret_cast<unsigned*>(fs.xrec),
                    static_cast<int>(fs.xrec_bytes / sizeof(unsigned)), st);
    d.cleared = true;
    d.epoch = 1024;
  }
  // no handle, so no status words: a lost exchange record ends the waits
  // at the default time limit without an error report
  if (d.clock_khz == 0) {
    int khz = 0;
    d.clock_khz = (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) ==
                       hipSuccess &&
                   khz > 0)
                      ? khz
                      : 100000;
  }
  return factor_loop<double>(n, b, d_A, lda, d_perm, d_piv, fs, nt, max_groups, &d.epoch, st,
                             nullptr, nullptr, static_cast<long long>(kDefaultWaitMs) * d.clock_khz);
}
