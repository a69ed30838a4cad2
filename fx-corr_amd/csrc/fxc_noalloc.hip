// Allocation-free sector entry point (no counterpart in the reference).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the
// sector path of fxc_sector_run_device with its scratch -- the geometric
// table, the workgroup partial slots and the reduction's chunk sums -- in a
// file-scope device array instead of plan-owned buffers.  The kernels are
// the same; the array is reached through the address hipGetSymbolAddress
// returns for the current device (the host-side name of a __device__ array
// is only a stand-in, not a device address).
#include <hip/hip_runtime.h>

#include <mutex>

#include "fxc_internal.h"
#include "fxcorr.h"

extern "C" fxc_status fxc_internal_fail(fxc_status s, const char *msg);

namespace {

constexpr size_t kScratchBytes = FXC_NOALLOC_SCRATCH_BYTES;   // per device, held until the process ends
constexpr size_t kTableBytes = (size_t)FXC_NOALLOC_MAX_TABLE * sizeof(double);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch_dev[kMaxDevices];   // device address of `scratch`, per device

// Device address of the scratch array on the current device, looked up once
// per device (with that device current) on the first call there.
fxc_status scratch_for(int dev, void **out) {
    if (dev < 0 || dev >= kMaxDevices)
        return fxc_internal_fail(FXC_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fxc_internal_fail(FXC_ERR_HIP, "hipGetSymbolAddress of the allocation-free scratch failed");
        g_scratch_dev[dev] = ptr;
    }
    *out = g_scratch_dev[dev];
    return FXC_OK;
}

}  // namespace

extern "C" fxc_status fxc_internal_sector_noalloc(fxc_plan *p, const uint8_t *d_raw1,
                                                  const uint8_t *d_raw2, int64_t first_frame,
                                                  int64_t nf, const fxc_delay_cfg *delay,
                                                  double *d_acc, void *stream, uint8_t *scratch_dev,
                                                  size_t scratch_bytes, size_t table_bytes,
                                                  int device);

extern "C" fxc_status fxc_sector_run_device_noalloc(fxc_plan *plan, const uint8_t *d_raw1,
                                                    const uint8_t *d_raw2, int64_t first_frame,
                                                    int64_t nf, const fxc_delay_cfg *delay,
                                                    double *d_acc, void *stream) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return fxc_internal_fail(FXC_ERR_HIP, "hipGetDevice failed");
    void *ptr = nullptr;
    fxc_status st = scratch_for(dev, &ptr);
    if (st != FXC_OK) return st;
    return fxc_internal_sector_noalloc(plan, d_raw1, d_raw2, first_frame, nf, delay, d_acc, stream,
                                       (uint8_t *)ptr, kScratchBytes, kTableBytes, dev);
}
