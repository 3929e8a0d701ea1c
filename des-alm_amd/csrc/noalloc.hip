// noalloc.hip -- desalm_gtsv_batch_noalloc: the batched SGTSV of
// desalm_gtsv_batch (same kernel, same results) whose device staging of the
// caller's host arrays lives in one fixed file-scope device array instead of
// per-call hipMalloc / hipFree, for callers that may not allocate device
// memory.
//
// The array is a host-side stand-in in host code: kernels and copies are
// given its device address, which hipGetSymbolAddress returns once per device
// (with that device current) on the first call there and which is kept per
// device.  Every call on one GPU shares that GPU's array; calls that overlap on
// one GPU race.  Nothing in it has to start at zero: every call copies its
// four arrays in and the kernel writes every INFO it reads back.
#include <mutex>

#include "desalm_internal.h"

namespace {

constexpr long kMaxEntries = DESALM_GTSV_NOALLOC_MAX;  // n * nsys
// dl, d, du, b (kMaxEntries doubles each), then kMaxEntries INFO ints
constexpr size_t kScratchBytes = 4 * (size_t)kMaxEntries * sizeof(double) + (size_t)kMaxEntries * sizeof(int);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

// Device address of the scratch array on the current device (dev), looked up
// once per device.
int scratch_ptr(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return desalm::host_fail(std::string("desalm_gtsv_batch_noalloc: hipGetSymbolAddress failed: ") +
                                     hipGetErrorString(e));
        g_ptr[dev] = (unsigned char*)p;
    }
    *out = g_ptr[dev];
    return 0;
}

}  // namespace

extern "C" int desalm_gtsv_batch_noalloc(int n, int nsys, double* dl, double* d, double* du, double* b,
                                         int* info_out) {
    using desalm::host_fail;
    if (n < 0 || nsys < 0) return host_fail("desalm_gtsv_batch_noalloc: negative size");
    if (n == 0 || nsys == 0) return 0;
    if (!dl || !d || !du || !b) return host_fail("desalm_gtsv_batch_noalloc: null argument");
    const long entries = (long)n * (long)nsys;
    if (entries > kMaxEntries)
        return host_fail("desalm_gtsv_batch_noalloc: n * nsys = " + std::to_string(entries) +
                         " exceeds the fixed scratch of " + std::to_string(kMaxEntries) +
                         " entries (DESALM_GTSV_NOALLOC_MAX); use desalm_gtsv_batch");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return host_fail("desalm_gtsv_batch_noalloc: no current HIP device");
    unsigned char* base = nullptr;
    if (scratch_ptr(dev, &base)) return -1;
    double* ddl = (double*)base;
    double* dd = ddl + kMaxEntries;
    double* ddu = dd + kMaxEntries;
    double* db = ddu + kMaxEntries;
    int* dinfo = (int*)(db + kMaxEntries);
    const size_t bytes = sizeof(double) * (size_t)entries;
    if (hipMemcpy(ddl, dl, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dd, d, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ddu, du, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) != hipSuccess)
        return host_fail("desalm_gtsv_batch_noalloc: host-to-device copy failed");
    desalm::launch_gtsv_batch(0, n, nsys, ddl, dd, ddu, db, dinfo);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return host_fail("desalm_gtsv_batch_noalloc: kernel failed");
    if (hipMemcpy(dl, ddl, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(d, dd, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(du, ddu, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(b, db, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
        (info_out && hipMemcpy(info_out, dinfo, sizeof(int) * (size_t)nsys, hipMemcpyDeviceToHost) != hipSuccess))
        return host_fail("desalm_gtsv_batch_noalloc: device-to-host copy failed");
    return 0;
}
