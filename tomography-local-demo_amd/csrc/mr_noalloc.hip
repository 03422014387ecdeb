// Entry points that allocate nothing: they work from a file-scope device array (one per
// GPU, held until the process ends) instead of the context's growable buffers, for
// callers that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// In host code the array's name is only a stand-in for its device address: the address
// of each GPU's copy comes from hipGetSymbolAddress, taken once per device (with that
// device current) when the first context on it is created, and every host use of the
// scratch (kernel arguments, copies) goes through that pointer.
#include <cmath>
#include <cstring>
#include <mutex>

#include "mr_internal.h"

namespace mr {

// The scratch of the entry points below: one copy per GPU.
static __device__ __align__(256) unsigned char scratch[MR_STATIC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu[kMaxDevices];          // one call at a time on each GPU's copy
void* g_scratch_ptr[kMaxDevices] = {};
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices] = {};

__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// Layout of the scratch for the clipped histogram: [bins] u64 counters, then the
// [bins + 1] float64 edges at the next 256-byte boundary.
constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }
constexpr size_t hist_scratch_bytes(int bins) {
    return align256(sizeof(unsigned long long) * bins) + sizeof(double) * (bins + 1);
}
static_assert(hist_scratch_bytes(MR_STATIC_HIST_MAX_BINS) <= MR_STATIC_SCRATCH_BYTES,
              "the static scratch holds the largest histogram");

}  // namespace

// Takes the device address of this GPU's scratch (the device must be current); once
// per device, from the context creation on that device.
int static_scratch_init(int device) {
    MR_CHECK(device >= 0 && device < kMaxDevices, MR_EVALUE, "device ordinal out of range");
    std::call_once(g_scratch_once[device], [device] {
        void* p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (g_scratch_err[device] == hipSuccess) g_scratch_ptr[device] = p;
    });
    if (g_scratch_err[device] != hipSuccess)
        return fail(MR_ERUNTIME, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(g_scratch_err[device]));
    MR_CHECK(g_scratch_ptr[device] != nullptr, MR_ERUNTIME, "static scratch has no address");
    return MR_OK;
}

}  // namespace mr

using namespace mr;

struct mr_ctx : Ctx {};

extern "C" int mr_clip_histogram_static(mr_ctx* c, const float* vol_device, int64_t n, double lo,
                                        double hi, int bins, const double* edges, int64_t* hist) {
    if (!c) return fail(MR_EVALUE, "null context");
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(n >= 0 && (n == 0 || vol_device), MR_EVALUE, "null volume");
    MR_CHECK(edges && hist, MR_EVALUE, "null edges or histogram");
    MR_CHECK(bins >= 1 && bins <= MR_STATIC_HIST_MAX_BINS, MR_EVALUE,
             "bins = " + std::to_string(bins) + " outside 1.." +
                 std::to_string(MR_STATIC_HIST_MAX_BINS) +
                 " (the static scratch of mr_clip_histogram_static holds " +
                 std::to_string(MR_STATIC_SCRATCH_BYTES) + " bytes)");
    MR_CHECK(std::isfinite(lo) && std::isfinite(hi) && lo < hi, MR_EVALUE,
             "range must be finite with lo < hi");
    const int dev = c->device;
    MR_HIP(hipSetDevice(dev));
    const int rs = static_scratch_init(dev);      // taken at context creation: its result
    if (rs) return rs;
    std::lock_guard<std::mutex> sk(g_scratch_mu[dev]);
    char* base = static_cast<char*>(g_scratch_ptr[dev]);
    auto* d_hist = reinterpret_cast<unsigned long long*>(base);
    auto* d_edges = reinterpret_cast<double*>(base + align256(sizeof(unsigned long long) * bins));
    // the counters keep the previous call's contents: clear them with a kernel
    hipLaunchKernelGGL(zero_u64_kernel, dim3((bins + 255) / 256), dim3(256), 0, c->s_req, d_hist,
                       bins);
    MR_HIP(hipGetLastError());
    MR_HIP(hipMemcpyAsync(d_edges, edges, sizeof(double) * (bins + 1), hipMemcpyHostToDevice,
                          c->s_req));
    if (n > 0) {
        const int rc = launch_clip_hist(vol_device, n, lo, hi, bins, d_edges, d_hist, c->s_req);
        if (rc) return rc;
    }
    // unsigned 64-bit counts straight into the caller's int64 array (same bits below 2^63)
    static_assert(sizeof(int64_t) == sizeof(unsigned long long), "count width");
    MR_HIP(hipMemcpyAsync(hist, d_hist, sizeof(unsigned long long) * bins, hipMemcpyDeviceToHost,
                          c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    return MR_OK;
}
