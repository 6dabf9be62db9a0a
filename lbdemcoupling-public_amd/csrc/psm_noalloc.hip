/*
 * Allocation-free form of lbdem_psm_subcell_counts, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).
 *
 * lbdem_psm_subcell_counts stages its queries and results in five device
 * buffers it allocates and frees on every call.  This form stages them in
 * `scratch`, a file-scope device array of fixed size (LBDEM_PSM_STATIC_MAX_QUERIES
 * queries, 48 bytes each), which each GPU the library runs on holds until
 * the process ends.  The host reaches the array through the device address
 * hipGetSymbolAddress gives for it, looked up once per device with that
 * device current and kept per device: the array's name in host code is a
 * host-side stand-in, not a device address.
 *
 * Every call on one GPU shares that GPU's array and leaves the last call's
 * data in it; calls that overlap on one GPU (two host threads, or two
 * handles' streams) race.  Every byte a call reads from the array is written
 * by that call first, so nothing needs clearing.
 */
#include "psm_common.h"

#include <mutex>

namespace psm {
constexpr int64_t STATIC_MAX = LBDEM_PSM_STATIC_MAX_QUERIES;
/* per query: offset (3 doubles), radius, solid fraction, count, branch */
constexpr size_t OFF_D3 = 0;
constexpr size_t OFF_R = OFF_D3 + 24 * (size_t)STATIC_MAX;
constexpr size_t OFF_SF = OFF_R + 8 * (size_t)STATIC_MAX;
constexpr size_t OFF_CNT = OFF_SF + 8 * (size_t)STATIC_MAX;
constexpr size_t OFF_BR = OFF_CNT + 4 * (size_t)STATIC_MAX;
constexpr size_t STATIC_BYTES = OFF_BR + 4 * (size_t)STATIC_MAX;
static_assert(OFF_R % 256 == 0 && OFF_SF % 256 == 0 && OFF_CNT % 256 == 0 && OFF_BR % 256 == 0,
              "scratch regions stay 256-byte aligned");
}  // namespace psm

static __device__ __align__(256) unsigned char scratch[psm::STATIC_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[MAX_DEVICES] = {};

/* The device address of `scratch` on the handle's device, looked up on the
 * first call on that device (with the device current) and kept. */
int scratch_for(lbdem_psm* h, unsigned char** out) {
    const int dev = h->desc.device;
    if (dev < 0 || dev >= MAX_DEVICES) return psm::set_err(h, LBDEM_PSM_ERR_NO_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        PSM_HIP(h, hipGetDevice(&cur));
        if (cur != dev) PSM_HIP(h, hipSetDevice(dev));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess) return psm::hip_check(h, e, "hipGetSymbolAddress(scratch)");
        if (!p) return psm::set_err(h, LBDEM_PSM_ERR_HIP, "no device address for the scratch array");
        g_scratch_ptr[dev] = (unsigned char*)p;
    }
    *out = g_scratch_ptr[dev];
    return LBDEM_PSM_OK;
}
}  // namespace

extern "C" {

int lbdem_psm_subcell_counts_static(lbdem_psm* h, int64_t n, const double* d3, const double* r, double* sf,
                                    int32_t* count, int32_t* branch) {
    if (!h || n < 0 || (n > 0 && (!d3 || !r || !sf || !count || !branch))) return LBDEM_PSM_ERR_INVALID_ARG;
    if (n > psm::STATIC_MAX)
        return psm::set_err(h, LBDEM_PSM_ERR_CAPACITY,
                            "subcell_counts_static: more than LBDEM_PSM_STATIC_MAX_QUERIES queries");
    if (n == 0) return LBDEM_PSM_OK;
    unsigned char* base = nullptr;
    int rc = scratch_for(h, &base);
    if (rc) return rc;
    double* dd = (double*)(base + psm::OFF_D3);
    double* dr = (double*)(base + psm::OFF_R);
    double* dsf = (double*)(base + psm::OFF_SF);
    int32_t* dc = (int32_t*)(base + psm::OFF_CNT);
    int32_t* db = (int32_t*)(base + psm::OFF_BR);
    PSM_HIP(h, hipMemcpyAsync(dd, d3, 3 * n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    PSM_HIP(h, hipMemcpyAsync(dr, r, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    rc = psm::launch_subcell(h, n, dd, dr, dsf, dc, db);
    if (rc) return rc;
    PSM_HIP(h, hipMemcpyAsync(sf, dsf, n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    PSM_HIP(h, hipMemcpyAsync(count, dc, n * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    PSM_HIP(h, hipMemcpyAsync(branch, db, n * sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
    PSM_HIP(h, hipStreamSynchronize(h->stream));
    return LBDEM_PSM_OK;
}

}  // extern "C"
