// cmg_residual_device_noalloc: cmg_residual_device without the context's element-force
// buffers, for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  Its scratch -- the corrupt-element
// flag, the per-element forces [E][3 npe] and the per-surface pressure forces [S][3 nps] --
// lives in one file-scope device array of fixed size; every call on a device shares that
// device's copy, so calls that overlap on one device (from any context) race.
#include <climits>
#include <cstring>
#include <mutex>

#include "cmg_ctx.h"

using namespace cmg;

namespace {

// 8 Mi doubles = 64 MiB: the flag area plus the element forces of up to ~279k T10
// elements (or ~699k T4) with their surface forces
constexpr int64_t kScratchDoubles = 8 * 1024 * 1024;
constexpr int64_t kFlagDoubles = 32;   // 256 bytes: the corrupt flag, keeps the force area aligned
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch` per device, resolved once

// device address of this device's copy of `scratch` (the current device must be `dev`)
int scratch_on(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail("cmg_residual_device_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CMG_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail("cmg_residual_device_noalloc: no device address for the scratch array", CMG_ERR_HIP);
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return CMG_SUCCESS;
}

}  // namespace

extern "C" {

int cmg_residual_device_noalloc(cmg_ctx* c, const double* x_dev, double time, int32_t include_pressure,
                                double* f_dev) {
    if (!c || !x_dev || !f_dev) return fail("cmg_residual_device_noalloc: null argument");
    const bool pressure = include_pressure != 0 && c->S > 0;
    const int64_t nfe = c->E * 3 * c->npe;
    const int64_t nfs = pressure ? c->S * 3 * c->nps : 0;
    if (kFlagDoubles + nfe + nfs > kScratchDoubles)
        return fail("cmg_residual_device_noalloc: the mesh needs " +
                    std::to_string((kFlagDoubles + nfe + nfs) * (int64_t)sizeof(double)) +
                    " bytes of scratch, more than the " + std::to_string(kScratchDoubles * sizeof(double)) +
                    " bytes this entry point has; use cmg_residual_device");
    CMG_HIP_CHECK(hipSetDevice(c->device));
    double* base = nullptr;
    CMG_TRY(scratch_on(c->device, &base));
    hipStream_t s = c->stream;
    int* corrupt = reinterpret_cast<int*>(base);
    double* fe = base + kFlagDoubles;
    double* fs = fe + nfe;
    launch_fill_int(1, INT_MAX, corrupt, s);   // a fill kernel: the array keeps the last call's flag
    ElemArgs a = elem_args(c, x_dev, time);
    a.fe = fe;
    a.corrupt = corrupt;
    launch_element_residual(c->npe, a, s);
    if (pressure) {
        SurfArgs sa = surf_args(c, x_dev, false);
        sa.fs = fs;
        sa.Ks = nullptr;
        launch_surface(sa, false, s);
    }
    CMG_HIP_CHECK(hipGetLastError());
    // the context's pinned staging word, allocated at cmg_create
    CMG_HIP_CHECK(hipMemcpyAsync(c->h_pinned, corrupt, sizeof(int), hipMemcpyDeviceToHost, s));
    CMG_HIP_CHECK(hipStreamSynchronize(s));
    int flag = INT_MAX;
    std::memcpy(&flag, c->h_pinned, sizeof(int));
    if (flag != INT_MAX) {
        set_error("element " + std::to_string(flag) + " is corrupt (det F <= 0)");
        return CMG_CORRUPT_ELEMENT;
    }
    launch_gather_vector(c->N, c->inc_ptr.as<int64_t>(), c->inc.as<int32_t>(), fe, 3 * c->npe,
                         pressure ? c->sinc_ptr.as<int64_t>() : nullptr, pressure ? c->sinc.as<int32_t>() : nullptr,
                         pressure ? fs : nullptr, 3 * c->nps, c->bcmask.as<uint8_t>(), f_dev, s);
    CMG_HIP_CHECK(hipGetLastError());
    return CMG_SUCCESS;
}

}  // extern "C"
