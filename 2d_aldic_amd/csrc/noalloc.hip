// Allocation-free image pre-pass (aldic_prep_noalloc) for callers that cannot allocate
// device memory: the same normalisation + gradient kernels as aldic_set_images
// (funNormalizeImg.m:22-40, funImgGradient.m:53-78), on caller-owned device buffers,
// with the kernels' work memory in a file-scope device array instead of a hipMalloc.
#include <mutex>

#include "aldic_internal.h"

namespace aldic {

// Work memory of the pre-pass: 2 x 2 kStatParts per-block ROI partial sums (sum and sum of
// squares of both images, written by k_stats), then {mean_f, 1/std_f, mean_g, 1/std_g}
// (written by k_stats_final).  Every call writes all of it before reading any of it, so
// nothing carries over from the previous call.  4 kStatParts + 4 doubles = 32,800 bytes per
// device, held until the process ends.
constexpr size_t kNoAllocScratchDoubles = 4 * (size_t)kStatParts + 4;
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
// Device address of `scratch` on each device (each device holds its own copy at its own
// address); the host-side name `scratch` is not a device address.
static double *g_scratch[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Called by aldic_create with `device` current.
hipError_t noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = static_cast<double *>(p);
    return hipSuccess;
}

static double *noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[device];
}

}  // namespace aldic

using namespace aldic;

extern "C" int aldic_prep_noalloc(aldic_ctx *ctx, const void *f, const void *g, int nx, int ny,
                                  int dtype, const uint8_t *mask, const int *roi, float *fo,
                                  float *go, float *gxo, float *gyo, void *stream)
{
    if (!ctx) return ALDIC_ERR_ARG;
    if (!f || !g || !fo || !go || !gxo || !gyo || nx < 8 || ny < 8) {
        ctx->err = "prep_noalloc: bad arguments";
        return ALDIC_ERR_ARG;
    }
    if ((size_t)nx * ny > kMaxPixels || nx >= (1 << 22)) {
        ctx->err = "prep_noalloc: more than 2^29 pixels or 2^22 columns";
        return ALDIC_ERR_ARG;
    }
    if (dtype < ALDIC_U8 || dtype > ALDIC_F64) {
        ctx->err = "prep_noalloc: unknown dtype";
        return ALDIC_ERR_ARG;
    }
    int r[4];
    r[0] = roi ? (roi[0] > 1 ? roi[0] : 1) : 1;
    r[1] = roi ? (roi[1] < nx ? roi[1] : nx) : nx;
    r[2] = roi ? (roi[2] > 1 ? roi[2] : 1) : 1;
    r[3] = roi ? (roi[3] < ny ? roi[3] : ny) : ny;
    if (r[1] - r[0] < 1 || r[3] - r[2] < 1) {
        ctx->err = "prep_noalloc: empty ROI";
        return ALDIC_ERR_ARG;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) {
        ctx->err = std::string("prep_noalloc: hipSetDevice: ") + hipGetErrorString(e);
        return ALDIC_ERR_HIP;
    }
    double *acc = noalloc_scratch(ctx->device);
    if (!acc) {
        ctx->err = "prep_noalloc: work array not set up on this device";
        return ALDIC_ERR_STATE;
    }
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    const double nroi = (double)(r[1] - r[0] + 1) * (r[3] - r[2] + 1);
    e = launch_stats(f, g, dtype, nx, ny, r, acc, st);
    if (e == hipSuccess)
        e = launch_prep(f, g, dtype, nx, ny, mask, acc, nroi, acc + 4 * kStatParts, fo, go, gxo,
                        gyo, st);
    if (e != hipSuccess) {
        ctx->err = std::string("prep_noalloc: launch: ") + hipGetErrorString(e);
        return ALDIC_ERR_HIP;
    }
    return ALDIC_OK;
}
