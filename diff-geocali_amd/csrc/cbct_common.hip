// Error state, shared checks and the small shared kernels.
#include "cbct_common.h"

#include <cmath>
#include <string>

namespace cbct {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_status(hipError_t err, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(err);
    return CBCT_EHIP;
}

int check_geometry(const float* const* g, int64_t A) {
    static const char* names[8] = {"angles", "sid", "sod", "det_u0", "det_v0",
                                   "src_x", "src_y", "src_z"};
    if (A < 1) return fail(CBCT_EINVAL, "angles must contain at least one view");
    for (int i = 0; i < 8; ++i)
        if (g[i] == nullptr) return fail(CBCT_EINVAL, std::string(names[i]) + " is a null pointer");
    return CBCT_OK;
}

int check_positive_finite(float x, const char* name) {
    if (!std::isfinite(x) || !(x > 0.0f))
        return fail(CBCT_EINVAL, std::string(name) + " must be finite and positive");
    return CBCT_OK;
}

int check_finite(float x, const char* name) {
    if (!std::isfinite(x)) return fail(CBCT_EINVAL, std::string(name) + " must be finite");
    return CBCT_OK;
}

int current_stream_device(hipStream_t stream, int* device) {
    int cur = -1;
    CBCT_HIP(hipGetDevice(&cur), "hipGetDevice");
    if (stream != nullptr) {
        hipDevice_t sd = -1;
        CBCT_HIP(hipStreamGetDevice(stream, &sd), "hipStreamGetDevice");
        if ((int)sd != cur)
            return fail(CBCT_EINVAL, "the stream belongs to device " + std::to_string((int)sd) + " but device " +
                                         std::to_string(cur) + " is current; make the stream's device current");
    }
    *device = cur;
    return CBCT_OK;
}

int check_step_resolution(float step, const FrameGuard& G) {
    if ((double)step < (2.0 * (double)G.R + 1.0) * 0x1p-22)
        return fail(CBCT_EINVAL, "step_size is below the FP32 resolution of the ray parameter over the volume "
                                 "(it must be at least 2^-22 * (volume diagonal + 1))");
    return CBCT_OK;
}

__global__ void frames_kernel(GeomPtrs g, int A, FrameGuard guard, ViewFrame* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < A) out[a] = guard_frame(g, a, guard);
}

__global__ void fill_kernel(float* __restrict__ p, int64_t n, float value) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = value;
}

int launch_frames(const GeomPtrs& g, int64_t A, ViewFrame* out, hipStream_t stream, const FrameGuard* guard) {
    FrameGuard none{};
    frames_kernel<<<(unsigned)ceil_div(A, 64), 64, 0, stream>>>(g, (int)A, guard ? *guard : none, out);
    CBCT_LAUNCHED("view-frame kernel launch");
    return CBCT_OK;
}

int launch_fill(float* p, int64_t n, float value, hipStream_t stream) {
    if (n <= 0) return CBCT_OK;
    const int64_t blocks = std::min<int64_t>(ceil_div(n, 256 * 4), 8192);
    fill_kernel<<<(unsigned)blocks, 256, 0, stream>>>(p, n, value);
    CBCT_LAUNCHED("fill kernel launch");
    return CBCT_OK;
}

}  // namespace cbct

extern "C" const char* cbct_last_error(void) { return cbct::g_last_error.c_str(); }

extern "C" const char* cbct_version(void) { return "cbct_hip 0.1.0 gfx950"; }
