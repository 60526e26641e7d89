// runtime.hip — device/stream/memory entry points, the integrator, host-side
// body set-up and the state hash of libdyphur_hip.so.
#include <cstring>
#include <vector>

#include "dy_internal.hpp"

namespace dy {

// Symplectic Euler with world-frame angular velocity and quaternion
// renormalisation (core/src/integrator.cpp:17-70).  A body's substeps only
// depend on its own state, so all substeps run in one launch.
__global__ __launch_bounds__(256) void integrate_kernel(dy_body_view b, float gx, float gy, float gz, float h,
                                                        int n_substeps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= b.n) return;
    if (b.inv_mass[i] == 0.f) return;
    if (b.flags[i] & kFlagSleeping) return;
    Kin k{b.pos_x[i], b.pos_y[i], b.pos_z[i], b.rot_w[i], b.rot_x[i], b.rot_y[i], b.rot_z[i],
          b.vel_x[i], b.vel_y[i], b.vel_z[i], b.ang_x[i], b.ang_y[i], b.ang_z[i]};
    integrate_body(k, gx, gy, gz, h, n_substeps);
    b.vel_x[i] = k.vx; b.vel_y[i] = k.vy; b.vel_z[i] = k.vz;
    b.pos_x[i] = k.px; b.pos_y[i] = k.py; b.pos_z[i] = k.pz;
    b.rot_w[i] = k.qw; b.rot_x[i] = k.qx; b.rot_y[i] = k.qy; b.rot_z[i] = k.qz;
}

}  // namespace dy

using namespace dy;

extern "C" {

const char *dy_status_string(int status) {
    switch (status) {
        case DY_OK: return "success";
        case DY_ERR_INVALID_ARG: return "invalid argument or capacity exceeded";
        case DY_ERR_UNSUPPORTED: return "unsupported input (convex-hull or triangle-mesh shapes)";
        case DY_ERR_NO_DEVICE: return "no HIP device";
        case DY_ERR_INTERNAL: return "internal consistency check failed";
        default: break;
    }
    if (status > 0) return hipGetErrorString(static_cast<hipError_t>(status));
    return "unknown status";
}

int dy_version(int *major, int *minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return DY_OK;
}

int dy_device_count(int *count) {
    if (!count) return DY_ERR_INVALID_ARG;
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) { *count = 0; return DY_OK; }
    DY_TRY(e);
    return DY_OK;
}

int dy_device_get_info(int device, dy_device_info *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    int n = 0;
    DY_TRY_STATUS(dy_device_count(&n));
    if (n == 0) return DY_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return DY_ERR_INVALID_ARG;
    hipDeviceProp_t p;
    DY_TRY(hipGetDeviceProperties(&p, device));
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, p.name, sizeof(out->name) - 1);
    out->is_gpu = 1;
    out->arch_major = p.major;
    out->arch_minor = p.minor;
    out->compute_units = static_cast<size_t>(p.multiProcessorCount);
    out->global_mem_bytes = p.totalGlobalMem;
    return DY_OK;
}

int dy_set_device(int device) {
    DY_TRY(hipSetDevice(device));
    return DY_OK;
}

int dy_stream_create(dy_stream_t *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    int n = 0;
    DY_TRY_STATUS(dy_device_count(&n));
    if (n == 0) return DY_ERR_NO_DEVICE;
    hipStream_t st;
    DY_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *out = reinterpret_cast<dy_stream_t>(st);
    return DY_OK;
}

int dy_stream_destroy(dy_stream_t s) {
    if (!s) return DY_OK;
    DY_TRY(hipStreamSynchronize(hs(s)));
    DY_TRY(hipStreamDestroy(hs(s)));
    return DY_OK;
}

int dy_stream_synchronize(dy_stream_t s) {
    DY_TRY(hipStreamSynchronize(hs(s)));
    return DY_OK;
}

int dy_malloc(void **ptr, size_t bytes) {
    if (!ptr) return DY_ERR_INVALID_ARG;
    *ptr = nullptr;
    DY_TRY(hipMalloc(ptr, bytes ? bytes : 1));
    return DY_OK;
}

int dy_free(void *ptr) {
    if (!ptr) return DY_OK;
    DY_TRY(hipFree(ptr));
    return DY_OK;
}

int dy_memcpy_h2d_async(void *dst, const void *src, size_t bytes, dy_stream_t s) {
    if (!bytes) return DY_OK;
    if (!dst || !src) return DY_ERR_INVALID_ARG;
    DY_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, hs(s)));
    return DY_OK;
}

int dy_memcpy_d2h(void *dst, const void *src, size_t bytes, dy_stream_t s) {
    if (!bytes) return DY_OK;
    if (!dst || !src) return DY_ERR_INVALID_ARG;
    DY_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    return DY_OK;
}

int dy_memcpy_d2d_async(void *dst, const void *src, size_t bytes, dy_stream_t s) {
    if (!bytes) return DY_OK;
    if (!dst || !src) return DY_ERR_INVALID_ARG;
    DY_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, hs(s)));
    return DY_OK;
}

int dy_memcpy_async(void *dst, const void *src, size_t bytes, dy_stream_t s) {
    if (!bytes) return DY_OK;
    if (!dst || !src) return DY_ERR_INVALID_ARG;
    DY_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, hs(s)));
    return DY_OK;
}

int dy_memset_async(void *dst, int value, size_t bytes, dy_stream_t s) {
    if (!bytes) return DY_OK;
    if (!dst) return DY_ERR_INVALID_ARG;
    DY_TRY(hipMemsetAsync(dst, value, bytes, hs(s)));
    return DY_OK;
}

int dy_event_create(dy_event_t *out) {
    if (!out) return DY_ERR_INVALID_ARG;
    hipEvent_t e;
    DY_TRY(hipEventCreate(&e));
    *out = reinterpret_cast<dy_event_t>(e);
    return DY_OK;
}

int dy_event_destroy(dy_event_t e) {
    if (!e) return DY_OK;
    DY_TRY(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)));
    return DY_OK;
}

int dy_event_record(dy_event_t e, dy_stream_t s) {
    if (!e) return DY_ERR_INVALID_ARG;
    DY_TRY(hipEventRecord(reinterpret_cast<hipEvent_t>(e), hs(s)));
    return DY_OK;
}

int dy_event_synchronize(dy_event_t e) {
    if (!e) return DY_ERR_INVALID_ARG;
    DY_TRY(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)));
    return DY_OK;
}

int dy_event_elapsed_ms(dy_event_t a, dy_event_t b, float *ms) {
    if (!a || !b || !ms) return DY_ERR_INVALID_ARG;
    DY_TRY(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(a), reinterpret_cast<hipEvent_t>(b)));
    return DY_OK;
}

struct dy_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int dy_graph_capture_begin(dy_stream_t s) {
    if (!s) return DY_ERR_INVALID_ARG;
    DY_TRY(hipStreamBeginCapture(hs(s), hipStreamCaptureModeThreadLocal));
    return DY_OK;
}

int dy_graph_capture_end(dy_stream_t s, dy_graph_t *out) {
    if (!s || !out) return DY_ERR_INVALID_ARG;
    *out = nullptr;
    auto *g = new dy_graph_s();
    hipError_t e = hipStreamEndCapture(hs(s), &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return static_cast<int>(e);
    }
    *out = g;
    return DY_OK;
}

int dy_graph_launch(dy_graph_t g, dy_stream_t s) {
    if (!g || !g->exec) return DY_ERR_INVALID_ARG;
    DY_TRY(hipGraphLaunch(g->exec, hs(s)));
    return DY_OK;
}

int dy_graph_destroy(dy_graph_t g) {
    if (!g) return DY_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return DY_OK;
}

int dy_body_dynamics_host(uint32_t n, const float *mass, const float *I, const uint32_t *flags_in, float *inv_mass,
                          float *iI, uint32_t *flags_out) {
    if (n && (!mass || !I || !flags_in || !inv_mass || !iI || !flags_out)) return DY_ERR_INVALID_ARG;
    for (uint32_t k = 0; k < n; ++k) {
        const bool is_static = (flags_in[k] & DY_BODY_STATIC) || mass[k] <= 0.f;
        inv_mass[k] = is_static ? 0.f : 1.f / mass[k];
        float *o = iI + 6 * static_cast<size_t>(k);
        if (is_static) {
            for (int j = 0; j < 6; ++j) o[j] = 0.f;
        } else {
            // Inverse of the row-major body inertia by adjugate / determinant
            // (core/include/core/math/mat3.hpp:61-79); the symmetric entries
            // (0,0) (1,1) (2,2) (0,1) (0,2) (1,2) are kept.
            const float *m = I + 9 * static_cast<size_t>(k);
            const float det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
                              m[2] * (m[3] * m[7] - m[4] * m[6]);
            o[0] = (m[4] * m[8] - m[5] * m[7]) / det;
            o[1] = (m[0] * m[8] - m[2] * m[6]) / det;
            o[2] = (m[0] * m[4] - m[1] * m[3]) / det;
            o[3] = -(m[1] * m[8] - m[2] * m[7]) / det;
            o[4] = (m[1] * m[5] - m[2] * m[4]) / det;
            o[5] = -(m[0] * m[5] - m[2] * m[3]) / det;
        }
        flags_out[k] = is_static ? (flags_in[k] | DY_BODY_STATIC) : flags_in[k];
    }
    return DY_OK;
}

int dy_integrate(dy_stream_t s, const dy_body_view *bodies, float gx, float gy, float gz, float dt, int n_substeps) {
    if (!bodies || n_substeps <= 0) return DY_ERR_INVALID_ARG;
    if (bodies->n == 0) return DY_OK;
    const float h = dt / static_cast<float>(n_substeps);
    integrate_kernel<<<ceil_div(bodies->n, 256), 256, 0, hs(s)>>>(*bodies, gx, gy, gz, h, n_substeps);
    DY_LAUNCH_CHECK();
    return DY_OK;
}

int dy_state_hash_fnv1a64(dy_stream_t s, const dy_body_view *b, uint64_t *out) {
    if (!b || !out) return DY_ERR_INVALID_ARG;
    const size_t n = b->n;
    std::vector<float> h(7 * n);
    const float *src[7] = {b->pos_x, b->pos_y, b->pos_z, b->rot_w, b->rot_x, b->rot_y, b->rot_z};
    for (int k = 0; k < 7 && n; ++k)
        DY_TRY(hipMemcpyAsync(h.data() + k * n, src[k], n * sizeof(float), hipMemcpyDeviceToHost, hs(s)));
    DY_TRY(hipStreamSynchronize(hs(s)));
    uint64_t x = 0xcbf29ce484222325ULL;
    const unsigned char *p = reinterpret_cast<const unsigned char *>(h.data());
    for (size_t i = 0; i < 7 * n * sizeof(float); ++i) { x ^= p[i]; x *= 0x00000100000001B3ULL; }
    *out = x;
    return DY_OK;
}

}  // extern "C"
