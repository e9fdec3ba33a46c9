// Test force hooks for soil_hip_set_force_hook: the harmonic soft-sphere force
// of the built-in test consumer, written as user hooks over the device view.
//
// F = k (ri + rj - d) / d * (xj - xi) for 0 < d < ri + rj, quantized to
// q = round_half_even(F * 2^32); the row atom i receives -q and j receives +q,
// summed exactly in int64 and added to the force once per atom. Integer sums
// do not depend on order, so both hooks give the consumer's forces bit for bit:
//   scatter: one thread per CSR row, int64 atomics into a per-atom accumulator;
//   gather:  one thread per atom over its row and its incoming pairs (the
//            transpose, requested with SOIL_HIP_HOOK_TRANSPOSE), no atomics.
// Counts and buffers are read from the device view at run time, so the hooks
// also work when recorded into a batch's step graph.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "soil_hip.h"

namespace {

constexpr int kBlock = 256;
constexpr double kScale = 4294967296.0;  // 2^32

struct HookState {
    unsigned long long* acc = nullptr;  // 3 per atom (scatter)
    size_t cap = 0;
    double k = 0.0;
    uint64_t calls = 0;
};

__device__ __forceinline__ void pair_force(const double4& pi, const double4& pj, double k, double& fx, double& fy,
                                           double& fz) {
    const double dx = pj.x - pi.x;
    const double dy = pj.y - pi.y;
    const double dz = pj.z - pi.z;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double rs = pi.w + pj.w;
    if (r2 < rs * rs && r2 > 0.0) {
        const double dist = sqrt(r2);
        const double s = k * (rs - dist) / dist;
        fx = s * dx;
        fy = s * dy;
        fz = s * dz;
    } else {
        fx = 0.0;
        fy = 0.0;
        fz = 0.0;
    }
}

__global__ void __launch_bounds__(kBlock) k_scatter_rows(const soil_hip_device_view* v, unsigned long long* acc,
                                                         double k) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= v->nlocal) return;
    const double4* xr = (const double4*)v->xr;
    const double4 pi = xr[i];
    long long sx = 0, sy = 0, sz = 0;
    for (uint32_t e = v->offsets[i]; e < v->offsets[i + 1]; ++e) {
        const uint32_t j = v->indices[e];
        double fx, fy, fz;
        pair_force(pi, xr[j], k, fx, fy, fz);
        const long long qx = __double2ll_rn(fx * kScale);
        const long long qy = __double2ll_rn(fy * kScale);
        const long long qz = __double2ll_rn(fz * kScale);
        sx -= qx;
        sy -= qy;
        sz -= qz;
        atomicAdd(&acc[3 * (size_t)j], (unsigned long long)qx);
        atomicAdd(&acc[3 * (size_t)j + 1], (unsigned long long)qy);
        atomicAdd(&acc[3 * (size_t)j + 2], (unsigned long long)qz);
    }
    atomicAdd(&acc[3 * (size_t)i], (unsigned long long)sx);
    atomicAdd(&acc[3 * (size_t)i + 1], (unsigned long long)sy);
    atomicAdd(&acc[3 * (size_t)i + 2], (unsigned long long)sz);
}

__global__ void __launch_bounds__(kBlock) k_scatter_apply(const soil_hip_device_view* v, unsigned long long* acc) {
    const uint32_t a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= v->nlocal + v->nghost) return;
    const double inv = 1.0 / kScale;
    v->fx[a] = v->fx[a] + (double)(long long)acc[3 * (size_t)a] * inv;
    v->fy[a] = v->fy[a] + (double)(long long)acc[3 * (size_t)a + 1] * inv;
    v->fz[a] = v->fz[a] + (double)(long long)acc[3 * (size_t)a + 2] * inv;
    acc[3 * (size_t)a] = 0ull;  // ready for the next step (no memset node in a graph)
    acc[3 * (size_t)a + 1] = 0ull;
    acc[3 * (size_t)a + 2] = 0ull;
}

__global__ void __launch_bounds__(kBlock) k_gather(const soil_hip_device_view* v, double k) {
    const uint32_t a = blockIdx.x * kBlock + threadIdx.x;
    if (a >= v->nlocal + v->nghost || v->tstart == nullptr) return;  // no transpose: nothing to read
    const double4* xr = (const double4*)v->xr;
    const double4 pa = xr[a];
    long long sx = 0, sy = 0, sz = 0;
    if (a < v->nlocal) {
        for (uint32_t e = v->offsets[a]; e < v->offsets[a + 1]; ++e) {
            double fx, fy, fz;
            pair_force(pa, xr[v->indices[e]], k, fx, fy, fz);
            sx -= __double2ll_rn(fx * kScale);
            sy -= __double2ll_rn(fy * kScale);
            sz -= __double2ll_rn(fz * kScale);
        }
    }
    for (uint32_t e = v->tstart[a]; e < v->tstart[a + 1]; ++e) {
        double fx, fy, fz;
        pair_force(xr[v->tsrc[e]], pa, k, fx, fy, fz);
        sx += __double2ll_rn(fx * kScale);
        sy += __double2ll_rn(fy * kScale);
        sz += __double2ll_rn(fz * kScale);
    }
    const double inv = 1.0 / kScale;
    v->fx[a] = v->fx[a] + (double)sx * inv;
    v->fy[a] = v->fy[a] + (double)sy * inv;
    v->fz[a] = v->fz[a] + (double)sz * inv;
}

unsigned blocks(uint32_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

extern "C" {

// State of one hooked box: an accumulator for up to `cap_atoms` atoms.
void* soil_test_hook_create(size_t cap_atoms, double k) {
    HookState* s = new HookState();
    s->cap = cap_atoms;
    s->k = k;
    if (hipMalloc((void**)&s->acc, 3 * cap_atoms * sizeof(unsigned long long) + 8) != hipSuccess ||
        hipMemset(s->acc, 0, 3 * cap_atoms * sizeof(unsigned long long) + 8) != hipSuccess) {
        delete s;
        return nullptr;
    }
    return s;
}

void soil_test_hook_destroy(void* user) {
    HookState* s = (HookState*)user;
    if (!s) return;
    hipFree(s->acc);
    delete s;
}

uint64_t soil_test_hook_calls(void* user) { return ((HookState*)user)->calls; }

int soil_test_hook_scatter(void* user, const soil_hip_device_view* view, const soil_hip_device_view* d_view,
                           void* stream) {
    HookState* s = (HookState*)user;
    if (!s || view->cap_atoms > s->cap) return 1;  // accumulator too small for this box
    s->calls += 1;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_scatter_rows, dim3(blocks(view->cap_local)), dim3(kBlock), 0, st, d_view, s->acc, s->k);
    hipLaunchKernelGGL(k_scatter_apply, dim3(blocks(view->cap_atoms)), dim3(kBlock), 0, st, d_view, s->acc);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

int soil_test_hook_gather(void* user, const soil_hip_device_view* view, const soil_hip_device_view* d_view,
                          void* stream) {
    HookState* s = (HookState*)user;
    if (!s) return 1;
    s->calls += 1;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_gather, dim3(blocks(view->cap_atoms)), dim3(kBlock), 0, st, d_view, s->k);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
