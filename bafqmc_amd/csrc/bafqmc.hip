// HIP back end: gfx950 kernels and the runtime hooks of the C ABI.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

#include "api.hpp"

namespace bafq {

static thread_local std::string g_err;

static int check(hipError_t e, const char* what) {
    if (e == hipSuccess) return 0;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return 1;
}

__global__ void __launch_bounds__(512) k_prepare(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem};
    body_prepare<false>(c, a);
}

__global__ void __launch_bounds__(512) k_sweep(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem};
    body_sweep<false>(c, a);
}

__global__ void __launch_bounds__(512) k_prepare_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem};
    body_prepare<true>(c, a);
}

__global__ void __launch_bounds__(512) k_sweep_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem};
    body_sweep<true>(c, a);
}

__global__ void __launch_bounds__(64) k_warmup(LaunchArgs a) {
    Ctx c{nullptr};
    body_warmup(c, a);
}

__global__ void __launch_bounds__(512) k_debug_gemm(cd* C, const cd* A, int oa, const cd* B, int ob, int n) {
    Ctx c{nullptr};
    gemm_cc(c, C, A, oa != 0, B, ob != 0, n);
}

// variant 0: workgroup-cooperative QR of A in global memory; 1: column-parallel
// QR of A copied into LDS (the LDS-path routine; needs 4 n <= threads).
__global__ void __launch_bounds__(512) k_debug_qrp(cd* A, int n, cd* tau, int* perm, double* d, int variant) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem};
    const Lds L = carve(smem, n, c.nwaves(), variant == 1);
    if (variant == 1) {
        for (int e = c.tid(); e < n * n; e += c.nthreads()) L.mat[e] = A[e];
        c.sync();
        qrp_cols(c, L.mat, n, L.qr, L.dtmp);
        for (int e = c.tid(); e < n * n; e += c.nthreads()) A[e] = L.mat[e];
    } else {
        qrp(c, A, n, L.qr, L.dtmp);
    }
    for (int k = c.tid(); k < n; k += c.nthreads()) {
        tau[k] = L.qr.tau[k];
        perm[k] = L.qr.perm[k];
        d[k] = L.dtmp[k];
    }
}

// Lane-layout probe of the FP64 MFMA: workgroup p (one wavefront) multiplies
// A = one-hot at lane p by B = (1000 + lane) and writes the 4 accumulators.
__global__ void __launch_bounds__(64) k_debug_mfma(double* out) {
    Ctx c{nullptr};
    const int p = blockIdx.x, l = threadIdx.x;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = c.mfma(l == p ? 1.0 : 0.0, 1000.0 + l, acc);
    for (int r = 0; r < 4; ++r) out[(p * 64 + l) * 4 + r] = acc[r];
}

// Cross-lane operations of one wavefront: out[0:64] row16_sum, [64:128] quad_sum,
// [128:192] lane_bcast(., 37), [192:256] / [256:320] wave_argmax value / index.
__global__ void __launch_bounds__(64) k_debug_lanes(const double* in, const int* idx, double* out) {
    Ctx c{nullptr};
    const int l = threadIdx.x;
    const double v = in[l];
    out[l] = c.row16_sum(v);
    out[64 + l] = c.quad_sum(v);
    out[128 + l] = c.lane_bcast(v, 37);
    double bv = v;
    int bi = idx[l];
    c.wave_argmax(bv, bi);
    out[192 + l] = bv;
    out[256 + l] = (double)bi;
}

int rt_debug_lanes(const double* in, const int32_t* idx, double* out) {
    double* d = nullptr;
    if (rt_malloc((void**)&d, sizeof(double) * (64 + 320) + sizeof(int) * 64)) return 1;
    double* din = d + 320;
    int* didx = (int*)(d + 384);
    int rc = rt_h2d(din, in, sizeof(double) * 64) || rt_h2d(didx, idx, sizeof(int) * 64);
    if (!rc) {
        hipLaunchKernelGGL(k_debug_lanes, dim3(1), dim3(64), 0, 0, din, didx, d);
        rc = check(hipGetLastError(), "k_debug_lanes") || rt_sync() || rt_d2h(out, d, sizeof(double) * 320);
    }
    rt_free(d);
    return rc;
}

int rt_debug_mfma(double* out) {
    double* d = nullptr;
    if (rt_malloc((void**)&d, sizeof(double) * 64 * 64 * 4)) return 1;
    hipLaunchKernelGGL(k_debug_mfma, dim3(64), dim3(64), 0, 0, d);
    int rc = check(hipGetLastError(), "k_debug_mfma") || rt_sync() || rt_d2h(out, d, sizeof(double) * 64 * 64 * 4);
    rt_free(d);
    return rc;
}

int rt_malloc(void** p, size_t bytes) { return check(hipMalloc(p, bytes), "hipMalloc"); }
void rt_free(void* p) { (void)hipFree(p); }
int rt_h2d(void* dst, const void* src, size_t bytes) {
    return check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
int rt_d2h(void* dst, const void* src, size_t bytes) {
    return check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
int rt_memset(void* dst, int v, size_t bytes) { return check(hipMemset(dst, v, bytes), "hipMemset"); }
int rt_sync() { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
const char* rt_error() { return g_err.c_str(); }

int rt_block_threads() {
    const char* s = std::getenv("BAFQ_BLOCK");
    int b = s ? std::atoi(s) : 256;
    if (b < 64 || b > 512 || (b % 64) != 0) b = 256;
    return b;
}

double rt_clock_hz() {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, 0) != hipSuccess || khz <= 0) return 1.0e8;
    return 1000.0 * khz;
}

int rt_lds_block_threads() { return kLdsWaves * 64; }

size_t rt_lds_limit() {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess || v <= 0) return 65536;
    return (size_t)v;
}

int rt_launch(KernelId k, PathId path, const LaunchArgs& a, int grid, int block, size_t smem, double* ms) {
    static hipEvent_t ev[2] = {nullptr, nullptr};
    if (!ev[0]) {
        if (check(hipEventCreate(&ev[0]), "hipEventCreate") || check(hipEventCreate(&ev[1]), "hipEventCreate"))
            return 1;
    }
    const void* fn = nullptr;
    if (k == K_PREPARE) fn = path == P_LDS ? (const void*)k_prepare_lds : (const void*)k_prepare;
    if (k == K_SWEEP) fn = path == P_LDS ? (const void*)k_sweep_lds : (const void*)k_sweep;
    if (fn && smem > 65536 &&
        check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem), "hipFuncSetAttribute"))
        return 1;
    if (check(hipEventRecord(ev[0], 0), "hipEventRecord")) return 1;
    if (k == K_WARMUP) {
        hipLaunchKernelGGL(k_warmup, dim3(grid), dim3(block), 0, 0, a);
    } else if (k == K_PREPARE) {
        if (path == P_LDS) hipLaunchKernelGGL(k_prepare_lds, dim3(grid), dim3(block), smem, 0, a);
        else hipLaunchKernelGGL(k_prepare, dim3(grid), dim3(block), smem, 0, a);
    } else {
        if (path == P_LDS) hipLaunchKernelGGL(k_sweep_lds, dim3(grid), dim3(block), smem, 0, a);
        else hipLaunchKernelGGL(k_sweep, dim3(grid), dim3(block), smem, 0, a);
    }
    if (check(hipGetLastError(), "kernel launch")) return 1;
    if (check(hipEventRecord(ev[1], 0), "hipEventRecord") || check(hipEventSynchronize(ev[1]), "hipEventSynchronize"))
        return 1;
    float f = 0.0f;
    if (check(hipEventElapsedTime(&f, ev[0], ev[1]), "hipEventElapsedTime")) return 1;
    if (ms) *ms = (double)f;
    return 0;
}

int rt_debug_gemm(int n, const double* a, int op_a, const double* b, int op_b, double* c) {
    if (n < 1 || n > 512) { g_err = "n out of range"; return 1; }
    const size_t bytes = sizeof(cd) * (size_t)n * n;
    cd *da = nullptr, *db = nullptr, *dc = nullptr;
    if (rt_malloc((void**)&da, bytes) || rt_malloc((void**)&db, bytes) || rt_malloc((void**)&dc, bytes)) return 1;
    int rc = rt_h2d(da, a, bytes) || rt_h2d(db, b, bytes);
    if (!rc) {
        hipLaunchKernelGGL(k_debug_gemm, dim3(1), dim3(rt_block_threads()), 0, 0, dc, da, op_a, db, op_b, n);
        rc = check(hipGetLastError(), "k_debug_gemm") || rt_sync() || rt_d2h(c, dc, bytes);
    }
    rt_free(da); rt_free(db); rt_free(dc);
    return rc;
}

int rt_debug_qrp(int n, double* a, double* tau, int32_t* perm, double* d, int variant) {
    if (n < 1 || n > 512) { g_err = "n out of range"; return 1; }
    if (variant == 1 && (n > kColMaxN || 4 * n > rt_lds_block_threads())) { g_err = "n too large for the column QR"; return 1; }
    const size_t bytes = sizeof(cd) * (size_t)n * n;
    cd *da = nullptr, *dt = nullptr;
    int* dp = nullptr;
    double* dd = nullptr;
    if (rt_malloc((void**)&da, bytes) || rt_malloc((void**)&dt, sizeof(cd) * n) ||
        rt_malloc((void**)&dp, sizeof(int) * n) || rt_malloc((void**)&dd, sizeof(double) * n)) return 1;
    int rc = rt_h2d(da, a, bytes);
    if (!rc) {
        const int block = variant == 1 ? rt_lds_block_threads() : rt_block_threads();
        const size_t lb = lds_bytes(n, block / 64, variant == 1);
        if (lb > rt_lds_limit()) g_err = "matrix does not fit in LDS";
        rc = lb > rt_lds_limit() || (lb > 65536 && check(hipFuncSetAttribute((const void*)k_debug_qrp,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb),
                                  "hipFuncSetAttribute"));
        if (!rc) {
            hipLaunchKernelGGL(k_debug_qrp, dim3(1), dim3(block), lb, 0, da, n, dt, dp, dd, variant);
            rc = check(hipGetLastError(), "k_debug_qrp") || rt_sync() || rt_d2h(a, da, bytes) ||
                 rt_d2h(tau, dt, sizeof(cd) * n) || rt_d2h(perm, dp, sizeof(int) * n) ||
                 rt_d2h(d, dd, sizeof(double) * n);
        }
    }
    rt_free(da); rt_free(dt); rt_free(dp); rt_free(dd);
    return rc;
}

}  // namespace bafq
