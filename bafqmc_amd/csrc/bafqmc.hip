// HIP back end: gfx950 kernels and the runtime hooks of the C ABI.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
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
    Ctx c{smem + kSlotBytes};
    body_prepare<false>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_sweep(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_sweep<false>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_prepare_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_prepare<true>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_sweep_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_sweep<true>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_grad(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_grad<false>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_grad_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_grad<true>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_measure(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_measure<false>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_measure_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_measure<true>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_sweep_tau(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_sweep<false, true>(c, a, (int*)smem);
}

__global__ void __launch_bounds__(512) k_sweep_tau_lds(LaunchArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem + kSlotBytes};
    body_sweep<true, true>(c, a, (int*)smem);
}

// Before each persistent kernel: its work counter and, for a paired launch,
// the pairs' barrier words.
__global__ void __launch_bounds__(64) k_reset_queue(LaunchArgs a) {
    if (threadIdx.x == 0) *a.queue = 0;
    if (a.pairs)
        for (int i = threadIdx.x; i < 2 * a.npairs; i += blockDim.x) a.pair_flags[i] = 0;
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

// ------------------------------------------------------------------ network
__global__ void __launch_bounds__(256) k_conv_fwd(const float* in, int cin, const float* W, const float* b, float* out,
                                                  int cout, int n, NetShape s, int relu) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const float y = conv_fwd_elem(in, cin, W, b, cout, s, e);
        out[e] = relu && y < 0.0f ? 0.0f : y;
    }
}
__global__ void __launch_bounds__(256) k_conv_bwd_data(const float* gout, int cout, const float* W, float* gin,
                                                       int cin, const float* act, int mask, int n, NetShape s) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
        gin[e] = conv_bwd_data_elem(gout, cout, W, cin, act, mask != 0, s, e);
}
// one workgroup per (output, input) channel pair computes the weights of all 27
// taps: each position's output gradient is read once and multiplied by the 27
// neighbours of its input (cached), instead of one workgroup re-reading both
// arrays for every tap; per weight the threads' partial sums and their tree
// reduction are those of a workgroup per weight
__global__ void __launch_bounds__(256) k_conv_bwd_weight(const float* gout, int cout, const float* in, int cin,
                                                         float* gW, int batch, NetShape s) {
    __shared__ float part[27][256];
    const int co = blockIdx.x / cin, ci = blockIdx.x - co * cin;
    const int v = s.vol(), plane = s.ly * s.lx;
    float acc[27];
#pragma unroll
    for (int d = 0; d < 27; ++d) acc[d] = 0.0f;
    for (int q = threadIdx.x; q < batch * v; q += blockDim.x) {
        const int bb = q / v, p = q - bb * v;
        const int t = p / plane, r1 = p - t * plane;
        const int y = r1 / s.lx, x = r1 - y * s.lx;
        const float g = gout[((size_t)bb * cout + co) * v + p];
        const float* src = in + ((size_t)bb * cin + ci) * v;
        int tw[3], yw[3], xw[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            tw[o] = wrap_idx(t + o - 1, s.lt) * plane;
            yw[o] = wrap_idx(y + o - 1, s.ly) * s.lx;
            xw[o] = wrap_idx(x + o - 1, s.lx);
        }
#pragma unroll
        for (int d = 0; d < 27; ++d) acc[d] += g * src[tw[d / 9] + yw[(d / 3) % 3] + xw[d % 3]];
    }
#pragma unroll
    for (int d = 0; d < 27; ++d) part[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int d = 0; d < 27; ++d) part[d][threadIdx.x] += part[d][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < 27) gW[(size_t)blockIdx.x * 27 + threadIdx.x] = part[threadIdx.x][0];
}
__global__ void __launch_bounds__(256) k_conv_bwd_bias(const float* gout, int cout, float* gb, int batch, NetShape s) {
    __shared__ float part[256];
    const int co = blockIdx.x, v = s.vol();
    float acc = 0.0f;
    for (int q = threadIdx.x; q < batch * v; q += blockDim.x) {
        const int bb = q / v, p = q - bb * v;
        acc += gout[((size_t)bb * cout + co) * v + p];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) gb[co] = part[0];
}
__global__ void __launch_bounds__(256) k_mse_grad(const float* y, const float* t, float* g, size_t n, float scale) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        g[e] = scale * (y[e] - t[e]);
}
__global__ void __launch_bounds__(256) k_adam(float* p, const float* g, float* m, float* v, float lr, int step,
                                              size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        adam_elem(p, g, m, v, lr, step, (int)e);
}

// ---- learned global moves (gmove.hpp)
__global__ void __launch_bounds__(256) k_gm_to_net(const double* phi, float* x, GmArgs a) {
    const size_t n = (size_t)a.batch * 2 * a.s.vol();
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        gm_to_net_elem(phi, x, a, (int)e);
}
__global__ void __launch_bounds__(256) k_gm_propose(double* phi, double* old, double* f0, const float* y, GmArgs a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        gm_propose_elem(phi, old, f0, y, a, (int)e);
}
// one workgroup per chain: strided partial sums, then a tree in LDS
__global__ void __launch_bounds__(256) k_gm_sums(const double* phi, const double* old, const double* f0,
                                                 const float* y, GmArgs a, double* out) {
    __shared__ double part[5][256];
    const int b = blockIdx.x, tid = threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    gm_sums(phi, old, f0, y, a, b, tid, gm_per(a.s), 256, acc);
    for (int k = 0; k < 5; ++k) part[k][tid] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 5; ++k) part[k][tid] += part[k][tid + h];
        __syncthreads();
    }
    if (tid < 5) out[(size_t)(a.c0 + b) * 5 + tid] = part[tid][0];
}
__global__ void __launch_bounds__(256) k_gm_restore(double* phi, const double* old, const int* accept, int per,
                                                    size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        gm_restore_elem(phi, old, accept, per, (int)e);
}

__global__ void __launch_bounds__(256) k_gm_restore_words(uint64_t* dst, const uint64_t* src, const int* accept,
                                                          size_t per, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        if (!accept[e / per]) dst[e] = src[e];
}

__global__ void __launch_bounds__(256) k_gm_hmc_init(const double* phi, double* old, double* p0, double* p, GmArgs a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        gm_hmc_init_elem(phi, old, p0, p, a, (int)e);
}
__global__ void __launch_bounds__(256) k_gm_hmc_kick(double* p, const float* y, GmArgs a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        gm_hmc_kick_elem(p, y, a, (int)e);
}
__global__ void __launch_bounds__(256) k_gm_hmc_rotate(double* phi, double* p, GmArgs a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        gm_hmc_rotate_elem(phi, p, a, (int)e);
}
__global__ void __launch_bounds__(256) k_gm_hmc_sums(const double* phi, const double* old, const double* p0,
                                                     const double* p, GmArgs a, double* out) {
    __shared__ double part[5][256];
    const int b = blockIdx.x, tid = threadIdx.x;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    gm_hmc_sums(phi, old, p0, p, a, b, tid, gm_per(a.s), 256, acc);
    for (int k = 0; k < 5; ++k) part[k][tid] = acc[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 5; ++k) part[k][tid] += part[k][tid + h];
        __syncthreads();
    }
    if (tid < 5) out[(size_t)(a.c0 + b) * 5 + tid] = part[tid][0];
}

__global__ void __launch_bounds__(256) k_gm_copy_accepted(uint64_t* dst, const uint64_t* src, const int* accept,
                                                          size_t per, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        if (accept[e / per]) dst[e] = src[e];
}

static int grid_for(size_t n) { return (int)((n + 255) / 256 < 65535 ? (n + 255) / 256 : 65535); }

int rt_gm_restore_words(uint64_t* dst, const uint64_t* src, const int* accept, size_t per, size_t n) {
    hipLaunchKernelGGL(k_gm_restore_words, dim3(grid_for(n)), dim3(256), 0, 0, dst, src, accept, per, n);
    return check(hipGetLastError(), "k_gm_restore_words");
}

int rt_gm_hmc_init(const double* phi, double* old, double* p0, double* p, const GmArgs& a) {
    hipLaunchKernelGGL(k_gm_hmc_init, dim3(grid_for((size_t)a.batch * gm_per(a.s))), dim3(256), 0, 0, phi, old, p0, p,
                       a);
    return check(hipGetLastError(), "k_gm_hmc_init");
}
int rt_gm_hmc_kick(double* p, const float* y, const GmArgs& a) {
    hipLaunchKernelGGL(k_gm_hmc_kick, dim3(grid_for((size_t)a.batch * gm_per(a.s))), dim3(256), 0, 0, p, y, a);
    return check(hipGetLastError(), "k_gm_hmc_kick");
}
int rt_gm_hmc_rotate(double* phi, double* p, const GmArgs& a) {
    hipLaunchKernelGGL(k_gm_hmc_rotate, dim3(grid_for((size_t)a.batch * gm_per(a.s))), dim3(256), 0, 0, phi, p, a);
    return check(hipGetLastError(), "k_gm_hmc_rotate");
}
int rt_gm_hmc_sums(const double* phi, const double* old, const double* p0, const double* p, const GmArgs& a,
                   double* out) {
    hipLaunchKernelGGL(k_gm_hmc_sums, dim3(a.batch), dim3(256), 0, 0, phi, old, p0, p, a, out);
    return check(hipGetLastError(), "k_gm_hmc_sums");
}

int rt_gm_copy_accepted(uint64_t* dst, const uint64_t* src, const int* accept, size_t per, size_t n) {
    hipLaunchKernelGGL(k_gm_copy_accepted, dim3(grid_for(n)), dim3(256), 0, 0, dst, src, accept, per, n);
    return check(hipGetLastError(), "k_gm_copy_accepted");
}

int rt_gm_to_net(const double* phi, float* x, const GmArgs& a) {
    hipLaunchKernelGGL(k_gm_to_net, dim3(grid_for((size_t)a.batch * 2 * a.s.vol())), dim3(256), 0, 0, phi, x, a);
    return check(hipGetLastError(), "k_gm_to_net");
}
int rt_gm_propose(double* phi, double* old, double* f0, const float* y, const GmArgs& a) {
    hipLaunchKernelGGL(k_gm_propose, dim3(grid_for((size_t)a.batch * gm_per(a.s))), dim3(256), 0, 0, phi, old, f0, y,
                       a);
    return check(hipGetLastError(), "k_gm_propose");
}
int rt_gm_sums(const double* phi, const double* old, const double* f0, const float* y, const GmArgs& a, double* out) {
    hipLaunchKernelGGL(k_gm_sums, dim3(a.batch), dim3(256), 0, 0, phi, old, f0, y, a, out);
    return check(hipGetLastError(), "k_gm_sums");
}
int rt_gm_restore(double* phi, const double* old, const int* accept, int per, size_t n) {
    hipLaunchKernelGGL(k_gm_restore, dim3(grid_for(n)), dim3(256), 0, 0, phi, old, accept, per, n);
    return check(hipGetLastError(), "k_gm_restore");
}

template <int NB>
__global__ void __launch_bounds__(kConvWaves * 64) k_conv_mfma(ConvArgs a) {
    extern __shared__ __align__(16) char smem[];
    Ctx c{smem};
    conv_mfma_body<NB>(c, a);
}

// The matrix-core convolution (conv.hpp) when the layer fits its plan, else
// 1 (the caller runs the per-element kernel).  BAFQ_CONV=elem selects the
// per-element kernels everywhere (for comparisons).
static int conv_mfma_launch(ConvArgs a, int batch, int* used) {
    *used = 0;
    const char* env = std::getenv("BAFQ_CONV");
    if (env && std::string(env) == "elem") return 0;
    const int plane = a.s.ly * a.s.lx;
    a.cinp = conv_pad4(a.cin);
    a.coutp = conv_pad16(a.cout);
    const int nb = a.coutp / 16;
    if (nb > 4) return 0;
    a.tt = conv_pick_tt(a.s.lt, plane, a.cinp, a.coutp, rt_lds_limit());
    if (a.tt == 0) return 0;
    a.tblocks = (a.s.lt + a.tt - 1) / a.tt;
    const size_t lb = conv_lds_bytes(a.tt, a.cinp, a.coutp, plane);
    const void* fn = nb == 1 ? (const void*)k_conv_mfma<1> : nb == 2 ? (const void*)k_conv_mfma<2>
                   : nb == 3 ? (const void*)k_conv_mfma<3> : (const void*)k_conv_mfma<4>;
    if (lb > 65536 && check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb),
                            "hipFuncSetAttribute"))
        return 1;
    const dim3 grid(batch * a.tblocks), block(kConvWaves * 64);
    if (nb == 1) hipLaunchKernelGGL(k_conv_mfma<1>, grid, block, lb, 0, a);
    else if (nb == 2) hipLaunchKernelGGL(k_conv_mfma<2>, grid, block, lb, 0, a);
    else if (nb == 3) hipLaunchKernelGGL(k_conv_mfma<3>, grid, block, lb, 0, a);
    else hipLaunchKernelGGL(k_conv_mfma<4>, grid, block, lb, 0, a);
    *used = 1;
    return check(hipGetLastError(), "k_conv_mfma");
}

int rt_conv_fwd(const float* in, int cin, const float* W, const float* b, float* out, int cout, int batch,
                NetShape s, int relu) {
    ConvArgs ca{};
    ca.in = in; ca.W = W; ca.bias = b; ca.mask = nullptr; ca.out = out;
    ca.cin = cin; ca.cout = cout; ca.relu = relu; ca.flip = 0; ca.s = s;
    int used = 0;
    if (conv_mfma_launch(ca, batch, &used)) return 1;
    if (used) return 0;
    const int n = batch * cout * s.vol();
    hipLaunchKernelGGL(k_conv_fwd, dim3(grid_for(n)), dim3(256), 0, 0, in, cin, W, b, out, cout, n, s, relu);
    return check(hipGetLastError(), "k_conv_fwd");
}
int rt_conv_bwd_data(const float* gout, int cout, const float* W, float* gin, int cin, const float* act, int mask,
                     int batch, NetShape s) {
    // the forward layer's transposed, reflected stencil from cout to cin channels
    ConvArgs ca{};
    ca.in = gout; ca.W = W; ca.bias = nullptr; ca.mask = mask ? act : nullptr; ca.out = gin;
    ca.cin = cout; ca.cout = cin; ca.relu = 0; ca.flip = 1; ca.s = s;
    int used = 0;
    if (conv_mfma_launch(ca, batch, &used)) return 1;
    if (used) return 0;
    const int n = batch * cin * s.vol();
    hipLaunchKernelGGL(k_conv_bwd_data, dim3(grid_for(n)), dim3(256), 0, 0, gout, cout, W, gin, cin, act, mask, n, s);
    return check(hipGetLastError(), "k_conv_bwd_data");
}
int rt_conv_bwd_weight(const float* gout, int cout, const float* in, int cin, float* gW, float* gb, int batch,
                       NetShape s) {
    hipLaunchKernelGGL(k_conv_bwd_weight, dim3(cout * cin), dim3(256), 0, 0, gout, cout, in, cin, gW, batch, s);
    if (check(hipGetLastError(), "k_conv_bwd_weight")) return 1;
    hipLaunchKernelGGL(k_conv_bwd_bias, dim3(cout), dim3(256), 0, 0, gout, cout, gb, batch, s);
    return check(hipGetLastError(), "k_conv_bwd_bias");
}
int rt_mse_grad(const float* y, const float* t, float* g, size_t n, float scale) {
    hipLaunchKernelGGL(k_mse_grad, dim3(grid_for(n)), dim3(256), 0, 0, y, t, g, n, scale);
    return check(hipGetLastError(), "k_mse_grad");
}
int rt_adam(float* p, const float* g, float* m, float* v, float lr, int step, size_t n) {
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(256), 0, 0, p, g, m, v, lr, step, n);
    return check(hipGetLastError(), "k_adam");
}

// Hidden activations of bafq_net_forward_noalloc: a file-scope device array,
// one copy per GPU, held until the process ends.  In host code its name is
// only a stand-in: the device address comes from hipGetSymbolAddress, taken
// once per device with that device current.
static __device__ __align__(256) float scratch[kNetScratchFloats];

int rt_net_scratch(float** ptr) {
    static std::mutex mu;
    static float* per_device[kMaxScratchDevices] = {};
    int dev = -1;
    if (check(hipGetDevice(&dev), "hipGetDevice")) return 1;
    if (dev < 0 || dev >= kMaxScratchDevices) {
        g_err = "device index beyond the scratch table";
        return 1;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!per_device[dev]) {
        void* p = nullptr;
        if (check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress")) return 1;
        per_device[dev] = (float*)p;
    }
    *ptr = per_device[dev];
    return 0;
}
int rt_launch_status() { return check(hipGetLastError(), "kernel launch"); }

// Lane-layout probe of the FP32 MFMA 16x16x4: as k_debug_mfma.
__global__ void __launch_bounds__(64) k_debug_mfma32(float* out) {
    Ctx c{nullptr};
    const int p = blockIdx.x, l = threadIdx.x;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = c.mfma32(l == p ? 1.0f : 0.0f, 1000.0f + l, acc);
    for (int r = 0; r < 4; ++r) out[(p * 64 + l) * 4 + r] = acc[r];
}

int rt_debug_mfma32(float* out) {
    float* d = nullptr;
    if (rt_malloc((void**)&d, sizeof(float) * 64 * 64 * 4)) return 1;
    hipLaunchKernelGGL(k_debug_mfma32, dim3(64), dim3(64), 0, 0, d);
    int rc = check(hipGetLastError(), "k_debug_mfma32") || rt_sync() || rt_d2h(out, d, sizeof(float) * 64 * 64 * 4);
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

struct Stopwatch {
    hipEvent_t ev[2] = {nullptr, nullptr};
};
int rt_timer_start(void** t) {
    if (!*t) {
        Stopwatch* w = new Stopwatch();
        *t = w;
        if (check(hipEventCreate(&w->ev[0]), "hipEventCreate") || check(hipEventCreate(&w->ev[1]), "hipEventCreate"))
            return 1;
    }
    return check(hipEventRecord(((Stopwatch*)*t)->ev[0], 0), "hipEventRecord");
}
int rt_timer_stop(void* t, double* ms) {
    Stopwatch* w = (Stopwatch*)t;
    if (check(hipEventRecord(w->ev[1], 0), "hipEventRecord") ||
        check(hipEventSynchronize(w->ev[1]), "hipEventSynchronize"))
        return 1;
    float f = 0.0f;
    if (check(hipEventElapsedTime(&f, w->ev[0], w->ev[1]), "hipEventElapsedTime")) return 1;
    *ms = (double)f;
    return 0;
}
void rt_timer_free(void* t) {
    if (!t) return;
    Stopwatch* w = (Stopwatch*)t;
    for (hipEvent_t& e : w->ev)
        if (e) (void)hipEventDestroy(e);
    delete w;
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
int rt_d2d(void* dst, const void* src, size_t bytes) {
    return check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}
int rt_sync() { return check(hipDeviceSynchronize(), "hipDeviceSynchronize"); }
const char* rt_error() { return g_err.c_str(); }

// Device-memory path: 8 wavefronts per workgroup (2 per SIMD) hide the HBM
// and L2 latency of its matrices better than 4 (21-28 % faster sweeps at
// N = 144 and 288).  BAFQ_BLOCK selects another multiple of 64 for experiments.
int rt_block_threads() {
    const char* s = std::getenv("BAFQ_BLOCK");
    int b = s ? std::atoi(s) : 512;
    if (b < 64 || b > 512 || (b % 64) != 0) b = 512;
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

static const void* kernel_of(KernelId k, PathId path) {
    if (k == K_PREPARE) return path == P_LDS ? (const void*)k_prepare_lds : (const void*)k_prepare;
    if (k == K_GRAD) return path == P_LDS ? (const void*)k_grad_lds : (const void*)k_grad;
    if (k == K_MEASURE) return path == P_LDS ? (const void*)k_measure_lds : (const void*)k_measure;
    if (k == K_SWEEP_TAU) return path == P_LDS ? (const void*)k_sweep_tau_lds : (const void*)k_sweep_tau;
    return path == P_LDS ? (const void*)k_sweep_lds : (const void*)k_sweep;
}


int rt_resident_groups(KernelId k, PathId path, int block, size_t smem) {
    const void* fn = kernel_of(k, path);
    if (smem > 65536 &&
        check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem), "hipFuncSetAttribute"))
        return -1;
    int per_cu = 0, cus = 0;
    if (check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, smem), "hipOccupancy") ||
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "hipDeviceGetAttribute"))
        return -1;
    return per_cu > 0 && cus > 0 ? per_cu * cus : 1;
}

// Launch state of one handle: its timing events and instantiated graphs, keyed
// by (kernel, sweeps, measure, time-displaced mode and recording); each graph holds, per path, a queue reset
// followed by the persistent kernel, the paths as independent branches that
// run concurrently.
struct GraphCache {
    hipEvent_t ev[2] = {nullptr, nullptr};
    struct Entry {
        int kind, nsweep, measure, tau, tau_record;
        std::vector<LaunchArgs> args;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Entry> entries;
    void clear() {
        for (Entry& e : entries) {
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
            if (e.graph) (void)hipGraphDestroy(e.graph);
        }
        entries.clear();
    }
};

void rt_release_launches(void* cache) {
    if (!cache) return;
    GraphCache* gc = (GraphCache*)cache;
    gc->clear();
    for (hipEvent_t& e : gc->ev)
        if (e) (void)hipEventDestroy(e);
    delete gc;
}

static GraphCache* launch_state(void** cache) {
    if (!*cache) *cache = new GraphCache();
    return (GraphCache*)*cache;
}

static int timed_begin(GraphCache* gc) {
    if (!gc->ev[0]) {
        if (check(hipEventCreate(&gc->ev[0]), "hipEventCreate") || check(hipEventCreate(&gc->ev[1]), "hipEventCreate"))
            return 1;
    }
    return check(hipEventRecord(gc->ev[0], 0), "hipEventRecord");
}

static int timed_end(GraphCache* gc, double* ms) {
    if (check(hipEventRecord(gc->ev[1], 0), "hipEventRecord") ||
        check(hipEventSynchronize(gc->ev[1]), "hipEventSynchronize"))
        return 1;
    float f = 0.0f;
    if (check(hipEventElapsedTime(&f, gc->ev[0], gc->ev[1]), "hipEventElapsedTime")) return 1;
    if (ms) *ms = (double)f;
    return 0;
}

int rt_launch_warmup(void** cache, const LaunchArgs& a, int grid, int block, double* ms) {
    GraphCache* gc = launch_state(cache);
    if (timed_begin(gc)) return 1;
    hipLaunchKernelGGL(k_warmup, dim3(grid), dim3(block), 0, 0, a);
    if (check(hipGetLastError(), "kernel launch")) return 1;
    return timed_end(gc, ms);
}

static int build_graph(GraphCache::Entry& e, KernelId k, const PathLaunch* pl, int n) {
    if (check(hipGraphCreate(&e.graph, 0), "hipGraphCreate")) return 1;
    e.args.resize(n);
    for (int i = 0; i < n; ++i) {
        e.args[i] = pl[i].a;
        const void* fn = kernel_of(k, pl[i].path);
        if (pl[i].smem > 65536 &&
            check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl[i].smem),
                  "hipFuncSetAttribute"))
            return 1;
        hipGraphNode_t reset, kern;
        void* rargs[] = {(void*)&e.args[i]};
        hipKernelNodeParams rp{};
        rp.func = (void*)k_reset_queue;
        rp.gridDim = dim3(1);
        rp.blockDim = dim3(64);
        rp.sharedMemBytes = 0;
        rp.kernelParams = rargs;
        rp.extra = nullptr;
        if (check(hipGraphAddKernelNode(&reset, e.graph, nullptr, 0, &rp), "hipGraphAddKernelNode")) return 1;
        void* kargs[] = {(void*)&e.args[i]};
        hipKernelNodeParams kp{};
        kp.func = (void*)fn;
        kp.gridDim = dim3(pl[i].grid);
        kp.blockDim = dim3(pl[i].block);
        kp.sharedMemBytes = (unsigned)pl[i].smem;
        kp.kernelParams = kargs;
        kp.extra = nullptr;
        if (check(hipGraphAddKernelNode(&kern, e.graph, &reset, 1, &kp), "hipGraphAddKernelNode")) return 1;
    }
    return check(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

int rt_launch_paths(void** cache, KernelId k, const PathLaunch* pl, int n, double* ms) {
    GraphCache* gc = launch_state(cache);
    const int nsweep = n ? pl[0].a.nsweep : 0, measure = n ? pl[0].a.measure : 0;
    const int tau = n ? pl[0].a.tau : 0, tau_record = n ? pl[0].a.tau_record : 0;
    GraphCache::Entry* hit = nullptr;
    for (GraphCache::Entry& e : gc->entries)
        if (e.kind == (int)k && e.nsweep == nsweep && e.measure == measure && e.tau == tau &&
            e.tau_record == tau_record)
            hit = &e;
    if (!hit) {
        if (gc->entries.size() >= 8) gc->clear();
        gc->entries.emplace_back();
        hit = &gc->entries.back();
        hit->kind = (int)k;
        hit->nsweep = nsweep;
        hit->measure = measure;
        hit->tau = tau;
        hit->tau_record = tau_record;
        if (build_graph(*hit, k, pl, n)) {
            gc->clear();
            return 1;
        }
    }
    if (timed_begin(gc)) return 1;
    if (check(hipGraphLaunch(hit->exec, 0), "hipGraphLaunch")) return 1;
    return timed_end(gc, ms);
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
