// Execution-context layer for the workgroup-cooperative solver code.
//
// The solver's device code (la.hpp, sweep.hpp) is written against a small
// context object `Ctx` that provides thread/wave indices, the workgroup
// barrier, 64-lane wavefront shuffles and the FP64 MFMA 16x16x4 tile
// operation.  Two back ends implement it:
//
//   * HipCtx  (default): CDNA4 / gfx950, compiled by hipcc.
//   * EmuCtx  (-DBAFQ_EMU): a CPU emulator in which every GPU thread of a
//     workgroup is a fiber (user-level context) on one OS thread, barriers
//     switch between fibers, and a wavefront's shuffles and MFMA are evaluated
//     from the 64 lanes' operands.  It exists so the same kernels can be run
//     and bounds-checked on a machine without a GPU.
//
// MFMA f64 16x16x4 operand layout (CDNA3/CDNA4): lane l holds
//   A[i][k] with i = l % 16, k = l / 16;  B[k][j] with j = l % 16, k = l / 16;
//   D[i][j] in accumulator register r with i = l / 16 + 4 * r, j = l % 16
// (checked on gfx950 by bafq_debug_mfma_probe / tests/test_gpu_parity.py).
#pragma once

#include <cstdint>
#include <cmath>

#ifndef BAFQ_EMU
#include <hip/hip_runtime.h>
#define DEV __device__ __forceinline__
#define DEVFN __device__ __noinline__
#define HOSTDEV __host__ __device__ inline
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
#else
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>
#define DEV inline
#define DEVFN inline
#define HOSTDEV inline
struct d4 {
    double v[4];
    double& operator[](int i) { return v[i]; }
    double operator[](int i) const { return v[i]; }
};
struct f4 {
    float v[4];
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};
#endif

namespace bafq {

constexpr int kWave = 64;

// Row of the 16x16 MFMA f64 result held by lane `lane` in accumulator register r.
HOSTDEV int mfma_row(int lane, int r) { return (lane >> 4) + 4 * r; }
// Row of the 16x16 MFMA f32 (16x16x4f32) result held by lane `lane` in
// accumulator register r; operands as for f64: A[l % 16][l / 16],
// B[l / 16][l % 16]; column l % 16 (checked on gfx950 by
// bafq_debug_mfma32_probe / tests/test_gpu_parity.py).
HOSTDEV int mfma32_row(int lane, int r) { return 4 * (lane >> 4) + r; }

#ifndef BAFQ_EMU
struct HipCtx {
    char* smem;
    DEV int tid() const { return (int)threadIdx.x; }
    DEV int nthreads() const { return (int)blockDim.x; }
    DEV int lane() const { return (int)(threadIdx.x & 63); }
    DEV int wave() const { return (int)(threadIdx.x >> 6); }
    DEV int nwaves() const { return (int)(blockDim.x >> 6); }
    DEV int block() const { return (int)blockIdx.x; }
    DEV int nblocks() const { return (int)gridDim.x; }
    DEV void sync() const { __syncthreads(); }
    DEV double shfl_xor(double v, int mask) const { return __shfl_xor(v, mask, 64); }
    DEV int shfl_xor_i(int v, int mask) const { return __shfl_xor(v, mask, 64); }
    DEV d4 mfma(double a, double b, d4 c) const {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    DEV f4 mfma32(float a, float b, f4 c) const {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    DEV int atomic_add(int* p, int v) const { return atomicAdd(p, v); }
    // constant-rate device clock (hipDeviceAttributeWallClockRate kHz)
    DEV uint64_t clock() const { return wall_clock64(); }
    // Ordering for workgroups that share data through device memory (agent
    // scope: also across the XCDs' L2 caches).  stores_done: this wavefront's
    // stores have reached L2 (the L1 is write-through); arrive: a release
    // increment (one L2 write-back for the workgroup); a waiter polls with
    // relaxed loads and takes one acquire fence after the last (an acquire
    // load per poll would invalidate the caches on every iteration).
    DEV void stores_done() const { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
    DEV void arrive(int* p) const { __hip_atomic_fetch_add(p, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }
    DEV int load_relaxed(int* p) const { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    DEV void acquire_device() const { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
    DEV void store_flag(int* p, int v) const { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    DEV void pause() const { __builtin_amdgcn_s_sleep(4); }
    // Pair and rank of workgroup b in a grid of g workgroups run as pairs.
    // Workgroups are dispatched to the 8 XCDs in turn (b % 8), so b and b + 8
    // of each group of 16 share an XCD and its L2; other grids pair b, b + 1.
    DEV static void pair_of(int b, int g, int& pair, int& rank) {
        if (g % 16 == 0) {
            pair = (b >> 4) * 8 + (b & 7);
            rank = (b >> 3) & 1;
        } else {
            pair = b >> 1;
            rank = b & 1;
        }
    }

    // DPP lane permutations (no LDS round trip): quad_perm [1,0,3,2] = 0xB1,
    // quad_perm [2,3,0,1] = 0x4E, row_half_mirror = 0x141, row_mirror = 0x140.
    // Every source lane of these permutations exists; callers keep all lanes of
    // the wavefront active (bound_ctrl would read 0 from an inactive lane).
    template <int C, class T>
    DEV static T dpp(T v) { return __builtin_amdgcn_update_dpp(T(0), v, C, 0xF, 0xF, true); }
    // Sum over the aligned quad of lanes; all four lanes get the same bits.
    DEV double quad_sum(double v) const {
        v += dpp<0xB1>(v);
        return v + dpp<0x4E>(v);
    }
    // Sum over the aligned group of 8 lanes; all eight lanes get the same bits.
    DEV double oct_sum(double v) const {
        v += dpp<0xB1>(v);
        v += dpp<0x4E>(v);
        return v + dpp<0x141>(v);
    }
    // Sum over the aligned row of 16 lanes; all sixteen lanes get the same bits.
    DEV double row16_sum(double v) const {
        v += dpp<0xB1>(v);
        v += dpp<0x4E>(v);
        v += dpp<0x141>(v);
        return v + dpp<0x140>(v);
    }
    // Value of lane l (uniform) in every lane.  v_readlane_b32 moves 32 bits:
    // a double goes through as its two halves.
    DEV static double readlane_d(double v, int l) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    }
    DEV double lane_bcast(double v, int l) const { return readlane_d(v, l); }
    DEV bool wave_any(bool p) const { return __any(p); }
    // A value that is the same in every lane, moved to a scalar register so that
    // branches on it are scalar branches.
    DEV static int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
    // Wavefront argmax of (v, i), ties to the smaller i; the result is uniform.
    DEV void wave_argmax(double& v, int& i) const {
        argmax_step(v, i, dpp<0xB1>(v), dpp<0xB1>(i));
        argmax_step(v, i, dpp<0x4E>(v), dpp<0x4E>(i));
        argmax_step(v, i, dpp<0x141>(v), dpp<0x141>(i));
        argmax_step(v, i, dpp<0x140>(v), dpp<0x140>(i));
        double bv = readlane_d(v, 0);
        int bi = __builtin_amdgcn_readlane(i, 0);
        for (int r = 16; r < 64; r += 16)
            argmax_step(bv, bi, readlane_d(v, r), __builtin_amdgcn_readlane(i, r));
        v = bv;
        i = bi;
    }
    DEV static void argmax_step(double& v, int& i, double ov, int oi) {
        if (ov > v || (ov == v && oi < i)) {
            v = ov;
            i = oi;
        }
    }
};
using Ctx = HipCtx;
#else
// ---------------------------------------------------------------- emulator
struct EmuGroup {
    int nthreads;
    std::vector<double> wbuf;            // [wave][3][64] exchange buffers
    char* smem;
    int block, nblocks;
    // barrier of the workgroup / of thread t's wavefront (emu.cpp)
    void (*group_barrier)(EmuGroup*, int t);
    void (*wave_barrier)(EmuGroup*, int t);
    void* sched;
};

struct EmuCtx {
    EmuGroup* g;
    int t;
    char* smem;
    int tid() const { return t; }
    int nthreads() const { return g->nthreads; }
    int lane() const { return t & 63; }
    int wave() const { return t >> 6; }
    int nwaves() const { return g->nthreads >> 6; }
    int block() const { return g->block; }
    int nblocks() const { return g->nblocks; }
    void sync() const { g->group_barrier(g, t); }
    void wsync() const { g->wave_barrier(g, t); }
    double* wb(int slot) const { return &g->wbuf[(size_t)(wave() * 3 + slot) * 64]; }
    double shfl_xor(double v, int mask) const {
        double* b = wb(0);
        b[lane()] = v;
        wsync();
        double r = b[lane() ^ mask];
        wsync();
        return r;
    }
    int shfl_xor_i(int v, int mask) const { return (int)shfl_xor((double)v, mask); }
    d4 mfma(double a, double b, d4 c) const {
        double* A = wb(0);
        double* B = wb(1);
        A[lane()] = a;
        B[lane()] = b;
        wsync();
        const int l = lane(), j = l & 15;
        for (int r = 0; r < 4; ++r) {
            const int i = mfma_row(l, r);
            double s = c[r];
            for (int k = 0; k < 4; ++k) s += A[i + 16 * k] * B[j + 16 * k];
            c[r] = s;
        }
        wsync();
        return c;
    }
    // f32 16x16x4: products accumulated in k order in FP32 (the hardware's
    // order of the four products may differ; tests compare with tolerances)
    f4 mfma32(float a, float b, f4 c) const {
        double* A = wb(0);
        double* B = wb(1);
        A[lane()] = a;
        B[lane()] = b;
        wsync();
        const int l = lane(), j = l & 15;
        for (int r = 0; r < 4; ++r) {
            const int i = mfma32_row(l, r);
            float s = c[r];
            for (int k = 0; k < 4; ++k) s += (float)A[i + 16 * k] * (float)B[j + 16 * k];
            c[r] = s;
        }
        wsync();
        return c;
    }
    int atomic_add(int* p, int v) const {
        return reinterpret_cast<std::atomic<int>*>(p)->fetch_add(v);
    }
    void stores_done() const { std::atomic_thread_fence(std::memory_order_release); }
    void arrive(int* p) const { reinterpret_cast<std::atomic<int>*>(p)->fetch_add(1, std::memory_order_release); }
    int load_relaxed(int* p) const { return reinterpret_cast<std::atomic<int>*>(p)->load(std::memory_order_relaxed); }
    void acquire_device() const { std::atomic_thread_fence(std::memory_order_acquire); }
    void store_flag(int* p, int v) const { reinterpret_cast<std::atomic<int>*>(p)->store(v, std::memory_order_relaxed); }
    void pause() const { std::this_thread::yield(); }
    // Workgroups run on a few OS threads in index order: pairs are b, b + 1,
    // so a workgroup that waits for its partner never holds back the partner's start.
    static void pair_of(int b, int, int& pair, int& rank) {
        pair = b >> 1;
        rank = b & 1;
    }
    // nanoseconds
    uint64_t clock() const {
        return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
                   std::chrono::steady_clock::now().time_since_epoch()).count();
    }
    double row16_sum(double v) const {
        for (int m = 1; m < 16; m <<= 1) v += shfl_xor(v, m);
        return v;
    }
    double oct_sum(double v) const {
        for (int m = 1; m < 8; m <<= 1) v += shfl_xor(v, m);
        return v;
    }
    double lane_bcast(double v, int l) const {
        double* b = wb(2);
        b[lane()] = v;
        wsync();
        const double r = b[l];
        wsync();
        return r;
    }
    static int uniform(int v) { return v; }
    bool wave_any(bool p) const {
        double v = p ? 1.0 : 0.0;
        for (int m = 1; m < 64; m <<= 1) v += shfl_xor(v, m);
        return v > 0.0;
    }
    double quad_sum(double v) const {
        v += shfl_xor(v, 1);
        return v + shfl_xor(v, 2);
    }
    void wave_argmax(double& v, int& i) const {
        for (int m = 1; m < 64; m <<= 1) {
            const double ov = shfl_xor(v, m);
            const int oi = shfl_xor_i(i, m);
            if (ov > v || (ov == v && oi < i)) {
                v = ov;
                i = oi;
            }
        }
    }
};
using Ctx = EmuCtx;
#endif

}  // namespace bafq
