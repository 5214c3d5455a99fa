// CPU emulator back end (-DBAFQ_EMU): runs the same workgroup kernels with one
// OS thread per GPU thread, for testing and bounds-checking without a GPU.
// Device memory is host memory; workgroups of a launch run one after another.
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

#include "api.hpp"

namespace bafq {

static std::string g_err;

template <class F>
static void run_groups(int grid, int block, size_t smem, F&& body) {
    const int nw = block / 64;
    for (int g = 0; g < grid; ++g) {
        std::vector<char> lds(smem + 64, 0);
        std::barrier<> bar(block);
        std::vector<std::barrier<>*> wb;
        for (int w = 0; w < nw; ++w) wb.push_back(new std::barrier<>(64));
        EmuGroup grp;
        grp.nthreads = block;
        grp.bar = &bar;
        grp.wbar = wb;
        grp.wbuf.assign((size_t)nw * 3 * 64, 0.0);
        grp.smem = lds.data();
        grp.block = g;
        grp.nblocks = grid;
        std::vector<std::thread> th;
        th.reserve(block);
        for (int t = 0; t < block; ++t)
            th.emplace_back([&, t]() {
                EmuCtx c{&grp, t, grp.smem};
                body(c);
            });
        for (auto& x : th) x.join();
        for (auto* p : wb) delete p;
    }
}

int rt_malloc(void** p, size_t bytes) {
    *p = std::calloc(1, bytes ? bytes : 1);
    if (!*p) { g_err = "out of memory"; return 1; }
    return 0;
}
void rt_free(void* p) { std::free(p); }
int rt_h2d(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); return 0; }
int rt_d2h(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); return 0; }
int rt_memset(void* dst, int v, size_t bytes) { std::memset(dst, v, bytes); return 0; }
int rt_sync() { return 0; }
const char* rt_error() { return g_err.c_str(); }

int rt_block_threads() {
    const char* s = std::getenv("BAFQ_BLOCK");
    int b = s ? std::atoi(s) : 128;
    if (b < 64 || b > 1024 || (b % 64) != 0) b = 128;
    return b;
}

double rt_clock_hz() { return 1.0e9; }

// The emulator may run the LDS path with fewer waves (small N only; the host
// checks that the waves cover every MFMA tile).
int rt_lds_block_threads() {
    const char* s = std::getenv("BAFQ_LDS_BLOCK");
    int b = s ? std::atoi(s) : 128;
    if (b < 64 || b > 512 || (b % 64) != 0) b = 128;
    return b;
}

size_t rt_lds_limit() { return 160 * 1024; }

int rt_launch(KernelId k, PathId path, const LaunchArgs& a, int grid, int block, size_t smem, double* ms) {
    if (ms) *ms = 0.0;
    if (k == K_WARMUP) {
        // one thread per chain; no workgroup cooperation
        for (int g = 0; g < grid; ++g)
            for (int t = 0; t < block; ++t) {
                EmuGroup grp;
                grp.nthreads = block;
                grp.block = g;
                grp.nblocks = grid;
                EmuCtx c{&grp, t, nullptr};
                body_warmup(c, a);
            }
        return 0;
    }
    run_groups(grid, block, smem, [&](EmuCtx& c) {
        if (k == K_PREPARE) {
            if (path == P_LDS) body_prepare<true>(c, a);
            else body_prepare<false>(c, a);
        } else {
            if (path == P_LDS) body_sweep<true>(c, a);
            else body_sweep<false>(c, a);
        }
    });
    return 0;
}

int rt_debug_lanes(const double* in, const int32_t* idx, double* out) {
    run_groups(1, 64, 0, [&](EmuCtx& c) {
        const int l = c.lane();
        const double v = in[l];
        out[l] = c.row16_sum(v);
        out[64 + l] = c.quad_sum(v);
        out[128 + l] = c.lane_bcast(v, 37);
        double bv = v;
        int bi = idx[l];
        c.wave_argmax(bv, bi);
        out[192 + l] = bv;
        out[256 + l] = (double)bi;
    });
    return 0;
}

int rt_debug_mfma(double* out) {
    for (int p = 0; p < 64; ++p)
        run_groups(1, 64, 0, [&](EmuCtx& c) {
            d4 acc = {0.0, 0.0, 0.0, 0.0};
            acc = c.mfma(c.lane() == p ? 1.0 : 0.0, 1000.0 + c.lane(), acc);
            for (int r = 0; r < 4; ++r) out[(p * 64 + c.lane()) * 4 + r] = acc[r];
        });
    return 0;
}

int rt_debug_gemm(int n, const double* a, int op_a, const double* b, int op_b, double* c) {
    run_groups(1, rt_block_threads(), 0, [&](EmuCtx& ctx) {
        gemm_cc(ctx, (cd*)c, (const cd*)a, op_a != 0, (const cd*)b, op_b != 0, n);
    });
    return 0;
}

int rt_debug_qrp(int n, double* a, double* tau, int32_t* perm, double* d) {
    const int block = rt_block_threads();
    run_groups(1, block, lds_bytes(n, block / 64, false), [&](EmuCtx& ctx) {
        const Lds L = carve(ctx.smem, n, ctx.nwaves(), false);
        qrp(ctx, (cd*)a, n, L.qr, L.dtmp);
        for (int k = ctx.tid(); k < n; k += ctx.nthreads()) {
            ((cd*)tau)[k] = L.qr.tau[k];
            perm[k] = L.qr.perm[k];
            d[k] = L.dtmp[k];
        }
    });
    return 0;
}

}  // namespace bafq
