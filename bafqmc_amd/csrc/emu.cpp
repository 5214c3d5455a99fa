// CPU emulator back end (-DBAFQ_EMU): runs the same workgroup kernels on the
// host, for testing and bounds-checking without a GPU.  Device memory is host
// memory.  The threads of a workgroup are fibers on one OS thread with their own
// stacks; a barrier hands control to the next fiber in thread order, so a
// workgroup barrier completes when the last thread arrives and a wavefront
// barrier when the wavefront's last lane does (every kernel here keeps barriers
// workgroup-uniform and shuffles wavefront-uniform).  Independent workgroups of
// a launch run in parallel on a few OS threads.  x86-64 only.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "api.hpp"

// Save the callee-saved registers, MXCSR and the x87 control word on the
// current stack, store the stack pointer in *from and resume the context whose
// stack pointer is `to`.
extern "C" void bafq_emu_switch(void** from, void* to);
asm(R"(
    .text
    .globl bafq_emu_switch
    .type bafq_emu_switch,@function
bafq_emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size bafq_emu_switch, .-bafq_emu_switch
)");

namespace bafq {

static std::string g_err;

namespace {

constexpr size_t kFiberStack = 256 * 1024;

struct Sched {
    std::vector<void*> sp;                 // saved stack pointer per fiber
    std::vector<std::unique_ptr<char[]>>* stacks;
    void* main_sp = nullptr;
    int cur = 0;
    int nthreads = 0;
    bool reverse = false;                  // run threads in descending order between barriers
    const std::function<void(EmuCtx&)>* body = nullptr;
    EmuGroup* grp = nullptr;
    std::vector<char> done;                // thread has returned from the body
    bool mismatch = false;
};

// Set when the threads of a workgroup passed different numbers of barriers
// (a thread waited at a barrier that a finished thread never reached): on the
// GPU that is undefined behaviour; here the launch fails with an error.
std::atomic<bool> g_barrier_mismatch{false};

thread_local Sched* tl_sched = nullptr;

void switch_to(Sched* s, int from, int to) {
    if (s->done[to]) {
        // abandon the workgroup: back to the launcher, never resumed
        s->mismatch = true;
        bafq_emu_switch(&s->sp[from], s->main_sp);
    }
    s->cur = to;
    bafq_emu_switch(&s->sp[from], s->sp[to]);
}

// Forward order: thread t hands over to t + 1; the last arrival restarts the
// round at thread 0.  Reverse order (BAFQ_EMU_REVERSE=1): t hands over to t - 1
// and rounds start at the last thread.  A result that depends on the order
// reveals a read that no barrier separates from the write it needs.
void group_barrier(EmuGroup* g, int t) {
    Sched* s = (Sched*)g->sched;
    if (!s->reverse) switch_to(s, t, t + 1 < s->nthreads ? t + 1 : 0);
    else switch_to(s, t, t > 0 ? t - 1 : s->nthreads - 1);
}

void wave_barrier(EmuGroup* g, int t) {
    Sched* s = (Sched*)g->sched;
    if (!s->reverse) switch_to(s, t, (t & 63) != 63 ? t + 1 : t - 63);
    else switch_to(s, t, (t & 63) != 0 ? t - 1 : t + 63);
}

[[noreturn]] void fiber_main() {
    Sched* s = tl_sched;
    const int t = s->cur;
    EmuCtx c{s->grp, t, s->grp->smem};
    (*s->body)(c);
    s->done[t] = 1;
    // finished: the next thread resumes; the last one returns to the launcher
    const int next = s->reverse ? t - 1 : t + 1;
    if (next >= 0 && next < s->nthreads) {
        switch_to(s, t, next);
    } else {
        bafq_emu_switch(&s->sp[t], s->main_sp);
    }
    std::abort();
}

// A fresh context that starts in fiber_main on the given stack.
void* make_context(char* stack, size_t size) {
    uintptr_t top = ((uintptr_t)(stack + size)) & ~(uintptr_t)15;
    uint64_t* p = (uint64_t*)top;
    *--p = 0;                                  // fake return address of fiber_main
    *--p = (uint64_t)(uintptr_t)&fiber_main;   // consumed by `ret`
    for (int i = 0; i < 6; ++i) *--p = 0;      // rbp rbx r12 r13 r14 r15
    --p;
    uint32_t* ctl = (uint32_t*)p;
    ctl[0] = 0x1F80;                           // default MXCSR
    ctl[1] = 0x037F;                           // default x87 control word
    return (void*)p;
}

void run_one_group(int g, int grid, int block, size_t smem, const std::function<void(EmuCtx&)>& body,
                   std::vector<std::unique_ptr<char[]>>& stacks) {
    const int nw = (block + 63) / 64;
    std::vector<char> lds(smem + 64, 0);
    EmuGroup grp;
    grp.nthreads = block;
    grp.wbuf.assign((size_t)nw * 3 * 64, 0.0);
    grp.smem = lds.data();
    grp.block = g;
    grp.nblocks = grid;
    grp.group_barrier = group_barrier;
    grp.wave_barrier = wave_barrier;
    Sched s;
    s.nthreads = block;
    s.body = &body;
    s.grp = &grp;
    s.sp.resize(block);
    s.done.assign(block, 0);
    while ((int)stacks.size() < block) stacks.emplace_back(new char[kFiberStack]);
    for (int t = 0; t < block; ++t) s.sp[t] = make_context(stacks[t].get(), kFiberStack);
    grp.sched = &s;
    const char* rev = std::getenv("BAFQ_EMU_REVERSE");
    s.reverse = rev && rev[0] == '1';
    tl_sched = &s;
    s.cur = s.reverse ? block - 1 : 0;
    bafq_emu_switch(&s.main_sp, s.sp[s.cur]);
    tl_sched = nullptr;
    for (int t = 0; t < block && !s.mismatch; ++t)
        if (!s.done[t]) s.mismatch = true;
    if (s.mismatch) g_barrier_mismatch = true;
}

}  // namespace

// min_threads: OS threads the launch needs at least (2 for a paired launch, whose
// workgroups wait for their partners).
template <class F>
static void run_groups(int grid, int block, size_t smem, F&& fn, int min_threads = 1) {
    const std::function<void(EmuCtx&)> body = fn;
    int nt = std::max(1, std::min<int>(grid, (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()))));
    nt = std::max(nt, std::min(grid, min_threads));
    std::atomic<int> next{0};
    auto worker = [&]() {
        std::vector<std::unique_ptr<char[]>> stacks;
        for (int g = next++; g < grid; g = next++) run_one_group(g, grid, block, smem, body, stacks);
    };
    if (nt == 1) {
        worker();
        return;
    }
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& x : th) x.join();
}

int rt_timer_start(void** t) {
    if (!*t) *t = new std::chrono::steady_clock::time_point();
    *(std::chrono::steady_clock::time_point*)*t = std::chrono::steady_clock::now();
    return 0;
}
int rt_timer_stop(void* t, double* ms) {
    const auto d = std::chrono::steady_clock::now() - *(std::chrono::steady_clock::time_point*)t;
    *ms = std::chrono::duration<double, std::milli>(d).count();
    return 0;
}
void rt_timer_free(void* t) { delete (std::chrono::steady_clock::time_point*)t; }

int rt_malloc(void** p, size_t bytes) {
    *p = std::calloc(1, bytes ? bytes : 1);
    if (!*p) { g_err = "out of memory"; return 1; }
    return 0;
}
void rt_free(void* p) { std::free(p); }
int rt_h2d(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); return 0; }
int rt_d2h(void* dst, const void* src, size_t bytes) { std::memcpy(dst, src, bytes); return 0; }
int rt_memset(void* dst, int v, size_t bytes) { std::memset(dst, v, bytes); return 0; }
int rt_d2d(void* dst, const void* src, size_t bytes) { std::memmove(dst, src, bytes); return 0; }
int rt_sync() { return 0; }
const char* rt_error() { return g_err.c_str(); }

// Workgroup sizes default to the GPU's (512 threads on both paths);
// BAFQ_BLOCK / BAFQ_LDS_BLOCK select others in tests (the host checks that
// the waves cover every MFMA tile).
int rt_block_threads() {
    const char* s = std::getenv("BAFQ_BLOCK");
    int b = s ? std::atoi(s) : 512;
    if (b < 64 || b > 1024 || (b % 64) != 0) b = 512;
    return b;
}

double rt_clock_hz() { return 1.0e9; }

int rt_lds_block_threads() {
    const char* s = std::getenv("BAFQ_LDS_BLOCK");
    int b = s ? std::atoi(s) : 512;
    if (b < 64 || b > 512 || (b % 64) != 0) b = 512;
    return b;
}

size_t rt_lds_limit() { return 160 * 1024; }

int rt_launch_warmup(void**, const LaunchArgs& a, int grid, int block, double* ms) {
    if (ms) *ms = 0.0;
    // one thread per chain; no workgroup cooperation
    for (int g = 0; g < grid; ++g)
        for (int t = 0; t < block; ++t) {
            EmuGroup grp;
            grp.nthreads = block;
            grp.block = g;
            grp.nblocks = grid;
            grp.smem = nullptr;
            grp.group_barrier = nullptr;
            grp.wave_barrier = nullptr;
            grp.sched = nullptr;
            EmuCtx c{&grp, t, nullptr};
            body_warmup(c, a);
        }
    return 0;
}

// A small persistent grid, so that every workgroup takes several chains from
// the queue as on the GPU.
int rt_resident_groups(KernelId, PathId, int, size_t) { return 3; }

void rt_release_launches(void*) {}

int rt_launch_paths(void**, KernelId k, const PathLaunch* pl, int n, double* ms) {
    if (ms) *ms = 0.0;
    g_barrier_mismatch = false;
    for (int i = 0; i < n; ++i) {
        const PathLaunch& l = pl[i];
        *l.a.queue = 0;
        if (l.a.pairs)
            for (int j = 0; j < 2 * l.a.npairs; ++j) l.a.pair_flags[j] = 0;
        run_groups(l.grid, l.block, l.smem, [&](EmuCtx& c0) {
            EmuCtx c{c0.g, c0.t, c0.smem + kSlotBytes};
            int* slot = (int*)c0.smem;
            if (k == K_PREPARE) {
                if (l.path == P_LDS) body_prepare<true>(c, l.a, slot);
                else body_prepare<false>(c, l.a, slot);
            } else if (k == K_GRAD) {
                if (l.path == P_LDS) body_grad<true>(c, l.a, slot);
                else body_grad<false>(c, l.a, slot);
            } else if (k == K_MEASURE) {
                if (l.path == P_LDS) body_measure<true>(c, l.a, slot);
                else body_measure<false>(c, l.a, slot);
            } else if (k == K_SWEEP_TAU) {
                if (l.path == P_LDS) body_sweep<true, true>(c, l.a, slot);
                else body_sweep<false, true>(c, l.a, slot);
            } else {
                if (l.path == P_LDS) body_sweep<true>(c, l.a, slot);
                else body_sweep<false>(c, l.a, slot);
            }
        }, l.a.pairs ? 2 : 1);
    }
    if (g_barrier_mismatch) {
        g_err = "emulator: the threads of a workgroup passed different numbers of barriers";
        return 1;
    }
    return 0;
}

// ------------------------------------------------------------------ learned global moves
int rt_gm_to_net(const double* phi, float* x, const GmArgs& a) {
    const size_t n = (size_t)a.batch * 2 * a.s.vol();
    for (size_t e = 0; e < n; ++e) gm_to_net_elem(phi, x, a, (int)e);
    return 0;
}
int rt_gm_propose(double* phi, double* old, double* f0, const float* y, const GmArgs& a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = 0; e < n; ++e) gm_propose_elem(phi, old, f0, y, a, (int)e);
    return 0;
}
int rt_gm_sums(const double* phi, const double* old, const double* f0, const float* y, const GmArgs& a, double* out) {
    for (int b = 0; b < a.batch; ++b) {
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        gm_sums(phi, old, f0, y, a, b, 0, gm_per(a.s), 1, acc);
        for (int k = 0; k < 5; ++k) out[(size_t)(a.c0 + b) * 5 + k] = acc[k];
    }
    return 0;
}
int rt_gm_hmc_init(const double* phi, double* old, double* p0, double* p, const GmArgs& a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = 0; e < n; ++e) gm_hmc_init_elem(phi, old, p0, p, a, (int)e);
    return 0;
}
int rt_gm_hmc_kick(double* p, const float* y, const GmArgs& a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = 0; e < n; ++e) gm_hmc_kick_elem(p, y, a, (int)e);
    return 0;
}
int rt_gm_hmc_rotate(double* phi, double* p, const GmArgs& a) {
    const size_t n = (size_t)a.batch * gm_per(a.s);
    for (size_t e = 0; e < n; ++e) gm_hmc_rotate_elem(phi, p, a, (int)e);
    return 0;
}
int rt_gm_hmc_sums(const double* phi, const double* old, const double* p0, const double* p, const GmArgs& a,
                   double* out) {
    for (int b = 0; b < a.batch; ++b) {
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        gm_hmc_sums(phi, old, p0, p, a, b, 0, gm_per(a.s), 1, acc);
        for (int k = 0; k < 5; ++k) out[(size_t)(a.c0 + b) * 5 + k] = acc[k];
    }
    return 0;
}
int rt_gm_restore(double* phi, const double* old, const int* accept, int per, size_t n) {
    for (size_t e = 0; e < n; ++e) gm_restore_elem(phi, old, accept, per, (int)e);
    return 0;
}
int rt_gm_copy_accepted(uint64_t* dst, const uint64_t* src, const int* accept, size_t per, size_t n) {
    for (size_t e = 0; e < n; ++e)
        if (accept[e / per]) dst[e] = src[e];
    return 0;
}
int rt_gm_restore_words(uint64_t* dst, const uint64_t* src, const int* accept, size_t per, size_t n) {
    for (size_t e = 0; e < n; ++e)
        if (!accept[e / per]) dst[e] = src[e];
    return 0;
}

// ------------------------------------------------------------------ network
// The matrix-core convolution (conv.hpp) on the emulator when the layer fits
// its plan; BAFQ_CONV=elem selects the per-element definition.
static int conv_mfma_run(ConvArgs a, int batch) {
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
    run_groups(batch * a.tblocks, kConvWaves * 64, conv_lds_bytes(a.tt, a.cinp, a.coutp, plane), [&](EmuCtx& c) {
        if (nb == 1) conv_mfma_body<1>(c, a);
        else if (nb == 2) conv_mfma_body<2>(c, a);
        else if (nb == 3) conv_mfma_body<3>(c, a);
        else conv_mfma_body<4>(c, a);
    });
    return 1;
}

int rt_conv_fwd(const float* in, int cin, const float* W, const float* b, float* out, int cout, int batch,
                NetShape s, int relu) {
    ConvArgs ca{};
    ca.in = in; ca.W = W; ca.bias = b; ca.mask = nullptr; ca.out = out;
    ca.cin = cin; ca.cout = cout; ca.relu = relu; ca.flip = 0; ca.s = s;
    if (conv_mfma_run(ca, batch)) return 0;
    const int n = batch * cout * s.vol();
    for (int e = 0; e < n; ++e) {
        const float y = conv_fwd_elem(in, cin, W, b, cout, s, e);
        out[e] = relu && y < 0.0f ? 0.0f : y;
    }
    return 0;
}
int rt_conv_bwd_data(const float* gout, int cout, const float* W, float* gin, int cin, const float* act, int mask,
                     int batch, NetShape s) {
    ConvArgs ca{};
    ca.in = gout; ca.W = W; ca.bias = nullptr; ca.mask = mask ? act : nullptr; ca.out = gin;
    ca.cin = cout; ca.cout = cin; ca.relu = 0; ca.flip = 1; ca.s = s;
    if (conv_mfma_run(ca, batch)) return 0;
    const int n = batch * cin * s.vol();
    for (int e = 0; e < n; ++e) gin[e] = conv_bwd_data_elem(gout, cout, W, cin, act, mask != 0, s, e);
    return 0;
}
int rt_conv_bwd_weight(const float* gout, int cout, const float* in, int cin, float* gW, float* gb, int batch,
                       NetShape s) {
    for (int e = 0; e < cout * cin * 27; ++e) gW[e] = conv_bwd_weight_elem(gout, cout, in, cin, 0, batch, s, e);
    for (int co = 0; co < cout; ++co) gb[co] = conv_bwd_bias_elem(gout, cout, batch, s, co);
    return 0;
}
int rt_mse_grad(const float* y, const float* t, float* g, size_t n, float scale) {
    for (size_t e = 0; e < n; ++e) g[e] = scale * (y[e] - t[e]);
    return 0;
}
int rt_adam(float* p, const float* g, float* m, float* v, float lr, int step, size_t n) {
    for (size_t e = 0; e < n; ++e) adam_elem(p, g, m, v, lr, step, (int)e);
    return 0;
}

// The emulator's "device" is host memory: the scratch is a static host array.
alignas(256) static float scratch[kNetScratchFloats];
int rt_net_scratch(float** ptr) {
    *ptr = scratch;
    return 0;
}
int rt_launch_status() { return 0; }

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

int rt_debug_mfma32(float* out) {
    for (int p = 0; p < 64; ++p)
        run_groups(1, 64, 0, [&](EmuCtx& c) {
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = c.mfma32(c.lane() == p ? 1.0f : 0.0f, 1000.0f + c.lane(), acc);
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

int rt_debug_qrp(int n, double* a, double* tau, int32_t* perm, double* d, int variant) {
    if (variant == 1 && (n > kColMaxN || 4 * n > rt_lds_block_threads())) { g_err = "n too large for the column QR"; return 1; }
    const int block = variant == 1 ? rt_lds_block_threads() : rt_block_threads();
    run_groups(1, block, lds_bytes(n, block / 64, variant == 1), [&](EmuCtx& ctx) {
        const Lds L = carve(ctx.smem, n, ctx.nwaves(), variant == 1);
        if (variant == 1) {
            for (int e = ctx.tid(); e < n * n; e += ctx.nthreads()) L.mat[e] = ((cd*)a)[e];
            ctx.sync();
            qrp_cols(ctx, L.mat, n, L.qr, L.dtmp);
            for (int e = ctx.tid(); e < n * n; e += ctx.nthreads()) ((cd*)a)[e] = L.mat[e];
        } else {
            qrp(ctx, (cd*)a, n, L.qr, L.dtmp);
        }
        for (int k = ctx.tid(); k < n; k += ctx.nthreads()) {
            ((cd*)tau)[k] = L.qr.tau[k];
            perm[k] = L.qr.perm[k];
            d[k] = L.dtmp[k];
        }
    });
    return 0;
}

}  // namespace bafq
