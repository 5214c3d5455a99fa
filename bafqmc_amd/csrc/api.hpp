// Host side of the C ABI (include/bafqmc.h), shared by the HIP build and the
// CPU emulator build.  A back end provides the rt_* functions below.
#pragma once

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bafqmc.h"
#include "sweep.hpp"
#include "net.hpp"
#include "gmove.hpp"
#include "conv.hpp"

namespace bafq {

enum KernelId { K_PREPARE = 0, K_SWEEP = 1, K_WARMUP = 2, K_GRAD = 3, K_MEASURE = 4, K_SWEEP_TAU = 5 };
enum PathId { P_GLOBAL = 0, P_LDS = 1 };

// back-end hooks (0 on success)
int rt_malloc(void** p, size_t bytes);
void rt_free(void* p);
int rt_h2d(void* dst, const void* src, size_t bytes);
int rt_d2h(void* dst, const void* src, size_t bytes);
int rt_memset(void* dst, int v, size_t bytes);
int rt_d2d(void* dst, const void* src, size_t bytes);
// One persistent launch per path: chains a.ids[0..a.count) taken from the
// queue a.queue (zeroed by the back end before the kernel) by `grid`
// workgroups.  rt_launch_paths launches the paths of a batch together; the
// launch functions keep per-handle back-end state (timing events, HIP graphs)
// in *cache, released by rt_release_launches.
struct PathLaunch {
    PathId path;
    LaunchArgs a;
    int grid, block;
    size_t smem;
};
int rt_launch_paths(void** cache, KernelId k, const PathLaunch* pl, int n, double* ms);
void rt_release_launches(void* cache);
int rt_launch_warmup(void** cache, const LaunchArgs& a, int grid, int block, double* ms);
// workgroups of a kernel that fit on the device at once (<= 0 on error)
int rt_resident_groups(KernelId k, PathId path, int block, size_t smem);

double rt_clock_hz();          // rate of Ctx::clock()
int rt_sync();
const char* rt_error();
int rt_block_threads();        // threads per workgroup for the global-path sweep kernels
int rt_lds_block_threads();    // threads per workgroup for the LDS-path sweep kernels
size_t rt_lds_limit();         // LDS bytes available to one workgroup
int rt_debug_gemm(int n, const double* a, int op_a, const double* b, int op_b, double* c);
int rt_debug_qrp(int n, double* a, double* tau, int32_t* perm, double* d, int variant);
int rt_debug_mfma(double* out);
int rt_debug_mfma32(float* out);
// device-time stopwatch on the default stream (HIP events); *t created on first use
int rt_timer_start(void** t);
int rt_timer_stop(void* t, double* ms);
void rt_timer_free(void* t);
int rt_debug_lanes(const double* in, const int32_t* idx, double* out);
// periodic 3D convolution network (net.hpp); device pointers
int rt_conv_fwd(const float* in, int cin, const float* W, const float* b, float* out, int cout, int batch,
                NetShape s, int relu);
int rt_conv_bwd_data(const float* gout, int cout, const float* W, float* gin, int cin, const float* act, int mask,
                     int batch, NetShape s);
int rt_conv_bwd_weight(const float* gout, int cout, const float* in, int cin, float* gW, float* gb, int batch,
                       NetShape s);
int rt_mse_grad(const float* y, const float* t, float* g, size_t n, float scale);
int rt_adam(float* p, const float* g, float* m, float* v, float lr, int step, size_t n);
// scratch of bafq_net_forward_noalloc: kNetScratchFloats floats of device memory
// owned by the library (a static device array, one per GPU), never allocated
constexpr size_t kNetScratchFloats = size_t(16) << 20;     // 64 MiB
constexpr int kMaxScratchDevices = 64;
int rt_net_scratch(float** ptr);
int rt_launch_status();    // error of the most recent kernel launches on this thread
// learned global moves (gmove.hpp); device pointers, y may be null (no network)
int rt_gm_to_net(const double* phi, float* x, const GmArgs& a);
int rt_gm_propose(double* phi, double* old, double* f0, const float* y, const GmArgs& a);
int rt_gm_sums(const double* phi, const double* old, const double* f0, const float* y, const GmArgs& a, double* out);
int rt_gm_restore(double* phi, const double* old, const int* accept, int per, size_t n);
// Hamiltonian moves (gmove.hpp gm_hmc_*)
int rt_gm_hmc_init(const double* phi, double* old, double* p0, double* p, const GmArgs& a);
int rt_gm_hmc_kick(double* p, const float* y, const GmArgs& a);
int rt_gm_hmc_rotate(double* phi, double* p, const GmArgs& a);
int rt_gm_hmc_sums(const double* phi, const double* old, const double* p0, const double* p, const GmArgs& a,
                   double* out);
// words [c * per, (c + 1) * per) of dst <- src for chains c with accept[c] == 0
int rt_gm_restore_words(uint64_t* dst, const uint64_t* src, const int* accept, size_t per, size_t n);
// words [c * per, (c + 1) * per) of dst <- src for chains c with accept[c] != 0
int rt_gm_copy_accepted(uint64_t* dst, const uint64_t* src, const int* accept, size_t per, size_t n);

constexpr double kZero = 1.0e-10;   // src/number_conserving/src/calc_basic.f90:5

}  // namespace bafq

struct bafq_handle {
    std::string err;
    std::vector<bafq::ModelDev> models;       // host copies with device pointers
    std::vector<bafq::ChainDev> chains;
    std::vector<int> chain_n, chain_ns, chain_lt;
    int obs_stride = 0;
    void* dev = nullptr;                      // one allocation for everything
    size_t dev_bytes = 0;
    bafq::ModelDev* d_models = nullptr;
    bafq::ChainDev* d_chains = nullptr;
    int* d_ids = nullptr;
    int* d_seeds = nullptr;
    double* d_phi = nullptr;
    double* d_obs = nullptr;
    double* d_stats = nullptr;
    size_t phi_total = 0;
    std::vector<size_t> phi_off;
    size_t max_lds[2] = {0, 0};
    int* d_path_ids[2] = {nullptr, nullptr};
    int path_count[2] = {0, 0};
    int* d_queue = nullptr;                   // one work counter per path
    int grid[2][2] = {{0, 0}, {0, 0}};        // persistent grid per [kernel][path]
    // plain sweeps of the device-memory path with two workgroups per chain:
    // npairs pairs (0: one workgroup per chain); d_pairs holds their barrier
    // words (2 per pair) and mailboxes (1 per pair)
    int npairs = 0;
    int* d_pairs = nullptr;
    double pair_wait_s = 20.0;                // longest wait at a pair barrier (BAFQ_PAIR_WAIT_MS)
    std::vector<bafq_schedule> schedule;      // in effect, per model
    void* launches = nullptr;                 // back-end launch cache
    // learned global moves: old fields, forward drift, per-chain sums, flags, subset ids
    double *gm_old = nullptr, *gm_f0 = nullptr, *gm_out = nullptr, *gm_stats = nullptr;
    double* gm_p = nullptr;                   // Hamiltonian moves: the momenta
    // Hamiltonian moves: network output at each chain's current fields and at
    // the proposal ([chain][2][Ltrot Ly Lx] floats); valid for one network
    // version while no sweep or other move has changed the fields
    float *gm_fcur = nullptr, *gm_fnew = nullptr;
    const void* gm_fnet = nullptr;
    uint64_t gm_fver = 0;
    int gm_fvalid = 0;
    int gm_last = 0;                          // last move: 1 learned drift move, 2 Hamiltonian move
    uint64_t* gm_state = nullptr;             // copy of every chain's matrices before the move
    double* d_grad = nullptr;                 // d log W / d phi, laid out like the fields
    int* gm_accept = nullptr;
    size_t chain_block_bytes = 0;             // matrix block of chain 0 (with one model: of every chain)
    double last_ms = 0.0;
    bafq::cd* tau_ws = nullptr;               // time-displaced workspaces (+ records), on first use
    bafq::cd* tau_rec = nullptr;
};

namespace bafq {

inline int fail(bafq_handle* h, const std::string& msg) {
    if (h) h->err = msg;
    return 1;
}

struct Bump {
    size_t off = 0;
    size_t take(size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    }
};

// Lattice helpers: x fastest, forward bonds (x+1,y), (x,y+1), (x-1,y+1)
// (src/number_conserving/src/lattice.f90:36-48,95-106); K = (4pi/3, 0) phase
// table exp(i K . a_imj) (lattice.f90:76-91, obser_equal.f90:123-125,150-164).
inline void lattice_tables(int lx, int ly, std::vector<int>& bonds, std::vector<cd>& phase, int& has_k) {
    const int ns = lx * ly;
    auto inv = [&](int x, int y) { return (((y % ly) + ly) % ly) * lx + (((x % lx) + lx) % lx); };
    bonds.assign(3 * ns, 0);
    for (int y = 0; y < ly; ++y)
        for (int x = 0; x < lx; ++x) {
            const int i = inv(x, y);
            bonds[3 * i + 0] = inv(x + 1, y);
            bonds[3 * i + 1] = inv(x, y + 1);
            bonds[3 * i + 2] = inv(x - 1, y + 1);
        }
    phase.assign(ns, cd{0.0, 0.0});
    has_k = (lx % 3 == 0 && ly % 3 == 0) ? 1 : 0;
    if (!has_k) return;
    const double pi = std::acos(-1.0);
    const int kx = 2 * lx / 3, ky = ly / 3;
    const double b1x = 2.0 * pi, b1y = -2.0 * pi / std::sqrt(3.0);
    const double b2x = 0.0, b2y = 4.0 * pi / std::sqrt(3.0);
    const double kvx = kx * b1x / lx + ky * b2x / ly, kvy = kx * b1y / lx + ky * b2y / ly;
    for (int r = 0; r < ns; ++r) {
        const int x = r % lx, y = r / lx;
        const double ax = x * 1.0 + y * 0.5, ay = y * std::sqrt(3.0) / 2.0;
        const double kr = kvx * ax + kvy * ay;
        phase[r] = cd{std::cos(kr), std::sin(kr)};
    }
}

}  // namespace bafq

extern "C" {

int bafq_create(const bafq_model_desc* models, int n_models, const int32_t* chain_model, int n_chains,
                bafq_handle** out) {
    return bafq_create_scheduled(models, n_models, chain_model, n_chains, nullptr, out);
}

int bafq_create_scheduled(const bafq_model_desc* models, int n_models, const int32_t* chain_model, int n_chains,
                          const bafq_schedule* schedule, bafq_handle** out) {
    using namespace bafq;
    *out = nullptr;
    bafq_handle* h = new bafq_handle();
    if (n_models <= 0 || n_chains <= 0) {
        h->err = "n_models and n_chains must be positive";
        *out = h;
        return 1;
    }
    // ---- validate models (calc_basic.f90:61-62,100-104)
    std::vector<std::vector<int>> bonds(n_models);
    std::vector<std::vector<cd>> phases(n_models);
    h->models.resize(n_models);
    for (int k = 0; k < n_models; ++k) {
        const bafq_model_desc& d = models[k];
        ModelDev& m = h->models[k];
        std::memset(&m, 0, sizeof(m));
        if (d.kind != BAFQ_NUMBER_CONSERVING && d.kind != BAFQ_PAIRING) { *out = h; return fail(h, "unknown solver kind"); }
        if (d.nlx < 1 || d.nly < 1 || d.ltrot < 1 || d.nwrap < 1) { *out = h; return fail(h, "lattice sizes must be positive"); }
        if (d.ltrot % d.nwrap != 0) { *out = h; return fail(h, "Ltrot is not a multiple of Nwrap"); }
        if (d.u1 < -kZero) { *out = h; return fail(h, "U1 must be non-negative (relative-density channel)"); }
        if (d.u2 > kZero) { *out = h; return fail(h, "U2 must be non-positive (total-density channel)"); }
        if (!(d.beta > 0.0)) { *out = h; return fail(h, "Beta must be positive"); }
        if (!d.expk_p || !d.expk_m) { *out = h; return fail(h, "kinetic propagators missing"); }
        m.kind = d.kind;
        m.ns = d.nlx * d.nly;
        m.n = d.kind == BAFQ_PAIRING ? 2 * m.ns : m.ns;
        m.nlx = d.nlx; m.nly = d.nly; m.ltrot = d.ltrot; m.nwrap = d.nwrap; m.nst = d.ltrot / d.nwrap;
        m.u1 = d.u1; m.u2 = d.u2; m.mu = d.mu; m.rdelta = d.kind == BAFQ_PAIRING ? d.rdelta : 0.0;
        m.shift_loc = d.shift_loc;
        const double dtau = d.beta / d.ltrot;
        // alpha = sqrt(-2 U dtau): operator_Hubbard.f90:36-43 (NC), :62-88 (pairing)
        m.alpha[0] = d.u1 > kZero ? cd{0.0, std::sqrt(2.0 * d.u1 * dtau)} : cd{0.0, 0.0};
        m.alpha[1] = d.u2 < -kZero ? cd{std::sqrt(-2.0 * d.u2 * dtau), 0.0} : cd{0.0, 0.0};
        m.active[0] = std::fabs(d.u1) > kZero;
        m.active[1] = std::fabs(d.u2) > kZero;
        lattice_tables(d.nlx, d.nly, bonds[k], phases[k], m.has_k);
        // ---- schedule: path and delayed-update width.  The LDS path needs every
        // MFMA output tile of an N x N matrix covered by the waves and G, the
        // work matrix and the update buffers within the LDS of a workgroup;
        // delayed updates use the widest buffers (<= 16 columns, a multiple of
        // 4) that fit unless the schedule fixes the width.
        {
            const bafq_schedule want = schedule ? schedule[k] : bafq_schedule{-1, -1};
            if (want.path < -1 || want.path > 1 || want.delayed < -1 || want.delayed > kMaxDkGlb ||
                (want.delayed > 0 && want.delayed % 4)) {
                *out = h;
                return fail(h, "invalid schedule (path -1/0/1, delayed -1/0/4/8/12/16, or up to 32 on the "
                               "device-memory path)");
            }
            const int lds_threads = rt_lds_block_threads();
            const int lds_waves = lds_threads / 64;
            const int glb_waves = rt_block_threads() / 64;
            const int ntile = (m.n + 15) / 16;
            const size_t limit = rt_lds_limit() - kSlotBytes;
            const bool lds_fits = m.n <= kLdsMaxN && ntile * ntile <= lds_waves * kMaxTiles &&
                                  lds_bytes(m.n, lds_waves, true, 0) <= limit;
            if (want.path == 1 && !lds_fits) { *out = h; return fail(h, "schedule asks for the LDS path but N does not fit"); }
            const int path = want.path == -1 ? (lds_fits ? 1 : 0) : want.path;
            // U and W of the delayed updates live in LDS; on the device-memory
            // path a width that does not fit there uses the chain's work
            // matrices w2, w3 instead (free during the local updates; N x dk
            // entries each, so dk <= N)
            auto fits = [&](int dk) {
                return path == 1 ? lds_bytes(m.n, lds_waves, true, dk) <= limit
                                 : lds_bytes(m.n, glb_waves, false, dk, dk) <= limit;
            };
            m.dk = 0;
            m.uw_glb = 0;
            if (want.delayed > 0) {
                if (want.delayed > 16 && path == 1) {
                    *out = h;
                    return fail(h, "delayed-update widths above 16 need the device-memory path");
                }
                if (want.delayed > 16 || !fits(want.delayed)) {
                    if (path == 1 || want.delayed > m.n) {
                        *out = h;
                        return fail(h, "schedule's delayed-update width does not fit");
                    }
                    m.uw_glb = 1;
                }
                m.dk = want.delayed;
            } else if (want.delayed == -1) {
                for (int dk = 16; dk >= 4; dk -= 4)
                    if (fits(dk)) {
                        m.dk = dk;
                        break;
                    }
                if (path == 0 && m.dk < 16 && m.n >= 16) {
                    // G streams from HBM at every flush: the widest buffers, in device memory
                    // (32 columns: blocks of 32 rows for the submatrix updates)
                    m.dk = m.n >= kMaxDkGlb ? kMaxDkGlb : 16;
                    m.uw_glb = 1;
                }
            }
            // the device-memory path keeps the submatrix of each block of sites in LDS
            if (path == 0 && lds_bytes(m.n, glb_waves, false, m.uw_glb ? 0 : m.dk, m.dk) > limit) {
                *out = h;
                return fail(h, "schedule's delayed-update width does not fit in LDS");
            }
            h->schedule.push_back(bafq_schedule{path, m.dk});
        }
    }
    // ---- layout of the device allocation
    Bump b;
    const size_t off_models = b.take(sizeof(ModelDev) * n_models);
    std::vector<size_t> off_ph(n_models), off_bd(n_models);
    std::vector<size_t> off_ez(n_models);
    std::vector<std::vector<double>> ez(n_models);
    for (int k = 0; k < n_models; ++k) {
        const int n = h->models[k].n, ns = h->models[k].ns;
        // MFMA operand order of e^{-+dtau A} (left: [ti][kk][lane] = E[16ti+l%16][4kk+l/16],
        // right: [tj][kk][lane] = E[4kk+l/16][16tj+l%16]), zero padded.
        const int nt = (n + 15) / 16, nk = (n + 3) / 4;
        const size_t per = (size_t)nt * nk * 64;
        ez[k].assign(4 * per, 0.0);
        for (int which = 0; which < 4; ++which) {
            const double* E = (which == 0 || which == 2) ? models[k].expk_p : models[k].expk_m;
            const bool left = which < 2;
            double* dst = ez[k].data() + which * per;
            for (int t = 0; t < nt; ++t)
                for (int kk = 0; kk < nk; ++kk)
                    for (int l = 0; l < 64; ++l) {
                        const int i = left ? 16 * t + (l & 15) : 4 * kk + (l >> 4);
                        const int j = left ? 4 * kk + (l >> 4) : 16 * t + (l & 15);
                        dst[((size_t)t * nk + kk) * 64 + l] = (i < n && j < n) ? E[(size_t)i * n + j] : 0.0;
                    }
        }
        off_ez[k] = b.take(sizeof(double) * 4 * per);
        off_ph[k] = b.take(sizeof(cd) * ns);
        off_bd[k] = b.take(sizeof(int) * 3 * ns);
    }
    const size_t off_chains = b.take(sizeof(ChainDev) * n_chains);
    const size_t off_ids = b.take(sizeof(int) * n_chains);
    const size_t off_pids0 = b.take(sizeof(int) * n_chains);
    const size_t off_pids1 = b.take(sizeof(int) * n_chains);
    const size_t off_seeds = b.take(sizeof(int) * n_chains);
    const size_t off_queue = b.take(sizeof(int) * 2);
    int max_ns = 0;
    for (int c = 0; c < n_chains; ++c) {
        if (chain_model[c] < 0 || chain_model[c] >= n_models) { *out = h; return fail(h, "chain_model out of range"); }
        max_ns = std::max(max_ns, h->models[chain_model[c]].ns);
    }
    h->obs_stride = O_SITE0 + max_ns;
    const size_t off_obs = b.take(sizeof(double) * (size_t)h->obs_stride * n_chains);
    const size_t off_stats = b.take(sizeof(double) * (size_t)S_LEN * n_chains);
    h->phi_off.resize(n_chains);
    h->phi_total = 0;
    for (int c = 0; c < n_chains; ++c) {
        const ModelDev& m = h->models[chain_model[c]];
        h->phi_off[c] = h->phi_total;
        h->phi_total += (size_t)m.ltrot * m.ns * 2;
    }
    const size_t off_phi = b.take(sizeof(double) * h->phi_total);
    struct COff { size_t g, ur, vr, ul, vl, dr, dl, url, vrl, ull, vll, drl, dll, wa, w1, w2, w3; };
    std::vector<COff> co(n_chains);
    for (int c = 0; c < n_chains; ++c) {
        const ModelDev& m = h->models[chain_model[c]];
        const size_t mat = sizeof(cd) * (size_t)m.n * m.n, vec = sizeof(double) * m.n;
        const size_t lists = (size_t)(m.nst + 1);
        COff& o = co[c];
        o.g = b.take(mat); o.ur = b.take(mat); o.vr = b.take(mat); o.ul = b.take(mat); o.vl = b.take(mat);
        o.dr = b.take(vec); o.dl = b.take(vec);
        o.url = b.take(mat * lists); o.vrl = b.take(mat * lists);
        o.ull = b.take(mat * lists); o.vll = b.take(mat * lists);
        o.drl = b.take(vec * lists); o.dll = b.take(vec * lists);
        o.wa = b.take(mat); o.w1 = b.take(mat); o.w2 = b.take(mat); o.w3 = b.take(mat);
        if (c == 0) h->chain_block_bytes = b.off - o.g;          // the same for every chain of one model
    }
    h->dev_bytes = b.off;
    if (rt_malloc(&h->dev, h->dev_bytes)) { *out = h; return fail(h, std::string("device allocation failed: ") + rt_error()); }
    char* base = (char*)h->dev;
    if (rt_memset(h->dev, 0, h->dev_bytes)) { *out = h; return fail(h, rt_error()); }
    // ---- model tables
    for (int k = 0; k < n_models; ++k) {
        ModelDev& m = h->models[k];
        rt_h2d(base + off_ph[k], phases[k].data(), sizeof(cd) * m.ns);
        rt_h2d(base + off_bd[k], bonds[k].data(), sizeof(int) * 3 * m.ns);
        rt_h2d(base + off_ez[k], ez[k].data(), sizeof(double) * ez[k].size());
        for (int w = 0; w < 4; ++w)
            m.ez[w] = (const double*)(base + off_ez[k]) + (size_t)w * (ez[k].size() / 4);
        m.phasek = (const cd*)(base + off_ph[k]);
        m.bonds = (const int*)(base + off_bd[k]);
    }
    h->d_models = (ModelDev*)(base + off_models);
    h->d_chains = (ChainDev*)(base + off_chains);
    h->d_ids = (int*)(base + off_ids);
    h->d_seeds = (int*)(base + off_seeds);
    h->d_queue = (int*)(base + off_queue);
    h->d_obs = (double*)(base + off_obs);
    h->d_stats = (double*)(base + off_stats);
    h->d_phi = (double*)(base + off_phi);
    rt_h2d(h->d_models, h->models.data(), sizeof(ModelDev) * n_models);
    h->chains.resize(n_chains);
    std::vector<int> ids(n_chains), pids[2];
    const int nw_glb = rt_block_threads() / 64;
    for (int c = 0; c < n_chains; ++c) {
        ChainDev& ch = h->chains[c];
        const COff& o = co[c];
        const ModelDev& m = h->models[chain_model[c]];
        ch.model = chain_model[c];
        ch.seed = h->d_seeds + c;
        ch.phi = h->d_phi + h->phi_off[c];
        ch.g = (cd*)(base + o.g); ch.ur = (cd*)(base + o.ur); ch.vr = (cd*)(base + o.vr);
        ch.ul = (cd*)(base + o.ul); ch.vl = (cd*)(base + o.vl);
        ch.dr = (double*)(base + o.dr); ch.dl = (double*)(base + o.dl);
        ch.urlist = (cd*)(base + o.url); ch.vrlist = (cd*)(base + o.vrl);
        ch.ullist = (cd*)(base + o.ull); ch.vllist = (cd*)(base + o.vll);
        ch.drlist = (double*)(base + o.drl); ch.dllist = (double*)(base + o.dll);
        ch.wa = (cd*)(base + o.wa); ch.w1 = (cd*)(base + o.w1); ch.w2 = (cd*)(base + o.w2); ch.w3 = (cd*)(base + o.w3);
        ch.obs = h->d_obs + (size_t)h->obs_stride * c;
        ch.stats = h->d_stats + (size_t)S_LEN * c;
        ch.tau = nullptr;
        ch.tau_rec = nullptr;
        ids[c] = c;
        h->chain_n.push_back(m.n);
        h->chain_ns.push_back(m.ns);
        h->chain_lt.push_back(m.ltrot);
        const int path = h->schedule[chain_model[c]].path;
        pids[path].push_back(c);
        const size_t lb = path == P_LDS ? lds_bytes(m.n, rt_lds_block_threads() / 64, true, m.dk)
                                        : lds_bytes(m.n, nw_glb, false, m.uw_glb ? 0 : m.dk, m.dk);
        h->max_lds[path] = std::max(h->max_lds[path], lb + kSlotBytes);
    }
    rt_h2d(h->d_chains, h->chains.data(), sizeof(ChainDev) * n_chains);
    rt_h2d(h->d_ids, ids.data(), sizeof(int) * n_chains);
    h->d_path_ids[P_GLOBAL] = (int*)(base + off_pids0);
    h->d_path_ids[P_LDS] = (int*)(base + off_pids1);
    for (int p = 0; p < 2; ++p) {
        // longest processing time first: a sweep costs ~ N^3 Ltrot
        std::stable_sort(pids[p].begin(), pids[p].end(), [&](int x, int y) {
            const double cx = (double)h->chain_n[x] * h->chain_n[x] * h->chain_n[x] * h->chain_lt[x];
            const double cy = (double)h->chain_n[y] * h->chain_n[y] * h->chain_n[y] * h->chain_lt[y];
            return cx > cy;
        });
        h->path_count[p] = (int)pids[p].size();
        if (!pids[p].empty()) rt_h2d(h->d_path_ids[p], pids[p].data(), sizeof(int) * pids[p].size());
        for (int k = 0; k < 2; ++k) {
            if (pids[p].empty()) continue;
            const int block = p == P_LDS ? rt_lds_block_threads() : rt_block_threads();
            const int res = rt_resident_groups((KernelId)k, (PathId)p, block, h->max_lds[p]);
            if (res <= 0) { *out = h; return fail(h, rt_error()); }
            h->grid[k][p] = std::min(res, h->path_count[p]);
        }
    }
    // Two workgroups per chain on the device-memory path when every chain then
    // still has its pair resident at once (or BAFQ_PAIRS=1: as many pairs as
    // fit) and every model on that path uses submatrix updates.  LDS-path
    // chains of the batch run in a kernel beside the paired ones: there the
    // two kernels must have the same resident capacity, and the LDS-path grid
    // is cut to what the pairs leave free, so every workgroup of both fits at
    // once.  BAFQ_PAIRS=0 turns it off; BAFQ_PAIR_WAIT_MS sets the longest
    // wait at a pair barrier (20000).
    {
        if (const char* w = std::getenv("BAFQ_PAIR_WAIT_MS")) h->pair_wait_s = std::max(0.0, std::atof(w) / 1000.0);
        const char* env = std::getenv("BAFQ_PAIRS");
        const int mode = env ? std::atoi(env) : -1;
        const bool mixed = h->path_count[P_LDS] > 0;
        bool ok = mode != 0 && h->path_count[P_GLOBAL] > 0;
        for (int c = 0; ok && c < n_chains; ++c)
            if (h->schedule[chain_model[c]].path == P_GLOBAL) ok = h->models[chain_model[c]].dk > 0;
        if (ok) {
            // the paired kernels: sweep, rebuild, measurement pass
            auto res_of = [&](PathId p) {
                const int block = p == P_LDS ? rt_lds_block_threads() : rt_block_threads();
                int r = rt_resident_groups(K_SWEEP, p, block, h->max_lds[p]);
                r = std::min(r, rt_resident_groups(K_PREPARE, p, block, h->max_lds[p]));
                return std::min(r, rt_resident_groups(K_MEASURE, p, block, h->max_lds[p]));
            };
            const int res = res_of(P_GLOBAL);
            const int cnt = h->path_count[P_GLOBAL];
            if (!mixed) {
                if (res >= 2 && (2 * cnt <= res || mode == 1)) h->npairs = std::min(cnt, res / 2);
            } else if (res_of(P_LDS) == res && 2 * cnt < res) {
                h->npairs = cnt;
                for (int k = 0; k < 2; ++k) h->grid[k][P_LDS] = std::min(h->grid[k][P_LDS], res - 2 * cnt);
            }
        }
        if (h->npairs > 0 && rt_malloc((void**)&h->d_pairs, sizeof(int) * 3 * (size_t)h->npairs)) {
            *out = h;
            return fail(h, rt_error());
        }
    }
    if (rt_sync()) { *out = h; return fail(h, rt_error()); }
    *out = h;
    return 0;
}

void bafq_destroy(bafq_handle* h) {
    if (!h) return;
    if (h->gm_old) bafq::rt_free(h->gm_old);
    if (h->gm_p) bafq::rt_free(h->gm_p);
    if (h->gm_fcur) bafq::rt_free(h->gm_fcur);
    if (h->gm_state) bafq::rt_free(h->gm_state);
    if (h->d_grad) bafq::rt_free(h->d_grad);
    if (h->tau_ws) bafq::rt_free(h->tau_ws);
    if (h->d_pairs) bafq::rt_free(h->d_pairs);
    bafq::rt_release_launches(h->launches);
    if (h->dev) bafq::rt_free(h->dev);
    delete h;
}

int bafq_get_schedule(const bafq_handle* h, int model, bafq_schedule* out) {
    if (!h || !out || model < 0 || model >= (int)h->schedule.size()) return 1;
    *out = h->schedule[model];
    return 0;
}

int bafq_pairs(const bafq_handle* h) { return h ? h->npairs : 0; }

const char* bafq_last_error(const bafq_handle* h) { return h ? h->err.c_str() : "null handle"; }

int bafq_dims(const bafq_handle* h, int chain, int32_t* n, int32_t* obs_stride, int32_t* stats_stride,
              int32_t* n_fields) {
    if (!h || chain < 0 || chain >= (int)h->chains.size()) return 1;
    if (n) *n = h->chain_n[chain];
    if (obs_stride) *obs_stride = h->obs_stride;
    if (stats_stride) *stats_stride = bafq::S_LEN;
    if (n_fields) *n_fields = h->chain_lt[chain] * h->chain_ns[chain] * 2;
    return 0;
}

int bafq_set_fields(bafq_handle* h, const double* phi, const int32_t* seeds) {
    using namespace bafq;
    if (!h) return 1;
    h->gm_fvalid = 0;
    if (rt_h2d(h->d_phi, phi, sizeof(double) * h->phi_total)) return fail(h, rt_error());
    for (size_t c = 0; c < h->chains.size(); ++c)
        if (seeds[c] <= 0) return fail(h, "RNG seeds must be positive");
    if (rt_h2d(h->d_seeds, seeds, sizeof(int) * h->chains.size())) return fail(h, rt_error());
    return rt_sync() ? fail(h, rt_error()) : 0;
}

int bafq_get_fields(bafq_handle* h, double* phi, int32_t* seeds) {
    using namespace bafq;
    if (!h) return 1;
    if (rt_sync()) return fail(h, rt_error());
    if (phi && rt_d2h(phi, h->d_phi, sizeof(double) * h->phi_total)) return fail(h, rt_error());
    if (seeds && rt_d2h(seeds, h->d_seeds, sizeof(int) * h->chains.size())) return fail(h, rt_error());
    return 0;
}

static int launch_all(bafq_handle* h, bafq::KernelId k, int nsweep, int measure, int tau = 0,
                      int tau_record = 0) {
    using namespace bafq;
    LaunchArgs a;
    a.models = h->d_models;
    a.chains = h->d_chains;
    a.nsweep = nsweep;
    a.measure = measure;
    a.tau = tau;
    a.tau_record = tau_record;
    a.queue = nullptr;
    a.grad = h->d_grad;
    a.phi0 = h->d_phi;
    a.pairs = 0;
    a.npairs = 0;
    a.pair_flags = nullptr;
    a.pair_mail = nullptr;
    a.pair_wait = 0;
    double total_ms = 0.0;
    const bool paired = (k == K_SWEEP || k == K_PREPARE || k == K_MEASURE) && h->npairs > 0;
    if (k == K_WARMUP) {
        a.ids = h->d_ids;
        a.count = (int)h->chains.size();
        const int block = 64, grid = (a.count + block - 1) / block;
        if (rt_launch_warmup(&h->launches, a, grid, block, &total_ms)) return fail(h, rt_error());
    } else {
        PathLaunch pl[2];
        int n = 0;
        for (int p = 0; p < 2; ++p) {
            if (h->path_count[p] == 0) continue;
            PathLaunch& l = pl[n++];
            l.path = (PathId)p;
            l.a = a;
            l.a.ids = h->d_path_ids[p];
            l.a.count = h->path_count[p];
            l.a.queue = h->d_queue + p;
            l.grid = h->grid[(k == K_GRAD || k == K_MEASURE || k == K_SWEEP_TAU) ? K_SWEEP : k][p];
            l.block = p == P_LDS ? rt_lds_block_threads() : rt_block_threads();
            l.smem = h->max_lds[p];
            if (paired && p == P_GLOBAL) {
                l.a.pairs = 1;
                l.a.npairs = h->npairs;
                l.a.pair_flags = h->d_pairs;
                l.a.pair_mail = h->d_pairs + 2 * h->npairs;
                l.a.pair_wait = (uint64_t)(h->pair_wait_s * rt_clock_hz());
                l.grid = 2 * h->npairs;
            }
        }
        if (rt_launch_paths(&h->launches, k, pl, n, &total_ms)) return fail(h, rt_error());
        if (paired) {
            std::vector<int> fl(2 * (size_t)h->npairs);
            if (rt_d2h(fl.data(), h->d_pairs, sizeof(int) * fl.size())) return fail(h, rt_error());
            for (int i = 0; i < h->npairs; ++i)
                if (fl[2 * i + 1])
                    return fail(h, "a workgroup waited longer than BAFQ_PAIR_WAIT_MS (default 20 s) for its partner "
                                   "(the pair was not resident at once); the launch's results are invalid, set "
                                   "BAFQ_PAIRS=0");
        }
    }
    h->last_ms = total_ms;
    if (rt_sync()) return fail(h, rt_error());
    return 0;
}

int bafq_warmup(bafq_handle* h, int n) {
    if (!h) return 1;
    h->gm_fvalid = 0;
    if (n <= 0) return 0;
    return launch_all(h, bafq::K_WARMUP, n, 0);
}

int bafq_prepare(bafq_handle* h) {
    if (!h) return 1;
    return launch_all(h, bafq::K_PREPARE, 0, 0);
}

int bafq_sweep(bafq_handle* h, int n_sweeps, int measure) {
    if (!h) return 1;
    h->gm_fvalid = 0;
    if (n_sweeps <= 0) return 0;
    return launch_all(h, bafq::K_SWEEP, n_sweeps, measure);
}

int bafq_sweep_tau(bafq_handle* h, int n_sweeps, int measure, int mode, int record) {
    using namespace bafq;
    if (!h) return 1;
    if (mode != TAU_BOSONIC && mode != TAU_REFERENCE)
        return fail(h, "time-displaced mode must be 1 (bosonic) or 2 (the reference's conventions)");
    h->gm_fvalid = 0;
    if (n_sweeps <= 0) return 0;
    const size_t nc = h->chains.size();
    const bool rec = record != 0;
    if (!h->tau_ws || (rec && !h->tau_rec)) {
        // one allocation per handle: the workspaces, then (if asked) the records
        std::vector<size_t> ws(nc), rs(nc);
        size_t total = 0, rtotal = 0;
        for (size_t c = 0; c < nc; ++c) {
            const int n = h->chain_n[c], lt = h->chain_lt[c];
            ws[c] = tau_ws_elems(n, lt, false);
            rs[c] = rec ? tau_ws_elems(n, lt, true) - ws[c] : 0;
            total += ws[c];
            rtotal += rs[c];
        }
        if (h->tau_ws) rt_free(h->tau_ws);
        h->tau_ws = nullptr;
        h->tau_rec = nullptr;
        if (rt_malloc((void**)&h->tau_ws, sizeof(cd) * (total + rtotal)))
            return fail(h, std::string("time-displaced workspace allocation failed: ") + rt_error());
        if (rec) h->tau_rec = h->tau_ws + total;
        size_t o = 0, ro = 0;
        for (size_t c = 0; c < nc; ++c) {
            h->chains[c].tau = h->tau_ws + o;
            h->chains[c].tau_rec = rec ? h->tau_rec + ro : nullptr;
            o += ws[c];
            ro += rs[c];
        }
        if (rt_h2d(h->d_chains, h->chains.data(), sizeof(ChainDev) * nc)) return fail(h, rt_error());
    }
    return launch_all(h, K_SWEEP_TAU, n_sweeps, measure, mode, rec ? 1 : 0);
}

int bafq_get_green_tau(bafq_handle* h, int chain, double* out) {
    using namespace bafq;
    if (!h || !out || chain < 0 || chain >= (int)h->chains.size()) return 1;
    if (!h->chains[chain].tau_rec)
        return fail(h, "no recorded time-displaced pass (bafq_sweep_tau with record = 1 first)");
    if (rt_sync()) return fail(h, rt_error());
    const size_t n = h->chain_n[chain], lt = h->chain_lt[chain];
    if (rt_d2h(out, h->chains[chain].tau_rec, sizeof(cd) * lt * 4 * n * n)) return fail(h, rt_error());
    return 0;
}

int bafq_measure_pass(bafq_handle* h, int n_passes, int mode) {
    if (!h) return 1;
    if (mode < 0 || mode > 2) return bafq::fail(h, "measurement mode must be 0, 1 or 2");
    if (n_passes <= 0) return 0;
    return launch_all(h, bafq::K_MEASURE, mode == 2 ? 1 : n_passes, mode + 1);
}

int bafq_get_stats(bafq_handle* h, double* stats) {
    using namespace bafq;
    if (!h) return 1;
    if (rt_sync()) return fail(h, rt_error());
    if (rt_d2h(stats, h->d_stats, sizeof(double) * S_LEN * h->chains.size())) return fail(h, rt_error());
    return 0;
}

int bafq_read_bin(bafq_handle* h, double* obs, double* stats) {
    using namespace bafq;
    if (!h) return 1;
    if (rt_sync()) return fail(h, rt_error());
    const size_t nc = h->chains.size();
    if (obs && rt_d2h(obs, h->d_obs, sizeof(double) * h->obs_stride * nc)) return fail(h, rt_error());
    if (stats && rt_d2h(stats, h->d_stats, sizeof(double) * S_LEN * nc)) return fail(h, rt_error());
    if (rt_memset(h->d_obs, 0, sizeof(double) * h->obs_stride * nc)) return fail(h, rt_error());
    // keep the last log|det G| (S_LOGDET) and the running drift maximum; reset counters
    std::vector<double> st(S_LEN * nc);
    if (rt_d2h(st.data(), h->d_stats, sizeof(double) * S_LEN * nc)) return fail(h, rt_error());
    for (size_t c = 0; c < nc; ++c)
        for (int k = 0; k < S_LEN; ++k)
            if (k != S_LOGDET && k != S_XMAX && !(k >= S_TAU_XMAX && k < S_TAU_XMAX + 3)) st[c * S_LEN + k] = 0.0;
    if (stats) {
        const double hz = rt_clock_hz();
        for (size_t c = 0; c < nc; ++c)
            for (int k = S_T_SITE; k < S_T_SITE + kNumTimers; ++k) stats[c * S_LEN + k] /= hz;
    }
    if (rt_h2d(h->d_stats, st.data(), sizeof(double) * S_LEN * nc)) return fail(h, rt_error());
    return rt_sync() ? fail(h, rt_error()) : 0;
}

int bafq_get_green(bafq_handle* h, int chain, double* g) {
    using namespace bafq;
    if (!h || chain < 0 || chain >= (int)h->chains.size()) return 1;
    if (rt_sync()) return fail(h, rt_error());
    const int n = h->chain_n[chain];
    if (rt_d2h(g, h->chains[chain].g, sizeof(cd) * n * n)) return fail(h, rt_error());
    return 0;
}

int bafq_last_kernel_ms(const bafq_handle* h, double* ms) {
    if (!h || !ms) return 1;
    *ms = h->last_ms;
    return 0;
}

int bafq_debug_gemm(int n, const double* a, int op_a, const double* b, int op_b, double* c) {
    return bafq::rt_debug_gemm(n, a, op_a, b, op_b, c);
}

int bafq_debug_qrp(int n, double* a, double* tau, int32_t* perm, double* d, int variant) {
    return bafq::rt_debug_qrp(n, a, tau, perm, d, variant);
}

int bafq_debug_mfma_probe(double* out) { return bafq::rt_debug_mfma(out); }
int bafq_debug_mfma32_probe(float* out) { return bafq::rt_debug_mfma32(out); }
int bafq_debug_lane_ops(const double* in, const int32_t* idx, double* out) { return bafq::rt_debug_lanes(in, idx, out); }

}  // extern "C"

// ------------------------------------------------------------------ network
struct bafq_net {
    bafq_net_desc d;
    int max_batch = 0;
    size_t nparam = 0;
    std::vector<size_t> woff, boff;      // parameter offsets per layer
    std::vector<int> cin, cout;
    float* dev = nullptr;                // one allocation
    float *params = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    std::vector<float*> act;             // act[0] input ... act[L] output
    float *g0 = nullptr, *g1 = nullptr, *target = nullptr;
    int step = 0;
    uint64_t version = 0;                // unique over all networks, renewed whenever the parameters change
    void* timer = nullptr;
    std::string err;
};

namespace bafq {
// Parameter versions, unique over every network of the process: a cached
// network output is valid only for the version that produced it.
inline uint64_t next_net_version() {
    static std::atomic<uint64_t> v{0};
    return ++v;
}
}  // namespace bafq

static int net_fail(bafq_net* h, const std::string& msg) {
    if (h) h->err = msg;
    return 1;
}

extern "C" {

int bafq_net_create(const bafq_net_desc* d, int max_batch, bafq_net** out) {
    using namespace bafq;
    *out = nullptr;
    bafq_net* h = new bafq_net();
    *out = h;
    if (!d || d->lt < 1 || d->ly < 1 || d->lx < 1 || d->channels < 1 || d->hidden < 1 || d->layers < 2 ||
        max_batch < 1)
        return net_fail(h, "invalid network description");
    h->d = *d;
    h->max_batch = max_batch;
    const int L = d->layers;
    for (int l = 0; l < L; ++l) {
        h->cin.push_back(l == 0 ? d->channels : d->hidden);
        h->cout.push_back(l == L - 1 ? d->channels : d->hidden);
        h->woff.push_back(h->nparam);
        h->nparam += (size_t)h->cout[l] * h->cin[l] * 27;
        h->boff.push_back(h->nparam);
        h->nparam += (size_t)h->cout[l];
    }
    const size_t vol = (size_t)d->lt * d->ly * d->lx;
    const int cmax = d->channels > d->hidden ? d->channels : d->hidden;
    size_t total = 4 * h->nparam + 2 * (size_t)max_batch * cmax * vol + (size_t)max_batch * d->channels * vol;
    for (int l = 0; l <= L; ++l) total += (size_t)max_batch * (l == 0 ? d->channels : h->cout[l - 1]) * vol;
    if (rt_malloc((void**)&h->dev, total * sizeof(float))) return net_fail(h, rt_error());
    if (rt_memset(h->dev, 0, total * sizeof(float))) return net_fail(h, rt_error());
    float* p = h->dev;
    auto take = [&](size_t n) { float* q = p; p += n; return q; };
    h->params = take(h->nparam);
    h->grad = take(h->nparam);
    h->m = take(h->nparam);
    h->v = take(h->nparam);
    for (int l = 0; l <= L; ++l) h->act.push_back(take((size_t)max_batch * (l == 0 ? d->channels : h->cout[l - 1]) * vol));
    h->g0 = take((size_t)max_batch * cmax * vol);
    h->g1 = take((size_t)max_batch * cmax * vol);
    h->target = take((size_t)max_batch * d->channels * vol);
    return 0;
}

void bafq_net_destroy(bafq_net* h) {
    if (!h) return;
    bafq::rt_timer_free(h->timer);
    if (h->dev) bafq::rt_free(h->dev);
    delete h;
}

const char* bafq_net_last_error(const bafq_net* h) { return h ? h->err.c_str() : "null handle"; }

int64_t bafq_net_param_count(const bafq_net* h) { return h ? (int64_t)h->nparam : -1; }

int bafq_net_set_params(bafq_net* h, const float* params) {
    using namespace bafq;
    if (!h) return 1;
    h->step = 0;
    h->version = bafq::next_net_version();
    if (rt_h2d(h->params, params, h->nparam * sizeof(float)) || rt_memset(h->m, 0, h->nparam * sizeof(float)) ||
        rt_memset(h->v, 0, h->nparam * sizeof(float)) || rt_sync())
        return net_fail(h, rt_error());
    return 0;
}

int bafq_net_get_params(bafq_net* h, float* params) {
    using namespace bafq;
    if (!h) return 1;
    if (rt_sync() || rt_d2h(params, h->params, h->nparam * sizeof(float))) return net_fail(h, rt_error());
    return 0;
}

static int net_forward_dev(bafq_net* h, int batch) {
    using namespace bafq;
    const NetShape s{h->d.lt, h->d.ly, h->d.lx};
    const int L = h->d.layers;
    for (int l = 0; l < L; ++l)
        if (rt_conv_fwd(h->act[l], h->cin[l], h->params + h->woff[l], h->params + h->boff[l], h->act[l + 1],
                        h->cout[l], batch, s, l < L - 1 ? 1 : 0))
            return net_fail(h, rt_error());
    return 0;
}

int bafq_net_forward(bafq_net* h, int batch, const float* x, float* y) {
    using namespace bafq;
    if (!h) return 1;
    if (batch < 1 || batch > h->max_batch) return net_fail(h, "batch out of range");
    const size_t n = (size_t)batch * h->d.channels * h->d.lt * h->d.ly * h->d.lx;
    if (rt_h2d(h->act[0], x, n * sizeof(float))) return net_fail(h, rt_error());
    if (net_forward_dev(h, batch)) return 1;
    if (rt_sync() || rt_d2h(y, h->act[h->d.layers], n * sizeof(float))) return net_fail(h, rt_error());
    return 0;
}

int bafq_net_forward_noalloc(const bafq_net_desc* d, const float* params, int batch, const float* x, float* y) {
    using namespace bafq;
    if (!d || !params || !x || !y || batch < 1 || d->lt < 1 || d->ly < 1 || d->lx < 1 || d->channels < 1 ||
        d->hidden < 1 || d->layers < 2)
        return BAFQ_NOALLOC_INVALID;
    const size_t vol = (size_t)d->lt * d->ly * d->lx;
    const size_t per_field = 2 * (size_t)d->hidden * vol;       // two ping-pong hidden activations
    if (per_field > kNetScratchFloats) return BAFQ_NOALLOC_TOO_LARGE;
    float* scr = nullptr;
    if (rt_net_scratch(&scr)) return BAFQ_NOALLOC_DEVICE;
    // the batch runs in chunks whose hidden activations fit in the scratch
    const int chunk = (int)std::min<size_t>((size_t)batch, kNetScratchFloats / per_field);
    float* hid[2] = {scr, scr + (size_t)chunk * d->hidden * vol};
    const NetShape s{d->lt, d->ly, d->lx};
    const int L = d->layers;
    const size_t io = (size_t)d->channels * vol;
    for (int b0 = 0; b0 < batch; b0 += chunk) {
        const int nb = std::min(chunk, batch - b0);
        const float* in = x + (size_t)b0 * io;
        size_t off = 0;                           // parameters: W[cout][cin][27], b[cout] per layer
        for (int l = 0; l < L; ++l) {
            const int cin = l == 0 ? d->channels : d->hidden;
            const int cout = l == L - 1 ? d->channels : d->hidden;
            float* out = l == L - 1 ? y + (size_t)b0 * io : hid[l & 1];
            const size_t nw = (size_t)cout * cin * 27;
            if (rt_conv_fwd(in, cin, params + off, params + off + nw, out, cout, nb, s, l < L - 1 ? 1 : 0))
                return BAFQ_NOALLOC_DEVICE;
            off += nw + cout;
            in = out;
        }
    }
    return rt_launch_status() ? BAFQ_NOALLOC_DEVICE : 0;
}

int bafq_net_time_forward(bafq_net* h, int batch, int reps, double* ms) {
    using namespace bafq;
    if (!h || !ms) return 1;
    if (batch < 1 || batch > h->max_batch || reps < 1) return net_fail(h, "batch or reps out of range");
    if (net_forward_dev(h, batch)) return 1;                  // warm (kernel attributes, caches)
    if (rt_sync() || rt_timer_start(&h->timer)) return net_fail(h, rt_error());
    for (int r = 0; r < reps; ++r)
        if (net_forward_dev(h, batch)) return 1;
    double t = 0.0;
    if (rt_timer_stop(h->timer, &t)) return net_fail(h, rt_error());
    *ms = t / reps;
    return 0;
}

// Loss, gradient in h->grad; returns the loss through *loss.
static int net_gradient_dev(bafq_net* h, int batch, double* loss) {
    using namespace bafq;
    const NetShape s{h->d.lt, h->d.ly, h->d.lx};
    const int L = h->d.layers;
    const size_t n = (size_t)batch * h->d.channels * s.vol();
    if (net_forward_dev(h, batch)) return 1;
    float* gcur = h->g0;
    float* gnext = h->g1;
    if (rt_mse_grad(h->act[L], h->target, gcur, n, 2.0f / (float)n)) return net_fail(h, rt_error());
    {
        std::vector<float> g(n);
        if (rt_sync() || rt_d2h(g.data(), gcur, n * sizeof(float))) return net_fail(h, rt_error());
        double acc = 0.0;
        for (size_t i = 0; i < n; ++i) acc += (double)g[i] * g[i];
        *loss = acc * (double)n / 4.0;     // g = 2 (y - t) / n
    }
    for (int l = L - 1; l >= 0; --l) {
        if (rt_conv_bwd_weight(gcur, h->cout[l], h->act[l], h->cin[l], h->grad + h->woff[l], h->grad + h->boff[l],
                               batch, s))
            return net_fail(h, rt_error());
        if (l > 0) {
            if (rt_conv_bwd_data(gcur, h->cout[l], h->params + h->woff[l], gnext, h->cin[l], h->act[l], 1, batch, s))
                return net_fail(h, rt_error());
            float* t = gcur; gcur = gnext; gnext = t;
        }
    }
    return 0;
}

int bafq_net_gradient(bafq_net* h, int batch, const float* x, const float* target, double* loss, float* grad) {
    using namespace bafq;
    if (!h) return 1;
    if (batch < 1 || batch > h->max_batch) return net_fail(h, "batch out of range");
    const size_t n = (size_t)batch * h->d.channels * h->d.lt * h->d.ly * h->d.lx;
    if (rt_h2d(h->act[0], x, n * sizeof(float)) || rt_h2d(h->target, target, n * sizeof(float)))
        return net_fail(h, rt_error());
    if (net_gradient_dev(h, batch, loss)) return 1;
    if (grad && (rt_sync() || rt_d2h(grad, h->grad, h->nparam * sizeof(float)))) return net_fail(h, rt_error());
    return 0;
}

int bafq_net_train(bafq_net* h, int batch, const float* x, const float* target, int steps, float lr,
                   double* loss) {
    using namespace bafq;
    if (!h) return 1;
    if (batch < 1 || batch > h->max_batch || steps < 1) return net_fail(h, "batch or steps out of range");
    const size_t n = (size_t)batch * h->d.channels * h->d.lt * h->d.ly * h->d.lx;
    if (rt_h2d(h->act[0], x, n * sizeof(float)) || rt_h2d(h->target, target, n * sizeof(float)))
        return net_fail(h, rt_error());
    for (int k = 0; k < steps; ++k) {
        if (net_gradient_dev(h, batch, loss)) return 1;
        ++h->step;
        h->version = next_net_version();
        if (rt_adam(h->params, h->grad, h->m, h->v, lr, h->step, h->nparam)) return net_fail(h, rt_error());
        // the input activation is overwritten by nothing: act[0] stays x
    }
    return rt_sync() ? net_fail(h, rt_error()) : 0;
}

}  // extern "C"

// ------------------------------------------------------------------ learned global moves
extern "C" {

// Buffers of the device global moves and the copy of every chain's state
// (G, UDV factors and checkpoint lists, work matrices: one block of the same
// size per chain, consecutive) and counters taken before a move, so that
// rejected chains get them back without a rebuild.
static int gm_snapshot(bafq_handle* h, size_t total, int nc) {
    using namespace bafq;
    if (!h->gm_old) {
        // old fields, forward drift (or initial momenta), 5 sums per chain (doubles); acceptance flags
        const size_t bytes = sizeof(double) * (2 * total + 5 * (size_t)nc) + sizeof(int) * (size_t)nc;
        if (rt_malloc((void**)&h->gm_old, bytes)) return fail(h, rt_error());
        h->gm_f0 = h->gm_old + total;
        h->gm_out = h->gm_f0 + total;
        h->gm_accept = (int*)(h->gm_out + 5 * (size_t)nc);
    }
    const size_t block = h->chain_block_bytes;
    if (!h->gm_state) {
        if (rt_malloc((void**)&h->gm_state, block * nc + sizeof(double) * S_LEN * (size_t)nc))
            return fail(h, rt_error());
        h->gm_stats = (double*)((char*)h->gm_state + block * nc);
    }
    if (rt_d2d(h->gm_state, h->chains[0].g, block * nc) ||
        rt_d2d(h->gm_stats, h->d_stats, sizeof(double) * S_LEN * (size_t)nc))
        return fail(h, rt_error());
    return 0;
}

int bafq_global_propose(bafq_handle* h, bafq_net* net, double lam, const double* sigma, uint64_t seed,
                        double* out) {
    using namespace bafq;
    if (!h) return 1;
    if (h->models.size() != 1) return fail(h, "device global moves need a handle with one model");
    if (!sigma || !out || !(lam >= 0.0)) return fail(h, "invalid global move arguments");
    const ModelDev& m = h->models[0];
    const int nc = (int)h->chains.size();
    const NetShape s{m.ltrot, m.nly, m.nlx};
    if (net && (net->d.lt != m.ltrot || net->d.ly != m.nly || net->d.lx != m.nlx || net->d.channels != 2))
        return fail(h, "the network's lattice is not the model's");
    const size_t per = (size_t)gm_per(s), total = per * nc;
    if (gm_snapshot(h, total, nc)) return 1;
    h->gm_fvalid = 0;
    h->gm_last = 1;
    GmArgs a{};
    a.s = s;
    a.lam = lam;
    a.seed = seed;
    for (int ch = 0; ch < 2; ++ch) {
        a.mask[ch] = m.active[ch] ? 1.0 : 0.0;
        a.sigma[ch] = m.active[ch] && sigma[ch] > 0.0 ? sigma[ch] : 0.0;
        a.w[ch] = a.sigma[ch] > 0.0 ? 1.0 / (2.0 * a.sigma[ch] * a.sigma[ch]) : 0.0;
    }
    const int chunk = net ? net->max_batch : nc;
    for (int c0 = 0; c0 < nc; c0 += chunk) {
        a.c0 = c0;
        a.batch = std::min(chunk, nc - c0);
        const float* y = nullptr;
        if (net) {
            if (rt_gm_to_net(h->d_phi, net->act[0], a)) return fail(h, rt_error());
            if (net_forward_dev(net, a.batch)) return fail(h, net->err);
            y = net->act[net->d.layers];
        }
        if (rt_gm_propose(h->d_phi, h->gm_old, h->gm_f0, y, a)) return fail(h, rt_error());
        if (net) {
            if (rt_gm_to_net(h->d_phi, net->act[0], a)) return fail(h, rt_error());
            if (net_forward_dev(net, a.batch)) return fail(h, net->err);
        }
        if (rt_gm_sums(h->d_phi, h->gm_old, h->gm_f0, y, a, h->gm_out)) return fail(h, rt_error());
    }
    if (rt_sync() || rt_d2h(out, h->gm_out, sizeof(double) * 5 * (size_t)nc)) return fail(h, rt_error());
    return launch_all(h, K_PREPARE, 0, 0);
}

int bafq_hmc_propose(bafq_handle* h, bafq_net* net, double angle, int steps, const double* center,
                     const double* force_scale, uint64_t seed, double* out) {
    using namespace bafq;
    if (!h) return 1;
    if (h->models.size() != 1) return fail(h, "device global moves need a handle with one model");
    if (!center || !force_scale || !out || !(angle > 0.0) || !(angle <= 3.2) || steps < 1 || steps > 64)
        return fail(h, "invalid Hamiltonian move arguments (0 < angle <= 3.2, 1 <= steps <= 64)");
    const ModelDev& m = h->models[0];
    const int nc = (int)h->chains.size();
    const NetShape s{m.ltrot, m.nly, m.nlx};
    if (net && (net->d.lt != m.ltrot || net->d.ly != m.nly || net->d.lx != m.nlx || net->d.channels != 2))
        return fail(h, "the network's lattice is not the model's");
    const size_t per = (size_t)gm_per(s), total = per * nc;
    if (gm_snapshot(h, total, nc)) return 1;
    if (!h->gm_p && rt_malloc((void**)&h->gm_p, sizeof(double) * total)) return fail(h, rt_error());
    GmArgs a{};
    a.s = s;
    a.lam = 0.5 * angle;                  // half kick
    a.seed = seed;
    a.cosh_ = std::cos(angle);
    a.sinh_ = std::sin(angle);
    for (int ch = 0; ch < 2; ++ch) {
        a.mask[ch] = m.active[ch] ? 1.0 : 0.0;
        a.center[ch] = center[ch];
        a.fscale[ch] = force_scale[ch];
    }
    // network output at the current fields: from the previous move when it is
    // still valid (same network version, fields unchanged since), else computed
    const size_t fper = 2 * (size_t)s.vol();
    if (net && !h->gm_fcur) {
        if (rt_malloc((void**)&h->gm_fcur, 2 * sizeof(float) * fper * nc)) return fail(h, rt_error());
        h->gm_fnew = h->gm_fcur + fper * nc;
        h->gm_fvalid = 0;
    }
    const bool cached = net && h->gm_fvalid && h->gm_fver == net->version;
    const int chunk = net ? net->max_batch : nc;
    for (int c0 = 0; c0 < nc; c0 += chunk) {
        a.c0 = c0;
        a.batch = std::min(chunk, nc - c0);
        if (rt_gm_hmc_init(h->d_phi, h->gm_old, h->gm_f0, h->gm_p, a)) return fail(h, rt_error());
        const float* yout = net ? net->act[net->d.layers] : nullptr;
        const float* y = nullptr;
        if (net) {
            float* cur = h->gm_fcur + fper * c0;
            if (!cached) {
                if (rt_gm_to_net(h->d_phi, net->act[0], a)) return fail(h, rt_error());
                if (net_forward_dev(net, a.batch)) return fail(h, net->err);
                if (rt_d2d(cur, yout, sizeof(float) * fper * a.batch)) return fail(h, rt_error());
            }
            y = cur;
        }
        for (int k = 0; k < steps; ++k) {
            if (net && rt_gm_hmc_kick(h->gm_p, y, a)) return fail(h, rt_error());
            if (rt_gm_hmc_rotate(h->d_phi, h->gm_p, a)) return fail(h, rt_error());
            if (net) {
                if (rt_gm_to_net(h->d_phi, net->act[0], a)) return fail(h, rt_error());
                if (net_forward_dev(net, a.batch)) return fail(h, net->err);
                y = yout;
                if (rt_gm_hmc_kick(h->gm_p, y, a)) return fail(h, rt_error());
            }
        }
        if (net && rt_d2d(h->gm_fnew + fper * c0, yout, sizeof(float) * fper * a.batch)) return fail(h, rt_error());
        if (rt_gm_hmc_sums(h->d_phi, h->gm_old, h->gm_f0, h->gm_p, a, h->gm_out)) return fail(h, rt_error());
    }
    h->gm_last = 2;
    h->gm_fnet = net;
    h->gm_fver = net ? net->version : 0;
    h->gm_fvalid = 0;                 // until bafq_global_finish settles which output is current
    if (rt_sync() || rt_d2h(out, h->gm_out, sizeof(double) * 5 * (size_t)nc)) return fail(h, rt_error());
    return launch_all(h, K_PREPARE, 0, 0);
}

int bafq_global_finish(bafq_handle* h, const int32_t* accept) {
    using namespace bafq;
    if (!h) return 1;
    if (!h->gm_old || !accept) return fail(h, "no device global move to finish");
    const ModelDev& m = h->models[0];
    const int nc = (int)h->chains.size();
    const int per = gm_per(NetShape{m.ltrot, m.nly, m.nlx});
    if (rt_h2d(h->gm_accept, accept, sizeof(int) * nc)) return fail(h, rt_error());
    if (rt_gm_restore(h->d_phi, h->gm_old, h->gm_accept, per, (size_t)per * nc)) return fail(h, rt_error());
    // matrices and counters of the rejected chains from the copy taken before the move
    const size_t words = h->chain_block_bytes / 8;
    if (rt_gm_restore_words((uint64_t*)h->chains[0].g, h->gm_state, h->gm_accept, words, words * nc) ||
        rt_gm_restore_words((uint64_t*)h->d_stats, (const uint64_t*)h->gm_stats, h->gm_accept, S_LEN,
                            (size_t)S_LEN * nc))
        return fail(h, rt_error());
    // Hamiltonian move with a network: the output at the proposal becomes the
    // current one where the move was accepted (2 Ltrot Ly Lx floats = Ltrot Ly Lx words)
    if (h->gm_last == 2 && h->gm_fnet) {
        const size_t fwords = (size_t)m.ltrot * m.nly * m.nlx;
        if (rt_gm_copy_accepted((uint64_t*)h->gm_fcur, (const uint64_t*)h->gm_fnew, h->gm_accept, fwords,
                                fwords * nc))
            return fail(h, rt_error());
        h->gm_fvalid = 1;
    } else {
        h->gm_fvalid = 0;
    }
    h->gm_last = 0;
    return rt_sync() ? fail(h, rt_error()) : 0;
}

}  // extern "C"

// ------------------------------------------------------------------ gradient of log W
extern "C" {

int bafq_grad_log_weight(bafq_handle* h, double* grad) {
    using namespace bafq;
    if (!h) return 1;
    if (!grad) return fail(h, "null output");
    if (!h->d_grad && rt_malloc((void**)&h->d_grad, sizeof(double) * h->phi_total)) return fail(h, rt_error());
    if (launch_all(h, K_GRAD, 1, 0)) return 1;
    if (rt_d2h(grad, h->d_grad, sizeof(double) * h->phi_total)) return fail(h, rt_error());
    return 0;
}

}  // extern "C"
