// The Markov-chain sweep of one chain, executed by one workgroup.
//
// Restates src/number_conserving/src/local_sweep.f90:55-161 (identical
// orchestration in src/pairing/src/local_sweep.f90) for both solvers in one
// "reduced" form:
//   * number conserving: N = Ns, one row per site, the c flavour is the
//     complex conjugate of the b flavour (src/number_conserving/src/obser_equal.f90:113-117);
//   * Nambu pairing: the reference's 4Ns matrices are block diagonal on the
//     sectors {b, c+} and {c, b+}, and for real t, real Delta, imaginary
//     alpha_1 and real alpha_2 the second block is the complex conjugate of
//     the first.  The solver propagates only the first block, N = 2 Ns, rows
//     (b_1..b_Ns, c+_1..c+_Ns); det P_4x4 = |det P_2x2|^2.
// HS diagonal factors of site i and channel nf are exp(s * alpha_nf * phi)
// with s = +1 on every row except the c+ row of the total-density channel
// (src/pairing/src/operator_Hubbard.f90:155-199).
//
// Two execution paths share the numerics:
//   * LDS path (N <= kLdsMaxN, 512 threads): G lives in LDS for the whole
//     sweep; the kinetic propagation runs on FP64 MFMA from LDS with the
//     propagators pre-ordered as MFMA operands; the B(tau) products of the U
//     factors are applied at each stabilisation point (the fields of those
//     slices are final by then), in LDS.
//   * global path (any N): G and all factors in global memory (L2/MALL).
#pragma once

#include "la.hpp"

namespace bafq {

// ---------------------------------------------------------------- layouts
enum ObsIndex {
    O_DENSITY_UP = 0, O_DENSITY_DO, O_DENSITY_TOTAL, O_DENSITY, O_KINETIC, O_DOUBLEOCC,
    O_SQUAREOCC, O_LOCAL_NUMSQUARE, O_PAIR_EQUAL, O_NUM_UP, O_NUM_DO, O_NUMSQUARE_UP,
    O_NUMSQUARE_DO, O_ONSITE_N2_UP, O_ONSITE_N2_DO, O_INTERACTION, O_PAIRING_E, O_CHEMICAL_E,
    O_GRAND_E, O_ENERGY, O_SF_K_RE, O_SF_K_IM, O_PSF_RE, O_PSF_IM, O_DW_K_RE, O_DW_K_IM,
    O_DEN_UPUP_RE, O_DEN_UPUP_IM, O_DEN_DODO_RE, O_DEN_DODO_IM, O_DEN_UPDO_RE, O_DEN_UPDO_IM,
    O_NOBS, O_SITE0 = 40
};
enum StatIndex {
    S_PROP_U1 = 0, S_ACC_U1, S_PROP_U2, S_ACC_U2, S_XMAX, S_XMEAN_SUM, S_NWRAP, S_LOGDET,
    S_THERM_PROP_U1, S_THERM_ACC_U1, S_THERM_PROP_U2, S_THERM_ACC_U2,
    // device-clock ticks spent per phase (thread 0 of the chain's workgroup)
    S_T_SITE, S_T_KIN, S_T_MEAS, S_T_WRAP, S_T_HS,
    // parts of S_T_WRAP: B-products of the U factor, the two pivoted QRs, the rest of stab_green
    S_T_WPROD, S_T_WQR, S_T_WGREEN,
    // Wrap_tau drifts of G(0,tau), G(tau,0), G(tau,tau): running maxima, sums, count
    S_TAU_XMAX = 20, S_TAU_XSUM = 23, S_TAU_NWRAP = 26, S_LEN = 27
};
constexpr int kNumTimers = 8;

constexpr int kLdsWaves = 8;      // LDS path: 512 threads
constexpr int kMaxDkGlb = 32;     // widest delayed-update buffers (in device memory)
constexpr int kLdsMaxN = 96;      // LDS path: N <= 96 (G in LDS, <= 36 MFMA tiles)
constexpr int kMaxTiles = 5;      // ceil(36 / 8) output tiles per wavefront

struct ModelDev {
    int kind;            // 0 number conserving, 1 Nambu pairing
    int n, ns, nlx, nly, ltrot, nwrap, nst, has_k;
    int dk;              // columns of the delayed-update buffers (0: per-site updates)
    int uw_glb;          // the buffers live in the chain's work matrices w2, w3 (device memory), not LDS
    int active[2];
    double u1, u2, mu, rdelta, shift_loc;
    cd alpha[2];
    // MFMA-operand order: [0] e^{-dtau A} as left factor ([ti][kk][64]),
    // [1] e^{+dtau A} left, [2] e^{-dtau A} right ([tj][kk][64]), [3] e^{+dtau A} right
    const double* ez[4];
    const cd* phasek;    // exp(i K . r_imj) indexed by imj
    const int* bonds;    // ns x 3 forward bonds
};

struct ChainDev {
    int model;
    int* seed;
    double* phi;                       // ltrot x ns x 2
    cd *g, *ur, *vr, *ul, *vl;
    double *dr, *dl;
    cd *urlist, *vrlist, *ullist, *vllist;
    double *drlist, *dllist;
    cd *wa, *w1, *w2, *w3;
    double* obs;
    double* stats;
    cd* tau;                           // time-displaced workspace (null until requested)
    cd* tau_rec;                       // Ltrot x 4 time-displaced functions of the last pass, or null
};

struct LaunchArgs {
    const ModelDev* models;
    const ChainDev* chains;
    const int* ids;
    int count;
    int nsweep;
    int measure;
    int* queue;          // work counter of the launch (zeroed before it), see next_item
    double* grad;        // K_GRAD: d log W / d phi, laid out like the fields starting at phi0
    const double* phi0;
    int tau;             // K_SWEEP_TAU: convention of the time-displaced pass (TauMode)
    int tau_record;      // K_SWEEP_TAU: keep the last pass's functions at every slice
    // K_SWEEP on the device-memory path: 1 = two workgroups per chain (the grid
    // holds 2 x count workgroups, all resident).  Per pair: pair_flags[2p] its
    // barrier arrivals, pair_flags[2p + 1] set when a wait timed out (both zeroed
    // before the kernel), pair_mail[p] the chain index the leader took.
    int pairs;
    int npairs;
    int* pair_flags;
    int* pair_mail;
    uint64_t pair_wait;  // longest wait at a pair barrier, in clock() ticks
};

// Bytes at the start of a workgroup's dynamic LDS that hold the work-queue
// slot; the per-chain LDS layout (carve) starts after them.
constexpr size_t kSlotBytes = 16;

// ---------------------------------------------------------------- RNG
constexpr int64_t kMod = 2147483648LL;
constexpr int64_t kMul = 48828125LL;
// seed <- 48828125 * seed mod 2^31, u = seed / 2^31 (src/common/random.f90:3-11).
DEV double ranf(int& seed) {
    const int64_t s = (kMul * (int64_t)seed) % kMod;
    seed = (int)s;
    return (double)s / 2147483648.0;
}
// 48828125^k mod 2^31 (jump ahead k draws).
DEV int64_t lcg_pow(int k) {
    int64_t r = 1, b = kMul;
    while (k) {
        if (k & 1) r = (r * b) % kMod;
        b = (b * b) % kMod;
        k >>= 1;
    }
    return r;
}
DEV int lcg_jump(int seed, int k) { return (int)((lcg_pow(k) * (int64_t)seed) % kMod); }

// ---------------------------------------------------------------- LDS plan
// One proposal of the local Metropolis step, fixed before the site loop:
// the RNG draws of a site do not depend on earlier accept decisions.
struct Proposal {
    double phi_new;
    double ratio0;   // Gaussian ratio (times the Nambu normal-ordering constant)
    double u;        // the uniform the ratio is compared with
    double pad;
    cd d0, d1;       // e^{s alpha (phi'-phi)} - 1 of the site's rows
};

struct Lds {
    QrScratch qr;
    double* dtmp;      // n
    cd* f;             // n
    cd* finv;          // n
    cd* col;           // 2 slots x 2 sectors x n
    cd* row;           // 2 slots x 2 sectors x n
    // the device-memory path's kinetic products stage X here (gemm_rc_panel):
    // the whole layout (panel_cap complex), none of which is live during those
    // products; with G in LDS only col .. prop (18 n + 256 >= 16 (n + 3))
    cd* panel;
    int panel_cap;
    double* red;       // nwaves x 64
    cd* site;          // 6 x ns (pairing measurement)
    Proposal* prop;    // ns
    cd* mat;           // n x n (LDS path only)
    cd* du;            // dk x n   delayed-update factors, U transposed: G -= du^T dw (LDS path, dk > 0)
    cd* dw;            // dk x n
    // submatrix updates (device-memory path, sb = dk > 0): G[R,R] of a block of
    // rows R (|R| <= sb) in two ping-pong buffers, its value at the block start,
    // the accepted rows' local indices and changes
    cd* sbuf;          // 2 x sb x sb
    cd* sg0;           // sb x sb
    int* sloc;         // sb
    cd* sd;            // sb
};

HOSTDEV size_t lds_align(size_t b) { return (b + 15) & ~(size_t)15; }
HOSTDEV size_t lds_bytes(int n, int nwaves, bool with_mat, int dk = 0, int sb = 0) {
    size_t b = 0;
    b += lds_align((size_t)n * sizeof(cd));           // tau
    b += lds_align((size_t)n * sizeof(int) + 16);     // perm
    b += lds_align((size_t)n * sizeof(double));       // nrm
    b += lds_align(16 * sizeof(double));              // bcast
    b += lds_align((size_t)n * sizeof(double));       // dtmp
    b += 2 * lds_align((size_t)n * sizeof(cd));       // f, finv
    b += 2 * lds_align(4 * (size_t)n * sizeof(cd));   // col, row: [slot][sector][n]
    b += lds_align((size_t)nwaves * 64 * sizeof(double));
    b += lds_align(6 * (size_t)n * sizeof(cd));
    b += lds_align((size_t)n * sizeof(Proposal));
    if (with_mat) b += lds_align((size_t)n * n * sizeof(cd));
    b += 2 * lds_align((size_t)n * dk * sizeof(cd));
    if (sb > 0) {
        b += lds_align((size_t)2 * sb * sb * sizeof(cd));
        b += lds_align((size_t)sb * sb * sizeof(cd));
        b += lds_align((size_t)sb * sizeof(int));
        b += lds_align((size_t)sb * sizeof(cd));
    }
    return b + 64;
}

DEV Lds carve(char* base, int n, int nwaves, bool with_mat, int dk = 0, int sb = 0) {
    Lds L;
    char* p = base;
    auto take = [&](size_t bytes) {
        char* q = p;
        p += lds_align(bytes);
        return q;
    };
    L.qr.tau = (cd*)take((size_t)n * sizeof(cd));
    L.qr.perm = (int*)take((size_t)n * sizeof(int) + 16);
    L.qr.nrm = (double*)take((size_t)n * sizeof(double));
    L.qr.bcast = (double*)take(16 * sizeof(double));
    L.dtmp = (double*)take((size_t)n * sizeof(double));
    L.f = (cd*)take((size_t)n * sizeof(cd));
    L.finv = (cd*)take((size_t)n * sizeof(cd));
    L.col = (cd*)take(4 * (size_t)n * sizeof(cd));
    L.qr.v = L.col;                              // free while a QR runs
    L.row = (cd*)take(4 * (size_t)n * sizeof(cd));
    L.panel = L.col;
    L.red = (double*)take((size_t)nwaves * 64 * sizeof(double));
    L.site = (cd*)take(6 * (size_t)n * sizeof(cd));
    L.prop = (Proposal*)take((size_t)n * sizeof(Proposal));
    L.panel_cap = (int)((p - (char*)L.panel) / sizeof(cd));
    L.mat = with_mat ? (cd*)take((size_t)n * n * sizeof(cd)) : nullptr;
    L.du = dk > 0 ? (cd*)take((size_t)n * dk * sizeof(cd)) : nullptr;
    L.dw = dk > 0 ? (cd*)take((size_t)n * dk * sizeof(cd)) : nullptr;
    L.sbuf = sb > 0 ? (cd*)take((size_t)2 * sb * sb * sizeof(cd)) : nullptr;
    L.sg0 = sb > 0 ? (cd*)take((size_t)sb * sb * sizeof(cd)) : nullptr;
    L.sloc = sb > 0 ? (int*)take((size_t)sb * sizeof(int)) : nullptr;
    L.sd = sb > 0 ? (cd*)take((size_t)sb * sizeof(cd)) : nullptr;
    if (!with_mat) {                             // the whole layout
        L.panel = (cd*)base;
        L.panel_cap = (int)((p - base) / sizeof(cd));
    }
    return L;
}

// ---------------------------------------------------------------- timers
struct Timer {
    uint64_t t0;
    bool on = true;      // false in the second workgroup of a pair (the leader keeps the times)
    DEV void start(const Ctx& c) { t0 = c.clock(); }
    DEV void stop(const Ctx& c, const ChainDev& ch, int slot) {
        const uint64_t t1 = c.clock();
        if (on && c.tid() == 0) ch.stats[slot] += (double)(t1 - t0);
        t0 = t1;
    }
};

// Workgroups that advance one chain together on the device-memory path (a
// pair, or a single workgroup: parts = 1).  Both members run the same control
// flow: every Metropolis decision is made by both from the same data, the
// matrix products and rank-k updates are split by 16-row tiles (Part), and
// the stabilisation and the estimators run in the leader (part 0) alone.
// team_sync orders every write to device memory before it for both members.
struct Team {
    int part = 0, parts = 1;
    int* flags = nullptr;    // [0] arrivals, [1] timed out (device memory)
    int* epoch = nullptr;    // barriers passed (LDS, thread 0)
    uint64_t wait = 0;       // longest wait, clock() ticks
    DEV bool lead() const { return part == 0; }
    DEV Part share() const { return Part{part, parts}; }
};

// Barrier of the team.  A member that waits longer than team.wait sets the
// timed-out flag, after which no member of that pair waits again: the kernel
// ends and the launch reports the error (its results are then invalid).
DEV void team_sync(const Ctx& c, const Team& team) {
    if (team.parts == 1) {
        c.sync();
        return;
    }
    c.stores_done();
    c.sync();
    if (c.tid() == 0) {
        const int target = team.parts * ++*team.epoch;
        c.arrive(team.flags);                    // releases the workgroup's completed stores
        if (!c.load_relaxed(team.flags + 1)) {
            const uint64_t t0 = c.clock();
            while (c.load_relaxed(team.flags) < target) {
                if (c.clock() - t0 > team.wait) {
                    c.store_flag(team.flags + 1, 1);
                    break;
                }
                c.pause();
            }
        }
        c.acquire_device();                      // the partner's stores, for the whole workgroup (one L1)
    }
    c.sync();
}

// ---------------------------------------------------------------- helpers
DEV int nsec(const ModelDev& m) { return m.kind == 1 ? 2 : 1; }
// Exponent sign of sector `sec` in channel nf (1: relative U1, 2: total U2).
DEV double sec_sign(const ModelDev& m, int nf, int sec) {
    return (m.kind == 1 && nf == 2 && sec == 1) ? -1.0 : 1.0;
}

// d log W / d phi_{t,i,nf} from the Green's function at the point of the
// local update of slice t, channel nf (where a field change e^{s alpha dphi}
// of site i multiplies det(1 - B) by det(1 + Delta (1 - G_SS))):
//   number conserving  -phi - 2 Re[alpha (1 - G_ii)]             (W ~ |det G|^2)
//   pairing            -phi - Re sum_sec s alpha (1 - G_rr) [- alpha_2 for nf = 2]
// (the normal-ordering constant of the U2 channel, W ~ |det G_2Ns|).
template <class GP>
DEV void grad_extract(const Ctx& c, const ModelDev& m, const double* phi_t, GP G, int t, int nf, double* grad) {
    const int ns = m.ns, n = m.n;
    const double cpow = m.kind == 1 ? 1.0 : 2.0;
    const cd a = m.alpha[nf - 1];
    for (int i = c.tid(); i < ns; i += c.nthreads()) {
        double det = 0.0;
        for (int sec = 0; sec < nsec(m); ++sec) {
            const int r = sec * ns + i;
            const cd g = ldc(G, r * n + r);
            det += sec_sign(m, nf, sec) * (a.r * (1.0 - g.r) + a.i * g.i);
        }
        double v = -phi_t[i * 2 + nf - 1] - cpow * det;
        if (m.kind == 1 && nf == 2) v -= m.alpha[1].r;
        grad[((size_t)t * ns + i) * 2 + nf - 1] = v;
    }
    c.sync();
}

// f[r] = exp(s_r alpha_nf phi_r), finv = 1 / f, for all rows (LDS); phi of
// field channel fld (0: the channel's own, nf).
DEV void hs_factors(const Ctx& c, const ModelDev& m, const Lds& L, const double* phi_t, int nf, int fld = 0) {
    const int ns = m.ns, k = nsec(m);
    auto f = as_lds(L.f);
    auto finv = as_lds(L.finv);
    for (int r = c.tid(); r < ns * k; r += c.nthreads()) {
        const int sec = r / ns, i = r - sec * ns;
        const double s = sec_sign(m, nf, sec);
        const double ph = phi_t[i * 2 + ((fld ? fld : nf) - 1)];
        const cd a = m.alpha[nf - 1];
        stc(f, r, cexp(cd{s * a.r * ph, s * a.i * ph}));
        stc(finv, r, cexp(cd{-s * a.r * ph, -s * a.i * ph}));
    }
    c.sync();
}

// M <- diag(fl) M diag(fr), fl/fr in LDS (or null); M in LDS or global (pointer type MP).
template <class MP>
DEV void scale_mat_t(const Ctx& c, MP M, int n, const cd* fl_, const cd* fr_, Part pt = Part{}) {
    auto fl = as_lds(fl_);
    auto fr = as_lds(fr_);
    const int nn = n * n;
    for (int e = c.tid() + pt.i * c.nthreads(); e < nn; e += pt.n * c.nthreads()) {
        const int r = e / n, k = e - r * n;
        cd x = ldc(M, e);
        if (fl_) x = ldc(fl, r) * x;
        if (fr_) x = x * ldc(fr, k);
        stc(M, e, x);
    }
    c.sync();
}
DEV void scale_mat(const Ctx& c, cd* M, int n, const cd* fl, const cd* fr, Part pt = Part{}) {
    scale_mat_t(c, M, n, fl, fr, pt);
}
DEV void scale_mat_lds(const Ctx& c, cd* M, int n, const cd* fl, const cd* fr) {
    scale_mat_t(c, as_lds(M), n, fl, fr);
}

// dst <- src between LDS and global memory (n x n complex).
DEV void lds_from_global(const Ctx& c, cd* dst, const cd* src, int n) {
    auto d = as_lds(dst);
    for (int e = c.tid(); e < n * n; e += c.nthreads()) stc(d, e, src[e]);
    c.sync();
}
DEV void global_from_lds(const Ctx& c, cd* dst, const cd* src, int n) {
    auto s = as_lds(src);
    for (int e = c.tid(); e < n * n; e += c.nthreads()) dst[e] = ldc(s, e);
    c.sync();
}

// ---------------------------------------------------------------- Metropolis
// Proposals of one channel on one slice (LocalU_metro's draws, NC
// src/number_conserving/src/localU.f90:44-61 and operator_Hubbard.f90:45-68;
// pairing src/pairing/src/localU.f90:46-71 and operator_Hubbard.f90:90-137).
// Site i is visited at position p (ascending: p = i, descending: p = Ns-1-i)
// and consumes draws 2p (shift) and 2p+1 (acceptance) of the chain stream.
DEVFN void make_proposals(const Ctx& c, const ModelDev& m, const Lds& L, const double* phi_t, int nf,
                        bool ascending, int seed) {
    const cd a = m.alpha[nf - 1];
    for (int i = c.tid(); i < m.ns; i += c.nthreads()) {
        const int pos = ascending ? i : m.ns - 1 - i;
        int s = lcg_jump(seed, 2 * pos);
        const double xflip = ranf(s);
        const double u = ranf(s);
        const double phi_old = phi_t[i * 2 + nf - 1];
        const double phi_new = phi_old + (xflip - 0.5) * fabs(m.shift_loc);
        const double rg = exp(-0.5 * phi_new * phi_new) / exp(-0.5 * phi_old * phi_old);
        Proposal P;
        P.phi_new = phi_new;
        P.u = u;
        P.pad = 0.0;
        if (m.kind == 0) {
            const cd eo = cexp(cd{a.r * phi_old, a.i * phi_old});
            const cd en = cexp(cd{a.r * phi_new, a.i * phi_new});
            P.d0 = cdiv(en, eo) - cd{1.0, 0.0};
            P.d1 = cd{0.0, 0.0};
            P.ratio0 = rg;
        } else {
            const double dphi = phi_new - phi_old;
            const cd ep = cexp(cd{a.r * dphi, a.i * dphi});
            const cd em = cexp(cd{-a.r * dphi, -a.i * dphi});
            P.d0 = ep - cd{1.0, 0.0};
            P.d1 = (nf == 2 ? em : ep) - cd{1.0, 0.0};
            double rconst = 1.0;
            if (nf == 2) {
                const cd eo = cexp(cd{a.r * phi_old, a.i * phi_old});
                const cd en = cexp(cd{a.r * phi_new, a.i * phi_new});
                rconst = cdiv(eo, en).r;
            }
            P.ratio0 = rg * rconst;
        }
        auto pp = as_lds(L.prop) + i;
        pp->phi_new = P.phi_new;
        pp->ratio0 = P.ratio0;
        pp->u = P.u;
        pp->pad = 0.0;
        pp->d0.r = P.d0.r; pp->d0.i = P.d0.i;
        pp->d1.r = P.d1.r; pp->d1.i = P.d1.i;
    }
    c.sync();
}

// Sequential Metropolis decisions and Sherman-Morrison updates of one channel
// on time slice t (0-based) in ascending (sweep right) or descending (sweep
// left) site order.  G must be in the frame in which the reference evaluates
// the ratio (all sites' HS factors of this channel applied for sweep right;
// none for sweep left) -- the per-site similarity transforms commute with the
// rank-k updates and leave G_SS unchanged.  Advances `seed` by 2 Ns draws.
template <class GP>
DEV void site_loop_t(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, GP G, int t,
                     int nf, bool ascending, int& seed) {
    const int n = m.n, ns = m.ns;
    const bool pairing = m.kind == 1;
    const int tid = c.tid(), nth = c.nthreads();
    double* phi_t = ch.phi + (size_t)t * ns * 2;
    make_proposals(c, m, L, phi_t, nf, ascending, seed);
    seed = lcg_jump(seed, 2 * ns);
    auto col0 = as_lds(L.col);
    auto col1 = as_lds(L.col) + n;
    auto row0 = as_lds(L.row);
    auto row1 = as_lds(L.row) + n;
    auto props = as_lds(L.prop);
    // fixed-column mapping of the rank-k update: thread -> column cc, rows r0, r0 + rstep, ...
    const bool fixed_col = n <= nth;
    const int rstep = fixed_col ? nth / n : 1;
    const int cc = fixed_col ? tid % n : 0;
    const int r0 = fixed_col ? tid / n : 0;
    int nacc = 0;
    for (int q = 0; q < ns; ++q) {
        const int i = ascending ? q : ns - 1 - q;
        const int s0 = i, s1 = ns + i;
        Proposal P;
        {
            auto pp = props + i;
            P.phi_new = pp->phi_new;
            P.ratio0 = pp->ratio0;
            P.u = pp->u;
            P.d0 = cd{pp->d0.r, pp->d0.i};
            P.d1 = cd{pp->d1.r, pp->d1.i};
        }
        cd m00, m01 = cd{0.0, 0.0}, m10 = cd{0.0, 0.0}, m11 = cd{0.0, 0.0};
        bool acc;
        if (!pairing) {
            const cd g00 = ldc(G, s0 * n + s0);
            const cd den = cd{1.0, 0.0} + P.d0 * (cd{1.0, 0.0} - g00);
            const cd rdet = cinv(den);
            const cd z = (cd{P.ratio0, 0.0} * rdet) * conj(rdet);
            acc = cabs(z) > P.u;
            m00 = rdet * P.d0;
        } else {
            const cd g00 = ldc(G, s0 * n + s0), g01 = ldc(G, s0 * n + s1);
            const cd g10 = ldc(G, s1 * n + s0), g11 = ldc(G, s1 * n + s1);
            const cd p00 = cd{1.0, 0.0} + P.d0 * (cd{1.0, 0.0} - g00);
            const cd p01 = -(P.d0 * g01);
            const cd p10 = -(P.d1 * g10);
            const cd p11 = cd{1.0, 0.0} + P.d1 * (cd{1.0, 0.0} - g11);
            const cd det = p00 * p11 - p01 * p10;
            acc = fabs(P.ratio0 / cabs(det)) > P.u;
            const cd idet = cinv(det);
            m00 = (p11 * idet) * P.d0;
            m01 = (-(p01 * idet)) * P.d1;
            m10 = (-(p10 * idet)) * P.d0;
            m11 = (p00 * idet) * P.d1;
        }
        if (!acc) continue;
        ++nacc;
        // snapshot the column(s) and row(s) of the site
        for (int r = tid; r < n; r += nth) {
            stc(col0, r, ldc(G, r * n + s0));
            stc(row0, r, ldc(G, s0 * n + r));
            if (pairing) {
                stc(col1, r, ldc(G, r * n + s1));
                stc(row1, r, ldc(G, s1 * n + r));
            }
        }
        c.sync();
        // every thread has read G_SS before this point
        if (tid == 0) phi_t[i * 2 + nf - 1] = P.phi_new;
        // G -= G[:,S] M (I_S - G[S,:])
        for (int cb = fixed_col ? cc : tid; cb < n; cb += fixed_col ? n : nth) {
            if (fixed_col && r0 >= rstep) break;
            cd e0 = -ldc(row0, cb);
            if (cb == s0) e0 = e0 + cd{1.0, 0.0};
            cd w0 = m00 * e0, w1 = cd{0.0, 0.0};
            if (pairing) {
                cd e1 = -ldc(row1, cb);
                if (cb == s1) e1 = e1 + cd{1.0, 0.0};
                w0 = w0 + m01 * e1;
                w1 = m10 * e0 + m11 * e1;
            }
            for (int r = fixed_col ? r0 : 0; r < n; r += rstep) {
                cd x = ldc(G, r * n + cb);
                x = x - ldc(col0, r) * w0;
                if (pairing) x = x - ldc(col1, r) * w1;
                stc(G, r * n + cb, x);
            }
        }
        c.sync();
    }
    if (tid == 0) {
        ch.stats[nf == 1 ? S_PROP_U1 : S_PROP_U2] += (double)ns;
        ch.stats[nf == 1 ? S_ACC_U1 : S_ACC_U2] += (double)nacc;
    }
}

DEVFN void site_loop(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G, int t, int nf,
                     bool ascending, int& seed) {
    site_loop_t(c, m, ch, L, G, t, nf, ascending, seed);
}
DEVFN void site_loop_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G, int t, int nf,
                         bool ascending, int& seed) {
    site_loop_t(c, m, ch, L, as_lds(G), t, nf, ascending, seed);
}

// ---------------------------------------------------------------- delayed local updates
// LDS path.  The accepted corrections of consecutive sites are collected as
// G = G0 - U W (U: n x cols, W: cols x n, cols <= dk) and applied to G0 with one
// MFMA product when the buffers are full and at the end of the slice.  Site i
// needs only G[S,S] = G0[S,S] - U[S,:] W[:,S] for its decision, which every
// wavefront evaluates itself (lanes over the k x k entries and the inner index,
// a DPP row sum), so a rejected proposal costs no barrier; an accepted one forms
// its own factors from the current columns G[:,S] and rows G[S,:] and costs one.
// Both factors are stored with the long index contiguous (Ut[j][r] = U[r][j],
// W[j][c]) so that consecutive lanes touch consecutive LDS addresses.

// G0 -= U[:, :cols] W[:cols, :] on FP64 MFMA 16x16x4 (each output tile owned by
// one wavefront); G in LDS or in device memory, U and W in LDS or device
// memory; cols <= kMaxCols (16 or 32).  A tile's operands for all its k-steps
// and its tile of G0 are loaded before the products (one round trip per tile).
// With a Part, only the row tiles ti = pt.i, pt.i + pt.n, ... (rows of Ut
// outside them are not read).
template <int kMaxCols = 16, class GP, class UP>
DEV void delayed_flush_t(const Ctx& c, GP G, UP Ut, UP W, int n, int cols, Part pt = Part{}) {
    constexpr int kQ = kMaxCols / 4;
    n = c.uniform(n);                 // scalar tests of the k-steps taken
    cols = c.uniform(cols);
    const int nt = (n + 15) >> 4;
    const int nrt = (nt - pt.i + pt.n - 1) / pt.n;    // row tiles of this part
    const int lane = c.lane(), li = lane & 15, lk = lane >> 4;
    for (int t = c.wave(); t < nrt * nt; t += c.nwaves()) {
        const int ti = pt.i + pt.n * (t / nt), tj = t % nt;
        const int ar = 16 * ti + li, bc = 16 * tj + li;
        cd a[kQ], b[kQ], g[4];
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            const int kk = 4 * q + lk;
            a[q] = cd{0.0, 0.0};
            b[q] = cd{0.0, 0.0};
            if (ar < n && kk < cols) a[q] = ldc(Ut, kk * n + ar);
            if (bc < n && kk < cols) b[q] = ldc(W, kk * n + bc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            g[r] = cd{0.0, 0.0};
            if (row < n && col < n) g[r] = ldc(G, row * n + col);
        }
        // three real products per complex one (Gauss): Re = ar br - ai bi,
        // Im = (ar + ai)(br + bi) - ar br - ai bi; three independent chains
        d4 rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ss = rr;
#pragma unroll
        for (int q = 0; q < kQ; ++q) {
            if (4 * q < cols) {
                rr = c.mfma(a[q].r, b[q].r, rr);
                ii = c.mfma(a[q].i, b[q].i, ii);
                ss = c.mfma(a[q].r + a[q].i, b[q].r + b[q].i, ss);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            if (row < n && col < n)
                stc(G, row * n + col, cd{g[r].r - (rr[r] - ii[r]), g[r].i - (ss[r] - rr[r] - ii[r])});
        }
    }
    c.sync();
}

DEVFN void delayed_flush(const Ctx& c, cd* G, const cd* Ut, const cd* W, int n, int cols) {
    delayed_flush_t(c, as_lds(G), as_lds(Ut), as_lds(W), n, cols);
}
DEVFN void delayed_flush_glb(const Ctx& c, cd* G, const cd* Ut, const cd* W, int n, int cols, Part pt = Part{}) {
    delayed_flush_t(c, as_global(G), as_lds(Ut), as_lds(W), n, cols, pt);
}
DEVFN void delayed_flush_glb_uw(const Ctx& c, cd* G, const cd* Ut, const cd* W, int n, int cols,
                                Part pt = Part{}) {
    delayed_flush_t<kMaxDkGlb>(c, as_global(G), as_global(Ut), as_global(W), n, cols, pt);
}

// Same decisions, RNG use and result as site_loop_t (LocalU_metro + LocalU_prop_L/R).
// G in LDS (kGlb false) or in device memory; U and W in LDS, or (kUW) in the
// chain's work matrices w2, w3 in device memory.  Requires dk <= n, and
// dk <= 16 with the buffers in LDS, dk <= kMaxDkGlb = 32 in device memory
// (fewer passes of the flush over a G in device memory).
template <bool kGlb, bool kUW = false>
DEV void site_loop_delayed_t(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G_, int t,
                             int nf, bool ascending, int& seed) {
    const int n = m.n, ns = m.ns, dk = m.dk;
    const bool pairing = m.kind == 1;
    const int ksec = pairing ? 2 : 1;
    const int tid = c.tid(), lane = c.lane();
    const int wbase = c.nthreads() >= 256 ? 128 : c.nthreads() / 2;
    auto G = [&] {
        if constexpr (kGlb) return as_global(G_);
        else return as_lds(G_);
    }();
    cd* const du = kUW ? ch.w2 : L.du;
    cd* const dw = kUW ? ch.w3 : L.dw;
    auto Ut = [&] {
        if constexpr (kUW) return as_global(du);
        else return as_lds(du);
    }();
    auto W = [&] {
        if constexpr (kUW) return as_global(dw);
        else return as_lds(dw);
    }();
    double* phi_t = ch.phi + (size_t)t * ns * 2;
    make_proposals(c, m, L, phi_t, nf, ascending, seed);
    seed = lcg_jump(seed, 2 * ns);
    auto props = as_lds(L.prop);
    // lane roles in the correction U[S,:] W[:,S]: entry (ea, eb), inner index ej
    const int ea = lane >> 5, eb = (lane >> 4) & 1, ej = lane & 15;
    int nacc = 0, cols = 0;
    for (int q = 0; q < ns; ++q) {
        const int i = ascending ? q : ns - 1 - q;
        const int s0 = i, s1 = ns + i;
        const int sa = (pairing && ea) ? s1 : s0, sb = (pairing && eb) ? s1 : s0;
        // everything the decision reads is loaded up front
        double cr = 0.0, ci = 0.0;
        if (ej < cols && (pairing || lane < 16)) {
            const cd u = ldc(Ut, ej * n + sa);
            const cd w = ldc(W, ej * n + sb);
            cr = u.r * w.r - u.i * w.i;
            ci = u.r * w.i + u.i * w.r;
        }
        if constexpr (kUW) {
            // inner indices 16..31 of buffers wider than 16 columns
            if (ej + 16 < cols && (pairing || lane < 16)) {
                const cd u = ldc(Ut, (ej + 16) * n + sa);
                const cd w = ldc(W, (ej + 16) * n + sb);
                cr += u.r * w.r - u.i * w.i;
                ci += u.r * w.i + u.i * w.r;
            }
        }
        const cd gss = ldc(G, sa * n + sb);
        Proposal P;
        {
            auto pp = props + i;
            P.phi_new = pp->phi_new;
            P.ratio0 = pp->ratio0;
            P.u = pp->u;
            P.d0 = cd{pp->d0.r, pp->d0.i};
            P.d1 = cd{pp->d1.r, pp->d1.i};
        }
        // lanes 16 e .. 16 e + 15 hold entry e = (ea, eb) of the current G[S,S]
        const cd ge = cd{gss.r - c.row16_sum(cr), gss.i - c.row16_sum(ci)};
        const cd g00 = cd{c.lane_bcast(ge.r, 0), c.lane_bcast(ge.i, 0)};
        bool acc;
        cd p00, p01 = cd{0.0, 0.0}, p10 = cd{0.0, 0.0}, p11 = cd{1.0, 0.0}, det;
        if (!pairing) {
            det = cd{1.0, 0.0} + P.d0 * (cd{1.0, 0.0} - g00);
            p00 = det;
            const double a2 = abs2(det);
            // |ratio0 / |den|^2| > u
            acc = fabs(P.ratio0 / a2) > P.u;
        } else {
            const cd g01 = cd{c.lane_bcast(ge.r, 16), c.lane_bcast(ge.i, 16)};
            const cd g10 = cd{c.lane_bcast(ge.r, 32), c.lane_bcast(ge.i, 32)};
            const cd g11 = cd{c.lane_bcast(ge.r, 48), c.lane_bcast(ge.i, 48)};
            p00 = cd{1.0, 0.0} + P.d0 * (cd{1.0, 0.0} - g00);
            p01 = -(P.d0 * g01);
            p10 = -(P.d1 * g10);
            p11 = cd{1.0, 0.0} + P.d1 * (cd{1.0, 0.0} - g11);
            det = p00 * p11 - p01 * p10;
            acc = fabs(P.ratio0 / sqrt(abs2(det))) > P.u;
        }
        if (!acc) continue;
        ++nacc;
        cd m00, m01 = cd{0.0, 0.0}, m10 = cd{0.0, 0.0}, m11 = cd{0.0, 0.0};
        const cd idet = cinv(det);
        if (!pairing) {
            m00 = idet * P.d0;
        } else {
            m00 = (p11 * idet) * P.d0;
            m01 = (-(p01 * idet)) * P.d1;
            m10 = (-(p10 * idet)) * P.d0;
            m11 = (p00 * idet) * P.d1;
        }
        if (tid == 0) (props + i)->pad = 1.0;                // accepted: phi is written after the loop
        // New factors: threads [0, wbase) form rows r of the new U columns, G[r,S] M
        // with G[r,S] = G0[r,S] - U[r,:] W[:,S]; threads [wbase, 2 wbase) columns cb
        // of the new W rows, I_S - G[S,cb] (wbase = 128: the two halves run on
        // different SIMDs; rows r, r + wbase, ... per thread).  The inner index
        // runs in chunks of four whose loads are issued together (rows beyond cols
        // are read from inside the buffer and replaced by zero).
        // In device memory (n up to 288 with wbase = 128) a thread has up to
        // three rows; their G0 loads, and the loads of each chunk of the inner
        // index, are issued together so that a factor step waits for one round
        // trip per chunk instead of one per row (same arithmetic, same order).
        constexpr int kRows = kGlb ? 3 : 1;
        const bool urow = tid < wbase;
        // urow: v = U[r,j], p = W[j,s];  else: v = W[j,cb], p = U[s,j]
        const auto vb = urow ? Ut : W;
        const auto pb = urow ? W : Ut;
        for (int xb = urow ? tid : tid - wbase; tid < 2 * wbase && xb < n; xb += kRows * wbase) {   // r or cb
            cd a0[kRows], a1[kRows];
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const int x = xb + r * wbase < n ? xb + r * wbase : xb;      // rows beyond n repeat row xb
                a0[r] = ldc(G, urow ? x * n + s0 : s0 * n + x);              // G0[r,s0] or G0[s0,cb]
                a1[r] = cd{0.0, 0.0};
                if (pairing) a1[r] = ldc(G, urow ? x * n + s1 : s1 * n + x);
            }
            for (int j0 = 0; j0 < cols; j0 += 4) {
                cd v[kRows][4], p0[4], p1[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = j0 + u < dk ? j0 + u : dk - 1;
                    p0[u] = ldc(pb, j * n + s0);
                    p1[u] = cd{0.0, 0.0};
                    if (pairing) p1[u] = ldc(pb, j * n + s1);
#pragma unroll
                    for (int r = 0; r < kRows; ++r) {
                        const int x = xb + r * wbase < n ? xb + r * wbase : xb;
                        v[r][u] = ldc(vb, j * n + x);
                    }
                }
#pragma unroll
                for (int r = 0; r < kRows; ++r)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const cd vv = csel(j0 + u < cols, v[r][u], cd{0.0, 0.0});
                        a0[r] = a0[r] - vv * p0[u];
                        if (pairing) a1[r] = a1[r] - vv * p1[u];
                    }
            }
#pragma unroll
            for (int r = 0; r < kRows; ++r) {
                const int x = xb + r * wbase;
                if (x >= n) break;
                if (urow) {
                    stc(Ut, cols * n + x, a0[r] * m00 + a1[r] * m10);
                    if (pairing) stc(Ut, (cols + 1) * n + x, a0[r] * m01 + a1[r] * m11);
                } else {
                    stc(W, cols * n + x, cd{(x == s0 ? 1.0 : 0.0) - a0[r].r, -a0[r].i});
                    if (pairing) stc(W, (cols + 1) * n + x, cd{(x == s1 ? 1.0 : 0.0) - a1[r].r, -a1[r].i});
                }
            }
        }
        cols += ksec;
        c.sync();
        if (cols + ksec > dk) {
            if (kUW) delayed_flush_glb_uw(c, G_, du, dw, n, cols);
            else if (kGlb) delayed_flush_glb(c, G_, du, dw, n, cols);
            else delayed_flush(c, G_, du, dw, n, cols);
            cols = 0;
        }
    }
    if (cols > 0) {
        // wavefronts still deciding the trailing rejected sites read G: they
        // must be done before any wavefront's flush writes its tiles
        c.sync();
        if (kUW) delayed_flush_glb_uw(c, G_, du, dw, n, cols);
        else if (kGlb) delayed_flush_glb(c, G_, du, dw, n, cols);
        else delayed_flush(c, G_, du, dw, n, cols);
    }
    for (int i = tid; i < ns; i += c.nthreads()) {
        auto pp = props + i;
        if (pp->pad != 0.0) phi_t[i * 2 + nf - 1] = pp->phi_new;
    }
    if (tid == 0) {
        ch.stats[nf == 1 ? S_PROP_U1 : S_PROP_U2] += (double)ns;
        ch.stats[nf == 1 ? S_ACC_U1 : S_ACC_U2] += (double)nacc;
    }
    c.sync();
}

DEVFN void site_loop_delayed(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G, int t,
                             int nf, bool ascending, int& seed) {
    site_loop_delayed_t<false>(c, m, ch, L, G, t, nf, ascending, seed);
}
DEVFN void site_loop_delayed_glb(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G, int t,
                                 int nf, bool ascending, int& seed) {
    if (m.uw_glb) site_loop_delayed_t<true, true>(c, m, ch, L, G, t, nf, ascending, seed);
    else site_loop_delayed_t<true, false>(c, m, ch, L, G, t, nf, ascending, seed);
}

// U = G0[:, A] X for the accepted rows of a block (device-memory path): Ut[j][x]
// = sum_a G[x][row_a] X[a][j], x < n, j < na, on FP64 MFMA 16x16x4 (complex by
// three real products); X (na x na, leading dimension sb) in LDS, the rows
// row_a = grow(loc[a]) of G gathered from device memory.
// With a Part, only the row tiles tx = pt.i, pt.i + pt.n, ... of U (those of
// the flush's share).
template <class GP, class UP, class RowOf>
DEV void sub_form_u(const Ctx& c, UP Ut, GP G, LDS_AS const cd* X, int sb, LDS_AS const int* loc, int n, int na,
                    RowOf grow, Part pt = Part{}) {
    const int lane = c.lane(), li = lane & 15, lk = lane >> 4;
    const int ntx = (n + 15) >> 4, ntj = (na + 15) >> 4, nk = (na + 3) >> 2;
    const int nrt = (ntx - pt.i + pt.n - 1) / pt.n;
    for (int t = c.wave(); t < nrt * ntj; t += c.nwaves()) {
        const int tx = pt.i + pt.n * (t / ntj), tj = t % ntj;
        const int ar = 16 * tx + li, bc = 16 * tj + li;
        d4 rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ss = rr;
        for (int k0 = 0; k0 < nk; k0 += 4) {
            cd a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = 4 * (k0 + u) + lk;
                a[u] = cd{0.0, 0.0};
                b[u] = cd{0.0, 0.0};
                if (k < na && ar < n) a[u] = ldc(G, ar * n + grow(loc[k]));
                if (k < na && bc < na) b[u] = ldc(X, k * sb + bc);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (k0 + u < nk) {
                    rr = c.mfma(a[u].r, b[u].r, rr);
                    ii = c.mfma(a[u].i, b[u].i, ii);
                    ss = c.mfma(a[u].r + a[u].i, b[u].r + b[u].i, ss);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = 16 * tx + mfma_row(lane, r), j = 16 * tj + li;
            if (x < n && j < na) stc(Ut, j * n + x, cd{rr[r] - ii[r], ss[r] - rr[r] - ii[r]});
        }
    }
}

// Local updates of one slice and channel on the device-memory path by
// submatrix delayed updates: the same decisions, RNG use and result as
// site_loop_t (LocalU_metro + LocalU_prop_L/R).  The sites are taken in blocks
// of sb / ksec in visiting order (sb = dk).  For a block with rows R (each site
// its row s0 = i, and for pairing s1 = Ns + i):
//   * G[R,R] is gathered into LDS and kept as G0[R,R];
//   * each site is decided from the current G[R,R] (every wavefront decides it
//     itself, so a rejected site costs no barrier), and an accepted site
//     updates only G[R,R] <- G[R,R] - G[R,S] M (1 - G)[S,R] (ping-pong
//     buffers, one barrier);
//   * at the end of the block the accepted rows A with their changes Delta_A
//     update G in one rank-|A| step, G <- G - G0[:,A] X (1 - G0)[A,:] with
//     X = (I + Delta_A (1 - G0)[A,A])^{-1} Delta_A (Woodbury for the combined
//     change of the accepted fields; the sequence of the sites' own updates
//     G <- G - G[:,S] M (1 - G)[S,:], M = (1 + Delta_S (1 - G)[S,S])^{-1} Delta_S,
//     gives the same G).  The inverse is a Gauss-Jordan sweep in acceptance
//     order, whose pivots are the accepted sites' P (nonzero), U = G0[:,A] X
//     runs on the matrix cores, then the flush G -= U W with W = (1 - G0)[A,:].
// The per-site work is O(|R|^2) in LDS instead of O(N |A|) in device memory.
// In a team (two workgroups on one chain) both members make every decision;
// U's row tiles and the flush are split between them, W is formed whole by
// each member when it is in LDS and in halves when in device memory, and
// team barriers separate the reads of G of a block from the flush and the
// flush from the next block's reads.  The caller has passed a team barrier
// since its last write to G.
template <bool kUW>
DEV void site_loop_sub_t(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G_, int t, int nf,
                         bool ascending, int& seed, const Team& team) {
    const int n = m.n, ns = m.ns, sb = m.dk;
    const bool pairing = m.kind == 1;
    const int ksec = pairing ? 2 : 1;
    const int bsites = sb / ksec;
    const int tid = c.tid(), nth = c.nthreads();
    auto G = as_global(G_);
    cd* const du = kUW ? ch.w2 : L.du;
    cd* const dw = kUW ? ch.w3 : L.dw;
    auto Ut = [&] {
        if constexpr (kUW) return as_global(du);
        else return as_lds(du);
    }();
    auto W = [&] {
        if constexpr (kUW) return as_global(dw);
        else return as_lds(dw);
    }();
    double* phi_t = ch.phi + (size_t)t * ns * 2;
    make_proposals(c, m, L, phi_t, nf, ascending, seed);
    seed = lcg_jump(seed, 2 * ns);
    auto props = as_lds(L.prop);
    LDS_AS cd* const g0 = as_lds(L.sg0);
    LDS_AS int* const loc = as_lds(L.sloc);
    LDS_AS cd* const dl = as_lds(L.sd);
    int nacc = 0;
    for (int q0 = 0; q0 < ns; q0 += bsites) {
        const int nbs = ns - q0 < bsites ? ns - q0 : bsites;
        const int nr = nbs * ksec;
        // global row of local row lr of R: sector lr % ksec of the block's site lr / ksec
        auto grow = [&](int lr) {
            const int qq = lr / ksec, sec = lr - qq * ksec, q = q0 + qq;
            return sec * ns + (ascending ? q : ns - 1 - q);
        };
        LDS_AS cd* cur = as_lds(L.sbuf);
        LDS_AS cd* nxt = cur + sb * sb;
        c.sync();                                    // the previous block's readers are done
        for (int e = tid; e < nr * nr; e += nth) {
            const int r = e / nr, k = e - r * nr;
            const cd v = ldc(G, grow(r) * n + grow(k));
            stc(g0, r * sb + k, v);
            stc(cur, r * sb + k, v);
        }
        c.sync();
        int na = 0;
        for (int qq = 0; qq < nbs; ++qq) {
            const int q = q0 + qq;
            const int i = ascending ? q : ns - 1 - q;
            const int l0 = ksec * qq, l1 = l0 + 1;
            Proposal P;
            {
                auto pp = props + i;
                P.phi_new = pp->phi_new;
                P.ratio0 = pp->ratio0;
                P.u = pp->u;
                P.d0 = cd{pp->d0.r, pp->d0.i};
                P.d1 = cd{pp->d1.r, pp->d1.i};
            }
            const cd g00 = ldc(cur, l0 * sb + l0);
            bool acc;
            cd p00, p01 = cd{0.0, 0.0}, p10 = cd{0.0, 0.0}, p11 = cd{1.0, 0.0}, det;
            if (!pairing) {
                det = cd{1.0, 0.0} + P.d0 * (cd{1.0, 0.0} - g00);
                p00 = det;
                acc = fabs(P.ratio0 / abs2(det)) > P.u;
            } else {
                const cd g01 = ldc(cur, l0 * sb + l1);
                const cd g10 = ldc(cur, l1 * sb + l0);
                const cd g11 = ldc(cur, l1 * sb + l1);
                p00 = cd{1.0, 0.0} + P.d0 * (cd{1.0, 0.0} - g00);
                p01 = -(P.d0 * g01);
                p10 = -(P.d1 * g10);
                p11 = cd{1.0, 0.0} + P.d1 * (cd{1.0, 0.0} - g11);
                det = p00 * p11 - p01 * p10;
                acc = fabs(P.ratio0 / sqrt(abs2(det))) > P.u;
            }
            if (!acc) continue;
            ++nacc;
            cd m00, m01 = cd{0.0, 0.0}, m10 = cd{0.0, 0.0}, m11 = cd{0.0, 0.0};
            const cd idet = cinv(det);
            if (!pairing) {
                m00 = idet * P.d0;
            } else {
                m00 = (p11 * idet) * P.d0;
                m01 = (-(p01 * idet)) * P.d1;
                m10 = (-(p10 * idet)) * P.d0;
                m11 = (p00 * idet) * P.d1;
            }
            if (tid == 0) {
                (props + i)->pad = 1.0;                // accepted: phi is written after the loop
                loc[na] = l0;
                stc(dl, na, P.d0);
                if (pairing) {
                    loc[na + 1] = l1;
                    stc(dl, na + 1, P.d1);
                }
            }
            // G[R,R] <- G[R,R] - G[R,S] M (1 - G)[S,R]
            for (int e = tid; e < nr * nr; e += nth) {
                const int r = e / nr, k = e - r * nr;
                const cd h0 = cd{l0 == k ? 1.0 : 0.0, 0.0} - ldc(cur, l0 * sb + k);
                cd v = ldc(cur, r * sb + k) - ldc(cur, r * sb + l0) * (m00 * h0);
                if (pairing) {
                    const cd h1 = cd{l1 == k ? 1.0 : 0.0, 0.0} - ldc(cur, l1 * sb + k);
                    v = v - ldc(cur, r * sb + l0) * (m01 * h1) - ldc(cur, r * sb + l1) * (m10 * h0 + m11 * h1);
                }
                stc(nxt, r * sb + k, v);
            }
            na += ksec;
            c.sync();
            LDS_AS cd* const tmp = cur;
            cur = nxt;
            nxt = tmp;
        }
        if (na == 0) continue;
        // ---- X = (I + Delta_A (1 - G0)[A,A])^{-1} Delta_A by a Gauss-Jordan sweep
        c.sync();                                    // every wavefront's decisions of the block are done
        LDS_AS cd* A = as_lds(L.sbuf);
        LDS_AS cd* B = A + sb * sb;
        for (int e = tid; e < na * na; e += nth) {
            const int a = e / na, b = e - a * na;
            const int la = loc[a], lb = loc[b];
            const cd h = cd{la == lb ? 1.0 : 0.0, 0.0} - ldc(g0, la * sb + lb);
            stc(A, a * sb + b, cd{a == b ? 1.0 : 0.0, 0.0} + ldc(dl, a) * h);
        }
        c.sync();
        for (int p = 0; p < na; ++p) {
            const cd ip = cinv(ldc(A, p * sb + p));
            for (int e = tid; e < na * na; e += nth) {
                const int a = e / na, b = e - a * na;
                cd v;
                if (a == p && b == p) v = ip;
                else if (a == p) v = ldc(A, p * sb + b) * ip;
                else if (b == p) v = -(ldc(A, a * sb + p) * ip);
                else v = ldc(A, a * sb + b) - ldc(A, a * sb + p) * (ldc(A, p * sb + b) * ip);
                stc(B, a * sb + b, v);
            }
            c.sync();
            LDS_AS cd* const tmp = A;
            A = B;
            B = tmp;
        }
        // X = inverse Delta_A (columns scaled), into B
        for (int e = tid; e < na * na; e += nth) {
            const int a = e / na, b = e - a * na;
            stc(B, a * sb + b, ldc(A, a * sb + b) * ldc(dl, b));
        }
        c.sync();
        // ---- U = G0[:,A] X and W = (1 - G0)[A,:] from G (still G0), then the flush
        sub_form_u(c, Ut, G, B, sb, loc, n, na, grow, team.share());
        const Part wp = kUW ? team.share() : Part{};
        for (int e = tid + wp.i * nth; e < na * n; e += wp.n * nth) {
            const int j = e / n, x = e - j * n;
            const int row = grow(loc[j]);
            const cd g = ldc(G, row * n + x);
            stc(W, j * n + x, cd{(x == row ? 1.0 : 0.0) - g.r, -g.i});
        }
        team_sync(c, team);                          // every read of G of this block is done
        if (kUW) delayed_flush_glb_uw(c, G_, du, dw, n, na, team.share());
        else delayed_flush_glb(c, G_, du, dw, n, na, team.share());
        if (team.parts > 1) team_sync(c, team);      // the flush is done before the next block's reads
    }
    if (team.lead()) {
        for (int i = tid; i < ns; i += nth) {
            auto pp = props + i;
            if (pp->pad != 0.0) phi_t[i * 2 + nf - 1] = pp->phi_new;
        }
        if (tid == 0) {
            ch.stats[nf == 1 ? S_PROP_U1 : S_PROP_U2] += (double)ns;
            ch.stats[nf == 1 ? S_ACC_U1 : S_ACC_U2] += (double)nacc;
        }
    }
    team_sync(c, team);                              // fields written, last reads of G done
}

DEVFN void site_loop_sub_glb(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* G, int t, int nf,
                             bool ascending, int& seed, const Team& team) {
    if (m.uw_glb) site_loop_sub_t<true>(c, m, ch, L, G, t, nf, ascending, seed, team);
    else site_loop_sub_t<false>(c, m, ch, L, G, t, nf, ascending, seed, team);
}

// ---------------------------------------------------------------- MFMA GEMM in LDS
// M (LDS, n x n complex) <- E M (left) or M E (right) with E real, read in
// MFMA operand order from global memory (coalesced 512 B per k-step).  Each
// wavefront keeps its output tiles in registers until every read of M is done.
DEVFN void lds_gemm(const Ctx& c, cd* M_, const double* Ez, int n, bool left) {
    auto M = as_lds(M_);
    const int nt = (n + 15) >> 4, nk = (n + 3) >> 2, ntt = nt * nt;
    const int lane = c.lane(), li = lane & 15, lk = lane >> 4, w = c.wave(), nw = c.nwaves();
    d4 ar[kMaxTiles], ai[kMaxTiles];
#pragma unroll
    for (int q = 0; q < kMaxTiles; ++q) {
        ar[q] = d4{0.0, 0.0, 0.0, 0.0};
        ai[q] = d4{0.0, 0.0, 0.0, 0.0};
        const int t = w + q * nw;
        if (t < ntt) {
            const int ti = t / nt, tj = t - ti * nt;
            // propagator fragments are fetched kChunk k-steps at a time so the
            // global-memory latency is paid once per chunk, not per MFMA
            constexpr int kChunk = 8;
            if (left) {
                const int col = 16 * tj + li;
                auto ez = as_global(Ez + (size_t)ti * nk * 64 + lane);
                for (int k0 = 0; k0 < nk; k0 += kChunk) {
                    double eb[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) eb[u] = (k0 + u < nk) ? ez[(size_t)(k0 + u) * 64] : 0.0;
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) {
                        if (k0 + u < nk) {
                            const int k = 4 * (k0 + u) + lk;
                            const cd b = (k < n && col < n) ? ldc(M, k * n + col) : cd{0.0, 0.0};
                            ar[q] = c.mfma(eb[u], b.r, ar[q]);
                            ai[q] = c.mfma(eb[u], b.i, ai[q]);
                        }
                    }
                }
            } else {
                const int row = 16 * ti + li;
                auto ez = as_global(Ez + (size_t)tj * nk * 64 + lane);
                for (int k0 = 0; k0 < nk; k0 += kChunk) {
                    double eb[kChunk];
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) eb[u] = (k0 + u < nk) ? ez[(size_t)(k0 + u) * 64] : 0.0;
#pragma unroll
                    for (int u = 0; u < kChunk; ++u) {
                        if (k0 + u < nk) {
                            const int k = 4 * (k0 + u) + lk;
                            const cd a = (row < n && k < n) ? ldc(M, row * n + k) : cd{0.0, 0.0};
                            ar[q] = c.mfma(a.r, eb[u], ar[q]);
                            ai[q] = c.mfma(a.i, eb[u], ai[q]);
                        }
                    }
                }
            }
        }
    }
    c.sync();
#pragma unroll
    for (int q = 0; q < kMaxTiles; ++q) {
        const int t = w + q * nw;
        if (t < ntt) {
            const int ti = t / nt, tj = t - ti * nt;
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
                if (row < n && col < n) stc(M, row * n + col, cd{ar[q][r], ai[q][r]});
            }
        }
    }
    c.sync();
}

// ---------------------------------------------------------------- stabilization
// QRP of A (already holding U D or U^H D) and the new factors; A may be LDS.
// stab_UR (src/number_conserving/src/stabilization.f90:69-93).
// ---- the stabilisation on a team (device-memory path): the pivoted QR runs
// in the leader, which leaves tau, D and the pivots (4n doubles) in the
// chain's work matrix w2 for its partner (w2 is free from the QR until the
// partner has read them); the products, the reflector blocks and the
// triangular solve are split by tiles or column tiles.  Each member's blocked
// routines use a work buffer of their own in w3 (blk_work_elems each), so the
// split form needs the blocked routines and both buffers in w3.
DEV bool team_wrap_fits(int n) { return n >= kBlkMinN && 2 * blk_work_elems(n) <= (size_t)n * n; }

DEV cd* team_work(const ChainDev& ch, const Team& team, int n) {
    return ch.w3 + (size_t)team.part * blk_work_elems(n);
}

DEV void qrp_team(const Ctx& c, const ChainDev& ch, const Lds& L, cd* A, int n, const Team& team) {
    double* pub = (double*)ch.w2;
    if (team.lead()) {
        qrp(c, A, n, L.qr, L.dtmp);
        for (int k = c.tid(); k < n; k += c.nthreads()) {
            pub[2 * k] = L.qr.tau[k].r;
            pub[2 * k + 1] = L.qr.tau[k].i;
            pub[2 * n + k] = L.dtmp[k];
            pub[3 * n + k] = (double)L.qr.perm[k];
        }
    }
    team_sync(c, team);
    if (!team.lead()) {
        for (int k = c.tid(); k < n; k += c.nthreads()) {
            L.qr.tau[k] = cd{pub[2 * k], pub[2 * k + 1]};
            L.dtmp[k] = pub[2 * n + k];
            L.qr.perm[k] = (int)pub[3 * n + k];
        }
    }
    c.sync();
}

DEVFN void stab_ur_team(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* A, const Team& team) {
    const int n = m.n;
    const Part sh = team.share();
    qrp_team(c, ch, L, A, n, team);
    trmm_perm_rows(c, ch.w1, A, ch.vr, L.qr.perm, n, sh);
    team_sync(c, team);
    mat_copy(c, ch.vr, ch.w1, n, false, sh);
    form_q(c, ch.ur, A, L.qr.tau, n, team_work(ch, team, n), sh);
    if (team.lead()) vec_copy(c, ch.dr, L.dtmp, n);
    team_sync(c, team);
}

DEVFN void stab_ul_team(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* A, const Team& team) {
    const int n = m.n;
    const Part sh = team.share();
    qrp_team(c, ch, L, A, n, team);
    perm_cols_times_vh(c, ch.w1, ch.vl, A, L.qr.perm, n, sh);
    team_sync(c, team);
    mat_copy(c, ch.vl, ch.w1, n, false, sh);
    team_sync(c, team);
    form_q(c, ch.w1, A, L.qr.tau, n, team_work(ch, team, n), sh);
    team_sync(c, team);
    mat_copy(c, ch.ul, ch.w1, n, true, sh);
    if (team.lead()) vec_copy(c, ch.dl, L.dtmp, n);
    team_sync(c, team);
}

DEVFN void stab_ur_from(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* A) {
    const int n = m.n;
    const bool cols = col_path(c, n, A == L.mat);
    if (cols) qrp_cols(c, A, n, L.qr, L.dtmp);
    else qrp(c, A, n, L.qr, L.dtmp);
    trmm_perm_rows(c, ch.w1, A, ch.vr, L.qr.perm, n);
    mat_copy(c, ch.vr, ch.w1, n, false);
    if (cols) form_q_cols(c, ch.ur, A, L.qr.tau, n, false);
    else form_q(c, ch.ur, A, L.qr.tau, n, ch.w3);
    vec_copy(c, ch.dr, L.dtmp, n);
}

// stab_UL (stabilization.f90:95-119).
DEVFN void stab_ul_from(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* A) {
    const int n = m.n;
    const bool cols = col_path(c, n, A == L.mat);
    if (cols) qrp_cols(c, A, n, L.qr, L.dtmp);
    else qrp(c, A, n, L.qr, L.dtmp);
    perm_cols_times_vh(c, ch.w1, ch.vl, A, L.qr.perm, n);
    mat_copy(c, ch.vl, ch.w1, n, false);
    if (cols) {
        form_q_cols(c, ch.ul, A, L.qr.tau, n, true);
    } else {
        form_q(c, ch.w1, A, L.qr.tau, n, ch.w3);
        mat_copy(c, ch.ul, ch.w1, n, true);
    }
    vec_copy(c, ch.dl, L.dtmp, n);
}

// A <- U diag(D) (conj_t = false) or U^H diag(D) (conj_t = true).
DEV void load_ud(const Ctx& c, cd* A, const cd* U, const double* D, int n, bool conj_t, Part pt = Part{}) {
    const int nn = n * n;
    for (int e = c.tid() + pt.i * c.nthreads(); e < nn; e += pt.n * c.nthreads()) {
        const int r = e / n, k = e - r * n;
        const cd u = conj_t ? conj(U[k * n + r]) : U[e];
        A[e] = u * cd{D[k], 0.0};
    }
    c.sync();
}

// stab_green (stabilization.f90:121-190): Gout = (1 - UR DR VR VL DL UL)^{-1}.
// For nt > Ltrot/2 the reference decomposes M^H; then G^H = UR P V^{-1} D^{-1} Q^H UL
// has the same structure as the nt <= Ltrot/2 branch with (UL^H, UR^H) -> (UR, UL).
// A is the QR work matrix (LDS or global).  Returns log|det G| = -sum log D.
DEVFN double stab_green(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt,
                      const cd* ur, const double* dr, const cd* vr, const cd* ul, const double* dl,
                      const cd* vl, cd* A, cd* Gout) {
    const int n = m.n, nn = n * n;
    const bool upper = nt > m.ltrot / 2;
    gemm_cc(c, ch.w1, vr, false, vl, false, n);          // VR VL
    gemm_cc(c, ch.w2, ur, true, ul, true, n);            // UR^H UL^H
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, k = e - r * n;
        if (!upper) {
            A[e] = ch.w2[e] - cd{dr[r] * dl[k], 0.0} * ch.w1[e];
        } else {
            const int t = k * n + r;
            A[e] = conj(ch.w2[t] - cd{dr[k] * dl[r], 0.0} * ch.w1[t]);
        }
    }
    c.sync();
    const bool cols = col_path(c, n, A == L.mat);
    if (cols) qrp_cols(c, A, n, L.qr, L.dtmp);
    else qrp(c, A, n, L.qr, L.dtmp);
    // Y = P V^{-1} D^{-1} Q^H Rt, Rt = UR^H (lower branch) or UL (upper branch)
    if (cols) {
        solve_cols(c, ch.w2, upper ? ul : ur, !upper, A, L.qr.tau, L.dtmp, L.qr.perm, n);
    } else {
        mat_copy(c, ch.w1, upper ? ul : ur, n, !upper);
        apply_qh_left(c, ch.w1, A, L.qr.tau, n, ch.w3);
        for (int e = c.tid(); e < nn; e += c.nthreads()) {
            const int r = e / n;
            ch.w1[e] = ch.w1[e] * cd{1.0 / L.dtmp[r], 0.0};
        }
        c.sync();
        trsm_upper_left(c, ch.w1, A, n, ch.w3);
        for (int e = c.tid(); e < nn; e += c.nthreads()) {
            const int r = e / n, k = e - r * n;
            ch.w2[L.qr.perm[r] * n + k] = ch.w1[e];
        }
        c.sync();
    }
    // G' = Lf Y, Lf = UL^H (lower) or UR (upper); G = G' or G'^H
    if (!upper) {
        gemm_cc(c, Gout, ul, true, ch.w2, false, n);
    } else {
        gemm_cc(c, ch.w3, ur, false, ch.w2, false, n);
        mat_copy(c, Gout, ch.w3, n, true);
    }
    double s = 0.0;
    for (int k = c.lane(); k < n; k += 64) s += log(L.dtmp[k]);
    s = wave_sum(c, s);
    return -s;
}

// stab_green on a team: the same steps as stab_green's device-memory form.
DEVFN double stab_green_team(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt,
                             const cd* ur, const double* dr, const cd* vr, const cd* ul, const double* dl,
                             const cd* vl, cd* A, cd* Gout, const Team& team) {
    const int n = m.n, nn = n * n;
    const Part sh = team.share();
    const bool upper = nt > m.ltrot / 2;
    gemm_cc(c, ch.w1, vr, false, vl, false, n, sh);
    gemm_cc(c, ch.w2, ur, true, ul, true, n, sh);
    team_sync(c, team);
    for (int e = c.tid() + sh.i * c.nthreads(); e < nn; e += sh.n * c.nthreads()) {
        const int r = e / n, k = e - r * n;
        if (!upper) {
            A[e] = ch.w2[e] - cd{dr[r] * dl[k], 0.0} * ch.w1[e];
        } else {
            const int t = k * n + r;
            A[e] = conj(ch.w2[t] - cd{dr[k] * dl[r], 0.0} * ch.w1[t]);
        }
    }
    team_sync(c, team);
    qrp_team(c, ch, L, A, n, team);
    mat_copy(c, ch.w1, upper ? ul : ur, n, !upper, sh);
    team_sync(c, team);
    cd* const work = team_work(ch, team, n);
    apply_qh_left(c, ch.w1, A, L.qr.tau, n, work, sh);
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, k = e - r * n;
        if (own_col(sh, k)) ch.w1[e] = ch.w1[e] * cd{1.0 / L.dtmp[r], 0.0};
    }
    c.sync();
    trsm_upper_left(c, ch.w1, A, n, work, sh);
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, k = e - r * n;
        if (own_col(sh, k)) ch.w2[L.qr.perm[r] * n + k] = ch.w1[e];
    }
    team_sync(c, team);
    if (!upper) {
        gemm_cc(c, Gout, ul, true, ch.w2, false, n, sh);
    } else {
        gemm_cc(c, ch.w3, ur, false, ch.w2, false, n, sh);
        team_sync(c, team);
        mat_copy(c, Gout, ch.w3, n, true, sh);
    }
    team_sync(c, team);
    double s = 0.0;
    for (int k = c.lane(); k < n; k += 64) s += log(L.dtmp[k]);
    s = wave_sum(c, s);
    return -s;
}

// Drift of the wrapped G against the recomputed one (compare_mat,
// stabilization.f90:256-266) and the last log|det G|.
DEV void store_drift_and_logdet(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L,
                                const cd* Gnew, const cd* Gold, double logdet) {
    const int nn = m.n * m.n;
    double v = 0.0;
    for (int e = c.tid(); e < nn; e += c.nthreads()) v += cabs(Gnew[e] - Gold[e]);
    v = wave_sum(c, v);
    if (c.lane() == 0) L.red[c.wave()] = v;
    c.sync();
    if (c.tid() == 0) {
        double s = 0.0;
        for (int w = 0; w < c.nwaves(); ++w) s += L.red[w];
        // the reference sums over its 4Ns Nambu G, whose {c, b+} block is the
        // complex conjugate of the reduced {b, c+} block held here
        if (m.kind == 1) s *= 2.0;
        if (s > ch.stats[S_XMAX]) ch.stats[S_XMAX] = s;
        ch.stats[S_XMEAN_SUM] += s;
        ch.stats[S_NWRAP] += 1.0;
        ch.stats[S_LOGDET] = logdet;
    }
    c.sync();
}

DEV cd* list_mat(cd* base, int st, int n) { return base + (size_t)st * n * n; }
DEV double* list_vec(double* base, int st, int n) { return base + (size_t)st * n; }

DEV void live_identity(const Ctx& c, int n, cd* u, double* d, cd* v) {
    mat_identity(c, u, n);
    mat_identity(c, v, n);
    vec_fill(c, d, n, 1.0);
}

DEV void store_right(const Ctx& c, const ModelDev& m, const ChainDev& ch, int st, Part pt = Part{}) {
    const int n = m.n;
    mat_copy(c, list_mat(ch.urlist, st, n), ch.ur, n, false, pt);
    mat_copy(c, list_mat(ch.vrlist, st, n), ch.vr, n, false, pt);
    if (pt.i == 0) vec_copy(c, list_vec(ch.drlist, st, n), ch.dr, n);
}

DEV void store_left(const Ctx& c, const ModelDev& m, const ChainDev& ch, int st, Part pt = Part{}) {
    const int n = m.n;
    mat_copy(c, list_mat(ch.ullist, st, n), ch.ul, n, false, pt);
    mat_copy(c, list_mat(ch.vllist, st, n), ch.vl, n, false, pt);
    if (pt.i == 0) vec_copy(c, list_vec(ch.dllist, st, n), ch.dl, n);
}

// ---------------------------------------------------------------- measurement
DEV int imj_index(const ModelDev& m, int i, int j) {
    const int xi = i % m.nlx, yi = i / m.nlx, xj = j % m.nlx, yj = j / m.nlx;
    int dx = xi - xj, dy = yi - yj;
    if (dx < 0) dx += m.nlx;
    if (dy < 0) dy += m.nly;
    return dy * m.nlx + dx;
}

// NC estimator, src/number_conserving/src/obser_equal.f90:95-198.
template <class GP>
DEV void measure_nc_t(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, GP G) {
    const int n = m.n, nn = n * n;
    auto red = as_lds(L.red);
    const double lq = (double)m.ns, norm = 1.0 / (lq * lq);
    enum { NV = 22 };
    double v[NV];
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    // v: 0 dens_up 1 dbl 2 sq 3 n2u 4 n2d 5 nsq_up 6 nsq_do 7-8 psf 9-10 sf 11-12 dw
    //    13-14 den_upup 15-16 den_dodo 17-18 den_updo 19 kinetic
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int i = e / n, j = e - i * n;
        const cd gij = ldc(G, e), gji = ldc(G, j * n + i);
        const cd gii = ldc(G, i * n + i), gjj = ldc(G, j * n + j);
        const cd du_i = gii - cd{1.0, 0.0}, du_j = gjj - cd{1.0, 0.0};
        const cd dd_i = conj(gii) - cd{1.0, 0.0}, dd_j = conj(gjj) - cd{1.0, 0.0};
        const double dl = (i == j) ? 1.0 : 0.0;
        const cd gupc = gji - cd{dl, 0.0};           // Grupc(i,j)
        const cd gdoc = conj(gji) - cd{dl, 0.0};     // Grdoc(i,j)
        const cd cu = du_i * du_j + gupc * gij;
        const cd cdn = dd_i * dd_j + gdoc * conj(gij);
        const cd cud = du_i * dd_j;
        v[5] += cu.r;
        v[6] += cdn.r;
        const cd psf = gupc * gdoc;
        v[7] += psf.r; v[8] += psf.i;
        if (m.has_k) {
            const cd ph = m.phasek[imj_index(m, i, j)];
            const cd sf = ph * (gupc + gdoc);
            v[9] += sf.r; v[10] += sf.i;
            const cd dw = ph * (cu + cdn + cud + dd_i * du_j);
            v[11] += dw.r; v[12] += dw.i;
        }
        v[13] += cu.r; v[14] += cu.i;
        v[15] += cdn.r; v[16] += cdn.i;
        v[17] += cud.r; v[18] += cud.i;
        if (i == j) {
            v[0] += du_i.r;
            v[1] += (du_i * dd_i).r;
            v[2] += (du_i * du_i + dd_i * dd_i).r;
            v[3] += (cd{2.0, 0.0} * du_i * du_i + du_i).r;
            v[4] += (cd{2.0, 0.0} * dd_i * dd_i + dd_i).r;
            ch.obs[O_SITE0 + i] += (du_i + dd_i).r;
        }
    }
    for (int e = c.tid(); e < m.ns * 3; e += c.nthreads()) {
        const int i = e / 3, j = m.bonds[e];
        const cd gij = ldc(G, i * n + j), gji = ldc(G, j * n + i);
        // Grupc(i,j)+Grupc(j,i)+Grdoc(i,j)+Grdoc(j,i), i != j
        v[19] += (gji + gij + conj(gji) + conj(gij)).r / lq;
    }
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(c, v[k]);
    if (c.lane() == 0)
        for (int k = 0; k < NV; ++k) red[c.wave() * NV + k] = v[k];
    c.sync();
    if (c.tid() == 0) {
        double s[NV];
        for (int k = 0; k < NV; ++k) {
            s[k] = 0.0;
            for (int w = 0; w < c.nwaves(); ++w) s[k] += red[w * NV + k];
        }
        double* o = ch.obs;
        const double dens = s[0] / lq;
        const double dbl = s[1] / lq;
        const double inter = (m.u1 + m.u2) * (s[3] + s[4]) / lq + 2.0 * (m.u2 - m.u1) * dbl;
        o[O_DENSITY_UP] += dens;
        o[O_DENSITY_DO] += dens;
        o[O_DENSITY_TOTAL] += 2.0 * dens;
        o[O_DOUBLEOCC] += dbl;
        o[O_SQUAREOCC] += s[2] / lq;
        o[O_NUM_UP] += s[0];
        o[O_NUM_DO] += s[0];
        o[O_NUMSQUARE_UP] += s[5];
        o[O_NUMSQUARE_DO] += s[6];
        o[O_ONSITE_N2_UP] += s[3];
        o[O_ONSITE_N2_DO] += s[4];
        o[O_PSF_RE] += s[7] * norm; o[O_PSF_IM] += s[8] * norm;
        o[O_SF_K_RE] += s[9] * norm; o[O_SF_K_IM] += s[10] * norm;
        o[O_DW_K_RE] += s[11] * norm; o[O_DW_K_IM] += s[12] * norm;
        o[O_DEN_UPUP_RE] += s[13] * norm; o[O_DEN_UPUP_IM] += s[14] * norm;
        o[O_DEN_DODO_RE] += s[15] * norm; o[O_DEN_DODO_IM] += s[16] * norm;
        o[O_DEN_UPDO_RE] += s[17] * norm; o[O_DEN_UPDO_IM] += s[18] * norm;
        o[O_KINETIC] += s[19];
        o[O_INTERACTION] += inter;
        o[O_ENERGY] += s[19] + inter;
        o[O_NOBS] += 1.0;
    }
    c.sync();
}

// Pairing estimator, src/pairing/src/obser_equal.f90:112-271, written in the
// reduced block X (basis b, c+) with the 4Ns blocks
//   G_b_b+ = Xbb, G_b_c = -Xbc, G_c+_b+ = Xcb, G_c+_c = -Xcc,
//   G_c_c+ = conj Xbb, G_c_b = -conj Xbc, G_b+_c+ = conj Xcb, G_b+_b = -conj Xcc,
// and the remaining eight blocks zero.
template <class GP>
DEV void measure_pairing_t(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, GP G) {
    const int n = m.n, ns = m.ns;
    const double nsd = (double)ns, norm = 1.0 / (nsd * nsd);
    auto X = [&](int bi, int bj, int i, int j) { return ldc(G, (bi * ns + i) * n + bj * ns + j); };
    auto red = as_lds(L.red);
    auto nbp = as_lds(L.site);
    auto ncp = as_lds(L.site) + ns;
    auto xbbp = as_lds(L.site) + 2 * ns;
    auto xbcp = as_lds(L.site) + 3 * ns;
    auto xcbp = as_lds(L.site) + 4 * ns;
    for (int i = c.tid(); i < ns; i += c.nthreads()) {
        const cd xcc = X(1, 1, i, i);
        stc(nbp, i, -conj(xcc));
        stc(ncp, i, -xcc);
        stc(xbbp, i, X(0, 0, i, i));
        stc(xbcp, i, X(0, 1, i, i));
        stc(xcbp, i, X(1, 0, i, i));
    }
    c.sync();
    enum { NV = 24 };
    double v[NV];
    for (int k = 0; k < NV; ++k) v[k] = 0.0;
    // v: 0 nb 1 nc 2 Dbc1 3 sq 4 Dbb1 5 Dcc1 6 pair 7 nsq_up 8 nsq_do 9-10 sf 11-12 dw
    //    13-14 psf 15-16 den_upup 17-18 den_dodo 19-20 den_updo 21 kinetic
    const int np = ns * ns;
    for (int e = c.tid(); e < np; e += c.nthreads()) {
        const int i = e / ns, j = e - i * ns;
        const cd Xbb = X(0, 0, i, j), Xbc = X(0, 1, i, j), Xcb = X(1, 0, i, j), Xcc = X(1, 1, i, j);
        const cd nbi = ldc(nbp, i), nbj = ldc(nbp, j), nci = ldc(ncp, i), ncj = ldc(ncp, j);
        const cd Dbb = nbi * nbj - conj(Xcc) * Xbb;
        const cd Dcc = nci * ncj - Xcc * conj(Xbb);
        const cd Dbc = nbi * ncj - conj(Xcb) * Xbc;
        const cd Dcb = nci * nbj - Xcb * conj(Xbc);
        v[7] += Dbb.r;
        v[8] += Dcc.r;
        v[15] += Dbb.r; v[16] += Dbb.i;
        v[17] += Dcc.r; v[18] += Dcc.i;
        v[19] += Dbc.r; v[20] += Dbc.i;
        if (m.has_k) {
            const cd ph = m.phasek[imj_index(m, i, j)];
            const cd sf = ph * (-(conj(Xcc) + Xcc));
            v[9] += sf.r; v[10] += sf.i;
            const cd dw = ph * (Dbb + Dcc + Dbc + Dcb);
            v[11] += dw.r; v[12] += dw.i;
        }
        const cd psf = -(conj(ldc(xcbp, i)) * conj(ldc(xbcp, j))) + conj(Xcc) * Xcc;
        v[13] += psf.r; v[14] += psf.i;
        if (i == j) {
            const cd xbbi = ldc(xbbp, i), xbci = ldc(xbcp, i), xcbi = ldc(xcbp, i);
            const cd Dbb1 = nbi * nbi + nbi * xbbi;
            const cd Dcc1 = nci * nci + nci * conj(xbbi);
            const cd Dbc1 = nbi * nci - conj(xcbi) * xbci;
            v[0] += nbi.r;
            v[1] += nci.r;
            v[2] += Dbc1.r;
            v[3] += (nbi * nbi + nci * nci).r;
            v[4] += Dbb1.r;
            v[5] += Dcc1.r;
            v[6] += (-xbci + conj(xcbi)).r;
            ch.obs[O_SITE0 + i] += (nbi + nci).r;
        }
    }
    for (int e = c.tid(); e < ns * 3; e += c.nthreads()) {
        const int i = e / 3, j = m.bonds[e];
        // G_b+_b(i,j)+G_b+_b(j,i)+G_c+_c(i,j)+G_c+_c(j,i) = -2 Re(Xcc_ij + Xcc_ji)
        v[21] += -2.0 * (X(1, 1, i, j) + X(1, 1, j, i)).r / nsd;
    }
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(c, v[k]);
    if (c.lane() == 0)
        for (int k = 0; k < NV; ++k) red[c.wave() * NV + k] = v[k];
    c.sync();
    if (c.tid() == 0) {
        double s[NV];
        for (int k = 0; k < NV; ++k) {
            s[k] = 0.0;
            for (int w = 0; w < c.nwaves(); ++w) s[k] += red[w * NV + k];
        }
        double* o = ch.obs;
        const double dens = (s[0] + s[1]) / nsd;
        const double dbl = s[2] / nsd;
        const double pair = s[6] / nsd;
        const double inter = (m.u1 + m.u2) * (s[4] + s[5]) / nsd + 2.0 * (m.u2 - m.u1) * dbl;
        const double pe = m.rdelta * pair;
        const double che = -m.mu * dens;
        const double kin = s[21];
        o[O_DENSITY_UP] += s[0] / nsd;
        o[O_DENSITY_DO] += s[1] / nsd;
        o[O_DENSITY] += dens;
        o[O_DENSITY_TOTAL] += dens;
        o[O_NUM_UP] += s[0];
        o[O_NUM_DO] += s[1];
        o[O_DOUBLEOCC] += dbl;
        o[O_SQUAREOCC] += s[3] / nsd;
        o[O_LOCAL_NUMSQUARE] += (s[4] + s[5]) / nsd;
        o[O_PAIR_EQUAL] += pair;
        o[O_NUMSQUARE_UP] += s[7];
        o[O_NUMSQUARE_DO] += s[8];
        o[O_SF_K_RE] += s[9] * norm; o[O_SF_K_IM] += s[10] * norm;
        o[O_DW_K_RE] += s[11] * norm; o[O_DW_K_IM] += s[12] * norm;
        o[O_PSF_RE] += s[13] * norm; o[O_PSF_IM] += s[14] * norm;
        o[O_DEN_UPUP_RE] += s[15] * norm; o[O_DEN_UPUP_IM] += s[16] * norm;
        o[O_DEN_DODO_RE] += s[17] * norm; o[O_DEN_DODO_IM] += s[18] * norm;
        o[O_DEN_UPDO_RE] += s[19] * norm; o[O_DEN_UPDO_IM] += s[20] * norm;
        o[O_KINETIC] += kin;
        o[O_ONSITE_N2_UP] += s[4];
        o[O_ONSITE_N2_DO] += s[5];
        o[O_INTERACTION] += inter;
        o[O_PAIRING_E] += pe;
        o[O_CHEMICAL_E] += che;
        o[O_ENERGY] += kin + inter + pe;
        o[O_GRAND_E] += kin + inter + pe + che;
        o[O_NOBS] += 1.0;
    }
    c.sync();
}

DEVFN void measure(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, const cd* G) {
    if (m.kind == 0) measure_nc_t(c, m, ch, L, G);
    else measure_pairing_t(c, m, ch, L, G);
}
DEVFN void measure_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, const cd* G) {
    if (m.kind == 0) measure_nc_t(c, m, ch, L, as_lds(G));
    else measure_pairing_t(c, m, ch, L, as_lds(G));
}

// ====================================================================== global path
// Wrap_R (stabilization.f90:330-366), nt % Nwrap == 0, nt != 0.
DEVFN void wrap_r_global(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt,
                         const Team& team = Team{}) {
    const int n = m.n, st = nt / m.nwrap;
    Timer tm;
    if (team.parts > 1) {
        const Part sh = team.share();
        tm.on = team.lead();
        tm.start(c);
        load_ud(c, ch.wa, ch.ur, ch.dr, n, false, sh);
        team_sync(c, team);
        stab_ur_team(c, m, ch, L, ch.wa, team);
        tm.stop(c, ch, S_T_WQR);
        const double ld = stab_green_team(c, m, ch, L, nt, ch.ur, ch.dr, ch.vr, list_mat(ch.ullist, st, n),
                                          list_vec(ch.dllist, st, n), list_mat(ch.vllist, st, n), ch.wa, ch.wa,
                                          team);
        tm.stop(c, ch, S_T_WGREEN);
        if (team.lead()) store_drift_and_logdet(c, m, ch, L, ch.wa, ch.g, ld);
        team_sync(c, team);
        mat_copy(c, ch.g, ch.wa, n, false, sh);
        store_right(c, m, ch, st, sh);
        team_sync(c, team);
        return;
    }
    tm.start(c);
    load_ud(c, ch.wa, ch.ur, ch.dr, n, false);
    stab_ur_from(c, m, ch, L, ch.wa);
    tm.stop(c, ch, S_T_WQR);
    const double ld = stab_green(c, m, ch, L, nt, ch.ur, ch.dr, ch.vr, list_mat(ch.ullist, st, n),
                                 list_vec(ch.dllist, st, n), list_mat(ch.vllist, st, n), ch.wa, ch.wa);
    tm.stop(c, ch, S_T_WGREEN);
    store_drift_and_logdet(c, m, ch, L, ch.wa, ch.g, ld);
    mat_copy(c, ch.g, ch.wa, n, false);
    store_right(c, m, ch, st);
}

// Wrap_L (stabilization.f90:292-328), nt % Nwrap == 0, nt != Ltrot.
DEVFN void wrap_l_global(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt,
                         const Team& team = Team{}) {
    const int n = m.n, st = nt / m.nwrap;
    Timer tm;
    if (team.parts > 1) {
        const Part sh = team.share();
        tm.on = team.lead();
        tm.start(c);
        load_ud(c, ch.wa, ch.ul, ch.dl, n, true, sh);
        team_sync(c, team);
        stab_ul_team(c, m, ch, L, ch.wa, team);
        tm.stop(c, ch, S_T_WQR);
        const double ld = stab_green_team(c, m, ch, L, nt, list_mat(ch.urlist, st, n), list_vec(ch.drlist, st, n),
                                          list_mat(ch.vrlist, st, n), ch.ul, ch.dl, ch.vl, ch.wa, ch.wa, team);
        tm.stop(c, ch, S_T_WGREEN);
        if (team.lead()) store_drift_and_logdet(c, m, ch, L, ch.wa, ch.g, ld);
        team_sync(c, team);
        mat_copy(c, ch.g, ch.wa, n, false, sh);
        store_left(c, m, ch, st, sh);
        team_sync(c, team);
        return;
    }
    tm.start(c);
    load_ud(c, ch.wa, ch.ul, ch.dl, n, true);
    stab_ul_from(c, m, ch, L, ch.wa);
    tm.stop(c, ch, S_T_WQR);
    const double ld = stab_green(c, m, ch, L, nt, list_mat(ch.urlist, st, n), list_vec(ch.drlist, st, n),
                                 list_mat(ch.vrlist, st, n), ch.ul, ch.dl, ch.vl, ch.wa, ch.wa);
    tm.stop(c, ch, S_T_WGREEN);
    store_drift_and_logdet(c, m, ch, L, ch.wa, ch.g, ld);
    mat_copy(c, ch.g, ch.wa, n, false);
    store_left(c, m, ch, st);
}

// A wrap on a team: split where team_wrap_fits, in the leader alone
// otherwise.  The caller's team barrier follows.
DEV void wrap_global_team(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt, bool right,
                          const Team& team) {
    if (team.parts > 1 && !team_wrap_fits(m.n)) {
        if (!team.lead()) return;
        if (right) wrap_r_global(c, m, ch, L, nt);
        else wrap_l_global(c, m, ch, L, nt);
        return;
    }
    if (right) wrap_r_global(c, m, ch, L, nt, team);
    else wrap_l_global(c, m, ch, L, nt, team);
}

// The G rebuild on a team (team_wrap_fits(n)): the slice products and
// scalings split by tiles and elements, the stabilisations as in the wraps.
// mat_copy, scale_mat and load_ud share one element split, so a member
// reads only its own elements between them.
DEV void prepare_global_team(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, const Team& team) {
    const int n = m.n;
    const Part sh = team.share();
    if (team.lead()) {
        live_identity(c, n, ch.ur, ch.dr, ch.vr);
        live_identity(c, n, ch.ul, ch.dl, ch.vl);
        store_right(c, m, ch, 0);
        store_left(c, m, ch, m.nst);
    }
    team_sync(c, team);
    for (int nt = 1; nt <= m.ltrot; ++nt) {
        const double* phi_t = ch.phi + (size_t)(nt - 1) * m.ns * 2;
        for (int nf = 1; nf <= 2; ++nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf);
            scale_mat(c, ch.ur, n, L.f, nullptr, sh);
        }
        team_sync(c, team);
        gemm_rc_panel(c, ch.w1, m.ez[0], ch.ur, 0, n, L.panel, L.panel_cap, sh);
        team_sync(c, team);
        mat_copy(c, ch.ur, ch.w1, n, false, sh);
        if (nt % m.nwrap == 0) {
            load_ud(c, ch.wa, ch.ur, ch.dr, n, false, sh);
            team_sync(c, team);
            stab_ur_team(c, m, ch, L, ch.wa, team);
            store_right(c, m, ch, nt / m.nwrap, sh);
        }
    }
    team_sync(c, team);
    const double ld = stab_green_team(c, m, ch, L, m.ltrot, ch.ur, ch.dr, ch.vr, ch.ul, ch.dl, ch.vl, ch.wa, ch.g,
                                      team);
    if (team.lead() && c.tid() == 0) ch.stats[S_LOGDET] = ld;
    team_sync(c, team);
}

DEV void prepare_global(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L) {
    const int n = m.n;
    live_identity(c, n, ch.ur, ch.dr, ch.vr);
    live_identity(c, n, ch.ul, ch.dl, ch.vl);
    store_right(c, m, ch, 0);
    store_left(c, m, ch, m.nst);
    for (int nt = 1; nt <= m.ltrot; ++nt) {
        const double* phi_t = ch.phi + (size_t)(nt - 1) * m.ns * 2;
        for (int nf = 1; nf <= 2; ++nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf);
            scale_mat(c, ch.ur, n, L.f, nullptr);
        }
        gemm_rc_panel(c, ch.w1, m.ez[0], ch.ur, 0, n, L.panel, L.panel_cap);
        mat_copy(c, ch.ur, ch.w1, n, false);
        if (nt % m.nwrap == 0) {
            load_ud(c, ch.wa, ch.ur, ch.dr, n, false);
            stab_ur_from(c, m, ch, L, ch.wa);
            store_right(c, m, ch, nt / m.nwrap);
        }
    }
    const double ld = stab_green(c, m, ch, L, m.ltrot, ch.ur, ch.dr, ch.vr, ch.ul, ch.dl, ch.vl, ch.wa, ch.g);
    if (c.tid() == 0) ch.stats[S_LOGDET] = ld;
    c.sync();
}

// With a team of two workgroups (submatrix updates only, dk > 0; no grad or
// frozen pass) the products and scalings are split by the team, the wraps
// split all but their pivoted QRs, the estimators run in the leader, and a
// team barrier follows every phase that writes device memory the other
// member reads next.  A single
// workgroup computes the same values in the same order.
DEV void sweep_left_global(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int& seed,
                           bool meas, double* grad = nullptr, bool frozen = false, const Team& team = Team{}) {
    const int n = m.n;
    const Part sh = team.share();
    Timer tm;
    tm.on = team.lead();
    if (team.lead()) live_identity(c, n, ch.ul, ch.dl, ch.vl);   // Wrap_L(Ltrot): L list slot Nst is the identity
    team_sync(c, team);
    tm.start(c);
    for (int nt = m.ltrot; nt >= 1; --nt) {
        if (nt % m.nwrap == 0 && nt != m.ltrot) {
            wrap_global_team(c, m, ch, L, nt, false, team);
            team_sync(c, team);
            tm.stop(c, ch, S_T_WRAP);
        }
        if (meas) {
            if (team.lead()) measure(c, m, ch, L, ch.g);
            tm.stop(c, ch, S_T_MEAS);
        }
        // propT_L: G <- e^{+dtau K} G e^{-dtau K}, UL <- UL e^{-dtau K}
        gemm_rc_panel(c, ch.w1, m.ez[2], ch.g, 1, n, L.panel, L.panel_cap, sh);
        gemm_rc_panel(c, ch.w2, m.ez[2], ch.ul, 1, n, L.panel, L.panel_cap, sh);
        team_sync(c, team);
        gemm_rc_panel(c, ch.g, m.ez[1], ch.w1, 0, n, L.panel, L.panel_cap, sh);
        mat_copy(c, ch.ul, ch.w2, n, false, sh);
        team_sync(c, team);
        tm.stop(c, ch, S_T_KIN);
        const int t = nt - 1;
        const double* phi_t = ch.phi + (size_t)t * m.ns * 2;
        for (int nf = 2; nf >= 1; --nf) {
            if (!m.active[nf - 1]) continue;
            if (grad) grad_extract(c, m, phi_t, as_global(ch.g), t, nf, grad);
            else if (frozen) {}
            else if (m.dk > 0) site_loop_sub_glb(c, m, ch, L, ch.g, t, nf, false, seed, team);
            else site_loop(c, m, ch, L, ch.g, t, nf, false, seed);
            tm.stop(c, ch, S_T_SITE);
            hs_factors(c, m, L, phi_t, nf);
            scale_mat(c, ch.g, n, L.finv, L.f, sh);      // G <- e^{-V} G e^{V}
            scale_mat(c, ch.ul, n, nullptr, L.f, sh);    // UL <- UL e^{V}
            team_sync(c, team);
            tm.stop(c, ch, S_T_HS);
        }
    }
    wrap_global_team(c, m, ch, L, 0, false, team);
    team_sync(c, team);
    tm.stop(c, ch, S_T_WRAP);
}

DEV void sweep_right_global(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int& seed,
                            bool meas, const Team& team = Team{}) {
    const int n = m.n;
    const Part sh = team.share();
    Timer tm;
    tm.on = team.lead();
    if (team.lead()) live_identity(c, n, ch.ur, ch.dr, ch.vr);   // Wrap_R(0): R list slot 0 is the identity
    team_sync(c, team);
    tm.start(c);
    for (int nt = 1; nt <= m.ltrot; ++nt) {
        const int t = nt - 1;
        const double* phi_t = ch.phi + (size_t)t * m.ns * 2;
        for (int nf = 1; nf <= 2; ++nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf);
            scale_mat(c, ch.g, n, L.f, L.finv, sh);      // G <- e^{V} G e^{-V} (old fields)
            team_sync(c, team);
            tm.stop(c, ch, S_T_HS);
            if (m.dk > 0) site_loop_sub_glb(c, m, ch, L, ch.g, t, nf, true, seed, team);
            else site_loop(c, m, ch, L, ch.g, t, nf, true, seed);
            tm.stop(c, ch, S_T_SITE);
            hs_factors(c, m, L, phi_t, nf);
            scale_mat(c, ch.ur, n, L.f, nullptr, sh);    // UR <- e^{V} UR (new fields)
            tm.stop(c, ch, S_T_HS);
        }
        team_sync(c, team);
        // propT_R: G <- e^{-dtau K} G e^{+dtau K}, UR <- e^{-dtau K} UR
        gemm_rc_panel(c, ch.w1, m.ez[0], ch.g, 0, n, L.panel, L.panel_cap, sh);
        gemm_rc_panel(c, ch.w2, m.ez[0], ch.ur, 0, n, L.panel, L.panel_cap, sh);
        team_sync(c, team);
        gemm_rc_panel(c, ch.g, m.ez[3], ch.w1, 1, n, L.panel, L.panel_cap, sh);
        mat_copy(c, ch.ur, ch.w2, n, false, sh);
        team_sync(c, team);
        tm.stop(c, ch, S_T_KIN);
        if (nt % m.nwrap == 0) {
            wrap_global_team(c, m, ch, L, nt, true, team);
            team_sync(c, team);
            tm.stop(c, ch, S_T_WRAP);
        }
        if (meas) {
            if (team.lead()) measure(c, m, ch, L, ch.g);
            if (team.parts > 1) team_sync(c, team);     // the next slice's scaling writes G
            tm.stop(c, ch, S_T_MEAS);
        }
    }
}

// ====================================================================== LDS path
// Right factor products of slices [t0, t1] (1-based, ascending) applied to the
// matrix in LDS: M <- B_t1 ... B_t0 M, B_t = e^{-dtau K} e^{V_1(t)} e^{V_2(t)}
// in the reference's order (propU_R U1 then U2, then propT_R; multiply.f90:80-86).
DEVFN void apply_slices_right(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* M,
                            int t0, int t1) {
    for (int nt = t0; nt <= t1; ++nt) {
        const double* phi_t = ch.phi + (size_t)(nt - 1) * m.ns * 2;
        for (int nf = 1; nf <= 2; ++nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf);
            scale_mat_lds(c, M, m.n, L.f, nullptr);
        }
        lds_gemm(c, M, m.ez[0], m.n, true);
    }
}

// Left factor products of slices t1 down to t0 (sweep-left order):
// M <- M B_t1 ... B_t0 with M B = (M e^{-dtau K}) e^{V_2} e^{V_1} (propT_L, then U2, U1).
DEVFN void apply_slices_left(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, cd* M,
                           int t0, int t1) {
    for (int nt = t1; nt >= t0; --nt) {
        lds_gemm(c, M, m.ez[2], m.n, false);
        const double* phi_t = ch.phi + (size_t)(nt - 1) * m.ns * 2;
        for (int nf = 2; nf >= 1; --nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf);
            scale_mat_lds(c, M, m.n, nullptr, L.f);
        }
    }
}

// Wrap_R at nt with G in LDS (L.mat): UR <- B(nt, nt-Nwrap] UR, stab_UR, stab_green.
DEVFN void wrap_r_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt) {
    const int n = m.n, st = nt / m.nwrap;
    Timer tm;
    tm.start(c);
    global_from_lds(c, ch.g, L.mat, n);                 // wrapped G, for the drift check
    lds_from_global(c, L.mat, ch.ur, n);
    apply_slices_right(c, m, ch, L, L.mat, nt - m.nwrap + 1, nt);
    scale_mat_cols_real(c, L.mat, n, ch.dr);
    tm.stop(c, ch, S_T_WPROD);
    stab_ur_from(c, m, ch, L, L.mat);
    tm.stop(c, ch, S_T_WQR);
    const double ld = stab_green(c, m, ch, L, nt, ch.ur, ch.dr, ch.vr, list_mat(ch.ullist, st, n),
                                 list_vec(ch.dllist, st, n), list_mat(ch.vllist, st, n), L.mat, ch.wa);
    tm.stop(c, ch, S_T_WGREEN);
    store_drift_and_logdet(c, m, ch, L, ch.wa, ch.g, ld);
    lds_from_global(c, L.mat, ch.wa, n);
    store_right(c, m, ch, st);
}

// Wrap_L at nt with G in LDS: UL <- UL B(nt+Nwrap, nt], stab_UL, stab_green.
DEVFN void wrap_l_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int nt) {
    const int n = m.n, st = nt / m.nwrap;
    Timer tm;
    tm.start(c);
    global_from_lds(c, ch.g, L.mat, n);
    lds_from_global(c, L.mat, ch.ul, n);
    apply_slices_left(c, m, ch, L, L.mat, nt + 1, nt + m.nwrap);
    global_from_lds(c, ch.ul, L.mat, n);
    load_ud(c, L.mat, ch.ul, ch.dl, n, true);
    tm.stop(c, ch, S_T_WPROD);
    stab_ul_from(c, m, ch, L, L.mat);
    tm.stop(c, ch, S_T_WQR);
    const double ld = stab_green(c, m, ch, L, nt, list_mat(ch.urlist, st, n), list_vec(ch.drlist, st, n),
                                 list_mat(ch.vrlist, st, n), ch.ul, ch.dl, ch.vl, L.mat, ch.wa);
    tm.stop(c, ch, S_T_WGREEN);
    store_drift_and_logdet(c, m, ch, L, ch.wa, ch.g, ld);
    lds_from_global(c, L.mat, ch.wa, n);
    store_left(c, m, ch, st);
}

DEV void prepare_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L) {
    const int n = m.n;
    live_identity(c, n, ch.ur, ch.dr, ch.vr);
    live_identity(c, n, ch.ul, ch.dl, ch.vl);
    store_right(c, m, ch, 0);
    store_left(c, m, ch, m.nst);
    for (int st = 1; st <= m.nst; ++st) {
        lds_from_global(c, L.mat, ch.ur, n);
        apply_slices_right(c, m, ch, L, L.mat, (st - 1) * m.nwrap + 1, st * m.nwrap);
        scale_mat_cols_real(c, L.mat, n, ch.dr);
        stab_ur_from(c, m, ch, L, L.mat);
        store_right(c, m, ch, st);
    }
    const double ld = stab_green(c, m, ch, L, m.ltrot, ch.ur, ch.dr, ch.vr, ch.ul, ch.dl, ch.vl, L.mat, ch.g);
    if (c.tid() == 0) ch.stats[S_LOGDET] = ld;
    c.sync();
}

DEV void sweep_left_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int& seed,
                        bool meas, double* grad = nullptr, bool frozen = false) {
    const int n = m.n;
    cd* G = L.mat;
    Timer tm;
    live_identity(c, n, ch.ul, ch.dl, ch.vl);     // Wrap_L(Ltrot): L list slot Nst is the identity
    tm.start(c);
    for (int nt = m.ltrot; nt >= 1; --nt) {
        if (nt % m.nwrap == 0 && nt != m.ltrot) {
            wrap_l_lds(c, m, ch, L, nt);
            tm.stop(c, ch, S_T_WRAP);
        }
        if (meas) {
            measure_lds(c, m, ch, L, G);
            tm.stop(c, ch, S_T_MEAS);
        }
        // propT_L: G <- e^{+dtau K} (G e^{-dtau K}); UL's product is applied at the next wrap
        lds_gemm(c, G, m.ez[2], n, false);
        lds_gemm(c, G, m.ez[1], n, true);
        tm.stop(c, ch, S_T_KIN);
        const int t = nt - 1;
        const double* phi_t = ch.phi + (size_t)t * m.ns * 2;
        for (int nf = 2; nf >= 1; --nf) {
            if (!m.active[nf - 1]) continue;
            if (grad) grad_extract(c, m, phi_t, as_lds(G), t, nf, grad);
            else if (frozen) {}
            else if (m.dk > 0) site_loop_delayed(c, m, ch, L, G, t, nf, false, seed);
            else site_loop_lds(c, m, ch, L, G, t, nf, false, seed);
            tm.stop(c, ch, S_T_SITE);
            hs_factors(c, m, L, phi_t, nf);
            scale_mat_lds(c, G, n, L.finv, L.f);         // G <- e^{-V} G e^{V}
            tm.stop(c, ch, S_T_HS);
        }
    }
    wrap_l_lds(c, m, ch, L, 0);
    tm.stop(c, ch, S_T_WRAP);
}

DEV void sweep_right_lds(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int& seed,
                         bool meas) {
    const int n = m.n;
    cd* G = L.mat;
    Timer tm;
    live_identity(c, n, ch.ur, ch.dr, ch.vr);     // Wrap_R(0): R list slot 0 is the identity
    tm.start(c);
    for (int nt = 1; nt <= m.ltrot; ++nt) {
        const int t = nt - 1;
        const double* phi_t = ch.phi + (size_t)t * m.ns * 2;
        for (int nf = 1; nf <= 2; ++nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf);
            scale_mat_lds(c, G, n, L.f, L.finv);         // G <- e^{V} G e^{-V} (old fields)
            tm.stop(c, ch, S_T_HS);
            if (m.dk > 0) site_loop_delayed(c, m, ch, L, G, t, nf, true, seed);
            else site_loop_lds(c, m, ch, L, G, t, nf, true, seed);
            tm.stop(c, ch, S_T_SITE);
        }
        // propT_R: G <- (e^{-dtau K} G) e^{+dtau K}; UR's product is applied at the wrap
        lds_gemm(c, G, m.ez[0], n, true);
        lds_gemm(c, G, m.ez[3], n, false);
        tm.stop(c, ch, S_T_KIN);
        if (nt % m.nwrap == 0) {
            wrap_r_lds(c, m, ch, L, nt);
            tm.stop(c, ch, S_T_WRAP);
        }
        if (meas) {
            measure_lds(c, m, ch, L, G);
            tm.stop(c, ch, S_T_MEAS);
        }
    }
}

}  // namespace bafq

#include "dyn.hpp"

namespace bafq {

// ---------------------------------------------------------------- kernel bodies
// Persistent workgroups take chains from a queue: lane 0 of the workgroup
// increments the launch's counter and shares the index through LDS.  The
// chain list is ordered by decreasing cost, so the first free workgroup always
// takes the most expensive remaining chain (longest-processing-time order).
DEV int next_item(const Ctx& c, int* queue, int* slot) {
    if (c.tid() == 0) *slot = c.atomic_add(queue, 1);
    c.sync();
    const int q = *slot;
    c.sync();
    return q;
}

// The same for a team: the leader takes the index and leaves it in the pair's
// mailbox for its partner.  After a timed-out wait both members stop.
DEV int next_item_team(const Ctx& c, const LaunchArgs& a, int* slot, const Team& team, int* mail) {
    if (team.parts == 1) return next_item(c, a.queue, slot);
    if (team.lead() && c.tid() == 0) c.store_flag(mail, c.atomic_add(a.queue, 1));
    team_sync(c, team);
    if (c.tid() == 0) *slot = c.load_relaxed(team.flags + 1) ? a.count : c.load_relaxed(mail);
    c.sync();
    const int q = *slot;
    c.sync();
    return q;
}

// The team of workgroup c.block() in a launch with a.pairs set.
DEV Team make_team(const Ctx& c, const LaunchArgs& a, int* slot, int** mail) {
    Team team;
    if (!a.pairs) return team;
    int pair, rank;
    Ctx::pair_of(c.block(), c.nblocks(), pair, rank);
    team.part = rank;
    team.parts = 2;
    team.flags = a.pair_flags + 2 * pair;
    team.epoch = slot + 1;
    team.wait = a.pair_wait;
    *mail = a.pair_mail + pair;
    if (c.tid() == 0) *team.epoch = 0;
    c.sync();
    return team;
}

// a.pairs (device-memory path): two workgroups per chain, splitting the work
// where team_wrap_fits(n), the leader alone otherwise.
template <bool kLds>
DEV void body_prepare(const Ctx& c, const LaunchArgs& a, int* slot) {
    int* mail = nullptr;
    const Team team = !kLds ? make_team(c, a, slot, &mail) : Team{};
    for (int q = next_item_team(c, a, slot, team, mail); q < a.count; q = next_item_team(c, a, slot, team, mail)) {
        const ChainDev& ch = a.chains[a.ids[q]];
        const ModelDev& m = a.models[ch.model];
        const Lds L = carve(c.smem, m.n, c.nwaves(), kLds, m.uw_glb ? 0 : m.dk, kLds ? 0 : m.dk);
        if (kLds) prepare_lds(c, m, ch, L);
        else if (team.parts > 1 && team_wrap_fits(m.n)) prepare_global_team(c, m, ch, L, team);
        else if (team.lead()) prepare_global(c, m, ch, L);
    }
}

// kTau: Local_sweep with toggle = .true. (local_sweep.f90:119-123): the
// time-displaced pass between sweep_L and sweep_R (a kernel of its own, so the
// plain sweep's code is not affected by it).
// a.pairs (device-memory path, plain sweep only): two workgroups per chain.
template <bool kLds, bool kTau = false>
DEV void body_sweep(const Ctx& c, const LaunchArgs& a, int* slot) {
    int* mail = nullptr;
    const Team team = (!kLds && !kTau) ? make_team(c, a, slot, &mail) : Team{};
    for (int q = next_item_team(c, a, slot, team, mail); q < a.count; q = next_item_team(c, a, slot, team, mail)) {
        const ChainDev& ch = a.chains[a.ids[q]];
        const ModelDev& m = a.models[ch.model];
        const Lds L = carve(c.smem, m.n, c.nwaves(), kLds, m.uw_glb ? 0 : m.dk, kLds ? 0 : m.dk);
        int seed = *ch.seed;
        if (kLds) lds_from_global(c, L.mat, ch.g, m.n);
        for (int s = 0; s < a.nsweep; ++s) {
            if (kLds) sweep_left_lds(c, m, ch, L, seed, a.measure != 0);
            else sweep_left_global(c, m, ch, L, seed, a.measure != 0, nullptr, false, team);
            if (kTau && ch.tau) {
                if (kLds) global_from_lds(c, ch.g, L.mat, m.n);    // G(0) for the pass; LDS keeps it
                tau_pass(c, m, ch, L, a.tau, a.tau_record != 0 && s == a.nsweep - 1);
            }
            if (kLds) sweep_right_lds(c, m, ch, L, seed, a.measure != 0);
            else sweep_right_global(c, m, ch, L, seed, a.measure != 0, team);
        }
        if (kLds) global_from_lds(c, ch.g, L.mat, m.n);
        if (team.lead() && c.tid() == 0) *ch.seed = seed;
        c.sync();
    }
}

// d log W / d phi of every chain's current fields: one left sweep without
// updates (the Green's function of every slice at the point of its local
// updates), into a.grad at the chain's field offset.
template <bool kLds>
DEV void body_grad(const Ctx& c, const LaunchArgs& a, int* slot) {
    for (int q = next_item(c, a.queue, slot); q < a.count; q = next_item(c, a.queue, slot)) {
        const ChainDev& ch = a.chains[a.ids[q]];
        const ModelDev& m = a.models[ch.model];
        const Lds L = carve(c.smem, m.n, c.nwaves(), kLds, m.uw_glb ? 0 : m.dk, kLds ? 0 : m.dk);
        int seed = *ch.seed;
        double* grad = a.grad + (ch.phi - a.phi0);
        if (kLds) {
            lds_from_global(c, L.mat, ch.g, m.n);
            sweep_left_lds(c, m, ch, L, seed, false, grad);
            global_from_lds(c, ch.g, L.mat, m.n);
        } else {
            sweep_left_global(c, m, ch, L, seed, false, grad);
        }
        c.sync();
    }
}

// Measurement at the stabilisation points only: the left pass with frozen
// fields keeps only what sweep_L needs at its checkpoints -- the product
// UL <- UL B(nt + Nwrap, nt] and Wrap_L's stab_UL and stab_green
// (stabilization.f90:292-328) -- and measures the Green's function stab_green
// returns there, G(nt) for nt = Ltrot (the current G), Ltrot - Nwrap, ...,
// Nwrap, the slices at which sweep_L measures a freshly stabilised G
// (local_sweep.f90:101-103).  The G wraps between checkpoints are not done.
// Ends, like sweep_L, with G(0), the left checkpoint lists and log|det G|.
template <bool kLds>
DEV void measure_checkpoints(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L) {
    const int n = m.n;
    cd* G = kLds ? L.mat : ch.g;
    if (kLds) measure_lds(c, m, ch, L, G);
    else measure(c, m, ch, L, G);
    live_identity(c, n, ch.ul, ch.dl, ch.vl);
    for (int st = m.nst - 1; st >= 0; --st) {
        const int nt = st * m.nwrap;
        double ld;
        if (kLds) {
            lds_from_global(c, L.mat, ch.ul, n);
            apply_slices_left(c, m, ch, L, L.mat, nt + 1, nt + m.nwrap);
            global_from_lds(c, ch.ul, L.mat, n);
            load_ud(c, L.mat, ch.ul, ch.dl, n, true);
            stab_ul_from(c, m, ch, L, L.mat);
            ld = stab_green(c, m, ch, L, nt, list_mat(ch.urlist, st, n), list_vec(ch.drlist, st, n),
                            list_mat(ch.vrlist, st, n), ch.ul, ch.dl, ch.vl, L.mat, ch.wa);
            lds_from_global(c, L.mat, ch.wa, n);
        } else {
            for (int t = nt + m.nwrap; t > nt; --t) {
                const double* phi_t = ch.phi + (size_t)(t - 1) * m.ns * 2;
                gemm_rc_panel(c, ch.w1, m.ez[2], ch.ul, 1, n, L.panel, L.panel_cap);   // UL <- UL e^{-dtau K}
                mat_copy(c, ch.ul, ch.w1, n, false);
                for (int nf = 2; nf >= 1; --nf) {
                    if (!m.active[nf - 1]) continue;
                    hs_factors(c, m, L, phi_t, nf);
                    scale_mat(c, ch.ul, n, nullptr, L.f);         // UL <- UL e^{V}
                }
            }
            load_ud(c, ch.wa, ch.ul, ch.dl, n, true);
            stab_ul_from(c, m, ch, L, ch.wa);
            ld = stab_green(c, m, ch, L, nt, list_mat(ch.urlist, st, n), list_vec(ch.drlist, st, n),
                            list_mat(ch.vrlist, st, n), ch.ul, ch.dl, ch.vl, ch.wa, ch.g);
        }
        store_left(c, m, ch, st);
        if (c.tid() == 0) ch.stats[S_LOGDET] = ld;
        c.sync();
        if (st > 0) {
            if (kLds) measure_lds(c, m, ch, L, G);
            else measure(c, m, ch, L, G);
        }
    }
}

// measure_checkpoints on a team (device-memory path, team_wrap_fits(n)): the
// estimators in the leader, everything else split as in prepare_global_team.
DEV void measure_checkpoints_team(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L,
                                  const Team& team) {
    const int n = m.n;
    const Part sh = team.share();
    if (team.lead()) {
        measure(c, m, ch, L, ch.g);
        live_identity(c, n, ch.ul, ch.dl, ch.vl);
    }
    team_sync(c, team);
    for (int st = m.nst - 1; st >= 0; --st) {
        const int nt = st * m.nwrap;
        for (int t = nt + m.nwrap; t > nt; --t) {
            const double* phi_t = ch.phi + (size_t)(t - 1) * m.ns * 2;
            gemm_rc_panel(c, ch.w1, m.ez[2], ch.ul, 1, n, L.panel, L.panel_cap, sh);   // UL <- UL e^{-dtau K}
            team_sync(c, team);
            mat_copy(c, ch.ul, ch.w1, n, false, sh);
            for (int nf = 2; nf >= 1; --nf) {
                if (!m.active[nf - 1]) continue;
                hs_factors(c, m, L, phi_t, nf);
                scale_mat(c, ch.ul, n, nullptr, L.f, sh);         // UL <- UL e^{V}
            }
            team_sync(c, team);
        }
        load_ud(c, ch.wa, ch.ul, ch.dl, n, true, sh);
        team_sync(c, team);
        stab_ul_team(c, m, ch, L, ch.wa, team);
        const double ld = stab_green_team(c, m, ch, L, nt, list_mat(ch.urlist, st, n), list_vec(ch.drlist, st, n),
                                          list_mat(ch.vrlist, st, n), ch.ul, ch.dl, ch.vl, ch.wa, ch.g, team);
        store_left(c, m, ch, st, sh);
        if (team.lead() && c.tid() == 0) ch.stats[S_LOGDET] = ld;
        team_sync(c, team);
        if (st > 0 && team.lead()) measure(c, m, ch, L, ch.g);
    }
    team_sync(c, team);
}

// Measurement pass: one left sweep with the fields frozen (no proposals, the
// RNG stream untouched) measuring the equal-time estimators at every slice,
// as sweep_L of Local_sweep measures them (local_sweep.f90:96-115) -- the
// measurement of a configuration reached by a global move; a.measure == 2
// measures at the stabilisation points only (measure_checkpoints), and
// a.measure == 3 measures the current Green's function once (no pass: the
// state is left as it is).  Otherwise G, the UDV factors and the checkpoint
// lists end in the state sweep_L leaves them in.
// a.pairs (device-memory path): two workgroups per chain as in body_prepare.
template <bool kLds>
DEV void body_measure(const Ctx& c, const LaunchArgs& a, int* slot) {
    int* mail = nullptr;
    const Team team = !kLds ? make_team(c, a, slot, &mail) : Team{};
    for (int q = next_item_team(c, a, slot, team, mail); q < a.count; q = next_item_team(c, a, slot, team, mail)) {
        const ChainDev& ch = a.chains[a.ids[q]];
        const ModelDev& m = a.models[ch.model];
        const Lds L = carve(c.smem, m.n, c.nwaves(), kLds, m.uw_glb ? 0 : m.dk, kLds ? 0 : m.dk);
        int seed = *ch.seed;
        if (a.measure == 3) {
            // the Green's function as it is (G(0) = G(beta) after a rebuild or a sweep)
            if (kLds) {
                lds_from_global(c, L.mat, ch.g, m.n);
                measure_lds(c, m, ch, L, L.mat);
            } else if (team.lead()) {
                measure(c, m, ch, L, ch.g);
            }
            c.sync();
            continue;
        }
        const bool split = team.parts > 1 && team_wrap_fits(m.n);
        for (int s = 0; s < a.nsweep; ++s) {
            if (kLds) lds_from_global(c, L.mat, ch.g, m.n);
            if (a.measure == 2) {
                if (split) measure_checkpoints_team(c, m, ch, L, team);
                else if (team.lead()) measure_checkpoints<kLds>(c, m, ch, L);
            } else if (kLds) {
                sweep_left_lds(c, m, ch, L, seed, true, nullptr, true);
            } else {
                sweep_left_global(c, m, ch, L, seed, true, nullptr, true, team);
            }
            if (kLds) global_from_lds(c, ch.g, L.mat, m.n);
        }
        c.sync();
    }
}

// Local_sweep_therm (local_sweep.f90:64-78, localU.f90:115-156): Gaussian
// weight only, one thread per chain.
DEV void body_warmup(const Ctx& c, const LaunchArgs& a) {
    const int q = c.block() * c.nthreads() + c.tid();
    if (q >= a.count) return;
    const ChainDev& ch = a.chains[a.ids[q]];
    const ModelDev& m = a.models[ch.model];
    int seed = *ch.seed;
    for (int s = 0; s < a.nsweep; ++s)
        for (int t = 0; t < m.ltrot; ++t)
            for (int nf = 1; nf <= 2; ++nf) {
                if (!m.active[nf - 1]) continue;
                double* phi_t = ch.phi + (size_t)t * m.ns * 2;
                for (int i = 0; i < m.ns; ++i) {
                    const double po = phi_t[i * 2 + nf - 1];
                    const double xflip = ranf(seed);
                    const double pn = po + (xflip - 0.5) * fabs(m.shift_loc);
                    const double rg = exp(-0.5 * pn * pn) / exp(-0.5 * po * po);
                    const double u = ranf(seed);
                    ch.stats[nf == 1 ? S_THERM_PROP_U1 : S_THERM_PROP_U2] += 1.0;
                    if (fabs(rg) > u) {
                        phi_t[i * 2 + nf - 1] = pn;
                        ch.stats[nf == 1 ? S_THERM_ACC_U1 : S_THERM_ACC_U2] += 1.0;
                    }
                }
            }
    *ch.seed = seed;
}

}  // namespace bafq
