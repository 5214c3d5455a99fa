// Workgroup-cooperative FP64 complex linear algebra for one Markov chain.
//
// Every routine here is executed by ALL threads of a workgroup, operates on
// one N x N complex matrix family of one chain, and ends with a workgroup
// barrier so its outputs are visible to every thread.  Matrices are complex
// row-major with leading dimension N and may live in global memory or LDS
// (generic pointers).  Routines that use wavefront operations (shuffles,
// MFMA) keep their control flow wavefront-uniform.
//
// These are the MI355X restatements of the reference's LAPACK/BLAS calls on
// the sweep path (src/number_conserving/src/stabilization.f90:41-190):
// ZGEQP3 -> qrp, ZUNGQR -> form_q, ZUNMQR -> apply_qh_left, ZTRSM ->
// trsm_upper_left, ZLAPMR+ZTRMM -> trmm_perm_rows, ZLAPMT+ZTRMM('C') ->
// perm_cols_times_vh, ZGEMM -> gemm_cc / gemm_rc (FP64 MFMA 16x16x4).
#pragma once

#include "platform.hpp"

namespace bafq {

struct alignas(16) cd {
    double r, i;
};

HOSTDEV cd mk(double r, double i = 0.0) { return cd{r, i}; }
HOSTDEV cd operator+(cd a, cd b) { return cd{a.r + b.r, a.i + b.i}; }
HOSTDEV cd operator-(cd a, cd b) { return cd{a.r - b.r, a.i - b.i}; }
HOSTDEV cd operator-(cd a) { return cd{-a.r, -a.i}; }
HOSTDEV cd operator*(cd a, cd b) { return cd{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
HOSTDEV cd operator*(double s, cd a) { return cd{s * a.r, s * a.i}; }
HOSTDEV cd conj(cd a) { return cd{a.r, -a.i}; }
HOSTDEV double abs2(cd a) { return a.r * a.r + a.i * a.i; }
HOSTDEV double cabs(cd a) { return hypot(a.r, a.i); }
HOSTDEV cd cdiv(cd a, cd b) {
    const double d = b.r * b.r + b.i * b.i;
    return cd{(a.r * b.r + a.i * b.i) / d, (a.i * b.r - a.r * b.i) / d};
}
HOSTDEV cd cinv(cd b) { return cdiv(cd{1.0, 0.0}, b); }
// Component-wise select.  A conditional expression of struct type can leave the
// operands in private memory (a select between their addresses); this cannot.
HOSTDEV cd csel(bool p, cd a, cd b) { return cd{p ? a.r : b.r, p ? a.i : b.i}; }
// exp of a complex number, Fortran semantics exp(re) * (cos(im), sin(im)).
HOSTDEV cd cexp(cd z) {
    const double e = exp(z.r);
    if (z.i == 0.0) return cd{e, 0.0};
    return cd{e * cos(z.i), e * sin(z.i)};
}

// ------------------------------------------------------------ LDS pointers
// Pointers known to address LDS are cast to the LDS address space so the
// compiler emits ds_read/ds_write instead of flat accesses.  Complex values are
// moved through their components (no struct copy across address spaces).
#ifndef BAFQ_EMU
#define LDS_AS __attribute__((address_space(3)))
#else
#define LDS_AS
#endif
template <class T>
DEV LDS_AS T* as_lds(T* p) { return (LDS_AS T*)(p); }
#ifndef BAFQ_EMU
#define GLB_AS __attribute__((address_space(1)))
#else
#define GLB_AS
#endif
template <class T>
DEV GLB_AS T* as_global(T* p) { return (GLB_AS T*)(p); }
template <class P>
DEV cd ldc(P p, int i) { return cd{p[i].r, p[i].i}; }
template <class P>
DEV void stc(P p, int i, cd v) {
    p[i].r = v.r;
    p[i].i = v.i;
}

// Matrices larger than this (the device-memory path) run the paired-tile
// GEMM variants.
constexpr int kPairMinN = 96;

// Share `i` of `n` of a routine's work when several workgroups split one
// matrix operation (the default: all of it).
struct Part {
    int i = 0, n = 1;
};

// ------------------------------------------------------------ reductions
DEV double wave_sum(const Ctx& c, double v) {
    for (int m = 32; m >= 1; m >>= 1) v += c.shfl_xor(v, m);
    return v;
}
// Sum over an aligned group of `gs` lanes (gs a power of two <= 64); groups
// of 8 and 16 reduce with DPP lane permutations (no LDS round trip), which
// need every lane of the wavefront active.
DEV double group_sum(const Ctx& c, double v, int gs) {
    if (gs == 8) return c.oct_sum(v);
    if (gs == 16) return c.row16_sum(v);
    for (int m = gs >> 1; m >= 1; m >>= 1) v += c.shfl_xor(v, m);
    return v;
}

// ------------------------------------------------------------ elementwise
DEV void mat_copy(const Ctx& c, cd* dst, const cd* src, int n, bool conj_t, Part pt = Part{}) {
    const int nn = n * n;
    for (int e = c.tid() + pt.i * c.nthreads(); e < nn; e += pt.n * c.nthreads()) {
        const int r = e / n, k = e - r * n;
        dst[e] = conj_t ? conj(src[k * n + r]) : src[e];
    }
    c.sync();
}

DEV void mat_identity(const Ctx& c, cd* dst, int n) {
    const int nn = n * n;
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, k = e - r * n;
        dst[e] = cd{r == k ? 1.0 : 0.0, 0.0};
    }
    c.sync();
}

DEV void vec_fill(const Ctx& c, double* dst, int n, double v) {
    for (int e = c.tid(); e < n; e += c.nthreads()) dst[e] = v;
    c.sync();
}

DEV void vec_copy(const Ctx& c, double* dst, const double* src, int n) {
    for (int e = c.tid(); e < n; e += c.nthreads()) dst[e] = src[e];
    c.sync();
}

// M[r][k] *= d[k] (real column scaling).
DEV void scale_mat_cols_real(const Ctx& c, cd* M, int n, const double* d) {
    const int nn = n * n;
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int k = e % n;
        M[e] = cd{M[e].r * d[k], M[e].i * d[k]};
    }
    c.sync();
}

// --------------------------------------------------------------- GEMMs
// C = op(A) * op(B), op in {N, C (conjugate transpose)}; C must not alias.
DEV cd load_op(const cd* M, bool conj_t, int row, int col, int n) {
    return conj_t ? conj(M[col * n + row]) : M[row * n + col];
}

// Complex n x n product on FP64 MFMA 16x16x4: C[r][j] = sum_k a(r, k) b(k, j).
// Each wavefront owns 16x16 output tiles; operand fragments for kChunk k-steps
// are fetched before they are consumed so their latencies overlap.  The
// loaders la(r, k), lb(k, j) apply any transposition, conjugation, masking or
// permutation; st(r, j, v) stores one result.
template <class LA, class LB, class ST>
DEV void mfma_gemm(const Ctx& c, int n, LA la, LB lb, ST st, Part pt = Part{}) {
    constexpr int kChunk = 4;
    const int nt = (n + 15) >> 4, nk = (n + 3) >> 2;
    const int lane = c.lane();
    const int li = lane & 15, lk = lane >> 4;
    for (int t = c.wave() + pt.i * c.nwaves(); t < nt * nt; t += pt.n * c.nwaves()) {
        const int ti = t / nt, tj = t - ti * nt;
        const int ar = 16 * ti + li, bc = 16 * tj + li;
        // three real products per complex one (Gauss): Re = sum ar br - sum ai bi,
        // Im = sum (ar + ai)(br + bi) - sum ar br - sum ai bi
        d4 rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ss = rr;
        for (int k0 = 0; k0 < nk; k0 += kChunk) {
            cd a[kChunk], b[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const int k = 4 * (k0 + u) + lk;
                a[u] = cd{0.0, 0.0};
                b[u] = cd{0.0, 0.0};
                if (k < n && ar < n) a[u] = la(ar, k);
                if (k < n && bc < n) b[u] = lb(k, bc);
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                if (k0 + u < nk) {
                    rr = c.mfma(a[u].r, b[u].r, rr);
                    ii = c.mfma(a[u].i, b[u].i, ii);
                    ss = c.mfma(a[u].r + a[u].i, b[u].r + b[u].i, ss);
                }
            }
        }
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            if (row < n && col < n) st(row, col, cd{rr[r] - ii[r], ss[r] - rr[r] - ii[r]});
        }
    }
    c.sync();
}

// mfma_gemm with two tiles of one tile row per wavefront iteration, sharing
// their A fragments, the loads of both issued together (for operands in device
// memory); the same products per tile.
template <class LA, class LB, class ST>
DEV void mfma_gemm2(const Ctx& c, int n, LA la, LB lb, ST st, Part pt = Part{}) {
    constexpr int kChunk = 4, kT = 2;
    const int nt = (n + 15) >> 4, nk = (n + 3) >> 2, nh = (nt + kT - 1) / kT;
    const int lane = c.lane();
    const int li = lane & 15, lk = lane >> 4;
    for (int t = c.wave() + pt.i * c.nwaves(); t < nt * nh; t += pt.n * c.nwaves()) {
        const int ti = t / nh, tj0 = kT * (t - ti * nh);
        const int ar = 16 * ti + li;
        // three real products per complex one (Gauss): Re = sum ar br - sum ai bi,
        // Im = sum (ar + ai)(br + bi) - sum ar br - sum ai bi
        d4 rr[kT], ii[kT], ss[kT];
#pragma unroll
        for (int u = 0; u < kT; ++u) {
            rr[u] = d4{0.0, 0.0, 0.0, 0.0};
            ii[u] = rr[u];
            ss[u] = rr[u];
        }
        for (int k0 = 0; k0 < nk; k0 += kChunk) {
            cd a[kChunk], b[kT][kChunk];
#pragma unroll
            for (int v = 0; v < kChunk; ++v) {
                const int k = 4 * (k0 + v) + lk;
                a[v] = cd{0.0, 0.0};
                if (k < n && ar < n) a[v] = la(ar, k);
#pragma unroll
                for (int u = 0; u < kT; ++u) {
                    const int bc = 16 * (tj0 + u) + li;
                    b[u][v] = cd{0.0, 0.0};
                    if (k < n && bc < n) b[u][v] = lb(k, bc);
                }
            }
#pragma unroll
            for (int v = 0; v < kChunk; ++v) {
                if (k0 + v < nk) {
#pragma unroll
                    for (int u = 0; u < kT; ++u) {
                        if (tj0 + u >= nt) break;
                        rr[u] = c.mfma(a[v].r, b[u][v].r, rr[u]);
                        ii[u] = c.mfma(a[v].i, b[u][v].i, ii[u]);
                        ss[u] = c.mfma(a[v].r + a[v].i, b[u][v].r + b[u][v].i, ss[u]);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kT; ++u) {
            if (tj0 + u >= nt) break;
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti + mfma_row(lane, r), col = 16 * (tj0 + u) + li;
                if (row < n && col < n) st(row, col, cd{rr[u][r] - ii[u][r], ss[u][r] - rr[u][r] - ii[u][r]});
            }
        }
    }
    c.sync();
}

DEVFN void gemm_cc(const Ctx& c, cd* C, const cd* A, bool opA, const cd* B, bool opB, int n, Part pt = Part{}) {
    auto la = [&](int r, int k) { return load_op(A, opA, r, k, n); };
    auto lb = [&](int k, int j) { return load_op(B, opB, k, j, n); };
    auto st = [&](int r, int j, cd v) { C[r * n + j] = v; };
    if (n > kPairMinN) mfma_gemm2(c, n, la, lb, st, pt);
    else mfma_gemm(c, n, la, lb, st, pt);
}

// C = E X (side 0) or X E (side 1) on FP64 MFMA 16x16x4, E real N x N in MFMA
// operand order (zero padded; side 0: Ez[ti][kk][lane] = E[16 ti + lane % 16][4 kk + lane / 16],
// side 1: Ez[tj][kk][lane] = E[4 kk + lane / 16][16 tj + lane % 16]), X and C
// complex N x N in device memory.  A wavefront computes two output tiles at a
// time that share their E fragments (side 0: one tile row and two tile
// columns; side 1 the transpose); the loads of eight k-steps of both are
// issued together before their products (the operands come from L2/HBM).
DEVFN void gemm_rc(const Ctx& c, cd* C, const double* Ez, const cd* X, int side, int n, Part pt = Part{}) {
    const int nt = (n + 15) >> 4, nk = (n + 3) >> 2, nh = (nt + 1) >> 1;
    const int lane = c.lane();
    const int li = lane & 15, lk = lane >> 4;
    auto Xg = as_global(X);
    auto Cg = as_global(C);
    for (int t = c.wave() + pt.i * c.nwaves(); t < nt * nh; t += pt.n * c.nwaves()) {
        const int sh = t / nh, p = t - sh * nh;        // shared tile index, pair of the other index
        const bool two = 2 * p + 1 < nt;               // wavefront-uniform
        int ti[2], tj[2];
        for (int u = 0; u < 2; ++u) {
            ti[u] = side == 0 ? sh : 2 * p + u;
            tj[u] = side == 0 ? 2 * p + u : sh;
        }
        auto ez = as_global(Ez + (size_t)sh * nk * 64 + lane);
        d4 accr[2], acci[2];
        for (int u = 0; u < 2; ++u) {
            accr[u] = d4{0.0, 0.0, 0.0, 0.0};
            acci[u] = accr[u];
        }
        for (int kk0 = 0; kk0 < nk; kk0 += 8) {
            double e[8];
            cd x[2][8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int kk = kk0 + v, k = 4 * kk + lk;
                e[v] = 0.0;
                if (kk < nk) e[v] = ez[(size_t)kk * 64];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    x[u][v] = cd{0.0, 0.0};
                    const int ar = 16 * ti[u] + li, bc = 16 * tj[u] + li;
                    if (u == 1 && !two) continue;
                    if (side == 0) {
                        if (k < n && bc < n) x[u][v] = ldc(Xg, k * n + bc);
                    } else {
                        if (ar < n && k < n) x[u][v] = ldc(Xg, ar * n + k);
                    }
                }
            }
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                if (kk0 + v < nk) {
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        if (u == 1 && !two) break;
                        if (side == 0) {
                            accr[u] = c.mfma(e[v], x[u][v].r, accr[u]);
                            acci[u] = c.mfma(e[v], x[u][v].i, acci[u]);
                        } else {
                            accr[u] = c.mfma(x[u][v].r, e[v], accr[u]);
                            acci[u] = c.mfma(x[u][v].i, e[v], acci[u]);
                        }
                    }
                }
            }
        }
        for (int u = 0; u < 2; ++u) {
            if (u == 1 && !two) break;
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * ti[u] + mfma_row(lane, r), col = 16 * tj[u] + li;
                if (row < n && col < n) stc(Cg, row * n + col, cd{accr[u][r], acci[u][r]});
            }
        }
    }
    c.sync();
}

// gemm_rc with X staged in LDS, blocks of 16 columns (side 0) or 16 rows
// (side 1) of X in `panel` (cap complex; two blocks at a time where they fit,
// each 4 ceil(n/4) x 16 complex, k-major),
// the wavefronts sharing out the output tiles of both blocks: each element of
// X is read from device memory once per product instead of once per output
// tile (kSt loads per thread issued together), E (shared by every chain of the
// model) comes from L2.  With a Part, the
// blocks b = pt.i (mod pt.n).  The same MFMA operands in the same order as
// gemm_rc, so the same bits.
DEVFN void gemm_rc_panel(const Ctx& c, cd* C, const double* Ez, const cd* X, int side, int n, cd* panel_, int cap,
                         Part pt = Part{}) {
    const int nt = (n + 15) >> 4, nk = (n + 3) >> 2, kp = 4 * nk;
    const int per = cap >= 2 * kp * 16 ? 2 : 1;       // blocks per round
    const int lane = c.lane(), li = lane & 15, lk = lane >> 4;
    auto Xg = as_global(X);
    auto Cg = as_global(C);
    LDS_AS cd* const panel = as_lds(panel_);
    const int nb = (nt - pt.i + pt.n - 1) / pt.n;     // this part's blocks
    for (int b0 = 0; b0 < nb; b0 += per) {
        const int nbb = nb - b0 < per ? nb - b0 : per; // blocks in this round
        // panel[h][k][j] = X[k][16 b + j] (side 0) or X[16 b + j][k] (side 1), zero beyond n
        constexpr int kSt = 10;
        const int tot = nbb * kp * 16, nth = c.nthreads();
        for (int e0 = 0; e0 < tot; e0 += kSt * nth) {
            cd v[kSt];
#pragma unroll
            for (int i = 0; i < kSt; ++i) {
                const int e = e0 + c.tid() + i * nth;
                const int h = e / (kp * 16), r = e - h * kp * 16;
                const int b = pt.i + pt.n * (b0 + h);
                const int k = r >> 4, g = 16 * b + (r & 15);
                v[i] = cd{0.0, 0.0};
                if (e < tot && k < n && g < n) v[i] = side == 0 ? ldc(Xg, k * n + g) : ldc(Xg, g * n + k);
            }
#pragma unroll
            for (int i = 0; i < kSt; ++i) {
                const int e = e0 + c.tid() + i * nth;
                if (e < tot) stc(panel, e, v[i]);
            }
        }
        c.sync();
        for (int u = c.wave(); u < nbb * nt; u += c.nwaves()) {
            const int h = u / nt, t = u - h * nt;
            const int b = pt.i + pt.n * (b0 + h);
            LDS_AS const cd* const pb = panel + (size_t)h * kp * 16;
            // output tile (t, b) (side 0) or (b, t) (side 1); E's tile t in operand order
            auto ez = as_global(Ez + (size_t)t * nk * 64 + lane);
            d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = accr;
            for (int kk0 = 0; kk0 < nk; kk0 += 8) {
                double e[8];
                cd x[8];
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    const int kk = kk0 + v;
                    e[v] = kk < nk ? ez[(size_t)kk * 64] : 0.0;
                    x[v] = kk < nk ? ldc(pb, (4 * kk + lk) * 16 + li) : cd{0.0, 0.0};
                }
#pragma unroll
                for (int v = 0; v < 8; ++v) {
                    if (kk0 + v < nk) {
                        if (side == 0) {
                            accr = c.mfma(e[v], x[v].r, accr);
                            acci = c.mfma(e[v], x[v].i, acci);
                        } else {
                            accr = c.mfma(x[v].r, e[v], accr);
                            acci = c.mfma(x[v].i, e[v], acci);
                        }
                    }
                }
            }
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (side == 0 ? t : b) + mfma_row(lane, r);
                const int col = 16 * (side == 0 ? b : t) + li;
                if (row < n && col < n) stc(Cg, row * n + col, cd{accr[r], acci[r]});
            }
        }
        c.sync();                                // the panel is read before the next round
    }
}

// ------------------------------------------------------ pivoted QR (ZGEQP3)
// Scratch in LDS: tau[n] (cd), perm[n] (int), nrm[n] (double), bcast (>= 8
// doubles), v[2n] (cd: the current reflector, then the device-memory QR's
// diagonal candidates).
struct QrScratch {
    cd* tau;
    int* perm;
    double* nrm;
    double* bcast;
    cd* v;
};

constexpr int kGroup = 8;   // lanes cooperating on one column

// H(k)^H = I - conj(tau) v v^H applied to columns j > k of A (device memory),
// their new squared norms below row k (nrm) and their entries in row k + 1
// (alpha: the next step's diagonal candidates).  A group of NG lanes owns a column;
// its rows k + gl + NG i are held in registers (slot i, R slots, n <= NG R)
// between the dot product and the update, so the column is read and written
// once; v comes from LDS.  With NG = kGroup the arithmetic and its order are
// those of the two-pass loop of qrp_apply_loop.
template <int R, int NG>
DEV void qrp_apply_regs(const Ctx& c, cd* A, int n, int k, LDS_AS const cd* v, LDS_AS double* nrm, cd ctau,
                        LDS_AS cd* alpha) {
    constexpr int kGroup = NG;
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    for (int jb = k + 1; jb < n; jb += ngrp) {
        const int j = jb + grp;
        const bool ok = j < n;
        const int jc = ok ? j : n - 1;           // idle groups read a valid column, store nothing
        cd x[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + gl + kGroup * i;
            x[i] = cd{0.0, 0.0};
            if (k + kGroup * i < n && r < n) x[i] = A[r * n + jc];
        }
        double wr = 0.0, wi = 0.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + gl + kGroup * i;
            if (k + kGroup * i < n && r < n) {
                const cd p2 = conj(ldc(v, r)) * x[i];
                wr += p2.r;
                wi += p2.i;
            }
        }
        wr = group_sum(c, wr, kGroup);
        wi = group_sum(c, wi, kGroup);
        const cd t = ctau * cd{wr, wi};
        double a = 0.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + gl + kGroup * i;
            if (k + kGroup * i < n && r < n) {
                const cd y = x[i] - ldc(v, r) * t;
                if (ok) A[r * n + j] = y;
                if (r > k) a += abs2(y);
                if (ok && r == k + 1) stc(alpha, j, y);      // the next step's diagonal entry
            }
        }
        a = group_sum(c, a, kGroup);
        if (ok && gl == 0) nrm[j] = a;
    }
}

DEV void qrp_apply_loop(const Ctx& c, cd* A, int n, int k, LDS_AS const cd* v, LDS_AS double* nrm, cd ctau,
                        LDS_AS cd* alpha) {
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    for (int jb = k + 1; jb < n; jb += ngrp) {
        const int j = jb + grp;
        const bool ok = j < n;
        double wr = 0.0, wi = 0.0;
        if (ok) {
            for (int r = k + gl; r < n; r += kGroup) {
                const cd p2 = conj(ldc(v, r)) * A[r * n + j];
                wr += p2.r;
                wi += p2.i;
            }
        }
        wr = group_sum(c, wr, kGroup);
        wi = group_sum(c, wi, kGroup);
        const cd t = ctau * cd{wr, wi};
        double a = 0.0;
        if (ok) {
            for (int r = k + gl; r < n; r += kGroup) {
                const cd x = A[r * n + j] - ldc(v, r) * t;
                A[r * n + j] = x;
                if (r > k) a += abs2(x);
                if (r == k + 1) stc(alpha, j, x);
            }
        }
        a = group_sum(c, a, kGroup);
        if (ok && gl == 0) nrm[j] = a;
    }
}

// In-place Householder QR with column pivoting, A * P = Q * R.
// On exit: strictly lower part of A holds the reflectors (v_k = 1 implied),
// the upper triangle holds V = diag(D)^{-1} R (unit-modulus diagonal),
// D[k] = |R_kk|, perm[k] = original index of column k, tau[k] the scalar
// factors with H(k) = I - tau_k v_k v_k^H, Q = H(0) H(1) ... H(n-1).
// Column norms are recomputed exactly after every step (no downdating).
// Two barriers per step: every wavefront finds the pivot itself from the
// norms in LDS, and every thread computes the step's ZLARFG scalars itself
// from the pivot column's norm from row k down (nrm) and its entry in row k
// (alpha), both left in LDS by the previous step's update; the column swap,
// the reflector's scaling and its copy to LDS are one pass over the rows.  The scratch's reflector buffer v holds 2n
// entries: v[0, n) the reflector, v[n, 2n) alpha.
DEVFN void qrp(const Ctx& c, cd* A, int n, const QrScratch& s, double* D) {
    // the scratch arrays are in LDS: typed pointers held in registers, so that
    // stores through A (which may alias any generic address) do not force
    // their reloading
    LDS_AS double* const nrm = as_lds(s.nrm);
    LDS_AS int* const perm = as_lds(s.perm);
    LDS_AS cd* const tau = as_lds(s.tau);
    LDS_AS cd* const vs = as_lds(s.v);
    LDS_AS cd* const alpha = vs + n;
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    // initial squared column norms and diagonal candidates
    for (int jb = 0; jb < n; jb += ngrp) {
        const int j = jb + grp;
        double a = 0.0;
        if (j < n)
            for (int r = gl; r < n; r += kGroup) a += abs2(A[r * n + j]);
        a = group_sum(c, a, kGroup);
        if (j < n && gl == 0) {
            nrm[j] = a;
            perm[j] = j;
            stc(alpha, j, A[j]);                 // row 0 of column j
        }
    }
    c.sync();
    for (int k = 0; k < n; ++k) {
        // ---- pivot: first index of the largest remaining column norm (every wavefront)
        double best = -1.0;
        int bi = n;
        for (int jb = k; jb < n; jb += 64) {
            const int j = jb + c.lane();
            if (j < n) {
                // a NaN norm (a chain that has blown up) wins, so the pivot is always a column
                const double v = nrm[j] == nrm[j] ? nrm[j] : INFINITY;
                if (v > best) { best = v; bi = j; }
            }
        }
        for (int m = 32; m >= 1; m >>= 1) {
            const double ob = c.shfl_xor(best, m);
            const int oi = c.shfl_xor_i(bi, m);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        const int p = bi;
        // ---- Householder reflector for the pivot column (ZLARFG conventions):
        // beta = -sign(Re alpha) sqrt(|alpha|^2 + |x|^2), |x|^2 = the norm below row k
        const cd al = ldc(alpha, p);
        const double a2 = al.r * al.r + al.i * al.i;
        const double nr2 = nrm[p];
        const double x2 = nr2 - a2 > 0.0 ? nr2 - a2 : (nr2 == nr2 ? 0.0 : nr2);
        double beta, tr, ti;
        cd scale;
        if (x2 == 0.0 && al.i == 0.0) {
            tr = 0.0; ti = 0.0; beta = al.r; scale = cd{0.0, 0.0};
        } else {
            const double nr = sqrt(a2 + x2);
            beta = al.r >= 0.0 ? -nr : nr;
            tr = (beta - al.r) / beta;
            ti = -al.i / beta;
            scale = cinv(al - cd{beta, 0.0});
        }
        // ---- columns k and p exchanged; column k becomes R above the diagonal,
        // beta on it and the scaled reflector below it (also into LDS)
        for (int r = tid; r < n; r += nth) {
            const cd xk = A[r * n + k];
            const cd xp = p != k ? A[r * n + p] : xk;
            cd y = xp;
            if (r == k) y = cd{beta, 0.0};
            else if (r > k) y = xp * scale;
            A[r * n + k] = y;
            if (p != k) A[r * n + p] = xk;
            if (r > k) stc(vs, r, y);
        }
        if (tid == 0) {
            stc(vs, k, cd{1.0, 0.0});
            stc(tau, k, cd{tr, ti});
            if (p != k) {
                const int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
            }
        }
        c.sync();
        // ---- apply H(k)^H = I - conj(tau) v v^H to columns j > k, new norms and
        // diagonal candidates (the exchanged column p included)
        if (k + 1 < n) {
            const cd ctau = cd{tr, -ti};
            if (n <= 18 * kGroup) qrp_apply_regs<18, kGroup>(c, A, n, k, vs, nrm, ctau, alpha);
            else if (n <= 18 * 2 * kGroup) qrp_apply_regs<18, 2 * kGroup>(c, A, n, k, vs, nrm, ctau, alpha);
            else qrp_apply_loop(c, A, n, k, vs, nrm, ctau, alpha);
            c.sync();
        }
    }
    // ---- D and row normalisation of R
    for (int k = tid; k < n; k += nth) D[k] = cabs(A[k * n + k]);
    c.sync();
    const int nn = n * n;
    for (int e = tid; e < nn; e += nth) {
        const int r = e / n, k = e - r * n;
        if (k >= r) {
            const double d = D[r];
            A[e] = cd{A[e].r / d, A[e].i / d};
        }
    }
    c.sync();
}

DEV cd refl(const cd* A, int n, int r, int k) {
    return r == k ? cd{1.0, 0.0} : A[r * n + k];
}

// X[:, j0:] <- (I - t v v^H) X[:, j0:], v = reflector k of A (unit at row k,
// A[r][k] below it), X in device memory.  A group of NG lanes owns a column
// and holds its rows k + gl + NG i (slot i, n - k <= NG R) in registers from
// the dot product to the update; the reflector's slots and the column's are
// each loaded in one batch (clamped rows, masked values).  The arithmetic and
// its order are those of the row loops of form_q and apply_qh_left.
template <int R, int NG>
DEV void refl_apply_glb(const Ctx& c, cd* X, const cd* A, int n, int k, cd t, int j0) {
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / NG, gl = tid % NG, ngrp = nth / NG;
    cd v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = k + gl + NG * i, rc = r < n ? r : n - 1;
        const cd a = A[rc * n + k];
        v[i] = r == k ? cd{1.0, 0.0} : (r < n ? a : cd{0.0, 0.0});
    }
    for (int jb = j0; jb < n; jb += ngrp) {
        const int j = jb + grp;
        const bool ok = j < n;
        const int jc = ok ? j : n - 1;
        cd x[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + gl + NG * i, rc = r < n ? r : n - 1;
            const cd a = X[rc * n + jc];
            x[i] = r < n ? a : cd{0.0, 0.0};
        }
        double wr = 0.0, wi = 0.0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            if (k + gl + NG * i < n) {
                const cd p2 = conj(v[i]) * x[i];
                wr += p2.r;
                wi += p2.i;
            }
        }
        wr = group_sum(c, wr, NG);
        wi = group_sum(c, wi, NG);
        const cd s = t * cd{wr, wi};
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = k + gl + NG * i;
            if (ok && r < n) X[r * n + j] = x[i] - v[i] * s;
        }
    }
}

// ------------------------- blocked reflector application and substitution (device memory)
// For n >= kBlkMinN and with a work buffer of 32 n + 512 complex, Q formation,
// With a Part, the blocked routines below touch only the 16-column tiles
// j / 16 = pt.i (mod pt.n) of the matrix they update (their columns are
// independent); each workgroup then needs a work buffer of its own.
DEV bool own_col(Part pt, int j) { return ((j >> 4) % pt.n) == pt.i; }

// Q^H X and the triangular solve run on blocks of kBlk reflectors or rows: a
// block of reflectors H(kb) ... H(kb + nb - 1) = I - V T V^H (ZLARFT, forward,
// columnwise) is applied as X <- X - V (op(T) (V^H X)), two FP64 MFMA
// products that read the rows of X once each, in place of nb passes over the
// whole matrix (ZLARFB).  Work layout: Vt[kBlk][n] (the block's reflectors as
// rows, unit on the diagonal, zero above it), W[kBlk][n], Gram[kBlk][kBlk],
// T[kBlk][kBlk].  Same mathematics as the per-reflector loops; the rounding
// differs.
constexpr int kBlk = 16;
constexpr int kBlkMinN = 64;

HOSTDEV size_t blk_work_elems(int n) { return (size_t)2 * kBlk * n + 2 * kBlk * kBlk; }

// Vt[j][r] = entry r of reflector kb + j (j < nb), zero rows for j >= nb.
DEV void blk_stage_reflectors(const Ctx& c, cd* Vt, const cd* A, int n, int kb, int nb) {
    for (int e = c.tid(); e < kBlk * n; e += c.nthreads()) {
        const int j = e / n, r = e - j * n;
        const int k = kb + j;
        cd v = cd{0.0, 0.0};
        if (j < nb) v = r == k ? cd{1.0, 0.0} : (r > k ? A[r * n + k] : cd{0.0, 0.0});
        Vt[e] = v;
    }
    c.sync();
}

// C[i][c0 + j] = sum_{r >= r0} conj(Vt[i][r]) B(r, c0 + j) for j < m (C with
// leading dimension ldc_), on FP64 MFMA 16x16x4: output tiles of 16 columns,
// one per wavefront in turn.
template <class LB>
DEV void blk_vh_times(const Ctx& c, cd* C, int ldc_, const cd* Vt, int n, int r0, int c0, int m, LB lb,
                      Part pt = Part{}) {
    const int lane = c.lane(), li = lane & 15, lk = lane >> 4;
    const int ntj = (m + 15) >> 4;
    // c0 is a multiple of 16: tile tj holds the columns of tile c0 / 16 + tj
    const int t0 = ((pt.i - (c0 >> 4)) % pt.n + pt.n) % pt.n;
    for (int tj = t0 + pt.n * c.wave(); tj < ntj; tj += pt.n * c.nwaves()) {
        const int col = c0 + 16 * tj + li;
        const bool cin = 16 * tj + li < m;
        d4 rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ri = rr, ir = rr;
        for (int k0 = r0; k0 < n; k0 += 16) {
            cd a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int ka = k0 + 4 * q + lk;        // A operand: row li, inner lk
                const int kr = k0 + 4 * q + lk;        // B operand: inner lk, column li
                a[q] = ka < n ? Vt[li * n + ka] : cd{0.0, 0.0};
                b[q] = (kr < n && cin) ? lb(kr, col) : cd{0.0, 0.0};
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                rr = c.mfma(a[q].r, b[q].r, rr);
                ii = c.mfma(a[q].i, b[q].i, ii);
                ri = c.mfma(a[q].r, b[q].i, ri);
                ir = c.mfma(a[q].i, b[q].r, ir);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = mfma_row(lane, r);
            if (cin) C[row * ldc_ + col] = cd{rr[r] + ii[r], ri[r] - ir[r]};   // conj(v) b
        }
    }
    c.sync();
}

// T of the block (upper triangular, H(kb) ... H(kb+nb-1) = I - V T V^H) from
// the Gram matrix Gm = V^H V: T[i][i] = tau_i, T[0:i][i] = -tau_i T[0:i][0:i] Gm[0:i][i]
// (ZLARFT 'F', 'C').  Lane a < 16 of wavefront 0 builds row a of T in
// registers (row a needs only itself and Gm), then stores it.
DEV void blk_build_t(const Ctx& c, cd* T, const cd* Gm, const cd* tau, int kb, int nb) {
    if (c.wave() == 0) {
        const int a = c.lane();
        cd trow[kBlk];
#pragma unroll
        for (int b = 0; b < kBlk; ++b) trow[b] = cd{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < kBlk; ++i) {
            if (i < nb) {
                cd z = cd{0.0, 0.0};          // sum_{b < i} T[a][b] Gm[b][i] (T[a][b] = 0 for b < a)
#pragma unroll
                for (int b = 0; b < i; ++b) z = z + trow[b] * Gm[b * kBlk + i];
                const cd ti = tau[kb + i];
                if (a < i) trow[i] = -(ti * z);
                else if (a == i) trow[i] = ti;
            }
        }
        if (a < kBlk)
#pragma unroll
            for (int b = 0; b < kBlk; ++b) T[a * kBlk + b] = trow[b];
    }
    c.sync();
}

// W[:, c0 : c0 + m] <- op(T) W, op(T) = T (herm false) or T^H; W with leading dimension n.
DEV void blk_t_times(const Ctx& c, cd* W, const cd* T, int n, int c0, int m, bool herm, Part pt = Part{}) {
    for (int j = c.tid(); j < m; j += c.nthreads()) {
        if (!own_col(pt, c0 + j)) continue;
        cd w[kBlk], o[kBlk];
#pragma unroll
        for (int b = 0; b < kBlk; ++b) w[b] = W[b * n + c0 + j];
#pragma unroll
        for (int i = 0; i < kBlk; ++i) {
            cd acc = cd{0.0, 0.0};
#pragma unroll
            for (int b = 0; b < kBlk; ++b) {
                // T upper: T[i][b] for b >= i; T^H[i][b] = conj(T[b][i]) for b <= i
                if (!herm && b >= i) acc = acc + T[i * kBlk + b] * w[b];
                if (herm && b <= i) acc = acc + conj(T[b * kBlk + i]) * w[b];
            }
            o[i] = acc;
        }
#pragma unroll
        for (int b = 0; b < kBlk; ++b) W[b * n + c0 + j] = o[b];
    }
    c.sync();
}

// X[r][col] -= sum_{k < cols} Ut[k][r] Wr[k][col] for rows r in [r_lo, r_hi) and
// columns col in [c_lo, n) (16-aligned tiles; Ut and Wr with leading dimension
// n, cols <= 16), on FP64 MFMA: the shape of the delayed-update flush.
DEV void blk_rank_update(const Ctx& c, cd* X, const cd* Ut, const cd* Wr, int n, int cols, int r_lo, int r_hi,
                         int c_lo, Part pt = Part{}) {
    const int lane = c.lane(), li = lane & 15, lk = lane >> 4;
    const int ti0 = r_lo >> 4, ti1 = (r_hi + 15) >> 4, tj1 = (n + 15) >> 4;
    // own column tiles tj = tjs, tjs + pt.n, ... (tjs the first >= c_lo / 16)
    const int tjs = (c_lo >> 4) + (((pt.i - (c_lo >> 4)) % pt.n + pt.n) % pt.n);
    const int nti = ti1 - ti0, ntj = tjs < tj1 ? (tj1 - tjs + pt.n - 1) / pt.n : 0;
    for (int t = c.wave(); t < nti * ntj; t += c.nwaves()) {
        const int ti = ti0 + t / ntj, tj = tjs + pt.n * (t % ntj);
        const int ar = 16 * ti + li, bc = 16 * tj + li;
        cd a[4], b[4], g[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int kk = 4 * q + lk;
            a[q] = (ar < r_hi && kk < cols) ? Ut[kk * n + ar] : cd{0.0, 0.0};
            b[q] = (bc < n && kk < cols) ? Wr[kk * n + bc] : cd{0.0, 0.0};
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            g[r] = (row < r_hi && col < n) ? X[row * n + col] : cd{0.0, 0.0};
        }
        d4 rr = {0.0, 0.0, 0.0, 0.0}, ii = rr, ss = rr;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (4 * q < cols) {
                rr = c.mfma(a[q].r, b[q].r, rr);
                ii = c.mfma(a[q].i, b[q].i, ii);
                ss = c.mfma(a[q].r + a[q].i, b[q].r + b[q].i, ss);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            if (row >= r_lo && row < r_hi && col >= c_lo && col < n)
                X[row * n + col] = cd{g[r].r - (rr[r] - ii[r]), g[r].i - (ss[r] - rr[r] - ii[r])};
        }
    }
    c.sync();
}

// One block of reflectors applied from the left to columns [c0, n) of X, rows
// [kb, n): X <- (I - V op(T) V^H) X with op(T) = T^H (herm, for Q^H) or T.
DEV void blk_apply(const Ctx& c, cd* X, const cd* A, const cd* tau, int n, int kb, int c0, bool herm, cd* work,
                   Part pt = Part{}) {
    const int nb = n - kb < kBlk ? n - kb : kBlk;
    cd* Vt = work;
    cd* W = work + (size_t)kBlk * n;
    cd* Gm = work + (size_t)2 * kBlk * n;
    cd* T = Gm + kBlk * kBlk;
    blk_stage_reflectors(c, Vt, A, n, kb, nb);
    blk_vh_times(c, Gm, kBlk, Vt, n, kb, 0, kBlk, [&](int r, int j) { return Vt[j * n + r]; });
    blk_build_t(c, T, Gm, tau, kb, nb);
    blk_vh_times(c, W, n, Vt, n, kb, c0, n - c0, [&](int r, int j) { return X[r * n + j]; }, pt);
    blk_t_times(c, W, T, n, c0, n - c0, herm, pt);
    blk_rank_update(c, X, Vt, W, n, nb, kb, n, c0, pt);
}

// Q = H(0) ... H(n-1) formed explicitly (ZUNGQR); blocked with a work buffer
// of blk_work_elems(n) complex for n >= kBlkMinN (Q <- Q_b Q from the last
// block to the first; block kb touches columns >= kb only, the others are
// still those of the identity there).
// A Part (blocked form only) forms the part's column tiles of Q.
DEVFN void form_q(const Ctx& c, cd* Q, const cd* A, const cd* tau, int n, cd* work = nullptr, Part pt = Part{}) {
    if (work && n >= kBlkMinN) {
        for (int e = c.tid(); e < n * n; e += c.nthreads()) {
            const int r = e / n, k = e - r * n;
            if (own_col(pt, k)) Q[e] = cd{r == k ? 1.0 : 0.0, 0.0};
        }
        c.sync();
        for (int kb = ((n - 1) / kBlk) * kBlk; kb >= 0; kb -= kBlk)
            blk_apply(c, Q, A, tau, n, kb, kb, false, work, pt);
        return;
    }
    mat_identity(c, Q, n);
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    for (int k = n - 1; k >= 0; --k) {
        const cd tk = tau[k];
        if (n <= 18 * kGroup) {
            refl_apply_glb<18, kGroup>(c, Q, A, n, k, tk, k);
            c.sync();
            continue;
        }
        if (n <= 18 * 2 * kGroup) {
            refl_apply_glb<18, 2 * kGroup>(c, Q, A, n, k, tk, k);
            c.sync();
            continue;
        }
        for (int jb = k; jb < n; jb += ngrp) {
            const int j = jb + grp;
            const bool ok = j < n;
            double wr = 0.0, wi = 0.0;
            if (ok)
                for (int r = k + gl; r < n; r += kGroup) {
                    const cd p2 = conj(refl(A, n, r, k)) * Q[r * n + j];
                    wr += p2.r; wi += p2.i;
                }
            wr = group_sum(c, wr, kGroup);
            wi = group_sum(c, wi, kGroup);
            const cd t = tk * cd{wr, wi};
            if (ok)
                for (int r = k + gl; r < n; r += kGroup)
                    Q[r * n + j] = Q[r * n + j] - refl(A, n, r, k) * t;
        }
        c.sync();
    }
}

// X <- Q^H X (ZUNMQR 'L','C'); blocked as form_q with a work buffer.
// A Part (blocked form only) updates the part's column tiles of X.
DEVFN void apply_qh_left(const Ctx& c, cd* X, const cd* A, const cd* tau, int n, cd* work = nullptr,
                         Part pt = Part{}) {
    if (work && n >= kBlkMinN) {
        for (int kb = 0; kb < n; kb += kBlk) blk_apply(c, X, A, tau, n, kb, 0, true, work, pt);
        return;
    }
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    for (int k = 0; k < n; ++k) {
        const cd tk = conj(tau[k]);
        if (n <= 18 * kGroup) {
            refl_apply_glb<18, kGroup>(c, X, A, n, k, tk, 0);
            c.sync();
            continue;
        }
        if (n <= 18 * 2 * kGroup) {
            refl_apply_glb<18, 2 * kGroup>(c, X, A, n, k, tk, 0);
            c.sync();
            continue;
        }
        for (int jb = 0; jb < n; jb += ngrp) {
            const int j = jb + grp;
            const bool ok = j < n;
            double wr = 0.0, wi = 0.0;
            if (ok)
                for (int r = k + gl; r < n; r += kGroup) {
                    const cd p2 = conj(refl(A, n, r, k)) * X[r * n + j];
                    wr += p2.r; wi += p2.i;
                }
            wr = group_sum(c, wr, kGroup);
            wi = group_sum(c, wi, kGroup);
            const cd t = tk * cd{wr, wi};
            if (ok)
                for (int r = k + gl; r < n; r += kGroup)
                    X[r * n + j] = X[r * n + j] - refl(A, n, r, k) * t;
        }
        c.sync();
    }
}

// X <- V^{-1} X, V = upper triangle of A (ZTRSM 'L','U','N','N').  With a
// work buffer (kBlk n complex) and n >= kBlkMinN: row blocks of kBlk from the
// bottom, each solved in registers (one thread per column of X) and then
// removed from the rows above by one MFMA product.
// A Part (blocked form only) solves for the part's column tiles of X.
DEVFN void trsm_upper_left(const Ctx& c, cd* X, const cd* A, int n, cd* work = nullptr, Part pt = Part{}) {
    if (work && n >= kBlkMinN) {
        for (int r0 = ((n - 1) / kBlk) * kBlk; r0 >= 0; r0 -= kBlk) {
            const int nb = n - r0 < kBlk ? n - r0 : kBlk;
            for (int j = c.tid(); j < n; j += c.nthreads()) {
                if (!own_col(pt, j)) continue;
                cd x[kBlk];
#pragma unroll
                for (int a = 0; a < kBlk; ++a) x[a] = a < nb ? X[(r0 + a) * n + j] : cd{0.0, 0.0};
#pragma unroll
                for (int a = kBlk - 1; a >= 0; --a) {
                    if (a < nb) {
                        x[a] = x[a] * cinv(A[(r0 + a) * n + r0 + a]);
#pragma unroll
                        for (int b = 0; b < a; ++b) x[b] = x[b] - A[(r0 + b) * n + r0 + a] * x[a];
                    }
                }
#pragma unroll
                for (int a = 0; a < kBlk; ++a)
                    if (a < nb) X[(r0 + a) * n + j] = x[a];
            }
            c.sync();
            if (r0 == 0) break;
            // Ut[k][r] = A[r][r0 + k] for the rows above the block
            for (int e = c.tid(); e < kBlk * n; e += c.nthreads()) {
                const int k = e / n, r = e - k * n;
                work[e] = (k < nb && r < r0) ? A[r * n + r0 + k] : cd{0.0, 0.0};
            }
            c.sync();
            blk_rank_update(c, X, work, X + (size_t)r0 * n, n, nb, 0, r0, 0, pt);
        }
        return;
    }
    const int tid = c.tid(), nth = c.nthreads();
    for (int r = n - 1; r >= 0; --r) {
        const cd inv = cinv(A[r * n + r]);
        for (int j = tid; j < n; j += nth) X[r * n + j] = X[r * n + j] * inv;
        c.sync();
        const int m = r * n;
        for (int e = tid; e < m; e += nth) {
            const int i = e / n, j = e - i * n;
            X[e] = X[e] - A[i * n + r] * X[r * n + j];
        }
        c.sync();
    }
}

// Y = V * X[perm, :]   (ZLAPMR forward + ZTRMM 'L','U','N'); V = upper triangle of A.
DEVFN void trmm_perm_rows(const Ctx& c, cd* Y, const cd* A, const cd* X, const int* perm, int n, Part pt = Part{}) {
    auto la = [&](int r, int k) { return csel(k >= r, A[r * n + k], cd{0.0, 0.0}); };
    auto lb = [&](int k, int j) { return X[perm[k] * n + j]; };
    auto st = [&](int r, int j, cd v) { Y[r * n + j] = v; };
    if (n > kPairMinN) mfma_gemm2(c, n, la, lb, st, pt);
    else mfma_gemm(c, n, la, lb, st, pt);
}

// Y = X[:, perm] * V^H  (ZLAPMT forward + ZTRMM 'R','U','C').
DEVFN void perm_cols_times_vh(const Ctx& c, cd* Y, const cd* X, const cd* A, const int* perm, int n,
                              Part pt = Part{}) {
    auto la = [&](int r, int k) { return X[r * n + perm[k]]; };
    auto lb = [&](int k, int j) { return csel(k >= j, conj(A[j * n + k]), cd{0.0, 0.0}); };
    auto st = [&](int r, int j, cd v) { Y[r * n + j] = v; };
    if (n > kPairMinN) mfma_gemm2(c, n, la, lb, st, pt);
    else mfma_gemm(c, n, la, lb, st, pt);
}

// ------------------------------------ column-parallel QR family (LDS, n <= 96)
// One aligned quad of lanes carries one column of length n <= 96 in registers:
// row r lives in lane r % 4 of the quad, register slot r / 4, and inner
// products over a column are reduced with two DPP quad permutations.  Each
// column goes through the whole chain of reflectors (or substitution steps)
// without a workgroup barrier or a memory round trip per step.  R is the
// number of register slots (4 R >= n); the factor matrix A is in LDS.  Loops
// over slots are branch-free: out-of-range rows are clamped for the load and
// masked to zero.
constexpr int kColGroup = 4;
constexpr int kColMaxN = 96;

// Compiler-only fence: LDS values loaded before it are loaded again after it
// instead of being kept live in registers.
DEV void reload_fence() {
#ifndef BAFQ_EMU
    asm volatile("" ::: "memory");
#endif
}

DEV bool col_path(const Ctx& c, int n, bool a_lds) {
    return a_lds && n <= kColMaxN && c.nthreads() >= kColGroup * n;
}

// Entry (r, k) of reflector k (unit at r == k, zero above the diagonal and
// beyond row n - 1).
DEV cd refl_entry(LDS_AS const cd* A, int n, int r, int k) {
    const int rc = r < n ? r : n - 1;
    const cd v = ldc(A, rc * n + k);
    const bool below = r > k && r < n;
    return cd{below ? v.r : (r == k ? 1.0 : 0.0), below ? v.i : 0.0};
}

// x <- (I - t v v^H) x, v = reflector k stored below the diagonal of A.
template <int R>
DEV void refl_apply_col(const Ctx& c, cd (&x)[R], LDS_AS const cd* A, int n, int k, cd t, int gl) {
    // Slots whose four rows all lie above row k or beyond row n - 1 are skipped
    // (k and n are uniform, so these are scalar branches).  The loads go in
    // batches of kB slots issued together (one LDS round trip per batch, not
    // per slot); a batch may read a slot that is then skipped.  The reflector
    // is read twice from LDS rather than held in registers.
    constexpr int kB = R % 6 == 0 ? 6 : (R % 4 == 0 ? 4 : 1);
    auto needed = [&](int i) { return kColGroup * i + kColGroup - 1 >= k && kColGroup * i < n; };
    double wr = 0.0, wi = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < R; i0 += kB) {
        if (kColGroup * (i0 + kB) - 1 >= k && kColGroup * i0 < n) {   // the batch holds a needed slot
            cd v[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) v[j] = refl_entry(A, n, gl + kColGroup * (i0 + j), k);
#pragma unroll
            for (int j = 0; j < kB; ++j) {
                if (needed(i0 + j)) {
                    wr += v[j].r * x[i0 + j].r + v[j].i * x[i0 + j].i;
                    wi += v[j].r * x[i0 + j].i - v[j].i * x[i0 + j].r;
                }
            }
        }
    }
    const cd s = t * cd{c.quad_sum(wr), c.quad_sum(wi)};
    reload_fence();
#pragma unroll
    for (int i0 = 0; i0 < R; i0 += kB) {
        if (kColGroup * (i0 + kB) - 1 >= k && kColGroup * i0 < n) {   // the batch holds a needed slot
            cd v[kB];
#pragma unroll
            for (int j = 0; j < kB; ++j) v[j] = refl_entry(A, n, gl + kColGroup * (i0 + j), k);
#pragma unroll
            for (int j = 0; j < kB; ++j)
                if (needed(i0 + j)) x[i0 + j] = x[i0 + j] - v[j] * s;
        }
    }
}

// ZLARFG scalars of the quad's register column as pivot of step k: alpha = x_k,
// beta = -sign(Re alpha) |x_{k..n-1}|, tau = (beta - alpha) / beta,
// scale = 1 / (alpha - beta); returns |x_{k..n-1}|^2 (the pivoting norm).
template <int R>
DEV double col_candidate(const Ctx& c, const cd (&x)[R], int n, int k, int gl, double& beta, cd& t, cd& scale) {
    double ar = 0.0, ai = 0.0, x2 = 0.0;
#pragma unroll
    for (int i = 0; i < R; ++i) {
        if (kColGroup * i + kColGroup - 1 >= k && kColGroup * i < n) {
            const int r = gl + kColGroup * i;
            ar += r == k ? x[i].r : 0.0;
            ai += r == k ? x[i].i : 0.0;
            x2 += r > k ? abs2(x[i]) : 0.0;
        }
    }
    ar = c.quad_sum(ar);
    ai = c.quad_sum(ai);
    x2 = c.quad_sum(x2);
    if (x2 == 0.0 && ai == 0.0) {
        beta = ar;
        t = cd{0.0, 0.0};
        scale = cd{0.0, 0.0};
    } else {
        const double nr = sqrt(ar * ar + ai * ai + x2);
        beta = ar >= 0.0 ? -nr : nr;
        t = cd{(beta - ar) / beta, -ai / beta};
        scale = cinv(cd{ar - beta, ai});
    }
    return ar * ar + ai * ai + x2;
}

// Pivoting norm of each original column for the next pivot search: -1 once
// the column has been chosen; a NaN norm (a chain that has blown up) becomes
// +inf and wins, so every step has an owner.
DEV void publish_norm(LDS_AS double* nrm, bool active, double a2, int col, int gl, int n) {
    if (gl == 0 && col < n) nrm[col] = active ? (a2 == a2 ? a2 : INFINITY) : -1.0;
}

// Householder QR with column pivoting (ZGEQP3 semantics, exact column norms)
// with the same outputs as qrp().  Quad q owns original column q for the whole
// factorization and keeps, besides the column, the ZLARFG scalars it would use
// as the next pivot.  Step k: every wavefront that holds columns picks the
// pivot from the norms published in step k - 1 (the largest, ties to the lower
// original column, as IDAMAX); the owner writes column k of the result (R above
// the diagonal, beta, the reflector below) -- barrier -- the other quads apply
// H(k)^H to their columns, compute their next norms and scalars, and publish
// the norm -- barrier.  One LDS slot per column replaces a per-wavefront
// reduction before the barrier, so each step runs one wavefront argmax.
template <int R>
DEV void qrp_cols_t(const Ctx& c, LDS_AS cd* A, int n, const QrScratch& qs, double* D_) {
    LDS_AS cd* tau = as_lds(qs.tau);
    LDS_AS int* perm = as_lds(qs.perm);
    LDS_AS double* nrm = as_lds(qs.nrm);        // pivoting norm per column (n <= 96)
    LDS_AS double* D = as_lds(D_);
    const int col = c.tid() / kColGroup, gl = c.tid() % kColGroup;
    const int colc = col < n ? col : n - 1;
    const bool holds = c.wave() * (64 / kColGroup) < n;   // the wavefront holds columns (uniform)
    bool active = col < n;
    cd x[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int r = gl + kColGroup * i;
        const cd v = ldc(A, (r < n ? r : n - 1) * n + colc);
        const bool in = active && r < n;
        x[i] = cd{in ? v.r : 0.0, in ? v.i : 0.0};
    }
    double beta;
    cd t, scale;
    double a2 = col_candidate<R>(c, x, n, 0, gl, beta, t, scale);
    publish_norm(nrm, active, a2, col, gl, n);
    if (c.tid() < n) perm[c.tid()] = c.tid();   // defined even if a step had no owner
    c.sync();                                    // columns in registers: A may be overwritten
#pragma unroll 1
    for (int k = 0; k < n; ++k) {
        // ---- pivot from the published norms: lane l looks at columns l, l + 64
        int bi = -1;                              // no owner in a wavefront without columns
        if (holds) {
            const int l0 = c.lane(), l1 = c.lane() + 64;
            double bv = l0 < n ? nrm[l0] : -1.0;
            bi = l0;
            const double v1 = l1 < n ? nrm[l1] : -1.0;
            if (v1 > bv) {
                bv = v1;
                bi = l1;
            }
            c.wave_argmax(bv, bi);
        }
        if (col == bi) {
            // column k of the result: R above the diagonal (normalised at the
            // end), beta on it, the reflector below
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = gl + kColGroup * i;
                const cd v = csel(r < k, x[i], csel(r == k, cd{beta, 0.0}, x[i] * scale));
                if (r < n) stc(A, r * n + k, v);
            }
            if (gl == 0) {
                stc(tau, k, t);
                perm[k] = col;
                D[k] = fabs(beta);
            }
            active = false;
        }
        c.sync();
        // ---- H(k)^H = I - conj(tau) v v^H on the remaining columns, next scalars
        if (k + 1 < n) {
            if (c.wave_any(active)) {
                const cd tk = ldc(tau, k);
                refl_apply_col<R>(c, x, A, n, k, csel(active, conj(tk), cd{0.0, 0.0}), gl);
                a2 = col_candidate<R>(c, x, n, k + 1, gl, beta, t, scale);
            }
            publish_norm(nrm, active, a2, col, gl, n);
            c.sync();
        }
    }
    c.sync();
    // V = D^{-1} R
    for (int e = c.tid(); e < n * n; e += c.nthreads()) {
        const int r = e / n, j = e - r * n;
        if (j >= r) {
            const double d = D[r];
            const cd a = ldc(A, e);
            stc(A, e, cd{a.r / d, a.i / d});
        }
    }
    c.sync();
}

// Q = H(0) ... H(n-1) (ZUNGQR), column j = H(0) ... H(n-1) e_j; writes Q, or
// Q^H when conj_t.
template <int R>
DEV void form_q_cols_t(const Ctx& c, cd* Q, LDS_AS const cd* A, const cd* tau_, int n, bool conj_t) {
    LDS_AS const cd* tau = as_lds(tau_);
    GLB_AS cd* Qg = as_global(Q);
    const int gl = c.tid() % kColGroup, ngrp = c.nthreads() / kColGroup;
    for (int jb = 0; jb < n; jb += ngrp) {
        const int j = jb + c.tid() / kColGroup;
        cd x[R];
#pragma unroll
        for (int i = 0; i < R; ++i) x[i] = cd{gl + kColGroup * i == j ? 1.0 : 0.0, 0.0};
        // H(k) e_j = e_j for k > j: start at the wavefront's largest column
        const int jmax = jb + (c.wave() * 64 + 63) / kColGroup;
        const int jmin = jb + c.wave() * 64 / kColGroup;
#pragma unroll 1
        for (int k = (jmax < n - 1 ? jmax : n - 1); k >= 0 && jmin < n; --k)
            refl_apply_col<R>(c, x, A, n, k, ldc(tau, k), gl);
        if (j < n) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = gl + kColGroup * i;
                if (r < n) {
                    if (conj_t) stc(Qg, j * n + r, conj(x[i]));
                    else stc(Qg, r * n + j, x[i]);
                }
            }
        }
    }
    c.sync();
}

// Y[perm[r], j] = (V^{-1} D^{-1} Q^H X)[r, j] with X = Xs (conj_t false) or
// Xs^H (conj_t true): ZUNMQR('L','C'), the 1/D row scaling, ZTRSM('L','U','N','N')
// and ZLAPMR backward of stabilization.f90:159-168, fused per column.
template <int R>
DEV void solve_cols_t(const Ctx& c, cd* Y, const cd* Xs, bool conj_t, LDS_AS const cd* A, const cd* tau_,
                      const double* D_, const int* perm_, int n) {
    LDS_AS const cd* tau = as_lds(tau_);
    LDS_AS const double* D = as_lds(D_);
    LDS_AS const int* perm = as_lds(perm_);
    GLB_AS const cd* Xg = as_global(Xs);
    GLB_AS cd* Yg = as_global(Y);
    const int gl = c.tid() % kColGroup, ngrp = c.nthreads() / kColGroup;
    for (int jb = 0; jb < n; jb += ngrp) {
        const int j = jb + c.tid() / kColGroup;
        const int jc = j < n ? j : n - 1;
        cd x[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = gl + kColGroup * i, rc = r < n ? r : n - 1;
            const cd v = csel(conj_t, conj(ldc(Xg, jc * n + rc)), ldc(Xg, rc * n + jc));
            x[i] = csel(r < n, v, cd{0.0, 0.0});
        }
        const bool idle = jb + c.wave() * 64 / kColGroup >= n;   // whole wavefront past column n - 1
#pragma unroll 1
        for (int k = 0; k < n && !idle; ++k) refl_apply_col<R>(c, x, A, n, k, conj(ldc(tau, k)), gl);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const int r = gl + kColGroup * i;
            const double dr = D[r < n ? r : n - 1];
            x[i] = cd{x[i].r / dr, x[i].i / dr};
        }
        // back substitution with V = upper triangle of A, column (axpy) form:
        // row r = 4 i + g is final once rows > r are done; its owner lane divides
        // by V_rr and the quad then removes V[q][r] y_r from the rows q < r.
#pragma unroll
        for (int i = R - 1; i >= 0; --i) {
#pragma unroll 1
            for (int g = kColGroup - 1; g >= 0; --g) {
                const int r = kColGroup * i + g;
                if (r >= n || idle) continue;
                const bool mine = gl == g;
                const cd y0 = cdiv(x[i], ldc(A, r * n + r));
                const cd y = csel(mine, y0, cd{0.0, 0.0});
                x[i] = csel(mine, y0, x[i]);
                const cd yb = cd{c.quad_sum(y.r), c.quad_sum(y.i)};
#pragma unroll
                for (int q = 0; q <= i; ++q) {
                    const int rq = gl + kColGroup * q;
                    const cd a = ldc(A, (rq < n ? rq : n - 1) * n + r);
                    const bool up = rq < r;
                    x[q] = x[q] - cd{up ? a.r : 0.0, up ? a.i : 0.0} * yb;
                }
            }
        }
        if (j < n) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int r = gl + kColGroup * i;
                if (r < n) stc(Yg, perm[r] * n + j, x[i]);
            }
        }
    }
    c.sync();
}

// One instance per register-slot count R (4 R >= n), inlined into the
// stabilisation routines that call them: as out-of-line calls each saved and
// restored the callee-saved registers it clobbers (some 400 scratch stores
// per lane on entry), which cost 10 % of a G rebuild and 6 % of a sweep at
// N = 81 (same results bit for bit).
template <int R>
DEV void qrp_cols_r(const Ctx& c, cd* A, int n, const QrScratch& s, double* D) {
    qrp_cols_t<R>(c, as_lds(A), n, s, D);
}
template <int R>
DEV void form_q_cols_r(const Ctx& c, cd* Q, const cd* A, const cd* tau, int n, bool conj_t) {
    form_q_cols_t<R>(c, Q, as_lds(A), tau, n, conj_t);
}
template <int R>
DEV void solve_cols_r(const Ctx& c, cd* Y, const cd* Xs, bool conj_t, const cd* A, const cd* tau,
                        const double* D, const int* perm, int n) {
    solve_cols_t<R>(c, Y, Xs, conj_t, as_lds(A), tau, D, perm, n);
}

#define BAFQ_COL_DISPATCH(n, F, ...)                      \
    do {                                                  \
        const int rs_ = ((n) + 3) / 4;                    \
        if (rs_ <= 6) F<6>(__VA_ARGS__);                  \
        else if (rs_ <= 12) F<12>(__VA_ARGS__);           \
        else if (rs_ <= 18) F<18>(__VA_ARGS__);           \
        else F<24>(__VA_ARGS__);                          \
    } while (0)

DEV void qrp_cols(const Ctx& c, cd* A, int n, const QrScratch& s, double* D) {
    BAFQ_COL_DISPATCH(n, qrp_cols_r, c, A, n, s, D);
}
DEV void form_q_cols(const Ctx& c, cd* Q, const cd* A, const cd* tau, int n, bool conj_t) {
    BAFQ_COL_DISPATCH(n, form_q_cols_r, c, Q, A, tau, n, conj_t);
}
DEV void solve_cols(const Ctx& c, cd* Y, const cd* Xs, bool conj_t, const cd* A, const cd* tau, const double* D,
                    const int* perm, int n) {
    BAFQ_COL_DISPATCH(n, solve_cols_r, c, Y, Xs, conj_t, A, tau, D, perm, n);
}

}  // namespace bafq
