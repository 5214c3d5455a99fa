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

// ------------------------------------------------------------ reductions
DEV double wave_sum(const Ctx& c, double v) {
    for (int m = 32; m >= 1; m >>= 1) v += c.shfl_xor(v, m);
    return v;
}
// Sum over an aligned group of `gs` lanes (gs a power of two <= 64).
DEV double group_sum(const Ctx& c, double v, int gs) {
    for (int m = gs >> 1; m >= 1; m >>= 1) v += c.shfl_xor(v, m);
    return v;
}

// ------------------------------------------------------------ elementwise
DEV void mat_copy(const Ctx& c, cd* dst, const cd* src, int n, bool conj_t) {
    const int nn = n * n;
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
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

DEVFN void gemm_cc(const Ctx& c, cd* C, const cd* A, bool opA, const cd* B, bool opB, int n) {
    const int nt = (n + 15) >> 4;
    const int lane = c.lane();
    const int li = lane & 15, lk = lane >> 4;
    for (int t = c.wave(); t < nt * nt; t += c.nwaves()) {
        const int ti = t / nt, tj = t - ti * nt;
        const int ar = 16 * ti + li, bc = 16 * tj + li;
        d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + lk;
            cd a = cd{0.0, 0.0}, b = cd{0.0, 0.0};
            if (ar < n && k < n) a = load_op(A, opA, ar, k, n);
            if (bc < n && k < n) b = load_op(B, opB, k, bc, n);
            accr = c.mfma(a.r, b.r, accr);
            accr = c.mfma(-a.i, b.i, accr);
            acci = c.mfma(a.r, b.i, acci);
            acci = c.mfma(a.i, b.r, acci);
        }
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            if (row < n && col < n) C[row * n + col] = cd{accr[r], acci[r]};
        }
    }
    c.sync();
}

// Real-by-complex: side 0: C = E * X ; side 1: C = X * E.  E real row-major.
DEVFN void gemm_rc(const Ctx& c, cd* C, const double* E, const cd* X, int side, int n) {
    const int nt = (n + 15) >> 4;
    const int lane = c.lane();
    const int li = lane & 15, lk = lane >> 4;
    for (int t = c.wave(); t < nt * nt; t += c.nwaves()) {
        const int ti = t / nt, tj = t - ti * nt;
        const int ar = 16 * ti + li, bc = 16 * tj + li;
        d4 accr = {0.0, 0.0, 0.0, 0.0}, acci = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < n; k0 += 4) {
            const int k = k0 + lk;
            if (side == 0) {
                double a = 0.0;
                cd b = cd{0.0, 0.0};
                if (ar < n && k < n) a = E[ar * n + k];
                if (bc < n && k < n) b = X[k * n + bc];
                accr = c.mfma(a, b.r, accr);
                acci = c.mfma(a, b.i, acci);
            } else {
                cd a = cd{0.0, 0.0};
                double b = 0.0;
                if (ar < n && k < n) a = X[ar * n + k];
                if (bc < n && k < n) b = E[k * n + bc];
                accr = c.mfma(a.r, b, accr);
                acci = c.mfma(a.i, b, acci);
            }
        }
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * ti + mfma_row(lane, r), col = 16 * tj + li;
            if (row < n && col < n) C[row * n + col] = cd{accr[r], acci[r]};
        }
    }
    c.sync();
}

// ------------------------------------------------------ pivoted QR (ZGEQP3)
// Scratch in LDS: tau[n] (cd), perm[n] (int), nrm[n] (double), bcast (>= 8 doubles).
struct QrScratch {
    cd* tau;
    int* perm;
    double* nrm;
    double* bcast;
};

constexpr int kGroup = 8;   // lanes cooperating on one column

// In-place Householder QR with column pivoting, A * P = Q * R.
// On exit: strictly lower part of A holds the reflectors (v_k = 1 implied),
// the upper triangle holds V = diag(D)^{-1} R (unit-modulus diagonal),
// D[k] = |R_kk|, perm[k] = original index of column k, tau[k] the scalar
// factors with H(k) = I - tau_k v_k v_k^H, Q = H(0) H(1) ... H(n-1).
// Column norms are recomputed exactly after every step (no downdating).
DEVFN void qrp(const Ctx& c, cd* A, int n, const QrScratch& s, double* D) {
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    // initial squared column norms
    for (int jb = 0; jb < n; jb += ngrp) {
        const int j = jb + grp;
        double a = 0.0;
        if (j < n)
            for (int r = gl; r < n; r += kGroup) a += abs2(A[r * n + j]);
        a = group_sum(c, a, kGroup);
        if (j < n && gl == 0) {
            s.nrm[j] = a;
            s.perm[j] = j;
        }
    }
    c.sync();
    for (int k = 0; k < n; ++k) {
        // ---- pivot: first index of the largest remaining column norm
        if (c.wave() == 0) {
            double best = -1.0;
            int bi = n;
            for (int jb = k; jb < n; jb += 64) {
                const int j = jb + c.lane();
                if (j < n) {
                    const double v = s.nrm[j];
                    if (v > best) { best = v; bi = j; }
                }
            }
            for (int m = 32; m >= 1; m >>= 1) {
                const double ob = c.shfl_xor(best, m);
                const int oi = c.shfl_xor_i(bi, m);
                if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
            }
            if (c.lane() == 0) s.bcast[0] = (double)bi;
        }
        c.sync();
        const int p = (int)s.bcast[0];
        if (p != k) {
            for (int r = tid; r < n; r += nth) {
                const cd t = A[r * n + k];
                A[r * n + k] = A[r * n + p];
                A[r * n + p] = t;
            }
            if (tid == 0) {
                const int ti = s.perm[k]; s.perm[k] = s.perm[p]; s.perm[p] = ti;
                const double tn = s.nrm[k]; s.nrm[k] = s.nrm[p]; s.nrm[p] = tn;
            }
        }
        c.sync();
        // ---- Householder reflector for column k (ZLARFG conventions)
        if (c.wave() == 0) {
            double x2 = 0.0;
            for (int rb = k + 1; rb < n; rb += 64) {
                const int r = rb + c.lane();
                if (r < n) x2 += abs2(A[r * n + k]);
            }
            x2 = wave_sum(c, x2);
            if (c.lane() == 0) {
                const cd alpha = A[k * n + k];
                double beta, tr, ti;
                cd scale;
                if (x2 == 0.0 && alpha.i == 0.0) {
                    tr = 0.0; ti = 0.0; beta = alpha.r; scale = cd{0.0, 0.0};
                } else {
                    const double nr = sqrt(alpha.r * alpha.r + alpha.i * alpha.i + x2);
                    beta = alpha.r >= 0.0 ? -nr : nr;
                    tr = (beta - alpha.r) / beta;
                    ti = -alpha.i / beta;
                    scale = cinv(alpha - cd{beta, 0.0});
                }
                s.tau[k] = cd{tr, ti};
                s.bcast[1] = scale.r;
                s.bcast[2] = scale.i;
                s.bcast[3] = beta;
            }
        }
        c.sync();
        {
            const cd scale = cd{s.bcast[1], s.bcast[2]};
            for (int r = k + 1 + tid; r < n; r += nth) A[r * n + k] = A[r * n + k] * scale;
            if (tid == 0) A[k * n + k] = cd{s.bcast[3], 0.0};
        }
        c.sync();
        // ---- apply H(k)^H = I - conj(tau) v v^H to columns j > k, new norms
        const cd ctau = conj(s.tau[k]);
        for (int jb = k + 1; jb < n; jb += ngrp) {
            const int j = jb + grp;
            const bool ok = j < n;
            double wr = 0.0, wi = 0.0;
            if (ok) {
                for (int r = k + gl; r < n; r += kGroup) {
                    const cd v = (r == k) ? cd{1.0, 0.0} : A[r * n + k];
                    const cd x = A[r * n + j];
                    const cd p2 = conj(v) * x;
                    wr += p2.r; wi += p2.i;
                }
            }
            wr = group_sum(c, wr, kGroup);
            wi = group_sum(c, wi, kGroup);
            const cd t = ctau * cd{wr, wi};
            double a = 0.0;
            if (ok) {
                for (int r = k + gl; r < n; r += kGroup) {
                    const cd v = (r == k) ? cd{1.0, 0.0} : A[r * n + k];
                    const cd x = A[r * n + j] - v * t;
                    A[r * n + j] = x;
                    if (r > k) a += abs2(x);
                }
            }
            a = group_sum(c, a, kGroup);
            if (ok && gl == 0) s.nrm[j] = a;
        }
        c.sync();
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

// Q = H(0) ... H(n-1) formed explicitly (ZUNGQR).
DEVFN void form_q(const Ctx& c, cd* Q, const cd* A, const cd* tau, int n) {
    mat_identity(c, Q, n);
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    for (int k = n - 1; k >= 0; --k) {
        const cd tk = tau[k];
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

// X <- Q^H X (ZUNMQR 'L','C').
DEVFN void apply_qh_left(const Ctx& c, cd* X, const cd* A, const cd* tau, int n) {
    const int tid = c.tid(), nth = c.nthreads();
    const int grp = tid / kGroup, gl = tid % kGroup, ngrp = nth / kGroup;
    for (int k = 0; k < n; ++k) {
        const cd tk = conj(tau[k]);
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

// X <- V^{-1} X, V = upper triangle of A (ZTRSM 'L','U','N','N').
DEVFN void trsm_upper_left(const Ctx& c, cd* X, const cd* A, int n) {
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

// Y = V * X[perm, :]   (ZLAPMR forward + ZTRMM 'L','U','N').
DEVFN void trmm_perm_rows(const Ctx& c, cd* Y, const cd* A, const cd* X, const int* perm, int n) {
    const int nn = n * n;
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, j = e - r * n;
        double sr = 0.0, si = 0.0;
        for (int k = r; k < n; ++k) {
            const cd p2 = A[r * n + k] * X[perm[k] * n + j];
            sr += p2.r; si += p2.i;
        }
        Y[e] = cd{sr, si};
    }
    c.sync();
}

// Y = X[:, perm] * V^H  (ZLAPMT forward + ZTRMM 'R','U','C').
DEVFN void perm_cols_times_vh(const Ctx& c, cd* Y, const cd* X, const cd* A, const int* perm, int n) {
    const int nn = n * n;
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, j = e - r * n;
        double sr = 0.0, si = 0.0;
        for (int k = j; k < n; ++k) {
            const cd p2 = X[r * n + perm[k]] * conj(A[j * n + k]);
            sr += p2.r; si += p2.i;
        }
        Y[e] = cd{sr, si};
    }
    c.sync();
}

}  // namespace bafq
