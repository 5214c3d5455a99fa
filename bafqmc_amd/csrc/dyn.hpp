// Time-displaced Green's functions of one chain, executed by one workgroup.
//
// Restates the reference's is_tau pass: Dyn_reset + Dyn_sweep_R
// (src/number_conserving/src/dynamics.f90:38-57; identical in
// src/pairing/src/dynamics.f90), run between sweep_L and sweep_R of a sweep
// (local_sweep.f90:119-123) with the fields frozen.  Starting from G(0) at the
// end of sweep_L (PropGreen reset, process_matrix.f90:78-85)
//   G(0,0) = G(tau,tau) = G(tau,0) = G,  G(0,tau) = G - 1,
// every slice propagates (propgrU_R, propgrT_R, multiply.f90:48-64,88-97)
//   G(tau,0) <- B G(tau,0),  G(0,tau) <- G(0,tau) B^{-1},  G(tau,tau) <- B G(tau,tau) B^{-1},
// and the right factor B(tau,0) = UR DR VR, and every Nwrap slices Wrap_tau
// (stabilization.f90:368-404) recomputes all four from the UDV factors of
// B(tau,0) and of B(beta,tau) = VL DL UL (the left checkpoint lists sweep_L
// left) through the 2N x 2N matrix of stab_green_big (stabilization.f90:192-254)
//   [[(VR VL)^{-1}, s DL], [-DR, (UL UR)^{-1}]],
//   [[G00, G0t], [Gt0, Gtt]] = diag(VR^{-1}, UL^{-1}) big^{-1} diag(VL^{-1}, UR^{-1}),
// recording the drift sum |G_new - G_propagated| of G0t, Gt0, Gtt.
//
// Two conventions (the `mode` of the pass):
//   * bosonic (mode 1): s = -1, so the block matrix is the inverse of
//     [[1, -B(beta,tau)], [-B(tau,0), 1]] and the recomputed functions are the
//     ones the propagation carries, G(0) = (1 - B(beta,0))^{-1};
//   * reference (mode 2): the reference's code as written -- s = +1 (the
//     fermionic (1 + B)^{-1} block form, stabilization.f90:206-208) and the U2
//     operator propagated with the U1 channel's field (nf = 1 for both
//     operators, dynamics.f90:50-52) -- which reproduces the reference's
//     info.txt drift lines.
// The pass works on device-memory matrices on both execution paths (on the
// LDS path G stays in LDS, untouched); its 2N x 2N pivoted QR takes its LDS
// scratch from regions the sweep does not use between its two halves.
// Included by sweep.hpp before the kernel bodies.
#pragma once

namespace bafq {

enum TauMode { TAU_OFF = 0, TAU_BOSONIC = 1, TAU_REFERENCE = 2 };

// Per-chain workspace (ChainDev::tau): G00, Gt0, G0t, Gtt (N x N each), the
// block matrix, its right-hand side and the solution (2N x 2N each), VR^{-1};
// with recording, then (G00, Gt0, G0t, Gtt) at tau = 0 .. Ltrot - 1.
struct TauWs {
    cd *g00, *gt0, *g0t, *gtt, *big, *rhs, *sol, *vri, *rec;
};

HOSTDEV size_t tau_ws_elems(int n, int ltrot, bool record) {
    const size_t nn = (size_t)n * n;
    return nn * (4 + 3 * 4 + 1) + (record ? (size_t)ltrot * 4 * nn : 0);
}

DEV TauWs tau_ws(const ModelDev& m, const ChainDev& ch) {
    const size_t nn = (size_t)m.n * m.n;
    TauWs T;
    T.g00 = ch.tau;
    T.gt0 = T.g00 + nn;
    T.g0t = T.gt0 + nn;
    T.gtt = T.g0t + nn;
    T.big = T.gtt + nn;
    T.rhs = T.big + 4 * nn;
    T.sol = T.rhs + 4 * nn;
    T.vri = T.sol + 4 * nn;
    T.rec = ch.tau_rec;
    return T;
}

// out = A^{-1} X from the pivoted QR of A held in F (qrp: A P = Q diag(D) V):
// A^{-1} = P V^{-1} D^{-1} Q^H.  X is overwritten; out (n x n, not X) is the
// work buffer of the blocked reflector application until the result lands there.
DEV void qrp_solve(const Ctx& c, cd* out, cd* X, const cd* F, const QrScratch& s, const double* D, int n) {
    apply_qh_left(c, X, F, s.tau, n, out);
    const int nn = n * n;
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n;
        X[e] = X[e] * cd{1.0 / D[r], 0.0};
    }
    c.sync();
    trsm_upper_left(c, X, F, n, out);
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, k = e - r * n;
        out[s.perm[r] * n + k] = X[e];
    }
    c.sync();
}

// out <- M^{-1} (N x N) through a pivoted QR; F and X are N x N work matrices.
DEV void inverse_qrp(const Ctx& c, cd* out, const cd* M, cd* F, cd* X, const Lds& L, int n) {
    mat_copy(c, F, M, n, false);
    qrp(c, F, n, L.qr, L.dtmp);
    mat_identity(c, X, n);
    qrp_solve(c, out, X, F, L.qr, L.dtmp, n);
}

// QR scratch for the 2N x 2N block matrix: tau in the rank-update row buffer
// (4N complex), perm in the proposal table (N x 64 bytes), norms and D in the
// measurement site table (12N doubles), the reflector in the column buffer.
DEV QrScratch big_qr_scratch(const Lds& L, int n, double** D) {
    QrScratch s;
    s.tau = L.row;
    s.perm = (int*)L.prop;
    s.nrm = (double*)L.site;
    s.bcast = L.qr.bcast;
    s.v = L.col;
    *D = (double*)L.site + 2 * (size_t)n;
    return s;
}

// C (N x N) = A (N x N, ld N) times the block (rb, cb) of the 2N x 2N matrix Y.
DEV void gemm_block(const Ctx& c, cd* C, const cd* A, bool opA, const cd* Y, int rb, int cb, int n) {
    const int n2 = 2 * n;
    mfma_gemm(
        c, n, [&](int r, int k) { return load_op(A, opA, r, k, n); },
        [&](int k, int j) { return Y[(size_t)(k + rb * n) * n2 + j + cb * n]; },
        [&](int r, int j, cd v) { C[r * n + j] = v; });
}

// Sum over the workgroup of |a - b| over N x N entries (compare_mat,
// stabilization.f90:256-266), on every thread.
DEV double drift_sum(const Ctx& c, const Lds& L, const cd* a, const cd* b, int n) {
    const int nn = n * n;
    double v = 0.0;
    for (int e = c.tid(); e < nn; e += c.nthreads()) v += cabs(a[e] - b[e]);
    v = wave_sum(c, v);
    if (c.lane() == 0) L.red[c.wave()] = v;
    c.sync();
    double s = 0.0;
    for (int w = 0; w < c.nwaves(); ++w) s += L.red[w];
    c.sync();
    return s;
}

// stab_green_big + the drift test and replacement of Wrap_tau, with the right
// factors in ch.ur/dr/vr (after stab_UR) and the left ones from list slot st.
DEVFN void wrap_tau(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, const TauWs& T, int st,
                  double sign) {
    const int n = m.n, n2 = 2 * n;
    const cd* ul = list_mat(ch.ullist, st, n);
    const cd* vl = list_mat(ch.vllist, st, n);
    const double* dl = list_vec(ch.dllist, st, n);
    // blocks of the 2N x 2N matrix: (VR VL)^{-1} (w2) and (UL UR)^{-1} = UR^H UL^H (w1)
    gemm_cc(c, ch.w1, ch.vr, false, vl, false, n);
    inverse_qrp(c, ch.w2, ch.w1, ch.wa, ch.w3, L, n);
    gemm_cc(c, ch.w1, ch.ur, true, ul, true, n);
    for (int e = c.tid(); e < n2 * n2; e += c.nthreads()) {
        const int r = e / n2, k = e - r * n2;
        cd v = cd{0.0, 0.0};
        if (r < n && k < n) v = ch.w2[r * n + k];
        else if (r >= n && k >= n) v = ch.w1[(r - n) * n + (k - n)];
        else if (r < n && k - n == r) v = cd{sign * dl[r], 0.0};
        else if (r >= n && r - n == k) v = cd{-ch.dr[k], 0.0};
        T.big[e] = v;
    }
    c.sync();
    // right-hand side diag(VL^{-1}, UR^{-1} = UR^H)
    inverse_qrp(c, ch.w2, vl, ch.wa, ch.w3, L, n);
    for (int e = c.tid(); e < n2 * n2; e += c.nthreads()) {
        const int r = e / n2, k = e - r * n2;
        cd v = cd{0.0, 0.0};
        if (r < n && k < n) v = ch.w2[r * n + k];
        else if (r >= n && k >= n) v = conj(ch.ur[(k - n) * n + (r - n)]);
        T.rhs[e] = v;
    }
    c.sync();
    double* D2 = nullptr;
    const QrScratch s2 = big_qr_scratch(L, n, &D2);
    qrp(c, T.big, n2, s2, D2);
    qrp_solve(c, T.sol, T.rhs, T.big, s2, D2, n2);
    // rows: diag(VR^{-1}, UL^H) times the solution, new functions into the free rhs
    inverse_qrp(c, T.vri, ch.vr, ch.wa, ch.w3, L, n);
    const size_t nn = (size_t)n * n;
    cd* n00 = T.rhs;
    cd* n0t = T.rhs + nn;
    cd* nt0 = T.rhs + 2 * nn;
    cd* ntt = T.rhs + 3 * nn;
    gemm_block(c, n00, T.vri, false, T.sol, 0, 0, n);
    gemm_block(c, n0t, T.vri, false, T.sol, 0, 1, n);
    gemm_block(c, nt0, ul, true, T.sol, 1, 0, n);
    gemm_block(c, ntt, ul, true, T.sol, 1, 1, n);
    // drifts of G0t, Gt0, Gtt (PropGr%Xmaxm(1..3)); the pairing reference sums
    // over its 4Ns functions, whose second sector is the conjugate of this one
    const double f = m.kind == 1 ? 2.0 : 1.0;
    const double d0 = f * drift_sum(c, L, n0t, T.g0t, n);
    const double d1 = f * drift_sum(c, L, nt0, T.gt0, n);
    const double d2 = f * drift_sum(c, L, ntt, T.gtt, n);
    if (c.tid() == 0) {
        const double d[3] = {d0, d1, d2};
        for (int k = 0; k < 3; ++k) {
            if (d[k] > ch.stats[S_TAU_XMAX + k]) ch.stats[S_TAU_XMAX + k] = d[k];
            ch.stats[S_TAU_XSUM + k] += d[k];
        }
        ch.stats[S_TAU_NWRAP] += 1.0;
    }
    mat_copy(c, T.g00, n00, n, false);
    mat_copy(c, T.g0t, n0t, n, false);
    mat_copy(c, T.gt0, nt0, n, false);
    mat_copy(c, T.gtt, ntt, n, false);
}

// The pass itself: G(0) in ch.g, the left checkpoint lists of sweep_L.
// Clobbers ch.ur/dr/vr and the work matrices (sweep_R starts from the
// identity right factor and reads the left lists), not G, the lists or the fields.
DEVFN void tau_pass(const Ctx& c, const ModelDev& m, const ChainDev& ch, const Lds& L, int mode, bool record) {
    const int n = m.n, nn = n * n;
    const TauWs T = tau_ws(m, ch);
    const double sign = mode == TAU_REFERENCE ? 1.0 : -1.0;
    const int u2_field = mode == TAU_REFERENCE ? 1 : 2;
    live_identity(c, n, ch.ur, ch.dr, ch.vr);      // Prop_d's right factor: URlist(0), the identity
    for (int e = c.tid(); e < nn; e += c.nthreads()) {
        const int r = e / n, k = e - r * n;
        const cd g = ch.g[e];
        T.g00[e] = g;
        T.gt0[e] = g;
        T.gtt[e] = g;
        T.g0t[e] = r == k ? g - cd{1.0, 0.0} : g;
    }
    c.sync();
    for (int nt = 1; nt <= m.ltrot; ++nt) {
        if (record && T.rec) {
            // Obs_tau_calc's view at tau = nt - 1
            cd* R = T.rec + (size_t)(nt - 1) * 4 * nn;
            for (int e = c.tid(); e < nn; e += c.nthreads()) {
                R[e] = T.g00[e];
                R[nn + e] = T.gt0[e];
                R[2 * nn + e] = T.g0t[e];
                R[3 * nn + e] = T.gtt[e];
            }
            c.sync();
        }
        const double* phi_t = ch.phi + (size_t)(nt - 1) * m.ns * 2;
        for (int nf = 1; nf <= 2; ++nf) {
            if (!m.active[nf - 1]) continue;
            hs_factors(c, m, L, phi_t, nf, nf == 2 ? u2_field : 1);
            scale_mat(c, T.gt0, n, L.f, nullptr);       // e^{V} G(tau,0)
            scale_mat(c, T.g0t, n, nullptr, L.finv);    // G(0,tau) e^{-V}
            scale_mat(c, T.gtt, n, L.f, L.finv);        // e^{V} G(tau,tau) e^{-V}
            scale_mat(c, ch.ur, n, L.f, nullptr);       // e^{V} UR
        }
        // propgrT_R: e^{-dtau K} from the left, e^{+dtau K} from the right
        gemm_rc(c, ch.w1, m.ez[0], T.gt0, 0, n);
        mat_copy(c, T.gt0, ch.w1, n, false);
        gemm_rc(c, ch.w1, m.ez[3], T.g0t, 1, n);
        mat_copy(c, T.g0t, ch.w1, n, false);
        gemm_rc(c, ch.w1, m.ez[0], T.gtt, 0, n);
        gemm_rc(c, T.gtt, m.ez[3], ch.w1, 1, n);
        gemm_rc(c, ch.w1, m.ez[0], ch.ur, 0, n);
        mat_copy(c, ch.ur, ch.w1, n, false);
        if (nt % m.nwrap == 0) {
            load_ud(c, ch.wa, ch.ur, ch.dr, n, false);
            stab_ur_from(c, m, ch, L, ch.wa);
            wrap_tau(c, m, ch, L, T, nt / m.nwrap, sign);
        }
    }
}

}  // namespace bafq
