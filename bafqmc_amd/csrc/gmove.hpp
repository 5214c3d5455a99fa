// Learned global moves on the device: the element-wise parts of a proposal
// phi' = phi + lam f(phi) + sigma xi over all chains of a batch, and the
// per-chain sums its exact Metropolis-Hastings acceptance needs.  Written per
// element (or per chain) so the same code runs as HIP grid-stride kernels and
// as loops on the CPU emulator.  No counterpart in the reference.
//
// Field layout: phi[c][t][site][ch] (site = x + Lx y), `per` = Ltrot Ns 2
// doubles per chain.  Network layout: x[b][ch][t][y][x] (net.hpp).
#pragma once

#include <cmath>
#include <cstdint>

#include "net.hpp"

namespace bafq {

struct GmArgs {
    NetShape s;          // lattice (lt = Ltrot, ly, lx)
    int c0, batch;       // chains c0 .. c0 + batch - 1
    double lam;          // drift scale (Hamiltonian moves: half the rotation angle, the kick scale)
    double sigma[2];     // noise per channel (0: channel not moved)
    double mask[2];      // 1 for channels whose field couples, else 0
    double w[2];         // 1 / (2 sigma^2) for moved channels, else 0
    uint64_t seed;
    double cosh_, sinh_; // Hamiltonian moves: cos and sin of the rotation angle
    double center[2];    // Hamiltonian moves: centre of the Gaussian part per channel
    double fscale[2];    // Hamiltonian moves: force = fscale[ch] x network output
};

HOSTDEV int gm_per(NetShape s) { return s.vol() * 2; }

// network input of chain c0 + b
HOSTDEV void gm_to_net_elem(const double* phi, float* x, const GmArgs& a, int e) {
    const int v = a.s.vol(), ns = a.s.ly * a.s.lx;
    const int b = e / (2 * v), r = e - b * 2 * v, ch = r / v, p = r - ch * v;
    const int t = p / ns, site = p - t * ns;
    x[e] = (float)phi[(size_t)(a.c0 + b) * gm_per(a.s) + ((size_t)t * ns + site) * 2 + ch];
}

// SplitMix64 (Steele, Lea, Flood, OOPSLA 2014): a counter-based 64-bit mixer.
HOSTDEV uint64_t gm_mix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// standard normal number i of stream `seed` (Box-Muller on two 32-bit uniforms)
HOSTDEV double gm_normal(uint64_t seed, uint64_t i) {
    const uint64_t z = gm_mix(seed ^ gm_mix(i));
    const double u1 = ((double)(uint32_t)(z >> 32) + 1.0) * (1.0 / 4294967296.0);   // (0, 1]
    const double u2 = (double)(uint32_t)z * (1.0 / 4294967296.0);                   // [0, 1)
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// drift of element q of chain c from the network output y (nullptr: no network)
HOSTDEV double gm_drift(const float* y, const GmArgs& a, int b, int q) {
    if (!y) return 0.0;
    const int v = a.s.vol();
    const int ch = q & 1, ts = q >> 1;                 // ts = t Ns + site = network position
    return a.lam * a.mask[ch] * (double)y[(size_t)b * 2 * v + (size_t)ch * v + ts];
}

// phi <- phi + f0 + sigma xi; the old field and f0 kept for the density ratio
HOSTDEV void gm_propose_elem(double* phi, double* old, double* f0, const float* y, const GmArgs& a, int e) {
    const int per = gm_per(a.s);
    const int b = e / per, q = e - b * per;
    const size_t k = (size_t)(a.c0 + b) * per + q;
    const int ch = q & 1;
    const double f = gm_drift(y, a, b, q);
    const double xi = a.sigma[ch] > 0.0 ? gm_normal(a.seed, (uint64_t)k) : 0.0;
    const double o = phi[k];
    old[k] = o;
    f0[k] = f;
    phi[k] = o + f + a.sigma[ch] * xi;
}

// per-chain sums of chain c0 + b over elements [q0, q1): log q(phi|phi') -
// log q(phi'|phi), sum phi^2, sum phi'^2, sum phi_2, sum phi'_2
HOSTDEV void gm_sums(const double* phi, const double* old, const double* f0, const float* y, const GmArgs& a,
                            int b, int q0, int q1, int qstep, double out[5]) {
    const size_t base = (size_t)(a.c0 + b) * gm_per(a.s);
    for (int q = q0; q < q1; q += qstep) {
        const int ch = q & 1;
        const double pn = phi[base + q], po = old[base + q];
        const double f1 = gm_drift(y, a, b, q);
        const double fw = pn - po - f0[base + q], bw = po - pn - f1;
        out[0] += a.w[ch] * (fw * fw - bw * bw);
        out[1] += po * po;
        out[2] += pn * pn;
        if (ch == 1) {
            out[3] += po;
            out[4] += pn;
        }
    }
}

// ---- Hamiltonian moves with a learned force (split HMC).
// log W(phi) = -sum (phi - c)^2 / 2 + R(phi) + const per channel centre c; the
// Gaussian part is integrated exactly (a rotation of (phi - c, p) by the angle
// h), the rest by half kicks p += (h/2) f(phi) with the network's f ~ grad R.
// Every step is a shear or a rotation, so the map is volume preserving and,
// with the momentum reversed, an involution for any deterministic f: the
// acceptance min(1, exp(-H(phi', p') + H(phi, p))) with the true log W is
// exact whatever the network has learned.

// old <- phi; p0 = p = N(0, 1) on moved channels, 0 elsewhere
HOSTDEV void gm_hmc_init_elem(const double* phi, double* old, double* p0, double* p, const GmArgs& a, int e) {
    const int per = gm_per(a.s);
    const int b = e / per, q = e - b * per;
    const size_t k = (size_t)(a.c0 + b) * per + q;
    const double v = a.mask[q & 1] > 0.0 ? gm_normal(a.seed, (uint64_t)k) : 0.0;
    old[k] = phi[k];
    p0[k] = v;
    p[k] = v;
}

// p <- p + lam fscale f(phi) on moved channels (y: network output of chains c0..)
HOSTDEV void gm_hmc_kick_elem(double* p, const float* y, const GmArgs& a, int e) {
    const int per = gm_per(a.s);
    const int b = e / per, q = e - b * per;
    p[(size_t)(a.c0 + b) * per + q] += a.fscale[q & 1] * gm_drift(y, a, b, q);
}

// (phi - c, p) <- rotation by the angle h on moved channels
HOSTDEV void gm_hmc_rotate_elem(double* phi, double* p, const GmArgs& a, int e) {
    const int per = gm_per(a.s);
    const int b = e / per, q = e - b * per;
    const int ch = q & 1;
    if (!(a.mask[ch] > 0.0)) return;
    const size_t k = (size_t)(a.c0 + b) * per + q;
    const double x = phi[k] - a.center[ch], v = p[k];
    phi[k] = a.center[ch] + x * a.cosh_ + v * a.sinh_;
    p[k] = v * a.cosh_ - x * a.sinh_;
}

// per-chain sums of chain c0 + b: (|p0|^2 - |p|^2) / 2 (the kinetic part of
// -dH), sum phi^2, sum phi'^2, sum phi_2 and sum phi'_2 -- the layout of gm_sums
HOSTDEV void gm_hmc_sums(const double* phi, const double* old, const double* p0, const double* p, const GmArgs& a,
                         int b, int q0, int q1, int qstep, double out[5]) {
    const size_t base = (size_t)(a.c0 + b) * gm_per(a.s);
    for (int q = q0; q < q1; q += qstep) {
        const double pn = phi[base + q], po = old[base + q];
        const double m0 = p0[base + q], m1 = p[base + q];
        out[0] += 0.5 * (m0 * m0 - m1 * m1);
        out[1] += po * po;
        out[2] += pn * pn;
        if (q & 1) {
            out[3] += po;
            out[4] += pn;
        }
    }
}

// rejected chains get their old fields back
HOSTDEV void gm_restore_elem(double* phi, const double* old, const int* accept, int per, int e) {
    if (!accept[e / per]) phi[e] = old[e];
}

}  // namespace bafq
