// Periodic 3D convolutional network over HS fields (learned global proposals).
//
// The network maps a field configuration phi (channels x Ltrot x Ly x Lx,
// periodic in all three directions) to a field-shaped output:
//   a_0 = x,  a_l = relu(conv_l(a_{l-1})) for l < L,  y = conv_L(a_{L-1}),
// each conv a 3x3x3 periodic stencil with bias.  FP32.  Parameters of layer l:
// W_l[cout][cin][3][3][3] then b_l[cout], concatenated over layers.
//
// Every routine is written per output element so the same code runs as a
// grid-stride HIP kernel and as a plain loop on the CPU emulator.  The reference
// has no network; this is part of the learned-proposal extension.
#pragma once

#include <cstdint>

#include "platform.hpp"

namespace bafq {

struct NetShape {
    int lt, ly, lx;   // lattice
    HOSTDEV int vol() const { return lt * ly * lx; }
};

HOSTDEV int wrap_idx(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// out[b][co][t][y][x] (before the activation)
HOSTDEV float conv_fwd_elem(const float* in, int cin, const float* W, const float* bias, int cout, NetShape s,
                            int e) {
    const int v = s.vol();
    const int b = e / (cout * v), r0 = e - b * cout * v;
    const int co = r0 / v, p = r0 - co * v;
    const int t = p / (s.ly * s.lx), r1 = p - t * s.ly * s.lx;
    const int y = r1 / s.lx, x = r1 - y * s.lx;
    float acc = bias[co];
    const float* wco = W + (size_t)co * cin * 27;
    for (int ci = 0; ci < cin; ++ci) {
        const float* src = in + ((size_t)b * cin + ci) * v;
        const float* w = wco + ci * 27;
        for (int dt = -1; dt <= 1; ++dt) {
            const int tt = wrap_idx(t + dt, s.lt);
            for (int dy = -1; dy <= 1; ++dy) {
                const int yy = wrap_idx(y + dy, s.ly);
                const float* row = src + ((size_t)tt * s.ly + yy) * s.lx;
                const float* wr = w + ((dt + 1) * 3 + (dy + 1)) * 3;
                acc += wr[0] * row[wrap_idx(x - 1, s.lx)] + wr[1] * row[x] + wr[2] * row[wrap_idx(x + 1, s.lx)];
            }
        }
    }
    return acc;
}

// gin[b][ci][p] = sum_{co,d} W[co][ci][d] gout[b][co][p - d], times the ReLU
// mask of the input activation when `mask` (the input was relu(.) > 0).
HOSTDEV float conv_bwd_data_elem(const float* gout, int cout, const float* W, int cin, const float* act, bool mask,
                                 NetShape s, int e) {
    const int v = s.vol();
    const int b = e / (cin * v), r0 = e - b * cin * v;
    const int ci = r0 / v, p = r0 - ci * v;
    if (mask && !(act[e] > 0.0f)) return 0.0f;
    const int t = p / (s.ly * s.lx), r1 = p - t * s.ly * s.lx;
    const int y = r1 / s.lx, x = r1 - y * s.lx;
    float acc = 0.0f;
    for (int co = 0; co < cout; ++co) {
        const float* g = gout + ((size_t)b * cout + co) * v;
        const float* w = W + ((size_t)co * cin + ci) * 27;
        for (int dt = -1; dt <= 1; ++dt) {
            const int tt = wrap_idx(t - dt, s.lt);
            for (int dy = -1; dy <= 1; ++dy) {
                const int yy = wrap_idx(y - dy, s.ly);
                const float* row = g + ((size_t)tt * s.ly + yy) * s.lx;
                const float* wr = w + ((dt + 1) * 3 + (dy + 1)) * 3;
                acc += wr[0] * row[wrap_idx(x + 1, s.lx)] + wr[1] * row[x] + wr[2] * row[wrap_idx(x - 1, s.lx)];
            }
        }
    }
    return acc;
}

// gW[co][ci][d] = sum_{b,p} gout[b][co][p] in[b][ci][p + d] over a slice
// [b0, b1) of the batch (the caller sums the slices).
HOSTDEV float conv_bwd_weight_elem(const float* gout, int cout, const float* in, int cin, int b0, int b1,
                                   NetShape s, int e) {
    const int v = s.vol();
    const int co = e / (cin * 27), r0 = e - co * cin * 27;
    const int ci = r0 / 27, d = r0 - ci * 27;
    const int dt = d / 9 - 1, dy = (d / 3) % 3 - 1, dx = d % 3 - 1;
    float acc = 0.0f;
    for (int b = b0; b < b1; ++b) {
        const float* g = gout + ((size_t)b * cout + co) * v;
        const float* src = in + ((size_t)b * cin + ci) * v;
        for (int t = 0; t < s.lt; ++t) {
            const int tt = wrap_idx(t + dt, s.lt);
            for (int y = 0; y < s.ly; ++y) {
                const int yy = wrap_idx(y + dy, s.ly);
                const float* grow = g + ((size_t)t * s.ly + y) * s.lx;
                const float* srow = src + ((size_t)tt * s.ly + yy) * s.lx;
                for (int x = 0; x < s.lx; ++x) acc += grow[x] * srow[wrap_idx(x + dx, s.lx)];
            }
        }
    }
    return acc;
}

// gb[co] = sum_{b,p} gout[b][co][p]
HOSTDEV float conv_bwd_bias_elem(const float* gout, int cout, int batch, NetShape s, int co) {
    const int v = s.vol();
    float acc = 0.0f;
    for (int b = 0; b < batch; ++b) {
        const float* g = gout + ((size_t)b * cout + co) * v;
        for (int p = 0; p < v; ++p) acc += g[p];
    }
    return acc;
}

// Adam (Kingma & Ba 2015) with bias correction; step >= 1.
HOSTDEV void adam_elem(float* p, const float* g, float* m, float* v, float lr, int step, int i) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m[i] = b1 * m[i] + (1.0f - b1) * g[i];
    v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
    const float mh = m[i] / (1.0f - powf(b1, (float)step));
    const float vh = v[i] / (1.0f - powf(b2, (float)step));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace bafq
