// Periodic 3D convolution as an implicit GEMM on the FP32 matrix cores.
//
// out[b][co][p] = act( bias[co] + sum_{ci, d} Wc(co, ci, d) in[b][ci][p (+) d] )
// over the 27 taps d = (dt, dy, dx) of a 3x3x3 periodic stencil on the
// (Ltrot x Ly x Lx) lattice, the layer of the proposal network (net.hpp holds
// the per-element definition it must equal).  As a GEMM: M = Cout, N = the
// lattice positions, K = Cin x 27.
//
// One workgroup (16 wavefronts) computes all output channels of one chain over
// a block of `tt` consecutive time slices (tt * Ly * Lx positions, in tiles of
// 16).  It stages in LDS
//   * the input slab: slices t0-1 .. t0+tt (periodic in t), every input
//     channel (padded with zero channels to a multiple of 4),
//   * the weights of one time-offset group of taps (9 taps, [tap][ci][co],
//     zero-padded to 16-channel output blocks), three groups in turn,
// and runs v_mfma_f32_16x16x4f32 with A = weights (16 output channels x 4
// input channels of one tap) and B = the input at the 16 positions of a tile
// shifted by the tap (4 input channels): each wavefront keeps the
// accumulators of up to kConvTiles position tiles x NB output blocks in
// registers for the whole K loop, so the slab and weights are read from LDS
// only.  The tap shift of a position is computed once per tap and reused for
// every input channel.
//
// The same routine runs the backward pass with respect to the input
// (flip != 0: Wc(co, ci, d) = W[ci][co][26 - d], the transposed, reflected
// stencil) with the ReLU mask of the forward activation (mask != nullptr).
#pragma once

#include "net.hpp"

namespace bafq {

constexpr int kConvWaves = 16;        // 1024 threads: four wavefronts per SIMD hide the LDS and
                                      // matrix-core latencies (3.74 vs 4.64 ms per forward pass of the
                                      // 4 x 32 network over 256 fields of the 9x9 x 80 lattice with 8)
constexpr int kConvTiles = 3;         // position tiles of 16 per wavefront (<= 48 per workgroup)
constexpr int kConvMaxPos = kConvWaves * kConvTiles * 16;
constexpr int kConvStage = 16;        // global loads in flight per thread while staging LDS

struct ConvArgs {
    const float* in;       // [batch][cin][lt][ly][lx]
    const float* W;        // [cout][cin][27] (flip: [cin][cout][27] of the forward layer)
    const float* bias;     // [cout] or nullptr
    const float* mask;     // [batch][cout][vol]: output zeroed where mask <= 0 (nullptr: none)
    float* out;            // [batch][cout][lt][ly][lx]
    int cin, cout, cinp, coutp;
    int relu, flip;
    int tt, tblocks;       // time slices per workgroup, workgroups per chain
    NetShape s;
};

HOSTDEV int conv_pad4(int c) { return (c + 3) & ~3; }
HOSTDEV int conv_pad16(int c) { return (c + 15) & ~15; }
HOSTDEV size_t conv_slab_floats(int tt, int cinp, int plane) { return (size_t)(tt + 2) * cinp * plane; }
HOSTDEV size_t conv_w_floats(int cinp, int coutp) { return (size_t)9 * cinp * coutp; }
HOSTDEV size_t conv_lds_bytes(int tt, int cinp, int coutp, int plane) {
    return sizeof(float) * (conv_slab_floats(tt, cinp, plane) + conv_w_floats(cinp, coutp));
}

// Largest time block whose slab and weights fit in `lds` bytes and whose
// positions fit the workgroup's accumulators; 0 if none.
HOSTDEV int conv_pick_tt(int lt, int plane, int cinp, int coutp, size_t lds) {
    int best = 0;
    for (int tt = 1; tt <= lt; ++tt) {
        if (tt * plane > kConvMaxPos) break;
        if (conv_lds_bytes(tt, cinp, coutp, plane) > lds) break;
        best = tt;
    }
    return best;
}

// Weight of output channel co, input channel ci, tap d of the convolution run.
HOSTDEV float conv_weight(const ConvArgs& a, int co, int ci, int d) {
    if (co >= a.cout || ci >= a.cin) return 0.0f;
    if (a.flip) return a.W[((size_t)ci * a.cout + co) * 27 + (26 - d)];
    return a.W[((size_t)co * a.cin + ci) * 27 + d];
}

template <int NB>
DEV void conv_mfma_body(const Ctx& c, const ConvArgs& a) {
    const int plane = a.s.ly * a.s.lx, lt = a.s.lt;
    const int b = c.block() / a.tblocks, tb = c.block() - b * a.tblocks;
    const int t0 = tb * a.tt;
    const int tt_eff = lt - t0 < a.tt ? lt - t0 : a.tt;
    const int npos = tt_eff * plane;
    const int cinp = a.cinp, coutp = a.coutp;
    float* slab = (float*)c.smem;
    float* wl = slab + conv_slab_floats(a.tt, cinp, plane);

    // input slab: slices t0-1 .. t0+tt (periodic), zero channels beyond cin;
    // each thread issues kConvStage loads before it stores any, so that many
    // global-memory round trips are in flight instead of one
    const int nslab = (a.tt + 2) * cinp * plane;
    const int nth = c.nthreads();
    for (int e0 = c.tid(); e0 < nslab; e0 += kConvStage * nth) {
        float v[kConvStage];
        for (int u = 0; u < kConvStage; ++u) {
            const int e = e0 + u * nth;
            v[u] = 0.0f;
            if (e < nslab) {
                const int ts = e / (cinp * plane), r = e - ts * cinp * plane;
                const int ci = r / plane, yx = r - ci * plane;
                int tg = t0 - 1 + ts;
                tg = tg < 0 ? tg + lt : tg;
                tg = tg >= lt ? tg % lt : tg;
                if (ci < a.cin) v[u] = a.in[(((size_t)b * a.cin + ci) * lt + tg) * plane + yx];
            }
        }
        for (int u = 0; u < kConvStage; ++u)
            if (e0 + u * nth < nslab) slab[e0 + u * nth] = v[u];
    }

    const int lane = c.lane(), wave = c.wave();
    const int jl = lane & 15, kl = lane >> 4;
    // this lane's position in each of its tiles (clamped into range; results
    // of positions >= npos are not stored)
    int pt[kConvTiles], py[kConvTiles], px[kConvTiles];
    for (int m = 0; m < kConvTiles; ++m) {
        int pos = (wave + m * kConvWaves) * 16 + jl;
        pos = pos < npos ? pos : npos - 1;
        pt[m] = pos / plane;
        const int yx = pos - pt[m] * plane;
        py[m] = yx / a.s.lx;
        px[m] = yx - py[m] * a.s.lx;
    }
    const int ntiles = (npos + 15) >> 4;
    f4 acc[kConvTiles][NB];
    for (int m = 0; m < kConvTiles; ++m)
        for (int nb = 0; nb < NB; ++nb) acc[m][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};

    for (int g = 0; g < 3; ++g) {
        c.sync();
        const int nw = 9 * cinp * coutp;
        for (int e0 = c.tid(); e0 < nw; e0 += kConvStage * nth) {
            float v[kConvStage];
            for (int u = 0; u < kConvStage; ++u) {
                const int e = e0 + u * nth;
                v[u] = 0.0f;
                if (e < nw) {
                    const int dd = e / (cinp * coutp), r = e - dd * cinp * coutp;
                    const int ci = r / coutp, co = r - ci * coutp;
                    v[u] = conv_weight(a, co, ci, g * 9 + dd);
                }
            }
            for (int u = 0; u < kConvStage; ++u)
                if (e0 + u * nth < nw) wl[e0 + u * nth] = v[u];
        }
        c.sync();
        for (int dd = 0; dd < 9; ++dd) {
            const int dy = dd / 3 - 1, dx = dd % 3 - 1;
            int off[kConvTiles];
            for (int m = 0; m < kConvTiles; ++m) {
                int y = py[m] + dy, x = px[m] + dx;
                y = y < 0 ? y + a.s.ly : (y >= a.s.ly ? y - a.s.ly : y);
                x = x < 0 ? x + a.s.lx : (x >= a.s.lx ? x - a.s.lx : x);
                off[m] = ((pt[m] + g) * cinp + kl) * plane + y * a.s.lx + x;    // slab slice pt + (g - 1) + 1
            }
            const float* wd = wl + (size_t)dd * cinp * coutp + kl * coutp + jl;
            for (int cg = 0; cg < cinp; cg += 4) {
                float av[NB];
                for (int nb = 0; nb < NB; ++nb) av[nb] = wd[cg * coutp + nb * 16];
                for (int m = 0; m < kConvTiles; ++m) {
                    if (wave + m * kConvWaves >= ntiles) break;     // uniform per wavefront
                    const float bv = slab[off[m] + cg * plane];
                    for (int nb = 0; nb < NB; ++nb) acc[m][nb] = c.mfma32(av[nb], bv, acc[m][nb]);
                }
            }
        }
    }

    const size_t vol = (size_t)lt * plane;
    for (int m = 0; m < kConvTiles; ++m) {
        const int tile = wave + m * kConvWaves;
        if (tile >= ntiles) break;
        const int pos = tile * 16 + jl;
        if (pos >= npos) continue;
        const size_t p = (size_t)t0 * plane + pos;
        for (int nb = 0; nb < NB; ++nb)
            for (int r = 0; r < 4; ++r) {
                const int co = nb * 16 + mfma32_row(lane, r);
                if (co >= a.cout) continue;
                float v = acc[m][nb][r] + (a.bias ? a.bias[co] : 0.0f);
                if (a.relu && v < 0.0f) v = 0.0f;
                const size_t o = ((size_t)b * a.cout + co) * vol + p;
                if (a.mask && !(a.mask[o] > 0.0f)) v = 0.0f;
                a.out[o] = v;
            }
    }
}

}  // namespace bafq
