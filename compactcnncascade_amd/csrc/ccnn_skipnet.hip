// Learned tile skipping: a small proposal network that predicts, from a 1/8-downsampled
// gray frame, which stage-1 tiles of which pyramid levels can hold a window that the
// cascade accepts, so that the others are not computed. No counterpart in the reference:
// the plain cascade is its fallback (tiles the network is not confident about run as
// usual) and the reference its recall is measured against.
//
// Network (SkipNet): input x0 = mean of each 8x8 block of the FP32 gray image / 255 - 0.5
// (H8 x W8 = ceil(H/8) x ceil(W/8)); three 3x3 convolutions with zero padding ("same"),
// 1 -> 8 -> 8 -> SLOTS channels, ReLU after the first two; output channel s = the logit
// that pyramid level with scale index s (InitScales order) has an accepted window whose
// centre lies in the 8x8 frame cell. Levels with scale index >= SLOTS or scale < 0.5 are
// always run.
//
// Kernels (FP32; the inference weights are the trained ones rounded to FP16):
//   k_sn_down8      the network input of every frame of a batch from its gray image
//   k_sn_conv       one 3x3 layer over a set of equally or differently sized frames
//                   (one thread per position, every output channel)
//   k_sn_tile_mask  keep flag of every stage-1 response tile of a batch
//   k_sn_l1_mask    keep flag of every L1 tile: set when a kept response tile reads it
//   k_sn_conv_bwd_data, k_sn_conv_bwd_w, k_sn_bce, k_sn_adam   training
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ccnn.h"
#include "ccnn_detector.h"

namespace ccnn {

__device__ __forceinline__ int sn_frame_of(const SnFrame* __restrict__ fr, int nf, int pos)
{
    int lo = 0, hi = nf - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (fr[mid].pos_off <= pos) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// x0 of every frame: levels' LevelDesc of the batch give each frame's gray image; one
// thread per cell. frames[f] = {h8, w8, off, pos_off} with off the frame's cell offset.
__global__ void __launch_bounds__(256) k_sn_down8(const float* __restrict__ gray, const FrameDesc* __restrict__ ft,
                                                  const SnFrame* __restrict__ fr, int nf, int total,
                                                  float* __restrict__ x0)
{
    const int pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= total) return;
    const int f = sn_frame_of(fr, nf, pos);
    const SnFrame F = fr[f];
    const int c = pos - F.pos_off;
    const int v = c / F.w8, u = c - v * F.w8;
    const FrameDesc& D = ft[f];
    const float* g = gray + D.gray_off;
    const int y0 = 8 * v, x0i = 8 * u;
    const int y1 = min(y0 + 8, D.h), x1 = min(x0i + 8, D.w);
    float s = 0.f;
    for (int y = y0; y < y1; ++y)
        for (int x = x0i; x < x1; ++x) s += g[(size_t)y * D.gp + x];
    const float n = (float)((y1 - y0) * (x1 - x0i));
    x0[F.off + c] = s / n * (1.f / 255.f) - 0.5f;
}

// out[f][co][y][x] = relu?(b[co] + sum_{ci,ky,kx} w[co][ci][ky][kx] in[f][ci][y+ky-1][x+kx-1])
// with zero padding. One thread per frame position computes every output channel from the
// 9 * CIN inputs it loads once; the weights are read from LDS.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_sn_conv(const float* __restrict__ in, const float* __restrict__ w,
                                                 const float* __restrict__ b, float* __restrict__ out,
                                                 const SnFrame* __restrict__ fr, int nf, int total, int relu)
{
    __shared__ float s_w[COUT * CIN * 9 + COUT];
    for (int i = threadIdx.x; i < COUT * CIN * 9; i += 256) s_w[i] = w[i];
    for (int i = threadIdx.x; i < COUT; i += 256) s_w[COUT * CIN * 9 + i] = b[i];
    __syncthreads();
    const int pos = blockIdx.x * 256 + threadIdx.x;
    if (pos >= total) return;
    const int f = sn_frame_of(fr, nf, pos);
    const SnFrame F = fr[f];
    const int c = pos - F.pos_off;
    const int y = c / F.w8, x = c - y * F.w8;
    const int plane = F.h8 * F.w8;
    const float* src = in + (size_t)F.off * CIN;
    float v[CIN][9];
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int yy = y + k / 3 - 1, xx = x + k % 3 - 1;
            v[ci][k] = (yy >= 0 && yy < F.h8 && xx >= 0 && xx < F.w8) ? src[(size_t)ci * plane + yy * F.w8 + xx] : 0.f;
        }
    float* dst = out + (size_t)F.off * COUT + c;
    for (int co = 0; co < COUT; ++co) {
        float acc = s_w[COUT * CIN * 9 + co];
        const float* k = s_w + co * CIN * 9;
#pragma unroll
        for (int ci = 0; ci < CIN; ++ci)
#pragma unroll
            for (int t = 0; t < 9; ++t) acc = __builtin_fmaf(k[ci * 9 + t], v[ci][t], acc);
        if (relu) acc = fmaxf(acc, 0.f);
        dst[(size_t)co * plane] = acc;
    }
}

// Gradient of the loss with respect to the layer input: gin[f][ci][y][x] =
// sum_{co,ky,kx} gpre[f][co][y-ky+1][x-kx+1] w[co][ci][ky][kx], times the ReLU mask of the
// input (act_in > 0) when the input is a ReLU output.
__global__ void __launch_bounds__(256) k_sn_conv_bwd_data(const float* __restrict__ gpre, const float* __restrict__ w,
                                                          const float* __restrict__ act_in, float* __restrict__ gin,
                                                          const SnFrame* __restrict__ fr, int nf, int total, int cin,
                                                          int cout)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= total * cin) return;
    const int ci = t / total, pos = t - ci * total;
    const int f = sn_frame_of(fr, nf, pos);
    const SnFrame F = fr[f];
    const int c = pos - F.pos_off;
    const int y = c / F.w8, x = c - y * F.w8;
    const int plane = F.h8 * F.w8;
    const float* g = gpre + (size_t)F.off * cout;
    float acc = 0.f;
    for (int co = 0; co < cout; ++co) {
        const float* k = w + (co * cin + ci) * 9;
        const float* p = g + (size_t)co * plane;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int yy = y - ky + 1;
            if (yy < 0 || yy >= F.h8) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int xx = x - kx + 1;
                if (xx < 0 || xx >= F.w8) continue;
                acc = __builtin_fmaf(k[ky * 3 + kx], p[yy * F.w8 + xx], acc);
            }
        }
    }
    const size_t o = (size_t)F.off * cin + (size_t)ci * plane + c;
    gin[o] = act_in[o] > 0.f ? acc : 0.f;
}

// Weight and bias gradients: dw[co][ci][k] += sum over positions of gpre[co] * in[ci] at
// the tap's offset; db[co] += sum gpre[co]. Grid (chunks of positions, cout * cin); each
// workgroup reduces its chunk and adds its partial sums atomically.
__global__ void __launch_bounds__(256) k_sn_conv_bwd_w(const float* __restrict__ gpre, const float* __restrict__ in,
                                                       float* __restrict__ dw, float* __restrict__ db,
                                                       const SnFrame* __restrict__ fr, int nf, int total, int cin,
                                                       int cout)
{
    __shared__ float s_red[10][4];
    const int co = blockIdx.y / cin, ci = blockIdx.y - co * cin;
    float acc[10];
#pragma unroll
    for (int k = 0; k < 10; ++k) acc[k] = 0.f;
    for (int pos = blockIdx.x * 256 + threadIdx.x; pos < total; pos += gridDim.x * 256) {
        const int f = sn_frame_of(fr, nf, pos);
        const SnFrame F = fr[f];
        const int c = pos - F.pos_off;
        const int y = c / F.w8, x = c - y * F.w8;
        const int plane = F.h8 * F.w8;
        const float g = gpre[(size_t)F.off * cout + (size_t)co * plane + c];
        const float* p = in + (size_t)F.off * cin + (size_t)ci * plane;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int yy = y + ky - 1;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int xx = x + kx - 1;
                const float v = (yy >= 0 && yy < F.h8 && xx >= 0 && xx < F.w8) ? p[yy * F.w8 + xx] : 0.f;
                acc[ky * 3 + kx] = __builtin_fmaf(g, v, acc[ky * 3 + kx]);
            }
        }
        acc[9] += g;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 10; ++k) {
        float v = acc[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) s_red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 10) {
        const int k = threadIdx.x;
        const float v = s_red[k][0] + s_red[k][1] + s_red[k][2] + s_red[k][3];
        if (k < 9) atomicAdd(&dw[(co * cin + ci) * 9 + k], v);
        else if (ci == 0) atomicAdd(&db[co], v);
    }
}

// Weighted binary cross-entropy with logits, averaged over n outputs: loss += w_i *
// softplus(-z) for a positive label, softplus(z) for a negative one, w_i = pos_weight for
// positives and 1 otherwise; dz = w_i (sigmoid(z) - y) / n.
__global__ void __launch_bounds__(256) k_sn_bce(const float* __restrict__ z, const unsigned char* __restrict__ y,
                                                float* __restrict__ dz, float* __restrict__ loss, int n,
                                                float pos_weight)
{
    __shared__ float s_l[4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float zi = z[i];
        const bool pos = y[i] != 0;
        const float wi = pos ? pos_weight : 1.f;
        const float sp = fmaxf(zi, 0.f) + log1pf(expf(-fabsf(zi)));      // softplus(z)
        l = wi * (pos ? sp - zi : sp);
        const float sg = 1.f / (1.f + expf(-zi));
        dz[i] = wi * (sg - (pos ? 1.f : 0.f)) / (float)n;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0) s_l[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (s_l[0] + s_l[1] + s_l[2] + s_l[3]) / (float)n);
}

// Adam (Kingma & Ba, 2015) with bias correction.
__global__ void __launch_bounds__(256) k_sn_adam(float* __restrict__ p, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                 float b1, float b2, float eps, int t)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// Keep flag of every response tile of a batch: the largest logit of the slots within
// `srad` of L.scl over the frame cells that hold the centres of the tile's windows (level
// position 4 * x + 11.5, 4 * y + 13.5, divided by the scale), widened by `margin` cells,
// against the logit threshold of slot L.scl. Levels whose scale index has no slot are kept. One wave
// per tile: the lanes share the tile's (slot, cell) pairs and reduce their maxima.
__global__ void __launch_bounds__(256) k_sn_tile_mask(const LevelDesc* __restrict__ levels, int nlevels, int ntiles,
                                                      int TX, int TY, const float* __restrict__ logits,
                                                      const SnFrame* __restrict__ fr, SkipThresholds th, int srad,
                                                      int margin,
                                                      unsigned char* __restrict__ keep)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= ntiles) return;
    int lo = 0, hi = nlevels - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (levels[mid].tile_off <= t) lo = mid; else hi = mid - 1;
    }
    const LevelDesc& L = levels[lo];
    int k = 1;
    // levels of scale < 0.5 are cheap (their pixel count falls with the square of the
    // scale) and their tiles span many cells: they always run
    if (L.scl < CCNN_SKIP_SLOTS && L.scale >= 0.5f) {
        const int tt = t - L.tile_off;
        const int tx = tt % L.tiles_x, ty = tt / L.tiles_x;
        const SnFrame F = fr[L.frame];
        // responses of the tile: x in [x0, x1), y in [y0, y1)
        const int x0 = tx * TX, x1 = min(x0 + TX, L.rw), y0 = ty * TY, y1 = min(y0 + TY, L.rh);
        const float inv = 1.f / (8.f * L.scale);
        const int u0 = max(0, (int)floorf((4.f * x0 + 11.5f) * inv) - margin);
        const int u1 = min(F.w8 - 1, (int)floorf((4.f * (x1 - 1) + 11.5f) * inv) + margin);
        const int v0 = max(0, (int)floorf((4.f * y0 + 13.5f) * inv) - margin);
        const int v1 = min(F.h8 - 1, (int)floorf((4.f * (y1 - 1) + 13.5f) * inv) + margin);
        const int s0 = max(0, L.scl - srad), s1 = min(CCNN_SKIP_SLOTS - 1, L.scl + srad);
        const size_t plane = (size_t)F.h8 * F.w8;
        const float* z = logits + (size_t)F.off * CCNN_SKIP_SLOTS + (size_t)s0 * plane;
        const int nu = u1 - u0 + 1, nv = v1 - v0 + 1;
        // a tile whose window centres fall outside the cell grid has no logit: it runs
        if (nu > 0 && nv > 0) {
            const int n = nu * nv, ns = s1 - s0 + 1;
            float mz = -INFINITY;
            for (int i = lane; i < n * ns; i += 64) {
                const int sl = i / n, c = i - sl * n;
                const int v = v0 + c / nu, u = u0 + c % nu;
                mz = fmaxf(mz, z[sl * plane + (size_t)v * F.w8 + u]);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) mz = fmaxf(mz, __shfl_xor(mz, off, 64));
            k = mz >= th.zt[L.scl];
        }
    }
    if (lane == 0) keep[t] = (unsigned char)k;
}

// Keep flag of every L1 tile: set when a kept response tile of its level reads it (the
// response tile (tx, ty) of TX x TY responses reads pooled L1 columns [2 TX tx, 2 TX tx +
// 2 TX + 8) and rows [2 TY ty, 2 TY ty + 2 TY + 10)).
__global__ void __launch_bounds__(256) k_sn_l1_mask(const LevelDesc* __restrict__ levels, int nlevels, int nl1tiles,
                                                    int ax, int ay, int tx_, int ty_,
                                                    const unsigned char* __restrict__ keep,
                                                    unsigned char* __restrict__ l1_keep)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nl1tiles) return;
    int lo = 0, hi = nlevels - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (levels[mid].l1_tile_off <= t) lo = mid; else hi = mid - 1;
    }
    const LevelDesc& L = levels[lo];
    const int tt = t - L.l1_tile_off;
    const int X0 = (tt % L.l1_tiles_x) * ax, Y0 = (tt / L.l1_tiles_x) * ay;
    // response tiles whose L1 window meets [X0, X0 + ax) x [Y0, Y0 + ay)
    const int rx0 = max(0, (X0 - (2 * tx_ + 8) + 1 + 2 * tx_ - 1) / (2 * tx_));
    const int rx1 = min(L.tiles_x - 1, (X0 + ax - 1) / (2 * tx_));
    const int ry0 = max(0, (Y0 - (2 * ty_ + 10) + 1 + 2 * ty_ - 1) / (2 * ty_));
    const int ry1 = min(L.tiles_y - 1, (Y0 + ay - 1) / (2 * ty_));
    int k = 0;
    for (int ry = ry0; ry <= ry1 && !k; ++ry)
        for (int rx = rx0; rx <= rx1 && !k; ++rx) k = keep[L.tile_off + ry * L.tiles_x + rx];
    l1_keep[t] = (unsigned char)k;
}

namespace {

constexpr int S = CCNN_SKIP_SLOTS;
// parameter layout: w1 [8][1][9], b1 [8], w2 [8][8][9], b2 [8], w3 [S][8][9], b3 [S]
constexpr int oW1 = 0, oB1 = oW1 + 8 * 9, oW2 = oB1 + 8, oB2 = oW2 + 8 * 8 * 9, oW3 = oB2 + 8, oB3 = oW3 + S * 8 * 9;
constexpr int NPAR = oB3 + S;
static_assert(NPAR == CCNN_SKIP_PARAMS, "parameter count");

template <class T>
struct DBuf {
    T* p = nullptr;
    ~DBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t n) { return hipMalloc((void**)&p, sizeof(T) * (n > 0 ? n : 1)) == hipSuccess; }
};

#define SN_OK(expr)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            note_error(e_);                                                                  \
            std::fprintf(stderr, "[ccnn] HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                                \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

// Device state of one training or gradient run over frames of one size h8 x w8.
struct SnRun {
    int n = 0, h8 = 0, w8 = 0, total = 0;
    DBuf<SnFrame> fr;
    DBuf<float> x0, a1, a2, z, g3, g2, g1, dz, par, grad, m, v, loss;
    DBuf<unsigned char> y;

    int init(int n_, int h8_, int w8_)
    {
        n = n_; h8 = h8_; w8 = w8_;
        total = n * h8 * w8;
        std::vector<SnFrame> f(n);
        for (int i = 0; i < n; ++i) f[i] = SnFrame{h8, w8, i * h8 * w8, i * h8 * w8};
        const size_t P = (size_t)total;
        if (!fr.alloc(n) || !x0.alloc(P) || !a1.alloc(8 * P) || !a2.alloc(8 * P) || !z.alloc(S * P) ||
            !g2.alloc(8 * P) || !g1.alloc(8 * P) || !dz.alloc(S * P) || !par.alloc(NPAR) || !grad.alloc(NPAR) ||
            !m.alloc(NPAR) || !v.alloc(NPAR) || !loss.alloc(1) || !y.alloc(S * P))
            return -1;
        SN_OK(hipMemcpy(fr.p, f.data(), sizeof(SnFrame) * n, hipMemcpyHostToDevice));
        return 0;
    }
    // forward and backward over the frames in x0 / y: grad and loss
    int step(float pos_weight)
    {
        const int T = total;
        auto blocks = [](long long k) { return (unsigned)((k + 255) / 256); };
        k_sn_conv<1, 8><<<blocks(T), 256>>>(x0.p, par.p + oW1, par.p + oB1, a1.p, fr.p, n, T, 1);
        k_sn_conv<8, 8><<<blocks(T), 256>>>(a1.p, par.p + oW2, par.p + oB2, a2.p, fr.p, n, T, 1);
        k_sn_conv<8, S><<<blocks(T), 256>>>(a2.p, par.p + oW3, par.p + oB3, z.p, fr.p, n, T, 0);
        SN_OK(hipMemsetAsync(grad.p, 0, sizeof(float) * NPAR, nullptr));
        SN_OK(hipMemsetAsync(loss.p, 0, sizeof(float), nullptr));
        k_sn_bce<<<blocks((long long)T * S), 256>>>(z.p, y.p, dz.p, loss.p, T * S, pos_weight);
        const unsigned wb = (unsigned)std::min<long long>(64, ((long long)T + 255) / 256);
        k_sn_conv_bwd_w<<<dim3(wb, S * 8), 256>>>(dz.p, a2.p, grad.p + oW3, grad.p + oB3, fr.p, n, T, 8, S);
        k_sn_conv_bwd_data<<<blocks((long long)T * 8), 256>>>(dz.p, par.p + oW3, a2.p, g2.p, fr.p, n, T, 8, S);
        k_sn_conv_bwd_w<<<dim3(wb, 8 * 8), 256>>>(g2.p, a1.p, grad.p + oW2, grad.p + oB2, fr.p, n, T, 8, 8);
        k_sn_conv_bwd_data<<<blocks((long long)T * 8), 256>>>(g2.p, par.p + oW2, a1.p, g1.p, fr.p, n, T, 8, 8);
        k_sn_conv_bwd_w<<<dim3(wb, 8 * 1), 256>>>(g1.p, x0.p, grad.p + oW1, grad.p + oB1, fr.p, n, T, 1, 8);
        SN_OK(hipGetLastError());
        return 0;
    }
};

}  // namespace

hipError_t launch_skipnet(const float* gray, const FrameDesc* ft, const SnFrame* fr, int nf, int cells,
                          const float* par, float* x0, float* a1, float* a2, float* z, const LevelDesc* levels, int nl,
                          int ntiles, int nl1tiles, int shape, const float* threshold, int srad, int margin,
                          unsigned char* keep, unsigned char* l1_keep, hipStream_t st)
{
    if (cells <= 0 || ntiles <= 0) return hipSuccess;
    const TileShape t = tile_shape(shape);
    SkipThresholds th;
    for (int k = 0; k < CCNN_SKIP_SLOTS; ++k) th.zt[k] = std::log(threshold[k] / (1.f - threshold[k]));
    auto blocks = [](long long k) { return (unsigned)((k + 255) / 256); };
    k_sn_down8<<<blocks(cells), 256, 0, st>>>(gray, ft, fr, nf, cells, x0);
    k_sn_conv<1, 8><<<blocks(cells), 256, 0, st>>>(x0, par + oW1, par + oB1, a1, fr, nf, cells, 1);
    k_sn_conv<8, 8><<<blocks(cells), 256, 0, st>>>(a1, par + oW2, par + oB2, a2, fr, nf, cells, 1);
    k_sn_conv<8, S><<<blocks(cells), 256, 0, st>>>(a2, par + oW3, par + oB3, z, fr, nf, cells, 0);
    k_sn_tile_mask<<<(unsigned)((ntiles + 3) / 4), 256, 0, st>>>(levels, nl, ntiles, t.tx, t.ty, z, fr, th, srad, margin, keep);
    k_sn_l1_mask<<<blocks(nl1tiles), 256, 0, st>>>(levels, nl, nl1tiles, t.ax, t.ay, t.tx, t.ty, keep, l1_keep);
    return hipGetLastError();
}

}  // namespace ccnn

using namespace ccnn;

extern "C" int ccnn_skipnet_grad(const float* params, const float* x0, const unsigned char* labels, int n, int h8,
                                 int w8, float pos_weight, float* grad, float* loss)
{
    if (!params || !x0 || !labels || !grad || n <= 0 || h8 <= 0 || w8 <= 0) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -1;
    SnRun r;
    if (r.init(n, h8, w8) < 0) return -1;
    const size_t P = (size_t)n * h8 * w8;
    SN_OK(hipMemcpy(r.par.p, params, sizeof(float) * NPAR, hipMemcpyHostToDevice));
    SN_OK(hipMemcpy(r.x0.p, x0, sizeof(float) * P, hipMemcpyHostToDevice));
    SN_OK(hipMemcpy(r.y.p, labels, S * P, hipMemcpyHostToDevice));
    if (r.step(pos_weight) < 0) return -1;
    SN_OK(hipDeviceSynchronize());
    SN_OK(hipMemcpy(grad, r.grad.p, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
    if (loss) SN_OK(hipMemcpy(loss, r.loss.p, sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ccnn_skipnet_forward(const float* params, const float* x0, int n, int h8, int w8, float* logits)
{
    if (!params || !x0 || !logits || n <= 0 || h8 <= 0 || w8 <= 0) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -1;
    SnRun r;
    if (r.init(n, h8, w8) < 0) return -1;
    const size_t P = (size_t)n * h8 * w8;
    SN_OK(hipMemcpy(r.par.p, params, sizeof(float) * NPAR, hipMemcpyHostToDevice));
    SN_OK(hipMemcpy(r.x0.p, x0, sizeof(float) * P, hipMemcpyHostToDevice));
    const int T = r.total;
    k_sn_conv<1, 8><<<(T + 255) / 256, 256>>>(r.x0.p, r.par.p + oW1, r.par.p + oB1, r.a1.p, r.fr.p, n, T, 1);
    k_sn_conv<8, 8><<<(T + 255) / 256, 256>>>(r.a1.p, r.par.p + oW2, r.par.p + oB2, r.a2.p, r.fr.p, n, T, 1);
    k_sn_conv<8, S><<<(T + 255) / 256, 256>>>(r.a2.p, r.par.p + oW3, r.par.p + oB3, r.z.p, r.fr.p, n, T, 0);
    SN_OK(hipGetLastError());
    SN_OK(hipDeviceSynchronize());
    SN_OK(hipMemcpy(logits, r.z.p, sizeof(float) * S * P, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int ccnn_skipnet_train(float* params, const float* x0, const unsigned char* labels, int n, int h8, int w8,
                                  int epochs, int batch, float lr, float pos_weight, unsigned seed,
                                  float* epoch_loss)
{
    if (!params || !x0 || !labels || n <= 0 || h8 <= 0 || w8 <= 0 || epochs <= 0 || batch <= 0) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -1;
    batch = std::min(batch, n);
    SnRun r;
    if (r.init(batch, h8, w8) < 0) return -1;
    const size_t F = (size_t)h8 * w8;
    // the whole set on the device; each step gathers its batch's frames
    DBuf<float> all_x;
    DBuf<unsigned char> all_y;
    if (!all_x.alloc(F * n) || !all_y.alloc(S * F * n)) return -1;
    SN_OK(hipMemcpy(all_x.p, x0, sizeof(float) * F * n, hipMemcpyHostToDevice));
    SN_OK(hipMemcpy(all_y.p, labels, S * F * n, hipMemcpyHostToDevice));
    SN_OK(hipMemcpy(r.par.p, params, sizeof(float) * NPAR, hipMemcpyHostToDevice));
    SN_OK(hipMemset(r.m.p, 0, sizeof(float) * NPAR));
    SN_OK(hipMemset(r.v.p, 0, sizeof(float) * NPAR));
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    unsigned long long st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto rnd = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (unsigned)(st >> 33);
    };
    int t = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int i = n - 1; i > 0; --i) std::swap(order[i], order[rnd() % (unsigned)(i + 1)]);
        double lsum = 0.0;
        int nb = 0;
        for (int b0 = 0; b0 + batch <= n; b0 += batch) {
            for (int k = 0; k < batch; ++k) {
                const int f = order[b0 + k];
                SN_OK(hipMemcpyAsync(r.x0.p + F * k, all_x.p + F * f, sizeof(float) * F, hipMemcpyDeviceToDevice,
                                     nullptr));
                // labels of the batch in [frame][slot][cell] order, as the logits
                SN_OK(hipMemcpyAsync(r.y.p + S * F * k, all_y.p + S * F * f, S * F, hipMemcpyDeviceToDevice, nullptr));
            }
            if (r.step(pos_weight) < 0) return -1;
            ++t;
            k_sn_adam<<<(NPAR + 255) / 256, 256>>>(r.par.p, r.grad.p, r.m.p, r.v.p, NPAR, lr, 0.9f, 0.999f, 1e-8f, t);
            float l = 0.f;
            SN_OK(hipMemcpy(&l, r.loss.p, sizeof(float), hipMemcpyDeviceToHost));
            lsum += l;
            ++nb;
        }
        if (epoch_loss) epoch_loss[e] = nb ? (float)(lsum / nb) : 0.f;
    }
    SN_OK(hipDeviceSynchronize());
    SN_OK(hipMemcpy(params, r.par.p, sizeof(float) * NPAR, hipMemcpyDeviceToHost));
    return 0;
}
