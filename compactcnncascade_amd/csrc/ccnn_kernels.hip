// gfx950 kernels of the compact-CNN-cascade detection path.
//
// Exact path: every floating-point expression is evaluated in the order the reference's
// CPU build evaluates it (built with -ffp-contract=off, so a*b + c stays a rounded
// multiply followed by a rounded add), which makes gray conversion, the pyramid, the
// response maps, the crops and the stage-2/3 responses bit-identical to the reference
// CPU path. References are to Bkmz21/CompactCNNCascade src/CNNObjectDetector/.
//
//   k_gray         u8 BGR/BGRA/gray -> FP32 gray             (image_proc.cpp:99-102,195-201,286-295)
//   k_level        one pyramid level, materialised (tests)   (image_resize.cpp:46-122,328-541)
//   k_stage1_l1    pyramid sampling + L1 + neighbour sums   (cnn_simd_cntk.cpp:640-725,
//                  of every level, once                       cnnpp_cplusplus.cpp)
//   k_stage1       per 16x16 response tile: L2, L3, hidden   (cnn_simd_cntk.cpp:726-963)
//                  and output layers; appends the candidate  (cnn_detector_v3.cpp:1918-1947)
//                  windows (score > t1) to the frame's list
//   k_check        per (candidate, crop variant): crop, NN   (cnn_detector_v3.cpp:1485-1818,
//                  resize to 49x51, u8, equalize, stage 2,    image_proc.cpp:357-368,720-800)
//                  stage 3 on the mirrored crop when needed
//   k_check_decide CPUCheckDetect's decision over mod 0 / mod 1 (stage-level entry; the
//                  frame path decides on the host, k_facial on its own)
//   k_resize_u8    8-bit bilinear pre-resize of 1-channel frames (image_resize.cpp:197-326)
//   k_facial       check decision + landmarks network on the      (cnn_detector_v3.cpp:1731-1754)
//                  accepted windows
//   k_landmarks    the landmarks network on one input (tests)
//   k_export       read-back of candidates and results into mapped host memory
//   k_selftest_tanh  the exact reciprocal against IEEE division over a range of floats

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include "ccnn_internal.h"

#include <type_traits>

namespace ccnn {

// Phase timestamps, compiled in only by profiling builds (CCNN_PHASE_STAMPS): thread 0
// of each of the first PHASE_BLOCKS workgroups of k_check and of k_stage1 records the
// 100 MHz wall clock (s_memrealtime) at fixed points of its work; ccnn_phase_stamps reads
// them. Slots, k_check (per workgroup, its last item): 0 item start, 1 crop prepared,
// 2 stage-2 input written, 3-8 after each barrier of the stage-2 network (L1, L2, sums,
// L3, hidden layer, output layer), 9-14 the same for stage 3, 15 item end, 16 stage-3
// input written; 20 hardware id, 21 item, 22 stage 3 ran. k_stage1 (per tile): 0 start,
// 1 L1-sum window in LDS, 2 L2 done, 3 L2 sums and tail weights in LDS, 4 response
// written (thread 0's), 20 hardware id.
// The stamps collect in LDS (so that no global address stays live in registers) and go
// to memory once, at the end of the workgroup's work (PHASE_FLUSH).
constexpr int PHASE_BLOCKS = 65536, PHASE_SLOTS = 24;
#ifdef CCNN_PHASE_STAMPS
__device__ unsigned long long g_phase[2][PHASE_BLOCKS][PHASE_SLOTS];
__shared__ unsigned long long s_phase[PHASE_SLOTS];
#define PHASE_SET(kern, slot, v)                   \
    do {                                           \
        if (threadIdx.x == 0) s_phase[slot] = (v); \
    } while (0)
#define PHASE_CLEAR()                                                       \
    do {                                                                    \
        if (threadIdx.x == 0)                                               \
            for (int i_ = 0; i_ < PHASE_SLOTS; ++i_) s_phase[i_] = 0;        \
    } while (0)
#define PHASE_FLUSH(kern)                                                                       \
    do {                                                                                        \
        if (threadIdx.x == 0 && blockIdx.x < PHASE_BLOCKS)                                      \
            for (int i_ = 0; i_ < PHASE_SLOTS; ++i_) g_phase[kern][blockIdx.x][i_] = s_phase[i_]; \
    } while (0)
#else
#define PHASE_SET(kern, slot, v) \
    do {                         \
    } while (0)
#define PHASE_CLEAR() \
    do {              \
    } while (0)
#define PHASE_FLUSH(kern) \
    do {                  \
    } while (0)
#endif
#define PHASE(kern, slot) PHASE_SET(kern, slot, wall_clock64())
// the stamps cost k_stage1 registers; the profiling build keeps its occupancy at the
// default build's 5 waves per SIMD so that the stamped phases are those of the product
#ifdef CCNN_PHASE_STAMPS
#define S1_WAVES __attribute__((amdgpu_waves_per_eu(5)))
#else
#define S1_WAVES
#endif
enum { PH_CHECK = 0, PH_STAGE1 = 1 };

// The reciprocal of the rational tanh's denominator d >= 1, as 1 - 1/d needs it: 1/d
// rounded to nearest for 1 <= d < 2^25, from the hardware reciprocal estimate (within
// 1 ulp) and one Newton-Raphson step r + r(1 - d r) evaluated with fused multiply-adds
// (Markstein's correction); 0 for d >= 2^25 (and +inf), where 1/d <= 2^-25 and 1 - 1/d
// rounds to 1 either way. A NaN stays a NaN. Five instructions instead of the general
// IEEE division's scale/fixup sequence; ccnn_selftest_tanh checks the reciprocal against
// IEEE division on every float in [1, 2^25) and the tanh on every float.
__device__ __forceinline__ float recip_ge1(float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, r, 1.f);
    const float q = __builtin_fmaf(e, r, r);
    return d >= 33554432.f ? 0.f : q;
}

template <bool IEEE_DIV = false>
__device__ __forceinline__ float rtanh(float c)
{
    // sgn * (1 - 1/(1 + |c| + c^2 + a c^4)), cnnpp_cplusplus.cpp:662-666, with the
    // association of the reference's -ffast-math build: ((|c| + c*c) + a*c*c*c*c) + 1.
    const float a = 1.41645f;
    const float den = ((fabsf(c) + c * c) + a * c * c * c * c) + 1.f;
    const float t = 1.f - (IEEE_DIV ? 1.f / den : recip_ge1(den));
    // the reference multiplies by sgn = -1 or 1, which is the negation or t itself
    return c < 0.f ? -t : t;
}

__device__ __forceinline__ float lrelu_bn(float conv, const float* __restrict__ p, int m, int M)
{
    // w1*c + w2*max(0,c) + bn_b with c = conv + conv_b (cnnpp_cplusplus.cpp:604-605,
    // 630-631), associated as the reference's build evaluates it.
    const float c = conv + p[m];
    return (p[2 * M + m] * fmaxf(0.f, c) + p[3 * M + m]) + p[M + m] * c;
}

// ---------------------------------------------------------------------------------
// Reduced-precision stage 1 (FP16 convolutions, FP32 accumulation): the inputs of every
// convolution and of the hidden layer are rounded to FP16 and multiplied pairwise with
// FP16 weights by v_dot2_f32_f16, which adds both exact products to an FP32 accumulator;
// activations, pooling, sums and the tanh stay in FP32 (fused multiply-adds allowed).
// The FP16 weights are laid out after the FP32 ones (Stage1H, filled by the host).
// ---------------------------------------------------------------------------------
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ h2_t pk16(float a, float b)
{
    h2_t r;
    r.x = (_Float16)a;
    r.y = (_Float16)b;
    return r;
}

__device__ __forceinline__ h2_t as_h2(uint32_t u) { return __builtin_bit_cast(h2_t, u); }

__device__ __forceinline__ float dot2(h2_t a, h2_t b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }

__device__ __forceinline__ float lrelu_bn_fma(float conv, const float* __restrict__ p, int m, int M)
{
    const float c = conv + p[m];
    return __builtin_fmaf(p[M + m], c, __builtin_fmaf(p[2 * M + m], fmaxf(0.f, c), p[3 * M + m]));
}

// Sample of pyramid level L at (lx, ly), from the FP32 gray image
// (ImageResizer::preprocessing + Nearest/BilinearInterpolation, scalar tails).
__device__ __forceinline__ float level_px(const float* __restrict__ g, int gp, int gw, int gh,
                                          const LevelDesc& L, int lx, int ly)
{
    if (L.mode == 2) return g[(size_t)ly * gp + lx];
    const uint32_t y = (uint32_t)ly * L.scly;
    uint32_t py = y >> 12;
    const uint32_t x = (uint32_t)lx * L.sclx;
    uint32_t px = x >> 12;
    if (L.mode == 0) {
        py = min(py, (uint32_t)(gh - 1));
        px = min(px, (uint32_t)(gw - 2));
        return g[(size_t)py * gp + px];
    }
    float fy = (float)(y - (py << 12));
    float cy = 4096.f - fy;
    fy = fy / 16777216.f;
    cy = cy / 16777216.f;
    py = min(py, (uint32_t)(gh - 1));
    if (py + 1 >= (uint32_t)gh) py--;
    const float fx = (float)(x - (px << 12));
    const float cx = 4096.f - fx;
    px = min(px, (uint32_t)(gw - 2));
    const float* r0 = g + (size_t)py * gp + px;
    const float* r1 = r0 + gp;
    const float p0 = r0[0], p1 = r0[1], p2 = r1[0], p3 = r1[1];
    return (p0 * cx + p1 * fx) * cy + (p2 * cx + p3 * fx) * fy;
}

// The gray image a level samples: the kernel's arguments on the single-frame path, the
// level's own frame in a batch (LevelDesc::gp > 0).
struct GrayRef {
    const float* g;
    int gp, gw, gh;
};
__device__ __forceinline__ GrayRef level_gray(const float* g, int gp, int gw, int gh, const LevelDesc& L)
{
    if (L.gp > 0) return GrayRef{g + L.gray_off, L.gp, L.gw, L.gh};
    return GrayRef{g, gp, gw, gh};
}

// ---------------------------------------------------------------------------------
// Gray conversion
// ---------------------------------------------------------------------------------
__device__ __forceinline__ float gray_of(float B, float G, float R)
{
    // 0.114 B + 0.587 G + 0.299 R, summed left to right (image_proc.cpp:99-102)
    return 0.114f * B + 0.587f * G + 0.299f * R;
}

__device__ __forceinline__ float ub(uint32_t w, int k) { return (float)((w >> (8 * k)) & 255u); }

// Gray of the quad (4 pixels) at column x of row s of a frame w pixels wide into d: 4-byte
// loads (16-byte for BGRA) and one 16-byte store when `aligned` and the quad is whole,
// byte loads otherwise.
__device__ __forceinline__ void gray_quad(const uint8_t* __restrict__ s, int x, int w, int ch, bool aligned,
                                          float* __restrict__ d)
{
    if (aligned && x + 4 <= w) {
        float4 v;
        if (ch == 3) {
            const uint32_t* p = reinterpret_cast<const uint32_t*>(s + (size_t)x * 3);
            const uint32_t a = p[0], b = p[1], c = p[2];
            v.x = gray_of(ub(a, 0), ub(a, 1), ub(a, 2));
            v.y = gray_of(ub(a, 3), ub(b, 0), ub(b, 1));
            v.z = gray_of(ub(b, 2), ub(b, 3), ub(c, 0));
            v.w = gray_of(ub(c, 1), ub(c, 2), ub(c, 3));
        } else if (ch == 4) {
            const uint4 p = *reinterpret_cast<const uint4*>(s + (size_t)x * 4);
            v.x = gray_of(ub(p.x, 0), ub(p.x, 1), ub(p.x, 2));
            v.y = gray_of(ub(p.y, 0), ub(p.y, 1), ub(p.y, 2));
            v.z = gray_of(ub(p.z, 0), ub(p.z, 1), ub(p.z, 2));
            v.w = gray_of(ub(p.w, 0), ub(p.w, 1), ub(p.w, 2));
        } else {
            const uint32_t a = *reinterpret_cast<const uint32_t*>(s + x);
            v.x = ub(a, 0); v.y = ub(a, 1); v.z = ub(a, 2); v.w = ub(a, 3);
        }
        *reinterpret_cast<float4*>(d) = v;
    } else {
        for (int i = 0; i < 4 && x + i < w; ++i) {
            const uint8_t* t = s + (size_t)(x + i) * ch;
            d[i] = ch == 1 ? (float)t[0] : gray_of((float)t[0], (float)t[1], (float)t[2]);
        }
    }
}

// Gray of the 16 pixels of a row from x (x a multiple of 16) into d. aligned: bit 0 as
// gray_quad's `aligned`; bit 1 rows 16-byte aligned (frame, row step and gray), so that a
// whole group is read with 16-byte loads (three for BGR, four for BGRA, one for gray) and
// written with four 16-byte stores. Otherwise quad by quad.
__device__ __forceinline__ void gray_group16(const uint8_t* __restrict__ s, int x, int w, int ch, int aligned,
                                             float* __restrict__ d)
{
    if ((aligned & 2) && x + 16 <= w) {
        float g[16];
        if (ch == 3) {
            const uint4* p = reinterpret_cast<const uint4*>(s + (size_t)x * 3);
            const uint4 a = p[0], b = p[1], c = p[2];
            const uint32_t u[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
#pragma unroll
            for (int j = 0; j < 16; ++j)
                g[j] = gray_of(ub(u[(3 * j) / 4], (3 * j) % 4), ub(u[(3 * j + 1) / 4], (3 * j + 1) % 4),
                               ub(u[(3 * j + 2) / 4], (3 * j + 2) % 4));
        } else if (ch == 4) {
            const uint4* p = reinterpret_cast<const uint4*>(s + (size_t)x * 4);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint4 a = p[k];
                g[4 * k] = gray_of(ub(a.x, 0), ub(a.x, 1), ub(a.x, 2));
                g[4 * k + 1] = gray_of(ub(a.y, 0), ub(a.y, 1), ub(a.y, 2));
                g[4 * k + 2] = gray_of(ub(a.z, 0), ub(a.z, 1), ub(a.z, 2));
                g[4 * k + 3] = gray_of(ub(a.w, 0), ub(a.w, 1), ub(a.w, 2));
            }
        } else {
            const uint4 a = *reinterpret_cast<const uint4*>(s + x);
            const uint32_t u[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
            for (int j = 0; j < 16; ++j) g[j] = ub(u[j / 4], j % 4);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            reinterpret_cast<float4*>(d)[k] = make_float4(g[4 * k], g[4 * k + 1], g[4 * k + 2], g[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (x + 4 * k < w) gray_quad(s, x + 4 * k, w, ch, (aligned & 1) != 0, d + 4 * k);
    }
}

// One thread per 4 pixels of a row (a quad), rows flattened over the grid: 4-byte loads
// of the frame, one 16-byte store of gray. Needs a 4-byte aligned frame and row step and
// a gray pitch that is a multiple of 4 floats; the last quad of a row may be partial.
__global__ void __launch_bounds__(256) k_gray(const uint8_t* __restrict__ img, int w, int h, int ch, int step,
                                              float* __restrict__ gray, int gp, int qw)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= qw * h) return;
    const int y = q / qw;
    const int x = (q - y * qw) * 4;
    gray_quad(img + (size_t)y * step, x, w, ch, true, gray + (size_t)y * gp + x);
}

// Gray of every frame of a batch: the frames' groups of 16 pixels numbered back to back
// (FrameDesc::quads_off), each thread's frame found by binary search of the table.
__global__ void __launch_bounds__(256) k_gray_batch(const FrameDesc* __restrict__ frames, int nframes,
                                                    int total_quads, float* __restrict__ gray)
{
    // the frame of the block's first group, by a search on block-uniform values (scalar
    // loads), then each thread steps to a later frame when the block straddles frames
    const int q0 = blockIdx.x * 256;
    int lo = 0, hi = nframes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (frames[mid].quads_off <= q0) lo = mid; else hi = mid - 1;
    }
    const int q = q0 + threadIdx.x;
    if (q >= total_quads) return;
    while (lo + 1 < nframes && frames[lo + 1].quads_off <= q) ++lo;
    const FrameDesc& F = frames[lo];
    const int qq = q - F.quads_off;
    const int y = qq / F.qw;
    const int x = (qq - y * F.qw) * 16;
    gray_group16(F.src + (size_t)y * F.step, x, F.w, F.ch, F.aligned, gray + F.gray_off + (size_t)y * F.gp + x);
}

// Any alignment: one thread per pixel, byte loads.
__global__ void __launch_bounds__(256) k_gray_bytes(const uint8_t* __restrict__ img, int w, int h, int ch,
                                                    int step, float* __restrict__ gray, int gp)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w || y >= h) return;
    const uint8_t* s = img + (size_t)y * step + (size_t)x * ch;
    gray[(size_t)y * gp + x] = ch == 1 ? (float)s[0] : gray_of((float)s[0], (float)s[1], (float)s[2]);
}

// 8-bit bilinear resize (ImageResizer::BilinearInterpolation for Image_8u, scalar tail,
// image_resize.cpp:198-326): the gray_image_only pre-resize of a 1-channel frame.
__global__ void __launch_bounds__(256) k_resize_u8(const uint8_t* __restrict__ src, int sw, int sh, int sstep,
                                                   uint8_t* __restrict__ dst, int dw, int dh, uint32_t sclx,
                                                   uint32_t scly)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= dw || iy >= dh) return;
    const uint32_t y = (uint32_t)iy * scly;
    uint32_t py = y >> 12;
    float fy = (float)(y - (py << 12));
    float cy = 4096.f - fy;
    fy = fy / 16777216.f;
    cy = cy / 16777216.f;
    py = min(py, (uint32_t)(sh - 1));
    if (py + 1 >= (uint32_t)sh) py--;
    const uint32_t x = (uint32_t)ix * sclx;
    uint32_t px = x >> 12;
    const float fx = (float)(x - (px << 12));
    const float cx = 4096.f - fx;
    px = min(px, (uint32_t)(sw - 2));
    const uint8_t* r0 = src + (size_t)py * sstep + px;
    const uint8_t* r1 = r0 + sstep;
    const float p0 = r0[0], p1 = r0[1], p2 = r1[0], p3 = r1[1];
    const float v = (p0 * cx + p1 * fx) * cy + (p2 * cx + p3 * fx) * fy;
    dst[(size_t)iy * dw + ix] = (uint8_t)(uint32_t)v;
}

// The pre-resize of frame f of a batch (FrameDesc::orig -> src), as k_resize_u8.
__global__ void __launch_bounds__(256) k_resize_u8_tab(const FrameDesc* __restrict__ frames, int f)
{
    const FrameDesc& F = frames[f];
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= F.w || iy >= F.h) return;
    const uint32_t sclx = ((uint32_t)F.ow << 12) / (uint32_t)F.w + 1;
    const uint32_t scly = ((uint32_t)F.oh << 12) / (uint32_t)F.h + 1;
    const uint32_t y = (uint32_t)iy * scly;
    uint32_t py = y >> 12;
    float fy = (float)(y - (py << 12));
    float cy = 4096.f - fy;
    fy = fy / 16777216.f;
    cy = cy / 16777216.f;
    py = min(py, (uint32_t)(F.oh - 1));
    if (py + 1 >= (uint32_t)F.oh) py--;
    const uint32_t x = (uint32_t)ix * sclx;
    uint32_t px = x >> 12;
    const float fx = (float)(x - (px << 12));
    const float cx = 4096.f - fx;
    px = min(px, (uint32_t)(F.ow - 2));
    const uint8_t* r0 = F.orig + (size_t)py * F.ostep + px;
    const uint8_t* r1 = r0 + F.ostep;
    const float p0 = r0[0], p1 = r0[1], p2 = r1[0], p3 = r1[1];
    const float v = (p0 * cx + p1 * fx) * cy + (p2 * cx + p3 * fx) * fy;
    const_cast<uint8_t*>(F.src)[(size_t)iy * F.w + ix] = (uint8_t)(uint32_t)v;
}

// One level image (tests and getGrayImage-style inspection; the detection path samples
// levels inside k_stage1 instead of storing them).
__global__ void __launch_bounds__(256) k_level(const float* __restrict__ g, int gp, int gw, int gh, LevelDesc L,
                                               float* __restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= L.lw || y >= L.lh) return;
    out[(size_t)y * L.lw + x] = level_px(g, gp, gw, gh, L, x, y);
}

// ---------------------------------------------------------------------------------
// Stage 1 (cnn_simd_cntk.cpp:640-963, cnnpp_cplusplus.cpp), two launches per frame:
//   k_stage1_l1: level sampling + L1 (conv 4x4, act, 2x2 max-pool) + the neighbour sums
//                feeding L2, for every pooled L1 position of every level, once;
//   k_stage1:    per 16x16 response tile, L2 (conv 3x3, act, pool) + sums + L3 (conv
//                4w x 5h, act) + hidden layer + output layer from the L1 sums.
// Each pooled L1 position depends only on the level samples at (2X..2X+4, 2Y..2Y+4),
// so the split computes the same values as one fused pass without recomputing the
// L1 halo that neighbouring response tiles share.
// ---------------------------------------------------------------------------------
namespace s1 {
using N = Stage1;
// a response tile of TX x TY responses reads P1W x P1H L1 sums and computes P2W x P2H
// pooled L2 positions; an L1 tile of AX x AY pooled positions samples AW x AH level pixels
template <int TX, int TY>
struct RespGeo {
    static constexpr int P1W = 2 * TX + 8, P1H = 2 * TY + 10;
    static constexpr int P2W = TX + 3, P2H = TY + 4;
    static constexpr int P1N = P1W * P1H, P2N = P2W * P2H;
};
}  // namespace s1

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2): workgroup b
// runs on XCD b % 8. Renumber them so that each XCD takes one contiguous run of tiles,
// whose input windows overlap and share that XCD's L2. A bijection on [0, n).
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    constexpr int NX = 8;
    const int xcd = b % NX, local = b / NX;
    const int q = n / NX, r = n % NX;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Level of global work item t: the last level whose first item (tile_off or
// l1_tile_off) is <= t.
enum { BY_TILE = 0, BY_L1_TILE = 1 };
template <int BY>
__device__ __forceinline__ int find_level(const LevelDesc* __restrict__ L, int n, int t)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int off = BY == BY_TILE ? L[mid].tile_off : L[mid].l1_tile_off;
        if (off <= t) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int AX, int AY, bool H>
__global__ void __launch_bounds__(S1A_THREADS) k_stage1_l1(const float* __restrict__ g0, int gp0, int gw0, int gh0,
                                                           const LevelDesc* __restrict__ levels, int nlevels,
                                                           const float* __restrict__ W, float* __restrict__ s1buf,
                                                           int* __restrict__ meta, const int2* __restrict__ order,
                                                           const unsigned char* __restrict__ l1_keep)
{
    using N = Stage1;
    constexpr int AW = 2 * AX + 3, AH = 2 * AY + 3;     // level window of one L1 tile
    __shared__ float s_in[AH * AW];
    __shared__ uint32_t s_row[AH], s_px[AW];
    __shared__ float s_fx[AW], s_cx[AW], s_fy[AH], s_cy[AH];
    static_assert(AW + AH <= S1A_THREADS, "one table entry per thread");
    const int tid = threadIdx.x;
    // the candidate count that k_stage1 accumulates starts at zero, and so does the work
    // queue of a persistent k_check
    if (meta && blockIdx.x == 0 && tid == 0) {
        meta[0] = 0;
        meta[1] = 0;
    }
    // order (optional): the tiles as {level, tile within the level}, arranged so that the
    // contiguous run each XCD takes covers one horizontal band of every level, whose
    // gray-image rows then stay in that XCD's L2 across the levels
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    int li, t;
    if (order) {
        const int2 o = order[tile];
        li = o.x;
        t = o.y;
    } else {
        li = find_level<BY_L1_TILE>(levels, nlevels, tile);
        t = tile - levels[li].l1_tile_off;
    }
    // l1_keep (optional): L1 tiles that no kept response tile reads are skipped
    if (l1_keep && !l1_keep[levels[li].l1_tile_off + t]) return;
    const LevelDesc L = levels[li];
    const GrayRef gr = level_gray(g0, gp0, gw0, gh0, L);
    const float* __restrict__ g = gr.g;
    const int gp = gr.gp, gw = gr.gw, gh = gr.gh;
    const int X0 = (t % L.l1_tiles_x) * AX, Y0 = (t / L.l1_tiles_x) * AY;
    const int x0 = 2 * X0, y0 = 2 * Y0;

    // level window (clamped at the level border; clamped samples only feed L1 positions
    // whose responses lie outside the response map). The resize's Q12 LUT is separable
    // (ImageResizer::preprocessing, image_resize.cpp:46-122): a table per window column
    // (source column, x weights) and per window row (source row, y weights) is built
    // first, then every sample is a table lookup, its source pixels and the bilinear
    // formula of level_px. Samples go in batches so that their loads are in flight
    // together.
    {
        if (tid < AW) {
            const int lx = min(x0 + tid, L.lw - 1);
            uint32_t px = (uint32_t)lx;
            float fx = 0.f, cx = 0.f;
            if (L.mode != 2) {
                const uint32_t x = (uint32_t)lx * L.sclx;
                px = min(x >> 12, (uint32_t)(gw - 2));
                fx = (float)(x - ((x >> 12) << 12));
                cx = 4096.f - fx;
            }
            s_px[tid] = px;
            s_fx[tid] = fx;
            s_cx[tid] = cx;
        } else if (tid < AW + AH) {
            const int iy = tid - AW;
            const int ly = min(y0 + iy, L.lh - 1);
            uint32_t py = (uint32_t)ly;
            float fy = 0.f, cy = 0.f;
            if (L.mode == 0) {
                py = min(((uint32_t)ly * L.scly) >> 12, (uint32_t)(gh - 1));
            } else if (L.mode == 1) {
                const uint32_t y = (uint32_t)ly * L.scly;
                py = y >> 12;
                fy = (float)(y - (py << 12));
                cy = 4096.f - fy;
                fy = fy / 16777216.f;
                cy = cy / 16777216.f;
                py = min(py, (uint32_t)(gh - 1));
                if (py + 1 >= (uint32_t)gh) py--;
            }
            s_row[iy] = py * (uint32_t)gp;
            s_fy[iy] = fy;
            s_cy[iy] = cy;
        }
        __syncthreads();
        // batches of B samples per thread; past NS the loads are clamped and the stores skipped
        constexpr int NS = AW * AH, PER = (NS + S1A_THREADS - 1) / S1A_THREADS, B = PER < 9 ? PER : 9;
        if (L.mode == 1) {
#pragma unroll 1
            for (int b0 = 0; b0 < PER; b0 += B) {
                float p[B][4];
                int ixs[B], iys[B];
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    const int i = min(tid + (b0 + k) * S1A_THREADS, NS - 1);
                    const int iy = i / AW, ix = i - iy * AW;
                    const float* r0 = g + (s_row[iy] + s_px[ix]);
                    const float* r1 = r0 + gp;
                    p[k][0] = r0[0];
                    p[k][1] = r0[1];
                    p[k][2] = r1[0];
                    p[k][3] = r1[1];
                    ixs[k] = ix;
                    iys[k] = iy;
                }
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    const float cx = s_cx[ixs[k]], fx = s_fx[ixs[k]], cy = s_cy[iys[k]], fy = s_fy[iys[k]];
                    const float v = (p[k][0] * cx + p[k][1] * fx) * cy + (p[k][2] * cx + p[k][3] * fx) * fy;
                    const int i = tid + (b0 + k) * S1A_THREADS;
                    if (i < NS) s_in[i] = v;
                }
            }
        } else {
#pragma unroll 1
            for (int b0 = 0; b0 < PER; b0 += B) {
                float v[B];
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    const int i = min(tid + (b0 + k) * S1A_THREADS, NS - 1);
                    const int iy = i / AW, ix = i - iy * AW;
                    v[k] = g[s_row[iy] + s_px[ix]];
                }
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    const int i = tid + (b0 + k) * S1A_THREADS;
                    if (i < NS) s_in[i] = v[k];
                }
            }
        }
    }
    __syncthreads();

    // L1: conv_4x4 chains + lrelu_bn_max for every map, then the sums feeding L2
    // (add / add2, cnnpp_cplusplus.cpp:916-950; wiring cnn_simd_cntk.cpp:705-725)
    const float* k1 = W + N::oK1;
    const float* b1 = W + N::oB1;
    const size_t plane = (size_t)L.s1_w * L.s1_h;
    float* out = s1buf + L.s1_off;
    // the positions responses inside the map read (vx x vy of the tile; fewer at the
    // level's right and bottom edges), numbered densely so that edge tiles issue fewer
    // waves rather than masked lanes (p / vx from a float reciprocal: exact here)
    const int vx = min(AX, 2 * L.rw + 8 - X0), vy = min(AY, 2 * L.rh + 10 - Y0);
    const float rvx = 1.f / (float)vx;
    for (int p = tid; p < vx * vy; p += S1A_THREADS) {
        const int py = (int)(((float)p + 0.5f) * rvx), px = p - py * vx;
        const int X = X0 + px, Y = Y0 + py;
        float win[5][5];
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int c = 0; c < 5; ++c) win[r][c] = s_in[(2 * py + r) * AW + 2 * px + c];
        float pm[N::M1];
        if constexpr (H) {
            // the window's adjacent pairs (c, c+1) of each row in FP16, less the centre
            // value mu (the differences are small in smooth regions, so they round far
            // closer than the raw values, and the taps' rounding error scales with them
            // too); tap pairs (k[r][0], k[r][1]) and (k[r][2], k[r][3]) of each map in
            // Stage1H; each map's sum starts from mu x its FP32 tap sum
            const float mu = win[2][2];
            h2_t hp[5][4];
#pragma unroll
            for (int r = 0; r < 5; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) hp[r][c] = pk16(win[r][c] - mu, win[r][c + 1] - mu);
            const uint32_t* kh = reinterpret_cast<const uint32_t*>(W + N::SIZE) + Stage1H::oL1;
            const float* ks = W + N::SIZE + Stage1H::oS1;
#pragma unroll
            for (int m = 0; m < N::M1; ++m) {
                float a[4];
                const float acc0 = mu * ks[m];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int dx = q & 1, dy = q >> 1;
                    float acc = acc0;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        acc = dot2(hp[dy + r][dx], as_h2(kh[m * 8 + 2 * r]), acc);
                        acc = dot2(hp[dy + r][dx + 2], as_h2(kh[m * 8 + 2 * r + 1]), acc);
                    }
                    a[q] = lrelu_bn_fma(acc, b1, m, N::M1);
                }
                pm[m] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
            }
        } else {
#pragma unroll
        for (int m = 0; m < N::M1; ++m) {
            const float* k = k1 + m * 16;
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int dx = q & 1, dy = q >> 1;
                float acc = win[dy][dx] * k[0];
#pragma unroll
                for (int j = 1; j < 16; ++j) acc = acc + win[dy + (j >> 2)][dx + (j & 3)] * k[j];
                a[q] = lrelu_bn(acc, b1, m, N::M1);
            }
            pm[m] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
        }
        }
        float* o = out + (size_t)Y * L.s1_w + X;
        o[0] = pm[0] + pm[1];
#pragma unroll
        for (int i = 1; i < N::M1 - 1; ++i) o[i * plane] = pm[i - 1] + pm[i] + pm[i + 1];
        o[(N::M1 - 1) * plane] = pm[N::M1 - 2] + pm[N::M1 - 1];
    }
}

template <int TX, int TY, bool H>
__global__ void __launch_bounds__(S1_THREADS) S1_WAVES k_stage1(const LevelDesc* __restrict__ levels, int nlevels,
                                                       const float* __restrict__ W,
                                                       const float* __restrict__ s1buf, float* __restrict__ resp,
                                                       float t1, int* __restrict__ meta, Candidate* __restrict__ cands,
                                                       Candidate* __restrict__ hcands, int hcap,
                                                       const int2* __restrict__ order,
                                                       const unsigned char* __restrict__ keep)
{
    using N = Stage1;
    using G = s1::RespGeo<TX, TY>;
    constexpr int P1W = G::P1W, P1H = G::P1H, P2W = G::P2W, P1N = G::P1N, P2N = G::P2N;
    static_assert(TX * TY <= S1_THREADS, "one response value per thread");
    __shared__ __align__(16) float s_s1[N::M1 * P1N];     // L1 sums; later the L2 sums and tail weights
    // the L3, hidden- and output-layer weights, which the per-response tail reads from
    // LDS; fetched into registers now, stored over the consumed L2 maps later
    constexpr int WT = N::SIZE - N::oK3, WPER = (WT + S1_THREADS - 1) / S1_THREADS;


    const int tid = threadIdx.x;
    float wt[WPER];
#pragma unroll
    for (int k = 0; k < WPER; ++k) wt[k] = W[N::oK3 + min(tid + k * S1_THREADS, WT - 1)];
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    PHASE_CLEAR();
    PHASE(PH_STAGE1, 0);
    PHASE_SET(PH_STAGE1, 20, __smid());
    // order (optional): each tile's {level, tile within the level}, one load instead of
    // the binary search over the level table (11 dependent loads for a 64-frame batch)
    int li, t;
    if (order) {
        const int2 o = order[tile];
        li = o.x;
        t = o.y;
    } else {
        li = find_level<BY_TILE>(levels, nlevels, tile);
        t = tile - levels[li].tile_off;
    }
    const LevelDesc L = levels[li];
    const int ox0 = (t % L.tiles_x) * TX, oy0 = (t / L.tiles_x) * TY;
    // the tile's responses inside the map (vw x vh; fewer at the map's right and bottom
    // edges) and the pooled L2 positions they read (qw x qh). Work items are numbered
    // over these only, so that edge tiles issue fewer waves rather than masked lanes;
    // small quotients come from float reciprocals (exact for these ranges: the margins
    // to the next integer are far above float rounding).
    const int vw = min(TX, L.rw - ox0), vh = min(TY, L.rh - oy0);
    const int qw = vw + 3, qh = vh + 4, nq = qw * qh;
    const float rqw = 1.f / (float)qw, rnq = 1.f / (float)nq, rvw = 1.f / (float)vw;
    // keep (optional): a skipped tile's responses are -inf, below every threshold
    if (keep && !keep[tile]) {
        if (tid < vw * vh) {
            const int oy = (int)(((float)tid + 0.5f) * rvw), ox = tid - oy * vw;
            resp[L.resp_off + (oy0 + oy) * L.rw + ox0 + ox] = -INFINITY;
        }
        return;
    }

    // the tile's L1 sums: rows 2*oy0 .. +P1H, columns 2*ox0 .. +P1W of each plane
    // (inside the planes: s1_w = 2*TX*tiles_x + 8, s1_h = 2*TY*tiles_y + 10). Every
    // plane row and level offset is a multiple of 8 floats and 2*ox0 of 32, so the rows
    // are copied as float4; the M1 planes stack as M1*P1H rows of the tile window.
    {
        static_assert(P1W % 4 == 0 && (2 * TX) % 4 == 0, "float4 rows");
        constexpr int RQ = P1W / 4, NQ = N::M1 * P1H * RQ;
        constexpr int PER = (NQ + S1_THREADS - 1) / S1_THREADS, B = PER;   // one batch
        const uint32_t plane = (uint32_t)L.s1_w * L.s1_h;
        const float* src = s1buf + L.s1_off + (size_t)(2 * oy0) * L.s1_w + 2 * ox0;
        float4 v[B];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int i = min(tid + k * S1_THREADS, NQ - 1);
            const int row = i / RQ, q = i - row * RQ;       // row of the stacked planes
            const int m = row / P1H, py = row - m * P1H;
            v[k] = *reinterpret_cast<const float4*>(src + m * plane + (uint32_t)py * L.s1_w + 4 * q);
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int i = tid + k * S1_THREADS;
            if (i < NQ) *reinterpret_cast<float4*>(s_s1 + 4 * i) = v[k];
        }
    }
    __syncthreads();
    PHASE(PH_STAGE1, 1);

    // L2, one pooled L2 position per thread: the eight maps (two conv_3x3 chains per L1
    // sum + lrelu_bn_max), then their neighbour sums, held in registers until every thread
    // has read the L1 sums and then written over them
    const float* k2 = W + N::oK2;
    const float* b2 = W + N::oB2;
    constexpr int RP = (P2W * G::P2H + S1_THREADS - 1) / S1_THREADS;   // positions per thread
    float sv[RP][N::M2];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        const int r = tid + k * S1_THREADS;
        if (r >= nq) break;
        const int qy = (int)(((float)r + 0.5f) * rqw), qx = r - qy * qw;
        float v[N::M2];
#pragma unroll
        for (int i = 0; i < N::M1; ++i) {
            const float* src = s_s1 + i * P1N + (2 * qy) * P1W + 2 * qx;
            float win[4][4];
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) win[r][c] = src[r * P1W + c];
            if constexpr (H) {
                // pairs (c, c+1) of each window row, and (c3, 0) for the last tap of the
                // right-hand quads; per map and row the tap pairs (k0, k1), (k2, 0)
                h2_t hp[4][3], hz[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) hp[r][c] = pk16(win[r][c], win[r][c + 1]);
                    hz[r] = pk16(win[r][3], 0.f);
                }
                const uint32_t* kh = reinterpret_cast<const uint32_t*>(W + N::SIZE) + Stage1H::oL2;
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int m = 2 * i + half;
                    float a[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int dx = q & 1, dy = q >> 1;
                        float acc = 0.f;
#pragma unroll
                        for (int r = 0; r < 3; ++r) {
                            acc = dot2(hp[dy + r][dx], as_h2(kh[m * 6 + 2 * r]), acc);
                            acc = dot2(dx ? hz[dy + r] : hp[dy + r][2], as_h2(kh[m * 6 + 2 * r + 1]), acc);
                        }
                        a[q] = lrelu_bn_fma(acc, b2, m, N::M2);
                    }
                    v[m] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
                }
            } else {
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const int m = 2 * i + half;
                    const float* kk = k2 + m * 9;
                    float a[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int dx = q & 1, dy = q >> 1;
                        float acc = win[dy][dx] * kk[0];
#pragma unroll
                        for (int j = 1; j < 9; ++j) acc = acc + win[dy + j / 3][dx + j % 3] * kk[j];
                        a[q] = lrelu_bn(acc, b2, m, N::M2);
                    }
                    v[m] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
                }
            }
        }
        sv[k][0] = v[0] + v[1];
#pragma unroll
        for (int i = 1; i < N::M2 - 1; ++i) sv[k][i] = v[i - 1] + v[i] + v[i + 1];
        sv[k][N::M2 - 1] = v[N::M2 - 2] + v[N::M2 - 1];
    }
    __syncthreads();
    PHASE(PH_STAGE1, 2);
    // the L2 sums over the (consumed) L1 sums, with rows P2S floats apart: an odd multiple
    // of 8, so that the rows of an 8 x 8 block of responses (below) start in distinct
    // groups of 8 LDS banks and the L3 window reads of a wave are free of bank conflicts;
    // the tail weights after them
    constexpr int P2S = (P2W + 7) / 16 * 16 + 8, P2NS = P2S * G::P2H;
    float* s_s2 = s_s1;
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        const int r = tid + k * S1_THREADS;
        if (r >= nq) break;
        const int qy = (int)(((float)r + 0.5f) * rqw), qx = r - qy * qw;
        const int ps = qy * P2S + qx;
#pragma unroll
        for (int m = 0; m < N::M2; ++m) s_s2[m * P2NS + ps] = sv[k][m];
    }
    float* s_wt = s_s1 + N::M2 * P2NS;
    static_assert(P2S >= P2W && N::M2 * P2NS + WT + Stage1H::SIZE <= N::M1 * P1N,
                  "L2 sums and tail weights fit in the L1-sum space");
#pragma unroll
    for (int k = 0; k < WPER; ++k)
        if (tid + k * S1_THREADS < WT) s_wt[tid + k * S1_THREADS] = wt[k];
    // FP16 path: the L3 and hidden-layer tap pairs follow the FP32 tail weights in LDS
    uint32_t* s_wh = reinterpret_cast<uint32_t*>(s_wt + WT);
    if constexpr (H) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(W + N::SIZE) + Stage1H::oL3;
        for (int k = tid; k < Stage1H::SIZE - Stage1H::oL3; k += S1_THREADS) s_wh[k] = src[k];
    }
    __syncthreads();
    PHASE(PH_STAGE1, 3);

    // L3 (conv_5x4: row sums accumulated, cnnpp_cplusplus.cpp:158-187) + lrelu_bn,
    // then the hidden and output layers in registers, one response value per thread. In a
    // whole tile each wave takes an 8 x 8 block of responses; in an edge tile the responses
    // inside the map are numbered densely and the threads past them stop here (no barrier
    // follows)
    static_assert(TX % 8 == 0 && TY % 8 == 0 && TX * TY % 64 == 0, "8 x 8 blocks of responses");
    int oy, ox;
    if (vw == TX && vh == TY) {
        if (tid >= TX * TY) return;
        const int b = tid >> 6, l = tid & 63;
        oy = (b / (TX / 8)) * 8 + (l >> 3);
        ox = (b % (TX / 8)) * 8 + (l & 7);
    } else {
        if (tid >= vw * vh) return;
        oy = (int)(((float)tid + 0.5f) * rvw);
        ox = tid - oy * vw;
    }
    const float* Wt = s_wt - N::oK3;   // indexed with the offsets into W
    const float* k3 = Wt + N::oK3;
    const float* b3 = Wt + N::oB3;
    float a3[N::M3];
#pragma unroll
    for (int i = 0; i < N::M2; ++i) {
        const float* src = s_s2 + i * P2NS + oy * P2S + ox;
        float win[N::K3H][N::K3W];
#pragma unroll
        for (int r = 0; r < N::K3H; ++r)
#pragma unroll
            for (int c = 0; c < N::K3W; ++c) win[r][c] = src[r * P2S + c];
        if constexpr (H) {
            static_assert(N::K3W == 4, "two tap pairs per L3 row");
            h2_t hp[N::K3H][2];
#pragma unroll
            for (int r = 0; r < N::K3H; ++r) {
                hp[r][0] = pk16(win[r][0], win[r][1]);
                hp[r][1] = pk16(win[r][2], win[r][3]);
            }
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int m = 2 * i + half;
                const uint32_t* kh = s_wh + (Stage1H::oL3 - Stage1H::oL3) + m * 2 * N::K3H;
                float d = 0.f;
#pragma unroll
                for (int r = 0; r < N::K3H; ++r) {
                    d = dot2(hp[r][0], as_h2(kh[2 * r]), d);
                    d = dot2(hp[r][1], as_h2(kh[2 * r + 1]), d);
                }
                a3[m] = lrelu_bn_fma(d, b3, m, N::M3);
            }
            continue;
        }
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int m = 2 * i + half;
            const float* k = k3 + m * N::K3W * N::K3H;
            float d = 0.f;
#pragma unroll
            for (int r = 0; r < N::K3H; ++r) {
                float rs = win[r][0] * k[r * N::K3W];
#pragma unroll
                for (int c = 1; c < N::K3W; ++c) rs = rs + win[r][c] * k[r * N::K3W + c];
                d = (r == 0) ? rs : d + rs;
            }
            a3[m] = lrelu_bn(d, b3, m, N::M3);
        }
    }
    // hidden layer: mulCN_add_tanhW (cnnpp_cplusplus.cpp:646-672, loop cnn_simd_cntk.cpp:874-883)
    float hv[N::HL];
    if constexpr (H) {
        static_assert(N::HLN == 4, "two input pairs per hidden unit");
        h2_t ha[N::M3 - 1];
#pragma unroll
        for (int j = 0; j < N::M3 - 1; ++j) ha[j] = pk16(a3[j], a3[j + 1]);
        const uint32_t* kh = s_wh + (Stage1H::oHL - Stage1H::oL3);
#pragma unroll
        for (int gi = 0; gi < N::HLG; ++gi)
#pragma unroll
            for (int tt = 0; tt < N::HLS; ++tt) {
                const int u = N::HLS * gi + tt;
                float c = dot2(ha[gi], as_h2(kh[2 * u]), Wt[N::oHB + u]);
                c = dot2(ha[gi + 2], as_h2(kh[2 * u + 1]), c);
                c = rtanh(c * Wt[N::oHT + gi]);
                hv[u] = __builtin_fmaf(Wt[N::oHBW + tt], __builtin_fmaf(c, 0.5f, 0.5f), Wt[N::oHBB + tt]);
            }
    } else {
#pragma unroll
    for (int gi = 0; gi < N::HLG; ++gi)
#pragma unroll
        for (int tt = 0; tt < N::HLS; ++tt) {
            const int u = N::HLS * gi + tt;
            float c = Wt[N::oHB + u];
#pragma unroll
            for (int j = 0; j < N::HLN; ++j) c = c + a3[gi + j] * Wt[N::oHW + u * N::HLN + j];
            c = c * Wt[N::oHT + gi];
            c = rtanh(c);
            c = (c + 1.f) * 0.5f;
            hv[u] = Wt[N::oHBW + tt] * c + Wt[N::oHBB + tt];
        }
    }
    // output layer: mulC / mulC1_add / tanhW (cnnpp_cplusplus.cpp:952-984, 673-690)
    float o = 0.f;
#pragma unroll
    for (int gi = 0; gi < N::HLG; ++gi)
#pragma unroll
        for (int tt = 0; tt < N::HLS; ++tt) {
            const float wv = Wt[N::oOW + tt * N::HLG + gi];
            if (gi == 0 && tt == 0) o = hv[0] * wv;
            else if (H) o = __builtin_fmaf(hv[N::HLS * gi + tt], wv, o);
            else o = hv[N::HLS * gi + tt] * wv + o;
        }
    float c = o + Wt[N::oOB];
    c = c * Wt[N::oOB + 1];
    c = rtanh(c);
    const float af = Wt[N::oOB + 2];
    const float out = (af == 0.f) ? (c + 1.f) * 0.5f : af * c;
    const int e = (oy0 + oy) * L.rw + ox0 + ox;
    resp[L.resp_off + e] = out;
    PHASE(PH_STAGE1, 4);
    PHASE_FLUSH(PH_STAGE1);
    // candidate windows (score > t1) go to the frame's list in no particular order; the
    // host replays them in the reference's order (level, then raster). Without meta (the
    // batch path) the response maps alone are written and the ordered compaction
    // (k_cmp_*) collects the candidates.
    if (meta && out > t1) {
        Candidate cd;
        cd.level = li;
        cd.x = ox0 + ox;
        cd.y = oy0 + oy;
        cd.score = out;
        const int idx = atomicAdd(&meta[0], 1);
        cands[idx] = cd;
        // hcands (optional): the same record in mapped host memory, for the first hcap
        if (hcands && idx < hcap) hcands[idx] = cd;
    }
}

// ---------------------------------------------------------------------------------
// Ordered compaction of the candidate windows (cnn_detector_v3.cpp:1918-1947: a window
// whose response exceeds t1 becomes a candidate; levels are visited in check order and
// each map in raster order). On the batch path the response maps of all frames' levels lie
// back to back in exactly that order, so the candidate list is the stream compaction of
// the flat response array: k_cmp_count counts the windows per block of CMP_BLOCK
// positions, k_cmp_scan turns the counts into each block's first output index, and
// k_cmp_emit writes the windows in position order. Position i = block * CMP_BLOCK +
// k * 256 + tid (k = 0..15): the loads are coalesced, and inside a block the output order
// is k, then wave, then lane, which per-wave ballots and their population counts give.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_cmp_count(const float* __restrict__ resp, int n, float t1,
                                                   int* __restrict__ counts)
{
    __shared__ int s_c[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int base = blockIdx.x * CMP_BLOCK;
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int i = base + k * 256 + tid;
        v[k] = i < n ? resp[i] : -INFINITY;
    }
    int c = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) c += __popcll(__ballot(v[k] > t1));
    if (lane == 0) s_c[wave] = c;
    __syncthreads();
    if (tid == 0) counts[blockIdx.x] = s_c[0] + s_c[1] + s_c[2] + s_c[3];
}

// One workgroup: exclusive prefix sums of the nb block counts into offs; meta[0] (and
// hmeta[0]) = the candidate total, meta[1] = 0 (the check kernel's work queue).
__global__ void __launch_bounds__(CMP_SCAN_THREADS) k_cmp_scan(const int* __restrict__ counts, int nb,
                                                               int* __restrict__ offs, int* __restrict__ meta,
                                                               int* __restrict__ hmeta)
{
    constexpr int NW = CMP_SCAN_THREADS / 64;
    __shared__ int s_w[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = (nb + CMP_SCAN_THREADS - 1) / CMP_SCAN_THREADS;
    const int b0 = min(tid * per, nb), b1 = min(b0 + per, nb);
    int sum = 0;
    for (int b = b0; b < b1; ++b) sum += counts[b];
    int v = sum;                                     // inclusive scan inside the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int a = __shfl_up(v, off, 64);
        if (lane >= off) v += a;
    }
    if (lane == 63) s_w[wave] = v;
    __syncthreads();
    int before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        before += k < wave ? s_w[k] : 0;
        total += s_w[k];
    }
    int run = before + v - sum;
    for (int b = b0; b < b1; ++b) {
        offs[b] = run;
        run += counts[b];
    }
    if (tid == 0) {
        meta[0] = total;
        meta[1] = 0;
        if (hmeta) hmeta[0] = total;
    }
}

__global__ void __launch_bounds__(256) k_cmp_emit(const float* __restrict__ resp, int n, float t1,
                                                  const int* __restrict__ offs, const LevelDesc* __restrict__ levels,
                                                  int nlevels, Candidate* __restrict__ cands, int cap,
                                                  Candidate* __restrict__ hcands, int hcap)
{
    __shared__ int s_c[16][4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int base = blockIdx.x * CMP_BLOCK;
    float v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const int i = base + k * 256 + tid;
        v[k] = i < n ? resp[i] : -INFINITY;
    }
    unsigned long long bal[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        bal[k] = __ballot(v[k] > t1);
        if (lane == 0) s_c[k][wave] = __popcll(bal[k]);
    }
    __syncthreads();
    int run = offs[blockIdx.x];
    const unsigned long long below = (1ull << lane) - 1;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        int at = run;
#pragma unroll
        for (int w = 0; w < 4; ++w) at += w < wave ? s_c[k][w] : 0;
        if (v[k] > t1) {
            const int idx = at + __popcll(bal[k] & below);
            const int i = base + k * 256 + tid;
            // the level holding position i: the last whose maps start at or before it
            int lo = 0, hi = nlevels - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (levels[mid].resp_off <= i) lo = mid; else hi = mid - 1;
            }
            const LevelDesc& L = levels[lo];
            const int e = i - L.resp_off;
            const int y = e / L.rw;
            Candidate cd;
            cd.level = lo;
            cd.x = e - y * L.rw;
            cd.y = y;
            cd.score = v[k];
            if (idx < cap) cands[idx] = cd;
            if (hcands && idx < hcap) hcands[idx] = cd;
        }
        run += s_c[k][0] + s_c[k][1] + s_c[k][2] + s_c[k][3];
    }
}

// ---------------------------------------------------------------------------------
// Stages 2/3: one workgroup evaluates one candidate's whole check
// (CPUCheckDetect without the order-dependent drop logic, which the host replays).
// ---------------------------------------------------------------------------------
namespace chk {
constexpr int CW = 49, CH = 51;          // crop after the NN resize (ext pattern + x offset)
constexpr int IW = 47, IH = 51;          // stage-2/3 input
constexpr int C1W = IW - 3, C1H = IH - 3;          // 44 x 48
constexpr int P1W = C1W / 2, P1H = C1H / 2;        // 22 x 24
constexpr int P1N = P1W * P1H;
constexpr int P2W = (P1W - 2) / 2, P2H = (P1H - 2) / 2;   // 10 x 11
constexpr int P2N = P2W * P2H;
constexpr int RW = P2W - 7 + 1, RH = P2H - 8 + 1;  // 4 x 4
constexpr int RN = RW * RH;
constexpr int HLMAX = 84;
constexpr int WL = Stage2::SIZE > Stage3::SIZE ? Stage2::SIZE : Stage3::SIZE;
}  // namespace chk

// LDS of one (candidate, crop variant). Regions are reused as their contents die:
//   in  the network input; after L1 the L2 maps; after the L2 sums the L3 maps
//       followed by the output
//   s1  the crop's histogram and LUT while the crop is prepared; then the L1 sums;
//       after L2 the L2 sums; after L3 the hidden layer
//   w   the weights of the running network (stage 2 or 3)
struct ChkSmem {
    alignas(16) float in[chk::IH * chk::IW];
    alignas(16) float s1[6 * chk::P1N];
    float w[chk::WL];
    float w3[Stage3::SIZE];  // k_check: the stage-3 weights beside the stage-2 ones in w
    __device__ float* p2() { return in; }
    __device__ float* a3() { return in; }
    __device__ float* out() { return in + 24 * chk::RN; }
    __device__ float* s2() { return s1; }
    __device__ float* h() { return s1; }
    __device__ int* hist() { return reinterpret_cast<int*>(s1); }
    __device__ int* lut() { return reinterpret_cast<int*>(s1) + 256; }
};
static_assert(12 * chk::P2N <= chk::IH * chk::IW && 24 * chk::RN + chk::RN <= chk::IH * chk::IW &&
                  12 * chk::P2N <= 6 * chk::P1N && chk::HLMAX * chk::RN <= 6 * chk::P1N && 512 <= 6 * chk::P1N,
              "region reuse fits");

// ---------------------------------------------------------------------------------
// Reduced-precision stages 2/3 on the matrix cores. L1 and L2 of each check window are
// GEMMs on v_mfma_f32_16x16x32_{bf16,fp8_fp8} (FP32 accumulation):
//   L1: rows = the 528 pooled L1 positions (33 tiles of 16), K = the 25 samples of the
//       5x5 window under a 2x2 pooling quad (padded to 32), columns = 6 maps x 4 quad
//       offsets (two tiles of 16); the weights matrix holds each map's 4x4 kernel at its
//       quad offset inside the 5x5 window, so one product row is the 24 conv outputs
//       that pool into one position;
//   L2: rows = the 110 pooled L2 positions (7 tiles), and per pair of L1 sums (a, b)
//       K = 16 samples of sum a's 4x4 window + 16 of sum b's, columns = sum a's two maps
//       x 4 quads then sum b's (block-diagonal weights), three products per row tile.
// The 4x4 quad columns of a map lie in four adjacent lanes of the accumulator, so the
// 2x2 max-pool is two lane swaps. The weights matrices are expanded per lane by the host
// (MfmaW) and rounded to the operand type in the kernel. FP8 (e4m3) operands are scaled
// by powers of two: the weights per layer (host), the L2 input per window (its largest
// magnitude), and the products are scaled back exactly. Activations and weights enter as
// two terms of the operand type each (value and rounding error), three products per
// pair. L3, the hidden and output layers run in FP32 as on the exact path.
// ---------------------------------------------------------------------------------
enum { RP_EXACT = 0, RP_BF16 = 1, RP_FP8 = 2 };
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int RP>
struct MOp;
template <>
struct MOp<RP_BF16> {
    typedef bf16x8_t T;
    static __device__ __forceinline__ T pack(const float (&v)[8])
    {
        T r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (__bf16)v[i];
        return r;
    }
    // the BF16 values of an operand, back in FP32 (exact)
    static __device__ __forceinline__ void unpack(T t, float (&v)[8])
    {
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = (float)t[i];
    }
    static __device__ __forceinline__ f32x4_t mma(T a, T b, f32x4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MOp<RP_FP8> {
    typedef long T;
    static __device__ __forceinline__ T pack(const float (&v)[8])
    {
        int lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], 0, false);
        lo = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], lo, true);
        int hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[4], v[5], 0, false);
        hi = __builtin_amdgcn_cvt_pk_fp8_f32(v[6], v[7], hi, true);
        return (long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
    // the FP8 values of an operand, back in FP32
    static __device__ __forceinline__ void unpack(T t, float (&v)[8])
    {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const int lo = (int)((unsigned long long)t & 0xffffffffull), hi = (int)((unsigned long long)t >> 32);
        const f2 a = __builtin_amdgcn_cvt_pk_f32_fp8(lo, false), b = __builtin_amdgcn_cvt_pk_f32_fp8(lo, true);
        const f2 c = __builtin_amdgcn_cvt_pk_f32_fp8(hi, false), d = __builtin_amdgcn_cvt_pk_f32_fp8(hi, true);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y; v[6] = d.x; v[7] = d.y;
    }
    static __device__ __forceinline__ f32x4_t mma(T a, T b, f32x4_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(a, b, c, 0, 0, 0);
    }
};

// Operands enter as two terms of the operand type each, the rounded value and its
// rounding error (BF16: about 16 significant bits together, FP8 e4m3 about 8), and each
// product pair as three MFMAs into one FP32 accumulator: hi x hi, lo x hi, hi x lo (lo x lo
// lies below the accumulator's rounding).
template <int RP>
struct WOp {
    typename MOp<RP>::T hi, lo;
};
template <int RP>
__device__ __forceinline__ typename MOp<RP>::T rest_of(const float (&w)[8], typename MOp<RP>::T hi)
{
    using Op = MOp<RP>;
    float v[8];
    Op::unpack(hi, v);
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = w[j] - v[j];
    return Op::pack(v);
}
template <int RP>
__device__ __forceinline__ WOp<RP> pack_w(const float (&w)[8])
{
    WOp<RP> r;
    r.hi = MOp<RP>::pack(w);
    r.lo = rest_of<RP>(w, r.hi);
    return r;
}

// The activation operand (eight values of one lane) times the weights.
template <int RP>
__device__ __forceinline__ f32x4_t mma_act(const float (&a)[8], const WOp<RP>& b, f32x4_t acc)
{
    using Op = MOp<RP>;
    const typename Op::T ah = Op::pack(a);
    acc = Op::mma(ah, b.hi, acc);
    acc = Op::mma(rest_of<RP>(a, ah), b.hi, acc);
    acc = Op::mma(ah, b.lo, acc);
    return acc;
}

// 2^e as a float (|e| < 126)
__device__ __forceinline__ float pow2f(int e) { return __int_as_float((e + 127) << 23); }

template <class N, int RP>
__device__ __forceinline__ void check_l1l2_mfma(const float* __restrict__ in, ChkSmem& s,
                                                const float* Wl, const int tid,
                                                const float* __restrict__ mw, int* s_red)
{
    using namespace chk;
    using Op = MOp<RP>;
    const int lane = tid & 63, wave = tid >> 6;
    const int row = lane & 15, kb = lane >> 4;
    // this lane's weight operands: L1 (two column tiles), L2 (three sum pairs)
    float wv[8];
    WOp<RP> b1[2], b2[3];
    const float sw1 = RP == RP_FP8 ? mw[MfmaW::oScale] : 1.f, sw2 = RP == RP_FP8 ? mw[MfmaW::oScale + 1] : 1.f;
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
#pragma unroll
        for (int j = 0; j < 8; ++j) wv[j] = mw[MfmaW::oL1 + (nt * 64 + lane) * 8 + j] * sw1;
        b1[nt] = pack_w<RP>(wv);
    }
#pragma unroll
    for (int jp = 0; jp < 3; ++jp) {
#pragma unroll
        for (int j = 0; j < 8; ++j) wv[j] = mw[MfmaW::oL2 + (jp * 64 + lane) * 8 + j] * sw2;
        b2[jp] = pack_w<RP>(wv);
    }
    // L1: 33 row tiles of 16 pooled positions over the 4 waves
    const float inv1 = 1.f / sw1;
    for (int mt = wave; mt < P1N / 16; mt += CHECK_THREADS / 64) {
        const int P = mt * 16 + row;
        const int py = P / P1W, px = P - py * P1W;
        float a[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int t = kb * 8 + j;                        // sample t of the 5x5 window
            a[j] = t < 25 ? in[(2 * py + t / 5) * IW + 2 * px + t % 5] : 0.f;
        }
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma_act<RP>(a, b1[nt], acc);
            const int col = row + 16 * nt, m = col >> 2;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = m < N::M1 ? lrelu_bn_fma(acc[e] * inv1, Wl + N::oB1, m, N::M1) : 0.f;
                v = fmaxf(v, __shfl_xor(v, 1, 64));
                v = fmaxf(v, __shfl_xor(v, 2, 64));
                if ((col & 3) == 0 && m < N::M1) s.s1[m * P1N + mt * 16 + kb * 4 + e] = v;
            }
        }
    }
    __syncthreads();
    // neighbour sums in place, and the largest magnitude (the FP8 scale of L2's input)
    float mx = 0.f;
    for (int p = tid; p < P1N; p += CHECK_THREADS) {
        float v[N::M1];
#pragma unroll
        for (int m = 0; m < N::M1; ++m) v[m] = s.s1[m * P1N + p];
        float o[N::M1];
        o[0] = v[0] + v[1];
#pragma unroll
        for (int i = 1; i < N::M1 - 1; ++i) o[i] = v[i - 1] + v[i] + v[i + 1];
        o[N::M1 - 1] = v[N::M1 - 2] + v[N::M1 - 1];
#pragma unroll
        for (int m = 0; m < N::M1; ++m) {
            s.s1[m * P1N + p] = o[m];
            mx = fmaxf(mx, fabsf(o[m]));
        }
    }
    float sx = 1.f;
    if (RP == RP_FP8) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
        if (lane == 0) s_red[wave] = __float_as_int(mx);
    }
    __syncthreads();
    if (RP == RP_FP8) {
        float m4 = 0.f;
#pragma unroll
        for (int w = 0; w < CHECK_THREADS / 64; ++w) m4 = fmaxf(m4, __int_as_float(s_red[w]));
        // the power of two that brings the largest input into [128, 256)
        int e = 0;
        if (m4 > 0.f) e = 8 - (((__float_as_int(m4) >> 23) & 255) - 126);
        sx = pow2f(min(max(e, -100), 100));
    }
    // L2: 7 row tiles of 16 pooled positions (the last one partly), three sum pairs each
    const float inv2 = 1.f / (sx * sw2);
    for (int mt = wave; mt < (P2N + 15) / 16; mt += CHECK_THREADS / 64) {
        const int Q = min(mt * 16 + row, P2N - 1);
        const int qy = Q / P2W, qx = Q - qy * P2W;
#pragma unroll
        for (int jp = 0; jp < 3; ++jp) {
            // k block kb: samples 8*kb .. 8*kb+7 of sum 2*jp (kb < 2) or 2*jp + 1
            const float* src = s.s1 + (2 * jp + (kb >> 1)) * P1N;
            float a[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int t = (kb & 1) * 8 + j;             // sample t of the 4x4 window
                a[j] = src[(2 * qy + (t >> 2)) * P1W + 2 * qx + (t & 3)] * sx;
            }
            f32x4_t acc = {0.f, 0.f, 0.f, 0.f};
            acc = mma_act<RP>(a, b2[jp], acc);
            const int m = 4 * jp + (row >> 2);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float v = lrelu_bn_fma(acc[e] * inv2, Wl + N::oB2, m, N::M2);
                v = fmaxf(v, __shfl_xor(v, 1, 64));
                v = fmaxf(v, __shfl_xor(v, 2, 64));
                const int q = mt * 16 + kb * 4 + e;
                if ((row & 3) == 0 && q < P2N) s.p2()[m * P2N + q] = v;
            }
        }
    }
    __syncthreads();
}

template <class N, int RP = RP_EXACT>
__device__ __forceinline__ void check_forward(const float* __restrict__ in, ChkSmem& s,
                                              const float* Wl, const int tid,
                                              const float* __restrict__ mw = nullptr, int* s_red = nullptr)
{
    using namespace chk;
    constexpr int PB = std::is_same<N, Stage3>::value ? 9 : 3;   // first phase slot
    // Wl: the network's weights, staged in LDS; read the same for every thread in L1,
    // per thread from L2 on
    const unsigned ut = (unsigned)tid;
    if constexpr (RP != RP_EXACT) {
        check_l1l2_mfma<N, RP>(in, s, Wl, tid, mw, s_red);
    } else {
    // L1 + sums: the first FULL positions one per lane (whole waves), the REM others by
    // one wave with a lane per 2x2-pool quad. Index arithmetic is unsigned throughout.
    constexpr unsigned FULL = P1N / 64 * 64, REM = P1N - FULL;
    static_assert(REM == 0 || REM * 4 == 64, "the remainder fills one wave, four lanes a position");
    for (unsigned p = ut; p < FULL; p += CHECK_THREADS) {
        const unsigned py = p / P1W, px = p - py * P1W;
        float win[5][5];
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int c = 0; c < 5; ++c) win[r][c] = in[(2 * py + r) * IW + 2 * px + c];
        float pm[N::M1];
#pragma unroll
        for (int m = 0; m < N::M1; ++m) {
            const float* k = Wl + N::oK1 + m * 16;
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int dx = q & 1, dy = q >> 1;
                float acc = win[dy][dx] * k[0];
#pragma unroll
                for (int j = 1; j < 16; ++j) acc = acc + win[dy + (j >> 2)][dx + (j & 3)] * k[j];
                a[q] = lrelu_bn(acc, Wl + N::oB1, m, N::M1);
            }
            pm[m] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
        }
        s.s1[p] = pm[0] + pm[1];
#pragma unroll
        for (int i = 1; i < N::M1 - 1; ++i) s.s1[i * P1N + p] = pm[i - 1] + pm[i] + pm[i + 1];
        s.s1[(N::M1 - 1) * P1N + p] = pm[N::M1 - 2] + pm[N::M1 - 1];
    }
    if (REM > 0 && ut < REM * 4) {
        const unsigned p = FULL + (ut >> 2), q = ut & 3u;
        const unsigned py = p / P1W, px = p - py * P1W;
        const unsigned dx = q & 1u, dy = q >> 1;
        float win[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) win[r][c] = in[(2 * py + dy + r) * IW + 2 * px + dx + c];
        float pm[N::M1];
#pragma unroll
        for (int m = 0; m < N::M1; ++m) {
            const float* k = Wl + N::oK1 + m * 16;
            float acc = win[0][0] * k[0];
#pragma unroll
            for (int j = 1; j < 16; ++j) acc = acc + win[j >> 2][j & 3] * k[j];
            const float a = lrelu_bn(acc, Wl + N::oB1, m, N::M1);
            // fmaxf(fmaxf(a0, a1), fmaxf(a2, a3)) across the position's four lanes
            const float m01 = fmaxf(a, __shfl_xor(a, 1, 64));
            pm[m] = fmaxf(m01, __shfl_xor(m01, 2, 64));
        }
        if (q == 0) {
            s.s1[p] = pm[0] + pm[1];
#pragma unroll
            for (int i = 1; i < N::M1 - 1; ++i) s.s1[i * P1N + p] = pm[i - 1] + pm[i] + pm[i + 1];
            s.s1[(N::M1 - 1) * P1N + p] = pm[N::M1 - 2] + pm[N::M1 - 1];
        }
    }
    __syncthreads();
    PHASE(PH_CHECK, PB + 0);
    // L2
    for (unsigned it = ut; it < N::M1 * P2N; it += CHECK_THREADS) {
        const unsigned i = it / P2N, p = it - i * P2N;
        const unsigned qy = p / P2W, qx = p - qy * P2W;
        const float* src = s.s1 + i * P1N + (2 * qy) * P1W + 2 * qx;
        float win[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) win[r][c] = src[r * P1W + c];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int m = 2 * i + half;
            const float* k = Wl + N::oK2 + m * 9;
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int dx = q & 1, dy = q >> 1;
                float acc = win[dy][dx] * k[0];
#pragma unroll
                for (int j = 1; j < 9; ++j) acc = acc + win[dy + j / 3][dx + j % 3] * k[j];
                a[q] = lrelu_bn(acc, Wl + N::oB2, m, N::M2);
            }
            s.p2()[m * P2N + p] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
        }
    }
    __syncthreads();
    PHASE(PH_CHECK, PB + 1);
    }
    float* s2 = s.s2();
    for (unsigned p = ut; p < P2N; p += CHECK_THREADS) {
        float v[N::M2];
#pragma unroll
        for (int m = 0; m < N::M2; ++m) v[m] = s.p2()[m * P2N + p];
        s2[p] = v[0] + v[1];
#pragma unroll
        for (int i = 1; i < N::M2 - 1; ++i) s2[i * P2N + p] = v[i - 1] + v[i] + v[i + 1];
        s2[(N::M2 - 1) * P2N + p] = v[N::M2 - 2] + v[N::M2 - 1];
    }
    __syncthreads();
    PHASE(PH_CHECK, PB + 2);
    // L3: conv_8x7 single chain (cnnpp_cplusplus.cpp:292-321) + lrelu_bn
    for (unsigned it = ut; it < N::M2 * RN; it += CHECK_THREADS) {
        const unsigned i = it / RN, p = it - i * RN;
        const unsigned y = p / RW, x = p - y * RW;
        const float* src = s2 + i * P2N + y * P2W + x;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int m = 2 * i + half;
            const float* k = Wl + N::oK3 + m * N::K3W * N::K3H;
            float acc = src[0] * k[0];
#pragma unroll
            for (int j = 1; j < N::K3W * N::K3H; ++j)
                acc = acc + src[(j / N::K3W) * P2W + (j % N::K3W)] * k[j];
            s.a3()[m * RN + p] = lrelu_bn(acc, Wl + N::oB3, m, N::M3);
        }
    }
    __syncthreads();
    PHASE(PH_CHECK, PB + 3);
    // hidden layer, each unit's value already multiplied by its output weight: unit
    // u = HLS * gi + tt of response p goes to s.h()[p * HLP + u], so that the output
    // layer's serial sum over u reads its terms with 16-byte loads issued together
    constexpr int HLP = (N::HL + 3) / 4 * 4;
    static_assert(HLP * RN <= 6 * P1N, "the output-layer terms fit the s1 region");
    for (unsigned it = ut; it < N::HL * RN; it += CHECK_THREADS) {
        const unsigned u = it / RN, p = it - u * RN;
        const unsigned gi = u / N::HLS, tt = u - gi * N::HLS;
        float c = Wl[N::oHB + u];
#pragma unroll
        for (int j = 0; j < N::HLN; ++j) c = c + s.a3()[(gi + j) * RN + p] * Wl[N::oHW + u * N::HLN + j];
        c = c * Wl[N::oHT + gi];
        c = rtanh(c);
        c = (c + 1.f) * 0.5f;
        const float hv = Wl[N::oHBW + tt] * c + Wl[N::oHBB + tt];
        s.h()[p * HLP + u] = hv * Wl[N::oOW + tt * N::HLG + gi];
    }
    __syncthreads();
    PHASE(PH_CHECK, PB + 4);
    // output layer: the products in unit order, summed serially (the first term alone,
    // then term + sum: IEEE addition commutes, so the order of the sum is what counts)
    if (tid < RN) {
        float t[HLP];
        const float4* src = reinterpret_cast<const float4*>(s.h() + tid * HLP);
#pragma unroll
        for (int q = 0; q < HLP / 4; ++q) {
            const float4 v = src[q];
            t[4 * q] = v.x; t[4 * q + 1] = v.y; t[4 * q + 2] = v.z; t[4 * q + 3] = v.w;
        }
        float o = t[0];
#pragma unroll
        for (int u = 1; u < N::HL; ++u) o = t[u] + o;
        float c = o + Wl[N::oOB];
        c = c * Wl[N::oOB + 1];
        c = rtanh(c);
        const float af = Wl[N::oOB + 2];
        s.out()[tid] = (af == 0.f) ? (c + 1.f) * 0.5f : af * c;
    }
    __syncthreads();
    PHASE(PH_CHECK, PB + 5);
}

// Copies n floats of weights into LDS with all of a thread's loads in flight together
// (load_weights issues them, store_weights writes them once they have arrived).
template <int n>
struct WeightRegs {
    static constexpr int PER = (n + CHECK_THREADS - 1) / CHECK_THREADS;
    float v[PER];
};

template <int n>
__device__ __forceinline__ void load_weights(WeightRegs<n>& r, const float* __restrict__ src, const int tid)
{
#pragma unroll
    for (int k = 0; k < WeightRegs<n>::PER; ++k) r.v[k] = src[min(tid + k * CHECK_THREADS, n - 1)];
}

template <int n>
__device__ __forceinline__ void store_weights(float* __restrict__ dst, const WeightRegs<n>& r, const int tid)
{
#pragma unroll
    for (int k = 0; k < WeightRegs<n>::PER; ++k)
        if (tid + k * CHECK_THREADS < n) dst[tid + k * CHECK_THREADS] = r.v[k];
}

template <int n>
__device__ __forceinline__ void stage_weights(float* __restrict__ dst, const float* __restrict__ src, const int tid)
{
    WeightRegs<n> r;
    load_weights<n>(r, src, tid);
    store_weights<n>(dst, r, tid);
}

// Crop geometry of one candidate (cnn_detector_v3.cpp:1491-1556): the window around the
// stage-1 position, shrunk by 15% for the retry (mod 1), clamped to the image.
struct Crop { int rx, ry, rcols, rrows; };

__device__ __forceinline__ Crop crop_geometry(int px, int py, float scale, int mod, int gw, int gh)
{
    const float inv = 1.f / scale;
    int x = (int)((float)(px - 6) * inv);      // (ext 11 + x offset 2) / 2
    int y = (int)((float)(py - 5) * inv);      // 11 / 2
    int width = (int)(36.f * inv);             // 23 + 11 + 2
    int height = (int)(38.f * inv);            // 27 + 11
    if (mod == 1) {
        // x += w*0.15/2; w -= w*0.15 (the reference's build folds w - w*0.15f into w * 0.85f)
        x = (int)((float)x + (float)width * 0.15f / 2.f);
        width = (int)((float)width * 0.85f);
        y = (int)((float)y + (float)height * 0.15f / 2.f);
        height = (int)((float)height * 0.85f);
    }
    Crop c;
    c.rx = min(max(x, 0), gw);
    c.ry = min(max(y, 0), gh);
    c.rcols = min(max(x + width, 0), gw) - c.rx;
    c.rrows = min(max(y + height, 0), gh) - c.ry;
    return c;
}

// Crop of one (candidate, crop variant): NN resize to 49x51 (a copy at that size,
// image_resize.cpp:328-392, 564-584), FloatToUChar (+0.5, truncate, image_proc.cpp:
// 357-368), then equalizeImage (image_proc.cpp:720-800: cumulative histogram and LUT
// int((cf - cf[0]) * (255 / (cf[255] - cf[0])) + 0.5), skipped if the histogram is
// flat). Each thread keeps its pixels (i = tid + k * CHECK_THREADS) in u[k]; when
// remap is set, s.lut maps them to the equalized values. With stage set (a workgroup's
// first item), the stage-2 and stage-3 weights W2 / W3 are loaded alongside the crop's
// gathers and stored to s.w / s.w3 before the first barrier, so that both latencies
// overlap.
namespace chk {
constexpr int PER = (CW * CH + CHECK_THREADS - 1) / CHECK_THREADS;
}

__device__ __forceinline__ void prepare_crop(const float* __restrict__ g, int gp, const Crop& cr, bool equalize,
                                             ChkSmem& s, int* s_w, uint32_t (&u)[chk::PER], bool& remap,
                                             const int tid, const float* __restrict__ W2,
                                             const float* __restrict__ W3, bool stage)
{
    using namespace chk;
    const bool copy = (cr.rcols == CW && cr.rrows == CH);
    const uint32_t sclx = ((uint32_t)cr.rcols << 12) / CW + 1;
    const uint32_t scly = ((uint32_t)cr.rrows << 12) / CH + 1;
    static_assert(CHECK_THREADS >= 256 && CHECK_THREADS % 64 == 0, "one histogram bin per thread");
    WeightRegs<Stage2::SIZE> r2;
    WeightRegs<Stage3::SIZE> r3;
    if (stage) {
        load_weights(r2, W2, tid);
        load_weights(r3, W3, tid);
    }
    // all of a thread's gathers are issued before any is used, so that their latencies
    // overlap; 32-bit offsets from the crop's corner keep one register per address
    const float* __restrict__ base = g + (size_t)cr.ry * gp + cr.rx;
    float v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = min(tid + k * CHECK_THREADS, CW * CH - 1);
        const int iy = i / CW, ix = i - iy * CW;
        uint32_t px = ix, py = iy;
        if (!copy) {
            px = min(((uint32_t)ix * sclx) >> 12, (uint32_t)(cr.rcols - 2));
            py = min(((uint32_t)iy * scly) >> 12, (uint32_t)(cr.rrows - 1));
        }
        v[k] = base[py * (uint32_t)gp + px];
        // its crop column and row in bits 8-15 and 16-31, for write_input
        u[k] = ((uint32_t)ix << 8) | ((uint32_t)iy << 16);
    }
    if (tid < 256) s.hist()[tid] = 0;
    if (stage) {
        store_weights(s.w, r2, tid);
        store_weights(s.w3, r3, tid);
    }
    __syncthreads();
    // the pixel value in bits 0-7
#pragma unroll
    for (int k = 0; k < PER; ++k) u[k] |= (uint32_t)(v[k] + 0.5f) & 255u;
    remap = false;
    if (equalize) {
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (tid + k * CHECK_THREADS < CW * CH) atomicAdd(&s.hist()[u[k] & 255u], 1);
        __syncthreads();
        // cumulative histogram and LUT by the first wave, four bins per lane
        if (tid < 64) {
            const int4 h = reinterpret_cast<const int4*>(s.hist())[tid];
            const int c0 = h.x, c1 = c0 + h.y, c2 = c1 + h.z, c3 = c2 + h.w;
            int v = c3;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int a = __shfl_up(v, off, 64);
                if (tid >= off) v += a;
            }
            const int b0 = v - c3;
            int4 cf = make_int4(b0 + c0, b0 + c1, b0 + c2, b0 + c3);
            const float bmin = (float)__shfl(cf.x, 0, 64), bmax = (float)__shfl(cf.w, 63, 64);
            if (bmin != bmax) {
                const float delt = 255.f / (bmax - bmin);
                cf.x = (int)(((float)cf.x - bmin) * delt + 0.5f);
                cf.y = (int)(((float)cf.y - bmin) * delt + 0.5f);
                cf.z = (int)(((float)cf.z - bmin) * delt + 0.5f);
                cf.w = (int)(((float)cf.w - bmin) * delt + 0.5f);
            }
            reinterpret_cast<int4*>(s.lut())[tid] = cf;
            if (tid == 0) s_w[0] = bmin != bmax;
        }
        __syncthreads();
        remap = s_w[0] != 0;
    }
}

// A network input from the thread's crop pixels: stage 2 reads UCharToFloat(offset 2)
// (image_proc.cpp:416-427, crop columns 2..48), stage 3 UCharToFloat_inv (mirror,
// :478-490, columns 46..0) or UCharToFloat(offset 0) (columns 0..46).
// With remap, the pixels are first replaced by their equalized values (in u, for the
// stage-3 input, as the LUT's LDS is reused by the network).
__device__ void write_input(ChkSmem& s, uint32_t (&u)[chk::PER], bool remap, bool stage3, bool reflection,
                            const int tid)
{
    using namespace chk;
    if (remap) {
        // every u[k] holds a crop pixel (the gather clamps the index), so all are mapped
#pragma unroll
        for (int k = 0; k < PER; ++k) u[k] = (u[k] & ~255u) | ((uint32_t)s.lut()[u[k] & 255u] & 255u);
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int i = tid + k * CHECK_THREADS;
        if (i < CW * CH) {
            const int ix = (u[k] >> 8) & 255u, iy = u[k] >> 16;
            const float f = (float)(u[k] & 255u);
            if (!stage3) {
                if (ix >= 2) s.in[iy * IW + ix - 2] = f;
            } else if (ix <= 46) {
                s.in[iy * IW + (reflection ? 46 - ix : ix)] = f;
            }
        }
    }
    __syncthreads();
}

// Stages 2 and 3 of one (candidate, crop variant), one pair per workgroup at a time:
// crop preparation, the stage-2 network on UCharToFloat(offset 2) (image_proc.cpp:416-427),
// then, unless the stage-2 outcome ends the check early (cnn_detector_v3.cpp:1704-1706),
// the stage-3 network on UCharToFloat_inv (mirror, image_proc.cpp:478-490) or offset 0.
// s2res / s3res[2*ci + mod] = {count above the threshold, max response bits};
// s3res {-1, -10.f} when stage 3 is skipped. hmeta / hs2 / hs3 (optional): the candidate
// count and, for the first hcap candidates, the same outcomes in mapped host memory, for
// decisions taken on the host.
template <int RP>
__global__ void __launch_bounds__(CHECK_THREADS, 4) k_check(const float* __restrict__ g0, int gp0, int gw0, int gh0,
                                                         const LevelDesc* __restrict__ levels,
                                                         const Candidate* __restrict__ cands,
                                                         const int* __restrict__ meta, CheckParams P,
                                                         const float* __restrict__ W2, const float* __restrict__ W3,
                                                         int2* __restrict__ s2res, int2* __restrict__ s3res,
                                                         int* __restrict__ hmeta, int2* __restrict__ hs2,
                                                         int2* __restrict__ hs3, int hcap, int cap,
                                                         int* __restrict__ queue, const float* __restrict__ MW2,
                                                         const float* __restrict__ MW3)
{
    using namespace chk;
    __shared__ ChkSmem s;
    __shared__ int s_w[CHECK_THREADS / 64];
    __shared__ int s_k1;
    __shared__ int s_item;
    // work items: (candidate, crop variant) pairs, both variants of a candidate next to
    // each other so that they are dispatched together. The candidates past cap (the
    // list's capacity) were not stored: the caller sees meta[0] > cap and redoes the work.
    const int nv = P.double_check ? 2 : 1;
    const int items = min(meta[0], cap) * nv;
    if (hmeta && blockIdx.x == 0 && threadIdx.x == 0) hmeta[0] = meta[0];
    // queue (optional): a counter, zero at launch, from which the workgroups take items
    // in turn (a persistent grid that balances the stage-3 tail); otherwise grid-stride:
    // round r covers items [r * G, r * G + n_r) with n_r = min(items - r * G, G), and in
    // each round the workgroups are renumbered (xcd_tile over n_r) so that every XCD takes
    // one contiguous run of items, whose candidates are neighbours with overlapping crops
    // in that XCD's L2. The workgroups past n_r have no item in this round or any later.
    int base = 0;
    auto stride_item = [&]() {
        const int nr = min(items - base, (int)gridDim.x);
        return (int)blockIdx.x < nr ? base + xcd_tile(blockIdx.x, nr) : items;
    };
    int item = stride_item();
    if (queue) {
        if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        item = s_item;
    }
    bool staged = false;
    PHASE_CLEAR();
    while (item < items) {
        const int ci = nv == 2 ? item >> 1 : item;
        const int mod = nv == 2 ? item & 1 : 0;
        // the thread index is made opaque per candidate: otherwise every per-thread index
        // of every phase is hoisted out of this loop and held in registers throughout,
        // which caps the occupancy
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const size_t slot = (size_t)ci * 2 + mod;
        PHASE(PH_CHECK, 0);
        const Candidate cd = cands[ci];
        const LevelDesc& Lv = levels[cd.level];
        const float scale = Lv.scale;
        const GrayRef gr = level_gray(g0, gp0, gw0, gh0, Lv);
        const float* __restrict__ g = gr.g;
        const int gp = gr.gp;
        const Crop cr = crop_geometry(cd.x * 4, cd.y * 4, scale, mod, gr.gw, gr.gh);
        // the stage-2 and stage-3 weights go to LDS with the workgroup's first item and
        // stay there for its later items
        uint32_t u[PER];
        bool remap;
        prepare_crop(g, gp, cr, P.equalize != 0, s, s_w, u, remap, tid, W2, W3, !staged);
        staged = true;
        PHASE(PH_CHECK, 1);
        write_input(s, u, remap, false, P.reflection != 0, tid);
        PHASE(PH_CHECK, 2);
        PHASE_SET(PH_CHECK, 20, __smid());
        PHASE_SET(PH_CHECK, 21, item);
        PHASE_SET(PH_CHECK, 22, 0);
        check_forward<Stage2, RP>(s.in, s, s.w, tid, MW2, s_w);
        if (tid == 0) {
            float d[RN];
#pragma unroll
            for (int q = 0; q < RN / 4; ++q) {
                const float4 v = reinterpret_cast<const float4*>(s.out())[q];
                d[4 * q] = v.x; d[4 * q + 1] = v.y; d[4 * q + 2] = v.z; d[4 * q + 3] = v.w;
            }
            float m1 = -10.f;
            int k1 = 0;
#pragma unroll
            for (int i = 0; i < RN; ++i) {
                m1 = (m1 < d[i]) ? d[i] : m1;
                if (d[i] > P.t2) k1++;
            }
            s2res[slot] = make_int2(k1, __float_as_int(m1));
            if (hs2 && ci < hcap) hs2[slot] = make_int2(k1, __float_as_int(m1));
            s_k1 = k1;
        }
        __syncthreads();
        const int k1 = s_k1;
        const int mn = (P.adapt && scale > 0.7f && P.min_neighbors > 1) ? P.min_neighbors - 1 : P.min_neighbors;
        const bool skip3 = (P.type_check == 0 && k1 >= mn) || (P.type_check > 0 && k1 <= 0) ||
                           (P.type_check == 2 && k1 < mn);
        if (skip3) {
            if (tid == 0) {
                s3res[slot] = make_int2(-1, __float_as_int(-10.f));
                if (hs3 && ci < hcap) hs3[slot] = make_int2(-1, __float_as_int(-10.f));
            }
        } else {
        write_input(s, u, false, true, P.reflection != 0, tid);
        PHASE(PH_CHECK, 16);
        PHASE_SET(PH_CHECK, 22, 1);
        check_forward<Stage3, RP>(s.in, s, s.w3, tid, MW3, s_w);
        if (tid == 0) {
            float d[RN];
#pragma unroll
            for (int q = 0; q < RN / 4; ++q) {
                const float4 v = reinterpret_cast<const float4*>(s.out())[q];
                d[4 * q] = v.x; d[4 * q + 1] = v.y; d[4 * q + 2] = v.z; d[4 * q + 3] = v.w;
            }
            float m2 = -10.f;
            int k2 = 0;
#pragma unroll
            for (int i = 0; i < RN; ++i) {
                m2 = (m2 < d[i]) ? d[i] : m2;
                if (d[i] > P.t3) k2++;
            }
            s3res[slot] = make_int2(k2, __float_as_int(m2));
            if (hs3 && ci < hcap) hs3[slot] = make_int2(k2, __float_as_int(m2));
        }
        }
        PHASE(PH_CHECK, 15);
        // the next item: every thread reads it after the barrier, and the previous item's
        // last LDS reads are done before it is written
        if (queue) {
            __syncthreads();
            if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
            __syncthreads();
            item = s_item;
        } else {
            __syncthreads();
            base += gridDim.x;
            item = base < items ? stride_item() : items;
        }
    }
    PHASE_FLUSH(PH_CHECK);
}

// Decision of CPUCheckDetect (cnn_detector_v3.cpp:1709-1725, 1808-1811) from the stage-2/3
// outcomes of both crop variants (s2/s3[mod] = {count above the threshold, max bits};
// s3 {-1, ..} when stage 3 was skipped): mod 0, then mod 1 if mod 0 is rejected. The host
// applies the same rule (check_decision) to the outcomes k_check writes to mapped memory.
__device__ CheckResult check_decision_dev(const CheckParams& P, float scale, const int2* __restrict__ s2,
                                          const int2* __restrict__ s3)
{
    const int mn = (P.adapt && scale > 0.7f && P.min_neighbors > 1) ? P.min_neighbors - 1 : P.min_neighbors;
    CheckResult res;
    res.n2 = 0;
    res.n3 = 0;
    for (int mod = 0; mod < 2; ++mod) {
        const int2 a = s2[mod];
        const int2 b = s3[mod];
        const int k1 = a.x, k2 = b.x < 0 ? 0 : b.x;
        const float m1 = __int_as_float(a.y), m2 = __int_as_float(b.y);
        res.n2 += 1;
        res.n3 += b.x < 0 ? 0 : 1;
        bool flag = false;
        if (P.type_check == 0) flag = (k1 >= mn || k2 >= mn);
        else if (P.type_check == 1) flag = ((k1 >= mn && k2 > 0) || (k1 > 0 && k2 >= mn));
        else if (P.type_check == 2) flag = (k1 >= mn && k2 >= mn);
        res.flag = flag; res.mod = mod; res.k1 = k1; res.k2 = k2; res.m1 = m1; res.m2 = m2;
        if (flag || !P.double_check) break;
    }
    res.fa.ok = 1;
    res.fa.gender = res.fa.smile = res.fa.glasses = -1;
    return res;
}

// The decisions of all candidates, to results (and, with hmeta, the count, candidates and
// results in mapped host memory for the first hcap).
__global__ void __launch_bounds__(256) k_check_decide(const LevelDesc* __restrict__ levels,
                                                      const Candidate* __restrict__ cands,
                                                      const int* __restrict__ meta, CheckParams P,
                                                      const int2* __restrict__ s2res, const int2* __restrict__ s3res,
                                                      CheckResult* __restrict__ results, int* __restrict__ hmeta,
                                                      Candidate* __restrict__ hcands,
                                                      CheckResult* __restrict__ hresults, int hcap)
{
    const int total = meta[0];
    if (hmeta && blockIdx.x == 0 && threadIdx.x == 0) hmeta[0] = total;
    for (int ci = blockIdx.x * blockDim.x + threadIdx.x; ci < total; ci += gridDim.x * blockDim.x) {
        const CheckResult res = check_decision_dev(P, levels[cands[ci].level].scale, s2res + (size_t)ci * 2,
                                                   s3res + (size_t)ci * 2);
        results[ci] = res;
        if (hresults && ci < hcap) {
            hresults[ci] = res;
            hcands[ci] = cands[ci];
        }
    }
}

// ---------------------------------------------------------------------------------
// Facial analysis (cnn_detector_v3.cpp:1731-1756): the landmarks network on the 49x51
// FP32 crop (the nearest-neighbour resize before the u8 conversion) of every window the
// checks accept. One workgroup of 512 threads (two per CU) per accepted candidate.
// ---------------------------------------------------------------------------------
namespace fa {
using N = Landmarks;
constexpr int W = 49, H = 51;                           // input
constexpr int P1W = (W - 3) / 2, P1H = (H - 3) / 2;     // 23 x 24 pooled L1
constexpr int P1N = P1W * P1H;
constexpr int P2W = (P1W - 2) / 2, P2H = (P1H - 2) / 2; // 10 x 11 pooled L2
constexpr int P2N = P2W * P2H;
constexpr int G1 = 6;                                   // L1 maps per work item
static_assert(P2W == N::K3W && P2H == N::K3H, "one L3 value per map");
}  // namespace fa

// LDS of one window (76 KB, small enough to share a CU with other kernels' workgroups):
//   big  pooled L1, then (in place) the L1 sums; after L2 the L3, hidden- and output-layer
//        products
//   c    the network input; after L1 the pooled L2, then (in place) the L2 sums
struct FaSmem {
    alignas(16) float big[Landmarks::M1 * fa::P1N];
    alignas(16) float c[Landmarks::M2 * fa::P2N > fa::H * fa::W ? Landmarks::M2 * fa::P2N : fa::H * fa::W];
    __device__ float* in() { return c; }
    float a3[Landmarks::M3];
    float h[Landmarks::HL];
    float out[16];
};

// Sums of neighbouring maps (add / add2: a[0]+a[1], (a[i-1]+a[i])+a[i+1], a[M-2]+a[M-1])
// of M maps of NP positions, in place: one work item per position reads all M maps of
// its position into registers, then writes the M sums there.
template <int M, int NP>
__device__ __forceinline__ void fa_map_sums(float* a, const int tid)
{
    for (int p = tid; p < NP; p += FA_THREADS) {
        float v[M];
#pragma unroll
        for (int k = 0; k < M; ++k) v[k] = a[k * NP + p];
        a[p] = v[0] + v[1];
#pragma unroll
        for (int i = 1; i < M - 1; ++i) a[i * NP + p] = v[i - 1] + v[i] + v[i + 1];
        a[(M - 1) * NP + p] = v[M - 2] + v[M - 1];
    }
}

// Dot-product terms x[xi(f)] * W[wi(f)], f < NF, formed in parallel: a thread's weights
// (f = tid + r * FA_THREADS) are loaded into registers by fa_load, one phase ahead of
// their use, and multiplied into dst[di(f)] by fa_mul.
template <int NF>
struct FaWeights {
    static constexpr int R = (NF + FA_THREADS - 1) / FA_THREADS;
    float w[R];
};

template <int NF, class WI>
__device__ __forceinline__ void fa_load(FaWeights<NF>& fw, const float* __restrict__ W, WI wi, const int tid)
{
#pragma unroll
    for (int r = 0; r < FaWeights<NF>::R; ++r) fw.w[r] = W[wi(min(tid + r * FA_THREADS, NF - 1))];
}

template <int NF, class XI, class DI>
__device__ __forceinline__ void fa_mul(const FaWeights<NF>& fw, const float* __restrict__ x, XI xi,
                                       float* __restrict__ dst, DI di, const int tid)
{
#pragma unroll
    for (int r = 0; r < FaWeights<NF>::R; ++r) {
        const int f = tid + r * FA_THREADS;
        if (f < NF) dst[di(f)] = x[xi(f)] * fw.w[r];
    }
}

// The landmarks network on s.in; all NOUT outputs to s.out. Layer arithmetic as in
// check_forward: conv_4x4 / conv_3x3 chains, lrelu_bn(_max), add / add2 sums; L3 is
// conv_11x10 (cnnpp_cplusplus.cpp:352-380), a chain that starts from 0.f; the hidden
// layer is mulCN_add_tanhW over all 96 maps; the outputs are mulC / mulC1_add / tanhW.
__device__ __noinline__ void landmarks_forward(const float* __restrict__ W, FaSmem& s)
{
    using namespace fa;
    const int tid = threadIdx.x;
    float* const A = s.big;
    float* const C = s.c;
    for (int it = tid; it < (N::M1 / G1) * P1N; it += FA_THREADS) {
        const int g = it / P1N, p = it - g * P1N;
        const int py = p / P1W, px = p - py * P1W;
        float win[5][5];
#pragma unroll
        for (int r = 0; r < 5; ++r)
#pragma unroll
            for (int c = 0; c < 5; ++c) win[r][c] = C[(2 * py + r) * fa::W + 2 * px + c];
#pragma unroll
        for (int mm = 0; mm < G1; ++mm) {
            const int m = g * G1 + mm;
            const float* k = W + N::oK1 + m * 16;
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int dx = q & 1, dy = q >> 1;
                float acc = win[dy][dx] * k[0];
#pragma unroll
                for (int j = 1; j < 16; ++j) acc = acc + win[dy + (j >> 2)][dx + (j & 3)] * k[j];
                a[q] = lrelu_bn(acc, W + N::oB1, m, N::M1);
            }
            A[m * P1N + p] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
        }
    }
    __syncthreads();
    fa_map_sums<N::M1, P1N>(A, tid);
    __syncthreads();
    for (int it = tid; it < N::M1 * P2N; it += FA_THREADS) {
        const int i = it / P2N, p = it - i * P2N;
        const int qy = p / P2W, qx = p - qy * P2W;
        const float* src = A + i * P1N + (2 * qy) * P1W + 2 * qx;
        float win[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) win[r][c] = src[r * P1W + c];
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int m = 2 * i + half;
            const float* k = W + N::oK2 + m * 9;
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int dx = q & 1, dy = q >> 1;
                float acc = win[dy][dx] * k[0];
#pragma unroll
                for (int j = 1; j < 9; ++j) acc = acc + win[dy + j / 3][dx + j % 3] * k[j];
                a[q] = lrelu_bn(acc, W + N::oB2, m, N::M2);
            }
            C[m * P2N + p] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
        }
    }
    __syncthreads();
    // L3, the hidden and the output layer are long dot products: all threads first form
    // their products (coalesced weight reads) in LDS, then one thread per output runs the
    // reference's chain over them in order (odd LDS strides: no bank conflicts). Each
    // phase's weights are loaded during the phase before it.
    constexpr int K3N = N::K3W * N::K3H, K3S = K3N + 1;
    static_assert(N::M3 * K3S <= N::M1 * P1N, "L3 products fit");
    constexpr int UPT = (N::HL + FA_THREADS - 1) / FA_THREADS;   // hidden units per thread
    constexpr int JC = N::HLN / 3, JS = JC + 1;                  // hidden dot products in 3 parts
    static_assert(N::HLN % 3 == 0 && N::HL * JS <= N::M1 * P1N, "hidden products fit");
    constexpr int OS = N::HL + 4;                                // output rows 16-byte aligned
    static_assert(N::HL % 4 == 0 && N::NOUT * OS <= N::M1 * P1N, "output products fit");
    // f = m * K3N + j: L3 map m's weight j, times sum map m / 2 at j (P2N == K3N)
    const auto l3_w = [](int f) { return N::oK3 + f; };
    const auto l3_x = [](int f) { return (f / (2 * K3N)) * K3N + f % K3N; };
    const auto l3_d = [](int f) { return f + f / K3N; };
    // f = u * JC + jj: hidden unit u's weight j0 + jj, times a3[u / HLS + j0 + jj]
    const auto hl_x = [](int f) { return (f / JC) / N::HLS + f % JC; };
    const auto hl_d = [](int f) { return f + f / JC; };
    // f = o * HL + q: output o's term q of its chain (q = HLS * gi + tt), times h[q]
    const auto ol_w = [](int f) {
        const int o = f / N::HL, q = f % N::HL;
        return N::oOW + o * N::HL + (q % N::HLS) * N::HLG + q / N::HLS;
    };
    FaWeights<N::M3 * K3N> w3;
    fa_load(w3, W, l3_w, tid);
    fa_map_sums<N::M2, P2N>(C, tid);
    __syncthreads();
    fa_mul(w3, C, l3_x, A, l3_d, tid);
    __syncthreads();
    FaWeights<N::HL * JC> wh[2];
    fa_load(wh[0], W, [](int f) { return N::oHW + (f / JC) * N::HLN + f % JC; }, tid);
    if (tid < N::M3) {
        const float* p = A + tid * K3S;
        float d = 0.f;
#pragma unroll 10
        for (int j = 0; j < K3N; ++j) d = d + p[j];
        s.a3[tid] = lrelu_bn(d, W + N::oB3, tid, N::M3);
    }
    __syncthreads();
    // hidden layer
    float c[UPT];
#pragma unroll
    for (int k = 0; k < UPT; ++k) c[k] = W[N::oHB + min(tid + k * FA_THREADS, N::HL - 1)];
    FaWeights<N::NOUT * N::HL> wo;
#pragma unroll
    for (int part = 0; part < 3; ++part) {
        const int j0 = part * JC;
        fa_mul(wh[part & 1], s.a3 + j0, hl_x, A, hl_d, tid);
        __syncthreads();
        if (part < 2)
            fa_load(wh[(part + 1) & 1], W,
                    [j0](int f) { return N::oHW + (f / JC) * N::HLN + j0 + JC + f % JC; }, tid);
        else
            fa_load(wo, W, ol_w, tid);
#pragma unroll
        for (int k = 0; k < UPT; ++k) {
            const int u = tid + k * FA_THREADS;
            if (u < N::HL) {
                const float* p = A + u * JS;
#pragma unroll 8
                for (int jj = 0; jj < JC; ++jj) c[k] = c[k] + p[jj];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < UPT; ++k) {
        const int u = tid + k * FA_THREADS;
        if (u < N::HL) {
            const int gi = u / N::HLS, tt = u - gi * N::HLS;
            float v = c[k] * W[N::oHT + gi];
            v = rtanh(v);
            v = (v + 1.f) * 0.5f;
            s.h[u] = W[N::oHBW + tt] * v + W[N::oHBB + tt];
        }
    }
    __syncthreads();
    // output layer
    fa_mul(wo, s.h, [](int f) { return f % N::HL; }, A, [](int f) { return f + 4 * (f / N::HL); }, tid);
    __syncthreads();
    if (tid < N::NOUT) {
        const float4* p4 = reinterpret_cast<const float4*>(A + tid * OS);
        float o = 0.f;
#pragma unroll 4
        for (int k = 0; k < N::HL / 4; ++k) {
            const float4 v = p4[k];
            o = (k == 0) ? v.x : v.x + o;
            o = v.y + o;
            o = v.z + o;
            o = v.w + o;
        }
        float c = o + W[N::oOB + tid];
        c = c * W[N::oOB + N::NOUT];
        c = rtanh(c);
        const float af = W[N::oOB + N::NOUT + 1];
        s.out[tid] = (af == 0.f) ? (c + 1.f) * 0.5f : af * c;
    }
    __syncthreads();
}

// Facial analysis of every accepted candidate, on the crop of its final crop variant.
// Every workgroup takes the check decisions of all candidates (k_check's outcomes), in
// chunks of FA_THREADS candidates, and numbers the accepted ones in candidate order;
// workgroup b runs the landmarks network on the accepted windows whose number is b modulo
// the grid size. The facial result goes to results[ci].fa and, for ci < hcap, to the
// mapped hresults[ci].fa.
__global__ void __launch_bounds__(FA_THREADS, 4) k_facial(const float* __restrict__ g0, int gp0, int gw0, int gh0,
                                                          const LevelDesc* __restrict__ levels,
                                                          const Candidate* __restrict__ cands,
                                                          const int* __restrict__ meta, CheckParams P,
                                                          const int2* __restrict__ s2res,
                                                          const int2* __restrict__ s3res,
                                                          const float* __restrict__ W,
                                                          CheckResult* __restrict__ results,
                                                          CheckResult* __restrict__ hresults, int hcap)
{
    using namespace fa;
    __shared__ FaSmem s;
    __shared__ int s_list[FA_THREADS];                    // accepted candidates of the chunk
    __shared__ int s_mod[FA_THREADS];                     // and their final crop variants
    __shared__ int s_wcnt[FA_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int total = meta[0];
    const int G = gridDim.x;
    int base = 0;                                         // accepted before this chunk
    for (int c0 = 0; c0 < total; c0 += FA_THREADS) {
        const int ci = c0 + tid;
        bool acc = false;
        int mod = 0;
        if (ci < total) {
            const CheckResult res = check_decision_dev(P, levels[cands[ci].level].scale, s2res + (size_t)ci * 2,
                                                       s3res + (size_t)ci * 2);
            acc = res.flag != 0;
            mod = res.mod;
        }
        const unsigned long long bal = __ballot(acc);
        if (lane == 0) s_wcnt[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, n = 0;
#pragma unroll
        for (int k = 0; k < FA_THREADS / 64; ++k) {
            off += k < wave ? s_wcnt[k] : 0;
            n += s_wcnt[k];
        }
        if (acc) {
            const int k = off + __popcll(bal & ((1ull << lane) - 1));
            s_list[k] = ci;
            s_mod[k] = mod;
        }
        __syncthreads();
        // this workgroup's windows of the chunk: numbers base + k with (base + k) % G == b
        for (int k = (int)(((unsigned)blockIdx.x + (unsigned)G - (unsigned)(base % G)) % (unsigned)G); k < n;
             k += G) {
            const int wi = s_list[k];
            const Candidate cd = cands[wi];
            const GrayRef gr = level_gray(g0, gp0, gw0, gh0, levels[cd.level]);
            const float* __restrict__ g = gr.g;
            const int gp = gr.gp;
            const Crop cr = crop_geometry(cd.x * 4, cd.y * 4, levels[cd.level].scale, s_mod[k], gr.gw, gr.gh);
            // nearest-neighbour resize of the crop to 49x51, a copy at that size
            // (image_resize.cpp:328-392, 564-584)
            const bool copy = (cr.rcols == fa::W && cr.rrows == fa::H);
            const uint32_t sclx = ((uint32_t)cr.rcols << 12) / fa::W + 1;
            const uint32_t scly = ((uint32_t)cr.rrows << 12) / fa::H + 1;
            constexpr int PER = (fa::W * fa::H + FA_THREADS - 1) / FA_THREADS;
            float v[PER];
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int i = min(tid + q * FA_THREADS, fa::W * fa::H - 1);
                const int iy = i / fa::W, ix = i - iy * fa::W;
                uint32_t px = ix, py = iy;
                if (!copy) {
                    px = min(((uint32_t)ix * sclx) >> 12, (uint32_t)(cr.rcols - 2));
                    py = min(((uint32_t)iy * scly) >> 12, (uint32_t)(cr.rrows - 1));
                }
                v[q] = g[(size_t)(cr.ry + py) * gp + cr.rx + px];
            }
            float* in = s.in();
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int i = tid + q * FA_THREADS;
                if (i < fa::W * fa::H) in[i] = v[q];
            }
            __syncthreads();
            landmarks_forward(W, s);
            if (tid == 0) {
                int zero = 0;
                for (int t = 0; t < 5; ++t) {
                    const int lx = (int)(s.out[2 * t] * (float)cr.rcols);
                    const int ly = (int)(s.out[2 * t + 1] * (float)cr.rrows);
                    if (lx < (cr.rcols >> 1) && ly < (cr.rrows >> 1)) zero++;
                }
                FaResult f;
                f.ok = zero != 5;
                f.gender = (int)(s.out[10] + 0.5f);
                f.smile = (int)(s.out[11] + 0.5f);
                f.glasses = (int)(s.out[12] + 0.5f);
                results[wi].fa = f;
                if (hresults && wi < hcap) hresults[wi].fa = f;
            }
            __syncthreads();
        }
        base += n;
    }
}

// The stage-2 or stage-3 network on n 47x51 inputs (tests): its 4x4 response map per
// input, one workgroup per input.
template <class N, int RP>
__global__ void __launch_bounds__(CHECK_THREADS, 4) k_net23(const float* __restrict__ in,
                                                         const float* __restrict__ W, float* __restrict__ out,
                                                         const float* __restrict__ MW)
{
    using namespace chk;
    __shared__ ChkSmem s;
    __shared__ int s_red[CHECK_THREADS / 64];
    const int tid = threadIdx.x;
    const float* src = in + (size_t)blockIdx.x * IW * IH;
    stage_weights<N::SIZE>(s.w, W, tid);
    for (int i = tid; i < IW * IH; i += CHECK_THREADS) s.in[i] = src[i];
    __syncthreads();
    check_forward<N, RP>(s.in, s, s.w, tid, MW, s_red);
    if (tid < RN) out[(size_t)blockIdx.x * RN + tid] = s.out()[tid];
}

// The landmarks network on one 49x51 input (tests).
__global__ void __launch_bounds__(FA_THREADS, 4) k_landmarks(const float* __restrict__ in,
                                                             const float* __restrict__ W, float* __restrict__ out)
{
    __shared__ FaSmem s;
    for (int i = threadIdx.x; i < fa::W * fa::H; i += FA_THREADS) s.in()[i] = in[i];
    __syncthreads();
    landmarks_forward(W, s);
    if (threadIdx.x < Landmarks::NOUT) out[threadIdx.x] = s.out[threadIdx.x];
}

// Self-test of the exact-path reciprocal: over float bit patterns [lo, hi) (NaNs
// skipped), counts the inputs where rtanh with recip_ge1 differs in any bit from rtanh
// with IEEE division (out[0]) and, for inputs in [1, 2^25), where recip_ge1 differs from
// 1/d (out[1]); out[2] keeps the largest mismatching pattern.
__global__ void __launch_bounds__(256) k_selftest_tanh(uint32_t lo, uint32_t hi,
                                                       unsigned long long* __restrict__ out)
{
    unsigned long long bad_t = 0, bad_r = 0, last = 0;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    for (uint64_t b = (uint64_t)lo + blockIdx.x * 256 + threadIdx.x; b < hi; b += stride) {
        const float c = __uint_as_float((uint32_t)b);
        if (c != c) continue;
        const float t0 = rtanh<false>(c), t1 = rtanh<true>(c);
        if (__float_as_uint(t0) != __float_as_uint(t1)) { ++bad_t; last = b; }
        if (c >= 1.f && c < 33554432.f) {
            const float r0 = recip_ge1(c), r1 = 1.f / c;
            if (__float_as_uint(r0) != __float_as_uint(r1)) { ++bad_r; last = b; }
        }
    }
    if (bad_t) atomicAdd(&out[0], bad_t);
    if (bad_r) atomicAdd(&out[1], bad_r);
    if (bad_t || bad_r) atomicMax(&out[2], last);
}

}  // namespace ccnn

// ---------------------------------------------------------------------------------
// Launchers (host side of this translation unit)
// ---------------------------------------------------------------------------------
namespace ccnn {

hipError_t launch_gray(const uint8_t* img, int w, int h, int ch, int step, float* gray, int gp, hipStream_t st)
{
    const bool aligned = ((uintptr_t)img % (ch == 4 ? 16 : 4)) == 0 && step % (ch == 4 ? 16 : 4) == 0 &&
                         ((uintptr_t)gray % 16) == 0 && gp % 4 == 0;
    // (one frame alone keeps more loads in flight with 4 pixels per thread than with the
    // batch kernel's 16: 9.4 against 11.3 us for a 1080p frame)
    if (aligned) {
        const int qw = (w + 3) / 4;
        const long long quads = (long long)qw * h;
        k_gray<<<(unsigned)((quads + 255) / 256), 256, 0, st>>>(img, w, h, ch, step, gray, gp, qw);
    } else {
        dim3 grid((w + 255) / 256, h);
        k_gray_bytes<<<grid, 256, 0, st>>>(img, w, h, ch, step, gray, gp);
    }
    return hipGetLastError();
}

hipError_t launch_resize_u8(const uint8_t* src, int sw, int sh, int sstep, uint8_t* dst, int dw, int dh,
                            hipStream_t st)
{
    if (dw == sw && dh == sh)
        return hipMemcpy2DAsync(dst, dw, src, sstep, dw, dh, hipMemcpyDeviceToDevice, st);
    const uint32_t sclx = ((uint32_t)sw << 12) / (uint32_t)dw + 1;
    const uint32_t scly = ((uint32_t)sh << 12) / (uint32_t)dh + 1;
    dim3 grid((dw + 255) / 256, dh);
    k_resize_u8<<<grid, 256, 0, st>>>(src, sw, sh, sstep, dst, dw, dh, sclx, scly);
    return hipGetLastError();
}

hipError_t launch_level(const float* gray, int gp, int gw, int gh, const LevelDesc& L, float* out, hipStream_t st)
{
    dim3 grid((L.lw + 255) / 256, L.lh);
    k_level<<<grid, 256, 0, st>>>(gray, gp, gw, gh, L, out);
    return hipGetLastError();
}

namespace {

template <int AX, int AY, bool H>
void launch_l1(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, int nlevels, int nl1tiles,
               const float* W1, float* s1buf, int* meta, const int2* l1_order, hipStream_t st, const hipEvent_t* kev,
               const unsigned char* l1_keep)
{
    if (kev)
        hipExtLaunchKernelGGL(k_stage1_l1<AX, AY, H>, dim3(nl1tiles), dim3(S1A_THREADS), 0, st, kev[0], kev[1], 0, gray,
                              gp, gw, gh, d_levels, nlevels, W1, s1buf, meta, l1_order, l1_keep);
    else
        k_stage1_l1<AX, AY, H><<<nl1tiles, S1A_THREADS, 0, st>>>(gray, gp, gw, gh, d_levels, nlevels, W1, s1buf, meta,
                                                                l1_order, l1_keep);
}

template <int TX, int TY, bool H>
void launch_resp(const LevelDesc* d_levels, int nlevels, int ntiles, const float* W1, const float* s1buf, float* resp,
                 float t1, int* meta, Candidate* cands, Candidate* hcands, int hcap, const int2* order,
                 hipStream_t st, const hipEvent_t* kev, const unsigned char* keep)
{
    if (kev)
        hipExtLaunchKernelGGL(k_stage1<TX, TY, H>, dim3(ntiles), dim3(S1_THREADS), 0, st, kev[2], kev[3], 0, d_levels,
                              nlevels, W1, s1buf, resp, t1, meta, cands, hcands, hcap, order, keep);
    else
        k_stage1<TX, TY, H><<<ntiles, S1_THREADS, 0, st>>>(d_levels, nlevels, W1, s1buf, resp, t1, meta, cands, hcands,
                                                           hcap, order, keep);
}

}  // namespace

template <bool H>
hipError_t launch_stage1_t(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, int nlevels,
                           int nl1tiles, int ntiles, const float* W1, float* s1buf, float* resp, float t1, int* meta,
                           Candidate* cands, Candidate* hcands, int hcap, const int2* l1_order, hipStream_t st,
                           const hipEvent_t* kev, int shape, const unsigned char* keep, const unsigned char* l1_keep)
{
    const TileShape t = tile_shape(shape);
#define L1_ARGS gray, gp, gw, gh, d_levels, nlevels, nl1tiles, W1, s1buf, meta, l1_order, st, kev, l1_keep
    if (t.ax == 64 && t.ay == 16) launch_l1<64, 16, H>(L1_ARGS);
    else if (t.ax == 32 && t.ay == 16) launch_l1<32, 16, H>(L1_ARGS);
    else if (t.ax == 32 && t.ay == 32) launch_l1<32, 32, H>(L1_ARGS);
    else if (t.ax == 64 && t.ay == 8) launch_l1<64, 8, H>(L1_ARGS);
    else return hipErrorInvalidValue;
#undef L1_ARGS
    // l1_order (optional) holds the nl1tiles L1 tiles, then the ntiles stage-1 tiles
    const int2* s1_order = l1_order ? l1_order + nl1tiles : nullptr;
#define RESP_ARGS d_levels, nlevels, ntiles, W1, s1buf, resp, t1, meta, cands, hcands, hcap, s1_order, st, kev, keep
    if (t.tx == 16 && t.ty == 16) launch_resp<16, 16, H>(RESP_ARGS);
    else if (t.tx == 32 && t.ty == 8) launch_resp<32, 8, H>(RESP_ARGS);
    else if (t.tx == 16 && t.ty == 8) launch_resp<16, 8, H>(RESP_ARGS);
    else return hipErrorInvalidValue;
#undef RESP_ARGS
    return hipGetLastError();
}

hipError_t launch_stage1(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, int nlevels,
                         int nl1tiles, int ntiles, const float* W1, float* s1buf, float* resp, float t1,
                         int* meta, Candidate* cands, Candidate* hcands, int hcap, const int2* l1_order,
                         hipStream_t st, const hipEvent_t* kev, int shape, bool half, const unsigned char* keep,
                         const unsigned char* l1_keep)
{
    if (ntiles <= 0) return hipSuccess;
    // kev (optional): start/stop events of the two launches, stamped when each kernel
    // starts and ends executing. shape: the tile shapes the plan was laid out for; half:
    // FP16 convolutions (W1 then carries the Stage1H pairs after the FP32 weights)
    if (half)
        return launch_stage1_t<true>(gray, gp, gw, gh, d_levels, nlevels, nl1tiles, ntiles, W1, s1buf, resp, t1, meta,
                                     cands, hcands, hcap, l1_order, st, kev, shape, keep, l1_keep);
    return launch_stage1_t<false>(gray, gp, gw, gh, d_levels, nlevels, nl1tiles, ntiles, W1, s1buf, resp, t1, meta,
                                  cands, hcands, hcap, l1_order, st, kev, shape, keep, l1_keep);
}

// Read-back of a frame's candidate list and check results straight into mapped host
// memory (no copy launches): the count to hmeta[0], the first min(count, cap) records.
__global__ void __launch_bounds__(256) k_export(const int* __restrict__ meta, const Candidate* __restrict__ cands,
                                                const CheckResult* __restrict__ results, int* __restrict__ hmeta,
                                                Candidate* __restrict__ hcands, CheckResult* __restrict__ hresults,
                                                int cap)
{
    const int total = meta[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) hmeta[0] = total;
    const int n = min(total, cap);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        hcands[i] = cands[i];
        if (results) hresults[i] = results[i];
    }
}

hipError_t launch_export(const int* meta, const Candidate* cands, const CheckResult* results, int* hmeta,
                         Candidate* hcands, CheckResult* hresults, int cap, hipStream_t st)
{
    k_export<<<32, 256, 0, st>>>(meta, cands, results, hmeta, hcands, hresults, cap);
    return hipGetLastError();
}

hipError_t launch_facial(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, const Candidate* cands,
                         const int* meta, const CheckParams& P, const int2* s2res, const int2* s3res, const float* W,
                         CheckResult* results, CheckResult* hresults, int hcap, int nblocks, hipStream_t st)
{
    k_facial<<<nblocks, FA_THREADS, 0, st>>>(gray, gp, gw, gh, d_levels, cands, meta, P, s2res, s3res, W, results,
                                             hresults, hcap);
    return hipGetLastError();
}

hipError_t launch_selftest_tanh(uint32_t lo, uint32_t hi, unsigned long long* out, hipStream_t st)
{
    k_selftest_tanh<<<8192, 256, 0, st>>>(lo, hi, out);
    return hipGetLastError();
}

hipError_t launch_net23(int stage, const float* in, int n, const float* W, float* out, hipStream_t st, int rp,
                        const float* MW)
{
    if (n <= 0) return hipSuccess;
    if (rp < RP_EXACT || rp > RP_FP8 || (rp != RP_EXACT && !MW)) return hipErrorInvalidValue;
#define NET23(NN, R) k_net23<NN, R><<<n, CHECK_THREADS, 0, st>>>(in, W, out, MW)
    if (stage == 1) {
        if (rp == RP_EXACT) NET23(Stage2, RP_EXACT); else if (rp == RP_BF16) NET23(Stage2, RP_BF16); else NET23(Stage2, RP_FP8);
    } else {
        if (rp == RP_EXACT) NET23(Stage3, RP_EXACT); else if (rp == RP_BF16) NET23(Stage3, RP_BF16); else NET23(Stage3, RP_FP8);
    }
#undef NET23
    return hipGetLastError();
}

hipError_t launch_landmarks(const float* in, const float* W, float* out, hipStream_t st)
{
    k_landmarks<<<1, FA_THREADS, 0, st>>>(in, W, out);
    return hipGetLastError();
}

hipError_t launch_resize_u8_tab(const FrameDesc* frames, int f, int dw, int dh, hipStream_t st)
{
    dim3 grid((dw + 255) / 256, dh);
    k_resize_u8_tab<<<grid, 256, 0, st>>>(frames, f);
    return hipGetLastError();
}

hipError_t launch_gray_batch(const FrameDesc* frames, int nframes, int total_quads, float* gray, hipStream_t st)
{
    if (total_quads <= 0) return hipSuccess;
    k_gray_batch<<<(total_quads + 255) / 256, 256, 0, st>>>(frames, nframes, total_quads, gray);
    return hipGetLastError();
}

hipError_t launch_compact(const float* resp, int n, float t1, int* counts, int* offs, const LevelDesc* d_levels,
                          int nlevels, Candidate* cands, int cap, Candidate* hcands, int hcap, int* meta, int* hmeta,
                          hipStream_t st)
{
    const int nb = (n + CMP_BLOCK - 1) / CMP_BLOCK;
    if (nb > 0) k_cmp_count<<<nb, 256, 0, st>>>(resp, n, t1, counts);
    k_cmp_scan<<<1, CMP_SCAN_THREADS, 0, st>>>(counts, nb, offs, meta, hmeta);
    if (nb > 0) k_cmp_emit<<<nb, 256, 0, st>>>(resp, n, t1, offs, d_levels, nlevels, cands, cap, hcands, hcap);
    return hipGetLastError();
}

hipError_t launch_check(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, const Candidate* cands,
                        const int* meta, const CheckParams& P, const float* W2, const float* W3, int2* s2res,
                        int2* s3res, CheckResult* results, int nblocks, int* hmeta, Candidate* hcands,
                        CheckResult* hresults, int2* hs2, int2* hs3, int hcap, hipStream_t st, const hipEvent_t* kev,
                        int cap, int* queue, int rp, const float* MW2, const float* MW3)
{
    if (rp < RP_EXACT || rp > RP_FP8 || (rp != RP_EXACT && (!MW2 || !MW3))) return hipErrorInvalidValue;
    // kev (optional): start/stop events of k_check, stamped when it starts and ends executing.
    // With hs2/hs3 the outcomes go to mapped host memory and the host takes the decisions
    // (no k_check_decide); otherwise k_check_decide writes results (and, with hmeta, the
    // read-back).
    int* km = hs2 ? hmeta : nullptr;
    auto go = [&](auto kern) {
        if (kev)
            hipExtLaunchKernelGGL(kern, dim3(nblocks), dim3(CHECK_THREADS), 0, st, kev[0], kev[1], 0, gray, gp, gw, gh,
                                  d_levels, cands, meta, P, W2, W3, s2res, s3res, km, hs2, hs3, hcap, cap, queue, MW2,
                                  MW3);
        else
            kern<<<nblocks, CHECK_THREADS, 0, st>>>(gray, gp, gw, gh, d_levels, cands, meta, P, W2, W3, s2res, s3res, km,
                                                    hs2, hs3, hcap, cap, queue, MW2, MW3);
    };
    if (rp == RP_EXACT) go(k_check<RP_EXACT>);
    else if (rp == RP_BF16) go(k_check<RP_BF16>);
    else go(k_check<RP_FP8>);
    if (!hs2)
        k_check_decide<<<(nblocks + 255) / 256, 256, 0, st>>>(d_levels, cands, meta, P, s2res, s3res, results, hmeta,
                                                              hcands, hresults, hcap);
    return hipGetLastError();
}

// The phase timestamps of kernel kern (PH_CHECK or PH_STAGE1): nblocks workgroups x
// PHASE_SLOTS values to out (when not null), then all of that kernel's are cleared. For
// an idle device; -1 without CCNN_PHASE_STAMPS or for arguments out of range.
int read_phase_stamps(int kern, unsigned long long* out, int nblocks)
{
#ifdef CCNN_PHASE_STAMPS
    if (kern < 0 || kern > 1 || nblocks < 0 || nblocks > PHASE_BLOCKS) return -1;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_phase)) != hipSuccess) return -1;
    unsigned long long* base = static_cast<unsigned long long*>(p) + (size_t)kern * PHASE_BLOCKS * PHASE_SLOTS;
    if (out && nblocks > 0 &&
        hipMemcpy(out, base, sizeof(unsigned long long) * PHASE_SLOTS * (size_t)nblocks, hipMemcpyDeviceToHost) !=
            hipSuccess)
        return -1;
    if (hipMemset(base, 0, sizeof(unsigned long long) * PHASE_SLOTS * (size_t)PHASE_BLOCKS) != hipSuccess) return -1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
#else
    (void)kern;
    (void)out;
    (void)nblocks;
    return -1;
#endif
}

}  // namespace ccnn
