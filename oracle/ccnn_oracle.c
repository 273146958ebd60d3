/*
 * ccnn_oracle.c -- CPU restatement of the reference detection path. TEST
 * INFRASTRUCTURE ONLY (see ccnn_oracle.h): the checker, never the product.
 *
 * Citations are reference file:line (src/CNNObjectDetector/ unless noted).
 * Floating-point expressions keep the reference's source order; the file is built
 * without contraction or reassociation so every rounding step matches the
 * reference's CMake-default (-ffast-math, unoptimised) build.
 *
 * Parity status per function is in the header; the third-party algorithm restated
 * here is libstdc++'s std::sort (GCC 11.4, bits/stl_algo.h / stl_heap.h), used by the
 * reference before NMS (cnn_detector_v3.cpp:2750).
 */
#include "ccnn_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OC_MIN(a, b) ((a) > (b) ? (b) : (a))   /* type.h:33 */
#define OC_MAX(a, b) ((a) < (b) ? (b) : (a))   /* type.h:34 */

static const float TANH_A = 1.41645f;          /* cnnpp_cplusplus.h:38 */

/* ------------------------------------------------------------------------------ */
/* Model loading: cnn_simd_cntk.cpp:42-404 (non-SIMD branches), format of App. B.  */
/* ------------------------------------------------------------------------------ */

typedef struct { const unsigned char* p; long n, off; int err; } rd_t;

static void rd(rd_t* r, void* dst, long sz)
{
    if (r->off + sz > r->n) { r->err = 1; memset(dst, 0, (size_t)sz); return; }
    memcpy(dst, r->p + r->off, (size_t)sz);
    r->off += sz;
}
static int rd_i32(rd_t* r) { int32_t v; rd(r, &v, 4); return v; }
static float rd_f32(rd_t* r) { float v; rd(r, &v, 4); return v; }
static unsigned char rd_u8(rd_t* r) { unsigned char v; rd(r, &v, 1); return v; }
static float* rd_arr(rd_t* r, long n)
{
    if (n <= 0 || n > (1L << 26)) { r->err = 1; return NULL; }
    float* a = (float*)malloc(sizeof(float) * (size_t)n);
    for (long i = 0; i < n; ++i) a[i] = rd_f32(r);
    return a;
}

void oc_model_free(oc_model* m)
{
    for (int l = 0; l < 3; ++l) {
        free(m->kern[l]); free(m->conv_b[l]); free(m->lrelu_w1[l]); free(m->lrelu_w2[l]);
        free(m->bn_w[l]); free(m->bn_b[l]);
    }
    free(m->hl_w); free(m->hl_b); free(m->hl_tanh_w); free(m->hl_bn_w); free(m->hl_bn_b);
    free(m->ol_w); free(m->ol_b);
    memset(m, 0, sizeof(*m));
}

int oc_model_load(oc_model* m, const unsigned char* bytes, long n, int index_output)
{
    memset(m, 0, sizeof(*m));
    rd_t r = { bytes, n, 0, 0 };
    float version = rd_f32(&r);                         /* :71-78 */
    if (r.err || version < 1.0f || version > 1.1f) return -1;
    int max_pool = rd_u8(&r);                           /* :81 */
    m->min_w = rd_i32(&r);                              /* :84-85 */
    m->min_h = rd_i32(&r);
    int layers = rd_i32(&r);                            /* :88 */
    if (r.err || layers != 3) { oc_model_free(m); return -1; }
    for (int l = 0; l < 3; ++l) m->maps[l] = rd_i32(&r);
    for (int l = 0; l < 3; ++l) {                       /* :100-277 */
        m->kw[l] = rd_i32(&r);
        m->kh[l] = rd_i32(&r);
        if (r.err || m->kw[l] <= 0 || m->kh[l] <= 0 || m->maps[l] <= 0 || m->maps[l] > 1024) {
            oc_model_free(m); return -1;
        }
        m->kern[l] = rd_arr(&r, (long)m->maps[l] * m->kw[l] * m->kh[l]);
    }
    m->af_scale = rd_f32(&r);                           /* :280 */
    for (int l = 0; l < 3; ++l) {                       /* :287-305 */
        int c = m->maps[l];
        m->conv_b[l] = (float*)malloc(sizeof(float) * c);
        m->lrelu_w1[l] = (float*)malloc(sizeof(float) * c);
        m->lrelu_w2[l] = (float*)malloc(sizeof(float) * c);
        m->bn_w[l] = (float*)malloc(sizeof(float) * c);
        m->bn_b[l] = (float*)malloc(sizeof(float) * c);
        for (int j = 0; j < c; ++j) {
            m->conv_b[l][j] = rd_f32(&r);
            m->lrelu_w1[l][j] = rd_f32(&r);
            m->lrelu_w2[l][j] = rd_f32(&r);
            m->bn_w[l][j] = rd_f32(&r);
            m->bn_b[l][j] = rd_f32(&r);
        }
    }
    m->full_connect = rd_u8(&r);                        /* :308-311 */
    m->hl = rd_i32(&r);
    m->connect = rd_i32(&r);
    m->hl_scale = rd_i32(&r);
    /* supported configurations only (:313-320) */
    if (r.err || !max_pool || m->kw[1] * m->kh[1] != 9 || m->hl <= 0 || m->hl_scale <= 0 ||
        m->connect <= 0 || m->hl % m->hl_scale != 0) {
        oc_model_free(m); return -1;
    }
    m->hl_nin = m->full_connect ? m->maps[2] : m->connect;   /* :322-347 */
    m->hl_w = rd_arr(&r, (long)m->hl * m->hl_nin);
    m->hl_b = rd_arr(&r, m->hl);                             /* :349-353 */
    m->hl_tanh_w = rd_arr(&r, m->hl / m->hl_scale);          /* :355-359 */
    m->hl_bn_w = (float*)malloc(sizeof(float) * m->hl_scale);/* :361-367 */
    m->hl_bn_b = (float*)malloc(sizeof(float) * m->hl_scale);
    for (int i = 0; i < m->hl_scale; ++i) {
        m->hl_bn_w[i] = rd_f32(&r);
        m->hl_bn_b[i] = rd_f32(&r);
    }
    m->ol = rd_i32(&r);                                      /* :369-379 */
    if (r.err || m->ol <= 0) { oc_model_free(m); return -1; }
    m->ol_w = rd_arr(&r, (long)m->ol * m->hl);
    m->ol_b = rd_arr(&r, m->ol);                             /* :381-385 */
    m->ol_tanh_w = rd_f32(&r);                               /* :387 */
    if (r.err) { oc_model_free(m); return -1; }
    m->index_output = OC_MIN(index_output, m->ol - 1);       /* :395-396 */
    if (m->index_output == 0) m->af_scale = -m->af_scale;
    /* index_output < 0: every output map is computed (cnn_simd_cntk.cpp:935-949) */
    /* the layer wiring (add/add2 + two convs per sum) needs maps[l+1] == 2*maps[l] */
    if (m->maps[1] != 2 * m->maps[0] || m->maps[2] != 2 * m->maps[1]) { oc_model_free(m); return -1; }
    if (!((m->kw[0] == 4 && m->kh[0] == 4) || (m->kw[0] == 3 && m->kh[0] == 3))) { oc_model_free(m); return -1; }
    if (m->hl / m->hl_scale + m->connect - 1 > m->maps[2] || m->connect > m->hl_nin) { oc_model_free(m); return -1; }
    return 0;
}

/* getOutputImgSize, cnn_simd_cntk.cpp:997-1012 */
void oc_output_size(const oc_model* m, int w, int h, int* ow, int* oh)
{
    int c1 = w - (m->kw[0] - 1), r1 = h - (m->kh[0] - 1);
    int c2 = (c1 >> 1) - (m->kw[1] - 1), r2 = (r1 >> 1) - (m->kh[1] - 1);
    *ow = (c2 >> 1) - (m->kw[2] - 1);
    *oh = (r2 >> 1) - (m->kh[2] - 1);
}

/* ------------------------------------------------------------------------------ */
/* Layer arithmetic: cnnpp_cplusplus.cpp                                          */
/* ------------------------------------------------------------------------------ */

/* Rational tanh of the model (USE_FAST_TANH, config.h:60; cnnpp_cplusplus.cpp:662-666):
 *   sgn * (1 - 1 / (1 + |c| + c*c + a*c*c*c*c)).
 * The reference's build (-ffast-math) folds the leading 1.f to the end of the sum;
 * that association is kept here: ((|c| + c*c) + a*c*c*c*c) + 1. */
static float rtanh(float c)
{
    float sgn = 1.f;
    if (c < 0.f) sgn = -1.f;
    float den = ((fabsf(c) + c * c) + TANH_A * c * c * c * c) + 1.f;
    return (1.f - 1.f / den) * sgn;
}

/* lrelu_bn / lrelu_bn_max element (cnnpp_cplusplus.cpp:604-605, 630-631):
 *   w1*c + w2*max(0, c) + bn_b, with c = conv + conv_b.
 * Association as the reference's build evaluates it: (w2*max(0,c) + bn_b) + w1*c. */
static float lrelu_bn(float conv, float b, float w1, float w2, float bnb)
{
    float c = conv + b;
    return (w2 * fmaxf(0.f, c) + bnb) + w1 * c;
}

/* Valid cross-correlation at (x, y) of `src` (row stride sstep) with a kh x kw kernel.
 * conv_5x4 (cnnpp_cplusplus.cpp:158-187) accumulates one row sum at a time; conv_11x10
 * (:352-380) is a chain that starts from 0.f; every other kernel (conv_3x3 :36-91,
 * conv_4x4 :92-157, conv_8x7 :292-321) is a single left-to-right chain over the rows. */
typedef float (*sample_fn)(const void* ctx, int x, int y);

static float conv_at(const float* k, int kw, int kh, sample_fn f, const void* ctx, int x, int y)
{
    if (kh == 5 && kw == 4) {
        float d = 0.f;
        for (int r = 0; r < 5; ++r) {
            float rs = f(ctx, x, y + r) * k[r * 4];
            for (int c = 1; c < 4; ++c) rs = rs + f(ctx, x + c, y + r) * k[r * 4 + c];
            d = (r == 0) ? rs : d + rs;
        }
        return d;
    }
    if (kh == 11 && kw == 10) {
        float d = 0.f;
        for (int r = 0; r < 11; ++r)
            for (int c = 0; c < 10; ++c) d = d + f(ctx, x + c, y + r) * k[r * 10 + c];
        return d;
    }
    float acc = f(ctx, x, y) * k[0];
    for (int r = 0; r < kh; ++r)
        for (int c = 0; c < kw; ++c) {
            if (r == 0 && c == 0) continue;
            acc = acc + f(ctx, x + c, y + r) * k[r * kw + c];
        }
    return acc;
}

typedef struct { const float* p; int step; } plane_t;
static float sample_plane(const void* ctx, int x, int y)
{
    const plane_t* pl = (const plane_t*)ctx;
    return pl->p[(long)y * pl->step + x];
}

/* add / add2 (cnnpp_cplusplus.cpp:916-950) with the wiring of cnn_simd_cntk.cpp:705-725 */
typedef struct { const float* const* maps; int nmaps; int idx; int step; } sum_t;
static float sample_sum(const void* ctx, int x, int y)
{
    const sum_t* s = (const sum_t*)ctx;
    long o = (long)y * s->step + x;
    int i = s->idx, t = s->nmaps;
    if (i > 0 && i < t - 1) return s->maps[i - 1][o] + s->maps[i][o] + s->maps[i + 1][o];
    if (i == 0) return s->maps[0][o] + s->maps[1][o];
    return s->maps[t - 2][o] + s->maps[t - 1][o];
}

int oc_cnn_forward(const oc_model* m, const float* img, int w, int h, int step, float* out)
{
    if (w < m->min_w || h < m->min_h) return -1;
    int M1 = m->maps[0], M2 = m->maps[1], M3 = m->maps[2];
    /* L1: conv + lrelu_bn_max (cnn_simd_cntk.cpp:649-663) over the valid region */
    int c1 = w - (m->kw[0] - 1), r1 = h - (m->kh[0] - 1);
    int p1w = c1 >> 1, p1h = r1 >> 1;
    int c2 = p1w - 2, r2 = p1h - 2;
    int p2w = c2 >> 1, p2h = r2 >> 1;
    int c3 = p2w - (m->kw[2] - 1), r3 = p2h - (m->kh[2] - 1);
    if (c3 <= 0 || r3 <= 0) return -1;

    float** P1 = (float**)malloc(sizeof(float*) * M1);
    plane_t in = { img, step };
    for (int k = 0; k < M1; ++k) {
        P1[k] = (float*)malloc(sizeof(float) * (size_t)p1w * p1h);
        const float* ker = m->kern[0] + (long)k * m->kw[0] * m->kh[0];
        for (int y = 0; y < p1h; ++y)
            for (int x = 0; x < p1w; ++x) {
                float a[4];
                for (int q = 0; q < 4; ++q) {
                    float cv = conv_at(ker, m->kw[0], m->kh[0], sample_plane, &in, 2 * x + (q & 1), 2 * y + (q >> 1));
                    a[q] = lrelu_bn(cv, m->conv_b[0][k], m->lrelu_w1[0][k], m->lrelu_w2[0][k], m->bn_b[0][k]);
                }
                P1[k][(long)y * p1w + x] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
            }
    }
    /* L2: sums of neighbouring pooled maps, two 3x3 convs per sum, lrelu_bn_max (:705-731) */
    float** P2 = (float**)malloc(sizeof(float*) * M2);
    for (int i = 0; i < M2 / 2; ++i) {
        sum_t s = { (const float* const*)P1, M1, i, p1w };
        for (int half = 0; half < 2; ++half) {
            int k = 2 * i + half;
            P2[k] = (float*)malloc(sizeof(float) * (size_t)p2w * p2h);
            const float* ker = m->kern[1] + (long)k * 9;
            for (int y = 0; y < p2h; ++y)
                for (int x = 0; x < p2w; ++x) {
                    float a[4];
                    for (int q = 0; q < 4; ++q) {
                        float cv = conv_at(ker, 3, 3, sample_sum, &s, 2 * x + (q & 1), 2 * y + (q >> 1));
                        a[q] = lrelu_bn(cv, m->conv_b[1][k], m->lrelu_w1[1][k], m->lrelu_w2[1][k], m->bn_b[1][k]);
                    }
                    P2[k][(long)y * p2w + x] = fmaxf(fmaxf(a[0], a[1]), fmaxf(a[2], a[3]));
                }
        }
    }
    /* L3: sums, two kw x kh convs per sum, lrelu_bn without pooling (:758-848) */
    float** A3 = (float**)malloc(sizeof(float*) * M3);
    for (int i = 0; i < M3 / 2; ++i) {
        sum_t s = { (const float* const*)P2, M2, i, p2w };
        for (int half = 0; half < 2; ++half) {
            int k = 2 * i + half;
            A3[k] = (float*)malloc(sizeof(float) * (size_t)c3 * r3);
            const float* ker = m->kern[2] + (long)k * m->kw[2] * m->kh[2];
            for (int y = 0; y < r3; ++y)
                for (int x = 0; x < c3; ++x) {
                    float cv = conv_at(ker, m->kw[2], m->kh[2], sample_sum, &s, x, y);
                    A3[k][(long)y * c3 + x] = lrelu_bn(cv, m->conv_b[2][k], m->lrelu_w1[2][k], m->lrelu_w2[2][k], m->bn_b[2][k]);
                }
        }
    }
    /* HL: mulCN_add_tanhW (cnnpp_cplusplus.cpp:646-672; loop cnn_simd_cntk.cpp:874-883)
     * OL: mulC / mulC1_add / tanhW (cnnpp_cplusplus.cpp:952-984, 673-690; loop :885-900) */
    int S = m->hl_scale, G = m->hl / m->hl_scale, N = m->connect, io = m->index_output;
    const int o0 = io < 0 ? 0 : io, o1 = io < 0 ? m->ol : io + 1;
    float* hv = (float*)malloc(sizeof(float) * m->hl);
    for (int y = 0; y < r3; ++y)
        for (int x = 0; x < c3; ++x) {
            long o = (long)y * c3 + x;
            for (int g = 0; g < G; ++g)
                for (int t = 0; t < S; ++t) {
                    int u = S * g + t;
                    float c = m->hl_b[u];
                    for (int j = 0; j < N; ++j) c += A3[g + j][o] * m->hl_w[(long)u * m->hl_nin + j];
                    c *= m->hl_tanh_w[g];
                    c = rtanh(c);
                    c = (c + 1.f) * 0.5f;               /* 0.5f*c + 0.5f, as folded by the build */
                    hv[u] = m->hl_bn_w[t] * c + m->hl_bn_b[t];
                }
            for (int oi = o0; oi < o1; ++oi) {
                float acc = 0.f;
                for (int g = 0; g < G; ++g)
                    for (int t = 0; t < S; ++t) {
                        float wv = m->ol_w[(long)oi * m->hl + t * G + g];
                        if (g == 0 && t == 0) acc = hv[0] * wv;
                        else acc = hv[S * g + t] * wv + acc;
                    }
                float c = acc + m->ol_b[oi];
                c *= m->ol_tanh_w;
                c = rtanh(c);
                /* 0.5f*c + 0.5f folded as in the hidden layer when af_scale is 0 */
                out[(long)(oi - o0) * c3 * r3 + o] = (m->af_scale == 0.f) ? (c + 1.f) * 0.5f : m->af_scale * c;
            }
        }
    free(hv);
    for (int k = 0; k < M1; ++k) free(P1[k]);
    for (int k = 0; k < M2; ++k) free(P2[k]);
    for (int k = 0; k < M3; ++k) free(A3[k]);
    free(P1); free(P2); free(A3);
    return 0;
}

/* ------------------------------------------------------------------------------ */
/* Image ops: image_proc.cpp, image_resize.cpp                                     */
/* ------------------------------------------------------------------------------ */

/* Img8uToImg32fGRAY (image_proc.cpp:299-317): 1 channel :99-102, BGR :195-201, BGRA :286-295 */
int oc_gray(const unsigned char* img, int w, int h, int ch, int step, float* out, int out_step)
{
    const float w0 = 0.114f, w1 = 0.587f, w2 = 0.299f;
    if (ch != 1 && ch != 3 && ch != 4) return -1;
    for (int y = 0; y < h; ++y) {
        const unsigned char* s = img + (long)y * step;
        float* d = out + (long)y * out_step;
        for (int x = 0; x < w; ++x) {
            if (ch == 1) { d[x] = (float)s[x]; continue; }
            const float B = (float)s[ch * x], G = (float)s[ch * x + 1], R = (float)s[ch * x + 2];
            d[x] = w0 * B + w1 * G + w2 * R;
        }
    }
    return 0;
}

/* ImageResizer::preprocessing (image_resize.cpp:46-122) */
typedef struct { unsigned* px; unsigned* py; float* fx; float* cx; float* fy; float* cy; } lut_t;

static void lut_build(lut_t* L, int sw, int sh, int dw, int dh)
{
    const unsigned QB = 12, Q = 1u << QB;
    const float Q2 = powf(2.0f, 24.0f);
    const unsigned sclx = ((unsigned)sw << QB) / (unsigned)dw + 1;
    const unsigned scly = ((unsigned)sh << QB) / (unsigned)dh + 1;
    L->px = (unsigned*)malloc(sizeof(unsigned) * dw);
    L->fx = (float*)malloc(sizeof(float) * dw);
    L->cx = (float*)malloc(sizeof(float) * dw);
    L->py = (unsigned*)malloc(sizeof(unsigned) * dh);
    L->fy = (float*)malloc(sizeof(float) * dh);
    L->cy = (float*)malloc(sizeof(float) * dh);
    for (unsigned iy = 0; iy < (unsigned)dh; ++iy) {
        unsigned y = iy * scly, py = y >> QB;
        float fy = (float)(y - (py << QB));
        float cy = (float)Q - fy;
        unsigned lim = (unsigned)(sh - 1);         /* MIN(py, h - 1) compares as unsigned */
        L->py[iy] = OC_MIN(py, lim);
        L->fy[iy] = fy / Q2;
        L->cy[iy] = cy / Q2;
    }
    for (unsigned ix = 0; ix < (unsigned)dw; ++ix) {
        unsigned x = ix * sclx, px = x >> QB;
        float fx = (float)(x - (px << QB));
        float cx = (float)Q - fx;
        unsigned lim = (unsigned)(sw - 2);
        L->px[ix] = OC_MIN(px, lim);
        L->fx[ix] = fx;
        L->cx[ix] = cx;
    }
}
static void lut_free(lut_t* L)
{
    free(L->px); free(L->py); free(L->fx); free(L->cx); free(L->fy); free(L->cy);
}

/* FastImageResize (32f) :564-584, NearestNeighborInterpolation :328-392 (scalar tail),
 * BilinearInterpolation :393-541 (scalar tail). type 0 = NN, 1 = bilinear. */
void oc_resize_f32(const float* src, int sw, int sh, int sstep,
                   float* dst, int dw, int dh, int dstep, int type)
{
    if (dw == sw && dh == sh) {
        for (int y = 0; y < dh; ++y) memcpy(dst + (long)y * dstep, src + (long)y * sstep, sizeof(float) * dw);
        return;
    }
    lut_t L;
    lut_build(&L, sw, sh, dw, dh);
    for (int iy = 0; iy < dh; ++iy) {
        unsigned py = L.py[iy];
        float* d = dst + (long)iy * dstep;
        if (type == 0) {
            const float* s0 = src + (long)py * sstep;
            for (int ix = 0; ix < dw; ++ix) d[ix] = s0[L.px[ix]];
        } else {
            if (py + 1 >= (unsigned)sh) py--;
            const float* s0 = src + (long)py * sstep;
            const float* s1 = src + (long)(py + 1) * sstep;
            const float fy = L.fy[iy], cy = L.cy[iy];
            for (int ix = 0; ix < dw; ++ix) {
                unsigned px = L.px[ix];
                const float fx = L.fx[ix], cx = L.cx[ix];
                const float p0 = s0[px], p1 = s0[px + 1], p2 = s1[px], p3 = s1[px + 1];
                d[ix] = (p0 * cx + p1 * fx) * cy + (p2 * cx + p3 * fx) * fy;
            }
        }
    }
    lut_free(&L);
}

/* FastImageResize (8u) :543-563; NN :160-197, bilinear :198-326 (scalar tails) */
void oc_resize_u8(const unsigned char* src, int sw, int sh, int sstep,
                  unsigned char* dst, int dw, int dh, int dstep, int type)
{
    if (dw == sw && dh == sh) {
        for (int y = 0; y < dh; ++y) memcpy(dst + (long)y * dstep, src + (long)y * sstep, (size_t)dw);
        return;
    }
    lut_t L;
    lut_build(&L, sw, sh, dw, dh);
    for (int iy = 0; iy < dh; ++iy) {
        unsigned py = L.py[iy];
        unsigned char* d = dst + (long)iy * dstep;
        if (type == 0) {
            const unsigned char* s0 = src + (long)py * sstep;
            for (int ix = 0; ix < dw; ++ix) d[ix] = s0[L.px[ix]];
        } else {
            if (py + 1 >= (unsigned)sh) py--;
            const unsigned char* s0 = src + (long)py * sstep;
            const unsigned char* s1 = src + (long)(py + 1) * sstep;
            const float fy = L.fy[iy], cy = L.cy[iy];
            for (int ix = 0; ix < dw; ++ix) {
                unsigned px = L.px[ix];
                const float fx = L.fx[ix], cx = L.cx[ix];
                const float p0 = s0[px], p1 = s0[px + 1], p2 = s1[px], p3 = s1[px + 1];
                const float v = (p0 * cx + p1 * fx) * cy + (p2 * cx + p3 * fx) * fy;
                d[ix] = (unsigned char)v;
            }
        }
    }
    lut_free(&L);
}

/* equalizeImage (image_proc.cpp:720-800, scalar branch) */
void oc_equalize(unsigned char* img, int w, int h, int step)
{
    int hist[256], cf[256];
    memset(hist, 0, sizeof(hist));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) hist[img[(long)y * step + x]]++;
    cf[0] = hist[0];
    for (int k = 1; k < 256; ++k) cf[k] = cf[k - 1] + hist[k];
    const float bmin = (float)cf[0], bmax = (float)cf[255];
    if (bmin == bmax) return;
    const float delt = 255.f / (bmax - bmin);
    for (int k = 0; k < 256; ++k) cf[k] = (int)(((float)cf[k] - bmin) * delt + 0.5f);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            unsigned char* p = img + (long)y * step + x;
            *p = (unsigned char)cf[*p];
        }
}

/* CPUCheckDetect image preparation (cnn_detector_v3.cpp:1558-1644): NN resize to 49x51
 * (check resizer, :829), FloatToUChar (image_proc.cpp:357-368), equalizeImage, then
 * UCharToFloat with offset 2 (:416-427) and UCharToFloat_inv (:478-490) or offset 0. */
void oc_check_prep(const float* crop, int cw, int ch, int cstep, int equalize, int reflection,
                   unsigned char* u8_out, float* s2_out, float* s3_out, float* f32_out)
{
    float r32[49 * 51];
    oc_resize_f32(crop, cw, ch, cstep, r32, 49, 51, 49, 0);
    if (f32_out) memcpy(f32_out, r32, sizeof(r32));
    for (int i = 0; i < 49 * 51; ++i) u8_out[i] = (unsigned char)(r32[i] + 0.5f);
    if (equalize) oc_equalize(u8_out, 49, 51, 49);
    for (int y = 0; y < 51; ++y)
        for (int x = 0; x < 47; ++x) {
            s2_out[y * 47 + x] = (float)u8_out[y * 49 + x + 2];
            s3_out[y * 47 + x] = reflection ? (float)u8_out[y * 49 + 46 - x] : (float)u8_out[y * 49 + x];
        }
}

/* ------------------------------------------------------------------------------ */
/* libstdc++ std::sort restated (GCC 11.4): introsort with median-of-three pivot,   */
/* heapsort fallback, final insertion sort with threshold 16.                      */
/* ------------------------------------------------------------------------------ */

typedef struct { int* a; const float* key; } srt_t;
/* comp(a, b) = b.scale < a.scale   (cnn_detector_v3.cpp:2750) */
static int s_comp(const srt_t* s, int va, int vb) { return s->key[vb] < s->key[va]; }
static void s_swap(int* a, long i, long j) { int t = a[i]; a[i] = a[j]; a[j] = t; }

static void s_push_heap(srt_t* s, long first, long hole, long top, int value)
{
    long parent = (hole - 1) / 2;
    while (hole > top && s_comp(s, s->a[first + parent], value)) {
        s->a[first + hole] = s->a[first + parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    s->a[first + hole] = value;
}
static void s_adjust_heap(srt_t* s, long first, long hole, long len, int value)
{
    const long top = hole;
    long second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (s_comp(s, s->a[first + second], s->a[first + second - 1])) second--;
        s->a[first + hole] = s->a[first + second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        s->a[first + hole] = s->a[first + second - 1];
        hole = second - 1;
    }
    s_push_heap(s, first, hole, top, value);
}
static void s_heap_sort(srt_t* s, long first, long last)
{
    long len = last - first;
    if (len >= 2) {                                   /* make_heap */
        long parent = (len - 2) / 2;
        for (;;) {
            int v = s->a[first + parent];
            s_adjust_heap(s, first, parent, len, v);
            if (parent == 0) break;
            parent--;
        }
    }
    while (last - first > 1) {                        /* sort_heap */
        --last;
        int v = s->a[last];
        s->a[last] = s->a[first];
        s_adjust_heap(s, first, 0, last - first, v);
    }
}
static void s_median_to_first(srt_t* s, long result, long a, long b, long c)
{
    int* v = s->a;
    if (s_comp(s, v[a], v[b])) {
        if (s_comp(s, v[b], v[c])) s_swap(v, result, b);
        else if (s_comp(s, v[a], v[c])) s_swap(v, result, c);
        else s_swap(v, result, a);
    } else if (s_comp(s, v[a], v[c])) s_swap(v, result, a);
    else if (s_comp(s, v[b], v[c])) s_swap(v, result, c);
    else s_swap(v, result, b);
}
static long s_unguarded_partition(srt_t* s, long first, long last, long pivot)
{
    for (;;) {
        while (s_comp(s, s->a[first], s->a[pivot])) ++first;
        --last;
        while (s_comp(s, s->a[pivot], s->a[last])) --last;
        if (!(first < last)) return first;
        s_swap(s->a, first, last);
        ++first;
    }
}
static void s_introsort_loop(srt_t* s, long first, long last, long depth)
{
    while (last - first > 16) {
        if (depth == 0) { s_heap_sort(s, first, last); return; }
        --depth;
        long mid = first + (last - first) / 2;
        s_median_to_first(s, first, first + 1, mid, last - 1);
        long cut = s_unguarded_partition(s, first + 1, last, first);
        s_introsort_loop(s, cut, last, depth);
        last = cut;
    }
}
static void s_unguarded_linear_insert(srt_t* s, long last)
{
    int val = s->a[last];
    long next = last - 1;
    while (s_comp(s, val, s->a[next])) {
        s->a[last] = s->a[next];
        last = next;
        --next;
    }
    s->a[last] = val;
}
static void s_insertion_sort(srt_t* s, long first, long last)
{
    if (first == last) return;
    for (long i = first + 1; i != last; ++i) {
        if (s_comp(s, s->a[i], s->a[first])) {
            int val = s->a[i];
            memmove(s->a + first + 1, s->a + first, sizeof(int) * (size_t)(i - first));
            s->a[first] = val;
        } else s_unguarded_linear_insert(s, i);
    }
}
void oc_sort_desc(int* perm, const float* key, int n)
{
    if (n <= 1) return;
    srt_t s = { perm, key };
    long lg = 63 - __builtin_clzl((unsigned long)n);
    s_introsort_loop(&s, 0, n, lg * 2);
    if (n > 16) {
        s_insertion_sort(&s, 0, 16);
        for (long i = 16; i != n; ++i) s_unguarded_linear_insert(&s, i);
    } else s_insertion_sort(&s, 0, n);
}

/* ------------------------------------------------------------------------------ */
/* Detector: cnn_detector_v3.cpp                                                   */
/* ------------------------------------------------------------------------------ */

typedef struct { int x, y, w, h, x2, y2; } rect_t;            /* type.h:107-137 */
/* Detection (cnn_detector_v3.h:185-216); fd holds the facial_data entries as
 * {gender, smile, glasses} triples (nfd of them, heap-owned by the detection). */
typedef struct { rect_t r; float score, scale; int knn, num_detect, checked; int nfd; int* fd; } det_t;
typedef struct { det_t* v; int n, cap; } dvec_t;

static rect_t mk_rect(int x, int y, int w, int h)
{
    rect_t r = { x, y, w, h, x + w, y + h };
    return r;
}
static void dpush(dvec_t* d, rect_t r, float score, float scale, int knn)
{
    if (d->n == d->cap) {
        d->cap = d->cap ? 2 * d->cap : 64;
        d->v = (det_t*)realloc(d->v, sizeof(det_t) * (size_t)d->cap);
    }
    det_t e = { r, score, scale, knn, 0, 0, 0, NULL };
    d->v[d->n++] = e;
}
static int* fd_dup(const int* fd, int n)
{
    if (n == 0) return NULL;
    int* a = (int*)malloc(sizeof(int) * 3 * (size_t)n);
    memcpy(a, fd, sizeof(int) * 3 * (size_t)n);
    return a;
}
static void fd_append(det_t* d, const int* fd, int n)
{
    if (n == 0) return;
    d->fd = (int*)realloc(d->fd, sizeof(int) * 3 * (size_t)(d->nfd + n));
    memcpy(d->fd + 3 * d->nfd, fd, sizeof(int) * 3 * (size_t)n);
    d->nfd += n;
}
static void dvec_free_fd(dvec_t* d)
{
    for (int i = 0; i < d->n; ++i) { free(d->v[i].fd); d->v[i].fd = NULL; d->v[i].nfd = 0; }
}
/* Rect::overlap (type.h:154-173): intersection over the bounding box of both */
static float rect_overlap(const rect_t* a, const rect_t* b)
{
    if (a->x > b->x2 || b->x > a->x2 || a->y > b->y2 || b->y > a->y2) return 0.f;
    int X0 = OC_MIN(a->x, b->x), X1 = OC_MAX(a->x, b->x), X2 = OC_MIN(a->x2, b->x2), X3 = OC_MAX(a->x2, b->x2);
    int Y0 = OC_MIN(a->y, b->y), Y1 = OC_MAX(a->y, b->y), Y2 = OC_MIN(a->y2, b->y2), Y3 = OC_MAX(a->y2, b->y2);
    float su = (float)((X3 - X0) * (Y3 - Y0));
    float si = (float)((X2 - X1) * (Y2 - Y1));
    return si / su;
}
/* Rect::intersects (type.h:140-153) */
static int rect_intersects(const rect_t* a, const rect_t* b)
{
    if (a->x > b->x2 || b->x > a->x2 || a->y > b->y2 || b->y > a->y2) return 0;
    int X1 = OC_MAX(a->x, b->x), X2 = OC_MIN(a->x2, b->x2);
    int Y1 = OC_MAX(a->y, b->y), Y2 = OC_MIN(a->y2, b->y2);
    return (X2 - X1) * (Y2 - Y1);
}
/* Rect::operator*(float) (type.h:185-192) */
static rect_t rect_shrink(const rect_t* r, float scale)
{
    scale = 1.f - scale;
    int sx = (int)((float)r->w * scale), sy = (int)((float)r->h * scale);
    return mk_rect(r->x + (sx >> 1), r->y + (sy >> 1), r->w - sx, r->h - sy);
}

/* Effective detector settings after FaceDetector::Init's conversion
 * (CompactCNNLibAPI_v2.cpp:89-138) and CNNDetector::Init (cnn_detector_v3.cpp:922-964). */
typedef struct {
    int type_check, min_num_detect, adapt_min_neighbors, drop_detect, merger, double_check;
    int equalize, reflection, detect_mode, min_neighbors;
    float t1, t2, t3;
} eff_t;

static eff_t effective(const oc_params* p)
{
    eff_t e;
    e.type_check = 1; e.min_num_detect = 1; e.adapt_min_neighbors = 0;
    e.drop_detect = p->drop_detect; e.equalize = p->equalize; e.reflection = p->reflection;
    e.double_check = 1;
    e.detect_mode = p->detect_mode;
    e.merger = p->detect_mode != 0;
    e.min_neighbors = p->min_neighbors;
    e.t1 = p->t1; e.t2 = p->t2; e.t3 = p->t3;
    if (p->precision == 0) {                       /* :923-937 */
        e.type_check = 1; e.equalize = 1; e.reflection = 1; e.adapt_min_neighbors = 1;
        e.double_check = 1; e.drop_detect = 0; e.min_num_detect = 1; e.merger = 1;
    }
    switch (p->precision) {                        /* :939-959 */
    case 1: e.type_check = 0; e.drop_detect = 0; break;
    case 2: e.type_check = 1; break;
    case 3: e.type_check = 2; break;
    case 4: e.type_check = 2; e.min_num_detect = 2; break;
    default: break;
    }
    if (e.detect_mode == 0 || p->facial_analysis) e.drop_detect = 0;      /* :961-964 */
    return e;
}

/* InitScales (cnn_detector_v3.cpp:289-346). pre_scale receives the gray_image_only
 * pre-resize factor (0 when absent); returns the scale count or -1. */
static int init_scales(const oc_params* p, int pattern_h, int gray_only, float* scales, int cap, float* pre_scale)
{
    float max_scale = 1.f;
    *pre_scale = 0.f;
    if (p->min_face > 0) {
        max_scale = (float)pattern_h / (float)p->min_face;
        max_scale = OC_MIN(3.f, max_scale);
        if (gray_only && max_scale < 0.8f) {
            *pre_scale = max_scale;
            max_scale = 1.f;
        }
    }
    float min_scale;
    if (p->max_face > 0) min_scale = (float)pattern_h / (float)p->max_face;
    else min_scale = (float)pattern_h / (float)p->max_h;
    min_scale = OC_MIN(max_scale, min_scale);
    float sf = p->scale_factor;
    if (sf < 1.f) sf = 1.1f;
    int n = 0;
    float scale = max_scale;
    if (sf != 1.f) {
        while (scale > min_scale) {
            if (n >= cap) return -1;
            scales[n++] = scale;
            scale /= sf;
        }
    } else scales[n++] = scale;
    return n > 0 ? n : -1;
}

/* Size::operator*(float) (type.h:63-66) */
static void size_mul(int w, int h, float s, int* ow, int* oh)
{
    *ow = (int)((float)w * s);
    *oh = (int)((float)h * s);
}

int oc_plan(const oc_params* p, int pattern_h, int w, int h, float* scales, int* n_scales,
            int* lvl_w, int* lvl_h, int* n_run)
{
    float pre;
    int n = init_scales(p, pattern_h, 0, scales, 512, &pre);
    if (n < 0) return -1;
    *n_scales = n;
    int run = 0;
    for (int scl = n - 1; scl >= 0; --scl) {
        size_mul(w, h, scales[scl], &lvl_w[scl], &lvl_h[scl]);
        if (!(lvl_w[scl] < 47 || lvl_h[scl] < 51)) run++;
    }
    *n_run = run;
    return 0;
}

typedef struct {
    const oc_model *m1, *m2, *m3, *mfa;
    const oc_params* p;
    eff_t e;
    const float* gray; int gw, gh, gstep;
    dvec_t rects;
    oc_trace* tr;
} dctx_t;

/* The stage-2/3 evaluation of one crop (cnn_detector_v3.cpp:1485-1818, from the crop
 * geometry onwards). Returns flag_object; fills k1, m1, k2, m2. */
static int check_eval(const dctx_t* c, float scale, int px, int py, int mod,
                      int* k1o, float* m1o, int* k2o, float* m2o, float* crop49, int* geo)
{
    const float inv = 1.f / scale;
    int x = (int)((float)(px - 6) * inv);          /* (ext 11 + x 2) / 2 = 6 */
    int y = (int)((float)(py - 5) * inv);          /* 11 / 2 = 5 */
    int width = (int)(36.f * inv);                 /* 23 + 11 + 2 */
    int height = (int)(38.f * inv);                /* 27 + 11 */
    if (mod == 1) {                                /* :1535-1542 */
        /* x += w*k/2; w -= w*k with k = 0.15f. The build folds w - w*k into
         * w * (1.f - 0.15f), a float constant equal to 0.85f. */
        const float k = 0.15f;
        x = (int)((float)x + (float)width * k / 2.f);
        width = (int)((float)width * 0.85f);
        y = (int)((float)y + (float)height * k / 2.f);
        height = (int)((float)height * 0.85f);
    }
    int rx = OC_MIN(OC_MAX(x, 0), c->gw);           /* :1553-1556 */
    int ry = OC_MIN(OC_MAX(y, 0), c->gh);
    int rcols = OC_MIN(OC_MAX(x + width, 0), c->gw) - rx;
    int rrows = OC_MIN(OC_MAX(y + height, 0), c->gh) - ry;
    unsigned char u8[49 * 51];
    float in2[47 * 51], in3[47 * 51], out[64];
    oc_check_prep(c->gray + (long)ry * c->gstep + rx, rcols, rrows, c->gstep,
                  c->e.equalize, c->e.reflection, u8, in2, in3, crop49);
    if (geo) { geo[0] = rx; geo[1] = ry; geo[2] = rcols; geo[3] = rrows; }
    const int mn = (c->e.adapt_min_neighbors && scale > 0.7f && c->e.min_neighbors > 1)
                   ? c->e.min_neighbors - 1 : c->e.min_neighbors;
    float max1 = -10.f, max2 = -10.f;
    int kn1 = 0, kn2 = 0;
    for (int icx = -1; icx <= 1; icx += 2) {
        const oc_model* m = icx == -1 ? c->m2 : c->m3;
        int ow, oh;
        oc_output_size(m, 47, 51, &ow, &oh);
        if (ow * oh > 64 || oc_cnn_forward(m, icx == -1 ? in2 : in3, 47, 51, 47, out) < 0) break;
        for (int i = 0; i < ow * oh; ++i) {
            const float d = out[i];
            if (icx == -1) { max1 = OC_MAX(max1, d); if (d > c->e.t2) kn1++; }
            else { max2 = OC_MAX(max2, d); if (d > c->e.t3) kn2++; }
        }
        if (c->e.type_check == 0 && kn1 >= mn) break;   /* :1704-1706 */
        if (c->e.type_check > 0 && kn1 <= 0) break;
        if (c->e.type_check == 2 && kn1 < mn) break;
    }
    int flag = 0;
    switch (c->e.type_check) {                          /* :1709-1725 */
    case 0: flag = (kn1 >= mn || kn2 >= mn); break;
    case 1: flag = ((kn1 >= mn && kn2 > 0) || (kn1 > 0 && kn2 >= mn)); break;
    case 2: flag = (kn1 >= mn && kn2 >= mn); break;
    }
    *k1o = kn1; *m1o = max1; *k2o = kn2; *m2o = max2;
    return flag;
}

static float check_score(float score0, int k1, float m1, int k2, float m2)
{
    /* :1758 */
    return score0 + (float)k1 * OC_MAX(-1.7159f, m1) + (float)k2 * OC_MAX(-1.7159f, m2);
}

/* Facial analysis of an accepted window (cnn_detector_v3.cpp:1731-1756): the landmarks
 * network on the 49x51 FP32 crop; five (x, y) landmarks scaled by the crop size, the
 * window rejected when all five fall in the crop's top-left quadrant, and gender /
 * smile / glasses rounded from outputs 10-12. Returns 0 when rejected, else 1 with
 * fd = {gender, smile, glasses}. */
static int facial_analysis(const oc_model* mfa, const float* crop49, int rcols, int rrows, int* fd)
{
    float o[16];
    if (mfa->ol > 16 || oc_cnn_forward(mfa, crop49, 49, 51, 49, o) < 0) return 1;
    int zero = 0;
    for (int t = 0; t < 5; ++t) {
        const int lx = (int)(o[2 * t] * (float)rcols);
        const int ly = (int)(o[2 * t + 1] * (float)rrows);
        if (lx < (rcols >> 1) && ly < (rrows >> 1)) zero++;
    }
    if (zero == 5) return 0;
    fd[0] = (int)(o[10] + 0.5f);
    fd[1] = (int)(o[11] + 0.5f);
    fd[2] = (int)(o[12] + 0.5f);
    return 1;
}

/* CPUCheckDetect with the drop check (cnn_detector_v3.cpp:1485-1818). */
static void cpu_check_detect(dctx_t* c, int rect_size, int px, int py, float score0, float scale,
                             int mod, int cand)
{
    const float inv = 1.f / scale;
    rect_t nr = mk_rect((int)((float)px * inv), (int)((float)py * inv), (int)(23.f * inv), (int)(27.f * inv));
    if (c->e.drop_detect && mod == 0) {                 /* :1513-1533 */
        for (int i = rect_size; i < c->rects.n; ++i) {
            if (rect_overlap(&nr, &c->rects.v[i].r) > 0.5f) {
                det_t* it = &c->rects.v[i];
                dpush(&c->rects, nr, it->score, scale, it->knn);
                if (cand >= 0) c->tr->c_status[cand] = 1;
                return;
            }
        }
    }
    int k1, k2; float m1, m2;
    float crop49[49 * 51];
    int geo[4];
    int flag = check_eval(c, scale, px, py, mod, &k1, &m1, &k2, &m2, crop49, geo);
    if (cand >= 0) {
        c->tr->c_mod[cand] = mod; c->tr->c_k1[cand] = k1; c->tr->c_m1[cand] = m1;
        c->tr->c_k2[cand] = k2; c->tr->c_m2[cand] = m2;
        c->tr->c_status[cand] = flag ? 2 : 3;
    }
    if (flag) {
        int fd[3];
        if (c->mfa) {
            if (!facial_analysis(c->mfa, crop49, geo[2], geo[3], fd)) {
                if (cand >= 0) c->tr->c_status[cand] = 4;
                return;                                /* :1749: no retry */
            }
        }
        dpush(&c->rects, nr, check_score(score0, k1, m1, k2, m2), scale, OC_MIN(k1, k2));
        if (c->mfa) fd_append(&c->rects.v[c->rects.n - 1], fd, 1);   /* :1762-1765 */
    } else if (c->e.double_check && mod == 0) {
        cpu_check_detect(c, rect_size, px, py, score0, scale, 1, cand);
    }
}

/* DropDetection (cnn_detector_v3.cpp:1820-1841), in == out == the detection list */
static int drop_detection(dctx_t* c, const rect_t* nr, float scale)
{
    rect_t exp = rect_shrink(nr, 0.7f);
    int n = c->rects.n;
    for (int i = 0; i < n; ++i) {
        det_t* it = &c->rects.v[i];
        const float ov = rect_overlap(nr, &it->r);
        if (ov == 0.f) continue;
        if (ov > 0.5f) {
            dpush(&c->rects, *nr, it->score, scale, it->knn);
            return 1;
        }
        if (rect_intersects(&it->r, &exp) == exp.w * exp.h) return 1;
    }
    return 0;
}

/* RunCheckDetect for one level (cnn_detector_v3.cpp:1842-2011, CPU branch) */
static void run_check_detect(dctx_t* c, int scl, float scale, const float* resp, int rw, int rh)
{
    const float inv = 1.f / scale;
    int np = 0, capp = 64;
    int* pts = (int*)malloc(sizeof(int) * 3 * capp);
    float* psc = (float*)malloc(sizeof(float) * capp);
    for (int j = 0; j < rh; ++j)
        for (int i = 0; i < rw; ++i) {
            const float score = resp[(long)j * rw + i];
            if (!(score > c->e.t1)) continue;
            const int px = i * 4, py = j * 4;            /* shift_pattern = 4 */
            int cand = -1;
            oc_trace* tr = c->tr;
            if (tr && tr->n_cand < tr->cap) {
                cand = tr->n_cand++;
                tr->c_scl[cand] = scl; tr->c_x[cand] = i; tr->c_y[cand] = j; tr->c_score[cand] = score;
                tr->c_status[cand] = 0; tr->c_mod[cand] = -1;
                tr->c_k1[cand] = 0; tr->c_k2[cand] = 0; tr->c_m1[cand] = 0.f; tr->c_m2[cand] = 0.f;
            }
            if (c->e.drop_detect) {
                rect_t nr = mk_rect((int)((float)px * inv), (int)((float)py * inv), (int)(23.f * inv), (int)(27.f * inv));
                if (drop_detection(c, &nr, scale)) continue;
            }
            if (np == capp) {
                capp *= 2;
                pts = (int*)realloc(pts, sizeof(int) * 3 * capp);
                psc = (float*)realloc(psc, sizeof(float) * capp);
            }
            pts[3 * np] = px; pts[3 * np + 1] = py; pts[3 * np + 2] = cand; psc[np] = score;
            np++;
            if (cand >= 0) tr->c_status[cand] = 3;
        }
    if (c->e.detect_mode != 0 && np > 0) {
        const int rect_size = c->rects.n;
        for (int k = 0; k < np; ++k)
            cpu_check_detect(c, rect_size, pts[3 * k], pts[3 * k + 1], psc[k], scale, 0, pts[3 * k + 2]);
    } else {
        for (int k = 0; k < np; ++k) {
            rect_t r = mk_rect((int)((float)pts[3 * k] * inv), (int)((float)pts[3 * k + 1] * inv),
                               (int)(23.f * inv), (int)(27.f * inv));
            dpush(&c->rects, r, psc[k], scale, 0);
            if (pts[3 * k + 2] >= 0) c->tr->c_status[pts[3 * k + 2]] = 2;
        }
    }
    free(pts); free(psc);
}

/* Merger (cnn_detector_v3.cpp:2814-2936) */
static void merger(dvec_t* out, dvec_t* in, float threshold, int del)
{
    for (int a = 0; a < in->n; ++a) {
        if (in->v[a].checked) continue;
        det_t mx = in->v[a], mn = in->v[a];
        mx.fd = fd_dup(in->v[a].fd, in->v[a].nfd);             /* the copy owns its facial_data */
        mn.fd = NULL; mn.nfd = 0;
        for (int b = a + 1; b < in->n; ++b) {
            det_t* it2 = &in->v[b];
            if (it2->checked) continue;
            int X[4] = { mx.r.x, mx.r.x2, it2->r.x, it2->r.x2 };
            int Y[4] = { mx.r.y, mx.r.y2, it2->r.y, it2->r.y2 };
            if (X[0] >= X[3] || X[1] <= X[2] || Y[0] >= Y[3] || Y[1] <= Y[2]) continue;
            const float S1 = (float)((X[1] - X[0]) * (Y[1] - Y[0]));
            const float S2 = (float)((X[3] - X[2]) * (Y[3] - Y[2]));
            for (int i = 1; i < 4; ++i) for (int j = i; j > 0 && X[j - 1] > X[j]; --j) { int t = X[j]; X[j] = X[j - 1]; X[j - 1] = t; }
            for (int i = 1; i < 4; ++i) for (int j = i; j > 0 && Y[j - 1] > Y[j]; --j) { int t = Y[j]; Y[j] = Y[j - 1]; Y[j - 1] = t; }
            const float Su = (float)((X[3] - X[0]) * (Y[3] - Y[0]));
            const float Si = (float)((X[2] - X[1]) * (Y[2] - Y[1]));
            const float r1 = Si / S1, r2 = Si / S2, r3 = Si / Su;
            float ratio = OC_MIN(r1, r2);
            const int CX1 = X[0] + ((X[3] - X[0]) >> 1), CY1 = Y[0] + ((Y[3] - Y[0]) >> 1);
            const int CX2 = X[1] + ((X[2] - X[1]) >> 1), CY2 = Y[1] + ((Y[2] - Y[1]) >> 1);
            const int thC = OC_MIN((X[3] - X[0]), (X[2] - X[1])) >> 1;
            if (r3 > 0.75f * threshold && OC_MAX(r1, r2) > 0.7f && abs(CX1 - CX2) < thC && abs(CY1 - CY2) < thC)
                ratio = 1.f;
            if (ratio > threshold) {
                if (del) {
                    if (S1 > S2) { it2->checked = 1; continue; }
                    mn.r.x2 = 0;
                    break;
                }
                mx.r.x = OC_MIN(mx.r.x, X[0]);
                mx.r.y = OC_MIN(mx.r.y, Y[0]);
                mx.r.x2 = OC_MAX(mx.r.x2, X[3]);
                mx.r.y2 = OC_MAX(mx.r.y2, Y[3]);
                mn.r.x = OC_MAX(mn.r.x, X[1]);
                mn.r.y = OC_MAX(mn.r.y, Y[1]);
                mn.r.x2 = OC_MIN(mn.r.x2, X[2]);
                mn.r.y2 = OC_MIN(mn.r.y2, Y[2]);
                mx.score = OC_MAX(mx.score, it2->score);
                mx.knn = OC_MAX(mx.knn, it2->knn);
                mx.scale = OC_MIN(mx.scale, it2->scale);
                fd_append(&mx, it2->fd, it2->nfd);             /* :2911-2914 */
                mx.num_detect++;
                it2->checked = 1;
            }
        }
        if (del && mn.r.x2 == 0) { free(mx.fd); continue; }
        mx.r.x = (mn.r.x + mx.r.x) >> 1;
        mx.r.y = (mn.r.y + mx.r.y) >> 1;
        mx.r.x2 = (mn.r.x2 + mx.r.x2) >> 1;
        mx.r.y2 = (mn.r.y2 + mx.r.y2) >> 1;
        mx.r.w = mx.r.x2 - mx.r.x;
        mx.r.h = mx.r.y2 - mx.r.y;
        mx.checked = 0;
        if (out->n == out->cap) {
            out->cap = out->cap ? 2 * out->cap : 64;
            out->v = (det_t*)realloc(out->v, sizeof(det_t) * (size_t)out->cap);
        }
        out->v[out->n++] = mx;
    }
}

int oc_check_candidate(const oc_model* m2, const oc_model* m3, const oc_params* p,
                       const float* gray, int gw, int gh, int gstep,
                       float scale, int px, int py, float* res)
{
    dctx_t c;
    memset(&c, 0, sizeof(c));
    c.m2 = m2; c.m3 = m3; c.p = p; c.e = effective(p);
    c.gray = gray; c.gw = gw; c.gh = gh; c.gstep = gstep;
    int k1, k2; float f1, f2;
    int mod = 0;
    int flag = check_eval(&c, scale, px, py, 0, &k1, &f1, &k2, &f2, NULL, NULL);
    if (!flag && c.e.double_check) { mod = 1; flag = check_eval(&c, scale, px, py, 1, &k1, &f1, &k2, &f2, NULL, NULL); }
    res[0] = (float)flag; res[1] = (float)mod; res[2] = (float)k1; res[3] = (float)k2;
    res[4] = f1; res[5] = f2;
    const float inv = 1.f / scale;
    res[6] = (float)(int)((float)px * inv); res[7] = (float)(int)((float)py * inv);
    res[8] = (float)(int)(23.f * inv); res[9] = (float)(int)(27.f * inv);
    return 0;
}

/* Median of one facial_data key as FaceDetector::Detect reports it
 * (CompactCNNLibAPI_v2.cpp:296-319): sorted descending, the element at size/2. */
static int fd_median(const det_t* d, int key)
{
    int* v = (int*)malloc(sizeof(int) * (size_t)d->nfd);
    for (int i = 0; i < d->nfd; ++i) v[i] = d->fd[3 * i + key];
    for (int i = 1; i < d->nfd; ++i)
        for (int j = i; j > 0 && v[j - 1] < v[j]; --j) { int t = v[j]; v[j] = v[j - 1]; v[j - 1] = t; }
    int r = v[d->nfd / 2];
    free(v);
    return r;
}

void oc_state_init(oc_state* s, const oc_params* p)
{
    s->max_w = p->max_w;
    s->max_h = p->max_h;
    s->num_scales = -1;
    s->gray_only = 1;
}

int oc_detect(const oc_model* m1, const oc_model* m2, const oc_model* m3, const oc_model* mfa,
              const oc_params* p, const unsigned char* img, int w, int h, int ch, int step,
              oc_face* out, int cap, oc_trace* trace)
{
    oc_state s;
    oc_state_init(&s, p);
    return oc_detect_state(m1, m2, m3, mfa, p, &s, img, w, h, ch, step, out, cap, trace);
}

int oc_detect_state(const oc_model* m1, const oc_model* m2, const oc_model* m3, const oc_model* mfa,
                    const oc_params* p, oc_state* st, const unsigned char* img, int w, int h, int ch, int step,
                    oc_face* out, int cap, oc_trace* trace)
{
    dctx_t c;
    memset(&c, 0, sizeof(c));
    c.m1 = m1; c.m2 = m2; c.m3 = m3; c.p = p; c.e = effective(p); c.tr = trace;
    c.mfa = p->facial_analysis ? mfa : NULL;
    if (trace) { trace->n_cand = 0; trace->n_levels = 0; trace->n_scales = 0; trace->n_raw = 0; }
    if (ch != 1 && ch != 3 && ch != 4) return -1;
    /* ext_pattern_size_cd (cnn_detector_v3.cpp:762-764, 879-882) */
    const int ext_w = c.e.detect_mode != 0 ? m2->min_w + 11 : m1->min_w;
    const int ext_h = c.e.detect_mode != 0 ? m2->min_h + 11 : m1->min_h;
    if (w < ext_w || h < ext_h) return 0;                          /* :2494-2497 */
    if (w > st->max_w || h > st->max_h) {                          /* :2499-2513: Clear, Init */
        st->max_w = w;
        st->max_h = h;
        if (st->gray_only && ch != 1) st->gray_only = 0;
        st->num_scales = -1;
    } else if (w != st->max_w || h != st->max_h) {
        st->num_scales = -1;                                        /* :2514-2518 */
    }
    oc_params pp = *p;
    pp.max_w = st->max_w;
    pp.max_h = st->max_h;
    float scales[512], pre = 0.f;
    int n_scales = init_scales(&pp, m1->min_h, st->gray_only, scales, 512, &pre);
    if (n_scales < 0) return -1;
    if (pre > 0.f && ch != 1) {
        /* a colour image turns gray_image_only off, Clears and detects again on a fresh
         * Init (cnn_detector_v3.cpp:2540-2546) */
        st->gray_only = 0;
        st->num_scales = -1;
        n_scales = init_scales(&pp, m1->min_h, 0, scales, 512, &pre);
        if (n_scales < 0) return -1;
    }
    /* the scales RunCPUDetect visits this time, and the count it leaves for the next call */
    const int ncap = st->num_scales < 0 ? n_scales : OC_MIN(st->num_scales, n_scales);
    int next_scales = ncap;

    const unsigned char* src = img;
    int sw = w, sh = h, sstep = step;
    unsigned char* pre_img = NULL;
    if (pre > 0.f) {                                                /* :2531-2539 */
        size_mul(w, h, pre, &sw, &sh);
        pre_img = (unsigned char*)malloc((size_t)sw * sh);
        oc_resize_u8(img, w, h, step, pre_img, sw, sh, sw, 1);
        src = pre_img; sstep = sw;
    }
    float* gray = (float*)malloc(sizeof(float) * (size_t)sw * sh);
    oc_gray(src, sw, sh, ch, sstep, gray, sw);                       /* :2631-2652 */
    c.gray = gray; c.gw = sw; c.gh = sh; c.gstep = sw;
    if (trace) { trace->n_scales = n_scales; for (int i = 0; i < n_scales; ++i) trace->scales[i] = scales[i]; }

    /* RunCPUDetect (:2063-2233): levels scl_max -> 0 */
    float* resp[512];
    int rw[512], rh[512], ran[512];
    for (int scl = n_scales - 1; scl >= 0; --scl) {
        resp[scl] = NULL; ran[scl] = 0;
        if (scl >= ncap) continue;
        int lw, lh;
        size_mul(sw, sh, scales[scl], &lw, &lh);
        if (lw < ext_w || lh < ext_h) {                             /* :2083-2088 */
            next_scales = OC_MIN(scl, next_scales);
            continue;
        }
        float* lvl = (float*)malloc(sizeof(float) * (size_t)lw * lh);
        int mode = scales[scl] != 1.f ? (scales[scl] > 0.7f ? 0 : 1) : 2;   /* :2093-2104 */
        if (mode == 2) oc_resize_f32(gray, sw, sh, sw, lvl, lw, lh, lw, 0);
        else oc_resize_f32(gray, sw, sh, sw, lvl, lw, lh, lw, mode);
        oc_output_size(m1, lw, lh, &rw[scl], &rh[scl]);
        resp[scl] = (float*)malloc(sizeof(float) * (size_t)(rw[scl] > 0 ? rw[scl] : 1) * (rh[scl] > 0 ? rh[scl] : 1));
        if (rw[scl] > 0 && rh[scl] > 0 && oc_cnn_forward(m1, lvl, lw, lh, lw, resp[scl]) == 0) ran[scl] = 1;
        else { rw[scl] = 0; rh[scl] = 0; ran[scl] = 1; }
        if (trace && trace->n_levels < 512) {
            int L = trace->n_levels++;
            trace->lvl_scl[L] = scl; trace->lvl_w[L] = lw; trace->lvl_h[L] = lh;
            trace->lvl_rw[L] = rw[scl]; trace->lvl_rh[L] = rh[scl]; trace->lvl_mode[L] = mode;
        }
        free(lvl);
    }
    st->num_scales = next_scales;
    /* RunCheckDetectAsync (:2012-2061): levels scl_max -> 0 */
    for (int scl = n_scales - 1; scl >= 0; --scl)
        if (ran[scl]) run_check_detect(&c, scl, scales[scl], resp[scl], rw[scl], rh[scl]);
    for (int scl = 0; scl < n_scales; ++scl) free(resp[scl]);
    if (trace) trace->n_raw = c.rects.n;

    /* Post-processing (:2745-2810) */
    int nout = 0;
    dvec_t raw = { 0, 0, 0 }, tmp = { 0, 0, 0 }, fin = { 0, 0, 0 };
    if (c.rects.n > 0) {
        raw.n = raw.cap = c.rects.n;
        raw.v = (det_t*)malloc(sizeof(det_t) * (size_t)raw.n);
        for (int i = 0; i < raw.n; ++i) raw.v[i] = c.rects.v[raw.n - 1 - i];    /* reverse_copy */
        int* perm = (int*)malloc(sizeof(int) * (size_t)raw.n);
        float* key = (float*)malloc(sizeof(float) * (size_t)raw.n);
        for (int i = 0; i < raw.n; ++i) { perm[i] = i; key[i] = raw.v[i].scale; }
        oc_sort_desc(perm, key, raw.n);
        det_t* sorted = (det_t*)malloc(sizeof(det_t) * (size_t)raw.n);
        for (int i = 0; i < raw.n; ++i) sorted[i] = raw.v[perm[i]];
        free(raw.v); raw.v = sorted;
        free(perm); free(key);
        dvec_t* res = &raw;
        if (c.e.merger) {
            merger(&tmp, &raw, 0.5f, 0);
            if (c.e.min_num_detect > 1) {
                int k = 0;
                for (int i = 0; i < tmp.n; ++i) {
                    if (!(tmp.v[i].num_detect < c.e.min_num_detect)) tmp.v[k++] = tmp.v[i];
                    else free(tmp.v[i].fd);
                }
                tmp.n = k;
            }
            if (tmp.n > 0) merger(&fin, &tmp, 0.2f, 1);
            res = &fin;
        }
        for (int i = 0; i < res->n; ++i) {
            rect_t r = res->v[i].r;
            if (pre > 0.f) {                                           /* :2783-2806 */
                const float s = (float)p->min_face / (float)m1->min_h;
                r = mk_rect((int)((float)r.x * s), (int)((float)r.y * s), (int)((float)r.w * s), (int)((float)r.h * s));
            }
            if (nout < cap) {
                out[nout].x = r.x; out[nout].y = r.y; out[nout].w = r.w; out[nout].h = r.h;
                out[nout].score = res->v[i].score;
                out[nout].gender = out[nout].smile = out[nout].glasses = -1;
                if (res->v[i].nfd > 0) {
                    out[nout].gender = fd_median(&res->v[i], 0);
                    out[nout].smile = fd_median(&res->v[i], 1);
                    out[nout].glasses = fd_median(&res->v[i], 2);
                }
            }
            nout++;
        }
    }
    /* raw holds shallow copies of c.rects; tmp and fin own their facial_data */
    dvec_free_fd(&c.rects); dvec_free_fd(&tmp); dvec_free_fd(&fin);
    free(raw.v); free(tmp.v); free(fin.v); free(c.rects.v);
    free(gray); free(pre_img);
    return nout;
}
