// Host side of the MI355X detection path. References are to Bkmz21/CompactCNNCascade
// src/CNNObjectDetector/.
#include "ccnn_detector.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ccnn {

// last HIP error seen by this library on the calling thread (0 if none since the reset)
thread_local int g_last_error = 0;

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            g_last_error = (int)e_;                                                          \
            std::fprintf(stderr, "[ccnn] HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                                \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

void note_error(hipError_t e)
{
    if (e != hipSuccess) g_last_error = (int)e;
}

int last_error(bool reset)
{
    const int e = g_last_error;
    if (reset) g_last_error = 0;
    return e;
}

// ---------------------------------------------------------------------------------
// Model loading: binary format v1.0-1.1 (cnn_simd_cntk.cpp:42-404)
// ---------------------------------------------------------------------------------
namespace {

struct Reader {
    const unsigned char* p;
    size_t n, off = 0;
    bool bad = false;
    template <class T>
    T get()
    {
        T v{};
        if (off + sizeof(T) > n) { bad = true; return v; }
        std::memcpy(&v, p + off, sizeof(T));
        off += sizeof(T);
        return v;
    }
};

}  // namespace

// Two FP32 values rounded to FP16 (to nearest even) and packed into one word, low first.
static uint32_t pack_half2(float a, float b)
{
    const _Float16 x = (_Float16)a, y = (_Float16)b;
    uint16_t ux, uy;
    std::memcpy(&ux, &x, 2);
    std::memcpy(&uy, &y, 2);
    return (uint32_t)ux | ((uint32_t)uy << 16);
}

void append_stage1_half(std::vector<float>& w)
{
    using N = Stage1;
    std::vector<uint32_t> h(Stage1H::SIZE, 0u);
    for (int m = 0; m < N::M1; ++m)
        for (int r = 0; r < 4; ++r) {
            const float* k = &w[N::oK1 + m * 16 + r * 4];
            h[Stage1H::oL1 + m * 8 + 2 * r] = pack_half2(k[0], k[1]);
            h[Stage1H::oL1 + m * 8 + 2 * r + 1] = pack_half2(k[2], k[3]);
        }
    for (int m = 0; m < N::M1; ++m) {
        float sum = 0.f;
        for (int t = 0; t < 16; ++t) sum += w[N::oK1 + m * 16 + t];
        std::memcpy(&h[Stage1H::oS1 + m], &sum, 4);
    }
    for (int m = 0; m < N::M2; ++m)
        for (int r = 0; r < 3; ++r) {
            const float* k = &w[N::oK2 + m * 9 + r * 3];
            h[Stage1H::oL2 + m * 6 + 2 * r] = pack_half2(k[0], k[1]);
            h[Stage1H::oL2 + m * 6 + 2 * r + 1] = pack_half2(k[2], 0.f);
        }
    for (int m = 0; m < N::M3; ++m)
        for (int r = 0; r < N::K3H; ++r) {
            const float* k = &w[N::oK3 + (m * N::K3H + r) * N::K3W];
            h[Stage1H::oL3 + (m * N::K3H + r) * 2] = pack_half2(k[0], k[1]);
            h[Stage1H::oL3 + (m * N::K3H + r) * 2 + 1] = pack_half2(k[2], k[3]);
        }
    for (int u = 0; u < N::HL; ++u) {
        const float* k = &w[N::oHW + u * N::HLN];
        h[Stage1H::oHL + 2 * u] = pack_half2(k[0], k[1]);
        h[Stage1H::oHL + 2 * u + 1] = pack_half2(k[2], k[3]);
    }
    const size_t base = w.size();
    w.resize(base + Stage1H::SIZE);
    std::memcpy(&w[base], h.data(), sizeof(uint32_t) * h.size());
}

// 2^(8 - E) for the smallest power of two 2^E above m: the power of two that brings m
// into [128, 256) (FP8 e4m3 operands reach 448).
static float fp8_scale(float m)
{
    if (!(m > 0.f)) return 1.f;
    int e;
    std::frexp(m, &e);
    return std::ldexp(1.f, 8 - e);
}

void mfma_weights(const std::vector<float>& w, std::vector<float>& out)
{
    using N = Stage2;              // the stage-3 convolutions have the same shape
    static_assert(Stage3::oK1 == N::oK1 && Stage3::oK2 == N::oK2 && Stage3::M1 == N::M1 && Stage3::M2 == N::M2,
                  "one layout for both check networks");
    out.assign(MfmaW::SIZE, 0.f);
    float m1 = 0.f, m2 = 0.f;
    for (int nt = 0; nt < 2; ++nt)
        for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < 8; ++j) {
                const int col = (lane & 15) + 16 * nt, m = col >> 2, q = col & 3, t = 8 * (lane >> 4) + j;
                if (m >= N::M1 || t >= 25) continue;
                const int rr = t / 5 - (q >> 1), cc = t % 5 - (q & 1);
                if (rr < 0 || rr > 3 || cc < 0 || cc > 3) continue;
                const float v = w[N::oK1 + m * 16 + rr * 4 + cc];
                out[MfmaW::oL1 + (nt * 64 + lane) * 8 + j] = v;
                m1 = std::max(m1, std::fabs(v));
            }
    for (int jp = 0; jp < 3; ++jp)
        for (int lane = 0; lane < 64; ++lane)
            for (int j = 0; j < 8; ++j) {
                const int col = lane & 15, kbk = lane >> 4, side = kbk >> 1, t = (kbk & 1) * 8 + j;
                if ((side == 0) != (col < 8)) continue;
                const int m = 4 * jp + (col >> 2), q = col & 3;
                const int rr = (t >> 2) - (q >> 1), cc = (t & 3) - (q & 1);
                if (rr < 0 || rr > 2 || cc < 0 || cc > 2) continue;
                const float v = w[N::oK2 + m * 9 + rr * 3 + cc];
                out[MfmaW::oL2 + (jp * 64 + lane) * 8 + j] = v;
                m2 = std::max(m2, std::fabs(v));
            }
    out[MfmaW::oScale] = fp8_scale(m1);
    out[MfmaW::oScale + 1] = fp8_scale(m2);
}

int load_model(const unsigned char* bytes, size_t n, int stage, int index_output, NetWeights& out, std::string& err)
{
    Reader r{bytes, n};
    const float version = r.get<float>();
    if (r.bad || version < 1.0f || version > 1.1f) {
        err = "Configuration file format is not supported!";
        return -1;
    }
    const int max_pool = r.get<unsigned char>();
    out.min_w = r.get<int32_t>();
    out.min_h = r.get<int32_t>();
    const int layers = r.get<int32_t>();
    if (r.bad || layers != 3) { err = "This configuration cnn models is not supported!"; return -1; }
    int maps[3], kw[3], kh[3];
    for (int l = 0; l < 3; ++l) maps[l] = r.get<int32_t>();
    std::vector<float> kern[3];
    for (int l = 0; l < 3; ++l) {
        kw[l] = r.get<int32_t>();
        kh[l] = r.get<int32_t>();
        if (r.bad || maps[l] <= 0 || maps[l] > 256 || kw[l] <= 0 || kh[l] <= 0 || kw[l] > 16 || kh[l] > 16) {
            err = "This configuration cnn models is not supported!";
            return -1;
        }
        kern[l].resize((size_t)maps[l] * kw[l] * kh[l]);
        for (auto& v : kern[l]) v = r.get<float>();
    }
    float af_scale = r.get<float>();
    std::vector<float> lay[3][5];   // conv_b, lrelu_w1, lrelu_w2, bn_w, bn_b
    for (int l = 0; l < 3; ++l) {
        for (int k = 0; k < 5; ++k) lay[l][k].resize(maps[l]);
        for (int j = 0; j < maps[l]; ++j)
            for (int k = 0; k < 5; ++k) lay[l][k][j] = r.get<float>();
    }
    const int full_connect = r.get<unsigned char>();
    const int hl = r.get<int32_t>();
    const int connect = r.get<int32_t>();
    const int hl_scale = r.get<int32_t>();
    if (r.bad || !max_pool || kw[1] * kh[1] != 9 || hl <= 0 || hl > 4096 || hl_scale <= 0 || connect <= 0 ||
        hl % hl_scale != 0) {
        err = "This configuration cnn models is not supported!";
        return -1;
    }
    // hidden-layer weights: `connect` per unit, or one per L3 map when fully connected
    // (:331-354); the forward pass reads the first `connect` of them (:881)
    const int nin = full_connect ? maps[2] : connect;
    if (connect > nin) { err = "This configuration cnn models is not supported!"; return -1; }
    std::vector<float> hl_w((size_t)hl * nin), hl_b(hl), hl_tw(hl / hl_scale), hl_bnw(hl_scale), hl_bnb(hl_scale);
    for (auto& v : hl_w) v = r.get<float>();
    for (auto& v : hl_b) v = r.get<float>();
    for (auto& v : hl_tw) v = r.get<float>();
    for (int i = 0; i < hl_scale; ++i) { hl_bnw[i] = r.get<float>(); hl_bnb[i] = r.get<float>(); }
    const int ol = r.get<int32_t>();
    if (r.bad || ol <= 0 || ol > 4096) { err = "This configuration cnn models is not supported!"; return -1; }
    std::vector<float> ol_w((size_t)ol * hl), ol_b(ol);
    for (auto& v : ol_w) v = r.get<float>();
    for (auto& v : ol_b) v = r.get<float>();
    const float ol_tw = r.get<float>();
    if (r.bad) { err = "Configuration file is truncated!"; return -1; }
    int io = std::min(index_output, ol - 1);                    // :395-396
    if (io == 0) af_scale = -af_scale;

    auto fill = [&](auto shape) -> int {
        using N = decltype(shape);
        // index_output < 0 computes every output (the landmarks network); otherwise one
        if (maps[0] != N::M1 || maps[1] != N::M2 || maps[2] != N::M3 || kw[0] != N::K1 || kh[0] != N::K1 ||
            kw[2] != N::K3W || kh[2] != N::K3H || hl / hl_scale != N::HLG || hl_scale != N::HLS ||
            connect != N::HLN || (io < 0 ? ol != N::NOUT : N::NOUT != 1)) {
            err = "This configuration cnn models is not supported!";
            return -1;
        }
        std::vector<float>& w = out.w;
        w.assign(N::SIZE, 0.f);
        std::copy(kern[0].begin(), kern[0].end(), w.begin() + N::oK1);
        std::copy(kern[1].begin(), kern[1].end(), w.begin() + N::oK2);
        std::copy(kern[2].begin(), kern[2].end(), w.begin() + N::oK3);
        const int ob[3] = {N::oB1, N::oB2, N::oB3};
        const int mm[3] = {N::M1, N::M2, N::M3};
        for (int l = 0; l < 3; ++l)
            for (int j = 0; j < mm[l]; ++j) {
                w[ob[l] + j] = lay[l][0][j];                 // conv_b
                w[ob[l] + mm[l] + j] = lay[l][1][j];         // lrelu_w1
                w[ob[l] + 2 * mm[l] + j] = lay[l][2][j];     // lrelu_w2
                w[ob[l] + 3 * mm[l] + j] = lay[l][4][j];     // bn_b (bn_w is folded into w1/w2)
            }
        for (int u = 0; u < hl; ++u)
            std::copy(hl_w.begin() + (size_t)u * nin, hl_w.begin() + (size_t)u * nin + connect,
                      w.begin() + N::oHW + (size_t)u * N::HLN);
        std::copy(hl_b.begin(), hl_b.end(), w.begin() + N::oHB);
        std::copy(hl_tw.begin(), hl_tw.end(), w.begin() + N::oHT);
        std::copy(hl_bnw.begin(), hl_bnw.end(), w.begin() + N::oHBW);
        std::copy(hl_bnb.begin(), hl_bnb.end(), w.begin() + N::oHBB);
        const int o0 = io < 0 ? 0 : io;
        std::copy(ol_w.begin() + (size_t)o0 * hl, ol_w.begin() + (size_t)(o0 + N::NOUT) * hl, w.begin() + N::oOW);
        for (int k = 0; k < N::NOUT; ++k) w[N::oOB + k] = ol_b[o0 + k];
        w[N::oOB + N::NOUT] = ol_tw;
        w[N::oOB + N::NOUT + 1] = af_scale;
        return 0;
    };
    if (stage == 0) {
        if (fill(Stage1{}) < 0) return -1;
        append_stage1_half(out.w);
        return 0;
    }
    if (stage == 1) return fill(Stage2{});
    if (stage == 2) return fill(Stage3{});
    return fill(Landmarks{});
}

size_t level_l1_layout(LevelDesc& L, int shape)
{
    const TileShape t = tile_shape(shape);
    L.tiles_x = L.rw > 0 ? (L.rw + t.tx - 1) / t.tx : 0;
    L.tiles_y = L.rh > 0 ? (L.rh + t.ty - 1) / t.ty : 0;
    L.s1_w = L.s1_h = 0;
    L.l1_tiles_x = L.l1_tiles_y = 0;
    if (L.tiles_x <= 0 || L.tiles_y <= 0) return 0;
    // the response tiles read pooled L1 rows 2*oy0 .. 2*oy0 + 2*TY + 9 and columns
    // 2*ox0 .. 2*ox0 + 2*TX + 7
    L.s1_w = 2 * t.tx * L.tiles_x + 8;
    L.s1_h = 2 * t.ty * L.tiles_y + 10;
    // the L1 tiles cover the positions that responses inside the map read: columns
    // < 2*rw + 8, rows < 2*rh + 10. The rest of the planes (read only by the tiles'
    // threads past the response map, whose results are dropped) is never written.
    L.l1_tiles_x = (2 * L.rw + 8 + t.ax - 1) / t.ax;
    L.l1_tiles_y = (2 * L.rh + 10 + t.ay - 1) / t.ay;
    return (size_t)Stage1::M1 * L.s1_w * L.s1_h;
}

// ---------------------------------------------------------------------------------
// Settings, scales
// ---------------------------------------------------------------------------------
Settings effective_settings(const ccnn_param& p)
{
    Settings s;
    s.drop_detect = p.drop_detect != 0;
    s.equalize = p.equalize != 0;
    s.reflection = p.reflection != 0;
    s.detect_mode = p.detect_mode;
    s.merger = p.detect_mode != CCNN_DETECT_DISABLE;            // CompactCNNLibAPI_v2.cpp:133
    s.min_neighbors = p.min_neighbors;
    s.t1 = p.treshold[0]; s.t2 = p.treshold[1]; s.t3 = p.treshold[2];
    s.double_check = 1;                                          // :129
    s.adapt = 0;                                                 // :130
    if (p.detect_precision == CCNN_PRECISION_DEF) {              // cnn_detector_v3.cpp:923-937
        s.type_check = 1; s.equalize = 1; s.reflection = 1; s.adapt = 1; s.double_check = 1;
        s.drop_detect = 0; s.min_num_detect = 1; s.merger = 1;
    }
    switch (p.detect_precision) {                                // :939-959
    case CCNN_PRECISION_LOW: s.type_check = 0; s.drop_detect = 0; break;
    case CCNN_PRECISION_NORMAL: s.type_check = 1; break;
    case CCNN_PRECISION_HIGH: s.type_check = 2; break;
    case CCNN_PRECISION_ULTRA: s.type_check = 2; s.min_num_detect = 2; break;
    default: break;
    }
    s.facial = p.facial_analysis != 0;
    if (s.detect_mode == CCNN_DETECT_DISABLE || s.facial) s.drop_detect = 0;   // :961-964
    return s;
}

std::vector<float> init_scales(const ccnn_param& p, int pattern_h, bool gray_only, float* pre_scale,
                               float* scale_factor_used)
{
    std::vector<float> scales;
    float max_scale = 1.f;
    *pre_scale = 0.f;
    if (p.min_face_height > 0) {
        max_scale = (float)pattern_h / (float)p.min_face_height;
        max_scale = std::min(3.f, max_scale);
        if (gray_only && max_scale < 0.8f) {
            *pre_scale = max_scale;
            max_scale = 1.f;
        }
    }
    float min_scale;
    if (p.max_face_height > 0) min_scale = (float)pattern_h / (float)p.max_face_height;
    else min_scale = (float)pattern_h / (float)p.max_image_height;
    min_scale = std::min(max_scale, min_scale);
    float sf = p.scale_factor;
    if (sf < 1.f) sf = 1.1f;
    *scale_factor_used = sf;
    float scale = max_scale;
    if (sf != 1.f) {
        while (scale > min_scale && scales.size() < (size_t)MAX_LEVELS) {
            scales.push_back(scale);
            scale /= sf;
        }
    } else {
        scales.push_back(scale);
    }
    return scales;
}

static void size_mul(int w, int h, float s, int* ow, int* oh)
{
    *ow = (int)((float)w * s);   // Size::operator* (type.h:63-66)
    *oh = (int)((float)h * s);
}

// ---------------------------------------------------------------------------------
// Rect helpers and Merger
// ---------------------------------------------------------------------------------
// Rect::overlap (type.h:154-173): intersection over the bounding box of both
static float rect_overlap(const Rect& a, const Rect& b)
{
    if (a.x > b.x2 || b.x > a.x2 || a.y > b.y2 || b.y > a.y2) return 0.f;
    const int X0 = std::min(a.x, b.x), X1 = std::max(a.x, b.x), X2 = std::min(a.x2, b.x2), X3 = std::max(a.x2, b.x2);
    const int Y0 = std::min(a.y, b.y), Y1 = std::max(a.y, b.y), Y2 = std::min(a.y2, b.y2), Y3 = std::max(a.y2, b.y2);
    const float su = (float)((X3 - X0) * (Y3 - Y0));
    const float si = (float)((X2 - X1) * (Y2 - Y1));
    return si / su;
}
// Rect::intersects (type.h:140-153)
static int rect_intersects(const Rect& a, const Rect& b)
{
    if (a.x > b.x2 || b.x > a.x2 || a.y > b.y2 || b.y > a.y2) return 0;
    return (std::min(a.x2, b.x2) - std::max(a.x, b.x)) * (std::min(a.y2, b.y2) - std::max(a.y, b.y));
}
// Rect::operator*(float) (type.h:185-192)
static Rect rect_shrink(const Rect& r, float scale)
{
    scale = 1.f - scale;
    const int sx = (int)((float)r.w * scale), sy = (int)((float)r.h * scale);
    return Rect(r.x + (sx >> 1), r.y + (sy >> 1), r.w - sx, r.h - sy);
}

void merger(std::vector<Det>& out, std::vector<Det>& in, float threshold, bool del)
{
    for (size_t a = 0; a < in.size(); ++a) {
        if (in[a].checked) continue;
        Det mx = in[a], mn = in[a];
        for (size_t b = a + 1; b < in.size(); ++b) {
            Det& it2 = in[b];
            if (it2.checked) continue;
            int X[4] = {mx.r.x, mx.r.x2, it2.r.x, it2.r.x2};
            int Y[4] = {mx.r.y, mx.r.y2, it2.r.y, it2.r.y2};
            if (X[0] >= X[3] || X[1] <= X[2] || Y[0] >= Y[3] || Y[1] <= Y[2]) continue;
            const float S1 = (float)((X[1] - X[0]) * (Y[1] - Y[0]));
            const float S2 = (float)((X[3] - X[2]) * (Y[3] - Y[2]));
            std::sort(X, X + 4);
            std::sort(Y, Y + 4);
            const float Su = (float)((X[3] - X[0]) * (Y[3] - Y[0]));
            const float Si = (float)((X[2] - X[1]) * (Y[2] - Y[1]));
            const float r1 = Si / S1, r2 = Si / S2, r3 = Si / Su;
            float ratio = std::min(r1, r2);
            const int CX1 = X[0] + ((X[3] - X[0]) >> 1), CY1 = Y[0] + ((Y[3] - Y[0]) >> 1);
            const int CX2 = X[1] + ((X[2] - X[1]) >> 1), CY2 = Y[1] + ((Y[2] - Y[1]) >> 1);
            const int thC = std::min(X[3] - X[0], X[2] - X[1]) >> 1;
            if (r3 > 0.75f * threshold && std::max(r1, r2) > 0.7f && std::abs(CX1 - CX2) < thC &&
                std::abs(CY1 - CY2) < thC)
                ratio = 1.f;
            if (ratio > threshold) {
                if (del) {
                    if (S1 > S2) { it2.checked = true; continue; }
                    mn.r.x2 = 0;
                    break;
                }
                mx.r.x = std::min(mx.r.x, X[0]);
                mx.r.y = std::min(mx.r.y, Y[0]);
                mx.r.x2 = std::max(mx.r.x2, X[3]);
                mx.r.y2 = std::max(mx.r.y2, Y[3]);
                mn.r.x = std::max(mn.r.x, X[1]);
                mn.r.y = std::max(mn.r.y, Y[1]);
                mn.r.x2 = std::min(mn.r.x2, X[2]);
                mn.r.y2 = std::min(mn.r.y2, Y[2]);
                mx.score = std::max(mx.score, it2.score);
                mx.knn = std::max(mx.knn, it2.knn);
                mx.scale = std::min(mx.scale, it2.scale);
                mx.fd.insert(mx.fd.end(), it2.fd.begin(), it2.fd.end());   // :2911-2914
                mx.num_detect++;
                it2.checked = true;
            }
        }
        if (del && mn.r.x2 == 0) continue;
        mx.r.x = (mn.r.x + mx.r.x) >> 1;
        mx.r.y = (mn.r.y + mx.r.y) >> 1;
        mx.r.x2 = (mn.r.x2 + mx.r.x2) >> 1;
        mx.r.y2 = (mn.r.y2 + mx.r.y2) >> 1;
        mx.r.w = mx.r.x2 - mx.r.x;
        mx.r.h = mx.r.y2 - mx.r.y;
        mx.checked = false;
        out.push_back(mx);
    }
}

std::vector<Det> post_process(const std::vector<Det>& raw, const Settings& s)
{
    std::vector<Det> dets;
    if (raw.empty()) return dets;
    std::vector<Det> rev(raw.rbegin(), raw.rend());                  // reverse_copy (:2746)
    std::sort(rev.begin(), rev.end(), [](const Det& a, const Det& b) { return b.scale < a.scale; });   // :2750
    if (!s.merger) return rev;
    std::vector<Det> tmp;
    tmp.reserve(rev.size() / 2);
    merger(tmp, rev, 0.5f, false);
    if (s.min_num_detect > 1) {
        const int th = s.min_num_detect;
        tmp.erase(std::remove_if(tmp.begin(), tmp.end(), [th](const Det& d) { return d.num_detect < th; }), tmp.end());
    }
    if (!tmp.empty()) merger(dets, tmp, 0.2f, true);
    return dets;
}


// ---------------------------------------------------------------------------------
// Detector
// ---------------------------------------------------------------------------------
namespace {

// the pinned read-back buffers hold at least this many candidates
constexpr size_t PREFETCH = 4096;
// and at most this many; larger candidate lists are copied to pageable memory
constexpr size_t PINNED_CAP = 1u << 20;

}  // namespace

Detector::Detector()
{
    ccnn_param_default(&param);
}

Detector::~Detector()
{
    release_batch();
    for (auto& s : slots_) release_slot(s);
    for (auto& w : d_w_) dfree(w);
    for (auto& w : d_mw_) dfree(w);
    dfree(d_skip_);
    for (Plan* p : plans_) delete p;
}

void Detector::release_slot(Slot& s)
{
    if (s.st) (void)hipStreamSynchronize(s.st);
    dfree(s.d_img); dfree(s.d_pre); dfree(s.d_gray); dfree(s.d_resp); dfree(s.d_s1); dfree(s.d_levels);
    dfree(s.d_l1_order);

    dfree(s.d_meta); dfree(s.d_cands); dfree(s.d_results); dfree(s.d_s2); dfree(s.d_s3);
    hfree(s.h_meta); hfree(s.h_cands); hfree(s.h_results); hfree(s.h_s23);
    for (auto& e : s.ev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    for (auto& e : s.kev)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (s.st) { (void)hipStreamDestroy(s.st); s.st = nullptr; }
    s = Slot();
}

int Detector::init_device()
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        std::printf("[ccnn] no HIP device available; this library runs the detection path on the GPU only\n");
        return -1;
    }
    device_ = (param.cuda_device_id >= 0 && param.cuda_device_id < count) ? param.cuda_device_id : 0;
    HIP_OK(hipSetDevice(device_));
    for (auto& s : slots_) {
        if (!s.st) HIP_OK(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        for (auto& e : s.ev)
            if (!e) HIP_OK(hipEventCreate(&e));
        for (auto& e : s.kev)
            if (!e) HIP_OK(hipEventCreate(&e));
    }
    if (!batch_.st) HIP_OK(hipStreamCreateWithFlags(&batch_.st, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_) == hipSuccess) {
        cus_ = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : cus_;
        if (param.device_info)
            std::printf("[ccnn] device %d: %s (%s), %d CUs\n", device_, prop.name, prop.gcnArchName,
                        prop.multiProcessorCount);
    }
    return 0;
}

int Detector::init(const ccnn_param& p)
{
    param = p;
    gray_only_ = true;          // forced on by the API (CompactCNNLibAPI_v2.cpp:131)
    return build();
}

// CNNDetector::Init chain (cnn_detector_v3.cpp:907-985) for the GPU path.
int Detector::build()
{
    built_ = false;
    empty_ = true;
    if (init_device() < 0) return -1;
    for (auto& s : slots_)
        if (s.busy) { (void)hipStreamSynchronize(s.st); s.busy = false; }
    set_ = effective_settings(param);
    // the landmarks network only with facial analysis (cnn_detector_v3.cpp:792-810); it
    // computes every output (AdvancedParam::index_output[3] = -1)
    for (int st = 0; st < (set_.facial ? 4 : 3); ++st) {
        const unsigned char* b;
        size_t n;
        embedded_model(st, &b, &n);
        std::string err;
        if (load_model(b, n, st, st < 3 ? param.index_output[st] : -1, net_[st], err) < 0) {
            std::printf("[ccnn] %s\n", err.c_str());
            return -1;
        }
        if (dalloc(d_w_[st], net_[st].w.size()) < 0) return -1;
        HIP_OK(hipMemcpy(d_w_[st], net_[st].w.data(), sizeof(float) * net_[st].w.size(), hipMemcpyHostToDevice));
        if (st == 1 || st == 2) {
            // the matrix-core form of the check networks (reduced precision)
            std::vector<float> mw;
            mfma_weights(net_[st].w, mw);
            if (dalloc(d_mw_[st - 1], mw.size()) < 0) return -1;
            HIP_OK(hipMemcpy(d_mw_[st - 1], mw.data(), sizeof(float) * mw.size(), hipMemcpyHostToDevice));
        }
    }
    float sf;
    scales_ = init_scales(param, net_[0].min_h, gray_only_, &pre_scale_, &sf);
    param.scale_factor = sf;                                     // :325-326 writes it back
    num_scales_ = (int)scales_.size();                           // :603
    if (scales_.empty()) {
        std::printf("[CNNDetector] Invalid configuration!\n");
        return -1;
    }
    // ext_pattern_size_cd (:762-764, 879-882)
    if (set_.detect_mode != CCNN_DETECT_DISABLE) {
        ext_w_ = net_[1].min_w + 11;
        ext_h_ = net_[1].min_h + 11;
        if (net_[1].min_w != net_[2].min_w || net_[1].min_h != net_[2].min_h) {
            std::printf("[CNNDetector] This configuration cnn models is not supported!\n");
            return -1;
        }
    } else {
        ext_w_ = net_[0].min_w;
        ext_h_ = net_[0].min_h;
    }
    for (Plan* pl : plans_) delete pl;
    plans_.clear();
    for (auto& s : slots_) s.levels_plan = nullptr;
    batch_.key.clear();          // the plans the batch graph was built from are gone
    batch_.gen++;
    built_ = true;
    empty_ = false;
    return 0;
}

// FaceDetector::Clear deletes the detector (CompactCNNLibAPI_v2.cpp:355-372): every device
// and pinned buffer, the weights and the cached plans are released; Init builds them again.
int Detector::clear()
{
    if (empty_) return -1;
    release_batch();
    for (auto& s : slots_) release_slot(s);
    for (auto& w : d_w_) dfree(w);
    for (auto& w : d_mw_) dfree(w);
    for (Plan* p : plans_) delete p;
    plans_.clear();
    last_slot_ = -1;
    last_gray_ = nullptr;
    last_resp_ = nullptr;
    std::vector<float>().swap(h_gray_);
    empty_ = true;
    built_ = false;
    return 0;
}

// Level table for a gray image of gw x gh: levels scl_max -> 0 (RunCPUDetect order,
// which is also the RunCheckDetectAsync order), skip rule (:2083-2088), resize mode
// (:2093-2104), response size (getOutputImgSize, cnn_simd_cntk.cpp:997-1012).
const Plan* Detector::plan_for(int gw, int gh, int ncap, int shape)
{
    ncap = std::min(std::max(ncap, 0), (int)scales_.size());
    for (const Plan* p : plans_)
        if (p->gw == gw && p->gh == gh && p->ncap == ncap && p->shape == shape) return p;
    Plan* P = new Plan();
    P->id = ++plan_ids_;
    P->shape = shape;
    P->gw = gw;
    P->gh = gh;
    P->ncap = ncap;
    P->next_num_scales = ncap;
    int tiles = 0, chunks = 0, l1tiles = 0;
    size_t roff = 0, s1off = 0;
    for (int scl = ncap - 1; scl >= 0; --scl) {
        LevelDesc L{};
        L.scale = scales_[scl];
        L.scl = scl;
        size_mul(gw, gh, L.scale, &L.lw, &L.lh);
        if (L.lw < ext_w_ || L.lh < ext_h_) {
            P->next_num_scales = std::min(P->next_num_scales, scl);
            continue;
        }
        L.mode = (L.scale != 1.f) ? (L.scale > 0.7f ? 0 : 1) : 2;
        if (L.lw == gw && L.lh == gh) L.mode = 2;       // FastImageResize copies equal sizes
        L.sclx = ((uint32_t)gw << 12) / (uint32_t)L.lw + 1;
        L.scly = ((uint32_t)gh << 12) / (uint32_t)L.lh + 1;
        const int c2 = ((L.lw - 3) >> 1) - 2, r2 = ((L.lh - 3) >> 1) - 2;
        L.rw = (c2 >> 1) - (Stage1::K3W - 1);
        L.rh = (r2 >> 1) - (Stage1::K3H - 1);
        if (L.rw <= 0 || L.rh <= 0) { L.rw = 0; L.rh = 0; }
        L.resp_off = (int)roff;
        roff += (size_t)L.rw * L.rh;
        L.tile_off = tiles;
        L.s1_off = (int)s1off;
        L.l1_tile_off = l1tiles;
        s1off += level_l1_layout(L, shape);
        tiles += L.tiles_x * L.tiles_y;
        l1tiles += L.l1_tiles_x * L.l1_tiles_y;
        L.chunk_off = chunks;
        L.chunks = (L.rw * L.rh + CMP_CHUNK - 1) / CMP_CHUNK;
        chunks += L.chunks;
        P->levels.push_back(L);
    }
    P->ntiles = tiles;
    P->nchunks = chunks;
    P->nl1tiles = l1tiles;
    P->resp_total = roff;
    P->s1_total = s1off;
    // L1 tiles band by band: band j holds tile rows [j, j+1) * tiles_y / 8 of every level,
    // so that XCD j (which runs one contiguous eighth of the grid) reads one eighth of
    // the gray image for all levels
    constexpr int NB = 8;
    P->l1_order.reserve(l1tiles);
    for (int j = 0; j < NB; ++j)
        for (int l = 0; l < (int)P->levels.size(); ++l) {
            const LevelDesc& L = P->levels[l];
            for (int ty = 0; ty < L.l1_tiles_y; ++ty) {
                if (ty * NB / L.l1_tiles_y != j) continue;
                for (int tx = 0; tx < L.l1_tiles_x; ++tx) P->l1_order.push_back(make_int2(l, ty * L.l1_tiles_x + tx));
            }
        }
    // then the stage-1 tiles in launch order, {level, tile within the level}
    P->l1_order.reserve(l1tiles + tiles);
    for (int l = 0; l < (int)P->levels.size(); ++l)
        for (int t = 0; t < P->levels[l].tiles_x * P->levels[l].tiles_y; ++t) P->l1_order.push_back(make_int2(l, t));
    if (plans_.size() >= 8) {          // a handful of frame sizes stay cached
        for (size_t k = 0; k < plans_.size(); ++k) {
            bool used = plans_[k]->pins > 0;
            for (auto& s : slots_) used |= s.busy && s.plan == plans_[k];
            if (used) continue;
            for (auto& s : slots_)
                if (s.levels_plan == plans_[k]) s.levels_plan = nullptr;
            delete plans_[k];
            plans_.erase(plans_.begin() + k);
            break;
        }
    }
    plans_.push_back(P);
    return P;
}

int Detector::ensure_slot(Slot& s, int w, int h, size_t img_bytes, const Plan& plan)
{
    if (img_bytes > s.img_bytes) {
        if (dalloc(s.d_img, img_bytes) < 0) return -1;
        s.img_bytes = img_bytes;
    }
    const size_t pre = (size_t)w * h;
    if (pre > s.pre_bytes) {
        if (dalloc(s.d_pre, pre) < 0) return -1;
        s.pre_bytes = pre;
    }
    const int gp = (w + 63) / 64 * 64;
    const size_t ge = (size_t)gp * (h + 1);
    // the stage-1 sampler addresses the gray image with 32-bit element offsets
    if (ge > 0xFFFFFFFFull) return -1;
    if (ge > s.gray_elems || gp > s.gpitch) {
        if (dalloc(s.d_gray, std::max(ge, s.gray_elems)) < 0) return -1;
        s.gray_elems = std::max(ge, s.gray_elems);
    }
    s.gpitch = gp;
    if (!s.d_levels && dalloc(s.d_levels, MAX_LEVELS) < 0) return -1;
    if (plan.l1_order.size() > s.l1_order_cap) {
        if (dalloc(s.d_l1_order, plan.l1_order.size()) < 0) return -1;
        s.l1_order_cap = plan.l1_order.size();
        s.levels_plan = nullptr;                 // re-upload the tables
    }
    if (!s.d_meta && dalloc(s.d_meta, 16) < 0) return -1;
    if (!s.h_meta && halloc(s.h_meta, s.hd_meta, 16) < 0) return -1;
    if (plan.resp_total > s.resp_cap) {
        const size_t n = plan.resp_total;
        if (dalloc(s.d_resp, n) < 0 || dalloc(s.d_cands, n) < 0 || dalloc(s.d_results, n) < 0 ||
            dalloc(s.d_s2, 2 * n) < 0 || dalloc(s.d_s3, 2 * n) < 0)
            return -1;
        s.resp_cap = s.cand_cap = n;
    }
    if (plan.s1_total > s.s1_cap) {
        if (dalloc(s.d_s1, plan.s1_total) < 0) return -1;
        // plane positions no L1 tile writes then hold zeros rather than whatever the
        // allocation held
        if (hipMemset(s.d_s1, 0, plan.s1_total * sizeof(float)) != hipSuccess) return -1;
        s.s1_cap = plan.s1_total;
    }
    const size_t hc = std::min(std::max<size_t>(plan.resp_total, PREFETCH), PINNED_CAP);
    if (hc > s.h_cap) {
        if (halloc(s.h_cands, s.hd_cands, hc) < 0 || halloc(s.h_results, s.hd_results, hc) < 0 ||
            halloc(s.h_s23, s.hd_s23, 4 * hc) < 0)
            return -1;
        s.h_cap = hc;
    }
    return 0;
}

// Size guards and re-initialisation at the start of CNNDetector::Detect
// (cnn_detector_v3.cpp:2483-2547). Returns 1 to run the frame, 0 for "no faces", -1 on error.
int Detector::prepare(const ccnn_image& img)
{
    if (empty_ || !built_) {
        if (build() < 0) {
            std::printf("[CNNDetector] CNN detector no initialized!\n");
            return -1;
        }
    }
    const int w = img.cols, h = img.rows, ch = img.channels;
    if (w < ext_w_ || h < ext_h_) return 0;                                  // :2494-2497
    if (w > param.max_image_width || h > param.max_image_height) {           // :2499-2513
        param.max_image_width = w;
        param.max_image_height = h;
        if (gray_only_ && ch != 1) gray_only_ = false;
        if (build() < 0) {
            std::printf("[CNNDetector] CNN detector no initialized!\n");
            return -1;
        }
    } else if (w != param.max_image_width || h != param.max_image_height) {
        num_scales_ = (int)scales_.size();                                   // :2514-2518
    }
    if (gray_only_ && pre_scale_ > 0.f && ch != 1) {                          // :2540-2546
        std::printf("[CNNDetector] gray_image_only flag set to false!\n");
        gray_only_ = false;
        if (build() < 0) return -1;
    }
    if (ch != 1 && ch != 3 && ch != 4) {
        std::printf("[CNNDetector] This type image is not supported!\n");
        return -1;
    }
    if (!img.data) return -1;
    return 1;
}

// Decision of CPUCheckDetect (cnn_detector_v3.cpp:1709-1725, 1808-1811) from the stage-2/3
// outcomes of both crop variants (s2/s3[mod] = {count above the threshold, max bits};
// s3 {-1, ..} when stage 3 was skipped): mod 0, then mod 1 if mod 0 is rejected. The same
// rule as k_check_decide.
CheckResult check_decision(const CheckParams& P, float scale, const int2* s2, const int2* s3)
{
    const int mn = (P.adapt && scale > 0.7f && P.min_neighbors > 1) ? P.min_neighbors - 1 : P.min_neighbors;
    CheckResult res{};
    res.n2 = 0;
    res.n3 = 0;
    for (int mod = 0; mod < 2; ++mod) {
        const int2 a = s2[mod], b = s3[mod];
        const int k1 = a.x, k2 = b.x < 0 ? 0 : b.x;
        float m1, m2;
        std::memcpy(&m1, &a.y, sizeof(float));
        std::memcpy(&m2, &b.y, sizeof(float));
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

CheckParams Detector::check_params() const
{
    CheckParams P;
    P.t2 = set_.t2; P.t3 = set_.t3; P.type_check = set_.type_check; P.min_neighbors = set_.min_neighbors;
    P.adapt = set_.adapt; P.double_check = set_.double_check; P.equalize = set_.equalize;
    P.reflection = set_.reflection;
    return P;
}

// Launches one frame on slot s: upload (host frames), optional 8-bit pre-resize, gray,
// stage 1 over all levels (which appends the candidate windows), stage-2/3 checks,
// asynchronous read-back.
const Plan* Detector::plan_frame(const ccnn_image& img, bool* pre, int* gw, int* gh, int shape)
{
    *gw = img.cols;
    *gh = img.rows;
    *pre = gray_only_ && pre_scale_ > 0.f && img.channels == 1;
    if (*pre) size_mul(img.cols, img.rows, pre_scale_, gw, gh);
    const Plan* plan = plan_for(*gw, *gh, num_scales_, shape);
    num_scales_ = plan->next_num_scales;
    return plan;
}

int Detector::enqueue(Slot& s, const ccnn_image& img, bool device_ptr)
{
    bool pre;
    int gw, gh;
    const Plan* plan = plan_frame(img, &pre, &gw, &gh, frame_shape_);
    return enqueue_plan(s, img, device_ptr, plan, pre, gw, gh);
}

int Detector::enqueue_plan(Slot& s, const ccnn_image& img, bool device_ptr, const Plan* plan, bool pre, int gw, int gh)
{
    HIP_OK(hipSetDevice(device_));
    const int w = img.cols, h = img.rows, ch = img.channels;
    int gch = ch;
    size_t gstep = img.step;
    const size_t bytes = img.step * (size_t)(h - 1) + (size_t)w * ch;
    if (ensure_slot(s, std::max(w, gw), std::max(h, gh), device_ptr ? 0 : bytes, *plan) < 0) return -1;
    hipStream_t st = s.st;
    const uint8_t* src = img.data;
    if (profiling_) HIP_OK(hipEventRecord(s.ev[0], st));
    if (!device_ptr) {
        HIP_OK(hipMemcpyAsync(s.d_img, img.data, bytes, hipMemcpyHostToDevice, st));
        src = s.d_img;
    }
    if (pre) {
        // 8-bit bilinear pre-resize of a gray frame (cnn_detector_v3.cpp:2531-2539)
        HIP_OK(launch_resize_u8(src, w, h, (int)img.step, s.d_pre, gw, gh, st));
        src = s.d_pre;
        gstep = gw;
        gch = 1;
    }
    HIP_OK(launch_gray(src, gw, gh, gch, (int)gstep, s.d_gray, s.gpitch, st));
    const int nl = (int)plan->levels.size();
    if (s.levels_plan != plan && nl > 0) {
        HIP_OK(hipMemcpyAsync(s.d_levels, plan->levels.data(), sizeof(LevelDesc) * nl, hipMemcpyHostToDevice, st));
        if (!plan->l1_order.empty())
            HIP_OK(hipMemcpyAsync(s.d_l1_order, plan->l1_order.data(), sizeof(int2) * plan->l1_order.size(),
                                  hipMemcpyHostToDevice, st));
        s.levels_plan = plan;
    }
    if (profiling_) HIP_OK(hipEventRecord(s.ev[1], st));
    // k_stage1_l1 zeroes the candidate count and k_stage1 appends the candidates
    s.checks = set_.detect_mode != CCNN_DETECT_DISABLE && nl > 0;
    // candidates and check outcomes go straight to the mapped host buffers (up to s.pre):
    // k_stage1 writes the candidates and k_check the stage-2/3 outcomes there, and the host
    // takes the decisions; k_facial takes the same decisions itself and writes the facial
    // results of the accepted candidates there; without checks k_export copies the
    // candidates
    s.pre = std::min(s.h_cap, s.cand_cap);
    s.host_decide = s.checks;
    HIP_OK(launch_stage1(s.d_gray, s.gpitch, gw, gh, s.d_levels, nl, plan->nl1tiles, plan->ntiles, d_w_[0], s.d_s1,
                         s.d_resp, set_.t1, s.d_meta, s.d_cands, s.host_decide ? s.hd_cands : nullptr, (int)s.pre,
                         s.d_l1_order, st, profiling_ ? s.kev : nullptr, plan->shape, half_stage1()));
    if (plan->ntiles <= 0) {
        HIP_OK(hipMemsetAsync(s.d_meta, 0, 2 * sizeof(int), st));
        if (profiling_)
            for (int k = 0; k < 4; ++k) HIP_OK(hipEventRecord(s.kev[k], st));
    }
    if (profiling_) HIP_OK(hipEventRecord(s.ev[2], st));
    if (s.checks) {
        HIP_OK(launch_check(s.d_gray, s.gpitch, gw, gh, s.d_levels, s.d_cands, s.d_meta, check_params(), d_w_[1],
                            d_w_[2], s.d_s2, s.d_s3, s.d_results,
                            check_frame_per_cu_ > 0 ? check_frame_per_cu_ * cus_ : check_blocks_,
                            s.host_decide ? s.hd_meta : nullptr, nullptr, nullptr,
                            s.host_decide ? s.hd_s23 : nullptr, s.host_decide ? s.hd_s23 + 2 * s.h_cap : nullptr,
                            (int)s.pre, st, profiling_ ? s.kev + 4 : nullptr, 0x7fffffff,
                            check_frame_per_cu_ > 0 ? s.d_meta + 1 : nullptr, check_precision(),
                            d_mw_[0], d_mw_[1]));
        if (set_.facial)
            HIP_OK(launch_facial(s.d_gray, s.gpitch, gw, gh, s.d_levels, s.d_cands, s.d_meta, check_params(), s.d_s2,
                                 s.d_s3, d_w_[3], s.d_results, s.hd_results, (int)s.pre, fa_blocks_, st));
    } else if (profiling_) {
        for (int k = 4; k < 6; ++k) HIP_OK(hipEventRecord(s.kev[k], st));
    }
    if (nl > 0 && !s.host_decide)
        HIP_OK(launch_export(s.d_meta, s.d_cands, s.checks ? s.d_results : nullptr, s.hd_meta, s.hd_cands,
                             s.hd_results, (int)s.pre, st));
    if (profiling_) HIP_OK(hipEventRecord(s.ev[3], st));
    s.plan = plan;
    s.gw = gw;
    s.gh = gh;
    s.rescale = pre;
    s.busy = true;
    return 0;
}

// Waits for slot s and replays the reference's order-dependent decisions.
int Detector::finish(Slot& s, std::vector<ccnn_face>& faces)
{
    faces.clear();
    HIP_OK(hipStreamSynchronize(s.st));
    s.busy = false;
    last_slot_ = (int)(&s - slots_);
    const std::vector<LevelDesc>& levels = s.plan->levels;
    const int nl = (int)levels.size();
    const int count = nl > 0 ? s.h_meta[0] : 0;
    const bool checks = s.checks;
    std::vector<Candidate> big_c;
    std::vector<CheckResult> big_r;
    const Candidate* cands = s.h_cands;
    const CheckResult* res = s.h_results;
    const bool more = (size_t)count > s.pre;
    if (more) {
        // more candidates than the mapped buffers hold: copy the whole list
        big_c.resize(count);
        HIP_OK(hipMemcpy(big_c.data(), s.d_cands, sizeof(Candidate) * count, hipMemcpyDeviceToHost));
        cands = big_c.data();
    }
    if (s.host_decide) {
        // CPUCheckDetect's decision from the stage-2/3 outcomes, as k_check_decide takes it
        std::vector<int2> big_s;
        const int2* s2 = s.h_s23;
        const int2* s3 = s.h_s23 + 2 * s.h_cap;
        if (more) {
            big_s.resize(4 * (size_t)count);
            HIP_OK(hipMemcpy(big_s.data(), s.d_s2, sizeof(int2) * 2 * count, hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(big_s.data() + 2 * count, s.d_s3, sizeof(int2) * 2 * count, hipMemcpyDeviceToHost));
            s2 = big_s.data();
            s3 = big_s.data() + 2 * count;
        }
        big_r.resize(count);
        const CheckParams P = check_params();
        for (int ci = 0; ci < count; ++ci)
            big_r[ci] = check_decision(P, levels[cands[ci].level].scale, s2 + 2 * (size_t)ci, s3 + 2 * (size_t)ci);
        if (set_.facial) {
            // k_facial's results of the accepted candidates
            std::vector<CheckResult> dev_r;
            const CheckResult* fr = s.h_results;
            if (more) {
                dev_r.resize(count);
                HIP_OK(hipMemcpy(dev_r.data(), s.d_results, sizeof(CheckResult) * count, hipMemcpyDeviceToHost));
                fr = dev_r.data();
            }
            for (int ci = 0; ci < count; ++ci)
                if (big_r[ci].flag) big_r[ci].fa = fr[ci].fa;
        }
        res = big_r.data();
    }
    if (profiling_) {
        // gray, k_stage1_l1, k_stage1, the gap to k_check, k_check, decisions (+ facial
        // analysis, read-back), the whole frame
        const hipEvent_t span[NT][2] = {{s.ev[0], s.ev[1]},  {s.kev[0], s.kev[1]}, {s.kev[2], s.kev[3]},
                                        {s.kev[3], s.ev[2]}, {s.kev[4], s.kev[5]}, {s.kev[5], s.ev[3]},
                                        {s.ev[0], s.ev[3]}};
        for (int i = 0; i < NT; ++i) (void)hipEventElapsedTime(&last_ms_[i], span[i][0], span[i][1]);
        for (int i = 0; i < NT; ++i) acc_ms_[i] += last_ms_[i];
        acc_frames_++;
    }
    // k_stage1 appends the candidate windows in no particular order; the reference visits
    // them level by level (check order) and in raster order inside a level. Bucket them
    // by chunk (CMP_CHUNK consecutive raster positions of one level, chunks in
    // check order), then order each bucket by raster position.
    {
        const int nchunks = s.plan->nchunks;
        std::vector<int> start(nchunks + 2, 0), ord(count);
        std::vector<int> key(count);
        for (int i = 0; i < count; ++i) {
            const LevelDesc& L = levels[cands[i].level];
            const int e = cands[i].y * L.rw + cands[i].x;
            key[i] = e;
            start[L.chunk_off + e / CMP_CHUNK + 2]++;
        }
        for (int c = 0; c < nchunks; ++c) start[c + 2] += start[c + 1];
        for (int i = 0; i < count; ++i) {
            const LevelDesc& L = levels[cands[i].level];
            ord[start[L.chunk_off + key[i] / CMP_CHUNK + 1]++] = i;
        }
        for (int c = 0; c < nchunks; ++c)
            std::sort(ord.begin() + (c > 0 ? start[c] : 0), ord.begin() + start[c + 1],
                      [&](int a, int b) { return key[a] < key[b]; });
        last_cands_.resize(count);
        for (int i = 0; i < count; ++i) last_cands_[i] = cands[ord[i]];
        if (checks) {
            last_results_.resize(count);
            for (int i = 0; i < count; ++i) last_results_[i] = res[ord[i]];
        } else {
            last_results_.clear();
        }
        cands = last_cands_.data();
        res = checks ? last_results_.data() : nullptr;
    }
    last_count_ = count;
    last_levels_ = levels;
    last_gray_ = s.d_gray;
    last_gp_ = s.gpitch;
    last_gw_ = s.gw;
    last_gh_ = s.gh;
    last_resp_ = s.d_resp;
    const std::vector<Det> rects = replay(levels.data(), nl, cands, res, count, checks, &last_status_);
    emit_faces(rects, s.rescale, faces);
    return (int)faces.size();
}

// Ordered replay of RunCheckDetect / CPUCheckDetect (cnn_detector_v3.cpp:1842-2011,
// 1485-1818) over one frame's candidates in check order (level, then raster): the
// scan-time DropDetection against all detections so far, then per surviving point the
// drop against this level's detections and the precomputed stage-2/3 outcome. Returns
// the detections before post-processing.
std::vector<Det> Detector::replay(const LevelDesc* levels, int nl, const Candidate* cands, const CheckResult* res,
                                  int count, bool checks, std::vector<int>* status) const
{
    std::vector<Det> rects;
    std::vector<int> points;
    if (status) status->assign(count, 0);
    int ci = 0;
    for (int l = 0; l < nl; ++l) {
        const float scale = levels[l].scale;
        const float inv = 1.f / scale;
        points.clear();
        for (; ci < count && cands[ci].level == l; ++ci) {
            const int px = cands[ci].x * 4, py = cands[ci].y * 4;
            if (set_.drop_detect) {
                const Rect nr((int)((float)px * inv), (int)((float)py * inv), (int)(23.f * inv), (int)(27.f * inv));
                // DropDetection (:1820-1841)
                const Rect exp = rect_shrink(nr, 0.7f);
                bool drop = false;
                const size_t n = rects.size();
                for (size_t k = 0; k < n; ++k) {
                    const float ov = rect_overlap(nr, rects[k].r);
                    if (ov == 0.f) continue;
                    if (ov > 0.5f) {
                        Det d;
                        d.r = nr; d.score = rects[k].score; d.scale = scale; d.knn = rects[k].knn;
                        rects.push_back(d);
                        drop = true;
                        break;
                    }
                    if (rect_intersects(rects[k].r, exp) == exp.w * exp.h) { drop = true; break; }
                }
                if (drop) continue;
            }
            points.push_back(ci);
        }
        if (checks && !points.empty()) {
            const size_t rect_size = rects.size();
            for (int pi : points) {
                const Candidate& c = cands[pi];
                const int px = c.x * 4, py = c.y * 4;
                const Rect nr((int)((float)px * inv), (int)((float)py * inv), (int)(23.f * inv), (int)(27.f * inv));
                if (set_.drop_detect) {                                      // :1513-1533
                    bool drop = false;
                    for (size_t k = rect_size; k < rects.size(); ++k) {
                        if (rect_overlap(nr, rects[k].r) > 0.5f) {
                            Det d;
                            d.r = nr; d.score = rects[k].score; d.scale = scale; d.knn = rects[k].knn;
                            rects.push_back(d);
                            drop = true;
                            break;
                        }
                    }
                    if (drop) {
                        if (status) (*status)[pi] = 1;
                        continue;
                    }
                }
                if (status) (*status)[pi] = 2;
                const CheckResult& r = res[pi];
                // facial analysis drops the window, without the retry (:1749)
                if (r.flag && set_.facial && !r.fa.ok) continue;
                if (r.flag) {
                    Det d;
                    d.r = nr;
                    // :1758
                    d.score = c.score + (float)r.k1 * std::max(-1.7159f, r.m1) + (float)r.k2 * std::max(-1.7159f, r.m2);
                    d.scale = scale;
                    d.knn = std::min(r.k1, r.k2);
                    if (set_.facial) d.fd.push_back({r.fa.gender, r.fa.smile, r.fa.glasses});   // :1762-1765
                    rects.push_back(d);
                }
            }
        } else {
            for (int pi : points) {
                const Candidate& c = cands[pi];
                if (status) (*status)[pi] = 3;
                Det d;
                d.r = Rect((int)((float)(c.x * 4) * inv), (int)((float)(c.y * 4) * inv), (int)(23.f * inv),
                           (int)(27.f * inv));
                d.score = c.score;
                d.scale = scale;
                d.knn = 0;
                rects.push_back(d);
            }
        }
    }
    return rects;
}

// Post-processing of Detect and FaceDetector's output conversion: Merger passes, the
// rescale of a pre-resized frame's boxes (cnn_detector_v3.cpp:2783-2806) and the facial
// attributes per face.
void Detector::emit_faces(const std::vector<Det>& rects, bool rescale, std::vector<ccnn_face>& faces) const
{
    faces.clear();
    std::vector<Det> dets = post_process(rects, set_);
    const float sc = rescale ? (float)param.min_face_height / (float)net_[0].min_h : 1.f;   // :2783-2806
    faces.reserve(dets.size());
    for (const Det& d : dets) {
        Rect r = d.r;
        if (rescale)
            r = Rect((int)((float)r.x * sc), (int)((float)r.y * sc), (int)((float)r.w * sc), (int)((float)r.h * sc));
        ccnn_face f;
        f.x = r.x; f.y = r.y; f.width = r.w; f.height = r.h; f.score = d.score;
        f.gender = -1; f.smile = -1; f.glasses = -1;
        if (!d.fd.empty()) {
            // FaceDetector::Detect reports, per attribute, the element at size/2 of the
            // entries sorted in descending order (CompactCNNLibAPI_v2.cpp:296-319)
            int* dst[3] = {&f.gender, &f.smile, &f.glasses};
            std::vector<int> v(d.fd.size());
            for (int k = 0; k < 3; ++k) {
                for (size_t i = 0; i < v.size(); ++i) v[i] = d.fd[i][k];
                std::sort(v.begin(), v.end(), [](int a, int b) { return a > b; });
                *dst[k] = v[v.size() / 2];
            }
        }
        faces.push_back(f);
    }
}

int Detector::detect(const ccnn_image& img, bool device_ptr, std::vector<ccnn_face>& faces)
{
    faces.clear();
    const int r = prepare(img);
    if (r <= 0) return r;
    Slot& s = slots_[0];
    if (s.busy) (void)hipStreamSynchronize(s.st);
    if (enqueue(s, img, device_ptr) < 0) return -1;
    return finish(s, faces);
}

int Detector::detect_batch(int n, const ccnn_image* imgs, bool device_ptr, std::vector<std::vector<ccnn_face>>& faces)
{
    if (batch_mode_ == 1 && !set_.facial && n > 0) return detect_batch_graph(n, imgs, device_ptr, faces);
    return detect_batch_pipelined(n, imgs, device_ptr, faces);
}

int Detector::detect_batch_pipelined(int n, const ccnn_image* imgs, bool device_ptr,
                                     std::vector<std::vector<ccnn_face>>& faces)
{
    faces.assign(std::max(n, 0), std::vector<ccnn_face>());
    // frames go to the slots round robin; when every slot holds a frame in flight, the
    // oldest is finished (host replay) while the others run on the GPU
    int owner[NSLOTS];
    for (int& o : owner) o = -1;
    int next = 0;
    auto finish_slot = [&](int k) -> int {
        const int f = owner[k];
        owner[k] = -1;
        return f >= 0 ? finish(slots_[k], faces[f]) : 0;
    };
    auto drain = [&]() -> int {
        for (int j = 0; j < NSLOTS; ++j)
            if (finish_slot((next + j) % NSLOTS) < 0) return -1;
        return 0;
    };
    for (int i = 0; i < n; ++i) {
        // the guards may re-initialise the detector: finish the frames in flight first
        const bool guard = empty_ || !built_ || imgs[i].cols > param.max_image_width ||
                           imgs[i].rows > param.max_image_height ||
                           (gray_only_ && pre_scale_ > 0.f && imgs[i].channels != 1);
        if (guard && drain() < 0) return -1;
        const int r = prepare(imgs[i]);
        if (r < 0) {
            faces[i].clear();
            continue;
        }
        if (r == 0) continue;
        if (finish_slot(next) < 0) return -1;
        Slot& s = slots_[next];
        if (s.busy) (void)hipStreamSynchronize(s.st);
        if (enqueue(s, imgs[i], device_ptr) < 0) return -1;
        owner[next] = i;
        next = (next + 1) % NSLOTS;
    }
    return drain();
}

int Detector::gray_image(ccnn_image* img, int pipeline)
{
    if (!img || empty_ || !last_gray_ || last_gw_ <= 0) return -1;
    for (auto& s : slots_)
        if (s.busy) (void)hipStreamSynchronize(s.st);
    if (pipeline == CCNN_PIPELINE_CPU) {
        h_gray_.resize((size_t)last_gp_ * last_gh_);
        HIP_OK(hipMemcpy(h_gray_.data(), last_gray_, sizeof(float) * h_gray_.size(), hipMemcpyDeviceToHost));
        img->data = (unsigned char*)h_gray_.data();
    } else {
        img->data = (unsigned char*)last_gray_;
    }
    img->cols = last_gw_;
    img->rows = last_gh_;
    img->channels = 1;
    img->step = (size_t)last_gp_;
    return 0;
}

int Detector::last_timing(float* ms) const
{
    for (int i = 0; i < NT; ++i) ms[i] = last_ms_[i];
    for (int i = 0; i < NT; ++i) ms[NT + i] = (float)acc_ms_[i];
    ms[2 * NT] = (float)acc_frames_;
    return 0;
}

int Detector::last_trace(int* n_levels, int* levels, int cap, float* cand, float* res) const
{
    if (n_levels) *n_levels = (int)last_levels_.size();
    if (levels)
        for (size_t l = 0; l < last_levels_.size(); ++l) {
            const LevelDesc& L = last_levels_[l];
            int* o = levels + 6 * l;
            o[0] = L.scl; o[1] = L.lw; o[2] = L.lh; o[3] = L.rw; o[4] = L.rh; o[5] = L.mode;
        }
    const int n = std::min(cap, last_count_);
    for (int i = 0; i < n && cand; ++i) {
        const Candidate& c = last_cands_[i];
        cand[4 * i] = (float)c.level; cand[4 * i + 1] = (float)c.x; cand[4 * i + 2] = (float)c.y;
        cand[4 * i + 3] = c.score;
    }
    for (int i = 0; i < n && res && !last_results_.empty(); ++i) {
        const CheckResult& r = last_results_[i];
        float* o = res + 8 * i;
        o[0] = (float)r.flag; o[1] = (float)r.mod; o[2] = (float)r.k1; o[3] = (float)r.k2;
        o[4] = r.m1; o[5] = r.m2; o[6] = (float)r.n2; o[7] = (float)r.n3;
    }
    return last_count_;
}

long Detector::last_responses(float* out, size_t cap)
{
    if (empty_ || !last_resp_) return -1;
    size_t n = 0;
    for (const LevelDesc& L : last_levels_) n += (size_t)L.rw * L.rh;
    if (!out) return (long)n;
    if (cap < n) return -1;
    for (auto& s : slots_)
        if (s.busy) HIP_OK(hipStreamSynchronize(s.st));
    HIP_OK(hipMemcpy(out, last_resp_, sizeof(float) * n, hipMemcpyDeviceToHost));
    return (long)n;
}

int Detector::last_facial(int* out, int cap) const
{
    const int n = std::min(cap, (int)last_results_.size());
    for (int i = 0; i < n && out; ++i) {
        const FaResult& f = last_results_[i].fa;
        out[4 * i] = f.ok; out[4 * i + 1] = f.gender; out[4 * i + 2] = f.smile; out[4 * i + 3] = f.glasses;
    }
    return (int)last_results_.size();
}

int Detector::stage_check(const float* gray, int gw, int gh, int n, const float* scale, const int* x, const int* y,
                          float* res)
{
    if (empty_ || !built_) if (build() < 0) return -1;
    if (n <= 0) return 0;
    HIP_OK(hipSetDevice(device_));
    const int gp = (gw + 63) / 64 * 64;
    float* d_g = nullptr;
    LevelDesc* d_l = nullptr;
    Candidate* d_c = nullptr;
    CheckResult* d_r = nullptr;
    int2* d_s2 = nullptr;
    int2* d_s3 = nullptr;
    int* d_m = nullptr;
    std::vector<LevelDesc> lv(n);
    std::vector<Candidate> cd(n);
    for (int i = 0; i < n; ++i) {
        lv[i] = LevelDesc{};
        lv[i].scale = scale[i];
        cd[i].level = i; cd[i].x = x[i]; cd[i].y = y[i]; cd[i].score = 0.f;
    }
    int rc = 0;
    std::vector<CheckResult> out(n);
    hipStream_t st = slots_[0].st;
    if (dalloc(d_g, (size_t)gp * (gh + 1)) < 0 || dalloc(d_l, n) < 0 || dalloc(d_c, n) < 0 || dalloc(d_r, n) < 0 ||
        dalloc(d_s2, 2 * (size_t)n) < 0 || dalloc(d_s3, 2 * (size_t)n) < 0 || dalloc(d_m, 4) < 0 ||
        hipMemcpy2D(d_g, sizeof(float) * gp, gray, sizeof(float) * gw, sizeof(float) * gw, gh, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(d_l, lv.data(), sizeof(LevelDesc) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_c, cd.data(), sizeof(Candidate) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_m, &n, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        launch_check(d_g, gp, gw, gh, d_l, d_c, d_m, check_params(), d_w_[1], d_w_[2], d_s2, d_s3, d_r,
                     std::min(2 * n, 2048), nullptr, nullptr, nullptr, nullptr, nullptr, 0, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess ||
        hipMemcpy(out.data(), d_r, sizeof(CheckResult) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -1;
    dfree(d_g); dfree(d_l); dfree(d_c); dfree(d_r); dfree(d_s2); dfree(d_s3); dfree(d_m);
    if (rc < 0) return -1;
    for (int i = 0; i < n; ++i) {
        float* o = res + 8 * i;
        o[0] = (float)out[i].flag; o[1] = (float)out[i].mod; o[2] = (float)out[i].k1; o[3] = (float)out[i].k2;
        o[4] = out[i].m1; o[5] = out[i].m2; o[6] = (float)out[i].n2; o[7] = (float)out[i].n3;
    }
    return 0;
}

}  // namespace ccnn
