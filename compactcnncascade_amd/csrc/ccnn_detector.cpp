// Host side of the MI355X detection path. References are to Bkmz21/CompactCNNCascade
// src/CNNObjectDetector/.
#include "ccnn_detector.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace ccnn {

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "[ccnn] HIP error %s at %s:%d\n", hipGetErrorString(e_),     \
                         __FILE__, __LINE__);                                                \
            return -1;                                                                       \
        }                                                                                    \
    } while (0)

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
        hl % hl_scale != 0 || full_connect) {
        err = "This configuration cnn models is not supported!";
        return -1;
    }
    std::vector<float> hl_w((size_t)hl * connect), hl_b(hl), hl_tw(hl / hl_scale), hl_bnw(hl_scale), hl_bnb(hl_scale);
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
    int io = std::min(index_output, ol - 1);                    // :389-390
    if (io < 0) { err = "Multi-output response maps (index_output < 0) are not supported!"; return -1; }
    if (io == 0) af_scale = -af_scale;

    auto fill = [&](auto shape) -> int {
        using N = decltype(shape);
        if (maps[0] != N::M1 || maps[1] != N::M2 || maps[2] != N::M3 || kw[0] != N::K1 || kh[0] != N::K1 ||
            kw[2] != N::K3W || kh[2] != N::K3H || hl / hl_scale != N::HLG || hl_scale != N::HLS ||
            connect != N::HLN) {
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
        std::copy(hl_w.begin(), hl_w.end(), w.begin() + N::oHW);
        std::copy(hl_b.begin(), hl_b.end(), w.begin() + N::oHB);
        std::copy(hl_tw.begin(), hl_tw.end(), w.begin() + N::oHT);
        std::copy(hl_bnw.begin(), hl_bnw.end(), w.begin() + N::oHBW);
        std::copy(hl_bnb.begin(), hl_bnb.end(), w.begin() + N::oHBB);
        std::copy(ol_w.begin() + (size_t)io * hl, ol_w.begin() + (size_t)(io + 1) * hl, w.begin() + N::oOW);
        w[N::oOB] = ol_b[io];
        w[N::oOB + 1] = ol_tw;
        w[N::oOB + 2] = af_scale;
        return 0;
    };
    if (stage == 0) return fill(Stage1{});
    if (stage == 1) return fill(Stage2{});
    return fill(Stage3{});
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
    if (s.detect_mode == CCNN_DETECT_DISABLE) s.drop_detect = 0; // :961-964
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
Detector::Detector()
{
    ccnn_param_default(&param);
}

Detector::~Detector()
{
    release();
    if (stream_) (void)hipStreamDestroy(stream_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
}

void Detector::release()
{
    auto F = [](auto*& p) { if (p) { (void)hipFree((void*)p); p = nullptr; } };
    auto FH = [](auto*& p) { if (p) { (void)hipHostFree((void*)p); p = nullptr; } };
    F(d_img_); F(d_pre_); F(d_gray_); F(d_resp_); F(d_levels_); F(d_counts_); F(d_offsets_); F(d_meta_);
    F(d_cands_); F(d_results_);
    for (auto& w : d_w_) F(w);
    FH(h_meta_); FH(h_cands_); FH(h_results_);
    cap_w_ = cap_h_ = 0;
    img_bytes_ = resp_cap_ = chunk_cap_ = cand_cap_ = h_cap_ = 0;
    plan_gw_ = plan_gh_ = -1;
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
    if (!stream_) HIP_OK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto& e : ev_)
        if (!e) HIP_OK(hipEventCreate(&e));
    if (param.device_info) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device_) == hipSuccess)
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
    set_ = effective_settings(param);
    for (int s = 0; s < 3; ++s) {
        const unsigned char* b;
        size_t n;
        embedded_model(s, &b, &n);
        std::string err;
        if (load_model(b, n, s, param.index_output[s], net_[s], err) < 0) {
            std::printf("[ccnn] %s\n", err.c_str());
            return -1;
        }
        if (!d_w_[s]) HIP_OK(hipMalloc(&d_w_[s], sizeof(float) * 4096));
        if (net_[s].w.size() > 4096) return -1;
        HIP_OK(hipMemcpy(d_w_[s], net_[s].w.data(), sizeof(float) * net_[s].w.size(), hipMemcpyHostToDevice));
    }
    float sf;
    scales_ = init_scales(param, net_[0].min_h, gray_only_, &pre_scale_, &sf);
    param.scale_factor = sf;                                     // :325-326 writes it back
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
    if (ensure_buffers(param.max_image_width, param.max_image_height) < 0) return -1;
    built_ = true;
    empty_ = false;
    return 0;
}

int Detector::clear()
{
    if (empty_) return -1;
    empty_ = true;
    built_ = false;
    plan_gw_ = plan_gh_ = -1;
    return 0;
}

int Detector::ensure_buffers(int w, int h)
{
    if (w <= cap_w_ && h <= cap_h_ && d_gray_) return 0;
    // device buffers only grow
    const int W = std::max(w, cap_w_), H = std::max(h, cap_h_);
    auto F = [](auto*& p) { if (p) { (void)hipFree((void*)p); p = nullptr; } };
    auto FH = [](auto*& p) { if (p) { (void)hipHostFree((void*)p); p = nullptr; } };
    F(d_img_); F(d_pre_); F(d_gray_); F(d_resp_); F(d_levels_); F(d_counts_); F(d_offsets_); F(d_meta_);
    F(d_cands_); F(d_results_);
    FH(h_meta_); FH(h_cands_); FH(h_results_);
    img_bytes_ = (size_t)W * H * 4;
    HIP_OK(hipMalloc(&d_img_, img_bytes_));
    HIP_OK(hipMalloc(&d_pre_, (size_t)W * H));
    gpitch_ = (W + 63) / 64 * 64;
    HIP_OK(hipMalloc(&d_gray_, sizeof(float) * (size_t)gpitch_ * (H + 1)));
    // response maps of all levels: bounded by the level pixel count / 16 (stride 4) with
    // the largest scale (3x) -> allow 9 * W * H / 16 * levels; compute exactly per plan
    // and grow lazily in plan_frame.
    resp_cap_ = 0;
    HIP_OK(hipMalloc(&d_levels_, sizeof(LevelDesc) * MAX_LEVELS));
    HIP_OK(hipMalloc(&d_meta_, sizeof(int) * 16));
    HIP_OK(hipHostMalloc(&h_meta_, sizeof(int) * 16));
    cap_w_ = W;
    cap_h_ = H;
    plan_gw_ = plan_gh_ = -1;
    return 0;
}

// Level table for a gray image of gw x gh: levels scl_max -> 0 (RunCPUDetect order,
// which is also the RunCheckDetectAsync order), skip rule (:2083-2088), resize mode
// (:2093-2104), response size (getOutputImgSize, cnn_simd_cntk.cpp:997-1012).
int Detector::plan_frame(int gw, int gh)
{
    if (gw == plan_gw_ && gh == plan_gh_) return 0;
    levels_.clear();
    int tiles = 0, chunks = 0;
    size_t roff = 0;
    for (int scl = (int)scales_.size() - 1; scl >= 0; --scl) {
        LevelDesc L{};
        L.scale = scales_[scl];
        L.scl = scl;
        size_mul(gw, gh, L.scale, &L.lw, &L.lh);
        if (L.lw < ext_w_ || L.lh < ext_h_) continue;
        L.mode = (L.scale != 1.f) ? (L.scale > 0.7f ? 0 : 1) : 2;
        if (L.mode != 2 && L.lw == gw && L.lh == gh) L.mode = 2;   // FastImageResize copies equal sizes
        L.sclx = ((uint32_t)gw << 12) / (uint32_t)L.lw + 1;
        L.scly = ((uint32_t)gh << 12) / (uint32_t)L.lh + 1;
        const int c1 = L.lw - 3, r1 = L.lh - 3;
        const int c2 = (c1 >> 1) - 2, r2 = (r1 >> 1) - 2;
        L.rw = (c2 >> 1) - (Stage1::K3W - 1);
        L.rh = (r2 >> 1) - (Stage1::K3H - 1);
        if (L.rw <= 0 || L.rh <= 0) { L.rw = 0; L.rh = 0; }
        L.resp_off = (int)roff;
        roff += (size_t)L.rw * L.rh;
        L.tile_off = tiles;
        L.tiles_x = (L.rw + S1_TX - 1) / S1_TX;
        L.tiles_y = (L.rh + S1_TY - 1) / S1_TY;
        tiles += L.tiles_x * L.tiles_y;
        L.chunk_off = chunks;
        L.chunks = (L.rw * L.rh + CMP_CHUNK - 1) / CMP_CHUNK;
        chunks += L.chunks;
        levels_.push_back(L);
    }
    if (levels_.size() > (size_t)MAX_LEVELS) return -1;
    ntiles_ = tiles;
    nchunks_ = chunks;
    resp_total_ = roff;
    auto F = [](auto*& p) { if (p) { (void)hipFree((void*)p); p = nullptr; } };
    auto FH = [](auto*& p) { if (p) { (void)hipHostFree((void*)p); p = nullptr; } };
    if (roff > resp_cap_) {
        F(d_resp_); F(d_cands_); F(d_results_);
        HIP_OK(hipMalloc(&d_resp_, sizeof(float) * std::max<size_t>(roff, 1)));
        HIP_OK(hipMalloc(&d_cands_, sizeof(Candidate) * std::max<size_t>(roff, 1)));
        HIP_OK(hipMalloc(&d_results_, sizeof(CheckResult) * std::max<size_t>(roff, 1)));
        resp_cap_ = roff;
        cand_cap_ = roff;
    }
    if ((size_t)chunks > chunk_cap_) {
        F(d_counts_); F(d_offsets_);
        HIP_OK(hipMalloc(&d_counts_, sizeof(int) * std::max(chunks, 1)));
        HIP_OK(hipMalloc(&d_offsets_, sizeof(int) * std::max(chunks, 1)));
        chunk_cap_ = chunks;
    }
    const size_t hc = std::min<size_t>(std::max<size_t>(roff, 1), 1u << 20);
    if (hc > h_cap_) {
        FH(h_cands_); FH(h_results_);
        HIP_OK(hipHostMalloc(&h_cands_, sizeof(Candidate) * hc));
        HIP_OK(hipHostMalloc(&h_results_, sizeof(CheckResult) * hc));
        h_cap_ = hc;
    }
    if (!levels_.empty())
        HIP_OK(hipMemcpy(d_levels_, levels_.data(), sizeof(LevelDesc) * levels_.size(), hipMemcpyHostToDevice));
    plan_gw_ = gw;
    plan_gh_ = gh;
    return 0;
}

int Detector::detect(const ccnn_image& img, bool device_ptr, std::vector<ccnn_face>& faces)
{
    faces.clear();
    if (empty_ || !built_) {
        if (build() < 0) {
            std::printf("[CNNDetector] CNN detector no initialized!\n");
            return -1;
        }
    }
    HIP_OK(hipSetDevice(device_));
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
    const uint8_t* src;
    if (device_ptr) {
        src = img.data;
    } else {
        const size_t bytes = img.step * (size_t)(h - 1) + (size_t)w * ch;
        if (bytes > img_bytes_) {
            if (d_img_) (void)hipFree(d_img_);
            d_img_ = nullptr;
            HIP_OK(hipMalloc(&d_img_, bytes));
            img_bytes_ = bytes;
        }
        HIP_OK(hipMemcpyAsync(d_img_, img.data, bytes, hipMemcpyHostToDevice, stream_));
        src = d_img_;
    }
    return run_frame(src, w, h, ch, img.step, faces);
}

int Detector::run_frame(const uint8_t* src, int w, int h, int ch, size_t step, std::vector<ccnn_face>& faces)
{
    hipStream_t st = stream_;
    int gw = w, gh = h, gch = ch;
    size_t gstep = step;
    if (profiling_) HIP_OK(hipEventRecord(ev_[0], st));
    if (gray_only_ && pre_scale_ > 0.f && ch == 1) {
        // 8-bit bilinear pre-resize of a gray frame (cnn_detector_v3.cpp:2531-2539)
        size_mul(w, h, pre_scale_, &gw, &gh);
        HIP_OK(launch_resize_u8(src, w, h, (int)step, d_pre_, gw, gh, st));
        src = d_pre_;
        gstep = gw;
        gch = 1;
    }
    HIP_OK(launch_gray(src, gw, gh, gch, (int)gstep, d_gray_, gpitch_, st));
    last_gw_ = gw;
    last_gh_ = gh;
    if (plan_frame(gw, gh) < 0) return -1;
    const int nl = (int)levels_.size();
    if (profiling_) HIP_OK(hipEventRecord(ev_[1], st));
    HIP_OK(launch_stage1(d_gray_, gpitch_, gw, gh, d_levels_, nl, ntiles_, d_w_[0], d_resp_, st));
    if (profiling_) HIP_OK(hipEventRecord(ev_[2], st));
    HIP_OK(launch_compact(d_resp_, d_levels_, nl, nchunks_, set_.t1, d_counts_, d_offsets_, d_meta_, d_cands_, st));
    if (profiling_) HIP_OK(hipEventRecord(ev_[3], st));
    const bool checks = set_.detect_mode != CCNN_DETECT_DISABLE;
    if (checks && nl > 0) {
        CheckParams P;
        P.t2 = set_.t2; P.t3 = set_.t3; P.type_check = set_.type_check; P.min_neighbors = set_.min_neighbors;
        P.adapt = set_.adapt; P.double_check = set_.double_check; P.equalize = set_.equalize;
        P.reflection = set_.reflection;
        HIP_OK(launch_check(d_gray_, gpitch_, gw, gh, d_levels_, d_cands_, d_meta_, P, d_w_[1], d_w_[2], d_results_,
                            check_blocks_, st));
    }
    if (profiling_) HIP_OK(hipEventRecord(ev_[4], st));
    // read back the candidate list and the check results
    HIP_OK(hipMemcpyAsync(h_meta_, d_meta_, sizeof(int), hipMemcpyDeviceToHost, st));
    const size_t pre = std::min<size_t>(h_cap_, 4096);
    HIP_OK(hipMemcpyAsync(h_cands_, d_cands_, sizeof(Candidate) * pre, hipMemcpyDeviceToHost, st));
    if (checks) HIP_OK(hipMemcpyAsync(h_results_, d_results_, sizeof(CheckResult) * pre, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const int count = nl > 0 ? h_meta_[0] : 0;
    std::vector<Candidate> big_c;
    std::vector<CheckResult> big_r;
    const Candidate* cands = h_cands_;
    const CheckResult* res = h_results_;
    if ((size_t)count > pre) {
        if ((size_t)count <= h_cap_) {
            HIP_OK(hipMemcpy(h_cands_ + pre, d_cands_ + pre, sizeof(Candidate) * (count - pre), hipMemcpyDeviceToHost));
            if (checks)
                HIP_OK(hipMemcpy(h_results_ + pre, d_results_ + pre, sizeof(CheckResult) * (count - pre),
                                 hipMemcpyDeviceToHost));
        } else {
            big_c.resize(count);
            HIP_OK(hipMemcpy(big_c.data(), d_cands_, sizeof(Candidate) * count, hipMemcpyDeviceToHost));
            cands = big_c.data();
            if (checks) {
                big_r.resize(count);
                HIP_OK(hipMemcpy(big_r.data(), d_results_, sizeof(CheckResult) * count, hipMemcpyDeviceToHost));
                res = big_r.data();
            }
        }
    }
    if (profiling_) {
        for (int i = 0; i < 4; ++i) (void)hipEventElapsedTime(&last_ms_[i == 0 ? 0 : i], ev_[i], ev_[i + 1]);
        (void)hipEventElapsedTime(&last_ms_[4], ev_[0], ev_[4]);
    }
    last_count_ = count;
    last_levels_ = levels_;
    last_cands_.assign(cands, cands + count);
    if (checks) last_results_.assign(res, res + count);
    else last_results_.clear();

    // Ordered replay of RunCheckDetect / CPUCheckDetect (cnn_detector_v3.cpp:1842-2011,
    // 1485-1818): the scan-time DropDetection against all detections so far, then per
    // surviving point the drop against this level's detections and the precomputed
    // stage-2/3 outcome.
    std::vector<Det> rects;
    std::vector<int> points;
    int ci = 0;
    for (int l = 0; l < nl; ++l) {
        const LevelDesc& L = levels_[l];
        const float scale = L.scale;
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
                    if (drop) continue;
                }
                const CheckResult& r = res[pi];
                if (r.flag) {
                    Det d;
                    d.r = nr;
                    // :1758
                    d.score = c.score + (float)r.k1 * std::max(-1.7159f, r.m1) + (float)r.k2 * std::max(-1.7159f, r.m2);
                    d.scale = scale;
                    d.knn = std::min(r.k1, r.k2);
                    rects.push_back(d);
                }
            }
        } else {
            for (int pi : points) {
                const Candidate& c = cands[pi];
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
    std::vector<Det> dets = post_process(rects, set_);
    const bool rescale = gray_only_ && pre_scale_ > 0.f && ch == 1;
    const float s = rescale ? (float)param.min_face_height / (float)net_[0].min_h : 1.f;   // :2783-2806
    faces.reserve(dets.size());
    for (const Det& d : dets) {
        Rect r = d.r;
        if (rescale)
            r = Rect((int)((float)r.x * s), (int)((float)r.y * s), (int)((float)r.w * s), (int)((float)r.h * s));
        ccnn_face f;
        f.x = r.x; f.y = r.y; f.width = r.w; f.height = r.h; f.score = d.score;
        f.gender = -1; f.smile = -1; f.glasses = -1;
        faces.push_back(f);
    }
    return (int)faces.size();
}

int Detector::gray_image(ccnn_image* img, int pipeline)
{
    if (!img || empty_ || !d_gray_ || last_gw_ <= 0) return -1;
    if (pipeline == CCNN_PIPELINE_CPU) {
        h_gray_.resize((size_t)gpitch_ * last_gh_);
        HIP_OK(hipMemcpy(h_gray_.data(), d_gray_, sizeof(float) * h_gray_.size(), hipMemcpyDeviceToHost));
        img->data = (unsigned char*)h_gray_.data();
    } else {
        img->data = (unsigned char*)d_gray_;
    }
    img->cols = last_gw_;
    img->rows = last_gh_;
    img->channels = 1;
    img->step = (size_t)gpitch_;
    return 0;
}

int Detector::last_timing(float* ms) const
{
    for (int i = 0; i < 5; ++i) ms[i] = last_ms_[i];
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
    int* d_m = nullptr;
    std::vector<LevelDesc> lv(n);
    std::vector<Candidate> cd(n);
    for (int i = 0; i < n; ++i) {
        lv[i] = LevelDesc{};
        lv[i].scale = scale[i];
        cd[i].level = i; cd[i].x = x[i]; cd[i].y = y[i]; cd[i].score = 0.f;
    }
    int rc = 0;
    CheckParams P;
    P.t2 = set_.t2; P.t3 = set_.t3; P.type_check = set_.type_check; P.min_neighbors = set_.min_neighbors;
    P.adapt = set_.adapt; P.double_check = set_.double_check; P.equalize = set_.equalize;
    P.reflection = set_.reflection;
    std::vector<CheckResult> out(n);
    if (hipMalloc(&d_g, sizeof(float) * (size_t)gp * (gh + 1)) != hipSuccess ||
        hipMalloc(&d_l, sizeof(LevelDesc) * n) != hipSuccess || hipMalloc(&d_c, sizeof(Candidate) * n) != hipSuccess ||
        hipMalloc(&d_r, sizeof(CheckResult) * n) != hipSuccess || hipMalloc(&d_m, sizeof(int) * 4) != hipSuccess ||
        hipMemcpy2D(d_g, sizeof(float) * gp, gray, sizeof(float) * gw, sizeof(float) * gw, gh, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemcpy(d_l, lv.data(), sizeof(LevelDesc) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_c, cd.data(), sizeof(Candidate) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_m, &n, sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        launch_check(d_g, gp, gw, gh, d_l, d_c, d_m, P, d_w_[1], d_w_[2], d_r, std::min(n, 2048), stream_) !=
            hipSuccess ||
        hipStreamSynchronize(stream_) != hipSuccess ||
        hipMemcpy(out.data(), d_r, sizeof(CheckResult) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -1;
    (void)hipFree(d_g); (void)hipFree(d_l); (void)hipFree(d_c); (void)hipFree(d_r); (void)hipFree(d_m);
    if (rc < 0) return -1;
    for (int i = 0; i < n; ++i) {
        float* o = res + 8 * i;
        o[0] = (float)out[i].flag; o[1] = (float)out[i].mod; o[2] = (float)out[i].k1; o[3] = (float)out[i].k2;
        o[4] = out[i].m1; o[5] = out[i].m2; o[6] = (float)out[i].n2; o[7] = (float)out[i].n3;
    }
    return 0;
}

}  // namespace ccnn
