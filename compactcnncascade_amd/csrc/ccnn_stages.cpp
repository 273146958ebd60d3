// Stage-level entry points of the C ABI (include/ccnn.h): each runs one kernel of the
// detection path on host inputs and copies the result back.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ccnn.h"
#include "ccnn_detector.h"

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool alloc(size_t n) { return hipMalloc(&p, n) == hipSuccess; }
    template <class T> T* as() { return static_cast<T*>(p); }
};

bool have_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        std::printf("[ccnn] no HIP device available\n");
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

int ccnn_stage_gray(const ccnn_image* img, float* out)
{
    if (!img || !img->data || !out || img->cols <= 0 || img->rows <= 0) return -1;
    if (img->channels != 1 && img->channels != 3 && img->channels != 4) return -1;
    if (!have_device()) return -1;
    const size_t bytes = img->step * (size_t)(img->rows - 1) + (size_t)img->cols * img->channels;
    // the detector's gray layout: pitch rounded up to 64 floats
    const int gp = (img->cols + 63) / 64 * 64;
    DevBuf src, dst;
    if (!src.alloc(bytes) || !dst.alloc(sizeof(float) * (size_t)gp * img->rows)) return -1;
    if (hipMemcpy(src.p, img->data, bytes, hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (ccnn::launch_gray(src.as<uint8_t>(), img->cols, img->rows, img->channels, (int)img->step, dst.as<float>(),
                          gp, nullptr) != hipSuccess)
        return -1;
    if (hipMemcpy2D(out, sizeof(float) * (size_t)img->cols, dst.p, sizeof(float) * (size_t)gp,
                    sizeof(float) * (size_t)img->cols, img->rows, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    return 0;
}

int ccnn_stage_level(const float* gray, int gw, int gh, int lw, int lh, int mode, float* out)
{
    if (!gray || !out || gw < 2 || gh < 2 || lw <= 0 || lh <= 0 || mode < 0 || mode > 2) return -1;
    if (mode == 2 && (lw != gw || lh != gh)) return -1;
    if (!have_device()) return -1;
    DevBuf g, o;
    if (!g.alloc(sizeof(float) * (size_t)gw * (gh + 1)) || !o.alloc(sizeof(float) * (size_t)lw * lh)) return -1;
    if (hipMemcpy(g.p, gray, sizeof(float) * (size_t)gw * gh, hipMemcpyHostToDevice) != hipSuccess) return -1;
    ccnn::LevelDesc L{};
    L.lw = lw; L.lh = lh; L.mode = mode;
    L.sclx = ((uint32_t)gw << 12) / (uint32_t)lw + 1;
    L.scly = ((uint32_t)gh << 12) / (uint32_t)lh + 1;
    if (ccnn::launch_level(g.as<float>(), gw, gw, gh, L, o.as<float>(), nullptr) != hipSuccess) return -1;
    if (hipMemcpy(out, o.p, sizeof(float) * (size_t)lw * lh, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

int ccnn_stage1_size(int lw, int lh, int* rw, int* rh)
{
    if (!rw || !rh) return -1;
    const int c2 = ((lw - 3) >> 1) - 2, r2 = ((lh - 3) >> 1) - 2;
    *rw = (c2 >> 1) - (ccnn::Stage1::K3W - 1);
    *rh = (r2 >> 1) - (ccnn::Stage1::K3H - 1);
    return (*rw > 0 && *rh > 0) ? 0 : -1;
}

int ccnn_selftest_tanh(unsigned int lo, unsigned int hi, unsigned long long* out)
{
    if (!out || hi < lo) return -1;
    if (!have_device()) return -1;
    DevBuf d;
    if (!d.alloc(3 * sizeof(unsigned long long))) return -1;
    if (hipMemset(d.p, 0, 3 * sizeof(unsigned long long)) != hipSuccess) return -1;
    if (ccnn::launch_selftest_tanh(lo, hi, d.as<unsigned long long>(), nullptr) != hipSuccess) return -1;
    if (hipMemcpy(out, d.p, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

int ccnn_stage_landmarks(const float* in, float* out)
{
    if (!in || !out) return -1;
    if (!have_device()) return -1;
    ccnn::NetWeights net;
    const unsigned char* b;
    size_t n;
    ccnn::embedded_model(3, &b, &n);
    std::string err;
    if (ccnn::load_model(b, n, 3, -1, net, err) < 0) return -1;
    DevBuf i, w, o;
    if (!i.alloc(sizeof(float) * 49 * 51) || !w.alloc(sizeof(float) * net.w.size()) ||
        !o.alloc(sizeof(float) * ccnn::Landmarks::NOUT))
        return -1;
    if (hipMemcpy(i.p, in, sizeof(float) * 49 * 51, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(w.p, net.w.data(), sizeof(float) * net.w.size(), hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    if (ccnn::launch_landmarks(i.as<float>(), w.as<float>(), o.as<float>(), nullptr) != hipSuccess) return -1;
    if (hipMemcpy(out, o.p, sizeof(float) * ccnn::Landmarks::NOUT, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

int ccnn_stage23(int stage, const float* in, int n, float* out)
{
    return ccnn_stage23_ex(stage, in, n, 0, out);
}

int ccnn_stage23_ex(int stage, const float* in, int n, int precision, float* out)
{
    if ((stage != 1 && stage != 2) || n < 0 || (n > 0 && (!in || !out)) || precision < 0 || precision > 2) return -1;
    if (n == 0) return 0;
    if (!have_device()) return -1;
    ccnn::NetWeights net;
    const unsigned char* b;
    size_t nb;
    ccnn::embedded_model(stage, &b, &nb);
    std::string err;
    // index_output defaults of FaceDetector::Param (CompactCNNLib.h:160)
    if (ccnn::load_model(b, nb, stage, 0, net, err) < 0) return -1;
    std::vector<float> mw;
    ccnn::mfma_weights(net.w, mw);
    DevBuf i, w, o, m;
    const size_t ni = (size_t)n * 47 * 51;
    if (!i.alloc(sizeof(float) * ni) || !w.alloc(sizeof(float) * net.w.size()) || !o.alloc(sizeof(float) * 16 * n) ||
        !m.alloc(sizeof(float) * mw.size()))
        return -1;
    if (hipMemcpy(i.p, in, sizeof(float) * ni, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(w.p, net.w.data(), sizeof(float) * net.w.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m.p, mw.data(), sizeof(float) * mw.size(), hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    if (ccnn::launch_net23(stage, i.as<float>(), n, w.as<float>(), o.as<float>(), nullptr, precision,
                           m.as<float>()) != hipSuccess)
        return -1;
    if (hipMemcpy(out, o.p, sizeof(float) * 16 * n, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

int ccnn_stage1(const float* level, int lw, int lh, float* out)
{
    return ccnn_stage1_ex(level, lw, lh, 0, out);
}

int ccnn_stage1_ex(const float* level, int lw, int lh, int precision, float* out)
{
    if (precision < 0 || precision > 2) return -1;
    int rw, rh;
    if (!level || !out || ccnn_stage1_size(lw, lh, &rw, &rh) < 0) return -1;
    if (lw < 23 || lh < 27) return -1;
    if (!have_device()) return -1;
    ccnn::NetWeights net;
    const unsigned char* b;
    size_t n;
    ccnn::embedded_model(0, &b, &n);
    std::string err;
    if (ccnn::load_model(b, n, 0, 1, net, err) < 0) return -1;
    // the level image is fed through the copy mode of the fused sampler
    ccnn::LevelDesc L{};
    L.lw = lw; L.lh = lh; L.rw = rw; L.rh = rh; L.mode = 2; L.scale = 1.f;
    L.tiles_x = (rw + ccnn::S1_TX - 1) / ccnn::S1_TX;
    L.tiles_y = (rh + ccnn::S1_TY - 1) / ccnn::S1_TY;
    const size_t s1n = ccnn::level_l1_layout(L);
    DevBuf g, w, r, l, s1;
    if (!g.alloc(sizeof(float) * (size_t)lw * (lh + 1)) || !w.alloc(sizeof(float) * net.w.size()) ||
        !r.alloc(sizeof(float) * (size_t)rw * rh) || !l.alloc(sizeof(L)) || !s1.alloc(sizeof(float) * s1n))
        return -1;
    if (hipMemset(s1.p, 0, sizeof(float) * s1n) != hipSuccess ||
        hipMemcpy(g.p, level, sizeof(float) * (size_t)lw * lh, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(w.p, net.w.data(), sizeof(float) * net.w.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(l.p, &L, sizeof(L), hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    if (ccnn::launch_stage1(g.as<float>(), lw, lw, lh, l.as<ccnn::LevelDesc>(), 1, L.l1_tiles_x * L.l1_tiles_y,
                            L.tiles_x * L.tiles_y, w.as<float>(), s1.as<float>(), r.as<float>(), INFINITY, nullptr,
                            nullptr, nullptr, 0, nullptr, nullptr, nullptr, 0, precision > 0) != hipSuccess)
        return -1;
    if (hipMemcpy(out, r.p, sizeof(float) * (size_t)rw * rh, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    return 0;
}

}  // extern "C"
