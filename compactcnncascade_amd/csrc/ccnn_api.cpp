// C ABI (include/ccnn.h) and the C++ FaceDetector facade (include/CompactCNNLib.h).
// Reference: Bkmz21/CompactCNNCascade src/CNNObjectDetector/CompactCNNLibAPI_v2.cpp.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <mutex>
#include <vector>

#include "CompactCNNLib.h"
#include "ccnn.h"
#include "ccnn_detector.h"

struct ccnn_detector {
    ccnn::Detector det;
    std::vector<ccnn_face> faces;
};

extern "C" {

void ccnn_param_default(ccnn_param* p)
{
    if (!p) return;
    p->max_image_width = 1920;
    p->max_image_height = 1080;
    p->min_face_height = 80;
    p->max_face_height = 0;
    p->scale_factor = 1.15f;
    p->min_neighbors = 2;
    p->detect_precision = CCNN_PRECISION_NORMAL;
    p->treshold[0] = -1.5f;
    p->treshold[1] = -1.25f;
    p->treshold[2] = -1.0f;
    p->drop_detect = 1;
    p->equalize = 1;
    p->reflection = 1;
    p->pipeline = CCNN_PIPELINE_GPU;
    p->detect_mode = CCNN_DETECT_ASYNC;
    p->num_threads = 2;
    p->models[0] = p->models[1] = p->models[2] = nullptr;
    p->index_output[0] = 1;
    p->index_output[1] = 0;
    p->index_output[2] = 0;
    p->device_info = 1;
    p->cuda_device_id = 0;
    p->cl_platform_id = -1;
    p->cl_device_id = -1;
    p->facial_analysis = 0;
}

ccnn_detector* ccnn_create(void)
{
    return new (std::nothrow) ccnn_detector();
}

void ccnn_destroy(ccnn_detector* d)
{
    delete d;
}

int ccnn_init(ccnn_detector* d, const ccnn_param* p)
{
    if (!d || !p) return -1;
    if (!d->det.empty()) d->det.clear();
    if (d->det.init(*p) < 0) {
        std::printf("[CompactCNNLibAPI] Error loading FaceDetector!\n");
        d->det.clear();
        return -1;
    }
    return 0;
}

static int detect_common(ccnn_detector* d, ccnn_face* faces, int capacity, const ccnn_image* img, bool device)
{
    if (!d || d->det.empty()) {
        std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
        return -1;
    }
    if (!faces || !img) return -1;
    // The reference's Detect ignores the detector's return code and reports the faces it
    // collected (CompactCNNLibAPI_v2.cpp:236-322): an unsupported image gives 0 faces.
    if (d->det.detect(*img, device, d->faces) < 0) return 0;
    const int n = (int)d->faces.size();
    const int m = capacity < 0 ? n : (n < capacity ? n : capacity);
    for (int i = 0; i < m; ++i) faces[i] = d->faces[i];
    return n;
}

int ccnn_detect(ccnn_detector* d, ccnn_face* faces, int capacity, const ccnn_image* img)
{
    return detect_common(d, faces, capacity, img, false);
}

int ccnn_detect_device(ccnn_detector* d, ccnn_face* faces, int capacity, const ccnn_image* img)
{
    return detect_common(d, faces, capacity, img, true);
}

int ccnn_detect_batch(ccnn_detector* d, int n, const ccnn_image* imgs, int device, ccnn_face* faces, int capacity,
                      int* counts)
{
    if (!d || d->det.empty()) {
        std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
        return -1;
    }
    if (n < 0 || (n > 0 && (!imgs || !counts || (capacity > 0 && !faces)))) return -1;
    std::vector<std::vector<ccnn_face>> out;
    if (d->det.detect_batch(n, imgs, device != 0, out) < 0) return -1;
    for (int i = 0; i < n; ++i) {
        counts[i] = (int)out[i].size();
        const int m = std::min(counts[i], std::max(capacity, 0));
        for (int k = 0; k < m; ++k) faces[(size_t)i * capacity + k] = out[i][k];
    }
    return 0;
}

int ccnn_set_batch_mode(ccnn_detector* d, int mode)
{
    if (!d || mode < 0 || mode > 1) return -1;
    d->det.set_batch_mode(mode);
    return 0;
}

int ccnn_set_compute_precision(ccnn_detector* d, int precision)
{
    if (!d) return -1;
    return d->det.set_precision(precision);
}

int ccnn_get_compute_precision(const ccnn_detector* d)
{
    return d ? d->det.precision() : -1;
}

int ccnn_set_skipnet(ccnn_detector* d, const float* params, float threshold, int min_face, float scale_factor)
{
    if (!d) return -1;
    return d->det.set_skipnet(params, threshold, min_face, scale_factor);
}

int ccnn_phase_stamps(int kernel, unsigned long long* out, int nblocks)
{
    return ccnn::read_phase_stamps(kernel, out, nblocks);
}

int ccnn_last_batch_timing(const ccnn_detector* d, float* ms)
{
    if (!d || !ms) return -1;
    d->det.batch_timing(ms);
    return 0;
}

int ccnn_set_check_grid(ccnn_detector* d, int frame_per_cu, int batch_per_cu)
{
    if (!d) return -1;
    return d->det.set_check_grid(frame_per_cu, batch_per_cu);
}

int ccnn_set_skip_slot_thresholds(ccnn_detector* d, const float* thresholds)
{
    if (!d || !thresholds) return -1;
    return d->det.set_skip_slot_thresholds(thresholds);
}

int ccnn_set_skip_window(ccnn_detector* d, int slot_radius, int cell_margin)
{
    if (!d) return -1;
    return d->det.set_skip_window(slot_radius, cell_margin);
}

int ccnn_skip_stats(ccnn_detector* d, long long* tiles, long long* kept)
{
    if (!d || !tiles || !kept) return -1;
    d->det.skip_stats(tiles, kept);
    return 0;
}

int ccnn_set_tile_shape(ccnn_detector* d, int shape)
{
    if (!d) return -1;
    return d->det.set_tile_shape(shape);
}

int ccnn_get_tile_shape(const ccnn_detector* d, int* ax, int* ay, int* tx, int* ty)
{
    if (!d) return -1;
    const int id = d->det.tile_shape_id();
    const ccnn::TileShape t = ccnn::tile_shape(id);
    if (ax) *ax = t.ax;
    if (ay) *ay = t.ay;
    if (tx) *tx = t.tx;
    if (ty) *ty = t.ty;
    return id;
}

int ccnn_autotune(ccnn_detector* d, int n, const ccnn_image* imgs, int device, int reps, float* shape_ms,
                  float* grid_ms)
{
    if (!d || d->det.empty() || n <= 0 || !imgs || reps <= 0) return -1;
    return d->det.autotune(n, imgs, device != 0, reps, shape_ms, grid_ms);
}

int ccnn_batch_stats(const ccnn_detector* d, int* launches, int* captures)
{
    if (!d) return -1;
    if (launches) *launches = d->det.batch_launches();
    if (captures) *captures = d->det.batch_captures();
    return 0;
}

int ccnn_clear(ccnn_detector* d)
{
    if (!d || d->det.empty()) return -1;
    return d->det.clear();
}

int ccnn_is_empty(const ccnn_detector* d)
{
    return !d || d->det.empty();
}

int ccnn_get_param(const ccnn_detector* d, ccnn_param* p)
{
    if (!d || !p) return -1;
    if (d->det.empty()) {
        std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
        ccnn_param_default(p);
        return -1;
    }
    *p = d->det.param;
    return 0;
}

void ccnn_set_param(ccnn_detector* d, const ccnn_param* p)
{
    if (!d || !p) return;
    if (d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return; }
    d->det.init(*p);
}

int ccnn_get_max_image_size(const ccnn_detector* d, int* w, int* h)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return -1; }
    if (w) *w = d->det.param.max_image_width;
    if (h) *h = d->det.param.max_image_height;
    return 0;
}

// The setters change one parameter and re-initialise lazily, as CNNDetector's setters
// Clear() and the next Detect re-Inits (cnn_detector_v3.h:404-436, cnn_detector_v3.cpp:2485-2492);
// gray_image_only keeps its state (only Init and setParam force it on).
void ccnn_set_max_image_size(ccnn_detector* d, int w, int h)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return; }
    ccnn_param p = d->det.param;
    p.max_image_width = w;
    p.max_image_height = h;
    d->det.set_fields(p);
}

int ccnn_get_min_object_height(const ccnn_detector* d)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return 0; }
    return d->det.param.min_face_height;
}

void ccnn_set_min_object_height(ccnn_detector* d, int v)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return; }
    ccnn_param p = d->det.param;
    p.min_face_height = v;
    d->det.set_fields(p);
}

int ccnn_get_max_object_height(const ccnn_detector* d)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return 0; }
    return d->det.param.max_face_height;
}

void ccnn_set_max_object_height(ccnn_detector* d, int v)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return; }
    ccnn_param p = d->det.param;
    p.max_face_height = v;
    d->det.set_fields(p);
}

int ccnn_get_num_threads(const ccnn_detector* d)
{
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return 0; }
    return d->det.param.num_threads;
}

void ccnn_set_num_threads(ccnn_detector* d, int n)
{
    // host threads do not take part in the GPU path; the value is recorded for getParam
    if (!d || d->det.empty()) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return; }
    d->det.param.num_threads = n;
}

int ccnn_get_gray_image(ccnn_detector* d, ccnn_image* img, int pipeline)
{
    if (!d) return -1;
    return d->det.gray_image(img, pipeline);
}

int ccnn_stage_check(const ccnn_param* p, const float* gray, int gw, int gh, int n, const float* scale, const int* x,
                     const int* y, float* res)
{
    if (!p || !gray || n < 0 || (n > 0 && (!scale || !x || !y || !res))) return -1;
    ccnn::Detector det;
    ccnn_param q = *p;
    q.device_info = 0;
    if (det.init(q) < 0) return -1;
    return det.stage_check(gray, gw, gh, n, scale, x, y, res);
}

int ccnn_last_trace(const ccnn_detector* d, int* n_levels, int* levels, int cap, float* cand, float* res)
{
    if (!d) return -1;
    return d->det.last_trace(n_levels, levels, cap, cand, res);
}

long ccnn_last_responses(ccnn_detector* d, float* out, size_t cap)
{
    if (!d) return -1;
    return d->det.last_responses(out, cap);
}

int ccnn_last_status(const ccnn_detector* d, int* out, int cap)
{
    if (!d) return -1;
    return d->det.last_status(out, cap);
}

int ccnn_last_facial(const ccnn_detector* d, int* out, int cap)
{
    if (!d) return -1;
    return d->det.last_facial(out, cap);
}

int ccnn_last_error(int reset)
{
    return ccnn::last_error(reset != 0);
}

void ccnn_set_profiling(ccnn_detector* d, int on)
{
    if (d) d->det.set_profiling(on != 0);
}

int ccnn_last_timing(const ccnn_detector* d, float* ms)
{
    if (!d || !ms) return -1;
    return d->det.last_timing(ms);
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// C++ facade: one process-wide detector and mutex (CompactCNNLibAPI_v2.cpp:32-33).
// ---------------------------------------------------------------------------------
namespace CompactCNNLib
{
namespace {
ccnn_detector* g_det = nullptr;
std::mutex g_mutex;

ccnn_param to_c(const FaceDetector::Param& p)
{
    ccnn_param c;
    ccnn_param_default(&c);
    c.max_image_width = p.max_image_size.width;
    c.max_image_height = p.max_image_size.height;
    c.min_face_height = p.min_face_height;
    c.max_face_height = p.max_face_height;
    c.scale_factor = p.scale_factor;
    c.min_neighbors = p.min_neighbors;
    c.detect_precision = (int)p.detect_precision;
    for (int i = 0; i < 3; ++i) {
        c.treshold[i] = p.treshold[i];
        c.models[i] = p.models[i];
        c.index_output[i] = p.index_output[i];
    }
    c.drop_detect = p.drop_detect;
    c.equalize = p.equalize;
    c.reflection = p.reflection;
    c.pipeline = (int)p.pipeline;
    c.detect_mode = (int)p.detect_mode;
    c.num_threads = p.num_threads;
    c.device_info = p.device_info;
    c.cuda_device_id = p.cuda_device_id;
    c.cl_platform_id = p.cl_platform_id;
    c.cl_device_id = p.cl_device_id;
    c.facial_analysis = p.facial_analysis;
    return c;
}

FaceDetector::Param from_c(const ccnn_param& c)
{
    FaceDetector::Param p;
    p.max_image_size = FaceDetector::Size(c.max_image_width, c.max_image_height);
    p.min_face_height = c.min_face_height;
    p.max_face_height = c.max_face_height;
    p.scale_factor = c.scale_factor;
    p.min_neighbors = c.min_neighbors;
    p.detect_precision = (FaceDetector::DetectPrecision)c.detect_precision;
    for (int i = 0; i < 3; ++i) {
        p.treshold[i] = c.treshold[i];
        p.models[i] = c.models[i];
        p.index_output[i] = c.index_output[i];
    }
    p.drop_detect = c.drop_detect != 0;
    p.equalize = c.equalize != 0;
    p.reflection = c.reflection != 0;
    p.pipeline = (FaceDetector::Pipeline)c.pipeline;
    p.detect_mode = (FaceDetector::DetectMode)c.detect_mode;
    p.num_threads = c.num_threads;
    p.device_info = c.device_info != 0;
    p.cuda_device_id = c.cuda_device_id;
    p.cl_platform_id = c.cl_platform_id;
    p.cl_device_id = c.cl_device_id;
    p.facial_analysis = c.facial_analysis != 0;
    return p;
}
}  // namespace

FaceDetector::FaceDetector()
{
    std::printf("[CompactCNNLibAPI] CNN Cascade, MI355X (HIP) implementation\n");
}

FaceDetector::~FaceDetector()
{
    Clear();
}

int FaceDetector::Init(Param& param)
{
    if (!isEmpty()) Clear();
    std::lock_guard<std::mutex> lk(g_mutex);
    ccnn_param c = to_c(param);
    if (!g_det) g_det = ccnn_create();
    if (!g_det || ccnn_init(g_det, &c) < 0) {
        ccnn_destroy(g_det);
        g_det = nullptr;
        return -1;
    }
    return 0;
}

int FaceDetector::Detect(Face* faces, ImageData& img)
{
    if (isEmpty()) {
        std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
        return -1;
    }
    if (faces == nullptr) return -1;
    std::lock_guard<std::mutex> lk(g_mutex);
    ccnn_image im{img.cols, img.rows, img.channels, img.data, img.step};
    // the reference writes every face without a capacity (CompactCNNLibAPI_v2.cpp:291-322)
    const int n = ccnn_detect(g_det, reinterpret_cast<ccnn_face*>(faces), -1, &im);
    static_assert(sizeof(Face) == sizeof(ccnn_face), "Face layout");
    return n;
}

int FaceDetector::Clear()
{
    if (isEmpty()) return -1;
    std::lock_guard<std::mutex> lk(g_mutex);
    ccnn_destroy(g_det);
    g_det = nullptr;
    return 0;
}

bool FaceDetector::isEmpty() const
{
    return g_det == nullptr;
}

FaceDetector::Param FaceDetector::getParam() const
{
    std::lock_guard<std::mutex> lk(g_mutex);
    ccnn_param c;
    ccnn_param_default(&c);
    if (g_det) ccnn_get_param(g_det, &c);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
    return from_c(c);
}

void FaceDetector::setParam(Param& param)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    ccnn_param c = to_c(param);
    if (g_det) ccnn_set_param(g_det, &c);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
}

FaceDetector::Size FaceDetector::getMaxImageSize() const
{
    std::lock_guard<std::mutex> lk(g_mutex);
    int w = 0, h = 0;
    if (g_det) ccnn_get_max_image_size(g_det, &w, &h);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
    return Size(w, h);
}

void FaceDetector::setMaxImageSize(Size s)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_det) ccnn_set_max_image_size(g_det, s.width, s.height);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
}

int FaceDetector::getMinObjectHeight() const
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_det) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return 0; }
    return ccnn_get_min_object_height(g_det);
}

void FaceDetector::setMinObjectHeight(int v)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_det) ccnn_set_min_object_height(g_det, v);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
}

int FaceDetector::getMaxObjectHeight() const
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_det) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return 0; }
    return ccnn_get_max_object_height(g_det);
}

void FaceDetector::setMaxObjectHeight(int v)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_det) ccnn_set_max_object_height(g_det, v);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
}

int FaceDetector::getNumThreads() const
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_det) { std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n"); return 0; }
    return ccnn_get_num_threads(g_det);
}

void FaceDetector::setNumThreads(int n)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_det) ccnn_set_num_threads(g_det, n);
    else std::printf("[CompactCNNLibAPI] FaceDetector no initialized!\n");
}

int FaceDetector::getGrayImage32F(ImageData* img, Pipeline pipeline)
{
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_det || !img) return -1;
    ccnn_image c{};
    if (ccnn_get_gray_image(g_det, &c, (int)pipeline) < 0) return -1;
    img->cols = c.cols;
    img->rows = c.rows;
    img->channels = c.channels;
    img->data = c.data;
    img->step = c.step;
    return 0;
}

void FaceDetector::CNTKDump2Binary(const char* binary_file, const char* cntk_model_dump)
{
    ccnn_cntk_dump_to_binary(binary_file, cntk_model_dump);
}

}  // namespace CompactCNNLib
