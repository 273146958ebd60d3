// Host side of the MI355X detection path: model loading, pyramid plan, device buffers,
// kernel dispatch and the ordered replay of the reference's order-dependent decisions.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "ccnn.h"
#include "ccnn_internal.h"

namespace ccnn {

// kernel launchers (ccnn_kernels.hip)
hipError_t launch_gray(const uint8_t* img, int w, int h, int ch, int step, float* gray, int gp, hipStream_t st);
hipError_t launch_resize_u8(const uint8_t* src, int sw, int sh, int sstep, uint8_t* dst, int dw, int dh,
                            hipStream_t st);
hipError_t launch_level(const float* gray, int gp, int gw, int gh, const LevelDesc& L, float* out, hipStream_t st);
hipError_t launch_stage1(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, int nlevels,
                         int nl1tiles, int ntiles, const float* W1, float* s1buf, float* resp, float t1,
                         int* meta, Candidate* cands, Candidate* hcands, int hcap, const int2* l1_order,
                         hipStream_t st, const hipEvent_t* kev = nullptr);
// L1-sum plane geometry and L1 tiles of a level whose response tiles are set
// (LevelDesc::tiles_x / tiles_y); returns the floats the level's planes take.
size_t level_l1_layout(LevelDesc& L);
hipError_t launch_check(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, const Candidate* cands,
                        const int* meta, const CheckParams& P, const float* W2, const float* W3, int2* s2res,
                        int2* s3res, CheckResult* results, int nblocks, int* hmeta, Candidate* hcands,
                        CheckResult* hresults, int2* hs2, int2* hs3, int hcap, hipStream_t st,
                        const hipEvent_t* kev = nullptr);
// k_facial takes the check decisions from the stage-2/3 outcomes itself and writes the facial
// results of the accepted candidates to results[].fa (and hresults[].fa below hcap)
hipError_t launch_facial(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, const Candidate* cands,
                         const int* meta, const CheckParams& P, const int2* s2res, const int2* s3res, const float* W,
                         CheckResult* results, CheckResult* hresults, int hcap, int nblocks, hipStream_t st);
hipError_t launch_landmarks(const float* in, const float* W, float* out, hipStream_t st);
hipError_t launch_selftest_tanh(uint32_t lo, uint32_t hi, unsigned long long* out, hipStream_t st);
hipError_t launch_export(const int* meta, const Candidate* cands, const CheckResult* results, int* hmeta,
                         Candidate* hcands, CheckResult* hresults, int cap, hipStream_t st);

// Last HIP error code seen by the library on this thread (0 if none); optionally reset.
int last_error(bool reset);

// A model of the cascade in the reference's binary format, repacked in NetShape layout.
struct NetWeights {
    std::vector<float> w;
    int min_w = 0, min_h = 0;
};
// Parses format v1.0-1.1 (cnn_simd_cntk.cpp:42-404) and checks the shape against the
// stage's prototype. Returns 0 or -1 (message in err).
int load_model(const unsigned char* bytes, size_t n, int stage, int index_output, NetWeights& out, std::string& err);
// Embedded copies of the shipped models (stage 0..2: cnn4face{1,2,3}_cntk.bin,
// 3: cnn4landmarks_cntk.bin).
void embedded_model(int stage, const unsigned char** bytes, size_t* n);

// Effective detector settings (FaceDetector::Init's conversion, CompactCNNLibAPI_v2.cpp:100-138,
// and CNNDetector::Init, cnn_detector_v3.cpp:922-964).
struct Settings {
    int type_check = 1, min_num_detect = 1, adapt = 0, drop_detect = 1, merger = 1, double_check = 1;
    int equalize = 1, reflection = 1, detect_mode = 2, min_neighbors = 2;
    float t1 = -1.5f, t2 = -1.25f, t3 = -1.0f;
    int facial = 0;
};
Settings effective_settings(const ccnn_param& p);

// InitScales (cnn_detector_v3.cpp:289-346). Returns the scales (largest first); pre_scale
// receives the gray_image_only pre-resize factor (0 if none). Empty on failure.
std::vector<float> init_scales(const ccnn_param& p, int pattern_h, bool gray_only, float* pre_scale,
                               float* scale_factor_used);

struct Rect {
    int x = 0, y = 0, w = 0, h = 0, x2 = 0, y2 = 0;
    Rect() = default;
    Rect(int x_, int y_, int w_, int h_) : x(x_), y(y_), w(w_), h(h_), x2(x_ + w_), y2(y_ + h_) {}
};
// Detection (cnn_detector_v3.h:185-216); fd: the facial_data entries {gender, smile, glasses}
struct Det {
    Rect r;
    float score = 0.f, scale = 0.f;
    int knn = 0, num_detect = 0;
    bool checked = false;
    std::vector<std::array<int, 3>> fd;
};
// Merger (cnn_detector_v3.cpp:2814-2936)
void merger(std::vector<Det>& out, std::vector<Det>& in, float threshold, bool del);
// Post-processing of Detect (cnn_detector_v3.cpp:2745-2781): reverse, sort by scale,
// Merger 0.5, min_num_detect filter, Merger 0.2 with deletion.
std::vector<Det> post_process(const std::vector<Det>& raw, const Settings& s);

// Pyramid plan of one gray-image size: levels in compute order (scl_max -> 0).
struct Plan {
    int gw = -1, gh = -1;
    std::vector<LevelDesc> levels;
    int ntiles = 0, nchunks = 0, nl1tiles = 0;
    size_t resp_total = 0;
    size_t s1_total = 0;         // floats of the L1-sum planes of all levels
    std::vector<int2> l1_order;  // L1 tiles {level, tile} in XCD-band order
};

// Device and pinned-host state of one frame in flight. Three slots on three streams let the
// GPU work on frame i+1 while the host replays the decisions of frame i.
struct Slot {
    hipStream_t st = nullptr;
    // profiling events: on the stream at the frame's start, after gray, after stage 1
    // and at the end (ev); start/stop of k_stage1_l1, k_stage1 and k_check,
    // stamped by the launches themselves (kev)
    static constexpr int NEV = 4, NKEV = 6;
    hipEvent_t ev[NEV] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t kev[NKEV] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    uint8_t* d_img = nullptr;    size_t img_bytes = 0;
    uint8_t* d_pre = nullptr;    size_t pre_bytes = 0;
    float* d_gray = nullptr;     size_t gray_elems = 0;  int gpitch = 0;
    float* d_resp = nullptr;     size_t resp_cap = 0;
    float* d_s1 = nullptr;       size_t s1_cap = 0;
    LevelDesc* d_levels = nullptr;
    int2* d_l1_order = nullptr;  size_t l1_order_cap = 0;
    int* d_meta = nullptr;
    Candidate* d_cands = nullptr;
    CheckResult* d_results = nullptr;
    int2* d_s2 = nullptr;        int2* d_s3 = nullptr;
    size_t cand_cap = 0;
    // mapped pinned host memory that the kernels write (h_*: host view, hd_*: device view)
    int* h_meta = nullptr;            int* hd_meta = nullptr;
    Candidate* h_cands = nullptr;     Candidate* hd_cands = nullptr;
    CheckResult* h_results = nullptr; CheckResult* hd_results = nullptr;
    int2* h_s23 = nullptr;            int2* hd_s23 = nullptr;   // stage-2 then stage-3 outcomes, 2 x h_cap each
    size_t h_cap = 0;
    const Plan* levels_plan = nullptr;   // plan whose level table is in d_levels
    // frame in flight
    const Plan* plan = nullptr;
    int gw = 0, gh = 0;
    bool rescale = false, checks = false, host_decide = false, busy = false;
    size_t pre = 0;
};

class Detector {
public:
    static constexpr int NSLOTS = 3;
    Detector();
    ~Detector();

    int init(const ccnn_param& p);
    int detect(const ccnn_image& img, bool device_ptr, std::vector<ccnn_face>& faces);
    // Detection of n frames with the per-frame launches of consecutive frames pipelined
    // over the slots' streams. faces[i] receives frame i's detections; returns 0, or -1 on a
    // HIP error or an uninitialised detector.
    int detect_batch(int n, const ccnn_image* imgs, bool device_ptr, std::vector<std::vector<ccnn_face>>& faces);
    int clear();
    bool empty() const { return empty_; }

    ccnn_param param;          // the API-level parameters (getParam)
    int gray_image(ccnn_image* img, int pipeline);
    void set_profiling(bool on)
    {
        profiling_ = on;
        for (auto& v : acc_ms_) v = 0.f;
        acc_frames_ = 0;
    }
    int last_timing(float* ms) const;
    int last_trace(int* n_levels, int* levels, int cap, float* cand, float* res) const;

    // stage-level access (tests)
    int stage_check(const float* gray, int gw, int gh, int n, const float* scale, const int* x, const int* y,
                    float* res);
    // facial analysis results of the last frame's candidates: {ok, gender, smile, glasses}
    int last_facial(int* out, int cap) const;

private:
    int init_device();
    int build();                                  // Init chain after parameter changes
    int prepare(const ccnn_image& img);           // size guards and re-initialisation of Detect
    const Plan* plan_for(int gw, int gh);
    int ensure_slot(Slot& s, int w, int h, size_t img_bytes, const Plan& plan);
    int enqueue(Slot& s, const ccnn_image& img, bool device_ptr);
    int finish(Slot& s, std::vector<ccnn_face>& faces);
    void release_slot(Slot& s);
    CheckParams check_params() const;

    bool empty_ = true;
    bool built_ = false;
    bool gray_only_ = true;        // AdvancedParam::gray_image_only as forced by the API
    int device_ = 0;
    Settings set_;
    NetWeights net_[4];
    float* d_w_[4] = {nullptr, nullptr, nullptr, nullptr};   // stages 1-3, landmarks
    std::vector<float> scales_;
    float pre_scale_ = 0.f;
    int ext_w_ = 47, ext_h_ = 51;
    std::vector<Plan*> plans_;
    Slot slots_[NSLOTS];
    int last_slot_ = -1;
    std::vector<float> h_gray_;

    // last frame record
    int last_count_ = 0;
    std::vector<Candidate> last_cands_;
    std::vector<CheckResult> last_results_;
    std::vector<LevelDesc> last_levels_;

    bool profiling_ = false;
    static constexpr int NT = 7;                 // 6 kernel spans + the whole frame
    float last_ms_[NT] = {};
    double acc_ms_[NT] = {};
    int acc_frames_ = 0;
    int check_blocks_ = 1024;      // (candidate, crop variant) workgroups of k_check, grid-stride
    int fa_blocks_ = 64;           // k_facial workgroups (76 KB LDS each); windows dealt round-robin
};

}  // namespace ccnn
