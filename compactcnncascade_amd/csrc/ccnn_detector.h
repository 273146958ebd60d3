// Host side of the MI355X detection path: model loading, pyramid plan, device buffers,
// kernel dispatch and the ordered replay of the reference's order-dependent decisions.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
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
                         hipStream_t st, const hipEvent_t* kev = nullptr, int shape = 0, bool half = false,
                         const unsigned char* keep = nullptr, const unsigned char* l1_keep = nullptr);
// Response tiles, L1-sum plane geometry and L1 tiles of a level whose response map size
// is set, for tile shape `shape`; returns the floats the level's planes take.
size_t level_l1_layout(LevelDesc& L, int shape = 0);
// cap: candidates stored in cands (the list may count more, then the caller redoes the
// work); queue (optional): a zeroed counter from which a persistent grid takes its items
// phase timestamps of the profiling build (ccnn_kernels.hip)
int read_phase_stamps(int kern, unsigned long long* out, int nblocks);
hipError_t launch_check(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, const Candidate* cands,
                        const int* meta, const CheckParams& P, const float* W2, const float* W3, int2* s2res,
                        int2* s3res, CheckResult* results, int nblocks, int* hmeta, Candidate* hcands,
                        CheckResult* hresults, int2* hs2, int2* hs3, int hcap, hipStream_t st,
                        const hipEvent_t* kev = nullptr, int cap = 0x7fffffff, int* queue = nullptr, int rp = 0,
                        const float* MW2 = nullptr, const float* MW3 = nullptr);
// pre-resize of frame f of a batch (FrameDesc::orig, ow x oh -> src, dw x dh)
hipError_t launch_resize_u8_tab(const FrameDesc* frames, int f, int dw, int dh, hipStream_t st);
// the tile-skipping network on every frame of a batch and the keep flags of its stage-1
// response tiles (keep) and L1 tiles (l1_keep); kept (optional) counts the kept tiles
hipError_t launch_skipnet(const float* gray, const FrameDesc* ft, const SnFrame* fr, int nf, int cells,
                          const float* par, float* x0, float* a1, float* a2, float* z, const LevelDesc* levels, int nl,
                          int ntiles, int nl1tiles, int shape, const float* threshold, int srad, int margin,
                          unsigned char* keep, unsigned char* l1_keep, hipStream_t st);
// gray of every frame of a batch (FrameDesc table in device memory)
hipError_t launch_gray_batch(const FrameDesc* frames, int nframes, int total_quads, float* gray, hipStream_t st);
// ordered compaction of the windows above t1 over n responses (all levels of a batch back
// to back): counts/offs hold one int per CMP_BLOCK positions; the first cap candidates go
// to cands (and the first hcap to hcands); meta[0] / hmeta[0] = total, meta[1] = 0
hipError_t launch_compact(const float* resp, int n, float t1, int* counts, int* offs, const LevelDesc* d_levels,
                          int nlevels, Candidate* cands, int cap, Candidate* hcands, int hcap, int* meta, int* hmeta,
                          hipStream_t st);
// k_facial takes the check decisions from the stage-2/3 outcomes itself and writes the facial
// results of the accepted candidates to results[].fa (and hresults[].fa below hcap)
hipError_t launch_facial(const float* gray, int gp, int gw, int gh, const LevelDesc* d_levels, const Candidate* cands,
                         const int* meta, const CheckParams& P, const int2* s2res, const int2* s3res, const float* W,
                         CheckResult* results, CheckResult* hresults, int hcap, int nblocks, hipStream_t st);
hipError_t launch_landmarks(const float* in, const float* W, float* out, hipStream_t st);
// stage 1 (stage-2 network) or 2 (stage-3 network) on n 47x51 inputs; 16 outputs each.
// rp: 0 exact FP32, 1 BF16 / 2 FP8 matrix-core L1 and L2 (MW: the network's MfmaW table)
hipError_t launch_net23(int stage, const float* in, int n, const float* W, float* out, hipStream_t st, int rp = 0,
                        const float* MW = nullptr);
// The MfmaW table of a stage-2/3 network (NetShape layout w)
void mfma_weights(const std::vector<float>& w, std::vector<float>& out);
hipError_t launch_selftest_tanh(uint32_t lo, uint32_t hi, unsigned long long* out, hipStream_t st);
hipError_t launch_export(const int* meta, const Candidate* cands, const CheckResult* results, int* hmeta,
                         Candidate* hcands, CheckResult* hresults, int cap, hipStream_t st);

// Last HIP error code seen by the library on this thread (0 if none); optionally reset.
int last_error(bool reset);
// Records a HIP error as last_error reports it.
void note_error(hipError_t e);

// Device and pinned-host allocation helpers (each frees what the pointer held first).
template <class T>
void dfree(T*& p)
{
    if (p) { (void)hipFree((void*)p); p = nullptr; }
}
template <class T>
void hfree(T*& p)
{
    if (p) { (void)hipHostFree((void*)p); p = nullptr; }
}
template <class T>
int dalloc(T*& p, size_t n)
{
    dfree(p);
    return hipMalloc((void**)&p, sizeof(T) * (n > 0 ? n : 1)) == hipSuccess ? 0 : -1;
}
// pinned host memory mapped into the device address space; dp receives the device view
template <class T>
int halloc(T*& p, T*& dp, size_t n)
{
    hfree(p);
    dp = nullptr;
    if (hipHostMalloc((void**)&p, sizeof(T) * (n > 0 ? n : 1), hipHostMallocMapped | hipHostMallocCoherent) !=
        hipSuccess)
        return -1;
    return hipHostGetDevicePointer((void**)&dp, p, 0) == hipSuccess ? 0 : -1;
}

// A model of the cascade in the reference's binary format, repacked in NetShape layout.
struct NetWeights {
    std::vector<float> w;
    int min_w = 0, min_h = 0;
};
// Parses format v1.0-1.1 (cnn_simd_cntk.cpp:42-404) and checks the shape against the
// stage's prototype. Stage 0 (stage 1 of the cascade) also gets its FP16 tap pairs
// (Stage1H) after the FP32 weights. Returns 0 or -1 (message in err).
int load_model(const unsigned char* bytes, size_t n, int stage, int index_output, NetWeights& out, std::string& err);
// Embedded copies of the shipped models (stage 0..2: cnn4face{1,2,3}_cntk.bin,
// 3: cnn4landmarks_cntk.bin).
void embedded_model(int stage, const unsigned char** bytes, size_t* n);
// Appends the Stage1H FP16 tap pairs of the stage-1 weights w (NetShape layout).
void append_stage1_half(std::vector<float>& w);

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

// Decision of CPUCheckDetect (cnn_detector_v3.cpp:1709-1725, 1808-1811) from the stage-2/3
// outcomes of both crop variants of a candidate.
CheckResult check_decision(const CheckParams& P, float scale, const int2* s2, const int2* s3);

// Pyramid plan of one gray-image size: levels in compute order (scl_max -> 0) among the
// first ncap scales, and the scale count the detector keeps for its next frame.
struct Plan {
    long long id = 0;            // unique per detector (the batch graph's key names plans by it)
    int shape = 0;               // stage-1 tile shape the levels are laid out for
    int pins = 0;                // frames of a batch being assembled that use the plan
    int gw = -1, gh = -1;
    int ncap = 0;                // scales visited (CNNDetector::num_scales on entry)
    int next_num_scales = 0;     // num_scales after the skip rule (cnn_detector_v3.cpp:2083-2088)
    std::vector<LevelDesc> levels;
    int ntiles = 0, nchunks = 0, nl1tiles = 0;
    size_t resp_total = 0;
    size_t s1_total = 0;         // floats of the L1-sum planes of all levels
    // {level, tile within the level}: the nl1tiles L1 tiles in XCD-band order, then the
    // ntiles stage-1 tiles in launch order
    std::vector<int2> l1_order;
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

// One frame of a batch, planned (ccnn_batch.cpp).
struct BatchFrame {
    int idx;                      // index in the caller's frame list
    const Plan* plan;
    const uint8_t* data;          // the caller's frame (host or device memory)
    int w, h, ch;
    size_t step;
    bool pre;                     // gray_image_only pre-resize to gw x gh
    int gw, gh;
};

// Device buffers and the captured graph of the batch path: every level of every frame of a
// batch goes through one graph launch (gray, stage 1, ordered compaction, checks).
struct BatchArena {
    hipStream_t st = nullptr;
    uint8_t* d_frames = nullptr;  size_t frames_cap = 0;   // host frames, uploaded
    uint8_t* d_pre = nullptr;     size_t pre_cap = 0;      // pre-resized frames
    float* d_gray = nullptr;      size_t gray_cap = 0;
    float* d_s1 = nullptr;        size_t s1_cap = 0;
    float* d_resp = nullptr;      size_t resp_cap = 0;
    LevelDesc* d_levels = nullptr; size_t levels_cap = 0;
    int2* d_l1_order = nullptr;   size_t l1_cap = 0;
    FrameDesc* d_ftab = nullptr;  FrameDesc* h_ftab = nullptr; FrameDesc* hd_ftab = nullptr; size_t ftab_cap = 0;
    int* d_counts = nullptr;      int* d_offs = nullptr;   size_t cnt_cap = 0;
    int* d_meta = nullptr;
    Candidate* d_cands = nullptr; int2* d_s2 = nullptr; int2* d_s3 = nullptr; size_t cand_cap = 0;
    int* h_meta = nullptr;        int* hd_meta = nullptr;
    Candidate* h_cands = nullptr; Candidate* hd_cands = nullptr;
    int2* h_s23 = nullptr;        int2* hd_s23 = nullptr;  size_t h_cap = 0;
    // tile skipping: per-frame cell table, network activations, keep flags
    SnFrame* d_snf = nullptr;     size_t snf_cap = 0;
    float* d_sn = nullptr;        size_t sn_cap = 0;       // x0, a1, a2, z back to back (1 + 8 + 8 + SLOTS per cell)
    unsigned char* d_keep = nullptr; size_t keep_cap = 0;  // response tiles, then L1 tiles
    long long tiles = 0;          // stage-1 response tiles of the last batch
    bool skipped = false;         // whether the last batch ran the proposal network
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<long long> key;   // what the captured graph was built for
    // profiled batches: events between the kernel groups, and the last batch's ms per
    // group (gray, proposal network, L1 sums, stage-1 tiles, compaction, checks), the
    // whole sequence, and its frame count
    hipEvent_t ev[11] = {};
    float ms[8] = {};
    int launches = 0;             // kernel launches in the graph
    int gen = 0;                  // bumped when buffers move (the graph is then recaptured)
};

// Host threads that replay the frames of a batch: started once, then woken per batch.
// run(n, fn) calls fn(i) for every i in [0, n) on the pool's threads and on the caller,
// whose own work `mine` runs first, and returns when all calls have returned.
class ReplayPool {
public:
    explicit ReplayPool(int threads);
    ~ReplayPool();
    void run(int n, const std::function<void(int)>& fn, const std::function<void()>& mine);

private:
    void loop();
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable wake_, done_;
    const std::function<void(int)>* fn_ = nullptr;
    std::atomic<int> next_{0};
    int n_ = 0, busy_ = 0;
    long long gen_ = 0;
    bool stop_ = false;
};

class Detector {
public:
    static constexpr int NSLOTS = 3;
    Detector();
    ~Detector();

    int init(const ccnn_param& p);
    // One parameter changed through a single-field setter: the detector is rebuilt at the
    // next Detect with gray_image_only as it stands (CNNDetector's setters only Clear and
    // change param, cnn_detector_v3.h:404-436; Init and setParam force it on again).
    void set_fields(const ccnn_param& p)
    {
        param = p;
        built_ = false;
    }
    int detect(const ccnn_image& img, bool device_ptr, std::vector<ccnn_face>& faces);
    // Detection of n frames with the per-frame launches of consecutive frames pipelined
    // over the slots' streams. faces[i] receives frame i's detections; returns 0, or -1 on a
    // HIP error or an uninitialised detector.
    int detect_batch(int n, const ccnn_image* imgs, bool device_ptr, std::vector<std::vector<ccnn_face>>& faces);
    // Batch mode of detect_batch: 1 (default) runs each batch of frames as one graph launch
    // (ccnn_batch.cpp); 0 pipelines per-frame launches over the slots' streams. Facial
    // analysis, and a batch whose candidates overflow the batch buffers, take the per-frame
    // path either way.
    void set_batch_mode(int m) { batch_mode_ = m; }
    int batch_mode() const { return batch_mode_; }
    // Compute precision: 0 the exact FP32 path (the drop-in, bit-exact with the reference);
    // 1 FP16 stage 1 and BF16 matrix-core stages 2/3; 2 FP16 stage 1 and FP8 stages 2/3;
    // 3 / 4 the exact stage 1 with BF16 / FP8 stages 2/3.
    int set_precision(int p)
    {
        if (p < 0 || p > 4) return -1;
        if (p != precision_) batch_.key.clear();
        precision_ = p;
        return 0;
    }
    int precision() const { return precision_; }
    bool half_stage1() const { return precision_ == 1 || precision_ == 2; }
    // operand type of the stage-2/3 kernels (RP_EXACT, RP_BF16, RP_FP8)
    int check_precision() const { return precision_ == 0 ? 0 : (precision_ == 1 || precision_ == 3) ? 1 : 2; }
    // Tile skipping (see ccnn_set_skipnet): params NULL switches it off.
    int set_skipnet(const float* params, float threshold, int min_face, float scale_factor);
    // Per-slot keep thresholds (CCNN_SKIP_SLOTS probabilities in (0, 1)). 0, or -1.
    int set_skip_slot_thresholds(const float* thr)
    {
        for (int k = 0; k < CCNN_SKIP_SLOTS; ++k)
            if (!(thr[k] > 0.f && thr[k] < 1.f)) return -1;
        for (int k = 0; k < CCNN_SKIP_SLOTS; ++k) skip_threshold_[k] = thr[k];
        batch_.key.clear();
        return 0;
    }
    // Keep window of the tile mask: scale slots within slot_radius of a level's own, cells
    // within cell_margin of its windows' centres (defaults 0 and 1). 0, or -1.
    int set_skip_window(int slot_radius, int cell_margin)
    {
        if (slot_radius < 0 || slot_radius >= CCNN_SKIP_SLOTS || cell_margin < 0 || cell_margin > 16) return -1;
        skip_srad_ = slot_radius;
        skip_margin_ = cell_margin;
        batch_.key.clear();
        return 0;
    }
    // Response tiles of the last batch and how many of them ran (counted from the keep
    // flags the batch left in device memory).
    void skip_stats(long long* tiles, long long* kept);
    // Grid of k_check: on the per-frame path frame_per_cu persistent workgroups per CU
    // taking (candidate, crop variant) items from a queue (0: one workgroup per item,
    // grid-stride over 1024); on the batch path batch_per_cu (1..8). 0, or -1.
    int set_check_grid(int frame_per_cu, int batch_per_cu)
    {
        if (frame_per_cu < 0 || frame_per_cu > 8 || batch_per_cu < 1 || batch_per_cu > 8) return -1;
        check_frame_per_cu_ = frame_per_cu;
        check_per_cu_ = batch_per_cu;
        return 0;
    }
    // Stage-1 tile shape of the plans made from now on (see TileShape); 0 or -1 for a bad id.
    int set_tile_shape(int id)
    {
        if (id < 0 || id >= N_TILE_SHAPES) return -1;
        shape_ = frame_shape_ = id;
        return 0;
    }
    int tile_shape_id() const { return shape_; }
    // Autotuner: runs the batch `reps` times with each tile shape (and each persistent grid
    // size of the check kernel in {2, 3, 4} workgroups per CU), keeps the fastest and
    // returns its shape id; ms[s] receives each shape's best batch time (N_TILE_SHAPES
    // entries, NaN for shapes that failed) and grid_ms[g] each grid's (3 entries).
    int autotune(int n, const ccnn_image* imgs, bool device_ptr, int reps, float* ms, float* grid_ms);
    // ms per kernel group of the last profiled batch (see BatchArena::ms): 8 floats
    void batch_timing(float* ms) const
    {
        for (int k = 0; k < 8; ++k) ms[k] = batch_.ms[k];
    }
    // kernel launches of the last batch graph, graph captures so far
    int batch_launches() const { return batch_.launches; }
    int batch_captures() const { return batch_captures_; }
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
    // stage-1 response maps of the last frame's levels (compute order, each rh x rw,
    // back to back); returns the float count, -1 if none or cap is too small
    long last_responses(float* out, size_t cap);

    // stage-level access (tests)
    int stage_check(const float* gray, int gw, int gh, int n, const float* scale, const int* x, const int* y,
                    float* res);
    // facial analysis results of the last frame's candidates: {ok, gender, smile, glasses}
    int last_facial(int* out, int cap) const;
    // replay status of the last frame's candidates (see replay)
    int last_status(int* out, int cap) const
    {
        const int n = std::min(cap, (int)last_status_.size());
        for (int i = 0; i < n && out; ++i) out[i] = last_status_[i];
        return (int)last_status_.size();
    }

private:
    int init_device();
    int build();                                  // Init chain after parameter changes
    int prepare(const ccnn_image& img);           // size guards and re-initialisation of Detect
    const Plan* plan_for(int gw, int gh, int ncap, int shape);
    int ensure_slot(Slot& s, int w, int h, size_t img_bytes, const Plan& plan);
    int enqueue(Slot& s, const ccnn_image& img, bool device_ptr);
    // launches one planned frame on slot s
    int enqueue_plan(Slot& s, const ccnn_image& img, bool device_ptr, const Plan* plan, bool pre, int gw, int gh);
    // the frame's pre-resize, gray size and plan; advances num_scales_
    const Plan* plan_frame(const ccnn_image& img, bool* pre, int* gw, int* gh, int shape);
    int detect_batch_pipelined(int n, const ccnn_image* imgs, bool device_ptr,
                               std::vector<std::vector<ccnn_face>>& faces);
    int detect_batch_graph(int n, const ccnn_image* imgs, bool device_ptr, std::vector<std::vector<ccnn_face>>& faces);
    // one batch of planned frames through the graph; 1 when the candidates overflowed the
    // batch buffers (nothing written to faces), 0 on success, -1 on error
    int run_batch(const std::vector<BatchFrame>& seg, bool device_ptr, std::vector<std::vector<ccnn_face>>& faces);
    void release_batch();
    int finish(Slot& s, std::vector<ccnn_face>& faces);
    void release_slot(Slot& s);
    CheckParams check_params() const;
    // status (optional): per candidate 0 dropped at the scan, 1 dropped before its check,
    // 2 checked, 3 a raw window (no checks)
    std::vector<Det> replay(const LevelDesc* levels, int nl, const Candidate* cands, const CheckResult* res, int count,
                            bool checks, std::vector<int>* status = nullptr) const;
    void emit_faces(const std::vector<Det>& rects, bool rescale, std::vector<ccnn_face>& faces) const;

    bool empty_ = true;
    bool built_ = false;
    bool gray_only_ = true;        // AdvancedParam::gray_image_only as forced by the API
    int device_ = 0;
    Settings set_;
    NetWeights net_[4];
    float* d_w_[4] = {nullptr, nullptr, nullptr, nullptr};   // stages 1-3, landmarks
    float* d_mw_[2] = {nullptr, nullptr};                     // MfmaW tables of stages 2, 3
    int precision_ = 0;
    float* d_skip_ = nullptr;      // skip-net weights (FP16-rounded), nullptr when off
    float skip_threshold_[CCNN_SKIP_SLOTS] = {};   // keep probability threshold per scale slot
    int skip_min_face_ = 0;
    float skip_sf_ = 0.f;
    int skip_srad_ = 0;
    int skip_margin_ = 1;
    std::vector<float> scales_;
    // CNNDetector::num_scales: the scales RunCPUDetect visits. Init sets it to the scale
    // count, each frame's skip rule lowers it to the first scale whose level is too small,
    // and Detect restores it only when the frame size differs from max_image_size
    // (cnn_detector_v3.cpp:603, 2083-2088, 2514-2518): a frame of the maximum size that
    // follows a smaller one skips the scales the smaller one could not run.
    int num_scales_ = 0;
    float pre_scale_ = 0.f;
    int ext_w_ = 47, ext_h_ = 51;
    std::vector<Plan*> plans_;
    Slot slots_[NSLOTS];
    int last_slot_ = -1;
    std::vector<float> h_gray_;

    // last frame record: its gray image and response maps (device memory of the slot or
    // batch that ran it), candidates and check results
    const float* last_gray_ = nullptr;
    int last_gp_ = 0, last_gw_ = 0, last_gh_ = 0;
    const float* last_resp_ = nullptr;
    int last_count_ = 0;
    std::vector<Candidate> last_cands_;
    std::vector<int> last_status_;
    std::vector<CheckResult> last_results_;
    std::vector<LevelDesc> last_levels_;

    bool profiling_ = false;
    static constexpr int NT = 7;                 // 6 kernel spans + the whole frame
    float last_ms_[NT] = {};
    double acc_ms_[NT] = {};
    int acc_frames_ = 0;
    int check_blocks_ = 1024;      // (candidate, crop variant) workgroups of k_check, grid-stride
    int cus_ = 256;                // compute units of the device
    int check_per_cu_ = 1024 / CHECK_THREADS;   // persistent k_check workgroups per CU on the batch path
    int check_frame_per_cu_ = 0;   // per-frame path: 0 grid-stride over check_blocks_, else persistent
    int batch_mode_ = 1;
    long long plan_ids_ = 0;
    // stage-1 tile shape of new plans of the batch path (set_tile_shape, autotune): 6 = L1
    // tiles of 32 x 32, response tiles of 16 x 16, the autotuner's pick on gfx950 for the
    // 1080p, 4K and mixed-size batches (bench.py --autotune); and of the per-frame path
    // (set_tile_shape): 0 = L1 tiles of 64 x 16, 2% faster than 6 for one 1080p frame
    // alone on the GPU
    int shape_ = 6;
    int frame_shape_ = 0;
    int batch_captures_ = 0;
    BatchArena batch_;
    std::unique_ptr<ReplayPool> pool_;   // host replay threads of the batch path
    int fa_blocks_ = 64;           // k_facial workgroups (76 KB LDS each); windows dealt round-robin
};

}  // namespace ccnn
