/*
 * ccnn_oracle.h -- CPU restatement of the reference's compact-CNN-cascade detection
 * path (Bkmz21/CompactCNNCascade, plain-C++ FP32 CPU build). TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker. The product never links or calls it.
 *
 * Every floating-point expression is evaluated in the reference's source order with
 * no contraction (built -O2 -ffp-contract=off -fno-fast-math), which reproduces the
 * reference's CMake-default build (-ffast-math without -O) bit for bit.
 *
 * Pinning (what checks each function, besides the compiled reference in oracle/_ref):
 *   oc_gray          closed form; exhaustive over all 2^24 (B,G,R) triples in tests.
 *   oc_resize_*      integer LUT closed form, brute-force small images, copy/clamp cases.
 *   oc_cnn_forward   float64 textbook restatement of each layer (numpy) within 1e-5;
 *                    4-pixel shift invariance; output-size rule.
 *   oc_equalize      closed form; the reference's own spread property test.
 *   oc_sort_*        libstdc++ std::sort (GCC 11) permutation, against std::sort.
 *   oc_detect        end-to-end detection sets against the compiled reference.
 */
#ifndef CCNN_ORACLE_H
#define CCNN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Model in the reference's binary format v1.0-1.1 (cnn_simd_cntk.cpp:42-404). */
typedef struct {
    int min_w, min_h;
    int maps[3];
    int kw[3], kh[3];
    float* kern[3];                 /* [map][kh*kw], row-major */
    float af_scale;                 /* sign already applied for index_output == 0 */
    float* conv_b[3];
    float* lrelu_w1[3];
    float* lrelu_w2[3];
    float* bn_w[3];
    float* bn_b[3];
    int full_connect, hl, connect, hl_scale;
    float* hl_w;                    /* [hl][nin] */
    int hl_nin;
    float* hl_b;                    /* [hl] */
    float* hl_tanh_w;               /* [hl / hl_scale] */
    float* hl_bn_w;                 /* [hl_scale] */
    float* hl_bn_b;                 /* [hl_scale] */
    int ol;
    float* ol_w;                    /* [ol][hl] */
    float* ol_b;                    /* [ol] */
    float ol_tanh_w;
    int index_output;
} oc_model;

int  oc_model_load(oc_model* m, const unsigned char* bytes, long n, int index_output);
void oc_model_free(oc_model* m);
void oc_output_size(const oc_model* m, int w, int h, int* ow, int* oh);
/* Dense output [oh][ow] ([ol][oh][ow] when index_output < 0: every output map); returns 0,
 * or -1 if the image is below the model's minimum. */
int  oc_cnn_forward(const oc_model* m, const float* img, int w, int h, int step, float* out);

/* Image ops */
int  oc_gray(const unsigned char* img, int w, int h, int ch, int step, float* out, int out_step);
void oc_resize_f32(const float* src, int sw, int sh, int sstep,
                   float* dst, int dw, int dh, int dstep, int type);
void oc_resize_u8(const unsigned char* src, int sw, int sh, int sstep,
                  unsigned char* dst, int dw, int dh, int dstep, int type);
void oc_equalize(unsigned char* img, int w, int h, int step);
/* The check crop pipeline of CPUCheckDetect (cnn_detector_v3.cpp:1565-1575, 1617, 1638):
 * nearest-neighbour resize to 49x51 (f32_out, may be NULL), +0.5-truncate to u8,
 * optional equalize, stage-2 input (cols 2..48) and stage-3 input (mirrored or plain). */
void oc_check_prep(const float* crop, int cw, int ch, int cstep, int equalize, int reflection,
                   unsigned char* u8_out, float* s2_out, float* s3_out, float* f32_out);

/* libstdc++ std::sort restated over an index permutation keyed by float, with the
 * detector's comparator comp(a, b) = key[b] < key[a]. perm is sorted in place. */
void oc_sort_desc(int* perm, const float* key, int n);

/* Detector parameters: the FaceDetector::Param fields the path reads
 * (prebuild/include/CompactCNNLib.h:98-165). */
typedef struct {
    int max_w, max_h;
    int min_face, max_face;
    int min_neighbors;
    int precision;      /* 0 def, 1 low, 2 normal, 3 high, 4 ultra */
    int drop_detect, equalize, reflection;
    int detect_mode;    /* 0 disable, 1 sync, 2 async */
    float scale_factor;
    float t1, t2, t3;
    int facial_analysis;
} oc_params;

/* FaceDetector::Face (CompactCNNLib.h:73-86); gender/smile/glasses -1 without facial data */
typedef struct { int x, y, w, h; float score; int gender, smile, glasses; } oc_face;

/* Per-level and per-candidate record of one detection, for stage-by-stage parity. */
typedef struct {
    int n_scales;                /* scales built by InitScales */
    float scales[512];
    int n_levels;                /* levels run, in compute order (smallest first) */
    int lvl_scl[512], lvl_w[512], lvl_h[512], lvl_rw[512], lvl_rh[512], lvl_mode[512];
    /* candidates: every response > t1, level compute order, raster order */
    int cap, n_cand;
    int* c_scl; int* c_x; int* c_y; float* c_score;
    int* c_status;               /* 0 scan-drop, 1 check-drop, 2 accepted, 3 rejected,
                                    4 accepted by the checks, rejected by facial analysis */
    int* c_mod;                  /* mod of the final evaluation (0 or 1), -1 if none */
    int* c_k1; float* c_m1; int* c_k2; float* c_m2;   /* of the final evaluation */
    int n_raw;                   /* detections before Merger */
} oc_trace;

/* Full detection of one frame with the reference's semantics, on a freshly initialised
 * detector. Returns the face count (<= cap written), 0 for images below 47x51, -1 on error. */
int oc_detect(const oc_model* m1, const oc_model* m2, const oc_model* m3, const oc_model* mfa,
              const oc_params* p, const unsigned char* img, int w, int h, int ch, int step,
              oc_face* out, int cap, oc_trace* trace);

/* Detector state that CNNDetector carries from one Detect call to the next:
 *   max_w, max_h  param.max_image_size, grown by a larger frame (cnn_detector_v3.cpp:2499-2513);
 *   num_scales    the scale count RunCPUDetect visits: Init sets it to the scale count
 *                 (:603), the skip rule lowers it to the first scale whose level is too
 *                 small (:2083-2088), and Detect resets it only when the frame's size
 *                 differs from max_image_size (:2514-2518). -1 stands for "all scales";
 *   gray_only     AdvancedParam::gray_image_only: forced on by the API's Init and setParam
 *                 (CompactCNNLibAPI_v2.cpp:131), turned off for good by a colour frame
 *                 that is larger than max_image_size (:2504-2507) or that meets the
 *                 pre-resize (:2540-2546). */
typedef struct { int max_w, max_h; int num_scales; int gray_only; } oc_state;

/* The state after FaceDetector::Init with p. */
void oc_state_init(oc_state* s, const oc_params* p);

/* As oc_detect, on a detector in state s (updated). */
int oc_detect_state(const oc_model* m1, const oc_model* m2, const oc_model* m3, const oc_model* mfa,
                    const oc_params* p, oc_state* s, const unsigned char* img, int w, int h, int ch, int step,
                    oc_face* out, int cap, oc_trace* trace);

/* One candidate's check (CPUCheckDetect without the drop logic): results of mod 0 and,
 * when mod 0 fails and double_check is on, mod 1. res[0..9] =
 * {accepted, mod, k1, k2, m1, m2, x, y, w, h} (floats). */
int oc_check_candidate(const oc_model* m2, const oc_model* m3, const oc_params* p,
                       const float* gray, int gw, int gh, int gstep,
                       float scale, int px, int py, float* res);

/* Scales and level plan (InitScales + skip rule) for an image of w x h. */
int oc_plan(const oc_params* p, int pattern_h, int w, int h, float* scales, int* n_scales,
            int* lvl_w, int* lvl_h, int* n_run);

#ifdef __cplusplus
}
#endif
#endif
