/*
 * ccnn.h -- C ABI of the MI355X compact-CNN-cascade face detector (libccnn.so).
 *
 * Drop-in boundary: every entry point below replaces one member of the reference's
 * public class CompactCNNLib::FaceDetector (Bkmz21/CompactCNNCascade,
 * prebuild/include/CompactCNNLib.h:36-194, implemented in
 * src/CNNObjectDetector/CompactCNNLibAPI_v2.cpp). The reference exports an MSVC C++
 * class; this library exports plain C functions over an opaque handle and POD structs,
 * and include/CompactCNNLib.h rebuilds the reference's C++ class on top of them.
 *
 * All detection work runs on the GPU (HIP, gfx950). There is no CPU fallback: on a
 * machine without a usable HIP device ccnn_init returns -1.
 */
#ifndef CCNN_H
#define CCNN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FaceDetector::Pipeline / DetectMode / DetectPrecision (CompactCNNLib.h:38-59) */
enum { CCNN_PIPELINE_CPU = 0, CCNN_PIPELINE_GPU = 1, CCNN_PIPELINE_GPU_CPU = 2 };
enum { CCNN_DETECT_DISABLE = 0, CCNN_DETECT_SYNC = 1, CCNN_DETECT_ASYNC = 2 };
enum { CCNN_PRECISION_DEF = 0, CCNN_PRECISION_LOW = 1, CCNN_PRECISION_NORMAL = 2,
       CCNN_PRECISION_HIGH = 3, CCNN_PRECISION_ULTRA = 4 };

/* FaceDetector::Param (CompactCNNLib.h:98-165), field for field. */
typedef struct ccnn_param {
    int max_image_width, max_image_height;   /* max_image_size */
    int min_face_height;
    int max_face_height;
    float scale_factor;
    int min_neighbors;
    int detect_precision;
    float treshold[3];
    int drop_detect;
    int equalize;
    int reflection;
    int pipeline;
    int detect_mode;
    int num_threads;
    const char* models[3];                   /* accepted and ignored, as the reference ignores
                                                them (cnn_detector_v3.cpp:917-920): the shipped
                                                models are embedded in the library */
    int index_output[3];
    int device_info;
    int cuda_device_id;                      /* HIP device ordinal */
    int cl_platform_id;
    int cl_device_id;
    int facial_analysis;
} ccnn_param;

/* FaceDetector::Face (CompactCNNLib.h:73-86) */
typedef struct ccnn_face {
    int x, y, width, height;
    float score;
    int gender, smile, glasses;
} ccnn_face;

/* FaceDetector::ImageData (CompactCNNLib.h:87-96): 8-bit gray, BGR or BGRA; step in bytes. */
typedef struct ccnn_image {
    int cols, rows, channels;
    unsigned char* data;
    size_t step;
} ccnn_image;

typedef struct ccnn_detector ccnn_detector;

/* FaceDetector::Param::Param() defaults (CompactCNNLib.h:148-164). */
void ccnn_param_default(ccnn_param* p);

/* FaceDetector::FaceDetector / ~FaceDetector (CompactCNNLibAPI_v2.cpp:142-180). */
ccnn_detector* ccnn_create(void);
void ccnn_destroy(ccnn_detector* d);

/* FaceDetector::Init (CompactCNNLibAPI_v2.cpp:182-216): 0, or -1 on failure. */
int ccnn_init(ccnn_detector* d, const ccnn_param* p);

/* FaceDetector::Detect (CompactCNNLibAPI_v2.cpp:218-353). img->data is host memory.
 * Writes min(count, capacity) faces and returns the face count; -1 if not initialised
 * or faces == NULL. capacity < 0 means unbounded (the reference passes no capacity). */
int ccnn_detect(ccnn_detector* d, ccnn_face* faces, int capacity, const ccnn_image* img);

/* As ccnn_detect, with img->data a device pointer already resident in GPU memory. */
int ccnn_detect_device(ccnn_detector* d, ccnn_face* faces, int capacity, const ccnn_image* img);

/* Detection of n frames (host images, or device pointers when device != 0), each with
 * the semantics of one ccnn_detect call, the launches of consecutive frames pipelined
 * over three GPU streams so that the host-side decisions of frame i overlap the GPU work
 * of frames i+1 and i+2. Frame i's faces go to faces[i*capacity ...] (at most capacity) and its
 * face count to counts[i]. Returns 0, or -1. No counterpart in the reference, whose
 * Detect takes one frame per call. */
int ccnn_detect_batch(ccnn_detector* d, int n, const ccnn_image* imgs, int device, ccnn_face* faces,
                      int capacity, int* counts);

/* FaceDetector::Clear (:355-372): 0, or -1 if empty. */
int ccnn_clear(ccnn_detector* d);
/* FaceDetector::isEmpty (:374-377) */
int ccnn_is_empty(const ccnn_detector* d);
/* FaceDetector::getParam / setParam (:379-404) */
int ccnn_get_param(const ccnn_detector* d, ccnn_param* p);
void ccnn_set_param(ccnn_detector* d, const ccnn_param* p);
/* get/set MaxImageSize, Min/MaxObjectHeight, NumThreads (:406-460) */
int ccnn_get_max_image_size(const ccnn_detector* d, int* width, int* height);
void ccnn_set_max_image_size(ccnn_detector* d, int width, int height);
int ccnn_get_min_object_height(const ccnn_detector* d);
void ccnn_set_min_object_height(ccnn_detector* d, int h);
int ccnn_get_max_object_height(const ccnn_detector* d);
void ccnn_set_max_object_height(ccnn_detector* d, int h);
int ccnn_get_num_threads(const ccnn_detector* d);
void ccnn_set_num_threads(ccnn_detector* d, int n);
/* FaceDetector::getGrayImage32F (:462-507): the FP32 gray image of the last frame.
 * pipeline CCNN_PIPELINE_CPU: host pointer (library-owned copy); otherwise a device
 * pointer. step is in elements, as the reference reports widthStep. 0 or -1. */
int ccnn_get_gray_image(ccnn_detector* d, ccnn_image* img, int pipeline);
/* FaceDetector::CNTKDump2Binary (:509-519; CNNModelsConverter::LoadCNTKModel and
 * SaveToBinaryFile, cnn_models_converter.cpp:270-742): converts the text dump of a
 * trained CNTK cascade network into the binary model format v1.1, byte-identical to the
 * reference's output (".bin" is appended to binary_file when it has none). Host only;
 * needs no detector and no GPU. 0, or -1 when the dump cannot be read or is not a
 * three-layer cascade network (then no file is written). */
int ccnn_cntk_dump_to_binary(const char* binary_file, const char* cntk_model_dump);

/* ---------------------------------------------------------------------------------
 * Stage-level entry points. They run one kernel of the detection path on host
 * inputs and copy the result back, so that each stage can be checked on its own.
 * All return 0 on success, -1 on bad arguments or a HIP error.
 * --------------------------------------------------------------------------------- */
/* Img8uToImg32fGRAY (image_proc.cpp:299-317); out: rows x cols floats. */
int ccnn_stage_gray(const ccnn_image* img, float* out);
/* One pyramid level (ImageResizer::FastImageResize, image_resize.cpp:564-584) of a gray
 * image, mode 0 nearest, 1 bilinear, 2 copy; out: lh x lw. */
int ccnn_stage_level(const float* gray, int gw, int gh, int lw, int lh, int mode, float* out);
/* Stage-1 response map of a level image (ConvNeuralNetwork::Forward, cnn_simd_cntk.cpp:964);
 * out: rh x rw with (rw, rh) from ccnn_stage1_size. */
int ccnn_stage1_size(int lw, int lh, int* rw, int* rh);
int ccnn_stage1(const float* level, int lw, int lh, float* out);
/* Stage-2/3 check of candidate windows on a gray image (CPUCheckDetect without the drop
 * logic): per candidate i, scale[i] and response-map position (x[i], y[i]); res[8*i..]
 * = {flag, mod, k1, k2, m1, m2, n2, n3}: acceptance, crop variant of the final evaluation,
 * stage-2/3 counts and maxima of that evaluation, and the number of stage-2 and stage-3
 * network evaluations performed. Settings from p. */
int ccnn_stage_check(const ccnn_param* p, const float* gray, int gw, int gh, int n,
                     const float* scale, const int* x, const int* y, float* res);

/* The facial-analysis network (cnn4landmarks_cntk.bin, run by CPUCheckDetect on accepted
 * windows when Param::facial_analysis is set, cnn_detector_v3.cpp:1731-1756) on one 49x51
 * FP32 input (row-major, 49 columns); out: its 13 outputs (landmarks x0,y0..x4,y4 as
 * fractions of the crop, gender, smile, glasses). */
int ccnn_stage_landmarks(const float* in, float* out);

/* Record of the last ccnn_detect* call on d: the levels run (compute order) and the
 * candidate windows with their check results. Returns the candidate count; levels[6*l..]
 * = {scale index, lw, lh, rw, rh, mode}, n_levels receives the level count;
 * cand[4*i..] = {level, x, y, score}; res[8*i..] as ccnn_stage_check. Arrays may be
 * NULL to query the counts. */
int ccnn_last_trace(const ccnn_detector* d, int* n_levels, int* levels, int cap,
                    float* cand, float* res);

/* Facial analysis of the last ccnn_detect* call's candidates (same order as
 * ccnn_last_trace): out[4*i..] = {ok, gender, smile, glasses}; ok 0 marks a window the
 * checks accepted and the landmarks rejected. Returns the candidate count; out may be NULL. */
int ccnn_last_facial(const ccnn_detector* d, int* out, int cap);

/* Self-test of the exact path's reciprocal (the rational tanh of the hidden and output
 * layers, cnnpp_cplusplus.cpp:662-666, divides by a denominator >= 1; the kernels take
 * that reciprocal from the hardware estimate and one fused Newton step instead of the
 * general IEEE division). Over the float bit patterns [lo, hi) (NaNs skipped) on the
 * GPU: out[0] counts inputs where the tanh differs in any bit from the one with IEEE
 * division, out[1] inputs in [1, 2^25) where the reciprocal differs from 1/d (beyond
 * 2^25, 1 - 1/d rounds to 1 and the kernels use 0), out[2] the largest mismatching
 * pattern. No counterpart in the reference. 0, or -1. */
int ccnn_selftest_tanh(unsigned int lo, unsigned int hi, unsigned long long* out);

/* The last HIP error code the library met on the calling thread (0 if none), for callers
 * that need to tell "no faces" from a failed call (Detect reports 0 faces either way, as
 * the reference's does); reset != 0 clears it. */
int ccnn_last_error(int reset);

/* Per-kernel timing from HIP events recorded on the stream each frame runs on, enabled
 * (and the totals reset) by ccnn_set_profiling(d, 1). ms[0..6] = gray (with upload and
 * pre-resize), stage-1 L1 sums (k_stage1_l1), stage-1 tiles (k_stage1, which also
 * appends the candidate list), the gap from k_stage1 to k_check, stage-2/3 networks
 * (k_check), what runs after k_check (facial analysis when on; the candidate read-back
 * when the checks are off), the whole frame,
 * for the last frame; ms[7..13] the same summed over the frames since profiling was
 * enabled; ms[14] that frame count. ms must hold 15 floats. Returns 0. */
void ccnn_set_profiling(ccnn_detector* d, int on);
int ccnn_last_timing(const ccnn_detector* d, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* CCNN_H */
