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

/* How ccnn_detect_batch runs a batch (no counterpart in the reference). mode 1 (the
 * default): every pyramid level of every frame goes through ONE captured HIP graph launch
 * (gray of all frames, stage 1 over the (frame, level, tile) work list, ordered compaction
 * of the candidate windows, stages 2/3 on a persistent grid), the graph captured once per
 * batch layout (frame sizes, channels) and replayed; mode 0: per-frame launches pipelined
 * over three streams. Both give each frame the detections of one ccnn_detect call in
 * turn. Facial analysis, and batches with more candidate windows than the batch buffers
 * hold, run per frame in either mode. 0, or -1 for a bad mode. */
int ccnn_set_batch_mode(ccnn_detector* d, int mode);
/* Compute precision of the detector (no counterpart in the reference):
 *   CCNN_PRECISION_FP32       the exact FP32 path (default): results bit-identical to the
 *                             reference's CPU path;
 *   CCNN_PRECISION_FP16_BF16  stage 1 with FP16 convolution inputs and weights (products
 *                             summed in FP32, v_dot2_f32_f16; L1 windows rounded less
 *                             their centre sample, which is added back in FP32); the
 *                             stage-2/3 L1 and L2 layers as BF16 GEMMs on the matrix
 *                             cores with FP32 accumulation;
 *   CCNN_PRECISION_FP16_FP8   the same with FP8 (e4m3) operands for stages 2/3, scaled
 *                             per layer by powers of two;
 *   CCNN_PRECISION_FP32_BF16  the exact FP32 stage 1 with the BF16 stages 2/3;
 *   CCNN_PRECISION_FP32_FP8   the exact FP32 stage 1 with the FP8 stages 2/3.
 * Reduced precisions change scores and may change the detection set; recall against the
 * FP32 path is stated by the project's tests and benchmark. 0, or -1 for a bad value. */
enum { CCNN_PRECISION_FP32 = 0, CCNN_PRECISION_FP16_BF16 = 1, CCNN_PRECISION_FP16_FP8 = 2,
       CCNN_PRECISION_FP32_BF16 = 3, CCNN_PRECISION_FP32_FP8 = 4 };
int ccnn_set_compute_precision(ccnn_detector* d, int precision);
int ccnn_get_compute_precision(const ccnn_detector* d);

/* Stage-1 tile shapes (no counterpart in the reference). The two stage-1 kernels are
 * built for CCNN_TILE_SHAPES combinations of an L1 tile (pooled L1 positions per
 * workgroup: 64x16, 32x16, 32x32, 64x8) and a response tile (responses per workgroup:
 * 16x16, 32x8, 16x8); shape id = L1 index * 3 + response index. Defaults: 6 (L1 32x32,
 * responses 16x16, the autotuner's pick on gfx950) on the batch path, 0 (L1 64x16,
 * responses 16x16) on the per-frame path. The results do not depend on the shape. set:
 * both paths, for the frames planned from now on; 0, or -1 for a bad id. get: the batch
 * path's id (the one ccnn_autotune sets), its sizes in ax/ay/tx/ty (may be NULL). */
#define CCNN_TILE_SHAPES 12
int ccnn_set_tile_shape(ccnn_detector* d, int shape);
/* Grid of the stage-2/3 check kernel (no counterpart in the reference): on the per-frame
 * path frame_per_cu persistent workgroups per compute unit that take (candidate, crop
 * variant) items from a queue, or 0 for one workgroup per item (grid-stride over 1024);
 * on the batch path batch_per_cu (1..8) persistent workgroups per compute unit.
 * Defaults 0 and 4. Results do not depend on it. 0, or -1. */
int ccnn_set_check_grid(ccnn_detector* d, int frame_per_cu, int batch_per_cu);
int ccnn_get_tile_shape(const ccnn_detector* d, int* ax, int* ay, int* tx, int* ty);
/* Autotuner of the batch path: runs ccnn_detect_batch over the n frames (host images, or
 * device pointers when device != 0) reps + 1 times per stage-1 tile shape, keeps the
 * fastest shape, then does the same over the persistent grid of the stage-2/3 kernel (2,
 * 3 or 4 workgroups per CU) and keeps the fastest. shape_ms[CCNN_TILE_SHAPES] and
 * grid_ms[3] (may be NULL) receive each one's best batch time in ms (NaN if it failed).
 * Returns the chosen shape id, or -1. */
int ccnn_autotune(ccnn_detector* d, int n, const ccnn_image* imgs, int device, int reps, float* shape_ms,
                  float* grid_ms);
/* Learned tile skipping (no counterpart in the reference). A proposal network (SkipNet:
 * three 3x3 convolutions, 1 -> 8 -> 8 -> CCNN_SKIP_SLOTS channels, ReLU, on the gray
 * frame downsampled by 8) gives per 8x8 frame cell and per pyramid scale index s <
 * CCNN_SKIP_SLOTS the logit that level s has a window the cascade accepts there. With a
 * network set, the batch path (ccnn_set_batch_mode 1) runs it on every frame and computes
 * only the stage-1 tiles whose largest probability over their windows reaches
 * `threshold`; the other tiles' windows are dropped. Levels with scale index >=
 * CCNN_SKIP_SLOTS, and detectors whose min_face_height / scale_factor differ from the
 * ones the network was trained for, run plain. params: CCNN_SKIP_PARAMS floats (w1
 * [8][1][3][3], b1 [8], w2 [8][8][3][3], b2 [8], w3 [SLOTS][8][3][3], b3 [SLOTS]), used
 * rounded to FP16; NULL switches skipping off. 0, or -1. */
#define CCNN_SKIP_SLOTS 24
#define CCNN_SKIP_PARAMS 2416
int ccnn_set_skipnet(ccnn_detector* d, const float* params, float threshold, int min_face, float scale_factor);
/* Keep window of the tile mask (no counterpart in the reference): a response tile runs
 * when some logit of the scale slots within slot_radius of its level's slot, over the
 * cells within cell_margin of its windows' centres, reaches the threshold. Defaults 0
 * and 1; slot_radius < CCNN_SKIP_SLOTS, cell_margin <= 16. 0, or -1. */
int ccnn_set_skip_window(ccnn_detector* d, int slot_radius, int cell_margin);
/* Keep threshold per scale slot (no counterpart in the reference): CCNN_SKIP_SLOTS
 * probabilities in (0, 1) replacing ccnn_set_skipnet's single threshold (which sets all
 * slots). 0, or -1. */
int ccnn_set_skip_slot_thresholds(ccnn_detector* d, const float* thresholds);
/* Stage-1 response tiles in the last batch and how many of them ran (read from the keep
 * flags in device memory: a synchronous copy). 0, or -1. */
int ccnn_skip_stats(ccnn_detector* d, long long* tiles, long long* kept);
/* Training and checking of the proposal network on the GPU with the library's own
 * kernels (forward, backward, weighted binary cross-entropy, Adam). x0: n inputs of h8 x
 * w8 floats; labels: n x CCNN_SKIP_SLOTS x h8 x w8 bytes (0/1); loss = mean over all
 * outputs of pos_weight * softplus(-z) for label 1, softplus(z) for label 0.
 * grad: one forward/backward pass, gradient (CCNN_SKIP_PARAMS floats) and loss;
 * forward: the logits (n x SLOTS x h8 x w8); train: `epochs` passes of Adam (lr) over
 * shuffled batches of `batch` frames, params updated in place, epoch_loss[epochs] (may be
 * NULL) the mean batch loss of each pass. All return 0, or -1. */
int ccnn_skipnet_grad(const float* params, const float* x0, const unsigned char* labels, int n, int h8, int w8,
                      float pos_weight, float* grad, float* loss);
int ccnn_skipnet_forward(const float* params, const float* x0, int n, int h8, int w8, float* logits);
int ccnn_skipnet_train(float* params, const float* x0, const unsigned char* labels, int n, int h8, int w8,
                       int epochs, int batch, float lr, float pos_weight, unsigned seed, float* epoch_loss);

/* Kernel launches in the last captured batch graph and the number of captures so far. */
int ccnn_batch_stats(const ccnn_detector* d, int* launches, int* captures);

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
/* As ccnn_stage1 in a compute precision of this library (no counterpart in the
 * reference): 0 the exact FP32 path (ccnn_stage1), 1 or 2 the reduced-precision stage 1
 * (FP16 convolution and hidden-layer inputs and weights, products summed in FP32 by
 * v_dot2_f32_f16; activations, pooling and tanh in FP32). */
int ccnn_stage1_ex(const float* level, int lw, int lh, int precision, float* out);

/* ccnn_stage1 on device buffers, allocating nothing. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). level and out are device pointers
 * (lh x lw floats in, rh x rw out, both dense); the stage-1 weights, the level
 * descriptor and the L1-sum planes live in a fixed device array of
 * CCNN_NOALLOC_SCRATCH_FLOATS floats (64 MiB) that each GPU holds until the process
 * ends, which limits the level to about 16 million pixels (5,376 x 3,024 at 16:9): larger levels are
 * rejected with -1 and a message. Runs on the null stream and returns after the
 * result is written. Calls on one GPU share its array, so calls that overlap on one GPU
 * (several host threads) race. 0, or -1. */
#define CCNN_NOALLOC_SCRATCH_FLOATS (16u << 20)
int ccnn_stage1_noalloc(const float* level, int lw, int lh, float* out);
/* Stage-2/3 check of candidate windows on a gray image (CPUCheckDetect without the drop
 * logic): per candidate i, scale[i] and response-map position (x[i], y[i]); res[8*i..]
 * = {flag, mod, k1, k2, m1, m2, n2, n3}: acceptance, crop variant of the final evaluation,
 * stage-2/3 counts and maxima of that evaluation, and the number of stage-2 and stage-3
 * network evaluations performed. Settings from p. */
int ccnn_stage_check(const ccnn_param* p, const float* gray, int gw, int gh, int n,
                     const float* scale, const int* x, const int* y, float* res);

/* The stage-2 (stage 1) or stage-3 (stage 2) network (ConvNeuralNetwork::Forward on a
 * 47x51 check input, cnn_simd_cntk.cpp:964, as CPUCheckDetect runs it,
 * cnn_detector_v3.cpp:1646-1700) on n inputs of 47x51 floats (row-major, 47 columns)
 * laid out back to back; out: the 4x4 response map of each (16 floats per input). */
int ccnn_stage23(int stage, const float* in, int n, float* out);
/* As ccnn_stage23 in a compute precision of this library (no counterpart in the
 * reference): 0 the exact FP32 path, 1 L1 and L2 on the matrix cores with BF16 operands,
 * 2 with FP8 (e4m3) operands scaled per layer by powers of two; FP32 accumulation, L3 and
 * the hidden and output layers in FP32. */
int ccnn_stage23_ex(int stage, const float* in, int n, int precision, float* out);

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

/* How the reference's ordered replay (RunCheckDetect, cnn_detector_v3.cpp:1842-2011) treated
 * each candidate of ccnn_last_trace, in the same order: 0 dropped at the scan
 * (DropDetection), 1 dropped before its check (the level's drop, :1513-1533), 2 checked by
 * stages 2/3, 3 a raw window (detect_mode disable). Returns the candidate count; out may
 * be NULL. */
int ccnn_last_status(const ccnn_detector* d, int* out, int cap);

/* Stage-1 response maps of the frame that ccnn_last_trace describes: the levels in its
 * order, each a dense rh x rw map (ConvNeuralNetwork::Forward's response map of that
 * level, cnn_detector_v3.cpp:2109-2112), back to back. Returns the float count (out may
 * be NULL to query it), or -1 when there is no such frame or cap is smaller. */
long ccnn_last_responses(ccnn_detector* d, float* out, size_t cap);

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
/* Batch path with profiling on (no counterpart in the reference): each batch runs its
 * kernels without the graph, with HIP events between the kernel groups. ms (8 floats)
 * receives the last such batch's milliseconds for: gray (and the 8-bit pre-resizes), the
 * proposal network and tile masks, k_stage1_l1, k_stage1, the compaction, k_check, the
 * whole sequence, and the batch's frame count. 0, or -1. */
int ccnn_last_batch_timing(const ccnn_detector* d, float* ms);
/* Phase timestamps of the check and stage-1 kernels (no counterpart in the reference; for
 * profiling builds, compiled with CCNN_PHASE_STAMPS: python -m compactcnncascade_amd.build
 * --phase-stamps writes libccnn_phases.so). Copies nblocks x 24 values per workgroup of
 * kernel 0 (k_check) or 1 (k_stage1) into out (NULL: none), then clears them; call with the
 * device idle. Values are 100 MHz clock stamps by the workgroup's thread 0, 0 where not
 * reached. k_check (its last item): 0 item start, 1 crop prepared, 2 stage-2 input
 * written, 3-8 after the stage-2 network's L1, L2, sums, L3, hidden and output layers,
 * 9-14 the same for stage 3, 15 item end, 16 stage-3 input written; 20 hardware id, 21
 * item, 22 stage 3 ran. k_stage1 (per tile): 0 start, 1 L1-sum window in LDS, 2 L2 done,
 * 3 L2 sums in LDS, 4 thread 0's response written; 20 hardware id. 0, or -1 for a library
 * built without the stamps or arguments out of range. */
int ccnn_phase_stamps(int kernel, unsigned long long* out, int nblocks);
int ccnn_last_timing(const ccnn_detector* d, float* ms);

#ifdef __cplusplus
}
#endif
#endif /* CCNN_H */
