// Test-infrastructure harness around the reference CPU path (Bkmz21/CompactCNNCascade).
//
// This file is compiled together with the reference's own sources, where they lie
// under the reference checkout, by oracle/Makefile into oracle/_ref/libccnn_ref*.so.
// It is a checker: only tests, smoke() and bench.py's cpu_baseline leg load it.
// Every function is a thin extern "C" wrapper over one public entry point of the
// reference, so the oracle restatement (oracle/ccnn_oracle.c) can be compared
// with the compiled reference on identical inputs:
//   ref_gray        -> SIMD::ImageConverter::Img8uToImg32fGRAY   (image_proc.cpp:299-317)
//   ref_resize      -> SIMD::ImageResizer::FastImageResize (32f) (image_resize.cpp:564-584)
//   ref_resize_u8   -> SIMD::ImageResizer::FastImageResize (8u)  (image_resize.cpp:543-563)
//   ref_cnn_*       -> SIMD::ConvNeuralNetwork Init/AllocateMemory/Forward (cnn_simd_cntk.cpp:42,405,964)
//   ref_check_prep  -> FloatToUChar + equalizeImage + UCharToFloat(_inv) (image_proc.cpp:328-491,720-800)
//   ref_detect      -> CNNDetector::Detect with FaceDetector's parameter conversion
//                      (CompactCNNLibAPI_v2.cpp:100-138, cnn_detector_v3.cpp:2483-2813)
//   ref_cntk_dump2binary -> FaceDetector::CNTKDump2Binary       (CompactCNNLibAPI_v2.cpp:509-519)

#include "CompactCNNLibAPI_v2.h"
#include "cnn_detector_v3.h"
#include "image_proc.h"
#include "image_resize.h"
#include "cnn_simd_cntk.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>
#include <string>
#include <fstream>
#include <sstream>

using namespace NeuralNetworksLib;

extern "C" {

int ref_gray(const unsigned char* img, int w, int h, int ch, int step, float* out, int out_step)
{
    SIMD::Image_8u src(w, h, ch, const_cast<unsigned char*>(img), step);
    SIMD::Image_32f dst(w, h, 1, out, out_step);
    return SIMD::ImageConverter::Img8uToImg32fGRAY(dst, src, 1);
}

// Resizer constructed for (dst size, src size) exactly as the detector's per-level
// resizers (cnn_detector_v3.cpp:526) and check resizer (cnn_detector_v3.cpp:829).
int ref_resize(const float* src, int sw, int sh, int sstep,
               float* dst, int dw, int dh, int dstep, int type, int lut_sw, int lut_sh)
{
    SIMD::Image_32f s(sw, sh, 1, const_cast<float*>(src), sstep);
    SIMD::Image_32f d(dw, dh, 1, dst, dstep);
    if (lut_sw <= 0) lut_sw = sw;
    if (lut_sh <= 0) lut_sh = sh;
    SIMD::ImageResizer rz(Size(dw, dh), Size(lut_sw, lut_sh));
    rz.FastImageResize(d, s, type, 1);
    return 0;
}

int ref_resize_u8(const unsigned char* src, int sw, int sh, int sstep,
                  unsigned char* dst, int dw, int dh, int dstep, int type)
{
    SIMD::Image_8u s(sw, sh, 1, const_cast<unsigned char*>(src), sstep);
    SIMD::Image_8u d(dw, dh, 1, dst, dstep);
    SIMD::ImageResizer rz(Size(dw, dh), Size(sw, sh));
    rz.FastImageResize(d, s, type, 1);
    return 0;
}

void* ref_cnn_create(const char* model_path, int index_output)
{
    SIMD::ConvNeuralNetwork* cnn = new SIMD::ConvNeuralNetwork();
    cnn->Init(std::string(model_path), index_output, 0);
    if (cnn->isEmpty()) { delete cnn; return nullptr; }
    return cnn;
}

void ref_cnn_destroy(void* h)
{
    delete static_cast<SIMD::ConvNeuralNetwork*>(h);
}

int ref_cnn_min_size(void* h, int* w, int* hh)
{
    Size s = static_cast<SIMD::ConvNeuralNetwork*>(h)->getMinInputImgSize();
    *w = s.width; *hh = s.height;
    return 0;
}

// Forward over a w x h image (row stride `step` floats). The network buffers are
// allocated at (alloc_w, alloc_h), as the detector allocates stage 1 at its largest
// level (cnn_detector_v3.cpp:620-625). Output copied densely into out[oh][ow].
int ref_cnn_forward(void* h, const float* img, int w, int hgt, int step,
                    int alloc_w, int alloc_h, float* out, int cap, int* ow, int* oh)
{
    SIMD::ConvNeuralNetwork* cnn = static_cast<SIMD::ConvNeuralNetwork*>(h);
    cnn->AllocateMemory(Size(alloc_w, alloc_h));
    SIMD::Image_32f in(w, hgt, 1, const_cast<float*>(img), step);
    SIMD::Image_32f resp;
    cnn->Forward(resp, in);
    *ow = resp.width; *oh = resp.height;
    if (resp.width * resp.height > cap) return -1;
    for (int y = 0; y < resp.height; ++y)
        for (int x = 0; x < resp.width; ++x)
            out[y * resp.width + x] = resp.data[y * resp.widthStep + x];
    return 0;
}

// The per-candidate image preparation of CPUCheckDetect (cnn_detector_v3.cpp:1558-1644):
// NN resize of the crop to 49x51, FloatToUChar, optional equalizeImage, then the two
// 47x51 float inputs of stage 2 (offset 2) and stage 3 (mirrored or offset 0).
int ref_check_prep(const float* crop, int cw, int chh, int cstep, int lut_sw, int lut_sh,
                   int equalize, int reflection,
                   unsigned char* u8_out, float* s2_out, float* s3_out)
{
    SIMD::Image_32f c(cw, chh, 1, const_cast<float*>(crop), cstep);
    std::vector<float> r32buf(49 * 51);
    SIMD::Image_32f r32(49, 51, 1, r32buf.data(), 49);
    SIMD::ImageResizer rz(Size(49, 51), Size(lut_sw, lut_sh));
    rz.FastImageResize(r32, c, 0, 1);
    SIMD::Image_8u u8(49, 51, 1, u8_out, 49);
    SIMD::ImageConverter::FloatToUChar(u8, r32, Rect(0, 0, 49, 51));
    if (equalize) SIMD::equalizeImage(u8);
    SIMD::Image_32f f2(47, 51, 1, s2_out, 47);
    SIMD::ImageConverter::UCharToFloat(f2, u8, 2);
    SIMD::Image_32f f3(47, 51, 1, s3_out, 47);
    if (reflection) SIMD::ImageConverter::UCharToFloat_inv(f3, u8);
    else SIMD::ImageConverter::UCharToFloat(f3, u8, 0);
    return 0;
}

int ref_equalize(unsigned char* img, int w, int h, int step)
{
    SIMD::Image_8u im(w, h, 1, img, step);
    SIMD::equalizeImage(im);
    return 0;
}

// Full detection with the public API's parameter conversion
// (CompactCNNLibAPI_v2.cpp:100-138 paramConverter, :218-353 Detect), driving
// CNNDetector directly. The API forces gray_image_only on and, for a colour image,
// the detector turns it off again and re-initialises (cnn_detector_v3.cpp:2540-2546);
// starting with it off for colour input is the same computation. It is left on for
// 1-channel input. (On Linux the API's own Init path with the pre-resize buffer
// aborts: TmpImage<uchar> asks posix_memalign for alignment 1, image.cpp:40.)
// iparams: [max_w, max_h, min_face, max_face, min_neighbors, precision,
//           drop_detect, equalize, reflection, detect_mode]
// fparams: [scale_factor, t1, t2, t3]
static CNNDetector* make_detector(const int* iparams, const float* fparams, int channels, bool facial = false)
{
    CNNDetector::Param cp;
    CNNDetector::AdvancedParam ap;
    cp.max_image_size = Size(iparams[0], iparams[1]);
    cp.min_obj_size.height = iparams[2];
    cp.max_obj_size.height = iparams[3];
    cp.min_neighbors = iparams[4];
    ap.detect_precision = static_cast<CNNDetector::DetectPrecision>(iparams[5]);
    ap.drop_detect = iparams[6] != 0;
    ap.equalize = iparams[7] != 0;
    ap.reflection = iparams[8] != 0;
    ap.detect_mode = static_cast<CNNDetector::DetectMode>(iparams[9]);
    cp.scale_factor = fparams[0];
    ap.treshold_1 = fparams[1];
    ap.treshold_2 = fparams[2];
    ap.treshold_3 = fparams[3];
    cp.pipeline = CNNDetector::Pipeline::CPU;
    cp.num_threads = 1;
    for (int i = 0; i < 3; ++i) ap.path_model[i] = "";
    ap.index_output[0] = 1; ap.index_output[1] = 0; ap.index_output[2] = 0;
    ap.device_info = false;
    ap.facial_analysis = facial;
    ap.double_check = true;
    ap.adapt_min_neighbors = false;
    ap.gray_image_only = channels == 1;
    ap.packet_detection = false;
    ap.merger_detect = ap.detect_mode != CNNDetector::DetectMode::disable;
    ap.blur = false;
    ap.uniform_noise = false;
    CNNDetector* d = new CNNDetector(&cp, &ap);
    if (d->isEmpty()) { delete d; return nullptr; }
    return d;
}

// out: 5 values per face (x, y, w, h as float, score), as FaceDetector::Detect copies them.
int ref_detect(const int* iparams, const float* fparams,
               const unsigned char* img, int w, int h, int ch, int step,
               float* out, int cap)
{
    CNNDetector* d = make_detector(iparams, fparams, ch);
    if (!d) return -1;
    std::vector<CNNDetector::Detection> dets;
    SIMD::Image_8u image(w, h, ch, const_cast<unsigned char*>(img), step);
    int rc = d->Detect(dets, image);
    delete d;
    if (rc < 0) return rc;
    int n = (int)dets.size();
    int m = n < cap ? n : cap;
    for (int i = 0; i < m; ++i)
    {
        out[5 * i + 0] = (float)dets[i].rect.x;
        out[5 * i + 1] = (float)dets[i].rect.y;
        out[5 * i + 2] = (float)dets[i].rect.width;
        out[5 * i + 3] = (float)dets[i].rect.height;
        out[5 * i + 4] = dets[i].score;
    }
    return n;
}

// 8 values per face (x, y, w, h, score, gender, smile, glasses), at most cap written.
// gender / smile / glasses are reported the way FaceDetector::Detect does
// (CompactCNNLibAPI_v2.cpp:296-319): the facial_data entries of the detection sorted by
// the key in descending order, the element at size/2; -1 when there are none.
static int write_faces(std::vector<CNNDetector::Detection>& dets, float* out, int cap);

// As ref_detect with facial analysis selectable (iparams[10]); out as write_faces.
int ref_detect_ex(const int* iparams, const float* fparams,
                  const unsigned char* img, int w, int h, int ch, int step,
                  float* out, int cap)
{
    CNNDetector* d = make_detector(iparams, fparams, ch, iparams[10] != 0);
    if (!d) return -1;
    std::vector<CNNDetector::Detection> dets;
    SIMD::Image_8u image(w, h, ch, const_cast<unsigned char*>(img), step);
    int rc = d->Detect(dets, image);
    delete d;
    if (rc < 0) return rc;
    return write_faces(dets, out, cap);
}

// Detect over a sequence of frames on ONE detector, whose state (max_image_size,
// num_scales, gray_image_only) carries from call to call as in the API's video loop
// (prebuild/examples/main.cpp:36-111). Frames lie back to back in `frames`;
// meta[5*i..] = {w, h, channels, step, min_face}: a min_face > 0 is set with
// setMinObjectSize before the frame (CompactCNNLibAPI_v2.cpp:426-431). Frame i's faces go
// to out + 8*cap*i (as write_faces), its face count to counts[i]. The detector is built
// for the first frame's channel count (colour first: gray_image_only off, see make_detector).
int ref_detect_seq(const int* iparams, const float* fparams, const unsigned char* frames, const int* meta, int n,
                   float* out, int cap, int* counts)
{
    if (n <= 0) return 0;
    CNNDetector* d = make_detector(iparams, fparams, meta[2], iparams[10] != 0);
    if (!d) return -1;
    size_t off = 0;
    for (int i = 0; i < n; ++i)
    {
        const int w = meta[5 * i], h = meta[5 * i + 1], ch = meta[5 * i + 2], step = meta[5 * i + 3];
        if (meta[5 * i + 4] > 0) d->setMinObjectSize(Size(meta[5 * i + 4], meta[5 * i + 4]));
        std::vector<CNNDetector::Detection> dets;
        SIMD::Image_8u image(w, h, ch, const_cast<unsigned char*>(frames + off), step);
        const int rc = d->Detect(dets, image);
        if (rc < 0) { delete d; return rc; }
        counts[i] = write_faces(dets, out + (size_t)8 * cap * i, cap);
        off += (size_t)step * h;
    }
    delete d;
    return 0;
}

static int write_faces(std::vector<CNNDetector::Detection>& dets, float* out, int cap)
{
    int n = (int)dets.size();
    int m = n < cap ? n : cap;
    for (int i = 0; i < m; ++i)
    {
        float* o = out + 8 * i;
        o[0] = (float)dets[i].rect.x;
        o[1] = (float)dets[i].rect.y;
        o[2] = (float)dets[i].rect.width;
        o[3] = (float)dets[i].rect.height;
        o[4] = dets[i].score;
        o[5] = o[6] = o[7] = -1.f;
        std::vector<CNNDetector::FacialData> fd = dets[i].facial_data;
        const size_t k = fd.size();
        if (k == 0) continue;
        std::sort(fd.begin(), fd.end(), [](const CNNDetector::FacialData& a, const CNNDetector::FacialData& b)
                  { return a.gender > b.gender; });
        o[5] = (float)fd[k / 2].gender;
        std::sort(fd.begin(), fd.end(), [](const CNNDetector::FacialData& a, const CNNDetector::FacialData& b)
                  { return a.smile > b.smile; });
        o[6] = (float)fd[k / 2].smile;
        std::sort(fd.begin(), fd.end(), [](const CNNDetector::FacialData& a, const CNNDetector::FacialData& b)
                  { return a.glasses > b.glasses; });
        o[7] = (float)fd[k / 2].glasses;
    }
    return n;
}

// Repeated detection on one initialised detector, for CPU-baseline timing: the
// detector is initialised once and Detect is called `reps` times over `nframes`
// frames laid out back to back. Returns the total number of faces found.
long ref_detect_loop(const int* iparams, const float* fparams,
                     const unsigned char* frames, int nframes, int w, int h, int ch, int step,
                     int reps)
{
    CNNDetector* d = make_detector(iparams, fparams, ch);
    if (!d) return -1;
    long total = 0;
    const size_t fbytes = (size_t)step * h;
    std::vector<CNNDetector::Detection> dets;
    for (int r = 0; r < reps; ++r)
        for (int f = 0; f < nframes; ++f)
        {
            dets.clear();
            SIMD::Image_8u image(w, h, ch, const_cast<unsigned char*>(frames + f * fbytes), step);
            if (d->Detect(dets, image) < 0) { delete d; return -1; }
            total += (long)dets.size();
        }
    delete d;
    return total;
}

void ref_cntk_dump2binary(const char* binary_file, const char* cntk_model_dump)
{
    CompactCNNLib::FaceDetector::CNTKDump2Binary(binary_file, cntk_model_dump);
}

}  // extern "C"
