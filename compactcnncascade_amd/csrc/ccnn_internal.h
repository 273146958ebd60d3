// Shared host/device definitions of the MI355X detection path.
#pragma once

#include <cstdint>
#include <cstddef>

namespace ccnn {

// ---------------------------------------------------------------------------------
// Network shapes. The detector supports the cascade's fixed prototypes (the shipped
// cnn4face{1,2,3}_cntk.bin models): three conv layers with 2x2 max-pooling after
// the first two, sums of neighbouring maps feeding two convolutions each, a hidden
// layer of groups over consecutive L3 maps, and NOUT output neurons (one for the cascade
// stages, whose index_output selects it; all 13 for the landmarks network).
// (Prototype: cntk/stage_1.cntk; loader: cnn_simd_cntk.cpp:42-404.)
// ---------------------------------------------------------------------------------
template <int M1_, int K1_, int M2_, int M3_, int K3W_, int K3H_, int HLG_, int HLS_, int HLN_, int NOUT_ = 1>
struct NetShape {
    static constexpr int M1 = M1_, K1 = K1_, M2 = M2_, M3 = M3_, K3W = K3W_, K3H = K3H_;
    static constexpr int HLG = HLG_, HLS = HLS_, HLN = HLN_, HL = HLG_ * HLS_, NOUT = NOUT_;
    // flat float layout of the weights in device memory
    static constexpr int oK1 = 0;
    static constexpr int oB1 = oK1 + M1 * K1 * K1;     // conv_b, lrelu_w1, lrelu_w2, bn_b per map
    static constexpr int oK2 = oB1 + 4 * M1;
    static constexpr int oB2 = oK2 + M2 * 9;
    static constexpr int oK3 = oB2 + 4 * M2;
    static constexpr int oB3 = oK3 + M3 * K3W * K3H;
    static constexpr int oHW = oB3 + 4 * M3;            // hl_w[HL][HLN]
    static constexpr int oHB = oHW + HL * HLN;          // hl_b[HL]
    static constexpr int oHT = oHB + HL;                // hl_tanh_w[HLG]
    static constexpr int oHBW = oHT + HLG;              // hl_bn_w[HLS]
    static constexpr int oHBB = oHBW + HLS;             // hl_bn_b[HLS]
    static constexpr int oOW = oHBB + HLS;              // ol_w[NOUT][HL] (row index_output if NOUT 1)
    static constexpr int oOB = oOW + NOUT * HL;         // ol_b[NOUT], ol_tanh_w, af_scale
    static constexpr int SIZE = oOB + NOUT + 2;
};

// stage 1: 23x27 window, maps 4/8/16, kernels 4x4 / 3x3 / 4w x 5h, HL 13 x 4 (N = 4)
using Stage1 = NetShape<4, 4, 8, 16, 4, 5, 13, 4, 4>;
// stage 2: 36x40 pattern on a 47x51 input, maps 6/12/24, 4x4 / 3x3 / 7w x 8h, HL 21 x 4
using Stage2 = NetShape<6, 4, 12, 24, 7, 8, 21, 4, 4>;
// stage 3: same convolutions, HL 17 x 2 (N = 8)
using Stage3 = NetShape<6, 4, 12, 24, 7, 8, 17, 2, 8>;
// facial analysis (cnn4landmarks_cntk.bin): 49x51 input, maps 24/48/96, 4x4 / 3x3 /
// 10w x 11h (one L3 value per map), fully connected HL of 384 units over the 96 maps,
// 13 outputs: five (x, y) landmarks, gender, smile, glasses
using Landmarks = NetShape<24, 4, 48, 96, 10, 11, 1, 384, 96, 13>;

// One pyramid level as the kernels see it.
struct LevelDesc {
    int lw, lh;           // level image size
    int rw, rh;           // response map size
    int mode;             // 0 nearest, 1 bilinear, 2 copy
    int scl;              // scale index (InitScales order)
    uint32_t sclx, scly;  // Q12 steps of the level's resize LUT
    int resp_off;         // offset of the response map in the frame's response buffer
    int tile_off;         // first stage-1 tile of the level
    int tiles_x, tiles_y; // stage-1 tiles
    int chunk_off;        // first candidate-ordering chunk of the level
    int chunks;           // candidate-ordering chunks
    float scale;          // pyramid scale
    int s1_off;           // offset of the level's L1-sum planes in the frame's L1 buffer
    int s1_w, s1_h;       // L1-sum plane size: every pooled L1 position the tiles read
    int l1_tile_off;      // first L1 tile of the level
    int l1_tiles_x, l1_tiles_y;
};

// Stage 1 runs in two launches. The first computes the four L1-sum maps (conv, act,
// pool, neighbour sums) of every level once, S1A_TX x S1A_TY pooled positions per
// workgroup; the second computes TX x TY response values per workgroup from them.
constexpr int S1A_TX = 64, S1A_TY = 16;
constexpr int S1A_THREADS = 256;
constexpr int S1_TX = 16, S1_TY = 16;
constexpr int S1_THREADS = 256;
// Chunk of raster positions of one level by which the host buckets the candidate list
// before ordering it.
constexpr int CMP_CHUNK = 1024;
constexpr int CHECK_THREADS = 256;
constexpr int FA_THREADS = 512;
constexpr int MAX_LEVELS = 512;

// Candidate window found by the stage-1 scan, in the reference's check order.
struct Candidate {
    int level;            // index into the frame's level table (compute order)
    int x, y;             // response-map coordinates
    float score;          // stage-1 response
};

// Facial analysis of a candidate the checks accept (cnn_detector_v3.cpp:1731-1756), on
// the crop of its final crop variant: ok 0 when all five landmarks fall in the crop's
// top-left quadrant (the window is then dropped); otherwise the rounded attributes.
struct FaResult {
    int ok;
    int gender, smile, glasses;
};

// Stage-2/3 outcome of one candidate's check (mod 0, then mod 1 if mod 0 fails).
struct CheckResult {
    int flag;             // object accepted
    int mod;              // crop variant of the final evaluation
    int k1, k2;           // responses above the stage-2/3 thresholds
    float m1, m2;         // maxima of the stage-2/3 responses
    int n2, n3;           // stage-2 and stage-3 network evaluations performed
    FaResult fa;          // facial analysis of an accepted window (ok 1, -1s when off)
};

// Scalar parameters of the check kernel.
struct CheckParams {
    float t2, t3;
    int type_check;
    int min_neighbors;    // already adapted if adapt_min_neighbors
    int adapt;            // adapt_min_neighbors (m-1 when scale > 0.7 and m > 1)
    int double_check;
    int equalize, reflection;
};

}  // namespace ccnn
