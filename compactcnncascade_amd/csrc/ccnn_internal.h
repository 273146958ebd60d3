// Shared host/device definitions of the MI355X detection path.
#pragma once

#include <cstdint>
#include <cstddef>

#include "ccnn.h"

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

// FP16 tap pairs of the stage-1 network for the reduced-precision path, as 32-bit words
// (two FP16 values each) after the FP32 weights: per L1 map and kernel row (k0, k1) and
// (k2, k3); per L2 map and row (k0, k1) and (k2, 0); per L3 map and row (k0, k1) and
// (k2, k3); per hidden unit (w0, w1) and (w2, w3); then, as FP32 bits, the sum of each
// L1 map's 16 FP32 taps (the FP16 L1 convolution runs on the window less its centre
// value and adds centre x sum back in FP32).
struct Stage1H {
    static constexpr int oL1 = 0;
    static constexpr int oL2 = oL1 + Stage1::M1 * Stage1::K1 * 2;
    static constexpr int oL3 = oL2 + Stage1::M2 * 3 * 2;
    static constexpr int oHL = oL3 + Stage1::M3 * Stage1::K3H * 2;
    static constexpr int oS1 = oHL + Stage1::HL * 2;
    static constexpr int SIZE = oS1 + Stage1::M1;
};

// Weights of the stage-2/3 networks for the matrix-core path, expanded per lane of the
// 16x16x32 MFMA operand (lane l holds column l % 16 and K samples 8*(l/16) .. +7):
// L1: two column tiles (6 maps x 4 pooling-quad offsets, padded to 32) x 64 lanes x 8;
// L2: three sum pairs x 64 lanes x 8; then the FP8 weight scales of L1 and L2 (powers of
// two), which the kernels apply before rounding to FP8.
struct MfmaW {
    static constexpr int oL1 = 0;
    static constexpr int oL2 = oL1 + 2 * 64 * 8;
    static constexpr int oScale = oL2 + 3 * 64 * 8;
    static constexpr int SIZE = oScale + 2;
};

// One frame of a skip-net batch (ccnn_skipnet.hip): its grid of 8x8 cells and where its
// planes start (a layer buffer holds the frames back to back, C planes of h8 x w8 each).
struct SnFrame {
    int h8, w8;
    int off;              // the frame's first cell in the batch (its planes start at off * C)
    int pos_off;          // the same in the flattened (frame, y, x) numbering of the launches
};

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
    // the level's frame in a batch of frames (gp > 0): frame index, offset (floats) of its
    // gray image in the batch's gray buffer, pitch and size. gp == 0 on the single-frame
    // path, where the kernels take the gray image's geometry as arguments.
    int frame;
    int gp, gw, gh;
    long long gray_off;
};

// A frame of a batch as the batched gray kernel sees it.
struct FrameDesc {
    const uint8_t* src;   // 8-bit frame (device memory): gray, BGR or BGRA rows of `step` bytes
    int w, h, ch, step;
    // a 1-channel frame that is pre-resized (gray_image_only): the original frame, which
    // the 8-bit bilinear resize writes into src (w x h, step w)
    const uint8_t* orig;
    int ow, oh, ostep;
    long long gray_off;   // offset (floats) of its FP32 gray image in the batch's gray buffer
    int gp;               // gray pitch (floats)
    int qw;               // groups of 16 pixels per row
    int quads_off;        // first group of the frame in the batch's group numbering
    int aligned;          // bit 0: 4-byte (16-byte for BGRA) aligned rows; bit 1: 16-byte aligned rows
};

// Stage 1 runs in two launches. The first computes the four L1-sum maps (conv, act,
// pool, neighbour sums) of every level once, AX x AY pooled positions per workgroup; the
// second computes TX x TY response values per workgroup from them. The kernels are built
// for the tile shapes below (shape id = L1 shape index * N_RESP_SHAPES + response shape
// index); a plan is laid out for one of them, which the autotuner picks per workload.
// Shape 0 (64 x 16 L1 tiles, 16 x 16 response tiles) is the default.
constexpr int N_L1_SHAPES = 4, N_RESP_SHAPES = 3, N_TILE_SHAPES = N_L1_SHAPES * N_RESP_SHAPES;
constexpr int L1_SHAPES[N_L1_SHAPES][2] = {{64, 16}, {32, 16}, {32, 32}, {64, 8}};
constexpr int RESP_SHAPES[N_RESP_SHAPES][2] = {{16, 16}, {32, 8}, {16, 8}};
struct TileShape {
    int ax, ay;           // pooled L1 positions per L1 tile
    int tx, ty;           // responses per response tile
};
inline TileShape tile_shape(int id)
{
    if (id < 0 || id >= N_TILE_SHAPES) id = 0;
    const int a = id / N_RESP_SHAPES, r = id % N_RESP_SHAPES;
    return TileShape{L1_SHAPES[a][0], L1_SHAPES[a][1], RESP_SHAPES[r][0], RESP_SHAPES[r][1]};
}
constexpr int S1A_TX = 64, S1A_TY = 16;      // shape 0
constexpr int S1A_THREADS = 256;
constexpr int S1_TX = 16, S1_TY = 16;
constexpr int S1_THREADS = 256;
// Chunk of raster positions of one level by which the host buckets the candidate list
// before ordering it.
constexpr int CMP_CHUNK = 1024;
constexpr int CHECK_THREADS = 256;
// Positions per workgroup of the ordered candidate compaction (256 threads x 16).
constexpr int CMP_BLOCK = 4096;
constexpr int CMP_SCAN_THREADS = 1024;
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

// Logit thresholds of the tile mask, one per scale slot (a tile of a level whose slot is s
// runs when its largest logit reaches zt[s]).
struct SkipThresholds {
    float zt[CCNN_SKIP_SLOTS];
};

}  // namespace ccnn
