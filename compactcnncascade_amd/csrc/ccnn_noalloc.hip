// Stage 1 of one pyramid level on device buffers with no allocation: the variant of
// ccnn_stage1 for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc). The intermediate data
// that ccnn_stage1 allocates per call -- the level descriptor, the stage-1 weights and
// the L1-sum planes k_stage1_l1 writes and k_stage1 reads -- live in one file-scope
// device array instead, of fixed size, that each GPU the library runs on holds until the
// process ends.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "ccnn.h"
#include "ccnn_detector.h"

namespace ccnn {
namespace {

// Layout of the scratch array (floats): the level descriptor, then the stage-1 weights,
// then the L1-sum planes.
constexpr size_t NA_DESC = 64;                                    // 256 bytes >= sizeof(LevelDesc)
constexpr size_t NA_W = (Stage1::SIZE + Stage1H::SIZE + 63) / 64 * 64;
constexpr size_t NA_PLANES = (size_t)CCNN_NOALLOC_SCRATCH_FLOATS - NA_DESC - NA_W;
static_assert(sizeof(LevelDesc) <= NA_DESC * sizeof(float), "descriptor slot");

__device__ __align__(256) float scratch[CCNN_NOALLOC_SCRATCH_FLOATS];

// Writes the level descriptor (a kernel argument) into the scratch array.
__global__ void k_put_level(LevelDesc L, LevelDesc* __restrict__ dst)
{
    if (threadIdx.x == 0) *dst = L;
}

// The array's device address on each GPU (hipGetSymbolAddress, once per device, with
// that device current): host code passes these to kernels and copies, never the array's
// name, which on the host is only a stand-in for the symbol.
constexpr int MAX_DEV = 64;
float* g_ptr[MAX_DEV] = {};
std::mutex g_mu;

// Scratch of the current device, its weights uploaded at set-up; nullptr on failure.
float* device_scratch()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[dev]) return g_ptr[dev];
    float* p = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        std::fprintf(stderr, "[ccnn] hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
        return nullptr;
    }
    NetWeights net;
    const unsigned char* b;
    size_t n;
    embedded_model(0, &b, &n);
    std::string err;
    if (load_model(b, n, 0, 1, net, err) < 0) return nullptr;
    // the weights are written once per device and never overwritten: every call reads them
    e = hipMemcpy(p + NA_DESC, net.w.data(), sizeof(float) * net.w.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        std::fprintf(stderr, "[ccnn] scratch weight upload failed: %s\n", hipGetErrorString(e));
        return nullptr;
    }
    g_ptr[dev] = p;
    return p;
}

}  // namespace
}  // namespace ccnn

extern "C" int ccnn_stage1_noalloc(const float* level, int lw, int lh, float* out)
{
    int rw, rh;
    if (!level || !out || ccnn_stage1_size(lw, lh, &rw, &rh) < 0) return -1;
    if (lw < 23 || lh < 27) return -1;
    ccnn::LevelDesc L{};
    L.lw = lw; L.lh = lh; L.rw = rw; L.rh = rh; L.mode = 2; L.scale = 1.f;
    L.tiles_x = (rw + ccnn::S1_TX - 1) / ccnn::S1_TX;
    L.tiles_y = (rh + ccnn::S1_TY - 1) / ccnn::S1_TY;
    const size_t planes = ccnn::level_l1_layout(L);
    if (planes > ccnn::NA_PLANES) {
        std::fprintf(stderr,
                     "[ccnn] ccnn_stage1_noalloc: a %dx%d level needs %zu floats of L1-sum planes, more than the "
                     "%zu its fixed scratch holds\n",
                     lw, lh, planes, ccnn::NA_PLANES);
        return -1;
    }
    float* p = ccnn::device_scratch();
    if (!p) return -1;
    ccnn::LevelDesc* d_desc = reinterpret_cast<ccnn::LevelDesc*>(p);
    // the descriptor goes in as a kernel argument, ahead of the stage-1 launches on the same
    // stream; the planes need no clearing: the positions k_stage1_l1 leaves unwritten are
    // read only by k_stage1 threads whose results fall outside the response map
    ccnn::k_put_level<<<1, 64, 0, nullptr>>>(L, d_desc);
    if (hipGetLastError() != hipSuccess) return -1;
    if (ccnn::launch_stage1(level, lw, lw, lh, d_desc, 1, L.l1_tiles_x * L.l1_tiles_y, L.tiles_x * L.tiles_y,
                            p + ccnn::NA_DESC, p + ccnn::NA_DESC + ccnn::NA_W, out, INFINITY, nullptr, nullptr, nullptr,
                            0, nullptr, nullptr) != hipSuccess)
        return -1;
    return hipStreamSynchronize(nullptr) == hipSuccess ? 0 : -1;
}
