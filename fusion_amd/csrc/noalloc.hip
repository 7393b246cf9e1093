// ICPStep without allocation: the tracking state and partial rows of fu_icp_step live in a
// fixed device array of this library instead of memory the library allocates on first use.
// For callers that cannot allocate device memory (code under graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  No reference counterpart.
//
// Each GPU holds its own copy of the array at its own address until the process ends; the
// address is taken once per device with hipGetSymbolAddress while that device is current.
// Every call on a GPU uses that GPU's one array, so calls that overlap on one GPU race; the
// state a call reads is written by the call itself before its kernels run.
#include <hip/hip_runtime.h>
#include <mutex>

#include "fu_api.h"
#include "fu_batch.h"

using namespace fu;

namespace {

constexpr size_t kStateBytes = (sizeof(TrackState) + 255) / 256 * 256;
constexpr size_t kPartialBytes = (size_t)kIcpMaxBlocks * 64 * sizeof(float);
constexpr size_t kScratchBytes = kStateBytes + kPartialBytes;  // 1 MiB + the state

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
unsigned char* g_scratch[kMaxDevices] = {};

// the device address of the current device's array, taken once per device
fu_status scratch_of_current_device(unsigned char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return api_error(FU_ERR_HIP, "fu_icp_step_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return api_error(FU_ERR_ARG, "fu_icp_step_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> g(g_lock);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return api_error(FU_ERR_HIP, "fu_icp_step_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch[dev] = (unsigned char*)p;
    }
    *out = g_scratch[dev];
    return FU_OK;
}

}  // namespace

extern "C" fu_status fu_icp_step_noalloc(const float* next_vmap, const float* last_vmap, const float* next_nmap,
                                         const float* last_nmap, int cols, int rows, const float Rcurr[9],
                                         const float tcurr[3], const float Rlast[9], const float RlastInv[9],
                                         const float tlast[3], float fx, float fy, float cx, float cy,
                                         float residual[2], double A[36], double b[6], fu_stream s) {
    if (!next_vmap || !last_vmap || !next_nmap || !last_nmap || !Rcurr || !tcurr || !Rlast || !RlastInv || !tlast)
        return api_error(FU_ERR_ARG, "fu_icp_step_noalloc: null argument");
    // the reduction writes at most kIcpMaxBlocks partial rows whatever the frame size (larger
    // frames loop over pixel chunks), so the array bounds the work-group count, not the frame;
    // the pixel count must fit the kernels' int indices
    if (cols <= 0 || rows <= 0 || (long long)cols * rows > (1ll << 30))
        return api_error(FU_ERR_ARG, "fu_icp_step_noalloc: frame %dx%d outside 1..2^30 pixels", cols, rows);
    unsigned char* base = nullptr;
    fu_status st = scratch_of_current_device(&base);
    if (st) return st;
    return icp_step_on((TrackState*)base, (float*)(base + kStateBytes), next_vmap, last_vmap, next_nmap, last_nmap, cols,
                       rows, Rcurr, tcurr, Rlast, RlastInv, tlast, fx, fy, cx, cy, residual, A, b, (hipStream_t)s);
}
