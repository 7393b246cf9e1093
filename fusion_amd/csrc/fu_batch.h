// Batched launches: one launch runs the same stage for the frames of many sessions.
//
// Every kernel body takes its work-group coordinates as a WG value instead of reading
// blockIdx/gridDim.  A single-session launch passes the hardware ids; a batched launch splits
// its grid into consecutive shares, one per session of the step, and passes the ids within
// the share.  A session therefore sees exactly the work-group decomposition it would get
// alone, so its results (including the ICP partial sums and their order) are bit-identical
// alone and in a batch.
#pragma once

#include "fu_device.h"
#include "fu_kernels.h"
#include "../../include/fusion_amd.h"

namespace fu {

constexpr int kMaxBatch = 64;

struct WG {
    int x, y, nx, ny;
};
__device__ __forceinline__ WG hw_wg() { return WG{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y}; }

// Everything a session's kernels need, resident in device memory (written when the session
// joins a batch context and read with scalar loads by every batched kernel).
struct SessDesc {
    MapView m;
    TrackState* st;
    float* partials;
    int cols, rows;  // level 0
    int lc[3], lr[3];
    float depth_scale_inv, cutoff;
    float Kl[3][4];    // fx, fy, cx, cy per level (Frame::SetK, K / 2^l)
    float invf[3][2];  // (float)(1.0 / (double)fx_l), same for fy
    float* raw;
    float* depth[3];
    f4* vmap[3];
    f4* nmap[3];
    f4* model_v[3];
    f4* model_n[3];
    // photometric term (null unless the session tracks with it): intensity pyramid of this
    // frame and of the last tracked frame, Sobel derivatives
    unsigned char* image[3];
    unsigned char* last_image[3];
    short* didx[3];
    short* didy[3];
    int use_rgb;
    IcpArgs icp[3];
    int icp_blocks[3];
    int iters[3];
    int last_level;
    FuseArgs fuse;
    VisArgs vis;
    ZRangeArgs zr;
    RaycastArgs rc;
};

// Kernel argument of a batched launch (passed by value; about 1.1 KB).
struct BatchLaunch {
    const SessDesc* desc;
    int n;                    // sessions in this step
    int level;                // pyramid / ICP level of the launch
    int iter;                 // ICP iteration of the launch
    int uniform;              // > 0: every share has this many work-groups
    int xcd;                  // != 0: hardware work-groups map to logical ones XCD by XCD (batch_wg)
    int wave;                 // FU_WAVE_*: pixel layout of a raycast wave
    unsigned long long first;  // bit i: slot i is its session's first frame
    unsigned short sess[kMaxBatch];  // slot -> session descriptor index
    unsigned short gx[kMaxBatch];    // work-groups in x of the slot's share (2-D tilings)
    int pre[kMaxBatch + 1];          // prefix of the shares' work-group counts
    const void* in[kMaxBatch];       // per-slot input of the step (depth or colour)
};

// Logical work-group of this hardware work-group.  The dispatcher hands consecutive
// work-groups to the 8 XCDs in turn (work-group i runs on XCD i mod 8), and each XCD has its
// own L2.  With b.xcd set, XCD x runs one contiguous range of logical work-groups instead
// (its i-th work-group is logical x*q + min(x, r) + i, q = G / 8, r = G mod 8: a
// bijection), so the work-groups of one session share an L2.  Logical ids, and so every
// result, are the same either way.
__device__ __forceinline__ int batch_wg(const BatchLaunch& b) {
    const int i = (int)blockIdx.x;
    if (!b.xcd) return i;
    const int G = (int)gridDim.x;
    const int x = i & 7, q = G >> 3, r = G & 7;
    return x * q + (x < r ? x : r) + (i >> 3);
}

// slot of hardware work-group `wg` and the work-group ids within its share
__device__ __forceinline__ int batch_slot(const BatchLaunch& b, int wg, WG& g) {
    int s;
    if (b.uniform > 0) {
        s = wg / b.uniform;
    } else {
        int lo = 0, hi = b.n;  // pre[lo] <= wg < pre[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (b.pre[mid] <= wg)
                lo = mid;
            else
                hi = mid;
        }
        s = lo;
    }
    const int local = wg - b.pre[s];
    const int gx = b.gx[s];
    const int cnt = b.pre[s + 1] - b.pre[s];
    g.x = local % gx;
    g.y = local / gx;
    g.nx = gx;
    g.ny = cnt / gx;
    return s;
}

// Copy of a descriptor part, read through the constant address space: the descriptors do not
// change while a kernel runs, so the loads are scalar and the values live in SGPRs.
template <class T>
__device__ __forceinline__ T cload(const T& src) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized");
    T v;
    typedef const __attribute__((address_space(4))) unsigned CU;
    CU* q = (CU*)(const unsigned*)&src;
    unsigned* d = reinterpret_cast<unsigned*>(&v);
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) d[i] = q[i];
    return v;
}

__device__ __forceinline__ bool slot_first(const BatchLaunch& b, int s) { return (b.first >> s) & 1ull; }

// ---- batched launchers (host) ------------------------------------------------------------
// `h` is the host mirror of the descriptor array; the launchers size each share from it.
enum BatchTiling : int {
    BT_TILES = 0,   // 16x16 tiles of the level-`level` image
    BT_PYR3 = 1,    // 16x16 tiles of all three levels, level 0 first
    BT_ICP = 2,     // icp_blocks[level] work-groups
    BT_ONE = 3,     // one work-group per session
    BT_FIXED = 4,   // a fixed number of work-groups per session (grid-stride bodies)
    BT_ZTILES = 5,  // 256-thread groups over the z-range tiles
    BT_ICP_PAIRS = 7,  // (icp_blocks[level] + 1) / 2 work-groups (kRgbPixels chunks)
};
// fills b.pre, b.gx and b.uniform for one launch; returns the total work-group count
int batch_layout(BatchLaunch& b, const SessDesc* h, BatchTiling t, int level, int fixed);
// the process-wide launch geometry (fu_set_tuning)
const fu_tuning& tuning();
inline int xcd_order(int stage) { return (tuning().xcd_order >> stage) & 1; }

// per-frame result of a batched step, gathered on the device and copied back in one transfer
struct FrameBrief {
    double pose[16];
    int status;
    float icp_error;
    int icp_count;
    unsigned n_visible;
    unsigned n_alloc;
    unsigned flags;
    int rgb_count;  // photometric residuals of the last iteration (0 without the term)
    unsigned pad;
};

// pyramid.hip
void launch_b_resize(BatchLaunch& b, const SessDesc* h, hipStream_t s);  // levels 1 and 2
void launch_b_filter(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_pyrdown(BatchLaunch& b, const SessDesc* h, int level, hipStream_t s);
void launch_b_nmap(BatchLaunch& b, const SessDesc* h, hipStream_t s);
// intensity pyramid and derivatives (sessions with the photometric term only), and the
// post-tracking copy of the pyramid into the last frame's
void launch_b_images(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_image_commit(BatchLaunch& b, const SessDesc* h, hipStream_t s);
// icp.hip
void launch_b_track_begin(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_icp(BatchLaunch& b, const SessDesc* h, int level, int iter, hipStream_t s);
void launch_b_rgb_lists(BatchLaunch& b, const SessDesc* h, hipStream_t s);  // once per frame
void launch_b_track_commit(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_results(BatchLaunch& b, FrameBrief* out, hipStream_t s);
// map.hip
void launch_b_alloc(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_visible(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_integrate(BatchLaunch& b, const SessDesc* h, int fmt, hipStream_t s);
// raycast.hip
void launch_b_zrange(BatchLaunch& b, const SessDesc* h, hipStream_t s);
void launch_b_raycast(BatchLaunch& b, const SessDesc* h, int fmt, hipStream_t s);

// capi.hip: batch_grab with an optional device result gather (brief[0..n)) and an optional
// event recorded on `s` once the step's kernels have read its frames (after the pyramid stage;
// after the whole step when it runs as a captured graph), and the error slot behind
// fu_last_error for the other translation units of the library
fu_status batch_grab(fu_batch* bt, int n, fu_system* const* sys, const uint16_t* const* depth,
                     const uint8_t* const* rgb, FrameBrief* brief, hipStream_t s, hipEvent_t inputs_read = nullptr);
fu_status api_error(fu_status code, const char* fmt, ...);
// mesh extraction of a system's map queued on `s` without synchronising; the triangle and
// occupied-block counts land in the map's counters, whose device address `counters` returns
fu_status mesh_async(fu_system* sys, float* vertex, float* normal, uint8_t* color, uint32_t max_triangles,
                     hipStream_t s);
const unsigned* system_counters(fu_system* sys);
// ICPStep on a caller-given tracking state and kIcpMaxBlocks x 64-float partial-row buffer
fu_status icp_step_on(TrackState* st, float* partials, const float* next_vmap, const float* last_vmap,
                      const float* next_nmap, const float* last_nmap, int cols, int rows, const float Rcurr[9],
                      const float tcurr[3], const float Rlast[9], const float RlastInv[9], const float tlast[3],
                      float fx, float fy, float cx, float cy, float residual[2], double A[36], double b[6],
                      hipStream_t s);

}  // namespace fu
