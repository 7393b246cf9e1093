// Argument structs and launchers of the HIP kernels (host <-> device interface inside the
// library; the public C ABI is include/fusion_amd.h).
#pragma once

#include <hip/hip_runtime.h>
#include "fu_math.h"

namespace fu {

struct MapView;

enum TrackStatus : int {
    TS_RUNNING = 0,
    TS_OK = 1,        // accepted (Tracking.cc:402-403)
    TS_FAILED = 2,    // NaN error or too few correspondences (:371-375)
    TS_REJECTED = 3,  // final error / motion test failed (:404-407)
    TS_FIRST = 4,     // first frame: identity pose, no tracking (Tracking.cc:66-71)
    TS_LOST = 5,      // lost (more than 9 missed frames, Tracking.cc:92-96): not tracked, nothing fused
    TS_RELOC = 6,     // lost, relocalised at a pose given from outside: no tracking, raycast there
};

// what a frame does after tracking (System::GrabImage, Core/System.cc:92-138)
enum FrameGate : int {
    GATE_FUSE = 1,    // allocation and integration (Mapping::FuseColor)
    GATE_RENDER = 2,  // visible list, z-range tiles and raycast (UpdateVisibility / RayTrace)
    GATE_SWAP = 4,    // the frame becomes the last frame (Tracker::SwapFrame)
};

// Per-session tracking state, resident in device memory.  Poses are 4x4 row-major FP64
// camera-to-world transforms; the float views are what the kernels read.
struct TrackState {
    double last_pose[16];
    double next_pose[16];
    double init_pose[16];
    double delta[16];
    double A[36], b[6], x[6];
    float Rcurr[9], RcurrInv[9], tcurr[3];
    float Rlast[9], RlastInv[9], tlast[3];
    float residual[2];
    float icp_error;
    int count;
    int status;
    int iters_done;
    float sums[32];
    // photometric term of the last iteration (RGBStep): the weighted-row sums, the residual
    // pass's {count, sum diff^2} and sigma, and sqrt(sum r^2) / count with its count
    float rgb_sums[32];
    int rgb_corresp[2];
    float sigma;
    float rgb_error;
    int rgb_count;
    // the tracker's state machine (Tracking.cc:76-125, System.cc:92-138)
    int sm_state;     // Tracker::state of the next frame: 1 before the first, 0 tracking, -1 lost
    int sm_prev;      // the state of the previous frame (Tracker::lastState after its swap)
    int missed;       // Tracker::noMissedFrames
    int after_reloc;  // System::num_frames_after_reloc (10 at start)
    int reloc;        // a relocalisation pose is in last_pose for the next (lost) frame
    int gates;        // FrameGate bits of the current frame
    int mapping_off;  // Tracker::mappingDisabled: frames are tracked and rendered, never fused
};

// camera pose as the kernels use it: R (camera->world), Rinv = R^T, t
struct PoseF {
    m3 R, Rinv;
    f3 t;
};

// When `st` is set the pose comes from the tracking state and the kernel does nothing unless
// the frame renders (a tracked or first frame, or one relocalised from outside); fusion
// kernels check GATE_FUSE as well (frame_fuses).
__device__ __forceinline__ bool frame_fuses(const TrackState* st_) {
#if defined(__HIP_DEVICE_COMPILE__)
    const auto st = (const __attribute__((address_space(1))) TrackState*)st_;
#else
    const TrackState* st = st_;
#endif
    return (st->gates & GATE_FUSE) != 0;
}
__device__ __forceinline__ bool pose_from_state(const TrackState* st_, PoseF& p) {
    // global view: the state pointer may come from a descriptor in memory (scalar loads)
#if defined(__HIP_DEVICE_COMPILE__)
    const auto st = (const __attribute__((address_space(1))) TrackState*)st_;
#else
    const TrackState* st = st_;
#endif
    if (!(st->gates & GATE_RENDER)) return false;
    const auto r = st->Rcurr;
    const auto ri = st->RcurrInv;
    p.R = m3{mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), mk3(r[6], r[7], r[8])};
    p.Rinv = m3{mk3(ri[0], ri[1], ri[2]), mk3(ri[3], ri[4], ri[5]), mk3(ri[6], ri[7], ri[8])};
    p.t = mk3(st->tcurr[0], st->tcurr[1], st->tcurr[2]);
    return true;
}

struct FuseArgs {
    PoseF pose;
    const TrackState* st;
    float fx, fy, cx, cy, invfx, invfy;
    int cols, rows;
    const float* depth;      // raw depth (metres, NaN = invalid)
    const f4* nmap;          // level-0 normal map of the same frame
    const unsigned char* rgb;  // 3 bytes per pixel
    // when set, integration gathers these per-frame packed inputs instead: (depth, n.x, n.y,
    // n.z) and r | g << 8 | b << 16 per pixel, written by the pack pass from depth, nmap and
    // rgb (nmap's w is 1 wherever x is a number, so it is implied)
    f4* px;
    unsigned* prgb;
};

struct VisArgs {
    PoseF pose;
    const TrackState* st;
    float fx, fy, cx, cy;
    int cols, rows;
    float depth_min, depth_max;
};

struct ZRangeArgs {
    PoseF pose;
    const TrackState* st;
    float fx, fy, cx, cy;
    float depth_min, depth_max;
    unsigned* zmin;  // float bit patterns, zcols x zrows
    unsigned* zmax;
    int zcols, zrows;
    unsigned max_fragments;
    int n_visible_override;  // >= 0: use this count instead of the device counter
};

struct RaycastArgs {
    PoseF pose;
    const TrackState* st;
    float invfx, invfy, cx, cy;
    float step_scale;
    int cols, rows;
    const unsigned* zmin;
    const unsigned* zmax;
    int zcols, zrows;
    f4* vmap;
    f4* nmap;
    // when set, a frame that produced no rendering fragments keeps these maps (the frame's
    // own level-0 maps) instead of a raycast, as Mapping::RayTrace skips Raycast
    const f4* keep_v;
    const f4* keep_n;
};

struct MeshOut {
    float* vertex;  // 9 floats per triangle
    float* normal;
    unsigned char* color;  // 9 bytes per triangle
    int* tri_meta;         // optional: 4 ints per triangle (voxel x, y, z, cube*16 + t)
    unsigned max_tri;
};

// work-group sizes of the ICP reduction (one pixel per lane) and of its one-group final sum
// most work-groups of an ICP reduction (larger frames loop over pixel chunks per lane)
constexpr int kIcpMaxBlocks = 4096;
constexpr int kIcpThreads = 1024;
// pixels per lane and per work-group chunk of the ICP-only reduction, and of the combined
// ICP + photometric pass (two per lane)
constexpr int kIcpLanePixels = 2;
constexpr int kIcpPixels = kIcpLanePixels * kIcpThreads;
constexpr int kRgbPixels = 2 * kIcpThreads;
constexpr int kIcpFinalThreads = 256;

// Inputs of the photometric term at one level (RGBReduction, Reduction.cu:377-407)
// The pose-independent tests of FindCorresp (Reduction.cu:420-433: the 5-pixel margin, a
// non-zero 4x4 window, a non-zero gradient, a valid vertex) are run once per frame and level;
// the pixels that pass form an ordered list, and the per-iteration passes run over the list.
struct RgbArgs {
    const unsigned char* inext;  // intensity of this frame
    const unsigned char* ilast;  // intensity of the last tracked frame
    const short* dx;             // Sobel derivatives of this frame's intensity
    const short* dy;
    int* list;    // eligible pixels of the level in pixel order
    int* count;   // their number
    int* wgcnt;   // per list-building work-group: eligible pixels of its 1024
    int* rec;     // per list entry, written by the residual pass: (j << 9) | (diff + 256), or -1
    int* corr;    // 2 ints per residual-pass work-group: count, sum of diff^2 (32-bit, wrapping)
};

struct IcpArgs {
    const f4* vcurr;
    const f4* ncurr;
    const f4* vlast;
    const f4* nlast;
    int cols, rows;
    float fx, fy, cx, cy;
    TrackState* st;
    float* partials;  // 32 floats per work-group (64 with the photometric term)
    int mode;         // 0: sums only; 1: sums, solve and pose update
    int min_count;
    int final_iter;
    float thresh;
    int use_rgb;      // 1: RGBStep is blended into every iteration (ComputeSE3 icpOnly = false)
    RgbArgs rgb;
};

// frames with the photometric term are limited to 2^22 pixels (the residual record packs
// the last-frame pixel index in 22 bits)
constexpr int kRgbMaxPixels = 1 << 22;

// pyramid.hip
void launch_filter_depth(const uint16_t* depth, float* raw, float* filt, int cols, int rows, float scale_inv,
                         float cutoff, hipStream_t s);
void launch_pyrdown(const float* src, int scols, int srows, float* dst, int dcols, int drows, hipStream_t s);
void launch_vmap(const float* depth, f4* vmap, int cols, int rows, float invfx, float invfy, float cx, float cy,
                 float cutoff, hipStream_t s);
void launch_nmap(const f4* vmap, f4* nmap, int cols, int rows, hipStream_t s);
void launch_resize(const f4* vsrc, const f4* nsrc, int scols, f4* vdst, f4* ndst, int dcols, int drows,
                   hipStream_t s);
void launch_intensity(const unsigned char* rgb, unsigned char* img, int cols, int rows, hipStream_t s);
void launch_pyrdown_u8(const unsigned char* src, int scols, int srows, unsigned char* dst, int dcols, int drows,
                       hipStream_t s);
void launch_derivative(const unsigned char* img, short* dx, short* dy, int cols, int rows, hipStream_t s);
// copies `n` bytes of each of up to three images when the frame was tracked (or is the first):
// Tracker::SwapFrame of the intensity pyramid
void launch_image_commit(const TrackState* st, const unsigned char* const src[3], unsigned char* const dst[3],
                         const int n[3], hipStream_t s);
// map.hip
void launch_reset_map(const MapView& m, hipStream_t s);
void launch_begin_pass(const MapView& m, int which, hipStream_t s);
void launch_alloc(const MapView& m, const FuseArgs& a, hipStream_t s);
void launch_visible(const MapView& m, const VisArgs& v, hipStream_t s);
void launch_integrate(const MapView& m, const FuseArgs& a, unsigned grid, hipStream_t s);
void launch_fuse_pack(const FuseArgs& a, hipStream_t s);  // fills a.px / a.prgb
// raycast.hip
void launch_zrange(const MapView& m, const ZRangeArgs& a, hipStream_t s);
void launch_raycast(const MapView& m, const RaycastArgs& a, hipStream_t s);
// mesh.hip
void upload_mc_tables();
void launch_mesh(const MapView& m, const MeshOut& o, unsigned* counts, unsigned* chunk, hipStream_t s);
// icp.hip
void launch_icp(const IcpArgs& a, int nblocks, hipStream_t s);  // with a.use_rgb: RGBStep too
// the list of eligible pixels of the photometric term for up to three levels (a[l].use_rgb)
void launch_rgb_lists(const IcpArgs* a, const int* nblocks, int nlevels, hipStream_t s);
void launch_track_begin(TrackState* st, hipStream_t s);
void launch_track_commit(TrackState* st, hipStream_t s);

}  // namespace fu
