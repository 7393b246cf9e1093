/*
 * fusion_amd.h - C ABI of the MI355X (gfx950) dense RGB-D pipeline.
 *
 * Drop-in boundary for the CUDA path of cnxyang/Fusion: the free functions declared in
 * Mapping/RenderScene.h:3-52 and Tracking/Reduction.h:18-72, the forwarding methods of
 * Mapping (Mapping/Mapping.h:20-70) and the per-frame sequence of System::GrabImage
 * (Core/System.cc:92-138).  Each entry point names the reference function it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Image arguments are device pointers to dense row-major
 *    images (pitch = width): depth u16, depth/range f32, vertex and normal maps as 4 floats
 *    per pixel (x,y,z,w), colour 3 bytes per pixel.
 *  - Every call takes a stream (a hipStream_t; NULL = the default stream) and is
 *    asynchronous, except where a host-memory output is requested: such calls synchronise
 *    the stream before returning, as the reference wrappers do (cudaDeviceSynchronize).
 *  - Errors are returned as fu_status codes (never exit()); fu_last_error() gives the text.
 *  - Poses are camera-to-world.  The float pose of the reference kernels (Matrix3f Rview,
 *    Matrix3f RviewInv, float3 tview, Core/Frame.cc:275-294) is fu_pose_f; 4x4 poses are
 *    row-major doubles (Core/Frame.h:78).
 */
#ifndef FUSION_AMD_H
#define FUSION_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int fu_status;
#define FU_OK 0
#define FU_ERR_ARG (-1)      /* invalid argument or size */
#define FU_ERR_HIP (-2)      /* HIP runtime error (text in fu_last_error) */
#define FU_ERR_NOMEM (-3)    /* device allocation failed */
#define FU_ERR_STATE (-4)    /* object not in a state that allows the call */

typedef void* fu_stream;

typedef struct fu_map fu_map;
typedef struct fu_tracker fu_tracker;
typedef struct fu_system fu_system;

/* Voxel storage formats (per voxel: sdf, weight, rgb) */
#define FU_VOXEL_F32_W8 0  /* f32 sdf + u8 weight: the reference Voxel (DeviceMap.h:61-93), 8 B */
#define FU_VOXEL_F32_W16 1 /* f32 sdf + u16 weight, 9 B */
#define FU_VOXEL_F16_W8 2  /* f16 sdf + u8 weight, 6 B */

/* Map capacities and constants: the compile-time constants of DeviceMap (Mapping/DeviceMap.h:135-151)
 * as run-time parameters. */
typedef struct fu_map_params {
    uint32_t num_buckets;        /* NumBuckets (:140) */
    uint32_t num_excess;         /* NumExcess (:139) */
    uint32_t num_blocks;         /* NumSdfBlocks (:141): voxel blocks in the pool */
    float voxel_size;            /* VoxelSize (:145), metres */
    float trunc_dist;            /* TruncateDist (:146), metres */
    float depth_min, depth_max;  /* DepthMin/DepthMax (:137-138) */
    uint32_t voxel_format;       /* FU_VOXEL_* */
    uint32_t max_weight;         /* weight cap (FuseMap.cu:186 uses 255) */
    uint32_t max_render_blocks;  /* MaxRenderingBlocks (:147) */
    uint32_t max_new_blocks;     /* per-frame capacity of newly allocated blocks (no reference counterpart) */
} fu_map_params;

typedef struct fu_map_stats {
    uint32_t n_allocated;  /* blocks taken from the pool (heapCounter in the reference) */
    uint32_t n_excess;     /* excess slots used (entryPtr - 1) */
    uint32_t n_visible;    /* noVisibleBlocks of the last visibility pass */
    uint32_t n_new;        /* blocks allocated by the last fusion pass */
    uint32_t n_fragments;  /* rendering fragments of the last z-range pass (noRenderingBlocks) */
    uint32_t n_triangles;  /* triangles of the last mesh pass (noTriangles) */
    uint32_t flags;        /* FU_FLAG_*: capacity overflows since the last reset */
    uint32_t capacity;     /* num_blocks */
} fu_map_stats;

#define FU_FLAG_HEAP_FULL 1u     /* CreateEntry found the heap empty (DeviceMap.cu:14-20) */
#define FU_FLAG_EXCESS_FULL 2u   /* excess list full (DeviceMap.cu:54) */
#define FU_FLAG_NEWSET_FULL 4u   /* more new blocks in one frame than max_new_blocks */
#define FU_FLAG_COORD_RANGE 8u   /* a block coordinate outside [-2^20, 2^20) was skipped */
#define FU_FLAG_RENDER_FULL 16u  /* rendering fragment list full (RenderScene.cu:144-146) */
#define FU_FLAG_TRI_FULL 32u     /* triangle output full (MeshScene.cu:263,281) */

typedef struct fu_pose_f {
    float R[9];    /* rotation, row-major (Frame::GpuRotation) */
    float Rinv[9]; /* its inverse (Frame::GpuInvRotation) */
    float t[3];    /* translation (Frame::GpuTranslation) */
} fu_pose_f;

/* ---------------------------------------------------------------- library */
const char* fu_last_error(void);
const char* fu_version(void);
/* The marching-cubes tables the library uses, in the reference's form (Mapping/Constant.h):
 * edge[256], tri[256*16] (-1 padded), vert[256]. Host memory; no device needed. */
fu_status fu_mc_tables(int32_t* edge, int32_t* tri, int32_t* vert);
void fu_map_params_default(fu_map_params* p);
/* bytes of device memory a map with these parameters allocates */
size_t fu_map_bytes(const fu_map_params* p);

/* ---------------------------------------------------------------- map (Mapping) */
/* Mapping::Create (Mapping/Mapping.cc:11-51): allocate every map buffer and reset the map. */
fu_status fu_map_create(const fu_map_params* p, fu_map** out);
fu_status fu_map_destroy(fu_map* m);
fu_status fu_map_get_params(const fu_map* m, fu_map_params* out);
/* ResetMap (Mapping/RenderScene.h:3, FuseMap.cu:352-364) */
fu_status fu_reset_map(fu_map* m, fu_stream s);
/* counters; synchronises */
fu_status fu_map_stats_get(fu_map* m, fu_map_stats* out, fu_stream s);

/* FuseMapColor (RenderScene.h:47-52, FuseMap.cu:251-319): allocate the blocks the frame's
 * truncation band touches, build the visible-block list, integrate depth and colour.
 * depth: raw range image (metres, NaN invalid), nmap: level-0 normals of the same frame.
 * As in the reference the map's depth_min/depth_max are used and depth_max/depth_min
 * arguments are ignored (FuseMap.cu:288-289).  n_visible (host, optional): synchronise
 * and return the visible count. */
fu_status fu_fuse_map_color(fu_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols, int rows,
                            const fu_pose_f* pose, float fx, float fy, float cx, float cy, float depth_max,
                            float depth_min, uint32_t* n_visible, fu_stream s);
/* CheckBlockVisibility (RenderScene.h:42-45, FuseMap.cu:206-249) */
fu_status fu_check_block_visibility(fu_map* m, const fu_pose_f* pose, int cols, int rows, float fx, float fy, float cx,
                                    float cy, float depth_max, float depth_min, uint32_t* n_visible, fu_stream s);
/* CreateRenderingBlocks (RenderScene.h:21-27, RenderScene.cu:199-276): z-range tile maps
 * (zcols x zrows floats, tile = 8x8 pixels) from the visible list.  n_visible: number of
 * visible entries to use (the reference's argument); UINT32_MAX = the map's own count.
 * n_fragments/rendered (host, optional): synchronise, return noRenderingBlocks and
 * whether anything is to be rendered. */
fu_status fu_create_rendering_blocks(fu_map* m, float* zrange_min, float* zrange_max, int zcols, int zrows,
                                     float depth_max, float depth_min, const fu_pose_f* pose, uint32_t n_visible,
                                     float fx, float fy, float cx, float cy, uint32_t* n_fragments, int* rendered,
                                     fu_stream s);
/* Raycast (RenderScene.h:13-19, RenderScene.cu:447-489) */
fu_status fu_raycast(fu_map* m, float* vmap, float* nmap, int cols, int rows, const float* zrange_min,
                     const float* zrange_max, int zcols, int zrows, const fu_pose_f* pose, float invfx, float invfy,
                     float cx, float cy, fu_stream s);
/* MeshScene (RenderScene.h:29-40, MeshScene.cu:306-355) + Mapping::CreateModel: triangle
 * soup of the whole map.  vertex/normal: 9 floats per triangle, color: 9 bytes per
 * triangle, tri_meta (optional): 4 ints per triangle (voxel x, y, z, cube_index*16 + t).
 * Synchronises; n_triangles = min(total, max_triangles) as the reference returns. */
fu_status fu_mesh_scene(fu_map* m, float* vertex, float* normal, uint8_t* color, int32_t* tri_meta,
                        uint32_t max_triangles, uint32_t* n_triangles, uint32_t* n_occupied, fu_stream s);

/* Inspection (no reference counterpart beyond the raw map dump of System::WriteMapToDisk,
 * Core/System.cc:195-234).  Host outputs; synchronise.
 *  export_blocks: block coordinates (3 ints each) of the allocated blocks, pool order.
 *  export_visible: block coordinates of the visible list.
 *  read_blocks: the 512 voxels (index z*64+y*8+x) of each listed block; absent blocks
 *  read as sdf NaN, weight 0, colour 0. */
fu_status fu_map_export_blocks(fu_map* m, int32_t* xyz, uint32_t cap, uint32_t* n, fu_stream s);
fu_status fu_map_export_visible(fu_map* m, int32_t* xyz, uint32_t cap, uint32_t* n, fu_stream s);
fu_status fu_map_read_blocks(fu_map* m, const int32_t* xyz, uint32_t n, float* sdf, uint16_t* weight, uint8_t* rgb,
                             fu_stream s);

/* ---------------------------------------------------------------- frame (Reduction.h) */
/* FilterDepth (Reduction.h:18-20, Pyrdown.cu:43-54) */
fu_status fu_filter_depth(const uint16_t* depth, float* raw, float* filtered, int cols, int rows, float depth_scale,
                          float depth_cutoff, fu_stream s);
/* PyrDownGauss (Reduction.h:29, Pyrdown.cu:160-169); dst is (cols/2) x (rows/2) */
fu_status fu_pyrdown_gauss(const float* src, int cols, int rows, float* dst, fu_stream s);
/* ComputeVMap (Reduction.h:22-24, Pyrdown.cu:80-91) */
fu_status fu_compute_vmap(const float* depth, float* vmap, int cols, int rows, float fx, float fy, float cx, float cy,
                          float depth_cutoff, fu_stream s);
/* ComputeNMap (Reduction.h:26-27, Pyrdown.cu:116-125) */
fu_status fu_compute_nmap(const float* vmap, float* nmap, int cols, int rows, fu_stream s);
/* ResizeMap (Reduction.h:40-42, Pyrdown.cu:281-291); destination (cols/2) x (rows/2) */
fu_status fu_resize_map(const float* vsrc, const float* nsrc, int cols, int rows, float* vdst, float* ndst,
                        fu_stream s);
/* ICPStep (Reduction.h:67-72, Reduction.cu:319-375): one point-to-plane reduction.
 * residual[2] = {sum r^2, count}, A (6x6 row-major) and b (6) in FP64 on the host.
 * Synchronises. */
fu_status fu_icp_step(const float* next_vmap, const float* last_vmap, const float* next_nmap, const float* last_nmap,
                      int cols, int rows, const float Rcurr[9], const float tcurr[3], const float Rlast[9],
                      const float RlastInv[9], const float tlast[3], float fx, float fy, float cx, float cy,
                      float residual[2], double A[36], double b[6], fu_stream s);
/* fu_icp_step without allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results; its tracking state and partial rows live
 * in a fixed device array of the library (1 MiB + 1.5 KB, held by each GPU it runs on until
 * the process ends) instead of memory allocated on first use.  Frames up to 2^30 pixels
 * (FU_ERR_ARG beyond).  All calls on one GPU share that array: calls that overlap on one GPU
 * race and must be ordered by the caller.  Synchronises. */
fu_status fu_icp_step_noalloc(const float* next_vmap, const float* last_vmap, const float* next_nmap,
                              const float* last_nmap, int cols, int rows, const float Rcurr[9], const float tcurr[3],
                              const float Rlast[9], const float RlastInv[9], const float tlast[3], float fx,
                              float fy, float cx, float cy, float residual[2], double A[36], double b[6],
                              fu_stream s);

/* ---------------------------------------------------------------- photometric term */
/* ImageToIntensity (Reduction.h:34-35, Pyrdown.cu:196-206): (int)(0.2125 c1 + 0.7154 c0 +
 * 0.0721 c2) in double per pixel of a 3-byte image */
fu_status fu_image_to_intensity(const uint8_t* rgb, uint8_t* image, int cols, int rows, fu_stream s);
/* PyrDownGauss on bytes (Reduction.h:31-32, Pyrdown.cu:171-181); dst is (cols/2) x (rows/2) */
fu_status fu_pyrdown_gauss_u8(const uint8_t* src, int cols, int rows, uint8_t* dst, fu_stream s);
/* ComputeDerivativeImage (Reduction.h:37-38, Pyrdown.cu:238-249): 3x3 Sobel as shorts */
fu_status fu_compute_derivative_image(const uint8_t* image, int16_t* dIdx, int16_t* dIdy, int cols, int rows,
                                      fu_stream s);
/* RGBStep (Reduction.h:74-84, Reduction.cu:537-627): one photometric reduction.
 * residual[2] = {sum of weighted squared differences, count}, A (6x6 row-major) and b (6) in
 * FP64 on the host; corresp[2] (may be NULL) = {count, sum of squared differences} of the
 * residual pass, from which sigma = sqrt(sum / count).  At most 2^22 pixels.  Synchronises. */
fu_status fu_rgb_step(const uint8_t* next_image, const uint8_t* last_image, const float* next_vmap,
                      const float* last_vmap, const int16_t* dIdx, const int16_t* dIdy, int cols, int rows,
                      const float Rcurr[9], const float RcurrInv[9], const float Rlast[9], const float RlastInv[9],
                      const float tcurr[3], const float tlast[3], float fx, float fy, float cx, float cy,
                      float residual[2], double A[36], double b[6], int32_t corresp[2], fu_stream s);

/* ---------------------------------------------------------------- tracker (Tracker::ComputeSE3) */
typedef struct fu_track_result {
    double pose[16];       /* NextFrame->pose after tracking (rolled back on failure) */
    double last_pose[16];  /* LastFrame->pose */
    double x[6];           /* last solved increment (upsilon, omega) */
    double A[36], b[6];    /* last normal equations */
    float icp_error;       /* sqrt(sum r^2) / count of the last iteration */
    int32_t count;         /* correspondences of the last iteration */
    int32_t status;        /* FU_TRACK_* */
    int32_t iterations;    /* iterations run */
    float rgb_error;       /* photometric term of the last iteration: sqrt(sum r^2) / count */
    int32_t rgb_count;
} fu_track_result;

#define FU_TRACK_RUNNING 0
#define FU_TRACK_OK 1       /* accepted (Tracking.cc:402-403) */
#define FU_TRACK_FAILED 2   /* NaN error or too few correspondences: pose rolled back (:371-375) */
#define FU_TRACK_REJECTED 3 /* final error or motion test failed: pose rolled back (:404-407) */
#define FU_TRACK_FIRST 4    /* first frame: identity pose (Tracking.cc:66-71) */
/* a system's tracker is lost after more than 9 frames in a row that were not tracked
 * (Tracking.cc:92-96); it then tracks nothing until it is relocalised: by graph matching
 * against a key map (fu_system_relocalise_graph / fu_system_validate_poses, on the lost frame
 * itself, as the reference's Relocalise), by a pose given for the next frame
 * (fu_system_relocalise, e.g. from fu_relocalise_keys), or by a reset */
#define FU_TRACK_LOST 5     /* lost: not tracked, nothing fused or raycast */
#define FU_TRACK_RELOC 6    /* relocalised (fu_system_relocalise's pose, or fu_system_validate_poses): raycast there */

/* Tracker state and ICP scratch (Tracker::Tracker, Tracking.cc:19-43) */
fu_status fu_tracker_create(int cols, int rows, fu_tracker** out);
fu_status fu_tracker_destroy(fu_tracker* t);
/* set LastFrame->pose (4x4 row-major) */
fu_status fu_tracker_set_last_pose(fu_tracker* t, const double pose[16], fu_stream s);
/* Tracker::ComputeSE3(icpOnly=true, iter, thresh) (Tracking.cc:296-409) run entirely on the
 * device: per iteration one reduction kernel whose last work-group solves the 6x6 system
 * and composes the pose in FP64.  Maps: 3 levels each, level l is (cols>>l) x (rows>>l);
 * K = {fx, fy, cx, cy} of level 0 (level l uses K/2^l). */
fu_status fu_compute_se3(fu_tracker* t, const float* const next_vmap[3], const float* const next_nmap[3],
                         const float* const last_vmap[3], const float* const last_nmap[3], const float K[4],
                         const int32_t iters[3], const int32_t min_count[3], float thresh, fu_stream s);
/* Tracker::ComputeSE3(icpOnly=false, iter, thresh) (Tracking.cc:296-409), the reference's
 * default tracking: every iteration runs RGBStep (Reduction.cu:537-627) on the intensity
 * pyramids (next/last, level l is (cols>>l) x (rows>>l) bytes) and the next frame's Sobel
 * derivatives, fails the frame when its residual count is below min_count or its error is
 * NaN (:341-346), and solves matA_rgb * 1e-4f + matA_icp (:379-383).  At most 2^22 pixels. */
fu_status fu_compute_se3_rgbd(fu_tracker* t, const float* const next_vmap[3], const float* const next_nmap[3],
                              const float* const last_vmap[3], const float* const last_nmap[3],
                              const uint8_t* const next_image[3], const uint8_t* const last_image[3],
                              const int16_t* const dIdx[3], const int16_t* const dIdy[3], const float K[4],
                              const int32_t iters[3], const int32_t min_count[3], float thresh, fu_stream s);
/* the result of the last compute_se3; synchronises */
fu_status fu_tracker_result(fu_tracker* t, fu_track_result* out, fu_stream s);

/* ---------------------------------------------------------------- per-frame pipeline (System) */
typedef struct fu_system_params {
    int32_t cols, rows;
    float fx, fy, cx, cy;
    float depth_scale;   /* Frame::mDepthScale: depth units per metre */
    float depth_cutoff;  /* Frame::mDepthCutoff, metres */
    int32_t iters[3];    /* ICP iterations per level (Tracking.h:100) */
    int32_t min_count[3]; /* MIN_ICP_COUNT (Tracking.h:102) */
    float icp_thresh;    /* THRESH_ICP_SE3 (Tracking.h:103) */
    int32_t plain;       /* 1: reference-style control (host solve, one synchronisation per
                            stage); 0: device-resident (on-device solve, no host round trip) */
    int32_t use_graph;   /* device-resident mode only: replay the frame as one hipGraph */
    int32_t use_rgb;     /* 1: track with the photometric term blended into ICP, as the
                            reference's default TrackFrame does (ComputeSE3(icpOnly=false),
                            Tracking.cc:162, 379-383); 0: ICP only (Tracking.cc:642-682) */
    fu_map_params map;
} fu_system_params;

typedef struct fu_frame_result {
    double pose[16];        /* camera-to-world pose of the frame (after tracking) */
    int32_t track_status;   /* FU_TRACK_* */
    int32_t fused;          /* 1 if the frame was fused and raycast */
    float icp_error;
    int32_t icp_count;
    uint32_t n_visible;
    uint32_t n_allocated;
    uint32_t n_fragments;
    uint32_t flags;
    uint32_t frame_index;
    int32_t rgb_count;      /* photometric residuals of the last iteration (use_rgb) */
} fu_frame_result;

/* maps owned by a system (device pointers) */
typedef struct fu_system_buffers {
    uint16_t* depth_in;  /* input slot: cols*rows u16 */
    uint8_t* rgb_in;     /* input slot: cols*rows*3 u8 */
    float* raw;          /* range image (metres) */
    float* depth[3];     /* filtered depth pyramid */
    float* vmap[3];      /* current frame maps */
    float* nmap[3];
    float* model_vmap[3]; /* raycast maps of the model (LastFrame after tracking) */
    float* model_nmap[3];
    float* zrange_min;
    float* zrange_max;
    int32_t zcols, zrows;
    /* photometric term (use_rgb only, else null): this frame's intensity pyramid and its
     * Sobel derivatives, and the pyramid of the last tracked frame (LastFrame->image) */
    uint8_t* image[3];
    uint8_t* last_image[3];
    int16_t* dIdx[3];
    int16_t* dIdy[3];
} fu_system_buffers;

void fu_system_params_default(fu_system_params* p);
fu_status fu_system_create(const fu_system_params* p, fu_system** out);
fu_status fu_system_destroy(fu_system* sys);
/* reset map and tracking (System request "reboot" -> Mapping::Reset, Tracker::ResetTracking) */
fu_status fu_system_reset(fu_system* sys, fu_stream s);
fu_status fu_system_get_buffers(fu_system* sys, fu_system_buffers* out);
fu_map* fu_system_map(fu_system* sys);
/* System::GrabImage (Core/System.cc:92-138) in ICP-only mode on the frame in the input slot
 * (copy the frame into depth_in/rgb_in first, on the same stream): pyramid, tracking
 * against the model raycast, fusion, raycast. Asynchronous. */
fu_status fu_system_grab_image(fu_system* sys, fu_stream s);
/* the same on caller-provided device images.  With use_graph, frames after the first run as
 * one hipGraph, captured once per system against the input slots: the images are copied into
 * depth_in/rgb_in on `s` (unless they are the slots) and the graph is replayed. */
fu_status fu_system_grab_image_from(fu_system* sys, const uint16_t* depth, const uint8_t* rgb, fu_stream s);
/* the same on host images, uploaded into the input slots on `s` as Frame::Frame does
 * (Core/Frame.cc:61-62); asynchronous for pinned (hipHostMalloc) memory, which must then stay
 * unchanged until the copy has run */
fu_status fu_system_grab_image_host(fu_system* sys, const uint16_t* depth, const uint8_t* rgb, fu_stream s);
/* Tracker::mappingDisabled (Tracking.h:45): with mapping off, frames are tracked, rendered and
 * raycast but never fused (System.cc:108-126), and fu_system_result reports fused = 0.
 * ReadMapFromDisk (System.cc:266-270) turns it off after a map load and restarts tracking;
 * here: fu_system_reset, fu_map_load on fu_system_map, fu_system_set_mapping(sys, 0).  The
 * setting survives fu_system_reset (as RebootSystem, System.cc:272-277) and is part of a
 * session snapshot.  Synchronises. */
fu_status fu_system_set_mapping(fu_system* sys, int enabled, fu_stream s);
/* graphs captured and replayed by this system (no reference counterpart) */
fu_status fu_system_graph_counts(fu_system* sys, uint32_t* captures, uint32_t* launches);
/* outcome of the last frame; synchronises */
fu_status fu_system_result(fu_system* sys, fu_frame_result* out, fu_stream s);
/* A lost system's relocalisation with a pose the caller found (for example with
 * fu_relocalise_keys, the reference's default Relocalise, Tracking.cc:411-467, from the lost
 * frame's keypoints; fu_system_relocalise_graph relocalises the lost frame itself).  The next
 * frame takes `pose` (camera to world, 4x4 row-major) as its pose:
 * status FU_TRACK_RELOC, the model raycast there, nothing fused; tracking resumes with the
 * frame after it, and fusion once 10 frames have passed (System.cc:97-111).  FU_ERR_STATE
 * unless the system is lost.  Synchronises. */
fu_status fu_system_relocalise(fu_system* sys, const double pose[16], fu_stream s);

/* ---------------------------------------------------------------- keypoint map */
/* SURF record of the keypoint map, in the reference's layout (Mapping/DeviceMap.h:99-108:
 * bool valid, float3 pos, float4 normal, float descriptor[64]; 288 bytes) */
typedef struct fu_surf {
    uint8_t valid;
    uint8_t pad_[3];
    float pos[3];
    float normal[4];
    float descriptor[64];
} fu_surf;
typedef struct fu_keymap fu_keymap;
/* KeyMap's constants (DeviceMap.h:163-170) as parameters: grid cell 0.01 m, 100000 buckets
 * of 5 slots */
typedef struct fu_keymap_params {
    float grid_size;
    uint32_t max_keys;
    uint32_t n_buckets;
} fu_keymap_params;
void fu_keymap_params_default(fu_keymap_params* p);
fu_status fu_keymap_create(const fu_keymap_params* p, fu_keymap** out); /* reset on creation */
fu_status fu_keymap_destroy(fu_keymap* k);
/* ResetKeyPoints (RenderScene.h:5, FuseMap.cu:366-380): every slot invalid.  Asynchronous. */
fu_status fu_reset_key_points(fu_keymap* k, fu_stream s);
/* InsertKeyPoints (RenderScene.h:7-8, FuseMap.cu:455-472; KeyMap::InsertKey,
 * DeviceMap.cu:258-287) of n device records: a key whose 1 cm cell already has a valid key in
 * its bucket takes that key's position and slot (index[i]); otherwise the bucket's first free
 * slot takes the key; a key meeting a full bucket is left out with index[i] unchanged.
 * Applied as one key at a time in index order (the reference's try-lock may drop a key that
 * meets another in its bucket in the same launch).  At most 2^24 keys.  Synchronises. */
fu_status fu_insert_key_points(fu_keymap* k, fu_surf* keys, int32_t* index, uint32_t n, fu_stream s);
/* CollectKeyPoints (RenderScene.h:10-11, FuseMap.cu:382-453): *n_keys = the number of valid
 * keys; valid key j (in slot order) is copied to out[j] for 0 < j < cap (the reference copies
 * only at offsets > 0, so out[0] is left as it was).  Synchronises. */
fu_status fu_collect_key_points(fu_keymap* k, fu_surf* out, uint32_t cap, uint32_t* n_keys, fu_stream s);
/* every slot of the map into host memory (max_keys * n_buckets records); synchronises */
fu_status fu_keymap_read(fu_keymap* k, fu_surf* host, fu_stream s);
/* BuildAdjacencyMatrix (Reduction.h:85-88, KeyPoints.cu:12-77): the n x n consistency scores
 * of n matched pairs (frame key i <-> map key i, descriptor distance dist[i]) into a dense
 * row-major float matrix (the reference's unused column sums are not computed).
 * Asynchronous. */
fu_status fu_build_adjacency_matrix(const fu_surf* frame_keys, const fu_surf* map_keys, const float* dist, int32_t n,
                                    float* matrix, fu_stream s);

/* ---------------------------------------------------------------- relocalisation
 * The detector-free parts of keypoint relocalisation.  The reference extracts SURF keypoints
 * with OpenCV (Frame::ExtractKeyPoints, Core/Frame.cc:168-212), which this library does not
 * have: the caller brings each keypoint's pixel position and 64-float descriptor. */
/* the keypoint test and back-projection of Frame::ExtractKeyPoints (Frame.cc:184-200, with
 * InterpDepth / InterpNormal, :93-165, and their bounds test as written: y against the width,
 * x against the height): for keypoint i at xy[2i], xy[2i+1] (device), the bilinear depth of
 * the level-0 filtered depth and normal of the level-0 normal map; ok[i] = 1 with pos (x, y, z
 * camera coordinates, 3 floats) and normal (4 floats) where both pass, else 0 and NaNs.
 * Asynchronous. */
fu_status fu_lift_keypoints(const float* xy, int32_t n, const float* depth, const float* nmap, int32_t cols,
                            int32_t rows, float fx, float fy, float cx, float cy, float depth_cutoff, float* pos,
                            float* normal, uint8_t* ok, fu_stream s);
/* the two nearest train descriptors of each query descriptor (knnMatch with k = 2 of
 * cv::cuda::BFMatcher, NORM_L2; Tracking.cc:436): idx[2i], idx[2i+1] and their distances
 * (sqrt of the sum of squared differences) in dist, nearest first; ties go to the lower train
 * index; -1 and +inf where there are fewer than two.  64-float descriptors, row-major, device
 * memory.  Asynchronous. */
fu_status fu_match_knn2(const float* query, int32_t nq, const float* train, int32_t nt, int32_t* idx, float* dist,
                        fu_stream s);
/* Solver::PoseEstimate (Optimization/Solver.cc:6-183) on the host: RANSAC over samples of
 * three pairs, Kabsch on each (rejecting reflections), the best model refitted on its
 * inliers (0.05 m), stopping after 20 iterations once the confidence reaches 0.95; then the
 * final refit on the inlier set as the reference forms it.  T (row-major 4x4) maps ref
 * points to src points; inlier[i] (optional) marks the final set; *ok = 0 when n < 3, no
 * sample was usable or the final rotation is a reflection.  The samples come from a
 * generator seeded by `seed` (the reference seeds rand() from the clock). */
fu_status fu_pose_estimate(const double* src, const double* ref, int32_t n, int32_t iterations, uint64_t seed,
                           double T[16], uint8_t* inlier, int32_t* ok);
/* Tracker::Relocalise with graph matching off, its default (Tracking.cc:411-467): the map's
 * keys (CollectKeyPoints, whose first record is never written and reads as zeros), the two
 * nearest map descriptors of each frame descriptor (device, nq x 64), the ratio test
 * d0 < 0.9 d1, then fu_pose_estimate over the pairs (frame_pos: host, nq x 3, the lifted
 * keypoints) with `iterations` (the reference's maxIterReloc is 100).  On success *ok = 1 and
 * pose is the frame's camera-to-world pose, to hand to fu_system_relocalise.  Synchronises. */
fu_status fu_relocalise_keys(fu_keymap* km, const float* frame_desc, const double* frame_pos, int32_t nq,
                             int32_t iterations, uint64_t seed, double pose[16], int32_t* n_pairs, int32_t* ok,
                             fu_stream s);

/* Graph matching (Tracker::GenerateGraph, Tracking.cc:469-629; the reference switches it on
 * from its viewer).  fu_select_subgraphs is its choice of mutually consistent pairs over the
 * n x n consistency matrix of fu_build_adjacency_matrix (device, row-major): rank_sum[x] =
 * the sum of column x (cv::cuda::reduce over rows; here in double in row order, rounded to
 * float), order = the pairs by rank, largest first (cv::sortIdx descending; ties to the lower
 * index, NaN ranks last); for each list i < n_lists, the i-th ranked pair heads it and the
 * pairs ranked after it join while matrix[head][pair] > min_score, up to max_select pairs
 * (head included; n_selected[i]); a list of three or more is refined: its k-th pair (k >= 1;
 * the head is not kept, as written) stays unless a later pair of the list has a score below
 * 5e-3 with it either way and a higher score with the head.  n_refined[i] pairs in
 * refined[i * max_select ...], in list order.  The reference's constants: 10 lists, 400
 * pairs, 0.1 (Tracking.h:129-132).  n <= 46340, max_select <= 1024; all buffers device
 * memory (rank_sum and order: n each).  Asynchronous. */
fu_status fu_select_subgraphs(const float* matrix, int32_t n, int32_t n_lists, int32_t max_select, float min_score,
                              float* rank_sum, int32_t* order, int32_t* n_selected, int32_t* n_refined,
                              int32_t* refined, fu_stream s);
/* Tracker::Relocalise with graph matching on, through GenerateGraph's PoseEstimate calls
 * (Tracking.cc:432-453, 469-629): the match list keeps both nearest map keys of a query that
 * fails the ratio test, pairs whose map key is not valid are dropped, the consistency matrix
 * and sub-graphs are built on the device, and PoseEstimate runs on the refined pairs of the
 * first list of three or more selected pairs, once for each such list up to five (the
 * reference orders the lists with cv::sortIdx on a one-column matrix, which leaves every
 * index at 0, so each call takes the first list), when it holds four or more pairs.  Call c
 * draws from seed + c.  poses: up to 5 camera-to-world candidates (row-major 4x4) in
 * *n_poses; *n_pairs the match list's size.  frame_desc: device, nq x 64; frame_pos (nq x 3)
 * and frame_normal (nq x 4): host floats, the lifted keypoints.  Synchronises. */
fu_status fu_relocalise_graph_candidates(fu_keymap* km, const float* frame_desc, const float* frame_pos,
                                         const float* frame_normal, int32_t nq, int32_t iterations, uint64_t seed,
                                         double poses[5 * 16], int32_t* n_poses, int32_t* n_pairs, fu_stream s);
/* Tracker::ValidatePose (Tracking.cc:636-682) on a lost system's current frame (the frame
 * last given to it, which came back FU_TRACK_LOST): for each candidate pose, the visible
 * blocks there (skipped below 512: accepted[i] = -1), the model raycast there and ComputeSE3
 * (ICP only) from it with iters_reloc and thresh_reloc (the reference's {3, 3, 2} and 1e-3;
 * accepted[i] = 1 if it passes, else 0); the first that passes (*chosen) is rendered again and
 * refined with the system's tracking iterations and threshold.  If that passes, *ok = 1 and
 * the frame is relocalised: its pose the refined one (fu_system_result: FU_TRACK_RELOC), the
 * model raycast there, nothing fused; tracking resumes with the next frame and fusion once 10
 * frames have passed, as after fu_system_relocalise.  A failed attempt leaves the model maps
 * rendered at the last candidate tried (the reference's LastFrame maps, likewise) and the
 * system lost.  FU_ERR_STATE unless the system is lost and has been given a frame since it
 * was reset or loaded.  Synchronises. */
fu_status fu_system_validate_poses(fu_system* sys, const double* poses, int32_t n, const int32_t iters_reloc[3],
                                   float thresh_reloc, int32_t* accepted, int32_t* chosen, int32_t* ok, fu_stream s);
/* Tracker::Relocalise with graph matching on, whole: fu_relocalise_graph_candidates, then
 * fu_system_validate_poses with {3, 3, 2} iterations and 1e-3 (Tracking.h:101-104).
 * *n_candidates (optional): the poses PoseEstimate gave.  Synchronises. */
fu_status fu_system_relocalise_graph(fu_system* sys, fu_keymap* km, const float* frame_desc, const float* frame_pos,
                                     const float* frame_normal, int32_t nq, int32_t iterations, uint64_t seed,
                                     int32_t* n_candidates, int32_t* ok, fu_stream s);

/* ---------------------------------------------------------------- self-test */
/* The integration and ICP kernels divide and take square roots with range-specialised forms
 * of the compiler's IEEE sequences (no reference counterpart).  This runs n random cases of
 * each range on the device and counts the results that differ from the IEEE operation:
 * mismatches[0] quotients by a 4-D norm, [1] sdf / truncation, [2] projections (their
 * truncated pixel coordinate), [3] square roots, [4] the ICP projection's reciprocal 1 / z
 * against the reference's double quotient rounded to float, [5] the
 * single-instruction float and double -> int conversions against their rule (truncation,
 * saturation, NaN -> 0), [6] the bilateral weights' exp against the general one.  All seven
 * must be 0.  Synchronises. */
fu_status fu_selftest_arith(uint64_t seed, uint32_t n, uint32_t mismatches[7], fu_stream s);

/* ---------------------------------------------------------------- on-disk formats */
/* Map snapshot: System::WriteMapToDisk / ReadMapFromDisk (Core/System.cc:195-271).  The
 * reference dumps the whole device map raw; the snapshot holds the used hash entries (at
 * their indices), the allocated blocks' keys and voxels and the map counters, in a versioned
 * little-endian file with an FNV-1a checksum.  Loading needs a map of the same parameters
 * (buckets, excess, blocks, voxel size/format, truncation, depth range, weight cap) and
 * rebuilds the same entries and pool slots, so the map continues exactly as before.  The
 * whole file is read and checked on the host before the device map is touched: the checksum,
 * then the structure the kernels rely on (entry keys in their buckets, excess links in range,
 * chains acyclic with each excess entry on one chain, block indices below the allocated count
 * and agreeing with the block keys, no key allocated twice).  A file of other parameters, or
 * one that fails a check, is refused with the map unchanged.  Both synchronise `s`. */
fu_status fu_map_save(fu_map* m, const char* path, fu_stream s);
fu_status fu_map_load(fu_map* m, const char* path, fu_stream s);
/* the checks of fu_map_load on the host alone, against a map of parameters `p` (no device
 * access; no reference counterpart) */
fu_status fu_map_check_file(const fu_map_params* p, const char* path);
/* Session snapshot (no reference counterpart; eviction and resume of a served session): the
 * map snapshot plus the tracking state, the frame count, the level-0 model maps the next frame
 * is tracked against and, with use_rgb, the last frame's intensity pyramid.  Loading needs a
 * system of the same frame size, intrinsics, depth scale/cut-off, tracking mode and map
 * parameters; the next frame then tracks and fuses exactly as it would have without the
 * round trip.  The session part is read into host memory and the map part checked as by
 * fu_map_load before anything is written, so a refused file leaves the session unchanged. */
fu_status fu_system_save(fu_system* sys, const char* path, fu_stream s);
fu_status fu_system_load(fu_system* sys, const char* path, fu_stream s);
/* System::WriteMeshToDisk (Core/System.cc:140-193): a triangle soup of 3 vertices per
 * triangle (x y z nx ny nz red green blue) and faces "3 i i+1 i+2", as the output of
 * fu_mesh_scene (device or host arrays).  binary = 0: the reference's ASCII text (floats as
 * std::ostream prints them, i.e. %g); 1: the same elements as binary_little_endian. */
fu_status fu_mesh_write_ply(const char* path, const float* vertex, const float* normal, const uint8_t* color,
                            uint32_t n_triangles, int binary, fu_stream s);

/* Per-stage timing with HIP events on the frame's stream (the reference defines a StopWatch,
 * Utility/StopWatch.h:8-28, but never calls it).  While enabled, every grab launches its
 * kernels directly (no graph), records an event at each stage boundary, synchronises at the
 * end of the frame and accumulates the stage times and the frame's map counters; the first
 * frame of a sequence is not accumulated. */
#define FU_STAGE_RESIZE 0    /* ResizeMap of the model maps */
#define FU_STAGE_PYRAMID 1   /* FilterDepth, PyrDownGauss, ComputeVMap/NMap */
#define FU_STAGE_ICP 2       /* all ICP iterations */
#define FU_STAGE_ALLOC 3     /* block allocation (both passes) */
#define FU_STAGE_VISIBLE 4   /* visible-block list */
#define FU_STAGE_INTEGRATE 5 /* TSDF integration */
#define FU_STAGE_ZRANGE 6    /* z-range tiles */
#define FU_STAGE_RAYCAST 7   /* ray march */
#define FU_NUM_STAGES 8
typedef struct fu_stage_times {
    double ms[FU_NUM_STAGES]; /* summed milliseconds per stage */
    uint32_t frames;          /* frames accumulated */
    double n_visible;         /* summed visible blocks */
    double n_new;             /* summed newly allocated blocks */
    double icp_iterations;    /* summed ICP iterations run */
    double icp_pixels;        /* summed pixel-iterations of ICP */
    double steps;             /* profiled launches of the stage sequence (frames, or batched steps) */
} fu_stage_times;
fu_status fu_system_set_profiling(fu_system* sys, int enable);
fu_status fu_system_get_stage_times(fu_system* sys, fu_stage_times* out);

/* ---- batched ICP of many frame pairs ------------------------------------------------------
 * Tracker::ComputeSE3 (Tracking/Tracking.cc:296-409) for n trackers at once: levels coarse to
 * fine, every iteration one reduction launch and one solve launch covering up to
 * FU_MAX_BATCH trackers (larger n runs in groups).  Each tracker's pyramids have its own size
 * (the tracker's cols x rows); K holds fx, fy, cx, cy of each pair's level 0.  Results are
 * those of fu_compute_se3 on each tracker alone (fu_tracker_result); asynchronous on `s`.
 * No reference counterpart for the batching itself. */
typedef struct fu_icp_batch fu_icp_batch;
fu_status fu_icp_batch_create(int capacity, fu_icp_batch** out);
fu_status fu_icp_batch_destroy(fu_icp_batch* b);
fu_status fu_compute_se3_batch(fu_icp_batch* b, int n, fu_tracker* const* trackers,
                               const float* const* next_vmap, const float* const* next_nmap,
                               const float* const* last_vmap, const float* const* last_nmap, const float* K,
                               const int32_t iters[3], const int32_t min_count[3], float thresh, fu_stream s);

/* ---- batched frames of many systems -------------------------------------------------------
 * A batch context runs one frame of each of up to FU_MAX_BATCH attached systems as one
 * sequence of launches (each stage of System::GrabImage, Core/System.cc:92-138, launched once
 * for all of them).  Systems keep their own map, pose, pyramid and model maps; each one's
 * results are bit-identical to running the same frames through fu_system_grab_image_from.
 * Systems of one context share a voxel format and may differ in frame size and intrinsics.
 * No reference counterpart: the reference processes one frame of one system at a time. */
#define FU_MAX_BATCH 64
typedef struct fu_batch fu_batch;
fu_status fu_batch_create(int capacity, fu_batch** out);  /* capacity = attachable systems */
fu_status fu_batch_destroy(fu_batch* b);                  /* detaches every system */
fu_status fu_batch_attach(fu_batch* b, fu_system* sys);
fu_status fu_batch_detach(fu_batch* b, fu_system* sys);
/* n distinct attached systems, one device-resident frame each (depth u16, RGB u8x3 of the
 * system's size); asynchronous on `s` */
fu_status fu_batch_grab(fu_batch* b, int n, fu_system* const* sys, const uint16_t* const* depth,
                        const uint8_t* const* rgb, fu_stream s);
/* per-stage HIP-event times of batched steps, as fu_system_set_profiling; a profiled step
 * synchronises at its end.  ms[] are per step, summed over steps; the counts sum over frames. */
fu_status fu_batch_set_profiling(fu_batch* b, int enable);
fu_status fu_batch_get_stage_times(fu_batch* b, fu_stage_times* out);
/* Captured steps (a batch's per-frame work as a hipGraph): with `enable`, a step whose
 * systems all have tracked a frame before runs its input stage (resize, filter, pyramid,
 * normals: the launches that read the frames) directly and everything after it -- tracking,
 * fusion, raycast, results -- as one hipGraph launch.  A graph is keyed by the step's systems
 * in slot order and the result buffer; it is captured the second time its key is met (so
 * one-off memberships run as direct launches) and replayed afterwards; the context keeps the
 * 32 most recently used.  Attaching or detaching a system, fu_set_tuning and disabling drop
 * the graphs.  Results are bit-identical to direct launches.  Profiled steps are never
 * captured. */
fu_status fu_batch_set_graph(fu_batch* b, int enable);
fu_status fu_batch_graph_counts(fu_batch* b, uint32_t* captures, uint32_t* replays);

/* ---- launch geometry of the batched stages ------------------------------------------------
 * Work-group counts, the pixel layout of a raycast wave and the placement of work-groups on
 * the XCDs for batched steps (fu_batch_grab and the server; the wave layout also for the raycast
 * of one system, where a captured frame graph keeps the layout of its capture).  The defaults
 * are FU_WAVE_8X8, no XCD ordering, 16384 / 8192 / 2048 work-groups.  None of them changes a result:
 * per-pixel and per-block work does not depend on them, and the reductions keep their fixed
 * decomposition (XCD placement only permutes which hardware work-group runs which logical
 * one).  Process-wide; applies to steps launched afterwards.  fusion_amd/autotune.py measures
 * candidates on a workload and caches the fastest per device.  No reference counterpart. */
#define FU_WAVE_16X4 0 /* a 64-lane wave covers 16x4 pixels of its 16x16 tile */
#define FU_WAVE_8X8 1  /* 8x8 pixels */
#define FU_WAVE_4X16 2 /* 4x16 pixels */
typedef struct fu_tuning {
    int32_t raycast_wave;     /* FU_WAVE_*: lane layout of the ray-march waves */
    int32_t xcd_order;        /* bit (1 << FU_STAGE_x): consecutive work-groups of the stage run on one XCD
                                 (applies to ICP, integrate, raycast) */
    int32_t integrate_groups; /* work-groups of one integration launch, split over its sessions, 256..262144 */
    int32_t visible_groups;   /* the same for the visible-block list, 256..262144 */
    int32_t alloc_groups;     /* the same for the allocation's second pass, 256..262144 */
} fu_tuning;
void fu_tuning_default(fu_tuning* t);
fu_status fu_set_tuning(const fu_tuning* t); /* FU_ERR_ARG for a value out of range */
fu_status fu_get_tuning(fu_tuning* t);

/* ---- serving many reconstruction sessions -------------------------------------------------
 * A server keeps every session's map, pose, pyramid and last raycast resident in HBM and
 * processes frames as they arrive (continuous batching): sessions are spread over `lanes`
 * HIP streams; on each lane, whenever fewer than `max_inflight` steps are queued, the oldest
 * pending frame of each session (up to `max_batch` sessions, oldest first) forms the next
 * batched step, so frames arriving while others run join the next step and the lanes' steps
 * overlap on the device.  Sessions are admitted while the sum of their map sizes
 * (fu_map_bytes) stays within `block_budget`.  Times are host-clock milliseconds since the
 * server was created.  Not thread-safe: one host thread drives a server.
 * No reference counterpart: the reference drives one System from its main loop
 * (Core/System.cc:92-138). */
#define FU_ERR_BUDGET (-5) /* session not admitted: voxel-block budget exhausted */
typedef struct fu_server fu_server;
typedef struct {
    int32_t max_batch;     /* sessions per step, 1..FU_MAX_BATCH */
    int32_t max_inflight;  /* steps queued per lane, >= 1 */
    int32_t lanes;         /* HIP streams, 1..8 */
    uint64_t block_budget; /* bytes of session maps admitted at once; 0 = no limit */
} fu_server_params;
typedef struct {
    uint64_t tag;          /* caller's tag of the frame */
    int32_t session;
    int32_t track_status;  /* FU_TRACK_* */
    double pose[16];       /* camera-to-world after the frame, row-major */
    float icp_error;
    uint32_t n_visible;
    uint32_t n_allocated;
    uint32_t flags;        /* FU_FLAG_* of the session's map */
    double t_submit_ms;    /* arrival time the latency is measured from */
    double t_done_ms;      /* when the host saw the frame's step complete */
    int32_t kind;          /* FU_REQ_FRAME or FU_REQ_MESH */
    uint32_t n_triangles;  /* mesh requests: triangles written (min(total, max_triangles)) */
    uint32_t n_triangles_total;
    uint32_t n_occupied;   /* mesh requests: blocks with at least one triangle */
    int32_t rgb_count;     /* frames of sessions with use_rgb: photometric residuals of the last iteration */
} fu_completion;
#define FU_REQ_FRAME 0
#define FU_REQ_MESH 1
typedef struct {
    int32_t session;
    const uint16_t* depth; /* frame of the session's size: device-resident, or host memory with `host` */
    const uint8_t* rgb;
    double arrival_ms;     /* offset from the start of the trace */
    uint64_t tag;
    int32_t kind;          /* FU_REQ_FRAME, or FU_REQ_MESH with the buffers below */
    uint32_t max_triangles;
    float* vertex;
    float* normal;
    uint8_t* color;
    int32_t host;          /* nonzero: depth and rgb are host memory, as fu_server_submit_host */
} fu_trace_frame;
void fu_server_params_default(fu_server_params* p);
fu_status fu_server_create(const fu_server_params* p, fu_server** out);
fu_status fu_server_destroy(fu_server* sv);
fu_status fu_server_open(fu_server* sv, const fu_system_params* p, int32_t* session); /* FU_ERR_BUDGET if full */
fu_status fu_server_close(fu_server* sv, int32_t session);  /* its submitted frames must have completed */
/* the session's system (owned by the server): for its results, map and snapshots, and to
 * relocalise a lost session (fu_system_relocalise_graph, fu_system_validate_poses,
 * fu_system_relocalise) once its submitted frames have completed (fu_server_poll) */
fu_system* fu_server_system(fu_server* sv, int32_t session);
double fu_server_now_ms(fu_server* sv);
/* queue one frame; t_submit_ms < 0 means now */
fu_status fu_server_submit(fu_server* sv, int32_t session, const uint16_t* depth, const uint8_t* rgb, uint64_t tag,
                           double t_submit_ms);
/* the same for a frame in host memory (pinned, hipHostMalloc, for an asynchronous copy): the
 * step that takes it uploads it into the session's input slots on its lane's stream, as
 * Frame::Frame uploads each frame (Core/Frame.cc:61-62); the host buffers must stay unchanged
 * until the frame's completion is returned */
fu_status fu_server_submit_host(fu_server* sv, int32_t session, const uint16_t* depth, const uint8_t* rgb,
                                uint64_t tag, double t_submit_ms);
/* queue a mesh extraction of the session's map (fu_mesh_scene), ordered after the session's
 * frames submitted before it; the output buffers (9 floats, 9 floats and 9 bytes per
 * triangle) must stay untouched until its completion is returned */
fu_status fu_server_submit_mesh(fu_server* sv, int32_t session, float* vertex, float* normal, uint8_t* color,
                                uint32_t max_triangles, uint64_t tag, double t_submit_ms);
/* launch steps where lanes have room and collect finished frames (at most cap) */
fu_status fu_server_poll(fu_server* sv, fu_completion* out, int32_t cap, int32_t* n_out);
fu_status fu_server_pending(fu_server* sv, int32_t* frames); /* submitted, not yet returned by poll */
/* Replays a trace in real time: open loop submits each request at its arrival offset; closed
 * loop submits a session's next request when its previous one completes (arrivals ignored).
 * Fills out[0..n) in completion order and returns when every frame has completed. */
fu_status fu_server_run_trace(fu_server* sv, const fu_trace_frame* frames, int32_t n, int32_t closed_loop,
                              fu_completion* out);
/* stage profiling of every lane's steps (fu_batch_set_profiling); times summed over lanes */
fu_status fu_server_set_profiling(fu_server* sv, int enable);
fu_status fu_server_get_stage_times(fu_server* sv, fu_stage_times* out);
/* captured steps on every lane (fu_batch_set_graph); counts summed over lanes */
fu_status fu_server_set_graph(fu_server* sv, int enable);
fu_status fu_server_graph_counts(fu_server* sv, uint32_t* captures, uint32_t* replays);

#ifdef __cplusplus
}
#endif
#endif /* FUSION_AMD_H */
