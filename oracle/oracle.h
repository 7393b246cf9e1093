/*
 * CPU oracle of the dense RGB-D pipeline - TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C++ restatement of the reference's CUDA algorithm, used by tests/ (as what the
 * GPU results are compared with), by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py.  The product library never links or calls it.  Every function cites the
 * reference file:line it follows.
 * Floating-point arithmetic is written in the reference's operation order and the library is
 * built with -ffp-contract=off, so single-precision results are those of a strict IEEE
 * evaluation of the reference's expressions.
 *
 * Defined behaviour where the reference has none (DEVIATIONS):
 *  D1 float->int conversions saturate and map NaN to 0 (the GPU conversion instruction).
 *  D2 the block hash reduces the wrapped int products modulo the *unsigned* bucket count
 *     (DeviceMap.cu:3-10: NumBuckets is `uint`).
 *  D3 allocation is sequential in pixel order; no insert is dropped (the reference's
 *     try-lock may drop one under contention, DeviceMap.cu:45-60).
 *  D4 the raycast's one-block cache starts empty (the reference's default cache "hits" block
 *     (0,0,0) with ptr -1, DeviceMap.cu:100-103).
 *  D5 rendering fragments are neither lost to overlapping writes nor read from stale slots
 *     (RenderScene.cu:109-116 writes ny*ny fragments into nx*ny slots; fillBlocksKernel
 *     covers ceil(n/4)*4 slots, :269); the loop-bound and corner quirks themselves are kept.
 *  D6 ResizeMap bounds its grid by the destination size (the reference checks the source
 *     size, Pyrdown.cu:256, which reads past the source at 1280x720).
 *  D7 invalid vertex-map pixels are all-NaN (the reference sets only x, Pyrdown.cu:75).
 *  D11 ImageToIntensity's double expression (Pyrdown.cu:191) is evaluated as written, with no
 *     fused multiply-add.  For 63 of the 2^24 colours (those whose exact luma
 *     (2125 c1 + 7154 c0 + 721 c2) / 10000 is an integer) the truncated value depends on that
 *     choice, so a contracting compile of the reference may differ there by one level.
 *  D12 RGBStep's integer sums (count, sum of squared differences) wrap at 32 bits instead of
 *     overflowing (undefined in C++).
 *
 * Pinning (tests/test_oracle.py checks every function against closed forms, brute force or
 * numpy/scipy; the reference has no tests or golden vectors of its own; the keypoint map,
 * lifting, matching, graph selection and ValidatePose in tests/test_oracle_keymap.py,
 * test_oracle_reloc.py and test_oracle_graph.py: independent numpy statements, a plane's
 * closed form, a consistency matrix worked by hand, the selection rules checked pair by pair,
 * and the pose tracking gives the relocalised frame).  PARITY UNPINNED:
 * the exact bits of the reference's third-party solve and exponential - Eigen's LDLT
 * (Tracking.cc:388, Eigen >= 3.2, unversioned) and Sophus::SE3d::exp (Tracking.cc:389,
 * unpinned) - which or_solve6 / or_se3_exp replace by a Cholesky solve and the closed-form
 * Rodrigues / left-Jacobian exponential; the tests bound them against numpy.linalg.solve and
 * scipy.linalg.expm to about 1e-15, not to the reference's own outputs.
 */
#ifndef FU_ORACLE_H
#define FU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_map or_map;

/* the bilateral filter's exp: the same published method as the library (Cody-Waite
 * reduction + degree-7 Taylor polynomial) */
float or_expf(float x);

/* Tracking/Pyrdown.cu */
void or_filter_depth(const uint16_t* depth, float* raw, float* filt, int cols, int rows, float depth_scale,
                     float cutoff);
void or_pyrdown_gauss(const float* src, int cols, int rows, float* dst);
void or_compute_vmap(const float* depth, float* vmap, int cols, int rows, float fx, float fy, float cx, float cy,
                     float cutoff);
void or_compute_nmap(const float* vmap, float* nmap, int cols, int rows);
void or_resize_map(const float* vsrc, const float* nsrc, int cols, int rows, float* vdst, float* ndst);

/* Tracking/Reduction.cu ICPReduce: sums32 = the 29 sums accumulated in FP32 in pixel order,
 * sums64 = the same FP32 row products accumulated in FP64 */
void or_icp_sums(const float* next_vmap, const float* last_vmap, const float* next_nmap, const float* last_nmap,
                 int cols, int rows, const float Rcurr[9], const float tcurr[3], const float Rlast[9],
                 const float RlastInv[9], const float tlast[3], float fx, float fy, float cx, float cy,
                 float sums32[29], double sums64[29]);
/* solve A x = b for symmetric positive definite A (Cholesky, FP64) */
int or_solve6(const double A[36], const double b[6], double x[6]);
/* SE(3) exponential of (upsilon, omega) */
void or_se3_exp(const double xi[6], double T[16]);
/* Tracker::ComputeSE3 in ICP-only mode (Tracking.cc:296-409) with FP64 accumulation.
 * Maps: 3 levels. last_pose in, pose out (rolled back on failure). Returns FU_TRACK_* code
 * (1 ok, 2 failed, 3 rejected); err/count of the last iteration. */
int or_compute_se3(const float* const next_vmap[3], const float* const next_nmap[3], const float* const last_vmap[3],
                   const float* const last_nmap[3], int cols, int rows, const float K[4], const int iters[3],
                   const int min_count[3], float thresh, const double last_pose[16], double pose[16],
                   float* err, int* count);
/* the same with every iteration recorded: 22 doubles each (the 4x4 current pose its sums were
 * taken at, then its solution x = (upsilon, omega)); trace holds sum(iters) * 22 doubles */
int or_compute_se3_trace(const float* const next_vmap[3], const float* const next_nmap[3],
                         const float* const last_vmap[3], const float* const last_nmap[3], int cols, int rows,
                         const float K[4], const int iters[3], const int min_count[3], float thresh,
                         const double last_pose[16], double pose[16], double* trace, int* n_trace);
/* the acceptance test's |sin(XYZ Euler angles)| of a motion (Tracking.cc:398-408), 4x4 row-major */
double or_euler_sine_norm(const double p[16]);
/* photometric term (Tracking/Pyrdown.cu:127-249, Tracking/Reduction.cu:377-627) */
void or_image_to_intensity(const uint8_t* rgb, uint8_t* img, int cols, int rows);
void or_pyrdown_u8(const uint8_t* src, int cols, int rows, uint8_t* dst);
void or_derivative_image(const uint8_t* img, int cols, int rows, int16_t* dx, int16_t* dy);
/* RGBStep: corresp = {count, sum diff^2} of the residual pass, sigma, and the 29 sums of the
 * weighted rows (FP32 in pixel order; the same FP32 products accumulated in FP64) */
void or_rgb_sums(const uint8_t* next_img, const uint8_t* last_img, const float* next_vmap, const float* last_vmap,
                 const int16_t* dx, const int16_t* dy, int cols, int rows, const float Rcurr[9],
                 const float RcurrInv[9], const float Rlast[9], const float RlastInv[9], const float tcurr[3],
                 const float tlast[3], float fx, float fy, float cx, float cy, int32_t corresp[2], float* sigma,
                 float sums32[29], double sums64[29]);
/* Tracker::ComputeSE3(icpOnly = false): RGBStep + ICPStep per iteration, A = A_rgb * 1e-4f + A_icp */
int or_compute_se3_rgbd(const float* const next_vmap[3], const float* const next_nmap[3],
                        const float* const last_vmap[3], const float* const last_nmap[3],
                        const uint8_t* const next_img[3], const uint8_t* const last_img[3],
                        const int16_t* const dx[3], const int16_t* const dy[3], int cols, int rows, const float K[4],
                        const int iters[3], const int min_count[3], float thresh, const double last_pose[16],
                        double pose[16], float* err, int* count, int* rgb_count);
/* float views of a 4x4 pose: R, Rinv (transpose in double, then float), t */
void or_pose_f(const double pose[16], float R[9], float Rinv[9], float t[3]);

/* map (Mapping/DeviceMap.*, FuseMap.cu) */
or_map* or_map_create(uint32_t num_buckets, uint32_t num_excess, uint32_t num_blocks, float voxel_size,
                      float trunc_dist, float depth_min, float depth_max, uint32_t max_weight, int precision);
void or_map_destroy(or_map* m);
void or_map_reset(or_map* m);
/* counters: [0] allocated, [1] excess used, [2] visible, [3] heap-full events, [4] excess-full events */
void or_map_counters(const or_map* m, uint32_t out[5]);
void or_create_blocks(or_map* m, const float* depth, int cols, int rows, const float R[9], const float t[3], float fx,
                      float fy, float cx, float cy);
uint32_t or_check_visibility(or_map* m, const float Rinv[9], const float t[3], int cols, int rows, float fx, float fy,
                             float cx, float cy, float depth_max, float depth_min);
void or_integrate(or_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols, int rows,
                  const float Rinv[9], const float t[3], float fx, float fy, float cx, float cy);
uint32_t or_fuse_map_color(or_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols, int rows,
                           const float R[9], const float Rinv[9], const float t[3], float fx, float fy, float cx,
                           float cy);
int or_create_rendering_blocks(or_map* m, float* zmin, float* zmax, int zcols, int zrows, float depth_max,
                               float depth_min, const float Rinv[9], const float t[3], uint32_t n_visible, float fx,
                               float fy, float cx, float cy, uint32_t* n_fragments);
void or_raycast(or_map* m, float* vmap, float* nmap, int cols, int rows, const float* zmin, const float* zmax,
                int zcols, int zrows, const float R[9], const float Rinv[9], const float t[3], float invfx,
                float invfy, float cx, float cy);
/* marching cubes over all allocated blocks; returns the total triangle count (outputs are
 * written for the first max_tri); meta: 4 ints per triangle (voxel x,y,z, cube*16+t) */
uint32_t or_mesh(or_map* m, float* vertex, float* normal, uint8_t* color, int32_t* meta, uint32_t max_tri);
uint32_t or_map_export_blocks(const or_map* m, int32_t* xyz, uint32_t cap);
uint32_t or_map_export_visible(const or_map* m, int32_t* xyz, uint32_t cap);
/* sdf as float (FP64 maps: rounded), weight, rgb of the listed blocks; absent -> NaN/0/0 */
void or_map_read_blocks(const or_map* m, const int32_t* xyz, uint32_t n, float* sdf, uint16_t* weight, uint8_t* rgb);
void or_map_read_blocks_f64(const or_map* m, const int32_t* xyz, uint32_t n, double* sdf);
/* test hook: the blocks within trunc of a sphere, holding its exact truncated SDF */
void or_map_set_sphere(or_map* m, float cx, float cy, float cz, float r);
/* the bucket index of a block (DeviceMap::Hash) */
uint32_t or_hash(int32_t x, int32_t y, int32_t z, uint32_t num_buckets);

/* the per-frame pipeline (System::GrabImage, ICP-only) */
typedef struct or_system or_system;
or_system* or_system_create(int cols, int rows, float fx, float fy, float cx, float cy, float depth_scale,
                            float depth_cutoff, const int iters[3], const int min_count[3], float icp_thresh,
                            uint32_t num_buckets, uint32_t num_excess, uint32_t num_blocks, float voxel_size,
                            float trunc_dist, uint32_t max_weight, int precision);
void or_system_destroy(or_system* s);
/* returns the track status (1 ok, 2 failed, 3 rejected, 4 first frame) */
int or_system_grab(or_system* s, const uint16_t* depth, const uint8_t* rgb, double pose_out[16]);
/* a lost system's relocalisation pose for its next frame (status 6); -1 unless it is lost */
int or_system_relocalise(or_system* s, const double pose[16]);
or_map* or_system_map(or_system* s);
/* 1: track with the photometric term (ComputeSE3 icpOnly = false), building the intensity
 * pyramid and derivatives each frame */
void or_system_set_rgb(or_system* s, int use_rgb);
/* which: 0 intensity (u8), 1 last frame's intensity, 2 dIdx (i16), 3 dIdy */
const void* or_system_image(or_system* s, int which, int level);
int or_system_rgb_count(const or_system* s);
/* level-0 maps: 0 raw depth, 1 vmap, 2 nmap, 3 model vmap, 4 model nmap */
const float* or_system_buffer(or_system* s, int which, int level);
void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
