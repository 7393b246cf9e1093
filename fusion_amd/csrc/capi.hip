// C ABI of the library (include/fusion_amd.h): object lifetime, argument checking, launch
// sequencing and the two per-frame drivers (reference-style plain control and the
// device-resident sequence, optionally replayed as a hipGraph).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <algorithm>

#include "../../include/fusion_amd.h"
#include "fu_device.h"
#include "fu_kernels.h"
#include "fu_track.h"
#include "fu_batch.h"
#include "mc_tables.h"
#include "fu_api.h"

using namespace fu;

// ------------------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

static fu_status set_err(fu_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define HIPCHK(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return set_err(FU_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                            __FILE__, __LINE__);                                  \
    } while (0)

#define LAUNCHCHK()                                                                               \
    do {                                                                                          \
        hipError_t e_ = hipGetLastError();                                                        \
        if (e_ != hipSuccess) return set_err(FU_ERR_HIP, "kernel launch: %s (%s:%d)", hipGetErrorString(e_), \
                                            __FILE__, __LINE__);                                  \
    } while (0)

fu_status fu::api_error(fu_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static hipStream_t S(fu_stream s) { return (hipStream_t)s; }

extern "C" const char* fu_last_error(void) { return g_err; }
extern "C" const char* fu_version(void) { return "fusion_amd 0.1 (gfx950)"; }

static bool g_tables_uploaded = false;
static fu_status ensure_tables() {
    if (!g_tables_uploaded) {
        upload_mc_tables();
        HIPCHK(hipGetLastError());
        g_tables_uploaded = true;
    }
    return FU_OK;
}

extern "C" fu_status fu_mc_tables(int32_t* edge, int32_t* tri, int32_t* vert) {
    if (!edge || !tri || !vert) return set_err(FU_ERR_ARG, "fu_mc_tables: null output");
    for (int c = 0; c < 256; ++c) {
        edge[c] = mc_edge_mask(c);
        vert[c] = mc_vertex_count(c);
        int i = 0;
        for (const char* p = kMcTriHex[c]; *p; ++p, ++i) tri[c * 16 + i] = mc_hex_digit(*p);
        for (; i < 16; ++i) tri[c * 16 + i] = -1;
    }
    return FU_OK;
}

// ------------------------------------------------------------------------------ poses
static m3 m3_from(const float* r) { return m3{mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), mk3(r[6], r[7], r[8])}; }
static PoseF posef_from(const fu_pose_f* p) {
    PoseF q;
    q.R = m3_from(p->R);
    q.Rinv = m3_from(p->Rinv);
    q.t = mk3(p->t[0], p->t[1], p->t[2]);
    return q;
}

// ------------------------------------------------------------------------------ map

extern "C" void fu_map_params_default(fu_map_params* p) {
    // Mapping/DeviceMap.h:135-151 with the voxel size and truncation of the 5 mm configuration
    p->num_buckets = 1u << 21;
    p->num_blocks = 1u << 20;
    p->num_excess = p->num_blocks;
    p->voxel_size = 0.005f;
    p->trunc_dist = 0.005f * 3.0f;
    p->depth_min = 0.1f;
    p->depth_max = 3.0f;
    p->voxel_format = FU_VOXEL_F32_W8;
    p->max_weight = 255;
    p->max_render_blocks = 260000;
    p->max_new_blocks = 1u << 18;
}

// occupancy bitmap: about 8 bits per pool block, 2^20..2^25 bits
static unsigned occ_bits(const fu_map_params* p) {
    unsigned n = 1u << 20;
    while (n < (1u << 25) && (unsigned long long)n < 8ull * p->num_blocks) n <<= 1;
    return n;
}

static unsigned newset_size(const fu_map_params* p) {
    unsigned n = 1024;
    while (n < 2u * p->max_new_blocks) n <<= 1;
    return n;
}

extern "C" size_t fu_map_bytes(const fu_map_params* p) {
    const size_t ne = (size_t)p->num_buckets + p->num_excess;
    const size_t chunks = (p->num_blocks + 2047) / 2048;
    return ne * sizeof(HEntry) + 64 + (size_t)p->num_blocks * (8 + 4 + 8 + 4) + newset_size(p) * 8ull +
           (size_t)p->max_new_blocks * 12 + chunks * 4 + occ_bits(p) / 8 +
           (size_t)p->num_blocks * vf_block_bytes((int)p->voxel_format);
}

static fu_status check_params(const fu_map_params* p) {
    if (!p) return set_err(FU_ERR_ARG, "null map params");
    if (p->num_buckets == 0 || p->num_blocks == 0 || p->max_new_blocks == 0)
        return set_err(FU_ERR_ARG, "map capacities must be positive");
    if ((size_t)p->num_buckets + p->num_excess > 0x7fffffffu) return set_err(FU_ERR_ARG, "too many hash entries");
    if (p->num_blocks > (0x7fffffffu >> 9)) return set_err(FU_ERR_ARG, "num_blocks too large for 32-bit voxel offsets");
    if (p->voxel_format > FU_VOXEL_F16_W8) return set_err(FU_ERR_ARG, "unknown voxel format %u", p->voxel_format);
    const unsigned wmax = p->voxel_format == FU_VOXEL_F32_W16 ? 65535u : 255u;
    if (p->max_weight == 0 || p->max_weight > wmax)
        return set_err(FU_ERR_ARG, "max_weight %u outside [1, %u] for this format", p->max_weight, wmax);
    if (!(p->voxel_size > 0) || !(p->trunc_dist > 0)) return set_err(FU_ERR_ARG, "voxel_size and trunc_dist must be > 0");
    if (!(p->depth_min < p->depth_max)) return set_err(FU_ERR_ARG, "depth_min must be < depth_max");
    return FU_OK;
}

template <typename T>
static fu_status dalloc(fu_map* m, T** out, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(FU_ERR_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
    }
    m->allocs.push_back(p);
    *out = (T*)p;
    return FU_OK;
}

extern "C" fu_status fu_map_destroy(fu_map* m) {
    if (!m) return FU_OK;
    for (void* p : m->allocs) (void)hipFree(p);
    delete m;
    return FU_OK;
}

extern "C" fu_status fu_map_create(const fu_map_params* p, fu_map** out) {
    if (!out) return set_err(FU_ERR_ARG, "null output");
    *out = nullptr;
    fu_status st = check_params(p);
    if (st) return st;
    if ((st = ensure_tables())) return st;
    fu_map* m = new fu_map();
    m->p = *p;
    MapView& v = m->v;
    const unsigned ns = newset_size(p);
    const size_t chunks = (p->num_blocks + 2047) / 2048;
    if ((st = dalloc(m, &v.entries, (size_t)p->num_buckets + p->num_excess)) ||
        (st = dalloc(m, &v.ctr, (size_t)C_COUNT)) || (st = dalloc(m, &v.block_key, p->num_blocks)) ||
        (st = dalloc(m, &v.visible, p->num_blocks)) || (st = dalloc(m, &v.visible_key, p->num_blocks)) ||
        (st = dalloc(m, &v.newset, ns)) ||
        (st = dalloc(m, &v.newkeys, p->max_new_blocks)) || (st = dalloc(m, &v.newslot, p->max_new_blocks)) ||
        (st = dalloc(m, &m->mesh_counts, p->num_blocks)) || (st = dalloc(m, &m->mesh_chunk, chunks)) ||
        (st = dalloc(m, &v.occ, occ_bits(p) / 32)) ||
        (st = dalloc(m, &v.pool, (size_t)p->num_blocks * vf_block_bytes((int)p->voxel_format)))) {
        fu_map_destroy(m);
        return st;
    }
    v.num_buckets = p->num_buckets;
    v.num_excess = p->num_excess;
    v.num_blocks = p->num_blocks;
    v.newset_mask = ns - 1;
    v.occ_mask = occ_bits(p) - 1;
    v.occ_shift = 32u - (unsigned)__builtin_ctz(occ_bits(p));
    v.bucket_mask = (p->num_buckets & (p->num_buckets - 1)) == 0 ? p->num_buckets - 1 : 0u;
    v.max_new = p->max_new_blocks;
    v.max_weight = p->max_weight;
    v.fmt = (int)p->voxel_format;
    v.voxel_size = p->voxel_size;
    v.voxel_inv = (float)(1.0 / (double)p->voxel_size);  // DeviceMap.h:148
    v.trunc = p->trunc_dist;
    v.depth_min = p->depth_min;
    v.depth_max = p->depth_max;
    launch_reset_map(v, nullptr);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        fu_map_destroy(m);
        return set_err(FU_ERR_HIP, "map reset failed");
    }
    *out = m;
    return FU_OK;
}

extern "C" fu_status fu_map_get_params(const fu_map* m, fu_map_params* out) {
    if (!m || !out) return set_err(FU_ERR_ARG, "null argument");
    *out = m->p;
    return FU_OK;
}

extern "C" fu_status fu_reset_map(fu_map* m, fu_stream s) {
    if (!m) return set_err(FU_ERR_ARG, "null map");
    launch_reset_map(m->v, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_map_stats_get(fu_map* m, fu_map_stats* out, fu_stream s) {
    if (!m || !out) return set_err(FU_ERR_ARG, "null argument");
    unsigned c[C_COUNT];
    HIPCHK(hipMemcpyAsync(c, m->v.ctr, sizeof(c), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    out->n_allocated = c[C_ALLOC] < m->p.num_blocks ? c[C_ALLOC] : m->p.num_blocks;
    out->n_excess = c[C_EXCESS] < m->p.num_excess ? c[C_EXCESS] : m->p.num_excess;
    out->n_visible = c[C_VISIBLE];
    out->n_new = c[C_NEW];
    out->n_fragments = c[C_RENDER];
    out->n_triangles = c[C_TRI];
    out->flags = c[C_FLAGS];
    out->capacity = m->p.num_blocks;
    return FU_OK;
}

static unsigned read_ctr(fu_map* m, int which, hipStream_t s, fu_status* st) {
    unsigned v = 0;
    if (hipMemcpyAsync(&v, m->v.ctr + which, 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        *st = set_err(FU_ERR_HIP, "counter read failed");
    return v;
}

static FuseArgs make_fuse(const fu_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols,
                          int rows, const PoseF& pose, const TrackState* st, float fx, float fy, float cx, float cy) {
    FuseArgs a;
    a.pose = pose;
    a.st = st;
    a.fx = fx;
    a.fy = fy;
    a.cx = cx;
    a.cy = cy;
    a.invfx = (float)(1.0 / (double)fx);  // FuseMap.cu:280-281
    a.invfy = (float)(1.0 / (double)fy);
    a.cols = cols;
    a.rows = rows;
    a.depth = depth;
    a.nmap = (const f4*)nmap;
    a.rgb = rgb;
    a.px = nullptr;
    a.prgb = nullptr;
    (void)m;
    return a;
}

static VisArgs make_vis(const fu_map* m, const PoseF& pose, const TrackState* st, int cols, int rows, float fx,
                        float fy, float cx, float cy, float dmin, float dmax) {
    VisArgs v;
    v.pose = pose;
    v.st = st;
    v.fx = fx;
    v.fy = fy;
    v.cx = cx;
    v.cy = cy;
    v.cols = cols;
    v.rows = rows;
    v.depth_min = dmin;
    v.depth_max = dmax;
    (void)m;
    return v;
}

static constexpr unsigned kIntegrateGrid = 8192;

extern "C" fu_status fu_fuse_map_color(fu_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols,
                                       int rows, const fu_pose_f* pose, float fx, float fy, float cx, float cy,
                                       float depth_max, float depth_min, uint32_t* n_visible, fu_stream s) {
    (void)depth_max;
    (void)depth_min;
    if (!m || !depth || !rgb || !nmap || !pose) return set_err(FU_ERR_ARG, "fu_fuse_map_color: null argument");
    if (cols <= 0 || rows <= 0) return set_err(FU_ERR_ARG, "fu_fuse_map_color: bad size %dx%d", cols, rows);
    const PoseF p = posef_from(pose);
    const FuseArgs a = make_fuse(m, depth, rgb, nmap, cols, rows, p, nullptr, fx, fy, cx, cy);
    // the fusion visibility test uses DeviceMap::DepthMin/DepthMax (FuseMap.cu:288-289)
    const VisArgs v = make_vis(m, p, nullptr, cols, rows, fx, fy, cx, cy, m->p.depth_min, m->p.depth_max);
    launch_begin_pass(m->v, 3, S(s));
    launch_alloc(m->v, a, S(s));
    launch_visible(m->v, v, S(s));
    launch_integrate(m->v, a, kIntegrateGrid, S(s));
    LAUNCHCHK();
    if (n_visible) {
        fu_status st = FU_OK;
        *n_visible = read_ctr(m, C_VISIBLE, S(s), &st);
        return st;
    }
    return FU_OK;
}

extern "C" fu_status fu_check_block_visibility(fu_map* m, const fu_pose_f* pose, int cols, int rows, float fx,
                                               float fy, float cx, float cy, float depth_max, float depth_min,
                                               uint32_t* n_visible, fu_stream s) {
    if (!m || !pose) return set_err(FU_ERR_ARG, "fu_check_block_visibility: null argument");
    const VisArgs v = make_vis(m, posef_from(pose), nullptr, cols, rows, fx, fy, cx, cy, depth_min, depth_max);
    launch_begin_pass(m->v, 2, S(s));
    launch_visible(m->v, v, S(s));
    LAUNCHCHK();
    if (n_visible) {
        fu_status st = FU_OK;
        *n_visible = read_ctr(m, C_VISIBLE, S(s), &st);
        return st;
    }
    return FU_OK;
}

static ZRangeArgs make_zr(const fu_map* m, float* zmin, float* zmax, int zc, int zr, float dmax, float dmin,
                          const PoseF& p, const TrackState* st, int nvis_override, float fx, float fy, float cx,
                          float cy) {
    ZRangeArgs a;
    a.pose = p;
    a.st = st;
    a.fx = fx;
    a.fy = fy;
    a.cx = cx;
    a.cy = cy;
    a.depth_min = dmin;
    a.depth_max = dmax;
    a.zmin = (unsigned*)zmin;
    a.zmax = (unsigned*)zmax;
    a.zcols = zc;
    a.zrows = zr;
    a.max_fragments = m->p.max_render_blocks;
    a.n_visible_override = nvis_override;
    return a;
}

extern "C" fu_status fu_create_rendering_blocks(fu_map* m, float* zrange_min, float* zrange_max, int zcols,
                                                int zrows, float depth_max, float depth_min, const fu_pose_f* pose,
                                                uint32_t n_visible, float fx, float fy, float cx, float cy,
                                                uint32_t* n_fragments, int* rendered, fu_stream s) {
    if (!m || !zrange_min || !zrange_max || !pose) return set_err(FU_ERR_ARG, "fu_create_rendering_blocks: null argument");
    if (zcols <= 0 || zrows <= 0 || zcols > 32767 || zrows > 32767) return set_err(FU_ERR_ARG, "bad z-range size");
    if (n_visible == 0) {  // RenderScene.cu:214-215
        if (n_fragments) *n_fragments = 0;
        if (rendered) *rendered = 0;
        return FU_OK;
    }
    int ov = n_visible == 0xffffffffu ? -1 : (int)(n_visible < m->p.num_blocks ? n_visible : m->p.num_blocks);
    const ZRangeArgs a =
        make_zr(m, zrange_min, zrange_max, zcols, zrows, depth_max, depth_min, posef_from(pose), nullptr, ov, fx, fy, cx, cy);
    launch_begin_pass(m->v, 4, S(s));
    launch_zrange(m->v, a, S(s));
    LAUNCHCHK();
    if (n_fragments || rendered) {
        fu_status st = FU_OK;
        const unsigned n = read_ctr(m, C_RENDER, S(s), &st);
        if (n_fragments) *n_fragments = n;
        if (rendered) *rendered = n > 0;
        return st;
    }
    return FU_OK;
}

static RaycastArgs make_rc(const fu_map* m, float* vmap, float* nmap, int cols, int rows, const float* zmin,
                           const float* zmax, int zc, int zr, const PoseF& p, const TrackState* st, float invfx,
                           float invfy, float cx, float cy) {
    RaycastArgs a;
    a.pose = p;
    a.st = st;
    a.invfx = invfx;
    a.invfy = invfy;
    a.cx = cx;
    a.cy = cy;
    // DeviceMap::stepScale = 0.5 * TruncateDist * voxelSizeInv, evaluated in double (DeviceMap.h:151)
    a.step_scale = (float)(0.5 * (double)m->p.trunc_dist * (double)m->v.voxel_inv);
    a.cols = cols;
    a.rows = rows;
    a.zmin = (const unsigned*)zmin;
    a.zmax = (const unsigned*)zmax;
    a.zcols = zc;
    a.zrows = zr;
    a.vmap = (f4*)vmap;
    a.nmap = (f4*)nmap;
    a.keep_v = nullptr;
    a.keep_n = nullptr;
    return a;
}

extern "C" fu_status fu_raycast(fu_map* m, float* vmap, float* nmap, int cols, int rows, const float* zrange_min,
                                const float* zrange_max, int zcols, int zrows, const fu_pose_f* pose, float invfx,
                                float invfy, float cx, float cy, fu_stream s) {
    if (!m || !vmap || !nmap || !zrange_min || !zrange_max || !pose) return set_err(FU_ERR_ARG, "fu_raycast: null argument");
    if (zcols < (cols + 7) / 8 || zrows < (rows + 7) / 8) return set_err(FU_ERR_ARG, "z-range maps smaller than the image tiles");
    const RaycastArgs a = make_rc(m, vmap, nmap, cols, rows, zrange_min, zrange_max, zcols, zrows, posef_from(pose),
                                  nullptr, invfx, invfy, cx, cy);
    launch_raycast(m->v, a, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_mesh_scene(fu_map* m, float* vertex, float* normal, uint8_t* color, int32_t* tri_meta,
                                   uint32_t max_triangles, uint32_t* n_triangles, uint32_t* n_occupied, fu_stream s) {
    if (!m || !vertex || !normal || !color) return set_err(FU_ERR_ARG, "fu_mesh_scene: null argument");
    MeshOut o;
    o.vertex = vertex;
    o.normal = normal;
    o.color = color;
    o.tri_meta = tri_meta;
    o.max_tri = max_triangles;
    launch_mesh(m->v, o, m->mesh_counts, m->mesh_chunk, S(s));
    LAUNCHCHK();
    fu_status st = FU_OK;
    const unsigned tot = read_ctr(m, C_TRI, S(s), &st);
    if (st) return st;
    if (tot > max_triangles) {
        unsigned f = FU_FLAG_TRI_FULL;
        HIPCHK(hipMemcpy(&f, m->v.ctr + C_FLAGS, 4, hipMemcpyDeviceToHost));
        f |= FU_FLAG_TRI_FULL;
        HIPCHK(hipMemcpy(m->v.ctr + C_FLAGS, &f, 4, hipMemcpyHostToDevice));
    }
    if (n_triangles) *n_triangles = tot < max_triangles ? tot : max_triangles;
    if (n_occupied) *n_occupied = read_ctr(m, C_OCC, S(s), &st);
    return st;
}

// ---- inspection
extern "C" fu_status fu_map_export_blocks(fu_map* m, int32_t* xyz, uint32_t cap, uint32_t* n, fu_stream s) {
    if (!m || !n) return set_err(FU_ERR_ARG, "null argument");
    fu_status st = FU_OK;
    unsigned na = read_ctr(m, C_ALLOC, S(s), &st);
    if (st) return st;
    if (na > m->p.num_blocks) na = m->p.num_blocks;
    std::vector<unsigned long long> keys(na);
    if (na) HIPCHK(hipMemcpy(keys.data(), m->v.block_key, na * 8ull, hipMemcpyDeviceToHost));
    unsigned k = 0;
    for (unsigned i = 0; i < na; ++i) {
        if (keys[i] == kEmptyKey) continue;
        if (xyz && k < cap) {
            const i3 p = unpack_key(keys[i]);
            xyz[3 * k] = p.x;
            xyz[3 * k + 1] = p.y;
            xyz[3 * k + 2] = p.z;
        }
        ++k;
    }
    *n = k;
    return FU_OK;
}

extern "C" fu_status fu_map_export_visible(fu_map* m, int32_t* xyz, uint32_t cap, uint32_t* n, fu_stream s) {
    if (!m || !n) return set_err(FU_ERR_ARG, "null argument");
    fu_status st = FU_OK;
    unsigned nv = read_ctr(m, C_VISIBLE, S(s), &st);
    if (st) return st;
    if (nv > m->p.num_blocks) nv = m->p.num_blocks;
    std::vector<unsigned> ids(nv);
    std::vector<unsigned long long> keys(m->p.num_blocks);
    if (nv) HIPCHK(hipMemcpy(ids.data(), m->v.visible, nv * 4ull, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(keys.data(), m->v.block_key, keys.size() * 8ull, hipMemcpyDeviceToHost));
    for (unsigned i = 0; i < nv && xyz; ++i) {
        if (i >= cap) break;
        const i3 p = unpack_key(keys[ids[i]]);
        xyz[3 * i] = p.x;
        xyz[3 * i + 1] = p.y;
        xyz[3 * i + 2] = p.z;
    }
    *n = nv;
    return FU_OK;
}

extern "C" fu_status fu_map_read_blocks(fu_map* m, const int32_t* xyz, uint32_t n, float* sdf, uint16_t* weight,
                                        uint8_t* rgb, fu_stream s) {
    if (!m || (n && !xyz)) return set_err(FU_ERR_ARG, "null argument");
    HIPCHK(hipStreamSynchronize(S(s)));
    // locate blocks on the host from the key array (inspection path, not performance critical)
    fu_status st = FU_OK;
    unsigned na = read_ctr(m, C_ALLOC, S(s), &st);
    if (st) return st;
    if (na > m->p.num_blocks) na = m->p.num_blocks;
    std::vector<unsigned long long> keys(na);
    if (na) HIPCHK(hipMemcpy(keys.data(), m->v.block_key, na * 8ull, hipMemcpyDeviceToHost));
    std::vector<std::pair<unsigned long long, unsigned>> idx;
    idx.reserve(na);
    for (unsigned i = 0; i < na; ++i)
        if (keys[i] != kEmptyKey) idx.push_back({keys[i], i});
    std::sort(idx.begin(), idx.end());
    const int F = m->v.fmt;
    const int bb = vf_block_bytes(F), sb = vf_sdf_bytes(F), wb = vf_w_bytes(F);
    std::vector<unsigned char> buf(bb);
    for (unsigned b = 0; b < n; ++b) {
        const i3 p = mki3(xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]);
        long heap = -1;
        if (key_in_range(p)) {
            const unsigned long long k = pack_key(p);
            auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(k, 0u));
            if (it != idx.end() && it->first == k) heap = it->second;
        }
        for (int i = 0; i < 512; ++i) {
            float sv = NAN;
            unsigned wv = 0, r = 0, g = 0, bl = 0;
            if (heap >= 0) {
                if (i == 0) HIPCHK(hipMemcpy(buf.data(), m->v.pool + (size_t)heap * bb, bb, hipMemcpyDeviceToHost));
                if (sb == 4) {
                    memcpy(&sv, &buf[i * 4], 4);
                } else {
                    uint16_t h;
                    memcpy(&h, &buf[i * 2], 2);
                    sv = __half2float(*reinterpret_cast<__half*>(&h));
                }
                if (wb == 2) {  // sdf | u16 weight | r | g | b planes
                    uint16_t w16;
                    memcpy(&w16, &buf[512 * sb + i * 2], 2);
                    wv = w16;
                    r = buf[512 * (sb + wb) + i];
                    g = buf[512 * (sb + wb) + 512 + i];
                    bl = buf[512 * (sb + wb) + 1024 + i];
                } else {  // sdf | u32 (weight | r << 8 | g << 16 | b << 24) plane
                    wv = buf[512 * sb + 4 * i];
                    r = buf[512 * sb + 4 * i + 1];
                    g = buf[512 * sb + 4 * i + 2];
                    bl = buf[512 * sb + 4 * i + 3];
                }
            }
            if (sdf) sdf[(size_t)b * 512 + i] = sv;
            if (weight) weight[(size_t)b * 512 + i] = (uint16_t)wv;
            if (rgb) {
                rgb[((size_t)b * 512 + i) * 3 + 0] = (uint8_t)r;
                rgb[((size_t)b * 512 + i) * 3 + 1] = (uint8_t)g;
                rgb[((size_t)b * 512 + i) * 3 + 2] = (uint8_t)bl;
            }
        }
    }
    return FU_OK;
}

// ------------------------------------------------------------------------------ frame functions
static fu_status check_img(const void* a, int cols, int rows, const char* what) {
    if (!a) return set_err(FU_ERR_ARG, "%s: null image", what);
    if (cols <= 0 || rows <= 0 || (long)cols * rows > (1l << 28)) return set_err(FU_ERR_ARG, "%s: bad size %dx%d", what, cols, rows);
    return FU_OK;
}

extern "C" fu_status fu_filter_depth(const uint16_t* depth, float* raw, float* filtered, int cols, int rows,
                                     float depth_scale, float depth_cutoff, fu_stream s) {
    fu_status st = check_img(depth, cols, rows, "fu_filter_depth");
    if (st) return st;
    if (!raw || !filtered) return set_err(FU_ERR_ARG, "fu_filter_depth: null output");
    // FilterDepth passes 1.0 / depthScale computed in double (Pyrdown.cu:50)
    launch_filter_depth(depth, raw, filtered, cols, rows, (float)(1.0 / (double)depth_scale), depth_cutoff, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_pyrdown_gauss(const float* src, int cols, int rows, float* dst, fu_stream s) {
    fu_status st = check_img(src, cols, rows, "fu_pyrdown_gauss");
    if (st) return st;
    if (!dst || cols < 2 || rows < 2) return set_err(FU_ERR_ARG, "fu_pyrdown_gauss: bad destination");
    launch_pyrdown(src, cols, rows, dst, cols / 2, rows / 2, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_compute_vmap(const float* depth, float* vmap, int cols, int rows, float fx, float fy, float cx,
                                     float cy, float depth_cutoff, fu_stream s) {
    fu_status st = check_img(depth, cols, rows, "fu_compute_vmap");
    if (st) return st;
    if (!vmap) return set_err(FU_ERR_ARG, "fu_compute_vmap: null output");
    // ComputeVMap passes 1.0 / fx computed in double (Pyrdown.cu:87)
    launch_vmap(depth, (f4*)vmap, cols, rows, (float)(1.0 / (double)fx), (float)(1.0 / (double)fy), cx, cy,
                depth_cutoff, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_compute_nmap(const float* vmap, float* nmap, int cols, int rows, fu_stream s) {
    fu_status st = check_img(vmap, cols, rows, "fu_compute_nmap");
    if (st) return st;
    if (!nmap) return set_err(FU_ERR_ARG, "fu_compute_nmap: null output");
    launch_nmap((const f4*)vmap, (f4*)nmap, cols, rows, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_resize_map(const float* vsrc, const float* nsrc, int cols, int rows, float* vdst, float* ndst,
                                   fu_stream s) {
    fu_status st = check_img(vsrc, cols, rows, "fu_resize_map");
    if (st) return st;
    if (!nsrc || !vdst || !ndst || cols < 2 || rows < 2) return set_err(FU_ERR_ARG, "fu_resize_map: bad argument");
    launch_resize((const f4*)vsrc, (const f4*)nsrc, cols, (f4*)vdst, (f4*)ndst, cols / 2, rows / 2, S(s));
    LAUNCHCHK();
    return FU_OK;
}

// ------------------------------------------------------------------------------ tracker

// Work-groups of the ICP reductions, at most 4096.  The photometric lists and residual pass
// take one pixel (list entry) per lane, a kIcpThreads-pixel work-group each (up to 4M
// pixels; the combined pass pairs these groups); the ICP-only reduction kIcpLanePixels
// pixels per lane, a kIcpPixels chunk each.  Every
// pixel's association is a dependent gather, so latency is hidden across waves rather than
// by a per-lane loop.  Larger frames loop over chunks.
static_assert(kIcpMaxBlocks == 4096, "the partial-row buffers hold kIcpMaxBlocks rows");
static int icp_blocks(int npix, int per_group = kIcpThreads) {
    int nb = (npix + per_group - 1) / per_group;
    if (nb < 1) nb = 1;
    if (nb > kIcpMaxBlocks) nb = kIcpMaxBlocks;
    return nb;
}
static int icp_groups(int npix) { return icp_blocks(npix, kIcpPixels); }

static void init_state(TrackState* h) {
    memset(h, 0, sizeof(*h));
    for (int i = 0; i < 16; ++i) {
        const double I = (i % 5 == 0) ? 1.0 : 0.0;
        h->last_pose[i] = h->next_pose[i] = h->init_pose[i] = h->delta[i] = I;
    }
    h->status = TS_FIRST;
    h->sm_state = 1;  // Tracker::state before the first frame
    h->sm_prev = 1;
    h->after_reloc = 10;  // System::num_frames_after_reloc (System.cc:72)
    refresh_float_pose(h);
}

extern "C" fu_status fu_tracker_create(int cols, int rows, fu_tracker** out) {
    if (!out) return set_err(FU_ERR_ARG, "null output");
    *out = nullptr;
    if (cols < 4 || rows < 4) return set_err(FU_ERR_ARG, "tracker size %dx%d too small", cols, rows);
    fu_tracker* t = new fu_tracker();
    t->cols = cols;
    t->rows = rows;
    size_t nlist = 0;
    for (int l = 0; l < 3; ++l) {
        t->nblocks[l] = icp_blocks((cols >> l) * (rows >> l));
        t->ngroups[l] = icp_groups((cols >> l) * (rows >> l));
        t->list_off[l] = nlist;
        nlist += (size_t)(cols >> l) * (rows >> l);
    }
    if (hipMalloc(&t->st, sizeof(TrackState)) != hipSuccess || hipMalloc(&t->partials, kIcpMaxBlocks * 64 * 4) != hipSuccess ||
        hipMalloc(&t->rgb_corr, kIcpMaxBlocks * 2 * 4) != hipSuccess ||
        hipMalloc(&t->rgb_rec, (size_t)cols * rows * 4 + 256) != hipSuccess ||
        hipMalloc(&t->rgb_list, nlist * 4 + 16) != hipSuccess ||
        hipMalloc(&t->rgb_aux, (size_t)3 * (kIcpMaxBlocks + 64) * 4) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(t->st);
        (void)hipFree(t->partials);
        (void)hipFree(t->rgb_corr);
        (void)hipFree(t->rgb_rec);
        (void)hipFree(t->rgb_list);
        (void)hipFree(t->rgb_aux);
        delete t;
        return set_err(FU_ERR_NOMEM, "tracker allocation failed");
    }
    TrackState h;
    init_state(&h);
    HIPCHK(hipMemcpy(t->st, &h, sizeof(h), hipMemcpyHostToDevice));
    *out = t;
    return FU_OK;
}

extern "C" fu_status fu_tracker_destroy(fu_tracker* t) {
    if (!t) return FU_OK;
    (void)hipFree(t->st);
    (void)hipFree(t->partials);
    (void)hipFree(t->rgb_corr);
    (void)hipFree(t->rgb_rec);
    (void)hipFree(t->rgb_list);
    (void)hipFree(t->rgb_aux);
    delete t;
    return FU_OK;
}

extern "C" fu_status fu_tracker_set_last_pose(fu_tracker* t, const double pose[16], fu_stream s) {
    if (!t || !pose) return set_err(FU_ERR_ARG, "null argument");
    TrackState h;
    HIPCHK(hipStreamSynchronize(S(s)));
    HIPCHK(hipMemcpy(&h, t->st, sizeof(h), hipMemcpyDeviceToHost));
    for (int i = 0; i < 16; ++i) h.last_pose[i] = h.next_pose[i] = pose[i];
    refresh_float_pose(&h);
    HIPCHK(hipMemcpy(t->st, &h, sizeof(h), hipMemcpyHostToDevice));
    return FU_OK;
}

static IcpArgs make_icp(fu_tracker* t, int level, const float* nv, const float* nn, const float* lv, const float* ln,
                        const float K[4], int mode, int min_count, int final_iter, float thresh) {
    IcpArgs a;
    a.vcurr = (const f4*)nv;
    a.ncurr = (const f4*)nn;
    a.vlast = (const f4*)lv;
    a.nlast = (const f4*)ln;
    a.cols = t->cols >> level;
    a.rows = t->rows >> level;
    // Intrinsics::operator()(scale): K / (1 << scale) in float (Utility/Intrinsics.h:13-16)
    const float sc = (float)(1 << level);
    a.fx = K[0] / sc;
    a.fy = K[1] / sc;
    a.cx = K[2] / sc;
    a.cy = K[3] / sc;
    a.st = t->st;
    a.partials = t->partials;
    a.mode = mode;
    a.min_count = min_count;
    a.final_iter = final_iter;
    a.thresh = thresh;
    a.use_rgb = 0;
    memset(&a.rgb, 0, sizeof(a.rgb));
    return a;
}

// the photometric inputs of one level (RGBStep's images, Tracking.cc:319-325)
static void set_rgb(IcpArgs& a, fu_tracker* t, int level, const uint8_t* inext, const uint8_t* ilast,
                    const int16_t* dx, const int16_t* dy) {
    a.use_rgb = 1;
    a.rgb.inext = inext;
    a.rgb.ilast = ilast;
    a.rgb.dx = dx;
    a.rgb.dy = dy;
    a.rgb.list = t->rgb_list + t->list_off[level];
    a.rgb.count = t->rgb_aux + (size_t)level * (kIcpMaxBlocks + 64);
    a.rgb.wgcnt = a.rgb.count + 64;
    a.rgb.rec = t->rgb_rec;
    a.rgb.corr = t->rgb_corr;
}

// the intensity pyramids and derivatives of ComputeSE3(icpOnly = false); null for ICP only
struct RgbLevels {
    const uint8_t* const* next_image;
    const uint8_t* const* last_image;
    const int16_t* const* dx;
    const int16_t* const* dy;
};

// queue the whole coarse-to-fine loop (device-resident)
static void queue_se3(fu_tracker* t, const float* const nv[3], const float* const nn[3], const float* const lv[3],
                      const float* const ln[3], const RgbLevels* rgb, const float K[4], const int32_t iters[3],
                      const int32_t minc[3], float thresh, hipStream_t s) {
    launch_track_begin(t->st, s);
    int last_level = -1;
    for (int l = 2; l >= 0; --l)
        if (iters[l] > 0) last_level = l;
    if (rgb) {  // the eligible pixels of each level, once per frame
        IcpArgs la[3];
        for (int l = 0; l < 3; ++l) {
            la[l] = make_icp(t, l, nv[l], nn[l], lv[l], ln[l], K, 1, minc[l], 0, thresh);
            set_rgb(la[l], t, l, rgb->next_image[l], rgb->last_image[l], rgb->dx[l], rgb->dy[l]);
            la[l].use_rgb = iters[l] > 0;
        }
        launch_rgb_lists(la, t->nblocks, 3, s);
    }
    for (int l = 2; l >= 0; --l)
        for (int j = 0; j < iters[l]; ++j) {
            const int fin = (l == last_level && j == iters[l] - 1);
            IcpArgs a = make_icp(t, l, nv[l], nn[l], lv[l], ln[l], K, 1, minc[l], fin, thresh);
            if (rgb) set_rgb(a, t, l, rgb->next_image[l], rgb->last_image[l], rgb->dx[l], rgb->dy[l]);
            launch_icp(a, a.use_rgb ? t->nblocks[l] : t->ngroups[l], s);
        }
}

extern "C" fu_status fu_compute_se3(fu_tracker* t, const float* const next_vmap[3], const float* const next_nmap[3],
                                    const float* const last_vmap[3], const float* const last_nmap[3], const float K[4],
                                    const int32_t iters[3], const int32_t min_count[3], float thresh, fu_stream s) {
    if (!t || !next_vmap || !next_nmap || !last_vmap || !last_nmap || !K || !iters || !min_count)
        return set_err(FU_ERR_ARG, "fu_compute_se3: null argument");
    for (int l = 0; l < 3; ++l)
        if (!next_vmap[l] || !next_nmap[l] || !last_vmap[l] || !last_nmap[l] || iters[l] < 0)
            return set_err(FU_ERR_ARG, "fu_compute_se3: bad level %d", l);
    queue_se3(t, next_vmap, next_nmap, last_vmap, last_nmap, nullptr, K, iters, min_count, thresh, S(s));
    LAUNCHCHK();
    return FU_OK;
}

extern "C" fu_status fu_compute_se3_rgbd(fu_tracker* t, const float* const next_vmap[3],
                                         const float* const next_nmap[3], const float* const last_vmap[3],
                                         const float* const last_nmap[3], const uint8_t* const next_image[3],
                                         const uint8_t* const last_image[3], const int16_t* const dIdx[3],
                                         const int16_t* const dIdy[3], const float K[4], const int32_t iters[3],
                                         const int32_t min_count[3], float thresh, fu_stream s) {
    if (!t || !next_vmap || !next_nmap || !last_vmap || !last_nmap || !next_image || !last_image || !dIdx || !dIdy ||
        !K || !iters || !min_count)
        return set_err(FU_ERR_ARG, "fu_compute_se3_rgbd: null argument");
    if ((long)t->cols * t->rows > kRgbMaxPixels)
        return set_err(FU_ERR_ARG, "fu_compute_se3_rgbd: %dx%d exceeds %d pixels", t->cols, t->rows, kRgbMaxPixels);
    for (int l = 0; l < 3; ++l)
        if (!next_vmap[l] || !next_nmap[l] || !last_vmap[l] || !last_nmap[l] || !next_image[l] || !last_image[l] ||
            !dIdx[l] || !dIdy[l] || iters[l] < 0)
            return set_err(FU_ERR_ARG, "fu_compute_se3_rgbd: bad level %d", l);
    const RgbLevels rgb{next_image, last_image, dIdx, dIdy};
    queue_se3(t, next_vmap, next_nmap, last_vmap, last_nmap, &rgb, K, iters, min_count, thresh, S(s));
    LAUNCHCHK();
    return FU_OK;
}

static void fill_result(const TrackState& h, fu_track_result* out) {
    memcpy(out->pose, h.next_pose, sizeof(out->pose));
    memcpy(out->last_pose, h.last_pose, sizeof(out->last_pose));
    memcpy(out->x, h.x, sizeof(out->x));
    memcpy(out->A, h.A, sizeof(out->A));
    memcpy(out->b, h.b, sizeof(out->b));
    out->icp_error = h.icp_error;
    out->count = h.count;
    out->status = h.status;
    out->iterations = h.iters_done;
    out->rgb_error = h.rgb_error;
    out->rgb_count = h.rgb_count;
}

extern "C" fu_status fu_tracker_result(fu_tracker* t, fu_track_result* out, fu_stream s) {
    if (!t || !out) return set_err(FU_ERR_ARG, "null argument");
    TrackState h;
    HIPCHK(hipMemcpyAsync(&h, t->st, sizeof(h), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    fill_result(h, out);
    return FU_OK;
}

// ICPStep's body on a caller-given state and partial-row buffer (kIcpMaxBlocks rows of 64
// floats): the pose into `st`, the reduction, the sums back to the host.  Synchronises.
fu_status fu::icp_step_on(TrackState* st, float* partials, const float* next_vmap, const float* last_vmap,
                          const float* next_nmap, const float* last_nmap, int cols, int rows, const float Rcurr[9],
                          const float tcurr[3], const float Rlast[9], const float RlastInv[9], const float tlast[3],
                          float fx, float fy, float cx, float cy, float residual[2], double A[36], double b[6],
                          hipStream_t s) {
    TrackState h;
    memset(&h, 0, sizeof(h));
    memcpy(h.Rcurr, Rcurr, 36);
    memcpy(h.tcurr, tcurr, 12);
    memcpy(h.Rlast, Rlast, 36);
    memcpy(h.RlastInv, RlastInv, 36);
    memcpy(h.tlast, tlast, 12);
    h.status = TS_RUNNING;
    HIPCHK(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    IcpArgs a;
    memset(&a, 0, sizeof(a));
    a.vcurr = (const f4*)next_vmap;
    a.ncurr = (const f4*)next_nmap;
    a.vlast = (const f4*)last_vmap;
    a.nlast = (const f4*)last_nmap;
    a.cols = cols;
    a.rows = rows;
    a.fx = fx;
    a.fy = fy;
    a.cx = cx;
    a.cy = cy;
    a.st = st;
    a.partials = partials;
    a.mode = 0;
    launch_icp(a, icp_groups(cols * rows), s);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // CreateMatrix<6,7> (Reduction.cu:6-16)
    int shift = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            const double v = (double)h.sums[shift++];
            if (j == 6) {
                if (b) b[i] = v;
            } else if (A) {
                A[j * 6 + i] = A[i * 6 + j] = v;
            }
        }
    if (residual) {
        residual[0] = h.sums[27];
        residual[1] = h.sums[28];
    }
    return FU_OK;
}

extern "C" fu_status fu_icp_step(const float* next_vmap, const float* last_vmap, const float* next_nmap,
                                 const float* last_nmap, int cols, int rows, const float Rcurr[9], const float tcurr[3],
                                 const float Rlast[9], const float RlastInv[9], const float tlast[3], float fx,
                                 float fy, float cx, float cy, float residual[2], double A[36], double b[6],
                                 fu_stream s) {
    if (!next_vmap || !last_vmap || !next_nmap || !last_nmap || !Rcurr || !tcurr || !Rlast || !RlastInv || !tlast)
        return set_err(FU_ERR_ARG, "fu_icp_step: null argument");
    if (cols <= 0 || rows <= 0) return set_err(FU_ERR_ARG, "fu_icp_step: bad size");
    // scratch state per call: the reference's Tracker owns sumSE3/outSE3 (Tracking.cc:29-30)
    static thread_local fu_tracker* scratch = nullptr;
    if (!scratch) {
        fu_status st = fu_tracker_create(8, 8, &scratch);
        if (st) return st;
    }
    return icp_step_on(scratch->st, scratch->partials, next_vmap, last_vmap, next_nmap, last_nmap, cols, rows, Rcurr,
                       tcurr, Rlast, RlastInv, tlast, fx, fy, cx, cy, residual, A, b, S(s));
}

// RGBStep (Reduction.h:74-84, Reduction.cu:537-627): the photometric reduction of one level.
// residual = {sum of the weighted squared differences, count}, A and b in FP64 on the host,
// corresp = {count, sum of squared differences} of the residual pass (may be null).
// Synchronises.
extern "C" fu_status fu_rgb_step(const uint8_t* next_image, const uint8_t* last_image, const float* next_vmap,
                                 const float* last_vmap, const int16_t* dIdx, const int16_t* dIdy, int cols, int rows,
                                 const float Rcurr[9], const float RcurrInv[9], const float Rlast[9],
                                 const float RlastInv[9], const float tcurr[3], const float tlast[3], float fx,
                                 float fy, float cx, float cy, float residual[2], double A[36], double b[6],
                                 int32_t corresp[2], fu_stream s) {
    if (!next_image || !last_image || !next_vmap || !last_vmap || !dIdx || !dIdy || !Rcurr || !RcurrInv || !Rlast ||
        !RlastInv || !tcurr || !tlast)
        return set_err(FU_ERR_ARG, "fu_rgb_step: null argument");
    if (cols <= 0 || rows <= 0 || (long)cols * rows > kRgbMaxPixels)
        return set_err(FU_ERR_ARG, "fu_rgb_step: bad size %dx%d", cols, rows);
    // scratch per thread, grown to the largest image seen (the reference allocates its
    // residual array on every call, Reduction.cu:572)
    static thread_local fu_tracker* scratch = nullptr;
    if (!scratch || (long)scratch->cols * scratch->rows < (long)cols * rows) {
        fu_tracker_destroy(scratch);
        scratch = nullptr;
        fu_status st = fu_tracker_create(cols < 4 ? 4 : cols, rows < 4 ? 4 : rows, &scratch);
        if (st) return st;
    }
    TrackState h;
    memset(&h, 0, sizeof(h));
    memcpy(h.Rcurr, Rcurr, 36);
    memcpy(h.RcurrInv, RcurrInv, 36);
    memcpy(h.tcurr, tcurr, 12);
    memcpy(h.Rlast, Rlast, 36);
    memcpy(h.RlastInv, RlastInv, 36);
    memcpy(h.tlast, tlast, 12);
    h.status = TS_RUNNING;
    HIPCHK(hipMemcpyAsync(scratch->st, &h, sizeof(h), hipMemcpyHostToDevice, S(s)));
    IcpArgs a;
    memset(&a, 0, sizeof(a));
    // no normal maps: the combined pass computes only the photometric rows
    a.vcurr = (const f4*)next_vmap;
    a.vlast = (const f4*)last_vmap;
    a.ncurr = nullptr;
    a.nlast = nullptr;
    a.cols = cols;
    a.rows = rows;
    a.fx = fx;
    a.fy = fy;
    a.cx = cx;
    a.cy = cy;
    a.st = scratch->st;
    a.partials = scratch->partials;
    a.mode = 0;
    set_rgb(a, scratch, 0, next_image, last_image, dIdx, dIdy);
    const int nb = icp_blocks(cols * rows);
    launch_rgb_lists(&a, &nb, 1, S(s));
    launch_icp(a, nb, S(s));
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(&h, scratch->st, sizeof(h), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    double Ah[36], bh[6];
    unpack_sums29(h.rgb_sums, Ah, bh);
    if (A) memcpy(A, Ah, sizeof(Ah));
    if (b) memcpy(b, bh, sizeof(bh));
    if (residual) {
        residual[0] = h.rgb_sums[27];
        residual[1] = h.rgb_sums[28];
    }
    if (corresp) {
        corresp[0] = h.rgb_corresp[0];
        corresp[1] = h.rgb_corresp[1];
    }
    return FU_OK;
}

static fu_status check_u8(const void* a, int cols, int rows, const char* what) { return check_img(a, cols, rows, what); }

// ImageToIntensity (Reduction.h:34-35, Pyrdown.cu:196-206)
extern "C" fu_status fu_image_to_intensity(const uint8_t* rgb, uint8_t* image, int cols, int rows, fu_stream s) {
    fu_status st = check_u8(rgb, cols, rows, "fu_image_to_intensity");
    if (st) return st;
    if (!image) return set_err(FU_ERR_ARG, "fu_image_to_intensity: null output");
    launch_intensity(rgb, image, cols, rows, S(s));
    LAUNCHCHK();
    return FU_OK;
}

// PyrDownGauss on bytes (Reduction.h:31-32, Pyrdown.cu:171-181); dst is (cols/2) x (rows/2)
extern "C" fu_status fu_pyrdown_gauss_u8(const uint8_t* src, int cols, int rows, uint8_t* dst, fu_stream s) {
    fu_status st = check_u8(src, cols, rows, "fu_pyrdown_gauss_u8");
    if (st) return st;
    if (!dst || cols < 4 || rows < 4) return set_err(FU_ERR_ARG, "fu_pyrdown_gauss_u8: bad destination");
    launch_pyrdown_u8(src, cols, rows, dst, cols / 2, rows / 2, S(s));
    LAUNCHCHK();
    return FU_OK;
}

// ComputeDerivativeImage (Reduction.h:37-38, Pyrdown.cu:238-249)
extern "C" fu_status fu_compute_derivative_image(const uint8_t* image, int16_t* dIdx, int16_t* dIdy, int cols,
                                                 int rows, fu_stream s) {
    fu_status st = check_u8(image, cols, rows, "fu_compute_derivative_image");
    if (st) return st;
    if (!dIdx || !dIdy) return set_err(FU_ERR_ARG, "fu_compute_derivative_image: null output");
    launch_derivative(image, dIdx, dIdy, cols, rows, S(s));
    LAUNCHCHK();
    return FU_OK;
}

// ------------------------------------------------------------------------------ system

extern "C" void fu_system_params_default(fu_system_params* p) {
    memset(p, 0, sizeof(*p));
    p->cols = 640;
    p->rows = 480;
    p->fx = p->fy = 525.0f;
    p->cx = 319.5f;
    p->cy = 239.5f;
    p->depth_scale = 5000.0f;
    p->depth_cutoff = 3.0f;
    p->iters[0] = 10;
    p->iters[1] = 5;
    p->iters[2] = 4;
    p->min_count[0] = 2000;
    p->min_count[1] = 1000;
    p->min_count[2] = 100;
    p->icp_thresh = 1e-4f;
    p->plain = 0;
    p->use_graph = 0;
    fu_map_params_default(&p->map);
}

template <typename T>
static fu_status salloc(fu_system* s, T** out, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, count * sizeof(T) + 16) != hipSuccess) {
        (void)hipGetLastError();
        return set_err(FU_ERR_NOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
    }
    s->allocs.push_back(p);
    *out = (T*)p;
    return FU_OK;
}

extern "C" fu_status fu_system_destroy(fu_system* sys) {
    if (!sys) return FU_OK;
    if (sys->batch) fu_batch_detach(sys->batch, sys);
    if (sys->graph) (void)hipGraphExecDestroy(sys->graph);
    for (int k = 0; k <= FU_NUM_STAGES; ++k)
        if (sys->ev[k]) (void)hipEventDestroy(sys->ev[k]);
    for (void* p : sys->allocs) (void)hipFree(p);
    fu_map_destroy(sys->map);
    fu_tracker_destroy(sys->trk);
    fu_tracker_destroy(sys->reloc_trk);
    delete sys;
    return FU_OK;
}

extern "C" fu_status fu_system_create(const fu_system_params* p, fu_system** out) {
    if (!p || !out) return set_err(FU_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->cols < 32 || p->rows < 32 || (p->cols % 4) || (p->rows % 4))
        return set_err(FU_ERR_ARG, "frame size %dx%d must be >= 32 and a multiple of 4", p->cols, p->rows);
    if (!(p->depth_scale > 0) || !(p->fx > 0) || !(p->fy > 0)) return set_err(FU_ERR_ARG, "bad camera parameters");
    if (p->use_rgb && (long)p->cols * p->rows > kRgbMaxPixels)
        return set_err(FU_ERR_ARG, "photometric tracking is limited to %d pixels", kRgbMaxPixels);
    fu_system* sys = new fu_system();
    sys->p = *p;
    fu_status st;
    if ((st = fu_map_create(&p->map, &sys->map)) || (st = fu_tracker_create(p->cols, p->rows, &sys->trk))) {
        fu_system_destroy(sys);
        return st;
    }
    fu_system_buffers& b = sys->b;
    memset(&b, 0, sizeof(b));
    const size_t n0 = (size_t)p->cols * p->rows;
    for (int l = 0; l < 3; ++l) {
        sys->lc[l] = p->cols >> l;
        sys->lr[l] = p->rows >> l;
    }
    b.zcols = (p->cols + 7) / 8;
    b.zrows = (p->rows + 7) / 8;
    if ((st = salloc(sys, &sys->fuse_px, n0)) || (st = salloc(sys, &sys->fuse_rgb, n0)) ||
        (st = salloc(sys, &b.depth_in, n0)) || (st = salloc(sys, &b.rgb_in, n0 * 3)) || (st = salloc(sys, &b.raw, n0)) ||
        (st = salloc(sys, &b.zrange_min, (size_t)b.zcols * b.zrows)) ||
        (st = salloc(sys, &b.zrange_max, (size_t)b.zcols * b.zrows))) {
        fu_system_destroy(sys);
        return st;
    }
    for (int l = 0; l < 3; ++l) {
        const size_t n = (size_t)sys->lc[l] * sys->lr[l];
        if ((st = salloc(sys, &b.depth[l], n)) || (st = salloc(sys, &b.vmap[l], n * 4)) ||
            (st = salloc(sys, &b.nmap[l], n * 4)) || (st = salloc(sys, &b.model_vmap[l], n * 4)) ||
            (st = salloc(sys, &b.model_nmap[l], n * 4))) {
            fu_system_destroy(sys);
            return st;
        }
        if (p->use_rgb && ((st = salloc(sys, &b.image[l], n)) || (st = salloc(sys, &b.last_image[l], n)) ||
                           (st = salloc(sys, &b.dIdx[l], n)) || (st = salloc(sys, &b.dIdy[l], n)))) {
            fu_system_destroy(sys);
            return st;
        }
    }
    *out = sys;
    return fu_system_reset(sys, nullptr);
}

extern "C" fu_status fu_system_reset(fu_system* sys, fu_stream s) {
    if (!sys) return set_err(FU_ERR_ARG, "null system");
    fu_status st = fu_reset_map(sys->map, s);
    if (st) return st;
    TrackState h;
    init_state(&h);
    h.mapping_off = sys->mapping_off ? 1 : 0;  // RebootSystem keeps mappingDisabled (System.cc:272-277)
    HIPCHK(hipMemcpyAsync(sys->trk->st, &h, sizeof(h), hipMemcpyHostToDevice, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    sys->frame = 0;
    sys->frame_at_load = 0;
    return FU_OK;
}

// Tracker::mappingDisabled (Tracking.h:45; set by ReadMapFromDisk, System.cc:266-270, and
// from the viewer): frames are still tracked, rendered and raycast, never fused
extern "C" fu_status fu_system_set_mapping(fu_system* sys, int enabled, fu_stream s) {
    if (!sys) return set_err(FU_ERR_ARG, "fu_system_set_mapping: null system");
    TrackState h;
    HIPCHK(hipMemcpyAsync(&h, sys->trk->st, sizeof(h), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    h.mapping_off = enabled ? 0 : 1;
    HIPCHK(hipMemcpyAsync(sys->trk->st, &h, sizeof(h), hipMemcpyHostToDevice, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    sys->mapping_off = !enabled;
    return FU_OK;
}

extern "C" fu_status fu_system_get_buffers(fu_system* sys, fu_system_buffers* out) {
    if (!sys || !out) return set_err(FU_ERR_ARG, "null argument");
    *out = sys->b;
    return FU_OK;
}

extern "C" fu_map* fu_system_map(fu_system* sys) { return sys ? sys->map : nullptr; }

// ImageToIntensity, PyrDownGauss of the intensity, ComputeDerivativeImage (Frame.cc:63-76)
static void queue_images(fu_system* sys, const uint8_t* rgb, hipStream_t s) {
    fu_system_buffers& b = sys->b;
    launch_intensity(rgb, b.image[0], sys->lc[0], sys->lr[0], s);
    for (int l = 1; l < 3; ++l) launch_pyrdown_u8(b.image[l - 1], sys->lc[l - 1], sys->lr[l - 1], b.image[l], sys->lc[l], sys->lr[l], s);
    for (int l = 0; l < 3; ++l) launch_derivative(b.image[l], b.dIdx[l], b.dIdy[l], sys->lc[l], sys->lr[l], s);
}

// Tracker::SwapFrame of the intensity pyramid: on the device, conditioned on the tracking state
static void queue_image_commit(fu_system* sys, hipStream_t s) {
    fu_system_buffers& b = sys->b;
    const unsigned char* src[3] = {b.image[0], b.image[1], b.image[2]};
    unsigned char* dst[3] = {b.last_image[0], b.last_image[1], b.last_image[2]};
    const int n[3] = {sys->lc[0] * sys->lr[0], sys->lc[1] * sys->lr[1], sys->lc[2] * sys->lr[2]};
    launch_image_commit(sys->trk->st, src, dst, n, s);
}

// Frame::FillImages (Core/Frame.cc:59-78); the intensity images only with the photometric term
static void queue_pyramid(fu_system* sys, const uint16_t* depth, hipStream_t s) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    launch_filter_depth(depth, b.raw, b.depth[0], p.cols, p.rows, (float)(1.0 / (double)p.depth_scale), p.depth_cutoff, s);
    for (int l = 1; l < 3; ++l) launch_pyrdown(b.depth[l - 1], sys->lc[l - 1], sys->lr[l - 1], b.depth[l], sys->lc[l], sys->lr[l], s);
    for (int l = 0; l < 3; ++l) {
        const float sc = (float)(1 << l);  // Frame::SetK: K / 2^l (Core/Frame.cc:229-237)
        const float fx = p.fx / sc, fy = p.fy / sc, cx = p.cx / sc, cy = p.cy / sc;
        launch_vmap(b.depth[l], (f4*)b.vmap[l], sys->lc[l], sys->lr[l], (float)(1.0 / (double)fx),
                    (float)(1.0 / (double)fy), cx, cy, p.depth_cutoff, s);
        launch_nmap((const f4*)b.vmap[l], (f4*)b.nmap[l], sys->lc[l], sys->lr[l], s);
    }
}

// LastFrame->ResizeImages (Core/Frame.cc:80-84) on the model maps
static void queue_resize_model(fu_system* sys, hipStream_t s) {
    fu_system_buffers& b = sys->b;
    for (int l = 1; l < 3; ++l)
        launch_resize((const f4*)b.model_vmap[l - 1], (const f4*)b.model_nmap[l - 1], sys->lc[l - 1],
                      (f4*)b.model_vmap[l], (f4*)b.model_nmap[l], sys->lc[l], sys->lr[l], s);
}

static void mark(fu_system* sys, int boundary, hipStream_t s) {
    if (sys->prof) (void)hipEventRecord(sys->ev[boundary], s);
}

// Mapping::FuseColor + Mapping::RayTrace (Mapping.cc:76-109), pose from the tracking state
static void queue_fuse_raycast(fu_system* sys, const uint8_t* rgb, hipStream_t s) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    fu_map* m = sys->map;
    const TrackState* st = sys->trk->st;
    PoseF dummy{};
    FuseArgs a = make_fuse(m, b.raw, rgb, b.nmap[0], p.cols, p.rows, dummy, st, p.fx, p.fy, p.cx, p.cy);
    a.px = sys->fuse_px;
    a.prgb = sys->fuse_rgb;
    const VisArgs v = make_vis(m, dummy, st, p.cols, p.rows, p.fx, p.fy, p.cx, p.cy, m->p.depth_min, m->p.depth_max);
    launch_begin_pass(m->v, 7, s);
    launch_fuse_pack(a, s);
    launch_alloc(m->v, a, s);
    mark(sys, FU_STAGE_ALLOC + 1, s);
    launch_visible(m->v, v, s);
    mark(sys, FU_STAGE_VISIBLE + 1, s);
    launch_integrate(m->v, a, kIntegrateGrid, s);
    mark(sys, FU_STAGE_INTEGRATE + 1, s);
    const ZRangeArgs z = make_zr(m, b.zrange_min, b.zrange_max, b.zcols, b.zrows, m->p.depth_max, m->p.depth_min, dummy,
                                 st, -1, p.fx, p.fy, p.cx, p.cy);
    launch_zrange(m->v, z, s);
    mark(sys, FU_STAGE_ZRANGE + 1, s);
    RaycastArgs r = make_rc(m, b.model_vmap[0], b.model_nmap[0], p.cols, p.rows, b.zrange_min, b.zrange_max, b.zcols,
                            b.zrows, dummy, st, (float)(1.0 / (double)p.fx), (float)(1.0 / (double)p.fy), p.cx, p.cy);
    r.keep_v = (const f4*)b.vmap[0];
    r.keep_n = (const f4*)b.nmap[0];
    launch_raycast(m->v, r, s);
    mark(sys, FU_STAGE_RAYCAST + 1, s);
}

static void queue_frame_device(fu_system* sys, const uint16_t* depth, const uint8_t* rgb, bool first, hipStream_t s) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    mark(sys, 0, s);
    if (!first) queue_resize_model(sys, s);
    mark(sys, FU_STAGE_RESIZE + 1, s);
    queue_pyramid(sys, depth, s);
    if (p.use_rgb) queue_images(sys, rgb, s);
    mark(sys, FU_STAGE_PYRAMID + 1, s);
    if (!first) {
        const float* nv[3] = {b.vmap[0], b.vmap[1], b.vmap[2]};
        const float* nn[3] = {b.nmap[0], b.nmap[1], b.nmap[2]};
        const float* lv[3] = {b.model_vmap[0], b.model_vmap[1], b.model_vmap[2]};
        const float* ln[3] = {b.model_nmap[0], b.model_nmap[1], b.model_nmap[2]};
        const float K[4] = {p.fx, p.fy, p.cx, p.cy};
        const uint8_t* ni[3] = {b.image[0], b.image[1], b.image[2]};
        const uint8_t* li[3] = {b.last_image[0], b.last_image[1], b.last_image[2]};
        const int16_t* gx[3] = {b.dIdx[0], b.dIdx[1], b.dIdx[2]};
        const int16_t* gy[3] = {b.dIdy[0], b.dIdy[1], b.dIdy[2]};
        const RgbLevels rgbl{ni, li, gx, gy};
        queue_se3(sys->trk, nv, nn, lv, ln, p.use_rgb ? &rgbl : nullptr, K, p.iters, p.min_count, p.icp_thresh, s);
    }
    launch_track_commit(sys->trk->st, s);
    if (p.use_rgb) queue_image_commit(sys, s);
    mark(sys, FU_STAGE_ICP + 1, s);
    queue_fuse_raycast(sys, rgb, s);
}

// Reference-style control: the same kernels, but every ICP iteration's sums come back to the
// host, which solves and uploads the new pose (Tracking.cc:316-396), and each fusion/raycast
// stage synchronises and reads its count as FuseMapColor/CreateRenderingBlocks do.
static fu_status run_frame_plain(fu_system* sys, const uint16_t* depth, const uint8_t* rgb, bool first, hipStream_t s) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    fu_tracker* t = sys->trk;
    if (!first) queue_resize_model(sys, s);
    queue_pyramid(sys, depth, s);
    if (p.use_rgb) queue_images(sys, rgb, s);
    HIPCHK(hipStreamSynchronize(s));
    TrackState h;
    HIPCHK(hipMemcpy(&h, t->st, sizeof(h), hipMemcpyDeviceToHost));
    if (!first) frame_begin(&h);  // a lost tracker does not track
    if (!first && h.status == TS_RUNNING) {
        const float K[4] = {p.fx, p.fy, p.cx, p.cy};
        int last_level = -1;
        for (int l = 2; l >= 0; --l)
            if (p.iters[l] > 0) last_level = l;
        if (p.use_rgb) {
            IcpArgs la[3];
            for (int l = 0; l < 3; ++l) {
                la[l] = make_icp(t, l, b.vmap[l], b.nmap[l], b.model_vmap[l], b.model_nmap[l], K, 0, 0, 0, 0);
                set_rgb(la[l], t, l, b.image[l], b.last_image[l], b.dIdx[l], b.dIdy[l]);
                la[l].use_rgb = p.iters[l] > 0;
            }
            launch_rgb_lists(la, t->nblocks, 3, s);
        }
        for (int l = 2; l >= 0 && h.status == TS_RUNNING; --l)
            for (int j = 0; j < p.iters[l] && h.status == TS_RUNNING; ++j) {
                HIPCHK(hipMemcpyAsync(t->st, &h, sizeof(h), hipMemcpyHostToDevice, s));
                IcpArgs a = make_icp(t, l, b.vmap[l], b.nmap[l], b.model_vmap[l], b.model_nmap[l], K, 0, 0, 0, 0);
                if (p.use_rgb) set_rgb(a, t, l, b.image[l], b.last_image[l], b.dIdx[l], b.dIdy[l]);
                launch_icp(a, a.use_rgb ? t->nblocks[l] : t->ngroups[l], s);
                LAUNCHCHK();
                HIPCHK(hipMemcpyAsync(&h, t->st, sizeof(h), hipMemcpyDeviceToHost, s));
                HIPCHK(hipStreamSynchronize(s));
                float sums[32], rsums[32];
                memcpy(sums, h.sums, sizeof(sums));
                memcpy(rsums, h.rgb_sums, sizeof(rsums));
                const int fin = (l == last_level && j == p.iters[l] - 1);
                icp_update(&h, sums, p.use_rgb ? rsums : nullptr, p.min_count[l], fin, p.icp_thresh);
            }
    }
    frame_commit(&h);  // what the frame does, and SwapFrame of its pose
    HIPCHK(hipMemcpyAsync(t->st, &h, sizeof(h), hipMemcpyHostToDevice, s));
    if (p.use_rgb && (h.gates & GATE_SWAP))  // Tracker::SwapFrame
        for (int l = 0; l < 3; ++l)
            HIPCHK(hipMemcpyAsync(b.last_image[l], b.image[l], (size_t)sys->lc[l] * sys->lr[l], hipMemcpyDeviceToDevice, s));
    if (!(h.gates & GATE_RENDER)) {
        HIPCHK(hipStreamSynchronize(s));
        return FU_OK;
    }
    // FuseMapColor with its synchronisations (FuseMap.cu:294-317), or UpdateVisibility alone
    // when the frame does not fuse
    const bool fuse = (h.gates & GATE_FUSE) != 0;
    fu_map* m = sys->map;
    PoseF dummy{};
    FuseArgs a = make_fuse(m, b.raw, rgb, b.nmap[0], p.cols, p.rows, dummy, t->st, p.fx, p.fy, p.cx, p.cy);
    a.px = sys->fuse_px;
    a.prgb = sys->fuse_rgb;
    const VisArgs v = make_vis(m, dummy, t->st, p.cols, p.rows, p.fx, p.fy, p.cx, p.cy, m->p.depth_min, m->p.depth_max);
    launch_begin_pass(m->v, 7, s);
    if (fuse) {
        launch_fuse_pack(a, s);
        launch_alloc(m->v, a, s);
        HIPCHK(hipStreamSynchronize(s));
    }
    launch_visible(m->v, v, s);
    fu_status st = FU_OK;
    const unsigned nvis = read_ctr(m, C_VISIBLE, s, &st);
    if (st) return st;
    if (nvis > 0 && fuse) {
        launch_integrate(m->v, a, nvis < m->p.num_blocks ? nvis : m->p.num_blocks, s);
        HIPCHK(hipStreamSynchronize(s));
    }
    // Mapping::RayTrace (Mapping.cc:98-109)
    unsigned nfrag = 0;
    if (nvis > 0) {
        const ZRangeArgs z = make_zr(m, b.zrange_min, b.zrange_max, b.zcols, b.zrows, m->p.depth_max, m->p.depth_min,
                                     dummy, t->st, -1, p.fx, p.fy, p.cx, p.cy);
        launch_zrange(m->v, z, s);
        nfrag = read_ctr(m, C_RENDER, s, &st);
        if (st) return st;
    }
    if (nfrag > 0) {
        const RaycastArgs r = make_rc(m, b.model_vmap[0], b.model_nmap[0], p.cols, p.rows, b.zrange_min,
                                      b.zrange_max, b.zcols, b.zrows, dummy, t->st, (float)(1.0 / (double)p.fx),
                                      (float)(1.0 / (double)p.fy), p.cx, p.cy);
        launch_raycast(m->v, r, s);
    } else {
        HIPCHK(hipMemcpyAsync(b.model_vmap[0], b.vmap[0], (size_t)p.cols * p.rows * 16, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(b.model_nmap[0], b.nmap[0], (size_t)p.cols * p.rows * 16, hipMemcpyDeviceToDevice, s));
    }
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(s));
    return FU_OK;
}

// The frame's work captured once per system against the system's own input slots
// (depth_in, rgb_in); later frames copy their images into the slots and replay it.
static fu_status capture_graph(fu_system* sys, hipGraphExec_t* out) {
    hipStream_t cs;
    HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return set_err(FU_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    }
    queue_frame_device(sys, sys->b.depth_in, sys->b.rgb_in, false, cs);
    e = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return set_err(FU_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(FU_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e));
    sys->graph_captures++;
    return FU_OK;
}

extern "C" fu_status fu_system_grab_image_from(fu_system* sys, const uint16_t* depth, const uint8_t* rgb, fu_stream s) {
    if (!sys || !depth || !rgb) return set_err(FU_ERR_ARG, "fu_system_grab_image: null argument");
    const bool first = sys->frame == 0;
    hipStream_t hs = S(s);
    fu_status st = FU_OK;
    if (sys->p.plain) {
        st = run_frame_plain(sys, depth, rgb, first, hs);
    } else if (sys->p.use_graph && !first && !sys->prof) {
        const size_t n0 = (size_t)sys->p.cols * sys->p.rows;
        if (depth != sys->b.depth_in)
            HIPCHK(hipMemcpyAsync(sys->b.depth_in, depth, n0 * 2, hipMemcpyDeviceToDevice, hs));
        if (rgb != sys->b.rgb_in) HIPCHK(hipMemcpyAsync(sys->b.rgb_in, rgb, n0 * 3, hipMemcpyDeviceToDevice, hs));
        if (!sys->graph && (st = capture_graph(sys, &sys->graph))) return st;
        HIPCHK(hipGraphLaunch(sys->graph, hs));
        sys->graph_launches++;
    } else {
        queue_frame_device(sys, depth, rgb, first, hs);
        LAUNCHCHK();
    }
    if (st) return st;
    if (sys->prof && !sys->p.plain && !first) {
        HIPCHK(hipEventSynchronize(sys->ev[FU_NUM_STAGES]));
        for (int k = 0; k < FU_NUM_STAGES; ++k) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, sys->ev[k], sys->ev[k + 1]));
            sys->times.ms[k] += ms;
        }
        unsigned c[C_COUNT];
        TrackState h;
        HIPCHK(hipMemcpy(c, sys->map->v.ctr, sizeof(c), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&h, sys->trk->st, sizeof(h), hipMemcpyDeviceToHost));
        sys->times.frames += 1;
        sys->times.steps += 1;
        sys->times.n_visible += c[C_VISIBLE];
        sys->times.n_new += c[C_NEW];
        if (!first) {
            sys->times.icp_iterations += h.iters_done;
            // pixel-iterations actually reduced (levels run to completion unless tracking failed)
            double px = 0;
            int left = h.iters_done;
            for (int l = 2; l >= 0 && left > 0; --l) {
                const int it = std::min(left, sys->p.iters[l]);
                px += (double)it * sys->lc[l] * sys->lr[l];
                left -= it;
            }
            sys->times.icp_pixels += px;
        }
    }
    sys->frame++;
    return FU_OK;
}

extern "C" fu_status fu_system_set_profiling(fu_system* sys, int enable) {
    if (!sys) return set_err(FU_ERR_ARG, "null system");
    if (enable && !sys->ev[0])
        for (int k = 0; k <= FU_NUM_STAGES; ++k) HIPCHK(hipEventCreate(&sys->ev[k]));
    sys->prof = enable != 0;
    memset(&sys->times, 0, sizeof(sys->times));
    return FU_OK;
}

extern "C" fu_status fu_system_get_stage_times(fu_system* sys, fu_stage_times* out) {
    if (!sys || !out) return set_err(FU_ERR_ARG, "null argument");
    *out = sys->times;
    return FU_OK;
}

extern "C" fu_status fu_system_grab_image(fu_system* sys, fu_stream s) {
    if (!sys) return set_err(FU_ERR_ARG, "null system");
    return fu_system_grab_image_from(sys, sys->b.depth_in, sys->b.rgb_in, s);
}

// System::GrabImage with the frame in host memory: the upload of Frame::Frame
// (Core/Frame.cc:61-62) on the stream, into the system's input slots, then the frame.  With
// pinned host memory the copy is asynchronous; the host buffers must stay unchanged until it
// has run.
extern "C" fu_status fu_system_grab_image_host(fu_system* sys, const uint16_t* depth, const uint8_t* rgb, fu_stream s) {
    if (!sys || !depth || !rgb) return set_err(FU_ERR_ARG, "fu_system_grab_image_host: null argument");
    const size_t n0 = (size_t)sys->p.cols * sys->p.rows;
    HIPCHK(hipMemcpyAsync(sys->b.depth_in, depth, n0 * 2, hipMemcpyHostToDevice, S(s)));
    HIPCHK(hipMemcpyAsync(sys->b.rgb_in, rgb, n0 * 3, hipMemcpyHostToDevice, S(s)));
    return fu_system_grab_image_from(sys, sys->b.depth_in, sys->b.rgb_in, s);
}

extern "C" fu_status fu_system_relocalise(fu_system* sys, const double pose[16], fu_stream s) {
    if (!sys || !pose) return set_err(FU_ERR_ARG, "fu_system_relocalise: null argument");
    TrackState h;
    HIPCHK(hipMemcpyAsync(&h, sys->trk->st, sizeof(h), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    if (h.sm_state != -1) return set_err(FU_ERR_STATE, "fu_system_relocalise: the system is not lost");
    for (int i = 0; i < 16; ++i) h.last_pose[i] = h.next_pose[i] = pose[i];
    h.reloc = 1;
    refresh_float_pose(&h);
    HIPCHK(hipMemcpyAsync(sys->trk->st, &h, sizeof(h), hipMemcpyHostToDevice, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    return FU_OK;
}

// ------------------------------------------------------- relocalisation by graph matching
static fu_pose_f pose_f_of(const double P[16]) {  // Frame::GpuRotation / GpuInvRotation / GpuTranslation
    fu_pose_f f;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            f.R[i * 3 + j] = (float)P[i * 4 + j];
            f.Rinv[i * 3 + j] = (float)P[j * 4 + i];
        }
        f.t[i] = (float)P[i * 4 + 3];
    }
    return f;
}

// Mapping::UpdateVisibility at P (Mapping.cc:60-66); unless fewer than min_vis blocks are
// visible, Mapping::RayTrace into the model maps (Mapping.cc:92-109: left as they are when no
// block renders, or, with keep_own, the frame's own maps as after a swap) and, with resize,
// LastFrame->ResizeImages.  *rendered tells whether the visibility test passed.
static fu_status render_at(fu_system* sys, const double P[16], unsigned min_vis, bool keep_own, bool resize,
                           bool* rendered, hipStream_t s) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    fu_map* m = sys->map;
    const fu_pose_f pf = pose_f_of(P);
    const PoseF pose = posef_from(&pf);
    *rendered = false;
    const VisArgs v = make_vis(m, pose, nullptr, p.cols, p.rows, p.fx, p.fy, p.cx, p.cy, m->p.depth_min, m->p.depth_max);
    launch_begin_pass(m->v, 2, s);
    launch_visible(m->v, v, s);
    LAUNCHCHK();
    fu_status st = FU_OK;
    const unsigned nvis = read_ctr(m, C_VISIBLE, s, &st);
    if (st) return st;
    if (nvis < min_vis) return FU_OK;
    *rendered = true;
    unsigned nfrag = 0;
    if (nvis > 0) {
        const ZRangeArgs z = make_zr(m, b.zrange_min, b.zrange_max, b.zcols, b.zrows, m->p.depth_max, m->p.depth_min,
                                     pose, nullptr, -1, p.fx, p.fy, p.cx, p.cy);
        launch_begin_pass(m->v, 4, s);
        launch_zrange(m->v, z, s);
        LAUNCHCHK();
        nfrag = read_ctr(m, C_RENDER, s, &st);
        if (st) return st;
    }
    if (nfrag > 0) {
        const RaycastArgs r = make_rc(m, b.model_vmap[0], b.model_nmap[0], p.cols, p.rows, b.zrange_min, b.zrange_max,
                                      b.zcols, b.zrows, pose, nullptr, (float)(1.0 / (double)p.fx),
                                      (float)(1.0 / (double)p.fy), p.cx, p.cy);
        launch_raycast(m->v, r, s);
    } else if (keep_own) {
        HIPCHK(hipMemcpyAsync(b.model_vmap[0], b.vmap[0], (size_t)p.cols * p.rows * 16, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(b.model_nmap[0], b.nmap[0], (size_t)p.cols * p.rows * 16, hipMemcpyDeviceToDevice, s));
    }
    if (resize) queue_resize_model(sys, s);
    LAUNCHCHK();
    return FU_OK;
}

// ComputeSE3(icpOnly = true, iters, thresh) of the system's current frame against the model
// maps, from P, on the scratch tracker t
static fu_status icp_from(fu_system* sys, fu_tracker* t, const double P[16], const int32_t iters[3], float thresh,
                          fu_track_result* out, hipStream_t s) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    fu_status st = fu_tracker_set_last_pose(t, P, s);
    if (st) return st;
    const float* nv[3] = {b.vmap[0], b.vmap[1], b.vmap[2]};
    const float* nn[3] = {b.nmap[0], b.nmap[1], b.nmap[2]};
    const float* lv[3] = {b.model_vmap[0], b.model_vmap[1], b.model_vmap[2]};
    const float* ln[3] = {b.model_nmap[0], b.model_nmap[1], b.model_nmap[2]};
    const float K[4] = {p.fx, p.fy, p.cx, p.cy};
    queue_se3(t, nv, nn, lv, ln, nullptr, K, iters, p.min_count, thresh, s);
    LAUNCHCHK();
    return fu_tracker_result(t, out, s);
}

// Tracker::ValidatePose (Tracking.cc:636-682) on a lost system's current frame
extern "C" fu_status fu_system_validate_poses(fu_system* sys, const double* poses, int32_t n,
                                              const int32_t iters_reloc[3], float thresh_reloc, int32_t* accepted,
                                              int32_t* chosen, int32_t* ok, fu_stream s) {
    if (!sys || !ok || !iters_reloc || n < 0 || (n > 0 && (!poses || !accepted)))
        return set_err(FU_ERR_ARG, "fu_system_validate_poses: bad argument");
    for (int l = 0; l < 3; ++l)
        if (iters_reloc[l] < 0) return set_err(FU_ERR_ARG, "fu_system_validate_poses: negative iterations");
    *ok = 0;
    if (chosen) *chosen = -1;
    hipStream_t hs = S(s);
    TrackState h;
    HIPCHK(hipMemcpyAsync(&h, sys->trk->st, sizeof(h), hipMemcpyDeviceToHost, hs));
    HIPCHK(hipStreamSynchronize(hs));
    if (h.sm_state != -1) return set_err(FU_ERR_STATE, "fu_system_validate_poses: the system is not lost");
    if (sys->frame == sys->frame_at_load)  // the pyramid to validate against is a grabbed frame's
        return set_err(FU_ERR_STATE, "fu_system_validate_poses: no frame since the system was reset or loaded");
    // the candidates' ICP runs on a tracker of the system's own, kept for later attempts
    fu_status st = FU_OK;
    if (!sys->reloc_trk && (st = fu_tracker_create(sys->p.cols, sys->p.rows, &sys->reloc_trk))) return st;
    fu_tracker* t = sys->reloc_trk;
    int first = -1;
    fu_track_result res;
    for (int i = 0; i < n && !st; ++i) {
        const double* P = poses + 16 * i;
        bool rendered = false;
        accepted[i] = -1;
        // below 512 visible blocks the candidate is skipped (:642-645)
        if ((st = render_at(sys, P, 512, false, true, &rendered, hs)) || !rendered) continue;
        if ((st = icp_from(sys, t, P, iters_reloc, thresh_reloc, &res, hs))) continue;
        accepted[i] = res.status == FU_TRACK_OK;
        if (accepted[i] && first < 0) first = i;
    }
    // the first candidate that passed (the reference sorts the passing errors with
    // cv::sortIdx on a one-column matrix, which sorts each one-element row and leaves index 0
    // first), rendered again and refined with the tracking iterations and threshold
    bool rendered = false;
    if (!st && first >= 0) st = render_at(sys, poses + 16 * first, 0, false, true, &rendered, hs);
    if (!st && first >= 0) st = icp_from(sys, t, poses + 16 * first, sys->p.iters, sys->p.icp_thresh, &res, hs);
    if (st) return st;
    if (chosen) *chosen = first;
    if (first < 0 || res.status != FU_TRACK_OK) return FU_OK;
    // relocalised (Tracking.cc:101-105): the frame takes the refined pose and is swapped in;
    // GrabImage in state -1 renders it without fusion (System.cc:99-120); tracking resumes
    // with the next frame, which is not swapped (Tracking.cc:76-86)
    for (int i = 0; i < 16; ++i) h.last_pose[i] = h.next_pose[i] = res.pose[i];
    h.sm_state = 0;
    h.sm_prev = -1;
    h.status = TS_RELOC;
    h.gates = GATE_RENDER | GATE_SWAP;
    h.reloc = 0;
    h.icp_error = res.icp_error;
    h.count = res.count;
    refresh_float_pose(&h);
    HIPCHK(hipMemcpyAsync(sys->trk->st, &h, sizeof(h), hipMemcpyHostToDevice, hs));
    if (sys->p.use_rgb)  // Tracker::SwapFrame of the intensity pyramid
        for (int l = 0; l < 3; ++l)
            HIPCHK(hipMemcpyAsync(sys->b.last_image[l], sys->b.image[l], (size_t)sys->lc[l] * sys->lr[l],
                                  hipMemcpyDeviceToDevice, hs));
    if ((st = render_at(sys, h.next_pose, 0, true, false, &rendered, hs))) return st;
    HIPCHK(hipStreamSynchronize(hs));
    *ok = 1;
    return FU_OK;
}

// Tracker::Relocalise with graph matching on (Tracking.cc:411-453, 469-682)
extern "C" fu_status fu_system_relocalise_graph(fu_system* sys, fu_keymap* km, const float* frame_desc,
                                                const float* frame_pos, const float* frame_normal, int32_t nq,
                                                int32_t iterations, uint64_t seed, int32_t* n_candidates,
                                                int32_t* ok, fu_stream s) {
    if (!sys || !km || !ok) return set_err(FU_ERR_ARG, "fu_system_relocalise_graph: null argument");
    *ok = 0;
    double poses[5 * 16];
    int32_t np = 0, accepted[5], chosen = -1;
    fu_status st = fu_relocalise_graph_candidates(km, frame_desc, frame_pos, frame_normal, nq, iterations, seed, poses,
                                                  &np, nullptr, s);
    if (st) return st;
    if (n_candidates) *n_candidates = np;
    const int32_t iters_reloc[3] = {3, 3, 2};  // Tracker::ITERATIONS_RELOC (Tracking.h:101)
    return fu_system_validate_poses(sys, poses, np, iters_reloc, 1e-3f, accepted, &chosen, ok, s);
}

extern "C" fu_status fu_system_graph_counts(fu_system* sys, uint32_t* captures, uint32_t* launches) {
    if (!sys || !captures || !launches) return set_err(FU_ERR_ARG, "null argument");
    *captures = sys->graph_captures;
    *launches = sys->graph_launches;
    return FU_OK;
}

extern "C" fu_status fu_system_result(fu_system* sys, fu_frame_result* out, fu_stream s) {
    if (!sys || !out) return set_err(FU_ERR_ARG, "null argument");
    TrackState h;
    unsigned c[C_COUNT];
    HIPCHK(hipMemcpyAsync(&h, sys->trk->st, sizeof(h), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipMemcpyAsync(c, sys->map->v.ctr, sizeof(c), hipMemcpyDeviceToHost, S(s)));
    HIPCHK(hipStreamSynchronize(S(s)));
    memcpy(out->pose, h.next_pose, sizeof(out->pose));
    out->track_status = h.status;
    out->fused = (h.gates & GATE_FUSE) ? 1 : 0;
    out->icp_error = h.icp_error;
    out->icp_count = h.count;
    out->n_visible = c[C_VISIBLE];
    out->n_allocated = c[C_ALLOC] < sys->p.map.num_blocks ? c[C_ALLOC] : sys->p.map.num_blocks;
    out->n_fragments = c[C_RENDER];
    out->flags = c[C_FLAGS];
    out->frame_index = sys->frame;
    out->rgb_count = sys->p.use_rgb ? h.rgb_count : 0;
    return FU_OK;
}

// ------------------------------------------------------------------------------ batches
namespace fu {
int batch_layout(BatchLaunch& b, const SessDesc* h, BatchTiling t, int level, int fixed) {
    int total = 0, uni = -1;
    for (int i = 0; i < b.n; ++i) {
        const SessDesc& d = h[b.sess[i]];
        int gx = 1, cnt = 1;
        switch (t) {
            case BT_TILES:
                gx = (d.lc[level] + 15) / 16;
                cnt = gx * ((d.lr[level] + 15) / 16);
                break;
            case BT_PYR3:
                cnt = 0;
                for (int l = 0; l < 3; ++l) cnt += ((d.lc[l] + 15) / 16) * ((d.lr[l] + 15) / 16);
                gx = cnt;
                break;
            case BT_ICP: gx = cnt = d.icp_blocks[level]; break;
            case BT_ONE: gx = cnt = 1; break;
            case BT_FIXED: gx = cnt = fixed; break;
            case BT_ZTILES: gx = cnt = (d.zr.zcols * d.zr.zrows + 255) / 256; break;
            case BT_ICP_PAIRS: gx = cnt = (d.icp_blocks[level] + 1) / 2; break;
        }
        b.pre[i] = total;
        b.gx[i] = (unsigned short)gx;
        total += cnt;
        uni = i == 0 ? cnt : (uni == cnt ? uni : 0);
    }
    b.pre[b.n] = total;
    b.uniform = uni > 0 ? uni : 0;
    return total;
}
}  // namespace fu

struct fu_batch {
    int cap = 0;
    int fmt = -1;
    int rgb = -1;  // tracking mode shared by the attached systems (use_rgb)
    SessDesc* d_desc = nullptr;
    std::vector<SessDesc> h_desc;
    std::vector<fu_system*> owner;
    int attached = 0;
    bool prof = false;
    hipEvent_t ev[FU_NUM_STAGES + 1] = {};
    fu_stage_times times{};
    // captured steps (fu_batch_set_graph): keyed by the step's sessions, inputs and result
    // slot, the context's membership generation and the launch geometry's generation
    struct Graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    static constexpr size_t kGraphs = 32, kSeen = 64;
    int use_graph = 0;
    uint64_t gen = 1, tuning_gen = 0, clock = 0;
    std::vector<Graph> graphs;
    std::vector<std::vector<uint64_t>> seen;  // keys met once, not captured
    hipStream_t cap_stream = nullptr;
    uint32_t captures = 0, replays = 0;
    void drop_graphs() {
        for (Graph& g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
        seen.clear();
        ++gen;
    }
};

// generation of the process-wide launch geometry (captured steps depend on it)
static uint64_t g_tuning_gen = 1;

extern "C" fu_status fu_batch_create(int capacity, fu_batch** out) {
    if (!out) return set_err(FU_ERR_ARG, "null output");
    *out = nullptr;
    if (capacity < 1 || capacity > 65535) return set_err(FU_ERR_ARG, "batch capacity %d out of range", capacity);
    fu_batch* bt = new fu_batch();
    bt->cap = capacity;
    bt->h_desc.resize(capacity);
    bt->owner.assign(capacity, nullptr);
    if (hipMalloc(&bt->d_desc, sizeof(SessDesc) * capacity) != hipSuccess) {
        (void)hipGetLastError();
        delete bt;
        return set_err(FU_ERR_NOMEM, "batch descriptor allocation failed");
    }
    *out = bt;
    return FU_OK;
}

extern "C" fu_status fu_batch_destroy(fu_batch* bt) {
    if (!bt) return FU_OK;
    for (fu_system* sys : bt->owner)
        if (sys) {
            sys->batch = nullptr;
            sys->batch_slot = -1;
        }
    for (int k = 0; k <= FU_NUM_STAGES; ++k)
        if (bt->ev[k]) (void)hipEventDestroy(bt->ev[k]);
    bt->drop_graphs();
    if (bt->cap_stream) (void)hipStreamDestroy(bt->cap_stream);
    (void)hipFree(bt->d_desc);
    delete bt;
    return FU_OK;
}

extern "C" fu_status fu_batch_set_graph(fu_batch* bt, int enable) {
    if (!bt) return set_err(FU_ERR_ARG, "null batch");
    bt->use_graph = enable != 0;
    if (!enable) bt->drop_graphs();
    return FU_OK;
}

extern "C" fu_status fu_batch_graph_counts(fu_batch* bt, uint32_t* captures, uint32_t* replays) {
    if (!bt || !captures || !replays) return set_err(FU_ERR_ARG, "null argument");
    *captures = bt->captures;
    *replays = bt->replays;
    return FU_OK;
}

extern "C" fu_status fu_batch_set_profiling(fu_batch* bt, int enable) {
    if (!bt) return set_err(FU_ERR_ARG, "null batch");
    if (enable && !bt->ev[0])
        for (int k = 0; k <= FU_NUM_STAGES; ++k) HIPCHK(hipEventCreate(&bt->ev[k]));
    bt->prof = enable != 0;
    memset(&bt->times, 0, sizeof(bt->times));
    return FU_OK;
}

extern "C" fu_status fu_batch_get_stage_times(fu_batch* bt, fu_stage_times* out) {
    if (!bt || !out) return set_err(FU_ERR_ARG, "null argument");
    *out = bt->times;
    return FU_OK;
}

// ------------------------------------------------------------------------------ tuning
static fu_tuning g_tuning = {FU_WAVE_8X8, 0, 16384, 8192, 2048};
const fu_tuning& fu::tuning() { return g_tuning; }

extern "C" void fu_tuning_default(fu_tuning* t) {
    if (t) *t = fu_tuning{FU_WAVE_8X8, 0, 16384, 8192, 2048};
}
extern "C" fu_status fu_set_tuning(const fu_tuning* t) {
    if (!t) return set_err(FU_ERR_ARG, "null tuning");
    if (t->raycast_wave < FU_WAVE_16X4 || t->raycast_wave > FU_WAVE_4X16)
        return set_err(FU_ERR_ARG, "fu_set_tuning: raycast_wave %d (0..2)", t->raycast_wave);
    if (t->xcd_order & ~((1 << FU_NUM_STAGES) - 1))
        return set_err(FU_ERR_ARG, "fu_set_tuning: xcd_order 0x%x has bits beyond the stages", t->xcd_order);
    const int32_t g[3] = {t->integrate_groups, t->visible_groups, t->alloc_groups};
    for (int i = 0; i < 3; ++i)
        if (g[i] < 256 || g[i] > 262144) return set_err(FU_ERR_ARG, "fu_set_tuning: work-group count %d (256..262144)", g[i]);
    g_tuning = *t;
    ++g_tuning_gen;
    return FU_OK;
}
extern "C" fu_status fu_get_tuning(fu_tuning* t) {
    if (!t) return set_err(FU_ERR_ARG, "null tuning");
    *t = g_tuning;
    return FU_OK;
}

static void fill_desc(fu_system* sys, SessDesc& d) {
    const fu_system_params& p = sys->p;
    fu_system_buffers& b = sys->b;
    fu_map* m = sys->map;
    fu_tracker* t = sys->trk;
    memset(&d, 0, sizeof(d));
    d.m = m->v;
    d.st = t->st;
    d.partials = t->partials;
    d.cols = p.cols;
    d.rows = p.rows;
    d.depth_scale_inv = (float)(1.0 / (double)p.depth_scale);
    d.cutoff = p.depth_cutoff;
    d.raw = b.raw;
    d.use_rgb = p.use_rgb ? 1 : 0;
    const float K[4] = {p.fx, p.fy, p.cx, p.cy};
    d.last_level = -1;
    for (int l = 2; l >= 0; --l)
        if (p.iters[l] > 0) d.last_level = l;
    for (int l = 0; l < 3; ++l) {
        d.lc[l] = sys->lc[l];
        d.lr[l] = sys->lr[l];
        const float sc = (float)(1 << l);
        d.Kl[l][0] = p.fx / sc;
        d.Kl[l][1] = p.fy / sc;
        d.Kl[l][2] = p.cx / sc;
        d.Kl[l][3] = p.cy / sc;
        d.invf[l][0] = (float)(1.0 / (double)d.Kl[l][0]);
        d.invf[l][1] = (float)(1.0 / (double)d.Kl[l][1]);
        d.depth[l] = b.depth[l];
        d.vmap[l] = (f4*)b.vmap[l];
        d.nmap[l] = (f4*)b.nmap[l];
        d.model_v[l] = (f4*)b.model_vmap[l];
        d.model_n[l] = (f4*)b.model_nmap[l];
        d.icp[l] = make_icp(t, l, b.vmap[l], b.nmap[l], b.model_vmap[l], b.model_nmap[l], K, 1, p.min_count[l], 0,
                            p.icp_thresh);
        if (p.use_rgb) {
            set_rgb(d.icp[l], t, l, b.image[l], b.last_image[l], b.dIdx[l], b.dIdy[l]);
            d.image[l] = b.image[l];
            d.last_image[l] = b.last_image[l];
            d.didx[l] = b.dIdx[l];
            d.didy[l] = b.dIdy[l];
        }
        d.icp_blocks[l] = p.use_rgb ? t->nblocks[l] : t->ngroups[l];
        d.iters[l] = p.iters[l];
    }
    PoseF dummy{};
    d.fuse = make_fuse(m, b.raw, nullptr, b.nmap[0], p.cols, p.rows, dummy, t->st, p.fx, p.fy, p.cx, p.cy);
    d.fuse.px = sys->fuse_px;
    d.fuse.prgb = sys->fuse_rgb;
    d.vis = make_vis(m, dummy, t->st, p.cols, p.rows, p.fx, p.fy, p.cx, p.cy, m->p.depth_min, m->p.depth_max);
    d.zr = make_zr(m, b.zrange_min, b.zrange_max, b.zcols, b.zrows, m->p.depth_max, m->p.depth_min, dummy, t->st, -1,
                   p.fx, p.fy, p.cx, p.cy);
    d.rc = make_rc(m, b.model_vmap[0], b.model_nmap[0], p.cols, p.rows, b.zrange_min, b.zrange_max, b.zcols, b.zrows,
                   dummy, t->st, d.invf[0][0], d.invf[0][1], p.cx, p.cy);
    d.rc.keep_v = (const f4*)b.vmap[0];
    d.rc.keep_n = (const f4*)b.nmap[0];
}

extern "C" fu_status fu_batch_attach(fu_batch* bt, fu_system* sys) {
    if (!bt || !sys) return set_err(FU_ERR_ARG, "null argument");
    if (sys->batch) return set_err(FU_ERR_STATE, "system is already attached to a batch context");
    if (bt->fmt >= 0 && bt->fmt != sys->p.map.voxel_format)
        return set_err(FU_ERR_ARG, "voxel format %d differs from the batch context's %d", sys->p.map.voxel_format, bt->fmt);
    if (bt->rgb >= 0 && bt->rgb != (sys->p.use_rgb ? 1 : 0))
        return set_err(FU_ERR_ARG, "tracking mode (use_rgb %d) differs from the batch context's", sys->p.use_rgb);
    int tiles = 0;
    for (int l = 0; l < 3; ++l) tiles += ((sys->lc[l] + 15) / 16) * ((sys->lr[l] + 15) / 16);
    if (tiles > 65535) return set_err(FU_ERR_ARG, "frame %dx%d too large for batched launches", sys->p.cols, sys->p.rows);
    int slot = -1;
    for (int i = 0; i < bt->cap && slot < 0; ++i)
        if (!bt->owner[i]) slot = i;
    if (slot < 0) return set_err(FU_ERR_STATE, "batch context full (%d systems)", bt->cap);
    fill_desc(sys, bt->h_desc[slot]);
    HIPCHK(hipMemcpy(bt->d_desc + slot, &bt->h_desc[slot], sizeof(SessDesc), hipMemcpyHostToDevice));
    bt->owner[slot] = sys;
    bt->drop_graphs();  // a slot's session (and its launch shares) changed
    bt->fmt = sys->p.map.voxel_format;
    bt->rgb = sys->p.use_rgb ? 1 : 0;
    bt->attached++;
    sys->batch = bt;
    sys->batch_slot = slot;
    return FU_OK;
}

extern "C" fu_status fu_batch_detach(fu_batch* bt, fu_system* sys) {
    if (!bt || !sys) return set_err(FU_ERR_ARG, "null argument");
    if (sys->batch != bt) return set_err(FU_ERR_STATE, "system is not attached to this batch context");
    bt->owner[sys->batch_slot] = nullptr;
    bt->drop_graphs();
    if (--bt->attached == 0) bt->fmt = bt->rgb = -1;
    sys->batch = nullptr;
    sys->batch_slot = -1;
    return FU_OK;
}

// A batched step on `hs`, in two parts.  The input part (model-map resize, filter, pyramid,
// normals with the packed integration inputs, intensities) reads the frames; the rest
// (tracking, fusion, raycast, the result gather) reads only what the input part made, so a
// captured step is the rest alone and the frames can sit anywhere.  With `prof`, the stage
// events ev[] are recorded between the stages.
static void launch_inputs(BatchLaunch& b, const SessDesc* h, bool track, int use_rgb, const void* const* depth,
                          const void* const* rgb, bool prof, hipEvent_t* ev, hipStream_t hs) {
    const int n = b.n;
    if (prof) (void)hipEventRecord(ev[0], hs);
    for (int i = 0; i < n; ++i) b.in[i] = depth[i];
    if (track) launch_b_resize(b, h, hs);
    if (prof) (void)hipEventRecord(ev[FU_STAGE_RESIZE + 1], hs);
    launch_b_filter(b, h, hs);
    launch_b_pyrdown(b, h, 1, hs);
    launch_b_pyrdown(b, h, 2, hs);
    for (int i = 0; i < n; ++i) b.in[i] = rgb[i];  // the normal pass packs the integration inputs
    launch_b_nmap(b, h, hs);
    if (use_rgb) launch_b_images(b, h, hs);
    if (prof) (void)hipEventRecord(ev[FU_STAGE_PYRAMID + 1], hs);
    for (int i = 0; i < n; ++i) b.in[i] = nullptr;  // the frames are not read after this part
}
static void launch_rest(BatchLaunch& b, const SessDesc* h, const int maxit[3], bool track, int use_rgb, int fmt,
                        FrameBrief* brief, bool prof, hipEvent_t* ev, hipStream_t hs) {
    auto mark = [&](int k) {
        if (prof) (void)hipEventRecord(ev[k], hs);
    };
    for (int i = 0; i < b.n; ++i) b.in[i] = nullptr;
    if (track) {
        launch_b_track_begin(b, h, hs);
        if (use_rgb) launch_b_rgb_lists(b, h, hs);
        for (int l = 2; l >= 0; --l)
            for (int j = 0; j < maxit[l]; ++j) launch_b_icp(b, h, l, j, hs);
    }
    launch_b_track_commit(b, h, hs);
    if (use_rgb) launch_b_image_commit(b, h, hs);
    mark(FU_STAGE_ICP + 1);
    launch_b_alloc(b, h, hs);
    mark(FU_STAGE_ALLOC + 1);
    launch_b_visible(b, h, hs);
    mark(FU_STAGE_VISIBLE + 1);
    launch_b_integrate(b, h, fmt, hs);
    mark(FU_STAGE_INTEGRATE + 1);
    launch_b_zrange(b, h, hs);
    mark(FU_STAGE_ZRANGE + 1);
    launch_b_raycast(b, h, fmt, hs);
    mark(FU_STAGE_RAYCAST + 1);
    if (brief) launch_b_results(b, brief, hs);
}

// The captured rest of a step: looked up by the step's key (sessions in slot order, result
// buffer, membership generation); a key met for the second time is captured, so one-off
// memberships of a served lane run as plain launches.  Null when the rest is to be launched
// directly.
static fu_status step_graph(fu_batch* bt, BatchLaunch& b, const SessDesc* h, const int maxit[3], FrameBrief* brief,
                            hipGraphExec_t* out) {
    *out = nullptr;
    if (bt->tuning_gen != g_tuning_gen) {
        bt->drop_graphs();
        bt->tuning_gen = g_tuning_gen;
    }
    std::vector<uint64_t> key;
    key.reserve(3 + (b.n + 3) / 4);
    key.push_back((uint64_t)b.n);
    key.push_back((uint64_t)(uintptr_t)brief);
    key.push_back(bt->gen);
    for (int i = 0; i < b.n; i += 4) {
        uint64_t w = 0;
        for (int k = 0; k < 4 && i + k < b.n; ++k) w |= (uint64_t)b.sess[i + k] << (16 * k);
        key.push_back(w);
    }
    ++bt->clock;
    for (fu_batch::Graph& g : bt->graphs)
        if (g.key == key) {
            g.used = bt->clock;
            *out = g.exec;
            return FU_OK;
        }
    auto seen = std::find(bt->seen.begin(), bt->seen.end(), key);
    if (seen == bt->seen.end()) {
        if (bt->seen.size() >= fu_batch::kSeen) bt->seen.erase(bt->seen.begin());
        bt->seen.push_back(std::move(key));
        return FU_OK;
    }
    bt->seen.erase(seen);
    if (!bt->cap_stream) HIPCHK(hipStreamCreateWithFlags(&bt->cap_stream, hipStreamNonBlocking));
    hipGraph_t g;
    hipError_t e = hipStreamBeginCapture(bt->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return set_err(FU_ERR_HIP, "batched step capture: %s", hipGetErrorString(e));
    launch_rest(b, h, maxit, true, bt->rgb == 1, bt->fmt, brief, false, nullptr, bt->cap_stream);
    e = hipStreamEndCapture(bt->cap_stream, &g);
    if (e != hipSuccess) return set_err(FU_ERR_HIP, "batched step capture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_err(FU_ERR_HIP, "batched step instantiate: %s", hipGetErrorString(e));
    if (bt->graphs.size() >= fu_batch::kGraphs) {  // evict the least recently used
        auto lru = std::min_element(bt->graphs.begin(), bt->graphs.end(),
                                    [](const fu_batch::Graph& x, const fu_batch::Graph& y) { return x.used < y.used; });
        (void)hipGraphExecDestroy(lru->exec);
        bt->graphs.erase(lru);
    }
    bt->graphs.push_back(fu_batch::Graph{key, exec, bt->clock});
    bt->captures++;
    *out = exec;
    return FU_OK;
}

// One frame of each of `n` systems as one sequence of batched launches: the same stages as
// queue_frame_device, each launched once for all sessions.  With `brief` set, the frames'
// results are gathered into brief[0..n).
fu_status fu::batch_grab(fu_batch* bt, int n, fu_system* const* sys, const uint16_t* const* depth,
                         const uint8_t* const* rgb, FrameBrief* brief, hipStream_t s, hipEvent_t inputs_read) {
    if (!bt || !sys || !depth || !rgb) return set_err(FU_ERR_ARG, "fu_batch_grab: null argument");
    if (n < 1 || n > kMaxBatch) return set_err(FU_ERR_ARG, "fu_batch_grab: %d frames (1..%d)", n, kMaxBatch);
    BatchLaunch b;
    memset(&b, 0, sizeof(b));
    b.desc = bt->d_desc;
    b.n = n;
    int maxit[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        fu_system* x = sys[i];
        if (!x || x->batch != bt) return set_err(FU_ERR_STATE, "fu_batch_grab: system %d not attached to this context", i);
        if (!depth[i] || !rgb[i]) return set_err(FU_ERR_ARG, "fu_batch_grab: null frame %d", i);
        for (int j = 0; j < i; ++j)
            if (sys[j] == x) return set_err(FU_ERR_ARG, "fu_batch_grab: system %d appears twice", i);
        b.sess[i] = (unsigned short)x->batch_slot;
        if (x->frame == 0) {
            b.first |= 1ull << i;
        } else {
            for (int l = 0; l < 3; ++l) maxit[l] = std::max(maxit[l], x->p.iters[l]);
        }
    }
    const bool track = b.first != (n == 64 ? ~0ull : ((1ull << n) - 1));
    const SessDesc* h = bt->h_desc.data();
    hipStream_t hs = s;
    const bool prof = bt->prof;
    const void* dp[kMaxBatch];
    const void* cp[kMaxBatch];
    for (int i = 0; i < n; ++i) {
        dp[i] = depth[i];
        cp[i] = rgb[i];
    }
    hipGraphExec_t exec = nullptr;
    if (bt->use_graph && !prof && b.first == 0) {
        fu_status st = step_graph(bt, b, h, maxit, brief, &exec);
        if (st) return st;
    }
    launch_inputs(b, h, track, bt->rgb == 1, dp, cp, prof, bt->ev, hs);
    // the frames are read: the rest uses the pyramid and the packed integration inputs
    if (inputs_read) HIPCHK(hipEventRecord(inputs_read, hs));
    if (exec) {
        HIPCHK(hipGraphLaunch(exec, hs));
        bt->replays++;
    } else {
        launch_rest(b, h, maxit, track, bt->rgb == 1, bt->fmt, brief, prof, bt->ev, hs);
    }
    LAUNCHCHK();
    for (int i = 0; i < n; ++i) sys[i]->frame++;
    if (prof) {
        // profiling steps synchronise: stage times, visible blocks and ICP work of the step
        HIPCHK(hipEventSynchronize(bt->ev[FU_NUM_STAGES]));
        for (int k = 0; k < FU_NUM_STAGES; ++k) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, bt->ev[k], bt->ev[k + 1]));
            bt->times.ms[k] += ms;
        }
        bt->times.steps += 1;
        for (int i = 0; i < n; ++i) {
            fu_system* x = sys[i];
            unsigned c[C_COUNT];
            TrackState t;
            HIPCHK(hipMemcpy(c, x->map->v.ctr, sizeof(c), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&t, x->trk->st, sizeof(t), hipMemcpyDeviceToHost));
            bt->times.frames += 1;
            bt->times.n_visible += c[C_VISIBLE];
            bt->times.n_new += c[C_NEW];
            bt->times.icp_iterations += t.iters_done;
            int left = t.iters_done;
            for (int l = 2; l >= 0 && left > 0; --l) {
                const int it = std::min(left, x->p.iters[l]);
                bt->times.icp_pixels += (double)it * x->lc[l] * x->lr[l];
                left -= it;
            }
        }
    }
    return FU_OK;
}


extern "C" fu_status fu_batch_grab(fu_batch* bt, int n, fu_system* const* sys, const uint16_t* const* depth,
                                   const uint8_t* const* rgb, fu_stream s) {
    return batch_grab(bt, n, sys, depth, rgb, nullptr, S(s));
}

fu_status fu::mesh_async(fu_system* sys, float* vertex, float* normal, uint8_t* color, uint32_t max_triangles,
                         hipStream_t s) {
    if (!sys || !vertex || !normal || !color) return set_err(FU_ERR_ARG, "mesh request: null argument");
    MeshOut o;
    o.vertex = vertex;
    o.normal = normal;
    o.color = color;
    o.tri_meta = nullptr;
    o.max_tri = max_triangles;
    launch_mesh(sys->map->v, o, sys->map->mesh_counts, sys->map->mesh_chunk, s);
    LAUNCHCHK();
    return FU_OK;
}

const unsigned* fu::system_counters(fu_system* sys) { return sys->map->v.ctr; }

// ------------------------------------------------------------------------------ batched ICP
struct fu_icp_batch {
    int cap = 0;
    SessDesc* d_desc = nullptr;
    std::vector<SessDesc> h_desc;
};

extern "C" fu_status fu_icp_batch_create(int capacity, fu_icp_batch** out) {
    if (!out) return set_err(FU_ERR_ARG, "null output");
    *out = nullptr;
    if (capacity < 1 || capacity > 65535) return set_err(FU_ERR_ARG, "capacity %d out of range", capacity);
    fu_icp_batch* b = new fu_icp_batch();
    b->cap = capacity;
    b->h_desc.resize(capacity);
    if (hipMalloc(&b->d_desc, sizeof(SessDesc) * capacity) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return set_err(FU_ERR_NOMEM, "descriptor allocation failed");
    }
    *out = b;
    return FU_OK;
}

extern "C" fu_status fu_icp_batch_destroy(fu_icp_batch* b) {
    if (!b) return FU_OK;
    (void)hipFree(b->d_desc);
    delete b;
    return FU_OK;
}

extern "C" fu_status fu_compute_se3_batch(fu_icp_batch* ib, int n, fu_tracker* const* t,
                                          const float* const* next_vmap, const float* const* next_nmap,
                                          const float* const* last_vmap, const float* const* last_nmap,
                                          const float* K, const int32_t iters[3], const int32_t min_count[3],
                                          float thresh, fu_stream s) {
    if (!ib || !t || !next_vmap || !next_nmap || !last_vmap || !last_nmap || !K || !iters || !min_count)
        return set_err(FU_ERR_ARG, "fu_compute_se3_batch: null argument");
    if (n < 1 || n > ib->cap) return set_err(FU_ERR_ARG, "fu_compute_se3_batch: %d pairs (capacity %d)", n, ib->cap);
    for (int l = 0; l < 3; ++l)
        if (iters[l] < 0) return set_err(FU_ERR_ARG, "fu_compute_se3_batch: negative iterations");
    int last_level = -1;
    for (int l = 2; l >= 0; --l)
        if (iters[l] > 0) last_level = l;
    for (int i = 0; i < n; ++i) {
        if (!t[i]) return set_err(FU_ERR_ARG, "fu_compute_se3_batch: null tracker %d", i);
        for (int j = 0; j < i; ++j)
            if (t[j] == t[i]) return set_err(FU_ERR_ARG, "fu_compute_se3_batch: tracker %d appears twice", i);
        SessDesc& d = ib->h_desc[i];
        memset(&d, 0, sizeof(d));
        d.st = t[i]->st;
        d.partials = t[i]->partials;
        d.cols = t[i]->cols;
        d.rows = t[i]->rows;
        d.last_level = last_level;
        for (int l = 0; l < 3; ++l) {
            const int k = i * 3 + l;
            if (!next_vmap[k] || !next_nmap[k] || !last_vmap[k] || !last_nmap[k])
                return set_err(FU_ERR_ARG, "fu_compute_se3_batch: pair %d level %d: null map", i, l);
            d.lc[l] = t[i]->cols >> l;
            d.lr[l] = t[i]->rows >> l;
            d.icp[l] = make_icp(t[i], l, next_vmap[k], next_nmap[k], last_vmap[k], last_nmap[k], K + 4 * i, 1,
                                min_count[l], 0, thresh);
            d.icp_blocks[l] = t[i]->ngroups[l];
            d.iters[l] = iters[l];
        }
    }
    hipStream_t hs = S(s);
    // the descriptors of this call; copied before any kernel of the call reads them
    HIPCHK(hipMemcpyAsync(ib->d_desc, ib->h_desc.data(), sizeof(SessDesc) * n, hipMemcpyHostToDevice, hs));
    HIPCHK(hipStreamSynchronize(hs));
    for (int g0 = 0; g0 < n; g0 += kMaxBatch) {
        BatchLaunch b;
        memset(&b, 0, sizeof(b));
        b.desc = ib->d_desc;
        b.n = std::min(kMaxBatch, n - g0);
        for (int i = 0; i < b.n; ++i) b.sess[i] = (unsigned short)(g0 + i);
        const SessDesc* h = ib->h_desc.data();
        launch_b_track_begin(b, h, hs);
        for (int l = 2; l >= 0; --l)
            for (int j = 0; j < iters[l]; ++j) launch_b_icp(b, h, l, j, hs);
    }
    LAUNCHCHK();
    return FU_OK;
}
