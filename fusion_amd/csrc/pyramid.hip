// Depth pyramid, bilateral filter, vertex/normal maps and map downsampling.
//
// Kernels follow Tracking/Pyrdown.cu (FilterDepthKernel :5-41, PyrDownGaussKernel
// :134-158, ComputeVMapKernel :56-78, ComputeNMapKernel :93-114, ResizeMapKernel :251-279)
// with the same window bounds, weights and IEEE operation order.  Images are dense
// row-major (pitch = width).  All of these are HBM-bound stencils; a 16x16 work-group
// covers a 16x16 output tile so each wave reads whole 64-byte row segments.
#include "fu_device.h"
#include "fu_kernels.h"
#include "fu_batch.h"

namespace fu {

// Pyrdown.cu:3-4: sigSpace = 0.5/(4*4), sigRange = 0.5/(0.5*0.5) (exact in float)
constexpr float kSigSpace = 0.5f / 16.0f;
constexpr float kSigRange = 2.0f;

__device__ __forceinline__ void filter_depth_body(WG g, const uint16_t* __restrict__ depth_, float* __restrict__ raw_,
                                                  float* __restrict__ filt_, int cols, int rows, float scale_inv,
                                                  float cutoff) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const auto depth = FU_G(const uint16_t, depth_);
    const auto raw = FU_G(float, raw_);
    const auto filt = FU_G(float, filt_);
    const float center = (float)depth[y * cols + x] * scale_inv;
    raw[y * cols + x] = center < cutoff ? center : fnan();
    // the window is [max(p-2,0), min(p+3, n-1)) in each axis: the last row/column of the
    // image never contributes (Pyrdown.cu:23-29)
    const int tx = min(x + 3, cols - 1);
    const int ty = min(y + 3, rows - 1);
    float sum1 = 0.f, sum2 = 0.f;
    for (int cy = max(y - 2, 0); cy < ty; ++cy) {
        for (int cx = max(x - 2, 0); cx < tx; ++cx) {
            const float val = (float)depth[cy * cols + cx] * scale_inv;
            const float space2 = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
            const float color2 = (center - val) * (center - val);
            const float weight = fu_expf(-(space2 * kSigSpace + color2 * kSigRange));
            sum1 += val * weight;
            sum2 += weight;
        }
    }
    const float fin = sum1 / sum2;
    filt[y * cols + x] = fin < cutoff ? fin : fnan();
}
__global__ __launch_bounds__(256) void k_filter_depth(const uint16_t* __restrict__ depth, float* __restrict__ raw,
                                                     float* __restrict__ filt, int cols, int rows, float scale_inv,
                                                     float cutoff) {
    filter_depth_body(hw_wg(), depth, raw, filt, cols, rows, scale_inv, cutoff);
}

// binomial [1 4 6 4 1] outer product, indexed (ty-cy-1)*5 + (tx-cx-1) as in Pyrdown.cu:151,
// which mirrors the weights at the right/bottom border
__constant__ float c_gauss[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};

__device__ __forceinline__ void pyrdown_body(WG g, const float* __restrict__ src_, int scols, int srows,
                                             float* __restrict__ dst_, int dcols, int drows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= dcols || y >= drows) return;
    const auto src = FU_G(const float, src_);
    const auto dst = FU_G(float, dst_);
    const int tx = min(2 * x + 3, scols - 1);
    const int ty = min(2 * y + 3, srows - 1);
    float sum = 0.f;
    int count = 0;
    for (int cy = max(0, 2 * y - 2); cy < ty; ++cy) {
        for (int cx = max(0, 2 * x - 2); cx < tx; ++cx) {
            const float v = src[cy * scols + cx];
            if (!isnanf_(v)) {
                const float g = c_gauss[(ty - cy - 1) * 5 + (tx - cx - 1)];
                sum += v * g;
                count += (int)g;
            }
        }
    }
    dst[y * dcols + x] = sum / (float)count;
}
__global__ __launch_bounds__(256) void k_pyrdown(const float* __restrict__ src, int scols, int srows,
                                                float* __restrict__ dst, int dcols, int drows) {
    pyrdown_body(hw_wg(), src, scols, srows, dst, dcols, drows);
}

__device__ __forceinline__ void vmap_body(WG g, const float* __restrict__ depth_, f4* __restrict__ vmap_, int cols,
                                          int rows, float invfx, float invfy, float cx, float cy, float cutoff) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const auto depth = FU_G(const float, depth_);
    const auto vmap = FU_G(f4, vmap_);
    const float z = depth[y * cols + x];
    f4 v;
    // `z > 0.1` compares in double (Pyrdown.cu:69)
    if (!isnanf_(z) && (double)z > 0.1 && z < cutoff) {
        v.x = z * ((float)x - cx) * invfx;
        v.y = z * ((float)y - cy) * invfy;
        v.z = z;
        v.w = 1.0f;
    } else {
        // the reference sets only x = NaN; every consumer tests x alone
        v = mk4(fnan(), fnan(), fnan(), fnan());
    }
    vmap[y * cols + x] = v;
}
__global__ __launch_bounds__(256) void k_vmap(const float* __restrict__ depth, f4* __restrict__ vmap, int cols,
                                             int rows, float invfx, float invfy, float cx, float cy, float cutoff) {
    vmap_body(hw_wg(), depth, vmap, cols, rows, invfx, invfy, cx, cy, cutoff);
}

__device__ __forceinline__ void nmap_body(WG g, const f4* __restrict__ vmap_, f4* __restrict__ nmap_, int cols, int rows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const auto vmap = FU_G(const f4, vmap_);
    const auto nmap = FU_G(f4, nmap_);
    const f4 nan4 = mk4(fnan(), fnan(), fnan(), fnan());
    if (x == cols - 1 || y == rows - 1) {
        nmap[y * cols + x] = nan4;
        return;
    }
    const f4 c = vmap[y * cols + x];
    const f4 r = vmap[y * cols + x + 1];
    const f4 d = vmap[(y + 1) * cols + x];
    if (!isnanf_(c.x) && !isnanf_(r.x) && !isnanf_(d.x)) {
        const f3 n = normalised(cross(sub(xyz(r), xyz(c)), sub(xyz(d), xyz(c))));
        nmap[y * cols + x] = mk4(n.x, n.y, n.z, 1.0f);
    } else {
        nmap[y * cols + x] = nan4;
    }
}
__global__ __launch_bounds__(256) void k_nmap(const f4* __restrict__ vmap, f4* __restrict__ nmap, int cols, int rows) {
    nmap_body(hw_wg(), vmap, nmap, cols, rows);
}

// 2x2 mean of the vertex map, 2x2 mean of the normal map normalised in 4-D (w included,
// VectorMath.h:340-342), NaN if any of the four is NaN
__device__ __forceinline__ void resize_body(WG g, const f4* __restrict__ vsrc_, const f4* __restrict__ nsrc_, int scols,
                                            f4* __restrict__ vdst_, f4* __restrict__ ndst_, int dcols, int drows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= dcols || y >= drows) return;
    const auto vsrc = FU_G(const f4, vsrc_);
    const auto nsrc = FU_G(const f4, nsrc_);
    const auto vdst = FU_G(f4, vdst_);
    const auto ndst = FU_G(f4, ndst_);
    const int s0 = (2 * y) * scols + 2 * x, s1 = (2 * y + 1) * scols + 2 * x;
    const f4 v00 = vsrc[s0], v01 = vsrc[s0 + 1], v10 = vsrc[s1], v11 = vsrc[s1 + 1];
    const f4 n00 = nsrc[s0], n01 = nsrc[s0 + 1], n10 = nsrc[s1], n11 = nsrc[s1 + 1];
    const f4 nan4 = mk4(fnan(), fnan(), fnan(), fnan());
    if (isnanf_(v00.x) || isnanf_(v01.x) || isnanf_(v10.x) || isnanf_(v11.x))
        vdst[y * dcols + x] = nan4;
    else
        vdst[y * dcols + x] = divs4(add4(add4(add4(v00, v01), v10), v11), 4.0f);
    if (isnanf_(n00.x) || isnanf_(n01.x) || isnanf_(n10.x) || isnanf_(n11.x))
        ndst[y * dcols + x] = nan4;
    else
        ndst[y * dcols + x] = normalised4(divs4(add4(add4(add4(n00, n01), n10), n11), 4.0f));
}
__global__ __launch_bounds__(256) void k_resize(const f4* __restrict__ vsrc, const f4* __restrict__ nsrc, int scols,
                                               f4* __restrict__ vdst, f4* __restrict__ ndst, int dcols, int drows) {
    resize_body(hw_wg(), vsrc, nsrc, scols, vdst, ndst, dcols, drows);
}

static inline dim3 grid16(int cols, int rows) { return dim3((cols + 15) / 16, (rows + 15) / 16); }

void launch_filter_depth(const uint16_t* depth, float* raw, float* filt, int cols, int rows, float scale_inv,
                         float cutoff, hipStream_t s) {
    hipLaunchKernelGGL(k_filter_depth, grid16(cols, rows), dim3(256), 0, s, depth, raw, filt, cols, rows, scale_inv,
                       cutoff);
}
void launch_pyrdown(const float* src, int scols, int srows, float* dst, int dcols, int drows, hipStream_t s) {
    hipLaunchKernelGGL(k_pyrdown, grid16(dcols, drows), dim3(256), 0, s, src, scols, srows, dst, dcols, drows);
}
void launch_vmap(const float* depth, f4* vmap, int cols, int rows, float invfx, float invfy, float cx, float cy,
                 float cutoff, hipStream_t s) {
    hipLaunchKernelGGL(k_vmap, grid16(cols, rows), dim3(256), 0, s, depth, vmap, cols, rows, invfx, invfy, cx, cy,
                       cutoff);
}
void launch_nmap(const f4* vmap, f4* nmap, int cols, int rows, hipStream_t s) {
    hipLaunchKernelGGL(k_nmap, grid16(cols, rows), dim3(256), 0, s, vmap, nmap, cols, rows);
}
void launch_resize(const f4* vsrc, const f4* nsrc, int scols, f4* vdst, f4* ndst, int dcols, int drows,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_resize, grid16(dcols, drows), dim3(256), 0, s, vsrc, nsrc, scols, vdst, ndst, dcols, drows);
}

// ------------------------------------------------------------------ batched launches
// Each work-group finds its session's share (batch_slot) and runs the body the single-session
// kernel runs, with the same work-group ids.

__global__ __launch_bounds__(256) void kb_resize(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    if (slot_first(b, s)) return;  // the first frame has no model maps
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = b.level;
    resize_body(g, d.model_v[l - 1], d.model_n[l - 1], d.lc[l - 1], d.model_v[l], d.model_n[l], d.lc[l], d.lr[l]);
}

__global__ __launch_bounds__(256) void kb_filter(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    filter_depth_body(g, (const uint16_t*)b.in[s], d.raw, d.depth[0], d.cols, d.rows, d.depth_scale_inv, d.cutoff);
}

__global__ __launch_bounds__(256) void kb_pyrdown(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = b.level;
    pyrdown_body(g, d.depth[l - 1], d.lc[l - 1], d.lr[l - 1], d.depth[l], d.lc[l], d.lr[l]);
}

// one launch over the tiles of all three levels; `local` -> (level, tile)
__device__ __forceinline__ int pyr3_level(const SessDesc& d, int local, WG& g) {
    int l = 0;
    while (l < 2) {
        const int tx = (d.lc[l] + 15) / 16, t = tx * ((d.lr[l] + 15) / 16);
        if (local < t) break;
        local -= t;
        ++l;
    }
    const int tx = (d.lc[l] + 15) / 16;
    g.x = local % tx;
    g.y = local / tx;
    g.nx = tx;
    g.ny = (d.lr[l] + 15) / 16;
    return l;
}

__global__ __launch_bounds__(256) void kb_vmap(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = pyr3_level(d, g.x, g);
    vmap_body(g, d.depth[l], d.vmap[l], d.lc[l], d.lr[l], d.invf[l][0], d.invf[l][1], d.Kl[l][2], d.Kl[l][3], d.cutoff);
}

__global__ __launch_bounds__(256) void kb_nmap(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = pyr3_level(d, g.x, g);
    nmap_body(g, d.vmap[l], d.nmap[l], d.lc[l], d.lr[l]);
}

static void blaunch(void (*k)(BatchLaunch), int grid, const BatchLaunch& b, hipStream_t s) {
    if (grid > 0) hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, s, b);
}
void launch_b_resize(BatchLaunch& b, const SessDesc* h, int level, hipStream_t s) {
    b.level = level;
    blaunch(kb_resize, batch_layout(b, h, BT_TILES, level, 0), b, s);
}
void launch_b_filter(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    b.level = 0;
    blaunch(kb_filter, batch_layout(b, h, BT_TILES, 0, 0), b, s);
}
void launch_b_pyrdown(BatchLaunch& b, const SessDesc* h, int level, hipStream_t s) {
    b.level = level;
    blaunch(kb_pyrdown, batch_layout(b, h, BT_TILES, level, 0), b, s);
}
void launch_b_vmap(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    blaunch(kb_vmap, batch_layout(b, h, BT_PYR3, 0, 0), b, s);
}
void launch_b_nmap(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    blaunch(kb_nmap, batch_layout(b, h, BT_PYR3, 0, 0), b, s);
}

}  // namespace fu
