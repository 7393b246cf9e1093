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

// ComputeVMapKernel's vertex of pixel (x, y) with depth z (Pyrdown.cu:56-78)
__device__ __forceinline__ f4 vertex_of(float z, int x, int y, float invfx, float invfy, float cx, float cy,
                                        float cutoff) {
    // `z > 0.1` compares in double (Pyrdown.cu:69)
    if (!isnanf_(z) && (double)z > 0.1 && z < cutoff)
        return mk4(z * ((float)x - cx) * invfx, z * ((float)y - cy) * invfy, z, 1.0f);
    // the reference sets only x = NaN; every consumer tests x alone
    return mk4(fnan(), fnan(), fnan(), fnan());
}
// where a depth kernel also writes the vertex map of its level (null vmap: it does not)
struct VertexOut {
    f4* vmap;
    float invfx, invfy, cx, cy;
};

__device__ __forceinline__ void filter_depth_body(WG g, const uint16_t* __restrict__ depth_, float* __restrict__ raw_,
                                                  float* __restrict__ filt_, int cols, int rows, float scale_inv,
                                                  float cutoff, VertexOut vo = VertexOut{}) {
    // the group's 16x16 pixels and their 2-pixel margin, converted to metres once per sample
    // ((float)d * scale_inv, the value every tap computes) and kept in LDS; samples outside
    // the image are never read (the window bounds below stay inside it)
    constexpr int T = 20;
    __shared__ float tile[T * T];
    const int x0 = g.x * 16 - 2, y0 = g.y * 16 - 2;
    const auto depth = FU_G(const uint16_t, depth_);
    for (int i = threadIdx.x; i < T * T; i += 256) {
        const int gx = x0 + i % T, gy = y0 + i / T;
        if (gx >= 0 && gy >= 0 && gx < cols && gy < rows) tile[i] = (float)depth[gy * cols + gx] * scale_inv;
    }
    __syncthreads();
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const auto raw = FU_G(float, raw_);
    const auto filt = FU_G(float, filt_);
    const float center = tile[(y - y0) * T + (x - x0)];
    raw[y * cols + x] = center < cutoff ? center : fnan();
    // the window is [max(p-2,0), min(p+3, n-1)) in each axis: the last row/column of the
    // image never contributes (Pyrdown.cu:23-29)
    const int tx = min(x + 3, cols - 1);
    const int ty = min(y + 3, rows - 1);
    float sum1 = 0.f, sum2 = 0.f;
    for (int cy = max(y - 2, 0); cy < ty; ++cy) {
        for (int cx = max(x - 2, 0); cx < tx; ++cx) {
            const float val = tile[(cy - y0) * T + (cx - x0)];
            const float space2 = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
            const float color2 = (center - val) * (center - val);
            // the argument is never positive or NaN: depths are finite
            const float weight = fu_expf_nonpos(-(space2 * kSigSpace + color2 * kSigRange));
            sum1 += val * weight;
            sum2 += weight;
        }
    }
    const float fin = sum1 / sum2;
    const float z = fin < cutoff ? fin : fnan();
    filt[y * cols + x] = z;
    if (vo.vmap) FU_G(f4, vo.vmap)[y * cols + x] = vertex_of(z, x, y, vo.invfx, vo.invfy, vo.cx, vo.cy, cutoff);
}
__global__ __launch_bounds__(256) void k_filter_depth(const uint16_t* __restrict__ depth, float* __restrict__ raw,
                                                     float* __restrict__ filt, int cols, int rows, float scale_inv,
                                                     float cutoff) {
    filter_depth_body(hw_wg(), depth, raw, filt, cols, rows, scale_inv, cutoff);
}

// binomial [1 4 6 4 1] outer product, indexed (ty-cy-1)*5 + (tx-cx-1) as in Pyrdown.cu:151,
// which mirrors the weights at the right/bottom border
__constant__ float c_gauss[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};
constexpr float kGauss[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};

__device__ __forceinline__ void pyrdown_body(WG g, const float* __restrict__ src_, int scols, int srows,
                                             float* __restrict__ dst_, int dcols, int drows,
                                             VertexOut vo = VertexOut{}, float cutoff = 0.f) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= dcols || y >= drows) return;
    const auto src = FU_G(const float, src_);
    const auto dst = FU_G(float, dst_);
    const int tx = min(2 * x + 3, scols - 1);
    const int ty = min(2 * y + 3, srows - 1);
    float sum = 0.f;
    int count = 0;
    if (x >= 1 && y >= 1 && 2 * x + 3 < scols && 2 * y + 3 < srows) {
        // the whole 5x5 window: its loads issue together; NaN samples are skipped as below
        const auto p = src + (2 * y - 2) * scols + (2 * x - 2);
        float v[25];
#pragma unroll
        for (int q = 0; q < 25; ++q) v[q] = p[(q / 5) * scols + q % 5];
#pragma unroll
        for (int q = 0; q < 25; ++q) {
            const float g = kGauss[(4 - q / 5) * 5 + (4 - q % 5)];
            const bool ok = !isnanf_(v[q]);
            sum = ok ? sum + v[q] * g : sum;
            count += ok ? (int)g : 0;
        }
    } else {
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
    }
    const float z = sum / (float)count;
    dst[y * dcols + x] = z;
    if (vo.vmap) FU_G(f4, vo.vmap)[y * dcols + x] = vertex_of(z, x, y, vo.invfx, vo.invfy, vo.cx, vo.cy, cutoff);
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
    vmap[y * cols + x] = vertex_of(depth[y * cols + x], x, y, invfx, invfy, cx, cy, cutoff);
}
__global__ __launch_bounds__(256) void k_vmap(const float* __restrict__ depth, f4* __restrict__ vmap, int cols,
                                             int rows, float invfx, float invfy, float cx, float cy, float cutoff) {
    vmap_body(hw_wg(), depth, vmap, cols, rows, invfx, invfy, cx, cy, cutoff);
}

// ComputeNMapKernel's normal of pixel (x, y) inside the image
__device__ __forceinline__ f4 normal_of(const f4* __restrict__ vmap_, int x, int y, int cols, int rows) {
    const auto vmap = FU_G(const f4, vmap_);
    const f4 nan4 = mk4(fnan(), fnan(), fnan(), fnan());
    if (x == cols - 1 || y == rows - 1) return nan4;
    const f4 c = vmap[y * cols + x];
    const f4 r = vmap[y * cols + x + 1];
    const f4 d = vmap[(y + 1) * cols + x];
    if (!isnanf_(c.x) && !isnanf_(r.x) && !isnanf_(d.x)) {
        const f3 n = normalised(cross(sub(xyz(r), xyz(c)), sub(xyz(d), xyz(c))));
        return mk4(n.x, n.y, n.z, 1.0f);
    }
    return nan4;
}
__device__ __forceinline__ void nmap_body(WG g, const f4* __restrict__ vmap_, f4* __restrict__ nmap_, int cols, int rows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    FU_G(f4, nmap_)[y * cols + x] = normal_of(vmap_, x, y, cols, rows);
}
__global__ __launch_bounds__(256) void k_nmap(const f4* __restrict__ vmap, f4* __restrict__ nmap, int cols, int rows) {
    nmap_body(hw_wg(), vmap, nmap, cols, rows);
}

// 2x2 mean of the vertex map, 2x2 mean of the normal map normalised in 4-D (w included,
// VectorMath.h:340-342), NaN if any of the four is NaN; sources in the order (2y, 2x),
// (2y, 2x+1), (2y+1, 2x), (2y+1, 2x+1)
__device__ __forceinline__ void resize_px(f4 v00, f4 v01, f4 v10, f4 v11, f4 n00, f4 n01, f4 n10, f4 n11, f4& vo,
                                          f4& no) {
    const f4 nan4 = mk4(fnan(), fnan(), fnan(), fnan());
    if (isnanf_(v00.x) || isnanf_(v01.x) || isnanf_(v10.x) || isnanf_(v11.x))
        vo = nan4;
    else
        vo = divs4(add4(add4(add4(v00, v01), v10), v11), 4.0f);
    if (isnanf_(n00.x) || isnanf_(n01.x) || isnanf_(n10.x) || isnanf_(n11.x))
        no = nan4;
    else
        no = normalised4(divs4(add4(add4(add4(n00, n01), n10), n11), 4.0f));
}
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
    f4 vo, no;
    resize_px(vsrc[s0], vsrc[s0 + 1], vsrc[s1], vsrc[s1 + 1], nsrc[s0], nsrc[s0 + 1], nsrc[s1], nsrc[s1 + 1], vo, no);
    vdst[y * dcols + x] = vo;
    ndst[y * dcols + x] = no;
}
__global__ __launch_bounds__(256) void k_resize(const f4* __restrict__ vsrc, const f4* __restrict__ nsrc, int scols,
                                               f4* __restrict__ vdst, f4* __restrict__ ndst, int dcols, int drows) {
    resize_body(hw_wg(), vsrc, nsrc, scols, vdst, ndst, dcols, drows);
}

// ------------------------------------------------------------------ intensity pyramid
// ImageToIntensityKernel (Pyrdown.cu:183-194): (int)(0.2125 c1 + 0.7154 c0 + 0.0721 c2) in
// double, evaluated as written (no fused multiply-add), stored as a byte
__device__ __forceinline__ void intensity_body(WG g, const unsigned char* __restrict__ rgb_,
                                               unsigned char* __restrict__ img_, int cols, int rows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const auto rgb = FU_G(const unsigned char, rgb_);
    const auto img = FU_G(unsigned char, img_);
    const int i = y * cols + x;
    const double v = (0.2125 * (double)rgb[3 * i + 1] + 0.7154 * (double)rgb[3 * i]) + 0.0721 * (double)rgb[3 * i + 2];
    img[i] = (unsigned char)(int)v;  // 0 <= v < 256
}
__global__ __launch_bounds__(256) void k_intensity(const unsigned char* __restrict__ rgb, unsigned char* __restrict__ img,
                                                  int cols, int rows) {
    intensity_body(hw_wg(), rgb, img, cols, rows);
}

// PyrDownGaussKernel<uchar, uchar> (Pyrdown.cu:133-158): the float kernel's window and
// mirrored weights over bytes (no byte is NaN), the float mean truncated to a byte.  Every
// product and partial sum is an integer below 2^24, so the float sums are exact.
__device__ __forceinline__ void pyrdown_u8_body(WG g, const unsigned char* __restrict__ src_, int scols, int srows,
                                                unsigned char* __restrict__ dst_, int dcols, int drows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= dcols || y >= drows) return;
    const auto src = FU_G(const unsigned char, src_);
    const auto dst = FU_G(unsigned char, dst_);
    const int tx = min(2 * x + 3, scols - 1);
    const int ty = min(2 * y + 3, srows - 1);
    float sum = 0.f;
    int count = 0;
    if (x >= 1 && y >= 1 && 2 * x + 3 < scols && 2 * y + 3 < srows) {
        // the whole 5x5 window: its loads issue together, the same products and sums in the
        // same order as the clipped loop below
        const auto p = src + (2 * y - 2) * scols + (2 * x - 2);
        float v[25];
#pragma unroll
        for (int q = 0; q < 25; ++q) v[q] = (float)p[(q / 5) * scols + q % 5];
#pragma unroll
        for (int q = 0; q < 25; ++q) {
            const float g = kGauss[(4 - q / 5) * 5 + (4 - q % 5)];
            sum += v[q] * g;
            count += (int)g;
        }
    } else {
        for (int cy = max(0, 2 * y - 2); cy < ty; ++cy) {
            for (int cx = max(0, 2 * x - 2); cx < tx; ++cx) {
                const float g = c_gauss[(ty - cy - 1) * 5 + (tx - cx - 1)];
                sum += (float)src[cy * scols + cx] * g;
                count += (int)g;
            }
        }
    }
    dst[y * dcols + x] = (unsigned char)f2i(sum / (float)count);
}
__global__ __launch_bounds__(256) void k_pyrdown_u8(const unsigned char* __restrict__ src, int scols, int srows,
                                                   unsigned char* __restrict__ dst, int dcols, int drows) {
    pyrdown_u8_body(hw_wg(), src, scols, srows, dst, dcols, drows);
}

// ComputeDerivativeImageKernel (Pyrdown.cu:210-236): 3x3 Sobel over the window clipped to
// the image, the weight index counting down from 8 across the clipped window (at the border
// the weights shift), float sums stored as shorts
__constant__ float c_sobel_x[9] = {1, 0, -1, 2, 0, -2, 1, 0, -1};
__constant__ float c_sobel_y[9] = {1, 2, 1, 0, 0, 0, -1, -2, -1};
__device__ __forceinline__ void derivative_body(WG g, const unsigned char* __restrict__ img_, short* __restrict__ dx_,
                                                short* __restrict__ dy_, int cols, int rows) {
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= rows) return;
    const auto img = FU_G(const unsigned char, img_);
    float a = 0.f, b = 0.f;
    if (x >= 1 && y >= 1 && x + 1 < cols && y + 1 < rows) {
        // the whole 3x3 window: its nine loads issue together, the same products and sums in
        // the same order as the clipped loop below (k = 8 .. 0)
        constexpr float sx[9] = {1, 0, -1, 2, 0, -2, 1, 0, -1};
        constexpr float sy[9] = {1, 2, 1, 0, 0, 0, -1, -2, -1};
        const auto p = img + (y - 1) * cols + (x - 1);
        float v[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) v[q] = (float)p[(q / 3) * cols + q % 3];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            a += v[q] * sx[8 - q];
            b += v[q] * sy[8 - q];
        }
    } else {
        int k = 8;
        for (int j = max(y - 1, 0); j <= min(y + 1, rows - 1); ++j) {
            for (int i = max(x - 1, 0); i <= min(x + 1, cols - 1); ++i) {
                const float v = (float)img[j * cols + i];
                a += v * c_sobel_x[k];
                b += v * c_sobel_y[k];
                --k;
            }
        }
    }
    FU_G(short, dx_)[y * cols + x] = (short)f2i(a);
    FU_G(short, dy_)[y * cols + x] = (short)f2i(b);
}
__global__ __launch_bounds__(256) void k_derivative(const unsigned char* __restrict__ img, short* __restrict__ dx,
                                                   short* __restrict__ dy, int cols, int rows) {
    derivative_body(hw_wg(), img, dx, dy, cols, rows);
}

// Tracker::SwapFrame for the intensity pyramid: after a swapped frame (GATE_SWAP) its images
// become the last frame's.  16 bytes per lane where both ends are 16-byte aligned.
__device__ __forceinline__ void image_commit_body(int lane, const TrackState* st_, const unsigned char* src_,
                                                  unsigned char* dst_, int n) {
    const auto st = FU_G(const TrackState, st_);
    if (!(st->gates & GATE_SWAP)) return;
    const auto src = FU_G(const unsigned char, src_);
    const auto dst = FU_G(unsigned char, dst_);
    const int i = lane * 16;
    if (i + 16 <= n) {
        *(FU_GPTR(uint4))(dst + i) = *(FU_GPTR(const uint4))(src + i);
    } else {
        for (int k = i; k < n; ++k) dst[k] = src[k];
    }
}
struct ImageCommit {
    const TrackState* st;
    const unsigned char* src[3];
    unsigned char* dst[3];
    int n[3];
    int pre[4];  // prefix of the levels' 256-lane groups
};
__global__ __launch_bounds__(256) void k_image_commit(ImageCommit c) {
    int l = 0;
    while (l < 2 && (int)blockIdx.x >= c.pre[l + 1]) ++l;
    image_commit_body(((int)blockIdx.x - c.pre[l]) * 256 + threadIdx.x, c.st, c.src[l], c.dst[l], c.n[l]);
}

static inline dim3 grid16(int cols, int rows) { return dim3((cols + 15) / 16, (rows + 15) / 16); }

void launch_intensity(const unsigned char* rgb, unsigned char* img, int cols, int rows, hipStream_t s) {
    hipLaunchKernelGGL(k_intensity, grid16(cols, rows), dim3(256), 0, s, rgb, img, cols, rows);
}
void launch_pyrdown_u8(const unsigned char* src, int scols, int srows, unsigned char* dst, int dcols, int drows,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_pyrdown_u8, grid16(dcols, drows), dim3(256), 0, s, src, scols, srows, dst, dcols, drows);
}
void launch_derivative(const unsigned char* img, short* dx, short* dy, int cols, int rows, hipStream_t s) {
    hipLaunchKernelGGL(k_derivative, grid16(cols, rows), dim3(256), 0, s, img, dx, dy, cols, rows);
}
void launch_image_commit(const TrackState* st, const unsigned char* const src[3], unsigned char* const dst[3],
                         const int n[3], hipStream_t s) {
    ImageCommit c;
    c.st = st;
    c.pre[0] = 0;
    for (int l = 0; l < 3; ++l) {
        c.src[l] = src[l];
        c.dst[l] = dst[l];
        c.n[l] = n[l];
        c.pre[l + 1] = c.pre[l] + (n[l] + 16 * 256 - 1) / (16 * 256);
    }
    hipLaunchKernelGGL(k_image_commit, dim3(c.pre[3]), dim3(256), 0, s, c);
}

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

// levels 1 and 2 of the model maps in one pass over 16x16 tiles of level 1: each lane makes
// its level-1 pixel as resize_body does and keeps it in LDS; after the barrier the lanes of
// even (x, y) make the level-2 pixel of their 2x2 quad from those four values, the ones a
// level-2 resize would read back.  A level-2 pixel (x2, y2) exists only when 2*x2 + 1 and
// 2*y2 + 1 are level-1 pixels, and its quad lies in the tile.
__global__ __launch_bounds__(256) void kb_resize(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    if (slot_first(b, s)) return;  // uniform: the first frame has no model maps
    const SessDesc& d = b.desc[b.sess[s]];
    __shared__ f4 tv[16][16], tn[16][16];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int x = g.x * 16 + tx, y = g.y * 16 + ty;
    const int c0 = d.lc[0], c1 = d.lc[1];
    if (x < c1 && y < d.lr[1]) {
        const auto vsrc = FU_G(const f4, d.model_v[0]);
        const auto nsrc = FU_G(const f4, d.model_n[0]);
        const int s0 = (2 * y) * c0 + 2 * x, s1 = (2 * y + 1) * c0 + 2 * x;
        f4 vo, no;
        resize_px(vsrc[s0], vsrc[s0 + 1], vsrc[s1], vsrc[s1 + 1], nsrc[s0], nsrc[s0 + 1], nsrc[s1], nsrc[s1 + 1], vo,
                  no);
        FU_G(f4, d.model_v[1])[y * c1 + x] = vo;
        FU_G(f4, d.model_n[1])[y * c1 + x] = no;
        tv[ty][tx] = vo;
        tn[ty][tx] = no;
    }
    __syncthreads();
    const int x2 = x >> 1, y2 = y >> 1, c2 = d.lc[2];
    if (!(tx & 1) && !(ty & 1) && x2 < c2 && y2 < d.lr[2]) {
        f4 vo, no;
        resize_px(tv[ty][tx], tv[ty][tx + 1], tv[ty + 1][tx], tv[ty + 1][tx + 1], tn[ty][tx], tn[ty][tx + 1],
                  tn[ty + 1][tx], tn[ty + 1][tx + 1], vo, no);
        FU_G(f4, d.model_v[2])[y2 * c2 + x2] = vo;
        FU_G(f4, d.model_n[2])[y2 * c2 + x2] = no;
    }
}

__global__ __launch_bounds__(256) void kb_filter(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    filter_depth_body(g, (const uint16_t*)b.in[s], d.raw, d.depth[0], d.cols, d.rows, d.depth_scale_inv, d.cutoff,
                      VertexOut{d.vmap[0], d.invf[0][0], d.invf[0][1], d.Kl[0][2], d.Kl[0][3]});
}

__global__ __launch_bounds__(256) void kb_pyrdown(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = b.level;
    pyrdown_body(g, d.depth[l - 1], d.lc[l - 1], d.lr[l - 1], d.depth[l], d.lc[l], d.lr[l],
                 VertexOut{d.vmap[l], d.invf[l][0], d.invf[l][1], d.Kl[l][2], d.Kl[l][3]}, d.cutoff);
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

__global__ __launch_bounds__(256) void kb_nmap(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = pyr3_level(d, g.x, g);
    const int cols = d.lc[l];
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    if (x >= cols || y >= d.lr[l]) return;
    const int i = y * cols + x;
    const f4 n = normal_of(d.vmap[l], x, y, cols, d.lr[l]);
    FU_G(f4, d.nmap[l])[i] = n;
    if (l == 0) {
        // uniform: the integration inputs of the pixel (k_fuse_pack) from the normal above,
        // the level-0 depth and the step's colour frame (b.in)
        const auto c = FU_G(const unsigned char, (const unsigned char*)b.in[s]) + 3 * i;
        FU_G(f4, d.fuse.px)[i] = mk4(FU_G(const float, d.fuse.depth)[i], n.x, n.y, n.z);
        FU_G(unsigned, d.fuse.prgb)[i] = (unsigned)c[0] | ((unsigned)c[1] << 8) | ((unsigned)c[2] << 16);
    }
}

// intensity pyramid of the sessions with the photometric term (the others have no share)
__global__ __launch_bounds__(256) void kb_intensity(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    intensity_body(g, (const unsigned char*)b.in[s], d.image[0], d.cols, d.rows);
}
__global__ __launch_bounds__(256) void kb_pyrdown_u8(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = b.level;
    pyrdown_u8_body(g, d.image[l - 1], d.lc[l - 1], d.lr[l - 1], d.image[l], d.lc[l], d.lr[l]);
}
__global__ __launch_bounds__(256) void kb_derivative(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = pyr3_level(d, g.x, g);
    derivative_body(g, d.image[l], d.didx[l], d.didy[l], d.lc[l], d.lr[l]);
}
// after tracking: one 256-lane group per 4 KB of the three levels (BT_PYR3 counts 16x16
// tiles, 256 pixels each, and the lanes of a tile copy 16 bytes: enough groups for 4096
// bytes per tile; the extra groups of a level find nothing to copy)
__global__ __launch_bounds__(256) void kb_image_commit(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = pyr3_level(d, g.x, g);
    const int tile = g.y * g.nx + g.x;
    image_commit_body(tile * 256 + threadIdx.x, d.st, d.image[l], d.last_image[l], d.lc[l] * d.lr[l]);
}

static void blaunch(void (*k)(BatchLaunch), int grid, const BatchLaunch& b, hipStream_t s) {
    if (grid > 0) hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, s, b);
}
void launch_b_resize(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    b.level = 1;
    blaunch(kb_resize, batch_layout(b, h, BT_TILES, 1, 0), b, s);
}
void launch_b_filter(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    b.level = 0;
    blaunch(kb_filter, batch_layout(b, h, BT_TILES, 0, 0), b, s);
}
void launch_b_pyrdown(BatchLaunch& b, const SessDesc* h, int level, hipStream_t s) {
    b.level = level;
    blaunch(kb_pyrdown, batch_layout(b, h, BT_TILES, level, 0), b, s);
}
void launch_b_nmap(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    blaunch(kb_nmap, batch_layout(b, h, BT_PYR3, 0, 0), b, s);
}
void launch_b_images(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    b.level = 0;
    blaunch(kb_intensity, batch_layout(b, h, BT_TILES, 0, 0), b, s);
    for (int l = 1; l < 3; ++l) {
        b.level = l;
        blaunch(kb_pyrdown_u8, batch_layout(b, h, BT_TILES, l, 0), b, s);
    }
    blaunch(kb_derivative, batch_layout(b, h, BT_PYR3, 0, 0), b, s);
}
void launch_b_image_commit(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    blaunch(kb_image_commit, batch_layout(b, h, BT_PYR3, 0, 0), b, s);
}

}  // namespace fu
