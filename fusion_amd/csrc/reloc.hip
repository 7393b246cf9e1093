// Keypoint relocalisation without the detector: the parts of Frame::ExtractKeyPoints
// (Core/Frame.cc:93-212) and Tracker::Relocalise (Tracking/Tracking.cc:411-467) that do not
// depend on OpenCV's SURF.  The caller supplies the keypoints (pixel positions) and their
// descriptors; the library lifts them to 3-D with the frame's depth and normal maps, matches
// them against the key map's records (the role of cv::cuda::BFMatcher's knnMatch, k = 2, L2)
// and estimates the pose with the reference's RANSAC / Kabsch loop (Optimization/Solver.cc).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <system_error>
#include <thread>
#include <vector>

#include "fu_api.h"

using namespace fu;

namespace {

// ---------------------------------------------------------------------------- lifting
// The bounds test of InterpDepth / InterpNormal (Frame.cc:96, :133) compares y with the width
// and x with the height, as written; where it would let the four neighbours fall outside the
// image (y >= rows - 1 on a landscape frame: the reference reads past the image) or the
// position is NaN, the keypoint is refused.
__device__ __forceinline__ bool in_frame(float x, float y, int cols, int rows) {
    if (x <= 1 || y <= 1 || y >= cols - 1 || x >= rows - 1) return false;
    return !(isnan(x) || isnan(y) || x >= cols - 1 || y >= rows - 1);
}

// Frame::InterpDepth (Core/Frame.cc:93-129): bilinear depth at (x, y) from the four
// neighbours, NaN when any is invalid or outside [0.3, cutoff] or the result strays more than
// 5 mm from the upper-left one.
__device__ __forceinline__ float interp_depth(const float* d, int cols, int rows, float x, float y, float cutoff) {
    const float nan = __int_as_float(0x7fffffff);
    if (!in_frame(x, y, cols, rows)) return nan;
    const float cx = x - floorf(x), cy = y - floorf(y);
    const int x0 = (int)floorf(x), x1 = (int)ceilf(x), y0 = (int)floorf(y), y1 = (int)ceilf(y);
    const float d00 = d[y0 * cols + x0];
    if (isnan(d00) || (double)d00 < 0.3 || d00 > cutoff) return nan;
    const float d10 = d[y1 * cols + x0];
    if (isnan(d10) || (double)d10 < 0.3 || d10 > cutoff) return nan;
    const float d01 = d[y0 * cols + x1];
    if (isnan(d01) || (double)d01 < 0.3 || d01 > cutoff) return nan;
    const float d11 = d[y1 * cols + x1];
    if (isnan(d11) || (double)d11 < 0.3 || d11 > cutoff) return nan;
    const float e0 = d01 * cx + d00 * (1 - cx);
    const float e1 = d11 * cx + d10 * (1 - cx);
    const float fin = (1 - cy) * e0 + cy * e1;
    return (double)fabsf(fin - d00) <= 0.005 ? fin : nan;
}

// Frame::InterpNormal (Frame.cc:131-165): the same blend of the four normals (4-D), NaN when
// it differs from the upper-left one by more than 0.1 (4-D norm)
__device__ __forceinline__ f4 interp_normal(const f4* n, int cols, int rows, float x, float y) {
    const float nan = __int_as_float(0x7fffffff);
    const f4 bad = mk4(nan, nan, nan, nan);
    if (!in_frame(x, y, cols, rows)) return bad;
    const float cx = x - floorf(x), cy = y - floorf(y);
    const int x0 = (int)floorf(x), x1 = (int)ceilf(x), y0 = (int)floorf(y), y1 = (int)ceilf(y);
    const f4 d00 = n[y0 * cols + x0], d10 = n[y1 * cols + x0], d01 = n[y0 * cols + x1], d11 = n[y1 * cols + x1];
    const float ax = 1 - cx, ay = 1 - cy;
    const f4 e0 = mk4(d01.x * cx + d00.x * ax, d01.y * cx + d00.y * ax, d01.z * cx + d00.z * ax, d01.w * cx + d00.w * ax);
    const f4 e1 = mk4(d11.x * cx + d10.x * ax, d11.y * cx + d10.y * ax, d11.z * cx + d10.z * ax, d11.w * cx + d10.w * ax);
    const f4 fin = mk4(e0.x * ay + e1.x * cy, e0.y * ay + e1.y * cy, e0.z * ay + e1.z * cy, e0.w * ay + e1.w * cy);
    const f4 df = mk4(fin.x - d00.x, fin.y - d00.y, fin.z - d00.z, fin.w - d00.w);
    const float nm = sqrtf(df.x * df.x + df.y * df.y + df.z * df.z + df.w * df.w);
    return (double)nm <= 0.1 ? fin : bad;
}

struct LiftArgs {
    const float* xy;     // n keypoints (x, y) in pixels
    const float* depth;  // level-0 filtered depth
    const f4* nmap;      // level-0 normals
    int cols, rows, n;
    float fx, fy, cx, cy, cutoff;
    float* pos;          // n x 3
    float* normal;       // n x 4
    unsigned char* ok;   // n
};

// the per-keypoint test and back-projection of ExtractKeyPoints (Frame.cc:184-200)
__global__ void k_lift_keys(LiftArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float x = a.xy[2 * i], y = a.xy[2 * i + 1];
    const float nan = __int_as_float(0x7fffffff);
    float p[3] = {nan, nan, nan};
    f4 nr = mk4(nan, nan, nan, nan);
    bool ok = false;
    const float dp = interp_depth(a.depth, a.cols, a.rows, x, y, a.cutoff);
    if (!isnan(dp) && (double)dp > 0.3 && dp < a.cutoff) {
        const f4 n = interp_normal(a.nmap, a.cols, a.rows, x, y);
        if (!isnan(n.x)) {
            p[0] = dp * (x - a.cx) / a.fx;
            p[1] = dp * (y - a.cy) / a.fy;
            p[2] = dp;
            nr = n;
            ok = true;
        }
    }
    a.pos[3 * i + 0] = p[0];
    a.pos[3 * i + 1] = p[1];
    a.pos[3 * i + 2] = p[2];
    a.normal[4 * i + 0] = nr.x;
    a.normal[4 * i + 1] = nr.y;
    a.normal[4 * i + 2] = nr.z;
    a.normal[4 * i + 3] = nr.w;
    a.ok[i] = ok;
}

// ---------------------------------------------------------------------------- matching
// The two nearest train descriptors of every query descriptor by L2 distance (sqrt of the sum
// of the 64 squared differences, in dimension order); ties go to the lower train index.
// Work-group: 256 queries, one per lane (its descriptor in registers), against a chunk of
// train descriptors staged through LDS kTile at a time (every lane reads the same record: LDS
// broadcast).  Each (query, chunk) pair leaves its best two; a second pass merges the chunks.
// The chunk length is chosen per call so that the grid holds about kTargetGroups work-groups
// (a few hundred queries against tens of thousands of keys would otherwise fill a dozen CUs).
// Both passes keep the two smallest (distance, index) pairs in lexicographic order, a total
// order over the candidates (NaN and +inf distances never enter), so the result is the same for
// every chunking and merge order: that of one lane scanning the train set in index order with
// strict comparisons.
constexpr int kDesc = 64;
constexpr int kTile = 128;
constexpr int kChunk = 8192;
constexpr int kTargetGroups = 4096;

struct Best2 {
    float d0, d1;
    int i0, i1;
};

__device__ __forceinline__ bool lex_less(float d, int i, float e, int j) { return d < e || (d == e && i < j); }

// (d, i) into the best two; the empty slots are (+inf, -1), which no candidate with a finite
// distance ties
__device__ __forceinline__ void push2(Best2& b, float d, int i) {
    if (lex_less(d, i, b.d0, b.i0)) {
        b.d1 = b.d0;
        b.i1 = b.i0;
        b.d0 = d;
        b.i0 = i;
    } else if (lex_less(d, i, b.d1, b.i1)) {
        b.d1 = d;
        b.i1 = i;
    }
}

__global__ __launch_bounds__(256) void k_match_chunk(const float* q, int nq, const float* t, int nt, int chunk,
                                                     Best2* part) {
    __shared__ float tile[kTile * kDesc];
    const int qi = blockIdx.x * 256 + threadIdx.x;
    const int c0 = blockIdx.y * chunk, c1 = min(nt, c0 + chunk);
    float mine[kDesc];
#pragma unroll
    for (int j = 0; j < kDesc; ++j) mine[j] = qi < nq ? q[(size_t)qi * kDesc + j] : 0.f;
    Best2 b{INFINITY, INFINITY, -1, -1};
    for (int base = c0; base < c1; base += kTile) {
        const int cnt = min(kTile, c1 - base);
        __syncthreads();
        for (int k = threadIdx.x; k < cnt * kDesc; k += 256) tile[k] = t[(size_t)base * kDesc + k];
        __syncthreads();
        for (int r = 0; r < cnt; ++r) {
            float s = 0.f;
#pragma unroll
            for (int j = 0; j < kDesc; ++j) {
                const float dj = mine[j] - tile[r * kDesc + j];
                s = s + dj * dj;
            }
            const float d = sqrtf(s);
            if (d < b.d1) push2(b, d, base + r);  // indices rise within the chunk: d < d1 is the test
        }
    }
    if (qi < nq) part[(size_t)blockIdx.y * nq + qi] = b;
}

// one wave per query: lane l merges chunks l, l + 64, ..., then the lanes' best two are merged
// across the wave
__global__ __launch_bounds__(256) void k_match_merge(const Best2* part, int nq, int nchunks, int* idx, float* dist) {
    const int qi = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (qi >= nq) return;
    Best2 b{INFINITY, INFINITY, -1, -1};
    for (int c = lane; c < nchunks; c += 64) {
        const Best2 p = part[(size_t)c * nq + qi];
        if (p.i0 >= 0) push2(b, p.d0, p.i0);
        if (p.i1 >= 0) push2(b, p.d1, p.i1);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float d0 = __shfl_xor(b.d0, off), d1 = __shfl_xor(b.d1, off);
        const int i0 = __shfl_xor(b.i0, off), i1 = __shfl_xor(b.i1, off);
        if (i0 >= 0) push2(b, d0, i0);
        if (i1 >= 0) push2(b, d1, i1);
    }
    if (lane == 0) {
        idx[2 * qi] = b.i0;
        idx[2 * qi + 1] = b.i1;
        dist[2 * qi] = b.d0;
        dist[2 * qi + 1] = b.d1;
    }
}

// train records per chunk: about kTargetGroups work-groups in all, at least one LDS tile and at
// most kChunk records each, a whole number of tiles
static int match_chunk_len(int nq, int nt) {
    const long long qgroups = (nq + 255) / 256;
    long long want = ((long long)nt * qgroups + kTargetGroups - 1) / kTargetGroups;
    want = (want + kTile - 1) / kTile * kTile;
    return (int)std::min<long long>(kChunk, std::max<long long>(kTile, want));
}

// ---------------------------------------------------------------------------- graph selection
// Tracker::GenerateGraph's choice of mutually consistent pairs (Tracking.cc:515-596) over the
// consistency matrix C (row-major n x n, C[a * n + b] = conMatrix(a, b)), without reading the
// matrix back to the host as the reference does:
//   k_rank_sums   the column sums (cv::cuda::reduce over rows): each column's rows in order,
//                 summed in double and rounded to float (OpenCV's order is its own).  A
//                 work-group takes 64 columns: all its lanes stage tiles of kSumRows rows in
//                 LDS, and one wave adds them, a lane per column;
//   k_rank_order  the pairs by rank, largest first (cv::sortIdx descending): a pair's position
//                 is the number of ranks above its own, ties going to the lower index
//                 (std::sort leaves them in no fixed order), counted against LDS tiles of the
//                 ranks by four waves a quarter each and added; a NaN rank counts as -inf, so
//                 the positions are a permutation;
//   k_subgraphs   one 1024-lane work-group per list: the head and the pairs ranked after it
//                 with C[head][pair] > min_score, kept in rank order by a block-wide ballot
//                 prefix until max_sel are selected; then the refinement test of each selected
//                 pair against every later pair of the list, one wave per pair, its lanes
//                 over the later pairs (the test is an "any", so their order is free).
constexpr int kSelLanes = 1024;
constexpr int kMaxSelect = 1024;

constexpr int kSumRows = 128;  // rows per LDS tile of k_rank_sums (32 KB)

// launch: ceil(n / 64) work-groups of 256 lanes.  Lane t loads columns 4 (t & 15) .. +3 of
// rows t >> 4, (t >> 4) + 16, ... of a tile (eight 16-byte loads when the row is whole and
// aligned), the next tile's loads issued before wave 0 adds the current one from LDS.
__global__ __launch_bounds__(256) void k_rank_sums(const float* C, int n, float* rank) {
    __shared__ __align__(16) float tile[kSumRows][64];  // written as 16-byte words
    const int tid = threadIdx.x, col = tid & 63;
    const int c0 = blockIdx.x * 64;
    const int q = (tid & 15) * 4, r_lane = tid >> 4;
    // whole 16-byte words: an aligned matrix and n a multiple of 4 keep every row start aligned
    const bool vec = ((uintptr_t)C & 15) == 0 && (n & 3) == 0 && c0 + q + 3 < n;
    float4 buf[kSumRows / 16];
    auto fetch = [&](int r0) {
#pragma unroll
        for (int k = 0; k < kSumRows / 16; ++k) {
            const int r = r0 + r_lane + 16 * k;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r < n) {
                const float* row = C + (size_t)r * n + c0 + q;
                if (vec) {
                    v = *reinterpret_cast<const float4*>(row);
                } else {
                    v.x = c0 + q < n ? row[0] : 0.f;
                    v.y = c0 + q + 1 < n ? row[1] : 0.f;
                    v.z = c0 + q + 2 < n ? row[2] : 0.f;
                    v.w = c0 + q + 3 < n ? row[3] : 0.f;
                }
            }
            buf[k] = v;
        }
    };
    double s = 0;  // wave 0: column c0 + col
    fetch(0);
    for (int r0 = 0; r0 < n; r0 += kSumRows) {
        __syncthreads();  // the previous tile is summed
#pragma unroll
        for (int k = 0; k < kSumRows / 16; ++k) *reinterpret_cast<float4*>(&tile[r_lane + 16 * k][q]) = buf[k];
        __syncthreads();
        if (r0 + kSumRows < n) fetch(r0 + kSumRows);
        // rows past n hold +0, which leaves the sum as it is (it starts at +0 and so is never
        // -0), so every tile is summed whole
        if (tid < 64) {
#pragma unroll 16
            for (int r = 0; r < kSumRows; ++r) s += (double)tile[r][col];
        }
    }
    if (tid < 64 && c0 + col < n) rank[c0 + col] = (float)s;
}

__device__ __forceinline__ float rank_key(float v) { return isnan(v) ? -INFINITY : v; }

// launch: ceil(n / 64) work-groups of 256 lanes; lane (w, l) counts for pair 64 * group + l
// against entries w * 256 .. w * 256 + 255 of every 1024-rank tile
__global__ __launch_bounds__(256) void k_rank_order(const float* rank, int n, int* order) {
    __shared__ float tile[1024];
    __shared__ int part[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const float kx = x < n ? rank_key(rank[x]) : 0.f;
    int pos = 0;
    for (int base = 0; base < n; base += 1024) {
        __syncthreads();
        // slots past n hold -inf: never above a rank, and equal to one only with an index
        // above it, so they never count
        for (int k = threadIdx.x; k < 1024; k += 256) tile[k] = base + k < n ? rank_key(rank[base + k]) : -INFINITY;
        __syncthreads();
        const int lo = w * 256;
#pragma unroll 16
        for (int k = lo; k < lo + 256; ++k) {
            const float ky = tile[k];
            pos += (ky > kx) || (ky == kx && base + k < x);
        }
    }
    part[w][lane] = pos;
    __syncthreads();
    if (w == 0 && x < n) order[part[0][lane] + part[1][lane] + part[2][lane] + part[3][lane]] = x;
}

struct SelArgs {
    const float* C;
    const int* order;
    int n, max_sel;
    float min_score;
    int* n_selected;  // per list
    int* n_refined;   // per list
    int* refined;     // max_sel per list
};

// the exclusive prefix of `flag` over the work-group (lanes in order) and the total
__device__ __forceinline__ int block_prefix(bool flag, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(flag);
    const int in_wave = __popcll(bal & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) wsum[wid] = __popcll(bal);
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < kSelLanes / 64; ++w) {
        const int c = wsum[w];
        before += w < wid ? c : 0;
        all += c;
    }
    total = all;
    return before + in_wave;
}

__global__ __launch_bounds__(kSelLanes) void k_subgraphs(SelArgs a) {
    __shared__ int sel[kMaxSelect];
    __shared__ float hs[kMaxSelect];  // C[head][sel[k]]
    __shared__ int wsum[kSelLanes / 64];
    const int i = blockIdx.x, tid = threadIdx.x;
    if (i >= a.n) {  // no pairs from this rank on (Tracking.cc:538)
        if (tid == 0) a.n_selected[i] = a.n_refined[i] = 0;
        return;
    }
    const int head = a.order[i];
    const float* Ch = a.C + (size_t)head * a.n;
    int count = 1;  // the head is always taken (:544-548)
    if (tid == 0) {
        sel[0] = head;
        hs[0] = Ch[head];
    }
    for (int base = i + 1; base < a.n && count < a.max_sel; base += kSelLanes) {
        const int j = base + tid;
        int idx = 0;
        float sc = 0.f;
        bool f = false;
        if (j < a.n) {
            idx = a.order[j];
            sc = Ch[idx];
            f = sc > a.min_score;  // :551-555
        }
        int total;
        const int slot = count + block_prefix(f, wsum, total);
        if (f && slot < a.max_sel) {
            sel[slot] = idx;
            hs[slot] = sc;
        }
        count = min(count + total, a.max_sel);  // the same in every lane
    }
    __syncthreads();
    if (tid == 0) a.n_selected[i] = count;
    if (count < 3) {  // no PoseEstimate from fewer than three (:565)
        if (tid == 0) a.n_refined[i] = 0;
        return;
    }
    // the t-th pair is kept unless a later one is nearly exclusive with it and scores higher
    // with the head (:567-582); the head (t = 0) is never kept.  Wave w tests pairs w + 1,
    // w + 17, ..., its lanes over the later pairs 64 at a time.
    __shared__ unsigned char kept[kMaxSelect];
    const int lane = tid & 63, wid = tid >> 6;
    for (int t = 1 + wid; t < count; t += kSelLanes / 64) {
        const int av = sel[t];
        const float ha = hs[t];
        const float* Ca = a.C + (size_t)av * a.n;
        bool drop = false;
#pragma unroll 4
        for (int l = t + 1 + lane; l < count; l += 64) {
            const int b = sel[l];
            const bool x = Ca[b] < 5e-3f || a.C[(size_t)b * a.n + av] < 5e-3f;
            drop = drop || (hs[l] > ha && x);
        }
        if (lane == 0) kept[t] = 0;
        if (__ballot(drop) == 0ull && lane == 0) kept[t] = 1;
    }
    __syncthreads();
    const bool keep = tid >= 1 && tid < count && kept[tid];
    int total;
    const int slot = block_prefix(keep, wsum, total);
    if (keep) a.refined[(size_t)i * a.max_sel + slot] = sel[tid];
    if (tid == 0) a.n_refined[i] = total;
}

// ---------------------------------------------------------------------------- pose
// 3x3 helpers in double, row-major
struct M3d {
    double a[9];
};
M3d mul(const M3d& x, const M3d& y) {
    M3d r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = x.a[i * 3] * y.a[j] + x.a[i * 3 + 1] * y.a[3 + j] + x.a[i * 3 + 2] * y.a[6 + j];
    return r;
}
M3d transpose(const M3d& x) {
    M3d r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.a[i * 3 + j] = x.a[j * 3 + i];
    return r;
}
double det(const M3d& m) {
    const double* a = m.a;
    return a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) + a[2] * (a[3] * a[7] - a[4] * a[6]);
}

// R = U V^T for Ab = U S V^T (the Kabsch rotation of Solver.cc, R = (V U^T)^T).  One-sided
// Jacobi on the columns of Ab gives Ab V = U S; a singular vector of a zero singular value
// (three points of a sample span a plane) is completed as the cross product of the other two,
// on both sides, so such a sample yields the proper rotation.
M3d kabsch(const M3d& Ab) {
    double A[9], V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(A, Ab.a, sizeof(A));
    for (int sweep = 0; sweep < 40; ++sweep) {
        double off = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double al = 0, be = 0, ga = 0;
                for (int i = 0; i < 3; ++i) {
                    al += A[i * 3 + p] * A[i * 3 + p];
                    be += A[i * 3 + q] * A[i * 3 + q];
                    ga += A[i * 3 + p] * A[i * 3 + q];
                }
                if (ga == 0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                off = fmax(off, fabs(ga) / sqrt(al * be));
                const double z = (be - al) / (2 * ga);
                const double t = (z >= 0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1 + z * z));
                const double c = 1 / sqrt(1 + t * t), s = c * t;
                for (int i = 0; i < 3; ++i) {
                    const double ap = A[i * 3 + p], aq = A[i * 3 + q];
                    A[i * 3 + p] = c * ap - s * aq;
                    A[i * 3 + q] = s * ap + c * aq;
                    const double vp = V[i * 3 + p], vq = V[i * 3 + q];
                    V[i * 3 + p] = c * vp - s * vq;
                    V[i * 3 + q] = s * vp + c * vq;
                }
            }
        if (off < 1e-15) break;
    }
    // singular values (column norms), largest first
    double sv[3];
    int ord[3] = {0, 1, 2};
    for (int k = 0; k < 3; ++k) sv[k] = sqrt(A[k] * A[k] + A[3 + k] * A[3 + k] + A[6 + k] * A[6 + k]);
    std::sort(ord, ord + 3, [&](int x, int y) { return sv[x] > sv[y]; });
    double U[9], W[9];
    for (int c = 0; c < 3; ++c) {
        const int k = ord[c];
        for (int i = 0; i < 3; ++i) W[i * 3 + c] = V[i * 3 + k];
        const double inv = sv[k] > 0 ? 1 / sv[k] : 0;
        for (int i = 0; i < 3; ++i) U[i * 3 + c] = A[i * 3 + k] * inv;
    }
    if (sv[ord[2]] <= 1e-12 * fmax(sv[ord[0]], 1e-300)) {
        // rank <= 2: third columns as cross products of the first two
        U[2] = U[3] * U[7] - U[6] * U[4];
        U[5] = U[6] * U[1] - U[0] * U[7];
        U[8] = U[0] * U[4] - U[3] * U[1];
        W[2] = W[3] * W[7] - W[6] * W[4];
        W[5] = W[6] * W[1] - W[0] * W[7];
        W[8] = W[0] * W[4] - W[3] * W[1];
    }
    M3d u, w;
    memcpy(u.a, U, sizeof(U));
    memcpy(w.a, W, sizeof(W));
    return mul(u, transpose(w));
}

struct V3d {
    double x, y, z;
};
V3d operator-(V3d a, V3d b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3d operator+(V3d a, V3d b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3d operator/(V3d a, double s) { return {a.x / s, a.y / s, a.z / s}; }
V3d rot(const M3d& R, V3d v) {
    return {R.a[0] * v.x + R.a[1] * v.y + R.a[2] * v.z, R.a[3] * v.x + R.a[4] * v.y + R.a[5] * v.z,
            R.a[6] * v.x + R.a[7] * v.y + R.a[8] * v.z};
}
double vnorm(V3d v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
V3d vcross(V3d a, V3d b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
void outer_add(M3d& m, V3d a, V3d b) {
    const double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) m.a[i * 3 + j] += av[i] * bv[j];
}

// the sample generator: splitmix64 in place of the reference's srand(clock) / rand()
struct Rng {
    uint64_t s;
    uint32_t next(uint32_t n) {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (uint32_t)((z >> 32) % n);
    }
};

}  // namespace

// ---------------------------------------------------------------------------- C ABI
extern "C" fu_status fu_lift_keypoints(const float* xy, int32_t n, const float* depth, const float* nmap, int32_t cols,
                                       int32_t rows, float fx, float fy, float cx, float cy, float depth_cutoff,
                                       float* pos, float* normal, uint8_t* ok, fu_stream s) {
    if (n < 0 || cols < 1 || rows < 1) return api_error(FU_ERR_ARG, "fu_lift_keypoints: bad sizes");
    if (n == 0) return FU_OK;
    if (!xy || !depth || !nmap || !pos || !normal || !ok) return api_error(FU_ERR_ARG, "fu_lift_keypoints: null argument");
    LiftArgs a{xy, depth, (const f4*)nmap, cols, rows, n, fx, fy, cx, cy, depth_cutoff, pos, normal, ok};
    hipLaunchKernelGGL(k_lift_keys, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_error(FU_ERR_HIP, "fu_lift_keypoints: %s", hipGetErrorString(e));
    return FU_OK;
}

extern "C" fu_status fu_match_knn2(const float* query, int32_t nq, const float* train, int32_t nt, int32_t* idx,
                                   float* dist, fu_stream s) {
    if (nq < 0 || nt < 0) return api_error(FU_ERR_ARG, "fu_match_knn2: bad sizes");
    if (nq == 0) return FU_OK;
    if (!query || !idx || !dist || (nt > 0 && !train)) return api_error(FU_ERR_ARG, "fu_match_knn2: null argument");
    if ((long long)nq * kDesc > (1ll << 31) || (long long)nt * kDesc > (1ll << 31))
        return api_error(FU_ERR_ARG, "fu_match_knn2: %d x %d descriptors (at most 2^25 each)", nq, nt);
    hipStream_t hs = (hipStream_t)s;
    const int chunk = match_chunk_len(nq, nt);
    const int nchunks = nt > 0 ? (nt + chunk - 1) / chunk : 1;
    Best2* part = nullptr;
    if (hipMallocAsync((void**)&part, sizeof(Best2) * (size_t)nchunks * nq, hs) != hipSuccess) {
        (void)hipGetLastError();
        return api_error(FU_ERR_NOMEM, "fu_match_knn2: %d partial lists", nchunks);
    }
    if (nt > 0)
        hipLaunchKernelGGL(k_match_chunk, dim3((nq + 255) / 256, nchunks), dim3(256), 0, hs, query, nq, train, nt, chunk,
                           part);
    hipLaunchKernelGGL(k_match_merge, dim3((nq + 3) / 4), dim3(256), 0, hs, part, nq, nt > 0 ? nchunks : 0, idx,
                       dist);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipFreeAsync(part, hs);
    if (e != hipSuccess) return api_error(FU_ERR_HIP, "fu_match_knn2: %s", hipGetErrorString(e));
    return FU_OK;
}

// Solver::PoseEstimate (Optimization/Solver.cc:6-183), checkAngle = false
extern "C" fu_status fu_pose_estimate(const double* src, const double* ref, int32_t n, int32_t iterations,
                                      uint64_t seed, double T[16], uint8_t* inlier, int32_t* ok) {
    if (!src || !ref || !T || !ok || n < 0 || iterations < 0) return api_error(FU_ERR_ARG, "fu_pose_estimate: bad argument");
    *ok = 0;
    for (int k = 0; k < 16; ++k) T[k] = (k % 5 == 0) ? 1.0 : 0.0;
    if (n < 3) return FU_OK;
    auto S = [&](int i) { return V3d{src[3 * i], src[3 * i + 1], src[3 * i + 2]}; };
    auto Rf = [&](int i) { return V3d{ref[3 * i], ref[3 * i + 1], ref[3 * i + 2]}; };
    M3d R_best{{1, 0, 0, 0, 1, 0, 0, 0, 1}};
    V3d t_best{0, 0, 0};
    int inliers_best = 0;
    const float thresh_inlier = 0.05f;
    const float thresh_confidence = 0.95f;
    const int min_iter = 20;
    float ratio = 0.f, confidence = 0.f;
    std::vector<char> outlier;  // the last evaluated hypothesis' outliers (Solver.cc:84-93)
    Rng rng{seed};
    int it = 0;
    // a reflection re-draws without counting the iteration (as the reference); a three-point
    // sample spans at most a plane and its completed bases are right-handed, so that does not
    // repeat, but the draws are bounded all the same
    long long draws = 0;
    const long long max_draws = 64ll * (iterations > 0 ? iterations : 1);
    while (it < iterations && draws++ < max_draws) {
        int smp[3];
        for (int k = 0; k < 3; ++k) smp[k] = (int)rng.next((uint32_t)n);
        const bool bad = smp[0] == smp[1] || smp[1] == smp[2] || smp[2] == smp[0];
        V3d sa = S(smp[0]), sb = S(smp[1]), sc = S(smp[2]);
        V3d ra = Rf(smp[0]), rb = Rf(smp[1]), rc = Rf(smp[2]);
        const float sd = (float)vnorm(vcross(sb - sa, sa - sc));  // stored in floats (Solver.cc:52-53)
        const float rd = (float)vnorm(vcross(rb - ra, ra - rc));
        if (bad || (double)sd < 1e-6 || (double)rd < 1e-6) {
            ++it;
            continue;
        }
        const V3d sm = (sa + sb + sc) / 3, rm = (ra + rb + rc) / 3;
        M3d Ab{{0, 0, 0, 0, 0, 0, 0, 0, 0}};
        outer_add(Ab, sa - sm, ra - rm);
        outer_add(Ab, sb - sm, rb - rm);
        outer_add(Ab, sc - sm, rc - rm);
        M3d R = kabsch(Ab);
        if (det(R) < 0) continue;  // as Solver.cc:82: a new sample, the iteration not counted
        const V3d t = sm - rot(R, rm);
        int nin = 0;
        outlier.assign(n, 1);
        for (int i = 0; i < n; ++i)
            if (vnorm(S(i) - (rot(R, Rf(i)) + t)) <= thresh_inlier) {
                ++nin;
                outlier[i] = 0;
            }
        if (nin > inliers_best) {
            V3d smn{0, 0, 0}, rmn{0, 0, 0};
            for (int i = 0; i < n; ++i)
                if (!outlier[i]) {
                    smn = smn + S(i);
                    rmn = rmn + Rf(i);
                }
            smn = smn / nin;
            rmn = rmn / nin;
            M3d Ab2{{0, 0, 0, 0, 0, 0, 0, 0, 0}};
            for (int i = 0; i < n; ++i)
                if (!outlier[i]) outer_add(Ab2, S(i) - smn, Rf(i) - rmn);
            R_best = kabsch(Ab2);
            t_best = smn - rot(R_best, rmn);
            inliers_best = nin;
            ratio = (float)nin / (float)n;
            // pow of a float and an int is the double pow (Solver.cc:124)
            confidence = (float)(1 - pow(1 - pow((double)ratio, 3), (double)(it + 1)));
            if (it >= min_iter && confidence >= thresh_confidence) break;
        }
        ++it;
    }
    if (outlier.empty()) return FU_OK;  // no hypothesis was evaluated
    // the final set: the best model's inliers added to the last evaluated hypothesis' (the
    // reference does not reset the flags first, Solver.cc:138-143)
    for (int i = 0; i < n; ++i)
        if (vnorm(S(i) - (rot(R_best, Rf(i)) + t_best)) <= thresh_inlier) outlier[i] = 0;
    int N = 0;
    V3d smn{0, 0, 0}, rmn{0, 0, 0};
    for (int i = 0; i < n; ++i)
        if (!outlier[i]) {
            smn = smn + S(i);
            rmn = rmn + Rf(i);
            ++N;
        }
    if (N == 0) return FU_OK;
    smn = smn / N;
    rmn = rmn / N;
    M3d Ab{{0, 0, 0, 0, 0, 0, 0, 0, 0}};
    for (int i = 0; i < n; ++i)
        if (!outlier[i]) outer_add(Ab, S(i) - smn, Rf(i) - rmn);
    R_best = kabsch(Ab);
    if (det(R_best) < 0) return FU_OK;
    t_best = smn - rot(R_best, rmn);
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T[i * 4 + j] = R_best.a[i * 3 + j];
    }
    T[3] = t_best.x;
    T[7] = t_best.y;
    T[11] = t_best.z;
    if (inlier)
        for (int i = 0; i < n; ++i) inlier[i] = !outlier[i];
    (void)inliers_best;
    *ok = 1;
    return FU_OK;
}

namespace {
// record j of `out` = keys[idx[j]], or zeros where idx[j] < 0 (a query without a neighbour);
// one lane per 16 bytes
__global__ void k_gather_keys(const fu_surf* keys, const int32_t* idx, int n, fu_surf* out) {
    constexpr int Q = (int)(sizeof(fu_surf) / 16);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * Q) return;
    const int j = t / Q, q = t % Q;
    const int k = idx[j];
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (k >= 0) v = reinterpret_cast<const uint4*>(&keys[k])[q];
    reinterpret_cast<uint4*>(&out[j])[q] = v;
}
static_assert(sizeof(fu_surf) % 16 == 0, "records move as 16-byte words");

// the map's keys (UpdateMapKeys, Mapping.cc:142-151; the first record is never written by
// CollectKeyPoints and reads as zeros here) and the two nearest of them to each frame
// descriptor: `nk` map keys; on the host, the neighbours' indices and distances (2 per query,
// -1 / +inf where the map has fewer keys) and their records in the same slots (zeros for -1).
// Device buffers are stream-ordered allocations; only the matched records come back.
fu_status collect_and_match(fu_keymap* km, const float* frame_desc, int32_t nq, uint32_t& nk,
                            std::vector<fu_surf>& matched, std::vector<int32_t>& hidx, std::vector<float>& hdist,
                            fu_stream s, const char* who) {
    hipStream_t hs = (hipStream_t)s;
    nk = 0;
    fu_status st = fu_collect_key_points(km, nullptr, 0, &nk, s);
    if (st) return st;
    matched.clear();
    if (nk == 0 || nq == 0) return FU_OK;  // no map keys (Tracking.cc:417-418)
    fu_surf *keys = nullptr, *mrec = nullptr;
    float* train = nullptr;
    int32_t* idx = nullptr;
    float* dist = nullptr;
    auto release = [&]() {
        if (keys) (void)hipFreeAsync(keys, hs);
        if (mrec) (void)hipFreeAsync(mrec, hs);
        if (train) (void)hipFreeAsync(train, hs);
        if (idx) (void)hipFreeAsync(idx, hs);
        if (dist) (void)hipFreeAsync(dist, hs);
    };
    const int nm = 2 * nq;
    if (hipMallocAsync((void**)&keys, sizeof(fu_surf) * nk, hs) != hipSuccess ||
        hipMallocAsync((void**)&mrec, sizeof(fu_surf) * nm, hs) != hipSuccess ||
        hipMallocAsync((void**)&train, sizeof(float) * 64 * nk, hs) != hipSuccess ||
        hipMallocAsync((void**)&idx, sizeof(int32_t) * nm, hs) != hipSuccess ||
        hipMallocAsync((void**)&dist, sizeof(float) * nm, hs) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return api_error(FU_ERR_NOMEM, "%s: %u keys", who, nk);
    }
    matched.resize(nm);
    hidx.resize(nm);
    hdist.resize(nm);
    hipError_t e = hipMemsetAsync(keys, 0, sizeof(fu_surf) * nk, hs);
    if (e == hipSuccess) {
        uint32_t again = 0;
        st = fu_collect_key_points(km, keys, nk, &again, s);
    }
    // descriptors packed for the matcher: 256 bytes of each 288-byte record
    if (!st && e == hipSuccess)
        e = hipMemcpy2DAsync(train, 64 * sizeof(float), (const char*)keys + offsetof(fu_surf, descriptor), sizeof(fu_surf),
                             64 * sizeof(float), nk, hipMemcpyDeviceToDevice, hs);
    if (!st && e == hipSuccess) st = fu_match_knn2(frame_desc, nq, train, (int32_t)nk, idx, dist, s);
    if (!st && e == hipSuccess) {
        const int lanes = nm * (int)(sizeof(fu_surf) / 16);
        hipLaunchKernelGGL(k_gather_keys, dim3((lanes + 255) / 256), dim3(256), 0, hs, keys, idx, nm, mrec);
        e = hipGetLastError();
    }
    if (!st && e == hipSuccess) e = hipMemcpyAsync(matched.data(), mrec, sizeof(fu_surf) * nm, hipMemcpyDeviceToHost, hs);
    if (!st && e == hipSuccess) e = hipMemcpyAsync(hidx.data(), idx, sizeof(int32_t) * nm, hipMemcpyDeviceToHost, hs);
    if (!st && e == hipSuccess) e = hipMemcpyAsync(hdist.data(), dist, sizeof(float) * nm, hipMemcpyDeviceToHost, hs);
    release();
    // also after a failure: nothing queued may still write into the host vectors
    const hipError_t es = hipStreamSynchronize(hs);
    if (e == hipSuccess) e = es;
    if (st) return st;
    if (e != hipSuccess) return api_error(FU_ERR_HIP, "%s: %s", who, hipGetErrorString(e));
    return FU_OK;
}

// the camera pose of a PoseEstimate transform T (frame point = R * map point + t): T^-1
void rigid_inverse(const double T[16], double pose[16]) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) pose[i * 4 + j] = T[j * 4 + i];
        pose[i * 4 + 3] = -(T[0 * 4 + i] * T[3] + T[1 * 4 + i] * T[7] + T[2 * 4 + i] * T[11]);
    }
    pose[12] = pose[13] = pose[14] = 0;
    pose[15] = 1;
}
}  // namespace

// Tracker::Relocalise with useGraphMatching off, its default (Tracking.cc:411-467): the map's
// keys, the two nearest map descriptors of every frame descriptor, the ratio test
// d0 < 0.9 d1, PoseEstimate over the surviving pairs (frame point <- map point) and the pose
// as the inverse of its transform.  frame_desc: nq x 64 floats in device memory; frame_pos:
// nq x 3 doubles on the host (the lifted keypoints).
extern "C" fu_status fu_relocalise_keys(fu_keymap* km, const float* frame_desc, const double* frame_pos, int32_t nq,
                                        int32_t iterations, uint64_t seed, double pose[16], int32_t* n_pairs,
                                        int32_t* ok, fu_stream s) {
    if (!km || !pose || !ok || nq < 0 || (nq > 0 && (!frame_desc || !frame_pos)))
        return api_error(FU_ERR_ARG, "fu_relocalise_keys: bad argument");
    *ok = 0;
    if (n_pairs) *n_pairs = 0;
    for (int k = 0; k < 16; ++k) pose[k] = (k % 5 == 0) ? 1.0 : 0.0;
    std::vector<fu_surf> matched;
    std::vector<int32_t> hidx;
    std::vector<float> hdist;
    uint32_t nk = 0;
    fu_status st = collect_and_match(km, frame_desc, nq, nk, matched, hidx, hdist, s, "fu_relocalise_keys");
    if (st) return st;
    if (matched.empty()) return FU_OK;
    // the ratio test (Tracking.cc:437-444); with one map key the second distance is infinite
    std::vector<double> src, ref;
    for (int i = 0; i < nq; ++i) {
        if (hidx[2 * i] < 0) continue;
        if ((double)hdist[2 * i] < 0.90 * (double)hdist[2 * i + 1]) {
            const fu_surf& k = matched[2 * i];
            src.insert(src.end(), {frame_pos[3 * i], frame_pos[3 * i + 1], frame_pos[3 * i + 2]});
            ref.insert(ref.end(), {(double)k.pos[0], (double)k.pos[1], (double)k.pos[2]});
        }
    }
    const int np = (int)(src.size() / 3);
    if (n_pairs) *n_pairs = np;
    if (np < 3) return FU_OK;
    double T[16];
    int32_t good = 0;
    if ((st = fu_pose_estimate(src.data(), ref.data(), np, iterations, seed, T, nullptr, &good))) return st;
    if (!good) return FU_OK;
    rigid_inverse(T, pose);
    *ok = 1;
    return FU_OK;
}

// GenerateGraph's selection on the device (Tracking.cc:515-596)
extern "C" fu_status fu_select_subgraphs(const float* matrix, int32_t n, int32_t n_lists, int32_t max_select,
                                         float min_score, float* rank_sum, int32_t* order, int32_t* n_selected,
                                         int32_t* n_refined, int32_t* refined, fu_stream s) {
    if (n < 0 || n > 46340 || n_lists < 1 || n_lists > 65535 || max_select < 1 || max_select > kMaxSelect)
        return api_error(FU_ERR_ARG, "fu_select_subgraphs: n %d (at most 46340), %d lists, %d per list (at most %d)", n,
                         n_lists, max_select, kMaxSelect);
    if (!n_selected || !n_refined || !refined || (n > 0 && (!matrix || !rank_sum || !order)))
        return api_error(FU_ERR_ARG, "fu_select_subgraphs: null argument");
    hipStream_t hs = (hipStream_t)s;
    if (n > 0) {
        hipLaunchKernelGGL(k_rank_sums, dim3((n + 63) / 64), dim3(256), 0, hs, matrix, n, rank_sum);
        hipLaunchKernelGGL(k_rank_order, dim3((n + 63) / 64), dim3(256), 0, hs, rank_sum, n, order);
    }
    const SelArgs a{matrix, order, n, max_select, min_score, n_selected, n_refined, refined};
    hipLaunchKernelGGL(k_subgraphs, dim3(n_lists), dim3(kSelLanes), 0, hs, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return api_error(FU_ERR_HIP, "fu_select_subgraphs: %s", hipGetErrorString(e));
    return FU_OK;
}

// Tracker::Relocalise with useGraphMatching on (Tracking.cc:432-453) through GenerateGraph's
// PoseEstimate calls (:469-629): the match list keeps both nearest map keys of a query that
// fails the ratio test; pairs whose map key is not valid are dropped; their consistency
// matrix (BuildAdjacencyMatrix) and sub-graph selection run on the device; then PoseEstimate
// on the pairs of the first list of three or more selected pairs, once per kept list up to
// five (the reference's cv::sortIdx over its one-column list of counts leaves every index at
// 0, so each call takes that first list), when it holds four or more refined pairs.  Call c
// draws its samples from seed + c.  frame_pos (nq x 3) and frame_normal (nq x 4): host
// floats, the lifted keypoints.
extern "C" fu_status fu_relocalise_graph_candidates(fu_keymap* km, const float* frame_desc, const float* frame_pos,
                                                    const float* frame_normal, int32_t nq, int32_t iterations,
                                                    uint64_t seed, double poses[5 * 16], int32_t* n_poses,
                                                    int32_t* n_pairs, fu_stream s) {
    if (!km || !poses || !n_poses || nq < 0 || iterations < 0 ||
        (nq > 0 && (!frame_desc || !frame_pos || !frame_normal)))
        return api_error(FU_ERR_ARG, "fu_relocalise_graph_candidates: bad argument");
    constexpr int kLists = 10, kSelect = 400, kPoseLists = 5;  // Tracking.h:129-132
    constexpr float kMinScore = 0.1f;
    *n_poses = 0;
    if (n_pairs) *n_pairs = 0;
    std::vector<fu_surf> matched;
    std::vector<int32_t> hidx;
    std::vector<float> hdist;
    uint32_t nk = 0;
    fu_status st =
        collect_and_match(km, frame_desc, nq, nk, matched, hidx, hdist, s, "fu_relocalise_graph_candidates");
    if (st) return st;
    if (matched.empty()) return FU_OK;
    // the match list (Tracking.cc:437-444), then the valid map keys (:485-505)
    std::vector<std::pair<int, int>> refd;  // (query, neighbour slot 2 * query + 0 or 1)
    std::vector<float> rdist;
    for (int i = 0; i < nq; ++i) {
        if (hidx[2 * i] < 0) continue;
        refd.push_back({i, 2 * i});
        rdist.push_back(hdist[2 * i]);
        if (!((double)hdist[2 * i] < 0.90 * (double)hdist[2 * i + 1]) && hidx[2 * i + 1] >= 0) {
            refd.push_back({i, 2 * i + 1});
            rdist.push_back(hdist[2 * i + 1]);
        }
    }
    if (n_pairs) *n_pairs = (int32_t)refd.size();
    if (refd.size() < 3) return FU_OK;
    std::vector<fu_surf> fk, mk;
    std::vector<float> dd;
    for (size_t r = 0; r < refd.size(); ++r) {
        const fu_surf& key = matched[refd[r].second];
        if (!key.valid) continue;
        fu_surf f;
        memset(&f, 0, sizeof(f));
        f.valid = 1;  // the reference leaves this record's flag unset
        memcpy(f.pos, frame_pos + 3 * refd[r].first, sizeof(f.pos));
        memcpy(f.normal, frame_normal + 4 * refd[r].first, sizeof(f.normal));
        fk.push_back(f);
        mk.push_back(key);
        dd.push_back(rdist[r]);
    }
    const int n = (int)fk.size();
    if (n == 0) return FU_OK;  // rank.cols == 0 (:523-524)
    if (n > 46340) return api_error(FU_ERR_ARG, "fu_relocalise_graph_candidates: %d pairs (at most 46340)", n);
    hipStream_t hs = (hipStream_t)s;
    fu_surf *dfk = nullptr, *dmk = nullptr;
    float *ddist = nullptr, *mat = nullptr, *rank = nullptr;
    int32_t *order = nullptr, *nsel = nullptr;
    auto release = [&]() {
        for (void* p : {(void*)dfk, (void*)dmk, (void*)ddist, (void*)mat, (void*)rank, (void*)order, (void*)nsel})
            if (p) (void)hipFreeAsync(p, hs);
    };
    const size_t nsel_words = 2 * kLists + (size_t)kLists * kSelect;
    if (hipMallocAsync((void**)&dfk, sizeof(fu_surf) * n, hs) != hipSuccess ||
        hipMallocAsync((void**)&dmk, sizeof(fu_surf) * n, hs) != hipSuccess ||
        hipMallocAsync((void**)&ddist, sizeof(float) * n, hs) != hipSuccess ||
        hipMallocAsync((void**)&mat, sizeof(float) * (size_t)n * n, hs) != hipSuccess ||
        hipMallocAsync((void**)&rank, sizeof(float) * n, hs) != hipSuccess ||
        hipMallocAsync((void**)&order, sizeof(int32_t) * n, hs) != hipSuccess ||
        hipMallocAsync((void**)&nsel, sizeof(int32_t) * nsel_words, hs) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return api_error(FU_ERR_NOMEM, "fu_relocalise_graph_candidates: %d pairs", n);
    }
    std::vector<int32_t> hsel(nsel_words);
    hipError_t e = hipMemcpyAsync(dfk, fk.data(), sizeof(fu_surf) * n, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(dmk, mk.data(), sizeof(fu_surf) * n, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) e = hipMemcpyAsync(ddist, dd.data(), sizeof(float) * n, hipMemcpyHostToDevice, hs);
    if (e == hipSuccess) st = fu_build_adjacency_matrix(dfk, dmk, ddist, n, mat, s);
    if (e == hipSuccess && !st)
        st = fu_select_subgraphs(mat, n, kLists, kSelect, kMinScore, rank, order, nsel, nsel + kLists,
                                 nsel + 2 * kLists, s);
    if (e == hipSuccess && !st)
        e = hipMemcpyAsync(hsel.data(), nsel, sizeof(int32_t) * nsel_words, hipMemcpyDeviceToHost, hs);
    release();
    const hipError_t es = hipStreamSynchronize(hs);  // also after a failure (host buffers)
    if (e == hipSuccess) e = es;
    if (st) return st;
    if (e != hipSuccess) return api_error(FU_ERR_HIP, "fu_relocalise_graph_candidates: %s", hipGetErrorString(e));
    const int32_t* n_selected = hsel.data();
    const int32_t* n_refined = hsel.data() + kLists;
    const int32_t* lists = hsel.data() + 2 * kLists;
    int kept = 0, first = -1;
    for (int i = 0; i < kLists; ++i)
        if (n_selected[i] >= 3) {
            if (first < 0) first = i;
            ++kept;
        }
    if (kept == 0) return FU_OK;  // "not enough points" (:591-594)
    const int M = std::min(kept, kPoseLists);
    const int R = n_refined[first];
    if (R <= 3) return FU_OK;  // every call skipped (:601-602)
    std::vector<double> src(3 * (size_t)R), ref(3 * (size_t)R);
    for (int k = 0; k < R; ++k) {
        const int p = lists[(size_t)first * kSelect + k];
        for (int c = 0; c < 3; ++c) {
            src[3 * k + c] = (double)fk[p].pos[c];
            ref[3 * k + c] = (double)mk[p].pos[c];
        }
    }
    // the M calls are independent (call c draws from seed + c alone), so calls 1 .. M-1 run on
    // host threads of their own beside call 0; the poses are kept in call order
    double T[kPoseLists][16];
    int32_t good[kPoseLists] = {};
    fu_status cst[kPoseLists] = {};
    auto call = [&](int c) {
        cst[c] = fu_pose_estimate(src.data(), ref.data(), R, iterations, seed + (uint64_t)c, T[c], nullptr, &good[c]);
    };
    std::vector<std::thread> workers;
    for (int c = 1; c < M; ++c) {
        try {
            workers.emplace_back(call, c);
        } catch (const std::system_error&) {
            call(c);  // no thread to be had: the call runs here
        }
    }
    call(0);
    for (std::thread& w : workers) w.join();
    for (int c = 0; c < M; ++c) {
        if (cst[c]) return api_error(cst[c], "fu_relocalise_graph_candidates: PoseEstimate call %d failed", c);
        if (good[c]) rigid_inverse(T[c], poses + 16 * (*n_poses)++);
    }
    return FU_OK;
}
