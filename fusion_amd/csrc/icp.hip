// Point-to-plane ICP: per-pixel association and Jacobian rows, compensated FP32
// reduction of the 29 sums, and (optionally) the FP64 solve and pose update on the device.
//
// Reference: Tracking/Reduction.cu ICPReduce::searchPoint (:227-257), getRow (:259-286),
// operator() (:288-312), Reduce<float,29> (:49-64), CreateMatrix<6,7> (:6-16), ICPStep
// (:319-375); Tracking/Tracking.cc Tracker::ComputeSE3 (:296-409).
//
// Each lane adds the products of its kIcpLanePixels pixels; a wave reduces its 29 sums with one
// reduce-scatter butterfly (wave_reduce_scatter32) and the waves are added through LDS into
// one partial row per work-group (frames with more pixel chunks than work-groups sum their
// chunks per lane with Neumaier-compensated FP32 adds first).
// A one-group kernel then sums the rows with the same compensated tree and (mode 1) unpacks
// A (6x6) and b (6) in FP64, solves by Cholesky, takes the SE(3) exponential and composes
// the pose on one lane: the only FP64 steps of the pipeline.
//
// With the photometric term (ComputeSE3 icpOnly = false) every iteration is three launches:
// the RGB residual pass (RGBReduction::ComputeResidual, Reduction.cu:469-483: one residual
// record per pixel and the integer sums of each work-group), one pass that computes both the
// ICP and the weighted RGB rows of each pixel (each work-group first adds the integer sums
// up to sigma itself) and reduces 64 values per lane, and the final group, which blends
// A_rgb * 1e-4f + A_icp in FP64 before the solve.
#include "fu_device.h"
#include "fu_kernels.h"
#include "fu_track.h"
#include "fu_batch.h"

namespace fu {

struct Kahan {
    float s, c;
    __device__ void add(float v) {
        const float t = s + v;
        if (fabsf(s) >= fabsf(v))
            c += (s - t) + v;
        else
            c += (v - t) + s;
        s = t;
    }
    __device__ float val() const { return s + c; }
};

// ------------------------------------------------------------------ reduction kernel
__device__ __forceinline__ m3 load_m3(FU_GPTR(const float) r) {
    return m3{mk3(r[0], r[1], r[2]), mk3(r[3], r[4], r[5]), mk3(r[6], r[7], r[8])};
}

// smallest float s with sqrtf(s) >= 0.6f, largest d with sqrtf(d) <= 0.1f, largest e with
// (double)sqrtf(e) < 0.1 (found by bisection over the floats with the host's IEEE sqrtf;
// tests/test_oracle.py checks all three)
constexpr float kSine2 = 0x1.70a3d8p-2f;
constexpr float kDist2 = 0x1.47ae16p-7f;
constexpr float kRgbDist2 = 0x1.47ae12p-7f;
constexpr bool angle_thresh_ok() { return kSine2 > 0.35f && kSine2 < 0.37f && kDist2 > 0.0099f && kDist2 < 0.0101f; }

struct IcpPose {
    m3 Rc, Rl, Rli;
    f3 tc, tl;
};

// ICPReduce::searchPoint + getRow (Reduction.cu:227-286) for pixel i (in: the lane has a
// pixel): the 29 products of the row [-n, n x v_l, -n.(v_c - v_l)] (last camera frame) and
// the count 1 into v[O .. O+28] when the pixel has a correspondence, else zeros (kAdd: added
// to v instead).  The reference's tests and arithmetic, as selects instead of early returns
// (the gathers read pixel 0 for lanes that fail the projection), so a wave runs one path; a
// wave without any valid vertex skips the rest.
template <int O, bool kAdd = false, int NV>
__device__ __forceinline__ void icp_row_values(const IcpArgs& a, const IcpPose& P, int i, bool in, float (&v)[NV]) {
    const int cols = a.cols, rows = a.rows;
    const auto vcurr = FU_G(const f4, a.vcurr);
    const auto ncurr = FU_G(const f4, a.ncurr);
    const auto vlast = FU_G(const f4, a.vlast);
    const auto nlast = FU_G(const f4, a.nlast);
    const int ic = in ? i : 0;
    const f3 vcurr_c = xyz(vcurr[ic]);
    bool ok = in && !(isnanf_(vcurr_c.x) || (double)vcurr_c.z < 1e-3);
    if (ballot(ok) == 0ull) {
        if constexpr (!kAdd) {
#pragma unroll
            for (int k = 0; k < 29; ++k) v[O + k] = 0.f;
        }
        return;
    }
    const f3 vcurr_g = add(mv(P.Rc, vcurr_c), P.tc);
    const f3 vcurr_p = mv(P.Rli, sub(vcurr_g, P.tl));
    // Reduction.cu:237-239: `float invz = 1.0 / z`, float products and sum, + 0.5 in double,
    // truncation.  The double quotient rounded to float is the float quotient 1.0f / z (double
    // rounding of a quotient of floats through a format of >= 2*24+2 bits is innocuous).  A lane
    // without a valid vertex divides by 1 (its result is discarded), so only a valid vertex
    // with |z| outside [2^-60, 2^30] takes the general division.
    const float zp = ok ? vcurr_p.z : 1.0f;
    const float azp = fabsf(zp);
    const float invz = (azp >= 0x1p-60f && azp <= 0x1p30f) ? fu_div_refined(1.0f, zp, fu_rcp_refined(zp)) : 1.0f / zp;
    const int u = d2i((double)(vcurr_p.x * invz * a.fx + a.cx) + 0.5);
    const int vv = d2i((double)(vcurr_p.y * invz * a.fy + a.cy) + 0.5);
    ok = ok && !(u < 0 || vv < 0 || u >= cols || vv >= rows);
    const int j = ok ? vv * cols + u : 0;
    const f3 vlast_g = add(mv(P.Rl, xyz(vlast[j])), P.tl);
    const f3 ncurr_c = xyz(ncurr[ic]);
    const f3 ncurr_g = mv(P.Rc, ncurr_c);
    const f3 nlast_c = xyz(nlast[j]);
    const f3 nlast_g = mv(P.Rl, nlast_c);
    const f3 dv = sub(vlast_g, vcurr_g);
    const f3 cr2 = cross(ncurr_g, nlast_g);
    const float dist2 = dot(dv, dv), sine2 = dot(cr2, cr2);
    ok = ok && sine2 < kSine2 && dist2 <= kDist2 && !isnanf_(ncurr_c.x) && !isnanf_(nlast_c.x);
    const f3 nl = mv(P.Rli, nlast_g);
    const f3 vcu = mv(P.Rli, sub(vcurr_g, P.tl));
    const f3 vla = mv(P.Rli, sub(vlast_g, P.tl));
    const f3 mn = neg(nl);
    const f3 cr = cross(nl, vla);
    // a lane without a correspondence contributes +0 to every sum, as the products of a zero row
    const float row[7] = {ok ? mn.x : 0.f, ok ? mn.y : 0.f, ok ? mn.z : 0.f, ok ? cr.x : 0.f,
                          ok ? cr.y : 0.f,  ok ? cr.z : 0.f, ok ? dot(mn, sub(vcu, vla)) : 0.f};
    int k = O;
#pragma unroll
    for (int r = 0; r < 7; ++r)
#pragma unroll
        for (int c = r; c < 7; ++c) {
            if constexpr (kAdd)
                v[k] += row[r] * row[c];
            else
                v[k] = row[r] * row[c];
            ++k;
        }
    if constexpr (kAdd)
        v[O + 28] += ok ? 1.0f : 0.f;
    else
        v[O + 28] = ok ? 1.0f : 0.f;
}

// sums of a work-group of NW waves: each wave's totals into LDS, then one 29-float row
// (v[29..31] must be 0).  Value j of a wave's totals is held by lanes 2j and 2j+1 after the
// reduce-scatter; the waves are added as a pairwise tree.
template <int NW, class Out>
__device__ __forceinline__ void block_rows29(float (&v)[32], float (&wsum)[NW][32], Out out) {
    const float t = wave_reduce_scatter32(v);
    const int lane = lane_id();
    if (!(lane & 1)) wsum[threadIdx.x >> 6][lane >> 1] = t;
    __syncthreads();
    if (threadIdx.x < 29) {
        const int j = threadIdx.x;
        float w[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) w[k] = wsum[k][j];
#pragma unroll
        for (int h = NW / 2; h >= 1; h /= 2)
#pragma unroll
            for (int k = 0; k < h; ++k) w[k] = w[2 * k] + w[2 * k + 1];
        out[j] = w[0];
    }
}

// the same 29 values into v[O .. O+28]
template <int O, int NV>
__device__ __forceinline__ void row_products(const float (&row)[7], float (&v)[NV]) {
    int k = O;
#pragma unroll
    for (int r = 0; r < 7; ++r)
#pragma unroll
        for (int c = r; c < 7; ++c) v[k++] = row[r] * row[c];
    v[O + 28] = 1.0f;
}

// kOne: the grid has a lane for every pixel (one pixel per lane, no per-lane sums); otherwise
// lanes loop over pixels with compensated sums seeded exactly by their first pixel.  Both
// give the same result when every lane has at most one pixel.
template <bool kOne>
__device__ __forceinline__ void icp_body(WG g, const IcpArgs& a) {
    const auto st = FU_G(const TrackState, a.st);
    if (a.mode != 0 && st->status != TS_RUNNING) return;  // uniform: a failed frame skips
    IcpPose P;
    P.Rc = load_m3(st->Rcurr);
    P.Rl = load_m3(st->Rlast);
    P.Rli = load_m3(st->RlastInv);
    P.tc = mk3(st->tcurr[0], st->tcurr[1], st->tcurr[2]);
    P.tl = mk3(st->tlast[0], st->tlast[1], st->tlast[2]);
    const int N = a.cols * a.rows;
    __shared__ float wsum[kIcpThreads / 64][32];
    float v[32];
    v[29] = v[30] = v[31] = 0.f;
    // a work-group takes chunks of kIcpPixels = kIcpLanePixels * kIcpThreads pixels: lane t
    // adds the products of pixels t, t + kIcpThreads, ... of the chunk, in that order, before
    // the wave reduction
    const int i0 = g.x * kIcpPixels + threadIdx.x;
    icp_row_values<0>(a, P, i0, i0 < N, v);
#pragma unroll
    for (int k = 1; k < kIcpLanePixels; ++k) icp_row_values<0, true>(a, P, i0 + k * kIcpThreads, i0 + k * kIcpThreads < N, v);
    if constexpr (!kOne) {
        // more chunks than work-groups: compensated per-lane sums over the chunks, seeded
        // exactly by the first (so both variants agree when every group has one chunk)
        Kahan acc[29];
#pragma unroll
        for (int j = 0; j < 29; ++j) acc[j] = Kahan{v[j], 0.f};
        const int stride = g.nx * kIcpPixels;
        for (int c = g.x * kIcpPixels + stride; c < N; c += stride) {
            const int i = c + threadIdx.x;
            icp_row_values<0>(a, P, i, i < N, v);
#pragma unroll
            for (int k = 1; k < kIcpLanePixels; ++k) icp_row_values<0, true>(a, P, i + k * kIcpThreads, i + k * kIcpThreads < N, v);
#pragma unroll
            for (int j = 0; j < 29; ++j) acc[j].add(v[j]);
        }
#pragma unroll
        for (int j = 0; j < 29; ++j) v[j] = acc[j].val();
    }
    block_rows29(v, wsum, FU_G(float, a.partials) + g.x * 32);
}
template <bool kOne>
__global__ __launch_bounds__(kIcpThreads) void k_icp(IcpArgs a) {
    icp_body<kOne>(hw_wg(), a);
}

// One work-group: sums the partial rows of k_icp (compensated per lane, then the same
// tree), stores the 29 sums and, in mode 1, runs the FP64 update on one lane.  Running it as
// its own kernel makes the partials visible through the kernel boundary; a last-group ticket
// inside k_icp would need a device-scope release per work-group, which on MI355X writes back
// the work-group's L2 (one per XCD) every time.
// The tracking state of a final kernel is staged through LDS: every lane loads a part of it
// at the start (overlapping the partial-row loads), lane 0 runs the serial FP64 update on the
// LDS copy, and every lane writes a part back, instead of one lane's long chains of global
// loads and stores.
static_assert(sizeof(TrackState) % 4 == 0, "dword-sized state");
constexpr int kStateWords = (int)(sizeof(TrackState) / 4);
__device__ __forceinline__ void state_to_lds(const TrackState* st, TrackState* sst) {
    const auto src = reinterpret_cast<const unsigned*>(st);
    auto dst = reinterpret_cast<unsigned*>(sst);
    for (int i = threadIdx.x; i < kStateWords; i += blockDim.x) dst[i] = src[i];
}
__device__ __forceinline__ void state_from_lds(const TrackState* sst, TrackState* st) {
    const auto src = reinterpret_cast<const unsigned*>(sst);
    auto dst = reinterpret_cast<unsigned*>(st);
    for (int i = threadIdx.x; i < kStateWords; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ void icp_final_body(const IcpArgs& a, int nparts) {
    TrackState* st = a.st;
    // the state and the partial rows are loaded together; a frame that is not running leaves
    // (uniformly) before anything is written
    __shared__ TrackState sst;
    state_to_lds(st, &sst);
    __shared__ float wsum[kIcpFinalThreads / 64][32];
    __shared__ float total[32];
    Kahan fin[29];
#pragma unroll
    for (int j = 0; j < 29; ++j) fin[j] = Kahan{0.f, 0.f};
    for (int g = threadIdx.x; g < nparts; g += blockDim.x) {
        const float4* row = reinterpret_cast<const float4*>(a.partials + g * 32);
        float v[32];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 t = row[q];
            v[4 * q + 0] = t.x;
            v[4 * q + 1] = t.y;
            v[4 * q + 2] = t.z;
            v[4 * q + 3] = t.w;
        }
#pragma unroll
        for (int j = 0; j < 29; ++j) fin[j].add(v[j]);
    }
    __syncthreads();
    if (a.mode != 0 && sst.status != TS_RUNNING) return;
    float v[32];
#pragma unroll
    for (int j = 0; j < 29; ++j) v[j] = fin[j].val();
    v[29] = v[30] = v[31] = 0.f;
    block_rows29(v, wsum, total);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j < 29; ++j) sst.sums[j] = total[j];
        if (a.mode == 1) icp_update(&sst, total, nullptr, a.min_count, a.final_iter, a.thresh);
    }
    __syncthreads();
    state_from_lds(&sst, st);
}
__global__ __launch_bounds__(kIcpFinalThreads) void k_icp_final(IcpArgs a, int nparts) { icp_final_body(a, nparts); }

// ------------------------------------------------------------------ photometric term
struct RgbPose {
    m3 Rc, Rci, Rl, Rli;
    f3 tc, tl;
};
__device__ __forceinline__ RgbPose rgb_pose(const TrackState* st_) {
    const auto st = FU_G(const TrackState, st_);
    RgbPose P;
    P.Rc = load_m3(st->Rcurr);
    P.Rci = load_m3(st->RcurrInv);
    P.Rl = load_m3(st->Rlast);
    P.Rli = load_m3(st->RlastInv);
    P.tc = mk3(st->tcurr[0], st->tcurr[1], st->tcurr[2]);
    P.tl = mk3(st->tlast[0], st->tlast[1], st->tlast[2]);
    return P;
}

// The pose-independent part of RGBReduction::FindCorresp (Reduction.cu:420-433) for pixel
// (x, y): inside the 5-pixel margin, the 4x4 window [y-2, y+2) x [x-2, x+2) non-zero (the
// margin keeps it in the image, so the reference's clamps are inactive), a non-zero gradient
// (sqrtf(gx * gx + gy * gy) > 0: the integer sum is positive unless both are zero) and a valid
// vertex
__device__ __forceinline__ bool rgb_eligible(const IcpArgs& a, int x, int y) {
    const int cols = a.cols, rows = a.rows;
    if (!(x >= 5 && x < cols - 5 && y >= 5 && y < rows - 5)) return false;
    const auto inext = FU_G(const unsigned char, a.rgb.inext);
    bool valid = true;
#pragma unroll
    for (int u = -2; u < 2; ++u)
#pragma unroll
        for (int v = -2; v < 2; ++v) valid &= inext[(y + u) * cols + x + v] > 0;
    if (!valid) return false;
    const int i = y * cols + x;
    const int gx = FU_G(const short, a.rgb.dx)[i], gy = FU_G(const short, a.rgb.dy)[i];
    if (gx == 0 && gy == 0) return false;
    return !isnanf_(FU_G(const f4, a.vcurr)[i].x);
}

// the rest of FindCorresp (Reduction.cu:434-463) for an eligible pixel i: the last-frame
// pixel j and the intensity difference, or false
__device__ __forceinline__ bool rgb_corresp(const IcpArgs& a, const RgbPose& P, int i, int& j, int& diff) {
    const int cols = a.cols, rows = a.rows;
    const f4 vc = FU_G(const f4, a.vcurr)[i];
    const f3 vg = add(mv(P.Rc, xyz(vc)), P.tc);
    const f3 vl = mv(P.Rli, sub(vg, P.tl));
    // __float2int_rd(fx * x / z + cx + 0.5): float quotient and sum, 0.5 added in double,
    // the double rounded to float, then floor
    const int u = f2i(floorf((float)((double)(a.fx * vl.x / vl.z + a.cx) + 0.5)));
    const int v = f2i(floorf((float)((double)(a.fy * vl.y / vl.z + a.cy) + 0.5)));
    if (!(u >= 0 && v >= 0 && u < cols && v < rows)) return false;
    const int k = v * cols + u;
    const f4 la = FU_G(const f4, a.vlast)[k];
    const int lk = FU_G(const unsigned char, a.rgb.ilast)[k];
    // (double)norm4(...) < 0.1 on the squared norm: see kRgbDist2
    const f4 d4 = sub4(la, vc);
    if (!(!isnanf_(la.x) && dot4(d4, d4) <= kRgbDist2 && lk != 0)) return false;
    j = k;
    diff = (int)FU_G(const unsigned char, a.rgb.inext)[i] - lk;
    return true;
}

// RGBReduction::GetRow (Reduction.cu:485-520) of a valid residual of pixel i
__device__ __forceinline__ void rgb_row(const IcpArgs& a, const RgbPose& P, int i, int j, int diff, float sigma,
                                        float (&row)[7]) {
    float w = sigma + (float)abs(diff);
    w = (double)w < 1e-3 ? 1.0f : 1.0f / w;
    if ((double)sigma < 1e-6) w = 1.0f;
    const f4 la = FU_G(const f4, a.vlast)[j];
    const f3 point = mv(P.Rci, sub(add(mv(P.Rl, xyz(la)), P.tl), P.tc));
    // (float)((double)s / 9.0) in the reference: for every short s this equals the correctly
    // rounded float quotient (s / 9 has a period-6 binary expansion, so rounding to double
    // first never lands on a float rounding boundary; checked for all 65536 values)
    const float gx = (float)FU_G(const short, a.rgb.dx)[i] / 9.0f;
    const float gy = (float)FU_G(const short, a.rgb.dy)[i] / 9.0f;
    const float invz = 1.0f / point.z;
    f3 left;
    left.x = gx * a.fx * invz;
    left.y = gy * a.fy * invz;
    left.z = -(left.x * point.x + left.y * point.y) * invz;
    const f3 c = cross(point, left);
    row[0] = w * left.x;
    row[1] = w * left.y;
    row[2] = w * left.z;
    row[3] = w * c.x;
    row[4] = w * c.y;
    row[5] = w * c.z;
    row[6] = -w * (float)diff;
}

// sum over the work-group of one int per lane (all lanes get it); `part` holds a slot per wave
__device__ __forceinline__ unsigned block_sum_u32(unsigned v, unsigned* part) {
    v = wave_sum_u32(v);
    if (lane_id() == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
    return t;
}

// list building, pass 1: the eligibility of each pixel as one 64-bit word per wave (kept in
// the record buffer until pass 2) and the eligible pixels of each 1024-pixel work-group
__device__ __forceinline__ void rgb_count_body(WG g, const IcpArgs& a) {
    const int N = a.cols * a.rows;
    const int i = g.x * kIcpThreads + threadIdx.x;
    bool e = false;
    if (i < N) {
        const int y = i / a.cols;
        e = rgb_eligible(a, i - y * a.cols, y);
    }
    const unsigned long long m = ballot(e);
    __shared__ unsigned part[kIcpThreads / 64];
    if (lane_id() == 0) {
        part[threadIdx.x >> 6] = (unsigned)__popcll(m);
        FU_G(unsigned long long, a.rgb.rec)[g.x * (kIcpThreads / 64) + (threadIdx.x >> 6)] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < kIcpThreads / 64; ++w) t += part[w];
        FU_G(int, a.rgb.wgcnt)[g.x] = (int)t;
    }
}
// pass 2: each work-group writes its eligible pixels at the prefix of the counts before it
// (pixel order); the last work-group stores the total
__device__ __forceinline__ void rgb_scatter_body(WG g, const IcpArgs& a) {
    __shared__ unsigned part[kIcpThreads / 64];
    __shared__ unsigned wbase[kIcpThreads / 64];
    const auto wgcnt = FU_G(const int, a.rgb.wgcnt);
    unsigned before = 0;
    for (int t = threadIdx.x; t < g.x; t += blockDim.x) before += (unsigned)wgcnt[t];
    const unsigned base = block_sum_u32(before, part);
    const int wave = threadIdx.x >> 6;
    const unsigned long long m = FU_G(const unsigned long long, a.rgb.rec)[g.x * (kIcpThreads / 64) + wave];
    __syncthreads();  // part is reused
    if (lane_id() == 0) part[wave] = (unsigned)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = base;
        for (int w = 0; w < kIcpThreads / 64; ++w) {
            wbase[w] = t;
            t += part[w];
        }
        if (g.x == g.nx - 1) *FU_G(int, a.rgb.count) = (int)t;
    }
    __syncthreads();
    if ((m >> lane_id()) & 1ull) FU_G(int, a.rgb.list)[wbase[wave] + lane_rank(m)] = g.x * kIcpThreads + threadIdx.x;
}

// residual pass over the list: record (j << 9) | (diff + 256) or -1 per entry, and the
// work-group's {count, sum diff^2}; one entry per lane
__device__ __forceinline__ void rgb_resid_body(WG g, const IcpArgs& a) {
    const auto st = FU_G(const TrackState, a.st);
    if (a.mode != 0 && st->status != TS_RUNNING) return;  // uniform: a failed frame skips
    const int n = *FU_G(const int, a.rgb.count);
    const int k = g.x * kIcpThreads + threadIdx.x;
    int rec = -1, diff = 0;
    if (g.x * kIcpThreads < n) {  // uniform
        const RgbPose P = rgb_pose(a.st);
        if (k < n) {
            int j;
            if (rgb_corresp(a, P, FU_G(const int, a.rgb.list)[k], j, diff)) rec = (j << 9) | (diff + 256);
            FU_G(int, a.rgb.rec)[k] = rec;
        }
    }
    const bool ok = rec >= 0;
    const unsigned c = (unsigned)__popcll(ballot(ok));
    const unsigned q = wave_sum_u32(ok ? (unsigned)(diff * diff) : 0u);
    __shared__ unsigned part[kIcpThreads / 64][2];
    if (lane_id() == 0) {
        part[threadIdx.x >> 6][0] = c;
        part[threadIdx.x >> 6][1] = q;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned t = 0;
        for (int w = 0; w < kIcpThreads / 64; ++w) t += part[w][threadIdx.x];
        FU_G(int, a.rgb.corr)[g.x * 2 + threadIdx.x] = (int)t;
    }
}
__global__ __launch_bounds__(kIcpThreads) void k_rgb_resid(IcpArgs a) { rgb_resid_body(hw_wg(), a); }

// totals of the residual pass's {count, sum diff^2} over its `nparts` work-groups (32-bit,
// wrapping, as the reference's int sums), and sigma = sqrt(sum / count) (Reduction.cu:605)
__device__ __forceinline__ float rgb_sigma(const IcpArgs& a, int nparts, int* corr_out) {
    __shared__ unsigned part[kIcpThreads / 64][2];
    const auto corr = FU_G(const int, a.rgb.corr);
    unsigned c = 0, q = 0;
    for (int t = threadIdx.x; t < nparts; t += blockDim.x) {
        c += (unsigned)corr[2 * t];
        q += (unsigned)corr[2 * t + 1];
    }
    c = wave_sum_u32(c);
    q = wave_sum_u32(q);
    if (lane_id() == 0) {
        part[threadIdx.x >> 6][0] = c;
        part[threadIdx.x >> 6][1] = q;
    }
    __syncthreads();
    unsigned tc = 0, tq = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        tc += part[w][0];
        tq += part[w][1];
    }
    const int ci = (int)tc, qi = (int)tq;
    if (corr_out && threadIdx.x == 0) {
        corr_out[0] = ci;
        corr_out[1] = qi;
    }
    return sqrtf((float)qi / (float)(ci == 0 ? 1 : ci));
}

// The 29 products of the photometric row of list entry i (n entries) into v[O .. O+28]
// (kAdd: added), zeros for a lane without a residual record.  The record of an entry outside
// the list is never read; such lanes evaluate rgb_row on entry 0's pixel and discard it.
template <int O, bool kAdd = false>
__device__ __forceinline__ void rgb_row_values(const IcpArgs& a, const RgbPose& Q, int i, int n, float sigma,
                                               float (&v)[64]) {
    const int rec = i < n ? FU_G(const int, a.rgb.rec)[i] : -1;
    const bool okr = rec >= 0;
    if (ballot(okr) == 0ull) {
        if constexpr (!kAdd) {
#pragma unroll
            for (int k = 0; k < 29; ++k) v[O + k] = 0.f;
        }
        return;
    }
    float row[7];
    rgb_row(a, Q, okr ? FU_G(const int, a.rgb.list)[i] : 0, okr ? rec >> 9 : 0, okr ? (rec & 511) - 256 : 0, sigma,
            row);
#pragma unroll
    for (int k = 0; k < 7; ++k) row[k] = okr ? row[k] : 0.f;
    int k = O;
#pragma unroll
    for (int r = 0; r < 7; ++r)
#pragma unroll
        for (int c = r; c < 7; ++c) {
            if constexpr (kAdd)
                v[k] += row[r] * row[c];
            else
                v[k] = row[r] * row[c];
            ++k;
        }
    if constexpr (kAdd)
        v[O + 28] += okr ? 1.0f : 0.f;
    else
        v[O + 28] = okr ? 1.0f : 0.f;
}

// ICP and photometric rows of a kRgbPixels chunk (two pixels and two list entries per lane,
// added before the wave reduction): values 0-28 are the ICP sums, 32-60 the RGB sums; the
// work-group's 64 totals go to partials[g.x * 64 ..].  nresid: work-groups of the residual
// pass, whose integer sums give sigma.
__device__ __forceinline__ void icp_rgb_body(WG g, const IcpArgs& a, int nresid) {
    const auto st = FU_G(const TrackState, a.st);
    if (a.mode != 0 && st->status != TS_RUNNING) return;
    IcpPose P;
    P.Rc = load_m3(st->Rcurr);
    P.Rl = load_m3(st->Rlast);
    P.Rli = load_m3(st->RlastInv);
    P.tc = mk3(st->tcurr[0], st->tcurr[1], st->tcurr[2]);
    P.tl = mk3(st->tlast[0], st->tlast[1], st->tlast[2]);
    const RgbPose Q = rgb_pose(a.st);
    const float sigma = rgb_sigma(a, nresid, nullptr);
    const int N = a.cols * a.rows;
    __shared__ float wsum[kIcpThreads / 64][64];
    float v[64];
    v[29] = v[30] = v[31] = 0.f;
    v[61] = v[62] = v[63] = 0.f;
    const int i0 = g.x * kRgbPixels + threadIdx.x, i1 = i0 + kIcpThreads;
    if (a.ncurr) {
        icp_row_values<0>(a, P, i0, i0 < N, v);
        icp_row_values<0, true>(a, P, i1, i1 < N, v);
    } else {  // no normals: RGB rows only
#pragma unroll
        for (int k = 0; k < 29; ++k) v[k] = 0.f;
    }
    const int n = *FU_G(const int, a.rgb.count);
    if (g.x * kRgbPixels < n) {
        rgb_row_values<32>(a, Q, i0, n, sigma, v);
        rgb_row_values<32, true>(a, Q, i1, n, sigma, v);
    } else {
#pragma unroll
        for (int k = 32; k < 61; ++k) v[k] = 0.f;
    }
    const float t = wave_reduce_scatter64(v);
    wsum[threadIdx.x >> 6][lane_id()] = t;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int j = threadIdx.x;
        float w[kIcpThreads / 64];
#pragma unroll
        for (int k = 0; k < kIcpThreads / 64; ++k) w[k] = wsum[k][j];
#pragma unroll
        for (int h = kIcpThreads / 128; h >= 1; h /= 2)
#pragma unroll
            for (int k = 0; k < h; ++k) w[k] = w[2 * k] + w[2 * k + 1];
        FU_G(float, a.partials)[g.x * 64 + j] = w[0];
    }
}
__global__ __launch_bounds__(kIcpThreads) void k_icp_rgb(IcpArgs a, int nresid) { icp_rgb_body(hw_wg(), a, nresid); }

// final group with the photometric term: each wave adds every fourth of the nparts partial
// rows (compensated, one value per lane), the waves are added pairwise, then the FP64 update;
// sigma and the correspondence counts from the nresid residual-pass partials
__device__ __forceinline__ void icp_rgb_final_body(const IcpArgs& a, int nparts, int nresid) {
    TrackState* st = a.st;
    __shared__ TrackState sst;
    state_to_lds(st, &sst);
    constexpr int NW = kIcpFinalThreads / 64;
    __shared__ float wsum[NW][64];
    __shared__ float total[64];
    __shared__ int corr[2];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    const float sigma = rgb_sigma(a, nresid, corr);
    // rows w, w + NW, w + 2 NW, ... in that order; eight loads in flight ahead of the adds
    Kahan k{0.f, 0.f};
    const auto parts = FU_G(const float, a.partials);
    for (int g = w; g < nparts; g += 8 * NW) {
        float x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = (g + u * NW < nparts) ? parts[(g + u * NW) * 64 + lane] : 0.f;
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (g + u * NW < nparts) k.add(x[u]);
    }
    __syncthreads();
    if (a.mode != 0 && sst.status != TS_RUNNING) return;  // uniform: nothing is written
    wsum[w][lane] = k.val();
    __syncthreads();
    if (threadIdx.x < 64) {
        float x[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) x[q] = wsum[q][threadIdx.x];
#pragma unroll
        for (int h = NW / 2; h >= 1; h /= 2)
#pragma unroll
            for (int q = 0; q < h; ++q) x[q] = x[2 * q] + x[2 * q + 1];
        total[threadIdx.x] = x[0];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = 0; j < 29; ++j) {
            sst.sums[j] = total[j];
            sst.rgb_sums[j] = total[32 + j];
        }
        sst.rgb_corresp[0] = corr[0];
        sst.rgb_corresp[1] = corr[1];
        sst.sigma = sigma;
        if (a.mode == 1) icp_update(&sst, total, total + 32, a.min_count, a.final_iter, a.thresh);
    }
    __syncthreads();
    state_from_lds(&sst, st);
}
__global__ __launch_bounds__(kIcpFinalThreads) void k_icp_rgb_final(IcpArgs a, int nparts, int nresid) {
    icp_rgb_final_body(a, nparts, nresid);
}

__global__ __launch_bounds__(kIcpThreads) void k_rgb_count(IcpArgs a) { rgb_count_body(hw_wg(), a); }
__global__ __launch_bounds__(kIcpThreads) void k_rgb_scatter(IcpArgs a) { rgb_scatter_body(hw_wg(), a); }
void launch_rgb_lists(const IcpArgs* a, const int* nblocks, int nlevels, hipStream_t s) {
    for (int l = 0; l < nlevels; ++l)
        if (a[l].use_rgb) {
            hipLaunchKernelGGL(k_rgb_count, dim3(nblocks[l]), dim3(kIcpThreads), 0, s, a[l]);
            hipLaunchKernelGGL(k_rgb_scatter, dim3(nblocks[l]), dim3(kIcpThreads), 0, s, a[l]);
        }
}

void launch_icp(const IcpArgs& a, int nblocks, hipStream_t s) {
    if (a.use_rgb) {  // one pixel per lane (the caller sizes nblocks to cover the level)
        hipLaunchKernelGGL(k_rgb_resid, dim3(nblocks), dim3(kIcpThreads), 0, s, a);
        const int pairs = (nblocks + 1) / 2;  // kRgbPixels chunks of the combined pass
        hipLaunchKernelGGL(k_icp_rgb, dim3(pairs), dim3(kIcpThreads), 0, s, a, nblocks);
        hipLaunchKernelGGL(k_icp_rgb_final, dim3(1), dim3(kIcpFinalThreads), 0, s, a, pairs, nblocks);
        return;
    }
    if ((long long)nblocks * kIcpPixels >= (long long)a.cols * a.rows)
        hipLaunchKernelGGL(k_icp<true>, dim3(nblocks), dim3(kIcpThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_icp<false>, dim3(nblocks), dim3(kIcpThreads), 0, s, a);
    hipLaunchKernelGGL(k_icp_final, dim3(1), dim3(kIcpFinalThreads), 0, s, a, nblocks);
}

// before tracking: a lost tracker does not track; otherwise the start of ComputeSE3
// (frame_begin)
__device__ __forceinline__ void track_begin_body(TrackState* st) {
    if (threadIdx.x != 0) return;
    frame_begin(st);
}
__global__ void k_track_begin(TrackState* st) { track_begin_body(st); }
// after tracking: the state machine decides what the frame does, and a swapped frame's pose
// becomes the last pose (frame_commit)
__device__ __forceinline__ void track_commit_body(TrackState* st) {
    if (threadIdx.x != 0) return;
    frame_commit(st);
}
__global__ void k_track_commit(TrackState* st) { track_commit_body(st); }
void launch_track_begin(TrackState* st, hipStream_t s) { hipLaunchKernelGGL(k_track_begin, dim3(1), dim3(64), 0, s, st); }
void launch_track_commit(TrackState* st, hipStream_t s) { hipLaunchKernelGGL(k_track_commit, dim3(1), dim3(64), 0, s, st); }

// ------------------------------------------------------------------ batched launches
// The iteration schedule is the union of the sessions' schedules: a session runs iteration
// `iter` of `level` only if it has that many, and its last iteration is final.
__device__ __forceinline__ bool icp_slot_args(const BatchLaunch& b, int s, IcpArgs& a, int& nblocks) {
    if (slot_first(b, s)) return false;
    const SessDesc& d = b.desc[b.sess[s]];
    const int l = b.level;
    if (b.iter >= d.iters[l]) return false;
    a = cload(d.icp[l]);
    a.final_iter = (l == d.last_level && b.iter == d.iters[l] - 1) ? 1 : 0;
    nblocks = d.icp_blocks[l];
    return true;
}

template <bool kOne>
__global__ __launch_bounds__(kIcpThreads) void kb_icp(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, batch_wg(b), g);
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, s, a, nb)) return;
    icp_body<kOne>(g, a);
}

__global__ __launch_bounds__(kIcpFinalThreads) void kb_icp_final(BatchLaunch b) {
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, blockIdx.x, a, nb)) return;
    icp_final_body(a, nb);
}

__global__ __launch_bounds__(kIcpThreads) void kb_rgb_count(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, s, a, nb)) return;
    rgb_count_body(g, a);
}
__global__ __launch_bounds__(kIcpThreads) void kb_rgb_scatter(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, s, a, nb)) return;
    rgb_scatter_body(g, a);
}
__global__ __launch_bounds__(kIcpThreads) void kb_rgb_resid(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, batch_wg(b), g);
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, s, a, nb)) return;
    rgb_resid_body(g, a);
}
__global__ __launch_bounds__(kIcpThreads) void kb_icp_rgb(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, batch_wg(b), g);
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, s, a, nb)) return;
    icp_rgb_body(g, a, nb);
}
__global__ __launch_bounds__(kIcpFinalThreads) void kb_icp_rgb_final(BatchLaunch b) {
    IcpArgs a;
    int nb;
    if (!icp_slot_args(b, blockIdx.x, a, nb)) return;
    icp_rgb_final_body(a, (nb + 1) / 2, nb);
}

__global__ void kb_track_begin(BatchLaunch b) {
    const int s = blockIdx.x;
    if (slot_first(b, s)) return;
    track_begin_body(b.desc[b.sess[s]].st);
}
__global__ void kb_track_commit(BatchLaunch b) { track_commit_body(b.desc[b.sess[blockIdx.x]].st); }

// tracking result and map counters of every slot, into out[slot]
__global__ void kb_results(BatchLaunch b, FrameBrief* out) {
    const int s = blockIdx.x;
    const SessDesc& d = b.desc[b.sess[s]];
    const TrackState* st = d.st;
    FrameBrief* o = out + s;
    const int t = threadIdx.x;
    if (t < 16) o->pose[t] = st->next_pose[t];
    if (t == 16) o->status = st->status;
    if (t == 17) o->icp_error = st->icp_error;
    if (t == 18) o->icp_count = st->count;
    if (t == 19) o->n_visible = d.m.ctr[C_VISIBLE];
    if (t == 20) o->n_alloc = min(d.m.ctr[C_ALLOC], d.m.num_blocks);
    if (t == 21) o->flags = d.m.ctr[C_FLAGS];
    if (t == 22) o->rgb_count = d.use_rgb ? st->rgb_count : 0;
}
void launch_b_results(BatchLaunch& b, FrameBrief* out, hipStream_t s) {
    hipLaunchKernelGGL(kb_results, dim3(b.n), dim3(64), 0, s, b, out);
}

void launch_b_track_begin(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    (void)h;
    hipLaunchKernelGGL(kb_track_begin, dim3(b.n), dim3(64), 0, s, b);
}
void launch_b_icp(BatchLaunch& b, const SessDesc* h, int level, int iter, hipStream_t s) {
    b.level = level;
    b.iter = iter;
    const int grid = batch_layout(b, h, BT_ICP, level, 0);
    b.xcd = xcd_order(FU_STAGE_ICP);
    if (h[b.sess[0]].use_rgb) {  // the sessions of a batch share the tracking mode
        hipLaunchKernelGGL(kb_rgb_resid, dim3(grid), dim3(kIcpThreads), 0, s, b);
        const int pairs = batch_layout(b, h, BT_ICP_PAIRS, level, 0);
        hipLaunchKernelGGL(kb_icp_rgb, dim3(pairs), dim3(kIcpThreads), 0, s, b);
        hipLaunchKernelGGL(kb_icp_rgb_final, dim3(b.n), dim3(kIcpFinalThreads), 0, s, b);
        return;
    }
    bool one = true;
    for (int i = 0; i < b.n; ++i) {
        const SessDesc& d = h[b.sess[i]];
        one = one && (long long)d.icp_blocks[level] * kIcpPixels >= (long long)d.lc[level] * d.lr[level];
    }
    if (one)
        hipLaunchKernelGGL(kb_icp<true>, dim3(grid), dim3(kIcpThreads), 0, s, b);
    else
        hipLaunchKernelGGL(kb_icp<false>, dim3(grid), dim3(kIcpThreads), 0, s, b);
    hipLaunchKernelGGL(kb_icp_final, dim3(b.n), dim3(kIcpFinalThreads), 0, s, b);
}
// the eligible-pixel lists of every level (iteration 0 of each level: a session without
// iterations at a level builds none)
void launch_b_rgb_lists(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    for (int l = 0; l < 3; ++l) {
        b.level = l;
        b.iter = 0;
        const int grid = batch_layout(b, h, BT_ICP, l, 0);
        hipLaunchKernelGGL(kb_rgb_count, dim3(grid), dim3(kIcpThreads), 0, s, b);
        hipLaunchKernelGGL(kb_rgb_scatter, dim3(grid), dim3(kIcpThreads), 0, s, b);
    }
}

void launch_b_track_commit(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    (void)h;
    hipLaunchKernelGGL(kb_track_commit, dim3(b.n), dim3(64), 0, s, b);
}

}  // namespace fu
