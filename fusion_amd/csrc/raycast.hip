// Raycast of the TSDF model: z-range tiles, then a per-pixel ray march.
//
// Reference: Mapping/RenderScene.cu.  Projection::projectBlock (:54-104) bounds each visible
// block on a grid of 8x8-pixel tiles; createRenderingBlockList/fillBlocks (:106-179) split the
// bound into 16x16-tile fragments and min/max the fragment z-ranges into the tile maps;
// Rendering::operator() (:354-440) marches each pixel between its tile's z-range.
//
// The fragment list is only a vehicle for the tile maps, so here one thread per visible block
// writes the union of its fragments straight into the maps with integer atomics (the ranges
// are positive floats, whose IEEE bit patterns order like integers).  The union reproduces the
// reference's fragment loop as written (inner bound `x < ny`, lower-right corner without the
// fragment offset, RenderScene.cu:109-116): the covered tiles are
//   [UL.x, min(LR.x, UL.x + e)] x [UL.y, min(LR.y, UL.y + e)],  e = 16 if ny >= 2 else 15.
// Tile maps are ceil(W/8) x ceil(H/8) (80x60 at 640x480, as RenderScene.cu:245-251).
#include "fu_device.h"
#include "fu_kernels.h"
#include "fu_batch.h"

namespace fu {

__device__ __forceinline__ void zrange_init_body(WG g, const TrackState* st, unsigned* zmin, unsigned* zmax, int n) {
    if (st && !(FU_G(const TrackState, st)->gates & GATE_RENDER)) return;
    const unsigned hundred = __float_as_uint(100.0f);
    for (int i = g.x * blockDim.x + threadIdx.x; i < n; i += g.nx * blockDim.x) {
        zmin[i] = hundred;
        zmax[i] = 0u;
    }
}
__global__ void k_zrange_init(const TrackState* st, unsigned* zmin, unsigned* zmax, int n) {
    zrange_init_body(hw_wg(), st, zmin, zmax, n);
}

// exclusive prefix sum of `v` over the lanes of a wave, and the wave total in `total`
__device__ __forceinline__ unsigned wave_exclusive_scan(unsigned v, unsigned& total) {
    const int lane = lane_id();
    unsigned x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

// Tile maps staged in LDS when they fit: every block's min/max goes to the work-group's copy,
// which is flushed into the global maps once, only where it changed.  Min and max are exact
// and order-free, so the maps are the same as with direct global atomics.
constexpr int kZTilesLds = 6144;  // 48 KB of LDS for both maps; 640x480 needs 4800 tiles
constexpr int kProjectGroups = 16;  // work-groups per session with LDS tile maps

template <bool kLds>
__device__ __forceinline__ void project_body(WG g, const MapView& m, const ZRangeArgs& a) {
    PoseF pose = a.pose;
    if (a.st && !pose_from_state(a.st, pose)) return;
    // a caller-given count never exceeds the list the visibility pass built
    const unsigned built = min(FU_G(const unsigned, m.ctr)[C_VISIBLE], m.num_blocks);
    const unsigned nvis = a.n_visible_override >= 0 ? min((unsigned)a.n_visible_override, built) : built;
    const int zc = a.zcols, zr = a.zrows;
    const int nt = zc * zr;
    __shared__ unsigned lmin[kLds ? kZTilesLds : 1], lmax[kLds ? kZTilesLds : 1];
    if constexpr (kLds) {
        for (int t = threadIdx.x; t < nt; t += blockDim.x) {
            lmin[t] = 0xFFFFFFFFu;
            lmax[t] = 0u;
        }
        __syncthreads();
    }
    // uniform trip count per wave: the fragment counter is advanced once per wave
    const unsigned stride = g.nx * blockDim.x;
    for (unsigned i0 = g.x * blockDim.x + (threadIdx.x & ~63u); i0 < nvis; i0 += stride) {
        const unsigned i = i0 + lane_id();
        bool live = i < nvis;
        int ulx = zc, uly = zr, lrx = -1, lry = -1;
        float zlo = a.depth_max, zhi = a.depth_min;
        if (live) {
            const i3 pos = unpack_key(FU_G(const unsigned long long, m.visible_key)[i]);
            // projectBlock (RenderScene.cu:54-104)
            for (int c = 0; c < 8; ++c) {
                const i3 q = mki3(pos.x + (c & 1), pos.y + ((c >> 1) & 1), pos.z + ((c >> 2) & 1));
                // int3 * uint * float: the integer product first, then to float (VectorMath.h:219-233)
                f3 p = i3f(mki3(q.x * 8, q.y * 8, q.z * 8), m.voxel_size);
                p = mv(pose.Rinv, sub(p, pose.t));
                if ((double)p.z < 2e-1) continue;
                const float px = (a.fx * p.x / p.z + a.cx) / 8.0f;
                const float py = (a.fy * p.y / p.z + a.cy) / 8.0f;
                const float flx = floorf(px), clx = ceilf(px), fly = floorf(py), cly = ceilf(py);
                if ((float)ulx > flx) ulx = (short)f2i(flx);
                if ((float)lrx < clx) lrx = (short)f2i(clx);
                if ((float)uly > fly) uly = (short)f2i(fly);
                if ((float)lry < cly) lry = (short)f2i(cly);
                if (zlo > p.z) zlo = p.z;
                if (zhi < p.z) zhi = p.z;
            }
            if (ulx < 0) ulx = 0;
            if (uly < 0) uly = 0;
            if (lrx >= zc) lrx = zc - 1;
            if (lry >= zr) lry = zr - 1;
            if (ulx > lrx || uly > lry) live = false;
            if (zlo < a.depth_min) zlo = a.depth_min;
            if (zhi < a.depth_min) live = false;
        }
        int nx = 0, ny = 0;
        unsigned need = 0;
        if (live) {
            nx = (int)ceilf(((float)lrx - (float)ulx + 1.0f) / 16.0f);
            ny = (int)ceilf(((float)lry - (float)uly + 1.0f) / 16.0f);
            need = (unsigned)(nx * ny);
        }
        // fragment slots (createRenderingBlockList, RenderScene.cu:106-154): one atomic per wave
        unsigned total;
        const unsigned pre = wave_exclusive_scan(need, total);
        unsigned base = 0;
        if (lane_id() == 0 && total) base = atomicAdd(&m.ctr[C_RENDER], total);
        base = __shfl(base, 0);
        if (!live) continue;
        if (base + pre + need >= a.max_fragments) {
            atomicOr(&m.ctr[C_FLAGS], (unsigned)FLAG_RENDER_FULL);
            continue;
        }
        const int e = ny >= 2 ? 16 : 15;
        const int x1 = min(lrx, ulx + e), y1 = min(lry, uly + e);
        const unsigned zl = __float_as_uint(zlo), zh = __float_as_uint(zhi);
        for (int ty = uly; ty <= y1; ++ty)
            for (int tx = ulx; tx <= x1; ++tx) {
                if constexpr (kLds) {
                    atomicMin(&lmin[ty * zc + tx], zl);
                    atomicMax(&lmax[ty * zc + tx], zh);
                } else {
                    __hip_atomic_fetch_min(FU_G(unsigned, a.zmin) + ty * zc + tx, zl, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_max(FU_G(unsigned, a.zmax) + ty * zc + tx, zh, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
    }
    if constexpr (kLds) {
        __syncthreads();
        for (int t = threadIdx.x; t < nt; t += blockDim.x) {
            const unsigned lo = lmin[t];
            if (lo != 0xFFFFFFFFu) {
                __hip_atomic_fetch_min(FU_G(unsigned, a.zmin) + t, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_max(FU_G(unsigned, a.zmax) + t, lmax[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}
template <bool kLds>
__global__ __launch_bounds__(256) void k_project_blocks(MapView m, ZRangeArgs a) {
    project_body<kLds>(hw_wg(), m, a);
}

void launch_zrange(const MapView& m, const ZRangeArgs& a, hipStream_t s) {
    const int n = a.zcols * a.zrows;
    hipLaunchKernelGGL(k_zrange_init, dim3((n + 255) / 256), dim3(256), 0, s, a.st, a.zmin, a.zmax, n);
    if (n <= kZTilesLds)
        hipLaunchKernelGGL(k_project_blocks<true>, dim3(kProjectGroups), dim3(256), 0, s, m, a);
    else
        hipLaunchKernelGGL(k_project_blocks<false>, dim3(512), dim3(256), 0, s, m, a);
}

// ------------------------------------------------------------------------------ ray march
// DeviceMap::FindVoxel(float3, cache, valid) (DeviceMap.cu:97-114) with the one-block cache
// kept in registers.  The reference's cache starts as a default HashEntry at block (0,0,0)
// with ptr -1 and would read voxel -1+idx on a "hit"; here an empty cache never hits.
template <int F>
struct Sampler {
    using IO = VoxelIO<F>;
    const MapView& m;
    i3 cbp;
    int cheap;
    __device__ Sampler(const MapView& mm) : m(mm), cbp(mki3(0, 0, 0)), cheap(-1) {}
    // returns sdf (NaN if the block is absent); `valid` as FindVoxel sets it, `w` the weight
    __device__ __forceinline__ float find(f3 pt, bool& valid, unsigned& w) {
        const i3 p = f2i3(pt);
        const i3 bp = voxel_to_block(p);
        int heap;
        if (cheap >= 0 && eqi(bp, cbp)) {
            heap = cheap;
        } else {
            heap = find_block(m, bp);
            if (heap < 0) {
                valid = false;
                w = 0;
                return fnan();
            }
            cbp = bp;
            cheap = heap;
        }
        valid = true;
        const unsigned char* blk = IO::block(m, heap);
        const int li = voxel_local_idx(p);
        w = IO::weight(blk, li);
        return IO::sdf(blk, li);
    }
    // Rendering::readSdf (RenderScene.cu:290-295)
    __device__ __forceinline__ float read_sdf(f3 pt, bool& valid) {
        unsigned w;
        const float s = find(pt, valid, w);
        if (w == 0) valid = false;
        return s;
    }
    __device__ __forceinline__ float sdf_at(f3 pt, bool& valid) {
        unsigned w;
        return find(pt, valid, w);
    }
    // Rendering::readSdfInterped (RenderScene.cu:297-319): fractions from floor, lookups by
    // truncation; `valid` ends as the last lookup left it.  The eight corner voxels are
    // computed exactly as the reference computes them (make_int3(pt + offset)); their blocks
    // are resolved first (through the cache, usually one probe) and the eight sdf loads are
    // then issued together.
    __device__ __forceinline__ float read_interp(f3 pt, bool& valid) {
        const f3 f = sub(pt, floor3(pt));
        // corner k = (x, y, z) bits sits at make_int3(pt + (kx, ky, kz)): two coordinates per axis
        const int x0 = f2i(pt.x), x1 = f2i(pt.x + 1.0f);
        const int y0 = f2i(pt.y), y1 = f2i(pt.y + 1.0f);
        const int z0 = f2i(pt.z), z1 = f2i(pt.z + 1.0f);
        const i3 b0 = mki3(x0 >> 3, y0 >> 3, z0 >> 3);
        const i3 b1 = mki3(x1 >> 3, y1 >> 3, z1 >> 3);
        int hp[8];
        if (eqi(b0, b1)) {
            // every corner in one block (about two samples in three): one resolution
            int h;
            if (cheap >= 0 && eqi(b0, cbp)) {
                h = cheap;
            } else {
                h = find_block(m, b0);
                if (h >= 0) {
                    cbp = b0;
                    cheap = h;
                }
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) hp[k] = h;
        } else {
            // corners in two or more blocks: each axis spans b0 or b0 + 1 (mask bit set), so the
            // corners' blocks are b0 + c for the subsets c of the mask.  Each distinct block is
            // resolved once (b0 through the cache) and corner k takes block k & mask, instead
            // of a lookup per corner, which alternates between the blocks and misses the
            // one-block cache every time.
            const int msk = (b1.x - b0.x) | ((b1.y - b0.y) << 1) | ((b1.z - b0.z) << 2);
            int hb[8];
            if (cheap >= 0 && eqi(b0, cbp)) {
                hb[0] = cheap;
            } else {
                hb[0] = find_block(m, b0);
                if (hb[0] >= 0) {
                    cbp = b0;
                    cheap = hb[0];
                }
            }
#pragma unroll
            for (int c = 1; c < 8; ++c) {
                hb[c] = -1;
                if ((c & ~msk) == 0) hb[c] = find_block(m, mki3(b0.x + (c & 1), b0.y + ((c >> 1) & 1), b0.z + (c >> 2)));
            }
            hp[0] = hb[0];
            hp[1] = (msk & 1) ? hb[1] : hb[0];
            hp[2] = (msk & 2) ? hb[2] : hb[0];
            hp[3] = (msk & 2) ? ((msk & 1) ? hb[3] : hb[2]) : hp[1];
            hp[4] = (msk & 4) ? hb[4] : hb[0];
            hp[5] = (msk & 4) ? ((msk & 1) ? hb[5] : hb[4]) : hp[1];
            hp[6] = (msk & 4) ? ((msk & 2) ? hb[6] : hb[4]) : hp[2];
            hp[7] = (msk & 4) ? ((msk & 2) ? ((msk & 1) ? hb[7] : hb[6]) : hp[5]) : hp[3];
        }
        // unconditional loads (an absent block reads pool block 0 and is masked to NaN), so
        // the eight loads issue back to back
        // corner k's index within its block: the index of corner 0 plus per-axis steps (the
        // local coordinate differences, negative where an axis crosses into the next block)
        const int l0 = ((z0 & 7) << 6) | ((y0 & 7) << 3) | (x0 & 7);
        const int ox = (x1 & 7) - (x0 & 7), oy = ((y1 & 7) - (y0 & 7)) << 3, oz = ((z1 & 7) - (z0 & 7)) << 6;
        const int li[8] = {l0, l0 + ox, l0 + oy, l0 + oy + ox, l0 + oz, l0 + oz + ox, l0 + oz + oy, l0 + oz + oy + ox};
        float sv[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float raw = IO::sdf(IO::block(m, hp[k] >= 0 ? hp[k] : 0), li[k]);
            sv[k] = hp[k] >= 0 ? raw : fnan();
        }
        valid = hp[7] >= 0;
        // corner k = (x, y, z) bits; same blend order as the reference
        const float r0 = (1.0f - f.x) * sv[0] + f.x * sv[1];
        const float r1 = (1.0f - f.x) * sv[2] + f.x * sv[3];
        const float r2 = (1.0f - f.y) * r0 + f.y * r1;
        const float r0b = (1.0f - f.x) * sv[4] + f.x * sv[5];
        const float r1b = (1.0f - f.x) * sv[6] + f.x * sv[7];
        const float r3 = (1.0f - f.y) * r0b + f.y * r1b;
        return (1.0f - f.z) * r2 + f.z * r3;
    }

    // Heap indices of the blocks lo + c, c = (dx, dy, dz) in {0, 1}^3 a subset of msk (bit 0:
    // x, 1: y, 2: z; the other entries stay -1 and are not read), probed together: the
    // occupancy words, then the bucket heads, as independent loads; only excess chains are
    // walked one step at a time.
    __device__ __forceinline__ void resolve8(i3 lo, int msk, int (&h)[8]) {
        uint32_t raw[8];
        bool maybe[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            maybe[k] = false;
            raw[k] = 0;
            if ((k & ~msk) == 0) {
                const i3 bp = mki3(lo.x + (k & 1), lo.y + ((k >> 1) & 1), lo.z + ((k >> 2) & 1));
                raw[k] = block_hash_raw(bp);
                maybe[k] = key_in_range(bp) && maybe_allocated(m, raw[k]);
            }
        }
        HEntry e[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (maybe[k]) e[k] = load_entry(&m.entries[bucket_of(raw[k], m.num_buckets, m.bucket_mask)]);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            h[k] = -1;
            if (!maybe[k]) continue;
            const i3 bp = mki3(lo.x + (k & 1), lo.y + ((k >> 1) & 1), lo.z + ((k >> 2) & 1));
            const unsigned long long key = pack_key(bp);
            HEntry c = e[k];
            while (true) {
                if (c.key == key && c.ptr >= 0) {
                    h[k] = c.ptr >> 9;
                    break;
                }
                if (c.next <= 0) break;
                c = load_entry(&m.entries[m.num_buckets + c.next - 1]);
            }
        }
    }

    // read_interp with every corner's block taken from a resolved neighbourhood (the corners
    // must lie in lo .. lo + 1 per axis)
    __device__ __forceinline__ float interp_nb(f3 pt, i3 lo, const int (&h)[8], bool& valid) const {
        const f3 f = sub(pt, floor3(pt));
        const int x0 = f2i(pt.x), x1 = f2i(pt.x + 1.0f);
        const int y0 = f2i(pt.y), y1 = f2i(pt.y + 1.0f);
        const int z0 = f2i(pt.z), z1 = f2i(pt.z + 1.0f);
        // corner k's block is h[x | y << 1 | z << 2], each bit saying whether that axis of the
        // corner lies in lo + 1: chosen by a tree of selects (z, then y, then x: 24 selects)
        // rather than indexing h per corner, which compiles to a compare-and-select chain
        const bool bx0 = (x0 >> 3) != lo.x, bx1 = (x1 >> 3) != lo.x;
        const bool by0 = (y0 >> 3) != lo.y, by1 = (y1 >> 3) != lo.y;
        const bool bz0 = (z0 >> 3) != lo.z, bz1 = (z1 >> 3) != lo.z;
        int pz[2][4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            pz[0][e] = bz0 ? h[4 + e] : h[e];
            pz[1][e] = bz1 ? h[4 + e] : h[e];
        }
        int row[2][2][2];  // [z][y][x]
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                row[c][0][x] = by0 ? pz[c][2 + x] : pz[c][x];
                row[c][1][x] = by1 ? pz[c][2 + x] : pz[c][x];
            }
        float sv[8];
        int hl = -1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int* r = row[k >> 2][(k >> 1) & 1];
            const int hk = ((k & 1) ? bx1 : bx0) ? r[1] : r[0];
            const int li = (((k & 4) ? z1 : z0) & 7) << 6 | (((k & 2) ? y1 : y0) & 7) << 3 | (((k & 1) ? x1 : x0) & 7);
            const float raw = IO::sdf(IO::block(m, hk >= 0 ? hk : 0), li);
            sv[k] = hk >= 0 ? raw : fnan();
            hl = hk;
        }
        valid = hl >= 0;
        const float r0 = (1.0f - f.x) * sv[0] + f.x * sv[1];
        const float r1 = (1.0f - f.x) * sv[2] + f.x * sv[3];
        const float r2 = (1.0f - f.y) * r0 + f.y * r1;
        const float r0b = (1.0f - f.x) * sv[4] + f.x * sv[5];
        const float r1b = (1.0f - f.x) * sv[6] + f.x * sv[7];
        const float r3 = (1.0f - f.y) * r0b + f.y * r1b;
        return (1.0f - f.z) * r2 + f.z * r3;
    }
};

// The per-pixel start of Rendering::operator() (RenderScene.cu:354-385): false when the
// pixel's outputs are already written (nothing rendered: the frame's own maps are kept, or an
// empty z-range tile: NaN).
__device__ __forceinline__ bool ray_setup(const MapView& m, const RaycastArgs& a, int x, int y, float& zr0, float& zr1) {
    const f4 nan4 = mk4(fnan(), fnan(), fnan(), fnan());
    const auto vout = FU_G(f4, a.vmap) + y * a.cols + x;
    const auto nout = FU_G(f4, a.nmap) + y * a.cols + x;
    if (a.keep_v && FU_G(const unsigned, m.ctr)[C_RENDER] == 0u) {
        // Mapping::RayTrace (Mapping.cc:102-108) does not raycast when there is nothing to
        // render; the frame keeps its own level-0 maps
        *vout = FU_G(const f4, a.keep_v)[y * a.cols + x];
        *nout = FU_G(const f4, a.keep_n)[y * a.cols + x];
        return false;
    }
    const int tile = (y / 8) * a.zcols + (x / 8);
    zr0 = __uint_as_float(FU_G(const unsigned, a.zmin)[tile]);
    zr1 = __uint_as_float(FU_G(const unsigned, a.zmax)[tile]);
    if ((double)zr1 < 1e-3 || (double)zr0 < 1e-3 || isnanf_(zr0) || isnanf_(zr1)) {
        *vout = nan4;
        *nout = nan4;
        return false;
    }
    return true;
}

// The march (RenderScene.cu:386-418): true on a hit, with the position in voxel units, the
// sdf there (<= 0) and the ray direction.
template <int F>
__device__ __forceinline__ bool ray_march(const MapView& m, const RaycastArgs& a, const PoseF& pose, int x, int y,
                                          float zr0, float zr1, Sampler<F>& smp, f3& result, float& sdf, f3& dir) {
    const float vinv = m.voxel_inv;
    const float step_scale = a.step_scale;
    sdf = 1.0f;
    f3 p3 = mk3(zr0 * ((float)x - a.cx) * a.invfx, zr0 * ((float)y - a.cy) * a.invfy, zr0);
    float dist_s = norm(p3) * vinv;
    const f3 block_s = mul(add(mv(pose.R, p3), pose.t), vinv);
    p3 = mk3(zr1 * ((float)x - a.cx) * a.invfx, zr1 * ((float)y - a.cy) * a.invfy, zr1);
    const float dist_e = norm(p3) * vinv;
    const f3 block_e = mul(add(mv(pose.R, p3), pose.t), vinv);
    dir = normalised(sub(block_e, block_s));
    result = block_s;
    bool valid = false;
    float step;
    while (dist_s < dist_e) {
        sdf = smp.read_sdf(result, valid);
        if (!valid) {
            step = 8.0f;
        } else {
            if (sdf <= 0.1f && sdf >= -0.5f) sdf = smp.read_interp(result, valid);
            if (sdf <= 0.0f) break;
            if (!isnanf_(sdf))
                step = fmaxf(sdf * step_scale, 1.0f);
            else
                step = 8.0f;
        }
        result = add(result, mul(step, dir));
        dist_s += step;
    }
    return sdf <= 0.0f;
}

// Two half-steps of refinement (RenderScene.cu:419-428) and Rendering::readNormal
// (RenderScene.cu:321-352): true with the camera-frame vertex and normal, false when a
// normal sample is NaN, 1.0 or invalid.
template <int F>
__device__ __forceinline__ bool ray_shade(const MapView& m, const RaycastArgs& a, const PoseF& pose, Sampler<F>& smp,
                                          f3 result, float sdf, f3 dir, f4& vo, f4& no) {
    const float step_scale = a.step_scale;
    bool valid;
    float step = sdf * step_scale;
    result = add(result, mul(step, dir));
    sdf = smp.read_interp(result, valid);
    step = sdf * step_scale;
    result = add(result, mul(step, dir));
    // the six samples' corners lie between make_int3(result - 1) and make_int3((result + 1) + 1)
    // per axis (truncation is monotonic), at most four voxels: lo or lo + 1 per axis, and only
    // the blocks of axes that do cross are probed
    const i3 lo = mki3(f2i(result.x - 1.0f) >> 3, f2i(result.y - 1.0f) >> 3, f2i(result.z - 1.0f) >> 3);
    const i3 hi = mki3(f2i((result.x + 1.0f) + 1.0f) >> 3, f2i((result.y + 1.0f) + 1.0f) >> 3,
                       f2i((result.z + 1.0f) + 1.0f) >> 3);
    int hb[8];
    smp.resolve8(lo, (hi.x - lo.x) | ((hi.y - lo.y) << 1) | ((hi.z - lo.z) << 2), hb);
    float s[6];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        // result + (+-1 on axis k / 2, 0 elsewhere): the other two coordinates as they are, which
        // is what adding +0.0 gives them for the corners and fractions read here (only a -0
        // would turn +0, and truncation, floor and x - floor(x) treat the two alike), so the
        // samples sharing a coordinate share its corners
        const float d = (k & 1) ? -1.0f : 1.0f;
        const f3 pt = mk3(k / 2 == 0 ? result.x + d : result.x, k / 2 == 1 ? result.y + d : result.y,
                          k / 2 == 2 ? result.z + d : result.z);
        bool v2;
        s[k] = smp.interp_nb(pt, lo, hb, v2);
        ok = ok && !(isnanf_(s[k]) || s[k] == 1.0f || !v2);
    }
    if (!ok) return false;
    const f3 n = normalised(mv(pose.Rinv, mk3(s[0] - s[1], s[2] - s[3], s[4] - s[5])));
    const f3 vtx = mv(pose.Rinv, sub(mul(result, m.voxel_size), pose.t));
    vo = mk4(vtx.x, vtx.y, vtx.z, 1.0f);
    no = mk4(n.x, n.y, n.z, 1.0f);
    return true;
}

// one lane per pixel of a 16x16 tile: set-up, march, refinement and normal.  `wave` picks the
// pixels of each 64-lane wave (FU_WAVE_*): 16x4 rows, 8x8 quadrants or 4x16 columns.
template <int F>
__device__ __forceinline__ void raycast_body(WG g, const MapView& m, const RaycastArgs& a, int wave) {
    PoseF pose = a.pose;
    if (a.st && !pose_from_state(a.st, pose)) return;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    int px, py;
    if (wave == FU_WAVE_8X8) {
        px = ((w & 1) << 3) | (l & 7);
        py = ((w >> 1) << 3) | (l >> 3);
    } else if (wave == FU_WAVE_4X16) {
        px = (w << 2) | (l & 3);
        py = l >> 2;
    } else {
        px = t & 15;
        py = t >> 4;
    }
    const int x = g.x * 16 + px;
    const int y = g.y * 16 + py;
    if (x >= a.cols || y >= a.rows) return;
    float zr0, zr1;
    if (!ray_setup(m, a, x, y, zr0, zr1)) return;
    Sampler<F> smp(m);
    f3 result, dir;
    float sdf;
    f4 vo = mk4(fnan(), fnan(), fnan(), fnan()), no = vo;
    if (ray_march<F>(m, a, pose, x, y, zr0, zr1, smp, result, sdf, dir)) {
        if (!ray_shade<F>(m, a, pose, smp, result, sdf, dir, vo, no)) vo = no = mk4(fnan(), fnan(), fnan(), fnan());
    }
    FU_G(f4, a.vmap)[y * a.cols + x] = vo;
    FU_G(f4, a.nmap)[y * a.cols + x] = no;
}
template <int F>
// eight waves per SIMD: the march is latency-bound, and the select trees of the normal samples
// would otherwise take the kernel past 64 registers per lane
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_raycast(MapView m, RaycastArgs a,
                                                                                            int wave) {
    raycast_body<F>(hw_wg(), m, a, wave);
}

template <int F>
static void raycast_f(const MapView& m, const RaycastArgs& a, hipStream_t s) {
    const dim3 g((a.cols + 15) / 16, (a.rows + 15) / 16);
    hipLaunchKernelGGL(k_raycast<F>, g, dim3(256), 0, s, m, a, (int)tuning().raycast_wave);
}
void launch_raycast(const MapView& m, const RaycastArgs& a, hipStream_t s) {
    switch (m.fmt) {
        case VF_F32_W8: raycast_f<VF_F32_W8>(m, a, s); break;
        case VF_F32_W16: raycast_f<VF_F32_W16>(m, a, s); break;
        default: raycast_f<VF_F16_W8>(m, a, s); break;
    }
}

// ------------------------------------------------------------------ batched launches
// Mapping::RayTrace for every session of a step: z-range tiles, block projection, ray march.
__global__ void kb_zrange_init(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    zrange_init_body(g, d.st, d.zr.zmin, d.zr.zmax, d.zr.zcols * d.zr.zrows);
}
template <bool kLds>
__global__ __launch_bounds__(256) void kb_project(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    project_body<kLds>(g, cload(d.m), cload(d.zr));
}
template <int F>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void kb_raycast(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, batch_wg(b), g);
    const SessDesc& d = b.desc[b.sess[s]];
    raycast_body<F>(g, cload(d.m), cload(d.rc), b.wave);
}

void launch_b_zrange(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    int grid = batch_layout(b, h, BT_ZTILES, 0, 0);
    hipLaunchKernelGGL(kb_zrange_init, dim3(grid), dim3(256), 0, s, b);
    bool lds = true;
    for (int i = 0; i < b.n; ++i) lds = lds && h[b.sess[i]].zr.zcols * h[b.sess[i]].zr.zrows <= kZTilesLds;
    if (lds) {
        grid = batch_layout(b, h, BT_FIXED, 0, kProjectGroups);
        hipLaunchKernelGGL(kb_project<true>, dim3(grid), dim3(256), 0, s, b);
    } else {
        const int per = b.n > 0 ? 4096 / b.n : 64;
        grid = batch_layout(b, h, BT_FIXED, 0, per < 32 ? 32 : (per > 512 ? 512 : per));
        hipLaunchKernelGGL(kb_project<false>, dim3(grid), dim3(256), 0, s, b);
    }
}
template <int F>
static void b_raycast_f(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    const int grid = batch_layout(b, h, BT_TILES, 0, 0);
    b.xcd = xcd_order(FU_STAGE_RAYCAST);
    b.wave = tuning().raycast_wave;
    hipLaunchKernelGGL(kb_raycast<F>, dim3(grid), dim3(256), 0, s, b);
}
void launch_b_raycast(BatchLaunch& b, const SessDesc* h, int fmt, hipStream_t s) {
    switch (fmt) {
        case VF_F32_W8: b_raycast_f<VF_F32_W8>(b, h, s); break;
        case VF_F32_W16: b_raycast_f<VF_F32_W16>(b, h, s); break;
        default: b_raycast_f<VF_F16_W8>(b, h, s); break;
    }
}

}  // namespace fu
