// Voxel-block hash table: reset, allocation, visible-block list and TSDF integration.
//
// Reference: Mapping/DeviceMap.cu:3-62 (Hash, CreateEntry, CreateBlock),
// Mapping/FuseMap.cu:36-138 (visibility, CreateBlocks, CheckFullVisibility),
// FuseMap.cu:140-191 (integrateColor), FuseMap.cu:321-364 (ResetMap).
//
// Allocation runs in two passes instead of the reference's per-bucket try-lock, which
// silently drops inserts under contention:
//   A  one thread per pixel marches the truncation band exactly as CreateBlocks does and
//      emits block keys; duplicates are removed first along the ray, then across the wave
//      with ballots, the survivors probe the table read-only, and keys not present are
//      claimed in a per-pass open-addressing set with a 64-bit CAS (one winner per key)
//      and appended to a compact list with one wave-aggregated atomic.
//   B  one thread per listed key takes a block from the pool and links the key into its
//      bucket: a CAS on the bucket head, else a CAS on the tail's `next` of the excess
//      chain with an excess slot reserved beforehand.  Keys are unique here, so the
//      lock-free append cannot duplicate, never waits on another lane and never leaks.
// The resulting block set is the sequential insertion set of the reference (within
// capacity); only the order of pool indices differs.
#include "fu_device.h"
#include "fu_kernels.h"
#include "fu_batch.h"

namespace fu {

__constant__ U8Div255 c_u8div255 = U8Div255();
__device__ __forceinline__ float u8div255(unsigned c) { return FU_G(const float, c_u8div255.v)[c]; }

// ---------------------------------------------------------------- reset (FuseMap.cu:321-364)
__global__ void k_reset_entries(HEntry* e, unsigned n) {
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        HEntry h;
        h.key = kEmptyKey;
        h.ptr = -1;
        h.next = 0;
        e[i] = h;
    }
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, unsigned n) {
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_zero_u32(unsigned* p, unsigned n) {
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

void launch_reset_map(const MapView& m, hipStream_t s) {
    const unsigned ne = m.num_buckets + m.num_excess;
    hipLaunchKernelGGL(k_reset_entries, dim3(2048), dim3(256), 0, s, m.entries, ne);
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, s, m.block_key, kEmptyKey, m.num_blocks);
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, s, m.newset, kEmptyKey, m.newset_mask + 1);
    hipLaunchKernelGGL(k_zero_u32, dim3(1024), dim3(256), 0, s, m.occ, (m.occ_mask + 1) / 32);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, m.ctr, (unsigned)C_COUNT);
}

// per-frame counters: the new-key count and the visible count start at zero
__device__ __forceinline__ void begin_pass_body(unsigned* ctr, int which) {
    if (threadIdx.x == 0) {
        if (which & 1) ctr[C_NEW] = 0;
        if (which & 2) ctr[C_VISIBLE] = 0;
        if (which & 4) ctr[C_RENDER] = 0;
    }
}
__global__ void k_begin_pass(unsigned* ctr, int which) { begin_pass_body(ctr, which); }
void launch_begin_pass(const MapView& m, int which, hipStream_t s) {
    hipLaunchKernelGGL(k_begin_pass, dim3(1), dim3(64), 0, s, m.ctr, which);
}

// ------------------------------------------------------------ allocation pass A (per pixel)
// claim `key` in the per-pass set; true if this lane is the one that inserted it
__device__ __forceinline__ bool newset_claim(const MapView& m, unsigned long long key, unsigned& slot_out) {
    unsigned slot = (unsigned)mix64(key) & m.newset_mask;
    for (unsigned probe = 0; probe <= m.newset_mask; ++probe) {
        unsigned long long old = atomicCAS(&m.newset[slot], kEmptyKey, key);
        if (old == kEmptyKey) {
            slot_out = slot;
            return true;
        }
        if (old == key) return false;
        slot = (slot + 1) & m.newset_mask;
    }
    atomicOr(&m.ctr[C_FLAGS], (unsigned)FLAG_NEWSET_FULL);
    return false;
}


// claim `key` (block bp) for allocation if it is not in the table (want: this lane has a
// block; every lane of the wave calls, for the wave-wide append)
__device__ __forceinline__ void claim_block(const MapView& m, bool want, unsigned long long key, i3 bp) {
    bool won = false;
    unsigned slot = 0;
    if (want && find_block(m, bp) < 0) won = newset_claim(m, key, slot);
    const unsigned idx = wave_append(&m.ctr[C_NEW], won);
    if (won) {
        if (idx < m.max_new) {
            m.newkeys[idx] = key;
            m.newslot[idx] = slot;
        } else {
            // list full: release the claim so a later pass can allocate this key
            atomicExch(&m.newset[slot], kEmptyKey);
            atomicOr(&m.ctr[C_FLAGS], (unsigned)FLAG_NEWSET_FULL);
        }
    }
}

// Fusion::CreateBlocks (FuseMap.cu:80-111), one lane per pixel, 16x16 pixels per group.  Each
// lane marches its ray's truncation band; the blocks its steps enter (a step in the block of
// the previous step is a no-op for CreateBlock) go into a work-group set in LDS, so a block
// the group's rays share is probed and claimed once, by the lanes that scan the set
// afterwards.  A block that finds no free LDS slot within kAllocProbes is probed and claimed
// at once by its lane; the global per-pass set de-duplicates across lanes and groups.
constexpr int kAllocSlots = 512;  // a 16x16 tile holds a few dozen blocks
constexpr int kAllocProbes = 8;
constexpr int kAllocThreads = 256;  // 16x16 pixels per group
static_assert(kAllocSlots % kAllocThreads == 0, "whole slot rows per lane");
__device__ __forceinline__ void alloc_a_body(WG g, const MapView& m, const FuseArgs& p) {
    __shared__ unsigned long long set[kAllocSlots];
    PoseF pose = p.pose;
    if (p.st && !(frame_fuses(p.st) && pose_from_state(p.st, pose))) return;  // uniform: the frame does not fuse
    for (int i = threadIdx.x; i < kAllocSlots; i += blockDim.x) set[i] = kEmptyKey;
    __syncthreads();
    const auto depth = FU_G(const float, p.depth);
    const int x = g.x * 16 + (threadIdx.x & 15);
    const int y = g.y * 16 + (threadIdx.x >> 4);
    bool live = x < p.cols && y < p.rows;
    f3 pt = mk3(0.f, 0.f, 0.f), dir = mk3(0.f, 0.f, 0.f);
    int nsteps = 0;
    if (live) {
        const float z = depth[y * p.cols + x];
        if (!(isnanf_(z) || z < m.depth_min || z > m.depth_max)) {
            const float thresh = m.trunc / 2.0f;
            const float z_near = fminf(m.depth_max, z - thresh);
            const float z_far = fminf(m.depth_max, z + thresh);
            if (z_near < z_far) {
                // unproject: Rview * (z(x-cx)/fx, z(y-cy)/fy, z) + tview, then to voxel units
                f3 a = mk3(z_near * ((float)x - p.cx) * p.invfx, z_near * ((float)y - p.cy) * p.invfy, z_near);
                f3 b = mk3(z_far * ((float)x - p.cx) * p.invfx, z_far * ((float)y - p.cy) * p.invfy, z_far);
                a = mul(add(mv(pose.R, a), pose.t), m.voxel_inv);
                b = mul(add(mv(pose.R, b), pose.t), m.voxel_inv);
                dir = sub(b, a);
                const float len = norm(dir);
                nsteps = d2i(ceil(2.0 * (double)len));
                dir = divs(dir, (float)(nsteps - 1));
                pt = a;
            }
        }
    }
    i3 prev = mki3(0x7fffffff, 0, 0);
    for (int i = 0; i < nsteps; ++i) {
        const i3 b = voxel_to_block(f2i3(pt));
        pt = add(pt, dir);
        if (eqi(b, prev)) continue;
        prev = b;
        if (!key_in_range(b)) {
            atomicOr(&m.ctr[C_FLAGS], (unsigned)FLAG_COORD_RANGE);
            continue;
        }
        const unsigned long long key = pack_key(b);
        unsigned h = (unsigned)mix64(key) & (kAllocSlots - 1);
        bool done = false;
        for (int probe = 0; probe < kAllocProbes && !done; ++probe) {
            const unsigned long long old = atomicCAS(&set[h], kEmptyKey, key);
            done = old == kEmptyKey || old == key;
            h = (h + 1) & (kAllocSlots - 1);
        }
        if (!done) claim_block(m, true, key, b);
    }
    __syncthreads();
    // the group's keys (a few dozen in a 16x16 tile) move to the front of the set, so the
    // table probes and claims run in as few wave passes as there are keys, not one per
    // occupied 64-slot row; the order of the claims does not matter
    __shared__ unsigned nkeys;
    unsigned long long held[kAllocSlots / kAllocThreads];
#pragma unroll
    for (int k = 0; k < kAllocSlots / kAllocThreads; ++k) held[k] = set[threadIdx.x + k * kAllocThreads];
    if (threadIdx.x == 0) nkeys = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kAllocSlots / kAllocThreads; ++k)
        if (held[k] != kEmptyKey) set[atomicAdd(&nkeys, 1u)] = held[k];
    __syncthreads();
    const int n = (int)nkeys;  // uniform: every lane runs the same passes
    for (int base = 0; base < n; base += kAllocThreads) {
        const int i = base + (int)threadIdx.x;
        const unsigned long long key = i < n ? set[i] : kEmptyKey;
        claim_block(m, key != kEmptyKey, key, unpack_key(key));
    }
}
__global__ __launch_bounds__(kAllocThreads) void k_alloc_pass_a(MapView m, FuseArgs p) { alloc_a_body(hw_wg(), m, p); }

// ------------------------------------------------------------ allocation pass B (per new key)
// one wave per new block: lane 0 links the key, all lanes initialise the block's voxels
// (sdf NaN, weight 0, colour 0: Voxel::release, DeviceMap.h:75-79)
__device__ __forceinline__ void alloc_b_body(WG g, const MapView& m, const TrackState* st) {
    if (st && !frame_fuses(st)) return;
    const unsigned nnew = min(m.ctr[C_NEW], m.max_new);
    const int wave = (g.x * blockDim.x + threadIdx.x) >> 6;
    const int nwaves = (g.nx * blockDim.x) >> 6;
    const int lane = lane_id();
    const int bbytes = vf_block_bytes(m.fmt);
    const int sb = vf_sdf_bytes(m.fmt);
    for (unsigned i = wave; i < nnew; i += nwaves) {
        const unsigned long long key = m.newkeys[i];
        int heap = -1;
        if (lane == 0) {
            m.newset[m.newslot[i]] = kEmptyKey;  // clear for the next pass
            const unsigned h = atomicAdd(&m.ctr[C_ALLOC], 1u);
            if (h >= m.num_blocks) {
                atomicOr(&m.ctr[C_FLAGS], (unsigned)FLAG_HEAP_FULL);
            } else {
                heap = (int)h;
                const i3 bp = unpack_key(key);
                HEntry* e = &m.entries[bucket_of(block_hash_raw(bp), m.num_buckets, m.bucket_mask)];
                const unsigned long long old = atomicCAS(&e->key, kEmptyKey, key);
                bool ok = true;
                if (old == kEmptyKey) {
                    e->ptr = heap * 512;
                } else {
                    // append to the excess chain (DeviceMap.cu:50-60 without the lock)
                    int slot = -1;
                    HEntry* tail = e;
                    while (true) {
                        int nx = __hip_atomic_load(&tail->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        while (nx != 0) {
                            tail = &m.entries[m.num_buckets + nx - 1];
                            nx = __hip_atomic_load(&tail->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        if (slot < 0) {
                            const unsigned sx = atomicAdd(&m.ctr[C_EXCESS], 1u);
                            if (sx >= m.num_excess) {
                                atomicOr(&m.ctr[C_FLAGS], (unsigned)FLAG_EXCESS_FULL);
                                ok = false;
                                break;
                            }
                            slot = (int)sx;
                            HEntry ne;
                            ne.key = key;
                            ne.ptr = heap * 512;
                            ne.next = 0;
                            m.entries[m.num_buckets + slot] = ne;
                            __threadfence();
                        }
                        const int prev = atomicCAS(&tail->next, 0, slot + 1);
                        if (prev == 0) break;
                        tail = &m.entries[m.num_buckets + prev - 1];
                    }
                }
                m.block_key[heap] = ok ? key : kEmptyKey;
                if (ok) {
                    const unsigned hb = occ_bit(m, block_hash_raw(bp));
                    atomicOr(&m.occ[hb >> 5], 1u << (hb & 31));
                }
            }
        }
        heap = __shfl(heap, 0);
        if (heap >= 0) {
            unsigned char* blk = m.pool + (size_t)heap * bbytes;
            // sdf plane
            if (sb == 4) {
                float* sp = reinterpret_cast<float*>(blk);
                for (int v = lane; v < 512; v += 64) sp[v] = fnan();
            } else {
                __half* sp = reinterpret_cast<__half*>(blk);
                for (int v = lane; v < 512; v += 64) sp[v] = __float2half_rn(fnan());
            }
            // weight and colour planes are zero
            unsigned* zp = reinterpret_cast<unsigned*>(blk + 512 * sb);
            const int nz = (bbytes - 512 * sb) / 4;
            for (int v = lane; v < nz; v += 64) zp[v] = 0u;
        }
    }
}
__global__ __launch_bounds__(256) void k_alloc_pass_b(MapView m, const TrackState* st) { alloc_b_body(hw_wg(), m, st); }

void launch_alloc(const MapView& m, const FuseArgs& p, hipStream_t s) {
    hipLaunchKernelGGL(k_alloc_pass_a, dim3((p.cols + 15) / 16, (p.rows + 15) / 16), dim3(kAllocThreads), 0, s, m, p);
    hipLaunchKernelGGL(k_alloc_pass_b, dim3(512), dim3(256), 0, s, m, p.st);
}

// ------------------------------------------------------------ visible list (FuseMap.cu:36-78,113-138)

__device__ __forceinline__ bool vertex_visible(const VisArgs& v, const PoseF& pose, f3 p) {
    p = mv(pose.Rinv, sub(p, pose.t));
    if (p.z < 1e-3f) return false;
    const float px = v.fx * p.x / p.z + v.cx;
    const float py = v.fy * p.y / p.z + v.cy;
    return px >= 0 && py >= 0 && px < (float)v.cols && py < (float)v.rows && p.z >= v.depth_min &&
           p.z <= v.depth_max;
}

// the eight corners are visited by repeated float adds/subtracts exactly as in
// Fusion::CheckBlockVisibility (FuseMap.cu:47-78)
__device__ __forceinline__ bool block_visible(const VisArgs& v, const PoseF& pose, i3 pos, float scale) {
    f3 c = i3f(pos, scale);
    if (vertex_visible(v, pose, c)) return true;
    c.z += scale;
    if (vertex_visible(v, pose, c)) return true;
    c.y += scale;
    if (vertex_visible(v, pose, c)) return true;
    c.x += scale;
    if (vertex_visible(v, pose, c)) return true;
    c.z -= scale;
    if (vertex_visible(v, pose, c)) return true;
    c.y -= scale;
    if (vertex_visible(v, pose, c)) return true;
    c.x -= scale;
    c.y += scale;
    if (vertex_visible(v, pose, c)) return true;
    c.x += scale;
    c.y -= scale;
    c.z += scale;
    if (vertex_visible(v, pose, c)) return true;
    return false;
}

// Scans the allocated-block array (N_alloc keys) rather than all N_E hash entries; the
// set is the same (every allocated block has exactly one live entry).  Compaction is a
// wave ballot plus one atomic per wave.
__device__ __forceinline__ void visible_body(WG g, const MapView& m, const VisArgs& v) {
    PoseF pose = v.pose;
    if (v.st && !pose_from_state(v.st, pose)) return;
    const unsigned nalloc = min(FU_G(const unsigned, m.ctr)[C_ALLOC], m.num_blocks);
    const float scale = m.voxel_size * 8.0f;  // DeviceMap::blockWidth
    const unsigned stride = g.nx * blockDim.x;
    // uniform trip count per wave so the ballot below sees every lane
    const unsigned base0 = g.x * blockDim.x + (threadIdx.x & ~63u);
    for (unsigned b0 = base0; b0 < nalloc; b0 += stride) {
        const unsigned i = b0 + lane_id();
        bool vis = false;
        unsigned long long key = kEmptyKey;
        if (i < nalloc) {
            key = FU_G(const unsigned long long, m.block_key)[i];
            if (key != kEmptyKey) vis = block_visible(v, pose, unpack_key(key), scale);
        }
        const unsigned idx = wave_append(&m.ctr[C_VISIBLE], vis);
        if (vis && idx < m.num_blocks) {
            FU_G(unsigned, m.visible)[idx] = i;
            FU_G(unsigned long long, m.visible_key)[idx] = key;
        }
    }
}
__global__ __launch_bounds__(256) void k_visible(MapView m, VisArgs v) { visible_body(hw_wg(), m, v); }

void launch_visible(const MapView& m, const VisArgs& v, hipStream_t s) {
    hipLaunchKernelGGL(k_visible, dim3(1024), dim3(256), 0, s, m, v);
}

// ------------------------------------------------------------ integration (FuseMap.cu:140-191)

// Fusion::integrateColor (FuseMap.cu:140-191).  A work-group takes kIntBlocks visible blocks per
// iteration, 128 lanes (two waves) per block.  Lane t of a block owns the voxels t + 128k,
// k = 0..3 (index z*64 + y*8 + x): one column (x, y) at z = t/64 + 2k, so the x and y terms of
// the voxel's position (VectorMath.h:364-366, evaluated ((r.x*X + r.y*Y) + r.z*Z) as the
// reference's float3 dot product) are computed once for its four voxels, and each wave's
// loads and stores cover 64 consecutive voxels.  The next iteration's block ids and keys are
// loaded before the current blocks are processed.
constexpr int kIntBlocks = 2;
constexpr int kIntThreads = 128 * kIntBlocks;  // the launch size the lane mapping assumes
template <int F, bool kPacked>
__device__ __forceinline__ void integrate_body(WG g, const MapView& m, const FuseArgs& a) {
    using IO = VoxelIO<F>;
    constexpr int H = 4;  // voxels per lane per iteration
    PoseF pose = a.pose;
    if (a.st && !(frame_fuses(a.st) && pose_from_state(a.st, pose))) return;  // uniform: the frame does not fuse
    // the byte / 255 quotients in LDS: a gather there instead of six global loads per voxel
    __shared__ float div255[256];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) div255[i] = u8div255(i);
    __syncthreads();
    const unsigned nvis = min(FU_G(const unsigned, m.ctr)[C_VISIBLE], m.num_blocks);
    const int cols = a.cols, rows = a.rows;
    const auto depth = FU_G(const float, a.depth);
    const auto nmap = FU_G(const f4, a.nmap);
    const auto rgb = FU_G(const unsigned char, a.rgb);
    const auto px = FU_G(const f4, a.px);
    const auto prgb = FU_G(const unsigned, a.prgb);
    const auto vis = FU_G(const unsigned, m.visible);
    const auto vkey = FU_G(const unsigned long long, m.visible_key);
    const float thresh = m.trunc;
    const float rthresh = fu_rcp_refined(thresh);
    const float vs = m.voxel_size;
    const unsigned stride = (unsigned)g.nx * kIntBlocks;
    // this lane's block of the iteration: uniform per wave, so its id, key and voxel pointer
    // live in scalar registers
    const unsigned mine = __builtin_amdgcn_readfirstlane(threadIdx.x >> 7);
    const int t = threadIdx.x & 127;
    const int lx = t & 7, ly = (t >> 3) & 7, lz = t >> 6;
    unsigned vb = (unsigned)g.x * kIntBlocks + mine;
    unsigned hn = vb < nvis ? vis[vb] : 0u;
    unsigned long long kn = vb < nvis ? vkey[vb] : 0ull;
    for (; vb < nvis; vb += stride) {
        const unsigned heap = hn;
        const i3 bp = unpack_key(kn);
        hn = vb + stride < nvis ? vis[vb + stride] : 0u;
        kn = vb + stride < nvis ? vkey[vb + stride] : 0ull;
        // the voxel column's x and y terms: (float)vp * voxel_size - t, times the first two
        // columns of Rinv, summed (the first addition of each row's dot product)
        const float X = (float)(bp.x * 8 + lx) * vs - pose.t.x;
        const float Y = (float)(bp.y * 8 + ly) * vs - pose.t.y;
        const float px0 = pose.Rinv.r0.x * X + pose.Rinv.r0.y * Y;
        const float py0 = pose.Rinv.r1.x * X + pose.Rinv.r1.y * Y;
        const float pz0 = pose.Rinv.r2.x * X + pose.Rinv.r2.y * Y;
        int pix[H];
        bool upd[H];
        float sdf[H];
        f3 pos[H];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const float Z = (float)(bp.z * 8 + lz + 2 * h) * vs - pose.t.z;
            pos[h] = mk3(px0 + pose.Rinv.r0.z * Z, py0 + pose.Rinv.r1.z * Z, pz0 + pose.Rinv.r2.z * Z);
            // f2i(fx * x / z + cx): both quotients share z's reciprocal in the range where
            // the division core is exact; below |a| = 2^-100, where it may not be, the
            // quotient is under 2^-60 and q + cx truncates to the same integer either way
            const float nu = a.fx * pos[h].x, nw = a.fy * pos[h].y, z = pos[h].z;
            float qu, qv;
            if (fabsf(z) >= 0x1p-60f && fmaxf(fmaxf(fabsf(nu), fabsf(nw)), fabsf(z)) <= 0x1p30f) {
                const float yz = fu_rcp_refined(z);
                qu = fu_div_refined(nu, z, yz);
                qv = fu_div_refined(nw, z, yz);
            } else {
                qu = nu / z;
                qv = nw / z;
            }
            const int u = f2i(qu + a.cx);
            const int v = f2i(qv + a.cy);
            upd[h] = !(u < 0 || v < 0 || u >= cols || v >= rows);
            pix[h] = upd[h] ? v * cols + u : 0;
        }
        float dp[H];
        f4 nl[H];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            if constexpr (kPacked) {
                const f4 q = px[pix[h]];
                dp[h] = q.x;
                nl[h] = mk4(q.y, q.z, q.w, 1.0f);
            } else {
                dp[h] = depth[pix[h]];
            }
        }
        bool any = false;
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const bool dv = upd[h] && !(isnanf_(dp[h]) || dp[h] > m.depth_max || dp[h] < m.depth_min);
            sdf[h] = dp[h] - pos[h].z;
            upd[h] = dv && (sdf[h] >= -thresh);
            if (!upd[h]) pix[h] = 0;
            any = any || upd[h];
        }
        if (!any) continue;
        unsigned char cr[H], cg[H], cb[H];
        float sold[H];
        VoxAux old[H];
        unsigned char* blk = IO::block(m, (int)heap);
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int li = t + 128 * h;
            if constexpr (kPacked) {
                const unsigned c = prgb[pix[h]];
                cr[h] = (unsigned char)c;
                cg[h] = (unsigned char)(c >> 8);
                cb[h] = (unsigned char)(c >> 16);
            } else {
                nl[h] = nmap[pix[h]];
                cr[h] = rgb[pix[h] * 3 + 0];
                cg[h] = rgb[pix[h] * 3 + 1];
                cb[h] = rgb[pix[h] * 3 + 2];
            }
            sold[h] = IO::sdf(blk, li);
            old[h] = IO::aux(blk, li);
        }
#pragma unroll
        for (int h = 0; h < H; ++h) {
            if (!upd[h] || isnanf_(nl[h].x)) continue;
            const int li = t + 128 * h;
            // sdf / thresh: |sdf| is 0 or at least ulp(depth_min), thresh a positive normal
            const float s = fminf(1.0f, fu_div_refined(sdf[h], thresh, rthresh));
            // 4-D dot with w = 1 on both sides (VectorMath.h:85-87,203-205,340-342); the
            // four quotients by the norm share its reciprocal when the components allow
            const f4 p4 = mk4(pos[h].x, pos[h].y, pos[h].z, 1.0f);
            f4 pn;
            if (fu_div_range3(p4.x, p4.y, p4.z)) {
                const float n = fu_sqrt_normal(dot4(p4, p4));  // norm4(p4): the sum is >= 1
                const float y = fu_rcp_refined(n);
                pn = mk4(fu_div_refined(p4.x, n, y), fu_div_refined(p4.y, n, y), fu_div_refined(p4.z, n, y),
                         fu_div_refined(1.0f, n, y));
            } else {
                pn = normalised4(p4);
            }
            const float w = dot4(nl[h], pn);
            VoxAux nv;
            if (old[h].w == 0) {
                IO::set_sdf(blk, li, s);
                nv.w = 1u;
                nv.r = cr[h];
                nv.g = cg[h];
                nv.b = cb[h];
            } else {
                const float aw = w * 0.2f;
                const float bw = 1.0f - w * 0.2f;
                const float fr = (aw * div255[cr[h]] + bw * div255[old[h].r]) * 255.f;
                const float fg = (aw * div255[cg[h]] + bw * div255[old[h].g]) * 255.f;
                const float fb = (aw * div255[cb[h]] + bw * div255[old[h].b]) * 255.f;
                const float fw = (float)old[h].w;
                IO::set_sdf(blk, li, (sold[h] * fw + w * s) / (fw + w));
                nv.w = min(m.max_weight, old[h].w + 1u);
                nv.r = f2u8(fr);
                nv.g = f2u8(fg);
                nv.b = f2u8(fb);
            }
            IO::set_aux(blk, li, nv);
        }
    }
}
template <int F, bool kPacked>
__global__ __launch_bounds__(kIntThreads) void k_integrate(MapView m, FuseArgs a) {
    integrate_body<F, kPacked>(hw_wg(), m, a);
}

// the packed per-pixel inputs of integration (FuseArgs::px, prgb)
__device__ __forceinline__ void fuse_pack_body(WG g, const FuseArgs& a) {
    const int i = (g.y * g.nx + g.x) * 256 + threadIdx.x;
    if (i >= a.cols * a.rows) return;
    const f4 n = FU_G(const f4, a.nmap)[i];
    const auto c = FU_G(const unsigned char, a.rgb) + 3 * i;
    FU_G(f4, a.px)[i] = mk4(FU_G(const float, a.depth)[i], n.x, n.y, n.z);
    FU_G(unsigned, a.prgb)[i] = (unsigned)c[0] | ((unsigned)c[1] << 8) | ((unsigned)c[2] << 16);
}
__global__ __launch_bounds__(256) void k_fuse_pack(FuseArgs a) { fuse_pack_body(hw_wg(), a); }

template <bool kPacked>
static void integrate_f(const MapView& m, const FuseArgs& a, unsigned grid, hipStream_t s) {
    switch (m.fmt) {
        case VF_F32_W8: hipLaunchKernelGGL((k_integrate<VF_F32_W8, kPacked>), dim3(grid), dim3(kIntThreads), 0, s, m, a); break;
        case VF_F32_W16: hipLaunchKernelGGL((k_integrate<VF_F32_W16, kPacked>), dim3(grid), dim3(kIntThreads), 0, s, m, a); break;
        default: hipLaunchKernelGGL((k_integrate<VF_F16_W8, kPacked>), dim3(grid), dim3(kIntThreads), 0, s, m, a); break;
    }
}
void launch_fuse_pack(const FuseArgs& a, hipStream_t s) {
    const int n = a.cols * a.rows;
    hipLaunchKernelGGL(k_fuse_pack, dim3((n + 255) / 256), dim3(256), 0, s, a);
}
void launch_integrate(const MapView& m, const FuseArgs& a, unsigned grid, hipStream_t s) {
    if (a.px)
        integrate_f<true>(m, a, grid, s);
    else
        integrate_f<false>(m, a, grid, s);
}

// ------------------------------------------------------------------ batched launches
// FuseMapColor for every session of a step: counters, allocation (both passes), visible
// list and integration, each stage one launch over all sessions.
__global__ void kb_begin_pass(BatchLaunch b) { begin_pass_body(b.desc[b.sess[blockIdx.x]].m.ctr, 7); }

__device__ __forceinline__ FuseArgs slot_fuse(const BatchLaunch& b, const SessDesc& d, int s) {
    FuseArgs a = d.fuse;
    a.rgb = (const unsigned char*)b.in[s];
    return a;
}

__global__ __launch_bounds__(kAllocThreads) void kb_alloc_a(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const MapView m = d.m;  // private copy: the body keeps the fields in scalar registers
    alloc_a_body(g, m, d.fuse);
}
__global__ __launch_bounds__(256) void kb_alloc_b(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const MapView m = d.m;  // private copy: the body keeps the fields in scalar registers
    alloc_b_body(g, m, d.st);
}
__global__ __launch_bounds__(256) void kb_visible(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, blockIdx.x, g);
    const SessDesc& d = b.desc[b.sess[s]];
    const MapView m = d.m;  // private copy: the body keeps the fields in scalar registers
    visible_body(g, m, d.vis);
}
template <int F>
__global__ __launch_bounds__(kIntThreads) void kb_integrate(BatchLaunch b) {
    WG g;
    const int s = batch_slot(b, batch_wg(b), g);
    const SessDesc& d = b.desc[b.sess[s]];
    const MapView m = d.m;  // private copy: the body keeps the fields in scalar registers
    integrate_body<F, true>(g, m, slot_fuse(b, d, s));
}

// work-groups per session for the grid-stride stages
static int per_slot(int total, int n, int lo, int hi) {
    int v = total / (n > 0 ? n : 1);
    return v < lo ? lo : (v > hi ? hi : v);
}

void launch_b_alloc(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    hipLaunchKernelGGL(kb_begin_pass, dim3(b.n), dim3(64), 0, s, b);
    // the integration inputs were packed by the normal pass (kb_nmap)
    int grid = batch_layout(b, h, BT_TILES, 0, 0);
    hipLaunchKernelGGL(kb_alloc_a, dim3(grid), dim3(kAllocThreads), 0, s, b);
    grid = batch_layout(b, h, BT_FIXED, 0, per_slot(tuning().alloc_groups, b.n, 16, 512));
    hipLaunchKernelGGL(kb_alloc_b, dim3(grid), dim3(256), 0, s, b);
}
void launch_b_visible(BatchLaunch& b, const SessDesc* h, hipStream_t s) {
    const int grid = batch_layout(b, h, BT_FIXED, 0, per_slot(tuning().visible_groups, b.n, 64, 1024));
    hipLaunchKernelGGL(kb_visible, dim3(grid), dim3(256), 0, s, b);
}
void launch_b_integrate(BatchLaunch& b, const SessDesc* h, int fmt, hipStream_t s) {
    const int grid = batch_layout(b, h, BT_FIXED, 0, per_slot(tuning().integrate_groups, b.n, 256, 8192));
    b.xcd = xcd_order(FU_STAGE_INTEGRATE);
    switch (fmt) {
        case VF_F32_W8: hipLaunchKernelGGL(kb_integrate<VF_F32_W8>, dim3(grid), dim3(kIntThreads), 0, s, b); break;
        case VF_F32_W16: hipLaunchKernelGGL(kb_integrate<VF_F32_W16>, dim3(grid), dim3(kIntThreads), 0, s, b); break;
        default: hipLaunchKernelGGL(kb_integrate<VF_F16_W8>, dim3(grid), dim3(kIntThreads), 0, s, b); break;
    }
}

}  // namespace fu
