// Two-pass marching cubes over all allocated blocks.
//
// Reference: Mapping/MeshScene.cu (checkBlocks :20-41, readNormal :43-103,
// readVertexAndColor :105-140, interp :142-150, buildVertexList :152-259,
// MarchingCube :261-295).  The reference runs one pass with a global atomic triangle
// counter and 32 hash lookups per voxel.  Here:
//   count  one work-group per block stages the block plus a one-voxel halo below and a
//          two-voxel halo above (11^3 sdf and colour values, from the 27 neighbouring
//          blocks looked up once) in LDS, classifies each voxel and sums its triangles;
//   scan   exclusive scan of the per-block counts (chunk sums, chunk scan, local scan);
//   emit   re-stages blocks that have triangles and writes them at their scanned offsets,
//          in voxel-index order within a block.
// Allocated blocks are read straight from the pool-ordered key array, so no compaction
// pass over the N_E hash entries is needed (checkBlocks).
#include "fu_device.h"
#include "fu_kernels.h"
#include "mc_tables.h"

namespace fu {

__constant__ unsigned long long c_mc_tri[256];
__constant__ unsigned short c_mc_edge[256];

void upload_mc_tables() {
    unsigned long long tri[256];
    unsigned short edge[256];
    for (int c = 0; c < 256; ++c) {
        tri[c] = mc_packed(c);
        edge[c] = (unsigned short)mc_edge_mask(c);
    }
    (void)hipMemcpyToSymbol(HIP_SYMBOL(c_mc_tri), tri, sizeof(tri));
    (void)hipMemcpyToSymbol(HIP_SYMBOL(c_mc_edge), edge, sizeof(edge));
}

// corners joined by edge e: edges 0-7 run around the bottom and top faces, 8-11 are the
// vertical edges (MeshScene.cu:185-256)
__device__ __forceinline__ int kMcEdgeCornerDev(int e, int k) {
    if (e < 8) return k == 0 ? e : ((e & 4) | ((e + 1) & 3));
    return k == 0 ? e - 8 : e - 4;
}

__global__ void k_zero_ctr2(unsigned* ctr) {
    if (threadIdx.x == 0) {
        ctr[C_TRI] = 0u;
        ctr[C_OCC] = 0u;
    }
}

constexpr int kHalo = 11;  // voxels -1 .. 9 relative to the block origin
constexpr int kHalo3 = kHalo * kHalo * kHalo;

struct MeshSmem {
    int nb[27];
    float sdf[kHalo3];
    unsigned rgb[kHalo3];
    int scan[256];
    int tot;
};

__device__ __forceinline__ int hidx(int x, int y, int z) { return ((z + 1) * kHalo + (y + 1)) * kHalo + (x + 1); }

template <int F>
__device__ void stage_block(const MapView& m, i3 bp, MeshSmem& s, bool with_rgb) {
    using IO = VoxelIO<F>;
    const int t = threadIdx.x;
    if (t < 27) {
        const i3 nbp = mki3(bp.x + t % 3 - 1, bp.y + (t / 3) % 3 - 1, bp.z + t / 9 - 1);
        s.nb[t] = find_block(m, nbp);
    }
    __syncthreads();
    for (int i = t; i < kHalo3; i += blockDim.x) {
        const int rx = i % kHalo - 1, ry = (i / kHalo) % kHalo - 1, rz = i / (kHalo * kHalo) - 1;
        const int ox = rx < 0 ? 0 : (rx >= 8 ? 2 : 1);
        const int oy = ry < 0 ? 0 : (ry >= 8 ? 2 : 1);
        const int oz = rz < 0 ? 0 : (rz >= 8 ? 2 : 1);
        const int heap = s.nb[oz * 9 + oy * 3 + ox];
        float sd = fnan();
        unsigned c = 0;
        if (heap >= 0) {
            const unsigned char* blk = IO::block(m, heap);
            const int li = ((rz & 7) * 8 + (ry & 7)) * 8 + (rx & 7);
            sd = IO::sdf(blk, li);
            if (with_rgb) {
                const VoxAux x = IO::aux(blk, li);
                c = (unsigned)x.r | ((unsigned)x.g << 8) | ((unsigned)x.b << 16);
            }
        }
        s.sdf[i] = sd;
        s.rgb[i] = c;
    }
    __syncthreads();
}

// corner c of the cell (MeshScene.cu:107-135 order)
__device__ __forceinline__ void corner_off(int c, int& dx, int& dy, int& dz) {
    dx = (c & 1) ^ ((c >> 1) & 1);
    dy = (c >> 1) & 1;
    dz = (c >> 2) & 1;
}

// classify the cell whose origin is local voxel (lx,ly,lz): returns the cube index, or -1
// when a corner is NaN or exactly 1.0, a central-difference neighbour is NaN, or the case
// has no edges (buildVertexList, MeshScene.cu:152-183)
__device__ __forceinline__ int classify(const MeshSmem& s, int lx, int ly, int lz, float sd[8]) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int dx, dy, dz;
        corner_off(c, dx, dy, dz);
        sd[c] = s.sdf[hidx(lx + dx, ly + dy, lz + dz)];
        if (sd[c] == 1.0f || isnanf_(sd[c])) return -1;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int dx, dy, dz;
        corner_off(c, dx, dy, dz);
        const int x = lx + dx, y = ly + dy, z = lz + dz;
        if (isnanf_(s.sdf[hidx(x - 1, y, z)]) || isnanf_(s.sdf[hidx(x, y - 1, z)]) ||
            isnanf_(s.sdf[hidx(x, y, z - 1)]) || isnanf_(s.sdf[hidx(x + 1, y, z)]) ||
            isnanf_(s.sdf[hidx(x, y + 1, z)]) || isnanf_(s.sdf[hidx(x, y, z + 1)]))
            return -1;
    }
    int cube = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (sd[c] < 0) cube |= 1 << c;
    if (c_mc_edge[cube] == 0) return -1;
    return cube;
}

__device__ __forceinline__ int cube_ntri(int cube) { return cube <= 0 ? 0 : (int)(c_mc_tri[cube] >> 60); }

template <int F>
__global__ __launch_bounds__(256) void k_mesh_count(MapView m, unsigned* counts) {
    __shared__ MeshSmem s;
    const unsigned nalloc = min(m.ctr[C_ALLOC], m.num_blocks);
    for (unsigned b = blockIdx.x; b < nalloc; b += gridDim.x) {
        const unsigned long long key = m.block_key[b];
        if (key == kEmptyKey) {
            if (threadIdx.x == 0) counts[b] = 0;
            continue;
        }
        const i3 bp = unpack_key(key);
        stage_block<F>(m, bp, s, false);
        int n = 0;
        for (int h = 0; h < 2; ++h) {
            const int li = threadIdx.x + h * 256;
            float sd[8];
            n += cube_ntri(classify(s, li & 7, (li >> 3) & 7, li >> 6, sd));
        }
        // work-group sum
        float fn = wave_sum((float)n);
        if (threadIdx.x == 0) s.tot = 0;
        __syncthreads();
        if (lane_id() == 0) atomicAdd(&s.tot, (int)fn);
        __syncthreads();
        if (threadIdx.x == 0) {
            counts[b] = (unsigned)s.tot;
            if (s.tot > 0) atomicAdd(&m.ctr[C_OCC], 1u);
        }
        __syncthreads();
    }
}

// ---- exclusive scan of counts[0..n) (n read from the allocation counter) ----------------
constexpr int kChunk = 2048;

__device__ __forceinline__ int block_exclusive_scan256(int v, int* sh, int& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const int a = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    total = sh[255];
    const int ex = sh[t] - v;
    __syncthreads();
    return ex;
}

__global__ __launch_bounds__(256) void k_scan_chunks(const MapView m, const unsigned* counts, unsigned* chunk_sum,
                                                    unsigned nchunks_cap) {
    __shared__ int sh[256];
    const unsigned n = min(m.ctr[C_ALLOC], m.num_blocks);
    for (unsigned c = blockIdx.x; c < nchunks_cap; c += gridDim.x) {
        const unsigned base = c * kChunk;
        int acc = 0;
        for (int k = 0; k < kChunk / 256; ++k) {
            const unsigned i = base + k * 256 + threadIdx.x;
            acc += i < n ? (int)counts[i] : 0;
        }
        int tot;
        block_exclusive_scan256(acc, sh, tot);
        if (threadIdx.x == 0) chunk_sum[c] = (unsigned)tot;
    }
}

__global__ __launch_bounds__(256) void k_scan_top(MapView m, unsigned* chunk_sum, unsigned nchunks_cap) {
    __shared__ int sh[256];
    unsigned carry = 0;
    for (unsigned b0 = 0; b0 < nchunks_cap; b0 += 256) {
        const unsigned i = b0 + threadIdx.x;
        const int v = i < nchunks_cap ? (int)chunk_sum[i] : 0;
        int tot;
        const int ex = block_exclusive_scan256(v, sh, tot);
        if (i < nchunks_cap) chunk_sum[i] = carry + (unsigned)ex;
        carry += (unsigned)tot;
    }
    if (threadIdx.x == 0) m.ctr[C_TRI] = carry;
}

__global__ __launch_bounds__(256) void k_scan_local(MapView m, unsigned* counts, const unsigned* chunk_base,
                                                   unsigned nchunks_cap) {
    __shared__ int sh[256];
    const unsigned n = min(m.ctr[C_ALLOC], m.num_blocks);
    for (unsigned c = blockIdx.x; c < nchunks_cap; c += gridDim.x) {
        unsigned carry = chunk_base[c];
        for (int k = 0; k < kChunk / 256; ++k) {
            const unsigned i = c * kChunk + k * 256 + threadIdx.x;
            const int v = i < n ? (int)counts[i] : 0;
            int tot;
            const int ex = block_exclusive_scan256(v, sh, tot);
            if (i < n) counts[i] = carry + (unsigned)ex;  // counts become offsets
            carry += (unsigned)tot;
        }
    }
}

// ---- emit --------------------------------------------------------------------------------
__device__ __forceinline__ float mc_interp(float v1, float v2) {
    // MeshScene.cu:142-150 (fabs of a float compared against the double 1e-6)
    if ((double)fabsf(0.0f - v1) < 1e-6) return 0.0f;
    if ((double)fabsf(0.0f - v2) < 1e-6) return 1.0f;
    if ((double)fabsf(v1 - v2) < 1e-6) return 0.0f;
    return (0.0f - v1) / (v2 - v1);
}

template <int F>
__global__ __launch_bounds__(256) void k_mesh_emit(MapView m, const unsigned* offsets, const unsigned* counts_dummy,
                                                  MeshOut o) {
    __shared__ MeshSmem s;
    (void)counts_dummy;
    const unsigned nalloc = min(m.ctr[C_ALLOC], m.num_blocks);
    const unsigned total = m.ctr[C_TRI];
    for (unsigned b = blockIdx.x; b < nalloc; b += gridDim.x) {
        const unsigned long long key = m.block_key[b];
        if (key == kEmptyKey) continue;
        const unsigned start = offsets[b];
        const unsigned end = (b + 1 < nalloc) ? offsets[b + 1] : total;
        if (end == start || start >= o.max_tri) continue;
        const i3 bp = unpack_key(key);
        stage_block<F>(m, bp, s, true);
        int cube[2];
        float sd[2][8];
        int nt[2];
        for (int h = 0; h < 2; ++h) {
            const int li = threadIdx.x + h * 256;
            cube[h] = classify(s, li & 7, (li >> 3) & 7, li >> 6, sd[h]);
            nt[h] = cube_ntri(cube[h]);
        }
        int tot0, tot1;
        const int ex0 = block_exclusive_scan256(nt[0], s.scan, tot0);
        const int ex1 = block_exclusive_scan256(nt[1], s.scan, tot1);
        for (int h = 0; h < 2; ++h) {
            if (nt[h] == 0) continue;
            const int li = threadIdx.x + h * 256;
            const int lx = li & 7, ly = (li >> 3) & 7, lz = li >> 6;
            const i3 vp = mki3(bp.x * 8 + lx, bp.y * 8 + ly, bp.z * 8 + lz);
            unsigned tri0 = start + (unsigned)(h == 0 ? ex0 : tot0 + ex1);
            // corner normals: central differences (MeshScene.cu:43-103)
            f3 nrm[8];
            unsigned col[8];
            for (int c = 0; c < 8; ++c) {
                int dx, dy, dz;
                corner_off(c, dx, dy, dz);
                const int x = lx + dx, y = ly + dy, z = lz + dz;
                nrm[c] = mk3(s.sdf[hidx(x + 1, y, z)] - s.sdf[hidx(x - 1, y, z)],
                             s.sdf[hidx(x, y + 1, z)] - s.sdf[hidx(x, y - 1, z)],
                             s.sdf[hidx(x, y, z + 1)] - s.sdf[hidx(x, y, z - 1)]);
                col[c] = s.rgb[hidx(x, y, z)];
            }
            const unsigned long long pk = c_mc_tri[cube[h]];
            const float* sv = sd[h];
            for (int t = 0; t < nt[h]; ++t) {
                const unsigned tid = tri0 + t;
                if (tid >= o.max_tri) break;
                for (int k = 0; k < 3; ++k) {
                    const int e = (int)((pk >> (4 * (3 * t + k))) & 0xF);
                    const int ca = kMcEdgeCornerDev(e, 0), cb = kMcEdgeCornerDev(e, 1);
                    const float val = mc_interp(sv[ca], sv[cb]);
                    int ax, ay, az, bx, by, bz;
                    corner_off(ca, ax, ay, az);
                    corner_off(cb, bx, by, bz);
                    const float ox = bx == ax ? (float)ax : (bx > ax ? (float)ax + val : (float)ax - val);
                    const float oy = by == ay ? (float)ay : (by > ay ? (float)ay + val : (float)ay - val);
                    const float oz = bz == az ? (float)az : (bz > az ? (float)az + val : (float)az - val);
                    const f3 vtx = mul(mk3((float)vp.x + ox, (float)vp.y + oy, (float)vp.z + oz), m.voxel_size);
                    const f3 nv = normalised(add(nrm[ca], mul(val, sub(nrm[cb], nrm[ca]))));
                    const int iv = f2i(val);
                    unsigned char cc[3];
                    for (int ch = 0; ch < 3; ++ch) {
                        const unsigned a8 = (col[ca] >> (8 * ch)) & 0xFF, b8 = (col[cb] >> (8 * ch)) & 0xFF;
                        const unsigned char d = (unsigned char)(b8 - a8);
                        const unsigned char prod = (unsigned char)(int)(d * iv);
                        cc[ch] = (unsigned char)(a8 + prod);
                    }
                    const size_t vi = (size_t)tid * 3 + k;
                    o.vertex[vi * 3 + 0] = vtx.x;
                    o.vertex[vi * 3 + 1] = vtx.y;
                    o.vertex[vi * 3 + 2] = vtx.z;
                    o.normal[vi * 3 + 0] = nv.x;
                    o.normal[vi * 3 + 1] = nv.y;
                    o.normal[vi * 3 + 2] = nv.z;
                    o.color[vi * 3 + 0] = cc[0];
                    o.color[vi * 3 + 1] = cc[1];
                    o.color[vi * 3 + 2] = cc[2];
                }
                if (o.tri_meta) {
                    o.tri_meta[(size_t)tid * 4 + 0] = vp.x;
                    o.tri_meta[(size_t)tid * 4 + 1] = vp.y;
                    o.tri_meta[(size_t)tid * 4 + 2] = vp.z;
                    o.tri_meta[(size_t)tid * 4 + 3] = cube[h] * 16 + t;
                }
            }
        }
        __syncthreads();
    }
}

void launch_mesh(const MapView& m, const MeshOut& o, unsigned* counts, unsigned* chunk, hipStream_t s) {
    const unsigned nchunks = (m.num_blocks + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_zero_ctr2, dim3(1), dim3(64), 0, s, m.ctr);
    switch (m.fmt) {
        case VF_F32_W8: hipLaunchKernelGGL(k_mesh_count<VF_F32_W8>, dim3(4096), dim3(256), 0, s, m, counts); break;
        case VF_F32_W16: hipLaunchKernelGGL(k_mesh_count<VF_F32_W16>, dim3(4096), dim3(256), 0, s, m, counts); break;
        default: hipLaunchKernelGGL(k_mesh_count<VF_F16_W8>, dim3(4096), dim3(256), 0, s, m, counts); break;
    }
    hipLaunchKernelGGL(k_scan_chunks, dim3(min(nchunks, 1024u)), dim3(256), 0, s, m, counts, chunk, nchunks);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, s, m, chunk, nchunks);
    hipLaunchKernelGGL(k_scan_local, dim3(min(nchunks, 1024u)), dim3(256), 0, s, m, counts, chunk, nchunks);
    switch (m.fmt) {
        case VF_F32_W8: hipLaunchKernelGGL(k_mesh_emit<VF_F32_W8>, dim3(4096), dim3(256), 0, s, m, counts, counts, o); break;
        case VF_F32_W16: hipLaunchKernelGGL(k_mesh_emit<VF_F32_W16>, dim3(4096), dim3(256), 0, s, m, counts, counts, o); break;
        default: hipLaunchKernelGGL(k_mesh_emit<VF_F16_W8>, dim3(4096), dim3(256), 0, s, m, counts, counts, o); break;
    }
}

}  // namespace fu
