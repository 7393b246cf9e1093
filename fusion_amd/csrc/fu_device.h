// Device-side views and helpers shared by the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include "fu_math.h"

namespace fu {

// Global-memory view of a pointer.  Pointers read from descriptors in memory are generic to
// the compiler (flat instructions, 64-bit VGPR addresses, no scalar loads); the explicit
// global address space gives global loads with a scalar base, and scalar loads for uniform
// addresses.
#if defined(__HIP_DEVICE_COMPILE__)
#define FU_GPTR(T) __attribute__((address_space(1))) T*
#else
#define FU_GPTR(T) T*
#endif
#define FU_G(T, p) ((FU_GPTR(T))(p))

// One hash-table record, 16 bytes, read with a single dwordx4 load.  It carries what the
// reference's 32-byte HashEntry carries (Mapping/DeviceMap.h:20-59): the block position
// (packed in `key`), `ptr` = voxel offset of the block (heap index * 512, -1 = free) and
// `next` = 1-based index into the excess area (0 ends the chain).
struct alignas(16) HEntry {
    unsigned long long key;
    int ptr;
    int next;
};

// Map-wide counters in device memory.
enum Counter : int {
    C_ALLOC = 0,    // blocks taken from the pool (bump index; may exceed capacity, clamp on read)
    C_EXCESS = 1,   // excess slots used
    C_VISIBLE = 2,  // entries of the visible list
    C_NEW = 3,      // new keys found by the current allocation pass
    C_FLAGS = 4,    // sticky overflow flags (FU_FLAG_*)
    C_RENDER = 5,   // rendering fragments of the last z-range pass
    C_TRI = 6,      // triangles of the last mesh pass
    C_DONE = 7,     // work-group completion counter (last-group pattern), left at 0
    C_OCC = 8,      // occupied blocks seen by the last mesh pass
    C_SPARE = 9,
    C_COUNT = 16
};

enum Flag : unsigned {
    FLAG_HEAP_FULL = 1u,
    FLAG_EXCESS_FULL = 2u,
    FLAG_NEWSET_FULL = 4u,
    FLAG_COORD_RANGE = 8u,
    FLAG_RENDER_FULL = 16u,
    FLAG_TRI_FULL = 32u,
};

// Voxel storage formats.  Each block of 8^3 voxels is stored as planes of 512 values in
// voxel order z*64 + y*8 + x:
//   u8 weights:  sdf[512] (f32 or f16) | wrgb[512] (u32: weight | r << 8 | g << 16 | b << 24)
//   u16 weights: sdf[512] (f32)        | weight[512] (u16) | r[512] | g[512] | b[512]
// so a wave touches 64 consecutive values of a plane, and an update is one sdf and one
// weight+colour access in the u8 formats.
enum VoxelFormat : int { VF_F32_W8 = 0, VF_F32_W16 = 1, VF_F16_W8 = 2 };

FU_HD int vf_sdf_bytes(int f) { return f == VF_F16_W8 ? 2 : 4; }
FU_HD int vf_w_bytes(int f) { return f == VF_F32_W16 ? 2 : 1; }
FU_HD int vf_block_bytes(int f) { return 512 * (vf_sdf_bytes(f) + vf_w_bytes(f) + 3); }

struct MapView {
    HEntry* entries;            // [num_buckets + num_excess]
    unsigned* ctr;              // [C_COUNT]
    unsigned long long* block_key;  // [num_blocks], key of each allocated block (kEmptyKey = hole)
    unsigned* visible;          // [num_blocks] block ids of the visible list
    unsigned long long* visible_key;  // [num_blocks] their keys (read without the id -> key hop)
    unsigned long long* newset; // [newset_mask + 1] per-pass set of new keys
    unsigned long long* newkeys;  // [max_new]
    unsigned* newslot;          // [max_new]
    unsigned char* pool;        // voxel blocks
    unsigned* occ;              // occupancy bitmap, one bit per hash of every allocated key
    unsigned occ_mask, occ_shift;
    unsigned bucket_mask;       // num_buckets - 1 when it is a power of two, else 0
    unsigned num_buckets, num_excess, num_blocks, newset_mask, max_new;
    unsigned max_weight;
    int fmt;
    float voxel_size, voxel_inv, trunc, depth_min, depth_max;
};

// ---- voxel plane access -----------------------------------------------------------------
struct VoxAux {
    unsigned w;
    unsigned char r, g, b;
};

template <int F>
struct VoxelIO {
    static constexpr int kSdfB = F == VF_F16_W8 ? 2 : 4;
    static constexpr bool kW16 = F == VF_F32_W16;
    static constexpr int kBlock = 512 * (kSdfB + (kW16 ? 2 : 1) + 3);
    __device__ static unsigned char* block(const MapView& m, int heap) {
        return m.pool + (size_t)heap * kBlock;
    }
    __device__ static float sdf(const unsigned char* b, int i) {
        if constexpr (kSdfB == 2)
            return __half2float(FU_G(const __half, b)[i]);
        else
            return FU_G(const float, b)[i];
    }
    __device__ static void set_sdf(unsigned char* b, int i, float v) {
        if constexpr (kSdfB == 2)
            FU_G(__half, b)[i] = __float2half_rn(v);
        else
            FU_G(float, b)[i] = v;
    }
    __device__ static VoxAux aux(const unsigned char* b, int i) {
        VoxAux a;
        if constexpr (kW16) {
            const auto c = FU_G(const unsigned char, b + 512 * 6);
            a.w = FU_G(const unsigned short, b + 512 * 4)[i];
            a.r = c[i];
            a.g = c[512 + i];
            a.b = c[1024 + i];
        } else {
            const unsigned v = FU_G(const unsigned, b + 512 * kSdfB)[i];
            a.w = v & 0xFFu;
            a.r = (unsigned char)(v >> 8);
            a.g = (unsigned char)(v >> 16);
            a.b = (unsigned char)(v >> 24);
        }
        return a;
    }
    __device__ static void set_aux(unsigned char* b, int i, VoxAux a) {
        if constexpr (kW16) {
            const auto c = FU_G(unsigned char, b + 512 * 6);
            FU_G(unsigned short, b + 512 * 4)[i] = (unsigned short)a.w;
            c[i] = a.r;
            c[512 + i] = a.g;
            c[1024 + i] = a.b;
        } else {
            FU_G(unsigned, b + 512 * kSdfB)[i] =
                (a.w & 0xFFu) | ((unsigned)a.r << 8) | ((unsigned)a.g << 16) | ((unsigned)a.b << 24);
        }
    }
    __device__ static unsigned weight(const unsigned char* b, int i) {
        if constexpr (kW16)
            return FU_G(const unsigned short, b + 512 * 4)[i];
        else
            return FU_G(const unsigned, b + 512 * kSdfB)[i] & 0xFFu;
    }
};

// ---- hash lookup (DeviceMap::FindEntry, DeviceMap.cu:122-135) ----------------------------
__device__ __forceinline__ HEntry load_entry(const HEntry* e) {
    // one 16-byte load
    const uint4 v = *FU_G(const uint4, e);
    HEntry h;
    h.key = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
    h.ptr = (int)v.z;
    h.next = (int)v.w;
    return h;
}

FU_HD unsigned long long mix64(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

// A clear occupancy bit proves the key is not allocated (bits are set for every allocated
// key; collisions only cost a probe).  The bitmap is small enough to stay in L2, so rays
// stepping through empty space skip the dependent hash-table load.  Bit index: the top bits
// of a multiplicative scramble of the 32-bit block hash.
FU_HD unsigned occ_bit(const MapView& m, uint32_t raw) { return (raw * 0x9E3779B1u) >> m.occ_shift; }
__device__ __forceinline__ bool maybe_allocated(const MapView& m, uint32_t raw) {
    const unsigned h = occ_bit(m, raw);
    return (FU_G(const unsigned, m.occ)[h >> 5] >> (h & 31)) & 1u;
}

// heap index of block `bp`, or -1 if it is not allocated
__device__ __forceinline__ int find_block(const MapView& m, i3 bp) {
    if (!key_in_range(bp)) return -1;
    const uint32_t raw = block_hash_raw(bp);
    if (!maybe_allocated(m, raw)) return -1;
    const unsigned long long key = pack_key(bp);
    HEntry h = load_entry(&m.entries[bucket_of(raw, m.num_buckets, m.bucket_mask)]);
    while (true) {
        if (h.key == key && h.ptr >= 0) return h.ptr >> 9;
        if (h.next <= 0) return -1;
        h = load_entry(&m.entries[m.num_buckets + h.next - 1]);
    }
}

// ---- wave64 helpers -------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int lane_rank(unsigned long long mask) {
    // number of set bits of `mask` below this lane
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}
__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int lane) {
    unsigned lo = __builtin_amdgcn_readlane((unsigned)v, lane);
    unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), lane);
    return (unsigned long long)lo | ((unsigned long long)hi << 32);
}

// wave-aggregated atomicAdd of 1 per lane with `pred`; returns this lane's slot
__device__ __forceinline__ unsigned wave_append(unsigned* counter, bool pred) {
    unsigned long long m = ballot(pred);
    if (m == 0) return 0;
    int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (unsigned)__popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    return base + lane_rank(m);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Pairwise tree sum over the 64 lanes of a wave: DPP within rows of 16 lanes
// (quad swaps, half-row and row mirrors), then cross-row swizzles.  Every lane ends
// with the wave total.
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_f<0x141>(v);  // row_half_mirror
    v += dpp_f<0x140>(v);  // row_mirror
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// Sums 32 per-lane values over the 64 lanes of a wave as a reduce-scatter: each step adds
// partial sums of disjoint lane sets (a pairwise tree) and halves the number of values a lane
// carries.  Step 1 swaps the wave halves (v_permlane32_swap), step 2 odd and even rows
// (v_permlane16_swap), steps 3-6 use DPP within rows (row_ror:8, row_half_mirror, quad
// permutes).  Afterwards lanes 2k and 2k+1 hold the total of value k; about 70 instructions
// instead of 32 separate wave trees.  `v` is clobbered.
__device__ __forceinline__ float wave_reduce_scatter32(float (&v)[32]) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {  // lanes 0-31 keep values 0-15, lanes 32-63 values 16-31
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[j]), __float_as_uint(v[j + 16]), false, false);
        v[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {  // even rows keep the lower 8 of their 16, odd rows the upper 8
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[j]), __float_as_uint(v[j + 8]), false, false);
        v[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const int lane = lane_id();
    bool hi = (lane & 8) != 0;  // partner lane ^ 8
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float keep = hi ? v[j + 4] : v[j], send = hi ? v[j] : v[j + 4];
        v[j] = keep + dpp_f<0x128>(send);  // row_ror:8
    }
    hi = (lane & 4) != 0;  // partner 7 - q within each half row
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float keep = hi ? v[j + 2] : v[j], send = hi ? v[j] : v[j + 2];
        v[j] = keep + dpp_f<0x141>(send);  // row_half_mirror
    }
    hi = (lane & 2) != 0;  // partner lane ^ 2
    {
        const float keep = hi ? v[1] : v[0], send = hi ? v[0] : v[1];
        v[0] = keep + dpp_f<0x4E>(send);  // quad_perm [2,3,0,1]
    }
    return v[0] + dpp_f<0xB1>(v[0]);  // quad_perm [1,0,3,2]
}

// The same reduce-scatter for 64 values: one more halving step, so that afterwards lane k
// holds the wave total of value k.  `v` is clobbered.
__device__ __forceinline__ float wave_reduce_scatter64(float (&v)[64]) {
#pragma unroll
    for (int j = 0; j < 32; ++j) {  // lanes 0-31 keep values 0-31, lanes 32-63 values 32-63
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[j]), __float_as_uint(v[j + 32]), false, false);
        v[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {  // even rows keep the lower 16 of their 32, odd rows the upper 16
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[j]), __float_as_uint(v[j + 16]), false, false);
        v[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const int lane = lane_id();
    bool hi = (lane & 8) != 0;  // partner lane ^ 8
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float keep = hi ? v[j + 8] : v[j], send = hi ? v[j] : v[j + 8];
        v[j] = keep + dpp_f<0x128>(send);  // row_ror:8
    }
    hi = (lane & 4) != 0;  // partner 7 - q within each half row
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float keep = hi ? v[j + 4] : v[j], send = hi ? v[j] : v[j + 4];
        v[j] = keep + dpp_f<0x141>(send);  // row_half_mirror
    }
    hi = (lane & 2) != 0;  // partner lane ^ 2
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float keep = hi ? v[j + 2] : v[j], send = hi ? v[j] : v[j + 2];
        v[j] = keep + dpp_f<0x4E>(send);  // quad_perm [2,3,0,1]
    }
    hi = (lane & 1) != 0;  // partner lane ^ 1
    const float keep = hi ? v[1] : v[0], send = hi ? v[0] : v[1];
    return keep + dpp_f<0xB1>(send);  // quad_perm [1,0,3,2]
}

// IEEE float division a / b without the range scaling and special-case steps.  HIP compiles
// `a / b` to v_div_scale (x2), v_rcp, a Newton refinement of the reciprocal, two corrections
// of the quotient, v_div_fmas and v_div_fixup; for finite operands with |a| and |b| in
// [2^-60, 2^30] (or a = +0), whose quotient is normal, the scale steps are the identity and
// the fixup returns its input, so the core below gives the same bits.  Several quotients
// with one denominator share its refined reciprocal (fu_rcp_refined).
__device__ __forceinline__ float fu_rcp_refined(float b) {
    const float y0 = __builtin_amdgcn_rcpf(b);
    const float e = fmaf(-b, y0, 1.0f);
    return fmaf(e, y0, y0);
}
__device__ __forceinline__ float fu_div_refined(float a, float b, float y) {
    const float m = a * y;
    const float f2 = fmaf(-b, m, a);
    const float f3 = fmaf(f2, y, m);
    const float f4 = fmaf(-b, f3, a);
    return fmaf(f4, y, f3);
}
// IEEE sqrtf without the range scaling and special-value steps: HIP compiles sqrtf to a
// scale for tiny inputs, v_sqrt, a one-ulp correction by the signs of s - y'y for the two
// neighbours y' of y, the unscale and a class test for 0 and inf; for finite s in
// [2^-90, 2^120] the scale, unscale and class steps are the identity.
__device__ __forceinline__ float fu_sqrt_normal(float s) {
    const float y = __builtin_amdgcn_sqrtf(s);
    const float dn = __uint_as_float(__float_as_uint(y) - 1u);
    const float up = __uint_as_float(__float_as_uint(y) + 1u);
    const float r = fmaf(-dn, y, s) <= 0.0f ? dn : y;
    return fmaf(-up, y, s) > 0.0f ? up : r;
}

// whether every component of (x, y, z) lies in the range of fu_div_refined by a denominator
// of at most 2^31 (the 4-D norm of (x, y, z, 1)): 2^-60 <= |c| <= 2^30
__device__ __forceinline__ bool fu_div_range3(float x, float y, float z) {
    const float lo = fminf(fminf(fabsf(x), fabsf(y)), fabsf(z));
    const float hi = fmaxf(fmaxf(fabsf(x), fabsf(y)), fabsf(z));
    return lo >= 0x1p-60f && hi <= 0x1p30f;
}

// wave total of an unsigned (wrapping) sum; every lane gets it
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v += (unsigned)__shfl_xor((int)v, m);
    return v;
}

}  // namespace fu
