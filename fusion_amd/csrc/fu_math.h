// Scalar and small-vector arithmetic with the exact semantics the reference pipeline
// relies on (Utility/VectorMath.h).  Every operation is written out as a named function
// so that the order of the IEEE float operations is explicit: the library is compiled
// with -ffp-contract=off, so each function below performs exactly the multiplies, adds
// and divides it spells, and the results are bit-identical to a strict-IEEE CPU build of
// the same expressions.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define FU_HD __host__ __device__ __forceinline__
#else
#define FU_HD inline
#endif

namespace fu {

struct f3 {
    float x, y, z;
};
struct alignas(16) f4 {
    float x, y, z, w;
};
struct i3 {
    int x, y, z;
};
// row-major 3x3, the reference's Matrix3f (VectorMath.h:356-371)
struct m3 {
    f3 r0, r1, r2;
};

FU_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
FU_HD f4 mk4(float x, float y, float z, float w) { return f4{x, y, z, w}; }
FU_HD i3 mki3(int x, int y, int z) { return i3{x, y, z}; }
FU_HD f3 xyz(const f4& a) { return f3{a.x, a.y, a.z}; }

FU_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
FU_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
FU_HD f3 neg(f3 a) { return f3{-a.x, -a.y, -a.z}; }
FU_HD f3 mul(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
FU_HD f3 mul(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
FU_HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
FU_HD f4 add4(f4 a, f4 b) { return f4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
FU_HD f4 sub4(f4 a, f4 b) { return f4{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
FU_HD f4 divs4(f4 a, float s) { return f4{a.x / s, a.y / s, a.z / s, a.w / s}; }
// float3 . float3 (VectorMath.h:191-193): ((x*x' + y*y') + z*z')
FU_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// float4 . float4 is a 4-D dot including w (VectorMath.h:203-205)
FU_HD float dot4(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
FU_HD f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
FU_HD float norm(f3 a) { return sqrtf(dot(a, a)); }
FU_HD float norm4(f4 a) { return sqrtf(dot4(a, a)); }
// normalised(float3) = a / |a| componentwise (VectorMath.h:336-338)
FU_HD f3 normalised(f3 a) { return divs(a, norm(a)); }
// normalised(float4) divides all four components by the 4-D norm (VectorMath.h:340-342)
FU_HD f4 normalised4(f4 a) { return divs4(a, norm4(a)); }
FU_HD f3 floor3(f3 a) { return f3{floorf(a.x), floorf(a.y), floorf(a.z)}; }
// Matrix3f * float3 (VectorMath.h:364-366)
FU_HD f3 mv(const m3& m, f3 a) { return f3{dot(m.r0, a), dot(m.r1, a), dot(m.r2, a)}; }
// int3 * float (VectorMath.h:231-233): each int converted to float first
FU_HD f3 i3f(i3 a, float s) { return f3{(float)a.x * s, (float)a.y * s, (float)a.z * s}; }
FU_HD i3 addi(i3 a, i3 b) { return i3{a.x + b.x, a.y + b.y, a.z + b.z}; }
FU_HD bool eqi(i3 a, i3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

// float -> int conversion with the saturating, NaN->0 behaviour of the GPU conversion
// instruction (the reference's make_int3/make_int2/(int) casts, VectorMath.h:28-38).  C++
// leaves out-of-range conversions undefined, so the rule is spelt out and shared with the
// CPU oracle.
FU_HD int f2i(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    // v_cvt_i32_f32 truncates, saturates out-of-range values and converts NaN to 0: the rule
    // below, in one instruction
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
#else
    if (!(v == v)) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
#endif
}
FU_HD int d2i(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    // v_cvt_i32_f64: truncation, saturation and NaN -> 0, as for f2i
    int r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(v));
    return r;
#else
    if (!(v == v)) return 0;
    if (v >= 2147483648.0) return 2147483647;
    if (v <= -2147483648.0) return (-2147483647 - 1);
    return (int)v;
#endif
}
FU_HD i3 f2i3(f3 a) { return i3{f2i(a.x), f2i(a.y), f2i(a.z)}; }
// (unsigned char)(int)v : truncation then wrap modulo 256 (VectorMath.h:16-18)
FU_HD uint8_t f2u8(float v) { return (uint8_t)(unsigned)f2i(v); }

FU_HD float fnan() { return __builtin_nanf(""); }

// (float)i / 255.f for every byte value, evaluated at compile time with IEEE division (the
// colour blend of integrateColor divides bytes by 255, FuseMap.cu:180-186)
struct U8Div255 {
    float v[256];
    constexpr U8Div255() : v() {
        for (int i = 0; i < 256; ++i) v[i] = (float)i / 255.f;
    }
};
FU_HD bool isnanf_(float v) { return !(v == v); }

// ---------------------------------------------------------------------------------------
// exp for the bilateral filter weights (Tracking/Pyrdown.cu:33 calls exp on a float).
// Own implementation, shared bit-for-bit by the GPU kernels and the CPU oracle, so the
// filtered depth is identical on both: Cody-Waite reduction x = k*ln2 + r with
// |r| <= ln2/2 (ln2 split into a 16-bit-exact high part and a low part), a degree-7
// Taylor polynomial for e^r in Horner form (truncation error about 5e-9 relative on the
// interval, below float rounding), and scaling by 2^k built in the exponent field.
// Accurate to about 1 ulp against a correctly rounded exp on [-87, 88].
// the reduction and polynomial of fu_expf: x = k*ln2 + r, p = e^r in [2^-1, 2^1)
FU_HD float fu_expf_poly(float x, int& k) {
    const float kLog2e = 1.44269502162933349609375f;
    const float kLn2Hi = 0.693145751953125f;         // 16 significant bits: k*kLn2Hi exact
    const float kLn2Lo = 1.428606765330187045e-06f;  // ln2 - kLn2Hi
    float kf = rintf(x * kLog2e);
    k = (int)kf;
    float r = (x - kf * kLn2Hi) - kf * kLn2Lo;
    float p = 1.9841270e-4f;              // 1/5040
    p = p * r + 1.3888889e-3f;            // 1/720
    p = p * r + 8.3333333e-3f;            // 1/120
    p = p * r + 4.1666668e-2f;            // 1/24
    p = p * r + 1.6666667e-1f;            // 1/6
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    return p;
}
FU_HD float fu_expf(float x) {
    if (!(x == x)) return x;
    if (x > 88.5f) return __builtin_inff();
    if (x < -103.0f) return 0.0f;
    int k;
    const float p = fu_expf_poly(x, k);
    // 2^k in two steps so that k down to -149 and up to 128 stay representable
    int k1 = k / 2, k2 = k - k1;
    union { uint32_t u; float f; } s1, s2;
    s1.u = (uint32_t)(k1 + 127) << 23;
    s2.u = (uint32_t)(k2 + 127) << 23;
    return (p * s1.f) * s2.f;
}
// fu_expf for x <= 0 that is not NaN (the bilateral weights), without the NaN and overflow
// tests and with the scaling in one exponent-field add and one multiply: k is in [-149, 0],
// p * 2^(k+64) is a normal float (exact) and the multiply by 2^-64 rounds once, as the
// second multiply of fu_expf does, so both give the same bits
FU_HD float fu_expf_nonpos(float x) {
    int k;
    const float p = fu_expf_poly(fmaxf(x, -103.0f), k);
    union { uint32_t u; float f; } s;
    union { float f; uint32_t u; } pb;
    pb.f = p;
    s.u = pb.u + ((uint32_t)(k + 64) << 23);
    const float e = s.f * 0x1p-64f;
    return x < -103.0f ? 0.0f : e;
}

// ---------------------------------------------------------------------------------------
// Voxel-block coordinates (Mapping/DeviceMap.cu:137-201, block edge 8 voxels).
// voxelPosToBlockPos (DeviceMap.cu:150-161) shifts negatives by 7 and divides by 8 with
// truncation, i.e. floor division by 8: an arithmetic shift right by 3 on two's complement
// ints.  voxelPosToLocalPos (:167-178) is the non-negative remainder, i.e. the low 3 bits.
FU_HD i3 voxel_to_block(i3 v) { return i3{v.x >> 3, v.y >> 3, v.z >> 3}; }
FU_HD i3 voxel_to_local(i3 v) { return i3{v.x & 7, v.y & 7, v.z & 7}; }
FU_HD int local_idx(i3 l) { return l.z * 64 + l.y * 8 + l.x; }  // DeviceMap.cu:180-182
FU_HD int voxel_local_idx(i3 v) { return ((v.z & 7) << 6) | ((v.y & 7) << 3) | (v.x & 7); }

// DeviceMap::Hash (DeviceMap.cu:3-10).  NumBuckets is an unsigned constant there, so the
// xor of the wrapped int products is reduced modulo it as an unsigned value.
FU_HD uint32_t block_hash_raw(i3 p) {
    return ((uint32_t)p.x * 73856093u) ^ ((uint32_t)p.y * 19349669u) ^ ((uint32_t)p.z * 83492791u);
}
FU_HD uint32_t block_hash(i3 p, uint32_t num_buckets) { return block_hash_raw(p) % num_buckets; }
// the same bucket with a mask when the bucket count is a power of two (mask != 0)
FU_HD uint32_t bucket_of(uint32_t raw, uint32_t num_buckets, uint32_t mask) {
    return mask ? (raw & mask) : (raw % num_buckets);
}

// Block coordinates packed into one 64-bit key, 21 bits per axis (two's complement), so a
// hash entry is a single 16-byte record.  Coordinates must lie in [-2^20, 2^20).
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int kCoordLimit = 1 << 20;
FU_HD bool key_in_range(i3 p) {
    return p.x >= -kCoordLimit && p.x < kCoordLimit && p.y >= -kCoordLimit && p.y < kCoordLimit &&
           p.z >= -kCoordLimit && p.z < kCoordLimit;
}
FU_HD uint64_t pack_key(i3 p) {
    return ((uint64_t)((uint32_t)p.x & 0x1FFFFFu) << 42) | ((uint64_t)((uint32_t)p.y & 0x1FFFFFu) << 21) |
           (uint64_t)((uint32_t)p.z & 0x1FFFFFu);
}
FU_HD int sext21(uint32_t v) { return (int)(v << 11) >> 11; }
FU_HD i3 unpack_key(uint64_t k) {
    return i3{sext21((uint32_t)(k >> 42) & 0x1FFFFFu), sext21((uint32_t)(k >> 21) & 0x1FFFFFu),
              sext21((uint32_t)k & 0x1FFFFFu)};
}

}  // namespace fu
