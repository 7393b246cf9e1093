// Self-test of the range-specialised IEEE arithmetic (fu_device.h: fu_rcp_refined,
// fu_div_refined, fu_sqrt_normal) against the compiler's IEEE division and square root on
// the device itself: the kernels rely on them being bit-identical inside their stated ranges.
// It also checks the single-instruction float/double -> int conversions (fu_math.h: f2i, d2i)
// against the rule they implement (truncation, saturation, NaN -> 0), and the bilateral
// weights' exp (fu_expf_nonpos) against fu_expf.
#include <hip/hip_runtime.h>

#include "fu_api.h"
#include "fu_device.h"

using namespace fu;

namespace {

__device__ __forceinline__ uint64_t mix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}
// the conversion rule of f2i / d2i spelt out with comparisons
template <class T>
__device__ __forceinline__ int conv_rule(T v) {
    if (!(v == v)) return 0;
    if (v >= (T)2147483648.0) return 2147483647;
    if (v <= (T)-2147483648.0) return (-2147483647 - 1);
    return (int)v;
}
// a float with a uniformly random mantissa and an exponent uniform in [lo, hi], random sign
__device__ __forceinline__ float rnd(uint64_t& s, int lo, int hi, bool sign) {
    s = mix(s + 0x9e3779b97f4a7c15ull);
    const unsigned m = (unsigned)s & 0x7fffffu;
    const int e = lo + (int)((s >> 23) % (uint64_t)(hi - lo + 1));
    const unsigned sg = (sign && ((s >> 40) & 1)) ? 0x80000000u : 0u;
    // one in eight mantissas all ones or zero: the reciprocal's hardest and easiest cases
    const unsigned mm = ((s >> 41) & 7) == 0 ? 0x7fffffu : (((s >> 41) & 7) == 1 ? 0u : m);
    return __uint_as_float(sg | ((unsigned)(e + 127) << 23) | mm);
}

__global__ void k_selftest(uint64_t seed, unsigned n, unsigned* bad) {
    unsigned b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0, b5 = 0, b6 = 0;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint64_t s = mix(seed ^ ((uint64_t)i * 0x2545f4914f6cdd1dull));
        // 1: quotients by a 4-D norm: numerators in [2^-60, 2^30], denominators in [1, 2^31]
        {
            const float a = rnd(s, -60, 29, true), d = rnd(s, 0, 30, false);
            const float y = fu_rcp_refined(d);
            b0 += __float_as_uint(fu_div_refined(a, d, y)) != __float_as_uint(a / d);
            b0 += __float_as_uint(fu_div_refined(1.0f, d, y)) != __float_as_uint(1.0f / d);
        }
        // 2: sdf / truncation: numerators 0 or in [2^-30, 2^20], denominators in [2^-12, 1]
        {
            const float a = ((s >> 50) & 15) == 0 ? 0.0f : rnd(s, -30, 19, true), d = rnd(s, -12, -1, false);
            b1 += __float_as_uint(fu_div_refined(a, d, fu_rcp_refined(d))) != __float_as_uint(a / d);
        }
        // 3: the projection: z in [2^-60, 2^30], numerators anywhere below 2^30 (tiny ones
        // included); the truncated q + c must agree
        {
            const float z = rnd(s, -60, 29, true), c = rnd(s, -8, 10, false);
            float a = rnd(s, -126, 29, true);
            if (((s >> 52) & 15) == 0) a = __uint_as_float((unsigned)(s >> 9) & 0x807fffffu);  // denormal
            const float q = fu_div_refined(a, z, fu_rcp_refined(z));
            b2 += f2i(q + c) != f2i(a / z + c);
            if (fabsf(a) >= 0x1p-100f) b2 += __float_as_uint(q) != __float_as_uint(a / z);
        }
        // 4: square roots in [2^-90, 2^120]
        {
            const float x = rnd(s, -90, 119, false);
            b3 += __float_as_uint(fu_sqrt_normal(x)) != __float_as_uint(sqrtf(x));
        }
        // 6: float and double -> int: any exponent (out of range, infinities and NaNs included)
        {
            s = mix(s + 0x9e3779b97f4a7c15ull);
            const float f = __uint_as_float((unsigned)s);
            const double d = __longlong_as_double((long long)mix(s));
            b5 += f2i(f) != conv_rule(f);
            b5 += d2i(d) != conv_rule(d);
            b5 += d2i((double)f) != conv_rule((double)f);
        }
        // 7: exp of a non-positive argument, uniform over [-110, 0] (subnormal and zero results
        // included) and over the exponents of small arguments
        {
            s = mix(s + 0x9e3779b97f4a7c15ull);
            const float u = -110.0f * (float)((s >> 40) & 0xffffff) * 0x1p-24f;
            const float t = -rnd(s, -30, 6, false);
            b6 += __float_as_uint(fu_expf_nonpos(u)) != __float_as_uint(fu_expf(u));
            b6 += __float_as_uint(fu_expf_nonpos(t)) != __float_as_uint(fu_expf(t));
        }
        // 5: the ICP projection's reciprocal: z in [2^-60, 2^30], either sign; the refined
        // float quotient 1 / z must equal the reference's double quotient rounded to float
        {
            const float z = rnd(s, -60, 29, true);
            const float r = fu_div_refined(1.0f, z, fu_rcp_refined(z));
            b4 += __float_as_uint(r) != __float_as_uint((float)(1.0 / (double)z));
            b4 += __float_as_uint(r) != __float_as_uint(1.0f / z);
        }
    }
    if (b6) atomicAdd(&bad[6], b6);
    if (b5) atomicAdd(&bad[5], b5);
    if (b0) atomicAdd(&bad[0], b0);
    if (b1) atomicAdd(&bad[1], b1);
    if (b2) atomicAdd(&bad[2], b2);
    if (b3) atomicAdd(&bad[3], b3);
    if (b4) atomicAdd(&bad[4], b4);
}

}  // namespace

extern "C" fu_status fu_selftest_arith(uint64_t seed, uint32_t n, uint32_t mismatches[7], fu_stream s) {
    if (!mismatches) return api_error(FU_ERR_ARG, "fu_selftest_arith: null output");
    unsigned* d = nullptr;
    if (hipMalloc(&d, 28) != hipSuccess) return api_error(FU_ERR_NOMEM, "fu_selftest_arith: allocation failed");
    hipStream_t hs = (hipStream_t)s;
    hipError_t e = hipMemsetAsync(d, 0, 28, hs);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_selftest, dim3(4096), dim3(256), 0, hs, seed, n, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(mismatches, d, 28, hipMemcpyDeviceToHost, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    (void)hipFree(d);
    return e == hipSuccess ? FU_OK : api_error(FU_ERR_HIP, "fu_selftest_arith: %s", hipGetErrorString(e));
}
