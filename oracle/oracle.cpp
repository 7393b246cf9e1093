// CPU oracle of the dense RGB-D pipeline - TEST INFRASTRUCTURE ONLY (see oracle.h).
// Scalar restatement of the reference's CUDA kernels; OpenMP is used only over independent
// pixels/blocks, and allocation is sequential in pixel order.
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <unordered_map>
#include <vector>

// OR_PLANT selects a variant with one known mistake, for the test that shows the oracle's
// pinning tests catch it (tests/test_oracle_planted.py); 0, the default, is the oracle itself.
//   1 ICP projection in double        2 ICP row with the current frame's normal
//   3 3-D dot for the integration weight   4 colour blend with old and new swapped
//   5 z-range fragments over nx (not ny) columns   6 Euler sine of X from R(2,1), not R(1,2)
#ifndef OR_PLANT
#define OR_PLANT 0
#endif
namespace {
constexpr int kPlant = OR_PLANT;
}  // namespace

namespace {

// ------------------------------------------------------------------ VectorMath.h semantics
struct V3 {
    float x, y, z;
};
struct V4 {
    float x, y, z, w;
};
struct I3 {
    int x, y, z;
};
struct M3 {
    V3 r[3];
};

inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }         // :191-193
inline float dot4(V4 a, V4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }  // :203-205
inline V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float norm(V3 a) { return sqrtf(dot(a, a)); }
inline V3 normalised(V3 a) { return a / norm(a); }
inline V4 normalised4(V4 a) {
    const float n = sqrtf(dot4(a, a));
    return V4{a.x / n, a.y / n, a.z / n, a.w / n};
}
inline V3 mul(const M3& m, V3 a) { return V3{dot(m.r[0], a), dot(m.r[1], a), dot(m.r[2], a)}; }  // :364-366
inline V3 xyz(V4 a) { return V3{a.x, a.y, a.z}; }
inline V3 i2f(I3 a, float s) { return V3{(float)a.x * s, (float)a.y * s, (float)a.z * s}; }  // :231-233
inline bool eq(I3 a, I3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
inline bool isn(float v) { return v != v; }
const float kNaN = __builtin_nanf("");

// D1: saturating conversion, NaN -> 0
inline int toi(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return -2147483647 - 1;
    return (int)v;
}
inline int toid(double v) {
    if (v != v) return 0;
    if (v >= 2147483648.0) return 2147483647;
    if (v <= -2147483648.0) return -2147483647 - 1;
    return (int)v;
}
inline I3 toi3(V3 a) { return I3{toi(a.x), toi(a.y), toi(a.z)}; }  // make_int3(float3), :36-38
inline unsigned char tou8(float v) { return (unsigned char)(unsigned)toi(v); }

M3 m3f(const float* r) {
    M3 m;
    m.r[0] = v3(r[0], r[1], r[2]);
    m.r[1] = v3(r[3], r[4], r[5]);
    m.r[2] = v3(r[6], r[7], r[8]);
    return m;
}

}  // namespace

// ------------------------------------------------------------------ exp (bilateral weights)
// Cody-Waite: x = k ln2 + r, |r| <= ln2/2; e^r by the degree-7 Taylor polynomial (Horner);
// result scaled by 2^k through the exponent field in two steps.
extern "C" float or_expf(float x) {
    if (x != x) return x;
    if (x > 88.5f) return INFINITY;
    if (x < -103.0f) return 0.0f;
    const float log2e = 1.44269502162933349609375f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.428606765330187045e-06f;
    const float kf = rintf(x * log2e);
    const int k = (int)kf;
    const float r = (x - kf * ln2_hi) - kf * ln2_lo;
    float p = 1.9841270e-4f;
    p = p * r + 1.3888889e-3f;
    p = p * r + 8.3333333e-3f;
    p = p * r + 4.1666668e-2f;
    p = p * r + 1.6666667e-1f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int k1 = k / 2, k2 = k - k1;
    uint32_t b1 = (uint32_t)(k1 + 127) << 23, b2 = (uint32_t)(k2 + 127) << 23;
    float s1, s2;
    memcpy(&s1, &b1, 4);
    memcpy(&s2, &b2, 4);
    return (p * s1) * s2;
}

// ------------------------------------------------------------------ Pyrdown.cu
// FilterDepthKernel (Pyrdown.cu:5-41)
extern "C" void or_filter_depth(const uint16_t* depth, float* raw, float* filt, int cols, int rows, float depth_scale,
                                float cutoff) {
    const float sinv = (float)(1.0 / (double)depth_scale);  // :50
    const float sig_space = 0.5f / (4 * 4), sig_range = 0.5f / (0.5f * 0.5f);  // :3-4
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            const float center = (float)depth[y * cols + x] * sinv;
            raw[y * cols + x] = center < cutoff ? center : kNaN;
            const int R = 2, D = R * 2 + 1;
            const int tx = std::min(x - D / 2 + D, cols - 1);
            const int ty = std::min(y - D / 2 + D, rows - 1);
            float sum1 = 0, sum2 = 0;
            for (int cy = std::max(y - D / 2, 0); cy < ty; ++cy)
                for (int cx = std::max(x - D / 2, 0); cx < tx; ++cx) {
                    const float val = (float)depth[cy * cols + cx] * sinv;
                    const float space2 = (float)((x - cx) * (x - cx) + (y - cy) * (y - cy));
                    const float color2 = (center - val) * (center - val);
                    const float weight = or_expf(-(space2 * sig_space + color2 * sig_range));
                    sum1 += val * weight;
                    sum2 += weight;
                }
            const float fin = sum1 / sum2;
            filt[y * cols + x] = fin < cutoff ? fin : kNaN;
        }
}

// PyrDownGaussKernel<float,float> (Pyrdown.cu:127-158)
extern "C" void or_pyrdown_gauss(const float* src, int cols, int rows, float* dst) {
    static const float g[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};
    const int dc = cols / 2, dr = rows / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < dr; ++y)
        for (int x = 0; x < dc; ++x) {
            const int D = 5;
            const int tx = std::min(2 * x - D / 2 + D, cols - 1);
            const int ty = std::min(2 * y - D / 2 + D, rows - 1);
            float sum = 0;
            int count = 0;
            for (int cy = std::max(0, 2 * y - D / 2); cy < ty; ++cy)
                for (int cx = std::max(0, 2 * x - D / 2); cx < tx; ++cx) {
                    const float v = src[cy * cols + cx];
                    if (!isn(v)) {
                        const float w = g[(ty - cy - 1) * 5 + (tx - cx - 1)];
                        sum += v * w;
                        count = (int)((float)count + w);
                    }
                }
            dst[y * dc + x] = sum / (float)count;
        }
}

// ComputeVMapKernel (Pyrdown.cu:56-91)
extern "C" void or_compute_vmap(const float* depth, float* vmap, int cols, int rows, float fx, float fy, float cx,
                                float cy, float cutoff) {
    const float invfx = (float)(1.0 / (double)fx), invfy = (float)(1.0 / (double)fy);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            const float z = depth[y * cols + x];
            V4 v;
            if (!isn(z) && (double)z > 0.1 && z < cutoff) {
                v.x = z * ((float)x - cx) * invfx;
                v.y = z * ((float)y - cy) * invfy;
                v.z = z;
                v.w = 1.0f;
            } else {
                v = V4{kNaN, kNaN, kNaN, kNaN};  // D7
            }
            memcpy(&vmap[(y * cols + x) * 4], &v, 16);
        }
}

// ComputeNMapKernel (Pyrdown.cu:93-125)
extern "C" void or_compute_nmap(const float* vmap, float* nmap, int cols, int rows) {
    auto at = [&](int x, int y) {
        V4 v;
        memcpy(&v, &vmap[(y * cols + x) * 4], 16);
        return v;
    };
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            V4 n{kNaN, kNaN, kNaN, kNaN};
            if (!(x == cols - 1 || y == rows - 1)) {
                const V4 c = at(x, y), r = at(x + 1, y), d = at(x, y + 1);
                if (!isn(c.x) && !isn(r.x) && !isn(d.x)) {
                    const V3 nn = normalised(cross(xyz(r) - xyz(c), xyz(d) - xyz(c)));
                    n = V4{nn.x, nn.y, nn.z, 1.0f};
                }
            }
            memcpy(&nmap[(y * cols + x) * 4], &n, 16);
        }
}

// ResizeMapKernel (Pyrdown.cu:251-291), grid bounded by the destination (D6)
extern "C" void or_resize_map(const float* vsrc, const float* nsrc, int cols, int rows, float* vdst, float* ndst) {
    const int dc = cols / 2, dr = rows / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < dr; ++y)
        for (int x = 0; x < dc; ++x) {
            V4 v[4], n[4];
            const int idx[4] = {(2 * y) * cols + 2 * x, (2 * y) * cols + 2 * x + 1, (2 * y + 1) * cols + 2 * x,
                                (2 * y + 1) * cols + 2 * x + 1};
            for (int k = 0; k < 4; ++k) {
                memcpy(&v[k], &vsrc[idx[k] * 4], 16);
                memcpy(&n[k], &nsrc[idx[k] * 4], 16);
            }
            V4 ov{kNaN, kNaN, kNaN, kNaN}, on{kNaN, kNaN, kNaN, kNaN};
            if (!(isn(v[0].x) || isn(v[1].x) || isn(v[2].x) || isn(v[3].x))) {
                ov = V4{((v[0].x + v[1].x) + v[2].x) + v[3].x, ((v[0].y + v[1].y) + v[2].y) + v[3].y,
                        ((v[0].z + v[1].z) + v[2].z) + v[3].z, ((v[0].w + v[1].w) + v[2].w) + v[3].w};
                ov = V4{ov.x / 4.0f, ov.y / 4.0f, ov.z / 4.0f, ov.w / 4.0f};
            }
            if (!(isn(n[0].x) || isn(n[1].x) || isn(n[2].x) || isn(n[3].x))) {
                V4 s{((n[0].x + n[1].x) + n[2].x) + n[3].x, ((n[0].y + n[1].y) + n[2].y) + n[3].y,
                     ((n[0].z + n[1].z) + n[2].z) + n[3].z, ((n[0].w + n[1].w) + n[2].w) + n[3].w};
                s = V4{s.x / 4.0f, s.y / 4.0f, s.z / 4.0f, s.w / 4.0f};
                on = normalised4(s);  // 4-D normalisation (VectorMath.h:340-342)
            }
            memcpy(&vdst[(y * dc + x) * 4], &ov, 16);
            memcpy(&ndst[(y * dc + x) * 4], &on, 16);
        }
}

// ------------------------------------------------------------------ Reduction.cu ICPReduce
// |sin(XYZ Euler angles)| of a motion's rotation (Tracking.cc:401: Eigen's eulerAngles(0, 1, 2)
// then .sin(), whose norm is the same for both Euler solutions): R = Rx(a) Ry(b) Rz(c) has
// sin b = R02, sin a = -R12 / |(R12, R22)|, sin c = -R01 / |(R01, R00)|
extern "C" double or_euler_sine_norm(const double p[16]) {
    const double r12 = kPlant == 6 ? p[9] : p[6];
    const double sa = -r12 / sqrt(r12 * r12 + p[10] * p[10]);
    const double sb = p[2];
    const double sc = -p[1] / sqrt(p[1] * p[1] + p[0] * p[0]);
    return sqrt(sa * sa + sb * sb + sc * sc);
}

namespace {
// searchPoint + getRow (Reduction.cu:227-286); returns found and the 7-vector row
inline bool icp_row(const float* vn, const float* vl, const float* nn, const float* nl, int cols, int rows, int x,
                    int y, const M3& Rc, V3 tc, const M3& Rl, const M3& Rli, V3 tl, float fx, float fy, float cx,
                    float cy, float row[7]) {
    const float angle_thresh = 0.6f, dist_thresh = 0.1f;  // :352-353
    V4 a;
    memcpy(&a, &vn[(y * cols + x) * 4], 16);
    const V3 vcurr_c = xyz(a);
    if (isn(vcurr_c.x) || (double)vcurr_c.z < 1e-3) return false;
    const V3 vcurr_g = mul(Rc, vcurr_c) + tc;
    const V3 vcurr_p = mul(Rli, vcurr_g - tl);
    // Reduction.cu:237-239: `float invz = 1.0 / z` (a double quotient stored in a float), the
    // products and the sum with cx in float, only the final + 0.5 in double, then truncation
    const float invz = (float)(1.0 / (double)vcurr_p.z);
    int u = toid((double)(vcurr_p.x * invz * fx + cx) + 0.5);
    int v = toid((double)(vcurr_p.y * invz * fy + cy) + 0.5);
    if constexpr (kPlant == 1) {
        const double dz = 1.0 / (double)vcurr_p.z;
        u = toid((double)vcurr_p.x * dz * (double)fx + (double)cx + 0.5);
        v = toid((double)vcurr_p.y * dz * (double)fy + (double)cy + 0.5);
    }
    if (u < 0 || v < 0 || u >= cols || v >= rows) return false;
    V4 b, c, d;
    memcpy(&b, &vl[(v * cols + u) * 4], 16);
    memcpy(&c, &nn[(y * cols + x) * 4], 16);
    memcpy(&d, &nl[(v * cols + u) * 4], 16);
    const V3 vlast_g = mul(Rl, xyz(b)) + tl;
    const V3 ncurr_c = xyz(c);
    const V3 ncurr_g = mul(Rc, ncurr_c);
    const V3 nlast_c = xyz(d);
    const V3 nlast_g = mul(Rl, nlast_c);
    const float dist = norm(vlast_g - vcurr_g);
    const float sine = norm(cross(ncurr_g, nlast_g));
    if (!(sine < angle_thresh && dist <= dist_thresh && !isn(ncurr_c.x) && !isn(nlast_c.x))) return false;
    const V3 nlr = mul(Rli, kPlant == 2 ? ncurr_g : nlast_g);
    const V3 vcr = mul(Rli, vcurr_g - tl);
    const V3 vlr = mul(Rli, vlast_g - tl);
    const V3 mn = -nlr;
    const V3 cr = cross(nlr, vlr);
    row[0] = mn.x;
    row[1] = mn.y;
    row[2] = mn.z;
    row[3] = cr.x;
    row[4] = cr.y;
    row[5] = cr.z;
    row[6] = dot(mn, vcr - vlr);
    return true;
}
}  // namespace

extern "C" void or_icp_sums(const float* next_vmap, const float* last_vmap, const float* next_nmap,
                            const float* last_nmap, int cols, int rows, const float Rcurr[9], const float tcurr[3],
                            const float Rlast[9], const float RlastInv[9], const float tlast[3], float fx, float fy,
                            float cx, float cy, float sums32[29], double sums64[29]) {
    const M3 Rc = m3f(Rcurr), Rl = m3f(Rlast), Rli = m3f(RlastInv);
    const V3 tc = v3(tcurr[0], tcurr[1], tcurr[2]), tl = v3(tlast[0], tlast[1], tlast[2]);
    std::vector<float> rowsbuf((size_t)cols * rows * 7);
    std::vector<unsigned char> found((size_t)cols * rows);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x)
            found[(size_t)y * cols + x] = icp_row(next_vmap, last_vmap, next_nmap, last_nmap, cols, rows, x, y, Rc, tc,
                                                  Rl, Rli, tl, fx, fy, cx, cy, &rowsbuf[((size_t)y * cols + x) * 7]);
    float s32[29] = {0};
    double s64[29] = {0};
    for (size_t i = 0; i < (size_t)cols * rows; ++i) {
        if (!found[i]) continue;
        const float* r = &rowsbuf[i * 7];
        int k = 0;
        for (int a = 0; a < 7; ++a)
            for (int b = a; b < 7; ++b) {
                const float p = r[a] * r[b];
                s32[k] += p;
                s64[k] += (double)p;
                ++k;
            }
        s32[28] += 1.0f;
        s64[28] += 1.0;
    }
    if (sums32) memcpy(sums32, s32, sizeof(s32));
    if (sums64) memcpy(sums64, s64, sizeof(s64));
}

extern "C" int or_solve6(const double A[36], const double b[6], double x[6]) {
    // Cholesky A = L L^T, then two triangular solves
    double L[36] = {0};
    for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
        if (!(s > 0)) return -1;
        L[j * 6 + j] = sqrt(s);
        for (int i = j + 1; i < 6; ++i) {
            double t = A[i * 6 + j];
            for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
            L[i * 6 + j] = t / L[j * 6 + j];
        }
    }
    double y[6];
    for (int i = 0; i < 6; ++i) {
        double t = b[i];
        for (int k = 0; k < i; ++k) t -= L[i * 6 + k] * y[k];
        y[i] = t / L[i * 6 + i];
    }
    for (int i = 5; i >= 0; --i) {
        double t = y[i];
        for (int k = i + 1; k < 6; ++k) t -= L[k * 6 + i] * x[k];
        x[i] = t / L[i * 6 + i];
    }
    return 0;
}

// exp of se(3): quaternion route for the rotation (unit quaternion (cos th/2, sin(th/2) w/th)),
// left Jacobian for the translation
extern "C" void or_se3_exp(const double xi[6], double T[16]) {
    const double w[3] = {xi[3], xi[4], xi[5]};
    const double th2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double th = sqrt(th2);
    double qw, f;
    if (th < 1e-8) {
        qw = 1.0 - th2 / 8.0;
        f = 0.5 - th2 / 48.0;
    } else {
        qw = cos(0.5 * th);
        f = sin(0.5 * th) / th;
    }
    const double qx = f * w[0], qy = f * w[1], qz = f * w[2];
    double R[9] = {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qw * qz),     2 * (qx * qz + qw * qy),
                   2 * (qx * qy + qw * qz),     1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qw * qx),
                   2 * (qx * qz - qw * qy),     2 * (qy * qz + qw * qx),     1 - 2 * (qx * qx + qy * qy)};
    double B, C;
    if (th < 1e-5) {
        B = 0.5 - th2 / 24.0;
        C = 1.0 / 6.0 - th2 / 120.0;
    } else {
        B = (1 - cos(th)) / th2;
        C = (th - sin(th)) / (th2 * th);
    }
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double V[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double w2 = 0;
            for (int k = 0; k < 3; ++k) w2 += W[i * 3 + k] * W[k * 3 + j];
            V[i * 3 + j] = (i == j ? 1.0 : 0.0) + B * W[i * 3 + j] + C * w2;
        }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T[i * 4 + j] = R[i * 3 + j];
        T[i * 4 + 3] = V[i * 3] * xi[0] + V[i * 3 + 1] * xi[1] + V[i * 3 + 2] * xi[2];
    }
    T[12] = T[13] = T[14] = 0;
    T[15] = 1;
}

namespace {
void mm4(const double* a, const double* b, double* o) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            t[i * 4 + j] = s;
        }
    memcpy(o, t, sizeof(t));
}
void inv4(const double* a, double* o) {
    double t[16] = {0};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = a[j * 4 + i];
        t[i * 4 + 3] = -(a[i] * a[3] + a[4 + i] * a[7] + a[8 + i] * a[11]);
    }
    t[15] = 1;
    memcpy(o, t, sizeof(t));
}
}  // namespace

extern "C" void or_pose_f(const double pose[16], float R[9], float Rinv[9], float t[3]) {
    // Frame::GpuRotation / GpuInvRotation / GpuTranslation (Core/Frame.cc:275-294)
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            R[i * 3 + j] = (float)pose[i * 4 + j];
            Rinv[i * 3 + j] = (float)pose[j * 4 + i];
        }
        t[i] = (float)pose[i * 4 + 3];
    }
}

// ------------------------------------------------------------------ photometric term
// ImageToIntensityKernel (Pyrdown.cu:183-194): (int)(0.2125 c1 + 0.7154 c0 + 0.0721 c2) in
// double, evaluated as written (two products summed, then the third; no fused multiply-add)
extern "C" void or_image_to_intensity(const uint8_t* rgb, uint8_t* img, int cols, int rows) {
    const size_t n = (size_t)cols * rows;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) {
        const uint8_t* c = rgb + 3 * i;
        const double v = (0.2125 * (double)c[1] + 0.7154 * (double)c[0]) + 0.0721 * (double)c[2];
        img[i] = (uint8_t)toid(v);
    }
}

// PyrDownGaussKernel<uchar,uchar> (Pyrdown.cu:133-158, launched by :171-181): the float
// kernel's window and mirrored weights, the mean truncated to a byte
extern "C" void or_pyrdown_u8(const uint8_t* src, int cols, int rows, uint8_t* dst) {
    static const float g[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6, 4, 16, 24, 16, 4, 1, 4, 6, 4, 1};
    const int dc = cols / 2, dr = rows / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < dr; ++y)
        for (int x = 0; x < dc; ++x) {
            const int tx = std::min(2 * x + 3, cols - 1);
            const int ty = std::min(2 * y + 3, rows - 1);
            float sum = 0;
            int count = 0;
            for (int cy = std::max(0, 2 * y - 2); cy < ty; ++cy)
                for (int cx = std::max(0, 2 * x - 2); cx < tx; ++cx) {
                    const float w = g[(ty - cy - 1) * 5 + (tx - cx - 1)];
                    sum += (float)src[cy * cols + cx] * w;
                    count = (int)((float)count + w);
                }
            dst[y * dc + x] = (uint8_t)toi(sum / (float)count);
        }
}

// ComputeDerivativeImageKernel (Pyrdown.cu:220-236): 3x3 Sobel over the window clipped to
// the image, the weight index counting down from 8 across the clipped window (so at the
// border the weights shift), float sums stored as shorts
extern "C" void or_derivative_image(const uint8_t* img, int cols, int rows, int16_t* dx, int16_t* dy) {
    static const float sx[9] = {1, 0, -1, 2, 0, -2, 1, 0, -1};
    static const float sy[9] = {1, 2, 1, 0, 0, 0, -1, -2, -1};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            float a = 0, b = 0;
            int k = 8;
            for (int j = std::max(y - 1, 0); j <= std::min(y + 1, rows - 1); ++j)
                for (int i = std::max(x - 1, 0); i <= std::min(x + 1, cols - 1); ++i) {
                    const float v = (float)img[j * cols + i];
                    a += v * sx[k];
                    b += v * sy[k];
                    --k;
                }
            dx[y * cols + x] = (int16_t)toi(a);
            dy[y * cols + x] = (int16_t)toi(b);
        }
}

namespace {
struct RgbPose {
    M3 Rc, Rci, Rl, Rli;
    V3 tc, tl;
};
RgbPose rgb_pose(const float* Rcurr, const float* RcurrInv, const float* Rlast, const float* RlastInv,
                 const float* tcurr, const float* tlast) {
    return RgbPose{m3f(Rcurr), m3f(RcurrInv), m3f(Rlast), m3f(RlastInv), v3(tcurr[0], tcurr[1], tcurr[2]),
                   v3(tlast[0], tlast[1], tlast[2])};
}

// RGBReduction::FindCorresp (Reduction.cu:414-467): the last-frame pixel j and the intensity
// difference of pixel (x, y), or false
inline bool rgb_corresp(const uint8_t* ni, const uint8_t* li, const float* nv, const float* lv, const int16_t* dx,
                        const int16_t* dy, int cols, int rows, int x, int y, const RgbPose& P, float fx, float fy,
                        float cx, float cy, int& j, int& diff) {
    if (!(x >= 5 && x < cols - 5 && y >= 5 && y < rows - 5)) return false;
    bool valid = true;  // the 4x4 window [y-2, y+2) x [x-2, x+2) is non-zero
    for (int u = std::max(y - 2, 0); u < std::min(y + 2, rows); ++u)
        for (int v = std::max(x - 2, 0); v < std::min(x + 2, cols); ++v) valid = valid && ni[u * cols + v] > 0;
    if (!valid) return false;
    const int i = y * cols + x;
    const int gx = dx[i], gy = dy[i];
    if (!(sqrtf((float)(gx * gx + gy * gy)) > 0)) return false;
    V4 vc;
    memcpy(&vc, &nv[(size_t)i * 4], 16);
    if (isn(vc.x)) return false;
    const V3 vg = mul(P.Rc, xyz(vc)) + P.tc;
    const V3 vl = mul(P.Rli, vg - P.tl);
    // __float2int_rd(fx * x / z + cx + 0.5): float quotient and sum, the 0.5 added in double,
    // the double rounded to float, then floor
    const int u = toi(floorf((float)((double)(fx * vl.x / vl.z + cx) + 0.5)));
    const int v = toi(floorf((float)((double)(fy * vl.y / vl.z + cy) + 0.5)));
    if (!(u >= 0 && v >= 0 && u < cols && v < rows)) return false;
    const int k = v * cols + u;
    V4 la;
    memcpy(&la, &lv[(size_t)k * 4], 16);
    const V4 d{la.x - vc.x, la.y - vc.y, la.z - vc.z, la.w - vc.w};
    if (!(!isn(la.x) && (double)sqrtf(dot4(d, d)) < 0.1 && li[k] != 0)) return false;
    j = k;
    diff = (int)ni[i] - (int)li[k];
    return true;
}

// RGBReduction::GetRow (Reduction.cu:485-520) of a valid residual
inline void rgb_row(const float* lv, const int16_t* dx, const int16_t* dy, int i, int j, int diff, float sigma,
                    const RgbPose& P, float fx, float fy, float row[7]) {
    float w = sigma + (float)abs(diff);
    w = (double)w < 1e-3 ? 1.0f : 1.0f / w;
    if ((double)sigma < 1e-6) w = 1.0f;
    V4 la;
    memcpy(&la, &lv[(size_t)j * 4], 16);
    const V3 point = mul(P.Rci, (mul(P.Rl, xyz(la)) + P.tl) - P.tc);
    const float gx = (float)((double)dx[i] / 9.0);
    const float gy = (float)((double)dy[i] / 9.0);
    const float invz = 1.0f / point.z;
    V3 left;
    left.x = gx * fx * invz;
    left.y = gy * fy * invz;
    left.z = -(left.x * point.x + left.y * point.y) * invz;
    const V3 a = w * left, b = w * cross(point, left);
    row[0] = a.x;
    row[1] = a.y;
    row[2] = a.z;
    row[3] = b.x;
    row[4] = b.y;
    row[5] = b.z;
    row[6] = -w * (float)diff;
}
}  // namespace

// RGBStep (Reduction.cu:537-627): residual pass with integer sums (count, sum of squared
// differences; 32-bit, wrapping), sigma = sqrt(sum / count), then the weighted rows.
// corresp = {count, sum diff^2}; sums32 in pixel order in FP32, sums64 the same products in FP64.
extern "C" void or_rgb_sums(const uint8_t* next_img, const uint8_t* last_img, const float* next_vmap,
                            const float* last_vmap, const int16_t* dx, const int16_t* dy, int cols, int rows,
                            const float Rcurr[9], const float RcurrInv[9], const float Rlast[9],
                            const float RlastInv[9], const float tcurr[3], const float tlast[3], float fx, float fy,
                            float cx, float cy, int32_t corresp[2], float* sigma_out, float sums32[29],
                            double sums64[29]) {
    const RgbPose P = rgb_pose(Rcurr, RcurrInv, Rlast, RlastInv, tcurr, tlast);
    const size_t n = (size_t)cols * rows;
    std::vector<int> jj(n, -1), dd(n, 0);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            int j, d;
            if (rgb_corresp(next_img, last_img, next_vmap, last_vmap, dx, dy, cols, rows, x, y, P, fx, fy, cx, cy, j, d)) {
                jj[(size_t)y * cols + x] = j;
                dd[(size_t)y * cols + x] = d;
            }
        }
    uint32_t cnt = 0, sq = 0;
    for (size_t i = 0; i < n; ++i)
        if (jj[i] >= 0) {
            cnt += 1u;
            sq += (uint32_t)(dd[i] * dd[i]);
        }
    const int c = (int)cnt, q = (int)sq;
    const float sigma = sqrtf((float)q / (float)(c == 0 ? 1 : c));
    if (corresp) {
        corresp[0] = c;
        corresp[1] = q;
    }
    if (sigma_out) *sigma_out = sigma;
    float s32[29] = {0};
    double s64[29] = {0};
    for (size_t i = 0; i < n; ++i) {
        if (jj[i] < 0) continue;
        float r[7];
        rgb_row(last_vmap, dx, dy, (int)i, jj[i], dd[i], sigma, P, fx, fy, r);
        int k = 0;
        for (int a = 0; a < 7; ++a)
            for (int b = a; b < 7; ++b) {
                const float p = r[a] * r[b];
                s32[k] += p;
                s64[k] += (double)p;
                ++k;
            }
        s32[28] += 1.0f;
        s64[28] += 1.0;
    }
    if (sums32) memcpy(sums32, s32, sizeof(s32));
    if (sums64) memcpy(sums64, s64, sizeof(s64));
}

namespace {
// CreateMatrix<6,7> (Reduction.cu:6-16)
void unpack29(const double* s, double* A, double* b) {
    int shift = 0;
    for (int i = 0; i < 6; ++i)
        for (int k = i; k < 7; ++k) {
            const double v = s[shift++];
            if (k == 6)
                b[i] = v;
            else
                A[k * 6 + i] = A[i * 6 + k] = v;
        }
}

struct RgbImages {
    const uint8_t* const* next_img;
    const uint8_t* const* last_img;
    const int16_t* const* dx;
    const int16_t* const* dy;
};

// Tracker::ComputeSE3 (Tracking.cc:296-409); rgb == nullptr is icpOnly = true
int compute_se3(const float* const next_vmap[3], const float* const next_nmap[3], const float* const last_vmap[3],
                const float* const last_nmap[3], const RgbImages* rgb, int cols, int rows, const float K[4],
                const int iters[3], const int min_count[3], float thresh, const double last_pose[16], double pose[16],
                float* err_out, int* count_out, int* rgb_count_out, double* trace = nullptr, int* n_trace = nullptr) {
    double next[16], delta[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (n_trace) *n_trace = 0;
    memcpy(next, last_pose, sizeof(next));
    float Rl[9], Rli[9], tl[3];
    or_pose_f(last_pose, Rl, Rli, tl);
    float icp_error = 0;
    int icp_count = 0;
    auto fail = [&]() {
        memcpy(pose, last_pose, 16 * sizeof(double));
        return 2;
    };
    for (int l = 2; l >= 0; --l) {
        const int c = cols >> l, r = rows >> l;
        const float sc = (float)(1 << l);  // Intrinsics::operator() (Utility/Intrinsics.h:13-16)
        const float fx = K[0] / sc, fy = K[1] / sc, cx = K[2] / sc, cy = K[3] / sc;
        for (int j = 0; j < iters[l]; ++j) {
            float Rc[9], Rci[9], tc[3];
            or_pose_f(next, Rc, Rci, tc);
            double A_rgb[36], b_rgb[6];
            if (rgb) {  // RGBStep and its failure test (Tracking.cc:318-346)
                double r64[29];
                or_rgb_sums(rgb->next_img[l], rgb->last_img[l], next_vmap[l], last_vmap[l], rgb->dx[l], rgb->dy[l], c,
                            r, Rc, Rci, Rl, Rli, tc, tl, fx, fy, cx, cy, nullptr, nullptr, nullptr, r64);
                unpack29(r64, A_rgb, b_rgb);
                const float rgb_error = (float)(sqrt(r64[27]) / r64[28]);
                const int rgb_count = (int)r64[28];
                if (rgb_count_out) *rgb_count_out = rgb_count;
                if (isn(rgb_error) || rgb_count < min_count[l]) return fail();
            }
            double s64[29];
            or_icp_sums(next_vmap[l], last_vmap[l], next_nmap[l], last_nmap[l], c, r, Rc, tc, Rl, Rli, tl, fx, fy, cx,
                        cy, nullptr, s64);
            double A[36], b[6];
            unpack29(s64, A, b);
            icp_error = (float)(sqrt(s64[27]) / s64[28]);
            icp_count = (int)s64[28];
            if (err_out) *err_out = icp_error;
            if (count_out) *count_out = icp_count;
            if (isn(icp_error) || icp_count < min_count[l]) return fail();
            if (rgb) {  // matA = matA_rgb * w + matA_icp with float w = 1e-4 (Tracking.cc:379-383)
                const double w = (double)1e-4f;
                for (int k = 0; k < 36; ++k) A[k] = A_rgb[k] * w + A[k];
                for (int k = 0; k < 6; ++k) b[k] = b_rgb[k] * w + b[k];
            }
            double x[6], dT[16], dinv[16];
            if (or_solve6(A, b, x)) return fail();
            if (trace) {  // the pose this iteration's sums were taken at, and its update x
                memcpy(trace, next, 16 * sizeof(double));
                memcpy(trace + 16, x, 6 * sizeof(double));
                trace += 22;
                ++*n_trace;
            }
            or_se3_exp(x, dT);
            mm4(dT, delta, delta);
            inv4(delta, dinv);
            mm4(last_pose, dinv, next);
        }
    }
    // acceptance (Tracking.cc:398-408): XYZ Euler sines of the motion against the initial guess
    double pinv[16], p[16];
    inv4(last_pose, pinv);
    mm4(pinv, next, p);
    const double an = or_euler_sine_norm(p);
    const double tn = sqrt(p[3] * p[3] + p[7] * p[7] + p[11] * p[11]);
    if (icp_error < thresh && an <= 0.1 && tn <= 0.1) {
        memcpy(pose, next, 16 * sizeof(double));
        return 1;
    }
    memcpy(pose, last_pose, 16 * sizeof(double));
    return 3;
}
}  // namespace

// Tracker::ComputeSE3(icpOnly = true) (Tracking.cc:296-409)
extern "C" int or_compute_se3(const float* const next_vmap[3], const float* const next_nmap[3],
                              const float* const last_vmap[3], const float* const last_nmap[3], int cols, int rows,
                              const float K[4], const int iters[3], const int min_count[3], float thresh,
                              const double last_pose[16], double pose[16], float* err_out, int* count_out) {
    return compute_se3(next_vmap, next_nmap, last_vmap, last_nmap, nullptr, cols, rows, K, iters, min_count, thresh,
                       last_pose, pose, err_out, count_out, nullptr);
}

// the same, recording every iteration: the current pose its sums were taken at (16 doubles)
// and its solution x = (upsilon, omega) (6 doubles); *n_trace iterations
extern "C" int or_compute_se3_trace(const float* const next_vmap[3], const float* const next_nmap[3],
                                    const float* const last_vmap[3], const float* const last_nmap[3], int cols, int rows,
                                    const float K[4], const int iters[3], const int min_count[3], float thresh,
                                    const double last_pose[16], double pose[16], double* trace, int* n_trace) {
    return compute_se3(next_vmap, next_nmap, last_vmap, last_nmap, nullptr, cols, rows, K, iters, min_count, thresh,
                       last_pose, pose, nullptr, nullptr, nullptr, trace, n_trace);
}

// Tracker::ComputeSE3(icpOnly = false): RGBStep blended into every ICP iteration
extern "C" int or_compute_se3_rgbd(const float* const next_vmap[3], const float* const next_nmap[3],
                                   const float* const last_vmap[3], const float* const last_nmap[3],
                                   const uint8_t* const next_img[3], const uint8_t* const last_img[3],
                                   const int16_t* const dx[3], const int16_t* const dy[3], int cols, int rows,
                                   const float K[4], const int iters[3], const int min_count[3], float thresh,
                                   const double last_pose[16], double pose[16], float* err_out, int* count_out,
                                   int* rgb_count_out) {
    const RgbImages im{next_img, last_img, dx, dy};
    return compute_se3(next_vmap, next_nmap, last_vmap, last_nmap, &im, cols, rows, K, iters, min_count, thresh,
                       last_pose, pose, err_out, count_out, rgb_count_out);
}

// ------------------------------------------------------------------ voxel-hash map
struct OrEntry {  // HashEntry (DeviceMap.h:20-59)
    I3 pos;
    int ptr;
    int offset;
};

struct or_map {
    uint32_t nb, nx, nblocks;
    float voxel, trunc, dmin, dmax, voxel_inv;
    uint32_t max_weight;
    int precision;  // 0: FP32 sdf, 1: FP64 sdf (FP64 twin of the update arithmetic)
    std::vector<OrEntry> entries;
    std::vector<int> heap;
    int heap_counter;
    int entry_ptr;
    std::vector<float> sdf;
    std::vector<double> sdf64;
    std::vector<uint16_t> weight;
    std::vector<unsigned char> rgb;
    std::vector<OrEntry> visible;
    uint32_t heap_full = 0, excess_full = 0;
};

extern "C" uint32_t or_hash(int32_t x, int32_t y, int32_t z, uint32_t num_buckets) {
    // DeviceMap::Hash (DeviceMap.cu:3-10), D2: int products wrap; `% NumBuckets` is unsigned
    const uint32_t h = ((uint32_t)x * 73856093u) ^ ((uint32_t)y * 19349669u) ^ ((uint32_t)z * 83492791u);
    return h % num_buckets;
}

namespace {
inline I3 voxel_to_block(I3 v) {  // DeviceMap.cu:150-161
    if (v.x < 0) v.x -= 7;
    if (v.y < 0) v.y -= 7;
    if (v.z < 0) v.z -= 7;
    return I3{v.x / 8, v.y / 8, v.z / 8};
}
inline int voxel_local_idx(I3 v) {  // DeviceMap.cu:167-201
    int lx = v.x % 8, ly = v.y % 8, lz = v.z % 8;
    if (lx < 0) lx += 8;
    if (ly < 0) ly += 8;
    if (lz < 0) lz += 8;
    return lz * 64 + ly * 8 + lx;
}

// CreateEntry (DeviceMap.cu:12-21)
OrEntry create_entry(or_map* m, I3 pos, int offset) {
    const int old = m->heap_counter--;
    if (old >= 0) {
        const int ptr = m->heap[old];
        if (ptr != -1) return OrEntry{pos, ptr * 512, offset};
    }
    m->heap_full++;
    return OrEntry{pos, -1, 0};
}

void init_block(or_map* m, int ptr) {
    for (int i = 0; i < 512; ++i) {
        m->sdf[ptr + i] = kNaN;
        if (m->precision) m->sdf64[ptr + i] = (double)kNaN;
        m->weight[ptr + i] = 0;
        m->rgb[(size_t)(ptr + i) * 3 + 0] = m->rgb[(size_t)(ptr + i) * 3 + 1] = m->rgb[(size_t)(ptr + i) * 3 + 2] = 0;
    }
}

// CreateBlock (DeviceMap.cu:23-62), sequential (D3)
void create_block(or_map* m, I3 bp) {
    int id = (int)or_hash(bp.x, bp.y, bp.z, m->nb);
    OrEntry* e = &m->entries[id];
    OrEntry* empty = nullptr;
    if (eq(e->pos, bp) && e->ptr != -1) return;
    if (e->ptr == -1 && !empty) empty = e;
    while (e->offset > 0) {
        id = (int)m->nb + e->offset - 1;
        e = &m->entries[id];
        if (eq(e->pos, bp) && e->ptr != -1) return;
        if (e->ptr == -1 && !empty) empty = e;
    }
    if (empty) {
        *empty = create_entry(m, bp, e->offset);  // the tail's offset, as DeviceMap.cu:47
        if (empty->ptr >= 0) init_block(m, empty->ptr);
    } else {
        const int offset = m->entry_ptr++;
        if (offset <= (int)m->nx) {
            OrEntry* ne = &m->entries[m->nb + offset - 1];
            *ne = create_entry(m, bp, 0);
            if (ne->ptr >= 0) init_block(m, ne->ptr);
            e->offset = offset;
        } else {
            m->excess_full++;
        }
    }
}

// FindEntry (DeviceMap.cu:122-135): the voxel offset of the block, -1 if absent
int find_ptr(const or_map* m, I3 bp) {
    int id = (int)or_hash(bp.x, bp.y, bp.z, m->nb);
    const OrEntry* e = &m->entries[id];
    if (e->ptr != -1 && eq(e->pos, bp)) return e->ptr;
    while (e->offset > 0) {
        id = (int)m->nb + e->offset - 1;
        e = &m->entries[id];
        if (eq(e->pos, bp) && e->ptr != -1) return e->ptr;
    }
    return -1;
}
}  // namespace

extern "C" or_map* or_map_create(uint32_t num_buckets, uint32_t num_excess, uint32_t num_blocks, float voxel_size,
                                 float trunc_dist, float depth_min, float depth_max, uint32_t max_weight,
                                 int precision) {
    or_map* m = new or_map();
    m->nb = num_buckets;
    m->nx = num_excess;
    m->nblocks = num_blocks;
    m->voxel = voxel_size;
    m->voxel_inv = (float)(1.0 / (double)voxel_size);
    m->trunc = trunc_dist;
    m->dmin = depth_min;
    m->dmax = depth_max;
    m->max_weight = max_weight;
    m->precision = precision;
    m->entries.resize((size_t)num_buckets + num_excess);
    m->heap.resize(num_blocks);
    m->sdf.resize((size_t)num_blocks * 512);
    if (precision) m->sdf64.resize((size_t)num_blocks * 512);
    m->weight.resize((size_t)num_blocks * 512);
    m->rgb.resize((size_t)num_blocks * 512 * 3);
    or_map_reset(m);
    return m;
}

extern "C" void or_map_destroy(or_map* m) { delete m; }

// ResetMap (FuseMap.cu:321-364); entries start zeroed as a fresh allocation would (the
// reference's release() keeps pos/offset)
extern "C" void or_map_reset(or_map* m) {
    for (auto& e : m->entries) e = OrEntry{I3{0, 0, 0}, -1, 0};
    for (uint32_t x = 0; x < m->nblocks; ++x) m->heap[x] = (int)(m->nblocks - x - 1);
    m->heap_counter = (int)m->nblocks - 1;
    m->entry_ptr = 1;
    m->visible.clear();
    m->heap_full = m->excess_full = 0;
}

extern "C" void or_map_counters(const or_map* m, uint32_t out[5]) {
    const int used = (int)m->nblocks - 1 - m->heap_counter;
    out[0] = (uint32_t)std::max(0, std::min(used, (int)m->nblocks));
    out[1] = (uint32_t)std::min<int>(m->entry_ptr - 1, (int)m->nx);
    out[2] = (uint32_t)m->visible.size();
    out[3] = m->heap_full;
    out[4] = m->excess_full;
}

// Fusion::CreateBlocks (FuseMap.cu:80-111); candidates per row in parallel, inserted in
// pixel order
extern "C" void or_create_blocks(or_map* m, const float* depth, int cols, int rows, const float Rr[9],
                                 const float tt[3], float fx, float fy, float cx, float cy) {
    const M3 R = m3f(Rr);
    const V3 t = v3(tt[0], tt[1], tt[2]);
    const float invfx = (float)(1.0 / (double)fx), invfy = (float)(1.0 / (double)fy);
    std::vector<std::vector<I3>> cand(rows);
#pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            const float z = depth[y * cols + x];
            if (isn(z) || z < m->dmin || z > m->dmax) continue;
            const float thresh = m->trunc / 2;
            const float z_near = fminf(m->dmax, z - thresh);
            const float z_far = fminf(m->dmax, z + thresh);
            if (z_near >= z_far) continue;
            auto unproject = [&](float zz) {  // Fusion::unproject (:28-34)
                const V3 p = v3(zz * ((float)x - cx) * invfx, zz * ((float)y - cy) * invfy, zz);
                return mul(R, p) + t;
            };
            V3 pn = unproject(z_near) * m->voxel_inv;
            const V3 pf = unproject(z_far) * m->voxel_inv;
            V3 dir = pf - pn;
            const float len = norm(dir);
            const int nsteps = toid(ceil(2.0 * (double)len));
            dir = dir / (float)(nsteps - 1);
            for (int i = 0; i < nsteps; ++i) {
                cand[y].push_back(voxel_to_block(toi3(pn)));
                pn = pn + dir;
            }
        }
    for (int y = 0; y < rows; ++y)
        for (const I3& b : cand[y]) create_block(m, b);
}

namespace {
// Fusion::CheckVertexVisibility / CheckBlockVisibility (FuseMap.cu:36-78)
bool vertex_visible(V3 p, const M3& Rinv, V3 t, int cols, int rows, float fx, float fy, float cx, float cy, float dmin,
                    float dmax) {
    p = mul(Rinv, p - t);
    if (p.z < 1e-3f) return false;
    const float px = fx * p.x / p.z + cx, py = fy * p.y / p.z + cy;
    return px >= 0 && py >= 0 && px < (float)cols && py < (float)rows && p.z >= dmin && p.z <= dmax;
}
bool block_visible(I3 pos, float scale, const M3& Rinv, V3 t, int cols, int rows, float fx, float fy, float cx,
                   float cy, float dmin, float dmax) {
    V3 c = i2f(pos, scale);
    auto vis = [&](V3 q) { return vertex_visible(q, Rinv, t, cols, rows, fx, fy, cx, cy, dmin, dmax); };
    if (vis(c)) return true;
    c.z += scale;
    if (vis(c)) return true;
    c.y += scale;
    if (vis(c)) return true;
    c.x += scale;
    if (vis(c)) return true;
    c.z -= scale;
    if (vis(c)) return true;
    c.y -= scale;
    if (vis(c)) return true;
    c.x -= scale;
    c.y += scale;
    if (vis(c)) return true;
    c.x += scale;
    c.y -= scale;
    c.z += scale;
    if (vis(c)) return true;
    return false;
}
}  // namespace

// CheckFullVisibility (FuseMap.cu:113-138): every entry with ptr != -1, any corner visible
extern "C" uint32_t or_check_visibility(or_map* m, const float Rinv_[9], const float tt[3], int cols, int rows,
                                        float fx, float fy, float cx, float cy, float depth_max, float depth_min) {
    const M3 Rinv = m3f(Rinv_);
    const V3 t = v3(tt[0], tt[1], tt[2]);
    const float scale = m->voxel * 8.0f;  // DeviceMap::blockWidth
    const size_t ne = m->entries.size();
    std::vector<unsigned char> vis(ne);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)ne; ++i) {
        const OrEntry& e = m->entries[i];
        vis[i] = e.ptr != -1 &&
                 block_visible(e.pos, scale, Rinv, t, cols, rows, fx, fy, cx, cy, depth_min, depth_max);
    }
    m->visible.clear();
    for (size_t i = 0; i < ne; ++i)
        if (vis[i]) m->visible.push_back(m->entries[i]);
    return (uint32_t)m->visible.size();
}

// Fusion::integrateColor (FuseMap.cu:140-191)
extern "C" void or_integrate(or_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols, int rows,
                             const float Rinv_[9], const float tt[3], float fx, float fy, float cx, float cy) {
    const M3 Rinv = m3f(Rinv_);
    const V3 t = v3(tt[0], tt[1], tt[2]);
    const long nv = (long)m->visible.size();
#pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < nv; ++b) {
        const OrEntry e = m->visible[b];
        if (e.ptr == -1) continue;
        const I3 base = I3{e.pos.x * 8, e.pos.y * 8, e.pos.z * 8};
        for (int lz = 0; lz < 8; ++lz)
            for (int ly = 0; ly < 8; ++ly)
                for (int lx = 0; lx < 8; ++lx) {
                    const int loc = lz * 64 + ly * 8 + lx;
                    V3 pos = i2f(I3{base.x + lx, base.y + ly, base.z + lz}, m->voxel);
                    pos = mul(Rinv, pos - t);
                    const int u = toi(fx * pos.x / pos.z + cx), v = toi(fy * pos.y / pos.z + cy);
                    if (u < 0 || v < 0 || u >= cols || v >= rows) continue;
                    const float dp = depth[v * cols + u];
                    if (isn(dp) || dp > m->dmax || dp < m->dmin) continue;
                    const float thresh = m->trunc;
                    float sdf = dp - pos.z;
                    if (!(sdf >= -thresh)) continue;
                    sdf = fminf(1.0f, sdf / thresh);
                    V4 nl;
                    memcpy(&nl, &nmap[(v * cols + u) * 4], 16);
                    if (isn(nl.x)) continue;
                    float w = dot4(nl, normalised4(V4{pos.x, pos.y, pos.z, 1.0f}));
                    if constexpr (kPlant == 3) w = dot(xyz(nl), normalised(pos));
                    const size_t vi = (size_t)e.ptr + loc;
                    const unsigned char* c = &rgb[(v * cols + u) * 3];
                    unsigned char* old = &m->rgb[vi * 3];
                    const unsigned wold = m->weight[vi];
                    if (wold == 0) {
                        m->sdf[vi] = sdf;
                        if (m->precision) m->sdf64[vi] = (double)sdf;
                        m->weight[vi] = 1;
                        old[0] = c[0];
                        old[1] = c[1];
                        old[2] = c[2];
                    } else {
                        for (int k = 0; k < 3; ++k) {
                            const float val = (float)c[k] / 255.f;
                            const float o = (float)old[k] / 255.f;
                            const float res = kPlant == 4 ? (w * 0.2f * o + (1 - w * 0.2f) * val) * 255.f
                                                          : (w * 0.2f * val + (1 - w * 0.2f) * o) * 255.f;
                            old[k] = tou8(res);
                        }
                        if (m->precision) {
                            const double s = (m->sdf64[vi] * (double)wold + (double)w * (double)sdf) / ((double)wold + (double)w);
                            m->sdf64[vi] = s;
                            m->sdf[vi] = (float)s;
                        } else {
                            m->sdf[vi] = (m->sdf[vi] * (float)wold + w * sdf) / ((float)wold + w);
                        }
                        m->weight[vi] = (uint16_t)std::min<unsigned>(m->max_weight, wold + 1);
                    }
                }
    }
}

// FuseMapColor (FuseMap.cu:251-319): the map's DepthMin/DepthMax for visibility
extern "C" uint32_t or_fuse_map_color(or_map* m, const float* depth, const uint8_t* rgb, const float* nmap, int cols,
                                      int rows, const float R[9], const float Rinv[9], const float t[3], float fx,
                                      float fy, float cx, float cy) {
    or_create_blocks(m, depth, cols, rows, R, t, fx, fy, cx, cy);
    const uint32_t n = or_check_visibility(m, Rinv, t, cols, rows, fx, fy, cx, cy, m->dmax, m->dmin);
    if (n) or_integrate(m, depth, rgb, nmap, cols, rows, Rinv, t, fx, fy, cx, cy);
    return n;
}

// ------------------------------------------------------------------ RenderScene.cu
// Projection::projectBlock + createRenderingBlockList + fillBlocks (RenderScene.cu:54-179,
// 199-276) with D5
extern "C" int or_create_rendering_blocks(or_map* m, float* zmin, float* zmax, int zcols, int zrows, float depth_max,
                                          float depth_min, const float Rinv_[9], const float tt[3],
                                          uint32_t n_visible, float fx, float fy, float cx, float cy,
                                          uint32_t* n_fragments) {
    if (n_fragments) *n_fragments = 0;
    if (n_visible == 0) return 0;
    const M3 Rinv = m3f(Rinv_);
    const V3 t = v3(tt[0], tt[1], tt[2]);
    for (int i = 0; i < zcols * zrows; ++i) {
        zmin[i] = 100.f;
        zmax[i] = 0.f;
    }
    uint32_t total = 0;
    const uint32_t nv = std::min<uint32_t>(n_visible, (uint32_t)m->visible.size());
    for (uint32_t k = 0; k < nv; ++k) {
        const OrEntry& e = m->visible[k];
        if (e.ptr == -1) continue;
        short ulx = (short)zcols, uly = (short)zrows, lrx = -1, lry = -1;
        float zlo = depth_max, zhi = depth_min;
        for (int corner = 0; corner < 8; ++corner) {
            I3 q = e.pos;
            q.x += (corner & 1) ? 1 : 0;
            q.y += (corner & 2) ? 1 : 0;
            q.z += (corner & 4) ? 1 : 0;
            V3 p = i2f(I3{(int)((unsigned)q.x * 8u), (int)((unsigned)q.y * 8u), (int)((unsigned)q.z * 8u)}, m->voxel);
            p = mul(Rinv, p - t);
            if ((double)p.z < 2e-1) continue;
            const float px = (fx * p.x / p.z + cx) / 8.0f, py = (fy * p.y / p.z + cy) / 8.0f;
            if ((float)ulx > floorf(px)) ulx = (short)toi(floorf(px));
            if ((float)lrx < ceilf(px)) lrx = (short)toi(ceilf(px));
            if ((float)uly > floorf(py)) uly = (short)toi(floorf(py));
            if ((float)lry < ceilf(py)) lry = (short)toi(ceilf(py));
            if (zlo > p.z) zlo = p.z;
            if (zhi < p.z) zhi = p.z;
        }
        if (ulx < 0) ulx = 0;
        if (uly < 0) uly = 0;
        if (lrx >= zcols) lrx = (short)(zcols - 1);
        if (lry >= zrows) lry = (short)(zrows - 1);
        if (ulx > lrx || uly > lry) continue;
        if (zlo < depth_min) zlo = depth_min;
        if (zhi < depth_min) continue;
        const float dx = (float)lrx - (float)ulx + 1, dy = (float)lry - (float)uly + 1;
        const int nx = (int)ceilf(dx / 16), ny = (int)ceilf(dy / 16);
        total += (uint32_t)(nx * ny);
        // fragments as createRenderingBlockList writes them (inner bound `x < ny`, lower
        // right without the fragment offset), then fillBlocks over each fragment's 16x16 tiles
        for (int fy_ = 0; fy_ < ny; ++fy_)
            for (int fx_ = 0; fx_ < (kPlant == 5 ? nx : ny); ++fx_) {
                int bulx = ulx + fx_ * 16, buly = uly + fy_ * 16;
                int blrx = ulx + 16, blry = uly + 16;
                if (blrx > lrx) blrx = lrx;
                if (blry > lry) blry = lry;
                for (int ty = 0; ty < 16; ++ty)
                    for (int tx = 0; tx < 16; ++tx) {
                        const int xp = bulx + tx, yp = buly + ty;
                        if (xp > blrx || xp >= zcols) continue;
                        if (yp > blry || yp >= zrows) continue;
                        zmin[yp * zcols + xp] = fminf(zmin[yp * zcols + xp], zlo);
                        zmax[yp * zcols + xp] = fmaxf(zmax[yp * zcols + xp], zhi);
                    }
            }
    }
    if (n_fragments) *n_fragments = total;
    return total > 0;
}

namespace {
struct Cache {
    I3 pos{0, 0, 0};
    int ptr = -1;  // D4: an empty cache never hits
};
// DeviceMap::FindVoxel(float3, cache, valid) (DeviceMap.cu:97-114): sdf and weight
inline float find_voxel(const or_map* m, V3 pt, Cache& c, bool& valid, unsigned& w) {
    const I3 p = toi3(pt);
    const I3 bp = voxel_to_block(p);
    int ptr;
    if (c.ptr >= 0 && eq(bp, c.pos)) {
        ptr = c.ptr;
    } else {
        ptr = find_ptr(m, bp);
        if (ptr < 0) {
            valid = false;
            w = 0;
            return kNaN;
        }
        c.pos = bp;
        c.ptr = ptr;
    }
    valid = true;
    const size_t vi = (size_t)ptr + voxel_local_idx(p);
    w = m->weight[vi];
    return m->sdf[vi];
}
inline float read_sdf(const or_map* m, V3 pt, Cache& c, bool& valid) {  // RenderScene.cu:290-295
    unsigned w;
    const float s = find_voxel(m, pt, c, valid, w);
    if (w == 0) valid = false;
    return s;
}
inline float sdf_at(const or_map* m, V3 pt, Cache& c, bool& valid) {
    unsigned w;
    return find_voxel(m, pt, c, valid, w);
}
// readSdfInterped (RenderScene.cu:297-319)
inline float read_interp(const or_map* m, V3 pt, Cache& c, bool& valid) {
    const V3 f = pt - v3(floorf(pt.x), floorf(pt.y), floorf(pt.z));
    float s0, s1, r0, r1, r2, r3;
    s0 = sdf_at(m, pt, c, valid);
    s1 = sdf_at(m, pt + v3(1, 0, 0), c, valid);
    r0 = (1.0f - f.x) * s0 + f.x * s1;
    s0 = sdf_at(m, pt + v3(0, 1, 0), c, valid);
    s1 = sdf_at(m, pt + v3(1, 1, 0), c, valid);
    r1 = (1.0f - f.x) * s0 + f.x * s1;
    r2 = (1.0f - f.y) * r0 + f.y * r1;
    s0 = sdf_at(m, pt + v3(0, 0, 1), c, valid);
    s1 = sdf_at(m, pt + v3(1, 0, 1), c, valid);
    r0 = (1.0f - f.x) * s0 + f.x * s1;
    s0 = sdf_at(m, pt + v3(0, 1, 1), c, valid);
    s1 = sdf_at(m, pt + v3(1, 1, 1), c, valid);
    r1 = (1.0f - f.x) * s0 + f.x * s1;
    r3 = (1.0f - f.y) * r0 + f.y * r1;
    return (1.0f - f.z) * r2 + f.z * r3;
}
}  // namespace

// Rendering::operator() (RenderScene.cu:354-440)
extern "C" void or_raycast(or_map* m, float* vmap, float* nmap, int cols, int rows, const float* zmin,
                           const float* zmax, int zcols, int zrows, const float R_[9], const float Rinv_[9],
                           const float tt[3], float invfx, float invfy, float cx, float cy) {
    (void)zrows;
    const M3 R = m3f(R_), Rinv = m3f(Rinv_);
    const V3 t = v3(tt[0], tt[1], tt[2]);
    const float step_scale = (float)(0.5 * (double)m->trunc * (double)m->voxel_inv);  // DeviceMap.h:151
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < rows; ++y)
        for (int x = 0; x < cols; ++x) {
            const V4 nan4{kNaN, kNaN, kNaN, kNaN};
            V4 vo = nan4, no = nan4;
            const float z0 = zmin[(y / 8) * zcols + x / 8], z1 = zmax[(y / 8) * zcols + x / 8];
            if (!((double)z1 < 1e-3 || (double)z0 < 1e-3 || isn(z0) || isn(z1))) {
                float sdf = 1.0f;
                V3 p3 = v3(z0 * ((float)x - cx) * invfx, z0 * ((float)y - cy) * invfy, z0);
                float dist_s = norm(p3) * m->voxel_inv;
                const V3 bs = (mul(R, p3) + t) * m->voxel_inv;
                p3 = v3(z1 * ((float)x - cx) * invfx, z1 * ((float)y - cy) * invfy, z1);
                const float dist_e = norm(p3) * m->voxel_inv;
                const V3 be = (mul(R, p3) + t) * m->voxel_inv;
                const V3 dir = normalised(be - bs);
                V3 result = bs;
                bool valid = false, found = false;
                float step;
                Cache c;
                while (dist_s < dist_e) {
                    sdf = read_sdf(m, result, c, valid);
                    if (!valid) {
                        step = 8.0f;
                    } else {
                        if (sdf <= 0.1f && sdf >= -0.5f) sdf = read_interp(m, result, c, valid);
                        if (sdf <= 0.0f) break;
                        step = !isn(sdf) ? fmaxf(sdf * step_scale, 1.0f) : 8.0f;
                    }
                    result = result + step * dir;
                    dist_s += step;
                }
                if (sdf <= 0.0f) {
                    step = sdf * step_scale;
                    result = result + step * dir;
                    sdf = read_interp(m, result, c, valid);
                    step = sdf * step_scale;
                    result = result + step * dir;
                    found = true;
                }
                if (found) {
                    // readNormal (RenderScene.cu:321-352)
                    const V3 off[6] = {v3(1, 0, 0), v3(-1, 0, 0), v3(0, 1, 0), v3(0, -1, 0), v3(0, 0, 1), v3(0, 0, -1)};
                    float s[6];
                    bool ok = true;
                    for (int k = 0; k < 6 && ok; ++k) {
                        bool v2;
                        s[k] = read_interp(m, result + off[k], c, v2);
                        if (isn(s[k]) || s[k] == 1.0f || !v2) ok = false;
                    }
                    if (ok) {
                        const V3 n = normalised(mul(Rinv, v3(s[0] - s[1], s[2] - s[3], s[4] - s[5])));
                        const V3 vt = mul(Rinv, result * m->voxel - t);
                        vo = V4{vt.x, vt.y, vt.z, 1.0f};
                        no = V4{n.x, n.y, n.z, 1.0f};
                    }
                }
            }
            memcpy(&vmap[(y * cols + x) * 4], &vo, 16);
            memcpy(&nmap[(y * cols + x) * 4], &no, 16);
        }
}

// ------------------------------------------------------------------ MeshScene.cu
namespace {
struct MeshCell {
    int cube;
    float sdf[8];
    V3 n[8];
    unsigned char c[8][3];
};
inline float vox_sdf(const or_map* m, I3 p, unsigned char* col) {  // FindVoxel(int3) (DeviceMap.cu:78-84)
    const int ptr = find_ptr(m, voxel_to_block(p));
    if (ptr < 0) {
        if (col) col[0] = col[1] = col[2] = 0;
        return kNaN;
    }
    const size_t vi = (size_t)ptr + voxel_local_idx(p);
    if (col) memcpy(col, &m->rgb[vi * 3], 3);
    return m->sdf[vi];
}
// readVertexAndColor + readNormal + cube index (MeshScene.cu:43-183); -1 when rejected
int build_cell(const or_map* m, I3 pos, MeshCell& cell) {
    static const int co[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int k = 0; k < 8; ++k) {
        cell.sdf[k] = vox_sdf(m, I3{pos.x + co[k][0], pos.y + co[k][1], pos.z + co[k][2]}, cell.c[k]);
        if (cell.sdf[k] == 1.0 || isn(cell.sdf[k])) return -1;
    }
    auto S = [&](int x, int y, int z) { return vox_sdf(m, I3{pos.x + x, pos.y + y, pos.z + z}, nullptr); };
    const float* s = cell.sdf;
    float v1, v2, v3_;
    v1 = S(-1, 0, 0), v2 = S(0, -1, 0), v3_ = S(0, 0, -1);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[0] = v3(s[1] - v1, s[3] - v2, s[4] - v3_);
    v1 = S(2, 0, 0), v2 = S(1, -1, 0), v3_ = S(1, 0, -1);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[1] = v3(v1 - s[0], s[2] - v2, s[5] - v3_);
    v1 = S(2, 1, 0), v2 = S(1, 2, 0), v3_ = S(1, 1, -1);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[2] = v3(v1 - s[3], v2 - s[1], s[6] - v3_);
    v1 = S(-1, 1, 0), v2 = S(0, 2, 0), v3_ = S(0, 1, -1);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[3] = v3(s[2] - v1, v2 - s[0], s[7] - v3_);
    v1 = S(-1, 0, 1), v2 = S(0, -1, 1), v3_ = S(0, 0, 2);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[4] = v3(s[5] - v1, s[7] - v2, v3_ - s[0]);
    v1 = S(2, 0, 1), v2 = S(1, -1, 1), v3_ = S(1, 0, 2);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[5] = v3(v1 - s[4], s[6] - v2, v3_ - s[1]);
    v1 = S(2, 1, 1), v2 = S(1, 2, 1), v3_ = S(1, 1, 2);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[6] = v3(v1 - s[7], v2 - s[5], v3_ - s[2]);
    v1 = S(-1, 1, 1), v2 = S(0, 2, 1), v3_ = S(0, 1, 2);
    if (isn(v1) || isn(v2) || isn(v3_)) return -1;
    cell.n[7] = v3(s[6] - v1, v2 - s[4], v3_ - s[3]);
    int cube = 0;
    for (int k = 0; k < 8; ++k)
        if (s[k] < 0) cube |= 1 << k;
    cell.cube = cube;
    return cube;
}
inline float mc_interp(float v1, float v2) {  // MeshScene.cu:142-150
    if ((double)fabsf(0 - v1) < 1e-6) return 0;
    if ((double)fabsf(0 - v2) < 1e-6) return 1;
    if ((double)fabsf(v1 - v2) < 1e-6) return 0;
    return (0 - v1) / (v2 - v1);
}
}  // namespace

extern int g_or_edge[256];
extern int g_or_tri[256][16];

// MeshEngine::MarchingCube (MeshScene.cu:261-295) over all allocated blocks, voxel order
// z,y,x within a block, blocks in pool order
extern "C" uint32_t or_mesh(or_map* m, float* vertex, float* normal, uint8_t* color, int32_t* meta, uint32_t max_tri) {
    std::vector<std::pair<int, I3>> blocks;  // (ptr, pos)
    for (const OrEntry& e : m->entries)
        if (e.ptr >= 0) blocks.push_back({e.ptr, e.pos});
    std::sort(blocks.begin(), blocks.end(), [](const std::pair<int, I3>& a, const std::pair<int, I3>& b) { return a.first < b.first; });
    const long nb = (long)blocks.size();
    std::vector<std::vector<std::pair<I3, MeshCell>>> cells(nb);
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < nb; ++b) {
        const I3 base{blocks[b].second.x * 8, blocks[b].second.y * 8, blocks[b].second.z * 8};
        for (int lz = 0; lz < 8; ++lz)
            for (int ly = 0; ly < 8; ++ly)
                for (int lx = 0; lx < 8; ++lx) {
                    MeshCell c;
                    const I3 p{base.x + lx, base.y + ly, base.z + lz};
                    const int cube = build_cell(m, p, c);
                    if (cube <= 0 || g_or_edge[cube] == 0) continue;
                    cells[b].push_back({p, c});
                }
    }
    uint32_t tri = 0;
    static const int ec[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    // vertex offset per edge as buildVertexList writes it (MeshScene.cu:185-256); -1 marks
    // the interpolated coordinate running from the first corner (val), -2 the reverse (1 - val)
    static const int eo[12][3] = {{-1, 0, 0}, {1, -1, 0}, {-2, 1, 0}, {0, -2, 0}, {-1, 0, 1}, {1, -1, 1},
                                  {-2, 1, 1}, {0, -2, 1}, {0, 0, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}};
    for (long b = 0; b < nb; ++b)
        for (auto& pc : cells[b]) {
            const I3 p = pc.first;
            const MeshCell& c = pc.second;
            for (int k = 0; k < 16 && g_or_tri[c.cube][k] >= 0; k += 3) {
                const uint32_t id = tri++;
                if (id >= max_tri) continue;
                for (int j = 0; j < 3; ++j) {
                    const int e = g_or_tri[c.cube][k + j];
                    const int a = ec[e][0], bb = ec[e][1];
                    const float val = mc_interp(c.sdf[a], c.sdf[bb]);
                    float o[3];
                    for (int q = 0; q < 3; ++q)
                        o[q] = eo[e][q] == -1 ? val : (eo[e][q] == -2 ? 1 - val : (float)eo[e][q]);
                    const V3 vp = v3((float)p.x + o[0], (float)p.y + o[1], (float)p.z + o[2]) * m->voxel;
                    const V3 nv = normalised(c.n[a] + val * (c.n[bb] - c.n[a]));
                    const int iv = toi(val);
                    unsigned char col[3];
                    for (int q = 0; q < 3; ++q) {
                        const unsigned char d = (unsigned char)(c.c[bb][q] - c.c[a][q]);
                        const unsigned char pr = (unsigned char)(d * iv);
                        col[q] = (unsigned char)(c.c[a][q] + pr);
                    }
                    const size_t vi = (size_t)id * 3 + j;
                    if (vertex) {
                        vertex[vi * 3] = vp.x;
                        vertex[vi * 3 + 1] = vp.y;
                        vertex[vi * 3 + 2] = vp.z;
                    }
                    if (normal) {
                        normal[vi * 3] = nv.x;
                        normal[vi * 3 + 1] = nv.y;
                        normal[vi * 3 + 2] = nv.z;
                    }
                    if (color) memcpy(&color[vi * 3], col, 3);
                }
                if (meta) {
                    meta[(size_t)id * 4] = p.x;
                    meta[(size_t)id * 4 + 1] = p.y;
                    meta[(size_t)id * 4 + 2] = p.z;
                    meta[(size_t)id * 4 + 3] = c.cube * 16 + k / 3;
                }
            }
        }
    return tri;
}

// test hook: allocate the blocks within `trunc` of a sphere (through CreateBlock) and store its
// exact truncated signed distance, weight 1, colour (100,150,200)
extern "C" void or_map_set_sphere(or_map* m, float cx, float cy, float cz, float r) {
    const double v = m->voxel, T = m->trunc;
    const int lo[3] = {(int)floor((cx - r - T) / v) - 1, (int)floor((cy - r - T) / v) - 1, (int)floor((cz - r - T) / v) - 1};
    const int hi[3] = {(int)ceil((cx + r + T) / v) + 1, (int)ceil((cy + r + T) / v) + 1, (int)ceil((cz + r + T) / v) + 1};
    for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const double d = sqrt((x * v - cx) * (x * v - cx) + (y * v - cy) * (y * v - cy) + (z * v - cz) * (z * v - cz)) - r;
                if (fabs(d) > T) continue;
                const I3 p{x, y, z};
                create_block(m, voxel_to_block(p));
                const int ptr = find_ptr(m, voxel_to_block(p));
                if (ptr < 0) continue;
                const size_t vi = (size_t)ptr + voxel_local_idx(p);
                m->sdf[vi] = (float)(d / T);
                if (m->precision) m->sdf64[vi] = d / T;
                m->weight[vi] = 1;
                m->rgb[vi * 3] = 100;
                m->rgb[vi * 3 + 1] = 150;
                m->rgb[vi * 3 + 2] = 200;
            }
}

extern "C" uint32_t or_map_export_blocks(const or_map* m, int32_t* xyz, uint32_t cap) {
    uint32_t k = 0;
    for (const OrEntry& e : m->entries)
        if (e.ptr >= 0) {
            if (xyz && k < cap) {
                xyz[3 * k] = e.pos.x;
                xyz[3 * k + 1] = e.pos.y;
                xyz[3 * k + 2] = e.pos.z;
            }
            ++k;
        }
    return k;
}

extern "C" uint32_t or_map_export_visible(const or_map* m, int32_t* xyz, uint32_t cap) {
    uint32_t k = 0;
    for (const OrEntry& e : m->visible) {
        if (xyz && k < cap) {
            xyz[3 * k] = e.pos.x;
            xyz[3 * k + 1] = e.pos.y;
            xyz[3 * k + 2] = e.pos.z;
        }
        ++k;
    }
    return k;
}

extern "C" void or_map_read_blocks(const or_map* m, const int32_t* xyz, uint32_t n, float* sdf, uint16_t* weight,
                                   uint8_t* rgb) {
    for (uint32_t b = 0; b < n; ++b) {
        const int ptr = find_ptr(m, I3{xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]});
        for (int i = 0; i < 512; ++i) {
            const size_t o = (size_t)b * 512 + i;
            if (ptr < 0) {
                if (sdf) sdf[o] = kNaN;
                if (weight) weight[o] = 0;
                if (rgb) rgb[o * 3] = rgb[o * 3 + 1] = rgb[o * 3 + 2] = 0;
            } else {
                if (sdf) sdf[o] = m->sdf[(size_t)ptr + i];
                if (weight) weight[o] = m->weight[(size_t)ptr + i];
                if (rgb) memcpy(&rgb[o * 3], &m->rgb[((size_t)ptr + i) * 3], 3);
            }
        }
    }
}

extern "C" void or_map_read_blocks_f64(const or_map* m, const int32_t* xyz, uint32_t n, double* sdf) {
    for (uint32_t b = 0; b < n; ++b) {
        const int ptr = find_ptr(m, I3{xyz[3 * b], xyz[3 * b + 1], xyz[3 * b + 2]});
        for (int i = 0; i < 512; ++i)
            sdf[(size_t)b * 512 + i] = ptr < 0 ? (double)kNaN
                                               : (m->precision ? m->sdf64[(size_t)ptr + i] : (double)m->sdf[(size_t)ptr + i]);
    }
}

// ------------------------------------------------------------------ System::GrabImage
struct or_system {
    int cols, rows;
    float K[4], depth_scale, cutoff, thresh;
    int iters[3], minc[3];
    or_map* map;
    std::vector<float> raw, depth[3], vmap[3], nmap[3], mvmap[3], mnmap[3], zmin, zmax;
    int zcols, zrows;
    double last_pose[16];
    int frame = 0;
    // photometric term (ComputeSE3 with icpOnly = false): this frame's intensity pyramid and
    // derivatives, and the pyramid of the last tracked frame (LastFrame->image)
    int use_rgb = 0;
    std::vector<uint8_t> img[3], last_img[3];
    std::vector<int16_t> dx[3], dy[3];
    int last_rgb_count = 0;
    float last_err = 0;  // final ICP error and correspondence count of the last tracked frame
    int last_cnt = 0;
    // Tracker::state / lastState / noMissedFrames (Tracking.cc:21) and
    // System::num_frames_after_reloc (System.cc:72); a relocalisation pose given before a lost
    // frame places that frame (or_system_relocalise), or or_system_validate_poses relocalises
    // the lost frame itself (Relocalise with graph matching, Tracking.cc:411-682)
    int state = 1, last_state = 1, missed = 0, after_reloc = 10;
    int reloc = 0;
    double reloc_pose[16];
    int mapping_off = 0;  // Tracker::mappingDisabled (Tracking.h:45)
};

extern "C" or_system* or_system_create(int cols, int rows, float fx, float fy, float cx, float cy, float depth_scale,
                                       float depth_cutoff, const int iters[3], const int min_count[3],
                                       float icp_thresh, uint32_t num_buckets, uint32_t num_excess,
                                       uint32_t num_blocks, float voxel_size, float trunc_dist, uint32_t max_weight,
                                       int precision) {
    or_system* s = new or_system();
    s->cols = cols;
    s->rows = rows;
    s->K[0] = fx;
    s->K[1] = fy;
    s->K[2] = cx;
    s->K[3] = cy;
    s->depth_scale = depth_scale;
    s->cutoff = depth_cutoff;
    s->thresh = icp_thresh;
    for (int i = 0; i < 3; ++i) {
        s->iters[i] = iters[i];
        s->minc[i] = min_count[i];
    }
    s->map = or_map_create(num_buckets, num_excess, num_blocks, voxel_size, trunc_dist, 0.1f, 3.0f, max_weight, precision);
    s->raw.resize((size_t)cols * rows);
    for (int l = 0; l < 3; ++l) {
        const size_t n = (size_t)(cols >> l) * (rows >> l);
        s->depth[l].resize(n);
        s->vmap[l].resize(n * 4);
        s->nmap[l].resize(n * 4);
        s->mvmap[l].resize(n * 4);
        s->mnmap[l].resize(n * 4);
        s->img[l].resize(n);
        s->last_img[l].resize(n);
        s->dx[l].resize(n);
        s->dy[l].resize(n);
    }
    s->zcols = (cols + 7) / 8;
    s->zrows = (rows + 7) / 8;
    s->zmin.resize((size_t)s->zcols * s->zrows);
    s->zmax.resize((size_t)s->zcols * s->zrows);
    for (int i = 0; i < 16; ++i) s->last_pose[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return s;
}

extern "C" void or_system_destroy(or_system* s) {
    if (!s) return;
    or_map_destroy(s->map);
    delete s;
}
extern "C" or_map* or_system_map(or_system* s) { return s->map; }
extern "C" const float* or_system_buffer(or_system* s, int which, int level) {
    switch (which) {
        case 0: return s->raw.data();
        case 1: return s->vmap[level].data();
        case 2: return s->nmap[level].data();
        case 3: return s->mvmap[level].data();
        case 4: return s->mnmap[level].data();
        case 5: return s->depth[level].data();
        default: return nullptr;
    }
}

extern "C" void or_system_set_rgb(or_system* s, int use_rgb) { s->use_rgb = use_rgb; }
extern "C" const void* or_system_image(or_system* s, int which, int level) {
    switch (which) {
        case 0: return s->img[level].data();
        case 1: return s->last_img[level].data();
        case 2: return s->dx[level].data();
        case 3: return s->dy[level].data();
        default: return nullptr;
    }
}
extern "C" int or_system_rgb_count(const or_system* s) { return s->last_rgb_count; }
extern "C" float or_system_icp_error(const or_system* s, int* count) {
    if (count) *count = s->last_cnt;
    return s->last_err;
}

extern "C" int or_system_grab(or_system* s, const uint16_t* depth, const uint8_t* rgb, double pose_out[16]) {
    const int c0 = s->cols, r0 = s->rows;
    const bool first = s->frame == 0;
    // LastFrame->ResizeImages (Core/Frame.cc:80-84)
    if (!first)
        for (int l = 1; l < 3; ++l)
            or_resize_map(s->mvmap[l - 1].data(), s->mnmap[l - 1].data(), c0 >> (l - 1), r0 >> (l - 1),
                          s->mvmap[l].data(), s->mnmap[l].data());
    // NextFrame->FillImages (Frame.cc:59-78)
    or_filter_depth(depth, s->raw.data(), s->depth[0].data(), c0, r0, s->depth_scale, s->cutoff);
    for (int l = 1; l < 3; ++l) or_pyrdown_gauss(s->depth[l - 1].data(), c0 >> (l - 1), r0 >> (l - 1), s->depth[l].data());
    for (int l = 0; l < 3; ++l) {
        const float sc = (float)(1 << l);
        or_compute_vmap(s->depth[l].data(), s->vmap[l].data(), c0 >> l, r0 >> l, s->K[0] / sc, s->K[1] / sc,
                        s->K[2] / sc, s->K[3] / sc, s->cutoff);
        or_compute_nmap(s->vmap[l].data(), s->nmap[l].data(), c0 >> l, r0 >> l);
    }
    if (s->use_rgb) {  // ImageToIntensity, PyrDownGauss<uchar>, ComputeDerivativeImage
        or_image_to_intensity(rgb, s->img[0].data(), c0, r0);
        for (int l = 1; l < 3; ++l)
            or_pyrdown_u8(s->img[l - 1].data(), c0 >> (l - 1), r0 >> (l - 1), s->img[l].data());
        for (int l = 0; l < 3; ++l) or_derivative_image(s->img[l].data(), c0 >> l, r0 >> l, s->dx[l].data(), s->dy[l].data());
    }
    s->frame++;
    double pose[16];
    int status = 4;
    bool swap = false;  // GrabFrame's return value: the frame becomes LastFrame
    // Tracker::GrabFrame (Tracking.cc:59-112)
    std::swap(s->state, s->last_state);
    if (s->state == 1) {
        memcpy(pose, s->last_pose, sizeof(pose));  // InitTracking: identity
        s->last_state = 0;
        swap = true;
    } else if (s->state == -1) {
        if (s->reloc) {
            s->reloc = 0;
            memcpy(pose, s->reloc_pose, sizeof(pose));
            status = 6;
            s->last_state = 0;
            swap = true;
        } else {
            memcpy(pose, s->last_pose, sizeof(pose));
            status = 5;
            s->last_state = -1;
        }
    } else {
        const float* nv[3] = {s->vmap[0].data(), s->vmap[1].data(), s->vmap[2].data()};
        const float* nn[3] = {s->nmap[0].data(), s->nmap[1].data(), s->nmap[2].data()};
        const float* lv[3] = {s->mvmap[0].data(), s->mvmap[1].data(), s->mvmap[2].data()};
        const float* ln[3] = {s->mnmap[0].data(), s->mnmap[1].data(), s->mnmap[2].data()};
        float err = 0;
        int cnt = 0;
        if (s->use_rgb) {
            const uint8_t* ni[3] = {s->img[0].data(), s->img[1].data(), s->img[2].data()};
            const uint8_t* li[3] = {s->last_img[0].data(), s->last_img[1].data(), s->last_img[2].data()};
            const int16_t* gx[3] = {s->dx[0].data(), s->dx[1].data(), s->dx[2].data()};
            const int16_t* gy[3] = {s->dy[0].data(), s->dy[1].data(), s->dy[2].data()};
            status = or_compute_se3_rgbd(nv, nn, lv, ln, ni, li, gx, gy, c0, r0, s->K, s->iters, s->minc, s->thresh,
                                         s->last_pose, pose, &err, &cnt, &s->last_rgb_count);
        } else {
            status = or_compute_se3(nv, nn, lv, ln, c0, r0, s->K, s->iters, s->minc, s->thresh, s->last_pose, pose,
                                    &err, &cnt);
        }
        s->last_err = err;
        s->last_cnt = cnt;
        if (status == 1) {
            s->missed = 0;
            swap = s->last_state != -1;  // the frame after a relocalisation is tracked, not swapped
            s->last_state = 0;
        } else {
            s->last_state = 0;
            if (++s->missed > 9) s->last_state = -1;
        }
    }
    // System::GrabImage (System.cc:92-111)
    if (s->state == 0)
        s->after_reloc++;
    else if (s->state == -1)
        s->after_reloc = 0;
    if (pose_out) memcpy(pose_out, pose, sizeof(pose));
    if (!swap) return status;
    const bool fuse = !s->mapping_off && s->state != -1 && s->after_reloc >= 10;  // System.cc:110-111
    memcpy(s->last_pose, pose, sizeof(pose));
    // Tracker::SwapFrame (Tracking.cc:245-247) after a swapped frame: its images
    // become LastFrame's
    if (s->use_rgb)
        for (int l = 0; l < 3; ++l) s->last_img[l] = s->img[l];
    float R[9], Rinv[9], t[3];
    or_pose_f(pose, R, Rinv, t);
    // Mapping::FuseColor + Mapping::RayTrace (Mapping.cc:76-109); without fusion the visible
    // blocks come from UpdateVisibility (the reference's count is left unset after a
    // relocalisation when it neither fuses nor updates the visibility)
    const uint32_t nvis =
        fuse ? or_fuse_map_color(s->map, s->raw.data(), rgb, s->nmap[0].data(), c0, r0, R, Rinv, t, s->K[0], s->K[1],
                                 s->K[2], s->K[3])
             : or_check_visibility(s->map, Rinv, t, c0, r0, s->K[0], s->K[1], s->K[2], s->K[3], s->map->dmax, s->map->dmin);
    uint32_t nfrag = 0;
    const int rendered =
        or_create_rendering_blocks(s->map, s->zmin.data(), s->zmax.data(), s->zcols, s->zrows, s->map->dmax,
                                   s->map->dmin, Rinv, t, nvis, s->K[0], s->K[1], s->K[2], s->K[3], &nfrag);
    if (rendered) {
        or_raycast(s->map, s->mvmap[0].data(), s->mnmap[0].data(), c0, r0, s->zmin.data(), s->zmax.data(), s->zcols,
                   s->zrows, R, Rinv, t, (float)(1.0 / (double)s->K[0]), (float)(1.0 / (double)s->K[1]), s->K[2],
                   s->K[3]);
    } else {
        s->mvmap[0] = s->vmap[0];
        s->mnmap[0] = s->nmap[0];
    }
    return status;
}

// Tracker::mappingDisabled: swapped frames are rendered (UpdateVisibility + RayTrace,
// System.cc:114-120) and never fused
extern "C" void or_system_set_mapping(or_system* s, int enabled) { s->mapping_off = enabled ? 0 : 1; }

// a relocalisation pose for a lost system's next frame; -1 unless the system is lost
extern "C" int or_system_relocalise(or_system* s, const double pose[16]) {
    if (s->last_state != -1) return -1;
    memcpy(s->reloc_pose, pose, sizeof(s->reloc_pose));
    s->reloc = 1;
    return 0;
}

extern "C" void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

// ------------------------------------------------------------------ keypoint map
// KeyMap (Mapping/DeviceMap.h:163-185, DeviceMap.cu:203-296) and its kernels
// ResetKeyPoints / CollectKeyPoints / InsertKeyPoints (Mapping/FuseMap.cu:366-472), with
// inserts applied one key at a time in index order (the reference's try-lock may drop a
// key that meets another in its bucket in the same launch).
struct OrSurf {  // SURF (DeviceMap.h:99-108): bool valid, float3 pos, float4 normal, float[64]
    uint8_t valid;
    uint8_t pad[3];
    float pos[3];
    float normal[4];
    float desc[64];
};
static_assert(sizeof(OrSurf) == 288, "SURF layout");

struct or_keymap {
    float grid;
    int max_keys, nbuckets;
    std::vector<OrSurf> keys;
    std::vector<int> mutex;
};

extern "C" or_keymap* or_keymap_create(float grid, int max_keys, int nbuckets) {
    or_keymap* k = new or_keymap();
    k->grid = grid;
    k->max_keys = max_keys;
    k->nbuckets = nbuckets;
    k->keys.assign((size_t)max_keys * nbuckets, OrSurf{});
    k->mutex.assign((size_t)max_keys * nbuckets, 0);
    return k;
}
extern "C" void or_keymap_destroy(or_keymap* k) { delete k; }

// ResetKeys (DeviceMap.cu:289-296): every mutex -1, every key invalid (the rest is kept)
extern "C" void or_keymap_reset(or_keymap* k) {
    for (int& m : k->mutex) m = -1;
    for (OrSurf& s : k->keys) s.valid = 0;
}

namespace {
// KeyMap::Hash (DeviceMap.cu:206-214): MaxKeys is an int, so the xor of the wrapped products
// is reduced with the signed remainder and made non-negative
int key_hash(I3 p, int max_keys) {
    const int h = (int)(((uint32_t)p.x * 73856093u) ^ ((uint32_t)p.y * 19349669u) ^ ((uint32_t)p.z * 83492791u));
    int r = h % max_keys;
    if (r < 0) r += max_keys;
    return r;
}
// make_int3(pos / GridSize): the float quotients truncated
I3 key_cell(const float* pos, float grid) { return I3{toi(pos[0] / grid), toi(pos[1] / grid), toi(pos[2] / grid)}; }
}  // namespace

// InsertKey (DeviceMap.cu:258-287) for keys[0..n) in order: a valid key of the same grid cell
// in the bucket gives the inserted key its position and its slot index; otherwise the first
// free slot of the bucket takes the key; a full bucket leaves the key out and its index as it was
extern "C" void or_keymap_insert(or_keymap* k, OrSurf* keys, int* index, int n) {
    for (int i = 0; i < n; ++i) {
        OrSurf& key = keys[i];
        const I3 p = key_cell(key.pos, k->grid);
        const int buck = key_hash(p, k->max_keys);
        int first = -1, found = -1;
        for (int j = 0; j < k->nbuckets; ++j) {
            const int b = buck * k->nbuckets + j;
            const OrSurf& s = k->keys[b];
            if (!s.valid && first == -1) first = b;
            if (s.valid && eq(key_cell(s.pos, k->grid), p)) {
                found = b;
                break;
            }
        }
        if (found >= 0) {
            index[i] = found;
            memcpy(key.pos, k->keys[found].pos, sizeof(key.pos));
        } else if (first != -1) {
            index[i] = first;
            k->keys[first] = key;
        }
    }
}

// CollectKeys (FuseMap.cu:384-406) in slot order: *n = the number of valid keys; valid key j
// (0-based) goes to out[j] for j >= 1 (the reference copies only at offsets > 0) and j < cap
extern "C" uint32_t or_keymap_collect(const or_keymap* k, OrSurf* out, uint32_t cap) {
    uint32_t n = 0;
    for (const OrSurf& s : k->keys)
        if (s.valid) {
            if (n > 0 && n < cap && out) out[n] = s;
            ++n;
        }
    return n;
}

extern "C" void or_keymap_read(const or_keymap* k, OrSurf* out) {
    memcpy(out, k->keys.data(), k->keys.size() * sizeof(OrSurf));
}

namespace {
inline float clamp1(float a) {  // KeyPoints.cu:6-10
    a = a > -1.f ? a : -1.f;
    a = a < 1.f ? a : 1.f;
    return a;
}
inline V4 s_normal(const OrSurf& s) { return V4{s.normal[0], s.normal[1], s.normal[2], s.normal[3]}; }
inline V3 s_pos(const OrSurf& s) { return V3{s.pos[0], s.pos[1], s.pos[2]}; }
}  // namespace

// BuildAdjecencyMatrixKernel (Tracking/KeyPoints.cu:12-57): n x n scores of the matched pairs
// (frame key i <-> map key i); a 4-D dot of the normals (float4 . float4) and of
// make_float4(direction) (w = 1) with a normal
extern "C" void or_adjacency(const OrSurf* frame, const OrSurf* map, const float* dist, int n, float* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            float score = 0;
            if (x == y) {
                score = expf(-dist[x]);
            } else {
                const OrSurf &m0 = map[x], &m1 = map[y], &f0 = frame[x], &f1 = frame[y];
                const float d00 = norm(s_pos(f0) - s_pos(f1));
                const float d01 = norm(s_pos(m0) - s_pos(m1));
                const V3 e10 = normalised(s_pos(f0) - s_pos(f1));
                const V3 e11 = normalised(s_pos(m0) - s_pos(m1));
                const V4 d10{e10.x, e10.y, e10.z, 1.0f}, d11{e11.x, e11.y, e11.z, 1.0f};
                if (d00 <= 1e-2 || d01 <= 1e-2) {
                    score = 0;
                } else {
                    const float alpha00 = acosf(clamp1(dot4(s_normal(f0), s_normal(f1))));
                    const float beta00 = acosf(clamp1(dot4(d10, s_normal(f0))));
                    const float gamma00 = acosf(clamp1(dot4(d10, s_normal(f1))));
                    const float alpha01 = acosf(clamp1(dot4(s_normal(m0), s_normal(m1))));
                    const float beta01 = acosf(clamp1(dot4(d11, s_normal(m0))));
                    const float gamma01 = acosf(clamp1(dot4(d11, s_normal(m1))));
                    score = expf(-(fabsf(d00 - d01) + fabsf(alpha00 - alpha01) + fabsf(beta00 - beta01) +
                                   fabsf(gamma00 - gamma01)));
                }
            }
            if (isn(score)) score = 0;
            out[(size_t)y * n + x] = score;
        }
}

// ------------------------------------------------------------------ keypoint relocalisation
// Frame::InterpDepth / InterpNormal / ExtractKeyPoints' per-keypoint test and back-projection
// (Core/Frame.cc:93-200).  The bounds test compares y with the width and x with the height,
// as written (:96, :133).  Parity pinned by an independent numpy statement of the same rules
// and the closed form of a plane (tests/test_oracle_reloc.py).
namespace {
// the test as written, and (where it would let the neighbours leave the image, which the
// reference then reads past) positions outside the image or NaN refused
bool or_in_frame(float x, float y, int cols, int rows) {
    if (x <= 1 || y <= 1 || y >= cols - 1 || x >= rows - 1) return false;
    return !(isn(x) || isn(y) || x >= cols - 1 || y >= rows - 1);
}
float or_interp_depth(const float* d, int cols, int rows, float x, float y, float cutoff) {
    const float nan = std::nanf("");
    if (!or_in_frame(x, y, cols, rows)) return nan;
    const float cx = x - std::floor(x), cy = y - std::floor(y);
    const int ulx = (int)std::floor(x), uly = (int)std::floor(y);
    const int llx = (int)std::floor(x), lly = (int)std::ceil(y);
    const int urx = (int)std::ceil(x), ury = (int)std::floor(y);
    const int lrx = (int)std::ceil(x), lry = (int)std::ceil(y);
    const float d00 = d[uly * cols + ulx];
    if (isn(d00) || d00 < 0.3 || d00 > cutoff) return nan;
    const float d10 = d[lly * cols + llx];
    if (isn(d10) || d10 < 0.3 || d10 > cutoff) return nan;
    const float d01 = d[ury * cols + urx];
    if (isn(d01) || d01 < 0.3 || d01 > cutoff) return nan;
    const float d11 = d[lry * cols + lrx];
    if (isn(d11) || d11 < 0.3 || d11 > cutoff) return nan;
    const float d0 = d01 * cx + d00 * (1 - cx);
    const float d1 = d11 * cx + d10 * (1 - cx);
    const float fin = (1 - cy) * d0 + cy * d1;
    return std::abs(fin - d00) <= 0.005 ? fin : nan;
}
V4 or_interp_normal(const float* nm, int cols, int rows, float x, float y) {
    const float nan = std::nanf("");
    const V4 bad{nan, nan, nan, nan};
    if (!or_in_frame(x, y, cols, rows)) return bad;
    const float cx = x - std::floor(x), cy = y - std::floor(y);
    auto at = [&](int xx, int yy) {
        const float* p = nm + 4 * ((size_t)yy * cols + xx);
        return V4{p[0], p[1], p[2], p[3]};
    };
    const V4 d00 = at((int)std::floor(x), (int)std::floor(y));
    const V4 d10 = at((int)std::floor(x), (int)std::ceil(y));
    const V4 d01 = at((int)std::ceil(x), (int)std::floor(y));
    const V4 d11 = at((int)std::ceil(x), (int)std::ceil(y));
    // float4 * float and float4 + float4 per component (VectorMath.h)
    auto sc = [](V4 a, float s) { return V4{a.x * s, a.y * s, a.z * s, a.w * s}; };
    auto ad = [](V4 a, V4 b) { return V4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; };
    const V4 d0 = ad(sc(d01, cx), sc(d00, 1 - cx));
    const V4 d1 = ad(sc(d11, cx), sc(d10, 1 - cx));
    const V4 fin = ad(sc(d0, 1 - cy), sc(d1, cy));
    const V4 df{fin.x - d00.x, fin.y - d00.y, fin.z - d00.z, fin.w - d00.w};
    const float n = std::sqrt(dot4(df, df));
    return n <= 0.1 ? fin : bad;
}
}  // namespace

extern "C" void or_lift_keypoints(const float* xy, int n, const float* depth, const float* nmap, int cols, int rows,
                                  float fx, float fy, float cx, float cy, float cutoff, float* pos, float* normal,
                                  uint8_t* ok) {
    for (int i = 0; i < n; ++i) {
        const float x = xy[2 * i], y = xy[2 * i + 1];
        const float nan = std::nanf("");
        float p[3] = {nan, nan, nan};
        V4 nr{nan, nan, nan, nan};
        bool good = false;
        const float dp = or_interp_depth(depth, cols, rows, x, y, cutoff);
        if (!isn(dp) && dp > 0.3 && dp < cutoff) {
            const V4 nn = or_interp_normal(nmap, cols, rows, x, y);
            if (!isn(nn.x)) {
                p[0] = dp * (x - cx) / fx;
                p[1] = dp * (y - cy) / fy;
                p[2] = dp;
                nr = nn;
                good = true;
            }
        }
        memcpy(pos + 3 * i, p, sizeof(p));
        normal[4 * i + 0] = nr.x;
        normal[4 * i + 1] = nr.y;
        normal[4 * i + 2] = nr.z;
        normal[4 * i + 3] = nr.w;
        ok[i] = good;
    }
}

// knnMatch, k = 2, L2 (the BFMatcher of Tracking.cc:436): every train descriptor in index
// order, the distance the square root of the sum of squared differences in dimension order;
// a later descriptor replaces one only when strictly nearer
extern "C" void or_match_knn2(const float* q, int nq, const float* t, int nt, int32_t* idx, float* dist) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nq; ++i) {
        float d0 = INFINITY, d1 = INFINITY;
        int i0 = -1, i1 = -1;
        for (int j = 0; j < nt; ++j) {
            float s = 0;
            for (int k = 0; k < 64; ++k) {
                const float e = q[(size_t)i * 64 + k] - t[(size_t)j * 64 + k];
                s = s + e * e;
            }
            const float d = std::sqrt(s);
            if (d < d0) {
                d1 = d0;
                i1 = i0;
                d0 = d;
                i0 = j;
            } else if (d < d1) {
                d1 = d;
                i1 = j;
            }
        }
        idx[2 * i] = i0;
        idx[2 * i + 1] = i1;
        dist[2 * i] = d0;
        dist[2 * i + 1] = d1;
    }
}

// ------------------------------------------------------------------ graph matching
// Tracker::GenerateGraph's selection of consistent pairs (Tracking/Tracking.cc:515-596) over
// the consistency matrix C of n matched pairs (row-major, C[a * n + b] = conMatrix(a, b)).
//   rank: the column sums (cv::cuda::reduce over the rows, CV_REDUCE_SUM, :520); OpenCV's
//     summation order is its own, so here the sum is taken in double in row order and
//     rounded to float;
//   order: the pairs by rank, largest first (cv::sortIdx descending, :527); ties, which
//     std::sort leaves in no fixed order, go to the lower index, and a NaN rank sorts as -inf;
//   list i (i < n_lists, :532-589): its head is the i-th pair in rank order; the pairs ranked
//     after it join while C[head][pair] > min_score, until max_sel pairs (the head included)
//     are selected; a list of three or more is refined: its k-th pair a (k >= 1; the head is
//     not kept, as written) stays unless a later pair b of the list is nearly exclusive with
//     it (C[a][b] or C[b][a] below 5e-3) and scores higher with the head (C[head][b] >
//     C[head][a]).
// Outputs per list: the number selected and the refined pairs (max_sel slots per list).
extern "C" void or_graph_select(const float* C, int n, int n_lists, int max_sel, float min_score, int32_t* n_selected,
                                int32_t* n_refined, int32_t* refined, float* rank_out, int32_t* order_out) {
    std::vector<float> rank(n);
    for (int x = 0; x < n; ++x) {
        double s = 0;
        for (int y = 0; y < n; ++y) s += (double)C[(size_t)y * n + x];
        rank[x] = (float)s;
    }
    std::vector<int> order(n);
    for (int x = 0; x < n; ++x) order[x] = x;
    auto key = [&](int x) { return isn(rank[x]) ? -INFINITY : rank[x]; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key(a) > key(b); });
    if (rank_out) memcpy(rank_out, rank.data(), sizeof(float) * n);
    if (order_out) memcpy(order_out, order.data(), sizeof(int32_t) * n);
    auto c = [&](int a, int b) { return C[(size_t)a * n + b]; };
    for (int i = 0; i < n_lists; ++i) {
        std::vector<int> sel;
        int head = -1;
        for (int j = i; j < n; ++j) {
            const int idx = order[j];
            if (sel.empty()) {
                sel.push_back(idx);
                head = idx;
            } else if (c(head, idx) > min_score) {
                sel.push_back(idx);
            }
            if ((int)sel.size() >= max_sel) break;
        }
        const int S = (int)sel.size();
        n_selected[i] = S;
        int R = 0;
        if (S >= 3)
            for (int k = 1; k < S; ++k) {
                const int a = sel[k];
                int l = k + 1;
                for (; l < S; ++l) {
                    const int b = sel[l];
                    if (c(a, b) < 5e-3f || c(b, a) < 5e-3f)
                        if (c(head, b) > c(head, a)) break;
                }
                if (l >= S) refined[(size_t)i * max_sel + R++] = a;
            }
        n_refined[i] = R;
    }
}

// Tracker::Relocalise with graph matching on (Tracking.cc:432-453) and GenerateGraph's list of
// pairs (:485-505): every query whose two nearest map keys pass the ratio test
// d0 < 0.9 d1 gives its nearest; every other gives both (in that order); pairs whose map key
// is not valid are dropped.  idx / dist: the knnMatch output (2 per query, -1 / +inf where
// missing); keys: the map's collected records.  The frame key of a pair takes the lifted
// position and normal of its query and is valid (the reference leaves `valid` of that record
// unset).  Returns the number of pairs (at most cap); *n_matches is the refined list's size
// before the validity filter (Relocalise gives up below 3).
extern "C" int or_graph_inputs(const int32_t* idx, const float* dist, int nq, const OrSurf* keys, int nk,
                               const float* frame_pos, const float* frame_normal, OrSurf* frame_out, OrSurf* map_out,
                               float* dist_out, int32_t* query_out, int cap, int* n_matches) {
    std::vector<std::pair<int, int>> ref;  // (query, slot 0 or 1)
    for (int i = 0; i < nq; ++i) {
        if (idx[2 * i] < 0) continue;
        if ((double)dist[2 * i] < 0.90 * (double)dist[2 * i + 1]) {
            ref.push_back({i, 0});
        } else {
            ref.push_back({i, 0});
            if (idx[2 * i + 1] >= 0) ref.push_back({i, 1});
        }
    }
    *n_matches = (int)ref.size();
    int n = 0;
    for (auto& r : ref) {
        const int t = idx[2 * r.first + r.second];
        if (t >= nk || !keys[t].valid) continue;
        if (n >= cap) break;
        OrSurf f{};
        f.valid = 1;
        memcpy(f.pos, frame_pos + 3 * r.first, sizeof(f.pos));
        memcpy(f.normal, frame_normal + 4 * r.first, sizeof(f.normal));
        frame_out[n] = f;
        map_out[n] = keys[t];
        dist_out[n] = dist[2 * r.first + r.second];
        query_out[n] = r.first;
        ++n;
    }
    return n;
}

// GenerateGraph's sub-graphs to PoseEstimate inputs (Tracking.cc:591-629), from the selection:
// the lists with three or more selected pairs are kept in list order and the number of
// PoseEstimate calls is M = min(n_select, kept).  The reference orders the kept lists with
// cv::sortIdx(cvNoSelected, tmp, CV_SORT_DESCENDING) on a one-column matrix, which sorts each
// one-element row, so tmp holds zeros and all M calls take the first kept list; a list of
// three or fewer refined pairs is skipped (:601-602).  Returns M when the first kept list has
// four or more refined pairs (its pairs in *list / *n_list), else 0.
extern "C" int or_graph_lists(const int32_t* n_selected, const int32_t* n_refined, const int32_t* refined, int n_lists,
                              int max_sel, int n_select, int32_t* list, int* n_list) {
    int kept = 0, first = -1;
    for (int i = 0; i < n_lists; ++i)
        if (n_selected[i] >= 3) {
            if (first < 0) first = i;
            ++kept;
        }
    *n_list = 0;
    if (kept == 0) return 0;
    const int M = kept < n_select ? kept : n_select;
    const int R = n_refined[first];
    if (R <= 3) return 0;
    memcpy(list, refined + (size_t)first * max_sel, sizeof(int32_t) * R);
    *n_list = R;
    return M;
}

// Tracker::ValidatePose (Tracking.cc:636-682) on a lost system's current frame: for each
// candidate pose, UpdateVisibility there (skipped below 512 visible blocks), RayTrace and
// ResizeImages of the model maps, and ComputeSE3(icpOnly, iters_reloc, thresh_reloc) from it;
// the first candidate that passes (the reference sorts the passing errors with
// cv::sortIdx on a one-column matrix, which leaves index 0 first) is rendered again and
// refined with the tracking iterations and threshold.  On success the frame is the
// relocalised one (Tracking.cc:101-105, System.cc:99-120): its pose the refined pose, swapped,
// the model raycast there, nothing fused; returns 1.  accepted[i]: -1 skipped for visibility,
// 0 rejected, 1 passed; *chosen the candidate refined (-1 if none).
namespace {
// UpdateVisibility at P, then (unless fewer than min_vis blocks are visible) RayTrace into the
// model maps (left as they are when nothing renders) and ResizeImages; false when skipped
bool or_render_at(or_system* s, const double P[16], uint32_t min_vis) {
    const int c0 = s->cols, r0 = s->rows;
    float R[9], Rinv[9], t[3];
    or_pose_f(P, R, Rinv, t);
    const uint32_t nvis =
        or_check_visibility(s->map, Rinv, t, c0, r0, s->K[0], s->K[1], s->K[2], s->K[3], s->map->dmax, s->map->dmin);
    if (nvis < min_vis) return false;
    uint32_t nfrag = 0;
    if (or_create_rendering_blocks(s->map, s->zmin.data(), s->zmax.data(), s->zcols, s->zrows, s->map->dmax,
                                   s->map->dmin, Rinv, t, nvis, s->K[0], s->K[1], s->K[2], s->K[3], &nfrag))
        or_raycast(s->map, s->mvmap[0].data(), s->mnmap[0].data(), c0, r0, s->zmin.data(), s->zmax.data(), s->zcols,
                   s->zrows, R, Rinv, t, (float)(1.0 / (double)s->K[0]), (float)(1.0 / (double)s->K[1]), s->K[2],
                   s->K[3]);
    for (int l = 1; l < 3; ++l)
        or_resize_map(s->mvmap[l - 1].data(), s->mnmap[l - 1].data(), c0 >> (l - 1), r0 >> (l - 1), s->mvmap[l].data(),
                      s->mnmap[l].data());
    return true;
}
}  // namespace

extern "C" int or_system_validate_poses(or_system* s, const double* poses, int n, const int iters_reloc[3],
                                        float thresh_reloc, int32_t* accepted, int32_t* chosen, double pose_out[16]) {
    *chosen = -1;
    if (s->last_state != -1) return -1;
    const float* nv[3] = {s->vmap[0].data(), s->vmap[1].data(), s->vmap[2].data()};
    const float* nn[3] = {s->nmap[0].data(), s->nmap[1].data(), s->nmap[2].data()};
    const float* lv[3] = {s->mvmap[0].data(), s->mvmap[1].data(), s->mvmap[2].data()};
    const float* ln[3] = {s->mnmap[0].data(), s->mnmap[1].data(), s->mnmap[2].data()};
    int first = -1;
    for (int i = 0; i < n; ++i) {
        const double* P = poses + 16 * i;
        accepted[i] = -1;
        if (!or_render_at(s, P, 512)) continue;
        double pose[16];
        float err = 0;
        int cnt = 0;
        const int st = or_compute_se3(nv, nn, lv, ln, s->cols, s->rows, s->K, iters_reloc, s->minc, thresh_reloc, P, pose,
                                      &err, &cnt);
        accepted[i] = st == 1;
        if (st == 1 && first < 0) first = i;
    }
    if (first < 0) return 0;
    *chosen = first;
    const double* P = poses + 16 * first;
    or_render_at(s, P, 0);
    double pose[16];
    float err = 0;
    int cnt = 0;
    const int st = or_compute_se3(nv, nn, lv, ln, s->cols, s->rows, s->K, s->iters, s->minc, s->thresh, P, pose, &err,
                                  &cnt);
    if (st != 1) return 0;
    // relocalised: Track() sets lastState = 0 and swaps the frame in; GrabImage (state -1)
    // renders it without fusion
    s->last_state = 0;
    s->last_err = err;
    s->last_cnt = cnt;
    memcpy(s->last_pose, pose, sizeof(pose));
    if (pose_out) memcpy(pose_out, pose, sizeof(pose));
    if (s->use_rgb)
        for (int l = 0; l < 3; ++l) s->last_img[l] = s->img[l];
    float R[9], Rinv[9], t[3];
    or_pose_f(pose, R, Rinv, t);
    const uint32_t nvis = or_check_visibility(s->map, Rinv, t, s->cols, s->rows, s->K[0], s->K[1], s->K[2], s->K[3],
                                              s->map->dmax, s->map->dmin);
    uint32_t nfrag = 0;
    if (or_create_rendering_blocks(s->map, s->zmin.data(), s->zmax.data(), s->zcols, s->zrows, s->map->dmax,
                                   s->map->dmin, Rinv, t, nvis, s->K[0], s->K[1], s->K[2], s->K[3], &nfrag)) {
        or_raycast(s->map, s->mvmap[0].data(), s->mnmap[0].data(), s->cols, s->rows, s->zmin.data(), s->zmax.data(),
                   s->zcols, s->zrows, R, Rinv, t, (float)(1.0 / (double)s->K[0]), (float)(1.0 / (double)s->K[1]),
                   s->K[2], s->K[3]);
    } else {
        s->mvmap[0] = s->vmap[0];
        s->mnmap[0] = s->nmap[0];
    }
    return 1;
}
