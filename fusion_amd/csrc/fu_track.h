// FP64 pose arithmetic of Tracker::ComputeSE3 (Tracking/Tracking.cc:368-408): unpacking of
// the 29 reduced sums (Reduction.cu:6-16), the 6x6 solve, the SE(3) exponential and the
// pose composition.  Host and device share this code: the device-resident tracker runs it
// on one lane of the last ICP work-group, the reference-style path on the host.
#pragma once

#include <math.h>
#include "fu_math.h"
#include "fu_kernels.h"

namespace fu {

FU_HD void mat4_mul(const double* a, const double* b, double* o) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
            t[i * 4 + j] = s;
        }
    for (int i = 0; i < 16; ++i) o[i] = t[i];
}
FU_HD void rigid_inv(const double* a, double* o) {
    double t[16];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = a[j * 4 + i];
        t[i * 4 + 3] = -(a[0 * 4 + i] * a[3] + a[1 * 4 + i] * a[7] + a[2 * 4 + i] * a[11]);
    }
    t[12] = t[13] = t[14] = 0;
    t[15] = 1;
    for (int i = 0; i < 16; ++i) o[i] = t[i];
}

// LDL^T with symmetric (diagonal) pivoting, solving A x = b; zero pivots give zero
// components, as a pivoted LDL^T solve of a semidefinite system does.
FU_HD void ldlt_solve6(const double* Ain, const double* b, double* x) {
    double A[36];
    int perm[6];
    for (int i = 0; i < 36; ++i) A[i] = Ain[i];
    for (int i = 0; i < 6; ++i) perm[i] = i;
    for (int k = 0; k < 6; ++k) {
        // pivot: largest remaining diagonal magnitude
        int p = k;
        double best = fabs(A[k * 6 + k]);
        for (int i = k + 1; i < 6; ++i)
            if (fabs(A[i * 6 + i]) > best) {
                best = fabs(A[i * 6 + i]);
                p = i;
            }
        if (p != k) {
            for (int j = 0; j < 6; ++j) {
                double t = A[k * 6 + j];
                A[k * 6 + j] = A[p * 6 + j];
                A[p * 6 + j] = t;
            }
            for (int i = 0; i < 6; ++i) {
                double t = A[i * 6 + k];
                A[i * 6 + k] = A[i * 6 + p];
                A[i * 6 + p] = t;
            }
            int t = perm[k];
            perm[k] = perm[p];
            perm[p] = t;
        }
        const double d = A[k * 6 + k];
        if (d != 0.0) {
            for (int i = k + 1; i < 6; ++i) A[i * 6 + k] /= d;
            for (int i = k + 1; i < 6; ++i)
                for (int j = k + 1; j <= i; ++j) {
                    A[i * 6 + j] -= A[i * 6 + k] * A[j * 6 + k] * d;
                    A[j * 6 + i] = A[i * 6 + j];
                }
        } else {
            for (int i = k + 1; i < 6; ++i) A[i * 6 + k] = 0.0;
        }
    }
    double y[6];
    for (int i = 0; i < 6; ++i) y[i] = b[perm[i]];
    for (int i = 0; i < 6; ++i)
        for (int k = 0; k < i; ++k) y[i] -= A[i * 6 + k] * y[k];
    // tolerance of the pivot test: smallest normal double
    for (int i = 0; i < 6; ++i) y[i] = fabs(A[i * 6 + i]) > 2.2250738585072014e-308 ? y[i] / A[i * 6 + i] : 0.0;
    for (int i = 5; i >= 0; --i)
        for (int k = i + 1; k < 6; ++k) y[i] -= A[k * 6 + i] * y[k];
    for (int i = 0; i < 6; ++i) x[perm[i]] = y[i];
}

// SE(3) exponential of (upsilon, omega): R by Rodrigues, t = V upsilon with the left
// Jacobian V = I + (1-cos th)/th^2 W + (th - sin th)/th^3 W^2 (series for small th)
FU_HD void se3_exp(const double* xi, double* T) {
    const double wx = xi[3], wy = xi[4], wz = xi[5];
    const double th2 = wx * wx + wy * wy + wz * wz;
    const double th = sqrt(th2);
    double A, B, C;  // sin/th, (1-cos)/th^2, (th-sin)/th^3
    if (th < 1e-5) {
        A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
        B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
        C = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
    } else {
        A = sin(th) / th;
        B = (1.0 - cos(th)) / th2;
        C = (th - sin(th)) / (th2 * th);
    }
    const double W[9] = {0, -wz, wy, wz, 0, -wx, -wy, wx, 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += W[i * 3 + k] * W[k * 3 + j];
            W2[i * 3 + j] = s;
        }
    double R[9], V[9];
    for (int i = 0; i < 9; ++i) {
        const double I = (i % 4 == 0) ? 1.0 : 0.0;
        R[i] = I + A * W[i] + B * W2[i];
        V[i] = I + B * W[i] + C * W2[i];
    }
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T[i * 4 + j] = R[i * 3 + j];
        T[i * 4 + 3] = V[i * 3 + 0] * xi[0] + V[i * 3 + 1] * xi[1] + V[i * 3 + 2] * xi[2];
    }
    T[12] = T[13] = T[14] = 0;
    T[15] = 1;
}

// float views of the poses used by the kernels (Frame::GpuRotation/GpuInvRotation/
// GpuTranslation, Core/Frame.cc:275-294)
FU_HD void refresh_float_pose(TrackState* st) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            st->Rcurr[i * 3 + j] = (float)st->next_pose[i * 4 + j];
            st->Rlast[i * 3 + j] = (float)st->last_pose[i * 4 + j];
            st->RlastInv[i * 3 + j] = (float)st->last_pose[j * 4 + i];
            st->RcurrInv[i * 3 + j] = (float)st->next_pose[j * 4 + i];
        }
        st->tcurr[i] = (float)st->next_pose[i * 4 + 3];
        st->tlast[i] = (float)st->last_pose[i * 4 + 3];
    }
}

FU_HD double sin_euler_norm(const double* p) {
    // XYZ Euler angles of R = Rx(a)Ry(b)Rz(c): sin b = R02, sin a = -R12/|(R12,R22)|,
    // sin c = -R01/|(R01,R00)|; the norm of the three sines does not depend on the branch
    const double r00 = p[0], r01 = p[1], r02 = p[2], r12 = p[6], r22 = p[10];
    const double sa = -r12 / sqrt(r12 * r12 + r22 * r22);
    const double sb = r02;
    const double sc = -r01 / sqrt(r01 * r01 + r00 * r00);
    return sqrt(sa * sa + sb * sb + sc * sc);
}

// One ComputeSE3 iteration after the reduction (Tracking.cc:368-394) and, on the last
// iteration, the acceptance test (:398-408).
FU_HD void icp_update(TrackState* st, const float* sums, int min_count, int final_iter, float thresh) {
    double A[36], b[6];
    int shift = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            const double v = (double)sums[shift++];
            if (j == 6)
                b[i] = v;
            else
                A[j * 6 + i] = A[i * 6 + j] = v;
        }
    for (int i = 0; i < 36; ++i) st->A[i] = A[i];
    for (int i = 0; i < 6; ++i) st->b[i] = b[i];
    st->residual[0] = sums[27];
    st->residual[1] = sums[28];
    const float err = sqrtf(sums[27]) / sums[28];
    const int cnt = f2i(sums[28]);
    st->icp_error = err;
    st->count = cnt;
    st->iters_done += 1;
    if (isnanf_(err) || cnt < min_count) {
        for (int i = 0; i < 16; ++i) st->next_pose[i] = st->last_pose[i];
        st->status = TS_FAILED;
        refresh_float_pose(st);
        return;
    }
    double x[6], dT[16], dinv[16];
    ldlt_solve6(A, b, x);
    for (int i = 0; i < 6; ++i) st->x[i] = x[i];
    se3_exp(x, dT);
    mat4_mul(dT, st->delta, st->delta);
    rigid_inv(st->delta, dinv);
    mat4_mul(st->last_pose, dinv, st->next_pose);
    if (final_iter) {
        double pinv[16], p[16];
        rigid_inv(st->init_pose, pinv);
        mat4_mul(pinv, st->next_pose, p);
        const double tn = sqrt(p[3] * p[3] + p[7] * p[7] + p[11] * p[11]);
        const double an = sin_euler_norm(p);
        if (err < thresh && an <= 0.1 && tn <= 0.1) {
            st->status = TS_OK;
        } else {
            for (int i = 0; i < 16; ++i) st->next_pose[i] = st->last_pose[i];
            st->status = TS_REJECTED;
        }
    }
    refresh_float_pose(st);
}


}  // namespace fu
