// FP64 pose arithmetic of Tracker::ComputeSE3 (Tracking/Tracking.cc:368-408): unpacking of
// the 29 reduced sums (Reduction.cu:6-16), the 6x6 solve, the SE(3) exponential and the
// pose composition.  Host and device share this code: the device-resident tracker runs it
// on one lane of the last ICP work-group, the reference-style path on the host.  Every loop
// has a constant trip count and is unrolled, so on the device all matrices live in
// registers (no private-memory spills on the critical path of each ICP iteration).
#pragma once

#include <math.h>
#include "fu_math.h"
#include "fu_kernels.h"

#if defined(__HIPCC__)
#define FU_UNROLL _Pragma("unroll")
#else
#define FU_UNROLL
#endif

namespace fu {

FU_HD void mat4_mul(const double* a, const double* b, double* o) {
    double t[16];
    FU_UNROLL for (int i = 0; i < 4; ++i) FU_UNROLL for (int j = 0; j < 4; ++j) {
        double s = 0;
        FU_UNROLL for (int k = 0; k < 4; ++k) s += a[i * 4 + k] * b[k * 4 + j];
        t[i * 4 + j] = s;
    }
    FU_UNROLL for (int i = 0; i < 16; ++i) o[i] = t[i];
}

// inverse of a rigid transform [R t; 0 1] = [R^T -R^T t; 0 1]
FU_HD void rigid_inv(const double* a, double* o) {
    double t[16];
    FU_UNROLL for (int i = 0; i < 3; ++i) {
        FU_UNROLL for (int j = 0; j < 3; ++j) t[i * 4 + j] = a[j * 4 + i];
        t[i * 4 + 3] = -(a[0 * 4 + i] * a[3] + a[1 * 4 + i] * a[7] + a[2 * 4 + i] * a[11]);
    }
    t[12] = t[13] = t[14] = 0;
    t[15] = 1;
    FU_UNROLL for (int i = 0; i < 16; ++i) o[i] = t[i];
}

// A x = b for the symmetric positive definite 6x6 normal matrix by Cholesky A = L L^T.
// Returns false when A is not positive definite (a degenerate set of correspondences); the
// tracker then fails the frame.  The reference solves with Eigen's LDLT, which agrees with
// Cholesky on positive definite A to rounding.
// Cholesky solve of the 6x6 normal equations; each diagonal's reciprocal is taken once and
// multiplies the column and both substitutions (one division per column on the serial chain
// instead of one per element)
FU_HD bool chol_solve6(const double* A, const double* b, double* x) {
    double L[6][6], r[6];
    bool ok = true;
    FU_UNROLL for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        FU_UNROLL for (int k = 0; k < j; ++k) s -= L[j][k] * L[j][k];
        ok = ok && (s > 0.0);
        const double d = sqrt(s > 0.0 ? s : 1.0);
        L[j][j] = d;
        r[j] = 1.0 / d;
        FU_UNROLL for (int i = j + 1; i < 6; ++i) {
            double t = A[i * 6 + j];
            FU_UNROLL for (int k = 0; k < j; ++k) t -= L[i][k] * L[j][k];
            L[i][j] = t * r[j];
        }
    }
    double y[6];
    FU_UNROLL for (int i = 0; i < 6; ++i) {
        double t = b[i];
        FU_UNROLL for (int k = 0; k < i; ++k) t -= L[i][k] * y[k];
        y[i] = t * r[i];
    }
    FU_UNROLL for (int i = 5; i >= 0; --i) {
        double t = y[i];
        FU_UNROLL for (int k = i + 1; k < 6; ++k) t -= L[k][i] * x[k];
        x[i] = t * r[i];
    }
    return ok;
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
    FU_UNROLL for (int i = 0; i < 3; ++i) FU_UNROLL for (int j = 0; j < 3; ++j) {
        double s = 0;
        FU_UNROLL for (int k = 0; k < 3; ++k) s += W[i * 3 + k] * W[k * 3 + j];
        W2[i * 3 + j] = s;
    }
    double R[9], V[9];
    FU_UNROLL for (int i = 0; i < 9; ++i) {
        const double I = (i % 4 == 0) ? 1.0 : 0.0;
        R[i] = I + A * W[i] + B * W2[i];
        V[i] = I + B * W[i] + C * W2[i];
    }
    FU_UNROLL for (int i = 0; i < 3; ++i) {
        FU_UNROLL for (int j = 0; j < 3; ++j) T[i * 4 + j] = R[i * 3 + j];
        T[i * 4 + 3] = V[i * 3 + 0] * xi[0] + V[i * 3 + 1] * xi[1] + V[i * 3 + 2] * xi[2];
    }
    T[12] = T[13] = T[14] = 0;
    T[15] = 1;
}

// float views of the poses used by the kernels (Frame::GpuRotation/GpuInvRotation/
// GpuTranslation, Core/Frame.cc:275-294)
// the float poses of the kernels from a next (current) and last pose
FU_HD void set_float_pose(TrackState* st, const double* next, const double* last) {
    FU_UNROLL for (int i = 0; i < 3; ++i) {
        FU_UNROLL for (int j = 0; j < 3; ++j) {
            st->Rcurr[i * 3 + j] = (float)next[i * 4 + j];
            st->Rlast[i * 3 + j] = (float)last[i * 4 + j];
            st->RlastInv[i * 3 + j] = (float)last[j * 4 + i];
            st->RcurrInv[i * 3 + j] = (float)next[j * 4 + i];
        }
        st->tcurr[i] = (float)next[i * 4 + 3];
        st->tlast[i] = (float)last[i * 4 + 3];
    }
}

FU_HD void refresh_float_pose(TrackState* st) {
    FU_UNROLL for (int i = 0; i < 3; ++i) {
        FU_UNROLL for (int j = 0; j < 3; ++j) {
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

// CreateMatrix<6,7> (Reduction.cu:6-16): the 29 sums -> A (6x6, symmetric) and b (6)
FU_HD void unpack_sums29(const float* sums, double* A, double* b) {
    int shift = 0;
    FU_UNROLL for (int i = 0; i < 6; ++i) FU_UNROLL for (int j = i; j < 7; ++j) {
        const double v = (double)sums[shift++];
        if (j == 6)
            b[i] = v;
        else
            A[j * 6 + i] = A[i * 6 + j] = v;
    }
}

// One ComputeSE3 iteration after the reduction (Tracking.cc:368-394) and, on the last
// iteration, the acceptance test (:398-408).  `rgb` holds the photometric sums of the same
// iteration, or is null (icpOnly): their failure test (:341-346) fails the frame like ICP's,
// and the systems are blended as matA = matA_rgb * w + matA_icp with float w = 1e-4 (:379-383).
// the poses an update reads: last, the accumulated delta and the initial guess
struct TrackPoses {
    double last[16], delta[16], init[16];
};
FU_HD void load_track_poses(const TrackState* st, TrackPoses& P) {
    FU_UNROLL for (int i = 0; i < 16; ++i) {
        P.last[i] = st->last_pose[i];
        P.delta[i] = st->delta[i];
        P.init[i] = st->init_pose[i];
    }
}

// One ComputeSE3 iteration's update from the reduced sums (and the photometric sums when
// rgb is set): unpack, blend, solve, compose, accept.  Reads the poses from P and writes the
// state without reading it back.
FU_HD void icp_update(TrackState* st, const TrackPoses& P, const float* sums, const float* rgb, int min_count,
                      int final_iter, float thresh) {
    double A[36], b[6];
    unpack_sums29(sums, A, b);
    bool rgb_ok = true;
    if (rgb) {
        double Ar[36], br[6];
        unpack_sums29(rgb, Ar, br);
        const float rerr = sqrtf(rgb[27]) / rgb[28];
        const int rcnt = f2i(rgb[28]);
        st->rgb_error = rerr;
        st->rgb_count = rcnt;
        rgb_ok = !(isnanf_(rerr) || rcnt < min_count);
        const double w = (double)1e-4f;
        FU_UNROLL for (int i = 0; i < 36; ++i) A[i] = Ar[i] * w + A[i];
        FU_UNROLL for (int i = 0; i < 6; ++i) b[i] = br[i] * w + b[i];
    }
    FU_UNROLL for (int i = 0; i < 36; ++i) st->A[i] = A[i];
    FU_UNROLL for (int i = 0; i < 6; ++i) st->b[i] = b[i];
    st->residual[0] = sums[27];
    st->residual[1] = sums[28];
    const float err = sqrtf(sums[27]) / sums[28];
    const int cnt = f2i(sums[28]);
    st->icp_error = err;
    st->count = cnt;
    st->iters_done += 1;
    double x[6];
    const bool solved = rgb_ok && !(isnanf_(err) || cnt < min_count) && chol_solve6(A, b, x);
    if (!solved) {
        FU_UNROLL for (int i = 0; i < 16; ++i) st->next_pose[i] = P.last[i];
        st->status = TS_FAILED;
        set_float_pose(st, P.last, P.last);
        return;
    }
    double dT[16], delta[16], dinv[16], next[16];
    FU_UNROLL for (int i = 0; i < 6; ++i) st->x[i] = x[i];
    se3_exp(x, dT);
    mat4_mul(dT, P.delta, delta);
    rigid_inv(delta, dinv);
    mat4_mul(P.last, dinv, next);
    FU_UNROLL for (int i = 0; i < 16; ++i) st->delta[i] = delta[i];
    if (final_iter) {
        double pinv[16], p[16];
        rigid_inv(P.init, pinv);
        mat4_mul(pinv, next, p);
        const double tn = sqrt(p[3] * p[3] + p[7] * p[7] + p[11] * p[11]);
        const double an = sin_euler_norm(p);
        if (err < thresh && an <= 0.1 && tn <= 0.1) {
            st->status = TS_OK;
        } else {
            FU_UNROLL for (int i = 0; i < 16; ++i) next[i] = P.last[i];
            st->status = TS_REJECTED;
        }
    }
    FU_UNROLL for (int i = 0; i < 16; ++i) st->next_pose[i] = next[i];
    set_float_pose(st, next, P.last);
}
FU_HD void icp_update(TrackState* st, const float* sums, const float* rgb, int min_count, int final_iter,
                      float thresh) {
    TrackPoses P;
    load_track_poses(st, P);
    icp_update(st, P, sums, rgb, min_count, final_iter, thresh);
}

// The state machine of Tracker::GrabFrame (Tracking.cc:76-125) and System::GrabImage
// (System.cc:92-138) around one frame, shared by the device (one lane) and the host.
//
// frame_begin, before tracking: a lost tracker (state -1) does not track.  The reference
// then calls Relocalise on the frame; here a relocalisation pose given before the frame
// (fu_system_relocalise) places it (status TS_RELOC, the frame at that pose), otherwise the
// frame is TS_LOST (fu_system_validate_poses may relocalise that frame afterwards, as the
// reference's graph-matching Relocalise does).  A tracker that is not lost starts ComputeSE3 (Tracking.cc:305-314): last
// pose = tracked pose of the previous frame, initial guess = last pose, delta = identity.
FU_HD void frame_begin(TrackState* st) {
    if (st->sm_state == -1) {
        st->status = st->reloc ? TS_RELOC : TS_LOST;
        FU_UNROLL for (int i = 0; i < 16; ++i) st->next_pose[i] = st->last_pose[i];
        st->iters_done = 0;
        refresh_float_pose(st);
        return;
    }
    FU_UNROLL for (int i = 0; i < 16; ++i) {
        st->next_pose[i] = st->last_pose[i];
        st->init_pose[i] = st->last_pose[i];
        st->delta[i] = (i % 5 == 0) ? 1.0 : 0.0;
    }
    st->status = TS_RUNNING;
    st->iters_done = 0;
    refresh_float_pose(st);
}

// frame_commit, after tracking: the next state, the counters, what the frame does (gates)
// and, when the frame becomes the last frame, its pose as the last pose.
//   state 1 (first frame): InitTracking + SwapFrame; fused and raycast.
//   state 0: num_frames_after_reloc + 1; a tracked frame resets the missed count and, unless
//     the previous frame was lost (then GrabFrame returns false and nothing happens), is
//     swapped and raycast, and fused once 10 frames have passed since a relocalisation; a
//     frame not tracked counts as missed, and more than 9 missed frames make the tracker lost.
//   state -1: num_frames_after_reloc = 0; relocalised: swapped and raycast (UpdateVisibility
//     + RayTrace, no fusion), tracking resumes; otherwise still lost, nothing happens.
FU_HD void frame_commit(TrackState* st) {
    const int cur = st->sm_state;
    int next = 0, gates = 0;
    if (cur == 1) {
        gates = GATE_FUSE | GATE_RENDER | GATE_SWAP;
        if (st->after_reloc < 10) gates &= ~GATE_FUSE;
    } else if (cur == 0) {
        if (st->after_reloc < (1 << 30)) st->after_reloc += 1;
        if (st->status == TS_OK) {
            st->missed = 0;
            if (st->sm_prev != -1) gates = GATE_RENDER | GATE_SWAP | (st->after_reloc >= 10 ? GATE_FUSE : 0);
        } else {
            st->missed += 1;
            if (st->missed > 9) next = -1;
        }
    } else {
        st->after_reloc = 0;
        if (st->reloc) {
            st->reloc = 0;
            gates = GATE_RENDER | GATE_SWAP;
        } else {
            next = -1;
        }
    }
    if (st->mapping_off) gates &= ~GATE_FUSE;  // System.cc:110-112: mapping disabled, no fusion
    st->sm_prev = cur;
    st->sm_state = next;
    st->gates = gates;
    if (gates & GATE_SWAP)
        FU_UNROLL for (int i = 0; i < 16; ++i) st->last_pose[i] = st->next_pose[i];
    refresh_float_pose(st);
}

}  // namespace fu
