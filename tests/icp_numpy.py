"""Point-to-plane ICP association and rows restated in numpy, independently of the oracle.

Follows Tracking/Reduction.cu ICPReduce::searchPoint (:227-257) and getRow (:259-286) with the
operator semantics of Utility/VectorMath.h: `Matrix3f * float3` is three left-to-right float
dots (:191-193, :364-366), `cross` as written (:308-312), `norm` = sqrt of the float dot
(:320-322).  numpy float32 element-wise operations are single IEEE operations (no fused
multiply-add), so every float below is the one the reference's expression gives.

The projection (Reduction.cu:237-239) is `float invz = 1.0 / z;` (a double quotient stored in a
float), `(int)(x * invz * fx + cx + 0.5)`: float products and sum, the + 0.5 in double, then
truncation.  `rule="double"` evaluates the whole projection in double instead, only so that a
test can show that the two rules associate different pixels on the same maps.
"""
import numpy as np

f32 = np.float32


def pose_f(pose):
    """Frame::GpuRotation / GpuInvRotation / GpuTranslation (Core/Frame.cc:275-294)"""
    pose = np.asarray(pose, np.float64)
    return pose[:3, :3].astype(f32), pose[:3, :3].T.astype(f32), pose[:3, 3].astype(f32)


def mv(R, v):
    x, y, z = v[..., 0], v[..., 1], v[..., 2]
    return np.stack([R[i, 0] * x + R[i, 1] * y + R[i, 2] * z for i in range(3)], -1)


def dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def trunc_int(d):
    """(int) of a double: truncation (NaN -> 0 and saturation, as the GPU's conversion)"""
    d = np.asarray(d, np.float64)
    out = np.where(np.isnan(d), 0.0, np.clip(np.trunc(d), -2147483648.0, 2147483647.0))
    return out.astype(np.int64)


def associate(vcurr, ncurr, vlast, nlast, pose_curr, pose_last, K, rule="reference"):
    """searchPoint for every pixel: (found mask, matched last-frame pixel index, row (H, W, 7))"""
    rows, cols, _ = vcurr.shape
    Rc, _, tc = pose_f(pose_curr)
    Rl, Rli, tl = pose_f(pose_last)
    fx, fy, cx, cy = (f32(k) for k in K)
    with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
        vc = vcurr[..., :3].astype(f32)
        ok = ~np.isnan(vc[..., 0]) & ~(vc[..., 2].astype(np.float64) < 1e-3)
        vg = mv(Rc, vc) + tc
        vp = mv(Rli, vg - tl)
        if rule == "reference":
            invz = (1.0 / vp[..., 2].astype(np.float64)).astype(f32)
            pu = (vp[..., 0] * invz * fx + cx).astype(np.float64) + 0.5
            pv = (vp[..., 1] * invz * fy + cy).astype(np.float64) + 0.5
        elif rule == "double":
            invz = 1.0 / vp[..., 2].astype(np.float64)
            pu = vp[..., 0].astype(np.float64) * invz * float(fx) + float(cx) + 0.5
            pv = vp[..., 1].astype(np.float64) * invz * float(fy) + float(cy) + 0.5
        else:
            raise ValueError(rule)
        u, v = trunc_int(pu), trunc_int(pv)
        ok &= (u >= 0) & (v >= 0) & (u < cols) & (v < rows)
        j = np.where(ok, v * cols + u, 0)
        vl = vlast.reshape(-1, 4)[j][..., :3].astype(f32)
        nl = nlast.reshape(-1, 4)[j][..., :3].astype(f32)
        nc = ncurr[..., :3].astype(f32)
        vlast_g = mv(Rl, vl) + tl
        ncurr_g = mv(Rc, nc)
        nlast_g = mv(Rl, nl)
        dist = np.sqrt(dot(vlast_g - vg, vlast_g - vg))
        sine = np.sqrt(dot(cross(ncurr_g, nlast_g), cross(ncurr_g, nlast_g)))
        ok &= (sine < f32(0.6)) & (dist <= f32(0.1)) & ~np.isnan(nc[..., 0]) & ~np.isnan(nl[..., 0])
        # getRow: everything mapped back into the last camera frame
        nlr = mv(Rli, nlast_g)
        vcr = mv(Rli, vg - tl)
        vlr = mv(Rli, vlast_g - tl)
        mn = -nlr
        row = np.concatenate([mn, cross(nlr, vlr), dot(mn, vcr - vlr)[..., None]], -1)
    row = np.where(ok[..., None], row, f32(0)).astype(f32)
    return ok, np.where(ok, j, -1), row


def sums64(ok, row):
    """the 28 upper-triangular float products of each found row, summed in FP64, and the count
    (the order of CreateMatrix<6,7>, Reduction.cu:6-16); also the sum of |products| per entry"""
    r = row[ok].astype(f32)
    s = np.zeros(29)
    a = np.zeros(29)
    k = 0
    for i in range(7):
        for j in range(i, 7):
            p = (r[:, i] * r[:, j]).astype(np.float64)
            s[k] = p.sum()
            a[k] = np.abs(p).sum()
            k += 1
    s[28] = a[28] = float(ok.sum())
    return s, a
