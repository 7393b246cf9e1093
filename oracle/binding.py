"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) - TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, as
what the GPU results are compared with.  Arrays are numpy, host memory.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FU_ORACLE_LIB names another build of the same sources (the planted-mistake variants that
# tests/test_oracle_planted.py builds)
LIB_PATH = os.environ.get("FU_ORACLE_LIB") or os.path.join(_HERE, "_build", "liboracle.so")

P = C.c_void_p
I = C.c_int
U = C.c_uint32
F = C.c_float
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)

_SIG = {
    "or_expf": (F, [F]),
    "or_filter_depth": (None, [P, P, P, I, I, F, F]),
    "or_pyrdown_gauss": (None, [P, I, I, P]),
    "or_compute_vmap": (None, [P, P, I, I, F, F, F, F, F]),
    "or_compute_nmap": (None, [P, P, I, I]),
    "or_resize_map": (None, [P, P, I, I, P, P]),
    "or_icp_sums": (None, [P, P, P, P, I, I, FP, FP, FP, FP, FP, F, F, F, F, FP, DP]),
    "or_solve6": (I, [DP, DP, DP]),
    "or_se3_exp": (None, [DP, DP]),
    "or_euler_sine_norm": (C.c_double, [DP]),
    "or_compute_se3": (I, [C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), I, I, FP,
                           C.POINTER(I), C.POINTER(I), F, DP, DP, FP, C.POINTER(I)]),
    "or_compute_se3_trace": (I, [C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), I, I, FP,
                                 C.POINTER(I), C.POINTER(I), F, DP, DP, DP, C.POINTER(I)]),
    "or_pose_f": (None, [DP, FP, FP, FP]),
    "or_image_to_intensity": (None, [P, P, I, I]),
    "or_pyrdown_u8": (None, [P, I, I, P]),
    "or_derivative_image": (None, [P, I, I, P, P]),
    "or_rgb_sums": (None, [P, P, P, P, P, P, I, I, FP, FP, FP, FP, FP, FP, F, F, F, F, C.POINTER(C.c_int32), FP,
                           FP, DP]),
    "or_compute_se3_rgbd": (I, [C.POINTER(P)] * 8 + [I, I, FP, C.POINTER(I), C.POINTER(I), F, DP, DP, FP,
                                                    C.POINTER(I), C.POINTER(I)]),
    "or_map_create": (P, [U, U, U, F, F, F, F, U, I]),
    "or_map_destroy": (None, [P]),
    "or_map_reset": (None, [P]),
    "or_map_counters": (None, [P, C.POINTER(U)]),
    "or_create_blocks": (None, [P, P, I, I, FP, FP, F, F, F, F]),
    "or_check_visibility": (U, [P, FP, FP, I, I, F, F, F, F, F, F]),
    "or_integrate": (None, [P, P, P, P, I, I, FP, FP, F, F, F, F]),
    "or_fuse_map_color": (U, [P, P, P, P, I, I, FP, FP, FP, F, F, F, F]),
    "or_create_rendering_blocks": (I, [P, P, P, I, I, F, F, FP, FP, U, F, F, F, F, C.POINTER(U)]),
    "or_raycast": (None, [P, P, P, I, I, P, P, I, I, FP, FP, FP, F, F, F, F]),
    "or_mesh": (U, [P, P, P, P, P, U]),
    "or_map_export_blocks": (U, [P, P, U]),
    "or_map_export_visible": (U, [P, P, U]),
    "or_map_read_blocks": (None, [P, P, U, P, P, P]),
    "or_map_read_blocks_f64": (None, [P, P, U, P]),
    "or_hash": (U, [C.c_int32, C.c_int32, C.c_int32, U]),
    "or_map_set_sphere": (None, [P, F, F, F, F]),
    "or_system_create": (P, [I, I, F, F, F, F, F, F, C.POINTER(I), C.POINTER(I), F, U, U, U, F, F, U, I]),
    "or_system_destroy": (None, [P]),
    "or_system_grab": (I, [P, P, P, DP]),
    "or_system_relocalise": (I, [P, DP]),
    "or_system_set_mapping": (None, [P, I]),
    "or_system_map": (P, [P]),
    "or_system_buffer": (FP, [P, I, I]),
    "or_system_set_rgb": (None, [P, I]),
    "or_system_image": (P, [P, I, I]),
    "or_system_rgb_count": (I, [P]),
    "or_system_icp_error": (C.c_float, [P, C.POINTER(C.c_int)]),
    "or_set_threads": (None, [I]),
    "or_keymap_create": (P, [F, I, I]),
    "or_keymap_destroy": (None, [P]),
    "or_keymap_reset": (None, [P]),
    "or_keymap_insert": (None, [P, P, P, I]),
    "or_keymap_collect": (U, [P, P, U]),
    "or_keymap_read": (None, [P, P]),
    "or_adjacency": (None, [P, P, P, I, P]),
    "or_lift_keypoints": (None, [P, I, P, P, I, I, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, P, P, P]),
    "or_match_knn2": (None, [P, I, P, I, P, P]),
    "or_graph_select": (None, [P, I, I, I, F, P, P, P, P, P]),
    "or_graph_inputs": (I, [P, P, I, P, I, P, P, P, P, P, P, I, C.POINTER(C.c_int)]),
    "or_graph_lists": (I, [P, P, P, I, I, I, P, C.POINTER(C.c_int)]),
    "or_system_validate_poses": (I, [P, DP, I, C.POINTER(I), F, P, P, DP]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: %s (make -C oracle)" % LIB_PATH)
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _fa(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.ctypes.data_as(FP), a


def euler_sine_norm(p):
    """|sin(XYZ Euler angles)| of a 4x4 motion, as the acceptance test of Tracking.cc:398-408"""
    p = np.ascontiguousarray(p, dtype=np.float64)
    return lib().or_euler_sine_norm(p.ctypes.data_as(DP))


def pose_f(pose):
    """(R, Rinv, t) float32 views of a 4x4 pose (Core/Frame.cc:275-294)"""
    pose = np.ascontiguousarray(pose, dtype=np.float64)
    R = np.zeros(9, np.float32)
    Ri = np.zeros(9, np.float32)
    t = np.zeros(3, np.float32)
    lib().or_pose_f(pose.ctypes.data_as(DP), R.ctypes.data_as(FP), Ri.ctypes.data_as(FP), t.ctypes.data_as(FP))
    return R, Ri, t


def filter_depth(depth_u16, depth_scale, cutoff):
    rows, cols = depth_u16.shape
    d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
    raw = np.zeros((rows, cols), np.float32)
    filt = np.zeros((rows, cols), np.float32)
    lib().or_filter_depth(_ptr(d), _ptr(raw), _ptr(filt), cols, rows, depth_scale, cutoff)
    return raw, filt


def pyrdown_gauss(src):
    rows, cols = src.shape
    s = np.ascontiguousarray(src, dtype=np.float32)
    dst = np.zeros((rows // 2, cols // 2), np.float32)
    lib().or_pyrdown_gauss(_ptr(s), cols, rows, _ptr(dst))
    return dst


def compute_vmap(depth, fx, fy, cx, cy, cutoff):
    rows, cols = depth.shape
    d = np.ascontiguousarray(depth, dtype=np.float32)
    v = np.zeros((rows, cols, 4), np.float32)
    lib().or_compute_vmap(_ptr(d), _ptr(v), cols, rows, fx, fy, cx, cy, cutoff)
    return v


def compute_nmap(vmap):
    rows, cols, _ = vmap.shape
    v = np.ascontiguousarray(vmap, dtype=np.float32)
    n = np.zeros_like(v)
    lib().or_compute_nmap(_ptr(v), _ptr(n), cols, rows)
    return n


def resize_map(vsrc, nsrc):
    rows, cols, _ = vsrc.shape
    v = np.ascontiguousarray(vsrc, dtype=np.float32)
    n = np.ascontiguousarray(nsrc, dtype=np.float32)
    vd = np.zeros((rows // 2, cols // 2, 4), np.float32)
    nd = np.zeros_like(vd)
    lib().or_resize_map(_ptr(v), _ptr(n), cols, rows, _ptr(vd), _ptr(nd))
    return vd, nd


def icp_sums(nv, lv, nn, ln, pose_curr, pose_last, K):
    rows, cols, _ = nv.shape
    Rc, _, tc = pose_f(pose_curr)
    Rl, Rli, tl = pose_f(pose_last)
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (nv, lv, nn, ln)]
    s32 = np.zeros(29, np.float32)
    s64 = np.zeros(29, np.float64)
    fp = lambda a: a.ctypes.data_as(FP)
    lib().or_icp_sums(*[_ptr(a) for a in arrs], cols, rows, fp(Rc), fp(tc), fp(Rl), fp(Rli), fp(tl),
                      K[0], K[1], K[2], K[3], fp(s32), s64.ctypes.data_as(DP))
    return s32, s64


def unpack_sums(s):
    """CreateMatrix<6,7> (Reduction.cu:6-16): A (6x6), b (6)"""
    A = np.zeros((6, 6))
    b = np.zeros(6)
    k = 0
    for i in range(6):
        for j in range(i, 7):
            if j == 6:
                b[i] = s[k]
            else:
                A[i, j] = A[j, i] = s[k]
            k += 1
    return A, b


def solve6(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(6)
    if lib().or_solve6(A.ctypes.data_as(DP), b.ctypes.data_as(DP), x.ctypes.data_as(DP)) != 0:
        raise np.linalg.LinAlgError("not positive definite")
    return x


def se3_exp(xi):
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    T = np.zeros(16)
    lib().or_se3_exp(xi.ctypes.data_as(DP), T.ctypes.data_as(DP))
    return T.reshape(4, 4)


def compute_se3(next_v, next_n, last_v, last_n, K, last_pose, iters=(10, 5, 4), min_count=(2000, 1000, 100),
                thresh=1e-4):
    keep = [np.ascontiguousarray(a, dtype=np.float32) for a in list(next_v) + list(next_n) + list(last_v) + list(last_n)]
    arr = lambda k: (C.c_void_p * 3)(*[keep[k * 3 + i].ctypes.data for i in range(3)])
    rows, cols, _ = keep[0].shape
    lp = np.ascontiguousarray(last_pose, dtype=np.float64).ravel()
    out = np.zeros(16)
    err = C.c_float()
    cnt = C.c_int()
    st = lib().or_compute_se3(arr(0), arr(1), arr(2), arr(3), cols, rows, (C.c_float * 4)(*K),
                              (C.c_int * 3)(*iters), (C.c_int * 3)(*min_count), thresh, lp.ctypes.data_as(DP),
                              out.ctypes.data_as(DP), C.byref(err), C.byref(cnt))
    return st, out.reshape(4, 4), err.value, cnt.value


def compute_se3_trace(next_v, next_n, last_v, last_n, K, last_pose, iters=(10, 5, 4), min_count=(2000, 1000, 100),
                      thresh=1e-4):
    """compute_se3 with every iteration recorded: (status, pose, [(level, pose_curr 4x4, x 6)])"""
    keep = [np.ascontiguousarray(a, dtype=np.float32) for a in list(next_v) + list(next_n) + list(last_v) + list(last_n)]
    arr = lambda k: (C.c_void_p * 3)(*[keep[k * 3 + i].ctypes.data for i in range(3)])
    rows, cols, _ = keep[0].shape
    lp = np.ascontiguousarray(last_pose, dtype=np.float64).ravel()
    out = np.zeros(16)
    tr = np.zeros((sum(iters), 22))
    n = C.c_int()
    st = lib().or_compute_se3_trace(arr(0), arr(1), arr(2), arr(3), cols, rows, (C.c_float * 4)(*K),
                                    (C.c_int * 3)(*iters), (C.c_int * 3)(*min_count), thresh, lp.ctypes.data_as(DP),
                                    out.ctypes.data_as(DP), tr.ctypes.data_as(DP), C.byref(n))
    levels = [l for l in (2, 1, 0) for _ in range(iters[l])]
    steps = [(levels[k], tr[k, :16].reshape(4, 4), tr[k, 16:]) for k in range(n.value)]
    return st, out.reshape(4, 4), steps


def image_to_intensity(rgb):
    """ImageToIntensity (Pyrdown.cu:183-194): (rows, cols, 3) u8 -> (rows, cols) u8"""
    c = np.ascontiguousarray(rgb, dtype=np.uint8)
    rows, cols = c.shape[:2]
    out = np.zeros((rows, cols), np.uint8)
    lib().or_image_to_intensity(_ptr(c), _ptr(out), cols, rows)
    return out


def pyrdown_u8(src):
    s = np.ascontiguousarray(src, dtype=np.uint8)
    rows, cols = s.shape
    dst = np.zeros((rows // 2, cols // 2), np.uint8)
    lib().or_pyrdown_u8(_ptr(s), cols, rows, _ptr(dst))
    return dst


def derivative_image(img):
    """ComputeDerivativeImage (Pyrdown.cu:220-249): (dIdx, dIdy) int16"""
    s = np.ascontiguousarray(img, dtype=np.uint8)
    rows, cols = s.shape
    dx = np.zeros((rows, cols), np.int16)
    dy = np.zeros((rows, cols), np.int16)
    lib().or_derivative_image(_ptr(s), cols, rows, _ptr(dx), _ptr(dy))
    return dx, dy


def rgb_sums(ni, li, nv, lv, dx, dy, pose_curr, pose_last, K):
    """RGBStep (Reduction.cu:537-627): ((count, sum diff^2), sigma, sums32, sums64)"""
    rows, cols = ni.shape
    Rc, Rci, tc = pose_f(pose_curr)
    Rl, Rli, tl = pose_f(pose_last)
    keep = [np.ascontiguousarray(ni, np.uint8), np.ascontiguousarray(li, np.uint8),
            np.ascontiguousarray(nv, np.float32), np.ascontiguousarray(lv, np.float32),
            np.ascontiguousarray(dx, np.int16), np.ascontiguousarray(dy, np.int16)]
    corr = np.zeros(2, np.int32)
    sigma = C.c_float()
    s32 = np.zeros(29, np.float32)
    s64 = np.zeros(29, np.float64)
    fp = lambda a: a.ctypes.data_as(FP)
    lib().or_rgb_sums(*[_ptr(a) for a in keep], cols, rows, fp(Rc), fp(Rci), fp(Rl), fp(Rli), fp(tc), fp(tl),
                      K[0], K[1], K[2], K[3], corr.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(sigma), fp(s32),
                      s64.ctypes.data_as(DP))
    return (int(corr[0]), int(corr[1])), sigma.value, s32, s64


def compute_se3_rgbd(next_v, next_n, last_v, last_n, next_i, last_i, dx, dy, K, last_pose, iters=(10, 5, 4),
                     min_count=(2000, 1000, 100), thresh=1e-4):
    """ComputeSE3(icpOnly=false): (status, pose, icp_error, icp_count, rgb_count)"""
    groups = [[np.ascontiguousarray(a, dtype=np.float32) for a in g] for g in (next_v, next_n, last_v, last_n)]
    groups += [[np.ascontiguousarray(a, dtype=np.uint8) for a in g] for g in (next_i, last_i)]
    groups += [[np.ascontiguousarray(a, dtype=np.int16) for a in g] for g in (dx, dy)]
    arrs = [(C.c_void_p * 3)(*[a.ctypes.data for a in g]) for g in groups]
    rows, cols = groups[4][0].shape
    lp = np.ascontiguousarray(last_pose, dtype=np.float64).ravel()
    out = np.zeros(16)
    err = C.c_float()
    cnt = C.c_int()
    rcnt = C.c_int()
    st = lib().or_compute_se3_rgbd(*arrs, cols, rows, (C.c_float * 4)(*K), (C.c_int * 3)(*iters),
                                   (C.c_int * 3)(*min_count), thresh, lp.ctypes.data_as(DP), out.ctypes.data_as(DP),
                                   C.byref(err), C.byref(cnt), C.byref(rcnt))
    return st, out.reshape(4, 4), err.value, cnt.value, rcnt.value


class Map:
    def __init__(self, num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
                 trunc_dist=None, depth_min=0.1, depth_max=3.0, max_weight=255, precision=0):
        if trunc_dist is None:
            trunc_dist = float(np.float32(voxel_size) * np.float32(3.0))
        self.h = lib().or_map_create(num_buckets, num_excess, num_blocks, voxel_size, trunc_dist, depth_min,
                                     depth_max, max_weight, precision)
        self.owned = True
        self.depth_min, self.depth_max = depth_min, depth_max

    def __del__(self):
        try:
            if self.owned and self.h:
                lib().or_map_destroy(self.h)
        except Exception:
            pass

    def reset(self):
        lib().or_map_reset(self.h)

    def counters(self):
        c = (U * 5)()
        lib().or_map_counters(self.h, c)
        return {"n_allocated": c[0], "n_excess": c[1], "n_visible": c[2], "heap_full": c[3], "excess_full": c[4]}

    def create_blocks(self, depth, pose, K):
        rows, cols = depth.shape
        R, _, t = pose_f(pose)
        d = np.ascontiguousarray(depth, dtype=np.float32)
        lib().or_create_blocks(self.h, _ptr(d), cols, rows, R.ctypes.data_as(FP), t.ctypes.data_as(FP), *K)

    def check_visibility(self, pose, cols, rows, K, depth_max=None, depth_min=None):
        _, Ri, t = pose_f(pose)
        return lib().or_check_visibility(self.h, Ri.ctypes.data_as(FP), t.ctypes.data_as(FP), cols, rows, *K,
                                         self.depth_max if depth_max is None else depth_max,
                                         self.depth_min if depth_min is None else depth_min)

    def integrate(self, depth, rgb, nmap, pose, K):
        rows, cols = depth.shape
        _, Ri, t = pose_f(pose)
        d = np.ascontiguousarray(depth, dtype=np.float32)
        c = np.ascontiguousarray(rgb, dtype=np.uint8)
        n = np.ascontiguousarray(nmap, dtype=np.float32)
        lib().or_integrate(self.h, _ptr(d), _ptr(c), _ptr(n), cols, rows, Ri.ctypes.data_as(FP),
                           t.ctypes.data_as(FP), *K)

    def fuse_map_color(self, depth, rgb, nmap, pose, K):
        rows, cols = depth.shape
        R, Ri, t = pose_f(pose)
        d = np.ascontiguousarray(depth, dtype=np.float32)
        c = np.ascontiguousarray(rgb, dtype=np.uint8)
        n = np.ascontiguousarray(nmap, dtype=np.float32)
        return lib().or_fuse_map_color(self.h, _ptr(d), _ptr(c), _ptr(n), cols, rows, R.ctypes.data_as(FP),
                                       Ri.ctypes.data_as(FP), t.ctypes.data_as(FP), *K)

    def create_rendering_blocks(self, zcols, zrows, pose, n_visible, K, depth_max=None, depth_min=None):
        _, Ri, t = pose_f(pose)
        zmin = np.zeros((zrows, zcols), np.float32)
        zmax = np.zeros((zrows, zcols), np.float32)
        nf = U()
        r = lib().or_create_rendering_blocks(self.h, _ptr(zmin), _ptr(zmax), zcols, zrows,
                                             self.depth_max if depth_max is None else depth_max,
                                             self.depth_min if depth_min is None else depth_min,
                                             Ri.ctypes.data_as(FP), t.ctypes.data_as(FP), n_visible, *K, C.byref(nf))
        return zmin, zmax, nf.value, bool(r)

    def raycast(self, cols, rows, zmin, zmax, pose, K):
        R, Ri, t = pose_f(pose)
        v = np.zeros((rows, cols, 4), np.float32)
        n = np.zeros_like(v)
        zmin = np.ascontiguousarray(zmin, np.float32)
        zmax = np.ascontiguousarray(zmax, np.float32)
        zrows, zcols = zmin.shape
        invfx = np.float32(1.0 / np.float64(np.float32(K[0])))
        invfy = np.float32(1.0 / np.float64(np.float32(K[1])))
        lib().or_raycast(self.h, _ptr(v), _ptr(n), cols, rows, _ptr(zmin), _ptr(zmax), zcols, zrows,
                         R.ctypes.data_as(FP), Ri.ctypes.data_as(FP), t.ctypes.data_as(FP), invfx, invfy, K[2], K[3])
        return v, n

    def mesh(self, max_tri=2_000_000, with_meta=True):
        total = lib().or_mesh(self.h, None, None, None, None, 0)
        k = min(total, max_tri)
        vert = np.zeros((k, 3, 3), np.float32)
        norm = np.zeros((k, 3, 3), np.float32)
        col = np.zeros((k, 3, 3), np.uint8)
        meta = np.zeros((k, 4), np.int32)
        lib().or_mesh(self.h, _ptr(vert), _ptr(norm), _ptr(col), _ptr(meta) if with_meta else None, k)
        return vert, norm, col, meta, total

    def set_sphere(self, c, r):
        lib().or_map_set_sphere(self.h, c[0], c[1], c[2], r)

    def export_blocks(self):
        n = lib().or_map_export_blocks(self.h, None, 0)
        out = np.zeros((n, 3), np.int32)
        lib().or_map_export_blocks(self.h, _ptr(out), n)
        return out

    def export_visible(self):
        n = lib().or_map_export_visible(self.h, None, 0)
        out = np.zeros((n, 3), np.int32)
        lib().or_map_export_visible(self.h, _ptr(out), n)
        return out

    def read_blocks(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        n = xyz.shape[0]
        sdf = np.zeros((n, 512), np.float32)
        w = np.zeros((n, 512), np.uint16)
        rgb = np.zeros((n, 512, 3), np.uint8)
        lib().or_map_read_blocks(self.h, _ptr(xyz), n, _ptr(sdf), _ptr(w), _ptr(rgb))
        return sdf, w, rgb

    def read_blocks_f64(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        sdf = np.zeros((xyz.shape[0], 512), np.float64)
        lib().or_map_read_blocks_f64(self.h, _ptr(xyz), xyz.shape[0], _ptr(sdf))
        return sdf


class System:
    """System::GrabImage on the CPU: ICP only, or with the photometric term (use_rgb)"""

    def __init__(self, cols=640, rows=480, K=(525.0, 525.0, 319.5, 239.5), depth_scale=5000.0, depth_cutoff=3.0,
                 iters=(10, 5, 4), min_count=(2000, 1000, 100), icp_thresh=1e-4, num_buckets=1 << 18,
                 num_excess=1 << 17, num_blocks=200_000, voxel_size=0.01, trunc_dist=None, max_weight=255,
                 precision=0, use_rgb=False):
        if trunc_dist is None:
            trunc_dist = float(np.float32(voxel_size) * np.float32(3.0))
        self.cols, self.rows = cols, rows
        self.h = lib().or_system_create(cols, rows, *K, depth_scale, depth_cutoff, (I * 3)(*iters),
                                        (I * 3)(*min_count), icp_thresh, num_buckets, num_excess, num_blocks,
                                        voxel_size, trunc_dist, max_weight, precision)
        self.map = Map.__new__(Map)
        self.map.h = lib().or_system_map(self.h)
        self.map.owned = False
        self.map.depth_min, self.map.depth_max = 0.1, 3.0
        lib().or_system_set_rgb(self.h, 1 if use_rgb else 0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def close(self):
        if self.h:
            self.map.h = None
            lib().or_system_destroy(self.h)
            self.h = None

    def grab(self, depth_u16, rgb):
        d = np.ascontiguousarray(depth_u16, dtype=np.uint16)
        c = np.ascontiguousarray(rgb, dtype=np.uint8)
        pose = np.zeros(16)
        st = lib().or_system_grab(self.h, _ptr(d), _ptr(c), pose.ctypes.data_as(DP))
        return st, pose.reshape(4, 4)

    def set_mapping(self, enabled):
        lib().or_system_set_mapping(self.h, 1 if enabled else 0)

    def validate_poses(self, poses, iters_reloc=(3, 3, 2), thresh_reloc=1e-3):
        """Tracker::ValidatePose on the current (lost) frame: (ok, accepted per candidate,
        chosen candidate, refined pose)"""
        P = np.ascontiguousarray(np.asarray(poses, dtype=np.float64).reshape(-1, 4, 4))
        n = len(P)
        acc = np.zeros(max(n, 1), np.int32)
        chosen = np.zeros(1, np.int32)
        out = np.eye(4)
        ok = lib().or_system_validate_poses(self.h, P.ctypes.data_as(DP), n, (I * 3)(*iters_reloc), thresh_reloc,
                                            _ptr(acc), _ptr(chosen), out.ctypes.data_as(DP))
        return ok, acc[:n], int(chosen[0]), out

    def relocalise(self, pose):
        p = np.ascontiguousarray(pose, dtype=np.float64)
        return lib().or_system_relocalise(self.h, p.ctypes.data_as(DP))

    def buffer(self, which, level=0):
        rows, cols = self.rows >> level, self.cols >> level
        shape = (rows, cols) if which in (0, 5) else (rows, cols, 4)
        ptr = lib().or_system_buffer(self.h, which, level)
        return np.ctypeslib.as_array(ptr, shape=shape).copy()


    def image(self, which, level=0):
        """0 intensity, 1 last frame's intensity (u8); 2 dIdx, 3 dIdy (i16)"""
        rows, cols = self.rows >> level, self.cols >> level
        ct = C.c_uint8 if which < 2 else C.c_int16
        ptr = C.cast(lib().or_system_image(self.h, which, level), C.POINTER(ct))
        return np.ctypeslib.as_array(ptr, shape=(rows, cols)).copy()

    def rgb_count(self):
        return lib().or_system_rgb_count(self.h)

    def icp_error(self):
        """final ICP error and correspondence count of the last tracked frame"""
        n = C.c_int(0)
        e = lib().or_system_icp_error(self.h, C.byref(n))
        return float(e), n.value


def set_threads(n):
    lib().or_set_threads(n)


# ---------------------------------------------------------------------- keypoint map
# SURF records (Mapping/DeviceMap.h:99-108) as a numpy structured array of 288 bytes
SURF = np.dtype([("valid", np.uint8), ("pad", np.uint8, 3), ("pos", np.float32, 3), ("normal", np.float32, 4),
                 ("descriptor", np.float32, 64)])
assert SURF.itemsize == 288


class KeyMap:
    """KeyMap (DeviceMap.h:163-185) with ResetKeyPoints / InsertKeyPoints / CollectKeyPoints"""

    def __init__(self, grid_size=0.01, max_keys=100000, n_buckets=5):
        self.n = max_keys * n_buckets
        self.h = lib().or_keymap_create(grid_size, max_keys, n_buckets)
        lib().or_keymap_reset(self.h)

    def __del__(self):
        try:
            lib().or_keymap_destroy(self.h)
        except Exception:
            pass

    def reset(self):
        lib().or_keymap_reset(self.h)

    def insert(self, keys, index):
        """keys: SURF array (positions rewritten for keys that met a key of their cell), index:
        int32 array (slot of each key; unchanged for a key that found its bucket full)"""
        assert keys.dtype == SURF and index.dtype == np.int32 and keys.flags.c_contiguous
        lib().or_keymap_insert(self.h, _ptr(keys), _ptr(index), len(keys))

    def collect(self, cap):
        out = np.zeros(cap, SURF)
        n = lib().or_keymap_collect(self.h, _ptr(out), cap)
        return out, n

    def read(self):
        out = np.zeros(self.n, SURF)
        lib().or_keymap_read(self.h, _ptr(out))
        return out


def adjacency(frame_keys, map_keys, dist):
    n = len(frame_keys)
    out = np.zeros((n, n), np.float32)
    fk = np.ascontiguousarray(frame_keys)
    mk = np.ascontiguousarray(map_keys)
    d = np.ascontiguousarray(dist, dtype=np.float32)
    lib().or_adjacency(_ptr(fk), _ptr(mk), _ptr(d), n, _ptr(out))
    return out


def lift_keypoints(xy, depth, nmap, K, cutoff=3.0):
    """Frame::ExtractKeyPoints' test and back-projection: (pos (n, 3), normal (n, 4), ok (n,))"""
    xy = np.ascontiguousarray(xy, dtype=np.float32)
    d = np.ascontiguousarray(depth, dtype=np.float32)
    nm = np.ascontiguousarray(nmap, dtype=np.float32)
    n = len(xy)
    rows, cols = d.shape
    pos = np.zeros((n, 3), np.float32)
    nrm = np.zeros((n, 4), np.float32)
    ok = np.zeros(n, np.uint8)
    lib().or_lift_keypoints(_ptr(xy), n, _ptr(d), _ptr(nm), cols, rows, K[0], K[1], K[2], K[3], cutoff, _ptr(pos),
                            _ptr(nrm), _ptr(ok))
    return pos, nrm, ok


def match_knn2(query, train):
    """knnMatch, k = 2, L2: (idx (n, 2), dist (n, 2))"""
    q = np.ascontiguousarray(query, dtype=np.float32)
    t = np.ascontiguousarray(train, dtype=np.float32)
    idx = np.zeros((len(q), 2), np.int32)
    dist = np.zeros((len(q), 2), np.float32)
    lib().or_match_knn2(_ptr(q), len(q), _ptr(t), len(t), _ptr(idx), _ptr(dist))
    return idx, dist


def graph_select(matrix, n_lists=10, max_sel=400, min_score=0.1):
    """GenerateGraph's sub-graph selection over a consistency matrix: (n_selected (n_lists,),
    refined lists (list of arrays), rank sums (n,), rank order (n,))"""
    C_ = np.ascontiguousarray(matrix, dtype=np.float32)
    n = C_.shape[0]
    nsel = np.zeros(n_lists, np.int32)
    nref = np.zeros(n_lists, np.int32)
    ref = np.zeros((n_lists, max_sel), np.int32)
    rank = np.zeros(max(n, 1), np.float32)
    order = np.zeros(max(n, 1), np.int32)
    lib().or_graph_select(_ptr(C_), n, n_lists, max_sel, min_score, _ptr(nsel), _ptr(nref), _ptr(ref), _ptr(rank),
                          _ptr(order))
    return nsel, [ref[i, : nref[i]].copy() for i in range(n_lists)], rank[:n], order[:n]


def graph_inputs(idx, dist, keys, frame_pos, frame_normal):
    """Relocalise's match list with graph matching on, filtered by map-key validity:
    (frame keys, map keys, distances, query of each pair, size of the list before the filter)"""
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    dist = np.ascontiguousarray(dist, dtype=np.float32)
    keys = np.ascontiguousarray(keys)
    fp = np.ascontiguousarray(frame_pos, dtype=np.float32)
    fn = np.ascontiguousarray(frame_normal, dtype=np.float32)
    nq = len(idx)
    cap = 2 * nq
    fk = np.zeros(max(cap, 1), SURF)
    mk = np.zeros(max(cap, 1), SURF)
    d = np.zeros(max(cap, 1), np.float32)
    q = np.zeros(max(cap, 1), np.int32)
    nm = C.c_int(0)
    n = lib().or_graph_inputs(_ptr(idx), _ptr(dist), nq, _ptr(keys), len(keys), _ptr(fp), _ptr(fn), _ptr(fk), _ptr(mk),
                              _ptr(d), _ptr(q), cap, C.byref(nm))
    return fk[:n], mk[:n], d[:n], q[:n], nm.value


def graph_lists(n_selected, refined, max_sel=400, n_select=5):
    """the number of PoseEstimate calls GenerateGraph makes and the pairs they take"""
    nl = len(n_selected)
    nsel = np.ascontiguousarray(n_selected, dtype=np.int32)
    nref = np.array([len(r) for r in refined], np.int32)
    ref = np.zeros((nl, max_sel), np.int32)
    for i, r in enumerate(refined):
        ref[i, : len(r)] = r
    out = np.zeros(max_sel, np.int32)
    nlist = C.c_int(0)
    M = lib().or_graph_lists(_ptr(nsel), _ptr(nref), _ptr(ref), nl, max_sel, n_select, _ptr(out), C.byref(nlist))
    return M, out[: nlist.value].copy()
