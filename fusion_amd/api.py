"""Python host mirror of the reference's host interface over the C ABI.

Names follow the reference: the free functions of Tracking/Reduction.h (FilterDepth,
PyrDownGauss, ComputeVMap, ComputeNMap, ResizeMap, ICPStep) and Mapping/RenderScene.h
(ResetMap, FuseMapColor, CheckBlockVisibility, CreateRenderingBlocks, Raycast, MeshScene),
the Mapping object (Mapping/Mapping.h) and System::GrabImage (Core/System.h).  Device
memory and streams come from PyTorch (ROCm); every call runs the HIP kernels of
libfusion_amd.so on the current stream.  Images are torch tensors on the GPU: depth u16 as
int16 (H, W), ranges float32 (H, W), vertex/normal maps float32 (H, W, 4), colour uint8
(H, W, 3).  Poses are 4x4 camera-to-world numpy float64 arrays.
"""
import ctypes as C
import os

import numpy as np
import torch

from . import native as N
from .native import check


def _addr(t):
    """device address of a dense tensor (the library reads images as dense row-major arrays)"""
    if not t.is_contiguous():
        raise N.FusionError("tensor of shape %s is not contiguous; pass .contiguous()" % (tuple(t.shape),))
    return t.data_ptr()


def _p(t):
    return C.c_void_p(_addr(t)) if t is not None else None


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev():
    if not torch.cuda.is_available():
        raise N.FusionError("fusion_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda")


def pose_f(pose):
    """Frame::GpuRotation/GpuInvRotation/GpuTranslation (Core/Frame.cc:275-294)."""
    pose = np.asarray(pose, dtype=np.float64).reshape(4, 4)
    p = N.PoseF()
    R = pose[:3, :3]
    for i in range(9):
        p.R[i] = float(np.float32(R.flat[i]))
        p.Rinv[i] = float(np.float32(R.T.flat[i]))
    for i in range(3):
        p.t[i] = float(np.float32(pose[i, 3]))
    return p


def level_k(K, level):
    """Intrinsics::operator()(scale): (fx, fy, cx, cy) / 2^level in float32."""
    s = np.float32(1 << level)
    return tuple(float(np.float32(k) / s) for k in K)


# --------------------------------------------------------------------- Reduction.h
def FilterDepth(depth_u16, depth_scale, depth_cutoff):
    rows, cols = depth_u16.shape
    raw = torch.empty((rows, cols), dtype=torch.float32, device=depth_u16.device)
    filt = torch.empty_like(raw)
    check(N.lib().fu_filter_depth(_p(depth_u16), _p(raw), _p(filt), cols, rows, depth_scale, depth_cutoff, _s()),
          "fu_filter_depth")
    return raw, filt


def PyrDownGauss(src):
    rows, cols = src.shape
    dst = torch.empty((rows // 2, cols // 2), dtype=torch.float32, device=src.device)
    check(N.lib().fu_pyrdown_gauss(_p(src), cols, rows, _p(dst), _s()), "fu_pyrdown_gauss")
    return dst


def ComputeVMap(depth, fx, fy, cx, cy, depth_cutoff):
    rows, cols = depth.shape
    vmap = torch.empty((rows, cols, 4), dtype=torch.float32, device=depth.device)
    check(N.lib().fu_compute_vmap(_p(depth), _p(vmap), cols, rows, fx, fy, cx, cy, depth_cutoff, _s()),
          "fu_compute_vmap")
    return vmap


def ComputeNMap(vmap):
    rows, cols, _ = vmap.shape
    nmap = torch.empty_like(vmap)
    check(N.lib().fu_compute_nmap(_p(vmap), _p(nmap), cols, rows, _s()), "fu_compute_nmap")
    return nmap


def ResizeMap(vsrc, nsrc):
    rows, cols, _ = vsrc.shape
    vdst = torch.empty((rows // 2, cols // 2, 4), dtype=torch.float32, device=vsrc.device)
    ndst = torch.empty_like(vdst)
    check(N.lib().fu_resize_map(_p(vsrc), _p(nsrc), cols, rows, _p(vdst), _p(ndst), _s()), "fu_resize_map")
    return vdst, ndst


def ICPStep(next_vmap, last_vmap, next_nmap, last_nmap, pose_curr, pose_last, K):
    """One ICP reduction: returns (A 6x6, b 6, residual[2]) as the reference's ICPStep."""
    rows, cols, _ = next_vmap.shape
    pc, pl = pose_f(pose_curr), pose_f(pose_last)
    res = (C.c_float * 2)()
    A = (C.c_double * 36)()
    b = (C.c_double * 6)()
    check(N.lib().fu_icp_step(_p(next_vmap), _p(last_vmap), _p(next_nmap), _p(last_nmap), cols, rows,
                              pc.R, pc.t, pl.R, pl.Rinv, pl.t, K[0], K[1], K[2], K[3], res, A, b, _s()),
          "fu_icp_step")
    return np.array(A[:]).reshape(6, 6), np.array(b[:]), np.array(res[:], dtype=np.float32)


# --------------------------------------------------------------------- photometric term
def ImageToIntensity(rgb):
    """ImageToIntensity (Pyrdown.cu:196-206): device uint8 (H, W, 3) -> uint8 (H, W)"""
    rows, cols = rgb.shape[:2]
    img = torch.empty((rows, cols), dtype=torch.uint8, device=rgb.device)
    check(N.lib().fu_image_to_intensity(_p(rgb), _p(img), cols, rows, _s()), "fu_image_to_intensity")
    return img


def PyrDownGaussU8(src):
    """PyrDownGauss on bytes (Pyrdown.cu:171-181)"""
    rows, cols = src.shape
    dst = torch.empty((rows // 2, cols // 2), dtype=torch.uint8, device=src.device)
    check(N.lib().fu_pyrdown_gauss_u8(_p(src), cols, rows, _p(dst), _s()), "fu_pyrdown_gauss_u8")
    return dst


def ComputeDerivativeImage(image):
    """ComputeDerivativeImage (Pyrdown.cu:238-249): (dIdx, dIdy) int16"""
    rows, cols = image.shape
    dx = torch.empty((rows, cols), dtype=torch.int16, device=image.device)
    dy = torch.empty_like(dx)
    check(N.lib().fu_compute_derivative_image(_p(image), _p(dx), _p(dy), cols, rows, _s()),
          "fu_compute_derivative_image")
    return dx, dy


def RGBStep(next_image, last_image, next_vmap, last_vmap, dIdx, dIdy, pose_curr, pose_last, K):
    """One photometric reduction (Reduction.cu:537-627): (A 6x6, b 6, residual[2],
    (count, sum diff^2))"""
    rows, cols = next_image.shape
    pc, pl = pose_f(pose_curr), pose_f(pose_last)
    res = (C.c_float * 2)()
    A = (C.c_double * 36)()
    b = (C.c_double * 6)()
    corr = (C.c_int32 * 2)()
    check(N.lib().fu_rgb_step(_p(next_image), _p(last_image), _p(next_vmap), _p(last_vmap), _p(dIdx), _p(dIdy),
                              cols, rows, pc.R, pc.Rinv, pl.R, pl.Rinv, pc.t, pl.t, K[0], K[1], K[2], K[3], res, A, b,
                              corr, _s()), "fu_rgb_step")
    return (np.array(A[:]).reshape(6, 6), np.array(b[:]), np.array(res[:], dtype=np.float32),
            (int(corr[0]), int(corr[1])))


def WriteMeshToDisk(path, vertex, normal, color, binary=False):
    """System::WriteMeshToDisk (Core/System.cc:140-193): a PLY triangle soup of MeshScene's
    output (device tensors (n, 3, 3) float, float, uint8)"""
    n = int(vertex.shape[0])
    check(N.lib().fu_mesh_write_ply(os.fsencode(path), _p(vertex) if n else None, _p(normal) if n else None,
                                    _p(color) if n else None, n, int(binary), _s()), "fu_mesh_write_ply")


# --------------------------------------------------------------------- map (Mapping)
def map_params(**kw):
    p = N.MapParams()
    N.lib().fu_map_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError("unknown map parameter %r" % k)
        setattr(p, k, v)
    return p


class Mapping:
    """Voxel-hash map: Mapping (Mapping/Mapping.h) over DeviceMap (Mapping/DeviceMap.h)."""

    def __init__(self, **params):
        _dev()
        self.params = map_params(**params)
        h = C.c_void_p()
        check(N.lib().fu_map_create(C.byref(self.params), C.byref(h)), "fu_map_create")
        self.handle = h

    def close(self):
        if self.handle:
            N.lib().fu_map_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def Reset(self):
        check(N.lib().fu_reset_map(self.handle, _s()), "fu_reset_map")

    def stats(self):
        st = N.MapStats()
        check(N.lib().fu_map_stats_get(self.handle, C.byref(st), _s()), "fu_map_stats_get")
        return {f: getattr(st, f) for f, _ in N.MapStats._fields_}

    def FuseMapColor(self, depth, rgb, nmap, pose, K):
        rows, cols = depth.shape
        n = C.c_uint32()
        check(N.lib().fu_fuse_map_color(self.handle, _p(depth), _p(rgb), _p(nmap), cols, rows,
                                        C.byref(pose_f(pose)), K[0], K[1], K[2], K[3], self.params.depth_max,
                                        self.params.depth_min, C.byref(n), _s()), "fu_fuse_map_color")
        return n.value

    def CheckBlockVisibility(self, pose, cols, rows, K, depth_max, depth_min):
        n = C.c_uint32()
        check(N.lib().fu_check_block_visibility(self.handle, C.byref(pose_f(pose)), cols, rows, K[0], K[1], K[2],
                                                K[3], depth_max, depth_min, C.byref(n), _s()),
              "fu_check_block_visibility")
        return n.value

    def CreateRenderingBlocks(self, zcols, zrows, pose, n_visible, K, depth_max=None, depth_min=None):
        dmax = self.params.depth_max if depth_max is None else depth_max
        dmin = self.params.depth_min if depth_min is None else depth_min
        zmin = torch.empty((zrows, zcols), dtype=torch.float32, device="cuda")
        zmax = torch.empty_like(zmin)
        nf = C.c_uint32()
        rendered = C.c_int32()
        check(N.lib().fu_create_rendering_blocks(self.handle, _p(zmin), _p(zmax), zcols, zrows, dmax, dmin,
                                                 C.byref(pose_f(pose)), n_visible, K[0], K[1], K[2], K[3],
                                                 C.byref(nf), C.byref(rendered), _s()), "fu_create_rendering_blocks")
        return zmin, zmax, nf.value, bool(rendered.value)

    def Raycast(self, cols, rows, zmin, zmax, pose, K):
        vmap = torch.empty((rows, cols, 4), dtype=torch.float32, device="cuda")
        nmap = torch.empty_like(vmap)
        zrows, zcols = zmin.shape
        invfx = float(np.float32(1.0 / np.float64(np.float32(K[0]))))
        invfy = float(np.float32(1.0 / np.float64(np.float32(K[1]))))
        check(N.lib().fu_raycast(self.handle, _p(vmap), _p(nmap), cols, rows, _p(zmin), _p(zmax), zcols, zrows,
                                 C.byref(pose_f(pose)), invfx, invfy, K[2], K[3], _s()), "fu_raycast")
        return vmap, nmap

    def RayTrace(self, n_visible, pose, cols, rows, K):
        """Mapping::RayTrace (Mapping.cc:98-109): z-range tiles then Raycast (None if nothing to render)."""
        zmin, zmax, nf, rendered = self.CreateRenderingBlocks((cols + 7) // 8, (rows + 7) // 8, pose, n_visible, K)
        if not rendered:
            return None
        return self.Raycast(cols, rows, zmin, zmax, pose, K)

    def MeshScene(self, max_triangles=2_000_000, with_meta=False):
        """MeshScene + Mapping::CreateModel: (vertex (n,3,3), normal (n,3,3), color (n,3,3), meta)."""
        vert = torch.empty((max_triangles, 9), dtype=torch.float32, device="cuda")
        norm = torch.empty_like(vert)
        col = torch.empty((max_triangles, 9), dtype=torch.uint8, device="cuda")
        meta = torch.empty((max_triangles, 4), dtype=torch.int32, device="cuda") if with_meta else None
        n = C.c_uint32()
        occ = C.c_uint32()
        check(N.lib().fu_mesh_scene(self.handle, _p(vert), _p(norm), _p(col), _p(meta), max_triangles, C.byref(n),
                                    C.byref(occ), _s()), "fu_mesh_scene")
        k = n.value
        out = (vert[:k].reshape(k, 3, 3), norm[:k].reshape(k, 3, 3), col[:k].reshape(k, 3, 3))
        return out + ((meta[:k] if meta is not None else None),)

    def save(self, path):
        """System::WriteMapToDisk (Core/System.cc:195-234) as a compact versioned snapshot"""
        check(N.lib().fu_map_save(self.handle, os.fsencode(path), _s()), "fu_map_save")

    def load(self, path):
        """System::ReadMapFromDisk (Core/System.cc:236-271) into a map of the same parameters"""
        check(N.lib().fu_map_load(self.handle, os.fsencode(path), _s()), "fu_map_load")

    def export_blocks(self):
        n = C.c_uint32()
        check(N.lib().fu_map_export_blocks(self.handle, None, 0, C.byref(n), _s()), "fu_map_export_blocks")
        out = np.zeros((n.value, 3), dtype=np.int32)
        check(N.lib().fu_map_export_blocks(self.handle, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n), _s()),
              "fu_map_export_blocks")
        return out

    def export_visible(self):
        n = C.c_uint32()
        check(N.lib().fu_map_export_visible(self.handle, None, 0, C.byref(n), _s()), "fu_map_export_visible")
        out = np.zeros((n.value, 3), dtype=np.int32)
        check(N.lib().fu_map_export_visible(self.handle, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n), _s()),
              "fu_map_export_visible")
        return out

    def read_blocks(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.int32)
        n = xyz.shape[0]
        sdf = np.zeros((n, 512), dtype=np.float32)
        w = np.zeros((n, 512), dtype=np.uint16)
        rgb = np.zeros((n, 512, 3), dtype=np.uint8)
        check(N.lib().fu_map_read_blocks(self.handle, xyz.ctypes.data_as(C.c_void_p), n,
                                         sdf.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                                         rgb.ctypes.data_as(C.c_void_p), _s()), "fu_map_read_blocks")
        return sdf, w, rgb


# --------------------------------------------------------------------- tracker
class Tracker:
    """Tracker::ComputeSE3 in ICP-only mode, run on the device (Tracking/Tracking.cc:296-409)."""

    def __init__(self, cols, rows):
        _dev()
        h = C.c_void_p()
        check(N.lib().fu_tracker_create(cols, rows, C.byref(h)), "fu_tracker_create")
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                N.lib().fu_tracker_destroy(self.handle)
        except Exception:
            pass

    def set_last_pose(self, pose):
        arr = (C.c_double * 16)(*np.asarray(pose, dtype=np.float64).ravel())
        check(N.lib().fu_tracker_set_last_pose(self.handle, arr, _s()), "fu_tracker_set_last_pose")

    def ComputeSE3(self, next_vmaps, next_nmaps, last_vmaps, last_nmaps, K, iters=(10, 5, 4),
                   min_count=(2000, 1000, 100), thresh=1e-4):
        arr = lambda ts: (C.c_void_p * 3)(*[_addr(t) for t in ts])
        check(N.lib().fu_compute_se3(self.handle, arr(next_vmaps), arr(next_nmaps), arr(last_vmaps),
                                     arr(last_nmaps), (C.c_float * 4)(*K), (C.c_int32 * 3)(*iters),
                                     (C.c_int32 * 3)(*min_count), thresh, _s()), "fu_compute_se3")
        return self.result()

    def ComputeSE3RGBD(self, next_vmaps, next_nmaps, last_vmaps, last_nmaps, next_images, last_images, dIdx, dIdy, K,
                       iters=(10, 5, 4), min_count=(2000, 1000, 100), thresh=1e-4):
        """ComputeSE3(icpOnly=false): the photometric term blended into every iteration"""
        arr = lambda ts: (C.c_void_p * 3)(*[_addr(t) for t in ts])
        check(N.lib().fu_compute_se3_rgbd(self.handle, arr(next_vmaps), arr(next_nmaps), arr(last_vmaps),
                                          arr(last_nmaps), arr(next_images), arr(last_images), arr(dIdx), arr(dIdy),
                                          (C.c_float * 4)(*K), (C.c_int32 * 3)(*iters), (C.c_int32 * 3)(*min_count),
                                          thresh, _s()), "fu_compute_se3_rgbd")
        return self.result()

    def result(self):
        r = N.TrackResult()
        check(N.lib().fu_tracker_result(self.handle, C.byref(r), _s()), "fu_tracker_result")
        return {"pose": np.array(r.pose[:]).reshape(4, 4), "last_pose": np.array(r.last_pose[:]).reshape(4, 4),
                "x": np.array(r.x[:]), "A": np.array(r.A[:]).reshape(6, 6), "b": np.array(r.b[:]),
                "icp_error": r.icp_error, "count": r.count, "status": r.status, "iterations": r.iterations,
                "rgb_error": r.rgb_error, "rgb_count": r.rgb_count}


class IcpBatch:
    """fu_compute_se3_batch: ComputeSE3 for many trackers, each iteration one launch pair"""

    def __init__(self, capacity=256):
        _dev()
        h = C.c_void_p()
        check(N.lib().fu_icp_batch_create(int(capacity), C.byref(h)), "fu_icp_batch_create")
        self.handle = h

    def __del__(self):
        try:
            if self.handle:
                N.lib().fu_icp_batch_destroy(self.handle)
        except Exception:
            pass

    def ComputeSE3(self, trackers, next_vmaps, next_nmaps, last_vmaps, last_nmaps, Ks, iters=(10, 5, 4),
                   min_count=(2000, 1000, 100), thresh=1e-4):
        """per pair: 3-level lists of device maps and (fx, fy, cx, cy); returns each tracker's result"""
        n = len(trackers)
        arr = lambda lists: (C.c_void_p * (3 * n))(*[_addr(t) for ts in lists for t in ts])
        Kf = (C.c_float * (4 * n))(*[float(v) for K in Ks for v in K])
        check(N.lib().fu_compute_se3_batch(self.handle, n, (C.c_void_p * n)(*[t.handle.value for t in trackers]),
                                           arr(next_vmaps), arr(next_nmaps), arr(last_vmaps), arr(last_nmaps), Kf,
                                           (C.c_int32 * 3)(*iters), (C.c_int32 * 3)(*min_count), thresh, _s()),
              "fu_compute_se3_batch")
        return [t.result() for t in trackers]


# --------------------------------------------------------------------- system
def system_params(cols=640, rows=480, K=(525.0, 525.0, 319.5, 239.5), plain=False, use_graph=False, map=None,
                  **kw):
    p = N.SystemParams()
    N.lib().fu_system_params_default(C.byref(p))
    p.cols, p.rows = cols, rows
    p.fx, p.fy, p.cx, p.cy = K
    p.plain = int(plain)
    p.use_graph = int(use_graph)
    for k, v in kw.items():
        if k in ("iters", "min_count"):
            for i in range(3):
                getattr(p, k)[i] = v[i]
        else:
            setattr(p, k, v)
    if map:
        for k, v in map.items():
            setattr(p.map, k, v)
    return p


class System:
    """System::GrabImage (Core/System.cc:92-138): ICP only, or with the photometric term
    (params.use_rgb, the reference's default TrackFrame)."""

    def __init__(self, params):
        _dev()
        self.params = params
        h = C.c_void_p()
        check(N.lib().fu_system_create(C.byref(params), C.byref(h)), "fu_system_create")
        self.handle = h
        b = N.SystemBuffers()
        check(N.lib().fu_system_get_buffers(h, C.byref(b)), "fu_system_get_buffers")
        self.buffers = b

    def close(self):
        if self.handle:
            N.lib().fu_system_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(N.lib().fu_system_reset(self.handle, _s()), "fu_system_reset")

    def map_handle(self):
        return N.lib().fu_system_map(self.handle)

    def GrabImage(self, depth, rgb):
        """depth: device int16 (H, W) holding u16 values; rgb: device uint8 (H, W, 3)."""
        check(N.lib().fu_system_grab_image_from(self.handle, _p(depth), _p(rgb), _s()), "fu_system_grab_image")

    def GrabImageHost(self, depth, rgb):
        """host frame (uint16 (H, W), uint8 (H, W, 3); numpy or pinned torch CPU tensors),
        uploaded into the system's input slots on the current stream, then the frame"""
        for a in (depth, rgb):
            if isinstance(a, torch.Tensor):
                if a.is_cuda or not a.is_contiguous():
                    raise N.FusionError("GrabImageHost: contiguous host tensors expected")
            elif not (isinstance(a, np.ndarray) and a.flags.c_contiguous):
                raise N.FusionError("GrabImageHost: contiguous host arrays expected")
        ptr = lambda a: C.c_void_p(a.data_ptr() if isinstance(a, torch.Tensor) else a.ctypes.data)
        check(N.lib().fu_system_grab_image_host(self.handle, ptr(depth), ptr(rgb), _s()), "fu_system_grab_image_host")

    def relocalise(self, pose):
        """a lost system's relocalisation at `pose` (found by the caller, e.g. KeyMap.relocalise
        on the lost frame's keypoints): the next frame is raycast there, tracking resumes after
        it and fusion 10 frames later"""
        arr = (C.c_double * 16)(*np.asarray(pose, dtype=np.float64).ravel())
        check(N.lib().fu_system_relocalise(self.handle, arr, _s()), "fu_system_relocalise")

    def validate_poses(self, poses, iters_reloc=(3, 3, 2), thresh_reloc=1e-3):
        """Tracker::ValidatePose (Tracking.cc:636-682) on a lost system's current frame:
        (ok, accepted per candidate: -1 skipped / 0 / 1, the candidate refined or -1)"""
        P = np.ascontiguousarray(np.asarray(poses, dtype=np.float64).reshape(-1, 16))
        n = len(P)
        acc = np.zeros(max(n, 1), np.int32)
        chosen = np.zeros(1, np.int32)
        ok = C.c_int()
        check(N.lib().fu_system_validate_poses(self.handle, P.ctypes.data_as(N.DP), n, (C.c_int * 3)(*iters_reloc),
                                               float(thresh_reloc), acc.ctypes.data_as(C.c_void_p),
                                               chosen.ctypes.data_as(C.c_void_p), C.byref(ok), _s()),
              "fu_system_validate_poses")
        return bool(ok.value), acc[:n].copy(), int(chosen[0])

    def relocalise_graph(self, keymap, frame_desc, frame_pos, frame_normal, iterations=100, seed=0):
        """Tracker::Relocalise with graph matching on: (ok, number of candidate poses)"""
        pos = np.ascontiguousarray(frame_pos, dtype=np.float32)
        nrm = np.ascontiguousarray(frame_normal, dtype=np.float32)
        n = int(frame_desc.shape[0])
        nc, ok = C.c_int(), C.c_int()
        check(N.lib().fu_system_relocalise_graph(self.handle, keymap.handle, _p(frame_desc),
                                                 pos.ctypes.data_as(C.c_void_p), nrm.ctypes.data_as(C.c_void_p), n,
                                                 int(iterations), int(seed), C.byref(nc), C.byref(ok), _s()),
              "fu_system_relocalise_graph")
        return bool(ok.value), nc.value

    def set_mapping(self, enabled):
        """Tracker::mappingDisabled: with mapping off frames are tracked and rendered, never fused"""
        check(N.lib().fu_system_set_mapping(self.handle, int(bool(enabled)), _s()), "fu_system_set_mapping")

    def graph_counts(self):
        """(graphs captured, graph launches) of this system"""
        c, n = C.c_uint32(), C.c_uint32()
        check(N.lib().fu_system_graph_counts(self.handle, C.byref(c), C.byref(n)), "fu_system_graph_counts")
        return c.value, n.value

    def result(self):
        r = N.FrameResult()
        check(N.lib().fu_system_result(self.handle, C.byref(r), _s()), "fu_system_result")
        d = {f: getattr(r, f) for f, _ in N.FrameResult._fields_ if f != "pose"}
        d["pose"] = np.array(r.pose[:]).reshape(4, 4)
        return d

    def save(self, path):
        """session snapshot: map, tracking state, model maps (and last intensity pyramid)"""
        check(N.lib().fu_system_save(self.handle, os.fsencode(path), _s()), "fu_system_save")

    def load(self, path):
        check(N.lib().fu_system_load(self.handle, os.fsencode(path), _s()), "fu_system_load")

    def set_profiling(self, enable=True):
        check(N.lib().fu_system_set_profiling(self.handle, int(enable)), "fu_system_set_profiling")

    def stage_times(self):
        t = N.StageTimes()
        check(N.lib().fu_system_get_stage_times(self.handle, C.byref(t)), "fu_system_get_stage_times")
        return _times(t)

    def view(self, name, level=0):
        """a torch view of one of the system's device maps"""
        b = self.buffers
        rows, cols = self.params.rows >> level, self.params.cols >> level
        ptr = getattr(b, name)
        if isinstance(ptr, int) or ptr is None:
            addr = ptr
        else:
            addr = ptr[level]
        if name in ("image", "last_image", "dIdx", "dIdy"):
            if not addr:
                raise N.FusionError("%s: the system tracks without the photometric term" % name)
            typestr = "|u1" if name in ("image", "last_image") else "<i2"
            return torch.as_tensor(_DevBuf(addr, (rows, cols), typestr), device="cuda")
        shape = {"raw": (rows, cols), "depth": (rows, cols)}.get(name, (rows, cols, 4))
        n = int(np.prod(shape))
        return _wrap(addr, n, shape)

    def mapping(self):
        """a Mapping view over the system's map (not owning)"""
        m = Mapping.__new__(Mapping)
        m.handle = C.c_void_p(self.map_handle())
        m.params = self.params.map
        m.close = lambda: None
        return m


TUNING_FIELDS = [f for f, _ in N.Tuning._fields_]


def default_tuning():
    """the library's default launch geometry of the batched stages, as a dict (fu_tuning)"""
    t = N.Tuning()
    N.lib().fu_tuning_default(C.byref(t))
    return {f: int(getattr(t, f)) for f in TUNING_FIELDS}


def get_tuning():
    t = N.Tuning()
    check(N.lib().fu_get_tuning(C.byref(t)), "fu_get_tuning")
    return {f: int(getattr(t, f)) for f in TUNING_FIELDS}


def set_tuning(**kw):
    """change fields of the process-wide launch geometry (fu_set_tuning); returns the new one.
    Results do not depend on it, only the speed of batched steps."""
    cur = get_tuning()
    for k, v in kw.items():
        if k not in cur:
            raise KeyError("no tuning field %r (fields: %s)" % (k, ", ".join(TUNING_FIELDS)))
        cur[k] = int(v)
    t = N.Tuning(**cur)
    check(N.lib().fu_set_tuning(C.byref(t)), "fu_set_tuning")
    return cur


class Batch:
    """A batch context: one frame of each of many attached Systems per call, every stage one
    launch for all of them (no reference counterpart; the reference runs one System)."""

    MAX = 64

    def __init__(self, capacity=256):
        _dev()
        h = C.c_void_p()
        check(N.lib().fu_batch_create(int(capacity), C.byref(h)), "fu_batch_create")
        self.handle = h
        self.systems = []

    def attach(self, system):
        check(N.lib().fu_batch_attach(self.handle, system.handle), "fu_batch_attach")
        self.systems.append(system)

    def detach(self, system):
        check(N.lib().fu_batch_detach(self.handle, system.handle), "fu_batch_detach")
        self.systems.remove(system)

    def close(self):
        if self.handle:
            N.lib().fu_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_profiling(self, enable=True):
        check(N.lib().fu_batch_set_profiling(self.handle, int(enable)), "fu_batch_set_profiling")

    def stage_times(self):
        t = N.StageTimes()
        check(N.lib().fu_batch_get_stage_times(self.handle, C.byref(t)), "fu_batch_get_stage_times")
        return _times(t)

    def set_graph(self, enable=True):
        """run steps as captured hipGraphs (captured on a step membership's second use)"""
        check(N.lib().fu_batch_set_graph(self.handle, int(enable)), "fu_batch_set_graph")

    def graph_counts(self):
        """(graphs captured, graph launches) of this context"""
        c, r = C.c_uint32(), C.c_uint32()
        check(N.lib().fu_batch_graph_counts(self.handle, C.byref(c), C.byref(r)), "fu_batch_graph_counts")
        return c.value, r.value

    def grab(self, systems, depths, rgbs):
        """one frame per system: device depth (H, W) u16-in-int16 and rgb (H, W, 3) u8 tensors"""
        n = len(systems)
        arr = C.c_void_p * n
        sy = arr(*[s.handle.value for s in systems])
        dp = arr(*[_addr(d) for d in depths])
        cp = arr(*[_addr(c) for c in rgbs])
        check(N.lib().fu_batch_grab(self.handle, n, sy, dp, cp, _s()), "fu_batch_grab")


class BudgetExceeded(N.FusionError):
    """a session was not admitted: the server's voxel-block budget is in use"""


class Server:
    """Continuous-batching server of many reconstruction sessions (fu_server_*)."""

    def __init__(self, max_batch=32, max_inflight=2, lanes=2, block_budget=0, use_graph=False):
        _dev()
        p = N.ServerParams()
        N.lib().fu_server_params_default(C.byref(p))
        p.max_batch, p.max_inflight, p.lanes, p.block_budget = max_batch, max_inflight, lanes, int(block_budget)
        h = C.c_void_p()
        check(N.lib().fu_server_create(C.byref(p), C.byref(h)), "fu_server_create")
        self.handle = h
        self.params = {}
        if use_graph:
            self.set_graph(True)

    def set_graph(self, enable=True):
        """run every lane's steps as captured hipGraphs (fu_server_set_graph)"""
        check(N.lib().fu_server_set_graph(self.handle, int(enable)), "fu_server_set_graph")

    def graph_counts(self):
        """(graphs captured, graph launches) summed over the lanes"""
        c, r = C.c_uint32(), C.c_uint32()
        check(N.lib().fu_server_graph_counts(self.handle, C.byref(c), C.byref(r)), "fu_server_graph_counts")
        return c.value, r.value

    def close(self):
        if self.handle:
            N.lib().fu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open(self, params):
        sid = C.c_int()
        st = N.lib().fu_server_open(self.handle, C.byref(params), C.byref(sid))
        if st == N.FU_ERR_BUDGET:
            raise BudgetExceeded(N.lib().fu_last_error().decode(errors="replace"))
        check(st, "fu_server_open")
        self.params[sid.value] = params
        return sid.value

    def close_session(self, sid):
        check(N.lib().fu_server_close(self.handle, int(sid)), "fu_server_close")
        del self.params[sid]

    def system(self, sid):
        """a non-owning System view of a session (result(), mapping(), view())"""
        s = System.__new__(System)
        s.handle = C.c_void_p(N.lib().fu_server_system(self.handle, int(sid)))
        s.params = self.params[sid]
        b = N.SystemBuffers()
        check(N.lib().fu_system_get_buffers(s.handle, C.byref(b)), "fu_system_get_buffers")
        s.buffers = b
        s.close = lambda: None
        return s

    def now_ms(self):
        return N.lib().fu_server_now_ms(self.handle)

    def set_profiling(self, enable=True):
        check(N.lib().fu_server_set_profiling(self.handle, int(enable)), "fu_server_set_profiling")

    def stage_times(self):
        t = N.StageTimes()
        check(N.lib().fu_server_get_stage_times(self.handle, C.byref(t)), "fu_server_get_stage_times")
        return _times(t)

    def submit(self, sid, depth, rgb, tag=0, t_submit_ms=-1.0):
        check(N.lib().fu_server_submit(self.handle, int(sid), _p(depth), _p(rgb), int(tag), float(t_submit_ms)),
              "fu_server_submit")

    def submit_host(self, sid, depth, rgb, tag=0, t_submit_ms=-1.0):
        """queue a frame held in (pinned) host tensors: uploaded inside the step that takes it"""
        if depth.is_cuda or rgb.is_cuda:
            raise N.FusionError("submit_host: host tensors expected")
        check(N.lib().fu_server_submit_host(self.handle, int(sid), _p(depth), _p(rgb), int(tag), float(t_submit_ms)),
              "fu_server_submit_host")

    def submit_mesh(self, sid, vertex, normal, color, max_triangles, tag=0, t_submit_ms=-1.0):
        """queue a mesh extraction (device buffers: float (T, 9), float (T, 9), uint8 (T, 9))"""
        check(N.lib().fu_server_submit_mesh(self.handle, int(sid), _p(vertex), _p(normal), _p(color),
                                            int(max_triangles), int(tag), float(t_submit_ms)), "fu_server_submit_mesh")

    def poll(self, cap=256):
        out = (N.Completion * cap)()
        n = C.c_int()
        check(N.lib().fu_server_poll(self.handle, out, cap, C.byref(n)), "fu_server_poll")
        return [_completion(out[i]) for i in range(n.value)]

    def pending(self):
        n = C.c_int()
        check(N.lib().fu_server_pending(self.handle, C.byref(n)), "fu_server_pending")
        return n.value

    def drain(self):
        out = []
        while self.pending():
            out += self.poll()
        return out

    def run_trace(self, frames, closed_loop=False):
        """frames: list of (session, depth, rgb, arrival_ms, tag) frame requests (device tensors,
        or pinned host tensors uploaded inside the step), or
        (session, (vertex, normal, color), None, arrival_ms, tag) mesh requests (buffers of
        shape (T, 9)); returns completions in completion order"""
        n = len(frames)
        tr = (N.TraceFrame * n)()
        for i, (sid, d, c, t, tag) in enumerate(frames):
            tr[i].session, tr[i].arrival_ms, tr[i].tag = sid, t, tag
            if c is None:
                v, nm, col = d
                tr[i].kind, tr[i].max_triangles = 1, v.shape[0]
                tr[i].vertex, tr[i].normal, tr[i].color = _addr(v), _addr(nm), _addr(col)
            else:
                tr[i].depth, tr[i].rgb = _addr(d), _addr(c)
                tr[i].host = 0 if d.is_cuda else 1  # host tensors are uploaded inside the step
        out = (N.Completion * n)()
        check(N.lib().fu_server_run_trace(self.handle, tr, n, int(closed_loop), out), "fu_server_run_trace")
        return [_completion(out[i]) for i in range(n)]


def _completion(c):
    d = {f: getattr(c, f) for f, _ in N.Completion._fields_ if f != "pose"}
    d["pose"] = np.array(c.pose[:]).reshape(4, 4)
    return d


def _times(t):
    return {"ms": dict(zip(N.STAGES, t.ms[:])), "frames": t.frames, "n_visible": t.n_visible,
            "n_new": t.n_new, "icp_iterations": t.icp_iterations, "icp_pixels": t.icp_pixels, "steps": t.steps}


class _DevBuf:
    """__cuda_array_interface__ over a device buffer owned by the library (zero-copy view)"""

    def __init__(self, addr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (int(addr), False),
                                         "version": 2, "strides": None}


def _wrap(addr, n, shape):
    del n
    return torch.as_tensor(_DevBuf(addr, shape, "<f4"), device="cuda")


# --------------------------------------------------------------------- keypoint map
# fu_surf records (SURF, Mapping/DeviceMap.h:99-108): 288 bytes each; on the device a uint8
# tensor of shape (n, 288)
SURF_DTYPE = np.dtype([("valid", np.uint8), ("pad", np.uint8, 3), ("pos", np.float32, 3), ("normal", np.float32, 4),
                       ("descriptor", np.float32, 64)])


def surf_to_device(records):
    a = np.ascontiguousarray(records, dtype=SURF_DTYPE)
    return torch.from_numpy(a.view(np.uint8).reshape(len(a), SURF_DTYPE.itemsize).copy()).to("cuda")


def surf_from_device(t):
    return t.cpu().numpy().reshape(-1).copy().view(SURF_DTYPE)


class KeyMap:
    """KeyMap (Mapping/DeviceMap.h:163-185) with ResetKeyPoints, InsertKeyPoints and
    CollectKeyPoints (Mapping/RenderScene.h:5-11) on the device"""

    def __init__(self, grid_size=0.01, max_keys=100000, n_buckets=5):
        _dev()
        p = N.KeyMapParams()
        N.lib().fu_keymap_params_default(C.byref(p))
        p.grid_size, p.max_keys, p.n_buckets = grid_size, max_keys, n_buckets
        self.n = max_keys * n_buckets
        h = C.c_void_p()
        check(N.lib().fu_keymap_create(C.byref(p), C.byref(h)), "fu_keymap_create")
        self.handle = h

    def close(self):
        if self.handle:
            N.lib().fu_keymap_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ResetKeyPoints(self):
        check(N.lib().fu_reset_key_points(self.handle, _s()), "fu_reset_key_points")

    def InsertKeyPoints(self, keys, index):
        """keys: device uint8 (n, 288) records (positions may be rewritten), index: device int32 (n,)"""
        check(N.lib().fu_insert_key_points(self.handle, _p(keys), _p(index), int(keys.shape[0]), _s()),
              "fu_insert_key_points")

    def CollectKeyPoints(self, cap):
        out = torch.zeros((max(cap, 1), SURF_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        n = C.c_uint32()
        check(N.lib().fu_collect_key_points(self.handle, _p(out), int(cap), C.byref(n), _s()), "fu_collect_key_points")
        return out[:cap], n.value

    def read(self):
        out = np.zeros(self.n, SURF_DTYPE)
        check(N.lib().fu_keymap_read(self.handle, out.ctypes.data_as(C.c_void_p), _s()), "fu_keymap_read")
        return out

    def relocalise(self, frame_desc, frame_pos, iterations=100, seed=0):
        """Tracker::Relocalise (Tracking.cc:411-467, graph matching off): frame_desc a device
        float32 (n, 64) tensor, frame_pos the lifted keypoints (n, 3); returns (pose or None,
        number of pairs that pass the ratio test)"""
        pos = np.ascontiguousarray(frame_pos, dtype=np.float64)
        n = int(frame_desc.shape[0])
        pose = np.zeros(16)
        npairs, ok = C.c_int(), C.c_int()
        check(N.lib().fu_relocalise_keys(self.handle, _p(frame_desc), pos.ctypes.data_as(N.DP), n, int(iterations),
                                         int(seed), pose.ctypes.data_as(N.DP), C.byref(npairs), C.byref(ok), _s()),
              "fu_relocalise_keys")
        return (pose.reshape(4, 4) if ok.value else None), npairs.value


    def graph_candidates(self, frame_desc, frame_pos, frame_normal, iterations=100, seed=0):
        """Tracker::Relocalise with graph matching on, through GenerateGraph's PoseEstimate
        calls (Tracking.cc:432-453, 469-629): (candidate camera poses (k, 4, 4), size of the
        match list)"""
        pos = np.ascontiguousarray(frame_pos, dtype=np.float32)
        nrm = np.ascontiguousarray(frame_normal, dtype=np.float32)
        n = int(frame_desc.shape[0])
        poses = np.zeros(5 * 16)
        npose, npairs = C.c_int(), C.c_int()
        check(N.lib().fu_relocalise_graph_candidates(self.handle, _p(frame_desc), pos.ctypes.data_as(C.c_void_p),
                                                     nrm.ctypes.data_as(C.c_void_p), n, int(iterations), int(seed),
                                                     poses.ctypes.data_as(N.DP), C.byref(npose), C.byref(npairs),
                                                     _s()), "fu_relocalise_graph_candidates")
        return poses.reshape(5, 4, 4)[: npose.value].copy(), npairs.value


def select_subgraphs(matrix, n_lists=10, max_select=400, min_score=0.1):
    """GenerateGraph's sub-graph selection (Tracking.cc:515-596) over a device float32 (n, n)
    consistency matrix: (n_selected (n_lists,), refined lists, rank sums (n,), rank order (n,)),
    host arrays"""
    n = int(matrix.shape[0])
    rank = torch.empty((max(n, 1),), dtype=torch.float32, device="cuda")
    order = torch.empty((max(n, 1),), dtype=torch.int32, device="cuda")
    nsel = torch.empty((n_lists,), dtype=torch.int32, device="cuda")
    nref = torch.empty((n_lists,), dtype=torch.int32, device="cuda")
    ref = torch.empty((n_lists, max_select), dtype=torch.int32, device="cuda")
    check(N.lib().fu_select_subgraphs(_p(matrix) if n else None, n, int(n_lists), int(max_select), float(min_score),
                                      _p(rank), _p(order), _p(nsel), _p(nref), _p(ref), _s()), "fu_select_subgraphs")
    nsel, nref, ref = nsel.cpu().numpy(), nref.cpu().numpy(), ref.cpu().numpy()
    return nsel, [ref[i, : nref[i]].copy() for i in range(n_lists)], rank.cpu().numpy()[:n], order.cpu().numpy()[:n]


def BuildAdjacencyMatrix(frame_keys, map_keys, dist):
    """BuildAdjacencyMatrix (Tracking/Reduction.h:85-88): the (n, n) consistency scores of n
    matched pairs (device records and float32 distances)"""
    n = int(frame_keys.shape[0])
    out = torch.empty((n, n), dtype=torch.float32, device="cuda")
    check(N.lib().fu_build_adjacency_matrix(_p(frame_keys), _p(map_keys), _p(dist), n, _p(out), _s()),
          "fu_build_adjacency_matrix")
    return out


def lift_keypoints(xy, depth, nmap, K, depth_cutoff=3.0):
    """Frame::ExtractKeyPoints' test and back-projection (Core/Frame.cc:184-200) of keypoints
    xy (device float32 (n, 2)) on a level-0 depth (H, W) and normal map (H, W, 4): (pos (n, 3),
    normal (n, 4), ok (n,) uint8) device tensors"""
    n = int(xy.shape[0])
    rows, cols = int(depth.shape[0]), int(depth.shape[1])
    pos = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    nrm = torch.empty((n, 4), dtype=torch.float32, device="cuda")
    ok = torch.empty((n,), dtype=torch.uint8, device="cuda")
    if n:
        check(N.lib().fu_lift_keypoints(_p(xy), n, _p(depth), _p(nmap), cols, rows, float(K[0]), float(K[1]),
                                        float(K[2]), float(K[3]), float(depth_cutoff), _p(pos), _p(nrm), _p(ok),
                                        _s()), "fu_lift_keypoints")
    return pos, nrm, ok


def match_knn2(query, train):
    """the two nearest train descriptors of every query descriptor (device float32 (n, 64)):
    (idx (n, 2) int32, dist (n, 2) float32) device tensors"""
    nq, nt = int(query.shape[0]), (0 if train is None else int(train.shape[0]))
    idx = torch.empty((nq, 2), dtype=torch.int32, device="cuda")
    dist = torch.empty((nq, 2), dtype=torch.float32, device="cuda")
    if nq:
        check(N.lib().fu_match_knn2(_p(query), nq, _p(train) if nt else None, nt, _p(idx), _p(dist), _s()),
              "fu_match_knn2")
    return idx, dist


def pose_estimate(src, ref, iterations=100, seed=0):
    """Solver::PoseEstimate (Optimization/Solver.cc): (T 4x4 mapping ref to src or None, inlier
    mask) for (n, 3) point arrays"""
    src = np.ascontiguousarray(src, dtype=np.float64)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    n = int(src.shape[0])
    T = np.zeros(16)
    inl = np.zeros(max(n, 1), np.uint8)
    ok = C.c_int()
    check(N.lib().fu_pose_estimate(src.ctypes.data_as(N.DP), ref.ctypes.data_as(N.DP), n, int(iterations), int(seed),
                                   T.ctypes.data_as(N.DP), inl.ctypes.data_as(C.c_void_p), C.byref(ok)),
          "fu_pose_estimate")
    return (T.reshape(4, 4) if ok.value else None), inl[:n].astype(bool)


# --------------------------------------------------------------------- synthetic scenes
class Scene:
    """Seeded synthetic scene (include/fusion_scene.h)."""

    def __init__(self, kind=0, size=(4.0, 4.0, 3.0), seed=0):
        h = C.c_void_p()
        if N.scene_lib().fu_scene_create(kind, size[0], size[1], size[2], seed, C.byref(h)) != 0:
            raise N.FusionError("fu_scene_create failed")
        self.handle = h

    def __del__(self):
        try:
            N.scene_lib().fu_scene_destroy(self.handle)
        except Exception:
            pass

    def pose(self, frame):
        p = (C.c_double * 16)()
        N.scene_lib().fu_scene_pose(self.handle, frame, p)
        return np.array(p[:]).reshape(4, 4)

    def render(self, frame, cols=640, rows=480, K=(525.0, 525.0, 319.5, 239.5), depth_scale=5000.0, noise=True):
        depth = np.zeros((rows, cols), dtype=np.uint16)
        rgb = np.zeros((rows, cols, 3), dtype=np.uint8)
        N.scene_lib().fu_scene_render(self.handle, frame, cols, rows, K[0], K[1], K[2], K[3], depth_scale,
                                      int(noise), depth.ctypes.data_as(C.c_void_p), rgb.ctypes.data_as(C.c_void_p))
        return depth, rgb

    def render_pose(self, pose, cols=640, rows=480, K=(525.0, 525.0, 319.5, 239.5), depth_scale=5000.0,
                    noise=False, noise_seed=0):
        depth = np.zeros((rows, cols), dtype=np.uint16)
        rgb = np.zeros((rows, cols, 3), dtype=np.uint8)
        arr = (C.c_double * 16)(*np.asarray(pose, dtype=np.float64).ravel())
        N.scene_lib().fu_scene_render_pose(self.handle, arr, cols, rows, K[0], K[1], K[2], K[3], depth_scale,
                                           int(noise), noise_seed, depth.ctypes.data_as(C.c_void_p),
                                           rgb.ctypes.data_as(C.c_void_p))
        return depth, rgb

    def sdf(self, x, y, z):
        return N.scene_lib().fu_scene_sdf(self.handle, x, y, z)


def to_device(depth_u16, rgb_u8):
    """host numpy frame -> device tensors (depth as int16 holding the u16 bits)"""
    d = torch.from_numpy(np.ascontiguousarray(depth_u16).view(np.int16)).to("cuda")
    c = torch.from_numpy(np.ascontiguousarray(rgb_u8)).to("cuda")
    return d, c
