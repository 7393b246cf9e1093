"""RGB-D frame sources: recorded sequences on disk and synthetic scenes.

Reference: MainTum.cc (the apt0 loop :62-96 and load_tum_dataset :100-134) and the OpenNI
camera (Core/Camera.cc:117-158, PrimeSense::FetchFrame).  A source yields host frames
(depth u16 (rows, cols), colour u8 (rows, cols, 3) in RGB order) with its camera constants;
`System.GrabImage` takes them after `api.to_device`.  Decoding is host work (Pillow's PNG and
JPEG decoders stand in for OpenCV's `imread`); nothing here touches the GPU.

Also: the TUM RGB-D trajectory format (timestamp tx ty tz qx qy qz qw) and the absolute
trajectory error against ground truth, for validating tracking on recorded data.
"""
import os

import numpy as np

TUM_FR1 = dict(K=(517.3, 516.5, 318.6, 255.3), depth_scale=5000.0, depth_cutoff=3.0)  # MainTum.cc:37-47
APT0 = dict(K=(583.0, 583.0, 320.0, 240.0), depth_scale=1000.0, depth_cutoff=3.0)  # MainTum.cc:62-71


def _pil():
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - the image ships Pillow
        raise RuntimeError("reading PNG/JPEG frames needs Pillow") from e
    return Image


def read_depth_png(path):
    """a 16-bit single-channel PNG as u16 (cv::imread(..., IMREAD_UNCHANGED))"""
    Image = _pil()
    with Image.open(path) as im:
        a = np.array(im)
    if a.dtype != np.uint16:
        if a.dtype in (np.int32, np.uint32, np.int64) and a.min() >= 0 and a.max() < 65536:
            a = a.astype(np.uint16)  # Pillow's mode "I" for 16-bit greyscale
        else:
            raise ValueError("%s: not a 16-bit depth image (%s)" % (path, a.dtype))
    if a.ndim != 2:
        raise ValueError("%s: depth must have one channel" % path)
    return np.ascontiguousarray(a)


def read_color(path):
    """an 8-bit colour image (PNG or JPEG) as RGB, as imread + cvtColor(BGR2RGB) give"""
    Image = _pil()
    with Image.open(path) as im:
        a = np.array(im.convert("RGB"))
    return np.ascontiguousarray(a, dtype=np.uint8)


def write_depth_png(path, depth_u16):
    Image = _pil()
    Image.fromarray(np.ascontiguousarray(depth_u16, dtype=np.uint16)).save(path)


def write_color_png(path, rgb_u8):
    Image = _pil()
    Image.fromarray(np.ascontiguousarray(rgb_u8, dtype=np.uint8), "RGB").save(path)


def tum_lists(path):
    """load_tum_dataset (MainTum.cc:100-134): depth.txt and rgb.txt, three header lines each,
    then the i-th depth line paired with the i-th colour line (by line, not by timestamp);
    returns [(timestamp of the depth line, depth path, colour path)].  The reference's loop
    tests end-of-file after appending, so when both lists end with a newline it appends the
    last pair a second time (with timestamp 0); that repeat is not reproduced."""
    root = path if path.endswith("/") else path + "/"

    def body(name):
        with open(root + name) as f:
            lines = f.read().split("\n")
        toks = " ".join(lines[3:]).split()  # `>>` skips every kind of whitespace
        return toks

    d = body("depth.txt")
    c = body("rgb.txt")
    n = min(len(d) // 2, len(c) // 2)
    return [(float(d[2 * i]), root + d[2 * i + 1], root + c[2 * i + 1]) for i in range(n)]


class TumSource:
    """a TUM RGB-D sequence directory (depth.txt, rgb.txt, depth/ and rgb/ PNGs)"""

    def __init__(self, path, K=TUM_FR1["K"], depth_scale=TUM_FR1["depth_scale"],
                 depth_cutoff=TUM_FR1["depth_cutoff"]):
        self.items = tum_lists(path)
        self.K, self.depth_scale, self.depth_cutoff = tuple(K), depth_scale, depth_cutoff
        d, _ = self.frame(0) if self.items else (np.zeros((480, 640), np.uint16), None)
        self.rows, self.cols = d.shape

    def __len__(self):
        return len(self.items)

    def frame(self, i):
        _, dp, cp = self.items[i]
        return read_depth_png(dp), read_color(cp)

    def timestamp(self, i):
        return self.items[i][0]

    def __iter__(self):
        for i in range(len(self)):
            yield self.frame(i)


class Apt0Source:
    """the apt0 layout of MainTum.cc:74-93: frame-%06d.depth.png and frame-%06d.color.jpg,
    numbered from 0 until the first missing pair"""

    def __init__(self, path, K=APT0["K"], depth_scale=APT0["depth_scale"], depth_cutoff=APT0["depth_cutoff"],
                 limit=None):
        self.root = path
        self.K, self.depth_scale, self.depth_cutoff = tuple(K), depth_scale, depth_cutoff
        n = 0
        while (limit is None or n < limit) and os.path.exists(self._paths(n)[0]) and os.path.exists(self._paths(n)[1]):
            n += 1
        self.n = n
        d, _ = self.frame(0) if n else (np.zeros((480, 640), np.uint16), None)
        self.rows, self.cols = d.shape

    def _paths(self, i):
        stem = os.path.join(self.root, "frame-%06d" % i)
        return stem + ".depth.png", stem + ".color.jpg"

    def __len__(self):
        return self.n

    def frame(self, i):
        dp, cp = self._paths(i)
        return read_depth_png(dp), read_color(cp)

    def timestamp(self, i):
        return float(i) / 30.0

    def __iter__(self):
        for i in range(len(self)):
            yield self.frame(i)


class SceneSource:
    """a seeded synthetic scene (fusion_amd.api.Scene) as a source"""

    def __init__(self, kind=0, size=(4.0, 4.0, 3.0), seed=0, frames=100, cols=640, rows=480,
                 K=(525.0, 525.0, 319.5, 239.5), noise=True):
        from .api import Scene

        self.scene = Scene(kind, size, seed)
        self.n, self.cols, self.rows, self.K, self.noise = frames, cols, rows, tuple(K), noise
        self.depth_scale, self.depth_cutoff = 5000.0, 3.0

    def __len__(self):
        return self.n

    def frame(self, i):
        return self.scene.render(i, self.cols, self.rows, self.K, self.depth_scale, self.noise)

    def timestamp(self, i):
        return float(i) / 30.0

    def ground_truth(self, i):
        return self.scene.pose(i)

    def __iter__(self):
        for i in range(len(self)):
            yield self.frame(i)


class StreamSource:
    """a live camera in the shape of PrimeSense::FetchFrame (Core/Camera.cc:117-141): `fetch`
    is any callable returning (depth u16, colour RGB u8) or None when the stream ends"""

    def __init__(self, fetch, cols=640, rows=480, K=(525.0, 525.0, 319.5, 239.5), depth_scale=1000.0,
                 depth_cutoff=3.0):
        self.fetch, self.cols, self.rows, self.K = fetch, cols, rows, tuple(K)
        self.depth_scale, self.depth_cutoff = depth_scale, depth_cutoff

    def __iter__(self):
        while True:
            f = self.fetch()
            if f is None:
                return
            d, c = f
            if d.shape != (self.rows, self.cols) or c.shape != (self.rows, self.cols, 3):
                raise ValueError("frame of shape %s / %s, expected %dx%d" % (d.shape, c.shape, self.cols, self.rows))
            yield np.ascontiguousarray(d, dtype=np.uint16), np.ascontiguousarray(c, dtype=np.uint8)


def prefetch(source, workers=8, depth=16):
    """the frames of an indexable source (`frame(i)`, `len`), decoded ahead by a thread pool
    (PNG and JPEG decoding release the interpreter lock), in order"""
    from concurrent.futures import ThreadPoolExecutor

    n = len(source)
    with ThreadPoolExecutor(max_workers=workers) as ex:
        futs = [ex.submit(source.frame, i) for i in range(min(depth, n))]
        for i in range(n):
            f = futs[i].result()
            if i + depth < n:
                futs.append(ex.submit(source.frame, i + depth))
            futs[i] = None
            yield f


def write_tum_sequence(path, frames, stamps):
    """a TUM-layout directory (depth.txt, rgb.txt with three comment lines, depth/ and rgb/
    PNGs) of host frames"""
    os.makedirs(os.path.join(path, "depth"), exist_ok=True)
    os.makedirs(os.path.join(path, "rgb"), exist_ok=True)
    dl = ["# depth maps", "# file: fusion_amd", "# timestamp filename"]
    cl = ["# color images", "# file: fusion_amd", "# timestamp filename"]
    for (d, c), t in zip(frames, stamps):
        name = "%.6f.png" % t
        write_depth_png(os.path.join(path, "depth", name), d)
        write_color_png(os.path.join(path, "rgb", name), c)
        dl.append("%.6f depth/%s" % (t, name))
        cl.append("%.6f rgb/%s" % (t, name))
    with open(os.path.join(path, "depth.txt"), "w") as f:
        f.write("\n".join(dl) + "\n")
    with open(os.path.join(path, "rgb.txt"), "w") as f:
        f.write("\n".join(cl) + "\n")


# --------------------------------------------------------------------- trajectories
def _quat_from_R(R):
    """unit quaternion (x, y, z, w) of a rotation matrix (w >= 0)"""
    R = np.asarray(R, dtype=np.float64)
    t = np.trace(R)
    if t > 0:
        s = np.sqrt(t + 1.0) * 2
        q = [(R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s, 0.25 * s]
    else:
        i = int(np.argmax(np.diag(R)))
        j, k = (i + 1) % 3, (i + 2) % 3
        s = np.sqrt(1.0 + R[i, i] - R[j, j] - R[k, k]) * 2
        q = [0.0, 0.0, 0.0, (R[k, j] - R[j, k]) / s]
        q[i] = 0.25 * s
        q[j] = (R[j, i] + R[i, j]) / s
        q[k] = (R[k, i] + R[i, k]) / s
    q = np.array(q)
    q /= np.linalg.norm(q)
    return q if q[3] >= 0 else -q


def _R_from_quat(q):
    x, y, z, w = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def write_tum_trajectory(path, stamps, poses):
    """one line per pose: timestamp tx ty tz qx qy qz qw (camera-to-world)"""
    with open(path, "w") as f:
        for t, P in zip(stamps, poses):
            P = np.asarray(P, dtype=np.float64)
            q = _quat_from_R(P[:3, :3])
            f.write("%.6f %.9f %.9f %.9f %.9f %.9f %.9f %.9f\n" % ((t,) + tuple(P[:3, 3]) + tuple(q)))


def read_tum_trajectory(path):
    stamps, poses = [], []
    with open(path) as f:
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            v = [float(x) for x in line.split()]
            P = np.eye(4)
            P[:3, :3] = _R_from_quat(np.array(v[4:8]))
            P[:3, 3] = v[1:4]
            stamps.append(v[0])
            poses.append(P)
    return np.array(stamps), poses


def ate_rmse(est, gt):
    """absolute trajectory error: RMSE of the positions after the least-squares rigid
    alignment of the estimated onto the ground-truth positions (Horn's closed form via SVD);
    returns (rmse, R, t)"""
    A = np.array([np.asarray(p)[:3, 3] for p in est], dtype=np.float64)
    B = np.array([np.asarray(p)[:3, 3] for p in gt], dtype=np.float64)
    ca, cb = A.mean(0), B.mean(0)
    H = (A - ca).T @ (B - cb)
    U, _, Vt = np.linalg.svd(H)
    D = np.eye(3)
    D[2, 2] = np.sign(np.linalg.det(Vt.T @ U.T))
    R = Vt.T @ D @ U.T
    t = cb - R @ ca
    err = (A @ R.T + t) - B
    return float(np.sqrt((err ** 2).sum(1).mean())), R, t
