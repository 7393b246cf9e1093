"""GPU parity of the depth pyramid (Tracking/Pyrdown.cu) against the CPU oracle.

Every kernel here is bit-exact with the oracle: same IEEE operation order, no FMA
contraction, and the same exp implementation in the bilateral filter.
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.helpers import mismatch, same_bits, scene_frames

pytestmark = pytest.mark.gpu

SIZES = [(640, 480), (320, 240), (1280, 720)]


def _pyr_gpu(d16, K, cutoff=3.0):
    d = torch.from_numpy(d16.view(np.int16)).cuda()
    raw, f0 = A.FilterDepth(d, 5000.0, cutoff)
    ds = [f0, A.PyrDownGauss(f0)]
    ds.append(A.PyrDownGauss(ds[1]))
    vs, ns = [], []
    for l in range(3):
        Kl = A.level_k(K, l)
        v = A.ComputeVMap(ds[l], *Kl, cutoff)
        vs.append(v)
        ns.append(A.ComputeNMap(v))
    torch.cuda.synchronize()
    c = lambda t: t.cpu().numpy()
    return c(raw), [c(x) for x in ds], [c(x) for x in vs], [c(x) for x in ns]


def _pyr_oracle(d16, K, cutoff=3.0):
    raw, f0 = O.filter_depth(d16, 5000.0, cutoff)
    ds = [f0, O.pyrdown_gauss(f0)]
    ds.append(O.pyrdown_gauss(ds[1]))
    vs, ns = [], []
    for l in range(3):
        v = O.compute_vmap(ds[l], *A.level_k(K, l), cutoff)
        vs.append(v)
        ns.append(O.compute_nmap(v))
    return raw, ds, vs, ns


@pytest.mark.parametrize("cols,rows", SIZES)
def test_pyramid_bit_exact(gpu, cols, rows):
    print("pyramid parity %dx%d" % (cols, rows), flush=True)
    _, K, frames = scene_frames(frames=(3,), cols=cols, rows=rows)
    d16, _ = frames[0]
    g = _pyr_gpu(d16, K)
    o = _pyr_oracle(d16, K)
    assert same_bits(g[0], o[0]), mismatch(g[0], o[0])
    for l in range(3):
        assert same_bits(g[1][l], o[1][l]), ("depth", l, mismatch(g[1][l], o[1][l]))
        assert same_bits(g[2][l], o[2][l]), ("vmap", l, mismatch(g[2][l], o[2][l]))
        assert same_bits(g[3][l], o[3][l]), ("nmap", l, mismatch(g[3][l], o[3][l]))


def test_pyramid_edge_frames(gpu):
    """all-dropped, all-beyond-cutoff and single-valid-pixel frames"""
    K = (525.0, 525.0, 319.5, 239.5)
    frames = [np.zeros((480, 640), np.uint16), np.full((480, 640), 60000, np.uint16)]
    one = np.zeros((480, 640), np.uint16)
    one[100, 200] = 7000
    frames.append(one)
    for d16 in frames:
        g = _pyr_gpu(d16, K)
        o = _pyr_oracle(d16, K)
        assert same_bits(g[0], o[0])
        for l in range(3):
            assert same_bits(g[1][l], o[1][l])
            assert same_bits(g[2][l], o[2][l])
            assert same_bits(g[3][l], o[3][l])


@pytest.mark.parametrize("cols,rows", SIZES)
def test_resize_map_bit_exact(gpu, cols, rows):
    _, K, frames = scene_frames(frames=(5,), cols=cols, rows=rows)
    _, _, vs, ns = _pyr_oracle(frames[0][0], K)
    vd, nd = A.ResizeMap(torch.from_numpy(vs[0]).cuda(), torch.from_numpy(ns[0]).cuda())
    ov, on = O.resize_map(vs[0], ns[0])
    assert same_bits(vd.cpu().numpy(), ov)
    assert same_bits(nd.cpu().numpy(), on)
