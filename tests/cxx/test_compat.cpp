// A System-shaped test program written against the reference's host API (Tracking/Reduction.h,
// Mapping/RenderScene.h; include/fusion_amd_compat.hpp): per frame FillImages, the ICP loop
// of Tracker::ComputeSE3 with its host solve (Tracking.cc:296-409), FuseMapColor,
// CreateRenderingBlocks and Raycast (System.cc:92-138), MeshScene at the end, and the keypoint
// map calls.  Every call is also made through the C ABI on the same inputs, and the two
// results must be bit-identical.  Prints one line per frame ("frame i status s pose ...") and
// "compat ok <checks>" at the end; exits non-zero on the first mismatch.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/fusion_amd_compat.hpp"
#include "../../include/fusion_scene.h"

static int g_checks = 0;

#define EXPECT(cond, ...)                    \
    do {                                     \
        if (!(cond)) {                       \
            fprintf(stderr, "MISMATCH: ");   \
            fprintf(stderr, __VA_ARGS__);    \
            fprintf(stderr, "\n");           \
            exit(1);                         \
        }                                    \
        ++g_checks;                          \
    } while (0)

template <class T>
static std::vector<unsigned char> bytes2d(const DeviceArray2D<T>& a) {
    std::vector<unsigned char> h((size_t)a.cols * a.rows * sizeof(T));
    a.download(h.data(), sizeof(T) * a.cols);
    return h;
}
static std::vector<unsigned char> bytes_dev(const void* p, size_t n) {
    std::vector<unsigned char> h(n);
    fusion_amd::check_hip(hipMemcpy(h.data(), p, n, hipMemcpyDeviceToHost), "download");
    return h;
}
template <class T>
static void same_as(const DeviceArray2D<T>& a, const void* ref, const char* what) {
    const auto x = bytes2d(a);
    const auto y = bytes_dev(ref, x.size());
    EXPECT(x == y, "%s", what);
}
// a mesh's triangles (vertices, normals, colours of each) as a sorted list: two maps built from
// the same frames hold their blocks in pool slots of atomic order, so their meshes list the
// same triangles in different orders
static std::vector<std::string> triangles(const void* v, const void* n, const void* c, size_t ntri) {
    const auto hv = bytes_dev(v, ntri * 36), hn = bytes_dev(n, ntri * 36), hc = bytes_dev(c, ntri * 9);
    std::vector<std::string> t(ntri);
    for (size_t i = 0; i < ntri; ++i) {
        t[i].assign((const char*)&hv[i * 36], 36);
        t[i].append((const char*)&hn[i * 36], 36);
        t[i].append((const char*)&hc[i * 9], 9);
    }
    std::sort(t.begin(), t.end());
    return t;
}
static void* dalloc(size_t n) {
    void* p = nullptr;
    fusion_amd::check_hip(hipMalloc(&p, n), "hipMalloc");
    return p;
}

// FP64 pose algebra of ComputeSE3: Cholesky solve of A x = b, SE(3) exp of (upsilon, omega)
static bool solve6(const double* A, const double* b, double* x) {
    double L[36] = {0};
    for (int j = 0; j < 6; ++j) {
        double s = A[j * 6 + j];
        for (int k = 0; k < j; ++k) s -= L[j * 6 + k] * L[j * 6 + k];
        if (!(s > 0)) return false;
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
    return true;
}
static void se3_exp(const double* xi, double* T) {
    const double w[3] = {xi[3], xi[4], xi[5]};
    const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const double A = th < 1e-9 ? 1.0 : sin(th) / th, B = th < 1e-9 ? 0.5 : (1 - cos(th)) / (th * th),
                 C = th < 1e-9 ? 1.0 / 6 : (th - sin(th)) / (th * th * th);
    const double W[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0};
    double W2[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            W2[i * 3 + j] = 0;
            for (int k = 0; k < 3; ++k) W2[i * 3 + j] += W[i * 3 + k] * W[k * 3 + j];
        }
    memset(T, 0, 16 * sizeof(double));
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) T[i * 4 + j] = (i == j) + A * W[i * 3 + j] + B * W2[i * 3 + j];
        for (int j = 0; j < 3; ++j) T[i * 4 + 3] += ((i == j) + B * W[i * 3 + j] + C * W2[i * 3 + j]) * xi[j];
    }
    T[15] = 1;
}
static void mul4(const double* a, const double* b, double* o) {
    double t[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            t[i * 4 + j] = 0;
            for (int k = 0; k < 4; ++k) t[i * 4 + j] += a[i * 4 + k] * b[k * 4 + j];
        }
    memcpy(o, t, sizeof(t));
}
static void inv4(const double* a, double* o) {  // rigid inverse
    double t[16] = {0};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) t[i * 4 + j] = a[j * 4 + i];
        t[i * 4 + 3] = -(a[0 * 4 + i] * a[3] + a[1 * 4 + i] * a[7] + a[2 * 4 + i] * a[11]);
    }
    t[15] = 1;
    memcpy(o, t, sizeof(t));
}
// Frame::GpuRotation / GpuInvRotation / GpuTranslation (Core/Frame.cc:275-294)
static Matrix3f rot(const double* p, bool inv) {
    Matrix3f m;
    float3* r[3] = {&m.rowx, &m.rowy, &m.rowz};
    for (int i = 0; i < 3; ++i)
        *r[i] = inv ? make_float3(p[0 * 4 + i], p[1 * 4 + i], p[2 * 4 + i]) : make_float3(p[i * 4 + 0], p[i * 4 + 1], p[i * 4 + 2]);
    return m;
}
static float3 trans(const double* p) { return make_float3(p[3], p[7], p[11]); }
static fu_pose_f cpose(const double* p) {
    fu_pose_f f;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            f.R[i * 3 + j] = (float)p[i * 4 + j];
            f.Rinv[i * 3 + j] = (float)p[j * 4 + i];
        }
        f.t[i] = (float)p[i * 4 + 3];
    }
    return f;
}

int main(int argc, char** argv) {
    const int frames = argc > 1 ? atoi(argv[1]) : 6;
    const int cols = 640, rows = 480;
    Intrinsics K(525.f, 525.f, 319.5f, 239.5f);
    fu_scene* scene = nullptr;
    if (fu_scene_create(FU_SCENE_ROOM, 4.f, 4.f, 3.f, 0, &scene)) return 2;
    fu_map_params mp;
    fu_map_params_default(&mp);
    mp.num_buckets = 1 << 17;
    mp.num_excess = 1 << 16;
    mp.num_blocks = 1 << 16;
    mp.voxel_size = 0.01f;
    mp.trunc_dist = 0.01f * 3.0f;
    DeviceMap map = CreateDeviceMap(mp);
    fu_map* ref = nullptr;  // the same map driven through the C ABI
    fusion_amd::check(fu_map_create(&mp, &ref), "fu_map_create");

    std::vector<uint16_t> hd((size_t)cols * rows);
    std::vector<uint8_t> hc((size_t)cols * rows * 3);
    DeviceArray2D<unsigned short> depth;
    DeviceArray2D<uchar3> color;
    DeviceArray2D<float> raw, filt[3];
    DeviceArray2D<float4> vmap[3], nmap[3], lastV[3], lastN[3];
    DeviceArray2D<float> zmin(cols / 8, rows / 8), zmax(cols / 8, rows / 8);
    lastV[0].create(cols, rows);  // the model view is the raycast's output (Frame.cc:44-48 sizes it)
    lastN[0].create(cols, rows);
    DeviceArray2D<float> sum;
    DeviceArray<float> out;
    DeviceArray<uint> noVisible, noRender;
    DeviceArray<RenderingBlock> renderList;
    // dense buffers of the C ABI side
    const size_t n0 = (size_t)cols * rows;
    float *r_raw = (float*)dalloc(n0 * 4), *r_f = (float*)dalloc(n0 * 4), *r_v = (float*)dalloc(n0 * 16),
          *r_n = (float*)dalloc(n0 * 16), *r_zx = (float*)dalloc(n0 / 64 * 4), *r_zy = (float*)dalloc(n0 / 64 * 4);
    const int iters[3] = {10, 5, 4}, minc[3] = {2000, 1000, 100};
    double last[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};

    for (int f = 0; f < frames; ++f) {
        fu_scene_render(scene, f, cols, rows, K.fx, K.fy, K.cx, K.cy, 5000.f, 1, hd.data(), hc.data());
        depth.upload(hd.data(), cols * 2, cols, rows);
        color.upload(hc.data(), cols * 3, cols, rows);
        // Frame::FillImages (Frame.cc:59-78)
        FilterDepth(depth, raw, filt[0], 5000.f, 3.f);
        fusion_amd::check(fu_filter_depth((const uint16_t*)depth.data, r_raw, r_f, cols, rows, 5000.f, 3.f, nullptr), "abi");
        same_as(raw, r_raw, "FilterDepth raw");
        same_as(filt[0], r_f, "FilterDepth filtered");
        for (int l = 1; l < 3; ++l) {
            PyrDownGauss(filt[l - 1], filt[l]);
            fusion_amd::check(fu_pyrdown_gauss((const float*)filt[l - 1].data, filt[l - 1].cols, filt[l - 1].rows, r_f, nullptr), "abi");
            same_as(filt[l], r_f, "PyrDownGauss");
        }
        for (int l = 0; l < 3; ++l) {
            Intrinsics k = K(l);
            ComputeVMap(filt[l], vmap[l], k.fx, k.fy, k.cx, k.cy, 3.f);
            ComputeNMap(vmap[l], nmap[l]);
            fusion_amd::check(fu_compute_vmap((const float*)filt[l].data, r_v, filt[l].cols, filt[l].rows, k.fx, k.fy, k.cx, k.cy, 3.f, nullptr), "abi");
            fusion_amd::check(fu_compute_nmap(r_v, r_n, filt[l].cols, filt[l].rows, nullptr), "abi");
            same_as(vmap[l], r_v, "ComputeVMap");
            same_as(nmap[l], r_n, "ComputeNMap");
        }
        int status = 4;
        double pose[16];
        memcpy(pose, last, sizeof(pose));
        if (f > 0) {
            // LastFrame->ResizeImages (Frame.cc:80-84)
            for (int l = 1; l < 3; ++l) {
                ResizeMap(lastV[l - 1], lastN[l - 1], lastV[l], lastN[l]);
                fusion_amd::check(fu_resize_map((const float*)lastV[l - 1].data, (const float*)lastN[l - 1].data, lastV[l - 1].cols,
                                                lastV[l - 1].rows, r_v, r_n, nullptr), "abi");
                same_as(lastV[l], r_v, "ResizeMap vertices");
                same_as(lastN[l], r_n, "ResizeMap normals");
            }
            // Tracker::ComputeSE3 (Tracking.cc:296-409), ICP only
            double delta[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}, next[16];
            memcpy(next, last, sizeof(next));
            float err = 0;
            status = 1;
            for (int l = 2; l >= 0 && status == 1; --l)
                for (int j = 0; j < iters[l] && status == 1; ++j) {
                    float res[2], res2[2];
                    double A[36], b[6], A2[36], b2[6];
                    ICPStep(vmap[l], lastV[l], nmap[l], lastN[l], rot(next, false), trans(next), rot(last, false),
                            rot(last, true), trans(last), K(l), sum, out, res, A, b);
                    const fu_pose_f pn = cpose(next), pl = cpose(last);
                    Intrinsics k = K(l);
                    fusion_amd::check(fu_icp_step((const float*)vmap[l].data, (const float*)lastV[l].data, (const float*)nmap[l].data,
                                                  (const float*)lastN[l].data, vmap[l].cols, vmap[l].rows, pn.R, pn.t, pl.R,
                                                  pl.Rinv, pl.t, k.fx, k.fy, k.cx, k.cy, res2, A2, b2, nullptr), "abi");
                    EXPECT(!memcmp(res, res2, 8) && !memcmp(A, A2, sizeof(A)) && !memcmp(b, b2, sizeof(b)), "ICPStep");
                    err = sqrtf(res[0]) / res[1];
                    double x[6], dT[16], dinv[16];
                    if (isnan(err) || (int)res[1] < minc[l] || !solve6(A, b, x)) {
                        status = 2;
                        break;
                    }
                    se3_exp(x, dT);
                    mul4(dT, delta, delta);
                    inv4(delta, dinv);
                    mul4(last, dinv, next);
                }
            if (status == 1) {
                double pinv[16], p[16];
                inv4(last, pinv);
                mul4(pinv, next, p);
                const double tn = sqrt(p[3] * p[3] + p[7] * p[7] + p[11] * p[11]);
                // |sin| of the XYZ Euler angles (Tracking.cc:401)
                const double sa = -p[6] / sqrt(p[6] * p[6] + p[10] * p[10]), sb = p[2],
                             sc = -p[1] / sqrt(p[1] * p[1] + p[0] * p[0]);
                const double an = sqrt(sa * sa + sb * sb + sc * sc);
                if (!(err < 1e-4 && an <= 0.1 && tn <= 0.1)) status = 3;
            }
            if (status == 1) memcpy(pose, next, sizeof(pose));
        }
        printf("frame %d status %d pose", f, status);
        for (int i = 0; i < 16; ++i) printf(" %.17g", pose[i]);
        printf("\n");
        if (status != 1 && status != 4) continue;
        memcpy(last, pose, sizeof(last));
        // Mapping::FuseColor + Mapping::RayTrace (Mapping.cc:76-109)
        uint nvis = 0, nvis2 = 0;
        const Matrix3f R = rot(pose, false), Ri = rot(pose, true);
        const float3 t = trans(pose);
        FuseMapColor(raw, color, nmap[0], noVisible, R, Ri, t, map, K.fx, K.fy, K.cx, K.cy, 3.f, 0.1f, &nvis);
        const fu_pose_f cp = cpose(pose);
        fusion_amd::check(fu_fuse_map_color(ref, (const float*)raw.data, (const uint8_t*)color.data, (const float*)nmap[0].data,
                                            cols, rows, &cp, K.fx, K.fy, K.cx, K.cy, 3.f, 0.1f, &nvis2, nullptr), "abi");
        EXPECT(nvis == nvis2 && nvis > 0, "FuseMapColor visible %u vs %u", nvis, nvis2);
        uint dev_nvis = 0;
        noVisible.download(&dev_nvis, 1);
        EXPECT(dev_nvis == nvis, "noVisibleBlocks");
        const bool rendered = CreateRenderingBlocks(map.visibleEntries, zmin, zmax, 3.f, 0.1f, renderList, noRender, Ri, t,
                                                    nvis, K.fx, K.fy, K.cx, K.cy);
        uint32_t nf2 = 0;
        int rendered2 = 0;
        fusion_amd::check(fu_create_rendering_blocks(ref, r_zx, r_zy, cols / 8, rows / 8, 3.f, 0.1f, &cp, nvis2, K.fx, K.fy,
                                                     K.cx, K.cy, &nf2, &rendered2, nullptr), "abi");
        EXPECT(rendered == (rendered2 != 0), "CreateRenderingBlocks");
        same_as(zmin, r_zx, "z-range minimum");
        same_as(zmax, r_zy, "z-range maximum");
        if (rendered) {
            Raycast(map, lastV[0], lastN[0], zmin, zmax, R, Ri, t, 1.f / K.fx, 1.f / K.fy, K.cx, K.cy);
            fusion_amd::check(fu_raycast(ref, r_v, r_n, cols, rows, r_zx, r_zy, cols / 8, rows / 8, &cp, 1.f / K.fx, 1.f / K.fy,
                                         K.cx, K.cy, nullptr), "abi");
            same_as(lastV[0], r_v, "Raycast vertices");
            same_as(lastN[0], r_n, "Raycast normals");
        } else {
            vmap[0].copyTo(lastV[0]);
            nmap[0].copyTo(lastN[0]);
        }
    }
    // MeshScene + Mapping::CreateModel (Mapping.cc:131-140)
    const size_t maxTri = 2000000;
    DeviceArray<float3> vert(3 * maxTri), nrm(3 * maxTri);
    DeviceArray<uchar3> col(3 * maxTri);
    DeviceArray<int3> blocks(1 << 16);
    DeviceArray<uint> nOcc, nTri;
    DeviceArray<int> edge, vtab;
    DeviceArray2D<int> tri;
    const uint ntri = MeshScene(nOcc, nTri, map, edge, vtab, tri, nrm, vert, col, blocks);
    float *rv = (float*)dalloc(maxTri * 36), *rn = (float*)dalloc(maxTri * 36);
    uint8_t* rc = (uint8_t*)dalloc(maxTri * 9);
    uint32_t ntri2 = 0, nocc2 = 0;
    fusion_amd::check(fu_mesh_scene(ref, rv, rn, rc, nullptr, (uint32_t)maxTri, &ntri2, &nocc2, nullptr), "abi");
    EXPECT(ntri == ntri2 && ntri > 1000, "MeshScene triangles %u vs %u", ntri, ntri2);
    EXPECT(triangles(vert.data, nrm.data, col.data, ntri) == triangles(rv, rn, rc, ntri), "MeshScene triangles");

    // keypoint map: InsertKeyPoints / CollectKeyPoints / ResetKeyPoints / BuildAdjacencyMatrix
    fu_keymap_params kp;
    fu_keymap_params_default(&kp);
    KeyMap km = CreateKeyMap(kp);
    fu_keymap* kref = nullptr;
    fusion_amd::check(fu_keymap_create(&kp, &kref), "abi");
    std::vector<SURF> hk(500);
    for (size_t i = 0; i < hk.size(); ++i) {
        hk[i].valid = true;
        hk[i].pos = make_float3(0.013f * (float)(i % 37), 0.021f * (float)(i % 11), 0.5f + 0.007f * (float)i);
        hk[i].normal = make_float4(0.f, 0.f, 1.f, 0.f);
        for (int j = 0; j < 64; ++j) hk[i].descriptor[j] = (float)((i * 31 + j) % 17);
    }
    DeviceArray<SURF> keys(hk), keys2(hk);
    std::vector<int> hidx(hk.size(), -1);
    DeviceArray<int> idx(hidx), idx2(hidx);
    InsertKeyPoints(km, keys, idx, hk.size());
    fusion_amd::check(fu_insert_key_points(kref, (fu_surf*)keys2.data, (int32_t*)idx2.data, (uint32_t)hk.size(), nullptr), "abi");
    EXPECT(bytes_dev(idx.data, hk.size() * 4) == bytes_dev(idx2.data, hk.size() * 4), "InsertKeyPoints indices");
    DeviceArray<SURF> got(1000), got2(1000);
    DeviceArray<uint> nk;
    CollectKeyPoints(km, got, nk);
    uint32_t n2 = 0;
    fusion_amd::check(fu_collect_key_points(kref, (fu_surf*)got2.data, 1000, &n2, nullptr), "abi");
    uint nkh = 0;
    nk.download(&nkh, 1);
    EXPECT(nkh == n2 && n2 > 0, "CollectKeyPoints count");
    EXPECT(bytes_dev((char*)got.data + 288, (n2 - 1) * 288) == bytes_dev((char*)got2.data + 288, (n2 - 1) * 288),
           "CollectKeyPoints records");
    DeviceArray2D<float> adj;
    std::vector<float> hdist(hk.size(), 0.25f);
    DeviceArray<float> dist(hdist);
    BuildAdjacencyMatrix(adj, keys, keys2, dist);
    float* radj = (float*)dalloc(hk.size() * hk.size() * 4);
    fusion_amd::check(fu_build_adjacency_matrix((const fu_surf*)keys.data, (const fu_surf*)keys2.data, (const float*)dist.data,
                                                (int)hk.size(), radj, nullptr), "abi");
    fusion_amd::check_hip(hipDeviceSynchronize(), "sync");
    same_as(adj, radj, "BuildAdjacencyMatrix");
    ResetKeyPoints(km);
    CollectKeyPoints(km, got, nk);
    nk.download(&nkh, 1);
    EXPECT(nkh == 0, "ResetKeyPoints");

    // shallow assignment shares the buffer and the last owner frees it (DeviceArray.h:236-251)
    {
        DeviceArray<float> a(16), b2;
        b2 = a;
        EXPECT(b2.data == a.data && a.ref->load() == 2, "DeviceArray sharing");
    }
    ReleaseKeyMap(km);
    fu_keymap_destroy(kref);
    ReleaseDeviceMap(map);
    fu_map_destroy(ref);
    fu_scene_destroy(scene);
    printf("compat ok %d\n", g_checks);
    return 0;
}
