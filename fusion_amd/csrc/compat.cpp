// The reference's free functions (Tracking/Reduction.h:18-96, Mapping/RenderScene.h:3-52)
// over the C ABI; types and the differences from the reference are described in
// include/fusion_amd_compat.hpp.  Host code only: every call goes through fusion_amd.h on the
// legacy default stream and synchronises before it returns.
#include "../../include/fusion_amd_compat.hpp"

#include <mutex>
#include <unordered_map>

namespace fusion_amd {
void check(fu_status st, const char* what) {
    if (st != FU_OK) throw Error(std::string(what) + ": " + fu_last_error());
}
void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace fusion_amd

using fusion_amd::check;
using fusion_amd::check_hip;

namespace {

// the maps behind DeviceMap::visibleEntries (CreateRenderingBlocks receives only the list)
std::mutex g_lock;
std::unordered_map<const void*, fu_map*> g_maps;

void rows9(const Matrix3f& m, float* o) {
    const float3 r[3] = {m.rowx, m.rowy, m.rowz};
    for (int i = 0; i < 3; ++i) {
        o[i * 3 + 0] = r[i].x;
        o[i * 3 + 1] = r[i].y;
        o[i * 3 + 2] = r[i].z;
    }
}
fu_pose_f pose_of(const Matrix3f& R, const Matrix3f& Rinv, float3 t) {
    fu_pose_f p;
    rows9(R, p.R);
    rows9(Rinv, p.Rinv);
    p.t[0] = t.x;
    p.t[1] = t.y;
    p.t[2] = t.z;
    return p;
}
// a pose given by its inverse rotation only: the rotation is its transpose (exact in float)
fu_pose_f pose_of_inv(const Matrix3f& Rinv, float3 t) {
    fu_pose_f p;
    rows9(Rinv, p.Rinv);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) p.R[i * 3 + j] = p.Rinv[j * 3 + i];
    p.t[0] = t.x;
    p.t[1] = t.y;
    p.t[2] = t.z;
    return p;
}

void sync() { check_hip(hipDeviceSynchronize(), "synchronise"); }

// a dense device view of a 2-D array: the array itself when its rows are dense, else a dense
// staging copy (copied in for inputs, back by commit() for outputs)
template <class T>
struct Dense {
    DeviceArray2D<T>* a;
    void* tmp = nullptr;
    T* p;
    Dense(const DeviceArray2D<T>& arr, bool in) : a(const_cast<DeviceArray2D<T>*>(&arr)) {
        const size_t row = sizeof(T) * (size_t)a->cols;
        if (a->step == row || a->rows <= 1) {
            p = (T*)a->data;
            return;
        }
        check_hip(hipMalloc(&tmp, row * a->rows), "staging copy");
        if (in) check_hip(hipMemcpy2D(tmp, row, a->data, a->step, row, a->rows, hipMemcpyDeviceToDevice), "staging copy");
        p = (T*)tmp;
    }
    void commit() {
        if (!tmp) return;
        const size_t row = sizeof(T) * (size_t)a->cols;
        check_hip(hipMemcpy2D(a->data, a->step, tmp, row, row, a->rows, hipMemcpyDeviceToDevice), "staging copy");
    }
    ~Dense() {
        if (tmp) (void)hipFree(tmp);
    }
    Dense(const Dense&) = delete;
    Dense& operator=(const Dense&) = delete;
};

template <class T>
void ensure(DeviceArray2D<T>& a, int cols, int rows) {
    if (a.cols != cols || a.rows != rows || !a.data) a.create(cols, rows);
}
void put_count(DeviceArray<uint>& a, uint v) {
    if (a.size < 1) a.create(1);
    check_hip(hipMemcpy(a.data, &v, 4, hipMemcpyHostToDevice), "count");
}

}  // namespace

// ------------------------------------------------------------------------ handles
DeviceMap CreateDeviceMap(const fu_map_params& params) {
    DeviceMap m;
    check(fu_map_create(&params, &m.map), "CreateDeviceMap");
    m.visibleEntries.create(1);
    std::lock_guard<std::mutex> g(g_lock);
    g_maps[m.visibleEntries.data] = m.map;
    return m;
}
void ReleaseDeviceMap(DeviceMap& m) {
    {
        std::lock_guard<std::mutex> g(g_lock);
        g_maps.erase(m.visibleEntries.data);
    }
    m.visibleEntries.release();
    fu_map_destroy(m.map);
    m.map = nullptr;
}
KeyMap CreateKeyMap(const fu_keymap_params& params) {
    KeyMap k;
    check(fu_keymap_create(&params, &k.map), "CreateKeyMap");
    return k;
}
void ReleaseKeyMap(KeyMap& k) {
    fu_keymap_destroy(k.map);
    k.map = nullptr;
}

// ------------------------------------------------------------------------ RenderScene.h
void ResetMap(DeviceMap map) {
    check(fu_reset_map(map.map, nullptr), "ResetMap");
    sync();
}

void ResetKeyPoints(KeyMap map) {
    check(fu_reset_key_points(map.map, nullptr), "ResetKeyPoints");
    sync();
}

void InsertKeyPoints(KeyMap map, DeviceArray<SURF>& keys, DeviceArray<int>& keyIndex, size_t size) {
    if (size == 0) return;
    if (size > keys.size || size > keyIndex.size) throw fusion_amd::Error("InsertKeyPoints: size beyond the arrays");
    check(fu_insert_key_points(map.map, (fu_surf*)keys.data, (int32_t*)keyIndex.data, (uint32_t)size, nullptr),
          "InsertKeyPoints");
}

void CollectKeyPoints(KeyMap map, DeviceArray<SURF>& keys, DeviceArray<uint>& noKeys) {
    uint32_t n = 0;
    check(fu_collect_key_points(map.map, (fu_surf*)keys.data, (uint32_t)keys.size, &n, nullptr), "CollectKeyPoints");
    put_count(noKeys, n);
}

void Raycast(DeviceMap map, DeviceArray2D<float4>& vmap, DeviceArray2D<float4>& nmap, DeviceArray2D<float>& zRangeX,
             DeviceArray2D<float>& zRangeY, Matrix3f Rview, Matrix3f RviewInv, float3 tview, float invfx, float invfy,
             float cx, float cy) {
    if (!vmap.data || !nmap.data || nmap.cols != vmap.cols || nmap.rows != vmap.rows)
        throw fusion_amd::Error("Raycast: vmap and nmap must be created with the image size");
    const fu_pose_f p = pose_of(Rview, RviewInv, tview);
    Dense<float4> v(vmap, false), n(nmap, false);
    Dense<float> zx(zRangeX, true), zy(zRangeY, true);
    check(fu_raycast(map.map, (float*)v.p, (float*)n.p, vmap.cols, vmap.rows, zx.p, zy.p, zRangeX.cols, zRangeX.rows,
                     &p, invfx, invfy, cx, cy, nullptr),
          "Raycast");
    v.commit();
    n.commit();
    sync();
}

bool CreateRenderingBlocks(const DeviceArray<HashEntry>& visibleBlocks, DeviceArray2D<float>& zRangeX,
                           DeviceArray2D<float>& zRangeY, const float& depthMax, const float& depthMin,
                           DeviceArray<RenderingBlock>& renderingBlockList, DeviceArray<uint>& noRenderingBlocks,
                           Matrix3f RviewInv, float3 tview, uint noVisibleBlocks, float fx, float fy, float cx,
                           float cy) {
    (void)renderingBlockList;  // the library folds the fragments straight into the tile maps
    fu_map* m = nullptr;
    {
        std::lock_guard<std::mutex> g(g_lock);
        auto it = g_maps.find(visibleBlocks.data);
        if (it != g_maps.end()) m = it->second;
    }
    if (!m) throw fusion_amd::Error("CreateRenderingBlocks: visibleBlocks is not a DeviceMap's visibleEntries");
    const fu_pose_f p = pose_of_inv(RviewInv, tview);
    Dense<float> zx(zRangeX, false), zy(zRangeY, false);
    uint32_t nfrag = 0;
    int rendered = 0;
    check(fu_create_rendering_blocks(m, zx.p, zy.p, zRangeX.cols, zRangeX.rows, depthMax, depthMin, &p, noVisibleBlocks,
                                     fx, fy, cx, cy, &nfrag, &rendered, nullptr),
          "CreateRenderingBlocks");
    zx.commit();
    zy.commit();
    put_count(noRenderingBlocks, nfrag);
    sync();
    return rendered != 0;
}

uint MeshScene(DeviceArray<uint>& noOccupiedBlocks, DeviceArray<uint>& noTotalTriangles, DeviceMap map,
               const DeviceArray<int>& edgeTable, const DeviceArray<int>& vertexTable,
               const DeviceArray2D<int>& triangleTable, DeviceArray<float3>& normal, DeviceArray<float3>& vertex,
               DeviceArray<uchar3>& color, DeviceArray<int3>& blockPoses) {
    (void)edgeTable;
    (void)vertexTable;
    (void)triangleTable;
    static_assert(sizeof(float3) == 12 && sizeof(uchar3) == 3 && sizeof(int3) == 12, "packed vector types");
    size_t cap = vertex.size;
    if (normal.size < cap) cap = normal.size;
    if (color.size < cap) cap = color.size;
    uint32_t ntri = 0, nocc = 0;
    check(fu_mesh_scene(map.map, (float*)vertex.data, (float*)normal.data, (uint8_t*)color.data, nullptr,
                        (uint32_t)(cap / 3), &ntri, &nocc, nullptr),
          "MeshScene");
    put_count(noOccupiedBlocks, nocc);
    put_count(noTotalTriangles, ntri);
    // the allocated blocks' coordinates (CheckBlockKernel's compaction, MeshScene.cu:20-41)
    uint32_t nb = 0;
    check(fu_map_export_blocks(map.map, nullptr, 0, &nb, nullptr), "MeshScene");
    if (nb && blockPoses.size >= nb) {
        std::vector<int3> xyz(nb);
        check(fu_map_export_blocks(map.map, (int32_t*)xyz.data(), nb, &nb, nullptr), "MeshScene");
        check_hip(hipMemcpy(blockPoses.data, xyz.data(), (size_t)nb * sizeof(int3), hipMemcpyHostToDevice), "MeshScene");
    }
    sync();
    return ntri;
}

void CheckBlockVisibility(DeviceMap map, DeviceArray<uint>& noVisibleBlocks, Matrix3f Rview, Matrix3f RviewInv,
                          float3 tview, int cols, int rows, float fx, float fy, float cx, float cy, float depthMax,
                          float depthMin, uint* host_data) {
    const fu_pose_f p = pose_of(Rview, RviewInv, tview);
    uint32_t n = 0;
    check(fu_check_block_visibility(map.map, &p, cols, rows, fx, fy, cx, cy, depthMax, depthMin, &n, nullptr),
          "CheckBlockVisibility");
    put_count(noVisibleBlocks, n);
    if (host_data) *host_data = n;
    sync();
}

void FuseMapColor(const DeviceArray2D<float>& depth, const DeviceArray2D<uchar3>& color,
                  const DeviceArray2D<float4>& nmap, DeviceArray<uint>& noVisibleBlocks, Matrix3f Rview,
                  Matrix3f RviewInv, float3 tview, DeviceMap map, float fx, float fy, float cx, float cy,
                  float depthMax, float depthMin, uint* host_data) {
    const fu_pose_f p = pose_of(Rview, RviewInv, tview);
    Dense<float> d(depth, true);
    Dense<uchar3> c(color, true);
    Dense<float4> n(nmap, true);
    uint32_t nv = 0;
    check(fu_fuse_map_color(map.map, d.p, (const uint8_t*)c.p, (const float*)n.p, depth.cols, depth.rows, &p, fx, fy,
                            cx, cy, depthMax, depthMin, &nv, nullptr),
          "FuseMapColor");
    put_count(noVisibleBlocks, nv);
    if (host_data) *host_data = nv;
    sync();
}

// ------------------------------------------------------------------------ Reduction.h
void FilterDepth(const DeviceArray2D<unsigned short>& depth, DeviceArray2D<float>& rawDepth,
                 DeviceArray2D<float>& filteredDepth, float depthScale, float depthCutoff) {
    ensure(rawDepth, depth.cols, depth.rows);
    ensure(filteredDepth, depth.cols, depth.rows);
    Dense<unsigned short> d(depth, true);
    Dense<float> r(rawDepth, false), f(filteredDepth, false);
    check(fu_filter_depth(d.p, r.p, f.p, depth.cols, depth.rows, depthScale, depthCutoff, nullptr), "FilterDepth");
    r.commit();
    f.commit();
    sync();
}

void ComputeVMap(const DeviceArray2D<float>& depth, DeviceArray2D<float4>& vmap, float fx, float fy, float cx, float cy,
                 float depthCutoff) {
    ensure(vmap, depth.cols, depth.rows);
    Dense<float> d(depth, true);
    Dense<float4> v(vmap, false);
    check(fu_compute_vmap(d.p, (float*)v.p, depth.cols, depth.rows, fx, fy, cx, cy, depthCutoff, nullptr), "ComputeVMap");
    v.commit();
    sync();
}

void ComputeNMap(const DeviceArray2D<float4>& vmap, DeviceArray2D<float4>& nmap) {
    ensure(nmap, vmap.cols, vmap.rows);
    Dense<float4> v(vmap, true), n(nmap, false);
    check(fu_compute_nmap((const float*)v.p, (float*)n.p, vmap.cols, vmap.rows, nullptr), "ComputeNMap");
    n.commit();
    sync();
}

void PyrDownGauss(const DeviceArray2D<float>& src, DeviceArray2D<float>& dst) {
    ensure(dst, src.cols / 2, src.rows / 2);
    Dense<float> s(src, true), d(dst, false);
    check(fu_pyrdown_gauss(s.p, src.cols, src.rows, d.p, nullptr), "PyrDownGauss");
    d.commit();
    sync();
}

void PyrDownGauss(const DeviceArray2D<unsigned char>& src, DeviceArray2D<unsigned char>& dst) {
    ensure(dst, src.cols / 2, src.rows / 2);
    Dense<unsigned char> s(src, true), d(dst, false);
    check(fu_pyrdown_gauss_u8(s.p, src.cols, src.rows, d.p, nullptr), "PyrDownGauss");
    d.commit();
    sync();
}

void ImageToIntensity(const DeviceArray2D<uchar3>& rgb, DeviceArray2D<unsigned char>& image) {
    ensure(image, rgb.cols, rgb.rows);
    Dense<uchar3> c(rgb, true);
    Dense<unsigned char> i(image, false);
    check(fu_image_to_intensity((const uint8_t*)c.p, i.p, rgb.cols, rgb.rows, nullptr), "ImageToIntensity");
    i.commit();
    sync();
}

void ComputeDerivativeImage(DeviceArray2D<unsigned char>& image, DeviceArray2D<short>& dx, DeviceArray2D<short>& dy) {
    ensure(dx, image.cols, image.rows);
    ensure(dy, image.cols, image.rows);
    Dense<unsigned char> i(image, true);
    Dense<short> x(dx, false), y(dy, false);
    check(fu_compute_derivative_image(i.p, x.p, y.p, image.cols, image.rows, nullptr), "ComputeDerivativeImage");
    x.commit();
    y.commit();
    sync();
}

void ResizeMap(const DeviceArray2D<float4>& vsrc, const DeviceArray2D<float4>& nsrc, DeviceArray2D<float4>& vdst,
               DeviceArray2D<float4>& ndst) {
    ensure(vdst, vsrc.cols / 2, vsrc.rows / 2);
    ensure(ndst, vsrc.cols / 2, vsrc.rows / 2);
    Dense<float4> vs(vsrc, true), ns(nsrc, true), vd(vdst, false), nd(ndst, false);
    check(fu_resize_map((const float*)vs.p, (const float*)ns.p, vsrc.cols, vsrc.rows, (float*)vd.p, (float*)nd.p, nullptr),
          "ResizeMap");
    vd.commit();
    nd.commit();
    sync();
}

void ICPStep(DeviceArray2D<float4>& nextVMap, DeviceArray2D<float4>& lastVMap, DeviceArray2D<float4>& nextNMap,
             DeviceArray2D<float4>& lastNMap, Matrix3f Rcurr, float3 tcurr, Matrix3f Rlast, Matrix3f RlastInv,
             float3 tlast, Intrinsics K, DeviceArray2D<float>& sum, DeviceArray<float>& out, float* residual,
             double* matrixA_host, double* vectorB_host) {
    (void)sum;
    (void)out;
    float rc[9], rl[9], rli[9];
    rows9(Rcurr, rc);
    rows9(Rlast, rl);
    rows9(RlastInv, rli);
    const float tc[3] = {tcurr.x, tcurr.y, tcurr.z}, tl[3] = {tlast.x, tlast.y, tlast.z};
    Dense<float4> nv(nextVMap, true), lv(lastVMap, true), nn(nextNMap, true), ln(lastNMap, true);
    check(fu_icp_step((const float*)nv.p, (const float*)lv.p, (const float*)nn.p, (const float*)ln.p, nextVMap.cols,
                      nextVMap.rows, rc, tc, rl, rli, tl, K.fx, K.fy, K.cx, K.cy, residual, matrixA_host, vectorB_host,
                      nullptr),
          "ICPStep");
}

void RGBStep(const DeviceArray2D<unsigned char>& nextImage, const DeviceArray2D<unsigned char>& lastImage,
             const DeviceArray2D<float4>& nextVMap, const DeviceArray2D<float4>& lastVMap,
             const DeviceArray2D<short>& dIdx, const DeviceArray2D<short>& dIdy, Matrix3f Rcurr, Matrix3f RcurrInv,
             Matrix3f Rlast, Matrix3f RlastInv, float3 tcurr, float3 tlast, Intrinsics K, DeviceArray2D<float>& sum,
             DeviceArray<float>& out, DeviceArray2D<int>& sumRes, DeviceArray<int>& outRes, float* residual,
             double* matrixA_host, double* vectorB_host) {
    (void)sum;
    (void)out;
    (void)sumRes;
    (void)outRes;
    float rc[9], rci[9], rl[9], rli[9];
    rows9(Rcurr, rc);
    rows9(RcurrInv, rci);
    rows9(Rlast, rl);
    rows9(RlastInv, rli);
    const float tc[3] = {tcurr.x, tcurr.y, tcurr.z}, tl[3] = {tlast.x, tlast.y, tlast.z};
    Dense<unsigned char> ni(nextImage, true), li(lastImage, true);
    Dense<float4> nv(nextVMap, true), lv(lastVMap, true);
    Dense<short> gx(dIdx, true), gy(dIdy, true);
    check(fu_rgb_step(ni.p, li.p, (const float*)nv.p, (const float*)lv.p, gx.p, gy.p, nextImage.cols, nextImage.rows, rc,
                      rci, rl, rli, tc, tl, K.fx, K.fy, K.cx, K.cy, residual, matrixA_host, vectorB_host, nullptr,
                      nullptr),
          "RGBStep");
}

void BuildAdjacencyMatrix(DeviceArray2D<float>& adjacencyMatrix, DeviceArray<SURF>& frameKeys,
                          DeviceArray<SURF>& mapKeys, DeviceArray<float>& dist) {
    const int n = (int)frameKeys.size;
    if (mapKeys.size < frameKeys.size || dist.size < frameKeys.size)
        throw fusion_amd::Error("BuildAdjacencyMatrix: map keys or distances shorter than the frame keys");
    ensure(adjacencyMatrix, n, n);
    Dense<float> m(adjacencyMatrix, false);
    check(fu_build_adjacency_matrix((const fu_surf*)frameKeys.data, (const fu_surf*)mapKeys.data, (const float*)dist.data,
                                    n, m.p, nullptr),
          "BuildAdjacencyMatrix");
    m.commit();
    sync();
}
