// C++ drop-in layer over the C ABI (fusion_amd.h): the reference's host-facing types and free
// functions with their signatures, so a System-shaped driver written against the reference's
// Tracking/Reduction.h and Mapping/RenderScene.h compiles unchanged against this header and
// libfusion_amd_cxx.so.
//
// Types: PtrSz / PtrStep / PtrStepSz and the ref-counted DeviceArray / DeviceArray2D
// (Utility/DeviceArray.h:9-394; shallow assignment shares the buffer), Matrix3f
// (Utility/VectorMath.h:356-371), Intrinsics (Utility/Intrinsics.h:4-18), HashEntry,
// RenderingBlock and SURF (Mapping/DeviceMap.h:11-108), DeviceMap and KeyMap as handles of
// the library's map and keypoint map (the reference's are structs of raw device pointers).
//
// Differences, each for a reason:
// * DeviceArray2D rows are dense (step = cols * sizeof(T)); the reference pitches them with
//   cudaMallocPitch, whose pitch is the implementation's choice.  A DeviceArray2D whose step
//   is not dense (set by hand) is staged through a dense copy.
// * Errors throw fusion_amd::Error; the reference prints and calls exit(0) (SafeCall.h:13-20).
// * Every function synchronises the legacy default stream before returning, as the
//   reference's do.
// * BuildAdjacencyMatrix takes a DeviceArray2D<float> (n x n) where the reference takes an
//   OpenCV GpuMat; the column sums the reference computes and discards are not computed.
// * CreateRenderingBlocks, MeshScene and ICPStep take the reference's work arrays (the
//   visible list, the rendering-block list, the marching-cubes tables, the partial sums) for
//   signature compatibility; the library keeps its own: the visible list inside the map, its
//   own tables (the reference's, checked by fingerprint) and reduction buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fusion_amd.h"

typedef unsigned int uint;
typedef unsigned char uchar;

namespace fusion_amd {
struct Error : std::runtime_error {
    explicit Error(const std::string& what) : std::runtime_error(what) {}
};
// throws Error with the library's message unless `st` is FU_OK
void check(fu_status st, const char* what);
void check_hip(hipError_t e, const char* what);
}  // namespace fusion_amd

// ---------------------------------------------------------------- Utility/DeviceArray.h
template <class T>
struct PtrSz {
    __host__ __device__ T& operator[](int x) const { return data[x]; }
    __host__ __device__ operator T*() const { return data; }
    T* data;
    size_t size;
};
template <class T>
struct PtrStep {
    __host__ __device__ T* ptr(int y = 0) const { return (T*)((char*)data + y * step); }
    T* data;
    size_t step;
};
template <class T>
struct PtrStepSz {
    __host__ __device__ T* ptr(int y = 0) const { return (T*)((char*)data + y * step); }
    T* data;
    int cols;
    int rows;
    size_t step;
};

template <class T>
struct DeviceArray {
    DeviceArray() : data(nullptr), ref(nullptr), size(0) {}
    explicit DeviceArray(size_t n) : DeviceArray() { create(n); }
    explicit DeviceArray(const std::vector<T>& v) : DeviceArray() {
        create(v.size());
        upload(v);
    }
    DeviceArray(const DeviceArray& o) : data(o.data), ref(o.ref), size(o.size) {
        if (ref) ref->fetch_add(1);
    }
    ~DeviceArray() { release(); }
    void create(size_t n) {
        if (n == size && data) return;
        release();
        if (n == 0) return;
        fusion_amd::check_hip(hipMalloc(&data, n * sizeof(T)), "DeviceArray::create");
        ref = new std::atomic<int>(1);
        size = n;
    }
    void upload(const void* host) { upload(host, size); }
    void upload(const std::vector<T>& v) { upload(v.data(), v.size()); }
    void upload(const void* host, size_t n) {
        create(n);
        if (n) fusion_amd::check_hip(hipMemcpy(data, host, n * sizeof(T), hipMemcpyHostToDevice), "DeviceArray::upload");
    }
    void download(void* host) const { download(host, size); }
    void download(std::vector<T>& v) const {
        v.resize(size);
        download(v.data(), size);
    }
    void download(void* host, size_t n) const {
        if (n) fusion_amd::check_hip(hipMemcpy(host, data, n * sizeof(T), hipMemcpyDeviceToHost), "DeviceArray::download");
    }
    void clear() {
        if (size) fusion_amd::check_hip(hipMemset(data, 0, size * sizeof(T)), "DeviceArray::clear");
    }
    void release() {
        if (ref && ref->fetch_sub(1) == 1) {
            (void)hipFree(data);
            delete ref;
        }
        data = nullptr;
        ref = nullptr;
        size = 0;
    }
    void copyTo(DeviceArray<T>& o) const {
        o.create(size);
        if (size) fusion_amd::check_hip(hipMemcpy(o.data, data, size * sizeof(T), hipMemcpyDeviceToDevice), "DeviceArray::copyTo");
    }
    DeviceArray& operator=(const DeviceArray& o) {
        if (this != &o) {
            if (o.ref) o.ref->fetch_add(1);
            release();
            data = o.data;
            ref = o.ref;
            size = o.size;
        }
        return *this;
    }
    operator T*() const { return (T*)data; }
    operator PtrSz<T>() const { return PtrSz<T>{(T*)data, size}; }
    T* ptr() const { return (T*)data; }
    void* data;
    std::atomic<int>* ref;
    size_t size;
};

template <class T>
struct DeviceArray2D {
    DeviceArray2D() : data(nullptr), ref(nullptr), step(0), cols(0), rows(0) {}
    DeviceArray2D(int c, int r) : DeviceArray2D() { create(c, r); }
    DeviceArray2D(const DeviceArray2D& o) : data(o.data), ref(o.ref), step(o.step), cols(o.cols), rows(o.rows) {
        if (ref) ref->fetch_add(1);
    }
    ~DeviceArray2D() { release(); }
    void create(int c, int r) {
        if (data && c == cols && r == rows) return;
        release();
        if (c <= 0 || r <= 0) return;
        fusion_amd::check_hip(hipMalloc(&data, (size_t)c * r * sizeof(T)), "DeviceArray2D::create");
        ref = new std::atomic<int>(1);
        cols = c;
        rows = r;
        step = (size_t)c * sizeof(T);
    }
    void upload(const void* host) { upload(host, sizeof(T) * cols, cols, rows); }
    void upload(const void* host, size_t host_step) { upload(host, host_step, cols, rows); }
    void upload(const void* host, size_t host_step, int c, int r) {
        create(c, r);
        fusion_amd::check_hip(hipMemcpy2D(data, step, host, host_step, sizeof(T) * cols, rows, hipMemcpyHostToDevice),
                              "DeviceArray2D::upload");
    }
    void download(void* host, size_t host_step) const {
        fusion_amd::check_hip(hipMemcpy2D(host, host_step, data, step, sizeof(T) * cols, rows, hipMemcpyDeviceToHost),
                              "DeviceArray2D::download");
    }
    void clear() {
        if (data) fusion_amd::check_hip(hipMemset2D(data, step, 0, sizeof(T) * cols, rows), "DeviceArray2D::clear");
    }
    void release() {
        if (ref && ref->fetch_sub(1) == 1) {
            (void)hipFree(data);
            delete ref;
        }
        data = nullptr;
        ref = nullptr;
        step = 0;
        cols = rows = 0;
    }
    void swap(DeviceArray2D& o) {
        std::swap(data, o.data);
        std::swap(ref, o.ref);
        std::swap(step, o.step);
        std::swap(cols, o.cols);
        std::swap(rows, o.rows);
    }
    void copyTo(DeviceArray2D& o) const {
        o.create(cols, rows);
        fusion_amd::check_hip(hipMemcpy2D(o.data, o.step, data, step, sizeof(T) * cols, rows, hipMemcpyDeviceToDevice),
                              "DeviceArray2D::copyTo");
    }
    DeviceArray2D& operator=(const DeviceArray2D& o) {
        if (this != &o) {
            if (o.ref) o.ref->fetch_add(1);
            release();
            data = o.data;
            ref = o.ref;
            step = o.step;
            cols = o.cols;
            rows = o.rows;
        }
        return *this;
    }
    operator T*() const { return (T*)data; }
    operator PtrStep<T>() const { return PtrStep<T>{(T*)data, step}; }
    operator PtrStepSz<T>() const { return PtrStepSz<T>{(T*)data, cols, rows, step}; }
    void* data;
    std::atomic<int>* ref;
    size_t step;
    int cols, rows;
};

// ---------------------------------------------------------------- VectorMath.h, Intrinsics.h
struct Matrix3f {
    float3 rowx, rowy, rowz;
    Matrix3f() : rowx(make_float3(0, 0, 0)), rowy(make_float3(0, 0, 0)), rowz(make_float3(0, 0, 0)) {}
};
struct Intrinsics {
    Intrinsics() : fx(0), fy(0), cx(0), cy(0) {}
    Intrinsics(float fx_, float fy_, float cx_, float cy_) : fx(fx_), fy(fy_), cx(cx_), cy(cy_) {}
    Intrinsics operator()(int scale) {
        const int i = 1 << scale;
        return Intrinsics(fx / i, fy / i, cx / i, cy / i);
    }
    float fx, fy, cx, cy;
};

// ---------------------------------------------------------------- DeviceMap.h
struct __align__(8) RenderingBlock {
    short2 upperLeft;
    short2 lowerRight;
    float2 zRange;
};
struct __align__(16) HashEntry {
    int3 pos;
    int ptr;
    int offset;
};
struct SURF {
    bool valid;
    float3 pos;
    float4 normal;
    float descriptor[64];
};
static_assert(sizeof(SURF) == sizeof(fu_surf), "SURF and fu_surf share a layout");
// the library's voxel-hash map (Mapping::operator DeviceMap in the reference).  The map keeps
// its visible list inside; `visibleEntries` stands for it where the reference passes the list
// (CreateRenderingBlocks finds the map from it).
struct DeviceMap {
    fu_map* map = nullptr;
    DeviceArray<HashEntry> visibleEntries;
};
struct KeyMap {
    fu_keymap* map = nullptr;
};
// creation and release of the handles (the reference's Mapping::Create / ~Mapping)
DeviceMap CreateDeviceMap(const fu_map_params& params);
void ReleaseDeviceMap(DeviceMap& map);
KeyMap CreateKeyMap(const fu_keymap_params& params);
void ReleaseKeyMap(KeyMap& map);

// ---------------------------------------------------------------- Mapping/RenderScene.h
void ResetMap(DeviceMap map);
void ResetKeyPoints(KeyMap map);
void InsertKeyPoints(KeyMap map, DeviceArray<SURF>& keys, DeviceArray<int>& keyIndex, size_t size);
void CollectKeyPoints(KeyMap map, DeviceArray<SURF>& keys, DeviceArray<uint>& noKeys);
void Raycast(DeviceMap map, DeviceArray2D<float4>& vmap, DeviceArray2D<float4>& nmap, DeviceArray2D<float>& zRangeX,
             DeviceArray2D<float>& zRangeY, Matrix3f Rview, Matrix3f RviewInv, float3 tview, float invfx, float invfy,
             float cx, float cy);
// `visibleBlocks` is a DeviceMap's visibleEntries: the map's own visible list is used
bool CreateRenderingBlocks(const DeviceArray<HashEntry>& visibleBlocks, DeviceArray2D<float>& zRangeX,
                           DeviceArray2D<float>& zRangeY, const float& depthMax, const float& depthMin,
                           DeviceArray<RenderingBlock>& renderingBlockList, DeviceArray<uint>& noRenderingBlocks,
                           Matrix3f RviewInv, float3 tview, uint noVisibleBlocks, float fx, float fy, float cx,
                           float cy);
uint MeshScene(DeviceArray<uint>& noOccupiedBlocks, DeviceArray<uint>& noTotalTriangles, DeviceMap map,
               const DeviceArray<int>& edgeTable, const DeviceArray<int>& vertexTable,
               const DeviceArray2D<int>& triangleTable, DeviceArray<float3>& normal, DeviceArray<float3>& vertex,
               DeviceArray<uchar3>& color, DeviceArray<int3>& blockPoses);
void CheckBlockVisibility(DeviceMap map, DeviceArray<uint>& noVisibleBlocks, Matrix3f Rview, Matrix3f RviewInv,
                          float3 tview, int cols, int rows, float fx, float fy, float cx, float cy, float depthMax,
                          float depthMin, uint* host_data);
void FuseMapColor(const DeviceArray2D<float>& depth, const DeviceArray2D<uchar3>& color,
                  const DeviceArray2D<float4>& nmap, DeviceArray<uint>& noVisibleBlocks, Matrix3f Rview,
                  Matrix3f RviewInv, float3 tview, DeviceMap map, float fx, float fy, float cx, float cy,
                  float depthMax, float depthMin, uint* host_data);

// ---------------------------------------------------------------- Tracking/Reduction.h
void FilterDepth(const DeviceArray2D<unsigned short>& depth, DeviceArray2D<float>& rawDepth,
                 DeviceArray2D<float>& filteredDepth, float depthScale, float depthCutoff);
void ComputeVMap(const DeviceArray2D<float>& depth, DeviceArray2D<float4>& vmap, float fx, float fy, float cx, float cy,
                 float depthCutoff);
void ComputeNMap(const DeviceArray2D<float4>& vmap, DeviceArray2D<float4>& nmap);
void PyrDownGauss(const DeviceArray2D<float>& src, DeviceArray2D<float>& dst);
void PyrDownGauss(const DeviceArray2D<unsigned char>& src, DeviceArray2D<unsigned char>& dst);
void ImageToIntensity(const DeviceArray2D<uchar3>& rgb, DeviceArray2D<unsigned char>& image);
void ComputeDerivativeImage(DeviceArray2D<unsigned char>& image, DeviceArray2D<short>& dx, DeviceArray2D<short>& dy);
void ResizeMap(const DeviceArray2D<float4>& vsrc, const DeviceArray2D<float4>& nsrc, DeviceArray2D<float4>& vdst,
               DeviceArray2D<float4>& ndst);
void ICPStep(DeviceArray2D<float4>& nextVMap, DeviceArray2D<float4>& lastVMap, DeviceArray2D<float4>& nextNMap,
             DeviceArray2D<float4>& lastNMap, Matrix3f Rcurr, float3 tcurr, Matrix3f Rlast, Matrix3f RlastInv,
             float3 tlast, Intrinsics K, DeviceArray2D<float>& sum, DeviceArray<float>& out, float* residual,
             double* matrixA_host, double* vectorB_host);
void RGBStep(const DeviceArray2D<unsigned char>& nextImage, const DeviceArray2D<unsigned char>& lastImage,
             const DeviceArray2D<float4>& nextVMap, const DeviceArray2D<float4>& lastVMap,
             const DeviceArray2D<short>& dIdx, const DeviceArray2D<short>& dIdy, Matrix3f Rcurr, Matrix3f RcurrInv,
             Matrix3f Rlast, Matrix3f RlastInv, float3 tcurr, float3 tlast, Intrinsics K, DeviceArray2D<float>& sum,
             DeviceArray<float>& out, DeviceArray2D<int>& sumRes, DeviceArray<int>& outRes, float* residual,
             double* matrixA_host, double* vectorB_host);
void BuildAdjacencyMatrix(DeviceArray2D<float>& adjacencyMatrix, DeviceArray<SURF>& frameKeys,
                          DeviceArray<SURF>& mapKeys, DeviceArray<float>& dist);
