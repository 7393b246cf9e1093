// The keypoint map: a bucketed hash of SURF records keyed by 1 cm grid cells, and the
// adjacency (consistency) matrix of matched keypoint pairs.
//
// Reference: Mapping/DeviceMap.h KeyMap (:163-185), DeviceMap.cu KeyMap::Hash / FindKey /
// InsertKey / ResetKeys (:203-296), Mapping/FuseMap.cu ResetKeyPoints / CollectKeyPoints /
// InsertKeyPoints (:366-472), Tracking/KeyPoints.cu BuildAdjacencyMatrix (:12-77).
//
// Inserts are applied as if one key at a time in index order: each round, every key still
// waiting claims its bucket with an atomic minimum of its index, and the owner of each bucket
// inserts (buckets are independent); the others wait for the next round.  The reference's
// per-bucket try-lock drops a key that meets another in its bucket in the same launch.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <string.h>

#include "fu_api.h"

using namespace fu;

namespace {

struct alignas(16) Surf {  // fu_surf / SURF: valid, pos, normal, descriptor (288 bytes)
    unsigned char valid;
    unsigned char pad[3];
    float pos[3];
    float normal[4];
    float desc[64];
};
static_assert(sizeof(Surf) == sizeof(fu_surf) && sizeof(Surf) == 288, "SURF layout");

// KeyMap::Hash (DeviceMap.cu:206-214): the signed remainder of the wrapped products' xor
__device__ __forceinline__ int key_hash(i3 p, int max_keys) {
    const int h = (int)(((uint32_t)p.x * 73856093u) ^ ((uint32_t)p.y * 19349669u) ^ ((uint32_t)p.z * 83492791u));
    int r = h % max_keys;
    return r < 0 ? r + max_keys : r;
}
// make_int3(pos / GridSize)
__device__ __forceinline__ i3 key_cell(const float* pos, float grid) {
    return i3{f2i(pos[0] / grid), f2i(pos[1] / grid), f2i(pos[2] / grid)};
}

struct KeyArgs {
    Surf* keys;        // the map: max_keys * nbuckets records
    int* mutex;        // KeyMap::Mutex (reset to -1; the inserts here do not lock)
    int* owner;        // per bucket: INT_MAX, or the lowest index of a waiting key this round
    int max_keys, nbuckets;
    float grid;
};

// ResetKeys (DeviceMap.cu:289-296)
__global__ void k_reset_keys(KeyArgs a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a.mutex[i] = -1;
    a.keys[i].valid = 0;
    if (i < a.max_keys) a.owner[i] = INT_MAX;
}

// one round of inserts: the claim, then the owners' inserts
__global__ void k_insert_claim(KeyArgs a, const Surf* in, const unsigned char* done, int* bucket, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || done[i]) return;
    const int b = key_hash(key_cell(in[i].pos, a.grid), a.max_keys);
    bucket[i] = b;
    atomicMin(&a.owner[b], i);
}
__global__ void k_insert_apply(KeyArgs a, Surf* in, int* index, unsigned char* done, const int* bucket, int n,
                               unsigned* left) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || done[i]) return;
    const int b = bucket[i];
    if (a.owner[b] != i) {
        atomicAdd(left, 1u);
        return;
    }
    // KeyMap::FindKey(pos, first, buck, hashIndex) and InsertKey
    const i3 p = key_cell(in[i].pos, a.grid);
    int first = -1, found = -1;
    for (int j = 0; j < a.nbuckets; ++j) {
        const int s = b * a.nbuckets + j;
        const Surf& k = a.keys[s];
        if (!k.valid && first == -1) first = s;
        if (k.valid && eqi(key_cell(k.pos, a.grid), p)) {
            found = s;
            break;
        }
    }
    if (found >= 0) {
        index[i] = found;
        in[i].pos[0] = a.keys[found].pos[0];
        in[i].pos[1] = a.keys[found].pos[1];
        in[i].pos[2] = a.keys[found].pos[2];
    } else if (first != -1) {
        index[i] = first;
        const uint4* src = reinterpret_cast<const uint4*>(&in[i]);
        uint4* dst = reinterpret_cast<uint4*>(&a.keys[first]);
        for (int q = 0; q < (int)(sizeof(Surf) / 16); ++q) dst[q] = src[q];
    }
    done[i] = 1;
    a.owner[b] = INT_MAX;
}

// CollectKeys (FuseMap.cu:384-406) in slot order: valid key j (0-based over the slots) goes to
// out[j] when 0 < j < cap (the reference copies only at offsets > 0).  Three passes: valid
// counts per 1024-slot group, their exclusive prefix (one group), the scatter.
constexpr int kCollect = 1024;
__global__ void k_collect_count(const Surf* keys, int n, unsigned* counts) {
    __shared__ unsigned c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    const int i = blockIdx.x * kCollect + threadIdx.x;
    const bool v = i < n && keys[i].valid;
    const unsigned long long m = __ballot(v);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(&c, (unsigned)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}
// the exclusive prefix of the group counts, in place, and their total: one work-group of
// kPrefixLanes lanes, lane t owning a run of consecutive counts, the runs' sums scanned across
// the work-group (integer sums: the same as one lane walking the counts in order)
constexpr int kPrefixLanes = 1024;
__global__ __launch_bounds__(kPrefixLanes) void k_collect_prefix(unsigned* counts, int groups, unsigned* total) {
    __shared__ unsigned wsum[kPrefixLanes / 64];
    const int per = (groups + kPrefixLanes - 1) / kPrefixLanes;
    const int b = min(groups, (int)threadIdx.x * per), e = min(groups, b + per);
    unsigned own = 0;
    for (int g = b; g < e; ++g) own += counts[g];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    unsigned x = own;  // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    unsigned before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kPrefixLanes / 64; ++k) {
        const unsigned c = wsum[k];
        before += k < w ? c : 0u;
        all += c;
    }
    unsigned run = before + x - own;
    for (int g = b; g < e; ++g) {
        const unsigned c = counts[g];
        counts[g] = run;
        run += c;
    }
    if (threadIdx.x == 0) *total = all;
}
__global__ void k_collect_scatter(const Surf* keys, int n, const unsigned* base, Surf* out, unsigned cap) {
    __shared__ unsigned wave_base[kCollect / 64];
    const int i = blockIdx.x * kCollect + threadIdx.x;
    const bool v = i < n && keys[i].valid;
    const unsigned long long m = __ballot(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wave_base[w] = (unsigned)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned s = 0;
        for (int k = 0; k < kCollect / 64; ++k) {
            const unsigned c = wave_base[k];
            wave_base[k] = s;
            s += c;
        }
    }
    __syncthreads();
    if (!v) return;
    const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    const unsigned j = base[blockIdx.x] + wave_base[w] + below;
    if (j > 0 && j < cap) {
        const uint4* src = reinterpret_cast<const uint4*>(&keys[i]);
        uint4* dst = reinterpret_cast<uint4*>(&out[j]);
        for (int q = 0; q < (int)(sizeof(Surf) / 16); ++q) dst[q] = src[q];
    }
}

__device__ __forceinline__ float clamp1(float a) {  // KeyPoints.cu:6-10
    a = a > -1.f ? a : -1.f;
    a = a < 1.f ? a : 1.f;
    return a;
}
__device__ __forceinline__ f3 spos(const Surf& s) { return mk3(s.pos[0], s.pos[1], s.pos[2]); }
__device__ __forceinline__ f4 snrm(const Surf& s) { return mk4(s.normal[0], s.normal[1], s.normal[2], s.normal[3]); }
__device__ __forceinline__ float dot4f(f4 a, f4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }

// BuildAdjecencyMatrixKernel (KeyPoints.cu:12-57), one lane per entry of the n x n matrix
__global__ void k_adjacency(const Surf* frame, const Surf* map, const float* dist, int n, float* out) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= n || y >= n) return;
    float score = 0;
    if (x == y) {
        score = expf(-dist[x]);
    } else {
        const Surf &m0 = map[x], &m1 = map[y], &f0 = frame[x], &f1 = frame[y];
        const float d00 = norm(sub(spos(f0), spos(f1)));
        const float d01 = norm(sub(spos(m0), spos(m1)));
        const f3 e10 = normalised(sub(spos(f0), spos(f1)));
        const f3 e11 = normalised(sub(spos(m0), spos(m1)));
        const f4 d10 = mk4(e10.x, e10.y, e10.z, 1.0f), d11 = mk4(e11.x, e11.y, e11.z, 1.0f);
        if ((double)d00 <= 1e-2 || (double)d01 <= 1e-2) {
            score = 0;
        } else {
            const float alpha00 = acosf(clamp1(dot4f(snrm(f0), snrm(f1))));
            const float beta00 = acosf(clamp1(dot4f(d10, snrm(f0))));
            const float gamma00 = acosf(clamp1(dot4f(d10, snrm(f1))));
            const float alpha01 = acosf(clamp1(dot4f(snrm(m0), snrm(m1))));
            const float beta01 = acosf(clamp1(dot4f(d11, snrm(m0))));
            const float gamma01 = acosf(clamp1(dot4f(d11, snrm(m1))));
            score = expf(-(fabsf(d00 - d01) + fabsf(alpha00 - alpha01) + fabsf(beta00 - beta01) +
                           fabsf(gamma00 - gamma01)));
        }
    }
    if (isnanf_(score)) score = 0;
    out[(size_t)y * n + x] = score;
}

}  // namespace

struct fu_keymap {
    fu_keymap_params p;
    KeyArgs a;
    unsigned* aux = nullptr;    // [0] keys left in a round, [1] collect total, then group counts
    unsigned char* done = nullptr;  // per key of an insert call, grown on demand
    int* bucket = nullptr;
    int cap = 0;
};

#define KM_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return api_error(FU_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

extern "C" void fu_keymap_params_default(fu_keymap_params* p) {
    // KeyMap's constants (DeviceMap.h:165-170)
    p->grid_size = 0.01f;
    p->max_keys = 100000;
    p->n_buckets = 5;
}

extern "C" fu_status fu_keymap_destroy(fu_keymap* k) {
    if (!k) return FU_OK;
    (void)hipFree(k->a.keys);
    (void)hipFree(k->a.mutex);
    (void)hipFree(k->a.owner);
    (void)hipFree(k->aux);
    (void)hipFree(k->done);
    (void)hipFree(k->bucket);
    delete k;
    return FU_OK;
}

static int n_slots(const fu_keymap* k) { return k->a.max_keys * k->a.nbuckets; }
static int n_groups(const fu_keymap* k) { return (n_slots(k) + kCollect - 1) / kCollect; }

extern "C" fu_status fu_keymap_create(const fu_keymap_params* p, fu_keymap** out) {
    if (!p || !out) return api_error(FU_ERR_ARG, "fu_keymap_create: null argument");
    *out = nullptr;
    if (p->max_keys < 1 || p->n_buckets < 1 || (long long)p->max_keys * p->n_buckets > (1ll << 28) || !(p->grid_size > 0))
        return api_error(FU_ERR_ARG, "fu_keymap_create: %u keys x %u buckets, grid %g", p->max_keys, p->n_buckets,
                         (double)p->grid_size);
    fu_keymap* k = new fu_keymap();
    k->p = *p;
    k->a.max_keys = (int)p->max_keys;
    k->a.nbuckets = (int)p->n_buckets;
    k->a.grid = p->grid_size;
    const size_t n = (size_t)n_slots(k);
    if (hipMalloc(&k->a.keys, n * sizeof(Surf)) != hipSuccess || hipMalloc(&k->a.mutex, n * 4) != hipSuccess ||
        hipMalloc(&k->a.owner, (size_t)p->max_keys * 4) != hipSuccess ||
        hipMalloc(&k->aux, (2 + (size_t)n_groups(k)) * 4) != hipSuccess) {
        (void)hipGetLastError();
        fu_keymap_destroy(k);
        return api_error(FU_ERR_NOMEM, "fu_keymap_create: allocation failed");
    }
    // records start zeroed (ResetKeys clears only the valid flags, so a record's other bytes
    // keep what the last key there left)
    if (hipMemset(k->a.keys, 0, n * sizeof(Surf)) != hipSuccess) {
        (void)hipGetLastError();
        fu_keymap_destroy(k);
        return api_error(FU_ERR_HIP, "fu_keymap_create: clearing the records failed");
    }
    fu_status st = fu_reset_key_points(k, nullptr);
    if (st) {
        fu_keymap_destroy(k);
        return st;
    }
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        fu_keymap_destroy(k);
        return api_error(FU_ERR_HIP, "fu_keymap_create: reset failed");
    }
    *out = k;
    return FU_OK;
}

extern "C" fu_status fu_reset_key_points(fu_keymap* k, fu_stream s) {
    if (!k) return api_error(FU_ERR_ARG, "fu_reset_key_points: null key map");
    const int n = n_slots(k);
    hipLaunchKernelGGL(k_reset_keys, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)s, k->a, n);
    KM_HIP(hipGetLastError());
    return FU_OK;
}

extern "C" fu_status fu_insert_key_points(fu_keymap* k, fu_surf* keys, int32_t* index, uint32_t n, fu_stream s) {
    if (!k || (n && (!keys || !index))) return api_error(FU_ERR_ARG, "fu_insert_key_points: null argument");
    if (n == 0) return FU_OK;  // InsertKeyPoints returns at once (FuseMap.cu:459-460)
    if (n > (1u << 24)) return api_error(FU_ERR_ARG, "fu_insert_key_points: %u keys (at most 2^24)", n);
    hipStream_t hs = (hipStream_t)s;
    if ((int)n > k->cap) {
        KM_HIP(hipStreamSynchronize(hs));
        (void)hipFree(k->done);
        (void)hipFree(k->bucket);
        k->done = nullptr;
        k->bucket = nullptr;
        k->cap = 0;
        if (hipMalloc(&k->done, n) != hipSuccess || hipMalloc(&k->bucket, (size_t)n * 4) != hipSuccess) {
            (void)hipGetLastError();
            return api_error(FU_ERR_NOMEM, "fu_insert_key_points: scratch allocation failed");
        }
        k->cap = (int)n;
    }
    KM_HIP(hipMemsetAsync(k->done, 0, n, hs));
    const dim3 grid((n + 255) / 256);
    // rounds until every key is placed; each round places at least the lowest waiting index
    for (uint32_t round = 0; round <= n; ++round) {
        KM_HIP(hipMemsetAsync(k->aux, 0, 4, hs));
        hipLaunchKernelGGL(k_insert_claim, grid, dim3(256), 0, hs, k->a, (const Surf*)keys, k->done, k->bucket, (int)n);
        hipLaunchKernelGGL(k_insert_apply, grid, dim3(256), 0, hs, k->a, (Surf*)keys, index, k->done, k->bucket, (int)n,
                           k->aux);
        KM_HIP(hipGetLastError());
        unsigned left = 0;
        KM_HIP(hipMemcpyAsync(&left, k->aux, 4, hipMemcpyDeviceToHost, hs));
        KM_HIP(hipStreamSynchronize(hs));
        if (left == 0) return FU_OK;
    }
    return api_error(FU_ERR_STATE, "fu_insert_key_points: inserts did not finish");
}

extern "C" fu_status fu_collect_key_points(fu_keymap* k, fu_surf* out, uint32_t cap, uint32_t* n_keys, fu_stream s) {
    if (!k || !n_keys || (cap && !out)) return api_error(FU_ERR_ARG, "fu_collect_key_points: null argument");
    hipStream_t hs = (hipStream_t)s;
    const int n = n_slots(k), g = n_groups(k);
    hipLaunchKernelGGL(k_collect_count, dim3(g), dim3(kCollect), 0, hs, (const Surf*)k->a.keys, n, k->aux + 2);
    hipLaunchKernelGGL(k_collect_prefix, dim3(1), dim3(kPrefixLanes), 0, hs, k->aux + 2, g, k->aux + 1);
    if (cap) hipLaunchKernelGGL(k_collect_scatter, dim3(g), dim3(kCollect), 0, hs, (const Surf*)k->a.keys, n, k->aux + 2,
                                (Surf*)out, cap);
    KM_HIP(hipGetLastError());
    KM_HIP(hipMemcpyAsync(n_keys, k->aux + 1, 4, hipMemcpyDeviceToHost, hs));
    KM_HIP(hipStreamSynchronize(hs));
    return FU_OK;
}

extern "C" fu_status fu_keymap_read(fu_keymap* k, fu_surf* host, fu_stream s) {
    if (!k || !host) return api_error(FU_ERR_ARG, "fu_keymap_read: null argument");
    KM_HIP(hipMemcpyAsync(host, k->a.keys, (size_t)n_slots(k) * sizeof(Surf), hipMemcpyDeviceToHost, (hipStream_t)s));
    KM_HIP(hipStreamSynchronize((hipStream_t)s));
    return FU_OK;
}

extern "C" fu_status fu_build_adjacency_matrix(const fu_surf* frame_keys, const fu_surf* map_keys, const float* dist,
                                               int32_t n, float* matrix, fu_stream s) {
    if (n < 0 || (n && (!frame_keys || !map_keys || !dist || !matrix)))
        return api_error(FU_ERR_ARG, "fu_build_adjacency_matrix: null argument or negative size");
    if (n > 46340) return api_error(FU_ERR_ARG, "fu_build_adjacency_matrix: %d pairs (the matrix index is 32-bit)", n);
    if (n == 0) return FU_OK;
    const dim3 grid((n + 15) / 16, (n + 15) / 16);
    hipLaunchKernelGGL(k_adjacency, grid, dim3(256), 0, (hipStream_t)s, (const Surf*)frame_keys, (const Surf*)map_keys,
                       dist, n, matrix);
    KM_HIP(hipGetLastError());
    return FU_OK;
}
