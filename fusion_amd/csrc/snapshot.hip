// On-disk formats: map and session snapshots, and the PLY triangle-soup writer.
//
// Reference: Core/System.cc System::WriteMapToDisk / ReadMapFromDisk (:195-271) dump the
// whole device map raw (every pool block, both entry arrays, the locks and the keypoint map,
// tens of GB at the reference's sizes).  A snapshot here keeps what defines the map and
// nothing else: the used hash entries, the allocated blocks' keys and voxels (the pool is
// filled from index 0 up, so they are one prefix of it) and the counters, in a versioned
// little-endian file with a checksum.  Loading rebuilds the device map with the same entries
// at the same indices and the same blocks at the same pool slots, so a map continues exactly
// as if it had never left the GPU.  A session snapshot adds the tracking state, the model
// maps the next frame is tracked against and (photometric tracking) the last frame's
// intensity pyramid.
//
// System::WriteMeshToDisk (:140-193) writes an ASCII PLY triangle soup; fu_mesh_write_ply
// writes the same text, or the same elements in binary.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <vector>

#include "fu_api.h"

using namespace fu;

namespace {

constexpr char kMapMagic[8] = {'F', 'U', 'S', 'N', 'M', 'A', 'P', '\0'};
constexpr char kSysMagic[8] = {'F', 'U', 'S', 'N', 'S', 'Y', 'S', '\0'};
constexpr uint32_t kVersion = 1;
constexpr size_t kChunk = 64ull << 20;  // staging size of large transfers

struct Fnv {  // FNV-1a, 64 bit
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* b = (const unsigned char*)p;
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
};

struct Out {
    FILE* f = nullptr;
    Fnv sum;
    bool ok = true;
    void put(const void* p, size_t n) {
        if (!ok || n == 0) return;
        sum.add(p, n);
        ok = fwrite(p, 1, n, f) == n;
    }
    template <class T>
    void val(T v) { put(&v, sizeof(v)); }
    void checksum() {
        const uint64_t h = sum.h;
        ok = ok && fwrite(&h, 1, 8, f) == 8;
        sum = Fnv();
    }
};

struct In {
    FILE* f = nullptr;
    Fnv sum;
    bool ok = true;
    void get(void* p, size_t n) {
        if (!ok || n == 0) return;
        ok = fread(p, 1, n, f) == n;
        if (ok) sum.add(p, n);
    }
    template <class T>
    T val() {
        T v{};
        get(&v, sizeof(v));
        return v;
    }
    bool checksum() {
        uint64_t h = 0;
        const uint64_t want = sum.h;
        ok = ok && fread(&h, 1, 8, f) == 8;
        sum = Fnv();
        return ok && h == want;
    }
};

struct File {
    FILE* f;
    explicit File(FILE* x) : f(x) {}
    ~File() {
        if (f) fclose(f);
    }
};

struct Pinned {
    void* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

// device <-> file through a pinned staging buffer, in chunks
bool dev_to_file(Out& o, const void* dev, size_t n, Pinned& buf) {
    const unsigned char* d = (const unsigned char*)dev;
    for (size_t off = 0; off < n && o.ok; off += kChunk) {
        const size_t c = n - off < kChunk ? n - off : kChunk;
        if (hipMemcpy(buf.p, d + off, c, hipMemcpyDeviceToHost) != hipSuccess) return false;
        o.put(buf.p, c);
    }
    return o.ok;
}
bool file_to_dev(In& in, void* dev, size_t n, Pinned& buf) {
    unsigned char* d = (unsigned char*)dev;
    for (size_t off = 0; off < n && in.ok; off += kChunk) {
        const size_t c = n - off < kChunk ? n - off : kChunk;
        in.get(buf.p, c);
        if (!in.ok || hipMemcpy(d + off, buf.p, c, hipMemcpyHostToDevice) != hipSuccess) return false;
    }
    return in.ok;
}

#define IOCHK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) return api_error(FU_ERR_ARG, __VA_ARGS__); \
    } while (0)
#define HIPOK(call)                                                                                \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return api_error(FU_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

// ---- map ------------------------------------------------------------------------------------
fu_status write_map(fu_map* m, Out& o) {
    const fu_map_params& p = m->p;
    const MapView& v = m->v;
    unsigned ctr[C_COUNT];
    HIPOK(hipMemcpy(ctr, v.ctr, sizeof(ctr), hipMemcpyDeviceToHost));
    const uint32_t n_alloc = ctr[C_ALLOC] < p.num_blocks ? ctr[C_ALLOC] : p.num_blocks;
    const uint32_t bb = (uint32_t)vf_block_bytes(v.fmt);
    const size_t ne = (size_t)p.num_buckets + p.num_excess;
    std::vector<HEntry> ent(ne);
    HIPOK(hipMemcpy(ent.data(), v.entries, ne * sizeof(HEntry), hipMemcpyDeviceToHost));
    uint32_t used = 0;
    for (const HEntry& e : ent) used += e.key != kEmptyKey;
    o.put(kMapMagic, 8);
    o.val<uint32_t>(kVersion);
    o.val<uint32_t>(p.num_buckets);
    o.val<uint32_t>(p.num_excess);
    o.val<uint32_t>(p.num_blocks);
    o.val<uint32_t>(p.voxel_format);
    o.val<uint32_t>(p.max_weight);
    o.val<float>(p.voxel_size);
    o.val<float>(p.trunc_dist);
    o.val<float>(p.depth_min);
    o.val<float>(p.depth_max);
    o.put(ctr, sizeof(ctr));
    o.val<uint32_t>(bb);
    o.val<uint32_t>(n_alloc);
    o.val<uint32_t>(used);
    for (size_t i = 0; i < ne; ++i)
        if (ent[i].key != kEmptyKey) {
            o.val<uint32_t>((uint32_t)i);
            o.val<uint64_t>(ent[i].key);
            o.val<int32_t>(ent[i].ptr);
            o.val<int32_t>(ent[i].next);
        }
    std::vector<unsigned long long> keys(n_alloc);
    if (n_alloc) HIPOK(hipMemcpy(keys.data(), v.block_key, n_alloc * 8ull, hipMemcpyDeviceToHost));
    o.put(keys.data(), n_alloc * 8ull);
    Pinned buf;
    HIPOK(hipHostMalloc(&buf.p, kChunk, hipHostMallocDefault));
    if (!dev_to_file(o, v.pool, (size_t)n_alloc * bb, buf)) return api_error(FU_ERR_HIP, "map snapshot: pool copy failed");
    o.checksum();
    return o.ok ? FU_OK : api_error(FU_ERR_ARG, "map snapshot: write failed");
}

// A map snapshot read and checked on the host before anything reaches the device: the
// checksum over the whole map part, then the structure the kernels rely on (every chain walk
// ends, every pool index is in range), so a damaged or inconsistent file is refused with the
// device map untouched.
struct MapImage {
    unsigned ctr[C_COUNT];
    uint32_t n_alloc = 0;
    std::vector<HEntry> ent;
    std::vector<unsigned long long> keys;
    off_t pool_off = 0;  // file offset of the pool bytes
    size_t pool_bytes = 0;
};

fu_status snapshot_error(const char* what) { return api_error(FU_ERR_ARG, "map snapshot: %s", what); }

// the consistency the device's lookups and allocation depend on
const char* check_map_image(const fu_map_params& p, const MapImage& img) {
    const size_t nb = p.num_buckets, nx = p.num_excess, ne = nb + nx;
    const uint32_t n_alloc = img.n_alloc;
    if (n_alloc != (img.ctr[C_ALLOC] < p.num_blocks ? img.ctr[C_ALLOC] : p.num_blocks))
        return "allocated-block count differs from the allocation counter";
    const size_t n_excess = img.ctr[C_EXCESS] < nx ? img.ctr[C_EXCESS] : nx;
    for (size_t i = 0; i < ne; ++i) {
        const HEntry& e = img.ent[i];
        if (e.key == kEmptyKey) {
            if (e.ptr != -1 || e.next != 0) return "empty entry with a block or a link";
            continue;
        }
        if (e.key >> 63) return "entry key out of range";
        if (i >= nb && i - nb >= n_excess) return "excess entry beyond the excess counter";
        if (e.ptr != -1) {
            if (e.ptr < 0 || (e.ptr & 511) || (uint32_t)(e.ptr >> 9) >= n_alloc) return "entry block index out of range";
            if (img.keys[e.ptr >> 9] != e.key) return "entry and block key disagree";
        }
        if (e.next < 0 || (size_t)e.next > n_excess) return "excess link out of range";
        if (e.next > 0 && img.ent[nb + e.next - 1].key == kEmptyKey) return "link to an empty excess entry";
    }
    // every bucket's chain: acyclic, each excess entry on exactly one chain, keys in their bucket
    std::vector<unsigned char> seen(nx, 0);
    for (size_t b = 0; b < nb; ++b) {
        const HEntry* e = &img.ent[b];
        if (e->key == kEmptyKey) continue;
        while (true) {
            if (block_hash_raw(unpack_key(e->key)) % p.num_buckets != b) return "entry in another key's bucket";
            if (e->next <= 0) break;
            const size_t k = (size_t)e->next - 1;
            if (seen[k]) return "excess chains share an entry or form a cycle";
            seen[k] = 1;
            e = &img.ent[nb + k];
        }
    }
    for (size_t k = 0; k < nx; ++k)
        if (img.ent[nb + k].key != kEmptyKey && !seen[k]) return "excess entry on no chain";
    // each allocated block's key maps to that block (holes are allowed: a block whose excess
    // slot could not be had)
    size_t mapped = 0;
    for (size_t i = 0; i < ne; ++i) mapped += img.ent[i].key != kEmptyKey && img.ent[i].ptr >= 0;
    std::vector<unsigned long long> keyed;
    for (uint32_t j = 0; j < n_alloc; ++j)
        if (img.keys[j] != kEmptyKey) keyed.push_back(img.keys[j]);
    if (mapped != keyed.size()) return "allocated blocks and entries disagree";
    std::sort(keyed.begin(), keyed.end());
    if (std::adjacent_find(keyed.begin(), keyed.end()) != keyed.end()) return "a block key allocated twice";
    return nullptr;
}

// header, counters, entries and keys into `img`; the pool bytes go through the checksum only
fu_status parse_map(const fu_map_params& p, In& in, MapImage& img) {
    char magic[8];
    in.get(magic, 8);
    if (!in.ok || memcmp(magic, kMapMagic, 8) != 0) return snapshot_error("not a map snapshot");
    const uint32_t ver = in.val<uint32_t>();
    if (!in.ok || ver != kVersion) return api_error(FU_ERR_ARG, "map snapshot: version %u (this library reads %u)", ver, kVersion);
    const uint32_t nb = in.val<uint32_t>(), nx = in.val<uint32_t>(), nk = in.val<uint32_t>();
    const uint32_t fmt = in.val<uint32_t>(), mw = in.val<uint32_t>();
    const float vs = in.val<float>(), tr = in.val<float>(), dmin = in.val<float>(), dmax = in.val<float>();
    if (!in.ok) return snapshot_error("truncated header");
    if (!(nb == p.num_buckets && nx == p.num_excess && nk == p.num_blocks && fmt == p.voxel_format &&
          mw == p.max_weight && vs == p.voxel_size && tr == p.trunc_dist && dmin == p.depth_min && dmax == p.depth_max))
        return api_error(FU_ERR_ARG,
                         "map snapshot: written by a map of other parameters (buckets %u/%u, excess %u/%u, blocks "
                         "%u/%u, format %u/%u)", nb, p.num_buckets, nx, p.num_excess, nk, p.num_blocks, fmt,
                         p.voxel_format);
    in.get(img.ctr, sizeof(img.ctr));
    const uint32_t bb = in.val<uint32_t>(), n_alloc = in.val<uint32_t>(), used = in.val<uint32_t>();
    const size_t ne = (size_t)p.num_buckets + p.num_excess;
    if (!in.ok || bb != (uint32_t)vf_block_bytes((int)p.voxel_format) || n_alloc > p.num_blocks || used > ne)
        return snapshot_error("inconsistent sizes");
    img.n_alloc = n_alloc;
    img.ent.assign(ne, HEntry{kEmptyKey, -1, 0});
    std::vector<unsigned char> taken(ne, 0);
    for (uint32_t k = 0; k < used && in.ok; ++k) {
        const uint32_t i = in.val<uint32_t>();
        HEntry e;
        e.key = in.val<uint64_t>();
        e.ptr = in.val<int32_t>();
        e.next = in.val<int32_t>();
        if (!in.ok) break;
        if (i >= ne || taken[i] || e.key == kEmptyKey) return snapshot_error("bad or repeated entry index");
        taken[i] = 1;
        img.ent[i] = e;
    }
    img.keys.resize(n_alloc);
    in.get(img.keys.data(), n_alloc * 8ull);
    if (!in.ok) return snapshot_error("truncated");
    img.pool_off = ftello(in.f);
    img.pool_bytes = (size_t)n_alloc * bb;
    std::vector<unsigned char> buf(img.pool_bytes < kChunk ? img.pool_bytes : kChunk);
    for (size_t off = 0; off < img.pool_bytes && in.ok; off += buf.size()) {
        const size_t c = img.pool_bytes - off < buf.size() ? img.pool_bytes - off : buf.size();
        in.get(buf.data(), c);
    }
    if (!in.ok || !in.checksum()) return snapshot_error("truncated or corrupt (checksum)");
    if (const char* why = check_map_image(p, img)) return snapshot_error(why);
    return FU_OK;
}

// the checked image onto the device: entries, keys, the occupancy bitmap (a function of the
// keys), the pool (read again from the file) and the map-wide counters; the per-frame
// counters start at zero as after a reset
fu_status commit_map(fu_map* m, FILE* f, const MapImage& img, hipStream_t s) {
    MapView& v = m->v;
    std::vector<unsigned> occ((v.occ_mask + 1) / 32, 0u);
    for (unsigned long long key : img.keys)
        if (key != kEmptyKey) {
            const unsigned hb = occ_bit(v, block_hash_raw(unpack_key(key)));
            occ[hb >> 5] |= 1u << (hb & 31);
        }
    Pinned buf;
    HIPOK(hipHostMalloc(&buf.p, kChunk, hipHostMallocDefault));
    if (fseeko(f, img.pool_off, SEEK_SET) != 0) return snapshot_error("cannot re-read the pool");
    HIPOK(hipStreamSynchronize(s));
    launch_reset_map(v, s);
    HIPOK(hipStreamSynchronize(s));
    HIPOK(hipMemcpy(v.entries, img.ent.data(), img.ent.size() * sizeof(HEntry), hipMemcpyHostToDevice));
    if (img.n_alloc) HIPOK(hipMemcpy(v.block_key, img.keys.data(), img.n_alloc * 8ull, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(v.occ, occ.data(), occ.size() * 4, hipMemcpyHostToDevice));
    In in;
    in.f = f;
    if (!file_to_dev(in, v.pool, img.pool_bytes, buf)) {
        launch_reset_map(v, s);
        (void)hipStreamSynchronize(s);
        return snapshot_error("the file changed while it was loaded (the map was left empty)");
    }
    unsigned c[C_COUNT] = {};
    c[C_ALLOC] = img.ctr[C_ALLOC];
    c[C_EXCESS] = img.ctr[C_EXCESS];
    c[C_FLAGS] = img.ctr[C_FLAGS];
    HIPOK(hipMemcpy(v.ctr, c, sizeof(c), hipMemcpyHostToDevice));
    return FU_OK;
}

FILE* open_file(const char* path, const char* mode) { return path ? fopen(path, mode) : nullptr; }

}  // namespace

extern "C" fu_status fu_map_save(fu_map* m, const char* path, fu_stream s) {
    if (!m || !path) return api_error(FU_ERR_ARG, "fu_map_save: null argument");
    HIPOK(hipStreamSynchronize((hipStream_t)s));
    File f(open_file(path, "wb"));
    if (!f.f) return api_error(FU_ERR_ARG, "fu_map_save: cannot open %s", path);
    Out o;
    o.f = f.f;
    return write_map(m, o);
}

extern "C" fu_status fu_map_load(fu_map* m, const char* path, fu_stream s) {
    if (!m || !path) return api_error(FU_ERR_ARG, "fu_map_load: null argument");
    File f(open_file(path, "rb"));
    if (!f.f) return api_error(FU_ERR_ARG, "fu_map_load: cannot open %s", path);
    In in;
    in.f = f.f;
    MapImage img;
    fu_status st = parse_map(m->p, in, img);
    return st ? st : commit_map(m, f.f, img, (hipStream_t)s);
}

// the checks of fu_map_load on the host alone
extern "C" fu_status fu_map_check_file(const fu_map_params* p, const char* path) {
    if (!p || !path) return api_error(FU_ERR_ARG, "fu_map_check_file: null argument");
    File f(open_file(path, "rb"));
    if (!f.f) return api_error(FU_ERR_ARG, "fu_map_check_file: cannot open %s", path);
    In in;
    in.f = f.f;
    MapImage img;
    return parse_map(*p, in, img);
}

// ---- session --------------------------------------------------------------------------------
extern "C" fu_status fu_system_save(fu_system* sys, const char* path, fu_stream s) {
    if (!sys || !path) return api_error(FU_ERR_ARG, "fu_system_save: null argument");
    HIPOK(hipStreamSynchronize((hipStream_t)s));
    File f(open_file(path, "wb"));
    if (!f.f) return api_error(FU_ERR_ARG, "fu_system_save: cannot open %s", path);
    const fu_system_params& p = sys->p;
    Out o;
    o.f = f.f;
    o.put(kSysMagic, 8);
    o.val<uint32_t>(kVersion);
    o.val<int32_t>(p.cols);
    o.val<int32_t>(p.rows);
    o.val<int32_t>(p.use_rgb);
    o.val<float>(p.fx);
    o.val<float>(p.fy);
    o.val<float>(p.cx);
    o.val<float>(p.cy);
    o.val<float>(p.depth_scale);
    o.val<float>(p.depth_cutoff);
    o.val<uint32_t>(sys->frame);
    TrackState st;
    HIPOK(hipMemcpy(&st, sys->trk->st, sizeof(st), hipMemcpyDeviceToHost));
    o.val<uint32_t>((uint32_t)sizeof(st));
    o.put(&st, sizeof(st));
    Pinned buf;
    HIPOK(hipHostMalloc(&buf.p, kChunk, hipHostMallocDefault));
    const size_t n0 = (size_t)p.cols * p.rows;
    bool ok = dev_to_file(o, sys->b.model_vmap[0], n0 * 16, buf) && dev_to_file(o, sys->b.model_nmap[0], n0 * 16, buf);
    if (p.use_rgb)
        for (int l = 0; l < 3 && ok; ++l) ok = dev_to_file(o, sys->b.last_image[l], (size_t)sys->lc[l] * sys->lr[l], buf);
    if (!ok) return api_error(FU_ERR_HIP, "fu_system_save: copy failed");
    o.checksum();
    return write_map(sys->map, o);
}

extern "C" fu_status fu_system_load(fu_system* sys, const char* path, fu_stream s) {
    if (!sys || !path) return api_error(FU_ERR_ARG, "fu_system_load: null argument");
    File f(open_file(path, "rb"));
    if (!f.f) return api_error(FU_ERR_ARG, "fu_system_load: cannot open %s", path);
    const fu_system_params& p = sys->p;
    In in;
    in.f = f.f;
    char magic[8];
    in.get(magic, 8);
    IOCHK(in.ok && memcmp(magic, kSysMagic, 8) == 0, "fu_system_load: not a session snapshot");
    const uint32_t ver = in.val<uint32_t>();
    IOCHK(in.ok && ver == kVersion, "fu_system_load: version %u (this library reads %u)", ver, kVersion);
    const int cols = in.val<int32_t>(), rows = in.val<int32_t>(), rgb = in.val<int32_t>();
    const float fx = in.val<float>(), fy = in.val<float>(), cx = in.val<float>(), cy = in.val<float>();
    const float ds = in.val<float>(), dc = in.val<float>();
    const uint32_t frame = in.val<uint32_t>();
    const uint32_t sts = in.val<uint32_t>();
    IOCHK(in.ok && sts == sizeof(TrackState), "fu_system_load: truncated header or other tracking-state layout");
    IOCHK(cols == p.cols && rows == p.rows && rgb == p.use_rgb && fx == p.fx && fy == p.fy && cx == p.cx &&
              cy == p.cy && ds == p.depth_scale && dc == p.depth_cutoff,
          "fu_system_load: written by a session of other parameters (%dx%d, use_rgb %d)", cols, rows, rgb);
    // the session part staged on the host and the map part checked before any device write,
    // so a refused file leaves the session as it was
    TrackState st;
    in.get(&st, sizeof(st));
    const size_t n0 = (size_t)p.cols * p.rows;
    std::vector<unsigned char> vm(n0 * 16), nm(n0 * 16), img[3];
    in.get(vm.data(), vm.size());
    in.get(nm.data(), nm.size());
    if (p.use_rgb)
        for (int l = 0; l < 3; ++l) {
            img[l].resize((size_t)sys->lc[l] * sys->lr[l]);
            in.get(img[l].data(), img[l].size());
        }
    IOCHK(in.ok && in.checksum(), "fu_system_load: truncated or corrupt session part");
    MapImage mi;
    fu_status r = parse_map(sys->map->p, in, mi);
    if (r) return r;
    hipStream_t hs = (hipStream_t)s;
    if ((r = commit_map(sys->map, f.f, mi, hs))) return r;
    HIPOK(hipMemcpy(sys->b.model_vmap[0], vm.data(), vm.size(), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(sys->b.model_nmap[0], nm.data(), nm.size(), hipMemcpyHostToDevice));
    if (p.use_rgb)
        for (int l = 0; l < 3; ++l) HIPOK(hipMemcpy(sys->b.last_image[l], img[l].data(), img[l].size(), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(sys->trk->st, &st, sizeof(st), hipMemcpyHostToDevice));
    sys->frame = frame;
    sys->frame_at_load = frame;
    sys->mapping_off = st.mapping_off != 0;
    return FU_OK;
}

// ---- PLY ------------------------------------------------------------------------------------
extern "C" fu_status fu_mesh_write_ply(const char* path, const float* vertex, const float* normal, const uint8_t* color,
                                       uint32_t n_triangles, int binary, fu_stream s) {
    if (!path || (n_triangles && (!vertex || !normal || !color)))
        return api_error(FU_ERR_ARG, "fu_mesh_write_ply: null argument");
    HIPOK(hipStreamSynchronize((hipStream_t)s));
    const size_t nv = (size_t)n_triangles * 3;
    std::vector<float> vx(nv * 3), nm(nv * 3);
    std::vector<uint8_t> cl(nv * 3);
    if (nv) {
        HIPOK(hipMemcpy(vx.data(), vertex, nv * 12, hipMemcpyDefault));
        HIPOK(hipMemcpy(nm.data(), normal, nv * 12, hipMemcpyDefault));
        HIPOK(hipMemcpy(cl.data(), color, nv * 3, hipMemcpyDefault));
    }
    File f(open_file(path, "wb"));
    if (!f.f) return api_error(FU_ERR_ARG, "fu_mesh_write_ply: cannot open %s", path);
    // the header of System::WriteMeshToDisk (:151-166)
    std::string h = "ply\n";
    h += binary ? "format binary_little_endian 1.0\n" : "format ascii 1.0\n";
    h += "element vertex " + std::to_string(nv) + "\n";
    h += "property float x\nproperty float y\nproperty float z\n";
    h += "property float nx\nproperty float ny\nproperty float nz\n";
    h += "property uchar red\nproperty uchar green\nproperty uchar blue\n";
    h += "element face " + std::to_string(n_triangles) + "\n";
    h += "property list uchar uint vertex_indices\n";
    h += "end_header\n";
    bool ok = fwrite(h.data(), 1, h.size(), f.f) == h.size();
    if (binary) {
        std::vector<unsigned char> rec(nv * 27);
        for (size_t i = 0; i < nv; ++i) {
            memcpy(&rec[i * 27], &vx[i * 3], 12);
            memcpy(&rec[i * 27 + 12], &nm[i * 3], 12);
            memcpy(&rec[i * 27 + 24], &cl[i * 3], 3);
        }
        ok = ok && fwrite(rec.data(), 1, rec.size(), f.f) == rec.size();
        std::vector<unsigned char> fr((size_t)n_triangles * 13);
        for (uint32_t t = 0; t < n_triangles; ++t) {
            fr[(size_t)t * 13] = 3;
            const uint32_t idx[3] = {3 * t, 3 * t + 1, 3 * t + 2};
            memcpy(&fr[(size_t)t * 13 + 1], idx, 12);
        }
        ok = ok && fwrite(fr.data(), 1, fr.size(), f.f) == fr.size();
    } else {
        // std::ostream's default float output is printf's %g with 6 significant digits
        std::string buf;
        buf.reserve(1 << 20);
        char line[256];
        for (size_t i = 0; i < nv && ok; ++i) {
            const int n = snprintf(line, sizeof(line), "%g %g %g %g %g %g %d %d %d\n", vx[i * 3], vx[i * 3 + 1],
                                   vx[i * 3 + 2], nm[i * 3], nm[i * 3 + 1], nm[i * 3 + 2], (int)cl[i * 3],
                                   (int)cl[i * 3 + 1], (int)cl[i * 3 + 2]);
            buf.append(line, n);
            if (buf.size() > (1u << 20)) {
                ok = fwrite(buf.data(), 1, buf.size(), f.f) == buf.size();
                buf.clear();
            }
        }
        for (uint32_t t = 0; t < n_triangles && ok; ++t) {
            const int n = snprintf(line, sizeof(line), "3 %d %d %d\n", (int)t * 3, (int)t * 3 + 1, (int)t * 3 + 2);
            buf.append(line, n);
            if (buf.size() > (1u << 20)) {
                ok = fwrite(buf.data(), 1, buf.size(), f.f) == buf.size();
                buf.clear();
            }
        }
        ok = ok && fwrite(buf.data(), 1, buf.size(), f.f) == buf.size();
    }
    return ok ? FU_OK : api_error(FU_ERR_ARG, "fu_mesh_write_ply: write failed");
}
