// Library-internal definitions of the objects behind the C ABI's opaque handles, shared by
// the translation units that implement the ABI (capi.hip, snapshot.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <vector>

#include "../../include/fusion_amd.h"
#include "fu_device.h"
#include "fu_kernels.h"

struct fu_map {
    fu_map_params p;
    fu::MapView v;
    unsigned* mesh_counts = nullptr;
    unsigned* mesh_chunk = nullptr;
    std::vector<void*> allocs;
};

struct fu_tracker {
    int cols, rows;
    fu::TrackState* st = nullptr;
    float* partials = nullptr;  // 64 floats per ICP work-group
    int* rgb_rec = nullptr;     // photometric residual records, cols * rows (and the eligibility
                                // words while a level's list is built: 8 B per 64 pixels)
    int* rgb_corr = nullptr;    // 2 ints per ICP work-group
    int* rgb_list = nullptr;    // eligible pixels of the three levels (level l at list_off[l])
    int* rgb_aux = nullptr;     // per level: the list length, then kIcpMaxBlocks work-group counts
    size_t list_off[3];
    int nblocks[3];  // work-groups per level of the photometric kernels (one pixel per lane)
    int ngroups[3];  // work-groups per level of the ICP-only reduction (two pixels per lane)
};

struct fu_system {
    fu_system_params p;
    fu_map* map = nullptr;
    fu_tracker* trk = nullptr;
    fu_tracker* reloc_trk = nullptr;  // ValidatePose's ICP, created on its first use
    std::vector<void*> allocs;
    fu_system_buffers b;
    int lc[3], lr[3];
    unsigned frame = 0;
    unsigned frame_at_load = 0;  // frame count at creation, reset or snapshot load
    bool mapping_off = false;    // Tracker::mappingDisabled, kept across fu_system_reset
    hipGraphExec_t graph = nullptr;
    bool prof = false;
    hipEvent_t ev[FU_NUM_STAGES + 1] = {};
    fu_stage_times times{};
    uint32_t graph_captures = 0;  // graphs captured (once per system unless profiling toggles)
    uint32_t graph_launches = 0;
    fu_batch* batch = nullptr;  // batch context the system is attached to
    int batch_slot = -1;
    fu::f4* fuse_px = nullptr;      // packed integration inputs (depth, normal) per pixel
    unsigned* fuse_rgb = nullptr;
};

namespace fu {
// error slot behind fu_last_error (defined in capi.hip)
fu_status api_error(fu_status code, const char* fmt, ...);
}  // namespace fu
