// Marching-cubes tables of the oracle, expanded into the 16-wide form of
// Mapping/Constant.h:39-296 from the library's table data header (the public-domain
// Bourke table; tests pin all three tables against the reference's fingerprints).
// TEST INFRASTRUCTURE ONLY.
#include "../fusion_amd/csrc/mc_tables.h"

int g_or_edge[256];
int g_or_tri[256][16];
int g_or_vert[256];

namespace {
struct Init {
    Init() {
        for (int c = 0; c < 256; ++c) {
            g_or_edge[c] = fu::mc_edge_mask(c);
            g_or_vert[c] = fu::mc_vertex_count(c);
            int i = 0;
            for (const char* p = fu::kMcTriHex[c]; *p; ++p, ++i) g_or_tri[c][i] = fu::mc_hex_digit(*p);
            for (; i < 16; ++i) g_or_tri[c][i] = -1;
        }
    }
} init_tables;
}  // namespace
