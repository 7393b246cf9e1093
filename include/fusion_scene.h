/*
 * fusion_scene.h - seeded synthetic RGB-D scenes (host library libfu_scene.so).
 *
 * Test and benchmark input for the dense pipeline; it replaces the reference's dataset
 * loaders (MainTum.cc:78-99, Core/Camera.cc), which read recorded sequences.  A scene is a
 * set of axis-aligned rooms (floor, walls, ceiling, interior walls with doorways) holding
 * seeded spheres and boxes.  Frames are exact ray casts of that geometry: depth along the
 * optical axis as u16 at `depth_scale` units per metre, with Gaussian noise of standard
 * deviation 0.0015 z^2 m and 1% of pixels dropped to 0, and Lambertian colour from per-
 * primitive albedo lit from the camera.  The camera follows a seeded closed Catmull-Rom
 * path at 0.3 m/s with a heading rate bounded by 20 deg/s (30 frames per second).
 * World frame: z up; camera frame: x right, y down, z forward.  Poses are camera-to-world,
 * 4x4 row-major doubles.  Output is bit-identical for the same arguments on any x86-64 host
 * with the same C library.
 */
#ifndef FUSION_SCENE_H
#define FUSION_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fu_scene fu_scene;

#define FU_SCENE_ROOM 0      /* one room of size sx x sy x sz */
#define FU_SCENE_MULTIROOM 1 /* sx x sy x sz split into rooms of about 4 m with doorways */

int fu_scene_create(int kind, float sx, float sy, float sz, uint64_t seed, fu_scene** out);
void fu_scene_destroy(fu_scene* s);
/* number of primitives (walls, spheres, boxes) */
int fu_scene_num_primitives(const fu_scene* s);
/* camera-to-world pose of frame `frame` of the scene's trajectory */
int fu_scene_pose(const fu_scene* s, int frame, double pose[16]);
/* ray cast at an arbitrary pose; noise != 0 adds depth noise and dropout seeded by
 * (scene seed, noise_seed).  depth: cols*rows u16, rgb: cols*rows*3 u8 (either may be NULL) */
int fu_scene_render_pose(const fu_scene* s, const double pose[16], int cols, int rows, float fx, float fy, float cx,
                         float cy, float depth_scale, int noise, uint64_t noise_seed, uint16_t* depth, uint8_t* rgb);
/* frame `frame` of the trajectory (noise seeded by the frame index) */
int fu_scene_render(const fu_scene* s, int frame, int cols, int rows, float fx, float fy, float cx, float cy,
                    float depth_scale, int noise, uint16_t* depth, uint8_t* rgb);
/* signed distance of a world point to the scene surface (negative inside solids) */
float fu_scene_sdf(const fu_scene* s, float x, float y, float z);

#ifdef __cplusplus
}
#endif
#endif
