// Seeded synthetic RGB-D scenes: axis-aligned rooms with spheres and boxes, ray cast in
// closed form (include/fusion_scene.h).  Host code, OpenMP over image rows.
#include "../../include/fusion_scene.h"

#include <math.h>
#include <string.h>
#include <algorithm>
#include <array>
#include <mutex>
#include <vector>

namespace {

struct Box {
    double lo[3], hi[3];
    unsigned char alb[3];
};
struct Sphere {
    double c[3], r;
    unsigned char alb[3];
};

uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double u01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }
double urange(uint64_t& s, double a, double b) { return a + (b - a) * u01(s); }

constexpr double kPi = 3.14159265358979323846;
constexpr double kFps = 30.0;
constexpr double kSpeed = 0.3;                  // m/s
constexpr double kMaxRate = 20.0 * kPi / 180.0;  // rad/s

}  // namespace

struct fu_scene {
    double lo[3], hi[3];
    std::vector<Box> solids;  // interior walls and boxes
    std::vector<Sphere> spheres;
    unsigned char face_alb[6][3];  // room faces: -x +x -y +y floor ceiling
    double light[3];               // point light fixed in the world (view-independent shading)
    std::vector<std::array<double, 3>> path;  // dense closed path samples
    std::vector<double> cum;                   // cumulative arc length
    double total = 0;
    uint64_t seed = 0;
    mutable std::vector<double> yaw;  // per-frame heading, filled on demand
    mutable std::mutex yaw_mu;
};

namespace {

void add_wall_segments(fu_scene* s, int axis, double pos, double a0, double a1, const std::vector<double>& doors,
                       double height, uint64_t& rng) {
    // wall perpendicular to `axis` at `pos`, spanning [a0, a1] along the other horizontal axis,
    // with 1.0 m wide, 2.1 m tall doorways centred at `doors`
    const int other = 1 - axis;
    std::vector<double> cuts = doors;
    std::sort(cuts.begin(), cuts.end());
    double start = a0;
    unsigned char alb[3] = {(unsigned char)urange(rng, 150, 230), (unsigned char)urange(rng, 150, 230),
                            (unsigned char)urange(rng, 150, 230)};
    auto seg = [&](double b0, double b1, double z0, double z1) {
        if (b1 - b0 < 1e-3 || z1 - z0 < 1e-3) return;
        Box b;
        b.lo[axis] = pos - 0.05;
        b.hi[axis] = pos + 0.05;
        b.lo[other] = b0;
        b.hi[other] = b1;
        b.lo[2] = z0;
        b.hi[2] = z1;
        memcpy(b.alb, alb, 3);
        s->solids.push_back(b);
    };
    for (double d : cuts) {
        seg(start, d - 0.5, 0.0, height);
        seg(d - 0.5, d + 0.5, 2.1, height);  // lintel
        start = d + 0.5;
    }
    seg(start, a1, 0.0, height);
}

void build_path(fu_scene* s, const std::vector<std::array<double, 3>>& wp) {
    // closed Catmull-Rom spline through the waypoints, sampled densely
    const int K = (int)wp.size();
    const int per = 400;
    s->path.clear();
    for (int i = 0; i < K; ++i) {
        const auto& p0 = wp[(i + K - 1) % K];
        const auto& p1 = wp[i];
        const auto& p2 = wp[(i + 1) % K];
        const auto& p3 = wp[(i + 2) % K];
        for (int j = 0; j < per; ++j) {
            const double t = (double)j / per, t2 = t * t, t3 = t2 * t;
            std::array<double, 3> q;
            for (int c = 0; c < 3; ++c)
                q[c] = 0.5 * ((2 * p1[c]) + (-p0[c] + p2[c]) * t + (2 * p0[c] - 5 * p1[c] + 4 * p2[c] - p3[c]) * t2 +
                              (-p0[c] + 3 * p1[c] - 3 * p2[c] + p3[c]) * t3);
            s->path.push_back(q);
        }
    }
    s->path.push_back(s->path[0]);
    s->cum.assign(s->path.size(), 0.0);
    for (size_t i = 1; i < s->path.size(); ++i) {
        double d = 0;
        for (int c = 0; c < 3; ++c) d += (s->path[i][c] - s->path[i - 1][c]) * (s->path[i][c] - s->path[i - 1][c]);
        s->cum[i] = s->cum[i - 1] + sqrt(d);
    }
    s->total = s->cum.back();
}

// position and unit tangent at arc length `a`
void path_at(const fu_scene* s, double a, double p[3], double tg[3]) {
    a = fmod(a, s->total);
    if (a < 0) a += s->total;
    size_t i = std::upper_bound(s->cum.begin(), s->cum.end(), a) - s->cum.begin();
    if (i == 0) i = 1;
    if (i >= s->cum.size()) i = s->cum.size() - 1;
    const double seg = s->cum[i] - s->cum[i - 1];
    const double f = seg > 0 ? (a - s->cum[i - 1]) / seg : 0.0;
    double n = 0;
    for (int c = 0; c < 3; ++c) {
        p[c] = s->path[i - 1][c] + f * (s->path[i][c] - s->path[i - 1][c]);
        tg[c] = s->path[i][c] - s->path[i - 1][c];
        n += tg[c] * tg[c];
    }
    n = sqrt(n);
    for (int c = 0; c < 3; ++c) tg[c] = n > 0 ? tg[c] / n : (c == 0 ? 1.0 : 0.0);
}

bool clear_of_path(const fu_scene* s, double x, double y, double rad) {
    for (size_t i = 0; i < s->path.size(); i += 8) {
        const double dx = s->path[i][0] - x, dy = s->path[i][1] - y;
        if (dx * dx + dy * dy < (0.75 + rad) * (0.75 + rad)) return false;
    }
    return true;
}

void place_objects(fu_scene* s, const double lo[3], const double hi[3], uint64_t& rng) {
    const int ns = 4 + (int)(u01(rng) * 3), nb = 4 + (int)(u01(rng) * 3);
    for (int k = 0; k < ns; ++k)
        for (int tries = 0; tries < 60; ++tries) {
            Sphere sp;
            sp.r = urange(rng, 0.1, 0.35);
            sp.c[0] = urange(rng, lo[0] + 0.2 + sp.r, hi[0] - 0.2 - sp.r);
            sp.c[1] = urange(rng, lo[1] + 0.2 + sp.r, hi[1] - 0.2 - sp.r);
            sp.c[2] = u01(rng) < 0.5 ? sp.r : urange(rng, 0.4, std::min(2.0, hi[2] - 0.4));
            for (int c = 0; c < 3; ++c) sp.alb[c] = (unsigned char)urange(rng, 40, 250);
            if (!clear_of_path(s, sp.c[0], sp.c[1], sp.r)) continue;
            s->spheres.push_back(sp);
            break;
        }
    for (int k = 0; k < nb; ++k)
        for (int tries = 0; tries < 60; ++tries) {
            Box b;
            const double hx = urange(rng, 0.1, 0.4), hy = urange(rng, 0.1, 0.4), hz = urange(rng, 0.1, 0.5);
            const double cx = urange(rng, lo[0] + 0.2 + hx, hi[0] - 0.2 - hx);
            const double cy = urange(rng, lo[1] + 0.2 + hy, hi[1] - 0.2 - hy);
            const double cz = u01(rng) < 0.7 ? hz : urange(rng, 0.5, std::min(2.0, hi[2] - 0.5));
            for (int c = 0; c < 3; ++c) b.alb[c] = (unsigned char)urange(rng, 40, 250);
            b.lo[0] = cx - hx;
            b.hi[0] = cx + hx;
            b.lo[1] = cy - hy;
            b.hi[1] = cy + hy;
            b.lo[2] = cz - hz;
            b.hi[2] = cz + hz;
            if (!clear_of_path(s, cx, cy, std::max(hx, hy) * 1.42)) continue;
            s->solids.push_back(b);
            break;
        }
}

double heading_target(const fu_scene* s, int frame) {
    const double t = frame / kFps;
    double p[3], tg[3];
    path_at(s, kSpeed * t, p, tg);
    return atan2(tg[1], tg[0]) + (20.0 * kPi / 180.0) * sin(2 * kPi * t / 8.0);
}

// heading of every frame up to `frame`: follows the target with the rate bounded by 20 deg/s
double heading(const fu_scene* s, int frame) {
    std::lock_guard<std::mutex> g(s->yaw_mu);
    if (s->yaw.empty()) s->yaw.push_back(heading_target(s, 0));
    while ((int)s->yaw.size() <= frame) {
        const int f = (int)s->yaw.size();
        const double prev = s->yaw.back();
        double d = heading_target(s, f) - prev;
        d = remainder(d, 2 * kPi);
        const double lim = kMaxRate / kFps;
        d = std::max(-lim, std::min(lim, d));
        s->yaw.push_back(prev + d);
    }
    return s->yaw[frame];
}

// intersections; t along d (not normalised); returns +inf if none
inline double hit_box(const Box& b, const double o[3], const double d[3], int& axis, int& sign) {
    double tn = -1e300, tf = 1e300;
    int an = -1, sn = 0;
    for (int c = 0; c < 3; ++c) {
        if (d[c] == 0.0) {
            if (o[c] < b.lo[c] || o[c] > b.hi[c]) return INFINITY;
            continue;
        }
        double t1 = (b.lo[c] - o[c]) / d[c], t2 = (b.hi[c] - o[c]) / d[c];
        int s1 = -1;
        if (t1 > t2) {
            std::swap(t1, t2);
            s1 = 1;
        }
        if (t1 > tn) {
            tn = t1;
            an = c;
            sn = s1;
        }
        if (t2 < tf) tf = t2;
    }
    if (tn > tf || tn <= 1e-9) return INFINITY;
    axis = an;
    sign = sn;
    return tn;
}

inline double hit_sphere(const Sphere& s, const double o[3], const double d[3]) {
    double oc[3] = {o[0] - s.c[0], o[1] - s.c[1], o[2] - s.c[2]};
    const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
    const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - s.r * s.r;
    const double disc = b * b - a * c;
    if (disc < 0) return INFINITY;
    const double t = (-b - sqrt(disc)) / a;
    return t > 1e-9 ? t : INFINITY;
}

// nearest surface along the ray: t, world normal, albedo
double trace(const fu_scene* s, const double o[3], const double d[3], double n[3], const unsigned char** alb) {
    // room interior: exit through a face
    double best = INFINITY;
    for (int c = 0; c < 3; ++c) {
        if (d[c] == 0.0) continue;
        const double bound = d[c] > 0 ? s->hi[c] : s->lo[c];
        const double t = (bound - o[c]) / d[c];
        if (t > 1e-9 && t < best) {
            best = t;
            n[0] = n[1] = n[2] = 0.0;
            n[c] = d[c] > 0 ? -1.0 : 1.0;
            const int face = c < 2 ? 2 * c + (d[c] > 0 ? 1 : 0) : (d[c] > 0 ? 5 : 4);
            *alb = s->face_alb[face];
        }
    }
    for (const Box& b : s->solids) {
        int ax, sg;
        const double t = hit_box(b, o, d, ax, sg);
        if (t < best) {
            best = t;
            n[0] = n[1] = n[2] = 0.0;
            n[ax] = (double)sg;
            *alb = b.alb;
        }
    }
    for (const Sphere& sp : s->spheres) {
        const double t = hit_sphere(sp, o, d);
        if (t < best) {
            best = t;
            double q = 0;
            for (int c = 0; c < 3; ++c) {
                n[c] = o[c] + t * d[c] - sp.c[c];
                q += n[c] * n[c];
            }
            q = sqrt(q);
            for (int c = 0; c < 3; ++c) n[c] /= q;
            *alb = sp.alb;
        }
    }
    return best;
}

}  // namespace

extern "C" int fu_scene_create(int kind, float sx, float sy, float sz, uint64_t seed, fu_scene** out) {
    if (!out || !(sx > 1.5f) || !(sy > 1.5f) || !(sz > 1.8f)) return -1;
    fu_scene* s = new fu_scene();
    s->seed = seed;
    uint64_t rng = seed * 0x2545F4914F6CDD1Dull + 0x1234567ull;
    s->lo[0] = s->lo[1] = s->lo[2] = 0.0;
    s->hi[0] = sx;
    s->hi[1] = sy;
    s->hi[2] = sz;
    s->light[0] = 0.5 * sx;
    s->light[1] = 0.5 * sy;
    s->light[2] = sz - 0.3;
    for (int f = 0; f < 6; ++f)
        for (int c = 0; c < 3; ++c) s->face_alb[f][c] = (unsigned char)urange(rng, 120, 235);
    const double camz = std::min(1.5, (double)sz - 0.7);
    std::vector<std::array<double, 3>> wp;
    int nx = 1, ny = 1;
    if (kind == FU_SCENE_MULTIROOM) {
        nx = std::max(1, (int)lround(sx / 4.0));
        ny = std::max(1, (int)lround(sy / 4.0));
    }
    const double cw = sx / nx, ch = sy / ny;
    if (nx == 1 && ny == 1) {
        // loop around the room centre, keeping 0.8 m from the walls
        const double rx = std::max(0.2, 0.5 * sx - 1.0), ry = std::max(0.2, 0.5 * sy - 1.0);
        const int K = 6;
        const double ph = u01(rng) * 2 * kPi;
        for (int k = 0; k < K; ++k) {
            const double a = ph + 2 * kPi * k / K;
            const double j = urange(rng, 0.6, 1.0);
            wp.push_back({0.5 * sx + j * rx * cos(a), 0.5 * sy + j * ry * sin(a), camz + urange(rng, -0.15, 0.15)});
        }
    } else {
        // tour of the rooms on the border of the grid, through the doorways between them
        std::vector<std::pair<int, int>> tour;
        if (ny == 1) {
            for (int i = 0; i < nx; ++i) tour.push_back({i, 0});
            for (int i = nx - 2; i > 0; --i) tour.push_back({i, 0});
        } else if (nx == 1) {
            for (int j = 0; j < ny; ++j) tour.push_back({0, j});
            for (int j = ny - 2; j > 0; --j) tour.push_back({0, j});
        } else {
            for (int i = 0; i < nx; ++i) tour.push_back({i, 0});
            for (int j = 1; j < ny; ++j) tour.push_back({nx - 1, j});
            for (int i = nx - 2; i >= 0; --i) tour.push_back({i, ny - 1});
            for (int j = ny - 2; j > 0; --j) tour.push_back({0, j});
        }
        for (size_t k = 0; k < tour.size(); ++k) {
            const auto a = tour[k], b = tour[(k + 1) % tour.size()];
            const double ax = (a.first + 0.5) * cw, ay = (a.second + 0.5) * ch;
            const double side = (ny == 1 || nx == 1) && k >= tour.size() / 2 ? 0.5 : 0.0;
            wp.push_back({ax + urange(rng, -0.3, 0.3), ay + urange(rng, -0.3, 0.3) + side, camz});
            if (a != b) {
                const double bx = (b.first + 0.5) * cw, by = (b.second + 0.5) * ch;
                // doorway centre on the shared wall
                wp.push_back({0.5 * (ax + bx), 0.5 * (ay + by) + side, camz});
            }
        }
        // interior walls with a doorway in front of every room centre
        for (int i = 1; i < nx; ++i) {
            std::vector<double> doors;
            for (int j = 0; j < ny; ++j) doors.push_back((j + 0.5) * ch + (nx == 1 || ny == 1 ? 0.25 : 0.0));
            add_wall_segments(s, 0, i * cw, 0.0, sy, doors, sz, rng);
        }
        for (int j = 1; j < ny; ++j) {
            std::vector<double> doors;
            for (int i = 0; i < nx; ++i) doors.push_back((i + 0.5) * cw);
            add_wall_segments(s, 1, j * ch, 0.0, sx, doors, sz, rng);
        }
    }
    build_path(s, wp);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const double lo[3] = {i * cw + 0.05, j * ch + 0.05, 0.0};
            const double hi[3] = {(i + 1) * cw - 0.05, (j + 1) * ch - 0.05, (double)sz};
            place_objects(s, lo, hi, rng);
        }
    *out = s;
    return 0;
}

extern "C" void fu_scene_destroy(fu_scene* s) { delete s; }

extern "C" int fu_scene_num_primitives(const fu_scene* s) { return s ? 6 + (int)(s->solids.size() + s->spheres.size()) : 0; }

extern "C" int fu_scene_pose(const fu_scene* s, int frame, double pose[16]) {
    if (!s || !pose || frame < 0) return -1;
    const double t = frame / kFps;
    double p[3], tg[3];
    path_at(s, kSpeed * t, p, tg);
    const double yaw = heading(s, frame);
    const double pitch = (-12.0 + 6.0 * sin(2 * kPi * t / 11.0)) * kPi / 180.0;
    const double f[3] = {cos(pitch) * cos(yaw), cos(pitch) * sin(yaw), sin(pitch)};
    const double r[3] = {sin(yaw), -cos(yaw), 0.0};
    // down = forward x right
    const double dn[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2], f[0] * r[1] - f[1] * r[0]};
    for (int i = 0; i < 3; ++i) {
        pose[i * 4 + 0] = r[i];
        pose[i * 4 + 1] = dn[i];
        pose[i * 4 + 2] = f[i];
        pose[i * 4 + 3] = p[i];
    }
    pose[12] = pose[13] = pose[14] = 0.0;
    pose[15] = 1.0;
    return 0;
}

extern "C" int fu_scene_render_pose(const fu_scene* s, const double pose[16], int cols, int rows, float fx, float fy,
                                    float cx, float cy, float depth_scale, int noise, uint64_t noise_seed,
                                    uint16_t* depth, uint8_t* rgb) {
    if (!s || !pose || cols <= 0 || rows <= 0) return -1;
    const double o[3] = {pose[3], pose[7], pose[11]};
#pragma omp parallel for schedule(static)
    for (int v = 0; v < rows; ++v) {
        for (int u = 0; u < cols; ++u) {
            const double dc[3] = {((double)u - cx) / fx, ((double)v - cy) / fy, 1.0};
            double d[3];
            for (int i = 0; i < 3; ++i) d[i] = pose[i * 4 + 0] * dc[0] + pose[i * 4 + 1] * dc[1] + pose[i * 4 + 2] * dc[2];
            double n[3] = {0, 0, 1};
            const unsigned char* alb = s->face_alb[4];
            const double t = trace(s, o, d, n, &alb);
            const size_t pix = (size_t)v * cols + u;
            // depth along the optical axis equals t because the camera-frame ray has z = 1
            double z = t;
            uint16_t q = 0;
            if (isfinite(z)) {
                if (noise) {
                    uint64_t st = s->seed * 0x9E3779B97F4A7C15ull ^ (noise_seed + 1) * 0xD1B54A32D192ED03ull ^
                                  (uint64_t)pix * 0x94D049BB133111EBull;
                    const double u1 = u01(st), u2 = u01(st), u3 = u01(st);
                    const double g = sqrt(-2.0 * log(u1 > 1e-300 ? u1 : 1e-300)) * cos(2 * kPi * u2);
                    z += 0.0015 * z * z * g;
                    if (u3 < 0.01) z = 0.0;
                }
                double qd = floor(z * depth_scale + 0.5);
                if (qd < 0) qd = 0;
                if (qd > 65535) qd = 65535;
                q = (uint16_t)qd;
            }
            if (depth) depth[pix] = q;
            if (rgb) {
                // Lambertian under a point light fixed in the world: a surface point has the
                // same colour from every viewpoint (brightness constancy), with smooth shading
                // gradients across flat faces
                double l[3];
                for (int i = 0; i < 3; ++i) l[i] = s->light[i] - (o[i] + t * d[i]);
                const double ll = sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
                double c = (n[0] * l[0] + n[1] * l[1] + n[2] * l[2]) / ll;
                if (!isfinite(t) || !(ll > 0)) c = 0;
                const double shade = 0.25 + 0.75 * (c > 0 ? c : 0.0);
                for (int k = 0; k < 3; ++k) rgb[pix * 3 + k] = (uint8_t)floor(alb[k] * shade + 0.5);
            }
        }
    }
    return 0;
}

extern "C" int fu_scene_render(const fu_scene* s, int frame, int cols, int rows, float fx, float fy, float cx, float cy,
                               float depth_scale, int noise, uint16_t* depth, uint8_t* rgb) {
    double pose[16];
    if (fu_scene_pose(s, frame, pose)) return -1;
    return fu_scene_render_pose(s, pose, cols, rows, fx, fy, cx, cy, depth_scale, noise, (uint64_t)frame, depth, rgb);
}

extern "C" float fu_scene_sdf(const fu_scene* s, float x, float y, float z) {
    if (!s) return NAN;
    const double p[3] = {x, y, z};
    double d = INFINITY;
    for (int c = 0; c < 3; ++c) d = std::min(d, std::min(p[c] - s->lo[c], s->hi[c] - p[c]));
    for (const Box& b : s->solids) {
        double q[3], out = 0, in = -INFINITY;
        for (int c = 0; c < 3; ++c) {
            q[c] = std::max(b.lo[c] - p[c], p[c] - b.hi[c]);
            if (q[c] > 0) out += q[c] * q[c];
            in = std::max(in, q[c]);
        }
        d = std::min(d, out > 0 ? sqrt(out) : in);
    }
    for (const Sphere& sp : s->spheres) {
        const double r = sqrt((p[0] - sp.c[0]) * (p[0] - sp.c[0]) + (p[1] - sp.c[1]) * (p[1] - sp.c[1]) +
                              (p[2] - sp.c[2]) * (p[2] - sp.c[2]));
        d = std::min(d, r - sp.r);
    }
    return (float)d;
}
