// Serving many reconstruction sessions with continuous batching.
//
// Each session is a system (map, pose, pyramid, last raycast, all resident in HBM).  Sessions
// are spread over a few lanes; a lane owns a HIP stream and a batch context.  A lane forms a
// step from the oldest pending frame of each of its sessions whenever fewer than
// `max_inflight` of its steps are on the device, so frames that arrive while steps run join
// the next step.  Steps of different lanes run concurrently on the device.  A step ends with
// a gather of every frame's result (pose, status, counters) into one device array and one
// copy to pinned host memory; an event per step tells the host when it is done.
//
// The per-session order of frames is the stream order of its lane.  The reference has no
// counterpart: it drives one System from its main loop (Core/System.cc:92-138).
#include <hip/hip_runtime.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <deque>
#include <vector>

#include "../../include/fusion_amd.h"
#include "fu_batch.h"

using namespace fu;

namespace {

struct Req {
    uint64_t tag;
    int session;
    const uint16_t* depth;
    const uint8_t* rgb;
    double t_submit;
    int kind;  // FU_REQ_FRAME or FU_REQ_MESH
    int host;  // frame in host memory: uploaded into the session's input slot inside the step
    float* mv;
    float* mn;
    uint8_t* mc;
    uint32_t max_tri;
};

struct Step {
    int n = 0;
    int slot = 0;
    int kind = FU_REQ_FRAME;  // a batched frame step, or one session's mesh extraction
    Req req[kMaxBatch];
};

struct Lane {
    hipStream_t stream = nullptr;
    // uploads of host frames run on the server's copy stream, so a step's frames are copied
    // while the lane's previous step computes: each upload waits until the lane's previous
    // step has read its frames (read), and the step waits for its uploads (up[slot])
    hipEvent_t read = nullptr;
    bool read_recorded = false;
    std::vector<hipEvent_t> up;
    fu_batch* batch = nullptr;
    std::vector<int> sessions;
    std::deque<Step> inflight;
    std::vector<hipEvent_t> ev;  // one per ring slot
    FrameBrief* d_res = nullptr;  // kMaxBatch results of the step being computed
    FrameBrief* h_res = nullptr;  // pinned, max_inflight x kMaxBatch: one slot per queued step
    uint64_t launched = 0;
    double pixels = 0;  // load measure for placing sessions
};

struct Session {
    fu_system* sys = nullptr;
    uint16_t* depth_in = nullptr;  // the system's input slots (host frames are uploaded there)
    uint8_t* rgb_in = nullptr;
    size_t npix = 0;
    int lane = -1;
    size_t bytes = 0;
    double pixels = 0;
    std::deque<Req> pending;  // submitted, not yet in a step
    int in_steps = 0;         // frames in launched steps
};

}  // namespace

struct fu_server {
    fu_server_params p;
    std::vector<Lane> lanes;
    hipStream_t copy = nullptr;  // uploads of host frames, for every lane (one DMA queue)
    std::vector<Session> sess;
    uint64_t used = 0;
    std::deque<fu_completion> done;
    int64_t outstanding = 0;
    std::chrono::steady_clock::time_point t0;
};

#define SV_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return api_error(FU_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

extern "C" void fu_server_params_default(fu_server_params* p) {
    p->max_batch = 32;
    p->max_inflight = 2;
    p->lanes = 2;
    p->block_budget = 0;
}

extern "C" double fu_server_now_ms(fu_server* sv) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - sv->t0).count();
}

extern "C" fu_status fu_server_destroy(fu_server* sv) {
    if (!sv) return FU_OK;
    for (Lane& L : sv->lanes)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (sv->copy) (void)hipStreamSynchronize(sv->copy);
    for (Session& s : sv->sess)
        if (s.sys) fu_system_destroy(s.sys);
    for (Lane& L : sv->lanes) {
        fu_batch_destroy(L.batch);
        for (hipEvent_t e : L.ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : L.up) (void)hipEventDestroy(e);
        if (L.read) (void)hipEventDestroy(L.read);
        (void)hipFree(L.d_res);
        (void)hipHostFree(L.h_res);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    if (sv->copy) (void)hipStreamDestroy(sv->copy);
    delete sv;
    return FU_OK;
}

extern "C" fu_status fu_server_create(const fu_server_params* p, fu_server** out) {
    if (!p || !out) return api_error(FU_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->max_batch < 1 || p->max_batch > kMaxBatch) return api_error(FU_ERR_ARG, "max_batch %d not in 1..%d", p->max_batch, kMaxBatch);
    if (p->max_inflight < 1 || p->max_inflight > 16) return api_error(FU_ERR_ARG, "max_inflight %d not in 1..16", p->max_inflight);
    if (p->lanes < 1 || p->lanes > 8) return api_error(FU_ERR_ARG, "lanes %d not in 1..8", p->lanes);
    fu_server* sv = new fu_server();
    sv->p = *p;
    sv->t0 = std::chrono::steady_clock::now();
    sv->lanes.resize(p->lanes);
    for (Lane& L : sv->lanes) {
        fu_status st = fu_batch_create(4096, &L.batch);
        if (st) {
            fu_server_destroy(sv);
            return st;
        }
        const size_t nb = sizeof(FrameBrief) * kMaxBatch * p->max_inflight;
        hipError_t e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
        if (e == hipSuccess && !sv->copy) e = hipStreamCreateWithFlags(&sv->copy, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&L.read, hipEventDisableTiming);
        if (e == hipSuccess) e = hipMalloc(&L.d_res, sizeof(FrameBrief) * kMaxBatch);
        if (e == hipSuccess) e = hipHostMalloc(&L.h_res, nb, hipHostMallocDefault);
        for (int k = 0; k < p->max_inflight && e == hipSuccess; ++k) {
            hipEvent_t ev, up;
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
            if (e == hipSuccess) L.ev.push_back(ev);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&up, hipEventDisableTiming);
            if (e == hipSuccess) L.up.push_back(up);
        }
        if (e != hipSuccess) {
            fu_server_destroy(sv);
            return api_error(FU_ERR_HIP, "server lane set-up: %s", hipGetErrorString(e));
        }
    }
    *out = sv;
    return FU_OK;
}

static Session* get_session(fu_server* sv, int32_t id) {
    if (!sv || id < 0 || id >= (int32_t)sv->sess.size() || !sv->sess[id].sys) return nullptr;
    return &sv->sess[id];
}

extern "C" fu_status fu_server_open(fu_server* sv, const fu_system_params* p, int32_t* session) {
    if (!sv || !p || !session) return api_error(FU_ERR_ARG, "null argument");
    const size_t bytes = fu_map_bytes(&p->map);
    if (bytes == 0) return api_error(FU_ERR_ARG, "bad map parameters");
    if (sv->p.block_budget && sv->used + bytes > sv->p.block_budget)
        return api_error(FU_ERR_BUDGET, "session needs %zu bytes, %llu of %llu in use", bytes,
                         (unsigned long long)sv->used, (unsigned long long)sv->p.block_budget);
    Session s;
    fu_status st = fu_system_create(p, &s.sys);
    if (st) return st;
    fu_system_buffers b;
    fu_system_get_buffers(s.sys, &b);
    s.depth_in = b.depth_in;
    s.rgb_in = b.rgb_in;
    s.npix = (size_t)p->cols * p->rows;
    // the lane with the least pixel load per step among those whose batch context takes the
    // session (a context holds one voxel format and one tracking mode)
    std::vector<int> order(sv->lanes.size());
    for (int i = 0; i < (int)order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return sv->lanes[a].pixels < sv->lanes[b].pixels; });
    int best = -1;
    for (int i : order)
        if ((st = fu_batch_attach(sv->lanes[i].batch, s.sys)) == FU_OK) {
            best = i;
            break;
        }
    if (best < 0) {
        fu_system_destroy(s.sys);
        return st;
    }
    s.lane = best;
    s.bytes = bytes;
    s.pixels = (double)p->cols * p->rows;
    sv->lanes[best].pixels += s.pixels;
    const int id = (int)sv->sess.size();
    sv->lanes[best].sessions.push_back(id);
    sv->sess.push_back(std::move(s));
    sv->used += bytes;
    *session = id;
    return FU_OK;
}

extern "C" fu_status fu_server_close(fu_server* sv, int32_t id) {
    Session* s = get_session(sv, id);
    if (!s) return api_error(FU_ERR_ARG, "no session %d", id);
    if (!s->pending.empty() || s->in_steps) return api_error(FU_ERR_STATE, "session %d has frames in flight", id);
    Lane& L = sv->lanes[s->lane];
    L.sessions.erase(std::find(L.sessions.begin(), L.sessions.end(), id));
    L.pixels -= s->pixels;
    fu_system_destroy(s->sys);  // detaches it from the lane's batch context
    s->sys = nullptr;
    sv->used -= s->bytes;
    return FU_OK;
}

extern "C" fu_system* fu_server_system(fu_server* sv, int32_t id) {
    Session* s = get_session(sv, id);
    return s ? s->sys : nullptr;
}

static fu_status submit_frame(fu_server* sv, int32_t id, const uint16_t* depth, const uint8_t* rgb, uint64_t tag,
                              double t_submit_ms, int host) {
    Session* s = get_session(sv, id);
    if (!s) return api_error(FU_ERR_ARG, "no session %d", id);
    if (!depth || !rgb) return api_error(FU_ERR_ARG, "null frame");
    Req r{};
    r.tag = tag;
    r.session = id;
    r.depth = depth;
    r.rgb = rgb;
    r.t_submit = t_submit_ms < 0 ? fu_server_now_ms(sv) : t_submit_ms;
    r.kind = FU_REQ_FRAME;
    r.host = host;
    s->pending.push_back(r);
    sv->outstanding++;
    return FU_OK;
}

extern "C" fu_status fu_server_submit(fu_server* sv, int32_t id, const uint16_t* depth, const uint8_t* rgb, uint64_t tag,
                                      double t_submit_ms) {
    return submit_frame(sv, id, depth, rgb, tag, t_submit_ms, 0);
}

extern "C" fu_status fu_server_submit_host(fu_server* sv, int32_t id, const uint16_t* depth, const uint8_t* rgb,
                                           uint64_t tag, double t_submit_ms) {
    return submit_frame(sv, id, depth, rgb, tag, t_submit_ms, 1);
}

extern "C" fu_status fu_server_submit_mesh(fu_server* sv, int32_t id, float* vertex, float* normal, uint8_t* color,
                                           uint32_t max_triangles, uint64_t tag, double t_submit_ms) {
    Session* s = get_session(sv, id);
    if (!s) return api_error(FU_ERR_ARG, "no session %d", id);
    if (!vertex || !normal || !color) return api_error(FU_ERR_ARG, "null mesh buffer");
    Req r{};
    r.tag = tag;
    r.session = id;
    r.t_submit = t_submit_ms < 0 ? fu_server_now_ms(sv) : t_submit_ms;
    r.kind = FU_REQ_MESH;
    r.mv = vertex;
    r.mn = normal;
    r.mc = color;
    r.max_tri = max_triangles;
    s->pending.push_back(r);
    sv->outstanding++;
    return FU_OK;
}

extern "C" fu_status fu_server_pending(fu_server* sv, int32_t* frames) {
    if (!sv || !frames) return api_error(FU_ERR_ARG, "null argument");
    *frames = (int32_t)sv->outstanding;
    return FU_OK;
}

// finished steps of a lane, in launch order
static fu_status retire(fu_server* sv, Lane& L) {
    while (!L.inflight.empty()) {
        Step& st = L.inflight.front();
        const hipError_t e = hipEventQuery(L.ev[st.slot]);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) return api_error(FU_ERR_HIP, "step completion: %s", hipGetErrorString(e));
        const double now = fu_server_now_ms(sv);
        const FrameBrief* r = L.h_res + (size_t)st.slot * kMaxBatch;
        if (st.kind == FU_REQ_MESH) {
            // the slot holds the map's counters after the extraction
            const unsigned* ctr = reinterpret_cast<const unsigned*>(r);
            fu_completion c;
            memset(&c, 0, sizeof(c));
            c.tag = st.req[0].tag;
            c.session = st.req[0].session;
            c.kind = FU_REQ_MESH;
            c.track_status = -1;
            c.n_triangles_total = ctr[C_TRI];
            c.n_triangles = ctr[C_TRI] < st.req[0].max_tri ? ctr[C_TRI] : st.req[0].max_tri;
            c.n_occupied = ctr[C_OCC];
            c.flags = ctr[C_FLAGS] | (ctr[C_TRI] > st.req[0].max_tri ? (unsigned)FU_FLAG_TRI_FULL : 0u);
            c.t_submit_ms = st.req[0].t_submit;
            c.t_done_ms = now;
            sv->done.push_back(c);
            sv->sess[c.session].in_steps--;
            L.inflight.pop_front();
            continue;
        }
        for (int i = 0; i < st.n; ++i) {
            fu_completion c;
            memset(&c, 0, sizeof(c));
            c.tag = st.req[i].tag;
            c.session = st.req[i].session;
            c.track_status = r[i].status;
            memcpy(c.pose, r[i].pose, sizeof(c.pose));
            c.icp_error = r[i].icp_error;
            c.n_visible = r[i].n_visible;
            c.n_allocated = r[i].n_alloc;
            c.flags = r[i].flags;
            c.rgb_count = r[i].rgb_count;
            c.t_submit_ms = st.req[i].t_submit;
            c.t_done_ms = now;
            sv->done.push_back(c);
            sv->sess[c.session].in_steps--;
        }
        L.inflight.pop_front();
    }
    return FU_OK;
}

// a step that could not be launched: its requests go back to the front of their sessions'
// queues, so the server stays consistent (the frames can be retried or the sessions closed)
static void unlaunch(fu_server* sv, const Step& st) {
    for (int i = st.n - 1; i >= 0; --i) {
        Session& s = sv->sess[st.req[i].session];
        s.pending.push_front(st.req[i]);
        s.in_steps--;
    }
}

// next steps of a lane while it has room: the oldest pending frame of each session, oldest
// sessions first
static fu_status launch(fu_server* sv, Lane& L) {
    std::vector<int> cand;
    while ((int)L.inflight.size() < sv->p.max_inflight) {
        // a mesh request at the head of a session's queue runs as a step of its own
        int mesh = -1;
        for (int id : L.sessions) {
            const Session& s = sv->sess[id];
            if (!s.pending.empty() && s.pending.front().kind == FU_REQ_MESH &&
                (mesh < 0 || s.pending.front().t_submit < sv->sess[mesh].pending.front().t_submit))
                mesh = id;
        }
        if (mesh >= 0) {
            Session& s = sv->sess[mesh];
            Step st;
            st.kind = FU_REQ_MESH;
            st.n = 1;
            st.slot = (int)(L.launched % (uint64_t)sv->p.max_inflight);
            st.req[0] = s.pending.front();
            s.pending.pop_front();
            s.in_steps++;
            fu_status e = mesh_async(s.sys, st.req[0].mv, st.req[0].mn, st.req[0].mc, st.req[0].max_tri, L.stream);
            hipError_t he = hipSuccess;
            if (!e) he = hipMemcpyAsync(L.h_res + (size_t)st.slot * kMaxBatch, system_counters(s.sys),
                                        sizeof(unsigned) * 16, hipMemcpyDeviceToHost, L.stream);
            if (!e && he == hipSuccess) he = hipEventRecord(L.ev[st.slot], L.stream);
            if (e || he != hipSuccess) {
                unlaunch(sv, st);
                return e ? e : api_error(FU_ERR_HIP, "mesh step: %s", hipGetErrorString(he));
            }
            L.launched++;
            L.inflight.push_back(st);
            continue;
        }
        cand.clear();
        for (int id : L.sessions)
            if (!sv->sess[id].pending.empty() && sv->sess[id].pending.front().kind == FU_REQ_FRAME) cand.push_back(id);
        if (cand.empty()) break;
        if ((int)cand.size() > sv->p.max_batch) {
            std::partial_sort(cand.begin(), cand.begin() + sv->p.max_batch, cand.end(), [&](int a, int b) {
                return sv->sess[a].pending.front().t_submit < sv->sess[b].pending.front().t_submit;
            });
            cand.resize(sv->p.max_batch);
        }
        // slots in session order: a lane whose pending sessions repeat launches the same step
        // (and replays its captured graph, fu_server_set_graph)
        std::sort(cand.begin(), cand.end());
        Step st;
        st.n = (int)cand.size();
        st.slot = (int)(L.launched % (uint64_t)sv->p.max_inflight);
        fu_system* sys[kMaxBatch];
        const uint16_t* dp[kMaxBatch];
        const uint8_t* cp[kMaxBatch];
        hipError_t he = hipSuccess;
        bool uploads = false;
        for (int i = 0; i < st.n; ++i) {
            Session& s = sv->sess[cand[i]];
            st.req[i] = s.pending.front();
            s.pending.pop_front();
            s.in_steps++;
            sys[i] = s.sys;
            dp[i] = st.req[i].depth;
            cp[i] = st.req[i].rgb;
            if (st.req[i].host) {  // the upload of Frame::Frame (Core/Frame.cc:61-62), for this step
                if (!uploads && L.read_recorded && he == hipSuccess) he = hipStreamWaitEvent(sv->copy, L.read, 0);
                uploads = true;
                if (he == hipSuccess) he = hipMemcpyAsync(s.depth_in, dp[i], s.npix * 2, hipMemcpyHostToDevice, sv->copy);
                if (he == hipSuccess) he = hipMemcpyAsync(s.rgb_in, cp[i], s.npix * 3, hipMemcpyHostToDevice, sv->copy);
                dp[i] = s.depth_in;
                cp[i] = s.rgb_in;
            }
        }
        if (uploads && he == hipSuccess) he = hipEventRecord(L.up[st.slot], sv->copy);
        if (uploads && he == hipSuccess) he = hipStreamWaitEvent(L.stream, L.up[st.slot], 0);
        // one device result buffer per lane: the lane's steps run in stream order, and each
        // step's results are copied to its host slot before the next step writes them
        FrameBrief* dres = L.d_res;
        FrameBrief* hres = L.h_res + (size_t)st.slot * kMaxBatch;
        fu_status e = FU_OK;
        if (he != hipSuccess) e = api_error(FU_ERR_HIP, "frame upload: %s", hipGetErrorString(he));
        if (!e) e = batch_grab(L.batch, st.n, sys, dp, cp, dres, L.stream, L.read);
        if (!e) L.read_recorded = true;
        if (!e) {
            he = hipMemcpyAsync(hres, dres, sizeof(FrameBrief) * st.n, hipMemcpyDeviceToHost, L.stream);
            if (he == hipSuccess) he = hipEventRecord(L.ev[st.slot], L.stream);
            if (he != hipSuccess) e = api_error(FU_ERR_HIP, "frame step: %s", hipGetErrorString(he));
        }
        if (e) {
            // batch_grab checks its arguments before any launch; after a HIP error the frames'
            // device work may be partly queued, and the requests are returned all the same
            unlaunch(sv, st);
            return e;
        }
        L.launched++;
        L.inflight.push_back(st);
    }
    return FU_OK;
}

static fu_status progress(fu_server* sv) {
    for (Lane& L : sv->lanes) {
        fu_status st = retire(sv, L);
        if (st) return st;
        if ((st = launch(sv, L))) return st;
    }
    return FU_OK;
}

static int take_done(fu_server* sv, fu_completion* out, int cap) {
    int k = 0;
    while (k < cap && !sv->done.empty()) {
        out[k++] = sv->done.front();
        sv->done.pop_front();
    }
    sv->outstanding -= k;
    return k;
}

extern "C" fu_status fu_server_poll(fu_server* sv, fu_completion* out, int32_t cap, int32_t* n_out) {
    if (!sv || !n_out || (cap > 0 && !out)) return api_error(FU_ERR_ARG, "null argument");
    fu_status st = progress(sv);
    *n_out = take_done(sv, out, cap);
    return st;
}

extern "C" fu_status fu_server_run_trace(fu_server* sv, const fu_trace_frame* frames, int32_t n, int32_t closed_loop,
                                         fu_completion* out) {
    if (!sv || (n > 0 && (!frames || !out))) return api_error(FU_ERR_ARG, "null argument");
    for (int32_t i = 0; i < n; ++i) {
        const fu_trace_frame& f = frames[i];
        const bool ok = f.kind == FU_REQ_MESH ? (f.vertex && f.normal && f.color) : (f.kind == FU_REQ_FRAME && f.depth && f.rgb);
        if (!get_session(sv, f.session) || !ok) return api_error(FU_ERR_ARG, "trace request %d: bad session or buffers", i);
    }
    if (sv->outstanding) return api_error(FU_ERR_STATE, "run_trace needs an idle server");
    std::vector<int32_t> order(n);
    for (int32_t i = 0; i < n; ++i) order[i] = i;
    // per-session queues keep the given order; open loop releases frames by arrival time
    std::vector<std::deque<int32_t>> per(sv->sess.size());
    for (int32_t i : order) per[frames[i].session].push_back(i);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return frames[a].arrival_ms < frames[b].arrival_ms; });
    const double t0 = fu_server_now_ms(sv);
    auto submit = [&](int32_t i, double t) {
        const fu_trace_frame& f = frames[i];
        if (f.kind == FU_REQ_MESH)
            return fu_server_submit_mesh(sv, f.session, f.vertex, f.normal, f.color, f.max_triangles, f.tag, t);
        return submit_frame(sv, f.session, f.depth, f.rgb, f.tag, t, f.host);
    };
    fu_status st = FU_OK;
    int32_t next = 0, got = 0;
    if (closed_loop)
        for (auto& q : per)
            if (!q.empty()) {
                if ((st = submit(q.front(), -1.0))) return st;
                q.pop_front();
            }
    std::vector<fu_completion> buf(256);
    while (got < n) {
        if (!closed_loop) {
            const double now = fu_server_now_ms(sv) - t0;
            while (next < n && frames[order[next]].arrival_ms <= now) {
                if ((st = submit(order[next], t0 + frames[order[next]].arrival_ms))) return st;
                ++next;
            }
        }
        int32_t k = 0;
        if ((st = fu_server_poll(sv, buf.data(), (int32_t)buf.size(), &k))) return st;
        for (int32_t j = 0; j < k; ++j) {
            out[got++] = buf[j];
            if (closed_loop) {
                auto& q = per[buf[j].session];
                if (!q.empty()) {
                    if ((st = submit(q.front(), -1.0))) return st;
                    q.pop_front();
                }
            }
        }
    }
    return FU_OK;
}

extern "C" fu_status fu_server_set_profiling(fu_server* sv, int enable) {
    if (!sv) return api_error(FU_ERR_ARG, "null server");
    for (Lane& L : sv->lanes) {
        fu_status st = fu_batch_set_profiling(L.batch, enable);
        if (st) return st;
    }
    return FU_OK;
}

extern "C" fu_status fu_server_set_graph(fu_server* sv, int enable) {
    if (!sv) return api_error(FU_ERR_ARG, "null server");
    for (Lane& L : sv->lanes) {
        fu_status st = fu_batch_set_graph(L.batch, enable);
        if (st) return st;
    }
    return FU_OK;
}

extern "C" fu_status fu_server_graph_counts(fu_server* sv, uint32_t* captures, uint32_t* replays) {
    if (!sv || !captures || !replays) return api_error(FU_ERR_ARG, "null argument");
    *captures = *replays = 0;
    for (Lane& L : sv->lanes) {
        uint32_t c = 0, r = 0;
        fu_status st = fu_batch_graph_counts(L.batch, &c, &r);
        if (st) return st;
        *captures += c;
        *replays += r;
    }
    return FU_OK;
}

extern "C" fu_status fu_server_get_stage_times(fu_server* sv, fu_stage_times* out) {
    if (!sv || !out) return api_error(FU_ERR_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    for (Lane& L : sv->lanes) {
        fu_stage_times t;
        fu_status st = fu_batch_get_stage_times(L.batch, &t);
        if (st) return st;
        for (int k = 0; k < FU_NUM_STAGES; ++k) out->ms[k] += t.ms[k];
        out->frames += t.frames;
        out->n_visible += t.n_visible;
        out->n_new += t.n_new;
        out->icp_iterations += t.icp_iterations;
        out->icp_pixels += t.icp_pixels;
        out->steps += t.steps;
    }
    return FU_OK;
}
