// capi.cpp -- host runtime behind gmapf_c.h: the gmapf::GMAPF object of the
// reference (include/gmapf/GMAPF.h:30-98, src/GMAPF.cu) re-designed around
// device-resident SoA state, device-side counts (no per-frame host readback
// of request or moving-set sizes) and, for the batched form, a captured
// hipGraph per frame.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <climits>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gmapf_c.h"
#include "device.h"
#include "kernels.h"
#include "navfile.h"
#include "navmesh.h"

using namespace gmapf_amd;

namespace {

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                         \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess)                                                                \
            throw ApiError(GMAPF_E_DEVICE, std::string("hip error: ") + hipGetErrorString(_e) + \
                                               " at " #expr);                                \
    } while (0)

using Clock = std::chrono::steady_clock;

// The caller's current HIP device for the duration of one entry point.  The
// reference pushes its CUDA context before device work and pops it after, so
// an embedding host keeps its own (src/GMAPF.cu:82,127,266,308); here the
// first device use of an entry point saves the caller's device, makes the
// handle's current, and the scope restores the caller's on every return.
struct DeviceScope {
    DeviceScope** owner;
    int prev = -1;  // the caller's device once saved
    explicit DeviceScope(DeviceScope** o) : owner(o) { *owner = this; }
    ~DeviceScope() {
        *owner = nullptr;
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) hipSetDevice(prev);
    }
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;
};

}  // namespace

struct gmapf_s {
    std::string err;
    bool inited = false, started = false, step_going = false;
    gmapf_config cfg{};

    // map build (GMAPF.cu:133-196)
    std::vector<float> cdt_v;
    std::vector<int> cdt_e;
    bool map_filled = false;
    Navmesh nav;

    // host agent tables; pos is the stable host mirror behind GetAgentPos
    int n_agents = 0;
    std::vector<float> pos;
    std::vector<int> state, curr, targ, gen;
    std::vector<float> targ_pos;
    // Target changes made after start, applied at the next AsyncStep, at
    // most one per agent (a later SetAgentTargPos overwrites the entry in
    // place).  Between steps they are written straight into the next frame's
    // pinned input (h_upd, h_upd_xy): entry upd_pos[id] holds agent id's
    // change when upd_epoch_of[id] == upd_epoch, in the bucket of agent id's
    // kUpdBucket agents; upd_n entries are in use, upd_buckets lists the
    // buckets with a count to clear; a finished step starts a new epoch.  Changes
    // made while a step runs (its input is in use) queue in upd_ids / upd_targ
    // / upd_xy (upd_qpos[id]: index, or -1) and move to the pinned input when
    // the step ends.
    std::vector<int> upd_ids, upd_targ, upd_qpos, upd_pos;
    std::vector<unsigned> upd_epoch_of;
    unsigned upd_epoch = 1;
    int upd_n = 0;
    std::vector<int> upd_buckets;
    std::vector<float> upd_xy;

    // device
    int dev = 0;
    hipStream_t s1 = nullptr, s2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // around the search launch of a timed frame (every search_timing-th
    // frame; recorded into the captured frame graph too).  Event nodes cost
    // the graph launch about 10 us each, so other frames carry none
    hipEvent_t ev_s[2] = {nullptr, nullptr};
    bool step_timed = false;  // the frame in flight records ev_s
    bool capturing = false;  // enqueue_frame runs under stream capture
    DevMap dmap{};
    DevAgents dag{};
    DevSlots dsl{};
    DevFrame dfr{};
    SearchParams sp{};
    CrowdParams cp{};
    int search_grid = 0;
    bool ident = false;  // slot i = agent i (P >= N)
    // this device's staging for gmapf_locate_device_noalloc
    float2* loc_pts = nullptr;
    int* loc_out = nullptr;
    int persistent = 0;
    std::vector<void*> dallocs;
    // pinned staging
    float* h_step = nullptr;
    float* h_pos_stage = nullptr;
    int* h_state_stage = nullptr;
    unsigned long long* h_stats = nullptr;
    unsigned long long* h_slab = nullptr;  // [0] slab ints in use, [1] deferred corridors
    int* h_group = nullptr;  // agents in Search and in Move after the last frame, per 64 agents
    int searching_after = -1;  // their sums (-1: no frame ran yet)
    int moving_after = -1;
    // 64-agent groups whose mirrored states the host changed (a target set):
    // the next readback copies them whatever the device exported
    std::vector<int> host_dirty;
    std::vector<unsigned char> host_dirty_flag;
    long long crowd_frames = 0;  // frames that ran without admission and search
    int* h_upd = nullptr;       // bucket counts [ceil(N / kUpdBucket)], then ids[N], targ[N]
    float2* h_upd_xy = nullptr;
    int upd_cap = 0;
    int upd_nb = 0;  // buckets of the update input
    // batched frame graphs, [compaction before the search][search timed]
    // [compaction before the search][search timed][crowd phase beside the search]
    hipGraph_t graph[2][2][2] = {};
    hipGraphExec_t gexec[2][2][2] = {};
    // agents in Move at the start of the next frame: the last frame's count
    // less those whose target was set since (their mirror state said Move)
    long long movers = 0;
    float step_dev = NAN;  // the step the device holds (NaN: none yet)
    bool compact_next = false;   // the slab was half full after the last frame
    long long compactions = 0, slab_in_use = 0, slab_deferred = 0;

    DeviceScope* scope = nullptr;  // set while an entry point runs

    // profiling
    Clock::time_point t_step;
    // [0] wall [1] admit [2] search [3] unused [4] crowd (device, profiler)
    // [5] host enqueue [6] host wait for the device [7] host readback
    double last_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double sum_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0}, max_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    Clock::time_point t_enq;
    hipEvent_t ev_done = nullptr;
    long long frames = 0;
    long long stats_host[8] = {0};

    ~gmapf_s() { release(); }

    template <class T>
    T* dalloc(size_t n, const char* what) {
        void* p = nullptr;
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        if (e != hipSuccess)
            throw ApiError(GMAPF_E_NOMEM, std::string("device allocation failed for ") + what + " (" +
                                              std::to_string(n * sizeof(T)) + " bytes)");
        dallocs.push_back(p);
        return static_cast<T*>(p);
    }

    // Saves the caller's device (once per entry point) and makes the handle's
    // device current.
    hipError_t enter_device() {
        if (scope && scope->prev < 0) {
            const hipError_t e = hipGetDevice(&scope->prev);
            if (e != hipSuccess) return e;
        }
        return hipSetDevice(dev);
    }
    void use_device() { HIPCHK(enter_device()); }

    void release() {
        if (!started) return;
        enter_device();
        if (s1) hipStreamSynchronize(s1);
        if (s2) hipStreamSynchronize(s2);
        for (int c = 0; c < 2; c++)
            for (int t = 0; t < 2; t++)
                for (int m = 0; m < 2; m++) {
                    if (gexec[c][t][m]) hipGraphExecDestroy(gexec[c][t][m]);
                    if (graph[c][t][m]) hipGraphDestroy(graph[c][t][m]);
                    gexec[c][t][m] = nullptr;
                    graph[c][t][m] = nullptr;
                }
        step_dev = NAN;
        for (void* p : dallocs) hipFree(p);
        dallocs.clear();
        if (h_step) hipHostFree(h_step);
        if (h_pos_stage) hipHostFree(h_pos_stage);
        if (h_state_stage) hipHostFree(h_state_stage);
        if (h_stats) hipHostFree(h_stats);
        if (h_slab) hipHostFree(h_slab);
        if (h_group) hipHostFree(h_group);
        h_group = nullptr;
        searching_after = -1;
        moving_after = -1;
        host_dirty.clear();
        host_dirty_flag.clear();
        if (h_upd) hipHostFree(h_upd);
        if (h_upd_xy) hipHostFree(h_upd_xy);
        h_step = h_pos_stage = nullptr;
        h_state_stage = nullptr;
        h_stats = nullptr;
        h_slab = nullptr;
        h_upd = nullptr;
        h_upd_xy = nullptr;
        for (auto& e : ev_p)
            if (e) hipEventDestroy(e);
        for (auto& e : ev_s)
            if (e) hipEventDestroy(e);
        if (ev_fork) hipEventDestroy(ev_fork);
        if (ev_join) hipEventDestroy(ev_join);
        if (ev_done) hipEventDestroy(ev_done);
        ev_done = nullptr;
        if (s1) hipStreamDestroy(s1);
        if (s2) hipStreamDestroy(s2);
        s1 = s2 = nullptr;
        started = false;
    }

    void require_init() const {
        if (!inited) throw ApiError(GMAPF_E_STATE, "not initialized");
    }

    // ------------------------------------------------------------ lifecycle
    void init(const gmapf_config& c) {
        if (inited) throw ApiError(GMAPF_E_STATE, "already initialized");
        if (c.agents_max_count < 0) throw ApiError(GMAPF_E_ARG, "AgentsMaxCount must be >= 0");
        cfg = c;
        pos.clear();
        pos.reserve(2 * (size_t)std::max(1, c.agents_max_count));
        inited = true;
    }

    void check_config() {
        int T = cfg.path_finder_threads_per_agents;
        if (T != 2 && T != 4 && T != 8 && T != 16 && T != 32 && T != 64)
            throw ApiError(GMAPF_E_ARG, "PathFinderThreadsPerAgents must be one of 2,4,8,16,32,64");
        int Q = cfg.path_finder_queue_capacity;
        if (Q < 1 || Q > 256) throw ApiError(GMAPF_E_ARG, "PathFinderQueueCapacity must be in [1, 256]");
        if (cfg.path_finder_parallel_agents < 1)
            throw ApiError(GMAPF_E_ARG, "PathFinderParallelAgents must be >= 1");
        // a negative weight would rank nodes farther from the other end first;
        // with K >= 0 every F is at least 1, which the search's keys rely on
        if (!std::isfinite(cfg.path_finder_heuristic_k) || cfg.path_finder_heuristic_k < 0.0f)
            throw ApiError(GMAPF_E_ARG, "PathFinderHeuristicK must be finite and >= 0");
        if (cfg.step_mode != GMAPF_MODE_PLAIN && cfg.step_mode != GMAPF_MODE_BATCHED)
            throw ApiError(GMAPF_E_ARG, "unknown step_mode");
        if (cfg.step_mode == GMAPF_MODE_BATCHED && cfg.round_budget < 1)
            throw ApiError(GMAPF_E_ARG, "batched mode needs round_budget >= 1");
        if (cfg.searches_per_wave > 64 / T)
            throw ApiError(GMAPF_E_ARG, "searches_per_wave must be at most 64 / PathFinderThreadsPerAgents");
        if (cfg.search_timing < 0) throw ApiError(GMAPF_E_ARG, "search_timing must be >= 0");
        if (cfg.search_workgroups < 0) throw ApiError(GMAPF_E_ARG, "search_workgroups must be >= 0");
    }

    void manual_start() {
        require_init();
        if (started) throw ApiError(GMAPF_E_STATE, "already started");
        if (!map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        check_config();
        // node ids, corridor offsets and slab indices are 32-bit
        if (nav.V >= (1 << 30)) throw ApiError(GMAPF_E_ARG, "navmesh has more than 2^30 - 1 nodes");
        dev = cfg.device;
        use_device();
        started = true;  // from here release() frees what was allocated
        try {
            start_device();
        } catch (...) {
            release();
            throw;
        }
    }

    void start_device() {
        const int V = nav.V;
        const int N = n_agents;
        const int P = cfg.path_finder_parallel_agents;
        const int T = cfg.path_finder_threads_per_agents;
        const int Q = cfg.path_finder_queue_capacity;
        HIPCHK(locate_scratch(dev, &loc_pts, &loc_out));
        HIPCHK(hipStreamCreateWithFlags(&s1, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming));
        for (auto& e : ev_p) HIPCHK(hipEventCreate(&e));
        for (auto& e : ev_s) HIPCHK(hipEventCreate(&e));

        // ---- navmesh
        std::vector<int4> nbr4(V);
        std::vector<float2> cen2(V), tri2(3 * (size_t)V);
        for (int i = 0; i < V; i++) {
            nbr4[i] = make_int4(nav.nbr[3 * (size_t)i], nav.nbr[3 * (size_t)i + 1], nav.nbr[3 * (size_t)i + 2], -1);
            cen2[i] = make_float2(nav.cen[2 * (size_t)i], nav.cen[2 * (size_t)i + 1]);
            for (int k = 0; k < 3; k++)
                tri2[3 * (size_t)i + k] = make_float2(nav.tri[6 * (size_t)i + 2 * k], nav.tri[6 * (size_t)i + 2 * k + 1]);
        }
        // search records: neighbours + their centroids (bit patterns) per node
        std::vector<int4> rec(3 * (size_t)V);
        auto fbits = [](float f) {
            int b;
            std::memcpy(&b, &f, sizeof(b));
            return b;
        };
        for (int i = 0; i < V; i++) {
            const int4 nb = nbr4[i];
            const int n[3] = {nb.x, nb.y, nb.z};
            float c[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            for (int k = 0; k < 3; k++)
                if (n[k] >= 0) {
                    c[2 * k] = cen2[n[k]].x;
                    c[2 * k + 1] = cen2[n[k]].y;
                }
            rec[3 * (size_t)i] = make_int4(n[0], n[1], n[2], -1);
            rec[3 * (size_t)i + 1] = make_int4(fbits(c[0]), fbits(c[1]), fbits(c[2]), fbits(c[3]));
            rec[3 * (size_t)i + 2] = make_int4(fbits(c[4]), fbits(c[5]), 0, 0);
        }
        int4* d_rec = dalloc<int4>(3 * (size_t)V, "navmesh search records");
        HIPCHK(hipMemcpy(d_rec, rec.data(), 3 * (size_t)V * sizeof(int4), hipMemcpyHostToDevice));
        int4* d_nbr = dalloc<int4>(V, "navmesh neighbours");
        float2* d_cen = dalloc<float2>(V, "navmesh centroids");
        float2* d_tri = dalloc<float2>(3 * (size_t)V, "navmesh triangles");
        int* d_cs = dalloc<int>(nav.cell_start.size(), "grid cells");
        int* d_ct = dalloc<int>(std::max<size_t>(1, nav.cell_tris.size()), "grid lists");
        int* d_gl = dalloc<int>(std::max<size_t>(1, nav.global_tris.size()), "grid global list");
        HIPCHK(hipMemcpy(d_nbr, nbr4.data(), V * sizeof(int4), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_cen, cen2.data(), V * sizeof(float2), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_tri, tri2.data(), 3 * (size_t)V * sizeof(float2), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_cs, nav.cell_start.data(), nav.cell_start.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!nav.cell_tris.empty())
            HIPCHK(hipMemcpy(d_ct, nav.cell_tris.data(), nav.cell_tris.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!nav.global_tris.empty())
            HIPCHK(hipMemcpy(d_gl, nav.global_tris.data(), nav.global_tris.size() * sizeof(int), hipMemcpyHostToDevice));
        dmap = DevMap{};
        dmap.V = V;
        dmap.nbr = d_nbr;
        dmap.cen = d_cen;
        dmap.tri = d_tri;
        dmap.rec = d_rec;
        dmap.gx0 = nav.gx0;
        dmap.gy0 = nav.gy0;
        dmap.inv_cs = nav.inv_cs;
        dmap.rx0 = nav.rx0;
        dmap.ry0 = nav.ry0;
        dmap.rx1 = nav.rx1;
        dmap.ry1 = nav.ry1;
        dmap.nx = nav.nx;
        dmap.ny = nav.ny;
        dmap.cell_start = d_cs;
        dmap.cell_tris = d_ct;
        dmap.global_tris = d_gl;
        dmap.n_global = (int)nav.global_tris.size();

        // ---- crowd constants (AgentsMover.h:221-224)
        cp = CrowdParams{};
        cp.cd = cfg.agent_radius * 2.5f;
        cp.cd2 = cp.cd * cp.cd;
        cp.slack = cfg.agent_radius;
        if (cfg.rot_cs == 0.0f && cfg.rot_sn == 0.0f) {
            // V2Float.h:39-47 with angle 10: 10 * pi_f / 180.0f in FP32, then
            // cos/sin evaluated in double and rounded once to FP32.
            const float pi_f = 3.14159265358979f;
            float ang = 10.0f * pi_f / 180.0f;
            cp.rot_cs = (float)std::cos((double)ang);
            cp.rot_sn = (float)std::sin((double)ang);
        } else {
            cp.rot_cs = cfg.rot_cs;
            cp.rot_sn = cfg.rot_sn;
        }
        // default cell: just above the avoidance distance cd, so every
        // neighbour within cd lies in the 3 x 3 cells around an agent (the
        // margin of 1/64 is far above the rounding of the cell coordinates)
        float cell = cfg.hash_cell > 0 ? cfg.hash_cell : cp.cd * (1.0f + 1.0f / 64.0f);
        if (!(cell > 0)) cell = 1.0f;
        dmap.hx0 = nav.bx0;
        dmap.hy0 = nav.by0;
        dmap.hinv = 1.0f / cell;
        dmap.hnx = (int)((nav.bx1 - nav.bx0) * dmap.hinv) + 1;
        dmap.hny = (int)((nav.by1 - nav.by0) * dmap.hinv) + 1;
        if ((long long)dmap.hnx * dmap.hny > (1LL << 26))
            throw ApiError(GMAPF_E_ARG, "hash_cell too small for the map size");
        dmap.hk = (int)std::floor(cp.cd / cell) + 1;

        // ---- agents
        const int Nn = std::max(N, 1);
        dag = DevAgents{};
        dag.n = N;
        dag.pos = dalloc<float2>(Nn, "agent positions");
        dag.curr = dalloc<int>(Nn, "agent nodes");
        dag.targ_pos = dalloc<float2>(Nn, "agent targets");
        dag.targ = dalloc<int>(Nn, "agent target nodes");
        dag.state = dalloc<int>(Nn, "agent states");
        dag.step_idx = dalloc<int>(Nn, "agent step idx");
        dag.step_node = dalloc<int>(Nn, "agent step node");
        dag.step_pos = dalloc<float2>(Nn, "agent step pos");
        dag.path_len = dalloc<int>(Nn, "agent path len");
        dag.path_off = dalloc<long long>(Nn, "agent corridor offsets");
        {
            // a corridor visits each node at most once, so V ints always
            // hold one; N * V hold every agent's
            const long long dense = (long long)Nn * V;
            long long cap = cfg.corridor_slab > 0 ? cfg.corridor_slab
                                                   : std::min(dense, std::max((long long)Nn * 1024, 16LL * V));
            cap = std::min(std::max(cap, (long long)V), (long long)INT_MAX - 1);  // offsets are 32-bit
            dag.slab_cap = cap;
            dag.slab[0] = dalloc<int>((size_t)cap, "corridor slab");
            dag.slab[1] = dalloc<int>((size_t)cap, "corridor slab (compaction copy)");
            dag.slab_now = dalloc<int*>(1, "corridor slab in use");
            dag.slab_top = dalloc<unsigned long long>(2, "corridor slab top");
            dag.slab_deferred = dag.slab_top + 1;
            HIPCHK(hipMemcpy(dag.slab_now, &dag.slab[0], sizeof(int*), hipMemcpyHostToDevice));
            HIPCHK(launch_fill_u64(dag.slab_top, 0, 2, s1));
        }
        dag.gen = dalloc<int>(Nn, "agent gen");
        dag.inflight = dalloc<int>(Nn, "agent inflight");
        dag.in_m = dalloc<int>(Nn, "agent moving flag");
        if (N > 0) {
            std::vector<float2> tp(N);
            for (int i = 0; i < N; i++) tp[i] = make_float2(targ_pos[2 * i], targ_pos[2 * i + 1]);
            HIPCHK(hipMemcpy(dag.pos, pos.data(), N * sizeof(float2), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dag.targ_pos, tp.data(), N * sizeof(float2), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dag.curr, curr.data(), N * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dag.targ, targ.data(), N * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dag.state, state.data(), N * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(dag.gen, gen.data(), N * sizeof(int), hipMemcpyHostToDevice));
        }
        HIPCHK(launch_fill(dag.step_idx, -1, Nn, s1));
        HIPCHK(launch_fill(dag.step_node, -1, Nn, s1));
        HIPCHK(launch_fill(dag.path_len, 0, Nn, s1));
        HIPCHK(hipMemsetAsync(dag.path_off, 0, Nn * sizeof(long long), s1));
        HIPCHK(launch_fill(dag.inflight, -1, Nn, s1));
        HIPCHK(launch_fill(dag.in_m, 0, Nn, s1));
        HIPCHK(hipMemsetAsync(dag.step_pos, 0, Nn * sizeof(float2), s1));

        // ---- search slots
        dsl = DevSlots{};
        dsl.P = P;
        dsl.agent = dalloc<int>(P, "slot agent");
        dsl.gen = dalloc<int>(P, "slot gen");
        dsl.S = dalloc<int>(P, "slot start");
        dsl.Tg = dalloc<int>(P, "slot target");
        dsl.stamp = dalloc<unsigned>(P, "slot stamp");
        dsl.clr = dalloc<unsigned>(P, "slot mark clear epoch");
        dsl.status = dalloc<int>(P, "slot status");
        dsl.mS = dalloc<int>(P, "slot meet start side");
        dsl.mT = dalloc<int>(P, "slot meet target side");
        dsl.corr = dalloc<int2>(P, "slot corridor in walk");
        dsl.walk = dalloc<int4>(P, "slot parked walk");
        dsl.VP = ((long long)V + 63) / 64 * 64;
        dsl.mark = dalloc<unsigned char>((size_t)P * dsl.VP, "visited marks");
        dsl.vis = dalloc<uint2>((size_t)P * V, "visited entries");
        dsl.qsave = dalloc<unsigned long long>((size_t)P * T * Q, "parked queues");
        dsl.qcnt = dalloc<int>((size_t)P * T, "parked queue counts");
        HIPCHK(launch_fill(dsl.agent, -1, P, s1));
        HIPCHK(launch_fill(dsl.status, 0, P, s1));
        HIPCHK(hipMemsetAsync(dsl.stamp, 0, P * sizeof(unsigned), s1));
        // the marks start zeroed, which is a clear of epoch 0 (stamps 0 .. 126)
        HIPCHK(hipMemsetAsync(dsl.clr, 0, P * sizeof(unsigned), s1));
        HIPCHK(hipMemsetAsync(dsl.mark, 0, (size_t)P * dsl.VP, s1));

        // ---- frame scratch
        dfr = DevFrame{};
        const int scan_block = 1024;
        dfr.blk = dalloc<int>(Nn / scan_block + 2, "scan partials");
        dfr.mblk = dalloc<int>(Nn / scan_block + 2, "moving partials");
        dfr.sblk = dalloc<int>(P / scan_block + 2, "slot scan partials");
        dfr.free_list = dalloc<int>(P, "free slots");
        dfr.adm_agent = dalloc<int>(P, "admitted agents");
        dfr.active = dalloc<int>(P, "active slots");
        dfr.counters = dalloc<int>(16, "counters");
        dfr.stats = dalloc<unsigned long long>(8, "stats");
        dfr.stats_part = dalloc<unsigned long long>(8 * kStatStripes, "striped frame stats");
        dfr.ncell = dmap.hnx * dmap.hny;
        dfr.hkey = dalloc<int>(Nn, "hash keys");
        dfr.hcount = dalloc<int>((size_t)dfr.ncell + 1, "hash counts");
        HIPCHK(launch_fill(dfr.hcount, 0, dfr.ncell + 1, s1));
        dfr.hstart = dalloc<int>((size_t)dfr.ncell + 2, "hash starts");
        HIPCHK(launch_fill(dfr.hstart, 0, dfr.ncell + 2, s1));  // no agent hashed before the first mover
        dfr.hslot = dalloc<int>(Nn, "hash arrival slots");
        dfr.horder = dalloc<int>(Nn, "hash order");
        dfr.hent = dalloc<float4>(Nn, "hash positions");
        dfr.hblk = dalloc<int>(dfr.ncell / scan_block + 2, "hash scan partials");
        // corridors of agents, then of slots whose corridor walk is parked
        dfr.slab_len = dalloc<int>((size_t)Nn + P + 1, "corridor lengths");
        dfr.slab_blk = dalloc<int>((Nn + P) / scan_block + 2, "corridor scan partials");
        dfr.stepv = dalloc<float>(1, "step");
        dfr.state_prev = dalloc<int>(Nn, "exported states");
        if (N > 0) HIPCHK(hipMemcpy(dfr.state_prev, state.data(), N * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(launch_fill(dfr.counters, 0, 16, s1));
        HIPCHK(launch_fill_u64(dfr.stats, 0, 8, s1));
        HIPCHK(launch_fill_u64(dfr.stats_part, 0, 8 * kStatStripes, s1));
        HIPCHK(launch_fill(dfr.hkey, -1, Nn, s1));

        // ---- pinned staging
        HIPCHK(hipHostMalloc((void**)&h_step, sizeof(float)));
        HIPCHK(hipHostMalloc((void**)&h_pos_stage, Nn * sizeof(float2)));
        HIPCHK(hipHostMalloc((void**)&h_state_stage, Nn * sizeof(int)));
        HIPCHK(hipHostMalloc((void**)&h_stats, 8 * sizeof(unsigned long long)));
        HIPCHK(hipHostMalloc((void**)&h_slab, 2 * sizeof(unsigned long long)));
        h_slab[0] = h_slab[1] = 0;
        HIPCHK(hipHostMalloc((void**)&h_group, (size_t)std::max(1, (N + 63) / 64) * sizeof(int)));
        std::memset(h_group, 0, (size_t)std::max(1, (N + 63) / 64) * sizeof(int));
        searching_after = -1;
        moving_after = -1;
        upd_cap = std::max(N, 1);
        upd_nb = (upd_cap + kUpdBucket - 1) / kUpdBucket;
        HIPCHK(hipHostMalloc((void**)&h_upd, ((size_t)upd_nb + 2 * (size_t)upd_cap) * sizeof(int)));
        HIPCHK(hipHostMalloc((void**)&h_upd_xy, (size_t)upd_cap * sizeof(float2)));
        std::memset(h_upd, 0, (size_t)upd_nb * sizeof(int));
        upd_buckets.clear();
        *h_step = 0.f;
        // the crowd's last kernel writes only the positions it sets, the
        // export pass only the states that changed
        if (N > 0) std::memcpy(h_pos_stage, pos.data(), (size_t)N * sizeof(float2));
        if (N > 0) std::memcpy(h_state_stage, state.data(), (size_t)N * sizeof(int));
        movers = std::count(state.begin(), state.begin() + N, (int)GMAPF_STATE_MOVE);
        dfr.h_step = h_step;
        dfr.h_upd_bcnt = h_upd;
        dfr.h_upd_ids = h_upd + upd_nb;
        dfr.h_upd_targ = h_upd + upd_nb + upd_cap;
        dfr.h_upd_xy = h_upd_xy;
        dfr.h_pos_out = reinterpret_cast<float2*>(h_pos_stage);
        dfr.h_state_out = h_state_stage;
        dfr.h_stats_out = h_stats;
        dfr.h_slab_out = h_slab;
        dfr.h_group_out = h_group;

        // ---- search launch shape
        const int rpw = cfg.searches_per_wave > 0 ? cfg.searches_per_wave : 64 / T;
        // P >= N: every request of a frame is admitted, so agent i can keep
        // slot i and admission needs no ranks (the slots are interchangeable:
        // a search's result does not depend on its slot)
        ident = P >= N && N > 0;
        sp = SearchParams{T, Q, cfg.path_finder_heuristic_k, 0, rpw, ident ? N : 0, cfg.search_wide_keys ? 1 : 0};
        const int Pe = ident ? N : P;  // slots that can be in flight
        HIPCHK(search_prepare(T, Q));
        if (cfg.step_mode == GMAPF_MODE_BATCHED) {
            int per_cu = 0, cus = 0;
            HIPCHK(search_occupancy(T, Q, &per_cu));
            HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            if (per_cu < 1) throw ApiError(GMAPF_E_ARG, "queue capacity does not fit in LDS");
            int waves_needed = (Pe + rpw - 1) / rpw;
            search_grid = std::max(1, std::min(per_cu * cus, waves_needed));
            if (cfg.search_workgroups > 0) search_grid = std::min(search_grid, cfg.search_workgroups);
            persistent = 1;
            sp.budget = cfg.round_budget;
        } else {
            search_grid = (Pe + rpw - 1) / rpw;
            persistent = 0;
            sp.budget = 0;
        }
        HIPCHK(hipStreamSynchronize(s1));
    }

    // ------------------------------------------------------------ frame
    // crowd: some agent moves, so the crowd phase runs on the second stream
    // beside the search.  A frame in which nobody moves leaves it out (every
    // crowd pass would find nothing to do): one stream, no fork and join, and
    // the search does not share the GPU with the crowd's launches (maze
    // workload: 14.95 against 14.60 M solves/s).
    void enqueue_frame(bool compact, bool timed, bool crowd) {
        HIPCHK(launch_admit(dmap, dag, dsl, dfr, ident, s1));
        if (compact) HIPCHK(launch_compact(dag, dsl, dfr, s1));
        if (crowd) {
            HIPCHK(hipEventRecord(ev_fork, s1));
            HIPCHK(hipStreamWaitEvent(s2, ev_fork, 0));
        }
        // inside a stream capture the records become external event nodes
        // of the graph (a plain record there only orders the capture and
        // leaves the events without timestamps)
        const unsigned evf = capturing ? hipEventRecordExternal : hipEventRecordDefault;
        if (timed) HIPCHK(hipEventRecordWithFlags(ev_s[0], s1, evf));
        HIPCHK(launch_search(dmap, dag, dsl, dfr, sp, search_grid, persistent, s1));
        if (timed) HIPCHK(hipEventRecordWithFlags(ev_s[1], s1, evf));
        if (crowd) {
            HIPCHK(launch_crowd(dmap, dag, dfr, cp, s2));
            HIPCHK(hipEventRecord(ev_join, s2));
            HIPCHK(hipStreamWaitEvent(s1, ev_join, 0));
        }
        HIPCHK(launch_export(dag, dfr, s1));
    }

    void debug_sync(const char* phase) {
        hipError_t e = hipStreamSynchronize(s1);
        if (e == hipSuccess) e = hipStreamSynchronize(s2);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess)
            throw ApiError(GMAPF_E_DEVICE, std::string("hip error after ") + phase + ": " + hipGetErrorString(e));
    }

    void enqueue_frame_profiled(bool compact) {
        HIPCHK(hipEventRecord(ev_p[0], s1));
        HIPCHK(launch_admit(dmap, dag, dsl, dfr, ident, s1));
        if (compact) HIPCHK(launch_compact(dag, dsl, dfr, s1));
        if (cfg.is_debug) debug_sync("admit");
        HIPCHK(hipEventRecord(ev_p[1], s1));
        HIPCHK(launch_search(dmap, dag, dsl, dfr, sp, search_grid, persistent, s1));
        if (cfg.is_debug) debug_sync("search");
        HIPCHK(hipEventRecord(ev_p[2], s1));
        HIPCHK(hipEventRecord(ev_p[3], s1));
        HIPCHK(launch_crowd(dmap, dag, dfr, cp, s1));
        if (cfg.is_debug) debug_sync("crowd");
        HIPCHK(hipEventRecord(ev_p[4], s1));
        HIPCHK(launch_export(dag, dfr, s1));
    }

    // A frame after which no agent searched and some moved, with no target
    // change since: admission and the search have nothing to do, so the frame
    // is the crowd phase alone, on one stream (its move pass takes the moving
    // set, its last pass exports).  new_step: the step differs from the one
    // the device holds.  Launched directly in both step forms: for these six
    // kernels on one stream a graph launch took about 4 us longer per frame
    // (0.0856 against 0.0817 ms on the 65,536-agent crowd).
    void enqueue_crowd_frame(bool new_step) { HIPCHK(launch_crowd(dmap, dag, dfr, cp, s1, true, new_step)); }

    // Captures one frame (with or without the slab compaction in front, with
    // or without events around the search, with or without the crowd phase)
    // into a graph and instantiates it.
    void build_graph(bool compact, bool timed, bool crowd) {
        hipGraph_t& gr = graph[compact][timed][crowd];
        hipGraphExec_t& ge = gexec[compact][timed][crowd];
        HIPCHK(hipStreamBeginCapture(s1, hipStreamCaptureModeThreadLocal));
        try {
            capturing = true;
            enqueue_frame(compact, timed, crowd);
            capturing = false;
        } catch (...) {
            capturing = false;
            hipGraph_t g = nullptr;
            hipStreamEndCapture(s1, &g);
            if (g) hipGraphDestroy(g);
            throw;
        }
        HIPCHK(hipStreamEndCapture(s1, &gr));
        HIPCHK(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
    }

    void async_step(float dt) {
        require_init();
        if (step_going) return;  // GMAPF.cu:260-261
        if (!started) manual_start();
        use_device();
        t_step = Clock::now();
        // frame inputs go to pinned host memory that the frame's first
        // kernels read directly (no copies; the captured graph stays fixed)
        *h_step = cfg.agent_speed * dt;  // AgentsMover.h:299
        // the changes are already in place (stage_update)
        // compact the corridor slab before this frame's search when the
        // last frame left it half full (the host reads the slab top that the
        // last frame exported; no device round trip)
        const bool compact = compact_next;
        // nothing to admit or search: no agent searched after the last frame
        // and no target changed since (a compaction waits for a full frame);
        // a frame in which nobody moves either runs in full, where every
        // pass finds nothing to do
        const bool crowd_only = searching_after == 0 && moving_after > 0 && upd_n == 0 && !compact &&
                                !cfg.is_profiler && !cfg.is_debug;
        if (compact) compactions++;
        const bool timed = !crowd_only && cfg.search_timing > 0 && frames % cfg.search_timing == 0;
        step_timed = timed;
        // every other frame copies the step to the device (k_apply_updates)
        const bool new_step = !(*h_step == step_dev);
        step_dev = *h_step;
        if (crowd_only) {
            crowd_frames++;
            enqueue_crowd_frame(new_step);
        } else if (cfg.step_mode == GMAPF_MODE_BATCHED && !cfg.is_profiler && !cfg.is_debug) {
            // every frame variant in use is captured on the first step, so a
            // later frame never pays for an instantiation
            for (int c = 0; c < 2; c++)
                for (int t = 0; t < (cfg.search_timing > 0 ? 2 : 1); t++)
                    for (int m = 0; m < 2; m++)
                        if (!gexec[c][t][m]) build_graph(c != 0, t != 0, m != 0);
            HIPCHK(hipGraphLaunch(gexec[compact][timed][movers != 0], s1));
        } else if (cfg.is_profiler || cfg.is_debug) {
            enqueue_frame_profiled(compact);
        } else {
            enqueue_frame(compact, timed, movers != 0);
        }
        HIPCHK(hipEventRecord(ev_done, s1));
        t_enq = Clock::now();
        step_going = true;
    }

    void wait_step_end() {
        require_init();
        if (!step_going) return;
        use_device();
        step_going = false;
        // poll the frame's completion event: lower wake-up latency than a
        // blocking stream synchronisation, like CUDA's default spin schedule
        hipError_t q;
        while ((q = hipEventQuery(ev_done)) == hipErrorNotReady) {
        }
        HIPCHK(q);
        HIPCHK(hipGetLastError());
        const Clock::time_point t_sync = Clock::now();
        if (n_agents > 0) std::memcpy(pos.data(), h_pos_stage, (size_t)n_agents * sizeof(float2));
        // states: the groups the export rewrote (it writes a group whole when
        // one of its states changed) and those whose mirror the host changed
        searching_after = 0;
        moving_after = 0;
        for (int w = 0; w < (n_agents + 63) / 64; w++) {
            const int word = h_group[w];
            searching_after += word & 0xff;
            moving_after += (word >> 8) & 0xff;
            if ((word >> 16) & 1 || (w < (int)host_dirty_flag.size() && host_dirty_flag[w])) {
                const int b = 64 * w, e = std::min(n_agents, b + 64);
                std::memcpy(state.data() + b, h_state_stage + b, (size_t)(e - b) * sizeof(int));
            }
        }
        for (const int w : host_dirty) host_dirty_flag[w] = 0;
        host_dirty.clear();
        for (int k = 0; k < 8; k++) stats_host[k] = (long long)h_stats[k];
        slab_in_use = (long long)h_slab[0];
        slab_deferred = (long long)h_slab[1];
        compact_next = 2 * slab_in_use > dag.slab_cap;
        frames++;
        stats_host[7] = frames;
        // the step has read its changes: a new epoch empties the pinned input;
        // changes made while the step ran stay pending for the next step
        upd_n = 0;
        upd_epoch++;
        for (const int b : upd_buckets) h_upd[b] = 0;
        upd_buckets.clear();
        movers = moving_after;
        for (size_t k = 0; k < upd_ids.size(); k++) {
            const int id = upd_ids[k];
            if (state[id] == GMAPF_STATE_MOVE) movers--;
            state[id] = GMAPF_STATE_SEARCH;
            mark_host_dirty(id);
            stage_update(id, upd_targ[k], upd_xy[2 * k], upd_xy[2 * k + 1]);
            upd_qpos[id] = -1;
        }
        upd_ids.clear();
        upd_targ.clear();
        upd_xy.clear();
        const Clock::time_point t_end = Clock::now();
        last_ms[0] = std::chrono::duration<double, std::milli>(t_end - t_step).count();
        last_ms[5] = std::chrono::duration<double, std::milli>(t_enq - t_step).count();
        last_ms[6] = std::chrono::duration<double, std::milli>(t_sync - t_enq).count();
        last_ms[7] = std::chrono::duration<double, std::milli>(t_end - t_sync).count();
        if (cfg.is_profiler || cfg.is_debug) {
            for (int k = 0; k < 4; k++) {
                float ms = 0;
                HIPCHK(hipEventElapsedTime(&ms, ev_p[k], ev_p[k + 1]));
                last_ms[k + 1] = ms;
            }
        } else if (step_timed) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev_s[0], ev_s[1]));
            last_ms[2] = ms;
        } else {
            last_ms[2] = -1.0;  // not a timed frame
        }
        for (int k = 0; k < 8; k++) {
            sum_ms[k] += last_ms[k];
            max_ms[k] = std::max(max_ms[k], last_ms[k]);
        }
    }

    // ------------------------------------------------------------ map
    void fill_map_start(float w, float h) {
        require_init();
        if (started) throw ApiError(GMAPF_E_STATE, "cannot modify map after start");
        cdt_v = {0.f, 0.f, 0.f, h, w, h, w, 0.f};  // GMAPF.cu:137-144
        cdt_e = {0, 1, 1, 2, 2, 3, 3, 0};
    }

    void fill_map_obstacle(const float* xy, int n) {
        require_init();
        if (started) throw ApiError(GMAPF_E_STATE, "cannot modify map after start");
        if (n < 1 || !xy) throw ApiError(GMAPF_E_ARG, "obstacle needs vertices");
        int start = (int)cdt_v.size() / 2;  // GMAPF.cu:147-159
        for (int i = 0; i < n; i++) {
            cdt_v.push_back(xy[2 * i]);
            cdt_v.push_back(xy[2 * i + 1]);
            if (i > 0) {
                cdt_e.push_back(start + i - 1);
                cdt_e.push_back(start + i);
            }
        }
        cdt_e.push_back(start + n - 1);
        cdt_e.push_back(start);
    }

    void fill_map_end() {
        require_init();
        if (started) throw ApiError(GMAPF_E_STATE, "cannot modify map after start");
        if (map_filled) throw ApiError(GMAPF_E_STATE, "map is already filled");
        if (cdt_v.empty()) throw ApiError(GMAPF_E_STATE, "FillMapStart was not called");
        std::vector<float> tri6;
        std::vector<int> nbr3;
        try {
            triangulate_map(cdt_v, cdt_e, 1.0f, tri6, nbr3);  // GMAPF.cu:165-169
        } catch (const std::exception& e) {
            throw ApiError(GMAPF_E_ARG, std::string("triangulation failed: ") + e.what());
        }
        fill_nodes((int)(tri6.size() / 6), tri6.data(), nbr3.data());
    }

    void fill_nodes(int n, const float* tri6, const int* nbr3) {
        require_init();
        if (started) throw ApiError(GMAPF_E_STATE, "cannot modify map after start");
        if (map_filled) throw ApiError(GMAPF_E_STATE, "map is already filled");
        try {
            nav.assign(n, tri6, nbr3);
        } catch (const std::exception& e) {
            throw ApiError(GMAPF_E_ARG, e.what());
        }
        map_filled = true;
    }

    // ------------------------------------------------------------ agents
    int add_agent(float x, float y) {
        require_init();
        if (started) throw ApiError(GMAPF_E_STATE, "cannot add agents after start");
        if (!map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        if (n_agents >= cfg.agents_max_count) throw ApiError(GMAPF_E_ARG, "AgentsMaxCount reached");
        int node = nav.locate(x, y);
        if (node < 0)
            throw ApiError(GMAPF_E_POS, "initial pos (" + std::to_string(x) + ", " + std::to_string(y) + ") is invalid");
        pos.push_back(x);
        pos.push_back(y);
        state.push_back(GMAPF_STATE_IDLE);
        curr.push_back(node);
        targ.push_back(-1);
        targ_pos.push_back(0.f);
        targ_pos.push_back(0.f);
        gen.push_back(0);
        return n_agents++;
    }

    // Agent id's target change into the next frame's pinned input (no step
    // is running, so nothing reads it now).
    void stage_update(int id, int node, float x, float y) {
        if ((int)upd_pos.size() < n_agents) {
            upd_pos.resize(n_agents, -1);
            upd_epoch_of.resize(n_agents, 0u);
        }
        int k = upd_pos[id];
        if (upd_epoch_of[id] != upd_epoch) {
            const int b = id / kUpdBucket;
            if (h_upd[b] == 0) upd_buckets.push_back(b);
            k = b * kUpdBucket + h_upd[b]++;
            upd_n++;
            upd_pos[id] = k;
            upd_epoch_of[id] = upd_epoch;
            h_upd[upd_nb + k] = id;
        }
        h_upd[upd_nb + upd_cap + k] = node;
        h_upd_xy[k] = make_float2(x, y);
    }

    void mark_host_dirty(int id) {
        const int w = id >> 6;
        if ((int)host_dirty_flag.size() <= w) host_dirty_flag.resize((size_t)w + 1, 0);
        if (!host_dirty_flag[w]) {
            host_dirty_flag[w] = 1;
            host_dirty.push_back(w);
        }
    }

    void set_target(int id, float x, float y) {
        require_init();
        if (!map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        if (id < 0 || id >= n_agents) throw ApiError(GMAPF_E_ARG, "agent id out of range");
        int node = nav.locate(x, y);
        if (node < 0)
            throw ApiError(GMAPF_E_POS, "target pos (" + std::to_string(x) + ", " + std::to_string(y) + ") is invalid");
        // a step in progress: the readback that ends it recounts the movers
        if (started && !step_going && state[id] == GMAPF_STATE_MOVE) movers--;
        state[id] = GMAPF_STATE_SEARCH;
        mark_host_dirty(id);
        if (!started) {
            targ[id] = node;
            targ_pos[2 * id] = x;
            targ_pos[2 * id + 1] = y;
            gen[id]++;
        } else if (!step_going) {
            stage_update(id, node, x, y);
        } else {
            if ((int)upd_qpos.size() < n_agents) upd_qpos.resize(n_agents, -1);
            int& k = upd_qpos[id];
            if (k >= 0) {  // an update of this agent is pending: the last one wins
                upd_targ[k] = node;
                upd_xy[2 * (size_t)k] = x;
                upd_xy[2 * (size_t)k + 1] = y;
            } else {
                k = (int)upd_ids.size();
                upd_ids.push_back(id);
                upd_targ.push_back(node);
                upd_xy.push_back(x);
                upd_xy.push_back(y);
            }
        }
    }

    void profiler_dump() const {
        require_init();
        if (!cfg.is_profiler) throw ApiError(GMAPF_E_STATE, "profiler disabled");  // GMAPF.cu:328-329
        const char* names[5] = {"step (wall)", "admit", "search", "finish", "crowd"};
        std::printf("[gmapf]----------------------\n");
        for (int k = 0; k < 5; k++)
            std::printf("[gmapf] %-12s avg: %.4f ms max: %.4f ms\n", names[k],
                        frames ? sum_ms[k] / frames : 0.0, max_ms[k]);
        std::printf("[gmapf] found %lld not found %lld expansions %lld rounds %lld\n", stats_host[0],
                    stats_host[1], stats_host[2], stats_host[5]);
        std::printf("[gmapf]----------------------\n");
        std::fflush(stdout);
    }
};

// ============================================================== C entry points
namespace {

template <class F>
int guarded(gmapf_handle h, F&& f) {
    if (!h) return GMAPF_E_ARG;
    DeviceScope ds(&h->scope);
    try {
        f();
        h->err.clear();
        return GMAPF_OK;
    } catch (const ApiError& e) {
        h->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        h->err = "host allocation failed";
        return GMAPF_E_NOMEM;
    } catch (const std::exception& e) {
        h->err = e.what();
        return GMAPF_E_INTERNAL;
    }
}

void check_id(gmapf_handle h, int id) {
    if (id < 0 || id >= h->n_agents) throw ApiError(GMAPF_E_ARG, "agent id out of range");
}

}  // namespace

extern "C" {

int gmapf_config_default(gmapf_config* c) {
    if (!c) return GMAPF_E_ARG;
    std::memset(c, 0, sizeof(*c));
    c->agents_max_count = 2048;
    c->agent_radius = 2.0f;
    c->agent_speed = 1.0f;
    c->path_finder_parallel_agents = 2048;
    c->path_finder_threads_per_agents = 4;
    c->path_finder_queue_capacity = 16;
    c->path_finder_heuristic_k = 1.0f;
    c->step_mode = GMAPF_MODE_PLAIN;
    c->round_budget = 256;
    return GMAPF_OK;
}

int gmapf_create(gmapf_handle* out) {
    if (!out) return GMAPF_E_ARG;
    try {
        *out = new gmapf_s();
    } catch (...) {
        return GMAPF_E_NOMEM;
    }
    return GMAPF_OK;
}

int gmapf_destroy(gmapf_handle h) {
    if (!h) return GMAPF_E_ARG;
    {
        DeviceScope ds(&h->scope);
        try {
            if (h->step_going) h->wait_step_end();
        } catch (...) {
        }
        h->release();
    }
    delete h;
    return GMAPF_OK;
}

const char* gmapf_last_error(gmapf_handle h) { return h ? h->err.c_str() : "null handle"; }

int gmapf_init(gmapf_handle h, const gmapf_config* c) {
    return guarded(h, [&] {
        if (!c) throw ApiError(GMAPF_E_ARG, "null config");
        h->init(*c);
    });
}

int gmapf_manual_start(gmapf_handle h) { return guarded(h, [&] { h->manual_start(); }); }

int gmapf_fill_map_start(gmapf_handle h, float w, float hh) {
    return guarded(h, [&] { h->fill_map_start(w, hh); });
}
int gmapf_fill_map_obstacle(gmapf_handle h, const float* xy, int n) {
    return guarded(h, [&] { h->fill_map_obstacle(xy, n); });
}
int gmapf_fill_map_end(gmapf_handle h) { return guarded(h, [&] { h->fill_map_end(); }); }

int gmapf_fill_map_manually(gmapf_handle h, const gmapf_node* nodes, int n) {
    return guarded(h, [&] {
        if (!nodes || n <= 0) throw ApiError(GMAPF_E_ARG, "no nodes");
        std::vector<float> tri6(6 * (size_t)n);
        std::vector<int> nbr3(3 * (size_t)n);
        for (int i = 0; i < n; i++) {
            const gmapf_node& nd = nodes[i];
            float* t = &tri6[6 * (size_t)i];
            t[0] = nd.p1[0];
            t[1] = nd.p1[1];
            t[2] = nd.p2[0];
            t[3] = nd.p2[1];
            t[4] = nd.p3[0];
            t[5] = nd.p3[1];
            for (int k = 0; k < 3; k++) nbr3[3 * (size_t)i + k] = nd.neibs_idx[k];
        }
        h->fill_nodes(n, tri6.data(), nbr3.data());
    });
}

int gmapf_add_agent(gmapf_handle h, float x, float y, int* id) {
    return guarded(h, [&] {
        int r = h->add_agent(x, y);
        if (id) *id = r;
    });
}

int gmapf_set_agent_targ_pos(gmapf_handle h, int id, float x, float y) {
    return guarded(h, [&] { h->set_target(id, x, y); });
}

int gmapf_set_agent_targ_pos_many(gmapf_handle h, int n, const int* ids, const float* xy, int* status) {
    return guarded(h, [&] {
        if (n < 0 || (n > 0 && (!ids || !xy))) throw ApiError(GMAPF_E_ARG, "bad arrays");
        int first_err = 0;
        std::string msg;
        for (int k = 0; k < n; k++) {
            int rc = GMAPF_OK;
            try {
                h->set_target(ids[k], xy[2 * k], xy[2 * k + 1]);
            } catch (const ApiError& e) {
                rc = e.code;
                if (!first_err) {
                    first_err = rc;
                    msg = e.what();
                }
            }
            if (status) status[k] = rc;
        }
        if (first_err && !status) throw ApiError(first_err, msg);
    });
}

int gmapf_async_step(gmapf_handle h, float dt) { return guarded(h, [&] { h->async_step(dt); }); }
int gmapf_wait_step_end(gmapf_handle h) { return guarded(h, [&] { h->wait_step_end(); }); }

int gmapf_get_agent_pos(gmapf_handle h, int id, float* xy) {
    return guarded(h, [&] {
        h->require_init();
        check_id(h, id);
        if (!xy) throw ApiError(GMAPF_E_ARG, "null output");
        xy[0] = h->pos[2 * id];
        xy[1] = h->pos[2 * id + 1];
    });
}

int gmapf_get_agent_state(gmapf_handle h, int id, int* st) {
    return guarded(h, [&] {
        h->require_init();
        check_id(h, id);
        if (!st) throw ApiError(GMAPF_E_ARG, "null output");
        *st = h->state[id];
    });
}

int gmapf_agent_positions(gmapf_handle h, const float** xy, int* n) {
    return guarded(h, [&] {
        h->require_init();
        if (xy) *xy = h->pos.data();
        if (n) *n = h->n_agents;
    });
}

int gmapf_profiler_dump(gmapf_handle h) { return guarded(h, [&] { h->profiler_dump(); }); }

int gmapf_agent_count(gmapf_handle h, int* n) {
    return guarded(h, [&] {
        h->require_init();
        if (n) *n = h->n_agents;
    });
}

int gmapf_get_config(gmapf_handle h, gmapf_config* out) {
    return guarded(h, [&] {
        h->require_init();
        if (!out) throw ApiError(GMAPF_E_ARG, "null output");
        *out = h->cfg;
    });
}

int gmapf_agent_states(gmapf_handle h, const int** st, int* n) {
    return guarded(h, [&] {
        h->require_init();
        if (st) *st = h->state.data();
        if (n) *n = h->n_agents;
    });
}

int gmapf_get_agents_soa(gmapf_handle h, float* xy, int* st, int* node) {
    return guarded(h, [&] {
        h->require_init();
        int N = h->n_agents;
        if (xy) std::memcpy(xy, h->pos.data(), (size_t)N * 2 * sizeof(float));
        if (st) std::memcpy(st, h->state.data(), (size_t)N * sizeof(int));
        if (node) {
            if (h->started && N > 0) {
                if (h->step_going) throw ApiError(GMAPF_E_STATE, "step is running");
                h->use_device();
                HIPCHK(hipMemcpy(node, h->dag.curr, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
            } else if (N > 0) {
                std::memcpy(node, h->curr.data(), (size_t)N * sizeof(int));
            }
        }
    });
}

int gmapf_get_agent_corridor(gmapf_handle h, int id, int* buf, int cap, int* len) {
    return guarded(h, [&] {
        h->require_init();
        check_id(h, id);
        if (!h->started) {
            if (len) *len = -1;
            return;
        }
        if (h->step_going) throw ApiError(GMAPF_E_STATE, "step is running");
        h->use_device();
        int L = 0;
        long long off = 0;
        int* slab = nullptr;
        HIPCHK(hipMemcpy(&L, h->dag.path_len + id, sizeof(int), hipMemcpyDeviceToHost));
        if (L <= 0) {
            if (len) *len = -1;
            return;
        }
        HIPCHK(hipMemcpy(&off, h->dag.path_off + id, sizeof(long long), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&slab, h->dag.slab_now, sizeof(int*), hipMemcpyDeviceToHost));
        if (slab != h->dag.slab[0] && slab != h->dag.slab[1])
            throw ApiError(GMAPF_E_INTERNAL, "corridor slab pointer is corrupt");
        if (off < 0 || off + L > h->dag.slab_cap) throw ApiError(GMAPF_E_INTERNAL, "corridor outside the slab");
        if (len) *len = L;
        if (buf && cap > 0)
            HIPCHK(hipMemcpy(buf, slab + off, (size_t)std::min(cap, L) * sizeof(int), hipMemcpyDeviceToHost));
    });
}

int gmapf_get_agent_steps(gmapf_handle h, int* step_idx, int* step_node, float* step_xy) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->started) throw ApiError(GMAPF_E_STATE, "not started");
        if (h->step_going) throw ApiError(GMAPF_E_STATE, "step is running");
        int N = h->n_agents;
        if (N == 0) return;
        h->use_device();
        if (step_idx) HIPCHK(hipMemcpy(step_idx, h->dag.step_idx, N * sizeof(int), hipMemcpyDeviceToHost));
        if (step_node) HIPCHK(hipMemcpy(step_node, h->dag.step_node, N * sizeof(int), hipMemcpyDeviceToHost));
        if (step_xy) HIPCHK(hipMemcpy(step_xy, h->dag.step_pos, N * sizeof(float2), hipMemcpyDeviceToHost));
    });
}

int gmapf_nav_encode(gmapf_handle h, void* buf, long long cap, long long* len) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        const std::vector<unsigned char> b =
            nav_encode(h->nav.V, h->nav.tri.data(), h->nav.nbr.data(), h->nav.cen.data());
        if (len) *len = (long long)b.size();
        if (buf) {
            if (cap < (long long)b.size()) throw ApiError(GMAPF_E_ARG, "buffer too small for the navmesh");
            std::memcpy(buf, b.data(), b.size());
        }
    });
}

int gmapf_fill_map_from_nav(gmapf_handle h, const void* buf, long long len) {
    return guarded(h, [&] {
        h->require_init();
        if (!buf || len <= 0) throw ApiError(GMAPF_E_ARG, "no navmesh bytes");
        int V = 0;
        std::vector<float> tri6;
        std::vector<int> nbr3;
        try {
            nav_decode(static_cast<const unsigned char*>(buf), (size_t)len, V, tri6, nbr3);
        } catch (const std::runtime_error& e) {
            throw ApiError(GMAPF_E_ARG, e.what());
        }
        h->fill_nodes(V, tri6.data(), nbr3.data());
    });
}

int gmapf_save_nav(gmapf_handle h, const char* path) {
    return guarded(h, [&] {
        h->require_init();
        if (!path) throw ApiError(GMAPF_E_ARG, "null path");
        if (!h->map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        const std::vector<unsigned char> b =
            nav_encode(h->nav.V, h->nav.tri.data(), h->nav.nbr.data(), h->nav.cen.data());
        FILE* f = std::fopen(path, "wb");
        if (!f) throw ApiError(GMAPF_E_ARG, std::string("cannot open ") + path + " for writing");
        const size_t w = std::fwrite(b.data(), 1, b.size(), f);
        const int rc = std::fclose(f);
        if (w != b.size() || rc != 0) throw ApiError(GMAPF_E_ARG, std::string("cannot write ") + path);
    });
}

int gmapf_load_nav(gmapf_handle h, const char* path) {
    return guarded(h, [&] {
        h->require_init();
        if (!path) throw ApiError(GMAPF_E_ARG, "null path");
        FILE* f = std::fopen(path, "rb");
        if (!f) throw ApiError(GMAPF_E_ARG, std::string("cannot open ") + path);
        std::vector<unsigned char> b;
        unsigned char chunk[1 << 16];
        size_t r;
        while ((r = std::fread(chunk, 1, sizeof(chunk), f)) > 0) b.insert(b.end(), chunk, chunk + r);
        const bool err = std::ferror(f) != 0;
        std::fclose(f);
        if (err) throw ApiError(GMAPF_E_ARG, std::string("cannot read ") + path);
        int V = 0;
        std::vector<float> tri6;
        std::vector<int> nbr3;
        try {
            nav_decode(b.data(), b.size(), V, tri6, nbr3);
        } catch (const std::runtime_error& e) {
            throw ApiError(GMAPF_E_ARG, std::string(path) + ": " + e.what());
        }
        h->fill_nodes(V, tri6.data(), nbr3.data());
    });
}

int gmapf_map_size(gmapf_handle h, int* nodes) {
    return guarded(h, [&] {
        h->require_init();
        if (nodes) *nodes = h->map_filled ? h->nav.V : 0;
    });
}

int gmapf_get_map(gmapf_handle h, float* tri6, int* nbr3, float* cen2) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        size_t V = h->nav.V;
        if (tri6) std::memcpy(tri6, h->nav.tri.data(), 6 * V * sizeof(float));
        if (nbr3) std::memcpy(nbr3, h->nav.nbr.data(), 3 * V * sizeof(int));
        if (cen2) std::memcpy(cen2, h->nav.cen.data(), 2 * V * sizeof(float));
    });
}

int gmapf_locate(gmapf_handle h, int n, const float* xy, int* out) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->map_filled) throw ApiError(GMAPF_E_STATE, "map is not filled");
        for (int i = 0; i < n; i++) out[i] = h->nav.locate(xy[2 * i], xy[2 * i + 1]);
    });
}

int gmapf_locate_device(gmapf_handle h, int n, const float* xy, int* out) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->started) throw ApiError(GMAPF_E_STATE, "not started");
        if (h->step_going) throw ApiError(GMAPF_E_STATE, "step is running");
        if (n <= 0) return;
        h->use_device();
        float2* dp = nullptr;
        int* dout = nullptr;
        HIPCHK(hipMalloc(&dp, n * sizeof(float2)));
        hipError_t e = hipMalloc(&dout, n * sizeof(int));
        if (e == hipSuccess) e = hipMemcpy(dp, xy, n * sizeof(float2), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = launch_locate(h->dmap, dp, n, dout, h->s1);
        if (e == hipSuccess) e = hipStreamSynchronize(h->s1);
        if (e == hipSuccess) e = hipMemcpy(out, dout, n * sizeof(int), hipMemcpyDeviceToHost);
        hipFree(dp);
        if (dout) hipFree(dout);
        HIPCHK(e);
    });
}

int gmapf_locate_device_noalloc(gmapf_handle h, int n, const float* xy, int* out) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->started) throw ApiError(GMAPF_E_STATE, "not started");
        if (h->step_going) throw ApiError(GMAPF_E_STATE, "step is running");
        if (n < 0 || n > kLocateScratchMax)
            throw ApiError(GMAPF_E_ARG, "gmapf_locate_device_noalloc takes at most " +
                                            std::to_string(kLocateScratchMax) + " points per call, got " +
                                            std::to_string(n));
        if (n == 0) return;
        if (!xy || !out) throw ApiError(GMAPF_E_ARG, "null points or output");
        h->use_device();
        HIPCHK(hipMemcpyAsync(h->loc_pts, xy, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, h->s1));
        HIPCHK(launch_locate(h->dmap, h->loc_pts, n, h->loc_out, h->s1));
        HIPCHK(hipMemcpyAsync(out, h->loc_out, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, h->s1));
        HIPCHK(hipStreamSynchronize(h->s1));
    });
}

int gmapf_get_stats(gmapf_handle h, long long* out8) {
    return guarded(h, [&] {
        h->require_init();
        for (int k = 0; k < 8; k++) out8[k] = h->stats_host[k];
    });
}

int gmapf_last_frame_ms(gmapf_handle h, double* out8) {
    return guarded(h, [&] {
        h->require_init();
        for (int k = 0; k < 8; k++) out8[k] = h->last_ms[k];
    });
}

int gmapf_get_hash(gmapf_handle h, int* keys, int* order, int* n) {
    return guarded(h, [&] {
        h->require_init();
        if (!h->started) throw ApiError(GMAPF_E_STATE, "not started");
        if (h->step_going) throw ApiError(GMAPF_E_STATE, "step is running");
        int N = h->n_agents;
        int total = 0;
        h->use_device();
        HIPCHK(hipMemcpy(&total, h->dfr.hstart + h->dfr.ncell, sizeof(int), hipMemcpyDeviceToHost));
        if (total < 0 || total > N) throw ApiError(GMAPF_E_INTERNAL, "hash total out of range");
        if (keys && N) HIPCHK(hipMemcpy(keys, h->dfr.hkey, N * sizeof(int), hipMemcpyDeviceToHost));
        if (order && total > 0) HIPCHK(hipMemcpy(order, h->dfr.horder, total * sizeof(int), hipMemcpyDeviceToHost));
        if (n) *n = total;
    });
}

int gmapf_get_slab_info(gmapf_handle h, long long* out4) {
    return guarded(h, [&] {
        h->require_init();
        if (!out4) throw ApiError(GMAPF_E_ARG, "null output");
        out4[0] = h->started ? h->dag.slab_cap : 0;
        out4[1] = std::min(h->slab_in_use, h->started ? h->dag.slab_cap : 0LL);
        out4[2] = h->compactions;
        out4[3] = h->slab_deferred;
    });
}

int gmapf_get_frame_kinds(gmapf_handle h, long long* out2) {
    return guarded(h, [&] {
        h->require_init();
        if (!out2) throw ApiError(GMAPF_E_ARG, "null output");
        out2[0] = h->frames;
        out2[1] = h->crowd_frames;
    });
}

int gmapf_get_launch_info(gmapf_handle h, int* out8) {
    return guarded(h, [&] {
        h->require_init();
        out8[0] = h->sp.T;
        out8[1] = h->sp.Q;
        out8[2] = h->cfg.path_finder_parallel_agents;
        out8[3] = h->sp.budget;
        out8[4] = h->search_grid;
        out8[5] = h->cfg.step_mode;
        out8[6] = h->map_filled ? h->nav.V : 0;
        out8[7] = h->persistent;
    });
}

}  // extern "C"
