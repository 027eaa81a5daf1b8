// GMAPF.h -- header-only C++ drop-in for the reference's gmapf::GMAPF
// (include/gmapf/GMAPF.h:30-98, Config.h:5-17, Agent.h:7-27,
// misc/CuNodesMap.h:14-21, misc/V2Float.h of the reference), implemented
// over the C ABI in gmapf_c.h.  No CUDA or HIP headers are needed by callers.
//
// Differences a caller can observe:
//   * errors are thrown as std::runtime_error with the library's message
//     (the reference throws MSVC's std::exception(const char*));
//   * SetAgentTargPos during a running step is applied at the next step
//     instead of being lost (README.md:91 of the reference);
//   * extra Config fields select the batched step and its knobs.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gmapf_c.h"

namespace gmapf {

struct V2Float {
    float X;
    float Y;
    V2Float() : X(0), Y(0) {}
    V2Float(float x, float y) : X(x), Y(y) {}
    float GetLengthSqr() const { return X * X + Y * Y; }
    float GetLength() const { return std::sqrt(X * X + Y * Y); }
    friend V2Float operator+(const V2Float& a, const V2Float& b) { return {a.X + b.X, a.Y + b.Y}; }
    friend V2Float operator-(const V2Float& a, const V2Float& b) { return {a.X - b.X, a.Y - b.Y}; }
    friend V2Float operator*(const V2Float& a, float k) { return {a.X * k, a.Y * k}; }
    friend V2Float operator/(const V2Float& a, float k) { return {a.X / k, a.Y / k}; }
};

enum class EAgentState { Idle = GMAPF_STATE_IDLE, Search = GMAPF_STATE_SEARCH, Move = GMAPF_STATE_MOVE };

struct Config {
    int AgentsMaxCount = 0;
    float AgentRadius = 0;
    float AgentSpeed = 0;
    int PathFinderParallelAgents = 0;
    int PathFinderThreadsPerAgents = 0;
    int PathFinderQueueCapacity = 0;
    float PathFinderHeuristicK = 1;
    bool IsDebug = false;
    bool IsProfiler = false;
    // extensions
    int StepMode = GMAPF_MODE_PLAIN;
    int RoundBudget = 256;
    float HashCell = 0;
    int Device = 0;
    long long CorridorSlab = 0;  // corridor storage in ints; 0: sized from agents and map
    int SearchesPerWave = 0;     // searches per 64-lane search workgroup; 0: 64 / T
    int SearchTiming = 0;        // time the search launch of every k-th frame; 0: none
    int SearchWorkgroups = 0;    // batched form: cap on the persistent search grid; 0: automatic
    bool SearchWideKeys = false;  // queue keys without G and side also on maps of <= 32,768 nodes
};

class CuNodesMap {
   public:
    static constexpr int INVALID = -1;
    static constexpr int NEIBS_MAX_COUNT = 3;
    struct Node {
        V2Float P1;
        V2Float P2;
        V2Float P3;
        V2Float PCenter;  // ignored, recomputed as (P1 + P2 + P3) / 3
        int NeibsIdx[NEIBS_MAX_COUNT];
    };
};
static_assert(sizeof(CuNodesMap::Node) == sizeof(gmapf_node), "Node layout must match gmapf_node");

class GMAPF {
   public:
    GMAPF() { check(gmapf_create(&h_), nullptr); }
    GMAPF(const GMAPF&) = delete;
    GMAPF(GMAPF&&) = delete;
    GMAPF& operator=(const GMAPF&) = delete;
    GMAPF& operator=(GMAPF&&) = delete;
    ~GMAPF() {
        if (h_) gmapf_destroy(h_);
    }

    void Init(const Config& c) {
        gmapf_config cc;
        gmapf_config_default(&cc);
        cc.agents_max_count = c.AgentsMaxCount;
        cc.agent_radius = c.AgentRadius;
        cc.agent_speed = c.AgentSpeed;
        cc.path_finder_parallel_agents = c.PathFinderParallelAgents;
        cc.path_finder_threads_per_agents = c.PathFinderThreadsPerAgents;
        cc.path_finder_queue_capacity = c.PathFinderQueueCapacity;
        cc.path_finder_heuristic_k = c.PathFinderHeuristicK;
        cc.is_debug = c.IsDebug;
        cc.is_profiler = c.IsProfiler;
        cc.step_mode = c.StepMode;
        cc.round_budget = c.RoundBudget;
        cc.hash_cell = c.HashCell;
        cc.device = c.Device;
        cc.corridor_slab = c.CorridorSlab;
        cc.searches_per_wave = c.SearchesPerWave;
        cc.search_timing = c.SearchTiming;
        cc.search_workgroups = c.SearchWorkgroups;
        cc.search_wide_keys = c.SearchWideKeys ? 1 : 0;
        check(gmapf_init(h_, &cc));
    }
    void ManualStart() { check(gmapf_manual_start(h_)); }

    int AddAgent(const V2Float& p) {
        int id = -1;
        check(gmapf_add_agent(h_, p.X, p.Y, &id));
        return id;
    }
    void SetAgentTargPos(int id, const V2Float& p) { check(gmapf_set_agent_targ_pos(h_, id, p.X, p.Y)); }
    void SetAgentTargetPos(int id, const V2Float& p) { SetAgentTargPos(id, p); }

    void FillMapStart(const V2Float& size) { check(gmapf_fill_map_start(h_, size.X, size.Y)); }
    void FillMapStart(float w, float h) { check(gmapf_fill_map_start(h_, w, h)); }
    void FillMapObstacle(const std::vector<V2Float>& v) {
        check(gmapf_fill_map_obstacle(h_, v.empty() ? nullptr : &v[0].X, (int)v.size()));
    }
    void FillMapEnd() { check(gmapf_fill_map_end(h_)); }
    void FillMapManually(const std::vector<CuNodesMap::Node>& nodes) {
        check(gmapf_fill_map_manually(h_, reinterpret_cast<const gmapf_node*>(nodes.data()), (int)nodes.size()));
    }
    // Map fixture (no reference counterpart): the filled map as a "GMAPFNAV"
    // byte string or file, loadable in place of FillMapManually.
    std::vector<unsigned char> SaveNav() const {
        long long n = 0;
        check(gmapf_nav_encode(h_, nullptr, 0, &n));
        std::vector<unsigned char> b((size_t)n);
        check(gmapf_nav_encode(h_, b.data(), n, &n));
        return b;
    }
    void SaveNav(const std::string& path) const { check(gmapf_save_nav(h_, path.c_str())); }
    void FillMapFromNav(const std::vector<unsigned char>& bytes) {
        check(gmapf_fill_map_from_nav(h_, bytes.data(), (long long)bytes.size()));
    }
    void LoadNav(const std::string& path) { check(gmapf_load_nav(h_, path.c_str())); }

    void AsyncStep(float dt) { check(gmapf_async_step(h_, dt)); }
    void WaitStepEnd() { check(gmapf_wait_step_end(h_)); }

    const V2Float& GetAgentPos(int id) {
        const float* xy = nullptr;
        int n = 0;
        check(gmapf_agent_positions(h_, &xy, &n));
        if (id < 0 || id >= n) throw std::runtime_error("agent id out of range");
        return *reinterpret_cast<const V2Float*>(xy + 2 * id);
    }
    EAgentState GetAgentState(int id) {
        int s = 0;
        check(gmapf_get_agent_state(h_, id, &s));
        return static_cast<EAgentState>(s);
    }
    void ProfilerDump() const { check(gmapf_profiler_dump(h_)); }

    gmapf_handle Handle() const { return h_; }

   private:
    gmapf_handle h_ = nullptr;

    void check(int rc) const { check(rc, h_); }
    static void check(int rc, gmapf_handle h) {
        if (rc != GMAPF_OK) throw std::runtime_error(std::string("[gmapf] ") + gmapf_last_error(h));
    }
};

}  // namespace gmapf
