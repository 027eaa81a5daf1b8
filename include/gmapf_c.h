/*
 * gmapf_c.h -- C ABI of the MI355X-native gmapf per-frame step.
 *
 * One opaque handle = one gmapf::GMAPF object of the reference
 * (include/gmapf/GMAPF.h:30-98).  Every function returns GMAPF_OK (0) or a
 * negative GMAPF_E_* code; gmapf_last_error() returns the message of the last
 * failure on that handle (the reference throws std::exception("msg"), e.g.
 * src/GMAPF.cu:75,202-206,222-228,241-245,257-258,329).  No CUDA, HIP or torch
 * types appear here.  Not thread-safe, like the reference (one host thread
 * per handle).
 */
#ifndef GMAPF_C_H
#define GMAPF_C_H

#ifdef __cplusplus
extern "C" {
#endif

#define GMAPF_OK 0
#define GMAPF_E_ARG -1      /* bad argument (null pointer, id out of range) */
#define GMAPF_E_STATE -2    /* call order violated (not initialised, started, ...) */
#define GMAPF_E_POS -3      /* position not on the navmesh */
#define GMAPF_E_DEVICE -4   /* HIP runtime error */
#define GMAPF_E_NOMEM -5    /* allocation failed */
#define GMAPF_E_INTERNAL -6

typedef struct gmapf_s* gmapf_handle;

/* Agent states: reference EAgentState (include/gmapf/Agent.h:7-12). */
#define GMAPF_STATE_IDLE 0
#define GMAPF_STATE_SEARCH 1
#define GMAPF_STATE_MOVE 2

/* Step forms (no reference counterpart; the reference has only the plain one). */
#define GMAPF_MODE_PLAIN 0   /* every admitted search runs to completion in its frame */
#define GMAPF_MODE_BATCHED 1 /* persistent work-queue search, round budget per frame,
                                whole frame replayed from a captured hipGraph */

/* Reference gmapf::Config (include/gmapf/Config.h:5-17) + build knobs. */
typedef struct gmapf_config {
    int agents_max_count;              /* Config::AgentsMaxCount */
    float agent_radius;                /* Config::AgentRadius */
    float agent_speed;                 /* Config::AgentSpeed */
    int path_finder_parallel_agents;   /* Config::PathFinderParallelAgents (P) */
    int path_finder_threads_per_agents;/* Config::PathFinderThreadsPerAgents (T: 2,4,..,64) */
    int path_finder_queue_capacity;    /* Config::PathFinderQueueCapacity (Q) */
    float path_finder_heuristic_k;     /* Config::PathFinderHeuristicK (K); finite and >= 0 */
    int is_debug;                      /* Config::IsDebug: sync + check after every phase */
    int is_profiler;                   /* Config::IsProfiler: per-phase event timing */
    /* build knobs, no reference counterpart */
    int step_mode;                     /* GMAPF_MODE_* */
    int round_budget;                  /* batched: search rounds per search per frame (>0) */
    float hash_cell;                   /* avoidance hash cell size; <= 0: just above 2.5 * agent_radius
                                          (a 3 x 3 cell stencil) */
    float rot_cs, rot_sn;              /* FP32 cos/sin of the 10-degree avoidance turn;
                                          both 0: computed from 10 * pi_f / 180 */
    int device;                        /* HIP device ordinal */
    long long corridor_slab;           /* corridor storage, ints (two buffers of this size are
                                          kept for compaction); <= 0: min(N * V, max(N * 1024,
                                          16 * V)); at least V.  A full slab delays a found
                                          search's corridor to a later frame; compaction keeps
                                          only the corridors of moving agents */
    int searches_per_wave;             /* searches a 64-lane search workgroup holds, 1 .. 64 / T;
                                          <= 0: 64 / T.  Fewer than 64 / T leaves lanes idle
                                          and puts more workgroups on each SIMD, which hides
                                          memory latency when searches are too few to fill it */
    int search_timing;                 /* HIP events around the search launch of every k-th
                                          frame (frames 0, k, 2k, ...), read by
                                          gmapf_last_frame_ms [2]; 0: no frame.  The two event
                                          nodes lengthen a graph-launched frame by about 20 us */
    int search_workgroups;             /* batched form: at most this many search workgroups
                                          (the persistent grid); <= 0: one per
                                          searches_per_wave searches in flight, up to what
                                          the GPU holds at once.  A smaller grid makes the
                                          workgroups take searches by atomic ticket */
    int search_wide_keys;              /* 1: queue keys hold (F, node) only and a popped node's G
                                          and side are read from its visited entry; 0: on maps
                                          of at most 32,768 nodes the keys also carry them (no
                                          visited-entry read per expansion).  Results are the
                                          same either way */
} gmapf_config;

/* Reference CuNodesMap::Node (include/gmapf/misc/CuNodesMap.h:14-21);
 * pcenter is ignored and recomputed (src/GMAPF.cu:215). */
typedef struct gmapf_node {
    float p1[2];
    float p2[2];
    float p3[2];
    float pcenter[2];
    int neibs_idx[3];
} gmapf_node;

/* ---- lifecycle: GMAPF() / ~GMAPF() / Init / ManualStart (GMAPF.h:33-41) */
int gmapf_config_default(gmapf_config* cfg);
int gmapf_create(gmapf_handle* out);
int gmapf_destroy(gmapf_handle h);
const char* gmapf_last_error(gmapf_handle h);
int gmapf_init(gmapf_handle h, const gmapf_config* cfg);          /* GMAPF::Init, GMAPF.cu:58 */
int gmapf_manual_start(gmapf_handle h);                           /* GMAPF::ManualStart, GMAPF.cu:72 */

/* ---- map build (GMAPF.h:46-50) */
int gmapf_fill_map_start(gmapf_handle h, float width, float height);   /* GMAPF.cu:133 */
int gmapf_fill_map_obstacle(gmapf_handle h, const float* xy, int n);   /* GMAPF.cu:147, n vertices */
int gmapf_fill_map_end(gmapf_handle h);                                /* GMAPF.cu:161 */
int gmapf_fill_map_manually(gmapf_handle h, const gmapf_node* nodes, int n); /* GMAPF.cu:199; n <= 2^26 */

/* ---- map fixture (no reference counterpart): a filled navmesh as one
 * versioned, checksummed byte string ("GMAPFNAV" v1: vertices, centroids,
 * CSR neighbour lists), loadable in place of FillMapManually (same call order
 * and errors as gmapf_fill_map_manually).  For callers that build a map once
 * with the triangulator and load it many times. */
/* bytes of the current map; buf NULL: only *len is set */
int gmapf_nav_encode(gmapf_handle h, void* buf, long long cap, long long* len);
int gmapf_fill_map_from_nav(gmapf_handle h, const void* buf, long long len);
int gmapf_save_nav(gmapf_handle h, const char* path);
int gmapf_load_nav(gmapf_handle h, const char* path);

/* ---- agents (GMAPF.h:43-44) */
int gmapf_add_agent(gmapf_handle h, float x, float y, int* id);        /* GMAPF.cu:219 */
int gmapf_set_agent_targ_pos(gmapf_handle h, int id, float x, float y);/* GMAPF.cu:238 */

/* ---- frame (GMAPF.h:52-53) */
int gmapf_async_step(gmapf_handle h, float dt);                        /* GMAPF.cu:255 */
int gmapf_wait_step_end(gmapf_handle h);                               /* GMAPF.cu:291 */

/* ---- readback (GMAPF.h:55-58) */
int gmapf_get_agent_pos(gmapf_handle h, int id, float* xy);            /* GMAPF.cu:316 */
int gmapf_get_agent_state(gmapf_handle h, int id, int* state);         /* GMAPF.cu:321 */
/* Host mirror of all positions (x0,y0,x1,y1,...), valid until destroy; the
 * reference's GetAgentPos returns a reference into the same kind of mirror. */
int gmapf_agent_positions(gmapf_handle h, const float** xy, int* n);
int gmapf_profiler_dump(gmapf_handle h);                               /* GMAPF.cu:326 */

/* ---- extensions (no reference counterpart): bulk, introspection, bench */
int gmapf_agent_count(gmapf_handle h, int* n);
/* the configuration given to gmapf_init */
int gmapf_get_config(gmapf_handle h, gmapf_config* out);
/* Host mirror of all states (GMAPF_STATE_*), valid until destroy, refreshed
 * by every WaitStepEnd. */
int gmapf_agent_states(gmapf_handle h, const int** state, int* n);
int gmapf_get_agents_soa(gmapf_handle h, float* xy, int* state, int* curr_node);
int gmapf_set_agent_targ_pos_many(gmapf_handle h, int n, const int* ids, const float* xy,
                                  int* status /* per item 0 or GMAPF_E_POS, may be NULL */);
int gmapf_get_agent_corridor(gmapf_handle h, int id, int* buf, int cap, int* len);
int gmapf_get_agent_steps(gmapf_handle h, int* step_idx, int* step_node, float* step_xy);
int gmapf_map_size(gmapf_handle h, int* nodes);
int gmapf_get_map(gmapf_handle h, float* tri6, int* nbr3, float* cen2);
int gmapf_locate(gmapf_handle h, int n, const float* xy, int* out);    /* host, exact */
int gmapf_locate_device(gmapf_handle h, int n, const float* xy, int* out); /* same, on GPU */
/* Same result as gmapf_locate_device without allocating: no reference
 * counterpart, for callers that cannot allocate device memory (code under
 * stream capture, serving loops that forbid allocation).  Points and results
 * are staged in a fixed per-GPU buffer of this library (12 MiB, held until
 * the process ends), so n is at most 1,048,576 (GMAPF_E_ARG beyond).  Calls
 * that overlap on one GPU, from any handle, share that buffer and race. */
int gmapf_locate_device_noalloc(gmapf_handle h, int n, const float* xy, int* out);
/* cumulative device counters: [0] found [1] not found [2] expansions
 * [3] discoveries [4] corridor nodes [5] rounds [6] admitted [7] frames */
int gmapf_get_stats(gmapf_handle h, long long* out8);
/* last frame, ms: [0] host wall (AsyncStep entry -> WaitStepEnd return);
 * with profiling, device phases [1] admit [2] search [4] crowd ([3] unused);
 * without, [2] is the search launch of a frame timed by search_timing (-1 for
 * other frames); host parts [5] enqueue [6] wait for the device [7] readback */
int gmapf_last_frame_ms(gmapf_handle h, double* out8);
/* spatial hash of the last frame that moved agents: cell key per agent (-1 for
 * a NaN position), and the stable counting sort of the agents by key (ascending
 * key, ascending agent index inside a cell; *n = agents with a key) */
int gmapf_get_hash(gmapf_handle h, int* keys, int* order, int* n);
/* resolved build knobs: [0] T [1] Q [2] P [3] budget [4] search grid [5] mode [6] V */
int gmapf_get_launch_info(gmapf_handle h, int* out8);
/* frames so far: [0] all, [1] those that ran without admission and search (after the
 * frame before no agent searched and some moved, and no target changed since: the
 * frame is the move, the spatial hash, the avoidance, progress and the readback);
 * no reference counterpart */
int gmapf_get_frame_kinds(gmapf_handle h, long long* out2);
/* corridor slab after the last frame: [0] capacity (ints per buffer) [1] ints in use
 * [2] compactions so far [3] corridors deferred for lack of space so far */
int gmapf_get_slab_info(gmapf_handle h, long long* out4);

#ifdef __cplusplus
}
#endif
#endif
