/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Single-threaded CPU restatement of gmapf's per-frame step (path requests,
 * bidirectional multi-lane search, corridor build, move, avoidance, progress)
 * with the canonical tie and ordering rules that make the reference's racy GPU
 * step deterministic.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker.  The product library
 * (gmapf_amd) never links or calls it.
 *
 * Every function cites the reference lines it restates (paths relative to the
 * reference repository root).
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   point location   brute force == reference scan (CuNodesMap.h:46-66,97-112);
 *                    grid accelerator == brute force on random + edge points
 *   queue            == brute-force "keep the Q smallest by (F, node)" model
 *   search           tiny hand graphs checked exhaustively; tree graphs must give
 *                    the unique tree path; K=0 gives level-synchronous BFS depths
 *   corridor         adjacency / endpoints / length = G(mS)+G(mT)+1 invariants
 *   move, avoidance  closed forms on 1-, 2- and 3-agent scenes
 *   hash avoidance   == brute-force O(N^2) avoidance bit for bit
 * The FP32 rotation constants (cos/sin of 10 degrees) are passed in by the
 * caller; their bit patterns versus CUDA's cosf/sinf are parity unpinned.
 */
#ifndef GMAPF_ORACLE_H
#define GMAPF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_map orc_map;
typedef struct orc_world orc_world;

/* ---- navmesh ---------------------------------------------------------- */
/* tri: 6*V floats (p1x,p1y,p2x,p2y,p3x,p3y); nbr: 3*V ints, compacted, -1 tail.
 * Centroids are computed as ((P1+P2)+P3)/3.0f (GMAPF.cu:215). */
orc_map* orc_map_create(int V, const float* tri, const int* nbr);
void orc_map_destroy(orc_map* m);
int orc_map_count(const orc_map* m);
void orc_map_centroids(const orc_map* m, float* cxy /* 2V */);
/* min index of a triangle containing (x,y) or -1; brute force O(V) scan */
int orc_locate_brute(const orc_map* m, float x, float y);
/* same result through the oracle's own exact grid accelerator */
int orc_locate_grid(const orc_map* m, float x, float y);
void orc_locate_many(const orc_map* m, int n, const float* xy, int* out, int use_grid);

/* ---- bounded queue model (CuQueue.h:33-70 with (F,node) total order) --- */
/* ops[i] = 0 push (F[i], node[i]); 1 pop (writes popped node or -1 to out[i]) */
void orc_queue_script(int Q, int nops, const int* ops, const float* F,
                      const int* node, int* out);

/* ---- one search (PathFinder.h:156-307) --------------------------------- */
/* returns 1 found, 0 not found; corridor written to buf (cap ints), *len.
 * stats[0]=rounds stats[1]=expansions stats[2]=discoveries stats[3]=corridor len
 * stats[4]=mS stats[5]=mT */
int orc_search(const orc_map* m, int S, int Tg, int T, int Q, float K,
               int* buf, int cap, int* len, long long* stats);

/* ---- world (GMAPF.cu:219-314, AgentsMover.h, PathFinder.h) ------------- */
typedef struct orc_config {
    int agents_max;
    float radius;
    float speed;
    int parallel;      /* P: max searches in flight */
    int threads;       /* T lanes per search */
    int queue_cap;     /* Q */
    float heur_k;      /* K */
    int round_budget;  /* rounds per search per frame; 0 = run to completion */
    float rot_cs;      /* FP32 cos(10 deg) */
    float rot_sn;      /* FP32 sin(10 deg) */
    float hash_cell;   /* >0: spatial-hash avoidance with this cell size; 0: O(N^2) */
    int use_grid_locate; /* 1: exact grid point location; 0: brute force */
} orc_config;

orc_world* orc_world_create(const orc_map* m, const orc_config* cfg);
void orc_world_destroy(orc_world* w);
/* returns id >= 0 or -1 when the position is off the navmesh / capacity full */
int orc_world_add_agent(orc_world* w, float x, float y);
/* returns 0, or -1 when off the navmesh / bad id */
int orc_world_set_target(orc_world* w, int id, float x, float y);
void orc_world_step(orc_world* w, float dt);
int orc_world_count(const orc_world* w);
/* xy: 2N floats, st: N ints (0 Idle, 1 Search, 2 Move), node: N ints (CurrNodeIdx) */
void orc_world_get(const orc_world* w, float* xy, int* st, int* node);
/* corridor of an agent (its current path) -> len, or -1 if none */
int orc_world_corridor(const orc_world* w, int id, int* buf, int cap);
/* per-agent path-step state: step_idx, step_node, step_pos xy */
void orc_world_get_steps(const orc_world* w, int* step_idx, int* step_node, float* step_xy);
/* cumulative counters: [0] found [1] not found [2] expansions [3] discoveries
 * [4] corridor nodes [5] rounds [6] admitted [7] frames */
void orc_world_stats(const orc_world* w, long long* out8);
/* counting-sort permutation used by the hash avoidance of the last frame */
int orc_world_last_sort(const orc_world* w, int* order, int* keys);

#ifdef __cplusplus
}
#endif
#endif
