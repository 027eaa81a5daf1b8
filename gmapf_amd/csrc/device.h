// device.h -- device-resident data layouts shared by the host engine and the
// HIP kernels of the per-frame step.
//
// Navmesh (CSR triangle graph, SoA):
//   nbr[V]   int4  {n0, n1, n2, unused}: compacted neighbours, -1 tail
//            (reference Node::NeibsIdx, CuNodesMap.h:14-21, GMAPF.cu:185-190)
//   cen[V]   float2 centroid ((P1+P2)+P3)/3 (GMAPF.cu:215)
//   rec[3V]  int4 x 3 per node, the search's view of it in one 48-byte read:
//            {n0, n1, n2, -1}, {bits of cen[n0], cen[n1]}, {bits of cen[n2], 0, 0}
//            (a missing neighbour's centroid is 0)
//   tri[V]   3 x float2 vertices, for point-in-triangle tests
//   point-location grid: cell_start/cell_tris (ascending triangle ids per
//   cell, triangles listed in every cell their margin-expanded bbox touches)
//   plus a short list of near-degenerate triangles that is always scanned.
// Agents (SoA, one entry per agent id):
//   pos float2, curr int, targ_pos float2, targ int, state int,
//   step_idx int, step_node int, step_pos float2, path_len int, path_off
//   long long (corridor = slab[path_off .. path_off + path_len)), gen int
//   (bumped by SetAgentTargPos), inflight int (search slot or -1), in_m int
//   (moving set of this frame).
// Corridor slab: two buffers of slab_cap ints; *slab_now is the one in use.
//   A found search takes path_len ints at *slab_top (atomic bump); a full
//   slab defers the corridor (the search keeps its slot, status 4, and
//   retries next frame).  Compaction (when the host sees the slab half
//   full) copies the corridors of moving agents into the other buffer in
//   agent order and drops the rest.
// Search slots (P of them, PathFinderParallelAgents):
//   mark[slot * VP + node] byte (m << 1) | dir, m = stamp % 127 + 1 of the
//   search that claimed the node: the probe a search makes for every
//   neighbour reads one byte (64 nodes per cache line); a node belongs to
//   the current search only when its m matches, so marks are cleared only
//   once per epoch of 127 stamps (by the epoch's first search to seed in
//   the slot, recorded in clr), and tables are never cleared per search
//   (replaces PathFinder.h:119-136).
//   vis[slot * V + node] uint2 {G << 1 | dir, parent}: read for the node a
//   lane pops and by the corridor walk, i.e. only for nodes of the current
//   search, so it needs no tag.
//   qsave[slot][lane][Q] uint64 keys + qcnt[slot][lane]: queues parked
//   between frames by the batched form.
#pragma once
#include <cstdint>

namespace gmapf_amd {

enum AgentState : int { kIdle = 0, kSearch = 1, kMove = 2 };

struct DevMap {
    int V;
    const int4* nbr;
    const float2* cen;
    const float2* tri;  // 3 per triangle
    const int4* rec;    // 3 per node: neighbours and their centroids
    // point location grid
    float gx0, gy0, inv_cs;
    float rx0, ry0, rx1, ry1;
    int nx, ny;
    const int* cell_start;
    const int* cell_tris;
    const int* global_tris;
    int n_global;
    // spatial hash (agents), cell keys over the navmesh bbox
    float hx0, hy0, hinv;
    int hnx, hny, hk;  // hk = stencil radius in cells
};

struct DevAgents {
    int n;
    float2* pos;
    int* curr;
    float2* targ_pos;
    int* targ;
    int* state;
    int* step_idx;
    int* step_node;
    float2* step_pos;
    int* path_len;
    long long* path_off;
    int** slab_now;                // device word: the slab buffer in use
    int* slab[2];
    long long slab_cap;            // ints per buffer
    unsigned long long* slab_top;  // ints allocated in the buffer in use
    unsigned long long* slab_deferred;  // corridors deferred for lack of slab space (total)
    int* gen;
    int* inflight;
    int* in_m;
};

// Per-slot search record (SoA over P slots).
struct DevSlots {
    int P;
    int* agent;     // -1 free
    int* gen;       // agent gen at admission
    int* S;
    int* Tg;
    unsigned* stamp;
    unsigned* clr;  // epoch (stamp / 127) of the slot's last mark clear
    int* status;    // 0 running, 1 found, 2 exhausted, 3 fresh (needs seeding),
                    // 4 found, corridor deferred (needs slab space),
                    // 5 found, corridor walk parked (batched form)
    int* mS;        // meet of a found search: start-side and target-side nodes
    int* mT;
    int2* corr;     // batched form, while the corridor is walked: its slab offset and length
    int4* walk;     // parked walk: {node, index - offset} of the start side, then the
                    // target side (node -1: side done)
    unsigned char* mark;    // P * VP  (m << 1) | dir
    uint2* vis;             // P * V   {G << 1 | dir, parent}
    long long VP;           // mark stride: V rounded up to 64
    unsigned long long* qsave;  // P * T * Q
    int* qcnt;              // P * T
};

// Per-frame scratch / counters (device).
// Device-scope atomics on one address serialise (about 35 ns each on this
// part): 1,024 search workgroups adding their counters to one set of words
// at the end of a launch made it 12% longer.  Striped over 64 sets, a word
// takes a few adds per launch.
constexpr int kStatStripes = 64;

// Agents per bucket of the target-update input (one admission workgroup's
// agents, so one workgroup can apply its agents' changes and then admit
// them, with no other workgroup involved).
constexpr int kUpdBucket = 1024;

struct DevFrame {
    int* blk;        // per-block request counts, then their exclusive prefixes
    int* mblk;       // per-block moving-agent counts
    int* sblk;       // per-block free-slot counts, then their exclusive prefixes
    int* free_list;  // P
    int* adm_agent;  // P: agent of the k-th admission this frame
    int* active;     // P: active slot list for the search kernel
    int* counters;   // [0] n_requests [1] n_active [2] work ticket [3] n_ended
                     // [4] n_free [5] moving count (0 between frames) [6] n_admitted
                     // [7] unused
                     // [8] unused
                     // [9] ticket stride: ticket t takes active[(t * stride) % n_active]
    unsigned long long* stats;  // [0] found [1] fail [2] expansions [3] discoveries
                                // [4] corridor nodes [5] rounds [6] admitted
    // this frame's additions to stats, in kStatStripes stripes of 8 (the
    // workgroups of a launch add to stripe blockIdx % kStatStripes); the
    // export folds them into stats and zeroes them
    unsigned long long* stats_part;
    // spatial hash
    int* hkey;        // N: cell key of the moved position, -1 for a NaN position
    int* hslot;       // N: arrival slot of the agent in its cell (histogram atomic)
    int* hcount;      // ncell + 1 (cell ncell: agents with a NaN position), zero between frames
    int* hstart;      // ncell + 2: first entry of each cell; [ncell] = agents hashed, [ncell + 1] = N
    int* horder;      // N: hashed agents grouped by cell, ascending index inside a cell
    float4* hent;     // N: {x, y, bits of the agent index, 0} grouped by cell,
                      // arrival order inside a cell, then the NaN agents
    int* hblk;        // scan partials over cells
    int ncell;
    float* stepv;                // device copy of speed * dt, made by the admission kernel
    // pinned host memory read / written by the frame's kernels (no copies,
    // so a captured frame is one graph launch)
    const float* h_step;         // speed * dt of the frame
    // target updates, bucketed by kUpdBucket agents: bucket b's changes are
    // entries b * kUpdBucket .. + h_upd_bcnt[b] - 1, each of an agent of b
    // (at most one per agent)
    const int* h_upd_bcnt;       // [ceil(N / kUpdBucket)] changes per bucket
    const int* h_upd_ids;        // [N] agent ids
    const int* h_upd_targ;       // [N] target nodes
    const float2* h_upd_xy;      // [N] target positions
    float2* h_pos_out;           // [N] positions after the frame
    int* h_state_out;            // [N] states after the frame (written where they changed)
    int* state_prev;             // [N] device copy of h_state_out
    unsigned long long* h_stats_out;  // [8] cumulative counters
    unsigned long long* h_slab_out;   // [2] slab top, deferred corridors
    int* h_group_out;                 // [ceil(N / 64)] per 64 agents after the frame: agents in
                                      // Search (bits 0-7) and in Move (bits 8-15), and bit 16
                                      // when the group's states in h_state_out were rewritten
    // corridor slab compaction scratch
    int* slab_len;                    // N + 1: live corridor lengths, then their offsets
    int* slab_blk;                    // scan partials
};

struct SearchParams {
    int T;
    int Q;
    float K;
    int budget;  // max rounds per search per launch; <= 0: unbounded
    int seg;     // searches per workgroup (segments 0 .. seg-1 take work), 1 .. 64 / T
    int ident;   // > 0: every agent has a slot of its own (slot i = agent i, P >= N);
                 // segment i takes slot i when it is in flight, and ident is the slot count
    int wide;    // 1: queue keys without G and side, also on maps that fit compact keys
};

struct CrowdParams {
    float cd;      // 2.5 * radius
    float cd2;
    float slack;   // radius
    float rot_cs;
    float rot_sn;
};

}  // namespace gmapf_amd
