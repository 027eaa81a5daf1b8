/*
 * gmapf_serve.h -- serving loop over several device-resident maps.
 *
 * No reference counterpart: the reference steps one map per GMAPF object and
 * has no request queue.  A server takes N engines (gmapf_handle, one per map,
 * each already initialised, filled and populated with agents) and owns their
 * frame loop:
 *
 *   submit(map, agent, target) x many    -- any time; requests queue on the host
 *   async_step(dt)                        -- admission, then AsyncStep on every map
 *   wait_step_end()                       -- WaitStepEnd on every map, retirement
 *   poll()                                -- completed requests
 *
 * Admission (at async_step): a request is admitted when
 *   - its agent has no request of the server in flight (requests for one
 *     agent run one after another), and
 *   - its map has a free search slot (PathFinderParallelAgents), and
 *   - the live search memory of all maps plus its footprint stays within
 *     queue_memory_budget.
 * A request held back by the first two conditions is skipped over.  The
 * order of the scan is the admission policy:
 *   GMAPF_ADMIT_FIFO (default): arrival order over all maps; a request held
 *     back by the memory budget stops the scan, so a large-map request is
 *     never overtaken indefinitely by small ones.  A burst on one map delays
 *     every later request of the other maps.
 *   GMAPF_ADMIT_FAIR: per-map arrival order, maps served by deficit round
 *     robin over bytes: at each admission every map with queued requests
 *     gains one quantum of credit (the largest footprint of all maps, so any
 *     map can afford one request per frame) and the maps take turns, one
 *     request per turn, from a start that rotates every frame, while their
 *     credit covers their next request's footprint.  A map whose credit covers its next
 *     request but the memory budget does not stops the scan (the budget
 *     drains towards it), so no map starves.  Budget left after the fair
 *     passes goes to the remaining requests in arrival order (charged to
 *     their maps' credit), so admission never idles while memory is free.
 *     Each map receives an equal share of the budget under contention,
 *     whatever the others' arrival rates.  Credit is kept within
 *     [-4, +2] quanta and reset when a map's queue empties.
 * Footprint of a request on a map with
 * V nodes, T lanes and queue capacity Q:
 *   T * Q * 8 (parked queue keys) + T * 4 (queue counts) + VP (visited
 *   marks, one byte per node, V rounded up to 64) + V * 8 (G, direction and
 *   parent per node),
 * i.e. the device memory its search slot keeps busy while it runs.  An
 * admitted request becomes the agent's SetAgentTargPos of the next frame
 * (reference include/gmapf/GMAPF.h:44); it is complete once the agent leaves
 * the Search state: Move = path found, Idle = no path (or start node equal to
 * the target node, which needs no search).  All maps are launched before any
 * is waited for, so their frames overlap on the device.
 *
 * While a server holds an engine, call only read-only functions on it
 * (positions, states, corridors); the server calls set_agent_targ_pos,
 * async_step and wait_step_end.  Agents already in the Search state when the
 * server first steps count as in flight (without a ticket) until they finish.
 * Not thread-safe (one host thread per server).
 */
#ifndef GMAPF_SERVE_H
#define GMAPF_SERVE_H

#include "gmapf_c.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GMAPF_E_FULL -7 /* pending-request queue is at its cap */

typedef struct gmapf_server_s* gmapf_server;

typedef struct gmapf_server_config {
    long long queue_memory_budget; /* bytes of live search memory, all maps; <= 0: unlimited */
    int max_pending;               /* cap of queued requests; <= 0: 4,194,304 */
    int record_start_nodes;        /* nonzero: fill gmapf_request_done.start_node (reads all
                                      agents' nodes of a map from the device in every frame
                                      that admits on it) */
    int admission_policy;          /* GMAPF_ADMIT_FIFO or GMAPF_ADMIT_FAIR */
} gmapf_server_config;

#define GMAPF_ADMIT_FIFO 0
#define GMAPF_ADMIT_FAIR 1

typedef struct gmapf_request_done {
    long long ticket;      /* returned by submit, increasing from 0 */
    int map, agent;
    int found;             /* 1: path found (agent moving); 0: no path / start node == target node */
    int start_node;        /* agent's node when admitted (-1 unless record_start_nodes) */
    long long submit_frame, admit_frame, done_frame; /* server frame counters */
    double latency_ms;     /* host clock, submit -> retirement at wait_step_end */
} gmapf_request_done;

int gmapf_server_create(gmapf_server* out, const gmapf_handle* maps, int n_maps,
                        const gmapf_server_config* cfg);
int gmapf_server_destroy(gmapf_server s);
const char* gmapf_server_last_error(gmapf_server s);

/* Queue one request; the target is checked on the map (GMAPF_E_POS). */
int gmapf_server_submit(gmapf_server s, int map, int agent, float x, float y, long long* ticket);
/* Queue n requests; status[k] (may be NULL) gets each one's code, tickets[k]
 * (may be NULL) its ticket or -1. */
int gmapf_server_submit_many(gmapf_server s, int n, const int* map, const int* agent,
                             const float* xy, long long* tickets, int* status);

int gmapf_server_async_step(gmapf_server s, float dt);
int gmapf_server_wait_step_end(gmapf_server s);

/* Completed requests since the last poll, oldest first; *n = number copied
 * (at most cap; the rest stay for the next poll).  At most max_pending
 * unpolled completions are kept (the oldest are dropped). */
int gmapf_server_poll(gmapf_server s, gmapf_request_done* out, int cap, int* n);

/* Requests the next async_step would admit, in admission order, without
 * admitting them (*n = count, tickets copied up to cap). */
int gmapf_server_plan(gmapf_server s, long long* tickets, int cap, int* n);

/* [0] frames [1] submitted [2] admitted [3] completed [4] found [5] no path
 * [6] pending [7] in flight [8] live search bytes [9] peak live search bytes
 * [10] rejected (queue full) [11] held back by the budget at the last admission
 * [12] number of maps [13..15] reserved */
int gmapf_server_get_stats(gmapf_server s, long long* out16);

/* Last frame, ms: [0] wall (async_step entry -> wait_step_end return)
 * [1] admission (host) [2] launch of all maps [3] wait for all maps
 * [4] retirement (host) [5..7] reserved */
int gmapf_server_last_frame_ms(gmapf_server s, double* out8);

/* Footprint in bytes of one request on map `map` (formula above). */
int gmapf_server_footprint(gmapf_server s, int map, long long* bytes);

#ifdef __cplusplus
}
#endif
#endif
