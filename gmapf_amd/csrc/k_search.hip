// k_search.hip -- bidirectional multi-lane bounded-queue search on gfx950.
//
// Semantics (reference: include/gmapf/PathFinder.h:138-307, misc/CuQueue.h:33-70,
// misc/CuVisitedMap.h:45-54, misc/CuNodesMap.h:237-243):
//   * a search owns T lanes; lane 0 is seeded with the start node (dir 1),
//     lane 1 with the target node (dir 0), both with F = dist2(S,Tg)*K;
//   * a round: every lane pops the minimum of its own queue and expands up to
//     three neighbours (stops at the first -1) with
//     h = dist2(centroid(n), centroid(opposite endpoint)), G' = G+1,
//     F' = (float)G' + h*K;
//   * claims are resolved in (lane, slot) order: the first candidate that
//     reaches an unvisited node claims it and appends it to the frontier; a
//     candidate that reaches a node owned by the other direction is a meet,
//     and the first meet in (lane, slot) order wins;
//   * frontier item i is pushed to the queue of lane i mod T; queues keep the
//     Q smallest items by (F, node);
//   * the search ends when a meet was found or every queue is empty; a found
//     search writes the corridor [start side reversed, start excluded] ++
//     [target side up to the target] into the agent's corridor buffer.
// The reference resolves claims and meets with atomics in arbitrary order;
// here the order is fixed, so results are deterministic and equal the
// oracle's serial restatement.
//
// Mapping: one workgroup = one wave64 = 64/T searches ("segments").  Each
// lane's queue is a 4-ary min-heap of 64-bit keys (F bits made monotone,
// node id) in LDS, entry i of lane L at q[i*64+L] (bank-conflict free);
// when full, a smaller key replaces the largest leaf and sifts up.  Claim
// prefix sums and the first-meet choice are wave ballots + popcounts;
// duplicate claims inside a round are found by DPP row shifts that compare
// each candidate with the segment's earlier ones (T <= 4) or by an LDS claim
// hash (T >= 8).  Visited
// state is split by use: a per-slot byte map (m << 1 | dir, m = the search's
// 7-bit generation, so maps are not cleared per search) answers the probe of
// every neighbour with one byte, 64 nodes per cache line; an 8-byte entry
// {G << 1 | dir, parent} is written for every claimed node and read by the
// corridor walk.  On maps of at most 2^15 nodes the queue keys also carry G
// and dir (compact keys), so a popped node's entry is not read; wider maps
// read it with the node's record.  A node's neighbours and their centroids
// come in one 48-byte record; the record of the node the next round will pop,
// and the marks of its neighbours, are requested while the current round's
// frontier is pushed.
//
// Forms: plain (one search per segment, run to completion; the agent of a
// found search moves at once and its corridor is walked beside the other
// searches, finished lane-parallel at the end) and persistent (segments take
// searches from an atomic ticket over the active list, or one each when the
// grid holds a segment per search, run up to `budget` rounds, park their
// queues in HBM and take the next ticket).  When every agent has a slot of
// its own (P >= N) there is no active list: the items are the slots, and
// a segment given a slot with no search in flight there has nothing to do
// (one each) or takes the next ticket.  In the persistent form a found
// search keeps its segment and walks its corridor inside the same budget,
// two parent steps per round, parking the walk in its slot when the budget
// runs out; the agent moves when the corridor is complete.
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace gmapf_amd {

namespace {

constexpr int kWave = 64;
// s_waitcnt operand: vmcnt(0), expcnt and lgkmcnt not waited for (gfx9
// layout: vmcnt[3:0] in bits 3:0 and [5:4] in bits 15:14, expcnt in 6:4,
// lgkmcnt in 11:8)
constexpr int kWaitVmem0 = 0x0f70;

// Queue key of an item: F's bits with the sign bit set (the order of
// non-negative floats as unsigned integers, above every bit pattern with the
// sign clear), then the low word.  ManualStart refuses a negative or
// non-finite HeuristicK, so F = (G + 1) + h * K >= 1 for every pushed item and
// the seeds' F = h * K >= +0; the general map of a signed float (all bits
// inverted when negative, -0 folded to +0) would give the same keys.
__device__ __forceinline__ unsigned long long make_key(float F, unsigned low) {
    return ((unsigned long long)(__float_as_uint(F) | 0x80000000u) << 32) | low;
}

__device__ __forceinline__ int key_node(unsigned long long k) { return (int)(unsigned)(k & 0xffffffffull); }

// Compact keys (maps of at most 2^15 nodes): the low word holds
// node << 17 | G << 1 | dir, so the key orders by (F, node) as before (a node
// enters a search's queues once, so G and dir never decide an order) and a
// popped key carries the G and direction the round needs: no visited-entry
// read per expansion.  G < V <= 2^15 fits its 16 bits.
constexpr int kKeyNodeShift = 17;
constexpr int kCompactMaxNodes = 1 << 15;
constexpr unsigned kKeyGdMask = (1u << kKeyNodeShift) - 1u;

template <bool KG>
__device__ __forceinline__ unsigned long long node_key(float F, int node, int G, int dir) {
    if constexpr (KG) {
        return make_key(F, ((unsigned)node << kKeyNodeShift) | ((unsigned)G << 1) | (unsigned)dir);
    } else {
        return make_key(F, (unsigned)node);
    }
}

template <bool KG>
__device__ __forceinline__ int key_node_of(unsigned long long k) {
    if constexpr (KG) {
        return (int)((unsigned)k >> kKeyNodeShift);
    } else {
        return key_node(k);
    }
}

// G << 1 | dir of a compact key (the layout of a visited entry's first word)
__device__ __forceinline__ unsigned key_gd(unsigned long long k) { return (unsigned)k & kKeyGdMask; }

__device__ __forceinline__ float dist2(float2 a, float2 b) {
    float x = a.x - b.x;
    float y = a.y - b.y;
    return x * x + y * y;
}

// ---- per-lane 4-ary min-heap in LDS, "keep the Q smallest" (CuQueue.h:33-70).
// Entry i of lane L is q[i * 64 + L]; node i's children are 4i+1 .. 4i+4
// (a 16-item heap is two levels deep).  Every entry at or beyond the lane's
// count holds kKeyNone, and so do rows Q .. Q+3 (the padding of a child group
// that runs past the heap), so a sift reads four children with no bounds
// test.  Row Q+4 is the target of the writes of a lane that has nothing to
// move.  The sifts run a fixed number of levels without branches and the wave
// leaves a sift as soon as no lane moves (a ballot per level).  All lanes of
// the wave call these functions; `on` selects the lanes that act.
constexpr unsigned long long kKeyNone = ~0ull;  // larger than every key
constexpr int kPadRows = 4;
constexpr int kSpareRows = kPadRows + 1;

// levels below the root of a Q-item 4-ary heap
__device__ __forceinline__ int heap_depth(int Q) {
    int d = 0;
    for (long long full = 1; full < Q; full = 4 * full + 1) ++d;
    return d;
}

// Sift key k up from hole i.
__device__ __forceinline__ void heap_sift_up(unsigned long long* q, int L, int Q, int depth, bool on, int i,
                                             unsigned long long k) {
    bool done = !on;
    for (int d = 0; d < depth; ++d) {
        if (__ballot(!done) == 0ull) break;
        const bool hp = i > 0;
        const int p = (i - 1) >> 2;
        const unsigned long long pv = q[(hp ? p : Q) * kWave + L];
        const bool move = !done && hp && pv > k;
        q[(move ? i : Q + kPadRows) * kWave + L] = pv;
        i = move ? p : i;
        done = !move;
    }
    q[(on ? i : Q + kPadRows) * kWave + L] = k;
}

// Sift key k down from a hole at the root (the entries past the heap's items
// hold kKeyNone).
__device__ __forceinline__ void heap_sift_down(unsigned long long* q, int L, int Q, int depth, bool on,
                                               unsigned long long k) {
    int i = 0;
    bool done = !on;
    for (int d = 0; d < depth; ++d) {
        if (__ballot(!done) == 0ull) break;
        const int c0 = 4 * i + 1;
        const int cb = c0 < Q ? c0 : Q;  // rows cb .. cb+3 lie within the heap and padding rows
        const unsigned long long* r = q + cb * kWave + L;
        const unsigned long long v0 = r[0], v1 = r[kWave], v2 = r[2 * kWave], v3 = r[3 * kWave];
        const bool b1 = v1 < v0, b3 = v3 < v2;
        const unsigned long long m01 = b1 ? v1 : v0, m23 = b3 ? v3 : v2;
        const bool b23 = m23 < m01;
        const unsigned long long cv = b23 ? m23 : m01;
        const int ci = c0 + (b23 ? (b3 ? 3 : 2) : (b1 ? 1 : 0));
        const bool move = !done && cv < k;
        q[(move ? i : Q + kPadRows) * kWave + L] = cv;
        i = move ? ci : i;
        done = !move;
    }
    q[(on ? i : Q + kPadRows) * kWave + L] = k;
}

// Push key k (lanes with k != kKeyNone) into a heap bounded to `cap` items:
// below the bound it is added; at the bound it replaces the largest item (a
// leaf: indices from parent(cap - 1) + 1 to cap - 1) when smaller, else it is
// dropped.
__device__ __forceinline__ void heap_push(unsigned long long* q, int L, int Q, int depth, int cap, int& cnt,
                                          unsigned long long k) {
    const bool on = k != kKeyNone;
    const bool full = cnt >= cap;
    bool go = on;
    int start = cnt;
    if (__ballot(on && full) != 0ull) {
        int mi = 0;
        unsigned long long mv = 0ull;  // with no leaf (cap 0) every key is dropped
        const int first_leaf = cap > 1 ? ((cap - 2) >> 2) + 1 : 0;
        for (int j = first_leaf; j < cap; ++j) {
            const unsigned long long v = q[j * kWave + L];
            mi = v > mv ? j : mi;
            mv = v > mv ? v : mv;
        }
        if (full) {
            go = on && k < mv;
            start = mi;
        }
    }
    heap_sift_up(q, L, Q, depth, go, start, k);
    cnt += (on && !full) ? 1 : 0;
}

// Ordering point between the lanes of the workgroup.  A workgroup is one
// wave (launch_search launches kWave threads), and the memory operations of
// one wavefront are performed in order with respect to each other (LDS and
// the CU's vector L1 alike), so a wavefront-scope fence plus a wave barrier
// (both compiler-level) orders the lanes' LDS and global accesses without
// the wait for outstanding loads and stores that a workgroup barrier
// implies.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int T>
__device__ __forceinline__ unsigned long long seg_mask(int sub) {
    if constexpr (T == 64) {
        return ~0ull;
    } else {
        return ((1ull << T) - 1ull) << (sub * T);
    }
}

// per-segment claim table: open addressing, 8T slots for the at most 3T
// candidates of a round (load factor <= 3/8; 4 KB per workgroup for any T)
template <int T>
constexpr int hash_bits() {
    int b = 1;
    while ((1 << b) < 8 * T) ++b;
    return b;
}
constexpr int kHashEmpty = -1;
constexpr int kHashNone = 0x7fffffff;

__device__ __forceinline__ unsigned claim_hash(int n, int bits) {
    return ((unsigned)n * 2654435761u) >> (32 - bits);
}

// First candidate of each of a lane's three neighbour slots among the
// candidates of its segment that come earlier in (lane, slot) order.  A
// candidate that is present (`ins`: the node exists and this search has not
// marked it) is keyed k = (node << SB | r) + 1, r the segment's index inside
// its 16-lane DPP row (SB bits), and packed as k << 1 | direction; an absent
// one has key -1 and packed value 0.  first[s] is the packed value of the
// earliest earlier candidate for the same node, or -1 when slot s's candidate
// is the node's first.  Lanes l - 1 .. l - (T - 1) are read with DPP row
// shifts (a segment of T <= 16 lanes lies inside one 16-lane row), nearest
// lane first so that the earliest match is the one kept.  A shift that
// reaches the previous segment brings keys with another r, and one that
// leaves the row brings 0 (bound_ctrl), so neither can match: no mask of
// the lanes below n is needed.
template <int T>
constexpr int seg_row_bits() {
    int b = 0;
    while ((T << b) < 16) ++b;
    return b;
}

template <int N>
__device__ __forceinline__ int row_shr(int v) {
    // row_shr:N; a lane with no source lane in its row gets 0 (absent)
    return __builtin_amdgcn_mov_dpp(v, 0x110 + N, 0xf, 0xf, true);
}

template <int T, int N>
__device__ __forceinline__ void seg_first_from(const int* key, const int* pk, int* first) {
    if constexpr (N < T) {
        int wp[3], ws[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            wp[s] = row_shr<N>(pk[s]);
            ws[s] = wp[s] >> 1;  // 0 when absent: never equal to a key
        }
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            first[s] = key[s] == ws[2] ? wp[2] : first[s];
            first[s] = key[s] == ws[1] ? wp[1] : first[s];
            first[s] = key[s] == ws[0] ? wp[0] : first[s];
        }
        seg_first_from<T, N + 1>(key, pk, first);
    }
}

template <int T>
__device__ __forceinline__ void seg_first_candidates(int sub, const int* nn, const bool* ins, int cdir, int* first) {
    constexpr int SB = seg_row_bits<T>();
    const int r = sub & ((1 << SB) - 1);
    int pk[3], key[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        key[s] = ins[s] ? ((nn[s] << SB) | r) + 1 : -1;
        pk[s] = ins[s] ? ((key[s] << 1) | cdir) : 0;
    }
    // earlier slots of the lane itself
    first[0] = -1;
    first[1] = key[1] == (pk[0] >> 1) ? pk[0] : -1;
    first[2] = key[2] == (pk[1] >> 1) ? pk[1] : -1;
    first[2] = key[2] == (pk[0] >> 1) ? pk[0] : first[2];
    seg_first_from<T, 1>(key, pk, first);
}

// Exclusive prefix `pre` of the claims `cl` over the lanes of the segment
// (lane order) and the segment's total `tot`.  T = 2 and 4: the segment is a
// DPP pair or quad; an inclusive scan of shifted adds, each shift fed only by
// lanes whose target lies in the same segment (the others contribute 0),
// then the segment's last lane broadcast with a quad permute.  Wider
// segments: two ballots of the claim count's bits and popcounts.
template <int T>
__device__ __forceinline__ void seg_scan(int l, int cl, unsigned long long below, unsigned long long smask, int& pre,
                                         int& tot) {
    if constexpr (T == 2 || T == 4) {
        int inc = cl + __builtin_amdgcn_mov_dpp(l == T - 1 ? 0 : cl, 0x111, 0xf, 0xf, true);  // row_shr:1
        if constexpr (T == 4) inc += __builtin_amdgcn_mov_dpp(l >= 2 ? 0 : inc, 0x112, 0xf, 0xf, true);  // row_shr:2
        pre = inc - cl;
        // quad_perm [1,1,3,3] (pairs) or [3,3,3,3] (quads): the segment's last lane
        tot = __builtin_amdgcn_mov_dpp(inc, T == 4 ? 0xff : 0xf5, 0xf, 0xf, false);
    } else {
        const unsigned long long b0 = __ballot(cl & 1);
        const unsigned long long b1 = __ballot(cl & 2);
        pre = __popcll(b0 & below) + 2 * __popcll(b1 & below);
        tot = __popcll(b0 & smask) + 2 * __popcll(b1 & smask);
    }
}

// Corridor walk of one side of a found search (PathFinder.h:276-306), one
// parent step per call so the walk of a search that ended overlaps the rounds
// of the other searches of the workgroup.  The corridor occupies
// slab[base .. base + G(mS) + G(mT) + 1).  Lane 0 of the segment walks the
// start side from mS down to the start node (exclusive), writing indices
// G(mS)-1 .. 0; lane 1 walks the target side from mT to the target node
// (inclusive), writing indices G(mS) .. G(mS)+G(mT).  `ent` holds the visited
// entry of `n`, requested by the previous step.
struct Walk {
    bool active = false;
    int n = 0;
    int idx = 0;  // slab index the next node goes to
    int term = 0;
    const uint2* vis = nullptr;  // the search slot's visited table
    uint2 ent = make_uint2(0u, 0u);
};

// The slab pointer is read from memory, so the compiler cannot tell its
// address space; stores through a generic pointer become flat stores, which
// count on the LDS counter as well and so make LDS waits wait for them.
typedef __attribute__((address_space(1))) int global_int;

// A step in two halves: `walk_advance` writes the node whose entry arrived
// and moves to its parent; `walk_request` asks for the parent's entry.
// Written with selects around the two stores, so a step is one short
// straight run with two masked stores.
__device__ __forceinline__ void walk_advance(Walk& w, int side, global_int* slab) {
    const bool mv = w.active && w.n != w.term;  // n is written, then its parent is next
    if (mv) slab[w.idx] = w.n;
    const int n = mv ? (int)w.ent.y : w.n;
    const int idx = mv ? w.idx + (side == 0 ? -1 : 1) : w.idx;
    const bool fin = w.active && n == w.term;  // the target side writes its terminal node
    if (fin && side == 1) slab[idx] = w.term;
    w.n = n;
    w.idx = idx;
    w.active = w.active && !fin;
}

__device__ __forceinline__ void walk_request(Walk& w) {
    if (w.active) w.ent = w.vis[w.n];
}

__device__ __forceinline__ void walk_step(Walk& w, int side, global_int* slab) {
    walk_advance(w, side, slab);
    walk_request(w);
}

// The corridor of a found search (meet mS, mT): lane 0 of the segment takes
// G(mS) + G(mT) + 1 ints of the slab (PathFinder.h:280-290); when they fit,
// lanes 0 and 1 start their walks, else the search keeps its slot with status
// 4 and tries again next frame.  Plain form (`hold` false): the agent moves
// now (state Move, path reset) and the walk runs beside later rounds.  Batched
// form (`hold`): the segment keeps the slot and walks the corridor within its
// round budget; the agent moves when the walk is done.  `share` is the
// segment's three LDS words (base or -1, G(mS), length).  Called by every lane
// of the segment; returns true when the corridor was placed, with its offset
// and length in cbase, clen.
__device__ __forceinline__ bool start_corridor(int l, int slot, int agent, int S, int Tg, int mS, int mT,
                                               const uint2* vis, global_int* slab, const DevAgents& ag,
                                               const DevSlots& slots, size_t V, int* share, Walk& walk, bool hold,
                                               int& cbase, int& clen, unsigned& st_found, unsigned& st_corr) {
    if (l < 2) {
        // a walk still pending on these lanes (their previous search in a
        // persistent segment) finishes first
        while (walk.active) walk_step(walk, l, slab);
    }
    if (l == 0) {
        const int gS = (int)(vis[mS].x >> 1), gT = (int)(vis[mT].x >> 1);
        const int len = gS + gT + 1;
        const unsigned long long o = atomicAdd(ag.slab_top, (unsigned long long)len);
        const bool ok = o + (unsigned long long)len <= (unsigned long long)ag.slab_cap;
        // A request that does not fit takes its add back when no allocation
        // came after it (the top still reads o + len), so a later, shorter
        // request can use the space left.  The compare-and-swap succeeds only
        // when the add was the last one, so no range is handed out twice;
        // otherwise the top stays past the end until the next compaction.
        // (A compare-and-swap loop for every allocation would be exact, but
        // under contention it completes one allocation per L2 round trip:
        // 40x slower at 16,384 searches.)
        if (!ok) atomicCAS(ag.slab_top, o + (unsigned long long)len, o);
        share[0] = ok ? (int)o : -1;
        share[1] = gS;
        share[2] = len;
        if (ok && hold) {
            slots.corr[slot] = make_int2((int)o, len);
        } else if (ok) {
            ag.path_off[agent] = (long long)o;
            ag.path_len[agent] = len;
            ag.state[agent] = kMove;
            ag.step_idx[agent] = -1;
            ag.inflight[agent] = -1;
            slots.agent[slot] = -1;
            slots.status[slot] = 1;
            ++st_found;
            st_corr += (unsigned)len;
        } else {
            slots.mS[slot] = mS;
            slots.mT[slot] = mT;
            slots.status[slot] = 4;
            atomicAdd(ag.slab_deferred, 1ull);
        }
    }
    wave_sync();
    const int base = share[0];
    cbase = base;
    clen = share[2];
    if (base >= 0 && l < 2) {
        const int gS = share[1];
        walk.active = true;
        walk.vis = vis;
        walk.n = l == 0 ? mS : mT;
        walk.term = l == 0 ? S : Tg;
        walk.idx = base + (l == 0 ? gS - 1 : gS);
        walk.ent = vis[walk.n];
    }
    return base >= 0;
}

// Batched form: lanes 0 and 1 store their side of a walk that stops at the
// end of the launch (node -1: side done; position relative to the corridor's
// offset, which compaction may move) and drop it.
__device__ __forceinline__ void park_walk(int l, int slot, int cbase, const DevSlots& slots, Walk& walk) {
    if (l < 2) {
        reinterpret_cast<int2*>(&slots.walk[slot])[l] = make_int2(walk.active ? walk.n : -1, walk.idx - cbase);
        walk.active = false;
    }
}

// Per-lane flags carried from one round to the next, kept as the bits of one
// word (a vector register).  As separate booleans each would live across the
// whole round as a 64-bit lane mask in scalar registers, which the round's
// compares and selects need.
enum : unsigned {
    kHasReq = 1u,       // the segment holds a search or a corridor walk
    kNoMore = 2u,       // the segment has no more work to take
    kTookPlain = 4u,    // the segment took its one item (one search per segment)
    kPoppedNext = 8u,   // the lane has a node to expand in the coming round (its record
                        // and the marks of its neighbours are requested)
    kWalking = 16u,     // batched form: the segment walks a found search's corridor
};
struct LaneFlags {
    unsigned v = 0u;
    __device__ __forceinline__ bool get(unsigned b) const { return (v & b) != 0u; }
    __device__ __forceinline__ void set(unsigned b, bool on) { v = on ? (v | b) : (v & ~b); }
};

// Resident waves per SIMD the launch needs to run one search per segment at
// 16,384 searches: T/4 (4 at T = 16: at most 128 VGPRs; 2 at T <= 8: at most
// 256).
template <int T>
constexpr int search_waves_per_eu() {
    return T >= 16 ? 4 : 2;
}

// QC > 0 is the queue capacity fixed at compile time (the heap loops unroll
// with constant bounds); QC = 0 takes it from the launch.  KG: compact keys
// (the map has at most 2^15 nodes; only with QC > 0).
template <int T, int QC, bool KG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(search_waves_per_eu<T>()))) void k_search(DevMap map, DevAgents ag, DevSlots slots, DevFrame fr,
                                               SearchParams sp, int persistent) {
    extern __shared__ unsigned long long q[];  // 64 * (Q + 4) keys: Q heap rows + the spare rows
    constexpr int RPW = kWave / T;
    constexpr int HB = hash_bits<T>();
    constexpr int HS = 1 << HB;
    __shared__ unsigned long long stage[kWave * 3];
    __shared__ int hkey[RPW * HS];  // node of the slot, kHashEmpty when free
    __shared__ int hval[RPW * HS];  // min over the node's candidates of (k << 1 | dir)
    // per-lane target of the table operations of an empty neighbour slot, so
    // the three slots' operations issue together without branches; it never
    // holds kHashEmpty, so a compare-and-swap on it never writes
    __shared__ int hdummy[kWave];
    __shared__ int hsink[kWave];  // per-lane target of the table resets of an empty slot
    __shared__ int wshare[RPW * 3];  // per segment: corridor base (or -1), G(mS), length
    for (int i = threadIdx.x; i < RPW * HS; i += kWave) {
        hkey[i] = kHashEmpty;
        hval[i] = kHashNone;
    }
    hdummy[threadIdx.x] = kHashNone;

    const int L = threadIdx.x;
    const int sub = L / T;
    const int l = L % T;
    const int segbase = sub * 3 * T;
    const unsigned long long smask = seg_mask<T>(sub);
    const unsigned long long below = smask & ((1ull << L) - 1ull);
    const int Q = QC > 0 ? QC : sp.Q;
    const int depth = heap_depth(Q);
    for (int r = 0; r < kPadRows; ++r) q[(Q + r) * kWave + threadIdx.x] = kKeyNone;
    const size_t V = (size_t)map.V;
    const int n_active = sp.ident > 0 ? sp.ident : fr.counters[1];
    // one segment per search when the grid holds that many, else tickets
    // (also when every agent has a slot of its own: with more agents than
    // the grid holds segments, the slots past the grid are reached by ticket)
    const bool one_each = (long long)n_active <= (long long)gridDim.x * sp.seg;
    const long long stride = fr.counters[9];
    // corridor slab in use (compaction runs between launches)
    global_int* const slab = (global_int*)*ag.slab_now;

    // per-segment search state (replicated in every lane of the segment)
    int slot = -1, agent = -1;
    LaneFlags fl;
    fl.set(kNoMore, sub >= sp.seg);
    // the node this lane expands in the coming round: taken from the heap
    // and the items dealt to the lane at the end of the previous round, or
    // popped at the top of the first round of a fresh search
    int c_next = -1;
    unsigned c_gd = 0u;  // compact keys: G << 1 | dir of c_next
    int S = 0, Tg = 0, cnt = 0;
    unsigned mval = 0;  // this search's mark generation, 1..127
    float2 cS = make_float2(0.f, 0.f), cT = make_float2(0.f, 0.f);
    // rounds_here counts the segment's rounds in this launch (search and
    // corridor walk, for the budget)
    int rounds_here = 0;
    // batched form: the segment walks its search's corridor (cbase, clen)
    int cbase = 0, clen = 0;
    int status = 0;
    uint2* vis = slots.vis;
    unsigned char* mark = slots.mark;
    // next round's pop, fetched while this round's frontier is pushed
    int pf_c = -1;
    int4 pf_r0 = make_int4(-1, -1, -1, -1), pf_r1 = make_int4(0, 0, 0, 0);
    int2 pf_r2 = make_int2(0, 0);
    uint2 pf_vc = make_uint2(0u, 0u);
    unsigned pf_vn[3] = {0u, 0u, 0u};
    // corridor walk of this lane's side of the segment's last found search
    Walk walk;
    // frame statistics, added to a stripe of fr.stats_part once per
    // workgroup at the end
    unsigned st_found = 0, st_fail = 0, st_exp = 0, st_disc = 0, st_corr = 0, st_rounds = 0;

    // Acquisition and the test for the end of the workgroup run only when a
    // segment may need work: in the first round, after a segment's search or
    // walk ended or parked, and while a ticket gave a segment nothing.
    bool acquire = true;
    while (true) {
        if (acquire) {
            // ---- acquire work for segments that have none
            if (__ballot(!fl.get(kHasReq) && !fl.get(kNoMore)) != 0ull) {
                if (!fl.get(kHasReq) && !fl.get(kNoMore)) {
                    int item = -1;
                    if (persistent && !one_each) {
                        if (l == 0) item = atomicAdd(&fr.counters[2], 1);
                        item = __shfl(item, sub * T, kWave);
                    } else if (!fl.get(kTookPlain)) {
                        // every search has a segment of its own: segment i takes
                        // item i, once (the plain form always does)
                        item = blockIdx.x * sp.seg + sub;
                        fl.set(kTookPlain, true);
                    }
                    // The item's slot.  Its fields, and the count and front of
                    // a queue parked there, are requested together; then the
                    // centroids and the first node's record; then (after the
                    // seeds) its neighbours' marks: three waits for memory.
                    int slot_i = -1;
                    if (item >= 0 && item < n_active)
                        slot_i = sp.ident > 0 ? item : fr.active[(int)(((long long)item * stride) % n_active)];
                    const size_t qb = (size_t)(slot_i >= 0 ? slot_i : 0) * T + l;
                    int agent_i = -1, S_i = 0, Tg_i = 0, status_i = 0, qcnt_i = 0;
                    unsigned stamp = 0u;
                    unsigned long long k0_i = kKeyNone;
                    if (slot_i >= 0) {
                        agent_i = slots.agent[slot_i];
                        S_i = slots.S[slot_i];
                        Tg_i = slots.Tg[slot_i];
                        stamp = slots.stamp[slot_i];
                        status_i = slots.status[slot_i];
                        qcnt_i = slots.qcnt[qb];
                        k0_i = slots.qsave[qb * (size_t)Q];
                    }
                    // slot = agent: nothing to do when no search is in flight
                    // there; with tickets the segment takes the next one
                    const bool idle_slot = sp.ident > 0 && slot_i >= 0 && agent_i < 0;
                    if (idle_slot) slot_i = -1;
                    if (slot_i < 0) {
                        fl.set(kNoMore, !idle_slot || !persistent || one_each);
                    } else {
                        slot = slot_i;
                        agent = agent_i;
                        S = S_i;
                        Tg = Tg_i;
                        mval = stamp % 127u + 1u;
                        status = status_i;
                        cS = map.cen[S];
                        cT = map.cen[Tg];
                        rounds_here = 0;
                        fl.set(kWalking, false);
                        vis = slots.vis + (size_t)slot * V;
                        mark = slots.mark + (size_t)slot * slots.VP;
                        fl.set(kHasReq, true);
                        const bool seeded = status == 3;
                        // The node the first round expands: a new search's
                        // seeds (start on lane 0 with G 0 and side 1, target on
                        // lane 1 with side 0, each its heap's only item and so
                        // popped at once: PathFinder.h:168-189) or a parked
                        // queue's front.  Its record is requested now.
                        fl.set(kPoppedNext, seeded ? l < 2 : status == 0 && k0_i != kKeyNone);
                        c_next = seeded ? (l == 0 ? S : Tg) : key_node_of<KG>(k0_i);
                        c_gd = seeded ? (l == 0 ? 1u : 0u) : key_gd(k0_i);
                        pf_c = fl.get(kPoppedNext) ? c_next : -1;
                        if (fl.get(kPoppedNext)) {
                            pf_r0 = map.rec[3 * (size_t)c_next];
                            pf_r1 = map.rec[3 * (size_t)c_next + 1];
                            pf_r2 = *reinterpret_cast<const int2*>(map.rec + 3 * (size_t)c_next + 2);
                            if constexpr (KG) {
                                pf_vc.x = c_gd;
                            } else {
                                pf_vc = seeded ? make_uint2(c_gd, 0xffffffffu) : vis[c_next];
                            }
                        }
                        if (status == 4) {
                            // found earlier, corridor deferred for lack of slab space
                            const bool hold = sp.budget > 0;
                            const bool ok =
                                start_corridor(l, slot, agent, S, Tg, slots.mS[slot], slots.mT[slot], vis, slab, ag,
                                               slots, V, &wshare[3 * sub], walk, hold, cbase, clen, st_found, st_corr);
                            fl.set(kWalking, hold && ok);
                            fl.set(kHasReq, fl.get(kWalking));
                            status = 0;
                        } else if (status == 5) {
                            // batched form: a corridor walk parked at the end of the
                            // last launch resumes
                            const int4 pw = slots.walk[slot];
                            const int2 cr = slots.corr[slot];
                            cbase = cr.x;
                            clen = cr.y;
                            if (l < 2) {
                                const int n = l == 0 ? pw.x : pw.z;
                                walk.active = n >= 0;
                                walk.vis = vis;
                                walk.n = n;
                                walk.term = l == 0 ? S : Tg;
                                walk.idx = cbase + (l == 0 ? pw.y : pw.w);
                                if (walk.active) walk.ent = vis[n];
                            }
                            fl.set(kWalking, true);
                            status = 0;
                        } else if (status == 3) {
                            // Stamps 127e .. 127e + 126 (epoch e) use generations
                            // 1 .. 127 once each.  The first search of an epoch to
                            // seed in this slot zeroes the slot's marks before its
                            // seeds, so only marks of the current epoch remain; a
                            // search admitted and dropped before it seeded (its
                            // agent re-targeted) cannot skip the clear.
                            const unsigned epoch = stamp / 127u;
                            if (slots.clr[slot] != epoch) {
                                uint4* m4 = reinterpret_cast<uint4*>(mark);
                                for (long long w = l; w < slots.VP / 16; w += T) m4[w] = make_uint4(0u, 0u, 0u, 0u);
                                wave_sync();
                                if (l == 0) slots.clr[slot] = epoch;
                            }
                            // the seeds (their queue keys are not needed: each is
                            // popped at once, above, and its heap stays empty)
                            cnt = 0;
                            for (int i = 0; i < Q; ++i) q[i * kWave + L] = kKeyNone;
                            if (l == 0) {
                                vis[S] = make_uint2(1u, 0xffffffffu);  // G 0, dir 1
                                mark[S] = (unsigned char)((mval << 1) | 1u);
                            }
                            if (l == 1) {
                                vis[Tg] = make_uint2(0u, 0xffffffffu);  // G 0, dir 0
                                mark[Tg] = (unsigned char)(mval << 1);
                            }
                            status = 0;
                        } else {
                            // parked: [node to expand next or kKeyNone, heap items]
                            // (at most Q - 1 heap items; the front and the count
                            // were read above).  Loads go in groups of 8 with
                            // clamped in-bounds indices and the count applied
                            // afterwards, so a group waits for memory once
                            cnt = qcnt_i;
                            const unsigned long long* src = slots.qsave + qb * (size_t)Q;
                            for (int i0 = 0; i0 < Q - 1; i0 += 8) {
                                unsigned long long x[8];
#pragma unroll
                                for (int j = 0; j < 8; ++j) x[j] = src[1 + min(i0 + j, Q - 2)];
#pragma unroll
                                for (int j = 0; j < 8; ++j)
                                    if (i0 + j < Q - 1) q[(i0 + j) * kWave + L] = i0 + j < cnt ? x[j] : kKeyNone;
                            }
                            q[(Q - 1) * kWave + L] = kKeyNone;
                        }
                        // the marks of the first node's neighbours, so that every
                        // round finds them prefetched (the barrier orders the
                        // seeds' marks before the loads of the segment's other
                        // lanes)
                        wave_sync();
                        if (fl.get(kPoppedNext)) {
                            const int a0 = pf_r0.x, a1 = a0 >= 0 ? pf_r0.y : -1, a2 = a1 >= 0 ? pf_r0.z : -1;
                            pf_vn[0] = mark[a0 >= 0 ? a0 : 0];
                            pf_vn[1] = mark[a1 >= 0 ? a1 : 0];
                            pf_vn[2] = mark[a2 >= 0 ? a2 : 0];
                        }
                    }
                }
            }
            // A segment whose ticket gave it nothing to run (a deferred corridor
            // that still does not fit) passes this round idle and takes the next
            // ticket, so the workgroup leaves only when no segment has work and
            // every segment's tickets are spent.  (A loop around the acquisition
            // would do the same without the idle round, but its code made every
            // round 9% slower.)
            if (__ballot(fl.get(kHasReq) || !fl.get(kNoMore)) == 0ull) break;
            acquire = __ballot(!fl.get(kHasReq) && !fl.get(kNoMore)) != 0ull;
            wave_sync();  // seeds, restored queues and visited writes visible
        }

        // ---- the node this lane expands: its record, its G and side and the
        // marks of its neighbours were requested at the end of the last round
        // (or when the search was taken)
        const bool popped = fl.get(kHasReq) && fl.get(kPoppedNext);
        const int c = popped ? c_next : -1;
        const int4 r0 = pf_r0, r1 = pf_r1;
        const int2 r2 = pf_r2;
        const uint2 vc = pf_vc;
        const int G = (int)(vc.x >> 1);
        const int cdir = (int)(vc.x & 1u);
        int nn[3];
        nn[0] = popped ? r0.x : -1;
        nn[1] = nn[0] >= 0 ? r0.y : -1;
        nn[2] = nn[1] >= 0 ? r0.z : -1;
        const float2 cnb[3] = {make_float2(__int_as_float(r1.x), __int_as_float(r1.y)),
                               make_float2(__int_as_float(r1.z), __int_as_float(r1.w)),
                               make_float2(__int_as_float(r2.x), __int_as_float(r2.y))};
        const unsigned mk[3] = {pf_vn[0], pf_vn[1], pf_vn[2]};  // marks of the neighbours
        // F of every slot (a missing neighbour's value is not used); h is
        // measured to the opposite endpoint of the popped node's direction
        const float2 cfar = cdir ? cT : cS;
        float Fn[3];
        bool prior[3];
        int pdir[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const float h = dist2(cnb[s], cfar);
            Fn[s] = (float)(G + 1) + h * sp.K;
            prior[s] = nn[s] >= 0 && (mk[s] >> 1) == mval;
            pdir[s] = (int)(mk[s] & 1u);
        }
        // corridor walks take two parent steps per round, each consuming an
        // entry whose wait the round already pays: the vector memory counter
        // is in order, so a wait for a load also covers every older one.
        // Step 1 (here) consumes the entry requested at the end of the last
        // round, after the marks just used; it requests the next entry,
        // which is older than this round's record loads and so has arrived
        // by the wait for them, where step 2 consumes it.  The marks are
        // waited for explicitly before the request (vmcnt(0); the compiler
        // would otherwise place their wait after the walk's code, where it
        // would also wait for the load just requested).
        __builtin_amdgcn_s_waitcnt(kWaitVmem0);
        if (__ballot(walk.active) != 0ull && l < 2) {
            walk_advance(walk, l, slab);
            walk_request(walk);
        }

        // ---- claims and meets.  A candidate claims an unmarked node when it
        // is the node's first candidate in (lane, slot) order; otherwise the
        // node's owner is the earlier marker (prior) or the first candidate,
        // and a different direction is a meet.
        bool ins[3], claim[3], meet[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) ins[s] = nn[s] >= 0 && !prior[s];
        [[maybe_unused]] unsigned hslot[3] = {0u, 0u, 0u};
        if constexpr (T <= 4) {
            // the segment's lanes compare their candidates directly: a
            // segment of at most 16 lanes lies in one 16-lane DPP row, so a
            // row shift by n brings lane l - n's candidates to lane l (at
            // T = 8 the 63 compares per lane cost more than the table)
            int first[3];
            seg_first_candidates<T>(sub, nn, ins, cdir, first);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                claim[s] = ins[s] && first[s] < 0;
                const int odir = ins[s] ? (first[s] >= 0 ? (first[s] & 1) : cdir) : pdir[s];
                meet[s] = nn[s] >= 0 && !claim[s] && odir != cdir;
            }
        } else {
            // place each node in the segment's table (linear probing, first
            // compare-and-swap wins the slot), then keep the smallest (lane,
            // slot) index per node
            int* hk = &hkey[sub * HS];
            int* hv = &hval[sub * HS];
            bool pend[3];
            int* kp[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                hslot[s] = claim_hash(nn[s] >= 0 ? nn[s] : 0, HB);
                kp[s] = ins[s] ? &hk[hslot[s]] : &hdummy[L];
            }
            {
                int old[3];
#pragma unroll
                for (int s = 0; s < 3; ++s) old[s] = atomicCAS(kp[s], kHashEmpty, nn[s]);
#pragma unroll
                for (int s = 0; s < 3; ++s) pend[s] = ins[s] && old[s] != kHashEmpty && old[s] != nn[s];
            }
            // slot taken by another node: probe on (rare)
            while (__ballot(pend[0] || pend[1] || pend[2]) != 0ull) {
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    if (pend[s]) {
                        hslot[s] = (hslot[s] + 1u) & (unsigned)(HS - 1);
                        const int old = atomicCAS(&hk[hslot[s]], kHashEmpty, nn[s]);
                        pend[s] = old != kHashEmpty && old != nn[s];
                    }
                }
            }
            int* vp[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                vp[s] = ins[s] ? &hv[hslot[s]] : &hdummy[L];
                atomicMin(vp[s], ((l * 3 + s) << 1) | cdir);
            }
            wave_sync();
            int ev[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) ev[s] = *vp[s];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                claim[s] = ins[s] && (ev[s] >> 1) == l * 3 + s;
                const int odir = ins[s] ? (ev[s] & 1) : pdir[s];
                meet[s] = nn[s] >= 0 && !claim[s] && odir != cdir;
            }
        }

        // ---- claims: frontier order = (lane, slot) order
        const int cl = (int)claim[0] + (int)claim[1] + (int)claim[2];
        int pre, tot;
        seg_scan<T>(l, cl, below, smask, pre, tot);
        {
            int f = pre;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                if (claim[s]) {
                    stage[segbase + f] = node_key<KG>(Fn[s], nn[s], G + 1, cdir);
                    mark[nn[s]] = (unsigned char)((mval << 1) | (unsigned)cdir);
                    vis[nn[s]] = make_uint2(((unsigned)(G + 1) << 1) | (unsigned)cdir, (unsigned)c);
                    ++f;
                }
            }
        }
        // ---- first meet in (lane, slot) order; the lanes exchange it only in
        // rounds where some search of the workgroup meets
        const int ms = meet[0] ? 0 : (meet[1] ? 1 : (meet[2] ? 2 : 3));
        const unsigned long long mball = __ballot(ms < 3);
        const unsigned long long mb = mball & smask;
        const bool found = fl.get(kHasReq) && !fl.get(kWalking) && (mb != 0ull);
        int mS = -1, mT = -1;
        if (mball != 0ull) {
            const int w = found ? __ffsll((long long)mb) - 1 : sub * T;
            const int n = ms < 3 ? nn[ms] : -1;
            const int a = cdir ? c : n, b = cdir ? n : c;
            mS = __shfl(a, w, kWave);
            mT = __shfl(b, w, kWave);
        }
        wave_sync();
        if constexpr (T > 4) {
            // empty this round's table slots for the next round
            int* hk = &hkey[sub * HS];
            int* hv = &hval[sub * HS];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                *(ins[s] ? &hk[hslot[s]] : &hsink[L]) = kHashEmpty;
                *(ins[s] ? &hv[hslot[s]] : &hsink[L]) = kHashNone;
            }
        }

        // ---- deal frontier item f to lane f mod T (at most three per lane),
        // push, then pop: the round pushes its items into a heap bounded to Q
        // and the next round pops the minimum.  That minimum m is min(heap
        // top, dealt items), known before any push, so next round's node data
        // is requested first and arrives while the heap is updated.  After
        // the pop the heap holds ranks 2 .. min(n + k, Q) of the heap's n and
        // the k dealt items; the same set results from keeping m out of the
        // heap and pushing the other items with bound Q - 1 (n <= Q - 1 here,
        // since every round ends with a pop).  When m is the top, the first
        // dealt item takes its place (one sift down instead of a push and a
        // pop).
        const bool cont = fl.get(kHasReq) && !fl.get(kWalking) && !found;
        unsigned long long d0 = stage[segbase + l], d1 = stage[segbase + l + T], d2 = stage[segbase + l + 2 * T];
        d0 = cont && l < tot ? d0 : kKeyNone;
        d1 = cont && l + T < tot ? d1 : kKeyNone;
        d2 = cont && l + 2 * T < tot ? d2 : kKeyNone;
        const unsigned long long top = cont ? q[L] : kKeyNone;  // kKeyNone when empty
        unsigned long long m = top < d0 ? top : d0;
        m = d1 < m ? d1 : m;
        m = d2 < m ? d2 : m;
        const bool has_next = m != kKeyNone;
        pf_c = has_next ? key_node_of<KG>(m) : -1;
        // issued by every lane (node 0 of slot 0 when there is no next node:
        // one shared cache line), so the loads and the wait for them below
        // lie on every path and the compiler can count the loads in flight
        const int pfa = has_next ? pf_c : 0;
        const uint2* const pvis = has_next ? vis : slots.vis;
        const unsigned char* const pmark = has_next ? mark : slots.mark;
        pf_r0 = map.rec[3 * (size_t)pfa];
        pf_r1 = map.rec[3 * (size_t)pfa + 1];
        pf_r2 = *reinterpret_cast<const int2*>(map.rec + 3 * (size_t)pfa + 2);
        if constexpr (KG) {
            pf_vc.x = key_gd(m);
        } else {
            pf_vc = pvis[pfa];
        }
        {
            const bool from_top = has_next && m == top;
            // the root's replacement: the first dealt item, or the last item
            const bool repl = from_top && d0 != kKeyNone;
            const bool shrink = from_top && !repl;
            cnt -= shrink ? 1 : 0;
            const int lrow = shrink ? cnt : Q + kPadRows;
            const unsigned long long last = q[lrow * kWave + L];
            q[lrow * kWave + L] = kKeyNone;  // the vacated entry
            heap_sift_down(q, L, Q, depth, repl || (shrink && cnt > 0), repl ? d0 : last);
            // corridor walk step 2: its entry (requested in step 1, before
            // the record loads) has arrived by now; the next entry is
            // requested before the mark loads, so the wait for the marks at
            // the top of the next round covers it without waiting longer
            if (__ballot(walk.active) != 0ull && l < 2) {
                walk_advance(walk, l, slab);
                walk_request(walk);
            }
            // marks of the next node's neighbours, as soon as its record is
            // in (it was requested before the sift): every write of this
            // round is visible (barrier above) and none can happen before the
            // next round reads them.  Every lane loads (node 0's neighbours
            // when there is no next node).
            {
                const int a0 = pf_r0.x, a1 = a0 >= 0 ? pf_r0.y : -1, a2 = a1 >= 0 ? pf_r0.z : -1;
                pf_vn[0] = pmark[a0 >= 0 ? a0 : 0];
                pf_vn[1] = pmark[a1 >= 0 ? a1 : 0];
                pf_vn[2] = pmark[a2 >= 0 ? a2 : 0];
            }
            // the items still to push, at most two: the dealt items other
            // than m (and d0 when it replaced the root)
            const bool skip0 = repl || m == d0;
            const unsigned long long pa = skip0 ? d1 : d0;
            const unsigned long long pb = (skip0 || m == d1) ? d2 : d1;
            heap_push(q, L, Q, depth, Q - 1, cnt, pa == m ? kKeyNone : pa);
            heap_push(q, L, Q, depth, Q - 1, cnt, pb == m ? kKeyNone : pb);
        }
        fl.set(kPoppedNext, has_next);
        c_next = pf_c;
        const unsigned long long nb = __ballot(has_next) & smask;
        const bool searching = fl.get(kHasReq) && !fl.get(kWalking);
        rounds_here += fl.get(kHasReq) ? 1 : 0;
        // frame statistics per lane (summed over the wave at the end): the
        // lane's expansion and claims, and the segment's round on lane 0
        st_exp += searching && popped ? 1u : 0u;
        st_disc += searching ? (unsigned)cl : 0u;
        st_rounds += searching && l == 0 ? 1u : 0u;
        status = !searching ? status : (found ? 1 : (nb == 0ull ? 2 : status));
        // batched form: a walking segment is done when both of its sides are
        const bool wdone = fl.get(kWalking) && (__ballot(walk.active) & smask) == 0ull;
        const bool ended = (searching && status != 0) || wdone;
        const bool park = fl.get(kHasReq) && !ended && sp.budget > 0 && rounds_here >= sp.budget;
        if (__ballot(ended || park) != 0ull) acquire = true;
        if (__ballot(ended || park) != 0ull && (ended || park)) {
            if (park && !fl.get(kWalking)) {
                // [node to expand next or kKeyNone, heap items]: at most Q keys
                const size_t qb = ((size_t)slot * T + l);
                slots.qcnt[qb] = cnt;
                unsigned long long* dst = slots.qsave + qb * (size_t)Q;
                dst[0] = m;
                for (int i = 0; i < cnt; ++i) dst[1 + i] = q[i * kWave + L];
            }
            if (park && fl.get(kWalking)) park_walk(l, slot, cbase, slots, walk);
            if (l == 0) {
                if (searching) {
                    slots.status[slot] = status;
                    if (ended && status == 2) {  // PathFinder.h:268-274
                        ag.state[agent] = kIdle;
                        ++st_fail;
                        slots.agent[slot] = -1;
                        ag.inflight[agent] = -1;
                    }
                } else if (park) {
                    slots.status[slot] = 5;
                }
                if (wdone) {  // batched form: corridor complete, the agent moves
                    ag.path_off[agent] = (long long)cbase;
                    ag.path_len[agent] = clen;
                    ag.state[agent] = kMove;
                    ag.step_idx[agent] = -1;
                    ag.inflight[agent] = -1;
                    slots.agent[slot] = -1;
                    slots.status[slot] = 1;
                    ++st_found;
                    st_corr += (unsigned)clen;
                }
            }
            bool walk_on = false;
            if (searching && ended && status == 1) {
                const bool hold = sp.budget > 0;
                const bool ok = start_corridor(l, slot, agent, S, Tg, mS, mT, vis, slab, ag, slots, V,
                                               &wshare[3 * sub], walk, hold, cbase, clen, st_found, st_corr);
                if (hold && ok) {
                    if (rounds_here >= sp.budget) {
                        // budget spent in the meeting round: the walk starts
                        // next launch
                        park_walk(l, slot, cbase, slots, walk);
                        if (l == 0) slots.status[slot] = 5;
                    } else {
                        walk_on = true;
                    }
                }
            }
            fl.set(kWalking, walk_on);
            fl.set(kHasReq, walk_on);
            if (walk_on) {
                status = 0;
                fl.set(kPoppedNext, false);
            }
        }
    }
    // walks still pending when the last search of the workgroup ended: the
    // lanes step them together
    if (l < 2)
        while (walk.active) walk_step(walk, l, slab);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        st_found += __shfl_xor(st_found, o, kWave);
        st_fail += __shfl_xor(st_fail, o, kWave);
        st_exp += __shfl_xor(st_exp, o, kWave);
        st_disc += __shfl_xor(st_disc, o, kWave);
        st_corr += __shfl_xor(st_corr, o, kWave);
        st_rounds += __shfl_xor(st_rounds, o, kWave);
    }
    if (L == 0) {
        unsigned long long* const part = fr.stats_part + (blockIdx.x % kStatStripes) * 8;
        if (st_found) atomicAdd(part + 0, (unsigned long long)st_found);
        if (st_fail) atomicAdd(part + 1, (unsigned long long)st_fail);
        if (st_exp) atomicAdd(part + 2, (unsigned long long)st_exp);
        if (st_disc) atomicAdd(part + 3, (unsigned long long)st_disc);
        if (st_corr) atomicAdd(part + 4, (unsigned long long)st_corr);
        if (st_rounds) atomicAdd(part + 5, (unsigned long long)st_rounds);
    }
}

// Compile-time queue capacities (the tuned ones) for T <= 8, where the
// unrolled heap fits in registers at the occupancy the launch needs; other
// (T, Q) use the generic form.
template <int T>
constexpr bool search_q_fixed(int Q) {
    return T <= 8 && (Q == 16 || Q == 32);
}

template <int T>
const void* search_fn_q(int Q, bool kg) {
    if constexpr (T <= 8) {
        if (search_q_fixed<T>(Q)) {
            if (kg) return Q == 16 ? (const void*)k_search<T, 16, true> : (const void*)k_search<T, 32, true>;
            return Q == 16 ? (const void*)k_search<T, 16, false> : (const void*)k_search<T, 32, false>;
        }
    }
    return (const void*)k_search<T, 0, false>;
}

const void* search_fn(int T, int Q, bool kg) {
    switch (T) {
        case 2: return search_fn_q<2>(Q, kg);
        case 4: return search_fn_q<4>(Q, kg);
        case 8: return search_fn_q<8>(Q, kg);
        case 16: return search_fn_q<16>(Q, kg);
        case 32: return search_fn_q<32>(Q, kg);
        case 64: return search_fn_q<64>(Q, kg);
        default: return nullptr;
    }
}

}  // namespace

int search_lds_bytes(int Q) { return kWave * (Q + kSpareRows) * (int)sizeof(unsigned long long); }

hipError_t search_prepare(int T, int Q) {
    for (const bool kg : {false, true}) {
        const void* fn = search_fn(T, Q, kg);
        if (!fn) return hipErrorInvalidValue;
        const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, search_lds_bytes(Q));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_search(const DevMap& map, const DevAgents& ag, const DevSlots& slots, const DevFrame& fr,
                         const SearchParams& sp, int grid, int persistent, hipStream_t st) {
    const size_t lds = (size_t)search_lds_bytes(sp.Q);
    // compact keys when every node id fits their 15-bit field
    const bool kg = map.V <= kCompactMaxNodes && sp.wide == 0;
#define GMAPF_SEARCH_LAUNCH(TT, QQ, KK)                                                                  \
    hipLaunchKernelGGL((k_search<TT, QQ, KK>), dim3(grid), dim3(kWave), lds, st, map, ag, slots, fr, sp, \
                       persistent)
#define GMAPF_SEARCH_CASE(TT)                          \
    case TT:                                           \
        if (!search_q_fixed<TT>(sp.Q))                 \
            GMAPF_SEARCH_LAUNCH(TT, 0, false);         \
        else if (sp.Q == 16 && kg)                     \
            GMAPF_SEARCH_LAUNCH(TT, 16, true);         \
        else if (sp.Q == 16)                           \
            GMAPF_SEARCH_LAUNCH(TT, 16, false);        \
        else if (kg)                                   \
            GMAPF_SEARCH_LAUNCH(TT, 32, true);         \
        else                                           \
            GMAPF_SEARCH_LAUNCH(TT, 32, false);        \
        return hipGetLastError();
#define GMAPF_SEARCH_CASE_GENERIC(TT)                  \
    case TT:                                           \
        GMAPF_SEARCH_LAUNCH(TT, 0, false);             \
        return hipGetLastError();
    switch (sp.T) {
        GMAPF_SEARCH_CASE(2)
        GMAPF_SEARCH_CASE(4)
        GMAPF_SEARCH_CASE(8)
        GMAPF_SEARCH_CASE_GENERIC(16)
        GMAPF_SEARCH_CASE_GENERIC(32)
        GMAPF_SEARCH_CASE_GENERIC(64)
        default:
            return hipErrorInvalidValue;
    }
#undef GMAPF_SEARCH_CASE_GENERIC
#undef GMAPF_SEARCH_CASE
#undef GMAPF_SEARCH_LAUNCH
}

hipError_t search_occupancy(int T, int Q, int* blocks_per_cu) {
    hipError_t e = search_prepare(T, Q);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, search_fn(T, Q, false), kWave,
                                                        (size_t)search_lds_bytes(Q));
}

}  // namespace gmapf_amd
