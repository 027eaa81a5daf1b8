// k_frame.hip -- per-frame kernels around the search: request admission,
// corridor build, move, spatial hash (integer counting sort), avoidance and
// path progress.  Every result is deterministic: orders come from prefix sums
// and ranks over agent / slot indices; where an atomic's arrival order places
// data (the hash's position copy inside a cell), the reader does not depend
// on it.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <mutex>

#include "device.h"
#include "kernels.h"

namespace gmapf_amd {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanBlock = kScanThreads * kScanItems;

__device__ __forceinline__ int block_reduce_sum(int v, int* sh) {
    // 256 threads = 4 waves
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x / 64, ln = threadIdx.x % 64;
    __syncthreads();
    if (ln == 0) sh[w] = v;
    __syncthreads();
    int t = 0;
    if (threadIdx.x == 0) t = sh[0] + sh[1] + sh[2] + sh[3];
    return t;  // valid in thread 0
}

// exclusive scan over the block; returns the exclusive prefix of v, *total = block sum
__device__ __forceinline__ int block_exscan(int v, int* sh, int* total) {
    const int ln = threadIdx.x % 64, w = threadIdx.x / 64;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (ln >= o) x += y;
    }
    __syncthreads();
    if (ln == 63) sh[w] = x;
    __syncthreads();
    int wpre = 0;
    for (int k = 0; k < w; ++k) wpre += sh[k];
    *total = sh[0] + sh[1] + sh[2] + sh[3];
    return wpre + x - v;
}

// per-block sums of vals[0, n); nothing when *gate is 0 (gate may be null)
__global__ __launch_bounds__(kScanThreads) void k_block_sum(const int* vals, int n, int* blk, const int* gate) {
    __shared__ int sh[4];
    if (gate && *gate == 0) return;
    const long long base = (long long)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += vals[base + k];
    int t = block_reduce_sum(s, sh);
    if (threadIdx.x == 0) blk[blockIdx.x] = t;
}

// exclusive scan of vals into out; *total = sum of all.  clear (null or
// vals itself): zero the values once read.  Nothing when *gate is 0.
__global__ __launch_bounds__(kScanThreads) void k_block_scan(const int* vals, int n, const int* blk,
                                                             int* out, int* total, int* clear, const int* gate) {
    __shared__ int sh[4];
    __shared__ int sh2[4];
    if (gate && *gate == 0) return;
    int o = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += kScanThreads) o += blk[b];
    int off = block_reduce_sum(o, sh);
    __shared__ int off_sh;
    if (threadIdx.x == 0) off_sh = off;
    __syncthreads();
    off = off_sh;
    const long long base = (long long)blockIdx.x * kScanBlock + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? vals[base + k] : 0;
        s += v[k];
    }
    if (clear) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k)
            if (base + k < n) clear[base + k] = 0;
    }
    int btot = 0;
    int pre = block_exscan(s, sh2, &btot);
    int run = off + pre;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0 && total) *total = off + btot;
}

hipError_t exscan(const int* vals, int n, int* blk, int* out, int* total, hipStream_t st, int* clear = nullptr,
                  const int* gate = nullptr) {
    if (n <= 0) return hipSuccess;
    int nb = (n + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_block_sum, dim3(nb), dim3(kScanThreads), 0, st, vals, n, blk, gate);
    hipLaunchKernelGGL(k_block_scan, dim3(nb), dim3(kScanThreads), 0, st, vals, n, (const int*)blk,
                       out, total, clear, gate);
    return hipGetLastError();
}

__global__ void k_fill(int* p, int v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, unsigned long long v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------- geometry
// CuNodesMap.h:272-275
__device__ __forceinline__ float get_sign(float p1x, float p1y, float p2x, float p2y, float p3x,
                                          float p3y) {
    return (p1x - p3x) * (p2y - p3y) - (p2x - p3x) * (p1y - p3y);
}

// CuNodesMap.h:260-270
__device__ __forceinline__ bool in_tri_v(float2 a, float2 b, float2 c, float px, float py) {
    float d1 = get_sign(px, py, a.x, a.y, b.x, b.y);
    float d2 = get_sign(px, py, b.x, b.y, c.x, c.y);
    float d3 = get_sign(px, py, c.x, c.y, a.x, a.y);
    bool neg = (d1 < 0) || (d2 < 0) || (d3 < 0);
    bool pos = (d1 > 0) || (d2 > 0) || (d3 > 0);
    return !(neg && pos);
}

__device__ __forceinline__ bool in_tri(const DevMap& m, int t, float px, float py) {
    return in_tri_v(m.tri[3 * (size_t)t], m.tri[3 * (size_t)t + 1], m.tri[3 * (size_t)t + 2], px, py);
}

__device__ __forceinline__ int cell_of(float v, float o, float inv, int n) {
    float f = floorf((v - o) * inv);
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)(n - 1)) return n - 1;
    return (int)f;
}

// CuNodesMap.h:216-229: first (minimum) triangle index containing p, through
// the margin grid; points outside the gridded region fall back to the scan.
__device__ int locate(const DevMap& m, float x, float y) {
    if (!(x >= m.rx0 && x <= m.rx1 && y >= m.ry0 && y <= m.ry1)) {
        for (int t = 0; t < m.V; ++t)
            if (in_tri(m, t, x, y)) return t;
        return -1;
    }
    const int cx = cell_of(x, m.gx0, m.inv_cs, m.nx), cy = cell_of(y, m.gy0, m.inv_cs, m.ny);
    const size_t c = (size_t)cy * m.nx + cx;
    int best = -1;
    // the cell's triangles in ascending id, four at a time (their loads go
    // out together; a short group repeats its first triangle, masked): the
    // first that contains the point is the smallest id
    for (int k0 = m.cell_start[c], e = m.cell_start[c + 1]; k0 < e && best < 0; k0 += 4) {
        int tt[4];
        bool in[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) tt[u] = m.cell_tris[k0 + u < e ? k0 + u : k0];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool r = in_tri(m, tt[u], x, y);  // every tt[u] is a valid triangle
            in[u] = (k0 + u < e) & r;
        }
#pragma unroll
        for (int u = 3; u >= 0; --u) best = in[u] ? tt[u] : best;
    }
    for (int k = 0; k < m.n_global; ++k) {
        int t = m.global_tris[k];
        if (best >= 0 && t > best) break;
        if (in_tri(m, t, x, y)) {
            best = t;
            break;
        }
    }
    return best;
}

// ------------------------------------------------------------- admission
// F1 (PathFinder.h:85-116) and F2 (AgentsMover.h:39-46) as two grid-wide
// passes, one agent or slot per thread (blocks [0, nbA) agents, the rest
// slots):
//   count:   a Search agent with no search in flight is a request, or turns
//            Idle when its node equals its target node (PathFinder.h:91-95);
//            Move agents form the moving set; a slot with no agent is free.
//            Block totals go to blk (requests), mblk (moving) and sblk (free
//            slots); the last block to finish turns blk and sblk into
//            exclusive prefixes and writes the frame's counts.
//   scatter: the r-th request in ascending agent index is admitted when
//            r < number of free slots (the rest stay Search,
//            PathFinder.h:97-102); free slots are listed ascending, occupied
//            ones start the active list.
// k_admit_slots initialises the admitted searches, the r-th request in the
// r-th free slot (PathFinder.h:104-115), and appends their slots to the
// active list.  A search whose agent got a new target was dropped before,
// by k_apply_updates.
constexpr int kAdmitThreads = 1024;
constexpr int kAdmitWaves = kAdmitThreads / 64;

// exclusive scan of one value per thread over the block; returns the
// thread's prefix and the block total.  Starts with a barrier, so
// consecutive calls may reuse sh.
__device__ __forceinline__ int admit_scan(int v, int* sh, int* total) {
    const int ln = threadIdx.x % 64, w = threadIdx.x / 64;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (ln >= o) x += y;
    }
    __syncthreads();
    if (ln == 63) sh[w] = x;
    __syncthreads();
    int wpre = 0, tot = 0;
    for (int k = 0; k < kAdmitWaves; ++k) {
        const int t = sh[k];
        wpre += (k < w) ? t : 0;
        tot += t;
    }
    *total = tot;
    return wpre + x - v;
}

// exclusive rank of a flag within the block and the block's flag count;
// wsum must not be in use by another call still being read
__device__ __forceinline__ int admit_rank(bool f, int* wsum, int* total) {
    const int ln = threadIdx.x % 64, w = threadIdx.x / 64;
    const unsigned long long bal = __ballot(f);
    if (ln == 0) wsum[w] = __popcll(bal);
    __syncthreads();
    int pre = 0, tot = 0;
    for (int k = 0; k < kAdmitWaves; ++k) {
        const int t = wsum[k];
        pre += (k < w) ? t : 0;
        tot += t;
    }
    *total = tot;
    return pre + __popcll(bal & ((1ull << ln) - 1ull));
}

__device__ __forceinline__ int load_agent_scope(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// exclusive prefix over vals[0..n) in place by one block; returns the total
__device__ int admit_prefix(int* vals, int n, int* sh) {
    int carry = 0;
    for (int base = 0; base < n; base += kAdmitThreads) {
        const int k = base + threadIdx.x;
        const int v = k < n ? load_agent_scope(vals + k) : 0;
        int tot;
        const int pre = admit_scan(v, sh, &tot);
        if (k < n) vals[k] = carry + pre;
        carry += tot;
    }
    return carry;
}

__global__ __launch_bounds__(kAdmitThreads) void k_admit_count(DevAgents ag, DevSlots sl, DevFrame fr, int nbA) {
    __shared__ int wsum_f[kAdmitWaves], wsum_m[kAdmitWaves];
    const int b = blockIdx.x;
    bool f = false, mv = false;
    if (b < nbA) {
        const int i = b * kAdmitThreads + threadIdx.x;
        if (i < ag.n) {
            const int st = ag.state[i];
            mv = st == kMove;
            ag.in_m[i] = mv ? 1 : 0;
            if (st == kSearch && ag.inflight[i] < 0) {
                if (ag.curr[i] == ag.targ[i])
                    ag.state[i] = kIdle;
                else
                    f = true;
            }
        }
    } else {
        const int s = (b - nbA) * kAdmitThreads + threadIdx.x;
        f = s < sl.P && sl.agent[s] < 0;
    }
    int nf, nm;
    admit_rank(f, wsum_f, &nf);
    admit_rank(mv, wsum_m, &nm);
    if (threadIdx.x == 0) {
        if (b < nbA) {
            fr.blk[b] = nf;
            fr.mblk[b] = nm;
        } else {
            fr.sblk[b - nbA] = nf;
        }
    }
}

// One block after k_admit_count: the blocks' prefixes, the frame's counters
// and the ticket stride.  (A last-block-done tail inside k_admit_count would
// save this launch, but every block's device-scope fence before the done
// count writes its L2 lines back, which costs more than the launch on a
// multi-XCD part.)
__global__ __launch_bounds__(kAdmitThreads) void k_admit_finish(DevSlots sl, DevFrame fr, int nbA, int nbS) {
    __shared__ int sh[kAdmitWaves];
    const int nreq = admit_prefix(fr.blk, nbA, sh);
    const int nfree = admit_prefix(fr.sblk, nbS, sh);
    int mv_part = 0;
    for (int k = threadIdx.x; k < nbA; k += kAdmitThreads) mv_part += load_agent_scope(fr.mblk + k);
    int moving;
    admit_scan(mv_part, sh, &moving);
    if (threadIdx.x == 0) {
        const int admitted = nreq < nfree ? nreq : nfree;
        fr.counters[0] = nreq;
        fr.counters[1] = (sl.P - nfree) + admitted;
        fr.counters[2] = 0;
        fr.counters[3] = 0;
        fr.counters[4] = nfree;
        fr.counters[5] = moving;
        fr.counters[6] = admitted;
        // Searches still running from earlier frames (the long ones) come
        // first in the active list and new ones last; a stride coprime with
        // the count, near 0.618 of it, deals both kinds to every workgroup,
        // so no workgroup holds only long searches (a launch lasts as long
        // as its slowest workgroup).
        const int na = fr.counters[1];
        int stride = 1;
        if (na > 2) {
            stride = (int)(((long long)na * 40503) >> 16) | 1;
            for (;; stride += 2) {
                int a = na, b = stride;
                while (b) {
                    const int t = a % b;
                    a = b;
                    b = t;
                }
                if (a == 1) break;
            }
        }
        fr.counters[9] = stride;
        fr.stats[6] += (unsigned long long)admitted;
    }
}

__global__ __launch_bounds__(kAdmitThreads) void k_admit_scatter(DevAgents ag, DevSlots sl, DevFrame fr, int nbA) {
    __shared__ int wsum[kAdmitWaves];
    const int b = blockIdx.x;
    int tot;
    if (b < nbA) {
        const int i = b * kAdmitThreads + threadIdx.x;
        const bool f = i < ag.n && ag.state[i] == kSearch && ag.inflight[i] < 0;
        const int r = fr.blk[b] + admit_rank(f, wsum, &tot);
        if (f && r < fr.counters[4]) fr.adm_agent[r] = i;
    } else {
        const int s = (b - nbA) * kAdmitThreads + threadIdx.x;
        const bool f = s < sl.P && sl.agent[s] < 0;
        const int r = fr.sblk[b - nbA] + admit_rank(f, wsum, &tot);
        if (s < sl.P) {
            if (f)
                fr.free_list[r] = s;
            else
                fr.active[s - r] = s;  // s - r = occupied slots before s
        }
    }
}

// Admission when every agent has a slot of its own (P >= N): every request
// is admitted (PathFinder.h:85-116 with no deferral), into slot i for agent i
// (PathFinder.h:104-115); the moving set is taken as in k_admit_count.  One
// pass, no ranks, and the frame's first kernel: the workgroup first applies
// the target changes of its own agents (their bucket of the input; a change
// frees only the agent's own slot), then admits them.  counters[5] (0 since
// the last export) is set when some agent moves, and the admissions are
// added to stats[6] through its stripes.
static_assert(kAdmitThreads == kUpdBucket, "a workgroup applies and admits one bucket's agents");
__device__ __forceinline__ void apply_update(const DevAgents& ag, const DevSlots& sl, const DevFrame& fr, int k);
__global__ __launch_bounds__(kAdmitThreads) void k_admit_identity(DevAgents ag, DevSlots sl, DevFrame fr) {
    __shared__ int wsum_m[kAdmitWaves], wsum_a[kAdmitWaves];
    const int i = blockIdx.x * kAdmitThreads + threadIdx.x;
    if (i == 0) {
        *fr.stepv = *fr.h_step;
        fr.counters[2] = 0;  // search tickets
    }
    apply_update(ag, sl, fr, i);
    __syncthreads();  // the workgroup's changes before its admission
    bool mv = false, adm = false;
    if (i < ag.n) {
        const int st = ag.state[i];
        mv = st == kMove;
        ag.in_m[i] = mv ? 1 : 0;
        if (st == kSearch && ag.inflight[i] < 0) {
            if (ag.curr[i] == ag.targ[i]) {
                ag.state[i] = kIdle;
            } else {
                sl.agent[i] = i;
                sl.gen[i] = ag.gen[i];
                sl.S[i] = ag.curr[i];
                sl.Tg[i] = ag.targ[i];
                sl.stamp[i] = sl.stamp[i] + 1u;
                sl.status[i] = 3;
                ag.inflight[i] = i;
                ag.path_len[i] = 0;  // old corridor released (PathFinder.h:104-108)
                adm = true;
            }
        }
    }
    int nm, na;
    admit_rank(mv, wsum_m, &nm);
    admit_rank(adm, wsum_a, &na);
    // the movers' count is only tested against 0, and nothing reads the
    // admitted count of this form: a flag and a striped add, no atomics on
    // one word from every block
    if (threadIdx.x == 0) {
        if (nm) fr.counters[5] = 1;
        if (na) atomicAdd(&fr.stats_part[(blockIdx.x % kStatStripes) * 8 + 6], (unsigned long long)na);
    }
}

// k-th admission: agent adm_agent[k] into slot free_list[k] (PathFinder.h:104-115)
__global__ void k_admit_slots(DevAgents ag, DevSlots sl, DevFrame fr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int admitted = fr.counters[6];
    if (k >= admitted) return;
    const int s = fr.free_list[k];
    const int i = fr.adm_agent[k];
    sl.agent[s] = i;
    sl.gen[s] = ag.gen[i];
    sl.S[s] = ag.curr[i];
    sl.Tg[s] = ag.targ[i];
    const unsigned st = sl.stamp[s] + 1u;
    sl.stamp[s] = st;
    sl.status[s] = 3;
    ag.inflight[i] = s;
    ag.path_len[i] = 0;  // old corridor released (PathFinder.h:104-108)
    fr.active[fr.counters[1] - admitted + k] = s;
}


// GMAPF.cu:247-250 for target changes made after start, read from pinned
// host memory; the host passes at most one update per agent (the last one).
// An agent whose search is in flight drops it: its slot is free for this
// frame's admission, and the agent requests again with the new target.
// Entry k of the bucketed input (thread k): every lane reads its bucket's
// count (one request per wave), then the lanes below it read their entry, so
// a frame with few updates reads little host memory.
__device__ __forceinline__ void apply_update(const DevAgents& ag, const DevSlots& sl, const DevFrame& fr, int k) {
    const int n = fr.h_upd_bcnt[k / kUpdBucket];
    if (k % kUpdBucket >= n || k >= ag.n) return;
    const int i = fr.h_upd_ids[k];
    const int tg = fr.h_upd_targ[k];
    const float2 xy = fr.h_upd_xy[k];
    ag.state[i] = kSearch;
    ag.targ[i] = tg;
    ag.targ_pos[i] = xy;
    ag.gen[i] = ag.gen[i] + 1;
    const int s = ag.inflight[i];
    if (s >= 0) {
        sl.agent[s] = -1;
        ag.inflight[i] = -1;
    }
}

__global__ void k_apply_updates(DevAgents ag, DevSlots sl, DevFrame fr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        *fr.stepv = *fr.h_step;
        fr.counters[2] = 0;  // search tickets (the ranked admission clears it too)
        fr.counters[5] = 0;  // counted by the admission pass that follows
        fr.counters[6] = 0;
    }
    apply_update(ag, sl, fr, k);
}

// States and counters of the finished frame into pinned host memory.  The
// positions are written there by the crowd's last kernel (the only one that
// moves agents; it runs beside the search), so this pass after the search
// writes the states of the 64-agent groups (one wave each) in which a state
// changed since the last export, whole, so the host writes stay full lines;
// the host buffer keeps the other groups' states (a crowd frame changes
// few).  No kernel of the search writes host memory: a
// store there inside the round loop would be waited for by the next round's
// wait for all memory operations (the counter is in order), and the stores
// would also hold up the end of the launch.
// Thread i of a grid of at least max(N, 8) threads, in whole waves.
// kFold: the frame ran a search or an admission, whose counters are in the
// stripes (blocks of 256 threads; block 0 folds them).
template <bool kFold>
__device__ __forceinline__ void export_agent(const DevAgents& ag, const DevFrame& fr, int i) {
    const int s = i < ag.n ? ag.state[i] : 0;
    const bool changed = i < ag.n && s != fr.state_prev[i];
    if (__ballot(changed) != 0ull && i < ag.n) {
        fr.h_state_out[i] = s;
        if (changed) fr.state_prev[i] = s;
    }
    // agents left searching and moving, per 64-agent group (the host adds
    // them up: a frame after which none searches and some move, and before
    // which no target changed, needs no admission or search), and whether
    // the group's states were written (the host copies only those groups)
    const unsigned long long srch = __ballot(i < ag.n && s == kSearch);
    const unsigned long long mov = __ballot(i < ag.n && s == kMove);
    const unsigned long long chg = __ballot(changed);
    if ((threadIdx.x & 63) == 0 && i < ag.n)
        fr.h_group_out[i >> 6] = __popcll(srch) | (__popcll(mov) << 8) | (chg != 0ull ? 1 << 16 : 0);
    if constexpr (kFold) {
        static_assert(kStatStripes == 64, "the fold takes two stripes per thread of 256");
        if (blockIdx.x == 0) {
            // the frame's additions, striped over 64 sets of 8 words, into
            // the cumulative counters (and the stripes back to 0): thread t
            // takes counter t % 8 of stripes t / 8 and t / 8 + 32, then a
            // tree over the threads of one counter
            __shared__ unsigned long long red[256];
            const int t = threadIdx.x;
            const int a = t, b = t + 256;  // word p * 8 + c of stripes p and p + 32
            const unsigned long long v = fr.stats_part[a] + fr.stats_part[b];
            fr.stats_part[a] = 0ull;
            fr.stats_part[b] = 0ull;
            red[t] = v;
            __syncthreads();
            for (int s = 128; s >= 8; s >>= 1) {
                if (t < s) red[t] += red[t + s];
                __syncthreads();
            }
            if (t < 8) {
                const unsigned long long tot = fr.stats[t] + red[t];
                fr.stats[t] = tot;
                fr.h_stats_out[t] = tot;
            }
        }
    } else {
        if (i < 8) fr.h_stats_out[i] = fr.stats[i];
    }
    if (i < 2) fr.h_slab_out[i] = ag.slab_top[i];  // [0] top, [1] deferred (adjacent words)
    // the movers of this frame are counted; the next frame counts from 0
    // (its admission pass, or the move pass of a frame without one, adds
    // them up again)
    if (i == 0) fr.counters[5] = 0;
}

__global__ void k_export(DevAgents ag, DevFrame fr) {
    export_agent<true>(ag, fr, blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------- corridor slab
// Live length per owner, owners 0 .. n-1 the agents and n .. n+P-1 the search
// slots: a moving agent keeps its corridor, a slot keeps the corridor of its
// parked walk (batched form; the walk stores its positions relative to the
// corridor's offset), everything else is released.
__global__ void k_slab_len(DevAgents ag, DevSlots sl, DevFrame fr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ag.n) {
        fr.slab_len[i] = (ag.state[i] == kMove && ag.path_len[i] > 0) ? ag.path_len[i] : 0;
    } else if (i < ag.n + sl.P) {
        const int s = i - ag.n;
        fr.slab_len[i] = (sl.status[s] == 5 && sl.agent[s] >= 0) ? sl.corr[s].y : 0;
    }
}

// after the scan slab_len[o] is owner o's new offset and slab_len[n + P] the
// total; one wave copies one owner's corridor
__global__ void k_slab_copy(DevAgents ag, DevSlots sl, DevFrame fr) {
    const int* src = *ag.slab_now;
    int* dst = src == ag.slab[0] ? ag.slab[1] : ag.slab[0];
    const int lane = threadIdx.x % 64;
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int nw = (gridDim.x * blockDim.x) / 64;
    for (int a = w; a < ag.n + sl.P; a += nw) {
        const int off = fr.slab_len[a];
        const int len = fr.slab_len[a + 1] - off;
        if (len == 0) {
            // a released corridor (the agent is not moving): the agent has
            // none from now on, so no reader follows its old offset into the
            // new buffer
            if (lane == 0 && a < ag.n) ag.path_len[a] = 0;
            continue;
        }
        const long long from = a < ag.n ? ag.path_off[a] : (long long)sl.corr[a - ag.n].x;
        for (int k = lane; k < len; k += 64) dst[off + k] = src[from + k];
        if (lane == 0) {
            if (a < ag.n) {
                ag.path_off[a] = off;
                ag.path_len[a] = len;
            } else {
                sl.corr[a - ag.n].x = off;
            }
        }
    }
}

__global__ void k_slab_finish(DevAgents ag, DevSlots sl, DevFrame fr) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int* cur = *ag.slab_now;
    *ag.slab_now = cur == ag.slab[0] ? ag.slab[1] : ag.slab[0];
    *ag.slab_top = (unsigned long long)fr.slab_len[ag.n + sl.P];
}

// ------------------------------------------------------------- crowd
// Spatial hash = stable counting sort of the agents by cell key (the order
// std::stable_sort by key gives):
//   k_move_key     move, key of the moved position, histogram; the atomic's
//                  return value (the agent's arrival slot in its cell) is kept
//   scan           exclusive prefix of the counts -> cell starts; the counts
//                  are zeroed as they are read, ready for the next frame
//   k_hash_place   each agent's position to start + arrival slot (cell
//                  groups complete, arrival order inside a cell); the
//                  avoidance pass reads only this copy
//   k_avoid_progress  while it reads its own cell, each agent counts the
//                  members with a smaller index: its rank, and its place in
//                  the final order (horder)
// AgentsMover.h:57-70 (move), then the agent's hash cell from its moved
// position, counted into hcount (all zero at the start of a frame).
// kAlone: a frame without admission, so this pass takes the moving set (the
// agents in Move).  The host runs such a frame only when some agent moves,
// and the passes after this one only test whether the movers' count is 0:
// it is set to 1.  (Counted with one atomic per wave on the one word, the
// pass took 16.2 us against 6.1 us for the pass after an admission.)
template <bool kAlone>
__global__ void k_move_key(DevMap map, DevAgents ag, DevFrame fr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool mv;
    if constexpr (kAlone) {
        if (i == 0) fr.counters[5] = 1;
        if (i >= ag.n) return;
        mv = ag.state[i] == kMove;
        ag.in_m[i] = mv ? 1 : 0;
    } else {
        if (i >= ag.n || fr.counters[5] == 0) return;
        mv = ag.in_m[i] != 0;
    }
    float2 p = ag.pos[i];
    if (mv && ag.step_idx[i] != -1) {
        const float step = *fr.stepv;
        const float2 sp = ag.step_pos[i];
        float dx = sp.x - p.x, dy = sp.y - p.y;
        float len = sqrtf(dx * dx + dy * dy);
        if (len <= step) {
            p = sp;
        } else {
            float ux = dx / len, uy = dy / len;
            p.x = p.x + ux * step;
            p.y = p.y + uy * step;
        }
        ag.pos[i] = p;
    }
    int key = -1;
    if (!(isnan(p.x) || isnan(p.y))) {
        int cx = cell_of(p.x, map.hx0, map.hinv, map.hnx);
        int cy = cell_of(p.y, map.hy0, map.hinv, map.hny);
        key = cy * map.hnx + cx;
    }
    // agents with a NaN position go to the extra cell ncell, after every
    // hashed agent (key -1: not in the hash's order)
    fr.hslot[i] = atomicAdd(&fr.hcount[key >= 0 ? key : fr.ncell], 1);
    fr.hkey[i] = key;
}

__device__ __forceinline__ int cell_start(const DevFrame& fr, int c) { return fr.hstart[c]; }

// The agent's entry at its cell's start plus its arrival slot: the cells are
// grouped in key order, the agents of a cell in arrival order (the avoidance
// pass takes a cell's agents by index itself).
__global__ void k_hash_place(DevAgents ag, DevFrame fr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ag.n || fr.counters[5] == 0) return;
    const int key = fr.hkey[i];
    const float2 p = ag.pos[i];
    fr.hent[cell_start(fr, key >= 0 ? key : fr.ncell) + fr.hslot[i]] = make_float4(p.x, p.y, __int_as_float(i), 0.f);
}

// AgentsMover.h:96-126 over the 3x3 (generally (2k+1)^2) stencil, Jacobi, the
// neighbour terms summed in ascending agent index, then AgentsMover.h:137-176
// (path progress) for the moving set.  Neighbours are read from the hash's
// position copy (hent), never from pos, so the pass writes each agent's new
// position in place and the sum still sees only post-move positions.
constexpr int kAvoidThreads = 256;
constexpr int kAvoidList = 16;  // neighbours an agent keeps in LDS; more are taken by rescans
constexpr int kAvoidRowsHalf = 2;  // stencils of up to 2 * 2 + 1 rows read their rows at once
constexpr int kAvoidRowLoad = 8;   // entries of each row requested at once

__device__ __forceinline__ void avoid_term(float2 p, float2 o, const CrowdParams& cp, float& sx, float& sy,
                                           int& cnt) {
    float dx = p.x - o.x, dy = p.y - o.y;
    float dist = sqrtf(dx * dx + dy * dy);
    float shift = cp.cd - dist;
    sx += (dx / dist) * shift;
    sy += (dy / dist) * shift;
    ++cnt;
}

// The neighbours within cd of agent i are summed in ascending agent index.
// One pass over the stencil puts them into a per-thread LDS list (entry m of
// thread t at m * kAvoidThreads + t: no bank conflicts); the sum then takes
// the list's entries by ascending index.  An agent with more neighbours than
// the list holds takes them by rescans of the stencil, one per neighbour,
// with the same result.
__global__ __launch_bounds__(kAvoidThreads) void k_avoid_progress(DevMap map, DevAgents ag, DevFrame fr,
                                                                  CrowdParams cp) {
    __shared__ int lj[kAvoidList * kAvoidThreads];
    __shared__ float2 lp[kAvoidList * kAvoidThreads];
    const int t = threadIdx.x;
    // thread q takes the agent of hash entry q: the threads of a wave take
    // agents of neighbouring cells and read the same stencil entries
    const int q_self = blockIdx.x * blockDim.x + t;
    if (q_self >= ag.n || fr.counters[5] == 0) return;
    const float4 me = fr.hent[q_self];
    const int i = __float_as_int(me.z);
    const float2 p = make_float2(me.x, me.y);
    // What does not depend on the avoidance is requested now, beside it: the
    // path progress inputs (the triangle of the agent's next corridor node,
    // the corridor node after it and its centroid) and the triangle of the
    // agent's current node (the first test of whether the avoided position is
    // on the navmesh).
    const int moving = ag.in_m[i];
    const int idx0 = ag.step_idx[i];
    const int node = ag.step_node[i];
    const int plen = ag.path_len[i];
    const int cur = ag.curr[i];
    const int tn = moving && idx0 != -1 ? node : 0;
    const float2 ta = map.tri[3 * (size_t)tn], tb = map.tri[3 * (size_t)tn + 1], tc = map.tri[3 * (size_t)tn + 2];
    const int cn = cur >= 0 ? cur : 0;
    const float2 ua = map.tri[3 * (size_t)cn], ub = map.tri[3 * (size_t)cn + 1], uc = map.tri[3 * (size_t)cn + 2];
    // corridor index the progress reads: 0 when the walk starts, else the one
    // after the current step (read only when it exists)
    const int nidx = idx0 == -1 ? 0 : idx0 + 1;
    const bool pread = moving && (idx0 == -1 || nidx < plen);
    const int* const slab = *ag.slab_now;
    const int pn = *(pread ? slab + ag.path_off[i] + nidx : slab);
    const float2 pc = map.cen[pread ? pn : 0];
    float sx = 0.f, sy = 0.f;
    int cnt = 0;
    if (q_self < fr.hstart[fr.ncell]) {  // a hashed agent (entries past them: NaN positions)
        const int cx = cell_of(p.x, map.hx0, map.hinv, map.hnx);
        const int cy = cell_of(p.y, map.hy0, map.hinv, map.hny);
        const int key = cy * map.hnx + cx;
        const int x0 = max(0, cx - map.hk), x1 = min(map.hnx - 1, cx + map.hk);
        const int y0 = max(0, cy - map.hk), y1 = min(map.hny - 1, cy + map.hk);
        // k: neighbours found (the first kAvoidList go to the LDS list); r: the
        // agent's rank in its cell, the number of the cell's agents with a
        // smaller index (its place in the stable counting sort's order)
        int k = 0, r = 0;
        const int ob = cell_start(fr, key), oe = cell_start(fr, key + 1);
        // entry q of the hash (q = -1: a padding entry, which holds the agent
        // itself and counts for nothing)
        auto take = [&](int q, const float4& o) {
            const int j = __float_as_int(o.z);
            r += (q >= ob && q < oe && j < i) ? 1 : 0;
            if (j == i) return;
            const float dx = p.x - o.x, dy = p.y - o.y;
            if (dx * dx + dy * dy < cp.cd2) {
                if (k < kAvoidList) {
                    lj[k * kAvoidThreads + t] = j;
                    lp[k * kAvoidThreads + t] = make_float2(o.x, o.y);
                }
                ++k;
            }
        };
        // a row's entries from q0 on, four at a time, so their loads go out
        // together; the list order does not matter, the sum below takes the
        // neighbours by ascending index
        auto take_rest = [&](int q0, int e) {
            for (; q0 < e; q0 += 4) {
                float4 ee[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) ee[u] = fr.hent[q0 + u < e ? q0 + u : q_self];
#pragma unroll
                for (int u = 0; u < 4; ++u) take(q0 + u < e ? q0 + u : -1, ee[u]);
            }
        };
        if (map.hk <= kAvoidRowsHalf) {
            // every row's bounds, then the first kAvoidRowLoad entries of
            // every row, each group requested at once: two waits for memory
            // where a row-by-row walk waits twice per row
            int rb[2 * kAvoidRowsHalf + 1], re[2 * kAvoidRowsHalf + 1];
#pragma unroll
            for (int rr = 0; rr <= 2 * kAvoidRowsHalf; ++rr) {
                const bool v = y0 + rr <= y1;
                const int row = (v ? y0 + rr : y0) * map.hnx;
                const int b = cell_start(fr, row + x0), e = cell_start(fr, row + x1 + 1);
                rb[rr] = v ? b : 0;
                re[rr] = v ? e : 0;
            }
            float4 ent[2 * kAvoidRowsHalf + 1][kAvoidRowLoad];
#pragma unroll
            for (int rr = 0; rr <= 2 * kAvoidRowsHalf; ++rr)
#pragma unroll
                for (int u = 0; u < kAvoidRowLoad; ++u) {
                    const int q = rb[rr] + u;
                    ent[rr][u] = fr.hent[q < re[rr] ? q : q_self];
                }
#pragma unroll
            for (int rr = 0; rr <= 2 * kAvoidRowsHalf; ++rr)
#pragma unroll
                for (int u = 0; u < kAvoidRowLoad; ++u) take(rb[rr] + u < re[rr] ? rb[rr] + u : -1, ent[rr][u]);
            // rows longer than that
#pragma unroll
            for (int rr = 0; rr <= 2 * kAvoidRowsHalf; ++rr) take_rest(rb[rr] + kAvoidRowLoad, re[rr]);
        } else {
            for (int yy = y0; yy <= y1; ++yy) {
                const int row = yy * map.hnx;
                take_rest(cell_start(fr, row + x0), cell_start(fr, row + x1 + 1));
            }
        }
        fr.horder[ob + r] = i;
        if (k <= kAvoidList) {
            int last = -1;
            for (int m = 0; m < k; ++m) {
                int best = INT_MAX;
                float2 bp = make_float2(0.f, 0.f);
                for (int r = 0; r < k; ++r) {
                    const int j = lj[r * kAvoidThreads + t];
                    if (j > last && j < best) {
                        best = j;
                        bp = lp[r * kAvoidThreads + t];
                    }
                }
                avoid_term(p, bp, cp, sx, sy, cnt);
                last = best;
            }
            k = -1;  // done
        }
        int last = -1;
        while (k >= 0) {
            int best = INT_MAX;
            float2 bp = make_float2(0.f, 0.f);
            for (int yy = y0; yy <= y1; ++yy) {
                const int row = yy * map.hnx;
                const int b = cell_start(fr, row + x0), e = cell_start(fr, row + x1 + 1);
                for (int q = b; q < e; ++q) {
                    const float4 o = fr.hent[q];
                    const int j = __float_as_int(o.z);
                    if (j <= last || j >= best || j == i) continue;
                    float dx = p.x - o.x, dy = p.y - o.y;
                    if (dx * dx + dy * dy < cp.cd2) {
                        best = j;
                        bp = make_float2(o.x, o.y);
                    }
                }
            }
            if (best == INT_MAX) break;
            avoid_term(p, bp, cp, sx, sy, cnt);
            last = best;
        }
    }
    float2 out = p;
    if (cnt > 0) {
        const float fc = (float)cnt;
        const float ax = sx / fc, ay = sy / fc;
        const float rx = ax * cp.rot_cs - ay * cp.rot_sn;
        const float ry = ax * cp.rot_sn + ay * cp.rot_cs;
        const float nx = p.x + rx, ny = p.y + ry;
        // on the navmesh when some triangle holds it: the current node's and
        // the next corridor node's are tried before the point location
        const bool on = (cur >= 0 && in_tri_v(ua, ub, uc, nx, ny)) ||
                        (moving && idx0 != -1 && in_tri_v(ta, tb, tc, nx, ny)) || locate(map, nx, ny) >= 0;
        if (on) out = make_float2(nx, ny);
    }
    ag.pos[i] = out;
    // AgentsMover.h:137-176 (after the Jacobi write-back)
    if (!moving) return;
    int idx = idx0;
    if (idx == -1) {
        ag.step_idx[i] = 0;
        ag.step_node[i] = pn;  // corridor node 0
        ag.step_pos[i] = pc;
        return;
    }
    if (!in_tri_v(ta, tb, tc, out.x, out.y)) return;
    ag.curr[i] = node;
    const int count = plen;
    if (idx < count) idx += 1;
    ag.step_idx[i] = idx;
    if (idx == count) {
        const float2 tp = ag.targ_pos[i];
        float dx = tp.x - out.x, dy = tp.y - out.y;
        if (sqrtf(dx * dx + dy * dy) > FLT_EPSILON + cp.slack)
            ag.step_pos[i] = tp;
        else
            ag.state[i] = kIdle;
    } else {
        ag.step_node[i] = pn;  // corridor node idx = idx0 + 1
        ag.step_pos[i] = pc;
    }
}

// The step of a frame without admission, when it differs from the last frame's.
__global__ void k_set_step(DevFrame fr) {
    if (threadIdx.x == 0) *fr.stepv = *fr.h_step;
}

// The frame's positions into pinned host memory in agent order (the host
// writes stay whole lines), after the avoidance pass has written them all.
// kExport: a frame without a search, whose export this pass also does (it
// clears the movers' count, so it does not test it: such a frame has movers).
template <bool kExport>
__global__ void k_pos_out(DevAgents ag, DevFrame fr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if constexpr (kExport) {
        if (i < ag.n) fr.h_pos_out[i] = ag.pos[i];
        export_agent<false>(ag, fr, i);  // no search or admission ran
    } else {
        if (i >= ag.n || fr.counters[5] == 0) return;
        fr.h_pos_out[i] = ag.pos[i];
    }
}

// Staging for bulk point location without allocation: the points, then the
// results.  One copy per GPU, held until the process ends; every call on a
// GPU shares it.  Host code reaches it only through the address that
// hipGetSymbolAddress gives for the current device (the array's own name is a
// host-side stand-in there).
constexpr size_t kLocateScratchBytes = (size_t)kLocateScratchMax * (sizeof(float2) + sizeof(int));
static __device__ __align__(256) unsigned char scratch[kLocateScratchBytes];

__global__ void k_locate(DevMap map, const float2* pts, int n, int* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = locate(map, pts[i].x, pts[i].y);
}

inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

hipError_t launch_fill(int* p, int v, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u64, dim3(nblk(n, 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

hipError_t launch_admit(const DevMap& map, const DevAgents& ag, const DevSlots& sl,
                        const DevFrame& fr, bool ident, hipStream_t st) {
    (void)map;
    const int nbA = nblk(ag.n, kAdmitThreads), nbS = nblk(sl.P, kAdmitThreads);
    if (ident) {
        hipLaunchKernelGGL(k_admit_identity, dim3(nbA > 0 ? nbA : 1), dim3(kAdmitThreads), 0, st, ag, sl, fr);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_apply_updates, dim3(nblk(ag.n > 0 ? ag.n : 1, 256)), dim3(256), 0, st, ag, sl, fr);
    hipLaunchKernelGGL(k_admit_count, dim3(nbA + nbS), dim3(kAdmitThreads), 0, st, ag, sl, fr, nbA);
    hipLaunchKernelGGL(k_admit_finish, dim3(1), dim3(kAdmitThreads), 0, st, sl, fr, nbA, nbS);
    hipLaunchKernelGGL(k_admit_scatter, dim3(nbA + nbS), dim3(kAdmitThreads), 0, st, ag, sl, fr, nbA);
    hipLaunchKernelGGL(k_admit_slots, dim3(nblk(sl.P, 256)), dim3(256), 0, st, ag, sl, fr);
    return hipGetLastError();
}

hipError_t launch_export(const DevAgents& ag, const DevFrame& fr, hipStream_t st) {
    hipLaunchKernelGGL(k_export, dim3(nblk(ag.n > 8 ? ag.n : 8, 256)), dim3(256), 0, st, ag, fr);
    return hipGetLastError();
}

hipError_t launch_compact(const DevAgents& ag, const DevSlots& sl, const DevFrame& fr, hipStream_t st) {
    const int owners = ag.n + sl.P;
    if (ag.n <= 0) return hipSuccess;
    hipError_t e;
    hipLaunchKernelGGL(k_slab_len, dim3(nblk(owners, 256)), dim3(256), 0, st, ag, sl, fr);
    if ((e = exscan(fr.slab_len, owners, fr.slab_blk, fr.slab_len, &fr.slab_len[owners], st)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_slab_copy, dim3(nblk(owners, 4) < 2048 ? nblk(owners, 4) : 2048), dim3(256), 0, st, ag, sl,
                       fr);
    hipLaunchKernelGGL(k_slab_finish, dim3(1), dim3(64), 0, st, ag, sl, fr);
    return hipGetLastError();
}

hipError_t launch_crowd(const DevMap& map, const DevAgents& ag, const DevFrame& fr,
                        const CrowdParams& cp, hipStream_t st, bool alone, bool new_step) {
    const int nb = nblk(ag.n, 256);
    if (alone && new_step) hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, st, fr);
    if (alone)
        hipLaunchKernelGGL(k_move_key<true>, dim3(nb), dim3(256), 0, st, map, ag, fr);
    else
        hipLaunchKernelGGL(k_move_key<false>, dim3(nb), dim3(256), 0, st, map, ag, fr);
    // the scan reads and zeroes the counts; nothing in a frame with no mover
    hipError_t e;
    if ((e = exscan(fr.hcount, fr.ncell + 1, fr.hblk, fr.hstart, &fr.hstart[fr.ncell + 1], st, fr.hcount,
                    &fr.counters[5])) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_hash_place, dim3(nb), dim3(256), 0, st, ag, fr);
    hipLaunchKernelGGL(k_avoid_progress, dim3(nblk(ag.n, kAvoidThreads)), dim3(kAvoidThreads), 0, st, map, ag, fr,
                       cp);
    if (alone)
        hipLaunchKernelGGL(k_pos_out<true>, dim3(nblk(ag.n > 8 ? ag.n : 8, 256)), dim3(256), 0, st, ag, fr);
    else
        hipLaunchKernelGGL(k_pos_out<false>, dim3(nb), dim3(256), 0, st, ag, fr);
    return hipGetLastError();
}

hipError_t locate_scratch(int device, float2** pts, int** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *pts = static_cast<float2*>(addr[device]);
    *out = reinterpret_cast<int*>(static_cast<unsigned char*>(addr[device]) + (size_t)kLocateScratchMax * sizeof(float2));
    return hipSuccess;
}

hipError_t launch_locate(const DevMap& map, const float2* pts, int n, int* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_locate, dim3(nblk(n, 256)), dim3(256), 0, st, map, pts, n, out);
    return hipGetLastError();
}

}  // namespace gmapf_amd
