// k_search.hip -- bidirectional multi-lane bounded-queue search on gfx950.
//
// Semantics (reference: include/gmapf/PathFinder.h:138-251, misc/CuQueue.h:33-70,
// misc/CuVisitedMap.h:45-54, misc/CuNodesMap.h:237-243):
//   * a search owns T lanes; lane 0 is seeded with the start node (dir 1),
//     lane 1 with the target node (dir 0), both with F = dist2(S,Tg)*K;
//   * a round: every lane pops the minimum of its own queue and expands up to
//     three neighbours (compacted list, stops at the first -1) with
//     h = dist2(centroid(n), centroid(opposite endpoint)), G' = G+1,
//     F' = (float)G' + h*K;
//   * claims are resolved in (lane, slot) order: the first candidate that
//     reaches an unvisited node claims it and appends it to the frontier; a
//     candidate that reaches a node owned by the other direction is a meet,
//     and the first meet in (lane, slot) order wins;
//   * frontier item i is pushed to the queue of lane i mod T; queues keep the
//     Q smallest items by (F, node);
//   * the search ends when a meet was found or every queue is empty.
// The reference resolves claims and meets with atomics in arbitrary order;
// here the order is fixed, so results are deterministic and equal the
// oracle's serial restatement.
//
// Mapping: one workgroup = one wave64 = 64/T searches.  Queues live in LDS
// (entry i of lane L at q[i*64+L], 64-bit keys (F,node) that compare as
// unsigned integers).  Claim prefix sums and the first-meet choice are wave
// ballots + popcounts; duplicate claims inside a round are found through a
// small LDS candidate table.  The visited table is per slot, stamp-tagged.
#include <hip/hip_runtime.h>

#include "device.h"
#include "kernels.h"

namespace gmapf_amd {

namespace {

constexpr int kWave = 64;

__device__ __forceinline__ unsigned long long make_key(float F, int node) {
    F = F + 0.0f;  // -0 -> +0 so equal floats give equal keys
    unsigned b = __float_as_uint(F);
    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((unsigned long long)b << 32) | (unsigned)node;
}

__device__ __forceinline__ int key_node(unsigned long long k) { return (int)(unsigned)(k & 0xffffffffull); }

__device__ __forceinline__ float dist2(float2 a, float2 b) {
    float x = a.x - b.x;
    float y = a.y - b.y;
    return x * x + y * y;
}

// Bounded queue stored descending (q[0] = max, q[cnt-1] = min), CuQueue.h:33-70
// with (F, node) ordering.
__device__ __forceinline__ void q_push(unsigned long long* q, int L, int& cnt, int Q,
                                       unsigned long long k) {
    if (cnt < Q) {
        int i = cnt - 1;
        while (i >= 0) {
            unsigned long long v = q[i * kWave + L];
            if (!(v < k)) break;
            q[(i + 1) * kWave + L] = v;
            --i;
        }
        q[(i + 1) * kWave + L] = k;
        ++cnt;
    } else {
        int p = 0;
        while (p < cnt && q[p * kWave + L] > k) ++p;
        if (p == 0) return;  // not smaller than the maximum: dropped
        for (int i = 0; i < p - 1; ++i) q[i * kWave + L] = q[(i + 1) * kWave + L];
        q[(p - 1) * kWave + L] = k;
    }
}

template <int T>
__device__ __forceinline__ unsigned long long seg_mask(int sub) {
    if constexpr (T == 64) {
        return ~0ull;
    } else {
        return ((1ull << T) - 1ull) << (sub * T);
    }
}

template <int T>
__global__ __launch_bounds__(64) void k_search(DevMap map, DevSlots slots, DevFrame fr,
                                               SearchParams sp, int persistent) {
    extern __shared__ unsigned long long q[];  // 64 * Q keys
    __shared__ int cand[kWave * 3];
    __shared__ unsigned long long stage[kWave * 3];

    const int L = threadIdx.x;
    const int sub = L / T;
    const int l = L % T;
    constexpr int RPW = kWave / T;
    const int segbase = sub * 3 * T;
    const unsigned long long smask = seg_mask<T>(sub);
    const unsigned long long below = smask & ((1ull << L) - 1ull);
    const int Q = sp.Q;
    const size_t V = (size_t)map.V;
    const int n_active = fr.counters[1];

    // per-segment request state (replicated in every lane of the segment)
    int slot = -1;
    bool has_req = false, no_more = false;
    int S = 0, Tg = 0, cnt = 0;
    unsigned stamp = 0;
    float2 cS = make_float2(0.f, 0.f), cT = make_float2(0.f, 0.f);
    int rounds_here = 0;
    long long rounds_tot = 0, exp_here = 0, disc_here = 0;
    int status = 0;
    bool took_plain = false;

    while (true) {
        // ---- acquire work for segments that have none
        if (!has_req && !no_more) {
            int item = -1;
            if (persistent) {
                if (l == 0) item = atomicAdd(&fr.counters[2], 1);
                item = __shfl(item, sub * T, kWave);
            } else if (!took_plain) {
                item = blockIdx.x * RPW + sub;
                took_plain = true;
            }
            if (item < 0 || item >= n_active) {
                no_more = true;
            } else {
                slot = fr.active[item];
                S = slots.S[slot];
                Tg = slots.Tg[slot];
                stamp = slots.stamp[slot];
                status = slots.status[slot];
                rounds_tot = slots.rounds[slot];
                cS = map.cen[S];
                cT = map.cen[Tg];
                rounds_here = 0;
                exp_here = 0;
                disc_here = 0;
                uint4* vis = slots.vis + (size_t)slot * V;
                if (status == 3) {
                    // PathFinder.h:168-189
                    float F0 = dist2(cS, cT) * sp.K;
                    cnt = 0;
                    if (l == 0) {
                        q[L] = make_key(F0, S);
                        cnt = 1;
                        vis[S] = make_uint4(stamp, 0xffffffffu, 0u, 1u);
                    }
                    if (l == 1) {
                        q[L] = make_key(F0, Tg);
                        cnt = 1;
                        vis[Tg] = make_uint4(stamp, 0xffffffffu, 0u, 0u);
                    }
                    status = 0;
                } else {
                    size_t qb = ((size_t)slot * T + l);
                    cnt = slots.qcnt[qb];
                    const unsigned long long* src = slots.qsave + qb * (size_t)Q;
                    for (int i = 0; i < cnt; ++i) q[i * kWave + L] = src[i];
                }
                has_req = true;
            }
        }
        if (__ballot(has_req) == 0ull) break;
        __syncthreads();  // seeds / restored queues and visited writes visible

        // ---- one round
        uint4* vis = slots.vis + (size_t)slot * V;
        const bool popped = has_req && cnt > 0;
        int c = -1, G = 0, cdir = 0;
        int n0 = -1, n1 = -1, n2 = -1;
        if (popped) {
            --cnt;
            c = key_node(q[cnt * kWave + L]);
            int4 nb = map.nbr[c];
            uint4 vc = vis[c];
            G = (int)vc.z;
            cdir = (int)vc.w;
            n0 = nb.x;
            n1 = n0 >= 0 ? nb.y : -1;
            n2 = n1 >= 0 ? nb.z : -1;
        }
        int nn[3] = {n0, n1, n2};
        float Fn[3];
        bool prior[3];
        int pdir[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            Fn[s] = 0.f;
            prior[s] = false;
            pdir[s] = 0;
            if (nn[s] >= 0) {
                float2 cn = map.cen[nn[s]];
                uint4 vn = vis[nn[s]];
                float h = cdir ? dist2(cn, cT) : dist2(cn, cS);
                Fn[s] = (float)(G + 1) + h * sp.K;
                prior[s] = (vn.x == stamp);
                pdir[s] = (int)vn.w;
            }
        }
#pragma unroll
        for (int s = 0; s < 3; ++s) cand[segbase + l * 3 + s] = nn[s] >= 0 ? (nn[s] * 2 + cdir) : -1;
        __syncthreads();

        bool claim[3], meet[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            claim[s] = false;
            meet[s] = false;
            if (nn[s] < 0) continue;
            int first = -1;
            const int kk = l * 3 + s;
            for (int k = 0; k < kk; ++k) {
                int cv = cand[segbase + k];
                if (cv >= 0 && (cv >> 1) == nn[s]) {
                    first = cv;
                    break;
                }
            }
            if (!prior[s] && first < 0) {
                claim[s] = true;
            } else {
                int odir = prior[s] ? pdir[s] : (first & 1);
                meet[s] = (odir != cdir);
            }
        }

        // claims: frontier order = (lane, slot) order
        const int cl = (int)claim[0] + (int)claim[1] + (int)claim[2];
        const unsigned long long b0 = __ballot(cl & 1);
        const unsigned long long b1 = __ballot(cl & 2);
        const int pre = __popcll(b0 & below) + 2 * __popcll(b1 & below);
        const int tot = __popcll(b0 & smask) + 2 * __popcll(b1 & smask);
        {
            int f = pre;
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                if (claim[s]) {
                    stage[segbase + f] = make_key(Fn[s], nn[s]);
                    vis[nn[s]] = make_uint4(stamp, (unsigned)c, (unsigned)(G + 1), (unsigned)cdir);
                    ++f;
                }
            }
        }
        // first meet in (lane, slot) order
        const int ms = meet[0] ? 0 : (meet[1] ? 1 : (meet[2] ? 2 : 3));
        const unsigned long long mb = __ballot(ms < 3) & smask;
        const bool found = has_req && (mb != 0ull);
        if (found && L == __ffsll((long long)mb) - 1) {
            int n = nn[ms];
            slots.mS[slot] = cdir ? c : n;
            slots.mT[slot] = cdir ? n : c;
        }
        const unsigned long long pb = __ballot(popped) & smask;
        __syncthreads();

        // deal frontier item f to lane f mod T
        if (has_req) {
            for (int f = l; f < tot; f += T) q_push(q, L, cnt, Q, stage[segbase + f]);
        }
        const unsigned long long nb = __ballot(cnt > 0) & smask;
        if (has_req) {
            ++rounds_here;
            ++rounds_tot;
            exp_here += __popcll(pb);
            disc_here += tot;
            if (found) {
                status = 1;
            } else if (nb == 0ull) {
                status = 2;
            }
            const bool ended = status != 0;
            const bool park = !ended && sp.budget > 0 && rounds_here >= sp.budget;
            if (ended || park) {
                if (park) {
                    size_t qb = ((size_t)slot * T + l);
                    slots.qcnt[qb] = cnt;
                    unsigned long long* dst = slots.qsave + qb * (size_t)Q;
                    for (int i = 0; i < cnt; ++i) dst[i] = q[i * kWave + L];
                }
                if (l == 0) {
                    slots.status[slot] = status;
                    slots.rounds[slot] = rounds_tot;
                    slots.expansions[slot] += exp_here;
                    slots.discoveries[slot] += disc_here;
                    atomicAdd(&fr.stats[2], (unsigned long long)exp_here);
                    atomicAdd(&fr.stats[3], (unsigned long long)disc_here);
                    atomicAdd(&fr.stats[5], (unsigned long long)rounds_here);
                }
                has_req = false;
            }
        }
    }
}

}  // namespace

int search_lds_bytes(int Q) { return kWave * Q * (int)sizeof(unsigned long long); }

namespace {
const void* search_fn(int T) {
    switch (T) {
        case 2: return (const void*)k_search<2>;
        case 4: return (const void*)k_search<4>;
        case 8: return (const void*)k_search<8>;
        case 16: return (const void*)k_search<16>;
        case 32: return (const void*)k_search<32>;
        case 64: return (const void*)k_search<64>;
        default: return nullptr;
    }
}
}  // namespace

hipError_t search_prepare(int T, int Q) {
    const void* fn = search_fn(T);
    if (!fn) return hipErrorInvalidValue;
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, search_lds_bytes(Q));
}

hipError_t launch_search(const DevMap& map, const DevSlots& slots, const DevFrame& fr,
                         const SearchParams& sp, int grid, int persistent, hipStream_t st) {
    size_t lds = (size_t)search_lds_bytes(sp.Q);
#define GMAPF_SEARCH_CASE(TT)                                                                      \
    case TT:                                                                                       \
        hipLaunchKernelGGL(k_search<TT>, dim3(grid), dim3(kWave), lds, st, map, slots, fr, sp,     \
                           persistent);                                                            \
        return hipGetLastError();
    switch (sp.T) {
        GMAPF_SEARCH_CASE(2)
        GMAPF_SEARCH_CASE(4)
        GMAPF_SEARCH_CASE(8)
        GMAPF_SEARCH_CASE(16)
        GMAPF_SEARCH_CASE(32)
        GMAPF_SEARCH_CASE(64)
        default:
            return hipErrorInvalidValue;
    }
#undef GMAPF_SEARCH_CASE
}

hipError_t search_occupancy(int T, int Q, int* blocks_per_cu) {
    hipError_t e = search_prepare(T, Q);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, search_fn(T), kWave,
                                                        (size_t)search_lds_bytes(Q));
}

}  // namespace gmapf_amd
