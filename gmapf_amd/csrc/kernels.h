// kernels.h -- host-callable launchers for the per-frame HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "device.h"

namespace gmapf_amd {

// k_search.hip
int search_lds_bytes(int Q);
// F3: search + corridor build + agent state for ended searches (PathFinder.h:138-307)
hipError_t launch_search(const DevMap& map, const DevAgents& ag, const DevSlots& slots, const DevFrame& fr,
                         const SearchParams& sp, int grid, int persistent, hipStream_t st);
hipError_t search_occupancy(int T, int Q, int* blocks_per_cu);
// sets the dynamic LDS limit of the search kernel for (T, Q); call before launching/capturing
hipError_t search_prepare(int T, int Q);

// k_frame.hip
// Fill n ints with v (used instead of hipMemsetAsync so frames can be captured).
hipError_t launch_fill(int* p, int v, long long n, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, unsigned long long v, long long n, hipStream_t st);

// F0 + F1 + F2: target updates from pinned host memory, then cancel stale
// searches, admit requests in ascending agent index into free slots in
// ascending slot index, moving set, active list (one workgroup).
hipError_t launch_admit(const DevMap& map, const DevAgents& ag, const DevSlots& slots, const DevFrame& fr,
                        bool ident, hipStream_t st);
// positions, states and counters into pinned host memory
hipError_t launch_export(const DevAgents& ag, const DevFrame& fr, hipStream_t st);
// corridor slab compaction: corridors of moving agents into the other buffer
// in agent order; the others are released
hipError_t launch_compact(const DevAgents& ag, const DevSlots& sl, const DevFrame& fr, hipStream_t st);
// F4: move, hash + avoidance (Jacobi), progress; all no-ops when the moving
// set of the frame is empty.  alone: the frame has no admission and no search
// (nobody searches, no target changed, some agent moves): the first kernel
// takes the frame's moving set and the last one also exports; new_step (with
// alone): the step differs from the last frame's and is copied first.
hipError_t launch_crowd(const DevMap& map, const DevAgents& ag, const DevFrame& fr,
                        const CrowdParams& cp, hipStream_t st, bool alone = false, bool new_step = false);

// bulk point location (min containing triangle or -1)
hipError_t launch_locate(const DevMap& map, const float2* pts, int n, int* out, hipStream_t st);
// Points a call of the allocation-free bulk location can stage at once.
constexpr int kLocateScratchMax = 1 << 20;
// Device addresses of this library's location staging on `device` (points,
// then results; kLocateScratchMax of each): looked up once per device, then
// cached.  Calls that overlap on one device share the staging and race.
hipError_t locate_scratch(int device, float2** pts, int** out);

}  // namespace gmapf_amd
