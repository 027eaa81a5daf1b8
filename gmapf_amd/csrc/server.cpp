// server.cpp -- serving loop over several device-resident maps (gmapf_serve.h).
//
// A client of the engine's C ABI only: each map is a gmapf_handle whose
// frame (admission, search, crowd) runs on the device; this layer keeps the
// request queue, admits requests by free search slots and by their search
// memory footprint, launches every map's frame before waiting for any (the
// frames overlap on the device: each engine has its own streams), and
// retires requests when their agent leaves the Search state.  Requests of
// one agent are serialised, so a request never supersedes another one in
// flight.  Nothing here has a reference counterpart: the reference steps one
// map per object (src/GMAPF.cu:255-314) and has no request queue.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gmapf_c.h"
#include "../../include/gmapf_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

struct ServeError : std::runtime_error {
    int code;
    ServeError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Pending {
    long long ticket;
    int map, agent;
    float x, y;
    long long submit_frame;
    Clock::time_point t_submit;
};

struct InFlight {
    long long ticket;  // -1: agent was searching before the server took over
    int map, agent, start_node;
    long long submit_frame, admit_frame;
    Clock::time_point t_submit;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct gmapf_server_s {
    std::string err;
    gmapf_server_config cfg{};
    std::vector<gmapf_handle> maps;
    std::vector<long long> foot;   // bytes per request
    std::vector<int> slots;        // search slots per map
    std::vector<int> live;         // requests in flight per map
    std::vector<std::vector<unsigned char>> busy;  // per map, per agent: in flight
    std::deque<Pending> pending;
    std::vector<InFlight> inflight;
    std::deque<gmapf_request_done> done;
    long long next_ticket = 0, frame = 0;
    long long n_submitted = 0, n_admitted = 0, n_completed = 0, n_found = 0, n_nopath = 0,
              n_rejected = 0, held_by_budget = 0;
    long long live_bytes = 0, peak_bytes = 0;
    std::vector<long long> credit;  // fair policy: per-map byte credit
    long long quantum = 0;          // fair policy: credit per admission (largest footprint)
    long long rr = 0;               // fair policy: admissions so far (rotates the first map)
    bool adopted = false, going = false;
    Clock::time_point t_frame;
    double last_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<int> node_buf;
    std::vector<size_t> launch_order;  // maps in the order their frames are launched

    void check(int rc, gmapf_handle h) {
        if (rc != GMAPF_OK) throw ServeError(rc, gmapf_last_error(h));
    }

    void create(const gmapf_handle* hs, int n, const gmapf_server_config* c) {
        if (!hs || n < 1) throw ServeError(GMAPF_E_ARG, "no maps");
        cfg = c ? *c : gmapf_server_config{};
        if (cfg.max_pending <= 0) cfg.max_pending = 1 << 22;
        if (cfg.admission_policy != GMAPF_ADMIT_FIFO && cfg.admission_policy != GMAPF_ADMIT_FAIR)
            throw ServeError(GMAPF_E_ARG, "unknown admission policy");
        for (int m = 0; m < n; m++) {
            gmapf_handle h = hs[m];
            if (!h) throw ServeError(GMAPF_E_ARG, "null map handle");
            for (int k = 0; k < m; k++)
                if (maps[k] == h) throw ServeError(GMAPF_E_ARG, "map handle given twice");
            gmapf_config gc{};
            check(gmapf_get_config(h, &gc), h);
            int V = 0;
            check(gmapf_map_size(h, &V), h);
            if (V <= 0) throw ServeError(GMAPF_E_STATE, "map " + std::to_string(m) + " is not filled");
            const long long T = gc.path_finder_threads_per_agents, Q = gc.path_finder_queue_capacity;
            maps.push_back(h);
            const long long VP = ((long long)V + 63) / 64 * 64;
            foot.push_back(T * Q * 8 + T * 4 + VP + (long long)V * 8);
            slots.push_back(gc.path_finder_parallel_agents);
            live.push_back(0);
            busy.emplace_back();
            credit.push_back(0);
            quantum = std::max(quantum, foot.back());
        }
    }

    int agents_of(int m) {
        int n = 0;
        check(gmapf_agent_count(maps[m], &n), maps[m]);
        if ((int)busy[m].size() < n) busy[m].resize(n, 0);
        return n;
    }

    long long submit(int m, int agent, float x, float y) {
        if (m < 0 || m >= (int)maps.size()) throw ServeError(GMAPF_E_ARG, "map index out of range");
        if (agent < 0 || agent >= agents_of(m)) throw ServeError(GMAPF_E_ARG, "agent id out of range");
        int node = -1;
        const float xy[2] = {x, y};
        check(gmapf_locate(maps[m], 1, xy, &node), maps[m]);
        if (node < 0)
            throw ServeError(GMAPF_E_POS, "target pos (" + std::to_string(x) + ", " + std::to_string(y) +
                                              ") is invalid");
        if (cfg.queue_memory_budget > 0 && foot[m] > cfg.queue_memory_budget)
            throw ServeError(GMAPF_E_ARG, "one request on map " + std::to_string(m) + " needs " +
                                              std::to_string(foot[m]) + " bytes, more than queue_memory_budget");
        if ((long long)pending.size() >= cfg.max_pending) {
            n_rejected++;
            throw ServeError(GMAPF_E_FULL, "pending-request queue is full");
        }
        pending.push_back(Pending{next_ticket, m, agent, x, y, frame, Clock::now()});
        n_submitted++;
        return next_ticket++;
    }

    // Agents that are searching when the server first steps keep their slots.
    void adopt() {
        if (adopted) return;
        adopted = true;
        for (int m = 0; m < (int)maps.size(); m++) {
            int n = agents_of(m);
            const int* st = nullptr;
            check(gmapf_agent_states(maps[m], &st, nullptr), maps[m]);
            for (int a = 0; a < n; a++) {
                if (st[a] != GMAPF_STATE_SEARCH) continue;
                busy[m][a] = 1;
                live[m]++;
                live_bytes += foot[m];
                inflight.push_back(InFlight{-1, m, a, -1, -1, -1, Clock::now()});
            }
        }
        peak_bytes = std::max(peak_bytes, live_bytes);
    }

    // Admission scan (see gmapf_serve.h); `take` receives admitted positions
    // in `pending`.  `commit` false: a dry run that changes no state.
    void plan(std::vector<size_t>& take, long long* held, bool commit) {
        if (cfg.admission_policy == GMAPF_ADMIT_FAIR) {
            plan_fair(take, held, commit);
            return;
        }
        std::vector<int> free_slots(maps.size());
        for (size_t m = 0; m < maps.size(); m++) free_slots[m] = slots[m] - live[m];
        long long bytes = live_bytes;
        const long long budget = cfg.queue_memory_budget;
        long long h = 0;
        for (size_t i = 0; i < pending.size(); i++) {
            const Pending& p = pending[i];
            if (busy[p.map][p.agent]) continue;  // in flight, or taken earlier in this scan
            if (free_slots[p.map] <= 0) continue;
            if (budget > 0 && bytes + foot[p.map] > budget) {
                h = (long long)(pending.size() - i);
                break;
            }
            bytes += foot[p.map];
            free_slots[p.map]--;
            busy[p.map][p.agent] = 2;
            take.push_back(i);
        }
        for (size_t k : take) busy[pending[k].map][pending[k].agent] = 0;
        if (held) *held = h;
    }

    void plan_fair(std::vector<size_t>& take, long long* held, bool commit) {
        const size_t M = maps.size();
        const long long budget = cfg.queue_memory_budget;
        std::vector<std::vector<size_t>> q(M);  // per map: positions in `pending`, arrival order
        for (size_t i = 0; i < pending.size(); i++) q[pending[i].map].push_back(i);
        std::vector<size_t> head(M, 0);
        std::vector<int> free_slots(M);
        std::vector<long long> d(credit);
        for (size_t m = 0; m < M; m++) {
            free_slots[m] = slots[m] - live[m];
            d[m] = q[m].empty() ? 0 : std::min(d[m] + quantum, 2 * quantum);
        }
        long long bytes = live_bytes;
        bool blocked = false;
        // map m's next request not in flight and not taken by this scan
        auto next = [&](size_t m) -> long long {
            while (head[m] < q[m].size() && busy[m][pending[q[m][head[m]]].agent]) head[m]++;
            return head[m] < q[m].size() ? (long long)q[m][head[m]] : -1;
        };
        auto admit = [&](size_t m, size_t i) {
            take.push_back(i);
            busy[m][pending[i].agent] = 2;
            head[m]++;
            bytes += foot[m];
            free_slots[m]--;
            d[m] -= foot[m];
        };
        const size_t start = M ? (size_t)(rr % (long long)M) : 0;
        // fair passes: turns in rotation while credit covers the next request
        for (bool progress = true; progress && !blocked;) {
            progress = false;
            for (size_t j = 0; j < M && !blocked; j++) {
                const size_t m = (start + j) % M;
                if (free_slots[m] <= 0 || (budget > 0 && d[m] < foot[m])) continue;
                const long long i = next(m);
                if (i < 0) continue;
                if (budget > 0 && bytes + foot[m] > budget) {
                    blocked = true;
                    break;
                }
                admit(m, (size_t)i);
                progress = true;
            }
        }
        // memory left: the rest in arrival order, charged to the maps' credit
        if (!blocked) {
            for (size_t i = 0; i < pending.size(); i++) {
                const size_t m = pending[i].map;
                if (busy[m][pending[i].agent] || free_slots[m] <= 0) continue;
                if (budget > 0 && bytes + foot[m] > budget) {
                    blocked = true;
                    break;
                }
                admit(m, i);
            }
        }
        for (size_t k : take) busy[pending[k].map][pending[k].agent] = 0;
        if (held) *held = blocked ? (long long)(pending.size() - take.size()) : 0;
        if (commit) {
            for (size_t m = 0; m < M; m++) credit[m] = std::max(d[m], -4 * quantum);
            rr++;
        }
    }

    void async_step(float dt) {
        if (going) return;  // like GMAPF::AsyncStep during a step (GMAPF.cu:260-261)
        t_frame = Clock::now();
        for (size_t m = 0; m < maps.size(); m++) agents_of((int)m);
        adopt();
        std::vector<size_t> take;
        plan(take, &held_by_budget, true);
        const Clock::time_point t_plan = Clock::now();
        // per map: ids and targets of this frame's admissions
        std::vector<std::vector<int>> ids(maps.size());
        std::vector<std::vector<float>> xy(maps.size());
        std::vector<std::vector<size_t>> which(maps.size());
        for (size_t k : take) {
            const Pending& p = pending[k];
            ids[p.map].push_back(p.agent);
            xy[p.map].push_back(p.x);
            xy[p.map].push_back(p.y);
            which[p.map].push_back(k);
        }
        // The maps' frames run side by side on their own streams, and each
        // launch costs tens of microseconds of host time, so the frame ends
        // with the map whose search starts last.  The map with the most
        // search memory live after this frame's admissions (the longest
        // expected search work) is admitted and launched first, each map's
        // launch right after its own admissions; the order changes no result.
        launch_order.resize(maps.size());
        for (size_t m = 0; m < maps.size(); m++) launch_order[m] = m;
        std::stable_sort(launch_order.begin(), launch_order.end(), [&](size_t a, size_t b) {
            return (long long)(live[a] + ids[a].size()) * foot[a] > (long long)(live[b] + ids[b].size()) * foot[b];
        });
        std::vector<unsigned char> gone(pending.size(), 0);
        Clock::time_point t_adm = Clock::now();
        double adm_ms = 0.0;
        for (size_t m : launch_order) {
            const Clock::time_point t_m = Clock::now();
            if (!ids[m].empty()) {
                if (cfg.record_start_nodes) {
                    int n = agents_of((int)m);
                    node_buf.assign(n, -1);
                    check(gmapf_get_agents_soa(maps[m], nullptr, nullptr, node_buf.data()), maps[m]);
                }
                std::vector<int> st(ids[m].size(), 0);
                check(gmapf_set_agent_targ_pos_many(maps[m], (int)ids[m].size(), ids[m].data(), xy[m].data(),
                                                    st.data()),
                      maps[m]);
                for (size_t j = 0; j < ids[m].size(); j++) {
                    const Pending& p = pending[which[m][j]];
                    gone[which[m][j]] = 1;
                    if (st[j] != GMAPF_OK) {  // cannot happen: targets were located at submit
                        done.push_back(gmapf_request_done{p.ticket, p.map, p.agent, 0, -1, p.submit_frame, frame,
                                                          frame, ms_between(p.t_submit, Clock::now())});
                        n_completed++;
                        n_nopath++;
                        continue;
                    }
                    busy[m][p.agent] = 1;
                    live[m]++;
                    live_bytes += foot[m];
                    n_admitted++;
                    inflight.push_back(InFlight{p.ticket, (int)m, p.agent,
                                                cfg.record_start_nodes ? node_buf[p.agent] : -1, p.submit_frame,
                                                frame, p.t_submit});
                }
            }
            t_adm = Clock::now();
            adm_ms += ms_between(t_m, t_adm);
            check(gmapf_async_step(maps[m], dt), maps[m]);
        }
        peak_bytes = std::max(peak_bytes, live_bytes);
        if (!take.empty()) {
            std::deque<Pending> keep;
            for (size_t i = 0; i < pending.size(); i++)
                if (!gone[i]) keep.push_back(pending[i]);
            pending.swap(keep);
        }
        going = true;
        // host parts: [1] admission (planning and every map's admissions),
        // [2] the launches
        const double total = ms_between(t_frame, Clock::now());
        last_ms[1] = ms_between(t_frame, t_plan) + adm_ms;
        last_ms[2] = total - last_ms[1];
    }

    void wait_step_end() {
        if (!going) return;
        going = false;
        const Clock::time_point t0 = Clock::now();
        int first_rc = GMAPF_OK;
        std::string first_msg;
        for (gmapf_handle h : maps) {  // wait for every map even if one failed
            int rc = gmapf_wait_step_end(h);
            if (rc != GMAPF_OK && first_rc == GMAPF_OK) {
                first_rc = rc;
                first_msg = gmapf_last_error(h);
            }
        }
        if (first_rc != GMAPF_OK) throw ServeError(first_rc, first_msg);
        const Clock::time_point t1 = Clock::now();
        std::vector<const int*> st(maps.size(), nullptr);
        for (size_t m = 0; m < maps.size(); m++) check(gmapf_agent_states(maps[m], &st[m], nullptr), maps[m]);
        size_t w = 0;
        for (size_t i = 0; i < inflight.size(); i++) {
            const InFlight& f = inflight[i];
            const int s = st[f.map][f.agent];
            if (s == GMAPF_STATE_SEARCH) {
                inflight[w++] = f;
                continue;
            }
            busy[f.map][f.agent] = 0;
            live[f.map]--;
            live_bytes -= foot[f.map];
            if (f.ticket < 0) continue;
            const int found = s == GMAPF_STATE_MOVE ? 1 : 0;
            done.push_back(gmapf_request_done{f.ticket, f.map, f.agent, found, f.start_node, f.submit_frame,
                                              f.admit_frame, frame, ms_between(f.t_submit, t1)});
            n_completed++;
            (found ? n_found : n_nopath)++;
        }
        inflight.resize(w);
        while ((long long)done.size() > cfg.max_pending) done.pop_front();  // never polled: keep the newest
        frame++;
        const Clock::time_point t2 = Clock::now();
        last_ms[0] = ms_between(t_frame, t2);
        last_ms[3] = ms_between(t0, t1);
        last_ms[4] = ms_between(t1, t2);
    }
};

namespace {

template <class F>
int guarded(gmapf_server s, F&& f) {
    if (!s) return GMAPF_E_ARG;
    try {
        f();
        s->err.clear();
        return GMAPF_OK;
    } catch (const ServeError& e) {
        s->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        s->err = "host allocation failed";
        return GMAPF_E_NOMEM;
    } catch (const std::exception& e) {
        s->err = e.what();
        return GMAPF_E_INTERNAL;
    }
}

}  // namespace

extern "C" {

int gmapf_server_create(gmapf_server* out, const gmapf_handle* maps, int n_maps, const gmapf_server_config* cfg) {
    if (!out) return GMAPF_E_ARG;
    *out = nullptr;
    gmapf_server s = nullptr;
    try {
        s = new gmapf_server_s();
    } catch (...) {
        return GMAPF_E_NOMEM;
    }
    int rc = guarded(s, [&] { s->create(maps, n_maps, cfg); });
    if (rc != GMAPF_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return GMAPF_OK;
}

int gmapf_server_destroy(gmapf_server s) {
    if (!s) return GMAPF_E_ARG;
    try {
        if (s->going) s->wait_step_end();
    } catch (...) {
    }
    delete s;
    return GMAPF_OK;
}

const char* gmapf_server_last_error(gmapf_server s) { return s ? s->err.c_str() : "null server"; }

int gmapf_server_submit(gmapf_server s, int map, int agent, float x, float y, long long* ticket) {
    return guarded(s, [&] {
        long long t = s->submit(map, agent, x, y);
        if (ticket) *ticket = t;
    });
}

int gmapf_server_submit_many(gmapf_server s, int n, const int* map, const int* agent, const float* xy,
                             long long* tickets, int* status) {
    return guarded(s, [&] {
        if (n < 0 || (n > 0 && (!map || !agent || !xy))) throw ServeError(GMAPF_E_ARG, "bad arrays");
        int first = GMAPF_OK;
        std::string msg;
        for (int k = 0; k < n; k++) {
            int rc = GMAPF_OK;
            long long t = -1;
            try {
                t = s->submit(map[k], agent[k], xy[2 * k], xy[2 * k + 1]);
            } catch (const ServeError& e) {
                rc = e.code;
                if (first == GMAPF_OK) {
                    first = rc;
                    msg = e.what();
                }
            }
            if (tickets) tickets[k] = t;
            if (status) status[k] = rc;
        }
        if (first != GMAPF_OK && !status) throw ServeError(first, msg);
    });
}

int gmapf_server_async_step(gmapf_server s, float dt) {
    return guarded(s, [&] { s->async_step(dt); });
}

int gmapf_server_wait_step_end(gmapf_server s) {
    return guarded(s, [&] { s->wait_step_end(); });
}

int gmapf_server_poll(gmapf_server s, gmapf_request_done* out, int cap, int* n) {
    return guarded(s, [&] {
        int k = 0;
        while (out && k < cap && !s->done.empty()) {
            out[k++] = s->done.front();
            s->done.pop_front();
        }
        if (n) *n = k;
    });
}

int gmapf_server_plan(gmapf_server s, long long* tickets, int cap, int* n) {
    return guarded(s, [&] {
        for (size_t m = 0; m < s->maps.size(); m++) s->agents_of((int)m);
        s->adopt();
        std::vector<size_t> take;
        s->plan(take, nullptr, false);
        for (int k = 0; tickets && k < cap && k < (int)take.size(); k++) tickets[k] = s->pending[take[k]].ticket;
        if (n) *n = (int)take.size();
    });
}

int gmapf_server_get_stats(gmapf_server s, long long* out16) {
    return guarded(s, [&] {
        if (!out16) throw ServeError(GMAPF_E_ARG, "null output");
        long long v[16] = {s->frame,
                           s->n_submitted,
                           s->n_admitted,
                           s->n_completed,
                           s->n_found,
                           s->n_nopath,
                           (long long)s->pending.size(),
                           (long long)s->inflight.size(),
                           s->live_bytes,
                           s->peak_bytes,
                           s->n_rejected,
                           s->held_by_budget,
                           (long long)s->maps.size(),
                           0,
                           0,
                           0};
        std::memcpy(out16, v, sizeof(v));
    });
}

int gmapf_server_last_frame_ms(gmapf_server s, double* out8) {
    return guarded(s, [&] {
        if (!out8) throw ServeError(GMAPF_E_ARG, "null output");
        for (int k = 0; k < 8; k++) out8[k] = s->last_ms[k];
    });
}

int gmapf_server_footprint(gmapf_server s, int map, long long* bytes) {
    return guarded(s, [&] {
        if (map < 0 || map >= (int)s->maps.size()) throw ServeError(GMAPF_E_ARG, "map index out of range");
        if (bytes) *bytes = s->foot[map];
    });
}

}  // extern "C"
