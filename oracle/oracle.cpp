// oracle.cpp -- TEST INFRASTRUCTURE ONLY (see oracle.h).
//
// Single-threaded CPU restatement of the reference's per-frame step with the
// canonical ordering rules:
//   * frame order F1..F4 serialises GMAPF.cu:255-314 (requests in ascending
//     agent index, moving set taken before the search, Jacobi avoidance);
//   * search rounds serialise PathFinder.h:163-250 in (lane, slot) order;
//   * queue order is (F asc, node asc) instead of CuQueue.h:38-45's FIFO ties.
// Build: -O2 -ffp-contract=off -fno-fast-math so FP32 results are the plain
// IEEE single-precision sequence the device kernels also produce.
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ---------------------------------------------------------------- navmesh
struct Map {
    int V = 0;
    std::vector<float> tri;  // 6V
    std::vector<int> nbr;    // 3V
    std::vector<float> cx, cy;
    // exact grid accelerator
    float gx0 = 0, gy0 = 0, inv_cs = 1;
    float rx0 = 0, ry0 = 0, rx1 = 0, ry1 = 0;  // region served by the grid
    int nx = 0, ny = 0;
    std::vector<int> cell_start, cell_tris, global_tris;
};

// CuNodesMap.h:272-275
inline float get_sign(float p1x, float p1y, float p2x, float p2y, float p3x, float p3y) {
    return (p1x - p3x) * (p2y - p3y) - (p2x - p3x) * (p1y - p3y);
}

// CuNodesMap.h:260-270 (boundary inclusive, NaN passes)
inline bool in_tri(const Map& m, int i, float px, float py) {
    const float* t = &m.tri[6 * (size_t)i];
    float d1 = get_sign(px, py, t[0], t[1], t[2], t[3]);
    float d2 = get_sign(px, py, t[2], t[3], t[4], t[5]);
    float d3 = get_sign(px, py, t[4], t[5], t[0], t[1]);
    bool neg = (d1 < 0) || (d2 < 0) || (d3 < 0);
    bool pos = (d1 > 0) || (d2 > 0) || (d3 > 0);
    return !(neg && pos);
}

// CuNodesMap.h:216-229 (first = min index)
int locate_brute(const Map& m, float x, float y) {
    for (int i = 0; i < m.V; i++)
        if (in_tri(m, i, x, y)) return i;
    return -1;
}

inline int cell_of(float v, float o, float inv, int n) {
    float f = floorf((v - o) * inv);
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)(n - 1)) return n - 1;
    return (int)f;
}

// Grid with per-triangle margins.  A point the FP32 sign test accepts lies
// within eps/sin(theta_min/2) of the triangle, where eps bounds the rounding
// error of get_sign divided by the edge length (4.25 * 2^-24 * D for points
// and vertices within a region of diameter D).  Triangles with a tiny inradius
// or a huge margin go to a list that is always scanned.
void build_grid(Map& m) {
    if (m.V == 0) return;
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (int i = 0; i < 3 * m.V; i++) {
        x0 = std::min(x0, (double)m.tri[2 * i]);
        x1 = std::max(x1, (double)m.tri[2 * i]);
        y0 = std::min(y0, (double)m.tri[2 * i + 1]);
        y1 = std::max(y1, (double)m.tri[2 * i + 1]);
    }
    double w = std::max(x1 - x0, 1e-3), h = std::max(y1 - y0, 1e-3);
    double cs = std::sqrt(w * h / m.V) * 1.5;
    cs = std::max(cs, std::max(w, h) / 2048.0);
    double pad = 2.0 * cs;
    double D = std::hypot(w + 2 * pad, h + 2 * pad);
    double eps = 4.25 * std::ldexp(1.0, -24) * D * 1.1;
    m.rx0 = (float)(x0 - pad);
    m.ry0 = (float)(y0 - pad);
    m.rx1 = (float)(x1 + pad);
    m.ry1 = (float)(y1 + pad);
    m.gx0 = m.rx0;
    m.gy0 = m.ry0;
    m.inv_cs = (float)(1.0 / cs);
    m.nx = (int)std::ceil((m.rx1 - m.rx0) / cs) + 1;
    m.ny = (int)std::ceil((m.ry1 - m.ry0) / cs) + 1;
    std::vector<std::vector<int>> cells((size_t)m.nx * m.ny);
    m.global_tris.clear();
    for (int i = 0; i < m.V; i++) {
        const float* t = &m.tri[6 * (size_t)i];
        double ax = t[0], ay = t[1], bx = t[2], by = t[3], qx = t[4], qy = t[5];
        double la = std::hypot(bx - qx, by - qy), lb = std::hypot(ax - qx, ay - qy),
               lc = std::hypot(ax - bx, ay - by);
        double area2 = std::fabs((bx - ax) * (qy - ay) - (qx - ax) * (by - ay));
        double per = la + lb + lc;
        double inr = per > 0 ? area2 / per : 0.0;
        // sin of half angles via the law of cosines
        auto half_sin = [](double opp, double s1, double s2) {
            if (s1 <= 0 || s2 <= 0) return 0.0;
            double c = (s1 * s1 + s2 * s2 - opp * opp) / (2 * s1 * s2);
            c = std::max(-1.0, std::min(1.0, c));
            return std::sqrt(std::max(0.0, (1 - c) / 2));
        };
        double smin = std::min({half_sin(la, lb, lc), half_sin(lb, la, lc), half_sin(lc, la, lb)});
        double margin = smin > 0 ? 1.25 * eps / smin + 1e-6 : 1e300;
        if (inr <= 4 * eps || margin > 4 * cs) {
            m.global_tris.push_back(i);
            continue;
        }
        float lo_x = std::nextafter((float)(std::min({ax, bx, qx}) - margin), -INFINITY);
        float hi_x = std::nextafter((float)(std::max({ax, bx, qx}) + margin), INFINITY);
        float lo_y = std::nextafter((float)(std::min({ay, by, qy}) - margin), -INFINITY);
        float hi_y = std::nextafter((float)(std::max({ay, by, qy}) + margin), INFINITY);
        int cx0 = cell_of(lo_x, m.gx0, m.inv_cs, m.nx), cx1 = cell_of(hi_x, m.gx0, m.inv_cs, m.nx);
        int cy0 = cell_of(lo_y, m.gy0, m.inv_cs, m.ny), cy1 = cell_of(hi_y, m.gy0, m.inv_cs, m.ny);
        for (int cy = cy0; cy <= cy1; cy++)
            for (int cx = cx0; cx <= cx1; cx++) cells[(size_t)cy * m.nx + cx].push_back(i);
    }
    m.cell_start.assign((size_t)m.nx * m.ny + 1, 0);
    for (size_t c = 0; c < cells.size(); c++) m.cell_start[c + 1] = m.cell_start[c] + (int)cells[c].size();
    m.cell_tris.resize(m.cell_start.back());
    for (size_t c = 0; c < cells.size(); c++)
        std::copy(cells[c].begin(), cells[c].end(), m.cell_tris.begin() + m.cell_start[c]);
}

int locate_grid(const Map& m, float x, float y) {
    if (m.V == 0) return -1;
    if (!(x >= m.rx0 && x <= m.rx1 && y >= m.ry0 && y <= m.ry1)) return locate_brute(m, x, y);
    int cx = cell_of(x, m.gx0, m.inv_cs, m.nx), cy = cell_of(y, m.gy0, m.inv_cs, m.ny);
    size_t c = (size_t)cy * m.nx + cx;
    int best = -1;
    for (int k = m.cell_start[c]; k < m.cell_start[c + 1]; k++) {
        int i = m.cell_tris[k];
        if (in_tri(m, i, x, y)) {
            best = i;
            break;
        }
    }
    for (int i : m.global_tris) {
        if (best >= 0 && i > best) break;
        if (in_tri(m, i, x, y)) {
            best = i;
            break;
        }
    }
    return best;
}

// V2Float.h:50-55 via CuNodesMap.h:237-243
inline float dist2(const Map& m, int a, int b) {
    float x = m.cx[a] - m.cx[b];
    float y = m.cy[a] - m.cy[b];
    return x * x + y * y;
}

// ------------------------------------------------------------------ queue
struct QItem {
    float F;
    int node;
};
inline bool item_less(const QItem& a, const QItem& b) {
    return a.F < b.F || (a.F == b.F && a.node < b.node);
}

// CuQueue.h:33-63: array sorted by descending key, pop takes the back.
// Full queue: an item not smaller than the maximum is dropped, otherwise the
// maximum is evicted.  With a total order this keeps the Q smallest items.
void q_push(std::vector<QItem>& q, int Q, const QItem& it) {
    auto pos = std::lower_bound(q.begin(), q.end(), it,
                                [](const QItem& a, const QItem& b) { return item_less(b, a); });
    if ((int)q.size() < Q) {
        q.insert(pos, it);
    } else {
        if (pos == q.begin()) return;
        q.insert(pos, it);
        q.erase(q.begin());
    }
}

// ------------------------------------------------------------ visited map
// CuVisitedMap.h:45-54 restated as a per-search open-addressing table: first
// claim wins, no reopening.
struct VisRec {
    int key;
    int prev;
    int G;
    int dir;
};
struct VisMap {
    std::vector<VisRec> tab;
    size_t count = 0, mask = 0;
    void init(size_t cap) {
        size_t n = 64;
        while (n < cap * 2) n <<= 1;
        tab.assign(n, VisRec{-1, -1, 0, 0});
        mask = n - 1;
        count = 0;
    }
    static size_t hash(int k) { return (size_t)((uint32_t)k * 2654435761u); }
    const VisRec* find(int k) const {
        size_t i = hash(k) & mask;
        while (true) {
            const VisRec& r = tab[i];
            if (r.key == k) return &r;
            if (r.key < 0) return nullptr;
            i = (i + 1) & mask;
        }
    }
    void insert(int k, int prev, int G, int dir) {
        if ((count + 1) * 2 > tab.size()) {
            std::vector<VisRec> old;
            old.swap(tab);
            tab.assign(old.size() * 2, VisRec{-1, -1, 0, 0});
            mask = tab.size() - 1;
            count = 0;
            for (auto& r : old)
                if (r.key >= 0) insert(r.key, r.prev, r.G, r.dir);
        }
        size_t i = hash(k) & mask;
        while (tab[i].key >= 0) i = (i + 1) & mask;
        tab[i] = VisRec{k, prev, G, dir};
        count++;
    }
};

// ----------------------------------------------------------------- search
struct Search {
    int agent = -1, S = -1, Tg = -1, gen = 0;
    int T = 2, Q = 16;
    float K = 1;
    std::vector<std::vector<QItem>> queues;
    VisMap vis;
    bool found = false, done = false;
    int mS = -1, mT = -1;
    long long rounds = 0, expansions = 0, discoveries = 0;
    std::vector<QItem> frontier;
    // batched form: rounds of corridor walk still to run after the meet
    long long walk_left = 0;
};

// PathFinder.h:168-189: lane 0 seeds the start, lane 1 the target.
void search_init(const Map& m, Search& s, int S, int Tg, int T, int Q, float K) {
    s.S = S;
    s.Tg = Tg;
    s.T = T;
    s.Q = Q;
    s.K = K;
    s.queues.assign(T, {});
    s.vis.init(64);
    s.found = s.done = false;
    s.mS = s.mT = -1;
    s.rounds = s.expansions = s.discoveries = 0;
    float F0 = dist2(m, S, Tg) * K;
    q_push(s.queues[0], Q, QItem{F0, S});
    q_push(s.queues[1], Q, QItem{F0, Tg});
    s.vis.insert(S, -1, 0, 1);
    s.vis.insert(Tg, -1, 0, 0);
}

// One round of PathFinder.h:191-250, serialised in (lane, slot) order.
// Returns true when the search has ended (found or exhausted).
bool search_round(const Map& m, Search& s) {
    if (s.found || s.done) return true;
    s.frontier.clear();
    for (int l = 0; l < s.T; l++) {
        auto& q = s.queues[l];
        if (q.empty()) continue;
        QItem c = q.back();
        q.pop_back();
        s.expansions++;
        const VisRec* vc = s.vis.find(c.node);
        int G = vc->G, dir = vc->dir;
        for (int k = 0; k < 3; k++) {
            int n = m.nbr[3 * (size_t)c.node + k];
            if (n < 0) break;  // PathFinder.h:200-201
            float h = dir ? dist2(m, n, s.Tg) : dist2(m, n, s.S);
            int G1 = G + 1;
            float F = (float)G1 + h * s.K;  // PathFinder.h:210-211
            const VisRec* vn = s.vis.find(n);
            if (!vn) {
                s.vis.insert(n, c.node, G1, dir);
                s.frontier.push_back(QItem{F, n});
                s.discoveries++;
            } else if (vn->dir != dir && !s.found) {  // PathFinder.h:220-228
                s.found = true;
                s.mS = dir ? c.node : n;
                s.mT = dir ? n : c.node;
            }
        }
    }
    // PathFinder.h:236-240: frontier item i goes to lane i mod T
    for (size_t i = 0; i < s.frontier.size(); i++) q_push(s.queues[i % s.T], s.Q, s.frontier[i]);
    s.rounds++;
    if (s.found) return true;
    bool any = false;
    for (auto& q : s.queues) any = any || !q.empty();
    if (!any) {
        s.done = true;
        return true;
    }
    return false;
}

// PathFinder.h:276-306: start side reversed (start excluded), then the target
// side up to and including the target.
// Batched form (no reference counterpart): the corridor is built by two parent
// steps per round on each side, from the round after the meet; the start side
// takes max(1, G(mS)) steps, the target side max(1, G(mT)) (its last step
// also writes the target).  The agent moves once both sides are done.
constexpr long long kWalkStepsPerRound = 2;
long long walk_rounds(const Search& s) {
    const long long a = std::max(1, s.vis.find(s.mS)->G), b = std::max(1, s.vis.find(s.mT)->G);
    return (std::max(a, b) + kWalkStepsPerRound - 1) / kWalkStepsPerRound;
}

void build_corridor(const Search& s, std::vector<int>& path) {
    int lenS = s.vis.find(s.mS)->G;
    int lenT = s.vis.find(s.mT)->G + 1;
    path.assign(lenS + lenT, -1);
    int idx = lenS - 1, n = s.mS;
    while (n != s.S) {
        path[idx--] = n;
        n = s.vis.find(n)->prev;
    }
    idx = lenS;
    n = s.mT;
    while (n != s.Tg) {
        path[idx++] = n;
        n = s.vis.find(n)->prev;
    }
    path[idx] = s.Tg;
}

// ------------------------------------------------------------------ world
enum { IDLE = 0, SEARCH = 1, MOVE = 2 };

struct Agent {
    float x = 0, y = 0;
    int curr = -1;
    float tx = 0, ty = 0;
    int targ = -1;
    int state = IDLE;
    int step_idx = -1, step_node = -1;
    float spx = 0, spy = 0;
    std::vector<int> path;
    bool has_path = false;
    int gen = 0;        // bumped by every set_target
    int inflight = -1;  // index into World::searches
};

struct World {
    const Map* m = nullptr;
    orc_config cfg{};
    float cd = 0, cd2 = 0, slack = 0;
    std::vector<Agent> agents;
    std::vector<Search*> searches;  // in admission order
    long long st[8] = {0};
    std::vector<int> last_order, last_keys;
    ~World() {
        for (auto* s : searches) delete s;
    }
};

int locate(const World& w, float x, float y) {
    return w.cfg.use_grid_locate ? locate_grid(*w.m, x, y) : locate_brute(*w.m, x, y);
}

// AgentsMover.h:57-70
void move_agents(World& w, const std::vector<int>& M, float step) {
    for (int i : M) {
        Agent& a = w.agents[i];
        if (a.step_idx == -1) continue;
        float dx = a.spx - a.x, dy = a.spy - a.y;
        float len = sqrtf(dx * dx + dy * dy);
        if (len <= step) {
            a.x = a.spx;
            a.y = a.spy;
        } else {
            float ux = dx / len, uy = dy / len;
            a.x = a.x + ux * step;
            a.y = a.y + uy * step;
        }
    }
}

// AgentsMover.h:96-112: one neighbour term
inline void push_term(float ax, float ay, float bx, float by, float cd, float cd2, float& sx,
                      float& sy, int& cnt) {
    float dx = ax - bx, dy = ay - by;
    float d2 = dx * dx + dy * dy;
    if (d2 < cd2) {
        float dist = sqrtf(dx * dx + dy * dy);
        float shift = cd - dist;
        sx += (dx / dist) * shift;
        sy += (dy / dist) * shift;
        cnt++;
    }
}

// AgentsMover.h:116-126
inline void finish_avoid(const World& w, float x, float y, float sx, float sy, int cnt, float& ox,
                         float& oy) {
    ox = x;
    oy = y;
    if (cnt == 0) return;
    float fc = (float)cnt;
    float ax = sx / fc, ay = sy / fc;
    float rx = ax * w.cfg.rot_cs - ay * w.cfg.rot_sn;
    float ry = ax * w.cfg.rot_sn + ay * w.cfg.rot_cs;
    float nx = x + rx, ny = y + ry;
    if (locate(w, nx, ny) >= 0) {
        ox = nx;
        oy = ny;
    }
}

// AgentsMover.h:73-129 for all agents, Jacobi, ascending neighbour index.
void avoid_brute(World& w) {
    size_t N = w.agents.size();
    std::vector<float> X(N), Y(N), OX(N), OY(N);
    for (size_t i = 0; i < N; i++) {
        X[i] = w.agents[i].x;
        Y[i] = w.agents[i].y;
    }
    for (size_t i = 0; i < N; i++) {
        float sx = 0, sy = 0;
        int cnt = 0;
        for (size_t j = 0; j < N; j++) {
            if (j == i) continue;
            push_term(X[i], Y[i], X[j], Y[j], w.cd, w.cd2, sx, sy, cnt);
        }
        finish_avoid(w, X[i], Y[i], sx, sy, cnt, OX[i], OY[i]);
    }
    for (size_t i = 0; i < N; i++) {
        w.agents[i].x = OX[i];
        w.agents[i].y = OY[i];
    }
}

// Same result through a uniform grid: integer cell keys, stable counting sort,
// stencil gather, candidates summed in ascending agent index.
void avoid_hash(World& w) {
    const Map& m = *w.m;
    size_t N = w.agents.size();
    std::vector<float> X(N), Y(N), OX(N), OY(N);
    for (size_t i = 0; i < N; i++) {
        X[i] = w.agents[i].x;
        Y[i] = w.agents[i].y;
    }
    float cs = w.cfg.hash_cell;
    float inv = 1.0f / cs;
    // grid over the navmesh bounding box; out-of-box agents clamp to border cells
    float x0 = 1e30f, y0 = 1e30f, x1 = -1e30f, y1 = -1e30f;
    for (int i = 0; i < 3 * m.V; i++) {
        x0 = std::min(x0, m.tri[2 * i]);
        x1 = std::max(x1, m.tri[2 * i]);
        y0 = std::min(y0, m.tri[2 * i + 1]);
        y1 = std::max(y1, m.tri[2 * i + 1]);
    }
    int nx = (int)((x1 - x0) * inv) + 1, ny = (int)((y1 - y0) * inv) + 1;
    int ncell = nx * ny;
    std::vector<int> key(N), cnt(ncell + 1, 0), order;
    for (size_t i = 0; i < N; i++) {
        if (std::isnan(X[i]) || std::isnan(Y[i])) {
            key[i] = -1;
            continue;
        }
        int cx = cell_of(X[i], x0, inv, nx), cy = cell_of(Y[i], y0, inv, ny);
        key[i] = cy * nx + cx;
        cnt[key[i] + 1]++;
    }
    for (int c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    order.assign(cnt[ncell], -1);
    for (size_t i = 0; i < N; i++)
        if (key[i] >= 0) order[fill[key[i]]++] = (int)i;  // stable: ascending i
    int k = (int)floorf(w.cd / cs) + 1;
    std::vector<int> cand;
    for (size_t i = 0; i < N; i++) {
        float sx = 0, sy = 0;
        int c = 0;
        if (key[i] >= 0) {
            int cx = key[i] % nx, cy = key[i] / nx;
            cand.clear();
            for (int yy = std::max(0, cy - k); yy <= std::min(ny - 1, cy + k); yy++)
                for (int xx = std::max(0, cx - k); xx <= std::min(nx - 1, cx + k); xx++) {
                    int cell = yy * nx + xx;
                    for (int p = cnt[cell]; p < cnt[cell + 1]; p++) {
                        int j = order[p];
                        if ((size_t)j == i) continue;
                        float dx = X[i] - X[j], dy = Y[i] - Y[j];
                        if (dx * dx + dy * dy < w.cd2) cand.push_back(j);
                    }
                }
            std::sort(cand.begin(), cand.end());
            for (int j : cand) push_term(X[i], Y[i], X[j], Y[j], w.cd, w.cd2, sx, sy, c);
        }
        finish_avoid(w, X[i], Y[i], sx, sy, c, OX[i], OY[i]);
    }
    for (size_t i = 0; i < N; i++) {
        w.agents[i].x = OX[i];
        w.agents[i].y = OY[i];
    }
    w.last_order = order;
    w.last_keys = key;
}

// AgentsMover.h:137-176
void progress_agents(World& w, const std::vector<int>& M) {
    const Map& m = *w.m;
    for (int i : M) {
        Agent& a = w.agents[i];
        if (a.step_idx == -1) {
            a.step_idx = 0;
            a.step_node = a.path[0];
            a.spx = m.cx[a.step_node];
            a.spy = m.cy[a.step_node];
        } else {
            if (!in_tri(m, a.step_node, a.x, a.y)) continue;
            a.curr = a.step_node;
            int count = (int)a.path.size();
            if (a.step_idx < count) a.step_idx += 1;
            if (a.step_idx == count) {
                float dx = a.tx - a.x, dy = a.ty - a.y;
                if (sqrtf(dx * dx + dy * dy) > FLT_EPSILON + w.slack) {
                    a.spx = a.tx;
                    a.spy = a.ty;
                } else {
                    a.state = IDLE;
                }
            } else {
                a.step_node = a.path[a.step_idx];
                a.spx = m.cx[a.step_node];
                a.spy = m.cy[a.step_node];
            }
        }
    }
}

void world_step(World& w, float dt) {
    const Map& m = *w.m;
    int N = (int)w.agents.size();
    // F1a: searches whose agent was given a new target since admission are dropped
    for (auto*& s : w.searches) {
        Agent& a = w.agents[s->agent];
        if (a.gen != s->gen) {
            a.inflight = -1;
            delete s;
            s = nullptr;
        }
    }
    w.searches.erase(std::remove(w.searches.begin(), w.searches.end(), nullptr), w.searches.end());
    for (size_t k = 0; k < w.searches.size(); k++) w.agents[w.searches[k]->agent].inflight = (int)k;
    // F1b: PathFinder.h:85-116 in ascending agent index
    for (int i = 0; i < N; i++) {
        Agent& a = w.agents[i];
        if (a.state != SEARCH || a.inflight >= 0) continue;
        if (a.curr == a.targ) {
            a.state = IDLE;
            continue;
        }
        if ((int)w.searches.size() >= w.cfg.parallel) continue;  // deferred
        a.path.clear();
        a.has_path = false;
        Search* s = new Search();
        s->agent = i;
        s->gen = a.gen;
        search_init(m, *s, a.curr, a.targ, w.cfg.threads, w.cfg.queue_cap, w.cfg.heur_k);
        a.inflight = (int)w.searches.size();
        w.searches.push_back(s);
        w.st[6]++;
    }
    // F2: AgentsMover.h:29-49, taken before the search writes states
    std::vector<int> M;
    for (int i = 0; i < N; i++)
        if (w.agents[i].state == MOVE) M.push_back(i);
    // F3: PathFinder.h:446-456 (+ round budget for the batched form, where a
    // found search also walks its corridor within the budget)
    const bool batched = w.cfg.round_budget > 0;
    long long budget = batched ? w.cfg.round_budget : (1LL << 62);
    for (auto*& s : w.searches) {
        bool ended = false;
        long long used = 0;
        if (s->walk_left > 0) {  // corridor walk carried over from an earlier frame
            used = std::min(s->walk_left, budget);
            s->walk_left -= used;
            if (s->walk_left > 0) continue;
            ended = true;
        } else {
            long long e0 = s->expansions, d0 = s->discoveries, r0 = s->rounds;
            for (long long r = 0; r < budget; r++) {
                used++;
                if (search_round(m, *s)) {
                    ended = true;
                    break;
                }
            }
            // work is counted in the frame it is done (also for searches
            // that are later dropped or still running)
            w.st[2] += s->expansions - e0;
            w.st[3] += s->discoveries - d0;
            w.st[5] += s->rounds - r0;
            if (!ended) continue;
            if (batched && s->found) {
                const long long need = walk_rounds(*s);
                if (need > budget - used) {
                    s->walk_left = need - (budget - used);
                    continue;
                }
            }
        }
        Agent& a = w.agents[s->agent];
        if (s->found) {  // PathFinder.h:280-290
            build_corridor(*s, a.path);
            a.has_path = true;
            a.state = MOVE;
            a.step_idx = -1;
            w.st[0]++;
            w.st[4] += (long long)a.path.size();
        } else {  // PathFinder.h:268-274
            a.state = IDLE;
            w.st[1]++;
        }
        a.inflight = -1;
        delete s;
        s = nullptr;
    }
    w.searches.erase(std::remove(w.searches.begin(), w.searches.end(), nullptr), w.searches.end());
    for (size_t k = 0; k < w.searches.size(); k++) w.agents[w.searches[k]->agent].inflight = (int)k;
    // F4: AgentsMover.h:294-308
    if (!M.empty()) {
        move_agents(w, M, w.cfg.speed * dt);
        if (w.cfg.hash_cell > 0)
            avoid_hash(w);
        else
            avoid_brute(w);
        progress_agents(w, M);
    }
    w.st[7]++;
}

}  // namespace

// ============================================================== C interface
struct orc_map {
    Map m;
};
struct orc_world {
    World w;
};

extern "C" {

orc_map* orc_map_create(int V, const float* tri, const int* nbr) {
    auto* om = new orc_map();
    Map& m = om->m;
    m.V = V;
    m.tri.assign(tri, tri + 6 * (size_t)V);
    m.nbr.assign(nbr, nbr + 3 * (size_t)V);
    m.cx.resize(V);
    m.cy.resize(V);
    for (int i = 0; i < V; i++) {  // GMAPF.cu:215
        const float* t = &m.tri[6 * (size_t)i];
        m.cx[i] = ((t[0] + t[2]) + t[4]) / 3.0f;
        m.cy[i] = ((t[1] + t[3]) + t[5]) / 3.0f;
    }
    build_grid(m);
    return om;
}

void orc_map_destroy(orc_map* m) { delete m; }
int orc_map_count(const orc_map* m) { return m->m.V; }
void orc_map_centroids(const orc_map* om, float* cxy) {
    for (int i = 0; i < om->m.V; i++) {
        cxy[2 * i] = om->m.cx[i];
        cxy[2 * i + 1] = om->m.cy[i];
    }
}
int orc_locate_brute(const orc_map* m, float x, float y) { return locate_brute(m->m, x, y); }
int orc_locate_grid(const orc_map* m, float x, float y) { return locate_grid(m->m, x, y); }
void orc_locate_many(const orc_map* m, int n, const float* xy, int* out, int use_grid) {
    for (int i = 0; i < n; i++)
        out[i] = use_grid ? locate_grid(m->m, xy[2 * i], xy[2 * i + 1])
                          : locate_brute(m->m, xy[2 * i], xy[2 * i + 1]);
}

void orc_queue_script(int Q, int nops, const int* ops, const float* F, const int* node, int* out) {
    std::vector<QItem> q;
    for (int i = 0; i < nops; i++) {
        out[i] = -1;
        if (ops[i] == 0) {
            q_push(q, Q, QItem{F[i], node[i]});
        } else if (!q.empty()) {
            out[i] = q.back().node;
            q.pop_back();
        }
    }
}

int orc_search(const orc_map* om, int S, int Tg, int T, int Q, float K, int* buf, int cap,
               int* len, long long* stats) {
    const Map& m = om->m;
    Search s;
    search_init(m, s, S, Tg, T, Q, K);
    while (!search_round(m, s)) {
    }
    std::vector<int> path;
    if (s.found) build_corridor(s, path);
    if (len) *len = (int)path.size();
    for (int i = 0; i < (int)path.size() && i < cap; i++) buf[i] = path[i];
    if (stats) {
        stats[0] = s.rounds;
        stats[1] = s.expansions;
        stats[2] = s.discoveries;
        stats[3] = (long long)path.size();
        stats[4] = s.mS;
        stats[5] = s.mT;
    }
    return s.found ? 1 : 0;
}

orc_world* orc_world_create(const orc_map* m, const orc_config* cfg) {
    auto* ow = new orc_world();
    World& w = ow->w;
    w.m = &m->m;
    w.cfg = *cfg;
    w.cd = cfg->radius * 2.5f;  // AgentsMover.h:222-224
    w.cd2 = w.cd * w.cd;
    w.slack = cfg->radius;
    return ow;
}
void orc_world_destroy(orc_world* w) { delete w; }

// GMAPF.cu:219-236
int orc_world_add_agent(orc_world* ow, float x, float y) {
    World& w = ow->w;
    if ((int)w.agents.size() >= w.cfg.agents_max) return -1;
    int node = locate(w, x, y);
    if (node < 0) return -1;
    Agent a;
    a.x = x;
    a.y = y;
    a.curr = node;
    a.state = IDLE;
    w.agents.push_back(a);
    return (int)w.agents.size() - 1;
}

// GMAPF.cu:238-253
int orc_world_set_target(orc_world* ow, int id, float x, float y) {
    World& w = ow->w;
    if (id < 0 || id >= (int)w.agents.size()) return -1;
    int node = locate(w, x, y);
    if (node < 0) return -1;
    Agent& a = w.agents[id];
    a.state = SEARCH;
    a.tx = x;
    a.ty = y;
    a.targ = node;
    a.gen++;
    return 0;
}

void orc_world_step(orc_world* ow, float dt) { world_step(ow->w, dt); }
int orc_world_count(const orc_world* ow) { return (int)ow->w.agents.size(); }

void orc_world_get(const orc_world* ow, float* xy, int* st, int* node) {
    const World& w = ow->w;
    for (size_t i = 0; i < w.agents.size(); i++) {
        if (xy) {
            xy[2 * i] = w.agents[i].x;
            xy[2 * i + 1] = w.agents[i].y;
        }
        if (st) st[i] = w.agents[i].state;
        if (node) node[i] = w.agents[i].curr;
    }
}

int orc_world_corridor(const orc_world* ow, int id, int* buf, int cap) {
    const World& w = ow->w;
    if (id < 0 || id >= (int)w.agents.size()) return -1;
    const Agent& a = w.agents[id];
    if (!a.has_path) return -1;
    for (int i = 0; i < (int)a.path.size() && i < cap; i++) buf[i] = a.path[i];
    return (int)a.path.size();
}

void orc_world_get_steps(const orc_world* ow, int* step_idx, int* step_node, float* step_xy) {
    const World& w = ow->w;
    for (size_t i = 0; i < w.agents.size(); i++) {
        step_idx[i] = w.agents[i].step_idx;
        step_node[i] = w.agents[i].step_node;
        step_xy[2 * i] = w.agents[i].spx;
        step_xy[2 * i + 1] = w.agents[i].spy;
    }
}

void orc_world_stats(const orc_world* ow, long long* out8) {
    for (int i = 0; i < 8; i++) out8[i] = ow->w.st[i];
}

int orc_world_last_sort(const orc_world* ow, int* order, int* keys) {
    const World& w = ow->w;
    if (order)
        for (size_t i = 0; i < w.last_order.size(); i++) order[i] = w.last_order[i];
    if (keys)
        for (size_t i = 0; i < w.last_keys.size(); i++) keys[i] = w.last_keys[i];
    return (int)w.last_order.size();
}

}  // extern "C"
