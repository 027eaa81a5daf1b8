// cdt.cpp -- constrained Delaunay triangulation for FillMapEnd.
//
// The reference triangulates with a third-party CDT library configured as
// Triangulation<float>(Auto order, TryResolve intersections, 1.0 tolerance)
// followed by eraseOuterTrianglesAndHoles (src/GMAPF.cu:161-196).  This file
// implements the same contract with its own algorithms:
//   * exact-duplicate vertices merged, edges remapped (first occurrence kept);
//   * Delaunay triangulation by incremental insertion in Hilbert order with
//     Lawson flips, inside a super triangle;
//   * constraint edges recovered by flipping the edges they cross
//     (Sloan 1993); an edge through an existing vertex (exactly collinear, or
//     within snap_dist of the line) is split at that vertex; an edge crossing
//     an earlier constraint splits both at their intersection (rounded to
//     FP32);
//   * every constraint insertion is counted; the depth of a triangle is the
//     minimum over paths from the outside of the summed counts of crossed
//     constraints, and triangles of even depth (outside the boundary, inside
//     obstacles, overlaps) are removed;
//   * output triangles are CCW, in creation order; neighbour lists are
//     compacted with a -1 tail (GMAPF.cu:185-190).
// Orientation tests are exact for FP32 inputs (two-product with fma);
// incircle tests use long double, which is exact for coordinates on coarse
// grids and only affects triangle quality otherwise.
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <deque>
#include <stdexcept>
#include <unordered_map>
#include <vector>

#include "navmesh.h"

namespace gmapf_amd {

namespace {

struct P2 {
    double x, y;
};

// exact sign of (a.x-c.x)*(b.y-c.y) - (a.y-c.y)*(b.x-c.x) for FP32-valued inputs
int orient(const P2& a, const P2& b, const P2& c) {
    double acx = a.x - c.x, bcy = b.y - c.y, acy = a.y - c.y, bcx = b.x - c.x;
    double p = acx * bcy, q = acy * bcx;
    if (p > q) return 1;
    if (p < q) return -1;
    double pl = std::fma(acx, bcy, -p), ql = std::fma(acy, bcx, -q);
    if (pl > ql) return 1;
    if (pl < ql) return -1;
    return 0;
}

double orient_val(const P2& a, const P2& b, const P2& c) {
    return (a.x - c.x) * (b.y - c.y) - (a.y - c.y) * (b.x - c.x);
}

// > 0 when d is inside the circumcircle of CCW triangle (a, b, c)
long double incircle(const P2& a, const P2& b, const P2& c, const P2& d) {
    long double adx = (long double)a.x - d.x, ady = (long double)a.y - d.y;
    long double bdx = (long double)b.x - d.x, bdy = (long double)b.y - d.y;
    long double cdx = (long double)c.x - d.x, cdy = (long double)c.y - d.y;
    long double ad = adx * adx + ady * ady, bd = bdx * bdx + bdy * bdy, cd = cdx * cdx + cdy * cdy;
    return adx * (bdy * cd - bd * cdy) - ady * (bdx * cd - bd * cdx) + ad * (bdx * cdy - bdy * cdx);
}

inline uint64_t ekey(int a, int b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

uint64_t hilbert(uint32_t x, uint32_t y, int order) {
    uint64_t d = 0;
    for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
        uint32_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - (x & (s - 1));
                y = s - 1 - (y & (s - 1));
            }
            std::swap(x, y);
        }
    }
    return d;
}

struct Tri {
    int v[3];
    int n[3];  // n[i]: triangle across the edge opposite v[i]
    bool alive;
};

class CDT {
   public:
    std::vector<P2> pts;
    std::vector<Tri> tris;
    std::vector<int> vtri;                       // one incident triangle per vertex
    std::unordered_map<uint64_t, int> fixed;     // constraint edge -> insertion count
    double snap = 0;
    int nsuper = 0;
    int last = 0;

    int add_point(double x, double y) {
        pts.push_back(P2{x, y});
        vtri.push_back(-1);
        return (int)pts.size() - 1;
    }

    int new_tri(int a, int b, int c, int na, int nb, int nc) {
        tris.push_back(Tri{{a, b, c}, {na, nb, nc}, true});
        int t = (int)tris.size() - 1;
        vtri[a] = vtri[b] = vtri[c] = t;
        return t;
    }

    void set_tri(int t, int a, int b, int c, int na, int nb, int nc) {
        tris[t] = Tri{{a, b, c}, {na, nb, nc}, true};
        vtri[a] = vtri[b] = vtri[c] = t;
    }

    int idx_of(const Tri& t, int v) const {
        for (int k = 0; k < 3; k++)
            if (t.v[k] == v) return k;
        return -1;
    }

    int nbr_idx(const Tri& t, int u) const {
        for (int k = 0; k < 3; k++)
            if (t.n[k] == u) return k;
        return -1;
    }

    void replace_nbr(int t, int old_n, int new_n) {
        if (t < 0) return;
        int k = nbr_idx(tris[t], old_n);
        if (k < 0) throw std::runtime_error("triangulation: broken adjacency");
        tris[t].n[k] = new_n;
    }

    bool is_fixed(int a, int b) const { return fixed.count(ekey(a, b)) != 0; }

    // walk to the triangle containing p; returns tri and sets loc:
    // 0 inside, 1 on edge (edge index in *ei), 2 on vertex (*ei = vertex)
    int locate(const P2& p, int* loc, int* ei) {
        int t = (last >= 0 && last < (int)tris.size() && tris[last].alive) ? last : -1;
        if (t < 0)
            for (int k = (int)tris.size() - 1; k >= 0; k--)
                if (tris[k].alive) {
                    t = k;
                    break;
                }
        uint32_t rng = 2463534242u;
        for (size_t steps = 0; steps < 4 * tris.size() + 64; steps++) {
            const Tri& T = tris[t];
            rng ^= rng << 13;
            rng ^= rng >> 17;
            rng ^= rng << 5;
            int s = (int)(rng % 3);
            int moved = -1;
            int zero = 0, zk = -1;
            for (int j = 0; j < 3; j++) {
                int k = (s + j) % 3;
                int o = orient(pts[T.v[(k + 1) % 3]], pts[T.v[(k + 2) % 3]], p);
                if (o < 0) {
                    moved = T.n[k];
                    break;
                }
                if (o == 0) {
                    zero++;
                    zk = k;
                }
            }
            if (moved >= 0) {
                t = moved;
                continue;
            }
            if (moved == -1) {
                // no outward edge: inside or on boundary
                last = t;
                if (zero == 0) {
                    *loc = 0;
                    return t;
                }
                for (int k = 0; k < 3; k++)
                    if (pts[T.v[k]].x == p.x && pts[T.v[k]].y == p.y) {
                        *loc = 2;
                        *ei = T.v[k];
                        return t;
                    }
                *loc = 1;
                *ei = zk;
                return t;
            }
        }
        throw std::runtime_error("triangulation: point location did not converge");
    }

    void legalize(int t, int k, std::vector<std::pair<int, int>>& stack) {
        stack.clear();
        stack.push_back({t, k});
        while (!stack.empty()) {
            auto [ti, ki] = stack.back();
            stack.pop_back();
            if (!tris[ti].alive) continue;
            Tri& T = tris[ti];
            int u = T.n[ki];
            if (u < 0) continue;
            int p1 = T.v[(ki + 1) % 3], p2 = T.v[(ki + 2) % 3];
            if (is_fixed(p1, p2)) continue;
            int a = T.v[ki];
            const Tri& U = tris[u];
            int j = nbr_idx(U, ti);
            int q = U.v[j];
            if (incircle(pts[T.v[0]], pts[T.v[1]], pts[T.v[2]], pts[q]) <= 0) continue;
            if (!(orient(pts[a], pts[q], pts[p1]) * orient(pts[a], pts[q], pts[p2]) < 0)) continue;
            flip(ti, ki);
            // after flip: ti = (a, p1, q), u = (q, p2, a); new outer edges
            stack.push_back({ti, idx_of(tris[ti], a)});
            stack.push_back({u, idx_of(tris[u], a)});
        }
    }

    // flip the edge opposite v[i] of triangle t.  t=(a,p1,p2), u=(q,p2,p1) ->
    // t=(a,p1,q), u=(q,p2,a)
    void flip(int t, int i) {
        Tri T = tris[t];
        int u = T.n[i];
        Tri U = tris[u];
        int j = nbr_idx(U, t);
        int a = T.v[i], p1 = T.v[(i + 1) % 3], p2 = T.v[(i + 2) % 3];
        int q = U.v[j];
        int N_ap1 = T.n[(i + 2) % 3], N_p2a = T.n[(i + 1) % 3];
        int N_p1q = U.n[(j + 1) % 3], N_qp2 = U.n[(j + 2) % 3];
        set_tri(t, a, p1, q, N_p1q, u, N_ap1);
        set_tri(u, q, p2, a, N_p2a, t, N_qp2);
        replace_nbr(N_p1q, u, t);
        replace_nbr(N_p2a, t, u);
    }

    // insert point index p; returns the vertex index actually used
    int insert_point(int p, std::vector<std::pair<int, int>>& stack) {
        int loc = 0, e = 0;
        int t = locate(pts[p], &loc, &e);
        if (loc == 2) return e;
        if (loc == 0) {
            Tri T = tris[t];
            int a = T.v[0], b = T.v[1], c = T.v[2];
            int na = T.n[0], nb = T.n[1], nc = T.n[2];
            int t1 = new_tri(b, c, p, -1, -1, -1);
            int t2 = new_tri(c, a, p, -1, -1, -1);
            // t = (a, b, p): opp a -> t1 (edge b,p), opp b -> t2 (edge p,a), opp p -> nc (edge a,b)
            set_tri(t, a, b, p, t1, t2, nc);
            tris[t1] = Tri{{b, c, p}, {t2, t, na}, true};
            tris[t2] = Tri{{c, a, p}, {t, t1, nb}, true};
            vtri[b] = t1;
            vtri[c] = t2;
            replace_nbr(na, t, t1);
            replace_nbr(nb, t, t2);
            legalize(t, 2, stack);
            legalize(t1, 2, stack);
            legalize(t2, 2, stack);
            return p;
        }
        // on edge opposite v[e] of t
        Tri T = tris[t];
        int a = T.v[e], p1 = T.v[(e + 1) % 3], p2 = T.v[(e + 2) % 3];
        int u = T.n[e];
        int fixed_count = 0;
        auto it = fixed.find(ekey(p1, p2));
        if (it != fixed.end()) {
            fixed_count = it->second;
            fixed.erase(it);
        }
        int N_ap1 = T.n[(e + 2) % 3], N_p2a = T.n[(e + 1) % 3];
        if (u < 0) throw std::runtime_error("triangulation: point on the outer hull");
        Tri U = tris[u];
        int j = nbr_idx(U, t);
        int q = U.v[j];
        int N_p1q = U.n[(j + 1) % 3], N_qp2 = U.n[(j + 2) % 3];
        // t=(a,p1,p), t3=(a,p,p2), u=(q,p2,p), t4=(q,p,p1)
        int t3 = new_tri(a, p, p2, -1, -1, -1);
        int t4 = new_tri(q, p, p1, -1, -1, -1);
        set_tri(t, a, p1, p, t4, t3, N_ap1);
        set_tri(u, q, p2, p, t3, t4, N_qp2);
        tris[t3] = Tri{{a, p, p2}, {u, N_p2a, t}, true};
        tris[t4] = Tri{{q, p, p1}, {t, N_p1q, u}, true};
        vtri[a] = t3;
        vtri[p2] = t3;
        vtri[q] = t4;
        vtri[p1] = t4;
        vtri[p] = t;
        replace_nbr(N_p2a, t, t3);
        replace_nbr(N_p1q, u, t4);
        if (fixed_count) {
            fixed[ekey(p1, p)] += fixed_count;
            fixed[ekey(p, p2)] += fixed_count;
        }
        legalize(t, 2, stack);
        legalize(t3, 1, stack);
        legalize(u, 2, stack);
        legalize(t4, 1, stack);
        return p;
    }

    // triangle with edge (a, b) (a -> b in CCW order of that triangle) or -1
    int tri_with_edge(int a, int b, int* k_out) {
        int t0 = vtri[a];
        int t = t0;
        for (int guard = 0; guard < 100000; guard++) {
            const Tri& T = tris[t];
            int k = idx_of(T, a);
            if (T.v[(k + 1) % 3] == b) {
                *k_out = (k + 2) % 3;  // edge (a,b) is opposite v[k+2]
                return t;
            }
            if (T.v[(k + 2) % 3] == b) {
                *k_out = (k + 1) % 3;
                return t;
            }
            t = T.n[(k + 1) % 3];  // rotate around a
            if (t < 0 || t == t0) return -1;
        }
        return -1;
    }

    bool has_edge(int a, int b) {
        int k;
        return tri_with_edge(a, b, &k) >= 0;
    }

    bool on_line(int a, int b, int o) const {
        if (orient(pts[a], pts[b], pts[o]) == 0) return true;
        if (snap <= 0) return false;
        double len = std::hypot(pts[b].x - pts[a].x, pts[b].y - pts[a].y);
        return std::fabs(orient_val(pts[a], pts[b], pts[o])) <= snap * len;
    }

    // Insert constraint (a, b) with multiplicity cnt.  Work items are pushed
    // onto `todo` when the edge has to be split.
    void insert_constraint(int a, int b, int cnt, std::vector<std::array<int, 3>>& todo,
                           std::vector<std::pair<int, int>>& stack) {
        if (a == b) return;
        if (has_edge(a, b)) {
            fixed[ekey(a, b)] += cnt;
            return;
        }
        const P2 A = pts[a], B = pts[b];
        // find the first crossed edge around a
        int t0 = vtri[a], t = t0;
        int l = -1, r = -1, tcur = -1;
        for (int guard = 0; guard < 100000; guard++) {
            const Tri& T = tris[t];
            int k = idx_of(T, a);
            int v1 = T.v[(k + 1) % 3], v2 = T.v[(k + 2) % 3];
            int o1 = orient(A, pts[v1], B), o2 = orient(A, pts[v2], B);
            // collinear vertex in the direction of b
            if (o1 == 0 && (pts[v1].x - A.x) * (B.x - A.x) + (pts[v1].y - A.y) * (B.y - A.y) > 0) {
                todo.push_back({a, v1, cnt});
                todo.push_back({v1, b, cnt});
                return;
            }
            if (o1 > 0 && o2 < 0) {
                r = v1;
                l = v2;
                tcur = t;
                break;
            }
            t = T.n[(k + 1) % 3];
            if (t < 0 || t == t0) break;
        }
        if (tcur < 0) throw std::runtime_error("triangulation: constraint start not found");
        // walk along the segment collecting crossed edges (l left, r right of a->b)
        std::vector<std::pair<int, int>> crossed;
        for (int guard = 0; guard < 10000000; guard++) {
            crossed.push_back({l, r});
            const Tri& T = tris[tcur];
            int kl = idx_of(T, l), kr = idx_of(T, r);
            int kopp = 3 - kl - kr;
            int u = T.n[kopp];
            if (u < 0) throw std::runtime_error("triangulation: constraint leaves the domain");
            const Tri& U = tris[u];
            int o = U.v[nbr_idx(U, tcur)];
            if (is_fixed(l, r)) {
                // crossing an earlier constraint: split both at the intersection
                const P2 L = pts[l], R = pts[r];
                double den = (B.x - A.x) * (R.y - L.y) - (B.y - A.y) * (R.x - L.x);
                double tt = ((L.x - A.x) * (R.y - L.y) - (L.y - A.y) * (R.x - L.x)) / den;
                float ix = (float)(A.x + tt * (B.x - A.x)), iy = (float)(A.y + tt * (B.y - A.y));
                int cl = fixed[ekey(l, r)];
                fixed.erase(ekey(l, r));
                int np = add_point(ix, iy);
                int used = insert_point(np, stack);
                if (used != np) {
                    pts.pop_back();
                    vtri.pop_back();
                }
                todo.push_back({l, used, cl});
                todo.push_back({used, r, cl});
                todo.push_back({a, used, cnt});
                todo.push_back({used, b, cnt});
                return;
            }
            if (o == b) break;
            if (on_line(a, b, o)) {
                todo.push_back({a, o, cnt});
                todo.push_back({o, b, cnt});
                return;
            }
            if (orient(A, B, pts[o]) > 0)
                l = o;
            else
                r = o;
            tcur = u;
        }
        // Sloan: flip crossed edges until none crosses a-b
        std::deque<std::pair<int, int>> q(crossed.begin(), crossed.end());
        std::vector<std::pair<int, int>> fresh;
        size_t stall = 0;
        while (!q.empty()) {
            auto [u0, w0] = q.front();
            q.pop_front();
            int k;
            int ti = tri_with_edge(u0, w0, &k);
            if (ti < 0) ti = tri_with_edge(w0, u0, &k);
            if (ti < 0) throw std::runtime_error("triangulation: lost a crossed edge");
            const Tri& T = tris[ti];
            int apex = T.v[k];
            int ui = T.n[k];
            const Tri& U = tris[ui];
            int qv = U.v[nbr_idx(U, ti)];
            int p1 = T.v[(k + 1) % 3], p2 = T.v[(k + 2) % 3];
            bool convex = orient(pts[apex], pts[qv], pts[p1]) * orient(pts[apex], pts[qv], pts[p2]) < 0;
            if (!convex) {
                q.push_back({u0, w0});
                if (++stall > 4 * (q.size() + 4) * (q.size() + 4))
                    throw std::runtime_error("triangulation: constraint recovery stalled");
                continue;
            }
            stall = 0;
            flip(ti, k);
            int oa = orient(A, B, pts[apex]), oq = orient(A, B, pts[qv]);
            if (oa * oq < 0 && !((apex == a && qv == b) || (apex == b && qv == a)))
                q.push_back({apex, qv});
            else
                fresh.push_back({apex, qv});
        }
        fixed[ekey(a, b)] += cnt;
        // restore the Delaunay property around the new edges
        bool changed = true;
        for (int pass = 0; changed && pass < 64; pass++) {
            changed = false;
            for (auto& e : fresh) {
                if (is_fixed(e.first, e.second)) continue;
                int k;
                int ti = tri_with_edge(e.first, e.second, &k);
                if (ti < 0) continue;
                const Tri& T = tris[ti];
                int ui = T.n[k];
                if (ui < 0) continue;
                const Tri& U = tris[ui];
                int qv = U.v[nbr_idx(U, ti)];
                int p1 = T.v[(k + 1) % 3], p2 = T.v[(k + 2) % 3];
                int apex = T.v[k];
                if (incircle(pts[T.v[0]], pts[T.v[1]], pts[T.v[2]], pts[qv]) > 0 &&
                    orient(pts[apex], pts[qv], pts[p1]) * orient(pts[apex], pts[qv], pts[p2]) < 0) {
                    flip(ti, k);
                    e = {apex, qv};
                    changed = true;
                }
            }
        }
    }
};

}  // namespace

void triangulate_map(const std::vector<float>& vxy, const std::vector<int>& edges, float snap_dist,
                     std::vector<float>& tri6, std::vector<int>& nbr3) {
    const int nin = (int)vxy.size() / 2;
    if (nin < 3) throw std::runtime_error("need at least 3 vertices");
    // merge exact duplicates (first occurrence kept), remap edges
    std::vector<int> remap(nin);
    std::unordered_map<uint64_t, int> seen;
    std::vector<float> ux;
    for (int i = 0; i < nin; i++) {
        float x = vxy[2 * i], y = vxy[2 * i + 1];
        if (!std::isfinite(x) || !std::isfinite(y)) throw std::runtime_error("non-finite vertex");
        if (x == 0.0f) x = 0.0f;  // -0 == +0
        if (y == 0.0f) y = 0.0f;
        uint32_t bx, by;
        std::memcpy(&bx, &x, 4);
        std::memcpy(&by, &y, 4);
        uint64_t k = ((uint64_t)bx << 32) | by;
        auto it = seen.find(k);
        if (it == seen.end()) {
            int id = (int)ux.size() / 2;
            seen.emplace(k, id);
            ux.push_back(x);
            ux.push_back(y);
            remap[i] = id;
        } else {
            remap[i] = it->second;
        }
    }
    const int nv = (int)ux.size() / 2;
    CDT cdt;
    cdt.snap = snap_dist;
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (int i = 0; i < nv; i++) {
        cdt.add_point(ux[2 * i], ux[2 * i + 1]);
        x0 = std::min(x0, (double)ux[2 * i]);
        x1 = std::max(x1, (double)ux[2 * i]);
        y0 = std::min(y0, (double)ux[2 * i + 1]);
        y1 = std::max(y1, (double)ux[2 * i + 1]);
    }
    // super triangle far outside the bbox (exactly representable coordinates)
    double cx = std::floor((x0 + x1) / 2), cy = std::floor((y0 + y1) / 2);
    double R = std::ldexp(1.0, (int)std::ceil(std::log2(std::max({x1 - x0, y1 - y0, 1.0}))) + 4);
    int s0 = cdt.add_point(cx - 3 * R, cy - 3 * R);
    int s1 = cdt.add_point(cx + 3 * R, cy - 3 * R);
    int s2 = cdt.add_point(cx, cy + 3 * R);
    cdt.nsuper = 3;
    cdt.new_tri(s0, s1, s2, -1, -1, -1);
    cdt.last = 0;
    // insertion order: Hilbert curve over the bbox
    std::vector<std::pair<uint64_t, int>> order(nv);
    double w = std::max(x1 - x0, 1e-9), h = std::max(y1 - y0, 1e-9);
    for (int i = 0; i < nv; i++) {
        uint32_t hx = (uint32_t)std::min(65535.0, (ux[2 * i] - x0) / w * 65535.0);
        uint32_t hy = (uint32_t)std::min(65535.0, (ux[2 * i + 1] - y0) / h * 65535.0);
        order[i] = {hilbert(hx, hy, 16), i};
    }
    std::sort(order.begin(), order.end());
    std::vector<std::pair<int, int>> stack;
    for (auto& o : order) cdt.insert_point(o.second, stack);
    // constraints in input order, with multiplicity
    std::vector<std::array<int, 3>> todo;
    for (size_t e = 0; e + 1 < edges.size(); e += 2) {
        int a = edges[e], b = edges[e + 1];
        if (a < 0 || a >= nin || b < 0 || b >= nin) throw std::runtime_error("edge index out of range");
        todo.push_back({remap[a], remap[b], 1});
        while (!todo.empty()) {
            auto it = todo.back();
            todo.pop_back();
            std::vector<std::array<int, 3>> more;
            cdt.insert_constraint(it[0], it[1], it[2], more, stack);
            for (auto r = more.rbegin(); r != more.rend(); ++r) todo.push_back(*r);
        }
    }
    // depth: Dijkstra from a triangle touching the super triangle, edge weight =
    // constraint count
    const int nt = (int)cdt.tris.size();
    std::vector<int> depth(nt, INT32_MAX);
    int seed = -1;
    for (int t = 0; t < nt && seed < 0; t++) {
        if (!cdt.tris[t].alive) continue;
        for (int k = 0; k < 3; k++)
            if (cdt.tris[t].v[k] >= nv && cdt.tris[t].v[k] < nv + 3) seed = t;
    }
    std::vector<std::vector<int>> buckets(1);
    depth[seed] = 0;
    buckets[0].push_back(seed);
    for (size_t d = 0; d < buckets.size(); d++) {
        for (size_t i = 0; i < buckets[d].size(); i++) {
            int t = buckets[d][i];
            if (depth[t] != (int)d) continue;
            const Tri& T = cdt.tris[t];
            for (int k = 0; k < 3; k++) {
                int u = T.n[k];
                if (u < 0) continue;
                int a = T.v[(k + 1) % 3], b = T.v[(k + 2) % 3];
                auto it = cdt.fixed.find(ekey(a, b));
                int wgt = it == cdt.fixed.end() ? 0 : it->second;
                int nd = (int)d + wgt;
                if (nd < depth[u]) {
                    depth[u] = nd;
                    if ((int)buckets.size() <= nd) buckets.resize(nd + 1);
                    buckets[nd].push_back(u);
                }
            }
        }
    }
    // keep odd depth; number the kept triangles along a Hilbert curve over
    // their centroids (creation order breaks ties), so that a triangle's
    // neighbours get nearby ids: a search probes the per-node visited bytes of
    // a node's neighbours, and nearby ids put them on the node's own cache line
    std::vector<std::pair<uint64_t, int>> kept;
    double cx0 = 0.0, cy0 = 0.0, cx1 = 1.0, cy1 = 1.0;
    for (int i = 0; i < nv; i++) {
        const P2& p = cdt.pts[i];
        if (i == 0 || p.x < cx0) cx0 = p.x;
        if (i == 0 || p.y < cy0) cy0 = p.y;
        if (i == 0 || p.x > cx1) cx1 = p.x;
        if (i == 0 || p.y > cy1) cy1 = p.y;
    }
    const double span = std::max(cx1 - cx0, cy1 - cy0) > 0.0 ? std::max(cx1 - cx0, cy1 - cy0) : 1.0;
    for (int t = 0; t < nt; t++) {
        if (!cdt.tris[t].alive || depth[t] == INT32_MAX || depth[t] % 2 == 0) continue;
        const Tri& T = cdt.tris[t];
        bool super = false;
        for (int k = 0; k < 3; k++) super = super || (T.v[k] >= nv && T.v[k] < nv + 3);
        if (super) continue;
        const double mx = (cdt.pts[T.v[0]].x + cdt.pts[T.v[1]].x + cdt.pts[T.v[2]].x) / 3.0;
        const double my = (cdt.pts[T.v[0]].y + cdt.pts[T.v[1]].y + cdt.pts[T.v[2]].y) / 3.0;
        const double fx = std::min(std::max((mx - cx0) / span, 0.0), 1.0);
        const double fy = std::min(std::max((my - cy0) / span, 0.0), 1.0);
        kept.push_back({hilbert((uint32_t)(fx * 65535.0), (uint32_t)(fy * 65535.0), 16), t});
    }
    std::sort(kept.begin(), kept.end());
    std::vector<int> newid(nt, -1);
    int nk = 0;
    for (const auto& kt : kept) newid[kt.second] = nk++;
    if (nk == 0) throw std::runtime_error("no walkable triangles");
    tri6.assign(6 * (size_t)nk, 0.f);
    nbr3.assign(3 * (size_t)nk, -1);
    for (int t = 0; t < nt; t++) {
        int id = newid[t];
        if (id < 0) continue;
        const Tri& T = cdt.tris[t];
        for (int k = 0; k < 3; k++) {
            tri6[6 * (size_t)id + 2 * k] = (float)cdt.pts[T.v[k]].x;
            tri6[6 * (size_t)id + 2 * k + 1] = (float)cdt.pts[T.v[k]].y;
        }
        int c = 0;
        for (int k = 0; k < 3; k++) {
            int u = T.n[k];
            if (u >= 0 && newid[u] >= 0) nbr3[3 * (size_t)id + c++] = newid[u];
        }
    }
}

}  // namespace gmapf_amd
