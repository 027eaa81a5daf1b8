// navmesh.cpp -- host navmesh container and exact point-location grid.
#include "navmesh.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace gmapf_amd {

namespace {

// CuNodesMap.h:272-275
inline float get_sign(float p1x, float p1y, float p2x, float p2y, float p3x, float p3y) {
    return (p1x - p3x) * (p2y - p3y) - (p2x - p3x) * (p1y - p3y);
}

inline int cell_of(float v, float o, float inv, int n) {
    float f = floorf((v - o) * inv);
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)(n - 1)) return n - 1;
    return (int)f;
}

}  // namespace

void Navmesh::assign(int v, const float* tri6, const int* nbr3) {
    if (v <= 0) throw std::runtime_error("navmesh has no nodes");
    // the search packs a node id with a few bits of segment and direction
    // into one 32-bit word (and every search slot holds 9 bytes per node)
    if (v > kMaxNodes) throw std::runtime_error("navmesh has more than 2^26 nodes");
    V = v;
    tri.assign(tri6, tri6 + 6 * (size_t)v);
    nbr.assign(3 * (size_t)v, -1);
    cen.resize(2 * (size_t)v);
    for (int i = 0; i < v; i++) {
        // kept as given: the search stops at the first -1 (PathFinder.h:200-201)
        for (int s = 0; s < 3; s++) {
            int n = nbr3[3 * (size_t)i + s];
            if (n != -1 && (n < 0 || n >= v)) throw std::runtime_error("neighbour index out of range");
            nbr[3 * (size_t)i + s] = n;
        }
        const float* t = &tri[6 * (size_t)i];
        cen[2 * (size_t)i] = ((t[0] + t[2]) + t[4]) / 3.0f;
        cen[2 * (size_t)i + 1] = ((t[1] + t[3]) + t[5]) / 3.0f;
    }
    bx0 = by0 = INFINITY;
    bx1 = by1 = -INFINITY;
    for (size_t k = 0; k < 3 * (size_t)v; k++) {
        float x = tri[2 * k], y = tri[2 * k + 1];
        if (!std::isfinite(x) || !std::isfinite(y)) throw std::runtime_error("non-finite vertex");
        bx0 = std::min(bx0, x);
        bx1 = std::max(bx1, x);
        by0 = std::min(by0, y);
        by1 = std::max(by1, y);
    }
    build_grid();
}

bool Navmesh::in_tri(int i, float px, float py) const {
    const float* t = &tri[6 * (size_t)i];
    float d1 = get_sign(px, py, t[0], t[1], t[2], t[3]);
    float d2 = get_sign(px, py, t[2], t[3], t[4], t[5]);
    float d3 = get_sign(px, py, t[4], t[5], t[0], t[1]);
    bool neg = (d1 < 0) || (d2 < 0) || (d3 < 0);
    bool pos = (d1 > 0) || (d2 > 0) || (d3 > 0);
    return !(neg && pos);
}

int Navmesh::locate_brute(float x, float y) const {
    for (int i = 0; i < V; i++)
        if (in_tri(i, x, y)) return i;
    return -1;
}

// A point accepted by the FP32 sign test of triangle t lies on the inner side
// of every edge line or within eps of it, eps = 4.25 * 2^-24 * D where D
// bounds |p - vertex| (rounding of the three differences, two products and
// one subtraction).  So it lies in the triangle grown by eps per edge, i.e.
// within eps / sin(theta_min / 2) of t.  A triangle whose inradius is below
// a few eps could also accept points with the opposite sign pattern; such
// triangles, and those whose margin exceeds a few cells, are kept in
// global_tris and always scanned.
void Navmesh::build_grid() {
    double w = std::max((double)bx1 - bx0, 1e-3), h = std::max((double)by1 - by0, 1e-3);
    double cs = std::sqrt(w * h / V) * 1.5;
    cs = std::max(cs, std::max(w, h) / 2048.0);
    double pad = 2.0 * cs;
    double D = std::hypot(w + 2 * pad, h + 2 * pad);
    double eps = 4.25 * std::ldexp(1.0, -24) * D * 1.1;
    rx0 = (float)(bx0 - pad);
    ry0 = (float)(by0 - pad);
    rx1 = (float)(bx1 + pad);
    ry1 = (float)(by1 + pad);
    gx0 = rx0;
    gy0 = ry0;
    inv_cs = (float)(1.0 / cs);
    nx = (int)std::ceil(((double)rx1 - rx0) / cs) + 1;
    ny = (int)std::ceil(((double)ry1 - ry0) / cs) + 1;
    std::vector<int> counts((size_t)nx * ny + 1, 0);
    struct Span {
        int x0, x1, y0, y1;
    };
    std::vector<Span> spans(V);
    global_tris.clear();
    for (int i = 0; i < V; i++) {
        const float* t = &tri[6 * (size_t)i];
        double ax = t[0], ay = t[1], bx = t[2], by = t[3], qx = t[4], qy = t[5];
        double la = std::hypot(bx - qx, by - qy), lb = std::hypot(ax - qx, ay - qy),
               lc = std::hypot(ax - bx, ay - by);
        double area2 = std::fabs((bx - ax) * (qy - ay) - (qx - ax) * (by - ay));
        double per = la + lb + lc;
        double inr = per > 0 ? area2 / per : 0.0;
        auto half_sin = [](double opp, double s1, double s2) {
            if (s1 <= 0 || s2 <= 0) return 0.0;
            double c = (s1 * s1 + s2 * s2 - opp * opp) / (2 * s1 * s2);
            c = std::max(-1.0, std::min(1.0, c));
            return std::sqrt(std::max(0.0, (1 - c) / 2));
        };
        double smin = std::min({half_sin(la, lb, lc), half_sin(lb, la, lc), half_sin(lc, la, lb)});
        double margin = smin > 0 ? 1.25 * eps / smin + 1e-6 : 1e300;
        if (inr <= 4 * eps || margin > 4 * cs) {
            global_tris.push_back(i);
            spans[i] = Span{1, 0, 1, 0};
            continue;
        }
        float lo_x = std::nextafter((float)(std::min({ax, bx, qx}) - margin), -INFINITY);
        float hi_x = std::nextafter((float)(std::max({ax, bx, qx}) + margin), INFINITY);
        float lo_y = std::nextafter((float)(std::min({ay, by, qy}) - margin), -INFINITY);
        float hi_y = std::nextafter((float)(std::max({ay, by, qy}) + margin), INFINITY);
        Span sp{cell_of(lo_x, gx0, inv_cs, nx), cell_of(hi_x, gx0, inv_cs, nx),
                cell_of(lo_y, gy0, inv_cs, ny), cell_of(hi_y, gy0, inv_cs, ny)};
        spans[i] = sp;
        for (int cy = sp.y0; cy <= sp.y1; cy++)
            for (int cx = sp.x0; cx <= sp.x1; cx++) counts[(size_t)cy * nx + cx + 1]++;
    }
    cell_start.assign((size_t)nx * ny + 1, 0);
    for (size_t c = 0; c < (size_t)nx * ny; c++) cell_start[c + 1] = cell_start[c] + counts[c + 1];
    cell_tris.assign(cell_start.back(), 0);
    std::vector<int> fill(cell_start.begin(), cell_start.end() - 1);
    for (int i = 0; i < V; i++) {  // ascending i -> each cell list sorted
        const Span& sp = spans[i];
        for (int cy = sp.y0; cy <= sp.y1; cy++)
            for (int cx = sp.x0; cx <= sp.x1; cx++) cell_tris[fill[(size_t)cy * nx + cx]++] = i;
    }
}

int Navmesh::locate(float x, float y) const {
    if (V == 0) return -1;
    if (!(x >= rx0 && x <= rx1 && y >= ry0 && y <= ry1)) return locate_brute(x, y);
    int cx = cell_of(x, gx0, inv_cs, nx), cy = cell_of(y, gy0, inv_cs, ny);
    size_t c = (size_t)cy * nx + cx;
    int best = -1;
    for (int k = cell_start[c]; k < cell_start[c + 1]; k++) {
        if (in_tri(cell_tris[k], x, y)) {
            best = cell_tris[k];
            break;
        }
    }
    for (int t : global_tris) {
        if (best >= 0 && t > best) break;
        if (in_tri(t, x, y)) {
            best = t;
            break;
        }
    }
    return best;
}

std::string Navmesh::validate() const {
    for (int i = 0; i < V; i++) {
        for (int s = 0; s < 3; s++) {
            int n = nbr[3 * (size_t)i + s];
            if (n < 0) continue;
            bool back = false;
            for (int k = 0; k < 3; k++) back = back || nbr[3 * (size_t)n + k] == i;
            if (!back) return "neighbour relation is not symmetric at node " + std::to_string(i);
        }
    }
    return "";
}

}  // namespace gmapf_amd
