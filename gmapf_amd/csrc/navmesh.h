// navmesh.h -- host-side navmesh: triangle SoA, compacted neighbour lists,
// centroids and the exact point-location grid shared with the device.
#pragma once
#include <string>
#include <vector>

namespace gmapf_amd {

// Largest navmesh the step takes (FillMap* and the GMAPFNAV loader reject more).
constexpr int kMaxNodes = 1 << 26;

struct Navmesh {
    int V = 0;
    std::vector<float> tri;  // 6V: p1x p1y p2x p2y p3x p3y
    std::vector<int> nbr;    // 3V, compacted, -1 tail
    std::vector<float> cen;  // 2V

    // point-location grid (see build_grid)
    float gx0 = 0, gy0 = 0, inv_cs = 1;
    float rx0 = 0, ry0 = 0, rx1 = -1, ry1 = -1;
    int nx = 0, ny = 0;
    std::vector<int> cell_start, cell_tris, global_tris;
    // bounding box of all vertices
    float bx0 = 0, by0 = 0, bx1 = 0, by1 = 0;

    // Build from triangles + neighbours; centroids ((P1+P2)+P3)/3.0f
    // (GMAPF.cu:212-216).  Neighbour lists are kept as given (-1 = none).
    // Throws std::runtime_error on malformed input.
    // at most kMaxNodes nodes
    void assign(int v, const float* tri6, const int* nbr3);
    void build_grid();
    // CuNodesMap.h:216-229: minimum index of a triangle containing (x, y), -1 if none.
    int locate(float x, float y) const;
    int locate_brute(float x, float y) const;
    bool in_tri(int t, float x, float y) const;
    // structural checks: neighbour symmetry, ids in range
    std::string validate() const;
};

// Constrained Delaunay triangulation of a map rectangle with polygonal
// obstacles (GMAPF.cu:133-196): boundary (0,0),(0,H),(W,H),(W,0); obstacles
// are closed polylines; duplicate vertices are merged; crossing constraint
// edges are split at their intersection; vertices closer than snap_dist to a
// constraint edge are put on it; triangles at even nesting depth (outside the
// boundary, inside obstacles) are removed.  Output: 6V floats, 3V neighbours.
void triangulate_map(const std::vector<float>& vxy, const std::vector<int>& edges, float snap_dist,
                     std::vector<float>& tri6, std::vector<int>& nbr3);

}  // namespace gmapf_amd
