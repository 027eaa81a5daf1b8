// ref_cdt_main.cpp -- TEST INFRASTRUCTURE ONLY.  Driver that runs the
// reference's vendored CDT library (compiled from the reference sources where
// they lie, see ref.mk) with the exact FillMapEnd recipe of the reference
// (src/GMAPF.cu:161-196), so tests can compare the product's own
// triangulator against it.  Input on stdin (text): nv, nv lines "x y",
// ne, ne lines "a b".  Output: ntri, then per triangle "x1 y1 x2 y2 x3 y3 n0 n1 n2"
// with neighbours compacted and -1 tail (GMAPF.cu:185-190).
#include <climits>
#include <cstdio>
#include <vector>

#include "CDT.h"

int main() {
    int nv = 0, ne = 0;
    if (std::scanf("%d", &nv) != 1) return 2;
    std::vector<CDT::V2d<float>> verts(nv);
    for (int i = 0; i < nv; i++) {
        float x, y;
        if (std::scanf("%f %f", &x, &y) != 2) return 2;
        verts[i] = CDT::V2d<float>::make(x, y);
    }
    if (std::scanf("%d", &ne) != 1) return 2;
    std::vector<CDT::Edge> edges;
    for (int i = 0; i < ne; i++) {
        int a, b;
        if (std::scanf("%d %d", &a, &b) != 2) return 2;
        edges.emplace_back(a, b);
    }
    CDT::RemoveDuplicatesAndRemapEdges(verts, edges);
    CDT::Triangulation<float> cdt(CDT::VertexInsertionOrder::Auto,
                                  CDT::IntersectingConstraintEdges::TryResolve, 1.0f);
    cdt.insertVertices(verts);
    cdt.insertEdges(edges);
    cdt.eraseOuterTrianglesAndHoles();
    std::printf("%zu\n", cdt.triangles.size());
    for (auto& tr : cdt.triangles) {
        for (int k = 0; k < 3; k++) {
            auto v = cdt.vertices[tr.vertices[k]];
            std::printf("%.9g %.9g ", v.x, v.y);
        }
        int c = 0, nb[3] = {-1, -1, -1};
        for (int k = 0; k < 3; k++)
            if (tr.neighbors[k] != UINT_MAX) nb[c++] = (int)tr.neighbors[k];
        std::printf("%d %d %d\n", nb[0], nb[1], nb[2]);
    }
    return 0;
}
