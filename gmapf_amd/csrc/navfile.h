// navfile.h -- binary navmesh fixture ("GMAPFNAV", version 1).
//
// A filled navmesh (the node list FillMapManually takes: reference
// CuNodesMap::Node, include/gmapf/misc/CuNodesMap.h:14-21) in one
// little-endian byte string, so a map built once by FillMapStart /
// FillMapObstacle / FillMapEnd (src/GMAPF.cu:133-217) can be stored and
// loaded again without the triangulator:
//
//   offset  type       field
//   0       char[8]    "GMAPFNAV"
//   8       u32        version = 1
//   12      u32        header bytes = 64
//   16      u32        V, nodes (triangles)
//   20      u32        E, neighbour entries in the CSR below
//   24      f32[4]     bounding box of all vertices: x0, y0, x1, y1
//   40      u32[6]     reserved, 0
//   64      f32[6V]    vertices P1, P2, P3 (x, y) of each node
//           f32[2V]    centroids ((P1 + P2) + P3) / 3 (informational; a
//                      loader recomputes them, as FillMapManually does)
//           u32[V+1]   CSR row starts: node i's neighbours are
//                      col[row[i] .. row[i+1]), at most 3
//           u32[E]     CSR neighbour ids in slot order (NeibsIdx without
//                      its -1 tail)
//           u64        FNV-1a 64 of every preceding byte
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace gmapf_amd {

constexpr uint32_t kNavVersion = 1;
constexpr uint32_t kNavHeaderBytes = 64;

// Encodes V nodes (tri6: 6 floats per node, nbr3: 3 ids per node with a -1
// tail, cen2: 2 floats per node).
std::vector<unsigned char> nav_encode(int V, const float* tri6, const int* nbr3, const float* cen2);

// Decodes and validates a byte string; throws std::runtime_error naming the
// first problem (size, magic, version, checksum, CSR shape, id range).
void nav_decode(const unsigned char* buf, size_t len, int& V, std::vector<float>& tri6, std::vector<int>& nbr3);

// FNV-1a 64 over n bytes.
uint64_t nav_checksum(const unsigned char* p, size_t n);

}  // namespace gmapf_amd
