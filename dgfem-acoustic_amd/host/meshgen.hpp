// Deterministic simplex mesh generators for the solver's workloads.
//
// The reference meshes its geometries with Gmsh from .geo scripts
// (doc/2d/*.geo, doc/3d/*.geo); Gmsh is not available here, so the host builds
// structured simplex meshes itself and emits them as an MSH model with
// high-order nodes shared between neighbouring elements (as Gmsh does) and
// boundary elements carrying the physical groups the config refers to.
#pragma once
#include <cstdint>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "meshio.hpp"

namespace dgh {

// Doubles from std::mt19937_64 as (r >> 11) * 2^-53, so sequences are
// reproducible across standard libraries.
struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }
};

struct BoundaryGroup {
  int tag;
  std::string name;
};

// Build an order-p MSH model from straight-sided simplices.
//  vxyz:   nv x 3 vertex coordinates
//  simp:   K x (dim+1) vertex indices, positively oriented
//  classify(face vertex indices, count) -> physical tag of a boundary face (0 = none)
MshModel build_model(int dim, int p, const std::vector<double>& vxyz, const std::vector<int64_t>& simp,
                     const std::vector<BoundaryGroup>& groups,
                     const std::function<int(const int64_t*, int)>& classify);

// [0,Lx]x[0,Ly], nx x ny cells, diagonal (i,j)->(i+1,j+1), two CCW triangles per cell.
// side_tags: tags for the sides {ymin, xmax, ymax, xmin}; groups names them.
MshModel gen_rect_tri(int nx, int ny, double Lx, double Ly, int p, const int side_tags[4],
                      const std::vector<BoundaryGroup>& groups);

// [0,Lx]x[0,Ly]x[0,Lz], nx x ny x nz cells, Kuhn 6-tet split (positive orientation).
// jitter: node jitter as a fraction of the cell size (interior nodes in 3D, boundary nodes
// tangentially only), seeded. block: optional removed cell range [b0,b1) per axis (empty if b0==b1);
// block faces take block_tag. side_tags: {xmin, xmax, ymin, ymax, zmin, zmax}.
MshModel gen_box_tet(int nx, int ny, int nz, double Lx, double Ly, double Lz, int p, double jitter, uint64_t seed,
                     const int block0[3], const int block1[3], int block_tag, const int side_tags[6],
                     const std::vector<BoundaryGroup>& groups);

// [0,L] split into n segments (1D, points carry the physical groups {xmin, xmax}).
MshModel gen_line(int n, double L, int p, const int side_tags[2], const std::vector<BoundaryGroup>& groups);

}  // namespace dgh
