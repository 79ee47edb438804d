// Gmsh MSH ingestion and output for the dgalerkin host.
//
// Replaces the Gmsh SDK calls the reference makes at setup
// (gmsh::open, dgalerkin.cpp:26; getDimension / getElementTypes /
// getElementsByType, src/Mesh.cpp:25-29; getPhysicalGroups / getPhysicalName,
// src/configParser.cpp:120-122; getNodesForPhysicalGroup, src/Mesh.cpp:371) and
// at output (view::add / addModelData / write, src/solver.cpp:185-191, 236-238,
// 288-291). Formats read: MSH 2.2 ASCII, 4.0 ASCII (the format of the meshes the
// reference ships), 4.1 ASCII and 4.1 binary. Written: MSH 4.1 ASCII.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace dgh {

struct ElemBlock {
  int dim = 0, entity = 0, type = 0;
  std::vector<int64_t> tags;   // element tags
  std::vector<int64_t> nodes;  // node tags, nodes_per_elem(type) per element
};

struct Entity {
  int dim = 0, tag = 0;
  std::vector<int> phys;  // physical tags
};

struct MshModel {
  std::vector<int64_t> node_tags;
  std::vector<double> node_xyz;  // 3 per node
  std::vector<Entity> entities;
  std::vector<ElemBlock> blocks;
  std::map<std::pair<int, int>, std::string> phys_names;  // (dim, tag) -> name
};

// Number of nodes of a Gmsh element type (simplices P1..P5 and the point), -1 if unknown.
int nodes_per_elem(int type);

// Throws std::runtime_error with a message naming the file and the problem.
MshModel read_msh(const std::string& path);
void write_msh41(const std::string& path, const MshModel& m);

// The element block the reference solver works on: the first element type of
// the model's top dimension (src/Mesh.cpp:25-29), in file order.
struct TopMesh {
  int dim = 0, order = 0, type = 0, Np = 0;
  int64_t K = 0;
  std::vector<int64_t> elem_tags;   // K
  std::vector<int64_t> node_tags;   // K*Np
  std::vector<double> node_xyz;     // K*Np*3
};
TopMesh extract_top(const MshModel& m);

int model_dimension(const MshModel& m);
// Physical groups of a dimension, ascending tag (gmsh::model::getPhysicalGroups order).
std::vector<std::pair<int, std::string>> physical_groups(const MshModel& m, int dim);
// Node tags of every element of the entities in physical group (dim, tag), sorted
// and unique (gmsh::model::mesh::getNodesForPhysicalGroup).
std::vector<int64_t> nodes_for_physical_group(const MshModel& m, int dim, int tag);

// Append ElementNodeData views (Pressure, Density, Velocity) for one saved step.
// Mirrors the per-save copy in src/solver.cpp:226-238.
struct ViewWriter {
  std::string path;
  bool header_written = false;
  void begin(const MshModel& m);  // writes the mesh once
  void add_step(const std::string& view, int step, double t, int ncomp, const std::vector<int64_t>& elem_tags,
                int Np, const double* data /* K*Np*ncomp */);
};

}  // namespace dgh
