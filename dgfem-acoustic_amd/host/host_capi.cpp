// C ABI of the host-side library (libdgkhost.so): MSH ingestion, config
// parsing, boundary-condition resolution and the mesh generators, for callers
// that are not C++ (the Python test suite and bench harness use it via ctypes).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>

#include "../../include/dgk_host.h"
#include "config.hpp"
#include "meshgen.hpp"
#include "meshio.hpp"
#include "../csrc/refelem.hpp"

struct dgh_model {
  dgh::MshModel m;
  dgh::TopMesh top;
  bool top_ok = false;
};

namespace {
thread_local std::string g_err;
int fail(const std::exception& e) {
  g_err = e.what();
  return 1;
}
std::vector<dgh::BoundaryGroup> parse_names(const char* names) {
  std::vector<dgh::BoundaryGroup> g;
  if (!names) return g;
  std::stringstream ss(names);
  std::string item;
  while (std::getline(ss, item, ';')) {
    if (item.empty()) continue;
    size_t c = item.find(':');
    g.push_back({std::stoi(item.substr(0, c)), item.substr(c + 1)});
  }
  return g;
}
dgh_model* finish(dgh::MshModel&& m) {
  dgh_model* h = new dgh_model();
  h->m = std::move(m);
  h->top = dgh::extract_top(h->m);
  h->top_ok = true;
  return h;
}
}  // namespace

extern "C" {

const char* dgh_last_error(void) { return g_err.c_str(); }

dgh_model* dgh_read(const char* path) {
  try {
    return finish(dgh::read_msh(path));
  } catch (std::exception& e) {
    fail(e);
    return nullptr;
  }
}

dgh_model* dgh_gen_rect(int nx, int ny, double Lx, double Ly, int p, const int* side_tags, const char* names) {
  try {
    return finish(dgh::gen_rect_tri(nx, ny, Lx, Ly, p, side_tags, parse_names(names)));
  } catch (std::exception& e) {
    fail(e);
    return nullptr;
  }
}

dgh_model* dgh_gen_box(int nx, int ny, int nz, double Lx, double Ly, double Lz, int p, double jitter, uint64_t seed,
                       const int* block0, const int* block1, int block_tag, const int* side_tags, const char* names) {
  try {
    return finish(dgh::gen_box_tet(nx, ny, nz, Lx, Ly, Lz, p, jitter, seed, block0, block1, block_tag, side_tags,
                                   parse_names(names)));
  } catch (std::exception& e) {
    fail(e);
    return nullptr;
  }
}

dgh_model* dgh_gen_line(int n, double L, int p, const int* side_tags, const char* names) {
  try {
    return finish(dgh::gen_line(n, L, p, side_tags, parse_names(names)));
  } catch (std::exception& e) {
    fail(e);
    return nullptr;
  }
}

void dgh_free(dgh_model* h) { delete h; }

int dgh_write(const dgh_model* h, const char* path) {
  try {
    dgh::write_msh41(path, h->m);
  } catch (std::exception& e) {
    return fail(e);
  }
  return 0;
}

int dgh_top_info(const dgh_model* h, int64_t* out4) {
  out4[0] = h->top.dim;
  out4[1] = h->top.order;
  out4[2] = h->top.K;
  out4[3] = h->top.Np;
  return 0;
}

int dgh_top_arrays(const dgh_model* h, double* node_xyz, int64_t* node_tags, int64_t* elem_tags) {
  if (node_xyz) memcpy(node_xyz, h->top.node_xyz.data(), h->top.node_xyz.size() * sizeof(double));
  if (node_tags) memcpy(node_tags, h->top.node_tags.data(), h->top.node_tags.size() * sizeof(int64_t));
  if (elem_tags) memcpy(elem_tags, h->top.elem_tags.data(), h->top.elem_tags.size() * sizeof(int64_t));
  return 0;
}

int dgh_config_json(const char* text, char* buf, int buflen) {
  try {
    dgh::Config c = dgh::parse_config_text(text);
    std::ostringstream o;
    o.precision(17);
    auto arr = [&](const std::vector<std::vector<double>>& v) {
      o << "[";
      for (size_t i = 0; i < v.size(); ++i) {
        o << (i ? "," : "") << "[";
        for (size_t k = 0; k < v[i].size(); ++k) o << (k ? "," : "") << v[i][k];
        o << "]";
      }
      o << "]";
    };
    o << "{\"timeStart\":" << c.timeStart << ",\"timeEnd\":" << c.timeEnd << ",\"timeStep\":" << c.timeStep
      << ",\"timeRate\":" << c.timeRate << ",\"elementType\":\"" << c.elementType << "\",\"timeIntMethod\":\""
      << c.timeIntMethod << "\",\"saveFile\":\"" << c.saveFile << "\",\"numThreads\":" << c.numThreads
      << ",\"v0\":[" << c.v0[0] << "," << c.v0[1] << "," << c.v0[2] << "],\"rho0\":" << c.rho0 << ",\"c0\":" << c.c0
      << ",\"precision\":\"" << c.precision << "\",\"fluxSign\":\"" << c.fluxSign << "\",\"sources\":";
    arr(c.sources);
    o << ",\"initConditions\":";
    arr(c.initConditions);
    o << ",\"receivers\":";
    arr(c.receivers);
    o << "}";
    std::string s = o.str();
    if ((int)s.size() + 1 > buflen) { g_err = "buffer too small"; return 2; }
    memcpy(buf, s.c_str(), s.size() + 1);
  } catch (std::exception& e) {
    return fail(e);
  }
  return 0;
}

int dgh_bc_groups(const dgh_model* h, const char* config_text, int max_groups, int* ngroups, int* tags,
                  int* reflecting, int64_t* sizes) {
  try {
    dgh::Config c = dgh::parse_config_text(config_text);
    dgh::resolve_bcs(c, h->m);
    int n = 0;
    int bcdim = dgh::model_dimension(h->m) - 1;
    for (auto& kv : c.physBCs) {
      if (n >= max_groups) break;
      tags[n] = kv.first;
      reflecting[n] = kv.second.first == "Reflecting";
      sizes[n] = (int64_t)dgh::nodes_for_physical_group(h->m, bcdim, kv.first).size();
      ++n;
    }
    *ngroups = n;
  } catch (std::exception& e) {
    return fail(e);
  }
  return 0;
}

int dgh_group_nodes(const dgh_model* h, int tag, int64_t* out) {
  int bcdim = dgh::model_dimension(h->m) - 1;
  auto v = dgh::nodes_for_physical_group(h->m, bcdim, tag);
  memcpy(out, v.data(), v.size() * sizeof(int64_t));
  return 0;
}

int dgh_node_bc(const dgh_model* h, const char* config_text, uint8_t* node_bc) {
  try {
    dgh::Config c = dgh::parse_config_text(config_text);
    dgh::resolve_bcs(c, h->m);
    dgh_node_bc_resolve(h, &c, node_bc);
  } catch (std::exception& e) {
    return fail(e);
  }
  return 0;
}

// Product reference element (csrc/refelem.hpp): nodes [Np*3], M [Np*Np], S [dim*Np*Np],
// Mf [Nfp*Nfp], W [dim*Np*Np], LIFT [Np*Nf*Nfp], fmask [Nf*Nfp]; any pointer may be NULL.
int dgh_ref_element(int dim, int p, double* nodes, double* M, double* S, double* Mf, double* W, double* LIFT,
                    int* fmask) {
  try {
    dgk::RefElem R = dgk::make_ref_elem(dim, p);
    if (nodes) for (int n = 0; n < R.Np; ++n) for (int c = 0; c < 3; ++c) nodes[3 * n + c] = (double)R.nodes[n][c];
    if (M) std::copy(R.M.begin(), R.M.end(), M);
    if (S) std::copy(R.S.begin(), R.S.end(), S);
    if (Mf) std::copy(R.Mf.begin(), R.Mf.end(), Mf);
    if (W) std::copy(R.W.begin(), R.W.end(), W);
    if (LIFT) std::copy(R.LIFT.begin(), R.LIFT.end(), LIFT);
    if (fmask) for (int f = 0; f < R.Nf; ++f) for (int k = 0; k < R.Nfp; ++k) fmask[f * R.Nfp + k] = R.fmask[f][k];
  } catch (std::exception& e) {
    return fail(e);
  }
  return 0;
}

}  // extern "C"

// Per element-node boundary code: the last honoured group (ascending physical tag)
// whose node set contains the node decides (1 reflecting, 0 absorbing); nodes in no
// group are absorbing. A boundary face takes the code of its first node, which
// reproduces src/Mesh.cpp:364-384.
void dgh_node_bc_resolve(const dgh_model* h, const void* cfg, uint8_t* node_bc) {
  const dgh::Config& c = *static_cast<const dgh::Config*>(cfg);
  int bcdim = dgh::model_dimension(h->m) - 1;
  const auto& tags = h->top.node_tags;
  std::fill(node_bc, node_bc + tags.size(), 0);
  for (auto& kv : c.physBCs) {
    auto nodes = dgh::nodes_for_physical_group(h->m, bcdim, kv.first);
    uint8_t v = kv.second.first == "Reflecting" ? 1 : 0;
    for (size_t i = 0; i < tags.size(); ++i)
      if (std::binary_search(nodes.begin(), nodes.end(), tags[i])) node_bc[i] = v;
  }
}
