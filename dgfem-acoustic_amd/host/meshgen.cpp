#include "meshgen.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <stdexcept>
#include <unordered_map>

#include "../csrc/refelem.hpp"

namespace dgh {

namespace {

struct Key {
  int64_t a, b, c;
  int i, j;
  bool operator==(const Key& o) const { return a == o.a && b == o.b && c == o.c && i == o.i && j == o.j; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) { h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); };
    mix((uint64_t)k.a); mix((uint64_t)k.b); mix((uint64_t)k.c); mix((uint64_t)k.i); mix((uint64_t)k.j);
    return (size_t)h;
  }
};

std::vector<std::array<int, 4>> elem_lattice(int dim, int p) {
  std::vector<std::array<int, 4>> L;
  if (dim == 1) for (auto a : dgk::line_lattice(p)) L.push_back({a[0], a[1], 0, 0});
  if (dim == 2) for (auto a : dgk::tri_lattice(p)) L.push_back({a[0], a[1], a[2], 0});
  if (dim == 3) for (auto a : dgk::tet_lattice(p)) L.push_back(a);
  return L;
}

}  // namespace

MshModel build_model(int dim, int p, const std::vector<double>& vxyz, const std::vector<int64_t>& simp,
                     const std::vector<BoundaryGroup>& groups,
                     const std::function<int(const int64_t*, int)>& classify) {
  const int nvpe = dim + 1;
  const int64_t nv = (int64_t)vxyz.size() / 3;
  const int64_t K = (int64_t)simp.size() / nvpe;
  MshModel m;
  // vertex nodes: tags 1..nv
  m.node_tags.resize(nv);
  m.node_xyz = vxyz;
  for (int64_t i = 0; i < nv; ++i) m.node_tags[i] = i + 1;
  std::unordered_map<Key, int64_t, KeyHash> shared;
  shared.reserve((size_t)K * 4);
  auto new_node = [&](const int64_t* verts, const int* idx, int n) {
    double x[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k)
      for (int c = 0; c < 3; ++c) x[c] += (double)idx[k] * vxyz[3 * verts[k] + c];
    for (int c = 0; c < 3; ++c) m.node_xyz.push_back(x[c] / p);
    int64_t tag = (int64_t)m.node_tags.size() + 1;
    m.node_tags.push_back(tag);
    return tag;
  };
  auto lat = elem_lattice(dim, p);
  const int Np = (int)lat.size();
  ElemBlock vol;
  vol.dim = dim; vol.entity = 1; vol.type = dgk::gmsh_type(dim, p);
  vol.tags.resize(K);
  vol.nodes.resize((size_t)K * Np);
  for (int64_t e = 0; e < K; ++e) {
    vol.tags[e] = e + 1;
    const int64_t* v = &simp[e * nvpe];
    for (int n = 0; n < Np; ++n) {
      // support vertices, sorted by vertex id (canonical order for shared nodes)
      std::array<std::pair<int64_t, int>, 4> sup;
      int ns = 0;
      for (int k = 0; k < nvpe; ++k)
        if (lat[n][k]) sup[ns++] = {v[k], lat[n][k]};
      std::sort(sup.begin(), sup.begin() + ns);
      int64_t tag;
      if (ns == 1) {
        tag = sup[0].first + 1;
      } else {
        int64_t vv[4] = {0, 0, 0, 0};
        int ii[4] = {0, 0, 0, 0};
        for (int k = 0; k < ns; ++k) { vv[k] = sup[k].first; ii[k] = sup[k].second; }
        bool interior = (ns == nvpe);  // nodes strictly inside the element are never shared
        if (interior) {
          tag = new_node(vv, ii, ns);
        } else {
          Key key{vv[0], vv[1], ns > 2 ? vv[2] : -1, ii[1], ns > 2 ? ii[2] : 0};
          auto it = shared.find(key);
          if (it != shared.end()) tag = it->second;
          else { tag = new_node(vv, ii, ns); shared.emplace(key, tag); }
        }
      }
      vol.nodes[(size_t)e * Np + n] = tag;
    }
  }
  // boundary faces: faces that appear once
  dgk::RefElem R = dgk::make_ref_elem(dim, p);
  std::unordered_map<Key, int, KeyHash> fcount;
  fcount.reserve((size_t)K * 4);
  for (int64_t e = 0; e < K; ++e)
    for (int f = 0; f < R.Nf; ++f) {
      int64_t fv[3] = {-1, -1, -1};
      for (int k = 0; k < dim; ++k) fv[k] = simp[e * nvpe + R.face_vertices[f][k]];
      std::sort(fv, fv + dim);
      fcount[Key{fv[0], fv[1], fv[2], 0, 0}]++;
    }
  std::map<int, ElemBlock> bblocks;
  int ftype = dim == 1 ? 15 : dgk::gmsh_type(dim - 1, p);
  int64_t next_tag = K + 1;
  for (int64_t e = 0; e < K; ++e)
    for (int f = 0; f < R.Nf; ++f) {
      int64_t fv[3] = {-1, -1, -1};
      for (int k = 0; k < dim; ++k) fv[k] = simp[e * nvpe + R.face_vertices[f][k]];
      int64_t sv[3] = {fv[0], fv[1], fv[2]};
      std::sort(sv, sv + dim);
      if (fcount[Key{sv[0], sv[1], sv[2], 0, 0}] != 1) continue;
      int tag = classify(fv, dim);
      if (!tag) continue;
      ElemBlock& b = bblocks[tag];
      b.dim = dim - 1; b.entity = tag; b.type = ftype;
      b.tags.push_back(next_tag++);
      for (int k : R.fmask[f]) b.nodes.push_back(vol.nodes[(size_t)e * Np + k]);
    }
  // entities and physical groups
  int dom_tag = 1;
  for (auto& g : groups) dom_tag = std::max(dom_tag, g.tag + 1);
  Entity ve; ve.dim = dim; ve.tag = 1; ve.phys = {dom_tag};
  m.entities.push_back(ve);
  m.phys_names[{dim, dom_tag}] = "Domain";
  for (auto& g : groups) {
    Entity be; be.dim = dim - 1; be.tag = g.tag; be.phys = {g.tag};
    m.entities.push_back(be);
    m.phys_names[{dim - 1, g.tag}] = g.name;
  }
  for (auto& kv : bblocks) m.blocks.push_back(std::move(kv.second));
  m.blocks.push_back(std::move(vol));
  return m;
}

MshModel gen_rect_tri(int nx, int ny, double Lx, double Ly, int p, const int side_tags[4],
                      const std::vector<BoundaryGroup>& groups) {
  std::vector<double> v;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) { v.push_back(Lx * i / nx); v.push_back(Ly * j / ny); v.push_back(0); }
  auto id = [&](int i, int j) { return (int64_t)j * (nx + 1) + i; };
  std::vector<int64_t> s;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int64_t a = id(i, j), b = id(i + 1, j), c = id(i + 1, j + 1), d = id(i, j + 1);
      s.insert(s.end(), {a, b, c});
      s.insert(s.end(), {a, c, d});
    }
  auto classify = [&](const int64_t* fv, int) {
    int i0 = (int)(fv[0] % (nx + 1)), j0 = (int)(fv[0] / (nx + 1));
    int i1 = (int)(fv[1] % (nx + 1)), j1 = (int)(fv[1] / (nx + 1));
    if (j0 == 0 && j1 == 0) return side_tags[0];
    if (i0 == nx && i1 == nx) return side_tags[1];
    if (j0 == ny && j1 == ny) return side_tags[2];
    if (i0 == 0 && i1 == 0) return side_tags[3];
    return 0;
  };
  return build_model(2, p, v, s, groups, classify);
}

MshModel gen_box_tet(int nx, int ny, int nz, double Lx, double Ly, double Lz, int p, double jitter, uint64_t seed,
                     const int block0[3], const int block1[3], int block_tag, const int side_tags[6],
                     const std::vector<BoundaryGroup>& groups) {
  const int n[3] = {nx, ny, nz};
  const double h[3] = {Lx / nx, Ly / ny, Lz / nz};
  bool has_block = block0 && block1 && block1[0] > block0[0] && block1[1] > block0[1] && block1[2] > block0[2];
  if (has_block && jitter != 0) throw std::runtime_error("gen_box_tet: jitter with an obstacle block is not supported");
  std::vector<double> v;
  v.reserve((size_t)(nx + 1) * (ny + 1) * (nz + 1) * 3);
  Rng rng(seed);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        int ijk[3] = {i, j, k};
        for (int c = 0; c < 3; ++c) {
          double x = h[c] * ijk[c];
          if (c == 0 && i == nx) x = Lx;
          if (c == 1 && j == ny) x = Ly;
          if (c == 2 && k == nz) x = Lz;
          if (jitter != 0) {
            double r = rng.uniform();  // drawn for every coordinate, so the stream does not depend on position
            if (ijk[c] > 0 && ijk[c] < n[c]) x += (2 * r - 1) * jitter * h[c];
          }
          v.push_back(x);
        }
      }
  auto id = [&](int i, int j, int k) { return ((int64_t)k * (ny + 1) + j) * (nx + 1) + i; };
  auto in_block = [&](int i, int j, int k) {
    return has_block && i >= block0[0] && i < block1[0] && j >= block0[1] && j < block1[1] && k >= block0[2] &&
           k < block1[2];
  };
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::vector<int64_t> s;
  s.reserve((size_t)nx * ny * nz * 24);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        if (in_block(i, j, k)) continue;
        for (auto& pm : perms) {
          int c[3] = {0, 0, 0};
          int64_t t[4];
          t[0] = id(i, j, k);
          for (int q = 0; q < 3; ++q) {
            c[pm[q]] = 1;
            t[q + 1] = id(i + c[0], j + c[1], k + c[2]);
          }
          // orientation from the unjittered lattice: odd permutations are negatively oriented
          bool odd = (pm[0] == 0 && pm[1] == 2) || (pm[0] == 1 && pm[1] == 0) || (pm[0] == 2 && pm[1] == 1);
          if (odd) std::swap(t[1], t[2]);
          double a[3], b[3], d[3];
          for (int q = 0; q < 3; ++q) {
            a[q] = v[3 * t[1] + q] - v[3 * t[0] + q];
            b[q] = v[3 * t[2] + q] - v[3 * t[0] + q];
            d[q] = v[3 * t[3] + q] - v[3 * t[0] + q];
          }
          double vol = a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                       a[2] * (b[0] * d[1] - b[1] * d[0]);
          if (!(vol > 0)) throw std::runtime_error("gen_box_tet: non-positive tetrahedron (jitter too large?)");
          s.insert(s.end(), t, t + 4);
        }
      }
  auto ijk_of = [&](int64_t vid, int* o) {
    o[0] = (int)(vid % (nx + 1));
    o[1] = (int)((vid / (nx + 1)) % (ny + 1));
    o[2] = (int)(vid / ((int64_t)(nx + 1) * (ny + 1)));
  };
  auto classify = [&](const int64_t* fv, int) {
    int a[3][3];
    for (int q = 0; q < 3; ++q) ijk_of(fv[q], a[q]);
    for (int c = 0; c < 3; ++c) {
      if (a[0][c] == 0 && a[1][c] == 0 && a[2][c] == 0) return side_tags[2 * c];
      if (a[0][c] == n[c] && a[1][c] == n[c] && a[2][c] == n[c]) return side_tags[2 * c + 1];
    }
    return has_block ? block_tag : 0;  // remaining boundary faces lie on the obstacle
  };
  return build_model(3, p, v, s, groups, classify);
}

MshModel gen_line(int n, double L, int p, const int side_tags[2], const std::vector<BoundaryGroup>& groups) {
  std::vector<double> v;
  for (int i = 0; i <= n; ++i) { v.push_back(L * i / n); v.push_back(0); v.push_back(0); }
  std::vector<int64_t> s;
  for (int i = 0; i < n; ++i) { s.push_back(i); s.push_back(i + 1); }
  auto classify = [&](const int64_t* fv, int) { return fv[0] == 0 ? side_tags[0] : fv[0] == n ? side_tags[1] : 0; };
  return build_model(1, p, v, s, groups, classify);
}

}  // namespace dgh
