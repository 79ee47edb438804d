#include "meshio.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

#include "../csrc/refelem.hpp"

namespace dgh {

int nodes_per_elem(int type) {
  if (type == 15) return 1;  // point
  int d, p;
  if (!dgk::gmsh_type_info(type, &d, &p)) return -1;
  return dgk::num_nodes(d, p);
}

namespace {

[[noreturn]] void fail(const std::string& path, const std::string& what) {
  throw std::runtime_error("MSH " + path + ": " + what);
}

// Token reader over an in-memory file that also supports raw binary reads.
struct Reader {
  std::string buf;
  size_t pos = 0;
  std::string path;
  bool eof() const { return pos >= buf.size(); }
  void skip_ws() { while (pos < buf.size() && isspace((unsigned char)buf[pos])) ++pos; }
  std::string token() {
    skip_ws();
    size_t s = pos;
    if (pos < buf.size() && buf[pos] == '"') {  // quoted string
      ++pos;
      while (pos < buf.size() && buf[pos] != '"') ++pos;
      std::string t = buf.substr(s + 1, pos - s - 1);
      ++pos;
      return t;
    }
    while (pos < buf.size() && !isspace((unsigned char)buf[pos])) ++pos;
    if (s == pos) fail(path, "unexpected end of file");
    return buf.substr(s, pos - s);
  }
  long long integer() {
    std::string t = token();
    char* e = nullptr;
    long long v = strtoll(t.c_str(), &e, 10);
    if (*e) fail(path, "expected an integer, got '" + t + "'");
    return v;
  }
  double real() {
    std::string t = token();
    char* e = nullptr;
    double v = strtod(t.c_str(), &e);
    if (*e) fail(path, "expected a number, got '" + t + "'");
    return v;
  }
  std::string line() {
    size_t s = pos;
    while (pos < buf.size() && buf[pos] != '\n') ++pos;
    std::string l = buf.substr(s, pos - s);
    if (pos < buf.size()) ++pos;
    if (!l.empty() && l.back() == '\r') l.pop_back();
    return l;
  }
  template <class T> T bin() {
    if (pos + sizeof(T) > buf.size()) fail(path, "truncated binary section");
    T v;
    memcpy(&v, buf.data() + pos, sizeof(T));
    pos += sizeof(T);
    return v;
  }
  void expect_end(const std::string& sec) {
    std::string t = token();
    if (t != "$End" + sec) fail(path, "expected $End" + sec + ", got '" + t + "'");
  }
  void skip_section(const std::string& sec) {
    std::string endtag = "$End" + sec;
    size_t f = buf.find(endtag, pos);
    if (f == std::string::npos) fail(path, "unterminated section $" + sec);
    pos = f + endtag.size();
  }
};

void read_entities_4(Reader& r, MshModel& m, bool v41, bool binary) {
  size_t counts[4];
  if (binary) for (auto& c : counts) c = r.bin<uint64_t>();
  else for (auto& c : counts) c = (size_t)r.integer();
  for (int dim = 0; dim < 4; ++dim)
    for (size_t i = 0; i < counts[dim]; ++i) {
      Entity e;
      e.dim = dim;
      if (binary) {
        e.tag = r.bin<int32_t>();
        int nbox = (dim == 0 && v41) ? 3 : 6;
        for (int k = 0; k < nbox; ++k) r.bin<double>();
        uint64_t np = r.bin<uint64_t>();
        for (uint64_t k = 0; k < np; ++k) e.phys.push_back(r.bin<int32_t>());
        if (dim > 0) {
          uint64_t nb = r.bin<uint64_t>();
          for (uint64_t k = 0; k < nb; ++k) r.bin<int32_t>();
        }
      } else {
        e.tag = (int)r.integer();
        int nbox = (dim == 0 && v41) ? 3 : 6;
        for (int k = 0; k < nbox; ++k) r.real();
        long long np = r.integer();
        for (long long k = 0; k < np; ++k) e.phys.push_back((int)r.integer());
        if (dim > 0) {
          long long nb = r.integer();
          for (long long k = 0; k < nb; ++k) r.integer();
        }
      }
      m.entities.push_back(e);
    }
}

void read_nodes_4(Reader& r, MshModel& m, bool v41, bool binary) {
  if (!v41) {  // 4.0 ASCII: numBlocks numNodes; block: tagEntity dimEntity parametric numNodes; tag x y z
    long long nb = r.integer(), nn = r.integer();
    m.node_tags.reserve(nn);
    m.node_xyz.reserve(3 * nn);
    for (long long b = 0; b < nb; ++b) {
      r.integer(); int dim = (int)r.integer(); int param = (int)r.integer(); long long n = r.integer();
      for (long long i = 0; i < n; ++i) {
        m.node_tags.push_back(r.integer());
        for (int c = 0; c < 3; ++c) m.node_xyz.push_back(r.real());
        if (param) for (int c = 0; c < dim; ++c) r.real();
      }
    }
    return;
  }
  size_t nb, nn;
  if (binary) { nb = r.bin<uint64_t>(); nn = r.bin<uint64_t>(); r.bin<uint64_t>(); r.bin<uint64_t>(); }
  else { nb = r.integer(); nn = r.integer(); r.integer(); r.integer(); }
  m.node_tags.reserve(nn);
  m.node_xyz.reserve(3 * nn);
  for (size_t b = 0; b < nb; ++b) {
    int dim, param; size_t n;
    if (binary) { dim = r.bin<int32_t>(); r.bin<int32_t>(); param = r.bin<int32_t>(); n = r.bin<uint64_t>(); }
    else { dim = (int)r.integer(); r.integer(); param = (int)r.integer(); n = r.integer(); }
    size_t base = m.node_tags.size();
    for (size_t i = 0; i < n; ++i) m.node_tags.push_back(binary ? (int64_t)r.bin<uint64_t>() : r.integer());
    m.node_xyz.resize(3 * (base + n));
    int extra = param ? dim : 0;
    for (size_t i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) m.node_xyz[3 * (base + i) + c] = binary ? r.bin<double>() : r.real();
      for (int c = 0; c < extra; ++c) binary ? r.bin<double>() : r.real();
    }
  }
}

void read_elements_4(Reader& r, MshModel& m, bool v41, bool binary) {
  size_t nb;
  if (!v41) { nb = r.integer(); r.integer(); }
  else if (binary) { nb = r.bin<uint64_t>(); r.bin<uint64_t>(); r.bin<uint64_t>(); r.bin<uint64_t>(); }
  else { nb = r.integer(); r.integer(); r.integer(); r.integer(); }
  for (size_t b = 0; b < nb; ++b) {
    ElemBlock blk;
    size_t n;
    if (!v41) { blk.entity = (int)r.integer(); blk.dim = (int)r.integer(); blk.type = (int)r.integer(); n = r.integer(); }
    else if (binary) { blk.dim = r.bin<int32_t>(); blk.entity = r.bin<int32_t>(); blk.type = r.bin<int32_t>(); n = r.bin<uint64_t>(); }
    else { blk.dim = (int)r.integer(); blk.entity = (int)r.integer(); blk.type = (int)r.integer(); n = r.integer(); }
    int npe = nodes_per_elem(blk.type);
    if (npe < 0) fail(r.path, "unsupported element type " + std::to_string(blk.type));
    blk.tags.resize(n);
    blk.nodes.resize(n * npe);
    for (size_t i = 0; i < n; ++i) {
      blk.tags[i] = binary ? (int64_t)r.bin<uint64_t>() : r.integer();
      for (int k = 0; k < npe; ++k) blk.nodes[i * npe + k] = binary ? (int64_t)r.bin<uint64_t>() : r.integer();
    }
    m.blocks.push_back(std::move(blk));
  }
}

void read_v2(Reader& r, MshModel& m, const std::string& sec) {
  if (sec == "Nodes") {
    long long n = r.integer();
    for (long long i = 0; i < n; ++i) {
      m.node_tags.push_back(r.integer());
      for (int c = 0; c < 3; ++c) m.node_xyz.push_back(r.real());
    }
    return;
  }
  // Elements: tag type ntags tags... nodes...; tags[0] physical, tags[1] elementary
  long long n = r.integer();
  std::map<std::tuple<int, int, int>, size_t> blk_of;  // (dim, elementary, type) -> block
  std::map<std::pair<int, int>, std::vector<int>> ent_phys;
  for (long long i = 0; i < n; ++i) {
    int64_t tag = r.integer();
    int type = (int)r.integer();
    int ntags = (int)r.integer();
    std::vector<int> tg(ntags);
    for (auto& t : tg) t = (int)r.integer();
    int npe = nodes_per_elem(type);
    if (npe < 0) fail(r.path, "unsupported element type " + std::to_string(type));
    int d = 0, p = 0;
    if (type != 15) dgk::gmsh_type_info(type, &d, &p);
    int phys = ntags > 0 ? tg[0] : 0, elem = ntags > 1 ? tg[1] : 0;
    auto key = std::make_tuple(d, elem, type);
    if (!blk_of.count(key)) {
      blk_of[key] = m.blocks.size();
      ElemBlock b; b.dim = d; b.entity = elem; b.type = type;
      m.blocks.push_back(b);
    }
    ElemBlock& b = m.blocks[blk_of[key]];
    b.tags.push_back(tag);
    for (int k = 0; k < npe; ++k) b.nodes.push_back(r.integer());
    auto& ph = ent_phys[{d, elem}];
    if (phys && std::find(ph.begin(), ph.end(), phys) == ph.end()) ph.push_back(phys);
  }
  for (auto& kv : ent_phys) {
    Entity e; e.dim = kv.first.first; e.tag = kv.first.second; e.phys = kv.second;
    m.entities.push_back(e);
  }
}

}  // namespace

MshModel read_msh(const std::string& path) {
  Reader r;
  r.path = path;
  {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail(path, "cannot open file");
    std::stringstream ss;
    ss << f.rdbuf();
    r.buf = ss.str();
  }
  MshModel m;
  double version = 0;
  bool binary = false, have_format = false;
  while (true) {
    r.skip_ws();
    if (r.eof()) break;
    std::string t = r.token();
    if (t.size() < 2 || t[0] != '$') fail(path, "expected a section header, got '" + t + "'");
    std::string sec = t.substr(1);
    if (sec == "MeshFormat") {
      version = r.real();
      binary = r.integer() != 0;
      int dsize = (int)r.integer();
      if (dsize != 8) fail(path, "data size must be 8");
      if (binary) {
        r.line();
        int one = r.bin<int32_t>();
        if (one != 1) fail(path, "binary file with foreign endianness");
      }
      r.expect_end(sec);
      have_format = true;
      if (!(version == 2.2 || version == 4.0 || version == 4.1 || version == 4))
        fail(path, "unsupported MSH version " + std::to_string(version));
      if (binary && version < 4.1) fail(path, "binary MSH is only read in version 4.1");
      continue;
    }
    if (!have_format) fail(path, "missing $MeshFormat");
    bool v4 = version >= 4.0, v41 = version >= 4.1;
    if (sec == "PhysicalNames") {
      long long n = r.integer();
      for (long long i = 0; i < n; ++i) {
        int dim = (int)r.integer(), tag = (int)r.integer();
        std::string name = r.token();
        m.phys_names[{dim, tag}] = name;
      }
      r.expect_end(sec);
    } else if (sec == "Entities" && v4) {
      if (binary) r.line();
      read_entities_4(r, m, v41, binary);
      r.expect_end(sec);
    } else if (sec == "Nodes") {
      if (binary) r.line();
      if (v4) read_nodes_4(r, m, v41, binary); else read_v2(r, m, sec);
      r.expect_end(sec);
    } else if (sec == "Elements") {
      if (binary) r.line();
      if (v4) read_elements_4(r, m, v41, binary); else read_v2(r, m, sec);
      r.expect_end(sec);
    } else {
      r.skip_section(sec);
    }
  }
  if (m.node_tags.empty()) fail(path, "no nodes");
  if (m.blocks.empty()) fail(path, "no elements");
  return m;
}

int model_dimension(const MshModel& m) {
  int d = 0;
  for (auto& b : m.blocks) d = std::max(d, b.dim);
  for (auto& e : m.entities) d = std::max(d, e.dim);
  return d;
}

TopMesh extract_top(const MshModel& m) {
  TopMesh t;
  t.dim = model_dimension(m);
  for (auto& b : m.blocks)
    if (b.dim == t.dim) { t.type = b.type; break; }
  if (!t.type) throw std::runtime_error("mesh has no element of the model dimension");
  if (!dgk::gmsh_type_info(t.type, &t.dim, &t.order))
    throw std::runtime_error("unsupported element type " + std::to_string(t.type) + " (simplices of order 1-5 only)");
  t.Np = dgk::num_nodes(t.dim, t.order);
  for (auto& b : m.blocks) {
    if (b.dim != t.dim) continue;
    if (b.type != t.type)
      throw std::runtime_error("mixed element types in dimension " + std::to_string(t.dim) + " are not supported");
    t.elem_tags.insert(t.elem_tags.end(), b.tags.begin(), b.tags.end());
    t.node_tags.insert(t.node_tags.end(), b.nodes.begin(), b.nodes.end());
  }
  t.K = (int64_t)t.elem_tags.size();
  std::unordered_map<int64_t, size_t> idx;
  idx.reserve(m.node_tags.size() * 2);
  for (size_t i = 0; i < m.node_tags.size(); ++i) idx[m.node_tags[i]] = i;
  t.node_xyz.resize(t.node_tags.size() * 3);
  for (size_t i = 0; i < t.node_tags.size(); ++i) {
    auto it = idx.find(t.node_tags[i]);
    if (it == idx.end()) throw std::runtime_error("element references unknown node " + std::to_string(t.node_tags[i]));
    for (int c = 0; c < 3; ++c) t.node_xyz[3 * i + c] = m.node_xyz[3 * it->second + c];
  }
  return t;
}

std::vector<std::pair<int, std::string>> physical_groups(const MshModel& m, int dim) {
  std::map<int, std::string> g;
  for (auto& kv : m.phys_names)
    if (kv.first.first == dim) g[kv.first.second] = kv.second;
  for (auto& e : m.entities)
    if (e.dim == dim)
      for (int p : e.phys)
        if (!g.count(p)) g[p] = "";
  return std::vector<std::pair<int, std::string>>(g.begin(), g.end());
}

std::vector<int64_t> nodes_for_physical_group(const MshModel& m, int dim, int tag) {
  std::vector<int64_t> out;
  for (auto& b : m.blocks) {
    if (b.dim != dim) continue;
    bool in = false;
    for (auto& e : m.entities)
      if (e.dim == dim && e.tag == b.entity)
        for (int p : e.phys) in |= (p == tag);
    if (in) out.insert(out.end(), b.nodes.begin(), b.nodes.end());
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

void write_msh41(const std::string& path, const MshModel& m) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  fprintf(f, "$MeshFormat\n4.1 0 8\n$EndMeshFormat\n");
  if (!m.phys_names.empty()) {
    fprintf(f, "$PhysicalNames\n%zu\n", m.phys_names.size());
    for (auto& kv : m.phys_names) fprintf(f, "%d %d \"%s\"\n", kv.first.first, kv.first.second, kv.second.c_str());
    fprintf(f, "$EndPhysicalNames\n");
  }
  // bounding box of all nodes (entities get the global box; readers do not depend on it)
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (size_t i = 0; i < m.node_tags.size(); ++i)
    for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], m.node_xyz[3 * i + c]); hi[c] = std::max(hi[c], m.node_xyz[3 * i + c]); }
  size_t cnt[4] = {0, 0, 0, 0};
  for (auto& e : m.entities) cnt[e.dim]++;
  fprintf(f, "$Entities\n%zu %zu %zu %zu\n", cnt[0], cnt[1], cnt[2], cnt[3]);
  for (int d = 0; d < 4; ++d)
    for (auto& e : m.entities) {
      if (e.dim != d) continue;
      if (d == 0) fprintf(f, "%d %.17g %.17g %.17g", e.tag, lo[0], lo[1], lo[2]);
      else fprintf(f, "%d %.17g %.17g %.17g %.17g %.17g %.17g", e.tag, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
      fprintf(f, " %zu", e.phys.size());
      for (int p : e.phys) fprintf(f, " %d", p);
      if (d > 0) fprintf(f, " 0");
      fprintf(f, "\n");
    }
  fprintf(f, "$EndEntities\n");
  // all nodes in one block on the first top-dimensional entity
  int topdim = model_dimension(m), topent = 1;
  for (auto& e : m.entities) if (e.dim == topdim) { topent = e.tag; break; }
  int64_t mn = m.node_tags.empty() ? 0 : *std::min_element(m.node_tags.begin(), m.node_tags.end());
  int64_t mx = m.node_tags.empty() ? 0 : *std::max_element(m.node_tags.begin(), m.node_tags.end());
  fprintf(f, "$Nodes\n1 %zu %lld %lld\n%d %d 0 %zu\n", m.node_tags.size(), (long long)mn, (long long)mx, topdim, topent,
          m.node_tags.size());
  for (auto t : m.node_tags) fprintf(f, "%lld\n", (long long)t);
  for (size_t i = 0; i < m.node_tags.size(); ++i)
    fprintf(f, "%.17g %.17g %.17g\n", m.node_xyz[3 * i], m.node_xyz[3 * i + 1], m.node_xyz[3 * i + 2]);
  fprintf(f, "$EndNodes\n");
  size_t ne = 0;
  int64_t emn = INT64_MAX, emx = 0;
  for (auto& b : m.blocks) {
    ne += b.tags.size();
    for (auto t : b.tags) { emn = std::min(emn, t); emx = std::max(emx, t); }
  }
  fprintf(f, "$Elements\n%zu %zu %lld %lld\n", m.blocks.size(), ne, (long long)emn, (long long)emx);
  for (auto& b : m.blocks) {
    int npe = nodes_per_elem(b.type);
    fprintf(f, "%d %d %d %zu\n", b.dim, b.entity, b.type, b.tags.size());
    for (size_t i = 0; i < b.tags.size(); ++i) {
      fprintf(f, "%lld", (long long)b.tags[i]);
      for (int k = 0; k < npe; ++k) fprintf(f, " %lld", (long long)b.nodes[i * npe + k]);
      fprintf(f, "\n");
    }
  }
  fprintf(f, "$EndElements\n");
  fclose(f);
}

void ViewWriter::begin(const MshModel& m) {
  write_msh41(path, m);
  header_written = true;
}

void ViewWriter::add_step(const std::string& view, int step, double t, int ncomp,
                          const std::vector<int64_t>& elem_tags, int Np, const double* data) {
  FILE* f = fopen(path.c_str(), "a");
  if (!f) throw std::runtime_error("cannot append to " + path);
  fprintf(f, "$ElementNodeData\n1\n\"%s\"\n1\n%.17g\n3\n%d\n%d\n%zu\n", view.c_str(), t, step, ncomp, elem_tags.size());
  for (size_t e = 0; e < elem_tags.size(); ++e) {
    fprintf(f, "%lld %d", (long long)elem_tags[e], Np);
    for (int k = 0; k < Np * ncomp; ++k) fprintf(f, " %.17g", data[e * (size_t)Np * ncomp + k]);
    fprintf(f, "\n");
  }
  fprintf(f, "$EndElementNodeData\n");
  fclose(f);
}

}  // namespace dgh
