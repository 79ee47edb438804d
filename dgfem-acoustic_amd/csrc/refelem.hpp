// Reference simplex elements for the nodal DG operators (host side, FP64 output).
//
// Node ordering follows Gmsh's high-order Lagrange convention (the ordering the
// reference solver receives from gmsh::model::mesh::getElementsByType,
// src/Mesh.cpp:29): vertices, then edge nodes edge by edge (each edge walked from
// its first vertex), then face-interior nodes face by face, then volume-interior
// nodes, each interior block ordered recursively the same way.
// Reference coordinates are Gmsh's: line u in [-1,1]; triangle (0,0),(1,0),(0,1);
// tetrahedron at the unit corner.
//
// All operator matrices are computed in long double from a Legendre-product
// Vandermonde and collapsed Gauss-Legendre quadrature (exact for the polynomial
// degrees involved), then rounded once to double. On affine elements the
// reference's quadrature-based operators (src/Mesh.cpp:454-489, 500-539) reduce
// exactly to these matrices scaled by per-element geometry.
#pragma once
#include <array>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace dgk {

typedef long double ld;

// ---------------------------------------------------------------- lattices
// Integer barycentric lattice indices (sum == p) in Gmsh order.
inline std::vector<std::array<int, 2>> line_lattice(int p) {
  std::vector<std::array<int, 2>> L;
  if (p == 0) { L.push_back({0, 0}); return L; }
  L.push_back({p, 0});
  L.push_back({0, p});
  for (int i = 1; i < p; ++i) L.push_back({p - i, i});
  return L;
}

inline std::vector<std::array<int, 3>> tri_lattice(int p) {
  std::vector<std::array<int, 3>> L;
  if (p == 0) { L.push_back({0, 0, 0}); return L; }
  L.push_back({p, 0, 0});
  L.push_back({0, p, 0});
  L.push_back({0, 0, p});
  for (int i = 1; i < p; ++i) L.push_back({p - i, i, 0});  // edge 0-1
  for (int i = 1; i < p; ++i) L.push_back({0, p - i, i});  // edge 1-2
  for (int i = 1; i < p; ++i) L.push_back({i, 0, p - i});  // edge 2-0
  if (p >= 3)
    for (auto t : tri_lattice(p - 3)) L.push_back({t[0] + 1, t[1] + 1, t[2] + 1});
  return L;
}

// Gmsh tetrahedron edge and face tables (vertex indices).
static const int kTetEdges[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
static const int kTetFaces[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
static const int kTriEdges[3][2] = {{0, 1}, {1, 2}, {2, 0}};

inline std::vector<std::array<int, 4>> tet_lattice(int p) {
  std::vector<std::array<int, 4>> L;
  if (p == 0) { L.push_back({0, 0, 0, 0}); return L; }
  for (int v = 0; v < 4; ++v) { std::array<int, 4> a{0, 0, 0, 0}; a[v] = p; L.push_back(a); }
  for (int e = 0; e < 6; ++e)
    for (int i = 1; i < p; ++i) {
      std::array<int, 4> a{0, 0, 0, 0};
      a[kTetEdges[e][0]] = p - i;
      a[kTetEdges[e][1]] = i;
      L.push_back(a);
    }
  if (p >= 3)
    for (int f = 0; f < 4; ++f)
      for (auto t : tri_lattice(p - 3)) {
        std::array<int, 4> a{0, 0, 0, 0};
        a[kTetFaces[f][0]] = t[0] + 1;
        a[kTetFaces[f][1]] = t[1] + 1;
        a[kTetFaces[f][2]] = t[2] + 1;
        L.push_back(a);
      }
  if (p >= 4)
    for (auto t : tet_lattice(p - 4)) L.push_back({t[0] + 1, t[1] + 1, t[2] + 1, t[3] + 1});
  return L;
}

// Gmsh element type ids for simplices of order 1..5 (Gmsh manual, "MSH file format").
inline int gmsh_type(int dim, int p) {
  static const int line[] = {0, 1, 8, 26, 27, 28};
  static const int tri[] = {0, 2, 9, 21, 23, 25};
  static const int tet[] = {0, 4, 11, 29, 30, 31};
  if (p < 1 || p > 5) return -1;
  return dim == 1 ? line[p] : dim == 2 ? tri[p] : dim == 3 ? tet[p] : -1;
}
// Inverse of gmsh_type: returns false if the type is not a supported simplex.
inline bool gmsh_type_info(int type, int* dim, int* p) {
  for (int d = 1; d <= 3; ++d)
    for (int q = 1; q <= 5; ++q)
      if (gmsh_type(d, q) == type) { *dim = d; *p = q; return true; }
  return false;
}
inline int num_nodes(int dim, int p) {
  return dim == 1 ? p + 1 : dim == 2 ? (p + 1) * (p + 2) / 2 : (p + 1) * (p + 2) * (p + 3) / 6;
}

// ---------------------------------------------------------------- small dense LA (long double)
// In-place inverse by Gauss-Jordan with partial pivoting; A is n x n row-major.
inline void invert(std::vector<ld>& A, int n) {
  std::vector<ld> I(n * n, 0);
  for (int i = 0; i < n; ++i) I[i * n + i] = 1;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabsl(A[r * n + c]) > fabsl(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0) throw std::runtime_error("singular reference matrix");
    if (piv != c)
      for (int k = 0; k < n; ++k) { std::swap(A[c * n + k], A[piv * n + k]); std::swap(I[c * n + k], I[piv * n + k]); }
    ld d = A[c * n + c];
    for (int k = 0; k < n; ++k) { A[c * n + k] /= d; I[c * n + k] /= d; }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      ld f = A[r * n + c];
      if (f == 0) continue;
      for (int k = 0; k < n; ++k) { A[r * n + k] -= f * A[c * n + k]; I[r * n + k] -= f * I[c * n + k]; }
    }
  }
  A.swap(I);
}

// ---------------------------------------------------------------- quadrature (long double)
// Gauss-Legendre on [-1,1], n points, by Newton iteration on P_n.
inline void gauss_legendre(int n, std::vector<ld>& x, std::vector<ld>& w) {
  x.assign(n, 0); w.assign(n, 0);
  const ld pi = 3.141592653589793238462643383279502884L;
  for (int i = 0; i < n; ++i) {
    ld z = cosl(pi * (i + 0.75L) / (n + 0.5L)), dp = 0;
    for (int it = 0; it < 100; ++it) {
      ld p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) { ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      if (n == 1) { p1 = z; p0 = 1; }
      dp = n * (z * p1 - p0) / (z * z - 1);
      ld dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-19L) break;
    }
    ld p0 = 1, p1 = z;
    for (int k = 2; k <= n; ++k) { ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    dp = n * (z * p1 - p0) / (z * z - 1);
    x[n - 1 - i] = z;  // ascending
    w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
  }
}

// Exact rule for polynomials of total degree <= deg on the reference simplex of
// dimension dim (Gmsh coordinates), as points (u,v,w) and weights.
inline void simplex_rule(int dim, int deg, std::vector<std::array<ld, 3>>& pts, std::vector<ld>& wts) {
  pts.clear(); wts.clear();
  int n = deg / 2 + 2;  // collapsed factors raise the degree by up to dim-1
  std::vector<ld> x, w;
  gauss_legendre(n, x, w);
  if (dim == 1) {
    for (int i = 0; i < n; ++i) { pts.push_back({x[i], 0, 0}); wts.push_back(w[i]); }
    return;
  }
  // map to [0,1]
  std::vector<ld> t(n), tw(n);
  for (int i = 0; i < n; ++i) { t[i] = (x[i] + 1) / 2; tw[i] = w[i] / 2; }
  if (dim == 2) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        ld a = t[i], b = t[j];
        pts.push_back({a, (1 - a) * b, 0});
        wts.push_back(tw[i] * tw[j] * (1 - a));
      }
    return;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        ld a = t[i], b = t[j], c = t[k];
        pts.push_back({a, (1 - a) * b, (1 - a) * (1 - b) * c});
        wts.push_back(tw[i] * tw[j] * tw[k] * (1 - a) * (1 - a) * (1 - b));
      }
}

// ---------------------------------------------------------------- basis
// Legendre P_n and derivative on [-1,1].
inline void legendre(int n, ld z, ld* P, ld* dP) {
  P[0] = 1; dP[0] = 0;
  if (n >= 1) { P[1] = z; dP[1] = 1; }
  for (int k = 2; k <= n; ++k) {
    P[k] = ((2 * k - 1) * z * P[k - 1] - (k - 1) * P[k - 2]) / k;
    dP[k] = dP[k - 2] + (2 * k - 1) * P[k - 1];
  }
}

struct ModalSet {
  int dim, p;
  std::vector<std::array<int, 3>> deg;  // multi-index of each modal function
};
inline ModalSet modal_set(int dim, int p) {
  ModalSet m{dim, p, {}};
  if (dim == 1) for (int i = 0; i <= p; ++i) m.deg.push_back({i, 0, 0});
  if (dim == 2) for (int i = 0; i <= p; ++i) for (int j = 0; i + j <= p; ++j) m.deg.push_back({i, j, 0});
  if (dim == 3)
    for (int i = 0; i <= p; ++i) for (int j = 0; i + j <= p; ++j) for (int k = 0; i + j + k <= p; ++k)
      m.deg.push_back({i, j, k});
  return m;
}
// Evaluate modal functions (value and reference gradient) at a point in Gmsh reference coords.
inline void modal_eval(const ModalSet& m, const std::array<ld, 3>& x, ld* val, ld (*grad)[3]) {
  ld P[3][8], dP[3][8];
  for (int a = 0; a < m.dim; ++a) {
    ld z = (m.dim == 1) ? x[0] : 2 * x[a] - 1;  // line already on [-1,1]
    legendre(m.p, z, P[a], dP[a]);
  }
  ld sc = (m.dim == 1) ? 1 : 2;  // d z / d u
  for (size_t f = 0; f < m.deg.size(); ++f) {
    ld v = 1, g[3] = {1, 1, 1};
    for (int a = 0; a < m.dim; ++a) {
      int d = m.deg[f][a];
      v *= P[a][d];
      for (int b = 0; b < m.dim; ++b) g[b] *= (a == b) ? sc * dP[a][d] : P[a][d];
    }
    val[f] = v;
    if (grad) for (int b = 0; b < 3; ++b) grad[f][b] = (b < m.dim) ? g[b] : 0;
  }
}

// ---------------------------------------------------------------- the element
struct RefElem {
  int dim = 0, p = 0, Np = 0, Nf = 0, Nfp = 0, Nv = 0;
  std::vector<std::array<int, 4>> lattice;    // barycentric indices of each node
  std::vector<std::array<ld, 3>> nodes;       // reference coordinates
  std::vector<std::vector<int>> face_vertices;  // Nf x dim (local vertex ids, Gmsh order)
  std::vector<std::vector<int>> fmask;        // Nf x Nfp element node of face node k (face Gmsh order)
  // Lagrange basis coefficients on the modal set: phi_n = sum_m C[n*Np+m] psi_m
  ModalSet modal;
  std::vector<ld> C;
  // operators (double, row-major)
  std::vector<double> M, Minv;   // Np x Np reference mass matrix (Gmsh reference measure)
  std::vector<double> S;         // dim x Np x Np : S[r][i][j] = int d_r phi_i phi_j
  std::vector<double> W;         // dim x Np x Np : W_r = M^-1 S_r
  std::vector<double> Mf;        // Nfp x Nfp reference face mass matrix (face reference measure)
  std::vector<double> LIFT;      // Np x (Nf*Nfp): M^-1 E_j Mf
  std::vector<std::array<double, 3>> ref_normals;  // outward reference-face normals (unnormalised)
  double ref_face_measure[4] = {0, 0, 0, 0};  // measure of reference face j in reference coords
  double face_ref_measure = 0;                 // measure of the face's own reference element

  void eval_basis(const std::array<ld, 3>& x, ld* phi, ld (*dphi)[3]) const {
    std::vector<ld> v(Np);
    std::vector<std::array<ld, 3>> g(Np);
    modal_eval(modal, x, v.data(), reinterpret_cast<ld(*)[3]>(g.data()));
    for (int n = 0; n < Np; ++n) {
      ld s = 0, gs[3] = {0, 0, 0};
      for (int m = 0; m < Np; ++m) {
        s += C[n * Np + m] * v[m];
        for (int b = 0; b < 3; ++b) gs[b] += C[n * Np + m] * g[m][b];
      }
      phi[n] = s;
      if (dphi) for (int b = 0; b < 3; ++b) dphi[n][b] = gs[b];
    }
  }
};

// Barycentric -> Gmsh reference coordinates.
inline std::array<ld, 3> bary_to_ref(int dim, const int* lat, int p) {
  if (p == 0) {
    if (dim == 1) return {0, 0, 0};
    if (dim == 2) return {1.0L / 3, 1.0L / 3, 0};
    return {0.25L, 0.25L, 0.25L};
  }
  if (dim == 1) return {(ld)(lat[1] - lat[0]) / p, 0, 0};  // v0 at -1, v1 at +1
  if (dim == 2) return {(ld)lat[1] / p, (ld)lat[2] / p, 0};
  return {(ld)lat[1] / p, (ld)lat[2] / p, (ld)lat[3] / p};
}

// Lagrange basis on an arbitrary node set (used for faces), returns coefficient matrix.
inline std::vector<ld> lagrange_coeffs(const ModalSet& m, const std::vector<std::array<ld, 3>>& nodes) {
  int n = (int)nodes.size();
  std::vector<ld> V(n * n);
  std::vector<ld> row(n);
  for (int i = 0; i < n; ++i) {
    modal_eval(m, nodes[i], row.data(), nullptr);
    for (int j = 0; j < n; ++j) V[i * n + j] = row[j];
  }
  invert(V, n);  // V^-1 ; phi_i = sum_m (V^-1)[m][i] psi_m
  std::vector<ld> C(n * n);
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < n; ++m) C[i * n + m] = V[m * n + i];
  return C;
}

inline RefElem make_ref_elem(int dim, int p) {
  if (dim < 1 || dim > 3 || p < 1 || p > 5) throw std::runtime_error("unsupported element (dim 1-3, order 1-5)");
  RefElem R;
  R.dim = dim; R.p = p; R.Np = num_nodes(dim, p); R.Nv = dim + 1;
  R.Nf = dim + 1;
  R.Nfp = dim == 1 ? 1 : dim == 2 ? p + 1 : (p + 1) * (p + 2) / 2;
  // nodes
  if (dim == 1) for (auto a : line_lattice(p)) R.lattice.push_back({a[0], a[1], 0, 0});
  if (dim == 2) for (auto a : tri_lattice(p)) R.lattice.push_back({a[0], a[1], a[2], 0});
  if (dim == 3) for (auto a : tet_lattice(p)) R.lattice.push_back(a);
  for (auto& a : R.lattice) R.nodes.push_back(bary_to_ref(dim, a.data(), p));
  // faces
  if (dim == 1) R.face_vertices = {{0}, {1}};
  if (dim == 2) R.face_vertices = {{0, 1}, {1, 2}, {2, 0}};
  if (dim == 3) R.face_vertices = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
  auto find_node = [&](const std::array<int, 4>& a) {
    for (int n = 0; n < R.Np; ++n) if (R.lattice[n] == a) return n;
    throw std::runtime_error("lattice node not found");
  };
  // face-local lattices (face Gmsh order) mapped into element nodes
  std::vector<std::array<int, 3>> flat;
  if (dim == 1) flat = {{0, 0, 0}};
  if (dim == 2) for (auto a : line_lattice(p)) flat.push_back({a[0], a[1], 0});
  if (dim == 3) flat = tri_lattice(p);
  R.fmask.resize(R.Nf);
  for (int f = 0; f < R.Nf; ++f)
    for (auto& fa : flat) {
      std::array<int, 4> a{0, 0, 0, 0};
      if (dim == 1) a[R.face_vertices[f][0]] = p;
      else for (int k = 0; k < dim; ++k) a[R.face_vertices[f][k]] = fa[k];
      R.fmask[f].push_back(find_node(a));
    }
  // basis
  R.modal = modal_set(dim, p);
  R.C = lagrange_coeffs(R.modal, R.nodes);
  // element quadrature
  std::vector<std::array<ld, 3>> q; std::vector<ld> qw;
  simplex_rule(dim, 2 * p, q, qw);
  int Np = R.Np;
  std::vector<ld> M(Np * Np, 0), S(dim * Np * Np, 0);
  std::vector<ld> phi(Np);
  std::vector<std::array<ld, 3>> dphi(Np);
  for (size_t g = 0; g < q.size(); ++g) {
    R.eval_basis(q[g], phi.data(), reinterpret_cast<ld(*)[3]>(dphi.data()));
    for (int i = 0; i < Np; ++i)
      for (int j = 0; j < Np; ++j) {
        M[i * Np + j] += qw[g] * phi[i] * phi[j];
        for (int r = 0; r < dim; ++r) S[(r * Np + i) * Np + j] += qw[g] * dphi[i][r] * phi[j];
      }
  }
  std::vector<ld> Mi = M;
  invert(Mi, Np);
  R.M.resize(Np * Np); R.Minv.resize(Np * Np); R.S.resize(dim * Np * Np); R.W.resize(dim * Np * Np);
  for (int i = 0; i < Np * Np; ++i) { R.M[i] = (double)M[i]; R.Minv[i] = (double)Mi[i]; }
  for (int i = 0; i < dim * Np * Np; ++i) R.S[i] = (double)S[i];
  for (int r = 0; r < dim; ++r)
    for (int i = 0; i < Np; ++i)
      for (int j = 0; j < Np; ++j) {
        ld s = 0;
        for (int k = 0; k < Np; ++k) s += Mi[i * Np + k] * S[(r * Np + k) * Np + j];
        R.W[(r * Np + i) * Np + j] = (double)s;
      }
  // face reference element and mass matrix
  int Nfp = R.Nfp;
  std::vector<ld> Mf(Nfp * Nfp, 0);
  if (dim == 1) {
    Mf[0] = 1;  // point face: unit measure
    R.face_ref_measure = 1;
  } else {
    int fd = dim - 1;
    std::vector<std::array<ld, 3>> fnodes;
    for (auto& fa : flat) {
      int lat[4] = {fa[0], fa[1], fa[2], 0};
      fnodes.push_back(bary_to_ref(fd, lat, p));
    }
    ModalSet fm = modal_set(fd, p);
    std::vector<ld> FC = lagrange_coeffs(fm, fnodes);
    std::vector<std::array<ld, 3>> fq; std::vector<ld> fw;
    simplex_rule(fd, 2 * p, fq, fw);
    std::vector<ld> mv(Nfp), fphi(Nfp);
    for (size_t g = 0; g < fq.size(); ++g) {
      modal_eval(fm, fq[g], mv.data(), nullptr);
      for (int i = 0; i < Nfp; ++i) { ld s = 0; for (int m = 0; m < Nfp; ++m) s += FC[i * Nfp + m] * mv[m]; fphi[i] = s; }
      for (int i = 0; i < Nfp; ++i) for (int j = 0; j < Nfp; ++j) Mf[i * Nfp + j] += fw[g] * fphi[i] * fphi[j];
    }
    R.face_ref_measure = (fd == 1) ? 2.0 : 0.5;
  }
  R.Mf.resize(Nfp * Nfp);
  for (int i = 0; i < Nfp * Nfp; ++i) R.Mf[i] = (double)Mf[i];
  R.LIFT.assign(Np * R.Nf * Nfp, 0);
  for (int f = 0; f < R.Nf; ++f)
    for (int i = 0; i < Np; ++i)
      for (int k = 0; k < Nfp; ++k) {
        ld s = 0;
        for (int m = 0; m < Nfp; ++m) s += Mi[i * Np + R.fmask[f][m]] * Mf[m * Nfp + k];
        R.LIFT[i * (R.Nf * Nfp) + f * Nfp + k] = (double)s;
      }
  // outward reference normals and face measures (reference coordinates)
  R.ref_normals.resize(R.Nf);
  for (int f = 0; f < R.Nf; ++f) {
    std::array<ld, 3> v[3];
    for (int k = 0; k < dim; ++k) {
      int lat[4] = {0, 0, 0, 0};
      lat[R.face_vertices[f][k]] = 1;
      v[k] = bary_to_ref(dim, lat, 1);
    }
    // opposite vertex: the one not on the face
    int opp = -1;
    for (int a = 0; a < R.Nv; ++a) {
      bool on = false;
      for (int k = 0; k < dim; ++k) on |= (R.face_vertices[f][k] == a);
      if (!on) opp = a;
    }
    int lat[4] = {0, 0, 0, 0}; lat[opp] = 1;
    std::array<ld, 3> o = bary_to_ref(dim, lat, 1);
    ld n[3] = {0, 0, 0}, meas = 1;
    if (dim == 1) { n[0] = (v[0][0] > o[0]) ? 1 : -1; meas = 1; }
    if (dim == 2) {
      ld tx = v[1][0] - v[0][0], ty = v[1][1] - v[0][1];
      n[0] = ty; n[1] = -tx; meas = sqrtl(tx * tx + ty * ty);
    }
    if (dim == 3) {
      ld a[3], b[3];
      for (int c = 0; c < 3; ++c) { a[c] = v[1][c] - v[0][c]; b[c] = v[2][c] - v[0][c]; }
      n[0] = a[1] * b[2] - a[2] * b[1]; n[1] = a[2] * b[0] - a[0] * b[2]; n[2] = a[0] * b[1] - a[1] * b[0];
      meas = sqrtl(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) / 2;
    }
    // make outward (away from the opposite vertex)
    ld dotp = 0;
    for (int c = 0; c < dim; ++c) dotp += n[c] * (v[0][c] - o[c]);
    if (dotp < 0) for (int c = 0; c < 3; ++c) n[c] = -n[c];
    ld nn = sqrtl(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    R.ref_normals[f] = {(double)(n[0] / nn), (double)(n[1] / nn), (double)(n[2] / nn)};
    R.ref_face_measure[f] = (double)meas;
  }
  return R;
}

}  // namespace dgk
