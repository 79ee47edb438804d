// CPU oracle for the DG hot path. TEST INFRASTRUCTURE ONLY: loaded by tests/,
// the smoke entry point and bench.py's CPU-baseline leg; never by the product.
// See oracle.h for the reference functions each part restates.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

// Mutation hooks for the oracle's own pin tests (tests/test_oracle_pins.py builds copies of this
// file with a planted mistake and checks that the pins reject them). The shipped build uses the
// reference's values: the absorbing boundary matrix RKR = 1/4 [[c, rho c^2 n^T], [n/rho, c n n^T]]
// (src/Mesh.cpp:397-415) and the mean flow v0 as given (src/Mesh.cpp:577-591).
#ifndef ORC_ABS_COEF
#define ORC_ABS_COEF 0.25
#endif
#ifndef ORC_V0_SIGN
#define ORC_V0_SIGN 1.0
#endif

namespace {

typedef long double ld;

// ------------------------------------------------------------------ dual numbers (value + 3 partials)
struct Du {
  ld v, d[3];
};
inline Du cst(ld a) { return {a, {0, 0, 0}}; }
inline Du operator+(Du a, Du b) { return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}}; }
inline Du operator-(Du a, Du b) { return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}}; }
inline Du operator*(Du a, Du b) {
  return {a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
inline Du operator/(Du a, Du b) {
  ld q = a.v / b.v;
  return {q, {(a.d[0] - q * b.d[0]) / b.v, (a.d[1] - q * b.d[1]) / b.v, (a.d[2] - q * b.d[2]) / b.v}};
}
inline Du operator*(ld s, Du a) { return {s * a.v, {s * a.d[0], s * a.d[1], s * a.d[2]}}; }

// Jacobi P_n^{(alpha,0)} by the three-term recurrence (unnormalised).
Du jacobi(int n, int alpha, Du x) {
  const int beta = 0;
  Du p0 = cst(1);
  if (n == 0) return p0;
  Du p1 = 0.5L * ((ld)(alpha + beta + 2) * x + cst((ld)(alpha - beta)));
  for (int k = 1; k < n; ++k) {
    ld a1 = 2.0L * (k + 1) * (k + alpha + beta + 1) * (2 * k + alpha + beta);
    ld a2 = (ld)(2 * k + alpha + beta + 1) * (ld)(alpha * alpha - beta * beta);
    ld a3 = (ld)(2 * k + alpha + beta) * (2 * k + alpha + beta + 1) * (2 * k + alpha + beta + 2);
    ld a4 = 2.0L * (k + alpha) * (k + beta) * (2 * k + alpha + beta + 2);
    Du p2 = (1.0L / a1) * ((cst(a2) + a3 * x) * p1 - a4 * p0);
    p0 = p1;
    p1 = p2;
  }
  return p1;
}
Du powi(Du x, int n) {
  Du r = cst(1);
  for (int i = 0; i < n; ++i) r = r * x;
  return r;
}

// Orthogonal (collapsed-coordinate) modal basis on the Gmsh reference simplex.
struct Modal {
  int dim, p;
  std::vector<std::array<int, 3>> idx;
  Modal(int d, int q) : dim(d), p(q) {
    if (d == 1) for (int i = 0; i <= q; ++i) idx.push_back({i, 0, 0});
    if (d == 2) for (int i = 0; i <= q; ++i) for (int j = 0; i + j <= q; ++j) idx.push_back({i, j, 0});
    if (d == 3)
      for (int i = 0; i <= q; ++i) for (int j = 0; i + j <= q; ++j) for (int k = 0; i + j + k <= q; ++k)
        idx.push_back({i, j, k});
  }
  // values and gradients (w.r.t. Gmsh reference coordinates) at x
  void eval(const ld* x, std::vector<Du>& out) const {
    Du u{x[0], {1, 0, 0}}, v{x[1], {0, 1, 0}}, w{x[2], {0, 0, 1}};
    out.resize(idx.size());
    if (dim == 1) {
      for (size_t m = 0; m < idx.size(); ++m) out[m] = jacobi(idx[m][0], 0, u);
      return;
    }
    Du r = 2.0L * u - cst(1), s = 2.0L * v - cst(1), t = 2.0L * w - cst(1);
    if (dim == 2) {
      Du a = (fabsl(1 - s.v) < 1e-15L) ? cst(-1) : (2.0L * (cst(1) + r)) / (cst(1) - s) - cst(1);
      Du b = s;
      for (size_t m = 0; m < idx.size(); ++m) {
        int i = idx[m][0], j = idx[m][1];
        out[m] = jacobi(i, 0, a) * jacobi(j, 2 * i + 1, b) * powi(cst(1) - b, i);
      }
      return;
    }
    Du den = cst(0) - s - t;
    Du a = (fabsl(den.v) < 1e-15L) ? cst(-1) : (2.0L * (cst(1) + r)) / den - cst(1);
    Du b = (fabsl(1 - t.v) < 1e-15L) ? cst(-1) : (2.0L * (cst(1) + s)) / (cst(1) - t) - cst(1);
    Du c = t;
    for (size_t m = 0; m < idx.size(); ++m) {
      int i = idx[m][0], j = idx[m][1], k = idx[m][2];
      out[m] = jacobi(i, 0, a) * jacobi(j, 2 * i + 1, b) * powi(cst(1) - b, i) * jacobi(k, 2 * i + 2 * j + 2, c) *
               powi(cst(1) - c, i + j);
    }
  }
};

// LU with partial pivoting, long double; returns inverse of A (n x n row-major).
std::vector<ld> inv_ld(std::vector<ld> A, int n) {
  std::vector<int> piv(n);
  for (int i = 0; i < n; ++i) piv[i] = i;
  for (int c = 0; c < n; ++c) {
    int pr = c;
    for (int r = c + 1; r < n; ++r) if (fabsl(A[r * n + c]) > fabsl(A[pr * n + c])) pr = r;
    if (A[pr * n + c] == 0) throw std::runtime_error("singular Vandermonde");
    if (pr != c) { for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[pr * n + k]); std::swap(piv[c], piv[pr]); }
    for (int r = c + 1; r < n; ++r) {
      ld f = A[r * n + c] / A[c * n + c];
      A[r * n + c] = f;
      for (int k = c + 1; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
    }
  }
  std::vector<ld> X(n * n, 0);
  for (int col = 0; col < n; ++col) {
    std::vector<ld> y(n);
    for (int i = 0; i < n; ++i) {
      ld s = (piv[i] == col) ? 1 : 0;
      for (int k = 0; k < i; ++k) s -= A[i * n + k] * y[k];
      y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
      ld s = y[i];
      for (int k = i + 1; k < n; ++k) s -= A[i * n + k] * X[k * n + col];
      X[i * n + col] = s / A[i * n + i];
    }
  }
  return X;
}

// Partial-pivot LU solve / inverse in double, as Eigen's PartialPivLU (src/utils.cpp:99-109).
void lu_solve_inplace(std::vector<double>& A, int n, double* b) {
  std::vector<int> piv(n);
  for (int c = 0; c < n; ++c) {
    int pr = c;
    for (int r = c + 1; r < n; ++r) if (fabs(A[r * n + c]) > fabs(A[pr * n + c])) pr = r;
    piv[c] = pr;
    if (pr != c) for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[pr * n + k]);
    for (int r = c + 1; r < n; ++r) {
      double f = A[r * n + c] / A[c * n + c];
      A[r * n + c] = f;
      for (int k = c + 1; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
    }
  }
  for (int c = 0; c < n; ++c) if (piv[c] != c) std::swap(b[c], b[piv[c]]);
  for (int i = 0; i < n; ++i) for (int k = 0; k < i; ++k) b[i] -= A[i * n + k] * b[k];
  for (int i = n - 1; i >= 0; --i) {
    for (int k = i + 1; k < n; ++k) b[i] -= A[i * n + k] * b[k];
    b[i] /= A[i * n + i];
  }
}
void inverse_inplace(double* M, int n) {
  std::vector<double> LU(M, M + n * n), col(n);
  std::vector<double> out(n * n);
  std::vector<double> work;
  for (int j = 0; j < n; ++j) {
    work = LU;
    for (int i = 0; i < n; ++i) col[i] = (i == j);
    lu_solve_inplace(work, n, col.data());
    for (int i = 0; i < n; ++i) out[i * n + j] = col[i];
  }
  memcpy(M, out.data(), sizeof(double) * n * n);
}

// ------------------------------------------------------------------ Gmsh-order nodes (coordinates)
typedef std::array<ld, 3> P3;
P3 comb(const std::vector<std::pair<ld, P3>>& terms) {
  P3 r{0, 0, 0};
  for (auto& t : terms) for (int c = 0; c < 3; ++c) r[c] += t.first * t.second[c];
  return r;
}
void line_nodes(int p, P3 A, P3 B, std::vector<P3>& out) {
  out.push_back(A);
  out.push_back(B);
  for (int i = 1; i < p; ++i) out.push_back(comb({{(ld)(p - i) / p, A}, {(ld)i / p, B}}));
}
void tri_nodes(int p, P3 A, P3 B, P3 C, std::vector<P3>& out) {
  if (p == 0) { out.push_back(comb({{1.0L / 3, A}, {1.0L / 3, B}, {1.0L / 3, C}})); return; }
  out.push_back(A); out.push_back(B); out.push_back(C);
  P3 V[3] = {A, B, C};
  for (int e = 0; e < 3; ++e)
    for (int i = 1; i < p; ++i) out.push_back(comb({{(ld)(p - i) / p, V[e]}, {(ld)i / p, V[(e + 1) % 3]}}));
  if (p >= 3) {
    P3 A2 = comb({{(ld)(p - 2) / p, A}, {1.0L / p, B}, {1.0L / p, C}});
    P3 B2 = comb({{1.0L / p, A}, {(ld)(p - 2) / p, B}, {1.0L / p, C}});
    P3 C2 = comb({{1.0L / p, A}, {1.0L / p, B}, {(ld)(p - 2) / p, C}});
    tri_nodes(p - 3, A2, B2, C2, out);
  }
}
const int kTE[6][2] = {{0, 1}, {1, 2}, {2, 0}, {3, 0}, {3, 2}, {3, 1}};
const int kTF[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}};
void tet_nodes(int p, const P3* V, std::vector<P3>& out) {
  if (p == 0) { out.push_back(comb({{0.25L, V[0]}, {0.25L, V[1]}, {0.25L, V[2]}, {0.25L, V[3]}})); return; }
  for (int v = 0; v < 4; ++v) out.push_back(V[v]);
  for (int e = 0; e < 6; ++e)
    for (int i = 1; i < p; ++i) out.push_back(comb({{(ld)(p - i) / p, V[kTE[e][0]]}, {(ld)i / p, V[kTE[e][1]]}}));
  if (p >= 3)
    for (int f = 0; f < 4; ++f) {
      P3 F[3] = {V[kTF[f][0]], V[kTF[f][1]], V[kTF[f][2]]};
      P3 G[3];
      for (int k = 0; k < 3; ++k)
        G[k] = comb({{(ld)(p - 2) / p, F[k]}, {1.0L / p, F[(k + 1) % 3]}, {1.0L / p, F[(k + 2) % 3]}});
      tri_nodes(p - 3, G[0], G[1], G[2], out);
    }
  if (p >= 4) {
    P3 W[4];
    for (int k = 0; k < 4; ++k) {
      std::vector<std::pair<ld, P3>> t;
      for (int m = 0; m < 4; ++m) t.push_back({m == k ? (ld)(p - 3) / p : 1.0L / p, V[m]});
      W[k] = comb(t);
    }
    tet_nodes(p - 4, W, out);
  }
}
std::vector<P3> ref_vertices(int dim) {
  if (dim == 1) return {{-1, 0, 0}, {1, 0, 0}};
  if (dim == 2) return {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}};
  return {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
}
std::vector<P3> elem_nodes(int dim, int p) {
  auto V = ref_vertices(dim);
  std::vector<P3> out;
  if (dim == 1) line_nodes(p, V[0], V[1], out);
  if (dim == 2) tri_nodes(p, V[0], V[1], V[2], out);
  if (dim == 3) tet_nodes(p, V.data(), out);
  return out;
}
// face node lists (element-node indices), Gmsh face order, as getElementFaceNodes/EdgeNodes return them
std::vector<std::vector<int>> face_node_lists(int dim, int p) {
  auto N = elem_nodes(dim, p);
  auto V = ref_vertices(dim);
  std::vector<std::vector<int>> fl;
  auto find = [&](const P3& x) {
    for (size_t n = 0; n < N.size(); ++n)
      if (fabsl(N[n][0] - x[0]) + fabsl(N[n][1] - x[1]) + fabsl(N[n][2] - x[2]) < 1e-12L) return (int)n;
    throw std::runtime_error("face node not found");
  };
  if (dim == 1) { fl.push_back({0}); fl.push_back({1}); return fl; }
  if (dim == 2) {
    int E[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (auto& e : E) {
      std::vector<P3> pts;
      line_nodes(p, V[e[0]], V[e[1]], pts);
      std::vector<int> l;
      for (auto& x : pts) l.push_back(find(x));
      fl.push_back(l);
    }
    return fl;
  }
  for (auto& f : kTF) {
    std::vector<P3> pts;
    tri_nodes(p, V[f[0]], V[f[1]], V[f[2]], pts);
    std::vector<int> l;
    for (auto& x : pts) l.push_back(find(x));
    fl.push_back(l);
  }
  return fl;
}

// Lagrange basis on a node set via the modal Vandermonde; returns value/gradient evaluator.
struct Lagrange {
  Modal modal;
  int n;
  std::vector<ld> Vinv;  // n x n
  Lagrange(int dim, int p, const std::vector<P3>& nodes) : modal(dim, p), n((int)nodes.size()) {
    std::vector<ld> V(n * n);
    std::vector<Du> out;
    for (int i = 0; i < n; ++i) {
      modal.eval(nodes[i].data(), out);
      for (int m = 0; m < n; ++m) V[i * n + m] = out[m].v;
    }
    Vinv = inv_ld(V, n);
  }
  // phi[i], grad[i][3] (Gmsh reference coordinates)
  void eval(const ld* x, double* phi, double (*grad)[3]) const {
    std::vector<Du> out;
    modal.eval(x, out);
    for (int i = 0; i < n; ++i) {
      ld s = 0, g[3] = {0, 0, 0};
      for (int m = 0; m < n; ++m) {
        ld c = Vinv[m * n + i];
        s += c * out[m].v;
        for (int k = 0; k < 3; ++k) g[k] += c * out[m].d[k];
      }
      phi[i] = (double)s;
      if (grad) for (int k = 0; k < 3; ++k) grad[i][k] = (double)g[k];
    }
  }
};

// Collapsed Gauss-Legendre rules on the Gmsh reference simplex; n points per direction,
// points listed with the first collapsed coordinate outermost, ascending.
void gl(int n, std::vector<ld>& x, std::vector<ld>& w) {
  x.resize(n); w.resize(n);
  for (int i = 0; i < n; ++i) {
    ld z = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
    for (int it = 0; it < 200; ++it) {
      ld p0 = 1, p1 = z;
      for (int k = 2; k <= n; ++k) { ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      if (n == 1) { p0 = 1; p1 = z; }
      ld dp = n * (z * p1 - p0) / (z * z - 1);
      ld dz = p1 / dp;
      z -= dz;
      if (fabsl(dz) < 1e-20L) {
        p0 = 1; p1 = z;
        for (int k = 2; k <= n; ++k) { ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
        if (n == 1) { p0 = 1; p1 = z; }
        dp = n * (z * p1 - p0) / (z * z - 1);
        break;
      }
    }
    ld p0 = 1, p1 = z;
    for (int k = 2; k <= n; ++k) { ld p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    if (n == 1) { p0 = 1; p1 = z; }
    ld dp = n * (z * p1 - p0) / (z * z - 1);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2 / ((1 - z * z) * dp * dp);
  }
}
void rule(int dim, int n, std::vector<P3>& pts, std::vector<ld>& wts) {
  pts.clear(); wts.clear();
  std::vector<ld> x, w;
  gl(n, x, w);
  if (dim == 0) { pts.push_back({0, 0, 0}); wts.push_back(1); return; }
  if (dim == 1) { for (int i = 0; i < n; ++i) { pts.push_back({x[i], 0, 0}); wts.push_back(w[i]); } return; }
  std::vector<ld> t(n), tw(n);
  for (int i = 0; i < n; ++i) { t[i] = (x[i] + 1) / 2; tw[i] = w[i] / 2; }
  if (dim == 2) {
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {
      pts.push_back({t[i], (1 - t[i]) * t[j], 0});
      wts.push_back(tw[i] * tw[j] * (1 - t[i]));
    }
    return;
  }
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) for (int k = 0; k < n; ++k) {
    pts.push_back({t[i], (1 - t[i]) * t[j], (1 - t[i]) * (1 - t[j]) * t[k]});
    wts.push_back(tw[i] * tw[j] * tw[k] * (1 - t[i]) * (1 - t[i]) * (1 - t[j]));
  }
}

struct VecHash {
  size_t operator()(const std::vector<int64_t>& v) const {
    uint64_t h = 1469598103934665603ull;
    for (auto x : v) { h ^= (uint64_t)x + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); }
    return (size_t)h;
  }
};

}  // namespace

// ==================================================================== the oracle state
struct orc_ctx {
  int dim, p, Np, Nf, Nfp, G, Gf, Gs, fc;
  bool lit;  // literal (quadrature-point) arrays present
  int64_t K, F;
  int sign_mode, nthreads;
  double c0, rho0, v0[3];
  std::vector<int64_t> elNodeTags;
  std::vector<double> nodeXYZ;  // K*Np*3
  std::unordered_map<int64_t, std::array<double, 3>> tagXYZ;
  // element quadrature data
  std::vector<double> w, basis, ugrad;  // G, G*Np, G*Np*3
  std::vector<double> elJac, elDet, grad;  // K*G*9, K*G, K*G*Np*3
  // faces
  std::vector<std::vector<int>> fl;       // local face node lists
  std::vector<int64_t> elFNodeTags;       // K*Nf*Nfp
  std::vector<int64_t> fNodeTags;         // F*Nfp
  std::vector<int> elFIds, elFOrient, elFOrientOrig;
  std::vector<std::vector<int>> fNbr, fNToEl;
  std::vector<double> fw, fbasis, fugrad;  // Gf, Gf*Nfp, Gf*Nfp*3
  std::vector<double> fJac, fDet, fNormals;  // F*Gf*9, F*Gf, F*Gf*3
  std::vector<char> fIsB;
  std::vector<int> fBC;
  std::vector<std::vector<double>> RKR;
  std::vector<double> Minv;  // K*Np*Np
  // factorized data
  std::vector<double> Mhat_inv, W, Mf;  // Np*Np, dim*Np*Np (W_r = Mhat^-1 S_r), Nfp*Nfp
  std::vector<double> Jinv, absdet;     // K*9, K
  // literal-mode scratch, as the reference's containers
  std::vector<std::vector<std::vector<double>>> Flux, FluxGhost;
  std::vector<std::vector<double>> uGhost;
  std::vector<double> fFlux;
};

namespace {

int64_t pos_of(const orc_ctx* c, int64_t el, int n) { return el * c->Np + n; }

void setup(orc_ctx* c, const int* grefl, const int64_t* const* gnodes, const int64_t* gsz, int ngroups,
           int literal_setup) {
  const int dim = c->dim, p = c->p, Np = c->Np;
  const int64_t K = c->K;
  auto nodes = elem_nodes(dim, p);
  if ((int)nodes.size() != Np) throw std::runtime_error("node count mismatch");
  Lagrange L(dim, p, nodes);
  // element rule: n = p + 3 points per collapsed direction (exact beyond degree 2p)
  std::vector<P3> q;
  std::vector<ld> qw;
  rule(dim, p + 3, q, qw);
  c->G = (int)q.size();
  const int G = c->G;
  c->w.resize(G); c->basis.resize(G * Np); c->ugrad.resize(G * Np * 3);
  for (int g = 0; g < G; ++g) {
    c->w[g] = (double)qw[g];
    L.eval(q[g].data(), &c->basis[g * Np], reinterpret_cast<double(*)[3]>(&c->ugrad[g * Np * 3]));
  }
  // Jacobians (isoparametric), determinants and physical gradients (src/Mesh.cpp:32-72)
  c->lit = !(literal_setup & 2);
  if (c->lit) {
    c->elJac.assign(K * G * 9, 0); c->elDet.assign(K * G, 0); c->grad.assign(K * G * Np * 3, 0);
  }
  #pragma omp parallel for schedule(static) num_threads(c->nthreads)
  for (int64_t el = 0; el < (c->lit ? K : 0); ++el) {
    std::vector<double> A(dim * dim);
    for (int g = 0; g < G; ++g) {
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[u][x] = d x_x / d u_u
      for (int n = 0; n < Np; ++n)
        for (int u = 0; u < dim; ++u)
          for (int x = 0; x < 3; ++x) J[u][x] += c->nodeXYZ[(el * Np + n) * 3 + x] * c->ugrad[(g * Np + n) * 3 + u];
      double det;
      if (dim == 3) {
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      } else if (dim == 2) {
        double n0 = J[0][1] * J[1][2] - J[0][2] * J[1][1], n1 = J[0][2] * J[1][0] - J[0][0] * J[1][2],
               n2 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        det = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      } else {
        det = sqrt(J[0][0] * J[0][0] + J[0][1] * J[0][1] + J[0][2] * J[0][2]);
      }
      for (int u = 0; u < 3; ++u) for (int x = 0; x < 3; ++x) c->elJac[(el * G + g) * 9 + u * 3 + x] = J[u][x];
      c->elDet[el * G + g] = det;
      for (int n = 0; n < Np; ++n) {
        // jacobian[i*dim+j] = J(x=i,u=j), mapped column-major => A(r,col) = J(x=col,u=r)
        for (int r = 0; r < dim; ++r) for (int col = 0; col < dim; ++col) A[r * dim + col] = J[r][col];
        double b[3] = {0, 0, 0};
        for (int u = 0; u < dim; ++u) b[u] = c->ugrad[(g * Np + n) * 3 + u];
        std::vector<double> Ac = A;
        lu_solve_inplace(Ac, dim, b);
        for (int x = 0; x < dim; ++x) c->grad[((el * G + g) * Np + n) * 3 + x] = b[x];
      }
    }
  }
  // ---------------------------------------------------------------- faces
  c->fl = face_node_lists(dim, p);
  c->Nf = (int)c->fl.size();
  c->Nfp = (int)c->fl[0].size();
  const int Nf = c->Nf, Nfp = c->Nfp;
  c->elFNodeTags.resize(K * Nf * Nfp);
  for (int64_t el = 0; el < K; ++el)
    for (int f = 0; f < Nf; ++f)
      for (int k = 0; k < Nfp; ++k) c->elFNodeTags[(el * Nf + f) * Nfp + k] = c->elNodeTags[el * Np + c->fl[f][k]];
  // unique faces, first occurrence kept (getUniqueFaceNodeTags, src/Mesh.cpp:679-713)
  std::vector<int64_t> ordered(c->elFNodeTags);
  for (int64_t i = 0; i < K * Nf; ++i) std::sort(ordered.begin() + i * Nfp, ordered.begin() + (i + 1) * Nfp);
  c->elFIds.assign(K * Nf, -1);
  if (literal_setup & 1) {
    // the reference's quadratic scan and erase (tiny meshes only)
    std::vector<int64_t> fo(ordered), fu(c->elFNodeTags);
    size_t io = 0, iu = 0;
    while (io < fo.size()) {
      size_t id = io + Nfp, idu = iu + Nfp;
      for (; id < fo.size(); id += Nfp, idu += Nfp)
        if (std::equal(fo.begin() + io, fo.begin() + io + Nfp, fo.begin() + id)) break;
      if (id < fo.size()) {
        fo.erase(fo.begin() + id, fo.begin() + id + Nfp);
        fu.erase(fu.begin() + idu, fu.begin() + idu + Nfp);
      } else { io += Nfp; iu += Nfp; }
    }
    c->F = (int64_t)fo.size() / Nfp;
    c->fNodeTags = fu;
    c->fNbr.assign(c->F, {});
    for (int64_t el = 0; el < K; ++el)
      for (int f = 0; f < Nf; ++f)
        for (int64_t g = 0; g < c->F; ++g)
          if (std::equal(fo.begin() + g * Nfp, fo.begin() + (g + 1) * Nfp, ordered.begin() + (el * Nf + f) * Nfp)) {
            c->elFIds[el * Nf + f] = (int)g;
            c->fNbr[g].push_back((int)el);
          }
  } else {
    std::unordered_map<std::vector<int64_t>, int, VecHash> id;
    id.reserve(K * Nf);
    c->fNodeTags.clear();
    std::vector<int64_t> key(Nfp);
    for (int64_t el = 0; el < K; ++el)
      for (int f = 0; f < Nf; ++f) {
        std::copy(ordered.begin() + (el * Nf + f) * Nfp, ordered.begin() + (el * Nf + f + 1) * Nfp, key.begin());
        auto it = id.find(key);
        int fid;
        if (it == id.end()) {
          fid = (int)(c->fNodeTags.size() / Nfp);
          id.emplace(key, fid);
          c->fNodeTags.insert(c->fNodeTags.end(), c->elFNodeTags.begin() + (el * Nf + f) * Nfp,
                              c->elFNodeTags.begin() + (el * Nf + f + 1) * Nfp);
          c->fNbr.push_back({});
        } else fid = it->second;
        c->elFIds[el * Nf + f] = fid;
        c->fNbr[fid].push_back((int)el);
      }
    c->F = (int64_t)c->fNodeTags.size() / Nfp;
  }
  const int64_t F = c->F;
  for (int64_t f = 0; f < F; ++f)
    if (c->fNbr[f].size() > 2) throw std::runtime_error("non-manifold face shared by more than two elements");
  // face-node -> element-node maps by tag equality (src/Mesh.cpp:253-263)
  c->fNToEl.assign(F, {});
  for (int64_t f = 0; f < F; ++f)
    for (int nf = 0; nf < Nfp; ++nf)
      for (int el : c->fNbr[f])
        for (int nel = 0; nel < Np; ++nel)
          if (c->fNodeTags[f * Nfp + nf] == c->elNodeTags[el * Np + nel]) c->fNToEl[f].push_back(nel);
  for (int64_t f = 0; f < F; ++f)
    if ((int)c->fNToEl[f].size() != Nfp * (int)c->fNbr[f].size())
      throw std::runtime_error("face nodes do not match element nodes by tag");
  // face quadrature, basis, Jacobians, gradients, normals (src/Mesh.cpp:124-194)
  const int fd = dim - 1;
  std::vector<P3> fq;
  std::vector<ld> fqw;
  rule(fd, p + 2, fq, fqw);
  c->Gf = (int)fq.size();
  const int Gf = c->Gf;
  c->fw.resize(Gf); c->fbasis.resize(Gf * Nfp); c->fugrad.assign(Gf * Nfp * 3, 0);
  if (fd == 0) {
    c->fw[0] = 1; c->fbasis[0] = 1;
  } else {
    auto fn = elem_nodes(fd, p);
    Lagrange FL(fd, p, fn);
    for (int g = 0; g < Gf; ++g) {
      c->fw[g] = (double)fqw[g];
      FL.eval(fq[g].data(), &c->fbasis[g * Nfp], reinterpret_cast<double(*)[3]>(&c->fugrad[g * Nfp * 3]));
    }
  }
  // without the literal arrays only the first face point is kept (normal and measure are
  // constant on straight faces and the normal is aligned to that point anyway)
  c->Gs = c->lit ? Gf : 1;
  const int Gs = c->Gs;
  c->fJac.assign(F * Gs * 9, 0); c->fDet.assign(F * Gs, 0); c->fNormals.assign(F * Gs * 3, 0);
  for (int64_t f = 0; f < F; ++f) {
    std::vector<std::array<double, 3>> fgrad(Gs * Nfp);
    for (int g = 0; g < Gs; ++g) {
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      double det = 1;
      if (fd >= 1) {
        for (int n = 0; n < Nfp; ++n) {
          auto& X = c->tagXYZ.at(c->fNodeTags[f * Nfp + n]);
          for (int u = 0; u < fd; ++u) for (int x = 0; x < 3; ++x) J[u][x] += X[x] * c->fugrad[(g * Nfp + n) * 3 + u];
        }
        if (fd == 1) {  // complete a line: in-plane normal, then their cross product
          double* a = J[0];
          double b[3];
          if ((fabs(a[0]) >= fabs(a[1]) && fabs(a[0]) >= fabs(a[2])) || (fabs(a[1]) >= fabs(a[0]) && fabs(a[1]) >= fabs(a[2]))) {
            b[0] = a[1]; b[1] = -a[0]; b[2] = 0;
          } else { b[0] = 0; b[1] = a[2]; b[2] = -a[1]; }
          double nb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
          for (int x = 0; x < 3; ++x) J[1][x] = b[x] / nb;
          double cc[3] = {a[1] * J[1][2] - a[2] * J[1][1], a[2] * J[1][0] - a[0] * J[1][2], a[0] * J[1][1] - a[1] * J[1][0]};
          double ncc = sqrt(cc[0] * cc[0] + cc[1] * cc[1] + cc[2] * cc[2]);
          for (int x = 0; x < 3; ++x) J[2][x] = cc[x] / ncc;
          det = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        } else {  // complete a triangle with its unit normal
          double n0 = J[0][1] * J[1][2] - J[0][2] * J[1][1], n1 = J[0][2] * J[1][0] - J[0][0] * J[1][2],
                 n2 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
          det = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
          J[2][0] = n0 / det; J[2][1] = n1 / det; J[2][2] = n2 / det;
        }
        for (int u = 0; u < 3; ++u) for (int x = 0; x < 3; ++x) c->fJac[(f * Gs + g) * 9 + u * 3 + x] = J[u][x];
        for (int n = 0; n < Nfp; ++n) {
          std::vector<double> A(dim * dim);
          for (int r = 0; r < dim; ++r) for (int col = 0; col < dim; ++col) A[r * dim + col] = J[r][col];
          double b[3] = {0, 0, 0};
          for (int u = 0; u < dim; ++u) b[u] = c->fugrad[(g * Nfp + n) * 3 + u];
          lu_solve_inplace(A, dim, b);
          fgrad[g * Nfp + n] = {dim > 0 ? b[0] : 0, dim > 1 ? b[1] : 0, dim > 2 ? b[2] : 0};
        }
      }
      c->fDet[f * Gs + g] = det;
      double nrm[3] = {0, 0, 0};
      if (fd == 0) { nrm[0] = 1; }
      else if (fd == 1) {
        auto& g0 = fgrad[g * Nfp + 0];
        // grad(phi_0) x (0,0,-1)
        nrm[0] = g0[1] * -1 - g0[2] * 0; nrm[1] = g0[2] * 0 - g0[0] * -1; nrm[2] = g0[0] * 0 - g0[1] * 0;
        auto& r0 = fgrad[0 * Nfp + 0];
        if (g0[0] * r0[0] + g0[1] * r0[1] + g0[2] * r0[2] < 0) for (int x = 0; x < 3; ++x) nrm[x] = -nrm[x];
      } else {
        auto& a = fgrad[g * Nfp + 0];
        auto& b = fgrad[g * Nfp + 1];
        nrm[0] = a[1] * b[2] - a[2] * b[1]; nrm[1] = a[2] * b[0] - a[0] * b[2]; nrm[2] = a[0] * b[1] - a[1] * b[0];
        if (g != 0) {
          double* n0 = &c->fNormals[(f * Gs + 0) * 3];
          if (n0[0] * nrm[0] + n0[1] * nrm[1] + n0[2] * nrm[2] < 0) for (int x = 0; x < 3; ++x) nrm[x] = -nrm[x];
        }
      }
      double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      for (int x = 0; x < 3; ++x) c->fNormals[(f * Gs + g) * 3 + x] = nrm[x] / nn;
    }
  }
  c->fc = (dim == 3 && p != 1) ? -1 : 1;  // src/Mesh.cpp:210-211
  // orientation: sign((x_faceNode0 - barycenter) . n_g0) (src/Mesh.cpp:278-294)
  c->elFOrient.assign(K * Nf, 0);
  for (int64_t el = 0; el < K; ++el) {
    double bc[3] = {0, 0, 0};
    for (int v = 0; v <= dim; ++v) for (int x = 0; x < 3; ++x) bc[x] += c->nodeXYZ[(el * Np + v) * 3 + x] / (dim + 1);
    for (int f = 0; f < Nf; ++f) {
      auto& X = c->tagXYZ.at(c->elFNodeTags[(el * Nf + f) * Nfp]);
      int fid = c->elFIds[el * Nf + f];
      double dp = 0;
      for (int x = 0; x < 3; ++x) dp += (X[x] - bc[x]) * c->fNormals[(fid * Gs + 0) * 3 + x];
      c->elFOrient[el * Nf + f] = dp >= 0 ? 1 : -1;
    }
  }
  c->elFOrientOrig = c->elFOrient;
  // (the neighbour reorder at src/Mesh.cpp:301-314 tests the face count, not the side count, and never fires)
  // boundary flags, outward boundary normals (src/Mesh.cpp:336-353)
  c->fIsB.assign(F, 0);
  for (int64_t f = 0; f < F; ++f) {
    if (c->fNbr[f].size() < 2) {
      c->fIsB[f] = 1;
      int el = c->fNbr[f][0];
      for (int lf = 0; lf < Nf; ++lf)
        if (c->elFIds[el * Nf + lf] == f) {
          int o = c->elFOrient[el * Nf + lf];
          for (int g = 0; g < Gs; ++g) for (int x = 0; x < 3; ++x) c->fNormals[(f * Gs + g) * 3 + x] *= o;
          c->elFOrient[el * Nf + lf] = 1;
        }
    }
  }
  // BC by first face node, ascending physical tag, last wins (src/Mesh.cpp:364-384)
  c->fBC.assign(F, 0);
  for (int gi = 0; gi < ngroups; ++gi) {
    const int64_t* nb = gnodes[gi];
    int64_t ns = gsz[gi];
    for (int64_t f = 0; f < F; ++f) {
      if (!c->fIsB[f]) continue;
      int64_t t = c->fNodeTags[f * Nfp];
      if (std::find(nb, nb + ns, t) != nb + ns) c->fBC[f] = grefl[gi] ? 1 : 0;
    }
  }
  // RKR (src/Mesh.cpp:391-418)
  c->RKR.assign(c->lit ? F * Gf : 0, {});
  for (int64_t f = 0; f < (c->lit ? F : 0); ++f) {
    if (c->fBC[f] != 0) continue;
    for (int g = 0; g < Gf; ++g) {
      const double* n = &c->fNormals[(f * Gf + g) * 3];
      const double cc = c->c0, r = c->rho0;
      const double a = ORC_ABS_COEF;
      c->RKR[f * Gf + g] = {a * cc, a * cc * cc * r * n[0], a * cc * cc * r * n[1], a * cc * cc * r * n[2],
                            a * n[0] / r, a * cc * n[0] * n[0], a * cc * n[0] * n[1], a * cc * n[0] * n[2],
                            a * n[1] / r, a * cc * n[1] * n[0], a * cc * n[1] * n[1], a * cc * n[1] * n[2],
                            a * n[2] / r, a * cc * n[2] * n[0], a * cc * n[2] * n[1], a * cc * n[2] * n[2]};
    }
  }
  // mass matrices and inverses (src/Mesh.cpp:440-466)
  c->Minv.assign(c->lit ? K * Np * Np : 0, 0);
  #pragma omp parallel for schedule(static) num_threads(c->nthreads)
  for (int64_t el = 0; el < (c->lit ? K : 0); ++el) {
    double* M = &c->Minv[el * Np * Np];
    for (int i = 0; i < Np; ++i)
      for (int j = 0; j < Np; ++j) {
        double s = 0;
        for (int g = 0; g < G; ++g) s += c->basis[g * Np + i] * c->basis[g * Np + j] * c->w[g] * c->elDet[el * G + g];
        M[i * Np + j] = s;
      }
    inverse_inplace(M, Np);
  }
  // factorized operators from the oracle's own basis
  {
    std::vector<ld> Mh(Np * Np, 0), Sh(dim * Np * Np, 0);
    for (int g = 0; g < G; ++g)
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
          Mh[i * Np + j] += qw[g] * (ld)c->basis[g * Np + i] * (ld)c->basis[g * Np + j];
          for (int r = 0; r < dim; ++r)
            Sh[(r * Np + i) * Np + j] += qw[g] * (ld)c->ugrad[(g * Np + i) * 3 + r] * (ld)c->basis[g * Np + j];
        }
    auto Mi = inv_ld(Mh, Np);
    c->Mhat_inv.resize(Np * Np);
    for (int i = 0; i < Np * Np; ++i) c->Mhat_inv[i] = (double)Mi[i];
    c->W.assign(dim * Np * Np, 0);
    for (int r = 0; r < dim; ++r)
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
          ld s = 0;
          for (int k = 0; k < Np; ++k) s += Mi[i * Np + k] * Sh[(r * Np + k) * Np + j];
          c->W[(r * Np + i) * Np + j] = (double)s;
        }
    c->Mf.assign(Nfp * Nfp, 0);
    for (int g = 0; g < Gf; ++g)
      for (int i = 0; i < Nfp; ++i)
        for (int j = 0; j < Nfp; ++j) c->Mf[i * Nfp + j] += c->fw[g] * c->fbasis[g * Nfp + i] * c->fbasis[g * Nfp + j];
    c->Jinv.assign(K * 9, 0);
    c->absdet.assign(K, 0);
    for (int64_t el = 0; el < K; ++el) {
      // affine map from the vertices: J[x][u] = X_{u+1} - X_0 (Gmsh reference coordinates)
      double J[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
      const double* X = &c->nodeXYZ[el * Np * 3];
      for (int u = 0; u < dim; ++u)
        for (int x = 0; x < dim; ++x) J[x][u] = (X[(u + 1) * 3 + x] - X[x]) / (dim == 1 ? 2.0 : 1.0);
      double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      double inv[3][3];
      inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
      inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
      inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
      inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
      for (int r = 0; r < 3; ++r) for (int x = 0; x < 3; ++x) c->Jinv[el * 9 + r * 3 + x] = inv[r][x];
      c->absdet[el] = fabs(det);
    }
  }
  // literal scratch (src/solver.cpp:180-182, src/Mesh.cpp:426-431)
  if (c->lit) {
    c->Flux.assign(4, std::vector<std::vector<double>>(K * Np, std::vector<double>(3)));
    c->FluxGhost.assign(4, std::vector<std::vector<double>>(F * Gf, std::vector<double>(3)));
    c->uGhost.assign(4, std::vector<double>(F * Gf));
    c->fFlux.assign(F * Nfp, 0);
  }
}

inline int nSides(const orc_ctx* c, int64_t f) { return (int)c->fNbr[f].size(); }
inline int fN2E(const orc_ctx* c, int64_t f, int nf, int s) { return c->fNToEl[f][nf * nSides(c, f) + s]; }

// penalty coefficient of an interior face
inline double face_fc(const orc_ctx* c, int64_t f) {
  if (c->sign_mode == ORC_SIGN_STABLE) {
    int up = c->fNbr[f][0];
    for (int lf = 0; lf < c->Nf; ++lf)
      if (c->elFIds[up * c->Nf + lf] == f) return c->elFOrientOrig[up * c->Nf + lf];
  }
  return c->fc;
}

// Mesh::updateFlux (src/Mesh.cpp:569-674), serial as in the reference
void update_flux(orc_ctx* c, std::vector<std::vector<double>>& u) {
  const double* v0 = c->v0;
  const double r = c->rho0, cc = c->c0;
  const int Np = c->Np, Nf = c->Nf, Gf = c->Gf, Nfp = c->Nfp;
  for (int64_t el = 0; el < c->K; ++el) {
    for (int n = 0; n < Np; ++n) {
      int64_t i = el * Np + n;
      c->Flux[0][i] = {v0[0] * u[0][i] + r * cc * cc * u[1][i], v0[1] * u[0][i] + r * cc * cc * u[2][i],
                       v0[2] * u[0][i] + r * cc * cc * u[3][i]};
      c->Flux[1][i] = {v0[0] * u[1][i] + u[0][i] / r, v0[1] * u[1][i], v0[2] * u[1][i]};
      c->Flux[2][i] = {v0[0] * u[2][i], v0[1] * u[2][i] + u[0][i] / r, v0[2] * u[2][i]};
      c->Flux[3][i] = {v0[0] * u[3][i], v0[1] * u[3][i], v0[2] * u[3][i] + u[0][i] / r};
    }
    for (int f = 0; f < Nf; ++f) {
      int64_t fId = c->elFIds[el * Nf + f];
      if (!c->fIsB[fId]) continue;
      for (int g = 0; g < Gf; ++g) {
        int64_t gId = fId * Gf + g;
        for (int q = 0; q < 4; ++q) c->uGhost[q][gId] = 0;
        for (int n = 0; n < Nfp; ++n) {
          int64_t nId = el * Np + fN2E(c, fId, n, 0);
          for (int q = 0; q < 4; ++q) c->uGhost[q][gId] += u[q][nId] * c->fbasis[g * Nfp + n];
        }
        const double* nr = &c->fNormals[gId * 3];
        if (c->fBC[fId] == 1) {
          double dot = nr[0] * c->uGhost[1][gId] + nr[1] * c->uGhost[2][gId] + nr[2] * c->uGhost[3][gId];
          c->uGhost[1][gId] -= dot * nr[0];
          c->uGhost[2][gId] -= dot * nr[1];
          c->uGhost[3][gId] -= dot * nr[2];
          double g0 = c->uGhost[0][gId], g1 = c->uGhost[1][gId], g2 = c->uGhost[2][gId], g3 = c->uGhost[3][gId];
          c->FluxGhost[0][gId] = {v0[0] * g0 + r * cc * cc * g1, v0[1] * g0 + r * cc * cc * g2, v0[2] * g0 + r * cc * cc * g3};
          c->FluxGhost[1][gId] = {v0[0] * g1 + g0 / r, v0[1] * g1, v0[2] * g1};
          c->FluxGhost[2][gId] = {v0[0] * g2, v0[1] * g2 + g0 / r, v0[2] * g2};
          c->FluxGhost[3][gId] = {v0[0] * g3, v0[1] * g3, v0[2] * g3 + g0 / r};
          for (int q = 0; q < 4; ++q) {  // eigen::dot over m_Dim = 3 components
            auto& Fg = c->FluxGhost[q][gId];
            Fg[0] = nr[0] * Fg[0] + nr[1] * Fg[1] + nr[2] * Fg[2];
          }
        } else {
          const auto& R = c->RKR[gId];
          double g0 = c->uGhost[0][gId], g1 = c->uGhost[1][gId], g2 = c->uGhost[2][gId], g3 = c->uGhost[3][gId];
          c->FluxGhost[0][gId][0] = R[0] * g0 + R[1] * g1 + R[2] * g2 + R[3] * g3;
          c->FluxGhost[1][gId][0] = R[4] * g0 + R[5] * g1 + R[6] * g2 + R[7] * g3;
          c->FluxGhost[2][gId][0] = R[8] * g0 + R[9] * g1 + R[10] * g2 + R[11] * g3;
          c->FluxGhost[3][gId][0] = R[12] * g0 + R[13] * g1 + R[14] * g2 + R[15] * g3;
        }
      }
    }
  }
}


// One face of Mesh::precomputeFlux (src/Mesh.cpp:510-537): quadrature form.
inline void face_flux_literal(orc_ctx* c, const std::vector<double>& u, const std::vector<std::vector<double>>& Flux,
                              int eq, int64_t f, double* FIntPts) {
  const int Gf = c->Gf, Nfp = c->Nfp, Np = c->Np;
  for (int g = 0; g < Gf; ++g) FIntPts[g] = 0;
  if (c->fIsB[f]) {
    for (int g = 0; g < Gf; ++g) FIntPts[g] = c->FluxGhost[eq][f * Gf + g][0];
  } else {
    double fcf = face_fc(c, f);
    for (int i = 0; i < Nfp; ++i) {
      int64_t up = (int64_t)c->fNbr[f][0] * Np + fN2E(c, f, i, 0);
      int64_t dn = (int64_t)c->fNbr[f][1] * Np + fN2E(c, f, i, 1);
      for (int g = 0; g < Gf; ++g) {
        const double* n = &c->fNormals[(f * Gf + g) * 3];
        double Fn[3];
        for (int x = 0; x < 3; ++x) Fn[x] = 0.5 * ((Flux[up][x] + Flux[dn][x]) + fcf * c->c0 * n[x] * (u[up] - u[dn]));
        FIntPts[g] += (n[0] * Fn[0] + n[1] * Fn[1] + n[2] * Fn[2]) * c->fbasis[g * Nfp + i];
      }
    }
  }
  for (int n = 0; n < Nfp; ++n) {
    double s = 0;
    for (int g = 0; g < Gf; ++g) s += c->fw[g] * c->fbasis[g * Nfp + n] * FIntPts[g] * c->fDet[f * Gf + g];
    c->fFlux[f * Nfp + n] = s;
  }
}

// Mesh::precomputeFlux. nested=true reproduces the reference's nested parallel region, in
// which every outer thread runs the whole face loop (src/Mesh.cpp:502,509).
void precompute_flux(orc_ctx* c, const std::vector<double>& u, const std::vector<std::vector<double>>& Flux, int eq,
                     bool nested) {
  if (nested) {
    #pragma omp parallel num_threads(c->nthreads)
    {
      std::vector<double> FIntPts(c->Gf, 0);
      #pragma omp parallel for schedule(static)
      for (int64_t f = 0; f < c->F; ++f) face_flux_literal(c, u, Flux, eq, f, FIntPts.data());
    }
  } else {
    #pragma omp parallel num_threads(c->nthreads)
    {
      std::vector<double> FIntPts(c->Gf, 0);
      #pragma omp for schedule(static)
      for (int64_t f = 0; f < c->F; ++f) face_flux_literal(c, u, Flux, eq, f, FIntPts.data());
    }
  }
}

// Mesh::getElFlux (src/Mesh.cpp:548-557)
inline void el_flux(const orc_ctx* c, int64_t el, double* Fo) {
  std::fill(Fo, Fo + c->Np, 0.0);
  for (int f = 0; f < c->Nf; ++f) {
    int fid = c->elFIds[el * c->Nf + f];
    int side = (el == c->fNbr[fid][0]) ? 0 : 1;
    for (int nf = 0; nf < c->Nfp; ++nf) Fo[fN2E(c, fid, nf, side)] += c->elFOrient[el * c->Nf + f] * c->fFlux[fid * c->Nfp + nf];
  }
}

// solver::numStep (src/solver.cpp:35-52) with getElStiffVector (src/Mesh.cpp:476-489) and linEq.
void num_step_literal(orc_ctx* c, std::vector<std::vector<double>>& u, double dt, double beta, bool nested) {
  const int Np = c->Np, G = c->G;
  for (int eq = 0; eq < 4; ++eq) {
    precompute_flux(c, u[eq], c->Flux[eq], eq, nested);
    std::vector<double> elFlux(Np), elStiff(Np);
    #pragma omp parallel for schedule(static) firstprivate(elFlux, elStiff) num_threads(c->nthreads)
    for (int64_t el = 0; el < c->K; ++el) {
      el_flux(c, el, elFlux.data());
      for (int i = 0; i < Np; ++i) {
        elStiff[i] = 0.0;
        for (int j = 0; j < Np; ++j) {
          const std::vector<double>& Fj = c->Flux[eq][el * Np + j];
          for (int g = 0; g < G; ++g) {
            const double* gr = &c->grad[((el * G + g) * Np + i) * 3];
            elStiff[i] += (Fj[0] * gr[0] + Fj[1] * gr[1] + Fj[2] * gr[2]) * c->basis[g * Np + j] * c->w[g] *
                          c->elDet[el * G + g];
          }
        }
      }
      for (int i = 0; i < Np; ++i) elStiff[i] -= elFlux[i];
      // Y = beta*Y + dt*A*X with the row-major inverse read column-major (src/utils.cpp:118-123)
      const double* A = &c->Minv[el * Np * Np];
      double* Y = &u[eq][el * Np];
      std::vector<double> AX(Np, 0.0);
      for (int col = 0; col < Np; ++col)
        for (int row = 0; row < Np; ++row) AX[row] += A[col * Np + row] * elStiff[col];
      for (int i = 0; i < Np; ++i) Y[i] = beta * Y[i] + dt * AX[i];
    }
  }
}

// Factorized evaluation of the same operator on affine elements: out = dt*L(u) (+ beta*u).
void num_step_factorized(orc_ctx* c, const double* u, double* out, double dt, double beta) {
  const int Np = c->Np, Nfp = c->Nfp, Nf = c->Nf, dim = c->dim;
  const int64_t K = c->K, F = c->F, KN = K * Np;
  const double r = c->rho0, cc = c->c0, *v0 = c->v0;
  auto nflux = [&](const double* n, const double* q, double* o) {  // n . F(q) for the 4 equations
    double vn = v0[0] * n[0] + v0[1] * n[1] + v0[2] * n[2];
    double un = q[1] * n[0] + q[2] * n[1] + q[3] * n[2];
    o[0] = vn * q[0] + r * cc * cc * un;
    for (int k = 0; k < 3; ++k) o[1 + k] = vn * q[1 + k] + q[0] / r * n[k];
  };
  std::vector<double> fF(4 * F * Nfp);
  #pragma omp parallel for schedule(static) num_threads(c->nthreads)
  for (int64_t f = 0; f < F; ++f) {
    const double* n = &c->fNormals[(f * c->Gs) * 3];
    double nodal[4][32];
    for (int i = 0; i < Nfp; ++i) {
      double qm[4], o[4];
      int64_t im = (int64_t)c->fNbr[f][0] * Np + fN2E(c, f, i, 0);
      for (int q = 0; q < 4; ++q) qm[q] = u[q * KN + im];
      if (!c->fIsB[f]) {
        double qp[4], op[4];
        int64_t ip = (int64_t)c->fNbr[f][1] * Np + fN2E(c, f, i, 1);
        for (int q = 0; q < 4; ++q) qp[q] = u[q * KN + ip];
        nflux(n, qm, o);
        nflux(n, qp, op);
        double fcf = face_fc(c, f), nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        for (int q = 0; q < 4; ++q) nodal[q][i] = 0.5 * (o[q] + op[q]) + 0.5 * fcf * cc * nn * (qm[q] - qp[q]);
      } else if (c->fBC[f] == 1) {
        double dot = n[0] * qm[1] + n[1] * qm[2] + n[2] * qm[3];
        for (int k = 0; k < 3; ++k) qm[1 + k] -= dot * n[k];
        nflux(n, qm, o);
        for (int q = 0; q < 4; ++q) nodal[q][i] = o[q];
      } else {
        double un = n[0] * qm[1] + n[1] * qm[2] + n[2] * qm[3];
        const double a = ORC_ABS_COEF;
        nodal[0][i] = a * cc * qm[0] + a * cc * cc * r * un;
        for (int k = 0; k < 3; ++k) nodal[1 + k][i] = a * n[k] / r * qm[0] + a * cc * n[k] * un;
      }
    }
    double sJ = c->fDet[f * c->Gs];
    for (int q = 0; q < 4; ++q)
      for (int a = 0; a < Nfp; ++a) {
        double s = 0;
        for (int b = 0; b < Nfp; ++b) s += c->Mf[a * Nfp + b] * nodal[q][b];
        fF[(q * F + f) * Nfp + a] = sJ * s;
      }
  }
  #pragma omp parallel for schedule(static) num_threads(c->nthreads)
  for (int64_t el = 0; el < K; ++el) {
    const double* Ji = &c->Jinv[el * 9];
    double Ft[3][4][64];  // contravariant fluxes, [r][eq][node]
    for (int n = 0; n < Np; ++n) {
      double q[4], Fx[4][3];
      for (int k = 0; k < 4; ++k) q[k] = u[k * KN + el * Np + n];
      for (int x = 0; x < 3; ++x) {
        Fx[0][x] = v0[x] * q[0] + r * cc * cc * q[1 + x];
        for (int k = 0; k < 3; ++k) Fx[1 + k][x] = v0[x] * q[1 + k] + (k == x ? q[0] / r : 0.0);
      }
      for (int rr = 0; rr < dim; ++rr)
        for (int k = 0; k < 4; ++k) Ft[rr][k][n] = Ji[rr * 3 + 0] * Fx[k][0] + Ji[rr * 3 + 1] * Fx[k][1] + Ji[rr * 3 + 2] * Fx[k][2];
    }
    for (int q = 0; q < 4; ++q) {
      double Fe[64];
      std::fill(Fe, Fe + Np, 0.0);
      for (int f = 0; f < Nf; ++f) {
        int fid = c->elFIds[el * Nf + f];
        int side = (el == c->fNbr[fid][0]) ? 0 : 1;
        for (int nf = 0; nf < Nfp; ++nf) Fe[fN2E(c, fid, nf, side)] += c->elFOrient[el * Nf + f] * fF[(q * F + fid) * Nfp + nf];
      }
      for (int i = 0; i < Np; ++i) {
        double vol = 0, lift = 0;
        for (int rr = 0; rr < dim; ++rr)
          for (int j = 0; j < Np; ++j) vol += c->W[(rr * Np + i) * Np + j] * Ft[rr][q][j];
        for (int j = 0; j < Np; ++j) lift += c->Mhat_inv[i * Np + j] * Fe[j];
        int64_t id = q * KN + el * Np + i;
        out[id] = beta * out[id] + dt * (vol - lift / c->absdet[el]);
      }
    }
  }
}

void to_nested(const orc_ctx* c, const double* u, std::vector<std::vector<double>>& v) {
  v.assign(4, std::vector<double>(c->K * c->Np));
  for (int q = 0; q < 4; ++q) std::copy(u + q * c->K * c->Np, u + (q + 1) * c->K * c->Np, v[q].begin());
}
void from_nested(const orc_ctx* c, const std::vector<std::vector<double>>& v, double* u) {
  for (int q = 0; q < 4; ++q) std::copy(v[q].begin(), v[q].end(), u + q * c->K * c->Np);
}

}  // namespace

// ==================================================================== C API
extern "C" {

orc_ctx* orc_create(int dim, int order, int64_t K, const double* node_xyz, const int64_t* node_tags, int ngroups,
                    const int* group_reflecting, const int64_t* const* group_nodes, const int64_t* group_sizes,
                    double c0, double rho0, const double* v0, int sign_mode, int nthreads, int literal_setup,
                    char* err, int errlen) {
  orc_ctx* c = new orc_ctx();
  try {
    if (dim < 1 || dim > 3 || order < 1 || order > 5 || K <= 0) throw std::runtime_error("bad mesh description");
    c->dim = dim; c->p = order; c->K = K;
    c->Np = dim == 1 ? order + 1 : dim == 2 ? (order + 1) * (order + 2) / 2 : (order + 1) * (order + 2) * (order + 3) / 6;
    if (c->Np > 64) throw std::runtime_error("order too high for the oracle");
    c->c0 = c0; c->rho0 = rho0;
    for (int k = 0; k < 3; ++k) c->v0[k] = v0 ? ORC_V0_SIGN * v0[k] : 0.0;
    c->sign_mode = sign_mode;
    c->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
    c->elNodeTags.assign(node_tags, node_tags + K * c->Np);
    c->nodeXYZ.assign(node_xyz, node_xyz + K * c->Np * 3);
    c->tagXYZ.reserve(K * c->Np);
    for (int64_t i = 0; i < K * c->Np; ++i) c->tagXYZ[node_tags[i]] = {node_xyz[3 * i], node_xyz[3 * i + 1], node_xyz[3 * i + 2]};
    setup(c, group_reflecting, group_nodes, group_sizes, ngroups, literal_setup);
  } catch (std::exception& e) {
    if (err && errlen > 0) snprintf(err, errlen, "%s", e.what());
    delete c;
    return nullptr;
  }
  return c;
}

void orc_destroy(orc_ctx* c) { delete c; }

void orc_sizes(const orc_ctx* c, int64_t* o) {
  int64_t nb = 0;
  for (char b : c->fIsB) nb += b;
  o[0] = c->Np; o[1] = c->Nf; o[2] = c->Nfp; o[3] = c->F; o[4] = nb; o[5] = c->G; o[6] = c->Gf; o[7] = c->fc;
}

void orc_connectivity(const orc_ctx* c, int32_t* elFIds, int32_t* elFOrientation, int32_t* fNbr, int32_t* fNToElNIds,
                      int32_t* fBC) {
  for (int64_t i = 0; i < c->K * c->Nf; ++i) { elFIds[i] = c->elFIds[i]; elFOrientation[i] = c->elFOrient[i]; }
  for (int64_t f = 0; f < c->F; ++f) {
    for (int s = 0; s < 2; ++s) fNbr[2 * f + s] = s < (int)c->fNbr[f].size() ? c->fNbr[f][s] : -1;
    for (int nf = 0; nf < c->Nfp; ++nf)
      for (int s = 0; s < 2; ++s)
        fNToElNIds[(f * c->Nfp + nf) * 2 + s] = s < (int)c->fNbr[f].size() ? fN2E(c, f, nf, s) : -1;
    fBC[f] = c->fIsB[f] ? c->fBC[f] : -1;
  }
}

void orc_face_normals(const orc_ctx* c, double* out) {
  for (int64_t f = 0; f < c->F; ++f)
    for (int x = 0; x < 3; ++x) out[f * 3 + x] = c->fNormals[(f * c->Gs) * 3 + x];
}

void orc_mass_inverse(const orc_ctx* c, double* out) { std::copy(c->Minv.begin(), c->Minv.end(), out); }

int orc_has_literal(const orc_ctx* c) { return c->lit ? 1 : 0; }

int orc_rhs(orc_ctx* c, const double* u, double* out, double dt, int mode) {
  if (mode & ORC_FACTORIZED) {
    std::fill(out, out + 4 * c->K * c->Np, 0.0);
    num_step_factorized(c, u, out, dt, 0.0);
    return 0;
  }
  if (!c->lit) return 2;
  std::vector<std::vector<double>> v;
  to_nested(c, u, v);
  update_flux(c, v);
  num_step_literal(c, v, dt, 0.0, (mode & 0x10) != 0);
  from_nested(c, v, out);
  return 0;
}

int orc_step(orc_ctx* c, double* u, int nsteps, double dt, const double* t_steps, int nsrc, const double* src, int mode,
             int euler) {
  const int64_t KN = c->K * c->Np;
  // source node lists (src/solver.cpp:197-210)
  std::vector<std::vector<int64_t>> sidx(nsrc);
  for (int s = 0; s < nsrc; ++s)
    for (int64_t n = 0; n < KN; ++n) {
      const double* X = &c->nodeXYZ[3 * n];
      const double* S = &src[9 * s];
      if (pow(X[0] - S[1], 2) + pow(X[1] - S[2], 2) + pow(X[2] - S[3], 2) < pow(S[4], 2)) sidx[s].push_back(n);
    }
  bool fact = (mode & ORC_FACTORIZED) != 0, nested = (mode & 0x10) != 0;
  if (!fact && !c->lit) return 2;
  std::vector<std::vector<double>> U, k1, k2, k3, k4;
  std::vector<double> uf, kf[4], tmp;
  if (fact) {
    uf.assign(u, u + 4 * KN);
    for (auto& k : kf) k.assign(4 * KN, 0.0);
  } else {
    to_nested(c, u, U);
  }
  for (int step = 0; step < nsteps; ++step) {
    double t = t_steps ? t_steps[step] : 0.0;
    for (int s = 0; s < nsrc; ++s) {
      const double* S = &src[9 * s];
      if (t < S[8])
        for (int64_t n : sidx[s]) {
          double val = S[5] * sin(2 * M_PI * S[6] * t + S[7]);
          if (fact) uf[n] = val; else U[0][n] = val;
        }
    }
    if (euler) {
      if (fact) {
        tmp = uf;
        num_step_factorized(c, tmp.data(), uf.data(), dt, 1.0);
      } else {
        update_flux(c, U);
        num_step_literal(c, U, dt, 1.0, nested);
      }
      continue;
    }
    if (fact) {
      // k_s = dt L(u + c_s k_{s-1}) with the reference's stage inputs
      const double cs[4] = {0, 0.5, 0.5, 1.0};
      tmp.resize(4 * KN);
      for (int s = 0; s < 4; ++s) {
        for (int64_t i = 0; i < 4 * KN; ++i) tmp[i] = s == 0 ? uf[i] : uf[i] + cs[s] * kf[s - 1][i];
        std::fill(kf[s].begin(), kf[s].end(), 0.0);
        num_step_factorized(c, tmp.data(), kf[s].data(), dt, 0.0);
      }
      for (int64_t i = 0; i < 4 * KN; ++i) uf[i] += (kf[0][i] + 2 * kf[1][i] + 2 * kf[2][i] + kf[3][i]) / 6.0;
    } else {
      // src/solver.cpp:261-285
      k1 = k2 = k3 = k4 = U;
      update_flux(c, k1);
      num_step_literal(c, k1, dt, 0, nested);
      for (int q = 0; q < 4; ++q) for (int64_t i = 0; i < KN; ++i) k2[q][i] += 0.5 * k1[q][i];
      update_flux(c, k2);
      num_step_literal(c, k2, dt, 0, nested);
      for (int q = 0; q < 4; ++q) for (int64_t i = 0; i < KN; ++i) k3[q][i] += 0.5 * k2[q][i];
      update_flux(c, k3);
      num_step_literal(c, k3, dt, 0, nested);
      for (int q = 0; q < 4; ++q) for (int64_t i = 0; i < KN; ++i) k4[q][i] += 1 * k3[q][i];
      update_flux(c, k4);
      num_step_literal(c, k4, dt, 0, nested);
      for (int q = 0; q < 4; ++q)
        for (int64_t i = 0; i < KN; ++i) U[q][i] += (k1[q][i] + 2 * k2[q][i] + 2 * k3[q][i] + k4[q][i]) / 6.0;
    }
  }
  if (fact) std::copy(uf.begin(), uf.end(), u);
  else from_nested(c, U, u);
  return 0;
}

int orc_ref_matrices(int dim, int order, double* M, double* S, double* Mf) {
  try {
    auto nodes = elem_nodes(dim, order);
    int Np = (int)nodes.size();
    Lagrange L(dim, order, nodes);
    std::vector<P3> q;
    std::vector<ld> qw;
    rule(dim, order + 3, q, qw);
    std::vector<double> phi(Np), gr(Np * 3);
    std::vector<ld> Ml(Np * Np, 0), Sl(dim * Np * Np, 0);
    for (size_t g = 0; g < q.size(); ++g) {
      L.eval(q[g].data(), phi.data(), reinterpret_cast<double(*)[3]>(gr.data()));
      for (int i = 0; i < Np; ++i)
        for (int j = 0; j < Np; ++j) {
          Ml[i * Np + j] += qw[g] * phi[i] * phi[j];
          for (int r = 0; r < dim; ++r) Sl[(r * Np + i) * Np + j] += qw[g] * gr[i * 3 + r] * phi[j];
        }
    }
    for (int i = 0; i < Np * Np; ++i) M[i] = (double)Ml[i];
    for (int i = 0; i < dim * Np * Np; ++i) S[i] = (double)Sl[i];
    if (dim > 1) {
      auto fn = elem_nodes(dim - 1, order);
      int Nfp = (int)fn.size();
      Lagrange FL(dim - 1, order, fn);
      rule(dim - 1, order + 2, q, qw);
      std::vector<double> fp(Nfp);
      std::vector<ld> Mfl(Nfp * Nfp, 0);
      for (size_t g = 0; g < q.size(); ++g) {
        FL.eval(q[g].data(), fp.data(), nullptr);
        for (int i = 0; i < Nfp; ++i) for (int j = 0; j < Nfp; ++j) Mfl[i * Nfp + j] += qw[g] * fp[i] * fp[j];
      }
      for (int i = 0; i < Nfp * Nfp; ++i) Mf[i] = (double)Mfl[i];
    } else {
      Mf[0] = 1;
    }
  } catch (...) {
    return 1;
  }
  return 0;
}

int orc_ref_nodes(int dim, int order, double* xyz) {
  auto nodes = elem_nodes(dim, order);
  for (size_t n = 0; n < nodes.size(); ++n) for (int c = 0; c < 3; ++c) xyz[3 * n + c] = (double)nodes[n][c];
  return 0;
}

}  // extern "C"
