// C ABI of libdgk (include/dgk.h): context setup, state transfer and time stepping.
#include <algorithm>
#include <array>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <cmath>
#include <cstring>
#include <unordered_map>

#include "dgk_internal.hpp"
#include "refelem.hpp"

namespace dgk {

void Ctx::free_receivers() {
  for (void* p : {(void*)rec.base, rec.w, rec.out, (void*)rec.counter})
    if (p) (void)hipFree(p);
  rec = Rec{};
}

void Ctx::drop_graphs() {
  for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  graphs.clear();
}

Ctx::~Ctx() {
  free_receivers();
  drop_graphs();
  if (tdev) (void)hipFree(tdev);
  for (auto& s : sources)
    if (s.idx) (void)hipFree(s.idx);
  for (void* p : allocs) (void)hipFree(p);
  if (stream) (void)hipStreamDestroy(stream);
}

thread_local std::string g_err;

namespace {

dgk_status fail(Ctx* c, dgk_status st, const std::string& msg) {
  g_err = msg;
  if (c) c->err = msg;
  return st;
}

}  // namespace

template <class F>
dgk_status guarded(Ctx* c, F&& f) {
  try {
    f();
    return DGK_OK;
  } catch (InvalidArg& e) {
    return fail(c, DGK_ERR_INVALID, e.what());
  } catch (Unsupported& e) {
    return fail(c, DGK_ERR_UNSUPPORTED, e.what());
  } catch (MeshError& e) {
    return fail(c, DGK_ERR_MESH, e.what());
  } catch (DeviceError& e) {
    return fail(c, DGK_ERR_DEVICE, e.what());
  } catch (std::bad_alloc&) {
    return fail(c, DGK_ERR_DEVICE, "host allocation failed");
  } catch (std::exception& e) {
    return fail(c, DGK_ERR_INVALID, e.what());
  }
}

namespace {

// derivative of the equispaced order-p vertex Lagrange function psi(l) = prod_k (p l - k)/(k+1)
double dpsi(int p, double l) {
  double s = 0;
  for (int m = 0; m < p; ++m) {
    double t = (double)p / (m + 1);
    for (int k = 0; k < p; ++k)
      if (k != m) t *= (p * l - k) / (k + 1);
    s += t;
  }
  return s;
}

// Sign of the reference normal at the first face quadrature point relative to the
// right-hand normal of the face's first vertices: grad(phi_0) x grad(phi_1) (3D) or
// grad(phi_0) x (0,0,-1) (2D) with phi_v = psi(lambda_v) (src/Mesh.cpp:173-188).
double normal_sigma(int dim, int p) {
  if (dim == 1) return 1;
  std::vector<ld> x, w;
  gauss_legendre(p + 2, x, w);
  double xi0 = (double)x[0];
  if (dim == 2) {
    double l0 = (1 - xi0) / 2;
    return dpsi(p, l0) >= 0 ? 1 : -1;
  }
  double t0 = (xi0 + 1) / 2;
  double u = t0, v = (1 - t0) * t0;
  double l0 = 1 - u - v, l1 = u;
  return dpsi(p, l0) * dpsi(p, l1) >= 0 ? 1 : -1;
}

template <typename T>
void upload_ops(Ctx& c, const RefElem& R, const std::vector<int>& vmap) {
  std::vector<T> W(R.W.begin(), R.W.end()), L(R.LIFT.begin(), R.LIFT.end());
  c.ops.W = c.alloc(W.size() * sizeof(T));
  c.ops.LIFT = c.alloc(L.size() * sizeof(T));
  DGK_HIP(hipMemcpy(c.ops.W, W.data(), W.size() * sizeof(T), hipMemcpyHostToDevice));
  DGK_HIP(hipMemcpy(c.ops.LIFT, L.data(), L.size() * sizeof(T), hipMemcpyHostToDevice));
  std::vector<int> fm;
  for (auto& f : R.fmask) fm.insert(fm.end(), f.begin(), f.end());
  c.ops.fmask = (int*)c.alloc(fm.size() * sizeof(int));
  DGK_HIP(hipMemcpy(c.ops.fmask, fm.data(), fm.size() * sizeof(int), hipMemcpyHostToDevice));
  c.ops.vmap = (int*)c.alloc(vmap.size() * sizeof(int));
  DGK_HIP(hipMemcpy(c.ops.vmap, vmap.data(), vmap.size() * sizeof(int), hipMemcpyHostToDevice));
  // MFMA A fragments: k-step s < d*NG covers W_r (r = s % d) at node 4*(s/d) + g for lane group g;
  // later k-steps cover face entry 4*(s - d*NG) + g of -LIFT. Rows beyond Np are zero. Layout
  // [block][k-step][lane]: MBF 16-row blocks (lane l: row 16*mb + l%16, lane group l/16), then the
  // NT 4-row tail pieces for the 4x4 MFMAs as [piece][k-step][lane group g][row r] (16 values; lane
  // l reads row l%4 at lane group l/16).
  int KS = 0, MB = 0, NT = 0;
  mfma_shape(R.dim, R.p, sizeof(T) == 4, c.oplo, &KS, &MB, &NT);
  const int Np = R.Np, NFF = R.Nf * R.Nfp, NG = (Np + 3) / 4, d = R.dim, MBF = NT ? MB - 1 : MB;
  auto coef = [&](int i, int s, int g) {  // operator entry of row i at k-step s, lane group g
    double v = 0;
    if (i < Np) {
      if (s < d * NG) {
        int r = s % d, n = 4 * (s / d) + g;
        if (n < Np) v = R.W[(r * Np + i) * Np + n];
      } else {
        int jk = 4 * (s - d * NG) + g;
        if (jk < NFF) v = -R.LIFT[i * NFF + jk];
      }
    }
    return v;
  };
  std::vector<T> A((size_t)(MBF * 64 + NT * 16) * KS, (T)0);
  for (int mb = 0; mb < MBF; ++mb)
    for (int s = 0; s < KS; ++s)
      for (int l = 0; l < 64; ++l) A[((size_t)mb * KS + s) * 64 + l] = (T)coef(16 * mb + (l % 16), s, l / 16);
  for (int t = 0; t < NT; ++t)
    for (int s = 0; s < KS; ++s)
      for (int v = 0; v < 16; ++v)
        A[(size_t)MBF * KS * 64 + ((size_t)t * KS + s) * 16 + v] = (T)coef(16 * MBF + 4 * t + v % 4, s, v / 4);
  c.ops.Amf = c.alloc(A.size() * sizeof(T));
  DGK_HIP(hipMemcpy(c.ops.Amf, A.data(), A.size() * sizeof(T), hipMemcpyHostToDevice));
  if (sizeof(T) == 4 && c.oplo) {
    // bf16 residuals: lane l of correction MFMA (mb, t) holds slots i = 0..7 for k-steps 8t+i
    const int NLO = (KS + 7) / 8;
    std::vector<uint16_t> lo((size_t)MB * NLO * 64 * 8, 0);
    auto to_bf16 = [](float f) {
      uint32_t u;
      memcpy(&u, &f, 4);
      u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
      return (uint16_t)(u >> 16);
    };
    for (int mb = 0; mb < MB; ++mb)
      for (int s = 0; s < KS; ++s)
        for (int l = 0; l < 64; ++l) {
          const double v = coef(16 * mb + (l % 16), s, l / 16);
          float hi = (float)v;
          lo[(((size_t)mb * NLO + s / 8) * 64 + l) * 8 + s % 8] = to_bf16((float)(v - (double)hi));
        }
    c.ops.Alo = c.alloc(lo.size() * sizeof(uint16_t));
    DGK_HIP(hipMemcpy(c.ops.Alo, lo.data(), lo.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  }
}

// vmap[(jn*nperm + perm)*Nfp + k]: neighbour element node matching own face node k when the
// own face's vertex slot m is the neighbour face's slot pi_perm[m].
std::vector<int> make_vmap(const RefElem& R, int nperm) {
  const int d = R.dim, p = R.p;
  std::vector<std::array<int, 3>> flat;
  if (d == 1) flat = {{0, 0, 0}};
  if (d == 2) for (auto a : line_lattice(p)) flat.push_back({a[0], a[1], 0});
  if (d == 3) flat = tri_lattice(p);
  static const int P6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const int P2[2][3] = {{0, 1, 2}, {1, 0, 2}};
  std::vector<int> vm(R.Nf * nperm * R.Nfp, 0);
  for (int jn = 0; jn < R.Nf; ++jn)
    for (int pm = 0; pm < nperm; ++pm)
      for (int k = 0; k < R.Nfp; ++k) {
        int kk = 0;
        if (d > 1) {
          const int* pi = d == 3 ? P6[pm] : P2[pm];
          std::array<int, 3> b{0, 0, 0};
          for (int m = 0; m < d; ++m) b[pi[m]] = flat[k][m];
          kk = -1;
          for (int q = 0; q < R.Nfp; ++q)
            if (flat[q] == b) kk = q;
          if (kk < 0) throw std::runtime_error("face lattice permutation not found");
        }
        vm[(jn * nperm + pm) * R.Nfp + k] = R.fmask[jn][kk];
      }
  return vm;
}

}  // namespace

void create_impl(Ctx& c, const dgk_mesh* m, const dgk_phys* ph, const dgk_opts* o) {
  if (!m || !ph) throw InvalidArg("mesh and phys must not be NULL");
  if (m->dim < 1 || m->dim > 3) throw Unsupported("element dimension must be 1, 2 or 3");
  if (m->order < 1 || m->order > 5) throw Unsupported("element order must be 1..5");
  if (m->num_elements <= 0 || !m->node_xyz || !m->node_tags) throw InvalidArg("empty mesh");
  // element and element-face indices are 32-bit on the device
  if ((m->num_elements + kTilePad) * (m->dim + 1) >= ((int64_t)1 << 31))
    throw Unsupported("more than 2^31 element faces");
  if (!(ph->c0 > 0) || !(ph->rho0 > 0)) throw InvalidArg("c0 and rho0 must be positive");
  c.dim = m->dim;
  c.p = m->order;
  c.K = m->num_elements;
  c.precision = o ? o->precision : DGK_FP64;
  c.sign_mode = o ? o->flux_sign : DGK_SIGN_REFERENCE;
  c.device = o ? o->device : 0;
  c.variant = o ? o->variant : 0;
  c.addr64 = o && o->addr64;
  Tuned tuned;
  const bool have_tuning = find_tuning(m->dim, m->order, o ? o->precision : DGK_FP64, &tuned);
  if (have_tuning) {
    c.grid_pct = tuned.grid_pct;
    c.tile_order = tuned.tile_order;
    if (c.variant == 0) c.variant = tuned.variant;
  }
  if (c.variant == 0) c.variant = kVariantMfma;
  if (c.variant != kVariantLds && c.variant != kVariantMfma) throw InvalidArg("unknown stage-kernel variant");
  const int flags = o ? o->fp32_flags : 0;
  if (flags & ~(DGK_FP32_PLAIN_STATE | DGK_FP32_OPERATOR_CORRECTION)) throw InvalidArg("unknown fp32_flags bits");
  // the thread-per-node kernel implements the plain FP32 form only
  const bool f32mfma = (o ? o->precision : DGK_FP64) == DGK_FP32 && c.variant == kVariantMfma;
  c.comp = f32mfma && !(flags & DGK_FP32_PLAIN_STATE);
  c.es = c.variant == kVariantMfma ? m->dim + 1 : 1;  // the state starts at zero: dim+1 live fields
  c.oplo = f32mfma && (flags & DGK_FP32_OPERATOR_CORRECTION);
  if (c.precision != DGK_FP64 && c.precision != DGK_FP32) throw InvalidArg("precision must be DGK_FP64 or DGK_FP32");
  if (c.sign_mode != DGK_SIGN_REFERENCE && c.sign_mode != DGK_SIGN_STABLE) throw InvalidArg("bad flux_sign");
  c.tsize = c.precision == DGK_FP64 ? 8 : 4;
  c.c0 = ph->c0;
  c.rho0 = ph->rho0;
  for (int k = 0; k < 3; ++k) c.v0[k] = ph->v0[k];
  c.fc = (c.dim == 3 && c.p != 1) ? -1 : 1;  // src/Mesh.cpp:210-211
  RefElem R = make_ref_elem(c.dim, c.p);
  c.Np = R.Np;
  c.Nf = R.Nf;
  c.Nfp = R.Nfp;
  c.nperm = c.dim == 3 ? 6 : c.dim == 2 ? 2 : 1;
  for (int j = 0; j < c.Nf; ++j) {
    for (int x = 0; x < 3; ++x) c.refn[j][x] = R.ref_normals[j][x];
    c.fsc[j] = R.ref_face_measure[j] / R.face_ref_measure;
  }
  const int D = c.dim, Np = c.Np, NV = D + 1;
  const int64_t K = c.K;
  // affine (straight-sided) check and vertex arrays
  std::vector<double> vxyz(K * NV * 3);
  for (int64_t e = 0; e < K; ++e) {
    const double* X = m->node_xyz + e * Np * 3;
    for (int v = 0; v < NV; ++v)
      for (int x = 0; x < 3; ++x) vxyz[(e * NV + v) * 3 + x] = X[v * 3 + x];
    double diam = 0;
    for (int v = 1; v < NV; ++v)
      for (int x = 0; x < 3; ++x) diam = std::max(diam, std::fabs(X[v * 3 + x] - X[x]));
    for (int n = NV; n < Np; ++n)
      for (int x = 0; x < 3; ++x) {
        double s = 0;
        for (int v = 0; v < NV; ++v) s += (double)R.lattice[n][v] * X[v * 3 + x];
        s /= c.p;
        if (std::fabs(s - X[n * 3 + x]) > 1e-9 * diam)
          throw Unsupported("element " + std::to_string(e) +
                            " is curved or its high-order nodes are not in Gmsh order; only straight-sided "
                            "elements are supported");
      }
  }
  // dense corner-vertex ids from node tags
  std::vector<int32_t> vid(K * NV);
  {
    std::unordered_map<int64_t, int32_t> ids;
    ids.reserve(K * 2);
    for (int64_t e = 0; e < K; ++e)
      for (int v = 0; v < NV; ++v) {
        int64_t t = m->node_tags[e * Np + v];
        auto it = ids.find(t);
        int32_t id;
        if (it == ids.end()) { id = (int32_t)ids.size(); ids.emplace(t, id); }
        else id = it->second;
        vid[e * NV + v] = id;
      }
    if (ids.size() >= ((size_t)1 << 31)) throw Unsupported("more than 2^31 mesh vertices");
  }
  std::vector<uint8_t> bcf(K * c.Nf, 0);
  if (m->node_bc)
    for (int64_t e = 0; e < K; ++e)
      for (int j = 0; j < c.Nf; ++j) bcf[e * c.Nf + j] = m->node_bc[e * Np + R.fmask[j][0]] ? 1 : 0;

  DGK_HIP(hipSetDevice(c.device));
  DGK_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  // device arrays cover Kpad elements; padding elements have zero state and geometry and an
  // absorbing "face" with zero normal, so kernels can treat every tile as full
  c.Kpad = (K + kTilePad - 1) / kTilePad * kTilePad;
  c.PS = c.Kpad * Np;
  const int64_t n = 4 * c.PS, KP = c.Kpad;
  c.U = c.alloc(n * c.tsize);
  c.A = c.alloc(n * c.tsize);
  c.V0 = c.alloc(n * c.tsize);
  c.V1 = c.alloc(n * c.tsize);
  c.Jinv = c.alloc(KP * 9 * c.tsize);
  c.fgeom = c.alloc(KP * c.Nf * 4 * c.tsize);
  c.nbr = (int32_t*)c.alloc(KP * c.Nf * sizeof(int32_t));
  c.code = (int32_t*)c.alloc(KP * c.Nf * sizeof(int32_t));
  c.nface = (int32_t*)c.alloc(KP * c.Nf * sizeof(int32_t));
  c.perm = (int32_t*)c.alloc(KP * c.Nf * sizeof(int32_t));
  c.flag = (int*)c.alloc(sizeof(int));
  for (void* p : {c.U, c.A, c.V0, c.V1}) DGK_HIP(hipMemsetAsync(p, 0, n * c.tsize, c.stream));
  c.nq = c.dim + 1;  // the state starts at zero
  if (c.comp) {
    c.Ulo = (uint16_t*)c.alloc(n * sizeof(uint16_t));
    DGK_HIP(hipMemsetAsync(c.Ulo, 0, n * sizeof(uint16_t), c.stream));
  }
  DGK_HIP(hipMemsetAsync(c.Jinv, 0, KP * 9 * c.tsize, c.stream));
  DGK_HIP(hipMemsetAsync(c.fgeom, 0, KP * c.Nf * 4 * c.tsize, c.stream));
  DGK_HIP(hipMemsetAsync(c.nbr, 0xff, KP * c.Nf * sizeof(int32_t), c.stream));
  DGK_HIP(hipMemsetAsync(c.code, 0, KP * c.Nf * sizeof(int32_t), c.stream));
  double* vx = nullptr;
  DGK_HIP(hipMalloc(&vx, vxyz.size() * sizeof(double)));
  try {
    DGK_HIP(hipMemcpyAsync(vx, vxyz.data(), vxyz.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    build_connectivity(c, vx, vid.data(), bcf.data(), normal_sigma(D, c.p));
    build_geometry(c, vx);
  } catch (...) {
    (void)hipFree(vx);
    throw;
  }
  DGK_HIP(hipFree(vx));
  auto vmap = make_vmap(R, c.nperm);
  if (c.precision == DGK_FP64) upload_ops<double>(c, R, vmap);
  else upload_ops<float>(c, R, vmap);
  // conformity: the device table must pair face nodes carrying equal tags (src/Mesh.cpp:253-263)
  std::vector<int32_t> nb(K * c.Nf), cd(K * c.Nf);
  DGK_HIP(hipMemcpy(nb.data(), c.nbr, nb.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  DGK_HIP(hipMemcpy(cd.data(), c.code, cd.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  int64_t nbnd = 0;
  for (int64_t e = 0; e < K; ++e)
    for (int j = 0; j < c.Nf; ++j) {
      int32_t o = nb[e * c.Nf + j];
      if (o < 0) { ++nbnd; continue; }
      int sel = cd[e * c.Nf + j] & 0xffff;
      for (int k = 0; k < c.Nfp; ++k)
        if (m->node_tags[e * Np + R.fmask[j][k]] != m->node_tags[(int64_t)o * Np + vmap[sel * c.Nfp + k]])
          throw MeshError("non-conforming face between elements " + std::to_string(e) + " and " + std::to_string(o));
    }
  c.Fb = nbnd;
  c.F = (K * c.Nf + nbnd) / 2;
  // the tuning table's tile shape, where this context's form has a kernel of that shape
  if (have_tuning && tuned.tile_shape != 0 && c.variant == kVariantMfma) {
    StageArgs a = base_args(c);
    if (tile_shape_supported(c, a, tuned.tile_shape)) c.tile_shape = tuned.tile_shape;
  }
}

StageArgs base_args(Ctx& c) {
  StageArgs a{};
  a.nq = c.variant == kVariantLds ? 4 : c.nq;
  a.tile_order = c.tile_order;
  a.addr64 = c.addr64 ? 1 : 0;
  a.K = c.K;
  a.Kpad = c.Kpad;
  a.PS = c.PS;
  a.U = c.U;
  a.A = c.A;
  a.Ulo = c.comp ? c.Ulo : nullptr;
  a.Jinv = c.Jinv;
  a.fgeom = c.fgeom;
  a.nbr = c.nbr;
  a.code = c.code;
  a.W = c.ops.W;
  a.LIFT = c.ops.LIFT;
  a.fmask = c.ops.fmask;
  a.vmap = c.ops.vmap;
  a.Amf = c.ops.Amf;
  a.Alo = c.ops.Alo;
  for (int j = 0; j < kMaxNF; ++j) {
    for (int x = 0; x < 3; ++x) a.refn[j][x] = c.refn[j][x];
    a.fsc[j] = c.fsc[j];
  }
  a.c0 = c.c0;
  a.rho0 = c.rho0;
  for (int k = 0; k < 3; ++k) a.v0[k] = c.v0[k];
  return a;
}

// Launch the four RK4 stages of one step on stream s.
void rk4_step(Ctx& c, double dt, hipStream_t s, hipEvent_t* ev) {
  StageArgs a = base_args(c);
  a.dt = dt;
  void* in[4] = {c.U, c.V0, c.V1, c.V0};
  void* out[4] = {c.V0, c.V1, c.V0, nullptr};
  for (int st = 0; st < 4; ++st) {
    a.stage = st;
    a.Vin = in[st];
    a.Vout = out[st];
    if (ev) DGK_HIP(hipEventRecord(ev[2 * st], s));
    launch_stage(c, a, s);
    if (ev) DGK_HIP(hipEventRecord(ev[2 * st + 1], s));
  }
  launch_record(c, s);
}

namespace {

// Barycentric coordinates of x in element e (normal equations of x - v0 = sum_a l_a (v_a - v0));
// returns the smallest coordinate (>= 0 inside).
double bary_score(int d, int Np, const double* node_xyz, int64_t e, const double* x, double* l) {
  const double* X = node_xyz + (size_t)e * Np * 3;
  double J[3][3] = {}, G[3][3] = {}, r[3] = {}, b[3] = {};
  for (int a = 0; a < d; ++a)
    for (int c = 0; c < 3; ++c) J[c][a] = X[3 * (a + 1) + c] - X[c];
  for (int c = 0; c < 3; ++c) r[c] = x[c] - X[c];
  for (int a = 0; a < d; ++a) {
    for (int bb = 0; bb < d; ++bb)
      for (int c = 0; c < 3; ++c) G[a][bb] += J[c][a] * J[c][bb];
    for (int c = 0; c < 3; ++c) b[a] += J[c][a] * r[c];
  }
  std::vector<ld> Gi(d * d);
  for (int a = 0; a < d; ++a)
    for (int bb = 0; bb < d; ++bb) Gi[a * d + bb] = G[a][bb];
  invert(Gi, d);
  double sum = 0;
  for (int a = 0; a < d; ++a) {
    ld t = 0;
    for (int bb = 0; bb < d; ++bb) t += Gi[a * d + bb] * b[bb];
    l[a + 1] = (double)t;
    sum += l[a + 1];
  }
  l[0] = 1 - sum;
  double score = l[0];
  for (int a = 1; a <= d; ++a) score = std::min(score, l[a]);
  return score;
}

}  // namespace

// Locate points xyz[n][3] in a mesh given by its element node coordinates (Gmsh order, vertices
// first): for each point the element whose smallest barycentric coordinate of the point is
// largest. Elements are binned by bounding box into a uniform grid of about K cells, so each
// point tests only the elements of its cell. Returns the elements; w receives, per point, the
// element's Lagrange basis at the point (pressure = sum_n w[n] p[e*Np + n]).
std::vector<int64_t> locate_points(const RefElem& R, const double* node_xyz, int64_t K, int n, const double* xyz,
                                   std::vector<double>& w) {
  const int d = R.dim, Np = R.Np;
  std::vector<int64_t> out(n, -1);
  w.assign((size_t)n * Np, 0.0);
  if (!n) return out;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t e = 0; e < K; ++e)
    for (int v = 0; v <= d; ++v)
      for (int c = 0; c < 3; ++c) {
        const double q = node_xyz[((size_t)e * Np + v) * 3 + c];
        lo[c] = std::min(lo[c], q);
        hi[c] = std::max(hi[c], q);
      }
  int g[3] = {1, 1, 1};
  const int gd = (int)std::max(1.0, std::floor(std::pow((double)K, 1.0 / d)));
  for (int c = 0; c < d; ++c) g[c] = gd;
  auto cell = [&](int c, double q) {
    const double span = hi[c] - lo[c];
    int i = span > 0 ? (int)std::floor((q - lo[c]) / span * g[c]) : 0;
    return std::min(std::max(i, 0), g[c] - 1);
  };
  // bin elements (CSR over cells)
  const int64_t ncell = (int64_t)g[0] * g[1] * g[2];
  std::vector<int64_t> cnt(ncell + 1, 0);
  std::vector<std::array<int, 6>> box(K);
  for (int64_t e = 0; e < K; ++e) {
    double a[3], b[3];
    for (int c = 0; c < 3; ++c) {
      a[c] = b[c] = node_xyz[((size_t)e * Np) * 3 + c];
      for (int v = 1; v <= d; ++v) {
        a[c] = std::min(a[c], node_xyz[((size_t)e * Np + v) * 3 + c]);
        b[c] = std::max(b[c], node_xyz[((size_t)e * Np + v) * 3 + c]);
      }
    }
    for (int c = 0; c < 3; ++c) {
      const double tol = 1e-9 * (1 + hi[c] - lo[c]);
      box[e][c] = cell(c, a[c] - tol);
      box[e][3 + c] = cell(c, b[c] + tol);
    }
    for (int i = box[e][0]; i <= box[e][3]; ++i)
      for (int j = box[e][1]; j <= box[e][4]; ++j)
        for (int k = box[e][2]; k <= box[e][5]; ++k) ++cnt[((int64_t)i * g[1] + j) * g[2] + k + 1];
  }
  for (int64_t c = 0; c < ncell; ++c) cnt[c + 1] += cnt[c];
  std::vector<int64_t> items(cnt[ncell]), fill(cnt.begin(), cnt.end() - 1);
  for (int64_t e = 0; e < K; ++e)
    for (int i = box[e][0]; i <= box[e][3]; ++i)
      for (int j = box[e][1]; j <= box[e][4]; ++j)
        for (int k = box[e][2]; k <= box[e][5]; ++k) items[fill[((int64_t)i * g[1] + j) * g[2] + k]++] = e;
  for (int p = 0; p < n; ++p) {
    const double* x = xyz + 3 * p;
    int64_t best = -1;
    double best_score = -1e300, best_l[4] = {0, 0, 0, 0}, l[4];
    bool in_box = true;
    for (int c = 0; c < d; ++c)
      if (x[c] < lo[c] - 1e-9 * (1 + hi[c] - lo[c]) || x[c] > hi[c] + 1e-9 * (1 + hi[c] - lo[c])) in_box = false;
    if (in_box) {
      const int64_t ci = ((int64_t)cell(0, x[0]) * g[1] + cell(1, x[1])) * g[2] + cell(2, x[2]);
      for (int64_t it = cnt[ci]; it < cnt[ci + 1]; ++it) {
        const double sc = bary_score(d, Np, node_xyz, items[it], x, l);
        if (sc > best_score || (sc == best_score && items[it] < best)) {
          best_score = sc;
          best = items[it];
          std::copy(l, l + 4, best_l);
        }
      }
    }
    if (best < 0 || best_score < -1e-6)
      throw InvalidArg("receiver point " + std::to_string(p) + " lies outside the mesh");
    std::array<ld, 3> ref{0, 0, 0};
    if (d == 1) ref[0] = best_l[1] - best_l[0];
    else
      for (int a = 0; a < d; ++a) ref[a] = best_l[a + 1];
    std::vector<ld> phi(Np);
    R.eval_basis(ref, phi.data(), nullptr);
    for (int k = 0; k < Np; ++k) w[(size_t)p * Np + k] = (double)phi[k];
    out[p] = best;
  }
  return out;
}

// Install receivers (element, weights) on the context, with room for cap recorded steps.
void install_receivers(Ctx& c, const std::vector<int64_t>& elem, const std::vector<double>& w, int cap) {
  DGK_HIP(hipStreamSynchronize(c.stream));
  c.free_receivers();
  c.drop_graphs();
  const int n = (int)elem.size();
  if (!n) return;
  if (cap < 1) throw InvalidArg("receiver capacity must be positive");
  std::vector<int64_t> base(n);
  for (int i = 0; i < n; ++i) base[i] = elem[i] * c.Np;
  DGK_HIP(hipMalloc(&c.rec.base, n * sizeof(int64_t)));
  DGK_HIP(hipMalloc(&c.rec.w, (size_t)n * c.Np * c.tsize));
  DGK_HIP(hipMalloc(&c.rec.out, (size_t)n * cap * c.tsize));
  DGK_HIP(hipMalloc(&c.rec.counter, sizeof(int)));
  c.rec.n = n;
  c.rec.cap = cap;
  DGK_HIP(hipMemcpy(c.rec.base, base.data(), n * sizeof(int64_t), hipMemcpyHostToDevice));
  if (c.precision == DGK_FP64) {
    DGK_HIP(hipMemcpy(c.rec.w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> wf(w.begin(), w.end());
    DGK_HIP(hipMemcpy(c.rec.w, wf.data(), wf.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  DGK_HIP(hipMemset(c.rec.counter, 0, sizeof(int)));
}

// Put rows x rec.n host values back into a freshly installed record and continue after them.
void restore_receivers(Ctx& c, const std::vector<double>& h, int rows) {
  const size_t m = (size_t)rows * c.rec.n;
  if (rows > c.rec.cap || h.size() < m) throw InvalidArg("restored receiver rows exceed the record");
  if (c.precision == DGK_FP64) {
    DGK_HIP(hipMemcpy(c.rec.out, h.data(), m * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> f(h.begin(), h.begin() + m);
    DGK_HIP(hipMemcpy(c.rec.out, f.data(), m * sizeof(float), hipMemcpyHostToDevice));
  }
  DGK_HIP(hipMemcpy(c.rec.counter, &rows, sizeof(int), hipMemcpyHostToDevice));
}

// Copy the recorded rows (all receivers) to host doubles; returns the row count.
int fetch_receivers(Ctx& c, std::vector<double>& out) {
  int rows = 0;
  if (!c.rec.n) { out.clear(); return 0; }
  DGK_HIP(hipStreamSynchronize(c.stream));
  DGK_HIP(hipMemcpy(&rows, c.rec.counter, sizeof(int), hipMemcpyDeviceToHost));
  const size_t m = (size_t)rows * c.rec.n;
  out.resize(m);
  if (c.precision == DGK_FP64) {
    DGK_HIP(hipMemcpy(out.data(), c.rec.out, m * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> f(m);
    DGK_HIP(hipMemcpy(f.data(), c.rec.out, m * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < m; ++i) out[i] = f[i];
  }
  return rows;
}

namespace {

void apply_sources(Ctx& c, double t) {
  for (auto& src : c.sources)
    if (t < src.duration) launch_source(c, src, src.amp * std::sin(2 * M_PI * src.freq * t + src.phase), c.stream);
}

bool sources_active(Ctx& c, double t) {
  for (auto& src : c.sources)
    if (t < src.duration && src.n) return true;
  return false;
}

// The graph of `steps` RK4 steps (each preceded by the timed source overwrite when src), captured
// once per (steps, src) and time step.
hipGraphExec_t rk4_graph(Ctx& c, int steps, double dt, bool src) {
  if (c.graph_dt != dt) c.drop_graphs();
  c.graph_dt = dt;
  auto it = c.graphs.find({steps, src});
  if (it != c.graphs.end()) return it->second;
  hipGraph_t g;
  DGK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  try {
    for (int k = 0; k < steps; ++k) {
      if (src) {
        for (auto& s : c.sources) launch_source_timed(c, s, c.stream);
        launch_tick(c, c.stream);
      }
      rk4_step(c, dt, c.stream);
    }
  } catch (...) {
    hipGraph_t junk = nullptr;
    if (hipStreamEndCapture(c.stream, &junk) == hipSuccess && junk) (void)hipGraphDestroy(junk);
    throw;
  }
  DGK_HIP(hipStreamEndCapture(c.stream, &g));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  DGK_HIP(e);
  c.graphs[{steps, src}] = ex;
  return ex;
}

// RK4 steps replayed from captured graphs of up to kChunk steps (the remainder gets a graph of its
// own length). With sources, the steps' start times go to the device first and the captured
// source kernels read the current step's time there, so source steps replay from graphs too.
void rk4_impl(Ctx& c, int nsteps, double dt, const double* ts) {
  if (nsteps < 0) throw InvalidArg("nsteps < 0");
  constexpr int kChunk = 25;
  bool src = false;
  if (ts && !c.sources.empty())
    for (int i = 0; i < nsteps && !src; ++i) src = sources_active(c, ts[i]);
  if (src) {
    if (nsteps > c.tcap) {  // the captures hold the old address
      DGK_HIP(hipStreamSynchronize(c.stream));
      c.drop_graphs();
      if (c.tdev) DGK_HIP(hipFree(c.tdev));
      c.tdev = nullptr;
      c.tcap = 0;
      DGK_HIP(hipMalloc(&c.tdev, (size_t)std::max(nsteps, 1024) * sizeof(double)));
      c.tcap = std::max(nsteps, 1024);
    }
    if (!c.tstep) c.tstep = (int*)c.alloc(sizeof(int));
    DGK_HIP(hipMemcpyAsync(c.tdev, ts, (size_t)nsteps * sizeof(double), hipMemcpyHostToDevice, c.stream));
    DGK_HIP(hipMemsetAsync(c.tstep, 0, sizeof(int), c.stream));
    DGK_HIP(hipStreamSynchronize(c.stream));  // ts is the caller's pageable buffer
  }
  for (int i = 0; i < nsteps;) {
    const int n = std::min(kChunk, nsteps - i);
    DGK_HIP(hipGraphLaunch(rk4_graph(c, n, dt, src), c.stream));
    i += n;
  }
}

void euler_impl(Ctx& c, int nsteps, double dt, const double* ts) {
  for (int i = 0; i < nsteps; ++i) {
    if (ts) apply_sources(c, ts[i]);
    StageArgs a = base_args(c);
    a.dt = dt;
    a.stage = 5;
    a.Vin = c.U;
    a.Vout = c.V0;
    launch_stage(c, a, c.stream);
    std::swap(c.U, c.V0);  // graphs capture buffer addresses; drop any cached one
    c.drop_graphs();
    launch_record(c, c.stream);
  }
}

}  // namespace

namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 3>, Tuned> g_tune;  // (dim, order, precision) -> launch settings
}  // namespace

bool find_tuning(int dim, int p, int precision, Tuned* t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find({dim, p, precision});
  if (it == g_tune.end()) return false;
  *t = it->second;
  return true;
}

int live_planes(int dim, const double* u, int64_t n) {
  for (int q = dim + 1; q < 4; ++q)
    for (int64_t i = 0; i < n; ++i)
      if (u[q * n + i] != 0.0) return 4;
  return dim + 1;
}

double* staging(Ctx& c) {
  if (!c.stage_d) {
    c.stage_d = (double*)c.alloc(4 * c.PS * sizeof(double));
    DGK_HIP(hipMemsetAsync(c.stage_d, 0, 4 * c.PS * sizeof(double), c.stream));
  }
  return c.stage_d;
}

namespace {

// Host u[4][K*Np] (FP64) <-> a device state buffer in the state precision and layout. FP64 field
// planes are copied directly; otherwise through the FP64 staging planes and a conversion kernel.
void upload_planes(Ctx& c, const double* u, void* dst) {
  const size_t row = (size_t)c.K * c.Np * sizeof(double);
  if (c.precision == DGK_FP64 && c.es == 1) {
    DGK_HIP(hipMemcpy2DAsync(dst, c.PS * sizeof(double), u, row, row, 4, hipMemcpyHostToDevice, c.stream));
    return;
  }
  double* s = staging(c);
  DGK_HIP(hipMemcpy2DAsync(s, c.PS * sizeof(double), u, row, row, 4, hipMemcpyHostToDevice, c.stream));
  state_from_f64(c, s, dst, dst == c.U && c.comp ? c.Ulo : nullptr, 0, c.K * c.Np);
}

void download_planes(Ctx& c, const void* src, double* u) {
  const size_t row = (size_t)c.K * c.Np * sizeof(double);
  if (c.precision == DGK_FP64 && c.es == 1) {
    DGK_HIP(hipMemcpy2DAsync(u, row, src, c.PS * sizeof(double), row, 4, hipMemcpyDeviceToHost, c.stream));
    return;
  }
  double* s = staging(c);
  state_to_f64(c, src, src == c.U && c.comp ? c.Ulo : nullptr, s, 0, c.K * c.Np);
  DGK_HIP(hipMemcpy2DAsync(u, row, s, c.PS * sizeof(double), row, 4, hipMemcpyDeviceToHost, c.stream));
}

}  // namespace
}  // namespace dgk

using namespace dgk;

extern "C" {

dgk_status dgk_create(const dgk_mesh* mesh, const dgk_phys* phys, const dgk_opts* opts, dgk_ctx** out) {
  if (!out) return fail(nullptr, DGK_ERR_INVALID, "out must not be NULL");
  *out = nullptr;
  dgk_ctx* c = new dgk_ctx();
  dgk_status st = guarded(c, [&] { create_impl(*c, mesh, phys, opts); });
  if (st != DGK_OK) {
    g_err = c->err;
    delete c;
    return st;
  }
  *out = c;
  return DGK_OK;
}

dgk_status dgk_destroy(dgk_ctx* c) {
  if (!c) return DGK_OK;
  delete c;
  return DGK_OK;
}

const char* dgk_last_error(const dgk_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

dgk_status dgk_info_get(const dgk_ctx* c, dgk_info* info) {
  if (!c || !info) return fail(nullptr, DGK_ERR_INVALID, "NULL argument");
  info->dim = c->dim; info->order = c->p; info->Np = c->Np; info->Nf = c->Nf; info->Nfp = c->Nfp;
  info->num_elements = c->K; info->num_faces = c->F; info->num_boundary_faces = c->Fb;
  info->precision = c->precision; info->variant = c->variant; info->fc = c->fc;
  info->device_bytes = c->device_bytes;
  return DGK_OK;
}

dgk_status dgk_set_state(dgk_ctx* c, const double* u) {
  if (!c || !u) return fail(c, DGK_ERR_INVALID, "NULL argument");
  return guarded(c, [&] {
    const int nq = live_planes(c->dim, u, c->K * c->Np);
    if (nq != c->nq) c->drop_graphs();  // captured launches carry the plane count
    c->nq = nq;
    if (c->es > 1) {
      c->es = nq;  // interleaved buffers hold exactly the live fields; the upload below relays U
    } else {
      // planes beyond nq are never written by the kernels: clear them in every buffer (Euler steps
      // swap U with a stage buffer that may hold values from an earlier 4-plane state)
      for (void* p : {c->U, c->A, c->V0, c->V1})
        DGK_HIP(hipMemsetAsync((char*)p + nq * c->PS * c->tsize, 0, (4 - nq) * c->PS * c->tsize, c->stream));
      if (c->Ulo) DGK_HIP(hipMemsetAsync(c->Ulo + nq * c->PS, 0, (4 - nq) * c->PS * sizeof(uint16_t), c->stream));
    }
    upload_planes(*c, u, c->U);
    DGK_HIP(hipStreamSynchronize(c->stream));
  });
}

dgk_status dgk_get_state(dgk_ctx* c, double* u) {
  if (!c || !u) return fail(c, DGK_ERR_INVALID, "NULL argument");
  return guarded(c, [&] {
    download_planes(*c, c->U, u);
    DGK_HIP(hipStreamSynchronize(c->stream));
  });
}

dgk_status dgk_set_sources(dgk_ctx* c, int n, const double* src, const double* node_xyz) {
  if (!c || n < 0 || (n > 0 && (!src || !node_xyz))) return fail(c, DGK_ERR_INVALID, "bad source arguments");
  return guarded(c, [&] {
    DGK_HIP(hipStreamSynchronize(c->stream));
    c->drop_graphs();  // source steps are captured with the source list
    for (auto& s : c->sources)
      if (s.idx) DGK_HIP(hipFree(s.idx));
    c->sources.clear();
    const int64_t KN = c->K * c->Np;
    for (int s = 0; s < n; ++s) {
      const double* S = src + 9 * s;
      std::vector<int64_t> idx;
      for (int64_t i = 0; i < KN; ++i) {
        const double* X = node_xyz + 3 * i;
        if (std::pow(X[0] - S[1], 2) + std::pow(X[1] - S[2], 2) + std::pow(X[2] - S[3], 2) < std::pow(S[4], 2))
          idx.push_back(i);  // pressure is equation 0: U[0*KN + i]
      }
      Ctx::Src d;
      d.amp = S[5]; d.freq = S[6]; d.phase = S[7]; d.duration = S[8];
      d.n = (int64_t)idx.size();
      if (d.n) {
        DGK_HIP(hipMalloc(&d.idx, idx.size() * sizeof(int64_t)));
        DGK_HIP(hipMemcpy(d.idx, idx.data(), idx.size() * sizeof(int64_t), hipMemcpyHostToDevice));
      }
      c->sources.push_back(d);
    }
  });
}

dgk_status dgk_set_receivers(dgk_ctx* c, int n, const double* xyz, const double* node_xyz, int capacity) {
  if (!c || n < 0 || (n > 0 && (!xyz || !node_xyz))) return fail(c, DGK_ERR_INVALID, "bad receiver arguments");
  return guarded(c, [&] {
    RefElem R = make_ref_elem(c->dim, c->p);
    std::vector<double> w;
    std::vector<int64_t> elem = locate_points(R, node_xyz, c->K, n, xyz, w);
    install_receivers(*c, elem, w, capacity);
  });
}

dgk_status dgk_get_receivers(dgk_ctx* c, double* out, int capacity_rows, int* rows) {
  if (!c || !rows) return fail(c, DGK_ERR_INVALID, "bad receiver arguments");
  return guarded(c, [&] {
    std::vector<double> h;
    int r = fetch_receivers(*c, h);
    *rows = r;
    if (out) {
      if (r > capacity_rows) throw InvalidArg("output buffer holds fewer rows than recorded");
      std::copy(h.begin(), h.end(), out);
    }
  });
}

dgk_status dgk_set_launch(dgk_ctx* c, int grid_pct, int tile_order) {
  if (!c || grid_pct < 1 || grid_pct > 10000 || tile_order < 0 || tile_order > 1)
    return fail(c, DGK_ERR_INVALID, "bad launch settings");
  return guarded(c, [&] {
    DGK_HIP(hipStreamSynchronize(c->stream));
    c->grid_pct = grid_pct;
    c->tile_order = tile_order;
    c->drop_graphs();
  });
}

dgk_status dgk_stage_addressing(int dim, int order, int precision, int64_t num_elements, int* addr64) {
  if (!addr64 || dim < 1 || dim > 3 || order < 1 || order > 5 || num_elements <= 0 ||
      (precision != DGK_FP64 && precision != DGK_FP32))
    return fail(nullptr, DGK_ERR_INVALID, "bad addressing query");
  if (num_elements >= ((int64_t)1 << 31) || (num_elements + kTilePad) * (dim + 1) >= ((int64_t)1 << 31))
    return fail(nullptr, DGK_ERR_UNSUPPORTED, "more than 2^31 element faces");
  int64_t np = 1;
  for (int k = 1; k <= dim; ++k) np = np * (order + k) / k;  // C(order + dim, dim)
  const int64_t kpad = (num_elements + kTilePad - 1) / kTilePad * kTilePad;
  *addr64 = stage_needs_addr64(kpad, (int)np, dim + 1, dim == 3 ? 4 : dim + 1, precision == DGK_FP64 ? 8 : 4);
  return DGK_OK;
}

dgk_status dgk_set_tile_shape(dgk_ctx* c, int shape) {
  if (!c || shape < 0 || shape >= kTileShapes) return fail(c, DGK_ERR_INVALID, "bad tile shape");
  return guarded(c, [&] {
    StageArgs a = base_args(*c);
    if (!tile_shape_supported(*c, a, shape))
      throw Unsupported("tile shape " + std::to_string(shape) + " has no kernel for this context's form");
    DGK_HIP(hipStreamSynchronize(c->stream));
    c->tile_shape = shape;
    c->drop_graphs();
  });
}

dgk_status dgk_load_tuning(const char* path) {
  if (!path) return fail(nullptr, DGK_ERR_INVALID, "NULL path");
  return guarded(nullptr, [&] {
    std::ifstream f(path);
    if (!f) throw InvalidArg(std::string("cannot open tuning table ") + path);
    std::map<std::array<int, 3>, Tuned> t;
    std::string line;
    while (std::getline(f, line)) {
      if (line.empty() || line[0] == '#') continue;
      std::istringstream ss(line);
      int d, p, prec;
      Tuned v;
      if (!(ss >> d >> p >> prec >> v.variant >> v.grid_pct >> v.tile_order))
        throw InvalidArg("malformed tuning line: " + line);
      // then either "tile_shape stage_us" or, in tables without the shape column, "stage_us"
      std::vector<std::string> rest;
      for (std::string w; ss >> w;) rest.push_back(w);
      if (rest.size() > 2) throw InvalidArg("malformed tuning line: " + line);
      if (rest.size() == 2) {
        size_t used = 0;
        try {
          v.tile_shape = std::stoi(rest[0], &used);
        } catch (const std::exception&) {
          used = 0;
        }
        if (used != rest[0].size()) throw InvalidArg("malformed tuning line: " + line);
      }
      if (d < 1 || d > 3 || p < 1 || p > 5 || (prec != DGK_FP64 && prec != DGK_FP32) ||
          (v.variant != kVariantLds && v.variant != kVariantMfma) || v.grid_pct < 1 || v.grid_pct > 10000 ||
          v.tile_order < 0 || v.tile_order > 1 || v.tile_shape < 0 || v.tile_shape >= kTileShapes)
        throw InvalidArg("out-of-range tuning line: " + line);
      t[{d, p, prec}] = v;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune = t;
  });
}

dgk_status dgk_rk4(dgk_ctx* c, int nsteps, double dt, const double* t_start) {
  if (!c) return fail(nullptr, DGK_ERR_INVALID, "NULL context");
  return guarded(c, [&] { rk4_impl(*c, nsteps, dt, t_start); });
}

dgk_status dgk_euler(dgk_ctx* c, int nsteps, double dt, const double* t_start) {
  if (!c) return fail(nullptr, DGK_ERR_INVALID, "NULL context");
  return guarded(c, [&] { euler_impl(*c, nsteps, dt, t_start); });
}

dgk_status dgk_rhs(dgk_ctx* c, const double* u, double* out, double dt) {
  if (!c || !u || !out) return fail(c, DGK_ERR_INVALID, "NULL argument");
  return guarded(c, [&] {
    const int nq = live_planes(c->dim, u, c->K * c->Np);
    const int es = c->es;
    if (es > 1) c->es = nq;  // the operand's own live fields, for this call only
    try {
      upload_planes(*c, u, c->V1);
      StageArgs a = base_args(*c);
      a.dt = dt;
      a.stage = 4;
      a.Vin = c->V1;
      a.Vout = c->V0;
      a.nq = nq;
      launch_stage(*c, a, c->stream);
      download_planes(*c, c->V0, out);
      DGK_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
      c->es = es;
      throw;
    }
    c->es = es;
    const int64_t n = c->K * c->Np;  // planes that are not live have a zero right-hand side
    std::fill(out + nq * n, out + 4 * n, 0.0);
  });
}

dgk_status dgk_sync(dgk_ctx* c) {
  if (!c) return fail(nullptr, DGK_ERR_INVALID, "NULL context");
  return guarded(c, [&] { DGK_HIP(hipStreamSynchronize(c->stream)); });
}

dgk_status dgk_check_finite(dgk_ctx* c) {
  if (!c) return fail(nullptr, DGK_ERR_INVALID, "NULL context");
  bool ok = true;
  dgk_status st = guarded(c, [&] { ok = check_finite(*c); });
  if (st != DGK_OK) return st;
  return ok ? DGK_OK : fail(c, DGK_ERR_NONFINITE, "state contains NaN or Inf");
}

dgk_status dgk_connectivity(dgk_ctx* c, int32_t* nbr, int32_t* nface, int32_t* perm, int32_t* sign) {
  if (!c) return fail(nullptr, DGK_ERR_INVALID, "NULL context");
  return guarded(c, [&] {
    size_t n = c->K * c->Nf;
    std::vector<int32_t> cd(n);
    if (nbr) DGK_HIP(hipMemcpy(nbr, c->nbr, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (nface) DGK_HIP(hipMemcpy(nface, c->nface, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (perm) DGK_HIP(hipMemcpy(perm, c->perm, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (sign) {
      std::vector<int32_t> nb(n);
      DGK_HIP(hipMemcpy(cd.data(), c->code, n * sizeof(int32_t), hipMemcpyDeviceToHost));
      DGK_HIP(hipMemcpy(nb.data(), c->nbr, n * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < n; ++i) sign[i] = nb[i] < 0 ? 0 : ((cd[i] >> 16) & 1 ? -1 : 1);
    }
  });
}

dgk_status dgk_rk4_timed(dgk_ctx* c, int nsteps, double dt, double* stage_ms, double* total_ms) {
  if (!c || nsteps < 1) return fail(c, DGK_ERR_INVALID, "bad arguments");
  return guarded(c, [&] {
    std::vector<hipEvent_t> ev(8 * nsteps + 2);
    for (auto& e : ev) DGK_HIP(hipEventCreate(&e));
    try {
      DGK_HIP(hipEventRecord(ev[8 * nsteps], c->stream));
      for (int i = 0; i < nsteps; ++i) rk4_step(*c, dt, c->stream, &ev[8 * i]);
      DGK_HIP(hipEventRecord(ev[8 * nsteps + 1], c->stream));
      DGK_HIP(hipEventSynchronize(ev[8 * nsteps + 1]));
      double sum = 0;
      for (int i = 0; i < 4 * nsteps; ++i) {
        float ms = 0;
        DGK_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        sum += ms;
      }
      float tot = 0;
      DGK_HIP(hipEventElapsedTime(&tot, ev[8 * nsteps], ev[8 * nsteps + 1]));
      if (stage_ms) *stage_ms = sum;
      if (total_ms) *total_ms = tot;
    } catch (...) {
      for (auto& e : ev) (void)hipEventDestroy(e);
      throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
  });
}

const char* dgk_variant_name(int v) { return variant_name(v); }

}  // extern "C"

// ------------------------------------------------------------------------------- batches
struct dgk_batch {
  struct Group {
    dgk::Ctx* ctx = nullptr;
    std::vector<int> sims;
    std::vector<int64_t> off;  // element offset of each member in the union
    // receivers per member: element (member-local) and basis weights; columns in the group record
    std::vector<std::vector<int64_t>> rel;
    std::vector<std::vector<double>> rw;
    std::vector<int> rstart;  // record row at which each member's series begins
    int rcap = 0;
  };
  std::vector<Group> groups;
  std::vector<std::pair<int, int>> where;  // sim -> (group, member)
  std::vector<int> nq;                     // live planes of each member's state
  std::vector<int64_t> K;                  // elements per sim
  hipStream_t main = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  // captured step sequences by (steps, schedule)
  std::map<std::pair<int, int>, hipGraphExec_t> graphs;
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
  std::string err;
  ~dgk_batch() {
    drop_graphs();
    for (auto& g : groups) delete g.ctx;
    for (auto e : joins) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (main) (void)hipStreamDestroy(main);
  }
};

namespace {

dgk_status bfail(dgk_batch* b, dgk_status st, const std::string& msg) {
  dgk::g_err = msg;
  if (b) b->err = msg;
  return st;
}

template <class F>
dgk_status bguard(dgk_batch* b, F&& f) {
  dgk::Ctx tmp;
  dgk_status st = dgk::guarded(&tmp, f);
  if (st != DGK_OK) return bfail(b, st, tmp.err);
  return DGK_OK;
}

void batch_create_impl(dgk_batch& B, int n, const dgk_mesh* meshes, const dgk_phys* phys, const double* dt,
                       const dgk_opts* opts) {
  using namespace dgk;
  if (n <= 0 || !meshes || !phys || !dt) throw InvalidArg("empty batch");
  B.where.resize(n);
  B.K.resize(n);
  B.nq.assign(n, 0);
  // Group key: dim, order, physics. A group closes when adding a member would break the per-context
  // limits (2^31 element faces, 2^31 vertices); the next member of that key opens a new one.
  std::vector<std::vector<int>> members;
  std::vector<std::array<double, 7>> keys;
  std::vector<int64_t> gK, gV;
  for (int i = 0; i < n; ++i) {
    const dgk_mesh& m = meshes[i];
    if (!(dt[i] > 0)) throw InvalidArg("batch time steps must be positive");
    if (m.dim < 1 || m.dim > 3) throw Unsupported("element dimension must be 1, 2 or 3");
    if (m.order < 1 || m.order > 5) throw Unsupported("element order must be 1..5");
    if (m.num_elements <= 0 || !m.node_xyz || !m.node_tags) throw InvalidArg("empty mesh in batch");
    const int Np = num_nodes(m.dim, m.order);
    std::vector<int64_t> corners;
    corners.reserve(m.num_elements * (m.dim + 1));
    for (int64_t e = 0; e < m.num_elements; ++e)
      for (int v = 0; v <= m.dim; ++v) corners.push_back(m.node_tags[e * Np + v]);
    std::sort(corners.begin(), corners.end());
    const int64_t nv = std::unique(corners.begin(), corners.end()) - corners.begin();
    std::array<double, 7> k{(double)m.dim, (double)m.order, phys[i].c0, phys[i].rho0,
                            phys[i].v0[0], phys[i].v0[1], phys[i].v0[2]};
    size_t gi = 0;
    while (gi < keys.size() &&
           (keys[gi] != k || (gK[gi] + m.num_elements + kTilePad) * (m.dim + 1) >= ((int64_t)1 << 31) ||
            gV[gi] + nv >= ((int64_t)1 << 31)))
      ++gi;
    if (gi == keys.size()) {
      keys.push_back(k);
      members.emplace_back();
      gK.push_back(0);
      gV.push_back(0);
    }
    members[gi].push_back(i);
    gK[gi] += m.num_elements;
    gV[gi] += nv;
  }
  for (size_t gi = 0; gi < keys.size(); ++gi) {
    dgk_batch::Group G;
    G.sims = members[gi];
    const dgk_mesh& m0 = meshes[G.sims[0]];
    const int Np = num_nodes(m0.dim, m0.order);
    int64_t Ktot = 0;
    for (int i : G.sims) Ktot += meshes[i].num_elements;
    std::vector<double> xyz;
    std::vector<int64_t> tags;
    std::vector<uint8_t> bc;
    xyz.reserve(Ktot * Np * 3);
    tags.reserve(Ktot * Np);
    bc.reserve(Ktot * Np);
    int64_t tag_base = 0;
    for (size_t k = 0; k < G.sims.size(); ++k) {
      const dgk_mesh& m = meshes[G.sims[k]];
      G.off.push_back((int64_t)tags.size() / Np);
      int64_t mx = 0;
      for (int64_t j = 0; j < m.num_elements * Np; ++j) {
        tags.push_back(m.node_tags[j] + tag_base);  // tags stay unique across members
        mx = std::max(mx, m.node_tags[j]);
        bc.push_back(m.node_bc ? m.node_bc[j] : 0);
      }
      xyz.insert(xyz.end(), m.node_xyz, m.node_xyz + m.num_elements * Np * 3);
      tag_base += mx + 1;
      B.where[G.sims[k]] = {(int)B.groups.size(), (int)k};
      B.K[G.sims[k]] = m.num_elements;
      B.nq[G.sims[k]] = m.dim + 1;
    }
    dgk_mesh um{m0.dim, m0.order, Ktot, xyz.data(), tags.data(), bc.data()};
    G.ctx = new Ctx();
    B.groups.push_back(G);  // owned by B from here on
    create_impl(*B.groups.back().ctx, &um, &phys[G.sims[0]], opts);
    Ctx& c = *B.groups.back().ctx;
    for (size_t k = 0; k < G.sims.size(); ++k) scale_geometry(c, G.off[k], B.K[G.sims[k]], dt[G.sims[k]]);
    DGK_HIP(hipStreamSynchronize(c.stream));
  }
  DGK_HIP(hipStreamCreateWithFlags(&B.main, hipStreamNonBlocking));
  DGK_HIP(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
  B.joins.resize(B.groups.size());
  for (auto& e : B.joins) DGK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// stage launches of nsteps RK4 steps for every group, on the groups' streams (parallel) or main
void batch_enqueue(dgk_batch& B, int nsteps, int schedule) {
  using namespace dgk;
  if (schedule == 0) {
    DGK_HIP(hipEventRecord(B.fork, B.main));
    for (size_t g = 0; g < B.groups.size(); ++g) {
      Ctx& c = *B.groups[g].ctx;
      DGK_HIP(hipStreamWaitEvent(c.stream, B.fork, 0));
      for (int i = 0; i < nsteps; ++i) rk4_step(c, 1.0, c.stream);
      DGK_HIP(hipEventRecord(B.joins[g], c.stream));
      DGK_HIP(hipStreamWaitEvent(B.main, B.joins[g], 0));
    }
  } else {
    for (int i = 0; i < nsteps; ++i)
      for (auto& G : B.groups) rk4_step(*G.ctx, 1.0, B.main);
  }
}

// The graph of `chunk` steps of every group under `schedule`, captured once.
hipGraphExec_t batch_graph(dgk_batch& B, int chunk, int schedule) {
  using namespace dgk;
  auto it = B.graphs.find({chunk, schedule});
  if (it != B.graphs.end()) return it->second;
  hipGraph_t gr;
  DGK_HIP(hipStreamBeginCapture(B.main, hipStreamCaptureModeGlobal));
  try {
    batch_enqueue(B, chunk, schedule);
  } catch (...) {
    hipGraph_t junk = nullptr;
    if (hipStreamEndCapture(B.main, &junk) == hipSuccess && junk) (void)hipGraphDestroy(junk);
    throw;
  }
  DGK_HIP(hipStreamEndCapture(B.main, &gr));
  hipGraphExec_t ex = nullptr;
  hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
  (void)hipGraphDestroy(gr);
  DGK_HIP(e);
  B.graphs[{chunk, schedule}] = ex;
  return ex;
}

// Steps per captured graph: a 256-simulation ensemble has about 32 stage launches per step, so a
// graph holds a few thousand kernel nodes and is replayed nsteps/100 times.
constexpr int kMaxGraphSteps = 100;

void batch_check(int nsteps, int schedule) {
  if (nsteps < 0 || schedule < 0 || schedule > 1) throw dgk::InvalidArg("bad batch step arguments");
}

void batch_rk4_impl(dgk_batch& B, int nsteps, int schedule) {
  using namespace dgk;
  batch_check(nsteps, schedule);
  for (auto& G : B.groups)  // the batch stream runs after anything queued on the contexts
    DGK_HIP(hipStreamSynchronize(G.ctx->stream));
  // Up to kMaxGraphSteps steps of every group are captured as one graph and replayed, the
  // remainder as a graph of its own length; captures are kept per (steps, schedule).
  for (int done = 0; done < nsteps;) {
    const int chunk = std::min(nsteps - done, kMaxGraphSteps);
    DGK_HIP(hipGraphLaunch(batch_graph(B, chunk, schedule), B.main));
    done += chunk;
  }
}

}  // namespace

extern "C" {

dgk_status dgk_batch_create(int n, const dgk_mesh* meshes, const dgk_phys* phys, const double* dt,
                            const dgk_opts* opts, dgk_batch** out) {
  if (!out) return bfail(nullptr, DGK_ERR_INVALID, "out must not be NULL");
  *out = nullptr;
  dgk_batch* b = new dgk_batch();
  dgk_status st = bguard(b, [&] { batch_create_impl(*b, n, meshes, phys, dt, opts); });
  if (st != DGK_OK) {
    delete b;
    return st;
  }
  *out = b;
  return DGK_OK;
}

dgk_status dgk_batch_destroy(dgk_batch* b) {
  delete b;
  return DGK_OK;
}

const char* dgk_batch_last_error(const dgk_batch* b) { return b ? b->err.c_str() : dgk::g_err.c_str(); }

}  // extern "C"

namespace {

// Copy simulation `sim`'s state between host u[4][K_i*Np] (FP64) and its slice of the group planes.
dgk_status batch_state_io(dgk_batch* b, int sim, double* u, bool to_device) {
  if (!b || !u || sim < 0 || sim >= (int)b->where.size()) return bfail(b, DGK_ERR_INVALID, "bad batch state arguments");
  return bguard(b, [&] {
    auto w = b->where[sim];
    dgk::Ctx& c = *b->groups[w.first].ctx;
    DGK_HIP(hipStreamSynchronize(b->main));
    const int64_t off = b->groups[w.first].off[w.second] * c.Np, n = b->K[sim] * c.Np;
    const bool f64 = c.precision == DGK_FP64;
    if (to_device) {  // the group's live planes: the most any member needs
      b->nq[sim] = dgk::live_planes(c.dim, u, n);
      int nq = c.dim + 1;
      for (int i : b->groups[w.first].sims) nq = std::max(nq, b->nq[i]);
      if (nq != c.nq && !b->graphs.empty()) {
        b->drop_graphs();
      }
      if (nq != c.nq && c.es > 1) {
        // interleaved buffers hold exactly the live fields: relay the whole group's state
        double* s = dgk::staging(c);
        uint16_t* lo = c.comp ? c.Ulo : nullptr;
        dgk::state_to_f64(c, c.U, lo, s, 0, c.K * c.Np);
        c.es = nq;
        dgk::state_from_f64(c, s, c.U, lo, 0, c.K * c.Np);
      }
      c.nq = nq;
    }
    // FP64 field planes are copied directly; otherwise through the FP64 staging planes
    const bool direct = f64 && c.es == 1;
    double* s = direct ? nullptr : dgk::staging(c);
    uint16_t* lo = c.comp ? c.Ulo : nullptr;
    if (!to_device && !direct) dgk::state_to_f64(c, c.U, lo, s, off, n);
    for (int q = 0; q < 4; ++q) {
      const int64_t o = q * c.PS + off;
      void* dev = direct ? (void*)((char*)c.U + o * c.tsize) : (void*)(s + o);
      double* host = u + q * n;
      if (to_device)
        DGK_HIP(hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
      else
        DGK_HIP(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    }
    if (to_device && !direct) dgk::state_from_f64(c, s, c.U, lo, off, n);
    DGK_HIP(hipStreamSynchronize(c.stream));
  });
}

}  // namespace

extern "C" {

dgk_status dgk_batch_set_state(dgk_batch* b, int sim, const double* u) {
  return batch_state_io(b, sim, const_cast<double*>(u), true);
}

dgk_status dgk_batch_get_state(dgk_batch* b, int sim, double* u) { return batch_state_io(b, sim, u, false); }

dgk_status dgk_batch_rk4(dgk_batch* b, int nsteps, int schedule) {
  if (!b) return bfail(nullptr, DGK_ERR_INVALID, "NULL batch");
  return bguard(b, [&] { batch_rk4_impl(*b, nsteps, schedule); });
}

dgk_status dgk_batch_set_receivers(dgk_batch* b, int sim, int n, const double* xyz, const double* node_xyz,
                                   int capacity) {
  if (!b || sim < 0 || sim >= (int)b->where.size() || n < 0 || (n > 0 && (!xyz || !node_xyz)))
    return bfail(b, DGK_ERR_INVALID, "bad receiver arguments");
  return bguard(b, [&] {
    DGK_HIP(hipStreamSynchronize(b->main));
    auto w = b->where[sim];
    auto& G = b->groups[w.first];
    dgk::Ctx& c = *G.ctx;
    const size_t nm = G.sims.size();
    G.rel.resize(nm);
    G.rw.resize(nm);
    G.rstart.resize(nm, 0);
    if (n > 0 && capacity < 1) throw dgk::InvalidArg("receiver capacity must be positive");
    // rows recorded so far for the other members are kept; this member's series restarts
    std::vector<double> old;
    const int rows = dgk::fetch_receivers(c, old);
    std::vector<size_t> col0(nm, 0);
    for (size_t k = 1; k < nm; ++k) col0[k] = col0[k - 1] + G.rel[k - 1].size();
    const size_t nold = c.rec.n;
    dgk::RefElem R = dgk::make_ref_elem(c.dim, c.p);
    std::vector<double> wt;
    std::vector<int64_t> el = dgk::locate_points(R, node_xyz, b->K[sim], n, xyz, wt);
    G.rel[w.second] = el;
    G.rw[w.second] = wt;
    G.rstart[w.second] = rows;
    G.rcap = std::max(G.rcap, rows + capacity);  // room for `capacity` more steps
    std::vector<int64_t> all;
    std::vector<double> allw;
    for (size_t k = 0; k < nm; ++k) {
      for (int64_t e : G.rel[k]) all.push_back(e + G.off[k]);
      allw.insert(allw.end(), G.rw[k].begin(), G.rw[k].end());
    }
    std::vector<double> keep;
    if (rows > 0 && !all.empty()) {
      keep.assign((size_t)rows * all.size(), std::nan(""));
      size_t c1 = 0;
      for (size_t k = 0; k < nm; ++k) {
        if ((int)k != w.second)
          for (int i = 0; i < rows; ++i)
            for (size_t j = 0; j < G.rel[k].size(); ++j) keep[(size_t)i * all.size() + c1 + j] = old[(size_t)i * nold + col0[k] + j];
        c1 += G.rel[k].size();
      }
    }
    dgk::install_receivers(c, all, allw, G.rcap);
    if (!keep.empty()) dgk::restore_receivers(c, keep, rows);
    b->drop_graphs();
  });
}

dgk_status dgk_batch_get_receivers(dgk_batch* b, int sim, double* out, int capacity_rows, int* rows) {
  if (!b || sim < 0 || sim >= (int)b->where.size() || !rows) return bfail(b, DGK_ERR_INVALID, "bad receiver arguments");
  return bguard(b, [&] {
    DGK_HIP(hipStreamSynchronize(b->main));
    auto w = b->where[sim];
    auto& G = b->groups[w.first];
    std::vector<double> h;
    const int r = dgk::fetch_receivers(*G.ctx, h);
    size_t col0 = 0, n = 0;
    if (w.second < (int)G.rel.size()) {
      for (int k = 0; k < w.second; ++k) col0 += G.rel[k].size();
      n = G.rel[w.second].size();
    }
    const int r0 = n ? std::min(G.rstart[w.second], r) : 0;  // the member's series begins at row r0
    *rows = n ? r - r0 : 0;
    if (out && n) {
      if (r - r0 > capacity_rows) throw dgk::InvalidArg("output buffer holds fewer rows than recorded");
      const size_t ng = G.ctx->rec.n;
      for (int i = r0; i < r; ++i)
        for (size_t j = 0; j < n; ++j) out[(size_t)(i - r0) * n + j] = h[(size_t)i * ng + col0 + j];
    }
  });
}

dgk_status dgk_batch_set_launch(dgk_batch* b, int group, int grid_pct, int tile_order) {
  if (!b || group < 0 || group >= (int)b->groups.size() || grid_pct < 1 || grid_pct > 10000 || tile_order < 0 ||
      tile_order > 1)
    return bfail(b, DGK_ERR_INVALID, "bad launch settings");
  return bguard(b, [&] {
    DGK_HIP(hipStreamSynchronize(b->main));
    dgk::Ctx& c = *b->groups[group].ctx;
    c.grid_pct = grid_pct;
    c.tile_order = tile_order;
    b->drop_graphs();
  });
}

dgk_status dgk_batch_set_tile_shape(dgk_batch* b, int group, int shape) {
  if (!b || group < 0 || group >= (int)b->groups.size() || shape < 0 || shape >= dgk::kTileShapes)
    return bfail(b, DGK_ERR_INVALID, "bad tile shape");
  return bguard(b, [&] {
    dgk::Ctx& c = *b->groups[group].ctx;
    dgk::StageArgs a = dgk::base_args(c);
    if (!dgk::tile_shape_supported(c, a, shape))
      throw dgk::Unsupported("tile shape " + std::to_string(shape) + " has no kernel for this group's form");
    DGK_HIP(hipStreamSynchronize(b->main));
    c.tile_shape = shape;
    b->drop_graphs();
  });
}

dgk_status dgk_batch_prepare(dgk_batch* b, int nsteps, int schedule) {
  if (!b) return bfail(nullptr, DGK_ERR_INVALID, "NULL batch");
  return bguard(b, [&] {
    batch_check(nsteps, schedule);
    if (nsteps > 0) batch_graph(*b, std::min(nsteps, kMaxGraphSteps), schedule);
    if (nsteps > kMaxGraphSteps && nsteps % kMaxGraphSteps) batch_graph(*b, nsteps % kMaxGraphSteps, schedule);
  });
}

dgk_status dgk_batch_rk4_timed(dgk_batch* b, int nsteps, double* group_stage_ms) {
  if (!b || nsteps < 1 || !group_stage_ms) return bfail(b, DGK_ERR_INVALID, "bad arguments");
  return bguard(b, [&] {
    DGK_HIP(hipStreamSynchronize(b->main));
    for (size_t g = 0; g < b->groups.size(); ++g) {
      dgk::Ctx& c = *b->groups[g].ctx;
      std::vector<hipEvent_t> ev(8 * nsteps);
      for (auto& e : ev) DGK_HIP(hipEventCreate(&e));
      try {
        for (int i = 0; i < nsteps; ++i) dgk::rk4_step(c, 1.0, c.stream, &ev[8 * i]);
        DGK_HIP(hipEventSynchronize(ev.back()));
        double sum = 0;
        for (int i = 0; i < 4 * nsteps; ++i) {
          float ms = 0;
          DGK_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
          sum += ms;
        }
        group_stage_ms[g] = sum;
      } catch (...) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        throw;
      }
      for (auto& e : ev) (void)hipEventDestroy(e);
    }
  });
}

dgk_status dgk_batch_sync(dgk_batch* b) {
  if (!b) return bfail(nullptr, DGK_ERR_INVALID, "NULL batch");
  return bguard(b, [&] {
    DGK_HIP(hipStreamSynchronize(b->main));
    for (auto& G : b->groups) DGK_HIP(hipStreamSynchronize(G.ctx->stream));
  });
}

dgk_status dgk_batch_groups(const dgk_batch* b, int* groups, int* dim, int* order, int64_t* elements, int* sims,
                            int64_t* faces, int64_t* boundary_faces) {
  if (!b) return bfail(nullptr, DGK_ERR_INVALID, "NULL batch");
  if (groups) *groups = (int)b->groups.size();
  for (size_t g = 0; g < b->groups.size(); ++g) {
    const dgk::Ctx& c = *b->groups[g].ctx;
    if (dim) dim[g] = c.dim;
    if (order) order[g] = c.p;
    if (elements) elements[g] = c.K;
    if (sims) sims[g] = (int)b->groups[g].sims.size();
    if (faces) faces[g] = c.F;
    if (boundary_faces) boundary_faces[g] = c.Fb;
  }
  return DGK_OK;
}

}  // extern "C"
