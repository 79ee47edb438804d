// Device-side setup: face connectivity built on the GPU and per-element affine geometry.
//
// Replaces the reference's quadratic face search and neighbour bookkeeping
// (src/Mesh.cpp:107-128, 235-314, 679-713) with: one key per element face (its sorted
// corner-vertex ids: 32 bits in 1D, 64 in 2D, 96 in 3D as a 32-bit high and a 64-bit low part),
// a device radix sort of (key, element-face) pairs (3D: two stable passes, low part first), a
// pairing pass over equal neighbouring keys and a finalisation pass that records the neighbour
// element, its local face, the vertex permutation between the two faces and the interior
// penalty sign. Radix sort is stable, so for a shared face the lower element index comes
// first, which is the reference's "up" side (src/Mesh.cpp:236-246).
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "dgk_internal.hpp"

namespace dgk {

namespace {

__constant__ int c_fv[4][3];

__global__ void k_face_keys(int64_t K, int D, int NF, const int32_t* __restrict__ vid, uint64_t* __restrict__ keys,
                            uint32_t* __restrict__ hi, int32_t* __restrict__ vals) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= K * NF) return;
  int64_t e = t / NF;
  int j = (int)(t % NF);
  uint64_t a[3] = {0, 0, 0};
  for (int m = 0; m < D; ++m) a[m] = (uint64_t)(uint32_t)vid[e * (D + 1) + c_fv[j][m]];
  // sort up to three ids
  if (D >= 2 && a[0] > a[1]) { uint64_t x = a[0]; a[0] = a[1]; a[1] = x; }
  if (D == 3) {
    if (a[1] > a[2]) { uint64_t x = a[1]; a[1] = a[2]; a[2] = x; }
    if (a[0] > a[1]) { uint64_t x = a[0]; a[0] = a[1]; a[1] = x; }
  }
  // 3D: keys holds the low part (a1, a2), hi the smallest id a0
  keys[t] = D == 3 ? (a[1] << 32) | a[2] : D == 2 ? (a[0] << 32) | a[1] : a[0];
  if (D == 3) hi[t] = (uint32_t)a[0];
  vals[t] = (int32_t)t;
}

// high key parts in the order of the first (low-part) sort
__global__ void k_gather_hi(int64_t n, const uint32_t* __restrict__ hi, const int32_t* __restrict__ vals,
                            uint32_t* __restrict__ out) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = hi[vals[i]];
}

// vals: element faces in key order; keys (and hi, 3D) indexed by element face
__global__ void k_face_pair(int64_t n, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ hi,
                            const int32_t* __restrict__ vals, int32_t* __restrict__ partner, int* __restrict__ err) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto same = [&](int64_t a, int64_t b) {
    return keys[vals[a]] == keys[vals[b]] && (!hi || hi[vals[a]] == hi[vals[b]]);
  };
  bool prev = i > 0 && same(i - 1, i);
  bool next = i + 1 < n && same(i + 1, i);
  if (prev && next) atomicOr(err, 1);  // a face shared by three or more elements
  partner[vals[i]] = prev ? vals[i - 1] : next ? vals[i + 1] : -1;
}

__global__ void k_face_finalize(int64_t K, int D, int NF, int NPERM, const int32_t* __restrict__ vid,
                                const double* __restrict__ vxyz, const int32_t* __restrict__ partner,
                                const uint8_t* __restrict__ bc_first, int sign_reference, int fc, double sigma,
                                int32_t* __restrict__ nbr, int32_t* __restrict__ code, int32_t* __restrict__ nface,
                                int32_t* __restrict__ perm, int* __restrict__ err) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= K * NF) return;
  int64_t e = t / NF;
  int j = (int)(t % NF);
  int32_t pt = partner[t];
  if (pt < 0) {
    nbr[t] = bc_first[t] ? -2 : -1;
    code[t] = 0;
    nface[t] = -1;
    perm[t] = -1;
    return;
  }
  int64_t nb = pt / NF;
  int jn = pt % NF;
  int A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, pi[3] = {0, 0, 0};
  for (int m = 0; m < D; ++m) {
    A[m] = vid[e * (D + 1) + c_fv[j][m]];
    B[m] = vid[nb * (D + 1) + c_fv[jn][m]];
  }
  for (int m = 0; m < D; ++m) {
    pi[m] = -1;
    for (int q = 0; q < D; ++q)
      if (B[q] == A[m]) pi[m] = q;
    if (pi[m] < 0) atomicOr(err, 2);
  }
  int pidx = 0;
  if (D == 3) {
    const int P6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    pidx = -1;
    for (int q = 0; q < 6; ++q)
      if (P6[q][0] == pi[0] && P6[q][1] == pi[1] && P6[q][2] == pi[2]) pidx = q;
    if (pidx < 0) { atomicOr(err, 2); pidx = 0; }
  } else if (D == 2) {
    pidx = pi[0] == 0 ? 0 : 1;
  }
  int sign = 1;
  if (sign_reference) {
    // fc * orientation of the lower-index element w.r.t. the face normal taken at the first
    // face quadrature point (src/Mesh.cpp:164-194, 210-211, 278-294)
    int64_t up = e < nb ? e : nb;
    int jup = e < nb ? j : jn;
    double X[3][3] = {}, bc[3] = {0, 0, 0};
    for (int m = 0; m < D; ++m)
      for (int x = 0; x < 3; ++x) X[m][x] = vxyz[(up * (D + 1) + c_fv[jup][m]) * 3 + x];
    for (int v = 0; v <= D; ++v)
      for (int x = 0; x < 3; ++x) bc[x] += vxyz[(up * (D + 1) + v) * 3 + x] / (D + 1);
    double n[3] = {1, 0, 0};
    if (D == 3) {
      double a[3], b[3];
      for (int x = 0; x < 3; ++x) { a[x] = X[1][x] - X[0][x]; b[x] = X[2][x] - X[0][x]; }
      n[0] = a[1] * b[2] - a[2] * b[1];
      n[1] = a[2] * b[0] - a[0] * b[2];
      n[2] = a[0] * b[1] - a[1] * b[0];
    } else if (D == 2) {
      n[0] = X[1][1] - X[0][1];
      n[1] = -(X[1][0] - X[0][0]);
      n[2] = 0;
    }
    double s = (D == 1) ? 1.0 : sigma;
    double dp = 0;
    for (int x = 0; x < 3; ++x) dp += (X[0][x] - bc[x]) * s * n[x];
    int o = dp >= 0 ? 1 : -1;
    sign = fc * o;
  }
  nbr[t] = (int32_t)nb;
  nface[t] = jn;
  perm[t] = pidx;
  code[t] = (jn * NPERM + pidx) | (sign < 0 ? (1 << 16) : 0);
}

// Affine geometry: Jinv[r][x] = d r / d x (Gmsh reference coordinates), stored 3x3 row-major.
struct FaceRef {
  double n[4][3];  // outward reference normals
  double fsc[4];   // reference face measure / face reference-element measure
};

// Affine geometry: Jinv[r][x] = d r / d x (Gmsh reference coordinates), stored 3x3 row-major,
// and per element face the unit outward normal J^-T n_ref / |J^-T n_ref| and the face scale
// sJ/|detJ| = |J^-T n_ref| * fsc, computed in FP64 and rounded once to T.
template <typename T>
__global__ void k_geometry(int64_t K, int D, const double* __restrict__ vxyz, T* __restrict__ Jinv,
                           T* __restrict__ fgeom, FaceRef fr, int* __restrict__ err) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= K) return;
  const double* X = vxyz + e * (D + 1) * 3;
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // J[x][u]
  double inv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (D == 1) {
    double j = (X[3] - X[0]) / 2;  // reference line is [-1, 1]
    if (j == 0) atomicOr(err, 4);
    inv[0][0] = 1.0 / j;
  } else if (D == 2) {
    for (int u = 0; u < 2; ++u)
      for (int x = 0; x < 2; ++x) J[x][u] = X[(u + 1) * 3 + x] - X[x];
    double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    if (det == 0) atomicOr(err, 4);
    inv[0][0] = J[1][1] / det;
    inv[0][1] = -J[0][1] / det;
    inv[1][0] = -J[1][0] / det;
    inv[1][1] = J[0][0] / det;
  } else {
    for (int u = 0; u < 3; ++u)
      for (int x = 0; x < 3; ++x) J[x][u] = X[(u + 1) * 3 + x] - X[x];
    double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (det == 0) atomicOr(err, 4);
    inv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    inv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    inv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  }
  for (int r = 0; r < 3; ++r)
    for (int x = 0; x < 3; ++x) Jinv[e * 9 + r * 3 + x] = (T)inv[r][x];
  for (int j = 0; j <= D; ++j) {
    double n[3];
    for (int x = 0; x < 3; ++x) {
      double s = 0;
      for (int r = 0; r < D; ++r) s += inv[r][x] * fr.n[j][r];
      n[x] = s;
    }
    double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int x = 0; x < 3; ++x) fgeom[(e * (D + 1) + j) * 4 + x] = (T)(n[x] / len);
    fgeom[(e * (D + 1) + j) * 4 + 3] = (T)(len * fr.fsc[j]);
  }
}

}  // namespace

void build_connectivity(Ctx& c, const double* vxyz_dev, const int32_t* vid_host, const uint8_t* bc_first_host,
                        double sigma) {
  const int D = c.dim, NF = c.Nf;
  const int64_t n = c.K * NF;
  if (n >= (int64_t)1 << 31) throw Unsupported("more than 2^31 element faces");
  int fv[4][3] = {};
  if (D == 1) { fv[0][0] = 0; fv[1][0] = 1; }
  if (D == 2) { int t[3][2] = {{0, 1}, {1, 2}, {2, 0}}; for (int j = 0; j < 3; ++j) for (int m = 0; m < 2; ++m) fv[j][m] = t[j][m]; }
  if (D == 3) { int t[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {3, 1, 2}}; for (int j = 0; j < 4; ++j) for (int m = 0; m < 3; ++m) fv[j][m] = t[j][m]; }
  DGK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_fv), fv, sizeof(fv)));
  hipStream_t s = c.stream;
  int32_t *vid = nullptr, *vals = nullptr, *vals2 = nullptr, *partner = nullptr;
  uint64_t *keys = nullptr, *keys2 = nullptr;
  uint32_t *hi = nullptr, *hi2 = nullptr, *hi3 = nullptr;
  uint8_t* bcf = nullptr;
  int* err = nullptr;
  void* tmp = nullptr;
  auto cleanup = [&]() {
    for (void* p : {(void*)vid, (void*)vals, (void*)vals2, (void*)partner, (void*)keys, (void*)keys2, (void*)hi,
                    (void*)hi2, (void*)hi3, (void*)bcf, (void*)err, tmp})
      if (p) (void)hipFree(p);
  };
  try {
    DGK_HIP(hipMalloc(&vid, sizeof(int32_t) * c.K * (D + 1)));
    DGK_HIP(hipMalloc(&vals, sizeof(int32_t) * n));
    DGK_HIP(hipMalloc(&vals2, sizeof(int32_t) * n));
    DGK_HIP(hipMalloc(&partner, sizeof(int32_t) * n));
    DGK_HIP(hipMalloc(&keys, sizeof(uint64_t) * n));
    DGK_HIP(hipMalloc(&keys2, sizeof(uint64_t) * n));
    if (D == 3) {
      DGK_HIP(hipMalloc(&hi, sizeof(uint32_t) * n));
      DGK_HIP(hipMalloc(&hi2, sizeof(uint32_t) * n));
      DGK_HIP(hipMalloc(&hi3, sizeof(uint32_t) * n));
    }
    DGK_HIP(hipMalloc(&bcf, n));
    DGK_HIP(hipMalloc(&err, sizeof(int)));
    DGK_HIP(hipMemcpyAsync(vid, vid_host, sizeof(int32_t) * c.K * (D + 1), hipMemcpyHostToDevice, s));
    DGK_HIP(hipMemcpyAsync(bcf, bc_first_host, n, hipMemcpyHostToDevice, s));
    DGK_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    int bs = 256;
    int64_t nb = (n + bs - 1) / bs;
    hipLaunchKernelGGL(k_face_keys, dim3((unsigned)nb), dim3(bs), 0, s, c.K, D, NF, vid, keys, hi, vals);
    DGK_HIP(hipGetLastError());
    size_t tmp_bytes = 0, tmp_hi = 0;
    const int end_bit = D == 1 ? 32 : 64;
    DGK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys2, vals, vals2, (size_t)n, 0, end_bit, s));
    if (D == 3) {
      DGK_HIP(rocprim::radix_sort_pairs(nullptr, tmp_hi, hi2, hi3, vals2, vals, (size_t)n, 0, 32, s));
      tmp_bytes = std::max(tmp_bytes, tmp_hi);
    }
    DGK_HIP(hipMalloc(&tmp, tmp_bytes));
    DGK_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys2, vals, vals2, (size_t)n, 0, end_bit, s));
    const int32_t* order = vals2;
    if (D == 3) {
      // second, stable pass on the high part: faces with equal (a0, a1, a2) end up adjacent, in
      // element-face order
      hipLaunchKernelGGL(k_gather_hi, dim3((unsigned)nb), dim3(bs), 0, s, n, hi, vals2, hi2);
      DGK_HIP(hipGetLastError());
      DGK_HIP(rocprim::radix_sort_pairs(tmp, tmp_bytes, hi2, hi3, vals2, vals, (size_t)n, 0, 32, s));
      order = vals;
    }
    hipLaunchKernelGGL(k_face_pair, dim3((unsigned)nb), dim3(bs), 0, s, n, keys, (const uint32_t*)hi, order, partner,
                       err);
    DGK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_face_finalize, dim3((unsigned)nb), dim3(bs), 0, s, c.K, D, NF, c.nperm, vid, vxyz_dev,
                       partner, bcf, c.sign_mode == DGK_SIGN_REFERENCE ? 1 : 0, (int)c.fc, sigma, c.nbr, c.code,
                       c.nface, c.perm, err);
    DGK_HIP(hipGetLastError());
    int herr = 0;
    DGK_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    DGK_HIP(hipStreamSynchronize(s));
    if (herr & 1) throw MeshError("a face is shared by more than two elements");
    if (herr & 2) throw MeshError("matching faces do not share their vertices");
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

void build_geometry(Ctx& c, const double* vxyz_dev) {
  int* err = nullptr;
  DGK_HIP(hipMalloc(&err, sizeof(int)));
  DGK_HIP(hipMemsetAsync(err, 0, sizeof(int), c.stream));
  int bs = 256;
  unsigned nb = (unsigned)((c.K + bs - 1) / bs);
  FaceRef fr{};
  for (int j = 0; j < 4; ++j) {
    for (int x = 0; x < 3; ++x) fr.n[j][x] = c.refn[j][x];
    fr.fsc[j] = c.fsc[j];
  }
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_geometry<double>, dim3(nb), dim3(bs), 0, c.stream, c.K, c.dim, vxyz_dev, (double*)c.Jinv,
                       (double*)c.fgeom, fr, err);
  else
    hipLaunchKernelGGL(k_geometry<float>, dim3(nb), dim3(bs), 0, c.stream, c.K, c.dim, vxyz_dev, (float*)c.Jinv,
                       (float*)c.fgeom, fr, err);
  int herr = 0;
  hipError_t e1 = hipGetLastError();
  hipError_t e2 = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c.stream);
  hipError_t e3 = hipStreamSynchronize(c.stream);
  (void)hipFree(err);
  DGK_HIP(e1);
  DGK_HIP(e2);
  DGK_HIP(e3);
  if (herr) throw MeshError("degenerate element (zero Jacobian)");
}

namespace {
template <typename T>
__global__ void k_scale_geometry(int64_t e0, int64_t n, int NF, double s, T* __restrict__ Jinv,
                                 T* __restrict__ fgeom) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t e = e0 + i;
  for (int c = 0; c < 9; ++c) Jinv[e * 9 + c] = (T)((double)Jinv[e * 9 + c] * s);
  for (int j = 0; j < NF; ++j) fgeom[(e * NF + j) * 4 + 3] = (T)((double)fgeom[(e * NF + j) * 4 + 3] * s);
}
}  // namespace

// Multiply the volume (J^-1) and face (sJ/|detJ|) scales of elements [e0, e0+n) by s: with s = dt
// the stage operator returns dt*L(u) for a launch time step of 1.
void scale_geometry(Ctx& c, int64_t e0, int64_t n, double s) {
  if (n <= 0) return;
  unsigned nb = (unsigned)((n + 255) / 256);
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_scale_geometry<double>, dim3(nb), dim3(256), 0, c.stream, e0, n, c.Nf, s, (double*)c.Jinv,
                       (double*)c.fgeom);
  else
    hipLaunchKernelGGL(k_scale_geometry<float>, dim3(nb), dim3(256), 0, c.stream, e0, n, c.Nf, s, (float*)c.Jinv,
                       (float*)c.fgeom);
  DGK_HIP(hipGetLastError());
}

}  // namespace dgk
