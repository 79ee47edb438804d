// Fused DG stage kernel: one launch evaluates k = dt*L(v) for a block of elements and
// applies the RK4 (or forward-Euler / plain rhs) update in its epilogue.
//
// Per element and equation the reference computes (src/solver.cpp:35-52,
// src/Mesh.cpp:476-557, 569-674)
//     k = dt * Minv^T (S(F(v)) - Fface(v))
// with quadrature-point gradients and a stored inverse mass matrix per element. On affine
// elements this equals, exactly in real arithmetic,
//     k = dt * ( sum_r W_r Ft_r  -  sum_j LIFT_j (sJ_j/|detJ|) (n.F*)_j )
// with W_r = Mhat^-1 Shat_r, LIFT_j = Mhat^-1 E_j Mhat_f, Ft_r = sum_x Jinv[r][x] F_x the
// contravariant nodal flux, and (n.F*)_j the nodal normal numerical flux on face j:
//   interior    1/2 n.(F(u-) + F(u+)) + 1/2 s c0 (u- - u+)      (Rusanov with speed c0, sign s)
//   reflecting  n.F(u~) with u~ = (p, v - (v.n) n)               (src/Mesh.cpp:616-647)
//   absorbing   1/4 [c0 p + rho c0^2 v.n ; n (p/rho + c0 v.n)]    (RKR, src/Mesh.cpp:391-418, 652-667)
// The face terms are exact at face nodes because every integrand is a polynomial of degree
// <= 2p on straight faces.
//
// Data layout in HBM is the reference's state layout u[eq][element][node]; a workgroup
// stages a tile of TE consecutive elements in LDS (coalesced loads of TE*Np contiguous
// values per equation), gathers neighbour traces through the device-built face table,
// and walks tiles in an XCD-aware order (each XCD sweeps one contiguous slab of the mesh,
// so neighbour traces are reused from that XCD's L2).
#include <algorithm>

#include "dgk_internal.hpp"

namespace dgk {

template <int D, int P>
struct Sz {
  static constexpr int NP = D == 1 ? P + 1 : D == 2 ? (P + 1) * (P + 2) / 2 : (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NF = D + 1;
  static constexpr int NFP = D == 1 ? 1 : D == 2 ? P + 1 : (P + 1) * (P + 2) / 2;
  static constexpr int NPERM = D == 3 ? 6 : D == 2 ? 2 : 1;
};

namespace {

constexpr int kBlock = 256;

template <typename T, int D, int P>
__global__ __launch_bounds__(kBlock) void k_stage_lds(StageArgs a) {
  using S = Sz<D, P>;
  constexpr int NP = S::NP, NF = S::NF, NFP = S::NFP, NPERM = S::NPERM;
  constexpr int TE = kBlock / NP;
  constexpr int NFF = NF * NFP;
  __shared__ T sU[4][TE * NP];
  __shared__ T sF[4][D][TE * NP];
  __shared__ T sFl[4][TE * NFF];
  __shared__ T sG[TE][9];
  __shared__ T sN[TE][NF][4];

  const int tid = threadIdx.x;
  const int64_t K = a.K, KN = a.K * NP;
  const T* __restrict__ Vin = (const T*)a.Vin;
  T* __restrict__ Vout = (T*)a.Vout;
  T* __restrict__ U = (T*)a.U;
  T* __restrict__ Acc = (T*)a.A;
  const T* __restrict__ Jinv = (const T*)a.Jinv;
  const T* __restrict__ W = (const T*)a.W;
  const T* __restrict__ LIFT = (const T*)a.LIFT;
  const T dt = (T)a.dt, c0 = (T)a.c0, rho = (T)a.rho0;
  const T rc2 = (T)(a.rho0 * a.c0 * a.c0), irho = (T)(1.0 / a.rho0);
  const T v0[3] = {(T)a.v0[0], (T)a.v0[1], (T)a.v0[2]};
  const int stage = a.stage;

  const int xcd = blockIdx.x % 8, lb = blockIdx.x / 8, bpx = gridDim.x / 8;
  for (int64_t it = lb; it < a.tiles_per_xcd; it += bpx) {
    const int64_t tile = (int64_t)xcd * a.tiles_per_xcd + it;
    if (tile >= a.tiles) break;
    const int64_t e0 = tile * TE;
    const int ne = (int)((K - e0) < TE ? (K - e0) : TE);
    __syncthreads();
    for (int idx = tid; idx < 4 * TE * NP; idx += kBlock) {
      int q = idx / (TE * NP), r = idx - q * (TE * NP);
      sU[q][r] = r < ne * NP ? Vin[q * KN + e0 * NP + r] : (T)0;
    }
    for (int idx = tid; idx < TE * 9; idx += kBlock) {
      int e = idx / 9, c = idx - e * 9;
      sG[e][c] = e < ne ? Jinv[(e0 + e) * 9 + c] : (T)0;
    }
    __syncthreads();
    // outward unit normals and face scale sJ/|detJ| from the inverse Jacobian
    for (int idx = tid; idx < TE * NF; idx += kBlock) {
      int e = idx / NF, j = idx - e * NF;
      T n[3];
      for (int x = 0; x < 3; ++x) {
        T s = 0;
        for (int r = 0; r < D; ++r) s += sG[e][r * 3 + x] * (T)a.refn[j][r];
        n[x] = s;
      }
      T len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      T il = len > 0 ? (T)1 / len : (T)0;
      sN[e][j][0] = n[0] * il;
      sN[e][j][1] = n[1] * il;
      sN[e][j][2] = n[2] * il;
      sN[e][j][3] = len * (T)a.fsc[j];
    }
    // contravariant nodal fluxes Ft_r = sum_x Jinv[r][x] F_x (src/Mesh.cpp:577-591)
    for (int idx = tid; idx < TE * NP; idx += kBlock) {
      int e = idx / NP;
      T p = sU[0][idx], v[3] = {sU[1][idx], sU[2][idx], sU[3][idx]};
      T F[4][3];
      for (int x = 0; x < 3; ++x) {
        F[0][x] = v0[x] * p + rc2 * v[x];
        for (int k = 0; k < 3; ++k) F[1 + k][x] = v0[x] * v[k] + (k == x ? p * irho : (T)0);
      }
      for (int r = 0; r < D; ++r)
        for (int q = 0; q < 4; ++q)
          sF[q][r][idx] = sG[e][r * 3 + 0] * F[q][0] + sG[e][r * 3 + 1] * F[q][1] + sG[e][r * 3 + 2] * F[q][2];
    }
    __syncthreads();
    // nodal normal numerical flux at every face node, scaled by sJ/|detJ|
    for (int idx = tid; idx < TE * NFF; idx += kBlock) {
      int e = idx / NFF, jk = idx - e * NFF, j = jk / NFP, k = jk - j * NFP;
      if (e >= ne) {
        for (int q = 0; q < 4; ++q) sFl[q][idx] = 0;
        continue;
      }
      const int on = a.fmask[j * NFP + k];
      T qm[4] = {sU[0][e * NP + on], sU[1][e * NP + on], sU[2][e * NP + on], sU[3][e * NP + on]};
      const T n[3] = {sN[e][j][0], sN[e][j][1], sN[e][j][2]};
      const T fs = sN[e][j][3];
      const int64_t eg = e0 + e;
      const int32_t nb = a.nbr[eg * NF + j];
      const T vn = v0[0] * n[0] + v0[1] * n[1] + v0[2] * n[2];
      T fl[4];
      if (nb >= 0) {
        const int32_t cd = a.code[eg * NF + j];
        const int node = a.vmap[(cd & 0xffff) * NFP + k];
        const T sgn = (cd >> 16) & 1 ? (T)-1 : (T)1;
        const int64_t pi = (int64_t)nb * NP + node;
        T qp[4] = {Vin[pi], Vin[KN + pi], Vin[2 * KN + pi], Vin[3 * KN + pi]};
        T unm = qm[1] * n[0] + qm[2] * n[1] + qm[3] * n[2];
        T unp = qp[1] * n[0] + qp[2] * n[1] + qp[3] * n[2];
        T hc = (T)0.5 * sgn * c0;
        fl[0] = (T)0.5 * ((vn * qm[0] + rc2 * unm) + (vn * qp[0] + rc2 * unp)) + hc * (qm[0] - qp[0]);
        for (int c = 0; c < 3; ++c)
          fl[1 + c] = (T)0.5 * ((vn * qm[1 + c] + qm[0] * irho * n[c]) + (vn * qp[1 + c] + qp[0] * irho * n[c])) +
                      hc * (qm[1 + c] - qp[1 + c]);
      } else if (nb == -2) {
        T un = qm[1] * n[0] + qm[2] * n[1] + qm[3] * n[2];
        T g[3] = {qm[1] - un * n[0], qm[2] - un * n[1], qm[3] - un * n[2]};
        T ung = g[0] * n[0] + g[1] * n[1] + g[2] * n[2];
        fl[0] = vn * qm[0] + rc2 * ung;
        for (int c = 0; c < 3; ++c) fl[1 + c] = vn * g[c] + qm[0] * irho * n[c];
      } else {
        T un = qm[1] * n[0] + qm[2] * n[1] + qm[3] * n[2];
        fl[0] = (T)0.25 * (c0 * qm[0] + rc2 * un);
        T t = (T)0.25 * (qm[0] * irho + c0 * un);
        for (int c = 0; c < 3; ++c) fl[1 + c] = t * n[c];
      }
      for (int q = 0; q < 4; ++q) sFl[q][idx] = fs * fl[q];
    }
    __syncthreads();
    // volume + lift contraction and the time-integration epilogue
    for (int idx = tid; idx < TE * NP; idx += kBlock) {
      int e = idx / NP, i = idx - e * NP;
      if (e >= ne) continue;
      T acc[4] = {0, 0, 0, 0};
      for (int r = 0; r < D; ++r)
        for (int n = 0; n < NP; ++n) {
          T w = W[(r * NP + i) * NP + n];
          for (int q = 0; q < 4; ++q) acc[q] += w * sF[q][r][e * NP + n];
        }
      for (int jk = 0; jk < NFF; ++jk) {
        T l = LIFT[i * NFF + jk];
        for (int q = 0; q < 4; ++q) acc[q] -= l * sFl[q][e * NFF + jk];
      }
      const int64_t gi = e0 * NP + idx;
      for (int q = 0; q < 4; ++q) {
        const T kq = dt * acc[q];
        const int64_t id = q * KN + gi;
        switch (stage) {
          case 0: { T u = sU[q][idx]; Acc[id] = u + kq / (T)6; Vout[id] = u + kq / (T)2; } break;
          case 1: { T u = U[id]; Acc[id] = Acc[id] + kq / (T)3; Vout[id] = u + kq / (T)2; } break;
          case 2: { T u = U[id]; Acc[id] = Acc[id] + kq / (T)3; Vout[id] = u + kq; } break;
          case 3: U[id] = Acc[id] + kq / (T)6; break;
          case 4: Vout[id] = kq; break;
          default: Vout[id] = sU[q][idx] + kq; break;
        }
      }
    }
  }
}

template <typename T>
__global__ void k_source(T* __restrict__ U, const int* __restrict__ idx, int64_t n, T value) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) U[idx[i]] = value;
}

template <typename T>
__global__ void k_finite(const T* __restrict__ x, int64_t n, int* __restrict__ flag) {
  int bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bad |= !isfinite(x[i]);
  if (__any(bad) && (threadIdx.x % 64) == 0) atomicOr(flag, 1);
}

__global__ void k_d2f(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (float)s[i];
}
__global__ void k_f2d(const float* __restrict__ s, double* __restrict__ d, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (double)s[i];
}

int num_cus() {
  static int n = 0;
  if (!n) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) n = prop.multiProcessorCount;
    if (n <= 0) n = 256;
  }
  return n;
}

template <typename T, int D, int P>
void launch_lds(StageArgs& a, hipStream_t s) {
  constexpr int TE = kBlock / Sz<D, P>::NP;
  static int occ = 0;
  if (!occ) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stage_lds<T, D, P>, kBlock, 0) != hipSuccess || occ < 1)
      occ = 1;
  }
  a.tiles = (a.K + TE - 1) / TE;
  a.tiles_per_xcd = (int)((a.tiles + 7) / 8);
  int64_t per_xcd_cap = (int64_t)(num_cus() / 8) * occ;
  int bpx = (int)std::min<int64_t>(a.tiles_per_xcd, per_xcd_cap);
  if (bpx < 1) bpx = 1;
  hipLaunchKernelGGL((k_stage_lds<T, D, P>), dim3(8 * bpx), dim3(kBlock), 0, s, a);
}

template <typename T>
void dispatch(Ctx& c, StageArgs& a, hipStream_t s) {
#define DGK_CASE(D, P) \
  if (c.dim == D && c.p == P) return launch_lds<T, D, P>(a, s);
  DGK_CASE(1, 1) DGK_CASE(1, 2) DGK_CASE(1, 3) DGK_CASE(1, 4) DGK_CASE(1, 5)
  DGK_CASE(2, 1) DGK_CASE(2, 2) DGK_CASE(2, 3) DGK_CASE(2, 4) DGK_CASE(2, 5)
  DGK_CASE(3, 1) DGK_CASE(3, 2) DGK_CASE(3, 3) DGK_CASE(3, 4) DGK_CASE(3, 5)
#undef DGK_CASE
  throw Unsupported("no stage kernel for this element type");
}

}  // namespace

void launch_stage(Ctx& c, StageArgs& a, hipStream_t s) {
  if (c.precision == DGK_FP64) dispatch<double>(c, a, s);
  else dispatch<float>(c, a, s);
  DGK_HIP(hipGetLastError());
}

void launch_source(Ctx& c, const Ctx::Src& src, double value, hipStream_t st) {
  if (!src.n) return;
  unsigned nb = (unsigned)((src.n + 255) / 256);
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_source<double>, dim3(nb), dim3(256), 0, st, (double*)c.U, src.idx, src.n, value);
  else
    hipLaunchKernelGGL(k_source<float>, dim3(nb), dim3(256), 0, st, (float*)c.U, src.idx, src.n, (float)value);
  DGK_HIP(hipGetLastError());
}

bool check_finite(Ctx& c) {
  DGK_HIP(hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
  int64_t n = 4 * c.K * c.Np;
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_finite<double>, dim3(1024), dim3(256), 0, c.stream, (const double*)c.U, n, c.flag);
  else
    hipLaunchKernelGGL(k_finite<float>, dim3(1024), dim3(256), 0, c.stream, (const float*)c.U, n, c.flag);
  DGK_HIP(hipGetLastError());
  int h = 0;
  DGK_HIP(hipMemcpyAsync(&h, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  DGK_HIP(hipStreamSynchronize(c.stream));
  return h == 0;
}

void convert_state(Ctx& c, const double* src, void* dst, int64_t n) {
  if (c.precision == DGK_FP64) {
    DGK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return;
  }
  hipLaunchKernelGGL(k_d2f, dim3(1024), dim3(256), 0, c.stream, src, (float*)dst, n);
  DGK_HIP(hipGetLastError());
}

void convert_state_back(Ctx& c, const void* src, double* dst, int64_t n) {
  if (c.precision == DGK_FP64) {
    DGK_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return;
  }
  hipLaunchKernelGGL(k_f2d, dim3(1024), dim3(256), 0, c.stream, (const float*)src, dst, n);
  DGK_HIP(hipGetLastError());
}

const char* variant_name(int v) {
  switch (v) {
    case 0: return "lds-tile";
    default: return "unknown";
  }
}

}  // namespace dgk
