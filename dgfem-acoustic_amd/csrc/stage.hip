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
// Data layout in HBM is the reference's state layout u[eq][element][node] for the
// thread-per-node kernel and the node-interleaved u[element][node][eq] (live fields only) for the
// MFMA kernel; a workgroup
// stages a tile of TE consecutive elements in LDS (coalesced loads of TE*Np contiguous
// values per equation, or of one contiguous block), gathers neighbour traces through the
// device-built face table, and walks tiles in an XCD-aware order (each XCD sweeps one
// contiguous slab of the mesh, so neighbour traces are reused from that XCD's L2) or
// round-robin.
#include <algorithm>
#include <type_traits>

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
  constexpr int NP = S::NP, NF = S::NF, NFP = S::NFP;
  constexpr int TE = kBlock / NP;
  constexpr int NFF = NF * NFP;
  __shared__ T sU[4][TE * NP];
  __shared__ T sF[4][D][TE * NP];
  __shared__ T sFl[4][TE * NFF];
  __shared__ T sG[TE][9];
  __shared__ T sN[TE][NF][4];

  const int tid = threadIdx.x;
  const int64_t K = a.K, KN = a.PS;  // KN: plane stride
  const T* __restrict__ Vin = (const T*)a.Vin;
  T* __restrict__ Vout = (T*)a.Vout;
  T* __restrict__ U = (T*)a.U;
  T* __restrict__ Acc = (T*)a.A;
  const T* __restrict__ Jinv = (const T*)a.Jinv;
  const T* __restrict__ W = (const T*)a.W;
  const T* __restrict__ LIFT = (const T*)a.LIFT;
  const T dt = (T)a.dt, c0 = (T)a.c0;
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


// ---------------------------------------------------------------------------------------------
// MFMA variant. One workgroup (4 waves) owns a tile of 16 elements; each wave owns 4 of them and
// computes R (Np x 16 columns = 4 elements x 4 fields) = Wcat (Np x KD) * X (KD x 16) with
// 16x16x4 MFMAs, where Wcat = [W_1 .. W_d | -LIFT] and X stacks the contravariant nodal fluxes
// and the scaled nodal face fluxes. The K dimension is permuted so that lane group g (lanes
// 16g..16g+15) always supplies nodes n = 4*m + g: at volume k-step s = m*d + r the lane builds
// Ft_r at node 4m+g from the LDS state tile with per-lane coefficients, at face k-step s the lane
// reads face entry 4*(s - d*NG) + g of its column from LDS. The matching A fragments are laid out
// on the host as [m-block][k-step][lane] so that every fragment is one coalesced 64-lane load.
// MFMA register layouts: A lane l <- A[l%16][l/16], B lane l <- B[l/16][l%16]; the
// accumulator row of register j in lane l is given by d_row<T> (column l%16).
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// bf16 16x16x32 MFMA into an FP32 accumulator; a no-op for other accumulator types (never called)
template <typename VV>
__device__ __forceinline__ VV mma_bf16(bf16x8, bf16x8, VV c) {
  return c;
}
template <>
__device__ __forceinline__ Mfma<float>::V mma_bf16(bf16x8 a, bf16x8 b, Mfma<float>::V c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// MFMA for a 4-row piece of the operator (layouts from in-bounds probes on gfx950):
// FP32 v_mfma_f32_4x4x1_16b_f32, 16 blocks of 4x4 with K = 1, 8 cycles: A lane 4b + i holds row i
//   of block b, B lane 4b + j column j, register i of lane 4b + j is D_b[i][j]. Block b = 4*kc + cg
//   takes the k-step's lane group kc and columns 4cg..4cg+3, which is exactly the B operand the lane
//   holds for the 16x16x4 MFMA; each lane group accumulates its own k residue, summed at the end.
// FP64 v_mfma_f64_4x4x4_4b_f64, 4 blocks of 4x4 with K = 4, 16 cycles: A lane l holds row l%4,
//   k = l/16 of block (l%16)/4; B lane l column l%4, k = l/16 of block (l%16)/4; lane L holds
//   D[row L/16][column L%4] of block (L%16)/4. Blocks = column groups, all 4 k of the k-step.
// The A fragment is the same for both: lane l holds row l%4 of the piece at lane group l/16.
template <typename T>
struct Tail;
template <>
struct Tail<float> {
  typedef Mfma<float>::V A;
  static __device__ __forceinline__ A mma(float a, float b, A c) {
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Tail<double> {
  typedef double A;
  static __device__ __forceinline__ A mma(double a, double b, A c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
  }
};

template <int D, int P>
struct MfmaShape {
  using S = Sz<D, P>;
  static constexpr int NP = S::NP, NFF = S::NF * S::NFP;
  static constexpr int NG = (NP + 3) / 4;     // node groups
  static constexpr int NFK = (NFF + 3) / 4;   // face k-steps
  static constexpr int KS = D * NG + NFK;     // k-steps
  static constexpr int MB = (NP + 15) / 16;   // 16-row blocks
  static constexpr int RB = NFF > NP ? NFF : NP;
  static constexpr int NLO = (KS + 7) / 8;    // bf16 low-order correction MFMAs per m-block
  // The last 16-row block holds only RT = Np - 16(MB-1) operator rows. Without the FP32 operator
  // correction, when they fit one 4-row piece (3D orders 1, 3, 4; 2D order 1; 1D up to order 3),
  // that piece runs on a 4x4 MFMA (Tail: FP32 8 cycles, FP64 16) instead of a 16x16x4 one (FP32 32
  // cycles, FP64 64) per k-step. Two or three pieces measured slower (one more operator read from
  // LDS per k-step, and at 3D order 5 register spills).
  static constexpr int RT = NP - 16 * (MB - 1), NT4 = (RT + 3) / 4;
  static constexpr int tail(bool /*fp32*/, bool lo) { return !lo && NT4 == 1 ? 1 : 0; }
};

// Output mapping of the 16x16x4 accumulator: row of register j in lane l (column is l % 16).
// v_mfma_f32_16x16x4_f32 puts rows 4*(l/16)..4*(l/16)+3 in one lane; v_mfma_f64_16x16x4_f64
// interleaves them as l/16 + 4j (both established with an in-bounds probe on gfx950).
template <typename T>
__device__ __forceinline__ int d_row(int lane, int j);
template <>
__device__ __forceinline__ int d_row<float>(int lane, int j) { return 4 * (lane >> 4) + j; }
template <>
__device__ __forceinline__ int d_row<double>(int lane, int j) { return (lane >> 4) + 4 * j; }

// LO: FP32 only. The operator is split as A = A_hi + A_lo with A_hi = fl32(A) and A_lo the
// residual rounded to bf16; after every 8 k-steps one 16x16x32 bf16 MFMA adds A_lo times the same
// 8 B values (packed to bf16), so the systematic rounding of the operator drops from 2^-24 to
// about 2^-33 relative at ~1/16 extra MFMA issue.
// CS: FP32 only, compensated solution. U holds fl32(u) and Ulo a bf16 low word with u = U + Ulo
// to about 2^-33 relative. The accumulator carries only the increment (stage 0 writes dt/6 k1
// instead of u + dt/6 k1), and stage 3 adds it to (U, Ulo) with an exact two-sum, so the
// per-step rounding of u (2^-24, accumulating over the run) is removed; stage inputs u + c k
// are still formed from U alone, which costs only a non-accumulating 2^-24 perturbation.
// Element access through a uniform base pointer plus a 32-bit byte offset, which maps to the
// global_load/store form with an SGPR base and a 32-bit VGPR offset (no 64-bit address arithmetic
// per access). Contexts whose buffers stay below 4 GiB use the buffers' own bases; larger ones
// (the BIG instantiation of the stage kernel) form a 64-bit base per tile on the scalar unit and
// the neighbour element's base per gather, with 32-bit offsets inside those.
template <typename P>
__device__ __forceinline__ P* byte_at(P* base, uint32_t off) {
  return (P*)((char*)base + off);
}
template <typename P>
__device__ __forceinline__ const P* byte_at(const P* base, uint32_t off) {
  return (const P*)((const char*)base + off);
}

__device__ __forceinline__ float bf16_bits_to_f32(uint32_t b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ uint32_t f32_to_bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;  // round to nearest even (finite inputs)
}

// Workgroup shape: 4 waves (16-element tiles) by default. 3D orders 4 and 5 run 8 waves (32-element
// tiles) so that two waves per SIMD share one operator copy in LDS: FP32 order 5 (64.5 KB
// operator) and FP64 order 4 (48.4 KB operator, 50 KB of tile buffers per 16 elements) have room
// for one workgroup per CU; FP32 order 4 (24 KB operator) is 2-4% faster than with two 4-wave
// workgroups.
// FP64 order 5 (129 KB operator) splits rows instead of elements: a 4-wave workgroup owns a tile of
// 4 elements (21 KB of tile buffers beside the operator in LDS) and wave w computes row block w of
// all 16 columns, so the operator never leaves LDS; every wave builds the same B operand.
// Tile shapes of the stage kernel (dgk_set_tile_shape): 0 the built-in choice per (precision,
// dim, order) below; 1 4-wave workgroups over 16-element tiles; 2 8-wave workgroups over
// 32-element tiles; 3 the row split, 4 waves over a 4-element tile with wave w computing a quarter
// of the operator's 16-row blocks (the whole operator in LDS); 4 4-wave workgroups over
// 16-element tiles with the operator fragments read from global memory; 5 the row split over
// 8 waves and an 8-element tile (two element groups of 4, two waves per SIMD); 6, 7, 8 4-wave
// workgroups over 16-element tiles compiled for 3, 4 or 5 waves per SIMD (at most 512/n registers
// per lane), so that as many workgroups share a CU where LDS allows; 9, 10, 11 as 7, 8, 6 with the
// RK epilogue operands loaded after the volume k-steps (fewer registers live across the chain).
// Built-in: 8-wave workgroups for 3D order 5 and FP64 3D order 4 (two waves per SIMD share one LDS
// copy of the operator); for FP64 3D order 5 the 8-wave row split (shape 5: its operator, 110 KB
// with the compact last row block, fits LDS only beside a small tile); shape 6 for FP32 3D order 4
// (three 52 KB workgroups per CU); 4-wave workgroups elsewhere.
template <typename T, int D, int P, int SH = 0>
struct MfmaLaunch {
  static constexpr bool DEF_ROWSPLIT = sizeof(T) == 8 && D == 3 && P == 5;
  static constexpr int DEF_WAVES = (D == 3 && (P == 5 || (P == 4 && sizeof(T) == 8))) ? 8 : 4;
  static constexpr int DEF_MINW = (sizeof(T) == 4 && D == 3 && P == 4) ? 3 : 1;
  static constexpr bool ROWSPLIT = SH == 3 || SH == 5 || (SH == 0 && DEF_ROWSPLIT);
  static constexpr int WAVES = SH == 0 ? DEF_WAVES : (SH == 2 || SH == 5) ? 8 : 4;
  static constexpr int MINW = SH == 0 ? DEF_MINW : (SH >= 6 && SH <= 8) ? SH - 3 : (SH == 9 || SH == 10) ? SH - 5 : SH == 11 ? 3 : 1;
  static constexpr bool DEFER = SH == 9 || SH == 10 || SH == 11;  // epilogue operands loaded after the volume k-steps  // waves per SIMD the registers are sized for
  static constexpr bool OPG = SH == 4;                // operator fragments from global memory
  static constexpr int KB = 64 * WAVES;              // threads per workgroup
  static constexpr int TE = ROWSPLIT ? WAVES : 4 * WAVES;  // elements per tile (row split: 4 row groups)
  static constexpr int RG = WAVES / (TE / 4);        // row groups: waves sharing one element group
};

// Pad (in values) of the LDS element stride NP*ES. A wave's 4 elements read and write node rows at
// the same offset, 16 lanes (one lane group) or 32 lanes per LDS pass: FP32 strides of 4 or 12 mod
// 16 words, FP64 strides of 16 or 48 mod 64 words (8 mod 16 doubles) put them in disjoint banks.
// None for the 2- and 3-field layouts.
template <typename T, int ES>
constexpr int lds_pad(int np) {
  if (ES != 4) return 0;
  if (sizeof(T) == 4) return np % 4 == 0 || np % 4 == 2 ? 4 : 0;
  return ((8 - 4 * np) % 16 + 16) % 16;
}

// LDS row stride of the face fluxes [element][field][row]: a face k-step's 64 lanes read value
// c*RB + g + const (column c < 16, lane group g < 4). 4-byte values: a stride of 4 x odd puts the
// 64 lanes in 64 distinct banks; 8-byte values: 2 x odd puts the 16 columns 4 words apart, so each
// pass of one or two lane groups is conflict-free.
template <typename T>
constexpr int rb_stride(int rb) {
  const int m = sizeof(T) == 4 ? 8 : 4, r = sizeof(T) == 4 ? 4 : 2;
  return rb + ((r - rb % m) % m + m) % m;
}

// ES consecutive field values of one node (ES = 4: one vector read), zero beyond ES
template <int ES, typename T>
__device__ __forceinline__ void node_fields(const T* p, T f[4]) {
  if constexpr (ES == 4) {
    typedef T V4 __attribute__((ext_vector_type(4)));
    const V4 w = *(const V4*)p;
#pragma unroll
    for (int c = 0; c < 4; ++c) f[c] = w[c];
  } else {
#pragma unroll
    for (int c = 0; c < 4; ++c) f[c] = c < ES ? p[c] : (T)0;
  }
}

// Footprint of one stage-kernel instantiation: tile, operator and LDS sizes (shared by the kernel
// and the host, which builds only the tile shapes that fit).
template <typename T, int D, int P, bool LO, int ES, int SH>
struct MfmaCfg {
  using S = Sz<D, P>;
  using M = MfmaShape<D, P>;
  using L = MfmaLaunch<T, D, P, SH>;
  static constexpr int NP = S::NP, NF = S::NF, NFP = S::NFP, NPERM = S::NPERM;
  static constexpr int NT4 = M::tail(sizeof(T) == 4, LO);  // 4-row tail pieces (0: all blocks 16x16x4)
  static constexpr int MBF = NT4 ? M::MB - 1 : M::MB;      // 16x16x4 row blocks
  // A fragments per k-step: MBF 16x16x4 blocks of 64 lane values, and the tail piece of 16
  // distinct values (row l%4 at lane group l/16; lanes 4 apart in l/4 read the same value)
  static constexpr int NAV = (MBF * 64 + NT4 * 16) * M::KS;  // operator values
  static constexpr int TE = L::TE, RG = L::RG;
  // Row split with one 16-row block per wave: when the last block holds at most 8 operator rows
  // (3D order 5: rows 48-55), its fragments are kept in LDS without the zero rows, 32 values per
  // k-step; lanes l of that wave read row l%8 (rows 56-63 of its result repeat 48-55 and are never
  // written out). NAVS: the operator values held in LDS.
  static constexpr int RT = S::NP - 16 * (M::MB - 1);
  static constexpr bool kCompact = RG > 1 && RG == MBF && NT4 == 0 && RT <= 8;
  static constexpr int NAVS = kCompact ? ((MBF - 1) * 64 + 32) * M::KS : NAV;
  // LDS element stride LES of the state tile and of the results: NP*ES values plus a pad (4-field
  // layouts only, whose element boundaries fall on whole 16-byte vectors) chosen so that the four
  // elements of a wave fall in different LDS banks. A volume k-step's lanes read node 4m+g of their
  // element (16 or 32 bytes) and the result write-out puts row 4g+j (FP32) or g+4j (FP64) of 4
  // elements x 4 fields: FP32 wants the stride in 4-byte words = 4 mod 16 (or 12), FP64 32 mod 64.
  static constexpr int PADV = lds_pad<T, ES>(NP);
  static constexpr int LES = NP * ES + PADV;
  // face fluxes [element][field][RBP], then the results in the state tile's layout
  static constexpr int RBP = rb_stride<T>(M::RB);
  static constexpr int NSB = 4 * TE * RBP > TE * LES ? 4 * TE * RBP : TE * LES;
  static constexpr size_t kLdsRest = sizeof(T) * (TE * LES + NSB + TE * NF * 4 + TE * 9) +
                                     sizeof(int) * (NF * NPERM * NFP + NF * NFP) + (LO ? 16 * M::MB * M::NLO * 64 : 0);
  static constexpr size_t kLdsMax = 160 * 1024 - 512;
  // Operator fragments are staged in LDS whenever the whole workgroup footprint fits the 160 KiB
  // of a gfx950 CU (fewer resident workgroups, but no per-k-step global loads).
  static constexpr bool kALds = !L::OPG && NAVS * sizeof(T) + kLdsRest <= kLdsMax;
  // the row split needs whole 16-row blocks per wave and the operator in LDS; the 8-wave tile shape
  // exists to share one LDS copy of the operator between two waves per SIMD (the built-in shape of
  // the FP32 3D order-5 operator correction runs 8 waves with the operator in global memory)
  // a register budget for MINW waves per SIMD needs that many workgroups' LDS on one CU
  static constexpr bool fits = kLdsRest <= kLdsMax && (RG == 1 || (NT4 == 0 && !LO && MBF % RG == 0 && kALds)) &&
                               (SH == 0 || L::WAVES == 4 || kALds) &&
                               (SH == 0 || L::MINW * 4 / L::WAVES * ((kALds ? NAVS * sizeof(T) : 0) + kLdsRest) <=
                                               160 * 1024);
};

template <typename T, int D, int P, bool LO, bool CS, int ES, bool BIG, int SH>
__global__ __launch_bounds__((MfmaLaunch<T, D, P, SH>::KB), (MfmaLaunch<T, D, P, SH>::MINW)) void k_stage_mfma(
    StageArgs a) {
  static_assert(!CS || sizeof(T) == 4, "the compensated state is an FP32 form");
  using S = Sz<D, P>;
  using M = MfmaShape<D, P>;
  using C = MfmaCfg<T, D, P, LO, ES, SH>;
  static_assert(C::fits, "tile shape does not fit this element type");
  constexpr int kBlock = MfmaLaunch<T, D, P, SH>::KB;
  constexpr int NP = S::NP, NF = S::NF, NFP = S::NFP, NPERM = S::NPERM, NFF = M::NFF;
  constexpr int NG = M::NG, NFK = M::NFK, KS = M::KS, MB = M::MB, RB = C::RBP, NLO = M::NLO;
  constexpr int NT4 = C::NT4, MBF = C::MBF, NAV = C::NAV, TE = C::TE;
  // row groups: RG waves share one group of 4 elements, wave rg computing row blocks
  // [rg*MBW, (rg+1)*MBW); with a single row block per wave the k-steps alternate between two
  // partial accumulators (summed before the write-out) so that consecutive MFMAs are independent
  constexpr int RG = C::RG;
  constexpr int MBW = MBF / RG;
  constexpr int NCH = (RG > 1 && MBW == 1) ? 2 : 1;
  constexpr int VW = 16 / sizeof(T);                        // values per 16-byte vector
  typedef T VT __attribute__((ext_vector_type(VW)));
  constexpr int PV = TE * NP / VW;                          // vectors per field of a tile (ES*PV in all)
  // The state is node-interleaved, [element][node][field] with the ES live fields (nq), so a tile
  // is one contiguous block of ES*PV vectors, a neighbour trace one load of ES values and a node's
  // fields one LDS read; the LDS tiles use the same layout.
  static_assert(ES == D + 1 || ES == 4, "ES: the live field count");
  constexpr int NE = (4 * PV + kBlock - 1) / kBlock;        // state vectors per thread
  constexpr int NI = (TE * NFF + kBlock - 1) / kBlock;      // face items per thread
  constexpr int NGEO = (TE * 9 + kBlock - 1) / kBlock;      // geometry values per thread
  constexpr int NFC = (TE * NF + kBlock - 1) / kBlock;      // element faces per thread
  typedef typename Mfma<T>::V V;
  constexpr int LES = C::LES;
  __shared__ __attribute__((aligned(32))) T sU[TE * LES];
  // face fluxes [element][field][RB] (RB: the padded row stride rb_stride), then the results in
  // sU's layout
  constexpr int NSB = C::NSB;
  __shared__ __attribute__((aligned(16))) T sB[NSB];
  __shared__ T sN[TE * NF][4];
  __shared__ T sG[TE][9];
  __shared__ int sVmap[NF * NPERM * NFP];
  __shared__ int sFmask[NF * NFP];
  constexpr bool kALds = C::kALds;
  constexpr bool kCompact = C::kCompact;
  __shared__ T sA[kALds ? C::NAVS : 1];
  __shared__ bf16x8 sAlo[LO ? MB * NLO * 64 : 1];
  T* const sR = sB;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // wave = rg * (TE/4) + element group; lane group g, column col = 4 * (element in group) + field q
  const int rg = RG > 1 ? wave / (TE / 4) : 0;
  const int g = lane >> 4, col = lane & 15, el = (wave % (TE / 4)) * 4 + (col >> 2), q = col & 3;
  // 32-bit byte offsets: a buffer holds 4*PS*sizeof(T) < 2^32 bytes (checked at context creation)
  const T* __restrict__ Vin = (const T*)a.Vin;
  T* __restrict__ Vout = (T*)a.Vout;
  T* __restrict__ U = (T*)a.U;
  T* __restrict__ Acc = (T*)a.A;
  uint16_t* __restrict__ Ulo = (uint16_t*)a.Ulo;
  const T* __restrict__ Jinv = (const T*)a.Jinv;
  const T* __restrict__ Fg = (const T*)a.fgeom;
  const T* __restrict__ Af = (const T*)a.Amf;
  const T dt = (T)a.dt, c0 = (T)a.c0;
  const T rc2 = (T)(a.rho0 * a.c0 * a.c0), irho = (T)(1.0 / a.rho0);
  const T v0[3] = {(T)a.v0[0], (T)a.v0[1], (T)a.v0[2]};
  const int stage = a.stage;
  constexpr int nq = ES;  // the buffers hold exactly the live fields
  for (int i = tid; i < NF * NPERM * NFP; i += kBlock) sVmap[i] = a.vmap[i];
  for (int i = tid; i < NF * NFP; i += kBlock) sFmask[i] = a.fmask[i];
  if (kALds) {
    constexpr int NFULL = kCompact ? (MBF - 1) * KS * 64 : NAV;
    for (int i = tid; i < NFULL; i += kBlock) sA[i] = Af[i];
    if constexpr (kCompact)  // the last block without its zero rows: value c of a k-step is row c%8, k c/8
      for (int i = tid; i < KS * 32; i += kBlock)
        sA[NFULL + i] = Af[NFULL + (i / 32) * 64 + ((i % 32) / 8) * 16 + (i % 8)];
  }
  static_assert(RG == 1 || kALds, "row split keeps the operator in LDS");
  // the wave of the compact last block (wave-uniform)
  const bool clast = kCompact && rg == MBF - 1;
  const T* __restrict__ pA = clast ? (const T*)sA + (MBF - 1) * KS * 64 + (lane >> 4) * 8 + (lane & 7)
                                   : (kALds ? (const T*)sA : Af) + lane + rg * (MBW * KS * 64);
  const T* __restrict__ pT = (kALds ? (const T*)sA : Af) + MBF * KS * 64 + 4 * (lane >> 4) + (lane & 3);
  if (LO)
    for (int i = tid; i < MB * NLO * 64; i += kBlock) sAlo[i] = ((const bf16x8*)a.Alo)[i];

  // Per-thread offsets that do not depend on the tile (every tile is full: the element count is
  // padded to a multiple of TE and padding elements carry zero state and geometry): the thread's
  // state vectors sit at v*VW values into the tile, in a state buffer (add e0*NP*ES) and in LDS.
  uint32_t voff[NE];
  int loff[NE];
#pragma unroll
  for (int t = 0; t < NE; ++t) {
    const int v = tid + t * kBlock;
    voff[t] = (uint32_t)(v * VW);
    loff[t] = v * VW + (v * VW / (NP * ES)) * C::PADV;
  }

  // Software pipeline: while tile t is processed, tile t+1's state, geometry and face geometry,
  // and the (neighbour, code) pairs of the thread's face items are in flight into registers; the
  // neighbour traces of tile t+1 are gathered right after tile t's flux phase.
  VT pst[NE];
  T pgeo[NGEO], pfg[NFC][4];
  int2 inc[NI], cnc[NI];
  T iqp[NI][4];
  // value `rel` of an array of `per` values per element, at the tile starting at element e0 (32-bit
  // arithmetic unless BIG)
  auto tile_at = [&](auto* base, int64_t e0, uint32_t per, uint32_t rel) {
    using E = std::remove_cv_t<std::remove_pointer_t<decltype(base)>>;
    if constexpr (BIG) return byte_at(base + e0 * per, rel * (uint32_t)sizeof(E));
    else return byte_at(base, ((uint32_t)e0 * per + rel) * (uint32_t)sizeof(E));
  };
  auto prefetch = [&](int64_t e0) {
#pragma unroll
    for (int t = 0; t < NE; ++t)
      if (tid + t * kBlock < nq * PV) pst[t] = *(const VT*)tile_at(Vin, e0, NP * ES, voff[t]);
#pragma unroll
    for (int t = 0; t < NGEO; ++t) {
      const int idx = tid + t * kBlock;
      if (idx < TE * 9) pgeo[t] = *tile_at(Jinv, e0, 9, (uint32_t)idx);
    }
#pragma unroll
    for (int t = 0; t < NFC; ++t) {
      const int idx = tid + t * kBlock;
      if (idx < TE * NF)
#pragma unroll
        for (int c = 0; c < 4; ++c) pfg[t][c] = *tile_at(Fg, e0, NF * 4, (uint32_t)(idx * 4 + c));
    }
#pragma unroll
    for (int t = 0; t < NI; ++t) {
      const int idx = tid + t * kBlock;
      const int ej = (idx / NFF) * NF + (idx % NFF) / NFP;
      inc[t] = idx < TE * NFF ? make_int2(*tile_at(a.nbr, e0, NF, (uint32_t)ej), *tile_at(a.code, e0, NF, (uint32_t)ej))
                              : make_int2(-1, 0);
    }
  };
  auto gather = [&]() {
#pragma unroll
    for (int t = 0; t < NI; ++t) {
      const int k = ((tid + t * kBlock) % NFF) % NFP;
      if (inc[t].x >= 0) {
        const uint32_t node = (uint32_t)sVmap[(inc[t].y & 0xffff) * NFP + k];
        if constexpr (BIG) {
          // the neighbour element's base in 64 bits, the node within it as a 32-bit offset
          const T* nb = Vin + (uint64_t)(uint32_t)inc[t].x * (uint32_t)(NP * ES);
          node_fields<ES>(byte_at(nb, node * (uint32_t)(ES * sizeof(T))), iqp[t]);
        } else {
          const uint32_t pi = (uint32_t)inc[t].x * (uint32_t)NP + node;
          node_fields<ES>(byte_at(Vin, pi * (uint32_t)(ES * sizeof(T))), iqp[t]);
        }
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) iqp[t][c] = (T)0;
      }
    }
  };

  // Tile walk: XCD-aware (workgroup b runs on XCD b%8; XCD x sweeps tiles [x*tpx, (x+1)*tpx), so
  // neighbour traces are reused from that XCD's L2) or round-robin over the whole grid.
  const bool rr = a.tile_order == 1;
  const int xcd = rr ? 0 : blockIdx.x % 8, lb = rr ? blockIdx.x : blockIdx.x / 8;
  const int bpx = rr ? gridDim.x : gridDim.x / 8;
  const int tpx = rr ? (int)a.tiles : a.tiles_per_xcd, ntiles = (int)a.tiles;
  auto valid = [&](int it) { return it < tpx && xcd * tpx + it < ntiles; };
  __syncthreads();  // sVmap ready for the first gathers
  if (valid(lb)) {
    prefetch((int64_t)(xcd * tpx + lb) * TE);
    gather();
  }
  for (int it = lb; valid(it); it += bpx) {
    const int64_t e0 = (int64_t)(xcd * tpx + it) * TE;
    __syncthreads();  // the previous tile is done with LDS
    // phase 1: commit the prefetched tile to LDS; epilogue operands and the next tile in flight
#pragma unroll
    for (int t = 0; t < NE; ++t)
      if (tid + t * kBlock < nq * PV) *(VT*)(sU + loff[t]) = pst[t];
#pragma unroll
    for (int t = 0; t < NGEO; ++t) {
      const int idx = tid + t * kBlock;
      if (idx < TE * 9) sG[idx / 9][idx % 9] = pgeo[t];
    }
#pragma unroll
    for (int t = 0; t < NFC; ++t) {
      const int idx = tid + t * kBlock;
      if (idx < TE * NF)
#pragma unroll
        for (int c = 0; c < 4; ++c) sN[idx][c] = pfg[t][c];
    }
    // epilogue operands: issued here, a tile ahead of their use, when registers allow; shapes with
    // the operator in global memory load them just before the epilogue instead
    VT pu[NE], pa[NE];
    uint2 plo[CS ? NE : 1];
    // u_n and the accumulator are touched once per stage per element, and on large meshes more
    // than the caches hold passes before their next use: their loads and stores carry the
    // non-temporal hint, which leaves the caches to the neighbour traces and stage inputs (1-3% per
    // stage on the HBM-bound shapes)
    auto load_epi = [&]() {
    if (stage >= 1 && stage <= 3) {
#pragma unroll
      for (int t = 0; t < NE; ++t)
        if (tid + t * kBlock < nq * PV) {
          if (stage != 3 || CS) pu[t] = __builtin_nontemporal_load((const VT*)tile_at(U, e0, NP * ES, voff[t]));
          if (CS && stage == 3) plo[CS ? t : 0] = *(const uint2*)tile_at(Ulo, e0, NP * ES, voff[t]);
          pa[t] = __builtin_nontemporal_load((const VT*)tile_at(Acc, e0, NP * ES, voff[t]));
        }
    }
    };
    // wide workgroups and operators read from global memory need the registers during the MFMA
    // chain: their epilogue operands are loaded later (below)
    constexpr bool kDefer = !kALds || MfmaLaunch<T, D, P, SH>::WAVES > 4 || MfmaLaunch<T, D, P, SH>::DEFER;
    if (!kDefer) load_epi();
#pragma unroll
    for (int t = 0; t < NI; ++t) cnc[t] = inc[t];
    const bool more = valid(it + bpx);
    if (more) prefetch((int64_t)(xcd * tpx + it + bpx) * TE);
    __syncthreads();
    // phase 2: nodal normal numerical fluxes scaled by sJ/|detJ| -> sB[(e*4 + q)*RB + jk]. All three face
    // kinds share one branch-free form (no divergence between interior and boundary face nodes):
    //   fl = A (F(qm) + F(qp)).n-with-speed-vne + H (qm - qp),   F(q).n = (vne p + rc2 v.n, vne v + p/rho n)
    // interior    qp = neighbour trace, A = 1/2, vne = v0.n, H = s c0/2      (Rusanov)
    // reflecting  qp = (p, v - 2 (v.n) n), A = 1/2, vne = v0.n, H = 0         (= n.F(p, v - (v.n) n))
    // absorbing   qp = (0, v - (v.n) n),   A = 1/4, vne = 0,    H = c0/4
    //             (= 1/4 [c0 p + rc2 v.n ; (p/rho + c0 v.n) n]; the qp part of F vanishes up to the
    //             rounding of v.n - (v.n)(n.n))
    // the face nodes' LDS addresses are recomputed each tile (a few integer operations) rather than
    // held across the tile loop: under register budgets of three or more waves per SIMD the
    // compiler spilled them to scratch and reloaded them every tile, and in 1D/2D the registers
    // saved raise the waves per SIMD (2D order 2 FP32: 5 -> 6, 10% faster); the 3D kernels without
    // a budget keep them (the extra integer work slows the issue-bound order-5 kernel by 1%)
    int ftid = tid;
    if constexpr (MfmaLaunch<T, D, P, SH>::MINW >= 3 || D < 3) asm volatile("" : "+v"(ftid));
#pragma unroll
    for (int t = 0; t < NI; ++t) {
      const int idx = ftid + t * kBlock;
      if ((t + 1) * kBlock > TE * NFF && idx >= TE * NFF) continue;  // only the last round is partial
      const int e = idx / NFF, jk = idx - e * NFF, j = jk / NFP;
      const int ej = e * NF + j;
      const int2 nc = cnc[t];
      const T n[3] = {sN[ej][0], sN[ej][1], sN[ej][2]};
      const T fs = sN[ej][3];
      T qm[4];
      node_fields<ES>(sU + e * LES + sFmask[jk] * ES, qm);
      const T vn = v0[0] * n[0] + v0[1] * n[1] + v0[2] * n[2];
      const T unm = qm[1] * n[0] + qm[2] * n[1] + qm[3] * n[2];
      const bool inter = nc.x >= 0, refl = nc.x == -2;
      const T gm = refl ? (T)2 : (T)1;
      const T qp0 = inter ? iqp[t][0] : (refl ? qm[0] : (T)0);
      T qpv[3];
      for (int c = 0; c < 3; ++c) qpv[c] = inter ? iqp[t][1 + c] : qm[1 + c] - gm * unm * n[c];
      const T unp = qpv[0] * n[0] + qpv[1] * n[1] + qpv[2] * n[2];
      // the face scale fs is folded into A and H; sums before products:
      //   fs fl_p = fs A (vne (pm + pp) + rc2 (un_m + un_p)) + fs H (pm - pp)
      //   fs fl_v = fs A (vne (vm + vp) + (pm + pp)/rho n) + fs H (vm - vp)
      const T A = fs * ((inter || refl) ? (T)0.5 : (T)0.25);
      const T H = fs * (inter ? (T)0.5 * ((nc.y >> 16) & 1 ? -c0 : c0) : (refl ? (T)0 : (T)0.25 * c0));
      const T Av = (inter || refl) ? A * vn : (T)0;
      const T ps = qm[0] + qp0, pd = qm[0] - qp0;
      const T Ap = A * irho * ps;
      T fl[4];
      fl[0] = Av * ps + (A * rc2) * (unm + unp) + H * pd;
      for (int c = 0; c < 3; ++c) fl[1 + c] = Av * (qm[1 + c] + qpv[c]) + Ap * n[c] + H * (qm[1 + c] - qpv[c]);
      for (int qq = 0; qq < 4; ++qq) sB[(e * 4 + qq) * RB + jk] = fl[qq];
    }
    if (more) gather();  // next tile's neighbour traces, in flight during the MFMA phase
    __syncthreads();
    // phase 3: per-lane coefficients Ft_r = alpha_r p + sum_x beta_rx v_x, then the MFMA chain
    T alpha[D], beta[D][3];
    {
      const T* G = sG[el];
      for (int r = 0; r < D; ++r) {
        const T at = G[r * 3 + 0] * v0[0] + G[r * 3 + 1] * v0[1] + G[r * 3 + 2] * v0[2];
        if (q == 0) {
          alpha[r] = at;
          for (int x = 0; x < 3; ++x) beta[r][x] = rc2 * G[r * 3 + x];
        } else {
          alpha[r] = G[r * 3 + (q - 1)] * irho;
          for (int x = 0; x < 3; ++x) beta[r][x] = (x == q - 1) ? at : (T)0;
        }
      }
    }
    V acc[MBW ? MBW * NCH : 1];
    typename Tail<T>::A acct[NT4 ? NT4 : 1];
#pragma unroll
    for (int mb = 0; mb < MBW * NCH; ++mb) acc[mb] = V{0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < NT4; ++t) acct[t] = typename Tail<T>::A{};
    bf16x8 blo;
    // Fragments read from global memory (shapes whose operator does not fit in LDS) are reloaded per
    // tile through an opaque pointer and the node loop is only partly unrolled: otherwise the
    // compiler hoists every fragment of the operator into registers and spills.
    const T* pAt = pA;
    const T* pTt = pT;
    if (!kALds) asm volatile("" : "+v"(pAt), "+v"(pTt));
    // the MFMA chain, with the A fragments AS values apart per k-step (64; 32 in the compact last
    // block of the row split)
    auto chain = [&](auto astride) {
      constexpr int AS = decltype(astride)::value;
      auto step = [&](const int s, const T b) {
#pragma unroll
        for (int mb = 0; mb < MBW; ++mb) {
          const int ai = NCH > 1 ? mb * NCH + (s & 1) : mb;
          acc[ai] = Mfma<T>::mma(pAt[(mb * KS + s) * AS], b, acc[ai]);
        }
#pragma unroll
        for (int t = 0; t < NT4; ++t) acct[t] = Tail<T>::mma(pTt[(t * KS + s) * 16], b, acct[t]);
        if constexpr (LO) {
          blo[s % 8] = (__bf16)b;
          if (s % 8 == 7 || s == KS - 1) {
            if (s % 8 != 7)
#pragma unroll
              for (int i = s % 8 + 1; i < 8; ++i) blo[i] = (__bf16)0.0f;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) acc[mb] = mma_bf16(sAlo[(mb * NLO + s / 8) * 64 + lane], blo, acc[mb]);
          }
        }
      };
      constexpr int kNodeUnroll = kALds ? NG : 2;
#pragma unroll kNodeUnroll
      for (int m = 0; m < NG; ++m) {
        const int n = 4 * m + g;
        T p = 0, vx = 0, vy = 0, vz = 0;
        if (n < NP) {
          T f[4];
          node_fields<ES>(sU + el * LES + n * ES, f);
          p = f[0];
          vx = f[1];
          vy = f[2];
          vz = f[3];
        }
#pragma unroll
        for (int r = 0; r < D; ++r)
          step(m * D + r, alpha[r] * p + beta[r][0] * vx + beta[r][1] * vy + beta[r][2] * vz);
      }
      // 8-wave workgroups: the epilogue operands are issued here, in flight during the face k-steps
      // (the volume coefficients are dead by now, which keeps the peak within 256 registers)
      if (kDefer && kALds) load_epi();
#pragma unroll
      for (int f = 0; f < NFK; ++f) {
        const int jk = 4 * f + g;
        step(D * NG + f, jk < NFF ? sB[(el * 4 + q) * RB + jk] : (T)0);
      }
    };
    if constexpr (kCompact) {
      if (clast) chain(std::integral_constant<int, 32>{});
      else chain(std::integral_constant<int, 64>{});
    } else {
      chain(std::integral_constant<int, 64>{});
    }
    __syncthreads();  // face fluxes consumed; sB now receives the results in the LDS tile's layout
#pragma unroll
    for (int mb = 0; mb < MBW; ++mb) {
      V r = acc[mb * NCH];
      if constexpr (NCH > 1) r += acc[mb * NCH + 1];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = 16 * (rg * MBW + mb) + d_row<T>(lane, j);
        if (i < NP && q < nq) sR[el * LES + i * ES + q] = r[j];
      }
    }
#pragma unroll
    for (int t = 0; t < NT4; ++t) {
      if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          // lane group g holds the partial sum over the k-steps' lane group g: add the four
          T v = acct[t][j];
          v += __shfl_xor(v, 16);
          v += __shfl_xor(v, 32);
          const int i = 16 * MBF + 4 * t + j;
          if (g == 0 && i < NP && q < nq) sR[el * LES + i * ES + q] = v;
        }
      } else {
        const int i = 16 * MBF + 4 * t + g;  // lane group g holds row g of the piece
        if (i < NP && q < nq) sR[el * LES + i * ES + q] = (T)acct[t];
      }
    }
    __syncthreads();
    if (kDefer && !kALds) load_epi();
    // phase 4: time-integration epilogue on 16-byte vectors
    const T s6 = (T)(1.0 / 6.0), s3 = (T)(1.0 / 3.0), s2 = (T)0.5;
#pragma unroll
    for (int t = 0; t < NE; ++t) {
      if (tid + t * kBlock >= nq * PV) break;
      const VT kq = dt * *(const VT*)(sR + loff[t]);
      VT* const pAcc = (VT*)tile_at(Acc, e0, NP * ES, voff[t]);
      VT* const pOut = (VT*)tile_at(Vout, e0, NP * ES, voff[t]);
      VT* const pU = (VT*)tile_at(U, e0, NP * ES, voff[t]);
      uint2* const pLo = (uint2*)tile_at(Ulo, e0, NP * ES, voff[t]);
      switch (stage) {
        case 0: {
          const VT u = *(const VT*)(sU + loff[t]);
          __builtin_nontemporal_store(CS ? kq * s6 : u + kq * s6, pAcc);
          *pOut = u + kq * s2;
        } break;
        case 1: __builtin_nontemporal_store(pa[t] + kq * s3, pAcc); *pOut = pu[t] + kq * s2; break;
        case 2: __builtin_nontemporal_store(pa[t] + kq * s3, pAcc); *pOut = pu[t] + kq; break;
        case 3:
          if constexpr (CS) {
            const VT d = pa[t] + kq * s6;  // dt/6 (k1 + 2 k2 + 2 k3 + k4)
            const uint2 lw = plo[CS ? t : 0];
            const uint32_t lb[4] = {lw.x & 0xffffu, lw.x >> 16, lw.y & 0xffffu, lw.y >> 16};
            VT hi;
            uint32_t ob[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const float u = pu[t][i], x = d[i];
              const float sum = u + x, bv = sum - u;
              const float err = (u - (sum - bv)) + (x - bv);  // two-sum: u + x = sum + err exactly
              const float lo = bf16_bits_to_f32(lb[i]) + err;
              const float h2 = sum + lo;
              hi[i] = h2;
              ob[i] = f32_to_bf16_bits(lo - (h2 - sum));
            }
            __builtin_nontemporal_store(hi, pU);
            *pLo = make_uint2(ob[0] | (ob[1] << 16), ob[2] | (ob[3] << 16));
          } else {
            __builtin_nontemporal_store(pa[t] + kq * s6, pU);
          }
          break;
        case 4: *pOut = kq; break;
        default:
          *pOut = *(const VT*)(sU + loff[t]) + kq;
          if (CS) *pLo = make_uint2(0, 0);  // Vout becomes U: an Euler step drops the low word
          break;
      }
    }
  }
}

// idx holds node offsets e*Np + node; the pressure sits at (e*Np + node)*es, es being the
// interleaved field count (1 for field planes)
template <typename T>
__global__ void k_source(T* __restrict__ U, uint16_t* __restrict__ Ulo, const int64_t* __restrict__ idx, int64_t n,
                         T value, int es) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const int64_t j = idx[i] * es;
    U[j] = value;
    if (Ulo) Ulo[j] = 0;
  }
}

// The same overwrite at the step whose start time is tdev[*tstep], when it lies before the
// source's end: amp*sin(2 pi freq t + phase) (src/solver.cpp:248-256), evaluated in FP64
template <typename T>
__global__ void k_source_timed(T* __restrict__ U, uint16_t* __restrict__ Ulo, const int64_t* __restrict__ idx,
                               int64_t n, double amp, double freq, double phase, double duration,
                               const double* __restrict__ tdev, const int* __restrict__ tstep, int es) {
  const double t = tdev[*tstep];
  if (!(t < duration)) return;
  const T value = (T)(amp * sin(2 * M_PI * freq * t + phase));
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = idx[i] * es;
    U[j] = value;
    if (Ulo) Ulo[j] = 0;
  }
}

__global__ void k_tick(int* __restrict__ tstep) {
  if (threadIdx.x == 0) *tstep += 1;
}

// One workgroup: every thread interpolates receivers r = tid, tid+256, ... from the pressure field
// (value (e*Np + node)*es of the state buffer) and writes row *counter of the record; the row index is read before and advanced after the
// writes, so a captured graph records consecutive rows on every replay. Rows beyond cap are dropped.
template <typename T>
__global__ void k_record(const T* __restrict__ U, const uint16_t* __restrict__ Ulo, const int64_t* __restrict__ base,
                         const T* __restrict__ w, int n, int Np, T* __restrict__ out, int* __restrict__ counter,
                         int cap, int es) {
  __shared__ int row;
  if (threadIdx.x == 0) row = *counter;
  __syncthreads();
  const int r0 = row;
  if (r0 < cap) {
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const int64_t b = base[r];
      T acc = 0;
      for (int k = 0; k < Np; ++k) {
        const int64_t j = (int64_t)(b + k) * es;  // pressure of node k
        T u = U[j];
        if (Ulo) u += (T)bf16_bits_to_f32(Ulo[j]);
        acc += w[r * Np + k] * u;
      }
      out[(int64_t)r0 * n + r] = acc;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && r0 < cap) *counter = r0 + 1;
}

template <typename T>
__global__ void k_finite(const T* __restrict__ x, int64_t n, int* __restrict__ flag) {
  int bad = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    bad |= !isfinite(x[i]);
  if (__any(bad) && (threadIdx.x % 64) == 0) atomicOr(flag, 1);
}

// FP64 field planes (plane stride PS; element nodes [off, off+n) of each plane) -> a state buffer in
// the context's precision and layout (es > 1: node-interleaved with es fields, the planes beyond es
// being zero; es = 1: planes), with the bf16 low word of the FP32 remainder when lo is given; and
// back (planes beyond es come back as zero).
__device__ __forceinline__ void field_node(int64_t i, int64_t n, int64_t off, int es, int64_t PS, int64_t* q,
                                           int64_t* node, int64_t* di) {
  if (es > 1) {
    *q = i & 3;
    *node = off + (i >> 2);
    *di = *node * es + *q;
  } else {
    *q = i / n;
    *node = off + (i - *q * n);
    *di = *q * PS + *node;
  }
}
template <typename T>
__global__ void k_from_f64(const double* __restrict__ s, int64_t PS, int64_t off, int64_t n, T* __restrict__ d,
                           uint16_t* __restrict__ lo, int es) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 4 * n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t q, node, di;
    field_node(i, n, off, es, PS, &q, &node, &di);
    if (es > 1 && q >= es) continue;
    const double x = s[q * PS + node];
    if constexpr (sizeof(T) == 8) {
      d[di] = x;
    } else {
      const float h = (float)x;
      d[di] = h;
      if (lo) lo[di] = (uint16_t)f32_to_bf16_bits((float)(x - (double)h));
    }
  }
}
template <typename T>
__global__ void k_to_f64(const T* __restrict__ d, const uint16_t* __restrict__ lo, int64_t PS, int64_t off, int64_t n,
                         double* __restrict__ s, int es) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 4 * n; i += (int64_t)gridDim.x * blockDim.x) {
    int64_t q, node, di;
    field_node(i, n, off, es, PS, &q, &node, &di);
    s[q * PS + node] = (es > 1 && q >= es) ? 0.0 : (double)d[di] + (lo ? (double)bf16_bits_to_f32(lo[di]) : 0.0);
  }
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

template <typename T, int D, int P, bool LO, bool CS, int ES, bool BIG, int SH>
void launch_mfma_kernel(StageArgs& a, hipStream_t s, int grid_pct) {
  constexpr int TE = MfmaLaunch<T, D, P, SH>::TE, KB = MfmaLaunch<T, D, P, SH>::KB;
  static int occ = 0;
  if (!occ) {
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_stage_mfma<T, D, P, LO, CS, ES, BIG, SH>, KB, 0) !=
            hipSuccess ||
        occ < 1)
      occ = 1;
  }
  a.tiles = a.Kpad / TE;
  a.tiles_per_xcd = (int)((a.tiles + 7) / 8);
  const int64_t per_xcd_cap = (int64_t)(num_cus() / 8) * occ * std::max(grid_pct, 1) / 100;
  int bpx = (int)std::min<int64_t>(a.tiles_per_xcd, std::max<int64_t>(per_xcd_cap, 1));
  if (bpx < 1) bpx = 1;
  hipLaunchKernelGGL((k_stage_mfma<T, D, P, LO, CS, ES, BIG, SH>), dim3(8 * bpx), dim3(KB), 0, s, a);
}

// One of the other tile shapes (1-4). A shape that resolves to the built-in one's workgroup, tile
// and operator placement runs the built-in kernel. Returns false, launching nothing, when the shape
// does not fit this element type (LDS or whole row blocks).
template <typename T, int D, int P, bool LO, bool CS, int ES, int SH>
bool launch_shape(StageArgs& a, hipStream_t s, int grid_pct, bool dry) {
  using L0 = MfmaLaunch<T, D, P, 0>;
  using L = MfmaLaunch<T, D, P, SH>;
  using C0 = MfmaCfg<T, D, P, LO, ES, 0>;
  using C = MfmaCfg<T, D, P, LO, ES, SH>;
  if constexpr (L::WAVES == L0::WAVES && L::TE == L0::TE && L::RG == L0::RG && L::MINW == L0::MINW &&
                L::DEFER == L0::DEFER && C::kALds == C0::kALds) {
    if (!dry) launch_mfma_kernel<T, D, P, LO, CS, ES, false, 0>(a, s, grid_pct);
    return true;
  } else if constexpr (C::fits) {
    if (!dry) launch_mfma_kernel<T, D, P, LO, CS, ES, false, SH>(a, s, grid_pct);
    return true;
  } else {
    return false;
  }
}

// 32-bit offsets from the buffers' own bases while every array the kernel addresses (state,
// J^-1, face geometry; the face tables are smaller) stays below 4 GiB, else per-tile 64-bit bases.
// Tile shapes other than the built-in one are built for the forms the workloads run -- FP64, and
// FP32 with the compensated solution, without the operator correction -- in 2D and 3D at orders
// 2-5 with the initial live fields (2D: 3, 3D: 4), below 4 GiB per buffer. Returns false, launching
// nothing, when `shape` has no kernel for this context's form.
template <typename T, int D, int P, bool LO, bool CS, int ES>
bool launch_mfma_es(StageArgs& a, hipStream_t s, int grid_pct, int shape, bool dry) {
  const bool big = a.addr64 || stage_needs_addr64(a.Kpad, Sz<D, P>::NP, Sz<D, P>::NF, ES, (int)sizeof(T));
  if (shape != 0) {
    constexpr bool kShapes =
        !LO && D >= 2 && P >= 2 && ES == (D == 3 ? 4 : D + 1) && (sizeof(T) == 8 ? !CS : CS);
    if constexpr (kShapes) {
      if (!big) switch (shape) {
          case 1: return launch_shape<T, D, P, LO, CS, ES, 1>(a, s, grid_pct, dry);
          case 2: return launch_shape<T, D, P, LO, CS, ES, 2>(a, s, grid_pct, dry);
          case 3: return launch_shape<T, D, P, LO, CS, ES, 3>(a, s, grid_pct, dry);
          case 4: return launch_shape<T, D, P, LO, CS, ES, 4>(a, s, grid_pct, dry);
          case 5: return launch_shape<T, D, P, LO, CS, ES, 5>(a, s, grid_pct, dry);
          case 6: return launch_shape<T, D, P, LO, CS, ES, 6>(a, s, grid_pct, dry);
          case 7: return launch_shape<T, D, P, LO, CS, ES, 7>(a, s, grid_pct, dry);
          case 8: return launch_shape<T, D, P, LO, CS, ES, 8>(a, s, grid_pct, dry);
          case 9: return launch_shape<T, D, P, LO, CS, ES, 9>(a, s, grid_pct, dry);
          case 10: return launch_shape<T, D, P, LO, CS, ES, 10>(a, s, grid_pct, dry);
          case 11: return launch_shape<T, D, P, LO, CS, ES, 11>(a, s, grid_pct, dry);
          default: break;
        }
    }
    return false;
  }
  if (dry) return true;
  if (!big) {
    launch_mfma_kernel<T, D, P, LO, CS, ES, false, 0>(a, s, grid_pct);
    return true;
  }
  if constexpr (LO) throw Unsupported("the FP32 operator correction is limited to contexts below 4 GiB per buffer");
  else launch_mfma_kernel<T, D, P, LO, CS, ES, true, 0>(a, s, grid_pct);
  return true;
}

// MFMA contexts keep the state node-interleaved with their live fields: ES = 4 in 3D, and in
// 1D/2D dim+1 (the extra velocity components identically zero) or 4.
template <typename T, int D, int P, bool LO, bool CS>
bool launch_mfma(StageArgs& a, hipStream_t s, int grid_pct, int es, int shape, bool dry) {
  if constexpr (D < 3)
    if (es == D + 1) return launch_mfma_es<T, D, P, LO, CS, D + 1>(a, s, grid_pct, shape, dry);
  if (es != 4) throw InvalidArg("stage kernel: interleaved state with an unsupported field count");
  return launch_mfma_es<T, D, P, LO, CS, 4>(a, s, grid_pct, shape, dry);
}

template <typename T, int D, int P>
bool launch_mfma_form(Ctx& c, StageArgs& a, hipStream_t s, int shape, bool dry) {
  if constexpr (std::is_same<T, float>::value) {
    if (c.comp && c.oplo) return launch_mfma<T, D, P, true, true>(a, s, c.grid_pct, c.es, shape, dry);
    if (c.comp) return launch_mfma<T, D, P, false, true>(a, s, c.grid_pct, c.es, shape, dry);
    if (c.oplo) return launch_mfma<T, D, P, true, false>(a, s, c.grid_pct, c.es, shape, dry);
  }
  return launch_mfma<T, D, P, false, false>(a, s, c.grid_pct, c.es, shape, dry);
}

// The context's tile shape, or the built-in one where the context's current form (live fields,
// buffer sizes) has no kernel of that shape. With `dry`, launches nothing and tells whether
// `shape` has a kernel for the context as it is.
template <typename T, int D, int P>
bool launch_variant(Ctx& c, StageArgs& a, hipStream_t s, int shape, bool dry) {
  if (c.variant == kVariantLds) {
    if (dry) return shape == 0;
    launch_lds<T, D, P>(a, s);
    return true;
  }
  if (launch_mfma_form<T, D, P>(c, a, s, shape, dry)) return true;
  if (dry) return false;
  return launch_mfma_form<T, D, P>(c, a, s, 0, false);
}

template <typename T>
bool dispatch(Ctx& c, StageArgs& a, hipStream_t s, int shape, bool dry) {
#define DGK_CASE(D, P) \
  if (c.dim == D && c.p == P) return launch_variant<T, D, P>(c, a, s, shape, dry);
  DGK_CASE(1, 1) DGK_CASE(1, 2) DGK_CASE(1, 3) DGK_CASE(1, 4) DGK_CASE(1, 5)
  DGK_CASE(2, 1) DGK_CASE(2, 2) DGK_CASE(2, 3) DGK_CASE(2, 4) DGK_CASE(2, 5)
  DGK_CASE(3, 1) DGK_CASE(3, 2) DGK_CASE(3, 3) DGK_CASE(3, 4) DGK_CASE(3, 5)
#undef DGK_CASE
  throw Unsupported("no stage kernel for this element type");
}

}  // namespace

void launch_stage(Ctx& c, StageArgs& a, hipStream_t s) {
  if (c.precision == DGK_FP64) dispatch<double>(c, a, s, c.tile_shape, false);
  else dispatch<float>(c, a, s, c.tile_shape, false);
  DGK_HIP(hipGetLastError());
}

bool tile_shape_supported(Ctx& c, StageArgs& a, int shape) {
  if (shape < 0 || shape >= kTileShapes) return false;
  return c.precision == DGK_FP64 ? dispatch<double>(c, a, nullptr, shape, true)
                                 : dispatch<float>(c, a, nullptr, shape, true);
}

void launch_source(Ctx& c, const Ctx::Src& src, double value, hipStream_t st) {
  if (!src.n) return;
  unsigned nb = (unsigned)((src.n + 255) / 256);
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_source<double>, dim3(nb), dim3(256), 0, st, (double*)c.U, (uint16_t*)nullptr, src.idx,
                       src.n, value, c.es);
  else
    hipLaunchKernelGGL(k_source<float>, dim3(nb), dim3(256), 0, st, (float*)c.U, (uint16_t*)(c.comp ? c.Ulo : nullptr),
                       src.idx, src.n, (float)value, c.es);
  DGK_HIP(hipGetLastError());
}

void launch_source_timed(Ctx& c, const Ctx::Src& src, hipStream_t st) {
  if (!src.n) return;
  const unsigned nb = (unsigned)std::min<int64_t>((src.n + 255) / 256, 1024);
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_source_timed<double>, dim3(nb), dim3(256), 0, st, (double*)c.U, (uint16_t*)nullptr, src.idx,
                       src.n, src.amp, src.freq, src.phase, src.duration, (const double*)c.tdev,
                       (const int*)c.tstep, c.es);
  else
    hipLaunchKernelGGL(k_source_timed<float>, dim3(nb), dim3(256), 0, st, (float*)c.U,
                       (uint16_t*)(c.comp ? c.Ulo : nullptr), src.idx, src.n, src.amp, src.freq, src.phase,
                       src.duration, (const double*)c.tdev, (const int*)c.tstep, c.es);
  DGK_HIP(hipGetLastError());
}

void launch_tick(Ctx& c, hipStream_t st) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, st, c.tstep);
  DGK_HIP(hipGetLastError());
}

void launch_record(Ctx& c, hipStream_t st) {
  if (!c.rec.n) return;
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_record<double>, dim3(1), dim3(256), 0, st, (const double*)c.U, (const uint16_t*)nullptr,
                       c.rec.base, (const double*)c.rec.w, c.rec.n, c.Np, (double*)c.rec.out, c.rec.counter,
                       c.rec.cap, c.es);
  else
    hipLaunchKernelGGL(k_record<float>, dim3(1), dim3(256), 0, st, (const float*)c.U,
                       (const uint16_t*)(c.comp ? c.Ulo : nullptr), c.rec.base, (const float*)c.rec.w, c.rec.n,
                       c.Np, (float*)c.rec.out, c.rec.counter, c.rec.cap, c.es);
  DGK_HIP(hipGetLastError());
}

bool check_finite(Ctx& c) {
  DGK_HIP(hipMemsetAsync(c.flag, 0, sizeof(int), c.stream));
  int64_t n = 4 * c.PS;
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

void state_from_f64(Ctx& c, const double* s, void* dst, uint16_t* lo, int64_t off, int64_t n) {
  const int es = c.es;
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_from_f64<double>, dim3(1024), dim3(256), 0, c.stream, s, c.PS, off, n, (double*)dst,
                       (uint16_t*)nullptr, es);
  else
    hipLaunchKernelGGL(k_from_f64<float>, dim3(1024), dim3(256), 0, c.stream, s, c.PS, off, n, (float*)dst, lo, es);
  DGK_HIP(hipGetLastError());
}

void state_to_f64(Ctx& c, const void* src, const uint16_t* lo, double* s, int64_t off, int64_t n) {
  const int es = c.es;
  if (c.precision == DGK_FP64)
    hipLaunchKernelGGL(k_to_f64<double>, dim3(1024), dim3(256), 0, c.stream, (const double*)src,
                       (const uint16_t*)nullptr, c.PS, off, n, s, es);
  else
    hipLaunchKernelGGL(k_to_f64<float>, dim3(1024), dim3(256), 0, c.stream, (const float*)src, lo, c.PS, off, n, s,
                       es);
  DGK_HIP(hipGetLastError());
}

const char* variant_name(int v) {
  switch (v) {
    case kVariantLds: return "lds-tile";
    case kVariantMfma: return "mfma16";
    default: return "unknown";
  }
}

void mfma_shape(int dim, int p, bool fp32, bool oplo, int* KS, int* MB, int* NT) {
#define DGK_SHAPE(D, P)                                                                   \
  if (dim == D && p == P) {                                                               \
    *KS = MfmaShape<D, P>::KS;                                                            \
    *MB = MfmaShape<D, P>::MB;                                                            \
    *NT = MfmaShape<D, P>::tail(fp32, oplo);                                              \
    return;                                                                               \
  }
  DGK_SHAPE(1, 1) DGK_SHAPE(1, 2) DGK_SHAPE(1, 3) DGK_SHAPE(1, 4) DGK_SHAPE(1, 5)
  DGK_SHAPE(2, 1) DGK_SHAPE(2, 2) DGK_SHAPE(2, 3) DGK_SHAPE(2, 4) DGK_SHAPE(2, 5)
  DGK_SHAPE(3, 1) DGK_SHAPE(3, 2) DGK_SHAPE(3, 3) DGK_SHAPE(3, 4) DGK_SHAPE(3, 5)
#undef DGK_SHAPE
  *KS = 0; *MB = 0; *NT = 0;
}

}  // namespace dgk
