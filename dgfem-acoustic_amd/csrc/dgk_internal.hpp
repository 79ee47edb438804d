// Internal declarations shared by the libdgk translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgk.h"

namespace dgk {

#define DGK_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw DeviceError(std::string(#call) + ": " + hipGetErrorString(e_));                    \
  } while (0)

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct MeshError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct InvalidArg : std::runtime_error {
  using std::runtime_error::runtime_error;
};

constexpr int kMaxNF = 4;
constexpr int kTilePad = 32;  // device arrays hold a multiple of this many elements (the widest tile)
// stage-kernel variants (dgk_opts.variant; 0 picks the default)
constexpr int kVariantLds = 1;   // thread per element node, operators read from memory
constexpr int kVariantMfma = 2;  // 16x16x4 MFMA contraction, one wave per 4 elements
// The MFMA stage kernel addresses its arrays with 32-bit byte offsets from the buffers' own bases
// while every one of them (state: Kpad*np*es values, J^-1: Kpad*9, face geometry: Kpad*nf*4) stays
// below 4 GiB, and through 64-bit per-tile bases otherwise.
inline bool stage_needs_addr64(int64_t Kpad, int np, int nf, int es, int tsize) {
  const uint64_t lim = (uint64_t)1 << 32, kp = (uint64_t)Kpad;
  return kp * np * es * tsize >= lim || kp * 9 * tsize >= lim || kp * nf * 4 * tsize >= lim;
}
// MFMA stage-kernel tile shapes (dgk_set_tile_shape): 0 built-in, 1 4 waves x 16 elements,
// 2 8 waves x 32 elements, 3 row split (4 waves x 4 elements), 4 4 waves x 16 elements with the
// operator read from global memory, 5 row split over 8 waves x 8 elements, 6-8 4 waves x 16
// elements compiled for 3-5 waves per SIMD, 9, 10, 11 as 7, 8, 6 with the epilogue operands
// loaded late
constexpr int kTileShapes = 12;
constexpr int kMaxNFP = 21;
constexpr int kMaxPerm = 6;

// Operator tables for one (dim, order), in the precision of the kernels.
struct DevOps {
  void* W = nullptr;     // [dim][Np][Np]  W_r = Mhat^-1 S_r
  void* LIFT = nullptr;  // [Np][Nf*Nfp]
  int* fmask = nullptr;  // [Nf][Nfp]
  int* vmap = nullptr;   // [Nf*nperm][Nfp] neighbour element node of own face node k
  void* Amf = nullptr;   // [MB][KS][64] MFMA A fragments of [W_1..W_d | -LIFT] in k-step order
  void* Alo = nullptr;   // FP32 only: [MB][ceil(KS/8)][64][8] bf16 residuals A - fl32(A)
};

// Per-stage kernel arguments (precision-agnostic pointers).
struct StageArgs {
  int64_t K;
  int64_t Kpad;  // element count padded to whole tiles
  int64_t PS;    // field plane stride in values (Kpad * Np)
  const void* Vin;
  void* Vout;
  void* U;
  void* A;
  void* Ulo;              // FP32 compensated form: bf16 low word of U (uint16 bits), else null
  const void* Jinv;       // [K][9]
  const void* fgeom;      // [K][Nf][4]: unit outward normal, sJ/|detJ|
  const int32_t* nbr;     // [K*Nf]
  const int32_t* code;    // [K*Nf]: (nbr_face*nperm + perm) | (penalty sign -1 ? 1<<16 : 0)
  const void* W;
  const void* LIFT;
  const int* fmask;
  const int* vmap;
  const void* Amf;
  const void* Alo;
  double refn[kMaxNF][3];  // outward reference-face normals
  double fsc[kMaxNF];      // reference face measure / face reference-element measure
  double dt, c0, rho0, v0[3];
  int stage;  // 0..3 RK4 stage, 4 = rhs (Vout = dt*L), 5 = forward Euler (Vout = Vin + dt*L)
  int nq;     // field planes that carry data: dim+1 when the velocity planes beyond dim are zero, else 4
  int tiles_per_xcd;
  int64_t tiles;
  int tile_order;  // 0: XCD-aware (XCD x sweeps a contiguous slab of tiles), 1: round-robin
  int addr64;      // 1: 64-bit per-tile bases even when every array fits 32-bit offsets
};

struct Ctx {
  int dim = 0, p = 0, Np = 0, Nf = 0, Nfp = 0, nperm = 1;
  int64_t K = 0, F = 0, Fb = 0;
  int64_t Kpad = 0, PS = 0;  // elements padded to kTilePad, field plane stride (values)
  int precision = DGK_FP64, sign_mode = DGK_SIGN_REFERENCE, device = 0, variant = 0;
  double fc = 1;
  double c0 = 1, rho0 = 1, v0[3] = {0, 0, 0};
  hipStream_t stream = nullptr;
  size_t tsize = 8;  // bytes per state value
  // device buffers
  void *U = nullptr, *A = nullptr, *V0 = nullptr, *V1 = nullptr, *Jinv = nullptr, *fgeom = nullptr;
  uint16_t* Ulo = nullptr;  // FP32 compensated form: bf16 low word of U
  bool comp = false;        // FP32: compensated solution (U, Ulo)
  bool oplo = false;        // FP32: bf16 operator-residual correction on MFMA
  // State layout of the device buffers U, A, V0, V1 (and Ulo): on the MFMA kernel node-interleaved
  // [Kpad][Np][es] with the es = nq live fields; on the thread-per-node kernel (es = 1) four field
  // planes [4][Kpad*Np] (plane stride PS). The host side always exchanges field planes.
  int es = 1;
  // Field planes that carry data. In 1D/2D the velocity components beyond dim stay identically
  // zero under the dynamics when they start at zero (no derivative or normal has those
  // components), so their planes are neither read nor written: nq = dim+1. A state with nonzero
  // values there keeps all 4 planes live.
  int nq = 4;
  // launch shape (tunable): persistent grid as a percentage of the occupancy-derived grid, the
  // tile order, and the MFMA kernel's tile shape (kTileShapes; 0 the built-in one)
  int grid_pct = 100, tile_order = 0, tile_shape = 0;
  bool addr64 = false;  // force the 64-bit addressing form of the stage kernel
  int32_t *nbr = nullptr, *code = nullptr, *nface = nullptr, *perm = nullptr;
  double* stage_d = nullptr;  // FP64 staging for state transfer in FP32 mode
  int* flag = nullptr;
  DevOps ops;
  double refn[kMaxNF][3] = {};
  double fsc[kMaxNF] = {};
  // sources
  struct Src {
    double amp, freq, phase, duration;
    int64_t* idx = nullptr;  // node offsets e*Np + node of the covered element nodes
    int64_t n = 0;
  };
  std::vector<Src> sources;
  // receivers: pressure interpolated at points (element Lagrange basis), recorded after every step
  struct Rec {
    int n = 0, cap = 0;
    int64_t* base = nullptr;  // [n] offset e*Np of the receiver's element in the pressure plane
    void* w = nullptr;       // [n][Np] basis weights in the state type
    void* out = nullptr;     // [cap][n] recorded pressure in the state type
    int* counter = nullptr;  // rows recorded so far (device)
  } rec;
  void free_receivers();
  int64_t device_bytes = 0;
  // Captured RK4 step sequences, by (steps, sources launched) for time step graph_dt. Captures
  // hold buffer addresses and launch settings: drop_graphs() whenever those change.
  std::map<std::pair<int, bool>, hipGraphExec_t> graphs;
  double graph_dt = 0;
  void drop_graphs();
  // Start times of the steps of the current dgk_rk4 call on the device (sources read
  // tdev[*tstep]; *tstep advances after every step)
  double* tdev = nullptr;
  int tcap = 0;
  int* tstep = nullptr;
  std::string err;
  std::vector<void*> allocs;

  void* alloc(size_t bytes) {
    void* p = nullptr;
    DGK_HIP(hipMalloc(&p, bytes));
    allocs.push_back(p);
    device_bytes += (int64_t)bytes;
    return p;
  }
  ~Ctx();
};

// setup.hip
void build_connectivity(Ctx& c, const double* vxyz_dev /* K*(dim+1)*3 */, const int32_t* vid_host /* K*(dim+1) */,
                        const uint8_t* bc_first_host /* K*Nf: code of each face's first node */, double sigma);
void build_geometry(Ctx& c, const double* vxyz_dev);
void scale_geometry(Ctx& c, int64_t e0, int64_t n, double s);
// stage.hip
void launch_stage(Ctx& c, StageArgs& a, hipStream_t s);
// whether the MFMA kernel has a kernel of tile shape `shape` for the context as it is (form, live
// fields, buffer sizes); `a` as for launch_stage, nothing is launched
bool tile_shape_supported(Ctx& c, StageArgs& a, int shape);
void launch_source(Ctx& c, const Ctx::Src& s, double value, hipStream_t st);
// hard source overwrite at the step whose start time is c.tdev[*c.tstep] (when t < duration)
void launch_source_timed(Ctx& c, const Ctx::Src& s, hipStream_t st);
void launch_tick(Ctx& c, hipStream_t st);  // ++*c.tstep
void launch_record(Ctx& c, hipStream_t st);  // no-op without receivers
bool check_finite(Ctx& c);
// FP64 field planes on the device (plane stride PS; element nodes [off, off+n) of each plane) -> a
// state buffer in the context's precision and layout; in the FP32 compensated form lo (if given)
// receives the bf16 low words. And back, adding the low words when lo is given.
void state_from_f64(Ctx& c, const double* s_dev, void* dst_dev, uint16_t* lo, int64_t off, int64_t n);
void state_to_f64(Ctx& c, const void* src_dev, const uint16_t* lo, double* s_dev, int64_t off, int64_t n);
const char* variant_name(int v);
// MFMA k-steps, 16-row blocks and 4-row tail pieces (0: none) of the stage kernel's operator
void mfma_shape(int dim, int p, bool fp32, bool oplo, int* KS, int* MB, int* NT);
// dgk_api.cpp
extern thread_local std::string g_err;
void create_impl(Ctx& c, const dgk_mesh* m, const dgk_phys* ph, const dgk_opts* o);
void rk4_step(Ctx& c, double dt, hipStream_t s, hipEvent_t* ev = nullptr);
StageArgs base_args(Ctx& c);  // stage-kernel arguments over the context's buffers
double* staging(Ctx& c);
int live_planes(int dim, const double* u, int64_t n);  // nq of a host state u[4][n]
// tuning table: launch settings per (dim, order, precision), loaded by dgk_load_tuning
struct Tuned {
  int variant = 0, grid_pct = 100, tile_order = 0, tile_shape = 0;
};
bool find_tuning(int dim, int p, int precision, Tuned* t);
struct RefElem;
std::vector<int64_t> locate_points(const RefElem& R, const double* node_xyz, int64_t K, int n, const double* xyz,
                                   std::vector<double>& w);
void install_receivers(Ctx& c, const std::vector<int64_t>& elem, const std::vector<double>& w, int cap);
int fetch_receivers(Ctx& c, std::vector<double>& out);
void restore_receivers(Ctx& c, const std::vector<double>& rows_host, int rows);


}  // namespace dgk

struct dgk_ctx : dgk::Ctx {};
