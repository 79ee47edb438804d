// Source-to-receiver surrogate: a Fourier-feature MLP trained and queried on the GPU with this
// file's kernels (BF16 MFMA GEMMs with FP32 accumulation, fused bias/activation epilogues, loss,
// bias reduction, Adam). No reference counterpart: the reference has no learned component.
//
// Network: x in [0,1]^in (normalised source and receiver coordinates)
//   z = [sin(2 pi B x), cos(2 pi B x)]                      B: F x in, fixed, N(0, sigma^2)
//   h_1 = silu(W_1 z + b_1), h_l = silu(W_l h_{l-1} + b_l)  width w, l = 1..depth
//   y = W_o h_depth + b_o                                   out samples
// GEMM operands are BF16 (weights keep an FP32 master copy), accumulation is FP32; activations
// are rounded to BF16 once, where they become the next GEMM's operand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgk_surrogate.h"

// Scratch of the allocation-free forward pass (dgs_forward_noalloc): one copy per GPU, held until
// the process ends. Host code never uses this name as an address; it uses the per-device pointer
// that hipGetSymbolAddress returns (dgs::scratch_base).
static __device__ __align__(256) unsigned char scratch[16u << 20];

namespace dgs {

constexpr size_t kScratchBytes = 16u << 20;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Error : std::runtime_error {
  dgs_status st;
  Error(dgs_status s, const std::string& m) : std::runtime_error(m), st(s) {}
};

#define DGS_HIP(x)                                                                              \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) throw ::dgs::Error(DGS_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite)
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline float sigm(float x) { return 1.0f / (1.0f + expf(-x)); }
__device__ inline float silu(float x) { return x * sigm(x); }
__device__ inline float dsilu(float x) {
  const float s = sigm(x);
  return s * (1.0f + x * (1.0f - s));
}

enum { EPI_STORE = 0, EPI_BIAS_ACT = 1, EPI_BIAS = 2, EPI_DACT = 3 };

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;  // element (m, k) at A[m*sam + k*sak]
  int64_t sam, sak;
  const uint16_t* B;  // element (k, n) at B[n*sbn + k*sbk]
  int64_t sbn, sbk;
  const float* bias;  // [N]
  float* pre;         // [M][ldc] pre-activation (written by BIAS_ACT, read by DACT)
  uint16_t* outb;     // [M][ldc] BF16 result
  float* outf;        // [M][ldc] FP32 result
  int ldc;
};

// C[M,N] = A[M,K] B[K,N] on v_mfma_f32_16x16x32_bf16. Workgroup: 4 waves, a 64x64 tile of C, each
// wave a 32x32 quarter as 2x2 MFMA tiles; K advances 32 per step through LDS, where both operands
// are stored k-contiguous ([row][k], rows padded to 40 values) so each lane's fragment is one
// 16-byte read. Operand lane layout (probed on gfx950): A lane l holds A[l%16][8(l/16)+i],
// B lane l holds B[8(l/16)+i][l%16]; accumulator register j of lane l is C[4(l/16)+j][l%16].
// AT / BT: the operand is m- / n-contiguous in memory (else k-contiguous).
// Shapes: M, N multiples of 64, K a multiple of 32 (checked on the host).
template <bool AT, bool BT, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  constexpr int LD = 40;
  __shared__ __attribute__((aligned(16))) uint16_t sA[64 * LD];
  __shared__ __attribute__((aligned(16))) uint16_t sB[64 * LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0, 0, 0, 0};
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    // stage the 64x32 operand tiles
    if (!AT) {
      const int r = tid >> 2, kq = (tid & 3) * 8;
      *(uint4*)&sA[r * LD + kq] = *(const uint4*)&g.A[(int64_t)(m0 + r) * g.sam + (k0 + kq)];
    } else {
      const int r = tid & 63, kq = (tid >> 6) * 8;
      for (int i = 0; i < 8; ++i) sA[r * LD + kq + i] = g.A[(int64_t)(m0 + r) * g.sam + (int64_t)(k0 + kq + i) * g.sak];
    }
    if (!BT) {
      const int r = tid >> 2, kq = (tid & 3) * 8;
      *(uint4*)&sB[r * LD + kq] = *(const uint4*)&g.B[(int64_t)(n0 + r) * g.sbn + (k0 + kq)];
    } else {
      const int r = tid & 63, kq = (tid >> 6) * 8;
      for (int i = 0; i < 8; ++i) sB[r * LD + kq + i] = g.B[(int64_t)(n0 + r) * g.sbn + (int64_t)(k0 + kq + i) * g.sbk];
    }
    __syncthreads();
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; ++i) {
      const uint4 va = *(const uint4*)&sA[(wr + 16 * i + (lane & 15)) * LD + 8 * (lane >> 4)];
      a[i] = __builtin_bit_cast(bf16x8, va);
      const uint4 vb = *(const uint4*)&sB[(wc + 16 * i + (lane & 15)) * LD + 8 * (lane >> 4)];
      b[i] = __builtin_bit_cast(bf16x8, vb);
    }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr + 16 * i + 4 * (lane >> 4) + r;
        const int col = n0 + wc + 16 * j + (lane & 15);
        const int64_t o = (int64_t)row * g.ldc + col;
        const float v = acc[i][j][r];
        if (EPI == EPI_STORE) {
          g.outf[o] = v;
        } else if (EPI == EPI_BIAS_ACT) {
          const float p = v + g.bias[col];
          g.pre[o] = p;
          g.outb[o] = f2bf(silu(p));
        } else if (EPI == EPI_BIAS) {
          g.outf[o] = v + g.bias[col];
        } else {
          g.outb[o] = f2bf(v * dsilu(g.pre[o]));
        }
      }
}

// z[r][f] = sin(2 pi B_f . x_r), z[r][F + f] = cos(2 pi B_f . x_r), rounded to BF16.
__global__ void k_features(int n, int in, int F, const float* __restrict__ x, const float* __restrict__ Bm,
                           uint16_t* __restrict__ z) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)n * F) return;
  const int r = (int)(i / F), f = (int)(i % F);
  float a = 0;
  for (int d = 0; d < in; ++d) a += Bm[f * in + d] * x[r * in + d];
  a *= 6.283185307179586f;
  z[(int64_t)r * 2 * F + f] = f2bf(sinf(a));
  z[(int64_t)r * 2 * F + F + f] = f2bf(cosf(a));
}

// Relative squared error per sample: loss_r = w_r sum_j (y - t)^2, w_r = 1 / (n |t_r|^2);
// dY = 2 w_r (y - t) rounded to BF16. One workgroup per sample row.
__global__ __launch_bounds__(256) void k_loss(int N, const float* __restrict__ y, const float* __restrict__ t,
                                              const float* __restrict__ w, uint16_t* __restrict__ dy,
                                              float* __restrict__ loss_row) {
  __shared__ float red[256];
  const int r = blockIdx.x;
  float s = 0;
  for (int j = threadIdx.x; j < N; j += 256) {
    const int64_t o = (int64_t)r * N + j;
    const float d = y[o] - t[o];
    s += d * d;
    dy[o] = f2bf(2.0f * w[r] * d);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_row[r] = w[r] * red[0];
}

// db[n] = sum_m d[m][n] over BF16 rows.
__global__ void k_colsum(int M, int N, const uint16_t* __restrict__ d, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0;
  for (int m = 0; m < M; ++m) s += bf2f(d[(int64_t)m * N + n]);
  db[n] = s;
}

// sum of n values into out[0] (one workgroup)
__global__ __launch_bounds__(256) void k_sum(int n, const float* __restrict__ v, float* __restrict__ out) {
  __shared__ float red[256];
  float s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// Adam with bias correction; refreshes the BF16 copy of the weights.
__global__ void k_adam(int64_t n, float* __restrict__ p, uint16_t* __restrict__ pb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2, float c1, float c2,
                       float eps) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float w = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = w;
    pb[i] = f2bf(w);
  }
}

__global__ void k_tobf(int64_t n, const float* __restrict__ p, uint16_t* __restrict__ pb) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    pb[i] = f2bf(p[i]);
}

// rows idx[r] of (x, t, w) -> batch buffers; the weight is divided by the batch size so that the
// minibatch loss is the mean of W_r |y_r - t_r|^2
__global__ void k_gather(int n, int in, int N, const int* __restrict__ idx, const float* __restrict__ X,
                         const float* __restrict__ T, const float* __restrict__ W, float* __restrict__ x,
                         float* __restrict__ t, float* __restrict__ w) {
  const int r = blockIdx.x;
  const int64_t s = idx[r];
  for (int j = threadIdx.x; j < N; j += blockDim.x) t[(int64_t)r * N + j] = T[s * N + j];
  if (threadIdx.x < in) x[r * in + threadIdx.x] = X[s * in + threadIdx.x];
  if (threadIdx.x == 0) w[r] = W[s] / (float)n;
}

}  // namespace

// ------------------------------------------------------------------------------- the network
struct Net {
  dgs_config cfg{};
  int device = 0;                  // HIP device the network lives on
  int L = 0;                       // weight layers (depth hidden + output)
  std::vector<int> din, dout;      // per layer
  std::vector<int64_t> woff, boff; // offsets into the flat parameter vector
  int64_t nparam = 0;
  hipStream_t s = nullptr;
  float* P = nullptr;   // FP32 master parameters [W_1 b_1 ... W_L b_L]
  uint16_t* Pb = nullptr;  // BF16 copy (only the W parts are read)
  float* G = nullptr;   // gradients, same layout
  float* Am = nullptr;  // Adam moments
  float* Av = nullptr;
  float* Bm = nullptr;  // Fourier matrix F x in
  int64_t adam_t = 0;
  // per-batch buffers (max_batch rows)
  float *x = nullptr, *t = nullptr, *w = nullptr, *y = nullptr, *loss_row = nullptr;
  uint16_t* z = nullptr;
  std::vector<uint16_t*> h;  // BF16 activations of hidden layers
  std::vector<float*> pre;   // FP32 pre-activations of hidden layers
  uint16_t *dA = nullptr, *dB = nullptr;  // ping-pong BF16 gradients of pre-activations
  // resident training set
  float *X = nullptr, *T = nullptr, *W = nullptr;
  int64_t ndata = 0;
  std::vector<void*> allocs;
  std::string err;

  template <typename U>
  U* alloc(size_t n) {
    void* p = nullptr;
    DGS_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(U)));
    allocs.push_back(p);
    return (U*)p;
  }
  ~Net() {
    for (void* p : allocs) (void)hipFree(p);
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

thread_local std::string g_err;

template <class F>
dgs_status guarded(Net* n, F&& f) {
  try {
    f();
    return DGS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    if (n) n->err = e.what();
    return e.st;
  } catch (const std::exception& e) {
    g_err = e.what();
    if (n) n->err = e.what();
    return DGS_ERR_INVALID;
  }
}

template <bool AT, bool BT, int EPI>
void gemm(Net& n, GemmArgs g) {
  if (g.M % 64 || g.N % 64 || g.K % 32 || g.M <= 0 || g.N <= 0 || g.K <= 0)
    throw Error(DGS_ERR_INVALID, "GEMM shape must be M, N multiples of 64 and K a multiple of 32");
  hipLaunchKernelGGL((k_gemm<AT, BT, EPI>), dim3(g.N / 64, g.M / 64), dim3(256), 0, n.s, g);
  DGS_HIP(hipGetLastError());
}

// Device buffers of one forward pass: inputs x [rows][in], features z, per hidden layer l the BF16
// activations h(l) and FP32 pre-activations pre(l), output y [rows][out].
struct Pass {
  const float* x;
  uint16_t* z;
  uint16_t* h[2];  // hidden layer l writes h[l % 2] (when not per layer)
  float* pre;      // one buffer for every hidden layer (when not per layer)
  const std::vector<uint16_t*>* hl = nullptr;  // per-layer buffers (training keeps them for backward)
  const std::vector<float*>* prel = nullptr;
  float* y;
};

// rows [0, mp) of dst take source rows min(r, m - 1) of x (padding with copies of the last row)
__global__ void k_pad_rows(const float* __restrict__ x, int m, int mp, int in, float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mp * in) return;
  const int r = i / in, c = i - r * in;
  dst[i] = x[(int64_t)min(r, m - 1) * in + c];
}

// Device address of `scratch` on device dev, looked up once per device with that device current.
std::mutex g_scratch_mu;
unsigned char* g_scratch[64] = {};
unsigned char* scratch_base(int dev) {
  if (dev < 0 || dev >= 64) throw Error(DGS_ERR_INVALID, "device ordinal out of range for the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    int cur = 0;
    DGS_HIP(hipGetDevice(&cur));
    if (cur != dev) DGS_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    DGS_HIP(e);
    g_scratch[dev] = (unsigned char*)p;
  }
  return g_scratch[dev];
}

// Layout of one forward pass of mp rows in the scratch array (256-byte aligned pieces); returns the
// bytes needed.
size_t scratch_layout(const Net& n, int mp, unsigned char* base, Pass* b) {
  const dgs_config& k = n.cfg;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    unsigned char* p = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  float* x = (float*)take((size_t)mp * k.in_dim * sizeof(float));
  b->x = x;
  b->z = (uint16_t*)take((size_t)mp * 2 * k.n_freq * sizeof(uint16_t));
  b->h[0] = (uint16_t*)take((size_t)mp * k.width * sizeof(uint16_t));
  b->h[1] = (uint16_t*)take((size_t)mp * k.width * sizeof(uint16_t));
  b->pre = (float*)take((size_t)mp * k.width * sizeof(float));
  b->y = (float*)take((size_t)mp * k.out_dim * sizeof(float));
  return off;
}

void check_batch(const Net& n, int rows) {
  if (rows <= 0 || rows > n.cfg.max_batch || rows % 64)
    throw Error(DGS_ERR_INVALID, "batch rows must be a positive multiple of 64 up to max_batch");
}

// forward pass of `rows` rows already in n.x; output in n.y
void forward(Net& n, int rows);
void forward(Net& n, int rows, const Pass& b) {
  const int F = n.cfg.n_freq, in = n.cfg.in_dim;
  const int64_t tot = (int64_t)rows * F;
  hipLaunchKernelGGL(k_features, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, n.s, rows, in, F, b.x, n.Bm, b.z);
  DGS_HIP(hipGetLastError());
  const uint16_t* a = b.z;
  for (int l = 0; l < n.L; ++l) {
    GemmArgs g{};
    g.M = rows;
    g.N = n.dout[l];
    g.K = n.din[l];
    g.A = a;
    g.sam = n.din[l];
    g.sak = 1;
    g.B = n.Pb + n.woff[l];
    g.sbn = n.din[l];
    g.sbk = 1;
    g.bias = n.P + n.boff[l];
    g.ldc = n.dout[l];
    if (l + 1 < n.L) {
      g.pre = b.prel ? (*b.prel)[l] : b.pre;
      g.outb = b.hl ? (*b.hl)[l] : b.h[l % 2];
      gemm<false, false, EPI_BIAS_ACT>(n, g);
      a = g.outb;
    } else {
      g.outf = b.y;
      gemm<false, false, EPI_BIAS>(n, g);
    }
  }
}

void forward(Net& n, int rows) {
  Pass b{};
  b.x = n.x;
  b.z = n.z;
  b.hl = &n.h;
  b.prel = &n.pre;
  b.y = n.y;
  forward(n, rows, b);
}

// loss and gradients of the rows in (n.x, n.t, n.w); loss_row filled; gradients in n.G
void backward(Net& n, int rows) {
  const int No = n.cfg.out_dim;
  hipLaunchKernelGGL(k_loss, dim3(rows), dim3(256), 0, n.s, No, n.y, n.t, n.w, n.dA, n.loss_row);
  DGS_HIP(hipGetLastError());
  uint16_t* d = n.dA;  // gradient of layer l's pre-activation, [rows][dout[l]]
  uint16_t* d2 = n.dB;
  for (int l = n.L - 1; l >= 0; --l) {
    const uint16_t* xin = l ? n.h[l - 1] : n.z;
    // dW_l = d^T xin
    GemmArgs g{};
    g.M = n.dout[l];
    g.N = n.din[l];
    g.K = rows;
    g.A = d;
    g.sam = 1;
    g.sak = n.dout[l];
    g.B = xin;
    g.sbn = 1;
    g.sbk = n.din[l];
    g.outf = n.G + n.woff[l];
    g.ldc = n.din[l];
    gemm<true, true, EPI_STORE>(n, g);
    hipLaunchKernelGGL(k_colsum, dim3((n.dout[l] + 255) / 256), dim3(256), 0, n.s, rows, n.dout[l], d,
                       n.G + n.boff[l]);
    DGS_HIP(hipGetLastError());
    if (l == 0) break;
    // d_{l-1} = (d W_l) * silu'(pre_{l-1})
    GemmArgs b{};
    b.M = rows;
    b.N = n.din[l];
    b.K = n.dout[l];
    b.A = d;
    b.sam = n.dout[l];
    b.sak = 1;
    b.B = n.Pb + n.woff[l];
    b.sbn = 1;
    b.sbk = n.din[l];
    b.pre = n.pre[l - 1];
    b.outb = d2;
    b.ldc = n.din[l];
    gemm<false, true, EPI_DACT>(n, b);
    std::swap(d, d2);
  }
}

void upload_rows(Net& n, int rows, const float* x, const float* t, const float* w) {
  DGS_HIP(hipMemcpyAsync(n.x, x, (size_t)rows * n.cfg.in_dim * sizeof(float), hipMemcpyHostToDevice, n.s));
  if (t) DGS_HIP(hipMemcpyAsync(n.t, t, (size_t)rows * n.cfg.out_dim * sizeof(float), hipMemcpyHostToDevice, n.s));
  if (w) DGS_HIP(hipMemcpyAsync(n.w, w, (size_t)rows * sizeof(float), hipMemcpyHostToDevice, n.s));
}

void refresh_bf16(Net& n) {
  hipLaunchKernelGGL(k_tobf, dim3(1024), dim3(256), 0, n.s, n.nparam, n.P, n.Pb);
  DGS_HIP(hipGetLastError());
}

void create_impl(Net& n, const dgs_config* c) {
  if (!c) throw Error(DGS_ERR_INVALID, "NULL config");
  n.cfg = *c;
  const dgs_config& k = n.cfg;
  if (k.in_dim < 1 || k.in_dim > 64 || k.n_freq < 32 || k.n_freq % 32 || k.width < 64 || k.width % 64 ||
      k.depth < 1 || k.out_dim < 64 || k.out_dim % 64 || k.max_batch < 64 || k.max_batch % 64)
    throw Error(DGS_ERR_INVALID, "unsupported network shape (width, out_dim, max_batch multiples of 64, n_freq of 32)");
  n.L = k.depth + 1;
  DGS_HIP(hipGetDevice(&n.device));
  (void)scratch_base(n.device);  // the allocation-free forward pass's scratch address on this device
  for (int l = 0; l < n.L; ++l) {
    n.din.push_back(l ? k.width : 2 * k.n_freq);
    n.dout.push_back(l + 1 < n.L ? k.width : k.out_dim);
  }
  for (int l = 0; l < n.L; ++l) {
    n.woff.push_back(n.nparam);
    n.nparam += (int64_t)n.din[l] * n.dout[l];
    n.boff.push_back(n.nparam);
    n.nparam += n.dout[l];
  }
  DGS_HIP(hipStreamCreateWithFlags(&n.s, hipStreamNonBlocking));
  n.P = n.alloc<float>(n.nparam);
  n.Pb = n.alloc<uint16_t>(n.nparam);
  n.G = n.alloc<float>(n.nparam);
  n.Am = n.alloc<float>(n.nparam);
  n.Av = n.alloc<float>(n.nparam);
  n.Bm = n.alloc<float>((size_t)k.n_freq * k.in_dim);
  const int mb = k.max_batch;
  n.x = n.alloc<float>((size_t)mb * k.in_dim);
  n.t = n.alloc<float>((size_t)mb * k.out_dim);
  n.w = n.alloc<float>(mb);
  n.y = n.alloc<float>((size_t)mb * k.out_dim);
  n.loss_row = n.alloc<float>(mb);
  n.z = n.alloc<uint16_t>((size_t)mb * 2 * k.n_freq);
  for (int l = 0; l + 1 < n.L; ++l) {
    n.h.push_back(n.alloc<uint16_t>((size_t)mb * k.width));
    n.pre.push_back(n.alloc<float>((size_t)mb * k.width));
  }
  const int dmax = std::max(k.width, k.out_dim);
  n.dA = n.alloc<uint16_t>((size_t)mb * dmax);
  n.dB = n.alloc<uint16_t>((size_t)mb * dmax);
  // initial parameters: He-normal weights, zero biases; Fourier matrix N(0, sigma^2)
  std::mt19937_64 rng(k.seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  std::vector<float> hp(n.nparam, 0.0f);
  for (int l = 0; l < n.L; ++l) {
    const double sd = std::sqrt(2.0 / n.din[l]) * (l + 1 < n.L ? 1.0 : 0.5);
    for (int64_t i = 0; i < (int64_t)n.din[l] * n.dout[l]; ++i) hp[n.woff[l] + i] = (float)(sd * nd(rng));
  }
  std::vector<float> hb((size_t)k.n_freq * k.in_dim);
  for (auto& v : hb) v = (float)(k.freq_sigma * nd(rng));
  DGS_HIP(hipMemcpy(n.P, hp.data(), hp.size() * sizeof(float), hipMemcpyHostToDevice));
  DGS_HIP(hipMemcpy(n.Bm, hb.data(), hb.size() * sizeof(float), hipMemcpyHostToDevice));
  DGS_HIP(hipMemset(n.Am, 0, n.nparam * sizeof(float)));
  DGS_HIP(hipMemset(n.Av, 0, n.nparam * sizeof(float)));
  DGS_HIP(hipMemset(n.G, 0, n.nparam * sizeof(float)));
  refresh_bf16(n);
  DGS_HIP(hipStreamSynchronize(n.s));
}

}  // namespace
}  // namespace dgs

using dgs::Net;

extern "C" {

struct dgs_net : Net {};

dgs_status dgs_create(const dgs_config* cfg, dgs_net** out) {
  if (!out) return DGS_ERR_INVALID;
  *out = nullptr;
  dgs_net* n = new dgs_net();
  dgs_status st = dgs::guarded(n, [&] { dgs::create_impl(*n, cfg); });
  if (st != DGS_OK) {
    delete n;
    return st;
  }
  *out = n;
  return DGS_OK;
}

dgs_status dgs_destroy(dgs_net* n) {
  delete n;
  return DGS_OK;
}

const char* dgs_last_error(const dgs_net* n) { return n ? n->err.c_str() : dgs::g_err.c_str(); }

int64_t dgs_num_params(const dgs_net* n) { return n ? n->nparam : 0; }

dgs_status dgs_get_params(dgs_net* n, float* params, float* fourier) {
  if (!n) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    DGS_HIP(hipStreamSynchronize(n->s));
    if (params) DGS_HIP(hipMemcpy(params, n->P, n->nparam * sizeof(float), hipMemcpyDeviceToHost));
    if (fourier)
      DGS_HIP(hipMemcpy(fourier, n->Bm, (size_t)n->cfg.n_freq * n->cfg.in_dim * sizeof(float), hipMemcpyDeviceToHost));
  });
}

dgs_status dgs_set_params(dgs_net* n, const float* params, const float* fourier) {
  if (!n) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    DGS_HIP(hipStreamSynchronize(n->s));
    if (params) DGS_HIP(hipMemcpy(n->P, params, n->nparam * sizeof(float), hipMemcpyHostToDevice));
    if (fourier)
      DGS_HIP(hipMemcpy(n->Bm, fourier, (size_t)n->cfg.n_freq * n->cfg.in_dim * sizeof(float), hipMemcpyHostToDevice));
    dgs::refresh_bf16(*n);
    DGS_HIP(hipMemset(n->Am, 0, n->nparam * sizeof(float)));
    DGS_HIP(hipMemset(n->Av, 0, n->nparam * sizeof(float)));
    n->adam_t = 0;
    DGS_HIP(hipStreamSynchronize(n->s));
  });
}

dgs_status dgs_forward(dgs_net* n, int rows, const float* x, float* y) {
  if (!n || rows < 0 || (rows && (!x || !y))) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    const int mb = n->cfg.max_batch, in = n->cfg.in_dim, No = n->cfg.out_dim;
    std::vector<float> xb((size_t)mb * in);
    for (int r0 = 0; r0 < rows; r0 += mb) {
      const int m = std::min(mb, rows - r0);
      const int mp = (m + 63) / 64 * 64;  // pad to whole 64-row tiles with copies of the last row
      for (int r = 0; r < mp; ++r)
        std::copy(x + (size_t)(r0 + std::min(r, m - 1)) * in, x + (size_t)(r0 + std::min(r, m - 1) + 1) * in,
                  xb.begin() + (size_t)r * in);
      dgs::upload_rows(*n, mp, xb.data(), nullptr, nullptr);
      dgs::forward(*n, mp);
      DGS_HIP(hipMemcpyAsync(y + (size_t)r0 * No, n->y, (size_t)m * No * sizeof(float), hipMemcpyDeviceToHost, n->s));
      DGS_HIP(hipStreamSynchronize(n->s));
    }
  });
}

dgs_status dgs_forward_noalloc(dgs_net* n, int rows, const float* x, float* y) {
  if (!n || rows < 0 || (rows && (!x || !y))) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    const int in = n->cfg.in_dim, No = n->cfg.out_dim;
    dgs::Pass b{};
    // rows per pass: the largest multiple of 64 (up to max_batch) whose buffers fit the scratch array
    int mb = 0;
    for (int r = 64; r <= n->cfg.max_batch; r += 64)
      if (dgs::scratch_layout(*n, r, nullptr, &b) <= dgs::kScratchBytes) mb = r;
    if (!mb)
      throw dgs::Error(DGS_ERR_INVALID, "network too wide for the 16 MiB forward scratch (64 rows do not fit)");
    unsigned char* base = dgs::scratch_base(n->device);
    for (int r0 = 0; r0 < rows; r0 += mb) {
      const int m = std::min(mb, rows - r0);
      const int mp = (m + 63) / 64 * 64;  // whole 64-row tiles, padded with copies of the last row
      dgs::scratch_layout(*n, mp, base, &b);
      hipLaunchKernelGGL(dgs::k_pad_rows, dim3((unsigned)((mp * in + 255) / 256)), dim3(256), 0, n->s,
                         x + (size_t)r0 * in, m, mp, in, (float*)b.x);
      DGS_HIP(hipGetLastError());
      dgs::forward(*n, mp, b);
      DGS_HIP(hipMemcpyAsync(y + (size_t)r0 * No, b.y, (size_t)m * No * sizeof(float), hipMemcpyDeviceToDevice, n->s));
    }
    DGS_HIP(hipStreamSynchronize(n->s));
  });
}

dgs_status dgs_forward_timed(dgs_net* n, int rows, int reps, double* ms_per_rep) {
  if (!n || reps < 1 || !ms_per_rep) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    dgs::check_batch(*n, rows);
    hipEvent_t a, b;
    DGS_HIP(hipEventCreate(&a));
    DGS_HIP(hipEventCreate(&b));
    DGS_HIP(hipEventRecord(a, n->s));
    for (int i = 0; i < reps; ++i) dgs::forward(*n, rows);
    DGS_HIP(hipEventRecord(b, n->s));
    DGS_HIP(hipEventSynchronize(b));
    float ms = 0;
    DGS_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_rep = ms / reps;
  });
}

dgs_status dgs_gradients(dgs_net* n, int rows, const float* x, const float* t, const float* w, float* grad,
                         double* loss) {
  if (!n || !x || !t || !w) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    dgs::check_batch(*n, rows);
    dgs::upload_rows(*n, rows, x, t, w);
    dgs::forward(*n, rows);
    dgs::backward(*n, rows);
    std::vector<float> lr(rows);
    DGS_HIP(hipMemcpyAsync(lr.data(), n->loss_row, rows * sizeof(float), hipMemcpyDeviceToHost, n->s));
    if (grad) DGS_HIP(hipMemcpyAsync(grad, n->G, n->nparam * sizeof(float), hipMemcpyDeviceToHost, n->s));
    DGS_HIP(hipStreamSynchronize(n->s));
    if (loss) {
      double s = 0;
      for (float v : lr) s += v;
      *loss = s;
    }
  });
}

dgs_status dgs_set_data(dgs_net* n, int64_t rows, const float* x, const float* t, const float* w) {
  if (!n || rows < 1 || !x || !t || !w) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    DGS_HIP(hipStreamSynchronize(n->s));
    for (void* p : {(void*)n->X, (void*)n->T, (void*)n->W}) {
      if (!p) continue;
      auto it = std::find(n->allocs.begin(), n->allocs.end(), p);
      if (it != n->allocs.end()) n->allocs.erase(it);
      DGS_HIP(hipFree(p));
    }
    n->X = n->alloc<float>((size_t)rows * n->cfg.in_dim);
    n->T = n->alloc<float>((size_t)rows * n->cfg.out_dim);
    n->W = n->alloc<float>(rows);
    DGS_HIP(hipMemcpy(n->X, x, (size_t)rows * n->cfg.in_dim * sizeof(float), hipMemcpyHostToDevice));
    DGS_HIP(hipMemcpy(n->T, t, (size_t)rows * n->cfg.out_dim * sizeof(float), hipMemcpyHostToDevice));
    DGS_HIP(hipMemcpy(n->W, w, rows * sizeof(float), hipMemcpyHostToDevice));
    n->ndata = rows;
  });
}

dgs_status dgs_train(dgs_net* n, int steps, int batch, float lr, uint64_t seed, float* losses) {
  if (!n || steps < 0) return DGS_ERR_INVALID;
  return dgs::guarded(n, [&] {
    dgs::check_batch(*n, batch);
    if (!n->ndata) throw dgs::Error(DGS_ERR_INVALID, "no training data (dgs_set_data)");
    if (steps == 0) return;
    // minibatch indices: each epoch is a seeded permutation of the rows
    std::mt19937_64 rng(seed);
    std::vector<int> idx((size_t)steps * batch), perm(n->ndata);
    for (int64_t i = 0; i < n->ndata; ++i) perm[i] = (int)i;
    size_t pos = perm.size();
    for (auto& v : idx) {
      if (pos == perm.size()) {
        std::shuffle(perm.begin(), perm.end(), rng);
        pos = 0;
      }
      v = perm[pos++];
    }
    int* didx = nullptr;
    float* dloss = nullptr;
    DGS_HIP(hipMalloc(&didx, idx.size() * sizeof(int)));
    DGS_HIP(hipMalloc(&dloss, steps * sizeof(float)));
    try {
      DGS_HIP(hipMemcpyAsync(didx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice, n->s));
      const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
      for (int it = 0; it < steps; ++it) {
        hipLaunchKernelGGL(dgs::k_gather, dim3(batch), dim3(256), 0, n->s, batch, n->cfg.in_dim, n->cfg.out_dim,
                           didx + (size_t)it * batch, n->X, n->T, n->W, n->x, n->t, n->w);
        DGS_HIP(hipGetLastError());
        dgs::forward(*n, batch);
        dgs::backward(*n, batch);
        hipLaunchKernelGGL(dgs::k_sum, dim3(1), dim3(256), 0, n->s, batch, n->loss_row, dloss + it);
        ++n->adam_t;
        const float c1 = 1.0f - std::pow(b1, (float)n->adam_t), c2 = 1.0f - std::pow(b2, (float)n->adam_t);
        hipLaunchKernelGGL(dgs::k_adam, dim3(1024), dim3(256), 0, n->s, n->nparam, n->P, n->Pb, n->G, n->Am, n->Av,
                           lr, b1, b2, c1, c2, eps);
        DGS_HIP(hipGetLastError());
      }
      if (losses) DGS_HIP(hipMemcpyAsync(losses, dloss, steps * sizeof(float), hipMemcpyDeviceToHost, n->s));
      DGS_HIP(hipStreamSynchronize(n->s));
    } catch (...) {
      (void)hipFree(didx);
      (void)hipFree(dloss);
      throw;
    }
    DGS_HIP(hipFree(didx));
    DGS_HIP(hipFree(dloss));
  });
}

}  // extern "C"
