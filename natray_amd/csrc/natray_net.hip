// The CG warm-start network: a 4-layer 3D convolution net on 16^3 tiles of cells, with
// hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16, FP32 accumulate) for the forward
// pass and both gradients, and the training step (masked MSE, Adam) on the GPU.
//
// Layers: conv 3^3 C0 -> 16, ReLU, conv 16 -> 16, ReLU, conv 16 -> 16, ReLU, conv 16 -> 1.
// Tiles are independent, zero-padded at their faces. Activations are NDHWC per tile: voxel
// v = 256 x + 16 y + z of the 16^3 tile, channels contiguous, stored as BF16 with 8 or 16
// channels (the 2 input channels and the 1 output channel are padded with zeros).
//
// A convolution is an implicit GEMM: rows are voxels, K runs over (tap, input channel) with
// k = tap * Cin + c (tap = 9 (dx + 1) + 3 (dy + 1) + (dz + 1)), columns are output channels.
// One wave computes the 16 voxels of a z-row against 16 output channels: lane l holds voxel
// l % 16 and the 8 consecutive k of group l / 16 of each 32-wide K step (8 channels of one
// tap: one 16-byte load), and gets rows 4 (l / 16) .. + 3 of column l % 16 back.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "natray.h"
#include "natray_net.h"

namespace natray {
namespace net {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int TILE = 16;
constexpr int VOX = TILE * TILE * TILE;  // 4096
constexpr int TAPS = 27;

__host__ __device__ constexpr int ksteps(int cin) { return (TAPS * cin + 31) / 32; }

// The neighbour of voxel (x, y, z) at tap `tap`, or -1 outside the tile.
__device__ __forceinline__ int tap_voxel(int x, int y, int z, int tap) {
  const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
  const int nx = x + dx, ny = y + dy, nz = z + dz;
  if ((unsigned)nx >= TILE || (unsigned)ny >= TILE || (unsigned)nz >= TILE) return -1;
  return (nx * TILE + ny) * TILE + nz;
}

// Y[t][v][n] = act(bias[n] + sum_k X[t][v + off(tap)][c] W[k][n]) for n < 16, with W in the
// MFMA B layout `wb` (ksteps(CIN) x 4 groups x 16 columns x 8 values, see k_pack). `yf` (FP32,
// 16 channels) or `yb` (BF16, 16 channels) receives the result.
//
// A work item is XP = 2 output x-planes of one tile (32 z-rows). Its input slab, planes
// x0 - 1 .. x0 + 2 with y and z from -1 to 16, is staged in LDS with zeros outside the tile, so
// the A fragments are plain 16-byte LDS reads at lane-constant offsets (no bounds tests); a
// staged voxel holds CIN channels padded to VS so that the 16 lanes of a row, 48 bytes apart
// for CIN = 16, fall in distinct banks. The weights stay in registers (B fragments, one per
// K step). Each wave runs two z-rows at a time, two independent MFMA chains. Work items are
// strided over the grid; the tile count is `tiles`, or min(*tiles_dev, tiles) when tiles_dev is
// given (a count only the device knows when the work is enqueued).
constexpr int XP = 2;
constexpr int SLAB = (XP + 2) * 18 * 18;

template <int CIN>
struct ConvShape {
  static constexpr int VS = CIN == 16 ? 24 : 8;  // bf16 per staged voxel
  static constexpr int KS = ksteps(CIN);
};

template <int CIN>
__global__ __launch_bounds__(256) void k_conv(const __bf16* __restrict__ X, const bf16x8* __restrict__ wb,
                                              const float* __restrict__ bias, int relu, __bf16* __restrict__ yb,
                                              float* __restrict__ yf, int tiles, const uint32_t* __restrict__ tiles_dev) {
  using Sh = ConvShape<CIN>;
  constexpr int VS = Sh::VS, KS = Sh::KS, CH = CIN / 8;  // 16-byte chunks of data per staged voxel
  constexpr int PER_THREAD = (SLAB * CH + 255) / 256;
  __shared__ __attribute__((aligned(16))) __bf16 slab[SLAB * VS];
  const int nt = tiles_dev ? (int)min(*tiles_dev, (uint32_t)tiles) : tiles;
  const int items = nt * (TILE / XP);
  if ((int)blockIdx.x >= items) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, col = lane & 15;
  bf16x8 w[KS];
  int off[KS];
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    w[s] = wb[s * 64 + lane];
    const int k = s * 32 + g * 8, tap = k / CIN, c0 = k % CIN;
    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
    // K padding (tap 27) reads a valid slab address and is replaced by zeros below.
    off[s] = tap < TAPS ? ((dx * 18 + dy) * 18 + dz) * VS + c0 : 0;
  }
  // Only the last step can hold K padding (tap 27).
  const bool last_ok = ((KS - 1) * 32 + g * 8) / CIN < TAPS;
  bf16x8 zero;
#pragma unroll
  for (int j = 0; j < 8; ++j) zero[j] = (__bf16)0.0f;
  const float bcol = bias ? bias[col] : 0.0f;
  for (int item = blockIdx.x; item < items; item += gridDim.x) {
    const int t = item / (TILE / XP), x0 = (item % (TILE / XP)) * XP;
    const __bf16* Xt = X + (size_t)t * VOX * CIN;
    // Every load of the slab is issued before the first LDS store (the padding of a staged
    // voxel beyond CIN channels is never read and stays unwritten).
    bf16x8 val[PER_THREAD];
#pragma unroll
    for (int k = 0; k < PER_THREAD; ++k) {
      const int i = threadIdx.x + 256 * k;
      const int v = i / CH, j = i % CH;
      const int px = v / 324, py = (v / 18) % 18, pz = v % 18;
      const int x = x0 - 1 + px, y = py - 1, z = pz - 1;
      val[k] = zero;
      if (i < SLAB * CH && (unsigned)x < TILE && (unsigned)y < TILE && (unsigned)z < TILE)
        val[k] = *reinterpret_cast<const bf16x8*>(Xt + (size_t)((x * TILE + y) * TILE + z) * CIN + 8 * j);
    }
    __syncthreads();  // the previous item's reads of the slab are done
#pragma unroll
    for (int k = 0; k < PER_THREAD; ++k) {
      const int i = threadIdx.x + 256 * k;
      if (i < SLAB * CH) *reinterpret_cast<bf16x8*>(slab + (i / CH) * VS + 8 * (i % CH)) = val[k];
    }
    __syncthreads();
    for (int r = wave; r < XP * TILE; r += 8) {
      f32x4 acc[2];
      int base[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int rr = r + 4 * h, xo = rr / TILE, y = rr % TILE;
        base[h] = (((xo + 1) * 18 + (y + 1)) * 18 + (col + 1)) * VS;
        acc[h] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      }
#pragma unroll
      for (int s = 0; s < KS; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          bf16x8 a = *reinterpret_cast<const bf16x8*>(slab + base[h] + off[s]);
          if (s == KS - 1 && !last_ok) a = zero;
          acc[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w[s], acc[h], 0, 0, 0);
        }
      }
      // acc[h][i] is voxel z = 4 g + i of row r + 4 h, output channel col.
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int rr = r + 4 * h, row = (x0 + rr / TILE) * TILE + rr % TILE;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float v = acc[h][i] + bcol;
          if (relu) v = v > 0.0f ? v : 0.0f;
          const size_t o = ((size_t)t * VOX + (size_t)row * TILE + 4 * g + i) * 16 + col;
          if (yf) yf[o] = v;
          if (yb) yb[o] = (__bf16)v;
        }
      }
    }
  }
}

// Packs FP32 weights W[tap][cin_real][cout_real] into the MFMA B layout of a layer with cin
// input and 16 output channels (zeros where the padding is), BF16. With `flip`, packs the
// weights of the input-gradient convolution instead: W'[tap][n][c] = W[26 - tap][c][n], whose
// input channels are the layer's outputs (cin = 16) and outputs its inputs.
__global__ void k_pack(const float* __restrict__ W, int cin_real, int cout_real, int cin, int flip, bf16x8* __restrict__ wb) {
  const int ks = ksteps(cin);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;  // (s, lane)
  if (i >= ks * 64) return;
  const int s = i / 64, lane = i % 64, g = lane >> 4, n = lane & 15;
  bf16x8 out;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = s * 32 + g * 8 + j, tap = k / cin, c = k % cin;
    float w = 0.0f;
    if (tap < TAPS) {
      if (!flip) {
        if (c < cin_real && n < cout_real) w = W[((size_t)tap * cin_real + c) * cout_real + n];
      } else {
        // c is the layer's output channel, n its input channel.
        if (c < cout_real && n < cin_real) w = W[((size_t)(TAPS - 1 - tap) * cin_real + n) * cout_real + c];
      }
    }
    out[j] = (__bf16)w;
  }
  wb[i] = out;
}

// dZ = dY where the layer's ReLU output was positive (all of dY without ReLU), as BF16.
__global__ void k_relu_back(const float* __restrict__ dY, const __bf16* __restrict__ Y, int relu, __bf16* __restrict__ dZ,
                            size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float g = dY[i];
    dZ[i] = (__bf16)((!relu || (float)Y[i] > 0.0f) ? g : 0.0f);
  }
}

// dW[tap][c][n] += sum over voxels of X[v + off(tap)][c] dZ[v][n] for one tile per workgroup
// column: rows are 16 consecutive (tap, c) of the layer's K, K of the MFMA runs over voxels.
// A: lane l holds (tap, c) row l % 16 at voxels 8 (l / 16) + j of the 32-voxel step; B: dZ at
// the same voxels for column (output channel) l % 16. Partial sums go to dW with atomics.
__global__ __launch_bounds__(256) void k_conv_wgrad(const __bf16* __restrict__ X, int cin, const __bf16* __restrict__ dZ,
                                                    float* __restrict__ dW, int cin_real, int cout_real, int tiles) {
  const int rowblocks = (TAPS * cin + 15) / 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rb = blockIdx.x % rowblocks, t = blockIdx.x / rowblocks;
  if (t >= tiles) return;
  const int g = lane >> 4, m = lane & 15;
  const int k = rb * 16 + m, tap = k / cin, c = k % cin;
  const __bf16* Xt = X + (size_t)t * VOX * cin;
  const __bf16* Zt = dZ + (size_t)t * VOX * 16;
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  // The 4096 voxels in 128 steps of 32, split over the 4 waves.
  for (int s = wave; s < VOX / 32; s += 4) {
    bf16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int v = s * 32 + g * 8 + j;
      const int x = v / 256, y = (v / 16) % 16, z = v % 16;
      const int nv = tap < TAPS ? tap_voxel(x, y, z, tap) : -1;
      a[j] = nv >= 0 ? Xt[(size_t)nv * cin + c] : (__bf16)0.0f;
      b[j] = Zt[(size_t)v * 16 + m];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  // acc[i]: row (tap, c) = rb * 16 + 4 g + i, column n = m.
  const int n = m;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int kk = rb * 16 + 4 * g + i, tp = kk / cin, cc = kk % cin;
    if (tp < TAPS && cc < cin_real && n < cout_real) atomicAdd(&dW[((size_t)tp * cin_real + cc) * cout_real + n], acc[i]);
  }
}

// db[n] += sum over voxels of dZ[v][n].
__global__ void k_bgrad(const __bf16* __restrict__ dZ, float* __restrict__ db, int cout_real, size_t voxels) {
  __shared__ float part[16];
  if (threadIdx.x < 16) part[threadIdx.x] = 0.0f;
  __syncthreads();
  const int n = threadIdx.x & 15;
  float s = 0.0f;
  for (size_t v = (size_t)blockIdx.x * (blockDim.x / 16) + threadIdx.x / 16; v < voxels; v += (size_t)gridDim.x * (blockDim.x / 16))
    s += (float)dZ[v * 16 + n];
  atomicAdd(&part[n], s);
  __syncthreads();
  if (threadIdx.x < cout_real) atomicAdd(&db[threadIdx.x], part[threadIdx.x]);
}

// Masked mean squared error of channel 0 of `pred` (FP32, 16 channels) against `target`:
// the loss goes to loss[0] (summed), dPred = 2 (pred - target) mask / count on channel 0.
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target, const float* __restrict__ mask,
                      float inv_count, float* __restrict__ dpred, float* __restrict__ loss, size_t voxels) {
  float s = 0.0f;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (size_t)gridDim.x * blockDim.x) {
    const float e = (pred[v * 16] - target[v]) * mask[v];
    s += e * e;
#pragma unroll
    for (int n = 0; n < 16; ++n) dpred[v * 16 + n] = n == 0 ? 2.0f * e * inv_count : 0.0f;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, s * inv_count);
}

// The scaled residual of a guess on each tile: e_v = x_v - (A o)_v on water voxels (diag_v > 0),
// with (A o)_v = diag_v o_v - sum of o_u over the water neighbours u of v inside the tile (bit
// 2 axis + 1 of nbm for the + side, 2 axis for the - side; axes x, y, z are voxel strides 256,
// 16, 1); x is input channel 0 (FP32, [voxel][2]) and o channel 0 of `pred`. Writes e and adds
// the mean of e^2 over the water voxels to loss[0].
__device__ __forceinline__ float tile_stencil(const float* __restrict__ f, int fstride, size_t v, uint32_t bits) {
  const int z = (int)(v & 15), y = (int)((v >> 4) & 15), x = (int)((v >> 8) & 15);
  float s = 0.0f;
  if ((bits & 1u) && x > 0) s += f[(v - 256) * fstride];
  if ((bits & 2u) && x < 15) s += f[(v + 256) * fstride];
  if ((bits & 4u) && y > 0) s += f[(v - 16) * fstride];
  if ((bits & 8u) && y < 15) s += f[(v + 16) * fstride];
  if ((bits & 16u) && z > 0) s += f[(v - 1) * fstride];
  if ((bits & 32u) && z < 15) s += f[(v + 1) * fstride];
  return s;
}
__global__ void k_res_e(const float* __restrict__ pred, const float* __restrict__ x, const float* __restrict__ diag,
                        const uint32_t* __restrict__ nbm, float inv_count, float* __restrict__ e, float* __restrict__ loss,
                        size_t voxels) {
  float s = 0.0f;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (size_t)gridDim.x * blockDim.x) {
    float ev = 0.0f;
    if (diag[v] > 0.0f) ev = x[v * 2] - (diag[v] * pred[v * 16] - tile_stencil(pred, 16, v, nbm[v]));
    e[v] = ev;
    s += ev * ev;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, s * inv_count);
}
// d loss / d o = -2 A e / count on water voxels (A is symmetric), on channel 0 of dpred.
__global__ void k_res_grad(const float* __restrict__ e, const float* __restrict__ diag, const uint32_t* __restrict__ nbm,
                           float inv_count, float* __restrict__ dpred, size_t voxels) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (size_t)gridDim.x * blockDim.x) {
    float g = 0.0f;
    if (diag[v] > 0.0f) g = -2.0f * inv_count * (diag[v] * e[v] - tile_stencil(e, 1, v, nbm[v]));
#pragma unroll
    for (int n = 0; n < 16; ++n) dpred[v * 16 + n] = n == 0 ? g : 0.0f;
  }
}

// Adam (Kingma and Ba, 2015) on n parameters with bias-corrected moments.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
}

}  // namespace net
}  // namespace natray

// ---- host side -------------------------------------------------------------------------------

namespace natray {
namespace net {

// FP32 [voxels][cin_real] -> BF16 [voxels][cin] with zero padding.
__global__ void k_to_bf16(const float* __restrict__ src, int cin_real, int cin, __bf16* __restrict__ dst, size_t voxels) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (size_t)gridDim.x * blockDim.x)
    for (int c = 0; c < cin; ++c) dst[v * cin + c] = (__bf16)(c < cin_real ? src[v * cin_real + c] : 0.0f);
}
// The first cout_real channels of FP32 [voxels][16] into [voxels][cout_real].
__global__ void k_take(const float* __restrict__ src, int cout_real, float* __restrict__ dst, size_t voxels) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (size_t)gridDim.x * blockDim.x)
    for (int c = 0; c < cout_real; ++c) dst[v * cout_real + c] = src[v * 16 + c];
}
// FP32 [voxels][cout_real] into [voxels][16], zero padded.
__global__ void k_pad16(const float* __restrict__ src, int cout_real, float* __restrict__ dst, size_t voxels) {
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < voxels; v += (size_t)gridDim.x * blockDim.x)
    for (int c = 0; c < 16; ++c) dst[v * 16 + c] = c < cout_real ? src[v * cout_real + c] : 0.0f;
}

struct Layer {
  int cin_real, cout_real, cin, relu;
  size_t w_off, b_off;  // into the parameter vector
};

// W1 27x2x16, b1 16, W2 27x16x16, b2 16, W3 27x16x16, b3 16, W4 27x16x1, b4 1.
constexpr Layer LAYERS[4] = {
    {2, 16, 8, 1, 0, 864},
    {16, 16, 16, 1, 880, 880 + 6912},
    {16, 16, 16, 1, 880 + 6928, 880 + 6928 + 6912},
    {16, 1, 16, 0, 880 + 2 * 6928, 880 + 2 * 6928 + 432},
};
static_assert(880 + 2 * 6928 + 433 == NATRAY_NET_PARAMS, "parameter layout");

}  // namespace net
}  // namespace natray

using namespace natray::net;

struct natray_net {
  uint32_t max_tiles = 0;
  float* p = nullptr;     // parameters (FP32 master copy)
  float* g = nullptr;     // their gradients
  float* m = nullptr;     // Adam moments
  float* v = nullptr;
  bf16x8* wb[4] = {};     // packed forward weights
  bf16x8* wt[4] = {};     // packed input-gradient weights (layers 2..4)
  __bf16* in = nullptr;   // [T][4096][8]
  __bf16* act[3] = {};    // [T][4096][16] outputs of layers 1..3
  float* out = nullptr;   // [T][4096][16] output of layer 4
  float* dy = nullptr;    // [T][4096][16] gradient w.r.t. a layer's output
  __bf16* dz = nullptr;   // [T][4096][16] the same through the ReLU, BF16
  float* target = nullptr;
  float* mask = nullptr;
  float* resid = nullptr;  // [T][4096] residual of the residual loss
  float* stage = nullptr;  // [T][4096][2] FP32 staging of inputs and outputs
  float* loss = nullptr;
  std::vector<void*> allocs;
  uint32_t steps = 0;
  bool packed = false;
  hipStream_t stream = nullptr;
};

namespace {

thread_local std::string g_net_err;

int nfail(int code, const std::string& msg) {
  g_net_err = msg;
  return code;
}

#define NCHK(x)                                                                                       \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) return nfail(NATRAY_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int nalloc(natray_net* n, T** ptr, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, (count ? count : 1) * sizeof(T));
  if (e != hipSuccess) return nfail(NATRAY_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  n->allocs.push_back(q);
  *ptr = static_cast<T*>(q);
  return NATRAY_OK;
}

unsigned grid_for(size_t items) { return (unsigned)std::min<size_t>((items + 255) / 256, 4096); }

int pack(natray_net* n) {
  for (int l = 0; l < 4; ++l) {
    const Layer& L = LAYERS[l];
    const int cnt = ksteps(L.cin) * 64;
    hipLaunchKernelGGL(k_pack, dim3((cnt + 255) / 256), dim3(256), 0, n->stream, n->p + L.w_off, L.cin_real, L.cout_real,
                       L.cin, 0, n->wb[l]);
    if (l > 0) {
      const int cntt = ksteps(16) * 64;
      hipLaunchKernelGGL(k_pack, dim3((cntt + 255) / 256), dim3(256), 0, n->stream, n->p + L.w_off, L.cin_real, L.cout_real,
                         16, 1, n->wt[l]);
    }
  }
  NCHK(hipGetLastError());
  n->packed = true;
  return NATRAY_OK;
}

// A convolution over `tiles` tiles: one workgroup per work item (2 x-planes), up to 2 per CU
// (the LDS slab of a 16-channel input takes 62 KB).
void launch_conv(int cin, uint32_t tiles, hipStream_t stream, const __bf16* X, const bf16x8* wb, const float* bias, int relu,
                 __bf16* yb, float* yf, const uint32_t* tiles_dev) {
  const unsigned grid = (unsigned)std::min<uint32_t>(tiles * (TILE / XP), 512u);
  if (cin == 8)
    hipLaunchKernelGGL(k_conv<8>, dim3(grid), dim3(256), 0, stream, X, wb, bias, relu, yb, yf, (int)tiles, tiles_dev);
  else
    hipLaunchKernelGGL(k_conv<16>, dim3(grid), dim3(256), 0, stream, X, wb, bias, relu, yb, yf, (int)tiles, tiles_dev);
}

int forward(natray_net* n, uint32_t tiles, const uint32_t* tiles_dev = nullptr) {
  if (!n->packed) {
    int e = pack(n);
    if (e) return e;
  }
  const __bf16* x = n->in;
  for (int l = 0; l < 4; ++l) {
    const Layer& L = LAYERS[l];
    launch_conv(L.cin, tiles, n->stream, x, n->wb[l], n->p + L.b_off, L.relu, l < 3 ? n->act[l] : nullptr,
                l == 3 ? n->out : nullptr, tiles_dev);
    if (l < 3) x = n->act[l];
  }
  NCHK(hipGetLastError());
  return NATRAY_OK;
}

// Gradients of the loss of the last forward: the masked MSE against n->target, or with
// `residual` the scaled residual with the operator in n->target (diagonal) and n->mask
// (neighbour bits). The parameters' gradients go to n->g.
int backward(natray_net* n, uint32_t tiles, float inv_count, bool residual = false) {
  const size_t vox = (size_t)tiles * VOX;
  NCHK(hipMemsetAsync(n->g, 0, NATRAY_NET_PARAMS * sizeof(float), n->stream));
  NCHK(hipMemsetAsync(n->loss, 0, sizeof(float), n->stream));
  if (!residual) {
    hipLaunchKernelGGL(k_mse, dim3(grid_for(vox)), dim3(256), 0, n->stream, n->out, n->target, n->mask, inv_count, n->dy,
                       n->loss, vox);
  } else {
    const uint32_t* bits = reinterpret_cast<const uint32_t*>(n->mask);
    hipLaunchKernelGGL(k_res_e, dim3(grid_for(vox)), dim3(256), 0, n->stream, n->out, n->stage, n->target, bits, inv_count,
                       n->resid, n->loss, vox);
    hipLaunchKernelGGL(k_res_grad, dim3(grid_for(vox)), dim3(256), 0, n->stream, n->resid, n->target, bits, inv_count, n->dy,
                       vox);
  }
  for (int l = 3; l >= 0; --l) {
    const Layer& L = LAYERS[l];
    const __bf16* y = l < 3 ? n->act[l] : nullptr;
    hipLaunchKernelGGL(k_relu_back, dim3(grid_for(vox * 16)), dim3(256), 0, n->stream, n->dy, y, L.relu, n->dz, vox * 16);
    const __bf16* xin = l > 0 ? n->act[l - 1] : n->in;
    const int rowblocks = (TAPS * L.cin + 15) / 16;
    hipLaunchKernelGGL(k_conv_wgrad, dim3(rowblocks * tiles), dim3(256), 0, n->stream, xin, L.cin, n->dz, n->g + L.w_off,
                       L.cin_real, L.cout_real, (int)tiles);
    hipLaunchKernelGGL(k_bgrad, dim3((unsigned)std::min<size_t>((vox + 15) / 16, 1024)), dim3(256), 0, n->stream, n->dz,
                       n->g + L.b_off, L.cout_real, vox);
    if (l > 0)
      launch_conv(16, tiles, n->stream, n->dz, n->wt[l], nullptr, 0, nullptr, n->dy, nullptr);
  }
  NCHK(hipGetLastError());
  return NATRAY_OK;
}

int upload_input(natray_net* n, const float* input, uint32_t tiles) {
  const size_t vox = (size_t)tiles * VOX;
  NCHK(hipMemcpyAsync(n->stage, input, vox * 2 * 4, hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(vox)), dim3(256), 0, n->stream, n->stage, 2, 8, n->in, vox);
  NCHK(hipGetLastError());
  return NATRAY_OK;
}

}  // namespace

namespace natray {
namespace net {
int forward_device(natray_net* n, uint32_t tiles, const uint32_t* tiles_dev, void* stream) {
  hipStream_t keep = n->stream;
  n->stream = static_cast<hipStream_t>(stream);
  int e = forward(n, std::min(tiles, n->max_tiles), tiles_dev);
  n->stream = keep;
  return e;
}
void* input_buffer(natray_net* n) { return n->in; }
const float* output_buffer(natray_net* n) { return n->out; }
uint32_t max_tiles(natray_net* n) { return n->max_tiles; }
}  // namespace net
}  // namespace natray

extern "C" {

const char* natray_net_last_error(void) { return g_net_err.c_str(); }

int natray_net_create(uint32_t max_tiles, uint64_t seed, natray_net** out) {
  if (!out || max_tiles == 0 || max_tiles > (1u << 16)) return nfail(NATRAY_E_ARG, "max_tiles must be 1..65536");
  natray_net* n = new natray_net();
  n->max_tiles = max_tiles;
  auto bail = [&](int code) {
    for (void* q : n->allocs) (void)hipFree(q);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return code;
  };
  const size_t vox = (size_t)max_tiles * VOX;
  int e = 0;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) return bail(nfail(NATRAY_E_HIP, "stream"));
  if ((e = nalloc(n, &n->p, NATRAY_NET_PARAMS)) || (e = nalloc(n, &n->g, NATRAY_NET_PARAMS)) ||
      (e = nalloc(n, &n->m, NATRAY_NET_PARAMS)) || (e = nalloc(n, &n->v, NATRAY_NET_PARAMS)) ||
      (e = nalloc(n, &n->in, vox * 8)) || (e = nalloc(n, &n->act[0], vox * 16)) || (e = nalloc(n, &n->act[1], vox * 16)) ||
      (e = nalloc(n, &n->act[2], vox * 16)) || (e = nalloc(n, &n->out, vox * 16)) || (e = nalloc(n, &n->dy, vox * 16)) ||
      (e = nalloc(n, &n->dz, vox * 16)) || (e = nalloc(n, &n->target, vox)) || (e = nalloc(n, &n->mask, vox)) ||
      (e = nalloc(n, &n->stage, vox * 2)) || (e = nalloc(n, &n->loss, 1)) || (e = nalloc(n, &n->resid, vox)))
    return bail(e);
  for (int l = 0; l < 4; ++l) {
    if ((e = nalloc(n, &n->wb[l], (size_t)ksteps(LAYERS[l].cin) * 64))) return bail(e);
    if (l > 0 && (e = nalloc(n, &n->wt[l], (size_t)ksteps(16) * 64))) return bail(e);
  }
  // He-uniform weights (bound sqrt(6 / fan_in)) from a splitmix64 stream; the last layer is
  // scaled by 0.1 so an untrained net predicts corrections near zero. Biases start at zero.
  std::vector<float> h(NATRAY_NET_PARAMS, 0.0f);
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < 4; ++l) {
    const Layer& L = LAYERS[l];
    const double bound = std::sqrt(6.0 / (TAPS * L.cin_real)) * (l == 3 ? 0.1 : 1.0);
    for (size_t i = 0; i < (size_t)TAPS * L.cin_real * L.cout_real; ++i) h[L.w_off + i] = (float)((2.0 * next() - 1.0) * bound);
  }
  if (hipMemcpy(n->p, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, NATRAY_NET_PARAMS * 4) != hipSuccess || hipMemset(n->v, 0, NATRAY_NET_PARAMS * 4) != hipSuccess ||
      hipMemset(n->in, 0, vox * 8 * sizeof(__bf16)) != hipSuccess)
    return bail(nfail(NATRAY_E_HIP, "initialising the network"));
  *out = n;
  return NATRAY_OK;
}

void natray_net_destroy(natray_net* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void* q : n->allocs) (void)hipFree(q);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

int natray_net_params(natray_net* n, float* out) {
  if (!n || !out) return nfail(NATRAY_E_ARG, "null argument");
  NCHK(hipStreamSynchronize(n->stream));
  NCHK(hipMemcpy(out, n->p, NATRAY_NET_PARAMS * 4, hipMemcpyDeviceToHost));
  return NATRAY_OK;
}

int natray_net_set_params(natray_net* n, const float* in) {
  if (!n || !in) return nfail(NATRAY_E_ARG, "null argument");
  NCHK(hipStreamSynchronize(n->stream));
  NCHK(hipMemcpy(n->p, in, NATRAY_NET_PARAMS * 4, hipMemcpyHostToDevice));
  n->packed = false;
  return NATRAY_OK;
}

int natray_net_grads(natray_net* n, float* out) {
  if (!n || !out) return nfail(NATRAY_E_ARG, "null argument");
  NCHK(hipStreamSynchronize(n->stream));
  NCHK(hipMemcpy(out, n->g, NATRAY_NET_PARAMS * 4, hipMemcpyDeviceToHost));
  return NATRAY_OK;
}

int natray_net_forward(natray_net* n, const float* input, uint32_t tiles, float* output) {
  if (!n || !input || !output) return nfail(NATRAY_E_ARG, "null argument");
  if (tiles == 0 || tiles > n->max_tiles) return nfail(NATRAY_E_ARG, "tiles must be 1..max_tiles");
  int e = upload_input(n, input, tiles);
  if (e) return e;
  if ((e = forward(n, tiles))) return e;
  const size_t vox = (size_t)tiles * VOX;
  hipLaunchKernelGGL(k_take, dim3(grid_for(vox)), dim3(256), 0, n->stream, n->out, 1, n->stage, vox);
  NCHK(hipMemcpyAsync(output, n->stage, vox * 4, hipMemcpyDeviceToHost, n->stream));
  NCHK(hipStreamSynchronize(n->stream));
  return NATRAY_OK;
}

int natray_net_train_step(natray_net* n, const float* input, const float* target, const float* mask, uint32_t tiles,
                          float lr, float* loss) {
  if (!n || !input || !target || !mask) return nfail(NATRAY_E_ARG, "null argument");
  if (tiles == 0 || tiles > n->max_tiles) return nfail(NATRAY_E_ARG, "tiles must be 1..max_tiles");
  const size_t vox = (size_t)tiles * VOX;
  double count = 0.0;
  for (size_t i = 0; i < vox; ++i) count += mask[i];
  int e = upload_input(n, input, tiles);
  if (e) return e;
  NCHK(hipMemcpyAsync(n->target, target, vox * 4, hipMemcpyHostToDevice, n->stream));
  NCHK(hipMemcpyAsync(n->mask, mask, vox * 4, hipMemcpyHostToDevice, n->stream));
  if ((e = forward(n, tiles))) return e;
  if ((e = backward(n, tiles, count > 0.0 ? (float)(1.0 / count) : 0.0f))) return e;
  if (lr > 0.0f) {
    n->steps += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->steps), c2 = 1.0f - std::pow(b2, (float)n->steps);
    hipLaunchKernelGGL(k_adam, dim3((NATRAY_NET_PARAMS + 255) / 256), dim3(256), 0, n->stream, n->p, n->g, n->m, n->v,
                       NATRAY_NET_PARAMS, lr, b1, b2, 1e-8f, c1, c2);
    NCHK(hipGetLastError());
    n->packed = false;
  }
  float h = 0.0f;
  NCHK(hipMemcpyAsync(&h, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
  NCHK(hipStreamSynchronize(n->stream));
  if (loss) *loss = h;
  return NATRAY_OK;
}

int natray_net_train_step_residual(natray_net* n, const float* input, const float* diag, const uint32_t* nbm,
                                   uint32_t tiles, float lr, float* loss) {
  if (!n || !input || !diag || !nbm) return nfail(NATRAY_E_ARG, "null argument");
  if (tiles == 0 || tiles > n->max_tiles) return nfail(NATRAY_E_ARG, "tiles must be 1..max_tiles");
  const size_t vox = (size_t)tiles * VOX;
  double count = 0.0;
  for (size_t i = 0; i < vox; ++i) count += diag[i] > 0.0f ? 1.0 : 0.0;
  int e = upload_input(n, input, tiles);
  if (e) return e;
  NCHK(hipMemcpyAsync(n->target, diag, vox * 4, hipMemcpyHostToDevice, n->stream));
  NCHK(hipMemcpyAsync(n->mask, nbm, vox * 4, hipMemcpyHostToDevice, n->stream));
  if ((e = forward(n, tiles))) return e;
  if ((e = backward(n, tiles, count > 0.0 ? (float)(1.0 / count) : 0.0f, true))) return e;
  if (lr > 0.0f) {
    n->steps += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)n->steps), c2 = 1.0f - std::pow(b2, (float)n->steps);
    hipLaunchKernelGGL(k_adam, dim3((NATRAY_NET_PARAMS + 255) / 256), dim3(256), 0, n->stream, n->p, n->g, n->m, n->v,
                       NATRAY_NET_PARAMS, lr, b1, b2, 1e-8f, c1, c2);
    NCHK(hipGetLastError());
    n->packed = false;
  }
  float h = 0.0f;
  NCHK(hipMemcpyAsync(&h, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
  NCHK(hipStreamSynchronize(n->stream));
  if (loss) *loss = h;
  return NATRAY_OK;
}

int natray_net_layer_check(uint32_t cin_real, uint32_t cout_real, int relu, const float* X, const float* W, const float* b,
                           const float* dY, uint32_t tiles, float* Y, float* dX, float* dW, float* db) {
  if (!X || !W || !b || !dY || !Y || !dX || !dW || !db || tiles == 0 || tiles > 64) return nfail(NATRAY_E_ARG, "bad argument");
  if (!((cin_real == 2 || cin_real == 16) && (cout_real == 16 || cout_real == 1)))
    return nfail(NATRAY_E_ARG, "layer shapes are (2 or 16) -> (16 or 1)");
  const int cin = cin_real == 2 ? 8 : 16;
  const size_t vox = (size_t)tiles * VOX, nw = (size_t)TAPS * cin_real * cout_real;
  std::vector<void*> a;
  auto al = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (hipMalloc(&q, bytes ? bytes : 16) != hipSuccess) return nullptr;
    a.push_back(q);
    return q;
  };
  auto done = [&](int code) {
    (void)hipDeviceSynchronize();
    for (void* q : a) (void)hipFree(q);
    return code;
  };
  float* dXf = (float*)al(vox * cin_real * 4);
  float* dWd = (float*)al(nw * 4);
  float* dbd = (float*)al(16 * 4);
  float* bd = (float*)al(16 * 4);
  float* stage = (float*)al(vox * 16 * 4);
  __bf16* xb = (__bf16*)al(vox * cin * 2);
  __bf16* yb = (__bf16*)al(vox * 16 * 2);
  float* yf = (float*)al(vox * 16 * 4);
  float* g = (float*)al(vox * 16 * 4);
  __bf16* dz = (__bf16*)al(vox * 16 * 2);
  float* dxf = (float*)al(vox * 16 * 4);
  bf16x8* wb = (bf16x8*)al((size_t)ksteps(cin) * 64 * 16);
  bf16x8* wt = (bf16x8*)al((size_t)ksteps(16) * 64 * 16);
  if (a.size() != 13) return done(nfail(NATRAY_E_NOMEM, "hipMalloc"));
  const unsigned grid = grid_for(vox);
  std::vector<float> bb(16, 0.0f);
  for (uint32_t i = 0; i < cout_real; ++i) bb[i] = b[i];
  if (hipMemcpy(dWd, W, nw * 4, hipMemcpyHostToDevice) || hipMemcpy(bd, bb.data(), 64, hipMemcpyHostToDevice) ||
      hipMemcpy(stage, X, vox * cin_real * 4, hipMemcpyHostToDevice))
    return done(nfail(NATRAY_E_HIP, "upload"));
  hipLaunchKernelGGL(k_to_bf16, dim3(grid), dim3(256), 0, 0, stage, (int)cin_real, cin, xb, vox);
  const int cnt = ksteps(cin) * 64, cntt = ksteps(16) * 64;
  hipLaunchKernelGGL(k_pack, dim3((cnt + 255) / 256), dim3(256), 0, 0, dWd, (int)cin_real, (int)cout_real, cin, 0, wb);
  hipLaunchKernelGGL(k_pack, dim3((cntt + 255) / 256), dim3(256), 0, 0, dWd, (int)cin_real, (int)cout_real, 16, 1, wt);
  launch_conv(cin, tiles, 0, xb, wb, bd, relu, yb, yf, nullptr);
  hipLaunchKernelGGL(k_take, dim3(grid), dim3(256), 0, 0, yf, (int)cout_real, stage, vox);
  if (hipMemcpy(Y, stage, vox * cout_real * 4, hipMemcpyDeviceToHost)) return done(nfail(NATRAY_E_HIP, "download Y"));
  // Backward from dY.
  if (hipMemcpy(stage, dY, vox * cout_real * 4, hipMemcpyHostToDevice)) return done(nfail(NATRAY_E_HIP, "upload dY"));
  hipLaunchKernelGGL(k_pad16, dim3(grid), dim3(256), 0, 0, stage, (int)cout_real, g, vox);
  hipLaunchKernelGGL(k_relu_back, dim3(grid_for(vox * 16)), dim3(256), 0, 0, g, yb, relu, dz, vox * 16);
  if (hipMemset(dWd, 0, nw * 4) || hipMemset(dbd, 0, 64)) return done(nfail(NATRAY_E_HIP, "memset"));
  hipLaunchKernelGGL(k_conv_wgrad, dim3(((TAPS * cin + 15) / 16) * tiles), dim3(256), 0, 0, xb, cin, dz, dWd, (int)cin_real,
                     (int)cout_real, (int)tiles);
  hipLaunchKernelGGL(k_bgrad, dim3((unsigned)std::min<size_t>((vox + 15) / 16, 1024)), dim3(256), 0, 0, dz, dbd,
                     (int)cout_real, vox);
  launch_conv(16, tiles, 0, dz, wt, nullptr, 0, nullptr, dxf, nullptr);
  hipLaunchKernelGGL(k_take, dim3(grid), dim3(256), 0, 0, dxf, (int)cin_real, dXf, vox);
  if (hipGetLastError() != hipSuccess) return done(nfail(NATRAY_E_HIP, "layer kernels"));
  if (hipMemcpy(dX, dXf, vox * cin_real * 4, hipMemcpyDeviceToHost) || hipMemcpy(dW, dWd, nw * 4, hipMemcpyDeviceToHost) ||
      hipMemcpy(db, dbd, cout_real * 4, hipMemcpyDeviceToHost))
    return done(nfail(NATRAY_E_HIP, "download gradients"));
  return done(NATRAY_OK);
}

}  // extern "C"
