// Learned per-block denoiser: BF16 MFMA 3D convolutions with FP32
// accumulation, forward and backward, Adam, and the write-back into the
// volume with a confidence fallback to plain fusion.  No reference
// counterpart (the reference has no learned component); the FP64 statement
// of the network is oracle/denoiser_ref.py.
//
// Network per 8^3 block (channels-last, bf16 activations):
//   X  [10][10][10][16]  tsdf_d (6), w_d (6), 4 zero channels; 1-voxel halo
//   h1 = relu(conv3x3x3(X) valid,  12->32)     [8][8][8][32]
//   h2 = relu(conv3x3x3(h1) same,  32->32)     [8][8][8][32]
//   y  = X.tsdf[centre] + conv3x3x3(h2) same, 32->6
// A convolution is an implicit GEMM: M = 512 output voxels, N = output
// channels, K = 27 taps x input channels, on gfx950's
// v_mfma_f32_16x16x32_bf16 (mfma32: a 16x16 output tile per instruction and
// wavefront, K = 32: one tap of 32 input channels, or two taps of conv1's
// 16).  Operand layout of the instruction, per lane l:
//   A[i][k]: i = l % 16, k = 8 (l / 16) + j   (8 bf16 per lane, j = 0..7)
//   B[k][n]: n = l % 16, k = 8 (l / 16) + j
//   D[i][n]: n = l % 16, i = 4 (l / 16) + r   (4 fp32 per lane, r = 0..3)
// (vhm_mfma_layout_probe_k32 checks this on the device).
//
// Forward: one workgroup of 16 wavefronts per block (4 per SIMD).  LDS holds
// the padded input grid and the padded h1 grid (ping-pong, 128 KiB); a lane
// reads 8 consecutive channels (16 bytes) of A from LDS and of B (weights,
// [tap][n][c] bf16) from L2 per instruction.
// Backward (training): per-layer kernels of 8 wavefronts, persistent over
// blocks, with the bf16 activations in HBM.  Weight gradients are GEMMs with
// K = voxels: A from a channel-major copy of the layer input (a lane's 8
// voxels are one x-row, one 16-byte LDS read), B gathered once per 32
// voxels and reused for every tap; they accumulate in each wavefront's
// registers across its blocks and are reduced in a fixed order
// (deterministic); the data gradient is the forward convolution with the
// flipped, transposed weights.
#include <cstring>

#include "vhm_internal.h"

#pragma clang fp contract(off)

namespace vhm {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

// v_mfma_f32_16x16x32_bf16: D (16 x 16 fp32) += A (16 x 32) B (32 x 16), bf16
// operands given as their bit patterns; per lane l, A[l % 16][8 (l / 16) + j]
// and B[8 (l / 16) + j][l % 16] (j = 0..7), D[4 (l / 16) + r][l % 16].
__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0, 0,
                                                 0);
}

constexpr int kDnIn = 16;      // input channels (12 used)
constexpr int kDnHid = 32;
constexpr int kDnOut = 6;
constexpr int kDnOutPad = 16;
constexpr int kDnOutPad8 = 8;  // dY channels in the layer-3 data gradient (4 taps per K = 32)
constexpr int kGrid = 10;      // padded grid edge
constexpr int kGridVox = 1000;
// master parameter offsets (oracle/denoiser_ref.py ORDER: W1 b1 W2 b2 W3 b3)
constexpr int kOffW1 = 0, kOffB1 = 27 * 12 * 32, kOffW2 = kOffB1 + 32, kOffB2 = kOffW2 + 27 * 32 * 32,
              kOffW3 = kOffB2 + 32, kOffB3 = kOffW3 + 27 * 32 * 6, kNumParams = kOffB3 + 6;  // 43270
static_assert(kNumParams == 43270, "parameter count");

VHM_HD uint16_t f2bf(float f) {  // round to nearest even (finite inputs)
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
VHM_HD float bf2f(uint16_t h) { return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16); }



namespace {

// Offset of channel c of voxel r in a channels-last padded LDS grid of C
// channels.  A 32-channel grid is kept as two planes of 16 channels (32
// bytes per voxel): an operand read takes one 16-byte chunk of 16
// consecutive voxels, which at a 64-byte voxel stride falls into the same
// LDS banks four at a time (about 1.6x the read time) and at 32 bytes does
// not; the tap's offset stays a constant added to the address.
template <int C>
__device__ __forceinline__ int gridOff(int r, int c) {
  if constexpr (C <= 16)
    return r * C + c;
  else
    return (c >> 4) * (kGridVox * 16) + r * 16 + (c & 15);
}

// Offset of weight (tap t, output n, input channel c) in a packed
// convolution weight tensor with NTOT outputs and CIN inputs: [27][NTOT][CIN],
// and for CIN = 32 [27][2][NTOT][16] (16-channel planes, 32 bytes per
// output row, the layout of the activation grids).
template <int CIN, int NTOT>
VHM_HD int wOff(int t, int n, int c) {
  if constexpr (CIN <= 16)
    return (t * NTOT + n) * CIN + c;
  else
    return ((t * 2 + (c >> 4)) * NTOT + n) * 16 + (c & 15);
}

// The 8 bf16 at base[e .. e + 8) of an LDS array whose base is 16-byte
// aligned, for any e: two aligned 16-byte reads and a funnel shift
// (v_alignbyte).  gfx950 serves a 16-byte LDS read that is not 16-byte
// aligned about 9x slower than an aligned one (measured: 16 against 139
// bytes per clock per CU), and the weight gradients' tap-shifted x-rows
// are misaligned for two taps in three.  The array needs 8 elements of
// slack after its end (read, not used).
constexpr int kLdsSlack = 8;
// Channel strides of the channel-major LDS buffers: 528 elements for 512
// voxels, 1,008 for a padded grid of 1,000.  An operand read takes 8
// voxels of each of 16 channels; a 16-byte read pattern at channel stride S
// measured (bytes per clock per CU, 16 wavefronts per CU): S = 1,024 bytes
// 31, 1,040 112, 1,056 185, 2,000 116, 2,016 191 -- conflict-free when S is
// 32 bytes off a multiple of 256.
constexpr int kChanStride = 528;
constexpr int kGridStride = 1008;
__device__ __forceinline__ v8s lds8(const uint16_t* base, int e) {
  const int a = e & ~7, s = e & 7;
  const uint4 lo = *reinterpret_cast<const uint4*>(base + a);
  const uint4 hi = *reinterpret_cast<const uint4*>(base + a + 8);
  // words 0..7 of the 32-byte window; skip (s >> 1) whole words in two
  // selects (by 1, then by 2), then (s & 1) halves in the byte shift
  // (scalar variables only: small arrays would be moved to LDS or scratch)
  const bool by1 = (s & 2) != 0, by2 = (s & 4) != 0;
  const uint32_t r = 2u * static_cast<uint32_t>(s & 1);
  const uint32_t y0 = by1 ? lo.y : lo.x, y1 = by1 ? lo.z : lo.y, y2 = by1 ? lo.w : lo.z, y3 = by1 ? hi.x : lo.w,
                 y4 = by1 ? hi.y : hi.x, y5 = by1 ? hi.z : hi.y, y6 = by1 ? hi.w : hi.z;
  const uint32_t x0 = by2 ? y2 : y0, x1 = by2 ? y3 : y1, x2 = by2 ? y4 : y2, x3 = by2 ? y5 : y3,
                 x4 = by2 ? y6 : y4;
  const uint4 o = make_uint4(__builtin_amdgcn_alignbyte(x1, x0, r), __builtin_amdgcn_alignbyte(x2, x1, r),
                             __builtin_amdgcn_alignbyte(x3, x2, r), __builtin_amdgcn_alignbyte(x4, x3, r));
  return __builtin_bit_cast(v8s, o);
}

__device__ __forceinline__ int padIndex(int m, int t) {
  // output voxel m of the 8^3 grid, tap t = kx + 3 (ky + 3 kz): index in the
  // 10^3 padded grid of the input it reads
  const int x = m & 7, y = (m >> 3) & 7, z = m >> 6;
  const int kx = t % 3, ky = (t / 3) % 3, kz = t / 9;
  return ((z + kz) * kGrid + (y + ky)) * kGrid + (x + kx);
}

// acc[i] (i < MT) += conv over the padded LDS grid `in` ([1000][CIN] bf16)
// for this wavefront's tiles: n-tile `nt` (16 output channels, weights
// W [27][NTOT][CIN]) and m-tiles mt = mt0 + i * mstride, on
// v_mfma_f32_16x16x32_bf16 (mfma32): K = 32 per instruction, one tap of 32
// input channels, for CIN = 16 two taps (k-groups 0-1: tap 2 st, 2-3: tap
// 2 st + 1) and for CIN = 8 four taps (k-group g: tap 4 st + g); taps past
// the 27th are zero operands.  A lane loads 8 consecutive channels
// (16 bytes) of A and of B.
template <int CIN, int NTOT, int MT, int NT = 1>
__device__ __forceinline__ void convTiles32(const uint16_t* in, const uint16_t* __restrict__ W, int nt, int mt0,
                                            int mstride, v4f (&acc)[NT][MT]) {
  static_assert(CIN == 8 || CIN == 16 || CIN == 32, "8, 16 or 32 input channels");
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, kg = lane >> 4;
  constexpr int kTapsPerStep = 32 / CIN;
  constexpr int kSteps = (27 + kTapsPerStep - 1) / kTapsPerStep;
  int m[MT];
#pragma unroll
  for (int i = 0; i < MT; ++i) m[i] = (mt0 + i * mstride) * 16 + li;
  // step st's tap and channel offset (taps past the 27th are zero operands)
  auto tapOf = [&](int st, int& t, int& c0) {
    bool live = true;
    if constexpr (CIN == 32) {
      t = st;
      c0 = kg * 8;
    } else if constexpr (CIN == 16) {
      t = 2 * st + (kg >> 1);
      c0 = (kg & 1) * 8;
      live = t < 27;
      t = live ? t : 26;
    } else {
      t = 4 * st + kg;
      c0 = 0;
      live = t < 27;
      t = live ? t : 26;
    }
    return live;
  };
  auto loadB = [&](int st, v8s (&b)[NT]) {
    int t, c0;
    const bool live = tapOf(st, t, c0);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const v8s v = *reinterpret_cast<const v8s*>(W + wOff<CIN, NTOT>(t, (nt + j) * 16 + li, c0));
      b[j] = live ? v : v8s{0, 0, 0, 0, 0, 0, 0, 0};
    }
  };
  // the weight fragments of step st + 1 are requested before step st's
  // products, so their latency overlaps them
  v8s b[NT];
  loadB(0, b);
  for (int st = 0; st < kSteps; ++st) {
    v8s bn[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) bn[j] = b[j];
    if (st + 1 < kSteps) loadB(st + 1, bn);
    int t, c0;
    tapOf(st, t, c0);
#pragma unroll
    for (int i = 0; i < MT; ++i) {
      const v8s av = *reinterpret_cast<const v8s*>(in + gridOff<CIN>(padIndex(m[i], t), c0));
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[j][i] = mfma32(av, b[j], acc[j][i]);
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) b[j] = bn[j];
  }
}

// Zero a padded grid's 488 halo voxels (the 8^3 interior is overwritten),
// 8 channels (16 bytes) per store.
template <int C>
__device__ void zeroHalo(uint16_t* g) {
  for (int r = threadIdx.x; r < kGridVox; r += blockDim.x) {
    const int x = r % kGrid, y = (r / kGrid) % kGrid, z = r / 100;
    if (x == 0 || y == 0 || z == 0 || x == 9 || y == 9 || z == 9)
#pragma unroll
      for (int c = 0; c < C; c += 8) *reinterpret_cast<uint4*>(g + gridOff<C>(r, c)) = make_uint4(0, 0, 0, 0);
  }
}

__device__ __forceinline__ int interiorIndex(int m) {
  const int x = m & 7, y = (m >> 3) & 7, z = m >> 6;
  return ((z + 1) * kGrid + (y + 1)) * kGrid + (x + 1);
}

// Writes relu(acc + bias) as bf16 into the padded grid (interior) and,
// optionally, to global memory [512][32]: n-tiles nt .. nt + NT - 1.
template <int MT, int NT = 1>
__device__ void storeHidden(const v4f (&acc)[NT][MT], const float* bias, int nt, int mt0, int mstride,
                            uint16_t* grid, uint16_t* gout) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < NT; ++j) {
    const int n = (nt + j) * 16 + (lane & 15);
    const float bb = bias[n];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = (mt0 + i * mstride) * 16 + (lane >> 4) * 4 + r;
        float v = acc[j][i][r] + bb;
        v = v > 0.0f ? v : 0.0f;
        const uint16_t h = f2bf(v);
        grid[gridOff<kDnHid>(interiorIndex(m), n)] = h;
        if (gout) gout[m * kDnHid + n] = h;
      }
  }
}

// Copies the interior of a channels-last padded grid to global [512][C]
// with 16-byte loads and stores (8 channels of one voxel each).
template <int C>
__device__ void copyOutInterior(const uint16_t* grid, uint16_t* __restrict__ gout) {
  constexpr int kChunks = 512 * C / 8;
  for (int i = threadIdx.x; i < kChunks; i += blockDim.x) {
    const int m = i / (C / 8), q = i % (C / 8);
    reinterpret_cast<uint4*>(gout)[i] = *reinterpret_cast<const uint4*>(grid + gridOff<C>(interiorIndex(m), q * 8));
  }
}

}  // namespace

// ---------------------------------------------------------------- gather --

namespace {

__device__ uint32_t dnLookup(const VolumeView& V, int x, int y, int z) {
  if (!keyInRange(x, y, z)) return kInvalid;
  const uint64_t key = packKey(x, y, z);
  uint32_t slot = hashSlot(key, V.log2cap);
  const int bound = (V.mask + 1u) < static_cast<uint32_t>(kMaxProbe) ? static_cast<int>(V.mask + 1u) : kMaxProbe;
  for (int probe = 0; probe < bound; ++probe) {
    const uint64_t cur = V.keys[slot];
    if (cur == key) {
      const uint32_t p = V.vals[slot];
      return p < V.max_blocks ? p : kInvalid;
    }
    if (cur == kEmptyKey) return kInvalid;
    slot = (slot + 1u) & V.mask;
  }
  return kInvalid;
}

// Heaviest observed slot of direction d (first on ties): tsdf at the voxel
// centre = planeSignedDistance / trunc clamped to [-1, 1], weight / max.
__device__ void voxelFeature(const uint32_t* blk, int vi, int d, F3 centre, float inv_trunc, float inv_maxw,
                             float& tsdf, float& w, int& slot_out) {
  float bw = 0.0f;
  int bk = -1;
#pragma unroll
  for (int k = 0; k < kSlots; ++k) {
    const float wk = __uint_as_float(blk[poolWord(d, 0, k, vi)]);
    if (wk > bw) {
      bw = wk;
      bk = k;
    }
  }
  slot_out = bk;
  if (bk < 0) {
    tsdf = 0.0f;
    w = 0.0f;
    return;
  }
  const F3 n = unpackNormalOct(blk[poolWord(d, 3, bk, vi)]);
  const float sd = dot3(n, centre) - __uint_as_float(blk[poolWord(d, 2, bk, vi)]);
  tsdf = sclamp(sd * inv_trunc, -1.0f, 1.0f);
  w = bw * inv_maxw;
}

struct DnGeom {
  float voxel, inv_trunc, inv_maxw;
};

// X[b][1000][16] for block keys[b] (with the 26 neighbours' halo voxels).
__global__ __launch_bounds__(256) void dnGatherKernel(VolumeView V, const int32_t* __restrict__ keys, DnGeom g,
                                                      uint16_t* __restrict__ X) {
  __shared__ uint32_t src[27];
  const int b = blockIdx.x;
  const int kx = keys[3 * b], ky = keys[3 * b + 1], kz = keys[3 * b + 2];
  if (threadIdx.x < 27) {
    const int o = threadIdx.x;
    src[o] = dnLookup(V, kx + o % 3 - 1, ky + (o / 3) % 3 - 1, kz + o / 9 - 1);
  }
  __syncthreads();
  uint16_t* out = X + static_cast<size_t>(b) * kGridVox * kDnIn;
  for (int r = threadIdx.x; r < kGridVox; r += 256) {
    const int x = r % kGrid - 1, y = (r / kGrid) % kGrid - 1, z = r / 100 - 1;  // -1..8 in block voxels
    const int bxo = x < 0 ? 0 : (x > 7 ? 2 : 1), byo = y < 0 ? 0 : (y > 7 ? 2 : 1), bzo = z < 0 ? 0 : (z > 7 ? 2 : 1);
    const uint32_t p = src[bxo + 3 * (byo + 3 * bzo)];
    uint16_t ch[kDnIn];
#pragma unroll
    for (int c = 0; c < kDnIn; ++c) ch[c] = 0;
    if (p != kInvalid) {
      const int vi = (x & 7) + 8 * ((y & 7) + 8 * (z & 7));
      const F3 centre{(static_cast<float>(kx * 8 + x) + 0.5f) * g.voxel, (static_cast<float>(ky * 8 + y) + 0.5f) * g.voxel,
                      (static_cast<float>(kz * 8 + z) + 0.5f) * g.voxel};
      const uint32_t* blk = V.pool + static_cast<size_t>(p) * kBlockWords;
#pragma unroll
      for (int d = 0; d < kDirs; ++d) {
        float ts, w;
        int s;
        voxelFeature(blk, vi, d, centre, g.inv_trunc, g.inv_maxw, ts, w, s);
        ch[d] = f2bf(ts);
        ch[6 + d] = f2bf(w);
      }
    }
#pragma unroll
    for (int c = 0; c < kDnIn; ++c) out[r * kDnIn + c] = ch[c];
  }
}

// Targets T[b][512][6] (tsdf) and masks M (1 where the target slot is observed).
__global__ __launch_bounds__(256) void dnTargetKernel(VolumeView V, const int32_t* __restrict__ keys, DnGeom g,
                                                      float* __restrict__ T, float* __restrict__ M) {
  const int b = blockIdx.x;
  const int kx = keys[3 * b], ky = keys[3 * b + 1], kz = keys[3 * b + 2];
  const uint32_t p = dnLookup(V, kx, ky, kz);
  for (int vi = threadIdx.x; vi < 512; vi += 256) {
    const int x = vi & 7, y = (vi >> 3) & 7, z = vi >> 6;
    const F3 centre{(static_cast<float>(kx * 8 + x) + 0.5f) * g.voxel, (static_cast<float>(ky * 8 + y) + 0.5f) * g.voxel,
                    (static_cast<float>(kz * 8 + z) + 0.5f) * g.voxel};
    for (int d = 0; d < kDirs; ++d) {
      float ts = 0.0f, w = 0.0f;
      int s = -1;
      if (p != kInvalid) voxelFeature(V.pool + static_cast<size_t>(p) * kBlockWords, vi, d, centre, g.inv_trunc,
                                      g.inv_maxw, ts, w, s);
      T[(static_cast<size_t>(b) * 512 + vi) * kDnOut + d] = ts;
      M[(static_cast<size_t>(b) * 512 + vi) * kDnOut + d] = s >= 0 ? 1.0f : 0.0f;
    }
  }
}

// ---------------------------------------------------------------- weights --

__global__ __launch_bounds__(256) void dnPackKernel(const float* __restrict__ P, uint16_t* W1t, uint16_t* W2t,
                                                    uint16_t* W3t, uint16_t* W2b, uint16_t* W3b) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  // W1t [27][32][16] from W1 (kz,ky,kx,12,32)
  if (i < 27 * 32 * 16) {
    const int c = i % 16, n = (i / 16) % 32, t = i / 512;
    W1t[i] = c < 12 ? f2bf(P[kOffW1 + (t * 12 + c) * 32 + n]) : 0;
  }
  if (i < 27 * 32 * 32) {
    const int c = i % 32, n = (i / 32) % 32, t = i / 1024;
    W2t[wOff<32, 32>(t, n, c)] = f2bf(P[kOffW2 + (t * 32 + c) * 32 + n]);
    // W2b[t'][n=c_in][k=o]: flipped tap 26 - t', input channel of the forward as output
    const int o = c, cin = n;
    W2b[wOff<32, 32>(t, cin, o)] = f2bf(P[kOffW2 + ((26 - t) * 32 + cin) * 32 + o]);
  }
  if (i < 27 * 16 * 32) {
    const int c = i % 32, n = (i / 32) % 16, t = i / 512;
    W3t[wOff<32, 16>(t, n, c)] = n < 6 ? f2bf(P[kOffW3 + (t * 32 + c) * 6 + n]) : 0;
  }
  if (i < 27 * 32 * kDnOutPad8) {
    // W3b[t'][n=cin 32][k=o 8]
    const int o = i % kDnOutPad8, cin = (i / kDnOutPad8) % 32, tt = i / (32 * kDnOutPad8);
    W3b[i] = o < 6 ? f2bf(P[kOffW3 + ((26 - tt) * 32 + cin) * 6 + o]) : 0;
  }
}

// ---------------------------------------------------------------- forward --

struct DnFwdArgs {
  const uint16_t* X;
  DnWeights w;
  float* Y;          // [n][512][6] prediction (tsdf units); null in training
  // training outputs (null for inference)
  uint16_t* H1;      // [n][512][32]
  uint16_t* H2;      // [n][512][32]
  const float* T;    // [n][512][6]
  const float* M;
  float* dY;         // [n][512][6]  dL/dy
  float* loss;       // [gridDim.x] partial sums of M (y - T)^2
  float inv_count;   // 1 / max(sum M, 1)
};

// One workgroup of 8 wavefronts per block with 64 KB of LDS, so two
// workgroups share a CU and one's memory round trips and barriers overlap
// the other's products.  One padded grid holds X (16 channels), then h1,
// then h2 (32 channels as two 16-channel planes): each convolution keeps its
// output tiles in registers until every wavefront has read its input.  The
// weight fragments come from L2.  conv1 and conv2: 64 output tiles (2
// n-tiles x 32 m-tiles), wavefront w takes both n-tiles of m-tiles w + 8 i
// (each A fragment feeds two products); conv3: 32 tiles, m-tiles w + 8 i.
constexpr int kFwdThreads = 512;
constexpr int kFwdWaves = kFwdThreads / 64;

__global__ __launch_bounds__(kFwdThreads) void dnForwardKernel(DnFwdArgs a) {
  __shared__ __align__(16) uint16_t grid[kGridVox * kDnHid];  // X, then h1, then h2
  __shared__ float red[kFwdWaves];
  const int b = blockIdx.x;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint16_t* Xb = a.X + static_cast<size_t>(b) * kGridVox * kDnIn;
  {  // X -> grid as [1000][16] (16-byte chunks; a lane's loads first)
    constexpr int kChunks = kGridVox * kDnIn / 8, kPer = (kChunks + kFwdThreads - 1) / kFwdThreads;
    uint4 q[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = threadIdx.x + k * kFwdThreads;
      if (i < kChunks) q[k] = reinterpret_cast<const uint4*>(Xb)[i];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = threadIdx.x + k * kFwdThreads;
      if (i < kChunks) reinterpret_cast<uint4*>(grid)[i] = q[k];
    }
  }
  __syncthreads();
  {  // conv1 (valid on X); h1 replaces X once every wavefront is done with it
    v4f acc[2][4] = {};
    convTiles32<kDnIn, 32, 4, 2>(grid, a.w.W1t, 0, wave, kFwdWaves, acc);
    __syncthreads();
    zeroHalo<kDnHid>(grid);
    storeHidden<4, 2>(acc, a.w.b + kOffB1, 0, wave, kFwdWaves, grid, nullptr);
  }
  __syncthreads();
  {  // conv2 (same on h1); h2 replaces h1 once every wavefront is done with it
    if (a.H1) copyOutInterior<kDnHid>(grid, a.H1 + static_cast<size_t>(b) * 512 * kDnHid);
    v4f acc[2][4] = {};
    convTiles32<kDnHid, 32, 4, 2>(grid, a.w.W2t, 0, wave, kFwdWaves, acc);
    __syncthreads();
    storeHidden<4, 2>(acc, a.w.b + kOffB2, 0, wave, kFwdWaves, grid, nullptr);
  }
  __syncthreads();
  // conv3: 1 n-tile (6 of 16 used) x 32 m-tiles
  float lsum = 0.0f;
  {
    if (a.H2) copyOutInterior<kDnHid>(grid, a.H2 + static_cast<size_t>(b) * 512 * kDnHid);
    v4f acc3[1][4] = {};
    convTiles32<kDnHid, kDnOutPad, 4>(grid, a.w.W3t, 0, wave, kFwdWaves, acc3);
    const v4f (&acc)[4] = acc3[0];
    const int n = lane & 15;
    if (n < kDnOut) {
      const float bb = a.w.b[kOffB3 + n];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = (wave + kFwdWaves * i) * 16 + (lane >> 4) * 4 + r;
          const float x0 = bf2f(Xb[interiorIndex(m) * kDnIn + n]);
          const float y = x0 + (acc[i][r] + bb);
          const size_t o = (static_cast<size_t>(b) * 512 + m) * kDnOut + n;
          if (a.Y) a.Y[o] = y;
          if (a.dY) {
            const float mk = a.M[o];
            const float diff = (y - a.T[o]) * mk;
            lsum += diff * (y - a.T[o]);
            a.dY[o] = 2.0f * diff * a.inv_count;
          }
        }
    }
  }
  if (a.loss) {  // workgroup partial of the loss numerator, in a fixed order
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) red[wave] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      a.loss[b] = ((red[0] + red[1]) + (red[2] + red[3])) + ((red[4] + red[5]) + (red[6] + red[7]));
    }
  }
}

// ------------------------------------------------------------- backward --
//
// Per layer one kernel of kBwdThreads lanes per workgroup, persistent over
// blocks (workgroup g takes blocks g, g + groups, ...), accumulating its
// share of the weight gradient in registers and writing one partial row;
// dnReduceKernel sums the rows in a fixed order (deterministic gradients).

constexpr int kBwdThreads = 512;
constexpr int kBwdWaves = kBwdThreads / 64;

// Weight-gradient tiles over one block, on mfma32: dW[t][c][o] += sum over
// the 512 voxels v of in[pad(v, t)][c] * dout[v][o].  A (i = c, k = v) comes
// from the channel-major padded grid inT ([CIN][1000], stride kGridStride):
// the 8 voxels of a lane's k-group are one x-row.  B (k = v, n = o) is
// gathered once per 32 voxels from dout (channels-last, [512][COUTP], or its
// padded grid's interior when kPadded) and reused for every tap the
// wavefront holds; kDoutT: dout is channel-major (channel stride
// kChanStride) instead, and a lane's 8 voxels of B are one 16-byte read.
//
// The taps are taken by rows: the three taps kx = 0, 1, 2 of row group
// g = ky + 3 kz read the same padded x-row of 10 entries at x = kx .. kx + 7,
// so one pair of aligned 16-byte reads (ldsRow3) yields all three A operands
// (one read per three taps instead of one per tap).  Work units: 18 per
// layer, unit u = (g = u >> 1, h = u & 1).  With 32 input channels h is
// the input-channel tile and the unit covers both output tiles; with 16
// input channels it is the output tile.  Wavefront w takes units w, w + 8,
// w + 16 (< 18), so its units share h = w & 1.  Every tile accumulates the
// same products in the same order as a tap-by-tap loop would.
constexpr int kDwUnits = 18;
constexpr int kDwUnitsPerWave = (kDwUnits + kBwdWaves - 1) / kBwdWaves;  // 3
template <int CIN>
constexpr int dwOutTiles() { return CIN == 32 ? 2 : 1; }  // output tiles per unit

// The three 8-element operands base[e + kx .. e + kx + 8), kx = 0, 1, 2, of an
// LDS array whose base is 16-byte aligned, for an even e: the two aligned
// 16-byte reads around them, a selection of 5 of their 8 words (e even: the
// row starts on a word), and one 2-byte funnel shift for kx = 1.  The array
// needs 8 elements of slack after its end (read, not used).
__device__ __forceinline__ void ldsRow3(const uint16_t* base, int e, v8s& f0, v8s& f1, v8s& f2) {
  const int a = e & ~7, s = e & 7;
  const uint4 lo = *reinterpret_cast<const uint4*>(base + a);
  const uint4 hi = *reinterpret_cast<const uint4*>(base + a + 8);
  // words (s >> 1) .. (s >> 1) + 4 of the 32-byte window: skip 2 words, then 1
  // (scalar variables only: small arrays would be moved to LDS or scratch)
  const bool by1 = (s & 2) != 0, by2 = (s & 4) != 0;
  const uint32_t u0 = by2 ? lo.z : lo.x, u1 = by2 ? lo.w : lo.y, u2 = by2 ? hi.x : lo.z, u3 = by2 ? hi.y : lo.w,
                 u4 = by2 ? hi.z : hi.x, u5 = by2 ? hi.w : hi.y;
  const uint32_t w0 = by1 ? u1 : u0, w1 = by1 ? u2 : u1, w2 = by1 ? u3 : u2, w3 = by1 ? u4 : u3,
                 w4 = by1 ? u5 : u4;
  f0 = __builtin_bit_cast(v8s, make_uint4(w0, w1, w2, w3));
  f1 = __builtin_bit_cast(v8s, make_uint4(__builtin_amdgcn_alignbyte(w1, w0, 2u), __builtin_amdgcn_alignbyte(w2, w1, 2u),
                                          __builtin_amdgcn_alignbyte(w3, w2, 2u), __builtin_amdgcn_alignbyte(w4, w3, 2u)));
  f2 = __builtin_bit_cast(v8s, make_uint4(w1, w2, w3, w4));
}

template <int CIN, int COUTP, bool kPadded, bool kDoutT = false>
__device__ __forceinline__ void accumulateDWRows(const uint16_t* inT, const uint16_t* dout, int wave,
                                                 v4f (&acc)[kDwUnitsPerWave][3][dwOutTiles<CIN>()]) {
  static_assert((CIN == 16 || CIN == 32) && COUTP == 32, "16 or 32 input channels, 32 output channels");
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, kg = lane >> 4;
  const int h = wave & 1;  // the h of every unit of this wavefront
  for (int kb = 0; kb < 16; ++kb) {
    const int v0 = kb * 32 + kg * 8;  // voxels v0 .. v0 + 7: x = 0..7 of row (y, z)
    const int y = (v0 >> 3) & 7, z = v0 >> 6;
    v8s bv[dwOutTiles<CIN>()];
#pragma unroll
    for (int q = 0; q < dwOutTiles<CIN>(); ++q) {
      const int ot = CIN == 32 ? q : h;
      if constexpr (kDoutT) {
        bv[q] = *reinterpret_cast<const v8s*>(dout + (ot * 16 + li) * kChanStride + v0);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int v = v0 + e;
          bv[q][e] = static_cast<short>(
              dout[kPadded ? gridOff<COUTP>(interiorIndex(v), ot * 16 + li) : v * COUTP + ot * 16 + li]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kDwUnitsPerWave; ++j) {
      const int u = wave + kBwdWaves * j;
      if (u >= kDwUnits) break;
      const int g = u >> 1, ky = g % 3, kz = g / 3;
      const int ct = CIN == 32 ? h : 0;
      v8s f0, f1, f2;
      ldsRow3(inT, (ct * 16 + li) * kGridStride + ((z + kz) * kGrid + (y + ky)) * kGrid, f0, f1, f2);
#pragma unroll
      for (int q = 0; q < dwOutTiles<CIN>(); ++q) {
        acc[j][0][q] = mfma32(f0, bv[q], acc[j][0][q]);
        acc[j][1][q] = mfma32(f1, bv[q], acc[j][1][q]);
        acc[j][2][q] = mfma32(f2, bv[q], acc[j][2][q]);
      }
    }
  }
}

template <int CIN>
__device__ __forceinline__ void zeroDWRows(v4f (&acc)[kDwUnitsPerWave][3][dwOutTiles<CIN>()]) {
#pragma unroll
  for (int j = 0; j < kDwUnitsPerWave; ++j)
#pragma unroll
    for (int kx = 0; kx < 3; ++kx)
#pragma unroll
      for (int q = 0; q < dwOutTiles<CIN>(); ++q) acc[j][kx][q] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
}

// Writes the wavefront's weight-gradient tiles into partial row P at offset
// off ([27][cin][cout] of the master layout): input channels c < cin and
// output channels o < cout are real.
template <int CIN>
__device__ __forceinline__ void storeDWRows(const v4f (&acc)[kDwUnitsPerWave][3][dwOutTiles<CIN>()], float* P,
                                            int off, int cin, int cout, int wave) {
  const int lane = threadIdx.x & 63;
  const int h = wave & 1;
#pragma unroll
  for (int j = 0; j < kDwUnitsPerWave; ++j) {
    const int u = wave + kBwdWaves * j;
    if (u >= kDwUnits) break;
    const int g = u >> 1, ct = CIN == 32 ? h : 0;
#pragma unroll
    for (int kx = 0; kx < 3; ++kx) {
      const int t = kx + 3 * g;
#pragma unroll
      for (int q = 0; q < dwOutTiles<CIN>(); ++q) {
        const int ot = CIN == 32 ? q : h;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int c = ct * 16 + (lane >> 4) * 4 + r, o = ot * 16 + (lane & 15);
          if (c < cin && o < cout) P[off + (t * cin + c) * cout + o] = acc[j][kx][q][r];
        }
      }
    }
  }
}

// Bias gradient of one block: for channel ch < C, the sum over the 512
// voxels of get(v, ch), as kBwdThreads / CS partial sums (CS = channel slots,
// a power of two >= C: lane t takes channel t % CS and voxels
// (t / CS) * (512 CS / kBwdThreads) ...) combined in order by lane ch, which
// adds it to db.  Workgroup-wide; part: kBwdThreads floats of LDS.
template <int C, int CS, class Get>
__device__ __forceinline__ void biasGrad(Get get, float* part, float& db) {
  constexpr int kParts = kBwdThreads / CS, kSpan = 512 / kParts;
  const int ch = threadIdx.x % CS, pt = threadIdx.x / CS;
  float s = 0.0f;
  if (ch < C)
    for (int q = 0; q < kSpan; ++q) s += get(pt * kSpan + q, ch);
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < C) {
    float t = 0.0f;
    for (int i = 0; i < kParts; ++i) t += part[i * CS + threadIdx.x];
    db += t;
  }
  __syncthreads();  // part is reused
}

struct DnBwdArgs {
  int n_blocks;
  const uint16_t* X;     // [n][1000][16]
  const uint16_t* H1;    // [n][512][32]
  const uint16_t* H2;
  const float* dY;       // [n][512][6]
  uint16_t* dA2;         // [n][512][32]   (b3 output, b2 input)
  uint16_t* dA1;         // [n][512][32]   (b2 output, b1 input)
  DnWeights w;
  float* partial;        // [gridDim.x][kNumParams]
};

// Copies n 16-byte chunks src[0..n) workgroup-wide: every load of a lane is
// issued before any of its LDS writes (put(i, chunk)), so a block's staging
// costs one memory round trip instead of one per chunk.
template <int kMaxPerLane, class Put>
__device__ __forceinline__ void stage16(const uint4* __restrict__ src, int n, Put put) {
  uint4 q[kMaxPerLane];
#pragma unroll
  for (int k = 0; k < kMaxPerLane; ++k) {
    const int i = threadIdx.x + k * kBwdThreads;
    if (i < n) q[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < kMaxPerLane; ++k) {
    const int i = threadIdx.x + k * kBwdThreads;
    if (i < n) put(i, q[k]);
  }
}

// The 8 bf16 of a 16-byte chunk, one per channel, at dst[e * stride]
// (e = 0..7): a transposing LDS write.
__device__ __forceinline__ void scatter8(uint16_t* dst, int stride, uint4 q) {
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    dst[(2 * e) * stride] = static_cast<uint16_t>(w[e] & 0xffffu);
    dst[(2 * e + 1) * stride] = static_cast<uint16_t>(w[e] >> 16);
  }
}

// Loads a [512][C] bf16 activation into the interior of a channel-major
// padded grid ([C][1000] at channel stride kGridStride).
template <int C>
__device__ void loadInteriorT(const uint16_t* __restrict__ src, uint16_t* gridT) {
  constexpr int kChunks = 512 * C / 8;
  stage16<(kChunks + kBwdThreads - 1) / kBwdThreads>(
      reinterpret_cast<const uint4*>(src), kChunks, [&](int i, uint4 q) {
        const int m = (i * 8) / C, c = (i * 8) % C;
        scatter8(gridT + c * kGridStride + interiorIndex(m), kGridStride, q);
      });
}

// Zeroes the halo of a channel-major padded grid ([C][1000] at stride kGridStride).
template <int C>
__device__ void zeroHaloT(uint16_t* gridT) {
  for (int r = threadIdx.x; r < kGridVox; r += kBwdThreads) {
    const int x = r % kGrid, y = (r / kGrid) % kGrid, z = r / 100;
    if (x == 0 || y == 0 || z == 0 || x == 9 || y == 9 || z == 9)
      for (int c = 0; c < C; ++c) gridT[c * kGridStride + r] = 0;
  }
}

// Loads a [512][C] bf16 activation into the interior of a channels-last
// padded grid ([1000][C]).
template <int C>
__device__ void loadInterior(const uint16_t* __restrict__ src, uint16_t* grid) {
  constexpr int kChunks = 512 * C / 8;
  stage16<(kChunks + kBwdThreads - 1) / kBwdThreads>(
      reinterpret_cast<const uint4*>(src), kChunks, [&](int i, uint4 q) {
        const int m = (i * 8) / C, c = (i * 8) % C;
        *reinterpret_cast<uint4*>(grid + gridOff<C>(interiorIndex(m), c)) = q;
      });
}

// Data gradient through a layer: g = conv(dout padded, flipped weights),
// masked by the layer input's ReLU (in > 0, read from the channel-major
// grid inT: padded [32][1000] when kMaskPadded, else 512 voxels per channel at
// stride kChanStride), written as
// bf16 [512][32].  32 m-tiles x 2 n-tiles: wavefront w takes both n-tiles
// of m-tiles w + 8 i (each A fragment feeds two products).
template <int CIN, bool kMaskPadded>
__device__ __forceinline__ void dataGrad(const uint16_t* doutPad, const uint16_t* __restrict__ Wb,
                                         const uint16_t* inT, uint16_t* __restrict__ out, int wave) {
  const int lane = threadIdx.x & 63;
  v4f g[2][4] = {};
  convTiles32<CIN, 32, 4, 2>(doutPad, Wb, 0, wave, kBwdWaves, g);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int n = j * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = (wave + kBwdWaves * i) * 16 + (lane >> 4) * 4 + r;
        const bool on = (kMaskPadded ? inT[n * kGridStride + interiorIndex(m)] : inT[n * kChanStride + m]) != 0;
        out[m * kDnHid + n] = on ? f2bf(g[j][i][r]) : 0;
      }
  }
}

// Layer 3's weight gradient with the taps in N.  dW3[t][c][o] = sum over
// the 512 voxels v of h2pad[v + k(t)][c] dY[v][o]; with w = v + k(t) - 1
// (the h2 voxel; h2pad is zero outside the interior) that is the sum over
// w of h2[w][c] dYpad[w + 2 - k(t)][o] (dYpad: dY with a zero 1-voxel
// halo).  So one GEMM per block: M = c (2 tiles), N = (t, o) (162 columns,
// 11 tiles of 16), K = w (16 steps of 32 voxels) -- 352 instructions where
// the tap-major form with N = o padded to 16 takes 864.  A[i = c][k = w]
// comes from h2T ([32][512] at stride kChanStride: a lane's 8 voxels are one x-row, one aligned
// 16-byte read); B[k = w][n = (t, o)] from the channel-major padded dY grid
// dyT ([6][1000]): the tap shifts the x-row, one 16-byte read.  Wavefront w
// takes K steps 8 (w >> 2) .. 8 (w >> 2) + 7 and N-tiles (w & 3) + 4 j
// (j < 3, tile < 11) for both M-tiles; the two K halves are summed at the
// end of the kernel (sumDW3Halves).
constexpr int kDw3Cols = 27 * kDnOut;                 // 162
constexpr int kDw3Tiles = (kDw3Cols + 15) / 16;       // 11
__device__ __forceinline__ void accumulateDW3(const uint16_t* h2T, const uint16_t* dyT, int wave,
                                              v4f (&acc)[3][2]) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, kg = lane >> 4;
  const int ng = wave & 3, kh = wave >> 2;
  int boff[3];
  bool bval[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int n = 16 * (ng + 4 * j) + li;
    bval[j] = n < kDw3Cols;
    const int tt = bval[j] ? n / kDnOut : 0, o = bval[j] ? n % kDnOut : 0;
    const int kx = tt % 3, ky = (tt / 3) % 3, kz = tt / 9;
    boff[j] = o * kGridStride + ((2 - kz) * kGrid + (2 - ky)) * kGrid + (2 - kx);
  }
  for (int kb = 8 * kh; kb < 8 * kh + 8; ++kb) {
    const int w0 = kb * 32 + kg * 8;  // voxels w0 .. w0 + 7: x = 0..7 of row (y, z)
    const int y = (w0 >> 3) & 7, z = w0 >> 6;
    const v8s a0 = *reinterpret_cast<const v8s*>(h2T + li * kChanStride + w0);
    const v8s a1 = *reinterpret_cast<const v8s*>(h2T + (16 + li) * kChanStride + w0);
    const int row = (z * kGrid + y) * kGrid;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (ng + 4 * j >= kDw3Tiles) break;  // wavefront-uniform
      v8s bv;
      bv = lds8(dyT, boff[j] + row);
      if (!bval[j]) bv = v8s{0, 0, 0, 0, 0, 0, 0, 0};
      acc[j][0] = mfma32(a0, bv, acc[j][0]);
      acc[j][1] = mfma32(a1, bv, acc[j][1]);
    }
  }
}

// Adds the K-upper-half wavefronts' (4..7) dW3 tiles into the lower half's
// (0..3), through LDS (scratch: at least 4 x 6 x 64 v4f).  Workgroup-wide.
__device__ __forceinline__ void sumDW3Halves(v4f (&acc)[3][2], int wave, v4f* scratch) {
  const int lane = threadIdx.x & 63;
  __syncthreads();
  if (wave >= 4)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) scratch[(((wave - 4) * 3 + j) * 2 + ct) * 64 + lane] = acc[j][ct];
  __syncthreads();
  if (wave < 4)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) acc[j][ct] += scratch[((wave * 3 + j) * 2 + ct) * 64 + lane];
}

// Layer 3: dW3, db3 and dA2 = conv(dY_pad, W3 flipped^T) * (h2 > 0).
__global__ __launch_bounds__(kBwdThreads) void dnBackward3Kernel(DnBwdArgs a) {
  __shared__ __align__(16) uint16_t h2T[kDnHid * kChanStride];      // h2, channel-major
  __shared__ __align__(16) uint16_t dyp[kGridVox * kDnOutPad8];      // padded dY, channels-last, 8 ch
  __shared__ __align__(16) uint16_t dyT[kDnOut * kGridStride + kLdsSlack];  // padded dY, channel-major
  __shared__ float part[kBwdThreads];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  v4f acc[3][2];
#pragma unroll
  for (int j = 0; j < 3; ++j) acc[j][0] = acc[j][1] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  float db = 0.0f;
  // halos (and dyp's channels 6, 7) stay zero: only interior channels < 6 are rewritten
  for (int i = threadIdx.x; i < kGridVox * kDnOutPad8; i += kBwdThreads) dyp[i] = 0;
  for (int i = threadIdx.x; i < kDnOut * kGridStride; i += kBwdThreads) dyT[i] = 0;
  for (int b = blockIdx.x; b < a.n_blocks; b += gridDim.x) {
    __syncthreads();
    const float* dYb = a.dY + static_cast<size_t>(b) * 512 * kDnOut;
    {  // dY [512][6] (fp32, 768 chunks of 4) -> both padded bf16 grids
      constexpr int kPer = 512 * kDnOut / 4 / kBwdThreads + 1;
      float4 q[kPer];
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = threadIdx.x + k * kBwdThreads;
        if (i < 512 * kDnOut / 4) q[k] = reinterpret_cast<const float4*>(dYb)[i];
      }
      // h2 [512][32] -> channel-major [32][512]
      stage16<512 * kDnHid / 8 / kBwdThreads>(
          reinterpret_cast<const uint4*>(a.H2 + static_cast<size_t>(b) * 512 * kDnHid), 512 * kDnHid / 8,
          [&](int i, uint4 v) {
            const int m = (i * 8) / kDnHid, c = (i * 8) % kDnHid;
            scatter8(h2T + c * kChanStride + m, kChanStride, v);
          });
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int i = threadIdx.x + k * kBwdThreads;
        if (i >= 512 * kDnOut / 4) continue;
        const float f[4] = {q[k].x, q[k].y, q[k].z, q[k].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int j = 4 * i + e, m = j / kDnOut, o = j % kDnOut;
          const uint16_t h = f2bf(f[e]);
          const int r = interiorIndex(m);
          dyp[r * kDnOutPad8 + o] = h;
          dyT[o * kGridStride + r] = h;
        }
      }
    }
    __syncthreads();
    // db3 in fp32 from dY
    biasGrad<kDnOut, 8>([&](int v, int ch) { return dYb[v * kDnOut + ch]; }, part, db);
    accumulateDW3(h2T, dyT, wave, acc);
    dataGrad<kDnOutPad8, false>(dyp, a.w.W3b, h2T, a.dA2 + static_cast<size_t>(b) * 512 * kDnHid, wave);
  }
  sumDW3Halves(acc, wave, reinterpret_cast<v4f*>(h2T));
  float* P = a.partial + static_cast<size_t>(blockIdx.x) * kNumParams;
  if (wave < 4)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int n = 16 * (wave + 4 * j) + (lane & 15);
      if (n >= kDw3Cols) continue;
      const int tt = n / kDnOut, o = n % kDnOut;
#pragma unroll
      for (int ct = 0; ct < 2; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int c = ct * 16 + (lane >> 4) * 4 + r;
          P[kOffW3 + (tt * kDnHid + c) * kDnOut + o] = acc[j][ct][r];
        }
    }
  if (threadIdx.x < kDnOut) P[kOffB3 + threadIdx.x] = db;
}

// Layer 2: dW2, db2 and dA1 = conv(dA2_pad, W2 flipped^T) * (h1 > 0).
__global__ __launch_bounds__(kBwdThreads) void dnBackward2Kernel(DnBwdArgs a) {
  __shared__ __align__(16) uint16_t h1T[kDnHid * kGridStride + kLdsSlack];  // h1, channel-major padded
  __shared__ __align__(16) uint16_t dap[kGridVox * kDnHid];  // dA2, channels-last padded
  __shared__ float part[kBwdThreads];
  const int wave = threadIdx.x >> 6;
  v4f acc[kDwUnitsPerWave][3][dwOutTiles<kDnHid>()];
  zeroDWRows<kDnHid>(acc);
  float db = 0.0f;
  zeroHaloT<kDnHid>(h1T);
  zeroHalo<kDnHid>(dap);
  // a block's h1 is fetched into registers one block ahead, so its memory
  // latency overlaps the previous block's products
  constexpr int kHPer = 512 * kDnHid / 8 / kBwdThreads;  // 4
  uint4 hq[kHPer];
  auto fetchH = [&](int b) {
    const uint4* hs = reinterpret_cast<const uint4*>(a.H1 + static_cast<size_t>(b) * 512 * kDnHid);
#pragma unroll
    for (int k = 0; k < kHPer; ++k) hq[k] = hs[threadIdx.x + k * kBwdThreads];
  };
  if (static_cast<int>(blockIdx.x) < a.n_blocks) fetchH(blockIdx.x);
  for (int b = blockIdx.x; b < a.n_blocks; b += gridDim.x) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kHPer; ++k) {
      const int i = threadIdx.x + k * kBwdThreads, m = (i * 8) / kDnHid, c = (i * 8) % kDnHid;
      scatter8(h1T + c * kGridStride + interiorIndex(m), kGridStride, hq[k]);
    }
    loadInterior<kDnHid>(a.dA2 + static_cast<size_t>(b) * 512 * kDnHid, dap);
    if (b + static_cast<int>(gridDim.x) < a.n_blocks) fetchH(b + gridDim.x);
    __syncthreads();
    biasGrad<kDnHid, 32>([&](int v, int ch) { return bf2f(dap[gridOff<kDnHid>(interiorIndex(v), ch)]); }, part, db);
    accumulateDWRows<kDnHid, kDnHid, true>(h1T, dap, wave, acc);
    dataGrad<kDnHid, true>(dap, a.w.W2b, h1T, a.dA1 + static_cast<size_t>(b) * 512 * kDnHid, wave);
  }
  float* P = a.partial + static_cast<size_t>(blockIdx.x) * kNumParams;
  storeDWRows<kDnHid>(acc, P, kOffW2, kDnHid, kDnHid, wave);
  if (threadIdx.x < kDnHid) P[kOffB2 + threadIdx.x] = db;
}

// Layer 1: dW1, db1 from X and dA1.
__global__ __launch_bounds__(kBwdThreads) void dnBackward1Kernel(DnBwdArgs a) {
  __shared__ __align__(16) uint16_t xT[kDnIn * kGridStride + kLdsSlack];  // X, channel-major (halo included)
  __shared__ __align__(16) uint16_t daT[kDnHid * kChanStride];  // dA1, channel-major
  __shared__ float part[kBwdThreads];
  const int wave = threadIdx.x >> 6;
  v4f acc[kDwUnitsPerWave][3][dwOutTiles<kDnIn>()];
  zeroDWRows<kDnIn>(acc);
  float db = 0.0f;
  // a block's inputs are fetched into registers one block ahead, so their
  // memory latency overlaps the previous block's products
  constexpr int kXChunks = kGridVox * kDnIn / 8, kDChunks = 512 * kDnHid / 8;  // 2000, 2048
  constexpr int kXPer = (kXChunks + kBwdThreads - 1) / kBwdThreads, kDPer = kDChunks / kBwdThreads;
  uint4 xq[kXPer], dq[kDPer];
  auto fetch = [&](int b) {
    const uint4* xs = reinterpret_cast<const uint4*>(a.X + static_cast<size_t>(b) * kGridVox * kDnIn);
    const uint4* ds = reinterpret_cast<const uint4*>(a.dA1 + static_cast<size_t>(b) * 512 * kDnHid);
#pragma unroll
    for (int k = 0; k < kXPer; ++k) {
      const int i = threadIdx.x + k * kBwdThreads;
      if (i < kXChunks) xq[k] = xs[i];
    }
#pragma unroll
    for (int k = 0; k < kDPer; ++k) dq[k] = ds[threadIdx.x + k * kBwdThreads];
  };
  if (static_cast<int>(blockIdx.x) < a.n_blocks) fetch(blockIdx.x);
  for (int b = blockIdx.x; b < a.n_blocks; b += gridDim.x) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kXPer; ++k) {
      const int i = threadIdx.x + k * kBwdThreads;
      if (i < kXChunks) scatter8(xT + ((i * 8) % kDnIn) * kGridStride + (i * 8) / kDnIn, kGridStride, xq[k]);
    }
#pragma unroll
    for (int k = 0; k < kDPer; ++k) {
      const int i = threadIdx.x + k * kBwdThreads;
      scatter8(daT + ((i * 8) % kDnHid) * kChanStride + (i * 8) / kDnHid, kChanStride, dq[k]);
    }
    if (b + static_cast<int>(gridDim.x) < a.n_blocks) fetch(b + gridDim.x);
    __syncthreads();
    biasGrad<kDnHid, 32>([&](int v, int ch) { return bf2f(daT[ch * kChanStride + v]); }, part, db);
    accumulateDWRows<kDnIn, kDnHid, false, true>(xT, daT, wave, acc);
  }
  float* P = a.partial + static_cast<size_t>(blockIdx.x) * kNumParams;
  storeDWRows<kDnIn>(acc, P, kOffW1, 12, kDnHid, wave);
  if (threadIdx.x < kDnHid) P[kOffB1 + threadIdx.x] = db;
}

// The packed bf16 copies of parameter i (the entries dnPackKernel writes
// for it): W1t; W2t and the flipped W2b; W3t and the flipped W3b.  Biases
// are read from P directly and have none.
__device__ __forceinline__ void packParam(int i, float p, const DnUpdate& u) {
  const uint16_t h = f2bf(p);
  if (i < kOffB1) {  // W1 (t, c < 12, n)
    const int t = i / (12 * 32), c = (i / 32) % 12, n = i % 32;
    u.W1t[(t * 32 + n) * 16 + c] = h;
  } else if (i >= kOffW2 && i < kOffB2) {  // W2 (t, c, n)
    const int j = i - kOffW2, t = j / (32 * 32), c = (j / 32) % 32, n = j % 32;
    u.W2t[wOff<32, 32>(t, n, c)] = h;
    u.W2b[wOff<32, 32>(26 - t, c, n)] = h;
  } else if (i >= kOffW3 && i < kOffB3) {  // W3 (t, c, n < 6)
    const int j = i - kOffW3, t = j / (32 * kDnOut), c = (j / kDnOut) % 32, n = j % kDnOut;
    u.W3t[wOff<32, 16>(t, n, c)] = h;
    u.W3b[((26 - t) * 32 + c) * kDnOutPad8 + n] = h;
  }
}

// grad[i] = sum over workgroups g (in order) of partial[g][i]: 4 slices per
// parameter (slice s takes rows s, s + 4, ...), combined in slice order.
// With an update, the same lane then takes Adam's step (Kingma & Ba, bias
// corrected, FP32 master parameters and moments) on parameter i and rewrites
// its packed copies, so a training step needs no separate update and pack
// launches.
constexpr int kReduceSlices = 4;
__global__ __launch_bounds__(256) void dnReduceKernel(const float* __restrict__ partial, int groups,
                                                      float* __restrict__ grad, DnUpdate u, int update) {
  __shared__ float part[kReduceSlices][64];
  const int pi = threadIdx.x & 63, sl = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + pi;
  float s = 0.0f;
  // unrolled so that 8 rows' loads are in flight together (the sum keeps
  // its order)
  if (i < kNumParams)
#pragma unroll 8
    for (int g = sl; g < groups; g += kReduceSlices) s += partial[static_cast<size_t>(g) * kNumParams + i];
  part[sl][pi] = s;
  __syncthreads();
  if (sl != 0 || i >= kNumParams) return;
  const float g = (part[0][pi] + part[1][pi]) + (part[2][pi] + part[3][pi]);
  grad[i] = g;
  if (!update) return;
  const float m = u.b1 * u.Mm[i] + (1.0f - u.b1) * g;
  const float v = u.b2 * u.Vv[i] + (1.0f - u.b2) * g * g;
  u.Mm[i] = m;
  u.Vv[i] = v;
  const float p = u.P[i] - u.lr * (m / u.c1) / (sqrtf(v / u.c2) + u.eps);
  u.P[i] = p;
  packParam(i, p, u);
}

// Write-back with confidence fallback.  For each voxel and direction whose
// heaviest slot is observed and whose input weight is at least min_w, and
// whose predicted change |y - x| is at most max_delta (in tsdf units), the
// slot's plane_offset is set so that the plane's signed distance at the
// voxel centre equals y * trunc.  Blocks whose mean |y - x| over such cells
// exceeds block_max_delta keep plain fusion entirely.
struct DnApplyArgs {
  const int32_t* keys;
  const float* Y;
  const uint16_t* X;
  DnGeom g;
  float trunc, min_w, max_delta, block_max_delta;
  unsigned long long* counters;  // [0] cells replaced, [1] blocks kept plain
};

__global__ __launch_bounds__(256) void dnApplyKernel(VolumeView V, DnApplyArgs a) {
  __shared__ float s_sum[256];
  __shared__ int s_cnt[256];
  const int b = blockIdx.x;
  const int kx = a.keys[3 * b], ky = a.keys[3 * b + 1], kz = a.keys[3 * b + 2];
  const uint32_t p = dnLookup(V, kx, ky, kz);
  float sum = 0.0f;
  int cnt = 0;
  for (int vi = threadIdx.x; vi < 512; vi += 256)
    for (int d = 0; d < kDirs; ++d) {
      const uint16_t* xr = a.X + (static_cast<size_t>(b) * kGridVox + interiorIndex(vi)) * kDnIn;
      if (bf2f(xr[6 + d]) <= 0.0f) continue;
      const float delta = fabsf(a.Y[(static_cast<size_t>(b) * 512 + vi) * kDnOut + d] - bf2f(xr[d]));
      sum += delta;
      ++cnt;
    }
  s_sum[threadIdx.x] = sum;
  s_cnt[threadIdx.x] = cnt;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
      s_cnt[threadIdx.x] += s_cnt[threadIdx.x + off];
    }
    __syncthreads();
  }
  const bool keep_plain = p == kInvalid || s_cnt[0] == 0 || s_sum[0] / static_cast<float>(s_cnt[0]) > a.block_max_delta;
  if (keep_plain) {
    if (threadIdx.x == 0) atomicAdd(&a.counters[1], 1ull);
    return;
  }
  unsigned long long replaced = 0;
  uint32_t* blk = V.pool + static_cast<size_t>(p) * kBlockWords;
  for (int vi = threadIdx.x; vi < 512; vi += 256) {
    const int x = vi & 7, y = (vi >> 3) & 7, z = vi >> 6;
    const F3 centre{(static_cast<float>(kx * 8 + x) + 0.5f) * a.g.voxel, (static_cast<float>(ky * 8 + y) + 0.5f) * a.g.voxel,
                    (static_cast<float>(kz * 8 + z) + 0.5f) * a.g.voxel};
    for (int d = 0; d < kDirs; ++d) {
      float ts, w;
      int s;
      voxelFeature(blk, vi, d, centre, a.g.inv_trunc, a.g.inv_maxw, ts, w, s);
      if (s < 0 || w < a.min_w) continue;
      const float yv = a.Y[(static_cast<size_t>(b) * 512 + vi) * kDnOut + d];
      const uint16_t* xr = a.X + (static_cast<size_t>(b) * kGridVox + interiorIndex(vi)) * kDnIn;
      if (fabsf(yv - bf2f(xr[d])) > a.max_delta) continue;
      const F3 n = unpackNormalOct(blk[poolWord(d, 3, s, vi)]);
      const float yc = sclamp(yv, -1.0f, 1.0f);
      blk[poolWord(d, 2, s, vi)] = __float_as_uint(dot3(n, centre) - yc * a.trunc);
      ++replaced;
    }
  }
  if (replaced) atomicAdd(&a.counters[0], replaced);
  if (threadIdx.x == 0) {
    V.dirty[p] = (1u << kDirs) - 1u;  // plane offsets of any direction may have changed
    uint32_t signs = 0;
    for (int d = 0; d < kDirs; ++d) signs |= signBitsBoth(d);
    atomicOr(&V.dirmask[p], signs);
  }
}

// 16x16x16 bf16 MFMA layout probe: A[i][k] = i + 16 k / 256, B = identity;
// out[lane * 4 + r] = D value held by (lane, r).
__global__ void dnMfmaProbeKernel(float* out) {
  const int lane = threadIdx.x;
  v4s a, b;
  for (int j = 0; j < 4; ++j) {
    const int i = lane & 15, k = (lane >> 4) * 4 + j;
    a[j] = static_cast<short>(f2bf(static_cast<float>(i) + static_cast<float>(k) / 16.0f));
    b[j] = static_cast<short>(f2bf((k == (lane & 15)) ? 1.0f : 0.0f));
  }
  v4f c = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = c[r];
}

// 16x16x32 bf16 MFMA (gfx950) layout probe: operands placed by the layout
// the kernels assume -- per lane l, A[i][k] and B[k][n] with i = n = l % 16
// and k = 8 (l / 16) + j for j = 0..7 -- from a (16 x 32) and b (32 x 16)
// row-major; out[4 l + r] = the D element (lane l, register r).
__global__ void dnMfma32ProbeKernel(const float* a, const float* b, float* out) {
  const int lane = threadIdx.x;
  v8s av, bv;
  for (int j = 0; j < 8; ++j) {
    const int i = lane & 15, k = (lane >> 4) * 8 + j;
    av[j] = static_cast<short>(f2bf(a[i * 32 + k]));
    bv[j] = static_cast<short>(f2bf(b[k * 16 + (lane & 15)]));
  }
  v4f c = v4f{0.0f, 0.0f, 0.0f, 0.0f};
  c = mfma32(av, bv, c);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = c[r];
}

}  // namespace

// ---------------------------------------------------------------- launchers --

hipError_t dnLaunchGather(const VolumeView& V, const int32_t* keys, uint32_t n, float voxel, float trunc,
                          float max_weight, uint16_t* X, hipStream_t s) {
  if (n == 0) return hipSuccess;
  dnGatherKernel<<<n, 256, 0, s>>>(V, keys, DnGeom{voxel, 1.0f / trunc, 1.0f / max_weight}, X);
  return hipGetLastError();
}
hipError_t dnLaunchTarget(const VolumeView& V, const int32_t* keys, uint32_t n, float voxel, float trunc,
                          float max_weight, float* T, float* M, hipStream_t s) {
  if (n == 0) return hipSuccess;
  dnTargetKernel<<<n, 256, 0, s>>>(V, keys, DnGeom{voxel, 1.0f / trunc, 1.0f / max_weight}, T, M);
  return hipGetLastError();
}
hipError_t dnLaunchPack(const float* P, uint16_t* W1t, uint16_t* W2t, uint16_t* W3t, uint16_t* W2b, uint16_t* W3b,
                        hipStream_t s) {
  dnPackKernel<<<(27 * 32 * 32 + 255) / 256, 256, 0, s>>>(P, W1t, W2t, W3t, W2b, W3b);
  return hipGetLastError();
}
hipError_t dnLaunchForward(const uint16_t* X, const DnWeights& w, uint32_t n, float* Y, uint16_t* H1, uint16_t* H2,
                           const float* T, const float* M, float* dY, float* loss, float inv_count, hipStream_t s) {
  if (n == 0) return hipSuccess;
  DnFwdArgs a{X, w, Y, H1, H2, T, M, dY, loss, inv_count};
  dnForwardKernel<<<n, kFwdThreads, 0, s>>>(a);
  return hipGetLastError();
}
hipError_t dnLaunchBackward(const uint16_t* X, const uint16_t* H1, const uint16_t* H2, const float* dY,
                            uint16_t* dA2, uint16_t* dA1, const DnWeights& w, uint32_t n, int groups, float* partial,
                            float* grad, const DnUpdate* update, int step, hipStream_t s) {
  if (n == 0) return hipSuccess;
  DnBwdArgs a{static_cast<int>(n), X, H1, H2, dY, dA2, dA1, w, partial};
  // each layer kernel writes every partial entry of its layer, so the
  // reduction never reads stale values
  dnBackward3Kernel<<<groups, kBwdThreads, 0, s>>>(a);
  dnBackward2Kernel<<<groups, kBwdThreads, 0, s>>>(a);
  dnBackward1Kernel<<<groups, kBwdThreads, 0, s>>>(a);
  DnUpdate u{};
  if (update) {
    u = *update;
    u.c1 = 1.0f - powf(u.b1, static_cast<float>(step));
    u.c2 = 1.0f - powf(u.b2, static_cast<float>(step));
  }
  dnReduceKernel<<<(kNumParams + 63) / 64, 256, 0, s>>>(partial, groups, grad, u, update ? 1 : 0);
  return hipGetLastError();
}
hipError_t dnLaunchApply(const VolumeView& V, const int32_t* keys, uint32_t n, const float* Y, const uint16_t* X,
                         float voxel, float trunc, float max_weight, float min_w, float max_delta,
                         float block_max_delta, unsigned long long* counters, hipStream_t s) {
  if (n == 0) return hipSuccess;
  DnApplyArgs a{keys, Y, X, DnGeom{voxel, 1.0f / trunc, 1.0f / max_weight}, trunc, min_w, max_delta,
                block_max_delta, counters};
  dnApplyKernel<<<n, 256, 0, s>>>(V, a);
  return hipGetLastError();
}
hipError_t dnLaunchProbe(float* out, hipStream_t s) {
  dnMfmaProbeKernel<<<1, 64, 0, s>>>(out);
  return hipGetLastError();
}
hipError_t dnLaunchProbe32(const float* a, const float* b, float* out, hipStream_t s) {
  dnMfma32ProbeKernel<<<1, 64, 0, s>>>(a, b, out);
  return hipGetLastError();
}
int dnNumParams() { return kNumParams; }

}  // namespace vhm
