// Directional multi-layer marching cubes.  Full extraction runs two passes
// (count, then emit at scanned offsets) with a deterministic output order;
// incremental extraction into device buffers and the batched step run one
// fused pass that counts, takes the item's output range with an atomic add
// and emits.
//
// Reference: VoxelHashTSDF::extractMesh / extractDirectionalMultiLayerMesh
// (voxel_hash_tsdf.h:99-107, 842-939) with selectClosestLayerForSurface
// (394-421), completeCubeValuesWithFallback (423-477), planeSignedDistance
// (336-339) and polygoniseCube (marching_cubes.h:34-94).  CUDA twin:
// countTrianglesKernel / extractTrianglesKernel launched once per direction
// with a host-built lookup table and an atomic output cursor
// (cuda_mesh_extractor.cu:553-720).
//
// One workgroup handles one (direction, block) item at a time: it finds the
// 7 positive neighbour blocks, stages the 9x9x9 corner region of that
// direction in LDS (unpacked slot normals and plane offsets, observed and
// sign bits -- each voxel's normals are unpacked once instead of once per
// cube corner), compacts the cubes whose corners may straddle iso over the
// lanes, and runs each such cube on one lane: its 24 candidate distances
// sorted in registers, greedy clusters (the centres within reach of iso in a
// per-workgroup global scratch column), and per cluster the corner values,
// completion and triangles.  Triangles of an item are written in cube order.
#include "vhm_internal.h"
#include "vhm_mc_tables.h"

#pragma clang fp contract(off)

namespace vhm {
namespace {

__device__ __constant__ McTables c_mc = makeMcTables();

constexpr int kR = 9;  // corner region edge (8 voxels + 1 from the + neighbours)
constexpr int kRegion = kR * kR * kR;
static_assert(kMeshCandWords == 8 * kSlots * 256, "one centre per (corner, slot) for each of 256 lanes");
// fused pass: per lane and cube slot (j >> 8), one word of cube | ncl << 16,
// then the 24 centre rows
constexpr size_t kFusedSetWords = (1 + 8 * kSlots) * 256;
static_assert(kMeshFusedCandWords == 2 * kFusedSetWords, "two cubes per lane");
static_assert(kMeshFusedCandWords >= kMeshCandWords, "one scratch buffer serves every mesh kernel");

// Per-voxel bits of the staged region.
constexpr uint8_t kBitAllocated = 8;   // the voxel's block is allocated
constexpr uint8_t kBitBelow = 16;      // some selectable slot's corner value may be < iso
constexpr uint8_t kBitAbove = 32;      // some selectable slot's corner value may be >= iso

struct MeshShared {
  float4 plane[kSlots][kRegion];  // slot k's unit normal (x, y, z) and plane offset (w): one 16-B read
  uint8_t num_tris[256];          // c_mc.num_tris, read once per cluster
  uint8_t bits[kRegion];       // bit k: slot k observed; kBit* above
  uint32_t src[8];        // pool index of self + 7 positive neighbours (kInvalid if absent)
  uint32_t any_dir;       // OR over the 8 source blocks' direction masks
  uint32_t scan[8];  // per-wavefront partial sums (fused pass: two halves)
  uint32_t nemit;    // fused pass: length of the emit list
  // cubes that may produce triangles, compacted so that busy lanes are dense
  uint16_t list_cube[kBlockVoxels];
  uint32_t list_start[kBlockVoxels];
  uint32_t nlist;
  uint8_t cnt[kBlockVoxels];  // fused pass: triangles per cube
  uint32_t base;              // fused pass: the item's output offset
};

__device__ uint32_t lookupPool(const VolumeView& V, int x, int y, int z) {
  if (!keyInRange(x, y, z)) return kInvalid;
  const uint64_t key = packKey(x, y, z);
  uint32_t slot = hashSlot(key, V.log2cap);
  const int bound = (V.mask + 1u) < static_cast<uint32_t>(kMaxProbe) ? static_cast<int>(V.mask + 1u) : kMaxProbe;
  for (int probe = 0; probe < bound; ++probe) {
    const uint64_t cur = V.keys[slot];
    if (cur == key) return V.vals[slot];
    if (cur == kEmptyKey) return kInvalid;
    slot = (slot + 1u) & V.mask;
  }
  return kInvalid;
}

// Stages the 9x9x9 corner region of direction `dir`.  Returns false (and
// stages nothing) when none of the 8 source blocks has an observed slot in
// that direction: then every corner is invalid and no cube yields a triangle.
//
// Besides the unpacked slots, each voxel gets its slots' plane distances at
// its own centre -- exactly the values a cube uses for that corner
// (voxel_hash_tsdf.h:925) -- and two bits saying whether any of them may lie
// below / at-or-above iso.  A cube's corner values, including the ones
// completeCubeValuesWithFallback fills in (means of valid values), are
// drawn from these, so a cube whose 8 corners never have both bits can
// produce no triangle.  The margin (1e-6 |iso| + 1e-30) exceeds the rounding
// of those means, and NaN sets both bits.
//
// src8 (optional): the 8 source blocks, already looked up (meshNeighbours),
// for an item the caller knows to have an observed source in `dir`.
__device__ bool stageRegion(MeshShared& S, const VolumeView& V, uint64_t key, uint32_t self, int dir,
                            const MeshArgs& a, const uint32_t* __restrict__ src8 = nullptr) {
  const int tid = threadIdx.x;
  if (src8) {
    if (tid < 8) S.src[tid] = src8[tid];
    if (tid == 0) {
      S.nlist = 0u;
      S.nemit = 0u;
    }
    __syncthreads();
  } else {
    if (tid == 0) {
      S.any_dir = 0u;
      S.nlist = 0u;
      S.nemit = 0u;
    }
    __syncthreads();
    if (tid < 8) {
      const int ox = tid & 1, oy = (tid >> 1) & 1, oz = tid >> 2;
      uint32_t p = tid == 0 ? self : lookupPool(V, keyX(key) + ox, keyY(key) + oy, keyZ(key) + oz);
      S.src[tid] = (p < V.max_blocks) ? p : kInvalid;
      if (p < V.max_blocks && (V.dirmask[p] & (1u << dir))) atomicOr(&S.any_dir, 1u);
    }
    __syncthreads();
    if (S.any_dir == 0u) return false;
  }
  const float margin = fabsf(a.iso) * 1e-6f + 1e-30f;
  const float below = a.iso + margin, above = a.iso - margin;
  const int bx = keyX(key) * kBlockSide, by = keyY(key) * kBlockSide, bz = keyZ(key) * kBlockSide;
  // All of a lane's loads first (its region voxels tid, tid + 256, tid + 512:
  // the voxel's 3 slots of direction dir, 48 contiguous bytes each), so that
  // they are in flight together.
  constexpr int kIters = (kRegion + 255) / 256;
  uint4 slots[kIters][kSlots];
  bool alloc[kIters];
#pragma unroll
  for (int it = 0; it < kIters; ++it) {
    const int r = tid + 256 * it;
    alloc[it] = false;
    if (r < kRegion) {
      const int x = r % kR, y = (r / kR) % kR, z = r / (kR * kR);
      const int sb = (x == 8 ? 1 : 0) | (y == 8 ? 2 : 0) | (z == 8 ? 4 : 0);
      const uint32_t p = S.src[sb];
      if (p != kInvalid) {
        alloc[it] = true;
        const int vi = (x & 7) + kBlockSide * ((y & 7) + kBlockSide * (z & 7));
        const uint4* rec = reinterpret_cast<const uint4*>(V.pool + static_cast<size_t>(p) * kBlockWords +
                                                          poolWord(dir, 0, 0, vi));
#pragma unroll
        for (int k = 0; k < kSlots; ++k) slots[it][k] = rec[k];
      }
    }
  }
#pragma unroll
  for (int it = 0; it < kIters; ++it) {
    const int r = tid + 256 * it;
    if (r >= kRegion) break;
    const int x = r % kR, y = (r / kR) % kR, z = r / (kR * kR);
    uint8_t b = 0;
    if (alloc[it]) {
      b = kBitAllocated;
      const F3 centre{(static_cast<float>(bx + x) + 0.5f) * a.voxel, (static_cast<float>(by + y) + 0.5f) * a.voxel,
                      (static_cast<float>(bz + z) + 0.5f) * a.voxel};
      // Only a slot that some cluster within reach of iso (processCube,
      // step 3) can select decides a triangle: its value at the cell centre
      // is within extract_tol of the cluster's, which is within reach
      // (extract_tol + 0.875 voxel + rounding) of iso, and its value here
      // differs from that by at most |n . (centre - cell centre)| <=
      // (sqrt(3) / 2) voxel < 0.875 voxel plus rounding.  The sign bits
      // count such slots only.
      const float near = 2.0f * a.extract_tol + 1.75f * a.voxel +
                         (fabsf(centre.x) + fabsf(centre.y) + fabsf(centre.z) + 4.0f * a.voxel) * 0x1.0p-18f;
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        const uint4 slot = slots[it][k];  // weight, surface_coord, plane_offset, packed_normal
        if (__uint_as_float(slot.x) > 0.0f) {
          b |= static_cast<uint8_t>(1u << k);
          const F3 n = unpackNormalOct(slot.w);
          const float po = __uint_as_float(slot.z);
          const float v = dot3(n, centre) - po;
          S.plane[k][r] = make_float4(n.x, n.y, n.z, po);
          if (!(fabsf(v - a.iso) > near)) {
            if (!(v >= below)) b |= kBitBelow;
            if (!(v <= above)) b |= kBitAbove;
          }
        }
      }
    }
    S.bits[r] = b;
  }
  __syncthreads();
  return true;
}

__device__ __forceinline__ float planeSD(const MeshShared& S, int k, int r, F3 p) {
  const float4 q = S.plane[k][r];
  return dot3(F3{q.x, q.y, q.z}, p) - q.w;
}

// completeCubeValuesWithFallback (voxel_hash_tsdf.h:423-477).
__device__ bool completeCube(float val[8], uint32_t& valid) {
  int n = 0;
  float sum = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (valid & (1u << c)) {
      ++n;
      sum += val[c];
    }
  if (n == 8) return true;
  if (n < 6) return false;
  const float global = n > 0 ? (sum / static_cast<float>(n)) : 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (valid & (1u << c)) continue;
    float ns = 0.0f;
    int nc = 0;
#pragma unroll
    for (int e = 0; e < 12; ++e) {
      const int c0 = kMcEdgeCorners[e][0], c1 = kMcEdgeCorners[e][1];
      int nb = -1;
      if (c0 == c)
        nb = c1;
      else if (c1 == c)
        nb = c0;
      if (nb < 0 || !(valid & (1u << nb))) continue;
      ns += val[nb];
      ++nc;
    }
    val[c] = nc > 0 ? (ns / static_cast<float>(nc)) : global;
    valid |= 1u << c;
  }
  return true;
}

__device__ __forceinline__ int cubeCase(const float val[8], float iso) {
  int ci = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (val[c] < iso) ci |= 1 << c;
  return ci;
}

// Corner c of the cube at offset ((kCornerX >> c) & 1, (kCornerY >> c) & 1,
// (kCornerZ >> c) & 1), and edge e joining corners (kEdgeC0 >> 4e) & 15 and
// (kEdgeC1 >> 4e) & 15: bit-packed forms of kMcCornerOffset / kMcEdgeCorners
// for lane-varying indices.
constexpr uint32_t packCornerAxis(int axis) {
  uint32_t m = 0;
  for (int c = 0; c < 8; ++c) m |= static_cast<uint32_t>(kMcCornerOffset[c][axis]) << c;
  return m;
}
constexpr uint64_t packEdgeCorner(int end) {
  uint64_t m = 0;
  for (int e = 0; e < 12; ++e) m |= static_cast<uint64_t>(kMcEdgeCorners[e][end]) << (4 * e);
  return m;
}
constexpr uint32_t kCornerX = packCornerAxis(0), kCornerY = packCornerAxis(1), kCornerZ = packCornerAxis(2);
constexpr uint64_t kEdgeC0 = packEdgeCorner(0), kEdgeC1 = packEdgeCorner(1);

// v[c] for a lane-varying c as a 3-level select tree on the bits of c (a
// compare chain would be folded back into an indexed access, which forces
// v out of registers).
__device__ __forceinline__ float pick8(const float v[8], int c) {
  const bool b0 = (c & 1) != 0, b1 = (c & 2) != 0, b2 = (c & 4) != 0;
  const float a01 = b0 ? v[1] : v[0], a23 = b0 ? v[3] : v[2], a45 = b0 ? v[5] : v[4], a67 = b0 ? v[7] : v[6];
  const float a03 = b1 ? a23 : a01, a47 = b1 ? a67 : a45;
  return b2 ? a47 : a03;
}

// Centre of the cube corner c of the cube at voxel (ox, oy, oz): the same
// expression as the corner positions of voxel_hash_tsdf.h:918-922.
__device__ __forceinline__ F3 cornerPos(int ox, int oy, int oz, int c, float voxel) {
  return F3{(static_cast<float>(ox + static_cast<int>((kCornerX >> c) & 1u)) + 0.5f) * voxel,
            (static_cast<float>(oy + static_cast<int>((kCornerY >> c) & 1u)) + 0.5f) * voxel,
            (static_cast<float>(oz + static_cast<int>((kCornerZ >> c) & 1u)) + 0.5f) * voxel};
}

// polygoniseCube (marching_cubes.h:48-94): writes the case's triangles.
// Each triangle vertex interpolates its edge (interpolateVertex,
// marching_cubes.h:34-46) on the spot, so no 12-edge vertex array is held.
__device__ __forceinline__ void emitCube(int ox, int oy, int oz, float voxel, const float val[8], float iso, float* out) {
  const int ci = cubeCase(val, iso);
  const int nt = c_mc.num_tris[ci];
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < 3; ++j) {
      const int e = c_mc.tri[ci][3 * i + j];
      const int c0 = static_cast<int>((kEdgeC0 >> (4 * e)) & 15u), c1 = static_cast<int>((kEdgeC1 >> (4 * e)) & 15u);
      const float v0 = pick8(val, c0), v1 = pick8(val, c1);
      const F3 p0 = cornerPos(ox, oy, oz, c0, voxel);
      F3 q = p0;
      const float delta = v1 - v0;
      if (!(fabsf(delta) < 1e-6f)) {
        const float t = sclamp((iso - v0) / delta, 0.0f, 1.0f);
        q = add3(p0, mul3(sub3(cornerPos(ox, oy, oz, c1, voxel), p0), t));
      }
      out[9 * i + 3 * j + 0] = q.x;
      out[9 * i + 3 * j + 1] = q.y;
      out[9 * i + 3 * j + 2] = q.z;
    }
}

// Whether `cube` may yield a triangle (see stageRegion): some corner may lie
// below iso and some at or above it, and at least 6 corners hold a slot that
// a cluster within reach of iso can select (a corner without one is never
// valid, and completeCubeValuesWithFallback drops a cube with fewer than 6
// valid corners).  A cube failing this yields no triangle.
__device__ __forceinline__ bool cubeMayCross(const MeshShared& S, int cube) {
  const int lx = cube & 7, ly = (cube >> 3) & 7, lz = cube >> 6;
  const int r0 = lx + kR * (ly + kR * lz);
  constexpr uint32_t kNear = kBitBelow | kBitAbove;
  uint32_t span = 0, usable = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t b = S.bits[r0 + (c & 1) + kR * ((c >> 1) & 1) + kR * kR * (c >> 2)];
    span |= b;
    usable += (b & kNear) != 0u ? 1u : 0u;
  }
  return (span & kNear) == kNear && usable >= 6u;
}

// Sorting network for the 24 candidate slots (8 corners x 3 slots): Batcher's
// odd-even merge sort of 32 inputs without the comparators that touch inputs
// 24..31 (those would hold +inf and never move).  Checked exhaustively by the
// 0-1 principle on the host when the table was generated.
constexpr int kSortPairs = 132;
constexpr int kSortPairs16 = 63;  // the prefix that sorts entries 0..15
constexpr int kSortPairs8 = 19;   // the prefix that sorts entries 0..7
constexpr uint8_t kSortNet[kSortPairs][2] = {
    {0,1}, {2,3}, {0,2}, {1,3}, {1,2}, {4,5}, {6,7}, {4,6}, {5,7}, {5,6}, {0,4}, {2,6}, {2,4}, {1,5}, {3,7}, {3,5},
    {1,2}, {3,4}, {5,6}, {8,9}, {10,11}, {8,10}, {9,11}, {9,10}, {12,13}, {14,15}, {12,14}, {13,15}, {13,14},
    {8,12}, {10,14}, {10,12}, {9,13}, {11,15}, {11,13}, {9,10}, {11,12}, {13,14}, {0,8}, {4,12}, {4,8}, {2,10},
    {6,14}, {6,10}, {2,4}, {6,8}, {10,12}, {1,9}, {5,13}, {5,9}, {3,11}, {7,15}, {7,11}, {3,5}, {7,9}, {11,13},
    {1,2}, {3,4}, {5,6}, {7,8}, {9,10}, {11,12}, {13,14}, {16,17}, {18,19}, {16,18}, {17,19}, {17,18}, {20,21},
    {22,23}, {20,22}, {21,23}, {21,22}, {16,20}, {18,22}, {18,20}, {17,21}, {19,23}, {19,21}, {17,18}, {19,20},
    {21,22}, {18,20}, {19,21}, {17,18}, {19,20}, {21,22}, {0,16}, {8,16}, {4,20}, {12,20}, {4,8}, {12,16}, {2,18},
    {10,18}, {6,22}, {14,22}, {6,10}, {14,18}, {2,4}, {6,8}, {10,12}, {14,16}, {18,20}, {1,17}, {9,17}, {5,21},
    {13,21}, {5,9}, {13,17}, {3,19}, {11,19}, {7,23}, {15,23}, {7,11}, {15,19}, {3,5}, {7,9}, {11,13}, {15,17},
    {19,21}, {1,2}, {3,4}, {5,6}, {7,8}, {9,10}, {11,12}, {13,14}, {15,16}, {17,18}, {19,20}, {21,22}};

// Compare-exchange pairs [kFrom, kTo) of kSortNet on v.
template <int kFrom, int kTo>
__device__ __forceinline__ void sortPairs(float* v) {
#pragma clang loop unroll(full)
  for (int q = kFrom; q < kTo; ++q) {
    const int i = kSortNet[q][0], j = kSortNet[q][1];
    const float x = v[i], y = v[j];
    const bool sw = y < x;
    v[i] = sw ? y : x;
    v[j] = sw ? x : y;
  }
}

// One cube of extractDirectionalMultiLayerMesh (voxel_hash_tsdf.h:853-935)
// that passed cubeMayCross.  Returns the triangle count; with kEmit also
// writes them at `out`.
//
// Three steps keep the wavefront's lanes in step: (1) the candidate plane
// distances at the cell centre, per (corner, slot) in registers (+inf for
// none), sorted by the network above (std::sort, voxel_hash_tsdf.h:887: only
// the values matter, equal values being interchangeable); (2) the greedy
// clustering (888-907), whose centres go to the lane's LDS column; (3) per
// cluster in ascending order, the closest slot of each corner, the corner
// values, completion and the triangles (909-934).
// The fused pass's count keeps, in the scratch column of a cube with at
// most kStoreMaxClusters clusters in reach (centres in rows 0..7), the
// completed corner values of its first two clusters with triangles: rows
// 16..23, then 8..15.  The emit then only triangulates them (emitStored).
constexpr int kStoreMaxClusters = 8;
__device__ __forceinline__ int storedRow(int e) { return 24 - 8 * (e + 1); }

// (3) of processCube: the triangles of the ncl clusters whose centres are
// centres[0], centres[256], ... (sc: the cube's candidates, rc: its corner
// voxels in the staged region, (ox, oy, oz): its first voxel).  kStore:
// also keep the corner values of clusters with triangles in the column
// `store` (the centres' own column, see storedRow); *nstored receives how
// many were kept, or 3 when they did not all fit.
template <bool kEmit, bool kStore = false>
__device__ __forceinline__ uint32_t clusterTriangles(const MeshShared& S, const float (&sc)[8][kSlots],
                                                     const int (&rc)[8], int ox, int oy, int oz, const MeshArgs& a,
                                                     const float* centres, int ncl, float* out,
                                                     float* store = nullptr, int* nstored = nullptr) {
  const float cx[2] = {(static_cast<float>(ox) + 0.5f) * a.voxel, (static_cast<float>(ox + 1) + 0.5f) * a.voxel};
  const float cy[2] = {(static_cast<float>(oy) + 0.5f) * a.voxel, (static_cast<float>(oy + 1) + 0.5f) * a.voxel};
  const float cz[2] = {(static_cast<float>(oz) + 0.5f) * a.voxel, (static_cast<float>(oz + 1) + 0.5f) * a.voxel};
  uint32_t ntri = 0;
  for (int j = 0; j < ncl; ++j) {
    const float cen = centres[256 * j];
    float val[8];
    uint32_t valid = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      val[c] = 0.0f;
      int best = -1;
      float best_delta = 3.402823466e+38f;
#pragma unroll
      for (int k = 0; k < kSlots; ++k) {
        const float delta = fabsf(sc[c][k] - cen);  // +inf for a missing candidate
        if (delta <= a.extract_tol && delta < best_delta) {
          best_delta = delta;
          best = k;
        }
      }
      if (best < 0) continue;
      // planeSD at the corner voxel's own centre, formed as the staging forms it
      val[c] = planeSD(S, best, rc[c], F3{cx[kMcCornerOffset[c][0]], cy[kMcCornerOffset[c][1]], cz[kMcCornerOffset[c][2]]});
      valid |= 1u << c;
    }
    if (completeCube(val, valid)) {
      const uint32_t t = S.num_tris[cubeCase(val, a.iso)];
      if (kEmit && t > 0) emitCube(ox, oy, oz, a.voxel, val, a.iso, out + 9u * ntri);
      if (kStore && t > 0) {
        if (ncl <= kStoreMaxClusters && *nstored < 2) {
          float* const row = store + 256 * storedRow(*nstored);
#pragma unroll
          for (int c = 0; c < 8; ++c) row[256 * c] = val[c];
          ++*nstored;
        } else {
          *nstored = 3;
        }
      }
      ntri += t;
    }
  }
  return ntri;
}

// Emits the triangles of a cube whose count kept the completed corner values
// of its ns (1 or 2) clusters with triangles in scratch column col.
__device__ __forceinline__ void emitStored(const MeshShared& S, uint64_t key, int cube, const MeshArgs& a,
                                           const float* __restrict__ col, int ns, float* out) {
  const int ox = keyX(key) * kBlockSide + (cube & 7), oy = keyY(key) * kBlockSide + ((cube >> 3) & 7),
            oz = keyZ(key) * kBlockSide + (cube >> 6);
  uint32_t ntri = 0;
  for (int e = 0; e < ns; ++e) {
    float val[8];
    const float* const row = col + 256 * storedRow(e);
#pragma unroll
    for (int c = 0; c < 8; ++c) val[c] = row[256 * c];
    emitCube(ox, oy, oz, a.voxel, val, a.iso, out + 9u * ntri);
    ntri += S.num_tris[cubeCase(val, a.iso)];
  }
}

// Emits the triangles of a cube whose in-reach cluster centres a count pass
// (processCube) left at centres[0], centres[256], ...: the candidates again,
// without the sort and the clustering.
__device__ __forceinline__ void emitFromCentres(const MeshShared& S, uint64_t key, int cube, const MeshArgs& a,
                                                const float* __restrict__ centres, int ncl, float* out) {
  const int lx = cube & 7, ly = (cube >> 3) & 7, lz = cube >> 6;
  const int ox = keyX(key) * kBlockSide + lx, oy = keyY(key) * kBlockSide + ly, oz = keyZ(key) * kBlockSide + lz;
  const float hv = 0.5f * a.voxel;
  const F3 cc{(static_cast<float>(ox) + 0.5f) * a.voxel + hv, (static_cast<float>(oy) + 0.5f) * a.voxel + hv,
              (static_cast<float>(oz) + 0.5f) * a.voxel + hv};
  int rc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c)
    rc[c] = (lx + kMcCornerOffset[c][0]) + kR * ((ly + kMcCornerOffset[c][1]) + kR * (lz + kMcCornerOffset[c][2]));
  float sc[8][kSlots];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t b = S.bits[rc[c]];
#pragma unroll
    for (int k = 0; k < kSlots; ++k)
      sc[c][k] = ((b & kBitAllocated) && (b & (1u << k))) ? planeSD(S, k, rc[c], cc) : __builtin_huge_valf();
  }
  clusterTriangles<true>(S, sc, rc, ox, oy, oz, a, centres, ncl, out);
}

// One cube: candidates, sort, clusters (the in-reach centres go to
// cand[256 j + lane], their number to ncl_out), and the triangles: counted,
// and written to out when kEmit; with kStore the corner values of clusters
// with triangles are kept (clusterTriangles), their number in *nstored.
template <bool kEmit, bool kStore = false>
__device__ __forceinline__ uint32_t processCube(MeshShared& S, uint64_t key, int cube, const MeshArgs& a,
                                                float* __restrict__ cand, float* out, int& ncl_out,
                                                int* nstored = nullptr) {
  const int tid = threadIdx.x;
  const int lx = cube & 7, ly = (cube >> 3) & 7, lz = cube >> 6;
  const int ox = keyX(key) * kBlockSide + lx, oy = keyY(key) * kBlockSide + ly, oz = keyZ(key) * kBlockSide + lz;
  const float hv = 0.5f * a.voxel;
  const F3 cc{(static_cast<float>(ox) + 0.5f) * a.voxel + hv, (static_cast<float>(oy) + 0.5f) * a.voxel + hv,
              (static_cast<float>(oz) + 0.5f) * a.voxel + hv};
  int rc[8];
#pragma unroll
  for (int c = 0; c < 8; ++c)
    rc[c] = (lx + kMcCornerOffset[c][0]) + kR * ((ly + kMcCornerOffset[c][1]) + kR * (lz + kMcCornerOffset[c][2]));

  // (1) candidates, slot-major: srt[8 k + c]
  constexpr float kNone = __builtin_huge_valf();
  float sc[8][kSlots];
  float srt[8 * kSlots];
  int nc = 0;
  bool slot1 = false, slot2 = false;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t b = S.bits[rc[c]];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) {
      float s = kNone;
      if ((b & kBitAllocated) && (b & (1u << k))) {
        s = planeSD(S, k, rc[c], cc);
        ++nc;
        slot1 |= k == 1;
        slot2 |= k == 2;
      }
      sc[c][k] = s;
      srt[8 * k + c] = s;
    }
  }
  ncl_out = 0;
  if (nc == 0) return 0;
  // Observed slots are contiguous from slot 0 (sortDirectionalLayers), so
  // when no lane of the wavefront has a slot-2 (or slot-1) candidate the
  // entries 16..23 (8..23) are +inf and the network's prefix that sorts
  // 0..15 (0..7) -- Batcher's construction sorts the halves first -- sorts
  // everything.
  // Each range is unrolled on its own so every index stays a constant and
  // srt stays in registers; the two branches are uniform.
  sortPairs<0, kSortPairs8>(srt);
  if (__any(slot1 || slot2)) {
    sortPairs<kSortPairs8, kSortPairs16>(srt);
    if (__any(slot2)) sortPairs<kSortPairs16, kSortPairs>(srt);
  }

  // A selected slot lies within extract_tol of its cluster's centre at the
  // cell centre, and its value at a corner centre differs from that by at
  // most |n . (corner - cell centre)| <= |n| (sqrt(3) / 2) voxel < 0.875
  // voxel (|n| = 1 up to rounding) plus the rounding of two dot products of
  // magnitude <= R (the cube's coordinate extent); a cluster whose centre
  // is farther than extract_tol + 0.875 voxel + 2^-20 R from iso has every
  // corner value (and every completion mean of them) on one side of iso and
  // yields no triangle.
  const float reach = a.extract_tol + 0.875f * a.voxel +
                      (fabsf(cc.x) + fabsf(cc.y) + fabsf(cc.z) + 2.0f * a.voxel) * 0x1.0p-20f;

  // (2) clusters: a candidate joins the open cluster unless it lies more than
  // extract_tol from its running mean.  Only centres within reach of iso are
  // kept (in ascending order, as the reference visits them).
  int ncl = 0;
  float centre = srt[0];
  int count = 1;
#pragma clang loop unroll(full)
  for (int i = 1; i < 8 * kSlots; ++i) {
    if (i < nc) {
      const float sk = srt[i];
      if (!(fabsf(sk - centre) > a.extract_tol)) {
        centre = (centre * static_cast<float>(count) + sk) / static_cast<float>(count + 1);
        ++count;
      } else {
        if (!(centre - reach > a.iso || centre + reach < a.iso)) cand[256 * ncl++ + tid] = centre;
        centre = sk;
        count = 1;
      }
    }
  }
  if (!(centre - reach > a.iso || centre + reach < a.iso)) cand[256 * ncl++ + tid] = centre;
  ncl_out = ncl;

  // (3) per cluster within reach
  return clusterTriangles<kEmit, kStore>(S, sc, rc, ox, oy, oz, a, cand + tid, ncl, out, cand + tid, nstored);
}

// Copies the per-case triangle counts into LDS (workgroup-wide, 256 lanes).
__device__ __forceinline__ void loadCaseCounts(MeshShared& S) {
  S.num_tris[threadIdx.x] = c_mc.num_tris[threadIdx.x];
  __syncthreads();
}

// Workgroup-wide exclusive scan of one value per lane (lane order); total
// receives the sum.  Each wavefront scans in registers, so the workgroup
// meets at two barriers only.
__device__ uint32_t blockExclusiveScan(MeshShared& S, uint32_t x, uint32_t& total) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t v = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(v, off);
    if (lane >= off) v += y;
  }
  if (lane == 63) S.scan[wave] = v;
  __syncthreads();
  const uint32_t s0 = S.scan[0], s1 = S.scan[1], s2 = S.scan[2], s3 = S.scan[3];
  __syncthreads();  // S.scan is rewritten by the next scan
  total = s0 + s1 + s2 + s3;
  return (wave > 0 ? s0 : 0u) + (wave > 1 ? s1 : 0u) + (wave > 2 ? s2 : 0u) + v - x;
}

// Appends `cube` (with `start`) to the compacted list when `take`; the
// whole wavefront calls it together.
__device__ __forceinline__ void listAppend(MeshShared& S, bool take, int cube, uint32_t start) {
  const int lane = threadIdx.x & 63;
  const unsigned long long mask = __ballot(take);
  uint32_t base = 0;
  if (lane == 0 && mask != 0ull) base = atomicAdd(&S.nlist, static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, 0);
  if (take) {
    const uint32_t i = base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    S.list_cube[i] = static_cast<uint16_t>(cube);
    S.list_start[i] = start;
  }
}

// Triangle count of (block b, direction dir); the workgroup takes part.
// Per-cube counts go to cube_counts[0..511] (one byte per cube: at most
// 24 clusters x 5 triangles) for the emit pass.  The cubes passing
// cubeMayCross are compacted first and then spread over the lanes.
__device__ uint32_t countItem(MeshShared& S, float* __restrict__ cand, const VolumeView& V, uint32_t b, int dir,
                              const MeshArgs& a, uint8_t* __restrict__ cube_counts, uint32_t& staged) {
  const uint64_t key = V.pool_keys[b];
  if (!stageRegion(S, V, key, b, dir, a)) {
    cube_counts[threadIdx.x] = 0;
    cube_counts[threadIdx.x + 256] = 0;
    __syncthreads();  // S.any_dir is rewritten by the next item
    return 0u;
  }
  ++staged;  // items staged, added to the header once per workgroup
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int cube = threadIdx.x + 256 * half;
    const bool take = cubeMayCross(S, cube);
    if (!take) cube_counts[cube] = 0;
    listAppend(S, take, cube, 0u);
  }
  __syncthreads();
  const uint32_t n = S.nlist;
  uint32_t mine = 0;
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    const int cube = S.list_cube[j];
    int ncl;
    const uint32_t c = processCube<false>(S, key, cube, a, cand, nullptr, ncl);
    cube_counts[cube] = static_cast<uint8_t>(c);
    mine += c;
  }
  // workgroup sum: wavefront reduction, then the 4 partial sums
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off);
  if ((threadIdx.x & 63) == 0) S.scan[threadIdx.x >> 6] = mine;
  __syncthreads();
  const uint32_t total = S.scan[0] + S.scan[1] + S.scan[2] + S.scan[3];
  __syncthreads();  // S.scan / S.nlist are reused by the next item
  return total;
}

// Emits the `total` triangles of (block b, direction dir) at tris[base ..] in
// cube order, running only the cubes the count pass found triangles in
// (compacted over the lanes); tri_owner (optional) receives owner for each
// triangle.
__device__ void emitItem(MeshShared& S, float* __restrict__ cand, const VolumeView& V, uint32_t b, int dir,
                         const MeshArgs& a, uint32_t base,
                         uint32_t total, const uint8_t* __restrict__ cube_counts, float* tris, uint32_t* tri_owner,
                         uint32_t owner, uint32_t capacity) {
  if (total == 0 || base >= capacity) return;  // workgroup-uniform: nothing of this item is written
  const uint64_t key = V.pool_keys[b];
  if (!stageRegion(S, V, key, b, dir, a)) {
    __syncthreads();
    return;
  }
  // offsets in the reference's cube order (lz, ly, lx: cubes 0..255, then 256..511)
  for (int half = 0; half < 2; ++half) {
    const int cube = threadIdx.x + 256 * half;
    const uint32_t c = cube_counts[cube];
    uint32_t sum;
    const uint32_t start = base + blockExclusiveScan(S, c, sum);
    listAppend(S, c > 0, cube, start);
    base += sum;
  }
  __syncthreads();
  const uint32_t n = S.nlist;
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    const int cube = S.list_cube[j];
    const uint32_t start = S.list_start[j];
    const uint32_t c = cube_counts[cube];
    if (start + c <= capacity) {
      int ncl;
      processCube<true>(S, key, cube, a, cand, tris + 9ull * start, ncl);
      if (tri_owner)
        for (uint32_t t = 0; t < c; ++t) tri_owner[start + t] = owner;
    }
  }
  __syncthreads();
}

// ---- fused count + emit (one staging per item) ----
//
// An item stages its region once, counts every candidate cube, takes its
// output range with one atomic add on the running total, and emits from the
// same staged region.  Items land in the order they finish (as in the
// reference's GPU extractor, whose triangles are appended atomically); within
// an item the triangles keep the two-pass order (cube, then cluster), and
// every triangle carries its item.

// One item: stage once, count every candidate cube, take the offset, emit.
__device__ uint32_t fusedItem(MeshShared& S, float* __restrict__ cand, const VolumeView& V, uint32_t b, int dir,
                              const MeshArgs& a, uint32_t item, uint32_t* total_out, float* tris, uint32_t* tri_owner,
                              uint32_t capacity, uint32_t& staged, const uint32_t* src8 = nullptr) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint64_t key = V.pool_keys[b];
  if (!stageRegion(S, V, key, b, dir, a, src8)) {  // workgroup-uniform
    __syncthreads();  // the item's LDS is reused by the next item
    return 0u;
  }
  ++staged;  // items staged, added to the header once per workgroup
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const int cube = tid + 256 * half;
    S.cnt[cube] = 0;
    listAppend(S, cubeMayCross(S, cube), cube, 0u);
  }
  __syncthreads();
  // count: a lane takes places j = tid and tid + 256 of the count list; the
  // centres of each go to scratch set j >> 8, column tid
  const uint32_t n = S.nlist;
  // No cube can cross (about 4 items in 10): the item has no triangle.
  // Workgroup-uniform; what the next item's staging writes before its
  // first barrier (the sources, and the list lengths, already zero) is
  // nothing this one still reads.
  if (n == 0u) return 0u;
  uint32_t cj0 = 0u, cj1 = 0u;
#pragma unroll 1
  for (uint32_t j = tid; j < n; j += 256) {  // one processCube in the code
    const int cube = S.list_cube[j];
    float* const set = cand + kFusedSetWords * (j >> 8);
    int ncl, nstored = 0;
    const uint32_t c = processCube<false, true>(S, key, cube, a, set + 256, nullptr, ncl, &nstored);
    S.cnt[cube] = static_cast<uint8_t>(c);
    set[tid] = __uint_as_float(static_cast<uint32_t>(cube) | (static_cast<uint32_t>(ncl) << 16) |
                               (static_cast<uint32_t>(nstored) << 24));
    if (j < 256u) cj0 = c; else cj1 = c;
  }
  __syncthreads();  // every count is in S.cnt, every lane has read its count-list places
  // offsets in the reference's cube order (cubes 0..255, then 256..511):
  // both halves scanned per wavefront in registers
  const uint32_t c0 = S.cnt[tid], c1 = S.cnt[tid + 256];
  uint32_t v0 = c0, v1 = c1;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y0 = __shfl_up(v0, off), y1 = __shfl_up(v1, off);
    if (lane >= off) {
      v0 += y0;
      v1 += y1;
    }
  }
  if (lane == 63) {
    S.scan[wave] = v0;
    S.scan[4 + wave] = v1;
  }
  // the count-list places j of the cubes with triangles (they name the
  // scratch columns holding the centres), compacted into list_cube
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const bool take = (half ? cj1 : cj0) > 0u;
    const unsigned long long mask = __ballot(take);
    uint32_t at = 0;
    if (lane == 0 && mask != 0ull) at = atomicAdd(&S.nemit, static_cast<uint32_t>(__popcll(mask)));
    at = __shfl(at, 0);
    if (take) S.list_cube[at + __popcll(mask & ((1ull << lane) - 1ull))] = static_cast<uint16_t>(tid + 256 * half);
  }
  __syncthreads();
  uint32_t t0 = 0, t1 = 0, b0 = 0, b1 = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t s0 = S.scan[w], s1 = S.scan[4 + w];
    t0 += s0;
    t1 += s1;
    if (w < wave) {
      b0 += s0;
      b1 += s1;
    }
  }
  const uint32_t total = t0 + t1;
  S.list_start[tid] = b0 + v0 - c0;  // relative to the item's first triangle
  S.list_start[tid + 256] = t0 + b1 + v1 - c1;
  if (tid == 0) S.base = total ? atomicAdd(total_out, total) : 0u;
  __syncthreads();
  const uint32_t base = S.base;
  const bool fits = static_cast<uint64_t>(base) + total <= capacity;  // workgroup-uniform
  if (total > 0 && fits) {
    const uint32_t m = S.nemit;
    for (uint32_t e = tid; e < m; e += 256) {
      const uint32_t j = S.list_cube[e];
      const float* const set = cand + kFusedSetWords * (j >> 8);
      const uint32_t meta = __float_as_uint(set[j & 255u]);
      const int cube = static_cast<int>(meta & 0xffffu), ncl = static_cast<int>((meta >> 16) & 0xffu);
      const int ns = static_cast<int>(meta >> 24);
      const uint32_t start = base + S.list_start[cube];
      const uint32_t c = S.cnt[cube];
      if (ns == 1 || ns == 2)
        emitStored(S, key, cube, a, set + 256 + (j & 255u), ns, tris + 9ull * start);
      else
        emitFromCentres(S, key, cube, a, set + 256 + (j & 255u), ncl, tris + 9ull * start);
      if (tri_owner)
        for (uint32_t t = 0; t < c; ++t) tri_owner[start + t] = item;
    }
  } else if (total > 0) {
    // The item does not fit the buffer: none of its triangles is written.
    // Its entries below capacity are marked invalid (item kInvalid, NaN
    // coordinates), and its direction is marked dirty again so that the next
    // incremental call lists it.
    const uint32_t end = static_cast<uint64_t>(base) + total < capacity ? base + total : capacity;
    for (uint32_t t = base + tid; t < end; t += 256) {
#pragma unroll
      for (int k = 0; k < 9; ++k) tris[9ull * t + k] = __uint_as_float(0x7fc00000u);
      if (tri_owner) tri_owner[t] = kInvalid;
    }
    if (tid == 0) atomicOr(&V.dirty[b], 1u << dir);
  }
  __syncthreads();  // the item's LDS is reused by the next item
  return total;
}

// Fused incremental pass, persistent over the 6 n items (item dir * n + i:
// listed block list[i], direction dir, skipped unless dirs[i] has bit dir).
// total_out: zeroed; receives the number of triangles.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void meshFusedKernel(VolumeView V, const uint32_t* __restrict__ list, uint32_t n,
                                                       MeshArgs a, const uint8_t* __restrict__ dirs,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint8_t* __restrict__ observed_dirs,
                                                       uint32_t* __restrict__ total_out, float* __restrict__ tris,
                                                       uint32_t* __restrict__ tri_owner, uint32_t capacity,
                                                       float* __restrict__ cand) {
  __shared__ MeshShared S;
  loadCaseCounts(S);
  float* const mine = cand + kMeshFusedCandWords * blockIdx.x;
  const uint32_t items = kDirs * n;
  uint32_t staged = 0;
  for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const uint32_t dir = item / n, i = item % n;
    // an item none of whose 8 sources observed the direction has no triangle
    if (dirs[i] & observed_dirs[i] & (1u << dir))  // workgroup-uniform
      fusedItem(S, mine, V, list[i], static_cast<int>(dir), a, item, total_out, tris, tri_owner, capacity, staged,
                nbr + 8ull * i);
  }
  if (threadIdx.x == 0 && staged) atomicAdd(&V.hdr->mesh_items, static_cast<unsigned long long>(staged));
}

// Source block k of listed block i (self and the 7 positive neighbours,
// kInvalid when absent) into nbr[8 i + k], and the OR of the 8 sources'
// observed-direction masks into observed_dirs[i].  Called by 8 consecutive
// lanes of a wavefront per block (all lanes of the wavefront take part).
//
// signs (iso 0 only): keep a direction only if the sources also hold both
// near-iso sign hints for it (kSignBelow / kSignAbove).
__device__ __forceinline__ void neighbourLane(const VolumeView& V, const uint32_t* __restrict__ list, bool active,
                                              uint32_t i, int k, uint32_t* __restrict__ nbr,
                                              uint8_t* __restrict__ observed_dirs, bool signs) {
  uint32_t dm = 0;
  if (active) {
    const uint32_t self = list[i];
    const uint64_t key = V.pool_keys[self];
    const uint32_t p = k == 0 ? self : lookupPool(V, keyX(key) + (k & 1), keyY(key) + ((k >> 1) & 1), keyZ(key) + (k >> 2));
    const bool ok = p < V.max_blocks;
    nbr[8ull * i + k] = ok ? p : kInvalid;
    dm = ok ? V.dirmask[p] : 0u;
  }
  // OR over the block's 8 lanes (aligned groups of 8 within the wavefront)
  dm |= __shfl_xor(dm, 1);
  dm |= __shfl_xor(dm, 2);
  dm |= __shfl_xor(dm, 4);
  if (signs) dm &= (dm >> kSignBelow) & (dm >> kSignAbove);
  if (active && k == 0) observed_dirs[i] = static_cast<uint8_t>(dm & 0x3Fu);
}

// The sources of listed blocks 0..n-1: eight lanes per block.
__global__ __launch_bounds__(256) void meshNeighboursKernel(VolumeView V, const uint32_t* __restrict__ list,
                                                            uint32_t n, uint32_t* __restrict__ nbr,
                                                            uint8_t* __restrict__ observed_dirs, bool signs) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  neighbourLane(V, list, (g >> 3) < n, g >> 3, static_cast<int>(g & 7u), nbr, observed_dirs, signs);
}

// Persistent over the 6 n items: item dir * n + i meshes block list[i] (or
// pool index i when list is null) in direction dir; total in counts[item],
// per-cube counts in cube_counts[512 * item ..].  dirs (optional): per
// listed block, the directions to mesh; the others count zero triangles.
// cand: kMeshCandWords floats of cluster centres per workgroup.
__global__ __launch_bounds__(256) void meshCountKernel(VolumeView V, const uint32_t* __restrict__ list, uint32_t n,
                                                       MeshArgs a, uint32_t* __restrict__ counts,
                                                       uint8_t* __restrict__ cube_counts,
                                                       const uint8_t* __restrict__ dirs, float* __restrict__ cand) {
  __shared__ MeshShared S;
  loadCaseCounts(S);
  float* const mine = cand + kMeshCandWords * blockIdx.x;
  const uint32_t items = kDirs * n;
  uint32_t staged = 0;
  for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const uint32_t dir = item / n, i = item % n;
    uint32_t total = 0;
    if (!dirs || (dirs[i] & (1u << dir)))  // workgroup-uniform
      total = countItem(S, mine, V, list ? list[i] : i, static_cast<int>(dir), a, cube_counts + 512ull * item, staged);
    if (threadIdx.x == 0) counts[item] = total;
  }
  if (threadIdx.x == 0 && staged) atomicAdd(&V.hdr->mesh_items, static_cast<unsigned long long>(staged));
}

// Writes the triangles of item dir * n + i at offsets[item] in cube order;
// tri_owner (optional) receives the item for each triangle.
__global__ __launch_bounds__(256) void meshEmitKernel(VolumeView V, const uint32_t* __restrict__ list, uint32_t n,
                                                      MeshArgs a, const uint32_t* __restrict__ counts,
                                                      const uint8_t* __restrict__ cube_counts,
                                                      const uint32_t* __restrict__ offsets, float* __restrict__ tris,
                                                      uint32_t* __restrict__ tri_owner, uint32_t capacity,
                                                      float* __restrict__ cand) {
  __shared__ MeshShared S;
  loadCaseCounts(S);
  float* const mine = cand + kMeshCandWords * blockIdx.x;
  const uint32_t items = kDirs * n;
  for (uint32_t item = blockIdx.x; item < items; item += gridDim.x) {
    const uint32_t dir = item / n, i = item % n;
    emitItem(S, mine, V, list ? list[i] : i, static_cast<int>(dir), a, offsets[item], counts[item],
             cube_counts + 512ull * item, tris, tri_owner, item, capacity);
  }
}

// (Block, direction) pairs whose cubes read a written direction of a dirty
// block: the dirty block itself and its -x/-y/-z neighbours (a cube at voxel
// v reads v + {0,1}^3), in the directions written.  ORs the dirty block's
// direction mask into flags[q] for each of those present in the table and
// clears the dirty mask.
__device__ void markRemeshRange(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, uint32_t start,
                                uint32_t stride) {
  for (uint32_t p = start; p < num_blocks; p += stride) {
    const uint32_t m = V.dirty[p];
    if (m == 0u) continue;
    V.dirty[p] = 0u;
    const uint64_t key = V.pool_keys[p];
    for (int o = 0; o < 8; ++o) {
      const uint32_t q = o == 0 ? p : lookupPool(V, keyX(key) - (o & 1), keyY(key) - ((o >> 1) & 1),
                                                 keyZ(key) - (o >> 2));
      if (q < V.max_blocks) atomicOr(&flags[q], m);
    }
  }
}

__global__ __launch_bounds__(256) void markRemeshKernel(VolumeView V, uint32_t num_blocks, uint32_t* __restrict__ flags) {
  markRemeshRange(V, num_blocks, flags, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

// flags[p] = 1 for blocks with any observed voxel (non-zero observed bitmask).
__global__ __launch_bounds__(256) void observedFlagsKernel(VolumeView V, uint32_t num_blocks,
                                                           uint32_t* __restrict__ flags) {
  for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < num_blocks; p += gridDim.x * 256u) {
    const uint32_t* o = V.obs + static_cast<size_t>(p) * 16u;
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) any |= o[i];
    flags[p] = any ? 1u : 0u;
  }
}

// list[pos[p]] = p for every p with flags[p] != 0 (pos: exclusive scan of
// flags != 0); dirs (optional) receives flags[p] at the same index; the
// flags are cleared.
__global__ __launch_bounds__(256) void compactFlagsKernel(uint32_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ pos, uint32_t n,
                                                          uint32_t* __restrict__ list, uint8_t* __restrict__ dirs) {
  for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < n; p += gridDim.x * 256u) {
    const uint32_t f = flags[p];
    if (f) {
      list[pos[p]] = p;
      if (dirs) dirs[pos[p]] = static_cast<uint8_t>(f);
      flags[p] = 0u;
    }
  }
}

__global__ __launch_bounds__(256) void listKeysKernel(VolumeView V, const uint32_t* __restrict__ list, uint32_t n,
                                                      int32_t* __restrict__ out) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const uint64_t key = V.pool_keys[list[i]];
    out[3 * i] = keyX(key);
    out[3 * i + 1] = keyY(key);
    out[3 * i + 2] = keyZ(key);
  }
}

// -- exclusive scan over u32 (three phases: tiles, partials, add) ----------

constexpr int kScanTile = 1024;

// kFlags: scan (in != 0) instead of in.
template <bool kFlags>
__global__ __launch_bounds__(256) void scanTilesKernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       uint32_t n, uint32_t* __restrict__ partials) {
  __shared__ uint32_t s[kScanTile];
  const uint32_t base = blockIdx.x * kScanTile;
  const int tid = threadIdx.x;
  for (int i = tid; i < kScanTile; i += 256) {
    const uint32_t x = (base + i < n) ? in[base + i] : 0u;
    s[i] = kFlags ? (x != 0u ? 1u : 0u) : x;
  }
  __syncthreads();
  // each lane scans 4 consecutive elements, then lanes are scanned
  uint32_t v[4], acc = 0;
  for (int j = 0; j < 4; ++j) {
    v[j] = acc;
    acc += s[tid * 4 + j];
  }
  __shared__ uint32_t lanes[256];
  lanes[tid] = acc;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const uint32_t y = tid >= off ? lanes[tid - off] : 0u;
    __syncthreads();
    lanes[tid] += y;
    __syncthreads();
  }
  const uint32_t lane_excl = lanes[tid] - acc;
  for (int j = 0; j < 4; ++j) {
    const uint32_t i = base + tid * 4 + j;
    if (i < n) out[i] = v[j] + lane_excl;
  }
  if (tid == 255) partials[blockIdx.x] = lanes[255];
}

__global__ __launch_bounds__(256) void scanPartialsKernel(uint32_t* __restrict__ partials, uint32_t m) {
  // single workgroup: sequential chunks of 256 with a running carry;
  // partials[m] receives the grand total.
  __shared__ uint32_t s[256];
  __shared__ uint32_t carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < m; base += 256) {
    const uint32_t x = (base + tid < m) ? partials[base + tid] : 0u;
    s[tid] = x;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const uint32_t y = tid >= off ? s[tid - off] : 0u;
      __syncthreads();
      s[tid] += y;
      __syncthreads();
    }
    if (base + tid < m) partials[base + tid] = carry + s[tid] - x;
    __syncthreads();
    if (tid == 0) carry += s[255];
    __syncthreads();
  }
  if (tid == 0) partials[m] = carry;
}

__global__ __launch_bounds__(256) void scanAddKernel(uint32_t* __restrict__ out, uint32_t n,
                                                     const uint32_t* __restrict__ partials) {
  const uint32_t add = partials[blockIdx.x];
  const uint32_t base = blockIdx.x * kScanTile;
  for (int i = threadIdx.x; i < kScanTile; i += 256)
    if (base + i < n) out[base + i] += add;
}

// ---- batched incremental meshing (all scenes of a batch step) ----

__device__ void markRemeshScene(const VolumeView& V, uint32_t* flags, uint32_t start, uint32_t stride) {
  const uint32_t nb = V.hdr->num_blocks < V.max_blocks ? V.hdr->num_blocks : V.max_blocks;
  markRemeshRange(V, nb, flags, start, stride);
}

__global__ __launch_bounds__(256) void markRemeshBatchKernel(const BatchScene* __restrict__ scenes) {
  const BatchScene& sc = scenes[blockIdx.y];
  markRemeshScene(sc.V, sc.remesh_flags, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

// Sequential-chunk exclusive scan by one workgroup: out[i] = sum(in[0..i)),
// returns the total (all lanes).  With `list`, instead compacts the indices of
// non-zero entries of `in` into list (in order) and clears them.
__device__ uint32_t workgroupScan(uint32_t* in, uint32_t n, uint32_t* out, uint32_t* list, uint8_t* dirs) {
  __shared__ uint32_t s[256];
  __shared__ uint32_t carry;
  const int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (uint32_t base = 0; base < n; base += 256) {
    const uint32_t i = base + tid;
    const uint32_t raw = i < n ? in[i] : 0u;
    const uint32_t x = list ? (raw ? 1u : 0u) : raw;
    s[tid] = x;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      const uint32_t y = tid >= off ? s[tid - off] : 0u;
      __syncthreads();
      s[tid] += y;
      __syncthreads();
    }
    const uint32_t excl = carry + s[tid] - x;
    if (i < n) {
      if (list) {
        if (x) {
          list[excl] = i;
          dirs[excl] = static_cast<uint8_t>(raw);
          in[i] = 0u;
        }
      } else {
        out[i] = excl;
      }
    }
    __syncthreads();
    if (tid == 255) carry += s[255];
    __syncthreads();
  }
  return carry;
}

__global__ __launch_bounds__(256) void compactRemeshBatchKernel(const BatchScene* __restrict__ scenes) {
  const BatchScene& sc = scenes[blockIdx.x];
  const VolumeView& V = sc.V;
  const uint32_t nb = V.hdr->num_blocks < V.max_blocks ? V.hdr->num_blocks : V.max_blocks;
  const uint32_t m = workgroupScan(sc.remesh_flags, nb, nullptr, sc.remesh_list, sc.remesh_dirs);
  if (threadIdx.x == 0) sc.out[0] = m;
}

// Item prefix over the scenes (6 items per listed block); also zeroes each
// scene's triangle total, which the fused pass accumulates.
__global__ void meshPrefixKernel(const BatchScene* __restrict__ scenes, int S, uint32_t* __restrict__ prefix) {
  if (threadIdx.x != 0) return;
  uint32_t acc = 0;
  for (int s = 0; s < S; ++s) {
    prefix[s] = acc;
    acc += 6u * scenes[s].out[0];
    scenes[s].out[1] = 0u;
  }
  prefix[S] = acc;
}

// The sources of every scene's listed blocks: a grid-stride loop (the grid
// is fixed for the captured step) over 8 lanes per block of the
// concatenation (block b of the concatenation is item 6 b of prefix's).
__global__ __launch_bounds__(256) void meshNeighboursBatchKernel(const BatchScene* __restrict__ scenes, int S,
                                                                 const uint32_t* __restrict__ prefix) {
  const uint32_t lanes = 8u * (prefix[S] / 6u);
  for (uint32_t base = blockIdx.x * blockDim.x; base < lanes; base += gridDim.x * blockDim.x) {
    const uint32_t g = base + threadIdx.x;  // the whole wavefront stays in the loop together
    const bool active = g < lanes;
    const uint32_t bg = (active ? g : lanes - 1u) >> 3;
    const int s = findSegment(prefix, S, 6u * bg);
    const BatchScene& sc = scenes[s];
    neighbourLane(sc.V, sc.remesh_list, active, bg - prefix[s] / 6u, static_cast<int>(g & 7u), sc.mesh_nbr,
                  sc.mesh_obs_dirs, sc.ma.iso == 0.0f);
  }
}

// Fused count + emit over the concatenated items of every scene: each item's
// triangles go to one range of its scene's buffer, taken with an atomic add
// on the scene's total (items in completion order).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void meshFusedBatchKernel(
    const BatchScene* __restrict__ scenes, int S, const uint32_t* __restrict__ prefix, float* __restrict__ cand) {
  __shared__ MeshShared M;
  loadCaseCounts(M);
  float* const mine = cand + kMeshFusedCandWords * blockIdx.x;
  const uint32_t total = prefix[S];
  // items staged, added to the scene's header when the scene changes and at the end
  uint32_t staged = 0;
  int staged_scene = -1;
  for (uint32_t i = blockIdx.x; i < total; i += gridDim.x) {
    const int s = __builtin_amdgcn_readfirstlane(findSegment(prefix, S, i));  // uniform: scene fields in SGPRs
    if (s != staged_scene) {
      if (threadIdx.x == 0 && staged)
        atomicAdd(&scenes[staged_scene].V.hdr->mesh_items, static_cast<unsigned long long>(staged));
      staged = 0;
      staged_scene = s;
    }
    const BatchScene& sc = scenes[s];
    const uint32_t m = sc.out[0], j = i - prefix[s];
    const uint32_t dir = j / m, idx = j % m;
    if (sc.remesh_dirs[idx] & sc.mesh_obs_dirs[idx] & (1u << dir))  // workgroup-uniform
      fusedItem(M, mine, sc.V, sc.remesh_list[idx], static_cast<int>(dir), sc.ma, dir * m + idx, sc.out + 1, sc.tris,
                sc.tri_owner, sc.tri_capacity, staged, sc.mesh_nbr + 8ull * idx);
  }
  if (threadIdx.x == 0 && staged)
    atomicAdd(&scenes[staged_scene].V.hdr->mesh_items, static_cast<unsigned long long>(staged));
}

}  // namespace

static uint32_t meshGrid(uint32_t items) { return items < kMeshGridMax ? items : kMeshGridMax; }

hipError_t launchMeshFused(const VolumeView& V, const uint32_t* list, uint32_t n, const MeshArgs& a,
                           const uint8_t* dirs, uint32_t* nbr, uint8_t* observed_dirs, uint32_t* total, float* tris,
                           uint32_t* tri_owner, uint32_t capacity, float* cand, hipStream_t s, uint32_t max_grid) {
  if (n == 0) return hipSuccess;
  meshNeighboursKernel<<<(8ull * n + 255) / 256, 256, 0, s>>>(V, list, n, nbr, observed_dirs, a.iso == 0.0f);
  uint32_t grid = meshGrid(kDirs * n);
  grid = grid < max_grid ? grid : max_grid;
  meshFusedKernel<<<grid, 256, 0, s>>>(V, list, n, a, dirs, nbr, observed_dirs, total, tris, tri_owner, capacity,
                                       cand);
  return hipGetLastError();
}

namespace {
__global__ void zeroWordsKernel(uint32_t* p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// A remesh list that is not meshed: every listed block's flags go back into
// its dirty mask (a superset of what was cleared: the next call lists these
// blocks and their -x/-y/-z neighbours) and the flags are cleared.
__global__ __launch_bounds__(256) void restoreDirtyKernel(VolumeView V, uint32_t num_blocks,
                                                          uint32_t* __restrict__ flags) {
  for (uint32_t p = blockIdx.x * 256u + threadIdx.x; p < num_blocks; p += gridDim.x * 256u) {
    const uint32_t f = flags[p];
    if (f) {
      V.dirty[p] |= f;
      flags[p] = 0u;
    }
  }
}
}  // namespace

static uint32_t gridFor(uint32_t n);

hipError_t launchZeroWords(uint32_t* p, uint32_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  zeroWordsKernel<<<(n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024, 256, 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launchRestoreDirty(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, hipStream_t s) {
  if (num_blocks == 0) return hipSuccess;
  restoreDirtyKernel<<<gridFor(num_blocks), 256, 0, s>>>(V, num_blocks, flags);
  return hipGetLastError();
}

hipError_t launchMeshCount(const VolumeView& V, const uint32_t* list, uint32_t n, const MeshArgs& a,
                           uint32_t* counts, uint8_t* cube_counts, const uint8_t* dirs, float* cand, hipStream_t s) {
  if (n == 0) return hipSuccess;
  meshCountKernel<<<meshGrid(kDirs * n), 256, 0, s>>>(V, list, n, a, counts, cube_counts, dirs, cand);
  return hipGetLastError();
}

size_t scanPartialsSize(uint32_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

// out[i] = sum(in[0..i)); partials[(n + 1023) / 1024] receives the total.
hipError_t launchScan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* partials, hipStream_t s,
                      bool flags) {
  const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles > 0) {
    if (flags)
      scanTilesKernel<true><<<tiles, 256, 0, s>>>(in, out, n, partials);
    else
      scanTilesKernel<false><<<tiles, 256, 0, s>>>(in, out, n, partials);
  }
  scanPartialsKernel<<<1, 256, 0, s>>>(partials, tiles);
  if (tiles > 0) scanAddKernel<<<tiles, 256, 0, s>>>(out, n, partials);
  return hipGetLastError();
}

hipError_t launchMeshEmit(const VolumeView& V, const uint32_t* list, uint32_t n, const MeshArgs& a,
                          const uint32_t* counts, const uint8_t* cube_counts, const uint32_t* offsets, float* tris,
                          uint32_t* tri_owner, uint32_t capacity, float* cand, hipStream_t s) {
  if (n == 0) return hipSuccess;
  meshEmitKernel<<<meshGrid(kDirs * n), 256, 0, s>>>(V, list, n, a, counts, cube_counts, offsets, tris, tri_owner,
                                                      capacity, cand);
  return hipGetLastError();
}

static uint32_t gridFor(uint32_t n) {
  const uint32_t g = (n + 255u) / 256u;
  return g == 0 ? 1u : (g < 4096u ? g : 4096u);
}

hipError_t launchMarkRemesh(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, hipStream_t s) {
  if (num_blocks == 0) return hipSuccess;
  markRemeshKernel<<<gridFor(num_blocks), 256, 0, s>>>(V, num_blocks, flags);
  return hipGetLastError();
}

hipError_t launchObservedFlags(const VolumeView& V, uint32_t num_blocks, uint32_t* flags, hipStream_t s) {
  if (num_blocks == 0) return hipSuccess;
  observedFlagsKernel<<<gridFor(num_blocks), 256, 0, s>>>(V, num_blocks, flags);
  return hipGetLastError();
}

hipError_t launchCompactFlags(uint32_t* flags, const uint32_t* pos, uint32_t n, uint32_t* list, uint8_t* dirs,
                              hipStream_t s) {
  if (n == 0) return hipSuccess;
  compactFlagsKernel<<<gridFor(n), 256, 0, s>>>(flags, pos, n, list, dirs);
  return hipGetLastError();
}

hipError_t launchMeshBatch(const BatchScene* scenes, int S, uint32_t* prefix, int grid, float* cand,
                           hipStream_t s) {
  const uint32_t g = meshGrid(grid > 0 ? static_cast<uint32_t>(grid) : 1u);
  markRemeshBatchKernel<<<dim3(64, S), 256, 0, s>>>(scenes);
  compactRemeshBatchKernel<<<S, 256, 0, s>>>(scenes);
  meshPrefixKernel<<<1, 64, 0, s>>>(scenes, S, prefix);
  meshNeighboursBatchKernel<<<1024, 256, 0, s>>>(scenes, S, prefix);
  meshFusedBatchKernel<<<g, 256, 0, s>>>(scenes, S, prefix, cand);
  return hipGetLastError();
}

hipError_t launchListKeys(const VolumeView& V, const uint32_t* list, uint32_t n, int32_t* out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  listKeysKernel<<<gridFor(n), 256, 0, s>>>(V, list, n, out);
  return hipGetLastError();
}

}  // namespace vhm
