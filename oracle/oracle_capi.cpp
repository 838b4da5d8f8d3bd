// CPU ORACLE C-ABI -- test infrastructure only (see cpu_ref.hpp header).
// Loaded by tests/ and bench.py's cpu_baseline leg through ctypes; never by
// the GPU product.  Each handle holds a float (parity) or double (FP64
// report) instantiation of the restated reference CPU pipeline.

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "cpu_ref.hpp"

using namespace oracle;

namespace {

struct Handle {
  int fp64;
  Volume<float>* f;
  Volume<double>* d;
};

template <typename T>
Frame<T> makeFrame(const float* depth, int w, int h, const float* K4) {
  Frame<T> fr;
  fr.K.width = w;
  fr.K.height = h;
  fr.K.fx = K4[0];
  fr.K.fy = K4[1];
  fr.K.cx = K4[2];
  fr.K.cy = K4[3];
  fr.depth.resize(static_cast<std::size_t>(w) * static_cast<std::size_t>(h));
  for (std::size_t i = 0; i < fr.depth.size(); ++i) fr.depth[i] = static_cast<T>(depth[i]);
  return fr;
}

template <typename T>
Pose<T> makePose(const float* p12) {
  Pose<T> p;
  for (int i = 0; i < 9; ++i) p.R[i] = static_cast<T>(p12[i]);
  for (int i = 0; i < 3; ++i) p.t[i] = static_cast<T>(p12[9 + i]);
  return p;
}

template <typename T>
void integrateImpl(Volume<T>& vol, const float* depth, int w, int h, const float* K4, const float* p12, int stride,
                   int apply_filter) {
  Frame<T> fr = makeFrame<T>(depth, w, h, K4);
  if (apply_filter) fr = bilateralFilter(fr, vol.P);
  vol.integrate(fr, makePose<T>(p12), stride);
}

template <typename T>
std::size_t prepareImpl(Volume<T>& vol, const float* depth, int w, int h, const float* K4, const float* p12,
                        int stride, int32_t* keys, std::size_t cap) {
  const std::vector<Key> k = vol.prepare(makeFrame<T>(depth, w, h, K4), makePose<T>(p12), stride);
  for (std::size_t i = 0; i < k.size() && i < cap; ++i) {
    keys[3 * i] = k[i].x;
    keys[3 * i + 1] = k[i].y;
    keys[3 * i + 2] = k[i].z;
  }
  return k.size();
}

template <typename T>
std::size_t keysImpl(const Volume<T>& vol, int32_t* keys, std::size_t cap) {
  std::size_t i = 0;
  for (const auto& e : vol.blocks) {
    if (i < cap) {
      keys[3 * i] = e.first.x;
      keys[3 * i + 1] = e.first.y;
      keys[3 * i + 2] = e.first.z;
    }
    ++i;
  }
  return vol.blocks.size();
}

// SoA dump: arrays indexed [block][voxel][dir][slot].
template <typename T>
void dumpSoaImpl(const Volume<T>& vol, const int32_t* keys, std::size_t n, double* w, double* sc, double* po,
                 uint32_t* pn) {
  const std::size_t per_block = static_cast<std::size_t>(kBlockVol) * kDirs * kSlots;
  for (std::size_t i = 0; i < n; ++i) {
    const auto it = vol.blocks.find(Key{keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]});
    for (int v = 0; v < kBlockVol; ++v)
      for (int d = 0; d < kDirs; ++d)
        for (int k = 0; k < kSlots; ++k) {
          const std::size_t o = i * per_block + (static_cast<std::size_t>(v) * kDirs + d) * kSlots + k;
          if (it == vol.blocks.end()) {
            w[o] = sc[o] = po[o] = 0.0;
            pn[o] = 0;
          } else {
            const LayerT<T>& l = it->second[static_cast<std::size_t>(v)].layers[d][k];
            w[o] = static_cast<double>(l.weight);
            sc[o] = static_cast<double>(l.surface_coord);
            po[o] = static_cast<double>(l.plane_offset);
            pn[o] = l.packed_normal;
          }
        }
  }
}

template <typename T>
std::size_t meshImpl(const Volume<T>& vol, float iso, float* tris, std::size_t cap) {
  const auto t = vol.extractMesh(static_cast<T>(iso));
  for (std::size_t i = 0; i < t.size() && i < cap; ++i)
    for (int j = 0; j < 3; ++j) {
      tris[9 * i + 3 * j + 0] = static_cast<float>(t[i][j].x);
      tris[9 * i + 3 * j + 1] = static_cast<float>(t[i][j].y);
      tris[9 * i + 3 * j + 2] = static_cast<float>(t[i][j].z);
    }
  return t.size();
}

template <typename T>
void pointCloudImpl(Volume<T>& vol, const float* pts, std::size_t n, const float* p12) {
  std::vector<V3<T>> pc(n);
  for (std::size_t i = 0; i < n; ++i) pc[i] = {T(pts[3 * i]), T(pts[3 * i + 1]), T(pts[3 * i + 2])};
  vol.integratePointCloud(pc, makePose<T>(p12));
}

}  // namespace

extern "C" {

std::size_t vor_sizeof_params() { return sizeof(Params); }

// integratePointCloud (voxel_hash_tsdf.h:75-97): n camera-space points.
void vor_integrate_point_cloud(void* hv, const float* pts, std::size_t n, const float* p12) {
  auto* hd = static_cast<Handle*>(hv);
  if (hd->fp64)
    pointCloudImpl(*hd->d, pts, n, p12);
  else
    pointCloudImpl(*hd->f, pts, n, p12);
}

void vor_default_params(Params* p) { *p = Params{}; }

void* vor_create(const Params* p, int fp64) {
  auto* h = new Handle{fp64, nullptr, nullptr};
  if (fp64)
    h->d = new Volume<double>(*p);
  else
    h->f = new Volume<float>(*p);
  return h;
}

void vor_destroy(void* hv) {
  auto* h = static_cast<Handle*>(hv);
  delete h->f;
  delete h->d;
  delete h;
}

// Bilateral filter (reconstruction_backend.cpp:96-189); fp64 computes in double.
void vor_filter(const Params* p, int fp64, const float* depth, int w, int h, const float* K4, float* out) {
  if (fp64) {
    const Frame<double> r = bilateralFilter(makeFrame<double>(depth, w, h, K4), *p);
    for (std::size_t i = 0; i < r.depth.size(); ++i) out[i] = static_cast<float>(r.depth[i]);
  } else {
    const Frame<float> r = bilateralFilter(makeFrame<float>(depth, w, h, K4), *p);
    std::memcpy(out, r.depth.data(), sizeof(float) * r.depth.size());
  }
}

void vor_integrate(void* hv, const float* depth, int w, int h, const float* K4, const float* p12, int stride,
                   int apply_filter) {
  auto* hd = static_cast<Handle*>(hv);
  if (hd->fp64)
    integrateImpl(*hd->d, depth, w, h, K4, p12, stride, apply_filter);
  else
    integrateImpl(*hd->f, depth, w, h, K4, p12, stride, apply_filter);
}

std::size_t vor_prepare(void* hv, const float* depth, int w, int h, const float* K4, const float* p12, int stride,
                        int32_t* keys, std::size_t cap) {
  auto* hd = static_cast<Handle*>(hv);
  return hd->fp64 ? prepareImpl(*hd->d, depth, w, h, K4, p12, stride, keys, cap)
                  : prepareImpl(*hd->f, depth, w, h, K4, p12, stride, keys, cap);
}

void vor_integrate_blocks(void* hv, const int32_t* keys, std::size_t n, const float* depth, int w, int h,
                          const float* K4, const float* p12) {
  auto* hd = static_cast<Handle*>(hv);
  std::vector<Key> active(n);
  for (std::size_t i = 0; i < n; ++i) active[i] = {keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]};
  if (hd->fp64)
    hd->d->integrateBlocks(active, makeFrame<double>(depth, w, h, K4), makePose<double>(p12));
  else
    hd->f->integrateBlocks(active, makeFrame<float>(depth, w, h, K4), makePose<float>(p12));
}

std::size_t vor_block_count(void* hv) {
  auto* hd = static_cast<Handle*>(hv);
  return hd->fp64 ? hd->d->blocks.size() : hd->f->blocks.size();
}

std::size_t vor_observed_voxel_count(void* hv) {
  auto* hd = static_cast<Handle*>(hv);
  return hd->fp64 ? hd->d->observedVoxelCount() : hd->f->observedVoxelCount();
}

uint64_t vor_update_count(void* hv) {
  auto* hd = static_cast<Handle*>(hv);
  return hd->fp64 ? hd->d->updates : hd->f->updates;
}

std::size_t vor_block_keys(void* hv, int32_t* keys, std::size_t cap) {
  auto* hd = static_cast<Handle*>(hv);
  return hd->fp64 ? keysImpl(*hd->d, keys, cap) : keysImpl(*hd->f, keys, cap);
}

void vor_dump_blocks_soa(void* hv, const int32_t* keys, std::size_t n, double* w, double* sc, double* po,
                         uint32_t* pn) {
  auto* hd = static_cast<Handle*>(hv);
  if (hd->fp64)
    dumpSoaImpl(*hd->d, keys, n, w, sc, po, pn);
  else
    dumpSoaImpl(*hd->f, keys, n, w, sc, po, pn);
}

// float oracle only: the reference's AoS voxel layout (288 B per voxel).
void vor_dump_blocks(void* hv, const int32_t* keys, std::size_t n, void* out) {
  auto* hd = static_cast<Handle*>(hv);
  auto* dst = static_cast<unsigned char*>(out);
  const std::size_t bb = sizeof(VoxelT<float>) * kBlockVol;
  for (std::size_t i = 0; i < n; ++i) {
    const auto it = hd->f->blocks.find(Key{keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]});
    if (it == hd->f->blocks.end())
      std::memset(dst + i * bb, 0, bb);
    else
      std::memcpy(dst + i * bb, it->second.data(), bb);
  }
}

// float oracle only: create/overwrite blocks from the AoS layout.
void vor_load_blocks(void* hv, const int32_t* keys, std::size_t n, const void* in) {
  auto* hd = static_cast<Handle*>(hv);
  const auto* src = static_cast<const unsigned char*>(in);
  const std::size_t bb = sizeof(VoxelT<float>) * kBlockVol;
  for (std::size_t i = 0; i < n; ++i) {
    auto& blk = hd->f->blocks[Key{keys[3 * i], keys[3 * i + 1], keys[3 * i + 2]}];
    blk.resize(kBlockVol);
    std::memcpy(blk.data(), src + i * bb, bb);
  }
}

std::size_t vor_extract_mesh(void* hv, float iso, float* tris, std::size_t cap) {
  auto* hd = static_cast<Handle*>(hv);
  return hd->fp64 ? meshImpl(*hd->d, iso, tris, cap) : meshImpl(*hd->f, iso, tris, cap);
}

// --- single helpers (float) -------------------------------------------------
uint32_t vor_pack_normal(const float* n3) { return packNormal(V3<float>{n3[0], n3[1], n3[2]}); }
void vor_unpack_normal(uint32_t p, float* n3) {
  const V3<float> n = unpackNormal<float>(p);
  n3[0] = n.x;
  n3[1] = n.y;
  n3[2] = n.z;
}
float vor_observation_weight(float depth, float sd, float trunc, float vaw) {
  return observationWeight(depth, sd, trunc, vaw, 0.15f);
}
int vor_dominant_direction(const float* n3) { return dominantDirection(V3<float>{n3[0], n3[1], n3[2]}); }
int vor_complete_cube(float* v8, uint8_t* ok8) {
  std::array<float, 8> v{};
  std::array<std::uint8_t, 8> ok{};
  for (int i = 0; i < 8; ++i) {
    v[i] = v8[i];
    ok[i] = ok8[i];
  }
  const bool r = Volume<float>::completeCube(v, ok);
  for (int i = 0; i < 8; ++i) {
    v8[i] = v[i];
    ok8[i] = ok[i];
  }
  return r ? 1 : 0;
}
int vor_polygonise(const float* pos24, const float* v8, float iso, float* tris45) {
  std::array<V3<float>, 8> pos{};
  std::array<float, 8> v{};
  for (int i = 0; i < 8; ++i) {
    pos[i] = {pos24[3 * i], pos24[3 * i + 1], pos24[3 * i + 2]};
    v[i] = v8[i];
  }
  std::vector<std::array<V3<float>, 3>> t;
  Volume<float>::polygonise(pos, v, iso, t);
  for (std::size_t i = 0; i < t.size(); ++i)
    for (int j = 0; j < 3; ++j) {
      tris45[9 * i + 3 * j] = t[i][j].x;
      tris45[9 * i + 3 * j + 1] = t[i][j].y;
      tris45[9 * i + 3 * j + 2] = t[i][j].z;
    }
  return static_cast<int>(t.size());
}
int vor_sample_bilinear(const Params* p, const float* depth, int w, int h, const float* K4, float u, float v,
                        float* out) {
  Volume<float> vol(*p);
  return vol.sampleBilinear(makeFrame<float>(depth, w, h, K4), u, v, *out) ? 1 : 0;
}
int vor_surface_normal(const Params* p, const float* depth, int w, int h, const float* K4, float u, float v,
                       float* n3) {
  Volume<float> vol(*p);
  V3<float> n{n3[0], n3[1], n3[2]};
  const bool ok = vol.surfaceNormal(makeFrame<float>(depth, w, h, K4), u, v, n);
  n3[0] = n.x;
  n3[1] = n.y;
  n3[2] = n.z;
  return ok ? 1 : 0;
}
void vor_integrate_sample(void* voxel288, int d, float coord, const float* p3, const float* n3, float w, float tol,
                          float min_cos) {
  VoxelT<float> vx;
  std::memcpy(&vx, voxel288, sizeof(vx));
  Volume<float>::integrateSample(vx, d, coord, V3<float>{p3[0], p3[1], p3[2]}, V3<float>{n3[0], n3[1], n3[2]}, w, tol,
                                 min_cos, 100.0f);
  std::memcpy(voxel288, &vx, sizeof(vx));
}

}  // extern "C"
