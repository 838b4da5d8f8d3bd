// CPU ORACLE -- test infrastructure only.
//
// A plain restatement of the reference CPU fusion path of HM-ZC/voxelhashmt,
// templated on the scalar type: T = float is the parity oracle (same operation
// order as the reference, so results are bit-identical to it), T = double is
// the FP64 oracle used to report how far float arithmetic drifts.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// the library built from this file (oracle/liboracle.so), and only as the
// checker or the timed CPU baseline.  The GPU product never links it.
//
// Every function cites the reference file:line it restates.  What pins each
// function beyond this restatement (the reference's own CPU code compiled
// from its sources, closed forms, brute force) is in the docstrings of
// tests/test_oracle_units.py, tests/test_golden.py and tests/test_point_cloud.py.

#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../voxelhashmt_amd/csrc/vhm_mc_tables.h"

namespace oracle {

// Algorithm constants.  The defaults are the reference CPU values
// (voxel_hash_tsdf.h:194-204, reconstruction_backend.cpp:26-32); the
// alternative preset holds the reference CUDA values
// (cuda_tsdf_integrator.cu:18-35).
struct Params {
  float voxel_size = 0.01f;
  float truncation = 0.04f;
  // depth filter (reconstruction_backend.cpp:26-32)
  int filter_radius = 2;
  float filter_sigma_spatial = 1.5f;
  float filter_sigma_range = 0.0012f;
  float filter_sigma_angular = 0.35f;
  float filter_disc_base = 0.0012f;
  float filter_disc_rel = 0.0025f;
  float filter_min_cos = 0.6f;
  // integrator (voxel_hash_tsdf.h:194-204)
  float max_weight = 100.0f;
  float min_view_angle_weight = 0.15f;
  float normal_sigma_spatial = 1.0f;
  float normal_sigma_range = 0.0012f;
  float normal_sigma_angular = 0.35f;
  float disc_base = 0.0012f;
  float disc_rel = 0.0025f;
  float layer_merge_scale = 0.25f;
  float layer_extract_scale = 0.55f;
  float layer_normal_cos = 0.75f;
};

template <typename T>
struct V3 {
  T x{}, y{}, z{};
  T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

// math_types.h:44-101 -- component-wise, left-to-right.
template <typename T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> T norm(V3<T> v) { return std::sqrt(dot(v, v)); }
template <typename T> V3<T> normalized(V3<T> v) {  // math_types.h:95-101
  const T len = norm(v);
  if (len <= T(1e-8f)) return {T(0), T(0), T(0)};
  return v / len;
}
// std::max / std::min / clampf exactly as the reference calls them
// (math_types.h:12-14): max(a,b) = a<b ? b : a, min(a,b) = b<a ? b : a.
template <typename T> T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T> T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> T clampv(T v, T lo, T hi) { return smax(lo, smin(hi, v)); }

template <typename T>
struct Pose {  // math_types.h:180-198 (R row-major)
  T R[9];
  T t[3];
  V3<T> transformPoint(V3<T> p) const {  // (R*p) + t, math_types.h:132-138, 184-186
    const V3<T> r{R[0] * p.x + R[1] * p.y + R[2] * p.z, R[3] * p.x + R[4] * p.y + R[5] * p.z,
                  R[6] * p.x + R[7] * p.y + R[8] * p.z};
    return r + V3<T>{t[0], t[1], t[2]};
  }
  V3<T> transformVector(V3<T> v) const {
    return {R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z,
            R[6] * v.x + R[7] * v.y + R[8] * v.z};
  }
  Pose inverse() const {  // math_types.h:192-197: R' = R^T, t' = R' * (t * -1)
    Pose o;
    const T Rt[9] = {R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8]};
    for (int i = 0; i < 9; ++i) o.R[i] = Rt[i];
    const V3<T> mt{t[0] * T(-1), t[1] * T(-1), t[2] * T(-1)};
    const V3<T> r = o.transformVector(mt);
    o.t[0] = r.x;
    o.t[1] = r.y;
    o.t[2] = r.z;
    return o;
  }
};

struct Intr {
  int width = 0, height = 0;
  float fx = 0, fy = 0, cx = 0, cy = 0;
};

template <typename T>
struct Frame {  // math_types.h:209-226
  Intr K;
  std::vector<T> depth;
  T at(int u, int v) const { return depth[static_cast<std::size_t>(v * K.width + u)]; }
};

// The scalar exp the reference calls: std::exp on its float argument (glibc
// expf for T=float); double exp for the FP64 oracle.
template <typename T> T texp(T x) { return std::exp(x); }

// voxel_hash_tsdf.h:552-554 / reconstruction_backend.cpp:34-36
template <typename T> T gaussian(T sq, T inv_two_sigma_sq) { return texp(-sq * inv_two_sigma_sq); }

template <typename T> T discThreshold(T d, T base, T rel) {  // voxel_hash_tsdf.h:226-228
  return smax(base, rel * smax(d, T(0)));
}

template <typename T> V3<T> backProjectInt(int u, int v, T d, const Intr& K) {  // math_types.h:228-232
  return {(static_cast<T>(u) - T(K.cx)) * d / T(K.fx), (static_cast<T>(v) - T(K.cy)) * d / T(K.fy), d};
}
template <typename T> V3<T> backProjectCont(T u, T v, T d, const Intr& K) {  // voxel_hash_tsdf.h:586-593
  return {(u - T(K.cx)) * d / T(K.fx), (v - T(K.cy)) * d / T(K.fy), d};
}
template <typename T> bool projectPoint(V3<T> p, const Intr& K, T& u, T& v) {  // math_types.h:234-244
  if (p.z <= T(1e-6f)) return false;
  u = T(K.fx) * p.x / p.z + T(K.cx);
  v = T(K.fy) * p.y / p.z + T(K.cy);
  return u >= T(0) && u <= static_cast<T>(K.width - 1) && v >= T(0) && v <= static_cast<T>(K.height - 1);
}
template <typename T> V3<T> orientTowardCamera(V3<T> n, V3<T> p) {  // voxel_hash_tsdf.h:230-237
  if (norm(n) <= T(1e-6f) || norm(p) <= T(1e-6f)) return n;
  const V3<T> ray = normalized(p);
  return dot(n, ray) > T(0) ? (n * T(-1)) : n;
}

// ---------------------------------------------------------------- filter ---

// reconstruction_backend.cpp:57-94
template <typename T>
bool pixelNormal(const Frame<T>& in, int u, int v, const Params& P, V3<T>& n) {
  if (u <= 0 || v <= 0 || u + 1 >= in.K.width || v + 1 >= in.K.height) return false;
  const T c = in.at(u, v), l = in.at(u - 1, v), r = in.at(u + 1, v), up = in.at(u, v - 1), dn = in.at(u, v + 1);
  if (c <= T(0) || l <= T(0) || r <= T(0) || up <= T(0) || dn <= T(0)) return false;
  const T thr = discThreshold(c, T(P.filter_disc_base), T(P.filter_disc_rel));
  if (std::fabs(l - c) > thr || std::fabs(r - c) > thr || std::fabs(up - c) > thr || std::fabs(dn - c) > thr)
    return false;
  const V3<T> pl = backProjectCont(static_cast<T>(u - 1), static_cast<T>(v), l, in.K);
  const V3<T> pr = backProjectCont(static_cast<T>(u + 1), static_cast<T>(v), r, in.K);
  const V3<T> pu = backProjectCont(static_cast<T>(u), static_cast<T>(v - 1), up, in.K);
  const V3<T> pd = backProjectCont(static_cast<T>(u), static_cast<T>(v + 1), dn, in.K);
  n = normalized(cross(pd - pu, pr - pl));
  if (norm(n) <= T(1e-6f)) return false;
  const V3<T> cp = backProjectCont(static_cast<T>(u), static_cast<T>(v), c, in.K);
  n = orientTowardCamera(n, cp);
  return norm(n) > T(1e-6f);
}

// reconstruction_backend.cpp:96-189.  The inverse two-sigma-squared constants
// are formed as the reference's constexprs are: 1 / (2 * s * s) in T.
template <typename T>
Frame<T> bilateralFilter(const Frame<T>& in, const Params& P) {
  Frame<T> out{in.K, std::vector<T>(in.depth.size(), T(0))};
  const int W = in.K.width, H = in.K.height;
  if (W <= 0 || H <= 0 || in.depth.empty()) return out;
  const T ss = T(P.filter_sigma_spatial), sr = T(P.filter_sigma_range), sa = T(P.filter_sigma_angular);
  const T inv_s = T(1) / (T(2) * ss * ss), inv_r = T(1) / (T(2) * sr * sr), inv_a = T(1) / (T(2) * sa * sa);
  std::vector<V3<T>> normals(in.depth.size());
  std::vector<std::uint8_t> valid(in.depth.size(), 0);
  for (int v = 0; v < H; ++v)
    for (int u = 0; u < W; ++u) {
      V3<T> n{};
      if (pixelNormal(in, u, v, P, n)) {
        normals[static_cast<std::size_t>(v * W + u)] = n;
        valid[static_cast<std::size_t>(v * W + u)] = 1;
      }
    }
  const int R = P.filter_radius;
  for (int v = 0; v < H; ++v)
    for (int u = 0; u < W; ++u) {
      const T c = in.at(u, v);
      if (c <= T(0)) continue;
      const std::size_t ci = static_cast<std::size_t>(v * W + u);
      const bool has_n = valid[ci] != 0;
      const V3<T> cn = has_n ? normals[ci] : V3<T>{};
      const T thr = discThreshold(c, T(P.filter_disc_base), T(P.filter_disc_rel));
      T wd = T(0), ws = T(0);
      for (int dv = -R; dv <= R; ++dv) {
        const int sv = v + dv;
        if (sv < 0 || sv >= H) continue;
        for (int du = -R; du <= R; ++du) {
          const int su = u + du;
          if (su < 0 || su >= W) continue;
          const T s = in.at(su, sv);
          if (s <= T(0)) continue;
          if (std::fabs(s - c) > thr) continue;
          const T sp = static_cast<T>(du * du + dv * dv);
          const T dd = s - c;
          T w = gaussian(sp, inv_s) * gaussian(dd * dd, inv_r);
          if (has_n) {
            const std::size_t si = static_cast<std::size_t>(sv * W + su);
            if (valid[si] == 0) continue;
            const T cosv = clampv(dot(cn, normals[si]), T(-1), T(1));
            if (cosv < T(P.filter_min_cos)) continue;
            const T ad = T(1) - cosv;
            w *= gaussian(ad * ad, inv_a);
          }
          wd += w * s;
          ws += w;
        }
      }
      out.depth[ci] = ws > T(1e-6f) ? (wd / ws) : c;
    }
  return out;
}

// ---------------------------------------------------------------- volume ---

struct Layer {  // voxel_hash_tsdf.h:18-23 (fields as T in the FP64 oracle)
  double weight = 0, surface_coord = 0, plane_offset = 0;
  std::uint32_t packed_normal = 0;
};

template <typename T>
struct LayerT {
  T weight{}, surface_coord{}, plane_offset{};
  std::uint32_t packed_normal = 0;
};

constexpr int kDirs = 6, kSlots = 3, kBlock = 8, kBlockVol = 512;

template <typename T>
struct VoxelT {  // voxel_hash_tsdf.h:25-31
  LayerT<T> layers[kDirs][kSlots];
};

struct Key {
  int x, y, z;
  bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
};
struct KeyHash {
  std::size_t operator()(const Key& k) const {
    std::uint64_t h = static_cast<std::uint32_t>(k.x);
    h = h * 0x9E3779B97F4A7C15ull ^ static_cast<std::uint32_t>(k.y);
    h = h * 0x9E3779B97F4A7C15ull ^ static_cast<std::uint32_t>(k.z);
    return static_cast<std::size_t>(h ^ (h >> 29));
  }
};

inline int floorDiv(int value, int divisor) {  // voxel_hash_tsdf.h:531-538
  int q = value / divisor;
  const int r = value % divisor;
  if (r != 0 && ((r < 0) != (divisor < 0))) --q;
  return q;
}
inline int positiveMod(int value, int divisor) {  // voxel_hash_tsdf.h:540-546
  int r = value % divisor;
  if (r < 0) r += divisor;
  return r;
}

template <typename T> T signNotZero(T v) { return v >= T(0) ? T(1) : T(-1); }

// voxel_hash_tsdf.h:268-289
template <typename T>
std::uint32_t packNormal(V3<T> raw) {
  const V3<T> n = normalized(raw);
  const T s = std::fabs(n.x) + std::fabs(n.y) + std::fabs(n.z);
  if (s <= T(1e-6f)) return 0u;
  V3<T> p = n / s;
  if (p.z < T(0)) {
    const T px = (T(1) - std::fabs(p.y)) * signNotZero(p.x);
    const T py = (T(1) - std::fabs(p.x)) * signNotZero(p.y);
    p.x = px;
    p.y = py;
  }
  auto q = [](T value) -> std::uint32_t {
    const T c = clampv(value * T(0.5f) + T(0.5f), T(0), T(1));
    return static_cast<std::uint32_t>(std::lround(c * T(65535.0f)));
  };
  return q(p.x) | (q(p.y) << 16u);
}

// voxel_hash_tsdf.h:291-309
template <typename T>
V3<T> unpackNormal(std::uint32_t packed) {
  auto dq = [](std::uint32_t c) { return (static_cast<T>(c & 0xffffu) / T(65535.0f)) * T(2) - T(1); };
  V3<T> n{dq(packed), dq(packed >> 16u), T(0)};
  n.z = T(1) - std::fabs(n.x) - std::fabs(n.y);
  if (n.z < T(0)) {
    const T px = (T(1) - std::fabs(n.y)) * signNotZero(n.x);
    const T py = (T(1) - std::fabs(n.x)) * signNotZero(n.y);
    n.x = px;
    n.y = py;
  }
  return normalized(n);
}

template <typename T> V3<T> dirUnit(int d) {  // voxel_hash_tsdf.h:311-326
  switch (d) {
    case 0: return {T(1), T(0), T(0)};
    case 1: return {T(-1), T(0), T(0)};
    case 2: return {T(0), T(1), T(0)};
    case 3: return {T(0), T(-1), T(0)};
    case 4: return {T(0), T(0), T(1)};
    default: return {T(0), T(0), T(-1)};
  }
}

template <typename T> int dominantDirection(V3<T> n) {  // voxel_hash_tsdf.h:239-254
  const T nx = std::fabs(n.x), ny = std::fabs(n.y), nz = std::fabs(n.z);
  if (nx <= T(1e-6f) && ny <= T(1e-6f) && nz <= T(1e-6f)) return -1;
  if (nx >= ny && nx >= nz) return n.x >= T(0) ? 0 : 1;
  if (ny >= nx && ny >= nz) return n.y >= T(0) ? 2 : 3;
  return n.z >= T(0) ? 4 : 5;
}

template <typename T> T planeSD(const LayerT<T>& l, V3<T> p) {  // voxel_hash_tsdf.h:336-339
  return dot(unpackNormal<T>(l.packed_normal), p) - l.plane_offset;
}

// voxel_hash_tsdf.h:556-584
template <typename T>
T observationWeight(T depth, T sd, T trunc, T view_angle_weight, T min_view) {
  if (depth <= T(0) || trunc <= T(1e-6f)) return T(0);
  const T ratio = depth / T(1.5f);
  const T dw = T(1) / (T(1) + ratio * ratio);
  const T r = std::fabs(sd) / trunc;
  T robust = T(1);
  if (r > T(0.35f)) robust = T(0.35f) / r;
  const T band = T(1) - T(0.5f) * clampv(r, T(0), T(1));
  const T base = smax(T(0.05f), dw * robust * band);
  return base * clampv(view_angle_weight, min_view, T(1));
}

template <typename T>
class Volume {
 public:
  explicit Volume(const Params& p) : P(p) {}
  Params P;
  std::unordered_map<Key, std::vector<VoxelT<T>>, KeyHash> blocks;
  std::uint64_t updates = 0;  // calls reaching integrateDirectionalSample

  T voxel() const { return T(P.voxel_size); }
  T trunc() const { return T(P.truncation); }

  // voxel_hash_tsdf.h:595-617
  std::array<int, 3> worldToVoxel(V3<T> p) const {
    return {static_cast<int>(std::floor(p.x / voxel())), static_cast<int>(std::floor(p.y / voxel())),
            static_cast<int>(std::floor(p.z / voxel()))};
  }
  V3<T> voxelToWorld(int x, int y, int z) const {
    return {(static_cast<T>(x) + T(0.5f)) * voxel(), (static_cast<T>(y) + T(0.5f)) * voxel(),
            (static_cast<T>(z) + T(0.5f)) * voxel()};
  }
  static Key voxelToKey(const std::array<int, 3>& v) {
    return {floorDiv(v[0], kBlock), floorDiv(v[1], kBlock), floorDiv(v[2], kBlock)};
  }

  // voxel_hash_tsdf.h:638-671 (with collectNeighborhood 626-636)
  std::vector<Key> collectBlocks(const Frame<T>& f, const Pose<T>& T_wc, int allocation_stride) const {
    const int stride = std::max(1, allocation_stride);
    const V3<T> origin{T_wc.t[0], T_wc.t[1], T_wc.t[2]};
    std::unordered_set<Key, KeyHash> active;
    auto add27 = [&](Key c) {
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) active.insert({c.x + dx, c.y + dy, c.z + dz});
    };
    for (int v = 0; v < f.K.height; v += stride)
      for (int u = 0; u < f.K.width; u += stride) {
        const T d = f.at(u, v);
        if (d <= T(0)) continue;
        const V3<T> pw = T_wc.transformPoint(backProjectInt(u, v, d, f.K));
        const V3<T> ray = normalized(pw - origin);
        const V3<T> front = pw - ray * trunc();
        const V3<T> back = pw + ray * trunc();
        add27(voxelToKey(worldToVoxel(pw)));
        add27(voxelToKey(worldToVoxel(front)));
        add27(voxelToKey(worldToVoxel(back)));
      }
    return std::vector<Key>(active.begin(), active.end());
  }

  std::vector<VoxelT<T>>& blockFor(const Key& k) {  // blocks_[key] (value-initialised voxels)
    auto it = blocks.find(k);
    if (it == blocks.end()) it = blocks.emplace(k, std::vector<VoxelT<T>>(kBlockVol)).first;
    return it->second;
  }

  std::vector<Key> prepare(const Frame<T>& f, const Pose<T>& T_wc, int stride) {  // :133-145
    std::vector<Key> keys = collectBlocks(f, T_wc, stride);
    for (const Key& k : keys)
      if (blocks.find(k) == blocks.end()) blocks.emplace(k, std::vector<VoxelT<T>>(kBlockVol));
    return keys;
  }

  // voxel_hash_tsdf.h:673-682
  bool sampleNearest(const Frame<T>& f, T u, T v, T& d) const {
    const int ui = static_cast<int>(std::lround(u)), vi = static_cast<int>(std::lround(v));
    if (ui < 0 || ui >= f.K.width || vi < 0 || vi >= f.K.height) return false;
    d = f.at(ui, vi);
    return d > T(0);
  }

  // voxel_hash_tsdf.h:684-748
  bool sampleBilinear(const Frame<T>& f, T u, T v, T& d) const {
    const int W = f.K.width, H = f.K.height;
    if (W <= 0 || H <= 0) return false;
    if (u < T(0) || v < T(0) || u > static_cast<T>(W - 1) || v > static_cast<T>(H - 1)) return false;
    const int x0 = static_cast<int>(std::floor(u)), y0 = static_cast<int>(std::floor(v));
    const int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, H - 1);
    const T tx = u - static_cast<T>(x0), ty = v - static_cast<T>(y0);
    const T d00 = f.at(x0, y0), d10 = f.at(x1, y0), d01 = f.at(x0, y1), d11 = f.at(x1, y1);
    if (d00 > T(0) && d10 > T(0) && d01 > T(0) && d11 > T(0)) {
      T ref = T(0);
      if (!sampleNearest(f, u, v, ref)) ref = d00;
      const T thr = discThreshold(ref, T(P.disc_base), T(P.disc_rel));
      const T mn = smin(smin(d00, d10), smin(d01, d11));
      const T mx = smax(smax(d00, d10), smax(d01, d11));
      const T top = d00 + (d10 - d00) * tx;
      const T bottom = d01 + (d11 - d01) * tx;
      if (mx - mn <= thr) {
        d = top + (bottom - top) * ty;
        return true;
      }
      const T nd[4] = {d00, d10, d01, d11};
      const T bw[4] = {(T(1) - tx) * (T(1) - ty), tx * (T(1) - ty), (T(1) - tx) * ty, tx * ty};
      T wds = T(0), ws = T(0);
      for (int i = 0; i < 4; ++i) {
        if (std::fabs(nd[i] - ref) > thr) continue;
        wds += bw[i] * nd[i];
        ws += bw[i];
      }
      if (ws > T(1e-6f)) {
        d = wds / ws;
        return true;
      }
    }
    return sampleNearest(f, u, v, d);
  }

  // voxel_hash_tsdf.h:750-771
  bool baseNormal(const Frame<T>& f, T u, T v, V3<T>& n) const {
    T dl = 0, dr = 0, du = 0, dd = 0;
    if (!sampleBilinear(f, u - T(1), v, dl) || !sampleBilinear(f, u + T(1), v, dr) ||
        !sampleBilinear(f, u, v - T(1), du) || !sampleBilinear(f, u, v + T(1), dd))
      return false;
    const V3<T> pl = backProjectCont(u - T(1), v, dl, f.K), pr = backProjectCont(u + T(1), v, dr, f.K);
    const V3<T> pu = backProjectCont(u, v - T(1), du, f.K), pd = backProjectCont(u, v + T(1), dd, f.K);
    n = normalized(cross(pd - pu, pr - pl));
    return norm(n) > T(1e-6f);
  }

  // voxel_hash_tsdf.h:773-840
  bool surfaceNormal(const Frame<T>& f, T u, T v, V3<T>& n) const {
    V3<T> cn{};
    if (!baseNormal(f, u, v, cn)) return false;
    T cd = 0;
    if (!sampleBilinear(f, u, v, cd)) return false;
    const V3<T> cp = backProjectCont(u, v, cd, f.K);
    cn = orientTowardCamera(cn, cp);
    const T ss = T(P.normal_sigma_spatial), sr = T(P.normal_sigma_range), sa = T(P.normal_sigma_angular);
    const T inv_s = T(1) / (T(2) * ss * ss), inv_r = T(1) / (T(2) * sr * sr), inv_a = T(1) / (T(2) * sa * sa);
    V3<T> acc{};
    T ws = T(0);
    for (int dv = -1; dv <= 1; ++dv)
      for (int du = -1; du <= 1; ++du) {
        const T su = u + static_cast<T>(du), sv = v + static_cast<T>(dv);
        V3<T> sn{};
        if (!baseNormal(f, su, sv, sn)) continue;
        T sd = 0;
        if (!sampleBilinear(f, su, sv, sd)) continue;
        const V3<T> sp = backProjectCont(su, sv, sd, f.K);
        sn = orientTowardCamera(sn, sp);
        const T spatial = static_cast<T>(du * du + dv * dv);
        const T delta = sd - cd;
        const T c = clampv(dot(cn, sn), T(-1), T(1));
        const T ad = T(1) - c;
        const T w = gaussian(spatial, inv_s) * gaussian(delta * delta, inv_r) * gaussian(ad * ad, inv_a);
        acc = acc + sn * w;
        ws += w;
      }
    if (ws <= T(1e-6f)) {
      n = cn;
      return true;
    }
    n = normalized(acc * (T(1) / ws));
    n = orientTowardCamera(n, cp);
    if (norm(n) <= T(1e-6f)) n = cn;
    return norm(n) > T(1e-6f);
  }

  // voxel_hash_tsdf.h:357-392
  static int findSlot(const VoxelT<T>& vx, int d, T coord, V3<T> n, T tol, T min_cos) {
    int empty = -1, best = -1;
    T best_delta = std::numeric_limits<T>::max();
    for (int k = 0; k < kSlots; ++k) {
      const LayerT<T>& l = vx.layers[d][k];
      if (!(l.weight > T(0))) {
        if (empty < 0) empty = k;
        continue;
      }
      if (dot(unpackNormal<T>(l.packed_normal), n) < min_cos) continue;
      const T delta = std::fabs(l.surface_coord - coord);
      if (delta < best_delta) {
        best_delta = delta;
        best = k;
      }
    }
    if (best >= 0 && best_delta <= tol) return best;
    return empty;
  }

  // voxel_hash_tsdf.h:341-355
  static void sortLayers(VoxelT<T>& vx, int d) {
    LayerT<T>* l = vx.layers[d];
    for (int pass = 0; pass < kSlots - 1; ++pass)
      for (int i = 0; i < kSlots - pass - 1; ++i) {
        const bool lv = l[i].weight > T(0), rv = l[i + 1].weight > T(0);
        if ((!lv && rv) || (lv && rv && l[i].surface_coord > l[i + 1].surface_coord)) std::swap(l[i], l[i + 1]);
      }
  }

  // voxel_hash_tsdf.h:479-529
  static void integrateSample(VoxelT<T>& vx, int d, T coord, V3<T> pw, V3<T> nw, T w, T tol, T min_cos, T max_w) {
    if (d < 0 || d >= kDirs || w <= T(0)) return;
    const V3<T> n = normalized(nw);
    if (norm(n) <= T(1e-6f)) return;
    const T offset = dot(n, pw);
    const int k = findSlot(vx, d, coord, n, tol, min_cos);
    if (k < 0) return;
    LayerT<T>& l = vx.layers[d][k];
    const T pw0 = l.weight;
    const T ew = smin(w, smax(T(0), max_w - pw0));
    if (ew <= T(0)) return;
    const T nw1 = pw0 + ew;
    V3<T> blended = n;
    l.surface_coord = pw0 > T(0) ? ((l.surface_coord * pw0 + coord * ew) / nw1) : coord;
    l.plane_offset = pw0 > T(0) ? ((l.plane_offset * pw0 + offset * ew) / nw1) : offset;
    if (pw0 > T(0)) {
      const V3<T> prev = unpackNormal<T>(l.packed_normal);
      blended = normalized((prev * pw0 + n * ew) / nw1);
      if (norm(blended) <= T(1e-6f)) blended = n;
    }
    l.weight = nw1;
    l.packed_normal = packNormal(blended);
    sortLayers(vx, d);
  }

  // voxel_hash_tsdf.h:985-1060
  void integrateBlocks(const std::vector<Key>& active, const Frame<T>& f, const Pose<T>& T_wc) {
    const Pose<T> T_cw = T_wc.inverse();
    const T tol = smax(T(1e-6f), voxel() * T(P.layer_merge_scale));
    for (const Key& key : active) {
      auto it = blocks.find(key);
      if (it == blocks.end()) continue;
      std::vector<VoxelT<T>>& blk = it->second;
      for (int lz = 0; lz < kBlock; ++lz)
        for (int ly = 0; ly < kBlock; ++ly)
          for (int lx = 0; lx < kBlock; ++lx) {
            const V3<T> pw = voxelToWorld(key.x * kBlock + lx, key.y * kBlock + ly, key.z * kBlock + lz);
            const V3<T> pc = T_cw.transformPoint(pw);
            if (pc.z <= T(0)) continue;
            T u = 0, v = 0;
            if (!projectPoint(pc, f.K, u, v)) continue;
            T depth = 0;
            if (!sampleBilinear(f, u, v, depth)) continue;
            const T sd = depth - pc.z;
            if (sd <= -trunc()) continue;
            const V3<T> ray = normalized(pc);
            T vaw = T(1);
            V3<T> nc = ray * T(-1);
            if (surfaceNormal(f, u, v, nc)) vaw = std::fabs(dot(nc, ray));
            const T w = observationWeight(depth, sd, trunc(), vaw, T(P.min_view_angle_weight));
            if (w <= T(0)) continue;
            const V3<T> spw = T_wc.transformPoint(backProjectCont(u, v, depth, f.K));
            const V3<T> nw = normalized(T_wc.transformVector(nc));
            const int d = dominantDirection(nw);
            if (d < 0) continue;
            const T coord = dot(spw, dirUnit<T>(d));  // coordinateAlongDirection, :328-330
            ++updates;
            integrateSample(blk[static_cast<std::size_t>(lx + kBlock * (ly + kBlock * lz))], d, coord, spw, nw, w, tol,
                            T(P.layer_normal_cos), T(P.max_weight));
          }
    }
  }

  void integrate(const Frame<T>& f, const Pose<T>& T_wc, int stride) {  // :70-73
    const std::vector<Key> active = prepare(f, T_wc, stride);
    integrateBlocks(active, f, T_wc);
  }

  // updateVoxelAtWorld (voxel_hash_tsdf.h:952-983): the block is created
  // before the weight test.
  void updateVoxelAtWorld(V3<T> pw, T depth, T sd, V3<T> spw, V3<T> nw) {
    const std::array<int, 3> v = worldToVoxel(pw);
    std::vector<VoxelT<T>>& blk = blockFor(voxelToKey(v));
    VoxelT<T>& vx = blk[static_cast<std::size_t>(positiveMod(v[0], kBlock) +
                                                 kBlock * (positiveMod(v[1], kBlock) + kBlock * positiveMod(v[2], kBlock)))];
    const T w = observationWeight(depth, sd, trunc(), T(1), T(P.min_view_angle_weight));
    if (w <= T(0)) return;
    const int d = dominantDirection(nw);
    if (d < 0) return;
    const T coord = dot(spw, dirUnit<T>(d));
    const T tol = smax(T(1e-6f), voxel() * T(P.layer_merge_scale));
    ++updates;
    integrateSample(vx, d, coord, spw, nw, w, tol, T(P.layer_normal_cos), T(P.max_weight));
  }

  // integratePointCloud (voxel_hash_tsdf.h:75-97): samples every
  // max(voxel / 2, 2.5 mm) along each point's ray across the truncation band.
  void integratePointCloud(const std::vector<V3<T>>& points_c, const Pose<T>& T_wc) {
    const V3<T> origin{T_wc.t[0], T_wc.t[1], T_wc.t[2]};
    const T step = smax(T(0.5f) * voxel(), T(0.0025f));
    for (const V3<T>& pc : points_c) {
      const T range = norm(pc);
      if (range <= T(1e-6f)) continue;
      const V3<T> ray_c = pc / range;
      const V3<T> ray_w = normalized(T_wc.transformVector(ray_c));
      const V3<T> spw = T_wc.transformPoint(pc);
      const T start = smax(T(0), range - trunc());
      const T end = range + trunc();
      for (T s = start; s <= end; s += step)
        updateVoxelAtWorld(origin + ray_w * s, range, range - s, spw, ray_w * T(-1));
    }
  }

  static bool observed(const VoxelT<T>& vx) {  // :215-224
    for (int d = 0; d < kDirs; ++d)
      for (int k = 0; k < kSlots; ++k)
        if (vx.layers[d][k].weight > T(0)) return true;
    return false;
  }

  std::size_t observedVoxelCount() const {  // :113-123
    std::size_t c = 0;
    for (const auto& e : blocks)
      for (const VoxelT<T>& vx : e.second) c += observed(vx) ? 1 : 0;
    return c;
  }

  const VoxelT<T>* findVoxel(int x, int y, int z) const {  // :941-950
    const auto it = blocks.find(voxelToKey({x, y, z}));
    if (it == blocks.end()) return nullptr;
    return &it->second[static_cast<std::size_t>(positiveMod(x, kBlock) +
                                               kBlock * (positiveMod(y, kBlock) + kBlock * positiveMod(z, kBlock)))];
  }

  // voxel_hash_tsdf.h:394-421
  static bool selectClosest(const VoxelT<T>& vx, int d, T key, V3<T> ref, T tol, LayerT<T>& out) {
    int best = -1;
    T best_delta = std::numeric_limits<T>::max();
    for (int k = 0; k < kSlots; ++k) {
      const LayerT<T>& l = vx.layers[d][k];
      if (!(l.weight > T(0))) continue;
      const T delta = std::fabs(planeSD(l, ref) - key);
      if (delta <= tol && delta < best_delta) {
        best_delta = delta;
        best = k;
      }
    }
    if (best < 0) return false;
    out = vx.layers[d][best];
    return true;
  }

  // voxel_hash_tsdf.h:423-477
  static bool completeCube(std::array<T, 8>& values, std::array<std::uint8_t, 8>& valid) {
    int n = 0;
    T sum = T(0);
    for (int c = 0; c < 8; ++c) {
      if (valid[c] == 0) continue;
      ++n;
      sum += values[c];
    }
    if (n == 8) return true;
    if (n < 6) return false;
    const T global = n > 0 ? (sum / static_cast<T>(n)) : T(0);
    for (int c = 0; c < 8; ++c) {
      if (valid[c] != 0) continue;
      T ns = T(0);
      int nc = 0;
      for (int e = 0; e < 12; ++e) {
        const int c0 = vhm::kMcEdgeCorners[e][0], c1 = vhm::kMcEdgeCorners[e][1];
        int nb = -1;
        if (c0 == c)
          nb = c1;
        else if (c1 == c)
          nb = c0;
        if (nb < 0 || valid[nb] == 0) continue;
        ns += values[nb];
        ++nc;
      }
      values[c] = nc > 0 ? (ns / static_cast<T>(nc)) : global;
      valid[c] = 1;
    }
    return true;
  }

  // marching_cubes.h:34-94
  static void polygonise(const std::array<V3<T>, 8>& pos, const std::array<T, 8>& val, T iso,
                         std::vector<std::array<V3<T>, 3>>& tris) {
    static constexpr vhm::McTables mc = vhm::makeMcTables();
    int ci = 0;
    for (int c = 0; c < 8; ++c)
      if (val[c] < iso) ci |= 1 << c;
    const int mask = mc.edge_mask[ci];
    if (mask == 0) return;
    std::array<V3<T>, 12> ev{};
    for (int e = 0; e < 12; ++e) {
      if ((mask & (1 << e)) == 0) continue;
      const int c0 = vhm::kMcEdgeCorners[e][0], c1 = vhm::kMcEdgeCorners[e][1];
      const T delta = val[c1] - val[c0];
      if (std::fabs(delta) < T(1e-6f)) {
        ev[e] = pos[c0];
        continue;
      }
      const T t = clampv((iso - val[c0]) / delta, T(0), T(1));
      ev[e] = pos[c0] + (pos[c1] - pos[c0]) * t;
    }
    for (int i = 0; i < 16; i += 3) {
      const int e0 = mc.tri[ci][i];
      if (e0 < 0) break;
      tris.push_back({ev[e0], ev[mc.tri[ci][i + 1]], ev[mc.tri[ci][i + 2]]});
    }
  }

  // voxel_hash_tsdf.h:842-939 (one direction); extractMesh :99-107 runs d=0..5.
  void extractDirection(int d, T iso, std::vector<std::array<V3<T>, 3>>& tris) const {
    const T tol = smax(T(1e-6f), voxel() * T(P.layer_extract_scale));
    for (const auto& e : blocks) {
      const Key& key = e.first;
      for (int lz = 0; lz < kBlock; ++lz)
        for (int ly = 0; ly < kBlock; ++ly)
          for (int lx = 0; lx < kBlock; ++lx) {
            const int ox = key.x * kBlock + lx, oy = key.y * kBlock + ly, oz = key.z * kBlock + lz;
            std::array<V3<T>, 8> pos{};
            const V3<T> cc = voxelToWorld(ox, oy, oz) + V3<T>{T(0.5f) * voxel(), T(0.5f) * voxel(), T(0.5f) * voxel()};
            std::array<T, 24> cand{};
            int nc = 0;
            const VoxelT<T>* corner_vox[8];
            for (int c = 0; c < 8; ++c) {
              const int sx = ox + vhm::kMcCornerOffset[c][0], sy = oy + vhm::kMcCornerOffset[c][1],
                        sz = oz + vhm::kMcCornerOffset[c][2];
              corner_vox[c] = findVoxel(sx, sy, sz);
              pos[c] = voxelToWorld(sx, sy, sz);
              if (corner_vox[c] == nullptr) continue;
              for (int k = 0; k < kSlots; ++k) {
                const LayerT<T>& l = corner_vox[c]->layers[d][k];
                if (!(l.weight > T(0))) continue;
                if (nc < 24) cand[nc++] = planeSD(l, cc);
              }
            }
            if (nc <= 0) continue;
            std::sort(cand.begin(), cand.begin() + nc);
            std::array<T, 24> centers{};
            std::array<int, 24> counts{};
            int ncl = 0;
            for (int i = 0; i < nc; ++i) {
              const T sk = cand[i];
              if (ncl == 0 || std::fabs(sk - centers[ncl - 1]) > tol) {
                centers[ncl] = sk;
                counts[ncl] = 1;
                ++ncl;
                continue;
              }
              const int cnt = counts[ncl - 1];
              centers[ncl - 1] = (centers[ncl - 1] * static_cast<T>(cnt) + sk) / static_cast<T>(cnt + 1);
              counts[ncl - 1] = cnt + 1;
            }
            for (int ci = 0; ci < ncl; ++ci) {
              std::array<T, 8> val{};
              std::array<std::uint8_t, 8> valid{};
              for (int c = 0; c < 8; ++c) {
                if (corner_vox[c] == nullptr) continue;
                LayerT<T> l{};
                if (!selectClosest(*corner_vox[c], d, centers[ci], cc, tol, l)) continue;
                val[c] = planeSD(l, pos[c]);
                valid[c] = 1;
              }
              if (!completeCube(val, valid)) continue;
              polygonise(pos, val, iso, tris);
            }
          }
    }
  }

  std::vector<std::array<V3<T>, 3>> extractMesh(T iso) const {
    std::vector<std::array<V3<T>, 3>> tris;
    for (int d = 0; d < kDirs; ++d) extractDirection(d, iso, tris);
    return tris;
  }
};

}  // namespace oracle
