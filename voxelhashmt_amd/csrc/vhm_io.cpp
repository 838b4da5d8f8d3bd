// Host-side companions of the fusion path (include/vhm_io.h): the reference
// demo's synthetic scene and PLY writer, and its dataset loader with a PNG
// decoder of our own (zlib inflate + the PNG filters and Adam7 interlace of
// the PNG specification, ISO/IEC 15948).
//
// Built with g++ -O2 -ffp-contract=off for x86-64 without FMA, the flags of
// the reference's own CPU build, so that the float operations below, written
// in the reference's evaluation order, round identically (the renderer's
// depths and the PLY normals are bit-for-bit the reference's).  The noise of
// the synthetic demo comes from the standard library's std::mt19937,
// std::normal_distribution<float> and std::uniform_real_distribution<float>,
// the generators the reference draws from, in the same order.

#include "../../include/vhm_io.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace fs = std::filesystem;

namespace {

thread_local std::string g_error;

struct IoError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_error.clear();
    return VHM_OK;
  } catch (const IoError& e) {
    g_error = e.what();
    return VHM_ERR_IO;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return VHM_ERR_INVALID_ARGUMENT;
  } catch (const std::out_of_range& e) {
    g_error = e.what();
    return VHM_ERR_IO;
  } catch (const std::exception& e) {
    g_error = e.what();
    return VHM_ERR_IO;
  }
}

// ------------------------------------------------------------ vectors ---
// Float operations in the order math_types.h evaluates them.

struct V3 {
  float x, y, z;
};

V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float dotv(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 crossv(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float normv(V3 v) { return std::sqrt(dotv(v, v)); }
V3 unit(V3 v) {  // normalized (math_types.h:97-103)
  const float len = normv(v);
  if (len <= 1e-8f) return {0.0f, 0.0f, 0.0f};
  return {v.x / len, v.y / len, v.z / len};
}
float axis(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

struct Rigid {  // row-major R, translation t
  float R[9];
  V3 t;
  V3 rotate(V3 v) const {
    return {R[0] * v.x + R[1] * v.y + R[2] * v.z, R[3] * v.x + R[4] * v.y + R[5] * v.z,
            R[6] * v.x + R[7] * v.y + R[8] * v.z};
  }
  V3 apply(V3 p) const { return rotate(p) + t; }
  Rigid inverse() const {  // Pose::inverse (math_types.h:190-195)
    Rigid r;
    const int tr[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};
    for (int i = 0; i < 9; ++i) r.R[i] = R[tr[i]];
    r.t = r.rotate(t * -1.0f);
    return r;
  }
};

Rigid fromPose(const vhm_pose& p) {
  Rigid r;
  std::memcpy(r.R, p.R, sizeof(r.R));
  r.t = {p.t[0], p.t[1], p.t[2]};
  return r;
}

// ------------------------------------------------------- demo scene ---

// Ray against the demo's sphere (main.cpp:189-212).
bool hitSphere(V3 o, V3 d, V3 c, float radius, float& t) {
  const V3 oc = o - c;
  const float b = dotv(oc, d);
  const float disc = b * b - (dotv(oc, oc) - radius * radius);
  if (disc < 0.0f) return false;
  const float s = std::sqrt(disc);
  const float near_t = -b - s, far_t = -b + s;
  const float pick = near_t > 1e-4f ? near_t : far_t;
  if (pick <= 1e-4f) return false;
  t = pick;
  return true;
}

// Slab test against the demo's box (main.cpp:214-253).
bool hitBox(V3 o, V3 d, V3 lo, V3 hi, float& t) {
  float enter = 0.0f, leave = std::numeric_limits<float>::max();
  for (int i = 0; i < 3; ++i) {
    const float di = axis(d, i), oi = axis(o, i);
    if (std::fabs(di) < 1e-6f) {
      if (oi < axis(lo, i) || oi > axis(hi, i)) return false;
      continue;
    }
    const float inv = 1.0f / di;
    float a = (axis(lo, i) - oi) * inv, b = (axis(hi, i) - oi) * inv;
    if (a > b) std::swap(a, b);
    enter = (enter < a) ? a : enter;  // std::max
    leave = (b < leave) ? b : leave;  // std::min
    if (leave < enter) return false;
  }
  if (leave <= 1e-4f) return false;
  t = enter > 1e-4f ? enter : leave;
  return true;
}

bool traceDemoScene(V3 o, V3 d, float& t) {
  constexpr float kFar = std::numeric_limits<float>::max();
  t = kFar;
  float s = 0.0f;
  if (hitSphere(o, d, V3{-0.18f, 0.02f, 0.0f}, 0.34f, s) && s < t) t = s;
  if (hitBox(o, d, V3{0.02f, -0.26f, -0.20f}, V3{0.42f, 0.14f, 0.18f}, s) && s < t) t = s;
  return t < kFar;
}

Rigid lookAt(V3 eye, V3 target) {  // makeLookAtPose (main.cpp:177-187)
  const V3 up{0.0f, 1.0f, 0.0f};
  const V3 fwd = unit(target - eye);
  const V3 right = unit(crossv(up, fwd));
  const V3 down = crossv(fwd, right);
  Rigid p;
  const float R[9] = {right.x, down.x, fwd.x, right.y, down.y, fwd.y, right.z, down.z, fwd.z};
  std::memcpy(p.R, R, sizeof(R));
  p.t = eye;
  return p;
}

// --------------------------------------------------------------- PNG ---

uint32_t be32(const uint8_t* p) {
  return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}

struct PngImage {
  uint32_t width = 0, height = 0;
  int bit_depth = 0, color_type = 0, interlace = 0;
  std::vector<uint8_t> idat;
};

std::string decodeError(const std::string& action, const std::string& path, const std::string& detail) {
  return action + ": " + path + " (detail=" + detail + ")";  // makeDecodeError (dataset_io.h:198-203)
}

std::vector<uint8_t> readFile(const std::string& path, const char* action) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw IoError(decodeError(action, path, "cannot open file"));
  return std::vector<uint8_t>(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

// Chunk walk; with headers_only the IDAT payload is not collected.
PngImage parsePng(const std::string& path, bool headers_only) {
  static const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  const std::vector<uint8_t> f = readFile(path, "Failed to open depth PNG");
  if (f.size() < 8 || std::memcmp(f.data(), kSig, 8) != 0)
    throw IoError(decodeError("Failed to open depth PNG", path, "not a PNG file"));
  PngImage img;
  bool have_header = false, have_end = false;
  size_t pos = 8;
  while (pos + 12 <= f.size()) {
    const uint32_t len = be32(&f[pos]);
    if (len > f.size() - pos - 12) break;
    const uint8_t* type = &f[pos + 4];
    const uint8_t* data = &f[pos + 8];
    const uint32_t crc = be32(&f[pos + 8 + len]);
    if (uint32_t(crc32(crc32(0L, Z_NULL, 0), type, len + 4)) != crc)
      throw IoError(decodeError("Failed to access depth PNG frame", path, "chunk CRC mismatch"));
    if (std::memcmp(type, "IHDR", 4) == 0) {
      if (len != 13) throw IoError(decodeError("Failed to access depth PNG frame", path, "bad IHDR"));
      img.width = be32(data);
      img.height = be32(data + 4);
      img.bit_depth = data[8];
      img.color_type = data[9];
      img.interlace = data[12];
      if (data[10] != 0 || data[11] != 0 || img.interlace > 1 || img.width == 0 || img.height == 0 ||
          img.width > (1u << 24) || img.height > (1u << 24))
        throw IoError(decodeError("Failed to access depth PNG frame", path, "unsupported IHDR"));
      have_header = true;
      if (headers_only) return img;
    } else if (std::memcmp(type, "IDAT", 4) == 0) {
      img.idat.insert(img.idat.end(), data, data + len);
    } else if (std::memcmp(type, "IEND", 4) == 0) {
      have_end = true;
      break;
    }
    pos += 12 + size_t(len);
  }
  if (!have_header || !have_end)
    throw IoError(decodeError("Failed to access depth PNG frame", path, "truncated PNG"));
  return img;
}

uint8_t paeth(int a, int b, int c) {
  const int p = a + b - c;
  const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return uint8_t(a);
  return uint8_t(pb <= pc ? b : c);
}

// Reverses the five PNG row filters in place over `rows` rows of `row_bytes`
// bytes each (every row preceded by its filter byte); `bpp` = bytes per
// complete pixel, at least 1.
void unfilter(uint8_t* buf, size_t rows, size_t row_bytes, size_t bpp, const std::string& path) {
  std::vector<uint8_t> zero(row_bytes, 0);
  const uint8_t* prev = zero.data();
  for (size_t y = 0; y < rows; ++y) {
    uint8_t* line = buf + y * (row_bytes + 1);
    const uint8_t ft = line[0];
    uint8_t* cur = line + 1;
    switch (ft) {
      case 0:
        break;
      case 1:
        for (size_t i = bpp; i < row_bytes; ++i) cur[i] = uint8_t(cur[i] + cur[i - bpp]);
        break;
      case 2:
        for (size_t i = 0; i < row_bytes; ++i) cur[i] = uint8_t(cur[i] + prev[i]);
        break;
      case 3:
        for (size_t i = 0; i < row_bytes; ++i) {
          const int left = i >= bpp ? cur[i - bpp] : 0;
          cur[i] = uint8_t(cur[i] + ((left + prev[i]) >> 1));
        }
        break;
      case 4:
        for (size_t i = 0; i < row_bytes; ++i) {
          const int left = i >= bpp ? cur[i - bpp] : 0;
          const int upleft = i >= bpp ? prev[i - bpp] : 0;
          cur[i] = uint8_t(cur[i] + paeth(left, prev[i], upleft));
        }
        break;
      default:
        throw IoError(decodeError("Failed to copy depth PNG pixels", path, "bad filter type"));
    }
    prev = cur;
  }
}

// One sample of a row (big-endian for 16 bit), widened to 16 bits the way a
// conversion to 16-bit grayscale does it: replicate the bits (an 8-bit v
// becomes v * 257, a 1-bit 1 becomes 65535).
uint16_t sample16(const uint8_t* row, size_t index, int bits) {
  if (bits == 16) return uint16_t((row[2 * index] << 8) | row[2 * index + 1]);
  if (bits == 8) return uint16_t(row[index] * 257u);
  const size_t bit = index * size_t(bits);
  const unsigned v = (row[bit >> 3] >> (8 - bits - (bit & 7))) & ((1u << bits) - 1u);
  return uint16_t(v * (65535u / ((1u << bits) - 1u)));
}

// Grayscale PNG (color type 0, 1-16 bit; type 4, gray + alpha, 8/16 bit,
// alpha ignored) to width*height 16-bit samples.
std::vector<uint16_t> decodeGray16(const std::string& path, uint32_t& width, uint32_t& height) {
  PngImage img = parsePng(path, false);
  width = img.width;
  height = img.height;
  int channels = 0;
  if (img.color_type == 0 &&
      (img.bit_depth == 1 || img.bit_depth == 2 || img.bit_depth == 4 || img.bit_depth == 8 || img.bit_depth == 16))
    channels = 1;
  else if (img.color_type == 4 && (img.bit_depth == 8 || img.bit_depth == 16))
    channels = 2;
  else
    throw IoError(decodeError("Failed to convert depth PNG to 16-bit grayscale", path,
                              "color type " + std::to_string(img.color_type) + ", bit depth " +
                                  std::to_string(img.bit_depth)));
  const size_t bits_pp = size_t(img.bit_depth) * channels;
  const size_t bpp = std::max<size_t>(1, bits_pp / 8);
  // Adam7 passes (x0, y0, dx, dy); a non-interlaced image is one pass.
  static const int kAdam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
  const int passes = img.interlace ? 7 : 1;
  struct Pass {
    size_t w, h, row_bytes, offset;
  };
  std::vector<Pass> pass(passes);
  size_t total = 0;
  for (int p = 0; p < passes; ++p) {
    const int x0 = img.interlace ? kAdam7[p][0] : 0, y0 = img.interlace ? kAdam7[p][1] : 0;
    const int dx = img.interlace ? kAdam7[p][2] : 1, dy = img.interlace ? kAdam7[p][3] : 1;
    const size_t w = img.width > uint32_t(x0) ? (img.width - x0 + dx - 1) / dx : 0;
    const size_t h = img.height > uint32_t(y0) ? (img.height - y0 + dy - 1) / dy : 0;
    const size_t rb = (w * bits_pp + 7) / 8;
    pass[p] = {w, h, rb, total};
    if (w && h) total += h * (rb + 1);
  }
  std::vector<uint8_t> raw(total);
  uLongf got = uLongf(total);
  const int zr = uncompress(raw.data(), &got, img.idat.data(), uLong(img.idat.size()));
  if (zr != Z_OK || got != total)
    throw IoError(decodeError("Failed to copy depth PNG pixels", path,
                              zr != Z_OK ? "zlib error " + std::to_string(zr) : std::string("short image data")));
  std::vector<uint16_t> out(size_t(img.width) * img.height, 0);
  for (int p = 0; p < passes; ++p) {
    const Pass& ps = pass[p];
    if (!ps.w || !ps.h) continue;
    uint8_t* base = raw.data() + ps.offset;
    unfilter(base, ps.h, ps.row_bytes, bpp, path);
    const int x0 = img.interlace ? kAdam7[p][0] : 0, y0 = img.interlace ? kAdam7[p][1] : 0;
    const int dx = img.interlace ? kAdam7[p][2] : 1, dy = img.interlace ? kAdam7[p][3] : 1;
    for (size_t y = 0; y < ps.h; ++y) {
      const uint8_t* row = base + y * (ps.row_bytes + 1) + 1;
      uint16_t* dst = out.data() + (y0 + y * dy) * size_t(img.width);
      for (size_t x = 0; x < ps.w; ++x) dst[x0 + x * dx] = sample16(row, x * channels, img.bit_depth);
    }
  }
  return out;
}

std::vector<uint16_t> loadDepthSamples(const char* path, const vhm_intrinsics* K) {
  if (!path || !K) throw std::invalid_argument("loadDepthPng failed: null argument");
  uint32_t w = 0, h = 0;
  std::vector<uint16_t> raw = decodeGray16(path, w, h);
  if (int64_t(w) != K->width || int64_t(h) != K->height)
    throw IoError(decodeError("Depth image resolution does not match dataset intrinsics metadata", path,
                              std::to_string(w) + "x" + std::to_string(h) + " vs " + std::to_string(K->width) +
                                  "x" + std::to_string(K->height)));
  return raw;
}

// ---------------------------------------------------------- datasets ---

// Whitespace-separated numbers up to the first token that is not one
// (readNumericValuesFromFile, dataset_io.h:38-56).
std::vector<double> readNumbers(const fs::path& path) {
  std::ifstream in(path);
  if (!in) throw IoError("Failed to open metadata file: " + path.string());
  std::vector<double> v;
  double x = 0.0;
  while (in >> x) v.push_back(x);
  if (v.empty()) throw IoError("Failed to parse numeric metadata from file: " + path.string());
  return v;
}

std::vector<int> frameIds(const fs::path& root) {
  std::vector<int> ids;
  for (const auto& e : fs::directory_iterator(root)) {
    if (!e.is_directory()) continue;
    const std::string name = e.path().filename().string();
    if (!std::all_of(name.begin(), name.end(), [](unsigned char c) { return std::isdigit(c) != 0; })) continue;
    ids.push_back(std::stoi(name));
  }
  std::sort(ids.begin(), ids.end());
  return ids;
}

fs::path intrinsicsSource(const fs::path& preferred) {  // dataset_io.h:65-82
  if (fs::exists(preferred)) return preferred;
  const fs::path root = preferred.parent_path();
  for (const int id : frameIds(root)) {
    const fs::path p = root / std::to_string(id) / "params.txt";
    if (fs::exists(p)) return p;
  }
  throw IoError("Failed to locate dataset intrinsics metadata. Expected " + preferred.string() +
                " or <frame>/params.txt under " + root.string());
}

fs::path firstDepthImage(const fs::path& root) {  // dataset_io.h:84-93
  for (const int id : frameIds(root)) {
    const fs::path p = root / std::to_string(id) / "depth.png";
    if (fs::exists(p)) return p;
  }
  throw IoError("Failed to locate any dataset depth.png under " + root.string());
}

void quaternionMatrix(float qw, float qx, float qy, float qz, float* R) {
  // normalized(Quaternionf) then the rotation matrix (math_types.h:143-178)
  const float len = std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
  float w = 1.0f, x = 0.0f, y = 0.0f, z = 0.0f;
  if (!(len <= 1e-8f)) {
    w = qw / len;
    x = qx / len;
    y = qy / len;
    z = qz / len;
  }
  const float ww = w * w, xx = x * x, yy = y * y, zz = z * z;
  const float wx = w * x, wy = w * y, wz = w * z, xy = x * y, xz = x * z, yz = y * z;
  const float m[9] = {ww + xx - yy - zz,  2.0f * (xy - wz),  2.0f * (xz + wy),
                      2.0f * (xy + wz),   ww - xx + yy - zz, 2.0f * (yz - wx),
                      2.0f * (xz - wy),   2.0f * (yz + wx),  ww - xx - yy + zz};
  std::memcpy(R, m, sizeof(m));
}

vhm_pose loadPose(const fs::path& path) {
  const std::vector<double> v = readNumbers(path);
  vhm_pose p;
  if (v.size() == 7) {
    quaternionMatrix(float(v[0]), float(v[1]), float(v[2]), float(v[3]), p.R);
    p.t[0] = float(v[4]);
    p.t[1] = float(v[5]);
    p.t[2] = float(v[6]);
    return p;
  }
  if (v.size() == 12 || v.size() == 16) {
    for (int r = 0; r < 3; ++r) {
      for (int c = 0; c < 3; ++c) p.R[3 * r + c] = float(v[4 * r + c]);
      p.t[r] = float(v[4 * r + 3]);
    }
    return p;
  }
  throw IoError("Unsupported pose format in " + path.string() + " (expected 7, 12, or 16 numeric values)");
}

size_t scaleDepths(const std::vector<uint16_t>& raw, float depth_scale, float* out) {
  size_t valid = 0;
  for (size_t i = 0; i < raw.size(); ++i) {
    out[i] = 0.0f;
    if (raw[i] == 0) continue;
    const float d = static_cast<float>(raw[i]) * depth_scale;
    if (d <= 0.0f) continue;
    out[i] = d;
    ++valid;
  }
  return valid;
}

}  // namespace

struct vhm_demo_renderer {
  std::mt19937 rng;
};

extern "C" {

const char* vhm_io_last_error(void) { return g_error.c_str(); }

void vhm_demo_intrinsics(int32_t width, int32_t height, vhm_intrinsics* K) {
  K->width = width;
  K->height = height;
  K->fx = static_cast<float>(width) * 1.05f;
  K->fy = static_cast<float>(height) * 1.15f;
  K->cx = (static_cast<float>(width) - 1.0f) * 0.5f;
  K->cy = (static_cast<float>(height) - 1.0f) * 0.5f;
}

void vhm_demo_pose(int32_t frame_index, int32_t frame_count, vhm_pose* T_wc) {
  constexpr float kPi = 3.14159265358979323846f;
  const float angle = 2.0f * kPi * static_cast<float>(frame_index) / static_cast<float>(std::max(1, frame_count));
  const float radius = 1.15f;
  const V3 eye{radius * std::cos(angle), 0.18f + 0.10f * std::sin(angle * 0.5f), radius * std::sin(angle)};
  const Rigid p = lookAt(eye, V3{0.08f, -0.02f, 0.0f});
  std::memcpy(T_wc->R, p.R, sizeof(p.R));
  T_wc->t[0] = p.t.x;
  T_wc->t[1] = p.t.y;
  T_wc->t[2] = p.t.z;
}

int vhm_demo_trace(const float origin[3], const float direction[3], float* hit_t) {
  float t = 0.0f;
  const bool hit = traceDemoScene(V3{origin[0], origin[1], origin[2]}, V3{direction[0], direction[1], direction[2]}, t);
  if (hit && hit_t) *hit_t = t;
  return hit ? 1 : 0;
}

int vhm_demo_renderer_create(uint32_t seed, vhm_demo_renderer** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("vhm_demo_renderer_create failed: null output");
    *out = new vhm_demo_renderer{std::mt19937(seed)};
  });
}

void vhm_demo_renderer_destroy(vhm_demo_renderer* r) { delete r; }

int vhm_demo_render(vhm_demo_renderer* r, const vhm_intrinsics* K, const vhm_pose* T_wc, int32_t frame_index,
                    uint32_t flags, float* depth_out) {
  return guarded([&] {
    if (!r || !K || !T_wc || !depth_out || K->width <= 0 || K->height <= 0)
      throw std::invalid_argument("renderSyntheticDepth failed: invalid argument");
    const Rigid wc = fromPose(*T_wc);
    const Rigid cw = wc.inverse();
    const bool noise = (flags & VHM_DEMO_NOISE) != 0;
    // Fresh distributions per frame over the persistent engine, as in the
    // reference (a pending second Gaussian of the polar method is dropped).
    std::normal_distribution<float> gauss(0.0f, 0.0025f);
    std::uniform_real_distribution<float> unif(0.0f, 1.0f);
    const float fi = static_cast<float>(frame_index);
    for (int v = 0; v < K->height; ++v) {
      for (int u = 0; u < K->width; ++u) {
        float& out = depth_out[size_t(v) * K->width + u];
        out = 0.0f;
        // backProject(u, v, 1, K) (math_types.h:226-230)
        const V3 ray_c = unit(V3{(static_cast<float>(u) - K->cx) * 1.0f / K->fx,
                                 (static_cast<float>(v) - K->cy) * 1.0f / K->fy, 1.0f});
        const V3 ray_w = unit(wc.rotate(ray_c));
        float t = 0.0f;
        if (!traceDemoScene(wc.t, ray_w, t)) continue;
        float depth = cw.apply(wc.t + ray_w * t).z;
        if (noise) {
          const float bias = 0.0015f * std::sin(0.12f * static_cast<float>(u) + 0.17f * fi) *
                             std::cos(0.10f * static_cast<float>(v) - 0.11f * fi);
          depth += bias + gauss(r->rng);
          if (unif(r->rng) < 0.03f) depth = 0.0f;
        }
        if (depth > 0.1f && depth < 4.0f) out = depth;
      }
    }
  });
}

int vhm_write_ply(const char* path, const float* triangles, size_t n_triangles, size_t* n_written) {
  return guarded([&] {
    if (!path || (n_triangles && !triangles)) throw std::invalid_argument("writeMeshAsPly failed: null argument");
    auto vert = [&](size_t i, int k) {
      const float* p = triangles + 9 * i + 3 * k;
      return V3{p[0], p[1], p[2]};
    };
    std::vector<uint8_t> keep(n_triangles);
    size_t n = 0;
    for (size_t i = 0; i < n_triangles; ++i) {
      const V3 nrm = crossv(vert(i, 1) - vert(i, 0), vert(i, 2) - vert(i, 0));
      keep[i] = normv(nrm) > 1e-8f;
      n += keep[i];
    }
    std::FILE* f = std::fopen(path, "wb");
    if (!f) throw IoError(std::string("Failed to write mesh to ") + path);
    std::ostringstream h;
    h << "ply\nformat binary_little_endian 1.0\ncomment generated by tsdf_voxel_hash_demo\n"
      << "element vertex " << n * 3 << "\nproperty float x\nproperty float y\nproperty float z\n"
      << "property float nx\nproperty float ny\nproperty float nz\n"
      << "element face " << n << "\nproperty list uchar int vertex_indices\nend_header\n";
    const std::string header = h.str();
    bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size();
    // x86-64 is little-endian: the float and int bytes go out as they are.
    std::vector<float> buf;
    buf.reserve(18 * 4096);
    for (size_t i = 0; ok && i < n_triangles; ++i) {
      if (!keep[i]) continue;
      const V3 nrm = unit(crossv(vert(i, 1) - vert(i, 0), vert(i, 2) - vert(i, 0)));
      for (int k = 0; k < 3; ++k) {
        const V3 p = vert(i, k);
        buf.insert(buf.end(), {p.x, p.y, p.z, nrm.x, nrm.y, nrm.z});
      }
      if (buf.size() >= 18 * 4096) {
        ok = std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
        buf.clear();
      }
    }
    if (ok && !buf.empty()) ok = std::fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
    std::vector<uint8_t> faces;
    faces.reserve(13 * 4096);
    for (size_t i = 0; ok && i < n; ++i) {
      const int32_t idx[3] = {int32_t(3 * i), int32_t(3 * i + 1), int32_t(3 * i + 2)};
      faces.push_back(3);
      const uint8_t* b = reinterpret_cast<const uint8_t*>(idx);
      faces.insert(faces.end(), b, b + sizeof(idx));
      if (faces.size() >= 13 * 4096) {
        ok = std::fwrite(faces.data(), 1, faces.size(), f) == faces.size();
        faces.clear();
      }
    }
    if (ok && !faces.empty()) ok = std::fwrite(faces.data(), 1, faces.size(), f) == faces.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw IoError(std::string("Failed to write mesh to ") + path);
    if (n_written) *n_written = n;
  });
}

int vhm_dataset_frame_ids(const char* root, int32_t* ids, size_t capacity, size_t* n) {
  return guarded([&] {
    if (!root || !n || (capacity && !ids)) throw std::invalid_argument("listDatasetFrameIds failed: null argument");
    if (!fs::is_directory(root)) throw IoError(std::string("Dataset directory not found: ") + root);
    const std::vector<int> v = frameIds(root);
    *n = v.size();
    std::copy_n(v.begin(), std::min(capacity, v.size()), ids);
  });
}

int vhm_dataset_intrinsics(const char* path, vhm_intrinsics* K, float* depth_scale) {
  return guarded([&] {
    if (!path || !K || !depth_scale) throw std::invalid_argument("loadDatasetIntrinsics failed: null argument");
    const fs::path src = intrinsicsSource(path);
    const std::vector<double> v = readNumbers(src);
    vhm_intrinsics k{0, 0, 0.0f, 0.0f, 0.0f, 0.0f};
    float scale = 0.0f;
    bool has_size = false;
    if (v.size() == 5 || v.size() == 6 || v.size() == 7) {
      k.fx = float(v[0]);
      k.fy = float(v[1]);
      k.cx = float(v[2]);
      k.cy = float(v[3]);
      if (v.size() == 5) {
        scale = float(v[4]);
      } else {
        k.width = static_cast<int>(v[4]);
        k.height = static_cast<int>(v[5]);
        scale = v.size() == 7 ? float(v[6]) : 1.0f;
        has_size = true;
      }
    } else {
      throw IoError("Unsupported intrinsics metadata format in " + src.string() +
                    " (expected 5, 6, or 7 numeric values)");
    }
    if (!has_size) {
      // dataset root of the metadata file (dataset_io.h:58-63)
      const fs::path root = src.filename() == "params.txt" ? src.parent_path().parent_path() : src.parent_path();
      const PngImage img = parsePng(firstDepthImage(root).string(), true);
      k.width = int32_t(img.width);
      k.height = int32_t(img.height);
    }
    *K = k;
    *depth_scale = scale;
  });
}

int vhm_load_pose(const char* path, vhm_pose* T_wc) {
  return guarded([&] {
    if (!path || !T_wc) throw std::invalid_argument("loadPoseFromFile failed: null argument");
    *T_wc = loadPose(path);
  });
}

void vhm_quaternion_to_matrix(float w, float x, float y, float z, float R[9]) { quaternionMatrix(w, x, y, z, R); }

int vhm_png_size(const char* path, int32_t* width, int32_t* height) {
  return guarded([&] {
    if (!path || !width || !height) throw std::invalid_argument("vhm_png_size failed: null argument");
    const PngImage img = parsePng(path, true);
    *width = int32_t(img.width);
    *height = int32_t(img.height);
  });
}

int vhm_load_depth_png(const char* path, const vhm_intrinsics* K, float depth_scale, float* depth_out,
                       size_t* valid) {
  return guarded([&] {
    const std::vector<uint16_t> raw = loadDepthSamples(path, K);
    if (!depth_out) throw std::invalid_argument("loadDepthPng failed: null output");
    const size_t n = scaleDepths(raw, depth_scale, depth_out);
    if (valid) *valid = n;
  });
}

int vhm_load_depth_png_u16(const char* path, const vhm_intrinsics* K, float depth_scale, uint16_t* raw_out,
                           size_t* valid) {
  return guarded([&] {
    const std::vector<uint16_t> raw = loadDepthSamples(path, K);
    if (!raw_out) throw std::invalid_argument("loadDepthPng failed: null output");
    std::copy(raw.begin(), raw.end(), raw_out);
    if (valid) {
      size_t n = 0;
      for (const uint16_t s : raw) n += (s != 0 && static_cast<float>(s) * depth_scale > 0.0f);
      *valid = n;
    }
  });
}

int vhm_load_dataset_frame(const char* root, int32_t frame_id, const char* pose_filename,
                           const vhm_intrinsics* K, float depth_scale, vhm_pose* T_wc, float* depth_out,
                           size_t* valid) {
  return guarded([&] {
    if (!root || !pose_filename || !T_wc) throw std::invalid_argument("loadDatasetDepthFrame failed: null argument");
    const fs::path dir = fs::path(root) / std::to_string(frame_id);
    *T_wc = loadPose(dir / pose_filename);
    const std::vector<uint16_t> raw = loadDepthSamples((dir / "depth.png").string().c_str(), K);
    if (!depth_out) throw std::invalid_argument("loadDatasetDepthFrame failed: null output");
    const size_t n = scaleDepths(raw, depth_scale, depth_out);
    if (valid) *valid = n;
  });
}

}  // extern "C"
