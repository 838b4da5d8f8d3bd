// Host build of the product's per-voxel integration arithmetic
// (voxelhashmt_amd/csrc/vhm_integrate_core.h) -- test infrastructure only.
// Lets the CPU test suite check, on real frames and without a GPU, that the
// deduplicated fast normal path equals the reference-order path, and that
// the whole per-voxel update equals the oracle.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../voxelhashmt_amd/csrc/vhm_integrate_core.h"

using namespace vhm;

namespace {

NormalParams makeParams(int W, int H, const float* K4, const float* p7) {
  NormalParams a;
  a.K = {W, H, K4[0], K4[1], K4[2], K4[3]};
  a.disc_base = p7[0];
  a.disc_rel = p7[1];
  a.spatial_w[0] = p7[2];
  a.spatial_w[1] = p7[3];
  a.spatial_w[2] = p7[4];
  a.inv_r = p7[5];
  a.inv_a = p7[6];
  a.inv_fx = 1.0f / a.K.fx;
  a.inv_fy = 1.0f / a.K.fy;
  a.exact_fx = reciprocalDivisionExact(a.K.fx) ? 1 : 0;
  a.exact_fy = reciprocalDivisionExact(a.K.fy) ? 1 : 0;
  a.exp_tab = kExp2Tab32;
  return a;
}

}  // namespace

extern "C" {

// For each (u, v): mode 0 = reference-order path, 1 = fast path (only where
// it applies; rows where it does not are reported with applies = 0).
// out rows: [applies, sample_ok, normal_ok, nx, ny, nz].
void vcheck_normals(const float* depth, int W, int H, const float* K4, const float* p7, const float* uv,
                    std::size_t n, int mode, float* out) {
  const NormalParams a = makeParams(W, H, K4, p7);
  const DepthView<float> D{depth, W, H};
  for (std::size_t i = 0; i < n; ++i) {
    const float u = uv[2 * i], v = uv[2 * i + 1];
    float* o = out + 6 * i;
    std::memset(o, 0, 6 * sizeof(float));
    float cd = 0.0f;
    const bool sok = sampleBilinear(D, a, u, v, cd);
    const bool applies = fastNormalApplies(u, v);
    o[0] = applies ? 1.0f : 0.0f;
    o[1] = sok ? 1.0f : 0.0f;
    if (!sok) continue;
    F3 nn{-1.0f, -2.0f, -3.0f};
    bool ok = false;
    if (mode == 0)
      ok = surfaceNormalReference(D, a, u, v, cd, nn);
    else if (applies)
      ok = surfaceNormalFast(D, a, u, v, cd, nn);
    else
      continue;
    o[2] = ok ? 1.0f : 0.0f;
    o[3] = nn.x;
    o[4] = nn.y;
    o[5] = nn.z;
  }
}

// Integrates one (filtered) frame into the given blocks held in the
// reference's AoS layout (n x 512 voxels x 6 dirs x 3 slots x 4 words),
// with the product's per-voxel code.  Returns the number of updates.
uint64_t vcheck_integrate_blocks(void* blocks_aos, const int32_t* keys, std::size_t n, const float* depth, int W,
                                 int H, const float* K4, const float* p7, const float* pose12, float voxel,
                                 float trunc, float merge_tol, float min_cos, float max_weight, float min_view) {
  const NormalParams a = makeParams(W, H, K4, p7);
  const DepthView<float> D{depth, W, H};
  Pose34 T_wc;
  for (int i = 0; i < 9; ++i) T_wc.R[i] = pose12[i];
  for (int i = 0; i < 3; ++i) T_wc.t[i] = pose12[9 + i];
  const Pose34 T_cw = inversePose(T_wc);
  auto* words = static_cast<uint32_t*>(blocks_aos);
  uint64_t updates = 0;
  for (std::size_t b = 0; b < n; ++b)
    for (int vi = 0; vi < 512; ++vi) {
      const int lx = vi & 7, ly = (vi >> 3) & 7, lz = vi >> 6;
      const F3 pw{(static_cast<float>(keys[3 * b] * 8 + lx) + 0.5f) * voxel,
                  (static_cast<float>(keys[3 * b + 1] * 8 + ly) + 0.5f) * voxel,
                  (static_cast<float>(keys[3 * b + 2] * 8 + lz) + 0.5f) * voxel};
      const VoxelObservation o = observeVoxel(D, a, T_wc, T_cw, pw, trunc, min_view);
      if (o.dir < 0) continue;
      ++updates;
      uint32_t* vox = words + (b * 512 + static_cast<std::size_t>(vi)) * 72;
      Slots L;
      for (int k = 0; k < 3; ++k) {
        const uint32_t* l = vox + (o.dir * 3 + k) * 4;
        std::memcpy(&L.w[k], l + 0, 4);
        std::memcpy(&L.sc[k], l + 1, 4);
        std::memcpy(&L.po[k], l + 2, 4);
        L.pn[k] = l[3];
      }
      if (integrateSample(L, o.coord, o.pw, o.nw, o.weight, merge_tol, min_cos, max_weight)) {
        for (int k = 0; k < 3; ++k) {
          uint32_t* l = vox + (o.dir * 3 + k) * 4;
          std::memcpy(l + 0, &L.w[k], 4);
          std::memcpy(l + 1, &L.sc[k], 4);
          std::memcpy(l + 2, &L.po[k], 4);
          l[3] = L.pn[k];
        }
      }
    }
  return updates;
}

// Largest |unpackNormalOct(packNormalOct(n)) - n| (Euclidean) over the given
// normals (each normalised first, as integrateSample's blended normal is):
// bounds the error of the integrator's near-iso sign hints.
float vic_oct_roundtrip_max_err(const float* n3, size_t count) {
  float worst = 0.0f;
  for (size_t i = 0; i < count; ++i) {
    const F3 n = normalized3(F3{n3[3 * i], n3[3 * i + 1], n3[3 * i + 2]});
    const F3 m = unpackNormalOct(packNormalOct(n));
    const float e = norm3(sub3(m, n));
    if (e > worst) worst = e;
  }
  return worst;
}

// The slot search's normal test: for each (packed, n, min_cos), the filtered
// decision octNormalCosBelowWith (with an rsqrt moved `rsqrt_ulps` ulp off
// the host's 1/sqrtf, as the GPU's may be) and the reference computation's
// decision; cos_out = the reference's rounded dot product.
void vic_oct_cos_below(const uint32_t* packed, const float* n3, const float* min_cos, size_t count, int rsqrt_ulps,
                       uint8_t* fast_out, uint8_t* exact_out, float* cos_out) {
  const auto rsqrt = [rsqrt_ulps](float x) {
    float r = 1.0f / sqrtf(x);
    for (int i = 0; i < rsqrt_ulps; ++i) r = std::nextafter(r, 2.0f * r);
    for (int i = 0; i < -rsqrt_ulps; ++i) r = std::nextafter(r, 0.0f);
    return r;
  };
  for (size_t i = 0; i < count; ++i) {
    const F3 n{n3[3 * i], n3[3 * i + 1], n3[3 * i + 2]};
    const float c = dot3(unpackNormalOct(packed[i]), n);
    cos_out[i] = c;
    exact_out[i] = c < min_cos[i];
    fast_out[i] = octNormalCosBelowWith(packed[i], n, min_cos[i], rsqrt);
  }
}

}  // extern "C"
