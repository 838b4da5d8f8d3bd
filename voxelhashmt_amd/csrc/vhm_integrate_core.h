// Per-voxel integration arithmetic (host/device), shared by the HIP kernel
// in vhm_integrate.hip and the host-side checker that compares its fast
// path with the reference-order path on real frames.
//
// Reference: integrateBlocks (voxel_hash_tsdf.h:985-1060) and the helpers
// it calls; every function keeps the reference's float operation order.
//
// Fast normal path.  estimateSurfaceNormalCamera (voxel_hash_tsdf.h:773-840)
// evaluates 45 bilinear depth samples at u + a, v + b for integer offsets
// |a|,|b| <= 2 (21 distinct offsets), each offset reached through one of a
// few float expressions ((u + 1) - 1, u + 1, (u - 1) - 1, ...).  When
// u >= 2, v >= 2 and u + 2, v + 2 are exact, every one of those expressions
// is exact, so every sample of the same offset has the same
// coordinates -- and therefore the same value -- and all share the
// fractional parts tx = u - floor(u), ty = v - floor(v).  The fast path then
// evaluates the 21 samples once from a 6x6 depth patch held in registers.
// Otherwise (about 1 % of voxels, next to pixel coordinates 2^k) the
// reference sequence is run as written.  A sample whose shifted coordinate
// is negative fails in both paths, so the u < 2 guard only keeps the
// equality argument simple.
#pragma once

#include "vhm_math.h"

#pragma clang fp contract(off)

namespace vhm {

struct NormalParams {
  Intrinsics K;
  float disc_base, disc_rel;
  float spatial_w[3];  // exp(-s * inv_s) for s = du^2 + dv^2 in {0, 1, 2}
  float inv_r, inv_a;
  // RN(1 / fx), RN(1 / fy) and whether divByConst by fx / fy is exact
  // (reciprocalDivisionExact); otherwise the plain division is used
  float inv_fx, inv_fy;
  int exact_fx, exact_fy;
  const uint64_t* exp_tab;  // kExp2Tab32 or a copy (the integrator keeps one in LDS)
};

// backProjectContinuous (voxel_hash_tsdf.h:586-593) with the divisions by the
// focal lengths through divByConst when that is exact for them.
VHM_HD F3 backProjectN(float u, float v, float d, const NormalParams& a) {
  const float xn = (u - a.K.cx) * d, yn = (v - a.K.cy) * d;
  return {a.exact_fx ? divByConst(xn, a.K.fx, a.inv_fx) : xn / a.K.fx,
          a.exact_fy ? divByConst(yn, a.K.fy, a.inv_fy) : yn / a.K.fy, d};
}

template <class DepthT>
struct DepthView {
  const DepthT* d;
  int W, H;
  VHM_HD float at(int u, int v) const { return d[static_cast<size_t>(v) * W + u]; }
};

// ------------------------------------------------ reference-order path --

// voxel_hash_tsdf.h:673-682
template <class D>
VHM_HD bool sampleNearest(const D& dv, float u, float v, float& out) {
  const int ui = lroundi(u), vi = lroundi(v);
  if (ui < 0 || ui >= dv.W || vi < 0 || vi >= dv.H) return false;
  out = dv.at(ui, vi);
  return out > 0.0f;
}

// Four-tap part of sampleDepthBilinear once the taps are known
// (voxel_hash_tsdf.h:707-747).  `nearest` is the depth at the rounded
// coordinate (one of the four taps) and `nearest_in_range` its bounds test.
VHM_HD bool bilinearFromTaps(float d00, float d10, float d01, float d11, float tx, float ty, float nearest,
                             bool nearest_in_range, float disc_base, float disc_rel, float& out) {
  if (d00 > 0.0f && d10 > 0.0f && d01 > 0.0f && d11 > 0.0f) {
    const float ref = (nearest_in_range && nearest > 0.0f) ? nearest : d00;
    const float thr = discontinuityThreshold(ref, disc_base, disc_rel);
    // all four taps are > 0 here (no NaN, no signed zero), where the IEEE
    // min/max instructions agree with std::min/std::max
    const float mn = fminf(fminf(d00, d10), fminf(d01, d11));
    const float mx = fmaxf(fmaxf(d00, d10), fmaxf(d01, d11));
    const float top = d00 + (d10 - d00) * tx;
    const float bottom = d01 + (d11 - d01) * tx;
    if (mx - mn <= thr) {
      out = top + (bottom - top) * ty;
      return true;
    }
    const float w00 = (1.0f - tx) * (1.0f - ty), w10 = tx * (1.0f - ty), w01 = (1.0f - tx) * ty, w11 = tx * ty;
    float wds = 0.0f, ws = 0.0f;
    if (!(fabsf(d00 - ref) > thr)) { wds += w00 * d00; ws += w00; }
    if (!(fabsf(d10 - ref) > thr)) { wds += w10 * d10; ws += w10; }
    if (!(fabsf(d01 - ref) > thr)) { wds += w01 * d01; ws += w01; }
    if (!(fabsf(d11 - ref) > thr)) { wds += w11 * d11; ws += w11; }
    if (ws > 1e-6f) {
      out = wds / ws;
      return true;
    }
  }
  if (!nearest_in_range) return false;
  out = nearest;
  return nearest > 0.0f;
}

// voxel_hash_tsdf.h:684-748
template <class D>
VHM_HD bool sampleBilinear(const D& dv, const NormalParams& a, float u, float v, float& out) {
  if (dv.W <= 0 || dv.H <= 0) return false;
  if (u < 0.0f || v < 0.0f || u > static_cast<float>(dv.W - 1) || v > static_cast<float>(dv.H - 1)) return false;
  const int x0 = static_cast<int>(floorf(u)), y0 = static_cast<int>(floorf(v));
  const int x1 = imin(x0 + 1, dv.W - 1), y1 = imin(y0 + 1, dv.H - 1);
  const float tx = u - static_cast<float>(x0), ty = v - static_cast<float>(y0);
  const float d00 = dv.at(x0, y0), d10 = dv.at(x1, y0), d01 = dv.at(x0, y1), d11 = dv.at(x1, y1);
  float nearest = 0.0f;
  const int ui = lroundi(u), vi = lroundi(v);
  const bool in_range = !(ui < 0 || ui >= dv.W || vi < 0 || vi >= dv.H);
  if (in_range) nearest = dv.at(ui, vi);
  return bilinearFromTaps(d00, d10, d01, d11, tx, ty, nearest, in_range, a.disc_base, a.disc_rel, out);
}

// voxel_hash_tsdf.h:750-771
template <class D>
VHM_HD bool baseNormal(const D& dv, const NormalParams& a, float u, float v, F3& n) {
  float dl = 0.0f, dr = 0.0f, du = 0.0f, dd = 0.0f;
  if (!sampleBilinear(dv, a, u - 1.0f, v, dl) || !sampleBilinear(dv, a, u + 1.0f, v, dr) ||
      !sampleBilinear(dv, a, u, v - 1.0f, du) || !sampleBilinear(dv, a, u, v + 1.0f, dd))
    return false;
  const F3 pl = backProjectN(u - 1.0f, v, dl, a), pr = backProjectN(u + 1.0f, v, dr, a);
  const F3 pu = backProjectN(u, v - 1.0f, du, a), pd = backProjectN(u, v + 1.0f, dd, a);
  n = normalized3(cross3(sub3(pd, pu), sub3(pr, pl)));
  return !normAtMost1em6(n);
}

// voxel_hash_tsdf.h:773-840.  `cd` = sampleDepthBilinear(u, v), already
// computed by the caller with the same arguments.
template <class D>
VHM_HD bool surfaceNormalReference(const D& dv, const NormalParams& a, float u, float v, float cd, F3& n) {
  F3 cn;
  if (!baseNormal(dv, a, u, v, cn)) return false;
  const F3 cp = backProjectN(u, v, cd, a);
  cn = orientTowardCamera(cn, cp);
  F3 acc{0.0f, 0.0f, 0.0f};
  float ws = 0.0f;
  for (int dvv = -1; dvv <= 1; ++dvv)
    for (int duu = -1; duu <= 1; ++duu) {
      const float su = u + static_cast<float>(duu), sv = v + static_cast<float>(dvv);
      F3 sn;
      if (!baseNormal(dv, a, su, sv, sn)) continue;
      float sdep = 0.0f;
      if (!sampleBilinear(dv, a, su, sv, sdep)) continue;
      const F3 sp = backProjectN(su, sv, sdep, a);
      sn = orientTowardCamera(sn, sp);
      const float delta = sdep - cd;
      const float c = sclamp(dot3(cn, sn), -1.0f, 1.0f);
      const float ad = 1.0f - c;
      // spatial_w[duu^2 + dvv^2] without a lane-varying index into the
      // parameter block (which would force it into scratch memory)
      const int s2 = duu * duu + dvv * dvv;
      const float sw = s2 == 0 ? a.spatial_w[0] : (s2 == 1 ? a.spatial_w[1] : a.spatial_w[2]);
      const float w = sw * gaussianWeightTab(delta * delta, a.inv_r, a.exp_tab) *
                      gaussianWeightTab(ad * ad, a.inv_a, a.exp_tab);
      acc = add3(acc, mul3(sn, w));
      ws += w;
    }
  if (ws <= 1e-6f) {
    n = cn;
    return true;
  }
  n = normalized3(mul3(acc, 1.0f / ws));
  n = orientTowardCamera(n, cp);
  if (normAtMost1em6(n)) n = cn;
  return !normAtMost1em6(n);
}

// --------------------------------------------------------- fast path ----

// u + 2 exact (checked through (u + 2) - u, itself exact for u >= 2 by
// Sterbenz) implies u + 1, (u + 1) + 1, (u + 1) - 1, u - 1, (u - 1) - 1 and
// (u - 1) + 1 are all exact too.
VHM_HD bool fastNormalApplies(float u, float v) {
  return u >= 2.0f && v >= 2.0f && (u + 2.0f) - u == 2.0f && (v + 2.0f) - v == 2.0f;
}

// 21 offsets (a, b), |a|, |b| <= 2 without the 4 corners, row-major in b.
constexpr int kTapCount = 21;
VHM_HD int tapIndex(int a, int b) {
  // row b = -2: a in -1..1 (3); b = -1..1: a in -2..2 (5 each); b = 2: a in -1..1 (3)
  return b == -2 ? (a + 1) : (b == 2 ? 18 + (a + 1) : 3 + (b + 1) * 5 + (a + 2));
}

// estimateSurfaceNormalCamera with the 21 distinct samples evaluated once.
// Requires fastNormalApplies(u, v).
template <class D>
VHM_HD bool surfaceNormalFast(const D& dv, const NormalParams& a, float u, float v, float cd, F3& n) {
  const int x0 = static_cast<int>(floorf(u)), y0 = static_cast<int>(floorf(v));
  const float tx = u - static_cast<float>(x0), ty = v - static_cast<float>(y0);
  // Coordinates exactly as the reference forms them (all exact here).
  const float X[5] = {(u - 1.0f) - 1.0f, u - 1.0f, u, u + 1.0f, (u + 1.0f) + 1.0f};
  const float Y[5] = {(v - 1.0f) - 1.0f, v - 1.0f, v, v + 1.0f, (v + 1.0f) + 1.0f};
  const float Wm1 = static_cast<float>(dv.W - 1), Hm1 = static_cast<float>(dv.H - 1);
  // 6x6 patch of depth, columns x0-2..x0+3, rows y0-2..y0+3, clamped to the
  // image (a clamped column is exactly the reference's x1 = min(x0+1, W-1)).
  float P[6][6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int yy = imin(y0 - 2 + j < 0 ? 0 : y0 - 2 + j, dv.H - 1);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int xx = imin(x0 - 2 + i < 0 ? 0 : x0 - 2 + i, dv.W - 1);
      P[j][i] = ((i == 0 || i == 5) && (j == 0 || j == 5)) ? 0.0f : dv.at(xx, yy);
    }
  }
  const bool rx = tx >= 0.5f, ry = ty >= 0.5f;  // lround picks the +1 tap
  float S[kTapCount];
  bool ok[kTapCount];
#pragma unroll
  for (int b = -2; b <= 2; ++b)
#pragma unroll
    for (int a2 = -2; a2 <= 2; ++a2) {
      if ((a2 == -2 || a2 == 2) && (b == -2 || b == 2)) continue;
      const int t = tapIndex(a2, b);
      const float xs = X[a2 + 2], ys = Y[b + 2];
      bool valid = !(xs < 0.0f || ys < 0.0f || xs > Wm1 || ys > Hm1);
      float out = 0.0f;
      if (valid) {
        const float d00 = P[b + 2][a2 + 2], d10 = P[b + 2][a2 + 3], d01 = P[b + 3][a2 + 2], d11 = P[b + 3][a2 + 3];
        const float nearest = rx ? (ry ? d11 : d10) : (ry ? d01 : d00);
        valid = bilinearFromTaps(d00, d10, d01, d11, tx, ty, nearest, true, a.disc_base, a.disc_rel, out);
      }
      S[t] = out;
      ok[t] = valid;
    }
  // back-projections of the 21 positions (backProjectContinuous)
  F3 Pt[kTapCount];
#pragma unroll
  for (int b = -2; b <= 2; ++b)
#pragma unroll
    for (int a2 = -2; a2 <= 2; ++a2) {
      if ((a2 == -2 || a2 == 2) && (b == -2 || b == 2)) continue;
      const int t = tapIndex(a2, b);
      Pt[t] = backProjectN(X[a2 + 2], Y[b + 2], S[t], a);
    }
  // base normals of the 9 neighbours (du, dv) = (i - 1, j - 1)
  F3 bn[9];
  bool bok[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int du = i - 1, dvv = j - 1;
      const int tl = tapIndex(du - 1, dvv), tr = tapIndex(du + 1, dvv), tu = tapIndex(du, dvv - 1),
                td = tapIndex(du, dvv + 1);
      bool good = ok[tl] && ok[tr] && ok[tu] && ok[td];
      F3 nn{0.0f, 0.0f, 0.0f};
      if (good) {
        nn = normalized3(cross3(sub3(Pt[td], Pt[tu]), sub3(Pt[tr], Pt[tl])));
        good = !normAtMost1em6(nn);
      }
      bn[j * 3 + i] = nn;
      bok[j * 3 + i] = good;
    }
  if (!bok[4]) return false;
  const F3 cp = Pt[tapIndex(0, 0)];  // backProjectCont(u, v, cd): S at (0,0) == cd
  const F3 cn = orientTowardCamera(bn[4], cp);
  F3 acc{0.0f, 0.0f, 0.0f};
  float ws = 0.0f;
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int du = i - 1, dvv = j - 1;
      const int t = tapIndex(du, dvv);
      if (!bok[j * 3 + i] || !ok[t]) continue;
      const F3 sn = orientTowardCamera(bn[j * 3 + i], Pt[t]);
      const float delta = S[t] - cd;
      const float c = sclamp(dot3(cn, sn), -1.0f, 1.0f);
      const float ad = 1.0f - c;
      const float w = a.spatial_w[du * du + dvv * dvv] * gaussianWeightTab(delta * delta, a.inv_r, a.exp_tab) *
                      gaussianWeightTab(ad * ad, a.inv_a, a.exp_tab);
      acc = add3(acc, mul3(sn, w));
      ws += w;
    }
  if (ws <= 1e-6f) {
    n = cn;
    return true;
  }
  n = normalized3(mul3(acc, 1.0f / ws));
  n = orientTowardCamera(n, cp);
  if (normAtMost1em6(n)) n = cn;
  return !normAtMost1em6(n);
}

template <class D>
VHM_HD bool surfaceNormal(const D& dv, const NormalParams& a, float u, float v, float cd, F3& n) {
  if (fastNormalApplies(u, v)) return surfaceNormalFast(dv, a, u, v, cd, n);
  return surfaceNormalReference(dv, a, u, v, cd, n);
}

// ------------------------------------------------------- layer update ---

// voxel_hash_tsdf.h:556-584
VHM_HD float observationWeight(float depth, float sd, float trunc, float vaw, float min_view) {
  if (depth <= 0.0f || trunc <= 1e-6f) return 0.0f;
  const float ratio = depth / 1.5f;
  const float dw = 1.0f / (1.0f + ratio * ratio);
  const float r = fabsf(sd) / trunc;
  float robust = 1.0f;
  if (r > 0.35f) robust = 0.35f / r;
  const float band = 1.0f - 0.5f * sclamp(r, 0.0f, 1.0f);
  const float base = smax(0.05f, dw * robust * band);
  return base * sclamp(vaw, min_view, 1.0f);
}

// voxel_hash_tsdf.h:239-254
VHM_HD int dominantDirection(F3 n) {
  const float nx = fabsf(n.x), ny = fabsf(n.y), nz = fabsf(n.z);
  if (nx <= 1e-6f && ny <= 1e-6f && nz <= 1e-6f) return -1;
  if (nx >= ny && nx >= nz) return n.x >= 0.0f ? 0 : 1;
  if (ny >= nx && ny >= nz) return n.y >= 0.0f ? 2 : 3;
  return n.z >= 0.0f ? 4 : 5;
}

// coordinateAlongDirection = dot(p, unit axis) (voxel_hash_tsdf.h:311-330).
VHM_HD float coordinateAlong(F3 p, int d) {
  const float s = (d & 1) ? -1.0f : 1.0f;
  const F3 e = d < 2 ? F3{s, 0.0f, 0.0f} : (d < 4 ? F3{0.0f, s, 0.0f} : F3{0.0f, 0.0f, s});
  return dot3(p, e);
}

struct Slots {
  float w[3], sc[3], po[3];
  uint32_t pn[3];
};

VHM_HD void swapSlot(Slots& L, int i, int j) {
  const float w = L.w[i], sc = L.sc[i], po = L.po[i];
  const uint32_t pn = L.pn[i];
  L.w[i] = L.w[j]; L.sc[i] = L.sc[j]; L.po[i] = L.po[j]; L.pn[i] = L.pn[j];
  L.w[j] = w; L.sc[j] = sc; L.po[j] = po; L.pn[j] = pn;
}

// integrateDirectionalSample (voxel_hash_tsdf.h:479-529) on one direction's
// three slots.  Returns true when a slot was written.
// n_out / po_out (optional): the written slot's unit normal before packing
// and its plane offset.
VHM_HD bool integrateSample(Slots& L, float coord, F3 pw, F3 nw, float w, float merge_tol, float min_cos,
                            float max_weight, F3* n_out = nullptr, float* po_out = nullptr) {
  if (w <= 0.0f) return false;
  const F3 n = normalized3(nw);
  if (normAtMost1em6(n)) return false;
  const float offset = dot3(n, pw);
  // findLayerSlotForObservation (357-392)
  int empty = -1, best = -1;
  float best_delta = 3.402823466e+38f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (!(L.w[k] > 0.0f)) {
      if (empty < 0) empty = k;
      continue;
    }
    if (octNormalCosBelow(L.pn[k], n, min_cos)) continue;
    const float delta = fabsf(L.sc[k] - coord);
    if (delta < best_delta) {
      best_delta = delta;
      best = k;
    }
  }
  const int k = (best >= 0 && best_delta <= merge_tol) ? best : empty;
  if (k < 0) return false;
  float pw0 = 0.0f, sc0 = 0.0f, po0 = 0.0f;
  uint32_t pn0 = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if (j == k) {
      pw0 = L.w[j]; sc0 = L.sc[j]; po0 = L.po[j]; pn0 = L.pn[j];
    }
  const float ew = smin(w, smax(0.0f, max_weight - pw0));
  if (ew <= 0.0f) return false;
  const float nw1 = pw0 + ew;
  F3 blended = n;
  const float sc1 = pw0 > 0.0f ? ((sc0 * pw0 + coord * ew) / nw1) : coord;
  const float po1 = pw0 > 0.0f ? ((po0 * pw0 + offset * ew) / nw1) : offset;
  if (pw0 > 0.0f) {
    const F3 prev = unpackNormalOct(pn0);
    blended = normalized3(div3(add3(mul3(prev, pw0), mul3(n, ew)), nw1));
    if (normAtMost1em6(blended)) blended = n;
  }
  const uint32_t pn1 = packNormalOct(blended);
  if (n_out) *n_out = blended;
  if (po_out) *po_out = po1;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    if (j == k) {
      L.w[j] = nw1; L.sc[j] = sc1; L.po[j] = po1; L.pn[j] = pn1;
    }
  // sortDirectionalLayers (341-355): observed first, ascending surface_coord.
#pragma unroll
  for (int pass = 0; pass < 2; ++pass)
#pragma unroll
    for (int i = 0; i < 2 - pass; ++i) {
      const bool lv = L.w[i] > 0.0f, rv = L.w[i + 1] > 0.0f;
      if ((!lv && rv) || (lv && rv && L.sc[i] > L.sc[i + 1])) swapSlot(L, i, i + 1);
    }
  return true;
}

struct VoxelObservation {
  int dir;       // -1: no update
  float coord;   // coordinateAlongDirection
  F3 pw, nw;     // surface point and normal in the world
  float weight;
};

// First half of a voxel's observation (voxel_hash_tsdf.h:1001-1021):
// camera-space centre, projection, bilinear depth and the truncation test.
// Returns false when the voxel is skipped; the outputs feed observeSurface.
template <class D>
VHM_HD bool observeDepth(const D& dv, const NormalParams& a, const Pose34& T_cw, F3 pw, float trunc, F3& pc,
                         float& u, float& v, float& dep) {
  pc = transformPoint(T_cw, pw);
  if (pc.z <= 0.0f) return false;
  u = 0.0f;
  v = 0.0f;
  if (!projectPoint(pc, a.K, u, v)) return false;
  dep = 0.0f;
  if (!sampleBilinear(dv, a, u, v, dep)) return false;
  const float sd = dep - pc.z;
  return !(sd <= -trunc);
}

// Second half (voxel_hash_tsdf.h:1023-1046): surface normal, weight,
// surface point and direction of a voxel that passed observeDepth.
// kFastNormal: the caller guarantees fastNormalApplies(u, v).
template <class D, bool kFastNormal = false>
VHM_HD VoxelObservation observeSurface(const D& dv, const NormalParams& a, const Pose34& T_wc, F3 pc, float u,
                                       float v, float dep, float trunc, float min_view) {
  VoxelObservation o;
  o.dir = -1;
  const float sd = dep - pc.z;
  const F3 ray = normalized3(pc);
  float vaw = 1.0f;
  F3 nc = mul3(ray, -1.0f);
  bool have;
  if constexpr (kFastNormal)
    have = surfaceNormalFast(dv, a, u, v, dep, nc);
  else
    have = surfaceNormal(dv, a, u, v, dep, nc);
  if (have) vaw = fabsf(dot3(nc, ray));
  const float w = observationWeight(dep, sd, trunc, vaw, min_view);
  if (w <= 0.0f) return o;
  const F3 spw = transformPoint(T_wc, backProjectN(u, v, dep, a));
  const F3 nw = normalized3(rotate(T_wc, nc));
  const int d = dominantDirection(nw);
  if (d < 0) return o;
  o.dir = d;
  o.coord = coordinateAlong(spw, d);
  o.pw = spw;
  o.nw = nw;
  o.weight = w;
  return o;
}

// Everything integrateBlocks does for one voxel centre before
// integrateDirectionalSample (voxel_hash_tsdf.h:1001-1046).
template <class D>
VHM_HD VoxelObservation observeVoxel(const D& dv, const NormalParams& a, const Pose34& T_wc, const Pose34& T_cw,
                                     F3 pw, float trunc, float min_view) {
  F3 pc;
  float u, v, dep;
  if (!observeDepth(dv, a, T_cw, pw, trunc, pc, u, v, dep)) {
    VoxelObservation o;
    o.dir = -1;
    return o;
  }
  return observeSurface(dv, a, T_wc, pc, u, v, dep, trunc, min_view);
}

}  // namespace vhm
