// Checker-only harness around the reference demo and dataset loader
// (HM-ZC/voxelhashmt src/main.cpp, include/tsdfmc/dataset_io.h).
//
// TEST INFRASTRUCTURE.  Only tests/ load the library built from this file,
// as the checker of voxelhashmt_amd/csrc/vhm_io.cpp -- never as part of the
// product.  The reference's sources are compiled where they lie
// (oracle/Makefile passes -I<reference>/include -I<reference>/src); nothing
// of them is copied here.  main.cpp is included with its `main` renamed so
// that the functions of its anonymous namespace (makeLookAtPose, traceScene,
// renderSyntheticDepth, writeMeshAsPly) are reachable from this
// translation unit; this file only adds an extern "C" surface over them.

#define main vhm_reference_demo_main
#include "main.cpp"
#undef main

#include <cstdio>

namespace {

void copyError(const std::exception& e, char* err, size_t cap) {
  if (err && cap) std::snprintf(err, cap, "%s", e.what());
}

Pose poseFrom(const float R[9], const float t[3]) {
  Pose p;
  for (int i = 0; i < 9; ++i) p.R.data[static_cast<std::size_t>(i)] = R[i];
  p.t = Vec3f{t[0], t[1], t[2]};
  return p;
}

}  // namespace

struct ref_demo_rng {
  std::mt19937 rng;
};

extern "C" {

// makeLookAtPose (main.cpp:177-187)
void ref_demo_look_at(const float eye[3], const float target[3], float R[9], float t[3]) {
  const Pose p = makeLookAtPose(Vec3f{eye[0], eye[1], eye[2]}, Vec3f{target[0], target[1], target[2]});
  for (int i = 0; i < 9; ++i) R[i] = p.R.data[static_cast<std::size_t>(i)];
  t[0] = p.t.x;
  t[1] = p.t.y;
  t[2] = p.t.z;
}

// traceScene (main.cpp:255-272)
int ref_demo_trace(const float o[3], const float d[3], float* hit_t) {
  float t = 0.0f;
  const bool hit = traceScene(Vec3f{o[0], o[1], o[2]}, Vec3f{d[0], d[1], d[2]}, t);
  if (hit) *hit_t = t;
  return hit ? 1 : 0;
}

ref_demo_rng* ref_demo_rng_create(uint32_t seed) { return new ref_demo_rng{std::mt19937(seed)}; }
void ref_demo_rng_destroy(ref_demo_rng* r) { delete r; }

// renderSyntheticDepth (main.cpp:295-335) with the caller's engine
void ref_demo_render(ref_demo_rng* r, int w, int h, const float K4[4], const float R[9], const float t[3],
                     int frame_index, float* out) {
  Intrinsics K;
  K.width = w;
  K.height = h;
  K.fx = K4[0];
  K.fy = K4[1];
  K.cx = K4[2];
  K.cy = K4[3];
  const DepthFrame f = renderSyntheticDepth(K, poseFrom(R, t), frame_index, r->rng);
  std::memcpy(out, f.depth.data(), sizeof(float) * f.depth.size());
}

// writeMeshAsPly (main.cpp:337-398)
int ref_write_ply(const char* path, const float* tris, size_t n) {
  std::vector<Triangle> v(n);
  for (size_t i = 0; i < n; ++i) {
    const float* p = tris + 9 * i;
    v[i].v0 = Vec3f{p[0], p[1], p[2]};
    v[i].v1 = Vec3f{p[3], p[4], p[5]};
    v[i].v2 = Vec3f{p[6], p[7], p[8]};
  }
  return writeMeshAsPly(path, v) ? 1 : 0;
}

// readNumericValuesFromFile + parseIntrinsicsMetadataValues
// (dataset_io.h:38-56, 95-133); 0 on success, -1 with err on a throw.
int ref_parse_intrinsics_file(const char* path, float K4[4], int wh[2], float* depth_scale, int* has_resolution,
                              char* err, size_t cap) {
  try {
    const std::vector<double> values = readNumericValuesFromFile(path);
    Intrinsics K;
    bool has = false;
    float scale = 0.0f;
    parseIntrinsicsMetadataValues(values, path, K, scale, has);
    K4[0] = K.fx;
    K4[1] = K.fy;
    K4[2] = K.cx;
    K4[3] = K.cy;
    wh[0] = K.width;
    wh[1] = K.height;
    *depth_scale = scale;
    *has_resolution = has ? 1 : 0;
    return 0;
  } catch (const std::exception& e) {
    copyError(e, err, cap);
    return -1;
  }
}

// loadPoseFromFile (dataset_io.h:165-196)
int ref_load_pose(const char* path, float R[9], float t[3], char* err, size_t cap) {
  try {
    const Pose p = loadPoseFromFile(path);
    for (int i = 0; i < 9; ++i) R[i] = p.R.data[static_cast<std::size_t>(i)];
    t[0] = p.t.x;
    t[1] = p.t.y;
    t[2] = p.t.z;
    return 0;
  } catch (const std::exception& e) {
    copyError(e, err, cap);
    return -1;
  }
}

// quaternionToMatrix (math_types.h:158-178)
void ref_quaternion_to_matrix(float w, float x, float y, float z, float R[9]) {
  Quaternionf q;
  q.w = w;
  q.x = x;
  q.y = y;
  q.z = z;
  const Mat3f m = quaternionToMatrix(q);
  for (int i = 0; i < 9; ++i) R[i] = m.data[static_cast<std::size_t>(i)];
}

// listDatasetFrameIds (dataset_io.h:146-163)
int ref_list_frame_ids(const char* root, int* ids, size_t cap, size_t* n, char* err, size_t ecap) {
  try {
    const std::vector<int> v = listDatasetFrameIds(root);
    *n = v.size();
    for (size_t i = 0; i < v.size() && i < cap; ++i) ids[i] = v[i];
    return 0;
  } catch (const std::exception& e) {
    copyError(e, err, ecap);
    return -1;
  }
}

// findDatasetIntrinsicsMetadataPath (dataset_io.h:65-82)
int ref_find_intrinsics_path(const char* preferred, char* out, size_t cap, char* err, size_t ecap) {
  try {
    const std::string p = findDatasetIntrinsicsMetadataPath(preferred).string();
    std::snprintf(out, cap, "%s", p.c_str());
    return 0;
  } catch (const std::exception& e) {
    copyError(e, err, ecap);
    return -1;
  }
}

}  // extern "C"
