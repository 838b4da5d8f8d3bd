// vhm_demo -- the reference demo's command line (src/main.cpp) over libvhm:
// the same options, the same synthetic scene and dataset reconstruction, the
// same printed lines, with filter -> allocate -> integrate -> marching cubes
// on the GPU.
//
//   vhm_demo --synthetic [--frames N] [--width W] [--height H]
//            [--voxel-size s] [--truncation t] [--output path.ply]
//   vhm_demo [--dataset-root path] [--frame-start N] [--frame-end N]
//            [--frame-step N] [--pose-file pose.txt|speckle_pose.txt] ...
//
// Options beyond the reference's: --max-blocks N (block pool capacity;
// default: half the free device memory) and --hash-log2 k (2^k hash
// buckets; default 22).  --backend accepts auto|cpu|gpu as the reference
// does; libvhm integrates on the GPU only, so cpu is refused.

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vhm.h"
#include "../../include/vhm_io.h"

namespace fs = std::filesystem;

namespace {

using Clock = std::chrono::steady_clock;
using Ms = std::chrono::duration<double, std::milli>;

#ifndef VHM_DEMO_DATASET_ROOT
#define VHM_DEMO_DATASET_ROOT "datasets"
#endif

struct Options {  // defaults of main.cpp:33-47
  std::string output_path = "output/reconstruction.ply";
  int frames = 28;
  int width = 128;
  int height = 96;
  float voxel_size = 0.0005f;
  float truncation = 0.0015f;
  std::string dataset_root = VHM_DEMO_DATASET_ROOT;
  int frame_start = 0;
  int frame_end = -1;
  int frame_step = 1;
  std::string pose_file = "pose.txt";
  std::string backend = "auto";
  bool synthetic = false;
  long long max_blocks = 0;
  int hash_log2 = 22;
};

std::string ms(Ms d) {
  std::ostringstream o;
  o << std::fixed << std::setprecision(3) << d.count() << " ms";
  return o.str();
}

void usage(const char* argv0) {
  std::cout << "Usage: " << argv0 << " [--output path] [--frames N] [--width W] [--height H]"
            << " [--voxel-size s] [--truncation t]\n"
            << "       " << argv0
            << " [--dataset-root path] [--frame-start N] [--frame-end N] [--frame-step N]"
            << " [--pose-file pose.txt|speckle_pose.txt] [--backend auto|cpu|gpu] [--output path]\n"
            << "       [--max-blocks N] [--hash-log2 k]\n";
  std::cout << "Default dataset root: " << VHM_DEMO_DATASET_ROOT << '\n';
  std::cout << "Use --synthetic to run the synthetic demo instead of dataset reconstruction.\n";
}

template <class T, class Parse>
bool parseNumber(const std::string& s, T& out, Parse parse) {
  try {
    out = parse(s);
    return true;
  } catch (...) {
    return false;
  }
}

Options parseArgs(int argc, char** argv) {  // main.cpp:83-152
  Options o;
  auto toInt = [](const std::string& s) { return std::stoi(s); };
  auto toFloat = [](const std::string& s) { return std::stof(s); };
  for (int i = 1; i < argc; ++i) {
    const std::string arg = argv[i];
    if (arg == "--help" || arg == "-h") {
      usage(argv[0]);
      std::exit(0);
    }
    if (arg == "--synthetic") {
      o.synthetic = true;
      continue;
    }
    if (i + 1 >= argc) {
      usage(argv[0]);
      throw std::runtime_error("Missing value for argument: " + arg);
    }
    const std::string v = argv[++i];
    if (arg == "--output") {
      o.output_path = v;
    } else if (arg == "--frames") {
      if (!parseNumber(v, o.frames, toInt)) throw std::runtime_error("Invalid integer for --frames");
    } else if (arg == "--width") {
      if (!parseNumber(v, o.width, toInt)) throw std::runtime_error("Invalid integer for --width");
    } else if (arg == "--height") {
      if (!parseNumber(v, o.height, toInt)) throw std::runtime_error("Invalid integer for --height");
    } else if (arg == "--voxel-size") {
      if (!parseNumber(v, o.voxel_size, toFloat)) throw std::runtime_error("Invalid float for --voxel-size");
    } else if (arg == "--truncation") {
      if (!parseNumber(v, o.truncation, toFloat)) throw std::runtime_error("Invalid float for --truncation");
    } else if (arg == "--dataset-root") {
      o.dataset_root = v;
    } else if (arg == "--frame-start") {
      if (!parseNumber(v, o.frame_start, toInt)) throw std::runtime_error("Invalid integer for --frame-start");
    } else if (arg == "--frame-end") {
      if (!parseNumber(v, o.frame_end, toInt)) throw std::runtime_error("Invalid integer for --frame-end");
    } else if (arg == "--frame-step") {
      if (!parseNumber(v, o.frame_step, toInt) || o.frame_step <= 0)
        throw std::runtime_error("Invalid integer for --frame-step");
    } else if (arg == "--pose-file") {
      o.pose_file = v;
    } else if (arg == "--backend") {
      std::string b = v;
      std::transform(b.begin(), b.end(), b.begin(), [](unsigned char c) { return char(std::tolower(c)); });
      if (b != "auto" && b != "cpu" && b != "gpu")
        throw std::runtime_error("Invalid value for --backend (expected auto|cpu|gpu)");
      o.backend = b;
    } else if (arg == "--max-blocks") {
      if (!parseNumber(v, o.max_blocks, [](const std::string& s) { return std::stoll(s); }) || o.max_blocks < 0)
        throw std::runtime_error("Invalid integer for --max-blocks");
    } else if (arg == "--hash-log2") {
      if (!parseNumber(v, o.hash_log2, toInt)) throw std::runtime_error("Invalid integer for --hash-log2");
    } else {
      usage(argv[0]);
      throw std::runtime_error("Unknown argument: " + arg);
    }
  }
  return o;
}

// TsdfIntegrationBackend (reconstruction_backend.cpp:222-314) with one
// libvhm volume; the CPU preference has no counterpart in libvhm.
class GpuFusion {
 public:
  GpuFusion(const Options& o, int max_width, int max_height) {
    char msg[256] = {0};
    const bool have_gpu = vhm_query_device(msg, sizeof(msg)) != 0;
    if (o.backend == "cpu")
      throw std::runtime_error("CPU backend requested, but libvhm integrates on the GPU only "
                               "(use --backend auto or gpu)");
    if (!have_gpu) {
      if (o.backend == "gpu") throw std::runtime_error(std::string("GPU backend requested but HIP is unavailable: ") + msg);
      throw std::runtime_error(std::string("No GPU available for fusion: ") + msg);
    }
    description_ = o.backend == "gpu" ? "GPU (HIP)" : "GPU (HIP, auto-selected)";
    vhm_params p;
    vhm_default_params(&p, VHM_PRESET_REFERENCE_CPU, o.voxel_size, o.truncation);
    long long blocks = o.max_blocks;
    if (blocks == 0) {
      size_t free_b = 0, total_b = 0;
      check(vhm_device_memory(&free_b, &total_b), "vhm_device_memory");
      blocks = static_cast<long long>(free_b / 2 / 147456u);
    }
    p.max_blocks = static_cast<uint32_t>(std::clamp<long long>(blocks, 1, 1ll << 31));
    p.hash_log2_buckets = static_cast<uint32_t>(o.hash_log2);
    p.max_width = max_width;
    p.max_height = max_height;
    const int rc = vhm_create(&p, &ctx_);
    if (rc != VHM_OK) throw std::runtime_error("vhm_create failed: status " + std::to_string(rc));
  }
  ~GpuFusion() {
    if (ctx_) vhm_destroy(ctx_);
  }
  GpuFusion(const GpuFusion&) = delete;
  GpuFusion& operator=(const GpuFusion&) = delete;

  const std::string& description() const { return description_; }

  void integrate(const float* depth, const vhm_intrinsics& K, const vhm_pose& T) {
    check(vhm_integrate_f32(ctx_, depth, VHM_HOST, &K, &T, 1, 0), "integrate");
    check(vhm_synchronize(ctx_), "integrate");
  }
  void integrateRaw(const uint16_t* raw, float scale, const vhm_intrinsics& K, const vhm_pose& T) {
    check(vhm_integrate_u16(ctx_, raw, scale, VHM_HOST, &K, &T, 1, 0), "integrate");
    check(vhm_synchronize(ctx_), "integrate");
  }
  size_t blockCount() {
    size_t n = 0;
    check(vhm_block_count(ctx_, &n), "block count");
    return n;
  }
  size_t observedVoxelCount() {
    size_t n = 0;
    check(vhm_observed_voxel_count(ctx_, &n), "observed voxel count");
    return n;
  }
  std::vector<float> extractMesh() {
    size_t n = 0;
    check(vhm_extract_mesh(ctx_, 0.0f, nullptr, 0, &n), "extract mesh");
    std::vector<float> tris(9 * n);
    if (n) check(vhm_extract_mesh(ctx_, 0.0f, tris.data(), n, &n), "extract mesh");
    tris.resize(9 * n);
    return tris;
  }

 private:
  void check(int rc, const char* what) {
    if (rc == VHM_OK) return;
    const std::string m = ctx_ ? vhm_last_error(ctx_) : std::string();
    throw std::runtime_error(m.empty() ? std::string(what) + " failed: status " + std::to_string(rc) : m);
  }
  vhm_ctx* ctx_ = nullptr;
  std::string description_;
};

void ioCheck(int rc) {
  if (rc != VHM_OK) throw std::runtime_error(vhm_io_last_error());
}

bool writeMesh(const fs::path& path, const std::vector<float>& tris) {
  if (path.has_parent_path()) fs::create_directories(path.parent_path());
  return vhm_write_ply(path.string().c_str(), tris.data(), tris.size() / 9, nullptr) == VHM_OK;
}

int runSynthetic(const Options& o) {  // runSyntheticDemo (main.cpp:400-480)
  const Clock::time_point start = Clock::now();
  vhm_intrinsics K;
  vhm_demo_intrinsics(o.width, o.height, &K);
  GpuFusion fusion(o, o.width, o.height);
  vhm_demo_renderer* rnd = nullptr;
  ioCheck(vhm_demo_renderer_create(42u, &rnd));
  std::unique_ptr<vhm_demo_renderer, void (*)(vhm_demo_renderer*)> guard(rnd, vhm_demo_renderer_destroy);

  std::cout << "Mode: synthetic demo\n";
  std::cout << "Frames=" << o.frames << ", resolution=" << o.width << "x" << o.height
            << ", voxel_size=" << o.voxel_size << ", truncation=" << o.truncation << '\n';
  std::cout << "Integration backend: " << fusion.description() << '\n';

  std::vector<float> depth(static_cast<size_t>(std::max(0, o.width)) * static_cast<size_t>(std::max(0, o.height)));
  Ms render_total{0.0}, integrate_total{0.0};
  for (int i = 0; i < o.frames; ++i) {
    const Clock::time_point f0 = Clock::now();
    vhm_pose T;
    vhm_demo_pose(i, o.frames, &T);
    const Clock::time_point r0 = Clock::now();
    ioCheck(vhm_demo_render(rnd, &K, &T, i, VHM_DEMO_NOISE, depth.data()));
    const Ms render = Clock::now() - r0;
    render_total += render;
    const Clock::time_point i0 = Clock::now();
    fusion.integrate(depth.data(), K, T);
    const Ms integ = Clock::now() - i0;
    integrate_total += integ;
    const Ms frame = Clock::now() - f0;
    std::cout << "Frame " << (i + 1) << "/" << o.frames << " rendered in " << ms(render) << ", integrated in "
              << ms(integ) << ", total=" << ms(frame) << ", active blocks=" << fusion.blockCount()
              << ", observed voxels=" << fusion.observedVoxelCount() << '\n';
  }

  const Clock::time_point m0 = Clock::now();
  const std::vector<float> tris = fusion.extractMesh();
  const Ms mesh = Clock::now() - m0;
  const fs::path out = o.output_path;
  const Clock::time_point w0 = Clock::now();
  if (!writeMesh(out, tris)) {
    std::cerr << "Failed to write mesh to " << out << '\n';
    return 1;
  }
  const Ms write = Clock::now() - w0;
  const Ms total = Clock::now() - start;
  std::cout << "Mesh written to " << out << '\n';
  std::cout << "Final stats: blocks=" << fusion.blockCount() << ", observed voxels=" << fusion.observedVoxelCount()
            << ", raw triangles=" << tris.size() / 9 << '\n';
  std::cout << "Timing summary: render=" << ms(render_total) << ", integrate=" << ms(integrate_total)
            << ", mesh_extract=" << ms(mesh) << ", mesh_write=" << ms(write) << ", total=" << ms(total) << '\n';
  return 0;
}

int runDataset(const Options& o) {  // runDatasetReconstruction (main.cpp:482-592)
  const Clock::time_point start = Clock::now();
  const fs::path root = o.dataset_root;
  const fs::path intrinsic_path = root / "intrinsic.txt";
  if (!fs::exists(root)) {
    std::cerr << "Dataset directory not found: " << root << '\n';
    return 1;
  }
  vhm_intrinsics K;
  float depth_scale = 0.0f;
  ioCheck(vhm_dataset_intrinsics(intrinsic_path.string().c_str(), &K, &depth_scale));
  size_t n_ids = 0;
  ioCheck(vhm_dataset_frame_ids(root.string().c_str(), nullptr, 0, &n_ids));
  std::vector<int32_t> ids(n_ids);
  ioCheck(vhm_dataset_frame_ids(root.string().c_str(), ids.data(), ids.size(), &n_ids));
  if (ids.empty()) {
    std::cerr << "No frame directories found under " << root << '\n';
    return 1;
  }
  const int last = ids.back();
  const int first = std::max(ids.front(), o.frame_start);
  const int end = o.frame_end >= 0 ? std::min(last, o.frame_end) : last;
  std::vector<int> chosen;
  for (const int id : ids)
    if (id >= first && id <= end && (id - first) % o.frame_step == 0) chosen.push_back(id);
  if (chosen.empty()) {
    std::cerr << "No frames selected with the current start/end/step options.\n";
    return 1;
  }

  GpuFusion fusion(o, K.width, K.height);
  std::cout << "Mode: dataset reconstruction\n";
  std::cout << "Dataset root: " << root << '\n';
  std::cout << "Intrinsics: fx=" << K.fx << ", fy=" << K.fy << ", cx=" << K.cx << ", cy=" << K.cy
            << ", width=" << K.width << ", height=" << K.height << ", depth_scale=" << depth_scale << '\n';
  std::cout << "Frame selection: start=" << first << ", end=" << end << ", step=" << o.frame_step
            << ", count=" << chosen.size() << '\n';
  std::cout << "Pose file: " << o.pose_file << '\n';
  std::cout << "Integration backend: " << fusion.description() << '\n';
  std::cout << "Depth input: <frame_id>/depth.png\n";

  const size_t pixels = static_cast<size_t>(K.width) * static_cast<size_t>(K.height);
  std::vector<uint16_t> raw(pixels);
  Ms load_total{0.0}, integrate_total{0.0};
  for (size_t i = 0; i < chosen.size(); ++i) {
    const Clock::time_point f0 = Clock::now();
    const int id = chosen[i];
    const Clock::time_point l0 = Clock::now();
    // loadDatasetDepthFrame (dataset_io.h:396-409): the pose, then the
    // depth; the raw samples go to the GPU, which applies depth_scale.
    const fs::path dir = root / std::to_string(id);
    vhm_pose T;
    size_t valid = 0;
    ioCheck(vhm_load_pose((dir / o.pose_file).string().c_str(), &T));
    ioCheck(vhm_load_depth_png_u16((dir / "depth.png").string().c_str(), &K, depth_scale, raw.data(), &valid));
    const Ms load = Clock::now() - l0;
    load_total += load;
    std::cout << "Loading frame " << id << " (" << (i + 1) << "/" << chosen.size() << "), valid_depth=" << valid
              << "/" << pixels << '\n';
    const Clock::time_point i0 = Clock::now();
    fusion.integrateRaw(raw.data(), depth_scale, K, T);
    const Ms integ = Clock::now() - i0;
    integrate_total += integ;
    const Ms frame = Clock::now() - f0;
    std::cout << "Integrated frame " << id << ", load=" << ms(load) << ", integrate=" << ms(integ)
              << ", total=" << ms(frame) << ", active blocks=" << fusion.blockCount()
              << ", observed voxels=" << fusion.observedVoxelCount() << '\n';
  }

  std::cout << "Extracting mesh with marching cubes...\n";
  const Clock::time_point m0 = Clock::now();
  const std::vector<float> tris = fusion.extractMesh();
  const Ms mesh = Clock::now() - m0;
  const fs::path out = o.output_path;
  const Clock::time_point w0 = Clock::now();
  if (!writeMesh(out, tris)) {
    std::cerr << "Failed to write mesh to " << out << '\n';
    return 1;
  }
  const Ms write = Clock::now() - w0;
  const Ms total = Clock::now() - start;
  std::cout << "Mesh written to " << out << '\n';
  std::cout << "Final stats: blocks=" << fusion.blockCount() << ", observed voxels=" << fusion.observedVoxelCount()
            << ", raw triangles=" << tris.size() / 9 << '\n';
  std::cout << "Timing summary: frame_load=" << ms(load_total) << ", integrate=" << ms(integrate_total)
            << ", mesh_extract=" << ms(mesh) << ", mesh_write=" << ms(write) << ", total=" << ms(total) << '\n';
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Options o = parseArgs(argc, argv);
    return o.synthetic ? runSynthetic(o) : runDataset(o);
  } catch (const std::exception& e) {
    std::cerr << e.what() << '\n';
    return 1;
  }
}
