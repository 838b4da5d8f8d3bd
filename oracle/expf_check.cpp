// Checker for the product's exp (voxelhashmt_amd/csrc/vhm_math.h,
// expf_glibc) compiled for the host -- test infrastructure only.  The same
// header is what the HIP kernels compile; here it is compared with the
// host's glibc expf (what the reference's std::exp(float) calls).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../voxelhashmt_amd/csrc/vhm_math.h"

extern "C" {

void vcheck_expf_product(const float* x, float* y, std::size_t n) {
  for (std::size_t i = 0; i < n; ++i) y[i] = vhm::expf_glibc(x[i]);
}

void vcheck_expf_libm(const float* x, float* y, std::size_t n) {
  for (std::size_t i = 0; i < n; ++i) y[i] = std::exp(x[i]);
}

// Counts inputs with bit patterns in [lo, hi] (step `stride`) where the
// product's exp and glibc expf differ; first mismatch pattern in *first.
uint64_t vcheck_expf_sweep(uint32_t lo, uint32_t hi, uint32_t stride, uint32_t* first) {
  uint64_t bad = 0;
  *first = 0;
  for (uint64_t b = lo; b <= hi; b += stride) {
    float x;
    const uint32_t bits = static_cast<uint32_t>(b);
    std::memcpy(&x, &bits, 4);
    const float a = vhm::expf_glibc(x), g = std::exp(x);
    if (std::memcmp(&a, &g, 4) != 0) {
      if (bad == 0) *first = bits;
      ++bad;
    }
  }
  return bad;
}

void vcheck_exp_table(uint64_t* out32) {
  for (int i = 0; i < 32; ++i) out32[i] = vhm::kExp2Tab32[i];
}

// The product's octahedral dequantisation for every 16-bit code, beside the
// plain correctly rounded division it replaces.
void vcheck_dequant_all(float* product, float* division) {
  for (uint32_t c = 0; c < 65536u; ++c) {
    product[c] = vhm::dequantOct(c);
    division[c] = (static_cast<float>(c) / 65535.0f) * 2.0f - 1.0f;
  }
}

}  // extern "C"
