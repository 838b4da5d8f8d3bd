// C-ABI of the learned block denoiser (include/vhm.h, vhm_denoiser_*):
// device training set, forward, training step (forward, backward, Adam),
// write-back with confidence fallback.  Kernels: vhm_denoise.hip.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vhm.h"
#include "vhm_internal.h"

using namespace vhm;


struct vhm_denoiser {
  int capacity = 0;  // blocks in the training/inference set
  int batch_capacity = 0;  // blocks per training step
  int count = 0;
  hipStream_t stream = nullptr;
  int groups = 256;  // backward workgroups (partial gradient rows): one per CU
  int step = 0;
  // parameters (fp32 master + Adam moments + gradient) and packed bf16 weights
  float* P = nullptr;
  float* G = nullptr;
  float* Mm = nullptr;
  float* Vv = nullptr;
  uint16_t* W1t = nullptr;
  uint16_t* W2t = nullptr;
  uint16_t* W3t = nullptr;
  uint16_t* W2b = nullptr;
  uint16_t* W3b = nullptr;
  float* partial = nullptr;
  // sample set
  uint16_t* X = nullptr;  // [cap][1000][16]
  float* T = nullptr;     // [cap][512][6]
  float* M = nullptr;
  int32_t* keys = nullptr;  // [cap][3]
  std::vector<uint32_t> mcount;  // per sample: number of target cells with M = 1 (host)
  // per-batch scratch
  float* Y = nullptr;
  float* dY = nullptr;
  uint16_t* H1 = nullptr;
  uint16_t* H2 = nullptr;
  uint16_t* dA1 = nullptr;
  uint16_t* dA2 = nullptr;
  float* loss = nullptr;
  unsigned long long* counters = nullptr;
  std::string err;
};

namespace {

constexpr size_t kXElems = 1000 * 16, kYElems = 512 * 6, kHElems = 512 * 32;

int dfail(vhm_denoiser* d, int code, const std::string& m) {
  if (d) d->err = m;
  return code;
}

#define DN_TRY(d, expr, what)                                                                        \
  do {                                                                                               \
    const hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess) return dfail(d, VHM_ERR_HIP, std::string(what) + " failed: " + hipGetErrorString(_e)); \
  } while (0)

DnWeights weights(const vhm_denoiser* d) { return {d->W1t, d->W2t, d->W3t, d->W2b, d->W3b, d->P}; }

int repack(vhm_denoiser* d) {
  DN_TRY(d, dnLaunchPack(d->P, d->W1t, d->W2t, d->W3t, d->W2b, d->W3b, d->stream), "weight pack launch");
  return VHM_OK;
}

}  // namespace

extern "C" {

int vhm_denoiser_num_params(void) { return dnNumParams(); }

int vhm_denoiser_create(int capacity, int batch_capacity, vhm_denoiser** out) {
  if (!out || capacity <= 0 || batch_capacity <= 0) return VHM_ERR_INVALID_ARGUMENT;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return VHM_ERR_NO_DEVICE;
  auto* d = new vhm_denoiser();
  d->capacity = capacity;
  d->batch_capacity = batch_capacity;
  const size_t np = dnNumParams();
  const size_t cap = capacity, bc = batch_capacity;
  hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  auto A = [&](void** p, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(p, bytes);
  };
  A(reinterpret_cast<void**>(&d->P), np * 4);
  A(reinterpret_cast<void**>(&d->G), np * 4);
  A(reinterpret_cast<void**>(&d->Mm), np * 4);
  A(reinterpret_cast<void**>(&d->Vv), np * 4);
  A(reinterpret_cast<void**>(&d->W1t), 27 * 32 * 16 * 2);
  A(reinterpret_cast<void**>(&d->W2t), 27 * 32 * 32 * 2);
  A(reinterpret_cast<void**>(&d->W3t), 27 * 16 * 32 * 2);
  A(reinterpret_cast<void**>(&d->W2b), 27 * 32 * 32 * 2);
  A(reinterpret_cast<void**>(&d->W3b), 27 * 32 * 8 * 2);
  A(reinterpret_cast<void**>(&d->partial), size_t(d->groups) * np * 4);
  A(reinterpret_cast<void**>(&d->X), cap * kXElems * 2);
  A(reinterpret_cast<void**>(&d->T), cap * kYElems * 4);
  A(reinterpret_cast<void**>(&d->M), cap * kYElems * 4);
  A(reinterpret_cast<void**>(&d->keys), cap * 3 * 4);
  A(reinterpret_cast<void**>(&d->Y), cap * kYElems * 4);
  A(reinterpret_cast<void**>(&d->dY), bc * kYElems * 4);
  A(reinterpret_cast<void**>(&d->H1), bc * kHElems * 2);
  A(reinterpret_cast<void**>(&d->H2), bc * kHElems * 2);
  A(reinterpret_cast<void**>(&d->dA1), bc * kHElems * 2);
  A(reinterpret_cast<void**>(&d->dA2), bc * kHElems * 2);
  A(reinterpret_cast<void**>(&d->loss), bc * 4);
  A(reinterpret_cast<void**>(&d->counters), 16);
  if (e == hipSuccess) e = hipMemset(d->P, 0, np * 4);
  if (e == hipSuccess) e = hipMemset(d->Mm, 0, np * 4);
  if (e == hipSuccess) e = hipMemset(d->Vv, 0, np * 4);
  if (e != hipSuccess) {
    vhm_denoiser_destroy(d);
    return VHM_ERR_HIP;
  }
  d->err.clear();
  *out = d;
  return VHM_OK;
}

int vhm_denoiser_destroy(vhm_denoiser* d) {
  if (!d) return VHM_OK;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  void* bufs[] = {d->P,  d->G,  d->Mm, d->Vv, d->W1t, d->W2t, d->W3t, d->W2b, d->W3b, d->partial, d->X, d->T,
                  d->M,  d->keys, d->Y, d->dY, d->H1, d->H2, d->dA1, d->dA2, d->loss, d->counters};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return VHM_OK;
}

const char* vhm_denoiser_last_error(const vhm_denoiser* d) { return d ? d->err.c_str() : "null denoiser"; }

int vhm_denoiser_set_params(vhm_denoiser* d, const float* params) {
  if (!d || !params) return VHM_ERR_INVALID_ARGUMENT;
  const size_t np = dnNumParams();
  DN_TRY(d, hipMemcpyAsync(d->P, params, np * 4, hipMemcpyHostToDevice, d->stream), "hipMemcpy params");
  DN_TRY(d, hipMemsetAsync(d->Mm, 0, np * 4, d->stream), "hipMemset adam m");
  DN_TRY(d, hipMemsetAsync(d->Vv, 0, np * 4, d->stream), "hipMemset adam v");
  d->step = 0;
  int rc = repack(d);
  if (rc) return rc;
  DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  return VHM_OK;
}

int vhm_denoiser_get_params(vhm_denoiser* d, float* params) {
  if (!d || !params) return VHM_ERR_INVALID_ARGUMENT;
  DN_TRY(d, hipMemcpyAsync(params, d->P, dnNumParams() * 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpy params");
  DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  return VHM_OK;
}

int vhm_denoiser_clear(vhm_denoiser* d) {
  if (!d) return VHM_ERR_INVALID_ARGUMENT;
  d->count = 0;
  return VHM_OK;
}

int vhm_denoiser_count(const vhm_denoiser* d) { return d ? d->count : 0; }

// Appends samples for blocks `keys`: inputs from `input_vol` and, when
// target_vol is not null, targets from it.  Both volumes must have been
// created with the same voxel size and truncation.
int vhm_denoiser_append(vhm_denoiser* d, vhm_ctx* input_vol, vhm_ctx* target_vol, const int32_t* keys, size_t n) {
  if (!d || !input_vol || (!keys && n)) return VHM_ERR_INVALID_ARGUMENT;
  if (d->count + n > static_cast<size_t>(d->capacity))
    return dfail(d, VHM_ERR_CAPACITY, "denoiser append failed: sample set full");
  if (n == 0) return VHM_OK;
  const vhm_params* P = vhmCtxParams(input_vol);
  // the volumes' queued work must finish before gathering from them
  DN_TRY(d, hipStreamSynchronize(vhmCtxStream(input_vol)), "hipStreamSynchronize input volume");
  if (target_vol) DN_TRY(d, hipStreamSynchronize(vhmCtxStream(target_vol)), "hipStreamSynchronize target volume");
  int32_t* k = d->keys + 3 * size_t(d->count);
  DN_TRY(d, hipMemcpyAsync(k, keys, 3 * n * 4, hipMemcpyHostToDevice, d->stream), "hipMemcpy keys");
  DN_TRY(d, dnLaunchGather(vhmCtxView(input_vol), k, static_cast<uint32_t>(n), P->voxel_size, P->truncation_distance,
                           P->max_weight, d->X + size_t(d->count) * kXElems, d->stream),
         "denoiser gather launch");
  if (d->mcount.size() < static_cast<size_t>(d->capacity)) d->mcount.resize(d->capacity);
  if (target_vol) {
    const vhm_params* Q = vhmCtxParams(target_vol);
    DN_TRY(d, dnLaunchTarget(vhmCtxView(target_vol), k, static_cast<uint32_t>(n), Q->voxel_size,
                             Q->truncation_distance, Q->max_weight, d->T + size_t(d->count) * kYElems,
                             d->M + size_t(d->count) * kYElems, d->stream),
           "denoiser target launch");
    // per-sample mask counts, once, for the loss normalisation of training steps
    std::vector<float> mh(n * kYElems);
    DN_TRY(d, hipMemcpyAsync(mh.data(), d->M + size_t(d->count) * kYElems, mh.size() * 4, hipMemcpyDeviceToHost,
                             d->stream),
           "hipMemcpy masks");
    DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < n; ++i) {
      uint32_t c = 0;
      for (size_t j = 0; j < kYElems; ++j) c += mh[i * kYElems + j] != 0.0f ? 1u : 0u;
      d->mcount[d->count + i] = c;
    }
  } else {
    for (size_t i = 0; i < n; ++i) d->mcount[d->count + i] = 0;
  }
  DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  d->count += static_cast<int>(n);
  return VHM_OK;
}

// Inference over samples [offset, offset + n): Y (device, kept for apply);
// y_out (host, n x 512 x 6) optional.
int vhm_denoiser_forward(vhm_denoiser* d, size_t offset, size_t n, float* y_out) {
  if (!d || offset + n > static_cast<size_t>(d->count)) return VHM_ERR_INVALID_ARGUMENT;
  DN_TRY(d, dnLaunchForward(d->X + offset * kXElems, weights(d), static_cast<uint32_t>(n), d->Y + offset * kYElems,
                            nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0.0f, d->stream),
         "denoiser forward launch");
  if (y_out)
    DN_TRY(d, hipMemcpyAsync(y_out, d->Y + offset * kYElems, n * kYElems * 4, hipMemcpyDeviceToHost, d->stream),
           "hipMemcpy predictions");
  DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  return VHM_OK;
}

// One training step on samples [offset, offset + n) (n <= batch capacity):
// forward, masked-MSE loss, backward, gradient reduction, Adam.
// lr <= 0 computes loss and gradient without updating the parameters.
int vhm_denoiser_train_step(vhm_denoiser* d, size_t offset, size_t n, float lr, float* loss_out, float* grad_out) {
  if (!d || n == 0 || offset + n > static_cast<size_t>(d->count)) return VHM_ERR_INVALID_ARGUMENT;
  if (n > static_cast<size_t>(d->batch_capacity))
    return dfail(d, VHM_ERR_INVALID_ARGUMENT, "denoiser train step failed: batch larger than batch_capacity");
  // mask count (counted per sample at append) -> 1 / max(sum M, 1)
  double msum = 0.0;
  for (size_t i = 0; i < n; ++i) msum += static_cast<double>(d->mcount[offset + i]);
  const float inv = 1.0f / static_cast<float>(msum > 1.0 ? msum : 1.0);
  const uint16_t* X = d->X + offset * kXElems;
  // the step's predictions are not kept (apply uses vhm_denoiser_forward's)
  DN_TRY(d, dnLaunchForward(X, weights(d), static_cast<uint32_t>(n), nullptr, d->H1, d->H2,
                            d->T + offset * kYElems, d->M + offset * kYElems, d->dY, d->loss, inv, d->stream),
         "denoiser training forward launch");
  // backward, gradient reduction and (lr > 0) Adam's step with the packed
  // weights rewritten, all on the stream
  const bool update = lr > 0.0f;
  if (update) ++d->step;
  const DnUpdate up{d->P, d->Mm, d->Vv, d->W1t, d->W2t, d->W3t, d->W2b, d->W3b, lr, 0.9f, 0.999f, 1e-8f, 0.0f, 0.0f};
  DN_TRY(d, dnLaunchBackward(X, d->H1, d->H2, d->dY, d->dA2, d->dA1, weights(d), static_cast<uint32_t>(n), d->groups,
                             d->partial, d->G, update ? &up : nullptr, d->step, d->stream),
         "denoiser backward launch");
  std::vector<float> lp(n);
  DN_TRY(d, hipMemcpyAsync(lp.data(), d->loss, n * 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpy loss");
  if (grad_out)
    DN_TRY(d, hipMemcpyAsync(grad_out, d->G, dnNumParams() * 4, hipMemcpyDeviceToHost, d->stream), "hipMemcpy grad");
  DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  double l = 0.0;
  for (float v : lp) l += v;
  if (loss_out) *loss_out = static_cast<float>(l * inv);
  return VHM_OK;
}

// Writes the predictions of samples [offset, offset + n) (from the last
// vhm_denoiser_forward over them) back into `vol`, with the confidence fallback described in
// vhm_denoise.hip (dnApplyKernel).
int vhm_denoiser_apply(vhm_denoiser* d, vhm_ctx* vol, size_t offset, size_t n, float min_weight, float max_delta,
                       float block_max_delta, uint64_t* cells_replaced, uint64_t* blocks_kept_plain) {
  if (!d || !vol || offset + n > static_cast<size_t>(d->count)) return VHM_ERR_INVALID_ARGUMENT;
  const vhm_params* P = vhmCtxParams(vol);
  DN_TRY(d, hipStreamSynchronize(vhmCtxStream(vol)), "hipStreamSynchronize volume");
  DN_TRY(d, hipMemsetAsync(d->counters, 0, 16, d->stream), "hipMemset counters");
  DN_TRY(d, dnLaunchApply(vhmCtxView(vol), d->keys + 3 * offset, static_cast<uint32_t>(n), d->Y + offset * kYElems,
                          d->X + offset * kXElems, P->voxel_size, P->truncation_distance, P->max_weight, min_weight,
                          max_delta, block_max_delta, d->counters, d->stream),
         "denoiser apply launch");
  unsigned long long c[2] = {0, 0};
  DN_TRY(d, hipMemcpyAsync(c, d->counters, 16, hipMemcpyDeviceToHost, d->stream), "hipMemcpy counters");
  DN_TRY(d, hipStreamSynchronize(d->stream), "hipStreamSynchronize");
  if (cells_replaced) *cells_replaced = c[0];
  if (blocks_kept_plain) *blocks_kept_plain = c[1];
  return VHM_OK;
}

// Copies sample inputs (bf16 bits, n x 1000 x 16) and targets/masks (n x 512 x 6).
int vhm_denoiser_samples(vhm_denoiser* d, size_t offset, size_t n, uint16_t* x_out, float* t_out, float* m_out) {
  if (!d || offset + n > static_cast<size_t>(d->count)) return VHM_ERR_INVALID_ARGUMENT;
  if (x_out) DN_TRY(d, hipMemcpy(x_out, d->X + offset * kXElems, n * kXElems * 2, hipMemcpyDeviceToHost), "hipMemcpy X");
  if (t_out) DN_TRY(d, hipMemcpy(t_out, d->T + offset * kYElems, n * kYElems * 4, hipMemcpyDeviceToHost), "hipMemcpy T");
  if (m_out) DN_TRY(d, hipMemcpy(m_out, d->M + offset * kYElems, n * kYElems * 4, hipMemcpyDeviceToHost), "hipMemcpy M");
  return VHM_OK;
}

int vhm_mfma_layout_probe(float* out256) {
  if (!out256) return VHM_ERR_INVALID_ARGUMENT;
  float* dev = nullptr;
  if (hipMalloc(&dev, 256 * 4) != hipSuccess) return VHM_ERR_HIP;
  hipError_t e = dnLaunchProbe(dev, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out256, dev, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  return e == hipSuccess ? VHM_OK : VHM_ERR_HIP;
}

int vhm_mfma_layout_probe_k32(const float* a16x32, const float* b32x16, float* out256) {
  if (!a16x32 || !b32x16 || !out256) return VHM_ERR_INVALID_ARGUMENT;
  float* dev = nullptr;
  if (hipMalloc(&dev, (512 + 512 + 256) * 4) != hipSuccess) return VHM_ERR_HIP;
  hipError_t e = hipMemcpy(dev, a16x32, 512 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dev + 512, b32x16, 512 * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = dnLaunchProbe32(dev, dev + 512, dev + 1024, nullptr);
  if (e == hipSuccess) e = hipMemcpy(out256, dev + 1024, 256 * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dev);
  return e == hipSuccess ? VHM_OK : VHM_ERR_HIP;
}

}  // extern "C"
