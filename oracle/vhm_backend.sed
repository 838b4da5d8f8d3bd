# The change INTEGRATION.md section 2 describes, as sed rules over the
# reference's src/reconstruction_backend.cpp: its CUDA dispatch branches
# call the libvhm host mirror (voxelhashmt_amd/host/tsdfmc_vhm_backend.hpp)
# instead, under TSDFMC_HAVE_VHM.  Applied by oracle/Makefile to a copy in a
# temporary directory; the patched file is never kept.
s|#if TSDFMC_HAVE_CUDA|#if TSDFMC_HAVE_VHM|
s|#include "cuda_mesh_extractor.h"|#include "tsdfmc_vhm_backend.hpp"|
\|#include "cuda_tsdf_integrator.h"|d
s|const CudaBackendStatus status = queryCudaBackendStatus();|const auto status = vhm_backend::queryGpuBackendStatus();|
s|but CUDA is unavailable|but HIP is unavailable|
s|"GPU (CUDA)"|"GPU (HIP, gfx950)"|
s|"GPU (CUDA, auto-selected)"|"GPU (HIP, auto-selected)"|
s|CUDA backend built but unavailable|HIP backend built but unavailable|
s|integrateDepthFrameCuda(volume, frame, T_wc, allocation_stride);|vhm_backend::integrateDepthFrameGpu(volume, frame, T_wc, allocation_stride);|
s|return extractMeshCuda(volume, iso_level);|return vhm_backend::extractMeshGpu<Triangle>(volume, iso_level);|
s|return gpuVolumeCacheBlockCount(volume);|return vhm_backend::gpuVolumeCacheBlockCount(volume);|
s|return gpuVolumeCacheObservedVoxelCount(volume);|return vhm_backend::gpuVolumeCacheObservedVoxelCount(volume);|
