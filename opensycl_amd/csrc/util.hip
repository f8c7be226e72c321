// Small device utilities shared by the launchers.
#include "internal.hpp"

namespace acpp_impl {

namespace {
// Zeroes n16 16-byte chunks and ntail trailing bytes; the chunk at index 1
// receives `chunk1` instead (the trip flag pointer of a sync region header,
// see sync_header) when `set_chunk1` is non-zero.
__global__ void zero_kernel(uint4 *p, uint64_t n16, unsigned char *tail, uint32_t ntail, uint4 chunk1,
                            int set_chunk1) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
    p[i] = (i == 1 && set_chunk1) ? chunk1 : make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

acpp_status launch_zero(void *p, size_t bytes, hipStream_t stream, uint4 chunk1, int set_chunk1) {
  if (bytes == 0) return ACPP_SUCCESS;
  if ((reinterpret_cast<uintptr_t>(p) & 15) != 0)
    return set_error(ACPP_ERROR_INVALID_ARGUMENT, "fill_zero: pointer not 16-byte aligned");
  const uint64_t n16 = bytes / 16;
  const uint32_t ntail = static_cast<uint32_t>(bytes % 16);
  uint64_t blocks = (n16 + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream,
                     static_cast<uint4 *>(p), n16, static_cast<unsigned char *>(p) + n16 * 16, ntail, chunk1,
                     set_chunk1);
  return check_launch("fill_zero");
}
}  // namespace

// Scratch regions are 256-byte aligned, so the bulk is cleared with 16-byte
// stores; a kernel (rather than a memset) also behaves inside graph capture.
acpp_status fill_zero(void *p, size_t bytes, hipStream_t stream) {
  return launch_zero(p, bytes, stream, make_uint4(0, 0, 0, 0), 0);
}

acpp_status reset_sync_region_flag(void *region, size_t bytes, uint32_t *trip_flag, hipStream_t stream) {
  if (bytes < sizeof(sync_header)) bytes = sizeof(sync_header);
  const uint64_t flag = reinterpret_cast<uintptr_t>(trip_flag);
  const uint4 chunk1 = make_uint4(static_cast<uint32_t>(flag), static_cast<uint32_t>(flag >> 32), 0, 0);
  return launch_zero(region, bytes, stream, chunk1, 1);
}

acpp_status reset_sync_region(void *region, size_t bytes, int device, hipStream_t stream) {
  return reset_sync_region_flag(region, bytes, host_trip_flag(device), stream);
}

bool stream_is_capturing(hipStream_t stream) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &st) != hipSuccess) return false;
  return st == hipStreamCaptureStatusActive;
}

}  // namespace acpp_impl
