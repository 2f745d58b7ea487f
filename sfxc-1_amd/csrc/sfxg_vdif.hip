// Raw VDIF ingest on the device: the frames stay in HBM as they arrive.
// k_vdif_words lifts the first four header words of every frame into a small
// table for the host's frame walk (the payload never leaves the device), and
// k_vdif_gather assembles the DATA payload of one channel's HEADER stream
// from the frames, piece by piece, in the order of the DATA records.
#include "sfxg_internal.h"

namespace sfxg {
namespace {

__global__ __launch_bounds__(256) void k_vdif_words(const uint8_t *__restrict__ frames, int64_t n_frames,
                                                    int frame_bytes, uint32_t *__restrict__ words) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_frames * 4) return;
  const int64_t f = i >> 2;
  const int w = (int)(i & 3);
  words[i] = reinterpret_cast<const uint32_t *>(frames + f * frame_bytes)[w];
}

// One workgroup per piece: bytes [src, src + n) of the frames to dst.
// Frame payloads start 8-byte aligned; pieces may start anywhere, so the
// copy goes byte by byte, four bytes per thread and pass.
__global__ __launch_bounds__(256) void k_vdif_gather(const uint8_t *__restrict__ frames,
                                                     const VdifPiece *__restrict__ pieces,
                                                     const int64_t *__restrict__ dst_off, uint8_t *__restrict__ dst) {
  const VdifPiece p = pieces[blockIdx.x];
  const uint8_t *s = frames + p.src;
  uint8_t *d = dst + dst_off[blockIdx.x];
  for (int64_t k = (int64_t)threadIdx.x * 4; k < p.n; k += 4 * 256) {
#pragma unroll
    for (int e = 0; e < 4; e++)
      if (k + e < p.n) d[k + e] = s[k + e];
  }
}

}  // namespace

hipError_t launch_vdif_words(const void *frames, int64_t n_frames, int frame_bytes, uint32_t *words,
                             hipStream_t s) {
  if (n_frames <= 0) return hipSuccess;
  const int64_t n = n_frames * 4;
  hipLaunchKernelGGL(k_vdif_words, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     static_cast<const uint8_t *>(frames), n_frames, frame_bytes, words);
  return hipGetLastError();
}

hipError_t launch_vdif_gather(const void *frames, const VdifPiece *pieces, const int64_t *dst_off, int n_pieces,
                              void *dst, hipStream_t s) {
  if (n_pieces <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_vdif_gather, dim3((unsigned)n_pieces), dim3(256), 0, s, static_cast<const uint8_t *>(frames),
                     pieces, dst_off, static_cast<uint8_t *>(dst));
  return hipGetLastError();
}

}  // namespace sfxg
