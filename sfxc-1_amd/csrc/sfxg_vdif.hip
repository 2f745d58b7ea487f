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

// Phase-cal accumulation (Input_node_data_writer::do_phasecal,
// input_node_data_writer.cc:232-306): sample j of a frame adds its value to
// bin (count0 + j) mod P of the frame's integration. One thread per bin
// walks a group of frames; its sum per integration is added to the bins
// with an integer atomic, so the result does not depend on the order.
__global__ __launch_bounds__(256) void k_phasecal(const uint8_t *__restrict__ frames,
                                                  const PcalFrame *__restrict__ table, int n_frames,
                                                  int frames_per_group, int samples_per_frame, int P, int bits,
                                                  int32_t *__restrict__ sums) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P) return;
  const int f0 = blockIdx.y * frames_per_group;
  const int f1 = min(n_frames, f0 + frames_per_group);
  const int spb = 8 / bits;
  int32_t acc = 0;
  int cur = -1;
  for (int f = f0; f < f1; f++) {
    const PcalFrame t = table[f];
    if (t.integration != cur) {
      if (cur >= 0 && acc != 0) atomicAdd(&sums[(int64_t)cur * P + k], acc);
      acc = 0;
      cur = t.integration;
    }
    const uint8_t *pay = frames + t.payload_offset;
    int j = k - t.count0;
    if (j < 0) j += P;
    for (; j < samples_per_frame; j += P) {
      const int byte = pay[j / spb];
      if (bits == 2) {
        const int code = (byte >> (2 * (j & 3))) & 3;
        acc += code == 0 ? -7 : code == 1 ? -2 : code == 2 ? 2 : 7;  // sample_value_2
      } else {
        acc += ((byte >> (j & 7)) & 1) ? 5 : -5;  // sample_value_1
      }
    }
  }
  if (cur >= 0 && acc != 0) atomicAdd(&sums[(int64_t)cur * P + k], acc);
}

}  // namespace

hipError_t launch_phasecal(const void *frames, const PcalFrame *table, int n_frames, int samples_per_frame, int P,
                           int bits, int32_t *sums, hipStream_t s) {
  if (n_frames <= 0) return hipSuccess;
  const int fpg = 16;
  dim3 grid((unsigned)((P + 255) / 256), (unsigned)((n_frames + fpg - 1) / fpg));
  hipLaunchKernelGGL(k_phasecal, grid, dim3(256), 0, s, static_cast<const uint8_t *>(frames), table, n_frames, fpg,
                     samples_per_frame, P, bits, sums);
  return hipGetLastError();
}

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
