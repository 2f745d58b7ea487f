// Phased-array core (Correlation_core_phased, correlation_core_phased.cc:13-126):
// the spectra of all streams summed per sub-integration,
// row[s][ch] = sum over segments of sub-integration s, in segment order, of
// sum over streams in stream order of X_stream[segment][ch].
//
// The work streams the spectra once (8 bytes per station-bin). A row can
// hold every segment of the slice, so the sum runs in two passes for
// parallelism: blocks of up to kPhasedBlock consecutive segments of one row
// are summed in the reference's order (segments, then streams) into
// partials, and one pass per row adds its partials to the row in segment
// order. Rows carry over chunks of the slice, so each chunk adds to them.
#include <algorithm>

#include "sfxg_internal.h"

namespace sfxg {
namespace {

constexpr int PH_THREADS = 256;

__global__ void __launch_bounds__(PH_THREADS)
    k_phased_partial(const float2 *__restrict__ spectra, int64_t stream_stride, int64_t seg_stride,
                     int64_t seg_base, int n_ch, int M, const PhasedBlock *__restrict__ blocks,
                     float2 *__restrict__ partial) {
  const int ch = blockIdx.y * PH_THREADS + threadIdx.x;
  if (ch >= n_ch) return;
  const PhasedBlock b = blocks[blockIdx.x];
  float2 acc = make_float2(0.0f, 0.0f);
  for (int64_t seg = b.seg0; seg < b.seg1; seg++) {
    const float2 *p = spectra + (seg - seg_base) * seg_stride + ch;
    for (int i = 0; i < M; i++) {
      const float2 v = p[(int64_t)i * stream_stride];
      acc.x += v.x;
      acc.y += v.y;
    }
  }
  partial[(int64_t)blockIdx.x * n_ch + ch] = acc;
}

__global__ void __launch_bounds__(PH_THREADS)
    k_phased_rows(const float2 *__restrict__ partial, int n_ch, const PhasedRow *__restrict__ rows,
                  float2 *__restrict__ out) {
  const int ch = blockIdx.y * PH_THREADS + threadIdx.x;
  if (ch >= n_ch) return;
  const PhasedRow r = rows[blockIdx.x];
  float2 acc = out[(int64_t)r.row * n_ch + ch];
  for (int b = r.b0; b < r.b1; b++) {
    const float2 v = partial[(int64_t)b * n_ch + ch];
    acc.x += v.x;
    acc.y += v.y;
  }
  out[(int64_t)r.row * n_ch + ch] = acc;
}

}  // namespace

void phased_tables(int64_t seg0, int64_t seg1, int n_sub, std::vector<PhasedBlock> &blocks,
                   std::vector<PhasedRow> &rows) {
  const int b_first = (int)blocks.size();
  for (int64_t s = seg0; s < seg1;) {
    const int row = (int)(s / n_sub);
    const int64_t row_end = std::min<int64_t>(seg1, (int64_t)(row + 1) * n_sub);
    const int b0 = (int)blocks.size() - b_first;
    for (int64_t a = s; a < row_end; a += kPhasedBlock)
      blocks.push_back({a, std::min<int64_t>(row_end, a + kPhasedBlock)});
    rows.push_back({row, b0, (int)blocks.size() - b_first, 0});
    s = row_end;
  }
}

hipError_t launch_phased(const float2 *spectra, int64_t stream_stride, int64_t seg_stride, int64_t seg_base,
                         int n_ch, int M, const PhasedBlock *blocks, int n_blocks, const PhasedRow *rows, int n_rows,
                         float2 *partial, float2 *out, hipStream_t s) {
  if (n_blocks <= 0 || n_rows <= 0) return hipSuccess;
  const unsigned gy = (unsigned)((n_ch + PH_THREADS - 1) / PH_THREADS);
  hipLaunchKernelGGL(k_phased_partial, dim3(n_blocks, gy), dim3(PH_THREADS), 0, s, spectra, stream_stride,
                     seg_stride, seg_base, n_ch, M, blocks, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_phased_rows, dim3(n_rows, gy), dim3(PH_THREADS), 0, s, partial, n_ch, rows, out);
  return hipGetLastError();
}

}  // namespace sfxg
