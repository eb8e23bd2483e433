// probe_fft_padding.hip -- the N = 4096 transform of the row kernels
// (fft_lds<float, 4096, 512>) as a standalone program, not part of the library:
// 1024 rows, each taken through 8 forward + inverse round trips in LDS, timed
// with HIP events and checked (round-trip error, a checksum that must not
// depend on the padding).  Built once per LDS pad period 2^FIL_LDS_PAD_SHIFT
// (make probe_fft_padding builds shifts 3..7); under rocprofv3 --pmc
// SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE SQ_INSTS_LDS it gives the conflicted
// share of the LDS cycles.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#include <cmath>
#include "fil_fft.h"

#define FIL_STR2(x) #x
#define FIL_STR(x) FIL_STR2(x)
using namespace fil;
#define CK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e_), __LINE__); return 1; } } while (0)
constexpr int N = 4096, NTH = 512;
__global__ __launch_bounds__(NTH) void fft_rows(cx<float>* data, const cx<float>* tw, int reps) {
  __shared__ cx<float> Z[lds_size(N)];
  const int tid = threadIdx.x;
  cx<float>* row = data + (size_t)blockIdx.x * N;
  for (int t = tid; t < N; t += NTH) Z[lds_index(t)] = row[t];
  __syncthreads();
  for (int r = 0; r < reps; r++) {
    fft_lds<float, N, NTH, -1>(Z, tw);
    fft_lds<float, N, NTH, +1>(Z, tw);
    for (int t = tid; t < N; t += NTH) Z[lds_index(t)] = (1.0f / N) * Z[lds_index(t)];
    __syncthreads();
  }
  for (int t = tid; t < N; t += NTH) row[t] = Z[lds_index(t)];
}
int main() {
  const int rows = 1024, reps = 8;
  // twiddles as the library builds them: per pass with twiddles, P x FIL_TW_STRIDE entries exp(-2 pi i p S k / N)
  std::vector<cx<float>> tw(fft_tw_size(N));
  for (int S = 1; S < N; S *= fft_radix(N, S)) {
    const int rad = fft_radix(N, S);
    if (S >= N / rad) continue;
    const int off = fft_tw_offset(N, S), P = N / (S * rad);
    for (int p = 0; p < P; p++) for (int k = 0; k < FIL_TW_STRIDE; k++) {
      const double a = -2.0 * M_PI * p * S * k / N;
      tw[off + p * FIL_TW_STRIDE + k] = mk((float)cos(a), (float)sin(a));
    }
  }
  std::vector<cx<float>> h((size_t)rows * N);
  for (size_t i = 0; i < h.size(); i++) h[i] = mk((float)sin(0.001 * i), (float)cos(0.0007 * i));
  cx<float> *d, *dtw;
  CK(hipMalloc(&d, h.size() * 8)); CK(hipMalloc(&dtw, tw.size() * 8));
  CK(hipMemcpy(dtw, tw.data(), tw.size() * 8, hipMemcpyHostToDevice));
  CK(hipMemcpy(d, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(fft_rows, dim3(rows), dim3(NTH), 0, 0, d, dtw, 1);
  CK(hipDeviceSynchronize());
  hipEvent_t e0, e1; CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  CK(hipEventRecord(e0));
  for (int it = 0; it < 20; it++) hipLaunchKernelGGL(fft_rows, dim3(rows), dim3(NTH), 0, 0, d, dtw, reps);
  CK(hipEventRecord(e1)); CK(hipEventSynchronize(e1));
  float ms; CK(hipEventElapsedTime(&ms, e0, e1));
  std::vector<cx<float>> o(h.size()); CK(hipMemcpy(o.data(), d, o.size() * 8, hipMemcpyDeviceToHost));
  double err = 0, nrm = 0;
  for (size_t i = 0; i < o.size(); i++) { double dx = o[i].x - h[i].x, dy = o[i].y - h[i].y; err += dx*dx + dy*dy; nrm += h[i].x*h[i].x + h[i].y*h[i].y; }
  double cs = 0; for (size_t i = 0; i < o.size(); i += 7) cs += o[i].x * 1.3 + o[i].y;
  printf("%s: %.2f us per launch (%d rows x %d round trips), %.3f us per 4096-point transform over all rows, roundtrip rel err %.2e, checksum %.9e\n",
         "pad 2^" FIL_STR(FIL_LDS_PAD_SHIFT), ms * 1e3 / 20, rows, reps, ms * 1e3 / 20 / (2 * reps), sqrt(err / nrm), cs);
  return 0;
}
