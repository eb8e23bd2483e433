// probe_lds_banks.hip -- measures the LDS cost of the lane -> address
// patterns the Hankel GEMMs use, as a standalone program (not part of the
// library): one wave per workgroup issues ds_read_b128 or ds_write_b128 with
// one pattern and row stride per kernel instantiation, and rocprofv3 reads
// SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE / SQ_INSTS_LDS per dispatch:
//   make -C filament.jl_amd/csrc probe_lds_banks
//   rocprofv3 --pmc SQ_LDS_BANK_CONFLICT SQ_LDS_IDX_ACTIVE SQ_INSTS_LDS SQ_WAVES \
//             --output-format csv -d <dir> -o run -- filament.jl_amd/csrc/probe_lds_banks
// Patterns (16 bytes per lane, row stride S bytes):
//   0  operand read of v_mfma_*_16x16x32: 16 rows x 4 k-chunks
//   1  staging store, 8 lanes per 128-byte row
//   2  staging store, 4 lanes per 64-byte row
//   3  contiguous (the reference point)
//   4  one lane per row (a transposing store)
// Every address stays below 20 KB of the 32 KB array (checked by static_assert).
#include <hip/hip_runtime.h>

#include <cstdio>

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int PAT, int S>
__host__ __device__ constexpr int addr_of(int l) {
  return PAT == 0 ? (l % 16) * S + (l / 16) * 16
       : PAT == 1 ? (l / 8) * S + (l % 8) * 16
       : PAT == 2 ? (l / 4) * S + (l % 4) * 16
       : PAT == 4 ? l * S
       : l * 16;
}
template <int PAT, int S>
constexpr int max_addr() {
  int m = 0;
  for (int l = 0; l < 64; l++) m = addr_of<PAT, S>(l) > m ? addr_of<PAT, S>(l) : m;
  return m;
}

template <int PAT, int S, bool WRITE>
__global__ __launch_bounds__(64) void lds_probe(u32x4* out, int iters) {
  static_assert(max_addr<PAT, S>() + 16 + 4096 <= 20480 && S % 16 == 0, "pattern stays inside the array, aligned");
  __shared__ __align__(16) unsigned char buf[32768];
  const int l = threadIdx.x;
  for (int i = l; i < 32768 / 16; i += 64) reinterpret_cast<u32x4*>(buf)[i] = u32x4{(unsigned)i, 0u, 0u, 0u};
  __syncthreads();
  const int a = addr_of<PAT, S>(l);
  u32x4 acc = {0, 0, 0, 0};
  for (int it = 0; it < iters; it++) {
    u32x4* p = reinterpret_cast<u32x4*>(buf + a + (it & 1) * 4096);
    if (WRITE) {
      *p = acc + u32x4{(unsigned)it, 1u, 2u, 3u};
    } else {
      acc += *p;
    }
    __builtin_amdgcn_s_waitcnt(0);
  }
  __syncthreads();
  out[blockIdx.x * 64 + l] = acc + reinterpret_cast<u32x4*>(buf)[l];
}

template <int PAT, int S, bool W>
static int run(u32x4* out) {
  hipLaunchKernelGGL((lds_probe<PAT, S, W>), dim3(256), dim3(64), 0, 0, out, 1024);
  return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}

int main() {
  u32x4* out;
  if (hipMalloc(&out, 256 * 64 * sizeof(u32x4)) != hipSuccess) return 1;
  int bad = 0;
#define P(PAT, S) bad |= run<PAT, S, false>(out); bad |= run<PAT, S, true>(out);
  P(3, 16)
  P(0, 128) P(0, 144) P(0, 160) P(0, 176) P(0, 192) P(0, 208) P(0, 224) P(0, 256) P(0, 272) P(0, 288)
  P(1, 128) P(1, 144) P(1, 160) P(1, 176) P(1, 192) P(1, 208) P(1, 256) P(1, 272)
  P(2, 64) P(2, 80) P(2, 96) P(2, 112) P(2, 144) P(2, 160)
  P(4, 128) P(4, 144) P(4, 160) P(4, 176) P(4, 192) P(4, 208) P(4, 224) P(4, 240)
#undef P
  (void)hipFree(out);
  printf("lds bank probe done, status %d\n", bad);
  return bad;
}
