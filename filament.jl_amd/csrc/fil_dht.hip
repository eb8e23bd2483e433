// fil_dht.hip -- the Hankel-transform phase of one z-step on the MFMA units.
//
// Reference: propagate!(E, LP::LinearPropagatorSpectral, dz)
// (src/modules/linear_propagators/spectral.jl:16-26): forward QDHT of the
// positive-frequency columns (HankelTransforms `P * E`, CUBLAS GEMM),
// E .*= exp(i KZ dz), spectral guard Kguard .* Wguard (Guards.jl:283-298),
// inverse QDHT.  With forward = diag(JV) T diag(RJ) and inverse =
// diag(1/RJ) T diag(1/JV), the JV factors cancel around the diagonal
// propagator, so the step is Out = M2 * (P .* (M1 * In)) with
// M1 = T diag(RJ), M2 = diag(1/RJ) T (folded on the host in FP64).
//
// GEMM shape per transform: M = Nr (k or r), N = Nt/2 complex columns,
// K = Nr.  The real matrix is the A operand; the complex field is the B
// operand with re and im planes split in LDS so that each lane holds the
// real and imaginary part of the same output element, and the propagator
// multiply runs in registers in the epilogue.
#include "fil_internal.h"
#include "fil_kernels.h"

namespace fil {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename R> struct Acc;
template <> struct Acc<float> {
  typedef floatx4 T;
  static __device__ __forceinline__ T mfma(float a, float b, T c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
};

// Tile: BM rows x BW complex columns per workgroup, BK = 16, 4 waves in 2x2,
// each wave BM/2 x BW/2 as (BM/32) x (BW/32) 16x16 MFMA tiles for re and im.
// MFMA 16x16x4 operand layout (lane l): a = A[l%16][l/16], b = B[l/16][l%16],
// c[i] = C[4*(l/16)+i][l%16].
template <typename R, int BM, int BW>
__global__ __launch_bounds__(256) void dht_gemm_kernel(const R* __restrict__ M, const cx<R>* __restrict__ In,
                                                       cx<R>* __restrict__ Out, const cx<R>* __restrict__ P,
                                                       int Nr, int Nh) {
  constexpr int BK = 16;
  constexpr int MT = BM / 32, WT = BW / 32;   // MFMA tiles per wave in m and w
  __shared__ R As[BK][BM + 4];
  __shared__ R Bre[BK][BW + 4];
  __shared__ R Bim[BK][BW + 4];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * BM, w0 = blockIdx.x * BW;
  const int wm = (wv >> 1) * (BM / 2), ww = (wv & 1) * (BW / 2);
  typename Acc<R>::T acc[MT][WT][2];
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < WT; b++) { acc[a][b][0] = {0, 0, 0, 0}; acc[a][b][1] = {0, 0, 0, 0}; }

  const int l16 = lane & 15, l4 = lane >> 4;
  for (int k0 = 0; k0 < Nr; k0 += BK) {
    // A tile: BM x BK, stored k-major
    for (int e = tid; e < BM * BK; e += 256) {
      const int row = e / BK, kk = e % BK;
      As[kk][row] = M[(size_t)(m0 + row) * Nr + k0 + kk];
    }
    // B tile: BK x BW complex, split into re / im planes
    for (int e = tid; e < BK * BW; e += 256) {
      const int kk = e / BW, c = e % BW;
      cx<R> v = In[(size_t)(k0 + kk) * Nh + w0 + c];
      Bre[kk][c] = v.x;
      Bim[kk][c] = v.y;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int k = ks + l4;
      R af[MT], br[WT], bi[WT];
#pragma unroll
      for (int a = 0; a < MT; a++) af[a] = As[k][wm + a * 16 + l16];
#pragma unroll
      for (int b = 0; b < WT; b++) { br[b] = Bre[k][ww + b * 16 + l16]; bi[b] = Bim[k][ww + b * 16 + l16]; }
#pragma unroll
      for (int a = 0; a < MT; a++)
#pragma unroll
        for (int b = 0; b < WT; b++) {
          acc[a][b][0] = Acc<R>::mfma(af[a], br[b], acc[a][b][0]);
          acc[a][b][1] = Acc<R>::mfma(af[a], bi[b], acc[a][b][1]);
        }
    }
    __syncthreads();
  }
  // epilogue: optional propagator multiply, interleaved store
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < WT; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = m0 + wm + a * 16 + 4 * l4 + i;
        const int w = w0 + ww + b * 16 + l16;
        cx<R> v = mk((R)acc[a][b][0][i], (R)acc[a][b][1][i]);
        const size_t o = (size_t)m * Nh + w;
        if (P) v = v * P[o];
        Out[o] = v;
      }
}

// FP64 variant on the vector ALUs (MI355X's FP64 vector and matrix peaks
// are equal): 64 x 64 complex tile, each thread a 4 x 4 block.
template <typename R>
__global__ __launch_bounds__(256) void dht_gemm_valu_kernel(const R* __restrict__ M, const cx<R>* __restrict__ In,
                                                            cx<R>* __restrict__ Out, const cx<R>* __restrict__ P,
                                                            int Nr, int Nh) {
  constexpr int BM = 64, BW = 64, BK = 16;
  __shared__ R As[BK][BM + 1];
  __shared__ cx<R> Bs[BK][BW];
  const int tid = threadIdx.x;
  const int m0 = blockIdx.y * BM, w0 = blockIdx.x * BW;
  const int tm = (tid / 16) * 4, tw = (tid % 16);
  cx<R> acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = mk((R)0, (R)0);
  for (int k0 = 0; k0 < Nr; k0 += BK) {
    for (int e = tid; e < BM * BK; e += 256) {
      const int row = e / BK, kk = e % BK;
      As[kk][row] = M[(size_t)(m0 + row) * Nr + k0 + kk];
    }
    for (int e = tid; e < BK * BW; e += 256) {
      const int kk = e / BW, c = e % BW;
      Bs[kk][c] = In[(size_t)(k0 + kk) * Nh + w0 + c];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      R a[4];
      cx<R> b[4];
#pragma unroll
      for (int q = 0; q < 4; q++) { a[q] = As[kk][tm + q]; b[q] = Bs[kk][tw + 16 * q]; }
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) { acc[x][y].x += a[x] * b[y].x; acc[x][y].y += a[x] * b[y].y; }
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const size_t o = (size_t)(m0 + tm + x) * Nh + w0 + tw + 16 * y;
      cx<R> v = acc[x][y];
      if (P) v = v * P[o];
      Out[o] = v;
    }
}

template <typename R>
hipError_t launch_dht_gemm(const R* M, const cx<R>* In, cx<R>* Out, const cx<R>* P, int Nr, int Nh, hipStream_t st) {
  constexpr int BM = 64, BW = 64;
  if (Nr % BM || Nh % BW) return hipErrorInvalidValue;
  dim3 grid(Nh / BW, Nr / BM);
  if (sizeof(R) == 8)
    hipLaunchKernelGGL((dht_gemm_valu_kernel<R>), grid, dim3(256), 0, st, M, In, Out, P, Nr, Nh);
  else
    hipLaunchKernelGGL((dht_gemm_kernel<float, BM, BW>), grid, dim3(256), 0, st, (const float*)M,
                       (const cx<float>*)In, (cx<float>*)Out, (const cx<float>*)P, Nr, Nh);
  return hipGetLastError();
}

// P = exp(i KZ dz) * KG, FP64 arithmetic, rounded to R.
template <typename R>
__global__ void propagator_kernel(const double2* __restrict__ KZ, const double* __restrict__ KG, double dz,
                                  cx<R>* __restrict__ P, size_t n) {
  size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  double2 kz = KZ[q];
  double amp = exp(-kz.y * dz) * KG[q];
  double s, c;
  sincos(kz.x * dz, &s, &c);
  P[q] = mk((R)(amp * c), (R)(amp * s));
}

template <typename R>
hipError_t launch_propagator(const double* KZ64, const double* KG64, double dz, cx<R>* P, int Nr, int Nh,
                             hipStream_t st) {
  size_t n = (size_t)Nr * Nh;
  hipLaunchKernelGGL(propagator_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const double2*>(KZ64), KG64, dz, P, n);
  return hipGetLastError();
}

template hipError_t launch_dht_gemm<float>(const float*, const cx<float>*, cx<float>*, const cx<float>*, int, int, hipStream_t);
template hipError_t launch_dht_gemm<double>(const double*, const cx<double>*, cx<double>*, const cx<double>*, int, int, hipStream_t);
template hipError_t launch_propagator<float>(const double*, const double*, double, cx<float>*, int, int, hipStream_t);
template hipError_t launch_propagator<double>(const double*, const double*, double, cx<double>*, int, int, hipStream_t);

hipError_t launch_dht_gemm_bf16x3(const uint16_t*, const uint16_t*, const cx<float>*, cx<float>*, const cx<float>*,
                                  int, int, hipStream_t) {
  return hipErrorNotSupported;
}

}  // namespace fil
