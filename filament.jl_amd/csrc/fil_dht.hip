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

// All FP32/FP64 kernels take the real matrix transposed, MT[k][m] = M[m][k],
// so an A tile (BK rows of k, BM consecutive m) is a contiguous copy.
//
// MFMA 16x16x4 f32 operand layout (lane l): a = A[l%16][l/16],
// b = B[l/16][l%16], c[i] = C[4*(l/16)+i][l%16].  Each wave owns a
// WM x WW (complex) sub-tile as (WM/16) x (WW/16) MFMA tiles for re and im.
// LDS row strides are padded to 16 (mod 64) floats so that the four k rows
// read by one MFMA operand load fall in disjoint bank groups.
template <int BM, int BW, int BK, int WM, int WW>
__global__ __launch_bounds__(64 * (BM / WM) * (BW / WW)) void dht_gemm_f32_kernel(
    const float* __restrict__ MT, const cx<float>* __restrict__ In, cx<float>* __restrict__ Out,
    const cx<float>* __restrict__ P, int Nr, int Nh, int Pw) {
  constexpr int MT_ = WM / 16, WT = WW / 16;
  constexpr int WAVES_M = BM / WM;
  constexpr int NT = 64 * (BM / WM) * (BW / WW);
  constexpr int SA = BM + 16, SB = BW + 16;
  __shared__ __align__(16) float As[2][BK][SA];
  __shared__ __align__(16) float Bre[2][BK][SB];
  __shared__ __align__(16) float Bim[2][BK][SB];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * BM, w0 = blockIdx.x * BW;
  const int wm = (wv % WAVES_M) * WM, ww = (wv / WAVES_M) * WW;
  const int l16 = lane & 15, l4 = lane >> 4;
  constexpr int NA = BK * BM / 4 / NT;    // float4 of A per thread
  constexpr int NB = BK * BW / 2 / NT;    // float4 (two complex) of B per thread
  static_assert(NA >= 1 && NB >= 1 && NA * 4 * NT == BK * BM && NB * 2 * NT == BK * BW, "tile shape");
  float4 ra[NA], rb[NB];
  // global -> registers (prefetch of the next k tile) and registers -> LDS,
  // written as macros so the prefetch arrays stay in registers
#define FIL_F32_GLOAD(k0_)                                                                      \
  _Pragma("unroll") for (int u = 0; u < NA; u++) {                                               \
    const int e = tid + NT * u, kk = e / (BM / 4), m4 = (e % (BM / 4)) * 4;                     \
    ra[u] = *reinterpret_cast<const float4*>(MT + (size_t)((k0_) + kk) * Nr + m0 + m4);           \
  }                                                                                              \
  _Pragma("unroll") for (int u = 0; u < NB; u++) {                                               \
    const int e = tid + NT * u, kk = e / (BW / 2), c2 = (e % (BW / 2)) * 2;                     \
    rb[u] = *reinterpret_cast<const float4*>(In + (size_t)((k0_) + kk) * Nh + w0 + c2);           \
  }
#define FIL_F32_LSTORE(buf_)                                                                    \
  _Pragma("unroll") for (int u = 0; u < NA; u++) {                                               \
    const int e = tid + NT * u, kk = e / (BM / 4), m4 = (e % (BM / 4)) * 4;                     \
    *reinterpret_cast<float4*>(&As[buf_][kk][m4]) = ra[u];                                       \
  }                                                                                              \
  _Pragma("unroll") for (int u = 0; u < NB; u++) {                                               \
    const int e = tid + NT * u, kk = e / (BW / 2), c2 = (e % (BW / 2)) * 2;                     \
    *reinterpret_cast<float2*>(&Bre[buf_][kk][c2]) = make_float2(rb[u].x, rb[u].z);              \
    *reinterpret_cast<float2*>(&Bim[buf_][kk][c2]) = make_float2(rb[u].y, rb[u].w);              \
  }
  floatx4 acc[MT_][WT][2];
#pragma unroll
  for (int a = 0; a < MT_; a++)
#pragma unroll
    for (int b = 0; b < WT; b++) { acc[a][b][0] = {0, 0, 0, 0}; acc[a][b][1] = {0, 0, 0, 0}; }

  const int nk = Nr / BK;
  FIL_F32_GLOAD(0);
  FIL_F32_LSTORE(0);
  __syncthreads();
  for (int kb = 0; kb < nk; kb++) {
    const int buf = kb & 1;
    // unconditional prefetch (the last iteration re-reads tile 0 into the idle buffer),
    // which keeps the prefetch registers out of scratch
    FIL_F32_GLOAD(((kb + 1) % nk) * BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int k = ks + l4;
      float af[MT_], br[WT], bi[WT];
#pragma unroll
      for (int a = 0; a < MT_; a++) af[a] = As[buf][k][wm + a * 16 + l16];
#pragma unroll
      for (int b = 0; b < WT; b++) { br[b] = Bre[buf][k][ww + b * 16 + l16]; bi[b] = Bim[buf][k][ww + b * 16 + l16]; }
#pragma unroll
      for (int a = 0; a < MT_; a++)
#pragma unroll
        for (int b = 0; b < WT; b++) {
          acc[a][b][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], br[b], acc[a][b][0], 0, 0, 0);
          acc[a][b][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[a], bi[b], acc[a][b][1], 0, 0, 0);
        }
    }
    FIL_F32_LSTORE(buf ^ 1);
    __syncthreads();
  }
#undef FIL_F32_GLOAD
#undef FIL_F32_LSTORE
#pragma unroll
  for (int a = 0; a < MT_; a++)
#pragma unroll
    for (int b = 0; b < WT; b++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = m0 + wm + a * 16 + 4 * l4 + i;
        const int w = w0 + ww + b * 16 + l16;
        cx<float> v = mk(acc[a][b][0][i], acc[a][b][1][i]);
        const size_t o = (size_t)m * Nh + w;
        if (P) v = v * P[(size_t)m * Pw + (w & (Pw - 1))];
        Out[o] = v;
      }
}

// FP64 variant on the vector ALUs (MI355X's FP64 vector and matrix peaks
// are equal): 64 x 64 complex tile, each thread a 4 x 4 block.
template <typename R>
__global__ __launch_bounds__(256) void dht_gemm_valu_kernel(const R* __restrict__ MT, const cx<R>* __restrict__ In,
                                                            cx<R>* __restrict__ Out, const cx<R>* __restrict__ P,
                                                            int Nr, int Nh, int Pw) {
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
      const int kk = e / BM, row = e % BM;
      As[kk][row] = MT[(size_t)(k0 + kk) * Nr + m0 + row];
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
      const int w = w0 + tw + 16 * y;
      const size_t o = (size_t)(m0 + tm + x) * Nh + w;
      cx<R> v = acc[x][y];
      if (P) v = v * P[(size_t)(m0 + tm + x) * Pw + (w & (Pw - 1))];
      Out[o] = v;
    }
}

// Tile shape by the size of one pulse's transform (Nr x Pw): a batch of
// pulses takes the same kernel, hence the same summation order, as each
// pulse on its own.
static bool wide_tiles(int Nr, int Pw, int tiling) {
  if (tiling == FIL_TILING_128) return true;
  if (tiling == FIL_TILING_64) return false;
  return Nr % 128 == 0 && (long)(Nr / 128) * (Pw / 64) >= 256;
}

template <typename R>
hipError_t launch_dht_gemm(const R* MT, const cx<R>* In, cx<R>* Out, const cx<R>* P, int Nr, int Nh, int Pw,
                           hipStream_t st, int tiling) {
  if (Nr % 64 || Nh % 64 || Pw < 1 || (Pw & (Pw - 1)) || Nh % Pw) return hipErrorInvalidValue;
  if (tiling == FIL_TILING_128 && Nr % 128) return hipErrorInvalidValue;
  if (sizeof(R) == 8) {
    fil_launch(dht_gemm_valu_kernel<R>, dim3(Nh / 64, Nr / 64), dim3(256), 0, st, MT, In, Out, P, Nr, Nh, Pw);
    return hipGetLastError();
  }
  const float* A = (const float*)MT;
  const cx<float>* B = (const cx<float>*)In;
  cx<float>* C = (cx<float>*)Out;
  const cx<float>* PP = (const cx<float>*)P;
  if (wide_tiles(Nr, Pw, tiling)) {
    // 8 waves of 32 x 32 (two per SIMD: one workgroup per CU at Nr = 1024)
    fil_launch(dht_gemm_f32_kernel<128, 64, 32, 32, 32>, dim3(Nh / 64, Nr / 128), dim3(512), 0, st, A, B, C, PP, Nr,
               Nh, Pw);
  } else {
    fil_launch(dht_gemm_f32_kernel<64, 64, 16, 32, 32>, dim3(Nh / 64, Nr / 64), dim3(256), 0, st, A, B, C, PP, Nr, Nh,
               Pw);
  }
  return hipGetLastError();
}

// P = exp(i KZ dz) * KG, FP64 arithmetic, rounded to R.
template <typename R>
__global__ void propagator_kernel(const double2* __restrict__ KZ, const double* __restrict__ KG, double dz_arg,
                                  const double* __restrict__ dz_dev, cx<R>* __restrict__ P, size_t n) {
  size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double dz = dz_dev ? *dz_dev : dz_arg;
  double2 kz = KZ[q];
  double amp = exp(-kz.y * dz) * KG[q];
  double s, c;
  sincos(kz.x * dz, &s, &c);
  P[q] = mk((R)(amp * c), (R)(amp * s));
}

template <typename R>
hipError_t launch_propagator(const double* KZ64, const double* KG64, double dz, cx<R>* P, int Nr, int Nh,
                             hipStream_t st, const double* dz_dev) {
  size_t n = (size_t)Nr * Nh;
  hipLaunchKernelGGL(propagator_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const double2*>(KZ64), KG64, dz, dz_dev, P, n);
  return hipGetLastError();
}

template hipError_t launch_dht_gemm<float>(const float*, const cx<float>*, cx<float>*, const cx<float>*, int, int, int,
                                           hipStream_t, int);
template hipError_t launch_dht_gemm<double>(const double*, const cx<double>*, cx<double>*, const cx<double>*, int, int,
                                            int, hipStream_t, int);
template hipError_t launch_propagator<float>(const double*, const double*, double, cx<float>*, int, int, hipStream_t,
                                             const double*);
template hipError_t launch_propagator<double>(const double*, const double*, double, cx<double>*, int, int, hipStream_t,
                                              const double*);

// ---------------------------------------------------------------------------
// Split-BF16 3-term form: with x = x_hi + x_lo (x_hi = bf16(x),
// x_lo = bf16(x - x_hi)), M*S ~ M_hi S_hi + M_hi S_lo + M_lo S_hi, each on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  M is split once on the
// host.  Operand layout (lane l): a[i] = A[l%16][8*(l/16)+i],
// b[i] = B[8*(l/16)+i][l%16], c[i] = C[4*(l/16)+i][l%16].
//
// The two products of a step run as
//   X = P .* (M1 S):  S (FP32, [k][w]) is split when its tile is staged; the
//                     epilogue multiplies by P and writes X already split, as
//                     four BF16 planes [re_hi, re_lo, im_hi, im_lo][w][m]
//                     (m contiguous: the k of the second product);
//   S = M2 X:         both operands arrive split and k-contiguous, so staging
//                     is a plain copy and the kernel is MFMA work only.
// The split values are the ones the FP32 X would give, so the result does not
// depend on which form X travels in.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t rsrc_t;

// raw buffer resource over [p, p + bytes) (gfx9 word 3: raw 32-bit access);
// loads outside the range return zero
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, size_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint4 load16(rsrc_t r, int voff, int soff) {
  const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ u32x2 load8(rsrc_t r, int voff, int soff) {
  return __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
}

// BM (m) x BW (complex w) x BK (k) tiles, NWM x NWW waves of (BM/NWM) x (BW/NWW)
// per k group; with KS = 2 two groups of waves take alternate 32-k halves of
// every k tile (two waves per SIMD at the LDS traffic of one group's register
// tiles) and add their partial sums once at the end; B stored
// [w][k] in LDS so each lane's MFMA operand (8 consecutive k) is one 16-byte
// LDS read; register-prefetch double buffering.
// BIN_PLANES: B is the four BF16 planes [4][Nh][Nr] (else FP32 complex [Nr][Nh]).
// OUT_PLANES: the output is written as the four planes [4][Nh][Nr] (else FP32
// complex [Nr][Nh]); P (optional) multiplies the output first.
template <int BM, int BW, int BK, int NWM, int NWW, int KS, bool BIN_PLANES, bool OUT_PLANES>
__global__ __launch_bounds__(64 * NWM * NWW * KS) void dht_gemm_bf16x3_kernel(const uint16_t* __restrict__ Mhi,
                                                              const uint16_t* __restrict__ Mlo,
                                                              const void* __restrict__ In_, void* __restrict__ Out_,
                                                              const cx<float>* __restrict__ P, int Nr, int Nh, int Pw) {
  // padded row of (BK + 16) bf16: a 160-byte row stride keeps both the
  // staging stores (8 lanes per 128-byte row) and the operand reads (16 rows x
  // 4 k-chunks per instruction) free of LDS bank conflicts on gfx950; a
  // 144-byte stride doubles the operand reads
  constexpr int SK = BK + 16;
  constexpr int NT = 64 * NWM * NWW * KS;
  constexpr int WM = BM / NWM, WW = BW / NWW;
  constexpr int MT = WM / 16, WT = WW / 16;
  static_assert((KS == 1 || KS == 2) && BK % (32 * KS) == 0 && (KS == 1 || MT % 2 == 0), "k groups");
  __shared__ __align__(16) uint16_t As[2][2][BM][SK];      // [buf][hi/lo][m][k]
  __shared__ __align__(16) uint16_t Bs[2][4][BW][SK];      // [buf][re_hi,re_lo,im_hi,im_lo][w][k]
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int m0 = blockIdx.y * BM, w0 = blockIdx.x * BW;
  const int kg = wv / (NWM * NWW), wq = wv % (NWM * NWW);   // k group, wave within the group
  const int wm = (wq / NWW) * WM, ww = (wq % NWW) * WW;
  const int l16 = lane & 15, l4 = lane >> 4;
  constexpr int NA = BM * BK / 8 / NT;      // 16-byte chunks of each A plane per thread
  constexpr int NB = BW * BK / 8 / NT;      // 16-byte chunks (one w, 8 consecutive k) of each B plane per thread
  static_assert(NA >= 1 && NB >= 1 && NA * 8 * NT == BM * BK && NB * 8 * NT == BW * BK && BK % 32 == 0 &&
                WM % 16 == 0 && WW % 16 == 0, "tile shape");
  const size_t plane = (size_t)Nh * Nr;     // elements of one BF16 plane of X
  // two sets of staging registers: tile kb + 2 is loaded while tile kb is
  // multiplied, so each load has two MFMA phases to arrive
  uint4 rah[2][NA], ral[2][NA];
  // B staging registers: 8 complex values (FP32 input) or 4 planes x 8 bf16
  uint4 rbp[2][BIN_PLANES ? NB : 1][4];
  u32x2 rbf[2][BIN_PLANES ? 1 : NB][8];
  // Buffer loads: a per-array resource (scalar registers) plus one 32-bit
  // byte offset per thread and chunk, the k position in the scalar offset.
  // Offsets stay below 2^31 (checked by the launcher).
  const rsrc_t rMh = make_rsrc(Mhi, (size_t)Nr * Nr * 2), rMl = make_rsrc(Mlo, (size_t)Nr * Nr * 2);
  rsrc_t rB[BIN_PLANES ? 4 : 1];
  if constexpr (BIN_PLANES) {
#pragma unroll
    for (int q = 0; q < 4; q++) rB[q] = make_rsrc(reinterpret_cast<const uint16_t*>(In_) + q * plane, plane * 2);
  } else {
    rB[0] = make_rsrc(In_, (size_t)Nr * Nh * 8);
  }
  int offA[NA], offB[NB];
#pragma unroll
  for (int u = 0; u < NA; u++) {
    const int e = tid + NT * u, row = e / (BK / 8), kq = (e % (BK / 8)) * 8;
    offA[u] = ((m0 + row) * Nr + kq) * 2;
  }
#pragma unroll
  for (int u = 0; u < NB; u++) {
    const int e = tid + NT * u;
    if constexpr (BIN_PLANES) {
      const int bw = e / (BK / 8), kq = (e % (BK / 8)) * 8;
      offB[u] = ((w0 + bw) * Nr + kq) * 2;
    } else {
      const int bw = e % BW, bkg = e / BW;
      offB[u] = (8 * bkg * Nh + w0 + bw) * 8;
    }
  }
#define FIL_BF_GLOAD(r_, k0_)                                                                     \
  _Pragma("unroll") for (int u = 0; u < NA; u++) {                                                 \
    rah[r_][u] = load16(rMh, offA[u], (k0_) * 2);                                                  \
    ral[r_][u] = load16(rMl, offA[u], (k0_) * 2);                                                  \
  }                                                                                                \
  _Pragma("unroll") for (int u = 0; u < NB; u++) {                                                 \
    if constexpr (BIN_PLANES) {                                                                    \
      _Pragma("unroll") for (int q = 0; q < 4; q++) rbp[r_][u][q] = load16(rB[q], offB[u], (k0_) * 2); \
    } else {                                                                                       \
      _Pragma("unroll") for (int i = 0; i < 8; i++) rbf[r_][u][i] = load8(rB[0], offB[u], ((k0_) + i) * Nh * 8); \
    }                                                                                              \
  }
#define FIL_BF_LSTORE(r_, buf_)                                                                   \
  _Pragma("unroll") for (int u = 0; u < NA; u++) {                                                 \
    const int e = tid + NT * u, row = e / (BK / 8), kq = (e % (BK / 8)) * 8;                      \
    *reinterpret_cast<uint4*>(&As[buf_][0][row][kq]) = rah[r_][u];                                 \
    *reinterpret_cast<uint4*>(&As[buf_][1][row][kq]) = ral[r_][u];                                 \
  }                                                                                                \
  _Pragma("unroll") for (int u = 0; u < NB; u++) {                                                 \
    const int e = tid + NT * u;                                                                   \
    if constexpr (BIN_PLANES) {                                                                    \
      const int bw = e / (BK / 8), kq = (e % (BK / 8)) * 8;                                        \
      _Pragma("unroll") for (int q = 0; q < 4; q++) *reinterpret_cast<uint4*>(&Bs[buf_][q][bw][kq]) = rbp[r_][u][q]; \
    } else {                                                                                       \
      const int bw = e % BW, bkg = e / BW;                                                         \
      bf16x8 rh, rl, ih, il;                                                                       \
      _Pragma("unroll") for (int i = 0; i < 8; i++) {                                              \
        const float xr = __uint_as_float(rbf[r_][u][i].x), xi = __uint_as_float(rbf[r_][u][i].y);  \
        __bf16 hh = (__bf16)xr;                                                                    \
        rh[i] = hh; rl[i] = (__bf16)(xr - (float)hh);                                              \
        hh = (__bf16)xi;                                                                           \
        ih[i] = hh; il[i] = (__bf16)(xi - (float)hh);                                              \
      }                                                                                            \
      *reinterpret_cast<bf16x8*>(&Bs[buf_][0][bw][8 * bkg]) = rh;                                  \
      *reinterpret_cast<bf16x8*>(&Bs[buf_][1][bw][8 * bkg]) = rl;                                  \
      *reinterpret_cast<bf16x8*>(&Bs[buf_][2][bw][8 * bkg]) = ih;                                  \
      *reinterpret_cast<bf16x8*>(&Bs[buf_][3][bw][8 * bkg]) = il;                                  \
    }                                                                                              \
  }
  floatx4 acc[MT][WT][2];
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < WT; b++) { acc[a][b][0] = {0, 0, 0, 0}; acc[a][b][1] = {0, 0, 0, 0}; }
#define FIL_BF_COMPUTE(buf_)                                                                      \
  _Pragma("unroll") for (int s_ = 0; s_ < BK / (32 * KS); s_++) {                                 \
    const int ks = (s_ * KS + kg) * 32;                                                            \
    bf16x8 ah[MT], al[MT];                                                                         \
    _Pragma("unroll") for (int a = 0; a < MT; a++) {                                               \
      ah[a] = *reinterpret_cast<const bf16x8*>(&As[buf_][0][wm + a * 16 + l16][ks + 8 * l4]);      \
      al[a] = *reinterpret_cast<const bf16x8*>(&As[buf_][1][wm + a * 16 + l16][ks + 8 * l4]);      \
    }                                                                                              \
    bf16x8 bp[WT][4];                                                                              \
    _Pragma("unroll") for (int b = 0; b < WT; b++) {                                               \
      const int col = ww + b * 16 + l16;                                                           \
      _Pragma("unroll") for (int q = 0; q < 4; q++)                                                \
        bp[b][q] = *reinterpret_cast<const bf16x8*>(&Bs[buf_][q][col][ks + 8 * l4]);               \
    }                                                                                              \
    /* the three terms in turn over all accumulators: consecutive MFMAs are   */                  \
    /* independent; each accumulator still sums lo*hi, hi*lo, hi*hi in order  */                  \
    _Pragma("unroll") for (int term = 0; term < 3; term++)                                         \
    _Pragma("unroll") for (int b = 0; b < WT; b++)                                                 \
    _Pragma("unroll") for (int a = 0; a < MT; a++)                                                 \
    _Pragma("unroll") for (int c = 0; c < 2; c++) {                                                \
      const bf16x8 av = term == 0 ? al[a] : ah[a];                                                 \
      const bf16x8 bv = term == 1 ? bp[b][2 * c + 1] : bp[b][2 * c];                               \
      acc[a][b][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b][c], 0, 0, 0);       \
    }                                                                                              \
  }
  // LDS buffer b holds tile kb while register set b receives tile kb + 2;
  // the prefetches past the last tile re-read valid tiles (kept unconditional
  // so the registers stay out of scratch)
  const int nk = Nr / BK;
  FIL_BF_GLOAD(0, 0);
  FIL_BF_GLOAD(1, (1 % nk) * BK);
  FIL_BF_LSTORE(0, 0);
  __syncthreads();
  const int nk2 = nk & ~1;
  for (int kb = 0; kb < nk2; kb += 2) {
    FIL_BF_GLOAD(0, ((kb + 2) % nk) * BK);
    FIL_BF_COMPUTE(0);
    FIL_BF_LSTORE(1, 1);
    __syncthreads();
    FIL_BF_GLOAD(1, ((kb + 3) % nk) * BK);
    FIL_BF_COMPUTE(1);
    FIL_BF_LSTORE(0, 0);
    __syncthreads();
  }
  if (nk & 1) FIL_BF_COMPUTE(0);   // the last tile of an odd count sits in buffer 0
#undef FIL_BF_COMPUTE
#undef FIL_BF_GLOAD
#undef FIL_BF_LSTORE
  if constexpr (KS == 2) {
    // The k groups add their partial sums: group 0 finalizes the row blocks
    // a < MT/2, group 1 the others, each handing the other group's wave with
    // the same (wm, ww) its half through LDS (the tiles are no longer read).
    constexpr int H = MT / 2, NV = H * WT * 2 * 4;
    static_assert(2 * NWM * NWW * NV * 64 * sizeof(float) <= sizeof(As), "exchange buffer");
    __syncthreads();
    float* xr = reinterpret_cast<float*>(&As[0][0][0][0]);
    float* mine = xr + (size_t)(kg * (NWM * NWW) + wq) * NV * 64 + lane;
    const float* other = xr + (size_t)((1 - kg) * (NWM * NWW) + wq) * NV * 64 + lane;
#pragma unroll
    for (int a = 0; a < H; a++)
#pragma unroll
      for (int b = 0; b < WT; b++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int v = ((a * WT + b) * 2 + c) * 4 + i;
            mine[v * 64] = kg == 0 ? acc[H + a][b][c][i] : acc[a][b][c][i];
          }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < H; a++)
#pragma unroll
      for (int b = 0; b < WT; b++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int v = ((a * WT + b) * 2 + c) * 4 + i;
            if (kg == 0) acc[a][b][c][i] += other[v * 64];
            else acc[H + a][b][c][i] += other[v * 64];
          }
  }
#pragma unroll
  for (int a = 0; a < MT; a++)
#pragma unroll
    for (int b = 0; b < WT; b++) {
      if (KS == 2 && ((a < MT / 2) != (kg == 0))) continue;   // the other group's row blocks
      const int w = w0 + ww + b * 16 + l16;
      const int mb = m0 + wm + a * 16 + 4 * l4;
      if constexpr (OUT_PLANES) {
        bf16x4 rh, rl, ih, il;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          cx<float> v = mk(acc[a][b][0][i], acc[a][b][1][i]);
          if (P) v = v * P[(size_t)(mb + i) * Pw + (w & (Pw - 1))];
          __bf16 hh = (__bf16)v.x;
          rh[i] = hh; rl[i] = (__bf16)(v.x - (float)hh);
          hh = (__bf16)v.y;
          ih[i] = hh; il[i] = (__bf16)(v.y - (float)hh);
        }
        __bf16* Xp = reinterpret_cast<__bf16*>(Out_) + (size_t)w * Nr + mb;
        *reinterpret_cast<bf16x4*>(Xp) = rh;
        *reinterpret_cast<bf16x4*>(Xp + plane) = rl;
        *reinterpret_cast<bf16x4*>(Xp + 2 * plane) = ih;
        *reinterpret_cast<bf16x4*>(Xp + 3 * plane) = il;
      } else {
        cx<float>* Out = reinterpret_cast<cx<float>*>(Out_);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          cx<float> v = mk(acc[a][b][0][i], acc[a][b][1][i]);
          if (P) v = v * P[(size_t)(mb + i) * Pw + (w & (Pw - 1))];
          Out[(size_t)(mb + i) * Nh + w] = v;
        }
      }
    }
}

template <bool BIN_PLANES, bool OUT_PLANES>
static hipError_t launch_bf16x3(const uint16_t* Mhi, const uint16_t* Mlo, const void* In, void* Out,
                                const cx<float>* P, int Nr, int Nh, int Pw, hipStream_t st, int tiling) {
  if (Nr % 64 || Nh % 64 || Pw < 1 || (Pw & (Pw - 1)) || Nh % Pw) return hipErrorInvalidValue;
  if (tiling == FIL_TILING_128 && Nr % 128) return hipErrorInvalidValue;
  // the kernels address their operands with 32-bit byte offsets
  if ((size_t)Nr * Nh * 8 >= ((size_t)1 << 31)) return hipErrorInvalidValue;
  // 8 waves (two per SIMD): one workgroup per CU at Nr = 1024 (the LDS stages
  // fill the CU), so the second wave per SIMD is what hides the LDS and MFMA
  // latencies; as two k groups of 2 x 2 waves their fragment reads from LDS
  // are those of 4 waves (43.0/34.3 -> 40.3/30.6 us at 1024 x 4096 against
  // one group of 4 x 2 waves)
  constexpr int KS = 2, NWM = 2, NWW = 2;
  if (wide_tiles(Nr, Pw, tiling)) {
    fil_launch(dht_gemm_bf16x3_kernel<128, 64, 64, NWM, NWW, KS, BIN_PLANES, OUT_PLANES>, dim3(Nh / 64, Nr / 128),
               dim3(64 * NWM * NWW * KS), 0, st, Mhi, Mlo, In, Out, P, Nr, Nh, Pw);
  } else {
    fil_launch(dht_gemm_bf16x3_kernel<64, 64, 64, NWM, NWW, KS, BIN_PLANES, OUT_PLANES>, dim3(Nh / 64, Nr / 64),
               dim3(64 * NWM * NWW * KS), 0, st, Mhi, Mlo, In, Out, P, Nr, Nh, Pw);
  }
  return hipGetLastError();
}

hipError_t launch_dht_gemm_bf16x3_to_planes(const uint16_t* Mhi, const uint16_t* Mlo, const cx<float>* In,
                                            uint16_t* Xp, const cx<float>* P, int Nr, int Nh, int Pw, hipStream_t st,
                                            int tiling) {
  return launch_bf16x3<false, true>(Mhi, Mlo, In, Xp, P, Nr, Nh, Pw, st, tiling);
}

hipError_t launch_dht_gemm_bf16x3_from_planes(const uint16_t* Mhi, const uint16_t* Mlo, const uint16_t* Xp,
                                              cx<float>* Out, int Nr, int Nh, int Pw, hipStream_t st, int tiling) {
  return launch_bf16x3<true, false>(Mhi, Mlo, Xp, Out, nullptr, Nr, Nh, Pw, st, tiling);
}

}  // namespace fil
