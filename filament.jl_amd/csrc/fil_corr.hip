// fil_corr.hip -- network of the learned coarse-step corrector (no counterpart
// in the reference): forward and backward of 3x3 convolutions, MSE loss and
// Adam, all on the device.
//
// Layer l maps [B][cin][H][W] -> [B][cout][H][W] with zero padding:
//   out[co][y][x] = b[co] + sum_ci sum_{dy,dx in -1..1} w[co][ci][(dy+1)*3+(dx+1)] in[ci][y+dy][x+dx]
// Weights are read as BF16 (the rounded copy of the FP32 master parameters)
// and every product is accumulated in FP32.  The forward pass runs on the
// matrix cores (FP32 MFMA); the backward passes own one pixel (data gradient)
// or one weight (weight gradient) per thread on the vector ALUs.
#include <algorithm>

#include "fil_internal.h"
#include "fil_kernels.h"

namespace fil {

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int CT = 128;   // threads per block of the convolution kernels
constexpr int CG = 8;     // channels per thread in the forward / data-gradient kernels

// Forward on the matrix cores: the layer is an implicit GEMM
//   out[co][p] = b[co] + sum_k A[co][k] X[k][p],   k = tap * CIN + ci,
//   A[co][k] = w[co][ci][tap] (BF16 values, exact in FP32),
//   X[k][p]  = in[ci][p + d(tap)]  (zero outside the image),
// on v_mfma_f32_16x16x4f32: FP32 operands, FP32 accumulation.  A block owns an
// MY x MX pixel tile of one sample and 16 output channels (blockIdx.z): their
// weight rows (copied as they are from the forward layout, whose row stride
// puts the 16 rows of an MFMA operand in distinct LDS banks) and the input
// tile with its one-pixel halo sit in LDS; wave w computes the 16 pixels of
// row w / 2, columns (w % 2) * 16 .. + 15.  k runs tap-major so a 4-wide k
// step is four consecutive input channels at one tap.
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int MY = 2, MX = 32;             // pixel tile of the forward kernel (one-row tiles measured slower)
constexpr int MT = 64 * (MY * MX / 16);    // threads: one wave per 16 pixels
// channel stride of the staged input tile: = 16 (mod 32), so the two k rows a
// half-wave reads (adjacent channels) fall in different LDS banks
constexpr int conv_x_stride() { return ((MY + 2) * (MX + 2) + 15) / 32 * 32 + 16; }

// POOL: the input is the per-row feature array Frow [H pr][CIN][W] of one
// sample, averaged over the pr rows of each pooled row while it is staged
// (the same sums, in the same order, as corr_pool_rows_kernel).
template <int CIN, int COUT, bool POOL = false>
__global__ __launch_bounds__(MT) void conv_fwd_kernel(const float* __restrict__ in, const int* __restrict__ idx,
                                                      const float* __restrict__ wfm, const float* __restrict__ bias,
                                                      float* __restrict__ out, int H, int W, int relu, int pr = 1) {
  static_assert(CIN % 4 == 0, "a k step covers four input channels");
  constexpr int K = 9 * CIN, AS = (K + 31) / 32 * 32 + 2;
  constexpr int HY = MY + 2, HX = MX + 2, XS = conv_x_stride();
  static_assert((16 * AS) % 4 == 0, "weight rows are copied as float4");
  __shared__ __align__(16) float wl[16 * AS];
  __shared__ float tin[CIN * XS];
  const int tid = threadIdx.x, mb = blockIdx.z;
  const int HW = H * W;
  const int b = blockIdx.y;
  const int ntx = (W + MX - 1) / MX;
  const int y0 = (blockIdx.x / ntx) * MY, x0 = (blockIdx.x % ntx) * MX;
  const float* inb = in + (size_t)(idx ? idx[b] : b) * CIN * HW;
  // staging: every global load of the thread is issued before the first LDS
  // store, so their latencies overlap
  constexpr int NW4 = 16 * AS / 4, PW = (NW4 + MT - 1) / MT;
  constexpr int NIN = CIN * HY * HX, PI = (NIN + MT - 1) / MT;
  const float4* src = reinterpret_cast<const float4*>(wfm + (size_t)mb * 16 * AS);
  float4 wreg[PW];
  float xv_[PI];
#pragma unroll
  for (int j = 0; j < PW; j++) {
    const int q = tid + j * MT;
    wreg[j] = src[q < NW4 ? q : NW4 - 1];   // clamped: the extra copies are not stored
  }
#pragma unroll
  for (int j = 0; j < PI; j++) {
    const int e = tid + j * MT;
    const int ci = e / (HY * HX), r = e % (HY * HX), yy = y0 - 1 + r / HX, xx = x0 - 1 + r % HX;
    const bool inside = e < NIN && yy >= 0 && yy < H && xx >= 0 && xx < W;
    if constexpr (POOL) {
      float s = 0.0f;
      if (inside) {
        const float* src = in + ((size_t)yy * pr * CIN + ci) * W + xx;
#pragma unroll 8
        for (int k = 0; k < pr; k++) s += src[(size_t)k * CIN * W];
      }
      xv_[j] = inside ? s / (float)pr : 0.0f;
    } else {
      xv_[j] = inside ? inb[(size_t)ci * HW + yy * W + xx] : 0.0f;
    }
  }
#pragma unroll
  for (int j = 0; j < PW; j++) {
    const int q = tid + j * MT;
    if (q < NW4) reinterpret_cast<float4*>(wl)[q] = wreg[j];
  }
#pragma unroll
  for (int j = 0; j < PI; j++) {
    const int e = tid + j * MT;
    if (e < NIN) tin[(e / (HY * HX)) * XS + e % (HY * HX)] = xv_[j];
  }
  __syncthreads();
  const int wv = tid >> 6, lane = tid & 63, l16 = lane & 15, l4 = lane >> 4;
  const int py = wv >> 1, px = (wv & 1) * 16 + l16;
  // two interleaved accumulation chains (even and odd k steps), fully
  // unrolled, so the LDS operand reads run ahead of the MFMA chain
  floatx4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
  const float* ar = wl + l16 * AS + l4;
  const float* xr = tin + l4 * XS + py * HX + px;
#pragma unroll
  for (int s = 0; s < K / 4; s++) {
    const int tap = s / (CIN / 4), c4 = (4 * s) % CIN;
    const float xv = xr[c4 * XS + (tap / 3) * HX + tap % 3];
    if (s & 1) acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[4 * s], xv, acc1, 0, 0, 0);
    else acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[4 * s], xv, acc0, 0, 0, 0);
  }
  const floatx4 acc = acc0 + acc1;
  const int y = y0 + py, x = x0 + px;
  if (y >= H || x >= W) return;
  float* ob = out + (size_t)b * COUT * HW + (size_t)y * W + x;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int co = mb * 16 + 4 * l4 + i;
    if (co < COUT) {
      const float v = acc[i] + bias[co];
      ob[(size_t)co * HW] = relu ? fmaxf(v, 0.0f) : v;
    }
  }
}

// wfm of every layer from the flat BF16 parameters (one thread per entry)
__global__ void conv_weight_layout_kernel(const uint16_t* __restrict__ wbf, float* __restrict__ wfm) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= corr_wfm_size()) return;
  int l = 0;
  while (l + 1 < CORR_LAYERS && q >= corr_wfm_off(l + 1)) l++;
  const int r = q - corr_wfm_off(l), AS = corr_a_stride(l), cin = corr_cin(l);
  const int co = r / AS, k = r % AS, tap = k / cin, ci = k % cin;
  float v = 0.0f;
  if (co < corr_cout(l) && k < 9 * cin) v = bf2f(wbf[corr_woff(l) + (co * cin + ci) * 9 + tap]);
  wfm[q] = v;
}

// din[ci][q] = sum_co sum_tap w[co][ci][tap] dout[co][q - d(tap)], times [act > 0];
// input channels in groups of G per thread (blockIdx.z)
template <int CIN, int COUT, int G>
__global__ __launch_bounds__(CT) void conv_bwd_data_kernel(const float* __restrict__ dout, const uint16_t* __restrict__ wbf,
                                                           const float* __restrict__ act, float* __restrict__ din,
                                                           int H, int W) {
  static_assert(CIN % G == 0, "group size must divide the input channels");
  __shared__ __align__(16) float wl[COUT * 9 * G];   // [co][tap][ci in group]
  const int ci0 = blockIdx.z * G;
  for (int q = threadIdx.x; q < COUT * 9 * G; q += CT) {
    const int ci = q % G, r = q / G, tap = r % 9, co = r / 9;
    wl[q] = bf2f(wbf[(co * CIN + ci0 + ci) * 9 + tap]);
  }
  __syncthreads();
  const int HW = H * W;
  const int b = blockIdx.y;
  const int p = blockIdx.x * CT + threadIdx.x;
  if (p >= HW) return;
  const int y = p / W, x = p % W;
  const float* gb = dout + (size_t)b * COUT * HW;
  float acc[G];
#pragma unroll
  for (int ci = 0; ci < G; ci++) acc[ci] = 0.0f;
#pragma unroll 1
  for (int co = 0; co < COUT; co++) {
    const float* gp = gb + (size_t)co * HW;
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
      const int yy = y - (tap / 3 - 1), xx = x - (tap % 3 - 1);
      const float gv = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? gp[yy * W + xx] : 0.0f;
      const float* wr = wl + (co * 9 + tap) * G;
#pragma unroll
      for (int ci = 0; ci < G; ci++) acc[ci] += wr[ci] * gv;
    }
  }
  float* db = din + ((size_t)b * CIN + ci0) * HW + p;
  const float* ab = act ? act + ((size_t)b * CIN + ci0) * HW + p : nullptr;
#pragma unroll
  for (int ci = 0; ci < G; ci++) {
    float g = acc[ci];
    if (ab && !(ab[(size_t)ci * HW] > 0.0f)) g = 0.0f;
    db[(size_t)ci * HW] = g;
  }
}

// Weight gradient in two deterministic passes.  Pass 1: one block per (ci, co,
// chunk of the batch's pixels); ci == CIN gives the bias gradient; it writes
// its nine partial sums.  Pass 2 adds the chunks in order.
template <int CIN, int COUT>
__global__ __launch_bounds__(CT) void conv_bwd_weight_kernel(const float* __restrict__ dout, const float* __restrict__ in,
                                                             const int* __restrict__ idx, float* __restrict__ part,
                                                             int B, int H, int W, int chunk) {
  __shared__ float red[9][CT / 64];
  const int ci = blockIdx.x, co = blockIdx.y, ch = blockIdx.z, tid = threadIdx.x;
  const int HW = H * W;
  const long total = (long)B * HW;
  const long p0 = (long)ch * chunk, p1 = p0 + chunk < total ? p0 + chunk : total;
  float s[9];
#pragma unroll
  for (int k = 0; k < 9; k++) s[k] = 0.0f;
  for (long pg = p0 + tid; pg < p1; pg += CT) {
    const int b = (int)(pg / HW), p = (int)(pg % HW);
    const float g = dout[((size_t)b * COUT + co) * HW + p];
    if (ci == CIN) {
      s[0] += g;
    } else {
      const float* ip = in + ((size_t)(idx ? idx[b] : b) * CIN + ci) * HW;
      const int y = p / W, x = p % W;
#pragma unroll
      for (int tap = 0; tap < 9; tap++) {
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) s[tap] += g * ip[yy * W + xx];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 9; k++) {
    float v = s[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((tid & 63) == 0) red[k][tid >> 6] = v;
  }
  __syncthreads();
  if (tid < 9) {
    float v = 0.0f;
    for (int w = 0; w < CT / 64; w++) v += red[tid][w];
    part[(((size_t)ch * COUT + co) * (CIN + 1) + ci) * 9 + tid] = v;
  }
}

template <int CIN, int COUT>
__global__ void conv_bwd_weight_reduce_kernel(const float* __restrict__ part, int nchunk, float* __restrict__ dW,
                                              float* __restrict__ dbias) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;     // over COUT x (CIN+1) x 9
  if (q >= COUT * (CIN + 1) * 9) return;
  float v = 0.0f;
  for (int c = 0; c < nchunk; c++) v += part[(size_t)c * COUT * (CIN + 1) * 9 + q];
  const int tap = q % 9, r = q / 9, ci = r % (CIN + 1), co = r / (CIN + 1);
  if (ci == CIN) { if (tap == 0) dbias[co] = v; }
  else dW[(co * CIN + ci) * 9 + tap] = v;
}

__global__ __launch_bounds__(256) void corr_loss_kernel(const float* __restrict__ out, const float* __restrict__ Y,
                                                        const int* __restrict__ idx, float inv_ys, float* dout,
                                                        double* part, int B, int HW) {
  __shared__ double red[4];
  const size_t n = (size_t)B * 2 * HW;
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  double s = 0;
  if (q < n) {
    const int b = (int)(q / (2 * (size_t)HW));
    const size_t r = q % (2 * (size_t)HW);
    const float y = Y[(size_t)idx[b] * 2 * HW + r] * inv_ys;
    const float d = out[q] - y;
    dout[q] = d * (2.0f / (float)n);
    s = (double)d * (double)d;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void sum_squares_kernel(const float* __restrict__ x, int n, double* part) {
  __shared__ double red[4];
  double s = 0;
  for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) s += (double)x[q] * (double)x[q];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void adam_kernel(float* w, const float* __restrict__ g, float* m, float* v, uint16_t* wbf, int n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  const float wq = w[q] - lr * (mq * c1) / (sqrtf(vq * c2) + eps);
  w[q] = wq;
  wbf[q] = f2bf(wq);
}

__global__ void round_bf16_kernel(const float* __restrict__ w, uint16_t* wbf, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) wbf[q] = f2bf(w[q]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define FIL_CORR_LAYER(l_, ...)                                                  \
  switch (l_) {                                                                  \
    case 0: { constexpr int CI = CORR_IN, CO = CORR_CH; __VA_ARGS__; } break;    \
    case 1: { constexpr int CI = CORR_CH, CO = CORR_CH; __VA_ARGS__; } break;    \
    case 2: { constexpr int CI = CORR_CH, CO = CORR_CH; __VA_ARGS__; } break;    \
    case 3: { constexpr int CI = CORR_CH, CO = CORR_OUT; __VA_ARGS__; } break;   \
    default: return hipErrorInvalidValue;                                        \
  }
static_assert(CORR_LAYERS == 4, "layer dispatch covers four layers");

hipError_t launch_conv_weight_layout(const uint16_t* wbf, float* wfm, hipStream_t st) {
  constexpr int n = corr_wfm_size();
  hipLaunchKernelGGL(conv_weight_layout_kernel, dim3((n + 255) / 256), dim3(256), 0, st, wbf, wfm);
  return hipGetLastError();
}

hipError_t launch_conv_fwd_pooled(const float* Frow, int pr, const float* wfm, const float* bias, float* out, int H,
                                  int W, hipStream_t st) {
  static_assert(corr_cin(0) == CORR_IN && corr_cout(0) == CORR_CH, "layer 0");
  const int tiles = ((H + MY - 1) / MY) * ((W + MX - 1) / MX);
  hipLaunchKernelGGL((conv_fwd_kernel<CORR_IN, CORR_CH, true>), dim3(tiles, 1, (CORR_CH + 15) / 16), dim3(MT), 0, st,
                     Frow, nullptr, wfm, bias, out, H, W, 1, pr);
  return hipGetLastError();
}

hipError_t launch_conv_fwd(int l, const float* in, const int* idx, const float* wfm, const float* bias, float* out,
                           int B, int H, int W, int relu, hipStream_t st) {
  FIL_CORR_LAYER(l, {
    const int tiles = ((H + MY - 1) / MY) * ((W + MX - 1) / MX);
    hipLaunchKernelGGL((conv_fwd_kernel<CI, CO>), dim3(tiles, B, (CO + 15) / 16), dim3(MT), 0, st, in, idx, wfm, bias,
                       out, H, W, relu);
  });
  return hipGetLastError();
}

hipError_t launch_conv_bwd_data(int l, const float* dout, const uint16_t* wbf, const float* act, float* din, int B,
                                int H, int W, hipStream_t st) {
  FIL_CORR_LAYER(l, {
    constexpr int G = CI < CG ? CI : CG;
    const dim3 grid((H * W + CT - 1) / CT, B, CI / G);
    hipLaunchKernelGGL((conv_bwd_data_kernel<CI, CO, G>), grid, dim3(CT), 0, st, dout, wbf, act, din, H, W);
  });
  return hipGetLastError();
}

hipError_t launch_conv_bwd_weight(int l, const float* dout, const float* in, const int* idx, float* dW, float* db,
                                  float* part, int B, int H, int W, hipStream_t st) {
  const long total = (long)B * H * W;
  int nchunk = (int)((total + CORR_WCHUNK_PIX - 1) / CORR_WCHUNK_PIX);
  if (nchunk > CORR_WCHUNK_MAX) nchunk = CORR_WCHUNK_MAX;
  const int chunk = (int)((total + nchunk - 1) / nchunk);
  FIL_CORR_LAYER(l, {
    hipLaunchKernelGGL((conv_bwd_weight_kernel<CI, CO>), dim3(CI + 1, CO, nchunk), dim3(CT), 0, st, dout, in, idx,
                       part, B, H, W, chunk);
    const int n = CO * (CI + 1) * 9;
    hipLaunchKernelGGL((conv_bwd_weight_reduce_kernel<CI, CO>), dim3((n + 255) / 256), dim3(256), 0, st, part, nchunk,
                       dW, db);
  });
  return hipGetLastError();
}

hipError_t launch_corr_loss(const float* out, const float* Y, const int* idx, float inv_ys, float* dout, double* part,
                            int B, int HW, int* nblk, hipStream_t st) {
  const size_t n = (size_t)B * 2 * HW;
  *nblk = (int)((n + 255) / 256);
  hipLaunchKernelGGL(corr_loss_kernel, dim3(*nblk), dim3(256), 0, st, out, Y, idx, inv_ys, dout, part, B, HW);
  return hipGetLastError();
}

hipError_t launch_sum_squares(const float* x, int n, double* part, int* nblk, hipStream_t st) {
  *nblk = std::min(1024, (n + 255) / 256);
  if (*nblk < 1) *nblk = 1;
  hipLaunchKernelGGL(sum_squares_kernel, dim3(*nblk), dim3(256), 0, st, x, n, part);
  return hipGetLastError();
}

hipError_t launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(adam_kernel, dim3((n + 255) / 256), dim3(256), 0, st, w, g, m, v, wbf, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_round_bf16(const float* w, uint16_t* wbf, int n, hipStream_t st) {
  hipLaunchKernelGGL(round_bf16_kernel, dim3((n + 255) / 256), dim3(256), 0, st, w, wbf, n);
  return hipGetLastError();
}

}  // namespace fil
