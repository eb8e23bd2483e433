// fil_corr.hip -- network of the learned coarse-step corrector (no counterpart
// in the reference): forward and backward of 3x3 convolutions, MSE loss and
// Adam, all on the device.
//
// Layer l maps [B][cin][H][W] -> [B][cout][H][W] with zero padding:
//   out[co][y][x] = b[co] + sum_ci sum_{dy,dx in -1..1} w[co][ci][(dy+1)*3+(dx+1)] in[ci][y+dy][x+dx]
// Weights are read as BF16 (the rounded copy of the FP32 master parameters)
// and every product is accumulated in FP32.  The forward pass runs the four
// layers in one launch on the BF16 matrix cores (exact products, below); the
// backward passes own one pixel (data gradient) or one weight (weight
// gradient) per thread on the vector ALUs.
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

constexpr int CT = 128;   // threads per block of the backward convolution kernels
constexpr int CG = 8;     // channels per thread in the data-gradient kernel

// ---------------------------------------------------------------------------
// Forward of the whole network in one launch, on v_mfma_f32_16x16x32_bf16.
//
// A workgroup owns a TY x TX tile of output pixels of one sample and carries
// it through the four layers in LDS: layer l (1..4) is computed on the tile
// grown by 4 - l pixels on every side (the region the later layers read), so
// no layer waits for another workgroup; pixels of a region outside the image
// are zero, which is the zero padding every layer sees.  Regions alternate
// between two LDS buffers (region 0 = input, 2 in buffer 0; 1, 3 in buffer 1).
//
// Exact products on BF16 matrix cores: every FP32 activation x is stored as
// three BF16 values x = hi + mid + lo (hi = bf16(x), mid = bf16(x - hi),
// lo = x - hi - mid, exact: the three take the 24 significand bits eight at a
// time, and both differences are exact by Sterbenz's lemma), and the weights
// are BF16 values, so each of the three MFMA terms is an exact product summed
// in FP32: the layer is an FP32 convolution up to the order of the sums.
//
// Layer l as a GEMM: out[co][p] = b[co] + sum_k A[co][k] X[k][p], k = tap cin + ci;
// one MFMA covers 16 output channels x 16 pixels x 32 k: for cin = 32 one
// k chunk is one tap (all 32 channels); for the 4-channel input two chunks
// cover the 36 k (taps 2 kg, 2 kg + 1 of k group kg).  Each wave takes 16-pixel
// tiles of the region in turn and keeps the layer's weights in registers.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int NET_T = 512;    // threads per workgroup (8 waves)
// bytes per pixel of a 32-channel region: planes hi, mid, lo of 64 bytes and
// 32 bytes of padding; 224 = 14 x 16 keeps the operand reads (16 pixels x 4
// 16-byte channel groups per instruction) free of bank conflicts
constexpr int NET_PS = 224;
constexpr int NET_PS0 = 32;   // input region: planes of 4 channels (8 bytes each) and 8 bytes of padding

__host__ __device__ constexpr int net_ry(int TY, int r) { return TY + 2 * (CORR_LAYERS - r); }
__host__ __device__ constexpr int net_buf0(int TY, int TX) {
  const int r0 = net_ry(TY, 0) * net_ry(TX, 0) * NET_PS0, r2 = net_ry(TY, 2) * net_ry(TX, 2) * NET_PS;
  return r0 > r2 ? r0 : r2;
}
__host__ __device__ constexpr int net_buf1(int TY, int TX) {
  const int r1 = net_ry(TY, 1) * net_ry(TX, 1) * NET_PS, r3 = net_ry(TY, 3) * net_ry(TX, 3) * NET_PS;
  return r1 > r3 ? r1 : r3;
}

// the network's weights (fragment layout) and biases, staged once per workgroup
constexpr int NET_WBYTES = corr_frag_size() * 2;
__host__ __device__ constexpr int net_bias_off(int l) {
  int o = 0;
  for (int k = 0; k < l; k++) o += corr_cout(k);
  return o;
}
constexpr int NET_BBYTES = (net_bias_off(CORR_LAYERS) * 4 + 15) / 16 * 16;
__host__ __device__ constexpr int net_smem(int TY, int TX) {
  return NET_WBYTES + NET_BBYTES + net_buf0(TY, TX) + net_buf1(TY, TX);
}

__device__ __forceinline__ void split3(float x, uint16_t* s) {
  s[0] = f2bf(x);
  const float r1 = x - bf2f(s[0]);
  s[1] = f2bf(r1);
  s[2] = f2bf(r1 - bf2f(s[1]));
}
__device__ __forceinline__ uint2 pack4(const uint16_t* v) {
  return make_uint2((uint32_t)v[0] | ((uint32_t)v[1] << 16), (uint32_t)v[2] | ((uint32_t)v[3] << 16));
}

struct NetArgs {
  const float* in;
  const int* idx;
  int pr;
  const uint16_t* wfr;
  const float* params;
  float* act[CORR_LAYERS - 1];
  float* out;
  int H, W;
};

// layer L (1..4): region L-1 in `src` -> region L in `dst` (L < 4) or the
// output (L = 4); (y0, x0): the image position of the output tile
template <int L, int TY, int TX>
__device__ __forceinline__ void net_layer(const unsigned char* src, unsigned char* dst, const bf16x8* wl,
                                          const float* bl, const NetArgs& a, int b, int y0, int x0) {
  constexpr int l = L - 1, COUT = corr_cout(l);
  constexpr int NG = corr_frag_ng(l), NC = corr_frag_nc(l);
  constexpr int RY = net_ry(TY, L), RX = net_ry(TX, L), RXI = RX + 2, G = CORR_LAYERS - L;
  constexpr int NPX = RY * RX, NTILE = (NPX + 15) / 16;
  constexpr int PSI = (L == 1) ? NET_PS0 : NET_PS;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, n = lane & 15, kg = lane >> 4;
  const int H = a.H, W = a.W;
  bf16x8 A[NG][NC];
  {
    const bf16x8* wf = wl + corr_frag_off(l) * 64 + lane;
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
      for (int c = 0; c < NC; c++) A[g][c] = wf[(g * NC + c) * 64];
    // held in registers for the whole layer (the compiler would otherwise
    // re-read them from LDS next to every MFMA and wait on each read)
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
      for (int c = 0; c < NC; c++) asm volatile("" : "+v"(A[g][c]));
  }
  float bias[NG][4];
#pragma unroll
  for (int g = 0; g < NG; g++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = 16 * g + 4 * kg + i;
      bias[g][i] = co < COUT ? bl[net_bias_off(l) + co] : 0.0f;
    }
  for (int t = wv; t < NTILE; t += NET_T / 64) {
    const int p = 16 * t + n;
    const int pc = p < NPX ? p : NPX - 1;       // lanes past the region read a valid pixel, store nothing
    const int pin = (pc / RX) * RXI + pc % RX;   // input pixel of tap 0 (dy = dx = -1)
    floatx4 acc[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) acc[g] = {0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NC; c++) {
      bf16x8 B[3];
      if constexpr (L == 1) {
        // k = 32 c + 8 kg + j: taps ta = 8 c + 2 kg (j < 4) and ta + 1 (j >= 4), channel j % 4
        const int ta = 8 * c + 2 * kg, tb = ta + 1;
        const unsigned char* pa = src + (pin + (ta / 3) * RXI + ta % 3) * PSI;
        const unsigned char* pb = src + (pin + (tb / 3) * RXI + tb % 3) * PSI;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const uint2 va = ta < 9 ? *reinterpret_cast<const uint2*>(pa + 8 * q) : make_uint2(0, 0);
          const uint2 vb = tb < 9 ? *reinterpret_cast<const uint2*>(pb + 8 * q) : make_uint2(0, 0);
          const u32x4 v = {va.x, va.y, vb.x, vb.y};
          B[q] = __builtin_bit_cast(bf16x8, v);
        }
      } else {
        const unsigned char* pp = src + (pin + (c / 3) * RXI + c % 3) * PSI + 16 * kg;
#pragma unroll
        for (int q = 0; q < 3; q++) B[q] = *reinterpret_cast<const bf16x8*>(pp + 64 * q);
      }
#pragma unroll
      for (int q = 0; q < 3; q++)
#pragma unroll
        for (int g = 0; g < NG; g++) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[g][c], B[q], acc[g], 0, 0, 0);
    }
    if (p >= NPX) continue;
    const int ry = p / RX, rx = p % RX, yy = y0 - G + ry, xx = x0 - G + rx;
    const bool inside = yy >= 0 && yy < H && xx >= 0 && xx < W;
    if constexpr (L == CORR_LAYERS) {
      if (inside && kg == 0) {
        float* o = a.out + (size_t)b * COUT * H * W + (size_t)yy * W + xx;
#pragma unroll
        for (int i = 0; i < COUT; i++) o[(size_t)i * H * W] = acc[0][i] + bias[0][i];
      }
    } else {
      // the tile's own pixels of this layer go to act[l] (one workgroup owns each)
      const bool own = a.act[l] && inside && ry >= G && ry < G + TY && rx >= G && rx < G + TX;
      float* ap = own ? a.act[l] + ((size_t)b * COUT + 4 * kg) * H * W + (size_t)yy * W + xx : nullptr;
#pragma unroll
      for (int g = 0; g < NG; g++) {
        uint16_t s[4][3];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const float v = inside ? fmaxf(acc[g][i] + bias[g][i], 0.0f) : 0.0f;
          split3(v, s[i]);
          if (own) ap[(size_t)(16 * g + i) * H * W] = v;
        }
        unsigned char* dp = dst + p * NET_PS + 2 * (16 * g + 4 * kg);
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const uint16_t v4[4] = {s[0][q], s[1][q], s[2][q], s[3][q]};
          *reinterpret_cast<uint2*>(dp + 64 * q) = pack4(v4);
        }
      }
    }
  }
}

// LDS admits one workgroup per CU (two waves per SIMD), so the register
// budget is that of two waves
template <int TY, int TX>
__global__ __launch_bounds__(NET_T) __attribute__((amdgpu_waves_per_eu(1, 2))) void conv_net_kernel(NetArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  bf16x8* wl = reinterpret_cast<bf16x8*>(smem);
  float* bl = reinterpret_cast<float*>(smem + NET_WBYTES);
  unsigned char* buf0 = smem + NET_WBYTES + NET_BBYTES;
  unsigned char* buf1 = buf0 + net_buf0(TY, TX);
  const int tid = threadIdx.x, b = blockIdx.y;
  const int H = a.H, W = a.W, HW = H * W;
  const int ntx = (W + TX - 1) / TX;
  const int y0 = (blockIdx.x / ntx) * TY, x0 = (blockIdx.x % ntx) * TX;
  constexpr int RY0 = net_ry(TY, 0), RX0 = net_ry(TX, 0), NE = (RY0 * RX0 + NET_T - 1) / NET_T;
  constexpr int NW16 = NET_WBYTES / 16, NWL = (NW16 + NET_T - 1) / NET_T;
  // every global load of the staging is issued before the first LDS store:
  // the weights and biases of all four layers, and region 0 (the four input
  // channels of the tile and its 4-pixel halo; Frow rows pooled on the way)
  uint4 wreg[NWL];
  const uint4* wsrc = reinterpret_cast<const uint4*>(a.wfr);
#pragma unroll
  for (int j = 0; j < NWL; j++) {
    const int q = tid + j * NET_T;
    wreg[j] = wsrc[q < NW16 ? q : NW16 - 1];   // clamped: the extra copies are not stored
  }
  float bv = 0.0f;
  if (tid < net_bias_off(CORR_LAYERS)) {
    int l = 0;
    while (l + 1 < CORR_LAYERS && tid >= net_bias_off(l + 1)) l++;
    bv = a.params[corr_woff(l) + corr_cout(l) * corr_cin(l) * 9 + tid - net_bias_off(l)];
  }
  const float* inb = a.in + (size_t)(a.idx ? a.idx[b] : b) * CORR_IN * HW;
  float v[NE][CORR_IN];
#pragma unroll
  for (int u = 0; u < NE; u++) {
    const int e = tid + u * NET_T;
    const int yy = y0 - CORR_LAYERS + e / RX0, xx = x0 - CORR_LAYERS + e % RX0;
#pragma unroll
    for (int ci = 0; ci < CORR_IN; ci++) v[u][ci] = 0.0f;
    if (e < RY0 * RX0 && yy >= 0 && yy < H && xx >= 0 && xx < W) {
      if (a.pr > 0) {
        // mean over the pr feature rows of pooled row yy (Frow [H pr][4][W]),
        // eight rows' loads in flight at a time
        const int pr = a.pr;
        const float* s = a.in + (size_t)yy * pr * CORR_IN * W + xx;
        for (int k0 = 0; k0 < pr; k0 += 8) {
          float t[8][CORR_IN];
#pragma unroll
          for (int k = 0; k < 8; k++)
#pragma unroll
            for (int ci = 0; ci < CORR_IN; ci++)
              t[k][ci] = (k0 + k < pr) ? s[((size_t)(k0 + k) * CORR_IN + ci) * W] : 0.0f;
#pragma unroll
          for (int k = 0; k < 8; k++)
#pragma unroll
            for (int ci = 0; ci < CORR_IN; ci++) v[u][ci] += t[k][ci];
        }
#pragma unroll
        for (int ci = 0; ci < CORR_IN; ci++) v[u][ci] /= (float)pr;
      } else {
#pragma unroll
        for (int ci = 0; ci < CORR_IN; ci++) v[u][ci] = inb[(size_t)ci * HW + yy * W + xx];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NWL; j++) {
    const int q = tid + j * NET_T;
    if (q < NW16) reinterpret_cast<uint4*>(wl)[q] = wreg[j];
  }
  if (tid < net_bias_off(CORR_LAYERS)) bl[tid] = bv;
#pragma unroll
  for (int u = 0; u < NE; u++) {
    const int e = tid + u * NET_T;
    if (e < RY0 * RX0) {
      uint16_t s[CORR_IN][3];
#pragma unroll
      for (int ci = 0; ci < CORR_IN; ci++) split3(v[u][ci], s[ci]);
#pragma unroll
      for (int q = 0; q < 3; q++) {
        const uint16_t v4[4] = {s[0][q], s[1][q], s[2][q], s[3][q]};
        *reinterpret_cast<uint2*>(buf0 + e * NET_PS0 + 8 * q) = pack4(v4);
      }
    }
  }
  __syncthreads();
  net_layer<1, TY, TX>(buf0, buf1, wl, bl, a, b, y0, x0);
  __syncthreads();
  net_layer<2, TY, TX>(buf1, buf0, wl, bl, a, b, y0, x0);
  __syncthreads();
  net_layer<3, TY, TX>(buf0, buf1, wl, bl, a, b, y0, x0);
  __syncthreads();
  net_layer<4, TY, TX>(buf1, nullptr, wl, bl, a, b, y0, x0);
}

// wfr of every layer from the flat BF16 parameters (one thread per value)
__global__ void conv_weight_layout_kernel(const uint16_t* __restrict__ wbf, uint16_t* __restrict__ wfr) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= corr_frag_size()) return;
  const int f = q / 512, lane = (q / 8) % 64, j = q % 8;
  int l = 0;
  while (l + 1 < CORR_LAYERS && f >= corr_frag_off(l + 1)) l++;
  const int fl = f - corr_frag_off(l), nc = corr_frag_nc(l), cin = corr_cin(l);
  const int g = fl / nc, c = fl % nc;
  const int co = 16 * g + lane % 16, k = 32 * c + 8 * (lane / 16) + j;
  const int tap = k / cin, ci = k % cin;
  wfr[q] = (co < corr_cout(l) && tap < 9) ? wbf[corr_woff(l) + (co * cin + ci) * 9 + tap] : (uint16_t)0;
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

hipError_t launch_conv_weight_layout(const uint16_t* wbf, uint16_t* wfr, hipStream_t st) {
  constexpr int n = corr_frag_size();
  hipLaunchKernelGGL(conv_weight_layout_kernel, dim3((n + 255) / 256), dim3(256), 0, st, wbf, wfr);
  return hipGetLastError();
}

// Output tile of the fused forward: 2 x 16 pixels for one sample (256
// workgroups at the 64 x 128 pooled grid: the per-workgroup latency of the
// four layers, not the halo's extra work, bounds a single sample), 4 x 16 for
// training batches (half the halo work per pixel).
template <int TY>
static hipError_t launch_conv_net_t(const NetArgs& a, int B, hipStream_t st) {
  constexpr int TX = 16;
  constexpr int sm = net_smem(TY, TX);
  static_assert(sm <= 160 * 1024, "LDS of one workgroup");
  auto k = conv_net_kernel<TY, TX>;
  hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, sm);
  if (e != hipSuccess) return e;
  const int tiles = ((a.H + TY - 1) / TY) * ((a.W + TX - 1) / TX);
  hipLaunchKernelGGL(k, dim3(tiles, B), dim3(NET_T), sm, st, a);
  return hipGetLastError();
}

hipError_t launch_conv_net(const float* in, const int* idx, int pr, const uint16_t* wfr, const float* params,
                           float* const* act, float* out, int B, int H, int W, hipStream_t st) {
  if (B < 1 || H < 1 || W < 1 || (pr > 0 && (B != 1 || idx))) return hipErrorInvalidValue;
  NetArgs a{};
  a.in = in; a.idx = idx; a.pr = pr; a.wfr = wfr; a.params = params; a.out = out; a.H = H; a.W = W;
  for (int l = 0; l < CORR_LAYERS - 1; l++) a.act[l] = act ? act[l] : nullptr;
  return B == 1 ? launch_conv_net_t<2>(a, B, st) : launch_conv_net_t<4>(a, B, st);
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
