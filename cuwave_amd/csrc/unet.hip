// The U-Net preconditioner's training and inference kernels for gfx950:
// 3x3 convolutions (stride 1, zero padding 1) as implicit GEMMs on the
// bfloat16 matrix cores with float32 accumulation, and the AdamW step. No
// reference counterpart: the reference has no learned preconditioner.
//
// Activations are NHWC bfloat16 (channels innermost, so the reduction axis
// of a GEMM tile is contiguous); a convolution's weights are (cout, 9 cin),
// row k = tap cin + ci with tap = 3 (dh + 1) + (dw + 1), in float32 for the
// optimiser and in bfloat16 for the kernels.
//
// - forward: y[p][co] = act(bias[co] + sum_{tap, ci} x[p + off(tap)][ci] w[co][tap][ci]),
//   a GEMM of M = pixels, N = cout, K = 9 cin; the x tile is gathered per
//   tap (zero off the image) and staged in LDS with the weight tile.
// - backward data: the same kernel on dy with the weights flipped in space
//   and transposed in channels (wd[ci][tap][co] = w[co][8 - tap][ci]).
// - backward weight: dw[co][tap][ci] = sum_p dy[p][co] x[p + off(tap)][ci],
//   a GEMM of M = 9 cin, N = cout reduced over the pixels: each block sums a
//   contiguous run of pixels into a float32 partial, and a second kernel
//   adds the partials in a fixed order (deterministic, no atomics).
//
// MFMA v_mfma_f32_32x32x16_bf16: lane l holds A[l % 32][8 (l / 32) + i] and
// B[8 (l / 32) + i][l % 32], i = 0 .. 7, and D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// in accumulator register r.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>

#include "../../include/cuwave_amd.h"

namespace cwh {
int fail(const std::string &what);
} // namespace cwh

namespace {

using cwh::fail;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kBM = 128; // pixels a forward block takes (4 waves x 32)
constexpr int kBK = 16;  // reduction step: 16 channels of one tap

__device__ __forceinline__ uint16_t to_bf16(float v) {
  const uint32_t bits = __float_as_uint(v);
  if ((bits & 0x7f800000u) == 0x7f800000u) // inf and nan keep their class
    return (uint16_t)((bits >> 16) | ((bits & 0xffffu) ? 0x40u : 0u));
  return (uint16_t)((bits + 0x7fffu + ((bits >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t b) {
  return __uint_as_float((uint32_t)b << 16);
}

struct Image {
  int n, h, w; // batch, height, width
  int64_t pixels;
};

// the pixel p + off(tap) of pixel p, -1 off the image
__device__ __forceinline__ int64_t shifted(const Image &im, int64_t p, int tap) {
  const int64_t hw = (int64_t)im.h * im.w;
  const int b = (int)(p / hw);
  const int rest = (int)(p - (int64_t)b * hw);
  const int y = rest / im.w + tap / 3 - 1;
  const int x = rest % im.w + tap % 3 - 1;
  if (y < 0 || y >= im.h || x < 0 || x >= im.w)
    return -1;
  return (int64_t)b * hw + (int64_t)y * im.w + x;
}

// ------------------------------------------------------------ forward
// one block: 128 pixels x BN output channels (BN = cout, 32 / 64 / 128);
// wave w takes pixels 32 w .. 32 w + 31 and every channel
template <int BN>
__global__ void __launch_bounds__(kThreads)
conv3x3_fwd(Image im, int cin, const uint4 *__restrict__ x, const uint4 *__restrict__ wt,
            const float *__restrict__ bias, int relu, uint16_t *__restrict__ y) {
  __shared__ uint4 sa[2][kBM * 2];  // [m][k]: 16 bf16 a row, two uint4
  __shared__ uint4 sb[2][BN * 2];   // [n][k]
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int64_t m0 = (int64_t)blockIdx.x * kBM;
  const int steps = 9 * (cin / kBK);
  const int cin8 = cin / 8;      // uint4 chunks a pixel
  const int krow8 = 9 * cin8;    // uint4 chunks a weight row
  // the chunk this thread stages: pixel am, half ah; weight row bn, half bh
  const int am = tid / 2, ah = tid % 2;
  const int bn = tid / 2, bh = tid % 2;
  uint4 ra = {0, 0, 0, 0}, rb = {0, 0, 0, 0};
  auto load = [&](int s) {
    const int tap = s / (cin / kBK), c0 = (s % (cin / kBK)) * kBK;
    const int64_t p = m0 + am;
    ra = uint4{0, 0, 0, 0};
    if (p < im.pixels) {
      const int64_t q = shifted(im, p, tap);
      if (q >= 0)
        ra = x[q * cin8 + c0 / 8 + ah];
    }
    if (bn < BN)
      rb = wt[(int64_t)bn * krow8 + tap * cin8 + c0 / 8 + bh];
  };
  f32x16 acc[BN / 32];
#pragma unroll
  for (int j = 0; j < BN / 32; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      acc[j][r] = 0.0f;
  load(0);
  for (int s = 0; s < steps; ++s) {
    const int buf = s & 1;
    sa[buf][am * 2 + ah] = ra;
    if (bn < BN)
      sb[buf][bn * 2 + bh] = rb;
    __syncthreads();
    if (s + 1 < steps)
      load(s + 1);
    const uint4 a = sa[buf][(wave * 32 + lane % 32) * 2 + lane / 32];
    const bf16x8 av = __builtin_bit_cast(bf16x8, a);
#pragma unroll
    for (int j = 0; j < BN / 32; ++j) {
      const uint4 b = sb[buf][(j * 32 + lane % 32) * 2 + lane / 32];
      acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, b),
                                                      acc[j], 0, 0, 0);
    }
  }
  // epilogue: bias, activation, bfloat16
#pragma unroll
  for (int j = 0; j < BN / 32; ++j) {
    const int n = j * 32 + lane % 32;
    const float b = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = wave * 32 + 8 * (r / 4) + 4 * (lane / 32) + r % 4;
      const int64_t p = m0 + m;
      if (p >= im.pixels)
        continue;
      float v = acc[j][r] + b;
      if (relu)
        v = v > 0.0f ? v : 0.0f;
      y[p * BN + n] = to_bf16(v);
    }
  }
}

// ------------------------------------------------------- backward weight
// one block: 128 rows of the (tap, ci) axis x BN output channels, summed
// over its run of pixels 32 at a time; the partial goes to part[split]
template <int BN>
__global__ void __launch_bounds__(kThreads)
conv3x3_wgrad(Image im, int cin, const uint4 *__restrict__ x, const uint4 *__restrict__ dy,
              int64_t run, float *__restrict__ part) {
  constexpr int P = 32; // pixels a step
  __shared__ __attribute__((aligned(16))) uint16_t sx[2][128][P + 8];
  __shared__ __attribute__((aligned(16))) uint16_t sd[2][BN][P + 8];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int krows = 9 * cin;
  const int r0 = blockIdx.x * 128;
  const int split = blockIdx.y;
  const int64_t p_begin = (int64_t)split * run;
  const int64_t p_end = p_begin + run < im.pixels ? p_begin + run : im.pixels;
  const int cin8 = cin / 8;
  constexpr int cout8 = BN / 8;
  constexpr int DPT = (P * cout8 + kThreads - 1) / kThreads; // dy chunks a thread
  // staging: x chunk (pixel xp, rows xr .. xr + 7) twice; dy chunks (pixel dp,
  // channels dc .. dc + 7)
  uint4 rx[2], rd[DPT];
  auto load = [&](int64_t q0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = tid + h * kThreads;       // 512 chunks: 32 pixels x 16
      const int xp = t / 16, xr = r0 + (t % 16) * 8;
      rx[h] = uint4{0, 0, 0, 0};
      const int64_t p = q0 + xp;
      if (p < p_end && xr < krows) {
        const int tap = xr / cin, ci = xr % cin;
        const int64_t q = shifted(im, p, tap);
        if (q >= 0)
          rx[h] = x[q * cin8 + ci / 8];
      }
    }
#pragma unroll
    for (int h = 0; h < DPT; ++h) {
      const int t = tid + h * kThreads;
      const int dp = t / cout8, dc = (t % cout8) * 8;
      rd[h] = uint4{0, 0, 0, 0};
      const int64_t p = q0 + dp;
      if (dp < P && p < p_end)
        rd[h] = dy[p * cout8 + dc / 8];
    }
  };
  f32x16 acc[BN / 32];
#pragma unroll
  for (int j = 0; j < BN / 32; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r)
      acc[j][r] = 0.0f;
  const int64_t steps = (p_end - p_begin + P - 1) / P;
  if (steps > 0)
    load(p_begin);
  for (int64_t s = 0; s < steps; ++s) {
    const int buf = (int)(s & 1);
    // transposed into LDS: [row][pixel], [channel][pixel]
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = tid + h * kThreads;
      const int xp = t / 16, xr = (t % 16) * 8;
      const uint16_t *v = reinterpret_cast<const uint16_t *>(&rx[h]);
#pragma unroll
      for (int i = 0; i < 8; ++i)
        sx[buf][xr + i][xp] = v[i];
    }
#pragma unroll
    for (int h = 0; h < DPT; ++h) {
      const int t = tid + h * kThreads;
      const int dp = t / cout8, dc = (t % cout8) * 8;
      if (dp < P) {
        const uint16_t *v = reinterpret_cast<const uint16_t *>(&rd[h]);
#pragma unroll
        for (int i = 0; i < 8; ++i)
          sd[buf][dc + i][dp] = v[i];
      }
    }
    __syncthreads();
    if (s + 1 < steps)
      load(p_begin + (s + 1) * P);
    // rows of P + 8 bfloat16 (80 bytes), so a lane's eight pixels are one
    // aligned 16-byte LDS read
#pragma unroll
    for (int k = 0; k < P; k += 16) {
      const bf16x8 av = __builtin_bit_cast(
          bf16x8, *reinterpret_cast<const uint4 *>(&sx[buf][wave * 32 + lane % 32][k + 8 * (lane / 32)]));
#pragma unroll
      for (int j = 0; j < BN / 32; ++j) {
        const bf16x8 bv = __builtin_bit_cast(
            bf16x8, *reinterpret_cast<const uint4 *>(&sd[buf][j * 32 + lane % 32][k + 8 * (lane / 32)]));
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[j], 0, 0, 0);
      }
    }
  }
  // the partial: part[split][row][co]
  float *out = part + (int64_t)split * krows * BN;
#pragma unroll
  for (int j = 0; j < BN / 32; ++j) {
    const int n = j * 32 + lane % 32;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = r0 + wave * 32 + 8 * (r / 4) + 4 * (lane / 32) + r % 4;
      if (row < krows)
        out[(int64_t)row * BN + n] = acc[j][r];
    }
  }
}

// dw[co][row] = sum over the splits in order of part[split][row][co]
__global__ void __launch_bounds__(kThreads)
conv3x3_wgrad_sum(int krows, int cout, int splits, const float *__restrict__ part,
                  float *__restrict__ dw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)krows * cout)
    return;
  const int row = (int)(t / cout), co = (int)(t % cout);
  float s = 0.0f;
  for (int k = 0; k < splits; ++k)
    s += part[((int64_t)k * krows + row) * cout + co];
  dw[(int64_t)co * krows + row] = s;
}

// the kernels' weights from the float32 master: wt (cout, 9 cin) as is and
// wd (cin, 9 cout) flipped in space and transposed in channels
__global__ void __launch_bounds__(kThreads)
conv3x3_weights(int cout, int cin, const float *__restrict__ w, uint16_t *__restrict__ wt,
                uint16_t *__restrict__ wd) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)cout * 9 * cin)
    return;
  const int co = (int)(t / (9 * cin)), k = (int)(t % (9 * cin));
  const int tap = k / cin, ci = k % cin;
  const uint16_t b = to_bf16(w[t]);
  wt[t] = b;
  if (wd)
    wd[(int64_t)ci * 9 * cout + (8 - tap) * cout + co] = b;
}

// AdamW (decoupled weight decay, bias-corrected moments)
__global__ void __launch_bounds__(kThreads)
adamw(int64_t count, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
      float *__restrict__ v, float lr, float beta1, float beta2, float eps, float decay,
      float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const float gi = g[i];
  float pi = p[i] * (1.0f - lr * decay);
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mhat = mi / c1, vhat = vi / c2;
  pi -= lr * mhat / (sqrtf(vhat) + eps);
  p[i] = pi;
}

int check_conv(const cw_conv3x3_args *a) {
  if (!a)
    return fail("args is NULL");
  if (a->n < 1 || a->h < 1 || a->w < 1)
    return fail("the image needs a positive batch, height and width");
  if (a->cin < 16 || a->cin % 16)
    return fail("cin must be a positive multiple of 16");
  if (a->cout != 32 && a->cout != 64 && a->cout != 128)
    return fail("cout must be 32, 64 or 128");
  if ((int64_t)a->n * a->h * a->w * (a->cin > a->cout ? a->cin : a->cout) >= (int64_t)1 << 31)
    return fail("the image overflows the int32 element index of a row");
  return 0;
}

int hip_status(const char *where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    fail(std::string(where) + ": " + hipGetErrorString(e));
    return 2;
  }
  return 0;
}

} // namespace

extern "C" {

int cw_conv3x3_bf16_fwd(const cw_conv3x3_args *a, void *stream) {
  if (int e = check_conv(a))
    return e;
  if (!a->x || !a->weight || !a->y)
    return fail("the forward needs x, weight and y");
  const Image im{a->n, a->h, a->w, (int64_t)a->n * a->h * a->w};
  const int blocks = (int)((im.pixels + kBM - 1) / kBM);
  const auto s = (hipStream_t)stream;
  const uint4 *x = (const uint4 *)a->x, *w = (const uint4 *)a->weight;
  const float *bias = (const float *)a->bias;
  uint16_t *y = (uint16_t *)a->y;
  switch (a->cout) {
  case 32: hipLaunchKernelGGL(conv3x3_fwd<32>, dim3(blocks), dim3(kThreads), 0, s, im, a->cin, x, w, bias, a->relu, y); break;
  case 64: hipLaunchKernelGGL(conv3x3_fwd<64>, dim3(blocks), dim3(kThreads), 0, s, im, a->cin, x, w, bias, a->relu, y); break;
  default: hipLaunchKernelGGL(conv3x3_fwd<128>, dim3(blocks), dim3(kThreads), 0, s, im, a->cin, x, w, bias, a->relu, y); break;
  }
  return hip_status("conv3x3_fwd");
}

int cw_conv3x3_bf16_bwd_data(const cw_conv3x3_args *a, void *stream) {
  // dy (cout channels) through the flipped, transposed weights into dx (cin)
  if (int e = check_conv(a))
    return e;
  if (!a->dy || !a->weight_flipped || !a->dx)
    return fail("the backward data pass needs dy, the flipped weights and dx");
  if (a->cin != 32 && a->cin != 64 && a->cin != 128)
    return fail("the backward data pass writes 32, 64 or 128 channels");
  cw_conv3x3_args b = *a;
  b.cin = a->cout;
  b.cout = a->cin;
  b.x = a->dy;
  b.weight = a->weight_flipped;
  b.bias = nullptr;
  b.relu = 0;
  b.y = a->dx;
  return cw_conv3x3_bf16_fwd(&b, stream);
}

int cw_conv3x3_bf16_bwd_weight(const cw_conv3x3_args *a, void *stream) {
  if (int e = check_conv(a))
    return e;
  if (!a->x || !a->dy || !a->dweight || !a->workspace || a->splits < 1)
    return fail("the backward weight pass needs x, dy, dweight, a workspace and splits");
  const Image im{a->n, a->h, a->w, (int64_t)a->n * a->h * a->w};
  const int krows = 9 * a->cin;
  const int64_t run = (im.pixels + a->splits - 1) / a->splits;
  const dim3 grid((krows + 127) / 128, a->splits);
  const auto s = (hipStream_t)stream;
  const uint4 *x = (const uint4 *)a->x, *dy = (const uint4 *)a->dy;
  float *part = (float *)a->workspace;
  switch (a->cout) {
  case 32: hipLaunchKernelGGL(conv3x3_wgrad<32>, grid, dim3(kThreads), 0, s, im, a->cin, x, dy, run, part); break;
  case 64: hipLaunchKernelGGL(conv3x3_wgrad<64>, grid, dim3(kThreads), 0, s, im, a->cin, x, dy, run, part); break;
  default: hipLaunchKernelGGL(conv3x3_wgrad<128>, grid, dim3(kThreads), 0, s, im, a->cin, x, dy, run, part); break;
  }
  if (int e = hip_status("conv3x3_wgrad"))
    return e;
  const int64_t total = (int64_t)krows * a->cout;
  hipLaunchKernelGGL(conv3x3_wgrad_sum, dim3((int)((total + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, krows, a->cout, a->splits, (const float *)part,
                     (float *)a->dweight);
  return hip_status("conv3x3_wgrad_sum");
}

int cw_conv3x3_weights(int32_t cout, int32_t cin, const void *w, void *wt, void *wd,
                       void *stream) {
  if (cout < 1 || cin < 1 || !w || !wt)
    return fail("the weight transform needs cout, cin, w and wt");
  const int64_t total = (int64_t)cout * 9 * cin;
  hipLaunchKernelGGL(conv3x3_weights, dim3((int)((total + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, (hipStream_t)stream, cout, cin, (const float *)w,
                     (uint16_t *)wt, (uint16_t *)wd);
  return hip_status("conv3x3_weights");
}

int cw_adamw(const cw_adamw_args *a, void *stream) {
  if (!a || a->count < 0 || (a->count && (!a->param || !a->grad || !a->m || !a->v)))
    return fail("AdamW needs param, grad, m and v");
  if (a->step < 1)
    return fail("AdamW steps count from 1");
  if (!a->count)
    return 0;
  const float c1 = 1.0f - powf(a->beta1, (float)a->step);
  const float c2 = 1.0f - powf(a->beta2, (float)a->step);
  hipLaunchKernelGGL(adamw, dim3((int)((a->count + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, (hipStream_t)stream, a->count, (float *)a->param, (const float *)a->grad,
                     (float *)a->m, (float *)a->v, a->lr, a->beta1, a->beta2, a->eps,
                     a->weight_decay, c1, c2);
  return hip_status("adamw");
}

} // extern "C"
