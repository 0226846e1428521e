// xg_mlp.hip -- the learned quadrature: a small MLP (I -> H -> H -> O, tanh)
// that maps the normalised P2 nodal level-set values of a cut tet to the
// fixed-node weights of its NEG part, plus the FP64 moment-residual check
// with fallback to the plain (subdivision) weights.
//
// No counterpart in the reference (it is the rebuild's learned rule); the
// plain rule it is trained on and falls back to is the subdivision rule of
// cutint/xintegration.cpp:1626-1754.
//
// Kernels
//  * training (FP32, VALU): dense layer forward, backward (input and weight
//    gradients, deterministic: one thread per weight reduces over the batch),
//    Adam.  Simple and exactly checkable against the FP64 oracle.
//  * inference (BF16 MFMA, FP32 accumulation), fused per 64-row tile:
//    layer 1 (K = 10) on VALU, written straight into the A-operand layout of
//    v_mfma_f32_16x16x32_bf16; layer 2 (128 x 128) on MFMA; tanh; the
//    16 x 128 tile goes through LDS once to re-layout it as the A operand of
//    layer 3 (128 x O) on MFMA.  W2 / W3 are pre-packed to BF16 in the
//    B-operand order (one 16-byte load per lane per MFMA).
#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace {

// ------------------------------------------------------------ FP32 layers
// Y[b][o] = act(sum_i X[b][i] W[o][i] + bias[o]); act: 0 identity, 1 tanh
__global__ void k_dense_fwd(int B, int I, int O, const float *__restrict__ X, const float *__restrict__ W,
                            const float *__restrict__ bias, int act, float *__restrict__ Y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)B * O) return;
  const int b = (int)(t / O), o = (int)(t - (long long)b * O);
  float a = bias[o];
  const float *x = X + (size_t)b * I;
  const float *w = W + (size_t)o * I;
  for (int i = 0; i < I; ++i) a += x[i] * w[i];
  Y[t] = act ? tanhf(a) : a;
}

// dZ = dY * (1 - Y^2) (tanh) in place
__global__ void k_tanh_bwd(long long n, const float *__restrict__ Y, float *__restrict__ dY) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dY[t] = dY[t] * (1.0f - Y[t] * Y[t]);
}

// dX[b][i] = sum_o dZ[b][o] W[o][i]
__global__ void k_dense_bwd_x(int B, int I, int O, const float *__restrict__ dZ, const float *__restrict__ W,
                              float *__restrict__ dX) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)B * I) return;
  const int b = (int)(t / I), i = (int)(t - (long long)b * I);
  float a = 0.0f;
  for (int o = 0; o < O; ++o) a += dZ[(size_t)b * O + o] * W[(size_t)o * I + i];
  dX[t] = a;
}

// dW[o][i] = sum_b dZ[b][o] X[b][i];  db[o] = sum_b dZ[b][o]   (i == I -> bias)
__global__ void k_dense_bwd_w(int B, int I, int O, const float *__restrict__ dZ, const float *__restrict__ X,
                              float *__restrict__ dW, float *__restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= O * (I + 1)) return;
  const int o = t / (I + 1), i = t - o * (I + 1);
  float a = 0.0f;
  if (i < I)
    for (int b = 0; b < B; ++b) a += dZ[(size_t)b * O + o] * X[(size_t)b * I + i];
  else
    for (int b = 0; b < B; ++b) a += dZ[(size_t)b * O + o];
  if (i < I) dW[(size_t)o * I + i] = a;
  else db[o] = a;
}

// dY = (Y - T) / B ; per-block partial loss 0.5 sum (Y - T)^2 / B
__global__ void k_mse(long long n, int B, const float *__restrict__ Y, const float *__restrict__ T,
                      float *__restrict__ dY, float *__restrict__ loss) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (t < n) {
    const float r = Y[t] - T[t];
    dY[t] = r / B;
    l = 0.5f * r * r / B;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

// ------------------------------------------------------------ BF16 inference
constexpr int MI = 10, MH = 128, MOP = 96;  // input, hidden, padded output

__device__ __forceinline__ short f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  const uint32_t r = ((u >> 16) & 1u) + 0x7fffu;
  u += r;
  return (short)(u >> 16);
}

// pack W2 (H x H, row o = output) and W3 (O x H) into the B-operand order of
// v_mfma_f32_16x16x32_bf16: for output tile t (16 columns), k-step s (32),
// lane l: 8 consecutive k = 32 s + 8 (l / 16) .. + 7 of column 16 t + l % 16.
__global__ void k_pack_b(int O, int H, const float *__restrict__ W, short *__restrict__ out, int ntiles) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over ntiles * (H/32) * 64 lanes
  const int nks = H / 32;
  if (idx >= ntiles * nks * 64) return;
  const int l = idx & 63, s = (idx >> 6) % nks, t = (idx >> 6) / nks;
  const int col = 16 * t + (l & 15);
  const int k0 = 32 * s + 8 * (l >> 4);
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const float w = col < O ? W[(size_t)col * H + k0 + q] : 0.0f;
    out[(size_t)idx * 8 + q] = f2bf(w);
  }
}

__device__ __forceinline__ floatx4 mfma_bf16(const shortx8 &a, const shortx8 &b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

// one wave = 16 rows; block = 4 waves = 64 rows
__global__ __launch_bounds__(256) void k_mlp_infer_bf16(int B, int O, const float *__restrict__ X,
                                                        const float *__restrict__ W1, const float *__restrict__ b1,
                                                        const shortx8 *__restrict__ W2p,
                                                        const float *__restrict__ b2,
                                                        const shortx8 *__restrict__ W3p,
                                                        const float *__restrict__ b3, float *__restrict__ Y) {
  __shared__ short sH[4][16][MH + 8];  // +8 shorts: 16-byte aligned rows, conflict-free
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int g = l >> 4, i = l & 15;
  const int row0 = (blockIdx.x * 4 + wv) * 16;
  if (row0 >= B) return;  // whole wave leaves together
  const int r = row0 + i;
  // ---- layer 1: lane computes H1[r][32 s + 8 g + q] for s < 4, q < 8
  float x[MI];
#pragma unroll
  for (int k = 0; k < MI; ++k) x[k] = r < B ? X[(size_t)r * MI + k] : 0.0f;
  shortx8 a1[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int n = 32 * s + 8 * g + q;
      float z = b1[n];
#pragma unroll
      for (int k = 0; k < MI; ++k) z += x[k] * W1[n * MI + k];
      a1[s][q] = f2bf(tanhf(z));
    }
  }
  // ---- layer 2: 8 output tiles x 4 k-steps
  floatx4 acc[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    acc[t] = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[t] = mfma_bf16(a1[s], W2p[(t * 4 + s) * 64 + l], acc[t]);
  }
  // tanh and store the 16 x 128 tile (row = 4 g + v, col = 16 t + i)
#pragma unroll
  for (int t = 0; t < 8; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int col = 16 * t + i;
      sH[wv][4 * g + v][col] = f2bf(tanhf(acc[t][v] + b2[col]));
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // ---- layer 3: A operand = H2[i][32 s + 8 g .. +7]
  shortx8 a2[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) a2[s] = *reinterpret_cast<const shortx8 *>(&sH[wv][i][32 * s + 8 * g]);
  const int nt3 = (O + 15) / 16;
#pragma unroll
  for (int t = 0; t < MOP / 16; ++t) {
    if (t >= nt3) break;
    floatx4 c = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) c = mfma_bf16(a2[s], W3p[(t * 4 + s) * 64 + l], c);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = row0 + 4 * g + v, col = 16 * t + i;
      if (rr < B && col < O) Y[(size_t)rr * O + col] = c[v] + b3[col];
    }
  }
}

// ------------------------------------------------------------ residual check
// rel = |w_pred - w_ref|_2 / |w_ref|_2 (FP64); accepted elements keep the
// prediction, the others fall back to the plain weights.
__global__ void k_residual(int nel, int nq, const float *__restrict__ pred, double scale,
                           const double *__restrict__ ref, double thr, double *__restrict__ rel,
                           uint8_t *__restrict__ accept, double *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double num = 0.0, den = 0.0;
  for (int k = 0; k < nq; ++k) {
    const double p = (double)pred[(size_t)e * nq + k] * scale;
    const double q = ref[(size_t)e * nq + k];
    num += (p - q) * (p - q);
    den += q * q;
  }
  const double rr = sqrt(num) / (sqrt(den) + 1e-300);
  if (rel) rel[e] = rr;
  const bool ok = rr <= thr;
  if (accept) accept[e] = ok;
  if (out)
    for (int k = 0; k < nq; ++k)
      out[(size_t)e * nq + k] = ok ? (double)pred[(size_t)e * nq + k] * scale : ref[(size_t)e * nq + k];
}

inline int nblk(long long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

// params layout (FP32): W1[H][I], b1[H], W2[H][H], b2[H], W3[O][H], b3[O]
extern "C" int xg_mlp_nparams(int I, int H, int O) { return H * I + H + H * H + H + O * H + O; }

extern "C" int xg_mlp_forward(xg_ctx *ctx, int I, int H, int O, const float *params, int B, const float *x,
                              float *y, float *work, int precision) {
  if (!ctx || I < 1 || H < 1 || O < 1 || B < 0 || !params || (B > 0 && (!x || !y)))
    return set_error(XG_ERR_ARG, "xg_mlp_forward: bad argument");
  if (B == 0) return XG_OK;
  const float *W1 = params, *b1 = W1 + H * I, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + O * H;
  if (precision == XG_PREC_FP32) {
    if (!work) return set_error(XG_ERR_ARG, "xg_mlp_forward: FP32 path needs work (2*B*H floats)");
    float *h1 = work, *h2 = work + (size_t)B * H;
    k_dense_fwd<<<nblk((long long)B * H, 256), 256, 0, ctx->stream>>>(B, I, H, x, W1, b1, 1, h1);
    k_dense_fwd<<<nblk((long long)B * H, 256), 256, 0, ctx->stream>>>(B, H, H, h1, W2, b2, 1, h2);
    k_dense_fwd<<<nblk((long long)B * O, 256), 256, 0, ctx->stream>>>(B, H, O, h2, W3, b3, 0, y);
    XG_LAUNCH_CHECK();
    return XG_OK;
  }
  if (precision != XG_PREC_BF16)
    return set_error(XG_ERR_UNSUPPORTED, "xg_mlp_forward: precision must be FP32 or BF16");
  if (I != MI || H != MH || O > MOP)
    return set_error(XG_ERR_UNSUPPORTED, "xg_mlp_forward: BF16 path is built for 10 -> 128 -> 128 -> <=96");
  // packed B operands live in the context scratch
  const size_t n2 = (size_t)8 * 4 * 64 * 8, n3 = (size_t)(MOP / 16) * 4 * 64 * 8;
  int rc = ensure_ws(ctx, sizeof(short) * (n2 + n3));
  if (rc) return rc;
  short *W2p = (short *)ctx->ws, *W3p = W2p + n2;
  k_pack_b<<<nblk(8 * 4 * 64, 256), 256, 0, ctx->stream>>>(H, H, W2, W2p, 8);
  k_pack_b<<<nblk((MOP / 16) * 4 * 64, 256), 256, 0, ctx->stream>>>(O, H, W3, W3p, MOP / 16);
  k_mlp_infer_bf16<<<nblk(B, 64), 256, 0, ctx->stream>>>(B, O, x, W1, b1, (const shortx8 *)W2p, b2,
                                                           (const shortx8 *)W3p, b3, y);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

// One Adam step on a minibatch (FP32).  work: 4*B*H + 2*B*O floats.
// grad (nullable): receives the gradient (params layout).  loss: device
// float[1], overwritten with 0.5/B sum (y - t)^2.
extern "C" int xg_mlp_train_step(xg_ctx *ctx, int I, int H, int O, float *params, float *adam_m, float *adam_v,
                                 int step, float lr, int B, const float *x, const float *target, float *loss,
                                 float *grad, float *work) {
  if (!ctx || I < 1 || H < 1 || O < 1 || B < 1 || !params || !x || !target || !work || step < 1)
    return set_error(XG_ERR_ARG, "xg_mlp_train_step: bad argument");
  const int np = xg_mlp_nparams(I, H, O);
  float *W1 = params, *b1 = W1 + H * I, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + O * H;
  float *h1 = work, *h2 = h1 + (size_t)B * H, *d2 = h2 + (size_t)B * H, *d1 = d2 + (size_t)B * H;
  float *y = d1 + (size_t)B * H, *dy = y + (size_t)B * O;
  int rc = ensure_ws(ctx, sizeof(float) * (size_t)np);
  if (rc) return rc;
  float *g = grad ? grad : (float *)ctx->ws;
  float *gW1 = g, *gb1 = gW1 + H * I, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + O * H;
  hipStream_t s = ctx->stream;
  k_dense_fwd<<<nblk((long long)B * H, 256), 256, 0, s>>>(B, I, H, x, W1, b1, 1, h1);
  k_dense_fwd<<<nblk((long long)B * H, 256), 256, 0, s>>>(B, H, H, h1, W2, b2, 1, h2);
  k_dense_fwd<<<nblk((long long)B * O, 256), 256, 0, s>>>(B, H, O, h2, W3, b3, 0, y);
  if (loss) XG_CHECK(hipMemsetAsync(loss, 0, sizeof(float), s));
  k_mse<<<nblk((long long)B * O, 256), 256, 0, s>>>((long long)B * O, B, y, target, dy, loss);
  // layer 3
  k_dense_bwd_w<<<nblk(O * (H + 1), 256), 256, 0, s>>>(B, H, O, dy, h2, gW3, gb3);
  k_dense_bwd_x<<<nblk((long long)B * H, 256), 256, 0, s>>>(B, H, O, dy, W3, d2);
  k_tanh_bwd<<<nblk((long long)B * H, 256), 256, 0, s>>>((long long)B * H, h2, d2);
  // layer 2
  k_dense_bwd_w<<<nblk(H * (H + 1), 256), 256, 0, s>>>(B, H, H, d2, h1, gW2, gb2);
  k_dense_bwd_x<<<nblk((long long)B * H, 256), 256, 0, s>>>(B, H, H, d2, W2, d1);
  k_tanh_bwd<<<nblk((long long)B * H, 256), 256, 0, s>>>((long long)B * H, h1, d1);
  // layer 1
  k_dense_bwd_w<<<nblk(H * (I + 1), 256), 256, 0, s>>>(B, I, H, d1, x, gW1, gb1);
  XG_LAUNCH_CHECK();
  if (adam_m && adam_v && lr > 0.0f) {
    const float b1c = 0.9f, b2c = 0.999f;
    const float c1 = 1.0f - powf(b1c, (float)step), c2 = 1.0f - powf(b2c, (float)step);
    k_adam<<<nblk(np, 256), 256, 0, s>>>(np, params, g, adam_m, adam_v, lr, b1c, b2c, 1e-8f, c1, c2);
    XG_LAUNCH_CHECK();
  }
  return XG_OK;
}

extern "C" int xg_moment_residual(xg_ctx *ctx, int nel, int nq, const float *pred, double scale,
                                  const double *ref, double thr, double *rel, uint8_t *accept, double *out) {
  if (!ctx || nel < 0 || nq < 1 || (nel > 0 && (!pred || !ref)))
    return set_error(XG_ERR_ARG, "xg_moment_residual: bad argument");
  if (nel == 0) return XG_OK;
  k_residual<<<nblk(nel, 256), 256, 0, ctx->stream>>>(nel, nq, pred, scale, ref, thr, rel, accept, out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}
