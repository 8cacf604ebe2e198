// h-predictor MLP for the h-rho iteration: 6 features -> 32 -> 32 (GELU) -> 1,
// BF16 weights with FP32 accumulation, trained online on the device by the
// kernels below (forward, backward, Adam).  Its output only seeds the
// Newton-Raphson iteration of k_nr_pass, which still converges every particle
// to tolh; the reference's seed is h = h_in + dt dh/dt
// (src/stepND_leapfrog_mhd.f90:153).
//
// Parameter vector (float[NPARAM]): W1[32][6] b1[32] W2[32][32] b2[32] W3[32] b3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "ndsp_mlp.h"

#define NIN MLP_NIN
#define NH MLP_NH
#define WB 64  // samples per training block (one wavefront)

__device__ __forceinline__ float bf16_to_f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ uint16_t f_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(r >> 16);
}

__device__ __forceinline__ float gelu(float x) {
  const float c = 0.7978845608028654f;
  return 0.5f * x * (1.0f + tanhf(c * (x + 0.044715f * x * x * x)));
}
__device__ __forceinline__ float dgelu(float x) {
  const float c = 0.7978845608028654f;
  const float u = c * (x + 0.044715f * x * x * x);
  const float t = tanhf(u);
  const float du = c * (1.0f + 3.0f * 0.044715f * x * x);
  return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * du;
}

__device__ __forceinline__ void load_weights(const uint16_t* __restrict__ wbf, float* sw) {
  for (int p = threadIdx.x; p < MLP_NPARAM; p += blockDim.x) sw[p] = bf16_to_f(wbf[p]);
  __syncthreads();
}

// forward of one sample; keeps z1, z2 (pre-activations) and a1, a2
__device__ __forceinline__ float mlp_fwd(const float* sw, const float* x, float* z1, float* a1, float* z2,
                                         float* a2) {
  const float* W1 = sw;
  const float* b1 = W1 + NH * NIN;
  const float* W2 = b1 + NH;
  const float* b2 = W2 + NH * NH;
  const float* W3 = b2 + NH;
  const float b3 = W3[NH];
  for (int o = 0; o < NH; o++) {
    float s = b1[o];
    for (int i = 0; i < NIN; i++) s += W1[o * NIN + i] * x[i];
    z1[o] = s;
    a1[o] = gelu(s);
  }
  for (int o = 0; o < NH; o++) {
    float s = b2[o];
    for (int i = 0; i < NH; i++) s += W2[o * NH + i] * a1[i];
    z2[o] = s;
    a2[o] = gelu(s);
  }
  float y = b3;
  for (int i = 0; i < NH; i++) y += W3[i] * a2[i];
  return y;
}

// inference on a feature matrix X[B][8] (columns 0..5 are the features)
__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t* __restrict__ wbf, const float* __restrict__ X,
                                                     int ldx, int B, float* __restrict__ y) {
  __shared__ float sw[MLP_NPARAM];
  load_weights(wbf, sw);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float x[NIN], z1[NH], a1[NH], z2[NH], a2[NH];
  for (int i = 0; i < NIN; i++) x[i] = X[(size_t)b * ldx + i];
  y[b] = mlp_fwd(sw, x, z1, a1, z2, a2);
}

// gradient of L = sum (y - t)^2 / B over the samples idx(b), b < B; sample b
// is row (off + b * stride) mod nrows of X (target in column tcol)
__global__ __launch_bounds__(WB) void k_mlp_grad(const uint16_t* __restrict__ wbf, const float* __restrict__ X,
                                                 int nrows, int tcol, int B, long long off, long long stride,
                                                 float* __restrict__ grad, float* __restrict__ loss) {
  __shared__ float sw[MLP_NPARAM];
  __shared__ float sX[WB][NIN + 1];
  __shared__ float sA1[WB][NH + 1];
  __shared__ float sA2[WB][NH + 1];
  __shared__ float sD1[WB][NH + 1];
  __shared__ float sD2[WB][NH + 1];
  __shared__ float sDy[WB];
  load_weights(wbf, sw);
  const int t = threadIdx.x;
  const int b = blockIdx.x * WB + t;
  const bool act = b < B;
  float x[NIN], z1[NH], a1[NH], z2[NH], a2[NH];
  float dy = 0.0f, r = 0.0f;
  if (act) {
    const long long row = (off + (long long)b * stride) % nrows;
    for (int i = 0; i < NIN; i++) x[i] = X[row * MLP_LDX + i];
    const float y = mlp_fwd(sw, x, z1, a1, z2, a2);
    r = y - X[row * MLP_LDX + tcol];
    dy = 2.0f * r / (float)B;
  } else {
    for (int i = 0; i < NIN; i++) x[i] = 0.0f;
    for (int o = 0; o < NH; o++) { z1[o] = a1[o] = z2[o] = a2[o] = 0.0f; }
  }
  const float* W2 = sw + NH * NIN + NH;
  const float* W3 = W2 + NH * NH + NH;
  float d2[NH];
  for (int o = 0; o < NH; o++) d2[o] = dy * W3[o] * dgelu(z2[o]);
  for (int i = 0; i < NH; i++) {
    float s = 0.0f;
    for (int o = 0; o < NH; o++) s += W2[o * NH + i] * d2[o];
    sD1[t][i] = act ? s * dgelu(z1[i]) : 0.0f;
    sA1[t][i] = a1[i];
    sA2[t][i] = a2[i];
    sD2[t][i] = act ? d2[i] : 0.0f;
  }
  for (int i = 0; i < NIN; i++) sX[t][i] = x[i];
  sDy[t] = dy;
  __syncthreads();
  // each thread sums whole parameter gradients over the block's samples
  for (int p = t; p < MLP_NPARAM; p += WB) {
    float s = 0.0f;
    int q = p;
    if (q < NH * NIN) {
      const int o = q / NIN, i = q % NIN;
      for (int k = 0; k < WB; k++) s += sD1[k][o] * sX[k][i];
    } else if ((q -= NH * NIN) < NH) {
      for (int k = 0; k < WB; k++) s += sD1[k][q];
    } else if ((q -= NH) < NH * NH) {
      const int o = q / NH, i = q % NH;
      for (int k = 0; k < WB; k++) s += sD2[k][o] * sA1[k][i];
    } else if ((q -= NH * NH) < NH) {
      for (int k = 0; k < WB; k++) s += sD2[k][q];
    } else if ((q -= NH) < NH) {
      for (int k = 0; k < WB; k++) s += sDy[k] * sA2[k][q];
    } else {
      for (int k = 0; k < WB; k++) s += sDy[k];
    }
    atomicAdd(&grad[p], s);
  }
  float l = act ? r * r / (float)B : 0.0f;
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
  if (t == 0 && loss) atomicAdd(loss, l);
}

// Adam update of the FP32 master weights, BF16 copy, gradient reset
__global__ void k_mlp_adam(float* __restrict__ w, uint16_t* __restrict__ wbf, float* __restrict__ m,
                           float* __restrict__ v, float* __restrict__ grad, float lr, float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = grad[p];
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  const float ww = w[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
  w[p] = ww;
  wbf[p] = f_to_bf16(ww);
  grad[p] = 0.0f;
}

__global__ void k_mlp_fill(float* __restrict__ a, int n, float v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = v;
}

__global__ void k_mlp_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < MLP_NPARAM) wbf[p] = f_to_bf16(w[p]);
}

// ---------------------------------------------------------------------------
// Inference on the matrix cores (v_mfma_f32_32x32x8bf16_1k): BF16 inputs and
// activations, FP32 accumulation.  A wavefront takes 64 samples as two
// blocks of 32.  Each layer is computed transposed, Z^T = W A^T: the weights
// are the A operand (row = output unit = lane % 32, four inputs 4 (lane / 32)
// + j per lane), the activations the B operand (column = sample = lane % 32,
// four inputs 4 (lane / 32) + j per lane), and the result D holds, in its
// register v, output unit 8 (v / 4) + 4 (lane / 32) + v % 4 of sample
// lane % 32 -- registers 4s..4s+3 are then exactly the B operand of the next
// layer's K-step s, so activations never leave the registers.  Layer 1 pads
// the 6 features to K = 8; layer 2 is four K-steps; layer 3 (32 -> 1) is a
// dot product per lane over its 16 units plus one exchange between the two
// half-waves.
// ---------------------------------------------------------------------------
typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x16 __attribute__((ext_vector_type(16)));

// tanh-form GELU with the hardware exponential and reciprocal:
// 0.5 x (1 + tanh u) = x / (1 + exp(-2u)); for very negative u the
// exponential overflows to inf and the quotient takes its limit 0
__device__ __forceinline__ float gelu_fast(float x) {
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  return x * __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

// per-lane weights of the transposed layers (see above)
struct MlpLaneW {
  mlp_bf16x4 a1, a2[4];
  float b1[16], b2[16], w3[16];
  float b3;
};
__device__ __forceinline__ void mlp_lane_weights(const uint16_t* __restrict__ wbf, MlpLaneW& w) {
  const int lane = threadIdx.x & 63, i = lane & 31, hb = lane >> 5;
  const uint16_t* W1 = wbf;
  const uint16_t* B1 = W1 + NH * NIN;
  const uint16_t* W2 = B1 + NH;
  const uint16_t* B2 = W2 + NH * NH;
  const uint16_t* W3 = B2 + NH;
  for (int j = 0; j < 4; j++) {
    const int kin = 4 * hb + j;
    w.a1[j] = kin < NIN ? (short)W1[i * NIN + kin] : (short)0;
    for (int s = 0; s < 4; s++) w.a2[s][j] = (short)W2[i * NH + 8 * s + 4 * hb + j];
  }
  for (int v = 0; v < 16; v++) {
    const int u = 8 * (v >> 2) + 4 * hb + (v & 3);
    w.b1[v] = bf16_to_f(B1[u]);
    w.b2[v] = bf16_to_f(B2[u]);
    w.w3[v] = bf16_to_f(W3[u]);
  }
  w.b3 = bf16_to_f(W3[NH]);
}

// the network on one block of 32 samples: x0..x3 are the lane's four inputs
// (features 4 (lane / 32) + j of sample lane % 32); returns the output of
// sample lane % 32 in both half-waves
__device__ __forceinline__ float mlp_mfma_block(const MlpLaneW& w, const float (&x)[4]) {
  mlp_bf16x4 b;
  for (int j = 0; j < 4; j++) b[j] = (short)f_to_bf16(x[j]);
  mlp_f32x16 zero;
  for (int v = 0; v < 16; v++) zero[v] = 0.0f;
  const mlp_f32x16 d1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w.a1, b, zero, 0, 0, 0);
  mlp_bf16x4 h[4];
  for (int v = 0; v < 16; v++) h[v >> 2][v & 3] = (short)f_to_bf16(gelu_fast(d1[v] + w.b1[v]));
  mlp_f32x16 d2 = zero;
  for (int s = 0; s < 4; s++) d2 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(w.a2[s], h[s], d2, 0, 0, 0);
  float part = 0.0f;
  for (int v = 0; v < 16; v++) part += w.w3[v] * gelu_fast(d2[v] + w.b2[v]);
  return part + __shfl_xor(part, 32) + w.b3;
}

// inference on caller rows X[B][ldx] (features in columns 0..5)
__global__ __launch_bounds__(256) void k_mlp_infer(const uint16_t* __restrict__ wbf, const float* __restrict__ X,
                                                   int ldx, int B, float* __restrict__ y) {
  MlpLaneW w;
  mlp_lane_weights(wbf, w);
  const int lane = threadIdx.x & 63, i = lane & 31, hb = lane >> 5;
  // each wave takes 64 samples at a time, strided over the grid (the loop
  // bound is uniform in the wave, so every lane reaches every MFMA)
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < B; base += step)
    for (int r = 0; r < 2; r++) {
      const long long p = base + 32 * r + i;
      float x[4] = {0.f, 0.f, 0.f, 0.f};
      if (p < B)
        for (int j = 0; j < 4; j++) {
          const int k = 4 * hb + j;
          if (k < NIN) x[j] = X[p * ldx + k];
        }
      const float out = mlp_mfma_block(w, x);
      if (hb == 0 && p < B) y[p] = out;
    }
}

// the seed of the h-rho iteration after the predictor: features formed from
// the particle fields, optionally stored as training rows, and (apply) the
// predictor's h scaled by exp(y / MLP_TSCALE), |y / MLP_TSCALE| <= 0.1;
// fixed boundary particles (itype 1) keep the h the predictor gave them
__global__ __launch_bounds__(256) void k_mlp_seed_mfma(const uint16_t* __restrict__ wbf, MlpSeedFields f) {
  MlpLaneW w;
  mlp_lane_weights(wbf, w);
  const int lane = threadIdx.x & 63, i = lane & 31, hb = lane >> 5;
  const float dtf = (float)f.dt, nexpf = (float)f.nexp, inv_nexp = (float)(1.0 / f.nexp);
  const float dndimf = (float)f.dndim, inv_hfact = (float)(1.0 / f.hfact);
  // 64 particles a wave at a time, strided over the grid, so that the lane
  // weights are loaded once per wave (the loop bound is uniform in the wave)
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < f.n; base += step)
    for (int r = 0; r < 2; r++) {
      const long long p = base + 32 * r + i;
      const bool ok = p < f.n;
      float x[4] = {0.f, 0.f, 0.f, 0.f};
      double hpred = 0.0;
      // features in FP32 (network inputs; the same arithmetic forms the
      // training rows), the half-wave taking features 0-3 or 4-5
      if (ok) {
        hpred = f.h[p];
        const float hin = (float)f.hin[p], rho = (float)f.rho[p];
        if (hb == 0) {
          x[0] = 100.0f * dtf * (float)f.dhdtin[p] * __builtin_amdgcn_rcpf(hin);
          x[1] = rho > 0.0f ? 100.0f * dtf * (float)f.drhodt[p] * __builtin_amdgcn_rcpf(rho) : 0.0f;
          x[2] = (float)(1000.0 * f.corr[p]);
          x[3] = ((float)f.numneigh[p] - nexpf) * inv_nexp;
        } else {
          x[0] = (float)f.nits[p] - 1.5f;
          x[1] = rho > 0.0f
                     ? 1000.0f * (hin * exp2f(dndimf * __log2f(rho * __builtin_amdgcn_rcpf((float)f.m[p]))) * inv_hfact -
                                  1.0f)
                     : 0.0f;
        }
      }
      const float out = mlp_mfma_block(w, x);
      if (!ok) continue;
      if (f.X) {
        float* row = f.X + p * MLP_LDX;
        if (hb == 0) {
          for (int j = 0; j < 4; j++) row[j] = x[j];
        } else {
          row[4] = x[0];
          row[5] = x[1];
          row[MLP_COL_HPRED] = (float)hpred;
        }
      }
      if (f.apply && hb == 0 && !(f.itype && f.itype[p] == 1)) {
        float yy = out / MLP_TSCALE;
        yy = fminf(fmaxf(yy, -0.1f), 0.1f);
        if (isfinite(yy) && yy != 0.0f && hpred > 0.0) f.h[p] = hpred * (double)__expf(yy);
      }
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static inline unsigned nb(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
// workgroups of the matrix-core inference: four waves each, a few per CU
static constexpr unsigned MLP_MFMA_BLOCKS = 1024;

hipError_t mlp_alloc(MlpState& s, int n) {
  hipError_t e;
  if ((e = hipMalloc(&s.w32, sizeof(float) * MLP_NPARAM))) return e;
  if ((e = hipMalloc(&s.wbf, sizeof(uint16_t) * MLP_NPARAM))) return e;
  if ((e = hipMalloc(&s.m, sizeof(float) * MLP_NPARAM))) return e;
  if ((e = hipMalloc(&s.v, sizeof(float) * MLP_NPARAM))) return e;
  if ((e = hipMalloc(&s.grad, sizeof(float) * MLP_NPARAM))) return e;
  if ((e = hipMalloc(&s.loss, sizeof(float) * 4))) return e;
  if ((e = hipMalloc(&s.X, sizeof(float) * (size_t)MLP_LDX * n))) return e;
  s.nrows = n;
  hipMemset(s.w32, 0, sizeof(float) * MLP_NPARAM);
  hipMemset(s.wbf, 0, sizeof(uint16_t) * MLP_NPARAM);
  hipMemset(s.m, 0, sizeof(float) * MLP_NPARAM);
  hipMemset(s.v, 0, sizeof(float) * MLP_NPARAM);
  hipMemset(s.grad, 0, sizeof(float) * MLP_NPARAM);
  hipMemset(s.X, 0, sizeof(float) * (size_t)MLP_LDX * n);
  return hipSuccess;
}

void mlp_free(MlpState& s) {
  void* p[] = {s.w32, s.wbf, s.m, s.v, s.grad, s.loss, s.X, s.Xext};
  for (void* q : p)
    if (q) (void)hipFree(q);
}

hipError_t mlp_set_params(MlpState& s, const float* host, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(s.w32, host, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, st);
  if (e) return e;
  hipLaunchKernelGGL(k_mlp_to_bf16, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, (const float*)s.w32, s.wbf);
  hipLaunchKernelGGL(k_mlp_fill, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, s.m, MLP_NPARAM, 0.0f);
  hipLaunchKernelGGL(k_mlp_fill, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, s.v, MLP_NPARAM, 0.0f);
  hipLaunchKernelGGL(k_mlp_fill, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, s.grad, MLP_NPARAM, 0.0f);
  s.adam_steps = 0;
  s.have_params = 1;
  return hipGetLastError();
}


void mlp_launch_seed_mfma(MlpState& s, const MlpSeedFields& f, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_seed_mfma, dim3(std::min(nb(f.n, 256), MLP_MFMA_BLOCKS)), dim3(256), 0, st,
                     (const uint16_t*)s.wbf, f);
}

void mlp_launch_infer(MlpState& s, const float* X, int ldx, int B, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_infer, dim3(std::min(nb(B, 256), MLP_MFMA_BLOCKS)), dim3(256), 0, st,
                     (const uint16_t*)s.wbf, X, ldx, B, y);
}

void mlp_launch_forward(MlpState& s, const float* X, int ldx, int B, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb(B, 256)), dim3(256), 0, st, (const uint16_t*)s.wbf, X, ldx, B, y);
}

void mlp_launch_grad(MlpState& s, const float* X, int nrows, int B, long long off, long long stride,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_fill, dim3(1), dim3(4), 0, st, s.loss, 4, 0.0f);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb(B, WB)), dim3(WB), 0, st, (const uint16_t*)s.wbf, X, nrows,
                     (int)MLP_COL_TARGET, B, off, stride, s.grad, s.loss);
}

void mlp_launch_adam(MlpState& s, float lr, hipStream_t st) {
  s.adam_steps++;
  const float bc1 = 1.0f - powf(0.9f, (float)s.adam_steps);
  const float bc2 = 1.0f - powf(0.999f, (float)s.adam_steps);
  hipLaunchKernelGGL(k_mlp_adam, dim3(nb(MLP_NPARAM, 256)), dim3(256), 0, st, s.w32, s.wbf, s.m, s.v, s.grad,
                     lr, bc1, bc2);
}
