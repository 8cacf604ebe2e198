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

// seed of the h-rho iteration: h = h_pred exp(y / scale), |y/scale| <= 0.1,
// applied to the predictor's FP64 h in place (a zero correction leaves it
// unchanged); fixed boundary particles (itype 1, when given) keep the h the
// predictor reset them to (src/stepND_leapfrog_mhd.f90:136)
__global__ __launch_bounds__(256) void k_mlp_seed(const uint16_t* __restrict__ wbf, const float* __restrict__ X,
                                                  int n, double* __restrict__ h, const int* __restrict__ itype) {
  __shared__ float sw[MLP_NPARAM];
  load_weights(wbf, sw);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (itype && itype[i] == 1) return;
  float x[NIN], z1[NH], a1[NH], z2[NH], a2[NH];
  for (int q = 0; q < NIN; q++) x[q] = X[(size_t)i * MLP_LDX + q];
  float y = mlp_fwd(sw, x, z1, a1, z2, a2) / MLP_TSCALE;
  y = fminf(fmaxf(y, -0.1f), 0.1f);
  const double hpred = h[i];
  if (isfinite(y) && y != 0.0f && hpred > 0.0) h[i] = hpred * exp((double)y);
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
// host launchers
// ---------------------------------------------------------------------------

static inline unsigned nb(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

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

void mlp_launch_seed(MlpState& s, int n, double* h, const int* itype, hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_seed, dim3(nb(n, 256)), dim3(256), 0, st, (const uint16_t*)s.wbf, (const float*)s.X, n, h,
                     itype);
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
