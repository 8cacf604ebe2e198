// Host-side interface of the h-predictor MLP kernels (ndsp_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLP_NIN 6
#define MLP_NH 32
#define MLP_NPARAM (MLP_NH * MLP_NIN + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1)
// feature rows: 6 features, predictor h, target
#define MLP_LDX 8
#define MLP_COL_HPRED 6
#define MLP_COL_TARGET 7
// the network predicts MLP_TSCALE * log(h_converged / h_predictor)
#define MLP_TSCALE 1000.0f

struct MlpState {
  float* w32 = nullptr;     // FP32 master weights (Adam)
  uint16_t* wbf = nullptr;  // BF16 weights used by every forward
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  float* loss = nullptr;
  float* X = nullptr;       // per-particle feature rows [n][MLP_LDX]
  float* Xext = nullptr;    // staging for caller-supplied batches
  int nrows = 0;
  int xext_rows = 0;
  int adam_steps = 0;
  int have_params = 0;
  // configuration
  int enabled = 0, train = 0, seed_after = 2, iters = 4, batch = 16384;
  float lr = 1.0e-3f;
  int steps_seen = 0;
  double last_loss = 0.0;
};

// the particle fields the seed's features are formed from (one evaluation
// after the predictor); h is scaled in place when apply is set, X receives
// the feature rows (6 features, predictor h) when not null
struct MlpSeedFields {
  const double* hin;
  const double* rho;
  const double* m;
  const double* dhdtin;
  const double* drhodt;
  const double* corr;
  const int* numneigh;
  const int* nits;
  const int* itype;  // null: no fixed particles
  double* h;
  float* X;
  double dt, nexp, hfact, dndim;
  int n, apply;
};

hipError_t mlp_alloc(MlpState& s, int n);
void mlp_free(MlpState& s);
hipError_t mlp_set_params(MlpState& s, const float* host, hipStream_t st);
void mlp_launch_forward(MlpState& s, const float* X, int ldx, int B, float* y, hipStream_t st);
// the same network on the matrix cores (BF16 inputs and activations)
void mlp_launch_infer(MlpState& s, const float* X, int ldx, int B, float* y, hipStream_t st);
void mlp_launch_seed_mfma(MlpState& s, const MlpSeedFields& f, hipStream_t st);
void mlp_launch_grad(MlpState& s, const float* X, int nrows, int B, long long off, long long stride,
                     hipStream_t st);
void mlp_launch_adam(MlpState& s, float lr, hipStream_t st);
