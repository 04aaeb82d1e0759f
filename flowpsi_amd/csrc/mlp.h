// Learned SGS initial guess: buffers and launchers (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace fpsi {

// offsets of the parameter vector (fp32 master copy and bf16 copy)
struct MlpLayout {
  int w0, b0, w1, b1, w2, b2, w3, b3, n;
};
MlpLayout mlp_layout();
int mlp_parts(int N);

struct MlpBufs {
  // graph: face neighbours of each cell (interior faces and periodic partners)
  const int *nb_off, *nb_idx;
  const float *geo;      // [N][4] shape factors and log volume
  float *raw;            // [N][32] X0, previous update, log|U_kk|, -, r0 = B - A X0, D^-1 r0
  float *norm;           // [32] sx[5], sp[5], mean log|U_kk|[5], -, sr[5], sd[5]
  float *x;              // [N][32] features of the current input
  float *y;              // [N][8] target
  float *h0, *a1, *h1, *a2, *h2, *out;  // activations
  float *dout;           // [N][32]
  float *g0, *g1, *g2, *da, *dcat;      // gradients
  float *w, *grad, *am, *av;            // fp32 parameters, gradient, Adam moments
  unsigned short *wb, *wt1, *wt2, *wt3; // bf16 copies
  const unsigned char *mask;            // trainable parameters
  const float *zero64;
  float *wpart, *bpart, *lpart, *loss;  // reduction scratch
};

void mlp_forward(const MlpBufs &m, int N, hipStream_t st);
// the guess X = X0 + sd * out[0:5] from the normalised features m.x (the
// forward pass of mlp_forward, fused; h0 and h1 are kept, nothing after)
void mlp_infer_guess(const MlpBufs &m, int N, double *X, hipStream_t st);
void mlp_backward(const MlpBufs &m, int N, hipStream_t st);
// raw features of the current system: X0 = X, the previous Newton update
// Xprev [N][5], the diagonal factors lu (real_fj packed pivoted LU with piv,
// or with lu_f64 the unpivoted FP64 factors and piv null), and
// r0 = B - A X0 [N][5]
void mlp_features(const MlpBufs &m, int N, int fp32, const double *X, const double *Xprev, const void *lu,
                  int lu_f64, const int *piv, const double *r0, hipStream_t st);
void mlp_feature_stats(const MlpBufs &m, int N, float *out32, hipStream_t st);
void mlp_normalise(const MlpBufs &m, int N, float *x, hipStream_t st);
void mlp_target(const MlpBufs &m, int N, const double *Xs, float *y, hipStream_t st);
void mlp_adam(const MlpBufs &m, float lr, int t, hipStream_t st);
void mlp_refresh_bf16(const MlpBufs &m, hipStream_t st);

}  // namespace fpsi
