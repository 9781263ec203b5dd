// spb6_mlp.h -- learned Newton guess: network constants and launchers (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include "spb6_kernels.h"

namespace spb {

constexpr int MLP_IN = 175, MLP_H = 128, MLP_KPAD = 224;   // layer-1 k padded for the MFMA slot order
// parameter vector layout (FP32): W1[175 x 128] (row k, column n), b1[128], W2[128 x 128], b2[128], w3[128], b3
constexpr int MLP_OFF_W1 = 0, MLP_OFF_B1 = MLP_IN * MLP_H, MLP_OFF_W2 = MLP_OFF_B1 + MLP_H, MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H,
              MLP_OFF_W3 = MLP_OFF_B2 + MLP_H, MLP_OFF_B3 = MLP_OFF_W3 + MLP_H, MLP_NPARAM = MLP_OFF_B3 + 1;

// device copy of a trained network for inference
struct MlpDev {
	const float* params;
	const __bf16* W1s; const __bf16* W1f; const __bf16* W2p;   // BF16 layer 1 for standardised / shifted inputs, layer 2
	const float* mub1;                                          // per-field input shift mu[7] and the folded bias b1f[128]
	float b3, mu_t, sd_t;
};

void mlp_feat_cell(const Dev& D, const double* Pplain, float* F, hipStream_t st);
void mlp_dump(const Dev& D, const float* F, double tau, const double* Pconv, const double* Pplain, const double* Gplain, double thr,
              int off, int stride, long cap, float* X, float* Y, unsigned long long* count, hipStream_t st);
void mlp_pack(const float* params, const float* mean, const float* istd, __bf16* W1s, __bf16* W1f, __bf16* W2p, float* mub1, hipStream_t st);
// the learned guess into Pout (D given) or the network's outputs for standardised rows Xin into yout; with an activity
// mask (Gplain) and list / count (device scratch of nrows / 16 + 1 ints), the active tiles are listed first
void mlp_infer(const Dev* D, float* F, double tau, int nrows, const float* Xin, const MlpDev& W, const double* Pplain,
               const double* Gplain, double thr, double* Pout, float* yout, hipStream_t st, int* list = nullptr, int* count = nullptr,
               const unsigned long long* nrm = nullptr, const int* flag = nullptr, double* thr_dev = nullptr, double* Gout = nullptr,
               int* rlist = nullptr, int* rcount = nullptr, unsigned long long* nrm2 = nullptr, int* flag2 = nullptr, bool feats = false);
// (nrm / flag / thr_dev: the activity threshold computed on the device from the plain residual's norm slots and flag;
// Gout ... flag2, with thr_dev: the guess's residual outside the cells it can change -- the plain residual -- into Gout,
// its norm into nrm2 and its failure into flag2, and the cells left to evaluate (launch_residual_list) into rlist;
// feats: the features F computed here (of the listed tiles' stencils where the tiles are staged), instead of by the
// caller's mlp_feat_cell)
// a cell takes the network's guess (and is dumped for training) when its plain-guess residual is at least this fraction
// of the plain residual's max norm: elsewhere the plain guess is kept, so quiescent cells are never perturbed
constexpr double GUESS_ACTIVE = 1e-3;
void mlp_fwd_bwd(const float* P, const float* Xb, const float* Yb, int Bn, float* ws, float* G, float* yout, float* loss, hipStream_t st);
void mlp_adam(float* P, const float* G, float* m, float* v, float lr, int t, hipStream_t st);
void mlp_gather(int Bn, const float* X, const float* Y, const int* idx, float* Xb, float* Yb, hipStream_t st);
inline long mlp_ws_floats(int Bn) { return (long)Bn * MLP_H * 6 + 2L * Bn; }

}   // namespace spb
