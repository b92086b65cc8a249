// Learned correction of the block-Jacobi preconditioner of the reduced camera system.
//
// For camera block b with diagonal block D = S_bb:  J = diag(D)^-1/2, J D J = L L^T
// (Cholesky), B = J L^-T (so B^T D B = I).  A 3-layer MLP (32 -> 128 -> 128 -> 21, tanh
// hidden layers) maps 32 features of the block to a lower-triangular G (diagonal
// exp(o_ii), below it o_ij, row-major order) and the preconditioner block is
//     M_b^-1 = B G G^T B^T,
// symmetric positive definite for any output; G = I (zero output) is block-Jacobi.
// Training minimizes, over RCS snapshots and probes w ~ N(0, I) with v_b = B_b w_b,
// u = S v, u~_b = B_b^T u_b, the whitened local error
//     loss_b = |G G^T u~_b - w_b|^2 / |w_b|^2,
// i.e. how far M^-1 S is from the identity on each block in the D-norm.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "pcg.hpp"

namespace sgpu {

constexpr int kPcFeat = 32;
constexpr int kPcHid = 128;
constexpr int kPcOut = 21;
constexpr int kPcOffW1 = 0;
constexpr int kPcOffB1 = kPcOffW1 + kPcHid * kPcFeat;
constexpr int kPcOffW2 = kPcOffB1 + kPcHid;
constexpr int kPcOffB2 = kPcOffW2 + kPcHid * kPcHid;
constexpr int kPcOffW3 = kPcOffB2 + kPcHid;
constexpr int kPcOffB3 = kPcOffW3 + kPcOut * kPcHid;
constexpr int kPcWeights = kPcOffB3 + kPcOut;

// FP32 -> bf16 (round to nearest even) for the inference copy of the weights.
std::vector<uint16_t> weights_to_bf16(const float* w, int64_t n);

// Inference buffers (one per optimizer).
struct DevLearned {
  const uint16_t* wbf16;  // kPcWeights bf16 weights
  float* feat;            // nb x 32
  float* h1;              // nb x 128
  float* h2;              // nb x 128
  float* out;             // nb x 21
};

// Features of every camera block (needs launch_pcg_setup's full rows first).
void launch_precond_features(hipStream_t s, const DevPcg& g, float* feat);
// MLP forward on n feature rows with bf16 (wbf16) or FP32 (w32) weights.
void launch_mlp_forward(hipStream_t s, int64_t n, const float* feat, const float* w32,
                        const uint16_t* wbf16, float* h1, float* h2, float* out);
// Replaces g.minv by the learned blocks (blocks whose D is not positive definite keep
// block-Jacobi).
void launch_learned_minv(hipStream_t s, const DevPcg& g, const float* out);
// The whole inference: features, forward, blocks.
void launch_learned_preconditioner(hipStream_t s, const DevPcg& g, const DevLearned& L);

// Stage-level pieces for the tests (host arrays in and out; synchronous).
void precond_blocks_host(int64_t n, const double* diag_blocks, const float* outputs, double* minv);
void precond_loss_grad_host(int64_t n, const float* features, const double* diag_blocks,
                            const double* u, const double* w, const float* weights,
                            double* loss, float* grad);
void precond_forward_host(int64_t n, const float* features, const float* weights, int use_bf16,
                          float* outputs);

}  // namespace sgpu
