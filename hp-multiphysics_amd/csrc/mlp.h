// Internal interface between the solver context and the preconditioner MLP.
#pragma once
#include <hip/hip_runtime.h>

struct hpg_mlp;

namespace hpg {
// outputs: log scalings of the vertex, side and interior rows per variable ([type][var])
constexpr int MLP_IN = 11, MLP_OUT = 9;
// Y[n][9] = MLP(X[n][11]) on device buffers, enqueued on `stream`
void mlp_forward_device(const hpg_mlp* m, const float* X, int n, float* Y, hipStream_t stream);
}  // namespace hpg
