// Host state of the band-tail closure network (rp_mlp.cpp); complete here so
// that rp_ctx can own it.
#pragma once
#include <cstdint>
#include <vector>

#include "rp_common.h"

namespace rp {

struct Mlp {
  std::vector<int> dims;           // layer widths: nin, hidden..., nout
  std::vector<size_t> woff, boff;  // offsets of W_l and b_l in the flat parameter vector
  size_t nparams = 0;
  int64_t step = 0;                // Adam steps taken
  int nmax = 0;                    // batch capacity of the activation buffers
  DevBuf<float> d_p, d_g, d_m1, d_m2;  // FP32 masters, gradient, Adam moments
  DevBuf<unsigned short> d_pb;         // BF16 copy of the parameters
  DevBuf<float> d_z, d_a;              // pre-activations / activations, per layer [n][width]
  DevBuf<float> d_dz0, d_dz1, d_part;  // gradient work buffers
  DevBuf<float> d_x, d_t, d_xb, d_tb;  // data (resident for training) and minibatch
  DevBuf<int> d_idx;
  DevBuf<double> d_loss;
  std::vector<size_t> aoff;            // offsets of layer l's [n][width] in d_z / d_a
};

}  // namespace rp
