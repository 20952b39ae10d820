// C ABI of the band-tail closure network (declarations in include/respack_gpu.h).
// The network has no counterpart in the reference: it predicts the part of
// chi_GG'(w) that a truncated empty-band sum leaves out, and the full-band
// method stays the fallback and the judge.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "rp_internal.h"
#include "rp_mlp_state.h"

using namespace rp;

namespace {

Mlp& mlp_of(rp_handle h, const char* what) {
  RP_CHECK(h && h->mlp, std::string(what) + ": no network (rp_mlp_open)");
  return *h->mlp;
}

int nlayers(const Mlp& m) { return (int)m.dims.size() - 1; }

// Activation buffers for batches of up to n samples.
void reserve_batch(Mlp& m, int n) {
  if (n <= m.nmax) return;
  const int L = nlayers(m);
  m.aoff.assign(L + 1, 0);
  size_t tot = 0;
  int wmax = 0;
  for (int l = 1; l <= L; ++l) {
    m.aoff[l] = tot;
    tot += (size_t)n * m.dims[l];
  }
  for (int d : m.dims) wmax = std::max(wmax, d);
  m.d_z.reserve(tot);
  m.d_a.reserve(tot);
  m.d_dz0.reserve((size_t)n * wmax);
  m.d_dz1.reserve((size_t)n * wmax);
  const size_t nchunks = ((size_t)n + mlp_chunk() - 1) / mlp_chunk();
  m.d_part.reserve(nchunks * (size_t)wmax * (wmax + 1));
  m.nmax = n;
}

// Forward pass over n samples at x (device, [n][nin]); returns the output (device).
const float* forward(Mlp& m, const float* x, int n, hipStream_t s) {
  const int L = nlayers(m);
  launch_to_bf16(m.d_p.p, m.nparams, m.d_pb.p, s);
  const float* a = x;
  for (int l = 1; l <= L; ++l) {
    float* z = m.d_z.p + m.aoff[l];
    float* an = m.d_a.p + m.aoff[l];
    const int act = l < L ? 1 : 0;
    launch_linear_fwd(a, n, m.dims[l - 1], m.d_pb.p + m.woff[l - 1], m.d_p.p + m.boff[l - 1], m.dims[l], act,
                      z, an, s);
    a = an;
  }
  return m.d_z.p + m.aoff[L];
}

// Mean-squared-error gradient into d_g after forward(); returns the loss.
double backward(Mlp& m, const float* x, const float* t, int n, hipStream_t s) {
  const int L = nlayers(m);
  const float* y = m.d_z.p + m.aoff[L];
  const size_t mo = (size_t)n * m.dims[L];
  const int nb = (int)std::min<size_t>((mo + 255) / 256, 1024);
  m.d_loss.reserve(nb);
  float* dz = m.d_dz0.p;
  float* dzn = m.d_dz1.p;
  launch_mse(y, t, mo, 1.0f / (float)mo, dz, m.d_loss.p, nb, s);
  for (int l = L; l >= 1; --l) {
    const float* ain = l == 1 ? x : m.d_a.p + m.aoff[l - 1];
    launch_linear_bwd_w(dz, ain, n, m.dims[l - 1], m.dims[l], m.d_part.p, m.d_g.p + m.woff[l - 1],
                        m.d_g.p + m.boff[l - 1], s);
    if (l > 1) {
      launch_linear_bwd_data(dz, n, m.dims[l], m.d_pb.p + m.woff[l - 1], m.dims[l - 1],
                             m.d_z.p + m.aoff[l - 1], 1, dzn, s);
      std::swap(dz, dzn);
    }
  }
  std::vector<double> lp(nb);
  RP_HIP(hipMemcpyAsync(lp.data(), m.d_loss.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  RP_HIP(hipStreamSynchronize(s));
  double sum = 0.0;
  for (double v : lp) sum += v;
  return sum / (double)mo;
}

// Uniform in [-1, 1) from the raw 64-bit stream (independent of the library's
// distribution implementations).
double uniform_pm1(std::mt19937_64& g) { return (double)(g() >> 11) * (2.0 / 9007199254740992.0) - 1.0; }

}  // namespace

extern "C" {

int rp_mlp_open(rp_handle h, int32_t nin, int32_t width, int32_t depth, int32_t nout, uint64_t seed) {
  return guarded([&] {
    RP_CHECK(h, "rp_mlp_open: no handle");
    RP_CHECK(nin >= 1 && width >= 1 && depth >= 1 && nout >= 1, "rp_mlp_open: bad sizes");
    RP_CHECK(nin <= mlp_max_width() && width <= mlp_max_width() && nout <= mlp_max_width(),
             "rp_mlp_open: layers are at most 128 wide");
    RP_HIP(hipSetDevice(h->device));
    h->mlp = std::make_unique<Mlp>();
    Mlp& m = *h->mlp;
    m.dims.push_back(nin);
    for (int i = 0; i < depth; ++i) m.dims.push_back(width);
    m.dims.push_back(nout);
    size_t off = 0;
    for (int l = 0; l + 1 < (int)m.dims.size(); ++l) {
      m.woff.push_back(off);
      off += (size_t)m.dims[l + 1] * m.dims[l];
      m.boff.push_back(off);
      off += (size_t)m.dims[l + 1];
    }
    m.nparams = off;
    // Glorot-uniform weights, zero biases
    std::vector<float> p(off, 0.0f);
    std::mt19937_64 g(seed);
    for (int l = 0; l + 1 < (int)m.dims.size(); ++l) {
      const double lim = std::sqrt(6.0 / (double)(m.dims[l] + m.dims[l + 1]));
      for (size_t i = 0; i < (size_t)m.dims[l + 1] * m.dims[l]; ++i) p[m.woff[l] + i] = (float)(lim * uniform_pm1(g));
    }
    hipStream_t s = h->stream;
    m.d_p.upload(p, s);
    m.d_g.reserve(off);
    m.d_pb.reserve(off);
    std::vector<float> z(off, 0.0f);
    m.d_m1.upload(z, s);
    m.d_m2.upload(z, s);
    RP_HIP(hipStreamSynchronize(s));
  });
}

int rp_mlp_nparams(rp_handle h, int64_t* n) {
  return guarded([&] {
    RP_CHECK(n, "rp_mlp_nparams: null output");
    *n = (int64_t)mlp_of(h, "rp_mlp_nparams").nparams;
  });
}

int rp_mlp_get_params(rp_handle h, float* p) {
  return guarded([&] {
    Mlp& m = mlp_of(h, "rp_mlp_get_params");
    RP_CHECK(p, "rp_mlp_get_params: null output");
    RP_HIP(hipMemcpy(p, m.d_p.p, sizeof(float) * m.nparams, hipMemcpyDeviceToHost));
  });
}

int rp_mlp_set_params(rp_handle h, const float* p) {
  return guarded([&] {
    Mlp& m = mlp_of(h, "rp_mlp_set_params");
    RP_CHECK(p, "rp_mlp_set_params: null input");
    RP_HIP(hipMemcpy(m.d_p.p, p, sizeof(float) * m.nparams, hipMemcpyHostToDevice));
    std::vector<float> z(m.nparams, 0.0f);
    RP_HIP(hipMemcpy(m.d_m1.p, z.data(), sizeof(float) * m.nparams, hipMemcpyHostToDevice));
    RP_HIP(hipMemcpy(m.d_m2.p, z.data(), sizeof(float) * m.nparams, hipMemcpyHostToDevice));
    m.step = 0;
  });
}

int rp_mlp_forward(rp_handle h, int64_t n, const float* x, float* y) {
  return guarded([&] {
    Mlp& m = mlp_of(h, "rp_mlp_forward");
    RP_CHECK(n >= 0 && (n == 0 || (x && y)), "rp_mlp_forward: bad arguments");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int nin = m.dims.front(), nout = m.dims.back();
    const int64_t blk = 65536;
    reserve_batch(m, (int)std::min<int64_t>(n, blk));
    m.d_xb.reserve((size_t)std::min<int64_t>(n, blk) * nin);
    for (int64_t r0 = 0; r0 < n; r0 += blk) {
      const int nb = (int)std::min<int64_t>(blk, n - r0);
      RP_HIP(hipMemcpyAsync(m.d_xb.p, x + (size_t)r0 * nin, sizeof(float) * nb * nin, hipMemcpyHostToDevice, s));
      const float* dy = forward(m, m.d_xb.p, nb, s);
      RP_HIP(hipMemcpyAsync(y + (size_t)r0 * nout, dy, sizeof(float) * nb * nout, hipMemcpyDeviceToHost, s));
      RP_HIP(hipStreamSynchronize(s));
    }
  });
}

int rp_mlp_gradient(rp_handle h, int64_t n, const float* x, const float* t, float* grad, double* loss) {
  return guarded([&] {
    Mlp& m = mlp_of(h, "rp_mlp_gradient");
    RP_CHECK(n >= 1 && n <= (1 << 24) && x && t && grad, "rp_mlp_gradient: bad arguments");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int nin = m.dims.front(), nout = m.dims.back();
    reserve_batch(m, (int)n);
    m.d_xb.reserve((size_t)n * nin);
    m.d_tb.reserve((size_t)n * nout);
    RP_HIP(hipMemcpyAsync(m.d_xb.p, x, sizeof(float) * n * nin, hipMemcpyHostToDevice, s));
    RP_HIP(hipMemcpyAsync(m.d_tb.p, t, sizeof(float) * n * nout, hipMemcpyHostToDevice, s));
    forward(m, m.d_xb.p, (int)n, s);
    const double l = backward(m, m.d_xb.p, m.d_tb.p, (int)n, s);
    RP_HIP(hipMemcpy(grad, m.d_g.p, sizeof(float) * m.nparams, hipMemcpyDeviceToHost));
    if (loss) *loss = l;
  });
}

int rp_mlp_train(rp_handle h, int64_t n, const float* x, const float* t, int32_t steps, int32_t batch,
                 double lr, uint64_t seed, double* losses) {
  return guarded([&] {
    Mlp& m = mlp_of(h, "rp_mlp_train");
    RP_CHECK(n >= 1 && x && t && steps >= 0 && batch >= 1 && batch <= (1 << 20) && lr > 0.0,
             "rp_mlp_train: bad arguments");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int nin = m.dims.front(), nout = m.dims.back();
    m.d_x.upload(x, (size_t)n * nin, s);
    m.d_t.upload(t, (size_t)n * nout, s);
    reserve_batch(m, batch);
    m.d_xb.reserve((size_t)batch * nin);
    m.d_tb.reserve((size_t)batch * nout);
    m.d_idx.reserve(batch);
    std::mt19937_64 g(seed);
    std::vector<int> idx(batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int st = 0; st < steps; ++st) {
      for (int j = 0; j < batch; ++j) idx[j] = (int)(g() % (uint64_t)n);
      RP_HIP(hipMemcpyAsync(m.d_idx.p, idx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, s));
      launch_gather_rows(m.d_x.p, m.d_t.p, m.d_idx.p, batch, nin, nout, m.d_xb.p, m.d_tb.p, s);
      forward(m, m.d_xb.p, batch, s);
      const double l = backward(m, m.d_xb.p, m.d_tb.p, batch, s);  // synchronizes: idx may be refilled
      ++m.step;
      const double c1 = 1.0 - std::pow((double)b1, (double)m.step);
      const double c2 = 1.0 / (1.0 - std::pow((double)b2, (double)m.step));
      launch_adam(m.d_p.p, m.d_g.p, m.d_m1.p, m.d_m2.p, m.nparams, (float)(lr / c1), b1, b2, (float)c2, eps, s);
      if (losses) losses[st] = l;
    }
    RP_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
