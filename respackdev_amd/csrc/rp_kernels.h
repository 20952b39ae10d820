// Host-side launchers of the gfx950 kernels (internal to librespack_gpu).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/respack_gpu.h"

namespace rp {

// rp_pairdens.hip
bool dft_size_supported(int n);
// LDS one workgroup may use (the bound of the LDS-resident transforms)
size_t lds_budget_bytes();
size_t twiddle_entries();
void launch_twiddles(double2* tab, hipStream_t s);
// Elements of FP64 scratch the band grids of ngrids (band, k) need when the
// grid is too large for the LDS (0 when it fits): batches of at most 1 GiB.
size_t band_grids_scratch_elems(int nx, int ny, int nz, size_t ngrids);
void launch_band_grids(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                       int nband, int nx, int ny, int nz, const double2* twtab, float2* ur,
                       double2* scratch, size_t scratch_elems, hipStream_t s);
// In-place 3D DFT (z, y, x) of nb grids in HBM as line passes; sign +1: e^{+i}
// unscaled (fft3_bw), -1: e^{-i} unscaled.
void launch_line_dft3(double2* g, long long nb, int nx, int ny, int nz, int sign,
                      const double2* tab, hipStream_t s);
// LDS of the pruned forward transform for nzz kept z planes
size_t pruned_transform_lds_bytes(int nx, int ny, int nz, int nzz);
// Plan of the pruned forward transform for a chi sphere lg (see rp_pairdens.hip);
// plan[0] is the number of kept z planes.
std::vector<int> pairdens_plan(const std::vector<int>& lg, int nx, int ny, int nz);
// Pair densities of every (pair, k) for the points lg of a G sphere (the plan
// of pairdens_plan): only the transform outputs the sphere needs are computed.
void launch_pair_densities_pruned(const float2* ur, int ntk, const int* pocc, const int* pvir,
                                  const int* ikq, const int* shift, int nx, int ny, int nz,
                                  const int* plan, int nzz, int ng, size_t lda, int nK,
                                  const double2* twtab, float2* A, hipStream_t s);
// calc_int: P[q][p][r] = sum_k wk[k] conj(u_iw,k+q(r) e^{-i 2pi r.dG/N}) u_jw,k(r)
// (FP64) for the q of ikq[nq][ntk] / shift[nq][ntk][3]; full: every pair
// p = iw + nwf*jw, else the diagonal p = w.  wk may be null (1).  kqs (may be
// null) packs the same per (q, k): pair_rsum_pack.
void launch_pair_rsum(const float2* ur, int ntk, int nwf, bool full, int nx, int ny, int nz,
                      const int* ikq, const int* shift, const unsigned* kqs, const double2* wk,
                      int nq, double2* P, hipStream_t s);
// k+q index (< 2^20) and umklapp vector (components -8..7) in one word, or
// false when they do not fit
inline bool pair_rsum_pack(int ikq, const int* d, unsigned* out) {
  if (ikq < 0 || ikq >= (1 << 20)) return false;
  unsigned v = (unsigned)ikq;
  for (int c = 0; c < 3; ++c) {
    if (d[c] < -8 || d[c] > 7) return false;
    v |= (unsigned)(d[c] + 8) << (20 + 4 * c);
  }
  *out = v;
  return true;
}
// rho[rho_off[q] + p*ngpsi[q] + i] from P of q = q_offset .. q_offset+nq-1:
// pruned forward transform read at -G of the psi sphere (plans at plan_off),
// / |q+G| (qg at qg_off), / NTK, zero at g0s[q].
void launch_rho_from_grid(const double2* P, int npair, int nx, int ny, int nz, const int* plan,
                          const long long* plan_off, int max_nzz, const double* qg,
                          const long long* qg_off, const int* g0s, const int* ngpsi,
                          const long long* rho_off, int q_offset, int nq, int ntk,
                          const double2* twtab, double2* rho, hipStream_t s);
// The same from grids transformed over HBM (grids beyond the LDS): P is
// transformed in place, then the psi-sphere points (lg: 3 ints per point at
// the offsets off of qg) are gathered.
void launch_rho_from_grid_hbm(double2* P, int npair, int nx, int ny, int nz, const int* lg,
                              const double* qg, const long long* off, const int* g0s,
                              const int* ngpsi, const long long* rho_off, int q_offset, int nq,
                              int ntk, const double2* twtab, double2* rho, hipStream_t s);
void launch_q0_heads(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                     const int* pocc, const int* pvir, const double* eband, const double* bvec,
                     double delt, int g0, int ng, size_t lda, size_t rows, int nK, const float2* A,
                     float2* A3, hipStream_t s);

// rp_tetra.hip
size_t tetra_lds_bytes(int nk3, int nf);
int tetra_freqs_per_block(int nk3, int nen);
void launch_tetra_weights(int ntk, int n1, int n2, int n3, const int* imt, const int* index_kpt,
                          const int* pocc, const int* pvir, const int* ikq, const double* eband,
                          const double2* pband, double ef, double delt, double dmna, double dmnr,
                          int nen, const double2* em, size_t ldd, int npair, int cuts, int* status,
                          double2* dout, hipStream_t s);
// q = 0 Drude plasma frequency per direction from the intraband bands -> wd[3]
void launch_drude(const double2* coef, const int* kg, const int* ng0, int ntk, int ntg,
                  const int* bands, int nbands, const double* sk0, const double* bvec,
                  const double* eband, double ef, double delt, double dmna, double dmnr, int n1,
                  int n2, int n3, const int* imt, const int* index_kpt, double volume,
                  int* status, double2* part, double* wd, hipStream_t s);

// rp_chi_gemm.hip
int chi_tile();
int chi_kchunk();
void launch_chi_gemm(int prec, const float2* A, size_t lda, const double2* dw, size_t ldd,
                     size_t kpad, int ng, int nen, int nsplit, double2* partial, hipStream_t s);
void launch_chi_reduce(const double2* partial, size_t n, int nsplit, double inv_vol, double2* chi,
                       hipStream_t s);
// zero `bytes` (a multiple of 4) at p (16-byte aligned) with a fill kernel
void launch_zero(void* p, size_t bytes, hipStream_t s);

// rp_chi_bf16.hip
size_t bf16x3_lda_multiple();
void launch_presplit(const float2* A, size_t lda, size_t kpad, void* Ao, hipStream_t s);
void launch_weights_to_f32(const double2* dw, size_t n, float2* df, hipStream_t s);
int bf16x3_nsplit(int ng, int nen, size_t kpad, int nwg);
int bf16x3_workgroups(int ncu);
size_t bf16x3_sync_words();
// default FP32 accumulation span (K) of the split-BF16 contraction
int bf16x3_default_span();
// the span that keeps eps^-1 within the mixed path's bar for a q whose eps^-1
// has 1-norm kappa (the amplification of chi's error)
int bf16x3_span_for_kappa(double kappa);
void launch_chi_bf16x3(const void* Ao, size_t lda, const float2* df, size_t ldd, size_t kpad, int ng,
                       int nen, int nsplit, int nwg, int span, double2* partial, unsigned* sync,
                       hipStream_t s);

// rp_dyson.hip
void launch_eps_build(const double2* chi, const double* qn, int n, int nbatch, int nen, int q0,
                      int g0, const double* wd, const double2* em, double delt, double2* eps,
                      hipStream_t s);
// kappa[b] = the 1-norm of matrix b of the batch (max column sum of |a|)
void launch_norm1(const double2* a, int n, int nbatch, double* kappa, hipStream_t s);
size_t gauss_jordan_work_ints(int nbatch);
void launch_gauss_jordan(double2* a, int n, int nbatch, int* ipiv, int* work, int* status,
                         hipStream_t s);

// rp_w3d.hip
int w3d_max_wannier();
// rows of a column of eps^-1 in HBM: NG_eps rounded up to 16, so that every
// 16-row block of the contractions starts on a 128-byte line
__host__ __device__ inline int w3d_eps_ld(int nge) { return (nge + 15) / 16 * 16; }
void launch_w_func(const double2* rho, const long long* rho_off, const int4* qinfo, int ntq,
                   int nwf, double2* func, double2* shell, hipStream_t s);
// q per group of launch_w_contract (q sharing one irreducible eps^-1), and
// the vectors of a group (nwf per q)
int w_contract_members(int nwf);
int w_contract_vectors(int nwf);
// groups g = 0 .. ngroups-1 of the q gmem[goff[g] .. goff[g+1]), each sharing
// one irreducible eps^-1 and at most w_contract_members(nwf) long; xw: scratch
// of w_contract_vectors(nwf) * (sum over q of NG_eps) complex values
void launch_w_contract(const float2* eps, const long long* eps_off, const double2* rho,
                       const long long* rho_off, const int4* qinfo, const int* minv,
                       const double2* pf1, const double2* pf2, const long long* map_off,
                       const double2* shell, const int* goff, const int* gmem, int ngroups,
                       int nwf, int nen, int max_nge, double2* xw, double2* funcw,
                       hipStream_t s);
// xg: scratch of j_contract_scratch(sum over q of NG_eps, npair) complex values
void launch_j_contract(const float2* eps, const long long* eps_off, const double2* rho,
                       const long long* rho_off, const int4* qinfo, const int* minv,
                       const double2* pf1, const double2* pf2, const long long* map_off,
                       double2* xg, const double2* shell, int ntq, int npair, int nen,
                       double2* funcw, hipStream_t s);
size_t j_contract_scratch(size_t sum_nge, int npair);
void launch_j_func(const double2* rho, const long long* rho_off, const int4* qinfo, int ntq,
                   int npair, double2* func, double2* shell, hipStream_t s);
void launch_j_sum(const double2* func, const double2* funcw, int ntq, int nwf, int nen, int r0,
                  double volume, double chead, const double2* cheadw, double2* xmat,
                  double2* jmat, size_t wstride, hipStream_t s);
void launch_w_fourier(const double2* func, const double2* funcw, const double* sq, int ntq, int nwf,
                      int nen, const int na[3], double volume, double chead, const double2* cheadw,
                      double2* vmat, double2* wmat, hipStream_t s);

// band-tail closure network (rp_mlp.hip)
int mlp_max_width();
int mlp_chunk();
void launch_to_bf16(const float* src, size_t n, unsigned short* dst, hipStream_t s);
void launch_linear_fwd(const float* A, int n, int nin, const unsigned short* W, const float* b, int nout,
                       int act, float* Z, float* Aout, hipStream_t s);
void launch_linear_bwd_data(const float* dZ, int n, int nout, const unsigned short* W, int nin,
                            const float* Zprev, int act, float* dA, hipStream_t s);
void launch_linear_bwd_w(const float* dZ, const float* A, int n, int nin, int nout, float* part, float* gW,
                         float* gb, hipStream_t s);
void launch_mse(const float* Y, const float* T, size_t m, float scale, float* dY, double* lossp, int nblocks,
                hipStream_t s);
void launch_adam(float* p, const float* g, float* m1, float* m2, size_t n, float lr_t, float b1, float b2,
                 float c2, float eps, hipStream_t s);
void launch_gather_rows(const float* x, const float* t, const int* idx, int nb, int nin, int nout, float* xb,
                        float* tb, hipStream_t s);

}  // namespace rp
