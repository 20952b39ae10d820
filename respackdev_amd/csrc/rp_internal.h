// Internal declarations shared by the host-side sources of librespack_gpu:
// the error-reporting wrapper of the C ABI, the restated setup routines of the
// reference (k-mesh, G spheres, grids) and the handle behind rp_handle.
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

// Message of the last failed C-ABI call on this thread (rp_last_error).
extern thread_local std::string g_err;

// Runs f; a thrown exception becomes status 1 with its message in g_err.
template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return 1;
  } catch (...) {
    g_err = "unknown error";
    return 1;
  }
}

constexpr double kAu = 27.21151;  // chiqw/config.h:19

int algn235(int n);                                   // est_nwx2.f90:1-42
void gen_grid(int nt, double e, double* em);          // gen_grid.f90:1-39
// make_KG0 (chiqw_sub.F:773-812): {G : |q+G|^2 <= gcut} in the reference's order
std::vector<int> make_kg0(const double b[3][3], const double a[3][3], double gcut, const double* q);
void kcheck(double* k, int* rw);                      // chiqw_sub.F:668-716
void kcheck_trs(double* k, int* rw);                  // chiqw_sub.F:718-771
void search_list(const double* in, double* sq);       // chiqw_sub.F:76-95
int search_kq(int ntk, const double* sk0, const double* q, int ik, int* sg);  // chiqw_sub.F:1-75
void mkidx(int n1, int n2, int n3, const double* b, int* idx);  // m_tetrahedron_20170325.F:505-584
inline size_t roundup(size_t x, size_t m) { return (x + m - 1) / m * m; }
// b_i = 2 pi (a_j x a_k) / volume (chiqw.F:145-156); returns the volume
double reciprocal(const double a[3][3], double b[3][3]);
// FFT grid of the reference (chiqw.F:389-401 / calc_intW.F:176-190)
void fft_grid(const double a[3][3], const double b[3][3], double ecut_psi, int nw[3]);
// Ecut_for_psi = max |k+G|^2 + 1e-8 over the irreducible k lists (chiqw.F:313-352)
double ecut_for_psi(const double b[3][3], int nk_irr, const double* ski, const int* ngi, int ntg,
                    const int* kgi);

struct W3d;  // calc_int state (rp_w3d.cpp)
struct Mlp;  // band-tail closure network (rp_mlp_state.h)

// Launch parameters of the chi contraction: K split and (split-BF16 only)
// persistent workgroups.
struct ChiLaunch {
  int nsplit = 1;
  int nwg = 0;
  int span = 0;  // split-BF16: FP32 accumulation span in K (0: the default)
};

// One autotuned contraction shape and the launch chosen for it.
struct TuneEntry {
  int prec = 0, ng = 0, nen = 0, ndir = 1;
  size_t kpad = 0;
  ChiLaunch best, heuristic;
  double best_ms = 0, heuristic_ms = 0;
  int ncand = 0;
};

// The per-q state of a handle, parked while another q is current (batched
// runs keep one per q; swap_q exchanges it with the handle's).
struct QSlot {
  bool qready = false, q0 = false, have_A = false, have_w = false, have_chi = false;
  double q[3] = {0, 0, 0};
  int ng = 0, g0 = -1, npair = 0;
  std::vector<int> lg, ikq, shift, pocc, pvir, intra, pdplan;
  std::vector<double> qn;
  size_t lda = 0, kpad = 0;
  double ecut_eps_win = 0;
  int n_vir_win = 0;
  DevBuf<int> d_lg, d_ikq, d_shift, d_pocc, d_pvir, d_intra, d_pdplan;
  DevBuf<double> d_qn;
  DevBuf<double2> d_eps;  // eps^-1 of this q (a batched run's output)
  ChiLaunch launch;
  int q_span = 0;
};

// Work buffers of the stages (shared by all q of one launch sequence).
struct Scratch {
  DevBuf<float2> d_A, d_A3, d_wf;
  DevBuf<unsigned char> d_Ao;
  DevBuf<unsigned> d_sync;
  DevBuf<double2> d_w, d_part, d_chi, d_drude;
  DevBuf<double> d_wd;
  DevBuf<int> d_status, d_ipiv, d_gjwork;
};

// A batched configuration: every q prepared, its launch sequence captured in
// one HIP graph over buffers that only the graph uses.
struct Batch {
  std::vector<QSlot> slots;
  Scratch scratch;
  int prec = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

}  // namespace rp

// ------------------------------------------------------------------ context
using rp::DevBuf;

struct rp_ctx {
  std::unique_ptr<rp::W3d> w3d;  // calc_int (Wannier-basis V/W), when opened
  std::unique_ptr<rp::Mlp> mlp;  // band-tail closure network, when opened
  int device = 0;
  int ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // a batch graph's pair densities and weights (default priority)
  hipEvent_t ev[6] = {};
  hipEvent_t gev[6] = {};  // around each main contraction launch (<= 3 directions)
  hipEvent_t bev[3] = {};  // the batch graph's fork and joins between the two streams
  double gemm_seconds = 0;
  int gemm_launches = 0;
  // system
  bool open = false;
  double a[3][3], b[3][3], volume = 0;
  int nk_irr = 0, ntg = 0, ntb = 0, nsym = 0, nnp = 1;
  double ef = 0, delt = 0, dmna = 0, dmnr = 0, delta_ex = 0, ecut_psi = 0, ecut_eps = 0, emax = 0;
  // per-q window (rp_chiqw_set_window): chi cutoff and number of empty bands in the pair list
  double ecut_eps_win = 0;
  int n_vir_win = 0;
  int nw[3] = {0, 0, 0};
  int ntk = 0, nkb[3] = {0, 0, 0};
  std::vector<double> sk0;
  std::vector<int> numirr, numrot, trs, rwv, index_kpt, imt, ng0, kg0;
  int n_calc = 0, n_docc = 0, n_pocc = 0, n_uocc = 0, n_occ = 0, n_vir = 0, nen = 0, flg_crpa = 0;
  std::vector<int> win, tindx;
  std::vector<double> eave, em;
  std::vector<double> eband;   // [band][k]
  std::vector<double> pband;   // complex [band][k]
  // device: system
  DevBuf<double2> d_coef;
  DevBuf<int> d_kg0, d_ng0, d_imt, d_index;
  DevBuf<float2> d_ur;
  DevBuf<double> d_eband, d_bvec, d_sk0;
  DevBuf<double2> d_pband, d_em;
  DevBuf<double2> d_tw;  // line-DFT twiddle table
  // per q
  bool qready = false, q0 = false, have_A = false, have_w = false, have_chi = false;
  double q[3] = {0, 0, 0};
  int ng = 0, g0 = -1, npair = 0;
  std::vector<int> lg, ikq, shift, pocc, pvir;
  std::vector<int> intra;   // q = 0, fRPA: partially occupied bands paired with themselves
  std::vector<int> pdplan;  // pruned pair-density transform of the chi sphere
  std::vector<double> qn;
  size_t lda = 0, kpad = 0;
  double prep_seconds = 0;
  DevBuf<int> d_lg, d_ikq, d_shift, d_pocc, d_pvir, d_intra, d_pdplan, d_status, d_ipiv, d_gjwork;
  DevBuf<double2> d_drude;  // per intraband band and direction: sum_k Im x(k) p(k)^2
  DevBuf<float2> d_A, d_A3;
  DevBuf<unsigned char> d_Ao;  // split-bf16 octets of the left operand
  DevBuf<float2> d_wf;         // single-precision copy of the weights
  DevBuf<unsigned> d_sync;     // per-XCD rendezvous counters of the split-BF16 contraction
  DevBuf<double2> d_w, d_part, d_chi, d_eps;
  DevBuf<double> d_qn, d_wd;
  // launch sequence being captured into a graph: no host syncs or status
  // read-back, buffers must already be large enough
  bool capturing = false;
  int autotune = 0;              // time the contraction's launch candidates per shape
  int span = 0;                  // split-BF16 FP32 accumulation span (K) set by the caller, 0: automatic
  int q_span = 0;                // the span chosen for the current q (0: the default)
  double last_kappa = 0;         // |eps^-1|_1 of the last q checked by adapt_span
  DevBuf<double> d_kappa;
  std::vector<rp::TuneEntry> tune;
  std::unique_ptr<rp::Batch> batch;

  rp_ctx();   // defined where rp::W3d and rp::Mlp are complete (rp_w3d.cpp)
  ~rp_ctx();
  void release_events() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : gev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : bev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
  }
  int nbatch() const { return q0 ? 3 * nen : nen; }
};

namespace rp {
// Stage bodies (rp_capi.cpp): enqueue only, on h->stream.
void check_status(rp_ctx* h, const char* what);
// Throws the error a non-zero status word of the kernels stands for.
void throw_status(int st, const std::string& what);
// Looks up this device's address of the fixed scratch of rp_dyson_noalloc
// (rp_noalloc.hip); called by rp_init with the device current.
void noalloc_init(int device);
void issue_pairdens(rp_ctx* h);
void issue_weights(rp_ctx* h);
// The handle's two streams exchanged for a scope (restored however it ends):
// the stage bodies issue on h->stream.
struct OnStream2 {
  rp_ctx* h;
  explicit OnStream2(rp_ctx* h_) : h(h_) { std::swap(h->stream, h->stream2); }
  ~OnStream2() { std::swap(h->stream, h->stream2); }
};
void issue_operands(rp_ctx* h, int prec);
void issue_accumulate(rp_ctx* h, int prec, const ChiLaunch& c, bool events,
                      bool operands_ready = false);
void issue_dyson(rp_ctx* h);
ChiLaunch chi_launch_heuristic(const rp_ctx* h, int prec);
// The launch for the current q: the tuned one for its shape, tuning it now
// when autotuning is on and the inputs are ready, else the heuristic.
ChiLaunch chi_launch(rp_ctx* h, int prec);
// The same without the span of the caller or of the q.
ChiLaunch chi_launch_tuned(rp_ctx* h, int prec);
// max over the batch of |eps^-1|_1 of the current q (after the Dyson step)
double epsinv_kappa(rp_ctx* h);
// The split-BF16 span the current q's eps^-1 calls for (bf16x3_span_for_kappa)
// when it is shorter than span_used (0: the default), else 0; automatic spans
// only (no caller's span), split-BF16 only.
int adapt_span(rp_ctx* h, int prec, int span_used);
void swap_q(rp_ctx* h, QSlot& s);
void swap_scratch(rp_ctx* h, Scratch& s);
}  // namespace rp
