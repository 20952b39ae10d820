// C ABI of the calc_int (calc_w3d) path: Wannier-basis bare and screened
// Coulomb integrals V_ij(R) and W_ij(R,w) (calc_intW.F), on gfx950.
//
// Host side: the reference's setup (k and q unfolding, G lists, the make_eps
// rotation maps and phases, the Hybertsen-Louie heads) restated so that every
// index the output depends on is the reference's.  Device side: the Wannier
// band grids, the pair densities rho_w(q+G), func / funcw and the Fourier sums
// (rp_w3d.hip, rp_pairdens.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "rp_internal.h"
#include "rp_mlp_state.h"

namespace rp {

struct W3d {
  bool open = false, have_wan = false, have_q = false, have_funcs = false;
  double a[3][3], b[3][3], volume = 0, ecut_psi = 0, ecut_eps = 0;
  int nw[3] = {0, 0, 0}, nkb[3] = {0, 0, 0}, na[3] = {0, 0, 0};
  int nk_irr = 0, ntg = 0, ntk = 0, nen = 0, nwf = 0, nsym = 0, nnp = 1;
  std::vector<int> rg, pg;
  std::vector<double> rginv, ski, sk0, em;
  std::vector<int> numirr, numrot, trs, rwv, ng0, kg0;
  // q mesh (calc_intW.F:420-545)
  int nq_irr = 0, ntq = 0, iqgm = -1, nog0 = -1;
  std::vector<double> sq;
  std::vector<int> numirrq, numrotq, trsq, rwq, nge, ngpsi;
  std::vector<std::vector<int>> lg0;
  // eps^-1 of every irreducible q lives in HBM only (d_eps, complex(4) as the
  // reference's epsirr, (ig,jg,ie) at eps_off[iq], columns padded to
  // w3d_eps_ld(NG_eps) rows); the host keeps which q
  // are set and the Gamma head element eps^-1_00(w) for the Hybertsen-Louie term
  std::vector<char> eps_have;
  std::vector<float2> eps_head;
  std::vector<long long> rho_off, map_off, eps_off;
  std::vector<int4> qinfo;
  // the q grouped by irreducible eps^-1 for the W contraction
  std::vector<int> wgoff, wgmem;
  int maps_nwf = 0;  // NWF the device copies of the maps and groups were made for (0: none)
  std::vector<int> minv;
  std::vector<double2> pf1, pf2;
  double chead = 0;
  std::vector<double2> cheadw;
  // device
  DevBuf<double2> d_coef, d_tw;
  DevBuf<int> d_kg0, d_ng0;
  DevBuf<float2> d_ur, d_eps;
  DevBuf<long long> d_eps_off, d_rho_off, d_map_off;
  DevBuf<int4> d_qinfo;
  DevBuf<int> d_minv, d_wgoff, d_wgmem;
  DevBuf<double2> d_wx;  // X vectors of the W contraction's q groups
  DevBuf<double> d_sq;
  // per-q geometry of the pair densities, every q at once (built on first use
  // after rp_w3d_set_q): |q+G|, k+q index and umklapp per k, pruned-transform plan
  bool geo_ready = false;
  std::vector<long long> geo_qg_off, geo_plan_off;
  std::vector<int> geo_g0, geo_q0, geo_nzz;
  int geo_max_nzz = 0;
  DevBuf<double> d_geo_qg;
  DevBuf<int> d_geo_ikq, d_geo_shift, d_geo_plan, d_geo_g0, d_ngpsi, d_geo_lg;
  DevBuf<unsigned> d_geo_kqs;  // ikq and shift packed (pair_rsum_pack; empty if they do not fit)
  DevBuf<long long> d_geo_qg_off, d_geo_plan_off;
  // summed pair-density grids P[q][p][r] of a chunk of q (launch_pair_rsum)
  DevBuf<double2> d_P;
  DevBuf<double2> d_rho, d_pf1, d_pf2, d_func, d_shell, d_funcw, d_cheadw, d_vmat, d_wmat;
  // device seconds of the last rp_w3d_funcs (rho, func, W contraction, total)
  // and rp_j3d_run (rho of every R, func and sums, J contraction, total),
  // from events on the handle's stream; ev grows to 2 per J launch
  double stage_s[8] = {};
  std::vector<hipEvent_t> ev;
  void events(size_t n) {
    while (ev.size() < n) {
      hipEvent_t e;
      RP_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
  }
  ~W3d() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  }
};

// seconds between events a and b (recorded, the stream synchronised)
double event_s(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  RP_HIP(hipEventElapsedTime(&ms, a, b));
  return 1e-3 * ms;
}

}  // namespace rp

using namespace rp;

rp_ctx::rp_ctx() = default;
rp_ctx::~rp_ctx() { release_events(); }

namespace {

W3d& w3d_of(rp_handle h, const char* what) {
  RP_CHECK(h, std::string(what) + ": null handle");
  RP_CHECK(h->w3d && h->w3d->open, std::string(what) + ": W3D system not open (rp_w3d_open)");
  return *h->w3d;
}

// make_LG0 (calc_int_sub.F:447-504): the eps sphere, then the shell
// Gcut_eps < |q+G|^2 <= Gcut_psi, each in make_KG0's loop order.
std::vector<int> make_lg0(const W3d& w, const double* q, int* n_eps) {
  std::vector<int> e = make_kg0(w.b, w.a, w.ecut_eps, q);
  std::vector<int> all = make_kg0(w.b, w.a, w.ecut_psi, q);
  *n_eps = (int)(e.size() / 3);
  std::vector<int> out = e;
  for (size_t i = 0; i < all.size(); i += 3) {
    double v[3];
    for (int d = 0; d < 3; ++d)
      v[d] = (q[0] + (double)all[i]) * w.b[0][d] + (q[1] + (double)all[i + 1]) * w.b[1][d] +
             (q[2] + (double)all[i + 2]) * w.b[2][d];
    const double s = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    if (s > w.ecut_eps && s <= w.ecut_psi) out.insert(out.end(), {all[i], all[i + 1], all[i + 2]});
  }
  return out;
}

// Hybertsen-Louie heads (calc_intW.F:782-795, calc_intJ.F:883-904)
void compute_heads(W3d& w) {
  const double pi = std::acos(-1.0);
  const double qsz = std::pow(6.0 * pi * pi / (double)w.ntq / w.volume, 1.0 / 3.0);
  w.chead = (2.0 / pi) * qsz;
  w.cheadw.assign(w.nen, make_double2(0, 0));
  for (int ie = 0; ie < w.nen; ++ie) {
    const float2 e = w.eps_head[ie];
    w.cheadw[ie] = make_double2((2.0 / pi) * (double)e.x * qsz, (2.0 / pi) * (double)e.y * qsz);
  }
}

void check_ready(const W3d& w, const char* what) {
  RP_CHECK(w.have_wan && w.have_q, std::string(what) + ": Wannier functions or q list missing");
  for (int iq = 0; iq < w.nq_irr; ++iq)
    RP_CHECK(w.eps_have[iq], std::string(what) + ": eps^-1 of irreducible q " +
                                        std::to_string(iq + 1) + " missing");
}

// the per-q rotation maps (eps^-1 itself is already in HBM), uploaded once per
// q list and NWF
void upload_eps_maps(W3d& w, hipStream_t s) {
  if (w.maps_nwf == w.nwf) return;
  w.d_eps_off.upload(w.eps_off, s);
  w.d_map_off.upload(w.map_off, s);
  w.d_qinfo.upload(w.qinfo, s);
  w.d_minv.upload(w.minv, s);
  // groups of at most w_contract_members q sharing one irreducible eps^-1, so
  // that the W contraction reads each eps^-1 once per group
  const size_t mx = (size_t)w_contract_members(std::max(1, w.nwf));
  std::vector<std::vector<int>> by(w.nq_irr);
  for (int q = 0; q < w.ntq; ++q) by[w.qinfo[q].z].push_back(q);
  w.wgoff.assign(1, 0);
  w.wgmem.clear();
  for (const auto& v : by)
    for (size_t i = 0; i < v.size(); i += mx) {
      for (size_t j = i; j < std::min(v.size(), i + mx); ++j) w.wgmem.push_back(v[j]);
      w.wgoff.push_back((int)w.wgmem.size());
    }
  w.d_wgoff.upload(w.wgoff, s);
  w.d_wgmem.upload(w.wgmem, s);
  w.d_pf1.upload(w.pf1, s);
  w.d_pf2.upload(w.pf2, s);
  w.maps_nwf = w.nwf;
}

// Per q: k+q map, umklapp shifts, |q+G| and the G = 0 slot at q = 0.
struct QGeom {
  bool q0;
  int g0;
  std::vector<int> ikq, shift;
  std::vector<double> qg;
};
QGeom q_geometry(const W3d& w, int iq) {
  QGeom r;
  const double* q = &w.sq[3 * (size_t)iq];
  r.q0 = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;  // calc_intW.F:650,710
  r.ikq.assign(w.ntk, 0);
  r.shift.assign(3 * (size_t)w.ntk, 0);
  for (int k = 0; k < w.ntk; ++k) {
    if (r.q0) { r.ikq[k] = k; continue; }
    int sg[3];
    const int kq = search_kq(w.ntk, w.sk0.data(), q, k, sg);
    RP_CHECK(kq >= 0, "k+q is not on the k mesh");
    r.ikq[k] = kq;
    for (int d = 0; d < 3; ++d) r.shift[3 * k + d] = sg[d];
  }
  const int np = w.ngpsi[iq];
  r.qg.assign(np, 0.0);
  r.g0 = -1;
  for (int g = 0; g < np; ++g) {
    double v[3];
    for (int d = 0; d < 3; ++d)
      v[d] = (q[0] + (double)w.lg0[iq][3 * g]) * w.b[0][d] +
             (q[1] + (double)w.lg0[iq][3 * g + 1]) * w.b[1][d] +
             (q[2] + (double)w.lg0[iq][3 * g + 2]) * w.b[2][d];
    r.qg[g] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (r.q0 && w.lg0[iq][3 * g] == 0 && w.lg0[iq][3 * g + 1] == 0 && w.lg0[iq][3 * g + 2] == 0)
      r.g0 = g;
  }
  return r;
}

// Every q's pair-density inputs in HBM, so that the loops over q only launch.
void ensure_geometry(W3d& w, hipStream_t s) {
  if (w.geo_ready) return;
  std::vector<double> qg;
  std::vector<int> ikq, shift, plan, lgs;
  w.geo_qg_off.assign(w.ntq, 0);
  w.geo_plan_off.assign(w.ntq, 0);
  w.geo_g0.assign(w.ntq, -1);
  w.geo_q0.assign(w.ntq, 0);
  w.geo_nzz.assign(w.ntq, 0);
  for (int iq = 0; iq < w.ntq; ++iq) {
    const QGeom geo = q_geometry(w, iq);
    const int np = w.ngpsi[iq];
    w.geo_qg_off[iq] = (long long)qg.size();
    qg.insert(qg.end(), geo.qg.begin(), geo.qg.end());
    lgs.insert(lgs.end(), w.lg0[iq].begin(), w.lg0[iq].begin() + 3 * (size_t)np);
    ikq.insert(ikq.end(), geo.ikq.begin(), geo.ikq.end());
    shift.insert(shift.end(), geo.shift.begin(), geo.shift.end());
    const std::vector<int> pl = pairdens_plan(
        std::vector<int>(w.lg0[iq].begin(), w.lg0[iq].begin() + 3 * (size_t)np), w.nw[0], w.nw[1],
        w.nw[2]);
    w.geo_plan_off[iq] = (long long)plan.size();
    w.geo_nzz[iq] = pl[0];
    plan.insert(plan.end(), pl.begin(), pl.end());
    w.geo_g0[iq] = geo.g0;
    w.geo_q0[iq] = geo.q0 ? 1 : 0;
  }
  w.geo_max_nzz = *std::max_element(w.geo_nzz.begin(), w.geo_nzz.end());
  w.d_geo_qg.upload(qg, s);
  w.d_geo_lg.upload(lgs, s);
  w.d_geo_ikq.upload(ikq, s);
  w.d_geo_shift.upload(shift, s);
  std::vector<unsigned> kqs(ikq.size());
  bool packed = true;
  for (size_t i = 0; i < ikq.size() && packed; ++i) packed = pair_rsum_pack(ikq[i], &shift[3 * i], &kqs[i]);
  if (packed) w.d_geo_kqs.upload(kqs, s);
  else w.d_geo_kqs.release();
  w.d_geo_plan.upload(plan, s);
  w.d_geo_qg_off.upload(w.geo_qg_off, s);
  w.d_geo_plan_off.upload(w.geo_plan_off, s);
  w.d_geo_g0.upload(w.geo_g0, s);
  w.d_ngpsi.upload(w.ngpsi, s);
  RP_HIP(hipStreamSynchronize(s));  // the host vectors end here
  w.geo_ready = true;
}

// Bound on the summed pair-density grids held at once (q chunks of launch_pair_rsum).
constexpr size_t kPBudgetBytes = size_t(1) << 30;

// rho of every q (calc_intW.F:647-767 with wk null, calc_intJ.F:618-720 with
// wk = e^{-i 2pi k.R}): the products summed over k on the grid, then one
// pruned transform per (q, pair), in q chunks of at most kPBudgetBytes.
void rho_all_q(W3d& w, bool full, const double2* wk, const long long* d_rho_off, double2* rho,
               hipStream_t s) {
  const int npair = full ? w.nwf * w.nwf : w.nwf;
  const size_t nr = (size_t)w.nw[0] * w.nw[1] * w.nw[2];
  const size_t per_q = (size_t)npair * nr * sizeof(double2);
  const int qchunk = (int)std::max<size_t>(1, std::min<size_t>(w.ntq, kPBudgetBytes / per_q));
  w.d_P.reserve((size_t)qchunk * npair * nr);
  for (int q0 = 0; q0 < w.ntq; q0 += qchunk) {
    const int nq = std::min(qchunk, w.ntq - q0);
    launch_pair_rsum(w.d_ur.p, w.ntk, w.nwf, full, w.nw[0], w.nw[1], w.nw[2],
                     w.d_geo_ikq.p + (size_t)q0 * w.ntk, w.d_geo_shift.p + (size_t)3 * q0 * w.ntk,
                     w.d_geo_kqs.n ? w.d_geo_kqs.p + (size_t)q0 * w.ntk : nullptr, wk, nq,
                     w.d_P.p, s);
    if (pruned_transform_lds_bytes(w.nw[0], w.nw[1], w.nw[2], w.geo_max_nzz) <= lds_budget_bytes())
      launch_rho_from_grid(w.d_P.p, npair, w.nw[0], w.nw[1], w.nw[2], w.d_geo_plan.p,
                           w.d_geo_plan_off.p, w.geo_max_nzz, w.d_geo_qg.p, w.d_geo_qg_off.p,
                           w.d_geo_g0.p, w.d_ngpsi.p, d_rho_off, q0, nq, w.ntk, w.d_tw.p, rho, s);
    else  // the psi sphere's z planes exceed the LDS: transform over HBM, then gather
      launch_rho_from_grid_hbm(w.d_P.p, npair, w.nw[0], w.nw[1], w.nw[2], w.d_geo_lg.p,
                               w.d_geo_qg.p, w.d_geo_qg_off.p, w.d_geo_g0.p, w.d_ngpsi.p, d_rho_off,
                               q0, nq, w.ntk, w.d_tw.p, rho, s);
  }
}

}  // namespace

extern "C" {

int rp_w3d_open(rp_handle h, const double* lattice, int32_t nsym, int32_t nnp, const int32_t* rg,
                const int32_t* pg, int32_t nk_irr, const double* ski, const int32_t* ngi,
                int32_t ntg, const int32_t* kgi, double ecut_for_eps, int32_t nen, const double* em) {
  return guarded([&] {
    RP_CHECK(h && lattice && rg && pg && ski && ngi && kgi && em, "rp_w3d_open: null argument");
    RP_HIP(hipSetDevice(h->device));
    h->w3d.reset(new W3d());
    W3d& w = *h->w3d;
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) w.a[i][d] = lattice[3 * i + d];
    w.volume = reciprocal(w.a, w.b);  // calc_intW.F:27-35
    w.nk_irr = nk_irr;
    w.ntg = ntg;
    w.nsym = nsym;
    w.nnp = nnp;
    w.nen = nen;
    RP_CHECK(nen >= 1, "rp_w3d_open: nen must be >= 1");
    w.em.assign(em, em + 2 * (size_t)nen);
    w.rg.assign(rg, rg + 9 * (size_t)nsym);
    w.pg.assign(pg, pg + 3 * (size_t)nsym);
    w.ski.assign(ski, ski + 3 * (size_t)nk_irr);
    // Ecut_for_psi (calc_intW.F:124-145), FFT grid (calc_intW.F:176-190)
    w.ecut_psi = ecut_for_psi(w.b, nk_irr, ski, ngi, ntg, kgi);
    w.ecut_eps = ecut_for_eps;
    RP_CHECK(ecut_for_eps > 0.0, "rp_w3d_open: Ecut_for_eps (dat.chi_cutoff) must be positive");
    fft_grid(w.a, w.b, w.ecut_psi, w.nw);
    for (int d = 0; d < 3; ++d)
      RP_CHECK(dft_size_supported(w.nw[d]), "FFT grid dimension " + std::to_string(w.nw[d]) +
                                                " is not supported (2,3,5-smooth sizes up to 32)");
    // rginv (invmat of rg; integer matrices of det +-1)
    w.rginv.assign(9 * (size_t)nsym, 0.0);
    for (int op = 0; op < nsym; ++op) {
      const int* r = rg + 9 * op;
      auto R = [&](int i, int j) { return (double)r[i + 3 * j]; };
      const double det = R(0, 0) * (R(1, 1) * R(2, 2) - R(1, 2) * R(2, 1)) -
                         R(0, 1) * (R(1, 0) * R(2, 2) - R(1, 2) * R(2, 0)) +
                         R(0, 2) * (R(1, 0) * R(2, 1) - R(1, 1) * R(2, 0));
      RP_CHECK(det != 0.0, "singular symmetry operation");
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          int j1 = (j + 1) % 3, j2 = (j + 2) % 3, i1 = (i + 1) % 3, i2 = (i + 2) % 3;
          w.rginv[9 * op + i + 3 * j] = (R(j1, i1) * R(j2, i2) - R(j1, i2) * R(j2, i1)) / det;
        }
    }
    // k unfolding (calc_intW.F:208-262): symmetry, then time reversal
    for (int ik = 0; ik < nk_irr; ++ik)
      for (int op = 0; op < nsym; ++op) {
        const int* r = rg + 9 * op;
        for (int pass = 0; pass < 2; ++pass) {
          double k[3];
          int rw[3] = {0, 0, 0};
          for (int i = 0; i < 3; ++i)
            k[i] = (double)r[i] * ski[3 * ik] + (double)r[i + 3] * ski[3 * ik + 1] +
                   (double)r[i + 6] * ski[3 * ik + 2];
          double sgn = 1.0;
          if (pass == 0) kcheck(k, rw);
          else { kcheck_trs(k, rw); sgn = -1.0; }
          bool found = false;
          for (size_t j = 0; j < w.numirr.size(); ++j)
            if (std::fabs(w.sk0[3 * j] - sgn * k[0]) < 1e-4 &&
                std::fabs(w.sk0[3 * j + 1] - sgn * k[1]) < 1e-4 &&
                std::fabs(w.sk0[3 * j + 2] - sgn * k[2]) < 1e-4) { found = true; break; }
          if (found) continue;
          for (int i = 0; i < 3; ++i) { w.sk0.push_back(sgn * k[i]); w.rwv.push_back(rw[i]); }
          w.numirr.push_back(ik);
          w.numrot.push_back(op);
          w.trs.push_back(pass == 0 ? 1 : -1);
        }
      }
    w.ntk = (int)w.numirr.size();
    // est_nkbi (est_NTK.f90:1-39); Na = nkb/2 (calc_intW.F:275)
    for (int d = 0; d < 3; ++d) {
      double x = 1.0;
      for (int i = 0; i < w.ntk; ++i) {
        const double v = std::fabs(w.sk0[3 * i + d]);
        if (v < 1.0e-7) continue;
        if (v < x) x = v;
      }
      w.nkb[d] = (int)std::lround(1.0 / x);
      w.na[d] = w.nkb[d] / 2;
    }
    // G lists of the full mesh (calc_intW.F:277-330)
    w.ng0.assign(w.ntk, 0);
    w.kg0.assign((size_t)3 * ntg * w.ntk, 0);
    for (int jk = 0; jk < w.ntk; ++jk) {
      const int ik = w.numirr[jk];
      const int* r = rg + 9 * w.numrot[jk];
      double k[3];
      for (int i = 0; i < 3; ++i)
        k[i] = (double)r[i] * ski[3 * ik] + (double)r[i + 3] * ski[3 * ik + 1] +
               (double)r[i + 6] * ski[3 * ik + 2] + (double)w.rwv[3 * jk + i];
      std::vector<int> kg = make_kg0(w.b, w.a, w.ecut_psi, k);
      const int n = (int)(kg.size() / 3);
      RP_CHECK(n == ngi[ik], "NG_for_psi should be NGI(ik) (calc_intW.F:296-299)");
      w.ng0[jk] = n;
      const int sgn = w.trs[jk] == 1 ? 1 : -1;
      for (int i = 0; i < 3 * n; ++i) w.kg0[(size_t)3 * ntg * jk + i] = sgn * kg[i];
    }
    w.open = true;
  });
}

int rp_w3d_info(rp_handle h, int32_t* iout, double* dout) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_info");
    if (iout) {
      int v[16] = {w.ntk, w.nkb[0], w.nkb[1], w.nkb[2], w.na[0], w.na[1], w.na[2], w.nw[0], w.nw[1],
                   w.nw[2], w.nen, w.ntg, w.nwf, w.nq_irr, w.ntq, w.nog0};
      std::memcpy(iout, v, sizeof v);
    }
    if (dout) {
      double v[8] = {w.ecut_psi, w.ecut_eps, w.volume, w.chead, 0, 0, 0, 0};
      std::memcpy(dout, v, sizeof v);
    }
  });
}

int rp_w3d_kmesh(rp_handle h, double* sk0, int32_t* ng0) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_kmesh");
    if (sk0) std::memcpy(sk0, w.sk0.data(), sizeof(double) * w.sk0.size());
    if (ng0) std::memcpy(ng0, w.ng0.data(), sizeof(int) * w.ng0.size());
  });
}

int rp_w3d_set_wannier(rp_handle h, int32_t nwf, const double* wan) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_set_wannier");
    RP_CHECK(wan && nwf >= 1, "rp_w3d_set_wannier: bad arguments");
    RP_CHECK(nwf <= w3d_max_wannier(), "rp_w3d_set_wannier: at most " +
                                           std::to_string(w3d_max_wannier()) + " Wannier functions");
    RP_HIP(hipSetDevice(h->device));
    w.nwf = nwf;
    const int ntk = w.ntk, ntg = w.ntg;
    // C0(ig, iw, ik) (dat.wan, calc_intW.F:332-341) -> coef[w][k][ntg]
    std::vector<double2> coef((size_t)nwf * ntk * ntg, make_double2(0, 0));
    for (int k = 0; k < ntk; ++k)
      for (int iw = 0; iw < nwf; ++iw)
        for (int ig = 0; ig < w.ng0[k]; ++ig) {
          const double* z = wan + 2 * ((size_t)ig + (size_t)ntg * (iw + (size_t)nwf * k));
          coef[((size_t)iw * ntk + k) * ntg + ig] = make_double2(z[0], z[1]);
        }
    hipStream_t s = h->stream;
    w.d_coef.upload(coef, s);
    w.d_kg0.upload(w.kg0, s);
    w.d_ng0.upload(w.ng0, s);
    const size_t nr = (size_t)w.nw[0] * w.nw[1] * w.nw[2];
    w.d_ur.reserve((size_t)nwf * ntk * nr);
    w.d_tw.reserve(twiddle_entries());
    launch_twiddles(w.d_tw.p, s);
    {
      DevBuf<double2> scratch;
      const size_t se = band_grids_scratch_elems(w.nw[0], w.nw[1], w.nw[2], (size_t)nwf * ntk);
      scratch.reserve(se);
      launch_band_grids(w.d_coef.p, w.d_kg0.p, w.d_ng0.p, ntk, ntg, nwf, w.nw[0], w.nw[1], w.nw[2],
                        w.d_tw.p, w.d_ur.p, scratch.p, se, s);
      RP_HIP(hipStreamSynchronize(s));
    }
    RP_HIP(hipStreamSynchronize(s));
    w.have_wan = true;
    w.have_funcs = false;
  });
}

int rp_w3d_set_q(rp_handle h, int32_t nq_irr, const double* sqi, int32_t* ng_eps) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_set_q");
    w.maps_nwf = 0;  // the maps change below
    RP_CHECK(sqi && nq_irr == w.nk_irr, "rp_w3d_set_q: one q per irreducible k (Nq_irr = Nk_irr)");
    w.nq_irr = nq_irr;
    w.sq.assign(sqi, sqi + 3 * (size_t)nq_irr);
    w.numirrq.clear(); w.numrotq.clear(); w.trsq.clear(); w.rwq.clear();
    for (int iq = 0; iq < nq_irr; ++iq) {
      w.numirrq.push_back(iq);
      w.numrotq.push_back(0);
      w.trsq.push_back(1);
      w.rwq.insert(w.rwq.end(), {0, 0, 0});
    }
    // q unfolding (calc_intW.F:420-482)
    for (int iq = 0; iq < nq_irr; ++iq)
      for (int op = 0; op < w.nsym; ++op) {
        const int* r = w.rg.data() + 9 * op;
        for (int pass = 0; pass < 2; ++pass) {
          double k[3];
          int rw[3] = {0, 0, 0};
          for (int i = 0; i < 3; ++i)
            k[i] = (double)r[i] * sqi[3 * iq] + (double)r[i + 3] * sqi[3 * iq + 1] +
                   (double)r[i + 6] * sqi[3 * iq + 2];
          double sgn = 1.0;
          if (pass == 0) kcheck(k, rw);
          else { kcheck_trs(k, rw); sgn = -1.0; }
          bool found = false;
          for (size_t j = 0; j < w.numirrq.size(); ++j)
            if (std::fabs(w.sq[3 * j] - sgn * k[0]) < 1e-4 &&
                std::fabs(w.sq[3 * j + 1] - sgn * k[1]) < 1e-4 &&
                std::fabs(w.sq[3 * j + 2] - sgn * k[2]) < 1e-4) { found = true; break; }
          if (found) continue;
          for (int i = 0; i < 3; ++i) { w.sq.push_back(sgn * k[i]); w.rwq.push_back(rw[i]); }
          w.numirrq.push_back(iq);
          w.numrotq.push_back(op);
          w.trsq.push_back(pass == 0 ? 1 : -1);
        }
      }
    w.ntq = (int)w.numirrq.size();
    RP_CHECK(w.ntq == w.ntk, "NTQ should be NTK (calc_intW.F:484-487)");
    // G lists per full q (calc_intW.F:500-545)
    w.lg0.assign(w.ntq, {});
    w.geo_ready = false;
    w.nge.assign(w.ntq, 0);
    w.ngpsi.assign(w.ntq, 0);
    for (int iq = 0; iq < w.ntq; ++iq) {
      const int ir = w.numirrq[iq];
      const int* r = w.rg.data() + 9 * w.numrotq[iq];
      double q[3];
      for (int i = 0; i < 3; ++i)
        q[i] = iq < nq_irr ? sqi[3 * iq + i]
                           : (double)r[i] * sqi[3 * ir] + (double)r[i + 3] * sqi[3 * ir + 1] +
                                 (double)r[i + 6] * sqi[3 * ir + 2] + (double)w.rwq[3 * iq + i];
      int ne = 0;
      std::vector<int> lg = make_lg0(w, q, &ne);
      if (w.trsq[iq] == -1)
        for (int& v : lg) v = -v;
      w.nge[iq] = ne;
      w.ngpsi[iq] = (int)(lg.size() / 3);
      RP_CHECK(w.ngpsi[iq] <= w.ntg, "NG_for_psi(q) exceeds NTG (LG0 is sized NTG in the reference)");
      w.lg0[iq] = lg;
    }
    // Gamma: the irreducible index (calc_intW.F:595-597) and G = 0 in its list
    w.iqgm = -1;
    for (int iq = 0; iq < nq_irr; ++iq)
      if (std::fabs(w.ski[3 * iq]) < 1e-5 && std::fabs(w.ski[3 * iq + 1]) < 1e-5 &&
          std::fabs(w.ski[3 * iq + 2]) < 1e-5)
        w.iqgm = iq;
    RP_CHECK(w.iqgm >= 0, "no Gamma point among the irreducible q");
    w.nog0 = -1;
    for (int iq = 0; iq < w.ntq; ++iq)
      if (w.sq[3 * iq] == 0.0 && w.sq[3 * iq + 1] == 0.0 && w.sq[3 * iq + 2] == 0.0)
        for (int g = 0; g < w.ngpsi[iq]; ++g)
          if (w.lg0[iq][3 * g] == 0 && w.lg0[iq][3 * g + 1] == 0 && w.lg0[iq][3 * g + 2] == 0)
            w.nog0 = g;
    RP_CHECK(w.nog0 >= 0, "G = 0 not found at q = 0");
    // make_eps maps (calc_int_sub.F:222-311): eps_q(G,G') = eps_irr(m(G), m(G')) pf1(G) pf2(G')
    // (trs = -1: eps_irr(m(G'), m(G))), with m(G) = packing of R^-1 (+-G + RW)
    std::vector<std::map<std::tuple<int, int, int>, int>> pack(nq_irr);
    for (int iq = 0; iq < nq_irr; ++iq)
      for (int g = 0; g < w.nge[iq]; ++g)
        pack[iq][{w.lg0[iq][3 * g], w.lg0[iq][3 * g + 1], w.lg0[iq][3 * g + 2]}] = g;
    const double tpi = 2.0 * std::acos(-1.0);
    w.minv.clear(); w.pf1.clear(); w.pf2.clear(); w.map_off.clear(); w.qinfo.clear();
    for (int iq = 0; iq < w.ntq; ++iq) {
      const int ir = w.numirrq[iq], op = w.numrotq[iq], sgn = w.trsq[iq];
      RP_CHECK(w.nge[iq] == w.nge[ir], "rotated eps sphere differs in size from the irreducible one");
      const double* ri = &w.rginv[9 * (size_t)op];
      const int* pgo = w.pg.data() + 3 * op;
      const size_t off = w.minv.size();
      w.map_off.push_back((long long)off);
      w.minv.resize(off + w.nge[iq], -1);
      for (int g = 0; g < w.nge[iq]; ++g) {
        const int i1 = sgn * w.lg0[iq][3 * g], j1 = sgn * w.lg0[iq][3 * g + 1], k1 = sgn * w.lg0[iq][3 * g + 2];
        const int i2 = i1 + w.rwq[3 * iq], j2 = j1 + w.rwq[3 * iq + 1], k2 = k1 + w.rwq[3 * iq + 2];
        const int i3 = (int)ri[0] * i2 + (int)ri[3] * j2 + (int)ri[6] * k2;
        const int j3 = (int)ri[1] * i2 + (int)ri[4] * j2 + (int)ri[7] * k2;
        const int k3 = (int)ri[2] * i2 + (int)ri[5] * j2 + (int)ri[8] * k2;
        auto it = pack[ir].find({i3, j3, k3});
        RP_CHECK(it != pack[ir].end(), "rotated G vector missing from the irreducible eps sphere");
        RP_CHECK(w.minv[off + it->second] < 0, "make_eps map is not one-to-one");
        w.minv[off + it->second] = g;
        const double ph = tpi * ((double)i1 * pgo[0] + (double)j1 * pgo[1] + (double)k1 * pgo[2]) / (double)w.nnp;
        w.pf1.push_back(make_double2(std::cos(ph), -std::sin(ph)));
        w.pf2.push_back(make_double2(std::cos(ph), std::sin(ph)));
      }
      w.qinfo.push_back(make_int4(w.nge[iq], w.ngpsi[iq], ir, sgn));
    }
    w.rho_off.assign(w.ntq, 0);
    long long ro = 0;
    for (int iq = 0; iq < w.ntq; ++iq) {
      w.rho_off[iq] = ro;
      ro += (long long)w.ngpsi[iq] * w3d_max_wannier();
    }
    // HBM for eps^-1 of every irreducible q, filled by rp_w3d_set_epsinv
    w.eps_off.assign(nq_irr, 0);
    size_t tot = 0;
    for (int iq = 0; iq < nq_irr; ++iq) {
      w.eps_off[iq] = (long long)tot;
      tot += (size_t)w3d_eps_ld(w.nge[iq]) * w.nge[iq] * w.nen;
    }
    RP_HIP(hipSetDevice(h->device));
    w.d_eps.reserve(tot);
    w.eps_have.assign(nq_irr, 0);
    w.eps_head.assign(w.nen, make_float2(0.f, 0.f));
    if (ng_eps)
      for (int iq = 0; iq < nq_irr; ++iq) ng_eps[iq] = w.nge[iq];
    w.have_q = true;
    w.have_funcs = false;
  });
}

int rp_w3d_set_epsinv(rp_handle h, int32_t iq, const double* epsinv) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_set_epsinv");
    RP_CHECK(w.have_q, "rp_w3d_set_epsinv: q list not set (rp_w3d_set_q)");
    RP_CHECK(iq >= 0 && iq < w.nq_irr && epsinv, "rp_w3d_set_epsinv: bad arguments");
    const size_t ng = w.nge[iq], ld = (size_t)w3d_eps_ld((int)ng), n = ld * ng * w.nen;
    // epsirr is complex(4) in the reference (calc_intW.F:586-630); columns
    // padded with zeros to ld rows
    std::vector<float2> e(n, make_float2(0.f, 0.f));
    for (size_t c = 0; c < ng * w.nen; ++c)
      for (size_t r = 0; r < ng; ++r) {
        const size_t i = r + ng * c;
        e[r + ld * c] = make_float2((float)epsinv[2 * i], (float)epsinv[2 * i + 1]);
      }
    if (iq == w.iqgm)
      for (int ie = 0; ie < w.nen; ++ie)
        w.eps_head[ie] = e[(size_t)w.nog0 + ld * ((size_t)w.nog0 + ng * ie)];
    RP_HIP(hipSetDevice(h->device));
    RP_HIP(hipMemcpy(w.d_eps.p + w.eps_off[iq], e.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
    w.eps_have[iq] = 1;
    w.have_funcs = false;
  });
}

int rp_w3d_funcs(rp_handle h, double* func_out, double* funcw_out) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_funcs");
    check_ready(w, "rp_w3d_funcs");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int ntk = w.ntk, nwf = w.nwf, nen = w.nen;
    upload_eps_maps(w, s);
    w.d_rho_off.upload(w.rho_off, s);
    w.d_rho.reserve((size_t)(w.rho_off.back() + (long long)w.ngpsi.back() * w3d_max_wannier()));
    // rho for every q (calc_intW.F:647-767): the diagonal Wannier pairs
    ensure_geometry(w, s);
    w.events(4);
    RP_HIP(hipEventRecord(w.ev[0], s));
    rho_all_q(w, false, nullptr, w.d_rho_off.p, w.d_rho.p, s);
    RP_HIP(hipEventRecord(w.ev[1], s));
    w.d_func.reserve((size_t)w.ntq * nwf * nwf);
    w.d_shell.reserve((size_t)w.ntq * nwf * nwf);
    w.d_funcw.reserve((size_t)w.ntq * nen * nwf * nwf);
    launch_w_func(w.d_rho.p, w.d_rho_off.p, w.d_qinfo.p, w.ntq, nwf, w.d_func.p, w.d_shell.p, s);
    RP_HIP(hipEventRecord(w.ev[2], s));
    w.d_wx.reserve((size_t)w_contract_vectors(nwf) * (size_t)w.minv.size());
    launch_w_contract(w.d_eps.p, w.d_eps_off.p, w.d_rho.p, w.d_rho_off.p, w.d_qinfo.p, w.d_minv.p,
                      w.d_pf1.p, w.d_pf2.p, w.d_map_off.p, w.d_shell.p, w.d_wgoff.p, w.d_wgmem.p,
                      (int)w.wgoff.size() - 1, nwf, nen,
                      *std::max_element(w.nge.begin(), w.nge.end()), w.d_wx.p, w.d_funcw.p, s);
    RP_HIP(hipEventRecord(w.ev[3], s));
    if (func_out)
      RP_HIP(hipMemcpyAsync(func_out, w.d_func.p, sizeof(double2) * w.ntq * nwf * nwf,
                            hipMemcpyDeviceToHost, s));
    if (funcw_out)
      RP_HIP(hipMemcpyAsync(funcw_out, w.d_funcw.p, sizeof(double2) * w.ntq * nen * nwf * nwf,
                            hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
    w.stage_s[0] = event_s(w.ev[0], w.ev[1]);
    w.stage_s[1] = event_s(w.ev[1], w.ev[2]);
    w.stage_s[2] = event_s(w.ev[2], w.ev[3]);
    w.stage_s[3] = event_s(w.ev[0], w.ev[3]);
    w.have_funcs = true;
  });
}

int rp_w3d_vw(rp_handle h, double* vmat, double* wmat, double* heads) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_vw");
    if (!w.have_funcs)
      if (rp_w3d_funcs(h, nullptr, nullptr)) throw Error(g_err);
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int nwf = w.nwf, nen = w.nen;
    compute_heads(w);
    const int nr = (2 * w.na[0] + 1) * (2 * w.na[1] + 1) * (2 * w.na[2] + 1);
    w.d_sq.upload(w.sq, s);
    w.d_cheadw.upload(w.cheadw, s);
    w.d_vmat.reserve((size_t)nr * nwf * nwf);
    w.d_wmat.reserve((size_t)nr * nwf * nwf * nen);
    launch_w_fourier(w.d_func.p, w.d_funcw.p, w.d_sq.p, w.ntq, nwf, nen, w.na, w.volume, w.chead,
                     w.d_cheadw.p, w.d_vmat.p, w.d_wmat.p, s);
    if (vmat)
      RP_HIP(hipMemcpyAsync(vmat, w.d_vmat.p, sizeof(double2) * nr * nwf * nwf, hipMemcpyDeviceToHost, s));
    if (wmat)
      RP_HIP(hipMemcpyAsync(wmat, w.d_wmat.p, sizeof(double2) * nr * nwf * nwf * nen,
                            hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
    if (heads) {
      heads[0] = w.chead;
      heads[1] = 0.0;
      for (int ie = 0; ie < nen; ++ie) {
        heads[2 + 2 * ie] = w.cheadw[ie].x;
        heads[3 + 2 * ie] = w.cheadw[ie].y;
      }
    }
  });
}

int rp_w3d_stage_times(rp_handle h, double* t) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_w3d_stage_times");
    RP_CHECK(t, "rp_w3d_stage_times: null output");
    for (int i = 0; i < 8; ++i) t[i] = w.stage_s[i];
  });
}

int rp_j3d_run(rp_handle h, int32_t nr, const int32_t* rvec, double* xmat, double* jmat,
               double* heads) {
  return guarded([&] {
    W3d& w = w3d_of(h, "rp_j3d_run");
    check_ready(w, "rp_j3d_run");
    RP_CHECK(nr >= 1 && rvec, "rp_j3d_run: empty R list");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int ntk = w.ntk, nwf = w.nwf, nen = w.nen, npair = nwf * nwf;
    upload_eps_maps(w, s);
    compute_heads(w);
    // pair p = iw + nwf*jw: C0_K = C0(:,jw,k), C0_KQ = C0(:,iw,k+q) (calc_intJ.F:651-657),
    // the pair order of launch_pair_rsum's full mode
    // e^{-i 2pi k.R} per R and k (calc_intJ.F:646-650)
    const double tpi = 2.0 * std::acos(-1.0);
    std::vector<double2> pfk((size_t)nr * ntk);
    for (int r = 0; r < nr; ++r)
      for (int k = 0; k < ntk; ++k) {
        const double ph = tpi * (w.sk0[3 * k] * rvec[3 * r] + w.sk0[3 * k + 1] * rvec[3 * r + 1] +
                                 w.sk0[3 * k + 2] * rvec[3 * r + 2]);
        pfk[(size_t)r * ntk + k] = make_double2(std::cos(ph), -std::sin(ph));
      }
    DevBuf<double2> d_pfk;
    d_pfk.upload(pfk, s);
    // rho per (R, q): [r][q][p][G]
    std::vector<long long> joff(w.ntq);
    long long tot = 0;
    for (int iq = 0; iq < w.ntq; ++iq) { joff[iq] = tot; tot += (long long)npair * w.ngpsi[iq]; }
    DevBuf<long long> d_joff;
    d_joff.upload(joff, s);
    DevBuf<double2> d_jrho, d_jfunc, d_jshell, d_jfuncw, d_x, d_j, d_cw, d_xg;
    d_jrho.reserve((size_t)tot * nr);
    ensure_geometry(w, s);
    w.events(2 + 2 * (size_t)nr);
    RP_HIP(hipEventRecord(w.ev[0], s));
    for (int r = 0; r < nr; ++r)
      rho_all_q(w, true, d_pfk.p + (size_t)r * ntk, d_joff.p, d_jrho.p + (size_t)r * tot, s);
    RP_HIP(hipEventRecord(w.ev[1], s));
    d_jfunc.reserve((size_t)w.ntq * npair);
    d_jshell.reserve((size_t)w.ntq * npair);
    d_jfuncw.reserve((size_t)w.ntq * nen * npair);
    d_x.reserve((size_t)nr * npair);
    d_j.reserve((size_t)nr * nen * npair);
    d_cw.upload(w.cheadw, s);
    d_xg.reserve(j_contract_scratch(w.minv.size(), npair));
    for (int r = 0; r < nr; ++r) {
      const double2* rho = d_jrho.p + (size_t)r * tot;
      launch_j_func(rho, d_joff.p, w.d_qinfo.p, w.ntq, npair, d_jfunc.p, d_jshell.p, s);
      RP_HIP(hipEventRecord(w.ev[2 + 2 * r], s));
      launch_j_contract(w.d_eps.p, w.d_eps_off.p, rho, d_joff.p, w.d_qinfo.p, w.d_minv.p, w.d_pf1.p,
                        w.d_pf2.p, w.d_map_off.p, d_xg.p, d_jshell.p, w.ntq, npair, nen, d_jfuncw.p,
                        s);
      RP_HIP(hipEventRecord(w.ev[3 + 2 * r], s));
      const int r0 = rvec[3 * r] == 0 && rvec[3 * r + 1] == 0 && rvec[3 * r + 2] == 0;
      launch_j_sum(d_jfunc.p, d_jfuncw.p, w.ntq, nwf, nen, r0, w.volume, w.chead, d_cw.p,
                   d_x.p + (size_t)r * npair, d_j.p + (size_t)r * npair, (size_t)nr * npair, s);
    }
    if (xmat)
      RP_HIP(hipMemcpyAsync(xmat, d_x.p, sizeof(double2) * nr * npair, hipMemcpyDeviceToHost, s));
    if (jmat)
      RP_HIP(hipMemcpyAsync(jmat, d_j.p, sizeof(double2) * nr * nen * npair, hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
    double tj = 0.0;
    for (int r = 0; r < nr; ++r) tj += event_s(w.ev[2 + 2 * r], w.ev[3 + 2 * r]);
    w.stage_s[4] = event_s(w.ev[0], w.ev[1]);
    w.stage_s[6] = tj;
    w.stage_s[7] = event_s(w.ev[0], w.ev[3 + 2 * (nr - 1)]);
    w.stage_s[5] = w.stage_s[7] - w.stage_s[4] - tj;
    if (heads) {
      heads[0] = w.chead;
      heads[1] = 0.0;
      for (int ie = 0; ie < nen; ++ie) {
        heads[2 + 2 * ie] = w.cheadw[ie].x;
        heads[3 + 2 * ie] = w.cheadw[ie].y;
      }
    }
  });
}

}  // extern "C"
