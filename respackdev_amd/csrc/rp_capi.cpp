// C ABI of librespack_gpu: host-side setup of the chiqw path and the per-q
// orchestration of the gfx950 kernels.  Declarations and argument layouts are
// in include/respack_gpu.h.
//
// The setup restates the reference's host logic (k-mesh unfolding, G spheres,
// band classification, frequency grid) so that every index and ordering the
// output depends on is the reference's; all per-q arithmetic of the hot path
// (pair densities, tetrahedron weights, contraction, Dyson) runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rp_internal.h"

using namespace rp;

extern "C" {

const char* rp_last_error(void) { return g_err.c_str(); }

int rp_init(int32_t device, rp_handle* out) {
  return guarded([&] {
    RP_CHECK(out, "rp_init: null output");
    int n = 0;
    RP_HIP(hipGetDeviceCount(&n));
    RP_CHECK(n > 0, "rp_init: no HIP device visible");
    RP_CHECK(device >= 0 && device < n, "rp_init: device ordinal out of range");
    RP_HIP(hipSetDevice(device));
    noalloc_init(device);
    hipDeviceProp_t prop;
    RP_HIP(hipGetDeviceProperties(&prop, device));
    RP_CHECK(std::string(prop.gcnArchName).rfind("gfx950", 0) == 0,
             std::string("rp_init: built for gfx950, device is ") + prop.gcnArchName);
    auto* h = new rp_ctx();
    h->device = device;
    h->ncu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    RP_HIP(hipStreamCreate(&h->stream));
    RP_HIP(hipStreamCreate(&h->stream2));
    for (auto& e : h->ev) RP_HIP(hipEventCreate(&e));
    for (auto& e : h->gev) RP_HIP(hipEventCreate(&e));
    for (auto& e : h->bev) RP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    h->d_status.reserve(1);
    *out = h;
  });
}

int rp_finalize(rp_handle h) {
  return guarded([&] {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int32_t rp_algn235(int32_t n) { return algn235(n); }

int rp_gen_grid(int32_t nt, double emax, double* em) {
  return guarded([&] {
    RP_CHECK(nt >= 1 && em, "rp_gen_grid: bad arguments");
    gen_grid(nt, emax, em);
  });
}

int rp_make_kg0(int32_t ntg, const double* b, double gcut, const double* q, int32_t* kg, int32_t* ng) {
  return guarded([&] {
    double bb[3][3], aa[3][3];
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) bb[i][d] = b[3 * i + d];
    // real-space vectors for the bounding box: a_i = 2pi (b^-1)^T rows
    double det = bb[0][0] * (bb[1][1] * bb[2][2] - bb[1][2] * bb[2][1]) -
                 bb[0][1] * (bb[1][0] * bb[2][2] - bb[1][2] * bb[2][0]) +
                 bb[0][2] * (bb[1][0] * bb[2][1] - bb[1][1] * bb[2][0]);
    RP_CHECK(std::fabs(det) > 0, "rp_make_kg0: singular reciprocal basis");
    double c[3][3];
    for (int i = 0; i < 3; ++i) {
      int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
      for (int d = 0; d < 3; ++d) {
        int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        c[i][d] = bb[i1][d1] * bb[i2][d2] - bb[i1][d2] * bb[i2][d1];
      }
    }
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) aa[i][d] = 2.0 * M_PI * c[i][d] / det;
    std::vector<int> v = make_kg0(bb, aa, gcut, q);
    *ng = (int32_t)(v.size() / 3);
    if (kg) std::memcpy(kg, v.data(), sizeof(int) * std::min<size_t>(v.size(), (size_t)3 * ntg));
  });
}

int rp_ttrhdrn_mkidx(int32_t n1, int32_t n2, int32_t n3, const double* b, int32_t* idx) {
  return guarded([&] { mkidx(n1, n2, n3, b, idx); });
}

int rp_chiqw_open(rp_handle h, const double* lattice, int32_t nsym, int32_t nnp, const int32_t* rg,
                  const int32_t* pg, int32_t nk_irr, const double* ski, const int32_t* ngi,
                  int32_t ntg, const int32_t* kgi, int32_t ntb, const double* eig,
                  const double* wfn, double ef, const rp_chiqw_params* prm) {
  return guarded([&] {
    RP_CHECK(h && lattice && rg && pg && ski && ngi && kgi && eig && wfn && prm,
             "rp_chiqw_open: null argument");
    RP_HIP(hipSetDevice(h->device));
    h->open = false;
    h->qready = false;
    // ---- lattice (chiqw.F:145-156)
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) h->a[i][d] = lattice[3 * i + d];
    auto cross = [](const double* x, const double* y, double* z) {
      z[0] = x[1] * y[2] - x[2] * y[1];
      z[1] = x[2] * y[0] - x[0] * y[2];
      z[2] = x[0] * y[1] - x[1] * y[0];
    };
    const double tpi = 2.0 * std::acos(-1.0);
    cross(h->a[1], h->a[2], h->b[0]);
    h->volume = h->a[0][0] * h->b[0][0] + h->a[0][1] * h->b[0][1] + h->a[0][2] * h->b[0][2];
    cross(h->a[2], h->a[0], h->b[1]);
    cross(h->a[0], h->a[1], h->b[2]);
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) h->b[i][d] = h->b[i][d] * tpi / h->volume;
    // ---- parameters (m_rdinput.f90:52-128, chiqw.F:67-69)
    h->nk_irr = nk_irr;
    h->ntg = ntg;
    h->ntb = ntb;
    h->nsym = nsym;
    h->nnp = nnp;
    h->ef = ef + prm->shift_ef / kAu;
    h->delt = prm->green_func_delt / kAu;
    h->delta_ex = prm->max_excitation_energy / kAu;
    h->dmna = prm->ttrhdrn_dmna;
    h->dmnr = prm->ttrhdrn_dmnr;
    h->nen = prm->num_freq_grid;
    h->flg_crpa = prm->flg_crpa;
    RP_CHECK(h->nen >= 1, "Num_freq_grid must be >= 1");
    // ---- Ecut_for_psi from the largest |k+G|^2 (chiqw.F:313-352)
    double emx = 0;
    for (int ik = 0; ik < nk_irr; ++ik)
      for (int ig = 0; ig < ngi[ik]; ++ig) {
        const int* g = kgi + 3 * ((size_t)ik * ntg + ig);
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = (ski[3 * ik] + (double)g[0]) * h->b[0][d] + (ski[3 * ik + 1] + (double)g[1]) * h->b[1][d] +
                 (ski[3 * ik + 2] + (double)g[2]) * h->b[2][d];
        emx = std::max(emx, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      }
    h->ecut_psi = emx + 1.0e-8;
    h->ecut_eps = prm->ecut_for_eps == 0.0 ? h->ecut_psi / 10.0 : prm->ecut_for_eps;
    // ---- packing of the irreducible G lists (chiqw.F:360-378)
    int L[3] = {0, 0, 0};
    for (size_t i = 0; i < (size_t)3 * ntg * nk_irr; ++i) L[i % 3] = std::max(L[i % 3], std::abs(kgi[i]));
    for (int d = 0; d < 3; ++d) L[d] += 1;
    const int P1 = 2 * L[0] + 1, P2 = 2 * L[1] + 1, P3 = 2 * L[2] + 1;
    std::vector<int> pack((size_t)P1 * P2 * P3 * nk_irr, -1);
    for (int ik = 0; ik < nk_irr; ++ik)
      for (int ig = 0; ig < ngi[ik]; ++ig) {
        const int* g = kgi + 3 * ((size_t)ik * ntg + ig);
        pack[(size_t)ik * P1 * P2 * P3 + (g[0] + L[0]) + P1 * ((g[1] + L[1]) + (size_t)P2 * (g[2] + L[2]))] = ig;
      }
    // ---- FFT grid (chiqw.F:389-401)
    const double qwf = 2.0 * std::sqrt(h->ecut_psi);
    for (int d = 0; d < 3; ++d) {
      double an = std::sqrt(h->a[d][0] * h->a[d][0] + h->a[d][1] * h->a[d][1] + h->a[d][2] * h->a[d][2]);
      double dd = std::fabs((h->b[d][0] * h->a[d][0] + h->b[d][1] * h->a[d][1] + h->b[d][2] * h->a[d][2]) / an);
      h->nw[d] = algn235((int)(qwf / dd) + 1);
      RP_CHECK(dft_size_supported(h->nw[d]),
               "FFT grid dimension " + std::to_string(h->nw[d]) + " is not supported (2,3,5-smooth sizes up to 32)");
    }
    // ---- symmetry (rginv by the adjugate: integer matrices of det +-1)
    std::vector<double> rginv(9 * (size_t)nsym);
    for (int op = 0; op < nsym; ++op) {
      const int* r = rg + 9 * op;  // r[i + 3j] = rg(i+1, j+1)
      auto R = [&](int i, int j) { return (double)r[i + 3 * j]; };
      double det = R(0, 0) * (R(1, 1) * R(2, 2) - R(1, 2) * R(2, 1)) -
                   R(0, 1) * (R(1, 0) * R(2, 2) - R(1, 2) * R(2, 0)) +
                   R(0, 2) * (R(1, 0) * R(2, 1) - R(1, 1) * R(2, 0));
      RP_CHECK(det != 0.0, "singular symmetry operation");
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          int j1 = (j + 1) % 3, j2 = (j + 2) % 3, i1 = (i + 1) % 3, i2 = (i + 2) % 3;
          // inverse(i,j) = cofactor(j,i)/det
          rginv[9 * op + i + 3 * j] = (R(j1, i1) * R(j2, i2) - R(j1, i2) * R(j2, i1)) / det;
        }
    }
    // ---- unfold the mesh (chiqw.F:437-480)
    h->sk0.clear(); h->numirr.clear(); h->numrot.clear(); h->trs.clear(); h->rwv.clear();
    for (int ik = 0; ik < nk_irr; ++ik)
      for (int op = 0; op < nsym; ++op) {
        const int* r = rg + 9 * op;
        for (int pass = 0; pass < 2; ++pass) {
          double k[3];
          int w[3] = {0, 0, 0};
          for (int i = 0; i < 3; ++i)
            k[i] = (double)r[i] * ski[3 * ik] + (double)r[i + 3] * ski[3 * ik + 1] + (double)r[i + 6] * ski[3 * ik + 2];
          double sgn = 1.0;
          if (pass == 0) kcheck(k, w);
          else { kcheck_trs(k, w); sgn = -1.0; }
          bool found = false;
          for (size_t j = 0; j < h->numirr.size(); ++j)
            if (std::fabs(h->sk0[3 * j] - sgn * k[0]) < 1e-4 && std::fabs(h->sk0[3 * j + 1] - sgn * k[1]) < 1e-4 &&
                std::fabs(h->sk0[3 * j + 2] - sgn * k[2]) < 1e-4) { found = true; break; }
          if (found) continue;
          for (int i = 0; i < 3; ++i) { h->sk0.push_back(sgn * k[i]); h->rwv.push_back(w[i]); }
          h->numirr.push_back(ik);
          h->numrot.push_back(op);
          h->trs.push_back(pass == 0 ? 1 : -1);
        }
      }
    h->ntk = (int)h->numirr.size();
    const int ntk = h->ntk;
    // ---- mesh size (est_NTK.f90:1-39) and the 3D index (chiqw_sub.F:324-384)
    for (int d = 0; d < 3; ++d) {
      double x = 1.0;
      for (int i = 0; i < ntk; ++i) {
        double v = std::fabs(h->sk0[3 * i + d]);
        if (v < 1.0e-7) continue;
        if (v < x) x = v;
      }
      h->nkb[d] = (int)std::lround(1.0 / x);
    }
    const int n1 = h->nkb[0], n2 = h->nkb[1], n3 = h->nkb[2];
    RP_CHECK(n1 * n2 * n3 == ntk, "k mesh is not a full Gamma-centred nkb1 x nkb2 x nkb3 mesh");
    h->index_kpt.assign((size_t)ntk, -1);
    for (int ik = 0; ik < ntk; ++ik) {
      double x = h->sk0[3 * ik] * n1, y = h->sk0[3 * ik + 1] * n2, z = h->sk0[3 * ik + 2] * n3;
      long ix, iy, iz;
      if (ntk % 2 != 0) {
        ix = std::lround(x + (n1 - 1.0) / 2.0) + 1; iy = std::lround(y + (n2 - 1.0) / 2.0) + 1;
        iz = std::lround(z + (n3 - 1.0) / 2.0) + 1;
      } else {
        ix = std::lround(x + n1 / 2.0); iy = std::lround(y + n2 / 2.0); iz = std::lround(z + n3 / 2.0);
      }
      RP_CHECK(ix >= 1 && ix <= n1 && iy >= 1 && iy <= n2 && iz >= 1 && iz <= n3, "k point off the mesh");
      h->index_kpt[(ix - 1) + n1 * ((iy - 1) + (size_t)n2 * (iz - 1))] = ik;
    }
    for (int v : h->index_kpt) RP_CHECK(v >= 0, "k mesh has holes");
    h->imt.assign((size_t)24 * ntk, 0);
    double bflat[9];
    for (int i = 0; i < 3; ++i)
      for (int d = 0; d < 3; ++d) bflat[3 * i + d] = h->b[i][d];
    mkidx(n1, n2, n3, bflat, h->imt.data());
    RP_CHECK(tetra_lds_bytes(ntk, 1) <= 160 * 1024,
             "k mesh x frequencies too large for the LDS accumulator of the tetrahedron kernel");
    // ---- G lists of the full mesh (chiqw.F:509-545)
    h->ng0.assign(ntk, 0);
    h->kg0.assign((size_t)3 * ntg * ntk, 0);
    for (int jk = 0; jk < ntk; ++jk) {
      const int ik = h->numirr[jk], op = h->numrot[jk];
      const int* r = rg + 9 * op;
      double k[3];
      for (int i = 0; i < 3; ++i)
        k[i] = (double)r[i] * ski[3 * ik] + (double)r[i + 3] * ski[3 * ik + 1] + (double)r[i + 6] * ski[3 * ik + 2] +
               (double)h->rwv[3 * jk + i];
      std::vector<int> kg = make_kg0(h->b, h->a, h->ecut_psi, k);
      int n = (int)(kg.size() / 3);
      RP_CHECK(n == ngi[ik], "NG_for_psi should be NGI(ik) (chiqw.F:521-524)");
      h->ng0[jk] = n;
      const double s = h->trs[jk] == 1 ? 1 : -1;
      for (int i = 0; i < 3 * n; ++i) h->kg0[(size_t)3 * ntg * jk + i] = (int)(s * kg[i]);
    }
    // ---- band classification (chiqw_sub.F:570-613; chiqw.F:941-967)
    h->n_calc = prm->n_calc_band == 0 ? ntb : prm->n_calc_band;
    RP_CHECK(h->n_calc >= 1 && h->n_calc <= ntb, "N_CALC_BAND out of range");
    h->win.assign(ntb, 0);
    h->tindx.assign(ntb, 0);
    h->eave.assign(ntb, 0);
    std::vector<double> bmax(ntb), bmin(ntb);
    for (int ib = 0; ib < ntb; ++ib) {
      double top = 0, btm = 0;
      for (int ik = 0; ik < nk_irr; ++ik) {
        double e = eig[ib + (size_t)ntb * ik];
        if (ik == 0) top = btm = e;
        if (e < btm) btm = e;
        if (e > top) top = e;
      }
      int fs = 0;
      if (btm > h->ef) fs = 2;
      if (btm <= h->ef && top >= h->ef) { fs = 1; h->tindx[ib] = ib + 1; }
      if (top < h->ef) fs = 0;
      h->win[ib] = fs;
      h->eave[ib] = (top + btm) / 2.0;
      bmax[ib] = top;
      bmin[ib] = btm;
    }
    h->n_docc = h->n_pocc = h->n_uocc = 0;
    for (int ib = 0; ib < h->n_calc; ++ib) {
      if (h->win[ib] == 0) ++h->n_docc;
      if (h->win[ib] == 1) ++h->n_pocc;
      if (h->win[ib] == 2) ++h->n_uocc;
    }
    h->n_occ = h->n_docc + h->n_pocc;
    h->n_vir = h->n_uocc + h->n_pocc;
    // ---- frequency grid (chiqw.F:1184-1192)
    double mx = -1e300, mn = 1e300;
    for (int ib = 0; ib < h->n_calc; ++ib) { mx = std::max(mx, bmax[ib]); mn = std::min(mn, bmin[ib]); }
    h->emax = std::fabs(mx - mn);
    h->em.assign(2 * (size_t)h->nen, 0.0);
    gen_grid(h->nen, h->emax, h->em.data());
    for (double v : h->em)
      RP_CHECK(std::isfinite(v), "the reference's gen_grid gives a non-finite frequency grid for this "
                                 "Num_freq_grid and band spread (choose another Num_freq_grid)");
    // ---- energies and (zero) projector weights per band and full-mesh k
    const int nb = h->n_calc;
    h->eband.assign((size_t)nb * ntk, 0.0);
    for (int ib = 0; ib < nb; ++ib)
      for (int k = 0; k < ntk; ++k) h->eband[(size_t)ib * ntk + k] = eig[ib + (size_t)ntb * h->numirr[k]];
    h->pband.assign(2 * (size_t)nb * ntk, 0.0);
    // ---- unfold the wavefunctions (make_C0, chiqw_sub.F:264-322) and upload
    std::vector<double2> coef((size_t)nb * ntk * ntg, make_double2(0, 0));
    for (int k = 0; k < ntk; ++k) {
      const int ik = h->numirr[k], op = h->numrot[k], itrs = h->trs[k];
      const double* ri = &rginv[9 * (size_t)op];
      const int* pgo = pg + 3 * op;
      const double sgn = itrs == 1 ? 1.0 : -1.0;
      for (int ig = 0; ig < h->ng0[k]; ++ig) {
        const int* g = &h->kg0[(size_t)3 * ntg * k + 3 * ig];
        int i1 = (int)sgn * g[0], j1 = (int)sgn * g[1], k1 = (int)sgn * g[2];
        int i2 = i1 + h->rwv[3 * k], j2 = j1 + h->rwv[3 * k + 1], k2 = k1 + h->rwv[3 * k + 2];
        int i3 = (int)ri[0] * i2 + (int)ri[3] * j2 + (int)ri[6] * k2;
        int j3 = (int)ri[1] * i2 + (int)ri[4] * j2 + (int)ri[7] * k2;
        int k3 = (int)ri[2] * i2 + (int)ri[5] * j2 + (int)ri[8] * k2;
        RP_CHECK(std::abs(i3) <= L[0] && std::abs(j3) <= L[1] && std::abs(k3) <= L[2],
                 "rotated G vector outside the packing box");
        int jg = pack[(size_t)ik * P1 * P2 * P3 + (i3 + L[0]) + P1 * ((j3 + L[1]) + (size_t)P2 * (k3 + L[2]))];
        RP_CHECK(jg >= 0, "rotated G vector missing from the irreducible G list");
        double ph = tpi * ((double)i1 * pgo[0] + (double)j1 * pgo[1] + (double)k1 * pgo[2]);
        double cr = std::cos(ph / (double)nnp), ci = -std::sin(ph / (double)nnp);
        for (int ib = 0; ib < nb; ++ib) {
          const double* z = wfn + 2 * ((size_t)jg + (size_t)ntg * (ib + (size_t)ntb * ik));
          double vr = z[0] * cr - z[1] * ci, vi = z[0] * ci + z[1] * cr;
          coef[((size_t)ib * ntk + k) * ntg + ig] = make_double2(vr, itrs == 1 ? vi : -vi);
        }
      }
    }
    hipStream_t s = h->stream;
    h->d_coef.upload(coef, s);
    h->d_kg0.upload(h->kg0, s);
    h->d_ng0.upload(h->ng0, s);
    const size_t nr = (size_t)h->nw[0] * h->nw[1] * h->nw[2];
    h->d_ur.reserve((size_t)nb * ntk * nr);
    h->d_tw.reserve(twiddle_entries());
    launch_twiddles(h->d_tw.p, s);
    {
      // grids too large for the LDS are transformed over HBM in batches
      DevBuf<double2> scratch;
      const size_t se = band_grids_scratch_elems(h->nw[0], h->nw[1], h->nw[2], (size_t)nb * ntk);
      scratch.reserve(se);
      launch_band_grids(h->d_coef.p, h->d_kg0.p, h->d_ng0.p, ntk, ntg, nb, h->nw[0], h->nw[1],
                        h->nw[2], h->d_tw.p, h->d_ur.p, scratch.p, se, s);
      RP_HIP(hipStreamSynchronize(s));
    }
    h->d_eband.upload(h->eband, s);
    h->d_pband.upload((const double2*)h->pband.data(), (size_t)nb * ntk, s);
    h->d_imt.upload(h->imt, s);
    h->d_index.upload(h->index_kpt, s);
    h->d_em.upload((const double2*)h->em.data(), (size_t)h->nen, s);
    h->d_bvec.upload(bflat, 9, s);
    h->d_sk0.upload(h->sk0, s);
    RP_HIP(hipStreamSynchronize(s));
    h->open = true;
  });
}

int rp_chiqw_set_crpa(rp_handle h, int32_t ntk, const int32_t* ns, const int32_t* nb, int32_t mb,
                      int32_t nwf, const double* umat) {
  return guarded([&] {
    RP_CHECK(h && h->open, "rp_chiqw_set_crpa: system not open");
    RP_CHECK(ntk == h->ntk, "rp_chiqw_set_crpa: NTK mismatch");
    // prob(Ns(k)+ib, k) = sum_w |U(ib,w,k)|^2  (chiqw.F:905-915)
    std::fill(h->pband.begin(), h->pband.end(), 0.0);
    for (int k = 0; k < ntk; ++k)
      for (int ib = 0; ib < nb[k]; ++ib) {
        double sr = 0, si = 0;
        for (int w = 0; w < nwf; ++w) {
          const double* u = umat + 2 * ((size_t)ib + (size_t)mb * (w + (size_t)nwf * k));
          // conj(U) U
          sr += u[0] * u[0] + u[1] * u[1];
          si += u[0] * u[1] - u[1] * u[0];
        }
        int band = ns[k] + ib;
        if (band >= 0 && band < h->n_calc) {
          h->pband[2 * ((size_t)band * ntk + k)] = sr;
          h->pband[2 * ((size_t)band * ntk + k) + 1] = si;
        }
      }
    h->d_pband.upload((const double2*)h->pband.data(), (size_t)h->n_calc * ntk, h->stream);
    RP_HIP(hipStreamSynchronize(h->stream));
  });
}

int rp_chiqw_info(rp_handle h, int32_t* iout, double* dout, double* em) {
  return guarded([&] {
    RP_CHECK(h && h->open, "rp_chiqw_info: system not open");
    if (iout) {
      int v[16] = {h->ntk, h->nkb[0], h->nkb[1], h->nkb[2], h->nw[0], h->nw[1], h->nw[2], h->nen,
                   h->n_occ, h->n_vir, h->n_docc, h->ntb, h->nk_irr, h->ntg, 0, 0};
      std::memcpy(iout, v, sizeof v);
    }
    if (dout) {
      double v[8] = {h->ecut_psi, h->ecut_eps, h->volume, h->emax, h->ef, h->delt, 0, 0};
      std::memcpy(dout, v, sizeof v);
    }
    if (em) std::memcpy(em, h->em.data(), sizeof(double) * h->em.size());
  });
}

int rp_chiqw_kmesh(rp_handle h, double* sk0, int32_t* numirr, int32_t* trs) {
  return guarded([&] {
    RP_CHECK(h && h->open, "rp_chiqw_kmesh: system not open");
    if (sk0) std::memcpy(sk0, h->sk0.data(), sizeof(double) * h->sk0.size());
    if (numirr) std::memcpy(numirr, h->numirr.data(), sizeof(int) * h->numirr.size());
    if (trs) std::memcpy(trs, h->trs.data(), sizeof(int) * h->trs.size());
  });
}

int rp_chiqw_prepare_q(rp_handle h, const double* qin, int32_t* ng, int32_t* lgout, int32_t lg_cap) {
  return guarded([&] {
    RP_CHECK(h && h->open, "rp_chiqw_prepare_q: system not open");
    auto t0 = std::chrono::steady_clock::now();
    double q[3];
    search_list(qin, q);  // chiqw.F:727-729
    const bool same = h->qready && q[0] == h->q[0] && q[1] == h->q[1] && q[2] == h->q[2];
    if (!same) {
      h->qready = h->have_A = h->have_w = h->have_chi = false;
      h->q_span = 0;
      std::memcpy(h->q, q, sizeof q);
      h->q0 = q[0] == 0.0 && q[1] == 0.0 && q[2] == 0.0;  // chiqw.F:1257
      h->lg = make_kg0(h->b, h->a, h->ecut_eps_win > 0.0 ? h->ecut_eps_win : h->ecut_eps, q);
      h->ng = (int)(h->lg.size() / 3);
      RP_CHECK(h->ng > 0, "empty chi G sphere");
      RP_CHECK(h->ng <= h->ntg, "NG_for_eps exceeds NTG (LG0 is sized NTG in the reference)");
      h->g0 = -1;
      for (int i = 0; i < h->ng; ++i)
        if (h->lg[3 * i] == 0 && h->lg[3 * i + 1] == 0 && h->lg[3 * i + 2] == 0) h->g0 = i;
      const int ntk = h->ntk;
      h->ikq.assign(ntk, 0);
      h->shift.assign(3 * (size_t)ntk, 0);
      for (int k = 0; k < ntk; ++k) {
        if (h->q0) { h->ikq[k] = k; continue; }
        int sg[3];
        int kq = search_kq(ntk, h->sk0.data(), q, k, sg);
        RP_CHECK(kq >= 0, "k+q is not on the k mesh (q must be a mesh point)");
        h->ikq[k] = kq;
        for (int d = 0; d < 3; ++d) h->shift[3 * k + d] = sg[d];
      }
      // pair list (chiqw.F:1267-1295; q = 0: chiqw.F:1582-1619)
      h->pocc.clear();
      h->pvir.clear();
      h->intra.clear();
      const int nvir = h->n_vir_win > 0 ? std::min(h->n_vir_win, h->n_vir) : h->n_vir;
      for (int jb = 0; jb < h->n_occ; ++jb)
        for (int ib = 0; ib < nvir; ++ib) {
          const int vb = h->n_docc + ib;
          if (std::fabs(h->eave[vb] - h->eave[jb]) >= h->delta_ex) continue;
          if (h->q0) {
            const int f1 = h->tindx[vb], f2 = h->tindx[jb];
            // intraband: the Drude term instead of a pair (none in cRPA, chiqw.F:1626-1629)
            if (f1 == f2 && f1 != 0) {
              if (h->flg_crpa == 0) h->intra.push_back(jb);
              continue;
            }
          }
          h->pocc.push_back(jb);
          h->pvir.push_back(vb);
        }
      h->npair = (int)h->pocc.size();
      RP_CHECK(h->npair > 0, "no band pairs");
      // |q+G| for the Dyson step (chiqw_sub.F:417-432)
      h->qn.assign(h->ng, 0.0);
      for (int i = 0; i < h->ng; ++i) {
        double v[3];
        for (int d = 0; d < 3; ++d)
          v[d] = (q[0] + (double)h->lg[3 * i]) * h->b[0][d] + (q[1] + (double)h->lg[3 * i + 1]) * h->b[1][d] +
                 (q[2] + (double)h->lg[3 * i + 2]) * h->b[2][d];
        h->qn[i] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      }
      h->lda = roundup((size_t)h->ng, std::max((size_t)chi_tile(), bf16x3_lda_multiple()));
      const size_t K = (size_t)h->npair * ntk;
      h->kpad = roundup(K, (size_t)chi_kchunk());
      hipStream_t s = h->stream;
      h->d_lg.upload(h->lg, s);
      h->d_ikq.upload(h->ikq, s);
      h->d_shift.upload(h->shift, s);
      h->d_pocc.upload(h->pocc, s);
      h->d_pvir.upload(h->pvir, s);
      if (!h->intra.empty()) h->d_intra.upload(h->intra, s);
      h->pdplan = pairdens_plan(h->lg, h->nw[0], h->nw[1], h->nw[2]);
      RP_CHECK(pruned_transform_lds_bytes(h->nw[0], h->nw[1], h->nw[2], h->pdplan[0]) <=
                   lds_budget_bytes(),
               "the chi sphere of this q spans " + std::to_string(h->pdplan[0]) +
                   " z planes of the FFT grid: more than the LDS-resident pair-density transform holds");
      h->d_pdplan.upload(h->pdplan, s);
      h->d_qn.upload(h->qn, s);
      RP_HIP(hipStreamSynchronize(s));
      h->qready = true;
    }
    h->prep_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (ng) *ng = h->ng;
    if (lgout) std::memcpy(lgout, h->lg.data(), sizeof(int) * std::min<size_t>(h->lg.size(), (size_t)3 * lg_cap));
  });
}

int rp_chiqw_set_window(rp_handle h, double ecut_for_eps, int32_t n_vir) {
  return guarded([&] {
    RP_CHECK(h && h->open, "rp_chiqw_set_window: system not open");
    RP_CHECK(ecut_for_eps >= 0.0 && n_vir >= 0, "rp_chiqw_set_window: negative window");
    RP_CHECK(ecut_for_eps <= h->ecut_psi, "rp_chiqw_set_window: chi cutoff above the wavefunction cutoff");
    if (ecut_for_eps != h->ecut_eps_win || n_vir != h->n_vir_win) h->qready = false;
    h->ecut_eps_win = ecut_for_eps;
    h->n_vir_win = n_vir;
  });
}

int rp_chiqw_set_span(rp_handle h, int32_t span) {
  return guarded([&] {
    RP_CHECK(h, "rp_chiqw_set_span: null handle");
    RP_CHECK(span == 0 || (span >= 32 && span % 32 == 0),
             "rp_chiqw_set_span: the span is 0 or a positive multiple of 32");
    h->span = span;
  });
}

int rp_chiqw_npair(rp_handle h, int32_t* npair) {
  return guarded([&] {
    RP_CHECK(h && h->qready, "no q prepared");
    *npair = h->npair;
  });
}

}  // extern "C"

namespace rp {

void throw_status(int st, const std::string& what) {
  if (st & 1) throw Error(what + ": a tetrahedron cut by E_F without partially occupied bands");
  if (st & 8) throw Error(what + ": too many sub-tetrahedra in a Fermi-surface cut");
  if (st & 2) throw Error(what + ": bad tetrahedron degeneracy class (errexec)");
  if (st & 4) throw Error(what + ": singular dielectric matrix (zero pivot)");
}

void check_status(rp_ctx* h, const char* what) {
  int st = 0;
  RP_HIP(hipMemcpyAsync(&st, h->d_status.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  RP_HIP(hipStreamSynchronize(h->stream));
  throw_status(st, what);
}

// ------------------------------------------------------------ stage bodies
// Enqueue only, on h->stream: no host synchronisation, no status read-back and
// no host copies, so that the same code runs eagerly and under graph capture.
// While capturing, every buffer must already be large enough (a reallocation
// would free memory the captured nodes use) and the status word is cleared
// once for the whole graph instead of per stage.
template <class T>
static void grow(const rp_ctx* h, DevBuf<T>& b, size_t m, const char* what) {
  if (h->capturing) {
    RP_CHECK(m == 0 || (b.p && b.n >= m), std::string("captured launch needs a larger ") + what);
    return;
  }
  b.reserve(m);
}

void issue_pairdens(rp_ctx* h) {
  hipStream_t s = h->stream;
  const size_t K = (size_t)h->npair * h->ntk;
  grow(h, h->d_A, h->kpad * h->lda, "pair-density buffer");
  if (h->kpad > K) launch_zero(h->d_A.p + K * h->lda, (h->kpad - K) * h->lda * sizeof(float2), s);
  launch_pair_densities_pruned(h->d_ur.p, h->ntk, h->d_pocc.p, h->d_pvir.p, h->d_ikq.p, h->d_shift.p,
                               h->nw[0], h->nw[1], h->nw[2], h->d_pdplan.p, h->pdplan[0], h->ng,
                               h->lda, (int)K, h->d_tw.p, h->d_A.p, s);
  if (h->q0) {
    grow(h, h->d_A3, 3 * h->kpad * h->lda, "q = 0 pair-density buffer");
    for (int ix = 0; ix < 3; ++ix)
      if (h->kpad > K)
        launch_zero(h->d_A3.p + (ix * h->kpad + K) * h->lda, (h->kpad - K) * h->lda * sizeof(float2), s);
    // three copies laid out [ix][kpad][lda]
    launch_q0_heads(h->d_coef.p, h->d_kg0.p, h->d_ng0.p, h->ntk, h->ntg, h->d_pocc.p, h->d_pvir.p,
                    h->d_eband.p, h->d_bvec.p, h->delt, h->g0, h->ng, h->lda, h->kpad, (int)K,
                    h->d_A.p, h->d_A3.p, s);
  }
}

void issue_weights(rp_ctx* h) {
  hipStream_t s = h->stream;
  grow(h, h->d_w, (size_t)h->nen * h->kpad, "weight buffer");
  launch_zero(h->d_w.p, sizeof(double2) * h->nen * h->kpad, s);
  if (!h->capturing) launch_zero(h->d_status.p, sizeof(int), s);
  launch_tetra_weights(h->ntk, h->nkb[0], h->nkb[1], h->nkb[2], h->d_imt.p, h->d_index.p, h->d_pocc.p,
                       h->d_pvir.p, h->d_ikq.p, h->d_eband.p, h->d_pband.p, h->ef, h->delt, h->dmna,
                       h->dmnr, h->nen, h->d_em.p, h->kpad, h->npair, h->n_pocc > 0 ? 1 : 0,
                       h->d_status.p, h->d_w.p, s);
}

ChiLaunch chi_launch_heuristic(const rp_ctx* h, int prec) {
  ChiLaunch c;
  if (prec == RP_PREC_BF16X3) {
    // 128 x 128 tiles, one persistent workgroup per CU; K split so that the
    // items fill every workgroup evenly
    c.nwg = bf16x3_workgroups(h->ncu);
    c.nsplit = bf16x3_nsplit(h->ng, h->nen, h->kpad, c.nwg);
  } else {
    const int ntile = (h->ng + chi_tile() - 1) / chi_tile();
    const size_t kchunks = h->kpad / chi_kchunk();
    const long wgs = (long)ntile * ntile * h->nen;
    c.nsplit = (int)std::min<long>((long)kchunks, std::max<long>(1, (2048 + wgs - 1) / wgs));
  }
  return c;
}

// The split-BF16 contraction's operands of direction 0: the single-precision
// weights and the presplit octets of the first (or only) pair-density set.
void issue_operands(rp_ctx* h, int prec) {
  if (prec != RP_PREC_BF16X3) return;
  hipStream_t s = h->stream;
  grow(h, h->d_Ao, h->kpad * h->lda * 8, "split operand buffer");
  grow(h, h->d_wf, (size_t)h->nen * h->kpad, "single-precision weight buffer");
  launch_weights_to_f32(h->d_w.p, (size_t)h->nen * h->kpad, h->d_wf.p, s);
  launch_presplit(h->q0 ? h->d_A3.p : h->d_A.p, h->lda, h->kpad, h->d_Ao.p, s);
}

void issue_accumulate(rp_ctx* h, int prec, const ChiLaunch& c, bool events, bool operands_ready) {
  hipStream_t s = h->stream;
  const int ng = h->ng, nen = h->nen, nc = h->q0 ? 3 : 1;
  const size_t per = (size_t)ng * ng * nen;
  grow(h, h->d_chi, (size_t)nc * per, "chi buffer");
  grow(h, h->d_part, (size_t)c.nsplit * per, "partial-sum buffer");
  if (prec == RP_PREC_BF16X3) {
    grow(h, h->d_sync, bf16x3_sync_words(), "rendezvous counters");
    if (!operands_ready) issue_operands(h, prec);
  }
  for (int ix = 0; ix < nc; ++ix) {
    const float2* A = h->q0 ? h->d_A3.p + (size_t)ix * h->kpad * h->lda : h->d_A.p;
    if (prec == RP_PREC_BF16X3 && ix > 0) launch_presplit(A, h->lda, h->kpad, h->d_Ao.p, s);
    if (events) RP_HIP(hipEventRecord(h->gev[2 * ix], s));
    if (prec == RP_PREC_BF16X3)
      launch_chi_bf16x3(h->d_Ao.p, h->lda, h->d_wf.p, h->kpad, h->kpad, ng, nen, c.nsplit, c.nwg,
                        c.span > 0 ? c.span : bf16x3_default_span(),
                        h->d_part.p, h->d_sync.p, s);
    else
      launch_chi_gemm(prec, A, h->lda, h->d_w.p, h->kpad, h->kpad, ng, nen, c.nsplit, h->d_part.p, s);
    if (events) RP_HIP(hipEventRecord(h->gev[2 * ix + 1], s));
    launch_chi_reduce(h->d_part.p, per, c.nsplit, 1.0 / h->volume, h->d_chi.p + (size_t)ix * per, s);
  }
}

void issue_dyson(rp_ctx* h) {
  hipStream_t s = h->stream;
  const int n = h->ng, nb = h->nbatch();
  const size_t per = (size_t)n * n;
  grow(h, h->d_eps, (size_t)nb * per, "eps buffer");
  grow(h, h->d_gjwork, gauss_jordan_work_ints(nb), "pivot work buffer");
  grow(h, h->d_ipiv, (size_t)nb * n, "pivot buffer");
  grow(h, h->d_wd, 3, "plasma-frequency buffer");
  if (!h->capturing) launch_zero(h->d_status.p, sizeof(int), s);
  if (h->q0 && !h->intra.empty()) {
    // Drude plasma frequency of the intraband bands (fRPA metals at q = 0)
    grow(h, h->d_drude, 3 * h->intra.size(), "Drude buffer");
    launch_drude(h->d_coef.p, h->d_kg0.p, h->d_ng0.p, h->ntk, h->ntg, h->d_intra.p,
                 (int)h->intra.size(), h->d_sk0.p, h->d_bvec.p, h->d_eband.p, h->ef, h->delt,
                 h->dmna, h->dmnr, h->nkb[0], h->nkb[1], h->nkb[2], h->d_imt.p, h->d_index.p,
                 h->volume, h->d_status.p, h->d_drude.p, h->d_wd.p, s);
  } else {
    launch_zero(h->d_wd.p, 3 * sizeof(double), s);  // no intraband pairs (cRPA, or an insulator)
  }
  launch_eps_build(h->d_chi.p, h->d_qn.p, n, nb, h->nen, h->q0 ? 1 : 0, h->g0, h->d_wd.p, h->d_em.p,
                   h->delt, h->d_eps.p, s);
  launch_gauss_jordan(h->d_eps.p, n, nb, h->d_ipiv.p, h->d_gjwork.p, h->d_status.p, s);
}

// ---------------------------------------------------------------- autotune
// Candidates around the heuristic: for the split-BF16 kernel the K splits
// whose work rounds per XCD slot (the heuristic's model) are within 15% of the
// best, plus half and double the heuristic's; for the FP64/FP32 kernel the
// splits that give 1024, 2048 and 4096 workgroups.  Each is timed on the
// current q's data (one warm-up, then the best of two), and the fastest is
// kept for every later q of the same shape.
static std::vector<ChiLaunch> chi_candidates(const rp_ctx* h, int prec) {
  std::vector<ChiLaunch> out;
  const ChiLaunch base = chi_launch_heuristic(h, prec);
  auto add = [&](ChiLaunch c) {
    for (const auto& o : out)
      if (o.nsplit == c.nsplit && o.nwg == c.nwg) return;
    out.push_back(c);
  };
  add(base);
  if (prec == RP_PREC_BF16X3) {
    const size_t nchunks = h->kpad / 32;
    const int ntile = (h->ng + 127) / 128;
    const long mypairs = ((long)ntile * ntile * h->nen + 7) / 8;
    const long nslot = std::max(1, base.nwg / 8);
    std::vector<std::pair<double, int>> r;
    for (int sp = 1; sp <= 64 && (size_t)sp <= nchunks; ++sp)
      r.push_back({(double)((mypairs * sp + nslot - 1) / nslot) / sp, sp});
    std::sort(r.begin(), r.end());
    for (const auto& e : r)
      if (e.first <= 1.15 * r[0].first && out.size() < 4) add({e.second, base.nwg});
    if (base.nsplit > 1) add({base.nsplit / 2, base.nwg});
    if ((size_t)base.nsplit * 2 <= nchunks && base.nsplit * 2 <= 64) add({base.nsplit * 2, base.nwg});
  } else {
    const int ntile = (h->ng + chi_tile() - 1) / chi_tile();
    const size_t kchunks = h->kpad / chi_kchunk();
    const long wgs = (long)ntile * ntile * h->nen;
    for (long target : {1024L, 2048L, 4096L})
      add({(int)std::min<long>((long)kchunks, std::max<long>(1, (target + wgs - 1) / wgs)), 0});
  }
  return out;
}

// seconds between two recorded (and completed) events
static double event_seconds(hipEvent_t a, hipEvent_t b) {
  float ms = 0.0f;
  RP_HIP(hipEventElapsedTime(&ms, a, b));
  return 1e-3 * ms;
}

static double time_launch(rp_ctx* h, int prec, const ChiLaunch& c) {
  issue_accumulate(h, prec, c, false);  // warm-up (and buffer growth)
  double best = 1e300;
  for (int rep = 0; rep < 2; ++rep) {
    RP_HIP(hipEventRecord(h->ev[4], h->stream));
    issue_accumulate(h, prec, c, false);
    RP_HIP(hipEventRecord(h->ev[5], h->stream));
    RP_HIP(hipEventSynchronize(h->ev[5]));
    float ms = 0;
    RP_HIP(hipEventElapsedTime(&ms, h->ev[4], h->ev[5]));
    best = std::min(best, (double)ms);
  }
  return best;
}

ChiLaunch chi_launch(rp_ctx* h, int prec) {
  ChiLaunch c = chi_launch_tuned(h, prec);
  if (prec == RP_PREC_BF16X3) c.span = h->span > 0 ? h->span : h->q_span;
  return c;
}

double epsinv_kappa(rp_ctx* h) {
  const int nb = h->nbatch();
  h->d_kappa.reserve(nb);
  launch_norm1(h->d_eps.p, h->ng, nb, h->d_kappa.p, h->stream);
  std::vector<double> k(nb);
  RP_HIP(hipMemcpyAsync(k.data(), h->d_kappa.p, sizeof(double) * nb, hipMemcpyDeviceToHost, h->stream));
  RP_HIP(hipStreamSynchronize(h->stream));
  return *std::max_element(k.begin(), k.end());
}

int adapt_span(rp_ctx* h, int prec, int span_used) {
  if (prec != RP_PREC_BF16X3 || h->span > 0) return 0;
  h->last_kappa = epsinv_kappa(h);
  const int sp = bf16x3_span_for_kappa(h->last_kappa);
  const int used = span_used > 0 ? span_used : bf16x3_default_span();
  return sp < used ? sp : 0;
}

ChiLaunch chi_launch_tuned(rp_ctx* h, int prec) {
  const int ndir = h->q0 ? 3 : 1;
  for (const auto& t : h->tune)
    if (t.prec == prec && t.ng == h->ng && t.nen == h->nen && t.kpad == h->kpad && t.ndir == ndir)
      return t.best;
  const ChiLaunch base = chi_launch_heuristic(h, prec);
  if (!h->autotune || h->capturing || !(h->have_A && h->have_w)) return base;
  TuneEntry t;
  t.prec = prec;
  t.ng = h->ng;
  t.nen = h->nen;
  t.kpad = h->kpad;
  t.ndir = ndir;
  t.heuristic = base;
  const auto cand = chi_candidates(h, prec);
  t.ncand = (int)cand.size();
  t.best_ms = 1e300;
  for (const auto& c : cand) {
    const double ms = time_launch(h, prec, c);
    if (c.nsplit == base.nsplit && c.nwg == base.nwg) t.heuristic_ms = ms;
    if (ms < t.best_ms) {
      t.best_ms = ms;
      t.best = c;
    }
  }
  h->tune.push_back(t);
  return t.best;
}

}  // namespace rp

extern "C" {

int rp_pairdens_batch(rp_handle h, float* m_out) {
  return guarded([&] {
    RP_CHECK(h && h->qready, "rp_pairdens_batch: no q prepared");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const size_t K = (size_t)h->npair * h->ntk;
    issue_pairdens(h);
    if (m_out) {
      const int nc = h->q0 ? 3 : 1;
      for (int ix = 0; ix < nc; ++ix) {
        const float2* src = h->q0 ? h->d_A3.p + (size_t)ix * h->kpad * h->lda : h->d_A.p;
        RP_HIP(hipMemcpy2DAsync(m_out + 2 * (size_t)ix * K * h->ng, sizeof(float2) * h->ng, src,
                                sizeof(float2) * h->lda, sizeof(float2) * h->ng, K, hipMemcpyDeviceToHost, s));
      }
    }
    RP_HIP(hipStreamSynchronize(s));
    h->have_A = true;
  });
}

int rp_tetra_weights_batch(rp_handle h, double* d_out) {
  return guarded([&] {
    RP_CHECK(h && h->qready, "rp_tetra_weights_batch: no q prepared");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const size_t K = (size_t)h->npair * h->ntk;
    issue_weights(h);
    check_status(h, "tetrahedron weights");
    if (d_out)
      RP_HIP(hipMemcpy2DAsync(d_out, sizeof(double2) * K, h->d_w.p, sizeof(double2) * h->kpad,
                              sizeof(double2) * K, h->nen, hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
    h->have_w = true;
  });
}

int rp_chi_accumulate(rp_handle h, int32_t precision, double* chi_out) {
  return guarded([&] {
    RP_CHECK(h && h->qready && h->have_A && h->have_w, "rp_chi_accumulate: pair densities/weights missing");
    RP_CHECK(precision >= RP_PREC_FP64 && precision <= RP_PREC_BF16X3, "unknown precision");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    const int nc = h->q0 ? 3 : 1;
    const ChiLaunch c = chi_launch(h, precision);
    issue_accumulate(h, precision, c, true);
    if (chi_out)
      RP_HIP(hipMemcpyAsync(chi_out, h->d_chi.p, sizeof(double2) * nc * (size_t)h->ng * h->ng * h->nen,
                            hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
    h->gemm_seconds = 0;
    h->gemm_launches = nc;
    for (int ix = 0; ix < nc; ++ix) {
      float ms;
      RP_HIP(hipEventElapsedTime(&ms, h->gev[2 * ix], h->gev[2 * ix + 1]));
      h->gemm_seconds += ms * 1e-3;
    }
    h->have_chi = true;
  });
}

int rp_chi_set(rp_handle h, const double* chi_in) {
  return guarded([&] {
    RP_CHECK(h && h->qready, "rp_chi_set: no q prepared");
    RP_CHECK(chi_in, "rp_chi_set: null input");
    RP_HIP(hipSetDevice(h->device));
    const size_t n = (size_t)h->ng * h->ng * h->nen * (h->q0 ? 3 : 1);
    h->d_chi.reserve(n);
    RP_HIP(hipMemcpy(h->d_chi.p, chi_in, sizeof(double2) * n, hipMemcpyHostToDevice));
    h->have_chi = true;
  });
}

int rp_dyson(rp_handle h, double* epsinv_out) {
  return guarded([&] {
    RP_CHECK(h && h->qready && h->have_chi, "rp_dyson: chi missing");
    RP_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    issue_dyson(h);
    check_status(h, "Dyson step");
    if (epsinv_out)
      RP_HIP(hipMemcpyAsync(epsinv_out, h->d_eps.p, sizeof(double2) * h->nbatch() * (size_t)h->ng * h->ng,
                            hipMemcpyDeviceToHost, s));
    RP_HIP(hipStreamSynchronize(s));
  });
}

int rp_chiqw_q(rp_handle h, const double* q, int32_t precision, double* epsinv_out, double* chi_out,
               double* timings) {
  int ng = 0;
  if (rp_chiqw_prepare_q(h, q, &ng, nullptr, 0)) return 1;
  return guarded([&] {
    hipStream_t s = h->stream;
    RP_HIP(hipEventRecord(h->ev[0], s));
    if (rp_pairdens_batch(h, nullptr)) throw Error(g_err);
    RP_HIP(hipEventRecord(h->ev[1], s));
    if (rp_tetra_weights_batch(h, nullptr)) throw Error(g_err);
    RP_HIP(hipEventRecord(h->ev[2], s));
    if (rp_chi_accumulate(h, precision, chi_out)) throw Error(g_err);
    RP_HIP(hipEventRecord(h->ev[3], s));
    if (rp_dyson(h, epsinv_out)) throw Error(g_err);
    RP_HIP(hipEventRecord(h->ev[4], s));
    // an ill-conditioned eps amplifies the FP32 accumulation error of the
    // split-BF16 contraction: run it again with the span its eps^-1 calls for
    // (its contraction and Dyson times are added to those stages)
    const int sp = adapt_span(h, precision, h->q_span);
    RP_HIP(hipEventSynchronize(h->ev[4]));
    double st[5];
    for (int i = 0; i < 4; ++i) st[i] = event_seconds(h->ev[i], h->ev[i + 1]);
    st[4] = event_seconds(h->ev[0], h->ev[4]);
    if (sp) {
      h->q_span = sp;
      RP_HIP(hipEventRecord(h->ev[1], s));
      if (rp_chi_accumulate(h, precision, chi_out)) throw Error(g_err);
      RP_HIP(hipEventRecord(h->ev[2], s));
      if (rp_dyson(h, epsinv_out)) throw Error(g_err);
      RP_HIP(hipEventRecord(h->ev[3], s));
      RP_HIP(hipEventSynchronize(h->ev[3]));
      const double ta = event_seconds(h->ev[1], h->ev[2]), td = event_seconds(h->ev[2], h->ev[3]);
      st[2] += ta;
      st[3] += td;
      st[4] += ta + td;
    }
    if (timings) {
      for (int i = 0; i < 5; ++i) timings[i] = st[i];
      timings[5] = h->prep_seconds;
      timings[6] = h->gemm_seconds;
      timings[7] = h->gemm_launches;
    }
  });
}

}  // extern "C"
