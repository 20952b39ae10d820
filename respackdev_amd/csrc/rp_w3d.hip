// calc_int (calc_w3d) on gfx950: the Wannier-basis bare and screened Coulomb
// integrals V_ij(R), W_ij(R,w) of calc_intW.F:636-985.
//
// rho_w(q+G) = (1/NTK) sum_k M_ww,k(G) / |q+G| comes from the pair-density
// kernels of rp_pairdens.hip (k_pair_rsum, k_rho_from_grid: the products summed
// over k on the grid, one transform per (q, pair)).  Then per full-mesh q (all
// q of one launch):
//   func_ij(q)     = sum_{G in psi} rho_i(G) conj(rho_j(G))                (k_w_func)
//   funcw_ij(q,w)  = sum_{G,G' in eps} rho_i(G) eps^-1_q(G,G',w) conj(rho_j(G'))
//                    + sum_{G in psi \ eps} rho_i(G) conj(rho_j(G))        (k_w_contract)
// where eps^-1_q is the irreducible matrix (complex64, as the reference stores
// it) seen through the rotation of make_eps (calc_int_sub.F:222-311): an index
// map and a phase per G, folded into the rho vectors so that eps^-1 is read
// once, coalesced, straight from its irreducible copy.  Finally the lattice
// Fourier sums with the Hybertsen-Louie heads (k_w_fourier).
#include "rp_common.h"
#include "rp_kernels.h"

namespace rp {

// Per q: func (all G of the psi list) and the shell part of funcw (G beyond the
// eps sphere).  grid: ntq blocks of 256; out [q][j][i].
__global__ void __launch_bounds__(256) k_w_func(const double2* __restrict__ rho,
                                                const long long* __restrict__ rho_off,
                                                const int4* __restrict__ qinfo, int nwf,
                                                double2* __restrict__ func,
                                                double2* __restrict__ shell) {
  const int q = blockIdx.x;
  const int4 qi = qinfo[q];  // (ng_eps, ng_psi, iq_irr, trs)
  const double2* r = rho + rho_off[q];
  const int ngpsi = qi.y, nge = qi.x;
  __shared__ double red[256];
  for (int ij = 0; ij < nwf * nwf; ++ij) {
    const int i = ij % nwf, j = ij / nwf;
    const double2* ri = r + (size_t)i * ngpsi;
    const double2* rj = r + (size_t)j * ngpsi;
    double f[4] = {0, 0, 0, 0};  // func re/im, shell re/im
    for (int g = threadIdx.x; g < ngpsi; g += blockDim.x) {
      const double2 a = ri[g], b = rj[g];
      const double pr = a.x * b.x + a.y * b.y, pi = a.y * b.x - a.x * b.y;  // a conj(b)
      f[0] += pr;
      f[1] += pi;
      if (g >= nge) { f[2] += pr; f[3] += pi; }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      red[threadIdx.x] = f[c];
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        double2* o = (c < 2 ? func : shell) + ((size_t)q * nwf + j) * nwf + i;
        if (c % 2 == 0) o->x = red[0];
        else o->y = red[0];
      }
      __syncthreads();
    }
  }
}

// funcw for one w and a group of q that share one irreducible eps^-1 E
// (complex64, column-major): per member q, Y(a,c) = sum_b E(a,b) X_c(b), then
// funcw = sum_a Z(a) Y(a,.) + shell, where
//   trs = +1: X_j(b) = conj(rho_j(m^-1 b)) pf2, Z = rho_i(m^-1 a) pf1
//   trs = -1: X_i(b) = rho_i(m^-1 b) pf1,       Z = conj(rho_j(m^-1 a)) pf2
// with the member's own map m and phases.  The vectors are the NV <= 16 Wannier
// functions of calc_w3d per member, 16 / NV members per group at most, and
// every (i,j) of a member is formed, out [q][w][j][i] (calc_j3d's diagonal
// pairs: k_j_contract).  A workgroup takes all the vectors of its group, so E
// is read once per (group, w).  Y = E X runs on the FP64 matrix cores as in
// k_j_contract (Re Y = Er Xr - Ei Xi, Im Y = Er Xi + Ei Xr, v_mfma_f64_16x16x4,
// the group's vectors as one 16-column block): lane l supplies A(l % 16, l / 16)
// and B(l / 16, l % 16) and holds D(l / 16 + 4 v, l % 16).  X of every group
// is formed once (k_w_xvec, independent of w).  A wave takes one 16-row block
// of E per pass over all columns; then each lane multiplies its four Y(a, c)
// by the Z_i(a) of c's member and keeps NV sums, reduced over the lanes of
// c and the waves at the end.
// grid: (nen, ngroups), 512 threads; group g has the q gmem[goff[g] .. goff[g+1]).
typedef double jd4 __attribute__((ext_vector_type(4)));
constexpr int kMaxWf = 16;     // Wannier functions of calc_w3d and calc_j3d
constexpr int kWcVec = 16;     // vectors of a group: one MFMA column block
constexpr int kWA = 3;         // row blocks per wave and pass of the W contraction
template <int NV>
constexpr int wc_members() { return kWcVec / NV; }
int w_contract_members(int nwf) { return kWcVec / nwf; }
int w_contract_vectors(int) { return kWcVec; }

// X_c(b) of every group, [nge][vb] at xw + vb * map_off[first member] (zero
// past the group's vectors).  grid: ngroups, 256 threads.
__global__ void __launch_bounds__(256) k_w_xvec(const double2* __restrict__ rho,
                                                const long long* __restrict__ rho_off,
                                                const int4* __restrict__ qinfo,
                                                const int* __restrict__ minv,
                                                const double2* __restrict__ pf1,
                                                const double2* __restrict__ pf2,
                                                const long long* __restrict__ map_off,
                                                const int* __restrict__ goff,
                                                const int* __restrict__ gmem, int nv, int vb,
                                                double2* __restrict__ xw) {
  const int grp = blockIdx.x;
  const int mbeg = goff[grp], nvt = (goff[grp + 1] - mbeg) * nv;
  const int q0 = gmem[mbeg];
  const int nge = qinfo[q0].x;
  double2* out = xw + map_off[q0] * vb;
  for (int e = threadIdx.x; e < vb * nge; e += blockDim.x) {
    const int b = e / vb, c = e % vb;
    double2 v = make_double2(0.0, 0.0);
    if (c < nvt) {
      const int q = gmem[mbeg + c / nv];
      const int4 qi = qinfo[q];
      const long long mo = map_off[q];
      const int g = minv[mo + b];
      const double2 rv = rho[rho_off[q] + (size_t)(c % nv) * qi.y + g];
      const double2 p = qi.w == 1 ? pf2[mo + g] : pf1[mo + g];
      // trs = +1: conj(rho) pf2; trs = -1: rho pf1
      const double xr = rv.x, xi = qi.w == 1 ? -rv.y : rv.y;
      v = make_double2(xr * p.x - xi * p.y, xr * p.y + xi * p.x);
    }
    out[e] = v;
  }
}

template <int NV>
__global__ void __launch_bounds__(512) k_w_contract(
    const float2* __restrict__ eps, const long long* __restrict__ eps_off,
    const double2* __restrict__ rho, const long long* __restrict__ rho_off,
    const int4* __restrict__ qinfo, const int* __restrict__ minv,
    const double2* __restrict__ pf1, const double2* __restrict__ pf2,
    const long long* __restrict__ map_off, const double2* __restrict__ shell,
    const int* __restrict__ goff, const int* __restrict__ gmem,
    const double2* __restrict__ xw, int nen, double2* __restrict__ funcw) {
  const int w = blockIdx.x, grp = blockIdx.y;
  const int mbeg = goff[grp], nm = goff[grp + 1] - mbeg;  // 1 <= nm <= 16 / NV
  const int nvt = nm * NV;
  const int q0m = gmem[mbeg];
  const int4 qi0 = qinfo[q0m];
  const int nge = qi0.x, iqir = qi0.z;
  const int ld = w3d_eps_ld(nge);
  const float2* E = eps + eps_off[iqir] + (size_t)w * ld * nge;
  const double2* Xg = xw + map_off[q0m] * kWcVec;  // (b, c) at Xg[b * 16 + c]
  __shared__ double red[kWcVec][NV][8][2];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nab = (nge + 15) / 16;
  // this lane's vector c = li: member li / NV, its NV sums Z_i Y_c
  const bool cv = li < nvt;
  const int qm = gmem[mbeg + (cv ? li / NV : 0)];
  const int4 qim = qinfo[qm];
  const long long mom = map_off[qm];
  const double2* rm = rho + rho_off[qm];
  double sre[NV], sim[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) sre[i] = sim[i] = 0.0;
  // up to kWA row blocks per wave and pass (ab = a0 + wave + 8 j) share each
  // load of X; blocks past the last are skipped (uniform per wave)
  for (int a0 = 0; a0 < nab; a0 += 8 * kWA) {
    jd4 accr[kWA], acci[kWA];
    int arow[kWA];
#pragma unroll
    for (int j = 0; j < kWA; ++j) {
      accr[j] = acci[j] = jd4{0.0, 0.0, 0.0, 0.0};
      // rows nge .. ld-1 are the zero padding; their sums are never used
      arow[j] = (a0 + wave + 8 * j) * 16 + li;
    }
    const int nj = min(kWA, (nab - a0 - wave + 7) / 8);  // live blocks of this wave
    // the next step's X and E are loaded during this step's MFMAs (the last
    // step loads itself again); columns past nge contribute zero
    int bn = lk < nge ? lk : nge - 1;
    double2 xn = Xg[(size_t)bn * kWcVec + li];
    float2 en[kWA];
#pragma unroll
    for (int j = 0; j < kWA; ++j) en[j] = j < nj ? E[(size_t)bn * ld + arow[j]] : make_float2(0.f, 0.f);
    for (int b0 = 0; b0 < nge; b0 += 4) {
      const bool bv = b0 + lk < nge;
      const double2 xv = xn;
      float2 ec[kWA];
#pragma unroll
      for (int j = 0; j < kWA; ++j) ec[j] = en[j];
      const int b1 = (b0 + 4 < nge ? b0 + 4 : b0) + lk;
      bn = b1 < nge ? b1 : nge - 1;
      xn = Xg[(size_t)bn * kWcVec + li];
#pragma unroll
      for (int j = 0; j < kWA; ++j)
        if (j < nj) en[j] = E[(size_t)bn * ld + arow[j]];
      const double xr = bv ? xv.x : 0.0, xi = bv ? xv.y : 0.0, nxi = -xi;
#pragma unroll
      for (int j = 0; j < kWA; ++j) {
        if (j >= nj) break;  // uniform
        const double er = ec[j].x, ei = ec[j].y;
        accr[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, xr, accr[j], 0, 0, 0);
        accr[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei, nxi, accr[j], 0, 0, 0);
        acci[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, xi, acci[j], 0, 0, 0);
        acci[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei, xr, acci[j], 0, 0, 0);
      }
    }
    // Z_i(a) Y(a, c) for this lane's rows a = 16 ab + lk + 4 v and vector c = li
    if (cv) {
#pragma unroll
      for (int j = 0; j < kWA; ++j) {
        if (j >= nj) break;  // uniform
        const int ab = a0 + wave + 8 * j;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ar = ab * 16 + lk + 4 * v;
          if (ar >= nge) continue;
          const int g = minv[mom + ar];
          const double yr = accr[j][v], yi = acci[j][v];
          const double2 pp = qim.w == 1 ? pf1[mom + g] : pf2[mom + g];
#pragma unroll
          for (int i = 0; i < NV; ++i) {
            const double2 rv = rm[(size_t)i * qim.y + g];
            double zr, zi;
            if (qim.w == 1) {  // Z = rho pf1
              zr = rv.x * pp.x - rv.y * pp.y;
              zi = rv.x * pp.y + rv.y * pp.x;
            } else {           // Z = conj(rho) pf2
              zr = rv.x * pp.x + rv.y * pp.y;
              zi = rv.x * pp.y - rv.y * pp.x;
            }
            sre[i] += zr * yr - zi * yi;
            sim[i] += zr * yi + zi * yr;
          }
        }
      }
    }
  }
  // lanes of one vector (li) hold their rows' sums: over lk, then the waves
#pragma unroll
  for (int i = 0; i < NV; ++i)
    for (int off = 16; off < 64; off <<= 1) {
      sre[i] += __shfl_xor(sre[i], off);
      sim[i] += __shfl_xor(sim[i], off);
    }
  if (lk == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      red[li][i][wave][0] = sre[i];
      red[li][i][wave][1] = sim[i];
    }
  }
  __syncthreads();
  for (int e = t; e < 2 * nvt * NV; e += blockDim.x) {
    const int part = e & 1, idx = e >> 1;
    const int c = idx / NV, zi = idx % NV;
    const int m = c / NV, yc = c % NV;
    const int q = gmem[mbeg + m];
    // trs = +1 pairs Z_i Y_j, trs = -1 pairs Z_j Y_i
    const bool tp = qinfo[q].w == 1;
    const int i = tp ? zi : yc, j = tp ? yc : zi;
    double sum = 0.0;
#pragma unroll
    for (int wv = 0; wv < 8; ++wv) sum += red[c][zi][wv][part];
    const double2 sh = shell[((size_t)q * NV + j) * NV + i];
    double2* o = funcw + (((size_t)q * nen + w) * NV + j) * NV + i;
    if (part == 0) o->x = sum + sh.x;
    else o->y = sum + sh.y;
  }
}

// calc_j3d's contraction on the FP64 matrix cores: funcw_pp of the diagonal
// pairs, by k_w_contract's formula, with Y = E X as real products v_mfma_f64_16x16x4 (Re Y =
// Er Xr - Ei Xi, Im Y = Er Xi + Ei Xr) over 16-row blocks of E and 16-column
// blocks of up to 32 pair vectors (blockIdx.z picks the 32).  X comes from
// k_j_xvec, [b][32 groups] in HBM (zero past the last pair).  A wave takes up
// to kJA row blocks per pass (wave + 8 j) and the passes cover every row, so
// a workgroup reads E once: E is read once per (q, w) for up to 32 pair vectors
// and once per (q, w, group of 32) beyond (NWF >= 6).  Lane
// l supplies A(l % 16, l / 16) and B(l / 16, l % 16) and holds
// D(l / 16 + 4 v, l % 16), v < 4 (checked on the GPU with row- and column-numbered
// operands).  grid: (nen, ntq, groups), 512 threads.
constexpr int kJA = 1;
__global__ void __launch_bounds__(256) k_j_xvec(const double2* __restrict__ rho,
                                                const long long* __restrict__ rho_off,
                                                const int4* __restrict__ qinfo,
                                                const int* __restrict__ minv,
                                                const double2* __restrict__ pf1,
                                                const double2* __restrict__ pf2,
                                                const long long* __restrict__ map_off, int nv,
                                                int g32, double2* __restrict__ xg) {
  const int q = blockIdx.x;
  const int4 qi = qinfo[q];
  const int nge = qi.x, ngpsi = qi.y, trs = qi.w;
  const double2* r = rho + rho_off[q];
  const int* mi = minv + map_off[q];
  const double2* p1 = pf1 + map_off[q];
  const double2* p2 = pf2 + map_off[q];
  double2* out = xg + map_off[q] * g32;
  for (int e = threadIdx.x; e < nge * g32; e += blockDim.x) {
    const int b = e / g32, c = e % g32;
    double2 v = make_double2(0.0, 0.0);
    if (c < nv) {
      const int g = mi[b];
      const double2 rv = r[(size_t)c * ngpsi + g];
      const double2 p = trs == 1 ? p2[g] : p1[g];
      // trs = +1: conj(rho) pf2; trs = -1: rho pf1
      const double xr = rv.x, xi = trs == 1 ? -rv.y : rv.y;
      v = make_double2(xr * p.x - xi * p.y, xr * p.y + xi * p.x);
    }
    out[e] = v;
  }
}

__global__ void __launch_bounds__(512) k_j_contract(
    const float2* __restrict__ eps, const long long* __restrict__ eps_off,
    const double2* __restrict__ rho, const long long* __restrict__ rho_off,
    const int4* __restrict__ qinfo, const int* __restrict__ minv,
    const double2* __restrict__ pf1, const double2* __restrict__ pf2,
    const long long* __restrict__ map_off, const double2* __restrict__ xg, int g32,
    const double2* __restrict__ shell, int nv, int nen, double2* __restrict__ funcw) {
  const int w = blockIdx.x, q = blockIdx.y, grp = blockIdx.z;
  const int4 qi = qinfo[q];
  const int nge = qi.x, ngpsi = qi.y, iqir = qi.z, trs = qi.w;
  const int ld = w3d_eps_ld(nge);
  const float2* E = eps + eps_off[iqir] + (size_t)w * ld * nge;
  const int c0 = grp * 32, nvb = min(32, nv - c0);
  const double2* Xq = xg + map_off[q] * g32 + c0;  // (b, c) at Xq[b * g32 + c]
  const double2* r = rho + rho_off[q] + (size_t)c0 * ngpsi;
  const int* mi = minv + map_off[q];
  const double2* p1 = pf1 + map_off[q];
  const double2* p2 = pf2 + map_off[q];
  __shared__ double red[32][8][2];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nab = (nge + 15) / 16;
  double sre[2] = {0.0, 0.0}, sim[2] = {0.0, 0.0};  // this lane's sums for c = 16 cb + li
  for (int a0 = 0; a0 < nab; a0 += 8 * kJA) {
    jd4 accr[kJA][2], acci[kJA][2];
    int arow[kJA];
#pragma unroll
    for (int j = 0; j < kJA; ++j) {
      // rows nge .. ld-1 are the zero padding, rows past ld (row blocks past
      // the last) read row 0; their sums are never used
      const int a = (a0 + wave + 8 * j) * 16 + li;
      arow[j] = a < ld ? a : 0;
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) accr[j][cb] = acci[j][cb] = jd4{0.0, 0.0, 0.0, 0.0};
    }
    // the next step's X and E are loaded during this step's MFMAs (the last
    // step loads itself again); columns past nge contribute zero
    int bn = lk < nge ? lk : nge - 1;
    double2 xn[2];
    float2 en[kJA];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) xn[cb] = Xq[(size_t)bn * g32 + cb * 16 + li];
#pragma unroll
    for (int j = 0; j < kJA; ++j) en[j] = E[(size_t)bn * ld + arow[j]];
    for (int b0 = 0; b0 < nge; b0 += 4) {
      const bool bv = b0 + lk < nge;
      double xr[2], xi[2], nxi[2];
      float2 ec[kJA];
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        xr[cb] = bv ? xn[cb].x : 0.0;
        xi[cb] = bv ? xn[cb].y : 0.0;
        nxi[cb] = -xi[cb];
      }
#pragma unroll
      for (int j = 0; j < kJA; ++j) ec[j] = en[j];
      const int b1 = (b0 + 4 < nge ? b0 + 4 : b0) + lk;
      bn = b1 < nge ? b1 : nge - 1;
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) xn[cb] = Xq[(size_t)bn * g32 + cb * 16 + li];
#pragma unroll
      for (int j = 0; j < kJA; ++j) en[j] = E[(size_t)bn * ld + arow[j]];
#pragma unroll
      for (int j = 0; j < kJA; ++j) {
        const double er = ec[j].x, ei = ec[j].y;
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          accr[j][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, xr[cb], accr[j][cb], 0, 0, 0);
          accr[j][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei, nxi[cb], accr[j][cb], 0, 0, 0);
          acci[j][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(er, xi[cb], acci[j][cb], 0, 0, 0);
          acci[j][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(ei, xr[cb], acci[j][cb], 0, 0, 0);
        }
      }
    }
    // Z_c(a) Y(a, c) for this lane's rows a = 16 (row block) + lk + 4 v and columns c
#pragma unroll
    for (int j = 0; j < kJA; ++j) {
      const int ab = a0 + wave + 8 * j;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int a = ab * 16 + lk + 4 * v;
        if (ab >= nab || a >= nge) continue;
        const int g = mi[a];
        const double2 pp = trs == 1 ? p1[g] : p2[g];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) {
          const int c = cb * 16 + li;
          if (c >= nvb) continue;
          const double2 rv = r[(size_t)c * ngpsi + g];
          double zr, zi;
          if (trs == 1) {  // Z = rho pf1
            zr = rv.x * pp.x - rv.y * pp.y;
            zi = rv.x * pp.y + rv.y * pp.x;
          } else {         // Z = conj(rho) pf2
            zr = rv.x * pp.x + rv.y * pp.y;
            zi = rv.x * pp.y - rv.y * pp.x;
          }
          const double yr = accr[j][cb][v], yi = acci[j][cb][v];
          sre[cb] += zr * yr - zi * yi;
          sim[cb] += zr * yi + zi * yr;
        }
      }
    }
  }
  // lanes of one column (li) hold their rows' sums: over lk, then the waves
#pragma unroll
  for (int cb = 0; cb < 2; ++cb)
    for (int off = 16; off < 64; off <<= 1) {
      sre[cb] += __shfl_xor(sre[cb], off);
      sim[cb] += __shfl_xor(sim[cb], off);
    }
  if (lk == 0) {
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      red[cb * 16 + li][wave][0] = sre[cb];
      red[cb * 16 + li][wave][1] = sim[cb];
    }
  }
  __syncthreads();
  for (int e = t; e < 2 * nvb; e += blockDim.x) {
    const int p = e >> 1, part = e & 1;
    double sum = 0.0;
#pragma unroll
    for (int wv = 0; wv < 8; ++wv) sum += red[p][wv][part];
    const double2 sh = shell[(size_t)q * nv + c0 + p];
    double2* o = funcw + ((size_t)q * nen + w) * nv + c0 + p;
    if (part == 0) o->x = sum + sh.x;
    else o->y = sum + sh.y;
  }
}

// calc_j3d func / shell (calc_intJ.F:680-690): per (q, p) sum_G |rho_p(G)|^2 as
// rho conj(rho).  grid: (ntq, npair), 256 threads; out [q][p].
__global__ void __launch_bounds__(256) k_j_func(const double2* __restrict__ rho,
                                                const long long* __restrict__ rho_off,
                                                const int4* __restrict__ qinfo, int np,
                                                double2* __restrict__ func,
                                                double2* __restrict__ shell) {
  const int q = blockIdx.x, p = blockIdx.y;
  const int4 qi = qinfo[q];
  const int nge = qi.x, ngpsi = qi.y;
  const double2* r = rho + rho_off[q] + (size_t)p * ngpsi;
  __shared__ double red[256];
  double f[4] = {0, 0, 0, 0};
  for (int g = threadIdx.x; g < ngpsi; g += blockDim.x) {
    const double2 a = r[g];
    const double pr = a.x * a.x + a.y * a.y, pi = a.y * a.x - a.x * a.y;
    f[0] += pr;
    f[1] += pi;
    if (g >= nge) { f[2] += pr; f[3] += pi; }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    red[threadIdx.x] = f[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      double2* o = (c < 2 ? func : shell) + (size_t)q * np + p;
      if (c % 2 == 0) o->x = red[0];
      else o->y = red[0];
    }
    __syncthreads();
  }
}

// X_ij(R) (slot w = nen) and J_ij(R, w) = 2*2pi sum_q f(q) / V / NTQ, heads on
// i = j at R = 0 (calc_intJ.F:830-904).  One R: grid (nen + 1), threads over p.
// out: X [p] and J [w * wstride + p] at this R's slot.
__global__ void __launch_bounds__(64) k_j_sum(const double2* __restrict__ func,
                                              const double2* __restrict__ funcw, int ntq, int nwf,
                                              int nen, int r0, double volume, double chead,
                                              const double2* __restrict__ cheadw,
                                              double2* __restrict__ xmat,
                                              double2* __restrict__ jmat, size_t wstride) {
  const int w = blockIdx.x;
  const bool isx = w == nen;
  const int np = nwf * nwf;
  const double tpi = 2.0 * 3.141592653589793238462643383279502884;
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int q = 0; q < ntq; ++q) {
      const double2 f = isx ? func[(size_t)q * np + p] : funcw[((size_t)q * nen + w) * np + p];
      sr += f.x;
      si += f.y;
    }
    double xr = 2.0 * tpi * sr / volume / (double)ntq, xi = 2.0 * tpi * si / volume / (double)ntq;
    const bool diag = r0 && (p % nwf) == (p / nwf);
    if (isx) {
      if (diag) xr += chead;
      xmat[p] = make_double2(xr, xi);
    } else {
      if (diag) { xr += cheadw[w].x; xi += cheadw[w].y; }
      jmat[(size_t)w * wstride + p] = make_double2(xr, xi);
    }
  }
}

// V_ij(R) (slot w = nen) and W_ij(R,w): 2*2pi sum_q f_ij(q) e^{-i 2pi q.R} / V / NTQ
// + head (calc_intW.F:942-985).  grid: (nR, nen + 1), 64 threads, up to four
// (i,j) each (NWF <= 16); the phases of a block's R are formed once per q in
// LDS.  out: V [R][j][i], W [w][R][j][i].
__global__ void __launch_bounds__(64) k_w_fourier(const double2* __restrict__ func,
                                                  const double2* __restrict__ funcw,
                                                  const double* __restrict__ sq, int ntq, int nwf,
                                                  int nen, int na1, int na2, int na3,
                                                  double volume, double chead,
                                                  const double2* __restrict__ cheadw,
                                                  double2* __restrict__ vmat,
                                                  double2* __restrict__ wmat) {
  const int r = blockIdx.x, w = blockIdx.y;
  const int n1 = 2 * na1 + 1, n2 = 2 * na2 + 1;
  const int ia1 = r % n1 - na1, ia2 = (r / n1) % n2 - na2, ia3 = r / (n1 * n2) - na3;
  const double tpi = 2.0 * 3.141592653589793238462643383279502884;
  const bool isv = w == nen;
  const int np = nwf * nwf;  // <= 256: four (i, j) per thread
  // e^{-i phase(q)} of kWfQ q at a time in LDS, shared by every (i, j)
  constexpr int kWfQ = 256;
  __shared__ double2 ph[kWfQ];
  double sr[4] = {0.0, 0.0, 0.0, 0.0}, si[4] = {0.0, 0.0, 0.0, 0.0};
  for (int qb = 0; qb < ntq; qb += kWfQ) {
    const int nq = min(kWfQ, ntq - qb);
    __syncthreads();
    for (int e = threadIdx.x; e < nq; e += blockDim.x) {
      const int q = qb + e;
      const double phase = tpi * (sq[3 * q] * (double)ia1 + sq[3 * q + 1] * (double)ia2 +
                                  sq[3 * q + 2] * (double)ia3);
      double sn, cs;
      sincos(phase, &sn, &cs);
      ph[e] = make_double2(cs, sn);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int ij = threadIdx.x + 64 * u;
      if (ij >= np) break;
      const int i = ij % nwf, j = ij / nwf;
      for (int e = 0; e < nq; ++e) {
        const int q = qb + e;
        const double2 f = isv ? func[((size_t)q * nwf + j) * nwf + i]
                              : funcw[(((size_t)q * nen + w) * nwf + j) * nwf + i];
        const double c = ph[e].x, sn = ph[e].y;
        // f * exp(-i phase)
        sr[u] += f.x * c + f.y * sn;
        si[u] += f.y * c - f.x * sn;
      }
    }
  }
  const double sc = 2.0 * tpi;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int ij = threadIdx.x + 64 * u;
    if (ij >= np) break;
    const int i = ij % nwf, j = ij / nwf;
    if (isv) {
      vmat[((size_t)r * nwf + j) * nwf + i] = make_double2(
          sc * sr[u] / volume / (double)ntq + chead, sc * si[u] / volume / (double)ntq);
    } else {
      const double2 h = cheadw[w];
      wmat[(((size_t)w * (n1 * n2 * (2 * na3 + 1)) + r) * nwf + j) * nwf + i] =
          make_double2(sc * sr[u] / volume / (double)ntq + h.x, sc * si[u] / volume / (double)ntq + h.y);
    }
  }
}

int w3d_max_wannier() { return kMaxWf; }

void launch_w_func(const double2* rho, const long long* rho_off, const int4* qinfo, int ntq,
                   int nwf, double2* func, double2* shell, hipStream_t s) {
  hipLaunchKernelGGL(k_w_func, dim3(ntq), dim3(256), 0, s, rho, rho_off, qinfo, nwf, func, shell);
  RP_HIP(hipGetLastError());
}

void launch_w_contract(const float2* eps, const long long* eps_off, const double2* rho,
                       const long long* rho_off, const int4* qinfo, const int* minv,
                       const double2* pf1, const double2* pf2, const long long* map_off,
                       const double2* shell, const int* goff, const int* gmem, int ngroups,
                       int nwf, int nen, int max_nge, double2* xw, double2* funcw,
                       hipStream_t s) {
  RP_CHECK(nwf >= 1 && nwf <= kMaxWf, "W contraction supports 1..16 Wannier functions");
  RP_CHECK(max_nge >= 1, "W contraction: empty eps sphere");
  if (ngroups <= 0) return;
  hipLaunchKernelGGL(k_w_xvec, dim3(ngroups), dim3(256), 0, s, rho, rho_off, qinfo, minv, pf1,
                     pf2, map_off, goff, gmem, nwf, w_contract_vectors(nwf), xw);
  switch (nwf) {
#define RP_WC(NV)                                                                             \
  case NV:                                                                                    \
    hipLaunchKernelGGL((k_w_contract<NV>), dim3(nen, ngroups), dim3(512), 0, s, eps, eps_off, \
                       rho, rho_off, qinfo, minv, pf1, pf2, map_off, shell, goff, gmem, xw,   \
                       nen, funcw);                                                           \
    break;
    RP_WC(1) RP_WC(2) RP_WC(3) RP_WC(4) RP_WC(5) RP_WC(6) RP_WC(7) RP_WC(8)
    RP_WC(9) RP_WC(10) RP_WC(11) RP_WC(12) RP_WC(13) RP_WC(14) RP_WC(15) RP_WC(16)
#undef RP_WC
  }
  RP_HIP(hipGetLastError());
}

void launch_j_contract(const float2* eps, const long long* eps_off, const double2* rho,
                       const long long* rho_off, const int4* qinfo, const int* minv,
                       const double2* pf1, const double2* pf2, const long long* map_off,
                       double2* xg, const double2* shell, int ntq, int npair, int nen,
                       double2* funcw, hipStream_t s) {
  const int groups = (npair + 31) / 32;
  const int g32 = 32 * groups;
  hipLaunchKernelGGL(k_j_xvec, dim3(ntq), dim3(256), 0, s, rho, rho_off, qinfo, minv, pf1, pf2,
                     map_off, npair, g32, xg);
  hipLaunchKernelGGL(k_j_contract, dim3(nen, ntq, groups), dim3(512), 0, s, eps, eps_off, rho,
                     rho_off, qinfo, minv, pf1, pf2, map_off, (const double2*)xg, g32, shell,
                     npair, nen, funcw);
  RP_HIP(hipGetLastError());
}

size_t j_contract_scratch(size_t sum_nge, int npair) { return sum_nge * 32 * ((npair + 31) / 32); }

void launch_j_func(const double2* rho, const long long* rho_off, const int4* qinfo, int ntq,
                   int npair, double2* func, double2* shell, hipStream_t s) {
  hipLaunchKernelGGL(k_j_func, dim3(ntq, npair), dim3(256), 0, s, rho, rho_off, qinfo, npair, func,
                     shell);
  RP_HIP(hipGetLastError());
}

void launch_j_sum(const double2* func, const double2* funcw, int ntq, int nwf, int nen, int r0,
                  double volume, double chead, const double2* cheadw, double2* xmat,
                  double2* jmat, size_t wstride, hipStream_t s) {
  hipLaunchKernelGGL(k_j_sum, dim3(nen + 1), dim3(64), 0, s, func, funcw, ntq, nwf, nen, r0, volume,
                     chead, cheadw, xmat, jmat, wstride);
  RP_HIP(hipGetLastError());
}

void launch_w_fourier(const double2* func, const double2* funcw, const double* sq, int ntq, int nwf,
                      int nen, const int na[3], double volume, double chead, const double2* cheadw,
                      double2* vmat, double2* wmat, hipStream_t s) {
  const int nr = (2 * na[0] + 1) * (2 * na[1] + 1) * (2 * na[2] + 1);
  hipLaunchKernelGGL(k_w_fourier, dim3(nr, nen + 1), dim3(64), 0, s, func, funcw, sq, ntq, nwf,
                     nen, na[0], na[1], na[2], volume, chead, cheadw, vmat, wmat);
  RP_HIP(hipGetLastError());
}

}  // namespace rp
