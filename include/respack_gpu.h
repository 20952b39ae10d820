/*
 * respack_gpu.h -- C ABI of librespack_gpu.so, the MI355X (gfx950) hot path of
 * RESPACK's cRPA polarisation (src/chiqw) and Wannier-basis interaction
 * integrals (src/calc_int).
 *
 * Conventions
 *   - Every entry point returns 0 on success and a non-zero status otherwise;
 *     rp_last_error() then returns the message (thread-local).  There is no
 *     silent fallback: a missing device or an unsupported size is an error.
 *   - Arrays are plain host pointers.  "complex" arrays are interleaved
 *     (re, im) doubles -- the bytes of Fortran complex(8) / C double _Complex.
 *     Index order is given per argument; "Fortran order (i,j,k)" means i fastest.
 *   - G vectors and k/q points use the reference's conventions: integer
 *     triplets g(3), fractional coordinates in units of the reciprocal vectors.
 *   - The library owns all device memory behind the handle; no host pointer is
 *     kept after a call returns.  One handle = one HIP device + one stream;
 *     calls on one handle must not overlap.
 */
#ifndef RESPACK_GPU_H
#define RESPACK_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rp_ctx* rp_handle;

/* Precision of the chi contraction (the rest of the path is FP64). */
enum {
  RP_PREC_FP64 = 0,   /* FP64 MFMA, FP64 accumulation                            */
  RP_PREC_FP32 = 1,   /* FP32 MFMA, FP32 accumulation per K-chunk, FP64 merge    */
  RP_PREC_BF16X3 = 2  /* split-BF16 3-term products on MFMA, FP32 accumulation    */
};

/* Lifecycle.  Replaces the reference's MPI_INIT/OpenMP setup
 * (chiqw.F:8-12); device = HIP ordinal. */
int rp_init(int32_t device, rp_handle* out);
int rp_finalize(rp_handle h);
const char* rp_last_error(void);

/* ------------------------------------------------------------------------
 * Setup helpers (host code; restate the reference routines cited).
 * ------------------------------------------------------------------------ */
/* est_nwx2.f90:1-13 */
int32_t rp_algn235(int32_t n);
/* gen_grid.f90:1-39; em = nt complex */
int rp_gen_grid(int32_t nt, double emax, double* em);
/* chiqw_sub.F:773-812; b = b1,b2,b3 (9 doubles); kg = 3*ntg ints; *ng = count
 * (kg written only while count <= ntg) */
int rp_make_kg0(int32_t ntg, const double* b, double gcut, const double* q,
                int32_t* kg, int32_t* ng);
/* m_tetrahedron_20170325.F:505-584; idx = 4*6*n1*n2*n3 zero-based grid
 * indices (ix + n1*(iy + n2*iz)) */
int rp_ttrhdrn_mkidx(int32_t n1, int32_t n2, int32_t n3, const double* b, int32_t* idx);

/* ------------------------------------------------------------------------
 * chiqw drop-in (replaces PROGRAM DIELECTRIC's hot path, chiqw.F:1250-1555
 * and the q = 0 branch chiqw.F:1557-2085).
 * ------------------------------------------------------------------------ */
typedef struct {
  int32_t n_calc_band;          /* N_CALC_BAND (0 -> NTB), m_rdinput.f90:54     */
  int32_t num_freq_grid;        /* Num_freq_grid (70)                           */
  int32_t flg_crpa;             /* 0 fRPA, 1 band-cRPA                          */
  int32_t precision;            /* RP_PREC_* for the contraction                */
  double ecut_for_eps;          /* Ry; 0 -> Ecut_for_psi/10 (chiqw.F:350-352)   */
  double shift_ef;              /* eV (chiqw.F:67)                              */
  double max_excitation_energy; /* eV (200)                                     */
  double green_func_delt;       /* eV (0.1)                                     */
  double ttrhdrn_dmna;          /* used in Hartree, as the reference does       */
  double ttrhdrn_dmnr;
} rp_chiqw_params;

/* Opens a chiqw system from the contents of dir-wfn (chiqw.F:88-329):
 *   lattice   9 doubles: a1, a2, a3 (bohr)
 *   rg        9*nsym ints, Fortran order rg(i,j,iop); pg 3*nsym ints
 *   ski       3*nk_irr doubles; ngi nk_irr ints
 *   kgi       3*ntg*nk_irr ints, Fortran order KGI(d,ig,ik)
 *   eig       ntb*nk_irr doubles (Ha), Fortran order E_EIGI(ib,ik)
 *   wfn       complex, Fortran order C(ig,ib,ik), ntg*ntb*nk_irr
 *   ef        Fermi energy (Ha) from dat.bandcalc
 * Unfolds the k mesh (symmetry + time reversal), classifies the bands,
 * builds the frequency grid and uploads the band grids u_nk(r) to HBM. */
int rp_chiqw_open(rp_handle h, const double* lattice, int32_t nsym, int32_t nnp,
                  const int32_t* rg, const int32_t* pg, int32_t nk_irr,
                  const double* ski, const int32_t* ngi, int32_t ntg,
                  const int32_t* kgi, int32_t ntb, const double* eig,
                  const double* wfn, double ef, const rp_chiqw_params* prm);

/* Band-cRPA projector (chiqw.F:870-934): per full-mesh k (NTK of them, in the
 * order rp_chiqw_info reports) the window offset ns[k] and width nb[k]; umat
 * complex, Fortran order UNT(jb, jw, k) with jb < mb (the max of nb). */
int rp_chiqw_set_crpa(rp_handle h, int32_t ntk, const int32_t* ns, const int32_t* nb,
                      int32_t mb, int32_t nwf, const double* umat);

/* Sizes of the opened system.  iout (16 ints): NTK, nkb1, nkb2, nkb3, nwx,
 * nwy, nwz, nen, N_occ, N_vir, N_docc, NTB, Nk_irr, ntg, 0, 0.
 * dout (8 doubles): Ecut_for_psi (Ry), Ecut_for_eps (Ry), volume, emax (Ha),
 * E_F (Ha, shifted), delt (Ha), 0, 0.  em: nen complex (may be NULL). */
int rp_chiqw_info(rp_handle h, int32_t* iout, double* dout, double* em);
/* Full-mesh k points (3*NTK doubles, SK0 order) and their irreducible index
 * (0-based) and time-reversal flag (+1/-1). */
int rp_chiqw_kmesh(rp_handle h, double* sk0, int32_t* numirr, int32_t* trs);

/* Per-q stages.  rp_chiqw_prepare_q sets q (fractional, folded as the
 * reference's search_list) and returns NG (size of the chi G sphere), and the
 * sphere itself in lg (3*NG ints, may be NULL; size it with a first call). */
int rp_chiqw_prepare_q(rp_handle h, const double* q, int32_t* ng, int32_t* lg, int32_t lg_cap);
/* Pair densities M_(pair,k)(G) of calc_InterStateMatrix (chiqw_sub.F:97-205)
 * for every pair and k into HBM.  m_out (optional, may be NULL): complex64
 * (interleaved floats) [K][NG] with K = pair*NTK + k; at q = 0, [3][K][NG]
 * (the G = 0 entry replaced per direction as chiqw.F:1787-1805). */
int rp_pairdens_batch(rp_handle h, float* m_out);
/* Tetrahedron weights (ttrhdrn_chi0 for dsgn = +1 and -1, chiqw.F:1350-1395)
 * times (1 - P_occ P_vir)/NTK into HBM.  d_out (optional): complex [nen][K]. */
int rp_tetra_weights_batch(rp_handle h, double* d_out);
/* chi(G,G',w) = sum_K conj(M_K(G)) d_K(w) M_K(G') / volume (chiqw.F:1397-1476)
 * in the given RP_PREC_*.  chi_out (optional): complex, Fortran order
 * chiqw(ig,jg,ie) (times 3 directions at q = 0). */
int rp_chi_accumulate(rp_handle h, int32_t precision, double* chi_out);
/* Replaces the chi of the prepared q (the layout rp_chi_accumulate writes:
 * complex (ig, jg, ie[, dir])) before rp_dyson, e.g. with the band-tail
 * closure applied.  No counterpart in the reference. */
int rp_chi_set(rp_handle h, const double* chi_in);
/* Dyson step calc_eps_rpa / calc_eps_rpa_q_0 (chiqw_sub.F:386-568) on the GPU
 * in FP64: eps^-1, Fortran order (ig,jg,ie) (times 3 at q = 0). */
int rp_dyson(rp_handle h, double* epsinv_out);
/* The Dyson step of rp_dyson for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  chi_dev: device memory
 * holding chi of the prepared q in the layout rp_chi_accumulate writes
 * (complex (ig, jg, ie[, dir])); epsinv_dev: device memory of the same size,
 * receives eps^-1 as rp_dyson's output.  The step's scratch is a fixed 4 MiB
 * device array per GPU (status, Drude terms, pivots, Gauss-Jordan move lists):
 * sizes that need more (about NG x 3 nen above 10^6) are rejected.  The call
 * makes no allocation and synchronises the handle's stream to check the
 * status.  Calls that overlap on one GPU share that scratch and race. */
int rp_dyson_noalloc(rp_handle h, const double* chi_dev, double* epsinv_dev);
/* All four stages for one q.  timings (optional, 8 doubles): seconds of pair
 * densities, weights, contraction, Dyson, total device time, host prepare,
 * the main contraction kernel alone (summed over its launches), and the
 * number of those launches (1, or 3 at q = 0). */
int rp_chiqw_q(rp_handle h, const double* q, int32_t precision, double* epsinv_out,
               double* chi_out, double* timings);
/* Device-resident number of pairs (after the Max_excitation_energy filter). */
int rp_chiqw_npair(rp_handle h, int32_t* npair);
/* Per-q window for the q points prepared after this call (the batched
 * configuration's uneven G counts and band windows): ecut_for_eps (Ry) replaces
 * the chi cutoff of rp_chiqw_open (chiqw.F:786-799 with a per-q Ecut_for_eps)
 * and n_vir keeps only the lowest n_vir empty bands in the pair list (a per-q
 * N_CALC_BAND, chiqw.F:1267-1295); 0 restores the value of rp_chiqw_open. */
int rp_chiqw_set_window(rp_handle h, double ecut_for_eps, int32_t n_vir);

/* FP32 accumulation span (in K, a multiple of 32) of the split-BF16 contraction
 * for the q points run after this call: the FP32 accumulators are added into
 * FP64 every span K.  Shorter spans lower the accumulation error at the cost
 * of more read-modify-writes of the FP64 partial tiles.  0 (the default):
 * 8192, shortened automatically for a q whose eps is ill-conditioned.  No
 * counterpart in the reference (its accumulation is FP64). */
int rp_chiqw_set_span(rp_handle h, int32_t span);

/* Batched configuration, captured (no counterpart in the reference, which runs
 * one q per iteration of its q loop, chiqw.F:662-687; for callers that run
 * many q of one system).  rp_chiqw_batch_plan prepares all nq q points
 * (qs[3*nq], fractional) with per-q ecut_for_eps[nq] and n_vir[nq] windows as
 * rp_chiqw_set_window (either may be NULL: the values of rp_chiqw_open), runs
 * every stage of every q once with its status checked (tuning the contraction
 * per shape when rp_chiqw_autotune is on), and captures the launch sequence of
 * all q and all frequencies into one HIP graph.  The batch owns its buffers;
 * the handle's own q state and buffers are left as they were.  A new plan
 * replaces the old one. */
int rp_chiqw_batch_plan(rp_handle h, int32_t nq, const double* qs, const double* ecut_for_eps,
                        const int32_t* n_vir, int32_t precision);
/* Replays the graph (all q recomputed from the resident wavefunctions);
 * seconds: the device time of the replay. */
int rp_chiqw_batch_run(rp_handle h, double* seconds);
/* Entry iq of the last replay: NG, number of directions (3 at q = 0, else 1),
 * the chi G list (lg[3*NG], capacity lg_cap triplets; may be NULL) and
 * eps^-1 as rp_dyson's output (may be NULL). */
int rp_chiqw_batch_result(rp_handle h, int32_t iq, int32_t* ng, int32_t* ndir, int32_t* lg,
                          int32_t lg_cap, double* epsinv_out);
int rp_chiqw_batch_free(rp_handle h);
/* Autotuner of the contraction (no counterpart in the reference): when on,
 * the first contraction of each shape (precision, NG, nen, padded K,
 * directions) times the launch candidates around the built-in heuristic on
 * that q's data and keeps the fastest for every later q of the shape. */
int rp_chiqw_autotune(rp_handle h, int32_t enable);
int rp_chiqw_tune_count(rp_handle h, int32_t* n);
/* Entry idx: iout[10] = precision, NG, nen, directions, padded K, candidates,
 * chosen K split, chosen workgroups, heuristic K split, heuristic workgroups;
 * dout[2] = chosen and heuristic time (ms, one contraction). */
int rp_chiqw_tune_info(rp_handle h, int32_t idx, int32_t* iout, double* dout);

/* ------------------------------------------------------------------------
 * calc_int drop-in, calc_w3d (replaces calc_intW.F:8-985): Wannier-basis bare
 * and screened Coulomb integrals V_ij(R) and W_ij(R, w).
 * ------------------------------------------------------------------------ */
/* Opens the W3D system from dir-wfn (calc_intW.F:8-330): the same arrays as
 * rp_chiqw_open (no eigenvalues or wavefunctions), the chi cutoff of
 * dir-eps/dat.chi_cutoff (Ry) and the frequencies of dir-eps/dat.wgrid
 * (nen complex, Ha).  Unfolds the k mesh and builds the G lists of every
 * full-mesh k. */
int rp_w3d_open(rp_handle h, const double* lattice, int32_t nsym, int32_t nnp,
                const int32_t* rg, const int32_t* pg, int32_t nk_irr, const double* ski,
                const int32_t* ngi, int32_t ntg, const int32_t* kgi, double ecut_for_eps,
                int32_t nen, const double* em);
/* Sizes.  iout (16 ints): NTK, nkb1, nkb2, nkb3, Na1, Na2, Na3, nwx, nwy, nwz,
 * nen, NTG, NWF, Nq_irr, NTQ, NoG0 (0-based, after rp_w3d_set_q).  dout (8):
 * Ecut_for_psi, Ecut_for_eps, volume, chead (after rp_w3d_vw), 0... */
int rp_w3d_info(rp_handle h, int32_t* iout, double* dout);
/* Full-mesh k points (3*NTK, SK0 order) and NG0(k): the record layout of
 * dir-wan/dat.wan. */
int rp_w3d_kmesh(rp_handle h, double* sk0, int32_t* ng0);
/* Wannier functions of dir-wan/dat.wan (calc_intW.F:332-341): complex
 * C0(NTG, NWF, NTK), Fortran order, G in make_KG0 order of SK0(k). */
int rp_w3d_set_wannier(rp_handle h, int32_t nwf, const double* wan);
/* The irreducible q list (dir-eps/qNNN/dat.sq, calc_intW.F:411-433; Nq_irr =
 * Nk_irr): unfolds the q mesh (calc_intW.F:420-487) and builds the eps/psi G
 * lists (make_LG0).  ng_eps (optional): the eps sphere size per irreducible q,
 * i.e. the leading dimensions of its dat.epsqw.600. */
int rp_w3d_set_q(rp_handle h, int32_t nq_irr, const double* sqi, int32_t* ng_eps);
/* eps^-1 of irreducible q iq (0-based), complex (ig,jg,ie) as read from
 * dat.epsqw.600 (at Gamma: the average of 600/601/602, calc_intW.F:595-612).
 * Stored in single precision, as the reference's epsirr. */
int rp_w3d_set_epsinv(rp_handle h, int32_t iq, const double* epsinv);
/* rho, func and funcw for every full-mesh q (calc_intW.F:647-912) on the GPU.
 * func_out (optional): complex func(i,j,q), Fortran order (NWF,NWF,NTQ);
 * funcw_out (optional): complex (NWF,NWF,nen,NTQ). */
int rp_w3d_funcs(rp_handle h, double* func_out, double* funcw_out);
/* V and W (calc_intW.F:782-795, 942-985).  vmat: complex V_MAT(NWF,NWF,
 * -Na1:Na1,-Na2:Na2,-Na3:Na3); wmat: complex W_MAT(...,nen); heads (optional,
 * 2+2*nen doubles): chead, 0, cheadw(nen) complex.  Runs rp_w3d_funcs first if
 * needed. */
int rp_w3d_vw(rp_handle h, double* vmat, double* wmat, double* heads);

/* calc_j3d (replaces calc_intJ.F:606-904) on the same opened system: the bare
 * and screened exchange X_ij(R), J_ij(R, w) for the nr lattice vectors rvec
 * (3*nr ints; the reference's IX/IY/IZ_INTJ ranges, ia1 outer, ia3 inner).
 * xmat: complex (NWF,NWF,nr); jmat: complex (NWF,NWF,nr,nen); heads as
 * rp_w3d_vw (added on i = j at R = 0 only). */
int rp_j3d_run(rp_handle h, int32_t nr, const int32_t* rvec, double* xmat, double* jmat,
               double* heads);
/* Device seconds (events on the handle's stream) of the last rp_w3d_funcs:
 * t[0] rho, t[1] func, t[2] the W contraction, t[3] total; and of the last
 * rp_j3d_run: t[4] rho of every R, t[5] func and sums, t[6] the J
 * contraction (all R), t[7] total (8 doubles).  No counterpart in the
 * reference (instrumentation). */
int rp_w3d_stage_times(rp_handle h, double* t);

/* ------------------------------------------------------------------------
 * Band-tail closure network.  No counterpart in the reference: a small MLP
 * (GELU hidden layers, linear output, BF16 weights with FP32 accumulation)
 * that predicts chi_full - chi_truncated for diagonal and near-diagonal
 * entries from features of the truncated chi; the full-band method stays the
 * fallback and the judge.  Parameters are one FP32 vector: per layer l the
 * weights W_l[out][in] then the biases b_l[out].
 * ------------------------------------------------------------------------ */
/* depth hidden layers of the given width (all layers at most 128 wide);
 * Glorot-uniform weights from seed, zero biases. */
int rp_mlp_open(rp_handle h, int32_t nin, int32_t width, int32_t depth, int32_t nout, uint64_t seed);
int rp_mlp_nparams(rp_handle h, int64_t* n);
int rp_mlp_get_params(rp_handle h, float* params);
/* Also resets the Adam state. */
int rp_mlp_set_params(rp_handle h, const float* params);
/* y[n][nout] = network(x[n][nin]) (host arrays). */
int rp_mlp_forward(rp_handle h, int64_t n, const float* x, float* y);
/* Gradient of the mean-squared error over the n samples with targets
 * t[n][nout] with respect to every parameter (same layout), and the loss. */
int rp_mlp_gradient(rp_handle h, int64_t n, const float* x, const float* t, float* grad, double* loss);
/* steps Adam steps (beta 0.9 / 0.999) on minibatches of `batch` samples drawn
 * with replacement from the n samples (uploaded once); losses[steps] optional. */
int rp_mlp_train(rp_handle h, int64_t n, const float* x, const float* t, int32_t steps, int32_t batch,
                 double lr, uint64_t seed, double* losses);

#ifdef __cplusplus
}
#endif
#endif
