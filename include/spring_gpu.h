/*
 * spring_gpu.h -- C ABI of libspring_gpu.so, the MI355X (gfx950) scattering-rate path.
 *
 * This library replaces the hot routines of SPRING's kappa.x / LineWidth.x:
 *   ScatterProbAllq1 + ScatterMatEl + ScatterProbability   src/Kappa/Src/ThirdOrder_m.f90:41-166, 278-439
 *   CalculateLinewidth (Gaussian delta)                     src/Kappa/Src/ThirdOrder_m.f90:442-543
 *   CalculateLwTetrahedron (tetrahedron delta)              src/Kappa/Src/ThirdOrder_m.f90:169-275
 *   LineWidth.x's off-mesh q0 path (spr_lw3_offmesh)        src/LineWidth/Src/Linewidth_m.f90:48-417
 *   ScatterProbq1q2_per + ScatterMatEl4 + ScatterProbability4  src/Kappa/Src/FourthOrder_m.f90:42-473
 *   TetrahedronIntegration4th (processes, Gaussian form)   src/Kappa/Src/Tetrahedron4th_sep.f90:37-425
 *   kappa_noIter / kappa_noIter_symm                        src/Kappa/Src/ThermalConductivity.f90:36-293
 *   the q0 loop that drives them                            src/Kappa/Src/main.f90:552-715
 *
 * Conventions.
 *   - Arrays are passed as the reference's column-major Fortran arrays, byte for byte
 *     (no host transposes).  complex(dp) arrays are interleaved (re, im) doubles.
 *   - Every index value stored in an array (basis atom, grid point) is 0-based; the
 *     Fortran interface module (fortran/spring_gpu_m.f90) subtracts 1 where the
 *     reference stores 1-based values.
 *   - Frequencies are angular, rad/ps; sigma is given in the same unit.
 *   - Every entry point returns 0 on success or a negative SPR_E_* code; the message is
 *     in spr_last_error() (thread-local).  Calls are synchronous: outputs are on the host
 *     when they return.  One context is used by one host thread at a time.
 *   - One GPU: the context binds to `device` (HIP ordinal) at creation.
 */
#ifndef SPRING_GPU_H
#define SPRING_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SPR_OK 0
#define SPR_E_ARG (-1)       /* bad argument / shape */
#define SPR_E_HIP (-2)       /* HIP runtime error */
#define SPR_E_OOM (-3)       /* device allocation failed */
#define SPR_E_STATE (-4)     /* invariant violation (e.g. no FC4 loaded) */
#define SPR_E_UNSUPPORTED (-5)

/* precision of the FT + eigenvector contraction (occupations, delta arguments and the
 * final reductions stay FP64 in every mode) */
#define SPR_PREC_FP64 0
#define SPR_PREC_FP32 1      /* FP32 phase sums + contraction, compensated FP32 delta sum */
#define SPR_PREC_BF16 2      /* BF16-operand MFMA contraction, FP32 accumulate */

/* delta function of the energy conservation */
#define SPR_DELTA_GAUSS 0
#define SPR_DELTA_TETRA 1    /* linear tetrahedron method (the reference's default) */

/* crystal: the `cell` fields the path uses (src/Kappa/Src/unit_mod.f90:9-38) */
typedef struct {
    int nb;                     /* basis atoms; Ndof = 3 nb */
    const double *mass;         /* (nb) amu */
    double vol;                 /* unit-cell volume, A^3 */
} spr_crystal;

/* anharmonic IFC, FC3type / FC4type (src/Kappa/Src/FC3_mod.f90:30-99, FC4_mod.f90:30-97) */
typedef struct {
    int order;                  /* 3 or 4 */
    int natm;                   /* leading neighbour dimension (Natm max) */
    const int *atm_num;         /* atmNum(nb) */
    const int *atm_indx;        /* atmIndx(4, natm, nb): cell (3 ints) + basis atom (0-based) */
    const double *fc;           /* fC(3,3,3,natm,natm,nb) or fC(3,3,3,3,natm,natm,natm,nb), eV/A^n */
} spr_ifc;

/* Gamma-centred mesh and irreducible wedge (q_tetrahedron, TetrahedronQ_m.f90:35-331) */
typedef struct {
    int mesh[3];
    int nq;                     /* mesh(1)*mesh(2)*mesh(3) */
    const int *q_int;           /* q_pnt_int(3, nq), folded addresses; grid index order
                                   i = sum_k mod(q_k, N_k) prod_{j<k} N_j */
    int nq0;
    const int *irr_grid;        /* (nq0) grid index of each irreducible q0 */
    const int *pos;             /* pos(nq): irreducible position of every grid point */
    const int *multiplicity;    /* multiplicity(nq0) */
    int nsym;                   /* point operations used by kappa symm (incl. time reversal) */
    const double *rot;          /* Rxyz(3,3,nsym) Cartesian rotations; may be NULL if nsym==0 */
    const int *tetra;           /* tetrahedrons(4, 6 nq) grid indices (TetrahedronQ_m.f90:172-188);
                                   may be NULL (then spr_lw3_tetra* fail) */
} spr_mesh;

/* harmonic phonons on the full mesh (Phon, phonon_mod.f90:33-46) */
typedef struct {
    int nq, ndof;
    const double *evec;         /* Evec(ndof, ndof, nq) complex(dp) */
    const double *omega;        /* omega(ndof, nq), rad/ps */
    const double *inv_omega;    /* OnebyOmega(ndof, nq) */
    const double *nbe;          /* nBE(ndof, nq) */
    const double *grp_vel;      /* grp_vel(3, ndof, nq), m/s (for kappa) */
} spr_phonons;

typedef struct spr_ctx spr_ctx;

/* Upload FC3 (and FC4 if not NULL), mesh tables and phonons once; replaces the data set-up
 * of main.f90:192-196, 279-300 and set_Phon_mp (phonon_mod.f90:156-315). */
int spr_ctx_create(const spr_crystal *cr, const spr_ifc *fc3, const spr_ifc *fc4,
                   const spr_mesh *mesh, const spr_phonons *ph, int device, spr_ctx **out);
int spr_ctx_destroy(spr_ctx *ctx);
const char *spr_last_error(void);

/* Parity path of ScatterProbAllq1 (ThirdOrder_m.f90:41-166): W(s2,s1,s0,q1) =
 * Sctrq1(Ndof,Ndof,Ndof,Nq) for irreducible q0 `i0`, every q1 (q2 = -(q0+q1) mod G),
 * written to W_out (Ndof^3 * Nq doubles). FP64. */
int spr_v3_w(spr_ctx *ctx, int i0, double *W_out);

/* The same W for a list of (irreducible q0 index, q1 grid index) pairs: W_out(Ndof^3, n),
 * W_out(s2,s1,s0,k) for pair k (q2 = -(q0+q1) mod G).  FP64.  Source of the training samples
 * of the learned proposal (spr_mlp_*). */
int spr_v3_w_pairs(spr_ctx *ctx, int n, const int *q0_index, const int *q1_grid, double *W_out);

/* Branch selection rules of the context (SelectionRule3 / SelectionRule4 of &AnhInfo, mapped by
 * PolarizationStrtEnd, src/Kappa/Src/Helper_m.f90:310-353, into Polarization3rd(2, 3) /
 * Polarization4th(2, 4)): pol[2 k], pol[2 k + 1] = first and last branch (1-based, inclusive) of
 * slot k (s0, s1, s2[, s3]).  Every later three-phonon (order 3) or four-phonon (order 4)
 * linewidth of the context sums over these ranges as CalculateLwTetrahedron
 * (ThirdOrder_m.f90:169-171, 207-214) and SumOverPolarization (Tetrahedron4th_sep.f90:213-224)
 * loop over them (s0 outside its range gives 0), and spr_kappa_rta / spr_lw3_kappa_all take their
 * mode range from slot 1 (default_strt..default_end of kappa_noIter, ThermalConductivity.f90:
 * 50-51, 106-111; the wider of the two when both orders are set).  The Gaussian sums apply the
 * same ranges.  Default: every branch.  The iterative BTE ignores them, as kappa_iter.x does. */
int spr_set_branch_rules(spr_ctx *ctx, int order, const int *pol);

/* Gaussian-delta three-phonon linewidth of irreducible q0 `i0` (ScatterProbAllq1 fused with
 * CalculateLinewidth, ThirdOrder_m.f90:442-543): lw(Ndof), rad/ps. */
int spr_lw3(spr_ctx *ctx, int i0, double sigma, int prec, double *lw);

/* All irreducible q0 in one persistent launch (the q0 loop of main.f90:552-715 restricted to
 * three-phonon scattering): LW_Allq0(Ndof, nq0). */
int spr_lw3_all(spr_ctx *ctx, double sigma, int prec, double *lw_all);

/* The same for the slice of irreducible q0 [q0_off, q0_off + q0_cnt): lw(Ndof, q0_cnt).  The
 * (offset, count) work split of DistributeQpoints (src/Kappa/Src/Helper_m.f90:356-434), by which
 * each coarray image takes a contiguous block of the q0 loop (main.f90:552-715). */
int spr_lw3_slice(spr_ctx *ctx, int q0_off, int q0_cnt, double sigma, int prec, double *lw);

/* Tetrahedron-method three-phonon linewidth of irreducible q0 `i0` (CalculateLwTetrahedron,
 * ThirdOrder_m.f90:169-275, the reference's default delta): lw(Ndof) rad/ps and P3_q0(Ndof);
 * the reference's P3 of the q0 is sum(P3_q0).  W of every q1 is built on the device (the W-output
 * mode of the small-cell kernel for 1- and 2-atom cells, of the FP64 matrix-core kernel for 3 to
 * 10 atoms, the generic W path otherwise); FP64 throughout.  Needs mesh->tetra. */
int spr_lw3_tetra(spr_ctx *ctx, int i0, double *lw, double *p3_q0);
/* every irreducible q0: lw_all(Ndof, nq0), p3_all(Ndof, nq0) */
int spr_lw3_tetra_all(spr_ctx *ctx, double *lw_all, double *p3_all);

/* LineWidth.x: the three-phonon linewidth of a q0 that is not a mesh point (a point of the
 * dispersion path), ScatterProbAllq1 + CalculateLwTetrahedron (delta = SPR_DELTA_TETRA, the path
 * src/LineWidth/Src/main.f90:293-379 runs) of src/LineWidth/Src/Linewidth_m.f90:48-417, or with
 * SPR_DELTA_GAUSS the Gaussian sum of CalculateLinewidth's formula (:234-323; main.f90 keeps that
 * call switched off) over every branch, sigma in rad/ps.  q0_cart (1/A) with its phonons as set_Phon_mp gives
 * them (evec0 = Evec(Ndof, Ndof) complex, omega0, inv_omega0 = OnebyOmega); the q1 are the
 * context's mesh (zero_centered selects the reference's tetraQZeroCentered mesh and BZ test,
 * main.f90:316-330); for every q1, q2 = G - (q0 + q1) with the reference's G search
 * (:120-184) and its phonons solved on the device from fc2 (order 2, fc(3,3,natm,nb)) at
 * temperature T; a q1 whose q2 is not found adds nothing.  lw(Ndof) rad/ps; W_out (optional,
 * Sctrq1(Ndof, Ndof, Ndof, Nq)) the W rows.  FP64.  latvec as spr_get_phonon. */
int spr_lw3_offmesh(spr_ctx *ctx, const double *q0_cart, const double *evec0, const double *omega0,
                    const double *inv_omega0, const spr_crystal *cr, const double *latvec, const spr_ifc *fc2,
                    double T, int zero_centered, int delta, double sigma, double *lw, double *W_out);

/* ---- Iterative Boltzmann transport (kappa_iter.x, src/Kappa/Src_iterative) ----
 * Works on the q points irreducible under time reversal only, as the reference does
 * (make_tetrahedron, Src_iterative/Irr_q_point_m.f90:206-330): representative = the lower grid
 * index of q and -q, F(-q) = -F(q).  The first call stores W of every representative q and
 * every grid q1 in device memory (QconsrvMomentum, Quasi_momentum_consrv.f90:118-451;
 * nirr x Nq x Ndof^3 doubles; each ordering of (q1, q2) evaluated on its own, by the W modes
 * of the small-cell (1-2 atoms) and matrix-core (3-10 atoms) kernels) and needs the context's
 * tetrahedra.  F arrays are
 * F[qi][s][3] (F_vecAll(3, s, qi)), 1/tau arrays [qi][s] (Inv_tauqs(s, qi)). */

/* nirr, and (if not NULL) the grid index of each irreducible point and the number of grid
 * points it stands for (1 or 2). */
int spr_bte_irr(spr_ctx *ctx, int *nirr, int *irr_grid, int *deg);

/* 1/tau of every irreducible (q, s): TetrahedronWplus + TetrahedronWminus
 * (Src_iterative/Quasi_momentum_consrv.f90:701-1208). */
int spr_bte_tau_inv(spr_ctx *ctx, double *inv_tau);

/* dF of one iteration for the given F, before the division by 1/tau: TetrahedronIterPlus +
 * TetrahedronIterMinus (Src_iterative/TetrahedronDeltaSum.f90:22-620). */
int spr_bte_delta_f(spr_ctx *ctx, const double *F, double *dF);

/* The self-consistent loop of Src_iterative/main.f90:243-361 at temperature T (K): F0 of
 * CreateFzero, then F = F0 + dF/(1/tau) (multiplyTaudFnew) until the largest change of a kappa
 * component is below eps_k (W/mK) or max_iter iterations ran.  kappa[9] (row-major 3x3, W/mK);
 * kappa0[9] (F = F0, may be NULL), iterations, F_out[nirr][nd][3] (may be NULL). */
int spr_bte_solve(spr_ctx *ctx, double T, double eps_k, int max_iter, double *kappa, double *kappa0,
                  int *iterations, double *F_out);

/* Harmonic phonons of nq q points on `device`, replacing get_phonon / DynMatgv_SR
 * (DynaMat_mod.f90:98-190, 486-709) as set_Phon_mp calls them (phonon_mod.f90:156-315) for a
 * non-polar crystal: D(q) from FC2 (fc2->order 2, fc(3,3,natm,nb) eV/A^2), eigenvalues and
 * eigenvectors (ascending, upper triangle of D as zheevd 'U'), dD/dq for the group velocity,
 * omega, 1/omega and n(T); the three lowest modes of a q with is_gamma[q] != 0 get
 * v = n = 1/omega = 0.  latvec[9] row-major (R = latvec . cell, Angstrom); q_cart[nq][3] in
 * 1/Angstrom.  Outputs in the layout of spr_phonons: evec[nq][ndof][ndof] complex (re, im),
 * evec[q][s][row]; omega, inv_omega, nbe [nq][ndof]; grp_vel[nq][ndof][3].  Ndof <= 32.
 * Eigenvectors of degenerate modes are any orthonormal basis of their subspace (as with
 * LAPACK), so per-mode quantities of degenerate modes are basis-dependent. */
int spr_get_phonon(const spr_crystal *cr, const double *latvec, const spr_ifc *fc2, int nq,
                   const double *q_cart, const int *is_gamma, double T, int device, double *evec,
                   double *omega, double *inv_omega, double *nbe, double *grp_vel);

/* The q mesh of kappa.x from the crystal alone (host code, no device), replacing make_tetrahedron
 * + FindIrrQ (src/Kappa/Src/TetrahedronQ_m.f90:56-331, zero-centred, shift 0) and the rotations
 * set_SpaceGrSPGlib takes from spglib: latvec[9] row-major as spr_get_phonon; basis_frac[nb][3]
 * (atm_pos(3, Nbasis) of &AtmInfo); typ[nb] species (typ_indx); mesh[3] (qmesh); time_rev (0/1)
 * adds -R to the reciprocal action, as spglib's is_time_reversal; symmetry = 0 uses the identity
 * only; symprec: distance tolerance (Angstrom) for mapping atoms.  Always sets nq0_out and
 * nsym_out; when every output pointer is NULL nothing else is written (size query).  Outputs
 * (any may be NULL): q_int[nq][3] folded into (-N/2, N/2] (q_pnt_int), q_cart[nq][3] 1/A
 * (q_pnt), irr_grid[nq0] (0-based grid index of each irreducible q, ascending), pos[nq]
 * (0-based), multiplicity[nq0], rot = Rxyz(3, 3, nsym) Cartesian rotations (column-major; with
 * time_rev the point group followed by its negatives), tetra[6 nq][4] (tetrahedrons(4, 6 Nq),
 * 0-based).  The point group is every integer matrix with entries in {-1, 0, 1} that keeps the
 * metric and maps the basis onto itself species by species up to a common translation; the
 * irreducible representative is the smallest grid index of each orbit (spglib's choice). */
int spr_mesh_build(const double *latvec, int nb, const double *basis_frac, const int *typ, const int *mesh,
                   int time_rev, int symmetry, double symprec, int *nq0_out, int *nsym_out, int *q_int,
                   double *q_cart, int *irr_grid, int *pos, int *multiplicity, double *rot, int *tetra);

/* Four-phonon Gaussian linewidth sum of irreducible q0 `i0` over all (q1,q2), the processes
 * and weights of TetrahedronIntegration4th (Tetrahedron4th_sep.f90:242-343) with a Gaussian
 * delta; main.f90:660 adds 0.5*lw to the linewidth. */
int spr_lw4(spr_ctx *ctx, int i0, double sigma, int prec, double *lw);
int spr_lw4_all(spr_ctx *ctx, double sigma, int prec, double *lw_all);

/* Four-phonon linewidth sum of irreducible q0 `i0` with the linear tetrahedron method, the
 * reference's delta for four-phonon scattering (TetrahedronIntegration4th,
 * Tetrahedron4th_sep.f90:37-425): for every q1 and every tetrahedron of the q2 mesh, the three
 * processes of spr_lw4 with g and vertex weights from FindgAndIver at Mq = omega(s0, q0), NaN
 * weights of degenerate vertices set to zero.  W4 of every ordered (q1, q2) from the FP64 W4
 * path, held one chunk of q1 rows at a time (at most 2 GiB); FP64 throughout.  Needs FC4 and
 * mesh->tetra; Ndof 3 or 6.  lw(Ndof) raw sum; main.f90:660 adds 0.5*lw. */
int spr_lw4_tetra(spr_ctx *ctx, int i0, double *lw);
/* every irreducible q0: lw_all(Ndof, nq0) */
int spr_lw4_tetra_all(spr_ctx *ctx, double *lw_all);

/* RTA thermal conductivity (kappa_noIter, ThermalConductivity.f90:36-164, symm = 0; or
 * kappa_noIter_symm + GrVelKronProdSymm, :167-293 and VelocityKronProd_symm.f90:34-93,
 * symm = 1) from LW_Allq0(Ndof, nq0), FP64: kappa(9) W/mK, mode_k(9, Ndof) optional. */
int spr_kappa_rta(spr_ctx *ctx, const double *lw_all, double T, int symm,
                  double *kappa, double *mode_k);

/* spr_kappa_rta plus the spectrally resolved kappa of -Spectral_k (SpectralResolve,
 * ThermalConductivity.f90:76-90, 147-160, completed and written as "SpectralResolvedKappa" at
 * main.f90:780-794): spectral_k(11, Ndof * nq0), column qi0 + s nq0 (0-based) = (s + 1, omega(s)
 * of irreducible q0 qi0, its nine kappa_ab contributions times kW_m_k / (T^2 V)); summed over
 * columns and divided by Nq they give kappa. */
int spr_kappa_rta_spectral(spr_ctx *ctx, const double *lw_all, double T, int symm, double *kappa,
                           double *mode_k, double *spectral_k);

/* spr_lw3_all followed by spr_kappa_rta with LW_Allq0 kept on the device in between (one
 * synchronisation instead of a download and an upload): lw_all(Ndof, nq0), kappa(9) W/mK, mode_k
 * (9, Ndof) optional.  The two calls of main.f90:552-715 and :750-760 (ThermalConductivity.f90:
 * 167-293) in one; same results. */
int spr_lw3_kappa_all(spr_ctx *ctx, double sigma, int prec, double T, int symm, double *lw_all,
                      double *kappa, double *mode_k);

/* Per-pair terms of the three-phonon Gaussian sum for caller-chosen (q0, q1) pairs:
 * out(Ndof, n): out(s0, k) = sum_{s1,s2} W(s2,s1,s0; q0_index(k), q1_grid(k)) [delta terms], the q1
 * term of CalculateLinewidth's sum (ThirdOrder_m.f90:490-541).  Used by the sampled (importance)
 * estimate of the linewidth.  FP64 needs Nb <= 2; FP32/BF16 any supported Nb. */
int spr_lw3_pairs(spr_ctx *ctx, int n, const int *q0_index, const int *q1_grid, double sigma,
                  int prec, double *out);

/* Weighted sums of those terms per irreducible q0, the importance-sampled estimate of the
 * linewidth: est(s0, qi) = sum over the n samples k with q0_index(k) == qi of
 * weight(k) * out(s0, k) of spr_lw3_pairs (est(Ndof, nq0), zero for q0 with no sample).  The
 * reference evaluates only the full sum over q1 (CalculateLinewidth, ThirdOrder_m.f90:442-543,
 * the q1 loop of ScatterProbAllq1 :98-164); this entry replaces that loop by the sampled sum the
 * learned proposal drives.  The samples are grouped by q0 into 64-pair work items that run on
 * the same kernels as the full sum, each pair weighted inside the kernel.  Weights must be finite. */
int spr_lw3_sampled(spr_ctx *ctx, int n, const int *q0_index, const int *q1_grid, const double *weight,
                    double sigma, int prec, double *est);

/* spr_lw3_pairs with no device allocation: the same terms for the same arguments (same kernels,
 * same summation order), with every device buffer it needs beyond the context's resident tables
 * held in a fixed library scratch array of each GPU (about 5 MB, kept until the process ends).
 * No counterpart in the reference: this entry is for callers that cannot allocate (code under
 * HIP graph capture, tight sampling or serving loops, harnesses that forbid hipMalloc).  At most
 * SPR_NOALLOC_MAX_PAIRS pairs per call (SPR_E_ARG beyond); Ndof <= 32.  Calls that overlap on one
 * GPU share its scratch array and race. */
#define SPR_NOALLOC_MAX_PAIRS 4096
int spr_lw3_pairs_noalloc(spr_ctx *ctx, int n, const int *q0_index, const int *q1_grid, double sigma,
                          int prec, double *out);

/* Learned proposal for importance sampling of the linewidth sum: a 3-layer MLP
 * (n_features -> 128 -> 128 -> 1, tanh) predicting log W of a branch triplet from
 * (q0, q1, omegas, branch ids).  No counterpart in the reference.  FP32 training with Adam
 * on the GPU; inference in FP32 or BF16 (MFMA).  Parameters: W1[F][128] b1 W2[128][128] b2
 * W3[128] b3, FP32. */
typedef struct spr_mlp spr_mlp;
int spr_mlp_create(int n_features, unsigned long long seed, int device, spr_mlp **out);
int spr_mlp_destroy(spr_mlp *m);
int spr_mlp_num_params(spr_mlp *m, long *n_params);
int spr_mlp_get_params(spr_mlp *m, float *params);
int spr_mlp_set_params(spr_mlp *m, const float *params);
/* gradient of mean((f(X) - y)^2) w.r.t. the parameters, and the loss (no update) */
int spr_mlp_grad(spr_mlp *m, const float *X, const float *y, long n, float *grad, double *loss);
/* Adam, mini-batches in the given order; loss_per_epoch[epochs] */
int spr_mlp_train(spr_mlp *m, const float *X, const float *y, long n, int epochs, int batch,
                  float lr, double *loss_per_epoch);
/* prec: SPR_PREC_FP32 or SPR_PREC_BF16 */
int spr_mlp_predict(spr_mlp *m, const float *X, long n, int prec, float *out);

/* Proposal scores for importance sampling of the linewidth sum of irreducible q0 q0_index(k),
 * k < n: S_out(nq, n), S(q1) = sum over branch triplets with min|w0 -+ w1 -+ w2| < window*sigma of
 * exp(y_std * f(x)) |G|, G the occupation-weighted Gaussians of CalculateLinewidth
 * (ThirdOrder_m.f90:514-537) and f the network's normalised log W prediction for the triplet's
 * feature row (reduced q0, q1; w0, w1, w2 / wmax; s0, s1, s2 / (Ndof - 1)).  Evaluated on the
 * device with BF16 matrix cores; n_eval (optional) receives the number of network evaluations.
 * m == NULL scores |G| alone (the proposal without a network).  No counterpart in the reference. */
int spr_mlp_score(spr_ctx *ctx, spr_mlp *m, int n, const int *q0_index, double sigma, double window,
                  float y_std, float wmax, double *S_out, long long *n_eval);

/* Pair-level proposal scores (the proposal draws q1 for every branch of a q0 at once).  For
 * each irreducible q0 q0_index(k), k < n, and every q1 (q2 = -(q0 + q1)): a scan of the branch
 * triplets within window*sigma of a resonance gives per s0 (skipping s0 with 1/w0 = 0)
 * g(s0) = sum |G| / (w1 w2) and its w1-, w2- and acoustic-weighted forms; with Gtot(s0) = sum over
 * q1 of g(s0), the pair's resonance weight is S = sum_s0 g(s0) / Gtot(s0).  X_out (optional,
 * [n][nq][12] float) receives the pair's feature row: reduced q0, q1; log(S Nq / ns0); the
 * S-weighted means of w1, w2, w0 over wmax; the acoustic share; log(1 + triplet count) /
 * log(Ndof^3).  S_out [n][nq]: S when m == NULL, else the network's (normalised) output for the
 * row, BF16 matrix cores.  The network must live on the context's device.  No counterpart in the
 * reference. */
int spr_pair_scores(spr_ctx *ctx, spr_mlp *m, int n, const int *q0_index, double sigma, double window,
                    float wmax, double *S_out, float *X_out);

/* Timing of the last launch of the dominant kernel (HIP events on the context's stream):
 * number of launches and summed milliseconds since the last reset. */
int spr_kernel_time(spr_ctx *ctx, int *launches, double *ms);
int spr_kernel_time_reset(spr_ctx *ctx);

/* The launch configuration the context's last work-item launch used: its kernel variant and tile
 * shape as text (shape, capacity cap bytes, NUL-terminated and truncated to fit), the launch mode
 * (0: one workgroup per item; 1, 2, 4: persistent grids of that many workgroups per CU) and the
 * work units per item.  Diagnostics only; no counterpart in the reference. */
int spr_last_launch(spr_ctx *ctx, char *shape, int cap, int *mode, int *units_per_item);

#ifdef __cplusplus
}
#endif
#endif
