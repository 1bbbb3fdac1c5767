/*
 * oracle.c -- FP64 CPU restatement of SPRING's three/four-phonon scattering-rate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (spring_amd/, libspring_gpu.so)
 * links, loads or calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg do, as the checker.
 *
 * Each function follows the reference routine named in its comment (paths relative to
 * the SPRING source tree).  Array layouts are the reference's column-major arrays with
 * the index order reversed (C row-major), e.g. fc3[mu][N1][N2][a][b][g] == fC(g,b,a,N2,N1,mu),
 * evec[q][s][row] == Evec(row,s,q), phi[s0][s1][s2] == Phi(s2,s1,s0).  All indices are
 * 0-based.
 *
 * Pinning (what checks each function, tests/test_oracle.py):
 *   orc_scatter_mat_el     literal six-deep loop; pinned by brute-force closed forms on
 *                          1-term FC3, linearity, gauge invariance, time reversal, ASR
 *   orc_v3_factorised      == orc_scatter_mat_el to 1e-13 on random inputs
 *   orc_scatter_prob       constants recomputed from their formulas; transpose identity
 *   orc_lw3_gauss          Gaussian against scipy.stats.norm.pdf on hand-built W;
 *                          pair-swap symmetry; brute force over all q1 on tiny meshes
 *   orc_lw3_gauss_pairs    == orc_lw3_gauss over all q1, with the reduced pair list, split lists
 *                          and branch rules (the transposed W stands for the swapped pair)
 *   orc_scatter_mat_el4 / orc_v4_factorised   as the 3-phonon ones
 *   orc_lw4_gauss          process factors and weights (tests/test_lw4_process_terms.py): on
 *                          flat integer bands, exactly on or off shell, the ordered sum equals
 *                          the sum over distinct physical processes (multisets, 1/m! for
 *                          identical modes) of the linearised golden-rule occupation
 *                          differences, which equal the compact factors only on shell
 *   orc_kappa_rta          closed-form single-mode model; symmetric tensor
 *   orc_lw3_tetra          sum of vertex weights == 1; linear-band density of states
 *   orc_w4_rows            the Gaussian four-phonon sum rebuilt from its W4 == orc_lw4_gauss
 *   orc_lw4_tetra_from_w   per process (disjoint energy ranges): integral over Mq == the grid sum
 *                          of F (the delta integrates to one), first moment == the grid mean
 *                          of M interpolated linearly; a single (q1, q2) spike picks n2, n3 of
 *                          q2 and q3 = -(q0+q1+q2); no NaN from degenerate vertices; Gamma skip;
 *                          window integrals on near-flat bands == the same physical-process sum
 *   orc_bte_sums           1/tau == 2 x the tetrahedron linewidth of orc_lw3_tetra; with
 *                          W+ = W- = 1 and F = a w on the full grid, dF == a w 1/tau exactly
 *                          (the vertex weights reproduce Mqs); linearity in F
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define SCTR3 (1.054571817 * (1.60217733 * 1.60217733) * \
               (6.02214076 * 6.02214076 * 6.02214076) * 1000000.0 * M_PI * 0.0625)

/* --------------------------------------------------------------------------------------
 * ScatterMatEl (src/Kappa/Src/ThirdOrder_m.f90:278-386), literally: six-deep loop,
 * phases dcmplx(dcos(q.R), dsin(q.R)) with Cartesian q (1/A) and R = latvec . cell,
 * no conjugation of eigenvectors, branches s < 4 skipped at Gamma (strt = 4).
 * latvec is row-major latvec[i*3+j] = component i of lattice vector j.
 * ------------------------------------------------------------------------------------*/
void orc_scatter_mat_el(int nb, int natm, const int *atm_num, const int *atm_indx,
                        const double *fc3, const double *latvec, const double *osm,
                        const double *q1, const double *q2, const int *q0chk,
                        const cplx *e0, const cplx *e1, const cplx *e2, cplx *phi)
{
    const int ndof = 3 * nb;
    int strt[3];
    for (int i = 0; i < 3; ++i) strt[i] = q0chk[i] ? 3 : 0;
    memset(phi, 0, sizeof(cplx) * ndof * ndof * ndof);
    for (int mu = 0; mu < nb; ++mu) {
        const int na = atm_num[mu];
        for (int n1 = 0; n1 < na; ++n1) {
            const int *ix1 = atm_indx + ((size_t)mu * natm + n1) * 4;
            const int nu = ix1[3];
            double R1[3];
            for (int i = 0; i < 3; ++i)
                R1[i] = latvec[i * 3 + 0] * ix1[0] + latvec[i * 3 + 1] * ix1[1] + latvec[i * 3 + 2] * ix1[2];
            const double qr1 = q1[0] * R1[0] + q1[1] * R1[1] + q1[2] * R1[2];
            const cplx p1 = cos(qr1) + I * sin(qr1);
            for (int n2 = 0; n2 < na; ++n2) {
                const int *ix2 = atm_indx + ((size_t)mu * natm + n2) * 4;
                const int eta = ix2[3];
                double R2[3];
                for (int i = 0; i < 3; ++i)
                    R2[i] = latvec[i * 3 + 0] * ix2[0] + latvec[i * 3 + 1] * ix2[1] + latvec[i * 3 + 2] * ix2[2];
                const double qr2 = q2[0] * R2[0] + q2[1] * R2[1] + q2[2] * R2[2];
                const cplx p2 = cos(qr2) + I * sin(qr2);
                const double mass = osm[mu] * osm[nu] * osm[eta];
                const cplx em = p1 * p2 * mass;
                for (int a = 0; a < 3; ++a) {
                    const int r0 = 3 * mu + a;
                    for (int s = strt[0]; s < ndof; ++s) {
                        const cplx w0 = e0[s * ndof + r0];
                        for (int b = 0; b < 3; ++b) {
                            const int r1 = 3 * nu + b;
                            for (int s1 = strt[1]; s1 < ndof; ++s1) {
                                const cplx w1 = e1[s1 * ndof + r1];
                                for (int g = 0; g < 3; ++g) {
                                    const int r2 = 3 * eta + g;
                                    const double f = fc3[((((size_t)mu * natm + n1) * natm + n2) * 27) + a * 9 + b * 3 + g];
                                    const cplx pm = f * em;
                                    for (int s2 = strt[2]; s2 < ndof; ++s2) {
                                        const cplx w2 = e2[s2 * ndof + r2];
                                        phi[((size_t)s * ndof + s1) * ndof + s2] += pm * w0 * w1 * w2;
                                    }
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

/* Factorised form of the same sum: T(a,b,c) = sum_{N1,N2} Phi p1 p2 m, then three mode
 * products.  Same phases and skips as orc_scatter_mat_el. */
void orc_v3_factorised(int nb, int natm, const int *atm_num, const int *atm_indx,
                       const double *fc3, const double *latvec, const double *osm,
                       const double *q1, const double *q2, const int *q0chk,
                       const cplx *e0, const cplx *e1, const cplx *e2, cplx *phi)
{
    const int nd = 3 * nb;
    cplx *T = calloc((size_t)nd * nd * nd, sizeof(cplx));
    cplx *U = calloc((size_t)nd * nd * nd, sizeof(cplx));
    for (int mu = 0; mu < nb; ++mu) {
        const int na = atm_num[mu];
        for (int n1 = 0; n1 < na; ++n1) {
            const int *ix1 = atm_indx + ((size_t)mu * natm + n1) * 4;
            const int nu = ix1[3];
            double R1[3], R2[3];
            for (int i = 0; i < 3; ++i)
                R1[i] = latvec[i * 3 + 0] * ix1[0] + latvec[i * 3 + 1] * ix1[1] + latvec[i * 3 + 2] * ix1[2];
            const double qr1 = q1[0] * R1[0] + q1[1] * R1[1] + q1[2] * R1[2];
            const cplx p1 = cos(qr1) + I * sin(qr1);
            for (int n2 = 0; n2 < na; ++n2) {
                const int *ix2 = atm_indx + ((size_t)mu * natm + n2) * 4;
                const int eta = ix2[3];
                for (int i = 0; i < 3; ++i)
                    R2[i] = latvec[i * 3 + 0] * ix2[0] + latvec[i * 3 + 1] * ix2[1] + latvec[i * 3 + 2] * ix2[2];
                const double qr2 = q2[0] * R2[0] + q2[1] * R2[1] + q2[2] * R2[2];
                const cplx em = p1 * (cos(qr2) + I * sin(qr2)) * (osm[mu] * osm[nu] * osm[eta]);
                const double *f = fc3 + (((size_t)mu * natm + n1) * natm + n2) * 27;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        for (int g = 0; g < 3; ++g)
                            T[((size_t)(3 * mu + a) * nd + 3 * nu + b) * nd + 3 * eta + g] += f[a * 9 + b * 3 + g] * em;
            }
        }
    }
    int strt[3];
    for (int i = 0; i < 3; ++i) strt[i] = q0chk[i] ? 3 : 0;
    /* U(a,b,s2) = sum_c T(a,b,c) e2(c,s2) */
    for (int a = 0; a < nd; ++a)
        for (int b = 0; b < nd; ++b)
            for (int s2 = 0; s2 < nd; ++s2) {
                cplx acc = 0;
                if (s2 >= strt[2])
                    for (int c = 0; c < nd; ++c) acc += T[((size_t)a * nd + b) * nd + c] * e2[s2 * nd + c];
                U[((size_t)a * nd + b) * nd + s2] = acc;
            }
    /* T(a,s1,s2) = sum_b e1(b,s1) U(a,b,s2) */
    for (int a = 0; a < nd; ++a)
        for (int s1 = 0; s1 < nd; ++s1)
            for (int s2 = 0; s2 < nd; ++s2) {
                cplx acc = 0;
                if (s1 >= strt[1])
                    for (int b = 0; b < nd; ++b) acc += e1[s1 * nd + b] * U[((size_t)a * nd + b) * nd + s2];
                T[((size_t)a * nd + s1) * nd + s2] = acc;
            }
    /* phi(s0,s1,s2) = sum_a e0(a,s0) T(a,s1,s2) */
    for (int s0 = 0; s0 < nd; ++s0)
        for (int s1 = 0; s1 < nd; ++s1)
            for (int s2 = 0; s2 < nd; ++s2) {
                cplx acc = 0;
                if (s0 >= strt[0])
                    for (int a = 0; a < nd; ++a) acc += e0[s0 * nd + a] * T[((size_t)a * nd + s1) * nd + s2];
                phi[((size_t)s0 * nd + s1) * nd + s2] = acc;
            }
    free(T);
    free(U);
}

/* ScatterProbability (ThirdOrder_m.f90:389-439): W(s2,s1,s0) = |Phi|^2 (1/w0 1/w1 1/w2) f,
 * W_per(s1,s2,s0) = the same value (transpose for the swapped pair). */
void orc_scatter_prob(int ndof, const int *q0chk, const double *iw0, const double *iw1,
                      const double *iw2, const cplx *phi, double factor, double *W, double *Wper)
{
    int strt[3];
    for (int i = 0; i < 3; ++i) strt[i] = q0chk[i] ? 3 : 0;
    const size_t n3 = (size_t)ndof * ndof * ndof;
    memset(W, 0, n3 * sizeof(double));
    memset(Wper, 0, n3 * sizeof(double));
    for (int s = strt[0]; s < ndof; ++s)
        for (int s1 = strt[1]; s1 < ndof; ++s1)
            for (int s2 = strt[2]; s2 < ndof; ++s2) {
                const cplx v = phi[((size_t)s * ndof + s1) * ndof + s2];
                const double abs2 = creal(v) * creal(v) + cimag(v) * cimag(v);
                const double el = abs2 * (iw0[s] * iw1[s1] * iw2[s2]);
                W[((size_t)s * ndof + s1) * ndof + s2] = el * factor;
                Wper[((size_t)s * ndof + s2) * ndof + s1] = el * factor;
            }
}

/* Gaussian-delta stage of CalculateLinewidth for one pair (ThirdOrder_m.f90:514-537):
 *   lw(s0) = sum_{s1,s2} W(s2,s1,s0) [ (n1+n2+1) d(w0-w1-w2) + (n1-n2)(d(w0+w1-w2) - d(w0-w1+w2)) ]
 *   d(x) = 1/(sigma sqrt(2 pi)) exp(-x^2/(2 sigma^2))        (:527-529) */
/* Branch selection (PolarizationStrtEnd, Helper_m.f90:310-353): pol = {lo0, hi0, lo1, hi1, ...},
 * 0-based inclusive ranges of s0, s1, s2[, s3]; NULL = every branch.  The sums run over these
 * ranges as CalculateLwTetrahedron's loops do (ThirdOrder_m.f90:207-214), and TetrahedronIntegration4th's
 * (Tetrahedron4th_sep.f90:213-224); s0 outside its range gives 0. */
static int pol_lo(const int *pol, int k) { return pol ? pol[2 * k] : 0; }
static int pol_hi(const int *pol, int k, int nd) { return pol ? pol[2 * k + 1] + 1 : nd; }

void orc_lw3_from_w(int nd, const double *W, const double *w0, const double *w1, const double *w2,
                    const double *n1, const double *n2, double sigma, const int *pol, double *lw)
{
    const double inv_s = 1.0 / sigma, inv_s2 = inv_s * inv_s;
    const double c0 = inv_s / sqrt(2.0 * M_PI);
    for (int s0 = 0; s0 < nd; ++s0) lw[s0] = 0.0;
    for (int s0 = pol_lo(pol, 0); s0 < pol_hi(pol, 0, nd); ++s0) {
        double acc = 0.0;
        for (int s1 = pol_lo(pol, 1); s1 < pol_hi(pol, 1, nd); ++s1)
            for (int s2 = pol_lo(pol, 2); s2 < pol_hi(pol, 2, nd); ++s2) {
                const double d1 = w0[s0] - w1[s1] - w2[s2];
                const double d2 = w0[s0] + w1[s1] - w2[s2];
                const double d3 = w0[s0] - w1[s1] + w2[s2];
                const double g1 = c0 * exp(-0.5 * inv_s2 * d1 * d1);
                const double g2 = c0 * exp(-0.5 * inv_s2 * d2 * d2);
                const double g3 = c0 * exp(-0.5 * inv_s2 * d3 * d3);
                acc += W[((size_t)s0 * nd + s1) * nd + s2] *
                       ((n1[s1] + n2[s2] + 1.0) * g1 + (n1[s1] - n2[s2]) * (g2 - g3));
            }
        lw[s0] = acc;
    }
}

/* --------------------------------------------------------------------------------------
 * Gaussian-delta three-phonon linewidth for one q0, summed over every q1 of the mesh
 * with q2 = -(q0+q1) mod G (CalculateLinewidth, ThirdOrder_m.f90:442-543, with the
 * scattering matrix from ScatterProbAllq1 :41-166):
 *   lw(s0) = sum_q1 sum_{s1,s2} W(s2,s1,s0,q1) [ (n1+n2+1) d(w0-w1-w2)
 *                                              + (n1-n2) (d(w0+w1-w2) - d(w0-w1+w2)) ]
 *   d(x) = 1/(sigma sqrt(2 pi)) exp(-x^2 / (2 sigma^2))
 * The reference's persistent strt skip (:483,500,509) is not reproduced: W is already 0
 * for Gamma acoustic modes (1/omega = 0).  q-points are grid integer addresses; the
 * Cartesian q used for the phases is q_cart[q].
 * ------------------------------------------------------------------------------------*/
void orc_lw3_gauss(int nb, int natm, const int *atm_num, const int *atm_indx,
                   const double *fc3, const double *latvec, const double *osm,
                   int nq, const int *mesh, const int *q_int, const double *q_cart,
                   const cplx *evec, const double *omega, const double *inv_omega,
                   const double *nbe, int i0, double sigma, int q1_lo, int q1_hi,
                   const int *pol, double *lw, double *W_out)
{
    const int nd = 3 * nb;
    const size_t n3 = (size_t)nd * nd * nd;
    cplx *phi = malloc(n3 * sizeof(cplx));
    double *W = malloc(n3 * sizeof(double));
    double *Wp = malloc(n3 * sizeof(double));
    const double factor = SCTR3 / (double)nq;
    for (int s = 0; s < nd; ++s) lw[s] = 0.0;
    const int *k0 = q_int + 3 * i0;
    for (int q1 = q1_lo; q1 < q1_hi; ++q1) {
        const int *k1 = q_int + 3 * q1;
        int k2[3];
        for (int d = 0; d < 3; ++d) {
            int v = -(k0[d] + k1[d]);
            v %= mesh[d];
            if (v < 0) v += mesh[d];
            k2[d] = v;
        }
        const int q2 = k2[0] + mesh[0] * (k2[1] + mesh[1] * k2[2]);
        int chk[3] = {k0[0] == 0 && k0[1] == 0 && k0[2] == 0, k1[0] == 0 && k1[1] == 0 && k1[2] == 0,
                      q2 == 0};
        orc_v3_factorised(nb, natm, atm_num, atm_indx, fc3, latvec, osm, q_cart + 3 * q1,
                          q_cart + 3 * q2, chk, evec + (size_t)i0 * nd * nd,
                          evec + (size_t)q1 * nd * nd, evec + (size_t)q2 * nd * nd, phi);
        orc_scatter_prob(nd, chk, inv_omega + (size_t)i0 * nd, inv_omega + (size_t)q1 * nd,
                         inv_omega + (size_t)q2 * nd, phi, factor, W, Wp);
        if (W_out) memcpy(W_out + (size_t)(q1 - q1_lo) * n3, W, n3 * sizeof(double));
        double lwq[64];
        orc_lw3_from_w(nd, W, omega + (size_t)i0 * nd, omega + (size_t)q1 * nd, omega + (size_t)q2 * nd,
                       nbe + (size_t)q1 * nd, nbe + (size_t)q2 * nd, sigma, pol, lwq);
        for (int s0 = 0; s0 < nd; ++s0) lw[s0] += lwq[s0];
    }
    free(phi);
    free(W);
    free(Wp);
}

/* --------------------------------------------------------------------------------------
 * The same linewidth over the pair-permutation reduced list, as ScatterProbAllq1 evaluates it
 * (ThirdOrder_m.f90:98-164): each listed pair (q1, q2), q2 = -(q0+q1) mod G, is evaluated once;
 * its W is the scattering matrix of q1 and its transpose W_per that of q2 (:134,160), so the
 * pair adds the Gaussian sum of both orderings (only one when q1 == q2).  With the list of every
 * q1 <= q2 this equals orc_lw3_gauss over all q1.  pairs[2k], pairs[2k+1] = (q1, q2).
 * ------------------------------------------------------------------------------------*/
void orc_lw3_gauss_pairs(int nb, int natm, const int *atm_num, const int *atm_indx,
                         const double *fc3, const double *latvec, const double *osm,
                         int nq, const int *q_int, const double *q_cart,
                         const cplx *evec, const double *omega, const double *inv_omega,
                         const double *nbe, int i0, double sigma, const int *pairs, int npairs,
                         const int *pol, double *lw)
{
    const int nd = 3 * nb;
    const size_t n3 = (size_t)nd * nd * nd;
    cplx *phi = malloc(n3 * sizeof(cplx));
    double *W = malloc(n3 * sizeof(double));
    double *Wp = malloc(n3 * sizeof(double));
    const double factor = SCTR3 / (double)nq;
    for (int s = 0; s < nd; ++s) lw[s] = 0.0;
    const int *k0 = q_int + 3 * i0;
    const int q0g = k0[0] == 0 && k0[1] == 0 && k0[2] == 0;
    for (int k = 0; k < npairs; ++k) {
        const int q1 = pairs[2 * k], q2 = pairs[2 * k + 1];
        const int *k1 = q_int + 3 * q1, *k2 = q_int + 3 * q2;
        int chk[3] = {q0g, k1[0] == 0 && k1[1] == 0 && k1[2] == 0, k2[0] == 0 && k2[1] == 0 && k2[2] == 0};
        orc_v3_factorised(nb, natm, atm_num, atm_indx, fc3, latvec, osm, q_cart + 3 * q1,
                          q_cart + 3 * q2, chk, evec + (size_t)i0 * nd * nd,
                          evec + (size_t)q1 * nd * nd, evec + (size_t)q2 * nd * nd, phi);
        orc_scatter_prob(nd, chk, inv_omega + (size_t)i0 * nd, inv_omega + (size_t)q1 * nd,
                         inv_omega + (size_t)q2 * nd, phi, factor, W, Wp);
        double lwq[64];
        orc_lw3_from_w(nd, W, omega + (size_t)i0 * nd, omega + (size_t)q1 * nd, omega + (size_t)q2 * nd,
                       nbe + (size_t)q1 * nd, nbe + (size_t)q2 * nd, sigma, pol, lwq);
        for (int s0 = 0; s0 < nd; ++s0) lw[s0] += lwq[s0];
        if (q1 != q2) {
            orc_lw3_from_w(nd, Wp, omega + (size_t)i0 * nd, omega + (size_t)q2 * nd, omega + (size_t)q1 * nd,
                           nbe + (size_t)q2 * nd, nbe + (size_t)q1 * nd, sigma, pol, lwq);
            for (int s0 = 0; s0 < nd; ++s0) lw[s0] += lwq[s0];
        }
    }
    free(phi);
    free(W);
    free(Wp);
}

/* --------------------------------------------------------------------------------------
 * Four-phonon: ScatterMatEl4 (FourthOrder_m.f90:286-422), literally.
 * fc4[mu][N1][N2][N3][a][b][g][d] == fC(d,g,b,a,N3,N2,N1,mu); phi[s0][s1][s2][s3].
 * ------------------------------------------------------------------------------------*/
void orc_scatter_mat_el4(int nb, int natm, const int *atm_num, const int *atm_indx,
                         const double *fc4, const double *latvec, const double *osm,
                         const double *q1, const double *q2, const double *q3, const int *q0chk,
                         const cplx *e0, const cplx *e1, const cplx *e2, const cplx *e3, cplx *phi)
{
    const int nd = 3 * nb;
    int strt[4];
    for (int i = 0; i < 4; ++i) strt[i] = q0chk[i] ? 3 : 0;
    memset(phi, 0, sizeof(cplx) * nd * nd * nd * nd);
    const double *qs[3] = {q1, q2, q3};
    for (int mu = 0; mu < nb; ++mu) {
        const int na = atm_num[mu];
        for (int n1 = 0; n1 < na; ++n1)
            for (int n2 = 0; n2 < na; ++n2)
                for (int n3 = 0; n3 < na; ++n3) {
                    const int nn[3] = {n1, n2, n3};
                    int at[3];
                    cplx em = osm[mu];
                    for (int k = 0; k < 3; ++k) {
                        const int *ix = atm_indx + ((size_t)mu * natm + nn[k]) * 4;
                        at[k] = ix[3];
                        double R[3];
                        for (int i = 0; i < 3; ++i)
                            R[i] = latvec[i * 3 + 0] * ix[0] + latvec[i * 3 + 1] * ix[1] + latvec[i * 3 + 2] * ix[2];
                        const double qr = qs[k][0] * R[0] + qs[k][1] * R[1] + qs[k][2] * R[2];
                        em *= (cos(qr) + I * sin(qr)) * osm[at[k]];
                    }
                    const double *f = fc4 + ((((size_t)mu * natm + n1) * natm + n2) * natm + n3) * 81;
                    for (int a = 0; a < 3; ++a)
                        for (int s = strt[0]; s < nd; ++s) {
                            const cplx w0 = e0[s * nd + 3 * mu + a];
                            for (int b = 0; b < 3; ++b)
                                for (int s1 = strt[1]; s1 < nd; ++s1) {
                                    const cplx w1 = e1[s1 * nd + 3 * at[0] + b];
                                    for (int g = 0; g < 3; ++g)
                                        for (int s2 = strt[2]; s2 < nd; ++s2) {
                                            const cplx w2 = e2[s2 * nd + 3 * at[1] + g];
                                            for (int d = 0; d < 3; ++d) {
                                                const cplx pm = f[a * 27 + b * 9 + g * 3 + d] * em;
                                                for (int s3 = strt[3]; s3 < nd; ++s3)
                                                    phi[(((size_t)s * nd + s1) * nd + s2) * nd + s3] +=
                                                        pm * w0 * w1 * w2 * e3[s3 * nd + 3 * at[2] + d];
                                            }
                                        }
                                }
                        }
                }
    }
}

/* Factorised four-phonon: T(a,b,c,d) by FT, then four mode products. */
void orc_v4_factorised(int nb, int natm, const int *atm_num, const int *atm_indx,
                       const double *fc4, const double *latvec, const double *osm,
                       const double *q1, const double *q2, const double *q3, const int *q0chk,
                       const cplx *e0, const cplx *e1, const cplx *e2, const cplx *e3, cplx *phi)
{
    const int nd = 3 * nb;
    const size_t n4 = (size_t)nd * nd * nd * nd;
    cplx *T = calloc(n4, sizeof(cplx));
    cplx *U = calloc(n4, sizeof(cplx));
    const double *qs[3] = {q1, q2, q3};
    for (int mu = 0; mu < nb; ++mu) {
        const int na = atm_num[mu];
        for (int n1 = 0; n1 < na; ++n1)
            for (int n2 = 0; n2 < na; ++n2)
                for (int n3 = 0; n3 < na; ++n3) {
                    const double *f = fc4 + ((((size_t)mu * natm + n1) * natm + n2) * natm + n3) * 81;
                    int nz = 0;
                    for (int i = 0; i < 81; ++i) nz |= f[i] != 0.0;
                    if (!nz) continue;
                    const int nn[3] = {n1, n2, n3};
                    int at[3];
                    cplx em = osm[mu];
                    for (int k = 0; k < 3; ++k) {
                        const int *ix = atm_indx + ((size_t)mu * natm + nn[k]) * 4;
                        at[k] = ix[3];
                        double R[3];
                        for (int i = 0; i < 3; ++i)
                            R[i] = latvec[i * 3 + 0] * ix[0] + latvec[i * 3 + 1] * ix[1] + latvec[i * 3 + 2] * ix[2];
                        const double qr = qs[k][0] * R[0] + qs[k][1] * R[1] + qs[k][2] * R[2];
                        em *= (cos(qr) + I * sin(qr)) * osm[at[k]];
                    }
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            for (int g = 0; g < 3; ++g)
                                for (int d = 0; d < 3; ++d)
                                    T[(((size_t)(3 * mu + a) * nd + 3 * at[0] + b) * nd + 3 * at[1] + g) * nd + 3 * at[2] + d] +=
                                        f[a * 27 + b * 9 + g * 3 + d] * em;
                }
    }
    int strt[4];
    for (int i = 0; i < 4; ++i) strt[i] = q0chk[i] ? 3 : 0;
    const size_t nd2 = (size_t)nd * nd, nd3 = nd2 * nd;
    /* contract the last index with e3, then e2, e1, e0 (each pass moves the new mode index
       to the back, so four passes give phi[s0][s1][s2][s3]) */
    const cplx *E[4] = {e3, e2, e1, e0};
    const int st[4] = {strt[3], strt[2], strt[1], strt[0]};
    cplx *src = T, *dst = U;
    for (int pass = 0; pass < 4; ++pass) {
        /* src index (x, y, z, c) with c the index being contracted (last); dst (s, x, y, z) */
        for (size_t xyz = 0; xyz < nd3; ++xyz)
            for (int s = 0; s < nd; ++s) {
                cplx acc = 0;
                if (s >= st[pass])
                    for (int c = 0; c < nd; ++c) acc += src[xyz * nd + c] * E[pass][s * nd + c];
                dst[(size_t)s * nd3 + xyz] = acc;
            }
        cplx *t = src; src = dst; dst = t;
    }
    /* after four passes src holds (s0, s1, s2, s3) */
    memcpy(phi, src, n4 * sizeof(cplx));
    free(T);
    free(U);
}

/* --------------------------------------------------------------------------------------
 * Gaussian-delta four-phonon linewidth for one q0: the tetrahedron routine's processes
 * (Tetrahedron4th_sep.f90:329-343, weights :242-266) with d_sigma(w0 - M) in place of the
 * tetrahedron integral, summed over every (q1, q2) with q3 = -(q0+q1+q2):
 *   lw(s0) = sum W4 [ 1/6 n1 n2 n3 / n0  d(w0 - w1 - w2 - w3)
 *                   + 1/2 (1+n1) n2 n3 / n0 d(w0 + w1 - w2 - w3)
 *                   + 1/2 (1+n1)(1+n2) n3 / n0 d(w0 + w1 + w2 - w3) ]
 * W4 = |V4|^2 prod(1/w) Sctr4THz2 / Nq^2 (ScatterProbability4, FourthOrder_m.f90:425-473).
 * Modes s0 < 3 are skipped at q0 = Gamma (:285-286); the returned lw is the raw sum
 * (main.f90:660 adds 0.5 lw to the linewidth).
 * ------------------------------------------------------------------------------------*/
void orc_lw4_gauss(int nb, int natm, const int *atm_num, const int *atm_indx,
                   const double *fc4, const double *latvec, const double *osm,
                   int nq, const int *mesh, const int *q_int, const double *q_cart,
                   const cplx *evec, const double *omega, const double *inv_omega,
                   const double *nbe, int i0, double sigma, int q1_lo, int q1_hi, const int *pol,
                   double *lw)
{
    const int nd = 3 * nb;
    const size_t n4 = (size_t)nd * nd * nd * nd;
    cplx *phi = malloc(n4 * sizeof(cplx));
    const double sctr4 = 0.125 * M_PI * pow(1.054571817 * 1.602176565, 2) * pow(6.022140857, 4) * 1000000.0;
    const double factor = sctr4 / ((double)nq * (double)nq);
    const double inv_s = 1.0 / sigma, inv_s2 = inv_s * inv_s;
    const double c0 = inv_s / sqrt(2.0 * M_PI);
    for (int s = 0; s < nd; ++s) lw[s] = 0.0;
    const int *k0 = q_int + 3 * i0;
    const int g0 = k0[0] == 0 && k0[1] == 0 && k0[2] == 0;
    for (int q1 = q1_lo; q1 < q1_hi; ++q1)
        for (int q2 = 0; q2 < nq; ++q2) {
            const int *k1 = q_int + 3 * q1, *k2 = q_int + 3 * q2;
            int k3[3];
            for (int d = 0; d < 3; ++d) {
                int v = -(k0[d] + k1[d] + k2[d]);
                v %= mesh[d];
                if (v < 0) v += mesh[d];
                k3[d] = v;
            }
            const int q3 = k3[0] + mesh[0] * (k3[1] + mesh[1] * k3[2]);
            const int chk[4] = {g0, q1 == 0, q2 == 0, q3 == 0};
            orc_v4_factorised(nb, natm, atm_num, atm_indx, fc4, latvec, osm, q_cart + 3 * q1,
                              q_cart + 3 * q2, q_cart + 3 * q3, chk, evec + (size_t)i0 * nd * nd,
                              evec + (size_t)q1 * nd * nd, evec + (size_t)q2 * nd * nd,
                              evec + (size_t)q3 * nd * nd, phi);
            const double *w0 = omega + (size_t)i0 * nd, *w1 = omega + (size_t)q1 * nd;
            const double *w2 = omega + (size_t)q2 * nd, *w3 = omega + (size_t)q3 * nd;
            const double *i0w = inv_omega + (size_t)i0 * nd, *i1w = inv_omega + (size_t)q1 * nd;
            const double *i2w = inv_omega + (size_t)q2 * nd, *i3w = inv_omega + (size_t)q3 * nd;
            const double *n0 = nbe + (size_t)i0 * nd, *n1 = nbe + (size_t)q1 * nd;
            const double *n2 = nbe + (size_t)q2 * nd, *n3 = nbe + (size_t)q3 * nd;
            const int s0a = pol_lo(pol, 0) > (g0 ? 3 : 0) ? pol_lo(pol, 0) : (g0 ? 3 : 0);
            for (int s0 = s0a; s0 < pol_hi(pol, 0, nd); ++s0) {
                double acc = 0.0;
                for (int s1 = pol_lo(pol, 1); s1 < pol_hi(pol, 1, nd); ++s1)
                    for (int s2 = pol_lo(pol, 2); s2 < pol_hi(pol, 2, nd); ++s2)
                        for (int s3 = pol_lo(pol, 3); s3 < pol_hi(pol, 3, nd); ++s3) {
                            const cplx v = phi[(((size_t)s0 * nd + s1) * nd + s2) * nd + s3];
                            const double W = (creal(v) * creal(v) + cimag(v) * cimag(v)) *
                                             (i0w[s0] * i1w[s1] * i2w[s2] * i3w[s3]) * factor;
                            const double x1 = w0[s0] - (w1[s1] + w2[s2] + w3[s3]);
                            const double x2 = w0[s0] - (w2[s2] + w3[s3] - w1[s1]);
                            const double x3 = w0[s0] - (w3[s3] - w1[s1] - w2[s2]);
                            const double d1 = c0 * exp(-0.5 * inv_s2 * x1 * x1);
                            const double d2 = c0 * exp(-0.5 * inv_s2 * x2 * x2);
                            const double d3 = c0 * exp(-0.5 * inv_s2 * x3 * x3);
                            const double F1 = n1[s1] * n2[s2] * n3[s3] / n0[s0];
                            const double F2 = (1.0 + n1[s1]) * n2[s2] * n3[s3] / n0[s0];
                            const double F3 = (1.0 + n1[s1]) * (1.0 + n2[s2]) * n3[s3] / n0[s0];
                            double t = W * (F1 * d1 / 6.0 + 0.5 * F2 * d2 + 0.5 * F3 * d3);
                            if (t != t) t = 0.0;   /* the reference zeroes NaNs (:238-263) */
                            acc += t;
                        }
                lw[s0] += acc;
            }
        }
    free(phi);
}

/* --------------------------------------------------------------------------------------
 * kappa_noIter (ThermalConductivity.f90:36-164): RTA thermal conductivity over the full
 * mesh, tau = 0.5 / lw(s, pos(q)), modes s < 3 skipped at Gamma:
 *   kappa_ab = kW_m_k / (T^2 V Nq) sum_{q,s} w^2 n (n+1) tau v_a v_b     (W/m K)
 * lw_irr[q0][s], pos[q] = irreducible position of grid point q, vel[q][s][3].  srange = {lo, hi}
 * (0-based inclusive) is the mode range default_strt..default_end of the s0 selection rule
 * (:106-111); at q = Gamma the loop starts at s = 3 (:123-124); NULL = every mode.
 * ------------------------------------------------------------------------------------*/
void orc_kappa_rta(int nd, int nq, const int *q_int, const int *pos, const double *omega,
                   const double *nbe, const double *vel, const double *lw_irr, double T,
                   double vol, const int *srange, double *kappa)
{
    const double kw = (1.054571817 * 1.054571817) * 0.001 / 1.380649;
    const double f2 = kw / (T * T * vol) / (double)nq;
    double k[9] = {0};
    for (int q = 0; q < nq; ++q) {
        const int g = q_int[3 * q] == 0 && q_int[3 * q + 1] == 0 && q_int[3 * q + 2] == 0;
        for (int s = g ? 3 : pol_lo(srange, 0); s < pol_hi(srange, 0, nd); ++s) {
            const double w = omega[(size_t)q * nd + s], n = nbe[(size_t)q * nd + s];
            const double tau = 0.5 / lw_irr[(size_t)pos[q] * nd + s];
            const double *v = vel + ((size_t)q * nd + s) * 3;
            const double c = w * w * n * (n + 1.0) * tau;
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) k[a * 3 + b] += c * v[a] * v[b];
        }
    }
    for (int i = 0; i < 9; ++i) kappa[i] = k[i] * f2;
}

/* kappa_noIter_symm + GrVelKronProdSymm (ThermalConductivity.f90:167-293,
 * VelocityKronProd_symm.f90:34-93): sum over irreducible q0 only, with
 * (mult/Nsym) sum_R (R v)(R v)^T in place of v v^T.  rot[r][3][3] Cartesian. */
void orc_kappa_rta_symm(int nd, int nq, int nq0, const int *irr_q_int, const int *mult,
                        int nsym, const double *rot, const double *omega0, const double *nbe0,
                        const double *vel0, const double *lw_irr, double T, double vol,
                        const int *srange, double *kappa)
{
    const double kw = (1.054571817 * 1.054571817) * 0.001 / 1.380649;
    const double f2 = kw / (T * T * vol) / (double)nq;
    double k[9] = {0};
    for (int q = 0; q < nq0; ++q) {
        const int *ki = irr_q_int + 3 * q;
        const int g = ki[0] == 0 && ki[1] == 0 && ki[2] == 0;
        const double p = 1.0 / (double)(nsym / mult[q]);
        for (int s = g ? 3 : pol_lo(srange, 0); s < pol_hi(srange, 0, nd); ++s) {
            const double *v = vel0 + ((size_t)q * nd + s) * 3;
            double kr[9] = {0};
            for (int r = 0; r < nsym; ++r) {
                const double *R = rot + r * 9;
                double rv[3];
                for (int a = 0; a < 3; ++a) rv[a] = R[a * 3] * v[0] + R[a * 3 + 1] * v[1] + R[a * 3 + 2] * v[2];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) kr[a * 3 + b] += rv[a] * rv[b];
            }
            const double w = omega0[(size_t)q * nd + s], n = nbe0[(size_t)q * nd + s];
            const double tau = 0.5 / lw_irr[(size_t)q * nd + s];
            const double c = w * w * n * (n + 1.0) * tau;
            for (int i = 0; i < 9; ++i) k[i] += c * kr[i] * p;
        }
    }
    for (int i = 0; i < 9; ++i) kappa[i] = k[i] * f2;
}

/* --------------------------------------------------------------------------------------
 * Linear tetrahedron method for one tetrahedron (FindgAndIver and FindIverxFver,
 * TetrahedronIntegration_m.f90:93-295): vertex values M(4) are sorted ascending (stable for
 * ties, the order my_sort's first-uncovered search produces), then the three cases give
 * g (density at Mqs) and the vertex weights I(4); returns g * sum(I F) / 6 in *gIF and g.
 * ------------------------------------------------------------------------------------*/
static int tetra_core(double Mqs, const double *M_in, const double *F_in, double *g_out, double *I_out,
                      double *F_out)
{
    double M[4], F[4];
    int idx[4] = {0, 1, 2, 3};
    for (int a = 1; a < 4; ++a)   /* stable insertion sort of the indices by value */
        for (int b = a; b > 0 && M_in[idx[b - 1]] > M_in[idx[b]]; --b) {
            int t = idx[b]; idx[b] = idx[b - 1]; idx[b - 1] = t;
        }
    for (int v = 0; v < 4; ++v) { M[v] = M_in[idx[v]]; F[v] = F_in[idx[v]]; }
    double g = 0.0, Iv[4] = {0, 0, 0, 0};
    int enter = 0;
    if (Mqs >= M[0] && Mqs < M[1]) {
        g = 3.0 * (Mqs - M[0]) * (Mqs - M[0]) / ((M[1] - M[0]) * (M[2] - M[0]) * (M[3] - M[0]));
        Iv[0] = ((Mqs - M[1]) / (M[0] - M[1]) + (Mqs - M[2]) / (M[0] - M[2]) + (Mqs - M[3]) / (M[0] - M[3])) / 3.0;
        for (int v = 1; v < 4; ++v) Iv[v] = ((Mqs - M[0]) / (M[v] - M[0])) / 3.0;
        enter = 1;
    } else if (Mqs >= M[1] && Mqs < M[2]) {
        g = (3.0 / ((M[1] - M[2]) * (M[3] - M[0]))) *
            (((Mqs - M[0]) * (Mqs - M[2])) / (M[2] - M[0]) - ((Mqs - M[1]) * (Mqs - M[3])) / (M[1] - M[3]));
        Iv[0] = ((Mqs - M[3]) / (M[0] - M[3])) / 3.0 +
               ((Mqs - M[2]) / (M[0] - M[2])) * ((Mqs - M[0]) / (M[2] - M[0])) * ((Mqs - M[2]) / (M[1] - M[2])) /
                   (g * (M[3] - M[0]));
        Iv[1] = ((Mqs - M[2]) / (M[1] - M[2])) / 3.0 +
               ((Mqs - M[3]) / (M[1] - M[3])) * ((Mqs - M[3]) / (M[1] - M[3])) * ((Mqs - M[1]) / (M[2] - M[1])) /
                   (g * (M[3] - M[0]));
        Iv[2] = ((Mqs - M[1]) / (M[2] - M[1])) / 3.0 +
               ((Mqs - M[0]) / (M[2] - M[0])) * ((Mqs - M[0]) / (M[2] - M[0])) * ((Mqs - M[2]) / (M[1] - M[2])) /
                   (g * (M[3] - M[0]));
        Iv[3] = ((Mqs - M[0]) / (M[3] - M[0])) / 3.0 +
               ((Mqs - M[1]) / (M[3] - M[1])) * ((Mqs - M[3]) / (M[1] - M[3])) * ((Mqs - M[1]) / (M[2] - M[1])) /
                   (g * (M[3] - M[0]));
        enter = 1;
    } else if (Mqs >= M[2] && Mqs <= M[3]) {
        g = 3.0 * (Mqs - M[3]) * (Mqs - M[3]) / ((M[3] - M[0]) * (M[3] - M[1]) * (M[3] - M[2]));
        for (int v = 0; v < 3; ++v) Iv[v] = ((Mqs - M[3]) / (M[v] - M[3])) / 3.0;
        Iv[3] = ((Mqs - M[0]) / (M[3] - M[0]) + (Mqs - M[1]) / (M[3] - M[1]) + (Mqs - M[2]) / (M[3] - M[2])) / 3.0;
        enter = 1;
    }
    *g_out = g;
    for (int v = 0; v < 4; ++v) { I_out[v] = Iv[v]; F_out[v] = F[v]; }
    return enter;
}

int orc_tetra_weights(double Mqs, const double *M_in, const double *F_in, double *g_out, double *I_out,
                      double *gIF_out)
{
    double g, Iv[4], F[4];
    const int enter = tetra_core(Mqs, M_in, F_in, &g, Iv, F);
    double ixf = 0.0;
    if (enter)
        for (int v = 0; v < 4; ++v) ixf += Iv[v] * F[v];
    ixf /= 6.0;
    if (g_out) *g_out = g;
    if (I_out) for (int v = 0; v < 4; ++v) I_out[v] = Iv[v];
    if (gIF_out) *gIF_out = g * ixf;
    return enter;
}

/* CalculateLwTetrahedron (ThirdOrder_m.f90:169-275) for one q0 given W(s2,s1,s0,q1) of every q1
 * (W[q1][s0][s1][s2]), with q2 = partner(q1) (Allq1q2(4,:)): processes M1 = w1+w2 with
 * F1 = W(n1+n2+1), M2 = w2-w1 with F2 = W(n1-n2), M3 = w1-w2 with F3 = F2 entering with a minus
 * sign; P3_q0 accumulates g of all three.  tetra[6 nq][4] grid indices. */
void orc_lw3_tetra(int nd, int nq, const int *tetra, const int *partner, const double *W,
                   const double *omega, const double *nbe, const double *omega0, const int *pol,
                   double *lw, double *p3q0)
{
    for (int s = 0; s < nd; ++s) { lw[s] = 0.0; p3q0[s] = 0.0; }
    for (int s0 = pol_lo(pol, 0); s0 < pol_hi(pol, 0, nd); ++s0) {
        const double Mq0 = omega0[s0];
        double acc = 0.0, pacc = 0.0;
        for (int t = 0; t < 6 * nq; ++t) {
            int v1[4], v2[4];
            for (int v = 0; v < 4; ++v) { v1[v] = tetra[4 * t + v]; v2[v] = partner[v1[v]]; }
            for (int s1 = pol_lo(pol, 1); s1 < pol_hi(pol, 1, nd); ++s1)
                for (int s2 = pol_lo(pol, 2); s2 < pol_hi(pol, 2, nd); ++s2) {
                    double F1[4], F2[4], F3[4], M1[4], M2[4], M3[4];
                    for (int v = 0; v < 4; ++v) {
                        const double w = W[(((size_t)v1[v] * nd + s0) * nd + s1) * nd + s2];
                        const double n1 = nbe[(size_t)v1[v] * nd + s1], n2 = nbe[(size_t)v2[v] * nd + s2];
                        const double w1 = omega[(size_t)v1[v] * nd + s1], w2 = omega[(size_t)v2[v] * nd + s2];
                        F1[v] = w * (n1 + n2 + 1.0);
                        F2[v] = w * (n1 - n2);
                        F3[v] = F2[v];
                        M1[v] = w1 + w2;
                        M2[v] = w2 - w1;
                        M3[v] = w1 - w2;
                    }
                    double g, gif;
                    orc_tetra_weights(Mq0, M1, F1, &g, 0, &gif); acc += gif; pacc += g;
                    orc_tetra_weights(Mq0, M2, F2, &g, 0, &gif); acc += gif; pacc += g;
                    orc_tetra_weights(Mq0, M3, F3, &g, 0, &gif); acc -= gif; pacc += g;
                }
        }
        lw[s0] = acc;
        p3q0[s0] = pacc;
    }
}


/* --------------------------------------------------------------------------------------
 * Four-phonon tetrahedron linewidth (TetrahedronIntegration4th, SumOverPolarization and
 * FindFandOmegaVer_4th, Tetrahedron4th_sep.f90:37-425).
 *
 * orc_w4_rows: W4(s3,s2,s1,s0; q1, q2) of irreducible q0 `i0` for q1 in [q1_lo, q1_hi) and
 * every q2, q3 = -(q0+q1+q2) mod G: W[q1-q1_lo][q2][s0][s1][s2][s3] =
 * |V4|^2 prod(1/w) Sctr4THz2/Nq^2 (ScatterProbability4, FourthOrder_m.f90:425-473).  The
 * reference stores W4 once per permutation class of (q1,q2,q3) and reads the other orders by
 * permuting (s1,s2,s3) (FindFandOmegaVer_4th :296-320, MomentumConsrv4th,
 * QuasiMomentumConsrv4th_m.f90:32-206); with an index-permutation-symmetric FC4 that is the
 * same number as W4 evaluated for the ordered triple, which is what is computed here.
 * ------------------------------------------------------------------------------------*/
static int grid_q3(const int *mesh, const int *k0, const int *k1, const int *k2)
{
    int k3[3];
    for (int d = 0; d < 3; ++d) {
        int v = -(k0[d] + k1[d] + k2[d]);
        v %= mesh[d];
        if (v < 0) v += mesh[d];
        k3[d] = v;
    }
    return k3[0] + mesh[0] * (k3[1] + mesh[1] * k3[2]);
}

void orc_w4_rows(int nb, int natm, const int *atm_num, const int *atm_indx, const double *fc4,
                 const double *latvec, const double *osm, int nq, const int *mesh, const int *q_int,
                 const double *q_cart, const cplx *evec, const double *inv_omega, int i0, int q1_lo,
                 int q1_hi, double *W)
{
    const int nd = 3 * nb;
    const size_t n4 = (size_t)nd * nd * nd * nd;
    cplx *phi = malloc(n4 * sizeof(cplx));
    const double sctr4 = 0.125 * M_PI * pow(1.054571817 * 1.602176565, 2) * pow(6.022140857, 4) * 1000000.0;
    const double factor = sctr4 / ((double)nq * (double)nq);
    const int *k0 = q_int + 3 * i0;
    const int g0 = k0[0] == 0 && k0[1] == 0 && k0[2] == 0;
    for (int q1 = q1_lo; q1 < q1_hi; ++q1)
        for (int q2 = 0; q2 < nq; ++q2) {
            const int q3 = grid_q3(mesh, k0, q_int + 3 * q1, q_int + 3 * q2);
            const int chk[4] = {g0, q1 == 0, q2 == 0, q3 == 0};
            orc_v4_factorised(nb, natm, atm_num, atm_indx, fc4, latvec, osm, q_cart + 3 * q1,
                              q_cart + 3 * q2, q_cart + 3 * q3, chk, evec + (size_t)i0 * nd * nd,
                              evec + (size_t)q1 * nd * nd, evec + (size_t)q2 * nd * nd,
                              evec + (size_t)q3 * nd * nd, phi);
            const double *i0w = inv_omega + (size_t)i0 * nd, *i1w = inv_omega + (size_t)q1 * nd;
            const double *i2w = inv_omega + (size_t)q2 * nd, *i3w = inv_omega + (size_t)q3 * nd;
            double *out = W + ((size_t)(q1 - q1_lo) * nq + q2) * n4;
            for (int s0 = 0; s0 < nd; ++s0)
                for (int s1 = 0; s1 < nd; ++s1)
                    for (int s2 = 0; s2 < nd; ++s2)
                        for (int s3 = 0; s3 < nd; ++s3) {
                            const size_t e = (((size_t)s0 * nd + s1) * nd + s2) * nd + s3;
                            const cplx v = phi[e];
                            out[e] = (creal(v) * creal(v) + cimag(v) * cimag(v)) *
                                     (i0w[s0] * i1w[s1] * i2w[s2] * i3w[s3]) * factor;
                        }
        }
    free(phi);
}

/* One process of SumOverPolarization (:242-263): FindgAndIver on the vertex values, NaN
 * weights and density set to zero, FindIverxFver (includes 1/6), returns g * IxF. */
static double tetra4_term(double Mq, const double *M, const double *F)
{
    double g, Iv[4], Fs[4];
    const int enter = tetra_core(Mq, M, F, &g, Iv, Fs);
    for (int v = 0; v < 4; ++v)
        if (Iv[v] != Iv[v]) Iv[v] = 0.0;
    if (g != g) g = 0.0;
    double ixf = 0.0;
    if (enter)
        for (int v = 0; v < 4; ++v) ixf += Iv[v] * Fs[v];
    ixf = (1.0 / 6.0) * ixf;
    return g * ixf;
}

/* orc_lw4_tetra_from_w: the q1 in [q1_lo, q1_hi) part of lw(s0) of q0 `i0` from
 * W[q1-q1_lo][q2][s0][s1][s2][s3] (orc_w4_rows).  For every tetrahedron of the q2 mesh
 * (tetra[6 nq][4]) the vertices carry q2_v and q3_v = -(q0+q1+q2_v), and
 *   M1 = w1 + w2 + w3,  F1 = W n1 n2 n3 / n0,              added with 1/6
 *   M2 = w2 + w3 - w1,  F2 = W (1+n1) n2 n3 / n0,          added with 1/2
 *   M3 = w3 - w1 - w2,  F3 = W (1+n1)(1+n2) n3 / n0,       added with 1/2
 * at Mq = omega0[s0] (:329-343).  omega0 and nbe0 are the q0's omega and n (the caller passes
 * omega[i0] and nbe[i0]); s0 < 3 skipped at q0 = Gamma (:285-286).  Raw sum; main.f90:660 adds
 * 0.5 lw to the linewidth. */
void orc_lw4_tetra_from_w(int nd, int nq, const int *mesh, const int *q_int, const int *tetra, int i0,
                          int q1_lo, int q1_hi, const double *W, const double *omega, const double *nbe,
                          const double *omega0, const double *nbe0, const int *pol, double *lw)
{
    const size_t n4 = (size_t)nd * nd * nd * nd;
    const int *k0 = q_int + 3 * i0;
    const int g0 = k0[0] == 0 && k0[1] == 0 && k0[2] == 0;
    for (int s = 0; s < nd; ++s) lw[s] = 0.0;
    for (int q1 = q1_lo; q1 < q1_hi; ++q1) {
        const double *Wq1 = W + (size_t)(q1 - q1_lo) * nq * n4;
        for (int t = 0; t < 6 * nq; ++t) {
            int q2v[4], q3v[4];
            for (int v = 0; v < 4; ++v) {
                q2v[v] = tetra[4 * t + v];
                q3v[v] = grid_q3(mesh, k0, q_int + 3 * q1, q_int + 3 * q2v[v]);
            }
            const int s0a = pol_lo(pol, 0) > (g0 ? 3 : 0) ? pol_lo(pol, 0) : (g0 ? 3 : 0);
            for (int s0 = s0a; s0 < pol_hi(pol, 0, nd); ++s0) {
                const double Mq0 = omega0[s0], n0 = nbe0[s0];
                for (int s1 = pol_lo(pol, 1); s1 < pol_hi(pol, 1, nd); ++s1) {
                    const double w1 = omega[(size_t)q1 * nd + s1], n1 = nbe[(size_t)q1 * nd + s1];
                    for (int s2 = pol_lo(pol, 2); s2 < pol_hi(pol, 2, nd); ++s2)
                        for (int s3 = pol_lo(pol, 3); s3 < pol_hi(pol, 3, nd); ++s3) {
                            double F1[4], F2[4], F3[4], M1[4], M2[4], M3[4];
                            for (int v = 0; v < 4; ++v) {
                                const double w = Wq1[(size_t)q2v[v] * n4 + (((size_t)s0 * nd + s1) * nd + s2) * nd + s3];
                                const double n2 = nbe[(size_t)q2v[v] * nd + s2], n3 = nbe[(size_t)q3v[v] * nd + s3];
                                const double w2 = omega[(size_t)q2v[v] * nd + s2], w3 = omega[(size_t)q3v[v] * nd + s3];
                                F1[v] = w * (n1 * n2 * n3 / n0);
                                M1[v] = w1 + w2 + w3;
                                F2[v] = w * ((1.0 + n1) * n2 * n3 / n0);
                                M2[v] = w2 + w3 - w1;
                                F3[v] = w * ((1.0 + n1) * (1.0 + n2) * n3 / n0);
                                M3[v] = w3 - w1 - w2;
                            }
                            lw[s0] += (1.0 / 6.0) * tetra4_term(Mq0, M1, F1);
                            lw[s0] += 0.5 * tetra4_term(Mq0, M2, F2);
                            lw[s0] += 0.5 * tetra4_term(Mq0, M3, F3);
                        }
                }
            }
        }
    }
}

/* --------------------------------------------------------------------------------------
 * Iterative BTE (src/Kappa/Src_iterative).  Linear tetrahedron weights of that code
 * (TetrahedronWplus, Quasi_momentum_consrv.f90:815-900): strict inequalities in all three
 * cases, the vertex weights must sum to one within zero_prec = 5e-5 (else nothing is added);
 * vertices sorted by M with their data (my_sort, :624-698, stable for ties).  Adds
 * mul * g/6 * sum_v I_v X_v[k] to acc[k] for k < K.
 * ------------------------------------------------------------------------------------*/
static void orc_tetra_iter(double Mqs, const double *M_in, const double *X_in, int K, double mul, double *acc)
{
    int idx[4] = {0, 1, 2, 3};
    for (int a = 1; a < 4; ++a)
        for (int b = a; b > 0 && M_in[idx[b - 1]] > M_in[idx[b]]; --b) {
            int t = idx[b]; idx[b] = idx[b - 1]; idx[b - 1] = t;
        }
    double M[4];
    for (int v = 0; v < 4; ++v) M[v] = M_in[idx[v]];
    double g, Iv[4];
    if (Mqs > M[0] && Mqs < M[1]) {
        g = 3.0 * (Mqs - M[0]) * (Mqs - M[0]) / ((M[1] - M[0]) * (M[2] - M[0]) * (M[3] - M[0]));
        Iv[0] = ((Mqs - M[1]) / (M[0] - M[1]) + (Mqs - M[2]) / (M[0] - M[2]) + (Mqs - M[3]) / (M[0] - M[3])) / 3.0;
        for (int v = 1; v < 4; ++v) Iv[v] = ((Mqs - M[0]) / (M[v] - M[0])) / 3.0;
    } else if (Mqs > M[1] && Mqs < M[2]) {
        g = (3.0 / ((M[1] - M[2]) * (M[3] - M[0]))) *
            (((Mqs - M[0]) * (Mqs - M[2])) / (M[2] - M[0]) - ((Mqs - M[1]) * (Mqs - M[3])) / (M[1] - M[3]));
        Iv[0] = ((Mqs - M[3]) / (M[0] - M[3])) / 3.0 +
               ((Mqs - M[2]) / (M[0] - M[2])) * ((Mqs - M[0]) / (M[2] - M[0])) * ((Mqs - M[2]) / (M[1] - M[2])) /
                   (g * (M[3] - M[0]));
        Iv[1] = ((Mqs - M[2]) / (M[1] - M[2])) / 3.0 +
               ((Mqs - M[3]) / (M[1] - M[3])) * ((Mqs - M[3]) / (M[1] - M[3])) * ((Mqs - M[1]) / (M[2] - M[1])) /
                   (g * (M[3] - M[0]));
        Iv[2] = ((Mqs - M[1]) / (M[2] - M[1])) / 3.0 +
               ((Mqs - M[0]) / (M[2] - M[0])) * ((Mqs - M[0]) / (M[2] - M[0])) * ((Mqs - M[2]) / (M[1] - M[2])) /
                   (g * (M[3] - M[0]));
        Iv[3] = ((Mqs - M[0]) / (M[3] - M[0])) / 3.0 +
               ((Mqs - M[1]) / (M[3] - M[1])) * ((Mqs - M[3]) / (M[1] - M[3])) * ((Mqs - M[1]) / (M[2] - M[1])) /
                   (g * (M[3] - M[0]));
    } else if (Mqs > M[2] && Mqs < M[3]) {
        g = 3.0 * (Mqs - M[3]) * (Mqs - M[3]) / ((M[3] - M[0]) * (M[3] - M[1]) * (M[3] - M[2]));
        for (int v = 0; v < 3; ++v) Iv[v] = ((Mqs - M[3]) / (M[v] - M[3])) / 3.0;
        Iv[3] = ((Mqs - M[0]) / (M[3] - M[0]) + (Mqs - M[1]) / (M[3] - M[1]) + (Mqs - M[2]) / (M[3] - M[2])) / 3.0;
    } else {
        return;
    }
    if (!(fabs(Iv[0] + Iv[1] + Iv[2] + Iv[3] - 1.0) < 5.0e-5)) return;
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int v = 0; v < 4; ++v) t += Iv[v] * X_in[idx[v] * K + k];
        acc[k] += mul * (1.0 / 6.0) * g * t;
    }
}

static int orc_grid(const int *mesh, int a0, int a1, int a2)
{
    a0 = ((a0 % mesh[0]) + mesh[0]) % mesh[0];
    a1 = ((a1 % mesh[1]) + mesh[1]) % mesh[1];
    a2 = ((a2 % mesh[2]) + mesh[2]) % mesh[2];
    return a0 + mesh[0] * (a1 + mesh[1] * a2);
}

/* TetrahedronWplus + TetrahedronWminus (Quasi_momentum_consrv.f90:701-1208) when F == NULL:
 *   out[qi][s] = 1/tau;
 * TetrahedronIterPlus + TetrahedronIterMinus (TetrahedronDeltaSum.f90:22-620) otherwise:
 *   out[qi][s][3] = dF before the division by 1/tau.
 * For the irreducible point q = irr_grid[qi] and every tetrahedron T (tetra[6 nq][4] grid
 * indices), every s, s1, s2, vertex q1 of T:
 *   +: M = w(q+q1, s2) - w(q1, s1), X = Wp[qi][q1][s][s1][s2] (F(q+q1, s2) - F(q1, s1))
 *   -: M = w(q1, s1) + w(q-q1, s2),  X = Wm[qi][-q1][s][s1][s2] (F(q-q1, s2) + F(q1, s1)), x 1/2
 * Wp / Wm[qi][q1] are W+ / W- of the pair (q1, -q-q1).  F of grid point g is
 * map_sign[g] * F[map_irr[g]] (F(-q) = -F(q)); X = W when F == NULL. */
void orc_bte_sums(int nd, int nq, const int *mesh, int nirr, const int *irr_grid, const int *map_irr,
                  const int *map_sign, const int *tetra, const double *Wp, const double *Wm,
                  const double *omega, const double *F, double *out)
{
    const int K = F ? 3 : 1;
    const size_t per_q = (size_t)nq * nd * nd * nd;
    memset(out, 0, sizeof(double) * (size_t)nirr * nd * K);
    for (int qi = 0; qi < nirr; ++qi) {
        const int q = irr_grid[qi];
        const int k0 = q % mesh[0], k1 = (q / mesh[0]) % mesh[1], k2 = q / (mesh[0] * mesh[1]);
        for (int t = 0; t < 6 * nq; ++t) {
            int vq1[4], vp[4], vm[4], vn[4];
            for (int v = 0; v < 4; ++v) {
                const int g = tetra[4 * t + v];
                const int a0 = g % mesh[0], a1 = (g / mesh[0]) % mesh[1], a2 = g / (mesh[0] * mesh[1]);
                vq1[v] = g;
                vp[v] = orc_grid(mesh, k0 + a0, k1 + a1, k2 + a2);
                vm[v] = orc_grid(mesh, k0 - a0, k1 - a1, k2 - a2);
                vn[v] = orc_grid(mesh, -a0, -a1, -a2);
            }
            for (int s = 0; s < nd; ++s) {
                const double Mqs = omega[(size_t)q * nd + s];
                for (int s1 = 0; s1 < nd; ++s1)
                    for (int s2 = 0; s2 < nd; ++s2) {
                        double Mp[4], Mm[4], Xp[12], Xm[12];
                        for (int v = 0; v < 4; ++v) {
                            const size_t off = (size_t)s * nd * nd + (size_t)s1 * nd + s2;
                            const double wp = Wp[(size_t)qi * per_q + (size_t)vq1[v] * nd * nd * nd + off];
                            const double wm = Wm[(size_t)qi * per_q + (size_t)vn[v] * nd * nd * nd + off];
                            Mp[v] = omega[(size_t)vp[v] * nd + s2] - omega[(size_t)vq1[v] * nd + s1];
                            Mm[v] = omega[(size_t)vq1[v] * nd + s1] + omega[(size_t)vm[v] * nd + s2];
                            if (!F) {
                                Xp[v] = wp;
                                Xm[v] = wm;
                            } else {
                                for (int k = 0; k < 3; ++k) {
                                    const double f1 = map_sign[vq1[v]] * F[((size_t)map_irr[vq1[v]] * nd + s1) * 3 + k];
                                    const double fp = map_sign[vp[v]] * F[((size_t)map_irr[vp[v]] * nd + s2) * 3 + k];
                                    const double fm = map_sign[vm[v]] * F[((size_t)map_irr[vm[v]] * nd + s2) * 3 + k];
                                    Xp[v * 3 + k] = wp * (fp - f1);
                                    Xm[v * 3 + k] = wm * (fm + f1);
                                }
                            }
                        }
                        double *o = out + ((size_t)qi * nd + s) * K;
                        orc_tetra_iter(Mqs, Mp, Xp, K, 1.0, o);
                        orc_tetra_iter(Mqs, Mm, Xm, K, 0.5, o);
                    }
            }
        }
    }
}
