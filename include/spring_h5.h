/* spring_h5.h -- HDF5 files of the scattering-rate path (host only, libspring_h5.so).
 *
 * Reads the force-constant files that kappa.x reads (FC2nd_/FC3rd_/FC4th_<prefix><T>K_F.h5,
 * src/Kappa/Src/hdf5_wrap_mod.f90:40-337) and writes the file it writes
 * (kappa_<prefix><T>K.h5, src/Kappa/Src/main.f90:773-798 with hdf5_wrap_mod.f90:650-765), with
 * the same dataset names, types and shapes.  Arrays cross this boundary in the reference's
 * column-major byte layout (C arrays with the Fortran index order reversed), the basis index
 * of atm_indx 1-based as stored in the files.
 *
 * File layout (Fortran dimensions; an HDF5 tool shows them reversed):
 *   NumAtoms               int    (Nb)                          atmNum
 *   atm_Indx               int    (4, Natm, Nb)                 cell (3) + basis atom (1-based)
 *   FC2                    double (L, 3, 3, Natm, Nb)           order 2; fC(1, beta, alpha, N, mu) used
 *   FC3rd/FC3_mu<k>        double (3, 3, 3, Natm, Natm)         order 3, one per basis atom k = 1..Nb
 *   FC4th/FC4_mu<k>        double (3, 3, 3, 3, Natm, Natm, Natm) order 4
 *
 * Every entry point returns 0 on success or a negative code; spr_h5_last_error() gives the
 * message (thread-local). */
#ifndef SPRING_H5_H
#define SPRING_H5_H

#ifdef __cplusplus
extern "C" {
#endif

/* Sizes of an FC file of `order` (2, 3 or 4): basis atoms, neighbour dimension Natm, and for
 * order 2 the leading dimension L of the FC2 dataset (1 in the files ReconstFC2 writes).
 * Replaces the size queries of r2ndFC / r3rdFC / r4thFC (hdf5_wrap_mod.f90:80-90,160-180,270-290). */
int spr_h5_fc_dims(const char *file, int order, int *nb, int *natm, int *lead);

/* Contents: atm_num[nb], atm_indx[nb][natm][4] (atm_Indx(4, Natm, Nb)), and fc:
 *   order 2: fc[nb][natm][3][3]  = FC2(1, beta, alpha, N, mu) (the slice kappa.x uses, DynaMat_mod.f90:157)
 *   order 3: fc[nb][natm][natm][27] = fC(gamma, beta, alpha, N2, N1, mu)
 *   order 4: fc[nb][natm][natm][natm][81]
 * r2ndFC / r3rdFC / r4thFC (hdf5_wrap_mod.f90:40-337). */
int spr_h5_read_fc(const char *file, int order, int nb, int natm, int *atm_num, int *atm_indx, double *fc);

/* Writes an FC file in the same layout (the writers of ReconstFC2/3 and ReconstFC4,
 * e.g. src/ReconstFC3/Src/hdf5_wrap_mod.f90:184-284); order 2 writes L = 1. */
int spr_h5_write_fc(const char *file, int order, int nb, int natm, const int *atm_num, const int *atm_indx,
                    const double *fc);

/* kappa_<prefix><T>K.h5: ThermalConductivityTensor (9), LineWidth (2 Ndof, Nq0) = [omega of the
 * irreducible q0; LW_Allq0], PhaseSpace3rd (2 Ndof, Nq0) when p3 is not NULL, ModeResolvedKappa
 * (9, Ndof) when mode_k is not NULL.  omega_q0, lw, p3: [nq0][ndof]; mode_k: [ndof][9]. */
int spr_h5_write_kappa(const char *file, const double *kappa, int ndof, int nq0, const double *omega_q0,
                       const double *lw, const double *p3, const double *mode_k);

/* Reads the datasets written by spr_h5_write_kappa back (NULL pointers are skipped). */
int spr_h5_read_kappa(const char *file, int ndof, int nq0, double *kappa, double *omega_q0, double *lw,
                      double *p3, double *mode_k);

/* Adds SpectralResolvedKappa (11, Ndof Nq0) to an existing kappa file (w2dh5_no_grp as
 * main.f90:780-794 writes the -Spectral_k output); spectral_k is spr_kappa_rta_spectral's
 * [Ndof nq0][11].  The read returns it back. */
int spr_h5_write_spectral(const char *file, int ndof, int nq0, const double *spectral_k);
int spr_h5_read_spectral(const char *file, int ndof, int nq0, double *spectral_k);

const char *spr_h5_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
