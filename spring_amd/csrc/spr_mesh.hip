// spr_mesh_build: the q mesh of kappa.x built on the host, so that a caller with only the crystal
// (&CrystalInfo / &AtmInfo) and the mesh (&AnhInfo qmesh, time_rev) gets every table
// spr_ctx_create takes.  Replaces make_tetrahedron + FindIrrQ (src/Kappa/Src/TetrahedronQ_m.f90:
// 56-331) and the rotations of set_SpaceGrSPGlib; spglib is not used:
//   * grid point i = sum_k mod(a_k, N_k) prod_{j<k} N_j, a folded into (-N/2, N/2] (fold_indx),
//     q = sum_k b_k a_k / N_k (reciprocal vectors b_k = 2 pi (A^-T)_{:,k});
//   * six tetrahedra per mesh cube, cubes in grid order, corner and vertex order as :103-110, 172-188;
//   * the point group: every integer matrix with entries in {-1, 0, 1} that keeps the lattice
//     metric and maps the basis onto itself, species by species, up to one common translation
//     (positions compared modulo lattice vectors within symprec Angstrom);
//   * irreducible points: orbits of the grid under k -> R^T k (and -R^T k with time reversal); the
//     representative is the smallest grid index of the orbit, as spg_get_ir_reciprocal_mesh maps
//     it, with pos and multiplicity as FindPositionAndMultiplicity gives them.
// Host code only (no device work).
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "spr_internal.h"

namespace {

double det3(const double m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
}

bool inv3(const double m[3][3], double out[3][3]) {
    const double d = det3(m);
    if (!(std::fabs(d) > 0.0)) return false;
    out[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / d;
    out[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / d;
    out[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / d;
    out[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / d;
    out[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / d;
    out[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / d;
    out[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / d;
    out[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / d;
    out[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / d;
    return true;
}

int fold1(int k, int n) {
    k %= n;
    if (k < 0) k += n;
    return k > n / 2 ? k - n : k;
}

// point-group rotations R (integer, fractional coordinates) of the crystal, in the order of the
// enumeration of the nine entries over (-1, 0, 1) with the last entry fastest
std::vector<std::array<int, 9>> point_group(const double A[3][3], int nb, const double *frac, const int *typ,
                                            double symprec) {
    double M[3][3], mmax = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            M[i][j] = A[0][i] * A[0][j] + A[1][i] * A[1][j] + A[2][i] * A[2][j];
            mmax = std::max(mmax, std::fabs(M[i][j]));
        }
    const double mtol = 1e-4 * mmax;
    std::vector<std::array<int, 9>> out;
    int e[9];
    for (int code = 0; code < 19683; ++code) {   // 3^9
        int c = code;
        for (int k = 8; k >= 0; --k) {
            e[k] = c % 3 - 1;
            c /= 3;
        }
        double R[3][3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R[i][j] = e[3 * i + j];
        if (std::fabs(std::fabs(det3(R)) - 1.0) > 1e-9) continue;
        bool metric_ok = true;
        for (int i = 0; i < 3 && metric_ok; ++i)
            for (int j = 0; j < 3 && metric_ok; ++j) {
                double s = 0.0;   // (R^T M R)_ij
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) s += R[k][i] * M[k][l] * R[l][j];
                metric_ok = std::fabs(s - M[i][j]) <= mtol;
            }
        if (!metric_ok) continue;
        // rotated positions rp_i = R frac_i; accept when some translation t = frac_j - rp_0 (j of
        // the species of atom 0) sends every rotated atom onto an atom of its species
        std::vector<double> rp((size_t)nb * 3);
        for (int i = 0; i < nb; ++i)
            for (int r = 0; r < 3; ++r)
                rp[i * 3 + r] = R[r][0] * frac[i * 3 + 0] + R[r][1] * frac[i * 3 + 1] + R[r][2] * frac[i * 3 + 2];
        bool any = false;
        for (int j = 0; j < nb && !any; ++j) {
            if (typ[j] != typ[0]) continue;
            double t[3];
            for (int r = 0; r < 3; ++r) t[r] = frac[j * 3 + r] - rp[r];
            bool ok = true;
            for (int i = 0; i < nb && ok; ++i) {
                bool hit = false;
                for (int l = 0; l < nb && !hit; ++l) {
                    if (typ[l] != typ[i]) continue;
                    double d[3];
                    for (int r = 0; r < 3; ++r) {
                        d[r] = rp[i * 3 + r] + t[r] - frac[l * 3 + r];
                        d[r] -= std::nearbyint(d[r]);
                    }
                    double dist2 = 0.0;
                    for (int r = 0; r < 3; ++r) {
                        const double x = A[r][0] * d[0] + A[r][1] * d[1] + A[r][2] * d[2];
                        dist2 += x * x;
                    }
                    hit = std::sqrt(dist2) < symprec;
                }
                ok = hit;
            }
            any = ok;
        }
        if (any) {
            std::array<int, 9> r;
            for (int k = 0; k < 9; ++k) r[k] = e[k];
            out.push_back(r);
        }
    }
    return out;
}

}  // namespace

using namespace spr;

extern "C" int spr_mesh_build(const double *latvec, int nb, const double *basis_frac, const int *typ,
                              const int *mesh, int time_rev, int symmetry, double symprec, int *nq0_out,
                              int *nsym_out, int *q_int, double *q_cart, int *irr_grid, int *pos,
                              int *multiplicity, double *rot, int *tetra) {
    if (!latvec || nb < 1 || !basis_frac || !typ || !mesh || !nq0_out || !nsym_out || !(symprec > 0.0)) {
        set_error("spr_mesh_build: bad argument");
        return SPR_E_ARG;
    }
    for (int k = 0; k < 3; ++k)
        if (mesh[k] < 1 || mesh[k] > 1024) {
            set_error("spr_mesh_build: mesh dimensions must be in 1..1024");
            return SPR_E_ARG;
        }
    const int N0 = mesh[0], N1 = mesh[1], N2 = mesh[2];
    const long nql = (long)N0 * N1 * N2;
    if (nql > (1L << 27)) {
        set_error("spr_mesh_build: mesh too large");
        return SPR_E_ARG;
    }
    const int nq = (int)nql;
    double A[3][3], Ainv[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) A[i][j] = latvec[i * 3 + j];
    if (!inv3(A, Ainv)) {
        set_error("spr_mesh_build: singular lattice");
        return SPR_E_ARG;
    }
    // rotations
    std::vector<std::array<int, 9>> R;
    if (symmetry) {
        R = point_group(A, nb, basis_frac, typ, symprec);
        if (R.empty()) {
            set_error("spr_mesh_build: no point-group operation found (symprec too small?)");
            return SPR_E_ARG;
        }
    } else {
        R.push_back({1, 0, 0, 0, 1, 0, 0, 0, 1});
    }
    const int ng = (int)R.size();
    const int nsym = time_rev ? 2 * ng : ng;
    // orbit minimum of every grid point under k -> R^T k (and -R^T k)
    std::vector<int> rep(nq);
    for (int i = 0; i < nq; ++i) {
        const int k[3] = {i % N0, (i / N0) % N1, i / (N0 * N1)};
        int best = i;
        for (int g = 0; g < ng; ++g) {
            const int *r = R[g].data();
            int kk[3];
            for (int c = 0; c < 3; ++c) kk[c] = r[0 * 3 + c] * k[0] + r[1 * 3 + c] * k[1] + r[2 * 3 + c] * k[2];
            for (int sgn = 0; sgn < (time_rev ? 2 : 1); ++sgn) {
                int m[3];
                for (int c = 0; c < 3; ++c) {
                    const int v = sgn ? -kk[c] : kk[c];
                    m[c] = ((v % mesh[c]) + mesh[c]) % mesh[c];
                }
                best = std::min(best, m[0] + N0 * (m[1] + N1 * m[2]));
            }
        }
        rep[i] = best;
    }
    for (;;) {   // closure (a no-op for a group, kept for meshes the rotations do not preserve)
        bool changed = false;
        for (int i = 0; i < nq; ++i)
            if (rep[rep[i]] != rep[i]) {
                rep[i] = rep[rep[i]];
                changed = true;
            }
        if (!changed) break;
    }
    std::vector<int> irr;
    for (int i = 0; i < nq; ++i)
        if (rep[i] == i) irr.push_back(i);
    *nq0_out = (int)irr.size();
    *nsym_out = nsym;
    if (!q_int && !q_cart && !irr_grid && !pos && !multiplicity && !rot && !tetra) return SPR_OK;

    double B[3][3];   // B[i][j] = component i of b_j = 2 pi (A^-1)_{j i}
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) B[i][j] = 2.0 * M_PI * Ainv[j][i];
    std::vector<int> posv(nq);
    for (int i = 0; i < nq; ++i) posv[i] = (int)(std::lower_bound(irr.begin(), irr.end(), rep[i]) - irr.begin());
    for (int i = 0; i < nq; ++i) {
        const int a[3] = {fold1(i % N0, N0), fold1((i / N0) % N1, N1), fold1(i / (N0 * N1), N2)};
        if (q_int)
            for (int c = 0; c < 3; ++c) q_int[(size_t)i * 3 + c] = a[c];
        if (q_cart) {
            const double x[3] = {(double)a[0] / N0, (double)a[1] / N1, (double)a[2] / N2};
            for (int r = 0; r < 3; ++r) q_cart[(size_t)i * 3 + r] = B[r][0] * x[0] + B[r][1] * x[1] + B[r][2] * x[2];
        }
        if (pos) pos[i] = posv[i];
    }
    if (irr_grid)
        for (size_t k = 0; k < irr.size(); ++k) irr_grid[k] = irr[k];
    if (multiplicity) {
        for (size_t k = 0; k < irr.size(); ++k) multiplicity[k] = 0;
        for (int i = 0; i < nq; ++i) ++multiplicity[posv[i]];
    }
    if (rot) {
        // Cartesian A R A^-1, stored as Rxyz(3, 3, nsym) column-major: rot[k*9 + j*3 + i] = (i, j)
        for (int g = 0; g < ng; ++g) {
            double AR[3][3], C[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    AR[i][j] = A[i][0] * R[g][0 * 3 + j] + A[i][1] * R[g][1 * 3 + j] + A[i][2] * R[g][2 * 3 + j];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) C[i][j] = AR[i][0] * Ainv[0][j] + AR[i][1] * Ainv[1][j] + AR[i][2] * Ainv[2][j];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    rot[(size_t)g * 9 + j * 3 + i] = C[i][j];
                    if (time_rev) rot[(size_t)(g + ng) * 9 + j * 3 + i] = -C[i][j];
                }
        }
    }
    if (tetra) {
        // cube corners (0,0,0) (1,0,0) (0,0,1) (0,1,1) (1,0,1) (1,1,1) (1,1,0) (0,1,0) and the six
        // tetrahedra (1,5,2,4) (2,4,6,5) (8,4,6,5) (8,4,7,5) (7,4,3,5) (1,4,3,5) in corner numbers
        static const int box[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {0, 1, 1},
                                      {1, 0, 1}, {1, 1, 1}, {1, 1, 0}, {0, 1, 0}};
        static const int tets[6][4] = {{0, 1, 2, 3}, {2, 3, 4, 1}, {5, 3, 4, 1},
                                       {5, 3, 6, 1}, {6, 3, 7, 1}, {0, 3, 7, 1}};
        for (int i = 0; i < nq; ++i) {
            const int k[3] = {i % N0, (i / N0) % N1, i / (N0 * N1)};
            int corner[8];
            for (int v = 0; v < 8; ++v) {
                const int m0 = (k[0] + box[v][0]) % N0, m1 = (k[1] + box[v][1]) % N1, m2 = (k[2] + box[v][2]) % N2;
                corner[v] = m0 + N0 * (m1 + N1 * m2);
            }
            for (int t = 0; t < 6; ++t)
                for (int v = 0; v < 4; ++v) tetra[((size_t)i * 6 + t) * 4 + v] = corner[tets[t][v]];
        }
    }
    return SPR_OK;
}
