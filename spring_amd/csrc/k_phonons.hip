// Harmonic phonons on the device (get_phonon / DynMatgv_SR, src/Kappa/Src/DynaMat_mod.f90:
// 98-190, 486-709; phonon_mod.f90:156-315), non-polar crystals (no Ewald term):
//   D(q)[(mu,a),(nu,b)] = (sum_{N of mu, basis(N) = nu} FC2(mu, N, a, b) e^{i q.R_N}) / sqrt(m_mu m_nu)
//   dD/dq_k              = the same with i R_N,k in every term
//   lambda, e = eig(D) from the upper triangle of D (zheevd 'V','U'), ascending
//   w = sqrt|lambda| (negative lambda beyond EPS gives -w), omega = w wTHz, 1/omega,
//   n = 1/(exp(BEConv |w| / T) - 1), v = Re <e| dD/dq |e> / (2 w) vg_ms;
//   at Gamma the three lowest modes get v = n = 1/omega = 0.
//
// One wave per q point.  D and the eigenvectors stay in LDS (2 Ndof^2 x 16 bytes); the
// eigensolve is a cyclic two-sided Jacobi method for Hermitian matrices: each step applies
// Ndof/2 rotations on disjoint index pairs at once (round-robin pairing, Ndof - 1 steps per
// sweep); each rotation is the phase diag(1, e^{-i phi}) that makes the (p, q) element real
// followed by the real Jacobi rotation that zeroes it (t = sign(z)/(|z| + sqrt(1+z^2)),
// z = (a_qq - a_pp) / (2 |a_pq|), c = 1/sqrt(1+t^2), s = t c; Golub & Van Loan, sec. 8.5).
// Sweeps run until the off-diagonal mass is below 1e-32 of the diagonal one.  The velocity
// <e|dD/dq|e> is summed term by term (3x3 blocks) instead of forming dD/dq.  FP64 throughout.
#include <algorithm>
#include <cmath>
#include <vector>

#include "spr_internal.h"

namespace spr {

namespace {

constexpr int kPhMaxNd = 32;
constexpr int kPhMaxSweeps = 40;

struct PhArgs {
    int nb, nd, nq;
    const double *blk;     // (nterm, 3, 3) FC2 block of each term, sorted by (mu, nu), eV/A^2
    const double *mfac;    // (nb*nb): 1/sqrt(m_mu m_nu)
    const double *R;       // (nterm, 3) Cartesian lattice vector of the term, Angstrom
    const int *pstart;     // (nb*nb + 1): first term of each (mu, nu) pair
    const double *q;       // (nq, 3) 1/Angstrom
    const int *gamma;      // (nq) 1 at Gamma
    double T, eps, wthz, vgms, beconv;
    double2 *evec;         // (nq, nd, nd): evec[q][s][row]
    double *omega, *inv_omega, *nbe, *vel;
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(64) void k_phonons(PhArgs A) {
    const int qi = blockIdx.x, l = threadIdx.x;
    const int nb = A.nb, nd = A.nd, nn = nd * nd;
    const int np = (nd + 1) & ~1;   // padded to even for the round-robin pairing
    extern __shared__ double2 sDyn[];   // D then V, nd*nd each
    double2 *sD = sDyn, *sV = sDyn + nn;
    __shared__ double sC[kPhMaxNd / 2], sS[kPhMaxNd / 2];
    __shared__ double2 sP[kPhMaxNd / 2];   // e^{-i phi} of each pair
    __shared__ int sPa[kPhMaxNd / 2], sPb[kPhMaxNd / 2];
    __shared__ double sLam[kPhMaxNd];
    __shared__ int sOrd[kPhMaxNd];
    __shared__ double sRed[2][64];
    const double q0 = A.q[3 * qi], q1 = A.q[3 * qi + 1], q2 = A.q[3 * qi + 2];

    // ---- upper triangle of D, mirrored (zheevd reads only 'U'; the diagonal's imaginary part
    // is ignored) ----
    for (int e = l; e < nn; e += 64) {
        const int r = e / nd, c = e % nd;
        if (r <= c) {
            const int mu = r / 3, a = r % 3, nu = c / 3, b = c % 3;
            double2 d = make_double2(0.0, 0.0);
            for (int t = A.pstart[mu * nb + nu]; t < A.pstart[mu * nb + nu + 1]; ++t) {
                const double f = A.blk[t * 9 + a * 3 + b];
                const double *R = A.R + 3 * t;
                double sn, cs;
                sincos(q0 * R[0] + q1 * R[1] + q2 * R[2], &sn, &cs);
                d.x = fma(f, cs, d.x);
                d.y = fma(f, sn, d.y);
            }
            const double m = A.mfac[mu * nb + nu];
            d.x *= m;
            d.y = r == c ? 0.0 : d.y * m;
            sD[r * nd + c] = d;
            sD[c * nd + r] = make_double2(d.x, -d.y);
        }
        sV[e] = make_double2(r == c ? 1.0 : 0.0, 0.0);
    }
    __syncthreads();

    // ---- cyclic parallel Jacobi ----
    for (int sweep = 0; sweep < kPhMaxSweeps; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int e = l; e < nn; e += 64) {
            const double2 a = sD[e];
            const double m = a.x * a.x + a.y * a.y;
            if (e / nd == e % nd) dia += m; else off += m;
        }
        sRed[0][l] = off;
        sRed[1][l] = dia;
        __syncthreads();
        for (int w = 32; w > 0; w >>= 1) {
            if (l < w) {
                sRed[0][l] += sRed[0][l + w];
                sRed[1][l] += sRed[1][l + w];
            }
            __syncthreads();
        }
        const bool done = !(sRed[0][0] > 1e-32 * sRed[1][0]);
        __syncthreads();
        if (done) break;
        for (int step = 0; step < np - 1; ++step) {
            // round robin: player 0 fixed, the others rotate by `step`
            if (l < np / 2) {
                const int k2 = np - 1 - l;
                int a = l == 0 ? 0 : 1 + (l - 1 + step) % (np - 1);
                int b = 1 + (k2 - 1 + step) % (np - 1);
                if (a > b) {
                    const int t = a;
                    a = b;
                    b = t;
                }
                double c = 1.0, s = 0.0;
                double2 ph = make_double2(1.0, 0.0);
                if (b < nd) {
                    const double2 apq = sD[a * nd + b];
                    const double r = sqrt(apq.x * apq.x + apq.y * apq.y);
                    if (r > 0.0) {
                        const double z = (sD[b * nd + b].x - sD[a * nd + a].x) / (2.0 * r);
                        const double t = (z >= 0.0 ? 1.0 : -1.0) / (fabs(z) + sqrt(1.0 + z * z));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                        ph = make_double2(apq.x / r, -apq.y / r);   // e^{-i phi}
                    }
                }
                sPa[l] = a;
                sPb[l] = b;
                sC[l] = c;
                sS[l] = s;
                sP[l] = ph;
            }
            __syncthreads();
            // columns of A and V times U: U_pp = c, U_pq = s, U_qp = -s e^{-i phi}, U_qq = c e^{-i phi}
            for (int i = l; i < (np / 2) * nd; i += 64) {
                const int k = i / nd, row = i % nd;
                const int a = sPa[k], b = sPb[k];
                if (b >= nd) continue;
                const double c = sC[k], s = sS[k];
                const double2 ph = sP[k];
                const double2 xa = sD[row * nd + a], xb = cmul(ph, sD[row * nd + b]);
                sD[row * nd + a] = make_double2(c * xa.x - s * xb.x, c * xa.y - s * xb.y);
                sD[row * nd + b] = make_double2(s * xa.x + c * xb.x, s * xa.y + c * xb.y);
                const double2 va = sV[row * nd + a], vb = cmul(ph, sV[row * nd + b]);
                sV[row * nd + a] = make_double2(c * va.x - s * vb.x, c * va.y - s * vb.y);
                sV[row * nd + b] = make_double2(s * va.x + c * vb.x, s * va.y + c * vb.y);
            }
            __syncthreads();
            // rows: U^H (A U); U^H_pp = c, U^H_pq = -s e^{i phi}, U^H_qp = s, U^H_qq = c e^{i phi}
            for (int i = l; i < (np / 2) * nd; i += 64) {
                const int k = i / nd, col = i % nd;
                const int a = sPa[k], b = sPb[k];
                if (b >= nd) continue;
                const double c = sC[k], s = sS[k];
                const double2 ph = make_double2(sP[k].x, -sP[k].y);
                const double2 xa = sD[a * nd + col], xb = cmul(ph, sD[b * nd + col]);
                // the rotated pair element is zero by construction: store it exactly
                sD[a * nd + col] = col == b ? make_double2(0.0, 0.0)
                                            : make_double2(c * xa.x - s * xb.x, c * xa.y - s * xb.y);
                sD[b * nd + col] = col == a ? make_double2(0.0, 0.0)
                                            : make_double2(s * xa.x + c * xb.x, s * xa.y + c * xb.y);
            }
            __syncthreads();
        }
    }

    // ---- ascending order of the eigenvalues (ties keep the column order) ----
    if (l < nd) sLam[l] = sD[l * nd + l].x;
    __syncthreads();
    if (l < nd) {
        int rank = 0;
        for (int j = 0; j < nd; ++j) rank += sLam[j] < sLam[l] || (sLam[j] == sLam[l] && j < l);
        sOrd[rank] = l;
    }
    __syncthreads();

    // ---- outputs ----
    for (int i = l; i < nn; i += 64) {
        const int s = i / nd, row = i % nd;
        A.evec[(size_t)qi * nn + i] = sV[row * nd + sOrd[s]];
    }
    if (l < nd) {
        const int s = l, col = sOrd[s];
        const double lam = sLam[col];
        double w = sqrt(fabs(lam));
        if (fabs(lam) > A.eps && lam < 0.0) w = -w;
        // <e| dD/dq_k |e> = sum_terms i R_k e^{i q.R} z_t,  z_t = sum_ab conj(e_{mu a}) f_ab e_{nu b}
        double g[3] = {0.0, 0.0, 0.0};
        for (int mu = 0; mu < nb; ++mu)
            for (int nu = 0; nu < nb; ++nu) {
                double gp[3] = {0.0, 0.0, 0.0};
                for (int t = A.pstart[mu * nb + nu]; t < A.pstart[mu * nb + nu + 1]; ++t) {
                    const double *f = A.blk + t * 9;
                    double2 z = make_double2(0.0, 0.0);
                    for (int a = 0; a < 3; ++a) {
                        double2 fe = make_double2(0.0, 0.0);
                        for (int b = 0; b < 3; ++b) {
                            const double2 e = sV[(3 * nu + b) * nd + col];
                            fe.x = fma(f[a * 3 + b], e.x, fe.x);
                            fe.y = fma(f[a * 3 + b], e.y, fe.y);
                        }
                        const double2 ea = sV[(3 * mu + a) * nd + col];
                        z.x += ea.x * fe.x + ea.y * fe.y;   // conj(e) fe
                        z.y += ea.x * fe.y - ea.y * fe.x;
                    }
                    const double *R = A.R + 3 * t;
                    double sn, cs;
                    sincos(q0 * R[0] + q1 * R[1] + q2 * R[2], &sn, &cs);
                    // Re(i e^{i qR} z) = -(sin z.x + cos z.y)
                    const double re = -(sn * z.x + cs * z.y);
                    for (int k = 0; k < 3; ++k) gp[k] = fma(R[k], re, gp[k]);
                }
                for (int k = 0; k < 3; ++k) g[k] = fma(gp[k], A.mfac[mu * nb + nu], g[k]);
            }
        double v[3];
        for (int k = 0; k < 3; ++k) v[k] = g[k] / (2.0 * w) * A.vgms;
        double n = 1.0 / (exp(A.beconv * fabs(w) / A.T) - 1.0);
        double iw = 1.0 / (w * A.wthz);
        if (A.gamma[qi] != 0 && s < 3) {
            v[0] = v[1] = v[2] = 0.0;
            n = 0.0;
            iw = 0.0;
        }
        const size_t o = (size_t)qi * nd + s;
        A.omega[o] = w * A.wthz;
        A.inv_omega[o] = iw;
        A.nbe[o] = n;
        for (int k = 0; k < 3; ++k) A.vel[o * 3 + k] = v[k];
    }
}

}  // namespace

}  // namespace spr

// Phonons of nq q points on `device`; declared in spring_gpu.h.
extern "C" int spr_get_phonon(const spr_crystal *cr, const double *latvec, const spr_ifc *fc2, int nq,
                           const double *q_cart, const int *is_gamma, double T, int device, double *evec,
                           double *omega, double *inv_omega, double *nbe, double *grp_vel) {
    using namespace spr;
    if (!cr || !latvec || !fc2 || !q_cart || !is_gamma || !evec || !omega || !inv_omega || !nbe || !grp_vel ||
        nq < 1 || !(T > 0.0)) {
        set_error("spr_get_phonon: bad argument");
        return SPR_E_ARG;
    }
    const int nb = cr->nb, nd = 3 * nb;
    if (nb < 1 || nd > kPhMaxNd || fc2->order != 2 || fc2->natm < 1 || !fc2->atm_num || !fc2->atm_indx ||
        !fc2->fc || !cr->mass) {
        set_error("spr_get_phonon: needs an order-2 IFC and 1 <= Ndof <= 32");
        return SPR_E_ARG;
    }
    // terms grouped by (mu, nu), each a 3x3 block; one mass factor per (mu, nu) (DynMatgv_SR)
    std::vector<std::vector<int>> pair_terms((size_t)nb * nb);
    for (int mu = 0; mu < nb; ++mu) {
        if (fc2->atm_num[mu] < 0 || fc2->atm_num[mu] > fc2->natm) {
            set_error("spr_get_phonon: atm_num out of range");
            return SPR_E_ARG;
        }
        for (int n = 0; n < fc2->atm_num[mu]; ++n) {
            const int nu = fc2->atm_indx[((size_t)mu * fc2->natm + n) * 4 + 3];
            if (nu < 0 || nu >= nb) {
                set_error("spr_get_phonon: basis index in atm_indx out of range (expects 0-based)");
                return SPR_E_ARG;
            }
            pair_terms[(size_t)mu * nb + nu].push_back(mu * fc2->natm + n);
        }
    }
    std::vector<double> blk, R, mfac;
    std::vector<int> pstart(1, 0);
    for (int mu = 0; mu < nb; ++mu)
        for (int nu = 0; nu < nb; ++nu) {
            mfac.push_back(1.0 / std::sqrt(cr->mass[mu] * cr->mass[nu]));
            for (int id : pair_terms[(size_t)mu * nb + nu]) {
                const double *f = fc2->fc + (size_t)id * 9;   // fC(1, b, a, N, mu)
                for (int ab = 0; ab < 9; ++ab) blk.push_back(f[ab]);
                const int *ix = fc2->atm_indx + (size_t)id * 4;
                for (int i = 0; i < 3; ++i)
                    R.push_back(latvec[i * 3 + 0] * ix[0] + latvec[i * 3 + 1] * ix[1] + latvec[i * 3 + 2] * ix[2]);
            }
            pstart.push_back((int)(R.size() / 3));
        }
    SPR_HIP(hipSetDevice(device));
    double *d_blk = nullptr, *d_R = nullptr, *d_q = nullptr, *d_out = nullptr, *d_mf = nullptr;
    int *d_g = nullptr, *d_ps = nullptr;
    double2 *d_evec = nullptr;
    const size_t nm = (size_t)nq * nd;
    auto cleanup = [&]() {
        (void)hipFree(d_blk);
        (void)hipFree(d_mf);
        (void)hipFree(d_R);
        (void)hipFree(d_ps);
        (void)hipFree(d_q);
        (void)hipFree(d_out);
        (void)hipFree(d_g);
        (void)hipFree(d_evec);
    };
    if (hipMalloc((void **)&d_blk, std::max<size_t>(1, blk.size()) * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_R, std::max<size_t>(1, R.size()) * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_ps, pstart.size() * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_mf, mfac.size() * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_q, (size_t)nq * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_g, (size_t)nq * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_out, nm * 6 * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_evec, nm * nd * sizeof(double2)) != hipSuccess) {
        cleanup();
        set_error("spr_get_phonon: hipMalloc failed");
        return SPR_E_OOM;
    }
    int rc = SPR_OK;
    if ((!blk.empty() && hipMemcpy(d_blk, blk.data(), blk.size() * sizeof(double), hipMemcpyHostToDevice)) ||
        (!R.empty() && hipMemcpy(d_R, R.data(), R.size() * sizeof(double), hipMemcpyHostToDevice)) ||
        hipMemcpy(d_ps, pstart.data(), pstart.size() * sizeof(int), hipMemcpyHostToDevice) ||
        hipMemcpy(d_mf, mfac.data(), mfac.size() * sizeof(double), hipMemcpyHostToDevice) ||
        hipMemcpy(d_q, q_cart, (size_t)nq * 3 * sizeof(double), hipMemcpyHostToDevice) ||
        hipMemcpy(d_g, is_gamma, (size_t)nq * sizeof(int), hipMemcpyHostToDevice)) {
        set_error("spr_get_phonon: upload failed");
        rc = SPR_E_HIP;
    }
    if (!rc) {
        PhArgs A;
        A.nb = nb;
        A.nd = nd;
        A.nq = nq;
        A.blk = d_blk;
        A.mfac = d_mf;
        A.R = d_R;
        A.pstart = d_ps;
        A.q = d_q;
        A.gamma = d_g;
        A.T = T;
        // constants.f90:40-55: EPS is the single-precision literal 1.0E-8 widened to double
        A.eps = 9.99999993922529e-09;
        A.wthz = std::sqrt(1.60217733 * 6.02214076 * 10.0) * 10.0;
        A.vgms = A.wthz * 100.0;
        A.beconv = std::sqrt(1.60217733 * 6.02214076 * 10.0) * 6.582119569 * 100.0 / 8.617333262;
        A.evec = d_evec;
        A.omega = d_out;
        A.inv_omega = d_out + nm;
        A.nbe = d_out + 2 * nm;
        A.vel = d_out + 3 * nm;
        hipLaunchKernelGGL(k_phonons, dim3(nq), dim3(64), 2 * (size_t)nd * nd * sizeof(double2), 0, A);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
            set_error("spr_get_phonon: kernel failed");
            rc = SPR_E_HIP;
        }
    }
    if (!rc && (hipMemcpy(evec, d_evec, nm * nd * sizeof(double2), hipMemcpyDeviceToHost) ||
                hipMemcpy(omega, d_out, nm * sizeof(double), hipMemcpyDeviceToHost) ||
                hipMemcpy(inv_omega, d_out + nm, nm * sizeof(double), hipMemcpyDeviceToHost) ||
                hipMemcpy(nbe, d_out + 2 * nm, nm * sizeof(double), hipMemcpyDeviceToHost) ||
                hipMemcpy(grp_vel, d_out + 3 * nm, nm * 3 * sizeof(double), hipMemcpyDeviceToHost))) {
        set_error("spr_get_phonon: download failed");
        rc = SPR_E_HIP;
    }
    cleanup();
    return rc;
}
