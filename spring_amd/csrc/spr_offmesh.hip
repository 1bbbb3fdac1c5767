// spr_lw3_offmesh: the three-phonon linewidth of LineWidth.x, for a q0 that is not a mesh point
// (a point of the dispersion path).  ScatterProbAllq1 of src/LineWidth/Src/Linewidth_m.f90:48-231:
// for every grid q1, q2 = G - (q0 + q1) with the first G of {0, 1, -1, 2, -2}^3 (c1 fastest) that
// puts q2 inside the Brillouin zone (CheckIfWithinBZ, :458-497: |x| <= 1/2 for the zero-centred
// mesh, 0 <= x <= 1 otherwise, x the reduced coordinates); q2's phonons solved on the fly
// (get_phonon_singleq); W = |V3|^2 / (w0 w1 w2 Nq) Sctr3THz2 (ScatterProbability, :612-656); a q1
// without such a q2 keeps W = 0 and w2 = n2 = 0.  Then the tetrahedron sum over the q1 mesh with
// q2's frequencies and occupations taken per q1 (CalculateLwTetrahedron, :326-417, the path
// main.f90:293-379 runs) or the Gaussian sum (CalculateLinewidth, :234-323).
//
// Device work: the q2 eigenproblems in one batched solve (k_phonons), W of every q1 in one launch
// of the generic kernel's off-mesh mode, then the delta stage.  Phases of q2 need no G: for an
// integer cell n, e^{2 pi i x2 . n} = e^{-2 pi i x0 . n} conj(e^{2 pi i x1 . n}), and D(q2 + G) =
// D(q2) for this Bloch convention, so G only decides whether q2 exists.
#include <algorithm>
#include <cmath>
#include <vector>

#include "spr_internal.h"

namespace spr {

constexpr int kOffBlk = 256;

// Gaussian sum of one q0 from W of every q1 (CalculateLinewidth of Linewidth_m.f90:234-323):
// lw(s0) = sum_{q1, s1, s2} W [(n1 + n2 + 1) d(w0 - w1 - w2) + (n1 - n2)(d(w0 + w1 - w2) - d(w0 - w1 + w2))]
// with q2's w2, n2 stored per q1.  One workgroup per s0, strided per-thread sums, fixed-order tree.
__global__ __launch_bounds__(kOffBlk) void k_lw3_off_gauss(const double *W, const double *omega, const double *nbe,
                                                           const double *omega2, const double *nbe2,
                                                           const double *omega0, int nq, int nd, double sigma,
                                                           double *lw) {
    const int s0 = blockIdx.x, tid = threadIdx.x;
    const int nd2 = nd * nd;
    const double c0 = 1.0 / (sigma * 2.5066282746310002), hinv = 0.5 / (sigma * sigma);
    const double w0 = omega0[s0];
    double acc = 0.0;
    for (long i = tid; i < (long)nq * nd2; i += kOffBlk) {
        const int q1 = (int)(i / nd2), p = (int)(i % nd2), s1 = p / nd, s2 = p % nd;
        const double w = W[((size_t)q1 * nd + s0) * nd2 + p];
        if (w == 0.0) continue;
        const double w1 = omega[(size_t)q1 * nd + s1], n1 = nbe[(size_t)q1 * nd + s1];
        const double w2 = omega2[(size_t)q1 * nd + s2], n2 = nbe2[(size_t)q1 * nd + s2];
        const double d1 = w0 - w1 - w2, d2 = w0 + w1 - w2, d3 = w0 - w1 + w2;
        acc += w * ((n1 + n2 + 1.0) * (c0 * exp(-hinv * d1 * d1)) +
                    (n1 - n2) * (c0 * exp(-hinv * d2 * d2) - c0 * exp(-hinv * d3 * d3)));
    }
    __shared__ double red[kOffBlk];
    red[tid] = acc;
    __syncthreads();
    for (int w = kOffBlk / 2; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) lw[s0] = red[0];
}

}  // namespace spr

using namespace spr;

extern "C" int spr_lw3_offmesh(spr_ctx *h, const double *q0_cart, const double *evec0, const double *omega0,
                               const double *inv_omega0, const spr_crystal *cr, const double *latvec,
                               const spr_ifc *fc2, double T, int zero_centered, int delta, double sigma, double *lw,
                               double *W_out) {
    if (!h || !q0_cart || !evec0 || !omega0 || !inv_omega0 || !cr || !latvec || !fc2 || !lw || !(T > 0.0) ||
        (delta != SPR_DELTA_GAUSS && delta != SPR_DELTA_TETRA) || (delta == SPR_DELTA_GAUSS && !(sigma > 0.0))) {
        set_error("spr_lw3_offmesh: bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    if (cr->nb != c.nb) {
        set_error("spr_lw3_offmesh: the crystal's basis differs from the context's");
        return SPR_E_ARG;
    }
    const int nd = c.nd, nq = c.nq;
    const int N[3] = {c.mesh[0], c.mesh[1], c.mesh[2]};
    // reduced coordinates x = A^T q / (2 pi) (A's columns the lattice vectors, latvec[i*3+j] =
    // component i of vector j); reciprocal vectors b_j = 2 pi (A^-T)_{:, j}
    const double tpi = 2.0 * M_PI;
    double x0[3], B[9];
    for (int j = 0; j < 3; ++j) x0[j] = (latvec[0 * 3 + j] * q0_cart[0] + latvec[1 * 3 + j] * q0_cart[1] +
                                         latvec[2 * 3 + j] * q0_cart[2]) / tpi;
    {
        const double *a = latvec;
        const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                           a[2] * (a[3] * a[7] - a[4] * a[6]);
        if (!(std::fabs(det) > 0.0)) {
            set_error("spr_lw3_offmesh: singular lattice");
            return SPR_E_ARG;
        }
        // inverse of A (row-major), then B = 2 pi inv(A)^T
        double inv[9];
        inv[0] = (a[4] * a[8] - a[5] * a[7]) / det;
        inv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
        inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
        inv[3] = (a[5] * a[6] - a[3] * a[8]) / det;
        inv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
        inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
        inv[6] = (a[3] * a[7] - a[4] * a[6]) / det;
        inv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
        inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) B[i * 3 + j] = tpi * inv[j * 3 + i];   // B[i][j] = component i of b_j
    }
    // q2 of every grid q1 (G search in the reference's order) and its Cartesian coordinates
    static const int sgn[5] = {0, 1, -1, 2, -2};
    std::vector<double> q2c((size_t)nq * 3, 0.0);
    std::vector<int> exists(nq, 0), is_gamma(nq, 0);
    for (int q1 = 0; q1 < nq; ++q1) {
        const int k[3] = {q1 % N[0], (q1 / N[0]) % N[1], q1 / (N[0] * N[1])};
        double x1[3];
        for (int d = 0; d < 3; ++d) {
            int kk = k[d];
            if (zero_centered && 2 * kk > N[d]) kk -= N[d];   // fold_indx to (-N/2, N/2]
            x1[d] = (double)kk / N[d];
        }
        double x2[3] = {0.0, 0.0, 0.0};
        bool found = false;
        for (int c3 = 0; c3 < 5 && !found; ++c3)
            for (int c2 = 0; c2 < 5 && !found; ++c2)
                for (int c1 = 0; c1 < 5 && !found; ++c1) {
                    const int g[3] = {sgn[c1], sgn[c2], sgn[c3]};
                    bool in = true;
                    for (int d = 0; d < 3; ++d) {
                        x2[d] = g[d] - x0[d] - x1[d];
                        in = in && (zero_centered ? std::fabs(x2[d]) <= 0.5 : (x2[d] >= 0.0 && x2[d] <= 1.0));
                    }
                    found = in;
                }
        if (!found)
            for (int d = 0; d < 3; ++d) x2[d] = -x0[d] - x1[d];   // phonons computed but W stays 0
        exists[q1] = found ? 1 : 0;
        double nrm = 0.0;
        for (int i = 0; i < 3; ++i) {
            q2c[(size_t)q1 * 3 + i] = B[i * 3 + 0] * x2[0] + B[i * 3 + 1] * x2[1] + B[i * 3 + 2] * x2[2];
            nrm += q2c[(size_t)q1 * 3 + i] * q2c[(size_t)q1 * 3 + i];
        }
        is_gamma[q1] = nrm < 1e-20 ? 1 : 0;
    }
    // q2 phonons (one batched device solve; the host arrays are its output)
    const size_t ne = (size_t)nq * nd * nd, nm = (size_t)nq * nd;
    std::vector<double> e2(2 * ne), w2(nm), iw2(nm), n2(nm), v2(nm * 3);
    int rc = spr_get_phonon(cr, latvec, fc2, nq, q2c.data(), is_gamma.data(), T, c.device, e2.data(), w2.data(),
                            iw2.data(), n2.data(), v2.data());
    if (rc) return rc;
    for (int q1 = 0; q1 < nq; ++q1)
        if (!exists[q1])
            for (int s = 0; s < nd; ++s) w2[(size_t)q1 * nd + s] = n2[(size_t)q1 * nd + s] = 0.0;
    // e^{-2 pi i x0 . n} of every FC3 site
    const int ns = c.fc3.nsite;
    std::vector<double> p0(2 * (size_t)ns);
    for (int s = 0; s < ns; ++s) {
        const int4 n = c.fc3.h_site_cell[s];
        double t = x0[0] * n.x + x0[1] * n.y + x0[2] * n.z;
        t -= std::floor(t);
        p0[2 * s] = std::cos(tpi * t);
        p0[2 * s + 1] = -std::sin(tpi * t);
    }
    // device buffers: one allocation for the inputs, the W rows in c.wout
    SPR_HIP(hipSetDevice(c.device));
    const size_t n_e0 = (size_t)nd * nd * 2, n_p0 = 2 * (size_t)ns;
    const size_t ndbl = n_e0 + 2 * ne + n_p0 + nd /*iw0*/ + nm /*iw2*/ + 2 * nm /*omega ext*/ + 2 * nm /*nbe ext*/ +
                        nd /*omega0*/ + 2 * (size_t)nd /*lw, p3*/;
    const size_t nbytes = ndbl * sizeof(double) + 2 * (size_t)nq * sizeof(int);
    if ((rc = ensure((void **)&c.off_buf, &c.off_cap, nbytes))) return rc;
    if ((rc = ensure((void **)&c.wout, &c.wout_cap, (size_t)nq * nd * nd * nd * sizeof(double)))) return rc;
    double *p = c.off_buf;
    double *d_e0 = p; p += n_e0;
    double *d_e2 = p; p += 2 * ne;
    double *d_p0 = p; p += n_p0;
    double *d_iw0 = p; p += nd;
    double *d_iw2 = p; p += nm;
    double *d_wext = p; p += 2 * nm;
    double *d_next = p; p += 2 * nm;
    double *d_w0 = p; p += nd;
    double *d_lw = p; p += 2 * nd;
    int *d_exists = reinterpret_cast<int *>(p);
    int *d_partner = d_exists + nq;
    std::vector<int> partner(nq);
    for (int q1 = 0; q1 < nq; ++q1) partner[q1] = nq + q1;   // q2 data of q1 follows the mesh's
    const hipStream_t st = c.stream;
    SPR_HIP(hipMemcpyAsync(d_e0, evec0, n_e0 * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_e2, e2.data(), 2 * ne * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_p0, p0.data(), n_p0 * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_iw0, inv_omega0, nd * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_iw2, iw2.data(), nm * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_wext, c.omega, nm * sizeof(double), hipMemcpyDeviceToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_wext + nm, w2.data(), nm * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_next, c.nbe, nm * sizeof(double), hipMemcpyDeviceToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_next + nm, n2.data(), nm * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_w0, omega0, nd * sizeof(double), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_exists, exists.data(), nq * sizeof(int), hipMemcpyHostToDevice, st));
    SPR_HIP(hipMemcpyAsync(d_partner, partner.data(), nq * sizeof(int), hipMemcpyHostToDevice, st));
    OffMeshDev om{reinterpret_cast<const double2 *>(d_e0), reinterpret_cast<const double2 *>(d_e2),
                  reinterpret_cast<const double2 *>(d_p0), d_iw0, d_iw2, d_exists};
    if ((rc = launch_w_offmesh(c, om, c.wout))) return rc;
    if (delta == SPR_DELTA_TETRA) {
        if ((rc = tetra_sums_from_w(c, c.wout, d_partner, d_wext, d_next, d_w0, d_lw, d_lw + nd))) return rc;
    } else {
        hipLaunchKernelGGL(k_lw3_off_gauss, dim3(nd), dim3(kOffBlk), 0, st, c.wout, c.omega, c.nbe, d_wext + nm,
                           d_next + nm, d_w0, nq, nd, sigma, d_lw);
        SPR_HIP(hipGetLastError());
    }
    SPR_HIP(hipMemcpyAsync(lw, d_lw, nd * sizeof(double), hipMemcpyDeviceToHost, st));
    if (W_out)
        SPR_HIP(hipMemcpyAsync(W_out, c.wout, (size_t)nq * nd * nd * nd * sizeof(double), hipMemcpyDeviceToHost, st));
    SPR_HIP(hipStreamSynchronize(st));
    return SPR_OK;
}
