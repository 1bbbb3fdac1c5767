// RTA thermal conductivity in FP64 (kappa_noIter / kappa_noIter_symm,
// src/Kappa/Src/ThermalConductivity.f90:36-293; GrVelKronProdSymm,
// src/Kappa/Src/VelocityKronProd_symm.f90:34-93):
//   tau = 0.5 / lw(s, pos(q));  kappa_ab = kW_m_k / (T^2 V Nq) sum_{q,s} w^2 n (n+1) tau v_a v_b
// symm = 1 sums the irreducible q0 only, with (mult / Nsym) sum_R (R v)(R v)^T for v v^T.
// Contributions are written per (q, s) and reduced per mode in a fixed order (deterministic).
#include <algorithm>

#include "spr_internal.h"

namespace spr {

__global__ void k_kappa_terms(int nq_terms, int nd, const int *qgrid, const int *lwpos,
                              const double *omega, const double *nbe, const double *vel,
                              const double *lw_all, const double *rot, int nsym,
                              const int *mult, int s_lo, int s_hi, double *terms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq_terms * nd) return;
    const int t = i / nd, s = i % nd;
    const int q = qgrid ? qgrid[t] : t;
    double *out = terms + (size_t)i * 9;
    const bool gamma = q == 0;
    // the mode loop default_strt..default_end, from s = 4 (1-based) at Gamma
    // (ThermalConductivity.f90:106-111, 123-124)
    if ((gamma && s < 3) || s < s_lo || s > s_hi) {
        for (int k = 0; k < 9; ++k) out[k] = 0.0;
        return;
    }
    const double w = omega[(size_t)q * nd + s], n = nbe[(size_t)q * nd + s];
    const double tau = 0.5 / lw_all[(size_t)lwpos[t] * nd + s];
    const double c = w * w * n * (n + 1.0) * tau;
    const double *v = vel + ((size_t)q * nd + s) * 3;
    double k9[9];
    if (rot) {
        for (int k = 0; k < 9; ++k) k9[k] = 0.0;
        for (int r = 0; r < nsym; ++r) {
            const double *R = rot + 9 * r;   // Fortran Rxyz(3,3,r): R(a,b) at a + 3b
            double rv[3];
            for (int a = 0; a < 3; ++a) rv[a] = R[a] * v[0] + R[a + 3] * v[1] + R[a + 6] * v[2];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) k9[a * 3 + b] += rv[a] * rv[b];
        }
        const double p = 1.0 / (double)(nsym / mult[t]);
        for (int k = 0; k < 9; ++k) k9[k] *= p;
    } else {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) k9[a * 3 + b] = v[a] * v[b];
    }
    for (int k = 0; k < 9; ++k) out[k] = c * k9[k];
}

// mode_k[s][k] = f * sum_t terms[t][s][k]; one workgroup per (s, k): strided partial sums per
// thread, then a tree over the 256 partials (the order is fixed, so the result is reproducible)
__global__ __launch_bounds__(256) void k_kappa_reduce(int nq_terms, int nd, const double *terms, double f,
                                                      double *mode_k) {
    const int i = blockIdx.x;
    const int s = i / 9, k = i % 9;
    double acc = 0.0;
    for (int t = threadIdx.x; t < nq_terms; t += 256) acc += terms[((size_t)t * nd + s) * 9 + k];
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) mode_k[i] = red[0] * f;
}

}  // namespace spr

namespace spr {

// the kappa mode range: slot 1 of the three-phonon rule, widened by the four-phonon one when the
// context holds FC4 (kappa_noIter's FindStrtEnd, ThermalConductivity.f90:106-111)
void kappa_mode_range(const Ctx &c, int &lo, int &hi) {
    lo = c.pol3[0];
    hi = c.pol3[1];
    if (c.fc4.order == 4) {
        lo = std::min(lo, c.pol4[0]);
        hi = std::max(hi, c.pol4[1]);
    }
}

int launch_kappa(Ctx &c, const double *d_lw_all, double T, int symm, double *d_kappa_modes) {
    const int nd = c.nd;
    if (!c.vel) {
        set_error("kappa: grp_vel was not given to spr_ctx_create");
        return SPR_E_STATE;
    }
    if (symm && (c.nsym < 1 || c.h_mult.empty() || c.h_rot.empty())) {
        set_error("kappa symm: mesh rotations / multiplicities were not given");
        return SPR_E_STATE;
    }
    if (!symm && c.h_pos.empty()) {
        set_error("kappa: mesh pos() was not given");
        return SPR_E_STATE;
    }
    const int nt = symm ? c.nq0 : c.nq;
    int *&d_pos = symm ? c.k_pos_irr : c.k_pos_full;
    if (!c.k_terms) SPR_HIP(hipMalloc((void **)&c.k_terms, (size_t)c.nq * nd * 9 * sizeof(double)));
    if (!d_pos) {
        std::vector<int> lwpos(nt);
        for (int t = 0; t < nt; ++t) lwpos[t] = symm ? t : c.h_pos[t];
        SPR_HIP(hipMalloc((void **)&d_pos, nt * sizeof(int)));
        SPR_HIP(hipMemcpy(d_pos, lwpos.data(), nt * sizeof(int), hipMemcpyHostToDevice));
    }
    if (symm && !c.k_mult) {
        SPR_HIP(hipMalloc((void **)&c.k_mult, nt * sizeof(int)));
        SPR_HIP(hipMalloc((void **)&c.k_rot, c.h_rot.size() * sizeof(double)));
        SPR_HIP(hipMemcpy(c.k_mult, c.h_mult.data(), nt * sizeof(int), hipMemcpyHostToDevice));
        SPR_HIP(hipMemcpy(c.k_rot, c.h_rot.data(), c.h_rot.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    const int n = nt * nd;
    int s_lo, s_hi;
    kappa_mode_range(c, s_lo, s_hi);
    hipLaunchKernelGGL(k_kappa_terms, dim3((n + 255) / 256), dim3(256), 0, c.stream, nt, nd,
                       symm ? c.irr_grid : nullptr, d_pos, c.omega, c.nbe, c.vel, d_lw_all,
                       symm ? c.k_rot : nullptr, c.nsym, symm ? c.k_mult : nullptr, s_lo, s_hi, c.k_terms);
    const double kw = (1.054571817 * 1.054571817) * 0.001 / 1.380649;   // constants.f90:65
    const double f = kw / (T * T * c.vol) / (double)c.nq;
    hipLaunchKernelGGL(k_kappa_reduce, dim3(nd * 9), dim3(256), 0, c.stream, nt, nd, c.k_terms, f, d_kappa_modes);
    SPR_HIP(hipGetLastError());
    return SPR_OK;
}

}  // namespace spr
