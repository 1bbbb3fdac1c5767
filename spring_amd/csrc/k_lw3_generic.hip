// Three-phonon scattering: generic fused kernel (any Ndof = 3..30), FP64 or FP32.
//
// One workgroup takes one irreducible q0, one chunk of q1 grid points and one block of
// s0 modes.  For every pair (q1, q2 = -(q0+q1) mod G) kept by the pair-permutation rule
// (q1 <= q2, weight 2 if q1 < q2) it
//   1. forms the phase sum T_a(b, c) = sum_{N1,N2} Phi(a,b,c;N1,N2) e^{i(q1 R1 + q2 R2)} / sqrt(m m m)
//      for one row a = (mu, alpha) at a time                      (ScatterMatEl, ThirdOrder_m.f90:316-384)
//   2. contracts it with the q2 and q1 eigenvectors: Y_a(s1, s2) = e1^T T_a e2
//   3. accumulates V(s0, s1, s2) += e0(a, s0) Y_a(s1, s2) in registers (thread owns (s1, s2))
//   4. forms W = |V|^2 /(w0 w1 w2) Sctr3THz2/Nq                       (ScatterProbability :389-439)
//      and adds w_pair W [(n1+n2+1) d(w0-w1-w2) + (n1-n2)(d(w0+w1-w2) - d(w0-w1+w2))]
//      to the thread's Gamma(s0) partial                               (CalculateLinewidth :442-543)
// Partials of each workgroup are reduced in a fixed order by k_reduce_partials, so results
// are bitwise reproducible.  With W_OUT the kernel writes W(s2,s1,s0,q1) for every q1 instead
// (the parity path of ScatterProbAllq1's Sctrq1 array).
//
// Phases use integer addresses: q.R = 2 pi sum_d k_d n_d / N_d, looked up in a table of
// exp(2 pi i m / L), L = lcm(N_d).
#include "cplx.h"
#include "spr_internal.h"

namespace spr {

struct Lw3Args {
    const int *q0_grid;
    int nq, L, m0, m1, m2;
    int chunk, nchunks;
    const int *blk_start;
    const int *tup_site;
    const void *fc;
    const int4 *site;
    int nsite;
    const void *phase;
    const void *evec;
    const double *omega, *inv_omega, *nbe;
    double sigma, factor;
    double *partial;   // [q0][s0b][chunk][S0B]
    double *w_out;     // [q1][s0][s1][s2]  (pair mode: [pair][s0][s1][s2])
    const int *pair_q0;   // pair mode (W_OUT only): block x = pair, q0 index and q1 per pair
    const int *pair_q1;
    const double *sel;    // branch selection (NULL: all): (nd) s0 mask, (nd, nd) pair mask [s1][s2]
    // off-mesh q0 (OFF instantiation, W output of one q0): its eigenvectors and 1/omega, the
    // eigenvectors and 1/omega of q2 = G - (q0 + q1) per grid q1, e^{-2 pi i x0 . n} per site
    // (x0 the fractional q0, n the site's cell), and whether q2 was found for each q1
    const void *off_e0, *off_e2, *off_p0;
    const double *off_iw0, *off_iw2;
    const int *off_exists;
};

__device__ __forceinline__ void grid_coords(int i, int m0, int m1, int &k0, int &k1, int &k2) {
    k0 = i % m0;
    k1 = (i / m0) % m1;
    k2 = i / (m0 * m1);
}

__device__ __forceinline__ double gauss_g(double w0, double w1, double w2, double n1, double n2,
                                          double c0, double hinv) {
    const double d1 = w0 - w1 - w2, d2 = w0 + w1 - w2, d3 = w0 - w1 + w2;
    const double g1 = c0 * exp(-hinv * d1 * d1);
    const double g2 = c0 * exp(-hinv * d2 * d2);
    const double g3 = c0 * exp(-hinv * d3 * d3);
    return (n1 + n2 + 1.0) * g1 + (n1 - n2) * (g2 - g3);
}

template <typename R, int ND, int S0B, int BLK, bool W_OUT, bool OFF = false>
__global__ __launch_bounds__(BLK) void k_lw3_generic(Lw3Args A) {
    static_assert(!OFF || W_OUT, "the off-mesh mode writes W");
    using Z = cplx_t<R>;
    constexpr int NB = ND / 3;
    constexpr int ND2 = ND * ND;
    constexpr int NS0B = (ND + S0B - 1) / S0B;
    __shared__ Z s_e0[ND2], s_e1[ND2], s_e2[ND2], s_T[ND2], s_Z[ND2];
    extern __shared__ __align__(16) int s_dyn[];   // m1[nsite], m2[nsite], phase[L]; OFF: + site phases
    int *s_m1 = s_dyn;
    int *s_m2 = s_dyn + A.nsite;
    Z *s_ph = reinterpret_cast<Z *>(s_dyn + 2 * ((A.nsite + 1) & ~1));
    Z *s_pa = s_ph + A.L;          // OFF: phase of q1 at each site
    Z *s_pb = s_pa + A.nsite;      // OFF: phase of q2 at each site

    const int tid = threadIdx.x;
    const int ch = blockIdx.x, s0b = blockIdx.y, qb = blockIdx.z;
    const bool pair_mode = W_OUT && A.pair_q0 != nullptr;
    const int i0 = OFF ? 0 : A.q0_grid[pair_mode ? A.pair_q0[ch] : qb];
    int k00, k01, k02;
    grid_coords(i0, A.m0, A.m1, k00, k01, k02);
    const Z *evec = reinterpret_cast<const Z *>(A.evec);
    const Z *phase = reinterpret_cast<const Z *>(A.phase);
    const R *fc = reinterpret_cast<const R *>(A.fc);

    const Z *e0src = OFF ? reinterpret_cast<const Z *>(A.off_e0) : evec + (size_t)i0 * ND2;
    for (int i = tid; i < ND2; i += BLK) s_e0[i] = e0src[i];
    for (int i = tid; i < A.L; i += BLK) s_ph[i] = phase[i];

    const bool own = tid < ND2;
    const int js1 = tid / ND, js2 = tid % ND;
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002);   // 1/(sigma sqrt(2 pi))
    const double hinv = 0.5 / (A.sigma * A.sigma);

    KSum<R> acc[S0B];
    double accd[S0B];
#pragma unroll
    for (int l = 0; l < S0B; ++l) accd[l] = 0.0;

    const int q1_lo = pair_mode ? A.pair_q1[ch] : ch * A.chunk;
    const int q1_hi = pair_mode ? q1_lo + 1 : min(A.nq, q1_lo + A.chunk);
    for (int q1 = q1_lo; q1 < q1_hi; ++q1) {
        int k10, k11, k12;
        grid_coords(q1, A.m0, A.m1, k10, k11, k12);
        int a0 = -(k00 + k10) % A.m0, a1 = -(k01 + k11) % A.m1, a2 = -(k02 + k12) % A.m2;
        a0 += a0 < 0 ? A.m0 : 0;
        a1 += a1 < 0 ? A.m1 : 0;
        a2 += a2 < 0 ? A.m2 : 0;
        const int i2 = a0 + A.m0 * (a1 + A.m1 * a2);
        if (!W_OUT && q1 > i2) continue;
        const double wp = W_OUT ? 1.0 : (q1 == i2 ? 1.0 : 2.0);
        __syncthreads();
        const Z *e2src = OFF ? reinterpret_cast<const Z *>(A.off_e2) + (size_t)q1 * ND2 : evec + (size_t)i2 * ND2;
        for (int i = tid; i < ND2; i += BLK) {
            s_e1[i] = evec[(size_t)q1 * ND2 + i];
            s_e2[i] = e2src[i];
        }
        for (int s = tid; s < A.nsite; s += BLK) {
            const int4 c = A.site[s];
            s_m1[s] = (k10 * c.x + k11 * c.y + k12 * c.z) % A.L;
            s_m2[s] = (a0 * c.x + a1 * c.y + a2 * c.z) % A.L;
            if constexpr (OFF) {
                // q2 = G - (q0 + q1): e^{2 pi i x2 . n} = e^{-2 pi i x0 . n} conj(e^{2 pi i x1 . n}) for
                // an integer cell n, whatever the G
                const Z p1 = s_ph[s_m1[s]];
                const Z p0 = reinterpret_cast<const Z *>(A.off_p0)[s];
                s_pa[s] = p1;
                s_pb[s] = cmk<R>(p0.x * p1.x + p0.y * p1.y, p0.y * p1.x - p0.x * p1.y);
            }
        }
        __syncthreads();

        Z V[S0B];
#pragma unroll
        for (int l = 0; l < S0B; ++l) V[l] = cmk<R>(0, 0);

        for (int a = 0; a < ND; ++a) {
            const int mu = a / 3, al = a % 3;
            if (own) {
                const int b = js1, c = js2;
                const int nu = b / 3, be = b % 3, eta = c / 3, ga = c % 3;
                const int blk = (mu * NB + nu) * NB + eta;
                Z t = cmk<R>(0, 0);
                const int t1 = A.blk_start[blk + 1];
                for (int u = A.blk_start[blk]; u < t1; ++u) {
                    Z ph;
                    if constexpr (OFF) {
                        const Z x = s_pa[A.tup_site[2 * u]], y = s_pb[A.tup_site[2 * u + 1]];
                        ph = cmk<R>(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
                    } else {
                        int m = s_m1[A.tup_site[2 * u]] + s_m2[A.tup_site[2 * u + 1]];
                        m -= m >= A.L ? A.L : 0;
                        ph = s_ph[m];
                    }
                    t = crfma(fc[(size_t)u * 27 + al * 9 + be * 3 + ga], ph, t);
                }
                s_T[tid] = t;
            }
            __syncthreads();
            if (own) {   // Z(b, s2) = sum_c T(b, c) e2(c, s2)
                const int b = js1, s2 = js2;
                Z z = cmk<R>(0, 0);
#pragma unroll 6
                for (int c = 0; c < ND; ++c) z = cfma(s_T[b * ND + c], s_e2[s2 * ND + c], z);
                s_Z[tid] = z;
            }
            __syncthreads();
            if (own) {   // Y(s1, s2) = sum_b e1(b, s1) Z(b, s2); V(s0, s1, s2) += e0(a, s0) Y
                Z y = cmk<R>(0, 0);
#pragma unroll 6
                for (int b = 0; b < ND; ++b) y = cfma(s_e1[js1 * ND + b], s_Z[b * ND + js2], y);
#pragma unroll
                for (int l = 0; l < S0B; ++l) {
                    const int s0 = s0b * S0B + l;
                    if (s0 < ND) V[l] = cfma(s_e0[s0 * ND + a], y, V[l]);
                }
            }
        }
        if (own) {
            const double w1 = A.omega[(size_t)q1 * ND + js1], w2 = A.omega[(size_t)i2 * ND + js2];
            const double n1 = A.nbe[(size_t)q1 * ND + js1], n2 = A.nbe[(size_t)i2 * ND + js2];
            const double iw12 = A.inv_omega[(size_t)q1 * ND + js1] *
                                (OFF ? A.off_iw2[(size_t)q1 * ND + js2] : A.inv_omega[(size_t)i2 * ND + js2]);
            const bool exists = !OFF || A.off_exists[q1];
#pragma unroll
            for (int l = 0; l < S0B; ++l) {
                const int s0 = s0b * S0B + l;
                if (s0 >= ND) continue;
                const double iw = (OFF ? A.off_iw0[s0] : A.inv_omega[(size_t)i0 * ND + s0]) * iw12;
                const double W = exists ? (double)cabs2(V[l]) * iw * A.factor : 0.0;
                if (W_OUT) {
                    A.w_out[(((size_t)(pair_mode ? ch : q1) * ND + s0) * ND + js1) * ND + js2] = W;
                } else {
                    double g = gauss_g(A.omega[(size_t)i0 * ND + s0], w1, w2, n1, n2, c0, hinv);
                    if (A.sel) g *= A.sel[s0] * A.sel[ND + js1 * ND + js2];
                    if (sizeof(R) == 8)
                        accd[l] += wp * W * g;
                    else
                        acc[l].add((R)(wp * W * g));
                }
            }
        }
    }
    if (W_OUT) return;
    // fixed-order block reduction of each s0 partial
    constexpr int P2 = BLK <= 64 ? 64 : BLK <= 128 ? 128 : BLK <= 256 ? 256 : BLK <= 512 ? 512 : 1024;
    __shared__ double s_red[P2];
    for (int l = 0; l < S0B; ++l) {
        s_red[tid] = sizeof(R) == 8 ? accd[l] : acc[l].value();
        if (tid + BLK < P2) s_red[tid + BLK] = 0.0;
        __syncthreads();
        for (int w = P2 / 2; w > 0; w >>= 1) {
            if (tid < w) s_red[tid] += s_red[tid + w];
            __syncthreads();
        }
        if (tid == 0)
            A.partial[(((size_t)qb * NS0B + s0b) * A.nchunks + ch) * S0B + l] = s_red[0];
        __syncthreads();
    }
}

// lw[q][s0] = sum over chunks of the partials, fixed order.
__global__ void k_reduce_partials(const double *partial, int nq0, int nd, int ns0b, int s0bs,
                                  int nchunks, double *lw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq0 * nd) return;
    const int q = i / nd, s0 = i % nd;
    const int b = s0 / s0bs, l = s0 % s0bs;
    const double *p = partial + (((size_t)q * ns0b + b) * nchunks) * s0bs + l;
    double s = 0.0;
    for (int c = 0; c < nchunks; ++c) s += p[(size_t)c * s0bs];
    lw[i] = s;
}

template <typename R, int ND>
static int launch_nd(Ctx &c, const int *d_q0, int nq0, double sigma, double *d_lw, double *d_w) {
    constexpr int S0B = ND <= 15 ? ND : 10;
    constexpr int NS0B = (ND + S0B - 1) / S0B;
    constexpr int BLK = ((ND * ND + 63) / 64) * 64;
    Lw3Args A;
    A.q0_grid = d_q0;
    A.nq = c.nq;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    // enough workgroups to fill 256 CUs several times over
    const long target = 4096;
    long per = (long)nq0 * NS0B;
    int nch = (int)std::min<long>(c.nq, std::max<long>(1, (target + per - 1) / per));
    A.chunk = (c.nq + nch - 1) / nch;
    A.nchunks = (c.nq + A.chunk - 1) / A.chunk;
    A.blk_start = c.fc3.blk_start;
    A.tup_site = c.fc3.tup_site;
    A.fc = sizeof(R) == 8 ? (const void *)c.fc3.fc64 : (const void *)c.fc3.fc32;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.phase = sizeof(R) == 8 ? (const void *)c.phase64 : (const void *)c.phase32;
    A.evec = sizeof(R) == 8 ? (const void *)c.evec : (const void *)c.evec32;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sel = (c.sel3 && !d_w) ? c.d_sel3 : nullptr;
    A.sigma = sigma;
    {
        // Sctr3THz2 from its defining factors (constants.f90:59-60)
        const double pi = 4.0 * std::atan(1.0);
        A.factor = 1.054571817 * (1.60217733 * 1.60217733) *
                   (6.02214076 * 6.02214076 * 6.02214076) * 1000000.0 * pi * 0.0625 / (double)c.nq;
    }
    A.w_out = d_w;
    const size_t dyn = (size_t)2 * ((A.nsite + 1) & ~1) * sizeof(int) + (size_t)c.L * sizeof(cplx_t<R>);
    A.pair_q0 = c.w_pair_q0;
    A.pair_q1 = c.w_pair_q1;
    dim3 grid(A.pair_q0 ? c.w_npairs : A.nchunks, NS0B, A.pair_q0 ? 1 : nq0);
    if (d_w) {
        A.partial = nullptr;
        hipLaunchKernelGGL((k_lw3_generic<R, ND, S0B, BLK, true>), grid, dim3(BLK), dyn, c.stream, A);
        SPR_HIP(hipGetLastError());
        return SPR_OK;
    }
    const size_t np = (size_t)nq0 * NS0B * A.nchunks * S0B;
    int rc;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, np * sizeof(double)))) return rc;
    A.partial = c.partial;
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    hipLaunchKernelGGL((k_lw3_generic<R, ND, S0B, BLK, false>), grid, dim3(BLK), dyn, c.stream, A);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    const int n = nq0 * ND;
    hipLaunchKernelGGL(k_reduce_partials, dim3((n + 255) / 256), dim3(256), 0, c.stream, c.partial,
                       nq0, ND, NS0B, S0B, A.nchunks, d_lw);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

// FP64 W of every grid q1 for an off-mesh q0 (the W rows of LineWidth's ScatterProbAllq1)
template <int ND>
static int launch_off_nd(Ctx &c, const OffMeshDev &om, double *d_w) {
    constexpr int S0B = ND <= 15 ? ND : 10;
    constexpr int NS0B = (ND + S0B - 1) / S0B;
    constexpr int BLK = ((ND * ND + 63) / 64) * 64;
    Lw3Args A{};
    A.q0_grid = nullptr;
    A.nq = c.nq;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    const long target = 4096;
    int nch = (int)std::min<long>(c.nq, std::max<long>(1, (target + NS0B - 1) / NS0B));
    A.chunk = (c.nq + nch - 1) / nch;
    A.nchunks = (c.nq + A.chunk - 1) / A.chunk;
    A.blk_start = c.fc3.blk_start;
    A.tup_site = c.fc3.tup_site;
    A.fc = c.fc3.fc64;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.phase = c.phase64;
    A.evec = c.evec;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sel = nullptr;
    A.sigma = 1.0;
    const double pi = 4.0 * std::atan(1.0);
    A.factor = 1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) * 1000000.0 * pi *
               0.0625 / (double)c.nq;   // Sctr3THz2 / Nq (constants.f90:59-60)
    A.w_out = d_w;
    A.pair_q0 = A.pair_q1 = nullptr;
    A.off_e0 = om.e0;
    A.off_e2 = om.e2;
    A.off_p0 = om.p0;
    A.off_iw0 = om.iw0;
    A.off_iw2 = om.iw2;
    A.off_exists = om.exists;
    const size_t dyn = (size_t)2 * ((A.nsite + 1) & ~1) * sizeof(int) + (size_t)(c.L + 2 * A.nsite) * sizeof(double2);
    if (dyn > 64 * 1024)
        SPR_HIP(hipFuncSetAttribute((const void *)k_lw3_generic<double, ND, S0B, BLK, true, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    hipLaunchKernelGGL((k_lw3_generic<double, ND, S0B, BLK, true, true>), dim3(A.nchunks, NS0B, 1), dim3(BLK), dyn,
                       c.stream, A);
    SPR_HIP(hipGetLastError());
    return SPR_OK;
}

int launch_w_offmesh(Ctx &c, const OffMeshDev &om, double *d_w) {
    switch (c.nd) {
        case 3: return launch_off_nd<3>(c, om, d_w);
        case 6: return launch_off_nd<6>(c, om, d_w);
        case 9: return launch_off_nd<9>(c, om, d_w);
        case 12: return launch_off_nd<12>(c, om, d_w);
        case 15: return launch_off_nd<15>(c, om, d_w);
        case 18: return launch_off_nd<18>(c, om, d_w);
        case 24: return launch_off_nd<24>(c, om, d_w);
        case 30: return launch_off_nd<30>(c, om, d_w);
        default:
            set_error("off-mesh three-phonon W: Ndof " + std::to_string(c.nd) + " not built (3..18, 24, 30)");
            return SPR_E_UNSUPPORTED;
    }
}

template <typename R>
static int dispatch_generic(Ctx &c, const int *d_q0, int nq0, double sigma, double *d_lw, double *d_w) {
    switch (c.nd) {
        case 3: return launch_nd<R, 3>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 6: return launch_nd<R, 6>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 9: return launch_nd<R, 9>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 12: return launch_nd<R, 12>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 15: return launch_nd<R, 15>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 18: return launch_nd<R, 18>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 24: return launch_nd<R, 24>(c, d_q0, nq0, sigma, d_lw, d_w);
        case 30: return launch_nd<R, 30>(c, d_q0, nq0, sigma, d_lw, d_w);
        default:
            set_error("three-phonon kernel: Ndof " + std::to_string(c.nd) + " not built (3..18, 24, 30)");
            return SPR_E_UNSUPPORTED;
    }
}

int launch_lw3_generic(Ctx &c, const int *d_q0, int nq0, double sigma, int prec, double *d_lw,
                       double *d_w) {
    if (d_w || prec == SPR_PREC_FP64) return dispatch_generic<double>(c, d_q0, nq0, sigma, d_lw, d_w);
    return dispatch_generic<float>(c, d_q0, nq0, sigma, d_lw, d_w);
}

}  // namespace spr
