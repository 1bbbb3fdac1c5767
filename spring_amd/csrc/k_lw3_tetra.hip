// Three-phonon linewidths with the linear tetrahedron method (CalculateLwTetrahedron,
// ThirdOrder_m.f90:169-275, with FindgAndIver / FindIverxFver, TetrahedronIntegration_m.f90:
// 96-256): for every s0, tetrahedron of the q1 mesh and branch pair (s1, s2), three processes
//   M1 = w1 + w2, F1 = W (n1 + n2 + 1)        (+)
//   M2 = w2 - w1, F2 = W (n1 - n2)            (+)
//   M3 = w1 - w2, F3 = F2                     (-)
// with q2 = -(q0 + q1) at each vertex, each contributing g sum_v I_v F_v / 6 at Mq = w0(s0);
// P3_q0(s0) accumulates g of all three processes.  W(q1, s0, s1, s2) of the q0 comes from the
// FP64 W path; all arithmetic is FP64.
//
// Mapping: one workgroup per (tetrahedron chunk, s0); a thread owns (s1 s2, tetrahedron) of
// its chunk, tetrahedron fastest across lanes, and reads W and n only when w0(s0) lies inside
// one of the three processes' vertex-energy ranges.  Per-workgroup partials are reduced in a
// fixed order by a second kernel (bitwise reproducible).
#include <cstdlib>
#include <string>
#include <vector>

#include "spr_internal.h"
#include "tetra_dev.h"

namespace spr {

namespace {

// g and sum_v I_v F_v / 6 for one process; M and F are sorted together by M (stable: equal M
// keep the vertex order, as the reference's sort keeps the first of equal values first)
__device__ __forceinline__ void tetra_gif(double mq, const double *Min, const double *Fin, double &g, double &gif) {
    int idx[4] = {0, 1, 2, 3};
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int b = a; b > 0; --b)
            if (Min[idx[b - 1]] > Min[idx[b]]) {
                const int t = idx[b];
                idx[b] = idx[b - 1];
                idx[b - 1] = t;
            }
    double M[4], F[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        M[v] = Min[idx[v]];
        F[v] = Fin[idx[v]];
    }
    double I0 = 0, I1 = 0, I2 = 0, I3 = 0;
    g = 0.0;
    gif = 0.0;
    if (mq >= M[0] && mq < M[1]) {
        g = 3.0 * (mq - M[0]) * (mq - M[0]) / ((M[1] - M[0]) * (M[2] - M[0]) * (M[3] - M[0]));
        I0 = ((mq - M[1]) / (M[0] - M[1]) + (mq - M[2]) / (M[0] - M[2]) + (mq - M[3]) / (M[0] - M[3])) / 3.0;
        I1 = ((mq - M[0]) / (M[1] - M[0])) / 3.0;
        I2 = ((mq - M[0]) / (M[2] - M[0])) / 3.0;
        I3 = ((mq - M[0]) / (M[3] - M[0])) / 3.0;
    } else if (mq >= M[1] && mq < M[2]) {
        g = (3.0 / ((M[1] - M[2]) * (M[3] - M[0]))) *
            (((mq - M[0]) * (mq - M[2])) / (M[2] - M[0]) - ((mq - M[1]) * (mq - M[3])) / (M[1] - M[3]));
        const double gd = g * (M[3] - M[0]);
        I0 = ((mq - M[3]) / (M[0] - M[3])) / 3.0 +
             ((mq - M[2]) / (M[0] - M[2])) * ((mq - M[0]) / (M[2] - M[0])) * ((mq - M[2]) / (M[1] - M[2])) / gd;
        I1 = ((mq - M[2]) / (M[1] - M[2])) / 3.0 +
             ((mq - M[3]) / (M[1] - M[3])) * ((mq - M[3]) / (M[1] - M[3])) * ((mq - M[1]) / (M[2] - M[1])) / gd;
        I2 = ((mq - M[1]) / (M[2] - M[1])) / 3.0 +
             ((mq - M[0]) / (M[2] - M[0])) * ((mq - M[0]) / (M[2] - M[0])) * ((mq - M[2]) / (M[1] - M[2])) / gd;
        I3 = ((mq - M[0]) / (M[3] - M[0])) / 3.0 +
             ((mq - M[1]) / (M[3] - M[1])) * ((mq - M[3]) / (M[1] - M[3])) * ((mq - M[1]) / (M[2] - M[1])) / gd;
    } else if (mq >= M[2] && mq <= M[3]) {
        g = 3.0 * (mq - M[3]) * (mq - M[3]) / ((M[3] - M[0]) * (M[3] - M[1]) * (M[3] - M[2]));
        I0 = ((mq - M[3]) / (M[0] - M[3])) / 3.0;
        I1 = ((mq - M[3]) / (M[1] - M[3])) / 3.0;
        I2 = ((mq - M[3]) / (M[2] - M[3])) / 3.0;
        I3 = ((mq - M[0]) / (M[3] - M[0]) + (mq - M[1]) / (M[3] - M[1]) + (mq - M[2]) / (M[3] - M[2])) / 3.0;
    } else {
        return;
    }
    gif = g * ((I0 * F[0] + I1 * F[1] + I2 * F[2] + I3 * F[3]) / 6.0);
}

constexpr int TBLK = 256;

// branch ranges of s0, s1, s2 (Polarization3rd, 0-based inclusive): the loops of
// CalculateLwTetrahedron (ThirdOrder_m.f90:207-214)
struct Ranges3 {
    int lo[3], hi[3];
};

__global__ __launch_bounds__(TBLK) void k_lw3_tetra(const double *W, const int *tetra, const int *partner,
                                                    const double *omega, const double *nbe, const double *omega0,
                                                    int nd, int ntet, int chunk, Ranges3 rg, double *partial) {
    const int ch = blockIdx.x, s0 = blockIdx.y, nch = gridDim.x;
    const bool s0_in = s0 >= rg.lo[0] && s0 <= rg.hi[0];
    const int nd2 = nd * nd;
    const int t_lo = ch * chunk, t_hi = min(ntet, t_lo + chunk);
    const double mq = omega0[s0];
    double lw = 0.0, p3 = 0.0;
    const int nt = t_hi - t_lo;
    for (int i = threadIdx.x; i < nt * nd2; i += TBLK) {
        // tetrahedron fastest across lanes: a wave holds neighbouring tetrahedra of one branch
        // pair, whose energy ranges are alike, so whole waves skip together below
        const int p = i / nt, t = t_lo + (i - p * nt);
        const int s1 = p / nd, s2 = p % nd;
        if (!s0_in || s1 < rg.lo[1] || s1 > rg.hi[1] || s2 < rg.lo[2] || s2 > rg.hi[2]) continue;
        double F1[4], F2[4], M1[4], M2[4], M3[4];
        int q1v[4], q2v[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q1 = tetra[4 * t + v], q2 = partner[q1];
            q1v[v] = q1;
            q2v[v] = q2;
            const double w1 = omega[(size_t)q1 * nd + s1], w2 = omega[(size_t)q2 * nd + s2];
            M1[v] = w1 + w2;
            M2[v] = w2 - w1;
            M3[v] = w1 - w2;
        }
        // a process adds g and weights only for M0 <= mq <= M3 of its vertex energies (the union
        // of the three cases of tetra_gif); outside every range W and n are not read
        const double lo1 = fmin(fmin(M1[0], M1[1]), fmin(M1[2], M1[3]));
        const double hi1 = fmax(fmax(M1[0], M1[1]), fmax(M1[2], M1[3]));
        const double lo2 = fmin(fmin(M2[0], M2[1]), fmin(M2[2], M2[3]));
        const double hi2 = fmax(fmax(M2[0], M2[1]), fmax(M2[2], M2[3]));
        if (!((mq >= lo1 && mq <= hi1) || (mq >= lo2 && mq <= hi2) || (mq >= -hi2 && mq <= -lo2))) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q1 = q1v[v], q2 = q2v[v];
            const double w = W[((size_t)q1 * nd + s0) * nd2 + p];
            const double n1 = nbe[(size_t)q1 * nd + s1], n2 = nbe[(size_t)q2 * nd + s2];
            F1[v] = w * (n1 + n2 + 1.0);
            F2[v] = w * (n1 - n2);
        }
        double g, gif;
        tetra_gif(mq, M1, F1, g, gif);
        lw += gif;
        p3 += g;
        tetra_gif(mq, M2, F2, g, gif);
        lw += gif;
        p3 += g;
        tetra_gif(mq, M3, F2, g, gif);
        lw -= gif;
        p3 += g;
    }
    __shared__ double s_lw[TBLK], s_p3[TBLK];
    s_lw[threadIdx.x] = lw;
    s_p3[threadIdx.x] = p3;
    __syncthreads();
    for (int w = TBLK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_lw[threadIdx.x] += s_lw[threadIdx.x + w];
            s_p3[threadIdx.x] += s_p3[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[((size_t)s0 * nch + ch) * 2] = s_lw[0];
        partial[((size_t)s0 * nch + ch) * 2 + 1] = s_p3[0];
    }
}

// Every s0 in one pass (Ndof 3 to 30).  A thread owns (s1 s2, tetrahedron), tetrahedron fastest
// across lanes; the vertex energies of the three processes do not depend on s0, so each process
// whose energy span meets [min w0, max w0] is sorted once (with its occupation factor and W
// offset as payload) and its six vertex-energy reciprocals are taken once, then every s0 with
// Mq inside the span reads its four W and adds g I F / 6.  partial[s0][workgroup][2].
template <int ND>
__global__ __launch_bounds__(TBLK) void k_lw3_tetra_s(const double *Wb, const int *tetra, const int *partner_b,
                                                      const double *omega, const double *nbe, const int *q0_grid,
                                                      int nq, int ntet, Ranges3 rg, double *partial_b) {
    constexpr int ND2 = ND * ND;
    const int tid = threadIdx.x;
    // blockIdx.y: the q0 of the batch (its W rows, partner table and partial rows)
    const int kb = blockIdx.y;
    const double *W = Wb + (size_t)kb * nq * ND * ND2;
    const int *partner = partner_b + (size_t)kb * nq;
    double *partial = partial_b + (size_t)kb * ND * gridDim.x * 2;
    const double *omega0 = omega + (size_t)q0_grid[kb] * ND;
    __shared__ double s_w0[ND];
    if (tid < ND) s_w0[tid] = omega0[tid];
    __syncthreads();
    double w0lo = s_w0[0], w0hi = s_w0[0];
#pragma unroll
    for (int s = 1; s < ND; ++s) {
        w0lo = fmin(w0lo, s_w0[s]);
        w0hi = fmax(w0hi, s_w0[s]);
    }
    double lw[ND], p3[ND];
#pragma unroll
    for (int s = 0; s < ND; ++s) lw[s] = p3[s] = 0.0;
    const int total = ntet * ND2;
    for (int it = blockIdx.x * TBLK + tid; it < total; it += gridDim.x * TBLK) {
        const int p = it / ntet, t = it - p * ntet;
        const int s1 = p / ND, s2 = p % ND;
        if (s1 < rg.lo[1] || s1 > rg.hi[1] || s2 < rg.lo[2] || s2 > rg.hi[2]) continue;
        double w1[4], w2[4];
        int q1v[4], q2v[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q1 = tetra[4 * t + v], q2 = partner[q1];
            q1v[v] = q1;
            q2v[v] = q2;
            w1[v] = omega[q1 * ND + s1];
            w2[v] = omega[q2 * ND + s2];
        }
#pragma unroll 1
        for (int k = 0; k < 3; ++k) {
            // M1 = w1 + w2 with F1 = W (n1 + n2 + 1); M2 = w2 - w1 and M3 = w1 - w2, both with
            // F2 = W (n1 - n2), the third subtracted
            Vx P[4];
            double mlo = 0.0, mhi = 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double m = k == 0 ? w1[v] + w2[v] : k == 1 ? w2[v] - w1[v] : w1[v] - w2[v];
                P[v] = Vx{m, 0.0, q1v[v] * ND * ND2 + p};
                mlo = v == 0 ? m : fmin(mlo, m);
                mhi = v == 0 ? m : fmax(mhi, m);
            }
            if (mhi < w0lo || mlo > w0hi) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double n1 = nbe[q1v[v] * ND + s1], n2 = nbe[q2v[v] * ND + s2];
                P[v].a = k == 0 ? n1 + n2 + 1.0 : n1 - n2;
            }
            sort4(P);
            const double M[4] = {P[0].M, P[1].M, P[2].M, P[3].M};
            const Rcp R = rcp6(M);
#pragma unroll 1
            for (int s0 = rg.lo[0]; s0 <= rg.hi[0]; ++s0) {
                const double mq = s_w0[s0];
                if (!(mq >= M[0] && mq <= M[3])) continue;
                double F[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) F[v] = W[P[v].off + s0 * ND2] * P[v].a;
                double g;
                const double gif = tetra_rcp<false>(mq, M, R, F, g);
#pragma unroll
                for (int s = 0; s < ND; ++s)
                    if (s == s0) {
                        lw[s] += k == 2 ? -gif : gif;
                        p3[s] += g;
                    }
            }
        }
    }
    __shared__ double s_lw[TBLK], s_p3[TBLK];
#pragma unroll
    for (int s = 0; s < ND; ++s) {
        __syncthreads();
        s_lw[tid] = lw[s];
        s_p3[tid] = p3[s];
        __syncthreads();
        for (int w = TBLK / 2; w > 0; w >>= 1) {
            if (tid < w) {
                s_lw[tid] += s_lw[tid + w];
                s_p3[tid] += s_p3[tid + w];
            }
            __syncthreads();
        }
        if (tid == 0) {
            partial[((size_t)s * gridDim.x + blockIdx.x) * 2] = s_lw[0];
            partial[((size_t)s * gridDim.x + blockIdx.x) * 2 + 1] = s_p3[0];
        }
    }
}

// lw[s0] and p3[s0]: the chunk partials of s0 summed in a fixed order, one workgroup per s0
// (strided per-thread sums, then a tree)
__global__ __launch_bounds__(256) void k_tetra_reduce(const double *partial_b, int nd, int nch, double *lw_b,
                                                      double *p3_b) {
    // blockIdx.y: the q0 of a batch (partial rows and outputs of q0 k at k * nd * nch * 2, k * nd)
    const int s0 = blockIdx.x, tid = threadIdx.x;
    const double *partial = partial_b + (size_t)blockIdx.y * nd * nch * 2;
    double *lw = lw_b + (size_t)blockIdx.y * nd, *p3 = p3_b + (size_t)blockIdx.y * nd;
    double a = 0.0, b = 0.0;
    for (int c = tid; c < nch; c += 256) {
        a += partial[((size_t)s0 * nch + c) * 2];
        b += partial[((size_t)s0 * nch + c) * 2 + 1];
    }
    __shared__ double ra[256], rb[256];
    ra[tid] = a;
    rb[tid] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            ra[tid] += ra[tid + w];
            rb[tid] += rb[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        lw[s0] = ra[0];
        p3[s0] = rb[0];
    }
}

}  // namespace

int ensure_tetra(Ctx &c) {
    if (c.tetra) return SPR_OK;
    if (c.h_tetra.empty()) {
        set_error("tetrahedron linewidth: the context was created without tetrahedra");
        return SPR_E_STATE;
    }
    SPR_HIP(hipMalloc((void **)&c.tetra, c.h_tetra.size() * sizeof(int)));
    SPR_HIP(hipMemcpy(c.tetra, c.h_tetra.data(), c.h_tetra.size() * sizeof(int), hipMemcpyHostToDevice));
    return SPR_OK;
}

int tetra_sums_from_w(Ctx &c, const double *d_w, const int *d_partner, const double *d_omega, const double *d_nbe,
                      const double *d_omega0, double *d_lw, double *d_p3) {
    int rc;
    if ((rc = ensure_tetra(c))) return rc;
    const int ntet = 6 * c.nq;
    const int chunk = std::max(1, (ntet + 255) / 256);
    const int nch = (ntet + chunk - 1) / chunk;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)nch * c.nd * 2 * sizeof(double)))) return rc;
    Ranges3 rg;
    for (int k = 0; k < 3; ++k) {
        rg.lo[k] = 0;
        rg.hi[k] = c.nd - 1;
    }
    hipLaunchKernelGGL(k_lw3_tetra, dim3(nch, c.nd), dim3(TBLK), 0, c.stream, d_w, c.tetra, d_partner, d_omega, d_nbe,
                       d_omega0, c.nd, ntet, chunk, rg, c.partial);
    SPR_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_tetra_reduce, dim3(c.nd, 1), dim3(256), 0, c.stream, c.partial, c.nd, nch, d_lw, d_p3);
    SPR_HIP(hipGetLastError());
    return SPR_OK;
}

// Tetrahedron linewidths of irreducible q0 [q_first, q_first + q_count) into d_lw[k][nd] and
// d_p3[k][nd] (device, k = q - q_first).  W of every q1 for up to 16 q0 at a time in c.wout,
// pair-reduced with both orderings written: the small-cell kernel for Nb <= 2, the matrix-core
// kernel's W mode for the larger cells it is built for; otherwise (or with SPR_KERNEL=generic)
// the generic W path one q0 at a time.  The partner tables of a batch are
// uploaded once.
int launch_lw3_tetra(Ctx &c, int q_first, int q_count, double *d_lw, double *d_p3) {
    int rc;
    const size_t nw = (size_t)c.nq * c.nd * c.nd * c.nd;
    const char *force = getenv("SPR_KERNEL");
    const bool generic = force && std::string(force) == "generic";
    const bool small = c.nb <= 2 && !c.override_items && !generic;
    const bool mmw = !small && !c.override_items && !generic && lw3_mm_w_built(c.nb);
    // W rows of up to 16 q0 per batch: a batch launch has 16x the work units of a one-q0 launch
    // (a 10-atom 20^3 q0 alone gives ~250 units, too few for 256 CUs); the buffer takes at most a
    // quarter of the free device memory, capped at 32 GiB, and at least 2 GiB
    size_t wcap = size_t(2) << 30;
    {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) wcap = std::max(wcap, std::min(size_t(32) << 30, fr / 4));
    }
    const int bmax = small || mmw ? (int)std::max<size_t>(1, std::min<size_t>(16, wcap / (nw * sizeof(double)))) : 1;
    const int B = std::min(bmax, q_count);
    if ((rc = ensure((void **)&c.wout, &c.wout_cap, (size_t)B * nw * sizeof(double)))) return rc;
    if ((rc = ensure_tetra(c))) return rc;
    if ((rc = ensure((void **)&c.partner, &c.partner_cap, (size_t)B * c.nq * sizeof(int)))) return rc;
    const int ntet = 6 * c.nq;
    // the Ndof built for k_lw3_tetra_s: all s0 per pass, one partial per workgroup; otherwise one
    // workgroup per (tetrahedron chunk, s0)
    const bool all_s0 = c.nd == 3 || c.nd == 6 || c.nd == 9 || c.nd == 12 || c.nd == 15 || c.nd == 18 ||
                        c.nd == 24 || c.nd == 30;
    const int chunk = std::max(1, (ntet + 255) / 256);
    const int nch = all_s0 ? (int)std::min<long>(8L * c.ncu, ((long)ntet * c.nd * c.nd + TBLK - 1) / TBLK)
                           : (ntet + chunk - 1) / chunk;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)B * nch * c.nd * 2 * sizeof(double)))) return rc;
    const int m0 = c.mesh[0], m1 = c.mesh[1], m2 = c.mesh[2];
    std::vector<int> part((size_t)B * c.nq);
    for (int b0 = q_first; b0 < q_first + q_count; b0 += B) {
        const int nb = std::min(B, q_first + q_count - b0);
        if (small) {
            if ((rc = launch_lw3_small_w(c, b0, nb, c.wout))) return rc;
        } else if (mmw) {
            if ((rc = launch_lw3_mm_w(c, b0, nb, c.wout))) return rc;
        } else if ((rc = launch_lw3(c, b0, 1, 0.1, SPR_PREC_FP64, nullptr, c.wout))) {
            return rc;
        }
        // partner tables of the batch: q2 = -(q0 + q1) on the grid
        for (int k = 0; k < nb; ++k) {
            const int i0 = c.h_irr_grid[b0 + k];
            for (int q1 = 0; q1 < c.nq; ++q1) {
                const int a0 = ((-(i0 % m0 + q1 % m0)) % m0 + m0) % m0;
                const int a1 = ((-((i0 / m0) % m1 + (q1 / m0) % m1)) % m1 + m1) % m1;
                const int a2 = ((-(i0 / (m0 * m1) + q1 / (m0 * m1))) % m2 + m2) % m2;
                part[(size_t)k * c.nq + q1] = a0 + m0 * (a1 + m1 * a2);
            }
        }
        SPR_HIP(hipMemcpyAsync(c.partner, part.data(), (size_t)nb * c.nq * sizeof(int), hipMemcpyHostToDevice,
                               c.stream));
        SPR_HIP(hipEventRecord(c.ev0, c.stream));
        Ranges3 rg;
        for (int k = 0; k < 3; ++k) {
            rg.lo[k] = c.pol3[2 * k];
            rg.hi[k] = c.pol3[2 * k + 1];
        }
        if (all_s0) {
            // one launch for the batch's q0 (grid.y), then one reduction launch
            const int *q0g = c.irr_grid + b0;
            switch (c.nd) {
#define SPR_TS(n)                                                                                                   \
    case n:                                                                                                         \
        hipLaunchKernelGGL(k_lw3_tetra_s<n>, dim3(nch, nb), dim3(TBLK), 0, c.stream, c.wout, c.tetra, c.partner, \
                           c.omega, c.nbe, q0g, c.nq, ntet, rg, c.partial);                                         \
        break;
                SPR_TS(3) SPR_TS(6) SPR_TS(9) SPR_TS(12) SPR_TS(15) SPR_TS(18) SPR_TS(24) SPR_TS(30)
#undef SPR_TS
            }
            SPR_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_tetra_reduce, dim3(c.nd, nb), dim3(256), 0, c.stream, c.partial, c.nd, nch,
                               d_lw + (size_t)(b0 - q_first) * c.nd, d_p3 + (size_t)(b0 - q_first) * c.nd);
            SPR_HIP(hipGetLastError());
        } else {
            for (int k = 0; k < nb; ++k) {
                const int qi = b0 + k, i0 = c.h_irr_grid[qi];
                hipLaunchKernelGGL(k_lw3_tetra, dim3(nch, c.nd), dim3(TBLK), 0, c.stream, c.wout + (size_t)k * nw,
                                   c.tetra, c.partner + (size_t)k * c.nq, c.omega, c.nbe,
                                   c.omega + (size_t)i0 * c.nd, c.nd, ntet, chunk, rg, c.partial);
                SPR_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_tetra_reduce, dim3(c.nd, 1), dim3(256), 0, c.stream, c.partial, c.nd, nch,
                                   d_lw + (size_t)(qi - q_first) * c.nd, d_p3 + (size_t)(qi - q_first) * c.nd);
                SPR_HIP(hipGetLastError());
            }
        }
        SPR_HIP(hipEventRecord(c.ev1, c.stream));
        // the host partner tables must outlive the copy
        SPR_HIP(hipStreamSynchronize(c.stream));
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess) {
            c.launches += nb;
            c.kernel_ms += ms;
        }
    }
    return SPR_OK;
}

}  // namespace spr
