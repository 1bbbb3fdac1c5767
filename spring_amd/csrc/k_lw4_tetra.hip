// Four-phonon linewidths with the linear tetrahedron method (TetrahedronIntegration4th,
// SumOverPolarization, FindFandOmegaVer_4th; Tetrahedron4th_sep.f90:37-425).  For q0, every q1
// of the mesh and every tetrahedron of the q2 mesh, vertex v carries q2_v and
// q3_v = -(q0 + q1 + q2_v); for every branch set (s0, s1, s2, s3) three processes
//   M1 = w1 + w2 + w3,  F1 = W n1 n2 n3 / n0              added with 1/6
//   M2 = w2 + w3 - w1,  F2 = W (1+n1) n2 n3 / n0          added with 1/2
//   M3 = w3 - w1 - w2,  F3 = W (1+n1)(1+n2) n3 / n0       added with 1/2
// each contribute g * sum_v I_v F_v / 6 at Mq = w0(s0) (FindgAndIver / FindIverxFver,
// TetrahedronIntegration_m.f90:96-256), with NaN weights and densities set to zero as the
// reference does for degenerate vertices (:238-263); s0 < 3 skipped at q0 = Gamma.  The sum is
// the raw lw; main.f90:660 adds 0.5 lw to the linewidth.
//
// W4(q1, q2) comes from the FP64 store mode of k_lw4 for a chunk of q1 rows at a time (every
// q2 of a row: the tetrahedra of a q1 touch all of them), so device memory stays bounded.
// The reference keeps W4 once per permutation class of (q1, q2, q3) and permutes (s1, s2, s3)
// to read the other orders; with an index-permutation-symmetric FC4 that is the same number as
// W4 of the ordered triple, computed here.
//
// Mapping: a thread owns (q1, tetrahedron, s1 s2 s3) and walks s0; the vertex energies do not
// depend on s0, so each process's vertices are sorted once (stable, with the vertex's n-product
// and W address as payload) and reused for every s0.  W reads are contiguous in (s1 s2 s3).
// Grid-stride over items; per-workgroup partials reduced in a fixed order (bitwise
// reproducible).  All arithmetic FP64.
#include <algorithm>
#include <cstdlib>
#include <string>

#include "spr_internal.h"

namespace spr {

namespace {

constexpr int T4BLK = 256;

struct Tet4Args {
    const double *W;            // [nq1][nq][ND^4] rows of q1 in [q1_lo, q1_lo + nq1)
    const int *tetra;           // [6 nq][4]
    const double *omega, *nbe;  // [nq][ND]
    int nq, m0, m1, m2;
    int i0, q1_lo, nq1;
    double *partial;            // [gridDim.x][ND] of this chunk
};

// one vertex record: energy, s0-independent part of F, W offset of s0 = 0
struct Vx {
    double M, a;
    long off;
};

__device__ __forceinline__ void cswap(Vx &x, Vx &y) {
    // strict: equal energies keep their order (the reference's sort keeps the first of equal
    // values first)
    if (x.M > y.M) {
        const Vx t = x;
        x = y;
        y = t;
    }
}

// stable insertion sort of four records by M, fully unrolled (compare-exchange of neighbours)
__device__ __forceinline__ void sort4(Vx *v) {
    cswap(v[0], v[1]);
    cswap(v[1], v[2]);
    cswap(v[0], v[1]);
    cswap(v[2], v[3]);
    cswap(v[1], v[2]);
    cswap(v[0], v[1]);
}

// g * (1/6) sum_v I_v F_v of one sorted tetrahedron at mq, NaN I_v and g set to zero
__device__ __forceinline__ double tetra4(double mq, const double *M, const double *F) {
    double g, I0, I1, I2, I3;
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
        return 0.0;
    }
    I0 = I0 != I0 ? 0.0 : I0;
    I1 = I1 != I1 ? 0.0 : I1;
    I2 = I2 != I2 ? 0.0 : I2;
    I3 = I3 != I3 ? 0.0 : I3;
    g = g != g ? 0.0 : g;
    double ixf = 0.0;
    ixf += I0 * F[0];
    ixf += I1 * F[1];
    ixf += I2 * F[2];
    ixf += I3 * F[3];
    return g * ((1.0 / 6.0) * ixf);
}

template <int ND>
__global__ __launch_bounds__(T4BLK) void k_lw4_tetra(Tet4Args A) {
    constexpr int ND2 = ND * ND, ND3 = ND2 * ND, ND4 = ND3 * ND;
    const int tid = threadIdx.x;
    const int ntet = 6 * A.nq;
    const long total = (long)A.nq1 * ntet * ND3;
    const int i0 = A.i0;
    const int k00 = i0 % A.m0, k01 = (i0 / A.m0) % A.m1, k02 = i0 / (A.m0 * A.m1);
    const int s0_lo = i0 == 0 ? 3 : 0;
    double acc[ND];
#pragma unroll
    for (int s = 0; s < ND; ++s) acc[s] = 0.0;
    for (long it = (long)blockIdx.x * T4BLK + tid; it < total; it += (long)gridDim.x * T4BLK) {
        const int p = (int)(it % ND3);
        const long r = it / ND3;
        const int t = (int)(r % ntet), q1l = (int)(r / ntet);
        const int q1 = A.q1_lo + q1l;
        const int s1 = p / ND2, s2 = (p / ND) % ND, s3 = p % ND;
        const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
        const double w1 = A.omega[(size_t)q1 * ND + s1], n1 = A.nbe[(size_t)q1 * ND + s1];
        Vx P1[4], P2[4], P3[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q2 = A.tetra[4 * t + v];
            const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
            int a0 = -(k00 + k10 + k20) % A.m0, a1 = -(k01 + k11 + k21) % A.m1, a2 = -(k02 + k12 + k22) % A.m2;
            a0 += a0 < 0 ? A.m0 : 0;
            a1 += a1 < 0 ? A.m1 : 0;
            a2 += a2 < 0 ? A.m2 : 0;
            const int q3 = a0 + A.m0 * (a1 + A.m1 * a2);
            const double w2 = A.omega[(size_t)q2 * ND + s2], n2 = A.nbe[(size_t)q2 * ND + s2];
            const double w3 = A.omega[(size_t)q3 * ND + s3], n3 = A.nbe[(size_t)q3 * ND + s3];
            const long off = ((long)q1l * A.nq + q2) * ND4 + p;
            P1[v] = Vx{w1 + w2 + w3, n1 * n2 * n3, off};
            P2[v] = Vx{w2 + w3 - w1, (1.0 + n1) * n2 * n3, off};
            P3[v] = Vx{w3 - w1 - w2, (1.0 + n1) * (1.0 + n2) * n3, off};
        }
        sort4(P1);
        sort4(P2);
        sort4(P3);
        const double M1[4] = {P1[0].M, P1[1].M, P1[2].M, P1[3].M};
        const double M2[4] = {P2[0].M, P2[1].M, P2[2].M, P2[3].M};
        const double M3[4] = {P3[0].M, P3[1].M, P3[2].M, P3[3].M};
#pragma unroll 1
        for (int s0 = s0_lo; s0 < ND; ++s0) {
            const double mq = A.omega[(size_t)i0 * ND + s0], n0 = A.nbe[(size_t)i0 * ND + s0];
            const long so = (long)s0 * ND3;
            double F1[4], F2[4], F3[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                F1[v] = A.W[P1[v].off + so] * (P1[v].a / n0);
                F2[v] = A.W[P2[v].off + so] * (P2[v].a / n0);
                F3[v] = A.W[P3[v].off + so] * (P3[v].a / n0);
            }
            const double t1 = (1.0 / 6.0) * tetra4(mq, M1, F1);
            const double t2 = 0.5 * tetra4(mq, M2, F2);
            const double t3 = 0.5 * tetra4(mq, M3, F3);
#pragma unroll
            for (int s = 0; s < ND; ++s)
                if (s == s0) {
                    acc[s] += t1;
                    acc[s] += t2;
                    acc[s] += t3;
                }
        }
    }
    __shared__ double s_red[T4BLK];
#pragma unroll
    for (int s = 0; s < ND; ++s) {
        __syncthreads();
        s_red[tid] = acc[s];
        __syncthreads();
        for (int w = T4BLK / 2; w > 0; w >>= 1) {
            if (tid < w) s_red[tid] += s_red[tid + w];
            __syncthreads();
        }
        if (tid == 0) A.partial[(size_t)blockIdx.x * ND + s] = s_red[0];
    }
}

// bytes of W4 rows held at once (a chunk of q1; a q1 row is nq * Ndof^4 doubles): 2 GiB, or
// SPR_W4_CHUNK_MB megabytes (at least one row)
size_t w4_chunk_bytes() {
    const char *e = getenv("SPR_W4_CHUNK_MB");
    const long mb = e ? atol(e) : 0;
    return mb > 0 ? (size_t)mb << 20 : size_t(2) << 30;
}

}  // namespace

int launch_lw4_tetra(Ctx &c, int qi, double *d_lw) {
    if (c.nd != 3 && c.nd != 6) {
        set_error("four-phonon tetrahedron path: Ndof " + std::to_string(c.nd) + " not built (3, 6)");
        return SPR_E_UNSUPPORTED;
    }
    int rc;
    if ((rc = ensure_tetra(c))) return rc;
    const size_t nd4 = (size_t)c.nd * c.nd * c.nd * c.nd;
    const size_t row = (size_t)c.nq * nd4 * sizeof(double);
    const int nq1c = (int)std::max<size_t>(1, std::min<size_t>((size_t)c.nq, w4_chunk_bytes() / row));
    const int ncalls = (c.nq + nq1c - 1) / nq1c;
    const long items = (long)nq1c * 6 * c.nq * c.nd * c.nd * c.nd;
    const int grid = (int)std::min<long>(8L * c.ncu, (items + T4BLK - 1) / T4BLK);
    if ((rc = ensure((void **)&c.w4buf, &c.w4buf_cap, (size_t)nq1c * row))) return rc;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)ncalls * grid * c.nd * sizeof(double))))
        return rc;
    const int i0 = c.h_irr_grid[qi];
    Tet4Args A;
    A.W = c.w4buf;
    A.tetra = c.tetra;
    A.omega = c.omega;
    A.nbe = c.nbe;
    A.nq = c.nq;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    A.i0 = i0;
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    for (int k = 0; k < ncalls; ++k) {
        const int lo = k * nq1c, n = std::min(nq1c, c.nq - lo);
        if ((rc = launch_w4_rows(c, qi, lo, n, c.w4buf))) return rc;
        A.q1_lo = lo;
        A.nq1 = n;
        A.partial = c.partial + (size_t)k * grid * c.nd;
        if (c.nd == 3)
            hipLaunchKernelGGL(k_lw4_tetra<3>, dim3(grid), dim3(T4BLK), 0, c.stream, A);
        else
            hipLaunchKernelGGL(k_lw4_tetra<6>, dim3(grid), dim3(T4BLK), 0, c.stream, A);
        SPR_HIP(hipGetLastError());
    }
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    if ((rc = launch_reduce_chunks(c, c.partial, 1, ncalls * grid, d_lw))) return rc;
    SPR_HIP(hipStreamSynchronize(c.stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess) {
        c.launches += 1;
        c.kernel_ms += ms;
    }
    return SPR_OK;
}

}  // namespace spr
