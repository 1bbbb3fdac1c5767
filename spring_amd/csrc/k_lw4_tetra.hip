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
// W4(q1, q2) comes from the FP64 store mode of k_lw4.  When every q1 row fits (a quarter of
// the free memory, at most 32 GiB), V4 is evaluated once per permutation class of (q1, q2, q3)
// and written to the place of each distinct ordering, as the reference keeps W4 once per class
// and permutes (s1, s2, s3) to read the other orders (with an index-permutation-symmetric FC4
// the numbers are those of the ordered triples); otherwise a chunk of q1 rows at a time (every q2
// of a row: the tetrahedra of a q1 touch all of them), every ordered triple evaluated.
//
// Mapping: a thread owns (q1, s1 s2 s3, tetrahedron), tetrahedron fastest across lanes, and
// walks the processes, then s0.  The vertex energies do not depend on s0, so each process's
// vertices are sorted once (stable, with the vertex's n-product and W address as payload) and
// reused for every s0; W is read only where Mq falls inside the tetrahedron's energy range.
// Neighbouring tetrahedra have alike energy ranges, so whole waves skip a process together,
// and they share vertices, so the W reads of a wave hit the same cache lines.
// Grid-stride over items; per-workgroup partials reduced in a fixed order (bitwise
// reproducible).  All arithmetic FP64.
#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#include "spr_internal.h"
#include "tetra_dev.h"

namespace spr {

namespace {

constexpr int T4BLK = 256;

struct Tet4Args {
    const double *W;            // [nq1][nq][ND^4] rows of q1 in [q1_lo, q1_lo + nq1)
    const int4 *tetra;          // [6 nq] grid indices of the four vertices
    const int *kpack;           // [nq] grid coordinates k0 | k1 << 10 | k2 << 20
    const double *omega, *nbe;  // [nq][ND]
    int nq, m0, m1, m2;
    int i0, q1_lo, nq1;
    double *partial;            // [gridDim.x][ND] of this chunk
    int rlo[4], rhi[4];         // branch ranges of s0..s3 (Polarization4th, 0-based inclusive)
};

template <int ND>
__global__ __launch_bounds__(T4BLK) void k_lw4_tetra(Tet4Args A) {
    constexpr int ND2 = ND * ND, ND3 = ND2 * ND, ND4 = ND3 * ND;
    const int tid = threadIdx.x;
    const int ntet = 6 * A.nq;
    // workgroups go to the eight XCDs round robin (workgroup b to XCD b mod 8): XCD x takes the
    // q1 rows x, x + 8, ..., so every W row is fetched into one XCD's L2 only
    const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, nlb = gridDim.x >> 3;
    const int nrows = (A.nq1 - xcd + 7) / 8;
    const int per_row = ntet * ND3;
    const int total = nrows * per_row;        // the host keeps nq1 x per_row below 2^31
    const int i0 = A.i0;
    const int kp0 = A.kpack[i0];
    // s0 from Polarization4th(1, 1), from 4 (1-based) at q0 = Gamma (Tetrahedron4th_sep.f90:213-216)
    const int s0_lo = max(i0 == 0 ? 3 : 0, A.rlo[0]), s0_hi = A.rhi[0];
    __shared__ double s_w0[ND], s_rn0[ND];
    if (tid < ND) {
        s_w0[tid] = A.omega[(size_t)i0 * ND + tid];
        s_rn0[tid] = 1.0 / A.nbe[(size_t)i0 * ND + tid];
    }
    __syncthreads();
    // range of Mq over the s0 that are summed: a process whose vertex energies miss it is skipped
    double w0lo = s_w0[s0_lo], w0hi = s_w0[s0_lo];
#pragma unroll
    for (int s = 0; s < ND; ++s)
        if (s >= s0_lo && s <= s0_hi) {
            w0lo = fmin(w0lo, s_w0[s]);
            w0hi = fmax(w0hi, s_w0[s]);
        }
    double acc[ND];
#pragma unroll
    for (int s = 0; s < ND; ++s) acc[s] = 0.0;
    for (int it = lb * T4BLK + tid; it < total; it += nlb * T4BLK) {
        // tetrahedron fastest: the lanes of a wave hold neighbouring tetrahedra of one branch set,
        // whose energy ranges are alike, so whole waves skip a process together
        const int row = it / per_row, rem = it - row * per_row;
        const int p = rem / ntet, t = rem - p * ntet;
        const int q1l = xcd + 8 * row;
        const int q1 = A.q1_lo + q1l;
        const int s1 = p / ND2, s2 = (p / ND) % ND, s3 = p % ND;
        if (s1 < A.rlo[1] || s1 > A.rhi[1] || s2 < A.rlo[2] || s2 > A.rhi[2] || s3 < A.rlo[3] || s3 > A.rhi[3])
            continue;
        // coordinates of -(q0 + q1) on the grid, before adding q2 and wrapping
        const int kp1 = A.kpack[q1];
        const int b0 = (kp0 & 1023) + (kp1 & 1023), b1 = ((kp0 >> 10) & 1023) + ((kp1 >> 10) & 1023),
                  b2 = (kp0 >> 20) + (kp1 >> 20);
        const double w1 = A.omega[(size_t)q1 * ND + s1], n1 = A.nbe[(size_t)q1 * ND + s1];
        double w2[4], w3[4], n2[4], n3[4];
        int off[4];
        const int4 tv = A.tetra[t];
        const int qv[4] = {tv.x, tv.y, tv.z, tv.w};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q2 = qv[v];
            const int kp2 = A.kpack[q2];
            // a = -(k0 + k1 + k2) mod m with 0 <= k0 + k1 + k2 <= 3 (m - 1)
            int a0 = -(b0 + (kp2 & 1023)), a1 = -(b1 + ((kp2 >> 10) & 1023)), a2 = -(b2 + (kp2 >> 20));
            a0 += a0 < 0 ? A.m0 : 0;
            a0 += a0 < 0 ? A.m0 : 0;
            a0 += a0 < 0 ? A.m0 : 0;
            a1 += a1 < 0 ? A.m1 : 0;
            a1 += a1 < 0 ? A.m1 : 0;
            a1 += a1 < 0 ? A.m1 : 0;
            a2 += a2 < 0 ? A.m2 : 0;
            a2 += a2 < 0 ? A.m2 : 0;
            a2 += a2 < 0 ? A.m2 : 0;
            const int q3 = a0 + A.m0 * (a1 + A.m1 * a2);
            w2[v] = A.omega[q2 * ND + s2];
            n2[v] = A.nbe[q2 * ND + s2];
            w3[v] = A.omega[q3 * ND + s3];
            n3[v] = A.nbe[q3 * ND + s3];
            off[v] = (q1l * A.nq + q2) * ND4 + p;
        }
        // one process at a time (fewer live registers); its vertices sorted once for all s0
#pragma unroll 1
        for (int k = 0; k < 3; ++k) {
            Vx P[4];
            double mlo = 0.0, mhi = 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const double m = k == 0 ? w1 + w2[v] + w3[v] : k == 1 ? w2[v] + w3[v] - w1 : w3[v] - w1 - w2[v];
                const double a = k == 0 ? n1 * n2[v] * n3[v]
                                        : k == 1 ? (1.0 + n1) * n2[v] * n3[v] : (1.0 + n1) * (1.0 + n2[v]) * n3[v];
                P[v] = Vx{m, a, off[v]};
                mlo = v == 0 ? m : fmin(mlo, m);
                mhi = v == 0 ? m : fmax(mhi, m);
            }
            if (mhi < w0lo || mlo > w0hi) continue;
            sort4(P);
            const double M[4] = {P[0].M, P[1].M, P[2].M, P[3].M};
            const Rcp R = rcp6(M);
            const double wk = k == 0 ? 1.0 / 6.0 : 0.5;
#pragma unroll 1
            for (int s0 = s0_lo; s0 <= s0_hi; ++s0) {
                const double mq = s_w0[s0];
                // a process contributes only for M0 <= mq <= M3 (the union of the three cases);
                // W and F are formed only then
                if (!(mq >= M[0] && mq <= M[3])) continue;
                const double rn0 = s_rn0[s0];
                const int so = s0 * ND3;
                double F[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) F[v] = A.W[P[v].off + so] * (P[v].a * rn0);
                double gk;
                const double tk = wk * tetra_rcp<true>(mq, M, R, F, gk);
#pragma unroll
                for (int s = 0; s < ND; ++s)
                    if (s == s0) acc[s] += tk;
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

// out[s] = sum over n partials[k][s] in a fixed order: one workgroup per s, strided partial sums
// per thread, then a tree (bitwise reproducible)
__global__ __launch_bounds__(256) void k_reduce_tree(const double *partial, int n, int nd, double *out) {
    const int s = blockIdx.x, tid = threadIdx.x;
    double a = 0.0;
    for (int k = tid; k < n; k += 256) a += partial[(size_t)k * nd + s];
    __shared__ double r[256];
    r[tid] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) r[tid] += r[tid + w];
        __syncthreads();
    }
    if (tid == 0) out[s] = r[0];
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
    if (c.mesh[0] > 1024 || c.mesh[1] > 1024 || c.mesh[2] > 1024) {
        set_error("four-phonon tetrahedron path: mesh dimensions above 1024");
        return SPR_E_UNSUPPORTED;
    }
    if (!c.kpack) {
        std::vector<int> kp(c.nq);
        for (int q = 0; q < c.nq; ++q)
            kp[q] = (q % c.mesh[0]) | ((q / c.mesh[0]) % c.mesh[1]) << 10 | (q / (c.mesh[0] * c.mesh[1])) << 20;
        SPR_HIP(hipMalloc((void **)&c.kpack, (size_t)c.nq * sizeof(int)));
        SPR_HIP(hipMemcpy(c.kpack, kp.data(), (size_t)c.nq * sizeof(int), hipMemcpyHostToDevice));
    }
    const size_t nd4 = (size_t)c.nd * c.nd * c.nd * c.nd;
    const size_t row = (size_t)c.nq * nd4 * sizeof(double);
    // chunk: the byte cap, W offsets below 2^31 and items (nq1 x 6 nq x Ndof^3) below 2^31
    const size_t item_row = (size_t)6 * c.nq * c.nd * c.nd * c.nd;
    const size_t lim = std::min<size_t>((((size_t)1 << 31) - 1) / (row / sizeof(double)),
                                        (((size_t)1 << 31) - 1) / item_row);
    // every row resident when it fits (a quarter of the free memory, at most 32 GiB, unless
    // SPR_W4_CHUNK_MB sets a cap): the store then evaluates V4 once per permutation class
    size_t cap = w4_chunk_bytes();
    if (!getenv("SPR_W4_CHUNK_MB")) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) cap = std::max(cap, std::min(size_t(32) << 30, fr / 4));
    }
    const bool classes = (size_t)c.nq <= cap / row && (size_t)c.nq <= lim;
    const int nq1c = classes ? c.nq : (int)std::max<size_t>(1, std::min<size_t>({(size_t)c.nq, w4_chunk_bytes() / row, lim}));
    const int ncalls = (c.nq + nq1c - 1) / nq1c;
    const long items = (long)nq1c * 6 * c.nq * c.nd * c.nd * c.nd;
    // a multiple of 8 workgroups (the XCD split of the kernel)
    const int grid = 8 * (int)std::max<long>(1, std::min<long>(c.ncu, (items + 8L * T4BLK - 1) / (8L * T4BLK)));
    if ((rc = ensure((void **)&c.w4buf, &c.w4buf_cap, (size_t)nq1c * row))) return rc;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)ncalls * grid * c.nd * sizeof(double))))
        return rc;
    const int i0 = c.h_irr_grid[qi];
    Tet4Args A;
    A.W = c.w4buf;
    A.tetra = reinterpret_cast<const int4 *>(c.tetra);
    A.kpack = c.kpack;
    A.omega = c.omega;
    A.nbe = c.nbe;
    A.nq = c.nq;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    A.i0 = i0;
    for (int k = 0; k < 4; ++k) {
        A.rlo[k] = c.pol4[2 * k];
        A.rhi[k] = c.pol4[2 * k + 1];
    }
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    for (int k = 0; k < ncalls; ++k) {
        const int lo = k * nq1c, n = std::min(nq1c, c.nq - lo);
        if ((rc = launch_w4_rows(c, qi, lo, n, c.w4buf, classes))) return rc;
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
    hipLaunchKernelGGL(k_reduce_tree, dim3(c.nd), dim3(256), 0, c.stream, c.partial, ncalls * grid, c.nd, d_lw);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipStreamSynchronize(c.stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess) {
        c.launches += 1;
        c.kernel_ms += ms;
    }
    return SPR_OK;
}

}  // namespace spr
