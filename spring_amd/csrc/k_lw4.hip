// Four-phonon scattering (Ndof <= 6), FP64 or FP32, Gaussian delta.
//
// One workgroup takes one irreducible q0 and a strided set of q1 grid points; for every q1 it
// runs over the q2 of the mesh with q3 = -(q0+q1+q2) mod G (MomentumConsrv4th,
// QuasiMomentumConsrv4th_m.f90:32-206), keeping one ordered triple per permutation class
// (q1 <= q2 <= q3), and
//   1. phase-sums FC4 into T(a,b,c,d) over the nonzero (N1,N2,N3) blocks
//      (ScatterMatEl4, FourthOrder_m.f90:286-422), integer-address phases from a table;
//   2. contracts d, c, b, a with e3, e2, e1, e0 (no conjugation);
//   3. W4 = |V4|^2 prod(1/w) Sctr4THz2/Nq^2 (ScatterProbability4 :425-473) and adds, for every
//      ordering (q1', q2', q3') of the class,
//        W4 [ 1/6 n1 n2 n3/n0 d(w0-w1-w2-w3) + 1/2 (1+n1) n2 n3/n0 d(w0+w1-w2-w3)
//           + 1/2 (1+n1)(1+n2) n3/n0 d(w0+w1+w2-w3) ]
//      the processes and weights of SumOverPolarization / FindFandOmegaVer_4th
//      (Tetrahedron4th_sep.f90:174-425) with a Gaussian in place of the tetrahedron
//      integral (seven distinct terms per class element); NaN terms are zeroed as there
//      (:238-263); s0 < 3 skipped at q0 = Gamma.  The store mode keeps every ordered triple.
// The returned sum is the raw lw of TetrahedronIntegration4th; main.f90:660 adds 0.5 lw.
//
// Store mode (WST, FP64): one workgroup per (q1, q2 chunk) writes W4 of every ordered (q1, q2)
// to W[q1 - q1_lo][q2][s0][s1][s2][s3] for the tetrahedron integration (k_lw4_tetra.hip)
// instead of the Gaussian processes; with every row resident (classes), V4 is evaluated once
// per permutation class and written to the place of each distinct ordering.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "cplx.h"
#include "spr_internal.h"

namespace spr {

struct Lw4Args {
    const int *q0_grid;
    int nq, L, m0, m1, m2;
    int chunk, nchunks;
    const int *blk_start, *tup_site;
    const void *fc;
    const int4 *site;
    int nsite;
    const void *phase, *evec;
    const double *omega, *inv_omega, *nbe;
    double sigma, factor;
    double *partial;   // [q0][chunk][ND]
    // store mode: q1 in [q1_lo, q1_lo + gridDim.x / nq2ch), q2 chunks of q2_chunk points
    double *wst;
    int q1_lo, q2_chunk, nq2ch;
    int classes;         // store mode with every row resident: one triple per permutation class,
                         // W4 written to each distinct ordering's place
    const double *sel;   // branch selection (NULL: all): (4, nd) masks of s0, s1, s2, s3
    int nterm;           // FC4 clusters (tup_site rows)
};

// SEL: a branch selection rule is set (A.sel), a compile-time flag: the runtime test costs the
// FP64 kernel about a fifth of its time (measured)
// STG: the cluster site table and the block starts are staged in LDS (when they fit), so the
// phase sum's index chain starts in LDS.
template <typename R, int ND, int BLK, bool WST, bool SEL = false, bool STG = false>
__global__ __launch_bounds__(BLK) void k_lw4(Lw4Args A) {
    using Z = cplx_t<R>;
    constexpr int NB = ND / 3;
    constexpr int ND2 = ND * ND, ND3 = ND2 * ND, ND4 = ND3 * ND;
    constexpr int PER = (ND4 + BLK - 1) / BLK;
    __shared__ Z s_a[ND4], s_b[ND4];
    __shared__ Z s_e[4][ND2];
    __shared__ double s_md[4][3][ND];   // omega, nbe, 1/omega of q0..q3
    extern __shared__ __align__(16) int s_dyn[];
    constexpr int NBLK = NB * NB * NB * NB;
    const int nm = (3 * A.nsite + 3) & ~3;
    const int nstg = STG ? ((3 * A.nterm + NBLK + 1 + 3) & ~3) : 0;
    int *s_m = s_dyn;                       // [3][nsite]
    int *s_tup = s_dyn + nm;                // STG: [nterm][3], then the NBLK + 1 block starts
    int *s_bst = s_tup + 3 * A.nterm;
    Z *s_ph = reinterpret_cast<Z *>(s_dyn + nm + nstg);

    const int tid = threadIdx.x;
    const int ch = blockIdx.x, qb = blockIdx.y;
    const int i0 = A.q0_grid[qb];
    const int k00 = i0 % A.m0, k01 = (i0 / A.m0) % A.m1, k02 = i0 / (A.m0 * A.m1);
    const bool g0 = i0 == 0;
    const Z *evec = reinterpret_cast<const Z *>(A.evec);
    const Z *phase = reinterpret_cast<const Z *>(A.phase);
    const R *fc = reinterpret_cast<const R *>(A.fc);
    for (int i = tid; i < A.L; i += BLK) s_ph[i] = phase[i];
    if constexpr (STG) {
        for (int i = tid; i < 3 * A.nterm; i += BLK) s_tup[i] = A.tup_site[i];
        for (int i = tid; i <= NBLK; i += BLK) s_bst[i] = A.blk_start[i];
    }
    const int *tup = STG ? s_tup : A.tup_site;
    const int *bst = STG ? s_bst : A.blk_start;
    for (int i = tid; i < ND2; i += BLK) s_e[0][i] = evec[(size_t)i0 * ND2 + i];
    if (tid < ND) {
        s_md[0][0][tid] = A.omega[(size_t)i0 * ND + tid];
        s_md[0][1][tid] = A.nbe[(size_t)i0 * ND + tid];
        s_md[0][2][tid] = A.inv_omega[(size_t)i0 * ND + tid];
    }
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002);
    const double hinv = 0.5 / (A.sigma * A.sigma);
    const double dcut = A.sigma * sqrt(200.0);
    double acc[ND];
#pragma unroll
    for (int s = 0; s < ND; ++s) acc[s] = 0.0;

    // store mode: the q2 chunk of one q1.  Gaussian mode: the q1 of this workgroup are
    // ch, ch + nchunks, ... (strided: the permutation classes below make the work of a q1 fall
    // with q1), every q2
    int q1_lo, q1_hi, q1_step, q2_lo, q2_hi;
    if constexpr (WST) {
        q1_lo = A.q1_lo + ch / A.nq2ch;
        q1_hi = q1_lo + 1;
        q1_step = 1;
        q2_lo = (ch % A.nq2ch) * A.q2_chunk;
        q2_hi = min(A.nq, q2_lo + A.q2_chunk);
    } else {
        q1_lo = ch;
        q1_hi = A.nq;
        q1_step = A.nchunks;
        q2_lo = 0;
        q2_hi = A.nq;
    }
    for (int q1 = q1_lo; q1 < q1_hi; q1 += q1_step) {
        const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
        for (int q2 = q2_lo; q2 < q2_hi; ++q2) {
            const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
            int a0 = -(k00 + k10 + k20) % A.m0, a1 = -(k01 + k11 + k21) % A.m1, a2 = -(k02 + k12 + k22) % A.m2;
            a0 += a0 < 0 ? A.m0 : 0;
            a1 += a1 < 0 ? A.m1 : 0;
            a2 += a2 < 0 ? A.m2 : 0;
            const int q3 = a0 + A.m0 * (a1 + A.m1 * a2);
            // Gaussian mode: one ordered triple per permutation class of (q1, q2, q3), the one with
            // q1 <= q2 <= q3 (QuasiMomentumConsrv4th_m.f90:224-312 stores one per class); its V4
            // serves every ordering of the class (V4 is symmetric under permutations of legs 1-3
            // with their branches), which the process terms below sum
            if (!WST || A.classes)
                if (!(q1 <= q2 && q2 <= q3)) continue;
            const int qs[3] = {q1, q2, q3};
            __syncthreads();
            for (int i = tid; i < 3 * ND2; i += BLK) {
                const int k = i / ND2, e = i % ND2;
                s_e[k + 1][e] = evec[(size_t)qs[k] * ND2 + e];
            }
            if (tid < 3 * ND) {
                const int k = tid / ND, s = tid % ND;
                s_md[k + 1][0][s] = A.omega[(size_t)qs[k] * ND + s];
                s_md[k + 1][1][s] = A.nbe[(size_t)qs[k] * ND + s];
                s_md[k + 1][2][s] = A.inv_omega[(size_t)qs[k] * ND + s];
            }
            for (int s = tid; s < A.nsite; s += BLK) {
                const int4 c = A.site[s];
                s_m[s] = (k10 * c.x + k11 * c.y + k12 * c.z) % A.L;
                s_m[A.nsite + s] = (k20 * c.x + k21 * c.y + k22 * c.z) % A.L;
                s_m[2 * A.nsite + s] = (a0 * c.x + a1 * c.y + a2 * c.z) % A.L;
            }
            __syncthreads();
            // 1. phase sum into s_a[a][b][c][d]
            for (int e = tid; e < ND4; e += BLK) {
                const int a = e / ND3, b = (e / ND2) % ND, c = (e / ND) % ND, d = e % ND;
                const int blk = ((a / 3 * NB + b / 3) * NB + c / 3) * NB + d / 3;
                const int comp = (a % 3) * 27 + (b % 3) * 9 + (c % 3) * 3 + (d % 3);
                Z t = cmk<R>(0, 0);
                const int u1 = bst[blk + 1];
                for (int u = bst[blk]; u < u1; ++u) {
                    // each site phase index is in [0, L), so the sum is in [0, 3L)
                    int m = s_m[tup[3 * u]] + s_m[A.nsite + tup[3 * u + 1]] + s_m[2 * A.nsite + tup[3 * u + 2]];
                    m -= m >= A.L ? A.L : 0;
                    m -= m >= A.L ? A.L : 0;
                    t = crfma(fc[(size_t)u * 81 + comp], s_ph[m], t);
                }
                s_a[e] = t;
            }
            // 2. four mode products; each pass contracts the last index and moves the new
            //    mode index to the front: (x,y,z,c) -> (s,x,y,z)
            Z *src = s_a, *dst = s_b;
#pragma unroll
            for (int pass = 0; pass < 4; ++pass) {
                __syncthreads();
                const Z *E = s_e[3 - pass];   // e3, e2, e1, e0
                // a thread per row xyz: the row is read once and gives all ND outputs (each summed
                // over c in order)
                for (int xyz = tid; xyz < ND3; xyz += BLK) {
                    Z a[ND];
#pragma unroll
                    for (int c = 0; c < ND; ++c) a[c] = src[xyz * ND + c];
#pragma unroll
                    for (int s = 0; s < ND; ++s) {
                        Z z = cmk<R>(0, 0);
#pragma unroll
                        for (int c = 0; c < ND; ++c) z = cfma(a[c], E[s * ND + c], z);
                        dst[s * ND3 + xyz] = z;
                    }
                }
                Z *t = src;
                src = dst;
                dst = t;
            }
            __syncthreads();
            if constexpr (WST) {
                // 3'. W4 only (every s0; the Gamma rows are zero through 1/omega = 0)
                if (A.classes) {
                    // every distinct ordering (legs i, j, l in places 1, 2, 3) of the class:
                    // W4[q_i][q_j][s0][s_i][s_j][s_l] = W4 of the canonical element
                    static constexpr int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
                    const int qq[3] = {q1, q2, q3};
                    bool use[6];
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        use[k] = true;
                        for (int k2 = 0; k2 < k; ++k2)
                            if (qq[perm[k][0]] == qq[perm[k2][0]] && qq[perm[k][1]] == qq[perm[k2][1]]) use[k] = false;
                    }
                    for (int e = tid; e < ND4; e += BLK) {
                        const int s0 = e / ND3, sl[3] = {(e / ND2) % ND, (e / ND) % ND, e % ND};
                        const double W = (double)cabs2(src[e]) * (s_md[0][2][s0] * s_md[1][2][sl[0]] *
                                                                  s_md[2][2][sl[1]] * s_md[3][2][sl[2]]) * A.factor;
#pragma unroll
                        for (int k = 0; k < 6; ++k) {
                            if (!use[k]) continue;
                            const int a = perm[k][0], b = perm[k][1], d = perm[k][2];
                            A.wst[((size_t)qq[a] * A.nq + qq[b]) * ND4 + ((s0 * ND + sl[a]) * ND + sl[b]) * ND + sl[d]] = W;
                        }
                    }
                    continue;
                }
                double *wrow = A.wst + ((size_t)(q1 - A.q1_lo) * A.nq + q2) * ND4;
                for (int e = tid; e < ND4; e += BLK) {
                    const int s0 = e / ND3, s1 = (e / ND2) % ND, s2 = (e / ND) % ND, s3 = e % ND;
                    wrow[e] = (double)cabs2(src[e]) * (s_md[0][2][s0] * s_md[1][2][s1] * s_md[2][2][s2] *
                                                      s_md[3][2][s3]) * A.factor;
                }
                continue;
            }
            // 3. W4 and the processes of every ordering of the class; src holds V[s0][s1][s2][s3].
            //    An ordering puts legs (j, k, l) of the triple in places (1, 2, 3) and adds
            //      W [ 1/6 F1 d(w0-w1-w2-w3) + 1/2 (1+n_j) n_k n_l/n0 d(w0+w_j-w_k-w_l)
            //          + 1/2 (1+n_j)(1+n_k) n_l/n0 d(w0+w_j+w_k-w_l) ]
            //    with the branch masks m(j,k,l) = S1[s_j] S2[s_k] S3[s_l]; over the six orderings
            //    each leg is first in two and last in two, and a class of c distinct orderings
            //    is met 6 / c times among the six, hence the factor c / 6 = 1 / |stabiliser|.
            const double istab = q1 == q3 ? 1.0 / 6.0 : (q1 == q2 || q2 == q3) ? 0.5 : 1.0;
#pragma unroll
            for (int j = 0; j < PER; ++j) {
                const int e = tid + j * BLK;
                if (e >= ND4) break;
                const int s0 = e / ND3, s1 = (e / ND2) % ND, s2 = (e / ND) % ND, s3 = e % ND;
                if (g0 && s0 < 3) continue;
                const double w0 = s_md[0][0][s0], n0 = s_md[0][1][s0];
                const double w1 = s_md[1][0][s1], n1 = s_md[1][1][s1];
                const double w2 = s_md[2][0][s2], n2 = s_md[2][1][s2];
                const double w3 = s_md[3][0][s3], n3 = s_md[3][1][s3];
                const double W = (double)cabs2(src[e]) * (s_md[0][2][s0] * s_md[1][2][s1] * s_md[2][2][s2] *
                                                          s_md[3][2][s3]) * A.factor;
                double m1 = 1.0, m2a = 1.0, m2b = 1.0, m2c = 1.0, m3a = 1.0, m3b = 1.0, m3c = 1.0;
                if constexpr (SEL) {
                    const double *S1 = A.sel + ND, *S2 = A.sel + 2 * ND, *S3 = A.sel + 3 * ND;
                    const double p123 = S1[s1] * S2[s2] * S3[s3], p132 = S1[s1] * S2[s3] * S3[s2];
                    const double p213 = S1[s2] * S2[s1] * S3[s3], p231 = S1[s2] * S2[s3] * S3[s1];
                    const double p312 = S1[s3] * S2[s1] * S3[s2], p321 = S1[s3] * S2[s2] * S3[s1];
                    m1 = (p123 + p132 + p213 + p231 + p312 + p321) / 6.0;
                    m2a = 0.5 * (p123 + p132);   // leg 1 first
                    m2b = 0.5 * (p213 + p231);   // leg 2 first
                    m2c = 0.5 * (p312 + p321);   // leg 3 first
                    m3c = 0.5 * (p123 + p213);   // leg 3 last
                    m3b = 0.5 * (p132 + p312);   // leg 2 last
                    m3a = 0.5 * (p231 + p321);   // leg 1 last
                }
                // a Gaussian below e^-100 of its peak (|x| >= sigma sqrt(200)) adds nothing
                // measurable and is not evaluated
                auto gs = [&](double x) { return fabs(x) < dcut ? c0 * exp(-hinv * x * x) : 0.0; };
                auto fin = [](double v) { return v != v ? 0.0 : v; };
                const double r0 = 1.0 / n0;
                double t = fin(m1 * W * n1 * n2 * n3 * r0 * gs(w0 - w1 - w2 - w3));
                t += fin(m2a * W * (1.0 + n1) * n2 * n3 * r0 * gs(w0 + w1 - w2 - w3));
                t += fin(m2b * W * (1.0 + n2) * n1 * n3 * r0 * gs(w0 + w2 - w1 - w3));
                t += fin(m2c * W * (1.0 + n3) * n1 * n2 * r0 * gs(w0 + w3 - w1 - w2));
                t += fin(m3c * W * (1.0 + n1) * (1.0 + n2) * n3 * r0 * gs(w0 + w1 + w2 - w3));
                t += fin(m3b * W * (1.0 + n1) * (1.0 + n3) * n2 * r0 * gs(w0 + w1 + w3 - w2));
                t += fin(m3a * W * (1.0 + n2) * (1.0 + n3) * n1 * r0 * gs(w0 + w2 + w3 - w1));
                t *= istab;
                if constexpr (SEL) t *= A.sel[s0];
#pragma unroll
                for (int s = 0; s < ND; ++s)
                    if (s == s0) acc[s] += t;
            }
        }
    }
    if constexpr (WST) return;
    __shared__ double s_red[BLK];
    for (int s = 0; s < ND; ++s) {
        __syncthreads();
        s_red[tid] = acc[s];
        __syncthreads();
        for (int w = BLK / 2; w > 0; w >>= 1) {
            if (tid < w) s_red[tid] += s_red[tid + w];
            __syncthreads();
        }
        if (tid == 0) A.partial[((size_t)qb * A.nchunks + ch) * ND + s] = s_red[0];
    }
}

__global__ void k_reduce_chunks(const double *partial, int nq0, int nchunks, int nd, double *lw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq0 * nd) return;
    const int q = i / nd, s = i % nd;
    double acc = 0.0;
    for (int c = 0; c < nchunks; ++c) acc += partial[((size_t)q * nchunks + c) * nd + s];
    lw[i] = acc;
}

// Sctr4THz2 (constants.f90:62-63) / Nq^2
// ints of the staged cluster site table and block starts (0: they stay in global memory, when
// more than 16 KB)
template <int ND>
static size_t lw4_stage_ints(int nterm) {
    constexpr int NB = ND / 3;
    const size_t n = ((size_t)3 * nterm + NB * NB * NB * NB + 1 + 3) & ~(size_t)3;
    return n * sizeof(int) <= 16384 ? n : 0;
}

double lw4_factor(const Ctx &c) {
    const double pi = 4.0 * std::atan(1.0);
    return 0.125 * pi * std::pow(1.054571817 * 1.602176565, 2) * std::pow(6.022140857, 4) * 1000000.0 /
           ((double)c.nq * (double)c.nq);
}

int launch_reduce_chunks(Ctx &c, const double *partial, int q0_count, int nchunks, double *d_lw) {
    hipLaunchKernelGGL(k_reduce_chunks, dim3((q0_count * c.nd + 255) / 256), dim3(256), 0, c.stream, partial,
                       q0_count, nchunks, c.nd, d_lw);
    SPR_HIP(hipGetLastError());
    return SPR_OK;
}

template <typename R, int ND>
static int launch_lw4_t(Ctx &c, int q0_first, int q0_count, double sigma, double *d_lw) {
    constexpr int BLK = 256;
    Lw4Args A;
    A.q0_grid = c.irr_grid + q0_first;
    A.nq = c.nq;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    const long target = 2048;
    int nch = (int)std::min<long>(c.nq, std::max<long>(1, (target + q0_count - 1) / q0_count));
    A.chunk = (c.nq + nch - 1) / nch;
    A.nchunks = (c.nq + A.chunk - 1) / A.chunk;
    A.blk_start = c.fc4.blk_start;
    A.tup_site = c.fc4.tup_site;
    A.fc = sizeof(R) == 8 ? (const void *)c.fc4.fc64 : (const void *)c.fc4.fc32;
    A.site = c.fc4.site;
    A.nsite = c.fc4.nsite;
    A.phase = sizeof(R) == 8 ? (const void *)c.phase64 : (const void *)c.phase32;
    A.evec = sizeof(R) == 8 ? (const void *)c.evec : (const void *)c.evec32;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sigma = sigma;
    A.factor = lw4_factor(c);
    int rc;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)q0_count * A.nchunks * ND * sizeof(double))))
        return rc;
    A.partial = c.partial;
    A.sel = c.sel4 ? c.d_sel4 : nullptr;
    A.nterm = c.fc4.nterm;
    const size_t nstg = lw4_stage_ints<ND>(A.nterm);
    const size_t dyn = ((size_t)((3 * A.nsite + 3) & ~3) + nstg) * sizeof(int) + (size_t)c.L * sizeof(cplx_t<R>);
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    A.wst = nullptr;
    A.q1_lo = A.q2_chunk = A.nq2ch = 0;
    A.classes = 0;
    const dim3 grid(A.nchunks, q0_count);
    if (A.sel) {
        if (nstg) hipLaunchKernelGGL((k_lw4<R, ND, BLK, false, true, true>), grid, dim3(BLK), dyn, c.stream, A);
        else hipLaunchKernelGGL((k_lw4<R, ND, BLK, false, true, false>), grid, dim3(BLK), dyn, c.stream, A);
    } else {
        if (nstg) hipLaunchKernelGGL((k_lw4<R, ND, BLK, false, false, true>), grid, dim3(BLK), dyn, c.stream, A);
        else hipLaunchKernelGGL((k_lw4<R, ND, BLK, false, false, false>), grid, dim3(BLK), dyn, c.stream, A);
    }
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    if ((rc = launch_reduce_chunks(c, c.partial, q0_count, A.nchunks, d_lw))) return rc;
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

template <int ND>
static int launch_w4_rows_t(Ctx &c, int qi, int q1_lo, int nq1, double *d_w, bool classes) {
    constexpr int BLK = 256;
    Lw4Args A{};
    A.q0_grid = c.irr_grid + qi;
    A.nq = c.nq;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    A.blk_start = c.fc4.blk_start;
    A.tup_site = c.fc4.tup_site;
    A.fc = c.fc4.fc64;
    A.site = c.fc4.site;
    A.nsite = c.fc4.nsite;
    A.phase = c.phase64;
    A.evec = c.evec;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.factor = lw4_factor(c);
    A.wst = d_w;
    A.q1_lo = q1_lo;
    A.classes = classes ? 1 : 0;
    // enough workgroups to fill the device: split the q2 range of every q1
    const long target = 4L * c.ncu;
    A.nq2ch = (int)std::min<long>(c.nq, std::max<long>(1, (target + nq1 - 1) / nq1));
    A.q2_chunk = (c.nq + A.nq2ch - 1) / A.nq2ch;
    A.nq2ch = (c.nq + A.q2_chunk - 1) / A.q2_chunk;
    A.nterm = c.fc4.nterm;
    const size_t nstg = lw4_stage_ints<ND>(A.nterm);
    const size_t dyn = ((size_t)((3 * A.nsite + 3) & ~3) + nstg) * sizeof(int) + (size_t)c.L * sizeof(cplx_t<double>);
    if (nstg)
        hipLaunchKernelGGL((k_lw4<double, ND, BLK, true, false, true>), dim3(nq1 * A.nq2ch, 1), dim3(BLK), dyn, c.stream, A);
    else
        hipLaunchKernelGGL((k_lw4<double, ND, BLK, true, false, false>), dim3(nq1 * A.nq2ch, 1), dim3(BLK), dyn, c.stream, A);
    SPR_HIP(hipGetLastError());
    return SPR_OK;
}

int launch_w4_rows(Ctx &c, int qi, int q1_lo, int nq1, double *d_w, bool classes) {
    if (classes && (q1_lo != 0 || nq1 != c.nq)) {
        set_error("four-phonon W4 store: the class mode needs every row resident");
        return SPR_E_ARG;
    }
    switch (c.nd) {
        case 3: return launch_w4_rows_t<3>(c, qi, q1_lo, nq1, d_w, classes);
        case 6: return launch_w4_rows_t<6>(c, qi, q1_lo, nq1, d_w, classes);
        default:
            set_error("four-phonon path: Ndof " + std::to_string(c.nd) + " not built (3, 6)");
            return SPR_E_UNSUPPORTED;
    }
}

int launch_lw4(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw) {
    if (prec == SPR_PREC_BF16) {
        set_error("four-phonon path: BF16 mode is not built (FP64 and FP32 are)");
        return SPR_E_UNSUPPORTED;
    }
    const bool f64 = prec == SPR_PREC_FP64;
    // FP32: the matrix-core kernel where its shape is built and the item fits in LDS
    // (SPR_KERNEL=generic keeps the vector kernel)
    if (!f64) {
        const char *f = getenv("SPR_KERNEL");
        if (!(f && std::string(f) == "generic")) {
            const int rc = launch_lw4_mm(c, q0_first, q0_count, sigma, d_lw);
            if (rc != SPR_E_UNSUPPORTED) return rc;
        }
    }
    switch (c.nd) {
        case 3: return f64 ? launch_lw4_t<double, 3>(c, q0_first, q0_count, sigma, d_lw)
                           : launch_lw4_t<float, 3>(c, q0_first, q0_count, sigma, d_lw);
        case 6: return f64 ? launch_lw4_t<double, 6>(c, q0_first, q0_count, sigma, d_lw)
                           : launch_lw4_t<float, 6>(c, q0_first, q0_count, sigma, d_lw);
        default:
            set_error("four-phonon path: Ndof " + std::to_string(c.nd) + " not built (3, 6)");
            return SPR_E_UNSUPPORTED;
    }
}

}  // namespace spr
