// Iterative solution of the phonon Boltzmann transport equation (kappa_iter.x,
// src/Kappa/Src_iterative/main.f90:243-361), on the set of q points irreducible under time
// reversal only, as the reference builds it (Irr_q_point_m.f90:206-330): q and -q share one
// F, F(-q) = -F(q).
//
//   W(q, q1, s, s1, s2) = |V3|^2 Sctr3THz2/Nq / (w w1 w2) of every grid q1 with q2 = -q-q1
//     (QconsrvMomentum / ScatterMatEl / ScatterProbability, Quasi_momentum_consrv.f90:118-622),
//     computed once by the three-phonon W path and kept in HBM;
//   W+ = W (n1 - n2), W- = W (1 + n1 + n2) of the pair (q1, q2);
//   1/tau(q,s) = sum_T 1/6 g+ sum_v I_v W+_v + 1/2 sum_T 1/6 g- sum_v I_v W-_v
//     (TetrahedronWplus / TetrahedronWminus, Quasi_momentum_consrv.f90:701-1208);
//   dF(q,s) = [sum_T 1/6 g+ sum_v I_v W+_v (F(q + q1_v, s2) - F(q1_v, s1))
//            + 1/2 sum_T 1/6 g- sum_v I_v W-_v (F(q - q1_v, s2) + F(q1_v, s1))] tau(q,s)
//     (TetrahedronIterPlus / TetrahedronIterMinus, TetrahedronDeltaSum.f90:22-620, then
//      multiplyTaudFnew, HelperSubroutines.f90:188-245).
// For a tetrahedron vertex q1_v the + process is q + q1_v -> q + q1_v with W+ of the pair
// (q1_v, -q-q1_v) and M+ = w(q+q1_v, s2) - w(q1_v, s1); the - process is q -> q1_v + (q - q1_v)
// with W- of the pair (-q1_v, q1_v - q) and M- = w(q1_v, s1) + w(q - q1_v, s2).  The linear
// tetrahedron weights are those of the iterative code: strict inequalities in the three cases
// and a vertex weight sum within 5e-5 of one, else the tetrahedron is skipped.
//
// Mapping: one workgroup per (tetrahedron chunk, s, irreducible q), a thread per
// (tetrahedron, s1, s2) walking its chunk; W of (q, q1, s) is contiguous in (s1, s2), so W
// reads of one vertex are coalesced; per-workgroup partial sums are reduced in a fixed order
// by a second kernel (bitwise reproducible).  FP64 throughout.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "spr_internal.h"

namespace spr {

namespace {

constexpr int kBteBlock = 256;
constexpr double kZeroPrec = 5.0e-5;   // zero_prec (Src_iterative/constants.f90:29)

struct BteArgs {
    const double *W;        // (nirr, nq, nd, nd, nd)
    const int *tetra;       // (6 nq, 4)
    const int4 *coord;      // (nq) mesh coordinates
    const int *irr_grid;    // (nirr)
    const int2 *map;        // (nq): irreducible index and sign of F
    const double *omega, *nbe;
    const double *F;        // (nirr, nd, 3); MODE 1 only
    int nq, nd, m0, m1, m2, ntet, chunk, nch;
    double scale;           // W of the solver / W of the linewidth path
    double *partial;        // (nirr, nd, nch, K)
};

__device__ __forceinline__ int grid_index(int a0, int a1, int a2, const BteArgs &A) {
    a0 %= A.m0;
    a1 %= A.m1;
    a2 %= A.m2;
    a0 += a0 < 0 ? A.m0 : 0;
    a1 += a1 < 0 ? A.m1 : 0;
    a2 += a2 < 0 ? A.m2 : 0;
    return a0 + A.m0 * (a1 + A.m1 * a2);
}

// Linear tetrahedron weights of the iterative code for one process: the four vertex values M
// are sorted (stable: idx[k] is the original vertex of sorted position k), then g and the vertex
// weights I (sorted order) of the case that holds Mqs strictly inside; false when no case holds
// or the I_v do not sum to one within kZeroPrec (the tetrahedron is skipped then).
__device__ __forceinline__ bool tetra_gI(double mq, const double *Min, int *idx, double &g, double *I) {
#pragma unroll
    for (int v = 0; v < 4; ++v) idx[v] = v;
#pragma unroll
    for (int a = 1; a < 4; ++a)
#pragma unroll
        for (int b = a; b > 0; --b)
            if (Min[idx[b - 1]] > Min[idx[b]]) {
                const int t = idx[b];
                idx[b] = idx[b - 1];
                idx[b - 1] = t;
            }
    double M[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) M[v] = Min[idx[v]];
    if (mq > M[0] && mq < M[1]) {
        g = 3.0 * (mq - M[0]) * (mq - M[0]) / ((M[1] - M[0]) * (M[2] - M[0]) * (M[3] - M[0]));
        I[0] = ((mq - M[1]) / (M[0] - M[1]) + (mq - M[2]) / (M[0] - M[2]) + (mq - M[3]) / (M[0] - M[3])) / 3.0;
        I[1] = ((mq - M[0]) / (M[1] - M[0])) / 3.0;
        I[2] = ((mq - M[0]) / (M[2] - M[0])) / 3.0;
        I[3] = ((mq - M[0]) / (M[3] - M[0])) / 3.0;
    } else if (mq > M[1] && mq < M[2]) {
        g = (3.0 / ((M[1] - M[2]) * (M[3] - M[0]))) *
            (((mq - M[0]) * (mq - M[2])) / (M[2] - M[0]) - ((mq - M[1]) * (mq - M[3])) / (M[1] - M[3]));
        const double gd = g * (M[3] - M[0]);
        I[0] = ((mq - M[3]) / (M[0] - M[3])) / 3.0 +
               ((mq - M[2]) / (M[0] - M[2])) * ((mq - M[0]) / (M[2] - M[0])) * ((mq - M[2]) / (M[1] - M[2])) / gd;
        I[1] = ((mq - M[2]) / (M[1] - M[2])) / 3.0 +
               ((mq - M[3]) / (M[1] - M[3])) * ((mq - M[3]) / (M[1] - M[3])) * ((mq - M[1]) / (M[2] - M[1])) / gd;
        I[2] = ((mq - M[1]) / (M[2] - M[1])) / 3.0 +
               ((mq - M[0]) / (M[2] - M[0])) * ((mq - M[0]) / (M[2] - M[0])) * ((mq - M[2]) / (M[1] - M[2])) / gd;
        I[3] = ((mq - M[0]) / (M[3] - M[0])) / 3.0 +
               ((mq - M[1]) / (M[3] - M[1])) * ((mq - M[3]) / (M[1] - M[3])) * ((mq - M[1]) / (M[2] - M[1])) / gd;
    } else if (mq > M[2] && mq < M[3]) {
        g = 3.0 * (mq - M[3]) * (mq - M[3]) / ((M[3] - M[0]) * (M[3] - M[1]) * (M[3] - M[2]));
        I[0] = ((mq - M[3]) / (M[0] - M[3])) / 3.0;
        I[1] = ((mq - M[3]) / (M[1] - M[3])) / 3.0;
        I[2] = ((mq - M[3]) / (M[2] - M[3])) / 3.0;
        I[3] = ((mq - M[0]) / (M[3] - M[0]) + (mq - M[1]) / (M[3] - M[1]) + (mq - M[2]) / (M[3] - M[2])) / 3.0;
    } else {
        return false;
    }
    return fabs(I[0] + I[1] + I[2] + I[3] - 1.0) < kZeroPrec;
}

// acc[k] += mul g/6 sum_v I_v X_v[k] for one process (X in the original vertex order)
template <int K>
__device__ __forceinline__ void tetra_acc(double mq, const double *Min, const double (*Xin)[K], double mul,
                                         double *acc) {
    int idx[4];
    double g, I[4];
    if (!tetra_gI(mq, Min, idx, g, I)) return;
    const double f = mul * g / 6.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) s += I[v] * Xin[idx[v]][k];
        acc[k] += f * s;
    }
}

// MODE 0: 1/tau (K = 1); MODE 1: dF before the division by 1/tau (K = 3)
template <int MODE>
__global__ __launch_bounds__(kBteBlock) void k_bte(BteArgs A) {
    constexpr int K = MODE == 0 ? 1 : 3;
    const int ch = blockIdx.x, s = blockIdx.y, qi = blockIdx.z;
    const int nd = A.nd, nd2 = nd * nd;
    const int q = A.irr_grid[qi];
    const int4 kq = A.coord[q];
    const double mq = A.omega[(size_t)q * nd + s];
    const double *Wq = A.W + (size_t)qi * A.nq * nd * nd2;
    const int t_lo = ch * A.chunk, t_hi = min(A.ntet, t_lo + A.chunk);
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    // each thread owns tetrahedra t_lo + tid + kBteBlock k of the chunk and walks every branch
    // pair of each, the vertex indices formed once per tetrahedron; at every step the lanes of a
    // wave hold neighbouring tetrahedra (alike energy ranges) of one branch pair
    for (int t = t_lo + (int)threadIdx.x; t < t_hi; t += kBteBlock) {
        int q1v[4], qpv[4], qmv[4], nq1v[4], qq2v[4], mq2v[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q1 = A.tetra[4 * t + v];
            const int4 k1 = A.coord[q1];
            q1v[v] = q1;
            qpv[v] = grid_index(kq.x + k1.x, kq.y + k1.y, kq.z + k1.z, A);        // q + q1
            qmv[v] = grid_index(kq.x - k1.x, kq.y - k1.y, kq.z - k1.z, A);        // q - q1
            qq2v[v] = grid_index(-kq.x - k1.x, -kq.y - k1.y, -kq.z - k1.z, A);   // -q - q1
            nq1v[v] = grid_index(-k1.x, -k1.y, -k1.z, A);                        // -q1
            mq2v[v] = grid_index(k1.x - kq.x, k1.y - kq.y, k1.z - kq.z, A);       // q1 - q
        }
        for (int s1 = 0; s1 < nd; ++s1) {
            double w1[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) w1[v] = A.omega[(size_t)q1v[v] * nd + s1];
            for (int s2 = 0; s2 < nd; ++s2) {
                const int p = s1 * nd + s2;
                double Mp[4], Mm[4], Xp[4][K], Xm[4][K];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Mp[v] = A.omega[(size_t)qpv[v] * nd + s2] - w1[v];
                    Mm[v] = w1[v] + A.omega[(size_t)qmv[v] * nd + s2];
                }
                // both processes need mq strictly inside their vertex-energy range (tetra_gI's
                // cases); outside it they add nothing, so W, n and F are not read
                {
                    const double lo_p = fmin(fmin(Mp[0], Mp[1]), fmin(Mp[2], Mp[3]));
                    const double hi_p = fmax(fmax(Mp[0], Mp[1]), fmax(Mp[2], Mp[3]));
                    const double lo_m = fmin(fmin(Mm[0], Mm[1]), fmin(Mm[2], Mm[3]));
                    const double hi_m = fmax(fmax(Mm[0], Mm[1]), fmax(Mm[2], Mm[3]));
                    if (!((mq > lo_p && mq < hi_p) || (mq > lo_m && mq < hi_m))) continue;
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // + process, pair (q1, -q-q1)
                    const double wp = Wq[((size_t)q1v[v] * nd + s) * nd2 + p] * A.scale *
                                      (A.nbe[(size_t)q1v[v] * nd + s1] - A.nbe[(size_t)qq2v[v] * nd + s2]);
                    // - process, pair (-q1, q1 - q)
                    const double wm = Wq[((size_t)nq1v[v] * nd + s) * nd2 + p] * A.scale *
                                      (1.0 + A.nbe[(size_t)nq1v[v] * nd + s1] + A.nbe[(size_t)mq2v[v] * nd + s2]);
                    if (MODE == 0) {
                        Xp[v][0] = wp;
                        Xm[v][0] = wm;
                    } else {
                        const int2 m1 = A.map[q1v[v]], mp = A.map[qpv[v]], mm = A.map[qmv[v]];
                        const double *F1 = A.F + ((size_t)m1.x * nd + s1) * 3;
                        const double *Fp = A.F + ((size_t)mp.x * nd + s2) * 3;
                        const double *Fm = A.F + ((size_t)mm.x * nd + s2) * 3;
#pragma unroll
                        for (int k = 0; k < K; ++k) {
                            const double f1 = m1.y * F1[k];
                            Xp[v][k] = wp * (mp.y * Fp[k] - f1);
                            Xm[v][k] = wm * (mm.y * Fm[k] + f1);
                        }
                    }
                }
                tetra_acc<K>(mq, Mp, Xp, 1.0, acc);
                tetra_acc<K>(mq, Mm, Xm, 0.5, acc);
            }
        }
    }
    __shared__ double s_red[K][kBteBlock];
#pragma unroll
    for (int k = 0; k < K; ++k) s_red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = kBteBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int k = 0; k < K; ++k) s_red[k][threadIdx.x] += s_red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < K)
        A.partial[(((size_t)qi * nd + s) * A.nch + ch) * K + threadIdx.x] = s_red[threadIdx.x][0];
}

// Vertex coefficients of one process for the collision matrix: coef[v] = mul g I_v W_v / 6 in
// the original vertex order; all zero when the tetrahedron is skipped (weights of tetra_gI).
__device__ __forceinline__ void tetra_coef(double mq, const double *Min, const double *Win, double mul,
                                           double *coef) {
    int idx[4];
    double g, I[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) coef[v] = 0.0;
    if (!tetra_gI(mq, Min, idx, g, I)) return;
    const double f = mul * g / 6.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) coef[idx[v]] = f * I[v] * Win[idx[v]];
}

// Collision matrix row r = (qi, s) of dF = A F: the + process adds c to column (q + q1, s2) and
// -c to (q1, s1), the - process adds c to (q - q1, s2) and to (q1, s1), each with the sign of
// the F map (F(-q) = -F(q)); the row is accumulated in LDS.  Also 1/tau(qi, s) = sum of c.
__global__ __launch_bounds__(kBteBlock) void k_bte_build(BteArgs A, int R, double *Amat, double *itau) {
    extern __shared__ double s_row[];
    __shared__ double s_red[kBteBlock];
    const int s = blockIdx.x, qi = blockIdx.y;
    const int nd = A.nd, nd2 = nd * nd;
    const int q = A.irr_grid[qi];
    const int4 kq = A.coord[q];
    const double mq = A.omega[(size_t)q * nd + s];
    const double *Wq = A.W + (size_t)qi * A.nq * nd * nd2;
    for (int i = threadIdx.x; i < R; i += kBteBlock) s_row[i] = 0.0;
    __syncthreads();
    double tau = 0.0;
    // each thread owns tetrahedra t = tid + kBteBlock k and walks every branch pair of each, so
    // the vertex indices are formed once per tetrahedron; at every step the lanes of a wave hold
    // neighbouring tetrahedra of one branch pair, whose energy ranges are alike, so whole waves
    // skip together below
    for (int t = threadIdx.x; t < A.ntet; t += kBteBlock) {
        int q1v[4], qpv[4], qmv[4], nq1v[4], qq2v[4], mq2v[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int q1 = A.tetra[4 * t + v];
            const int4 k1 = A.coord[q1];
            q1v[v] = q1;
            qpv[v] = grid_index(kq.x + k1.x, kq.y + k1.y, kq.z + k1.z, A);
            qmv[v] = grid_index(kq.x - k1.x, kq.y - k1.y, kq.z - k1.z, A);
            qq2v[v] = grid_index(-kq.x - k1.x, -kq.y - k1.y, -kq.z - k1.z, A);
            nq1v[v] = grid_index(-k1.x, -k1.y, -k1.z, A);
            mq2v[v] = grid_index(k1.x - kq.x, k1.y - kq.y, k1.z - kq.z, A);
        }
        for (int s1 = 0; s1 < nd; ++s1) {
            double w1[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) w1[v] = A.omega[(size_t)q1v[v] * nd + s1];
            for (int s2 = 0; s2 < nd; ++s2) {
                const int p = s1 * nd + s2;
                double Mp[4], Mm[4], Wp[4], Wm[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Mp[v] = A.omega[(size_t)qpv[v] * nd + s2] - w1[v];
                    Mm[v] = w1[v] + A.omega[(size_t)qmv[v] * nd + s2];
                }
                // both processes need Mq strictly inside their vertex-energy range (tetra_gI's
                // cases); outside it they add nothing, so W and the occupations are not read
                const double lo_p = fmin(fmin(Mp[0], Mp[1]), fmin(Mp[2], Mp[3]));
                const double hi_p = fmax(fmax(Mp[0], Mp[1]), fmax(Mp[2], Mp[3]));
                const double lo_m = fmin(fmin(Mm[0], Mm[1]), fmin(Mm[2], Mm[3]));
                const double hi_m = fmax(fmax(Mm[0], Mm[1]), fmax(Mm[2], Mm[3]));
                if (!((mq > lo_p && mq < hi_p) || (mq > lo_m && mq < hi_m))) continue;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    Wp[v] = Wq[((size_t)q1v[v] * nd + s) * nd2 + p] * A.scale *
                            (A.nbe[(size_t)q1v[v] * nd + s1] - A.nbe[(size_t)qq2v[v] * nd + s2]);
                    Wm[v] = Wq[((size_t)nq1v[v] * nd + s) * nd2 + p] * A.scale *
                            (1.0 + A.nbe[(size_t)nq1v[v] * nd + s1] + A.nbe[(size_t)mq2v[v] * nd + s2]);
                }
                double kp[4], km[4];
                tetra_coef(mq, Mp, Wp, 1.0, kp);
                tetra_coef(mq, Mm, Wm, 0.5, km);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    tau += kp[v] + km[v];
                    if (kp[v] != 0.0 || km[v] != 0.0) {
                        const int2 m1 = A.map[q1v[v]];
                        const int c1 = m1.x * nd + s1;
                        if (kp[v] != 0.0) {
                            const int2 mp = A.map[qpv[v]];
                            atomicAdd(&s_row[mp.x * nd + s2], mp.y * kp[v]);
                            atomicAdd(&s_row[c1], -m1.y * kp[v]);
                        }
                        if (km[v] != 0.0) {
                            const int2 mm = A.map[qmv[v]];
                            atomicAdd(&s_row[mm.x * nd + s2], mm.y * km[v]);
                            atomicAdd(&s_row[c1], m1.y * km[v]);
                        }
                    }
                }
            }
        }
    }
    s_red[threadIdx.x] = tau;
    __syncthreads();
    const size_t row = (size_t)qi * nd + s;
    for (int i = threadIdx.x; i < R; i += kBteBlock) Amat[row * R + i] = s_row[i];
    for (int w = kBteBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_red[threadIdx.x] += s_red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) itau[row] = s_red[0];
}

// One F update on the collision matrix: a wave per four consecutive rows (each F value loaded
// once for the four), dF = sum_c A[r][c] F[c] in a fixed order per lane then a fixed tree,
// F_new[r] = F0[r] + dF / (1/tau[r]) where `divide` is set (multiplyTaudFnew), else + dF.
constexpr int kGemvRows = 4;

__global__ void k_bte_gemv(const double *Amat, const double *F, const double *F0, const double *itau,
                           const unsigned char *divide, int R, double *Fnew) {
    const int wv = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    const int r0 = wv * kGemvRows;
    if (r0 >= R) return;   // whole waves
    const int nr = min(kGemvRows, R - r0);
    double acc[kGemvRows][3];
#pragma unroll
    for (int j = 0; j < kGemvRows; ++j) acc[j][0] = acc[j][1] = acc[j][2] = 0.0;
    for (int c = lane; c < R; c += 64) {
        const double f0 = F[3 * c], f1 = F[3 * c + 1], f2 = F[3 * c + 2];
#pragma unroll
        for (int j = 0; j < kGemvRows; ++j) {
            const double a = j < nr ? Amat[(size_t)(r0 + j) * R + c] : 0.0;
            acc[j][0] = fma(a, f0, acc[j][0]);
            acc[j][1] = fma(a, f1, acc[j][1]);
            acc[j][2] = fma(a, f2, acc[j][2]);
        }
    }
#pragma unroll
    for (int j = 0; j < kGemvRows; ++j)
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc[j][k] += __shfl_down(acc[j][k], off, 64);
    if (lane == 0)
        for (int j = 0; j < nr; ++j) {
            const int r = r0 + j;
#pragma unroll
            for (int k = 0; k < 3; ++k) Fnew[3 * r + k] = F0[3 * r + k] + (divide[r] ? acc[j][k] / itau[r] : acc[j][k]);
        }
}

// out[(qi, s)][k] = sum over chunks in order
__global__ void k_bte_reduce(const double *partial, int nrow, int nch, int K, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrow * K) return;
    const int r = i / K, k = i % K;
    double a = 0.0;
    for (int c = 0; c < nch; ++c) a += partial[((size_t)r * nch + c) * K + k];
    out[i] = a;
}

}  // namespace

// Time-reversal irreducible set (representative = the lower grid index of q and -q), the F map,
// and W of every representative q for every grid q1 (FP64 W path).
int bte_setup(Ctx &c) {
    BteState &b = c.bte;
    if (b.ready) return SPR_OK;
    if (c.h_tetra.empty()) {
        set_error("BTE: the context was created without tetrahedra");
        return SPR_E_STATE;
    }
    if (c.fc3.order != 3) {
        set_error("BTE: the context was created without FC3");
        return SPR_E_STATE;
    }
    const int nq = c.nq, m0 = c.mesh[0], m1 = c.mesh[1], m2 = c.mesh[2];
    std::vector<int4> coord(nq);
    std::vector<int2> map(nq, make_int2(-1, 0));
    b.h_irr.clear();
    b.h_deg.clear();
    for (int g = 0; g < nq; ++g) coord[g] = make_int4(g % m0, (g / m0) % m1, g / (m0 * m1), 0);
    for (int g = 0; g < nq; ++g) {
        if (map[g].x >= 0) continue;
        const int n = ((m0 - coord[g].x) % m0) + m0 * (((m1 - coord[g].y) % m1) + m1 * ((m2 - coord[g].z) % m2));
        const int k = (int)b.h_irr.size();
        b.h_irr.push_back(g);
        map[g] = make_int2(k, 1);
        b.h_deg.push_back(1);
        if (n != g) {
            map[n] = make_int2(k, -1);
            b.h_deg.back() = 2;
        }
    }
    b.nirr = (int)b.h_irr.size();
    b.h_map = map;
    const size_t per_q = (size_t)nq * c.nd * c.nd * c.nd;
    const size_t wbytes = (size_t)b.nirr * per_q * sizeof(double);
    if (wbytes > ((size_t)96 << 30)) {
        set_error("BTE: W storage of " + std::to_string(wbytes >> 30) + " GiB exceeds the 96 GiB budget");
        return SPR_E_UNSUPPORTED;
    }
    auto up = [&](auto **dst, const auto *src, size_t n) -> int {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*dst)>>;
        if (hipMalloc((void **)dst, n * sizeof(T)) != hipSuccess) {
            set_error("BTE: hipMalloc failed");
            return SPR_E_OOM;
        }
        if (src && hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("BTE: upload failed");
            return SPR_E_HIP;
        }
        return SPR_OK;
    };
    int rc;
    if ((rc = up(&b.d_coord, coord.data(), (size_t)nq))) return rc;
    if ((rc = up(&b.d_map, map.data(), (size_t)nq))) return rc;
    if ((rc = up(&b.d_irr, b.h_irr.data(), (size_t)b.nirr))) return rc;
    if ((rc = up(&b.d_tetra, c.h_tetra.data(), c.h_tetra.size()))) return rc;
    if ((rc = up(&b.d_W, (const double *)nullptr, (size_t)b.nirr * per_q))) return rc;
    if ((rc = up(&b.d_F, (const double *)nullptr, (size_t)b.nirr * c.nd * 3))) return rc;
    if ((rc = up(&b.d_out, (const double *)nullptr, (size_t)b.nirr * c.nd * 3))) return rc;
    // W of every q1 of each representative, each ordering evaluated on its own (as the oracle and
    // Src_iterative evaluate them), from items of 64 consecutive q1 with their partners
    // q2 = -(q0 + q1): the W mode of the small-cell kernel (1 and 2 atoms) or of the matrix-core
    // kernel (3 to 10 atoms); otherwise, or with SPR_KERNEL=generic, the generic kernel one
    // representative at a time
    const char *force = getenv("SPR_KERNEL");
    const bool generic = force && std::string(force) == "generic";
    if (!generic && (c.nb <= 2 || lw3_mm_w_built(c.nb))) {
        std::vector<int2> pairs((size_t)b.nirr * nq);
        std::vector<int> iq0, ioff, icnt;
        for (int k = 0; k < b.nirr; ++k) {
            const int g0 = b.h_irr[k], a0 = g0 % m0, a1 = (g0 / m0) % m1, a2 = g0 / (m0 * m1);
            for (int q = 0; q < nq; ++q) {
                const int x = q % m0, y = (q / m0) % m1, z = q / (m0 * m1);
                const int p = (2 * m0 - a0 - x) % m0 + m0 * ((2 * m1 - a1 - y) % m1 + m1 * ((2 * m2 - a2 - z) % m2));
                pairs[(size_t)k * nq + q] = make_int2(q, p);
            }
            for (int o = 0; o < nq; o += 64) {
                iq0.push_back(k);
                ioff.push_back(k * nq + o);
                icnt.push_back(std::min(64, nq - o));
            }
        }
        const int nit = (int)iq0.size();
        int2 *d_pairs = nullptr;
        int *d_iq0 = nullptr, *d_ioff = nullptr, *d_icnt = nullptr;
        double *d_part = nullptr;
        auto done = [&](int r) {
            c.override_items = nullptr;
            for (void *q : {(void *)d_pairs, (void *)d_iq0, (void *)d_ioff, (void *)d_icnt, (void *)d_part})
                if (q) (void)hipFree(q);
            return r;
        };
        if ((rc = up(&d_pairs, pairs.data(), pairs.size())) || (rc = up(&d_iq0, iq0.data(), iq0.size())) ||
            (rc = up(&d_ioff, ioff.data(), ioff.size())) || (rc = up(&d_icnt, icnt.data(), icnt.size())) ||
            (rc = up(&d_part, (const double *)nullptr, (size_t)nit * kMaxUnitsPerItem * c.nd)))
            return done(rc);
        ItemSet set{d_iq0, d_ioff, d_icnt, d_pairs, nit, d_part};
        set.dense = true;
        c.override_items = &set;
        rc = c.nb <= 2 ? launch_lw3_small_w_items(c, b.d_irr, b.nirr, b.d_W)
                       : launch_lw3_mm_w_items(c, b.d_irr, b.nirr, b.d_W);
        if (rc == SPR_OK && hipStreamSynchronize(c.stream) != hipSuccess) rc = SPR_E_HIP;
        if (rc) return done(rc);
        done(SPR_OK);
    } else {
        for (int k = 0; k < b.nirr; ++k)
            if ((rc = launch_lw3_generic(c, b.d_irr + k, 1, 1.0, SPR_PREC_FP64, nullptr, b.d_W + (size_t)k * per_q)))
                return rc;
    }
    SPR_HIP(hipStreamSynchronize(c.stream));
    b.ready = true;
    return SPR_OK;
}

// MODE 0: d_out = 1/tau (nirr, nd); MODE 1: d_out = dF (nirr, nd, 3) of b.d_F before the division
int bte_sums(Ctx &c, int mode) {
    BteState &b = c.bte;
    BteArgs A;
    A.W = b.d_W;
    A.tetra = b.d_tetra;
    A.coord = b.d_coord;
    A.irr_grid = b.d_irr;
    A.map = b.d_map;
    A.omega = c.omega;
    A.nbe = c.nbe;
    A.F = b.d_F;
    A.nq = c.nq;
    A.nd = c.nd;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    A.ntet = 6 * c.nq;
    // Sctr3THz2 of the iterative code (pi / 4, Src_iterative/constants.f90:53-54) over the pi / 16
    // of the linewidth path's W
    A.scale = 4.0;
    const long rows = (long)b.nirr * c.nd;
    // chunks of at least one tetrahedron per thread (each thread owns whole tetrahedra), enough
    // of them for about 4096 workgroups
    A.nch = (int)std::max<long>(1, std::min<long>(A.ntet / kBteBlock, (4096 + rows - 1) / rows));
    A.chunk = (A.ntet + A.nch - 1) / A.nch;
    A.nch = (A.ntet + A.chunk - 1) / A.chunk;
    const int K = mode == 0 ? 1 : 3;
    int rc;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)rows * A.nch * K * sizeof(double)))) return rc;
    A.partial = c.partial;
    const dim3 grid(A.nch, c.nd, b.nirr);
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    if (mode == 0)
        hipLaunchKernelGGL(k_bte<0>, grid, dim3(kBteBlock), 0, c.stream, A);
    else
        hipLaunchKernelGGL(k_bte<1>, grid, dim3(kBteBlock), 0, c.stream, A);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    hipLaunchKernelGGL(k_bte_reduce, dim3((unsigned)((rows * K + 255) / 256)), dim3(256), 0, c.stream, c.partial,
                       (int)rows, A.nch, K, b.d_out);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

// Collision matrix A (rows and columns (irreducible q, s)) and 1/tau in one pass over the
// tetrahedra; SPR_E_UNSUPPORTED (no error text) when a row does not fit in LDS or A exceeds the
// memory budget, and the caller keeps the per-iteration tetrahedron sums.
int bte_build(Ctx &c) {
    BteState &b = c.bte;
    // SPR_BTE_NO_MATRIX=1 keeps the per-iteration tetrahedron sums (the path of cells whose
    // collision matrix does not fit), so that path can be run on any mesh
    if (const char *e = std::getenv("SPR_BTE_NO_MATRIX"))
        if (std::atoi(e) != 0) return SPR_E_UNSUPPORTED;
    if (b.a_ready) return SPR_OK;
    const int R = b.nirr * c.nd;
    const size_t lds = (size_t)R * sizeof(double);
    if (lds > 150 * 1024 || (size_t)R * R * sizeof(double) > ((size_t)32 << 30)) return SPR_E_UNSUPPORTED;
    BteArgs A;
    A.W = b.d_W;
    A.tetra = b.d_tetra;
    A.coord = b.d_coord;
    A.irr_grid = b.d_irr;
    A.map = b.d_map;
    A.omega = c.omega;
    A.nbe = c.nbe;
    A.F = nullptr;
    A.nq = c.nq;
    A.nd = c.nd;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    A.ntet = 6 * c.nq;
    A.scale = 4.0;
    A.chunk = A.ntet;
    A.nch = 1;
    A.partial = nullptr;
    if (hipMalloc((void **)&b.d_A, (size_t)R * R * sizeof(double)) != hipSuccess) {
        set_error("BTE: hipMalloc of the collision matrix failed");
        return SPR_E_OOM;
    }
    SPR_HIP(hipMalloc((void **)&b.d_itau, (size_t)R * sizeof(double)));
    SPR_HIP(hipMalloc((void **)&b.d_F0, (size_t)R * 3 * sizeof(double)));
    SPR_HIP(hipMalloc((void **)&b.d_Fb, (size_t)R * 3 * sizeof(double)));
    SPR_HIP(hipMalloc((void **)&b.d_div, (size_t)R));
    SPR_HIP(hipFuncSetAttribute((const void *)k_bte_build, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    hipLaunchKernelGGL(k_bte_build, dim3(c.nd, b.nirr), dim3(kBteBlock), lds, c.stream, A, R, b.d_A, b.d_itau);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    b.build_ms = ms;
    b.a_ready = true;
    return SPR_OK;
}

// F_new = F0 + A F_cur / (1/tau) on the device (one launch)
int bte_gemv(Ctx &c, const double *d_F, double *d_Fnew) {
    BteState &b = c.bte;
    const int R = b.nirr * c.nd;
    const int waves = (R + kGemvRows - 1) / kGemvRows;
    const int blocks = (waves + 3) / 4;   // four waves per block
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    hipLaunchKernelGGL(k_bte_gemv, dim3(blocks), dim3(256), 0, c.stream, b.d_A, d_F, b.d_F0, b.d_itau, b.d_div, R,
                       d_Fnew);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

}  // namespace spr

using spr::Ctx;

namespace {

int bte_ready(spr_ctx *h) {
    if (!h) {
        spr::set_error("BTE: NULL context");
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(h->c.device));
    return spr::bte_setup(h->c);
}

}  // namespace

extern "C" int spr_bte_irr(spr_ctx *h, int *nirr, int *irr_grid, int *deg) {
    int rc;
    if ((rc = bte_ready(h))) return rc;
    const spr::BteState &b = h->c.bte;
    if (nirr) *nirr = b.nirr;
    if (irr_grid) std::memcpy(irr_grid, b.h_irr.data(), b.h_irr.size() * sizeof(int));
    if (deg) std::memcpy(deg, b.h_deg.data(), b.h_deg.size() * sizeof(int));
    return SPR_OK;
}

extern "C" int spr_bte_tau_inv(spr_ctx *h, double *inv_tau) {
    int rc;
    if ((rc = bte_ready(h))) return rc;
    if (!inv_tau) {
        spr::set_error("spr_bte_tau_inv: NULL output");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    rc = spr::bte_build(c);
    if (rc == SPR_OK) {
        SPR_HIP(hipMemcpy(inv_tau, c.bte.d_itau, (size_t)c.bte.nirr * c.nd * sizeof(double), hipMemcpyDeviceToHost));
        return SPR_OK;
    }
    if (rc != SPR_E_UNSUPPORTED) return rc;
    if ((rc = spr::bte_sums(c, 0))) return rc;
    SPR_HIP(hipMemcpy(inv_tau, c.bte.d_out, (size_t)c.bte.nirr * c.nd * sizeof(double), hipMemcpyDeviceToHost));
    return SPR_OK;
}

extern "C" int spr_bte_delta_f(spr_ctx *h, const double *F, double *dF) {
    int rc;
    if ((rc = bte_ready(h))) return rc;
    if (!F || !dF) {
        spr::set_error("spr_bte_delta_f: NULL argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    const size_t n = (size_t)c.bte.nirr * c.nd * 3;
    SPR_HIP(hipMemcpy(c.bte.d_F, F, n * sizeof(double), hipMemcpyHostToDevice));
    if ((rc = spr::bte_sums(c, 1))) return rc;
    SPR_HIP(hipMemcpy(dF, c.bte.d_out, n * sizeof(double), hipMemcpyDeviceToHost));
    return SPR_OK;
}

// kappa of F (Helper kappa, HelperSubroutines.f90:274-347): kW_mK/(Nq vol) sum_q deg sum_s
// w n (n + 1) v (x) F, acoustic branches at Gamma skipped
static void bte_kappa(const Ctx &c, const std::vector<double> &F, double *k) {
    const spr::BteState &b = c.bte;
    const int nd = c.nd;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int qi = 0; qi < b.nirr; ++qi) {
        const int g = b.h_irr[qi];
        for (int s = g == 0 ? 3 : 0; s < nd; ++s) {
            const double w = c.h_omega[(size_t)g * nd + s], n = c.h_nbe[(size_t)g * nd + s];
            const double f = b.h_deg[qi] * w * n * (n + 1.0);
            const double *v = &c.h_vel[((size_t)g * nd + s) * 3];
            const double *Fv = &F[((size_t)qi * nd + s) * 3];
            for (int a = 0; a < 3; ++a)
                for (int bb = 0; bb < 3; ++bb) acc[3 * a + bb] += f * v[a] * Fv[bb];
        }
    }
    const double kw_mk = 1.054571817 * 0.001;   // kW_mK, Src_iterative/constants.f90:58
    for (int i = 0; i < 9; ++i) k[i] = kw_mk / ((double)c.nq * c.vol) * acc[i];
}

extern "C" int spr_bte_solve(spr_ctx *h, double T, double eps_k, int max_iter, double *kappa, double *kappa0,
                             int *iterations, double *F_out) {
    int rc;
    if ((rc = bte_ready(h))) return rc;
    if (!kappa || !(T > 0.0) || max_iter < 0) {
        spr::set_error("spr_bte_solve: bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    spr::BteState &b = c.bte;
    const int nd = c.nd;
    const size_t nrow = (size_t)b.nirr * nd;
    if (c.h_vel.empty()) {
        spr::set_error("spr_bte_solve: the context has no group velocities (grp_vel at spr_ctx_create)");
        return SPR_E_STATE;
    }
    // the collision matrix where it fits (F updates are then one matrix-vector product each),
    // else the tetrahedron sums every iteration, as the reference does
    rc = spr::bte_build(c);
    if (rc != SPR_OK && rc != SPR_E_UNSUPPORTED) return rc;
    const bool matrix = rc == SPR_OK;
    std::vector<double> itau(nrow), F0(nrow * 3, 0.0), F(nrow * 3), dF(nrow * 3);
    if (matrix) {
        SPR_HIP(hipMemcpy(itau.data(), b.d_itau, nrow * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        if ((rc = spr::bte_sums(c, 0))) return rc;
        SPR_HIP(hipMemcpy(itau.data(), b.d_out, nrow * sizeof(double), hipMemcpyDeviceToHost));
    }
    // CreateFzero (HelperSubroutines.f90:118-186): F0 = hbar/kB w v / (1/tau T^2)
    const double funit = 1.054571817 / 1.380649;   // FunitE_11m_K, Src_iterative/constants.f90:56
    auto skipped = [&](int qi, int s) { return b.h_irr[qi] == 0 && s < 3; };
    for (int qi = 0; qi < b.nirr; ++qi)
        for (int s = 0; s < nd; ++s) {
            if (skipped(qi, s) || !(std::fabs(itau[(size_t)qi * nd + s]) > spr::kZeroPrec)) continue;
            const int g = b.h_irr[qi];
            for (int a = 0; a < 3; ++a)
                F0[((size_t)qi * nd + s) * 3 + a] = funit * c.h_omega[(size_t)g * nd + s] *
                                                     c.h_vel[((size_t)g * nd + s) * 3 + a] /
                                                     (itau[(size_t)qi * nd + s] * T * T);
        }
    F = F0;
    double k_old[9], k_new[9];
    bte_kappa(c, F, k_old);
    if (kappa0) std::memcpy(kappa0, k_old, sizeof(k_old));
    // multiplyTaudFnew: divide by 1/tau where it is nonzero, acoustic branches at Gamma untouched
    std::vector<unsigned char> divide(nrow);
    for (int qi = 0; qi < b.nirr; ++qi)
        for (int s = 0; s < nd; ++s)
            divide[(size_t)qi * nd + s] =
                !skipped(qi, s) && std::fabs(itau[(size_t)qi * nd + s]) > spr::kZeroPrec ? 1 : 0;
    double *d_cur = b.d_F, *d_next = b.d_Fb;
    if (matrix) {
        SPR_HIP(hipMemcpy(b.d_F0, F0.data(), F0.size() * sizeof(double), hipMemcpyHostToDevice));
        SPR_HIP(hipMemcpy(b.d_F, F0.data(), F0.size() * sizeof(double), hipMemcpyHostToDevice));
        SPR_HIP(hipMemcpy(b.d_div, divide.data(), nrow, hipMemcpyHostToDevice));
    }
    int it = 0;
    std::memcpy(k_new, k_old, sizeof(k_old));
    while (it < max_iter) {
        ++it;
        if (matrix) {
            if ((rc = spr::bte_gemv(c, d_cur, d_next))) return rc;
            SPR_HIP(hipMemcpy(F.data(), d_next, F.size() * sizeof(double), hipMemcpyDeviceToHost));
            std::swap(d_cur, d_next);
        } else {
            SPR_HIP(hipMemcpy(b.d_F, F.data(), F.size() * sizeof(double), hipMemcpyHostToDevice));
            if ((rc = spr::bte_sums(c, 1))) return rc;
            SPR_HIP(hipMemcpy(dF.data(), b.d_out, dF.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t r = 0; r < nrow; ++r)
                for (int a = 0; a < 3; ++a)
                    F[r * 3 + a] = F0[r * 3 + a] + (divide[r] ? dF[r * 3 + a] / itau[r] : dF[r * 3 + a]);
        }
        bte_kappa(c, F, k_new);
        double diff = 0.0;
        for (int i = 0; i < 9; ++i) diff = std::max(diff, std::fabs(k_new[i] - k_old[i]));
        std::memcpy(k_old, k_new, sizeof(k_new));
        if (diff < eps_k) break;
    }
    std::memcpy(kappa, k_new, sizeof(k_new));
    if (iterations) *iterations = it;
    if (F_out) std::memcpy(F_out, F.data(), F.size() * sizeof(double));
    return SPR_OK;
}
