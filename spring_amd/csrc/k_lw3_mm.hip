// Three-phonon Gaussian linewidths for cells with Nb >= 3 (Ndof = 3 Nb up to 30) on the
// matrix cores, in FP64, FP32 or split BF16.
//
// Work unit: a batch of up to 64 (q1, q2) pairs of one irreducible q0 (pair-permutation
// reduced list, ScatterProbAllq1 ThirdOrder_m.f90:98-164) and one block of s2 branches (the
// s2 split keeps V of a unit in the accumulator registers; only the FT is repeated per block).
// The block size is a tile shape: shape 0 keeps V within 64 registers per lane (Ndof = 30: two
// blocks in FP32/BF16, four in FP64), shape 1 (FP64) takes the next larger block (two blocks at
// Ndof = 30, 254 registers); the launcher times both on the problem and keeps the faster.  One
// workgroup (8 waves) runs the pairs of its unit one after another.  Per pair, per basis atom mu
// (rows a = 3 mu + alpha):
//   FT      T(alpha, b, c) = sum_{(N1,N2) of mu} Phi(alpha,beta,gamma) e^{2 pi i m/L}   (VALU)
//   GEMM B  U1(alpha, b, s2) = sum_c  T(alpha, b, c) e2(c, s2)
//   GEMM C  U2(alpha, s1, s2) = sum_b U1(alpha, b, s2) e1(b, s1)
//   GEMM D  V(s1 s2, s0) += sum_alpha U2(alpha, s1 s2) e0(3 mu + alpha, s0)
// with V held in the accumulator registers of the waves across mu (ScatterMatEl,
// ThirdOrder_m.f90:278-386, factorised; no eigenvector is conjugated).  After the last mu,
// W = |V|^2 /(w0 w1 w2) Sctr3THz2/Nq and the occupation-weighted Gaussians
// (CalculateLinewidth, ThirdOrder_m.f90:514-537) are summed per s0.
//
// Complex GEMMs run as real GEMMs: with re/im interleaved along K and N,
//   [Y_re Y_im] = [X_re X_im] x [[E_re, E_im], [-E_im, E_re]],
// the A operand is the interleaved intermediate (rows in LDS, K contiguous) and the B operand
// is expanded from an eigenvector matrix held as complex E[s][c] in LDS.
// Matrix instructions (16x16 output tiles, lane l owns column l%16 and four rows):
//   FP64   v_mfma_f64_16x16x4_f64: A lane l -> A[l%16][k], B lane l -> B[k][l%16]; lane group
//          g = l/16 walks K in its own contiguous quarter (any consistent K order sums the
//          same), so each lane's fragment is one contiguous LDS vector.  Output rows of lane l:
//          4 r + l/16 (r = 0..3; measured on gfx950).
//   FP32   v_mfma_f32_16x16x4_f32, same operand scheme; output rows 4 (l/16) + r.
//   BF16   v_mfma_f32_16x16x32_bf16 (lane l: A[l%16][8 g + t], B[8 g + t][l%16], t < 8) on
//          split operands x = hi + lo (hi = bf16(x), lo = bf16(x - hi)): X Y ~ Xh Yh + Xh Yl +
//          Xl Yh, FP32 accumulation, so GEMMs B and C keep about 16 mantissa bits; GEMM D
//          (K = 3 complex per atom) runs on the FP32 instruction.
// Delta arguments w0 -+ w1 -+ w2 are formed in FP64 in every mode; FP32 and BF16 evaluate only
// the triplets within sqrt(100) sigma of a resonance (packed per wave, every dropped term is below
// e^-50 of its peak) and sum them with a compensated FP32 sum per lane; the per-s0 totals are
// reduced in FP64 in a fixed order (bitwise reproducible).
#include <vector>

#include "k_lw3_mm.h"

namespace spr {

// FP64 phase-sum task lists (built by ft2_schedule): per atom mu a head (offset, rounds) and the
// tasks as [round][lane] over the 512 lanes of the workgroup.  Task = block (nu, eta) | component
// j << 7 | first cluster within the atom << 12 | cluster count << 23; j = 31 clears the block;
// kFt2Empty is no task.
constexpr int kFt2Rounds = 4;
constexpr unsigned kFt2Empty = 0xFFFFFFFFu;

// SEL: a branch selection rule is set (A.sel), a compile-time flag so that the common case
// carries no test of it (measured 1.6 % of the FP64 launch at Ndof = 30)
template <int MODE, int NB, int VAR, bool WOUT, bool SEL>
__device__ __forceinline__ void mm_unit(const MmArgs &A, const int unit) {
    static_assert(!WOUT || MODE == MM_F64, "W output is FP64");
    using Sh = MmShape<MODE, NB, VAR>;
    using L = MmLds<MODE, NB, VAR>;
    using R = typename Sh::R;
    using Z = typename Sh::Z;
    using ACC = typename Sh::ACC;
    using F = Frag<MODE, NB, VAR>;
    constexpr int ND = Sh::ND, K2 = Sh::K2, PA = Sh::PA, PE = Sh::PE, SB = Sh::SB;
    // GEMM D runs on the FP64 (FP64 mode) or FP32 instruction
    using DACC = std::conditional_t<MODE == MM_F64, v4d, v4f>;
    using DR = std::conditional_t<MODE == MM_F64, double, float>;

    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *p = smem;
    R *sT = reinterpret_cast<R *>(p);                 // T rows (alpha, b), K = (c, p)
    p += L::tu();
    Z *sU2 = reinterpret_cast<Z *>(p);                // U2[s1][s2 local][alpha], s1 stride U2S1
    p += L::u2();
    R *sAC = reinterpret_cast<R *>(p);                // U1 as rows (alpha, s2 local), K = (b, p)
    double *sRed = reinterpret_cast<double *>(p);     //   (end of unit only)
    p += L::ac();
    Z *sE1 = reinterpret_cast<Z *>(p);
    p += L::e();
    Z *sE2 = reinterpret_cast<Z *>(p);
    p += L::e();
    Z *sP = reinterpret_cast<Z *>(p);                 // cluster phases (all atoms or atom mu)
    p += L::p_bytes(A.npmax, A.nterm);
    Z *sPh = reinterpret_cast<Z *>(p);
    p += (size_t)A.L * sizeof(Z);
    double *sMode = reinterpret_cast<double *>(p);    // w0 iw0f | w1 n1 iw1 | w2 n2 iw2 | -
    p += 8 * ND * sizeof(double);
    float *sModeF = reinterpret_cast<float *>(p);     // n1 iw1 n2 iw2 in FP32
    p += 4 * ND * sizeof(float);
    int *sM1 = reinterpret_cast<int *>(p);
    int *sM2 = sM1 + A.nsite;
    int *sBlk = sM2 + A.nsite;                        // blk_start of every (mu, nu, eta) block
    Z *sE0 = reinterpret_cast<Z *>(((uintptr_t)(sBlk + NB * NB * NB + 1) + 15) & ~(uintptr_t)15);   // e0[s0][a]
    R *sFC = reinterpret_cast<R *>(sE0 + ND * ND);    // FC of atom mu (FP32 / BF16 modes)
    int *sFt2 = reinterpret_cast<int *>(sE0 + ND * ND);   // FP64: task-list heads of every atom

    const int tid = threadIdx.x, l = tid & 63, g = l >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int item = unit / Sh::NSPLIT, h = unit % Sh::NSPLIT;
    const int qi = A.item_q0[item];
    const int i0 = A.q0_grid[qi];
    const int cnt = A.item_cnt[item];
    const int poff = A.item_off[item];
    const R *fc = reinterpret_cast<const R *>(A.fc);
    const Z *phase = reinterpret_cast<const Z *>(A.phase);
    const Z *evec = reinterpret_cast<const Z *>(A.evec);
    __syncthreads();   // the previous unit of a persistent workgroup is done with LDS

    // ---- per unit: zero padding, phase table, q0 modes ----
    for (int i = tid; i < Sh::MC * PA; i += 512) sAC[i] = (R)0;
    for (int i = tid; i < Sh::MB * PA; i += 512) sT[i] = (R)0;          // padding rows and columns stay 0
    for (int i = tid; i < Sh::U2ROWS * Sh::U2S1; i += 512) sU2[i] = Z{0, 0};
    for (int i = tid; i < ND * PE; i += 512) {
        if (i % PE >= ND) {
            sE1[i] = Z{0, 0};
            sE2[i] = Z{0, 0};
        }
    }
    for (int i = tid; i < A.L; i += 512) sPh[i] = phase[i];
    for (int i = tid; i <= NB * NB * NB; i += 512) sBlk[i] = A.blk_start[i];
    for (int i = tid; i < ND * ND; i += 512) sE0[i] = evec[(size_t)i0 * ND * ND + i];
    if (!Sh::PALL && A.ft2)
        for (int i = tid; i < 2 * NB; i += 512) sFt2[i] = A.ft2[i];
    if (tid < ND) {
        sMode[tid] = A.omega[(size_t)i0 * ND + tid];
        sMode[ND + tid] = A.inv_omega[(size_t)i0 * ND + tid] * A.factor;
    }
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002);   // 1/(sigma sqrt(2 pi))
    const double hinv = 0.5 / (A.sigma * A.sigma);
    const float hneg = (float)(-hinv);

    // GEMM B: N tiles h NTS + (w % NTSB), M tiles w / NTSB + (8 / NTSB) j
    const int ntB = h * Sh::NTS + (w % Sh::NTSB);
    const bool bOK = (w % Sh::NTSB) < Sh::NTS && ntB < Sh::NT;
    const int mtB0 = w / Sh::NTSB;
    constexpr int MSTEPB = 8 / Sh::NTSB;
    // GEMMs C and D: N tile w % NTC, M tiles w / NTC + (8 / NTC) j
    const int ntC = w % Sh::NTC;
    const bool cOK = ntC < Sh::NT;
    const int mtC0 = w / Sh::NTC;
    constexpr int MSTEPC = 8 / Sh::NTC;
    const int npD = ntC * 16 + (l & 15);      // this lane's column (s0, q) in GEMM D
    const int s0 = npD >> 1, qD = npD & 1;
    const bool s0ok = cOK && s0 < ND;

    KSum<float> accf;
    double accd = 0.0;
    // FT outputs of this thread, e = tid + 512 k -> (block nu eta, component alpha beta gamma):
    // the block index within the atom, the component and the destination in T, the same for
    // every pair and atom of the unit
    int ft_bl[Sh::KE], ft_j[Sh::KE], ft_dst[Sh::KE];
#pragma unroll
    for (int k = 0; k < Sh::KE; ++k) {
        const int e = tid + 512 * k;
        ft_bl[k] = -1;
        ft_j[k] = 0;
        ft_dst[k] = 0;
        if (e < Sh::NE) {
            const int bl = e / 27, j = e % 27;
            const int nu = bl / NB, eta = bl % NB;
            const int al = j / 9, be = (j / 3) % 3, ga = j % 3;
            ft_bl[k] = bl;
            ft_j[k] = j;
            ft_dst[k] = (al * ND + 3 * nu + be) * PA + 2 * (3 * eta + ga);
        }
    }

    for (int pi = 0; pi < cnt; ++pi) {
        const int2 pr = A.pairs[poff + pi];
        const int q1 = pr.x, q2 = pr.y;
        const double wpair = A.pair_w ? A.pair_w[poff + pi] : q1 == q2 ? 1.0 : 2.0;
        __syncthreads();   // the previous pair is done with sE1, sE2, sMode, sM1, sM2
        for (int i = tid; i < ND * ND; i += 512) {
            const int s = i / ND, c = i % ND;
            sE1[s * PE + c] = evec[(size_t)q1 * ND * ND + i];
            sE2[s * PE + c] = evec[(size_t)q2 * ND * ND + i];
        }
        if (tid < ND) {
            sMode[2 * ND + tid] = A.omega[(size_t)q1 * ND + tid];
            sMode[3 * ND + tid] = A.nbe[(size_t)q1 * ND + tid];
            sMode[4 * ND + tid] = A.inv_omega[(size_t)q1 * ND + tid];
            sMode[5 * ND + tid] = A.omega[(size_t)q2 * ND + tid];
            sMode[6 * ND + tid] = A.nbe[(size_t)q2 * ND + tid];
            sMode[7 * ND + tid] = A.inv_omega[(size_t)q2 * ND + tid];
            sModeF[tid] = (float)A.nbe[(size_t)q1 * ND + tid];
            sModeF[ND + tid] = (float)A.inv_omega[(size_t)q1 * ND + tid];
            sModeF[2 * ND + tid] = (float)A.nbe[(size_t)q2 * ND + tid];
            sModeF[3 * ND + tid] = (float)A.inv_omega[(size_t)q2 * ND + tid];
        }
        {
            const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
            const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
            for (int s = tid; s < A.nsite; s += 512) {
                const int4 c = A.site[s];
                sM1[s] = (k10 * c.x + k11 * c.y + k12 * c.z) % A.L;
                sM2[s] = (k20 * c.x + k21 * c.y + k22 * c.z) % A.L;
            }
        }
        DACC vacc[Sh::VT];
#pragma unroll
        for (int j = 0; j < Sh::VT; ++j) vacc[j] = DACC{0, 0, 0, 0};
        __syncthreads();
        // FC of atom mu is staged in LDS (FP32 / BF16) by direct global-to-LDS loads (no
        // registers), issued for the next atom while the GEMMs of this one run
        auto stage = [&](int mu_n) {
            const int a = sBlk[mu_n * NB * NB] * 27, n = sBlk[(mu_n + 1) * NB * NB] * 27 - a;
            for (int c = w; c * 64 < n; c += 8)
                if (c * 64 + l < n) __builtin_amdgcn_global_load_lds(fc + a + c * 64 + l, sFC + c * 64, 4, 0, 0);
        };
        if constexpr (Sh::PALL) {
            stage(0);
            // cluster phases of every atom: P(u) = e^{2 pi i (m1(N1) + m2(N2)) / L}
            for (int u = tid; u < A.nterm; u += 512) {
                int m = sM1[A.tup_site[2 * u]] + sM2[A.tup_site[2 * u + 1]];
                m -= m >= A.L ? A.L : 0;
                sP[u] = sPh[m];
            }
            __builtin_amdgcn_s_waitcnt(0);   // the staged FC has landed
            __syncthreads();
        }

        // cluster phases of atom mu (FP64 mode; the other modes hold every atom's phases)
        auto phases = [&](int mu) {
            const int u0 = sBlk[mu * NB * NB], u1 = sBlk[(mu + 1) * NB * NB];
            for (int u = u0 + tid; u < u1; u += 512) {
                int m = sM1[A.tup_site[2 * u]] + sM2[A.tup_site[2 * u + 1]];
                m -= m >= A.L ? A.L : 0;
                sP[u - u0] = sPh[m];
            }
        };
        // ---- FT: T[(alpha, 3 nu + beta)][2 (3 eta + gamma) + p]; lanes run along the 27
        //      components of a block, so FC reads are contiguous ----
        auto ft = [&](int mu) {
            const int u0 = Sh::PALL ? 0 : sBlk[mu * NB * NB];
            const R *fcm = Sh::PALL ? sFC : fc + (size_t)u0 * 27;
            const int ub0 = sBlk[mu * NB * NB];   // first cluster of atom mu
            if constexpr (Sh::PALL) {
                // FP32 / BF16: index arithmetic of this thread's outputs precomputed per unit
#pragma unroll
                for (int k = 0; k < Sh::KE; ++k) {
                    if (ft_bl[k] < 0) continue;
                    const int blk = mu * NB * NB + ft_bl[k];
                    const int ua = sBlk[blk], ub = sBlk[blk + 1];
                    const R *fj = fcm + (size_t)(ua - ub0) * 27 + ft_j[k];
                    const Z *pj = sP + (ua - u0);
                    R re = 0, im = 0;
#pragma unroll 4
                    for (int u = 0; u < ub - ua; ++u) {
                        const R f = fj[(size_t)u * 27];
                        const Z ph = pj[u];
                        re = fma(f, ph.x, re);
                        im = fma(f, ph.y, im);
                    }
                    R *dst = sT + ft_dst[k];
                    dst[0] = re;
                    dst[1] = im;
                }
            } else {
                // FP64 (FC read from global memory): the registers the precomputed indices would
                // take cost more than the arithmetic they save (measured)
                for (int e = tid; e < Sh::NE; e += 512) {
                    const int bl = e / 27, j = e % 27;
                    const int nu = bl / NB, eta = bl % NB;
                    const int al = j / 9, be = (j / 3) % 3, ga = j % 3;
                    const int blk = (mu * NB + nu) * NB + eta;
                    const int ua = sBlk[blk], ub = sBlk[blk + 1];
                    R re = 0, im = 0;
#pragma unroll 4
                    for (int u = ua; u < ub; ++u) {
                        const R f = fcm[(size_t)(u - ub0) * 27 + j];
                        const Z ph = sP[u - u0];
                        re = fma(f, ph.x, re);
                        im = fma(f, ph.y, im);
                    }
                    R *dst = sT + (al * ND + 3 * nu + be) * PA + 2 * (3 * eta + ga);
                    dst[0] = re;
                    dst[1] = im;
                }
            }
        };
        // FP64 phase sum of atom mu from its task list (one task per lane and round; within a
        // round the lanes of a wave hold tasks of equal or next-to-equal cluster count, and the
        // lane totals are even).  Each lane issues the FC loads of four clusters before their FMAs
        // (clusters past the count add exact zeros, so each output is summed in cluster order).
        auto ft_list = [&](int mu) {
            const int off = sFt2[2 * mu], nr = sFt2[2 * mu + 1];
            const unsigned *cs = reinterpret_cast<const unsigned *>(A.ft2) + off + tid;
            const R *fcm = fc + (size_t)sBlk[mu * NB * NB] * 27;
            unsigned code[kFt2Rounds];
#pragma unroll
            for (int r = 0; r < kFt2Rounds; ++r) code[r] = r < nr ? cs[(size_t)r * 512] : kFt2Empty;
#pragma unroll
            for (int r = 0; r < kFt2Rounds; ++r) {
                const unsigned cd = code[r];
                if (cd == kFt2Empty) continue;
                const int bl = cd & 127, j = (cd >> 7) & 31, uo = (cd >> 12) & 2047, nc = (cd >> 23) & 255;
                const int nu = bl / NB, eta = bl % NB;
                if (j == 31) {   // a block the previous atom wrote and this one does not
#pragma unroll
                    for (int rr = 0; rr < 9; ++rr)
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc) {
                            R *dst = sT + ((rr / 3) * ND + 3 * nu + rr % 3) * PA + 2 * (3 * eta + cc);
                            dst[0] = 0;
                            dst[1] = 0;
                        }
                    continue;
                }
                const int al = j / 9, be = (j / 3) % 3, ga = j % 3;
                const R *fj = fcm + (size_t)uo * 27 + j;
                const Z *pj = sP + uo;
                R re = 0, im = 0;
                for (int u = 0; u < nc; u += 4) {
                    R f[4];
#pragma unroll
                    for (int k = 0; k < 4; ++k) f[k] = u + k < nc ? fj[(size_t)(u + k) * 27] : (R)0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const Z ph = u + k < nc ? pj[u + k] : Z{0, 0};
                        re = fma(f[k], ph.x, re);
                        im = fma(f[k], ph.y, im);
                    }
                }
                R *dst = sT + (al * ND + 3 * nu + be) * PA + 2 * (3 * eta + ga);
                dst[0] = re;
                dst[1] = im;
            }
        };
        const bool tasks = !Sh::PALL && A.ft2 != nullptr;   // FP64 task lists
        if constexpr (!Sh::PALL) {
            phases(0);
            __syncthreads();
        }
        if (tasks) ft_list(0);
        else ft(0);
        __syncthreads();

        // Per atom, three barrier-separated segments: [GEMM B; stage the next atom's FC (FP32 /
        // BF16) or its cluster phases (FP64 task lists)] | [GEMM C] | [GEMM D; the next atom's phase
        // sum, before GEMM D on waves 0-3 and after it on waves 4-7, so that the phase sum of one
        // wave of a SIMD runs beside the matrix products of the other].
#pragma unroll 1
        for (int mu = 0; mu < NB; ++mu) {
            if constexpr (Sh::PALL)
                if (mu + 1 < NB) stage(mu + 1);   // sFC of atom mu was read by the last FT
            // ---- GEMM B: U1[(alpha, b)][(s2, q)] = T x E2', s2 in this unit's block
            //      -> sAC[(alpha, s2 - h SB)][(b, q)] ----
            if (bOK) {
                typename F::BF bf;
                if constexpr (MODE != MM_F64) bf = F::load_b(sE2, ntB * 16);
                const int np = ntB * 16 + (l & 15);
                const int s2 = np >> 1, q = np & 1;
                // two M tiles per pass: independent accumulation chains
                for (int mt = mtB0; mt < Sh::MT_B; mt += 2 * MSTEPB) {
                    const int mt2 = mt + MSTEPB < Sh::MT_B ? mt + MSTEPB : mt;
                    ACC d[2];
                    if (mt2 == mt) {   // one tile left (wave-uniform)
                        if constexpr (MODE == MM_F64) d[0] = F::mma1_e(sT, mt * 16, sE2, ntB * 16);
                        else d[0] = F::mma(sT, mt * 16, bf, ACC{0, 0, 0, 0});
                    } else if constexpr (MODE == MM_F64) F::mma2_e(sT, mt * 16, mt2 * 16, sE2, ntB * 16, d);
                    else F::mma2(sT, mt * 16, mt2 * 16, bf, d);
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int mtt = t ? mt2 : mt;
                        if (t && mt2 == mt) break;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int m = mtt * 16 + tile_row<MODE == MM_F64>(g, r);
                            if (m < 3 * ND && s2 < ND) {
                                const int al = m / ND, b = m % ND;
                                sAC[(al * SB + s2 - h * SB) * PA + 2 * b + q] = (R)d[t][r];
                            }
                        }
                    }
                }
            }
            if (tasks && mu + 1 < NB) phases(mu + 1);   // sP of atom mu was read by the last phase sum
            __syncthreads();
            // ---- GEMM C: U2[(alpha, s2 local)][(s1, q)] = sAC x E1' -> sU2[(s1, s2 local)][alpha] ----
            if (cOK) {
                typename F::BF bf;
                if constexpr (MODE != MM_F64) bf = F::load_b(sE1, ntC * 16);
                const int np = ntC * 16 + (l & 15);
                const int s1 = np >> 1, q = np & 1;
                for (int mt = mtC0; mt < Sh::MT_C; mt += 2 * MSTEPC) {
                    const int mt2 = mt + MSTEPC < Sh::MT_C ? mt + MSTEPC : mt;
                    ACC d[2];
                    if (mt2 == mt) {
                        if constexpr (MODE == MM_F64) d[0] = F::mma1_e(sAC, mt * 16, sE1, ntC * 16);
                        else d[0] = F::mma(sAC, mt * 16, bf, ACC{0, 0, 0, 0});
                    } else if constexpr (MODE == MM_F64) F::mma2_e(sAC, mt * 16, mt2 * 16, sE1, ntC * 16, d);
                    else F::mma2(sAC, mt * 16, mt2 * 16, bf, d);
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int mtt = t ? mt2 : mt;
                        if (t && mt2 == mt) break;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int m = mtt * 16 + tile_row<MODE == MM_F64>(g, r);
                            if (m < 3 * SB && s1 < ND) {
                                const int al = m / SB, sl = m % SB;
                                R *z = reinterpret_cast<R *>(sU2 + s1 * Sh::U2S1 + sl * 3 + al);
                                z[q] = (R)d[t][r];
                            }
                        }
                    }
                }
            }
            if constexpr (Sh::PALL) {
                __builtin_amdgcn_s_waitcnt(0);   // the next atom's FC has landed
            } else if (!tasks) {
                if (mu + 1 < NB) phases(mu + 1);   // sP of atom mu was read by the last FT
            }
            __syncthreads();
            // ---- GEMM D: V[(s1, s2 local)][(s0, q)] += U2[.][(alpha, p)] E0'[(alpha, p)][(s0, q)]
            //      K' = 8: step i, lane group g -> (alpha = g, p = i); alpha = 3 is padding ----
            if (w < 4 && mu + 1 < NB) {
                if (tasks) ft_list(mu + 1);
                else ft(mu + 1);
            }
            if (cOK) {
                DR b0 = 0, b1 = 0;
                if (g < 3 && s0 < ND) {
                    const Z e = sE0[s0 * ND + 3 * mu + g];
                    b0 = (DR)(qD ? e.y : e.x);
                    b1 = (DR)(qD ? e.x : -e.y);
                }
                // every tile index in range for each wave when the M tiles divide evenly; the U2 row of
                // tile j at a fixed stride from tile 0 when the s2 block divides 16; lane group 3
                // reads a neighbouring element (in range) and zeroes it, branch-free
                constexpr bool kDFull = Sh::MT_D % MSTEPC == 0 && Sh::VT * MSTEPC == Sh::MT_D;
                const int m0 = mtC0 * 16 + (l & 15);
                const Z *dA = sU2 + (m0 / SB) * Sh::U2S1 + (m0 % SB) * 3 + g;
#pragma unroll
                for (int j = 0; j < Sh::VT; ++j) {
                    const int mt = mtC0 + MSTEPC * j;
                    if (kDFull || mt < Sh::MT_D) {
                        const int m = mt * 16 + (l & 15);
                        Z a;
                        if constexpr (16 % SB == 0) a = dA[j * (MSTEPC * 16 / SB) * Sh::U2S1];
                        else a = sU2[(m / SB) * Sh::U2S1 + (m % SB) * 3 + g];
                        if (g >= 3) a = Z{0, 0};
                        if constexpr (MODE == MM_F64) {
                            vacc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b0, vacc[j], 0, 0, 0);
                            vacc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b1, vacc[j], 0, 0, 0);
                        } else {
                            vacc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b0, vacc[j], 0, 0, 0);
                            vacc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b1, vacc[j], 0, 0, 0);
                        }
                    }
                }
            }
            if (w >= 4 && mu + 1 < NB) {   // T of atom mu was read by GEMM B
                if (tasks) ft_list(mu + 1);
                else ft(mu + 1);
            }
            __syncthreads();
        }

        // ---- W and the Gaussian-delta weights (FP32 and BF16: on the resonant triplets only).  Lanes l and l^1
        //      hold Re and Im of the same (row, s0); the lane with q = 0 owns rows r = 0, 1 of
        //      every tile and the lane with q = 1 rows 2, 3 ("its elements").  Per chunk of
        //      kDvTiles tiles, each wave (1) stages its V tiles in LDS, (2) tests its elements: a
        //      triplet whose three Gaussians are all below exp(-TCUT) of their peak adds nothing
        //      measurable (FP64: e^-100; FP32 and BF16: e^-50) and is dropped, (3) packs the others
        //      into a list (ballot order), (4) evaluates the list with every lane, writing each term
        //      over the staged Re of its element, and (5) each owner lane adds the terms of its own
        //      elements in element order, so the sums do not depend on the packing.  The delta
        //      arguments are formed in FP64 in every mode (they cancel near resonance); FP32 and
        //      BF16 evaluate the rest in FP32 and sum per pair, then into a compensated sum. ----
        if constexpr (WOUT) {
            // W of every element, both orderings: W(q1; s0, s1, s2) and W(q2; s0, s2, s1) (the map
            // q -> -(q0 + q) is an involution, so every grid q1 is written exactly once).  W is
            // staged in the dead LDS prefix as sW[s0][s1][s2 local] for a chunk of S0C s0 at a time,
            // then written out in runs of SB (first ordering) and Ndof (second) contiguous values.
            constexpr size_t PRE = L::TU + L::U2 + L::AC + 2 * L::E;
            constexpr int PER_S0 = ND * SB;
            constexpr int S0C = (int)(PRE / (PER_S0 * sizeof(double))) < ND ? (int)(PRE / (PER_S0 * sizeof(double))) : ND;
            static_assert(S0C >= 1, "the W staging chunk must hold one s0");
            double *sW = reinterpret_cast<double *>(smem);
            double *wq = A.w_out + (size_t)(qi - A.w_q0_first) * A.w_nq * ND * ND * ND;
            for (int c0 = 0; c0 < ND; c0 += S0C) {
                __syncthreads();   // GEMM D (first chunk) or the last write-out is done with the prefix
                if (s0ok && s0 >= c0 && s0 < c0 + S0C) {
                    const double iw0 = sMode[ND + s0];
#pragma unroll
                    for (int j = 0; j < Sh::VT; ++j) {
                        const int mt = mtC0 + MSTEPC * j;
                        if (mt >= Sh::MT_D) continue;
#pragma unroll
                        for (int k = 0; k < 2; ++k) {
                            const double x = qD ? vacc[j][2 + k] : vacc[j][k];
                            const double y = __shfl_xor(qD ? vacc[j][k] : vacc[j][2 + k], 1, 64);
                            const int m = mt * 16 + tile_row<true>(g, 2 * qD + k);
                            const int s1 = m / SB, sl = m % SB, s2 = h * SB + sl;
                            if (s1 < ND && s2 < ND)
                                sW[((s0 - c0) * ND + s1) * SB + sl] =
                                    (x * x + y * y) * iw0 * sMode[4 * ND + s1] * sMode[7 * ND + s2];
                        }
                    }
                }
                __syncthreads();
                const int n0 = (ND - c0 < S0C ? ND - c0 : S0C) * PER_S0;
                for (int i = tid; i < n0; i += 512) {   // W(q1; s0, s1, s2): runs of SB in s2
                    const int s0l = i / PER_S0, r = i % PER_S0, s1 = r / SB, sl = r % SB;
                    if (h * SB + sl < ND)
                        wq[(((size_t)q1 * ND + c0 + s0l) * ND + s1) * ND + h * SB + sl] = sW[i];
                }
                if (q1 != q2 && !A.w_one)
                    for (int i = tid; i < n0; i += 512) {   // W(q2; s0, s2, s1): runs of Ndof in s1
                        const int s0l = i / PER_S0, r = i % PER_S0, sl = r / ND, s1 = r % ND;
                        if (h * SB + sl < ND)
                            wq[(((size_t)q2 * ND + c0 + s0l) * ND + h * SB + sl) * ND + s1] =
                                sW[(s0l * ND + s1) * SB + sl];
                    }
            }
            // the staging overwrote zero padding of the GEMM operand buffers (set once per unit):
            // clear what it used once every thread is done with it
            __syncthreads();
            constexpr size_t USED = (size_t)S0C * PER_S0 * sizeof(double);
            for (int i = tid; i < (int)((USED + 15) / 16); i += 512)
                reinterpret_cast<uint4 *>(smem)[i] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            // FP64 evaluates the list in FP64 and sums each pair's terms in FP64 (element order);
            // FP32 and BF16 evaluate in FP32 and add each pair's sum into a compensated FP32 sum
            using DT = typename L::DT;
            constexpr bool F64 = MODE == MM_F64;
            constexpr double TCUT = F64 ? 100.0 : 50.0;
            const double dcut = A.sigma * sqrt(2.0 * TCUT);
            DT *sV = reinterpret_cast<DT *>(smem) + (size_t)w * (kDvTiles * 256);
            unsigned short *sList = reinterpret_cast<unsigned short *>(smem + L::DV) + (size_t)w * (kDvTiles * 128);
            const double w0 = s0ok ? sMode[s0] : 0.0;
            // s2 blocks of 16: an element's s1 is its tile's M index (the same for the whole wave)
            // and its s2 depends only on the lane and k, so the test needs w0 - w2 and w0 + w2 of
            // the lane's two s2 once per pair and w1 once per tile
            constexpr bool kRowS2 = SB == 16;
            double dA[2] = {0.0, 0.0}, dB[2] = {0.0, 0.0};
            bool okA[2] = {false, false};
            if constexpr (kRowS2) {
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const int s2 = h * SB + tile_row<F64>(g, 2 * qD + k);
                    okA[k] = s0ok && s2 < ND;
                    const double w2 = okA[k] ? sMode[5 * ND + s2] : 0.0;
                    dA[k] = w0 - w2;
                    dB[k] = w0 + w2;
                }
            }
            DT pairv = 0;
#pragma unroll
            for (int j0 = 0; j0 < Sh::VT; j0 += kDvTiles) {
                // (1) stage this wave's tiles
#pragma unroll
                for (int t = 0; t < kDvTiles; ++t)
                    if (j0 + t < Sh::VT) *reinterpret_cast<DACC *>(sV + (t * 64 + l) * 4) = vacc[j0 + t];
                mm_wave_sync();
                // (2) + (3): test this lane's elements, pack the resonant ones
                int cnt = 0;
                unsigned act = 0;
#pragma unroll
                for (int t = 0; t < kDvTiles; ++t) {
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const int mt = mtC0 + MSTEPC * (j0 + t);
                        const int m = mt * 16 + tile_row<F64>(g, 2 * qD + k);
                        const int s1 = m / SB, s2 = h * SB + m % SB;
                        bool near = false;
                        if constexpr (kRowS2) {
                            if (okA[k] && j0 + t < Sh::VT && mt < Sh::MT_D && (!SEL || A.sel[ND + s1 * ND + s2] != 0.0)) {
                                const double w1 = sMode[2 * ND + mt];
                                near = fabs(dA[k] - w1) < dcut || fabs(dA[k] + w1) < dcut || fabs(dB[k] - w1) < dcut;
                            }
                        } else if (s0ok && j0 + t < Sh::VT && mt < Sh::MT_D && s1 < ND && s2 < ND &&
                                   (!SEL || A.sel[ND + s1 * ND + s2] != 0.0)) {
                            const double w1 = sMode[2 * ND + s1], w2 = sMode[5 * ND + s2];
                            const double ap = w1 + w2, am = w1 - w2;
                            near = fabs(w0 - ap) < dcut || fabs(w0 + am) < dcut || fabs(w0 - am) < dcut;
                        }
                        const unsigned long long msk = __ballot(near);
                        if (near) {
                            const int pos = cnt + (int)__builtin_amdgcn_mbcnt_hi(
                                                      (unsigned)(msk >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0u));
                            sList[pos] = (unsigned short)(t * 128 + l * 2 + k);
                            act |= 1u << (2 * t + k);
                        }
                        cnt += __popcll(msk);
                    }
                }
                mm_wave_sync();
                // (4): every lane evaluates list entries; the term replaces the staged Re of its
                //      element (read by this lane only, before the write)
                for (int idx = l; idx < cnt; idx += 64) {
                    const int code = sList[idx];
                    const int t = code >> 7, lo = (code >> 1) & 63, k = code & 1;
                    const int r = 2 * (lo & 1) + k;
                    const DT x = sV[(t * 64 + lo) * 4 + r], y = sV[(t * 64 + (lo ^ 1)) * 4 + r];
                    const int mt = mtC0 + MSTEPC * (j0 + t);
                    const int m = mt * 16 + tile_row<F64>(lo >> 4, r);
                    const int s1 = m / SB, s2 = h * SB + m % SB;
                    const double we = sMode[(ntC * 16 + (lo & 15)) >> 1];
                    const double w1 = sMode[2 * ND + s1], w2 = sMode[5 * ND + s2];
                    const double ap = w1 + w2, am = w1 - w2;
                    if constexpr (F64) {
                        const double d1 = we - ap, d2 = we + am, d3 = we - am;
                        const double W = (x * x + y * y) * (sMode[4 * ND + s1] * sMode[7 * ND + s2]);
                        const double n1 = sMode[3 * ND + s1], n2 = sMode[6 * ND + s2];
                        const double G = (n1 + n2 + 1.0) * exp(-hinv * d1 * d1) +
                                         (n1 - n2) * (exp(-hinv * d2 * d2) - exp(-hinv * d3 * d3));
                        sV[(t * 64 + lo) * 4 + r] = SEL ? W * G * A.sel[ND + s1 * ND + s2] : W * G;
                    } else {
                        const float f1 = (float)(we - ap), f2 = (float)(we + am), f3 = (float)(we - am);
                        const float n1 = sModeF[s1], iw1 = sModeF[ND + s1];
                        const float n2 = sModeF[2 * ND + s2], iw2 = sModeF[3 * ND + s2];
                        const float W = (x * x + y * y) * (iw1 * iw2);
                        const float e1 = __expf(hneg * f1 * f1), e2 = __expf(hneg * f2 * f2), e3 = __expf(hneg * f3 * f3);
                        const float sm = SEL ? (float)A.sel[ND + s1 * ND + s2] : 1.f;
                        sV[(t * 64 + lo) * 4 + r] = W * sm * ((n1 + n2 + 1.f) * e1 + (n1 - n2) * (e2 - e3));
                    }
                }
                mm_wave_sync();
                // (5): the owner adds its elements' terms in element order
#pragma unroll
                for (int t = 0; t < kDvTiles; ++t)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        if (act & (1u << (2 * t + k))) pairv += sV[(t * 64 + l) * 4 + 2 * qD + k];
                mm_wave_sync();   // this wave has read the slice before the next chunk overwrites it
            }
            if (s0ok) {
                if constexpr (F64) accd += pairv * (wpair * c0 * sMode[ND + s0]);
                else accf.add(pairv * (float)(wpair * c0 * sMode[ND + s0]));
            }
            // the staging overwrote zero padding of the GEMM operand buffers (set once per unit):
            // clear the prefix again once every wave is done with it
            __syncthreads();
            for (int i = tid; i < (int)((L::DV + L::DL) / 16); i += 512)
                reinterpret_cast<uint4 *>(smem)[i] = make_uint4(0u, 0u, 0u, 0u);
        }
    }
    // ---- reduce per s0 in a fixed order: wave w holds the columns of N tile w % NTC ----
    __syncthreads();
    sRed[tid] = s0ok ? (MODE == MM_F64 ? accd : accf.value()) : 0.0;
    __syncthreads();
    if (tid < ND) {
        const int s = tid;
        const int nt = (2 * s) >> 4, sj = (2 * s) & 15;
        double tot = 0.0;
        for (int ww = nt; ww < 8; ww += Sh::NTC)
            for (int gg = 0; gg < 4; ++gg)
                for (int q = 0; q < 2; ++q) tot += sRed[ww * 64 + gg * 16 + sj + q];
        A.ic.partial[(size_t)unit * ND + s] = SEL ? tot * A.sel[s] : tot;
    }
}

template <int MODE, int NB, int VAR, bool WOUT, bool SEL>
__global__ __launch_bounds__(512) void k_lw3_mm(MmArgs A) {
    if (!A.ic.counter) {
        mm_unit<MODE, NB, VAR, WOUT, SEL>(A, blockIdx.x);
        return;
    }
    __shared__ int s_unit;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_unit = atomicAdd(A.ic.counter, 1);
        __syncthreads();
        const int unit = s_unit;
        if (unit >= A.ic.nitems) break;
        mm_unit<MODE, NB, VAR, WOUT, SEL>(A, unit);
    }
}

// The FP64 phase-sum task lists of every atom (format above the kernel), built once per context:
// the outputs of the atom's non-empty blocks and a clear task per block the previous atom wrote
// and this one does not, sorted by cluster count and laid out in rounds over the 512 lanes in
// alternating directions, so that the lanes of a wave run equal counts and the lane totals (and
// the shares of the two waves of a SIMD, which run theirs at different times) even out.  false
// when the layout does not fit the task format or kFt2Rounds rounds.
static bool ft2_schedule(Ctx &c, int nb) {
    DevIFC &f = c.fc3;
    if (f.ft2_tried) return f.ft2_ok;
    f.ft2_tried = true;
    const std::vector<int> &bs = f.h_blk_start;
    const int nb2 = nb * nb;
    if ((int)bs.size() < nb * nb2 + 1 || nb2 > 127) return false;
    typedef std::pair<int, unsigned> Task;   // (cluster count, code)
    std::vector<int> h(2 * nb);
    for (int mu = 0; mu < nb; ++mu) {
        const int pm = (mu + nb - 1) % nb, u0 = bs[mu * nb2];
        std::vector<Task> v;
        for (int bl = 0; bl < nb2; ++bl) {
            const int b = mu * nb2 + bl, cnt = bs[b + 1] - bs[b], uo = bs[b] - u0;
            if (cnt > 255 || uo + cnt > 2048) return false;
            if (cnt == 0) {
                if (bs[pm * nb2 + bl + 1] > bs[pm * nb2 + bl]) v.push_back({0, (unsigned)bl | 31u << 7});
                continue;
            }
            for (int j = 0; j < 27; ++j)
                v.push_back({cnt, (unsigned)bl | (unsigned)j << 7 | (unsigned)uo << 12 | (unsigned)cnt << 23});
        }
        std::stable_sort(v.begin(), v.end(), [](const Task &x, const Task &y) { return x.first > y.first; });
        const int nr = ((int)v.size() + 511) / 512;
        if (nr > kFt2Rounds) return false;
        const size_t off = h.size();
        h.resize(off + (size_t)nr * 512, (int)kFt2Empty);
        for (size_t i = 0; i < v.size(); ++i) {
            const int r = (int)(i / 512), k = (int)(i % 512);
            h[off + (size_t)r * 512 + (r % 2 ? 511 - k : k)] = (int)v[i].second;
        }
        h[2 * mu] = (int)off;
        h[2 * mu + 1] = nr;
    }
    if (hipMalloc((void **)&f.ft2, h.size() * sizeof(int)) != hipSuccess) {
        f.ft2 = nullptr;
        return false;
    }
    if (hipMemcpy(f.ft2, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return false;
    f.ft2_ok = true;
    return true;
}

// ItemLaunch of tile shape VAR, or false when it cannot run here (LDS, registers)
template <int MODE, int NB, int VAR, bool WOUT>
static bool mm_launch_of(MmArgs &A, int nitems, ItemLaunch &L) {
    using Sh = MmShape<MODE, NB, VAR>;
    const size_t dyn = MmLds<MODE, NB, VAR>::total(A.npmax, A.nterm, A.nsite, A.L);
    if (dyn > 160 * 1024) return false;
    const void *fn = A.sel ? (const void *)k_lw3_mm<MODE, NB, VAR, WOUT, !WOUT>
                           : (const void *)k_lw3_mm<MODE, NB, VAR, WOUT, false>;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) return false;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 512, dyn) != hipSuccess || per_cu < 1) return false;
    static const char *names[3] = {"mm_f64", "mm_f32", "mm_bf16"};
    L = ItemLaunch{fn, &A, sizeof(A), 512, dyn, nitems * Sh::NSPLIT, per_cu,
                   std::string(names[MODE]) + (WOUT ? "_w" : "")};
    L.upi = Sh::NSPLIT;
    L.shape = "s2blk" + std::to_string(Sh::SB);
    return true;
}

template <int MODE, int NB, bool WOUT = false>
static int launch_mm_t(Ctx &c, MmArgs A, int nitems, int q0_first, int q0_count, int it0, double *d_lw) {
    ItemLaunch Ls[3];
    int n = 0;
    // FP64: the phase sum over the task lists (SPR_FT_SCHEDULE=0 keeps the loop over every block
    // of the atom)
    A.ft2 = nullptr;
    if constexpr (MODE == MM_F64) {
        const char *fe = getenv("SPR_FT_SCHEDULE");
        if (!(fe && std::string(fe) == "0") && ft2_schedule(c, NB)) A.ft2 = c.fc3.ft2;
    }
    if (mm_launch_of<MODE, NB, 0, WOUT>(A, nitems, Ls[n])) ++n;
    // the larger s2 block is built for FP64 only: in FP32 / BF16 its accumulators spill
    if constexpr (MODE == MM_F64 && MmShape<MODE, NB, 1>::NTS != MmShape<MODE, NB, 0>::NTS)
        if (mm_launch_of<MODE, NB, 1, WOUT>(A, nitems, Ls[n])) ++n;
    // FP32 / BF16: the s2-per-wave kernel where it is built
    if constexpr (MODE != MM_F64 && !WOUT)
        if (mx_launch_of(c, MODE, NB, A, nitems, Ls[n])) ++n;
    if (n == 0) {
        set_error("three-phonon MFMA kernel: no tile shape fits this device (LDS " +
                  std::to_string(MmLds<MODE, NB, 0>::total(A.npmax, A.nterm, A.nsite, A.L)) + " B)");
        return SPR_E_UNSUPPORTED;
    }
    for (int k = 0; k < n; ++k) Ls[k].direct = c.override_items != nullptr;
    // caller-supplied items run directly on one candidate: for items of one pair each the first
    // (the per-atom kernel, four units per pair in FP64); for whole 64-pair items the last built
    // (the s2-per-wave kernel in FP32 / BF16, the larger s2 block in FP64), the shapes the tuner
    // picks for the context's own items at Ndof = 30
    if (c.override_items && c.override_items->dense) return run_items_shapes(c, Ls + (n - 1), 1, q0_first, q0_count, it0, d_lw);
    return run_items_shapes(c, Ls, n, q0_first, q0_count, it0, d_lw);
}

int launch_lw3_mm(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw) {
    int rc;
    // caller-supplied items (override_items) need none of the context's pair lists
    if (!c.override_items && (rc = ensure_pairs(c))) return rc;
    const ItemSet *ov = c.override_items;
    const int it0 = ov ? 0 : c.h_item_first[q0_first];
    const int nitems = ov ? ov->nitems : c.h_item_first[q0_first + q0_count] - it0;
    const bool f64 = prec == SPR_PREC_FP64;
    MmArgs A;
    A.w_out = nullptr;
    A.w_q0_first = A.w_nq = A.w_one = 0;
    A.ft_sched = nullptr;
    A.ft_nmax = 0;
    A.item_q0 = ov ? ov->item_q0 : c.d_item_q0 + it0;
    A.item_off = ov ? ov->item_off : c.d_item_off + it0;
    A.item_cnt = ov ? ov->item_cnt : c.d_item_cnt + it0;
    A.pairs = ov ? ov->pairs : c.d_pairs;
    A.pair_w = ov ? ov->pair_w : nullptr;
    A.q0_grid = c.irr_grid;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.blk_start = c.fc3.blk_start;
    A.tup_site = c.fc3.tup_site;
    A.fc = f64 ? (const void *)c.fc3.fc64 : (const void *)c.fc3.fc32;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.phase = f64 ? (const void *)c.phase64 : (const void *)c.phase32;
    A.evec = f64 ? (const void *)c.evec : (const void *)c.evec32;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sel = c.sel3 ? c.d_sel3 : nullptr;
    A.sigma = sigma;
    {
        // Sctr3THz2 from its defining factors (constants.f90:59-60)
        const double pi = 4.0 * std::atan(1.0);
        A.factor = 1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) *
                   1000000.0 * pi * 0.0625 / (double)c.nq;
    }
    int npmax = 0;
    const int nb2 = c.nb * c.nb;
    for (int mu = 0; mu < c.nb; ++mu)
        npmax = std::max(npmax, c.fc3.h_blk_start[(mu + 1) * nb2] - c.fc3.h_blk_start[mu * nb2]);
    A.npmax = std::max(npmax, 1);
    A.nterm = std::max(c.fc3.h_blk_start[c.nb * nb2], 1);
    A.ic.partial = ov ? ov->partial : c.ipartial;
    const int mode = prec == SPR_PREC_FP64 ? MM_F64 : prec == SPR_PREC_FP32 ? MM_F32 : MM_BF16;
    switch (c.nb) {
#define SPR_NB(n)                                                                                    \
    case n:                                                                                          \
        return mode == MM_F64   ? launch_mm_t<MM_F64, n>(c, A, nitems, q0_first, q0_count, it0, d_lw) \
               : mode == MM_F32 ? launch_mm_t<MM_F32, n>(c, A, nitems, q0_first, q0_count, it0, d_lw) \
                                : launch_mm_t<MM_BF16, n>(c, A, nitems, q0_first, q0_count, it0, d_lw);
        SPR_NB(2) SPR_NB(3) SPR_NB(4) SPR_NB(5) SPR_NB(6) SPR_NB(8) SPR_NB(10)
#undef SPR_NB
        default:
            set_error("three-phonon MFMA kernel: Nb " + std::to_string(c.nb) + " not built (2-6, 8, 10)");
            return SPR_E_UNSUPPORTED;
    }
}

bool lw3_mm_w_built(int nb) { return nb == 2 || nb == 3 || nb == 4 || nb == 5 || nb == 6 || nb == 8 || nb == 10; }

// The W-output launch for the items in A (item lists, q0 grid table): FP64, no delta stage.
static int launch_mm_w_args(Ctx &c, MmArgs &A, int nitems, int q_first, int q_count, double *d_w) {
    int rc;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.blk_start = c.fc3.blk_start;
    A.tup_site = c.fc3.tup_site;
    A.fc = (const void *)c.fc3.fc64;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.phase = (const void *)c.phase64;
    A.evec = (const void *)c.evec;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sel = nullptr;   // W output: the tetrahedron and BTE sums apply the branch ranges themselves
    A.sigma = 1.0;     // unused without the delta stage
    const double pi = 4.0 * std::atan(1.0);
    A.factor = 1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) * 1000000.0 * pi *
               0.0625 / (double)c.nq;   // Sctr3THz2 / Nq (constants.f90:59-60)
    int npmax = 0;
    const int nb2 = c.nb * c.nb;
    for (int mu = 0; mu < c.nb; ++mu)
        npmax = std::max(npmax, c.fc3.h_blk_start[(mu + 1) * nb2] - c.fc3.h_blk_start[mu * nb2]);
    A.npmax = std::max(npmax, 1);
    A.nterm = std::max(c.fc3.h_blk_start[c.nb * nb2], 1);
    A.ft_sched = nullptr;
    A.ft_nmax = 0;
    A.w_out = d_w;
    A.w_nq = c.nq;
    // the per-unit sums of this launch are zero; they are reduced into a scratch row per q0
    if ((rc = ensure((void **)&c.wsink, &c.wsink_cap, (size_t)q_count * c.nd * sizeof(double)))) return rc;
    switch (c.nb) {
#define SPR_NB(n) \
    case n: return launch_mm_t<MM_F64, n, true>(c, A, nitems, q_first, q_count, 0, c.wsink);
        SPR_NB(2) SPR_NB(3) SPR_NB(4) SPR_NB(5) SPR_NB(6) SPR_NB(8) SPR_NB(10)
#undef SPR_NB
        default: return SPR_E_UNSUPPORTED;
    }
}

// W of every grid q1 for irreducible q0 [q_first, q_first + q_count) into
// d_w[q0 - q_first][q1][s0][s1][s2] (FP64), from the pair-reduced items of the context: the
// W-output mode of the matrix-core kernel, both orderings written, no underflow skip.
int launch_lw3_mm_w(Ctx &c, int q_first, int q_count, double *d_w) {
    if (c.override_items || !lw3_mm_w_built(c.nb)) {
        set_error("matrix-core W path: Nb in 2-6, 8, 10 and no item override");
        return SPR_E_UNSUPPORTED;
    }
    int rc;
    if ((rc = ensure_pairs(c))) return rc;
    const int it0 = c.h_item_first[q_first];
    const int nitems = c.h_item_first[q_first + q_count] - it0;
    MmArgs A;
    A.item_q0 = c.d_item_q0 + it0;
    A.item_off = c.d_item_off + it0;
    A.item_cnt = c.d_item_cnt + it0;
    A.pairs = c.d_pairs;
    A.pair_w = nullptr;
    A.q0_grid = c.irr_grid;
    A.ic.partial = c.ipartial;
    A.w_q0_first = q_first;
    A.w_one = 0;
    return launch_mm_w_args(c, A, nitems, q_first, q_count, d_w);
}

// W of every grid q1 for the q0 of caller-supplied items (c.override_items, whole items of up
// to 64 pairs, every q1 of a q0 listed with its partner): d_w[k][q1][s0][s1][s2] for the q0 of
// grid index q0_grid[k], k < nq0 (item_q0 holds k); each ordering evaluated on its own.
int launch_lw3_mm_w_items(Ctx &c, const int *d_q0_grid, int nq0, double *d_w) {
    const ItemSet *ov = c.override_items;
    if (!ov || !ov->dense || !lw3_mm_w_built(c.nb)) {
        set_error("matrix-core W path for listed items: Nb in 2-6, 8, 10 and whole items");
        return SPR_E_UNSUPPORTED;
    }
    MmArgs A;
    A.item_q0 = ov->item_q0;
    A.item_off = ov->item_off;
    A.item_cnt = ov->item_cnt;
    A.pairs = ov->pairs;
    A.pair_w = nullptr;
    A.q0_grid = d_q0_grid;
    A.ic.partial = ov->partial;
    A.w_q0_first = 0;
    A.w_one = 1;
    return launch_mm_w_args(c, A, ov->nitems, 0, nq0, d_w);
}

}  // namespace spr
