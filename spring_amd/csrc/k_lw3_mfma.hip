// Three-phonon scattering for large cells (Ndof = 3 Nb up to 30) on the matrix cores.
//
// Work unit: a batch of up to 64 (q1, q2) pairs of one irreducible q0 (pair-permutation
// reduced list, ScatterProbAllq1 ThirdOrder_m.f90:98-164).  One workgroup (8 waves) runs
// the pairs of its batch one after another.  Per pair, per basis atom mu (rows a = 3mu+alpha):
//   FT      T(a, b, c)   = sum_{(N1,N2)} Phi(alpha,beta,gamma; mu N1 N2) e^{2 pi i m/L}  (VALU FP32)
//   GEMM B  U1(a, b, s2) = sum_c T(a, b, c) e2(c, s2)
//   GEMM C  U2(a, s2, s1) = sum_b U1(a, b, s2) e1(b, s1)        -> stash rows a
//   GEMM D  V(s1 s2, s0) += sum_{a in stash} U2(a, s1 s2) e0(a, s0)   (accumulators persist)
// then W = |V|^2 /(w0 w1 w2) Sctr3THz2/Nq and the Gaussian-delta occupation weights are
// summed per s0 with a compensated FP32 sum (CalculateLinewidth, ThirdOrder_m.f90:514-537).
//
// Complex GEMMs run as real GEMMs on MFMA: with interleaved complex columns,
//   [C_re C_im] = [A_re A_im] x [[B_re, B_im], [-B_im, B_re]]
// so the A operand is the interleaved intermediate and the B operand is built on the fly
// from an eigenvector matrix.  FP32 mode: v_mfma_f32_16x16x4_f32, FP32 intermediates.
// BF16 mode: v_mfma_f32_16x16x32_bf16, BF16 intermediates and eigenvectors, FP32 accumulation
// (BF16 keeps FP32's exponent range, so its error is the 8-bit mantissa rounding of the
// stored intermediates, about 2^-9 relative per element).
#include <algorithm>
#include <cmath>
#include <vector>

#include "cplx.h"
#include "spr_internal.h"

namespace spr {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct MfmaArgs {
    ItemCommon ic;
    const int *item_q0, *item_off, *item_cnt;
    const int2 *pairs;
    const int *q0_grid;
    int L, m0, m1;
    const int *blk_start, *tup_site;
    const float *fc;
    const int4 *site;
    int nsite;
    const float2 *phase;
    const float2 *evec;
    const float *omega, *inv_omega, *nbe;
    float sigma, factor;
    float *scratch;          // FP32 mode: per-workgroup U2 stash slots (persistent grid only)
    int nslots;
};

template <bool BF> struct Store;
template <> struct Store<false> { using S = float; };
template <> struct Store<true> { using S = __bf16; };
template <bool BF> using S_of = typename Store<BF>::S;

// B' element of the real-expanded complex eigenvector matrix E(k, n) = evec row k, mode n,
// stored as E[n * ND + k] (interleaved re/im in S): rows k' = 2k + p, columns n' = 2n + q.
template <typename S, int ND>
__device__ __forceinline__ S bprime(const S *E, int kp, int np) {
    const int k = kp >> 1, p = kp & 1, n = np >> 1, q = np & 1;
    if (k >= ND || n >= ND) return (S)0.0f;
    const S re = E[2 * (n * ND + k)], im = E[2 * (n * ND + k) + 1];
    if (p == q) return re;
    return p == 0 ? im : (S)(-(float)im);
}

// One 16x16 output tile, K' = KP (multiple of 4 / 32): acc += A(m, k') B(k', n)
template <bool BF, int KP, class FA, class FB>
__device__ __forceinline__ v4f mma_tile(FA A, FB B, int m_base, int n_base, v4f acc) {
    const int l = threadIdx.x & 63;
    if constexpr (!BF) {
#pragma unroll
        for (int kk = 0; kk < KP; kk += 4) {
            const float a = (float)A(m_base + (l & 15), kk + (l >> 4));
            const float b = (float)B(kk + (l >> 4), n_base + (l & 15));
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int kk = 0; kk < KP; kk += 32) {
            bf16x8 a, b;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                a[t] = A(m_base + (l & 15), kk + 8 * (l >> 4) + t);
                b[t] = B(kk + 8 * (l >> 4) + t, n_base + (l & 15));
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
    }
    return acc;
}

template <bool BF, int NB>
__device__ __forceinline__ void mfma_item(const MfmaArgs &A, const int item, S_of<BF> *st_global) {
    using S = typename Store<BF>::S;
    constexpr int ND = 3 * NB;
    constexpr int ND2 = ND * ND;
    constexpr int W2 = 2 * ND;                         // interleaved complex row length
    constexpr int KP = BF ? 64 : ((W2 + 3) / 4) * 4;  // padded K' of GEMMs B and C
    constexpr int MT_BC = (3 * ND + 15) / 16;          // m tiles of GEMMs B and C
    constexpr int NT = (W2 + 15) / 16;                 // n tiles (<= 4)
    constexpr int MT_D = (ND2 + 15) / 16;              // m tiles of GEMM D
    constexpr int KP_D = BF ? ((2 * ND + 31) / 32) * 32 : ((2 * ND + 3) / 4) * 4;
    constexpr int NWAVE = 8;
    constexpr int DT_PER_WAVE = (MT_D + 1) / 2;   // wave w: n tile w % 4, m tiles w / 4 + 2 t
    static_assert(NT <= 4, "Ndof must be <= 32");

    // The stash U2[a][s1 s2] (all ND rows) lives in LDS for BF16 and in a per-workgroup
    // global slot for FP32 (FP32 U2 of Ndof = 30 needs 216 KB).
    extern __shared__ __align__(16) unsigned char smem[];
    S *s_T = reinterpret_cast<S *>(smem);                       // [3 ND][W2]
    S *s_U1 = s_T + 3 * ND * W2;                                // [3 ND][W2]
    S *s_stl = s_U1 + 3 * ND * W2;                              // [ND][ND2][2] (BF16 only)
    S *s_st = BF ? s_stl : st_global;
    S *s_e0 = s_stl + (BF ? ND * ND2 * 2 : 0);                  // [ND][ND][2]  (mode, row)
    S *s_e1 = s_e0 + ND2 * 2;
    S *s_e2 = s_e1 + ND2 * 2;
    float *s_mode = reinterpret_cast<float *>(s_e2 + ND2 * 2 + ((ND2 * 2 * 3) & 1));  // 8 x ND
    float2 *s_ph = reinterpret_cast<float2 *>(s_mode + 8 * ND);
    int *s_m1 = reinterpret_cast<int *>(s_ph + A.L);
    int *s_m2 = s_m1 + A.nsite;
    double *s_red = reinterpret_cast<double *>(((uintptr_t)(s_m2 + A.nsite) + 15) & ~(uintptr_t)15);

    const int tid = threadIdx.x, l = tid & 63;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int qi = A.item_q0[item];
    const int i0 = A.q0_grid[qi];
    const int cnt = A.item_cnt[item];
    const int poff = A.item_off[item];
    __syncthreads();   // the previous item of a persistent workgroup is done with LDS

    for (int i = tid; i < A.L; i += 512) s_ph[i] = A.phase[i];
    for (int i = tid; i < ND2; i += 512) {
        const float2 e = A.evec[(size_t)i0 * ND2 + i];
        s_e0[2 * i] = (S)e.x;
        s_e0[2 * i + 1] = (S)e.y;
    }
    if (tid < ND) {
        s_mode[0 * ND + tid] = A.omega[(size_t)i0 * ND + tid];
        s_mode[1 * ND + tid] = A.inv_omega[(size_t)i0 * ND + tid] * A.factor;
    }
    const float c0 = 1.0f / (A.sigma * 2.50662827f);
    const float hinv = 0.5f / (A.sigma * A.sigma);
    const int nt_d = w & 3;                 // GEMM D: fixed n tile per wave
    const int mt_d0 = w >> 2;               // m tiles mt_d0, mt_d0 + 2, ...
    const int j = l & 15;
    const int s0 = (nt_d * 16 + j) >> 1;    // this lane's s0 in GEMM D / step E
    KSum<float> acc;

    for (int pi = 0; pi < cnt; ++pi) {
        const int2 pr = A.pairs[poff + pi];
        const int q1 = pr.x, q2 = pr.y;
        const float wpair = q1 == q2 ? 1.0f : 2.0f;
        __syncthreads();   // previous pair finished with every buffer
        for (int i = tid; i < ND2; i += 512) {
            const float2 a1 = A.evec[(size_t)q1 * ND2 + i], a2 = A.evec[(size_t)q2 * ND2 + i];
            s_e1[2 * i] = (S)a1.x;
            s_e1[2 * i + 1] = (S)a1.y;
            s_e2[2 * i] = (S)a2.x;
            s_e2[2 * i + 1] = (S)a2.y;
        }
        if (tid < ND) {
            s_mode[2 * ND + tid] = A.omega[(size_t)q1 * ND + tid];
            s_mode[3 * ND + tid] = A.nbe[(size_t)q1 * ND + tid];
            s_mode[4 * ND + tid] = A.inv_omega[(size_t)q1 * ND + tid];
            s_mode[5 * ND + tid] = A.omega[(size_t)q2 * ND + tid];
            s_mode[6 * ND + tid] = A.nbe[(size_t)q2 * ND + tid];
            s_mode[7 * ND + tid] = A.inv_omega[(size_t)q2 * ND + tid];
        }
        {
            const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
            const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
            for (int s = tid; s < A.nsite; s += 512) {
                const int4 c = A.site[s];
                s_m1[s] = (k10 * c.x + k11 * c.y + k12 * c.z) % A.L;
                s_m2[s] = (k20 * c.x + k21 * c.y + k22 * c.z) % A.L;
            }
        }
        __syncthreads();

        for (int mu = 0; mu < NB; ++mu) {
            const int srow0 = 3 * mu;   // stash rows of this atom
            // ---- FT: T(alpha, b, c) for the three rows of atom mu ----
            for (int o = tid; o < 3 * ND2; o += 512) {
                const int al = o / ND2, b = (o / ND) % ND, c = o % ND;
                const int nu = b / 3, be = b % 3, eta = c / 3, ga = c % 3;
                const int blk = (mu * NB + nu) * NB + eta;
                float re = 0.f, im = 0.f;
                const int u1 = A.blk_start[blk + 1];
                for (int u = A.blk_start[blk]; u < u1; ++u) {
                    int m = s_m1[A.tup_site[2 * u]] + s_m2[A.tup_site[2 * u + 1]];
                    m -= m >= A.L ? A.L : 0;
                    const float2 ph = s_ph[m];
                    const float f = A.fc[(size_t)u * 27 + al * 9 + be * 3 + ga];
                    re = fmaf(f, ph.x, re);
                    im = fmaf(f, ph.y, im);
                }
                s_T[(al * ND + b) * W2 + 2 * c] = (S)re;
                s_T[(al * ND + b) * W2 + 2 * c + 1] = (S)im;
            }
            __syncthreads();
            // ---- GEMM B: U1[(al,b)][(s2,q)] = T[(al,b)][(c,p)] x E2'[(c,p)][(s2,q)] ----
            for (int t = w; t < MT_BC * NT; t += NWAVE) {
                const int mt = t / NT, nt = t % NT;
                auto fa = [&](int m, int kp) -> S { return (m < 3 * ND && kp < W2) ? s_T[m * W2 + kp] : (S)0.0f; };
                auto fb = [&](int kp, int np) -> S { return bprime<S, ND>(s_e2, kp, np); };
                v4f d = mma_tile<BF, KP>(fa, fb, mt * 16, nt * 16, v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mt * 16 + 4 * (l >> 4) + r, np = nt * 16 + (l & 15);
                    if (m < 3 * ND && np < W2) s_U1[m * W2 + np] = (S)d[r];
                }
            }
            __syncthreads();
            // ---- GEMM C: U2[(al,s2)][(s1,q)] = U1^T[(al,s2)][(b,p)] x E1'[(b,p)][(s1,q)] ----
            for (int t = w; t < MT_BC * NT; t += NWAVE) {
                const int mt = t / NT, nt = t % NT;
                auto fa = [&](int m, int kp) -> S {
                    if (m >= 3 * ND || kp >= W2) return (S)0.0f;
                    const int al = m / ND, s2 = m % ND, b = kp >> 1;
                    return s_U1[(al * ND + b) * W2 + 2 * s2 + (kp & 1)];
                };
                auto fb = [&](int kp, int np) -> S { return bprime<S, ND>(s_e1, kp, np); };
                v4f d = mma_tile<BF, KP>(fa, fb, mt * 16, nt * 16, v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int m = mt * 16 + 4 * (l >> 4) + r, np = nt * 16 + (l & 15);
                    if (m < 3 * ND && np < W2) {
                        const int al = m / ND, s2 = m % ND, s1 = np >> 1;
                        s_st[((srow0 + al) * ND2 + s1 * ND + s2) * 2 + (np & 1)] = (S)d[r];
                    }
                }
            }
        }
        // ---- GEMM D fused with step E, one 16x16 tile at a time:
        //      V[(s1 s2)][(s0,q)] = U2^T[(s1 s2)][(a,p)] x E0'[(a,p)][(s0,q)], then W and the
        //      Gaussian-delta weights (each lane: one s0 and rows r) ----
        __syncthreads();
        {
            auto fa = [&](int m, int kp) -> S {
                const int a = kp >> 1;
                return (m < ND2 && a < ND) ? s_st[(a * ND2 + m) * 2 + (kp & 1)] : (S)0.0f;
            };
            auto fb = [&](int kp, int np) -> S { return bprime<S, ND>(s_e0, kp, np); };
            const bool s0ok = s0 < ND;
            const float w0 = s0ok ? s_mode[s0] : 0.f;
            const float iw0 = s0ok ? s_mode[ND + s0] : 0.f;
#pragma unroll 1
            for (int t = 0; t < DT_PER_WAVE; ++t) {
                const int mt = mt_d0 + 2 * t;
                if (mt >= MT_D || nt_d >= NT) continue;
                const v4f vt = mma_tile<BF, KP_D>(fa, fb, mt * 16, nt_d * 16, v4f{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float x = vt[r];
                    const float y = __shfl_xor(x, 1, 64);
                    const float v2 = x * x + y * y;
                    // lanes with q = 0 take rows 0,1, lanes with q = 1 rows 2,3
                    if (((r >> 1) != (j & 1)) || !s0ok) continue;
                    const int m = mt * 16 + 4 * (l >> 4) + r;
                    if (m >= ND2) continue;
                    const int s1 = m / ND, s2 = m % ND;
                    const float w1 = s_mode[2 * ND + s1], n1 = s_mode[3 * ND + s1], iw1 = s_mode[4 * ND + s1];
                    const float w2 = s_mode[5 * ND + s2], n2 = s_mode[6 * ND + s2], iw2 = s_mode[7 * ND + s2];
                    const float W = v2 * iw0 * iw1 * iw2;
                    const float d1 = w0 - w1 - w2, d2 = w0 + w1 - w2, d3 = w0 - w1 + w2;
                    const float g1 = c0 * __expf(-hinv * d1 * d1);
                    const float g2 = c0 * __expf(-hinv * d2 * d2);
                    const float g3 = c0 * __expf(-hinv * d3 * d3);
                    acc.add(wpair * W * ((n1 + n2 + 1.0f) * g1 + (n1 - n2) * (g2 - g3)));
                }
            }
        }
    }
    // ---- reduce per s0 in a fixed order ----
    __syncthreads();
    s_red[tid] = acc.value();
    __syncthreads();
    if (tid < ND) {
        const int sj = (2 * tid) & 15, nt = (2 * tid) >> 4;
        double tot = 0.0;
        for (int ww = nt; ww < NWAVE; ww += 4)
            for (int g = 0; g < 4; ++g)
                for (int q = 0; q < 2; ++q) tot += s_red[ww * 64 + g * 16 + sj + q];
        A.ic.partial[(size_t)item * ND + tid] = tot;
    }
}

template <bool BF, int NB>
__global__ __launch_bounds__(512) void k_lw3_mfma(MfmaArgs A) {
    constexpr int ND = 3 * NB;
    if (!BF && (int)blockIdx.x >= A.nslots) return;   // persistent grid: the others take the items
    S_of<BF> *slot = BF ? nullptr : reinterpret_cast<S_of<BF> *>(A.scratch) + (size_t)blockIdx.x * ND * ND * ND * 2;
    if (!A.ic.counter) {
        mfma_item<BF, NB>(A, blockIdx.x, slot);
        return;
    }
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(A.ic.counter, 1);
        __syncthreads();
        const int item = s_item;
        if (item >= A.ic.nitems) break;
        mfma_item<BF, NB>(A, item, slot);
    }
}

template <bool BF, int NB>
static size_t mfma_lds(int nsite, int L) {
    using S = typename Store<BF>::S;
    constexpr int ND = 3 * NB, ND2 = ND * ND, W2 = 2 * ND;
    size_t nS = (size_t)2 * 3 * ND * W2 + (BF ? (size_t)ND * ND2 * 2 : 0) + 3 * ND2 * 2;
    size_t b = nS * sizeof(S) + ((ND2 * 2 * 3) & 1) * sizeof(S) + 8 * ND * sizeof(float) +
               (size_t)L * sizeof(float2) + 2 * (size_t)nsite * sizeof(int);
    b = (b + 15) & ~(size_t)15;
    return b + 512 * sizeof(double);
}

template <bool BF, int NB>
static int launch_mfma_t(Ctx &c, MfmaArgs A, int nitems, int q0_first, int q0_count, int it0, double *d_lw) {
    const size_t dyn = mfma_lds<BF, NB>(A.nsite, A.L);
    if (dyn > 160 * 1024) {
        set_error("MFMA kernel: LDS need " + std::to_string(dyn) + " B exceeds 160 KB");
        return SPR_E_UNSUPPORTED;
    }
    const void *fn = (const void *)k_lw3_mfma<BF, NB>;
    SPR_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 512, dyn) != hipSuccess || per_cu < 1) per_cu = 1;
    ItemLaunch L{fn, &A, sizeof(A), 512, dyn, nitems, per_cu, std::string(BF ? "mfma_bf16" : "mfma_f32")};
    L.direct = c.override_items != nullptr;
    if (!BF) {
        // FP32: the U2 stash of each resident workgroup lives in its own global slot, so the
        // grid is persistent with exactly one slot per workgroup (mode 1: ncu x per_cu)
        constexpr int ND = 3 * NB;
        const int slots = c.ncu * per_cu;
        const size_t bytes = (size_t)slots * ND * ND * ND * 2 * sizeof(float);
        if (bytes > c.mfma_scratch_bytes) {
            if (c.mfma_scratch) SPR_HIP(hipFree(c.mfma_scratch));
            c.mfma_scratch = nullptr;
            c.mfma_scratch_bytes = 0;
            drop_graphs(c);   // captured graphs hold the old slot address
            if (hipMalloc((void **)&c.mfma_scratch, bytes) != hipSuccess) {
                set_error("MFMA kernel: cannot allocate " + std::to_string(bytes) + " B of stash slots");
                return SPR_E_OOM;
            }
            c.mfma_scratch_bytes = bytes;
        }
        A.scratch = c.mfma_scratch;
        A.nslots = slots;
        L.fixed_mode = 1;
    }
    return run_items(c, L, q0_first, q0_count, it0, d_lw);
}

int launch_lw3_mfma(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw) {
    int rc;
    if ((rc = ensure_pairs(c))) return rc;
    if (!c.omega32) {
        const size_t nm = (size_t)c.nq * c.nd;
        std::vector<float> o(nm), iw(nm), nb(nm);
        for (size_t i = 0; i < nm; ++i) {
            o[i] = (float)c.h_omega[i];
            nb[i] = (float)c.h_nbe[i];
        }
        SPR_HIP(hipMalloc((void **)&c.omega32, 3 * nm * sizeof(float)));
        std::vector<double> hiw(nm);
        SPR_HIP(hipMemcpy(hiw.data(), c.inv_omega, nm * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nm; ++i) iw[i] = (float)hiw[i];
        SPR_HIP(hipMemcpy(c.omega32, o.data(), nm * sizeof(float), hipMemcpyHostToDevice));
        SPR_HIP(hipMemcpy(c.omega32 + nm, iw.data(), nm * sizeof(float), hipMemcpyHostToDevice));
        SPR_HIP(hipMemcpy(c.omega32 + 2 * nm, nb.data(), nm * sizeof(float), hipMemcpyHostToDevice));
    }
    const ItemSet *ov = c.override_items;
    const int it0 = ov ? 0 : c.h_item_first[q0_first];
    const int nitems = ov ? ov->nitems : c.h_item_first[q0_first + q0_count] - it0;
    const size_t nm = (size_t)c.nq * c.nd;
    MfmaArgs A{};
    A.item_q0 = ov ? ov->item_q0 : c.d_item_q0 + it0;
    A.item_off = ov ? ov->item_off : c.d_item_off + it0;
    A.item_cnt = ov ? ov->item_cnt : c.d_item_cnt + it0;
    A.pairs = ov ? ov->pairs : c.d_pairs;
    A.q0_grid = c.irr_grid;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.blk_start = c.fc3.blk_start;
    A.tup_site = c.fc3.tup_site;
    A.fc = c.fc3.fc32;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.phase = c.phase32;
    A.evec = c.evec32;
    A.omega = c.omega32;
    A.inv_omega = c.omega32 + nm;
    A.nbe = c.omega32 + 2 * nm;
    A.sigma = (float)sigma;
    const double pi = 4.0 * std::atan(1.0);
    A.factor = (float)(1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) *
                       1000000.0 * pi * 0.0625 / (double)c.nq);
    A.ic.partial = ov ? ov->partial : c.ipartial;
    const bool bf = prec == SPR_PREC_BF16;
    switch (c.nb) {
#define SPR_NB(n)                                                                                    \
    case n:                                                                                          \
        return bf ? launch_mfma_t<true, n>(c, A, nitems, q0_first, q0_count, it0, d_lw)               \
                  : launch_mfma_t<false, n>(c, A, nitems, q0_first, q0_count, it0, d_lw);
        SPR_NB(2) SPR_NB(3) SPR_NB(4) SPR_NB(5) SPR_NB(6) SPR_NB(8) SPR_NB(10)
#undef SPR_NB
        default:
            set_error("MFMA kernel: Nb " + std::to_string(c.nb) + " not built (2-6, 8, 10)");
            return SPR_E_UNSUPPORTED;
    }
}

}  // namespace spr
