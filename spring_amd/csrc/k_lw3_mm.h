// Pieces shared by the three-phonon matrix-core kernels (k_lw3_mm.hip, k_lw3_mx.hip): the
// argument block, the tile shapes of the per-atom GEMM chain, its LDS layout and the MFMA
// fragment helpers.  See k_lw3_mm.hip for the method.
#pragma once
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "cplx.h"
#include "spr_internal.h"

namespace spr {

enum { MM_F64 = 0, MM_F32 = 1, MM_BF16 = 2 };

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

struct MmArgs {
    ItemCommon ic;
    const int *item_q0, *item_off, *item_cnt;
    const int2 *pairs;
    const int *q0_grid;
    int L, m0, m1;
    const int *blk_start, *tup_site;
    const void *fc;              // (nterm, 27) R
    const int4 *site;
    int nsite;
    const void *phase;           // (L) complex R
    const void *evec;            // (nq, nd, nd) complex R
    const double *omega, *inv_omega, *nbe;
    double sigma, factor;
    int npmax;                   // most clusters of one basis atom
    int nterm;                   // clusters of all basis atoms
    const double *sel;           // branch selection (NULL: all): (nd) s0 mask, (nd, nd) pair mask [s1][s2]
    // W output (WOUT instantiation, FP64): w_out[q0 - w_q0_first][q1][s0][s1][s2] of every grid
    // q1, both orderings of each reduced pair written
    double *w_out;
    int w_q0_first, w_nq;
    // W output: only the first ordering W(q1; s0, s1, s2) of each pair (items that list every q1
    // of their q0, so that each ordering is evaluated on its own)
    int w_one;
    // s2-per-wave kernel: phase-sum schedule ((Nb + 1) atom offsets, then the entries) and the
    // largest entry count of one atom
    const int *ft_sched;
    int ft_nmax;
    // FP64 per-atom kernel: phase-sum task lists (NULL: every block of the atom by one loop)
    const int *ft2;
    // per-pair weights of caller-supplied items (NULL: 1 for q1 == q2, else 2)
    const double *pair_w;
};

constexpr int pow2ceil(int x) { return x <= 1 ? 1 : x <= 2 ? 2 : x <= 4 ? 4 : 8; }

// V tiles per wave when each unit covers s2 blocks of 8 nts values
constexpr int v_tiles(int nd, int nt, int nts) { return ((nd * 8 * nts + 15) / 16 * pow2ceil(nt) + 7) / 8; }
// largest s2 block (in N tiles of 8 s2) whose V accumulators fit the register budget per lane
// (64; 48 when three N tiles leave a quarter of the waves idle in GEMMs C and D)
constexpr int pick_nts(int nd, int nt, int words) {
    const int budget = nt == 3 ? 48 : 64;
    return v_tiles(nd, nt, nt) * words <= budget ? nt : v_tiles(nd, nt, 2) * words <= budget && nt >= 2 ? 2 : 1;
}
// the next larger s2 block that cuts the number of blocks (fewer repeated phase sums, more
// accumulator registers); equal to `nts` when there is none
constexpr int next_nts(int nt, int nts) {
    for (int k = nts + 1; k <= nt; ++k)
        if ((nt + k - 1) / k < (nt + nts - 1) / nts) return k;
    return nts;
}

// Tile shape VAR: 0 = the register-budget choice above, 1 = next_nts of it.  Both are built;
// the launcher times them on the problem and keeps the faster (k_launch.hip).
template <int MODE, int NB, int VAR> struct MmShape {
    using R = std::conditional_t<MODE == MM_F64, double, float>;
    using Z = cplx_t<R>;
    using ACC = std::conditional_t<MODE == MM_F64, v4d, v4f>;
    static constexpr int ND = 3 * NB, ND2 = ND * ND, K2 = 2 * ND;
    static constexpr int KP = MODE == MM_BF16 ? ((K2 + 31) / 32) * 32 : ((K2 + 15) / 16) * 16;
    static constexpr int S = KP / 4;                        // K per lane group (FP64 / FP32)
    static constexpr int PA = KP + (MODE == MM_F64 ? 2 : 4);    // A row pitch (R)
    static constexpr int PE = KP / 2 + 2;                   // eigenvector row pitch (complex)
    static constexpr int NT = (K2 + 15) / 16;               // N tiles of any (s, q) column set
    static constexpr int MT_B = (3 * ND + 15) / 16;         // GEMM B: rows (alpha, b)
    static constexpr int MB = MT_B * 16;
    // s2 split: a unit covers s2 in [h SB, (h + 1) SB), i.e. NTS N tiles of GEMM B
    static constexpr int NTS0 = pick_nts(ND, NT, MODE == MM_F64 ? 8 : 4);
    static constexpr int NTS = VAR == 0 ? NTS0 : next_nts(NT, NTS0);
    static constexpr int SB = 8 * NTS;
    static constexpr int NSPLIT = (NT + NTS - 1) / NTS;
    static constexpr int NTSB = pow2ceil(NTS);              // waves per M tile, GEMM B
    static constexpr int MT_C = (3 * SB + 15) / 16;         // GEMM C: rows (alpha, s2 local)
    static constexpr int MC = MT_C * 16;
    static constexpr int NTC = pow2ceil(NT);                // waves per M tile, GEMMs C and D
    static constexpr int MT_D = (ND * SB + 15) / 16;        // GEMM D: rows (s1, s2 local)
    static constexpr int VT = (MT_D * NTC + 7) / 8;         // V tiles per wave
    // U2 rows (s1, s2 local) hold 3 complex (alpha); each s1 block of SB rows is padded by one
    // complex so that the GEMM C stores of consecutive s1 fall in different LDS banks
    static constexpr int U2S1 = 3 * SB + 1;
    static constexpr int U2ROWS = (MT_D * 16 + SB - 1) / SB;   // s1 blocks, padding rows included
    static_assert(NT <= 4, "Ndof must be <= 32");
    static_assert(NSPLIT <= kMaxUnitsPerItem, "too many s2 blocks");
    // cluster phases of every basis atom per pair in LDS (FP32 / BF16), else per atom
    static constexpr bool PALL = MODE != MM_F64;
    static constexpr int NE = NB * NB * 27;                 // FT outputs per atom (block, j)
    static constexpr int KE = (NE + 511) / 512;             // FT outputs per thread
};

// LDS layout (bytes), shared by the kernel and the host-side size check.  After the atom loop of
// a pair, everything before the phase table is dead; the FP32 / BF16 delta stage then stages V
// there (kDvTiles 16 x 16 tiles per wave) followed by each wave's list of resonant elements.
constexpr int kDvTiles = 3;

template <int MODE, int NB, int VAR> struct MmLds {
    using Sh = MmShape<MODE, NB, VAR>;
    using R = typename Sh::R;
    using Z = typename Sh::Z;
    static constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
    static constexpr size_t TU = (size_t)Sh::MB * Sh::PA * sizeof(R);
    static constexpr size_t U2 = (size_t)Sh::U2ROWS * Sh::U2S1 * sizeof(Z);
    static constexpr size_t AC = cmax((size_t)Sh::MC * Sh::PA * sizeof(R), (size_t)512 * sizeof(double));
    static constexpr size_t E = (size_t)Sh::ND * Sh::PE * sizeof(Z);
    // delta stage: staged V (in the accumulator type) and the element lists
    using DT = std::conditional_t<MODE == MM_F64, double, float>;
    static constexpr size_t DV = (size_t)8 * kDvTiles * 256 * sizeof(DT);
    static constexpr size_t DL = (size_t)8 * kDvTiles * 128 * sizeof(unsigned short);
    __host__ __device__ static constexpr size_t tu() { return TU; }
    __host__ __device__ static constexpr size_t u2() { return U2; }
    __host__ __device__ static constexpr size_t ac() { return AC; }
    __host__ __device__ static constexpr size_t e() { return E; }
    // cluster phases (all atoms in the FP32 / BF16 modes, atom mu in FP64), padded so that the
    // dead prefix holds the delta stage's buffers
    __host__ __device__ static size_t p_bytes(int npmax, int nterm) {
        size_t b = (size_t)(Sh::PALL ? nterm : npmax) * sizeof(Z);
        const size_t pre = TU + U2 + AC + 2 * E;
        if (pre + b < DV + DL) b = DV + DL - pre;
        return b;
    }
    static size_t total(int npmax, int nterm, int nsite, int L) {
        size_t b = tu() + u2() + ac() + 2 * e();
        b += p_bytes(npmax, nterm) + (size_t)L * sizeof(Z) + 8 * Sh::ND * sizeof(double) +
             4 * Sh::ND * sizeof(float);
        if (Sh::PALL) b += (size_t)npmax * 27 * sizeof(R);      // FC of one atom (staged)
        b += (size_t)Sh::ND * Sh::ND * sizeof(Z) + 16;           // e0 of the unit's q0
        if (!Sh::PALL) b += (size_t)NB * 2 * sizeof(int);        // FP64 phase-sum task-list heads
        b += 2 * (size_t)nsite * sizeof(int) + (size_t)(NB * NB * NB + 1) * sizeof(int);
        return (b + 15) & ~(size_t)15;
    }
};

// orders the LDS accesses of one wave (a wave's LDS operations complete in issue order)
__device__ __forceinline__ void mm_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// output row (within a 16 x 16 tile) of accumulator element r in lane group g
template <bool F64OUT>
__device__ __forceinline__ int tile_row(int g, int r) { return F64OUT ? 4 * r + g : 4 * g + r; }

// 16 x 16 output tile of A (rows in LDS, K contiguous, pitch PA) times the B' expansion of
// the complex eigenvector matrix E (LDS, [s][c] pitch PE): acc[m][n'] += sum_k' A[m][k'] B'[k'][n'].
// The B' fragment is built once per call site and passed in `bf`.
template <int MODE, int NB, int VAR>
struct Frag {
    using Sh = MmShape<MODE, NB, VAR>;
    using R = typename Sh::R;
    using Z = typename Sh::Z;
    // FP64 / FP32: S values per lane; BF16: KP/32 slabs of (hi, lo) bf16x8
    static constexpr int NSLAB = Sh::KP / 32;
    struct BF {
        std::conditional_t<MODE == MM_BF16, bf16x8, R> h[MODE == MM_BF16 ? NSLAB : Sh::S];
        bf16x8 lo[MODE == MM_BF16 ? NSLAB : 1];
    };

    // B'[k'][n'] with n' = (s, q), k' = (c, p): p == q -> Re E(c,s); p=0,q=1 -> Im; p=1,q=0 -> -Im
    __device__ static BF load_b(const Z *sE, int n_base) {
        const int l = threadIdx.x & 63, g = l >> 4;
        const int np = n_base + (l & 15);
        const int s = np >> 1, q = np & 1;
        const bool ok = np < Sh::K2;
        BF b;
        if constexpr (MODE != MM_BF16) {
#pragma unroll
            for (int i = 0; i < Sh::S / 2; ++i) {
                const int c = g * (Sh::S / 2) + i;
                Z e = ok ? sE[s * Sh::PE + c] : Z{0, 0};
                b.h[2 * i] = q ? e.y : e.x;          // p = 0
                b.h[2 * i + 1] = q ? e.x : -e.y;     // p = 1
            }
        } else {
#pragma unroll
            for (int sl = 0; sl < NSLAB; ++sl) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = 16 * sl + 4 * g + i;
                    Z e = ok ? sE[s * Sh::PE + c] : Z{0, 0};
                    const float v0 = q ? e.y : e.x, v1 = q ? e.x : -e.y;
                    const __bf16 h0 = (__bf16)v0, h1 = (__bf16)v1;
                    b.h[sl][2 * i] = h0;
                    b.h[sl][2 * i + 1] = h1;
                    b.lo[sl][2 * i] = (__bf16)(v0 - (float)h0);
                    b.lo[sl][2 * i + 1] = (__bf16)(v1 - (float)h1);
                }
            }
        }
        return b;
    }

    // FP64: two tiles as mma2, with the B' fragment built from E in two halves of K inside the
    // product (16 instead of 32 registers live; the K order, and so the sums, are unchanged)
    __device__ static void mma2_e(const R *sA, int m0, int m1, const Z *sE, int n_base, typename Sh::ACC *d) {
        static_assert(MODE == MM_F64, "FP64 only");
        const int l = threadIdx.x & 63, g = l >> 4;
        const int np = n_base + (l & 15);
        const int s = np >> 1, q = np & 1;
        const bool ok = np < Sh::K2;
        const double2 *p0 = reinterpret_cast<const double2 *>(sA + (m0 + (l & 15)) * Sh::PA + g * Sh::S);
        const double2 *p1 = reinterpret_cast<const double2 *>(sA + (m1 + (l & 15)) * Sh::PA + g * Sh::S);
        typename Sh::ACC a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
        constexpr int HQ = Sh::S / 4;   // complex K values per lane group and half
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            double bv[2 * HQ];
#pragma unroll
            for (int i = 0; i < HQ; ++i) {
                const int c = g * (Sh::S / 2) + hh * HQ + i;
                const Z e = ok ? sE[s * Sh::PE + c] : Z{0, 0};
                bv[2 * i] = q ? e.y : e.x;          // p = 0
                bv[2 * i + 1] = q ? e.x : -e.y;     // p = 1
            }
#pragma unroll
            for (int i = 0; i < HQ; ++i) {
                const double2 x = p0[hh * HQ + i], y = p1[hh * HQ + i];
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, bv[2 * i], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.x, bv[2 * i], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, bv[2 * i + 1], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.y, bv[2 * i + 1], a1, 0, 0, 0);
            }
        }
        d[0] = a0;
        d[1] = a1;
    }

    // FP64: one tile, the same K order as each chain of mma2_e (a wave with an odd tile count
    // issues no product for a repeated tile)
    __device__ static typename Sh::ACC mma1_e(const R *sA, int m0, const Z *sE, int n_base) {
        static_assert(MODE == MM_F64, "FP64 only");
        const int l = threadIdx.x & 63, g = l >> 4;
        const int np = n_base + (l & 15);
        const int s = np >> 1, q = np & 1;
        const bool ok = np < Sh::K2;
        const double2 *p0 = reinterpret_cast<const double2 *>(sA + (m0 + (l & 15)) * Sh::PA + g * Sh::S);
        typename Sh::ACC a0 = {0, 0, 0, 0};
        constexpr int HQ = Sh::S / 4;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            double bv[2 * HQ];
#pragma unroll
            for (int i = 0; i < HQ; ++i) {
                const int c = g * (Sh::S / 2) + hh * HQ + i;
                const Z e = ok ? sE[s * Sh::PE + c] : Z{0, 0};
                bv[2 * i] = q ? e.y : e.x;
                bv[2 * i + 1] = q ? e.x : -e.y;
            }
#pragma unroll
            for (int i = 0; i < HQ; ++i) {
                const double2 x = p0[hh * HQ + i];
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, bv[2 * i], a0, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, bv[2 * i + 1], a0, 0, 0, 0);
            }
        }
        return a0;
    }

    // two tiles (rows m0.., m1..) with the same B fragment; the two chains interleave
    __device__ static void mma2(const R *sA, int m0, int m1, const BF &b, typename Sh::ACC *d) {
        const int l = threadIdx.x & 63, g = l >> 4;
        const R *r0 = sA + (m0 + (l & 15)) * Sh::PA;
        const R *r1 = sA + (m1 + (l & 15)) * Sh::PA;
        typename Sh::ACC a0 = {0, 0, 0, 0}, a1 = {0, 0, 0, 0};
        if constexpr (MODE == MM_F64) {
            const double2 *p0 = reinterpret_cast<const double2 *>(r0 + g * Sh::S);
            const double2 *p1 = reinterpret_cast<const double2 *>(r1 + g * Sh::S);
#pragma unroll
            for (int i = 0; i < Sh::S / 2; ++i) {
                const double2 x = p0[i], y = p1[i];
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.x, b.h[2 * i], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.x, b.h[2 * i], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f64_16x16x4f64(x.y, b.h[2 * i + 1], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f64_16x16x4f64(y.y, b.h[2 * i + 1], a1, 0, 0, 0);
            }
        } else if constexpr (MODE == MM_F32) {
            const float4 *p0 = reinterpret_cast<const float4 *>(r0 + g * Sh::S);
            const float4 *p1 = reinterpret_cast<const float4 *>(r1 + g * Sh::S);
#pragma unroll
            for (int i = 0; i < Sh::S / 4; ++i) {
                const float4 x = p0[i], y = p1[i];
                a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, b.h[4 * i], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.x, b.h[4 * i], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, b.h[4 * i + 1], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.y, b.h[4 * i + 1], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.z, b.h[4 * i + 2], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.z, b.h[4 * i + 2], a1, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x.w, b.h[4 * i + 3], a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_16x16x4f32(y.w, b.h[4 * i + 3], a1, 0, 0, 0);
            }
        } else {
            a0 = mma(sA, m0, b, a0);
            a1 = mma(sA, m1, b, a1);
        }
        d[0] = a0;
        d[1] = a1;
    }

    __device__ static typename Sh::ACC mma(const R *sA, int m_base, const BF &b, typename Sh::ACC acc) {
        const int l = threadIdx.x & 63, g = l >> 4;
        const R *row = sA + (m_base + (l & 15)) * Sh::PA;
        if constexpr (MODE == MM_F64) {
            const double2 *p = reinterpret_cast<const double2 *>(row + g * Sh::S);
#pragma unroll
            for (int i = 0; i < Sh::S / 2; ++i) {
                const double2 a = p[i];
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, b.h[2 * i], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, b.h[2 * i + 1], acc, 0, 0, 0);
            }
        } else if constexpr (MODE == MM_F32) {
            const float4 *p = reinterpret_cast<const float4 *>(row + g * Sh::S);
#pragma unroll
            for (int i = 0; i < Sh::S / 4; ++i) {
                const float4 a = p[i];
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.x, b.h[4 * i], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.y, b.h[4 * i + 1], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.z, b.h[4 * i + 2], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a.w, b.h[4 * i + 3], acc, 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int sl = 0; sl < NSLAB; ++sl) {
                const float4 *p = reinterpret_cast<const float4 *>(row + 32 * sl + 8 * g);
                const float4 x0 = p[0], x1 = p[1];
                const float xs[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
                bf16x8 ah, al;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const __bf16 h = (__bf16)xs[t];
                    ah[t] = h;
                    al[t] = (__bf16)(xs[t] - (float)h);
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.h[sl], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b.lo[sl], acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b.h[sl], acc, 0, 0, 0);
            }
        }
        return acc;
    }
};

// The s2-per-wave FP32 / BF16 kernel (k_lw3_mx.hip) as one more tile-shape candidate of the
// three-phonon launch: false when it does not apply (mode, Nb) or does not fit.
bool mx_launch_of(Ctx &c, int mode, int nb, MmArgs &A, int nitems, ItemLaunch &L);

}  // namespace spr
