// Three-phonon Gaussian linewidths on the matrix cores for the 10-atom cell (Ndof = 30), FP32 and
// split BF16: the same per-atom GEMM chain as k_lw3_mm.hip (FT, then U1 = T E2, U2 = U1 E1,
// V += U2 e0 per basis atom mu; ScatterMatEl, ThirdOrder_m.f90:278-386, factorised), with the
// work split among the waves by s2 instead of by s2 blocks of separate work units:
//
//   - a unit is a whole item (up to 64 pairs of one q0, all s2): the phase sum T of each atom is
//     formed once per pair;
//   - eight waves, two per SIMD; GEMM B (U1 = T E2, all s2) is shared (wave w computes three of
//     the 6 x 4 output tiles, N tile w % 4) and written to LDS as the A operand of GEMM C;
//   - wave w owns s2 in [4w, 4w + 4): GEMM C (rows (alpha, s2 local), 12) and GEMM D (rows
//     (s1, s2 local), 120) run on its own rows only, with U2 in a private LDS slice and V of its s2
//     (128 values per lane) in its accumulator registers across the atoms;
//   - two barriers per atom: [GEMM B(mu) + staging of the next atom's FC, sites, phases] |
//     [FT(mu + 1) + GEMM C(mu) + GEMM D(mu)]; in the second segment the two waves of a SIMD take
//     their parts in opposite orders (waves 0-3: phase sum first, waves 4-7: GEMMs first), so that
//     the phase sum of one runs beside the matrix products of the other;
//   - the phase sum runs over a schedule built once per context: per atom the outputs of its
//     non-empty (nu, eta) blocks sorted by cluster count (the lanes of a wave loop over equal
//     counts), then one entry per block the previous atom wrote and this one does not (cleared);
//   - in BF16 mode the producers of T and U1 (the phase sum, GEMM B) store the split operand
//     (hi and lo bf16 planes), so the GEMMs load their A fragments without conversions; GEMM C
//     loads its A fragment once per atom for all four N tiles.
//
// After the last atom each wave applies the Gaussian delta to its V (W = |V|^2 / (w0 w1 w2) ...;
// CalculateLinewidth, ThirdOrder_m.f90:514-537) exactly as the per-atom kernel's FP32 delta
// stage does: arguments in FP64, triplets beyond sqrt(100) sigma of every resonance dropped, terms
// in FP32 summed per pair and then into a compensated FP32 sum per lane, and the per-s0 totals
// reduced in FP64 in a fixed order.
#include <algorithm>
#include <vector>

#include "k_lw3_mm.h"

namespace spr {

typedef float v16f __attribute__((ext_vector_type(16)));

template <int MODE, int NB> struct MxShape {
    static_assert(MODE != MM_F64, "FP32 / split BF16 only");
    using Sh = MmShape<MODE, NB, 0>;   // K padding, pitches and fragments of the per-atom kernel
    static constexpr int ND = 3 * NB, K2 = 2 * ND, KP = Sh::KP, PA = Sh::PA, PE = Sh::PE, S = Sh::S;
    static constexpr int NW = 8;                       // waves, two per SIMD
    static constexpr int NT = NW * 64;                 // threads
    static constexpr int SBW = 4;                      // s2 per wave
    static constexpr int MT_B = 6;                     // GEMM B: rows alpha * 32 + b
    static constexpr int TROWS = MT_B * 16;            // T rows; b >= Ndof rows stay zero
    static constexpr int U1ROWS = 3 * 32;              // U1 rows alpha * 32 + s2
    static constexpr int MT_C = (3 * SBW + 15) / 16;   // GEMM C: rows alpha * 4 + s2 local
    static constexpr int MT_D = (ND * SBW + 31) / 32;  // GEMM D (32 x 32 tiles): rows s1 * 4 + s2 local
    static constexpr int U2ROWS = MT_D * 32;
    static constexpr int NE = NB * NB * 27;            // FT outputs per atom
    // GEMM operand buffers T and U1 (96 rows each): FP32 rows of pitch PA, or in BF16 mode the
    // split operand x = hi + lo already formed by the producer, as a hi plane and a lo plane of
    // bf16 rows of pitch PB
    static constexpr bool SPLIT = MODE == MM_BF16;
    static constexpr int PB = KP + 8;
    static constexpr int ROWB = SPLIT ? PB * 2 : PA * 4;     // bytes per row
    static constexpr int PLANE = 96 * ROWB;                  // bytes per plane
    static constexpr int OPB = SPLIT ? 2 * PLANE : PLANE;    // bytes per buffer
    static_assert(ND <= NW * SBW && ND <= 32 && MT_C == 1 && KP == 64, "built for 27 <= Ndof <= 32");
};

// LDS byte offsets (kernel and host-side size check)
struct MxLds {
    size_t t, u1, u2, e1, e2, e0, p, fc, tup, sched, ph, mode, modef, m1, m2, atoms, total;
};

template <int MODE, int NB>
__host__ __device__ inline MxLds mx_lds(int npmax, int nsched, int nsite, int L) {
    using X = MxShape<MODE, NB>;
    MxLds o;
    size_t b = 0;
    auto take = [&](size_t n) {
        const size_t at = b;
        b = (b + n + 15) & ~(size_t)15;
        return at;
    };
    o.t = take((size_t)X::OPB);
    o.u1 = take((size_t)X::OPB);
    o.u2 = take((size_t)X::NW * X::U2ROWS * 3 * 8);
    o.e1 = take((size_t)X::ND * X::PE * 8);
    o.e2 = take((size_t)X::ND * X::PE * 8);
    o.e0 = take((size_t)X::ND * X::ND * 8);
    o.p = take((size_t)npmax * 8);
    o.fc = take((size_t)(npmax * 27 + 8) * 4);   // + the aligned superset's extra elements
    o.tup = take((size_t)(npmax * 2 + 8) * 4);
    o.sched = take((size_t)(nsched + 8) * 4);
    o.ph = take((size_t)L * 8);
    o.mode = take((size_t)8 * X::ND * 8);
    o.modef = take((size_t)7 * X::ND * 4);
    o.m1 = take((size_t)nsite * 4);
    o.m2 = take((size_t)nsite * 4);
    o.atoms = take((size_t)2 * (NB + 1) * 4);
    o.total = b;
    return o;
}

// Operand buffer element stores: one value, or two adjacent (k, k + 1); `at` is the byte
// address of the element in the FP32 buffer or in the hi plane
template <int MODE, int NB>
__device__ __forceinline__ void mx_st1(unsigned char *at, float v) {
    using X = MxShape<MODE, NB>;
    if constexpr (X::SPLIT) {
        const __bf16 h = (__bf16)v;
        *reinterpret_cast<__bf16 *>(at) = h;
        *reinterpret_cast<__bf16 *>(at + X::PLANE) = (__bf16)(v - (float)h);
    } else {
        *reinterpret_cast<float *>(at) = v;
    }
}
template <int MODE, int NB>
__device__ __forceinline__ void mx_st2(unsigned char *at, float a, float b) {
    using X = MxShape<MODE, NB>;
    if constexpr (X::SPLIT) {
        typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
        const __bf16 ha = (__bf16)a, hb = (__bf16)b;
        *reinterpret_cast<bf16x2 *>(at) = bf16x2{ha, hb};
        *reinterpret_cast<bf16x2 *>(at + X::PLANE) = bf16x2{(__bf16)(a - (float)ha), (__bf16)(b - (float)hb)};
    } else {
        *reinterpret_cast<float2 *>(at) = make_float2(a, b);
    }
}

// A fragments of NTILE 16-row tiles over the whole K (this lane's row of each given by its byte
// address), loaded once and used with several B fragments: FP32 values, or the hi and lo bf16
// slabs of the split operand
template <int MODE, int NB, int NTILE> struct MxA {
    using X = MxShape<MODE, NB>;
    static constexpr int NSL = X::KP / 32;
    float f[MODE == MM_F32 ? NTILE : 1][MODE == MM_F32 ? X::S : 1];
    bf16x8 h[X::SPLIT ? NTILE : 1][NSL], lo[X::SPLIT ? NTILE : 1][NSL];
    __device__ __forceinline__ void load(const unsigned char *const *row) {
        const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
        for (int t = 0; t < NTILE; ++t) {
            if constexpr (!X::SPLIT) {
#pragma unroll
                for (int i = 0; i < X::S / 4; ++i) {
                    const float4 x = reinterpret_cast<const float4 *>(row[t] + 4 * g * X::S)[i];
                    f[t][4 * i] = x.x;
                    f[t][4 * i + 1] = x.y;
                    f[t][4 * i + 2] = x.z;
                    f[t][4 * i + 3] = x.w;
                }
            } else {
#pragma unroll
                for (int sl = 0; sl < NSL; ++sl) {
                    h[t][sl] = *reinterpret_cast<const bf16x8 *>(row[t] + 2 * (32 * sl + 8 * g));
                    lo[t][sl] = *reinterpret_cast<const bf16x8 *>(row[t] + X::PLANE + 2 * (32 * sl + 8 * g));
                }
            }
        }
    }
    __device__ __forceinline__ void mma(const typename Frag<MODE, NB, 0>::BF &b, v4f *acc) const {
        if constexpr (!X::SPLIT) {
#pragma unroll
            for (int k = 0; k < X::S; ++k)
#pragma unroll
                for (int t = 0; t < NTILE; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(f[t][k], b.h[k], acc[t], 0, 0, 0);
        } else {
#pragma unroll
            for (int sl = 0; sl < NSL; ++sl) {
#pragma unroll
                for (int t = 0; t < NTILE; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h[t][sl], b.h[sl], acc[t], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < NTILE; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h[t][sl], b.lo[sl], acc[t], 0, 0, 0);
#pragma unroll
                for (int t = 0; t < NTILE; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lo[t][sl], b.h[sl], acc[t], 0, 0, 0);
            }
        }
    }
};

// SEL: a branch selection rule is set (A.sel), a compile-time flag so that the common case
// carries no test of it
template <int MODE, int NB, bool SEL>
__device__ __forceinline__ void mx_unit(const MmArgs &A, const int unit) {
    using X = MxShape<MODE, NB>;
    using F = Frag<MODE, NB, 0>;
    using Z = cplx_t<float>;
    constexpr int ND = X::ND, PA = X::PA, PE = X::PE;

    extern __shared__ __align__(16) unsigned char smem[];
    const MxLds o = mx_lds<MODE, NB>(A.npmax, A.ft_nmax, A.nsite, A.L);
    unsigned char *sT = smem + o.t;      // T rows alpha * 32 + b, K = (c, p)
    unsigned char *sU1 = smem + o.u1;    // U1 rows alpha * 32 + s2, K = (b, p)
    Z *sE1 = reinterpret_cast<Z *>(smem + o.e1);            // E1[s1][b], pitch PE
    Z *sE2 = reinterpret_cast<Z *>(smem + o.e2);            // E2[s2][c], pitch PE
    Z *sE0 = reinterpret_cast<Z *>(smem + o.e0);            // e0[s0][a]
    Z *sP = reinterpret_cast<Z *>(smem + o.p);              // cluster phases of one atom
    float *sFC = reinterpret_cast<float *>(smem + o.fc);    // FC of one atom
    int *sTup = reinterpret_cast<int *>(smem + o.tup);      // site pairs (N1, N2) of its clusters
    Z *sPh = reinterpret_cast<Z *>(smem + o.ph);
    double *sMode = reinterpret_cast<double *>(smem + o.mode);   // w0 iw0f | w1 n1 iw1 | w2 n2 iw2 | -
    float *sModeF = reinterpret_cast<float *>(smem + o.modef);   // n1 iw1 n2 iw2 w0 w1 w2 (FP32)
    int *sM1 = reinterpret_cast<int *>(smem + o.m1);
    int *sM2 = reinterpret_cast<int *>(smem + o.m2);
    unsigned *sSched = reinterpret_cast<unsigned *>(smem + o.sched);   // phase-sum schedule of one atom
    // per atom: first cluster and first schedule entry (Nb + 1 each), read from LDS so that the
    // staging issues no global load of its own (a wait for that load would also wait for the
    // direct-to-LDS loads issued before it)
    int *aU = reinterpret_cast<int *>(smem + o.atoms);
    int *schOff = aU + NB + 1;
    const unsigned *sch = reinterpret_cast<const unsigned *>(A.ft_sched + NB + 1);
    double *sRed = reinterpret_cast<double *>(smem + o.t);   // end of unit only

    constexpr int NT = X::NT;
    const int tid = threadIdx.x, l = tid & 63, g = l >> 4;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    Z *sU2 = reinterpret_cast<Z *>(smem + o.u2) + (size_t)w * X::U2ROWS * 3;   // this wave's U2[(s1, s2l)][alpha]
    const int qi = A.item_q0[unit];
    const int i0 = A.q0_grid[qi];
    const int cnt = A.item_cnt[unit];
    const int poff = A.item_off[unit];
    const float *fc = reinterpret_cast<const float *>(A.fc);
    const Z *phase = reinterpret_cast<const Z *>(A.phase);
    const Z *evec = reinterpret_cast<const Z *>(A.evec);
    __syncthreads();   // the previous unit of a persistent workgroup is done with LDS

    // ---- per unit: zero the operand buffers (padding stays zero), phase table, q0 modes ----
    for (int i = tid; i < (int)((o.e1 - o.t) / 16); i += NT)
        reinterpret_cast<uint4 *>(smem + o.t)[i] = make_uint4(0u, 0u, 0u, 0u);   // sT, sU1, sU2
    for (int i = tid; i < ND * PE; i += NT) {
        if (i % PE >= ND) {
            sE1[i] = Z{0.f, 0.f};
            sE2[i] = Z{0.f, 0.f};
        }
    }
    for (int i = tid; i < A.L; i += NT) sPh[i] = phase[i];
    for (int i = tid; i <= NB; i += NT) {
        aU[i] = A.blk_start[i * NB * NB];
        schOff[i] = A.ft_sched[i];
    }
    for (int i = tid; i < ND * ND; i += NT) sE0[i] = evec[(size_t)i0 * ND * ND + i];
    if (tid < ND) {
        sMode[tid] = A.omega[(size_t)i0 * ND + tid];
        sMode[ND + tid] = A.inv_omega[(size_t)i0 * ND + tid] * A.factor;
        sModeF[4 * ND + tid] = (float)A.omega[(size_t)i0 * ND + tid];
    }

    // GEMM B: N tile w % 4 of (s2, q), M tiles w / 4 + 2 t (t = 0, 1, 2) of (alpha, b); this
    // lane's A row in tile 0 and its U1 store base (row alpha * 32 + s2, column 2 b + q)
    const int nB = w & 3, mB0 = w >> 2;
    const int npB = nB * 16 + (l & 15);
    const int s2B = npB >> 1, qB = npB & 1;
    constexpr int RB = X::ROWB, EB = X::SPLIT ? 2 : 4;   // bytes per row, per element
    const unsigned char *rowB = sT + (l & 15) * RB;
    unsigned char *u1st = sU1 + s2B * RB + EB * (8 * g + qB);
    // GEMM C: row (alpha, s2 local) = l % 16 of this wave; alpha = 3 reads a zero row of T
    const unsigned char *rowC[1];
    {
        const int r = l & 15, al = r >> 2, s2 = 4 * w + (r & 3);
        rowC[0] = al < 3 ? sU1 + (al * 32 + s2) * RB : sT + 31 * RB;   // T row 31 (b = 31) is zero
    }
    // GEMM C stores: U2 float ((s1 * 4 + s2l) * 3 + alpha) * 2 + q
    float *u2st = reinterpret_cast<float *>(sU2) + ((l & 15) >> 1) * 24 + 2 * g + (l & 1);
    const int qD = l & 1;
    KSum<float> accf[2];

    // Elements [first, first + n) of a 4-byte array (readable for 3 elements past its end) into
    // dst by 16-byte direct global-to-LDS loads of the aligned superset starting at first & ~3
    // (1 KB per wave and chunk; wave w takes chunks w, w + 8, ...); element first lands at
    // dst[first & 3]
    auto dma16 = [&](const void *src, int first, int n, void *dst) {
        const int a0 = first & ~3, m = first + n - a0;
        const unsigned char *gs = reinterpret_cast<const unsigned char *>(src) + (size_t)a0 * 4;
        unsigned char *ld = reinterpret_cast<unsigned char *>(dst);
        for (int c = w; c * 256 < m; c += X::NW)
            if (c * 256 + 4 * l < m) __builtin_amdgcn_global_load_lds(gs + c * 1024 + 16 * l, ld + c * 1024, 16, 0, 0);
    };
    // FC, cluster sites and schedule of atom mu -> sFC, sTup, sSched
    auto stage = [&](int mu) {
        const int u0 = aU[mu], nu = aU[mu + 1] - u0;
        dma16(fc, u0 * 27, nu * 27, sFC);
        dma16(A.tup_site, 2 * u0, 2 * nu, sTup);
        dma16(sch, schOff[mu], schOff[mu + 1] - schOff[mu], sSched);
    };
    // P(u) = e^{2 pi i (m1(N1) + m2(N2)) / L} of the clusters whose sites this wave staged (its chunk
    // c holds sites 256 c .. 256 c + 255 of the superset, i.e. clusters (256 c - d) / 2 + 0 .. 127,
    // d = (2 u0) & 3), once its own staging has landed: no barrier needed
    auto phases = [&](int mu) {
        const int u0 = aU[mu], nu = aU[mu + 1] - u0, d = (2 * u0) & 3;
        const int *tp = sTup + d;
        for (int c = w; c * 256 < 2 * nu + d; c += X::NW) {
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int u = ((256 * c - d) >> 1) + l + 64 * k;
                if (u >= 0 && u < nu) {
                    int m = sM1[tp[2 * u]] + sM2[tp[2 * u + 1]];
                    m -= m >= A.L ? A.L : 0;
                    sP[u] = sPh[m];
                }
            }
        }
    };
    // T[(alpha, 3 nu + beta)][2 (3 eta + gamma) + p] of atom mu from its schedule: entry = block
    // (nu, eta) | component j << 7 | first cluster (within the atom) << 12 | cluster count << 23;
    // lanes run along the components of a block (contiguous FC reads); component 31 clears all 27
    // outputs of the block
    auto ft = [&](int mu) {
        const int ns = schOff[mu + 1] - schOff[mu];
        const unsigned *ss = sSched + (schOff[mu] & 3);
        const float *sf = sFC + ((aU[mu] * 27) & 3);
        for (int e = tid; e < ns; e += NT) {
            const unsigned code = ss[e];
            const int bl = code & 127, j = (code >> 7) & 31, uo = (code >> 12) & 2047, nc = code >> 23;
            const int nu = bl / NB, eta = bl % NB;
            if (j == 31) {
#pragma unroll
                for (int r = 0; r < 9; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        mx_st2<MODE, NB>(sT + ((r / 3) * 32 + 3 * nu + r % 3) * RB + EB * 2 * (3 * eta + c), 0.f, 0.f);
                continue;
            }
            const int al = j / 9, be = (j / 3) % 3, ga = j % 3;
            const float *fj = sf + uo * 27 + j;
            const Z *pj = sP + uo;
            float re = 0.f, im = 0.f;
#pragma unroll 4
            for (int u = 0; u < nc; ++u) {
                const float f = fj[u * 27];
                const Z ph = pj[u];
                re = fmaf(f, ph.x, re);
                im = fmaf(f, ph.y, im);
            }
            mx_st2<MODE, NB>(sT + (al * 32 + 3 * nu + be) * RB + EB * 2 * (3 * eta + ga), re, im);
        }
    };

    for (int pi = 0; pi < cnt; ++pi) {
        const int2 pr = A.pairs[poff + pi];
        const int q1 = pr.x, q2 = pr.y;
        __syncthreads();   // the previous pair is done with sE1, sE2, sMode, sM1, sM2, sP, sFC, sT
        // the previous pair's delta stage staged its tiles over T and U1: restore their zero padding
        for (int i = tid; i < X::NW * 1280 / 4; i += NT)
            reinterpret_cast<uint4 *>(smem + o.t)[i] = make_uint4(0u, 0u, 0u, 0u);
        for (int i = tid; i < ND * ND; i += NT) {
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
            sModeF[5 * ND + tid] = (float)A.omega[(size_t)q1 * ND + tid];
            sModeF[6 * ND + tid] = (float)A.omega[(size_t)q2 * ND + tid];
        }
        {
            const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
            const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
            for (int s = tid; s < A.nsite; s += NT) {
                const int4 c = A.site[s];
                sM1[s] = (k10 * c.x + k11 * c.y + k12 * c.z) % A.L;
                sM2[s] = (k20 * c.x + k21 * c.y + k22 * c.z) % A.L;
            }
        }
        stage(0);
        __builtin_amdgcn_s_waitcnt(0);   // this wave's staging has landed
        __syncthreads();                 // sM1, sM2 are complete
        phases(0);
        __syncthreads();
        ft(0);
        v16f V[X::MT_D][2];
#pragma unroll
        for (int i = 0; i < X::MT_D; ++i)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt)
#pragma unroll
                for (int e = 0; e < 16; ++e) V[i][nt][e] = 0.f;
        __syncthreads();   // T of atom 0 is complete

#pragma unroll 1
        for (int mu = 0; mu < NB; ++mu) {
            // ---- GEMM B (shared): U1[(alpha, b)][(s2, q)] = T x E2' -> sU1[(alpha, s2)][(b, q)];
            //      the next atom's FC, sites and schedule are staged meanwhile ----
            {
                const unsigned char *rb[3];
#pragma unroll
                for (int t = 0; t < 3; ++t) rb[t] = rowB + (mB0 + 2 * t) * 16 * RB;
                v4f d[3] = {v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}};
                {
                    // the operands are read before the next atom's staging is issued: an LDS read
                    // after a direct-to-LDS load waits for that load
                    const typename F::BF bE2 = F::load_b(sE2, nB * 16);
                    MxA<MODE, NB, 3> a;
                    a.load(rb);
                    if (mu + 1 < NB) stage(mu + 1);
                    a.mma(bE2, d);
                }
                // output rows mt 16 + 4 g + r (mt = mB0 + 2 t): alpha = mt / 2, b = 16 (mt & 1) + 4 g + r
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int mt = mB0 + 2 * t, al = mt >> 1, bh = 16 * (mt & 1);
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (s2B < ND && bh + 4 * g + r < ND)
                            mx_st1<MODE, NB>(u1st + al * 32 * RB + EB * 2 * (bh + r), d[t][r]);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);   // this wave's staging of the next atom has landed
            if (mu + 1 < NB) phases(mu + 1);
            __syncthreads();
            // ---- GEMM C (this wave's s2): U2[(alpha, s2l)][(s1, q)] = U1' x E1' -> sU2[(s1, s2l)][alpha];
            //      GEMM D (this wave's s2): V[(s1, s2l)][(s0, q)] += U2[.][(alpha, p)] E0'[(alpha, p)][(s0, q)] ----
            auto gemm_cd = [&]() {
                {
                    MxA<MODE, NB, 1> a;
                    a.load(rowC);
#pragma unroll
                    for (int n = 0; n < 4; ++n) {
                        const typename F::BF bE1 = F::load_b(sE1, n * 16);
                        v4f d[1] = {v4f{0.f, 0.f, 0.f, 0.f}};
                        a.mma(bE1, d);
                        // output rows 4 g + r: alpha = g, s2l = r; column (s1, q), s1 = 8 n + (l % 16) / 2
                        if (g < 3 && 8 * n + ((l & 15) >> 1) < ND)
#pragma unroll
                            for (int r = 0; r < 4; ++r) u2st[n * 8 * 24 + r * 6] = d[0][r];
                    }
                }
                mm_wave_sync();   // this wave's U2 stores precede its GEMM D loads
                // v_mfma_f32_32x32x2f32, K = 2 per instruction: one instruction per alpha with
                // K = (alpha, p), p = l / 32; lane l: A[m = l % 32][p], B[p][n = l % 32]; B' =
                // [[Re e0, Im e0], [-Im e0, Re e0]] on (p, q) with column (s0, q), s0 = 16 nt + (l % 32) / 2
                const int hp = l >> 5;
                float bd[3][2];
#pragma unroll
                for (int al = 0; al < 3; ++al)
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt) {
                        const int s0 = 16 * nt + ((l & 31) >> 1), q = l & 1;
                        bd[al][nt] = 0.f;
                        if (s0 < ND) {
                            const Z e = sE0[s0 * ND + 3 * mu + al];
                            bd[al][nt] = hp == 0 ? (q ? e.y : e.x) : (q ? e.x : -e.y);
                        }
                    }
#pragma unroll
                for (int i = 0; i < X::MT_D; ++i) {
                    const int m = i * 32 + (l & 31);
#pragma unroll
                    for (int al = 0; al < 3; ++al) {
                        const Z a = sU2[m * 3 + al];
                        const float av = hp ? a.y : a.x;
#pragma unroll
                        for (int nt = 0; nt < 2; ++nt)
                            V[i][nt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bd[al][nt], V[i][nt], 0, 0, 0);
                    }
                }
                mm_wave_sync();   // GEMM D's loads of U2 precede the next atom's GEMM C stores
            };
            // the two waves of a SIMD (w and w + 4) take the phase sum of the next atom and their
            // GEMMs in opposite orders (one loop, so that each part is emitted once)
#pragma unroll 1
            for (int ph = 0; ph < 2; ++ph) {
                if ((ph == 0) == (w < 4)) {
                    if (mu + 1 < NB) ft(mu + 1);
                } else {
                    gemm_cd();
                }
            }
            __syncthreads();   // T of atom mu + 1 is complete; sU1 is free
        }

        // ---- Gaussian-delta weights of this wave's V (T and U1 are free: every wave is past its
        //      last GEMM C).  Tile (i, nt), element 4 a + b of lane l: row m = 32 i + 8 a + 4 h + b
        //      (h = l / 32): s1 = 8 i + 2 a + h, s2 = 4 w + b; column (s0, q), s0 = 16 nt + (l % 32) / 2,
        //      q = l & 1.  Lanes l and l^1 hold Re and Im of the same element and the lane with
        //      q = qD owns b = 2 qD + k ("its elements").  Per tile each wave (1) stages the tile in
        //      LDS, (2) tests its elements in FP32 with a margin over the FP32 rounding of the
        //      frequencies, (3) packs the candidates into a list (ballot order), (4) evaluates the
        //      list with every lane -- arguments in FP64, a triplet with all three Gaussians below
        //      e^-50 of their peak adds nothing and is dropped, the rest in FP32 -- writing each term
        //      over the staged value of its element, and (5) each owner adds its elements' terms in
        //      element order, so the sums do not depend on the packing ----
        const double c0 = 1.0 / (A.sigma * 2.5066282746310002);   // 1/(sigma sqrt(2 pi))
        const double hinv = 0.5 / (A.sigma * A.sigma);
        const float hneg = (float)(-hinv);
        const double dcut = A.sigma * 10.0;                 // sqrt(2 x 50) sigma
        const float dcutf = (float)dcut + 1e-3f;            // FP32 pre-test: a superset of the FP64 test
        float *sV = reinterpret_cast<float *>(smem + o.t) + w * 1280;   // 4 x 64 x 4 staged values (40 KB of T, U1)
        unsigned short *sList = reinterpret_cast<unsigned short *>(sV + 1024);   // up to 512 codes
        const int hh = l >> 5;
        float pairf[2] = {0.f, 0.f};
        float w2f[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int s2 = 4 * w + 2 * qD + k;
            w2f[k] = s2 < ND ? sModeF[6 * ND + s2] : 1e30f;   // 1e30: never near
        }
#pragma unroll
        for (int i = 0; i < X::MT_D; ++i)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                const int s0 = 16 * nt + ((l & 31) >> 1);
                const float w0f = s0 < ND ? sModeF[4 * ND + s0] : 1e30f;
                // (1) stage
#pragma unroll
                for (int a = 0; a < 4; ++a)
                    *reinterpret_cast<v4f *>(sV + (a * 64 + l) * 4) =
                        v4f{V[i][nt][4 * a], V[i][nt][4 * a + 1], V[i][nt][4 * a + 2], V[i][nt][4 * a + 3]};
                mm_wave_sync();
                // (2) + (3) test and pack
                int cnt = 0;
                unsigned act = 0;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int s1 = 8 * i + 2 * a + hh;
                    const float w1f = s1 < ND ? sModeF[5 * ND + s1] : 1e30f;
#pragma unroll
                    for (int k = 0; k < 2; ++k) {
                        const float ap = w1f + w2f[k], am = w1f - w2f[k];
                        const float dm = fminf(fabsf(w0f - ap), fminf(fabsf(w0f + am), fabsf(w0f - am)));
                        const bool near = dm < dcutf;
                        const unsigned long long msk = __ballot(near);
                        if (near) {
                            const int pos = cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(msk >> 32),
                                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)msk, 0u));
                            sList[pos] = (unsigned short)((a << 7) | (l << 1) | k);
                            act |= 1u << (2 * a + k);
                        }
                        cnt += __popcll(msk);
                    }
                }
                mm_wave_sync();
                // (4) evaluate: the term replaces the staged value of its element (read by this lane only)
                for (int idx = l; idx < cnt; idx += 64) {
                    const int code = sList[idx];
                    const int a = code >> 7, lo = (code >> 1) & 63, k = code & 1;
                    const int b = 2 * (lo & 1) + k;
                    const float x = sV[(a * 64 + lo) * 4 + b], y = sV[(a * 64 + (lo ^ 1)) * 4 + b];
                    const int s0e = 16 * nt + ((lo & 31) >> 1);
                    const int s1e = 8 * i + 2 * a + (lo >> 5);
                    const int s2 = 4 * w + b;
                    const double w0 = sMode[s0e], w1 = sMode[2 * ND + s1e], w2 = sMode[5 * ND + s2];
                    const double ap = w1 + w2, am = w1 - w2;
                    const double d1 = w0 - ap, d2 = w0 + am, d3 = w0 - am;
                    float term = 0.f;
                    if (fabs(d1) < dcut || fabs(d2) < dcut || fabs(d3) < dcut) {
                        const float f1 = (float)d1, f2 = (float)d2, f3 = (float)d3;
                        const float n1 = sModeF[s1e], iw1 = sModeF[ND + s1e];
                        const float n2 = sModeF[2 * ND + s2], iw2 = sModeF[3 * ND + s2];
                        const float W = (x * x + y * y) * (iw1 * iw2);
                        const float e1 = __expf(hneg * f1 * f1), e2 = __expf(hneg * f2 * f2), e3 = __expf(hneg * f3 * f3);
                        const float sm = SEL ? (float)A.sel[ND + s1e * ND + s2] : 1.f;
                        term = W * sm * ((n1 + n2 + 1.f) * e1 + (n1 - n2) * (e2 - e3));
                    }
                    sV[(a * 64 + lo) * 4 + b] = term;
                }
                mm_wave_sync();
                // (5) the owner adds its elements' terms in element order
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        if (act & (1u << (2 * a + k))) pairf[nt] += sV[(a * 64 + l) * 4 + 2 * qD + k];
                mm_wave_sync();   // this wave has read the tile before the next one overwrites it
            }
        const double wpair = A.pair_w ? A.pair_w[poff + pi] : q1 == q2 ? 1.0 : 2.0;
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            const int s0 = 16 * nt + ((l & 31) >> 1);
            if (s0 < ND) accf[nt].add(pairf[nt] * (float)(wpair * c0 * sMode[ND + s0]));
        }
    }
    // ---- reduce per s0 in a fixed order: lane l of every wave holds s0 = 16 nt + (l % 32) / 2 ----
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) sRed[nt * NT + tid] = accf[nt].value();
    __syncthreads();
    if (tid < ND) {
        const int s = tid;
        const int nt = s >> 4, sj = s & 15;
        double tot = 0.0;
        for (int ww = 0; ww < X::NW; ++ww)
            for (int h2 = 0; h2 < 2; ++h2)
                for (int q = 0; q < 2; ++q) tot += sRed[nt * NT + ww * 64 + h2 * 32 + 2 * sj + q];
        A.ic.partial[(size_t)unit * ND + s] = SEL ? tot * A.sel[s] : tot;
    }
}

template <int MODE, int NB, bool SEL>
__global__ __launch_bounds__(512) void k_lw3_mx(MmArgs A) {
    // one unit per workgroup, or (persistent) units from the atomic counter until none is left
    __shared__ int s_unit;
    for (int it = 0;; ++it) {
        int unit;
        if (!A.ic.counter) {
            if (it > 0) break;
            unit = blockIdx.x;
        } else {
            __syncthreads();
            if (threadIdx.x == 0) s_unit = atomicAdd(A.ic.counter, 1);
            __syncthreads();
            unit = s_unit;
            if (unit >= A.ic.nitems) break;
        }
        mx_unit<MODE, NB, SEL>(A, unit);
    }
}

// The phase-sum schedule of every atom (see the kernel's header); false when the FC3 layout does
// not fit its entry format (2048 clusters per atom, 255 per block).
static bool mx_schedule(Ctx &c, int nb) {
    DevIFC &f = c.fc3;
    if (f.ft_tried) return f.ft_ok;
    f.ft_tried = true;
    const std::vector<int> &bs = f.h_blk_start;
    const int nb2 = nb * nb;
    if ((int)bs.size() < nb * nb2 + 1 || nb2 > 128) return false;
    std::vector<int> off(nb + 1), ent;
    int nmax = 0;
    for (int mu = 0; mu < nb; ++mu) {
        off[mu] = (int)ent.size();
        const int pm = (mu + nb - 1) % nb;   // the atom before (for atom 0: the last atom of the previous pair)
        const int u0 = bs[mu * nb2];
        std::vector<std::pair<int, unsigned>> v;
        for (int bl = 0; bl < nb2; ++bl) {
            const int b = mu * nb2 + bl, cnt = bs[b + 1] - bs[b], uo = bs[b] - u0;
            const bool prev = bs[pm * nb2 + bl + 1] > bs[pm * nb2 + bl];
            if (cnt > 255 || uo > 2047) return false;
            if (cnt == 0) {
                if (prev) v.push_back({0, (unsigned)bl | 31u << 7});   // clear the whole block
                continue;
            }
            for (int j = 0; j < 27; ++j)
                v.push_back({cnt, (unsigned)bl | (unsigned)j << 7 | (unsigned)uo << 12 | (unsigned)cnt << 23});
        }
        std::sort(v.begin(), v.end(), [](const std::pair<int, unsigned> &x, const std::pair<int, unsigned> &y) {
            return x.first != y.first ? x.first > y.first : x.second < y.second;
        });
        for (auto &e : v) ent.push_back((int)e.second);
        nmax = std::max(nmax, (int)v.size());
    }
    off[nb] = (int)ent.size();
    std::vector<int> h(off);
    h.insert(h.end(), ent.begin(), ent.end());
    h.resize(h.size() + 4, 0);   // readable past the end by the 16-byte staging
    if (hipMalloc((void **)&f.ft_sched, h.size() * sizeof(int)) != hipSuccess) {
        f.ft_sched = nullptr;
        return false;
    }
    if (hipMemcpy(f.ft_sched, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return false;
    f.ft_nmax = std::max(nmax, 1);
    f.ft_ok = true;
    return true;
}

template <int MODE, int NB>
static bool mx_launch_of_t(Ctx &c, MmArgs &A, int nitems, ItemLaunch &L) {
    if (!mx_schedule(c, NB)) return false;
    A.ft_sched = c.fc3.ft_sched;
    A.ft_nmax = c.fc3.ft_nmax;
    const size_t dyn = mx_lds<MODE, NB>(A.npmax, A.ft_nmax, A.nsite, A.L).total;
    if (dyn > 160 * 1024 - 64) return false;
    const void *fn = A.sel ? (const void *)k_lw3_mx<MODE, NB, true> : (const void *)k_lw3_mx<MODE, NB, false>;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) return false;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, MxShape<MODE, NB>::NT, dyn) != hipSuccess || per_cu < 1)
        return false;
    L = ItemLaunch{fn, &A, sizeof(A), MxShape<MODE, NB>::NT, dyn, nitems, per_cu,
                   std::string(MODE == MM_F32 ? "mx_f32" : "mx_bf16")};
    L.upi = 1;
    L.shape = "s2wave4";
    return true;
}

bool mx_launch_of(Ctx &c, int mode, int nb, MmArgs &A, int nitems, ItemLaunch &L) {
    if (nb != 10 || (c.override_items && !c.override_items->dense)) return false;
    return mode == MM_F32 ? mx_launch_of_t<MM_F32, 10>(c, A, nitems, L)
           : mode == MM_BF16 ? mx_launch_of_t<MM_BF16, 10>(c, A, nitems, L)
                             : false;
}

}  // namespace spr
