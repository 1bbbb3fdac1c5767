// Four-phonon Gaussian linewidths on the matrix cores in FP32 (built for Ndof = 6).
//
// Work item: one irreducible q0 and one q1 (every q1 of the mesh).  The 8 waves of the
// workgroup take the q2 of the mesh in turn, q3 = -(q0 + q1 + q2) (MomentumConsrv4th,
// QuasiMomentumConsrv4th_m.f90:32-206), one triple per permutation class (q1 <= q2 <= q3) when
// no branch rule is set; each wave works on its own (q0, q1, q2) with no workgroup barrier
// between them.  With the index pairs ab = (a, b), cd = (c, d),
// s01 = (s0, s1), s23 = (s2, s3), per (q0, q1, q2):
//   FT4     T[cd][ab]   = sum_u Phi_u(a, b, c, d) e^{2 pi i m_u / L}     (ScatterMatEl4,
//           FourthOrder_m.f90:286-422; VALU, lanes along the FC4 components of a cluster, the
//           cluster phases of the triple computed once per wave)
//   GEMM 1  Y[cd][s01]  = sum_ab T[cd][ab] e0(a, s0) e1(b, s1)           (K = Ndof^2 complex)
//   GEMM 2  V[s23][s01] = sum_cd e2(c, s2) e3(d, s3) Y[cd][s01]          (K = Ndof^2 complex)
// (no eigenvector conjugated), then W4 = |V|^2 prod(1/w) Sctr4THz2/Nq^2 and the three processes
//   W4 [ 1/6 n1 n2 n3/n0 d(w0-w1-w2-w3) + 1/2 (1+n1) n2 n3/n0 d(w0+w1-w2-w3)
//        + 1/2 (1+n1)(1+n2) n3/n0 d(w0+w1+w2-w3) ]
// of SumOverPolarization / FindFandOmegaVer_4th (Tetrahedron4th_sep.f90:174-425) with Gaussian
// deltas; s0 < 3 is skipped at q0 = Gamma.  Same sum as k_lw4, which runs it in FP64 on the
// vector ALU.
//
// The complex GEMMs run as real GEMMs on v_mfma_f32_16x16x4_f32 (lane l: A[l%16][k] and
// B[k][l%16], lane group g = l/16 walks its own contiguous quarter of K; output rows 4 g + r)
// with re/im interleaved along K, the eigenvector operand expanded: in GEMM 1 as the B operand
// [[Re, Im], [-Im, Re]] (output columns (s01, re/im)), in GEMM 2 as the A operand
// [[Re, -Im], [Im, Re]] (output rows (s23, re/im), so a lane holds whole complex values).
// GEMM 1 consumes T one M tile (16 rows cd) at a time and writes those rows of Y in place,
// so T and Y share one tile per wave.  e0 x e1 is the same for every triple of the item: its
// B fragments are tabulated once per item in LDS; e2 x e3 is formed while the A fragments
// are built.  The delta
// arguments are formed in FP64, the weighted terms summed with a compensated FP32 sum per
// lane, and the per-s0 totals reduced in FP64 in a fixed order.
#include <algorithm>
#include <cmath>
#include <string>

#include "cplx.h"
#include "spr_internal.h"

namespace spr {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

struct Lw4MmArgs {
    const int *q0_grid;
    int nq, L, m0, m1, m2;
    const int *blk_start, *tup_site;
    const float *fc;           // (nterm, 81)
    int nterm;
    const int4 *site;
    int nsite;
    const float2 *phase;       // (L)
    const float2 *evec;        // (nq, nd, nd): e(c, s) = evec[q][s nd + c]
    const double *omega, *inv_omega, *nbe;
    double sigma, factor;
    double *partial;           // [q0][q1][nd]
    const double *sel;         // branch selection (NULL: all): (4, nd) masks of s0, s1, s2, s3
};

template <int ND> struct M4 {
    static constexpr int NB = ND / 3;
    static constexpr int NN = ND * ND;                 // complex K of both GEMMs
    static constexpr int K2 = 2 * NN;                  // real K, and real N
    static constexpr int KP = (K2 + 15) / 16 * 16;     // K padded to the 4-lane-group split
    static constexpr int S = KP / 4;                   // K values per lane group
    static constexpr int PA = KP + 4;                  // row pitch of T / Y (floats)
    static constexpr int MT1 = (NN + 15) / 16;         // GEMM 1: rows cd
    static constexpr int NT1 = (K2 + 15) / 16;         //         columns (s01, re/im)
    static constexpr int MT2 = (K2 + 15) / 16;         // GEMM 2: rows (s23, re/im)
    static constexpr int NT2 = (NN + 15) / 16;         //         columns s01
    static constexpr int NW = 8;                       // waves per workgroup
    static constexpr int NBLK = NB * NB * NB * NB;
    static_assert(S % 4 == 0, "A fragments are read as float4");
};

constexpr size_t al16(size_t b) { return (b + 15) & ~(size_t)15; }

// LDS layout, shared by the kernel and the host-side size check
template <int ND> struct M4Lds {
    using Sh = M4<ND>;
    size_t tile, e01, e, md, m1, ph, blk, red, wave, total;
    // e01: the GEMM 1 B fragments of e0 x e1, per (N tile, lane) S contiguous floats
    __host__ __device__ M4Lds(int nterm, int nsite, int L) {
        // per-wave T / Y tiles (NN rows each)
        tile = al16((size_t)Sh::NW * Sh::NN * Sh::PA * sizeof(float));
        e01 = al16((size_t)Sh::NT1 * 64 * Sh::S * sizeof(float));
        e = al16((size_t)2 * Sh::NN * sizeof(float2));
        md = al16((size_t)6 * ND * sizeof(double));
        m1 = al16((size_t)nsite * sizeof(int));
        ph = al16((size_t)L * sizeof(float2));
        blk = al16((size_t)(Sh::NBLK + 1) * sizeof(int));
        red = al16((size_t)Sh::NW * ND * sizeof(double));
        // per wave: e2 e3 | column constants of s23 (double2, float4) | site phase indices of
        // q2 q3 | cluster phases
        wave = e + al16((size_t)Sh::NN * 32) + al16((size_t)2 * nsite * sizeof(int)) +
               al16((size_t)nterm * sizeof(float2));
        total = tile + e01 + e + md + m1 + ph + blk + red + Sh::NW * wave;
    }
};

// orders the LDS accesses of one wave (a wave's LDS operations complete in issue order)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// B fragment element j of lane l, N tile nt, of the expansion of E[(s0, s1)][(a, b)] =
// e0(a, s0) e1(b, s1): B'[(ab, p)][(s01, q)] = Re E for p == q, Im E for (p, q) = (0, 1),
// -Im E for (1, 0)
template <int ND>
__device__ __forceinline__ float b_e01(const float2 *e0, const float2 *e1, int nt, int l, int j) {
    using Sh = M4<ND>;
    const int g = l >> 4;
    const int np = nt * 16 + (l & 15);
    const int s = np >> 1, q = np & 1;
    const int c = g * (Sh::S / 2) + j / 2, p = j & 1;
    if (np >= Sh::K2 || c >= Sh::NN) return 0.f;
    const float2 e = cmulf(e0[(s / ND) * ND + c / ND], e1[(s % ND) * ND + c % ND]);
    return p == q ? e.x : (q ? e.y : -e.y);
}

// the tabulated B fragment of N tile nt for this lane
template <int ND>
__device__ __forceinline__ void b_table(const float *sB, int nt, float (&b)[M4<ND>::S]) {
    using Sh = M4<ND>;
    const int l = threadIdx.x & 63;
    const float4 *p = reinterpret_cast<const float4 *>(sB + ((size_t)nt * 64 + l) * Sh::S);
#pragma unroll
    for (int i = 0; i < Sh::S / 4; ++i) {
        const float4 v = p[i];
        b[4 * i] = v.x;
        b[4 * i + 1] = v.y;
        b[4 * i + 2] = v.z;
        b[4 * i + 3] = v.w;
    }
}

// A fragment (rows mt*16.., lane row l%16) of the expansion of E[(s2, s3)][(c, d)] =
// e2(c, s2) e3(d, s3): A'[(s23, po)][(cd, pi)] = Re E for po == pi, -Im E for (0, 1), Im E for (1, 0)
template <int ND>
__device__ __forceinline__ void a_kron(const float2 *e2, const float2 *e3, int mt, float (&a)[M4<ND>::S]) {
    using Sh = M4<ND>;
    const int l = threadIdx.x & 63, g = l >> 4;
    const int m = mt * 16 + (l & 15);
    const int s = m >> 1, po = m & 1;
    const bool ok = m < Sh::K2;
    const int s2 = ok ? s / ND : 0, s3 = ok ? s % ND : 0;
#pragma unroll
    for (int i = 0; i < Sh::S / 2; ++i) {
        // padding (rows past K2, K past NN) reads a valid entry and is zeroed by a select
        const int c = g * (Sh::S / 2) + i, cc = min(c, Sh::NN - 1);
        float2 e = cmulf(e2[s2 * ND + cc / ND], e3[s3 * ND + cc % ND]);
        if (!(ok && c < Sh::NN)) e = make_float2(0.f, 0.f);
        a[2 * i] = po ? e.y : e.x;
        a[2 * i + 1] = po ? e.x : -e.y;
    }
}

// A fragment of rows mt*16.. of a tile in LDS (pitch PA, K contiguous)
template <int ND>
__device__ __forceinline__ void a_rows(const float *sA, int mt, float (&a)[M4<ND>::S]) {
    using Sh = M4<ND>;
    const int l = threadIdx.x & 63, g = l >> 4;
    // rows past NN (the last M tile) repeat row NN - 1; their outputs are not used
    const int row = min(mt * 16 + (l & 15), Sh::NN - 1);
    const float4 *p = reinterpret_cast<const float4 *>(sA + row * Sh::PA + g * Sh::S);
#pragma unroll
    for (int i = 0; i < Sh::S / 4; ++i) {
        const float4 v = p[i];
        a[4 * i] = v.x;
        a[4 * i + 1] = v.y;
        a[4 * i + 2] = v.z;
        a[4 * i + 3] = v.w;
    }
}

// B fragment of Y[cd][(s01, pi)] (LDS, pitch PA) with K = (cd, pi), column s01 = nt*16 + l%16
template <int ND>
__device__ __forceinline__ void b_y(const float *Y, int nt, float (&b)[M4<ND>::S]) {
    using Sh = M4<ND>;
    const int l = threadIdx.x & 63, g = l >> 4;
    const int s01 = nt * 16 + (l & 15), sc = min(s01, Sh::NN - 1);
#pragma unroll
    for (int i = 0; i < Sh::S / 2; ++i) {
        // padding reads a valid entry and is zeroed by a select
        const int cd = g * (Sh::S / 2) + i, cc = min(cd, Sh::NN - 1);
        float2 y = *reinterpret_cast<const float2 *>(Y + cc * Sh::PA + 2 * sc);
        if (!(s01 < Sh::NN && cd < Sh::NN)) y = make_float2(0.f, 0.f);
        b[2 * i] = y.x;
        b[2 * i + 1] = y.y;
    }
}

template <int S>
__device__ __forceinline__ v4f mma(const float (&a)[S], const float (&b)[S]) {
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < S; ++k) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[k], b[k], acc, 0, 0, 0);
    return acc;
}

template <int ND>
__global__ __launch_bounds__(512) void k_lw4_mm(Lw4MmArgs A) {
    using Sh = M4<ND>;
    constexpr int NN = Sh::NN, PA = Sh::PA, NW = Sh::NW, NB = Sh::NB;
    constexpr int MT1 = Sh::MT1, NT1 = Sh::NT1, MT2 = Sh::MT2, NT2 = Sh::NT2;
    const M4Lds<ND> Ly(A.nterm, A.nsite, A.L);
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned char *p = smem;
    float *sTile = reinterpret_cast<float *>(p);
    p += Ly.tile;
    float *sB01 = reinterpret_cast<float *>(p);        // GEMM 1 B fragments of e0 x e1
    p += Ly.e01;
    const float *__restrict__ fcv = A.fc;             // FC4 from global memory (cached)
    float2 *sE = reinterpret_cast<float2 *>(p);        // e0 | e1
    p += Ly.e;
    double *sMd = reinterpret_cast<double *>(p);       // w n 1/w of q0, then of q1
    p += Ly.md;
    int *sM1 = reinterpret_cast<int *>(p);             // site phase indices of q1
    p += Ly.m1;
    float2 *sPh = reinterpret_cast<float2 *>(p);
    p += Ly.ph;
    int *sBlk = reinterpret_cast<int *>(p);
    p += Ly.blk;
    double *sRed = reinterpret_cast<double *>(p);
    p += Ly.red;

    const int tid = threadIdx.x, l = tid & 63, g = l >> 4, q = l & 1;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
    unsigned char *pw = p + (size_t)w * Ly.wave;
    float2 *wE = reinterpret_cast<float2 *>(pw);       // e2 | e3
    pw += Ly.e;
    double2 *wCd = reinterpret_cast<double2 *>(pw);    // per s23: w2 + w3, w3 - w2
    float4 *wCf = reinterpret_cast<float4 *>(pw + (size_t)NN * 16);   // n2 n3, (1 + n2) n3, 1/(w2 w3)
    pw += al16((size_t)NN * 32);
    int *wMi = reinterpret_cast<int *>(pw);            // site phase indices of q2, then of q3
    pw += al16((size_t)2 * A.nsite * sizeof(int));
    float2 *wPh = reinterpret_cast<float2 *>(pw);      // cluster phases of the triple
    float *T = sTile + (size_t)w * NN * PA;

    const int q1 = blockIdx.x, qb = blockIdx.y;
    const int i0 = A.q0_grid[qb];
    const int k00 = i0 % A.m0, k01 = (i0 / A.m0) % A.m1, k02 = i0 / (A.m0 * A.m1);
    const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);

    // ---- per item: zero the tiles (K padding stays zero), q0 / q1 data, FC4, e0 x e1 ----
    for (int i = tid; i < (int)(Ly.tile / sizeof(float)); i += 512) sTile[i] = 0.f;
    for (int i = tid; i < 2 * NN; i += 512) sE[i] = A.evec[(size_t)(i < NN ? i0 : q1) * NN + i % NN];
    if (tid < 6 * ND) {
        const int k = tid / (3 * ND), what = (tid / ND) % 3, s = tid % ND;
        const double *src = what == 0 ? A.omega : what == 1 ? A.nbe : A.inv_omega;
        sMd[tid] = src[(size_t)(k ? q1 : i0) * ND + s];
    }
    for (int s = tid; s < A.nsite; s += 512) {
        const int4 c = A.site[s];
        sM1[s] = (k10 * c.x + k11 * c.y + k12 * c.z) % A.L;
    }
    for (int i = tid; i < A.L; i += 512) sPh[i] = A.phase[i];
    for (int i = tid; i <= Sh::NBLK; i += 512) sBlk[i] = A.blk_start[i];
    __syncthreads();
    for (int i = tid; i < NT1 * 64 * Sh::S; i += 512) {
        const int nt = i / (64 * Sh::S), ll = (i / Sh::S) % 64, j = i % Sh::S;
        sB01[i] = b_e01<ND>(sE, sE + NN, nt, ll, j);
    }
    __syncthreads();

    // ---- per lane: constants of its GEMM 2 columns s01 = nt*16 + l%16 ----
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002);   // 1/(sigma sqrt(2 pi))
    // exp(-x^2 / (2 sigma^2)) = 2^(hl2 x^2) on the exponential unit
    const float hl2 = (float)(-0.5 / (A.sigma * A.sigma) * 1.4426950408889634);
    const bool g0 = i0 == 0;
    double Ra[NT2], Rb[NT2];
    float r1[NT2], r2[NT2];
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
        const int s01 = nt * 16 + (l & 15);
        Ra[nt] = Rb[nt] = 0.0;
        r1[nt] = r2[nt] = 0.f;
        if (s01 < NN) {
            const int s0 = s01 / ND, s1 = s01 % ND;
            const double w0 = sMd[s0], n0 = sMd[ND + s0];
            const double w1 = sMd[3 * ND + s1], n1 = sMd[4 * ND + s1];
            Ra[nt] = w0 - w1;
            Rb[nt] = w0 + w1;
            if (!(g0 && s0 < 3)) {
                r1[nt] = (float)(n1 / n0);
                r2[nt] = (float)((1.0 + n1) / n0);
            }
        }
    }
    KSum<float> acc[NT2];

    // Without branch rules, one ordered triple per permutation class of (q1, q2, q3): the one
    // with q1 <= q2 <= q3, whose V4 serves every ordering (seven distinct process terms per
    // element, see below); with branch rules every ordered triple, each with its masks.
    const bool classes = A.sel == nullptr;
    for (int q2 = (classes ? q1 : 0) + w; q2 < A.nq; q2 += NW) {
        const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
        int a0 = -(k00 + k10 + k20) % A.m0, a1 = -(k01 + k11 + k21) % A.m1, a2 = -(k02 + k12 + k22) % A.m2;
        a0 += a0 < 0 ? A.m0 : 0;
        a1 += a1 < 0 ? A.m1 : 0;
        a2 += a2 < 0 ? A.m2 : 0;
        const int q3 = a0 + A.m0 * (a1 + A.m1 * a2);
        if (classes && q3 < q2) continue;
        // 1 / |stabiliser| of the class (the six orderings meet each distinct one 6 / c times)
        const float istab = !classes ? 1.f : q1 == q3 ? 1.f / 6.f : (q1 == q2 || q2 == q3) ? 0.5f : 1.f;
        wave_sync();   // the previous triple of this wave is done with its LDS
        for (int i = l; i < 2 * NN; i += 64) wE[i] = A.evec[(size_t)(i < NN ? q2 : q3) * NN + i % NN];
        if (l < NN) {   // column constants of s23 = l
            const int s2 = l / ND, s3 = l % ND;
            const double w2 = A.omega[(size_t)q2 * ND + s2], w3 = A.omega[(size_t)q3 * ND + s3];
            const double n2 = A.nbe[(size_t)q2 * ND + s2], n3 = A.nbe[(size_t)q3 * ND + s3];
            wCd[l] = make_double2(w2 + w3, w3 - w2);
            const double m23 = A.sel ? A.sel[2 * ND + s2] * A.sel[3 * ND + s3] : 1.0;
            wCf[l] = make_float4((float)(n2 * n3), (float)((1.0 + n2) * n3),
                                 (float)(A.inv_omega[(size_t)q2 * ND + s2] * A.inv_omega[(size_t)q3 * ND + s3] * m23),
                                 (float)(n2 * (1.0 + n3)));
        }
        for (int i = l; i < 2 * A.nsite; i += 64) {
            const int4 c = A.site[i % A.nsite];
            wMi[i] = i < A.nsite ? (k20 * c.x + k21 * c.y + k22 * c.z) % A.L : (a0 * c.x + a1 * c.y + a2 * c.z) % A.L;
        }
        wave_sync();
        for (int u = l; u < A.nterm; u += 64) {
            int m = sM1[A.tup_site[3 * u]] + wMi[A.tup_site[3 * u + 1]] + wMi[A.nsite + A.tup_site[3 * u + 2]];
            m -= m >= A.L ? A.L : 0;
            m -= m >= A.L ? A.L : 0;
            wPh[u] = sPh[m];
        }
        wave_sync();

        // ---- FT4: T[cd][(ab, re/im)]; lanes run along the 81 components of a block ----
#pragma unroll 1
        for (int blk = 0; blk < Sh::NBLK; ++blk) {
            const int mu = blk / (NB * NB * NB), nu = (blk / (NB * NB)) % NB, et = (blk / NB) % NB, rh = blk % NB;
            const int u0 = sBlk[blk], u1 = sBlk[blk + 1];
            // component l and (lanes < 17) 64 + l
            const bool two = l < 81 - 64;
            float re0 = 0.f, im0 = 0.f, re1 = 0.f, im1 = 0.f;
#pragma unroll 4
            for (int u = u0; u < u1; ++u) {
                const float2 ph = wPh[u];
                const float f0 = fcv[u * 81 + l];
                const float f1 = two ? fcv[u * 81 + 64 + l] : 0.f;
                re0 = fmaf(f0, ph.x, re0);
                im0 = fmaf(f0, ph.y, im0);
                re1 = fmaf(f1, ph.x, re1);
                im1 = fmaf(f1, ph.y, im1);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int comp = h * 64 + l;
                if (comp < 81) {
                    const int a = 3 * mu + comp / 27, b = 3 * nu + (comp / 9) % 3;
                    const int c = 3 * et + (comp / 3) % 3, d = 3 * rh + comp % 3;
                    float *dst = T + (c * ND + d) * PA + 2 * (a * ND + b);
                    dst[0] = h ? re1 : re0;
                    dst[1] = h ? im1 : im0;
                }
            }
        }
        wave_sync();

        // ---- GEMM 1 per M tile of rows cd: Y[cd][(s01, re/im)] written over the rows just read ----
#pragma unroll 1
        for (int mt = 0; mt < MT1; ++mt) {
            float a[Sh::S];
            a_rows<ND>(T, mt, a);
            wave_sync();   // the rows of this M tile are in registers; Y may overwrite them
#pragma unroll 1
            for (int nt = 0; nt < NT1; ++nt) {
                float b[Sh::S];
                b_table<ND>(sB01, nt, b);
                const v4f y = mma<Sh::S>(a, b);
                const int col = nt * 16 + (l & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int cd = mt * 16 + 4 * g + r;
                    if (cd < NN && col < Sh::K2) T[cd * PA + col] = y[r];
                }
            }
        }
        wave_sync();

        // ---- GEMM 2 per M tile of rows (s23, re/im), then W4 and the processes on each tile:
        //      rows 4g, 4g+1 are (re, im) of s23 = mt*8 + 2g, rows 4g+2, 4g+3 of s23 + 1 ----
        float tri[NT2];
#pragma unroll
        for (int nt = 0; nt < NT2; ++nt) tri[nt] = 0.f;
#pragma unroll 1
        for (int mt = 0; mt < MT2; ++mt) {
            asm volatile("" ::: "memory");   // keeps the B fragments of the N tiles inside the loop
            float a[Sh::S];
            a_kron<ND>(wE, wE + NN, mt, a);
#pragma unroll
            for (int nt = 0; nt < NT2; ++nt) {
                float b[Sh::S];
                b_y<ND>(T, nt, b);
                const v4f v = mma<Sh::S>(a, b);
                if (nt * 16 + (l & 15) >= NN) continue;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int s23 = mt * 8 + 2 * g + h;
                    if (s23 >= NN) continue;
                    const double2 cd2 = wCd[s23];
                    const float4 cf = wCf[s23];
                    float inner;
                    if (classes) {
                        // every ordering of the class: with P = n2 n3, Q = (1+n2) n3, U = n2 (1+n3),
                        // X = (1+n2)(1+n3), a = n1/n0, b = (1+n1)/n0 (weights 1/6, 1/2, 1/2 summed
                        // over the six orderings):
                        //   a [P g(w0-w1-w2-w3) + Q g(w0-w1+w2-w3) + U g(w0-w1-w2+w3) + X g(w0-w1+w2+w3)]
                        //   + b [P g(w0+w1-w2-w3) + Q g(w0+w1+w2-w3) + U g(w0+w1-w2+w3)]
                        const float P = cf.x, Q = cf.y, U = cf.w, X = 1.f + Q + U - P;
                        const double ra = Ra[nt], rb = Rb[nt];
                        const float xa1 = (float)(ra - cd2.x), xa2 = (float)(ra - cd2.y), xa3 = (float)(ra + cd2.y),
                                    xa4 = (float)(ra + cd2.x);
                        const float xb1 = (float)(rb - cd2.x), xb2 = (float)(rb - cd2.y), xb3 = (float)(rb + cd2.y);
                        const float sa = P * __builtin_amdgcn_exp2f(hl2 * xa1 * xa1) +
                                         Q * __builtin_amdgcn_exp2f(hl2 * xa2 * xa2) +
                                         U * __builtin_amdgcn_exp2f(hl2 * xa3 * xa3) +
                                         X * __builtin_amdgcn_exp2f(hl2 * xa4 * xa4);
                        const float sb = P * __builtin_amdgcn_exp2f(hl2 * xb1 * xb1) +
                                         Q * __builtin_amdgcn_exp2f(hl2 * xb2 * xb2) +
                                         U * __builtin_amdgcn_exp2f(hl2 * xb3 * xb3);
                        inner = r1[nt] * sa + r2[nt] * sb;
                    } else {
                        const float x1 = (float)(Ra[nt] - cd2.x), x2 = (float)(Rb[nt] - cd2.x), x3 = (float)(Rb[nt] - cd2.y);
                        const float e1 = __builtin_amdgcn_exp2f(hl2 * x1 * x1), e2 = __builtin_amdgcn_exp2f(hl2 * x2 * x2),
                                    e3 = __builtin_amdgcn_exp2f(hl2 * x3 * x3);
                        inner = cf.x * (r1[nt] * e1 * (1.f / 6.f) + r2[nt] * e2 * 0.5f) + r2[nt] * cf.y * e3 * 0.5f;
                    }
                    tri[nt] = fmaf((v[2 * h] * v[2 * h] + v[2 * h + 1] * v[2 * h + 1]) * cf.z, inner, tri[nt]);
                }
            }
        }
#pragma unroll
        for (int nt = 0; nt < NT2; ++nt) acc[nt].add(tri[nt] * istab);
    }

    // ---- per s0: the lane's columns times their constants, then waves in a fixed order ----
    double tot[ND];
#pragma unroll
    for (int s = 0; s < ND; ++s) tot[s] = 0.0;
#pragma unroll
    for (int nt = 0; nt < NT2; ++nt) {
        const int s01 = nt * 16 + (l & 15);
        if (s01 < NN) {
            const int s0 = s01 / ND, s1 = s01 % ND;
            double v = sMd[2 * ND + s0] * sMd[5 * ND + s1] * c0 * A.factor * acc[nt].value();
            if (A.sel) v *= A.sel[s0] * A.sel[ND + s1];
#pragma unroll
            for (int s = 0; s < ND; ++s)
                if (s == s0) tot[s] += v;
        }
    }
#pragma unroll
    for (int s = 0; s < ND; ++s) {
        double v = tot[s];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (l == 0) sRed[w * ND + s] = v;
    }
    __syncthreads();
    if (tid < ND) {
        double sum = 0.0;
        for (int ww = 0; ww < NW; ++ww) sum += sRed[ww * ND + tid];
        A.partial[((size_t)qb * A.nq + q1) * ND + tid] = sum;
    }
}

template <int ND>
int launch_lw4_mm_t(Ctx &c, int q0_first, int q0_count, double sigma, double *d_lw) {
    Lw4MmArgs A;
    A.q0_grid = c.irr_grid + q0_first;
    A.nq = c.nq;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.m2 = c.mesh[2];
    A.blk_start = c.fc4.blk_start;
    A.tup_site = c.fc4.tup_site;
    A.fc = c.fc4.fc32;
    A.nterm = c.fc4.nterm;
    A.site = c.fc4.site;
    A.nsite = c.fc4.nsite;
    A.phase = reinterpret_cast<const float2 *>(c.phase32);
    A.evec = reinterpret_cast<const float2 *>(c.evec32);
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sigma = sigma;
    A.factor = lw4_factor(c);
    const size_t dyn = M4Lds<ND>(A.nterm, A.nsite, A.L).total;
    if (dyn > 160 * 1024) return SPR_E_UNSUPPORTED;   // the caller falls back to the vector kernel
    int rc;
    if ((rc = ensure((void **)&c.partial, &c.partial_cap, (size_t)q0_count * c.nq * ND * sizeof(double))))
        return rc;
    A.partial = c.partial;
    A.sel = c.sel4 ? c.d_sel4 : nullptr;
    const void *fn = (const void *)k_lw4_mm<ND>;
    SPR_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    SPR_HIP(hipEventRecord(c.ev0, c.stream));
    hipLaunchKernelGGL(k_lw4_mm<ND>, dim3(c.nq, q0_count), dim3(512), dyn, c.stream, A);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipEventRecord(c.ev1, c.stream));
    if ((rc = launch_reduce_chunks(c, c.partial, q0_count, c.nq, d_lw))) return rc;
    SPR_HIP(hipEventSynchronize(c.ev1));
    float ms = 0.f;
    SPR_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
    c.launches += 1;
    c.kernel_ms += ms;
    return SPR_OK;
}

}  // namespace

// FP32 four-phonon linewidths on the matrix cores; SPR_E_UNSUPPORTED (no error text) when the
// shape is not built or the item's data do not fit in LDS.
int launch_lw4_mm(Ctx &c, int q0_first, int q0_count, double sigma, double *d_lw) {
    if (c.nd == 6) return launch_lw4_mm_t<6>(c, q0_first, q0_count, sigma, d_lw);
    return SPR_E_UNSUPPORTED;
}

}  // namespace spr
