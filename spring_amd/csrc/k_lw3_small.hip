// Three-phonon scattering for small cells (Nb <= 2, Ndof <= 6): the benchmark kernel.
//
// Work unit: a batch of up to 64 (q1, q2) pairs of one irreducible q0, from the
// pair-permutation reduced list (q1 <= q2, weight 2 if q1 < q2; ScatterProbAllq1's
// q1q2_permuted, ThirdOrder_m.f90:98-164).  Workgroup = Ndof waves; wave `a` owns row
// a = (mu, alpha) of the phase-summed tensor, lane `l` owns pair l of the batch.
//
//   1. phase sum (ScatterMatEl, ThirdOrder_m.f90:316-384, grouped by basis atoms):
//        T_a(b, c) = sum_{(N1,N2) in block(mu, nu, eta)} Phi(alpha, beta, gamma) e^{2 pi i m / L}
//      the FC values are wave-uniform (scalar loads), the phase index m = m1[N1] + m2[N2]
//      is per lane (integer addresses, table lookup in LDS), T_a lives in registers.
//   2. T_a <- T_a e2 then e1^T T_a (per lane, eigenvectors staged in LDS lane-major).
//   3. the e0 contraction crosses waves: per s1, every wave writes U(a, s1, :) to LDS and
//      wave (s2, block of three s0) reads U(:, s1, s2) once and forms V(s0, s1, s2) =
//      sum_a e0(a, s0) U(a, s1, s2) for its three s0.
//   4. W = |V|^2 /(w0 w1 w2) Sctr3THz2/Nq and the Gaussian-delta occupation weights
//      (CalculateLinewidth, ThirdOrder_m.f90:514-537) are accumulated per lane (a wave skips an
//      s0 whose three Gaussians underflow to exact zero on every lane), then
//      reduced over the wave: partial[item][s0].  A second kernel sums the items of each
//      q0 in a fixed order (bitwise reproducible).
#include <algorithm>
#include <cmath>
#include <thread>
#include <type_traits>
#include <vector>

#include "cplx.h"
#include "spr_internal.h"

namespace spr {

struct SmallArgs {
    ItemCommon ic;
    const int *item_q0;      // (nitems) irreducible index of each batch
    const int *item_off;     // (nitems) offset into the pair list
    const int *item_cnt;     // (nitems) pairs in the batch (<= 64)
    const int2 *pairs;       // (npairs) (q1, q2)
    const int *q0_grid;      // (nq0) grid index of each irreducible q0
    int L, m0, m1;
    const int *blk_start;
    const int *tup_site;
    const void *fc;
    const int4 *site;
    int nsite;
    const void *phase;
    const void *evec;
    const double *omega, *inv_omega, *nbe;
    double sigma, factor;
    // W output (WOUT instantiation): w_out[q0 - w_q0_first][q1][s0][s1][s2] for every grid q1,
    // both orientations of each pair written
    double *w_out;
    int w_q0_first, w_nq;
    // W output: only the first orientation W(q1; s0, s1, s2) of each pair (items that list every
    // q1 of their q0, so that each orientation is evaluated on its own)
    int w_one;
    // branch selection (NULL: every branch): (nd) s0 mask, then the (nd, nd) pair mask [s1][s2]
    const double *sel;
    // per-pair weights of caller-supplied items (NULL: 1 for q1 == q2, else 2)
    const double *pair_w;
};

// SEL: a branch selection rule is set (A.sel); a compile-time flag, because a runtime test in the
// delta loop costs this register-bound kernel about 40 % (measured in FP64)
template <typename R, int NB, bool WOUT, bool SEL>
__device__ __forceinline__ void small_item(const SmallArgs &A, const int item, const R *__restrict__ fc,
                                           const int *__restrict__ blk_start, const int *__restrict__ tup_site) {
    using Z = cplx_t<R>;
    constexpr int ND = 3 * NB;
    constexpr int ND2 = ND * ND;
    constexpr int NW = ND * NB;            // waves: row a x half nu
    extern __shared__ __align__(16) unsigned char smem[];
    // layout: e1[ND2][64] | x[NW][ND][64] (exchange; the e2[ND2][64] staging aliases its start)
    //         | m1[nsite][64], m2[nsite][64] (int) | md[6 ND][64] (double) | phase[2L] | e0[ND2]
    // m1 holds the LDS byte address of phase[m1] and m2 the byte offset 16 m2 (8 m2 in FP32):
    // the sum addresses phase[m1 + m2], and the table holds two periods so that no mod L is needed
    Z *s_e1 = reinterpret_cast<Z *>(smem);
    Z *s_x = s_e1 + ND2 * 64;
    Z *s_e2 = s_x;
    int *s_m1 = reinterpret_cast<int *>(s_x + NW * ND * 64);
    int *s_m2 = s_m1 + A.nsite * 64;
    double *s_md = reinterpret_cast<double *>(s_m2 + A.nsite * 64);
    Z *s_ph = reinterpret_cast<Z *>(s_md + 6 * ND * 64);
    Z *s_e0 = s_ph + 2 * A.L;

    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int a = w / NB, nu = w % NB;             // phase-sum row a, rows b = 3 nu + beta
    const int mu = a / 3, al = a % 3;
    const int qi = A.item_q0[item];
    const int i0 = A.q0_grid[qi];
    const int cnt = A.item_cnt[item];
    const int2 pr = A.pairs[A.item_off[item] + (lane < cnt ? lane : 0)];
    const int q1 = pr.x, q2 = pr.y;
    const double wpair = lane >= cnt ? 0.0 : A.pair_w ? A.pair_w[A.item_off[item] + lane] : q1 == q2 ? 1.0 : 2.0;
    const Z *evec = reinterpret_cast<const Z *>(A.evec);
    const Z *phase = reinterpret_cast<const Z *>(A.phase);

    // ---- stage per-lane data: eigenvectors (lane-major), phase indices, mode data ----
    for (int i = w; i < ND2; i += NW) {
        s_e1[i * 64 + lane] = evec[(size_t)q1 * ND2 + i];
        s_e2[i * 64 + lane] = evec[(size_t)q2 * ND2 + i];
    }
    {
        const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
        const int k20 = q2 % A.m0, k21 = (q2 / A.m0) % A.m1, k22 = q2 / (A.m0 * A.m1);
        const int ph_base = (int)(reinterpret_cast<const unsigned char *>(s_ph) - smem);
        for (int s = w; s < A.nsite; s += NW) {
            const int4 c = A.site[s];
            s_m1[s * 64 + lane] = ph_base + (int)sizeof(Z) * ((k10 * c.x + k11 * c.y + k12 * c.z) % A.L);
            s_m2[s * 64 + lane] = (int)sizeof(Z) * ((k20 * c.x + k21 * c.y + k22 * c.z) % A.L);
        }
    }
    for (int k = w; k < 6 * ND; k += NW) {   // omega, nBE, 1/omega of q1 (k < 3 ND), then of q2
        const int qq = k < 3 * ND ? q1 : q2, kk = k % (3 * ND), what = kk / ND, s = kk % ND;
        const double *src = what == 0 ? A.omega : what == 1 ? A.nbe : A.inv_omega;
        s_md[k * 64 + lane] = src[(size_t)qq * ND + s];
    }
    for (int i = threadIdx.x; i < 2 * A.L; i += NW * 64) s_ph[i] = phase[i < A.L ? i : i - A.L];
    for (int i = threadIdx.x; i < ND2; i += NW * 64) s_e0[i] = evec[(size_t)i0 * ND2 + i];
    __syncthreads();

    // ---- 1. phase sum: T(beta, c) = T_a(3 nu + beta, c), FC values wave-uniform ----
    Z T[3][ND];
#pragma unroll
    for (int be = 0; be < 3; ++be)
#pragma unroll
        for (int c = 0; c < ND; ++c) T[be][c] = cmk<R>(0, 0);
    // tuples in batches of four (every block is padded to a multiple of four with zero FC
    // values): the site pairs and FC values of a batch arrive in one scalar-load round, its
    // four phases in one LDS round, so the load latency is paid once per 72 FMAs
#pragma unroll
    for (int eta = 0; eta < NB; ++eta) {
        const int blk = (mu * NB + nu) * NB + eta;
        const int u0 = blk_start[blk], u1 = blk_start[blk + 1];
        for (int u = u0; u < u1; u += kClusterBatch) {
            // the batch's four (site 1, site 2) pairs in one scalar load (u is a multiple of four)
            const int4 *ts = reinterpret_cast<const int4 *>(tup_site + 2 * u);
            const int4 t01 = ts[0], t23 = ts[1];
            const int sa[kClusterBatch] = {t01.x, t01.z, t23.x, t23.z};
            const int sb[kClusterBatch] = {t01.y, t01.w, t23.y, t23.w};
            Z ph[kClusterBatch];
#pragma unroll
            for (int k = 0; k < kClusterBatch; ++k) {
                const int addr = s_m1[sa[k] * 64 + lane] + s_m2[sb[k] * 64 + lane];
                ph[k] = *reinterpret_cast<const Z *>(smem + addr);
            }
#pragma unroll
            for (int k = 0; k < kClusterBatch; ++k) {
                const R *f = fc + (size_t)(u + k) * 27 + al * 9;
#pragma unroll
                for (int be = 0; be < 3; ++be)
#pragma unroll
                    for (int ga = 0; ga < 3; ++ga)
                        T[be][3 * eta + ga] = crfma(f[be * 3 + ga], ph[k], T[be][3 * eta + ga]);
            }
        }
    }
    // ---- 2. U1(beta, s2) = sum_c T(beta, c) e2(c, s2), in place per row ----
#pragma unroll
    for (int be = 0; be < 3; ++be) {
        Z t[ND];
#pragma unroll
        for (int s2 = 0; s2 < ND; ++s2) {
            Z z = cmk<R>(0, 0);
#pragma unroll
            for (int c = 0; c < ND; ++c) z = cfma(T[be][c], s_e2[(s2 * ND + c) * 64 + lane], z);
            t[s2] = z;
        }
#pragma unroll
        for (int s2 = 0; s2 < ND; ++s2) T[be][s2] = t[s2];
    }

    // ---- 3 + 4. per s1: every wave writes its partial U2(a, nu; s1, s2) = sum_{b of nu} e1(b, s1)
    //      U1(b, s2) for all s2 to LDS; wave (s2, s0 block of three) then reads the partials of its
    //      s2 once, sums them against e0(a, s0) for its three s0, and forms W and the delta sum ----
    const int s2 = w / NB, s0b = 3 * (w % NB);
    double w0[3], iw0f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w0[k] = A.omega[(size_t)i0 * ND + s0b + k];
        iw0f[k] = A.inv_omega[(size_t)i0 * ND + s0b + k] * A.factor;
    }
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002);
    const double hinv = 0.5 / (A.sigma * A.sigma);
    KSum<R> kacc[3];
    double dacc[3] = {0.0, 0.0, 0.0};
    __syncthreads();   // every wave is done with e2 before the exchange buffer overwrites it
    const double w2 = s_md[(3 * ND + s2) * 64 + lane];
    const double n2 = s_md[(4 * ND + s2) * 64 + lane];
    const double iw2 = s_md[(5 * ND + s2) * 64 + lane];
#pragma unroll 1
    for (int s1 = 0; s1 < ND; ++s1) {
        Z e1v[3];
#pragma unroll
        for (int be = 0; be < 3; ++be) e1v[be] = s_e1[(s1 * ND + 3 * nu + be) * 64 + lane];
#pragma unroll
        for (int sx = 0; sx < ND; ++sx) {
            Z z = cmk<R>(0, 0);
#pragma unroll
            for (int be = 0; be < 3; ++be) z = cfma(e1v[be], T[be][sx], z);
            s_x[(w * ND + sx) * 64 + lane] = z;
        }
        __syncthreads();
        const double w1 = s_md[(0 * ND + s1) * 64 + lane];
        const double n1 = s_md[(1 * ND + s1) * 64 + lane];
        const double iw1 = s_md[(2 * ND + s1) * 64 + lane];
        // exp(-t) is exactly zero in FP64 beyond t = 745.2, so a wave whose lanes are all that far
        // from resonance for this s0 (consecutive q1 of one q0 have close frequencies, so most
        // are) adds exact zeros: it skips V and the exponentials, with the same result bit for bit
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double d1 = w0[k] - w1 - w2, d2 = w0[k] + w1 - w2, d3 = w0[k] - w1 + w2;
            const double t1 = hinv * d1 * d1, t2 = hinv * d2 * d2, t3 = hinv * d3 * d3;
            if (!WOUT && !__ballot(fmin(t1, fmin(t2, t3)) < 746.0)) continue;
            Z v = cmk<R>(0, 0);
#pragma unroll
            for (int ar = 0; ar < ND; ++ar) {
                Z u = s_x[((ar * NB) * ND + s2) * 64 + lane];
#pragma unroll
                for (int h = 1; h < NB; ++h) u = cadd(u, s_x[((ar * NB + h) * ND + s2) * 64 + lane]);
                v = cfma(s_e0[(s0b + k) * ND + ar], u, v);
            }
            const double W = (double)cabs2(v) * (iw0f[k] * iw1) * iw2;
            if constexpr (WOUT) {
                // W(s2, s1, s0; q1) and, for the partner ordering, W(s1, s2, s0; q2)
                if (lane < cnt) {
                    const int s0 = s0b + k;
                    double *wq = A.w_out + (size_t)(qi - A.w_q0_first) * A.w_nq * ND * ND * ND;
                    wq[(((size_t)q1 * ND + s0) * ND + s1) * ND + s2] = W;
                    if (q1 != q2 && !A.w_one) wq[(((size_t)q2 * ND + s0) * ND + s2) * ND + s1] = W;
                }
                continue;
            }
            double g = (n1 + n2 + 1.0) * (c0 * exp(-t1)) + (n1 - n2) * (c0 * exp(-t2) - c0 * exp(-t3));
            if constexpr (SEL) g *= A.sel[ND + s1 * ND + s2];
            if (sizeof(R) == 8)
                dacc[k] += wpair * W * g;
            else
                kacc[k].add((R)(wpair * W * g));
        }
        __syncthreads();
    }
    // wave reductions, then for each s0 the ND waves (one per s2) sharing its block, in a fixed order
    double *s_red = reinterpret_cast<double *>(s_x);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double tot = sizeof(R) == 8 ? dacc[k] : kacc[k].value();
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) tot += __shfl_down(tot, off, 64);
        if (lane == 0) s_red[w * 3 + k] = tot;
    }
    __syncthreads();
    if (threadIdx.x < ND) {
        const int h = threadIdx.x / 3, k = threadIdx.x % 3;
        double sum = 0.0;
        for (int sx = 0; sx < ND; ++sx) sum += s_red[(sx * NB + h) * 3 + k];
        A.ic.partial[(size_t)item * ND + threadIdx.x] = SEL ? sum * A.sel[threadIdx.x] : sum;
    }
}

// One workgroup per item, or (counter set) a persistent grid pulling items until none are left.
template <typename R, int NB, bool WOUT, bool SEL>
__global__ __launch_bounds__(NB * NB * 3 * 64) void k_lw3_small(SmallArgs A, const R *__restrict__ fc,
                                                               const int *__restrict__ blk_start,
                                                               const int *__restrict__ tup_site) {
    if (!A.ic.counter) {
        small_item<R, NB, WOUT, SEL>(A, blockIdx.x, fc, blk_start, tup_site);
        return;
    }
    __shared__ int s_item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(A.ic.counter, 1);
        __syncthreads();
        const int item = s_item;
        if (item >= A.ic.nitems) break;
        small_item<R, NB, WOUT, SEL>(A, item, fc, blk_start, tup_site);
    }
}

// lw[q][s] = sum of partial[unit][s] over the units of items [first[q], first[q+1]) in order
// (upi units per item, unit = item * upi + h)
__global__ void k_reduce_items(const double *partial, const int *first, int base, int upi, int nq0, int nd,
                               double *lw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq0 * nd) return;
    const int q = i / nd, s = i % nd;
    double acc = 0.0;
    for (int u = (first[q] - base) * upi; u < (first[q + 1] - base) * upi; ++u) acc += partial[(size_t)u * nd + s];
    lw[i] = acc;
}

template <typename R, int NB>
static int launch_small_t(Ctx &c, const SmallArgs &A0, int nitems, int q0_first, int q0_count, int it0,
                          double *d_lw, bool wout = false) {
    constexpr int ND = 3 * NB;
    SmallArgs A = A0;
    A.fc = sizeof(R) == 8 ? (const void *)c.fc3.fc64_p : (const void *)c.fc3.fc32_p;
    A.phase = sizeof(R) == 8 ? (const void *)c.phase64 : (const void *)c.phase32;
    A.evec = sizeof(R) == 8 ? (const void *)c.evec : (const void *)c.evec32;
    constexpr int NW = ND * NB;
    const size_t dyn = (size_t)(ND * ND + NW * ND) * 64 * sizeof(cplx_t<R>) + (size_t)2 * A.nsite * 64 * sizeof(int) +
                       (size_t)6 * ND * 64 * sizeof(double) + (size_t)(2 * c.L + ND * ND) * sizeof(cplx_t<R>);
    if (dyn > 160 * 1024) {
        set_error("small-cell kernel: LDS need exceeds 160 KB (too many neighbour sites)");
        return SPR_E_UNSUPPORTED;
    }
    const void *fn = wout     ? (const void *)k_lw3_small<R, NB, true, false>
                     : A.sel ? (const void *)k_lw3_small<R, NB, false, true>
                             : (const void *)k_lw3_small<R, NB, false, false>;
    if (dyn > 64 * 1024) SPR_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    int per_cu = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NW * 64, dyn) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const R *fcp = reinterpret_cast<const R *>(A.fc);
    const int *bsp = A.blk_start, *tsp = A.tup_site;
    ItemLaunch L{fn, &A, sizeof(A), NW * 64, dyn, nitems, per_cu,
                 std::string("small") + (sizeof(R) == 8 ? "64" : "32") + (wout ? "w" : A.sel ? "s" : "")};
    L.extra[0] = (void *)&fcp;
    L.extra[1] = (void *)&bsp;
    L.extra[2] = (void *)&tsp;
    L.nextra = 3;
    L.direct = c.override_items != nullptr;
    return run_items(c, L, q0_first, q0_count, it0, d_lw);
}

// Builds (once per context) the pair lists of every irreducible q0 and the 64-pair batches.
int ensure_pairs(Ctx &c) {
    if (c.pairs_built) return SPR_OK;
    const int nq = c.nq, m0 = c.mesh[0], m1 = c.mesh[1], m2 = c.mesh[2];
    // q2 = -(q0 + q1) mod G per grid coordinate from three small tables per q0 (no divisions);
    // the q0 are split over host threads, first counting the pairs of each q0, then filling
    // each q0's range of the list, so the list is the same as a single pass in q0 order
    auto for_pairs = [&](int qi, auto &&emit) {
        const int i0 = c.h_irr_grid[qi];
        const int k0 = i0 % m0, k1 = (i0 / m0) % m1, k2 = i0 / (m0 * m1);
        std::vector<int> n0(m0), n1(m1), n2(m2);
        for (int x = 0; x < m0; ++x) n0[x] = (2 * m0 - k0 - x) % m0;
        for (int y = 0; y < m1; ++y) n1[y] = m0 * ((2 * m1 - k1 - y) % m1);
        for (int z = 0; z < m2; ++z) n2[z] = m0 * m1 * ((2 * m2 - k2 - z) % m2);
        int q1 = 0;
        for (int z = 0; z < m2; ++z)
            for (int y = 0; y < m1; ++y) {
                const int b = n1[y] + n2[z];
                for (int x = 0; x < m0; ++x, ++q1) {
                    const int q2 = n0[x] + b;
                    if (q1 <= q2) emit(q1, q2);
                }
            }
    };
    const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    auto parallel_q0 = [&](auto &&body) {
        std::vector<std::thread> th;
        for (int t = 0; t < nth; ++t)
            th.emplace_back([&, t]() {
                for (int qi = t; qi < c.nq0; qi += nth) body(qi);
            });
        for (auto &x : th) x.join();
    };
    std::vector<size_t> pstart(c.nq0 + 1, 0);
    parallel_q0([&](int qi) {
        size_t n = 0;
        for_pairs(qi, [&](int, int) { ++n; });
        pstart[qi + 1] = n;
    });
    for (int qi = 0; qi < c.nq0; ++qi) pstart[qi + 1] += pstart[qi];
    std::vector<int2> pairs(pstart[c.nq0]);
    parallel_q0([&](int qi) {
        size_t at = pstart[qi];
        for_pairs(qi, [&](int q1, int q2) { pairs[at++] = make_int2(q1, q2); });
    });
    std::vector<int> iq0, ioff, icnt, first(c.nq0 + 1, 0);
    for (int qi = 0; qi < c.nq0; ++qi) {
        first[qi] = (int)iq0.size();
        for (size_t o = pstart[qi]; o < pstart[qi + 1]; o += 64) {
            iq0.push_back(qi);
            ioff.push_back((int)o);
            icnt.push_back((int)std::min<size_t>(64, pstart[qi + 1] - o));
        }
    }
    first[c.nq0] = (int)iq0.size();
    c.n_pairs = (long)pairs.size();
    c.n_items = (int)iq0.size();
    auto up = [&](auto **dst, const auto &v) -> int {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        if (hipMalloc((void **)dst, v.size() * sizeof(T)) != hipSuccess) {
            set_error("pair list: hipMalloc failed");
            return SPR_E_OOM;
        }
        if (hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("pair list: upload failed");
            return SPR_E_HIP;
        }
        return SPR_OK;
    };
    int rc;
    if ((rc = up(&c.d_pairs, pairs))) return rc;
    if ((rc = up(&c.d_item_q0, iq0))) return rc;
    if ((rc = up(&c.d_item_off, ioff))) return rc;
    if ((rc = up(&c.d_item_cnt, icnt))) return rc;
    if ((rc = up(&c.d_item_first, first))) return rc;
    if (hipMalloc((void **)&c.ipartial, (size_t)c.n_items * kMaxUnitsPerItem * c.nd * sizeof(double)) !=
        hipSuccess) {
        set_error("item partials: hipMalloc failed");
        return SPR_E_OOM;
    }
    c.h_item_first = first;
    c.pairs_built = true;
    return SPR_OK;
}

int launch_reduce_items(Ctx &c, int q0_first, int q0_count, int it0, int upi, double *d_lw) {
    hipLaunchKernelGGL(k_reduce_items, dim3((q0_count * c.nd + 255) / 256), dim3(256), 0, c.stream,
                       c.ipartial, c.d_item_first + q0_first, it0, upi, q0_count, c.nd, d_lw);
    SPR_HIP(hipGetLastError());
    return SPR_OK;
}

// Linewidths of irreducible q0 [first, first + count) with the small-cell kernel.
int launch_lw3_small(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw) {
    int rc;
    // caller-supplied items (override_items) need none of the context's pair lists
    if (!c.override_items && (rc = ensure_pairs(c))) return rc;
    const ItemSet *ov = c.override_items;
    const int it0 = ov ? 0 : c.h_item_first[q0_first];
    const int nitems = ov ? ov->nitems : c.h_item_first[q0_first + q0_count] - it0;
    SmallArgs A;
    A.item_q0 = ov ? ov->item_q0 : c.d_item_q0 + it0;
    A.item_off = ov ? ov->item_off : c.d_item_off + it0;
    A.item_cnt = ov ? ov->item_cnt : c.d_item_cnt + it0;
    A.pairs = ov ? ov->pairs : c.d_pairs;
    A.q0_grid = c.irr_grid;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.blk_start = c.fc3.blk_start_p;
    A.tup_site = c.fc3.tup_site_p;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sigma = sigma;
    const double pi = 4.0 * std::atan(1.0);
    A.factor = 1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) *
               1000000.0 * pi * 0.0625 / (double)c.nq;   // Sctr3THz2 / Nq, constants.f90:59-60
    A.ic.partial = ov ? ov->partial : c.ipartial;
    A.w_out = nullptr;
    A.w_q0_first = A.w_nq = A.w_one = 0;
    A.sel = c.sel3 ? c.d_sel3 : nullptr;
    A.pair_w = ov ? ov->pair_w : nullptr;
    const bool f64 = prec == SPR_PREC_FP64;
    if (c.nb == 2)
        return f64 ? launch_small_t<double, 2>(c, A, nitems, q0_first, q0_count, it0, d_lw)
                   : launch_small_t<float, 2>(c, A, nitems, q0_first, q0_count, it0, d_lw);
    if (c.nb == 1)
        return f64 ? launch_small_t<double, 1>(c, A, nitems, q0_first, q0_count, it0, d_lw)
                   : launch_small_t<float, 1>(c, A, nitems, q0_first, q0_count, it0, d_lw);
    set_error("small-cell kernel supports Nb <= 2");
    return SPR_E_UNSUPPORTED;
}

// FP64 W of every grid q1 for irreducible q0 [q_first, q_first + q_count) into
// d_w[q0 - q_first][q1][s0][s1][s2] (Nb <= 2): the pair-reduced items of those q0, each pair
// writing both orderings (q -> -(q0 + q) is an involution, so every q1 is written exactly once),
// no underflow skip.
int launch_lw3_small_w(Ctx &c, int q_first, int q_count, double *d_w) {
    if (c.nb > 2 || c.override_items) {
        set_error("small-cell W path: Nb <= 2 and no item override");
        return SPR_E_UNSUPPORTED;
    }
    int rc;
    if ((rc = ensure_pairs(c))) return rc;
    const int it0 = c.h_item_first[q_first];
    const int nitems = c.h_item_first[q_first + q_count] - it0;
    SmallArgs A;
    A.item_q0 = c.d_item_q0 + it0;
    A.item_off = c.d_item_off + it0;
    A.item_cnt = c.d_item_cnt + it0;
    A.pairs = c.d_pairs;
    A.q0_grid = c.irr_grid;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.blk_start = c.fc3.blk_start_p;
    A.tup_site = c.fc3.tup_site_p;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sigma = 1.0;   // unused without the delta stage
    const double pi = 4.0 * std::atan(1.0);
    A.factor = 1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) *
               1000000.0 * pi * 0.0625 / (double)c.nq;   // Sctr3THz2 / Nq, constants.f90:59-60
    A.ic.partial = c.ipartial;
    A.w_out = d_w;
    A.sel = nullptr;   // W output: the tetrahedron sums apply the branch ranges themselves
    A.pair_w = nullptr;
    A.w_q0_first = q_first;
    A.w_nq = c.nq;
    A.w_one = 0;
    // the per-item sums of this launch are zero; they are reduced into a scratch row per q0
    if ((rc = ensure((void **)&c.wsink, &c.wsink_cap, (size_t)q_count * c.nd * sizeof(double)))) return rc;
    if (c.nb == 2) return launch_small_t<double, 2>(c, A, nitems, q_first, q_count, it0, c.wsink, true);
    return launch_small_t<double, 1>(c, A, nitems, q_first, q_count, it0, c.wsink, true);
}

// FP64 W of every grid q1 for the q0 of caller-supplied items (c.override_items, every q1 of a
// q0 listed with its partner), Nb <= 2: d_w[k][q1][s0][s1][s2] for the q0 of grid index
// q0_grid[k], k < nq0 (item_q0 holds k); each orientation evaluated on its own.
int launch_lw3_small_w_items(Ctx &c, const int *d_q0_grid, int nq0, double *d_w) {
    const ItemSet *ov = c.override_items;
    if (c.nb > 2 || !ov) {
        set_error("small-cell W path for listed items: Nb <= 2 and an item list");
        return SPR_E_UNSUPPORTED;
    }
    int rc;
    SmallArgs A;
    A.item_q0 = ov->item_q0;
    A.item_off = ov->item_off;
    A.item_cnt = ov->item_cnt;
    A.pairs = ov->pairs;
    A.q0_grid = d_q0_grid;
    A.L = c.L;
    A.m0 = c.mesh[0];
    A.m1 = c.mesh[1];
    A.blk_start = c.fc3.blk_start_p;
    A.tup_site = c.fc3.tup_site_p;
    A.site = c.fc3.site;
    A.nsite = c.fc3.nsite;
    A.omega = c.omega;
    A.inv_omega = c.inv_omega;
    A.nbe = c.nbe;
    A.sigma = 1.0;   // unused without the delta stage
    const double pi = 4.0 * std::atan(1.0);
    A.factor = 1.054571817 * (1.60217733 * 1.60217733) * (6.02214076 * 6.02214076 * 6.02214076) *
               1000000.0 * pi * 0.0625 / (double)c.nq;   // Sctr3THz2 / Nq, constants.f90:59-60
    A.ic.partial = ov->partial;
    A.w_out = d_w;
    A.sel = nullptr;
    A.pair_w = nullptr;
    A.w_q0_first = 0;
    A.w_nq = c.nq;
    A.w_one = 1;
    if ((rc = ensure((void **)&c.wsink, &c.wsink_cap, (size_t)nq0 * c.nd * sizeof(double)))) return rc;
    if (c.nb == 2) return launch_small_t<double, 2>(c, A, ov->nitems, 0, nq0, 0, c.wsink, true);
    return launch_small_t<double, 1>(c, A, ov->nitems, 0, nq0, 0, c.wsink, true);
}

}  // namespace spr
