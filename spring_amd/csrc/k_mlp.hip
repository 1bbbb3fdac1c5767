// Learned importance-sampling proposal: a 3-layer MLP (F -> 128 -> 128 -> 1, tanh) that
// maps (q0, q1, w0, w1, w2, branch ids) to log W, with hand-written training (FP32, Adam)
// and inference (FP32 VALU, or BF16 operands on v_mfma_f32_16x16x32_bf16 for the 128x128
// layer with FP32 accumulation).  No counterpart in the reference; it serves the sampled
// estimate of the linewidth sum whose deterministic form is CalculateLinewidth
// (src/Kappa/Src/ThirdOrder_m.f90:442-543).
//
// Parameter vector (FP32): W1[F][H] b1[H] W2[H][H] b2[H] W3[H] b3.
// Training step: forward + backward per 32-row tile (per-tile gradient partials), then a
// fixed-order reduction over tiles fused with the Adam update, so training is deterministic.
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "spr_internal.h"

namespace spr {

constexpr int MH = 128;      // hidden width
constexpr int MROWS = 32;    // rows per training tile
constexpr int MFMAX = 32;    // max input features

typedef float v4f_m __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_m __attribute__((ext_vector_type(8)));

struct MlpParamsView {
    const float *W1, *b1, *W2, *b2, *W3, *b3;
};

__host__ __device__ inline size_t mlp_nparam(int F) { return (size_t)F * MH + MH + MH * MH + MH + MH + 1; }

__device__ inline MlpParamsView mlp_view(const float *p, int F) {
    MlpParamsView v;
    v.W1 = p;
    v.b1 = v.W1 + F * MH;
    v.W2 = v.b1 + MH;
    v.b2 = v.W2 + MH * MH;
    v.W3 = v.b2 + MH;
    v.b3 = v.W3 + MH;
    return v;
}

// Forward (and, with TRAIN, backward) over one tile of MROWS rows.  256 threads: row
// r = tid / 8, hidden slice [16 (tid % 8), +16).
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp_tile(const float *param, int F, const float *X, const float *y,
                                                  int n, float *pred, float *grad_part, double *loss_part) {
    extern __shared__ __align__(16) float sm[];
    float *s_W2 = sm;                        // [MH][MH]
    float *s_h1 = s_W2 + MH * MH;            // [MROWS][MH]
    float *s_h2 = s_h1 + MROWS * MH;         // [MROWS][MH]
    float *s_dz2 = s_h2 + MROWS * MH;        // [MROWS][MH]
    float *s_x = s_dz2 + MROWS * MH;         // [MROWS][MFMAX]
    float *s_dy = s_x + MROWS * MFMAX;       // [MROWS]
    const int tid = threadIdx.x;
    const int r = tid >> 3, js = (tid & 7) * 16;
    const int row = blockIdx.x * MROWS + r;
    const bool valid = row < n;
    const MlpParamsView P = mlp_view(param, F);
    for (int i = tid; i < MH * MH; i += 256) s_W2[i] = P.W2[i];
    for (int i = tid; i < MROWS * F; i += 256) {
        const int rr = i / F, ff = i % F;
        const int gr = blockIdx.x * MROWS + rr;
        s_x[rr * MFMAX + ff] = gr < n ? X[(size_t)gr * F + ff] : 0.f;
    }
    __syncthreads();
    // layer 1
    for (int jj = 0; jj < 16; ++jj) {
        const int j = js + jj;
        float z = P.b1[j];
        for (int f = 0; f < F; ++f) z = fmaf(s_x[r * MFMAX + f], P.W1[f * MH + j], z);
        s_h1[r * MH + j] = tanhf(z);
    }
    __syncthreads();
    // layer 2
    float h2[16];
    float ypart = 0.f;
    for (int jj = 0; jj < 16; ++jj) {
        const int j = js + jj;
        float z = P.b2[j];
        for (int k = 0; k < MH; ++k) z = fmaf(s_h1[r * MH + k], s_W2[k * MH + j], z);
        h2[jj] = tanhf(z);
        s_h2[r * MH + j] = h2[jj];
        ypart = fmaf(h2[jj], P.W3[j], ypart);
    }
    // sum over the 8 threads of the row (consecutive lanes of one wave)
    ypart += __shfl_xor(ypart, 1, 64);
    ypart += __shfl_xor(ypart, 2, 64);
    ypart += __shfl_xor(ypart, 4, 64);
    const float yp = ypart + P.b3[0];
    if (!TRAIN) {
        if (valid && (tid & 7) == 0) pred[row] = yp;
        return;
    }
    const float err = valid ? yp - y[row] : 0.f;
    const float dy = 2.f * err / (float)n;
    if ((tid & 7) == 0) s_dy[r] = dy;
    for (int jj = 0; jj < 16; ++jj) {
        const int j = js + jj;
        s_dz2[r * MH + j] = dy * P.W3[j] * (1.f - h2[jj] * h2[jj]);
    }
    __syncthreads();
    float *g = grad_part + (size_t)blockIdx.x * mlp_nparam(F);
    float *gW1 = g, *gb1 = gW1 + F * MH, *gW2 = gb1 + MH, *gb2 = gW2 + MH * MH, *gW3 = gb2 + MH, *gb3 = gW3 + MH;
    // dW2[k][j] = sum_r h1[r][k] dz2[r][j]: thread owns an 8 x 8 block
    {
        const int k0 = (tid >> 4) * 8, j0 = (tid & 15) * 8;
        float acc[8][8];
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) acc[a][b] = 0.f;
        for (int rr = 0; rr < MROWS; ++rr)
            for (int a = 0; a < 8; ++a) {
                const float h = s_h1[rr * MH + k0 + a];
                for (int b = 0; b < 8; ++b) acc[a][b] = fmaf(h, s_dz2[rr * MH + j0 + b], acc[a][b]);
            }
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) gW2[(k0 + a) * MH + j0 + b] = acc[a][b];
    }
    // db2, dW3, db3
    if (tid < MH) {
        float s2 = 0.f, s3 = 0.f;
        for (int rr = 0; rr < MROWS; ++rr) {
            s2 += s_dz2[rr * MH + tid];
            s3 = fmaf(s_h2[rr * MH + tid], s_dy[rr], s3);
        }
        gb2[tid] = s2;
        gW3[tid] = s3;
    }
    if (tid == 0) {
        float s = 0.f;
        for (int rr = 0; rr < MROWS; ++rr) s += s_dy[rr];
        gb3[0] = s;
    }
    // dz1[r][k] = (sum_j dz2[r][j] W2[k][j]) (1 - h1^2); reuse s_h2 for dz1
    __syncthreads();
    for (int kk = 0; kk < 16; ++kk) {
        const int k = js + kk;
        float d = 0.f;
        for (int j = 0; j < MH; ++j) d = fmaf(s_dz2[r * MH + j], s_W2[k * MH + j], d);
        const float h = s_h1[r * MH + k];
        s_h2[r * MH + k] = d * (1.f - h * h);
    }
    __syncthreads();
    for (int e = tid; e < F * MH; e += 256) {
        const int f = e / MH, k = e % MH;
        float s = 0.f;
        for (int rr = 0; rr < MROWS; ++rr) s = fmaf(s_x[rr * MFMAX + f], s_h2[rr * MH + k], s);
        gW1[e] = s;
    }
    if (tid < MH) {
        float s = 0.f;
        for (int rr = 0; rr < MROWS; ++rr) s += s_h2[rr * MH + tid];
        gb1[tid] = s;
    }
    if (tid == 0) {
        double l = 0.0;
        for (int rr = 0; rr < MROWS; ++rr) {
            const int gr = blockIdx.x * MROWS + rr;
            if (gr < n) {
                const double e = (double)s_dy[rr] * (double)n * 0.5;
                l += e * e;
            }
        }
        loss_part[blockIdx.x] = l;
    }
}

// g = sum over tiles (fixed order); optional Adam step on the parameters.
__global__ void k_mlp_reduce_adam(const float *grad_part, int ntiles, size_t np, float *grad, float *param,
                                  float *m, float *v, int step, float lr, int do_step) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    float g = 0.f;
    for (int t = 0; t < ntiles; ++t) g += grad_part[(size_t)t * np + i];
    grad[i] = g;
    if (!do_step) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)step));
    const float vh = vi / (1.f - powf(b2, (float)step));
    param[i] -= lr * mh / (sqrtf(vh) + eps);
}

// BF16 inference: 64 rows per workgroup (4 waves).  Layer 1 and the output layer on VALU
// (FP32, inputs rounded to BF16), layer 2 as a 64 x 128 x 128 GEMM on BF16 MFMA.
__global__ __launch_bounds__(256) void k_mlp_infer_bf16(const float *param, int F, const float *X, int n,
                                                        float *pred) {
    constexpr int R = 64;
    extern __shared__ __align__(16) unsigned char smb[];
    __bf16 *s_W2 = reinterpret_cast<__bf16 *>(smb);          // [MH][MH] (k, j)
    __bf16 *s_h1 = s_W2 + MH * MH;                            // [R][MH]
    float *s_y = reinterpret_cast<float *>(s_h1 + R * MH);    // [R][8] partials
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
    const MlpParamsView P = mlp_view(param, F);
    for (int i = tid; i < MH * MH; i += 256) s_W2[i] = (__bf16)P.W2[i];
    // layer 1: thread -> (row tid/4, 32 hidden units)
    {
        const int r = tid >> 2, j0 = (tid & 3) * 32;
        const int row = blockIdx.x * R + r;
        float x[MFMAX];
        for (int f = 0; f < F; ++f) x[f] = row < n ? (float)(__bf16)X[(size_t)row * F + f] : 0.f;
        for (int jj = 0; jj < 32; ++jj) {
            const int j = j0 + jj;
            float z = (float)(__bf16)P.b1[j];
            for (int f = 0; f < F; ++f) z = fmaf(x[f], (float)(__bf16)P.W1[f * MH + j], z);
            s_h1[r * MH + j] = (__bf16)tanhf(z);
        }
    }
    __syncthreads();
    // layer 2 on MFMA: wave w takes row tile w (16 rows) x all 8 column tiles
    float yacc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int nt = 0; nt < MH / 16; ++nt) {
        v4f_m acc = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < MH; kk += 32) {
            bf16x8_m a, b;
            for (int t = 0; t < 8; ++t) {
                a[t] = s_h1[(w * 16 + (l & 15)) * MH + kk + 8 * (l >> 4) + t];
                b[t] = s_W2[(kk + 8 * (l >> 4) + t) * MH + nt * 16 + (l & 15)];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
        }
        const int j = nt * 16 + (l & 15);
        const float b2 = (float)(__bf16)P.b2[j], w3 = (float)(__bf16)P.W3[j];
        for (int q = 0; q < 4; ++q) yacc[q] = fmaf((float)(__bf16)tanhf(acc[q] + b2), w3, yacc[q]);
    }
    // reduce over the 16 columns held by lanes with the same row group
    for (int q = 0; q < 4; ++q) {
        float v = yacc[q];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        yacc[q] = v;
    }
    if ((l & 15) == 0)
        for (int q = 0; q < 4; ++q) {
            const int r = w * 16 + 4 * (l >> 4) + q;
            const int row = blockIdx.x * R + r;
            if (row < n) pred[row] = yacc[q] + P.b3[0];
        }
}


// ---- proposal scores on the device ------------------------------------------------------
// For each requested q0 and every q1 (q2 = -(q0 + q1)):
//   S(q0, q1) = sum over branch triplets with min_k |d_k| < window sigma of exp(sd f(x)) |G|,
// d1 = w0 - w1 - w2, d2 = w0 + w1 - w2, d3 = w0 - w1 + w2, G = (n1 + n2 + 1) g(d1) +
// (n1 - n2) (g(d2) - g(d3)) with g the normalised Gaussian, and f(x) the network's normalised
// log W prediction for the feature row x of the triplet (the rows of spring_amd.learned.features);
// without a network (param == NULL) the weight is |G| alone.
// A workgroup (4 waves) owns a work item (q0, 4 consecutive q1); each wave works alone on its
// share of the (q1, s1, s2) combinations: a lane tests every s0 of its combination (a bit mask),
// resonant triplets are appended in lane order to the wave's LDS queue, and every 16 queued rows
// run through the network on the wave's own tile (layer 1, K = 32, and layer 2, 128 x 128, on
// BF16 MFMA with FP32 accumulation, tanh(x) = 1 - 2 / (1 + e^{2x})).  Per-q1 sums are formed in a
// fixed order (lane, wave), so the scores are reproducible.  Persistent workgroups loop over the
// work items.
constexpr int SC_QPB = 4;       // q1 per work item
constexpr int SC_NT = 256;      // threads (4 waves)
constexpr int SC_NW = SC_NT / 64;
constexpr int SC_WQ = 16 + 64;  // per-wave queue capacity
constexpr int SC_PW1 = 40;      // row pitch (bf16) of X and W1^T
constexpr int SC_PW2 = 136;     // row pitch (bf16) of H1 and W2^T

struct ScoreArgs {
    const float *param;          // network parameters, or NULL for |G| alone
    int F;
    const int *q0_grid;
    int nq0s, nq, m0, m1, m2;
    const double *omega, *nbe;
    double sigma, window;
    float sd, inv_wmax;
    int *counter;
    int nwork, nchunk;
    double *S;
    unsigned long long *n_eval;
};

__device__ __forceinline__ float fast_tanh(float x) { return 1.f - __fdividef(2.f, 1.f + __expf(2.f * x)); }

__device__ __forceinline__ int fold_coord(int k, int n) { return k > n / 2 ? k - n : k; }

template <int ND>
__global__ __launch_bounds__(SC_NT) void k_mlp_score(ScoreArgs A) {
    constexpr int ND2 = ND * ND, ND3 = ND2 * ND;
    static_assert(ND <= 32, "s0 mask is 32 bits");
    extern __shared__ __align__(16) unsigned char smsc[];
    __bf16 *sW1T = reinterpret_cast<__bf16 *>(smsc);         // [MH][SC_PW1]  (j, f)
    __bf16 *sW2T = sW1T + MH * SC_PW1;                        // [MH][SC_PW2]  (j2, j1)
    float *sB1 = reinterpret_cast<float *>(sW2T + MH * SC_PW2);
    float *sB2 = sB1 + MH, *sW3 = sB2 + MH;
    double *sMd = reinterpret_cast<double *>(sW3 + MH);      // w0[32] | (w1 n1 w2 n2)[QPB][32]
    double *sWS = sMd + (1 + 4 * SC_QPB) * 32;                // [SC_NW][SC_QPB]
    int *sItem = reinterpret_cast<int *>(sWS + SC_NW * SC_QPB);
    unsigned char *wbase = reinterpret_cast<unsigned char *>(sItem + 4);
    const int tid = threadIdx.x, l = tid & 63, w = tid >> 6, g = l >> 4;
    // per-wave region: X[16][SC_PW1] bf16 | H1[16][SC_PW2] bf16 | key[SC_WQ] | G[SC_WQ] | out[16]
    constexpr size_t WREG = (16 * SC_PW1 + 16 * SC_PW2) * 2 + SC_WQ * 8 + 16 * 4;
    __bf16 *wX = reinterpret_cast<__bf16 *>(wbase + w * WREG);
    __bf16 *wH = wX + 16 * SC_PW1;
    int *wK = reinterpret_cast<int *>(wH + 16 * SC_PW2);
    float *wG = reinterpret_cast<float *>(wK + SC_WQ);
    float *wO = wG + SC_WQ;
    const bool net = A.param != nullptr;
    if (net) {
        const MlpParamsView P = mlp_view(A.param, A.F);
        for (int i = tid; i < MH * SC_PW1; i += SC_NT) {
            const int j = i / SC_PW1, f = i % SC_PW1;
            sW1T[i] = (__bf16)(f < A.F ? P.W1[f * MH + j] : 0.f);
        }
        for (int i = tid; i < MH * MH; i += SC_NT) sW2T[(i % MH) * SC_PW2 + i / MH] = (__bf16)P.W2[i];
        for (int i = tid; i < MH; i += SC_NT) {
            sB1[i] = P.b1[i];
            sB2[i] = P.b2[i];
            sW3[i] = P.W3[i];
        }
    }
    const float b3 = net ? A.param[mlp_nparam(A.F) - 1] : 0.f;
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002), hinv = 0.5 / (A.sigma * A.sigma);
    const double win = A.window * A.sigma;
    const float fnd = ND > 1 ? 1.f / (float)(ND - 1) : 0.f;

    for (;;) {
        __syncthreads();
        if (tid == 0) sItem[0] = atomicAdd(A.counter, 1);
        __syncthreads();
        const int wi = sItem[0];
        if (wi >= A.nwork) break;
        const int qi = wi / A.nchunk, q1lo = (wi % A.nchunk) * SC_QPB;
        const int nql = min(SC_QPB, A.nq - q1lo);
        const int i0 = A.q0_grid[qi];
        const int k00 = i0 % A.m0, k01 = (i0 / A.m0) % A.m1, k02 = i0 / (A.m0 * A.m1);
        if (tid < ND) sMd[tid] = A.omega[(size_t)i0 * ND + tid];
        for (int i = tid; i < nql * ND; i += SC_NT) {
            const int ql = i / ND, s = i % ND, q1 = q1lo + ql;
            const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
            const int a0 = ((-(k00 + k10)) % A.m0 + A.m0) % A.m0, a1 = ((-(k01 + k11)) % A.m1 + A.m1) % A.m1,
                      a2 = ((-(k02 + k12)) % A.m2 + A.m2) % A.m2;
            const int q2 = a0 + A.m0 * (a1 + A.m1 * a2);
            sMd[32 + (4 * ql + 0) * 32 + s] = A.omega[(size_t)q1 * ND + s];
            sMd[32 + (4 * ql + 1) * 32 + s] = A.nbe[(size_t)q1 * ND + s];
            sMd[32 + (4 * ql + 2) * 32 + s] = A.omega[(size_t)q2 * ND + s];
            sMd[32 + (4 * ql + 3) * 32 + s] = A.nbe[(size_t)q2 * ND + s];
        }
        __syncthreads();
        const float k0x = (float)fold_coord(k00, A.m0) / A.m0, k0y = (float)fold_coord(k01, A.m1) / A.m1,
                    k0z = (float)fold_coord(k02, A.m2) / A.m2;
        // w0 of every s0 in registers (uniform across lanes)
        double w0r[ND];
#pragma unroll
        for (int s = 0; s < ND; ++s) w0r[s] = sMd[s];

        double accq[SC_QPB] = {0.0, 0.0, 0.0, 0.0};   // lane 0: this wave's per-q1 sums
        unsigned long long nres = 0;
        int qn = 0;                                    // wave queue length (uniform)

        // the wave's network tile over queue rows [0, nv), nv <= 16; adds to accq
        auto tile = [&](int nv) {
            if (l < 16) {
                float xs[32];
#pragma unroll
                for (int f = 0; f < 32; ++f) xs[f] = 0.f;
                if (l < nv) {
                    const int key = wK[l];
                    const int ql = key / ND3, r = key % ND3;
                    const int s0 = r / ND2, s1 = (r / ND) % ND, s2 = r % ND;
                    const int q1 = q1lo + ql;
                    const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
                    xs[0] = k0x;
                    xs[1] = k0y;
                    xs[2] = k0z;
                    xs[3] = (float)fold_coord(k10, A.m0) / A.m0;
                    xs[4] = (float)fold_coord(k11, A.m1) / A.m1;
                    xs[5] = (float)fold_coord(k12, A.m2) / A.m2;
                    xs[6] = (float)sMd[s0] * A.inv_wmax;
                    xs[7] = (float)sMd[32 + (4 * ql + 0) * 32 + s1] * A.inv_wmax;
                    xs[8] = (float)sMd[32 + (4 * ql + 2) * 32 + s2] * A.inv_wmax;
                    xs[9] = s0 * fnd;
                    xs[10] = s1 * fnd;
                    xs[11] = s2 * fnd;
                }
#pragma unroll
                for (int f = 0; f < 32; ++f) wX[l * SC_PW1 + f] = (__bf16)xs[f];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            {
                bf16x8_m a;
#pragma unroll
                for (int t = 0; t < 8; ++t) a[t] = wX[(l & 15) * SC_PW1 + 8 * g + t];
#pragma unroll
                for (int nt = 0; nt < MH / 16; ++nt) {
                    bf16x8_m b;
#pragma unroll
                    for (int t = 0; t < 8; ++t) b[t] = sW1T[(nt * 16 + (l & 15)) * SC_PW1 + 8 * g + t];
                    v4f_m z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, v4f_m{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    const int j = nt * 16 + (l & 15);
#pragma unroll
                    for (int r = 0; r < 4; ++r) wH[(4 * g + r) * SC_PW2 + j] = (__bf16)fast_tanh(z[r] + sB1[j]);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            float yacc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int nt = 0; nt < MH / 16; ++nt) {
                v4f_m z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kk = 0; kk < MH; kk += 32) {
                    bf16x8_m a, b;
#pragma unroll
                    for (int t = 0; t < 8; ++t) {
                        a[t] = wH[(l & 15) * SC_PW2 + kk + 8 * g + t];
                        b[t] = sW2T[(nt * 16 + (l & 15)) * SC_PW2 + kk + 8 * g + t];
                    }
                    z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, z, 0, 0, 0);
                }
                const int j = nt * 16 + (l & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) yacc[r] = fmaf(fast_tanh(z[r] + sB2[j]), sW3[j], yacc[r]);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = yacc[r];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                v += __shfl_xor(v, 8, 64);
                yacc[r] = v;
            }
            if ((l & 15) == 0)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 4 * g + r;
                    wO[row] = row < nv ? __expf(A.sd * (yacc[r] + b3)) * wG[row] : 0.f;
                }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // per-q1 sums of the tile (lanes 0..15 hold the rows; fixed reduction order)
            const int ql = l < nv ? wK[l] / ND3 : -1;
            const double o = l < 16 ? (double)wO[l] : 0.0;
#pragma unroll
            for (int q = 0; q < SC_QPB; ++q) {
                double v = ql == q ? o : 0.0;
#pragma unroll
                for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
                accq[q] += v;   // every lane holds the same tile sum (lane 0 is the one used)
            }
        };

        for (int pb = w * 64; pb < nql * ND2; pb += SC_NT) {
            const int pidx = pb + l;
            const bool pv = pidx < nql * ND2;
            const int ql = pv ? pidx / ND2 : 0, s1 = pv ? (pidx / ND) % ND : 0, s2 = pv ? pidx % ND : 0;
            const double w1 = sMd[32 + (4 * ql + 0) * 32 + s1], n1 = sMd[32 + (4 * ql + 1) * 32 + s1];
            const double w2 = sMd[32 + (4 * ql + 2) * 32 + s2], n2 = sMd[32 + (4 * ql + 3) * 32 + s2];
            const double ap = w1 + w2, am = w1 - w2;   // d1 = w0 - ap, d2 = w0 + am, d3 = w0 - am
            unsigned mask = 0;
            if (pv) {
#pragma unroll
                for (int s0 = 0; s0 < ND; ++s0) {
                    const double w0 = w0r[s0];
                    const double dm = fmin(fabs(w0 - ap), fmin(fabs(w0 + am), fabs(w0 - am)));
                    mask |= (dm < win ? 1u : 0u) << s0;
                }
            }
            // emit the resonant triplets in (lane, s0) order
            while (__ballot(mask != 0)) {
                const bool has = mask != 0;
                const int s0 = has ? __ffs(mask) - 1 : 0;
                mask &= has ? mask - 1 : mask;
                float G = 0.f;
                if (has) {
                    const double w0 = sMd[s0];
                    const double d1 = w0 - ap, d2 = w0 + am, d3 = w0 - am;
                    G = (float)fabs((n1 + n2 + 1.0) * c0 * exp(-hinv * d1 * d1) +
                                    (n1 - n2) * (c0 * exp(-hinv * d2 * d2) - c0 * exp(-hinv * d3 * d3)));
                }
                const unsigned long long bal = __ballot(has);
                const int cnt = __popcll(bal);
                nres += (unsigned long long)cnt;
                if (!net) {
                    // no network: the weight is |G|, summed per lane in emission order
#pragma unroll
                    for (int q = 0; q < SC_QPB; ++q)
                        if (has && ql == q) accq[q] += (double)G;
                    continue;
                }
                const int rank = __popcll(bal & ((1ull << l) - 1ull));
                if (has) {
                    wK[qn + rank] = ql * ND3 + s0 * ND2 + s1 * ND + s2;
                    wG[qn + rank] = G;
                }
                qn += cnt;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                while (qn >= 16) {
                    tile(16);
                    // shift the remaining entries (< 64) to the front, in order
                    int kq = 0;
                    float gq = 0.f;
                    const bool mv = l + 16 < qn;
                    if (mv) {
                        kq = wK[l + 16];
                        gq = wG[l + 16];
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (mv) {
                        wK[l] = kq;
                        wG[l] = gq;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    qn -= 16;
                }
            }
        }
        if (net && qn > 0) tile(qn);
        // this wave's per-q1 totals: with a network every lane holds them; without, reduce lanes
#pragma unroll
        for (int q = 0; q < SC_QPB; ++q) {
            double v = accq[q];
            if (!net) {
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            }
            if (l == 0) sWS[w * SC_QPB + q] = v;
        }
        if (l == 0) atomicAdd(A.n_eval, net ? nres : 0ull);
        __syncthreads();
        if (tid < nql) {
            double t = 0.0;
            for (int ww = 0; ww < SC_NW; ++ww) t += sWS[ww * SC_QPB + tid];
            A.S[(size_t)qi * A.nq + q1lo + tid] = t;
        }
    }
}

size_t score_lds() {
    const size_t wreg = (16 * SC_PW1 + 16 * SC_PW2) * 2 + SC_WQ * 8 + 16 * 4;
    return (size_t)(MH * SC_PW1 + MH * SC_PW2) * 2 + 3 * MH * 4 + ((1 + 4 * SC_QPB) * 32 + SC_NW * SC_QPB) * 8 +
           4 * 4 + SC_NW * wreg + 16;
}

}  // namespace spr

using namespace spr;

struct spr_mlp {
    int F = 0;
    size_t np = 0;
    float *param = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *gpart = nullptr;
    float *X = nullptr, *y = nullptr, *pred = nullptr;
    double *lpart = nullptr;
    size_t cap_rows = 0, cap_tiles = 0;
    int step = 0;
    hipStream_t stream = nullptr;
};

static int mlp_reserve(spr_mlp *h, size_t rows) {
    if (rows <= h->cap_rows) return SPR_OK;
    (void)hipFree(h->X);
    (void)hipFree(h->y);
    (void)hipFree(h->pred);
    (void)hipFree(h->gpart);
    (void)hipFree(h->lpart);
    const size_t tiles = (rows + MROWS - 1) / MROWS;
    if (hipMalloc((void **)&h->X, rows * h->F * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->y, rows * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->pred, rows * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->gpart, tiles * h->np * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->lpart, tiles * sizeof(double)) != hipSuccess) {
        set_error("spr_mlp: hipMalloc failed");
        return SPR_E_OOM;
    }
    h->cap_rows = rows;
    h->cap_tiles = tiles;
    return SPR_OK;
}

extern "C" int spr_mlp_create(int n_features, unsigned long long seed, int device, spr_mlp **out) {
    if (!out || n_features < 1 || n_features > MFMAX) {
        set_error("spr_mlp_create: n_features must be 1..32");
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(device));
    auto *h = new spr_mlp();
    h->F = n_features;
    h->np = mlp_nparam(n_features);
    std::vector<float> p(h->np, 0.f);
    std::mt19937_64 g(seed);
    auto init = [&](size_t off, int fan_in, size_t count) {
        std::normal_distribution<float> d(0.f, 1.f / std::sqrt((float)fan_in));
        for (size_t i = 0; i < count; ++i) p[off + i] = d(g);
    };
    const int F = n_features;
    init(0, F, (size_t)F * MH);                                   // W1
    init((size_t)F * MH + MH, MH, (size_t)MH * MH);               // W2
    init((size_t)F * MH + MH + MH * MH + MH, MH, MH);             // W3
    if (hipMalloc((void **)&h->param, h->np * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->m, h->np * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->v, h->np * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&h->grad, h->np * sizeof(float)) != hipSuccess) {
        delete h;
        set_error("spr_mlp_create: hipMalloc failed");
        return SPR_E_OOM;
    }
    SPR_HIP(hipMemcpy(h->param, p.data(), h->np * sizeof(float), hipMemcpyHostToDevice));
    SPR_HIP(hipMemset(h->m, 0, h->np * sizeof(float)));
    SPR_HIP(hipMemset(h->v, 0, h->np * sizeof(float)));
    SPR_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    *out = h;
    return SPR_OK;
}

extern "C" int spr_mlp_destroy(spr_mlp *h) {
    if (!h) return SPR_OK;
    for (void *p : {(void *)h->param, (void *)h->m, (void *)h->v, (void *)h->grad, (void *)h->gpart,
                    (void *)h->X, (void *)h->y, (void *)h->pred, (void *)h->lpart})
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return SPR_OK;
}

extern "C" int spr_mlp_num_params(spr_mlp *h, long *np) {
    if (!h || !np) return SPR_E_ARG;
    *np = (long)h->np;
    return SPR_OK;
}

extern "C" int spr_mlp_get_params(spr_mlp *h, float *out) {
    if (!h || !out) return SPR_E_ARG;
    SPR_HIP(hipMemcpy(out, h->param, h->np * sizeof(float), hipMemcpyDeviceToHost));
    return SPR_OK;
}

extern "C" int spr_mlp_set_params(spr_mlp *h, const float *in) {
    if (!h || !in) return SPR_E_ARG;
    SPR_HIP(hipMemcpy(h->param, in, h->np * sizeof(float), hipMemcpyHostToDevice));
    return SPR_OK;
}

static const size_t kTileLds = (size_t)(MH * MH + 3 * MROWS * MH + MROWS * MFMAX + MROWS) * sizeof(float);

// Mean-squared-error gradient of the loss over n rows (no update): grad(np), loss.
extern "C" int spr_mlp_grad(spr_mlp *h, const float *X, const float *y, long n, float *grad, double *loss) {
    if (!h || !X || !y || n < 1 || !grad) {
        set_error("spr_mlp_grad: bad argument");
        return SPR_E_ARG;
    }
    int rc;
    if ((rc = mlp_reserve(h, n))) return rc;
    SPR_HIP(hipMemcpy(h->X, X, n * h->F * sizeof(float), hipMemcpyHostToDevice));
    SPR_HIP(hipMemcpy(h->y, y, n * sizeof(float), hipMemcpyHostToDevice));
    const int tiles = (int)((n + MROWS - 1) / MROWS);
    SPR_HIP(hipFuncSetAttribute((const void *)k_mlp_tile<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTileLds));
    hipLaunchKernelGGL(k_mlp_tile<true>, dim3(tiles), dim3(256), kTileLds, h->stream, h->param, h->F, h->X, h->y,
                       (int)n, h->pred, h->gpart, h->lpart);
    hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((h->np + 255) / 256), dim3(256), 0, h->stream, h->gpart, tiles,
                       h->np, h->grad, h->param, h->m, h->v, 1, 0.f, 0);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipMemcpyAsync(grad, h->grad, h->np * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    std::vector<double> lp(tiles);
    SPR_HIP(hipMemcpyAsync(lp.data(), h->lpart, tiles * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SPR_HIP(hipStreamSynchronize(h->stream));
    if (loss) {
        double s = 0.0;
        for (double v : lp) s += v;
        *loss = s / (double)n;
    }
    return SPR_OK;
}

// Mini-batch Adam over the data set for `epochs` passes (batches in order, data shuffled
// by the caller); loss_per_epoch[epochs] gets the mean training loss of each epoch.
extern "C" int spr_mlp_train(spr_mlp *h, const float *X, const float *y, long n, int epochs, int batch, float lr,
                             double *loss_per_epoch) {
    if (!h || !X || !y || n < 1 || epochs < 1 || batch < MROWS) {
        set_error("spr_mlp_train: bad argument (batch must be >= 32)");
        return SPR_E_ARG;
    }
    int rc;
    if ((rc = mlp_reserve(h, n))) return rc;
    SPR_HIP(hipMemcpy(h->X, X, n * h->F * sizeof(float), hipMemcpyHostToDevice));
    SPR_HIP(hipMemcpy(h->y, y, n * sizeof(float), hipMemcpyHostToDevice));
    SPR_HIP(hipFuncSetAttribute((const void *)k_mlp_tile<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTileLds));
    std::vector<double> lp;
    for (int ep = 0; ep < epochs; ++ep) {
        double lsum = 0.0;
        for (long b0 = 0; b0 < n; b0 += batch) {
            const long nb = std::min<long>(batch, n - b0);
            const int tiles = (int)((nb + MROWS - 1) / MROWS);
            h->step += 1;
            hipLaunchKernelGGL(k_mlp_tile<true>, dim3(tiles), dim3(256), kTileLds, h->stream, h->param, h->F,
                               h->X + b0 * h->F, h->y + b0, (int)nb, h->pred, h->gpart, h->lpart);
            hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((h->np + 255) / 256), dim3(256), 0, h->stream, h->gpart,
                               tiles, h->np, h->grad, h->param, h->m, h->v, h->step, lr, 1);
            lp.resize(tiles);
            SPR_HIP(hipMemcpyAsync(lp.data(), h->lpart, tiles * sizeof(double), hipMemcpyDeviceToHost, h->stream));
            SPR_HIP(hipStreamSynchronize(h->stream));
            for (double v : lp) lsum += v;
        }
        SPR_HIP(hipGetLastError());
        if (loss_per_epoch) loss_per_epoch[ep] = lsum / (double)n;
    }
    return SPR_OK;
}

// Predictions for n rows; prec SPR_PREC_FP32 (VALU) or SPR_PREC_BF16 (MFMA layer 2).
extern "C" int spr_mlp_predict(spr_mlp *h, const float *X, long n, int prec, float *out) {
    if (!h || !X || n < 1 || !out || (prec != SPR_PREC_FP32 && prec != SPR_PREC_BF16)) {
        set_error("spr_mlp_predict: bad argument (prec must be FP32 or BF16)");
        return SPR_E_ARG;
    }
    int rc;
    if ((rc = mlp_reserve(h, n))) return rc;
    SPR_HIP(hipMemcpy(h->X, X, n * h->F * sizeof(float), hipMemcpyHostToDevice));
    if (prec == SPR_PREC_FP32) {
        const int tiles = (int)((n + MROWS - 1) / MROWS);
        SPR_HIP(hipFuncSetAttribute((const void *)k_mlp_tile<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)kTileLds));
        hipLaunchKernelGGL(k_mlp_tile<false>, dim3(tiles), dim3(256), kTileLds, h->stream, h->param, h->F, h->X,
                           nullptr, (int)n, h->pred, nullptr, nullptr);
    } else {
        const size_t lds = (size_t)MH * MH * 2 + 64 * MH * 2 + 64 * 8 * sizeof(float);
        SPR_HIP(hipFuncSetAttribute((const void *)k_mlp_infer_bf16, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((n + 63) / 64), dim3(256), lds, h->stream, h->param, h->F, h->X,
                           (int)n, h->pred);
    }
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipMemcpyAsync(out, h->pred, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    SPR_HIP(hipStreamSynchronize(h->stream));
    return SPR_OK;
}

extern "C" int spr_mlp_score(spr_ctx *ctx, spr_mlp *m, int n, const int *q0_index, double sigma, double window,
                             float y_std, float wmax, double *S_out, long long *n_eval) {
    if (!ctx || n < 1 || !q0_index || !S_out || sigma <= 0.0 || window <= 0.0 || wmax <= 0.f) {
        set_error("spr_mlp_score: bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = ctx->c;
    if (m && m->F != 12) {
        set_error("spr_mlp_score: the network must take the 12 proposal features");
        return SPR_E_ARG;
    }
    if (c.nd > 32) {
        set_error("spr_mlp_score: Ndof must be <= 32");
        return SPR_E_UNSUPPORTED;
    }
    std::vector<int> grid(n);
    for (int k = 0; k < n; ++k) {
        if (q0_index[k] < 0 || q0_index[k] >= c.nq0) {
            set_error("spr_mlp_score: q0 index out of range");
            return SPR_E_ARG;
        }
        grid[k] = c.h_irr_grid[q0_index[k]];
    }
    SPR_HIP(hipSetDevice(c.device));
    int *d_grid = nullptr, *d_cnt = nullptr;
    double *d_S = nullptr;
    unsigned long long *d_ne = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(d_grid);
        (void)hipFree(d_cnt);
        (void)hipFree(d_S);
        (void)hipFree(d_ne);
    };
    if (hipMalloc((void **)&d_grid, n * sizeof(int)) != hipSuccess || hipMalloc((void **)&d_cnt, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_S, (size_t)n * c.nq * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_ne, sizeof(unsigned long long)) != hipSuccess) {
        cleanup();
        set_error("spr_mlp_score: hipMalloc failed");
        return SPR_E_OOM;
    }
    int rc = SPR_OK;
    do {
        if (hipMemcpy(d_grid, grid.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(d_cnt, 0, sizeof(int)) != hipSuccess || hipMemset(d_ne, 0, sizeof(unsigned long long)) != hipSuccess) {
            set_error("spr_mlp_score: upload failed");
            rc = SPR_E_HIP;
            break;
        }
        ScoreArgs A;
        A.param = m ? m->param : nullptr;
        A.F = m ? m->F : 12;
        A.q0_grid = d_grid;
        A.nq0s = n;
        A.nq = c.nq;
        A.m0 = c.mesh[0];
        A.m1 = c.mesh[1];
        A.m2 = c.mesh[2];
        A.omega = c.omega;
        A.nbe = c.nbe;
        A.sigma = sigma;
        A.window = window;
        A.sd = y_std;
        A.inv_wmax = 1.f / wmax;
        A.counter = d_cnt;
        A.nchunk = (c.nq + SC_QPB - 1) / SC_QPB;
        A.nwork = n * A.nchunk;
        A.S = d_S;
        A.n_eval = d_ne;
        const size_t lds = score_lds();
        const void *fn = nullptr;
        switch (c.nd) {
            case 6: fn = (const void *)k_mlp_score<6>; break;
            case 9: fn = (const void *)k_mlp_score<9>; break;
            case 12: fn = (const void *)k_mlp_score<12>; break;
            case 15: fn = (const void *)k_mlp_score<15>; break;
            case 18: fn = (const void *)k_mlp_score<18>; break;
            case 24: fn = (const void *)k_mlp_score<24>; break;
            case 30: fn = (const void *)k_mlp_score<30>; break;
            default: break;
        }
        if (!fn) {
            set_error("spr_mlp_score: Ndof " + std::to_string(c.nd) + " not built (6-18, 24, 30)");
            rc = SPR_E_UNSUPPORTED;
            break;
        }
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
            set_error("spr_mlp_score: LDS attribute failed");
            rc = SPR_E_HIP;
            break;
        }
        const int grid_n = std::min(A.nwork, 2 * c.ncu);
        (void)hipEventRecord(c.ev0, c.stream);
        {
            void *args[] = {&A};
            if (hipLaunchKernel(fn, dim3(grid_n), dim3(SC_NT), args, lds, c.stream) != hipSuccess) {
                set_error("spr_mlp_score: launch failed");
                rc = SPR_E_HIP;
                break;
            }
        }
        (void)hipEventRecord(c.ev1, c.stream);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c.stream) != hipSuccess) {
            set_error("spr_mlp_score: kernel failed");
            rc = SPR_E_HIP;
            break;
        }
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c.ev0, c.ev1) == hipSuccess) {
            c.launches += 1;
            c.kernel_ms += ms;
        }
        unsigned long long ne = 0;
        if (hipMemcpy(S_out, d_S, (size_t)n * c.nq * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&ne, d_ne, sizeof(ne), hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("spr_mlp_score: download failed");
            rc = SPR_E_HIP;
            break;
        }
        if (n_eval) *n_eval = (long long)ne;
    } while (false);
    cleanup();
    return rc;
}

// ---- pair-level proposal: features of (q0, q1) and the network's score of the pair ---------
// The proposal draws q1 for all branches of a q0 at once, so it scores pairs, not triplets.  For
// each (q0, q1), q2 = -(q0 + q1), a scan over the branch triplets within window sigma of a
// resonance gives, per s0 (skipping s0 with 1/w0 = 0),
//   g(s0) = sum |G| / (w1 w2),  h1(s0), h2(s0) = the same weighted by w1 / w2,
//   a(s0) = the same restricted to s1 < 3 or s2 < 3 (acoustic branches), and the triplet count;
// with Gtot(s0) = sum over q1 of g(s0) the pair's resonance weight is S = sum_s0 g / Gtot (the
// proposal without a network; sum over q1 of S = number of s0 with Gtot > 0).  The network maps
// the pair's feature row (reduced q0, q1; log(S Nq / ns0); the S-weighted means of w1, w2, w0 over
// wmax; the acoustic share; log(1 + count) / log(Ndof^3)) to the log of the pair's measured share
// of the linewidths (sum_s0 |t(q1, s0)| / sum_q1 |t(q1, s0)|, normalised the same way).
namespace spr {

constexpr int PF_NF = 12;

struct PairScanArgs {
    const int *q0_grid;   // the batch's q0 (grid indices)
    int nq, m0, m1, m2;
    const double *omega, *nbe, *inv_omega;
    double sigma, window;
    float *acc;           // [batch][nq][Ndof][4]: g, h1, h2, a
    int *cnt;             // [batch][nq]
};

// one workgroup per (q1, q0 of the batch); thread (s0 = tid % 32, slice = tid / 32) scans the
// (s1, s2) combinations slice, slice + 8, ... in registers; slices are summed in a fixed order
template <int ND>
__global__ __launch_bounds__(256) void k_pair_scan(PairScanArgs A) {
    const int q1 = blockIdx.x, kb = blockIdx.y, tid = threadIdx.x;
    const int i0 = A.q0_grid[kb];
    const int k00 = i0 % A.m0, k01 = (i0 / A.m0) % A.m1, k02 = i0 / (A.m0 * A.m1);
    const int k10 = q1 % A.m0, k11 = (q1 / A.m0) % A.m1, k12 = q1 / (A.m0 * A.m1);
    const int a0 = ((-(k00 + k10)) % A.m0 + A.m0) % A.m0, a1 = ((-(k01 + k11)) % A.m1 + A.m1) % A.m1,
              a2 = ((-(k02 + k12)) % A.m2 + A.m2) % A.m2;
    const int q2 = a0 + A.m0 * (a1 + A.m1 * a2);
    __shared__ double sm[8][32];   // w0 iw0 | w1 n1 iw1 | w2 n2 iw2
    __shared__ float red[8][32][4];
    __shared__ int redc[256];
    if (tid < ND) {
        sm[0][tid] = A.omega[(size_t)i0 * ND + tid];
        sm[1][tid] = A.inv_omega[(size_t)i0 * ND + tid];
        sm[2][tid] = A.omega[(size_t)q1 * ND + tid];
        sm[3][tid] = A.nbe[(size_t)q1 * ND + tid];
        sm[4][tid] = A.inv_omega[(size_t)q1 * ND + tid];
        sm[5][tid] = A.omega[(size_t)q2 * ND + tid];
        sm[6][tid] = A.nbe[(size_t)q2 * ND + tid];
        sm[7][tid] = A.inv_omega[(size_t)q2 * ND + tid];
    }
    __syncthreads();
    const int s0 = tid & 31, sl = tid >> 5;
    const double c0 = 1.0 / (A.sigma * 2.5066282746310002), hinv = 0.5 / (A.sigma * A.sigma);
    const double win = A.window * A.sigma;
    float g = 0.f, h1 = 0.f, h2 = 0.f, ac = 0.f;
    int c = 0;
    if (s0 < ND && sm[1][s0] != 0.0) {
        const double w0 = sm[0][s0];
        for (int p = sl; p < ND * ND; p += 8) {
            const int s1 = p / ND, s2 = p - s1 * ND;
            const double w1 = sm[2][s1], w2 = sm[5][s2];
            const double d1 = w0 - w1 - w2, d2 = w0 + w1 - w2, d3 = w0 - w1 + w2;
            if (fmin(fabs(d1), fmin(fabs(d2), fabs(d3))) >= win) continue;
            const double n1 = sm[3][s1], n2 = sm[6][s2];
            const double G = (n1 + n2 + 1.0) * (c0 * exp(-hinv * d1 * d1)) +
                             (n1 - n2) * (c0 * exp(-hinv * d2 * d2) - c0 * exp(-hinv * d3 * d3));
            const float wgt = (float)(fabs(G) * sm[4][s1] * sm[7][s2]);
            g += wgt;
            h1 += wgt * (float)w1;
            h2 += wgt * (float)w2;
            if (s1 < 3 || s2 < 3) ac += wgt;
            ++c;
        }
    }
    red[sl][s0][0] = g;
    red[sl][s0][1] = h1;
    red[sl][s0][2] = h2;
    red[sl][s0][3] = ac;
    redc[tid] = c;
    __syncthreads();
    if (tid < ND * 4) {
        const int s = tid >> 2, k = tid & 3;
        float t = 0.f;
        for (int j = 0; j < 8; ++j) t += red[j][s][k];
        A.acc[(((size_t)kb * A.nq + q1) * ND + s) * 4 + k] = t;
    }
    if (tid == 0) {
        int t = 0;
        for (int j = 0; j < 256; ++j) t += redc[j];
        A.cnt[(size_t)kb * A.nq + q1] = t;
    }
}

// Gtot(kb, s0) = sum over q1 of g (strided per-thread sums, fixed tree), one workgroup per q0
__global__ __launch_bounds__(256) void k_pair_gtot(const float *acc, int nq, int nd, double *gtot) {
    const int kb = blockIdx.x, tid = threadIdx.x;
    __shared__ double r[256];
    for (int s = 0; s < nd; ++s) {
        double t = 0.0;
        for (int q = tid; q < nq; q += 256) t += acc[(((size_t)kb * nq + q) * nd + s) * 4];
        r[tid] = t;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) r[tid] += r[tid + w];
            __syncthreads();
        }
        if (tid == 0) gtot[(size_t)kb * nd + s] = r[0];
        __syncthreads();
    }
}

// feature rows X[kb][q1][12] and the resonance weight S[kb][q1]
__global__ __launch_bounds__(256) void k_pair_feat(const float *acc, const int *cnt, const double *gtot,
                                                   const int *q0_grid, const double *omega, int nq, int nd, int m0,
                                                   int m1, int m2, float inv_wmax, float *X, double *S) {
    const int kb = blockIdx.y, q1 = blockIdx.x * 256 + threadIdx.x;
    if (q1 >= nq) return;
    const int i0 = q0_grid[kb];
    double s = 0.0, e1 = 0.0, e2 = 0.0, e0 = 0.0, ea = 0.0;
    int nact = 0;
    for (int s0 = 0; s0 < nd; ++s0) {
        const double gt = gtot[(size_t)kb * nd + s0];
        if (!(gt > 0.0)) continue;
        ++nact;
        const float *a = acc + (((size_t)kb * nq + q1) * nd + s0) * 4;
        s += a[0] / gt;
        e1 += a[1] / gt;
        e2 += a[2] / gt;
        e0 += a[0] / gt * omega[(size_t)i0 * nd + s0];
        ea += a[3] / gt;
    }
    auto fold = [](int k, int n) { return (float)(k > n / 2 ? k - n : k) / (float)n; };
    float *x = X + ((size_t)kb * nq + q1) * PF_NF;
    x[0] = fold(i0 % m0, m0);
    x[1] = fold((i0 / m0) % m1, m1);
    x[2] = fold(i0 / (m0 * m1), m2);
    x[3] = fold(q1 % m0, m0);
    x[4] = fold((q1 / m0) % m1, m1);
    x[5] = fold(q1 / (m0 * m1), m2);
    x[6] = (float)log(fmax(s * nq / (nact > 0 ? nact : 1), 1e-12));
    const bool any = s > 0.0;
    x[7] = any ? (float)(e1 / s) * inv_wmax : 0.f;
    x[8] = any ? (float)(e2 / s) * inv_wmax : 0.f;
    x[9] = any ? (float)(e0 / s) * inv_wmax : 0.f;
    x[10] = any ? (float)(ea / s) : 0.f;
    x[11] = (float)(log1p((double)cnt[(size_t)kb * nq + q1]) / log((double)nd * nd * nd));
    S[(size_t)kb * nq + q1] = s;
}

}  // namespace spr

extern "C" int spr_pair_scores(spr_ctx *ctx, spr_mlp *m, int n, const int *q0_index, double sigma, double window,
                               float wmax, double *S_out, float *X_out) {
    if (!ctx || n < 1 || !q0_index || !S_out || !(sigma > 0.0) || !(window > 0.0) || !(wmax > 0.f)) {
        set_error("spr_pair_scores: bad argument");
        return SPR_E_ARG;
    }
    if (m && m->F != PF_NF) {
        set_error("spr_pair_scores: the network must take the 12 pair features");
        return SPR_E_ARG;
    }
    Ctx &c = ctx->c;
    const int nd = c.nd, nq = c.nq;
    const void *scan = nullptr;
    switch (nd) {
        case 3: scan = (const void *)k_pair_scan<3>; break;
        case 6: scan = (const void *)k_pair_scan<6>; break;
        case 9: scan = (const void *)k_pair_scan<9>; break;
        case 12: scan = (const void *)k_pair_scan<12>; break;
        case 15: scan = (const void *)k_pair_scan<15>; break;
        case 18: scan = (const void *)k_pair_scan<18>; break;
        case 24: scan = (const void *)k_pair_scan<24>; break;
        case 30: scan = (const void *)k_pair_scan<30>; break;
        default:
            set_error("spr_pair_scores: Ndof " + std::to_string(nd) + " not built (3-18, 24, 30)");
            return SPR_E_UNSUPPORTED;
    }
    std::vector<int> grid(n);
    for (int k = 0; k < n; ++k) {
        if (q0_index[k] < 0 || q0_index[k] >= c.nq0) {
            set_error("spr_pair_scores: q0 index out of range");
            return SPR_E_ARG;
        }
        grid[k] = c.h_irr_grid[q0_index[k]];
    }
    SPR_HIP(hipSetDevice(c.device));
    // q0 in batches of B: per batch acc (B nq Ndof 4 floats), counts, Gtot, X, S, predictions
    const int B = (int)std::max<size_t>(1, std::min<size_t>((size_t)n, ((size_t)1 << 30) / ((size_t)nq * (nd * 16 + 64))));
    float *d_acc = nullptr, *d_X = nullptr, *d_pred = nullptr;
    int *d_cnt = nullptr, *d_grid = nullptr;
    double *d_gt = nullptr, *d_S = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(d_acc);
        (void)hipFree(d_X);
        (void)hipFree(d_pred);
        (void)hipFree(d_cnt);
        (void)hipFree(d_grid);
        (void)hipFree(d_gt);
        (void)hipFree(d_S);
    };
    const size_t rows = (size_t)B * nq;
    if (hipMalloc((void **)&d_acc, rows * nd * 4 * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&d_X, rows * PF_NF * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&d_pred, rows * sizeof(float)) != hipSuccess ||
        hipMalloc((void **)&d_cnt, rows * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_grid, (size_t)n * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_gt, (size_t)B * nd * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_S, rows * sizeof(double)) != hipSuccess) {
        cleanup();
        set_error("spr_pair_scores: hipMalloc failed");
        return SPR_E_OOM;
    }
    int rc = SPR_OK;
    std::vector<float> pred;
    do {
        if (hipMemcpy(d_grid, grid.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("spr_pair_scores: upload failed");
            rc = SPR_E_HIP;
            break;
        }
        for (int b0 = 0; b0 < n && rc == SPR_OK; b0 += B) {
            const int nb = std::min(B, n - b0);
            PairScanArgs A;
            A.q0_grid = d_grid + b0;
            A.nq = nq;
            A.m0 = c.mesh[0];
            A.m1 = c.mesh[1];
            A.m2 = c.mesh[2];
            A.omega = c.omega;
            A.nbe = c.nbe;
            A.inv_omega = c.inv_omega;
            A.sigma = sigma;
            A.window = window;
            A.acc = d_acc;
            A.cnt = d_cnt;
            void *args[] = {&A};
            if (hipLaunchKernel(scan, dim3(nq, nb), dim3(256), args, 0, c.stream) != hipSuccess) {
                set_error("spr_pair_scores: scan launch failed");
                rc = SPR_E_HIP;
                break;
            }
            hipLaunchKernelGGL(k_pair_gtot, dim3(nb), dim3(256), 0, c.stream, d_acc, nq, nd, d_gt);
            hipLaunchKernelGGL(k_pair_feat, dim3((nq + 255) / 256, nb), dim3(256), 0, c.stream, d_acc, d_cnt, d_gt,
                               d_grid + b0, c.omega, nq, nd, c.mesh[0], c.mesh[1], c.mesh[2], 1.f / wmax, d_X, d_S);
            if (m) {
                const size_t lds = (size_t)MH * MH * 2 + 64 * MH * 2 + 64 * 8 * sizeof(float);
                SPR_HIP(hipFuncSetAttribute((const void *)k_mlp_infer_bf16,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(k_mlp_infer_bf16, dim3((unsigned)((size_t)nb * nq + 63) / 64), dim3(256), lds,
                                   c.stream, m->param, PF_NF, d_X, nb * nq, d_pred);
            }
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c.stream) != hipSuccess) {
                set_error("spr_pair_scores: kernel failed");
                rc = SPR_E_HIP;
                break;
            }
            const size_t nrow = (size_t)nb * nq;
            bool ok = true;
            if (m) {
                pred.resize(nrow);
                ok = hipMemcpy(pred.data(), d_pred, nrow * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess;
                for (size_t i = 0; ok && i < nrow; ++i) S_out[(size_t)b0 * nq + i] = pred[i];
            } else {
                ok = hipMemcpy(S_out + (size_t)b0 * nq, d_S, nrow * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
            }
            if (ok && X_out)
                ok = hipMemcpy(X_out + (size_t)b0 * nq * PF_NF, d_X, nrow * PF_NF * sizeof(float),
                               hipMemcpyDeviceToHost) == hipSuccess;
            if (!ok) {
                set_error("spr_pair_scores: download failed");
                rc = SPR_E_HIP;
            }
        }
    } while (false);
    cleanup();
    return rc;
}
