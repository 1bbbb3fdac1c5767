// spr_lw3_pairs_noalloc: the per-pair three-phonon terms of spr_lw3_pairs without any device
// allocation, for callers that cannot allocate (code under HIP graph capture, tight sampling or
// serving loops, harnesses that forbid hipMalloc).
//
// Every device buffer the call needs beyond the context's resident tables (the uploaded pair
// indices, the per-pair work items, the partial rows of every work unit and the reduced output)
// lives in `scratch`, a file-scope device array of kNoallocBytes that each GPU the library runs
// on holds until the process ends.  In host code the array's name is only a host-side stand-in:
// its device address is taken once per device with hipGetSymbolAddress when the first context on
// that device is created (spr_ctx_create calls noalloc_setup with the device current), and only
// that address is ever passed to kernels or copies.
//
// Calls share their GPU's array: two calls that overlap on one GPU (two host threads, or two
// contexts on one device used concurrently) race.  Nothing in the array has to start at zero:
// every region is written before it is read within one call.
#include <cstdint>
#include <mutex>

#include "spr_internal.h"

namespace spr {

int launch_lw3_small(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);
int launch_lw3_mm(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);

namespace {

constexpr int kNoallocMaxPairs = SPR_NOALLOC_MAX_PAIRS;
constexpr int kNoallocMaxNd = 32;
// layout: q0_index, q1_grid, item_off, item_cnt (ints), pairs (int2), partial rows
// (pairs x kMaxUnitsPerItem x Ndof doubles), reduced output (pairs x Ndof doubles)
constexpr size_t kOffQ0 = 0;
constexpr size_t kOffQ1 = kOffQ0 + (size_t)kNoallocMaxPairs * sizeof(int);
constexpr size_t kOffOff = kOffQ1 + (size_t)kNoallocMaxPairs * sizeof(int);
constexpr size_t kOffCnt = kOffOff + (size_t)kNoallocMaxPairs * sizeof(int);
constexpr size_t kOffPairs = kOffCnt + (size_t)kNoallocMaxPairs * sizeof(int);
constexpr size_t kOffRows = kOffPairs + (size_t)kNoallocMaxPairs * sizeof(int2);
constexpr size_t kOffOut = kOffRows + (size_t)kNoallocMaxPairs * kMaxUnitsPerItem * kNoallocMaxNd * sizeof(double);
constexpr size_t kNoallocBytes = kOffOut + (size_t)kNoallocMaxPairs * kNoallocMaxNd * sizeof(double);

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

// Resolves this device's copy of `scratch` (the device must be current).
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("noalloc: device ordinal beyond the scratch table");
        return SPR_E_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return SPR_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        set_error(std::string("noalloc: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        return SPR_E_HIP;
    }
    g_scratch[device] = p;
    return SPR_OK;
}

// item k = pair k alone: offset k, one pair, (q1, q2 = -(q0 + q1)) from the grid coordinates
__global__ void k_noalloc_items(int n, const int *q0_index, const int *q1_grid, const int *irr_grid, int m0, int m1,
                                int m2, int *item_off, int *item_cnt, int2 *pairs) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int i0 = irr_grid[q0_index[k]], q1 = q1_grid[k];
    const int a0 = ((-(i0 % m0 + q1 % m0)) % m0 + m0) % m0;
    const int a1 = ((-((i0 / m0) % m1 + (q1 / m0) % m1)) % m1 + m1) % m1;
    const int a2 = ((-(i0 / (m0 * m1) + q1 / (m0 * m1))) % m2 + m2) % m2;
    item_off[k] = k;
    item_cnt[k] = 1;
    pairs[k] = make_int2(q1, a0 + m0 * (a1 + m1 * a2));
}

// out[k][s] = sum over the pair's units h of rows[k * upi + h][s], halved when q1 != q2 (the
// kernels weight such pairs by 2); the same order and arithmetic as spr_lw3_pairs' host loop
__global__ void k_noalloc_reduce(int n, int nd, int upi, const double *rows, const int2 *pairs, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * nd) return;
    const int k = i / nd, s = i % nd;
    double t = 0.0;
    for (int h = 0; h < upi; ++h) t += rows[((size_t)k * upi + h) * nd + s];
    if (pairs[k].x != pairs[k].y) t *= 0.5;
    out[i] = t;
}

}  // namespace spr

using namespace spr;

extern "C" int spr_lw3_pairs_noalloc(spr_ctx *h, int n, const int *q0_index, const int *q1_grid, double sigma,
                                     int prec, double *out) {
    if (!h || n < 0 || (n > 0 && (!q0_index || !q1_grid || !out))) {
        set_error("spr_lw3_pairs_noalloc: bad argument");
        return SPR_E_ARG;
    }
    if (n > kNoallocMaxPairs) {
        set_error("spr_lw3_pairs_noalloc: " + std::to_string(n) + " pairs exceed the scratch limit of " +
                  std::to_string(kNoallocMaxPairs) + " (SPR_NOALLOC_MAX_PAIRS); split the call");
        return SPR_E_ARG;
    }
    if (prec != SPR_PREC_FP64 && prec != SPR_PREC_FP32 && prec != SPR_PREC_BF16) {
        set_error("unknown precision mode");
        return SPR_E_ARG;
    }
    if (!(sigma > 0.0)) {
        set_error("sigma must be positive");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    if (c.nd > kNoallocMaxNd) {
        set_error("spr_lw3_pairs_noalloc: Ndof above 32");
        return SPR_E_UNSUPPORTED;
    }
    for (int k = 0; k < n; ++k)
        if (q0_index[k] < 0 || q0_index[k] >= c.nq0 || q1_grid[k] < 0 || q1_grid[k] >= c.nq) {
            set_error("spr_lw3_pairs_noalloc: index out of range");
            return SPR_E_ARG;
        }
    if (n == 0) return SPR_OK;
    SPR_HIP(hipSetDevice(c.device));
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = static_cast<unsigned char *>(g_scratch[c.device]);
    }
    if (!base) {
        set_error("spr_lw3_pairs_noalloc: scratch of this device not resolved (no context created on it)");
        return SPR_E_STATE;
    }
    int *d_q0 = reinterpret_cast<int *>(base + kOffQ0);
    int *d_q1 = reinterpret_cast<int *>(base + kOffQ1);
    int *d_off = reinterpret_cast<int *>(base + kOffOff);
    int *d_cnt = reinterpret_cast<int *>(base + kOffCnt);
    int2 *d_pairs = reinterpret_cast<int2 *>(base + kOffPairs);
    double *d_rows = reinterpret_cast<double *>(base + kOffRows);
    double *d_out = reinterpret_cast<double *>(base + kOffOut);
    SPR_HIP(hipMemcpyAsync(d_q0, q0_index, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c.stream));
    SPR_HIP(hipMemcpyAsync(d_q1, q1_grid, (size_t)n * sizeof(int), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_noalloc_items, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, d_q0, d_q1, c.irr_grid,
                       c.mesh[0], c.mesh[1], c.mesh[2], d_off, d_cnt, d_pairs);
    SPR_HIP(hipGetLastError());
    ItemSet set{d_q0, d_off, d_cnt, d_pairs, n, d_rows};
    c.override_items = &set;
    int rc = (prec == SPR_PREC_BF16 || c.nb > 2) ? launch_lw3_mm(c, 0, 0, sigma, prec, nullptr)
                                                 : launch_lw3_small(c, 0, 0, sigma, prec, nullptr);
    c.override_items = nullptr;
    if (rc) return rc;
    const int upi = c.last_upi;
    hipLaunchKernelGGL(k_noalloc_reduce, dim3((n * c.nd + 255) / 256), dim3(256), 0, c.stream, n, c.nd, upi,
                       (const double *)d_rows, (const int2 *)d_pairs, d_out);
    SPR_HIP(hipGetLastError());
    SPR_HIP(hipMemcpyAsync(out, d_out, (size_t)n * c.nd * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SPR_HIP(hipStreamSynchronize(c.stream));
    return SPR_OK;
}
