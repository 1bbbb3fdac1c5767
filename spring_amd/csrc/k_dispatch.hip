// Chooses the kernel variant for each entry point.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>
#include <string>

#include "spr_internal.h"

namespace spr {

int launch_lw3_small(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);
int launch_lw3_mm(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);

// Kernel choice.
//   W parity path (FP64 W of every q1)          generic kernel
//   FP64:  Nb <= 2 small-cell batch kernel, else MFMA kernel (FP64 MFMA)
//   FP32:  Nb <= 2 small-cell batch kernel (FP32 VALU), else MFMA kernel (FP32 MFMA)
//   BF16:  MFMA kernel (split-BF16 MFMA)
// SPR_KERNEL=generic|small|mfma overrides the choice where the variant supports the mode.
int launch_lw3(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw,
               double *d_w_out) {
    if (d_w_out) return launch_lw3_generic(c, c.irr_grid + q0_first, q0_count, sigma, prec, d_lw, d_w_out);
    const char *f = getenv("SPR_KERNEL");
    const std::string force = f ? f : "";
    if (force == "generic" && prec != SPR_PREC_BF16)
        return launch_lw3_generic(c, c.irr_grid + q0_first, q0_count, sigma, prec, d_lw, nullptr);
    if (force == "mfma") return launch_lw3_mm(c, q0_first, q0_count, sigma, prec, d_lw);
    if (force == "small" && c.nb <= 2 && prec != SPR_PREC_BF16)
        return launch_lw3_small(c, q0_first, q0_count, sigma, prec, d_lw);
    if (prec != SPR_PREC_BF16 && c.nb <= 2) return launch_lw3_small(c, q0_first, q0_count, sigma, prec, d_lw);
    return launch_lw3_mm(c, q0_first, q0_count, sigma, prec, d_lw);
}


}  // namespace spr

namespace spr {

// Per-pair linewidth contributions for caller-chosen (q0, q1) pairs: out[k][s0] =
// sum_{s1,s2} W [delta terms] of pair k alone (the q1 term of CalculateLinewidth's sum).
int launch_lw3_pairs(Ctx &c, int n, const int *q0_index, const int *q1_grid, double sigma, int prec,
                     double *h_out) {
    if (n < 1) return SPR_OK;
    int rc;
    if ((rc = ensure_pairs(c))) return rc;
    std::vector<int> iq0(n), ioff(n), icnt(n, 1);
    std::vector<int2> pr(n);
    const int m0 = c.mesh[0], m1 = c.mesh[1], m2 = c.mesh[2];
    for (int k = 0; k < n; ++k) {
        if (q0_index[k] < 0 || q0_index[k] >= c.nq0 || q1_grid[k] < 0 || q1_grid[k] >= c.nq) {
            set_error("spr_lw3_pairs: index out of range");
            return SPR_E_ARG;
        }
        const int i0 = c.h_irr_grid[q0_index[k]], q1 = q1_grid[k];
        const int a0 = ((-(i0 % m0 + q1 % m0)) % m0 + m0) % m0;
        const int a1 = ((-((i0 / m0) % m1 + (q1 / m0) % m1)) % m1 + m1) % m1;
        const int a2 = ((-(i0 / (m0 * m1) + q1 / (m0 * m1))) % m2 + m2) % m2;
        iq0[k] = q0_index[k];
        ioff[k] = k;
        pr[k] = make_int2(q1, a0 + m0 * (a1 + m1 * a2));
    }
    int *d_i = nullptr;
    int2 *d_p = nullptr;
    double *d_out = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(d_i);
        (void)hipFree(d_p);
        (void)hipFree(d_out);
        c.override_items = nullptr;
    };
    if (hipMalloc((void **)&d_i, 3 * (size_t)n * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_p, (size_t)n * sizeof(int2)) != hipSuccess ||
        hipMalloc((void **)&d_out, (size_t)n * kMaxUnitsPerItem * c.nd * sizeof(double)) != hipSuccess) {
        cleanup();
        set_error("spr_lw3_pairs: hipMalloc failed");
        return SPR_E_OOM;
    }
    if (hipMemcpy(d_i, iq0.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_i + n, ioff.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_i + 2 * n, icnt.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_p, pr.data(), n * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        set_error("spr_lw3_pairs: upload failed");
        return SPR_E_HIP;
    }
    ItemSet set{d_i, d_i + n, d_i + 2 * n, d_p, n, d_out};
    c.override_items = &set;
    rc = (prec == SPR_PREC_BF16 || c.nb > 2) ? launch_lw3_mm(c, 0, 0, sigma, prec, nullptr)
                                             : launch_lw3_small(c, 0, 0, sigma, prec, nullptr);
    c.override_items = nullptr;
    // every unit of a pair wrote its own partial row: rows k * upi + h, h < upi
    const int upi = c.last_upi;
    std::vector<double> rows((size_t)n * upi * c.nd);
    if (!rc && hipMemcpy(rows.data(), d_out, rows.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("spr_lw3_pairs: download failed");
        rc = SPR_E_HIP;
    }
    cleanup();
    if (rc) return rc;
    for (int k = 0; k < n; ++k)
        for (int s = 0; s < c.nd; ++s) {
            double t = 0.0;
            for (int h = 0; h < upi; ++h) t += rows[((size_t)k * upi + h) * c.nd + s];
            h_out[(size_t)k * c.nd + s] = t;
        }
    for (int k = 0; k < n; ++k)
        if (pr[k].x != pr[k].y)
            for (int s = 0; s < c.nd; ++s) h_out[(size_t)k * c.nd + s] *= 0.5;   // kernels weight q1 != q2 by 2
    return SPR_OK;
}

}  // namespace spr

namespace spr {

// Weighted sums of q1 terms per irreducible q0: est[qi][s0] = sum over the samples k with
// q0_index[k] == qi of weight[k] * t(q0_k, q1_k; s0), t being the q1 term of CalculateLinewidth's
// sum (ThirdOrder_m.f90:442-543), i.e. what spr_lw3_pairs returns for the pair.  The samples are
// grouped (stably) by q0 into items of up to 64 pairs (q1, q2 = -(q0 + q1)), each carrying its
// weight into the kernel in place of the 1-or-2 pair weight of the full sum, so a sampled estimate
// runs on the same whole-item kernels as the full sum; the per-item sums are added per q0 in item
// order on the host (bitwise reproducible).
int launch_lw3_sampled(Ctx &c, int n, const int *q0_index, const int *q1_grid, const double *weight, double sigma,
                       int prec, double *h_est) {
    for (size_t i = 0; i < (size_t)c.nq0 * c.nd; ++i) h_est[i] = 0.0;
    if (n < 1) return SPR_OK;
    const int m0 = c.mesh[0], m1 = c.mesh[1], m2 = c.mesh[2];
    std::vector<int> order(n), cnt_q0(c.nq0 + 1, 0);
    for (int k = 0; k < n; ++k) {
        if (q0_index[k] < 0 || q0_index[k] >= c.nq0 || q1_grid[k] < 0 || q1_grid[k] >= c.nq) {
            set_error("spr_lw3_sampled: index out of range");
            return SPR_E_ARG;
        }
        if (!std::isfinite(weight[k])) {
            set_error("spr_lw3_sampled: weight not finite");
            return SPR_E_ARG;
        }
        ++cnt_q0[q0_index[k] + 1];
    }
    for (int q = 0; q < c.nq0; ++q) cnt_q0[q + 1] += cnt_q0[q];
    {
        std::vector<int> at(cnt_q0.begin(), cnt_q0.end() - 1);
        for (int k = 0; k < n; ++k) order[at[q0_index[k]]++] = k;   // stable by q0
    }
    std::vector<int2> pr(n);
    std::vector<double> pw(n);
    std::vector<int> iq0, ioff, icnt;
    for (int q = 0; q < c.nq0; ++q) {
        const int i0 = c.h_irr_grid[q];
        const int k0 = i0 % m0, k1 = (i0 / m0) % m1, k2 = i0 / (m0 * m1);
        for (int j = cnt_q0[q]; j < cnt_q0[q + 1]; ++j) {
            const int k = order[j], q1 = q1_grid[k];
            const int a0 = ((-(k0 + q1 % m0)) % m0 + m0) % m0;
            const int a1 = ((-(k1 + (q1 / m0) % m1)) % m1 + m1) % m1;
            const int a2 = ((-(k2 + q1 / (m0 * m1))) % m2 + m2) % m2;
            pr[j] = make_int2(q1, a0 + m0 * (a1 + m1 * a2));
            pw[j] = weight[k];
        }
        for (int j = cnt_q0[q]; j < cnt_q0[q + 1]; j += 64) {
            iq0.push_back(q);
            ioff.push_back(j);
            icnt.push_back(std::min(64, cnt_q0[q + 1] - j));
        }
    }
    const int nitems = (int)iq0.size();
    int *d_i = nullptr;
    int2 *d_p = nullptr;
    double *d_w = nullptr, *d_out = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(d_i);
        (void)hipFree(d_p);
        (void)hipFree(d_w);
        (void)hipFree(d_out);
        c.override_items = nullptr;
    };
    if (hipMalloc((void **)&d_i, 3 * (size_t)nitems * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&d_p, (size_t)n * sizeof(int2)) != hipSuccess ||
        hipMalloc((void **)&d_w, (size_t)n * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&d_out, (size_t)nitems * kMaxUnitsPerItem * c.nd * sizeof(double)) != hipSuccess) {
        cleanup();
        set_error("spr_lw3_sampled: hipMalloc failed");
        return SPR_E_OOM;
    }
    if (hipMemcpy(d_i, iq0.data(), nitems * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_i + nitems, ioff.data(), nitems * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_i + 2 * nitems, icnt.data(), nitems * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_p, pr.data(), n * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_w, pw.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        cleanup();
        set_error("spr_lw3_sampled: upload failed");
        return SPR_E_HIP;
    }
    ItemSet set{d_i, d_i + nitems, d_i + 2 * nitems, d_p, nitems, d_out, d_w, true};
    c.override_items = &set;
    int rc = (prec == SPR_PREC_BF16 || c.nb > 2) ? launch_lw3_mm(c, 0, 0, sigma, prec, nullptr)
                                                 : launch_lw3_small(c, 0, 0, sigma, prec, nullptr);
    c.override_items = nullptr;
    // every unit of an item wrote its own partial row: rows i * upi + h, h < upi
    const int upi = c.last_upi;
    std::vector<double> rows((size_t)nitems * upi * c.nd);
    if (!rc && hipMemcpy(rows.data(), d_out, rows.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("spr_lw3_sampled: download failed");
        rc = SPR_E_HIP;
    }
    cleanup();
    if (rc) return rc;
    for (int i = 0; i < nitems; ++i)
        for (int h = 0; h < upi; ++h)
            for (int s = 0; s < c.nd; ++s) h_est[(size_t)iq0[i] * c.nd + s] += rows[((size_t)i * upi + h) * c.nd + s];
    return SPR_OK;
}

}  // namespace spr
