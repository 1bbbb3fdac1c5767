// Chooses the kernel variant for each entry point.
#include <cstdlib>
#include <string>

#include "spr_internal.h"

namespace spr {

int launch_lw3_generic(Ctx &c, const int *d_q0, int nq0, double sigma, int prec, double *d_lw,
                       double *d_w);

int launch_lw3_small(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);
int launch_lw3_mfma(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);

// Kernel choice.
//   W parity path (FP64 W of every q1)          generic kernel
//   FP64:  Nb <= 2 small-cell batch kernel, else generic kernel
//   FP32:  Nb <= 2 small-cell batch kernel (FP32 VALU), else MFMA kernel (FP32 MFMA)
//   BF16:  MFMA kernel (BF16 MFMA)
// SPR_KERNEL=generic|small|mfma overrides the choice where the variant supports the mode.
int launch_lw3(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw,
               double *d_w_out) {
    if (d_w_out) return launch_lw3_generic(c, c.irr_grid + q0_first, q0_count, sigma, prec, d_lw, d_w_out);
    const char *f = getenv("SPR_KERNEL");
    const std::string force = f ? f : "";
    if (force == "generic" && prec != SPR_PREC_BF16)
        return launch_lw3_generic(c, c.irr_grid + q0_first, q0_count, sigma, prec, d_lw, nullptr);
    if (force == "mfma" && prec != SPR_PREC_FP64) return launch_lw3_mfma(c, q0_first, q0_count, sigma, prec, d_lw);
    if (force == "small" && c.nb <= 2 && prec != SPR_PREC_BF16)
        return launch_lw3_small(c, q0_first, q0_count, sigma, prec, d_lw);
    if (prec == SPR_PREC_BF16) return launch_lw3_mfma(c, q0_first, q0_count, sigma, prec, d_lw);
    if (c.nb <= 2) return launch_lw3_small(c, q0_first, q0_count, sigma, prec, d_lw);
    if (prec == SPR_PREC_FP32) return launch_lw3_mfma(c, q0_first, q0_count, sigma, prec, d_lw);
    return launch_lw3_generic(c, c.irr_grid + q0_first, q0_count, sigma, prec, d_lw, nullptr);
}


}  // namespace spr
