// Chooses the kernel variant for each entry point.
#include "spr_internal.h"

namespace spr {

int launch_lw3_generic(Ctx &c, const int *d_q0, int nq0, double sigma, int prec, double *d_lw,
                       double *d_w);

int launch_lw3(Ctx &c, const int *d_q0_grid, int nq0, double sigma, int prec, double *d_lw,
               double *d_w_out) {
    return launch_lw3_generic(c, d_q0_grid, nq0, sigma, prec, d_lw, d_w_out);
}

int launch_lw4(Ctx &c, const int *, int, double, int, double *) {
    (void)c;
    set_error("four-phonon kernel not built yet");
    return SPR_E_UNSUPPORTED;
}

}  // namespace spr
