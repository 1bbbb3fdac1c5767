// Internal declarations of libspring_gpu.so (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/spring_gpu.h"

namespace spr {

void set_error(const std::string &msg);

#define SPR_HIP(call)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::spr::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return SPR_E_HIP;                                                             \
        }                                                                                 \
    } while (0)

// First member of every item-kernel argument struct: the launcher sets these.
struct ItemCommon {
    int *counter;       // persistent mode: atomic work counter (nullptr: one workgroup per item)
    int nitems;
    double *partial;    // (nitems, ndof) per-item partial linewidths
};

// One item-kernel launch as the launcher (k_launch.hip) sees it.
struct ItemLaunch {
    const void *func;
    void *args;             // points to the kernel's argument struct (ItemCommon first)
    size_t args_size;
    int block;
    size_t dyn;
    int nitems;
    int per_cu;             // resident workgroups per compute unit
    std::string name;
    bool direct = false;    // one workgroup per item, no graph, no tuning, no reduction
    void *extra[4] = {nullptr, nullptr, nullptr, nullptr};   // further kernel arguments
    int nextra = 0;
    int upi = 1;            // work units per item (nitems counts units); partial rows per item
    std::string shape;      // tile-shape label of this variant (tuning table, last_shape)
};

// Caller-supplied work items (spr_lw3_pairs): partial[item] receives the item's sum.
struct ItemSet {
    const int *item_q0, *item_off, *item_cnt;
    const int2 *pairs;
    int nitems;
    double *partial;
    // per-pair weight in place of the kernels' 1 (q1 == q2) or 2 (NULL: those)
    const double *pair_w = nullptr;
    // items hold up to 64 pairs of one q0 as the context's own lists do (the kernels that take
    // whole items are preferred), rather than one pair each
    bool dense = false;
};

// Anharmonic IFC re-laid out for the phase sum.  The sum over (N1, N2[, N3]) is grouped by
// the basis atoms of the sites: for every (mu, nu, eta[, rho]) "block" the list of
// contributing site tuples, each with its 27 (81) FC values premultiplied by
// 1/sqrt(m_mu m_nu m_eta [m_rho]).  Only tuples with a nonzero FC block are kept.
struct DevIFC {
    int order = 0;
    int nblk = 0;            // Nb^(order)
    int nterm = 0;           // number of (mu, N1, N2[, N3]) tuples kept
    int nsite = 0;           // sum_mu atm_num[mu]
    int *blk_start = nullptr;  // (nblk + 1) offsets into the tuple list
    int *tup_site = nullptr;   // (nterm, order-1) global site index of N1, N2[, N3]
    double *fc64 = nullptr;    // (nterm, 3^order) mass-scaled FC values, (a,b,c[,d]) order
    float *fc32 = nullptr;
    int4 *site = nullptr;      // (nsite): phase coefficients (c1,c2,c3) and basis atom
    std::vector<int4> h_site_cell;   // (nsite): the site's lattice cell (n1,n2,n3) and basis atom
    std::vector<int> h_blk_start;
    // the same list with every block padded to a multiple of kClusterBatch tuples (zero FC
    // values on a valid site), for kernels that load the tuples of a batch together
    int *blk_start_p = nullptr;
    int *tup_site_p = nullptr;
    double *fc64_p = nullptr;
    float *fc32_p = nullptr;
    // phase-sum schedule of the s2-per-wave kernel (k_lw3_mx.hip), built on first use: per basis
    // atom the outputs of its non-empty blocks sorted by cluster count, then the outputs to clear
    int *ft_sched = nullptr;
    int ft_nmax = 0;
    bool ft_tried = false, ft_ok = false;
    // phase-sum task lists of the FP64 per-atom kernel (k_lw3_mm.hip), built on first use
    int *ft2 = nullptr;
    bool ft2_tried = false, ft2_ok = false;
};

constexpr int kClusterBatch = 4;

// Iterative BTE state (time-reversal irreducible set, F map, W of every representative q)
struct BteState {
    bool ready = false;
    int nirr = 0;
    std::vector<int> h_irr, h_deg;        // (nirr) grid index, number of grid points it stands for
    std::vector<int2> h_map;              // (nq) irreducible index and sign of F
    int4 *d_coord = nullptr;
    int2 *d_map = nullptr;
    int *d_irr = nullptr, *d_tetra = nullptr;
    double *d_W = nullptr;                // (nirr, nq, nd, nd, nd)
    double *d_F = nullptr, *d_out = nullptr;
    // collision matrix dF = A F over rows / columns (irreducible q, s), built once
    bool a_ready = false;
    double *d_A = nullptr, *d_itau = nullptr, *d_F0 = nullptr, *d_Fb = nullptr;
    unsigned char *d_div = nullptr;
    double build_ms = 0.0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nb = 0, nd = 0, nq = 0, nq0 = 0;
    int mesh[3] = {0, 0, 0};
    int L = 0;                     // phase table length: lcm of the mesh dimensions
    double vol = 0.0;
    DevIFC fc3, fc4;
    double2 *phase64 = nullptr;    // (L) exp(2 pi i m / L)
    float2 *phase32 = nullptr;
    double2 *evec = nullptr;       // (nq, nd, nd): evec[q][s][row] = Evec(row, s, q)
    float2 *evec32 = nullptr;
    double *omega = nullptr, *inv_omega = nullptr, *nbe = nullptr, *vel = nullptr;
    int *irr_grid = nullptr;       // (nq0)
    std::vector<int> h_irr_grid, h_pos, h_mult;
    std::vector<double> h_rot;
    int nsym = 0;
    std::vector<double> h_omega, h_nbe, h_vel;
    // scratch
    double *partial = nullptr;
    size_t partial_cap = 0;
    double *out = nullptr;
    size_t out_cap = 0;
    int *counter = nullptr;
    // pair-permutation reduced (q1, q2) lists of every irreducible q0, in 64-pair batches
    bool pairs_built = false;
    long n_pairs = 0;
    int n_items = 0;
    int2 *d_pairs = nullptr;
    int *d_item_q0 = nullptr, *d_item_off = nullptr, *d_item_cnt = nullptr, *d_item_first = nullptr;
    std::vector<int> h_item_first;
    int ncu = 256;                 // compute units of the device
    int last_mode = -1;            // launch mode of the last item launch (0: one WG per item)
    int last_upi = 1;              // work units per item of the last item launch
    std::string last_shape;        // tile shape of the last item launch
    std::string last_name;         // kernel variant of the last item launch
    std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
    double *ipartial = nullptr;    // (n_items, nd) item partials (stable: graphs keep the pointer)
    double *wout = nullptr;        // W parity output (also W of one q0 for the tetrahedron path)
    int *tetra = nullptr, *partner = nullptr;   // (6 nq, 4) tetrahedra; q2 partner of every q1
    std::vector<int> h_tetra;
    double *tout = nullptr;        // tetrahedron outputs: lw then P3_q0, (2, nq0, nd)
    double *w4buf = nullptr;       // W4 rows of one q1 chunk (four-phonon tetrahedron path)
    int *kpack = nullptr;          // grid coordinates of every q, k0 | k1 << 10 | k2 << 20
    double *wsink = nullptr;       // (q0 count, nd) reduction target of the small-cell W-output launch
    size_t wsink_cap = 0, partner_cap = 0;
    size_t w4buf_cap = 0;
    // kappa buffers, allocated on first use: per-mode terms, lw / mode sums, index tables
    double *k_terms = nullptr, *k_io = nullptr, *k_rot = nullptr;
    int *k_pos_full = nullptr, *k_pos_irr = nullptr, *k_mult = nullptr;
    const ItemSet *override_items = nullptr;
    const int *w_pair_q0 = nullptr, *w_pair_q1 = nullptr;   // W parity path over a pair list
    int w_npairs = 0;
    size_t wout_cap = 0;
    BteState bte;
    // branch selection rules (PolarizationStrtEnd, Helper_m.f90:310-353): 0-based inclusive
    // (lo, hi) of s0, s1, s2 (three-phonon) and s0..s3 (four-phonon); every branch by default
    int pol3[6] = {0, 0, 0, 0, 0, 0};
    int pol4[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool sel3 = false, sel4 = false;   // a rule other than "every branch" is set
    double *d_sel3 = nullptr;          // (nd) s0 mask, then (nd, nd) pair mask [s1][s2] (see spr_api.hip)
    double *d_sel4 = nullptr;          // (4, nd) masks of s0, s1, s2, s3
    double *off_buf = nullptr;         // inputs of the off-mesh (LineWidth) path
    size_t off_cap = 0;
    // timing of the dominant kernel
    int launches = 0;
    double kernel_ms = 0.0;
    // run_items leaves its events unread (no host wait) while defer_timing is set;
    // collect_timing reads them after the caller's own synchronisation
    bool defer_timing = false, timing_pending = false;
};

int ensure(void **p, size_t *cap, size_t bytes);

// kernels launched from the API layer
// irreducible q0 [q0_first, q0_first + q0_count); d_w_out (FP64 W of every q1) needs count 1
int launch_lw3(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw,
               double *d_w_out);
int launch_lw4(Ctx &c, int q0_first, int q0_count, double sigma, int prec, double *d_lw);
int launch_lw4_mm(Ctx &c, int q0_first, int q0_count, double sigma, double *d_lw);
// W of every grid q1 for the grid points d_q0[0..nq0) (d_w, nq0 = 1) or their Gaussian linewidths
int launch_lw3_generic(Ctx &c, const int *d_q0, int nq0, double sigma, int prec, double *d_lw, double *d_w);
int bte_setup(Ctx &c);
int bte_sums(Ctx &c, int mode);
int bte_build(Ctx &c);
int bte_gemv(Ctx &c, const double *d_F, double *d_Fnew);
double lw4_factor(const Ctx &c);
// d_lw[q][s] = sum over chunks in order of partial[(q, chunk)][s]
int launch_reduce_chunks(Ctx &c, const double *partial, int q0_count, int nchunks, double *d_lw);
int ensure_pairs(Ctx &c);
int run_items(Ctx &c, ItemLaunch &L, int q0_first, int q0_count, int it0, double *d_lw, int mode = -1);
// the same over alternative tile shapes of one kernel (same items, each its own units per item):
// the launcher times every (shape, launch mode) on a sample of the items, keeps the fastest in the
// tuning table and runs it
int run_items_shapes(Ctx &c, ItemLaunch *Ls, int n, int q0_first, int q0_count, int it0, double *d_lw);
void drop_graphs(Ctx &c);
int collect_timing(Ctx &c);
int launch_reduce_items(Ctx &c, int q0_first, int q0_count, int it0, int upi, double *d_lw);
constexpr int kMaxUnitsPerItem = 4;   // partial rows reserved per item
int launch_kappa(Ctx &c, const double *d_lw_all, double T, int symm, double *d_mode_k);
void kappa_mode_range(const Ctx &c, int &lo, int &hi);
// tetrahedron linewidths of irreducible q0 [q_first, q_first + q_count): d_lw[k][nd], d_p3[k][nd]
int launch_lw3_tetra(Ctx &c, int q_first, int q_count, double *d_lw, double *d_p3);
// FP64 W of every grid q1 of irreducible q0 [q_first, q_first + q_count) with the small-cell
// kernel (Nb <= 2): d_w[q0 - q_first][q1][s0][s1][s2]
int launch_lw3_small_w(Ctx &c, int q_first, int q_count, double *d_w);
int launch_lw3_mm_w(Ctx &c, int q_first, int q_count, double *d_w);
int launch_lw3_mm_w_items(Ctx &c, const int *d_q0_grid, int nq0, double *d_w);
int launch_lw3_small_w_items(Ctx &c, const int *d_q0_grid, int nq0, double *d_w);
bool lw3_mm_w_built(int nb);
// device tetrahedra (6 nq, 4) of the context (uploaded on first use)
int ensure_tetra(Ctx &c);
// FP64 W4 of irreducible q0 `qi` for q1 in [q1_lo, q1_lo + nq1) and every q2:
// d_w[q1 - q1_lo][q2][s0][s1][s2][s3]
int launch_w4_rows(Ctx &c, int qi, int q1_lo, int nq1, double *d_w, bool classes = false);
int launch_lw4_tetra(Ctx &c, int qi, double *d_lw);
// device inputs of the off-mesh q0 W (LineWidth): eigenvectors (nd, nd) and 1/omega (nd) of q0,
// eigenvectors (nq, nd, nd) and 1/omega (nq, nd) of q2 per grid q1, e^{-2 pi i x0 . n} per FC3 site,
// and whether q2 exists per q1
struct OffMeshDev {
    const double2 *e0, *e2, *p0;
    const double *iw0, *iw2;
    const int *exists;
};
int launch_w_offmesh(Ctx &c, const OffMeshDev &om, double *d_w);
// tetrahedron sums of one q0 from W of every q1 (k_lw3_tetra): q2 data at omega[partner[q1]],
// nbe[partner[q1]]; every branch; d_lw(nd), d_p3(nd)
int tetra_sums_from_w(Ctx &c, const double *d_w, const int *d_partner, const double *d_omega, const double *d_nbe,
                      const double *d_omega0, double *d_lw, double *d_p3);
// resolves the device address of the no-allocation scratch array of `device` (current device)
int noalloc_setup(int device);

}  // namespace spr

// the opaque handle of include/spring_gpu.h
struct spr_ctx {
    spr::Ctx c;
};
