// C ABI of libspring_gpu.so: context set-up (uploads), entry points, error reporting.
// Declarations and reference citations: include/spring_gpu.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "spr_internal.h"

namespace spr {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int ensure(void **p, size_t *cap, size_t bytes) {
    if (*cap >= bytes && *p) return SPR_OK;
    if (*p) {
        SPR_HIP(hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    if (hipMalloc(p, bytes) != hipSuccess) {
        set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        return SPR_E_OOM;
    }
    *cap = bytes;
    return SPR_OK;
}

template <typename T>
static int upload(T **dst, const T *src, size_t n) {
    if (n == 0) {
        *dst = nullptr;
        return SPR_OK;
    }
    if (hipMalloc((void **)dst, n * sizeof(T)) != hipSuccess) {
        set_error("hipMalloc failed in upload");
        return SPR_E_OOM;
    }
    SPR_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return SPR_OK;
}

static long lcm(long a, long b) { return a / std::gcd(a, b) * b; }

// Build the block-grouped tuple list of an IFC (see DevIFC).
static int build_ifc(Ctx &c, const spr_ifc *f, const double *osm, DevIFC &d) {
    const int nb = c.nb, order = f->order, natm = f->natm;
    if ((order != 3 && order != 4) || natm <= 0) {
        set_error("spr_ifc: order must be 3 or 4 and natm > 0");
        return SPR_E_ARG;
    }
    const int ncomp = order == 3 ? 27 : 81;
    const int nsub = order - 1;
    std::vector<int> site_off(nb + 1, 0);
    for (int mu = 0; mu < nb; ++mu) {
        if (f->atm_num[mu] < 1 || f->atm_num[mu] > natm) {
            set_error("spr_ifc: atm_num out of range");
            return SPR_E_ARG;
        }
        site_off[mu + 1] = site_off[mu] + f->atm_num[mu];
    }
    const int nsite = site_off[nb];
    std::vector<int4> site(nsite);
    for (int mu = 0; mu < nb; ++mu)
        for (int n = 0; n < f->atm_num[mu]; ++n) {
            const int *ix = f->atm_indx + ((size_t)mu * natm + n) * 4;
            if (ix[3] < 0 || ix[3] >= nb) {
                set_error("spr_ifc: basis index in atm_indx out of range (expects 0-based)");
                return SPR_E_ARG;
            }
            int cc[3];
            for (int k = 0; k < 3; ++k) {
                long m = ((long)ix[k] % c.mesh[k] + c.mesh[k]) % c.mesh[k];
                cc[k] = (int)((m * (c.L / c.mesh[k])) % c.L);
            }
            site[site_off[mu] + n] = make_int4(cc[0], cc[1], cc[2], ix[3]);
            d.h_site_cell.push_back(make_int4(ix[0], ix[1], ix[2], ix[3]));
        }
    const int nblk = order == 3 ? nb * nb * nb : nb * nb * nb * nb;
    std::vector<std::vector<int>> tup(nblk);
    std::vector<std::vector<double>> val(nblk);
    auto basis = [&](int mu, int n) { return f->atm_indx[((size_t)mu * natm + n) * 4 + 3]; };
    for (int mu = 0; mu < nb; ++mu) {
        const int na = f->atm_num[mu];
        const int n3max = order == 4 ? na : 1;
        for (int n1 = 0; n1 < na; ++n1)
            for (int n2 = 0; n2 < na; ++n2)
                for (int n3 = 0; n3 < n3max; ++n3) {
                    const double *src = order == 3
                        ? f->fc + (((size_t)mu * natm + n1) * natm + n2) * 27
                        : f->fc + ((((size_t)mu * natm + n1) * natm + n2) * natm + n3) * 81;
                    bool nz = false;
                    for (int i = 0; i < ncomp; ++i) nz |= src[i] != 0.0;
                    if (!nz) continue;
                    const int nu = basis(mu, n1), eta = basis(mu, n2);
                    double m = osm[mu] * osm[nu] * osm[eta];
                    int b;
                    if (order == 3) {
                        b = (mu * nb + nu) * nb + eta;
                    } else {
                        const int rho = basis(mu, n3);
                        m *= osm[rho];
                        b = ((mu * nb + nu) * nb + eta) * nb + rho;
                    }
                    tup[b].push_back(site_off[mu] + n1);
                    tup[b].push_back(site_off[mu] + n2);
                    if (order == 4) tup[b].push_back(site_off[mu] + n3);
                    for (int i = 0; i < ncomp; ++i) val[b].push_back(src[i] * m);
                }
    }
    std::vector<int> start(nblk + 1, 0), all_tup;
    std::vector<double> all_val;
    for (int b = 0; b < nblk; ++b) {
        start[b + 1] = start[b] + (int)(tup[b].size() / nsub);
        all_tup.insert(all_tup.end(), tup[b].begin(), tup[b].end());
        all_val.insert(all_val.end(), val[b].begin(), val[b].end());
    }
    d.order = order;
    d.nblk = nblk;
    d.nterm = start[nblk];
    d.nsite = nsite;
    d.h_blk_start = start;
    std::vector<float> v32(all_val.begin(), all_val.end());
    int rc;
    if ((rc = upload(&d.blk_start, start.data(), start.size()))) return rc;
    // tup_site and fc32 carry 4 zero elements past their end: kernels that stage a range with
    // 16-byte direct-to-LDS loads read the aligned superset of it
    const size_t ntup = all_tup.size(), n32 = v32.size();
    all_tup.resize(ntup + 4, 0);
    v32.resize(n32 + 4, 0.f);
    if ((rc = upload(&d.tup_site, all_tup.data(), all_tup.size()))) return rc;
    if ((rc = upload(&d.fc64, all_val.data(), all_val.size()))) return rc;
    if ((rc = upload(&d.fc32, v32.data(), v32.size()))) return rc;
    all_tup.resize(ntup);
    v32.resize(n32);
    if ((rc = upload(&d.site, site.data(), site.size()))) return rc;
    // padded copy: block b of leading atom mu gets zero-valued tuples on site site_off[mu]
    std::vector<int> pstart(nblk + 1, 0), ptup;
    std::vector<double> pval;
    const int per_mu = nblk / nb;
    for (int b = 0; b < nblk; ++b) {
        const int n = start[b + 1] - start[b];
        const int np = (n + kClusterBatch - 1) / kClusterBatch * kClusterBatch;
        pstart[b + 1] = pstart[b] + np;
        ptup.insert(ptup.end(), all_tup.begin() + (size_t)start[b] * nsub, all_tup.begin() + (size_t)start[b + 1] * nsub);
        pval.insert(pval.end(), all_val.begin() + (size_t)start[b] * ncomp, all_val.begin() + (size_t)start[b + 1] * ncomp);
        ptup.insert(ptup.end(), (size_t)(np - n) * nsub, site_off[b / per_mu]);
        pval.insert(pval.end(), (size_t)(np - n) * ncomp, 0.0);
    }
    std::vector<float> p32(pval.begin(), pval.end());
    if ((rc = upload(&d.blk_start_p, pstart.data(), pstart.size()))) return rc;
    if ((rc = upload(&d.tup_site_p, ptup.data(), ptup.size()))) return rc;
    if ((rc = upload(&d.fc64_p, pval.data(), pval.size()))) return rc;
    if ((rc = upload(&d.fc32_p, p32.data(), p32.size()))) return rc;
    return SPR_OK;
}

static void free_ifc(DevIFC &d) {
    (void)hipFree(d.blk_start);
    (void)hipFree(d.tup_site);
    (void)hipFree(d.fc64);
    (void)hipFree(d.fc32);
    (void)hipFree(d.site);
    (void)hipFree(d.blk_start_p);
    (void)hipFree(d.tup_site_p);
    (void)hipFree(d.fc64_p);
    (void)hipFree(d.fc32_p);
    (void)hipFree(d.ft_sched);
    (void)hipFree(d.ft2);
    d = DevIFC();
}

}  // namespace spr

using namespace spr;

extern "C" const char *spr_last_error(void) { return spr::g_err.c_str(); }

extern "C" int spr_ctx_create(const spr_crystal *cr, const spr_ifc *fc3, const spr_ifc *fc4,
                              const spr_mesh *mesh, const spr_phonons *ph, int device,
                              spr_ctx **out) {
    if (!out || !cr || !mesh || !ph || !fc3) {
        set_error("spr_ctx_create: NULL argument");
        return SPR_E_ARG;
    }
    *out = nullptr;
    if (cr->nb < 1 || ph->ndof != 3 * cr->nb || ph->nq != mesh->nq ||
        mesh->nq != mesh->mesh[0] * mesh->mesh[1] * mesh->mesh[2] || mesh->nq0 < 1) {
        set_error("spr_ctx_create: inconsistent sizes (ndof != 3 nb, or nq != prod(mesh))");
        return SPR_E_ARG;
    }
    int ndev = 0;
    SPR_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) {
        set_error("spr_ctx_create: no HIP device " + std::to_string(device));
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(device));
    {
        const int rc0 = noalloc_setup(device);
        if (rc0) return rc0;
    }
    auto *h = new spr_ctx();
    Ctx &c = h->c;
    c.device = device;
    c.nb = cr->nb;
    c.nd = 3 * cr->nb;
    c.nq = mesh->nq;
    c.nq0 = mesh->nq0;
    c.vol = cr->vol;
    for (int k = 0; k < 3; ++k) c.mesh[k] = mesh->mesh[k];
    c.L = (int)lcm(lcm(c.mesh[0], c.mesh[1]), c.mesh[2]);
    for (int k = 0; k < 3; ++k) c.pol3[2 * k + 1] = c.nd - 1;
    for (int k = 0; k < 4; ++k) c.pol4[2 * k + 1] = c.nd - 1;
    int rc = SPR_OK;
    auto fail = [&](int r) {
        spr_ctx_destroy(h);
        return r;
    };
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c.ev0) != hipSuccess || hipEventCreate(&c.ev1) != hipSuccess) {
        set_error("spr_ctx_create: stream/event creation failed");
        return fail(SPR_E_HIP);
    }
    std::vector<double> osm(c.nb);
    for (int i = 0; i < c.nb; ++i) {
        if (!(cr->mass[i] > 0.0)) {
            set_error("spr_ctx_create: masses must be positive");
            return fail(SPR_E_ARG);
        }
        osm[i] = 1.0 / std::sqrt(cr->mass[i]);
    }
    if ((rc = build_ifc(c, fc3, osm.data(), c.fc3))) return fail(rc);
    if (fc4 && (rc = build_ifc(c, fc4, osm.data(), c.fc4))) return fail(rc);
    // phase table exp(2 pi i m / L)
    std::vector<double2> tab(c.L);
    std::vector<float2> tab32(c.L);
    for (int m = 0; m < c.L; ++m) {
        const double x = 2.0 * M_PI * (double)m / (double)c.L;
        tab[m] = make_double2(std::cos(x), std::sin(x));
        tab32[m] = make_float2((float)tab[m].x, (float)tab[m].y);
    }
    if ((rc = upload(&c.phase64, tab.data(), tab.size()))) return fail(rc);
    if ((rc = upload(&c.phase32, tab32.data(), tab32.size()))) return fail(rc);
    const size_t ne = (size_t)c.nq * c.nd * c.nd;
    if ((rc = upload(&c.evec, reinterpret_cast<const double2 *>(ph->evec), ne))) return fail(rc);
    {
        std::vector<float2> e32(ne);
        for (size_t i = 0; i < ne; ++i) e32[i] = make_float2((float)ph->evec[2 * i], (float)ph->evec[2 * i + 1]);
        if ((rc = upload(&c.evec32, e32.data(), ne))) return fail(rc);
    }
    const size_t nm = (size_t)c.nq * c.nd;
    if ((rc = upload(&c.omega, ph->omega, nm))) return fail(rc);
    if ((rc = upload(&c.inv_omega, ph->inv_omega, nm))) return fail(rc);
    if ((rc = upload(&c.nbe, ph->nbe, nm))) return fail(rc);
    c.h_omega.assign(ph->omega, ph->omega + nm);
    c.h_nbe.assign(ph->nbe, ph->nbe + nm);
    if (ph->grp_vel) {
        if ((rc = upload(&c.vel, ph->grp_vel, nm * 3))) return fail(rc);
        c.h_vel.assign(ph->grp_vel, ph->grp_vel + nm * 3);
    }
    for (int i = 0; i < c.nq0; ++i)
        if (mesh->irr_grid[i] < 0 || mesh->irr_grid[i] >= c.nq) {
            set_error("spr_ctx_create: irr_grid out of range");
            return fail(SPR_E_ARG);
        }
    c.h_irr_grid.assign(mesh->irr_grid, mesh->irr_grid + c.nq0);
    if (mesh->pos) c.h_pos.assign(mesh->pos, mesh->pos + c.nq);
    if (mesh->multiplicity) c.h_mult.assign(mesh->multiplicity, mesh->multiplicity + c.nq0);
    c.nsym = mesh->nsym;
    if (mesh->nsym > 0 && mesh->rot) c.h_rot.assign(mesh->rot, mesh->rot + 9 * mesh->nsym);
    if (mesh->tetra) {
        for (size_t i = 0; i < (size_t)24 * c.nq; ++i)
            if (mesh->tetra[i] < 0 || mesh->tetra[i] >= c.nq) {
                set_error("spr_ctx_create: tetra index out of range");
                return fail(SPR_E_ARG);
            }
        c.h_tetra.assign(mesh->tetra, mesh->tetra + (size_t)24 * c.nq);
    }
    if ((rc = upload(&c.irr_grid, mesh->irr_grid, c.nq0))) return fail(rc);
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
            c.ncu = prop.multiProcessorCount;
    }
    if (hipMalloc((void **)&c.out, (size_t)c.nq0 * c.nd * sizeof(double)) != hipSuccess) {
        set_error("hipMalloc output failed");
        return fail(SPR_E_OOM);
    }
    c.out_cap = (size_t)c.nq0 * c.nd * sizeof(double);
    if (hipMalloc((void **)&c.counter, 64 * sizeof(int)) != hipSuccess) {
        set_error("hipMalloc counter failed");
        return fail(SPR_E_OOM);
    }
    *out = h;
    return SPR_OK;
}

extern "C" int spr_ctx_destroy(spr_ctx *h) {
    if (!h) return SPR_OK;
    Ctx &c = h->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    drop_graphs(c);
    free_ifc(c.fc3);
    free_ifc(c.fc4);
    for (void *p : {(void *)c.phase64, (void *)c.phase32, (void *)c.evec, (void *)c.evec32,
                    (void *)c.omega, (void *)c.inv_omega, (void *)c.nbe, (void *)c.vel,
                    (void *)c.irr_grid, (void *)c.partial, (void *)c.out, (void *)c.counter,
                    (void *)c.d_pairs, (void *)c.d_item_q0, (void *)c.d_item_off, (void *)c.d_item_cnt,
                    (void *)c.d_item_first,
                    (void *)c.ipartial, (void *)c.wout, (void *)c.tetra, (void *)c.partner, (void *)c.tout, (void *)c.w4buf, (void *)c.kpack, (void *)c.wsink,
                    (void *)c.k_terms, (void *)c.k_io, (void *)c.k_rot, (void *)c.k_pos_full, (void *)c.k_pos_irr,
                    (void *)c.k_mult, (void *)c.bte.d_coord, (void *)c.bte.d_map, (void *)c.bte.d_irr,
                    (void *)c.bte.d_tetra, (void *)c.bte.d_W, (void *)c.bte.d_F, (void *)c.bte.d_out,
                    (void *)c.bte.d_A, (void *)c.bte.d_itau, (void *)c.bte.d_F0, (void *)c.bte.d_Fb,
                    (void *)c.bte.d_div, (void *)c.d_sel3, (void *)c.d_sel4, (void *)c.off_buf})
        if (p) (void)hipFree(p);
    if (c.ev0) (void)hipEventDestroy(c.ev0);
    if (c.ev1) (void)hipEventDestroy(c.ev1);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete h;
    return SPR_OK;
}

static int check_i0(Ctx &c, int i0) {
    if (i0 < 0 || i0 >= c.nq0) {
        set_error("irreducible index i0 out of range");
        return SPR_E_ARG;
    }
    return SPR_OK;
}

static int check_prec(int prec) {
    if (prec != SPR_PREC_FP64 && prec != SPR_PREC_FP32 && prec != SPR_PREC_BF16) {
        set_error("unknown precision mode");
        return SPR_E_ARG;
    }
    return SPR_OK;
}

extern "C" int spr_v3_w(spr_ctx *h, int i0, double *W_out) {
    if (!h || !W_out) {
        set_error("spr_v3_w: NULL argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    int rc;
    if ((rc = check_i0(c, i0))) return rc;
    SPR_HIP(hipSetDevice(c.device));
    const size_t n = (size_t)c.nd * c.nd * c.nd * c.nq;
    if ((rc = ensure((void **)&c.wout, &c.wout_cap, n * sizeof(double)))) return rc;
    if ((rc = launch_lw3(c, i0, 1, 1.0, SPR_PREC_FP64, nullptr, c.wout))) return rc;
    SPR_HIP(hipMemcpyAsync(W_out, c.wout, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SPR_HIP(hipStreamSynchronize(c.stream));
    return SPR_OK;
}

static int lw_common(spr_ctx *h, int first, int count, double sigma, int prec, double *lw, int order) {
    if (!h || !lw) {
        set_error("linewidth: NULL argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    int rc;
    if ((rc = check_prec(prec))) return rc;
    if (!(sigma > 0.0)) {
        set_error("sigma must be positive");
        return SPR_E_ARG;
    }
    if (order == 4 && c.fc4.nterm == 0 && c.fc4.order == 0) {
        set_error("no FC4 was given to spr_ctx_create");
        return SPR_E_STATE;
    }
    SPR_HIP(hipSetDevice(c.device));
    // c.out (nq0 x Ndof) is allocated once; captured graphs keep pointers into it
    const size_t n = (size_t)count * c.nd;
    double *d_lw = c.out + (size_t)first * c.nd;
    rc = order == 3 ? launch_lw3(c, first, count, sigma, prec, d_lw, nullptr)
                    : launch_lw4(c, first, count, sigma, prec, d_lw);
    if (rc) return rc;
    SPR_HIP(hipMemcpyAsync(lw, d_lw, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    SPR_HIP(hipStreamSynchronize(c.stream));
    return SPR_OK;
}

extern "C" int spr_lw3(spr_ctx *h, int i0, double sigma, int prec, double *lw) {
    if (!h) return SPR_E_ARG;
    int rc;
    if ((rc = check_i0(h->c, i0))) return rc;
    return lw_common(h, i0, 1, sigma, prec, lw, 3);
}

extern "C" int spr_lw3_all(spr_ctx *h, double sigma, int prec, double *lw_all) {
    if (!h) return SPR_E_ARG;
    return lw_common(h, 0, h->c.nq0, sigma, prec, lw_all, 3);
}

extern "C" int spr_lw3_slice(spr_ctx *h, int q0_off, int q0_cnt, double sigma, int prec, double *lw) {
    if (!h) return SPR_E_ARG;
    if (q0_off < 0 || q0_cnt < 1 || q0_off + q0_cnt > h->c.nq0) {
        set_error("spr_lw3_slice: q0 slice out of range");
        return SPR_E_ARG;
    }
    return lw_common(h, q0_off, q0_cnt, sigma, prec, lw, 3);
}

extern "C" int spr_lw4(spr_ctx *h, int i0, double sigma, int prec, double *lw) {
    if (!h) return SPR_E_ARG;
    int rc;
    if ((rc = check_i0(h->c, i0))) return rc;
    return lw_common(h, i0, 1, sigma, prec, lw, 4);
}

extern "C" int spr_lw4_all(spr_ctx *h, double sigma, int prec, double *lw_all) {
    if (!h) return SPR_E_ARG;
    return lw_common(h, 0, h->c.nq0, sigma, prec, lw_all, 4);
}

extern "C" int spr_set_branch_rules(spr_ctx *h, int order, const int *pol) {
    if (!h || !pol || (order != 3 && order != 4)) {
        set_error("spr_set_branch_rules: order must be 3 or 4 and pol non-NULL");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    const int nd = c.nd;
    int r[8];
    bool all = true;
    for (int k = 0; k < order; ++k) {
        const int lo = pol[2 * k], hi = pol[2 * k + 1];
        if (lo < 1 || hi > nd || lo > hi) {
            set_error("spr_set_branch_rules: range of slot " + std::to_string(k + 1) + " outside 1.." + std::to_string(nd));
            return SPR_E_ARG;
        }
        r[2 * k] = lo - 1;
        r[2 * k + 1] = hi - 1;
        all = all && lo == 1 && hi == nd;
    }
    SPR_HIP(hipSetDevice(c.device));
    auto in = [&](int k, int s) { return s >= r[2 * k] && s <= r[2 * k + 1] ? 1.0 : 0.0; };
    if (order == 3) {
        // s0 mask, then the pair mask of the pair-reduced Gaussian sums: an unordered pair {q1, q2}
        // stands for both orderings, so branch (s1, s2) of the stored ordering takes
        // (m1(s1) m2(s2) + m1(s2) m2(s1)) / 2 (the reduced sum weights q1 != q2 by 2)
        std::vector<double> m((size_t)nd + (size_t)nd * nd);
        for (int s = 0; s < nd; ++s) m[s] = in(0, s);
        for (int a = 0; a < nd; ++a)
            for (int b = 0; b < nd; ++b) m[nd + (size_t)a * nd + b] = 0.5 * (in(1, a) * in(2, b) + in(1, b) * in(2, a));
        if (!c.d_sel3) SPR_HIP(hipMalloc((void **)&c.d_sel3, m.size() * sizeof(double)));
        SPR_HIP(hipMemcpy(c.d_sel3, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
        std::copy(r, r + 6, c.pol3);
        c.sel3 = !all;
    } else {
        std::vector<double> m((size_t)4 * nd);
        for (int k = 0; k < 4; ++k)
            for (int s = 0; s < nd; ++s) m[(size_t)k * nd + s] = in(k, s);
        if (!c.d_sel4) SPR_HIP(hipMalloc((void **)&c.d_sel4, m.size() * sizeof(double)));
        SPR_HIP(hipMemcpy(c.d_sel4, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice));
        std::copy(r, r + 8, c.pol4);
        c.sel4 = !all;
    }
    return SPR_OK;
}

extern "C" int spr_kernel_time(spr_ctx *h, int *launches, double *ms) {
    if (!h) return SPR_E_ARG;
    if (launches) *launches = h->c.launches;
    if (ms) *ms = h->c.kernel_ms;
    return SPR_OK;
}

extern "C" int spr_last_launch(spr_ctx *h, char *shape, int cap, int *mode, int *units_per_item) {
    if (!h) return SPR_E_ARG;
    const Ctx &c = h->c;
    if (shape && cap > 0) {
        const std::string t = c.last_name + (c.last_shape.empty() ? "" : ":" + c.last_shape);
        const size_t n = std::min(t.size(), (size_t)cap - 1);
        std::memcpy(shape, t.data(), n);
        shape[n] = 0;
    }
    if (mode) *mode = c.last_mode;
    if (units_per_item) *units_per_item = c.last_upi;
    return SPR_OK;
}

extern "C" int spr_kernel_time_reset(spr_ctx *h) {
    if (!h) return SPR_E_ARG;
    h->c.launches = 0;
    h->c.kernel_ms = 0.0;
    return SPR_OK;
}

extern "C" int spr_kappa_rta(spr_ctx *h, const double *lw_all, double T, int symm, double *kappa,
                             double *mode_k) {
    if (!h || !lw_all || !kappa) {
        set_error("spr_kappa_rta: NULL argument");
        return SPR_E_ARG;
    }
    if (!(T > 0.0)) {
        set_error("spr_kappa_rta: T must be positive");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    SPR_HIP(hipSetDevice(c.device));
    const size_t nl = (size_t)c.nq0 * c.nd, nm = (size_t)c.nd * 9;
    if (!c.k_io) SPR_HIP(hipMalloc((void **)&c.k_io, (nl + nm) * sizeof(double)));
    double *d = c.k_io;
    int rc = SPR_OK;
    std::vector<double> modes(nm);
    if (hipMemcpyAsync(d, lw_all, nl * sizeof(double), hipMemcpyHostToDevice, c.stream) != hipSuccess) {
        set_error("spr_kappa_rta: upload failed");
        rc = SPR_E_HIP;
    }
    if (!rc) rc = launch_kappa(c, d, T, symm, d + nl);
    if (!rc && (hipMemcpyAsync(modes.data(), d + nl, nm * sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
                hipStreamSynchronize(c.stream) != hipSuccess)) {
        set_error("spr_kappa_rta: download failed");
        rc = SPR_E_HIP;
    }
    if (rc) return rc;
    for (int k = 0; k < 9; ++k) {
        double s = 0.0;
        for (int m = 0; m < c.nd; ++m) s += modes[(size_t)m * 9 + k];
        kappa[k] = s;
    }
    if (mode_k) std::memcpy(mode_k, modes.data(), nm * sizeof(double));
    return SPR_OK;
}

extern "C" int spr_kappa_rta_spectral(spr_ctx *h, const double *lw_all, double T, int symm, double *kappa,
                                      double *mode_k, double *spectral_k) {
    if (!h || !spectral_k) {
        set_error("spr_kappa_rta_spectral: NULL argument");
        return SPR_E_ARG;
    }
    int rc;
    if ((rc = spr_kappa_rta(h, lw_all, T, symm, kappa, mode_k))) return rc;
    Ctx &c = h->c;
    const int nd = c.nd, nt = symm ? c.nq0 : c.nq;
    // the per-(q, s) terms spr_kappa_rta left in k_terms (without the kW_m_k / (T^2 V) factor)
    std::vector<double> terms((size_t)nt * nd * 9);
    SPR_HIP(hipMemcpy(terms.data(), c.k_terms, terms.size() * sizeof(double), hipMemcpyDeviceToHost));
    const double kw = (1.054571817 * 1.054571817) * 0.001 / 1.380649;   // constants.f90:65
    const double f1 = kw / (T * T * c.vol);
    const size_t ncol = (size_t)nd * c.nq0;
    std::fill(spectral_k, spectral_k + 11 * ncol, 0.0);
    // column (qi0, s) = qi0 + s nq0 (indx2num over (Nq0, Ndof)); grid q adds to its irreducible
    // point in q order (kappa_noIter's q loop) or the irreducible point holds its own term (symm)
    for (int t = 0; t < nt; ++t) {
        const int qi = symm ? t : c.h_pos[t];
        for (int s = 0; s < nd; ++s) {
            double *col = spectral_k + 11 * ((size_t)s * c.nq0 + qi);
            for (int k = 0; k < 9; ++k) col[2 + k] += terms[((size_t)t * nd + s) * 9 + k] * f1;
        }
    }
    for (int qi = 0; qi < c.nq0; ++qi)
        for (int s = 0; s < nd; ++s) {
            double *col = spectral_k + 11 * ((size_t)s * c.nq0 + qi);
            col[0] = (double)(s + 1);
            col[1] = c.h_omega[(size_t)c.h_irr_grid[qi] * nd + s];
        }
    return SPR_OK;
}

extern "C" int spr_lw3_kappa_all(spr_ctx *h, double sigma, int prec, double T, int symm, double *lw_all,
                                 double *kappa, double *mode_k) {
    if (!h || !lw_all || !kappa) {
        set_error("spr_lw3_kappa_all: NULL argument");
        return SPR_E_ARG;
    }
    if (!(T > 0.0)) {
        set_error("spr_lw3_kappa_all: T must be positive");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    int rc;
    if ((rc = check_prec(prec))) return rc;
    if (!(sigma > 0.0)) {
        set_error("sigma must be positive");
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(c.device));
    const size_t nl = (size_t)c.nq0 * c.nd, nm = (size_t)c.nd * 9;
    if (!c.k_io) SPR_HIP(hipMalloc((void **)&c.k_io, (nl + nm) * sizeof(double)));
    // LW_Allq0 stays on the device between the linewidths and kappa, and the host waits once
    c.defer_timing = true;
    rc = launch_lw3(c, 0, c.nq0, sigma, prec, c.out, nullptr);
    c.defer_timing = false;
    if (!rc) rc = launch_kappa(c, c.out, T, symm, c.k_io + nl);
    std::vector<double> modes(nm);
    if (!rc && (hipMemcpyAsync(lw_all, c.out, nl * sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
                hipMemcpyAsync(modes.data(), c.k_io + nl, nm * sizeof(double), hipMemcpyDeviceToHost, c.stream) !=
                    hipSuccess ||
                hipStreamSynchronize(c.stream) != hipSuccess)) {
        set_error("spr_lw3_kappa_all: download failed");
        rc = SPR_E_HIP;
    }
    if (rc) {
        c.timing_pending = false;
        return rc;
    }
    if ((rc = collect_timing(c))) return rc;
    for (int k = 0; k < 9; ++k) {
        double acc = 0.0;
        for (int m = 0; m < c.nd; ++m) acc += modes[(size_t)m * 9 + k];
        kappa[k] = acc;
    }
    if (mode_k) std::memcpy(mode_k, modes.data(), nm * sizeof(double));
    return SPR_OK;
}

namespace spr {
int launch_lw3_pairs(Ctx &c, int n, const int *q0_index, const int *q1_grid, double sigma, int prec, double *h_out);
int launch_lw3_sampled(Ctx &c, int n, const int *q0_index, const int *q1_grid, const double *weight, double sigma,
                       int prec, double *h_est);
}

extern "C" int spr_lw3_sampled(spr_ctx *h, int n, const int *q0_index, const int *q1_grid, const double *weight,
                               double sigma, int prec, double *est) {
    if (!h || n < 0 || !est || (n > 0 && (!q0_index || !q1_grid || !weight))) {
        set_error("spr_lw3_sampled: bad argument");
        return SPR_E_ARG;
    }
    int rc;
    if ((rc = check_prec(prec))) return rc;
    if (!(sigma > 0.0)) {
        set_error("sigma must be positive");
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(h->c.device));
    return launch_lw3_sampled(h->c, n, q0_index, q1_grid, weight, sigma, prec, est);
}

extern "C" int spr_lw3_pairs(spr_ctx *h, int n, const int *q0_index, const int *q1_grid, double sigma, int prec,
                             double *out) {
    if (!h || n < 0 || (n > 0 && (!q0_index || !q1_grid || !out))) {
        set_error("spr_lw3_pairs: bad argument");
        return SPR_E_ARG;
    }
    int rc;
    if ((rc = check_prec(prec))) return rc;
    if (!(sigma > 0.0)) {
        set_error("sigma must be positive");
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(h->c.device));
    return launch_lw3_pairs(h->c, n, q0_index, q1_grid, sigma, prec, out);
}

extern "C" int spr_v3_w_pairs(spr_ctx *h, int n, const int *q0_index, const int *q1_grid, double *W_out) {
    if (!h || n < 1 || !q0_index || !q1_grid || !W_out) {
        set_error("spr_v3_w_pairs: bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    for (int k = 0; k < n; ++k)
        if (q0_index[k] < 0 || q0_index[k] >= c.nq0 || q1_grid[k] < 0 || q1_grid[k] >= c.nq) {
            set_error("spr_v3_w_pairs: index out of range");
            return SPR_E_ARG;
        }
    SPR_HIP(hipSetDevice(c.device));
    const size_t nw = (size_t)c.nd * c.nd * c.nd * n;
    int rc;
    if ((rc = ensure((void **)&c.wout, &c.wout_cap, nw * sizeof(double)))) return rc;
    int *d_idx = nullptr;
    if (hipMalloc((void **)&d_idx, 2 * (size_t)n * sizeof(int)) != hipSuccess) {
        set_error("spr_v3_w_pairs: hipMalloc failed");
        return SPR_E_OOM;
    }
    rc = SPR_OK;
    if (hipMemcpy(d_idx, q0_index, n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_idx + n, q1_grid, n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
        set_error("spr_v3_w_pairs: upload failed");
        rc = SPR_E_HIP;
    }
    if (!rc) {
        c.w_pair_q0 = d_idx;
        c.w_pair_q1 = d_idx + n;
        c.w_npairs = n;
        rc = launch_lw3(c, 0, 1, 1.0, SPR_PREC_FP64, nullptr, c.wout);
        c.w_pair_q0 = c.w_pair_q1 = nullptr;
        c.w_npairs = 0;
    }
    if (!rc && hipMemcpy(W_out, c.wout, nw * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("spr_v3_w_pairs: download failed");
        rc = SPR_E_HIP;
    }
    (void)hipFree(d_idx);
    return rc;
}

extern "C" int spr_lw3_tetra(spr_ctx *h, int i0, double *lw, double *p3_q0) {
    if (!h || !lw || !p3_q0) {
        set_error("spr_lw3_tetra: bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    if (i0 < 0 || i0 >= c.nq0) {
        set_error("spr_lw3_tetra: q0 index out of range");
        return SPR_E_ARG;
    }
    SPR_HIP(hipSetDevice(c.device));
    int rc;
    if (!c.tout) SPR_HIP(hipMalloc((void **)&c.tout, (size_t)2 * c.nq0 * c.nd * sizeof(double)));
    if ((rc = spr::launch_lw3_tetra(c, i0, 1, c.tout, c.tout + c.nd))) return rc;
    SPR_HIP(hipMemcpy(lw, c.tout, c.nd * sizeof(double), hipMemcpyDeviceToHost));
    SPR_HIP(hipMemcpy(p3_q0, c.tout + c.nd, c.nd * sizeof(double), hipMemcpyDeviceToHost));
    return SPR_OK;
}

static int lw4_tetra_common(spr_ctx *h, int first, int count, double *lw, const char *who) {
    if (!h || !lw) {
        set_error(std::string(who) + ": bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    if (first < 0 || count < 1 || first + count > c.nq0) {
        set_error(std::string(who) + ": q0 index out of range");
        return SPR_E_ARG;
    }
    if (c.fc4.nterm == 0 && c.fc4.order == 0) {
        set_error(std::string(who) + ": no FC4 was given to spr_ctx_create");
        return SPR_E_STATE;
    }
    SPR_HIP(hipSetDevice(c.device));
    int rc;
    if (!c.tout) SPR_HIP(hipMalloc((void **)&c.tout, (size_t)2 * c.nq0 * c.nd * sizeof(double)));
    for (int qi = first; qi < first + count; ++qi)
        if ((rc = spr::launch_lw4_tetra(c, qi, c.tout + (size_t)(qi - first) * c.nd))) return rc;
    SPR_HIP(hipMemcpy(lw, c.tout, (size_t)count * c.nd * sizeof(double), hipMemcpyDeviceToHost));
    return SPR_OK;
}

extern "C" int spr_lw4_tetra(spr_ctx *h, int i0, double *lw) {
    return lw4_tetra_common(h, i0, 1, lw, "spr_lw4_tetra");
}

extern "C" int spr_lw4_tetra_all(spr_ctx *h, double *lw_all) {
    return lw4_tetra_common(h, 0, h ? h->c.nq0 : 0, lw_all, "spr_lw4_tetra_all");
}

extern "C" int spr_lw3_tetra_all(spr_ctx *h, double *lw_all, double *p3_all) {
    if (!h || !lw_all || !p3_all) {
        set_error("spr_lw3_tetra_all: bad argument");
        return SPR_E_ARG;
    }
    Ctx &c = h->c;
    SPR_HIP(hipSetDevice(c.device));
    int rc;
    if (!c.tout) SPR_HIP(hipMalloc((void **)&c.tout, (size_t)2 * c.nq0 * c.nd * sizeof(double)));
    const size_t n = (size_t)c.nq0 * c.nd;
    if ((rc = spr::launch_lw3_tetra(c, 0, c.nq0, c.tout, c.tout + n))) return rc;
    SPR_HIP(hipMemcpy(lw_all, c.tout, n * sizeof(double), hipMemcpyDeviceToHost));
    SPR_HIP(hipMemcpy(p3_all, c.tout + n, n * sizeof(double), hipMemcpyDeviceToHost));
    return SPR_OK;
}
