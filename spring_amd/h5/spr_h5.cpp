// HDF5 files of the scattering-rate path: the FC2/FC3/FC4 files kappa.x reads and the kappa
// file it writes, in the reference's layout (dataset names, types, shapes).  Host code over
// the HDF5 C API; the HDF5 Fortran API the reference uses lists dimensions in the reverse
// order of the C API, so a Fortran (d1, ..., dn) dataset is created here as C (dn, ..., d1)
// and the bytes are the same.
#include <hdf5.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/spring_h5.h"

namespace {

thread_local std::string g_err;

int fail(const std::string &msg) {
    g_err = msg;
    return -1;
}

// closes an HDF5 identifier with the matching call when it leaves scope
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    Hid(const Hid &) = delete;
    ~Hid() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
};

void quiet() {
    static bool done = false;
    if (!done) {
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors are reported through spr_h5_last_error
        done = true;
    }
}

// C-order dimensions of a dataset (empty on failure)
std::vector<hsize_t> dims_of(hid_t dset) {
    Hid sp(H5Dget_space(dset), H5Sclose);
    if (!sp.ok()) return {};
    const int r = H5Sget_simple_extent_ndims(sp.id);
    if (r <= 0) return {};
    std::vector<hsize_t> d(r);
    H5Sget_simple_extent_dims(sp.id, d.data(), nullptr);
    return d;
}

std::string dims_str(const std::vector<hsize_t> &d) {
    std::string s = "(";
    for (size_t i = d.size(); i-- > 0;) s += std::to_string((unsigned long long)d[i]) + (i ? ", " : ")");
    return s;
}

int read_ds(hid_t loc, const char *name, hid_t memtype, const std::vector<hsize_t> &want, void *buf,
            const std::string &file) {
    Hid ds(H5Dopen2(loc, name, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(file + ": no dataset " + name);
    const std::vector<hsize_t> d = dims_of(ds.id);
    if (d != want) return fail(file + ": dataset " + name + " has Fortran shape " + dims_str(d) + ", expected " + dims_str(want));
    if (H5Dread(ds.id, memtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) return fail(file + ": reading " + name + " failed");
    return 0;
}

int write_ds(hid_t loc, const char *name, hid_t type, const std::vector<hsize_t> &d, const void *buf) {
    Hid sp(H5Screate_simple((int)d.size(), d.data(), nullptr), H5Sclose);
    if (!sp.ok()) return fail(std::string("dataspace for ") + name);
    Hid ds(H5Dcreate2(loc, name, type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    if (!ds.ok()) return fail(std::string("creating dataset ") + name + " failed");
    if (H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) < 0) return fail(std::string("writing ") + name + " failed");
    return 0;
}

// per basis atom FC block: group and dataset name, C-order dims (hdf5_wrap_mod.f90:140-210, 250-320)
const char *group_of(int order) { return order == 3 ? "FC3rd" : "FC4th"; }
std::string mu_name(int order, int mu) { return std::string(order == 3 ? "FC3_mu" : "FC4_mu") + std::to_string(mu + 1); }
std::vector<hsize_t> block_dims(int order, int natm) {
    std::vector<hsize_t> d(order - 1, (hsize_t)natm);
    for (int i = 0; i < order; ++i) d.push_back(3);
    return d;
}
size_t block_size(int order, int natm) {
    size_t n = 1;
    for (int i = 0; i < order - 1; ++i) n *= (size_t)natm;
    for (int i = 0; i < order; ++i) n *= 3;
    return n;
}

}  // namespace

extern "C" const char *spr_h5_last_error(void) { return g_err.c_str(); }

extern "C" int spr_h5_fc_dims(const char *file, int order, int *nb, int *natm, int *lead) {
    quiet();
    if (!file || !nb || !natm || (order < 2 || order > 4)) return fail("spr_h5_fc_dims: bad argument");
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot open as HDF5");
    Hid ix(H5Dopen2(f.id, "atm_Indx", H5P_DEFAULT), H5Dclose);
    if (!ix.ok()) return fail(std::string(file) + ": no dataset atm_Indx");
    const std::vector<hsize_t> d = dims_of(ix.id);
    if (d.size() != 3 || d[2] != 4) return fail(std::string(file) + ": atm_Indx is not (4, Natm, Nb)");
    *nb = (int)d[0];
    *natm = (int)d[1];
    int L = 1;
    if (order == 2) {
        Hid fc(H5Dopen2(f.id, "FC2", H5P_DEFAULT), H5Dclose);
        if (!fc.ok()) return fail(std::string(file) + ": no dataset FC2");
        const std::vector<hsize_t> e = dims_of(fc.id);
        if (e.size() != 5) return fail(std::string(file) + ": FC2 is not rank 5");
        L = (int)e[4];
    }
    if (lead) *lead = L;
    return 0;
}

extern "C" int spr_h5_read_fc(const char *file, int order, int nb, int natm, int *atm_num, int *atm_indx, double *fc) {
    quiet();
    if (!file || !atm_num || !atm_indx || !fc || nb < 1 || natm < 1 || order < 2 || order > 4)
        return fail("spr_h5_read_fc: bad argument");
    int nb_f, natm_f, L;
    if (spr_h5_fc_dims(file, order, &nb_f, &natm_f, &L)) return -1;
    if (nb_f != nb || natm_f != natm)
        return fail(std::string(file) + ": file holds Nb = " + std::to_string(nb_f) + ", Natm = " + std::to_string(natm_f));
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot open as HDF5");
    if (read_ds(f.id, "NumAtoms", H5T_NATIVE_INT, {(hsize_t)nb}, atm_num, file)) return -1;
    if (read_ds(f.id, "atm_Indx", H5T_NATIVE_INT, {(hsize_t)nb, (hsize_t)natm, 4}, atm_indx, file)) return -1;
    if (order == 2) {
        // FC2(L, 3, 3, Natm, Nb): kappa.x uses the slice FC2(1, :, :, :, :)
        std::vector<double> all((size_t)nb * natm * 9 * L);
        if (read_ds(f.id, "FC2", H5T_NATIVE_DOUBLE, {(hsize_t)nb, (hsize_t)natm, 3, 3, (hsize_t)L}, all.data(), file))
            return -1;
        for (size_t i = 0; i < (size_t)nb * natm * 9; ++i) fc[i] = all[i * L];
        return 0;
    }
    Hid g(H5Gopen2(f.id, group_of(order), H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(std::string(file) + ": no group " + group_of(order));
    const size_t bs = block_size(order, natm);
    for (int mu = 0; mu < nb; ++mu)
        if (read_ds(g.id, mu_name(order, mu).c_str(), H5T_NATIVE_DOUBLE, block_dims(order, natm), fc + (size_t)mu * bs, file))
            return -1;
    return 0;
}

extern "C" int spr_h5_write_fc(const char *file, int order, int nb, int natm, const int *atm_num, const int *atm_indx,
                               const double *fc) {
    quiet();
    if (!file || !atm_num || !atm_indx || !fc || nb < 1 || natm < 1 || order < 2 || order > 4)
        return fail("spr_h5_write_fc: bad argument");
    Hid f(H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot create");
    if (write_ds(f.id, "NumAtoms", H5T_NATIVE_INT, {(hsize_t)nb}, atm_num)) return -1;
    if (write_ds(f.id, "atm_Indx", H5T_NATIVE_INT, {(hsize_t)nb, (hsize_t)natm, 4}, atm_indx)) return -1;
    if (order == 2) return write_ds(f.id, "FC2", H5T_NATIVE_DOUBLE, {(hsize_t)nb, (hsize_t)natm, 3, 3, 1}, fc);
    Hid g(H5Gcreate2(f.id, group_of(order), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return fail(std::string(file) + ": cannot create group " + group_of(order));
    const size_t bs = block_size(order, natm);
    for (int mu = 0; mu < nb; ++mu)
        if (write_ds(g.id, mu_name(order, mu).c_str(), H5T_NATIVE_DOUBLE, block_dims(order, natm), fc + (size_t)mu * bs))
            return -1;
    return 0;
}

// (2 Ndof, Nq0) = [omega(:, q0); x(:, q0)] per q0 (w2dh5_no_grp_type2, hdf5_wrap_mod.f90:722-765)
static std::vector<double> stack2(int ndof, int nq0, const double *omega, const double *x) {
    std::vector<double> a((size_t)2 * ndof * nq0);
    for (int q = 0; q < nq0; ++q) {
        std::memcpy(&a[(size_t)q * 2 * ndof], omega + (size_t)q * ndof, ndof * sizeof(double));
        std::memcpy(&a[(size_t)q * 2 * ndof + ndof], x + (size_t)q * ndof, ndof * sizeof(double));
    }
    return a;
}

extern "C" int spr_h5_write_kappa(const char *file, const double *kappa, int ndof, int nq0, const double *omega_q0,
                                  const double *lw, const double *p3, const double *mode_k) {
    quiet();
    if (!file || !kappa || !omega_q0 || !lw || ndof < 1 || nq0 < 1) return fail("spr_h5_write_kappa: bad argument");
    Hid f(H5Fcreate(file, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot create");
    if (write_ds(f.id, "ThermalConductivityTensor", H5T_NATIVE_DOUBLE, {9}, kappa)) return -1;
    std::vector<double> a = stack2(ndof, nq0, omega_q0, lw);
    if (write_ds(f.id, "LineWidth", H5T_NATIVE_DOUBLE, {(hsize_t)nq0, (hsize_t)2 * ndof}, a.data())) return -1;
    if (p3) {
        a = stack2(ndof, nq0, omega_q0, p3);
        if (write_ds(f.id, "PhaseSpace3rd", H5T_NATIVE_DOUBLE, {(hsize_t)nq0, (hsize_t)2 * ndof}, a.data())) return -1;
    }
    if (mode_k && write_ds(f.id, "ModeResolvedKappa", H5T_NATIVE_DOUBLE, {(hsize_t)ndof, 9}, mode_k)) return -1;
    return 0;
}

extern "C" int spr_h5_read_kappa(const char *file, int ndof, int nq0, double *kappa, double *omega_q0, double *lw,
                                 double *p3, double *mode_k) {
    quiet();
    if (!file || ndof < 1 || nq0 < 1) return fail("spr_h5_read_kappa: bad argument");
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot open as HDF5");
    if (kappa && read_ds(f.id, "ThermalConductivityTensor", H5T_NATIVE_DOUBLE, {9}, kappa, file)) return -1;
    std::vector<double> a((size_t)2 * ndof * nq0);
    auto split = [&](double *x) {
        for (int q = 0; q < nq0; ++q) {
            if (omega_q0) std::memcpy(omega_q0 + (size_t)q * ndof, &a[(size_t)q * 2 * ndof], ndof * sizeof(double));
            if (x) std::memcpy(x + (size_t)q * ndof, &a[(size_t)q * 2 * ndof + ndof], ndof * sizeof(double));
        }
    };
    if (lw || omega_q0) {
        if (read_ds(f.id, "LineWidth", H5T_NATIVE_DOUBLE, {(hsize_t)nq0, (hsize_t)2 * ndof}, a.data(), file)) return -1;
        split(lw);
    }
    if (p3) {
        if (read_ds(f.id, "PhaseSpace3rd", H5T_NATIVE_DOUBLE, {(hsize_t)nq0, (hsize_t)2 * ndof}, a.data(), file)) return -1;
        for (int q = 0; q < nq0; ++q) std::memcpy(p3 + (size_t)q * ndof, &a[(size_t)q * 2 * ndof + ndof], ndof * sizeof(double));
    }
    if (mode_k && read_ds(f.id, "ModeResolvedKappa", H5T_NATIVE_DOUBLE, {(hsize_t)ndof, 9}, mode_k, file)) return -1;
    return 0;
}

// SpectralResolvedKappa (11, Ndof Nq0) added to an existing kappa file (w2dh5_no_grp,
// hdf5_wrap_mod.f90:650-721, as main.f90:780-794 writes it)
extern "C" int spr_h5_write_spectral(const char *file, int ndof, int nq0, const double *spectral_k) {
    quiet();
    if (!file || !spectral_k || ndof < 1 || nq0 < 1) return fail("spr_h5_write_spectral: bad argument");
    Hid f(H5Fopen(file, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot open for writing (write the kappa file first)");
    return write_ds(f.id, "SpectralResolvedKappa", H5T_NATIVE_DOUBLE, {(hsize_t)ndof * nq0, 11}, spectral_k);
}

extern "C" int spr_h5_read_spectral(const char *file, int ndof, int nq0, double *spectral_k) {
    quiet();
    if (!file || !spectral_k || ndof < 1 || nq0 < 1) return fail("spr_h5_read_spectral: bad argument");
    Hid f(H5Fopen(file, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return fail(std::string(file) + ": cannot open as HDF5");
    return read_ds(f.id, "SpectralResolvedKappa", H5T_NATIVE_DOUBLE, {(hsize_t)ndof * nq0, 11}, spectral_k, file);
}
