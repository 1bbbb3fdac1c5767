// Device helpers of the linear tetrahedron method shared by the three- and four-phonon kernels:
// a stable sort of four vertex records by energy, the reciprocals of the six vertex-energy
// differences (shared by every Mq of one tetrahedron), and g and the vertex weights of
// FindgAndIver (TetrahedronIntegration_m.f90:96-256) written with those reciprocals.
#pragma once

namespace spr {

// one vertex record: energy, s0-independent part of F, W offset of s0 = 0 (a W buffer holds at most
// 2^31 doubles)
struct Vx {
    double M, a;
    int off;
};

__device__ __forceinline__ void cswap(Vx &x, Vx &y) {
    // strict: equal energies keep their order (the reference's sort keeps the first of equal
    // values first)
    if (x.M > y.M) {
        const Vx t = x;
        x = y;
        y = t;
    }
}

// stable insertion sort of four records by M, fully unrolled (compare-exchange of neighbours)
__device__ __forceinline__ void sort4(Vx *v) {
    cswap(v[0], v[1]);
    cswap(v[1], v[2]);
    cswap(v[0], v[1]);
    cswap(v[2], v[3]);
    cswap(v[1], v[2]);
    cswap(v[0], v[1]);
}

// Reciprocals of the six differences of one sorted tetrahedron's vertex energies; they do not
// depend on s0.  A zero difference gives inf, which turns the reference's 0/0 into 0 * inf =
// NaN and its x/0 into x * inf = inf, so the NaN handling below sees the same cases.
struct Rcp {
    double r10, r20, r30, r21, r31, r32;
};

__device__ __forceinline__ Rcp rcp6(const double *M) {
    return Rcp{1.0 / (M[1] - M[0]), 1.0 / (M[2] - M[0]), 1.0 / (M[3] - M[0]),
               1.0 / (M[2] - M[1]), 1.0 / (M[3] - M[1]), 1.0 / (M[3] - M[2])};
}

// g * (1/6) sum_v I_v F_v of one sorted tetrahedron at mq, and g (FindgAndIver / FindIverxFver,
// TetrahedronIntegration_m.f90:96-256, with the divisions by vertex-energy differences as
// products with R).  NAN0: NaN I_v and g set to zero, as the four-phonon routine does
// (Tetrahedron4th_sep.f90:238-263); the three-phonon routine keeps them.
template <bool NAN0>
__device__ __forceinline__ double tetra_rcp(double mq, const double *M, const Rcp &R, const double *F, double &g_out) {
    constexpr double third = 1.0 / 3.0;
    double g, I0, I1, I2, I3;
    if (mq >= M[0] && mq < M[1]) {
        const double x0 = mq - M[0];
        g = 3.0 * x0 * x0 * (R.r10 * R.r20 * R.r30);
        I0 = ((M[1] - mq) * R.r10 + (M[2] - mq) * R.r20 + (M[3] - mq) * R.r30) * third;
        I1 = x0 * R.r10 * third;
        I2 = x0 * R.r20 * third;
        I3 = x0 * R.r30 * third;
    } else if (mq >= M[1] && mq < M[2]) {
        const double x0 = mq - M[0], x1 = mq - M[1], x2 = mq - M[2], x3 = mq - M[3];
        // (M1 - M2) = -d21 and (M1 - M3) = -d31
        g = (-3.0 * R.r21 * R.r30) * (x0 * x2 * R.r20 + x1 * x3 * R.r31);
        const double igd = 1.0 / (g * (M[3] - M[0]));
        I0 = (-x3 * R.r30) * third + (-x2 * R.r20) * (x0 * R.r20) * (-x2 * R.r21) * igd;
        I1 = (-x2 * R.r21) * third + (-x3 * R.r31) * (-x3 * R.r31) * (x1 * R.r21) * igd;
        I2 = (x1 * R.r21) * third + (x0 * R.r20) * (x0 * R.r20) * (-x2 * R.r21) * igd;
        I3 = (x0 * R.r30) * third + (x1 * R.r31) * (-x3 * R.r31) * (x1 * R.r21) * igd;
    } else if (mq >= M[2] && mq <= M[3]) {
        const double x3 = mq - M[3];
        g = 3.0 * x3 * x3 * (R.r30 * R.r31 * R.r32);
        I0 = -x3 * R.r30 * third;
        I1 = -x3 * R.r31 * third;
        I2 = -x3 * R.r32 * third;
        I3 = ((mq - M[0]) * R.r30 + (mq - M[1]) * R.r31 + (mq - M[2]) * R.r32) * third;
    } else {
        g_out = 0.0;
        return 0.0;
    }
    if constexpr (NAN0) {
        I0 = I0 != I0 ? 0.0 : I0;
        I1 = I1 != I1 ? 0.0 : I1;
        I2 = I2 != I2 ? 0.0 : I2;
        I3 = I3 != I3 ? 0.0 : I3;
        g = g != g ? 0.0 : g;
    }
    g_out = g;
    double ixf = 0.0;
    ixf += I0 * F[0];
    ixf += I1 * F[1];
    ixf += I2 * F[2];
    ixf += I3 * F[3];
    return g * ((1.0 / 6.0) * ixf);
}


}  // namespace spr
