// hex_plastic.hpp -- the plastic solid-shell element update in FP64, with its strain operator
// formed from the element's rest coordinates rather than read from memory.
//
// elementPlasticUpdate (engine/sim/solid_shell.cpp:826-1029) on the forms of
// computeForms (solid_shell.cpp:237-365), evaluated where they are used:
//   - covariant strain rows [E11, E22, E33, 2E12, 2E23, 2E31] of the rest base vectors g_k
//     and the displacement's natural derivatives (solid_shell.cpp:140-158);
//   - assumed natural transverse shear from (xi, eta) = (-1, 0), (+1, 0) for 2E23 and
//     (0, -1), (0, +1) for 2E31, assumed natural thickness strain from the four in-plane
//     corners, each at the Gauss point's zeta (solid_shell.cpp:181-235);
//   - the push to Cartesian strain with A = J^-1, eps = A^T E A (solid_shell.cpp:116-138);
//   - seven enhanced modes: column k at point g is (det J0/det J_g) m_k(g) T0[:, r_k],
//     normalised in the weighted L2 norm over the points (solid_shell.cpp:160-179, 313-365).
// The force is the same operator transposed: the Cartesian stress is pulled back to the
// covariant frame, Q = A S A^T, and spread over the nodes through the same rows, the
// assumed rows through their sampling points. Everything works on corner offsets from
// corner 0 (X_0 = 0, d_0 = 0): the operator annihilates translations.
//
// Geometry and displacement enter as the coefficients of their trilinear fields over the
// seven non-constant monomials (basis() below), so every natural derivative is a handful of
// multiply-adds and the force is assembled on those coefficients, then mapped to corners once.
// Shared by the device pass and by host code; the caller supplies the basis coefficients of
// the rest and co-rotated corner offsets and receives the internal force of corners 1..7 in
// the co-rotated frame.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace shipfem::hexp {

struct Material {
    double mu, kappa, yield, hard;
    double fail, sens, eref, ecut;  // failure strain (<= 0: none) and Rice-Tracey constants
    bool damage;
    // constants of the return, formed once (derive()): 1 / (2 mu + 2/3 H),
    // 1 / (1 + H / (3 mu)), 1 / (2 mu), 1 / (3 kappa), 1 / mu
    double igam, thb0, i2mu, i3k, imu;
};

__host__ __device__ inline Material derive(Material m) {
    m.igam = 1.0 / (2.0 * m.mu + (2.0 / 3.0) * m.hard);
    m.thb0 = 1.0 / (1.0 + m.hard / (3.0 * m.mu));
    m.i2mu = 1.0 / (2.0 * m.mu);
    m.i3k = 1.0 / (3.0 * m.kappa);
    m.imu = 1.0 / m.mu;
    return m;
}

// 1/x and 1/sqrt(x) for the well-scaled arguments of this update (Jacobian determinants,
// Cholesky pivots, stress norms). On the device: the hardware estimate and two Newton
// steps, each of which at least squares the relative error, so the result is within an
// ulp or so of the quotient without the IEEE division's scaling and fix-up sequence. On
// the host: the plain quotient.
__host__ __device__ inline double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
#else
    return 1.0 / x;
#endif
}
__host__ __device__ inline double invSqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    // y <- y + y (1 - x y^2) / 2
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
#else
    return 1.0 / sqrt(x);
#endif
}

struct Result {
    double dissipation = 0.0;  // sum w sigma : d eps_p
    double released = 0.0;     // stored energy of the points that failed in this update
    double strain = 0.0;       // sum w 1/2 sigma : C^-1 : sigma (when asked)
    int failedStart = 0, failedEnd = 0;
};

// The trilinear field of values v_a at the corners, written over the monomials
//   phi = (xi, eta, zeta, xi eta, eta zeta, zeta xi, xi eta zeta)
// as v(xi, eta, zeta) = v_centre + sum_m phi_m c_m with c_m = 1/8 sum_a s_m(a) v_a, s_m(a) the
// monomial at corner a's natural coordinates. A natural derivative anywhere is then one
// coefficient plus three scaled ones, and a force conjugate to the c_m maps back to the
// corners through the same signs: f_a = 1/8 sum_m s_m(a) F_m.
__host__ __device__ constexpr double cornerSign(int m, int a) {
    const int x = (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : -1;
    const int y = (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : -1;
    const int z = a >= 4 ? 1 : -1;
    const int s[7] = {x, y, z, x * y, y * z, z * x, x * y * z};
    return s[m];
}

// c[3 m + i] of values v[3 (a - 1) + i] at corners 1..7, corner 0 at zero.
__host__ __device__ inline void basis(const double* v, double c[21]) {
#pragma unroll
    for (int m = 0; m < 7; ++m)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double t = 0.0;
#pragma unroll
            for (int a = 1; a < 8; ++a) t += cornerSign(m, a) * v[3 * (a - 1) + i];
            c[3 * m + i] = 0.125 * t;
        }
}

// Coefficients held with a stride (a lane's column of a block-shared array on the device).
struct Strided {
    const double* p;
    uint32_t s;
    __host__ __device__ double operator[](int j) const { return p[j * s]; }
};

// Column l of the natural derivative at (xi, eta, zeta): out[i] = d v_i / d xi_l.
//   l = 0: c_0 + eta c_3 + zeta c_5 + eta zeta c_6
//   l = 1: c_1 + xi c_3 + zeta c_4 + xi zeta c_6
//   l = 2: c_2 + eta c_4 + xi c_5 + xi eta c_6
template <class C>
__host__ __device__ inline void column(const C& c, int l, double xi, double eta, double zeta,
                                       double out[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (l == 0)
            out[i] = c[i] + eta * c[9 + i] + zeta * c[15 + i] + eta * zeta * c[18 + i];
        else if (l == 1)
            out[i] = c[3 + i] + xi * c[9 + i] + zeta * c[12 + i] + xi * zeta * c[18 + i];
        else
            out[i] = c[6 + i] + eta * c[12 + i] + xi * c[15 + i] + xi * eta * c[18 + i];
    }
}

// J[i][k] = d v_i / d xi_k at a point.
template <class C>
__host__ __device__ inline void jacobian(const C& c, double xi, double eta, double zeta,
                                         double J[3][3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double col[3];
        column(c, k, xi, eta, zeta, col);
#pragma unroll
        for (int i = 0; i < 3; ++i) J[i][k] = col[i];
    }
}

// F_m += P d phi_m / d xi_l at (xi, eta, zeta).
__host__ __device__ inline void addGrad(double F[21], const double P[3], int l, double xi,
                                        double eta, double zeta) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (l == 0) {
            F[0 + i] += P[i];
            F[9 + i] += eta * P[i];
            F[15 + i] += zeta * P[i];
            F[18 + i] += eta * zeta * P[i];
        } else if (l == 1) {
            F[3 + i] += P[i];
            F[9 + i] += xi * P[i];
            F[12 + i] += zeta * P[i];
            F[18 + i] += xi * zeta * P[i];
        } else {
            F[6 + i] += P[i];
            F[12 + i] += eta * P[i];
            F[15 + i] += xi * P[i];
            F[18 + i] += xi * eta * P[i];
        }
    }
}

__host__ __device__ inline double inverse3(const double m[3][3], double r[3][3], double* inv = nullptr) {
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    const double id = recip(det);
    r[0][0] = c00 * id;
    r[1][0] = c01 * id;
    r[2][0] = c02 * id;
    r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    if (inv) *inv = id;
    return det;
}

// Cartesian Voigt strain (engineering shears) of a covariant Voigt strain: A^T E A.
__host__ __device__ inline void push(const double A[3][3], const double e[6], double out[6]) {
    const double E[3][3] = {{e[0], 0.5 * e[3], 0.5 * e[5]},
                            {0.5 * e[3], e[1], 0.5 * e[4]},
                            {0.5 * e[5], 0.5 * e[4], e[2]}};
    double EA[3][3], M[3][3];
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) EA[k][j] = E[k][0] * A[0][j] + E[k][1] * A[1][j] + E[k][2] * A[2][j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[i][j] = A[0][i] * EA[0][j] + A[1][i] * EA[1][j] + A[2][i] * EA[2][j];
    out[0] = M[0][0];
    out[1] = M[1][1];
    out[2] = M[2][2];
    out[3] = 2.0 * M[0][1];
    out[4] = 2.0 * M[1][2];
    out[5] = 2.0 * M[2][0];
}

// The work-conjugate of push: covariant Voigt stress tau = T^T sigma, Q = A S A^T.
__host__ __device__ inline void pull(const double A[3][3], const double s[6], double out[6]) {
    const double S[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
    double SA[3][3], Q[3][3];
    for (int i = 0; i < 3; ++i)
        for (int l = 0; l < 3; ++l) SA[i][l] = S[i][0] * A[l][0] + S[i][1] * A[l][1] + S[i][2] * A[l][2];
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) Q[k][l] = A[k][0] * SA[0][l] + A[k][1] * SA[1][l] + A[k][2] * SA[2][l];
    out[0] = Q[0][0];
    out[1] = Q[1][1];
    out[2] = Q[2][2];
    out[3] = Q[0][1];
    out[4] = Q[1][2];
    out[5] = Q[2][0];
}

__host__ __device__ inline double gauss(int g, int axis) {
    const double q = 0.57735026918962576451;
    return (g >> axis) & 1 ? q : -q;
}

// the enhanced mode k's natural-coordinate monomial and its covariant row
__host__ __device__ inline double monomial(int k, double xi, double eta, double zeta) {
    return k == 0 ? xi : k == 1 ? eta : k == 2 ? zeta : k == 3 ? xi * zeta : k == 4 ? eta * zeta
                                                              : k == 5 ? xi : eta;
}
__host__ __device__ inline int modeRow(int k) { return k < 2 ? k : k < 5 ? 2 : 3; }
// the axes of mode k's monomial as bits (xi 1, eta 2, zeta 4) and its degree
__host__ __device__ constexpr int modeAxes(int k) {
    return k == 0 || k == 5 ? 1 : k == 1 || k == 6 ? 2 : k == 2 ? 4 : k == 3 ? 5 : 6;
}
__host__ __device__ constexpr int modeDegree(int k) { return k == 3 || k == 4 ? 2 : 1; }

// Assumed-strain samples of one zeta level: 2E23 at (-1, 0), (+1, 0); 2E31 at (0, -1),
// (0, +1); E33 at the corners (-1,-1), (1,-1), (1,1), (-1,1).
constexpr double kShear[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
constexpr double kCorner[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};

// eqps grows with every plastic increment, so a point whose eqps is +0 has never yielded:
// its plastic strain and damage are zero and are not read. (A history handed over with
// plastic strain but zero eqps carries -0, which reads.)
__host__ __device__ inline bool hasHistory(double eqps) { return __builtin_bit_cast(uint64_t, eqps) != 0; }

// Point g's history (plastic strain 6, eqps, damage) at the start of the step. hmask has bit
// g set when point g had yielded before the step (its eqps is not +0): the others have a zero
// history and are not read.
__host__ __device__ inline void loadPoint(const double* hist, int g, uint32_t stride, int hmask, double h0[8]) {
    if ((hmask >> g) & 1) {
        for (int k = 0; k < 8; ++k) h0[k] = hist[(8 * g + k) * stride];
    } else {
        for (int k = 0; k < 8; ++k) h0[k] = 0.0;
    }
}

__host__ __device__ inline double dot3(const double a[3], const double b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// The covariant strain at Gauss point g of the displacement with coefficients cd, over the
// rest geometry cX (rows 0, 1, 3 at the point, rows 2, 4, 5 from the samples of its level).
// J is the rest Jacobian at the point.
__host__ __device__ inline void covariantStrain(const double J[3][3], const double* cd, int g,
                                                const double smp[8], double e[6]) {
    const double xi = gauss(g, 0), eta = gauss(g, 1), zeta = gauss(g, 2);
    double u0[3], u1[3];
    column(cd, 0, xi, eta, zeta, u0);
    column(cd, 1, xi, eta, zeta, u1);
    const double g0[3] = {J[0][0], J[1][0], J[2][0]}, g1[3] = {J[0][1], J[1][1], J[2][1]};
    e[0] = dot3(g0, u0);
    e[1] = dot3(g1, u1);
    e[3] = dot3(g0, u1) + dot3(g1, u0);
    e[4] = 0.5 * (1 - xi) * smp[0] + 0.5 * (1 + xi) * smp[1];
    e[5] = 0.5 * (1 - eta) * smp[2] + 0.5 * (1 + eta) * smp[3];
    double t = 0.0;
    for (int s = 0; s < 4; ++s) t += 0.25 * (1 + xi * kCorner[s][0]) * (1 + eta * kCorner[s][1]) * smp[4 + s];
    e[2] = t;
}

// The samples of one zeta level for the displacement with coefficients cd.
template <class CX>
__host__ __device__ inline void samples(const CX& cX, const double* cd, double zeta, double smp[8]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const double xi = kShear[s][0], eta = kShear[s][1];
        // 2E23 = g_1 . d,zeta + g_2 . d,eta; 2E31 = g_2 . d,xi + g_0 . d,zeta
        const int l = s < 2 ? 1 : 0;
        double ga[3], gz[3], ua[3], uz[3];
        column(cX, l, xi, eta, zeta, ga);
        column(cX, 2, xi, eta, zeta, gz);
        column(cd, l, xi, eta, zeta, ua);
        column(cd, 2, xi, eta, zeta, uz);
        smp[s] = dot3(ga, uz) + dot3(gz, ua);
        column(cX, 2, kCorner[s][0], kCorner[s][1], zeta, gz);
        column(cd, 2, kCorner[s][0], kCorner[s][1], zeta, uz);
        smp[4 + s] = dot3(gz, uz);
    }
}

// One point's return from its history h0 (plastic strain 6, eqps, damage) for strain e
// (plasticity.cpp:319-505, linear hardening, no back stress). Returns bit 0 yielded, bit 1
// failed now, bit 2 failed (no stress, no tangent).
__host__ __device__ inline int pointReturn(const Material& m, const double e[6], const double h0[8],
                                           double h[8], double sig[6], double& diss,
                                           double& theta, double& thetabar, double n[6],
                                           double& released) {
    diss = 0.0;
    released = 0.0;
    theta = 1.0;
    thetabar = 0.0;
    for (int k = 0; k < 8; ++k) h[k] = h0[k];
    for (int i = 0; i < 6; ++i) n[i] = 0.0;
    if (h0[7] >= 1.0) {
        for (int i = 0; i < 6; ++i) sig[i] = 0.0;
        return 4;
    }
    const double mu = m.mu, kappa = m.kappa;
    double el[6];
    for (int i = 0; i < 6; ++i) el[i] = e[i] - h0[i];
    const double vol = el[0] + el[1] + el[2];
    const double p = kappa * vol;
    double tr[6];
    for (int i = 0; i < 3; ++i) {
        tr[i] = 2.0 * mu * (el[i] - vol * (1.0 / 3.0));
        tr[3 + i] = mu * el[3 + i];
    }
    const double n2 = tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2] +
                      2.0 * (tr[3] * tr[3] + tr[4] * tr[4] + tr[5] * tr[5]);
    const double root23 = 0.81649658092772603273;
    const double radius = root23 * (m.yield + m.hard * fmax(h0[6], 0.0));
    // inside the surface by the squared norms first: an elastic point takes no root
    double nrm = 0.0, inrm = 0.0;
    if (n2 > radius * radius) {
        inrm = invSqrt(n2);
        nrm = n2 * inrm;
    }
    const double excess = nrm - radius;
    if (!(excess > 0.0)) {
        for (int i = 0; i < 3; ++i) {
            sig[i] = p + tr[i];
            sig[3 + i] = tr[3 + i];
        }
        return 0;
    }
    const double gamma = excess * m.igam;
    double dep[6];
    for (int i = 0; i < 6; ++i) n[i] = tr[i] * inrm;
    for (int i = 0; i < 3; ++i) {
        dep[i] = gamma * n[i];
        dep[3 + i] = 2.0 * gamma * n[3 + i];
        sig[i] = p + tr[i] - 2.0 * mu * gamma * n[i];
        sig[3 + i] = tr[3 + i] - 2.0 * mu * gamma * n[3 + i];
    }
    double dd = 0.0;
    for (int i = 0; i < 6; ++i) {
        dd += sig[i] * dep[i];
        h[i] = h0[i] + dep[i];
    }
    diss = dd;
    const double deq = root23 * gamma;
    h[6] = h0[6] + deq;
    theta = 1.0 - 2.0 * mu * gamma * inrm;
    thetabar = m.thb0 - (1.0 - theta);
    if (m.damage && m.fail > 0.0) {
        const double q = 1.22474487139158904909 * (nrm - 2.0 * mu * gamma);
        const double eta = q > 0.0 ? p * recip(q) : 0.0;
        if (eta > m.ecut) {
            const double critical = m.fail * exp(-m.sens * (eta - m.eref));
            if (critical > 0.0 && critical <= 1.7976931348623157e308)  // positive and finite
                h[7] = h0[7] + deq / critical;
        }
        if (h[7] >= 1.0) {
            h[7] = 1.0;
            const double pm = (sig[0] + sig[1] + sig[2]) * (1.0 / 3.0);
            double u = 0.0;
            for (int i = 0; i < 3; ++i) u += 0.5 * sig[i] * ((sig[i] - pm) * m.i2mu + pm * m.i3k);
            for (int i = 3; i < 6; ++i) u += 0.5 * sig[i] * (sig[i] * m.imu);
            released = u;
            for (int i = 0; i < 6; ++i) sig[i] = 0.0;
            return 1 | 2 | 4;
        }
    }
    return 1;
}

// The element's points at the start of the step: which have a history (bit g) and how many
// have failed. Read once per step: the Newton passes and the commit test the mask.
struct Start {
    int hmask = 0, failed = 0;
};
// (With the points' eqps already loaded: only a point with a history has its damage read.)
__host__ __device__ inline Start startOf(const double eqps[8], const double* hist, uint32_t stride) {
    Start s;
    for (int g = 0; g < 8; ++g) s.hmask |= hasHistory(eqps[g]) ? 1 << g : 0;
    if (s.hmask)
        for (int g = 0; g < 8; ++g)
            if ((s.hmask >> g) & 1) s.failed += hist[(8 * g + 7) * stride] >= 1.0;
    return s;
}
__host__ __device__ inline Start startOf(const double* hist, uint32_t stride) {
    double eqps[8];
    for (int g = 0; g < 8; ++g) eqps[g] = hist[(8 * g + 6) * stride];
    return startOf(eqps, hist, stride);
}

// The element update. cXs, cX: basis coefficients of the rest corner offsets (corner 0 at
// zero), the same values as the set-up reads them (an array) and as the later passes read
// them (on the device, this lane's column of a block-shared array); cd: those of the
// co-rotated displacement offsets; hist: 8 points x (plastic strain 6, eqps, damage) at
// hist[(8 g + k) * stride]; enh: 7 enhanced parameters at enh[k * stride]. Writes the internal force of corners 1..7 (corner 0 carries minus their
// sum) in the co-rotated frame. A torn element (all points failed) is left as it is and
// returns a zero force. start: startOf(hist, stride) at the step's start; alpha0, when
// given, the enhanced parameters as already read from enh.
template <class CX>
__host__ __device__ inline Result update(const Material& m, const double* cXs, const CX& cX, const double* cd,
                                         double* hist, double* enh, uint32_t stride,
                                         bool wantStrain, double f[21], double* eu,
                                         uint32_t euStride, const Start& start,
                                         const double* alpha0 = nullptr) {
    Result res;
    for (int j = 0; j < 21; ++j) f[j] = 0.0;
    const int hmask = start.hmask;
    res.failedStart = start.failed;
    if (m.damage && res.failedStart == 8) {
        res.failedEnd = 8;
        return res;
    }
    // the centre's map T0 (columns 0-3): the centre Jacobian's columns are c_0, c_1, c_2
    double J0[3][3], A0[3][3];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) J0[i][k] = cXs[3 * k + i];
    const double det0 = inverse3(J0, A0);
    double t0[4][6];
    for (int r = 0; r < 4; ++r) {
        double unit[6] = {0, 0, 0, 0, 0, 0};
        unit[r] = 1.0;
        push(A0, unit, t0[r]);
    }
    // T0[:, p]^T C T0[:, q] = kappa tr_p tr_q + mu theta D_pq - 2 mu thetaBar nv_p nv_q for
    // C = kappa 1(x)1 + 2 mu theta I_dev - 2 mu thetaBar n(x)n, engineering shears: tr and D
    // are the element's, nv = n . T0[:, q] the point's
    double tr4[4], D[10];
    for (int p = 0; p < 4; ++p) tr4[p] = t0[p][0] + t0[p][1] + t0[p][2];
    for (int p = 0; p < 4; ++p)
        for (int q = 0; q <= p; ++q) {
            double t = 0.0;
            for (int i = 0; i < 3; ++i) t += 2.0 * t0[p][i] * (t0[q][i] - tr4[q] * (1.0 / 3.0));
            for (int i = 3; i < 6; ++i) t += t0[p][i] * t0[q][i];
            D[p * (p + 1) / 2 + q] = t;
        }
    // eu: the base strain of point g, row i at eu[(6 g + i) * euStride] (the caller's
    // buffer: on the device, this lane's column of a block-shared array)
    double w[8], sc[8];
    {
        double smp[2][8];
        samples(cXs, cd, -0.57735026918962576451, smp[0]);
        samples(cXs, cd, 0.57735026918962576451, smp[1]);
        for (int g = 0; g < 8; ++g) {
            double J[3][3], A[3][3], ec[6];
            jacobian(cXs, gauss(g, 0), gauss(g, 1), gauss(g, 2), J);
            double idg;
            w[g] = inverse3(J, A, &idg);
            sc[g] = det0 * idg;
            covariantStrain(J, cd, g, smp[(g >> 2) & 1], ec);
            double e6[6];
            push(A, ec, e6);
            for (int i = 0; i < 6; ++i) eu[(6 * g + i) * euStride] = e6[i];
        }
    }
    double inv[7];
    for (int k = 0; k < 7; ++k) {
        const int r = modeRow(k);
        double c2 = 0.0, s = 0.0;
        for (int i = 0; i < 6; ++i) c2 += t0[r][i] * t0[r][i];
        for (int g = 0; g < 8; ++g) {
            const double mk = monomial(k, gauss(g, 0), gauss(g, 1), gauss(g, 2));
            s += w[g] * sc[g] * sc[g] * mk * mk;
        }
        inv[k] = c2 * s > 0.0 ? invSqrt(c2 * s) : 1.0;
    }
    // Column k of G at point g is c_k(g) T0[:, r_k], c_k(g) = (det J0/det J)_g m_k(g) / |mode k|:
    // every column is a multiple of one of four vectors, so G alpha needs four sums and
    // G^T C G needs only the 4 x 4 form T0[:, r]^T C T0[:, s] of each point.
    int eas = res.failedStart > 0 ? 0 : 7;
    double alpha[7];
    for (int k = 0; k < 7; ++k) alpha[k] = eas ? (alpha0 ? alpha0[k] : enh[k * stride]) : 0.0;
    auto coef = [&](int g, int k) {
        return sc[g] * monomial(k, gauss(g, 0), gauss(g, 1), gauss(g, 2)) * inv[k];
    };
    auto strainAt = [&](int g, int n, double st[6]) {
        double a4[4] = {0, 0, 0, 0};
        for (int k = 0; k < n; ++k) a4[modeRow(k)] += coef(g, k) * alpha[k];
        for (int i = 0; i < 6; ++i)
            st[i] = eu[(6 * g + i) * euStride] +
                    (a4[0] * t0[0][i] + a4[1] * t0[1][i] + a4[2] * t0[2][i] + a4[3] * t0[3][i]);
    };
    double vol = 0.0;
    for (int g = 0; g < 8; ++g) vol += w[g];
    const double scale = m.yield * vol;
    // The commit at the final alpha: history, energies, and the force through the same
    // rows, conjugate to the basis coefficients (F), the assumed rows through their
    // samples. With onlyElastic it gives up (false) at the first point that yields,
    // before writing anything: a point that does not yield keeps its history bit for bit,
    // so the points before it have changed nothing.
    double F[21], cs[2][8];
    auto commit = [&](bool onlyElastic) -> bool {
        for (int j = 0; j < 21; ++j) F[j] = 0.0;
        for (int l = 0; l < 2; ++l)
            for (int s = 0; s < 8; ++s) cs[l][s] = 0.0;
        double dsum = 0.0, rsum = 0.0, usum = 0.0;
        int failed = 0;
        // the points level by level, so each level's assumed-row sums cs[lv] are fixed registers
#pragma unroll
        for (int lv = 0; lv < 2; ++lv)
#pragma unroll 1
        for (int gq = 0; gq < 4; ++gq) {
            const int g = 4 * lv + gq;
            double st[6], h0[8], h[8], sig[6], n6[6], dg, th, thb, rel;
            strainAt(g, eas, st);
            loadPoint(hist, g, stride, hmask, h0);
            const int fl = pointReturn(m, st, h0, h, sig, dg, th, thb, n6, rel);
            if (onlyElastic && fl != 0) return false;
            if (fl & 1)  // only a yielding point changes its history
                for (int k = 0; k < 8; ++k) hist[(8 * g + k) * stride] = h[k];
            failed += (fl & 4) != 0;
            dsum += w[g] * dg;
            rsum += w[g] * rel;
            if (wantStrain) {
                const double pm = (sig[0] + sig[1] + sig[2]) * (1.0 / 3.0);
                double u = 0.0;
                for (int i = 0; i < 3; ++i) u += 0.5 * sig[i] * ((sig[i] - pm) * m.i2mu + pm * m.i3k);
                for (int i = 3; i < 6; ++i) u += 0.5 * sig[i] * (sig[i] * m.imu);
                usum += w[g] * u;
            }
            if (fl & 4) continue;  // a failed point carries no stress
            const double xi = gauss(g, 0), eta = gauss(g, 1), zeta = gauss(g, 2);
            double J[3][3], A[3][3], tau[6];
            jacobian(cX, xi, eta, zeta, J);
            inverse3(J, A);
            pull(A, sig, tau);
            // rows E11, E22, 2E12 at the point: tau_0 g_0 + tau_3 g_1 against d phi / d xi and
            // tau_1 g_1 + tau_3 g_0 against d phi / d eta
            double p0[3], p1[3];
            for (int i = 0; i < 3; ++i) {
                p0[i] = w[g] * (tau[0] * J[i][0] + tau[3] * J[i][1]);
                p1[i] = w[g] * (tau[1] * J[i][1] + tau[3] * J[i][0]);
            }
            addGrad(F, p0, 0, xi, eta, zeta);
            addGrad(F, p1, 1, xi, eta, zeta);
            cs[lv][0] += w[g] * tau[4] * 0.5 * (1 - xi);
            cs[lv][1] += w[g] * tau[4] * 0.5 * (1 + xi);
            cs[lv][2] += w[g] * tau[5] * 0.5 * (1 - eta);
            cs[lv][3] += w[g] * tau[5] * 0.5 * (1 + eta);
            for (int s = 0; s < 4; ++s)
                cs[lv][4 + s] += w[g] * tau[2] * 0.25 * (1 + xi * kCorner[s][0]) * (1 + eta * kCorner[s][1]);
        }
        res.failedEnd = failed;
        res.dissipation = dsum;
        res.released = rsum;
        res.strain = usum;
        return true;
    };
    // The enhanced Newton (and, when a point fails in it, the pass without enhanced modes),
    // then the commit. When every point answered elastically the residual is affine in
    // alpha and the first correction zeroes it: the commit is tried at once, and when a
    // point yields there the Newton goes on from that alpha as it would have.
    int iter = 0;
    for (;;) {
        bool died = false, tryFast = false;
        for (;;) {
            double r[7] = {0, 0, 0, 0, 0, 0, 0}, kaa[28];
            for (int k = 0; k < 28; ++k) kaa[k] = 0.0;
            died = false;
            bool linear = true;
            // moments of w sc^2 over the points that answered elastically: mo[b] is the sum of
            // w sc^2 chi_b(g), chi_b(g) the product of point g's coordinate signs over the axes in b
            double mo[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            bool anyElastic = false;
#pragma unroll 1
            for (int g = 0; g < 8; ++g) {
                double st[6], h0[8], h[8], sig[6], n6[6], dg, th, thb, rel;
                strainAt(g, eas, st);
                loadPoint(hist, g, stride, hmask, h0);
                const int fl = pointReturn(m, st, h0, h, sig, dg, th, thb, n6, rel);
                died = died || (fl & 2);
                linear = linear && fl == 0;
                if (!eas) continue;
                double ts[4];
                for (int q = 0; q < 4; ++q) {
                    double t = 0.0;
                    for (int i = 0; i < 6; ++i) t += t0[q][i] * sig[i];
                    ts[q] = t;
                }
                double cg[7];
                for (int k = 0; k < 7; ++k) {
                    cg[k] = coef(g, k);
                    r[k] += w[g] * cg[k] * ts[modeRow(k)];
                }
                if (fl == 0) {
                    anyElastic = true;
                    const double sg = w[g] * sc[g] * sc[g];
#pragma unroll
                    for (int b = 0; b < 8; ++b) mo[b] += __builtin_popcount(b & ~g) & 1 ? -sg : sg;
                } else if (!(fl & 4)) {
                    // a yielding point's own algorithmic tangent (a failed point adds none)
                    double nv[4], M[10];
                    for (int q = 0; q < 4; ++q) {
                        double u = 0.0;
                        for (int i = 0; i < 6; ++i) u += t0[q][i] * n6[i];
                        nv[q] = u;
                    }
                    const double cm = m.mu * th, cn = 2.0 * m.mu * thb;
                    for (int p = 0; p < 4; ++p)
                        for (int q = 0; q <= p; ++q) {
                            const int i = p * (p + 1) / 2 + q;
                            M[i] = m.kappa * tr4[p] * tr4[q] + cm * D[i] - cn * nv[p] * nv[q];
                        }
                    for (int p = 0; p < 7; ++p)
                        for (int q = 0; q <= p; ++q) {
                            const int a = modeRow(p), b = modeRow(q);
                            const double mpq = a >= b ? M[a * (a + 1) / 2 + b] : M[b * (b + 1) / 2 + a];
                            kaa[p * (p + 1) / 2 + q] += w[g] * cg[p] * cg[q] * mpq;
                        }
                }
            }
            if (eas && anyElastic) {
                // The elastic points' share of Kaa. Their tangent is the element's elastic form
                // Me = kappa tr tr^T + mu D, so they add inv_p inv_q Me_ab sum_g w sc^2 m_p m_q,
                // and at a Gauss point m_p m_q = q^(deg p + deg q) chi_(e_p ^ e_q)(g).
                const double qq = 0.57735026918962576451 * 0.57735026918962576451;
                const double qn[3] = {qq, qq * 0.57735026918962576451, qq * qq};  // q^2, q^3, q^4
#pragma unroll
                for (int p = 0; p < 7; ++p)
#pragma unroll
                    for (int q = 0; q <= p; ++q) {
                        const int a = modeRow(p), b = modeRow(q);
                        const int ab = a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
                        const double me = m.kappa * tr4[a] * tr4[b] + m.mu * D[ab];
                        kaa[p * (p + 1) / 2 + q] +=
                            inv[p] * inv[q] * (qn[modeDegree(p) + modeDegree(q) - 2] * me) * mo[modeAxes(p) ^ modeAxes(q)];
                    }
            }
            if (!eas) break;
            double nrm = 0.0;
            for (int k = 0; k < 7; ++k) nrm += r[k] * r[k];
            const double tol = 1e-12 * scale;
            if (nrm <= tol * tol) break;
            if (iter + 1 >= 40) break;
            // Cholesky with the pivots' reciprocals (one division per column)
            double L[28], id[7];
            bool ok = true;
            for (int i = 0; i < 7; ++i)
                for (int j = 0; j <= i; ++j) {
                    double t = kaa[i * (i + 1) / 2 + j];
                    for (int k = 0; k < j; ++k) t -= L[i * (i + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
                    if (i == j) {
                        ok = ok && t > 0.0;
                        const double tt = fmax(t, 1e-300), y = invSqrt(tt);
                        L[i * (i + 1) / 2 + i] = tt * y;
                        id[i] = y;
                    } else {
                        L[i * (i + 1) / 2 + j] = t * id[j];
                    }
                }
            if (!ok) break;
            double del[7];
            for (int i = 0; i < 7; ++i) {
                double t = -r[i];
                for (int k = 0; k < i; ++k) t -= L[i * (i + 1) / 2 + k] * del[k];
                del[i] = t * id[i];
            }
            for (int i = 6; i >= 0; --i) {
                double t = del[i];
                for (int k = i + 1; k < 7; ++k) t -= L[k * (k + 1) / 2 + i] * del[k];
                del[i] = t * id[i];
            }
            double work = 0.0;
            for (int k = 0; k < 7; ++k) work += del[k] * r[k];
            if (fabs(work) <= 1e-16 * scale) break;
            for (int k = 0; k < 7; ++k) alpha[k] += del[k];
            ++iter;
            if (linear && iter == 1) {
                tryFast = true;
                break;
            }
        }
        if (!tryFast && died && eas) {
            // a point failed in this step: rerun it without the enhanced modes
            eas = 0;
            for (int k = 0; k < 7; ++k) alpha[k] = 0.0;
            continue;
        }
        if (commit(tryFast)) break;
    }
    // the assumed rows through their samples
#pragma unroll 1
    for (int lv = 0; lv < 2; ++lv) {
        const double zeta = lv ? 0.57735026918962576451 : -0.57735026918962576451;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const double xi = kShear[s][0], eta = kShear[s][1];
            // 2E23 (s < 2) couples g_1 with d,zeta and g_2 with d,eta; 2E31 couples g_2 with
            // d,xi and g_0 with d,zeta
            const int l = s < 2 ? 1 : 0;
            double ga[3], gz[3], pa[3], pz[3];
            column(cX, l, xi, eta, zeta, ga);
            column(cX, 2, xi, eta, zeta, gz);
            for (int i = 0; i < 3; ++i) {
                pz[i] = cs[lv][s] * ga[i];
                pa[i] = cs[lv][s] * gz[i];
            }
            addGrad(F, pz, 2, xi, eta, zeta);
            addGrad(F, pa, l, xi, eta, zeta);
            // E33 at the corner sample
            column(cX, 2, kCorner[s][0], kCorner[s][1], zeta, gz);
            for (int i = 0; i < 3; ++i) pz[i] = cs[lv][4 + s] * gz[i];
            addGrad(F, pz, 2, kCorner[s][0], kCorner[s][1], zeta);
        }
    }
    // back to corners 1..7
#pragma unroll
    for (int a = 1; a < 8; ++a)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double t = 0.0;
#pragma unroll
            for (int mm = 0; mm < 7; ++mm) t += cornerSign(mm, a) * F[3 * mm + i];
            f[3 * (a - 1) + i] = 0.125 * t;
        }
    for (int k = 0; k < 7; ++k) enh[k * stride] = alpha[k];
    return res;
}

}  // namespace shipfem::hexp
