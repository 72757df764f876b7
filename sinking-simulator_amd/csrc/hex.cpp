// hex.cpp -- the solid-shell hexahedron in FP64 on the host: rest forms, condensed
// elastic stiffness, lumped mass, stable step and the co-rotational force.
//
// The element is sim::solidshell's (engine/sim/solid_shell.hpp:111-194, solid_shell.cpp):
// an 8-node trilinear hex whose zeta axis is the shell's thickness, with
//   - assumed natural transverse shear (Dvorkin-Bathe: 2E23 sampled at (xi, eta) =
//     (-1, 0), (+1, 0) and interpolated linearly in xi, 2E31 at (0, -1), (0, +1) and
//     linear in eta), and assumed natural thickness strain (Betsch-Stein: E33 sampled
//     at the four in-plane corners and interpolated bilinearly), both at the Gauss
//     point's own zeta (solid_shell.cpp:181-235);
//   - seven enhanced assumed strains (xi on E11, eta on E22, zeta, xi zeta, eta zeta on
//     E33, xi and eta on 2E12), pushed forward with the centre Jacobian and scaled by
//     det J0 / det J, each mode normalised in the weighted L2 norm over the Gauss
//     points (solid_shell.cpp:160-179, 313-365);
//   - 2x2x2 Gauss, and static condensation K = Kuu - Kua Kaa^-1 Kua^T by Cholesky
//     (solid_shell.cpp:384-458, 500-526).
// The same steps are kept so the stiffness is the reference's to rounding; the code is
// this library's own. Mass is the consistent row sum rho int N_a dV (solid_shell.cpp:
// 655-663); the force is f = -R K (R^T x - X) with R the polar factor of the deformation
// gradient at the element centre (solid_shell.cpp:680-766).
#include "hex_plastic.hpp"
#include "shipfem.h"
#include "shipfem_host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace shipfem {
namespace {

constexpr int kN = 8, kD = 24, kE = 7, kG = 8;
constexpr double kXi[kN] = {-1, 1, 1, -1, -1, 1, 1, -1};
constexpr double kEta[kN] = {-1, -1, 1, 1, -1, -1, 1, 1};
constexpr double kZeta[kN] = {-1, -1, -1, -1, 1, 1, 1, 1};

struct Natural {
    double dN[kN][3];  // d N_a / d(xi, eta, zeta)
    double J[3][3];    // J[i][k] = d x_i / d xi_k
};

void natural(const double* x, double xi, double eta, double zeta, Natural& p) {
    for (int a = 0; a < kN; ++a) {
        const double sx = 1 + xi * kXi[a], sy = 1 + eta * kEta[a], sz = 1 + zeta * kZeta[a];
        p.dN[a][0] = 0.125 * kXi[a] * sy * sz;
        p.dN[a][1] = 0.125 * kEta[a] * sx * sz;
        p.dN[a][2] = 0.125 * kZeta[a] * sx * sy;
    }
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int a = 0; a < kN; ++a) s += p.dN[a][k] * x[3 * a + i];
            p.J[i][k] = s;
        }
}

double inverse3(const double m[3][3], double r[3][3]) {
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    if (det == 0.0) return 0.0;
    const double id = 1.0 / det;
    r[0][0] = c00 * id;
    r[1][0] = c01 * id;
    r[2][0] = c02 * id;
    r[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
    r[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
    r[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
    r[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
    r[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
    r[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    return det;
}

// Covariant strain rows [E11, E22, E33, 2E12, 2E23, 2E31] per DOF 3a + i:
// E_kl = (g_k . u,l + g_l . u,k) / 2 with g_k the k-th column of J.
void covariantB(const Natural& p, double b[6][kD]) {
    for (int a = 0; a < kN; ++a)
        for (int i = 0; i < 3; ++i) {
            const int c = 3 * a + i;
            const double* n = p.dN[a];
            const double g0 = p.J[i][0], g1 = p.J[i][1], g2 = p.J[i][2];
            b[0][c] = n[0] * g0;
            b[1][c] = n[1] * g1;
            b[2][c] = n[2] * g2;
            b[3][c] = n[1] * g0 + n[0] * g1;
            b[4][c] = n[2] * g1 + n[1] * g2;
            b[5][c] = n[0] * g2 + n[2] * g0;
        }
}

// Voigt map of eps = A^T E A (A = J^-1) from covariant to Cartesian strain; column j is
// the image of the j-th unit covariant strain (engineering shears halve into the tensor).
void toCartesian(const double A[3][3], double t[6][6]) {
    constexpr int r_[6] = {0, 1, 2, 0, 1, 2}, c_[6] = {0, 1, 2, 1, 2, 0};
    for (int j = 0; j < 6; ++j) {
        const double v = j < 3 ? 1.0 : 0.5;
        double e[3][3] = {};
        e[r_[j]][c_[j]] = v;
        e[c_[j]][r_[j]] = v;
        double m[3][3];
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double s = 0;
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < 3; ++l) s += A[k][r] * e[k][l] * A[l][c];
                m[r][c] = s;
            }
        t[0][j] = m[0][0];
        t[1][j] = m[1][1];
        t[2][j] = m[2][2];
        t[3][j] = 2 * m[0][1];
        t[4][j] = 2 * m[1][2];
        t[5][j] = 2 * m[2][0];
    }
}

struct Forms {
    double B[kG][6][kD];
    double G[kG][6][kE];
    double w[kG];
    double Jinv0[3][3];
    bool ok = false;
};

bool restForms(const double* x, Forms& f) {
    Natural c;
    natural(x, 0, 0, 0, c);
    const double det0 = inverse3(c.J, f.Jinv0);
    if (!(det0 > 0)) return false;
    double t0[6][6];
    toCartesian(f.Jinv0, t0);
    const double q = 1.0 / std::sqrt(3.0);
    // assumed-strain samples at the two zeta levels of the rule
    constexpr double shearAt[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
    constexpr double cornerAt[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    double shear[2][4][kD], thick[2][4][kD];
    for (int lv = 0; lv < 2; ++lv) {
        const double zeta = lv ? q : -q;
        for (int s = 0; s < 4; ++s) {
            double b[6][kD];
            Natural p;
            natural(x, shearAt[s][0], shearAt[s][1], zeta, p);
            covariantB(p, b);
            std::memcpy(shear[lv][s], b[s < 2 ? 4 : 5], sizeof(double) * kD);
            natural(x, cornerAt[s][0], cornerAt[s][1], zeta, p);
            covariantB(p, b);
            std::memcpy(thick[lv][s], b[2], sizeof(double) * kD);
        }
    }
    for (int g = 0; g < kG; ++g) {
        const double xi = (g & 1) ? q : -q, eta = (g & 2) ? q : -q, zeta = (g & 4) ? q : -q;
        const int lv = (g & 4) ? 1 : 0;
        Natural p;
        natural(x, xi, eta, zeta, p);
        double A[3][3];
        const double det = inverse3(p.J, A);
        if (!(det > 0)) return false;
        f.w[g] = det;
        double b[6][kD];
        covariantB(p, b);
        for (int c = 0; c < kD; ++c) {
            b[4][c] = 0.5 * (1 - xi) * shear[lv][0][c] + 0.5 * (1 + xi) * shear[lv][1][c];
            b[5][c] = 0.5 * (1 - eta) * shear[lv][2][c] + 0.5 * (1 + eta) * shear[lv][3][c];
            double e33 = 0;
            for (int s = 0; s < 4; ++s)
                e33 += 0.25 * (1 + xi * cornerAt[s][0]) * (1 + eta * cornerAt[s][1]) * thick[lv][s][c];
            b[2][c] = e33;
        }
        double t[6][6];
        toCartesian(A, t);
        for (int i = 0; i < 6; ++i)
            for (int c = 0; c < kD; ++c) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += t[i][k] * b[k][c];
                f.B[g][i][c] = s;
            }
        double m[6][kE] = {};
        m[0][0] = xi;
        m[1][1] = eta;
        m[2][2] = zeta;
        m[2][3] = xi * zeta;
        m[2][4] = eta * zeta;
        m[3][5] = xi;
        m[3][6] = eta;
        const double scale = det0 / det;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < kE; ++j) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += t0[i][k] * m[k][j];
                f.G[g][i][j] = scale * s;
            }
    }
    for (int j = 0; j < kE; ++j) {  // each enhanced mode to unit weighted L2 norm
        double n2 = 0;
        for (int g = 0; g < kG; ++g)
            for (int i = 0; i < 6; ++i) n2 += f.w[g] * f.G[g][i][j] * f.G[g][i][j];
        if (!(n2 > 0)) continue;
        const double s = 1.0 / std::sqrt(n2);
        for (int g = 0; g < kG; ++g)
            for (int i = 0; i < 6; ++i) f.G[g][i][j] *= s;
    }
    f.ok = true;
    return true;
}

// In-place Cholesky solve of Kaa X = R for `cols` right-hand sides (rows of R).
bool choleskySolve(const double a[kE][kE], double r[kE][kD]) {
    double l[kE][kE] = {};
    for (int i = 0; i < kE; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = a[i][j];
            for (int k = 0; k < j; ++k) s -= l[i][k] * l[j][k];
            if (i == j) {
                if (!(s > 0)) return false;
                l[i][i] = std::sqrt(s);
            } else {
                l[i][j] = s / l[j][j];
            }
        }
    for (int c = 0; c < kD; ++c) {
        for (int i = 0; i < kE; ++i) {
            double s = r[i][c];
            for (int k = 0; k < i; ++k) s -= l[i][k] * r[k][c];
            r[i][c] = s / l[i][i];
        }
        for (int i = kE - 1; i >= 0; --i) {
            double s = r[i][c];
            for (int k = i + 1; k < kE; ++k) s -= l[k][i] * r[k][c];
            r[i][c] = s / l[i][i];
        }
    }
    return true;
}

}  // namespace

bool hexStiffness(const double* x, double E, double nu, double K[kD * kD]) {
    std::fill(K, K + kD * kD, 0.0);
    Forms f;
    if (!restForms(x, f)) return false;
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    double C[6][6] = {};
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) C[i][j] = lam;
        C[i][i] += 2 * mu;
        C[3 + i][3 + i] = mu;
    }
    double kua[kD][kE] = {}, kaa[kE][kE] = {};
    for (int g = 0; g < kG; ++g) {
        double CB[6][kD], CG[6][kE];
        for (int i = 0; i < 6; ++i) {
            for (int c = 0; c < kD; ++c) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += C[i][k] * f.B[g][k][c];
                CB[i][c] = s;
            }
            for (int j = 0; j < kE; ++j) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += C[i][k] * f.G[g][k][j];
                CG[i][j] = s;
            }
        }
        const double w = f.w[g];
        for (int r = 0; r < kD; ++r) {
            for (int c = 0; c < kD; ++c) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += f.B[g][k][r] * CB[k][c];
                K[r * kD + c] += w * s;
            }
            for (int j = 0; j < kE; ++j) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += f.B[g][k][r] * CG[k][j];
                kua[r][j] += w * s;
            }
        }
        for (int i = 0; i < kE; ++i)
            for (int j = 0; j < kE; ++j) {
                double s = 0;
                for (int k = 0; k < 6; ++k) s += f.G[g][k][i] * CG[k][j];
                kaa[i][j] += w * s;
            }
    }
    double x_[kE][kD];  // Kaa^-1 Kua^T
    for (int i = 0; i < kE; ++i)
        for (int c = 0; c < kD; ++c) x_[i][c] = kua[c][i];
    if (!choleskySolve(kaa, x_)) return false;
    for (int r = 0; r < kD; ++r)
        for (int c = 0; c < kD; ++c) {
            double s = 0;
            for (int k = 0; k < kE; ++k) s += kua[r][k] * x_[k][c];
            K[r * kD + c] -= s;
        }
    return true;
}

bool hexMass(const double* x, double rho, double m[kN]) {
    std::fill(m, m + kN, 0.0);
    const double q = 1.0 / std::sqrt(3.0);
    for (int g = 0; g < kG; ++g) {
        const double xi = (g & 1) ? q : -q, eta = (g & 2) ? q : -q, zeta = (g & 4) ? q : -q;
        Natural p;
        natural(x, xi, eta, zeta, p);
        double A[3][3];
        const double det = inverse3(p.J, A);
        if (!(det > 0)) return false;
        for (int a = 0; a < kN; ++a)
            m[a] += rho * det * 0.125 * (1 + xi * kXi[a]) * (1 + eta * kEta[a]) * (1 + zeta * kZeta[a]);
    }
    return true;
}

bool hexCentreGradient(const double* x, double b0[kN][3]) {
    Natural c;
    natural(x, 0, 0, 0, c);
    double A[3][3];
    if (!(inverse3(c.J, A) > 0)) return false;
    for (int a = 0; a < kN; ++a)
        for (int j = 0; j < 3; ++j) b0[a][j] = c.dN[a][0] * A[0][j] + c.dN[a][1] * A[1][j] + c.dN[a][2] * A[2][j];
    return true;
}

// solidshell::elementSize (solid_shell.cpp:770-806): the in-plane size is the square root
// of the mid-surface quad's area (two triangles, exact on a warped element), the thickness
// the element's volume (its Gauss volumes) over that area. Zeros when either is not positive.
void hexElementSize(const double* x, double& inPlane, double& thickness) {
    inPlane = thickness = 0.0;
    double mid[4][3];
    for (int a = 0; a < 4; ++a)
        for (int i = 0; i < 3; ++i) mid[a][i] = 0.5 * (x[3 * a + i] + x[3 * (a + 4) + i]);
    double area = 0.0;
    for (int t = 1; t <= 2; ++t) {
        double e1[3], e2[3];
        for (int i = 0; i < 3; ++i) {
            e1[i] = mid[t][i] - mid[0][i];
            e2[i] = mid[t + 1][i] - mid[0][i];
        }
        const double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2],
                     c2 = e1[0] * e2[1] - e1[1] * e2[0];
        area += 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    }
    if (!(area > 0.0)) return;
    double m[kN];
    if (!hexMass(x, 1.0, m)) return;
    double volume = 0.0;
    for (int a = 0; a < kN; ++a) volume += m[a];
    inPlane = std::sqrt(area);
    thickness = volume / area;
}

// plasticity::regularisedFailureStrain (plasticity.cpp:158-169) at an element's own size.
double hexFailureStrain(const double* x, double uniform, double fracture) {
    double l = 0.0, t = 0.0;
    hexElementSize(x, l, t);
    const double share = (l > 0.0 && t > 0.0) ? std::min(1.0, t / l) : 1.0;
    return uniform + (fracture - uniform) * share;
}

double hexOmegaMax(const double K[kD * kD], const double m[kN]) {
    // largest omega^2 of K v = omega^2 M v with M the element's lumped mass, by power
    // iteration on M^-1/2 K M^-1/2 from a fixed start, to a relative 1e-12 change
    double s[kD], v[kD], y[kD];
    for (int i = 0; i < kD; ++i) {
        s[i] = 1.0 / std::sqrt(m[i / 3]);
        v[i] = 1.0 + 0.1 * i;
    }
    double lam = 0;
    for (int it = 0; it < 500; ++it) {
        double n = 0;
        for (int i = 0; i < kD; ++i) n += v[i] * v[i];
        n = std::sqrt(n);
        for (int i = 0; i < kD; ++i) v[i] /= n;
        for (int r = 0; r < kD; ++r) {
            double a = 0;
            for (int c = 0; c < kD; ++c) a += K[r * kD + c] * s[c] * v[c];
            y[r] = s[r] * a;
        }
        double l = 0;
        for (int i = 0; i < kD; ++i) l += v[i] * y[i];
        std::memcpy(v, y, sizeof v);
        if (it > 8 && std::fabs(l - lam) <= 1e-12 * std::fabs(l)) {
            lam = l;
            break;
        }
        lam = l;
    }
    return std::sqrt(std::max(lam, 0.0));
}

bool hexRotation(const double* X, const double* x, double R[3][3]) {
    // R: polar factor of F = J(x) J(X)^-1 at the centre, Newton-Higham to convergence
    Natural cx, cX;
    natural(x, 0, 0, 0, cx);
    natural(X, 0, 0, 0, cX);
    double A[3][3];
    if (inverse3(cX.J, A) == 0.0) {
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R[i][j] = i == j ? 1.0 : 0.0;
        return false;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) R[i][j] = cx.J[i][0] * A[0][j] + cx.J[i][1] * A[1][j] + cx.J[i][2] * A[2][j];
    for (int it = 0; it < 64; ++it) {
        double inv[3][3];
        if (inverse3(R, inv) == 0.0) break;
        double change = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double nx = 0.5 * (R[i][j] + inv[j][i]);
                change = std::max(change, std::fabs(nx - R[i][j]));
                R[i][j] = nx;
            }
        if (change < 1e-16) break;
    }
    return true;
}

void hexForce(const double K[kD * kD], const double* X, const double* x, double out[kD]) {
    double R[3][3];
    hexRotation(X, x, R);
    double d[kD], kd[kD];
    for (int a = 0; a < kN; ++a)
        for (int i = 0; i < 3; ++i)
            d[3 * a + i] = R[0][i] * x[3 * a] + R[1][i] * x[3 * a + 1] + R[2][i] * x[3 * a + 2] - X[3 * a + i];
    for (int r = 0; r < kD; ++r) {
        double s = 0;
        for (int c = 0; c < kD; ++c) s += K[r * kD + c] * d[c];
        kd[r] = s;
    }
    for (int a = 0; a < kN; ++a)
        for (int i = 0; i < 3; ++i)
            out[3 * a + i] = -(R[i][0] * kd[3 * a] + R[i][1] * kd[3 * a + 1] + R[i][2] * kd[3 * a + 2]);
}

// The device's per-element data of a hex case. Every element's stiffness, mass and
// frequency is independent of the others, so the loop splits over host threads; the
// nodal masses are summed afterwards in element order, so they do not depend on the
// thread count.
std::string hexRestState(const shipfem_case_desc& d, HostCase& hc, std::vector<double>& mass) {
    const uint32_t T = hc.T;
    const uint64_t tp = hc.Tpad;
    const bool plastic = d.material.plastic != 0;
    if (plastic) {
        hc.hgeod.assign(kHexGeo * tp, 0.0);
    } else {
        hc.hk.assign(kHexPacked * tp, 0.0);
    }
    hc.hgeo.assign(kHexGeo * tp, 0.0f);
    // per-element failure strain, resolved from each element's size (initialisePlasticState,
    // solid_shell.cpp:808-815), when the case gives a local fracture strain
    const bool perElement = plastic && d.material.failure_strain > 0.0 && d.fracture_strain > 0.0;
    if (perElement) hc.hfail.assign(tp, 0.0);
    std::vector<double> em(8ull * T), omega(T, 0.0), volume(T, 0.0);
    std::vector<int> bad(T, 0);
    const double* X = d.rest_position;
    auto work = [&](uint32_t e0, uint32_t e1) {
        double x[kD], K[kD * kD], m[kN], b0[kN][3];
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t* v = &d.index[8ull * e];
            for (int a = 0; a < kN; ++a)
                for (int i = 0; i < 3; ++i) x[3 * a + i] = X[3ull * v[a] + i];
            if (!hexStiffness(x, hc.E, hc.nu, K) || !hexMass(x, hc.rho, m) ||
                !hexCentreGradient(x, b0)) {
                bad[e] = 1;
                continue;
            }
            for (int a = 0; a < kN; ++a) {
                em[8ull * e + a] = m[a];
                volume[e] += m[a];
            }
            volume[e] /= hc.rho;
            omega[e] = hexOmegaMax(K, m);
            if (perElement) hc.hfail[e] = hexFailureStrain(x, d.material.failure_strain, d.fracture_strain);
            // corners 1..7: rows and columns 3..23 of K
            int p = 0;
            for (int i = 0; i < kHexDof && !plastic; ++i)
                for (int j = i; j < kHexDof; ++j) hc.hk[(p++) * tp + e] = K[(3 + i) * kD + 3 + j];
            double off[kHexDof], cX[kHexDof];
            for (int a = 1; a < kN; ++a)
                for (int i = 0; i < 3; ++i) {
                    off[3 * (a - 1) + i] = x[3 * a + i] - x[i];
                    hc.hgeo[(3 * (a - 1) + i) * tp + e] = static_cast<float>(b0[a][i]);
                    hc.hgeo[(kHexDof + 3 * (a - 1) + i) * tp + e] = static_cast<float>(off[3 * (a - 1) + i]);
                }
            if (plastic) {
                // the plastic pass reads the rest geometry as its basis coefficients
                hexp::basis(off, cX);
                for (int j = 0; j < kHexDof; ++j) {
                    hc.hgeod[j * tp + e] = b0[1 + j / 3][j % 3];
                    hc.hgeod[(kHexDof + j) * tp + e] = cX[j];
                }
            }
        }
    };
    const uint32_t threads =
        std::max(1u, std::min<uint32_t>(std::thread::hardware_concurrency(), std::min(16u, T / 256 + 1)));
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t)
        pool.emplace_back(work, static_cast<uint32_t>(uint64_t(T) * t / threads),
                          static_cast<uint32_t>(uint64_t(T) * (t + 1) / threads));
    for (auto& th : pool) th.join();
    hc.hex_omega = 0;
    hc.hvol = volume;
    for (uint32_t e = 0; e < T; ++e) {
        if (bad[e]) return "hex element " + std::to_string(e) + " is degenerate or inverted";
        hc.hex_omega = std::max(hc.hex_omega, omega[e]);
        hc.total_volume += volume[e];
        for (int a = 0; a < kN; ++a) mass[d.index[8ull * e + a]] += em[8ull * e + a];
    }
    return {};
}

}  // namespace shipfem

extern "C" {

shipfem_status shipfem_hex_stiffness(const double nodes[24], double youngs_modulus,
                                     double poisson_ratio, double out[576]) {
    if (!nodes || !out || !(youngs_modulus > 0) || !(poisson_ratio > -1 && poisson_ratio < 0.5))
        return SHIPFEM_INVALID_ARG;
    return shipfem::hexStiffness(nodes, youngs_modulus, poisson_ratio, out) ? SHIPFEM_OK
                                                                           : SHIPFEM_INVALID_ARG;
}

shipfem_status shipfem_hex_failure_strain(const double nodes[24], double uniform, double fracture,
                                          double* out) {
    if (!nodes || !out || !std::isfinite(uniform) || !std::isfinite(fracture)) return SHIPFEM_INVALID_ARG;
    *out = shipfem::hexFailureStrain(nodes, uniform, fracture);
    return SHIPFEM_OK;
}

shipfem_status shipfem_hex_mass(const double nodes[24], double density, double out[8]) {
    if (!nodes || !out || !(density > 0)) return SHIPFEM_INVALID_ARG;
    return shipfem::hexMass(nodes, density, out) ? SHIPFEM_OK : SHIPFEM_INVALID_ARG;
}

shipfem_status shipfem_hex_critical_timestep(const double nodes[24], double youngs_modulus,
                                            double poisson_ratio, double density, double safety,
                                            double* dt) {
    if (!nodes || !dt || !(density > 0) || !(safety > 0)) return SHIPFEM_INVALID_ARG;
    double K[576], m[8];
    const shipfem_status st = shipfem_hex_stiffness(nodes, youngs_modulus, poisson_ratio, K);
    if (st != SHIPFEM_OK) return st;
    if (!shipfem::hexMass(nodes, density, m)) return SHIPFEM_INVALID_ARG;
    const double w = shipfem::hexOmegaMax(K, m);
    if (!(w > 0)) return SHIPFEM_INVALID_ARG;
    *dt = safety * 2.0 / w;
    return SHIPFEM_OK;
}

shipfem_status shipfem_hex_plate_mesh(double lx, double ly, double thickness, int32_t nx,
                                      int32_t ny, int32_t nz, uint32_t* node_count,
                                      uint32_t* element_count, double* rest_position,
                                      uint32_t* index) {
    if (!(lx > 0) || !(ly > 0) || !(thickness > 0) || nx < 1 || ny < 1 || nz < 1)
        return SHIPFEM_INVALID_ARG;
    const uint64_t sx = nx + 1ull, sy = ny + 1ull, sz = nz + 1ull;
    if (sx * sy * sz > 0xFFFFFFF0ull || uint64_t(nx) * ny * nz > 0x1FFFFFF0ull) return SHIPFEM_INVALID_ARG;
    if (node_count) *node_count = static_cast<uint32_t>(sx * sy * sz);
    if (element_count) *element_count = static_cast<uint32_t>(uint64_t(nx) * ny * nz);
    // node (i, j, k) with the thickness index fastest, then y, then x; mid-surface z = 0
    auto at = [&](uint64_t i, uint64_t j, uint64_t k) { return static_cast<uint32_t>((i * sy + j) * sz + k); };
    if (rest_position)
        for (uint64_t i = 0; i < sx; ++i)
            for (uint64_t j = 0; j < sy; ++j)
                for (uint64_t k = 0; k < sz; ++k) {
                    double* x = rest_position + 3ull * at(i, j, k);
                    x[0] = lx * static_cast<double>(i) / nx;
                    x[1] = ly * static_cast<double>(j) / ny;
                    x[2] = -0.5 * thickness + thickness * static_cast<double>(k) / nz;
                }
    if (index) {
        uint32_t* q = index;
        for (uint64_t i = 0; i < uint64_t(nx); ++i)
            for (uint64_t j = 0; j < uint64_t(ny); ++j)
                for (uint64_t k = 0; k < uint64_t(nz); ++k) {
                    const uint32_t v[8] = {at(i, j, k),         at(i + 1, j, k),
                                           at(i + 1, j + 1, k), at(i, j + 1, k),
                                           at(i, j, k + 1),     at(i + 1, j, k + 1),
                                           at(i + 1, j + 1, k + 1), at(i, j + 1, k + 1)};
                    std::memcpy(q, v, sizeof v);
                    q += 8;
                }
    }
    return SHIPFEM_OK;
}

shipfem_status shipfem_hex_pressure_load(const double* rest_position, uint32_t node_count,
                                         const uint32_t* index, uint32_t element_count,
                                         double pressure, double* load) {
    if (!rest_position || !index || !load) return SHIPFEM_INVALID_ARG;
    for (uint64_t q = 0; q < 8ull * element_count; ++q)
        if (index[q] >= node_count) return SHIPFEM_INVALID_ARG;
    // a face is on the boundary when one element alone carries it (node sets compared
    // sorted, so two elements' opposite windings match)
    static constexpr int kFace[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                        {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};
    std::map<std::array<uint32_t, 4>, int> count;
    auto key = [&](uint32_t e, int f) {
        std::array<uint32_t, 4> k;
        for (int c = 0; c < 4; ++c) k[c] = index[8ull * e + kFace[f][c]];
        std::sort(k.begin(), k.end());
        return k;
    };
    for (uint32_t e = 0; e < element_count; ++e)
        for (int f = 0; f < 6; ++f) ++count[key(e, f)];
    // the +zeta face (corners 4-7, counter-clockwise seen from outside): consistent
    // load -p n dA by 2 x 2 Gauss on the bilinear face
    const double g = 1.0 / std::sqrt(3.0);
    constexpr double cs[4] = {-1, 1, 1, -1}, ct[4] = {-1, -1, 1, 1};
    for (uint32_t e = 0; e < element_count; ++e) {
        if (count[key(e, 1)] != 1) continue;
        const uint32_t* n = index + 8ull * e + 4;
        for (int p = 0; p < 4; ++p) {
            const double s = (p & 1) ? g : -g, t = (p & 2) ? g : -g;
            double xs[3] = {0, 0, 0}, xt[3] = {0, 0, 0}, N[4];
            for (int c = 0; c < 4; ++c) {
                N[c] = 0.25 * (1 + s * cs[c]) * (1 + t * ct[c]);
                const double ds = 0.25 * cs[c] * (1 + t * ct[c]), dt = 0.25 * ct[c] * (1 + s * cs[c]);
                for (int i = 0; i < 3; ++i) {
                    xs[i] += ds * rest_position[3ull * n[c] + i];
                    xt[i] += dt * rest_position[3ull * n[c] + i];
                }
            }
            const double a[3] = {xs[1] * xt[2] - xs[2] * xt[1], xs[2] * xt[0] - xs[0] * xt[2],
                                 xs[0] * xt[1] - xs[1] * xt[0]};
            for (int c = 0; c < 4; ++c)
                for (int i = 0; i < 3; ++i) load[3ull * n[c] + i] -= pressure * N[c] * a[i];
        }
    }
    return SHIPFEM_OK;
}

shipfem_status shipfem_hex_rotation(const double rest[24], const double current[24],
                                    double out[9]) {
    if (!rest || !current || !out) return SHIPFEM_INVALID_ARG;
    double R[3][3];
    const bool ok = shipfem::hexRotation(rest, current, R);
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) out[3 * c + r] = R[r][c];
    return ok ? SHIPFEM_OK : SHIPFEM_INVALID_ARG;
}

shipfem_status shipfem_hex_plastic_update(const double rest[24], const double current[24],
                                          double youngs_modulus, double poisson_ratio,
                                          double yield_strength, double hardening,
                                          double failure_strain, double sensitivity,
                                          double reference_triaxiality, double cutoff_triaxiality,
                                          double point_state[64], double enhanced[7],
                                          double force[24], double* dissipation,
                                          int32_t* failed_points) {
    if (!rest || !current || !point_state || !enhanced || !force || !(youngs_modulus > 0) ||
        !(poisson_ratio > -1 && poisson_ratio < 0.5) || !(yield_strength > 0) || !(hardening >= 0))
        return SHIPFEM_INVALID_ARG;
    double b0[8][3];
    if (!shipfem::hexCentreGradient(rest, b0)) return SHIPFEM_INVALID_ARG;
    double R[3][3];
    shipfem::hexRotation(rest, current, R);
    // corner offsets from corner 0: rest X, co-rotated displacement d = R^T (x - x_0) - X
    double X[21], d[21];
    for (int a = 1; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
            X[3 * (a - 1) + i] = rest[3 * a + i] - rest[i];
            double t = 0.0;
            for (int k = 0; k < 3; ++k) t += R[k][i] * (current[3 * a + k] - current[k]);
            d[3 * (a - 1) + i] = t - X[3 * (a - 1) + i];
        }
    const double mu = youngs_modulus / (2.0 * (1.0 + poisson_ratio));
    const double lam = youngs_modulus * poisson_ratio / ((1.0 + poisson_ratio) * (1.0 - 2.0 * poisson_ratio));
    const shipfem::hexp::Material m = shipfem::hexp::derive(
        {mu, lam + (2.0 / 3.0) * mu, yield_strength, hardening, failure_strain, sensitivity,
         reference_triaxiality, cutoff_triaxiality, failure_strain > 0.0});
    double cX[21], cd[21], fi[21], eu[48];
    shipfem::hexp::basis(X, cX);
    shipfem::hexp::basis(d, cd);
    const shipfem::hexp::Result res =
        shipfem::hexp::update(m, cX, cX, cd, point_state, enhanced, 1, false, fi, eu, 1,
                              shipfem::hexp::startOf(point_state, 1));
    // the force on the nodes, -R f, corner 0 carrying minus the others
    double f0[3] = {0.0, 0.0, 0.0};
    for (int a = 1; a < 8; ++a)
        for (int i = 0; i < 3; ++i) f0[i] -= fi[3 * (a - 1) + i];
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
            const double* v = a == 0 ? f0 : fi + 3 * (a - 1);
            force[3 * a + i] = -(R[i][0] * v[0] + R[i][1] * v[1] + R[i][2] * v[2]);
        }
    if (dissipation) *dissipation = res.dissipation;
    if (failed_points) *failed_points = res.failedEnd;
    return SHIPFEM_OK;
}

shipfem_status shipfem_hex_force(const double stiffness[576], const double rest[24],
                                 const double current[24], double out[24]) {
    if (!stiffness || !rest || !current || !out) return SHIPFEM_INVALID_ARG;
    shipfem::hexForce(stiffness, rest, current, out);
    return SHIPFEM_OK;
}

}  // extern "C"
