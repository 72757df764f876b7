// omega.hpp -- the largest eigenvalue of B^T C B of a linear tet, in FP64 (device).
//
// For the lumped-mass linear tet, K_e = V B^T C B and each corner carries rho V / 4, so the
// element frequency bound of Irons and Hughes is omega_e^2 = 4 lambda_max(B^T C B) / rho and
// the stable step is 2 / max_e omega_e. B holds the gradients of the corner shape
// functions: rows 1..3 of Dm^-1 for corners 1..3, minus their sum for corner 0.
#pragma once

#include <hip/hip_runtime.h>

namespace shipfem {

// lambda_max(B^T C B) = lambda_max(Q (B B^T) Q), Q = C^{1/2} in Voigt form with
// engineering shears: C^{1/2} = sqrt(3K) P_vol + sqrt(2 mu) P_dev on the normal
// block and sqrt(mu) on the shear block. B: Dm^-1 row-major, entry 3 * row + col.
__device__ inline double lambda_max_dminv(const double B[9], double lam, double mu) {
    double g[4][3];
#pragma unroll
    for (int cc = 1; cc < 4; ++cc)
#pragma unroll
        for (int a = 0; a < 3; ++a) g[cc][a] = B[(cc - 1) * 3 + a];
#pragma unroll
    for (int a = 0; a < 3; ++a) g[0][a] = -(g[1][a] + g[2][a] + g[3][a]);
    double M[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) M[k] = 0.0;
#pragma unroll
    for (int nd = 0; nd < 4; ++nd) {
        const double gx = g[nd][0], gy = g[nd][1], gz = g[nd][2];
        // rows of B_a: [gx,0,0] [0,gy,0] [0,0,gz] [gy,gx,0] [0,gz,gy] [gz,0,gx]
        const double Ba[6][3] = {{gx, 0, 0}, {0, gy, 0}, {0, 0, gz},
                                 {gy, gx, 0}, {0, gz, gy}, {gz, 0, gx}};
#pragma unroll
        for (int i = 0; i < 6; ++i)
#pragma unroll
            for (int j = 0; j < 6; ++j)
                M[i * 6 + j] += Ba[i][0] * Ba[j][0] + Ba[i][1] * Ba[j][1] + Ba[i][2] * Ba[j][2];
    }
    const double K3 = 3.0 * lam + 2.0 * mu;
    const double qa = sqrt(K3) / 3.0 + sqrt(2.0 * mu) * (2.0 / 3.0);
    const double qb = sqrt(K3) / 3.0 - sqrt(2.0 * mu) / 3.0;
    const double qs = sqrt(mu);
    // S = Q M Q with Q block-diagonal
    double Q[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) Q[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) Q[i * 6 + j] = i == j ? qa : qb;
        Q[(3 + i) * 6 + 3 + i] = qs;
    }
    double QM[36], S[36];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += Q[i * 6 + k] * M[k * 6 + j];
            QM[i * 6 + j] = s;
        }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) s += QM[i * 6 + k] * Q[k * 6 + j];
            S[i * 6 + j] = s;
        }
    // cyclic Jacobi, fixed sweep count
    for (int sweep = 0; sweep < 12; ++sweep) {
#pragma unroll
        for (int p = 0; p < 5; ++p)
#pragma unroll
            for (int q = p + 1; q < 6; ++q) {
                const double apq = S[p * 6 + q];
                if (apq != 0.0) {
                    const double theta = (S[q * 6 + q] - S[p * 6 + p]) / (2.0 * apq);
                    const double tt = (theta >= 0.0 ? 1.0 : -1.0) /
                                      (fabs(theta) + sqrt(theta * theta + 1.0));
                    const double cs = 1.0 / sqrt(tt * tt + 1.0), sn = tt * cs;
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        const double akp = S[k * 6 + p], akq = S[k * 6 + q];
                        S[k * 6 + p] = cs * akp - sn * akq;
                        S[k * 6 + q] = sn * akp + cs * akq;
                    }
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        const double apk = S[p * 6 + k], aqk = S[q * 6 + k];
                        S[p * 6 + k] = cs * apk - sn * aqk;
                        S[q * 6 + k] = sn * apk + cs * aqk;
                    }
                }
            }
    }
    double mx = S[0];
#pragma unroll
    for (int k = 1; k < 6; ++k) mx = fmax(mx, S[k * 6 + k]);
    return mx;
}

}  // namespace shipfem
