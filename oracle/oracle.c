/*
 * oracle.c -- FP64 CPU oracle and FP32 reference restatement (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, never by the product.
 *
 * Parity pins (what checks each function). Golden vectors from the compiled reference
 * are in test_oracle_golden.py; checks fixed by the mathematics in test_oracle_pins.py.
 *   ora_box_mesh / ora_box_mesh_f32  golden connectivity and positions of the compiled
 *                                    makeBoxTetMesh (fem.cpp:217-263)
 *   ora_rest_state                   golden restInv/volume/mass/CSR of the compiled
 *                                    computeRestState (fem.cpp:119-158); exact volume sum
 *   ora_edge_timestep                golden criticalTimestep (fem.cpp:183-185)
 *   ora_max_element_omega            eigenvalues of a stiffness assembled by central
 *                                    differences of ora_tet_force; leapfrog stable at
 *                                    0.99 of 2/omega and divergent at 1.03 (single tet)
 *   ora_tet_force                    compiled tetForces on the twisted fem_spike mesh;
 *                                    rigid rotation => zero force; zero net moment;
 *                                    patch test: uniform strain gives the consistent
 *                                    nodal loads of the exact traction at every node,
 *                                    also rotated
 *   ora_j2_update                    compiled plasticity::update on strain paths;
 *                                    pure-shear and uniaxial-strain closed forms with and
 *                                    without hardening, step-size independence on radial
 *                                    paths, hydrostatic no-yield, zero plastic trace,
 *                                    return onto the surface, idempotence
 *   ora_face_traction                total force = traction * face area (closed form)
 *   ora_run                          compiled stepCpu trajectory (FP32 reference vs FP64
 *                                    oracle to FP32 rounding); rest and free translation;
 *                                    energy closes to 1e-9 for load, drive + damping;
 *                                    settled drive stores 1/2 E A d^2/L with nu contraction;
 *                                    contact impulse = momentum, contact work closes
 *   ora_panel_mesh                   positive volumes, closed-form volume and surface area,
 *                                    conforming (no face shared by three tets), web lines
 *   ora_box_jitter / ora_uniform     faces kept planar, amplitude bound, seed dependence;
 *                                    SplitMix64 published outputs
 *   ref32_*                          bit-identical to the compiled reference
 *                                    (fem.cpp, -ffp-contract=off), 1 and N threads
 */
#define _POSIX_C_SOURCE 200809L /* pthread barriers */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================== */
/* small FP64 3x3 helpers, column-major m[col*3+row] like fem.cpp:10-16      */
/* ======================================================================== */
#define AT(m, r, c) ((m)[(c) * 3 + (r)])

static void mul3(const double* a, const double* b, double* r) {
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += AT(a, i, k) * AT(b, k, c);
            AT(r, i, c) = s;
        }
}

static void transpose3(const double* a, double* r) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) AT(r, i, j) = AT(a, j, i);
}

static double det3(const double* a) {
    return AT(a, 0, 0) * (AT(a, 1, 1) * AT(a, 2, 2) - AT(a, 1, 2) * AT(a, 2, 1)) -
           AT(a, 0, 1) * (AT(a, 1, 0) * AT(a, 2, 2) - AT(a, 1, 2) * AT(a, 2, 0)) +
           AT(a, 0, 2) * (AT(a, 1, 0) * AT(a, 2, 1) - AT(a, 1, 1) * AT(a, 2, 0));
}

/* cofactor inverse with the reference's singular guard (fem.cpp:42-56) */
static void inv3(const double* a, double* r) {
    const double det = det3(a);
    const double id = fabs(det) > 1e-30 ? 1.0 / det : 0.0;
    AT(r, 0, 0) = (AT(a, 1, 1) * AT(a, 2, 2) - AT(a, 1, 2) * AT(a, 2, 1)) * id;
    AT(r, 0, 1) = (AT(a, 0, 2) * AT(a, 2, 1) - AT(a, 0, 1) * AT(a, 2, 2)) * id;
    AT(r, 0, 2) = (AT(a, 0, 1) * AT(a, 1, 2) - AT(a, 0, 2) * AT(a, 1, 1)) * id;
    AT(r, 1, 0) = (AT(a, 1, 2) * AT(a, 2, 0) - AT(a, 1, 0) * AT(a, 2, 2)) * id;
    AT(r, 1, 1) = (AT(a, 0, 0) * AT(a, 2, 2) - AT(a, 0, 2) * AT(a, 2, 0)) * id;
    AT(r, 1, 2) = (AT(a, 0, 2) * AT(a, 1, 0) - AT(a, 0, 0) * AT(a, 1, 2)) * id;
    AT(r, 2, 0) = (AT(a, 1, 0) * AT(a, 2, 1) - AT(a, 1, 1) * AT(a, 2, 0)) * id;
    AT(r, 2, 1) = (AT(a, 0, 1) * AT(a, 2, 0) - AT(a, 0, 0) * AT(a, 2, 1)) * id;
    AT(r, 2, 2) = (AT(a, 0, 0) * AT(a, 1, 1) - AT(a, 0, 1) * AT(a, 1, 0)) * id;
}

/* Higham's Newton iteration R <- (R + R^-T)/2, exactly four times (fem.cpp:62-69) */
static void polar4(const double* f, double* r) {
    double inv[9], it[9];
    memcpy(r, f, sizeof(double) * 9);
    for (int i = 0; i < 4; ++i) {
        inv3(r, inv);
        transpose3(inv, it);
        for (int k = 0; k < 9; ++k) r[k] = 0.5 * (r[k] + it[k]);
    }
}

/* Kuhn subdivision of a cell into six tets on the 0-7 diagonal (fem.cpp:236-237);
 * cell corner c sits at (i + (c&1), j + ((c>>1)&1), k + ((c>>2)&1)). */
static const int kKuhn[6][4] = {{0, 1, 3, 7}, {0, 1, 7, 5}, {0, 5, 7, 4},
                                {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 4, 6, 7}};

/* ======================================================================== */
/* meshes                                                                    */
/* ======================================================================== */

/* Appends the six Kuhn tets of one cell with corner node ids `corner`, swapping
 * v1,v2 when the double-precision orientation determinant is negative
 * (fem.cpp:246-258). */
static int64_t emit_cell(const double* X, const uint32_t corner[8], uint32_t* index, int64_t t) {
    for (int k = 0; k < 6; ++k) {
        uint32_t v[4];
        for (int c = 0; c < 4; ++c) v[c] = corner[kKuhn[k][c]];
        double dm[9];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) AT(dm, a, c) = X[v[c + 1] * 3 + a] - X[v[0] * 3 + a];
        if (det3(dm) < 0) {
            uint32_t s = v[1];
            v[1] = v[2];
            v[2] = s;
        }
        if (index) memcpy(&index[t * 4], v, sizeof v);
        ++t;
    }
    return t;
}

void ora_box_mesh(double lx, double ly, double lz, int nx, int ny, int nz, int64_t* nodes,
                  int64_t* tets, double* X, uint32_t* index) {
    const int sx = nx + 1, sy = ny + 1, sz = nz + 1;
    if (nodes) *nodes = (int64_t)sx * sy * sz;
    if (tets) *tets = (int64_t)nx * ny * nz * 6;
    if (!X || !index) return;
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                double* p = &X[((int64_t)(k * sy + j) * sx + i) * 3];
                p[0] = lx * (double)i / (double)nx;
                p[1] = ly * (double)j / (double)ny;
                p[2] = lz * (double)k / (double)nz;
            }
    int64_t t = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                uint32_t corner[8];
                for (int c = 0; c < 8; ++c)
                    corner[c] = (uint32_t)(((k + ((c >> 2) & 1)) * sy + (j + ((c >> 1) & 1))) * sx +
                                           (i + (c & 1)));
                t = emit_cell(X, corner, index, t);
            }
}

void ora_box_mesh_f32(float lx, float ly, float lz, int nx, int ny, int nz, float* X,
                      uint32_t* index) {
    const int sx = nx + 1, sy = ny + 1, sz = nz + 1;
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                float* p = &X[((int64_t)(k * sy + j) * sx + i) * 3];
                p[0] = lx * (float)i / (float)nx;
                p[1] = ly * (float)j / (float)ny;
                p[2] = lz * (float)k / (float)nz;
            }
    int64_t t = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                uint32_t corner[8];
                for (int c = 0; c < 8; ++c)
                    corner[c] = (uint32_t)(((k + ((c >> 2) & 1)) * sy + (j + ((c >> 1) & 1))) * sx +
                                           (i + (c & 1)));
                for (int q = 0; q < 6; ++q) {
                    uint32_t v[4];
                    for (int c = 0; c < 4; ++c) v[c] = corner[kKuhn[q][c]];
                    float d[9];
                    for (int c = 0; c < 3; ++c)
                        for (int a = 0; a < 3; ++a)
                            d[c * 3 + a] = X[v[c + 1] * 3 + a] - X[v[0] * 3 + a];
                    /* float determinant, same expression as fem.cpp:36-40 */
                    float det = d[0] * (d[4] * d[8] - d[7] * d[5]) -
                                d[3] * (d[1] * d[8] - d[7] * d[2]) +
                                d[6] * (d[1] * d[5] - d[4] * d[2]);
                    if (det < 0) {
                        uint32_t s = v[1];
                        v[1] = v[2];
                        v[2] = s;
                    }
                    memcpy(&index[t * 4], v, sizeof v);
                    ++t;
                }
            }
}

double ora_uniform(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void ora_box_jitter(double* X, int nx, int ny, int nz, double lx, double ly, double lz,
                    double amp, uint64_t seed) {
    const int sx = nx + 1, sy = ny + 1, sz = nz + 1;
    const double h[3] = {lx / nx, ly / ny, lz / nz};
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                const int64_t n = (int64_t)(k * sy + j) * sx + i;
                const int onface[3] = {i == 0 || i == nx, j == 0 || j == ny, k == 0 || k == nz};
                for (int a = 0; a < 3; ++a) {
                    if (onface[a]) continue;
                    const double r = ora_uniform(seed, (uint64_t)n * 3 + (uint64_t)a);
                    X[n * 3 + a] += amp * h[a] * (2.0 * r - 1.0);
                }
            }
}

/* y grid lines of the stiffened panel; returns count of lines, fills web column
 * cell indices. */
static int panel_y_lines(const ora_panel_spec* s, double* y, int* web_col) {
    int n = 0;
    if (y) y[n] = 0.0;
    ++n;
    for (int b = 0; b <= s->n_stiff; ++b) {
        const double ys_prev = s->ly * (double)b / (double)(s->n_stiff + 1);
        const double y0 = b == 0 ? 0.0 : ys_prev + 0.5 * s->web_t;
        const double y1 = b == s->n_stiff
                              ? s->ly
                              : s->ly * (double)(b + 1) / (double)(s->n_stiff + 1) - 0.5 * s->web_t;
        for (int j = 1; j <= s->ny_bay; ++j) {
            if (y) y[n] = j == s->ny_bay ? y1 : y0 + (y1 - y0) * (double)j / (double)s->ny_bay;
            ++n;
        }
        if (b < s->n_stiff) {
            if (web_col) web_col[b] = n - 1; /* the cell between line n-1 and n */
            if (y) y[n] = s->ly * (double)(b + 1) / (double)(s->n_stiff + 1) + 0.5 * s->web_t;
            ++n;
        }
    }
    return n;
}

void ora_panel_mesh(const ora_panel_spec* s, int64_t* nodes, int64_t* tets, double* X,
                    uint32_t* index) {
    const int ny_lines = panel_y_lines(s, NULL, NULL);
    const int ny = ny_lines - 1, nx = s->nx, nzc = 1 + s->nz_web;
    const int sx = nx + 1, sy = ny + 1, sz = nzc + 1;
    double* y = (double*)malloc(sizeof(double) * ny_lines);
    int* webc = (int*)malloc(sizeof(int) * (s->n_stiff > 0 ? s->n_stiff : 1));
    panel_y_lines(s, y, webc);
    unsigned char* colweb = (unsigned char*)calloc((size_t)ny, 1);
    for (int b = 0; b < s->n_stiff; ++b) colweb[webc[b]] = 1;
#define CELL_ACTIVE(i, j, k) ((k) == 0 || colweb[(j)])
    /* node used flags over the full lattice, then lexicographic numbering */
    int64_t* id = (int64_t*)malloc(sizeof(int64_t) * (size_t)sx * sy * sz);
    for (int64_t q = 0; q < (int64_t)sx * sy * sz; ++q) id[q] = -1;
    int64_t cells = 0;
    for (int k = 0; k < nzc; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                if (!CELL_ACTIVE(i, j, k)) continue;
                ++cells;
                for (int c = 0; c < 8; ++c)
                    id[((int64_t)(k + ((c >> 2) & 1)) * sy + (j + ((c >> 1) & 1))) * sx +
                       (i + (c & 1))] = 0;
            }
    int64_t nn = 0;
    for (int k = 0; k < sz; ++k)
        for (int j = 0; j < sy; ++j)
            for (int i = 0; i < sx; ++i) {
                int64_t q = ((int64_t)k * sy + j) * sx + i;
                if (id[q] < 0) continue;
                id[q] = nn;
                if (X) {
                    X[nn * 3 + 0] = s->lx * (double)i / (double)nx;
                    X[nn * 3 + 1] = y[j];
                    X[nn * 3 + 2] = k == 0 ? 0.0
                                           : s->t_plate + s->web_h * (double)(k - 1) / (double)s->nz_web;
                }
                ++nn;
            }
    if (nodes) *nodes = nn;
    if (tets) *tets = cells * 6;
    if (X && index) {
        int64_t t = 0;
        for (int k = 0; k < nzc; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    if (!CELL_ACTIVE(i, j, k)) continue;
                    uint32_t corner[8];
                    for (int c = 0; c < 8; ++c)
                        corner[c] = (uint32_t)id[((int64_t)(k + ((c >> 2) & 1)) * sy +
                                                  (j + ((c >> 1) & 1))) * sx + (i + (c & 1))];
                    t = emit_cell(X, corner, index, t);
                }
    }
#undef CELL_ACTIVE
    free(id);
    free(colweb);
    free(webc);
    free(y);
}

/* ======================================================================== */
/* rest state, time step, loads                                             */
/* ======================================================================== */

void ora_rest_state(const double* X, int64_t N, const uint32_t* index, int64_t T, double density,
                    double* rest_inv, double* volume, double* mass, uint32_t* adj_offset,
                    uint32_t* adj_entry) {
    memset(mass, 0, sizeof(double) * (size_t)N);
    for (int64_t t = 0; t < T; ++t) {
        const uint32_t* idx = &index[t * 4];
        double dm[9];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) AT(dm, a, c) = X[idx[c + 1] * 3 + a] - X[idx[0] * 3 + a];
        volume[t] = fabs(det3(dm) / 6.0);
        inv3(dm, &rest_inv[t * 9]);
        const double quarter = density * volume[t] * 0.25;
        for (int c = 0; c < 4; ++c) mass[idx[c]] += quarter;
    }
    /* CSR node -> (tet*4 + corner) by counting sort, entries in (tet, corner)
     * order (fem.cpp:147-157) */
    memset(adj_offset, 0, sizeof(uint32_t) * (size_t)(N + 1));
    for (int64_t t = 0; t < T; ++t)
        for (int c = 0; c < 4; ++c) ++adj_offset[index[t * 4 + c] + 1];
    for (int64_t n = 0; n < N; ++n) adj_offset[n + 1] += adj_offset[n];
    uint32_t* cursor = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N);
    memcpy(cursor, adj_offset, sizeof(uint32_t) * (size_t)N);
    for (int64_t t = 0; t < T; ++t)
        for (uint32_t c = 0; c < 4; ++c) adj_entry[cursor[index[t * 4 + c]]++] = (uint32_t)t * 4 + c;
    free(cursor);
}

double ora_edge_timestep(const double* X, const uint32_t* index, int64_t T, double E, double rho,
                         double safety) {
    double shortest = INFINITY;
    for (int64_t t = 0; t < T; ++t) {
        const uint32_t* idx = &index[t * 4];
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                double d2 = 0;
                for (int k = 0; k < 3; ++k) {
                    const double d = X[idx[a] * 3 + k] - X[idx[b] * 3 + k];
                    d2 += d * d;
                }
                if (sqrt(d2) < shortest) shortest = sqrt(d2);
            }
    }
    return safety * shortest / sqrt(E / rho);
}

/* Cyclic Jacobi on a symmetric n x n (row-major), returns the largest eigenvalue. */
static double jacobi_max_eig(double* a, int n) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, diag = 0;
        for (int p = 0; p < n; ++p) {
            diag += a[p * n + p] * a[p * n + p];
            for (int q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
        }
        if (off <= 1e-30 * diag) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
            }
    }
    double m = a[0];
    for (int p = 1; p < n; ++p)
        if (a[p * n + p] > m) m = a[p * n + p];
    return m;
}

double ora_max_element_omega(const double* rest_inv, int64_t T, double E, double nu, double rho) {
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
    double best = 0;
    for (int64_t t = 0; t < T; ++t) {
        const double* di = &rest_inv[t * 9];
        double g[4][3];
        /* grad N_c = row (c-1) of Dm^-1 for c = 1..3, grad N_0 = -sum */
        for (int c = 1; c < 4; ++c)
            for (int a = 0; a < 3; ++a) g[c][a] = AT(di, c - 1, a);
        for (int a = 0; a < 3; ++a) g[0][a] = -(g[1][a] + g[2][a] + g[3][a]);
        /* A = B^T C B by columns: A e_(b,j) is the nodal force of the unit
         * displacement of node b along j, per unit volume */
        double A[144];
        for (int b = 0; b < 4; ++b)
            for (int j = 0; j < 3; ++j) {
                double grad[3][3] = {{0}}; /* grad u (row i: component, col: derivative) */
                for (int k = 0; k < 3; ++k) grad[j][k] = g[b][k];
                double eps[3][3];
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) eps[p][q] = 0.5 * (grad[p][q] + grad[q][p]);
                const double tr = eps[0][0] + eps[1][1] + eps[2][2];
                double sig[3][3];
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) sig[p][q] = 2 * mu * eps[p][q] + (p == q ? lam * tr : 0);
                for (int a = 0; a < 4; ++a)
                    for (int i = 0; i < 3; ++i) {
                        double f = 0;
                        for (int k = 0; k < 3; ++k) f += sig[i][k] * g[a][k];
                        A[(a * 3 + i) * 12 + (b * 3 + j)] = f;
                    }
            }
        const double lmax = jacobi_max_eig(A, 12);
        const double omega = sqrt(4.0 * lmax / rho);
        if (omega > best) best = omega;
    }
    return best;
}

void ora_face_traction(const double* X, const uint32_t* index, int64_t T, int axis, double value,
                       double tol, const double traction[3], double* load) {
    static const int faces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    for (int64_t t = 0; t < T; ++t) {
        const uint32_t* idx = &index[t * 4];
        for (int f = 0; f < 4; ++f) {
            int on = 1;
            for (int q = 0; q < 3; ++q)
                if (fabs(X[idx[faces[f][q]] * 3 + axis] - value) > tol) on = 0;
            if (!on) continue;
            const double* a = &X[idx[faces[f][0]] * 3];
            const double* b = &X[idx[faces[f][1]] * 3];
            const double* c = &X[idx[faces[f][2]] * 3];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                  e1[0] * e2[1] - e1[1] * e2[0]};
            const double area = 0.5 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int q = 0; q < 3; ++q)
                for (int k = 0; k < 3; ++k)
                    load[idx[faces[f][q]] * 3 + k] += traction[k] * area / 3.0;
        }
    }
}

/* ======================================================================== */
/* element: co-rotational force, elastic or J2                              */
/* ======================================================================== */

/* Stress from the co-rotated small strain (tensor eps, col-major), returning the
 * symmetric stress tensor and, for J2, updating the point's plastic state. */
static int j2_return(const ora_j2* m, const double strain[6], double eps_p[6], double* eqps,
                     double* damage, double stress[6], double* dissipation, double normal[3],
                     double* eta);

static void stress_of(const double* eps, double lam, double mu, const ora_j2* j2, double* ep,
                      double* eqps, double* damage, double* sig, double* diss,
                      double* energy_density, int* torn, double stress_torn[6], double* released) {
    if (torn) *torn = 0;
    if (!j2) {
        const double tr = AT(eps, 0, 0) + AT(eps, 1, 1) + AT(eps, 2, 2);
        for (int k = 0; k < 9; ++k) sig[k] = 2.0 * mu * eps[k];
        for (int i = 0; i < 3; ++i) AT(sig, i, i) += lam * tr;
        if (diss) *diss = 0;
        if (energy_density) {
            double s = 0;
            for (int k = 0; k < 9; ++k) s += sig[k] * eps[k];
            *energy_density = 0.5 * s;
        }
        return;
    }
    const double strain[6] = {AT(eps, 0, 0), AT(eps, 1, 1), AT(eps, 2, 2),
                              2 * AT(eps, 0, 1), 2 * AT(eps, 1, 2), 2 * AT(eps, 0, 2)};
    double s6[6], d = 0;
    const int r = j2_return(j2, strain, ep, eqps, damage, s6, &d, NULL, NULL);
    if (diss) *diss = d;
    if (r & 2) {
        /* torn in this step: its stored energy leaves the account as fracture; the
         * returned stress (co-rotated frame) goes out for the failure plane and, kept
         * here, gives the forces the caller needs to finish the release and zeroes */
        double u = 0;
        for (int i = 0; i < 6; ++i) u += s6[i] * (strain[i] - ep[i]);
        if (released) *released = 0.5 * u;
        if (torn) *torn = 1;
        if (stress_torn)
            for (int i = 0; i < 6; ++i) stress_torn[i] = s6[i];
    }
    AT(sig, 0, 0) = s6[0];
    AT(sig, 1, 1) = s6[1];
    AT(sig, 2, 2) = s6[2];
    AT(sig, 0, 1) = AT(sig, 1, 0) = s6[3];
    AT(sig, 1, 2) = AT(sig, 2, 1) = s6[4];
    AT(sig, 0, 2) = AT(sig, 2, 0) = s6[5];
    if (energy_density) {
        /* 1/2 sigma : (eps - eps_p); engineering shears pair with tensor stress; a
         * torn point holds none */
        double s = 0;
        for (int i = 0; i < 6; ++i) s += s6[i] * (strain[i] - ep[i]);
        *energy_density = (r & 2) ? 0.0 : 0.5 * s;
    }
}

/* Force of one tet from current node positions x (4 x xyz). Mirrors
 * fem.cpp:73-117 in double, with the stress law pluggable. */
static void tet_force_general(const double x[12], const double* dmi, double volume, double lam,
                              double mu, const ora_j2* j2, double* ep, double* eqps,
                              double* damage, double out[12], double* diss, double* energy,
                              int* torn, double normal_world[3], double* released) {
    double ds[9], f[9], r[9], rt[9], s[9], eps[9], sig[9], piola[9], dmit[9], h[9];
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) AT(ds, a, c) = x[(c + 1) * 3 + a] - x[a];
    mul3(ds, dmi, f);
    polar4(f, r);
    transpose3(r, rt);
    mul3(rt, f, s);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) AT(eps, i, j) = 0.5 * (AT(s, i, j) + AT(s, j, i));
    for (int i = 0; i < 3; ++i) AT(eps, i, i) -= 1.0;
    double ed = 0, rel = 0, sloc[6] = {0, 0, 0, 0, 0, 0};
    int t = 0;
    stress_of(eps, lam, mu, j2, ep, eqps, damage, sig, diss, energy ? &ed : NULL, &t, sloc, &rel);
    if (diss) *diss *= volume;
    if (energy) *energy = volume * ed;
    if (torn) *torn = t;
    if (t) {
        /* failure plane: maximum principal direction of the world-frame stress
         * R sigma R^T at the tear, largest component positive */
        if (released) *released = volume * rel;
        if (normal_world) {
            double sl[9] = {sloc[0], sloc[3], sloc[5], sloc[3], sloc[1], sloc[4], sloc[5], sloc[4],
                            sloc[2]};
            double tmp[9], rt2[9], sw[9];
            mul3(r, sl, tmp);
            transpose3(r, rt2);
            mul3(tmp, rt2, sw);
            ora_world_stress_normal(sw, normal_world);
        }
    }
    mul3(r, sig, piola);
    transpose3(dmi, dmit);
    mul3(piola, dmit, h);
    double f0[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            const double v = -volume * AT(h, a, c);
            out[(c + 1) * 3 + a] = v;
            f0[a] -= v;
        }
    for (int a = 0; a < 3; ++a) out[a] = f0[a];
}

void ora_tet_force(const double x[12], const double rest_inv[9], double volume, double lambda,
                   double mu, double out[12]) {
    tet_force_general(x, rest_inv, volume, lambda, mu, NULL, NULL, NULL, NULL, out, NULL, NULL,
                      NULL, NULL, NULL);
}

double ora_regularised_failure_strain(double uniform, double fracture, double length,
                                      double thickness) {
    /* an element no larger than the plate is thick resolves the neck itself */
    double share = 1.0;
    if (length > 0.0 && thickness > 0.0) share = thickness / length < 1.0 ? thickness / length : 1.0;
    return uniform + (fracture - uniform) * share;
}

double ora_triaxiality_factor(double sensitivity, double reference, double cutoff, double eta) {
    if (eta <= cutoff) return INFINITY;
    return exp(-sensitivity * (eta - reference));
}

/* Maximum principal direction of a symmetric stress (Voigt, tensor shears) by
 * cyclic Jacobi, sign fixed so the largest component is positive. */
static void max_principal_direction(const double s6[6], double out[3]) {
    double a[9] = {s6[0], s6[3], s6[5], s6[3], s6[1], s6[4], s6[5], s6[4], s6[2]};
    double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        const double diag = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (off <= 1e-30 * diag || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = a[p * 3 + q];
                if (apq == 0.0) continue;
                const double theta = (a[q * 3 + q] - a[p * 3 + p]) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double akp = a[k * 3 + p], akq = a[k * 3 + q];
                    a[k * 3 + p] = c * akp - sn * akq;
                    a[k * 3 + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double apk = a[p * 3 + k], aqk = a[q * 3 + k];
                    a[p * 3 + k] = c * apk - sn * aqk;
                    a[q * 3 + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double vkp = v[k * 3 + p], vkq = v[k * 3 + q];
                    v[k * 3 + p] = c * vkp - sn * vkq;
                    v[k * 3 + q] = sn * vkp + c * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 3; ++i)
        if (a[i * 3 + i] > a[best * 3 + best]) best = i;
    for (int i = 0; i < 3; ++i) out[i] = v[i * 3 + best];
    double norm = sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
    if (!(norm > 0.0)) {
        out[0] = 1.0;
        out[1] = out[2] = 0.0;
        return;
    }
    int largest = 0;
    for (int i = 1; i < 3; ++i)
        if (fabs(out[i]) > fabs(out[largest])) largest = i;
    if (out[largest] < 0.0) norm = -norm;
    for (int i = 0; i < 3; ++i) out[i] /= norm;
}

/* The return with damage; at failure `stress` keeps the returned (pre-failure)
 * stress and bit 1 is set: the caller zeroes it. */
static int j2_return_t(const ora_j2* m, const double strain[6], double eps_p[6], double* eqps,
                       double* damage, double stress[6], double* dissipation, double normal[3],
                       double* eta, double tangent[36]);

static int j2_return(const ora_j2* m, const double strain[6], double eps_p[6], double* eqps,
                     double* damage, double stress[6], double* dissipation, double normal[3],
                     double* eta) {
    return j2_return_t(m, strain, eps_p, eqps, damage, stress, dissipation, normal, eta, NULL);
}

/* The elastic moduli in Voigt form with engineering shears (plasticity.cpp elasticModuli). */
static void elastic_moduli(double mu, double kappa, double c[36]) {
    for (int k = 0; k < 36; ++k) c[k] = 0.0;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) c[i * 6 + j] = kappa + 2.0 * mu * ((i == j) - 1.0 / 3.0);
        c[(3 + i) * 6 + 3 + i] = mu;
    }
}

/* With `tangent`, also the algorithmic tangent of the return (plasticity.cpp:481-505):
 * kappa 1(x)1 + 2 mu theta I_dev - 2 mu thetaBar n(x)n, theta = 1 - 2 mu dgamma/|s_trial|,
 * thetaBar = 1/(1 + H/(3 mu)) - (1 - theta); the elastic moduli when the point does not
 * yield, zero for a failed point. */
static int j2_return_t(const ora_j2* m, const double strain[6], double eps_p[6], double* eqps,
                       double* damage, double stress[6], double* dissipation, double normal[3],
                       double* eta, double tangent[36]) {
    if (eta) *eta = 0.0;
    if (dissipation) *dissipation = 0.0;
    /* a failed point carries no stress and its history is frozen (plasticity.cpp:326) */
    if (damage && *damage >= 1.0) {
        for (int i = 0; i < 6; ++i) stress[i] = 0.0;
        if (tangent)
            for (int k = 0; k < 36; ++k) tangent[k] = 0.0;
        return 0;
    }
    const double mu = m->E / (2.0 * (1.0 + m->nu));
    const double kappa = m->E / (3.0 * (1.0 - 2.0 * m->nu));
    const double root23 = sqrt(2.0 / 3.0);
    double el[6];
    for (int i = 0; i < 6; ++i) el[i] = strain[i] - eps_p[i];
    const double vol = el[0] + el[1] + el[2];
    const double p = kappa * vol;
    double tr[6];
    for (int i = 0; i < 3; ++i) {
        tr[i] = 2.0 * mu * (el[i] - vol / 3.0);
        tr[3 + i] = mu * el[3 + i];
    }
    const double norm = sqrt(tr[0] * tr[0] + tr[1] * tr[1] + tr[2] * tr[2] +
                             2.0 * (tr[3] * tr[3] + tr[4] * tr[4] + tr[5] * tr[5]));
    const double radius = root23 * (m->yield + m->hardening * (*eqps > 0 ? *eqps : 0.0));
    const double excess = norm - radius;
    if (!(excess > 0.0)) {
        for (int i = 0; i < 3; ++i) {
            stress[i] = p + tr[i];
            stress[3 + i] = tr[3 + i];
        }
        if (tangent) elastic_moduli(mu, kappa, tangent);
        return 0;
    }
    /* linear hardening makes the consistency condition affine in dgamma: the
     * Newton root of plasticity.cpp:370-413 in closed form */
    const double gamma = excess / (2.0 * mu + (2.0 / 3.0) * m->hardening);
    double n[6], dep[6];
    for (int i = 0; i < 6; ++i) n[i] = tr[i] / norm;
    for (int i = 0; i < 3; ++i) {
        dep[i] = gamma * n[i];
        dep[3 + i] = 2.0 * gamma * n[3 + i];
        stress[i] = p + tr[i] - 2.0 * mu * gamma * n[i];
        stress[3 + i] = tr[3 + i] - 2.0 * mu * gamma * n[3 + i];
    }
    double d = 0;
    for (int i = 0; i < 6; ++i) d += stress[i] * dep[i];
    if (dissipation) *dissipation = d;
    for (int i = 0; i < 6; ++i) eps_p[i] += dep[i];
    const double deq = root23 * gamma;
    *eqps += deq;
    /* triaxiality of the returned stress: mean over von Mises, sqrt(3/2)|s| */
    const double q = sqrt(1.5) * (norm - 2.0 * mu * gamma);
    const double e = q > 0.0 ? p / q : 0.0;
    if (eta) *eta = e;
    if (damage && m->failure_strain > 0.0) {
        const double critical = m->failure_strain * ora_triaxiality_factor(
            m->sensitivity, m->reference_triaxiality, m->cutoff_triaxiality, e);
        if (critical > 0.0 && isfinite(critical)) *damage += deq / critical;
        if (*damage >= 1.0) {
            *damage = 1.0;
            if (normal) max_principal_direction(stress, normal);
            if (tangent)
                for (int k = 0; k < 36; ++k) tangent[k] = 0.0;
            return 3;
        }
    }
    if (tangent) {
        const double theta = 1.0 - 2.0 * mu * gamma / norm;
        const double thetabar = 1.0 / (1.0 + m->hardening / (3.0 * mu)) - (1.0 - theta);
        elastic_moduli(mu * theta, kappa, tangent);
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) tangent[i * 6 + j] -= 2.0 * mu * thetabar * n[i] * n[j];
    }
    return 1;
}

void ora_world_stress_normal(const double s9[9], double normal[3]) {
    const double v[6] = {s9[0], s9[4], s9[8], s9[1], s9[5], s9[2]};
    max_principal_direction(v, normal);
}

int ora_j2_update_damage(const ora_j2* m, const double strain[6], double eps_p[6],
                         double* eqps, double* damage, double stress[6], double* dissipation,
                         double normal[3], double* eta) {
    const int r = j2_return(m, strain, eps_p, eqps, damage, stress, dissipation, normal, eta);
    if (r & 2)
        for (int i = 0; i < 6; ++i) stress[i] = 0.0;
    return r;
}

int ora_j2_update(const ora_j2* m, const double strain[6], double eps_p[6], double* eqps,
                  double stress[6], double* dissipation) {
    return ora_j2_update_damage(m, strain, eps_p, eqps, NULL, stress, dissipation, NULL, NULL) & 1;
}

/* ======================================================================== */
/* explicit solver                                                          */
/* ======================================================================== */

static int hex_case(const ora_case* c) { return c->corners == 8; }

int ora_case_init(ora_case* c) {
    if (hex_case(c)) {
        /* nodal masses and the (element, corner) CSR of the hex solve (zone.cpp:978-980) */
        const int64_t N = c->N, CT = 8 * c->T;
        c->rest_inv = NULL;
        c->volume = (double*)calloc((size_t)c->T, sizeof(double));
        c->mass = (double*)calloc((size_t)N, sizeof(double));
        c->adj_offset = (uint32_t*)calloc((size_t)(N + 1), sizeof(uint32_t));
        c->adj_entry = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)CT);
        uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(N + 1));
        if (!c->volume || !c->mass || !c->adj_offset || !c->adj_entry || !cur) {
            free(cur);
            return 1;
        }
        for (int64_t q = 0; q < CT; ++q) {
            c->mass[c->index[q]] += c->hex_mass[q];
            c->volume[q / 8] += c->hex_mass[q] / c->rho;
            ++c->adj_offset[c->index[q] + 1];
        }
        for (int64_t i = 0; i < N; ++i) c->adj_offset[i + 1] += c->adj_offset[i];
        memcpy(cur, c->adj_offset, sizeof(uint32_t) * (size_t)N);
        for (int64_t q = 0; q < CT; ++q) c->adj_entry[cur[c->index[q]]++] = (uint32_t)q;
        free(cur);
        return 0;
    }
    c->rest_inv = (double*)malloc(sizeof(double) * 9 * (size_t)c->T);
    c->volume = (double*)malloc(sizeof(double) * (size_t)c->T);
    c->mass = (double*)malloc(sizeof(double) * (size_t)c->N);
    c->adj_offset = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(c->N + 1));
    c->adj_entry = (uint32_t*)malloc(sizeof(uint32_t) * 4 * (size_t)c->T);
    if (!c->rest_inv || !c->volume || !c->mass || !c->adj_offset || !c->adj_entry) return 1;
    ora_rest_state(c->X, c->N, c->index, c->T, c->rho, c->rest_inv, c->volume, c->mass,
                   c->adj_offset, c->adj_entry);
    return 0;
}

void ora_case_free(ora_case* c) {
    free(c->rest_inv);
    free(c->volume);
    free(c->mass);
    free(c->adj_offset);
    free(c->adj_entry);
    c->rest_inv = c->volume = c->mass = NULL;
    c->adj_offset = c->adj_entry = NULL;
}

static double load_scale(const ora_case* c, double t) {
    if (!(c->load_ramp > 0)) return 1.0;
    const double s = t / c->load_ramp;
    return s < 1.0 ? (s > 0.0 ? s : 0.0) : 1.0;
}

static double drive_fraction(const ora_case* c, double t) {
    if (!(c->drive_ramp > 0)) return 1.0;
    double s = t / c->drive_ramp;
    s = s < 0 ? 0 : (s > 1 ? 1 : s);
    return s * s * (3.0 - 2.0 * s);
}

/* Per-step sums of one range of elements or nodes. */
typedef struct {
    double strain, diss, fracture;                    /* element range */
    double w_ext, w_con, w_bnd, damp;                 /* node range, running terms */
    double kinetic, half_ext, half_con, half_bnd, half_damp, fcon[3]; /* sample terms */
} ora_sums;

/* Co-rotational force of one solid-shell hex (solidshell::internalForce,
 * solid_shell.cpp:680-766): R the polar factor of F = J(x) J(X)^-1 at the element centre by
 * R <- (R + R^-T)/2 until no entry moves by 1e-16 (at most 64 steps), d = R^T x - X,
 * f = -R K d; adds 1/2 d.K d to *energy when it is not NULL (zone.cpp:1040-1060). */
static void hex_force(const double* K, const double X[24], const double x[24], double f[24],
                      double* energy) {
    static const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double Jx[9], JX[9], JXi[9], r[9], inv[9], it[9];
    for (int i = 0; i < 3; ++i) {
        double gx[3] = {0, 0, 0}, gX[3] = {0, 0, 0};
        for (int a = 0; a < 8; ++a) {
            const double dn[3] = {0.125 * sx[a], 0.125 * sy[a], 0.125 * sz[a]};
            for (int k = 0; k < 3; ++k) {
                gx[k] += dn[k] * x[3 * a + i];
                gX[k] += dn[k] * X[3 * a + i];
            }
        }
        for (int k = 0; k < 3; ++k) {
            AT(Jx, i, k) = gx[k];
            AT(JX, i, k) = gX[k];
        }
    }
    inv3(JX, JXi);
    mul3(Jx, JXi, r);
    for (int step = 0; step < 64; ++step) {
        inv3(r, inv);
        transpose3(inv, it);
        double change = 0;
        for (int k = 0; k < 9; ++k) {
            const double nx = 0.5 * (r[k] + it[k]);
            change = fmax(change, fabs(nx - r[k]));
            r[k] = nx;
        }
        if (change < 1e-16) break;
    }
    double d[24], kd[24];
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += AT(r, k, i) * x[3 * a + k];
            d[3 * a + i] = s - X[3 * a + i];
        }
    double e = 0;
    for (int p = 0; p < 24; ++p) {
        double s = 0;
        for (int q = 0; q < 24; ++q) s += K[p * 24 + q] * d[q];
        kd[p] = s;
        e += 0.5 * d[p] * s;
    }
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
            double s = 0;
            for (int k = 0; k < 3; ++k) s += AT(r, i, k) * kd[3 * a + k];
            f[3 * a + i] = -s;
        }
    if (energy) *energy += e;
}

void ora_hex_force(const double K[576], const double X[24], const double x[24], double out[24]) {
    hex_force(K, X, x, out, NULL);
}

/* Polar factor of the centre deformation gradient of a hex, to convergence (column-major),
 * as hex_force forms it. */
static void hex_rotation(const double X[24], const double x[24], double r[9]) {
    static const double sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
    static const double sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
    static const double sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
    double Jx[9], JX[9], JXi[9], inv[9], it[9];
    for (int i = 0; i < 3; ++i) {
        double gx[3] = {0, 0, 0}, gX[3] = {0, 0, 0};
        for (int a = 0; a < 8; ++a) {
            const double dn[3] = {0.125 * sx[a], 0.125 * sy[a], 0.125 * sz[a]};
            for (int k = 0; k < 3; ++k) {
                gx[k] += dn[k] * x[3 * a + i];
                gX[k] += dn[k] * X[3 * a + i];
            }
        }
        for (int k = 0; k < 3; ++k) {
            AT(Jx, i, k) = gx[k];
            AT(JX, i, k) = gX[k];
        }
    }
    inv3(JX, JXi);
    mul3(Jx, JXi, r);
    for (int step = 0; step < 64; ++step) {
        inv3(r, inv);
        transpose3(inv, it);
        double change = 0;
        for (int k = 0; k < 9; ++k) {
            const double nx = 0.5 * (r[k] + it[k]);
            change = fmax(change, fabs(nx - r[k]));
            r[k] = nx;
        }
        if (change < 1e-16) break;
    }
}

/* In-place Cholesky solve of the n x n system a (row-major, stride 7) for b; 0 when a
 * pivot is not positive. */
static int chol_solve7(const double* a, int n, double* b) {
    double l[49];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double t = a[i * 7 + j];
            for (int k = 0; k < j; ++k) t -= l[i * 7 + k] * l[j * 7 + k];
            if (i == j) {
                if (!(t > 0.0)) return 0;
                l[i * 7 + i] = sqrt(t);
            } else {
                l[i * 7 + j] = t / l[j * 7 + j];
            }
        }
    for (int i = 0; i < n; ++i) {
        double t = b[i];
        for (int k = 0; k < i; ++k) t -= l[i * 7 + k] * b[k];
        b[i] = t / l[i * 7 + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double t = b[i];
        for (int k = i + 1; k < n; ++k) t -= l[k * 7 + i] * b[k];
        b[i] = t / l[i * 7 + i];
    }
    return 1;
}


/* Stored energy density 1/2 s : C^-1 : s of a Voigt stress (tensor shears), as
 * zone.cpp:1022-1036 forms it. */
static double stored_density(const double s[6], double mu, double kappa) {
    const double p = (s[0] + s[1] + s[2]) / 3.0;
    double d = 0;
    for (int i = 0; i < 3; ++i) d += 0.5 * s[i] * ((s[i] - p) / (2.0 * mu) + p / (3.0 * kappa));
    for (int i = 3; i < 6; ++i) d += 0.5 * s[i] * (s[i] / mu);
    return d;
}

static int hex_plastic(const double* B, const double* G, const double* w, const double X[24],
                       const double x[24], const ora_j2* m, double* point, double enhanced[7],
                       double force[24], double stress[48], double info[6], double* released);

int ora_hex_plastic_update(const double* B, const double* G, const double* w,
                           const double X[24], const double x[24], const ora_j2* m,
                           double* point, double enhanced[7], double force[24],
                           double stress[48], double info[6]) {
    return hex_plastic(B, G, w, X, x, m, point, enhanced, force, stress, info, NULL);
}

/* With `released`: the stored energy w 1/2 s : C^-1 : s of every point that fails in this
 * update, s its returned stress at failure (that energy leaves the element with it). */
static int hex_plastic(const double* B, const double* G, const double* w, const double X[24],
                       const double x[24], const ora_j2* m, double* point, double enhanced[7],
                       double force[24], double stress[48], double info[6], double* released) {
    const double mu = m->E / (2.0 * (1.0 + m->nu));
    const double kappa = m->E / (3.0 * (1.0 - 2.0 * m->nu));
    for (int k = 0; k < 24; ++k) force[k] = 0.0;
    for (int k = 0; k < 6; ++k) info[k] = 0.0;
    double r[9], d[24];
    hex_rotation(X, x, r);
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
            double t = 0;
            for (int k = 0; k < 3; ++k) t += AT(r, k, i) * x[3 * a + k];
            d[3 * a + i] = t - X[3 * a + i];
        }
    /* an element with a failed point has dropped its enhanced modes (solid_shell.cpp:860-866) */
    int eas = 7;
    for (int g = 0; g < 8; ++g)
        if (point[9 * g + 7] >= 1.0) eas = 0;
    double volume = 0;
    for (int g = 0; g < 8; ++g) volume += w[g];
    const double scale = m->yield * volume;
    /* the base strain B d of each point, and the history at the start of the step */
    double eu[8][6], start[8][8], trial[8][8], sg[8][6], cg[8][36], diss[8], rel[8];
    int flags[8];
    for (int g = 0; g < 8; ++g) {
        for (int i = 0; i < 6; ++i) {
            double t = 0;
            for (int j = 0; j < 24; ++j) t += B[(g * 6 + i) * 24 + j] * d[j];
            eu[g][i] = t;
        }
        for (int k = 0; k < 8; ++k) start[g][k] = point[9 * g + k];
    }
    double alpha[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < eas; ++k) alpha[k] = enhanced[k];
    int converged = 0, iterations = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        converged = 0;
        for (int iter = 0;; ++iter) {
            for (int g = 0; g < 8; ++g) {
                double e[6];
                for (int i = 0; i < 6; ++i) {
                    double t = eu[g][i];
                    for (int k = 0; k < eas; ++k) t += G[(g * 6 + i) * 7 + k] * alpha[k];
                    e[i] = t;
                }
                for (int k = 0; k < 8; ++k) trial[g][k] = start[g][k];
                flags[g] = j2_return_t(m, e, trial[g], &trial[g][6], &trial[g][7], sg[g], &diss[g],
                                       NULL, NULL, cg[g]);
                rel[g] = 0.0;
                if (flags[g] & 2) {
                    rel[g] = w[g] * stored_density(sg[g], mu, kappa);
                    for (int i = 0; i < 6; ++i) sg[g][i] = 0.0;
                }
            }
            iterations = iter + 1;
            if (eas == 0) {
                converged = 1;
                break;
            }
            double res[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int g = 0; g < 8; ++g)
                for (int k = 0; k < eas; ++k) {
                    double t = 0;
                    for (int i = 0; i < 6; ++i) t += G[(g * 6 + i) * 7 + k] * sg[g][i];
                    res[k] += w[g] * t;
                }
            double nrm = 0;
            for (int k = 0; k < eas; ++k) nrm += res[k] * res[k];
            if (sqrt(nrm) <= 1e-12 * scale) {
                converged = 1;
                break;
            }
            if (iter + 1 >= 40) break;
            double kaa[49];
            for (int k = 0; k < 49; ++k) kaa[k] = 0.0;
            for (int g = 0; g < 8; ++g) {
                double cgm[6][7];
                for (int i = 0; i < 6; ++i)
                    for (int k = 0; k < eas; ++k) {
                        double t = 0;
                        for (int j = 0; j < 6; ++j) t += cg[g][i * 6 + j] * G[(g * 6 + j) * 7 + k];
                        cgm[i][k] = t;
                    }
                for (int p = 0; p < eas; ++p)
                    for (int q = 0; q < eas; ++q) {
                        double t = 0;
                        for (int i = 0; i < 6; ++i) t += G[(g * 6 + i) * 7 + p] * cgm[i][q];
                        kaa[p * 7 + q] += w[g] * t;
                    }
            }
            double delta[7];
            for (int k = 0; k < eas; ++k) delta[k] = -res[k];
            if (!chol_solve7(kaa, eas, delta)) break;
            double work = 0;
            for (int k = 0; k < eas; ++k) work += delta[k] * res[k];
            if (fabs(work) <= 1e-16 * scale) {
                converged = 1;
                break;
            }
            for (int k = 0; k < eas; ++k) alpha[k] += delta[k];
        }
        int died = 0;
        for (int g = 0; g < 8; ++g) died |= trial[g][7] >= 1.0;
        if (!(died && eas > 0)) break;
        /* a point died inside this step: redo the element without the enhanced modes */
        eas = 0;
        for (int k = 0; k < 7; ++k) alpha[k] = 0.0;
    }
    int yielded = 0, failed = 0;
    double dissipation = 0;
    for (int g = 0; g < 8; ++g) {
        for (int k = 0; k < 8; ++k) point[9 * g + k] = trial[g][k];
        point[9 * g + 8] = trial[g][7] >= 1.0 ? 1.0 : 0.0;
        dissipation += w[g] * diss[g];
        yielded += (flags[g] & 1) != 0;
        failed += trial[g][7] >= 1.0;
        if (stress)
            for (int i = 0; i < 6; ++i) stress[6 * g + i] = sg[g][i];
    }
    for (int k = 0; k < 7; ++k) enhanced[k] = alpha[k];
    double internal[24];
    for (int j = 0; j < 24; ++j) {
        double t = 0;
        for (int g = 0; g < 8; ++g) {
            double u = 0;
            for (int i = 0; i < 6; ++i) u += B[(g * 6 + i) * 24 + j] * sg[g][i];
            t += w[g] * u;
        }
        internal[j] = t;
    }
    for (int a = 0; a < 8; ++a)
        for (int i = 0; i < 3; ++i) {
            double t = 0;
            for (int k = 0; k < 3; ++k) t += AT(r, i, k) * internal[3 * a + k];
            force[3 * a + i] = -t;
        }
    info[0] = dissipation;
    info[1] = iterations;
    info[2] = converged;
    info[3] = yielded;
    info[4] = failed;
    info[5] = failed == 8;
    if (released) {
        *released = 0.0;
        for (int g = 0; g < 8; ++g) *released += rel[g];
    }
    return failed == 8;
}

/* element pass at u_n over elements [e0, e1) */
static void run_elements(const ora_case* c, ora_state* st, int64_t e0, int64_t e1, int64_t n,
                         int sample, ora_sums* out) {
    if (hex_case(c)) {
        const ora_j2 j2 = {c->E, c->nu, c->yield, c->hardening, c->failure_strain, c->sensitivity,
                           c->reference_triaxiality, c->cutoff_triaxiality};
        const double mu = c->E / (2 * (1 + c->nu)), kappa = c->E / (3 * (1 - 2 * c->nu));
        for (int64_t e = e0; e < e1; ++e) {
            const uint32_t* idx = &c->index[e * 8];
            double* f = &st->force[e * 24];
            double X[24], x[24];
            for (int q = 0; q < 8; ++q)
                for (int a = 0; a < 3; ++a) {
                    X[q * 3 + a] = c->X[idx[q] * 3 + a];
                    x[q * 3 + a] = X[q * 3 + a] + st->u[idx[q] * 3 + a];
                }
            if (!c->plastic) {
                hex_force(&c->hex_k[e * 576], X, x, f, sample ? &out->strain : NULL);
                continue;
            }
            /* zone.cpp:935-962: a torn element carries nothing and is not updated */
            double* point = &st->hex_point[e * 72];
            int torn = 1;
            for (int g = 0; g < 8; ++g) torn &= point[9 * g + 7] >= 1.0;
            if (torn) {
                for (int k = 0; k < 24; ++k) f[k] = 0.0;
                continue;
            }
            double stress[48], info[6], rel = 0.0;
            const double* w = &c->hex_w[e * 8];
            /* the element's own failure strain when the case resolves one per element */
            ora_j2 je = j2;
            if (c->hex_failure) je.failure_strain = c->hex_failure[e];
            const int now = hex_plastic(&c->hex_B[e * 1152], &c->hex_G[e * 336], w, X, x, &je, point,
                                        &st->hex_enhanced[e * 7], f, stress, info, &rel);
            out->diss += info[0];
            out->fracture += rel;
            if (now && st->tear_step) st->tear_step[e] = n;
            if (sample)
                for (int g = 0; g < 8; ++g) out->strain += w[g] * stored_density(&stress[6 * g], mu, kappa);
        }
        return;
    }
    const double lam = c->E * c->nu / ((1 + c->nu) * (1 - 2 * c->nu));
    const double mu = c->E / (2 * (1 + c->nu));
    const ora_j2 j2 = {c->E, c->nu, c->yield, c->hardening, c->failure_strain, c->sensitivity,
                       c->reference_triaxiality, c->cutoff_triaxiality};
    const int dmg = c->plastic && c->failure_strain > 0.0 && st->damage;
    for (int64_t e = e0; e < e1; ++e) {
        const uint32_t* idx = &c->index[e * 4];
        double x[12];
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 3; ++a)
                x[q * 3 + a] = c->X[idx[q] * 3 + a] + st->u[idx[q] * 3 + a];
        double de = 0, ue = 0, rel = 0, nrm[3] = {0, 0, 0};
        int torn = 0;
        tet_force_general(x, &c->rest_inv[e * 9], c->volume[e], lam, mu,
                          c->plastic ? &j2 : NULL, c->plastic ? &st->eps_p[e * 6] : NULL,
                          c->plastic ? &st->eqps[e] : NULL, dmg ? &st->damage[e] : NULL,
                          &st->force[e * 12], &de, sample ? &ue : NULL, &torn, nrm, &rel);
        out->diss += de;
        out->strain += ue;
        if (torn) {
            /* The tet's force vanishes from step n on. Against the shadow kinetic
             * energy 1/2 m v(n-1/2).v(n+1/2) the potential it carried at u_n is
             * U_e(u_n) + 1/2 f_e(u_n).(u_n - u_(n-1)), f_e its intact nodal forces:
             * that whole amount leaves as fracture, then the forces are zeroed. */
            double* fe = &st->force[e * 12];
            double half = 0;
            for (int q = 0; q < 4; ++q)
                for (int a = 0; a < 3; ++a) half += fe[q * 3 + a] * st->v[idx[q] * 3 + a];
            out->fracture += rel + 0.5 * c->dt * half;
            for (int k = 0; k < 12; ++k) fe[k] = 0.0;
            if (st->tear_step) st->tear_step[e] = n;
            if (st->tear_normal)
                for (int a = 0; a < 3; ++a) st->tear_normal[e * 3 + a] = nrm[a];
        }
    }
}

/* node pass of step n over nodes [i0, i1): gather in CSR order, loads, contact, pins,
 * drive, damping; every impulse is worked at the mid-step velocity (vo + vn)/2, the
 * form under which the shadow kinetic energy makes the account close exactly */
static void run_nodes(const ora_case* c, ora_state* st, int64_t i0, int64_t i1, int64_t n,
                      int sample, ora_sums* out) {
    const double dt = c->dt;
    const double t = (double)n * dt;
    const double sl = load_scale(c, t);
    const double sd = drive_fraction(c, t + dt);
    for (int64_t i = i0; i < i1; ++i) {
        const uint32_t fl = c->flags ? c->flags[i] : 0u;
        const double m = c->mass[i];
        if ((fl & (ORA_PIN_X | ORA_PIN_Y | ORA_PIN_Z)) == (ORA_PIN_X | ORA_PIN_Y | ORA_PIN_Z)) {
            for (int a = 0; a < 3; ++a) st->v[i * 3 + a] = 0.0;
            continue;
        }
        double f[3] = {0, 0, 0};
        for (uint32_t q = c->adj_offset[i]; q < c->adj_offset[i + 1]; ++q) {
            const uint32_t en = c->adj_entry[q];
            const double* s = &st->force[(size_t)en * 3]; /* slot of (element, corner) */
            for (int a = 0; a < 3; ++a) f[a] += s[a];
        }
        double fe[3] = {0, 0, -c->gravity * m};
        if (c->load)
            for (int a = 0; a < 3; ++a) fe[a] += sl * c->load[i * 3 + a];
        double fc[3] = {0, 0, 0};
        if (c->indenter && (fl & ORA_CONTACT)) {
            double d[3], d2 = 0;
            for (int a = 0; a < 3; ++a) {
                d[a] = c->X[i * 3 + a] + st->u[i * 3 + a] - (c->ind_center[a] + t * c->ind_velocity[a]);
                d2 += d[a] * d[a];
            }
            const double dist = sqrt(d2);
            if (dist < c->ind_radius && dist > 0) {
                const double g = c->ind_radius - dist;
                const double k = c->ind_kappa * m / (dt * dt);
                for (int a = 0; a < 3; ++a) fc[a] = k * g * d[a] / dist;
            }
        }
        if (sample)
            for (int a = 0; a < 3; ++a) out->fcon[a] += fc[a];
        for (int a = 0; a < 3; ++a) {
            double* v = &st->v[i * 3 + a];
            if (fl & (ORA_PIN_X << a)) {
                *v = 0.0;
                continue;
            }
            const double vo = *v;
            double vn;
            if (c->prescribed && (fl & (ORA_DRIVE_X << a))) {
                /* the constraint impulse m (vn - vfree) of the drive (zone.cpp:1134-1145);
                 * the reference works it at the end velocity, which leaves an O(dt)
                 * residual against this account */
                const double vfree = vo + (f[a] + fe[a] + fc[a]) / m * dt;
                vn = (c->prescribed[i * 3 + a] * sd - st->u[i * 3 + a]) / dt;
                const double J = m * (vn - vfree);
                out->w_bnd += J * (vo + vn) * 0.5;
                if (sample) out->half_bnd += J * vn * 0.5;
            } else {
                vn = vo + (f[a] + fe[a] + fc[a]) / m * dt;
                if (c->damping != 1.0) {
                    /* velocity scaling (zone.cpp:1113-1121) as the impulse m (d - 1) v */
                    const double J = m * (1.0 - c->damping) * vn;
                    out->damp += J * (vo + c->damping * vn) * 0.5;
                    if (sample) out->half_damp += J * c->damping * vn * 0.5;
                    vn *= c->damping;
                }
            }
            out->w_ext += fe[a] * (vo + vn) * 0.5 * dt;
            out->w_con += fc[a] * (vo + vn) * 0.5 * dt;
            if (sample) {
                out->half_ext += fe[a] * vn * 0.5 * dt;
                out->half_con += fc[a] * vn * 0.5 * dt;
                out->kinetic += 0.5 * m * vo * vn;
            }
            *v = vn;
            st->u[i * 3 + a] += vn * dt;
        }
    }
}

static void add_sums(ora_sums* a, const ora_sums* b) {
    a->strain += b->strain;
    a->diss += b->diss;
    a->fracture += b->fracture;
    a->w_ext += b->w_ext;
    a->w_con += b->w_con;
    a->w_bnd += b->w_bnd;
    a->damp += b->damp;
    a->kinetic += b->kinetic;
    a->half_ext += b->half_ext;
    a->half_con += b->half_con;
    a->half_bnd += b->half_bnd;
    a->half_damp += b->half_damp;
    for (int k = 0; k < 3; ++k) a->fcon[k] += b->fcon[k];
}

static int sample_now(int64_t n, int64_t sample_every, const ora_energy* rec, int64_t max_rec,
                      const int64_t* nrec) {
    return sample_every > 0 && n % sample_every == 0 && rec && nrec && *nrec < max_rec;
}

/* fold one step's sums into the running account and, on sample steps, a record */
static void close_step(const ora_case* c, ora_state* st, int64_t n, int sample, const ora_sums* s,
                       ora_energy* rec, int64_t* nrec) {
    st->dissipation += s->diss;
    st->fracture += s->fracture;
    st->w_ext += s->w_ext;
    st->w_contact += s->w_con;
    st->w_boundary += s->w_bnd;
    st->damping_loss += s->damp;
    if (sample) {
        ora_energy* r = &rec[(*nrec)++];
        r->step = n;
        r->time = (double)n * c->dt;
        r->strain = s->strain;
        r->kinetic = s->kinetic;
        r->dissipation = st->dissipation;
        r->w_ext = st->w_ext - s->half_ext;
        r->w_contact = st->w_contact - s->half_con;
        r->w_boundary = st->w_boundary - s->half_bnd;
        r->damping_loss = st->damping_loss - s->half_damp;
        r->fracture = st->fracture;
        for (int a = 0; a < 3; ++a) r->contact_force[a] = s->fcon[a];
    }
    st->step = n + 1;
}

void ora_run(const ora_case* c, ora_state* st, int64_t steps, int64_t sample_every,
             ora_energy* rec, int64_t max_rec, int64_t* nrec) {
    for (int64_t k = 0; k < steps; ++k) {
        const int64_t n = st->step;
        const int sample = sample_now(n, sample_every, rec, max_rec, nrec);
        ora_sums s;
        memset(&s, 0, sizeof s);
        run_elements(c, st, 0, c->T, n, sample, &s);
        run_nodes(c, st, 0, c->N, n, sample, &s);
        close_step(c, st, n, sample, &s, rec, nrec);
    }
}

/* The same steps on `threads` threads: a static split of the element and node loops
 * (each element and node computes exactly as above), per-thread sums folded in
 * thread order. Deterministic for a given thread count. */
typedef struct {
    const ora_case* c;
    ora_state* st;
    int64_t steps, sample_every, max_rec;
    ora_energy* rec;
    int64_t* nrec;
    int threads;
    pthread_barrier_t* bar;
    ora_sums* sums; /* [threads] */
} ora_mt_shared;

typedef struct {
    ora_mt_shared* sh;
    int id;
} ora_mt_arg;

static void* ora_mt_worker(void* p) {
    const ora_mt_arg* a = (const ora_mt_arg*)p;
    ora_mt_shared* sh = a->sh;
    const ora_case* c = sh->c;
    const int id = a->id, nt = sh->threads;
    const int64_t e0 = c->T * id / nt, e1 = c->T * (id + 1) / nt;
    const int64_t i0 = c->N * id / nt, i1 = c->N * (id + 1) / nt;
    for (int64_t k = 0; k < sh->steps; ++k) {
        const int64_t n = sh->st->step;
        const int sample = sample_now(n, sh->sample_every, sh->rec, sh->max_rec, sh->nrec);
        ora_sums* mine = &sh->sums[id];
        memset(mine, 0, sizeof *mine);
        run_elements(c, sh->st, e0, e1, n, sample, mine);
        pthread_barrier_wait(sh->bar);
        run_nodes(c, sh->st, i0, i1, n, sample, mine);
        pthread_barrier_wait(sh->bar);
        if (id == 0) {
            ora_sums tot;
            memset(&tot, 0, sizeof tot);
            for (int t = 0; t < nt; ++t) add_sums(&tot, &sh->sums[t]);
            close_step(c, sh->st, n, sample, &tot, sh->rec, sh->nrec);
        }
        pthread_barrier_wait(sh->bar);
    }
    return NULL;
}

void ora_run_mt(const ora_case* c, ora_state* st, int64_t steps, int64_t sample_every,
                ora_energy* rec, int64_t max_rec, int64_t* nrec, int threads) {
    if (threads <= 1 || c->T < 2 * (int64_t)threads) {
        ora_run(c, st, steps, sample_every, rec, max_rec, nrec);
        return;
    }
    pthread_barrier_t bar;
    pthread_barrier_init(&bar, NULL, (unsigned)threads);
    ora_sums* sums = (ora_sums*)calloc((size_t)threads, sizeof(ora_sums));
    ora_mt_arg* args = (ora_mt_arg*)calloc((size_t)threads, sizeof(ora_mt_arg));
    pthread_t* tid = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    ora_mt_shared sh = {c, st, steps, sample_every, max_rec, rec, nrec, threads, &bar, sums};
    for (int t = 0; t < threads; ++t) {
        args[t].sh = &sh;
        args[t].id = t;
    }
    for (int t = 1; t < threads; ++t) pthread_create(&tid[t], NULL, ora_mt_worker, &args[t]);
    ora_mt_worker(&args[0]);
    for (int t = 1; t < threads; ++t) pthread_join(tid[t], NULL);
    pthread_barrier_destroy(&bar);
    free(tid);
    free(args);
    free(sums);
}

/* ======================================================================== */
/* FP32 reference restatement: the float operation order of fem.cpp          */
/* ======================================================================== */
typedef struct {
    float m[9];
} M3f;
#define AF(M, r, c) ((M).m[(c) * 3 + (r)])

static M3f f_mul(M3f a, M3f b) {
    M3f r;
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) {
            float s = 0;
            for (int k = 0; k < 3; ++k) s += AF(a, i, k) * AF(b, k, c);
            AF(r, i, c) = s;
        }
    return r;
}
static M3f f_tr(M3f a) {
    M3f r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) AF(r, i, j) = AF(a, j, i);
    return r;
}
static float f_det(M3f a) {
    return AF(a, 0, 0) * (AF(a, 1, 1) * AF(a, 2, 2) - AF(a, 1, 2) * AF(a, 2, 1)) -
           AF(a, 0, 1) * (AF(a, 1, 0) * AF(a, 2, 2) - AF(a, 1, 2) * AF(a, 2, 0)) +
           AF(a, 0, 2) * (AF(a, 1, 0) * AF(a, 2, 1) - AF(a, 1, 1) * AF(a, 2, 0));
}
static M3f f_inv(M3f a) {
    const float det = f_det(a);
    const float id = fabsf(det) > 1e-30f ? 1.0f / det : 0.0f;
    M3f r;
    AF(r, 0, 0) = (AF(a, 1, 1) * AF(a, 2, 2) - AF(a, 1, 2) * AF(a, 2, 1)) * id;
    AF(r, 0, 1) = (AF(a, 0, 2) * AF(a, 2, 1) - AF(a, 0, 1) * AF(a, 2, 2)) * id;
    AF(r, 0, 2) = (AF(a, 0, 1) * AF(a, 1, 2) - AF(a, 0, 2) * AF(a, 1, 1)) * id;
    AF(r, 1, 0) = (AF(a, 1, 2) * AF(a, 2, 0) - AF(a, 1, 0) * AF(a, 2, 2)) * id;
    AF(r, 1, 1) = (AF(a, 0, 0) * AF(a, 2, 2) - AF(a, 0, 2) * AF(a, 2, 0)) * id;
    AF(r, 1, 2) = (AF(a, 0, 2) * AF(a, 1, 0) - AF(a, 0, 0) * AF(a, 1, 2)) * id;
    AF(r, 2, 0) = (AF(a, 1, 0) * AF(a, 2, 1) - AF(a, 1, 1) * AF(a, 2, 0)) * id;
    AF(r, 2, 1) = (AF(a, 0, 1) * AF(a, 2, 0) - AF(a, 0, 0) * AF(a, 2, 1)) * id;
    AF(r, 2, 2) = (AF(a, 0, 0) * AF(a, 1, 1) - AF(a, 0, 1) * AF(a, 1, 0)) * id;
    return r;
}

void ref32_rest_state(float* position, int64_t N, const uint32_t* index, int64_t T, float density,
                      float* rest_inv, float* volume, float* mass, uint32_t* adj_offset,
                      uint32_t* adj_entry) {
    for (int64_t n = 0; n < N; ++n) mass[n] = 0.0f;
    for (int64_t t = 0; t < T; ++t) {
        const uint32_t* idx = &index[t * 4];
        M3f dm;
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a)
                AF(dm, a, c) = position[idx[c + 1] * 3 + a] - position[idx[0] * 3 + a];
        const float vol = f_det(dm) / 6.0f;
        volume[t] = fabsf(vol);
        const M3f inv = f_inv(dm);
        for (int k = 0; k < 9; ++k) rest_inv[t * 9 + k] = inv.m[k];
        const float quarter = density * volume[t] * 0.25f;
        for (int c = 0; c < 4; ++c) mass[idx[c]] += quarter;
    }
    memset(adj_offset, 0, sizeof(uint32_t) * (size_t)(N + 1));
    for (int64_t t = 0; t < T; ++t)
        for (int c = 0; c < 4; ++c) ++adj_offset[index[t * 4 + c] + 1];
    for (int64_t n = 0; n < N; ++n) adj_offset[n + 1] += adj_offset[n];
    uint32_t* cursor = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)N);
    memcpy(cursor, adj_offset, sizeof(uint32_t) * (size_t)N);
    for (int64_t t = 0; t < T; ++t)
        for (uint32_t c = 0; c < 4; ++c) adj_entry[cursor[index[t * 4 + c]]++] = (uint32_t)t * 4 + c;
    free(cursor);
}

void ref32_tet_forces(const float* p, const uint32_t* index, const float* rest_inv,
                      const float* volume, int64_t t, float lambda, float mu, float out[12]) {
    const uint32_t* idx = &index[t * 4];
    M3f ds, dmi;
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) AF(ds, a, c) = p[idx[c + 1] * 3 + a] - p[idx[0] * 3 + a];
    for (int k = 0; k < 9; ++k) dmi.m[k] = rest_inv[t * 9 + k];
    const M3f f = f_mul(ds, dmi);
    M3f r = f;
    for (int i = 0; i < 4; ++i) {
        const M3f rit = f_tr(f_inv(r));
        for (int k = 0; k < 9; ++k) r.m[k] = 0.5f * (r.m[k] + rit.m[k]);
    }
    const M3f s = f_mul(f_tr(r), f);
    M3f strain;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) AF(strain, i, j) = 0.5f * (AF(s, i, j) + AF(s, j, i));
    for (int i = 0; i < 3; ++i) AF(strain, i, i) -= 1.0f;
    const float trace = AF(strain, 0, 0) + AF(strain, 1, 1) + AF(strain, 2, 2);
    M3f stress;
    for (int k = 0; k < 9; ++k) stress.m[k] = 2.0f * mu * strain.m[k];
    for (int i = 0; i < 3; ++i) AF(stress, i, i) += lambda * trace;
    const M3f piola = f_mul(r, stress);
    const M3f h = f_mul(piola, f_tr(dmi));
    const float scale = -volume[t];
    float f0[3] = {0, 0, 0};
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a) {
            const float v = scale * AF(h, a, c);
            out[(c + 1) * 3 + a] = v;
            f0[a] -= v;
        }
    for (int a = 0; a < 3; ++a) out[a] = f0[a];
}

static void ref32_lame(float E, float nu, float* lam, float* mu) {
    *lam = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
    *mu = E / (2.0f * (1.0f + nu));
}

typedef struct {
    float *position, *velocity, *tet_force;
    const float *mass, *rest_inv, *volume;
    const uint32_t *fixed, *index, *adj_offset, *adj_entry;
    float lambda, mu, dt, gravity, damping;
    int64_t lo, hi; /* element or node range */
} ref32_job;

static void ref32_elements(const ref32_job* j) {
    for (int64_t t = j->lo; t < j->hi; ++t)
        ref32_tet_forces(j->position, j->index, j->rest_inv, j->volume, t, j->lambda, j->mu,
                         &j->tet_force[t * 12]);
}

static void ref32_nodes(const ref32_job* j) {
    for (int64_t n = j->lo; n < j->hi; ++n) {
        if (j->fixed[n]) {
            for (int k = 0; k < 3; ++k) j->velocity[n * 3 + k] = 0.0f;
            continue;
        }
        float force[3] = {0.0f, 0.0f, -j->gravity * j->mass[n]};
        for (uint32_t e = j->adj_offset[n]; e < j->adj_offset[n + 1]; ++e) {
            const uint32_t entry = j->adj_entry[e];
            const float* f = &j->tet_force[(entry >> 2) * 12 + (entry & 3u) * 3];
            for (int k = 0; k < 3; ++k) force[k] += f[k];
        }
        const float inv_mass = 1.0f / j->mass[n];
        for (int k = 0; k < 3; ++k) {
            float v = j->velocity[n * 3 + k] + force[k] * inv_mass * j->dt;
            v *= j->damping;
            j->velocity[n * 3 + k] = v;
            j->position[n * 3 + k] += v * j->dt;
        }
    }
}

void ref32_step(float* position, float* velocity, const float* mass, const uint32_t* fixed,
                int64_t N, const uint32_t* index, const float* rest_inv, const float* volume,
                int64_t T, float* tet_force, const uint32_t* adj_offset,
                const uint32_t* adj_entry, float E, float nu, float dt, float gravity,
                float damping) {
    ref32_job j = {position, velocity, tet_force, mass, rest_inv, volume, fixed, index,
                   adj_offset, adj_entry, 0, 0, dt, gravity, damping, 0, T};
    ref32_lame(E, nu, &j.lambda, &j.mu);
    ref32_elements(&j);
    j.lo = 0;
    j.hi = N;
    ref32_nodes(&j);
}

static void* ref32_elem_thread(void* p) {
    ref32_elements((const ref32_job*)p);
    return NULL;
}
static void* ref32_node_thread(void* p) {
    ref32_nodes((const ref32_job*)p);
    return NULL;
}

void ref32_step_mt(float* position, float* velocity, const float* mass, const uint32_t* fixed,
                   int64_t N, const uint32_t* index, const float* rest_inv, const float* volume,
                   int64_t T, float* tet_force, const uint32_t* adj_offset,
                   const uint32_t* adj_entry, float E, float nu, float dt, float gravity,
                   float damping, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    ref32_job jobs[256];
    ref32_job base = {position, velocity, tet_force, mass, rest_inv, volume, fixed, index,
                      adj_offset, adj_entry, 0, 0, dt, gravity, damping, 0, 0};
    ref32_lame(E, nu, &base.lambda, &base.mu);
    /* element pass: each element writes only its own slot */
    for (int w = 0; w < threads; ++w) {
        jobs[w] = base;
        jobs[w].lo = T * w / threads;
        jobs[w].hi = T * (w + 1) / threads;
        pthread_create(&tid[w], NULL, ref32_elem_thread, &jobs[w]);
    }
    for (int w = 0; w < threads; ++w) pthread_join(tid[w], NULL);
    /* node pass: each node gathers in the fixed CSR order */
    for (int w = 0; w < threads; ++w) {
        jobs[w] = base;
        jobs[w].lo = N * w / threads;
        jobs[w].hi = N * (w + 1) / threads;
        pthread_create(&tid[w], NULL, ref32_node_thread, &jobs[w]);
    }
    for (int w = 0; w < threads; ++w) pthread_join(tid[w], NULL);
}
