// mesh.cpp -- host-side case building: Kuhn box, jitter, stiffened panel, node
// flags and consistent face loads (the shipfem_mesh_* entry points).
//
// makeBoxTetMesh's numbering and tet order are kept exactly (engine/sim/fem.cpp:217-263)
// because the CSR gather order, and therefore the FP32 summation order, follows them.
#include "shipfem_host.hpp"

#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <new>

using namespace shipfem;

namespace {

// Kuhn subdivision on the cell's 0-7 diagonal (fem.cpp:236-237); corner c of a cell
// sits at (i + (c&1), j + ((c>>1)&1), k + ((c>>2)&1)).
constexpr int kKuhn[6][4] = {{0, 1, 3, 7}, {0, 1, 7, 5}, {0, 5, 7, 4},
                             {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 4, 6, 7}};

double det3cols(const double* a, const double* b, const double* c) {
    // det of the matrix with columns a, b, c
    return a[0] * (b[1] * c[2] - c[1] * b[2]) - b[0] * (a[1] * c[2] - c[1] * a[2]) +
           c[0] * (a[1] * b[2] - b[1] * a[2]);
}

// Six tets of one cell, v1/v2 swapped when the orientation is negative (fem.cpp:246-258).
void emitCell(const std::vector<double>& X, const uint32_t corner[8], std::vector<uint32_t>& index) {
    for (const auto& tet : kKuhn) {
        uint32_t v[4];
        for (int c = 0; c < 4; ++c) v[c] = corner[tet[c]];
        double e[3][3];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) e[c][a] = X[v[c + 1] * 3 + a] - X[v[0] * 3 + a];
        if (det3cols(e[0], e[1], e[2]) < 0) std::swap(v[1], v[2]);
        index.insert(index.end(), {v[0], v[1], v[2], v[3]});
    }
}

uint64_t splitmix(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

double shipfem::uniform01(uint64_t seed, uint64_t counter) {
    return static_cast<double>(splitmix(seed, counter) >> 11) * (1.0 / 9007199254740992.0);
}

extern "C" {

shipfem_status shipfem_mesh_box(double lx, double ly, double lz, int32_t nx, int32_t ny,
                                int32_t nz, shipfem_mesh** out) {
    if (!out || nx < 1 || ny < 1 || nz < 1 || !(lx > 0) || !(ly > 0) || !(lz > 0))
        return SHIPFEM_INVALID_ARG;
    const int64_t sx = nx + 1, sy = ny + 1, sz = nz + 1;
    if (sx * sy * sz > 0xFFFFFFF0ll || int64_t(nx) * ny * nz * 6 > 0x7FFFFFF0ll)
        return SHIPFEM_INVALID_ARG;
    auto* m = new (std::nothrow) shipfem_mesh();
    if (!m) return SHIPFEM_OUT_OF_MEMORY;
    m->kind = shipfem_mesh::Box;
    m->box = {lx, ly, lz, nx, ny, nz};
    const size_t N = static_cast<size_t>(sx * sy * sz);
    m->X.resize(N * 3);
    for (int64_t k = 0; k < sz; ++k)
        for (int64_t j = 0; j < sy; ++j)
            for (int64_t i = 0; i < sx; ++i) {
                double* p = &m->X[((k * sy + j) * sx + i) * 3];
                p[0] = lx * static_cast<double>(i) / nx;
                p[1] = ly * static_cast<double>(j) / ny;
                p[2] = lz * static_cast<double>(k) / nz;
            }
    m->index.reserve(static_cast<size_t>(nx) * ny * nz * 24);
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                uint32_t corner[8];
                for (int c = 0; c < 8; ++c)
                    corner[c] = static_cast<uint32_t>(((k + ((c >> 2) & 1)) * sy + (j + ((c >> 1) & 1))) * sx +
                                                      (i + (c & 1)));
                emitCell(m->X, corner, m->index);
            }
    m->flags.assign(N, 0u);
    m->load.assign(N * 3, 0.0);
    *out = m;
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_jitter(shipfem_mesh* m, double amp, uint64_t seed) {
    if (!m || m->kind != shipfem_mesh::Box || !(amp >= 0) || amp >= 0.25) return SHIPFEM_INVALID_ARG;
    const auto& b = m->box;
    const int64_t sx = b.nx + 1, sy = b.ny + 1, sz = b.nz + 1;
    const double h[3] = {b.lx / b.nx, b.ly / b.ny, b.lz / b.nz};
    std::vector<double> X = m->X;
    for (int64_t k = 0; k < sz; ++k)
        for (int64_t j = 0; j < sy; ++j)
            for (int64_t i = 0; i < sx; ++i) {
                const int64_t n = (k * sy + j) * sx + i;
                const bool onface[3] = {i == 0 || i == b.nx, j == 0 || j == b.ny, k == 0 || k == b.nz};
                for (int a = 0; a < 3; ++a) {
                    if (onface[a]) continue;
                    const double r = uniform01(seed, static_cast<uint64_t>(n) * 3 + a);
                    X[n * 3 + a] += amp * h[a] * (2.0 * r - 1.0);
                }
            }
    // refuse a jitter that inverts or flattens a tet
    const size_t T = m->index.size() / 4;
    for (size_t t = 0; t < T; ++t) {
        const uint32_t* v = &m->index[t * 4];
        double e[3][3];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 3; ++a) e[c][a] = X[v[c + 1] * 3 + a] - X[v[0] * 3 + a];
        if (!(det3cols(e[0], e[1], e[2]) > 0)) return SHIPFEM_INVALID_ARG;
    }
    m->X.swap(X);
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_panel(double lx, double ly, double t_plate, int32_t n_stiff,
                                  double web_h, double web_t, int32_t nx, int32_t ny_bay,
                                  int32_t nz_web, shipfem_mesh** out) {
    if (!out || !(lx > 0) || !(ly > 0) || !(t_plate > 0) || n_stiff < 0 || nx < 1 || ny_bay < 1 ||
        (n_stiff > 0 && (nz_web < 1 || !(web_h > 0) || !(web_t > 0))))
        return SHIPFEM_INVALID_ARG;
    // y grid lines: bays split in ny_bay cells, each web one cell of width web_t
    std::vector<double> y{0.0};
    std::vector<int> webCol;
    for (int b = 0; b <= n_stiff; ++b) {
        const double y0 = b == 0 ? 0.0 : ly * b / (n_stiff + 1) + 0.5 * web_t;
        const double y1 = b == n_stiff ? ly : ly * (b + 1) / (n_stiff + 1) - 0.5 * web_t;
        if (!(y1 > y0)) return SHIPFEM_INVALID_ARG;
        for (int j = 1; j <= ny_bay; ++j) y.push_back(j == ny_bay ? y1 : y0 + (y1 - y0) * j / ny_bay);
        if (b < n_stiff) {
            webCol.push_back(static_cast<int>(y.size()) - 1);
            y.push_back(ly * (b + 1) / (n_stiff + 1) + 0.5 * web_t);
        }
    }
    const int64_t ny = static_cast<int64_t>(y.size()) - 1, nzc = 1 + (n_stiff > 0 ? nz_web : 0);
    const int64_t sx = nx + 1, sy = ny + 1, sz = nzc + 1;
    std::vector<uint8_t> colweb(static_cast<size_t>(ny), 0);
    for (int c : webCol) colweb[c] = 1;
    auto active = [&](int64_t j, int64_t k) { return k == 0 || colweb[j]; };
    std::vector<int64_t> id(static_cast<size_t>(sx * sy * sz), -1);
    int64_t cells = 0;
    for (int64_t k = 0; k < nzc; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                if (!active(j, k)) continue;
                ++cells;
                for (int c = 0; c < 8; ++c)
                    id[((k + ((c >> 2) & 1)) * sy + (j + ((c >> 1) & 1))) * sx + (i + (c & 1))] = 0;
            }
    if (cells * 6 > 0x7FFFFFF0ll) return SHIPFEM_INVALID_ARG;
    auto* m = new (std::nothrow) shipfem_mesh();
    if (!m) return SHIPFEM_OUT_OF_MEMORY;
    m->kind = shipfem_mesh::Panel;
    int64_t nn = 0;
    for (int64_t k = 0; k < sz; ++k)
        for (int64_t j = 0; j < sy; ++j)
            for (int64_t i = 0; i < sx; ++i) {
                int64_t& q = id[(k * sy + j) * sx + i];
                if (q < 0) continue;
                q = nn++;
                m->X.push_back(lx * static_cast<double>(i) / nx);
                m->X.push_back(y[j]);
                m->X.push_back(k == 0 ? 0.0 : t_plate + web_h * static_cast<double>(k - 1) / nz_web);
            }
    m->index.reserve(static_cast<size_t>(cells) * 24);
    for (int64_t k = 0; k < nzc; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                if (!active(j, k)) continue;
                uint32_t corner[8];
                for (int c = 0; c < 8; ++c)
                    corner[c] = static_cast<uint32_t>(
                        id[((k + ((c >> 2) & 1)) * sy + (j + ((c >> 1) & 1))) * sx + (i + (c & 1))]);
                emitCell(m->X, corner, m->index);
            }
    m->flags.assign(static_cast<size_t>(nn), 0u);
    m->load.assign(static_cast<size_t>(nn) * 3, 0.0);
    *out = m;
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_counts(const shipfem_mesh* m, uint32_t* nodes, uint32_t* tets) {
    if (!m) return SHIPFEM_INVALID_ARG;
    if (nodes) *nodes = static_cast<uint32_t>(m->X.size() / 3);
    if (tets) *tets = static_cast<uint32_t>(m->index.size() / 4);
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_arrays(const shipfem_mesh* m, const double** rest_position,
                                   const uint32_t** index, const uint32_t** node_flags,
                                   const double** load) {
    if (!m) return SHIPFEM_INVALID_ARG;
    if (rest_position) *rest_position = m->X.data();
    if (index) *index = m->index.data();
    if (node_flags) *node_flags = m->flags.data();
    if (load) *load = m->load.data();
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_flag_plane(shipfem_mesh* m, int32_t axis, double value, double tol,
                                       uint32_t flags) {
    if (!m || axis < 0 || axis > 2 || !(tol >= 0)) return SHIPFEM_INVALID_ARG;
    const size_t N = m->X.size() / 3;
    for (size_t n = 0; n < N; ++n)
        if (std::fabs(m->X[n * 3 + axis] - value) <= tol) m->flags[n] |= flags;
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_flag_disc(shipfem_mesh* m, int32_t axis, double value, double tol,
                                      double c0, double c1, double radius, uint32_t flags) {
    if (!m || axis < 0 || axis > 2 || !(tol >= 0) || !(radius >= 0)) return SHIPFEM_INVALID_ARG;
    const int a0 = (axis + 1) % 3, a1 = (axis + 2) % 3;
    // (c0, c1) are the coordinates along the lower- and higher-numbered other axes
    const int lo = std::min(a0, a1), hi = std::max(a0, a1);
    const size_t N = m->X.size() / 3;
    for (size_t n = 0; n < N; ++n) {
        const double* p = &m->X[n * 3];
        if (std::fabs(p[axis] - value) > tol) continue;
        const double d0 = p[lo] - c0, d1 = p[hi] - c1;
        if (d0 * d0 + d1 * d1 <= radius * radius) m->flags[n] |= flags;
    }
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_traction(shipfem_mesh* m, int32_t axis, double value, double tol,
                                     const double traction[3]) {
    if (!m || axis < 0 || axis > 2 || !traction || !(tol >= 0)) return SHIPFEM_INVALID_ARG;
    static const int kFaces[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    const size_t T = m->index.size() / 4;
    for (size_t t = 0; t < T; ++t) {
        const uint32_t* v = &m->index[t * 4];
        for (const auto& f : kFaces) {
            bool on = true;
            for (int q = 0; q < 3; ++q)
                if (std::fabs(m->X[v[f[q]] * 3 + axis] - value) > tol) on = false;
            if (!on) continue;
            const double* a = &m->X[v[f[0]] * 3];
            const double* b = &m->X[v[f[1]] * 3];
            const double* c = &m->X[v[f[2]] * 3];
            const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
            const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
            const double cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                                  e1[0] * e2[1] - e1[1] * e2[0]};
            const double area = 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            for (int q = 0; q < 3; ++q) {
                for (int k = 0; k < 3; ++k) m->load[v[f[q]] * 3 + k] += traction[k] * area / 3.0;
                m->flags[v[f[q]]] |= SHIPFEM_LOADED;
            }
        }
    }
    return SHIPFEM_OK;
}

shipfem_status shipfem_mesh_fill_case(const shipfem_mesh* m, shipfem_case_desc* d) {
    if (!m || !d) return SHIPFEM_INVALID_ARG;
    d->node_count = static_cast<uint32_t>(m->X.size() / 3);
    d->tet_count = static_cast<uint32_t>(m->index.size() / 4);
    d->rest_position = m->X.data();
    d->index = m->index.data();
    d->node_flags = m->flags.data();
    bool anyLoad = false;
    for (double v : m->load)
        if (v != 0.0) { anyLoad = true; break; }
    d->load = anyLoad ? m->load.data() : nullptr;
    return SHIPFEM_OK;
}

void shipfem_mesh_destroy(shipfem_mesh* m) { delete m; }

double shipfem_edge_timestep(const double* X, const uint32_t* index, uint32_t tet_count,
                             double E, double rho, double safety) {
    if (!X || !index || !(E > 0) || !(rho > 0)) return 0.0;
    double shortest = INFINITY;
    for (uint32_t t = 0; t < tet_count; ++t) {
        const uint32_t* v = &index[static_cast<size_t>(t) * 4];
        for (int a = 0; a < 4; ++a)
            for (int b = a + 1; b < 4; ++b) {
                double d2 = 0;
                for (int k = 0; k < 3; ++k) {
                    const double d = X[v[a] * 3 + k] - X[v[b] * 3 + k];
                    d2 += d * d;
                }
                shortest = std::min(shortest, std::sqrt(d2));
            }
    }
    return safety * shortest / std::sqrt(E / rho);
}

double shipfem_regularised_failure_strain(double uniform, double fracture, double length,
                                          double thickness) {
    // an element no larger than the plate is thick resolves the neck itself; degenerate
    // sizes fall to the unregularised end (plasticity.cpp:158-169)
    double share = 1.0;
    if (length > 0.0 && thickness > 0.0) share = std::min(1.0, thickness / length);
    return uniform + (fracture - uniform) * share;
}

shipfem_status shipfem_torn_groups(const double* X, const uint32_t* index, uint32_t T,
                                   const uint32_t* group, const float* damage, double fraction,
                                   uint32_t* out, uint32_t max, uint32_t* count) {
    if (!X || !index || !group || !damage || (!out && max)) return SHIPFEM_INVALID_ARG;
    // per group: total and torn rest volume, then the groups past `fraction`
    std::map<uint32_t, std::pair<double, double>> vol;
    for (uint32_t t = 0; t < T; ++t) {
        const double* p0 = &X[3ull * index[4ull * t]];
        double e[3][3];
        for (int c = 1; c < 4; ++c)
            for (int a = 0; a < 3; ++a) e[c - 1][a] = X[3ull * index[4ull * t + c] + a] - p0[a];
        const double det = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
                           e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                           e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
        auto& g = vol[group[t]];
        g.first += std::fabs(det) / 6.0;
        if (damage[t] >= 1.0f) g.second += std::fabs(det) / 6.0;
    }
    uint32_t n = 0;
    for (const auto& [id, v] : vol) {
        if (!(v.second > 0.0) || v.second < fraction * v.first) continue;
        if (out && n < max) out[n] = id;
        ++n;
    }
    if (count) *count = n;
    return SHIPFEM_OK;
}

}  // extern "C"
