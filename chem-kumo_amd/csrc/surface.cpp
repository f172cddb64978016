// Molecular surfaces: density isosurfaces, the fields mapped on them and
// their statistics (include/chem/kumo/surface.hpp, src/surface.cpp).
//
// The density on the construction grid comes from a GPU grid evaluation, and
// each round of the vertex placement evaluates the density and its gradient
// at every unconverged vertex in one GPU point evaluation, as do the fields
// mapped on the finished mesh. The marching-tetrahedra polygonisation walks
// the cells in Field order and numbers the vertices in the order their edges
// are first met, and the statistics sum over the triangles in order, as the
// reference does, so that the meshes and the statistics are the reference's.
#include "surface.hpp"

#include "runtime.hpp"

#include <algorithm>
#include <cctype>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <limits>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>

namespace kumo {
namespace {

constexpr double kRootTolerance = 1e-10; // bohr, along a vertex's edge
constexpr int kMaxRootIterations = 200;
constexpr double kHartreeInKcal = 627.5094740631; // include/chem/kumo/surface.hpp:17
constexpr double kMinimumIsovalue = 1e-8;         // include/chem/kumo/surface.hpp:90

struct V3 {
    double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }

// Point (i, j, k) of a grid, origin + steps (i, j, k), with the arithmetic
// order of the GPU point source.
V3 grid_point(const kumo_hip_grid& g, long long linear) {
    const long long ny = g.counts[1], nz = g.counts[2];
    const double fi = static_cast<double>(linear / (ny * nz));
    const double fj = static_cast<double>((linear / nz) % ny);
    const double fk = static_cast<double>(linear % nz);
    const double* s = g.steps;
    return {g.origin[0] + ((s[0] * fi + s[3] * fj) + s[6] * fk), g.origin[1] + ((s[1] * fi + s[4] * fj) + s[7] * fk),
            g.origin[2] + ((s[2] * fi + s[5] * fj) + s[8] * fk)};
}

long long grid_linear(const kumo_hip_grid& g, int i, int j, int k) {
    return (static_cast<long long>(i) * g.counts[1] + j) * g.counts[2] + k;
}

// Corners of a cell by 3-bit masks (bit 0: i + 1, bit 1: j + 1, bit 2: k + 1)
// and the six tetrahedra around the main diagonal 0-7, one per permutation of
// the axes (src/surface.cpp:64-73).
constexpr int kTetra[6][4] = {{0, 1, 3, 7}, {0, 1, 5, 7}, {0, 2, 3, 7}, {0, 2, 6, 7}, {0, 4, 5, 7}, {0, 4, 6, 7}};

struct CrossedEdge {
    long long inside, outside; // linear indices of the end points
    std::array<int, 3> cell;   // grid indices of the lower end
};

// Where the isosurface crosses one tetrahedron: the edges (inside corner,
// outside corner) of its 3 or 4 polygon corners in cyclic order, oriented
// from the inside corners to the outside ones (tetrahedron_polygon,
// src/surface.cpp:86-126).
int tetra_polygon(const int* masks, const bool* inside, const V3* pos, int ends[4][2]) {
    int in[4], out[4], nin = 0, nout = 0;
    for (int c = 0; c < 4; ++c) {
        if (inside[masks[c]])
            in[nin++] = masks[c];
        else
            out[nout++] = masks[c];
    }
    if (nin == 0 || nout == 0) return 0;
    int count;
    if (nin == 1) {
        const int e[3][2] = {{in[0], out[0]}, {in[0], out[1]}, {in[0], out[2]}};
        for (int c = 0; c < 3; ++c) ends[c][0] = e[c][0], ends[c][1] = e[c][1];
        count = 3;
    } else if (nout == 1) {
        const int e[3][2] = {{in[0], out[0]}, {in[1], out[0]}, {in[2], out[0]}};
        for (int c = 0; c < 3; ++c) ends[c][0] = e[c][0], ends[c][1] = e[c][1];
        count = 3;
    } else {
        const int e[4][2] = {{in[0], out[0]}, {in[0], out[1]}, {in[1], out[1]}, {in[1], out[0]}};
        for (int c = 0; c < 4; ++c) ends[c][0] = e[c][0], ends[c][1] = e[c][1];
        count = 4;
    }
    // Orientation from the edge midpoints (they depend on the cell only).
    V3 mid[4];
    for (int c = 0; c < count; ++c) mid[c] = 0.5 * (pos[ends[c][0]] + pos[ends[c][1]]);
    const V3 normal = count == 3 ? cross(mid[1] - mid[0], mid[2] - mid[0]) : cross(mid[2] - mid[0], mid[3] - mid[1]);
    V3 outward{0, 0, 0};
    for (int c = 0; c < nout; ++c) outward = outward + (1.0 / nout) * pos[out[c]];
    for (int c = 0; c < nin; ++c) outward = outward - (1.0 / nin) * pos[in[c]];
    if (dot(normal, outward) < 0)
        for (int a = 0, b = count - 1; a < b; ++a, --b) std::swap(ends[a], ends[b]);
    return count;
}

struct Polygons {
    std::vector<CrossedEdge> edges;          // one per vertex, in order of first use
    std::vector<std::array<long long, 4>> poly; // vertex ids; [3] = -1 for a triangle
};

// Marching tetrahedra over the cells, k fastest (polygonize, src/surface.cpp:150-199).
Polygons polygonize(const kumo_hip_grid& g, const std::vector<double>& f, double iso) {
    Polygons r;
    std::unordered_map<std::uint64_t, long long> ids; // (lower end, direction) -> vertex
    for (int i = 0; i + 1 < g.counts[0]; ++i)
        for (int j = 0; j + 1 < g.counts[1]; ++j)
            for (int k = 0; k + 1 < g.counts[2]; ++k) {
                long long corner[8];
                bool inside[8];
                int nin = 0;
                for (int m = 0; m < 8; ++m) {
                    corner[m] = grid_linear(g, i + (m & 1), j + ((m >> 1) & 1), k + (m >> 2));
                    inside[m] = f[static_cast<std::size_t>(corner[m])] >= iso;
                    nin += inside[m];
                }
                if (nin == 0 || nin == 8) continue;
                V3 pos[8];
                for (int m = 0; m < 8; ++m) pos[m] = grid_point(g, corner[m]);
                // The corner masks of a tetrahedron are nested bit sets, so
                // on each of its edges the smaller mask is the lower end and
                // their XOR the direction: (lower end, direction) names the
                // edge in every cell that shares it.
                const auto vertex = [&](int a, int b) {
                    const int lo = std::min(a, b), hi = std::max(a, b);
                    const std::uint64_t key = static_cast<std::uint64_t>(corner[lo]) * 8 + static_cast<std::uint64_t>(hi ^ lo);
                    const auto [it, added] = ids.try_emplace(key, static_cast<long long>(r.edges.size()));
                    if (added)
                        r.edges.push_back({corner[a], corner[b], {i + (lo & 1), j + ((lo >> 1) & 1), k + (lo >> 2)}});
                    return it->second;
                };
                for (const auto& masks : kTetra) {
                    int ends[4][2];
                    const int count = tetra_polygon(masks, inside, pos, ends);
                    if (count == 0) continue;
                    std::array<long long, 4> p{-1, -1, -1, -1};
                    for (int c = 0; c < count; ++c) p[static_cast<std::size_t>(c)] = vertex(ends[c][0], ends[c][1]);
                    r.poly.push_back(p);
                }
            }
    return r;
}

// Safeguarded Newton steps along each vertex's edge, every unconverged
// vertex in one evaluation per round (place_vertices, src/surface.cpp:204-253).
std::vector<double> place_vertices(const kumo_hip_grid& g, const std::vector<double>& f, double iso,
                                   const std::vector<CrossedEdge>& edges, const SurfaceFunction& fn) {
    const std::size_t nv = edges.size();
    std::vector<double> t(nv), lo(nv, 0.0), hi(nv, 1.0);
    for (std::size_t v = 0; v < nv; ++v) {
        const double a = f[static_cast<std::size_t>(edges[v].inside)], b = f[static_cast<std::size_t>(edges[v].outside)];
        t[v] = std::clamp((a - iso) / (a - b), 0.0, 1.0);
    }
    std::vector<std::size_t> active(nv);
    std::iota(active.begin(), active.end(), std::size_t{0});
    std::vector<double> pts, values, grads;
    for (int it = 0; !active.empty(); ++it) {
        if (it == kMaxRootIterations) runtime("isosurface vertices did not converge");
        pts.resize(3 * active.size());
        for (std::size_t p = 0; p < active.size(); ++p) {
            const std::size_t v = active[p];
            const V3 a = grid_point(g, edges[v].inside), b = grid_point(g, edges[v].outside);
            const V3 x = a + t[v] * (b - a);
            pts[3 * p] = x.x, pts[3 * p + 1] = x.y, pts[3 * p + 2] = x.z;
        }
        values.assign(active.size(), 0.0);
        grads.assign(3 * active.size(), 0.0);
        fn(pts, values, grads);
        std::vector<std::size_t> next;
        for (std::size_t p = 0; p < active.size(); ++p) {
            const std::size_t v = active[p];
            const double r = values[p] - iso;
            const V3 gr{grads[3 * p], grads[3 * p + 1], grads[3 * p + 2]};
            if (!std::isfinite(r) || !std::isfinite(gr.x) || !std::isfinite(gr.y) || !std::isfinite(gr.z))
                runtime("isosurface function is not finite");
            if (r == 0.0) continue;
            (r > 0.0 ? lo[v] : hi[v]) = t[v];
            const V3 a = grid_point(g, edges[v].inside), b = grid_point(g, edges[v].outside);
            const double length = norm(b - a);
            double step = t[v] - r / dot(gr, b - a);
            if (!(step > lo[v] && step < hi[v])) step = 0.5 * (lo[v] + hi[v]);
            const double moved = std::abs(step - t[v]) * length;
            t[v] = step;
            if (moved > kRootTolerance && (hi[v] - lo[v]) * length > kRootTolerance) next.push_back(v);
        }
        active = std::move(next);
    }
    std::vector<double> xyz(3 * nv);
    for (std::size_t v = 0; v < nv; ++v) {
        const V3 a = grid_point(g, edges[v].inside), b = grid_point(g, edges[v].outside);
        const V3 x = a + t[v] * (b - a);
        xyz[3 * v] = x.x, xyz[3 * v + 1] = x.y, xyz[3 * v + 2] = x.z;
    }
    return xyz;
}

V3 vertex_at(const double* xyz, long long v) { return {xyz[3 * v], xyz[3 * v + 1], xyz[3 * v + 2]}; }

// Clips a triangle with vertex values v (linear inside) to the part >= level
// (upper) or < level, split into triangles from its first corner; calls
// piece(area, w0, w1, w2) with the interpolated values at their corners
// (clip, src/surface.cpp:377-403).
template <typename Piece>
void clip(const V3 (&x)[3], const double (&v)[3], double level, bool upper, const Piece& piece) {
    V3 px[4];
    double pv[4];
    int n = 0;
    const auto keep = [&](double value) { return upper ? value >= level : value < level; };
    for (int i = 0; i < 3; ++i) {
        const int j = (i + 1) % 3;
        const bool ki = keep(v[i]), kj = keep(v[j]);
        if (ki) {
            px[n] = x[i];
            pv[n++] = v[i];
        }
        if (ki != kj) {
            const double s = (level - v[i]) / (v[j] - v[i]);
            px[n] = x[i] + s * (x[j] - x[i]);
            pv[n++] = level;
        }
    }
    for (int i = 1; i + 1 < n; ++i) piece(0.5 * norm(cross(px[i] - px[0], px[i + 1] - px[0])), pv[0], pv[i], pv[i + 1]);
}

// Integral of (w - s)^2 over a triangle of the given area, w linear with
// corner values w0, w1, w2.
double square_integral(double area, double w0, double w1, double w2, double s) {
    const double a = w0 - s, b = w1 - s, c = w2 - s;
    return area / 6.0 * (a * a + b * b + c * c + a * b + b * c + c * a);
}

} // namespace

bool isosurface_reaches_boundary(const kumo_hip_grid& g, const std::vector<double>& f, double iso) {
    const int nx = g.counts[0], ny = g.counts[1], nz = g.counts[2];
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            const bool face = i == 0 || i == nx - 1 || j == 0 || j == ny - 1;
            for (int k = 0; k < nz; k += face || nz == 1 ? 1 : nz - 1)
                if (f[static_cast<std::size_t>(grid_linear(g, i, j, k))] >= iso) return true;
        }
    return false;
}

Mesh extract_isosurface(const kumo_hip_grid& g, const std::vector<double>& f, double iso, const SurfaceFunction& fn) {
    const long long n = validated_point_count(g);
    if (std::min({g.counts[0], g.counts[1], g.counts[2]}) < 2)
        invalid("isosurface grid needs at least two points along each axis");
    if (static_cast<long long>(f.size()) != n) invalid("isosurface samples do not match the grid");
    if (!std::isfinite(iso)) invalid("isovalue must be finite");
    if (!std::all_of(f.begin(), f.end(), [](double v) { return std::isfinite(v); }))
        invalid("isosurface samples must be finite");
    if (isosurface_reaches_boundary(g, f, iso)) invalid("the isosurface reaches the grid boundary, so it would not be closed");
    const Polygons p = polygonize(g, f, iso);
    Mesh m;
    m.xyz = place_vertices(g, f, iso, p.edges, fn);
    m.cell.reserve(p.edges.size());
    for (const auto& e : p.edges) m.cell.push_back(e.cell);
    // Quadrilaterals split along their shorter diagonal (triangulate, src/surface.cpp:257-275).
    for (const auto& q : p.poly) {
        if (q[3] < 0) {
            m.tri.insert(m.tri.end(), {q[0], q[1], q[2]});
            continue;
        }
        const V3 d02 = vertex_at(m.xyz.data(), q[0]) - vertex_at(m.xyz.data(), q[2]);
        const V3 d13 = vertex_at(m.xyz.data(), q[1]) - vertex_at(m.xyz.data(), q[3]);
        if (dot(d02, d02) <= dot(d13, d13))
            m.tri.insert(m.tri.end(), {q[0], q[1], q[2], q[0], q[2], q[3]});
        else
            m.tri.insert(m.tri.end(), {q[0], q[1], q[3], q[1], q[2], q[3]});
    }
    return m;
}

double mesh_area(const double* xyz, const long long* tri, long long nt) {
    double total = 0.0;
    for (long long t = 0; t < nt; ++t) {
        const V3 a = vertex_at(xyz, tri[3 * t]);
        total += 0.5 * norm(cross(vertex_at(xyz, tri[3 * t + 1]) - a, vertex_at(xyz, tri[3 * t + 2]) - a));
    }
    return total;
}

double mesh_volume(const double* xyz, long long nv, const long long* tri, long long nt) {
    if (nv == 0) return 0.0;
    V3 c{0, 0, 0};
    for (long long v = 0; v < nv; ++v) c = c + vertex_at(xyz, v);
    c = (1.0 / static_cast<double>(nv)) * c;
    double total = 0.0;
    for (long long t = 0; t < nt; ++t)
        total += dot(vertex_at(xyz, tri[3 * t]) - c,
                     cross(vertex_at(xyz, tri[3 * t + 1]) - c, vertex_at(xyz, tri[3 * t + 2]) - c));
    return total / 6.0;
}

SurfaceStats surface_statistics(const double* xyz, long long nv, const long long* tri, long long nt,
                                const double* values, double polar) {
    for (long long v = 0; v < nv; ++v)
        if (!std::isfinite(values[v])) invalid("surface values must be finite");
    if (!std::isfinite(polar) || polar < 0.0) invalid("polar threshold must be finite and nonnegative");
    for (long long q = 0; q < 3 * nt; ++q)
        if (tri[q] < 0 || tri[q] >= nv) invalid("surface triangle refers to a missing vertex");
    SurfaceStats r;
    kumo_hip_surface_stats& s = r.s;
    const auto corners = [&](long long t, V3 (&x)[3], double (&v)[3]) {
        for (int c = 0; c < 3; ++c) {
            x[c] = vertex_at(xyz, tri[3 * t + c]);
            v[c] = values[tri[3 * t + c]];
        }
    };
    const auto mean = [](double a, double b, double c) { return (a + b + c) / 3.0; };
    double pos_int = 0.0, neg_int = 0.0, above = 0.0, below = 0.0;
    for (long long t = 0; t < nt; ++t) {
        V3 x[3];
        double v[3];
        corners(t, x, v);
        clip(x, v, 0.0, true, [&](double a, double w0, double w1, double w2) {
            s.positive_area += a;
            pos_int += a * mean(w0, w1, w2);
        });
        clip(x, v, 0.0, false, [&](double a, double w0, double w1, double w2) {
            s.negative_area += a;
            neg_int += a * mean(w0, w1, w2);
        });
        clip(x, v, polar, true, [&](double a, double, double, double) { above += a; });
        clip(x, v, -polar, false, [&](double a, double, double, double) { below += a; });
    }
    s.area = s.positive_area + s.negative_area;
    const double integral = pos_int + neg_int;
    if (!(s.area > 0.0) || !std::isfinite(s.area)) invalid("surface area must be positive");
    const double nan = std::numeric_limits<double>::quiet_NaN();
    s.average = integral / s.area;
    s.positive_average = s.positive_area > 0.0 ? pos_int / s.positive_area : nan;
    s.negative_average = s.negative_area > 0.0 ? neg_int / s.negative_area : nan;
    s.polarity_index = (pos_int - neg_int) / s.area;
    s.nonpolar_area = std::max(0.0, s.area - above - below);
    double pos_sq = 0.0, neg_sq = 0.0, dev = 0.0;
    for (long long t = 0; t < nt; ++t) {
        V3 x[3];
        double v[3];
        corners(t, x, v);
        if (s.positive_area > 0.0)
            clip(x, v, 0.0, true, [&](double a, double w0, double w1, double w2) {
                pos_sq += square_integral(a, w0, w1, w2, s.positive_average);
            });
        if (s.negative_area > 0.0)
            clip(x, v, 0.0, false, [&](double a, double w0, double w1, double w2) {
                neg_sq += square_integral(a, w0, w1, w2, s.negative_average);
            });
        clip(x, v, s.average, true, [&](double a, double w0, double w1, double w2) { dev += a * (mean(w0, w1, w2) - s.average); });
        clip(x, v, s.average, false, [&](double a, double w0, double w1, double w2) { dev += a * (s.average - mean(w0, w1, w2)); });
    }
    s.positive_variance = s.positive_area > 0.0 ? pos_sq / s.positive_area : 0.0;
    s.negative_variance = s.negative_area > 0.0 ? neg_sq / s.negative_area : 0.0;
    s.total_variance = s.positive_variance + s.negative_variance;
    s.balance = s.total_variance > 0.0 ? s.positive_variance * s.negative_variance / (s.total_variance * s.total_variance) : nan;
    s.charge_separation = dev / s.area;
    const auto [lowest, highest] = std::minmax_element(values, values + nv);
    s.minimum_vertex = lowest - values;
    s.maximum_vertex = highest - values;
    s.minimum = *lowest;
    s.maximum = *highest;
    // Local extrema: below (above) every vertex within two triangle edges.
    std::vector<std::vector<long long>> nb(static_cast<std::size_t>(nv));
    for (long long t = 0; t < nt; ++t)
        for (int a = 0; a < 3; ++a) {
            auto& list = nb[static_cast<std::size_t>(tri[3 * t + a])];
            list.push_back(tri[3 * t + (a + 1) % 3]);
            list.push_back(tri[3 * t + (a + 2) % 3]);
        }
    for (auto& list : nb) {
        std::sort(list.begin(), list.end());
        list.erase(std::unique(list.begin(), list.end()), list.end());
    }
    const auto extremum = [&](long long i, auto beyond) {
        const auto& ni = nb[static_cast<std::size_t>(i)];
        if (ni.empty()) return false;
        for (long long j : ni) {
            if (j == i) continue;
            if (!beyond(values[i], values[j])) return false;
            for (long long k : nb[static_cast<std::size_t>(j)])
                if (k != i && !beyond(values[i], values[k])) return false;
        }
        return true;
    };
    for (long long i = 0; i < nv; ++i) {
        if (extremum(i, std::less<double>{})) r.local_minima.push_back(i);
        if (extremum(i, std::greater<double>{})) r.local_maxima.push_back(i);
    }
    std::sort(r.local_minima.begin(), r.local_minima.end(),
              [&](long long a, long long b) { return values[a] < values[b] || (values[a] == values[b] && a < b); });
    std::sort(r.local_maxima.begin(), r.local_maxima.end(),
              [&](long long a, long long b) { return values[a] > values[b] || (values[a] == values[b] && a < b); });
    s.local_minima = static_cast<int64_t>(r.local_minima.size());
    s.local_maxima = static_cast<int64_t>(r.local_maxima.size());
    return r;
}

} // namespace kumo

// ------------------------------------------------------------ the C-ABI object

struct kumo_hip_surface {
    kumo::Mesh mesh;
    std::vector<std::vector<double>> fields;
    std::vector<kumo::SurfaceStats> stats;
    int backend_used = -1;
};

namespace kumo {
namespace {

// Grid::around (src/grid.cpp:15-38).
kumo_hip_grid grid_around(const kumo_hip_molecule* mol, double margin, double spacing) {
    if (mol->atoms.empty() || !std::isfinite(margin) || margin < 0 || !std::isfinite(spacing) || spacing <= 0)
        invalid("invalid grid parameters");
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) lo[k] = hi[k] = mol->atoms.front().xyz[k];
    for (const auto& a : mol->atoms) {
        if (a.z <= 0 || !std::isfinite(a.xyz[0]) || !std::isfinite(a.xyz[1]) || !std::isfinite(a.xyz[2]))
            invalid("invalid atom in grid");
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], a.xyz[k]);
            hi[k] = std::max(hi[k], a.xyz[k]);
        }
    }
    kumo_hip_grid g{};
    for (int k = 0; k < 3; ++k) {
        g.origin[k] = lo[k] - margin;
        const double count = std::ceil((hi[k] - lo[k] + 2 * margin) / spacing) + 1;
        if (!std::isfinite(count) || count > 2147483647.0) invalid("grid too large");
        g.counts[k] = std::max(2, static_cast<int>(count));
        g.steps[3 * k + k] = spacing;
    }
    return g;
}

kumo_hip_options defaults() {
    kumo_hip_options o{};
    kumo_hip_default_options(&o);
    return o;
}

// Total density and gradient at points (density_derivatives_shared with the
// default block and screening, AO-P: SurfaceAccess::total_density,
// src/surface.cpp:33-42), on the GPU.
SurfaceFunction density_function(kumo_hip_context* ctx, const kumo_hip_molecule* mol) {
    return [ctx, mol](const std::vector<double>& pts, std::vector<double>& values, std::vector<double>& grads) {
        const long long n = static_cast<long long>(values.size());
        const kumo_hip_property props[2] = {{0, 0, 0, 0}, {1, 0, 0, 0}};
        std::vector<double> gx(static_cast<std::size_t>(n)), gy(gx.size()), gz(gx.size());
        double* outs[4] = {values.data(), gx.data(), gy.data(), gz.data()};
        kumo_hip_options o = defaults();
        o.backend = KUMO_HIP_AO_P;
        PointsArg pa;
        pa.points = pts.data();
        pa.n = n;
        evaluate(ctx, mol, pa, props, 2, o, outs, nullptr);
        for (long long p = 0; p < n; ++p) {
            grads[3 * p] = gx[static_cast<std::size_t>(p)];
            grads[3 * p + 1] = gy[static_cast<std::size_t>(p)];
            grads[3 * p + 2] = gz[static_cast<std::size_t>(p)];
        }
    };
}

// density_surface (src/surface.cpp:341-369): the grid grows from a 4-bohr
// margin in 2-bohr steps until the isosurface stays inside it.
Mesh density_surface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double iso, double spacing,
                     kumo_hip_grid* used = nullptr) {
    if (!std::isfinite(iso) || iso < kMinimumIsovalue) invalid("density isovalue must be finite and at least 1e-8 e/bohr^3");
    if (!std::isfinite(spacing) || spacing <= 0.0) invalid("surface grid spacing must be positive and finite");
    const SurfaceFunction fn = density_function(ctx, mol);
    for (double margin = 4.0; margin <= 16.0; margin += 2.0) {
        const kumo_hip_grid g = grid_around(mol, margin, spacing);
        const long long n = validated_point_count(g);
        std::vector<double> rho(static_cast<std::size_t>(n));
        const kumo_hip_property p{0, 0, 0, 0};
        double* out = rho.data();
        kumo_hip_options o = defaults();
        o.backend = KUMO_HIP_AO_P; // densities_blocked: the AO-P grid path
        PointsArg pa;
        pa.grid = &g;
        evaluate(ctx, mol, pa, &p, 1, o, &out, nullptr);
        if (isosurface_reaches_boundary(g, rho, iso)) continue;
        if (used) *used = g;
        return extract_isosurface(g, rho, iso, fn);
    }
    runtime("the density isosurface extends more than 16 bohr beyond the atoms; use a larger isovalue");
    return {};
}

// Values of one-component properties at the mesh vertices (GPU point path).
std::vector<std::vector<double>> at_vertices(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const Mesh& m,
                                             const std::vector<kumo_hip_property>& props, const kumo_hip_options& o,
                                             int32_t* backend) {
    const long long n = m.vertices();
    std::vector<std::vector<double>> v(props.size(), std::vector<double>(static_cast<std::size_t>(n)));
    std::vector<double*> outs;
    for (auto& x : v) outs.push_back(x.data());
    PointsArg pa;
    pa.points = m.xyz.data();
    pa.n = n;
    evaluate(ctx, mol, pa, props.data(), props.size(), o, outs.data(), backend);
    return v;
}

// average_local_ionization_energy (src/surface.cpp:563-592): the ALIE at
// points, undefined where the density of the occupied orbitals vanishes.
std::vector<double> alie_at(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* pts, long long n) {
    for (long long q = 0; q < 3 * n; ++q)
        if (!std::isfinite(pts[q])) invalid("ionization energy points must be finite");
    // The orbital checks run before anything else (an evaluation of no points).
    const kumo_hip_property alie{19, 0, 0, 0}, rho{0, 0, 0, 0};
    const kumo_hip_property props[2] = {alie, rho};
    kumo_hip_options o = defaults();
    std::vector<double> a(static_cast<std::size_t>(n)), d(static_cast<std::size_t>(n));
    double* outs[2] = {a.data(), d.data()};
    PointsArg pa;
    pa.points = pts;
    pa.n = n;
    evaluate(ctx, mol, pa, props, 1, o, outs, nullptr);
    // The denominator of the ALIE is the occupied-orbital density, which is
    // zero exactly where the ALIE kernel writes 0 and that density is 0.
    if (n) {
        kumo_hip_property occ{0, 0, 0, 0};
        kumo_hip_options od = defaults();
        od.backend = KUMO_HIP_AO_MO;
        od.screening_threshold = 0.0;
        evaluate(ctx, mol, pa, &occ, 1, od, &outs[1], nullptr);
        for (long long p = 0; p < n; ++p)
            if (!(d[static_cast<std::size_t>(p)] > 0.0))
                invalid("the average local ionization energy is undefined where the density vanishes");
    }
    return a;
}

kumo_hip_surface* finish(Mesh&& m, std::vector<std::vector<double>>&& fields, double polar) {
    auto* s = new kumo_hip_surface();
    s->mesh = std::move(m);
    for (auto& f : fields) {
        s->stats.push_back(surface_statistics(s->mesh.xyz.data(), s->mesh.vertices(), s->mesh.tri.data(),
                                              s->mesh.triangles(), f.data(), polar));
        s->fields.push_back(std::move(f));
    }
    return s;
}

void put_text(std::FILE* f, const std::string& s) {
    if (std::fwrite(s.data(), 1, s.size(), f) != s.size()) runtime("PLY write failed");
}

} // namespace
} // namespace kumo

// ------------------------------------------------------------------ C-ABI

namespace {
template <typename F> int surface_call(kumo_hip_context* ctx, F&& f) {
    if (!ctx) return KUMO_HIP_INVALID_ARGUMENT;
    try {
        std::lock_guard<std::mutex> lock(ctx->mutex);
        KUMO_HIP(hipSetDevice(ctx->device));
        f();
        return KUMO_HIP_OK;
    } catch (const kumo::Error& e) {
        ctx->last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->last_error = "host memory exhausted";
        return KUMO_HIP_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        ctx->last_error = e.what();
        return KUMO_HIP_RUNTIME_ERROR;
    }
}
} // namespace

extern "C" {

int kumo_hip_density_isosurface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double isovalue, double spacing,
                                kumo_hip_surface** out) {
    if (!out || !mol) return KUMO_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    return surface_call(ctx, [&] { *out = kumo::finish(kumo::density_surface(ctx, mol, isovalue, spacing), {}, 0.0); });
}

int kumo_hip_electrostatic_potential_surface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double isovalue,
                                             double spacing, double tolerance, kumo_hip_surface** out) {
    if (!out || !mol) return KUMO_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    return surface_call(ctx, [&] {
        if (!std::isfinite(tolerance) || tolerance < 0.0) kumo::invalid("potential tolerance must be finite and nonnegative");
        kumo::Mesh m = kumo::density_surface(ctx, mol, isovalue, spacing);
        kumo_hip_options o = kumo::defaults();
        o.potential_tolerance = tolerance;
        auto v = kumo::at_vertices(ctx, mol, m, {{9, 0, 0, 0}}, o, nullptr);
        *out = kumo::finish(std::move(m), std::move(v), 10.0 / kumo::kHartreeInKcal);
    });
}

int kumo_hip_local_ionization_energy_surface(kumo_hip_context* ctx, const kumo_hip_molecule* mol, double isovalue,
                                             double spacing, kumo_hip_surface** out) {
    if (!out || !mol) return KUMO_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    return surface_call(ctx, [&] {
        (void)kumo::alie_at(ctx, mol, nullptr, 0); // the orbital checks before the surface
        kumo::Mesh m = kumo::density_surface(ctx, mol, isovalue, spacing);
        std::vector<std::vector<double>> v{kumo::alie_at(ctx, mol, m.xyz.data(), m.vertices())};
        *out = kumo::finish(std::move(m), std::move(v), 10.0 / kumo::kHartreeInKcal);
    });
}

int kumo_hip_map_on_density_isosurface(kumo_hip_context* ctx, const kumo_hip_molecule* mol,
                                       const kumo_hip_property* props, size_t nprops, double isovalue, double spacing,
                                       const kumo_hip_options* options, kumo_hip_surface** out) {
    if (!out || !mol || (nprops && !props)) return KUMO_HIP_INVALID_ARGUMENT;
    *out = nullptr;
    return surface_call(ctx, [&] {
        if (nprops == 0) kumo::invalid("no property to map on the surface");
        std::vector<kumo_hip_property> list;
        for (size_t i = 0; i < nprops; ++i) {
            const auto& p = props[i];
            if (kumo::channels(p.kind) != 1)
                kumo::invalid("only one-component properties can be mapped on a surface (not a gradient, a "
                              "Hessian or the Cartesian terms of G or K)");
            const bool seen = std::any_of(list.begin(), list.end(), [&](const kumo_hip_property& q) {
                return q.kind == p.kind && q.component == p.component && q.channel == p.channel &&
                       q.orbital_index == p.orbital_index;
            });
            if (!seen) list.push_back(p);
        }
        kumo_hip_options o = options ? *options : kumo::defaults();
        o.device_outputs = 0;
        o.progress = nullptr;
        // The request is checked (components, orbitals) before the surface is built.
        kumo::Mesh none;
        int32_t backend = -1;
        (void)kumo::at_vertices(ctx, mol, none, list, o, &backend);
        kumo::Mesh m = kumo::density_surface(ctx, mol, isovalue, spacing);
        auto v = kumo::at_vertices(ctx, mol, m, list, o, &backend);
        *out = kumo::finish(std::move(m), std::move(v), 0.0);
        (*out)->backend_used = backend;
    });
}

int kumo_hip_average_local_ionization_energy(kumo_hip_context* ctx, const kumo_hip_molecule* mol, const double* points,
                                             size_t n, double* out) {
    if (!mol || (n && (!points || !out))) return KUMO_HIP_INVALID_ARGUMENT;
    return surface_call(ctx, [&] {
        const auto a = kumo::alie_at(ctx, mol, points, static_cast<long long>(n));
        std::copy(a.begin(), a.end(), out);
    });
}

int kumo_hip_surface_sizes(const kumo_hip_surface* s, int64_t* nv, int64_t* nt, int32_t* nf, int32_t* backend) {
    if (!s) return KUMO_HIP_INVALID_ARGUMENT;
    if (nv) *nv = s->mesh.vertices();
    if (nt) *nt = s->mesh.triangles();
    if (nf) *nf = static_cast<int32_t>(s->fields.size());
    if (backend) *backend = s->backend_used;
    return KUMO_HIP_OK;
}

int kumo_hip_surface_mesh(const kumo_hip_surface* s, double* vertices, int64_t* triangles) {
    if (!s) return KUMO_HIP_INVALID_ARGUMENT;
    if (vertices) std::copy(s->mesh.xyz.begin(), s->mesh.xyz.end(), vertices);
    if (triangles) std::copy(s->mesh.tri.begin(), s->mesh.tri.end(), triangles);
    return KUMO_HIP_OK;
}

int kumo_hip_surface_field(const kumo_hip_surface* s, int32_t field, double* values, kumo_hip_surface_stats* stats,
                           int64_t* minima, int64_t* maxima) {
    if (!s || field < 0 || static_cast<size_t>(field) >= s->fields.size()) return KUMO_HIP_INVALID_ARGUMENT;
    const auto& f = s->fields[static_cast<size_t>(field)];
    const auto& st = s->stats[static_cast<size_t>(field)];
    if (values) std::copy(f.begin(), f.end(), values);
    if (stats) *stats = st.s;
    if (minima) std::copy(st.local_minima.begin(), st.local_minima.end(), minima);
    if (maxima) std::copy(st.local_maxima.begin(), st.local_maxima.end(), maxima);
    return KUMO_HIP_OK;
}

int kumo_hip_surface_release(kumo_hip_surface* s) {
    delete s;
    return KUMO_HIP_OK;
}

int kumo_hip_surface_statistics(kumo_hip_context* ctx, const double* vertices, int64_t nv, const int64_t* triangles,
                                int64_t nt, const double* values, double polar, kumo_hip_surface_stats* stats,
                                int64_t* minima, int64_t* maxima, double* area, double* volume) {
    if (!ctx || nv < 0 || nt < 0 || (nv && (!vertices || !values)) || (nt && !triangles)) return KUMO_HIP_INVALID_ARGUMENT;
    return surface_call(ctx, [&] {
        const auto* tri = reinterpret_cast<const long long*>(triangles);
        if (area) *area = kumo::mesh_area(vertices, tri, nt);
        if (volume) *volume = kumo::mesh_volume(vertices, nv, tri, nt);
        if (!stats) return;
        const auto r = kumo::surface_statistics(vertices, nv, tri, nt, values, polar);
        *stats = r.s;
        if (minima) std::copy(r.local_minima.begin(), r.local_minima.end(), minima);
        if (maxima) std::copy(r.local_maxima.begin(), r.local_maxima.end(), maxima);
    });
}

int kumo_hip_write_ply(kumo_hip_context* ctx, const char* path, const double* xyz, int64_t nv, const int64_t* tri,
                       int64_t nt, int32_t nfields, const char* const* names, const double* const* values,
                       int32_t ncomments, const char* const* comments) {
    if (!ctx || !path || nv < 0 || nt < 0 || nfields < 0 || ncomments < 0) return KUMO_HIP_INVALID_ARGUMENT;
    return surface_call(ctx, [&] {
        // write_ply (src/surface.cpp:701-720).
        std::vector<std::string> cols{"x", "y", "z"};
        for (int32_t f = 0; f < nfields; ++f) {
            const std::string name = names[f] ? names[f] : "";
            if (name.empty() || !std::all_of(name.begin(), name.end(), [](unsigned char c) { return std::isalnum(c) || c == '_'; }))
                kumo::invalid("invalid PLY property name '" + name + "'");
            if (std::find(cols.begin(), cols.end(), name) != cols.end())
                kumo::invalid("duplicate PLY property name '" + name + "'");
            cols.push_back(name);
            for (int64_t v = 0; v < nv; ++v)
                if (!std::isfinite(values[f][v])) kumo::invalid("nonfinite value in PLY property '" + name + "'");
        }
        for (int32_t c = 0; c < ncomments; ++c)
            if (std::string(comments[c]).find_first_of("\r\n") != std::string::npos)
                kumo::invalid("PLY comments must not contain line breaks");
        for (int64_t q = 0; q < 3 * nv; ++q)
            if (!std::isfinite(xyz[q])) kumo::invalid("nonfinite PLY vertex");
        for (int64_t q = 0; q < 3 * nt; ++q)
            if (tri[q] < 0 || tri[q] >= nv) kumo::invalid("PLY triangle refers to a missing vertex");
        std::FILE* f = std::fopen(path, "w");
        if (!f) kumo::runtime(std::string("cannot open PLY output ") + path);
        std::string text = "ply\nformat ascii 1.0\n";
        for (int32_t c = 0; c < ncomments; ++c) text += std::string("comment ") + comments[c] + "\n";
        text += "element vertex " + std::to_string(nv) + "\n";
        for (const auto& c : cols) text += "property double " + c + "\n";
        text += "element face " + std::to_string(nt) + "\nproperty list uchar int vertex_indices\nend_header\n";
        char buf[64];
        for (int64_t v = 0; v < nv; ++v) {
            for (int k = 0; k < 3; ++k) {
                std::snprintf(buf, sizeof buf, k ? " %.9e" : "%.9e", xyz[3 * v + k]);
                text += buf;
            }
            for (int32_t fi = 0; fi < nfields; ++fi) {
                std::snprintf(buf, sizeof buf, " %.9e", values[fi][v]);
                text += buf;
            }
            text += '\n';
        }
        for (int64_t t = 0; t < nt; ++t) {
            std::snprintf(buf, sizeof buf, "3 %" PRId64 " %" PRId64 " %" PRId64 "\n", tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]);
            text += buf;
        }
        try {
            kumo::put_text(f, text);
        } catch (...) {
            std::fclose(f);
            throw;
        }
        if (std::fclose(f) != 0) kumo::runtime(std::string("PLY write failed: ") + path);
    });
}

} // extern "C"
