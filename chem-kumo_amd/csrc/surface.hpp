// Molecular surfaces (include/chem/kumo/surface.hpp, src/surface.cpp): the
// density isosurface by marching tetrahedra with vertices placed on the
// exactly computed density, and the Politzer-Murray statistics of vertex
// values on it. The density, its gradient and the mapped properties are
// evaluated on the GPU through the point path; the polygonisation and the
// statistics are host code over the mesh.
#pragma once

#include "../../include/kumo_hip.h"

#include <array>
#include <functional>
#include <vector>

namespace kumo {

struct Mesh {
    std::vector<double> xyz;          // 3 per vertex, bohr
    std::vector<long long> tri;       // 3 per triangle, counter-clockwise seen from outside
    std::vector<std::array<int, 3>> cell; // per vertex: grid indices of its edge's lower end
    long long vertices() const { return static_cast<long long>(xyz.size() / 3); }
    long long triangles() const { return static_cast<long long>(tri.size() / 3); }
};

struct SurfaceStats {
    kumo_hip_surface_stats s{};
    std::vector<long long> local_minima, local_maxima;
};

// Values and gradients (3 per point) of a scalar function at n points.
using SurfaceFunction = std::function<void(const std::vector<double>& points, std::vector<double>& values,
                                           std::vector<double>& gradients)>;

// extract_isosurface (src/surface.cpp:309-331): grid samples in Field order.
Mesh extract_isosurface(const kumo_hip_grid& grid, const std::vector<double>& samples, double isovalue,
                        const SurfaceFunction& function);
bool isosurface_reaches_boundary(const kumo_hip_grid& grid, const std::vector<double>& samples, double isovalue);

// surface_statistics (src/surface.cpp:418-529).
SurfaceStats surface_statistics(const double* xyz, long long nvertices, const long long* tri, long long ntriangles,
                                const double* values, double polar_threshold);
double mesh_area(const double* xyz, const long long* tri, long long ntriangles);
double mesh_volume(const double* xyz, long long nvertices, const long long* tri, long long ntriangles);

} // namespace kumo
