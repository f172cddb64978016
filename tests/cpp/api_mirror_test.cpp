// The C++ mirror of the reference Evaluator API, used exactly as a reference
// user writes it (include/chem/kumo/evaluator.hpp:474-486 usage block).
// Prints one line per check; exits 1 on a failed check. Needs a GPU.
#include <chem/kumo/hip.hpp>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

using namespace chem::kumo;

int failures = 0;
void check(bool ok, const char* what) {
    std::printf("%s: %s\n", ok ? "ok" : "FAILED", what);
    if (!ok) ++failures;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const Molecule molecule = Molecule::from_fchk(argv[1]);
    check(molecule.basis_size() == 43, "water def2-TZVP has 43 AOs");
    const Grid grid = Grid::around(molecule.atoms(), 3.0, 0.3);
    Evaluator evaluator(molecule);
    Request request;
    request.add(density::total.value()).add(density::total.gradient()).add(density::total.elf());
    const Results results = evaluator.evaluate(grid, request);
    const auto rho = results.get(density::total.value()).component();
    double sum = 0.0;
    for (double v : rho)
        sum += v;
    const double electrons = sum * 0.3 * 0.3 * 0.3;
    std::printf("grid %d %d %d origin %.17g %.17g %.17g\n", grid.counts[0], grid.counts[1], grid.counts[2],
                grid.origin[0], grid.origin[1], grid.origin[2]);
    std::printf("density_sum=%.17g electrons=%.6f\n", sum, electrons);
    check(rho.size() == grid.size(), "one density value per grid point");
    check(results.backend_used() == Backend::AO_MO, "AUTO picks AO-MO for an eligible molecule");
    check(results.get(density::total.gradient()).components.size() == 3, "gradient has three components");
    bool threw = false;
    try {
        (void)evaluator.evaluate(grid, Request().add(density::difference.elf()));
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw, "ELF of the spin difference throws std::invalid_argument");
    threw = false;
    try {
        (void)evaluator.evaluate(grid, Request().add(orbital::alpha(100000).value()));
    } catch (const std::out_of_range&) {
        threw = true;
    }
    check(threw, "an MO index past the last MO throws std::out_of_range");
    std::vector<Vector3> points{{0.0, 0.0, 0.0}, {0.5, 0.2, -0.1}};
    const auto at = evaluator.evaluate(points, Request().add(density::total.value()));
    check(at.get(density::total.value()).component().size() == 2, "evaluation at explicit points");

    // The ESP on the grid: exact sums and the far field at tolerance 1e-6.
    EvaluationOptions far;
    far.potential_tolerance = 1e-6;
    far.far_field = true;
    const auto exact_esp = evaluator.evaluate(grid, Request().add(density::total.electrostatic_potential()));
    const auto far_esp = Evaluator(molecule, far).evaluate(grid, Request().add(density::total.electrostatic_potential()));
    const auto ve = exact_esp.get(density::total.electrostatic_potential()).component();
    const auto vf = far_esp.get(density::total.electrostatic_potential()).component();
    double worst = 0.0;
    for (std::size_t q = 0; q < ve.size(); ++q)
        worst = std::max(worst, std::abs(ve[q] - vf[q]));
    std::printf("far-field ESP: largest difference from the exact sums %.3g\n", worst);
    check(vf.size() == ve.size() && worst <= 1e-6, "far-field ESP within its tolerance of the exact sums");

    // Cube output as tests/atom_charge_cube_test.cpp writes it: an iron atom
    // (Z 26) with an ECP charge of 16, two points per axis.
    Grid small;
    small.counts = {2, 2, 2};
    const Evaluator::Field iron{small, {{26, 16.0, {0.0, 0.0, 0.0}}}, std::vector<double>(8, 0.5)};
    const std::string dir = argc > 2 ? argv[2] : "/tmp";
    iron.write_cube(dir + "/kumo-charge-density-test.cube");
    iron.write_orbital_cube(dir + "/kumo-charge-orbital-test.cube", 1);
    for (const char* name : {"/kumo-charge-density-test.cube", "/kumo-charge-orbital-test.cube"}) {
        std::FILE* f = std::fopen((dir + name).c_str(), "r");
        char line[256];
        int natoms = 0, z = 0, count = 0;
        double charge = 0.0, x = 0.0;
        bool ok = f != nullptr;
        for (int i = 0; ok && i < 2; ++i) ok = std::fgets(line, sizeof line, f) != nullptr;
        ok = ok && std::fscanf(f, "%d %lf %lf %lf", &natoms, &x, &x, &x) == 4;
        for (int a = 0; ok && a < 3; ++a) ok = std::fscanf(f, "%d %lf %lf %lf", &count, &x, &x, &x) == 4 && count == 2;
        ok = ok && std::fscanf(f, "%d %lf %lf %lf %lf", &z, &charge, &x, &x, &x) == 5;
        if (f) std::fclose(f);
        const bool orbital = std::string(name).find("orbital") != std::string::npos;
        check(ok && natoms == (orbital ? -1 : 1) && z == 26 && std::abs(charge - 16.0) < 1e-6,
              orbital ? "orbital Cube keeps Z 26 and charge 16, atom count -1"
                      : "density Cube keeps Z 26 and charge 16, atom count 1");
    }
    threw = false;
    try {
        iron.write_orbital_cube(dir + "/kumo-charge-orbital-test.cube", 0);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw, "orbital index 0 throws std::invalid_argument");
    Evaluator::Field broken = iron;
    broken.values[3] = std::nan("");
    threw = false;
    try {
        broken.write_cube(dir + "/kumo-charge-broken-test.cube");
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw, "a nonfinite sample throws std::invalid_argument");

    // Molecular surfaces (include/chem/kumo/surface.hpp) as a reference user
    // calls them.
    const PotentialSurface esp = electrostatic_potential_surface(molecule);
    std::printf("esp_surface vertices=%zu triangles=%zu area=%.17g minimum=%.17g\n", esp.mesh.vertices.size(),
                esp.mesh.triangles.size(), esp.statistics.area, esp.statistics.minimum);
    check(esp.potential.size() == esp.mesh.vertices.size(), "one potential value per surface vertex");
    check(std::abs(esp.statistics.area - esp.mesh.area()) <= 1e-12 * esp.mesh.area(),
          "statistics area equals TriangleMesh::area");
    check(esp.mesh.volume() > 0.0, "the isosurface encloses a positive volume");
    check(!esp.statistics.local_minima.empty() && esp.statistics.local_minima.size() <= esp.mesh.vertices.size(),
          "ESP surface has local minima");
    const TriangleMesh mesh = density_isosurface(molecule);
    check(mesh.vertices.size() == esp.mesh.vertices.size() && mesh.triangles == esp.mesh.triangles,
          "density_isosurface is the ESP surface's mesh");
    const IonizationEnergySurface alie = local_ionization_energy_surface(molecule);
    check(alie.statistics.negative_area == 0.0 && alie.statistics.minimum > 0.0, "ALIE is positive on the surface");
    const auto at_points = average_local_ionization_energy(molecule, points);
    check(at_points.size() == 2 && at_points[0] > 0.0, "average_local_ionization_energy at points");
    const MappedSurface mapped =
        map_on_density_isosurface(molecule, {density::total.value(), density::total.elf(), density::total.value()});
    bool on_iso = mapped.values.size() == 2 && mapped.properties.size() == 2;
    for (double v : on_iso ? mapped.values[0] : std::vector<double>{})
        on_iso = on_iso && std::abs(v - 1e-3) <= 1e-9;
    check(on_iso, "mapped density equals the isovalue; a repeated property is evaluated once");
    // The reference test's octahedron with V = x: exact statistics.
    TriangleMesh octahedron;
    octahedron.vertices = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    octahedron.triangles = {{0, 2, 4}, {0, 5, 2}, {0, 4, 3}, {0, 3, 5}, {1, 4, 2}, {1, 2, 5}, {1, 3, 4}, {1, 5, 3}};
    const std::vector<double> x{1, -1, 0, 0, 0, 0};
    const SurfaceStatistics oct = surface_statistics(octahedron, x, 0.5);
    check(std::abs(oct.area - 4.0 * std::sqrt(3.0)) < 1e-13 && std::abs(octahedron.volume() - 4.0 / 3.0) < 1e-14 &&
              std::abs(oct.positive_variance - 1.0 / 18.0) < 1e-15 && oct.local_minima == std::vector<std::size_t>{1} &&
              oct.local_maxima == std::vector<std::size_t>{0},
          "octahedron statistics are exact");
    write_ply(dir + "/kumo-surface-test.ply", esp.mesh, {{"esp", esp.potential}}, {"water ESP"});
    std::FILE* ply = std::fopen((dir + "/kumo-surface-test.ply").c_str(), "r");
    char head[64] = {0};
    const bool ply_ok = ply && std::fgets(head, sizeof head, ply) && std::string(head) == "ply\n";
    if (ply) std::fclose(ply);
    check(ply_ok, "write_ply writes an ASCII PLY file");
    threw = false;
    try {
        (void)density_isosurface(molecule, 1e-9);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    check(threw, "an isovalue below 1e-8 throws std::invalid_argument");
    return failures ? 1 : 0;
}
