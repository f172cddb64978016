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
    return failures ? 1 : 0;
}
