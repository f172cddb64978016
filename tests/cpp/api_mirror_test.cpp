// The C++ mirror of the reference Evaluator API, used exactly as a reference
// user writes it (include/chem/kumo/evaluator.hpp:474-486 usage block).
// Prints one line per check; exits 1 on a failed check. Needs a GPU.
#include <chem/kumo/hip.hpp>

#include <cmath>
#include <cstdio>
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
    return failures ? 1 : 0;
}
