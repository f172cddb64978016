"""The reference's surface checks against Multiwfn's quantitative molecular
surface analyses (tests/multiwfn_surface_reference_test.cpp, arguments of
meson.build:746-782), restated over the committed reports so that they run on
the oracle (CPU) and on the HIP path (GPU). A surface is a dict with
"vertices" (n x 3, bohr), "triangles" (m x 3), "values" (n) and "stats" (the
surface_statistics fields, local_minima / local_maxima as vertex lists)."""
import os
import re

import numpy as np

from kumo_hip.api import PropertyKind as K

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "golden")

# Multiwfn's conversion factors (define.f90), as the reference test uses them.
MULTIWFN_KCAL = 627.51
MULTIWFN_EV = 27.2113838
BOHR_IN_ANGSTROM = 0.529177210903

# (molecule, report, mode): esp / alie / leae / lea, "-no-extrema" skips the local extrema.
SURFACE_CASES = [("water", "h2o_esp", "esp"), ("o2", "o2_esp", "esp-no-extrema"),
                 ("acrolein", "acrolein_esp", "esp"), ("water", "h2o_alie", "alie"),
                 ("o2", "o2_alie", "alie-no-extrema"), ("acrolein", "acrolein_alie", "alie"),
                 ("acrolein", "acrolein_leae", "leae"), ("o2", "o2_leae", "leae-no-extrema"),
                 ("acrolein", "acrolein_lea", "lea")]
# Mapped properties: (kind, isovalue, tolerances global, minima, average, variance).
MAPPED = {"leae": (K.LocalElectronAttachmentEnergy, 4e-3, (1.5e-6, 6e-5, 1e-3, 2.5e-2)),
          "lea": (K.LocalElectronAffinity, 1e-3, (4e-5, 6e-3, 1.5e-4, 5e-3))}


def read_report(name):
    with open(os.path.join(GOLDEN, "multiwfn", name + "_surface.txt")) as fh:
        return fh.read().split("\n")


def numbers(text):
    out = []
    for tok in re.sub(r"[()*]", " ", text).split():
        try:
            out.append(float(tok))
        except ValueError:
            pass
    return out


def value_after(lines, label, index):
    for line in lines:
        at = line.find(label)
        if at < 0:
            continue
        values = numbers(line[at + len(label):])
        assert index < len(values), f"reference line '{line}' lacks a value"
        return values[index]
    raise AssertionError(f"reference lacks '{label}'")


def extrema(lines, kind):
    label = f"The number of surface {kind}:"
    for i, line in enumerate(lines):
        if label not in line:
            continue
        count = int(value_after(lines, label, 0))
        out = []
        for row in lines[i + 2:i + 2 + count]:
            v = numbers(row)
            assert len(v) == 7, f"malformed extremum: {row}"
            out.append((v[1], np.array(v[4:7]) / BOHR_IN_ANGSTROM))
        return out
    raise AssertionError(f"reference lacks '{label}'")


def close(actual, expected, rel, ab, label, worst):
    """kumo_test::check_close: |a - e| <= max(abs, rel max(|a|, |e|))."""
    tol = max(ab, rel * max(abs(actual), abs(expected)))
    diff = abs(actual - expected)
    assert np.isfinite(actual) and np.isfinite(expected) and diff <= tol, \
        f"{label}: expected {expected!r}, got {actual!r} (difference {diff:.3g}, tolerance {tol:.3g})"
    worst.append((diff / tol if tol > 0 else 0.0, label))


def compare_extrema(surface, found, expected, tolerance, label, worst):
    assert len(found) == len(expected), f"{label}: {len(expected)} local extrema (got {len(found)})"
    used = [False] * len(found)
    v = surface["vertices"]
    for value, pos in expected:
        best, nearest = None, 0.0
        for f, idx in enumerate(found):
            d = float(np.linalg.norm(v[idx] - pos))
            if not used[f] and (best is None or d < nearest):
                best, nearest = f, d
        used[best] = True
        close(nearest, 0.0, 0.0, 0.4, f"{label} near {pos}: distance", worst)
        close(float(surface["values"][found[best]]), value, 0.0, tolerance, f"{label} near {pos}: value", worst)


def check_surface(surface, mode, lines, label, volume):
    """One report against a surface; volume is the mesh volume. Returns the
    list of (fraction of tolerance, check)."""
    worst = []
    s = surface["stats"]
    kind = mode.replace("-no-extrema", "")
    local = not mode.endswith("-no-extrema")
    close(volume, value_after(lines, "Volume:", 0), 5e-3, 0, label + ": volume", worst)
    close(s["area"], value_after(lines, "Overall surface area:", 0), 5e-3, 0, label + ": area", worst)
    if kind == "esp":
        kcal = lambda key, i: value_after(lines, key, i) / MULTIWFN_KCAL  # noqa: E731
        kcal2 = lambda key: value_after(lines, key, 1) / MULTIWFN_KCAL ** 2  # noqa: E731
        close(s["positive_area"], value_after(lines, "Positive surface area:", 0), 5e-3, 0, label + ": A+", worst)
        close(s["negative_area"], value_after(lines, "Negative surface area:", 0), 5e-3, 0, label + ": A-", worst)
        close(s["nonpolar_area"] / s["area"],
              value_after(lines, "Nonpolar surface area (|ESP| <= 10 kcal/mol):", 1) / 100, 0, 2e-3,
              label + ": nonpolar fraction", worst)
        close(s["minimum"], kcal("Minimal value:", 0), 0, 4e-5, label + ": global minimum", worst)
        close(s["maximum"], kcal("Minimal value:", 1), 0, 4e-5, label + ": global maximum", worst)
        rel, ab = 5e-3, 2e-6
        close(s["average"], value_after(lines, "Overall average value:", 0), rel, ab, label + ": average", worst)
        close(s["positive_average"], value_after(lines, "Positive average value:", 0), rel, ab, label + ": avg+", worst)
        close(s["negative_average"], value_after(lines, "Negative average value:", 0), rel, ab, label + ": avg-", worst)
        close(s["charge_separation"], kcal("Internal charge separation (Pi):", 1), rel, ab, label + ": Pi", worst)
        close(s["polarity_index"], kcal("Molecular polarity index (MPI):", 1), rel, ab, label + ": MPI", worst)
        vr, va = 1e-2, 2e-8
        close(s["total_variance"], kcal2("Overall variance (sigma^2_tot):"), vr, va, label + ": var", worst)
        close(s["positive_variance"], kcal2("Positive variance:"), vr, va, label + ": var+", worst)
        close(s["negative_variance"], kcal2("Negative variance:"), vr, va, label + ": var-", worst)
        close(s["balance"], value_after(lines, "Balance of charges (nu):", 0), 0, 1e-3, label + ": nu", worst)
        if local:
            compare_extrema(surface, s["local_minima"], extrema(lines, "minima"), 4e-5, label + " minimum", worst)
            compare_extrema(surface, s["local_maxima"], extrema(lines, "maxima"), 4e-5, label + " maximum", worst)
    elif kind == "alie":
        close(s["positive_area"], s["area"], 1e-14, 0, label + ": A+ = A", worst)
        close(s["minimum"], value_after(lines, "Global surface minimum:", 0), 0, 4e-5, label + ": minimum", worst)
        close(s["maximum"], value_after(lines, "Global surface maximum:", 0), 0, 4e-4, label + ": maximum", worst)
        close(s["average"], value_after(lines, "Average value:", 0), 5e-5, 0, label + ": average", worst)
        close(s["positive_average"], s["average"], 1e-14, 0, label + ": avg+ = avg", worst)
        close(s["total_variance"], value_after(lines, "Variance:", 0), 2e-2, 0, label + ": variance", worst)
        assert s["negative_area"] == 0.0, label + ": negative area 0"
        if local:
            compare_extrema(surface, s["local_minima"], extrema(lines, "minima"), 8e-4, label + " minimum", worst)
    else:
        _, _, (tg, tm, ta, tv) = MAPPED[kind]
        close(s["negative_area"], value_after(lines, "Negative surface area:", 0), 5e-3, 0, label + ": A-", worst)
        close(s["minimum"], value_after(lines, "Minimal value:", 0) / MULTIWFN_EV, 0, tg, label + ": minimum", worst)
        close(s["maximum"], value_after(lines, "Minimal value:", 1) / MULTIWFN_EV, 0, tg, label + ": maximum", worst)
        close(s["average"], value_after(lines, "Overall average value:", 2) / MULTIWFN_KCAL, ta, 0,
              label + ": average", worst)
        close(s["total_variance"], value_after(lines, "Overall variance:", 2) / MULTIWFN_KCAL ** 2, tv, 0,
              label + ": variance", worst)
        if local:
            compare_extrema(surface, s["local_minima"], extrema(lines, "minima"), tm, label + " minimum", worst)
    return worst
