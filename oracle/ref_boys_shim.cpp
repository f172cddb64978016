// ORACLE — test infrastructure only. C entry point over the reference's own
// Boys function (src/boys_function.cpp, compiled from the reference sources
// where they lie by oracle/Makefile), used to pin orc_boys bit for bit.
#include "boys_function.hpp"

#include <span>

extern "C" int ref_boys(double T, int m, double* out) {
    try {
        chem::kumo::detail::boys_function(T, std::span<double>(out, static_cast<std::size_t>(m + 1)));
        return 0;
    } catch (...) {
        return 1;
    }
}
