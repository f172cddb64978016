// FCHK reader (see fchk.hpp). Each record header holds the label in its first
// 43 columns, then the type letter and, for arrays, "N=" and the count
// (src/fchk.cpp:120-187). Only integer and real records are kept.
#include "fchk.hpp"

#include "common.hpp"

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

namespace kumo {
namespace {

std::string read_file(const std::string& path) {
    gzFile f = gzopen(path.c_str(), "rb"); // reads plain files transparently
    if (!f) runtime("cannot open formatted checkpoint: " + path);
    std::string out;
    char buf[1 << 16];
    int n;
    while ((n = gzread(f, buf, sizeof(buf))) > 0)
        out.append(buf, static_cast<std::size_t>(n));
    const bool failed = n < 0;
    gzclose(f);
    if (failed) runtime("cannot decompress formatted checkpoint: " + path);
    return out;
}

std::string rtrim(std::string s) {
    while (!s.empty() && (s.back() == ' ' || s.back() == '\t' || s.back() == '\r'))
        s.pop_back();
    return s;
}

double parse_real(std::string tok, const std::string& label) {
    for (auto& c : tok)
        if (c == 'D' || c == 'd') c = 'E';
    char* end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0' || !std::isfinite(v))
        runtime("invalid real token in fchk record " + label + ": " + tok);
    return v;
}

long long parse_int(const std::string& tok, const std::string& label) {
    char* end = nullptr;
    const long long v = std::strtoll(tok.c_str(), &end, 10);
    if (end == tok.c_str() || *end != '\0') runtime("invalid integer token in fchk record " + label + ": " + tok);
    return v;
}

} // namespace

Fchk::Fchk(const std::string& path) {
    const std::string text = read_file(path);
    std::istringstream in(text);
    std::string line;
    if (!std::getline(in, line) || !std::getline(in, line))
        runtime("formatted checkpoint is missing its two-line header: " + path);
    while (std::getline(in, line)) {
        if (line.size() < 44) continue;
        const std::string label = rtrim(line.substr(0, 43));
        if (label.empty()) continue;
        std::istringstream meta(line.substr(43));
        char type = 0;
        meta >> type;
        if (type != 'I' && type != 'R') continue;
        std::string marker;
        if (!(meta >> marker)) runtime("malformed fchk record metadata: " + label);
        FchkRecord rec;
        rec.real = type == 'R';
        if (marker == "N=") {
            long long count = -1;
            if (!(meta >> count) || count < 0 || static_cast<unsigned long long>(count) > text.size())
                runtime("invalid array size in fchk: " + label);
            rec.array = true;
            std::string tok;
            for (long long i = 0; i < count; ++i) {
                if (!(in >> tok)) runtime("truncated array in fchk: " + label);
                if (rec.real) rec.reals.push_back(parse_real(tok, label));
                else rec.integers.push_back(parse_int(tok, label));
            }
            std::getline(in, line); // rest of the last data line
        } else if (rec.real) {
            rec.value = parse_real(marker, label);
        } else {
            rec.integer = parse_int(marker, label);
        }
        records_.emplace(label, std::move(rec));
    }
}

const FchkRecord& Fchk::get(const std::string& label) const {
    const auto it = records_.find(label);
    if (it == records_.end()) runtime("missing fchk record: " + label);
    return it->second;
}

long long Fchk::integer(const std::string& label) const {
    const auto& r = get(label);
    if (r.array || r.real) runtime("fchk record is not an integer scalar: " + label);
    return r.integer;
}

const std::vector<long long>& Fchk::integers(const std::string& label) const {
    const auto& r = get(label);
    if (!r.array || r.real) runtime("fchk record is not an integer array: " + label);
    return r.integers;
}

const std::vector<double>& Fchk::reals(const std::string& label) const {
    const auto& r = get(label);
    if (!r.array || !r.real) runtime("fchk record is not a real array: " + label);
    return r.reals;
}

} // namespace kumo
