// Gaussian formatted checkpoint (FCHK) reader, plain or gzip
// (src/fchk.cpp:78-187, src/gzip_input.cpp).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace kumo {

struct FchkRecord {
    bool real = false, array = false;
    long long integer = 0;
    double value = 0.0;
    std::vector<long long> integers;
    std::vector<double> reals;
};

class Fchk {
public:
    explicit Fchk(const std::string& path);
    bool contains(const std::string& label) const { return records_.count(label) != 0; }
    long long integer(const std::string& label) const;
    const std::vector<long long>& integers(const std::string& label) const;
    const std::vector<double>& reals(const std::string& label) const;

private:
    const FchkRecord& get(const std::string& label) const;
    std::map<std::string, FchkRecord> records_;
};

} // namespace kumo
