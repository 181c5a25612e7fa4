// utils.hpp -- run-time configuration from `.in` files.  Same grammar and defaults as the
// reference's Config singleton (src/utils/utils.hpp:12-22, src/utils/read_config.cpp:18-63):
// `key = value  # comment`, '#' lines and blank lines skipped, values parsed as Real, a missing
// key returns the caller's default (0.0 when none is given).
#pragma once

#include <map>
#include <string>

#include "../gaukuk.hpp"

namespace Gaukuk {

class Config {
public:
    static Config& getInstance();
    void loadFromFile(const std::string& filename);  // throws std::invalid_argument if unreadable
    void loadFromString(const std::string& text);
    Real get(const std::string& key, Real defaultValue = 0.0);
    bool has(const std::string& key) const { return data.count(key) != 0; }
    void set(const std::string& key, Real v) { data[key] = v; }
    void clear() { data.clear(); }

private:
    Config() = default;
    std::map<std::string, Real> data;
    static std::string trim(const std::string& s);
};

namespace GMath {
inline Real SQR(Real a) { return a * a; }
}  // namespace GMath

}  // namespace Gaukuk
