// read_config.cpp -- `.in` parsing (grammar of the reference's src/utils/read_config.cpp:18-63).
#include <cctype>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "utils.hpp"

namespace Gaukuk {

Config& Config::getInstance() {
    static Config instance;
    return instance;
}

std::string Config::trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) b++;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) e--;
    return s.substr(b, e - b);
}

void Config::loadFromString(const std::string& text) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t eq = line.find('=');
        if (eq != std::string::npos) line[eq] = ' ';
        std::istringstream fields(line);
        std::string key;
        Real value;
        if (fields >> key >> value) {
            data[key] = value;
        } else {
            std::cerr << "Warning: Invalid line in config file: " << line << std::endl;
        }
    }
}

void Config::loadFromFile(const std::string& filename) {
    std::ifstream f(filename);
    if (!f) throw std::invalid_argument("Could not open the input file");
    std::stringstream ss;
    ss << f.rdbuf();
    loadFromString(ss.str());
}

Real Config::get(const std::string& key, Real defaultValue) {
    auto it = data.find(key);
    return it == data.end() ? defaultValue : it->second;
}

}  // namespace Gaukuk
