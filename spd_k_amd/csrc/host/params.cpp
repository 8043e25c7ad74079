#include "params.hpp"
#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace spdk_host {

static std::string strip(const std::string& s) {
    const char* ws = " \t\r\n";
    size_t a = s.find_first_not_of(ws);
    if (a == std::string::npos) return std::string();
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

void Params::load_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Params: cannot open file '" + path + "'");
    std::string raw, block;
    for (int line = 1; std::getline(in, raw); line++) {
        std::string s = strip(raw.substr(0, raw.find('#')));
        if (s.empty()) continue;
        if (s[0] == '<') {
            size_t close = s.find('>');
            if (close == std::string::npos)
                throw std::runtime_error("Params: unterminated block name at line " + std::to_string(line));
            block = strip(s.substr(1, close - 1));
            if (block.empty()) throw std::runtime_error("Params: empty block name at line " + std::to_string(line));
            continue;
        }
        size_t eq = s.find('=');
        if (eq == std::string::npos)
            throw std::runtime_error("Params: expected 'name = value' at line " + std::to_string(line) + ": '" + s + "'");
        if (block.empty()) throw std::runtime_error("Params: parameter outside any <block> at line " + std::to_string(line));
        std::string name = strip(s.substr(0, eq)), value = strip(s.substr(eq + 1));
        if (name.empty() || value.empty())
            throw std::runtime_error("Params: empty name or value at line " + std::to_string(line));
        table_[block][name] = value;
    }
}

void Params::apply_override(const std::string& a) {
    size_t slash = a.find('/'), eq = a.find('=');
    if (slash == std::string::npos || eq == std::string::npos || eq < slash)
        throw std::runtime_error("Params: override must be block/name=value, got '" + a + "'");
    std::string block = strip(a.substr(0, slash)), name = strip(a.substr(slash + 1, eq - slash - 1)),
                value = strip(a.substr(eq + 1));
    if (block.empty() || name.empty() || value.empty()) throw std::runtime_error("Params: bad override '" + a + "'");
    table_[block][name] = value;
}

bool Params::has(const std::string& b, const std::string& n) const {
    auto it = table_.find(b);
    return it != table_.end() && it->second.count(n) > 0;
}

const std::string& Params::lookup(const std::string& b, const std::string& n) const {
    auto it = table_.find(b);
    if (it != table_.end()) {
        auto jt = it->second.find(n);
        if (jt != it->second.end()) return jt->second;
    }
    throw std::runtime_error("Params: missing parameter '" + b + "/" + n + "'");
}

std::string Params::get_string(const std::string& b, const std::string& n) const { return lookup(b, n); }
int Params::get_int(const std::string& b, const std::string& n) const { return std::stoi(lookup(b, n)); }
double Params::get_real(const std::string& b, const std::string& n) const { return std::stod(lookup(b, n)); }
bool Params::get_bool(const std::string& b, const std::string& n) const {
    std::string v = lookup(b, n);
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    if (v == "true" || v == "yes" || v == "on" || v == "1") return true;
    if (v == "false" || v == "no" || v == "off" || v == "0") return false;
    throw std::runtime_error("Params: '" + b + "/" + n + "' is not a boolean: '" + v + "'");
}

void Params::set(const std::string& b, const std::string& n, const std::string& v) { table_[b][n] = v; }

int Params::int_or(const std::string& b, const std::string& n, int def) {
    if (!has(b, n)) set(b, n, std::to_string(def));
    return get_int(b, n);
}
double Params::real_or(const std::string& b, const std::string& n, double def) {
    if (!has(b, n)) {
        std::ostringstream ss;
        ss.precision(17);
        ss << def;
        set(b, n, ss.str());
    }
    return get_real(b, n);
}
bool Params::bool_or(const std::string& b, const std::string& n, bool def) {
    if (!has(b, n)) set(b, n, def ? "true" : "false");
    return get_bool(b, n);
}
std::string Params::string_or(const std::string& b, const std::string& n, const std::string& def) {
    if (!has(b, n)) set(b, n, def);
    return get_string(b, n);
}

void Params::dump(std::ostream& os) const {
    for (const auto& b : table_) {
        os << "<" << b.first << ">\n";
        for (const auto& kv : b.second) os << kv.first << " = " << kv.second << "\n";
        os << "\n";
    }
}

}  // namespace spdk_host
