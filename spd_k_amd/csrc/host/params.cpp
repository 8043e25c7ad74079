#include "params.hpp"

#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <fstream>

namespace spdk_host {

namespace {

bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }

// The characters of [first, last) without leading and trailing white space.
std::string clip(const char* first, const char* last) {
    while (first < last && blank(*first)) ++first;
    while (last > first && blank(last[-1])) --last;
    return std::string(first, last);
}

// One line of an input file, after its comment is removed.
struct Line {
    enum Kind { kEmpty, kHeader, kAssign } kind = kEmpty;
    std::string first, second;  // header: block; assignment: name, value
};

Line lex(const std::string& raw, const std::string& where) {
    const char* begin = raw.c_str();
    const char* end = begin;
    while (*end && *end != '#') ++end;  // '#' starts a comment
    const std::string body = clip(begin, end);
    Line out;
    if (body.empty()) return out;
    const char* s = body.c_str();
    const char* stop = s + body.size();
    if (*s == '<') {
        const char* gt = s + 1;
        while (gt < stop && *gt != '>') ++gt;
        if (gt == stop) throw ConfigError(where + ": block header '" + body + "' has no closing '>'");
        out.kind = Line::kHeader;
        out.first = clip(s + 1, gt);
        if (out.first.empty()) throw ConfigError(where + ": block header '" + body + "' names no block");
        return out;
    }
    const char* eq = s;
    while (eq < stop && *eq != '=') ++eq;
    if (eq == stop) throw ConfigError(where + ": '" + body + "' is neither a <block> header nor 'name = value'");
    out.kind = Line::kAssign;
    out.first = clip(s, eq);
    out.second = clip(eq + 1, stop);
    if (out.first.empty() || out.second.empty())
        throw ConfigError(where + ": '" + body + "' needs both a name and a value");
    return out;
}

// Shortest decimal text that reads back as exactly `v`.
std::string exact_text(double v) {
    char buf[40];
    for (int digits = 15; digits <= 17; digits++) {
        std::snprintf(buf, sizeof(buf), "%.*g", digits, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    return buf;
}

std::string lowercase(std::string s) {
    for (char& c : s)
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
    return s;
}

}  // namespace

void Params::set(const std::string& key, const std::string& text, const std::string& origin) {
    auto it = index_.find(key);
    if (it != index_.end()) {
        entries_[it->second].text = text;
        entries_[it->second].origin = origin;
        return;
    }
    const size_t slash = key.find('/');
    index_.emplace(key, entries_.size());
    entries_.push_back({key.substr(0, slash), slash == std::string::npos ? std::string() : key.substr(slash + 1),
                        text, origin});
}

void Params::load_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw ConfigError("input file '" + path + "' cannot be read");
    std::string raw, block;
    int number = 0;
    while (std::getline(in, raw)) {
        const std::string where = path + ":" + std::to_string(++number);
        const Line line = lex(raw, where);
        if (line.kind == Line::kHeader) {
            block = line.first;
        } else if (line.kind == Line::kAssign) {
            if (block.empty()) throw ConfigError(where + ": '" + line.first + "' is assigned before any <block> header");
            set(block + "/" + line.first, line.second, where);
        }
    }
}

void Params::apply_override(const std::string& a) {
    const size_t slash = a.find('/'), eq = a.find('=');
    std::string block, name, value;
    if (slash != std::string::npos && eq != std::string::npos && slash < eq) {
        block = clip(a.c_str(), a.c_str() + slash);
        name = clip(a.c_str() + slash + 1, a.c_str() + eq);
        value = clip(a.c_str() + eq + 1, a.c_str() + a.size());
    }
    if (block.empty() || name.empty() || value.empty())
        throw ConfigError("command-line assignment '" + a + "' is not of the form block/name=value");
    set(block + "/" + name, value, "command line");
}

const Params::Entry& Params::entry(const std::string& key) const {
    auto it = index_.find(key);
    if (it == index_.end()) throw ConfigError("parameter " + key + " is not set");
    return entries_[it->second];
}

template <> std::string Params::get<std::string>(const std::string& key) const { return entry(key).text; }

template <> int Params::get<int>(const std::string& key) const {
    const Entry& e = entry(key);
    char* stop = nullptr;
    errno = 0;
    const long v = std::strtol(e.text.c_str(), &stop, 10);
    if (stop == e.text.c_str() || errno == ERANGE || v < INT_MIN || v > INT_MAX)
        throw ConfigError("parameter " + key + " = '" + e.text + "' (" + e.origin + ") is not an integer");
    return (int)v;
}

template <> double Params::get<double>(const std::string& key) const {
    const Entry& e = entry(key);
    char* stop = nullptr;
    const double v = std::strtod(e.text.c_str(), &stop);
    if (stop == e.text.c_str())
        throw ConfigError("parameter " + key + " = '" + e.text + "' (" + e.origin + ") is not a number");
    return v;
}

template <> bool Params::get<bool>(const std::string& key) const {
    const Entry& e = entry(key);
    static const char* const yes[] = {"true", "yes", "on", "1"};
    static const char* const no[] = {"false", "no", "off", "0"};
    const std::string t = lowercase(e.text);
    for (int i = 0; i < 4; i++) {
        if (t == yes[i]) return true;
        if (t == no[i]) return false;
    }
    throw ConfigError("parameter " + key + " = '" + e.text + "' (" + e.origin +
                      ") is not a boolean (true/false, yes/no, on/off, 1/0)");
}

template <> int Params::get_or<int>(const std::string& key, const int& fallback) {
    if (!has(key)) set(key, std::to_string(fallback), "default");
    return get<int>(key);
}
template <> double Params::get_or<double>(const std::string& key, const double& fallback) {
    if (!has(key)) set(key, exact_text(fallback), "default");
    return get<double>(key);
}
template <> bool Params::get_or<bool>(const std::string& key, const bool& fallback) {
    if (!has(key)) set(key, fallback ? "true" : "false", "default");
    return get<bool>(key);
}
template <> std::string Params::get_or<std::string>(const std::string& key, const std::string& fallback) {
    if (!has(key)) set(key, fallback, "default");
    return get<std::string>(key);
}

std::string Params::get_or(const std::string& key, const char* fallback) {
    return get_or<std::string>(key, std::string(fallback));
}

void Params::dump(std::ostream& os) const {
    std::vector<std::string> blocks;
    for (const Entry& e : entries_) {
        bool seen = false;
        for (const std::string& b : blocks) seen = seen || b == e.block;
        if (!seen) blocks.push_back(e.block);
    }
    os << "# effective parameters of this run; each line names where its value came from\n";
    for (const std::string& b : blocks) {
        os << "\n<" << b << ">\n";
        for (const Entry& e : entries_)
            if (e.block == b) os << e.name << " = " << e.text << "    # " << e.origin << "\n";
    }
}

}  // namespace spdk_host
