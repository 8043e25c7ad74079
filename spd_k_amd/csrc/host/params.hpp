// Run configuration: an input file in the Athena++ format followed by
// command-line assignments (the interface of src/parameter_input.hpp:1-49).
//
//   <block>                 opens a block
//   name = value   # note   assigns a key inside the open block
//   block/name=value        command-line form, applied after the file
//
// Keys are addressed as "block/name". The last assignment to a key wins.
// Values are kept as text and converted when read; every failure (file,
// syntax, missing key, text that is not of the requested type) throws
// ConfigError, which the driver maps to the configuration-error exit.
#pragma once
#include <ostream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace spdk_host {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Params {
  public:
    void load_file(const std::string& path);
    void apply_override(const std::string& assignment);

    bool has(const std::string& key) const { return index_.count(key) != 0; }
    // Typed read of a key that must be present (T = int, double, bool,
    // std::string).
    template <class T> T get(const std::string& key) const;
    // Typed read; an absent key is first recorded with the default, so the
    // provenance dump lists every value the run used.
    template <class T> T get_or(const std::string& key, const T& fallback);
    std::string get_or(const std::string& key, const char* fallback);
    void set(const std::string& key, const std::string& text, const std::string& origin);
    // Writes the effective parameters as a loadable input file, blocks and
    // keys in the order they were first assigned, each with its origin.
    void dump(std::ostream& os) const;

  private:
    struct Entry {
        std::string block, name, text, origin;
    };
    std::vector<Entry> entries_;                      // first-assignment order
    std::unordered_map<std::string, size_t> index_;  // "block/name" -> entries_
    const Entry& entry(const std::string& key) const;
};

}  // namespace spdk_host
