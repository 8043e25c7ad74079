// Athena++-style runtime parameters ("<block>" headers, "name = value  #
// comment" lines, "block/name=value" overrides; last assignment wins) with
// the access semantics of the reference's ParameterInput
// (src/parameter_input.hpp:1-49, src/parameter_input.cpp:24-152).
#pragma once
#include <map>
#include <ostream>
#include <string>

namespace spdk_host {

class Params {
  public:
    void load_file(const std::string& path);             // throws std::runtime_error
    void apply_override(const std::string& assignment);  // "block/name=value"
    bool has(const std::string& block, const std::string& name) const;
    std::string get_string(const std::string& block, const std::string& name) const;
    int get_int(const std::string& block, const std::string& name) const;
    double get_real(const std::string& block, const std::string& name) const;
    bool get_bool(const std::string& block, const std::string& name) const;
    // return the stored value, storing the default when absent
    int int_or(const std::string& block, const std::string& name, int def);
    double real_or(const std::string& block, const std::string& name, double def);
    bool bool_or(const std::string& block, const std::string& name, bool def);
    std::string string_or(const std::string& block, const std::string& name, const std::string& def);
    void set(const std::string& block, const std::string& name, const std::string& value);
    void dump(std::ostream& os) const;  // provenance echo (parameters.txt)

  private:
    std::map<std::string, std::map<std::string, std::string>> table_;
    const std::string& lookup(const std::string& block, const std::string& name) const;
};

}  // namespace spdk_host
