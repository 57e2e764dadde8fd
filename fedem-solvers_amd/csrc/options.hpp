// Command-line option registry of the drop-in reducer library: the reference
// declares its options through FFaCmdLineArg (reducerInterface.C:52-130,
// 165-204, absent fedem-foundation); the same names and defaults are
// registered here and read by the Fortran host driver through
// fgpu_opt_get_* (C ABI, bind(C) on the Fortran side).
#pragma once
#include <map>
#include <string>
#include <vector>

namespace fgpu {

class Options {
 public:
  static Options& instance();
  void define(const std::string& name, const std::string& def, const std::string& help);
  // "-key=value", "-key value", "-key" (boolean true); "-fco=file"/"-fop=file"
  // name option files holding more such tokens
  int parse(int argc, char** argv, std::string& err);
  bool is_set(const std::string& name) const;
  std::string get(const std::string& name) const;
  int get_int(const std::string& name) const;
  double get_double(const std::string& name) const;
  bool get_bool(const std::string& name) const;
  std::vector<double> get_doubles(const std::string& name) const;
  void clear_values();
  bool defined() const { return !defs_.empty(); }

 private:
  std::map<std::string, std::string> defs_, vals_;
};

}  // namespace fgpu
