#include "options.hpp"
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace fgpu {

Options& Options::instance() {
  static Options o;
  return o;
}

void Options::define(const std::string& name, const std::string& def, const std::string&) { defs_[name] = def; }

static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

static bool is_number(const std::string& s) {
  if (s.empty()) return false;
  char* end = nullptr;
  std::strtod(s.c_str(), &end);
  return end && *end == 0;
}

int Options::parse(int argc, char** argv, std::string& err) {
  std::vector<std::string> tok;
  for (int i = 0; i < argc; i++)
    if (argv && argv[i]) tok.push_back(argv[i]);
  for (size_t i = 0; i < tok.size(); i++) {
    std::string t = tok[i];
    if (t.size() < 2 || t[0] != '-' || is_number(t)) continue;  // program name or stray value
    std::string key = t.substr(1), val = "true";
    size_t eq = key.find('=');
    if (eq != std::string::npos) {
      val = key.substr(eq + 1);
      key = key.substr(0, eq);
    } else if (i + 1 < tok.size() && (tok[i + 1][0] != '-' || is_number(tok[i + 1]))) {
      val = tok[++i];
      while (i + 1 < tok.size() && is_number(tok[i + 1]) && is_number(val.substr(val.rfind(' ') + 1)))
        val += " " + tok[++i];  // numeric vectors: -gvec 0 9.81 0
    }
    if (key == "fco" || key == "fop") {
      std::ifstream f(val);
      if (!f) { err = "cannot open option file " + val; return -1; }
      std::vector<std::string> ft;
      std::string line;
      while (std::getline(f, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string w;
        while (is >> w) ft.push_back(w);
      }
      std::vector<char*> av;
      for (auto& w : ft) av.push_back(&w[0]);
      if (parse((int)av.size(), av.data(), err)) return -1;
      continue;
    }
    vals_[key] = trim(val);
  }
  return 0;
}

bool Options::is_set(const std::string& n) const { return vals_.count(n) > 0; }

std::string Options::get(const std::string& n) const {
  auto it = vals_.find(n);
  if (it != vals_.end()) return it->second;
  auto d = defs_.find(n);
  return d != defs_.end() ? d->second : std::string();
}

int Options::get_int(const std::string& n) const { return std::atoi(get(n).c_str()); }
double Options::get_double(const std::string& n) const { return std::atof(get(n).c_str()); }
bool Options::get_bool(const std::string& n) const {
  std::string v = get(n);
  return v == "true" || v == "1" || v == "yes" || v == "on" || v == "TRUE" || v == "True";
}
std::vector<double> Options::get_doubles(const std::string& n) const {
  std::vector<double> r;
  std::string v = get(n);
  for (char& c : v)
    if (c == ',') c = ' ';
  std::istringstream is(v);
  double x;
  while (is >> x) r.push_back(x);
  return r;
}
void Options::clear_values() { vals_.clear(); }

}  // namespace fgpu
