// Nastran bulk-data subset reader for the part input (-linkfile).  The
// reference parses parts with FFl (absent fedem-foundation); this reader
// covers the cards of the element types on the device path: GRID, CHEXA,
// CQUAD4, CBAR, MAT1, PSHELL, PSOLID, PBAR, PBARL (TUBE), ASET/ASET1, SPC1 (both with
// "G1 THRU G2" ranges), RBE2, RBE3 and the
// FORCE/MOMENT nodal loads (basic coordinate system only), in
// free (comma), small (8 column) and large (16 column, '*') field formats.
// Nodes are numbered in ascending GRID id order, as SAM numbers them.
#include "nastran.hpp"
#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

namespace fgpu {
namespace {

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// Nastran real: 1.5E-3, 1.5D-3 or the implied-exponent form 1.5-3
bool to_double(std::string s, double& v) {
  s = trim(s);
  if (s.empty()) return false;
  for (char& c : s) c = (char)std::toupper(c);
  std::replace(s.begin(), s.end(), 'D', 'E');
  if (s.find('E') == std::string::npos) {
    size_t p = s.find_last_of("+-");
    if (p != std::string::npos && p > 0) s.insert(p, "E");
  }
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == 0;
}

struct Card {
  std::string name;
  std::vector<std::string> f;
};

// free-field fields of a line, empty ones included (also after a trailing comma)
std::vector<std::string> split_free(const std::string& ln) {
  std::vector<std::string> r;
  size_t a = 0;
  while (true) {
    size_t b = ln.find(',', a);
    r.push_back(trim(ln.substr(a, b == std::string::npos ? std::string::npos : b - a)));
    if (b == std::string::npos) break;
    a = b + 1;
  }
  return r;
}

std::vector<Card> read_cards(std::istream& in) {
  std::vector<Card> cards;
  std::string ln;
  bool large = false;
  while (std::getline(in, ln)) {
    if (!ln.empty() && ln.back() == '\r') ln.pop_back();
    if (trim(ln).empty() || ln[0] == '$') continue;
    auto split_fixed = [](const std::string& l, int w) {
      std::vector<std::string> r;
      std::string body = l.size() > 8 ? l.substr(8, std::min<size_t>(l.size(), 72) - 8) : "";
      for (size_t i = 0; i < body.size(); i += w) r.push_back(trim(body.substr(i, w)));
      return r;
    };
    if (ln[0] == '*' || ln[0] == '+') {
      if (cards.empty()) continue;
      std::vector<std::string> add;
      if (ln.find(',') != std::string::npos) {
        std::vector<std::string> all = split_free(ln);
        add.assign(all.begin() + 1, all.end());
      } else {
        add = split_fixed(ln, large ? 16 : 8);
      }
      cards.back().f.insert(cards.back().f.end(), add.begin(), add.end());
      continue;
    }
    Card c;
    if (ln.find(',') != std::string::npos) {
      std::vector<std::string> all = split_free(ln);
      c.name = all[0];
      c.f.assign(all.begin() + 1, all.end());
      large = !c.name.empty() && c.name.back() == '*';
    } else {
      c.name = trim(ln.substr(0, std::min<size_t>(8, ln.size())));
      large = !c.name.empty() && c.name.back() == '*';
      c.f = split_fixed(ln, large ? 16 : 8);
    }
    if (!c.name.empty() && c.name.back() == '*') c.name.pop_back();
    for (char& ch : c.name) ch = (char)std::toupper(ch);
    cards.push_back(c);
  }
  return cards;
}

bool to_int(const std::string& s, int& v) {
  if (s.empty()) return false;
  char* end = nullptr;
  long x = std::strtol(s.c_str(), &end, 10);
  if (!end || *end) return false;
  v = (int)x;
  return true;
}

// Grid ids of the fields f[k0..] of an SPC1 / ASET1 card: blank fields are
// skipped, "G1 THRU G2" is the range G1..G2 (its generated ids need not
// exist: those go to `ranged`, the listed ones to `listed`).  False with a
// message for any other field.
bool grid_list(const std::vector<std::string>& f, size_t k0, const std::string& card, std::vector<int>& listed,
               std::vector<int>& ranged, std::string& err) {
  for (size_t k = k0; k < f.size(); k++) {
    std::string t = trim(f[k]);
    if (t.empty()) continue;
    std::string u = t;
    for (char& ch : u) ch = (char)std::toupper(ch);
    int g = 0;
    if (u == "THRU") {
      int a = listed.empty() ? 0 : listed.back(), b = 0;
      const bool ok = k > k0 && !listed.empty() && k + 1 < f.size() && to_int(trim(f[k + 1]), b) && b >= a;
      if (!ok) {
        err = card + ": THRU needs grid ids before and after it, the first not larger";
        return false;
      }
      for (int q = a + 1; q < b; q++) ranged.push_back(q);
      listed.push_back(b);
      k++;
    } else if (to_int(t, g) && g > 0) {
      listed.push_back(g);
    } else {
      err = card + ": field '" + t + "' is not a grid id";
      return false;
    }
  }
  return true;
}

}  // namespace

int read_nastran(const std::string& path, NastranPart& out, std::string& err) {
  std::ifstream in(path);
  if (!in) { err = "cannot open link file " + path; return -1; }
  std::vector<Card> cards = read_cards(in);
  struct Grid { double x[3]; };
  struct Elem { int type, eid, pid; std::vector<int> nodes; double orient[3]; };
  struct Mat { double E = 0, G = 0, nu = 0, rho = 0; bool hasG = false; };
  struct Prop { int kind = 0, mid = 0; double t = 0, A = 0, I1 = 0, I2 = 0, J = 0, nsm = 0, K1 = 0, K2 = 0; };
  std::map<int, Grid> grids;
  std::vector<Elem> elems;
  std::map<int, Mat> mats;
  std::map<int, Prop> props;
  std::vector<int> aset, aset_range;  // listed ASET grids; grids inside ASET1 THRU ranges
  std::map<int, int> spc;  // grid -> fixed-DOF bitmask
  std::vector<std::pair<int, std::pair<std::string, std::vector<int>>>> rbe2;
  struct R3Group {
    double wt;
    int comps;  // bitmask over the six components
    std::vector<int> grids;
  };
  struct R3 {
    int eid = 0, ref = 0, refc = 0;
    std::vector<R3Group> groups;
  };
  std::vector<R3> rbe3;
  struct Load { int sid, grid, dof0; double v[3]; };
  std::vector<Load> loads;
  auto I = [](const std::vector<std::string>& f, size_t k) -> int {
    return k < f.size() && !f[k].empty() ? std::atoi(f[k].c_str()) : 0;
  };
  auto D = [](const std::vector<std::string>& f, size_t k, double def = 0.0) -> double {
    double v;
    return (k < f.size() && to_double(f[k], v)) ? v : def;
  };
  for (const Card& c : cards) {
    const auto& f = c.f;
    if (c.name == "GRID") {
      Grid g{{D(f, 2), D(f, 3), D(f, 4)}};
      grids[I(f, 0)] = g;
    } else if (c.name == "CHEXA" || c.name == "CQUAD4" || c.name == "CBAR") {
      Elem e;
      e.type = c.name == "CHEXA" ? HEX8 : c.name == "CQUAD4" ? ANDES4 : BEAM2;
      e.eid = I(f, 0);
      e.pid = I(f, 1);
      int nn = elem_nodes(e.type);
      for (int k = 0; k < nn; k++) e.nodes.push_back(I(f, 2 + k));
      e.orient[0] = e.orient[1] = e.orient[2] = 0.0;
      if (e.type == BEAM2)
        for (int k = 0; k < 3; k++) e.orient[k] = D(f, 4 + k);
      elems.push_back(e);
    } else if (c.name == "MAT1") {
      Mat m;
      m.E = D(f, 1);
      double g;
      m.hasG = f.size() > 2 && to_double(f[2], g);
      m.G = m.hasG ? g : 0.0;
      m.nu = D(f, 3);
      m.rho = D(f, 4);
      mats[I(f, 0)] = m;
    } else if (c.name == "PSHELL") {
      Prop p;
      p.kind = 1; p.mid = I(f, 1); p.t = D(f, 2);
      props[I(f, 0)] = p;
    } else if (c.name == "PSOLID") {
      Prop p;
      p.kind = 2; p.mid = I(f, 1);
      props[I(f, 0)] = p;
    } else if (c.name == "PBAR") {
      Prop p;
      p.kind = 3; p.mid = I(f, 1); p.A = D(f, 2); p.I1 = D(f, 3); p.I2 = D(f, 4); p.J = D(f, 5); p.nsm = D(f, 6);
      p.K1 = D(f, 16); p.K2 = D(f, 17);
      props[I(f, 0)] = p;
    } else if (c.name == "PBARL") {
      // TUBE only: DIM1 outer radius, DIM2 inner radius; thin-tube shear area A/2
      std::string type = f.size() > 3 ? trim(f[3]) : std::string();
      for (char& ch : type) ch = (char)std::toupper(ch);
      if (type != "TUBE") { err = "PBARL section type " + type + " is not supported"; return -2; }
      const double R = D(f, 8), r = D(f, 9), pi = 3.14159265358979323846;
      Prop p;
      p.kind = 3; p.mid = I(f, 1);
      p.A = pi * (R * R - r * r);
      p.I1 = p.I2 = pi * (R * R * R * R - r * r * r * r) / 4.0;
      p.J = 2.0 * p.I1;
      p.K1 = p.K2 = 0.5;
      props[I(f, 0)] = p;
    } else if (c.name == "ASET") {
      for (size_t k = 0; k + 1 < f.size() + 1; k += 2)
        if (I(f, k)) aset.push_back(I(f, k));
    } else if (c.name == "ASET1") {
      std::vector<int> ranged;
      if (!grid_list(f, 1, c.name, aset, ranged, err)) return -2;
      aset_range.insert(aset_range.end(), ranged.begin(), ranged.end());
    } else if (c.name == "SPC1") {
      int mask = 0;
      for (char ch : f.size() > 1 ? f[1] : std::string())
        if (ch >= '1' && ch <= '6') mask |= 1 << (ch - '1');
      std::vector<int> listed, ranged;
      if (!grid_list(f, 2, c.name, listed, ranged, err)) return -2;
      for (int g : listed) spc[g] |= mask;
      for (int g : ranged) spc[g] |= mask;
    } else if (c.name == "FORCE" || c.name == "MOMENT") {
      if (I(f, 2) != 0) { err = c.name + ": only the basic coordinate system (CID 0) is supported"; return -2; }
      double F = D(f, 3);
      loads.push_back({I(f, 0), I(f, 1), c.name == "FORCE" ? 0 : 3, {F * D(f, 4), F * D(f, 5), F * D(f, 6)}});
    } else if (c.name == "RBE3") {
      // EID, blank, REFGRID, REFC, then groups WT, C, G...; a real field
      // starts a group.  The UM (extra dependent DOFs) section is not
      // supported; ALPHA (thermal expansion) does not enter the reduction.
      R3 r;
      r.eid = I(f, 0);
      r.ref = I(f, 2);
      r.refc = I(f, 3);
      int state = 0;  // 0 expect weight, 1 expect components, 2 grids
      for (size_t k = 4; k < f.size(); k++) {
        std::string t = trim(f[k]);
        if (t.empty()) continue;
        std::string up = t;
        for (char& ch : up) ch = (char)std::toupper(ch);
        if (up == "UM") { err = "RBE3 " + std::to_string(r.eid) + ": UM dependent DOFs are not supported"; return -2; }
        if (up == "ALPHA" || up == "TREF") break;
        int iv = 0;
        double dv = 0.0;
        const bool is_int = to_int(t, iv);
        if (!is_int && to_double(t, dv)) {
          r.groups.push_back({dv, 0, {}});
          state = 1;
        } else if (is_int && state == 1) {
          for (char ch : t)
            if (ch >= '1' && ch <= '6') r.groups.back().comps |= 1 << (ch - '1');
          state = 2;
        } else if (is_int && state == 2) {
          r.groups.back().grids.push_back(iv);
        } else {
          err = "RBE3 " + std::to_string(r.eid) + ": unexpected field '" + t + "'";
          return -2;
        }
      }
      rbe3.push_back(r);
    } else if (c.name == "RBE2") {
      std::vector<int> dep;
      for (size_t k = 3; k < f.size(); k++)
        if (I(f, k)) dep.push_back(I(f, k));
      rbe2.push_back({I(f, 1), {f.size() > 2 ? f[2] : std::string(), dep}});
    }
  }
  NastranPart& np = out;
  np = NastranPart();
  Part& p = np.part;
  std::map<int, int> idx;
  for (auto& g : grids) {
    idx[g.first] = p.nnod++;
    np.node_ids.push_back(g.first);
    for (int k = 0; k < 3; k++) p.xyz.push_back(g.second.x[k]);
  }
  p.status.assign(p.nnod, 0);
  for (auto& s : spc)
    if (idx.count(s.first)) p.status[idx[s.first]] = -s.second;
  for (int g : aset) {
    if (!idx.count(g)) { err = "ASET grid " + std::to_string(g) + " not defined"; return -2; }
    p.status[idx[g]] = 1;
  }
  for (int g : aset_range)
    if (idx.count(g)) p.status[idx[g]] = 1;
  p.eptr.push_back(0);
  for (auto& e : elems) {
    if (!props.count(e.pid)) { err = "element " + std::to_string(e.eid) + ": missing property"; return -2; }
    const Prop& pr = props[e.pid];
    if (!mats.count(pr.mid)) { err = "element " + std::to_string(e.eid) + ": missing material"; return -2; }
    const Mat& m = mats[pr.mid];
    p.etype.push_back(e.type);
    for (int n : e.nodes) {
      if (!idx.count(n)) { err = "element " + std::to_string(e.eid) + ": undefined grid"; return -2; }
      p.enodes.push_back(idx[n]);
    }
    p.eptr.push_back((int)p.enodes.size());
    std::vector<double> ep(NPROP, 0.0);
    ep[P_E] = m.E;
    ep[P_NU] = m.nu;
    ep[P_RHO] = m.rho;
    ep[P_G] = m.hasG ? m.G : m.E / (2.0 * (1.0 + m.nu));
    if (e.type == ANDES4)
      for (int k = 0; k < 4; k++) ep[P_T1 + k] = pr.t;
    if (e.type == BEAM2) {
      ep[P_A] = pr.A; ep[P_IY] = pr.I1; ep[P_IZ] = pr.I2; ep[P_IT] = pr.J;
      ep[P_DX] = e.orient[0]; ep[P_DY] = e.orient[1]; ep[P_DZ] = e.orient[2];
      ep[P_KY] = pr.K1; ep[P_KZ] = pr.K2; ep[P_NSM] = pr.nsm;
    }
    p.eprop.insert(p.eprop.end(), ep.begin(), ep.end());
    p.nel++;
  }
  for (auto& r : rbe2) {
    Rbe2 c;
    if (!idx.count(r.first)) { err = "RBE2: undefined independent grid"; return -2; }
    c.indep = idx[r.first];
    c.dofs = 0;
    for (char ch : r.second.first)
      if (ch >= '1' && ch <= '6') c.dofs |= 1 << (ch - '1');
    for (int g : r.second.second) {
      if (!idx.count(g)) { err = "RBE2: undefined dependent grid"; return -2; }
      c.dep.push_back(idx[g]);
    }
    np.rbe2.push_back(c);
  }
  for (auto& r : rbe3) {
    const std::string tag = "RBE3 " + std::to_string(r.eid);
    if (!idx.count(r.ref)) { err = tag + ": undefined reference grid"; return -2; }
    Rbe3 c;
    c.ref = idx[r.ref];
    c.refc = r.refc;
    // independent nodes in order of first appearance; per component the
    // weight of each node (summed over the groups that list it with that
    // component), components that share weights share one weight set
    std::map<int, int> pos;
    for (auto& g : r.groups)
      for (int gid : g.grids) {
        if (!idx.count(gid)) { err = tag + ": undefined independent grid"; return -2; }
        if (!pos.count(gid)) {
          pos[gid] = (int)c.nodes.size();
          c.nodes.push_back(idx[gid]);
        }
      }
    const int nM = (int)c.nodes.size();
    if (nM == 0) { err = tag + ": no independent grids"; return -2; }
    for (int comp = 0; comp < 6; comp++) {
      std::vector<double> w(nM, 0.0);
      bool any = false;
      for (auto& g : r.groups)
        if ((g.comps >> comp) & 1)
          for (int gid : g.grids) {
            w[pos[gid]] += g.wt;
            any = true;
          }
      if (!any) continue;
      for (int prev = 0; prev < comp && !c.indc[comp]; prev++)
        if (c.indc[prev] > 0 && std::equal(w.begin(), w.end(), c.w.begin() + c.indc[prev] - 1)) c.indc[comp] = c.indc[prev];
      if (!c.indc[comp]) {
        c.indc[comp] = (int)c.w.size() + 1;
        c.w.insert(c.w.end(), w.begin(), w.end());
      }
    }
    np.rbe3.push_back(std::move(c));
  }
  for (auto& l : loads) {
    if (!idx.count(l.grid)) { err = "load on undefined grid " + std::to_string(l.grid); return -2; }
    for (int k = 0; k < 3; k++)
      if (l.v[k] != 0.0) np.loads.push_back({l.sid, idx[l.grid], l.dof0 + k, l.v[k]});
  }
  return 0;
}

}  // namespace fgpu
