// CHARMM files on the host: PSF, CRD, parameter / stream files, restart.
//
// Reference behaviour restated (file formats and key-matching rules):
//   CharmmPSF::readCharmmPSF   src/CharmmPSF.cu:387+ (title, atom section,
//                              bond/angle/dihedral/improper/cross-term lists)
//   CharmmCrd                  src/CharmmCrd.cu:19-59 (title, count, atoms)
//   CharmmParameters           src/CharmmParameters.cu:57-335 (sections),
//                              getBondedParamsAndLists :357-714,
//                              getVdwParamsAndTypes :718-844
//   RestartSubscriber::update  src/RestartSubscriber.cu:73-445
// Type names are matched upper-case; the first entry of a bond, angle,
// improper or NBFIX key wins, a repeated NONBONDED type takes the last entry,
// every line of a dihedral key adds a Fourier term.  Angles with a
// Urey-Bradley constant of at most 0.01 get no Urey-Bradley term.
#include "charmm_files.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <set>
#include <sstream>
#include <stdexcept>

#include "common.hpp"

namespace apo {
namespace files {

namespace {

constexpr double kDeg = 3.14159265358979323846 / 180.0;

std::vector<std::string> read_lines(const std::string &path) {
  std::ifstream f(path);
  if (!f.is_open()) throw RuntimeError("could not open file \"" + path + "\"");
  std::vector<std::string> out;
  std::string line;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    out.push_back(line);
  }
  return out;
}

std::vector<std::string> split(const std::string &s) {
  std::istringstream is(s);
  std::vector<std::string> out;
  std::string t;
  while (is >> t) out.push_back(t);
  return out;
}

std::string upper(std::string s) {
  for (char &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

std::string strip(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

double to_d(const std::string &s, const std::string &where) {
  try {
    size_t used = 0;
    double v = std::stod(s, &used);
    if (used != s.size()) throw std::invalid_argument(s);
    return v;
  } catch (const std::exception &) {
    throw RuntimeError("not a number \"" + s + "\" in " + where);
  }
}

int to_i(const std::string &s, const std::string &where) {
  try {
    size_t used = 0;
    long v = std::stol(s, &used);
    if (used != s.size()) throw std::invalid_argument(s);
    return (int)v;
  } catch (const std::exception &) {
    throw RuntimeError("not an integer \"" + s + "\" in " + where);
  }
}

// Dihedral / improper key: smaller outer type first; inner pair sorted when
// the outer types are equal.
Key4 dihedral_key(std::string a, std::string b, std::string c, std::string d) {
  if (a > d) {
    std::swap(a, d);
    std::swap(b, c);
  } else if (a == d && b > c) {
    std::swap(b, c);
  }
  return {a, b, c, d};
}

}  // namespace

// ---------------------------------------------------------------------------
std::vector<int32_t> Psf::waters() const {
  std::vector<int32_t> out;
  const int n = natom();
  for (int i = 0; i + 2 < n;) {
    if (upper(types[i]) == "OT" && upper(types[i + 1]) == "HT" && upper(types[i + 2]) == "HT") {
      out.insert(out.end(), {i, i + 1, i + 2});
      i += 3;
    } else {
      i++;
    }
  }
  return out;
}

Psf read_psf(const std::string &path) {
  const auto lines = read_lines(path);
  Psf p;
  p.file_name = path;
  // section headers: "<count> !NAME[: comment]"
  std::map<std::string, std::pair<size_t, long>> sec;
  for (size_t i = 0; i < lines.size(); i++) {
    const auto tok = split(lines[i]);
    if (tok.size() >= 2 && tok[1].rfind("!N", 0) == 0) {
      std::string key = tok[1];
      const size_t colon = key.find(':');
      if (colon != std::string::npos) key = key.substr(0, colon);
      sec[key] = {i + 1, std::stol(tok[0])};
    }
  }
  auto it = sec.find("!NATOM");
  if (it == sec.end()) throw RuntimeError("no !NATOM section in PSF \"" + path + "\"");
  const size_t a0 = it->second.first;
  const long n = it->second.second;
  if (n < 0 || a0 + (size_t)n > lines.size())
    throw RuntimeError("truncated atom section in PSF \"" + path + "\"");
  for (long k = 0; k < n; k++) {
    const auto f = split(lines[a0 + k]);
    if (f.size() < 8)
      throw RuntimeError("short atom line " + std::to_string(k + 1) + " in PSF \"" + path + "\"");
    p.segids.push_back(f[1]);
    p.resids.push_back(to_i(f[2], "PSF atom line"));
    p.resnames.push_back(f[3]);
    p.names.push_back(f[4]);
    p.types.push_back(f[5]);
    p.charges.push_back(to_d(f[6], "PSF atom line"));
    p.masses.push_back(to_d(f[7], "PSF atom line"));
  }
  auto ints = [&](const char *key, int per, std::vector<int32_t> &out) {
    auto s = sec.find(key);
    if (s == sec.end()) return;
    const long want = s->second.second * per;
    size_t j = s->second.first;
    while ((long)out.size() < want) {
      if (j >= lines.size())
        throw RuntimeError(std::string("truncated ") + key + " section in PSF \"" + path + "\"");
      for (const auto &t : split(lines[j])) {
        if ((long)out.size() == want) break;
        const int v = to_i(t, key) - 1;
        if (v < 0 || v >= n) throw RuntimeError(std::string("atom index out of range in ") + key);
        out.push_back(v);
      }
      j++;
    }
  };
  ints("!NBOND", 2, p.bonds);
  ints("!NTHETA", 3, p.angles);
  ints("!NPHI", 4, p.dihedrals);
  ints("!NIMPHI", 4, p.impropers);
  ints("!NCRTERM", 8, p.crossterms);
  return p;
}

Crd read_crd(const std::string &path) {
  std::vector<std::string> lines;
  for (const auto &l : read_lines(path))
    if (!l.empty() && l[0] != '*' && !strip(l).empty()) lines.push_back(l);
  if (lines.empty()) throw RuntimeError("empty CRD file \"" + path + "\"");
  const int n = to_i(split(lines[0]).at(0), "CRD atom count");
  if (n < 0 || (size_t)n + 1 > lines.size())
    throw RuntimeError("truncated CRD file \"" + path + "\"");
  Crd c;
  c.xyz.resize(3 * (size_t)n);
  for (int k = 0; k < n; k++) {
    const auto f = split(lines[1 + k]);
    if (f.size() < 7) throw RuntimeError("short CRD atom line in \"" + path + "\"");
    for (int d = 0; d < 3; d++) c.xyz[3 * (size_t)k + d] = to_d(f[4 + d], "CRD atom line");
  }
  return c;
}

// ---------------------------------------------------------------------------
void Params::read(const std::string &path) {
  auto lines = read_lines(path);
  // a stream file carries topology first; parameters follow "READ PARA"
  for (size_t k = 0; k < lines.size(); k++) {
    const std::string ln = upper(strip(lines[k].substr(0, lines[k].find('!'))));
    if (ln.rfind("READ", 0) == 0 && ln.find("PARA") != std::string::npos) {
      lines.erase(lines.begin(), lines.begin() + (long)k + 1);
      break;
    }
  }
  static const std::pair<const char *, const char *> heads[] = {
      {"ATOM", "atoms"},     {"BOND", "bonds"},     {"ANGL", "angles"},    {"THET", "angles"},
      {"DIHE", "dihedrals"}, {"PHI", "dihedrals"},  {"IMPR", "impropers"}, {"IMPH", "impropers"},
      {"CMAP", "cmap"},      {"NONB", "nonbonded"}, {"NBON", "nonbonded"}, {"NBFI", "nbfix"},
      {"HBON", "hbond"}};
  std::string state;
  for (size_t k = 0; k < lines.size();) {
    std::string ln = upper(strip(lines[k].substr(0, lines[k].find('!'))));
    k++;
    if (ln.empty() || ln[0] == '*') continue;
    auto tok = split(ln);
    const std::string &head = tok[0];
    if (head == "END" || head == "RETURN") {
      state.clear();
      continue;
    }
    std::string next;
    for (const auto &h : heads) {
      const std::string st = h.second;
      if (head.rfind(h.first, 0) == 0 &&
          (tok.size() == 1 || st == "nonbonded" || st == "hbond")) {
        next = st;
        break;
      }
    }
    if (!next.empty()) {
      state = next;
      if (state == "nonbonded")  // the option list may continue over lines ending in "-"
        while (!ln.empty() && ln.back() == '-' && k < lines.size()) {
          ln = upper(strip(lines[k].substr(0, lines[k].find('!'))));
          k++;
        }
      continue;
    }
    const std::string where = "parameter file \"" + path + "\"";
    if (state == "bonds" && tok.size() >= 4) {
      Key2 key{tok[0], tok[1]};
      std::sort(key.begin(), key.end());
      bonds.emplace(key, std::array<double, 2>{to_d(tok[3], where), to_d(tok[2], where)});
    } else if (state == "angles" && tok.size() >= 5) {
      Key3 key{tok[0], tok[1], tok[2]};
      if (key[0] > key[2]) std::swap(key[0], key[2]);
      angles.emplace(key, std::array<double, 2>{to_d(tok[4], where) * kDeg, to_d(tok[3], where)});
      ureyb.emplace(key, tok.size() >= 7
                             ? std::array<double, 2>{to_d(tok[6], where), to_d(tok[5], where)}
                             : std::array<double, 2>{0.0, 0.0});
    } else if (state == "dihedrals" && tok.size() >= 7) {
      const Key4 key = dihedral_key(tok[0], tok[1], tok[2], tok[3]);
      dihedrals[key].push_back({(double)to_i(tok[5], where), to_d(tok[4], where),
                                to_d(tok[6], where) * kDeg});
    } else if (state == "impropers" && tok.size() >= 7) {
      const Key4 key = dihedral_key(tok[0], tok[1], tok[2], tok[3]);
      impropers.emplace(key, std::array<double, 3>{0.0, to_d(tok[4], where),
                                                   to_d(tok[6], where) * kDeg});
    } else if (state == "nonbonded") {
      if (tok.size() == 4) {
        vdw[tok[0]] = {to_d(tok[2], where), to_d(tok[3], where)};
      } else if (tok.size() == 7) {
        vdw[tok[0]] = {to_d(tok[2], where), to_d(tok[3], where)};
        vdw14[tok[0]] = {to_d(tok[5], where), to_d(tok[6], where)};
      } else {
        throw RuntimeError("unexpected NONBONDED line in " + where + ": " + ln);
      }
    } else if (state == "nbfix" && tok.size() >= 4) {
      Key2 key{tok[0], tok[1]};
      std::sort(key.begin(), key.end());
      nbfix.emplace(key, std::array<double, 2>{std::fabs(to_d(tok[2], where)), to_d(tok[3], where)});
    }
  }
}

// ---------------------------------------------------------------------------
Resolved resolve(const Psf &psf, const Params &prm) {
  Resolved r;
  const int n = psf.natom();
  std::vector<std::string> t(n);
  for (int i = 0; i < n; i++) t[i] = upper(psf.types[i]);

  // vdW types: the sorted set of atom types (getVdwParamsAndTypes)
  std::set<std::string> tset(t.begin(), t.end());
  r.type_names.assign(tset.begin(), tset.end());
  r.ntypes = (int)r.type_names.size();
  std::map<std::string, int> index;
  for (int k = 0; k < r.ntypes; k++) index[r.type_names[k]] = k;
  std::string missing;
  for (const auto &nm : r.type_names)
    if (!prm.vdw.count(nm)) missing += " " + nm;
  if (!missing.empty()) throw InvalidArgument("no NONBONDED entry for types:" + missing);
  const int npair = r.ntypes * (r.ntypes + 1) / 2;
  std::vector<double> tab(2 * (size_t)npair), tab14(2 * (size_t)npair);
  auto par = [&](const std::string &nm, bool one4) {
    if (one4) {
      auto f = prm.vdw14.find(nm);
      if (f != prm.vdw14.end()) return f->second;
    }
    return prm.vdw.at(nm);
  };
  for (int a = 0; a < r.ntypes; a++)
    for (int b = 0; b <= a; b++) {
      const int row = a * (a + 1) / 2 + b;
      for (int one4 = 0; one4 < 2; one4++) {
        const auto pa = par(r.type_names[a], one4), pb = par(r.type_names[b], one4);
        const double e = std::sqrt(pa[0] * pb[0]), rm = pa[1] + pb[1];
        double *dst = one4 ? tab14.data() : tab.data();
        dst[2 * row] = 2.0 * e * std::pow(rm, 6);
        dst[2 * row + 1] = e * std::pow(rm, 12);
      }
    }
  for (const auto &kv : prm.nbfix) {
    auto ia = index.find(kv.first[0]), ib = index.find(kv.first[1]);
    if (ia == index.end() || ib == index.end()) continue;
    const int hi = std::max(ia->second, ib->second), lo = std::min(ia->second, ib->second);
    const int row = hi * (hi + 1) / 2 + lo;
    const double emin = kv.second[0], rmin = kv.second[1];
    for (double *dst : {tab.data(), tab14.data()}) {
      dst[2 * row] = 2.0 * emin * std::pow(rmin, 6);
      dst[2 * row + 1] = emin * std::pow(rmin, 12);
    }
  }
  r.vdwtab.assign(tab.begin(), tab.end());
  r.vdwtab14.assign(tab14.begin(), tab14.end());
  r.vdwtype.resize(n);
  for (int i = 0; i < n; i++) r.vdwtype[i] = index[t[i]];

  // bonded lists (getBondedParamsAndLists)
  std::map<Key2, double> bond_b0;
  for (size_t k = 0; k + 1 < psf.bonds.size(); k += 2) {
    const int i = psf.bonds[k], j = psf.bonds[k + 1];
    Key2 key{t[i], t[j]};
    std::sort(key.begin(), key.end());
    auto f = prm.bonds.find(key);
    if (f == prm.bonds.end()) throw InvalidArgument("bond not found: " + key[0] + " " + key[1]);
    r.bonds.insert(r.bonds.end(), {i, j});
    r.bond_par.insert(r.bond_par.end(), {(float)f->second[0], (float)f->second[1]});
    bond_b0[key] = f->second[0];
  }
  for (size_t k = 0; k + 2 < psf.angles.size(); k += 3) {
    const int i = psf.angles[k], j = psf.angles[k + 1], l = psf.angles[k + 2];
    Key3 key{t[i], t[j], t[l]};
    if (key[0] > key[2]) std::swap(key[0], key[2]);
    auto f = prm.angles.find(key);
    if (f == prm.angles.end())
      throw InvalidArgument("angle not found: " + key[0] + " " + key[1] + " " + key[2]);
    r.angles.insert(r.angles.end(), {i, j, l});
    r.angle_par.insert(r.angle_par.end(), {(float)f->second[0], (float)f->second[1]});
    auto u = prm.ureyb.find(key);
    if (u != prm.ureyb.end() && std::fabs(u->second[1]) > 0.01) {
      r.ureyb.insert(r.ureyb.end(), {i, l});
      r.ub_par.insert(r.ub_par.end(), {(float)u->second[0], (float)u->second[1]});
    }
  }
  for (size_t k = 0; k + 3 < psf.dihedrals.size(); k += 4) {
    const int *q = &psf.dihedrals[k];
    const Key4 key = dihedral_key(t[q[0]], t[q[1]], t[q[2]], t[q[3]]);
    auto f = prm.dihedrals.find(key);
    if (f == prm.dihedrals.end()) {
      std::string b = t[q[1]], c = t[q[2]];
      if (b > c) std::swap(b, c);
      f = prm.dihedrals.find(Key4{"X", b, c, "X"});
    }
    if (f == prm.dihedrals.end())
      throw InvalidArgument("dihedral not found: " + t[q[0]] + " " + t[q[1]] + " " + t[q[2]] +
                            " " + t[q[3]]);
    for (const auto &term : f->second) {
      r.dihe.insert(r.dihe.end(), q, q + 4);
      r.dihe_par.insert(r.dihe_par.end(), {(float)term[0], (float)term[1], (float)term[2]});
    }
  }
  for (size_t k = 0; k + 3 < psf.impropers.size(); k += 4) {
    const int *q = &psf.impropers[k];
    Key4 key = dihedral_key(t[q[0]], t[q[1]], t[q[2]], t[q[3]]);
    auto f = prm.impropers.find(key);
    if (f == prm.impropers.end()) f = prm.impropers.find(Key4{key[0], "X", "X", key[3]});
    if (f == prm.impropers.end())
      throw InvalidArgument("improper not found: " + t[q[0]] + " " + t[q[1]] + " " + t[q[2]] +
                            " " + t[q[3]]);
    r.imdihe.insert(r.imdihe.end(), q, q + 4);
    r.imdihe_par.insert(r.imdihe_par.end(),
                        {(float)f->second[0], (float)f->second[1], (float)f->second[2]});
  }
  r.waters = psf.waters();

  // SHAKE groups: bonds with a hydrogen (type name starting with H), except
  // water O-H and H-H bonds, grouped by their heavy atom
  std::vector<std::vector<int>> hyd(n);
  for (size_t k = 0; k + 1 < psf.bonds.size(); k += 2) {
    const int i = psf.bonds[k], j = psf.bonds[k + 1];
    const bool hi = t[i][0] == 'H', hj = t[j][0] == 'H';
    if (!(hi || hj)) continue;
    if ((t[i] == "OT" && t[j] == "HT") || (t[i] == "HT" && t[j] == "OT") || (hi && hj)) continue;
    const int heavy = hi ? j : i, h = hi ? i : j;
    hyd[heavy].push_back(h);
  }
  for (int i = 0; i < n; i++) {
    if (hyd[i].empty()) continue;
    if (hyd[i].size() > 3)
      throw InvalidArgument("atom " + std::to_string(i + 1) +
                            " carries more than three hydrogens (SHAKE groups hold three)");
    int g[4] = {i, -1, -1, -1};
    double b0 = 0.0;
    for (size_t k = 0; k < hyd[i].size(); k++) {
      g[k + 1] = hyd[i][k];
      Key2 key{t[i], t[hyd[i][k]]};
      std::sort(key.begin(), key.end());
      b0 = bond_b0.at(key);
    }
    r.shake_groups.insert(r.shake_groups.end(), g, g + 4);
    r.shake_r2.push_back(b0 * b0);
  }
  return r;
}

// ---------------------------------------------------------------------------
namespace {

// Fortran-style scientific text: 15 digits after the point, exponent with D
// (apo::cDoubleToFortSciStr, src/str_utils.cpp:138-144), right-aligned in 22.
std::string fsci(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.15e", v);
  std::string s(buf);
  for (char &c : s)
    if (c == 'e') c = 'D';
  if (s.size() < 22) s = std::string(22 - s.size(), ' ') + s;
  return s;
}

double fsci_read(const std::string &line, size_t at) {
  if (line.size() < at + 1) throw RuntimeError("short restart line");
  std::string s = line.substr(at, 22);
  for (char &c : s)
    if (c == 'D' || c == 'd') c = 'e';
  return to_d(strip(s), "restart file");
}

std::string pad_left(const std::string &s, size_t w) {
  return s.size() >= w ? s : std::string(w - s.size(), ' ') + s;
}

}  // namespace

void write_restart(const std::string &path, const Restart &r) {
  std::ofstream f(path);
  if (!f.is_open()) throw RuntimeError("could not open restart file \"" + path + "\"");
  const bool cubic = r.box[0] == r.box[1] && r.box[0] == r.box[2];
  const std::string xtl = cubic ? "CUBI" : (r.box[0] == r.box[1] ? "TETR" : "ORTH");
  f << "REST" << pad_left("50", 6) << pad_left("1", 6) << "  " << xtl << "  "
    << "    "
    << "  "
    << "APO" << '\n';
  f << '\n';
  f << pad_left("2", 8) << " !NTITLE followed by title\n";
  f << "* APOCHARMM RESTART FILE                                                        \n";
  f << "* USED FOR CONTINUING MOLECULAR DYNAMICS TRAJECTORY                             \n";
  f << '\n';
  f << " !CRYSTAL PARAMETERS\n";
  // XTLABC (lower triangle a, 0, b, 0, 0, c), HDOT, PNH/PNHV/PNHF, UC1A..UC2B,
  // the four GRAD values: zero for the Langevin thermostat
  f << fsci(r.box[0]) << fsci(0.0) << fsci(r.box[1]) << '\n';
  f << fsci(0.0) << fsci(0.0) << fsci(r.box[2]) << '\n';
  for (int row = 0; row < 11; row++) f << fsci(0.0) << fsci(0.0) << fsci(0.0) << '\n';
  f << fsci(0.0) << '\n';
  f << '\n';
  f << " !NATOM,NPRIV,NSTEP,NSAVC,NSAVV,JHSTRT,NDEGF,SEED,NSAVL\n";
  f << pad_left(std::to_string(r.natom), 12) << pad_left(std::to_string(r.npriv), 12)
    << pad_left(std::to_string(r.nstep), 12) << pad_left(std::to_string(r.nsavc), 12)
    << pad_left("0", 12) << pad_left("0", 12) << pad_left(std::to_string(r.ndegf), 12)
    << pad_left(std::to_string(r.seed), 22) << "  " << r.rng_state << '\n';
  f << '\n';
  // energies and statistics: zero, as the reference writes them
  constexpr int kLenEnp = 60, kLenEnt = 128, kLenEnv = 50;
  f << " !ENERGIES and STATISTICS\n";
  f << std::string(kLenEnp, 'T') << '\n' << std::string(kLenEnt, 'T') << '\n';
  f << pad_left("0", 8) << fsci(0.0) << fsci(0.0) << fsci(r.avetem) << '\n';
  auto zero_rows = [&](int rows, int cols) {
    for (int i = 0; i < rows; i++) {
      for (int c = 0; c < cols; c++) f << fsci(0.0);
      f << '\n';
    }
  };
  zero_rows(kLenEnp, 3);
  zero_rows(kLenEnp, 2);
  zero_rows(kLenEnt, 3);
  zero_rows(kLenEnt, 2);
  zero_rows(kLenEnv, 3);
  zero_rows(kLenEnv, 2);
  f << '\n';
  auto section = [&](const char *title, const std::vector<double> &a) {
    f << title << '\n';
    for (int i = 0; i < r.natom; i++)
      f << fsci(a[3 * (size_t)i]) << fsci(a[3 * (size_t)i + 1]) << fsci(a[3 * (size_t)i + 2])
        << '\n';
  };
  section(" !XOLD, YOLD, ZOLD", r.x);
  f << '\n';
  section(" !VX, VY, VZ", r.v);
  f << '\n';
  section(" !X, Y, Z", r.dprev);
  if (!f.good()) throw RuntimeError("writing restart file \"" + path + "\" failed");
}

Restart read_restart(const std::string &path) {
  const auto lines = read_lines(path);
  Restart r;
  if (lines.empty()) throw RuntimeError("empty restart file \"" + path + "\"");
  r.apo = lines[0].size() >= 33 && lines[0].substr(30, 3) == "APO";
  size_t k = 0;
  auto find = [&](const std::string &title) {
    while (k < lines.size() && lines[k] != title) k++;
    if (k == lines.size())
      throw RuntimeError("could not find " + strip(title) + " section in \"" + path + "\"");
    k++;
  };
  find(" !CRYSTAL PARAMETERS");
  if (k + 1 >= lines.size()) throw RuntimeError("truncated restart file \"" + path + "\"");
  double xtl[6];
  for (int row = 0; row < 2; row++)
    for (int c = 0; c < 3; c++) xtl[3 * row + c] = fsci_read(lines[k + row], 22 * (size_t)c);
  r.box[0] = xtl[0];
  r.box[1] = xtl[2];
  r.box[2] = xtl[5];
  find(" !NATOM,NPRIV,NSTEP,NSAVC,NSAVV,JHSTRT,NDEGF,SEED,NSAVL");
  if (k >= lines.size()) throw RuntimeError("truncated restart file \"" + path + "\"");
  const std::string &ln = lines[k];
  if (ln.size() < 106) throw RuntimeError("short integer line in restart file \"" + path + "\"");
  r.natom = to_i(strip(ln.substr(0, 12)), "restart NATOM");
  r.npriv = std::stoull(strip(ln.substr(12, 12)));
  r.nstep = to_i(strip(ln.substr(24, 12)), "restart NSTEP");
  r.nsavc = to_i(strip(ln.substr(36, 12)), "restart NSAVC");
  r.ndegf = to_i(strip(ln.substr(72, 12)), "restart NDEGF");
  r.seed = std::stoull(strip(ln.substr(84, 22)));
  const std::string rng = strip(ln.size() > 106 ? ln.substr(106) : std::string());
  r.rng_state = r.apo && !rng.empty() ? std::stoull(rng) : 0ull;
  if (r.natom <= 0) throw RuntimeError("restart file has no atoms");
  auto read_xyz = [&](const std::string &title, std::vector<double> &out) {
    find(title);
    if (k + (size_t)r.natom > lines.size())
      throw RuntimeError("truncated" + title + " section in \"" + path + "\"");
    out.resize(3 * (size_t)r.natom);
    for (int i = 0; i < r.natom; i++)
      for (int c = 0; c < 3; c++) out[3 * (size_t)i + c] = fsci_read(lines[k + i], 22 * (size_t)c);
    k += (size_t)r.natom;
  };
  read_xyz(" !XOLD, YOLD, ZOLD", r.x);
  read_xyz(" !VX, VY, VZ", r.v);
  read_xyz(" !X, Y, Z", r.dprev);
  return r;
}

}  // namespace files
}  // namespace apo
