// Native readers for the wavefunction formats besides fchk that the
// reference's loader accepts through iodata.load_one (src/iodata.cpp:17-20):
//
//   * AIM .wfn and AIMAll .wfx: molecular orbitals over uncontracted
//     Cartesian Gaussian primitives x^a y^b z^c exp(-alpha r^2) with
//     UNnormalised coefficients, one type code per primitive (1 s, 2-4 p,
//     5-10 d, 11-20 f, 21-35 g);
//   * Molden: contracted shells per atom ([GTO]), MO coefficients over
//     normalised contracted functions ([MO]), [5D]/[7F]/[9G] flags for pure
//     shells.
//
// All produce the FchkData the fchk reader produces: the basis in the
// reference's conventions (shells with L2-normalised primitives, Cartesian
// components in the fchk order, pure components c0, c1, s1, ...), bohr
// coordinates, nuclear charges and P = sum_j occ_j C_j C_j^T over every
// listed orbital (both spins when the file lists them separately).

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "mol.h"

namespace ct {

namespace {

constexpr double kPi = 3.14159265358979323846;
// bohr per angstrom, 1 / 0.52917721092 (CODATA 2010)
constexpr double kAngstrom = 1.0 / 0.52917721092;

std::string lower(std::string s) {
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// Every number in s, Fortran D exponents accepted; fixed-width fields that
// run together ("0.1D+01-0.2D+01") are split where the next sign starts.
std::vector<double> numbers(std::string s, const std::string& where) {
  for (char& c : s)
    if (c == 'D' || c == 'd') c = 'E';
  std::vector<double> out;
  const char* p = s.c_str();
  while (*p) {
    while (*p && std::isspace((unsigned char)*p)) ++p;
    if (!*p) break;
    char* e = nullptr;
    const double v = std::strtod(p, &e);
    if (e == p) throw Error(CT_E_IO, "cannot parse a number in '" + trim(s) + "' (" + where + ")");
    out.push_back(v);
    p = e;
  }
  return out;
}

double number(const std::string& s, const std::string& where) {
  const auto v = numbers(s, where);
  if (v.size() != 1) throw Error(CT_E_IO, "expected one number in '" + trim(s) + "' (" + where + ")");
  return v[0];
}

const char* const kSymbols[] = {
    "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar",
    "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr",
    "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe",
    "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf",
    "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn"};
constexpr int kNumSymbols = sizeof(kSymbols) / sizeof(kSymbols[0]);

// Atomic number from an atom label such as "O", "O1", "cl12"; 0 if unknown.
int z_from_label(const std::string& label) {
  std::string letters;
  for (char c : label) {
    if (!std::isalpha((unsigned char)c)) break;
    letters += c;
  }
  if (letters.empty()) return 0;
  for (int len : {2, 1}) {
    if ((int)letters.size() < len) continue;
    const std::string cand = lower(letters.substr(0, len));
    for (int z = 0; z < kNumSymbols; ++z)
      if (lower(kSymbols[z]) == cand) return z + 1;
  }
  return 0;
}

// Cartesian exponents in the reference's component order (fchk order:
// src/evaluate_density.cu:80 for d, :143 for f, :235-236 for g).
std::vector<std::array<int, 3>> cart_order(int L) {
  switch (L) {
    case 0: return {{0, 0, 0}};
    case 1: return {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    case 2: return {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    case 3:
      return {{3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {1, 2, 0}, {2, 1, 0}, {2, 0, 1}, {1, 0, 2}, {0, 1, 2}, {0, 2, 1}, {1, 1, 1}};
    case 4:
      return {{0, 0, 4}, {0, 1, 3}, {0, 2, 2}, {0, 3, 1}, {0, 4, 0}, {1, 0, 3}, {1, 1, 2}, {1, 2, 1},
              {1, 3, 0}, {2, 0, 2}, {2, 1, 1}, {2, 2, 0}, {3, 0, 1}, {3, 1, 0}, {4, 0, 0}};
  }
  throw Error(CT_E_UNSUPPORTED, "angular momentum above g");
}

int cart_index(int L, int a, int b, int c) {
  const auto o = cart_order(L);
  for (size_t k = 0; k < o.size(); ++k)
    if (o[k][0] == a && o[k][1] == b && o[k][2] == c) return (int)k;
  throw Error(CT_E_IO, "no Cartesian component x^a y^b z^c of this order");
}

double dfact(int n) {  // n!!, (-1)!! = 1
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

// L2 norm of the Cartesian primitive x^a y^b z^c exp(-alpha r^2) is 1 / this.
double cart_norm(double alpha, int a, int b, int c) {
  const int L = a + b + c;
  return std::sqrt(std::pow(2.0 * alpha / kPi, 1.5) * std::pow(4.0 * alpha, L) /
                   (dfact(2 * a - 1) * dfact(2 * b - 1) * dfact(2 * c - 1)));
}

void accumulate_density(std::vector<double>& P, int M, const double* c, double occ) {
  if (occ == 0.0) return;
  for (int i = 0; i < M; ++i) {
    const double ci = occ * c[i];
    if (ci == 0.0) continue;
    double* row = &P[(size_t)i * M];
    for (int j = 0; j < M; ++j) row[j] += ci * c[j];
  }
}

// ---------------------------------------------------------------- wfn / wfx

// Exponents (a, b, c) of the AIM primitive type codes 1..35.
constexpr int kTypeExp[36][3] = {
    {0, 0, 0},                                                                                   // (unused)
    {0, 0, 0},                                                                                   // 1 s
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},                                                             // 2-4 p
    {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},                            // 5-10 d
    {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {2, 0, 1}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2},      // 11-18 f
    {0, 1, 2}, {1, 1, 1},                                                                        // 19-20 f
    {4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {3, 1, 0}, {3, 0, 1}, {1, 3, 0}, {0, 3, 1}, {1, 0, 3},      // 21-28 g
    {0, 1, 3}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};                // 29-35 g

struct PrimitiveWfn {
  std::vector<double> atom_xyz, charges;
  std::vector<int32_t> numbers;
  std::vector<int> center, type;  // per primitive (center 0-based)
  std::vector<double> exps;
  std::vector<double> occ;        // per orbital
  std::vector<double> coef;       // orbital-major: coef[j * nprim + p]
};

// Primitives become Cartesian shells of one primitive each: consecutive
// primitives with the same centre, exponent and L form one shell (the
// components a file leaves out get zero coefficients).  The coefficient of
// our normalised component is the file's coefficient divided by the
// primitive's normalisation constant.
FchkData from_primitives(const PrimitiveWfn& w, const std::string& path) {
  const int nprim = (int)w.exps.size();
  const int nmo = (int)w.occ.size();
  if ((int)w.center.size() != nprim || (int)w.type.size() != nprim || (size_t)nmo * nprim != w.coef.size())
    throw Error(CT_E_IO, "primitive / orbital counts disagree in " + path);
  FchkData d;
  d.atom_xyz = w.atom_xyz;
  d.charges = w.charges;
  d.numbers = w.numbers;
  const int natom = (int)d.numbers.size();
  HostBasis& b = d.basis;
  std::vector<int> func(nprim);  // basis function of each primitive
  std::vector<double> scale(nprim);
  int M = 0, cur_l = -1;
  std::vector<char> used;
  for (int p = 0; p < nprim; ++p) {
    const int t = w.type[p];
    if (t < 1 || t > 35) throw Error(CT_E_UNSUPPORTED, "primitive type " + std::to_string(t) + " in " + path);
    if (w.center[p] < 0 || w.center[p] >= natom) throw Error(CT_E_IO, "primitive centre out of range in " + path);
    const int a = kTypeExp[t][0], bb = kTypeExp[t][1], c = kTypeExp[t][2];
    const int L = a + bb + c;
    const int comp = cart_index(L, a, bb, c);
    const bool same = cur_l == L && b.shell_atom.back() == w.center[p] && b.exps.back() == w.exps[p] && !used[comp];
    if (!same) {
      b.shell_atom.push_back(w.center[p]);
      b.shell_l.push_back(L);
      b.shell_nprim.push_back(1);
      b.shell_prim_off.push_back((int32_t)b.exps.size());
      b.exps.push_back(w.exps[p]);
      b.coefs.push_back(1.0);
      used.assign(cart_order(L).size(), 0);
      cur_l = L;
      M += (int)used.size();
    }
    used[comp] = 1;
    func[p] = M - (int)used.size() + comp;
    scale[p] = 1.0 / cart_norm(w.exps[p], a, bb, c);
  }
  b.nbasis = M;
  d.P.assign((size_t)M * M, 0.0);
  std::vector<double> cj(M);
  for (int j = 0; j < nmo; ++j) {
    std::fill(cj.begin(), cj.end(), 0.0);
    for (int p = 0; p < nprim; ++p) cj[func[p]] += w.coef[(size_t)j * nprim + p] * scale[p];
    accumulate_density(d.P, M, cj.data(), w.occ[j]);
  }
  return d;
}

// Integer fields of "CENTRE ASSIGNMENTS" / "TYPE ASSIGNMENTS" lines: 20I3
// after a 20-character label (fixed width, so three-digit values may touch).
void fixed_ints(const std::string& line, std::vector<int>& out) {
  for (size_t pos = 20; pos < line.size(); pos += 3) {
    const std::string f = trim(line.substr(pos, 3));
    if (!f.empty()) out.push_back(std::atoi(f.c_str()));
  }
}

// After "KEY =" on a line: the number that follows.
double value_after(const std::string& line, const std::string& key, const std::string& where) {
  const size_t k = line.find(key);
  if (k == std::string::npos) throw Error(CT_E_IO, "'" + key + "' missing in " + where);
  const size_t eq = line.find('=', k + key.size());
  std::istringstream is(line.substr(eq == std::string::npos ? k + key.size() : eq + 1));
  std::string tok;
  is >> tok;
  return number(tok, where);
}

}  // namespace

FchkData load_wfn(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error(CT_E_IO, "cannot open wfn file: " + path);
  std::string line;
  std::getline(f, line);  // title
  std::getline(f, line);  // GAUSSIAN  nmo MOL ORBITALS  nprim PRIMITIVES  nnuc NUCLEI
  int nmo = 0, nprim = 0, nnuc = 0;
  {
    const size_t a = line.find("MOL ORBITALS"), b = line.find("PRIMITIVES"), c = line.find("NUCLEI");
    if (a == std::string::npos || b == std::string::npos || c == std::string::npos)
      throw Error(CT_E_IO, "not a wfn header line in " + path + ": " + line);
    if (trim(line).rfind("GAUSSIAN", 0) != 0)
      throw Error(CT_E_UNSUPPORTED, "only Gaussian-type wfn files are supported: " + path);
    auto last_int = [&](size_t from, size_t to) {
      std::istringstream is(line.substr(from, to - from));
      std::string tok, lastnum;
      while (is >> tok)
        if (std::isdigit((unsigned char)tok[0])) lastnum = tok;
      return std::atoi(lastnum.c_str());
    };
    nmo = last_int(0, a);
    nprim = last_int(a, b);
    nnuc = last_int(b, c);
  }
  if (nmo <= 0 || nprim <= 0 || nnuc <= 0) throw Error(CT_E_IO, "bad counts in the wfn header of " + path);
  PrimitiveWfn w;
  for (int i = 0; i < nnuc; ++i) {
    if (!std::getline(f, line)) throw Error(CT_E_IO, "truncated atom list in " + path);
    // "  O    1    (CENTRE  1)   0.00000000  0.00000000  0.22143053  CHARGE =  8.0"
    const size_t close = line.find(')'), ch = line.find("CHARGE");
    if (close == std::string::npos || ch == std::string::npos) throw Error(CT_E_IO, "bad atom line in " + path);
    const auto xyz = numbers(line.substr(close + 1, ch - close - 1), path);
    if (xyz.size() != 3) throw Error(CT_E_IO, "bad atom coordinates in " + path);
    w.atom_xyz.insert(w.atom_xyz.end(), xyz.begin(), xyz.end());
    const double q = value_after(line, "CHARGE", path);
    w.charges.push_back(q);
    std::istringstream is(line);
    std::string label;
    is >> label;
    int z = z_from_label(label);
    w.numbers.push_back(z > 0 ? z : (int32_t)std::lround(q));
  }
  std::string exps_text;
  while ((int)w.center.size() < nprim || (int)w.type.size() < nprim || (int)w.exps.size() < nprim) {
    if (!std::getline(f, line)) throw Error(CT_E_IO, "truncated primitive section in " + path);
    if (line.rfind("CENTRE ASSIGNMENTS", 0) == 0) {
      fixed_ints(line, w.center);
    } else if (line.rfind("TYPE ASSIGNMENTS", 0) == 0) {
      fixed_ints(line, w.type);
    } else if (line.rfind("EXPONENTS", 0) == 0) {
      const auto v = numbers(line.substr(9), path);
      w.exps.insert(w.exps.end(), v.begin(), v.end());
    } else if (!trim(line).empty()) {
      throw Error(CT_E_IO, "unexpected line in the primitive section of " + path + ": " + line);
    }
  }
  if ((int)w.center.size() != nprim || (int)w.type.size() != nprim || (int)w.exps.size() != nprim)
    throw Error(CT_E_IO, "primitive section length mismatch in " + path);
  for (int& c : w.center) c -= 1;
  std::vector<double> cur;
  while (std::getline(f, line)) {
    if (line.rfind("END DATA", 0) == 0) break;
    if (line.rfind("MO", 0) == 0 && line.find("OCC NO") != std::string::npos) {
      if ((int)w.occ.size() > 0 && (int)cur.size() != nprim) throw Error(CT_E_IO, "short orbital in " + path);
      w.coef.insert(w.coef.end(), cur.begin(), cur.end());
      cur.clear();
      w.occ.push_back(value_after(line, "OCC NO", path));
    } else if (!trim(line).empty()) {
      const auto v = numbers(line, path);
      cur.insert(cur.end(), v.begin(), v.end());
    }
  }
  if ((int)cur.size() != nprim) throw Error(CT_E_IO, "short last orbital in " + path);
  w.coef.insert(w.coef.end(), cur.begin(), cur.end());
  if ((int)w.occ.size() != nmo) throw Error(CT_E_IO, "orbital count mismatch in " + path);
  return from_primitives(w, path);
}

FchkData load_wfx(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error(CT_E_IO, "cannot open wfx file: " + path);
  // Sections <Name> ... </Name>; the orbital coefficients section holds one
  // "<MO Number> k </MO Number>" line before each orbital's values.
  std::map<std::string, std::string> sec;
  std::vector<std::vector<double>> mos;
  std::string line, open;
  std::string body;
  while (std::getline(f, line)) {
    const std::string t = trim(line);
    if (t.empty() || t[0] == '#') continue;
    if (open.empty()) {
      if (t.size() > 2 && t[0] == '<' && t[1] != '/') {
        const size_t e = t.find('>');
        open = t.substr(1, e - 1);
        body = t.substr(e + 1);
        const std::string close = "</" + open + ">";
        const size_t c = body.find(close);
        if (c != std::string::npos) {  // one-line section
          sec[open] = body.substr(0, c);
          open.clear();
        }
      }
      continue;
    }
    if (t == "</" + open + ">") {
      sec[open] = body;
      open.clear();
      continue;
    }
    if (open == "Molecular Orbital Primitive Coefficients") {
      if (t.rfind("<MO Number>", 0) == 0) {
        mos.emplace_back();
        continue;
      }
      if (mos.empty()) throw Error(CT_E_IO, "coefficients before <MO Number> in " + path);
      const auto v = numbers(t, path);
      mos.back().insert(mos.back().end(), v.begin(), v.end());
      continue;
    }
    body += " " + t;
  }
  auto need = [&](const std::string& k) -> const std::string& {
    auto it = sec.find(k);
    if (it == sec.end()) throw Error(CT_E_IO, "wfx section <" + k + "> missing in " + path);
    return it->second;
  };
  for (const char* k : {"Additional Electron Density Function (EDF)", "Number of EDF Primitives"})
    if (sec.count(k)) throw Error(CT_E_UNSUPPORTED, "wfx core-density (EDF) sections are not supported: " + path);
  PrimitiveWfn w;
  const int nnuc = (int)number(need("Number of Nuclei"), path);
  const int nprim = (int)number(need("Number of Primitives"), path);
  w.atom_xyz = numbers(need("Nuclear Cartesian Coordinates"), path);
  w.charges = numbers(need("Nuclear Charges"), path);
  for (double z : numbers(need("Atomic Numbers"), path)) w.numbers.push_back((int32_t)std::lround(z));
  for (double c : numbers(need("Primitive Centers"), path)) w.center.push_back((int)std::lround(c) - 1);
  for (double t : numbers(need("Primitive Types"), path)) w.type.push_back((int)std::lround(t));
  w.exps = numbers(need("Primitive Exponents"), path);
  w.occ = numbers(need("Molecular Orbital Occupation Numbers"), path);
  if ((int)w.atom_xyz.size() != 3 * nnuc || (int)w.charges.size() != nnuc || (int)w.numbers.size() != nnuc)
    throw Error(CT_E_IO, "nucleus count mismatch in " + path);
  if ((int)w.exps.size() != nprim) throw Error(CT_E_IO, "primitive count mismatch in " + path);
  if (mos.size() != w.occ.size()) throw Error(CT_E_IO, "orbital count mismatch in " + path);
  for (const auto& m : mos) {
    if ((int)m.size() != nprim) throw Error(CT_E_IO, "orbital length mismatch in " + path);
    w.coef.insert(w.coef.end(), m.begin(), m.end());
  }
  return from_primitives(w, path);
}

// ------------------------------------------------------------------ Molden

namespace {

// Molden's Cartesian g order (the d and f orders equal the reference's).
constexpr int kMoldenG[15][3] = {{4, 0, 0}, {0, 4, 0}, {0, 0, 4}, {3, 1, 0}, {3, 0, 1}, {1, 3, 0}, {0, 3, 1}, {1, 0, 3},
                                 {0, 1, 3}, {2, 2, 0}, {2, 0, 2}, {0, 2, 2}, {2, 1, 1}, {1, 2, 1}, {1, 1, 2}};

int shell_l_of(const std::string& s, const std::string& path) {
  const std::string t = lower(s);
  if (t == "s") return 0;
  if (t == "p") return 1;
  if (t == "d") return 2;
  if (t == "f") return 3;
  if (t == "g") return 4;
  throw Error(CT_E_UNSUPPORTED, "Molden shell type '" + s + "' in " + path);
}

}  // namespace

FchkData load_molden(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error(CT_E_IO, "cannot open Molden file: " + path);
  std::vector<std::string> lines;
  for (std::string line; std::getline(f, line);) lines.push_back(line);

  FchkData d;
  struct RawShell {
    int atom, L;  // L >= 0; purity decided by the flags
    std::vector<double> exps, coefs;
  };
  std::vector<RawShell> raw;
  bool pure_d = false, pure_f = false, pure_g = false, have_atoms = false, have_gto = false;
  std::vector<std::vector<double>> mo;  // coefficient lists (index -> value, 1-based)
  std::vector<double> occ;

  size_t i = 0;
  auto is_section = [&](size_t k) { return trim(lines[k]).rfind('[', 0) == 0; };
  while (i < lines.size()) {
    const std::string t = trim(lines[i]);
    if (t.empty() || t[0] != '[') {
      ++i;
      continue;
    }
    const size_t e = t.find(']');
    const std::string name = lower(t.substr(1, e - 1));
    const std::string rest = lower(trim(t.substr(e + 1)));
    ++i;
    if (name == "atoms") {
      const double unit = rest.find("ang") != std::string::npos ? kAngstrom : 1.0;
      for (; i < lines.size() && !is_section(i); ++i) {
        std::istringstream is(lines[i]);
        std::string label;
        int idx = 0, z = 0;
        double x, y, zz;
        if (!(is >> label)) continue;
        if (!(is >> idx >> z >> x >> y >> zz)) throw Error(CT_E_IO, "bad [Atoms] line in " + path + ": " + lines[i]);
        d.numbers.push_back(z);
        d.charges.push_back((double)z);
        d.atom_xyz.insert(d.atom_xyz.end(), {x * unit, y * unit, zz * unit});
      }
      have_atoms = true;
    } else if (name == "gto") {
      have_gto = true;
      while (i < lines.size() && !is_section(i)) {
        std::istringstream hdr(lines[i]);
        int atom = 0;
        if (!(hdr >> atom)) {  // blank line between atoms
          ++i;
          continue;
        }
        ++i;
        while (i < lines.size() && !is_section(i) && !trim(lines[i]).empty()) {
          if (std::isdigit((unsigned char)trim(lines[i])[0])) break;  // next atom without a blank line
          std::istringstream sh(lines[i]);
          std::string type;
          int np = 0;
          sh >> type >> np;
          if (np <= 0) throw Error(CT_E_IO, "bad [GTO] shell line in " + path + ": " + lines[i]);
          ++i;
          const bool sp = lower(type) == "sp";
          RawShell s0{atom - 1, sp ? 0 : shell_l_of(type, path), {}, {}};
          RawShell s1{atom - 1, 1, {}, {}};
          for (int k = 0; k < np; ++k, ++i) {
            if (i >= lines.size()) throw Error(CT_E_IO, "truncated [GTO] shell in " + path);
            const auto v = numbers(lines[i], path);
            if (v.size() < (sp ? 3u : 2u)) throw Error(CT_E_IO, "bad primitive line in " + path + ": " + lines[i]);
            s0.exps.push_back(v[0]);
            s0.coefs.push_back(v[1]);
            if (sp) {
              s1.exps.push_back(v[0]);
              s1.coefs.push_back(v[2]);
            }
          }
          raw.push_back(s0);
          if (sp) raw.push_back(s1);
        }
      }
    } else if (name == "5d" || name == "5d7f") {
      pure_d = pure_f = true;
    } else if (name == "5d10f") {
      pure_d = true;
    } else if (name == "7f") {
      pure_f = true;
    } else if (name == "9g") {
      pure_g = true;
    } else if (name == "sto") {
      throw Error(CT_E_UNSUPPORTED, "Slater-type Molden basis in " + path);
    } else if (name == "mo") {
      while (i < lines.size() && !is_section(i)) {
        const std::string l = trim(lines[i]);
        if (l.empty()) {
          ++i;
          continue;
        }
        if (l.find('=') != std::string::npos) {  // Sym= / Ene= / Spin= / Occup= of a new orbital
          if (mo.empty() || !mo.back().empty()) {
            mo.emplace_back();
            occ.push_back(0.0);
          }
          if (lower(l).rfind("occup", 0) == 0) occ.back() = number(l.substr(l.find('=') + 1), path);
          ++i;
          continue;
        }
        if (mo.empty()) throw Error(CT_E_IO, "[MO] coefficients before an orbital header in " + path);
        const auto v = numbers(l, path);
        if (v.size() != 2 || v[0] < 1) throw Error(CT_E_IO, "bad [MO] line in " + path + ": " + l);
        const size_t idx = (size_t)std::lround(v[0]);
        if (mo.back().size() < idx) mo.back().resize(idx, 0.0);
        mo.back()[idx - 1] = v[1];
        ++i;
      }
    }
  }
  if (!have_atoms || !have_gto || mo.empty()) throw Error(CT_E_IO, "[Atoms], [GTO] or [MO] missing in " + path);
  const int natom = (int)d.numbers.size();

  // Basis in the reference's conventions; perm[k]: the Molden function that
  // becomes our function k.  Contractions are renormalised so that each
  // contracted function has unit norm (Molden's MO coefficients refer to
  // normalised contracted functions; the overlap of two normalised
  // primitives with exponents a, b of the same component is
  // (2 sqrt(ab) / (a + b))^(L + 3/2), whatever the component).
  HostBasis& b = d.basis;
  std::vector<int> perm;
  int M = 0;
  for (const RawShell& s : raw) {
    if (s.atom < 0 || s.atom >= natom) throw Error(CT_E_IO, "[GTO] atom index out of range in " + path);
    const bool pure = (s.L == 2 && pure_d) || (s.L == 3 && pure_f) || (s.L == 4 && pure_g);
    const int l = pure ? -s.L : s.L;
    double norm2 = 0.0;
    for (size_t p = 0; p < s.exps.size(); ++p)
      for (size_t q = 0; q < s.exps.size(); ++q)
        norm2 += s.coefs[p] * s.coefs[q] *
                 std::pow(2.0 * std::sqrt(s.exps[p] * s.exps[q]) / (s.exps[p] + s.exps[q]), s.L + 1.5);
    if (!(norm2 > 0.0)) throw Error(CT_E_IO, "contracted shell of zero norm in " + path);
    const double sc = 1.0 / std::sqrt(norm2);
    b.shell_atom.push_back(s.atom);
    b.shell_l.push_back(l);
    b.shell_nprim.push_back((int32_t)s.exps.size());
    b.shell_prim_off.push_back((int32_t)b.exps.size());
    for (size_t p = 0; p < s.exps.size(); ++p) {
      b.exps.push_back(s.exps[p]);
      b.coefs.push_back(s.coefs[p] * sc);
    }
    const int nf = shell_nfunc(l);
    if (s.L == 4 && !pure) {
      // Molden's g order -> the reference's
      const auto ours = cart_order(4);
      for (int k = 0; k < nf; ++k) {
        int src = -1;
        for (int m = 0; m < 15; ++m)
          if (kMoldenG[m][0] == ours[k][0] && kMoldenG[m][1] == ours[k][1] && kMoldenG[m][2] == ours[k][2]) src = m;
        if (src < 0) throw Error(CT_E_INVALID, "Molden g component table incomplete");
        perm.push_back(M + src);
      }
    } else {
      for (int k = 0; k < nf; ++k) perm.push_back(M + k);  // s, p, Cartesian d/f and pure orders coincide
    }
    M += nf;
  }
  b.nbasis = M;
  d.P.assign((size_t)M * M, 0.0);
  std::vector<double> cj(M);
  for (size_t j = 0; j < mo.size(); ++j) {
    if ((int)mo[j].size() > M) throw Error(CT_E_IO, "[MO] index beyond the basis in " + path);
    for (int k = 0; k < M; ++k) cj[k] = perm[k] < (int)mo[j].size() ? mo[j][perm[k]] : 0.0;
    accumulate_density(d.P, M, cj.data(), occ[j]);
  }
  return d;
}

FchkData load_wavefunction(const std::string& path) {
  const std::string p = lower(path);
  auto ends = [&](const char* suf) {
    const size_t n = std::strlen(suf);
    return p.size() >= n && p.compare(p.size() - n, n, suf) == 0;
  };
  if (ends(".fchk") || ends(".fch")) return load_fchk(path);
  if (ends(".wfn")) return load_wfn(path);
  if (ends(".wfx")) return load_wfx(path);
  if (ends(".molden") || ends(".molden.input") || ends(".mold")) return load_molden(path);
  throw Error(CT_E_UNSUPPORTED, "unknown wavefunction file type (expected .fchk, .wfn, .wfx or .molden): " + path);
}

}  // namespace ct
