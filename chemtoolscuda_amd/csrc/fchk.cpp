// Native reader for Gaussian formatted checkpoint (.fchk) files.
//
// Replaces the reference's embedded-Python loader
// (src/iodata.cpp:15-211, which imports the `iodata` package and builds P with
// numpy).  Same conventions:
//   * shell types 0 s, 1 p, -1 sp, 2/3/4 Cartesian d/f/g, -2/-3/-4 pure; an
//     SP shell becomes an s and a p shell sharing exponents
//     (src/iodata.cpp:57-133);
//   * coordinates in bohr, nuclear (core) charges from "Nuclear charges"
//     (atcorenums, src/iodata.cpp:167-169);
//   * P = sum over spin channels of C occ C^T with aufbau occupations from the
//     alpha/beta electron counts (src/iodata.cpp:139-160); restricted files
//     use the alpha orbitals for both channels.

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

#include "mol.h"

namespace ct {

namespace {

struct Field {
  char kind = 0;  // 'I', 'R', 'C'
  bool is_array = false;
  std::vector<double> values;
};

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

std::map<std::string, Field> parse(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw Error(CT_E_IO, "cannot open fchk file: " + path);
  std::map<std::string, Field> out;
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    if (lineno <= 2) continue;  // title and job lines
    if (line.size() < 44 || line[0] == ' ') continue;
    const std::string name = trim(line.substr(0, 40));
    const char kind = line[43];
    if (kind != 'I' && kind != 'R' && kind != 'C' && kind != 'L') continue;
    const std::string rest = line.size() > 44 ? line.substr(44) : "";
    Field fld;
    fld.kind = kind;
    const size_t neq = rest.find("N=");
    if (neq == std::string::npos) {
      if (kind == 'I' || kind == 'R') fld.values.push_back(std::strtod(trim(rest).c_str(), nullptr));
      out[name] = fld;
      continue;
    }
    fld.is_array = true;
    const long count = std::strtol(rest.c_str() + neq + 2, nullptr, 10);
    if (kind == 'C' || kind == 'L') {
      const long per_line = kind == 'C' ? 5 : 72;
      long nlines = (count + per_line - 1) / per_line;
      while (nlines-- > 0 && std::getline(f, line)) ++lineno;
      out[name] = fld;
      continue;
    }
    fld.values.reserve(count);
    while ((long)fld.values.size() < count && std::getline(f, line)) {
      ++lineno;
      std::istringstream is(line);
      std::string tok;
      while (is >> tok) fld.values.push_back(std::strtod(tok.c_str(), nullptr));
    }
    if ((long)fld.values.size() != count)
      throw Error(CT_E_IO, "truncated array '" + name + "' in " + path);
    out[name] = fld;
  }
  return out;
}

const Field& need(const std::map<std::string, Field>& m, const std::string& key, const std::string& path) {
  auto it = m.find(key);
  if (it == m.end() || it->second.values.empty())
    throw Error(CT_E_IO, "fchk field '" + key + "' missing or empty in " + path);
  return it->second;
}

// an integer-valued entry of a field, range-checked
int as_int(double v, int lo, int hi, const std::string& what, const std::string& path) {
  if (!(v >= lo && v <= hi) || v != (double)(long)v)
    throw Error(CT_E_IO, "fchk: " + what + " out of range (" + std::to_string(v) + ") in " + path);
  return (int)v;
}

}  // namespace

FchkData load_fchk(const std::string& path) {
  const auto m = parse(path);
  FchkData d;
  const int nbasis = as_int(need(m, "Number of basis functions", path).values[0], 1, 1 << 20, "basis size", path);
  const int na = as_int(need(m, "Number of alpha electrons", path).values[0], 0, 1 << 20, "electron count", path);
  const int nb = as_int(need(m, "Number of beta electrons", path).values[0], 0, 1 << 20, "electron count", path);
  const auto& nums = need(m, "Atomic numbers", path).values;
  const int natom = (int)nums.size();
  d.atom_xyz = need(m, "Current cartesian coordinates", path).values;
  if ((int)d.atom_xyz.size() != 3 * natom) throw Error(CT_E_IO, "coordinate count mismatch in " + path);
  d.numbers.resize(natom);
  for (int a = 0; a < natom; ++a) d.numbers[a] = (int32_t)as_int(nums[a], 0, 200, "atomic number", path);
  d.charges = need(m, "Nuclear charges", path).values;
  if ((int)d.charges.size() != natom) throw Error(CT_E_IO, "nuclear charge count mismatch in " + path);

  const auto& types = need(m, "Shell types", path).values;
  const auto& nprims = need(m, "Number of primitives per shell", path).values;
  const auto& s2a = need(m, "Shell to atom map", path).values;
  const auto& exps = need(m, "Primitive exponents", path).values;
  const auto& cc = need(m, "Contraction coefficients", path).values;
  const std::vector<double>* spc = nullptr;
  auto it = m.find("P(S=P) Contraction coefficients");
  if (it != m.end()) spc = &it->second.values;
  if (nprims.size() != types.size() || s2a.size() != types.size())
    throw Error(CT_E_IO, "shell array length mismatch in " + path);
  if (cc.size() != exps.size() || (spc && spc->size() != exps.size()))
    throw Error(CT_E_IO, "contraction coefficient count mismatch in " + path);
  for (double a : exps)
    if (!(a > 0.0) || !std::isfinite(a)) throw Error(CT_E_IO, "non-positive primitive exponent in " + path);

  HostBasis& b = d.basis;
  size_t off = 0;
  for (size_t s = 0; s < types.size(); ++s) {
    const int t = as_int(types[s], -1000, 1000, "shell type", path);
    const int np = as_int(nprims[s], 1, 1 << 20, "primitives per shell", path);
    const int atom = as_int(s2a[s], 1, natom, "shell centre", path) - 1;
    if (off + np > exps.size()) throw Error(CT_E_IO, "primitive count mismatch in " + path);
    auto add = [&](int l, const std::vector<double>& coef) {
      b.shell_atom.push_back(atom);
      b.shell_l.push_back(l);
      b.shell_nprim.push_back(np);
      b.shell_prim_off.push_back((int32_t)b.exps.size());
      for (int k = 0; k < np; ++k) {
        b.exps.push_back(exps[off + k]);
        b.coefs.push_back(coef[off + k]);
      }
    };
    if (t == -1) {
      if (!spc) throw Error(CT_E_IO, "SP shell without P(S=P) coefficients in " + path);
      add(0, cc);
      add(1, *spc);
    } else {
      if (t < -4 || t > 4) throw Error(CT_E_UNSUPPORTED, "shell type " + std::to_string(t) + " unsupported");
      add(t, cc);
    }
    off += np;
  }
  int M = 0;
  for (int l : b.shell_l) M += shell_nfunc(l);
  if (M != nbasis) throw Error(CT_E_IO, "basis function count mismatch in " + path);
  b.nbasis = M;

  const auto& ca = need(m, "Alpha MO coefficients", path).values;
  const int nmo = (int)(ca.size() / nbasis);
  if ((size_t)nmo * nbasis != ca.size() || nmo < std::max(na, nb))
    throw Error(CT_E_IO, "MO coefficient count mismatch in " + path);
  const std::vector<double>* cb = &ca;
  auto itb = m.find("Beta MO coefficients");
  const bool unrestricted = itb != m.end();
  if (unrestricted) {
    cb = &itb->second.values;
    if (cb->size() % nbasis != 0 || (int)(cb->size() / nbasis) < nb)
      throw Error(CT_E_IO, "beta MO coefficient count mismatch in " + path);
  }
  d.P.assign((size_t)nbasis * nbasis, 0.0);
  // MO k occupies ca[k*nbasis .. (k+1)*nbasis)
  auto accumulate = [&](const std::vector<double>& C, int nocc, double occ) {
    for (int k = 0; k < nocc; ++k) {
      const double* c = &C[(size_t)k * nbasis];
      for (int i = 0; i < nbasis; ++i) {
        const double ci = occ * c[i];
        double* row = &d.P[(size_t)i * nbasis];
        for (int j = 0; j < nbasis; ++j) row[j] += ci * c[j];
      }
    }
  };
  accumulate(ca, na, 1.0);
  accumulate(*cb, nb, 1.0);
  return d;
}

}  // namespace ct
