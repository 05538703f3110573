// Readers of the reference's input formats (host only, no GPU):
//  - the input file molecules/<name>.in: "key = value" settings up to a line
//    "system:", then the atom count and one "Element x y z" line per atom
//    (Settings::read_settings_file / set_default_settings, src/settings.cpp:39-187;
//    Molecule::read_molecule_from_file, src/molecule.cpp:61-126);
//  - the basis-set file basis/<name>.dat (Molecule::set_basis_set,
//    src/molecule.cpp:135-238), expanded into the shell list of dg_basis in the
//    reference's CGF order: elements in the order of the file (ascending Z),
//    atoms of that element in input order, shells in file order.
// The line splitting follows the reference's boost::split calls: settings lines
// on '=' with adjacent separators merged, atom and basis lines on blanks and
// tabs with adjacent separators merged (a leading separator gives an empty
// first field, which is why the basis files' primitive lines, indented by a
// tab, carry exponent and coefficient in fields 1 and 2).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/dftgrid.h"
#include "grid_tables.h"

struct dg_input {
  std::map<std::string, std::string> kv;  // settings, first occurrence of a key kept
  std::string name, basis_name, units;
  int n_radial = 0, n_angular = 0, lmax = 0, lebedev_index = 0, hartree = DG_HARTREE_BECKE;
  std::vector<int> Z;
  std::vector<double> xyz;
  std::vector<int> sh_atom, sh_L, sh_nprim, sh_off;
  std::vector<double> alpha, coef;
};

namespace {

int fail(int code, const std::string& msg) {
  dg::set_last_error(msg);
  return code;
}

// The reference's Lebedev enum (src/settings.h:63-77) as point counts.
const int kLebedevPoints[11] = {6, 14, 26, 38, 50, 74, 86, 110, 146, 170, 194};
const char* kElements[18] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
                             "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"};

// boost::split(pieces, s, is_any_of(seps), token_compress_on): fields between
// separators, a run of separators counting as one; a leading or trailing run
// gives an empty first or last field.
std::vector<std::string> split_compress(const std::string& s, const char* seps) {
  std::vector<std::string> out;
  std::string cur;
  bool in_sep = false;
  for (char ch : s) {
    const bool sep = std::strchr(seps, ch) != nullptr && ch != '\0';
    if (sep) {
      if (!in_sep) {
        out.push_back(cur);
        cur.clear();
      }
      in_sep = true;
    } else {
      cur.push_back(ch);
      in_sep = false;
    }
  }
  out.push_back(cur);
  return out;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

// Whole-string conversions, as boost::lexical_cast (no surrounding blanks).
bool to_uint(const std::string& s, unsigned long& v) {
  if (s.empty() || !std::isdigit((unsigned char)s[0])) return false;
  errno = 0;
  char* end = nullptr;
  v = std::strtoul(s.c_str(), &end, 10);
  return errno == 0 && end && *end == '\0' && v <= UINT_MAX;
}

bool to_double(const std::string& s, double& v) {
  if (s.empty() || std::isspace((unsigned char)s[0])) return false;
  errno = 0;
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return errno == 0 && end && *end == '\0' && std::isfinite(v);
}

bool is_system_line(const std::string& line) {
  // ^system:\s*$ (src/settings.cpp:46, 66)
  if (line.compare(0, 7, "system:") != 0) return false;
  for (size_t i = 7; i < line.size(); ++i)
    if (!std::isspace((unsigned char)line[i])) return false;
  return true;
}

int element_number(const std::string& el) {
  for (int i = 0; i < 18; ++i)
    if (el == kElements[i]) return i + 1;
  return 0;
}

// Settings::set_grid_fineness (src/settings.cpp:172-204)
bool grid_fineness(const std::string& g, int& nrad, int& leb_index, int& lmax) {
  if (g == "coarse") {
    nrad = 10, leb_index = 4, lmax = 5;
  } else if (g == "medium") {
    nrad = 15, leb_index = 7, lmax = 8;
  } else if (g == "fine") {
    nrad = 20, leb_index = 8, lmax = 10;
  } else if (g == "ultrafine") {
    nrad = 30, leb_index = 10, lmax = 11;
  } else {
    return false;
  }
  return true;
}

int read_settings(std::ifstream& in, dg_input* d) {
  std::string line;
  while (std::getline(in, line)) {
    const std::vector<std::string> p = split_compress(line, "=");
    if (p.size() == 2) d->kv.emplace(trim(p[0]), trim(p[1]));
    if (is_system_line(line)) break;
  }
  auto get = [&](const char* key, std::string& v) {
    auto it = d->kv.find(key);
    if (it == d->kv.end()) return false;
    v = it->second;
    return true;
  };
  if (!get("name", d->name)) return fail(DG_ERR_INVALID, "input: Could not find name");
  if (!get("basis", d->basis_name)) return fail(DG_ERR_INVALID, "input: Could not find basis");
  if (!get("units", d->units)) d->units = "";
  std::string he;
  if (get("hartree_evaluation", he) && he == "two_electron_integrals") d->hartree = DG_HARTREE_ERI;
  std::string g = "medium";
  (void)get("grid", g);
  if (!grid_fineness(g, d->n_radial, d->lebedev_index, d->lmax))
    return fail(DG_ERR_INVALID, "input: unknown grid fineness '" + g +
                                    "' (coarse, medium, fine or ultrafine; the reference leaves the grid unset)");
  // overrides; a value that does not convert is ignored (src/settings.cpp:136-152)
  std::string v;
  unsigned long u = 0;
  if (get("radial_points", v) && to_uint(v, u)) d->n_radial = (int)u;
  if (get("lebedev_order", v) && to_uint(v, u)) {
    if (u > 10)
      return fail(DG_ERR_INVALID, "input: lebedev_order " + v + " is not an index 0..10 of the reference's rules");
    d->lebedev_index = (int)u;
  }
  if (get("lmax", v) && to_uint(v, u)) d->lmax = (int)u;
  d->n_angular = kLebedevPoints[d->lebedev_index];
  return DG_OK;
}

int read_atoms(const char* path, dg_input* d) {
  std::ifstream in(path);
  if (!in) return fail(DG_ERR_INVALID, std::string("input: Cannot open ") + path + "!");
  const bool angstrom = d->units == "angstrom";
  const double a2b = 1.889725989;  // src/molecule.cpp:65
  std::string line;
  bool found = false;
  while (std::getline(in, line))
    if (line.compare(0, 7, "system:") == 0) {
      found = true;
      break;
    }
  if (!found || !std::getline(in, line)) return fail(DG_ERR_INVALID, "input: no 'system:' block with an atom count");
  unsigned long nat = 0;
  if (!to_uint(line, nat) || nat == 0 || nat > 100000)
    return fail(DG_ERR_INVALID, "input: bad atom count '" + line + "'");
  for (unsigned long i = 0; i < nat; ++i) {
    if (!std::getline(in, line)) return fail(DG_ERR_INVALID, "input: fewer atom lines than the atom count");
    const std::vector<std::string> p = split_compress(line, " \t");
    if (p.size() < 4) return fail(DG_ERR_INVALID, "input: atom line '" + line + "' has fewer than four fields");
    const int z = element_number(p[0]);
    if (z == 0) return fail(DG_ERR_INVALID, "input: Unknown element: " + p[0]);
    double x[3];
    for (int c = 0; c < 3; ++c)
      if (!to_double(p[1 + c], x[c])) return fail(DG_ERR_INVALID, "input: bad coordinate '" + p[1 + c] + "'");
    d->Z.push_back(z);
    for (int c = 0; c < 3; ++c) d->xyz.push_back(angstrom ? x[c] * a2b : x[c]);
  }
  return DG_OK;
}

struct Shell {
  int L;
  std::vector<double> a, c;
};

int read_basis(const std::string& path, dg_input* d) {
  std::ifstream in(path);
  if (!in) return fail(DG_ERR_INVALID, "input: Cannot open " + path + "!");
  int zmax = 0;
  for (int z : d->Z) zmax = std::max(zmax, z);
  std::vector<bool> seen(19, false);
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line[0] == '#') continue;
    std::vector<std::string> p = split_compress(line, " \t");
    unsigned long z = 0, nsh = 0;
    if (p.size() < 2 || !to_uint(p[0], z) || !to_uint(p[1], nsh) || z < 1 || z > 18)
      return fail(DG_ERR_INVALID, "input: bad element header '" + line + "' in " + path);
    std::vector<Shell> shells;
    for (unsigned long s = 0; s < nsh; ++s) {
      if (!std::getline(in, line)) return fail(DG_ERR_INVALID, "input: " + path + " ends inside an element");
      p = split_compress(line, " \t");
      unsigned long np = 0;
      if (p.size() < 2 || p[0].empty() || !to_uint(p[1], np) || np == 0)
        return fail(DG_ERR_INVALID, "input: bad shell header '" + line + "' in " + path);
      Shell sh;
      switch (p[0][0]) {
        case 'S': sh.L = 0; break;
        case 'P': sh.L = 1; break;
        case 'D': sh.L = 2; break;
        default: return fail(DG_ERR_INVALID, "input: shell type '" + p[0] + "' (S, P or D) in " + path);
      }
      for (unsigned long k = 0; k < np; ++k) {
        if (!std::getline(in, line)) return fail(DG_ERR_INVALID, "input: " + path + " ends inside a shell");
        p = split_compress(line, " \t");
        double a = 0.0, c = 0.0;
        if (p.size() < 3 || !to_double(p[1], a) || !to_double(p[2], c))
          return fail(DG_ERR_INVALID, "input: bad primitive line '" + line + "' in " + path);
        sh.a.push_back(a);
        sh.c.push_back(c);
      }
      shells.push_back(sh);
    }
    // the element's functions on every atom of that element, in input order
    for (size_t ia = 0; ia < d->Z.size(); ++ia) {
      if (d->Z[ia] != (int)z) continue;
      for (const Shell& sh : shells) {
        d->sh_atom.push_back((int)ia);
        d->sh_L.push_back(sh.L);
        d->sh_nprim.push_back((int)sh.a.size());
        d->sh_off.push_back((int)d->alpha.size());
        d->alpha.insert(d->alpha.end(), sh.a.begin(), sh.a.end());
        d->coef.insert(d->coef.end(), sh.c.begin(), sh.c.end());
      }
    }
    seen[z] = true;
    if ((int)z == zmax) break;  // the highest element of the molecule ends the read
  }
  for (int z : d->Z)
    if (!seen[z])
      return fail(DG_ERR_INVALID, std::string("input: ") + path + " has no functions for element " + kElements[z - 1]);
  return DG_OK;
}

}  // namespace

extern "C" {

int dg_input_read(const char* path, const char* basis_dir, dg_input** out) {
  if (!path || !basis_dir || !out) return fail(DG_ERR_INVALID, "dg_input_read: null argument");
  *out = nullptr;
  dg_input* d = new (std::nothrow) dg_input;
  if (!d) return fail(DG_ERR_ALLOC, "dg_input_read: out of host memory");
  int rc = DG_OK;
  try {
    std::ifstream in(path);
    if (!in) {
      rc = fail(DG_ERR_INVALID, std::string("input: Cannot open ") + path + "!");
    } else {
      rc = read_settings(in, d);
      if (rc == DG_OK) rc = read_atoms(path, d);
      if (rc == DG_OK) rc = read_basis(std::string(basis_dir) + "/" + d->basis_name + ".dat", d);
    }
  } catch (const std::exception& e) {
    rc = fail(DG_ERR_ALLOC, std::string("dg_input_read: ") + e.what());
  }
  if (rc != DG_OK) {
    delete d;
    return rc;
  }
  *out = d;
  return DG_OK;
}

int dg_input_destroy(dg_input* d) {
  delete d;
  return DG_OK;
}

int dg_input_molecule(const dg_input* d, dg_molecule* mol) {
  if (!d || !mol) return fail(DG_ERR_INVALID, "dg_input_molecule: null argument");
  mol->natoms = (int)d->Z.size();
  mol->Z = d->Z.data();
  mol->xyz = d->xyz.data();
  return DG_OK;
}

int dg_input_basis(const dg_input* d, dg_basis* basis) {
  if (!d || !basis) return fail(DG_ERR_INVALID, "dg_input_basis: null argument");
  basis->nshell = (int)d->sh_atom.size();
  basis->atom = d->sh_atom.data();
  basis->L = d->sh_L.data();
  basis->nprim = d->sh_nprim.data();
  basis->offset = d->sh_off.data();
  basis->alpha = d->alpha.data();
  basis->coef = d->coef.data();
  return DG_OK;
}

int dg_input_grid(const dg_input* d, dg_grid_params* grid, int* lebedev_index, int* hartree) {
  if (!d) return fail(DG_ERR_INVALID, "dg_input_grid: null input");
  if (grid) {
    grid->n_radial = d->n_radial;
    grid->n_angular = d->n_angular;
    grid->lmax = d->lmax;
    grid->precision = DG_PREC_FP64;
    grid->screen = 0.0;
  }
  if (lebedev_index) *lebedev_index = d->lebedev_index;
  if (hartree) *hartree = d->hartree;
  return DG_OK;
}

int dg_input_value(const dg_input* d, const char* key, const char** value) {
  if (!d || !key || !value) return fail(DG_ERR_INVALID, "dg_input_value: null argument");
  auto it = d->kv.find(key);
  if (it == d->kv.end()) {
    *value = nullptr;
    return fail(DG_ERR_INVALID, std::string("Could not find ") + key);
  }
  *value = it->second.c_str();
  return DG_OK;
}

}  // extern "C"
