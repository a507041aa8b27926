// LAMMPS data files (atom_style full, units real) for the systems this path evaluates — the on-disk input of a run and
// its checkpoint (core/src/io_lammps.cpp:58-297 reader, :505-583 writer; checkpoints as cli/main.cpp:7820-7846 writes
// them: the state with velocities, written whole to a temporary file and renamed). The force field travels in the
// coefficient sections: Pair Coeffs lj/cut (ε, σ per type, geometric mixing) or PairIJ Coeffs (every pair), Bond Coeffs
// harmonic (K, r0: E = K (r − r0)²) and fene (K, R0, ε, σ), Angle Coeffs harmonic (K, θ0 in degrees). Numbers are written
// with 17 significant digits, so a file read back gives the same doubles.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/capsgpu.h"
#include "../../include/capsgpu_build.h"
#include "system.hpp"

namespace capsgpu {
void set_last_error(const std::string& m);
}

namespace {

using capsgpu::exclusions;
using capsgpu::Sys;

constexpr double kPi = 3.14159265358979323846;

double radians(double deg) { return deg * kPi / 180.0; }

// θ0 in degrees for the file: of the doubles next to θ0·180/π, one the reader turns back into θ0 exactly, when there is one
double degrees(double rad) {
  const double d = rad * 180.0 / kPi;
  double up = d, down = d;
  for (int k = 0; k < 64; ++k) {
    if (radians(up) == rad) return up;
    if (radians(down) == rad) return down;
    up = std::nextafter(up, HUGE_VAL), down = std::nextafter(down, -HUGE_VAL);
  }
  return d;
}

std::string trim(const std::string& s) {
  const size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

std::vector<std::string> words(const std::string& s) {
  std::istringstream in(s);
  std::vector<std::string> w;
  for (std::string t; in >> t;) w.push_back(t);
  return w;
}

double num(const std::string& t, const std::string& where) {
  try {
    size_t used = 0;
    const double v = std::stod(t, &used);
    if (used != t.size()) throw std::invalid_argument(t);
    return v;
  } catch (const std::exception&) {
    throw std::invalid_argument(where + ": '" + t + "' is not a number");
  }
}

// the mixed ε, σ tables factorise as geometric means of their diagonals (LAMMPS pair_modify mix geometric)
bool geometric(int nt, const double* eps, const double* sig) {
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const double e = std::sqrt(eps[a * nt + a] * eps[b * nt + b]), s = std::sqrt(sig[a * nt + a] * sig[b * nt + b]);
      if (std::fabs(e - eps[a * nt + b]) > 1e-12 * std::max(1.0, std::fabs(e)) || std::fabs(s - sig[a * nt + b]) > 1e-12 * std::max(1.0, s))
        return false;
    }
  return true;
}

// the position a row of the Atoms section stands for: the wrapped position plus the image flags' cell vectors, added
// in this order by the reader and the writer alike
double unwrap(const double* w, const double* cell, const int* im, int k) {
  return w[k] + cell[3 + k] * im[0] + cell[6 + k] * im[1] + cell[9 + k] * im[2];
}

Sys* read_data(const std::string& path, int ex13) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument("cannot open " + path);
  static const char* kSections[] = {"Masses", "Atoms", "Velocities", "Bonds", "Angles", "Dihedrals", "Impropers", "Pair Coeffs",
                                    "PairIJ Coeffs", "Bond Coeffs", "Angle Coeffs", "Dihedral Coeffs", "Improper Coeffs"};
  std::string line, section, style;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, xy = 0, xz = 0, yz = 0;
  int64_t natoms = -1, ntypes = -1;
  std::map<int, double> masses;
  std::map<int, std::pair<double, double>> pair_ii;
  std::map<std::pair<int, int>, std::pair<double, double>> pair_ij;
  struct Coeff {
    std::string style;
    std::vector<double> p;
    double rad = -1;   // Angle Coeffs: θ0 in radians from a "# theta0 <rad> rad" comment, when it agrees with the degrees
  };
  std::map<int, Coeff> bond_coeff, angle_coeff;
  struct AtomRow {
    int type;
    double q, r[3];
    int im[3];
  };
  std::map<int64_t, AtomRow> atoms;
  std::map<int64_t, std::array<double, 3>> vel;
  std::vector<std::array<int64_t, 3>> bonds;      // type, i, j (ids)
  std::vector<std::array<int64_t, 4>> angles;     // type, i, j, k
  size_t lineno = 1;
  std::getline(in, line);   // the title
  while (std::getline(in, line)) {
    ++lineno;
    const std::string where = path + ":" + std::to_string(lineno);
    std::string comment;
    const size_t h = line.find('#');
    if (h != std::string::npos) comment = trim(line.substr(h + 1)), line = line.substr(0, h);
    const std::string body = trim(line);
    if (body.empty()) continue;
    bool header = false;
    for (const char* s : kSections)
      if (body == s) header = true, section = s, style = comment;
    if (header) continue;
    const auto t = words(body);
    if (section.empty()) {
      if (t.size() >= 2 && t[1] == "atoms") natoms = int64_t(num(t[0], where));
      else if (t.size() >= 3 && t[1] == "atom" && t[2] == "types") ntypes = int64_t(num(t[0], where));
      else if (t.size() >= 4 && t[2] == "xlo") lo[0] = num(t[0], where), hi[0] = num(t[1], where);
      else if (t.size() >= 4 && t[2] == "ylo") lo[1] = num(t[0], where), hi[1] = num(t[1], where);
      else if (t.size() >= 4 && t[2] == "zlo") lo[2] = num(t[0], where), hi[2] = num(t[1], where);
      else if (t.size() >= 6 && t[3] == "xy") xy = num(t[0], where), xz = num(t[1], where), yz = num(t[2], where);
      continue;
    }
    if (section == "Masses" && t.size() >= 2) {
      masses[int(num(t[0], where))] = num(t[1], where);
    } else if (section == "Pair Coeffs" && t.size() >= 3) {
      pair_ii[int(num(t[0], where))] = {num(t[1], where), num(t[2], where)};
    } else if (section == "PairIJ Coeffs" && t.size() >= 4) {
      pair_ij[{int(num(t[0], where)), int(num(t[1], where))}] = {num(t[2], where), num(t[3], where)};
    } else if ((section == "Bond Coeffs" || section == "Angle Coeffs") && t.size() >= 2) {
      Coeff c;
      size_t p = 1;
      c.style = style.empty() ? "harmonic" : words(style)[0];
      if (c.style == "hybrid") c.style = t[p++];
      for (; p < t.size(); ++p) c.p.push_back(num(t[p], where));
      const auto cw = words(comment);
      if (section == "Angle Coeffs" && c.p.size() >= 2 && cw.size() >= 3 && cw[0] == "theta0" && cw[2] == "rad") {
        const double r = num(cw[1], where);
        if (std::fabs(r - radians(c.p[1])) <= 1e-12 * std::max(1.0, std::fabs(r))) c.rad = r;
      }
      (section == "Bond Coeffs" ? bond_coeff : angle_coeff)[int(num(t[0], where))] = c;
    } else if (section == "Atoms") {
      if (t.size() != 7 && t.size() != 10) throw std::invalid_argument(where + ": atom_style full has 7 or 10 columns");
      AtomRow a{};
      a.type = int(num(t[2], where));
      a.q = num(t[3], where);
      for (int k = 0; k < 3; ++k) a.r[k] = num(t[4 + size_t(k)], where);
      for (int k = 0; k < 3; ++k) a.im[k] = t.size() == 10 ? int(num(t[7 + size_t(k)], where)) : 0;
      atoms[int64_t(num(t[0], where))] = a;
    } else if (section == "Velocities" && t.size() >= 4) {
      vel[int64_t(num(t[0], where))] = {num(t[1], where), num(t[2], where), num(t[3], where)};
    } else if (section == "Bonds" && t.size() >= 4) {
      bonds.push_back({int64_t(num(t[1], where)), int64_t(num(t[2], where)), int64_t(num(t[3], where))});
    } else if (section == "Angles" && t.size() >= 5) {
      angles.push_back({int64_t(num(t[1], where)), int64_t(num(t[2], where)), int64_t(num(t[3], where)), int64_t(num(t[4], where))});
    } else if (section == "Dihedrals" || section == "Impropers" || section == "Dihedral Coeffs" || section == "Improper Coeffs") {
      throw std::invalid_argument(where + ": dihedral and improper terms are outside this path");
    }
  }
  if (natoms >= 0 && natoms != int64_t(atoms.size()))
    throw std::invalid_argument(path + ": the header says " + std::to_string(natoms) + " atoms, the Atoms section has " + std::to_string(atoms.size()));
  if (atoms.empty()) throw std::invalid_argument(path + ": no atoms");
  auto S = std::make_unique<Sys>();
  const int n = int(atoms.size());
  int nt = int(ntypes);
  for (const auto& kv : atoms) nt = std::max(nt, kv.second.type);
  S->s.n_atoms = n, S->s.n_types = nt;
  // the cell (io_lammps.cpp:58-65)
  const double cell[12] = {lo[0], lo[1], lo[2], hi[0] - lo[0], 0, 0, xy, hi[1] - lo[1], 0, xz, yz, hi[2] - lo[2]};
  std::copy(cell, cell + 12, S->s.cell);
  std::map<int64_t, int32_t> index;
  for (const auto& kv : atoms) {
    const AtomRow& a = kv.second;
    if (a.type < 1 || a.type > nt) throw std::invalid_argument(path + ": atom " + std::to_string(kv.first) + " has a type out of range");
    index[kv.first] = int32_t(S->type.size());
    S->type.push_back(a.type - 1);
    S->charge.push_back(a.q);
    auto m = masses.find(a.type);
    if (m == masses.end() || !(m->second > 0)) throw std::invalid_argument(path + ": no mass for atom type " + std::to_string(a.type));
    S->mass.push_back(m->second);
    // image flags make the molecules whole (io_lammps.cpp:284-289)
    for (int k = 0; k < 3; ++k) S->x.push_back(unwrap(a.r, cell, a.im, k));
  }
  if (!vel.empty()) {
    S->v.assign(3 * size_t(n), 0.0);
    for (const auto& kv : vel) {
      auto it = index.find(kv.first);
      if (it != index.end())
        for (int k = 0; k < 3; ++k) S->v[3 * size_t(it->second) + size_t(k)] = kv.second[size_t(k)];
    }
  }
  // LJ: per-type ε, σ mixed geometrically, or every pair given
  S->eps.assign(size_t(nt) * nt, 0.0), S->sigma.assign(size_t(nt) * nt, 1.0);
  for (int a = 1; a <= nt; ++a)
    for (int b = 1; b <= nt; ++b) {
      double e = 0, s = 1;
      auto ij = pair_ij.find({std::min(a, b), std::max(a, b)});
      if (ij != pair_ij.end()) {
        e = ij->second.first, s = ij->second.second;
      } else {
        auto pa = pair_ii.find(a), pb = pair_ii.find(b);
        if (pa != pair_ii.end() && pb != pair_ii.end()) e = std::sqrt(pa->second.first * pb->second.first), s = std::sqrt(pa->second.second * pb->second.second);
      }
      S->eps[size_t(a - 1) * nt + size_t(b - 1)] = e, S->sigma[size_t(a - 1) * nt + size_t(b - 1)] = s;
    }
  auto id = [&](int64_t i) {
    auto it = index.find(i);
    if (it == index.end()) throw std::invalid_argument(path + ": a bonded term names atom " + std::to_string(i) + ", which is not defined");
    return it->second;
  };
  for (const auto& b : bonds) {
    auto c = bond_coeff.find(int(b[0]));
    if (c == bond_coeff.end()) throw std::invalid_argument(path + ": no coefficients for bond type " + std::to_string(b[0]));
    if (c->second.style == "harmonic" && c->second.p.size() >= 2) {
      S->bond_atoms.insert(S->bond_atoms.end(), {id(b[1]), id(b[2])});
      S->bond_param.insert(S->bond_param.end(), {c->second.p[0], c->second.p[1]});
    } else if (c->second.style == "fene" && c->second.p.size() >= 4) {
      S->fene_atoms.insert(S->fene_atoms.end(), {id(b[1]), id(b[2])});
      S->fene_param.insert(S->fene_param.end(), {c->second.p[0], c->second.p[1], c->second.p[2], c->second.p[3]});
    } else {
      throw std::invalid_argument(path + ": bond style '" + c->second.style + "' is outside this path (harmonic, fene)");
    }
  }
  for (const auto& a : angles) {
    auto c = angle_coeff.find(int(a[0]));
    if (c == angle_coeff.end() || c->second.style != "harmonic" || c->second.p.size() < 2)
      throw std::invalid_argument(path + ": angle type " + std::to_string(a[0]) + " needs harmonic coefficients");
    S->angle_atoms.insert(S->angle_atoms.end(), {id(a[1]), id(a[2]), id(a[3])});
    S->angle_param.insert(S->angle_param.end(), {c->second.p[0], c->second.rad >= 0 ? c->second.rad : radians(c->second.p[1])});
  }
  exclusions(*S, n, ex13 < 0 ? !angles.empty() : ex13 != 0);
  S->publish();
  return S.release();
}

std::string g17(double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// LAMMPS's restricted triclinic frame (io_lammps.cpp:505-553 writes positions in it): a along x, b in the xy plane, c
// with a positive z. A cell already in that form keeps its numbers (the rotation is the identity); another is rotated
// about its origin, positions and velocities with it.
struct Frame {
  double q[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double cell[12];
  bool identity = true;
  explicit Frame(const double* c) {
    std::copy(c, c + 12, cell);
    if (c[4] == 0 && c[5] == 0 && c[8] == 0 && c[3] > 0 && c[7] > 0 && c[11] > 0) return;
    auto dot = [](const double* u, const double* w) { return u[0] * w[0] + u[1] * w[1] + u[2] * w[2]; };
    const double *a = c + 3, *b = c + 6, *cc = c + 9;
    const double la = std::sqrt(dot(a, a));
    if (!(la > 0)) throw std::invalid_argument("the cell has a zero vector");
    double e1[3], e2[3], e3[3];
    for (int k = 0; k < 3; ++k) e1[k] = a[k] / la;
    const double bx = dot(b, e1);
    for (int k = 0; k < 3; ++k) e2[k] = b[k] - bx * e1[k];
    const double lb = std::sqrt(dot(e2, e2));
    if (!(lb > 0)) throw std::invalid_argument("the cell's a and b are parallel");
    for (int k = 0; k < 3; ++k) e2[k] /= lb;
    e3[0] = e1[1] * e2[2] - e1[2] * e2[1], e3[1] = e1[2] * e2[0] - e1[0] * e2[2], e3[2] = e1[0] * e2[1] - e1[1] * e2[0];
    if (!(dot(cc, e3) > 0)) throw std::invalid_argument("the cell is left-handed or flat; LAMMPS needs c with a positive z");
    for (int k = 0; k < 3; ++k) q[k] = e1[k], q[3 + k] = e2[k], q[6 + k] = e3[k];
    identity = false;
    for (int v = 0; v < 3; ++v) rotate(c + 3 + 3 * v, cell + 3 + 3 * v);
    cell[4] = cell[5] = cell[8] = 0;   // exactly, not to rounding
  }
  void rotate(const double* in, double* out) const {
    for (int r = 0; r < 3; ++r) out[r] = q[3 * r] * in[0] + q[3 * r + 1] * in[1] + q[3 * r + 2] * in[2];
  }
  void position(const double* x, const double* origin, double* out) const {
    if (identity) return (void)std::copy(x, x + 3, out);
    const double d[3] = {x[0] - origin[0], x[1] - origin[1], x[2] - origin[2]};
    rotate(d, out);
    for (int k = 0; k < 3; ++k) out[k] += origin[k];
  }
};

// the box's upper bound: of the doubles next to lo + length, one whose difference from lo is the length exactly (the
// reader takes hi - lo), when there is one
double upper(double lo, double length) {
  const double h = lo + length;
  double up = h, down = h;
  for (int k = 0; k < 64; ++k) {
    if (up - lo == length) return up;
    if (down - lo == length) return down;
    up = std::nextafter(up, HUGE_VAL), down = std::nextafter(down, -HUGE_VAL);
  }
  return h;
}

void write_data(const std::string& path, const capsgpu_system& s, const double* v, const std::string& title) {
  const int n = s.n_atoms, nt = s.n_types;
  const Frame frame(s.cell);
  const double* c = frame.cell;
  // molecules: connected components of the bonds
  std::vector<int> mol(size_t(n), -1);
  {
    std::vector<std::vector<int>> nb(static_cast<size_t>(n));
    for (int b = 0; b < s.n_bonds; ++b) nb[size_t(s.bond_atoms[2 * b])].push_back(s.bond_atoms[2 * b + 1]), nb[size_t(s.bond_atoms[2 * b + 1])].push_back(s.bond_atoms[2 * b]);
    for (int b = 0; b < s.n_fene; ++b) nb[size_t(s.fene_atoms[2 * b])].push_back(s.fene_atoms[2 * b + 1]), nb[size_t(s.fene_atoms[2 * b + 1])].push_back(s.fene_atoms[2 * b]);
    int m = 0;
    for (int i = 0; i < n; ++i) {
      if (mol[size_t(i)] >= 0) continue;
      std::vector<int> stack = {i};
      mol[size_t(i)] = ++m;
      while (!stack.empty()) {
        const int a = stack.back();
        stack.pop_back();
        for (int b : nb[size_t(a)])
          if (mol[size_t(b)] < 0) mol[size_t(b)] = m, stack.push_back(b);
      }
    }
  }
  // one bond type per distinct parameter set, harmonic first, then FENE; one angle type per distinct (K, θ0)
  std::map<std::vector<double>, int> btype, atype;
  std::vector<std::pair<std::string, std::vector<double>>> bcoeff, acoeff;
  std::vector<int> bt(size_t(s.n_bonds + s.n_fene)), at(size_t(s.n_angles));
  for (int b = 0; b < s.n_bonds + s.n_fene; ++b) {
    const bool fene = b >= s.n_bonds;
    std::vector<double> key = fene ? std::vector<double>(s.fene_param + 4 * (b - s.n_bonds), s.fene_param + 4 * (b - s.n_bonds) + 4)
                                   : std::vector<double>(s.bond_param + 2 * b, s.bond_param + 2 * b + 2);
    key.insert(key.begin(), fene ? 1.0 : 0.0);
    auto it = btype.find(key);
    if (it == btype.end()) {
      it = btype.emplace(key, int(btype.size()) + 1).first;
      bcoeff.push_back({fene ? "fene" : "harmonic", std::vector<double>(key.begin() + 1, key.end())});
    }
    bt[size_t(b)] = it->second;
  }
  for (int a = 0; a < s.n_angles; ++a) {
    const std::vector<double> key = {s.angle_param[2 * a], s.angle_param[2 * a + 1]};
    auto it = atype.find(key);
    if (it == atype.end()) it = atype.emplace(key, int(atype.size()) + 1).first, acoeff.push_back({"harmonic", {key[0], degrees(key[1]), key[1]}});
    at[size_t(a)] = it->second;
  }
  const bool harm = s.n_bonds > 0, fen = s.n_fene > 0;
  const std::string tmp = path + ".tmp";
  {
    std::ofstream out(tmp);
    if (!out) throw std::invalid_argument("cannot write " + tmp);
    out << "CAPS-AMD · " << (title.empty() ? std::string("structure") : title) << " · atom_style full · units real\n\n";
    out << n << " atoms\n" << s.n_bonds + s.n_fene << " bonds\n" << s.n_angles << " angles\n\n";
    out << nt << " atom types\n" << bcoeff.size() << " bond types\n" << acoeff.size() << " angle types\n\n";
    out << g17(c[0]) << " " << g17(upper(c[0], c[3])) << " xlo xhi\n" << g17(c[1]) << " " << g17(upper(c[1], c[7])) << " ylo yhi\n"
        << g17(c[2]) << " " << g17(upper(c[2], c[11])) << " zlo zhi\n";
    if (c[6] != 0 || c[9] != 0 || c[10] != 0) out << g17(c[6]) << " " << g17(c[9]) << " " << g17(c[10]) << " xy xz yz\n";
    out << "\nMasses\n\n";
    for (int t = 0; t < nt; ++t) {
      double m = 0;
      for (int i = 0; i < n; ++i)
        if (s.type[i] == t) {
          m = s.mass[i];
          break;
        }
      out << t + 1 << " " << g17(m) << "\n";
    }
    if (geometric(nt, s.eps, s.sigma)) {
      out << "\nPair Coeffs  # lj/cut\n\n";
      for (int t = 0; t < nt; ++t) out << t + 1 << " " << g17(s.eps[t * nt + t]) << " " << g17(s.sigma[t * nt + t]) << "\n";
    } else {
      out << "\nPairIJ Coeffs  # lj/cut\n\n";
      for (int a = 0; a < nt; ++a)
        for (int b = a; b < nt; ++b) out << a + 1 << " " << b + 1 << " " << g17(s.eps[a * nt + b]) << " " << g17(s.sigma[a * nt + b]) << "\n";
    }
    if (!bcoeff.empty()) {
      out << "\nBond Coeffs  # " << (harm && fen ? "hybrid" : fen ? "fene" : "harmonic") << "\n\n";
      for (size_t k = 0; k < bcoeff.size(); ++k) {
        out << k + 1;
        if (harm && fen) out << " " << bcoeff[k].first;
        for (double p : bcoeff[k].second) out << " " << g17(p);
        out << "\n";
      }
    }
    if (!acoeff.empty()) {
      out << "\nAngle Coeffs  # harmonic\n\n";
      // θ0 in degrees as LAMMPS reads it; the radians it came from in a comment, since not every double in radians has a
      // double in degrees that converts back to it
      for (size_t k = 0; k < acoeff.size(); ++k)
        out << k + 1 << " " << g17(acoeff[k].second[0]) << " " << g17(acoeff[k].second[1]) << "  # theta0 " << g17(acoeff[k].second[2]) << " rad\n";
    }
    // positions wrapped into the cell with image flags (io_lammps.cpp:561-570). An atom whose wrapped position would not
    // give its own position back to the last bit is written unwrapped with zero flags (LAMMPS maps it into the box), so a
    // file read back holds the same doubles.
    out << "\nAtoms  # full\n\n";
    const double *A = c + 3, *B = c + 6, *Cc = c + 9;
    for (int i = 0; i < n; ++i) {
      double r[3];
      frame.position(s.x + 3 * i, c, r);
      const double d[3] = {r[0] - c[0], r[1] - c[1], r[2] - c[2]};
      // an upper-triangular cell gives the fractional coordinates by back substitution
      const double fz = d[2] / Cc[2];
      const double fy = (d[1] - fz * Cc[1]) / B[1];
      const double fx = (d[0] - fy * B[0] - fz * Cc[0]) / A[0];
      int im[3] = {int(std::floor(fx)), int(std::floor(fy)), int(std::floor(fz))};
      double w[3];
      for (int k = 0; k < 3; ++k) w[k] = r[k] - (A[k] * im[0] + B[k] * im[1] + Cc[k] * im[2]);
      bool exact = true;
      for (int k = 0; k < 3; ++k) exact = exact && unwrap(w, c, im, k) == r[k];
      if (!exact) std::copy(r, r + 3, w), im[0] = im[1] = im[2] = 0;
      out << i + 1 << " " << mol[size_t(i)] << " " << s.type[i] + 1 << " " << g17(s.charge[i]) << " " << g17(w[0]) << " " << g17(w[1]) << " "
          << g17(w[2]) << " " << im[0] << " " << im[1] << " " << im[2] << "\n";
    }
    if (v) {
      out << "\nVelocities\n\n";
      for (int i = 0; i < n; ++i) {
        double u[3];
        if (frame.identity) std::copy(v + 3 * i, v + 3 * i + 3, u);
        else frame.rotate(v + 3 * i, u);
        out << i + 1 << " " << g17(u[0]) << " " << g17(u[1]) << " " << g17(u[2]) << "\n";
      }
    }
    if (s.n_bonds + s.n_fene > 0) {
      out << "\nBonds\n\n";
      for (int b = 0; b < s.n_bonds + s.n_fene; ++b) {
        const int32_t* p = b < s.n_bonds ? s.bond_atoms + 2 * b : s.fene_atoms + 2 * (b - s.n_bonds);
        out << b + 1 << " " << bt[size_t(b)] << " " << p[0] + 1 << " " << p[1] + 1 << "\n";
      }
    }
    if (s.n_angles > 0) {
      out << "\nAngles\n\n";
      for (int a = 0; a < s.n_angles; ++a)
        out << a + 1 << " " << at[size_t(a)] << " " << s.angle_atoms[3 * a] + 1 << " " << s.angle_atoms[3 * a + 1] + 1 << " " << s.angle_atoms[3 * a + 2] + 1 << "\n";
    }
    if (!out) throw std::invalid_argument("writing " + tmp + " failed");
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::invalid_argument("cannot rename " + tmp + " to " + path);
}

}  // namespace

extern "C" {

int32_t capsgpu_read_lammps(const char* path, int32_t exclude13, capsgpu_system** out) {
  try {
    if (!path || !out) throw std::invalid_argument("null argument");
    *out = &read_data(path, exclude13)->s;
    return CAPSGPU_OK;
  } catch (const std::exception& e) {
    capsgpu::set_last_error(e.what());
    return CAPSGPU_EINVAL;
  }
}

int32_t capsgpu_write_lammps(const char* path, const capsgpu_system* s, const double* v, const char* title) {
  try {
    if (!path || !s) throw std::invalid_argument("null argument");
    write_data(path, *s, v, title ? title : "");
    return CAPSGPU_OK;
  } catch (const std::exception& e) {
    capsgpu::set_last_error(e.what());
    return CAPSGPU_EINVAL;
  }
}

}  // extern "C"
