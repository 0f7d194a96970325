// INCAR front-end: the reference's driver (main.cpp:233-298 reading INCAR,
// initializer.cpp:19-215 running the method) for Method HF and CCSD, with the
// CCSD iteration on the GPU through the drop-in class.
//
// Keys (one per line, "Key value", as the reference reads them):
//   Basis_Set <name>            STO-3G or 6-31G (H and O in this build)
//   Method <HF|CCSD>
//   convergance_criteria <x>    tolerance 10^x
//   Freeze_Core <true|false>    absent = false
//   Relax_Pos <true|false>      true is rejected (geometry relaxation is not on this path)
//   use_angstrom <true|false>   positions in angstrom, scaled by 1.889725989 (initializer.cpp:56-59)
//   print_stuffies <true|false>
// Atoms: "<symbol> x y z" lines between a "#ATOMS START" line and the next
// line starting with '#'; without the marker every such line in the file.
// Extensions of this build:
//   Precision <fp64|fp32|bf16x3>   amplitude-update precision before the FP64 sweeps (default fp64)
//   Checkpoint <path>              amplitude checkpoint / restart file
//   Spin_Adapted <auto|true|false> closed-shell spin-adapted amplitudes (auto: with fp64)
//   DIIS <dim>                     DIIS subspace (0 = plain Jacobi as the reference; default 8)
// Usage: ccsd_incar [INCAR]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ccsd_memory_optimized.h"
#include "ccsd_rhf.h"

namespace {

int charge_of(const std::string& sym) {
  static const char* table[] = {"H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg",
                                "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr",
                                "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};
  for (int i = 0; i < (int)(sizeof(table) / sizeof(table[0])); ++i)
    if (sym == table[i]) return i + 1;
  return -1;
}

bool is_true(const std::string& v) { return v.compare(0, 4, "true") == 0; }

struct Input {
  std::string basis, method, precision = "fp64", checkpoint, spin_adapted = "auto";
  double toler = 1e-8;
  bool freeze = false, relax = false, angstrom = false, print = false;
  int diis = 8;
  std::vector<int> Z;
  std::vector<double> R;
};

int parse(const std::string& path, Input& in) {
  std::ifstream f(path);
  if (!f) {
    std::fprintf(stderr, "Failed to open file %s\n", path.c_str());
    return 1;
  }
  std::vector<std::string> lines;
  for (std::string l; std::getline(f, l);) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
  }
  bool have_marker = false;
  for (auto& l : lines)
    if (l.rfind("#ATOMS START", 0) == 0) have_marker = true;
  bool in_atoms = !have_marker;
  for (auto& l : lines) {
    if (have_marker && l.rfind("#ATOMS START", 0) == 0) { in_atoms = true; continue; }
    if (have_marker && in_atoms && !l.empty() && l[0] == '#') { in_atoms = false; continue; }
    std::istringstream ss(l);
    std::string key, val;
    if (!(ss >> key)) continue;
    std::getline(ss >> std::ws, val);
    if (key == "Basis_Set") in.basis = val;
    else if (key == "Method") in.method = val;
    else if (key == "convergance_criteria") in.toler = std::pow(10.0, std::atof(val.c_str()));
    else if (key == "Freeze_Core") in.freeze = is_true(val);
    else if (key == "Relax_Pos") in.relax = is_true(val);
    else if (key == "use_angstrom") in.angstrom = is_true(val);
    else if (key == "print_stuffies") in.print = is_true(val);
    else if (key == "Precision") in.precision = val;
    else if (key == "Checkpoint") in.checkpoint = val;
    else if (key == "Spin_Adapted") in.spin_adapted = val;
    else if (key == "DIIS") in.diis = std::atoi(val.c_str());
    else if (in_atoms) {
      const int z = charge_of(key);
      std::istringstream vs(val);
      double x, y, zz;
      if (z > 0 && (vs >> x >> y >> zz)) {
        in.Z.push_back(z);
        in.R.insert(in.R.end(), {x, y, zz});
      }
    }
  }
  if (in.basis.empty() || in.method.empty() || in.Z.empty()) {
    std::fprintf(stderr, "INCAR needs Basis_Set, Method and at least one atom\n");
    return 1;
  }
  if (in.angstrom)
    for (double& r : in.R) r *= 1.889725989;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Input in;
  if (parse(argc > 1 ? argv[1] : "INCAR", in)) return 1;
  if (in.relax) {
    std::fprintf(stderr, "Relax_Pos true: geometry relaxation is not supported by this build\n");
    return 1;
  }
  if (in.method != "HF" && in.method != "CCSD") {
    std::fprintf(stderr, "Method %s: this build runs HF and CCSD\n", in.method.c_str());
    return 1;
  }
  const int n_nuc = (int)in.Z.size();
  int n_elec = 0;
  ccsd_compat::vec Z(n_nuc);
  ccsd_compat::mat R(n_nuc, 3);
  for (int a = 0; a < n_nuc; ++a) {
    Z(a) = in.Z[a];
    n_elec += in.Z[a];
    for (int x = 0; x < 3; ++x) R(a, x) = in.R[3 * a + x];
  }
  if (in.print) {
    std::printf("Method used: %s\nBasis set used: %s\nNumber of electrons: %d\n", in.method.c_str(),
                in.basis.c_str(), n_elec);
  }
  RHF_Hartree_Fock hf(n_nuc, Z, R, in.basis, n_elec, in.print, 0, 1, in.freeze);
  double energy;
  if (in.method == "HF") {
    energy = hf.get_Energy(in.toler, 0);
    if (hf.status() != CCSD_OK) {
      std::fprintf(stderr, "HF failed: %s\n", ccsd_strerror(hf.status()));
      return 1;
    }
    std::printf("Energy = %.12f med %g som convergens kriterie i Hartree Fock Metoden\n", energy, in.toler);
    std::printf("E_HF %.12f iterations %d\n", energy, hf.result().iterations);
    return 0;
  }
  CCSD_Memory_optimized cc(n_nuc, Z, R, in.basis, n_elec, 0, 1, &hf, in.freeze);
  ccsd_options o;
  ccsd_default_options(&o);
  if (in.precision == "fp32") o.precision = CCSD_PREC_FP32;
  else if (in.precision == "bf16x3") o.precision = CCSD_PREC_BF16X3;
  else if (in.precision != "fp64") {
    std::fprintf(stderr, "Precision %s: use fp64, fp32 or bf16x3\n", in.precision.c_str());
    return 1;
  }
  o.diis_dim = in.diis;
  cc.set_options(o);
  if (in.spin_adapted == "true") cc.set_spin_adapted(1);
  else if (in.spin_adapted == "false") cc.set_spin_adapted(0);
  else if (in.spin_adapted != "auto") {
    std::fprintf(stderr, "Spin_Adapted %s: use auto, true or false\n", in.spin_adapted.c_str());
    return 1;
  }
  if (!in.checkpoint.empty()) cc.set_checkpoint(in.checkpoint);
  energy = cc.CCSD(in.toler, in.print);
  if (cc.last_status() != CCSD_OK) {
    std::fprintf(stderr, "CCSD failed: %s\n", ccsd_strerror(cc.last_status()));
    return 1;
  }
  std::printf("Energy = %.12f med CCSD metoden\n", energy);
  std::printf("E_HF %.12f E_corr %.12f E_total %.12f iterations %d restarted %d closed_shell %d\n",
              energy - cc.correlation_energy(), cc.correlation_energy(), energy, cc.last_result().iterations,
              cc.restarted() ? 1 : 0, cc.closed_shell_used() ? 1 : 0);
  return 0;
}
