// Basis sets, shells and molecules for the RHF two-electron path.
//
// Ordering rules follow hfscf so that matrices line up element by element:
//   * atoms in input order, shells per atom in .gbs file order, an SP block is
//     split into an S shell followed by a P shell sharing the exponents
//     (hfscf: src/basis/hfscf_basis.cpp:175-184, 245-275);
//   * Cartesian components of a shell of angular momentum L are ordered
//     for i in 0..L: lx = L-i; for j in 0..i: ly = i-j, lz = j
//     (hfscf: src/integrals/hfscf_shell.cpp:27-40);
//   * spherical components are ordered m = -l..+l (hfscf: src/math/solid_harmonics.cpp:21-25);
//   * contraction coefficients are renormalised so that the x^L component is
//     unit-normalised (hfscf: src/integrals/hfscf_shell.cpp:44-54);
//   * coordinates are in bohr; angstrom input is multiplied by 1.8897261254535
//     (hfscf: src/molecule/hfscf_constants.hpp:5).
#pragma once
#include <array>
#include <map>
#include <string>
#include <vector>

namespace hfscf {

constexpr double kAngstromToBohr = 1.8897261254535;

inline int ncart(int L) { return (L + 1) * (L + 2) / 2; }
inline int nsph(int L) { return 2 * L + 1; }

// One contracted shell as read from a .gbs file (raw coefficients).
struct GbsShell {
    int L = 0;
    std::vector<double> alpha;
    std::vector<double> coef;
};

struct BasisLibrary {
    bool spherical = false;                     // header line of the .gbs file
    std::map<int, std::vector<GbsShell>> by_z;  // Z -> shells in file order
};

// Parse a Psi4-format .gbs file (D exponents accepted, SP split into S then P).
BasisLibrary read_gbs(const std::string& path);

struct Shell {
    int L = 0;
    int atom = 0;
    std::array<double, 3> r{};
    std::vector<double> alpha;
    std::vector<double> c;      // normalised: c_j * norm * alpha_j^((L+1.5)/2)
    std::vector<double> c_raw;  // as read from the basis file
    int idx = 0;                // Cartesian function offset
    int ids = 0;                // spherical function offset
    int ncart() const { return hfscf::ncart(L); }
    int nsph() const { return hfscf::nsph(L); }
};

// Build a shell with the hfscf normalisation (hfscf: src/integrals/hfscf_shell.cpp:44-54).
Shell make_shell(int L, int atom, const std::array<double, 3>& r, const std::vector<double>& alpha,
                 const std::vector<double>& coef, int idx, int ids);

// Cartesian -> spherical coefficient matrix, ncart(l) rows x nsph(l) columns, row-major,
// column l+m (hfscf: src/math/solid_harmonics.cpp:16-109).
std::vector<double> ylm_matrix(int l);

struct Molecule {
    std::vector<int> Z;
    std::vector<std::array<double, 3>> xyz;  // bohr
    std::vector<Shell> shells;
    bool pure = false;
    int ncart_total = 0;
    int nsph_total = 0;
    int natom() const { return (int)Z.size(); }
    int nbf() const { return pure ? nsph_total : ncart_total; }
    int nelectrons(int charge = 0) const;
    double nuclear_repulsion() const;
    // offset of shell s in the output (pure or Cartesian) basis
    int out_offset(int s) const { return pure ? shells[s].ids : shells[s].idx; }
    int out_size(int s) const { return pure ? shells[s].nsph() : shells[s].ncart(); }
    void set_geometry(const std::vector<std::array<double, 3>>& xyz_bohr);
};

// pure_override: -1 = use the basis file header, 0 = Cartesian, 1 = spherical
Molecule build_molecule(const std::vector<int>& Z, const std::vector<std::array<double, 3>>& xyz_bohr,
                        const BasisLibrary& lib, int pure_override = -1);

int element_z(const std::string& symbol);

// Read an .xyz file (angstrom) into Z and bohr coordinates.
void read_xyz(const std::string& path, std::vector<int>& Z, std::vector<std::array<double, 3>>& xyz_bohr);

}  // namespace hfscf
