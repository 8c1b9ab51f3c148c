// Host one-electron integrals for RHF: overlap S, kinetic T, nuclear attraction V and
// their first-derivative contractions for the analytic gradient.
//
// These stay on the host CPU (they are O(N^2 * natom)); they mirror what hfscf computes in
// OSOverlap/OSNuclear (hfscf: src/integrals/hfscf_osoverlap.cpp:10, hfscf_osnuclear.cpp:17-83,
// hfscf_osoverlapderiv1.cpp:8, hfscf_oskineticderiv1.cpp:8, hfscf_osnuclearderiv1.cpp:15),
// computed here with McMurchie-Davidson Hermite expansions. All matrices are nbf x nbf,
// row-major, in the molecule's output basis (spherical when mol.pure).
#pragma once
#include <vector>

#include "basis.hpp"

namespace hfscf {

// Boys function F_m(T), m = 0..mmax, accurate to ~1e-15 relative.
void boys_host(int mmax, double T, double* F);

void overlap_kinetic(const Molecule& mol, std::vector<double>& S, std::vector<double>& T);
void nuclear_attraction(const Molecule& mol, std::vector<double>& V);

// One-electron part of the RHF gradient (natom x 3, added into grad):
//   2 sum_ij D_ij d(T+V)_ij/dX  -  2 sum_ij Q_ij dS_ij/dX
// with D the half density and Q = D F D (hfscf: src/gradient/hfscf_grad_rhf.cpp:20,143-167).
void oneint_gradient(const Molecule& mol, const double* D, const double* Q, double* grad);

// Nuclear-repulsion gradient (natom x 3, added into grad).
void nuclear_repulsion_gradient(const Molecule& mol, double* grad);

}  // namespace hfscf
