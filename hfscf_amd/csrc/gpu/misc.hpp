// Launchers of the small kernels of misc_kernels.hip that only the host side of the library calls
// (kept out of kernels.hpp, which every class-kernel translation unit includes).
#pragma once
#include <hip/hip_runtime.h>

#include "gather.hpp"
#include "kernels.hpp"
#include "layout.hpp"

namespace hfg {

// Plain copy of n 16-byte elements (a kernel of known HBM traffic).
hipError_t launch_copy(const double2* src, double2* dst, size_t n, hipStream_t s);
// z = a x + b y over n doubles.
hipError_t launch_axpby(double a, const double* x, double b, const double* y, double* z, size_t n, hipStream_t s);
// Largest |D| of every shell block of every molecule into the density-screening table.
hipError_t launch_block_absmax(const double* D, const MolDev* mols, const int* shell_ao, const int* shell_nf,
                               const int* shell_mol, int nshell, float* dmax, hipStream_t s);
// Two-pass gather: the chunk sums of every tile's J and K gather lists into part, then per
// tile the chunk sums in chunk order (cJ_off / cK_off: CSR offsets of each tile's chunks) into
// the symmetric J and K. Deterministic: the partition and the order of every sum are fixed.
hipError_t launch_gather_chunks(const GatherArgs& a, const GatherChunk* chunks, const int* cJ_off, const int* cK_off,
                                int nchunks, int ntiles, double* part, hipStream_t s);
// F_0..F_m at n points through the device Boys functions (test hook).
hipError_t launch_boys_probe(int m, const double* T, const double* tab, double* out, int n, int closed_form,
                             hipStream_t s);

// Writes the address of the hardware queue that runs stream s's kernels into out[k] (device).
hipError_t launch_queue_probe(unsigned long long* out, int k, hipStream_t s);

}  // namespace hfg
