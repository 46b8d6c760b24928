// Result and checkpoint files (raw little-endian binary).
//
// The reference stores results through h5py in its examples
// (examples/hdf5_storage.py:5-13: the Lanczos energy and eigenvector, the
// Davidson eigenvalues and eigenvectors) and keeps its Lanczos basis only in
// memory (src/solvers/lanczos.cpp:123-126, 169-171).  Here both go to plain
// files with a fixed layout:
//
//   header, 64 bytes
//     char   magic[8]   "QKRYLOV" + NUL
//     u32    version    1
//     u32    kind       1 eigenpairs, 2 Lanczos checkpoint
//     u64    dim        basis dimension
//     u64    fingerprint  of the Hamiltonian (0: none recorded)
//     u64    a, b, c    kind 1: values, vectors present (0/1), 0
//                       kind 2: next step, alphas, betas
//     f64    e          kind 2: energy of the last completed step
//   payload (8-byte words)
//     kind 1: a doubles (eigenvalues), then a*dim complex doubles when b = 1
//     kind 2: alphas, betas, then (next step + 1)*dim complex doubles
//   trailer: u64 checksum of header and payload (64-bit words, see .cpp)
#pragma once

#include "qk_hamiltonian.hpp"

#include <string>
#include <vector>

namespace qk {

// Hash of what defines H: site kind, size, sector, dimension, and every
// OpSum term (coefficient bits, factor names and sites) in order.
u64 hamiltonian_fingerprint(const Hamiltonian& H);

struct LanczosState {
    int next_step = 0;              // the step the run continues with
    std::vector<double> alphas;     // next_step values
    std::vector<double> betas;      // next_step values (beta of each completed step)
    double e_old = 1e100;           // the stopping test's previous energy
    std::vector<cplx> vectors;      // (next_step + 1) * dim on load
};

// Writes st and the device vectors dev[0..next_step] (dim each).
void save_lanczos_checkpoint(const std::string& path, const Hamiltonian& H, const LanczosState& st,
                             const cplx* const* dev, hipStream_t stream);
// Reads a checkpoint written for a Hamiltonian with the same fingerprint
// and dimension (anything else is an error).
LanczosState load_lanczos_checkpoint(const std::string& path, const Hamiltonian& H);

struct EigenFile {
    u64 dim = 0;
    u64 fingerprint = 0;
    std::vector<double> values;
    bool has_vectors = false;
    std::vector<cplx> vectors;      // values.size() * dim when loaded
};

void save_eigenpairs(const std::string& path, u64 dim, u64 fingerprint, int n, const double* values,
                     const cplx* vectors);
// want_vectors = false reads the header and the values only.
EigenFile load_eigenpairs(const std::string& path, bool want_vectors);

}  // namespace qk
