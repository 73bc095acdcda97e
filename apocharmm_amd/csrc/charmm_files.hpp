// CHARMM input and restart files on the host: PSF topology, CRD coordinates,
// parameter / stream files, and the resolution of a PSF against its
// parameters into the arrays apo_nb_create takes.  Restart files follow the
// layout RestartSubscriber writes (src/RestartSubscriber.cu:73-445).
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace apo {
namespace files {

struct Psf {
  std::string file_name;
  std::vector<std::string> segids, resnames, names, types;  // types as written
  std::vector<int> resids;
  std::vector<double> charges, masses;
  // 0-based atom indices, 2/3/4/4/8 per term
  std::vector<int32_t> bonds, angles, dihedrals, impropers, crossterms;
  int natom() const { return (int)types.size(); }
  // (OT, HT, HT) triples of consecutive atoms (CharmmPSF::getWaterMolecules)
  std::vector<int32_t> waters() const;
};
Psf read_psf(const std::string &path);

struct Crd {
  std::vector<double> xyz;  // 3 * natom
  int natom() const { return (int)(xyz.size() / 3); }
};
Crd read_crd(const std::string &path);

using Key2 = std::array<std::string, 2>;
using Key3 = std::array<std::string, 3>;
using Key4 = std::array<std::string, 4>;

// Force-field parameters keyed by upper-case type names.
struct Params {
  std::map<Key2, std::array<double, 2>> bonds;   // (b0, kb)
  std::map<Key3, std::array<double, 2>> angles;  // (theta0 [rad], k)
  std::map<Key3, std::array<double, 2>> ureyb;   // (s0, kub)
  std::map<Key4, std::vector<std::array<double, 3>>> dihedrals;  // (n, k, delta [rad]) per term
  std::map<Key4, std::array<double, 3>> impropers;               // (0, k, psi0 [rad])
  std::map<std::string, std::array<double, 2>> vdw, vdw14;       // (eps, rmin/2)
  std::map<Key2, std::array<double, 2>> nbfix;                   // (|emin|, rmin)
  void read(const std::string &path);
};

// A PSF resolved against its parameters: the arrays of apo_nb_topology
// (vdW types = sorted set of the PSF's types, (c6, c12) tables with NBFIX,
// bonded lists with their parameters, one dihedral row per Fourier term).
struct Resolved {
  int ntypes = 0;
  std::vector<std::string> type_names;
  std::vector<int32_t> vdwtype;
  std::vector<float> vdwtab, vdwtab14;
  std::vector<int32_t> bonds, ureyb, angles, dihe, imdihe;
  std::vector<float> bond_par, ub_par, angle_par, dihe_par, imdihe_par;
  std::vector<int32_t> waters;
  // SHAKE groups of hydrogens bonded to one heavy atom (outside waters):
  // (heavy, h1, h2 or -1, h3 or -1) and one squared target length per group,
  // the b0 of the group's last hydrogen bond, as
  // ForceManager::initializeHolonomicConstraintsVariables builds them
  // (src/ForceManager.cu:844-918)
  std::vector<int32_t> shake_groups;
  std::vector<double> shake_r2;
};
Resolved resolve(const Psf &psf, const Params &prm);

// Restart file (RestartSubscriber::update, src/RestartSubscriber.cu:73-445;
// read as CudaLangevinThermostatIntegrator::initializeFromRestartFile,
// src/CudaLangevinThermostatIntegrator.cu:264-520).  x / v / dprev are
// 3 * natom in atom order: the "XOLD" section holds the coordinates, "VX"
// the velocities and "X" the previous half-step displacements, as the
// reference writes them.  rng_state is the step count of the noise stream.
struct Restart {
  double box[3] = {0, 0, 0};
  int natom = 0;
  unsigned long long npriv = 0;  // steps run since the start of the trajectory
  int nstep = 0;                 // steps of the last propagate call
  int nsavc = 0;
  int ndegf = 0;
  uint64_t seed = 0;
  unsigned long long rng_state = 0;
  bool apo = true;  // written by this library or the reference (APO flag)
  double avetem = 0.0;
  std::vector<double> x, v, dprev;
};
void write_restart(const std::string &path, const Restart &r);
Restart read_restart(const std::string &path);

}  // namespace files
}  // namespace apo
