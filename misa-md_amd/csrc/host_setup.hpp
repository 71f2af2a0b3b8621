// Host-side set-up for the MI355X EAM library: setfl parsing and
// interpolation, neighbour offsets, Wigner-Seitz helpers and the world
// builder.  Everything here runs once per simulation (or once per PKA) on the
// host; the per-step work is on the GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace misa {

// Units and species (src/types/pre_define.h:11-21, atom_types.h:17-19,61-72).
constexpr double kBoltz = 8.617343e-5;
constexpr double kMvv2e = 1.0364269e-4;
constexpr double kFtm2v = 1.0 / kMvv2e;
constexpr int kTypes = 3;
inline double typeMass(int t) { return t == 0 ? 55.845 : (t == 1 ? 63.546 : (t == 2 ? 58.6934 : 0.0)); }
inline int typeZ(int t) { return t == 0 ? 26 : (t == 1 ? 29 : (t == 2 ? 28 : 0)); }

// The reference's 104-byte AtomElement (src/atom/atom_element.h:18-41).
struct AtomElement {
  uint64_t id;
  int32_t type;
  double x[3], v[3], f[3];
  double rho, df;
};
static_assert(sizeof(AtomElement) == 104, "AtomElement layout");

// Parsed DYNAMO setfl file.
struct Setfl {
  int nel = 0;
  std::vector<int> Z;
  std::vector<double> mass;
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, cutoff = 0;
  std::vector<std::vector<double>> F, rho, z2;  // z2 index i*(i+1)/2 + j, j <= i
};
bool parseSetfl(const std::string& path, Setfl& out, std::string& err);

// eam/alloy interpolation: 7 coefficients per knot, c[0..2] the derivative
// cubic (already divided by the knot spacing), c[3..6] the value cubic in the
// local coordinate p in [0, 1].  Knot slopes are fourth-order central
// differences, second-order one-sided at both ends.
void interpolate7(const std::vector<double>& y, double delta, std::vector<double>& c7);

// Ghost-extended doubled-x lattice of the single-rank domain.
struct Box {
  int64_t n[3];        // cells per dimension
  int g;               // ghost cells per face: ceil(f) + 1
  int cut_lattice;     // ceil(f)
  double a, f, rc;     // rc = a * f (atom.cpp:14-19)
  int64_t sx, sy, sz;  // ghost-extended dims of the reference array (x doubled)
  void init(const int64_t phase[3], double a, double f);
  int64_t refIndex(int64_t X, int64_t y, int64_t z) const { return (z * sy + y) * sx + X; }
  int64_t slots() const { return sx * sy * sz; }
};

// NeighbourIndex::make (neighbour_index.inl:13-76): four linear-offset lists.
struct Offsets {
  std::vector<int64_t> even, odd, half_even, half_odd;
};
void makeOffsets(const Box& b, Offsets& o);

// One neighbour offset resolved into the parity-plane layout: target plane
// parity, cell deltas and the ideal separation in lattice units.
struct PlaneOffset {
  int dp, dcx, dy, dz;  // target parity minus source parity; cell deltas
  double len;           // |ideal separation| / a
};
// Decompose linear offset `off` of the reference array for a centre of
// parity `p`; returns false if it does not decode to a lattice neighbour.
bool decodeOffset(const Box& b, int p, int64_t off, PlaneOffset& out);

// WorldBuilder::build (world_builder.cpp:64-103) into the reference AoS array.
void buildWorld(const Box& b, uint32_t seed, double tset, const int32_t ratio[3], uint32_t alloy_seed,
                std::vector<AtomElement>& atoms);

}  // namespace misa
