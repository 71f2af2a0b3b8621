// CPU oracle for the EAM step of MISA-MD: TEST INFRASTRUCTURE ONLY.
//
// This is a restatement, in plain C++, of the reference's single-rank EAM
// molecular-dynamics step (MISA-MD v0.4.0).  It exists to check the HIP
// implementation and to serve as the timed CPU baseline.  Nothing in the
// product library links, loads or calls it; only tests/, the smoke check in
// __graft_entry__.py and the cpu_baseline leg of bench.py may use it.
//
// Every function cites the reference file:line it restates.  Layout and
// loop orders follow the reference exactly: a dense ghost-extended array of
// 104-byte AtomElement slots with doubled x index, half neighbour lists with
// both-sided accumulation into ghost slots, reverse halos for rho and force,
// a forward halo for F', the inter-atom list, decide(), and velocity Verlet.
//
// The EAM arithmetic lives in the external package `pot` v0.1.0
// (git.hpcer.dev/HPCer/CrystalMD/potential, pkg.yaml:16), which is not in the
// reference tree.  Its call sites fix the interface (chargeDensity(Z, r^2),
// dEmbedEnergy(Z, rho), toForce(Zi, Zj, r^2, df_i, df_j)); its interpolation
// is restated here from the public eam/alloy convention (cubic Hermite on the
// tabulated knots with fourth-order central-difference knot slopes, as
// published with LAMMPS pair_style eam/alloy).  Parity of that convention
// against pot itself is unpinned; the tests pin it against the closed forms
// the synthetic tables are sampled from.
#pragma once
#include <cstdint>
#include <list>
#include <map>
#include <string>
#include <vector>

namespace oracle {

// src/types/pre_define.h:11-21, atom_types.h:17-19,61-72
constexpr double kBoltz = 8.617343e-5;
constexpr double kMvv2e = 1.0364269e-4;
constexpr double kFtm2v = 1.0 / kMvv2e;
constexpr int kNumTypes = 3;
constexpr int kInvalid = -1;  // atom_type::INVALID

inline double atomMass(int t) {  // atom_types.h:38-48
  switch (t) {
    case 0: return 55.845;
    case 1: return 63.546;
    case 2: return 58.6934;
    default: return 0;
  }
}
inline int typeZ(int t) {  // atom_types.h:61-72
  switch (t) {
    case 0: return 26;
    case 1: return 29;
    case 2: return 28;
    default: return 0;
  }
}

// src/atom/atom_element.h:18-41 (104 bytes; x@16, v@40, f@64, rho@88, df@96)
struct AtomElement {
  uint64_t id;
  int32_t type;
  double x[3];
  double v[3];
  double f[3];
  double rho;
  double df;
  bool isInter() const { return type == kInvalid; }
};
static_assert(sizeof(AtomElement) == 104, "AtomElement must be 104 bytes");

// Setfl tables with interpolation coefficients (restated pot; see header).
struct Spline {
  int n = 0;
  double delta = 0, inv = 0;
  // per knot: c[0..2] derivative cubic, c[3..6] value cubic (eam/alloy order)
  std::vector<double> c;  // n * 7
  void build(const std::vector<double>& y, double delta);
  // index arithmetic shared by r- and rho-tables
  double value(double x, bool clamp_low) const;
  double deriv(double x, bool clamp_low) const;
};

struct Potential {
  int nel = 0;
  std::vector<int> Z;
  std::vector<double> mass;
  int nrho = 0, nr = 0;
  double drho = 0, dr = 0, cutoff = 0, rhomax = 0;
  std::vector<Spline> F, rho, z2;  // z2 index: i*(i+1)/2 + j, j <= i
  bool load(const std::string& path, std::string& err);
  int elemOfZ(int z) const;
  // pot interface at the reference call sites (atom.cpp:181-184, 281, 303, 341)
  double chargeDensity(int Z, double r2) const;
  double dEmbedEnergy(int Z, double rho) const;
  double toForce(int Zi, int Zj, double r2, double dfi, double dfj) const;
  // energies for the potential-energy reduction (no reference counterpart)
  double embedEnergy(int Z, double rho) const;
  double pairPhi(int Zi, int Zj, double r2) const;
};

// Single-rank BccDomain geometry (libcomm v0.3.3 fields used on the path,
// simulation.cpp:40-47; ghost = ceil(f)+1, cut_lattice = ceil(f)).
struct Domain {
  int64_t n[3];  // phase_space (cells)
  double a, f;   // lattice_const, cutoff_radius_factor
  int cut_lattice, ghost;
  int64_t sub_low[3];  // sub_box_lattice_region low (cells)
  int64_t sub[3];      // sub_box_lattice_size (cells)
  int64_t ext[3];      // ghost_extended_lattice_size (cells)
  // doubled-x ("dbx") variants
  int64_t dbx_sub[3], dbx_ghost[3], dbx_ext[3], dbx_sub_low[3], dbx_ext_low[3];
  double L[3];  // meas_global_length
  void init(const int64_t phase[3], double a, double f);
};

// src/lattice/lattice.h:48-67 with doubled x
struct Lattice {
  int64_t sx = 0, sy = 0, sz = 0, size = 0;
  int64_t index(int64_t x, int64_t y, int64_t z) const { return (z * sy + y) * sx + x; }
  void coords(int64_t i, int64_t& x, int64_t& y, int64_t& z) const {
    x = i % sx;
    i /= sx;
    y = i % sy;
    z = i / sy;
  }
};

struct NeighbourIndex {  // src/atom/neighbour_index.inl:13-76
  std::vector<int64_t> even, odd, half_even, half_odd;
  void make(const Lattice& lat, int cut_lattice, double f);
  static bool isPositive(double x, double y, double z);
  const std::vector<int64_t>& list(bool half, int64_t x) const {
    const bool ev = (x % 2 == 0);
    return half ? (ev ? half_even : half_odd) : (ev ? even : odd);
  }
};

namespace ws {  // src/lattice/ws_utils.cpp, box.h
constexpr int64_t kIndexNotExists = -1;
void voronoy(const double x[3], double a, int64_t out[3]);  // global doubled-x lattice coords
unsigned isOutBox(const double x[3], double a, const int64_t sub_low[3], const int64_t sub[3]);
}  // namespace ws

struct Sim {
  Domain d;
  Lattice lat;
  NeighbourIndex nei;
  Potential pot;
  std::vector<AtomElement> atoms;  // ghost-extended
  std::list<AtomElement> inter_list, inter_ghost_list;
  std::vector<std::vector<AtomElement*>> intersendlist, interrecvlist;
  std::vector<std::vector<int64_t>> sendlist, recvlist;
  std::map<int64_t, std::vector<AtomElement*>> inter_map;  // makeIndex (bucket order = insertion order)
  double rc = 0;
  double dt = 0, dt_inv_m[kNumTypes] = {0, 0, 0};
  // per-phase timers of the step loop (simulation.cpp:173-186)
  double t_motion = 0, t_decide = 0, t_comm = 0, t_compute = 0;

  bool init(const int64_t phase[3], double a, double f, const std::string& setfl, double dt, std::string& err);
  AtomElement& subBox(int64_t i, int64_t j, int64_t k) {
    return atoms[lat.index(d.dbx_ghost[0] + i, d.dbx_ghost[1] + j, d.dbx_ghost[2] + k)];
  }
  // world builder (world_builder.cpp:64-199)
  void buildWorld(uint32_t seed, double tset, const int ratio[3], uint32_t alloy_seed);
  void vcm(double p[4]);
  // ws helpers bound to this domain
  void nearLatCoord(const double x[3], int64_t c[3]) const;  // ghost-extended coords
  int64_t nearLatIndexInSubBox(const double x[3]) const;
  unsigned isOutBox(const double x[3]) const;
  // step phases
  void setDt(double dt);
  void firstStep();
  void secondStep();
  int decide();
  // Planted mistakes in decide, for the tests that show its checks catch
  // them (0 = the reference's rule): 1 threshold 0.2a left unsquared, 2 >=
  // in place of >, 3 refill without the stale-position box test, 4 the last
  // claimant of a vacancy wins, 5 refills into ghost sites.
  int plant = 0;
  void exchangeAtomFirst();
  void exchangeAtom();
  void exchangeInter();
  void borderInter();
  void clearForce();
  void computeEam();
  void makeIndex();
  void latRho();
  void interRho();
  void latDf();
  void latForce();
  void interForce();
  void rhoReverse();
  void dfForward();
  void forceReverse();
  void prepareForStart();
  void step();
  void setv(const int lat4[4], const double dir[3], double energy);
  void collisionStep(const int lat4[4], const double dir[3], double energy);
  // system configuration (system_configuration.cpp:26-126)
  double mvv();
  double temperature();
  // Adaptive step length (no reference counterpart): dt = min(dt_max,
  // max_disp / max|v|) over lattice and inter atoms, then setDt(dt).
  double adaptiveDt(double dt_max, double max_disp);
  void rescale(double T);
  int64_t globalAtoms() const { return 2 * d.n[0] * d.n[1] * d.n[2]; }
  // potential energy and virial (no reference counterpart)
  void energy(double out[4]);
  void systemForce(double out[3]);
};

}  // namespace oracle
