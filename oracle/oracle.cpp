// CPU oracle for the EAM step of MISA-MD: TEST INFRASTRUCTURE ONLY.
// See oracle.hpp for scope.  Compiled with -ffp-contract=off so that every
// expression rounds the way the reference's x86-64 build (no FMA) rounds it.
#include "oracle.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

namespace oracle {

// ---------------------------------------------------------------------------
// pot (external, v0.1.0): interpolation restated from the eam/alloy convention
// ---------------------------------------------------------------------------

// Knot slopes by fourth-order central differences (second-order one-sided at
// the two ends), cubic Hermite segments; coefficients stored per knot as
// c[0..2] = derivative cubic / delta, c[3..6] = value cubic (c[6] = y_k).
void Spline::build(const std::vector<double>& y, double d) {
  n = static_cast<int>(y.size());
  delta = d;
  inv = 1.0 / d;
  c.assign(static_cast<size_t>(n) * 7, 0.0);
  auto C = [&](int m, int k) -> double& { return c[static_cast<size_t>(m) * 7 + k]; };
  for (int m = 0; m < n; m++) C(m, 6) = y[m];
  C(0, 5) = C(1, 6) - C(0, 6);
  C(1, 5) = 0.5 * (C(2, 6) - C(0, 6));
  C(n - 2, 5) = 0.5 * (C(n - 1, 6) - C(n - 3, 6));
  C(n - 1, 5) = C(n - 1, 6) - C(n - 2, 6);
  for (int m = 2; m <= n - 3; m++)
    C(m, 5) = ((C(m - 2, 6) - C(m + 2, 6)) + 8.0 * (C(m + 1, 6) - C(m - 1, 6))) / 12.0;
  for (int m = 0; m <= n - 2; m++) {
    C(m, 4) = 3.0 * (C(m + 1, 6) - C(m, 6)) - 2.0 * C(m, 5) - C(m + 1, 5);
    C(m, 3) = C(m, 5) + C(m + 1, 5) - 2.0 * (C(m + 1, 6) - C(m, 6));
  }
  C(n - 1, 4) = 0.0;
  C(n - 1, 3) = 0.0;
  for (int m = 0; m < n; m++) {
    C(m, 2) = C(m, 5) / d;
    C(m, 1) = 2.0 * C(m, 4) / d;
    C(m, 0) = 3.0 * C(m, 3) / d;
  }
}

// Knot search: p = x/delta + 1 (one-based knot), m = int(p) clamped to
// [1 or lo, n-1], local coordinate p - m clamped to at most 1.
static inline const double* locate(const Spline& s, double x, bool clamp_low, double& p) {
  p = x * s.inv + 1.0;
  double pm = p < static_cast<double>(s.n - 1) ? p : static_cast<double>(s.n - 1);
  int m = static_cast<int>(pm);
  if (clamp_low && m < 1) m = 1;
  p -= m;
  if (p > 1.0) p = 1.0;
  return &s.c[static_cast<size_t>(m - 1) * 7];
}

double Spline::value(double x, bool clamp_low) const {
  double p;
  const double* k = locate(*this, x, clamp_low, p);
  return ((k[3] * p + k[4]) * p + k[5]) * p + k[6];
}

double Spline::deriv(double x, bool clamp_low) const {
  double p;
  const double* k = locate(*this, x, clamp_low, p);
  return (k[0] * p + k[1]) * p + k[2];
}

bool Potential::load(const std::string& path, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "open potential file " + path + " failed";  // simulation.cpp:98-103
    return false;
  }
  std::string line;
  for (int i = 0; i < 3; i++) std::getline(in, line);
  in >> nel;
  if (!in || nel < 1 || nel > 8) {
    err = "bad setfl element count";
    return false;
  }
  for (int i = 0; i < nel; i++) {
    std::string name;
    in >> name;
  }
  in >> nrho >> drho >> nr >> dr >> cutoff;
  if (!in || nrho < 5 || nr < 5) {
    err = "bad setfl grid line";
    return false;
  }
  Z.resize(nel);
  mass.resize(nel);
  F.resize(nel);
  rho.resize(nel);
  z2.resize(nel * (nel + 1) / 2);
  std::vector<double> buf;
  for (int e = 0; e < nel; e++) {
    double a;
    std::string latt;
    in >> Z[e] >> mass[e] >> a >> latt;
    buf.resize(nrho);
    for (int k = 0; k < nrho; k++) in >> buf[k];
    F[e].build(buf, drho);
    buf.resize(nr);
    for (int k = 0; k < nr; k++) in >> buf[k];
    rho[e].build(buf, dr);
  }
  for (int i = 0; i < nel; i++)
    for (int j = 0; j <= i; j++) {
      buf.resize(nr);
      for (int k = 0; k < nr; k++) in >> buf[k];
      z2[i * (i + 1) / 2 + j].build(buf, dr);
    }
  if (!in) {
    err = "truncated setfl file";
    return false;
  }
  rhomax = (nrho - 1) * drho;
  return true;
}

int Potential::elemOfZ(int z) const {
  for (int e = 0; e < nel; e++)
    if (Z[e] == z) return e;
  return -1;
}

static inline int pairIndex(int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

double Potential::chargeDensity(int Zs, double r2) const {
  return rho[elemOfZ(Zs)].value(std::sqrt(r2), false);
}

double Potential::dEmbedEnergy(int Zi, double r) const { return F[elemOfZ(Zi)].deriv(r, true); }

double Potential::embedEnergy(int Zi, double r) const {
  const Spline& s = F[elemOfZ(Zi)];
  double e = s.value(r, true);
  if (r > rhomax) e += s.deriv(r, true) * (r - rhomax);
  return e;
}

double Potential::toForce(int Zi, int Zj, double r2, double dfi, double dfj) const {
  const int ei = elemOfZ(Zi), ej = elemOfZ(Zj);
  const double r = std::sqrt(r2);
  const double rhoip = rho[ei].deriv(r, false);
  const double rhojp = rho[ej].deriv(r, false);
  const Spline& zs = z2[pairIndex(ei, ej)];
  const double z2p = zs.deriv(r, false);
  const double z2v = zs.value(r, false);
  const double recip = 1.0 / r;
  const double phi = z2v * recip;
  const double phip = z2p * recip - phi * recip;
  const double psip = dfi * rhojp + dfj * rhoip + phip;
  return -psip * recip;
}

double Potential::pairPhi(int Zi, int Zj, double r2) const {
  const double r = std::sqrt(r2);
  return z2[pairIndex(elemOfZ(Zi), elemOfZ(Zj))].value(r, false) / r;
}

// ---------------------------------------------------------------------------
// domain, lattice, neighbour offsets, Wigner-Seitz
// ---------------------------------------------------------------------------

void Domain::init(const int64_t phase[3], double a_, double f_) {
  a = a_;
  f = f_;
  cut_lattice = static_cast<int>(std::ceil(f));
  ghost = static_cast<int>(std::ceil(f)) + 1;  // simulation.cpp:46
  for (int k = 0; k < 3; k++) {
    n[k] = phase[k];
    sub_low[k] = 0;
    sub[k] = phase[k];
    ext[k] = phase[k] + 2 * ghost;
    L[k] = phase[k] * a;
  }
  dbx_sub[0] = 2 * sub[0];
  dbx_sub[1] = sub[1];
  dbx_sub[2] = sub[2];
  dbx_ghost[0] = 2 * ghost;
  dbx_ghost[1] = ghost;
  dbx_ghost[2] = ghost;
  dbx_ext[0] = 2 * ext[0];
  dbx_ext[1] = ext[1];
  dbx_ext[2] = ext[2];
  dbx_sub_low[0] = 2 * sub_low[0];
  dbx_sub_low[1] = sub_low[1];
  dbx_sub_low[2] = sub_low[2];
  dbx_ext_low[0] = 2 * (sub_low[0] - ghost);
  dbx_ext_low[1] = sub_low[1] - ghost;
  dbx_ext_low[2] = sub_low[2] - ghost;
}

bool NeighbourIndex::isPositive(double x, double y, double z) {  // neighbour_index.inl:78-92
  if (z > 0) return true;
  if (z == 0) {
    if (y > 0) return true;
    if (y == 0 && x > 0) return true;
  }
  return false;
}

void NeighbourIndex::make(const Lattice& lat, int cl, double f) {  // neighbour_index.inl:13-76
  even.clear();
  odd.clear();
  half_even.clear();
  half_odd.clear();
  const double clf = f + 2 * 0.5;  // LAT_CUTOFF 0.5 (md_building_config.h.in:24-30)
  for (int parity = 0; parity < 2; parity++) {
    const double s = parity == 0 ? 1.0 : -1.0;
    for (int64_t z = -cl - 1; z <= cl + 1; z++)
      for (int64_t y = -cl - 1; y <= cl + 1; y++)
        for (int64_t x = -2 * cl - 2; x <= 2 * cl + 2; x++) {
          const double h = static_cast<double>(x % 2) / 2;
          const double zz = static_cast<double>(z) + s * h;
          const double yy = static_cast<double>(y) + s * h;
          const double xx = static_cast<double>(x) / 2;
          const double r = xx * xx + yy * yy + zz * zz;
          if (r < clf * clf && r > 0) {
            const bool neg_odd = (x < 0 && x % 2 != 0);
            const int64_t iy = neg_odd ? (parity == 0 ? y - 1 : y + 1) : y;
            const int64_t iz = neg_odd ? (parity == 0 ? z - 1 : z + 1) : z;
            const int64_t off = lat.index(x, iy, iz);
            (parity == 0 ? even : odd).push_back(off);
            if (isPositive(xx, yy, zz)) (parity == 0 ? half_even : half_odd).push_back(off);
          }
        }
  }
}

namespace ws {
// ws_utils.cpp:15-58 (offset table, bisector normals, d = -3/4) and :66-89
static const int64_t kOff[8][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {1, 0, -1},
                                    {-1, -1, 0}, {1, -1, 0}, {-1, 0, 0}, {1, 0, 0}};
static const double kNv[8][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
                                  {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}};

void voronoy(const double x[3], double a, int64_t out[3]) {
  int64_t lx = static_cast<int64_t>(std::lround(x[0] / a));
  int64_t ly = static_cast<int64_t>(std::lround(x[1] / a));
  int64_t lz = static_cast<int64_t>(std::lround(x[2] / a));
  const double dx = x[0] / a - lx;
  const double dy = x[1] / a - ly;
  const double dz = x[2] / a - lz;
  const unsigned fl = (dz > 0 ? 4u : 0u) | (dy > 0 ? 2u : 0u) | (dx > 0 ? 1u : 0u);
  lx = 2 * lx;
  if (kNv[fl][0] * dx + kNv[fl][1] * dy + kNv[fl][2] * dz + (-3.0 / 4.0) >= 0.0) {
    lx += kOff[fl][0];
    ly += kOff[fl][1];
    lz += kOff[fl][2];
  }
  out[0] = lx;
  out[1] = ly;
  out[2] = lz;
}

unsigned isOutBox(const double x[3], double a, const int64_t low[3], const int64_t sub[3]) {  // :91-115
  int64_t c[3];
  voronoy(x, a, c);
  c[0] -= 2 * low[0];
  c[1] -= low[1];
  c[2] -= low[2];
  unsigned flag = 0;
  if (c[0] < 0) flag |= 1u;
  else if (c[0] >= 2 * sub[0]) flag |= 2u;
  if (c[1] < 0) flag |= 4u;
  else if (c[1] >= sub[1]) flag |= 8u;
  if (c[2] < 0) flag |= 16u;
  else if (c[2] >= sub[2]) flag |= 32u;
  return flag;
}
}  // namespace ws

void Sim::nearLatCoord(const double x[3], int64_t c[3]) const {  // ws_utils.cpp:154-161
  ws::voronoy(x, d.a, c);
  c[0] -= 2 * (d.sub_low[0] - d.ghost);
  c[1] -= d.sub_low[1] - d.ghost;
  c[2] -= d.sub_low[2] - d.ghost;
}

int64_t Sim::nearLatIndexInSubBox(const double x[3]) const {  // ws_utils.cpp:131-152
  int64_t c[3];
  ws::voronoy(x, d.a, c);
  int64_t j = c[0] - 2 * d.sub_low[0], k = c[1] - d.sub_low[1], l = c[2] - d.sub_low[2];
  if (j < 0 || k < 0 || l < 0 || j >= 2 * d.sub[0] || k >= d.sub[1] || l >= d.sub[2]) return ws::kIndexNotExists;
  return lat.index(j + 2 * d.ghost, k + d.ghost, l + d.ghost);
}

unsigned Sim::isOutBox(const double x[3]) const { return ws::isOutBox(x, d.a, d.sub_low, d.sub); }

// ---------------------------------------------------------------------------
// simulation set-up
// ---------------------------------------------------------------------------

bool Sim::init(const int64_t phase[3], double a, double f, const std::string& setfl, double dt_, std::string& err) {
  d.init(phase, a, f);
  lat.sx = d.dbx_ext[0];
  lat.sy = d.dbx_ext[1];
  lat.sz = d.dbx_ext[2];
  lat.size = lat.sx * lat.sy * lat.sz;
  rc = a * f;  // atom.cpp:14-19
  atoms.assign(static_cast<size_t>(lat.size), AtomElement{});
  for (auto& e : atoms) e.type = kInvalid;  // ghost slots are filled by the halo; sub-box by the builder
  nei.make(lat, d.cut_lattice, f);
  intersendlist.assign(6, {});
  interrecvlist.assign(6, {});
  setDt(dt_);
  if (!setfl.empty() && !pot.load(setfl, err)) return false;
  return true;
}

void Sim::setDt(double dt_) {  // newton_motion.cpp:13-28
  dt = dt_;
  for (int i = 0; i < kNumTypes; i++) {
    const double dt_halve = 0.5 * dt * kFtm2v;
    dt_inv_m[i] = dt_halve / atomMass(i);
  }
}

void Sim::vcm(double p[4]) {  // world_builder.cpp:158-178
  p[0] = p[1] = p[2] = p[3] = 0;
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        const double m = atomMass(at.type);
        p[0] += at.v[0] * m;
        p[1] += at.v[1] * m;
        p[2] += at.v[2] * m;
        p[3] += m;
      }
}

// world_builder.cpp:64-199, release build (MD_RAND = MT): velocities from
// std::mt19937 seeded with `seed`, alloy types from glibc rand().  The
// reference never seeds rand() (misa_md.cpp:102), so a fresh process draws
// from the state srand(1) leaves; `alloy_seed` restores that state explicitly.
void Sim::buildWorld(uint32_t seed, double tset, const int ratio[3], uint32_t alloy_seed) {
  std::mt19937 rng;
  rng.seed(seed);
  srand(alloy_seed);
  const double inv_max = 1.0 / static_cast<double>(std::mt19937::max());
  int ratio_total = 0;
  for (int i = 0; i < kNumTypes; i++) ratio_total += ratio[i];
  uint64_t id_pre = 0;
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        at.id = ++id_pre;
        const int r = rand() % ratio_total;
        int acc = 0, t = 0;
        for (int e = 0; e < kNumTypes; e++) {
          acc += ratio[e];
          if (r < acc) {
            t = e;
            break;
          }
        }
        at.type = t;
        const double m = atomMass(t);
        at.x[0] = (d.dbx_sub_low[0] + i) * 0.5 * d.a;
        at.x[1] = (d.dbx_sub_low[1] + j) * d.a + (i % 2) * (d.a / 2);
        at.x[2] = (d.dbx_sub_low[2] + k) * d.a + (i % 2) * (d.a / 2);
        at.v[0] = (rng() * inv_max - 0.5) / m;
        at.v[1] = (rng() * inv_max - 0.5) / m;
        at.v[2] = (rng() * inv_max - 0.5) / m;
        at.f[0] = at.f[1] = at.f[2] = 0;
        at.rho = at.df = 0;
      }
  double p[4];
  vcm(p);
  const double n_atoms = static_cast<double>(globalAtoms());
  if (p[3] > 0.0) {
    p[0] /= n_atoms;
    p[1] /= n_atoms;
    p[2] /= n_atoms;
  }
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)  // zeroMomentum, world_builder.cpp:143-156
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        const double m = atomMass(at.type);
        at.v[0] -= p[0] / m;
        at.v[1] -= p[1] / m;
        at.v[2] -= p[2] / m;
      }
  if (tset != 0.0) rescale(tset);
}

// ---------------------------------------------------------------------------
// configuration (system_configuration.cpp)
// ---------------------------------------------------------------------------

double Sim::mvv() {  // :61-84
  double e = 0.0;
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        if (at.type != kInvalid)
          e += (at.v[0] * at.v[0] + at.v[1] * at.v[1] + at.v[2] * at.v[2]) * atomMass(at.type);
      }
  for (auto& at : inter_list) e += (at.v[0] * at.v[0] + at.v[1] * at.v[1] + at.v[2] * at.v[2]) * atomMass(at.type);
  return e;
}

double Sim::temperature() {  // :51-59
  const uint64_t dof = 3 * static_cast<uint64_t>(globalAtoms()) - 3;
  return mvv() * kMvv2e / (dof * kBoltz);
}

double Sim::adaptiveDt(double dt_max, double max_disp) {
  double vmax2 = 0.0;
  auto visit = [&](const AtomElement& at) {
    const double v2 = (at.v[0] * at.v[0] + at.v[1] * at.v[1]) + at.v[2] * at.v[2];
    if (v2 > vmax2) vmax2 = v2;
  };
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        if (at.type != kInvalid) visit(at);
      }
  for (auto& at : inter_list) visit(at);
  const double vmax = std::sqrt(vmax2);
  const double dt_ = vmax > 0.0 ? std::fmin(dt_max, max_disp / vmax) : dt_max;
  setDt(dt_);
  return dt_;
}

void Sim::rescale(double T) {  // :86-111
  const double factor = std::sqrt(T / temperature());
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        at.v[0] *= factor;
        at.v[1] *= factor;
        at.v[2] *= factor;
      }
  for (auto& at : inter_list) {
    at.v[0] *= factor;
    at.v[1] *= factor;
    at.v[2] *= factor;
  }
}

void Sim::systemForce(double out[3]) {  // system_configuration.cpp:8-24
  out[0] = out[1] = out[2] = 0;
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        if (at.type != kInvalid)
          for (int c = 0; c < 3; c++) out[c] += at.f[c];
      }
  for (auto& at : inter_list)
    for (int c = 0; c < 3; c++) out[c] += at.f[c];
}

// ---------------------------------------------------------------------------
// integrator (newton_motion.cpp:30-74)
// ---------------------------------------------------------------------------

void Sim::firstStep() {
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        if (at.isInter()) continue;
        const double c = dt_inv_m[at.type];
        at.v[0] = at.v[0] + c * at.f[0];
        at.v[1] = at.v[1] + c * at.f[1];
        at.v[2] = at.v[2] + c * at.f[2];
        at.x[0] += dt * at.v[0];
        at.x[1] += dt * at.v[1];
        at.x[2] += dt * at.v[2];
      }
  for (auto& at : inter_list)
    for (int c = 0; c < 3; c++) {
      at.v[c] = at.v[c] + dt_inv_m[at.type] * at.f[c];
      at.x[c] += dt * at.v[c];
    }
}

void Sim::secondStep() {
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        if (at.isInter()) continue;
        const double c = dt_inv_m[at.type];
        at.v[0] += c * at.f[0];
        at.v[1] += c * at.f[1];
        at.v[2] += c * at.f[2];
      }
  for (auto& at : inter_list)
    for (int c = 0; c < 3; c++) at.v[c] += dt_inv_m[at.type] * at.f[c];
}

// ---------------------------------------------------------------------------
// ownership: decide (atom.cpp:21-84)
// ---------------------------------------------------------------------------

int Sim::decide() {
  inter_ghost_list.clear();  // clearGhost, inter_atom_list.cpp:71-74
  int nflag = 0;
  const double thr = plant == 1 ? 0.2 * d.a : std::pow(0.2 * d.a, 2.0);
  for (int64_t k = 0; k < d.dbx_sub[2]; k++)
    for (int64_t j = 0; j < d.dbx_sub[1]; j++)
      for (int64_t i = 0; i < d.dbx_sub[0]; i++) {
        AtomElement& at = subBox(i, j, k);
        if (at.isInter()) continue;
        const double xt = (i + d.dbx_sub_low[0]) * 0.5 * d.a;
        const double yt = (j + d.dbx_sub_low[1] + (i % 2) * 0.5) * d.a;
        const double zt = (k + d.dbx_sub_low[2] + (i % 2) * 0.5) * d.a;
        double dist = (at.x[0] - xt) * (at.x[0] - xt);
        dist += (at.x[1] - yt) * (at.x[1] - yt);
        dist += (at.x[2] - zt) * (at.x[2] - zt);
        if (plant == 2 ? dist >= thr : dist > thr) {
          inter_list.push_back(at);
          at.type = kInvalid;
          at.v[0] = at.v[1] = at.v[2] = 0;
          nflag = 1;
        }
      }
  if (plant == 4) inter_list.reverse();
  for (auto it = inter_list.begin(); it != inter_list.end();) {
    int64_t n = nearLatIndexInSubBox(it->x);
    if (plant == 5) {
      int64_t c[3];
      nearLatCoord(it->x, c);
      n = lat.index(c[0], c[1], c[2]);
    }
    AtomElement* near = n == ws::kIndexNotExists ? nullptr : &atoms[n];
    if (near != nullptr && near->isInter() && (plant == 3 || isOutBox(near->x) == 0u)) {
      near->id = it->id;
      near->type = it->type;
      for (int c = 0; c < 3; c++) {
        near->x[c] = it->x[c];
        near->v[c] = it->v[c];
      }
      it = inter_list.erase(it);
    } else {
      ++it;
    }
  }
  if (plant == 4) inter_list.reverse();
  return nflag;
}

// ---------------------------------------------------------------------------
// halos at one rank (libcomm neiSendReceive with the rank as its own
// neighbour in every direction; packers in src/pack/)
// ---------------------------------------------------------------------------

namespace {
struct Region {
  int64_t lo[3], hi[3];
  bool isIn(const int64_t c[3]) const {
    return c[0] >= lo[0] && c[0] < hi[0] && c[1] >= lo[1] && c[1] < hi[1] && c[2] >= lo[2] && c[2] < hi[2];
  }
};

// Forwarding (send) region for (dimension, direction) in ghost-extended dbx
// coordinates: the ghost-wide layer of the sub-box next to that face; the
// dimensions already exchanged span the full ghost-extended range, the later
// ones the sub-box only.  This is the shape LatPackerFirst::onReceive
// (lat_particle_packer.cpp:54-160) fills on the receiving side.
Region fwRegion(const Domain& d, int dim, int dir) {
  Region r;
  for (int k = 0; k < 3; k++) {
    const int64_t g = d.dbx_ghost[k], s = d.dbx_sub[k], e = d.dbx_ext[k];
    if (k == dim) {
      r.lo[k] = dir == 0 ? g : s;
      r.hi[k] = dir == 0 ? 2 * g : s + g;
    } else if (k < dim) {
      r.lo[k] = 0;
      r.hi[k] = e;
    } else {
      r.lo[k] = g;
      r.hi[k] = g + s;
    }
  }
  return r;
}

// Receive region: the ghost layer on the far side (lat_particle_packer.cpp:54-160).
Region recvRegion(const Domain& d, int dim, int dir) {
  Region r = fwRegion(d, dim, dir);
  const int64_t g = d.dbx_ghost[dim], s = d.dbx_sub[dim];
  r.lo[dim] = dir == 0 ? g + s : 0;
  r.hi[dim] = dir == 0 ? 2 * g + s : g;
  return r;
}

double shiftOf(const Domain& d, int dim, int dir) {  // lat_particle_packer.cpp:22-32
  return dir == 0 ? d.L[dim] : -d.L[dim];
}
}  // namespace

void Sim::exchangeAtomFirst() {  // atom_list.cpp:25-49
  sendlist.assign(6, {});
  recvlist.assign(6, {});
  for (int dim = 0; dim < 3; dim++)
    for (int dir = 0; dir < 2; dir++) {
      const Region r = fwRegion(d, dim, dir);
      auto& sl = sendlist[2 * dim + dir];
      for (int64_t z = r.lo[2]; z < r.hi[2]; z++)
        for (int64_t y = r.lo[1]; y < r.hi[1]; y++)
          for (int64_t x = r.lo[0]; x < r.hi[0]; x++) sl.push_back(lat.index(x, y, z));
      const Region q = recvRegion(d, dim, dir);
      auto& rl = recvlist[2 * dim + dir];
      for (int64_t z = q.lo[2]; z < q.hi[2]; z++)
        for (int64_t y = q.lo[1]; y < q.hi[1]; y++)
          for (int64_t x = q.lo[0]; x < q.hi[0]; x++) rl.push_back(lat.index(x, y, z));
    }
  exchangeAtom();
}

void Sim::exchangeAtom() {  // lat_particle_packer.cpp:163-194
  struct P {
    int32_t type;
    double r[3];
  };
  std::vector<P> buf[2];
  for (int dim = 0; dim < 3; dim++) {
    for (int dir = 0; dir < 2; dir++) {
      double off[3] = {0.0, 0.0, 0.0};
      off[dim] = shiftOf(d, dim, dir);
      auto& sl = sendlist[2 * dim + dir];
      buf[dir].resize(sl.size());
      for (size_t i = 0; i < sl.size(); i++) {
        const AtomElement& at = atoms[sl[i]];
        buf[dir][i].type = at.type;
        for (int c = 0; c < 3; c++) buf[dir][i].r[c] = at.x[c] + off[c];
      }
    }
    for (int dir = 0; dir < 2; dir++) {
      auto& rl = recvlist[2 * dim + dir];
      for (size_t i = 0; i < rl.size(); i++) {
        AtomElement& at = atoms[rl[i]];
        at.type = buf[dir][i].type;
        for (int c = 0; c < 3; c++) at.x[c] = buf[dir][i].r[c];
      }
    }
  }
}

void Sim::rhoReverse() {  // rho_packer.cpp:13-46, reversed dimension order (atom.cpp:113-115)
  std::vector<double> buf[2];
  for (int dim = 2; dim >= 0; dim--) {
    for (int dir = 0; dir < 2; dir++) {
      auto& rl = recvlist[2 * dim + (dir == 0 ? 1 : 0)];
      buf[dir].resize(rl.size());
      for (size_t i = 0; i < rl.size(); i++) buf[dir][i] = atoms[rl[i]].rho;
    }
    for (int dir = 0; dir < 2; dir++) {
      auto& sl = sendlist[2 * dim + (dir == 0 ? 1 : 0)];
      for (size_t i = 0; i < sl.size(); i++) atoms[sl[i]].rho += buf[dir][i];
    }
  }
}

void Sim::forceReverse() {  // force_packer.cpp:11-40
  std::vector<double> buf[2];
  for (int dim = 2; dim >= 0; dim--) {
    for (int dir = 0; dir < 2; dir++) {
      auto& rl = recvlist[2 * dim + (dir == 0 ? 1 : 0)];
      buf[dir].resize(3 * rl.size());
      for (size_t i = 0; i < rl.size(); i++)
        for (int c = 0; c < 3; c++) buf[dir][3 * i + c] = atoms[rl[i]].f[c];
    }
    for (int dir = 0; dir < 2; dir++) {
      auto& sl = sendlist[2 * dim + (dir == 0 ? 1 : 0)];
      for (size_t i = 0; i < sl.size(); i++)
        for (int c = 0; c < 3; c++) atoms[sl[i]].f[c] += buf[dir][3 * i + c];
    }
  }
}

void Sim::dfForward() {  // df_embed_packer.cpp:17-68
  std::vector<double> buf[2];
  for (int dim = 0; dim < 3; dim++) {
    for (int dir = 0; dir < 2; dir++) {
      const int idx = 2 * dim + dir;
      buf[dir].clear();
      for (int64_t s : sendlist[idx]) buf[dir].push_back(atoms[s].df);
      for (AtomElement* p : intersendlist[idx]) buf[dir].push_back(p->df);
    }
    for (int dir = 0; dir < 2; dir++) {
      const int idx = 2 * dim + dir;
      if (buf[dir].size() != recvlist[idx].size() + interrecvlist[idx].size()) {
        fprintf(stderr, "wrong number of dfembed recv!!!\n");  // df_embed_packer.cpp:50-53
        abort();
      }
      size_t m = 0;
      for (int64_t s : recvlist[idx]) atoms[s].df = buf[dir][m++];
      for (AtomElement* p : interrecvlist[idx]) {
        if (p == nullptr) {
          m++;
          continue;
        }
        p->df = buf[dir][m++];
      }
    }
  }
}

void Sim::exchangeInter() {  // inter_particle_packer.cpp:17-121, inter_atom_list.cpp:19-25
  static const unsigned flags[3][2] = {{1u, 2u}, {4u, 8u}, {16u, 32u}};
  std::vector<AtomElement> buf[2];
  for (int dim = 0; dim < 3; dim++) {
    for (int dir = 0; dir < 2; dir++) {
      buf[dir].clear();
      double off[3] = {0.0, 0.0, 0.0};
      off[dim] = shiftOf(d, dim, dir);
      for (auto it = inter_list.begin(); it != inter_list.end();) {
        if (isOutBox(it->x) & flags[dim][dir]) {
          AtomElement e;
          std::memset(&e, 0, sizeof(e));
          e.id = it->id;
          e.type = it->type;
          for (int c = 0; c < 3; c++) {
            e.x[c] = it->x[c] + off[c];
            e.v[c] = it->v[c];
          }
          buf[dir].push_back(e);
          it = inter_list.erase(it);
        } else {
          ++it;
        }
      }
    }
    for (int dir = 0; dir < 2; dir++)
      for (auto& e : buf[dir]) inter_list.push_back(e);
  }
}

void Sim::borderInter() {  // inter_border_packer.cpp:12-106
  intersendlist.assign(6, {});
  interrecvlist.assign(6, {});
  struct P {
    int32_t type;
    double r[3];
  };
  std::vector<P> buf[2];
  for (int dim = 0; dim < 3; dim++) {
    for (int dir = 0; dir < 2; dir++) {
      const int idx = 2 * dim + dir;
      const Region r = fwRegion(d, dim, dir);
      int64_t c[3];
      for (auto& at : inter_list) {
        nearLatCoord(at.x, c);
        if (r.isIn(c)) intersendlist[idx].push_back(&at);
      }
      for (auto& at : inter_ghost_list) {
        nearLatCoord(at.x, c);
        if (r.isIn(c)) intersendlist[idx].push_back(&at);
      }
      const double shift = shiftOf(d, dim, dir);
      buf[dir].resize(intersendlist[idx].size());
      for (size_t i = 0; i < intersendlist[idx].size(); i++) {
        const AtomElement* p = intersendlist[idx][i];
        buf[dir][i].type = p->type;
        for (int k = 0; k < 3; k++) buf[dir][i].r[k] = p->x[k] + (k == dim ? shift : 0.0);
      }
    }
    for (int dir = 0; dir < 2; dir++) {
      const int idx = 2 * dim + dir;
      interrecvlist[idx].resize(buf[dir].size());
      for (size_t i = 0; i < buf[dir].size(); i++) {
        AtomElement e;
        std::memset(&e, 0, sizeof(e));
        e.type = buf[dir][i].type;
        for (int k = 0; k < 3; k++) e.x[k] = buf[dir][i].r[k];
        inter_ghost_list.push_back(e);
        interrecvlist[idx][i] = &inter_ghost_list.back();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// EAM (atom.cpp:86-473)
// ---------------------------------------------------------------------------

void Sim::clearForce() {  // atom.cpp:86-100
  for (auto& at : atoms) {
    at.f[0] = at.f[1] = at.f[2] = 0;
    at.rho = 0;
  }
  for (auto& at : inter_list) {
    at.f[0] = at.f[1] = at.f[2] = 0;
    at.rho = 0;
  }
}

void Sim::makeIndex() {  // inter_atom_list.cpp:27-45
  inter_map.clear();
  int64_t c[3];
  for (auto& at : inter_list) {
    nearLatCoord(at.x, c);
    inter_map[lat.index(c[0], c[1], c[2])].push_back(&at);
  }
  for (auto& at : inter_ghost_list) {
    nearLatCoord(at.x, c);
    inter_map[lat.index(c[0], c[1], c[2])].push_back(&at);
  }
}

static inline double dist2Of(const double a[3], const double b[3], double del[3]) {
  del[0] = a[0] - b[0];
  del[1] = a[1] - b[1];
  del[2] = a[2] - b[2];
  return del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
}

void Sim::latRho() {  // atom.cpp:151-192
  const double rc2 = rc * rc;
  const int64_t xs = d.dbx_ghost[0], ys = d.dbx_ghost[1], zs = d.dbx_ghost[2];
  for (int64_t k = zs; k < d.dbx_sub[2] + zs; k++)
    for (int64_t j = ys; j < d.dbx_sub[1] + ys; j++)
      for (int64_t i = xs; i < d.dbx_sub[0] + xs; i++) {
        const int64_t ci = lat.index(i, j, k);
        AtomElement& c = atoms[ci];
        if (c.isInter()) continue;
        for (int64_t off : nei.list(true, i)) {
          AtomElement& n = atoms[ci + off];
          if (n.isInter()) continue;
          double del[3];
          const double dist2 = dist2Of(c.x, n.x, del);
          if (dist2 < rc2) {
            c.rho += pot.chargeDensity(typeZ(n.type), dist2);
            n.rho += pot.chargeDensity(typeZ(c.type), dist2);
          }
        }
      }
}

void Sim::interRho() {  // atom.cpp:194-284
  const double rc2 = rc * rc;
  for (auto& I : inter_list) {
    const int64_t ni = nearLatIndexInSubBox(I.x);
    if (ni == ws::kIndexNotExists) continue;
    double del[3];
    AtomElement& near = atoms[ni];
    double dist2 = dist2Of(I.x, near.x, del);
    if (!near.isInter() && dist2 < rc2) {
      I.rho += pot.chargeDensity(typeZ(near.type), dist2);
      near.rho += pot.chargeDensity(typeZ(I.type), dist2);
    }
    int64_t x, y, z;
    lat.coords(ni, x, y, z);
    const auto& full = nei.list(false, x);
    for (int64_t off : full) {
      AtomElement& n = atoms[ni + off];
      if (n.isInter()) continue;
      dist2 = dist2Of(I.x, n.x, del);
      if (dist2 < rc2) {
        I.rho += pot.chargeDensity(typeZ(n.type), dist2);
        n.rho += pot.chargeDensity(typeZ(I.type), dist2);
      }
    }
    auto b = inter_map.find(ni);
    if (b != inter_map.end())
      for (AtomElement* J : b->second) {
        if (J->id == I.id) continue;
        dist2 = dist2Of(I.x, J->x, del);
        if (dist2 < rc2) I.rho += pot.chargeDensity(typeZ(J->type), dist2);
      }
    for (int64_t off : full) {
      auto nb = inter_map.find(ni + off);
      if (nb == inter_map.end()) continue;
      for (AtomElement* J : nb->second) {
        dist2 = dist2Of(I.x, J->x, del);
        if (dist2 < rc2) I.rho += pot.chargeDensity(typeZ(J->type), dist2);
      }
    }
    I.df = pot.dEmbedEnergy(typeZ(I.type), I.rho);
  }
}

void Sim::latDf() {  // atom.cpp:286-309
  const int64_t xs = d.dbx_ghost[0], ys = d.dbx_ghost[1], zs = d.dbx_ghost[2];
  for (int64_t k = zs; k < d.dbx_sub[2] + zs; k++)
    for (int64_t j = ys; j < d.dbx_sub[1] + ys; j++)
      for (int64_t i = xs; i < d.dbx_sub[0] + xs; i++) {
        AtomElement& at = atoms[lat.index(i, j, k)];
        if (at.isInter()) continue;
        at.df = pot.dEmbedEnergy(typeZ(at.type), at.rho);
      }
}

void Sim::latForce() {  // atom.cpp:311-358
  const double rc2 = rc * rc;
  const int64_t xs = d.dbx_ghost[0], ys = d.dbx_ghost[1], zs = d.dbx_ghost[2];
  for (int64_t k = zs; k < d.dbx_sub[2] + zs; k++)
    for (int64_t j = ys; j < d.dbx_sub[1] + ys; j++)
      for (int64_t i = xs; i < d.dbx_sub[0] + xs; i++) {
        const int64_t ci = lat.index(i, j, k);
        AtomElement& c = atoms[ci];
        if (c.isInter()) continue;
        for (int64_t off : nei.list(true, i)) {
          AtomElement& n = atoms[ci + off];
          double del[3];
          const double dist2 = dist2Of(c.x, n.x, del);
          if (dist2 < rc2 && !n.isInter()) {
            const double fpair = pot.toForce(typeZ(c.type), typeZ(n.type), dist2, c.df, n.df);
            c.f[0] += del[0] * fpair;
            c.f[1] += del[1] * fpair;
            c.f[2] += del[2] * fpair;
            n.f[0] -= del[0] * fpair;
            n.f[1] -= del[1] * fpair;
            n.f[2] -= del[2] * fpair;
          }
        }
      }
}

void Sim::interForce() {  // atom.cpp:360-473
  const double rc2 = rc * rc;
  for (auto& I : inter_list) {
    const int64_t ni = nearLatIndexInSubBox(I.x);
    if (ni == ws::kIndexNotExists) continue;
    double del[3];
    AtomElement& c = atoms[ni];
    double dist2 = dist2Of(I.x, c.x, del);
    if (dist2 < rc2 && !c.isInter()) {
      const double fpair = pot.toForce(typeZ(I.type), typeZ(c.type), dist2, I.df, c.df);
      for (int k = 0; k < 3; k++) {
        I.f[k] += del[k] * fpair;
        c.f[k] -= del[k] * fpair;
      }
    }
    int64_t x, y, z;
    lat.coords(ni, x, y, z);
    const auto& full = nei.list(false, x);
    for (int64_t off : full) {
      AtomElement& n = atoms[ni + off];
      dist2 = dist2Of(I.x, n.x, del);
      if (dist2 < rc2 && !n.isInter()) {
        const double fpair = pot.toForce(typeZ(I.type), typeZ(n.type), dist2, I.df, n.df);
        for (int k = 0; k < 3; k++) {
          I.f[k] += del[k] * fpair;
          n.f[k] -= del[k] * fpair;
        }
      }
    }
    auto b = inter_map.find(ni);
    if (b != inter_map.end())
      for (AtomElement* J : b->second) {
        if (J->id == I.id) continue;
        dist2 = dist2Of(I.x, J->x, del);
        if (dist2 < rc2) {
          const double fpair = pot.toForce(typeZ(I.type), typeZ(J->type), dist2, I.df, J->df);
          for (int k = 0; k < 3; k++) I.f[k] += del[k] * fpair;
        }
      }
    for (int64_t off : full) {
      auto nb = inter_map.find(ni + off);
      if (nb == inter_map.end()) continue;
      for (AtomElement* J : nb->second) {
        dist2 = dist2Of(I.x, J->x, del);
        if (dist2 < rc2) {
          const double fpair = pot.toForce(typeZ(I.type), typeZ(J->type), dist2, I.df, J->df);
          for (int k = 0; k < 3; k++) I.f[k] += del[k] * fpair;
        }
      }
    }
  }
}

void Sim::computeEam() {  // atom.cpp:102-149
  makeIndex();
  latRho();
  interRho();
  rhoReverse();
  latDf();
  dfForward();
  latForce();
  interForce();
  forceReverse();
}

void Sim::prepareForStart() {  // simulation.cpp:87-153 (after the potential is loaded)
  exchangeAtomFirst();
  clearForce();
  computeEam();
}

void Sim::step() {  // simulation.cpp:164-194
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  firstStep();
  auto t1 = clk::now();
  decide();
  auto t2 = clk::now();
  exchangeInter();
  borderInter();
  exchangeAtom();
  auto t3 = clk::now();
  clearForce();
  computeEam();
  auto t4 = clk::now();
  secondStep();
  auto t5 = clk::now();
  auto s = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  t_motion += s(t0, t1) + s(t4, t5);
  t_decide += s(t1, t2);
  t_comm += s(t2, t3);
  t_compute += s(t3, t4);
}

void Sim::setv(const int l[4], const double dir[3], double energy) {  // atom.cpp:475-494
  if (l[0] * 2 >= d.dbx_sub_low[0] && l[0] * 2 < d.dbx_sub_low[0] + d.dbx_sub[0] && l[1] >= d.dbx_sub_low[1] &&
      l[1] < d.dbx_sub_low[1] + d.dbx_sub[1] && l[2] >= d.dbx_sub_low[2] && l[2] < d.dbx_sub_low[2] + d.dbx_sub[2]) {
    const int64_t kk = lat.index(l[0] * 2 - d.dbx_ext_low[0], l[1] - d.dbx_ext_low[1], l[2] - d.dbx_ext_low[2]) + l[3];
    AtomElement& at = atoms[kk];
    const double v_ = std::sqrt(2 * energy / atomMass(at.type) / kMvv2e);
    const double d_ = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    at.v[0] += v_ * dir[0] / d_;
    at.v[1] += v_ * dir[1] / d_;
    at.v[2] += v_ * dir[2] / d_;
  }
}

void Sim::collisionStep(const int l[4], const double dir[3], double energy) {  // simulation.cpp:208-217
  setv(l, dir, energy);
  exchangeInter();
  borderInter();
  exchangeAtom();
  clearForce();
  computeEam();
}

// Compensated (Neumaier) sum: the totals below add millions of terms, and a
// plain running sum would carry a rounding error of order N eps |E|, far
// above the per-atom bar the device's reduction is held to.
struct CompSum {
  double s = 0.0, c = 0.0;
  CompSum& operator+=(double x) {
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
    return *this;
  }
  double value() const { return s + c; }
};

// Potential energy E = sum_i F_i(rho_i) + 1/2 sum_{i != j} phi_ij(r_ij) and
// pair virial W = sum_pairs r_ij . f_ij over the state left by computeEam().
// Each lattice pair is visited once through the half lists, lattice-inter
// pairs once from the inter side, inter-inter pairs from both sides at 1/2.
void Sim::energy(double out[4]) {
  const double rc2 = rc * rc;
  CompSum e_embed, e_pair, w;
  const int64_t xs = d.dbx_ghost[0], ys = d.dbx_ghost[1], zs = d.dbx_ghost[2];
  for (int64_t k = zs; k < d.dbx_sub[2] + zs; k++)
    for (int64_t j = ys; j < d.dbx_sub[1] + ys; j++)
      for (int64_t i = xs; i < d.dbx_sub[0] + xs; i++) {
        const int64_t ci = lat.index(i, j, k);
        AtomElement& c = atoms[ci];
        if (c.isInter()) continue;
        e_embed += pot.embedEnergy(typeZ(c.type), c.rho);
        for (int64_t off : nei.list(true, i)) {
          AtomElement& n = atoms[ci + off];
          double del[3];
          const double dist2 = dist2Of(c.x, n.x, del);
          if (dist2 < rc2 && !n.isInter()) {
            e_pair += pot.pairPhi(typeZ(c.type), typeZ(n.type), dist2);
            w += dist2 * pot.toForce(typeZ(c.type), typeZ(n.type), dist2, c.df, n.df);
          }
        }
      }
  for (auto& I : inter_list) {
    const int64_t ni = nearLatIndexInSubBox(I.x);
    if (ni == ws::kIndexNotExists) continue;
    e_embed += pot.embedEnergy(typeZ(I.type), I.rho);
    int64_t x, y, z;
    lat.coords(ni, x, y, z);
    std::vector<int64_t> sites{0};
    for (int64_t off : nei.list(false, x)) sites.push_back(off);
    for (int64_t off : sites) {
      AtomElement& n = atoms[ni + off];
      double del[3];
      const double dist2 = dist2Of(I.x, n.x, del);
      if (dist2 < rc2 && !n.isInter()) {
        e_pair += pot.pairPhi(typeZ(I.type), typeZ(n.type), dist2);
        w += dist2 * pot.toForce(typeZ(I.type), typeZ(n.type), dist2, I.df, n.df);
      }
      auto nb = inter_map.find(ni + off);
      if (nb == inter_map.end()) continue;
      for (AtomElement* J : nb->second) {
        if (off == 0 && J->id == I.id) continue;
        const double d2 = dist2Of(I.x, J->x, del);
        if (d2 < rc2) {
          e_pair += 0.5 * pot.pairPhi(typeZ(I.type), typeZ(J->type), d2);
          w += 0.5 * d2 * pot.toForce(typeZ(I.type), typeZ(J->type), d2, I.df, J->df);
        }
      }
    }
  }
  out[0] = e_embed.value();
  out[1] = e_pair.value();
  out[2] = out[0] + out[1];
  out[3] = w.value();
}

}  // namespace oracle
