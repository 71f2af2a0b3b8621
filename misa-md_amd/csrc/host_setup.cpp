// Host-side set-up for the MI355X EAM library.  Compiled with
// -ffp-contract=off so the interpolation tables and the world builder round
// every operation separately, as the reference's x86-64 build does.
#include "host_setup.hpp"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <random>

namespace misa {

bool parseSetfl(const std::string& path, Setfl& s, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "open potential file " + path + " failed";  // simulation.cpp:98-103
    return false;
  }
  std::string line;
  for (int i = 0; i < 3; i++) std::getline(in, line);  // comment lines
  in >> s.nel;
  if (!in || s.nel < 1 || s.nel > 8) {
    err = "setfl: bad element count";
    return false;
  }
  for (int i = 0; i < s.nel; i++) {
    std::string name;
    in >> name;
  }
  in >> s.nrho >> s.drho >> s.nr >> s.dr >> s.cutoff;
  if (!in || s.nrho < 5 || s.nr < 5 || !(s.drho > 0) || !(s.dr > 0)) {
    err = "setfl: bad grid line";
    return false;
  }
  s.Z.assign(s.nel, 0);
  s.mass.assign(s.nel, 0);
  s.F.assign(s.nel, {});
  s.rho.assign(s.nel, {});
  s.z2.assign(s.nel * (s.nel + 1) / 2, {});
  for (int e = 0; e < s.nel; e++) {
    double lat_const;
    std::string lat_type;
    in >> s.Z[e] >> s.mass[e] >> lat_const >> lat_type;
    s.F[e].resize(s.nrho);
    for (auto& v : s.F[e]) in >> v;
    s.rho[e].resize(s.nr);
    for (auto& v : s.rho[e]) in >> v;
  }
  for (int i = 0; i < s.nel; i++)
    for (int j = 0; j <= i; j++) {
      auto& t = s.z2[i * (i + 1) / 2 + j];
      t.resize(s.nr);
      for (auto& v : t) in >> v;
    }
  if (!in) {
    err = "setfl: file ends before all tables were read";
    return false;
  }
  return true;
}

void interpolate7(const std::vector<double>& y, double delta, std::vector<double>& c) {
  const int n = static_cast<int>(y.size());
  c.assign(static_cast<size_t>(n) * 7, 0.0);
  double* k = c.data();
  auto at = [k](int m, int j) -> double& { return k[7 * m + j]; };
  for (int m = 0; m < n; m++) at(m, 6) = y[m];
  // knot slopes (in units of the knot index)
  at(0, 5) = at(1, 6) - at(0, 6);
  at(1, 5) = 0.5 * (at(2, 6) - at(0, 6));
  at(n - 2, 5) = 0.5 * (at(n - 1, 6) - at(n - 3, 6));
  at(n - 1, 5) = at(n - 1, 6) - at(n - 2, 6);
  for (int m = 2; m <= n - 3; m++)
    at(m, 5) = ((at(m - 2, 6) - at(m + 2, 6)) + 8.0 * (at(m + 1, 6) - at(m - 1, 6))) / 12.0;
  // Hermite cubic of segment [m, m+1]
  for (int m = 0; m <= n - 2; m++) {
    const double dy = at(m + 1, 6) - at(m, 6);
    at(m, 4) = 3.0 * dy - 2.0 * at(m, 5) - at(m + 1, 5);
    at(m, 3) = at(m, 5) + at(m + 1, 5) - 2.0 * dy;
  }
  at(n - 1, 4) = 0.0;
  at(n - 1, 3) = 0.0;
  for (int m = 0; m < n; m++) {
    at(m, 2) = at(m, 5) / delta;
    at(m, 1) = 2.0 * at(m, 4) / delta;
    at(m, 0) = 3.0 * at(m, 3) / delta;
  }
}

void Box::init(const int64_t phase[3], double a_, double f_) {
  a = a_;
  f = f_;
  rc = a * f;
  cut_lattice = static_cast<int>(std::ceil(f));
  g = cut_lattice + 1;  // simulation.cpp:46
  for (int k = 0; k < 3; k++) n[k] = phase[k];
  sx = 2 * (n[0] + 2 * g);
  sy = n[1] + 2 * g;
  sz = n[2] + 2 * g;
}

// Every site within (f + 1) lattice constants of a centre, enumerated over
// the doubled-x box [-2cl-2, 2cl+2] x [-cl-1, cl+1]^2; odd x offsets land on
// the other sublattice, whose cell index shifts by one for negative odd x.
void makeOffsets(const Box& b, Offsets& o) {
  o = Offsets();
  const int cl = b.cut_lattice;
  const double reach = b.f + 1.0;  // f + 2 * LAT_CUTOFF (md_building_config.h.in:24-30)
  for (int parity = 0; parity < 2; parity++) {
    const double sgn = parity == 0 ? 1.0 : -1.0;
    auto& full = parity == 0 ? o.even : o.odd;
    auto& half = parity == 0 ? o.half_even : o.half_odd;
    for (int64_t z = -cl - 1; z <= cl + 1; z++)
      for (int64_t y = -cl - 1; y <= cl + 1; y++)
        for (int64_t X = -2 * cl - 2; X <= 2 * cl + 2; X++) {
          const double h = static_cast<double>(X % 2) / 2;
          const double rz = static_cast<double>(z) + sgn * h;
          const double ry = static_cast<double>(y) + sgn * h;
          const double rx = static_cast<double>(X) / 2;
          const double r2 = rx * rx + ry * ry + rz * rz;
          if (!(r2 < reach * reach && r2 > 0)) continue;
          const bool shifted = X < 0 && X % 2 != 0;
          const int64_t dy = shifted ? y - static_cast<int64_t>(sgn) : y;
          const int64_t dz = shifted ? z - static_cast<int64_t>(sgn) : z;
          const int64_t off = (dz * b.sy + dy) * b.sx + X;
          full.push_back(off);
          const bool positive = rz > 0 || (rz == 0 && (ry > 0 || (ry == 0 && rx > 0)));
          if (positive) half.push_back(off);
        }
  }
}

static int64_t centred(int64_t v, int64_t m) {
  int64_t r = ((v % m) + m) % m;
  if (r > m / 2) r -= m;
  return r;
}

bool decodeOffset(const Box& b, int p, int64_t off, PlaneOffset& out) {
  const int64_t X = centred(off, b.sx);
  const int64_t rest = (off - X) / b.sx;
  const int64_t dy = centred(rest, b.sy);
  const int64_t dz = (rest - dy) / b.sy;
  if ((dz * b.sy + dy) * b.sx + X != off) return false;
  const int pt = static_cast<int>((p + X) & 1);
  out.dp = pt - p;
  out.dcx = static_cast<int>((p + X - pt) / 2);
  out.dy = static_cast<int>(dy);
  out.dz = static_cast<int>(dz);
  const double h = 0.5 * out.dp;
  const double rx = out.dcx + h, ry = out.dy + h, rz = out.dz + h;
  out.len = std::sqrt(rx * rx + ry * ry + rz * rz);
  return true;
}

void buildWorld(const Box& b, uint32_t seed, double tset, const int32_t ratio[3], uint32_t alloy_seed,
                std::vector<AtomElement>& atoms) {
  atoms.assign(static_cast<size_t>(b.slots()), AtomElement{});
  for (auto& at : atoms) at.type = -1;
  std::mt19937 rng;
  rng.seed(seed);   // md_rand::initSeed, random.cpp:14-20 (MD_RAND = MT)
  srand(alloy_seed);  // randomAtomsType draws rand() % sum(ratio), world_builder.cpp:180-199
  const double inv_max = 1.0 / static_cast<double>(std::mt19937::max());
  int total = 0;
  for (int t = 0; t < kTypes; t++) total += ratio[t];
  const int64_t nx2 = 2 * b.n[0];
  uint64_t id = 0;
  for (int64_t k = 0; k < b.n[2]; k++)
    for (int64_t j = 0; j < b.n[1]; j++)
      for (int64_t i = 0; i < nx2; i++) {
        AtomElement& at = atoms[b.refIndex(i + 2 * b.g, j + b.g, k + b.g)];
        at.id = ++id;
        const int r = rand() % total;
        int acc = 0, t = 0;
        for (int e = 0; e < kTypes; e++) {
          acc += ratio[e];
          if (r < acc) {
            t = e;
            break;
          }
        }
        at.type = t;
        const double m = typeMass(t);
        at.x[0] = i * 0.5 * b.a;
        at.x[1] = j * b.a + (i % 2) * (b.a / 2);
        at.x[2] = k * b.a + (i % 2) * (b.a / 2);
        at.v[0] = (rng() * inv_max - 0.5) / m;
        at.v[1] = (rng() * inv_max - 0.5) / m;
        at.v[2] = (rng() * inv_max - 0.5) / m;
      }
  // zero momentum: subtract (sum m v) / N / m_i (world_builder.cpp:76-92,143-156)
  double p[4] = {0, 0, 0, 0};
  auto each = [&](auto&& fn) {
    for (int64_t k = 0; k < b.n[2]; k++)
      for (int64_t j = 0; j < b.n[1]; j++)
        for (int64_t i = 0; i < nx2; i++) fn(atoms[b.refIndex(i + 2 * b.g, j + b.g, k + b.g)]);
  };
  each([&](AtomElement& at) {
    const double m = typeMass(at.type);
    for (int c = 0; c < 3; c++) p[c] += at.v[c] * m;
    p[3] += m;
  });
  const double natoms = static_cast<double>(2 * b.n[0] * b.n[1] * b.n[2]);
  if (p[3] > 0.0)
    for (int c = 0; c < 3; c++) p[c] /= natoms;
  each([&](AtomElement& at) {
    const double m = typeMass(at.type);
    for (int c = 0; c < 3; c++) at.v[c] -= p[c] / m;
  });
  if (tset != 0.0) {  // configuration::rescale, system_configuration.cpp:86-111
    double mvv = 0;
    each([&](AtomElement& at) {
      if (at.type >= 0) mvv += (at.v[0] * at.v[0] + at.v[1] * at.v[1] + at.v[2] * at.v[2]) * typeMass(at.type);
    });
    const uint64_t dof = 3 * static_cast<uint64_t>(natoms) - 3;
    const double T = mvv * kMvv2e / (dof * kBoltz);
    const double s = std::sqrt(tset / T);
    each([&](AtomElement& at) {
      for (int c = 0; c < 3; c++) at.v[c] *= s;
    });
  }
}

}  // namespace misa
