// C entry points of the CPU oracle for the test suite (ctypes): TEST
// INFRASTRUCTURE ONLY.  See oracle.hpp.
#include <cstring>
#include <string>

#include "oracle.hpp"

using namespace oracle;

static void copyErr(const std::string& e, char* err, int errlen) {
  if (err && errlen > 0) {
    std::strncpy(err, e.c_str(), errlen - 1);
    err[errlen - 1] = 0;
  }
}

extern "C" {

void* oracle_create(const int64_t* n, double a, double f, const char* setfl, double dt, char* err, int errlen) {
  auto* s = new Sim();
  std::string e;
  if (!s->init(n, a, f, setfl ? setfl : "", dt, e)) {
    copyErr(e, err, errlen);
    delete s;
    return nullptr;
  }
  return s;
}

void oracle_destroy(void* h) { delete static_cast<Sim*>(h); }

void oracle_build_world(void* h, uint32_t seed, double tset, const int* ratio, uint32_t alloy_seed) {
  static_cast<Sim*>(h)->buildWorld(seed, tset, ratio, alloy_seed);
}

void oracle_prepare(void* h) { static_cast<Sim*>(h)->prepareForStart(); }

void oracle_run(void* h, int64_t steps) {
  Sim* s = static_cast<Sim*>(h);
  for (int64_t i = 0; i < steps; i++) s->step();
}

void oracle_set_dt(void* h, double dt) { static_cast<Sim*>(h)->setDt(dt); }

void oracle_plant(void* h, int which) { static_cast<Sim*>(h)->plant = which; }

void oracle_setv(void* h, const int* lat4, const double* dir, double e) { static_cast<Sim*>(h)->setv(lat4, dir, e); }

void oracle_collision(void* h, const int* lat4, const double* dir, double e) {
  static_cast<Sim*>(h)->collisionStep(lat4, dir, e);
}

int oracle_phase(void* h, int which) {
  Sim* s = static_cast<Sim*>(h);
  switch (which) {
    case 0: s->firstStep(); break;
    case 1: return s->decide();
    case 2: s->exchangeInter(); break;
    case 3: s->borderInter(); break;
    case 4: s->exchangeAtom(); break;
    case 5: s->clearForce(); break;
    case 6: s->computeEam(); break;
    case 7: s->secondStep(); break;
    case 8: s->makeIndex(); break;
    case 9: s->latRho(); break;
    case 10: s->interRho(); break;
    case 11: s->rhoReverse(); break;
    case 12: s->latDf(); break;
    case 13: s->dfForward(); break;
    case 14: s->latForce(); break;
    case 15: s->interForce(); break;
    case 16: s->forceReverse(); break;
    case 17: s->exchangeAtomFirst(); break;
    default: return -1;
  }
  return 0;
}

int64_t oracle_num_slots(void* h) { return static_cast<Sim*>(h)->lat.size; }
void* oracle_atoms(void* h) { return static_cast<Sim*>(h)->atoms.data(); }
void oracle_lattice_dims(void* h, int64_t* out) {
  Sim* s = static_cast<Sim*>(h);
  out[0] = s->lat.sx;
  out[1] = s->lat.sy;
  out[2] = s->lat.sz;
}
int oracle_ghost(void* h) { return static_cast<Sim*>(h)->d.ghost; }
double oracle_cutoff(void* h) { return static_cast<Sim*>(h)->rc; }

int64_t oracle_inter_count(void* h) { return static_cast<int64_t>(static_cast<Sim*>(h)->inter_list.size()); }
void oracle_inter_get(void* h, AtomElement* out) {
  for (auto& a : static_cast<Sim*>(h)->inter_list) *out++ = a;
}
void oracle_inter_set(void* h, const AtomElement* in, int64_t n) {
  Sim* s = static_cast<Sim*>(h);
  s->inter_list.clear();
  for (int64_t i = 0; i < n; i++) s->inter_list.push_back(in[i]);
}
int64_t oracle_inter_ghost_count(void* h) {
  return static_cast<int64_t>(static_cast<Sim*>(h)->inter_ghost_list.size());
}
void oracle_inter_ghost_get(void* h, AtomElement* out) {
  for (auto& a : static_cast<Sim*>(h)->inter_ghost_list) *out++ = a;
}

double oracle_mvv(void* h) { return static_cast<Sim*>(h)->mvv(); }
double oracle_adaptive_dt(void* h, double dt_max, double max_disp) {
  return static_cast<Sim*>(h)->adaptiveDt(dt_max, max_disp);
}
double oracle_temperature(void* h) { return static_cast<Sim*>(h)->temperature(); }
void oracle_rescale(void* h, double T) { static_cast<Sim*>(h)->rescale(T); }
void oracle_energy(void* h, double* out) { static_cast<Sim*>(h)->energy(out); }
void oracle_system_force(void* h, double* out) { static_cast<Sim*>(h)->systemForce(out); }
void oracle_vcm(void* h, double* out) { static_cast<Sim*>(h)->vcm(out); }
void oracle_timers(void* h, double* out) {
  Sim* s = static_cast<Sim*>(h);
  out[0] = s->t_motion;
  out[1] = s->t_decide;
  out[2] = s->t_comm;
  out[3] = s->t_compute;
}

int oracle_nei_make(int64_t sx, int64_t sy, int64_t sz, int cut_lattice, double f, int64_t* even, int64_t* odd,
                    int64_t* heven, int64_t* hodd, int64_t cap, int64_t* counts) {
  Lattice l;
  l.sx = sx;
  l.sy = sy;
  l.sz = sz;
  l.size = sx * sy * sz;
  NeighbourIndex n;
  n.make(l, cut_lattice, f);
  counts[0] = static_cast<int64_t>(n.even.size());
  counts[1] = static_cast<int64_t>(n.odd.size());
  counts[2] = static_cast<int64_t>(n.half_even.size());
  counts[3] = static_cast<int64_t>(n.half_odd.size());
  if (counts[0] > cap || counts[1] > cap) return -1;
  std::memcpy(even, n.even.data(), n.even.size() * sizeof(int64_t));
  std::memcpy(odd, n.odd.data(), n.odd.size() * sizeof(int64_t));
  std::memcpy(heven, n.half_even.data(), n.half_even.size() * sizeof(int64_t));
  std::memcpy(hodd, n.half_odd.data(), n.half_odd.size() * sizeof(int64_t));
  return 0;
}

int oracle_nei_is_positive(double x, double y, double z) { return NeighbourIndex::isPositive(x, y, z) ? 1 : 0; }

void oracle_ws_voronoy(const double* x, double a, int64_t* out) { ws::voronoy(x, a, out); }
unsigned oracle_ws_is_out_box(const double* x, double a, const int64_t* low, const int64_t* sub) {
  return ws::isOutBox(x, a, low, sub);
}

void* oracle_pot_load(const char* path, char* err, int errlen) {
  auto* p = new Potential();
  std::string e;
  if (!p->load(path, e)) {
    copyErr(e, err, errlen);
    delete p;
    return nullptr;
  }
  return p;
}
void oracle_pot_free(void* p) { delete static_cast<Potential*>(p); }

double oracle_pot_eval(void* pp, int what, int Zi, int Zj, double x, double dfi, double dfj) {
  const Potential* p = static_cast<const Potential*>(pp);
  switch (what) {
    case 0: return p->chargeDensity(Zi, x);
    case 1: return p->dEmbedEnergy(Zi, x);
    case 2: return p->toForce(Zi, Zj, x, dfi, dfj);
    case 3: return p->embedEnergy(Zi, x);
    case 4: return p->pairPhi(Zi, Zj, x);
    default: return 0;
  }
}

// Coefficients (7 per knot) of table `idx` of `kind` (0 F, 1 rho, 2 r*phi).
int64_t oracle_pot_table(void* pp, int kind, int idx, double* out, int64_t cap) {
  const Potential* p = static_cast<const Potential*>(pp);
  const std::vector<Spline>& v = kind == 0 ? p->F : (kind == 1 ? p->rho : p->z2);
  if (idx < 0 || idx >= static_cast<int>(v.size())) return -1;
  const int64_t n = static_cast<int64_t>(v[idx].c.size());
  if (out && n <= cap) std::memcpy(out, v[idx].c.data(), n * sizeof(double));
  return n;
}

void* oracle_sim_pot(void* h) { return &static_cast<Sim*>(h)->pot; }

}  // extern "C"
