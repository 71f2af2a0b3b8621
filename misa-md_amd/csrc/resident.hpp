// Resident-mode device state and the host launchers of its kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "host_setup.hpp"

namespace misa {

// Inter atoms (atoms that left their lattice site, inter_atom_list.h:30-80).
// Local list in SoA form with a fixed capacity; ghost images and the
// nearest-site bucket index are rebuilt every step.
struct InterDev {
  int64_t cap = 0;
  uint64_t* id = nullptr;
  int8_t* type = nullptr;
  double* x[3] = {nullptr, nullptr, nullptr};
  double* v[3] = {nullptr, nullptr, nullptr};
  double* f[3] = {nullptr, nullptr, nullptr};
  double* rho = nullptr;
  double* df = nullptr;
  // scratch of the same capacity for list maintenance (order-preserving moves)
  uint64_t* id2 = nullptr;
  int8_t* type2 = nullptr;
  double* x2[3] = {nullptr, nullptr, nullptr};
  double* v2[3] = {nullptr, nullptr, nullptr};
  // new inter atoms found by the Verlet kernels (unordered, sorted later)
  int64_t* stage_key = nullptr;  // reference linear index of the vacated site
  int64_t* stage_slot = nullptr;
  uint64_t* stage_id = nullptr;
  int8_t* stage_type = nullptr;
  double* stage_x[3] = {nullptr, nullptr, nullptr};
  double* stage_v[3] = {nullptr, nullptr, nullptr};
  // ghost images: source index into the local list, image position
  int64_t gcap = 0;
  int32_t* g_src = nullptr;
  double* g_x[3] = {nullptr, nullptr, nullptr};
  // bucket index over local + ghost entries: sorted (slot << 24 | entry) keys
  uint64_t* keys = nullptr;
  uint64_t* keys_tmp = nullptr;
  int32_t* bucket_head = nullptr;  // per slot, first sorted position or -1
  int64_t* prev_slots = nullptr;   // nearest slots of the previous step's entries (flag clearing)
};

// Control block in device memory (counters and error flags).
struct Ctl {
  int64_t n_inter;      // local inter atoms
  int64_t n_ghost;      // ghost images
  int64_t n_stage;      // new inter atoms staged by the Verlet kernels
  int64_t n_prev;       // entries recorded in prev_slots
  int64_t err_flags;    // bit 0: inter capacity, bit 1: ghost capacity, bit 2: non-finite, bit 3: species
  int64_t n_keys;
  int64_t pad[2];
};

struct DevState {
  Geom G{};
  double* x[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int cur = 0;
  double* v[3] = {nullptr, nullptr, nullptr};
  double* f[3] = {nullptr, nullptr, nullptr};
  double* rho = nullptr;
  double* df64 = nullptr;
  float* df32 = nullptr;
  int8_t* type = nullptr;
  uint64_t* id = nullptr;
  InterDev in;
  Ctl* ctl = nullptr;
  double* partials = nullptr;  // reduction scratch
  int64_t npartials = 0;
};

struct VerletParams {
  double dt;
  double c[3];  // 0.5 dt ftm2v / m per species (newton_motion.cpp:19-28)
};

// Host launchers (kernels_resident.hip).
void launchUpload(const DevState& s, const AtomElement* aos_dev, int64_t sx, int64_t sy, hipStream_t st);
void launchDownload(const DevState& s, AtomElement* aos_dev, int64_t sx, int64_t sy, bool mixed, hipStream_t st);
void launchHaloAll(const DevState& s, int buf, hipStream_t st);
void launchRhoDf(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                 bool mixed, int acc, bool store_rho, hipStream_t st);
void launchForce(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                 bool mixed, int acc, int mode, const VerletParams& vp, hipStream_t st);
void launchFirst(const DevState& s, const VerletParams& vp, hipStream_t st);
// Partial sums per block: [mvv, n_lat, n_vac, fx, fy, fz, vmax2] (7 per block)
int64_t launchThermo(const DevState& s, hipStream_t st);
// Partial sums per block: [e_embed, e_pair, virial] (3 per block)
int64_t launchEnergy(const DevState& s, const LatOffsets& o, const Tables<double>& t64, hipStream_t st);
void launchScaleV(const DevState& s, double factor, hipStream_t st);
void launchSetv(const DevState& s, int64_t slot, double dv0, double dv1, double dv2, hipStream_t st);

}  // namespace misa
