// Resident-mode device state and the host launchers of its kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "host_setup.hpp"
#include "inter_common.hpp"

namespace misa {

// Inter atoms (atoms that left their lattice site, inter_atom_list.h:30-80).
// The local list is SoA in list order with a fixed capacity; the entries
// (local atoms first, then their periodic ghost images) and the nearest-site
// bucket index are rebuilt every step by the decide kernel.
struct InterDev {
  int64_t cap = 0;   // local list capacity
  int64_t ecap = 0;  // entry capacity (local + ghost images)
  int64_t kcap = 0;  // sort-key capacity (power of two >= ecap)
  uint64_t* id = nullptr;
  int8_t* type = nullptr;
  double* x[3] = {nullptr, nullptr, nullptr};
  double* v[3] = {nullptr, nullptr, nullptr};
  double* f[3] = {nullptr, nullptr, nullptr};
  double* rho = nullptr;
  double* df64 = nullptr;
  float* df32 = nullptr;
  // scratch of the same capacity for order-preserving moves
  uint64_t* id2 = nullptr;
  int8_t* type2 = nullptr;
  double* x2[3] = {nullptr, nullptr, nullptr};
  double* v2[3] = {nullptr, nullptr, nullptr};
  int32_t* tmp = nullptr;  // per local atom: recombination target / group
  // lattice atoms farther than 0.2a from their site (refilled vacancies):
  // found for the next index, and as flagged by the last one
  int64_t* far_new = nullptr;
  int64_t* far_prev = nullptr;
  // new inter atoms found by the lattice Verlet kernels (unordered)
  int64_t* stage_key = nullptr;  // reference linear index of the vacated site
  int64_t* stage_slot = nullptr;
  uint64_t* stage_id = nullptr;
  int8_t* stage_type = nullptr;
  double* stage_x[3] = {nullptr, nullptr, nullptr};
  double* stage_v[3] = {nullptr, nullptr, nullptr};
  // entries: local atoms (0..n_inter-1) then ghost images
  int32_t* e_src = nullptr;
  double* e_x[3] = {nullptr, nullptr, nullptr};
  int64_t* e_slot = nullptr;
  uint64_t* keys = nullptr;
  int32_t* head = nullptr;        // per plane slot
  int64_t* prev_slot = nullptr;   // bucket slots of the previous index (reset)
  int32_t* site_claim = nullptr;  // per plane slot, INT_MAX when idle
  FullPlaneOffsets* full = nullptr;
};

// Choice of the LDS-table force pass (the values of MISA_LDS_AUTO / ON / OFF).
constexpr int kLdsAuto = 0, kLdsOn = 1, kLdsOff = 2;

struct DevState {
  Geom G{};
  double* x[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
  int cur = 0;
  double* v[3] = {nullptr, nullptr, nullptr};
  double* f[3] = {nullptr, nullptr, nullptr};
  double* rho = nullptr;
  double* df64 = nullptr;
  int8_t* type = nullptr;
  uint64_t* id = nullptr;
  // mixed precision: site records {u, F'} per slot (ping-pong with x) and the
  // lattice vector of every pruned offset per centre parity, twice: with the
  // packed plane delta in w (2 x kMaxLatOffsets), then with the record
  // shift * 16 in w (2 x kMaxLatOffsets)
  float4* rec[2] = {nullptr, nullptr};
  float4* dl = nullptr;
  // FP64 precision: site records per slot ({x, y}, {z, w}; ping-pong with
  // x), site_common.hpp Rec64
  double2* pxy[2] = {nullptr, nullptr};
  double2* pzw[2] = {nullptr, nullptr};
  // per wavefront of the site kernels: the pruned offsets within the cutoff
  // for some lane, from the mixed density pass for the force pass (2 words)
  uint64_t* wmask = nullptr;
  // gathers in flight per wave in the mixed passes (chosen by misa_autotune)
  int uRho = 4, uForce = 4;
  // compute units of the device (persistent launches of the LDS-table force
  // pass) and the choice of that pass (MISA_LDS_AUTO / ON / OFF)
  int ncu = 0;
  int lds = 0;  // kLdsAuto, kLdsOn, kLdsOff
  // neighbour skin of the mixed density pass (A; 0: every step scans all
  // pruned offsets) and, per wavefront, the pruned offsets within cutoff +
  // skin for some lane at the last full scan (2 words)
  double skin = 0.2;
  uint64_t* smask = nullptr;
  // per z-plane: the LDS force pass's wavefront cycles, collected on the
  // launches made with zmeasure set (the XCD slab balance reads them)
  unsigned long long* zwork = nullptr;
  bool zmeasure = false;
  InterDev in;
  Ctl* ctl = nullptr;
  double* partials = nullptr;  // reduction scratch
  int64_t npartials = 0;
};

// Blocks of the grid-stride inter-atom kernels (and of their partial sums).
constexpr int kInterBlocks = 64;

struct VerletParams {
  double dt;
  double c[3];  // 0.5 dt ftm2v / m per species (newton_motion.cpp:19-28)
  // adaptive time step: the step length lives in device memory (Ctl::dt) and
  // dt, c are derived from it in the kernels; null for a fixed step
  const double* dtp = nullptr;
  // when set, the secondstep pass (MODE 0) also reduces max |v|^2 of the
  // sites it updates into this word (the adaptive step's velocity maximum)
  unsigned long long* vmax = nullptr;
};

// Host launchers (kernels_resident.hip).
void launchUpload(const DevState& s, const AtomElement* aos_dev, int64_t sx, int64_t sy, hipStream_t st);
void launchDownload(const DevState& s, AtomElement* aos_dev, int64_t sx, int64_t sy, bool mixed, hipStream_t st);
void launchHaloAll(const DevState& s, int buf, hipStream_t st);
void launchRhoDf(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                 bool mixed, int acc, bool store_rho, hipStream_t st);
// Returns the force pass launched: 0 kForce (FP64), 1 kForceM, 2 kForceL (the
// values of MISA_FORCE_PASS_*).
int launchForce(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                bool mixed, int acc, int mode, const VerletParams& vp, hipStream_t st);
void launchFirst(const DevState& s, bool mixed, const VerletParams& vp, hipStream_t st);
// Site records of buffer `buf` from x and type over every slot (the mixed
// records, or with fp64 the FP64 records).
void launchPackRec(const DevState& s, int buf, bool fp64, hipStream_t st);
// Partial sums per block: [mvv, n_lat, n_vac, fx, fy, fz, vmax2] (7 per block)
int64_t launchThermo(const DevState& s, hipStream_t st);
// Partial sums per block: [e_embed, e_pair, virial] (3 per block)
// all: visit every pruned offset (else the wavefronts' skin masks, which
// must fit the current slab cut and positions).
int64_t launchEnergy(const DevState& s, const LatOffsets& o, const Tables<double>& t64, bool all, hipStream_t st);
// The energy pass proper (kernels_hot.hip), on FP64 records and F' that
// launchEnergy has just packed.
int64_t launchEnergyPass(const DevState& s, const LatOffsets& o, const Tables<double>& t64, bool all,
                         hipStream_t st);
void launchScaleV(const DevState& s, double factor, hipStream_t st);
// Adaptive step: max |v|^2 over lattice and inter atoms into Ctl::vmax2, then
// Ctl::dt = min(dt_max, max_disp / max|v|) and Ctl::t_sim += Ctl::dt; with
// sites false the lattice sites' maximum is already in Ctl::vmax2 (reduced
// by the secondstep pass).
void launchAdaptiveDt(const DevState& s, double dt_max, double max_disp, bool sites, hipStream_t st);
// Inter-atom work (kernels_inter.hip).
// decide mode 0: the step's ownership pass (append new inter atoms, vacancy
// recombination, migration, ghost images, index, flags); mode 1: migration,
// ghost images, index and flags only (first force, PKA step, upload).
void launchDecide(const DevState& s, int mode, bool mixed, hipStream_t st);
// Ask the next decide for a full scan of the mixed density pass.
void launchRescan(const DevState& s, hipStream_t st);
InterView interView(const DevState& s, bool mixed);
// Lattice atoms farther than 0.2a from their site -> far list (before a
// mode-1 decide; in a step only refilled vacancies can be far).
void launchFarScan(const DevState& s, hipStream_t st);
void launchInterRho(const DevState& s, const Tables<double>& t64, const Tables<float>& t32, bool mixed,
                    hipStream_t st);
// mode 0: f stored, v += c f (secondstep); mode 2: f stored only
void launchInterForce(const DevState& s, const Tables<double>& t64, const Tables<float>& t32, bool mixed, int mode,
                      const VerletParams& vp, hipStream_t st);
void launchInterFirst(const DevState& s, const VerletParams& vp, hipStream_t st);
void launchInterUpload(const DevState& s, const AtomElement* aos_dev, int64_t n, hipStream_t st);
void launchInterDownload(const DevState& s, AtomElement* aos_dev, bool mixed, hipStream_t st);
void launchInterScaleV(const DevState& s, double factor, hipStream_t st);
// partials per block: [mvv, fx, fy, fz, vmax2] (5 per block), 64 blocks
int64_t launchInterThermo(const DevState& s, hipStream_t st);
// partials per block: [e_embed, e_pair, virial] (3 per block), 64 blocks
int64_t launchInterEnergy(const DevState& s, const Tables<double>& t64, hipStream_t st);
void launchSetv(const DevState& s, int64_t slot, double dv0, double dv1, double dv2, hipStream_t st);

// Batches whose step passes run as one launch for all boxes (the boxes'
// records, site_common.hpp BatchBox, in device memory).
namespace site {
struct BatchBox;
}
// Inter-atom work of every box: 0 density and F', 1 force and half kick,
// 2 first half step, 3 the decide (clear flags, list work, set flags) on
// buffer cur ^ flip, with decideArgs holding two records per box (for
// either buffer) of decideArgsBytes() each, filled by fillDecideArgs.
size_t decideArgsBytes();
void fillDecideArgs(const site::BatchBox& b, int cur, void* out);
void launchInterBatch(int which, const site::BatchBox* boxes, int nbox, int flip, const void* decideArgs,
                      hipStream_t st);
// The mixed density pass of every box (rhoStart: blocks per box, prefix
// sums) and the LDS force pass with the next half step and decide staging
// (MODE 1; vstart: virtual blocks per XCD, prefix sums; dl: the shared
// lattice vectors and packed deltas).  false when the device refuses the
// force pass's LDS size.
void launchRhoDfBatch(const site::BatchBox* boxes, const uint32_t* rhoStart, int nbox, unsigned blocks, int acc,
                      int flip, hipStream_t st);
bool launchForceBatch(const site::BatchBox* boxes, const uint32_t* vstart, int nbox, const float4* dl, int nr,
                      int ncu, int acc, int flip, hipStream_t st);

}  // namespace misa
