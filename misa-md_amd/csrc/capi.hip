// C-ABI of the MI355X EAM library (declared in include/misa_eam.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/misa_eam.h"
#include "device_common.hpp"
#include "hook.hpp"
#include "host_setup.hpp"
#include "resident.hpp"
#include "site_common.hpp"

using namespace misa;

struct misa_handle {
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  Box box{};
  DevState s{};
  Offsets ref{};
  bool have_offsets = false;
  LatOffsets lat{};
  FullOffsets full{};
  FullPlaneOffsets fullPlane{};
  float4 dlHost[2 * kMaxLatOffsets]{};  // lattice vectors of the pruned offsets (mixed records)
  void* interBlock = nullptr;  // one allocation carved into the inter-atom arrays
  AtomElement* interAos = nullptr;
  int64_t interAosCap = 0;
  Setfl setfl{};
  bool have_pot = false;
  std::vector<std::vector<double>> cF, cR, cZ;  // 7-coefficient tables by file element
  bool species_present[3] = {false, false, false};
  double* tab64 = nullptr;
  float* tab32 = nullptr;
  Tables<double> t64{};
  Tables<float> t32{};
  int precision = MISA_PREC_FP64;
  int acc = 0;
  double dt = 0.001;
  // adaptive step (max_disp > 0): dt = min(dt_max, max_disp / max|v|) per step
  double dt_max = 0.001, max_disp = 0.0;
  double t_fixed = 0.0;  // simulated time of the fixed-step calls (ps)
  bool state_ready = false;
  bool forces_ready = false;
  uint64_t gen = 0;  // bumped by every change a captured step graph would bake in
  uint64_t skin_gen = ~0ull;  // gen at which the density pass's skin masks were last invalidated
  uint64_t mask_gen = ~0ull;  // gen of the last density pass (its wave masks fit the current slab cut)
  AtomElement* aos_dev = nullptr;
  std::vector<void*> allocs;
  bool profile = false;
  std::vector<hipEvent_t> events;
  std::vector<int> event_cls;
  double timings[4] = {0, 0, 0, 0};
  int64_t launches[3] = {0, 0, 0};
  int force_pass = -1;       // MISA_FORCE_PASS_* of the last force pass launched
  bool range_check = false;  // a density beyond the F(rho) table fails the call
  int64_t flags_seen = 0;    // the device flags at the last check
  // XCD slabs cut by measured work (misa_set_xcd_balance); virtual blocks
  // per parity per XCD the wave-mask and partial-sum arrays are sized for
  bool balance = true;
  int64_t bpx_cap = 0;
};

namespace {

int fail(misa_handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

#define HIPCHECK(h, call)                                                                          \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return fail(h, MISA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
int devAlloc(misa_handle* h, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, count * sizeof(T) + 64);
  if (e != hipSuccess) return fail(h, MISA_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  h->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return MISA_OK;
}

VerletParams verletOf(double dt) {  // NewtonMotion::preComputeDtInv2m (newton_motion.cpp:19-28)
  VerletParams vp;
  vp.dt = dt;
  for (int t = 0; t < kTypes; t++) {
    const double dt_halve = 0.5 * dt * kFtm2v;
    vp.c[t] = dt_halve / typeMass(t);
  }
  return vp;
}

int buildPlaneOffsets(misa_handle* h) {
  const Box& b = h->box;
  const double keep = b.f + 0.41;  // rc/a + 2 * 0.2 (max displacement) + 0.01 margin
  for (int p = 0; p < 2; p++) {
    const std::vector<int64_t>& L = p == 0 ? h->ref.even : h->ref.odd;
    if (static_cast<int>(L.size()) > kMaxFullOffsets) return fail(h, MISA_E_ARG, "offset list too long");
    int nl = 0, nw = 0;
    h->full.n[p] = static_cast<int>(L.size());
    for (size_t k = 0; k < L.size(); k++) {
      PlaneOffset po;
      if (!decodeOffset(b, p, L[k], po)) return fail(h, MISA_E_ARG, "offset does not decode to a lattice neighbour");
      if (std::abs(po.dcx) > b.g || std::abs(po.dy) > b.g || std::abs(po.dz) > b.g)
        return fail(h, MISA_E_ARG, "offset reaches beyond the ghost layer");
      h->full.off[p][k] = static_cast<int32_t>(L[k]);
      {
        const int64_t NX = b.n[0] + 2 * b.g, NY = b.n[1] + 2 * b.g, NZ = b.n[2] + 2 * b.g;
        h->fullPlane.off[p][k] =
            static_cast<int32_t>(po.dp * NX * NY * NZ + (static_cast<int64_t>(po.dz) * NY + po.dy) * NX + po.dcx);
      }
      if (po.len >= keep) {
        const int64_t NX = b.n[0] + 2 * b.g, NY = b.n[1] + 2 * b.g, NZ = b.n[2] + 2 * b.g;
        h->fullPlane.woff[p][nw++] =
            static_cast<int32_t>(po.dp * NX * NY * NZ + (static_cast<int64_t>(po.dz) * NY + po.dy) * NX + po.dcx);
      } else {
        if (nl >= kPadOffset) return fail(h, MISA_E_ARG, "pruned offset list too long");
        const int64_t NX = b.n[0] + 2 * b.g, NY = b.n[1] + 2 * b.g, NZ = b.n[2] + 2 * b.g;
        const int64_t plane = NX * NY * NZ;
        const int64_t d = po.dp * plane + (static_cast<int64_t>(po.dz) * NY + po.dy) * NX + po.dcx;
        // lattice vector from the centre's site to the neighbour's site
        const double hdp = 0.5 * po.dp;
        const uint32_t packed = packDelta(po.dp, po.dz, po.dy, po.dcx);
        float packedBits;
        std::memcpy(&packedBits, &packed, sizeof(packedBits));
        h->dlHost[p * kMaxLatOffsets + nl] =
            make_float4(static_cast<float>((po.dcx + hdp) * b.a), static_cast<float>((po.dy + hdp) * b.a),
                        static_cast<float>((po.dz + hdp) * b.a), packedBits);
        h->lat.off[p][nl++] = static_cast<int32_t>(d);
      }
    }
    h->lat.n[p] = nl;
    for (int k = nl; k < kMaxLatOffsets; k++)  // the padding entries (kPadOffset among them)
      h->dlHost[p * kMaxLatOffsets + k] = make_float4(1e18f, 1e18f, 1e18f, 0.0f);
    h->fullPlane.n[p] = static_cast<int>(L.size());
    h->fullPlane.nw[p] = nw;
    for (int k = static_cast<int>(L.size()); k < kMaxFullOffsets; k++) h->fullPlane.off[p][k] = 0;
    for (int k = nw; k < kMaxFullOffsets; k++) h->fullPlane.woff[p][k] = 0;
  }
  if (h->fullPlane.n[0] != h->fullPlane.n[1]) return fail(h, MISA_E_ARG, "even and odd offset lists differ in length");
  if (h->s.in.full) {
    hipError_t e = hipMemcpy(h->s.in.full, &h->fullPlane, sizeof(FullPlaneOffsets), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(h, MISA_E_HIP, std::string("hipMemcpy offsets: ") + hipGetErrorString(e));
  }
  // non-negative form for the buffer loads: sh = off - lo, lo = most negative delta
  int32_t lo = 0;
  for (int p = 0; p < 2; p++)
    for (int k = 0; k < h->lat.n[p]; k++) lo = std::min(lo, h->lat.off[p][k]);
  h->lat.lo = lo;
  for (int p = 0; p < 2; p++) {
    for (int k = 0; k < h->lat.n[p]; k++) h->lat.sh[p][k] = static_cast<uint32_t>(h->lat.off[p][k] - lo);
    // the padding entries read the site's own record (delta 0): a valid
    // address for either parity, whereas lo alone lies below the array for
    // the even sites
    for (int k = h->lat.n[p]; k < kMaxLatOffsets; k++) h->lat.sh[p][k] = static_cast<uint32_t>(-lo);
  }
  {
    // every entry, the padding included, must address a slot of the array
    // from every sub-box site of its parity (slots grow with cx, cy, cz, so
    // the first and the last site bound them)
    const int64_t NX = b.n[0] + 2 * b.g, NY = b.n[1] + 2 * b.g, NZ = b.n[2] + 2 * b.g;
    const int64_t plane = NX * NY * NZ;
    for (int p = 0; p < 2; p++) {
      const int64_t first = p * plane + (static_cast<int64_t>(b.g) * NY + b.g) * NX + b.g;
      const int64_t last = p * plane + (static_cast<int64_t>(b.n[2] - 1 + b.g) * NY + (b.n[1] - 1 + b.g)) * NX +
                           (b.n[0] - 1 + b.g);
      for (int k = 0; k < kMaxLatOffsets; k++) {
        const int64_t delta = static_cast<int64_t>(h->lat.sh[p][k]) + lo;
        if (first + delta < 0 || last + delta >= 2 * plane)
          return fail(h, MISA_E_ARG, "a pruned offset reaches beyond the site arrays");
      }
    }
  }
  if (h->s.dl) {
    // device table: the lattice vectors with the packed deltas in w, then
    // the same vectors with the record shifts * 16 in w
    std::vector<float4> dev(4 * kMaxLatOffsets);
    for (int i = 0; i < 2 * kMaxLatOffsets; i++) {
      dev[i] = h->dlHost[i];
      const int p = i / kMaxLatOffsets, k = i % kMaxLatOffsets;
      const uint32_t sh16 = h->lat.sh[p][k] * 16u;
      float w;
      std::memcpy(&w, &sh16, sizeof(w));
      dev[2 * kMaxLatOffsets + i] = make_float4(h->dlHost[i].x, h->dlHost[i].y, h->dlHost[i].z, w);
    }
    hipError_t e = hipMemcpy(h->s.dl, dev.data(), dev.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(h, MISA_E_HIP, std::string("hipMemcpy offsets: ") + hipGetErrorString(e));
  }
  h->have_offsets = true;
  h->gen++;
  return MISA_OK;
}

// Traversal of the site kernels: z-slab per XCD, y-tiles of up to 32 rows
// (on the 200^3 cascade box, tiles of 20-25 rows step about 1 % faster than
// tiles of 10, and tiles of 40 rows or more are slower again).
// Slab bounds cut at whole planes, and the virtual blocks per parity of the
// largest slab.
void setSlabs(Geom& G, const int* zlo) {
  Walk& W = G.W;
  int64_t bmax = 0;
  for (int c = 0; c <= 8; c++) W.zlo[c] = zlo[c];
  for (int c = 0; c < 8; c++) {
    const int64_t sites = static_cast<int64_t>(W.zlo[c + 1] - W.zlo[c]) * G.n1 * G.n0;
    bmax = std::max(bmax, (sites + 255) / 256);
  }
  W.bpx = bmax;
}
void evenSlabs(Geom& G) {
  int zlo[9];
  for (int c = 0; c <= 8; c++) zlo[c] = static_cast<int>((static_cast<int64_t>(c) * G.n2) / 8);
  setSlabs(G, zlo);
}

void makeWalk(Geom& G) {
  Walk& W = G.W;
  W.ty = 1;
  for (int t = std::min(32, G.n1); t >= 1; t--)
    if (G.n1 % t == 0) {
      W.ty = t;
      break;
    }
  evenSlabs(G);
}

int syncAndCheck(misa_handle* h) {
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  Ctl c;
  HIPCHECK(h, hipMemcpy(&c, h->s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  h->flags_seen = c.err_flags;
  if (c.err_flags & kErrInterCap) return fail(h, MISA_E_CAPACITY, "inter-atom list capacity exceeded");
  if (c.err_flags & kErrGhostCap) return fail(h, MISA_E_CAPACITY, "ghost inter-atom capacity exceeded");
  if (c.err_flags & kErrNonFinite) return fail(h, MISA_E_DEVICE, "non-finite force or electron density");
  if (c.err_flags & kErrLeftBox) return fail(h, MISA_E_DEVICE, "an inter atom left the ghost-extended box");
  if ((c.err_flags & kErrRhoRange) && h->range_check)
    return fail(h, MISA_E_DEVICE, "electron density beyond the embedding-function table");
  if (c.err_flags & ~int64_t(kErrInterCap | kErrGhostCap | kErrNonFinite | kErrLeftBox | kErrRhoRange))
    return fail(h, MISA_E_DEVICE, "device error flag set");
  return MISA_OK;
}

// Inter-atom arrays for `cap` local atoms in one allocation; per-slot bucket
// heads and recombination claims are sized by the box.
int allocInter(misa_handle* h, int64_t cap) {
  InterDev& in = h->s.in;
  if (h->interBlock) {
    (void)hipFree(h->interBlock);
    h->interBlock = nullptr;
  }
  const int64_t ecap = 4 * cap;
  int64_t kcap = 1;
  while (kcap < ecap) kcap <<= 1;
  const int64_t slots = 2 * h->s.G.plane;
  size_t bytes = 0;
  auto need = [&](size_t n) {
    const size_t off = bytes;
    bytes += (n + 255) & ~size_t(255);
    return off;
  };
  struct Part {
    void** p;
    size_t off;
  };
  std::vector<Part> parts;
  auto add = [&](auto** ptr, size_t n, size_t elem) { parts.push_back({reinterpret_cast<void**>(ptr), need(n * elem)}); };
  add(&in.id, cap, 8);
  add(&in.type, cap, 1);
  add(&in.id2, cap, 8);
  add(&in.type2, cap, 1);
  add(&in.rho, cap, 8);
  add(&in.df64, cap, 8);
  add(&in.df32, cap, 4);
  add(&in.tmp, cap, 4);
  add(&in.far_new, cap, 8);
  add(&in.far_prev, cap, 8);
  add(&in.stage_key, cap, 8);
  add(&in.stage_slot, cap, 8);
  add(&in.stage_id, cap, 8);
  add(&in.stage_type, cap, 1);
  for (int k = 0; k < 3; k++) {
    add(&in.x[k], cap, 8);
    add(&in.v[k], cap, 8);
    add(&in.f[k], cap, 8);
    add(&in.x2[k], cap, 8);
    add(&in.v2[k], cap, 8);
    add(&in.stage_x[k], cap, 8);
    add(&in.stage_v[k], cap, 8);
    add(&in.e_x[k], ecap, 8);
  }
  add(&in.e_src, ecap, 4);
  add(&in.e_slot, ecap, 8);
  add(&in.prev_slot, ecap, 8);
  add(&in.keys, kcap, 8);
  add(&in.head, slots, 4);
  add(&in.site_claim, slots, 4);
  add(&in.full, 1, sizeof(FullPlaneOffsets));
  char* base = nullptr;
  hipError_t e = hipMalloc(&base, bytes);
  if (e != hipSuccess) return fail(h, MISA_E_HIP, std::string("hipMalloc inter: ") + hipGetErrorString(e));
  h->interBlock = base;
  for (auto& pt : parts) *pt.p = base + pt.off;
  in.cap = cap;
  in.ecap = ecap;
  in.kcap = kcap;
  h->gen++;
  if (hipMemset(base, 0, bytes) != hipSuccess || hipMemset(in.head, 0xff, slots * 4) != hipSuccess ||
      hipMemset(in.site_claim, 0x7f, slots * 4) != hipSuccess ||
      hipMemcpy(in.full, &h->fullPlane, sizeof(FullPlaneOffsets), hipMemcpyHostToDevice) != hipSuccess)
    return fail(h, MISA_E_HIP, "initialising the inter-atom arrays failed");
  return MISA_OK;
}

struct Timer {
  misa_handle* h;
  int cls;
  size_t e0 = 0;
  Timer(misa_handle* hh, int c) : h(hh), cls(c) {
    if (!h->profile) return;
    e0 = h->events.size();
    h->event_cls.push_back(c);
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    h->events.push_back(a);
    h->events.push_back(b);
    (void)hipEventRecord(a, h->stream);
  }
  ~Timer() {
    if (!h->profile) return;
    (void)hipEventRecord(h->events[e0 + 1], h->stream);
  }
};

// Staging array of misa_lat_hook_noalloc: a fixed device array per GPU in
// place of the buffer the first hook call allocates.  Host code never names
// it in a kernel argument or copy: `scratch` on the host is only a stand-in
// symbol, and the array's device address on each GPU comes from
// hipGetSymbolAddress, looked up once per device at misa_create.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* scratchDev[kMaxDevices] = {};
static std::mutex scratchLock;

// The device address of `scratch` on the current device `device` (looked up
// on the first call for that device).
static hipError_t scratchOf(int device, void** ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(scratchLock);
  if (!scratchDev[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    scratchDev[device] = p;
  }
  *ptr = scratchDev[device];
  return hipSuccess;
}

}  // namespace

extern "C" {

int misa_create(const misa_geom* g, int device, misa_handle** out) {
  if (!g || !out) return MISA_E_ARG;
  *out = nullptr;
  if (g->phase_space[0] < 1 || g->phase_space[1] < 1 || g->phase_space[2] < 1 || !(g->lattice_const > 0) ||
      !(g->cutoff_radius_factor > 0))
    return MISA_E_ARG;
  auto* h = new misa_handle();
  h->device = device;
  void* stage = nullptr;  // the staging array of misa_lat_hook_noalloc on this device
  if (hipSetDevice(device) != hipSuccess || scratchOf(device, &stage) != hipSuccess) {
    delete h;
    return MISA_E_HIP;
  }
  h->box.init(g->phase_space, g->lattice_const, g->cutoff_radius_factor);
  const Box& b = h->box;
  if (b.n[0] < b.g || b.n[1] < b.g || b.n[2] < b.g) {
    delete h;
    return MISA_E_ARG;  // a periodic image must not be its own neighbour's ghost source twice
  }
  Geom& G = h->s.G;
  G.n0 = static_cast<int>(b.n[0]);
  G.n1 = static_cast<int>(b.n[1]);
  G.n2 = static_cast<int>(b.n[2]);
  G.g = b.g;
  G.NX = G.n0 + 2 * b.g;
  G.NY = G.n1 + 2 * b.g;
  G.NZ = G.n2 + 2 * b.g;
  G.plane = static_cast<int64_t>(G.NX) * G.NY * G.NZ;
  G.nsub = b.n[0] * b.n[1] * b.n[2];
  G.a = b.a;
  G.rc2 = b.rc * b.rc;
  G.L0 = b.n[0] * b.a;
  G.L1 = b.n[1] * b.a;
  G.L2 = b.n[2] * b.a;
  G.thr = std::pow(0.2 * b.a, 2.0);
  makeWalk(G);
  // buffer loads address the planes with 32-bit byte offsets (site + shifted delta)
  if (G.nsub > (int64_t(1) << 31) || 4 * G.plane * 8 >= (int64_t(1) << 31)) {
    delete h;
    return MISA_E_ARG;
  }
  if (hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return MISA_E_HIP;
  }
  const size_t slots = static_cast<size_t>(2 * G.plane);
  DevState& s = h->s;
  int rc = MISA_OK;
  for (int k = 0; k < 3 && rc == MISA_OK; k++) {
    rc |= devAlloc(h, &s.x[0][k], slots);
    rc |= devAlloc(h, &s.x[1][k], slots);
    rc |= devAlloc(h, &s.v[k], slots);
    rc |= devAlloc(h, &s.f[k], slots);
  }
  if (rc == MISA_OK) rc |= devAlloc(h, &s.rho, slots);
  if (rc == MISA_OK) rc |= devAlloc(h, &s.df64, slots);
  if (rc == MISA_OK) rc |= devAlloc(h, &s.rec[0], slots);
  if (rc == MISA_OK) rc |= devAlloc(h, &s.rec[1], slots);
  for (int b = 0; b < 2 && rc == MISA_OK; b++) {
    rc |= devAlloc(h, &s.pxy[b], slots);
    rc |= devAlloc(h, &s.pzw[b], slots);
  }
  if (rc == MISA_OK) rc |= devAlloc(h, &s.dl, 4 * kMaxLatOffsets);
  // a balanced slab may hold up to twice the planes of an even cut
  h->bpx_cap = 2 * G.W.bpx;
  if (rc == MISA_OK) rc |= devAlloc(h, &s.wmask, static_cast<size_t>(2 * 4 * 16 * h->bpx_cap));
  if (rc == MISA_OK) rc |= devAlloc(h, &s.smask, static_cast<size_t>(2 * 4 * 16 * h->bpx_cap));
  if (rc == MISA_OK) rc |= devAlloc(h, &s.zwork, static_cast<size_t>(G.n2));
  if (rc == MISA_OK) rc |= devAlloc(h, &s.type, slots);
  if (rc == MISA_OK) rc |= devAlloc(h, &s.id, slots);
  if (rc == MISA_OK) rc |= devAlloc(h, &s.ctl, 1);
  if (rc == MISA_OK) {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device) == hipSuccess) s.ncu = ncu;
  }
  const int64_t blocks = 16 * h->bpx_cap;
  // room for the site kernels' 7 partials per block, the inter kernels' 5
  // and the persistent energy pass's 3 per compute unit
  s.npartials = std::max<int64_t>(std::max<int64_t>(8 * blocks, 8 * kInterBlocks), 3 * static_cast<int64_t>(s.ncu));
  if (rc == MISA_OK) rc |= devAlloc(h, &s.partials, static_cast<size_t>(s.npartials));
  if (rc == MISA_OK) rc |= allocInter(h, 65536);
  if (rc != MISA_OK) {
    misa_destroy(h);
    return MISA_E_HIP;
  }
  for (int k = 0; k < 3; k++) {
    (void)hipMemset(s.x[0][k], 0, slots * 8);
    (void)hipMemset(s.x[1][k], 0, slots * 8);
    (void)hipMemset(s.v[k], 0, slots * 8);
    (void)hipMemset(s.f[k], 0, slots * 8);
  }
  (void)hipMemset(s.rho, 0, slots * 8);
  (void)hipMemset(s.df64, 0, slots * 8);
  (void)hipMemset(s.rec[0], 0, slots * sizeof(float4));
  (void)hipMemset(s.rec[1], 0, slots * sizeof(float4));
  (void)hipMemset(s.type, 0xff, slots);
  (void)hipMemset(s.id, 0, slots * 8);
  (void)hipMemset(s.ctl, 0, sizeof(Ctl));
  if (hipDeviceSynchronize() != hipSuccess) {
    misa_destroy(h);
    return MISA_E_HIP;
  }
  *out = h;
  return MISA_OK;
}

int misa_destroy(misa_handle* h) {
  if (!h) return MISA_E_ARG;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void* p : h->allocs) (void)hipFree(p);
  if (h->aos_dev) (void)hipFree(h->aos_dev);
  if (h->interBlock) (void)hipFree(h->interBlock);
  if (h->interAos) (void)hipFree(h->interAos);
  for (auto e : h->events) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return MISA_OK;
}

const char* misa_last_error(const misa_handle* h) { return h ? h->err.c_str() : "null handle"; }

int misa_set_offsets(misa_handle* h, const int64_t* even, int64_t n_even, const int64_t* odd, int64_t n_odd,
                     const int64_t* half_even, int64_t n_half_even, const int64_t* half_odd, int64_t n_half_odd) {
  if (!h) return MISA_E_ARG;
  if (!even || !odd || n_even <= 0 || n_odd <= 0) return fail(h, MISA_E_ARG, "empty offset lists");
  h->ref.even.assign(even, even + n_even);
  h->ref.odd.assign(odd, odd + n_odd);
  h->ref.half_even.assign(half_even, half_even + (half_even ? n_half_even : 0));
  h->ref.half_odd.assign(half_odd, half_odd + (half_odd ? n_half_odd : 0));
  return buildPlaneOffsets(h);
}

int misa_make_offsets(misa_handle* h) {
  if (!h) return MISA_E_ARG;
  makeOffsets(h->box, h->ref);
  return buildPlaneOffsets(h);
}

int misa_get_offsets(const misa_handle* h, int which, int64_t* out, int64_t cap, int64_t* n) {
  if (!h || !n) return MISA_E_ARG;
  const std::vector<int64_t>* v = nullptr;
  switch (which) {
    case 0: v = &h->ref.even; break;
    case 1: v = &h->ref.odd; break;
    case 2: v = &h->ref.half_even; break;
    case 3: v = &h->ref.half_odd; break;
    case 4:
    case 5: {  // pruned plane deltas (4: even, 5: odd)
      const int p = which - 4;
      *n = h->lat.n[p];
      if (out && *n <= cap)
        for (int k = 0; k < h->lat.n[p]; k++) out[k] = h->lat.off[p][k];
      return MISA_OK;
    }
    default: return MISA_E_ARG;
  }
  *n = static_cast<int64_t>(v->size());
  if (out && *n <= cap) std::memcpy(out, v->data(), v->size() * sizeof(int64_t));
  return MISA_OK;
}

int misa_set_potential_file(misa_handle* h, const char* path) {
  if (!h || !path) return MISA_E_ARG;
  std::string e;
  Setfl& sf = h->setfl;
  if (!parseSetfl(path, sf, e)) return fail(h, MISA_E_IO, e);
  h->gen++;
  h->cF.assign(sf.nel, {});
  h->cR.assign(sf.nel, {});
  h->cZ.assign(sf.z2.size(), {});
  for (int i = 0; i < sf.nel; i++) {
    interpolate7(sf.F[i], sf.drho, h->cF[i]);
    interpolate7(sf.rho[i], sf.dr, h->cR[i]);
  }
  for (size_t i = 0; i < sf.z2.size(); i++) interpolate7(sf.z2[i], sf.dr, h->cZ[i]);
  // species -> file element by atomic number (atom_types.h:61-72)
  int el[3];
  for (int t = 0; t < kTypes; t++) {
    el[t] = -1;
    for (int e2 = 0; e2 < sf.nel; e2++)
      if (sf.Z[e2] == typeZ(t)) el[t] = e2;
    h->species_present[t] = el[t] >= 0;
  }
  if (el[0] < 0 && el[1] < 0 && el[2] < 0) return fail(h, MISA_E_IO, "setfl holds none of Fe, Cu, Ni");
  const int64_t sF = static_cast<int64_t>(sf.nrho) * kKnot, sR = static_cast<int64_t>(sf.nr) * kKnot;
  const int64_t nr = sf.nr;
  const int64_t oRhoV = 3 * sF + 9 * sR;            // [3][nr][4] density value cubics
  const int64_t oPair = oRhoV + 3 * nr * kRhoRec;  // [9][nr][8] force records per ordered pair
  const int64_t oRhoH = oPair + 9 * nr * kPairRec;  // [3][nr][2] compact density knots {y, s}
  const int64_t oRhoD = oRhoH + 3 * nr * 2;         // [3][nr][4] density slope cubics / dr (FP32 force pass)
  const int64_t oPairZ = oRhoD + 3 * nr * kRhoRec;  // [9][nr][4] r*phi value cubics (FP32 force pass)
  const int64_t total = oPairZ + 9 * nr * kRhoRec;
  std::vector<double> h64(static_cast<size_t>(total), 0.0);
  auto put = [&](int64_t base, const std::vector<double>& c7, int n) {
    for (int m = 0; m < n; m++) {
      double* k = &h64[base + static_cast<int64_t>(m) * kKnot];
      const double* q = &c7[static_cast<size_t>(m) * 7];
      k[0] = q[0];
      k[1] = q[1];
      k[2] = q[2];
      k[3] = 0.0;
      k[4] = q[3];
      k[5] = q[4];
      k[6] = q[5];
      k[7] = q[6];
    }
  };
  for (int t = 0; t < 3; t++) {
    if (el[t] < 0) continue;
    put(t * sF, h->cF[el[t]], sf.nrho);
    put(3 * sF + t * sR, h->cR[el[t]], sf.nr);
  }
  for (int a = 0; a < 3; a++)
    for (int b2 = 0; b2 <= a; b2++) {
      if (el[a] < 0 || el[b2] < 0) continue;
      const int ea = std::max(el[a], el[b2]), eb = std::min(el[a], el[b2]);
      put(3 * sF + 3 * sR + (a * (a + 1) / 2 + b2) * sR, h->cZ[ea * (ea + 1) / 2 + eb], sf.nr);
    }
  // packed records: the same coefficients, regrouped for one gather per pair
  for (int t = 0; t < 3; t++) {
    if (el[t] < 0) continue;
    const std::vector<double>& R = h->cR[el[t]];
    for (int64_t m = 0; m < nr; m++)
      for (int q = 0; q < 4; q++) h64[oRhoV + (t * nr + m) * kRhoRec + q] = R[m * 7 + 3 + q];
  }
  for (int a = 0; a < 3; a++)
    for (int b2 = 0; b2 < 3; b2++) {
      if (el[a] < 0 || el[b2] < 0) continue;
      const int ea = std::max(el[a], el[b2]), eb = std::min(el[a], el[b2]);
      const std::vector<double>& Rb = h->cR[el[b2]];
      const std::vector<double>& Z = h->cZ[ea * (ea + 1) / 2 + eb];
      for (int64_t m = 0; m < nr; m++) {
        double* r = &h64[oPair + ((a * 3 + b2) * nr + m) * kPairRec];
        double* z = &h64[oPairZ + ((a * 3 + b2) * nr + m) * kRhoRec];
        for (int q = 0; q < 4; q++) {
          r[q] = Z[m * 7 + 3 + q];
          r[4 + q] = Rb[m * 7 + 3 + q];
          z[q] = Z[m * 7 + 3 + q];
        }
      }
    }
  std::vector<float> h32(h64.size());
  for (size_t i = 0; i < h64.size(); i++) h32[i] = static_cast<float>(h64[i]);
  const double rdr = 1.0 / sf.dr;
  // FP32 density knots {y, s} of the density pass (Hermite form of the same
  // segments: c[5] is the knot slope s, c[6] the value y), and the density
  // slope cubics of the force pass divided by the knot spacing
  for (int t = 0; t < 3; t++) {
    if (el[t] < 0) continue;
    const std::vector<double>& R = h->cR[el[t]];
    for (int64_t m = 0; m < nr; m++) {
      float* kh = &h32[oRhoH + (t * nr + m) * 2];
      kh[0] = static_cast<float>(R[m * 7 + 6]);
      kh[1] = static_cast<float>(R[m * 7 + 5]);
      float* kd = &h32[oRhoD + (t * nr + m) * kRhoRec];
      kd[0] = static_cast<float>(3.0 * R[m * 7 + 3] * rdr);
      kd[1] = static_cast<float>(2.0 * R[m * 7 + 4] * rdr);
      kd[2] = static_cast<float>(R[m * 7 + 5] * rdr);
      kd[3] = 0.0f;
    }
  }
  if (!h->tab64) {
    if (devAlloc(h, &h->tab64, static_cast<size_t>(total)) != MISA_OK) return MISA_E_HIP;
    if (devAlloc(h, &h->tab32, static_cast<size_t>(total)) != MISA_OK) return MISA_E_HIP;
  } else if (h->t64.nr != sf.nr || h->t64.nrho != sf.nrho) {
    return fail(h, MISA_E_ARG, "potential grid changed; create a new handle");
  }
  HIPCHECK(h, hipMemcpy(h->tab64, h64.data(), h64.size() * 8, hipMemcpyHostToDevice));
  HIPCHECK(h, hipMemcpy(h->tab32, h32.data(), h32.size() * 4, hipMemcpyHostToDevice));
  Tables<double>& a64 = h->t64;
  a64.F = h->tab64;
  a64.rho = h->tab64 + 3 * sF;
  a64.z2 = h->tab64 + 3 * sF + 3 * sR;
  a64.sF = sF;
  a64.sR = sR;
  a64.nrho = sf.nrho;
  a64.nr = sf.nr;
  a64.rdrho = 1.0 / sf.drho;
  a64.rdr = 1.0 / sf.dr;
  a64.rhomax = (sf.nrho - 1) * sf.drho;
  a64.rhoV = h->tab64 + oRhoV;
  a64.pairF = h->tab64 + oPair;
  a64.rhoH = a64.rhoD = a64.pairZ = nullptr;  // FP32 passes only
  Tables<float>& a32 = h->t32;
  a32.F = h->tab32;
  a32.rho = h->tab32 + 3 * sF;
  a32.z2 = h->tab32 + 3 * sF + 3 * sR;
  a32.sF = sF;
  a32.sR = sR;
  a32.nrho = sf.nrho;
  a32.nr = sf.nr;
  a32.rdrho = static_cast<float>(1.0 / sf.drho);
  a32.rdr = static_cast<float>(1.0 / sf.dr);
  a32.rhomax = static_cast<float>(a64.rhomax);
  a32.rhoV = h->tab32 + oRhoV;
  a32.pairF = h->tab32 + oPair;
  a32.rhoH = h->tab32 + oRhoH;
  a32.rhoD = h->tab32 + oRhoD;
  a32.pairZ = h->tab32 + oPairZ;
  h->have_pot = true;
  return MISA_OK;
}

int misa_get_potential_table(const misa_handle* h, int kind, int idx, double* out, int64_t cap, int64_t* n) {
  if (!h || !n || !h->have_pot) return MISA_E_ARG;
  const auto& v = kind == 0 ? h->cF : (kind == 1 ? h->cR : h->cZ);
  if (idx < 0 || idx >= static_cast<int>(v.size())) return MISA_E_ARG;
  *n = static_cast<int64_t>(v[idx].size());
  if (out && *n <= cap) std::memcpy(out, v[idx].data(), v[idx].size() * 8);
  return MISA_OK;
}

static int hookCall(misa_handle* h, int which, void* atoms, double rc) {
  if (!h || !atoms) return MISA_E_ARG;
  if (!h->have_offsets || !h->have_pot) return fail(h, MISA_E_ARG, "offsets and potential must be set first");
  HIPCHECK(h, hipSetDevice(h->device));
  const size_t bytes = static_cast<size_t>(h->box.slots()) * sizeof(AtomElement);
  if (!h->aos_dev) HIPCHECK(h, hipMalloc(&h->aos_dev, bytes));
  HIPCHECK(h, hipMemcpyAsync(h->aos_dev, atoms, bytes, hipMemcpyHostToDevice, h->stream));
  launchHook(which, h->box, rc, h->full, h->t64, h->aos_dev, h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipMemcpyAsync(atoms, h->aos_dev, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return MISA_OK;
}

int misa_lat_hook_noalloc(misa_handle* h, int which, void* atoms, double rc) {
  if (!h || !atoms || which < 0 || which > 2) return MISA_E_ARG;
  if (!h->have_offsets || !h->have_pot) return fail(h, MISA_E_ARG, "offsets and potential must be set first");
  const size_t bytes = static_cast<size_t>(h->box.slots()) * sizeof(AtomElement);
  if (bytes > kScratchBytes)
    return fail(h, MISA_E_ARG, "box of " + std::to_string(h->box.slots()) + " slots (" + std::to_string(bytes) +
                                   " bytes) exceeds the " + std::to_string(kScratchBytes) +
                                   "-byte staging array of misa_lat_hook_noalloc");
  HIPCHECK(h, hipSetDevice(h->device));
  void* stage = nullptr;
  HIPCHECK(h, scratchOf(h->device, &stage));
  AtomElement* a = static_cast<AtomElement*>(stage);
  // the whole array is overwritten by the copy: nothing in it needs clearing
  HIPCHECK(h, hipMemcpyAsync(a, atoms, bytes, hipMemcpyHostToDevice, h->stream));
  launchHook(which, h->box, rc, h->full, h->t64, a, h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipMemcpyAsync(atoms, a, bytes, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return MISA_OK;
}

int misa_lat_rho(misa_handle* h, void* atoms, double rc) { return hookCall(h, 0, atoms, rc); }
int misa_lat_df(misa_handle* h, void* atoms, double rc) { return hookCall(h, 1, atoms, rc); }
int misa_lat_force(misa_handle* h, void* atoms, double rc) { return hookCall(h, 2, atoms, rc); }

int misa_upload_state(misa_handle* h, const void* atoms, const void* inter, int64_t n_inter) {
  if (!h || !atoms || n_inter < 0 || (n_inter > 0 && !inter)) return MISA_E_ARG;
  if (n_inter > h->s.in.cap) return fail(h, MISA_E_CAPACITY, "inter list larger than the inter-atom capacity");
  HIPCHECK(h, hipSetDevice(h->device));
  const AtomElement* a = static_cast<const AtomElement*>(atoms);
  const Box& b = h->box;
  // species present in the potential?
  auto checkType = [&](int t) -> bool {
    if (t > 2 || t < -1) return false;
    return !(t >= 0 && h->have_pot && !h->species_present[t]);
  };
  for (int64_t k = 0; k < b.n[2]; k++)
    for (int64_t j = 0; j < b.n[1]; j++)
      for (int64_t i = 0; i < 2 * b.n[0]; i++)
        if (!checkType(a[b.refIndex(i + 2 * b.g, j + b.g, k + b.g)].type))
          return fail(h, MISA_E_ARG, "atom type unknown or missing from the potential file");
  const AtomElement* ia = static_cast<const AtomElement*>(inter);
  for (int64_t i = 0; i < n_inter; i++)
    if (ia[i].type < 0 || !checkType(ia[i].type)) return fail(h, MISA_E_ARG, "bad inter-atom type");
  const size_t bytes = static_cast<size_t>(b.slots()) * sizeof(AtomElement);
  if (!h->aos_dev) HIPCHECK(h, hipMalloc(&h->aos_dev, bytes));
  HIPCHECK(h, hipMemcpyAsync(h->aos_dev, atoms, bytes, hipMemcpyHostToDevice, h->stream));
  h->s.cur = 0;
  launchUpload(h->s, h->aos_dev, b.sx, b.sy, h->stream);
  launchHaloAll(h->s, h->s.cur, h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipMemsetAsync(h->s.ctl, 0, sizeof(Ctl), h->stream));
  h->flags_seen = 0;
  h->t_fixed = 0.0;
  HIPCHECK(h, hipMemsetAsync(h->s.in.head, 0xff, static_cast<size_t>(2 * h->s.G.plane) * 4, h->stream));
  if (n_inter > 0) {
    if (h->interAosCap < n_inter) {
      if (h->interAos) HIPCHECK(h, hipFree(h->interAos));
      h->interAos = nullptr;
      HIPCHECK(h, hipMalloc(&h->interAos, static_cast<size_t>(h->s.in.cap) * sizeof(AtomElement)));
      h->interAosCap = h->s.in.cap;
    }
    HIPCHECK(h, hipMemcpyAsync(h->interAos, inter, n_inter * sizeof(AtomElement), hipMemcpyHostToDevice, h->stream));
  }
  launchInterUpload(h->s, h->interAos, n_inter, h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  h->state_ready = true;
  h->forces_ready = false;
  return MISA_OK;
}

int misa_download_state(misa_handle* h, void* atoms, void* inter, int64_t inter_cap, int64_t* n_inter) {
  if (!h || !atoms) return MISA_E_ARG;
  if (!h->state_ready) return fail(h, MISA_E_ARG, "no state uploaded");
  HIPCHECK(h, hipSetDevice(h->device));
  const Box& b = h->box;
  const bool mixed = h->precision == MISA_PREC_MIXED;
  const size_t bytes = static_cast<size_t>(b.slots()) * sizeof(AtomElement);
  if (!h->aos_dev) HIPCHECK(h, hipMalloc(&h->aos_dev, bytes));
  launchDownload(h->s, h->aos_dev, b.sx, b.sy, mixed, h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipMemcpyAsync(atoms, h->aos_dev, bytes, hipMemcpyDeviceToHost, h->stream));
  Ctl c;
  HIPCHECK(h, hipMemcpyAsync(&c, h->s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  if (n_inter) *n_inter = c.n_inter;
  if (c.n_inter > 0) {
    if (!inter || c.n_inter > inter_cap) return fail(h, MISA_E_CAPACITY, "inter buffer too small");
    if (h->interAosCap < c.n_inter) {
      if (h->interAos) HIPCHECK(h, hipFree(h->interAos));
      h->interAos = nullptr;
      HIPCHECK(h, hipMalloc(&h->interAos, static_cast<size_t>(h->s.in.cap) * sizeof(AtomElement)));
      h->interAosCap = h->s.in.cap;
    }
    launchInterDownload(h->s, h->interAos, mixed, h->stream);
    HIPCHECK(h, hipGetLastError());
    HIPCHECK(h, hipMemcpyAsync(inter, h->interAos, c.n_inter * sizeof(AtomElement), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
  }
  return MISA_OK;
}

// The handle's state, downloaded for the frontend's dumps.
static int downloadForDump(misa_handle* h, std::vector<AtomElement>& atoms, std::vector<AtomElement>& inter,
                           int64_t& n_inter, misa_geom& g) {
  atoms.resize(static_cast<size_t>(h->box.slots()));
  inter.resize(static_cast<size_t>(std::max<int64_t>(h->s.in.cap, 1)));
  const int rc = misa_download_state(h, atoms.data(), inter.data(), h->s.in.cap, &n_inter);
  if (rc != MISA_OK) return rc;
  for (int k = 0; k < 3; k++) g.phase_space[k] = h->box.n[k];
  g.lattice_const = h->box.a;
  g.cutoff_radius_factor = h->box.f;
  return MISA_OK;
}

int misa_dump_frame_state(misa_dump* d, misa_handle* h, uint64_t time_step) {
  if (!d || !h) return MISA_E_ARG;
  std::vector<AtomElement> atoms, inter;
  int64_t n = 0;
  misa_geom g;
  const int rc = downloadForDump(h, atoms, inter, n, g);
  if (rc != MISA_OK) return rc;
  return misa_dump_frame(d, &g, atoms.data(), inter.data(), n, time_step);
}

int misa_dump_text_state(misa_handle* h, const char* path) {
  if (!h || !path) return MISA_E_ARG;
  std::vector<AtomElement> atoms, inter;
  int64_t n = 0;
  misa_geom g;
  const int rc = downloadForDump(h, atoms, inter, n, g);
  if (rc != MISA_OK) return rc;
  return misa_dump_text(path, &g, atoms.data(), inter.data(), n);
}

int misa_build_world(misa_handle* h, uint32_t seed, double tset, const int32_t ratio[3], uint32_t alloy_seed) {
  if (!h || !ratio) return MISA_E_ARG;
  if (ratio[0] < 0 || ratio[1] < 0 || ratio[2] < 0 || ratio[0] + ratio[1] + ratio[2] <= 0)
    return fail(h, MISA_E_ARG, "alloy ratio must be non-negative with a positive sum");
  std::vector<AtomElement> atoms;
  buildWorld(h->box, seed, tset, ratio, alloy_seed, atoms);
  return misa_upload_state(h, atoms.data(), nullptr, 0);
}

int misa_set_precision(misa_handle* h, int mode) {
  if (!h) return MISA_E_ARG;
  h->gen++;
  if (mode == MISA_PREC_FP64) {
    h->precision = mode;
    h->acc = 0;
  } else if (mode == MISA_PREC_MIXED) {
    h->precision = mode;
    h->acc = 0;
  } else if (mode == MISA_PREC_MIXED_KAHAN) {
    h->precision = MISA_PREC_MIXED;
    h->acc = 1;
  } else if (mode == MISA_PREC_MIXED_F32SUM) {
    h->precision = MISA_PREC_MIXED;
    h->acc = 2;
  } else {
    return fail(h, MISA_E_ARG, "unknown precision mode");
  }
  h->forces_ready = false;
  return MISA_OK;
}

int misa_set_timestep(misa_handle* h, double dt) {
  if (!h || !(dt > 0)) return MISA_E_ARG;
  h->gen++;
  h->dt = dt;
  return MISA_OK;
}

int misa_set_adaptive_dt(misa_handle* h, double dt_max, double max_disp) {
  if (!h || !(dt_max > 0) || max_disp < 0) return MISA_E_ARG;
  h->gen++;
  h->dt_max = dt_max;
  h->max_disp = max_disp;
  return MISA_OK;
}

int misa_set_neighbour_skin(misa_handle* h, double skin) {
  if (!h || !(skin >= 0.0) || skin > 1.0) return MISA_E_ARG;
  h->s.skin = skin;
  h->gen++;
  return MISA_OK;
}

int misa_sim_time(misa_handle* h, double* t_ps, double* dt_last) {
  if (!h || !t_ps) return MISA_E_ARG;
  HIPCHECK(h, hipSetDevice(h->device));
  Ctl c;
  HIPCHECK(h, hipMemcpy(&c, h->s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  *t_ps = h->t_fixed + c.t_sim;
  if (dt_last) *dt_last = h->max_disp > 0.0 ? c.dt : h->dt;
  return MISA_OK;
}

int misa_set_range_check(misa_handle* h, int32_t on) {
  if (!h || (on != 0 && on != 1)) return MISA_E_ARG;
  h->range_check = on != 0;
  return MISA_OK;
}

int misa_device_flags(const misa_handle* h, int64_t* flags) {
  if (!h || !flags) return MISA_E_ARG;
  *flags = h->flags_seen;
  return MISA_OK;
}

int misa_set_inter_capacity(misa_handle* h, int64_t cap) {
  if (!h || cap < 1 || cap >= (int64_t(1) << 21)) return MISA_E_ARG;
  if (h->state_ready) return fail(h, MISA_E_ARG, "set the inter-atom capacity before uploading a state");
  HIPCHECK(h, hipSetDevice(h->device));
  return allocInter(h, cap);
}

// clearForce + computeEam (atom.cpp:86-149) on the resident state, after the
// inter-atom migration, ghost images and index (mode 1 of the decide kernel).
static void forcePass(misa_handle* h, int lattice_mode, int inter_mode, bool store_rho, const VerletParams& vp) {
  const bool mixed = h->precision == MISA_PREC_MIXED;
  {
    Timer t(h, 2);
    launchInterRho(h->s, h->t64, h->t32, mixed, h->stream);
  }
  {
    Timer t(h, 0);
    launchRhoDf(h->s, h->lat, h->t64, h->t32, mixed, h->acc, store_rho, h->stream);
    h->mask_gen = h->gen;
  }
  {
    Timer t(h, 2);
    launchInterForce(h->s, h->t64, h->t32, mixed, inter_mode, vp, h->stream);
  }
  {
    Timer t(h, 1);
    h->force_pass = launchForce(h->s, h->lat, h->t64, h->t32, mixed, h->acc, lattice_mode, vp, h->stream);
  }
}

int misa_compute_forces(misa_handle* h) {
  if (!h) return MISA_E_ARG;
  if (!h->state_ready || !h->have_pot || !h->have_offsets)
    return fail(h, MISA_E_ARG, "state, potential and offsets required");
  HIPCHECK(h, hipSetDevice(h->device));
  const VerletParams vp = verletOf(h->dt);
  const bool prof = h->profile;
  h->profile = false;
  const bool mixed = h->precision == MISA_PREC_MIXED;
  launchPackRec(h->s, h->s.cur, !mixed, h->stream);
  launchFarScan(h->s, h->stream);
  launchDecide(h->s, 1, mixed, h->stream);
  forcePass(h, 2, 2, true, vp);
  h->profile = prof;
  HIPCHECK(h, hipGetLastError());
  const int rc = syncAndCheck(h);
  if (rc != MISA_OK) return rc;
  h->forces_ready = true;
  return MISA_OK;
}

}  // extern "C"

namespace {

// The loop body of simulation::simulate (simulation.cpp:164-194) as launch
// sequences on the handle's stream, without synchronisation:
//   prologue: (adaptive: step length) firststep of step 0, decide
//   body(i):  forces of step i and its secondstep; unless last, the
//             firststep of step i+1 (fused into the force pass with a fixed
//             step) and its decide
struct StepPlan {
  bool adaptive, mixed;
  VerletParams vp;
};

StepPlan planOf(misa_handle* h) {
  StepPlan p;
  p.adaptive = h->max_disp > 0.0;
  p.mixed = h->precision == MISA_PREC_MIXED;
  p.vp = verletOf(h->dt);
  if (p.adaptive) p.vp.dtp = &h->s.ctl->dt;
  return p;
}

// Any setter change (offsets, potential, shape, precision, ...) may leave
// the density pass's skin masks stale: the next decide scans in full.
void enqRescanIfChanged(misa_handle* h) {
  if (h->skin_gen == h->gen) return;
  launchRescan(h->s, h->stream);
  h->skin_gen = h->gen;
}

void enqFirstAndDecide(misa_handle* h, const StepPlan& p, bool lattice, bool vmax_done = false) {
  {
    Timer t(h, 2);
    if (p.adaptive) launchAdaptiveDt(h->s, h->dt_max, h->max_disp, !vmax_done, h->stream);
    if (lattice) launchFirst(h->s, p.mixed, p.vp, h->stream);
    launchInterFirst(h->s, p.vp, h->stream);
  }
  h->s.cur = 1 - h->s.cur;
  {
    Timer t(h, 2);
    launchDecide(h->s, 0, p.mixed, h->stream);
  }
}

void enqBody(misa_handle* h, const StepPlan& p, bool last) {
  const bool fuse = !last && !p.adaptive;
  VerletParams vp = p.vp;
  // adaptive: the secondstep pass reduces the lattice sites' max |v|^2 itself
  if (p.adaptive && !last) vp.vmax = &h->s.ctl->vmax2;
  forcePass(h, fuse ? 1 : 0, 0, last, vp);
  if (!last) enqFirstAndDecide(h, p, !fuse, p.adaptive);
}

// Error flags of the step kernels; timings of the profiled launches.
int finishSteps(misa_handle* h, int64_t nsteps, hipEvent_t e0, hipEvent_t e1) {
  if (h->max_disp <= 0.0) h->t_fixed += static_cast<double>(nsteps) * h->dt;
  const int rc = syncAndCheck(h);
  float ms = 0;
  if (e0 && e1) {
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  h->timings[0] = h->timings[1] = h->timings[2] = 0;
  h->launches[0] = h->launches[1] = h->launches[2] = 0;
  for (size_t k = 0; k < h->event_cls.size(); k++) {
    float t = 0;
    (void)hipEventElapsedTime(&t, h->events[2 * k], h->events[2 * k + 1]);
    h->timings[h->event_cls[k]] += t;
    h->launches[h->event_cls[k]]++;
  }
  h->timings[3] = ms;
  if (rc != MISA_OK) h->forces_ready = false;
  return rc;
}

// Re-cut the XCD z-slabs so that each holds about an eighth of the work the
// LDS force pass measured per plane (a cascade's disordered core costs
// several times a perfect lattice's planes, and one slab holding it keeps
// the other XCDs waiting).  Each slab keeps at least one plane and at most
// the planes the wave-mask arrays are sized for; a cut that lowers the
// busiest slab's work by less than 3 % is not taken.  A new cut bumps the
// generation, so the next step scans the density pass's offsets in full.
int balanceSlabs(misa_handle* h) {
  Geom& G = h->s.G;
  Walk& W = G.W;
  const int n2 = G.n2;
  if (n2 < 16) return MISA_OK;
  std::vector<unsigned long long> w(static_cast<size_t>(n2));
  HIPCHECK(h, hipMemcpy(w.data(), h->s.zwork, w.size() * 8, hipMemcpyDeviceToHost));
  std::vector<double> P(static_cast<size_t>(n2) + 1, 0.0);
  for (int z = 0; z < n2; z++) P[z + 1] = P[z] + static_cast<double>(w[z]);
  if (!(P[n2] > 0.0)) return MISA_OK;
  const int64_t row = static_cast<int64_t>(G.n1) * G.n0;
  const int zmax = static_cast<int>(std::min<int64_t>(n2, h->bpx_cap * 256 / row));
  int zl[9];
  zl[0] = 0;
  zl[8] = n2;
  for (int c = 1; c < 8; c++) {
    const double target = P[n2] * c / 8.0;
    const int lo = std::max(zl[c - 1] + 1, n2 - (8 - c) * zmax);
    const int hi = std::min(zl[c - 1] + zmax, n2 - (8 - c));
    int z = lo;
    while (z < hi && P[z + 1] <= target) z++;
    if (z < hi && P[z + 1] - target < target - P[z]) z++;
    zl[c] = z;
  }
  auto busiest = [&](const int* cut) {
    double m = 0.0;
    for (int c = 0; c < 8; c++) m = std::max(m, P[cut[c + 1]] - P[cut[c]]);
    return m;
  };
  if (busiest(zl) > 0.97 * busiest(W.zlo)) return MISA_OK;
  setSlabs(G, zl);
  h->gen++;
  return MISA_OK;
}

void clearEvents(misa_handle* h) {
  for (auto e : h->events) (void)hipEventDestroy(e);
  h->events.clear();
  h->event_cls.clear();
}

}  // namespace

extern "C" {

int misa_step(misa_handle* h, int64_t nsteps) {
  if (!h || nsteps < 0) return MISA_E_ARG;
  if (!h->forces_ready) return fail(h, MISA_E_ARG, "misa_compute_forces must run before misa_step");
  if (nsteps == 0) return MISA_OK;
  HIPCHECK(h, hipSetDevice(h->device));
  const StepPlan p = planOf(h);
  clearEvents(h);
  hipEvent_t e0, e1;
  HIPCHECK(h, hipEventCreate(&e0));
  HIPCHECK(h, hipEventCreate(&e1));
  HIPCHECK(h, hipEventRecord(e0, h->stream));
  enqRescanIfChanged(h);
  enqFirstAndDecide(h, p, true);
  // the first force pass of the call measures its work per plane
  const bool measure = h->balance && p.mixed && h->s.ncu >= 8;
  for (int64_t i = 0; i < nsteps; i++) {
    if (i == 0 && measure) {
      HIPCHECK(h, hipMemsetAsync(h->s.zwork, 0, static_cast<size_t>(h->s.G.n2) * 8, h->stream));
      h->s.zmeasure = true;
    }
    enqBody(h, p, i == nsteps - 1);
    h->s.zmeasure = false;
  }
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipEventRecord(e1, h->stream));
  const int rc = finishSteps(h, nsteps, e0, e1);
  if (rc != MISA_OK || !measure || h->force_pass != MISA_FORCE_PASS_MIXED_LDS) return rc;
  return balanceSlabs(h);
}

int misa_setv(misa_handle* h, const int32_t lat[4], const double dir[3], double energy) {
  if (!h || !lat || !dir) return MISA_E_ARG;
  if (!h->state_ready) return fail(h, MISA_E_ARG, "no state");
  const Box& b = h->box;
  // atom::setv only acts when the site is in this sub-box (atom.cpp:477-482)
  if (!(lat[0] >= 0 && lat[0] < b.n[0] && lat[1] >= 0 && lat[1] < b.n[1] && lat[2] >= 0 && lat[2] < b.n[2]) ||
      lat[3] < 0 || lat[3] > 1)
    return MISA_OK;
  HIPCHECK(h, hipSetDevice(h->device));
  const Geom& G = h->s.G;
  const int64_t slot = lat[3] * G.plane + (static_cast<int64_t>(lat[2] + G.g) * G.NY + (lat[1] + G.g)) * G.NX +
                       (lat[0] + G.g);
  int8_t t;
  HIPCHECK(h, hipMemcpy(&t, h->s.type + slot, 1, hipMemcpyDeviceToHost));
  const double m = t < 0 ? 0.0 : typeMass(t & 3);
  const double v_ = std::sqrt(2 * energy / m / kMvv2e);
  const double d_ = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
  launchSetv(h->s, slot, v_ * dir[0] / d_, v_ * dir[1] / d_, v_ * dir[2] / d_, h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return MISA_OK;
}

int misa_collision_step(misa_handle* h, const int32_t lat[4], const double dir[3], double energy) {
  int rc = misa_setv(h, lat, dir, energy);
  if (rc != MISA_OK) return rc;
  return misa_compute_forces(h);
}

int misa_thermo_compute(misa_handle* h, misa_thermo* out) {
  if (!h || !out) return MISA_E_ARG;
  if (!h->state_ready) return fail(h, MISA_E_ARG, "no state");
  HIPCHECK(h, hipSetDevice(h->device));
  std::memset(out, 0, sizeof(*out));
  const int64_t nb = launchThermo(h->s, h->stream);
  HIPCHECK(h, hipGetLastError());
  std::vector<double> p(static_cast<size_t>(7 * nb));
  HIPCHECK(h, hipMemcpyAsync(p.data(), h->s.partials, p.size() * 8, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  double vmax2 = 0;
  for (int64_t b = 0; b < nb; b++) {
    out->mvv += p[7 * b + 0];
    out->n_lattice += static_cast<int64_t>(p[7 * b + 1]);
    out->n_vacancy += static_cast<int64_t>(p[7 * b + 2]);
    out->force_sum[0] += p[7 * b + 3];
    out->force_sum[1] += p[7 * b + 4];
    out->force_sum[2] += p[7 * b + 5];
    vmax2 = std::max(vmax2, p[7 * b + 6]);
  }
  {
    const int64_t ni = launchInterThermo(h->s, h->stream);
    HIPCHECK(h, hipGetLastError());
    std::vector<double> q(static_cast<size_t>(5 * ni));
    Ctl c;
    HIPCHECK(h, hipMemcpyAsync(q.data(), h->s.partials, q.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipMemcpyAsync(&c, h->s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    for (int64_t b = 0; b < ni; b++) {
      out->mvv += q[5 * b + 0];
      out->force_sum[0] += q[5 * b + 1];
      out->force_sum[1] += q[5 * b + 2];
      out->force_sum[2] += q[5 * b + 3];
      vmax2 = std::max(vmax2, q[5 * b + 4]);
    }
    out->n_inter = c.n_inter;
  }
  out->max_speed = std::sqrt(vmax2);
  out->kinetic = out->mvv / 2;
  const uint64_t n = static_cast<uint64_t>(2 * h->box.n[0] * h->box.n[1] * h->box.n[2]);
  const uint64_t dof = 3 * n - 3;
  out->temperature = out->mvv * kMvv2e / (dof * kBoltz);
  if (h->have_pot && h->have_offsets && h->forces_ready) {
    // the energy pass visits the skin masks' offsets only while they fit
    // the current slab cut (a re-cut since the last density pass moves sites
    // between wavefronts) and hold a skin
    const bool all = !(h->s.skin > 0.0) || h->mask_gen != h->gen;
    const int64_t ne = launchEnergy(h->s, h->lat, h->t64, all, h->stream);
    HIPCHECK(h, hipGetLastError());
    std::vector<double> q(static_cast<size_t>(3 * ne));
    HIPCHECK(h, hipMemcpyAsync(q.data(), h->s.partials, q.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    for (int64_t b = 0; b < ne; b++) {
      out->e_embed += q[3 * b + 0];
      out->e_pair += q[3 * b + 1];
      out->virial += q[3 * b + 2];
    }
    const int64_t ni = launchInterEnergy(h->s, h->t64, h->stream);
    HIPCHECK(h, hipGetLastError());
    std::vector<double> r(static_cast<size_t>(3 * ni));
    HIPCHECK(h, hipMemcpyAsync(r.data(), h->s.partials, r.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(h, hipStreamSynchronize(h->stream));
    for (int64_t b = 0; b < ni; b++) {
      out->e_embed += r[3 * b + 0];
      out->e_pair += r[3 * b + 1];
      out->virial += r[3 * b + 2];
    }
    out->e_potential = out->e_embed + out->e_pair;
  }
  return MISA_OK;
}

int misa_rescale(misa_handle* h, double tset) {
  misa_thermo t;
  int rc = misa_thermo_compute(h, &t);
  if (rc != MISA_OK) return rc;
  if (!(t.temperature > 0)) return fail(h, MISA_E_ARG, "zero temperature cannot be rescaled");
  launchScaleV(h->s, std::sqrt(tset / t.temperature), h->stream);
  launchInterScaleV(h->s, std::sqrt(tset / t.temperature), h->stream);
  HIPCHECK(h, hipGetLastError());
  HIPCHECK(h, hipStreamSynchronize(h->stream));
  return MISA_OK;
}

int misa_last_timings(const misa_handle* h, double out[4]) {
  if (!h || !out) return MISA_E_ARG;
  for (int k = 0; k < 4; k++) out[k] = h->timings[k];
  return MISA_OK;
}

// Per-kernel event timing of misa_step (HIP events on the handle's stream).
int misa_set_profiling(misa_handle* h, int on) {
  if (!h) return MISA_E_ARG;
  h->profile = on != 0;
  return MISA_OK;
}

int misa_last_launch_counts(const misa_handle* h, int64_t out[3]) {
  if (!h || !out) return MISA_E_ARG;
  for (int k = 0; k < 3; k++) out[k] = h->launches[k];
  return MISA_OK;
}

// ---- host-only set-up (no device access) ----

// ---- batches of independent boxes: captured step graphs ----

}  // extern "C"

struct misa_batch {
  std::vector<misa_handle*> hs;
  int device = 0;
  std::string err;
  hipStream_t origin = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> done;  // per box: its stream reached the graph
  hipEvent_t after = nullptr;    // the origin stream finished the replays
  hipGraphExec_t exec = nullptr;
  std::vector<uint64_t> key;  // per box: generation and position buffer the graph was captured for
  int64_t replays = 0;
  // Boxes alike in precision, fixed step length, potential and neighbour
  // offsets run every pass of a step as one launch for all boxes, from box
  // records in device memory (one allocation: records, block starts, decide
  // arguments); the graph then holds these launches instead of a branch per box.
  bool fused = false;
  void* fusedBlock = nullptr;
  site::BatchBox* boxes = nullptr;
  uint32_t* rhoStart = nullptr;
  uint32_t* vstart = nullptr;
  void* decideArgs = nullptr;
  unsigned rhoBlocks = 0;
};

namespace {

int bfail(misa_batch* b, int code, const std::string& msg) {
  b->err = msg;
  return code;
}

#define BHIP(b, call)                                                                               \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return bfail(b, MISA_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

std::vector<uint64_t> batchKey(const misa_batch* b) {
  std::vector<uint64_t> k;
  for (const misa_handle* h : b->hs) {
    k.push_back(h->gen);
    k.push_back(static_cast<uint64_t>(h->s.cur));
  }
  return k;
}

// Can the boxes' passes run as one launch each?  Same mixed precision and
// accumulation, fixed step length, identical potential tables and pruned
// offsets (the lattice vectors and packed deltas), and the LDS force pass
// available.
bool fusable(const misa_batch* b) {
  const misa_handle* h0 = b->hs[0];
  if (h0->precision != MISA_PREC_MIXED || h0->s.ncu < 8) return false;
  for (const misa_handle* h : b->hs) {
    if (h->precision != MISA_PREC_MIXED || h->acc != h0->acc || h->max_disp > 0.0 || !h->have_pot ||
        !h->have_offsets)
      return false;
    if (h->setfl.nr != h0->setfl.nr || h->setfl.nrho != h0->setfl.nrho || h->setfl.dr != h0->setfl.dr ||
        h->setfl.drho != h0->setfl.drho || h->cF != h0->cF || h->cR != h0->cR || h->cZ != h0->cZ)
      return false;
    for (int t = 0; t < 3; t++)
      if (h->species_present[t] != h0->species_present[t]) return false;
    if (h->lat.n[0] != h0->lat.n[0] || h->lat.n[1] != h0->lat.n[1] ||
        std::memcmp(h->dlHost, h0->dlHost, sizeof(h->dlHost)) != 0)
      return false;
  }
  return true;
}

// The box records of a fused batch, with the buffers as they stand now.
int buildFused(misa_batch* b) {
  const int n = static_cast<int>(b->hs.size());
  std::vector<site::BatchBox> hb(n);
  std::vector<uint32_t> rs(n + 1, 0), vs(n + 1, 0);
  for (int i = 0; i < n; i++) {
    misa_handle* h = b->hs[i];
    const DevState& s = h->s;
    site::BatchBox& B = hb[i];
    std::memset(&B, 0, sizeof(B));
    B.G = s.G;
    B.O = h->lat;
    B.t32 = h->t32;
    B.iv = interView(s, true);
    B.in = s.in;
    for (int c = 0; c < 2; c++)
      for (int k = 0; k < 3; k++) B.x[c][k] = s.x[c][k];
    for (int k = 0; k < 3; k++) {
      B.v[k] = s.v[k];
      B.f[k] = s.f[k];
    }
    B.type = s.type;
    B.id = s.id;
    B.rec[0] = s.rec[0];
    B.rec[1] = s.rec[1];
    B.dl = s.dl;
    B.wmask = s.wmask;
    B.smask = s.smask;
    const double rcs = std::sqrt(s.G.rc2) + s.skin;
    B.rcs2 = static_cast<float>(rcs * rcs);
    B.skin = s.skin;
    B.vp = verletOf(h->dt);
    B.st = site::stageOf(s);
    B.ctl = s.ctl;
    B.cur = s.cur;
    rs[i + 1] = rs[i] + site::siteGrid(s.G);
    vs[i + 1] = vs[i] + static_cast<uint32_t>(2 * s.G.W.bpx);
  }
  const size_t da = decideArgsBytes();
  auto up = [](size_t v) { return (v + 255) & ~size_t(255); };
  const size_t oBox = 0, oRs = up(n * sizeof(site::BatchBox)), oVs = oRs + up((n + 1) * 4),
               oDa = oVs + up((n + 1) * 4), total = oDa + up(2 * n * da);
  std::vector<char> host(total, 0);
  std::memcpy(host.data() + oBox, hb.data(), n * sizeof(site::BatchBox));
  std::memcpy(host.data() + oRs, rs.data(), (n + 1) * 4);
  std::memcpy(host.data() + oVs, vs.data(), (n + 1) * 4);
  for (int i = 0; i < n; i++)
    for (int f = 0; f < 2; f++) fillDecideArgs(hb[i], hb[i].cur ^ f, host.data() + oDa + (2 * i + f) * da);
  if (b->fusedBlock) {
    BHIP(b, hipStreamSynchronize(b->origin));
    BHIP(b, hipFree(b->fusedBlock));
    b->fusedBlock = nullptr;
  }
  char* base = nullptr;
  BHIP(b, hipMalloc(&base, total));
  b->fusedBlock = base;
  BHIP(b, hipMemcpy(base, host.data(), total, hipMemcpyHostToDevice));
  b->boxes = reinterpret_cast<site::BatchBox*>(base + oBox);
  b->rhoStart = reinterpret_cast<uint32_t*>(base + oRs);
  b->vstart = reinterpret_cast<uint32_t*>(base + oVs);
  b->decideArgs = base + oDa;
  b->rhoBlocks = rs[n];
  return MISA_OK;
}

// Two loop bodies of every box of a fused batch as one launch per pass:
// the inter atoms' density, the lattice density pass, the inter atoms'
// forces, the LDS force pass with the next half step and decide staging,
// the inter atoms' first half step and the decide (on the other buffer).
int captureFused(misa_batch* b) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  const misa_handle* h0 = b->hs[0];
  const int n = static_cast<int>(b->hs.size());
  BHIP(b, hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
  bool lds = true;
  for (int f = 0; f < 2; f++) {
    launchInterBatch(0, b->boxes, n, f, nullptr, b->origin);
    launchRhoDfBatch(b->boxes, b->rhoStart, n, b->rhoBlocks, h0->acc, f, b->origin);
    launchInterBatch(1, b->boxes, n, f, nullptr, b->origin);
    lds = lds && launchForceBatch(b->boxes, b->vstart, n, h0->s.dl, h0->t32.nr, h0->s.ncu, h0->acc, f, b->origin);
    launchInterBatch(2, b->boxes, n, f, nullptr, b->origin);
    launchInterBatch(3, b->boxes, n, f ^ 1, b->decideArgs, b->origin);
  }
  const hipError_t e0 = hipGetLastError();
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->origin, &g);
  if (e0 != hipSuccess) return bfail(b, MISA_E_HIP, std::string("capturing the fused step: ") + hipGetErrorString(e0));
  if (e2 != hipSuccess) return bfail(b, MISA_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
  if (!lds) {
    (void)hipGraphDestroy(g);
    return bfail(b, MISA_E_HIP, "the device refused the LDS size of the fused force pass");
  }
  const hipError_t e3 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) return bfail(b, MISA_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
  b->key = batchKey(b);
  return MISA_OK;
}

// Capture two loop bodies of every box, each box a branch of its own stream
// forked from and joined back into the origin stream.
int captureBatch(misa_batch* b, const std::vector<StepPlan>& plans) {
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  BHIP(b, hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed));
  hipError_t e = hipEventRecord(b->fork, b->origin);
  for (size_t i = 0; i < b->hs.size() && e == hipSuccess; i++) {
    misa_handle* h = b->hs[i];
    e = hipStreamWaitEvent(h->stream, b->fork, 0);
    if (e != hipSuccess) break;
    enqBody(h, plans[i], false);
    enqBody(h, plans[i], false);
    e = hipEventRecord(b->done[i], h->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->origin, b->done[i], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(b->origin, &g);
  if (e != hipSuccess) return bfail(b, MISA_E_HIP, std::string("capturing the step graph: ") + hipGetErrorString(e));
  if (e2 != hipSuccess) return bfail(b, MISA_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
  const hipError_t e3 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e3 != hipSuccess) return bfail(b, MISA_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
  b->key = batchKey(b);
  return MISA_OK;
}

}  // namespace

extern "C" {

int misa_batch_create(misa_handle* const* handles, int n, misa_batch** out) {
  if (!handles || n < 1 || !out) return MISA_E_ARG;
  *out = nullptr;
  misa_batch* b = new misa_batch();
  b->device = handles[0]->device;
  for (int i = 0; i < n; i++) {
    if (!handles[i] || handles[i]->device != b->device) {
      delete b;
      return MISA_E_ARG;
    }
    for (int j = 0; j < i; j++)
      if (handles[j] == handles[i]) {
        delete b;
        return MISA_E_ARG;
      }
    b->hs.push_back(handles[i]);
  }
  b->done.assign(n, nullptr);
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->after, hipEventDisableTiming);
  for (int i = 0; i < n && e == hipSuccess; i++) e = hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    misa_batch_destroy(b);
    return MISA_E_HIP;
  }
  *out = b;
  return MISA_OK;
}

int misa_batch_destroy(misa_batch* b) {
  if (!b) return MISA_E_ARG;
  (void)hipSetDevice(b->device);
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->fusedBlock) (void)hipFree(b->fusedBlock);
  for (hipEvent_t e : b->done)
    if (e) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->after) (void)hipEventDestroy(b->after);
  if (b->origin) (void)hipStreamDestroy(b->origin);
  delete b;
  return MISA_OK;
}

const char* misa_batch_last_error(const misa_batch* b) { return b ? b->err.c_str() : "null batch"; }

int misa_batch_step(misa_batch* b, int64_t nsteps, int64_t* graph_replays) {
  if (!b || nsteps < 0) return MISA_E_ARG;
  for (misa_handle* h : b->hs)
    if (!h->forces_ready) return bfail(b, MISA_E_ARG, "misa_compute_forces must run on every box before stepping");
  if (nsteps == 0) return MISA_OK;
  BHIP(b, hipSetDevice(b->device));
  const size_t n = b->hs.size();
  std::vector<StepPlan> plans;
  std::vector<bool> prof(n);
  for (size_t i = 0; i < n; i++) {
    misa_handle* h = b->hs[i];
    prof[i] = h->profile;
    h->profile = false;  // no per-launch events inside graphs
    clearEvents(h);
    plans.push_back(planOf(h));
  }
  // prologue of every box on its own stream
  for (size_t i = 0; i < n; i++) {
    enqRescanIfChanged(b->hs[i]);
    enqFirstAndDecide(b->hs[i], plans[i], true);
  }
  const int64_t pairs = (nsteps - 1) / 2;
  int rc = MISA_OK;
  if (pairs > 0) {
    if (!b->exec || b->key != batchKey(b)) {
      b->fused = fusable(b);
      if (b->fused) {
        rc = buildFused(b);
        if (rc == MISA_OK) rc = captureFused(b);
      } else {
        rc = captureBatch(b, plans);
      }
    }
    if (rc == MISA_OK) {
      for (size_t i = 0; i < n && rc == MISA_OK; i++) {
        if (hipEventRecord(b->done[i], b->hs[i]->stream) != hipSuccess ||
            hipStreamWaitEvent(b->origin, b->done[i], 0) != hipSuccess)
          rc = bfail(b, MISA_E_HIP, "ordering the graph after the prologue failed");
      }
      for (int64_t r = 0; r < pairs && rc == MISA_OK; r++)
        if (hipGraphLaunch(b->exec, b->origin) != hipSuccess) rc = bfail(b, MISA_E_HIP, "hipGraphLaunch failed");
      if (rc == MISA_OK && hipEventRecord(b->after, b->origin) != hipSuccess)
        rc = bfail(b, MISA_E_HIP, "hipEventRecord failed");
      for (size_t i = 0; i < n && rc == MISA_OK; i++)
        if (hipStreamWaitEvent(b->hs[i]->stream, b->after, 0) != hipSuccess)
          rc = bfail(b, MISA_E_HIP, "ordering the epilogue after the graph failed");
      b->replays += pairs;
    }
  }
  // the steps the graph replays did not cover, then the last body
  const int64_t eager = (nsteps - 1) - 2 * (rc == MISA_OK ? pairs : 0);
  for (size_t i = 0; i < n; i++) {
    misa_handle* h = b->hs[i];
    for (int64_t k = 0; k < eager; k++) enqBody(h, plans[i], false);
    enqBody(h, plans[i], true);
  }
  for (size_t i = 0; i < n; i++) {
    misa_handle* h = b->hs[i];
    const int r = finishSteps(h, nsteps, nullptr, nullptr);
    h->profile = prof[i];
    if (r != MISA_OK && rc == MISA_OK) rc = bfail(b, r, "box " + std::to_string(i) + ": " + h->err);
  }
  if (graph_replays) *graph_replays = b->replays;
  return rc;
}

// Launch shape of the site kernels: every y-tile height that divides the
// box (up to 32 rows) and, in mixed precision, the gathers each wave keeps in
// flight in the density pass (4 or 8) and in the force pass (2 or 4).  Each
// candidate runs `steps` steps (real steps of the simulation); the fastest
// is kept.  Only the assignment of sites to threads and the grouping of the
// loads change, not the arithmetic of a site.
int misa_autotune(misa_handle* h, int64_t steps, int32_t* best_ty, double* best_ms) {
  if (!h || steps < 1) return MISA_E_ARG;
  if (!h->forces_ready) return fail(h, MISA_E_ARG, "misa_compute_forces must run before misa_autotune");
  HIPCHECK(h, hipSetDevice(h->device));
  const bool prof = h->profile;
  h->profile = false;
  struct Shape {
    int ty, uRho, uForce;
  };
  std::vector<Shape> cands;
  const bool mixed = h->precision == MISA_PREC_MIXED;
  for (int ty = 1; ty <= std::min(32, h->s.G.n1); ty++) {
    if (h->s.G.n1 % ty) continue;
    if (!mixed) {
      cands.push_back({ty, 4, 4});
      continue;
    }
    for (int ur : {4, 8})
      for (int uf : {2, 4}) cands.push_back({ty, ur, uf});
  }
  Shape best{h->s.G.W.ty, h->s.uRho, h->s.uForce};
  double bestMs = 1e300;
  for (const Shape& c : cands) {
    h->s.G.W.ty = c.ty;
    h->s.uRho = c.uRho;
    h->s.uForce = c.uForce;
    h->gen++;
    const int rc = misa_step(h, steps);
    if (rc != MISA_OK) {
      h->profile = prof;
      return rc;
    }
    const double ms = h->timings[3] / static_cast<double>(steps);
    if (ms < bestMs) {
      bestMs = ms;
      best = c;
    }
  }
  h->s.G.W.ty = best.ty;
  h->s.uRho = best.uRho;
  h->s.uForce = best.uForce;
  h->gen++;
  h->profile = prof;
  if (best_ty) *best_ty = best.ty;
  if (best_ms) *best_ms = bestMs;
  return MISA_OK;
}

int misa_set_launch_shape(misa_handle* h, int32_t ty, int32_t gathers_rho, int32_t gathers_force) {
  if (!h) return MISA_E_ARG;
  if (ty < 1 || h->s.G.n1 % ty || (gathers_rho != 4 && gathers_rho != 8) ||
      (gathers_force != 2 && gathers_force != 4))
    return fail(h, MISA_E_ARG, "y-tile must divide the box; gathers 4|8 (density), 2|4 (force)");
  h->s.G.W.ty = ty;
  h->s.uRho = gathers_rho;
  h->s.uForce = gathers_force;
  h->gen++;
  return MISA_OK;
}

static_assert(MISA_LDS_AUTO == kLdsAuto && MISA_LDS_ON == kLdsOn && MISA_LDS_OFF == kLdsOff, "LDS modes");
int misa_set_lds_tables(misa_handle* h, int32_t mode) {
  if (!h) return MISA_E_ARG;
  if (mode != MISA_LDS_AUTO && mode != MISA_LDS_ON && mode != MISA_LDS_OFF)
    return fail(h, MISA_E_ARG, "LDS table mode must be MISA_LDS_AUTO, MISA_LDS_ON or MISA_LDS_OFF");
  h->s.lds = mode;
  h->gen++;
  return MISA_OK;
}

int misa_set_xcd_balance(misa_handle* h, int32_t on) {
  if (!h || (on != 0 && on != 1)) return MISA_E_ARG;
  h->balance = on != 0;
  if (!h->balance) {
    evenSlabs(h->s.G);
    h->gen++;
  }
  return MISA_OK;
}

int misa_set_xcd_slabs(misa_handle* h, const int32_t zlo[9]) {
  if (!h || !zlo) return MISA_E_ARG;
  const Geom& G = h->s.G;
  const int64_t row = static_cast<int64_t>(G.n1) * G.n0;
  bool ok = zlo[0] == 0 && zlo[8] == G.n2;
  for (int c = 0; c < 8 && ok; c++)
    ok = zlo[c + 1] >= zlo[c] && ((zlo[c + 1] - zlo[c]) * row + 255) / 256 <= h->bpx_cap;
  if (!ok)
    return fail(h, MISA_E_ARG,
                "slab bounds must run from 0 to the box's planes, non-decreasing, each slab at most twice an "
                "eighth of the box");
  int cut[9];
  for (int c = 0; c <= 8; c++) cut[c] = zlo[c];
  setSlabs(h->s.G, cut);
  h->balance = false;
  h->gen++;
  return MISA_OK;
}

int misa_xcd_slabs(const misa_handle* h, int32_t zlo[9]) {
  if (!h || !zlo) return MISA_E_ARG;
  for (int c = 0; c <= 8; c++) zlo[c] = h->s.G.W.zlo[c];
  return MISA_OK;
}

int misa_defect_counters(misa_handle* h, int64_t out[3]) {
  if (!h || !out) return MISA_E_ARG;
  HIPCHECK(h, hipSetDevice(h->device));
  Ctl c;
  HIPCHECK(h, hipMemcpy(&c, h->s.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  out[0] = c.staged_total;
  out[1] = c.refill_total;
  out[2] = c.far_total;
  return MISA_OK;
}

int misa_force_pass_used(const misa_handle* h, int32_t* kind) {
  if (!h || !kind) return MISA_E_ARG;
  *kind = h->force_pass;
  return MISA_OK;
}

int misa_get_launch_shape(const misa_handle* h, int32_t* ty, int32_t* gathers_rho, int32_t* gathers_force) {
  if (!h || !ty || !gathers_rho || !gathers_force) return MISA_E_ARG;
  *ty = h->s.G.W.ty;
  *gathers_rho = h->s.uRho;
  *gathers_force = h->s.uForce;
  return MISA_OK;
}

int misa_host_offsets(const misa_geom* g, int which, int64_t* out, int64_t cap, int64_t* n) {
  if (!g || !n || which < 0 || which > 5) return MISA_E_ARG;
  misa_handle h;
  h.box.init(g->phase_space, g->lattice_const, g->cutoff_radius_factor);
  makeOffsets(h.box, h.ref);
  if (buildPlaneOffsets(&h) != MISA_OK) return MISA_E_ARG;
  return misa_get_offsets(&h, which, out, cap, n);
}

int misa_host_potential_table(const char* path, int kind, int idx, double* out, int64_t cap, int64_t* n) {
  if (!path || !n) return MISA_E_ARG;
  Setfl sf;
  std::string e;
  if (!parseSetfl(path, sf, e)) return MISA_E_IO;
  const std::vector<std::vector<double>>& src = kind == 0 ? sf.F : (kind == 1 ? sf.rho : sf.z2);
  if (idx < 0 || idx >= static_cast<int>(src.size())) return MISA_E_ARG;
  std::vector<double> c;
  interpolate7(src[idx], kind == 0 ? sf.drho : sf.dr, c);
  *n = static_cast<int64_t>(c.size());
  if (out && *n <= cap) std::memcpy(out, c.data(), c.size() * 8);
  return MISA_OK;
}

int misa_host_build_world(const misa_geom* g, uint32_t seed, double tset, const int32_t ratio[3], uint32_t alloy_seed,
                          void* atoms) {
  if (!g || !ratio || !atoms) return MISA_E_ARG;
  if (ratio[0] < 0 || ratio[1] < 0 || ratio[2] < 0 || ratio[0] + ratio[1] + ratio[2] <= 0) return MISA_E_ARG;
  Box b;
  b.init(g->phase_space, g->lattice_const, g->cutoff_radius_factor);
  std::vector<AtomElement> v;
  buildWorld(b, seed, tset, ratio, alloy_seed, v);
  std::memcpy(atoms, v.data(), v.size() * sizeof(AtomElement));
  return MISA_OK;
}

}  // extern "C"
