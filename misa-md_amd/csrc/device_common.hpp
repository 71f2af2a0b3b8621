// Device-side layout and inline arithmetic shared by the EAM kernels.
//
// Resident layout ("parity planes").  The reference stores the BCC lattice as
// one ghost-extended array whose x index is doubled: even x = cube corner,
// odd x = body centre (atom_set.cpp:17-26).  On the GPU the two sublattices
// live in two separate planes, each a dense (NZ, NY, NX) array of cells with
// the same ghost width g as the reference:
//
//     slot(p, cx, cy, cz) = p * plane + (cz * NY + cy) * NX + cx,  padded coords
//
// so 64 consecutive sites of one parity form one wavefront, every lane of it
// uses the same neighbour offset list, and offset k is one uniform slot delta:
// neighbour loads are unit-stride across lanes.  Reference linear offsets map
// one-to-one onto (target parity, dcx, dy, dz) (decodeOffset in
// host_setup.cpp), so the pairs visited are the reference's pairs.
//
// State is SoA: x (two ping-pong buffers per component), v, f, df, rho, a
// type byte and the atom id; mixed precision adds 16-byte site records
// (site_common.hpp).  The type byte holds the species in bits 0-1, bit 6
// marks a site with an inter atom in its 229-site neighbourhood, bit 5 a site
// near a lattice atom far from its site (kWide), bit 4 that far atom itself
// (kFar), and a negative value marks a vacancy (atom_type::INVALID).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace misa {

constexpr int kMaxLatOffsets = 128;   // pruned lattice-lattice list per parity
// Entry of the pruned-offset tables past every list (at most kPadOffset
// offsets): a lattice vector of 1e18 A and a zero record shift.  The mixed
// passes visit offsets in full groups and point a group's entries past the
// list there, so those pairs are beyond the cutoff and add nothing.
constexpr int kPadOffset = kMaxLatOffsets - 1;
constexpr int kMaxFullOffsets = 256;  // reference full list per parity
constexpr int kKnot = 8;              // values per knot record: c0 c1 c2 _ c3 c4 c5 c6
constexpr int8_t kVacant = -1;
constexpr int8_t kNearInter = 0x40;
// A site within the full neighbourhood of a lattice atom that sits farther
// than 0.2a from its site (an inter atom that has just refilled a vacancy
// can sit anywhere in that site's Wigner-Seitz cell): its sums also visit
// the full-list offsets the pruned list leaves out.
constexpr int8_t kWide = 0x20;
// The lattice atom farther than 0.2a from its site itself (and its periodic
// images): the only targets beyond the pruned list that a kWide site's sums
// can reach, unless the site carries kFar itself.
constexpr int8_t kFar = 0x10;

// Site traversal of the resident kernels: each XCD takes a z-slab of both
// parity planes and walks it in y-tiles of `ty` rows (z inner), so the
// five-plane stencil window of the rows in flight stays in that XCD's L2.
struct Walk {
  int ty;          // rows per y-tile (divides n1)
  int zlo[9];      // z-slab bounds per XCD: [zlo[c], zlo[c+1])
  int64_t bpx;     // blocks per parity per XCD (max over XCDs)
};

struct Geom {
  int n0, n1, n2;  // cells per dimension of the periodic box
  int g;           // ghost cells per face
  int NX, NY, NZ;  // padded plane dimensions (n + 2g)
  int64_t plane;   // NX * NY * NZ
  int64_t nsub;    // n0 * n1 * n2 (sites per parity)
  double a;        // lattice constant
  double rc2;      // (a*f)^2
  double L0, L1, L2;
  double thr;      // (0.2 a)^2, decide() threshold (atom.cpp:42)
  Walk W;          // traversal order of the site kernels
};

// Control block in device memory (counters and error flags).
struct Ctl {
  int64_t n_inter;      // local inter atoms
  int64_t n_ghost;      // ghost images
  int64_t n_stage;      // new inter atoms staged by the Verlet kernels
  int64_t n_prev;       // entries recorded in prev_slots
  int64_t err_flags;    // kErr* bits below
  int64_t n_far;        // lattice atoms farther than 0.2a from their site, found for the next index
  int64_t n_far_prev;   // the same, as flagged by the last decide
  double dt;            // adaptive step: length of the current step (ps)
  double t_sim;         // adaptive step: simulated time accumulated on the device (ps)
  unsigned long long vmax2;  // adaptive step: max |v|^2 (bits of a non-negative double)
  // neighbour skin of the mixed density pass: largest displacement of a
  // lattice atom summed over the drifts since the last full scan (A), the
  // current drift's max |dx|^2 (bits), whether this step's density pass
  // visits only the offsets within cutoff + skin of the last full scan, and
  // a request for a full scan
  double disp_acc;
  unsigned long long dispmax2;
  int32_t skin_mode;
  int32_t rescan;
  // since the upload: lattice atoms taken off their sites, inter atoms put
  // back on a vacant site, and of those the ones farther than 0.2a from it
  int64_t staged_total;
  int64_t refill_total;
  int64_t far_total;
  // staged atoms put back on their own site ahead of the decide's list work
  // in this step (kQuickRefill), whose stage entries the decide skips
  int64_t n_quick;
};
static_assert(sizeof(Ctl) == 17 * 8, "Ctl layout");

// Bits of Ctl::err_flags, set by the kernels and read after every call.
constexpr int64_t kErrInterCap = 1;   // inter-atom list capacity exceeded
constexpr int64_t kErrGhostCap = 2;   // ghost inter-atom capacity exceeded
constexpr int64_t kErrNonFinite = 4;  // a force (or density) came out NaN or infinite
constexpr int64_t kErrLeftBox = 8;    // an inter atom left the ghost-extended box
constexpr int64_t kErrRhoRange = 16;  // an electron density beyond the last knot of F(rho)

__device__ __forceinline__ void raiseFlag(const Ctl* ctl, int64_t bits) {
  atomicOr(reinterpret_cast<unsigned long long*>(&const_cast<Ctl*>(ctl)->err_flags),
           static_cast<unsigned long long>(bits));
}
// Non-finite sums raise kErrNonFinite (the sum of the three is NaN or
// infinite when any of them is; a finite overflow of the sum is as bad).
__device__ __forceinline__ void checkFinite(const Ctl* ctl, double a, double b, double c) {
  if (!__builtin_isfinite(a + b + c)) raiseFlag(ctl, kErrNonFinite);
}
// The density of a site against the embedding table: NaN or infinite raises
// kErrNonFinite, beyond the last knot kErrRhoRange (F' then comes from the
// last segment, as the interpolation clamps it).
__device__ __forceinline__ void checkRho(const Ctl* ctl, double rho, double rhomax) {
  if (!__builtin_isfinite(rho)) raiseFlag(ctl, kErrNonFinite);
  else if (rho > rhomax) raiseFlag(ctl, kErrRhoRange);
}

// Masses and the force-to-velocity unit factor (atom_types.h:17-19,
// pre_define.h:11-21) for the kernels that derive the half-kick factors from
// a step length held in device memory.
__device__ __forceinline__ double speciesMass(int t) { return t == 0 ? 55.845 : (t == 1 ? 63.546 : 58.6934); }
constexpr double kFtm2vDev = 1.0 / 1.0364269e-4;

// Offset (target parity dp, dz, dy, dcx in cells) of a pruned neighbour packed
// into one word, one signed byte each (in the w of the lattice-vector table),
// so that every box of a batch can turn it into its own slot delta.
__host__ __device__ inline uint32_t packDelta(int dp, int dz, int dy, int dcx) {
  return (static_cast<uint32_t>(dp) & 0xffu) | ((static_cast<uint32_t>(dz) & 0xffu) << 8) |
         ((static_cast<uint32_t>(dy) & 0xffu) << 16) | ((static_cast<uint32_t>(dcx) & 0xffu) << 24);
}

// Pruned lattice-lattice neighbour list: slot deltas per centre parity, and
// the same deltas shifted by `lo` (the most negative delta) so that a buffer
// load can take them as its non-negative scalar byte offset from a base that
// sits `lo` slots below the array.
struct LatOffsets {
  int n[2];
  int32_t off[2][kMaxLatOffsets];
  int32_t lo;
  uint32_t sh[2][kMaxLatOffsets];  // off - lo
};


// Spline tables: one knot record of kKnot values per knot, all tables of one
// kind in one allocation with a fixed stride; index by species (0 Fe, 1 Cu,
// 2 Ni) and by pair (max*(max+1)/2 + min).
template <typename T>
struct Tables {
  const T* F;    // 3 tables
  const T* rho;  // 3 tables
  const T* z2;   // 6 tables
  int64_t sF, sR;  // strides (values) of one F table / one r table
  int nrho, nr;
  T rdrho, rdr;   // 1/drho, 1/dr
  T rhomax;
  // packed records for the resident kernels
  const T* rhoV;   // [3][nr][4]: value cubic c3..c6 of rho_e
  const T* pairF;  // [9][nr][8]: value cubics of (r phi)_eiej and rho_ej (c3..c6 each)
  // Compact knots of the mixed passes (FP32 only).  A cubic Hermite segment
  // m is fixed by the values y and slopes s (per unit knot coordinate) at its
  // two ends and one more coefficient, so a knot need only hold what segment
  // m cannot take from knot m+1, and one load of 12 or 16 bytes at knot m
  // reads a whole segment.  The narrower the knot, the fewer cache lines the
  // gathers of a wavefront touch.
  const T* rhoH;   // [3][nr][2]: {y, s} of rho_e (density pass)
  // FP32 force pass (global memory and LDS alike): 16-byte knots that need
  // no coefficient arithmetic and sit 16-byte aligned for LDS reads
  const T* rhoD;   // [3][nr][4]: slope cubic of rho_e divided by dr {3 c3, 2 c4, c5, 0} / dr
  const T* pairZ;  // [9][nr][4]: value cubic of (r phi)_eiej {c3, c4, c5, c6}
};
constexpr int kPairRec = 8;
constexpr int kRhoRec = 4;

__device__ __forceinline__ bool isVacant(int8_t t) { return t < 0; }
// A full-list offset beyond the pruned list can hold a pair within the
// cutoff only when the site or the occupied target is a far lattice atom.
__device__ __forceinline__ bool wideTarget(int8_t ti, int8_t tj) { return !(tj < 0) && ((ti | tj) & kFar) != 0; }
__device__ __forceinline__ int species(int8_t t) { return t & 3; }
__device__ __forceinline__ int pairIdx(int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

// Knot search of the mixed-precision passes for r below the cutoff: knot
// floor(r/dr) (at most n-2) and the local coordinate r/dr - knot (at most 1).
// The same knot and coordinate as knotOf in exact arithmetic, without the
// one-based detour.
__device__ __forceinline__ int knotM(float r, float rdr, int n, float& q) {
#pragma clang fp contract(off)
  const float t = r * rdr;
  const int m = min(static_cast<int>(t), n - 2);
  q = fminf(t - static_cast<float>(m), 1.0f);
  return m;
}

// Knot search of the eam/alloy convention: p = x/delta + 1 (one-based knot),
// m = int(p) clamped to n-1 (and to >= 1 for rho-tables), p -= m, p <= 1.
// Returns the zero-based knot; p is the local coordinate.
template <typename T>
__device__ __forceinline__ int knotOf(T x, T rinv, int n, bool clamp_low, T& p) {
  p = x * rinv + T(1);
  const T pm = p < T(n - 1) ? p : T(n - 1);
  int m = static_cast<int>(pm);
  if (clamp_low && m < 1) m = 1;
  p -= T(m);
  p = p < T(1) ? p : T(1);
  return m - 1;
}

template <typename T>
__device__ __forceinline__ T splineVal(const T* k, T p) {
  return ((k[4] * p + k[5]) * p + k[6]) * p + k[7];
}
template <typename T>
__device__ __forceinline__ T splineDer(const T* k, T p) {
  return (k[0] * p + k[1]) * p + k[2];
}

// rho contributed by species e at distance r (pot eam::chargeDensity).
template <typename T>
__device__ __forceinline__ T rhoOf(const Tables<T>& tb, int e, T r) {
  T p;
  const int m = knotOf(r, tb.rdr, tb.nr, false, p);
  return splineVal(tb.rho + e * tb.sR + static_cast<int64_t>(m) * kKnot, p);
}

// F'(rho) (pot eam::dEmbedEnergy).
template <typename T>
__device__ __forceinline__ T dEmbed(const Tables<T>& tb, int e, T rho) {
  T p;
  const int m = knotOf(rho, tb.rdrho, tb.nrho, true, p);
  return splineDer(tb.F + e * tb.sF + static_cast<int64_t>(m) * kKnot, p);
}

// F(rho) with linear extension above the table.
template <typename T>
__device__ __forceinline__ T embed(const Tables<T>& tb, int e, T rho) {
  T p;
  const int m = knotOf(rho, tb.rdrho, tb.nrho, true, p);
  const T* k = tb.F + e * tb.sF + static_cast<int64_t>(m) * kKnot;
  T v = splineVal(k, p);
  if (rho > tb.rhomax) v += splineDer(k, p) * (rho - tb.rhomax);
  return v;
}

__device__ __forceinline__ double rsqrtRefined(double x) {
  double y = __builtin_amdgcn_rsq(x);
  // two Newton steps on y = x^-1/2: y <- y (3 - x y^2) / 2
  double h = x * y;
  double e = fma(-h, y, 1.0);
  y = fma(0.5 * y, e, y);
  h = x * y;
  e = fma(-h, y, 1.0);
  y = fma(0.5 * y, e, y);
  return y;
}

// Pair force factor (pot eam::toForce): fpair = -(F'_i rho'_j + F'_j rho'_i +
// phi') / r with phi = z2 / r; f_i += (x_i - x_j) * fpair.  FP64 version.
__device__ __forceinline__ double pairForce(const Tables<double>& tb, int ei, int ej, double r2, double dfi,
                                            double dfj) {
  const double rinv = rsqrtRefined(r2);
  const double r = r2 * rinv;
  double p;
  const int m = knotOf(r, tb.rdr, tb.nr, false, p);
  const int64_t km = static_cast<int64_t>(m) * kKnot;
  const double rhoip = splineDer(tb.rho + ei * tb.sR + km, p);
  const double rhojp = splineDer(tb.rho + ej * tb.sR + km, p);
  const double* kz = tb.z2 + pairIdx(ei, ej) * tb.sR + km;
  const double z2p = splineDer(kz, p);
  const double z2 = splineVal(kz, p);
  const double phi = z2 * rinv;
  const double phip = z2p * rinv - phi * rinv;
  const double psip = dfi * rhojp + dfj * rhoip + phip;
  return -psip * rinv;
}

// FP32 version: splines and pair terms in FP32 from an FP64 r^2.
__device__ __forceinline__ float pairForce(const Tables<float>& tb, int ei, int ej, double r2d, float dfi,
                                           float dfj) {
  const float r2 = static_cast<float>(r2d);
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * rinv;
  float p;
  const int m = knotOf(r, tb.rdr, tb.nr, false, p);
  const int64_t km = static_cast<int64_t>(m) * kKnot;
  const float rhoip = splineDer(tb.rho + ei * tb.sR + km, p);
  const float rhojp = splineDer(tb.rho + ej * tb.sR + km, p);
  const float* kz = tb.z2 + pairIdx(ei, ej) * tb.sR + km;
  const float z2p = splineDer(kz, p);
  const float z2 = splineVal(kz, p);
  const float phi = z2 * rinv;
  const float phip = z2p * rinv - phi * rinv;
  const float psip = dfi * rhojp + dfj * rhoip + phip;
  return -psip * rinv;
}

__device__ __forceinline__ double rhoPair(const Tables<double>& tb, int e, double r2) {
  return rhoOf(tb, e, r2 * rsqrtRefined(r2));
}
__device__ __forceinline__ float rhoPair(const Tables<float>& tb, int e, double r2d) {
  const float r2 = static_cast<float>(r2d);
  return rhoOf(tb, e, r2 * __builtin_amdgcn_rsqf(r2));
}

}  // namespace misa
