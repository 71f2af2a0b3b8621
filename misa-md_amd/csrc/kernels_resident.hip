// Resident-mode kernels of the EAM step on MI355X (gfx950).
//
// Per step the lattice work is two global passes over the SoA state, one
// thread per site, one wavefront per 64 consecutive sites of one parity:
//
//   kRhoDf  : rho_i = sum_j rho_{t_j}(r_ij) over the pruned full neighbour
//             list, F'(rho_i) -> df, df copied into the ghost images
//             (atom::latRho + latDf + the F' forward halo, atom.cpp:102-149)
//   kForce  : f_i = sum_j (x_i - x_j) fpair_ij, then the second half kick
//             and, in the fused mode, the next step's half kick, drift,
//             the lattice part of atom::decide and the position halo
//             (latForce + NewtonMotion + decide, simulation.cpp:164-194)
//
// A full list (each lane owns its site's sums) replaces the reference's half
// list with ghost partials and reverse halos: no atomics, deterministic sums.
// The list keeps only the reference offsets whose ideal separation is below
// rc + 0.4a (+1% of a): every lattice atom stays within 0.2a of its site (or
// decide() turns it into an inter atom), so every dropped offset is beyond
// the cutoff for every lattice pair and the visited pair set is unchanged.
//
// Workgroups are numbered XCD-aware: the hardware deals workgroups to the 8
// XCDs round-robin, so block b runs on XCD b % 8; the launch maps XCD c to the
// c-th z-slab of both parity planes so that each XCD's L2 holds its own slab
// and its halo instead of all XCDs streaming the whole box.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "inter_common.hpp"
#include "resident.hpp"
#include "site_common.hpp"

namespace misa {

namespace {

using namespace site;

// firststep from stored forces + lattice decide + position halo.  A
// streaming pass (no site reads another's data), so its grid walks the sites
// in slot order, x fastest, for long runs of consecutive addresses in every
// array; it strides over them with a few blocks per compute unit, so that few
// wavefronts add their largest drift to the one word every wavefront of the
// pass feeds.
__global__ __launch_bounds__(kBlock) void kFirst(Geom G, CXPtr x, XPtr xn, const int8_t* __restrict__ type,
                                                 XPtr v, CXPtr f, const uint64_t* __restrict__ id, VerletParams vp,
                                                 StageArgs st, Ctl* ctl, float4* __restrict__ recn, Rec64 rn64) {
#pragma clang fp contract(off)  // the half kick and drift round as newton_motion.cpp does
  unsigned long long db = 0;  // |dx|^2 bits of the drift, largest over the lane's sites
  const int64_t nsites = 2 * G.nsub;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; t < nsites;
       t += static_cast<int64_t>(gridDim.x) * kBlock) {
    int64_t s;
    int p, cx, cy, cz;
    siteInOrder(G, t, s, p, cx, cy, cz);
    const int8_t ti = type[s];
    double y0 = x.c[0][s], y1 = x.c[1][s], y2 = x.c[2][s];
    if (!isVacant(ti)) {
      const StepDt sd = stepDt(vp, species(ti));
      const double c = sd.c;
      double v0 = v.c[0][s] + c * f.c[0][s];
      double v1 = v.c[1][s] + c * f.c[1][s];
      double v2 = v.c[2][s] + c * f.c[2][s];
      y0 = y0 + sd.dt * v0;
      y1 = y1 + sd.dt * v1;
      y2 = y2 + sd.dt * v2;
      const unsigned long long d = speed2Bits(y0 - x.c[0][s], y1 - x.c[1][s], y2 - x.c[2][s]);
      db = d > db ? d : db;
      if (leftSite(G, p, cx, cy, cz, y0, y1, y2)) {
        stageInter(st, ctl, refIndex(G, p, cx, cy, cz), s, id[s], ti, y0, y1, y2, v0, v1, v2);
        v0 = v1 = v2 = 0.0;
      }
      v.c[0][s] = v0;
      v.c[1][s] = v1;
      v.c[2][s] = v2;
    }
    xn.c[0][s] = y0;
    xn.c[1][s] = y1;
    xn.c[2][s] = y2;
    forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
      xn.c[0][q] = y0 + sx * G.L0;
      xn.c[1][q] = y1 + sy * G.L1;
      xn.c[2][q] = y2 + sz * G.L2;
    });
    if (recn) {
      const float4 r = isVacant(ti) ? vacantRec() : siteRec(G, p, cx, cy, cz, y0, y1, y2, species(ti));
      recn[s] = r;
      forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) { recn[q] = r; });
    }
    if (rn64.xy) {
      putRec64(rn64, s, ti, y0, y1, y2);
      forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
        putRec64(rn64, q, ti, y0 + sx * G.L0, y1 + sy * G.L1, y2 + sz * G.L2);
      });
    }
  }
  waveMaxTo(db, &ctl->dispmax2);
}

// FP64 site records from x and type over every slot, ghosts included.
__global__ __launch_bounds__(kBlock) void kPackRec64(Geom G, CXPtr x, const int8_t* __restrict__ type, Rec64 r) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= 2 * G.plane) return;
  putRec64(r, t, type[t], x.c[0][t], x.c[1][t], x.c[2][t]);
}

// Site records from x and type over every slot, ghosts included.
__global__ __launch_bounds__(kBlock) void kPackRec(Geom G, CXPtr x, const int8_t* __restrict__ type,
                                                   float4* __restrict__ rec) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= 2 * G.plane) return;
  const int p = t >= G.plane ? 1 : 0;
  int64_t r = t - p * G.plane;
  const int cx = static_cast<int>(r % G.NX) - G.g;
  r /= G.NX;
  const int cy = static_cast<int>(r % G.NY) - G.g;
  const int cz = static_cast<int>(r / G.NY) - G.g;
  const int8_t ti = type[t];
  rec[t] = isVacant(ti) ? vacantRec() : siteRec(G, p, cx, cy, cz, x.c[0][t], x.c[1][t], x.c[2][t], species(ti));
}

// ---------------------------------------------------------------------------
// upload / download / full halo
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void kUpload(Geom G, const AtomElement* __restrict__ aos, int64_t sx,
                                                  int64_t sy, XPtr x, XPtr v, XPtr f, int8_t* type, uint64_t* id) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int64_t r = (static_cast<int64_t>(cz + G.g) * sy + (cy + G.g)) * sx + (2 * (cx + G.g) + p);
  const AtomElement& a = aos[r];
  for (int k = 0; k < 3; k++) {
    x.c[k][s] = a.x[k];
    v.c[k][s] = a.v[k];
    f.c[k][s] = a.f[k];
  }
  type[s] = a.type < 0 ? kVacant : static_cast<int8_t>(a.type);
  id[s] = a.id;
}

// Fill every ghost slot from its periodic source: x + shift, type, df.
__global__ __launch_bounds__(kBlock) void kHaloAll(Geom G, XPtr x, int8_t* type, double* df64, float* df32) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= 2 * G.plane) return;
  const int p = t >= G.plane ? 1 : 0;
  int64_t r = t - p * G.plane;
  const int ix = static_cast<int>(r % G.NX);
  r /= G.NX;
  const int iy = static_cast<int>(r % G.NY);
  const int iz = static_cast<int>(r / G.NY);
  int c[3] = {ix - G.g, iy - G.g, iz - G.g};
  const int n[3] = {G.n0, G.n1, G.n2};
  int sh[3] = {0, 0, 0};
  bool ghost = false;
  for (int k = 0; k < 3; k++) {
    if (c[k] < 0) {
      c[k] += n[k];
      sh[k] = -1;
      ghost = true;
    } else if (c[k] >= n[k]) {
      c[k] -= n[k];
      sh[k] = 1;
      ghost = true;
    }
  }
  if (!ghost) return;
  const int64_t src = slotOf(G, p, c[0], c[1], c[2]);
  x.c[0][t] = x.c[0][src] + sh[0] * G.L0;
  x.c[1][t] = x.c[1][src] + sh[1] * G.L1;
  x.c[2][t] = x.c[2][src] + sh[2] * G.L2;
  type[t] = type[src];
  if (df64) df64[t] = df64[src];
  if (df32) df32[t] = df32[src];
}

// Plane state -> reference AoS array (every slot, ghosts from their images).
__global__ __launch_bounds__(kBlock) void kDownload(Geom G, AtomElement* __restrict__ aos, int64_t sx, int64_t sy,
                                                    CXPtr x, CXPtr v, CXPtr f, const int8_t* type,
                                                    const uint64_t* id, const double* rho, const double* df64,
                                                    const float* dfrec) {  // mixed: w of the site records
  const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (r >= 2 * G.plane) return;
  const int64_t X = r % sx;
  const int64_t rest = r / sx;
  const int64_t y = rest % sy, z = rest / sy;
  const int p = static_cast<int>(X & 1);
  const int64_t s = p * G.plane + (z * G.NY + y) * G.NX + (X >> 1);
  const bool inside = (X >> 1) >= G.g && (X >> 1) < G.g + G.n0 && y >= G.g && y < G.g + G.n1 && z >= G.g &&
                      z < G.g + G.n2;
  AtomElement a;
  a.id = inside ? id[s] : 0;
  a.type = type[s] < 0 ? -1 : species(type[s]);
  for (int k = 0; k < 3; k++) {
    a.x[k] = x.c[k][s];
    a.v[k] = inside ? v.c[k][s] : 0.0;
    a.f[k] = inside ? f.c[k][s] : 0.0;
  }
  a.rho = inside ? rho[s] : 0.0;
  a.df = df64 ? df64[s] : static_cast<double>(dfrec[4 * s]);
  aos[r] = a;
}

// ---------------------------------------------------------------------------
// reductions (FP64, fixed order: block partials summed on the host in block order)
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void blockSum(double (&vals)[N], double* out) {
  __shared__ double red[N][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; q++) {
    double v = vals[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[q][w] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; q++) {
      double v = 0;
      for (int k = 0; k < kBlock / 64; k++) v += red[q][k];
      out[q] = v;
    }
  }
}

__global__ __launch_bounds__(kBlock) void kThermo(Geom G, CXPtr v, CXPtr f, const int8_t* type, double* partials) {
  int64_t s;
  int p, cx, cy, cz;
  double vals[7] = {0, 0, 0, 0, 0, 0, 0};
  if (siteOfThread(G, s, p, cx, cy, cz)) {
    const int8_t t = type[s];
    if (isVacant(t)) {
      vals[2] = 1;
    } else {
      const double m = species(t) == 0 ? 55.845 : (species(t) == 1 ? 63.546 : 58.6934);
      const double a = v.c[0][s], b = v.c[1][s], c = v.c[2][s];
      const double vv = a * a + b * b + c * c;
      vals[0] = vv * m;
      vals[1] = 1;
      vals[3] = f.c[0][s];
      vals[4] = f.c[1][s];
      vals[5] = f.c[2][s];
      vals[6] = vv;
    }
  }
  // max over the block for the speed, sums for the rest
  __shared__ double vmax[kBlock];
  vmax[threadIdx.x] = vals[6];
  vals[6] = 0;
  blockSum<7>(vals, partials + 7 * blockIdx.x);
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) vmax[threadIdx.x] = fmax(vmax[threadIdx.x], vmax[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[7 * blockIdx.x + 6] = vmax[0];
}

// F'(rho) with the FP64 tables for every occupied slot, ghosts included,
// from the stored densities, over every slot: what the FP64 density pass
// leaves in df and in the w of the FP64 site records, for the energy pass.
__global__ __launch_bounds__(kBlock) void kDfOfRho(Geom G, Tables<double> tb, const int8_t* __restrict__ type,
                                                   const double* __restrict__ rho, double* __restrict__ df,
                                                   double2* __restrict__ zw) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (t >= 2 * G.plane) return;
  const int8_t ti = type[t];
  if (isVacant(ti)) return;
  const double d = dEmbed(tb, species(ti), rho[t]);
  df[t] = d;
  zw[t].y = packW64(d, species(ti));
}

__global__ __launch_bounds__(kBlock) void kScaleV(Geom G, XPtr v, double factor) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  v.c[0][s] *= factor;
  v.c[1][s] *= factor;
  v.c[2][s] *= factor;
}

__global__ void kAddV(XPtr v, int64_t s, double a, double b, double c) {
  v.c[0][s] += a;
  v.c[1][s] += b;
  v.c[2][s] += c;
}

}  // namespace

void launchUpload(const DevState& s, const AtomElement* aos_dev, int64_t sx, int64_t sy, hipStream_t st) {
  kUpload<<<siteGrid(s.G), kBlock, 0, st>>>(s.G, aos_dev, sx, sy, xp(s.x[s.cur]), xp(s.v), xp(s.f), s.type, s.id);
}

void launchHaloAll(const DevState& s, int buf, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((2 * s.G.plane + kBlock - 1) / kBlock);
  kHaloAll<<<grid, kBlock, 0, st>>>(s.G, xp(s.x[buf]), s.type, s.df64, nullptr);
}

void launchDownload(const DevState& s, AtomElement* aos_dev, int64_t sx, int64_t sy, bool mixed, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((2 * s.G.plane + kBlock - 1) / kBlock);
  kDownload<<<grid, kBlock, 0, st>>>(s.G, aos_dev, sx, sy, cxp(s.x[s.cur]), cxp(s.v), cxp(s.f), s.type, s.id,
                                     s.rho, mixed ? nullptr : s.df64,
                                     reinterpret_cast<const float*>(s.rec[s.cur]) + 3);
}

void launchPackRec(const DevState& s, int buf, bool fp64, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((2 * s.G.plane + kBlock - 1) / kBlock);
  if (fp64)
    kPackRec64<<<grid, kBlock, 0, st>>>(s.G, cxp(s.x[buf]), s.type, Rec64{s.pxy[buf], s.pzw[buf]});
  else
    kPackRec<<<grid, kBlock, 0, st>>>(s.G, cxp(s.x[buf]), s.type, s.rec[buf]);
}

void launchFirst(const DevState& s, bool mixed, const VerletParams& vp, hipStream_t st) {
  kFirst<<<strideGrid(s.G, s.ncu), kBlock, 0, st>>>(s.G, cxp(s.x[s.cur]), xp(s.x[1 - s.cur]), s.type, xp(s.v),
                                                    cxp(s.f), s.id, vp, stageOf(s), s.ctl,
                                                    mixed ? s.rec[1 - s.cur] : nullptr,
                                                    mixed ? Rec64{nullptr, nullptr}
                                                          : Rec64{s.pxy[1 - s.cur], s.pzw[1 - s.cur]});
}

int64_t launchThermo(const DevState& s, hipStream_t st) {
  const unsigned grid = siteGrid(s.G);
  kThermo<<<grid, kBlock, 0, st>>>(s.G, cxp(s.v), cxp(s.f), s.type, s.partials);
  return grid;
}

int64_t launchEnergy(const DevState& s, const LatOffsets& o, const Tables<double>& t64, bool all, hipStream_t st) {
  // the FP64 site records of the current positions with F' of the stored
  // densities in w, whichever precision and step left the state (a mixed
  // step keeps no FP64 records; in FP64 the values come out the same)
  const unsigned g = static_cast<unsigned>((2 * s.G.plane + kBlock - 1) / kBlock);
  const Rec64 r64{s.pxy[s.cur], s.pzw[s.cur]};
  kPackRec64<<<g, kBlock, 0, st>>>(s.G, cxp(s.x[s.cur]), s.type, r64);
  kDfOfRho<<<g, kBlock, 0, st>>>(s.G, t64, s.type, s.rho, s.df64, r64.zw);
  return launchEnergyPass(s, o, t64, all, st);
}

void launchScaleV(const DevState& s, double factor, hipStream_t st) {
  kScaleV<<<siteGrid(s.G), kBlock, 0, st>>>(s.G, xp(s.v), factor);
}

void launchSetv(const DevState& s, int64_t slot, double a, double b, double c, hipStream_t st) {
  kAddV<<<1, 1, 0, st>>>(xp(s.v), slot, a, b, c);
}

}  // namespace misa
