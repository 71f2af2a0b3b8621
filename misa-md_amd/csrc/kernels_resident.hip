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
#include "resident.hpp"

namespace misa {

namespace {

constexpr int kBlock = 256;

struct XPtr {
  double* c[3];
};
struct CXPtr {
  const double* c[3];
};

__device__ __forceinline__ int64_t slotOf(const Geom& G, int p, int cx, int cy, int cz) {
  return p * G.plane + (static_cast<int64_t>(cz + G.g) * G.NY + (cy + G.g)) * G.NX + (cx + G.g);
}

// Hardware block -> (parity, site) with the XCD-aware slab order.
__device__ __forceinline__ bool siteOfThread(const Geom& G, int64_t& s, int& p, int& cx, int& cy, int& cz) {
  const int64_t B = (G.nsub + kBlock - 1) / kBlock;  // blocks per parity
  const int64_t C = (B + 7) / 8;                      // blocks per parity per XCD
  const int64_t b = blockIdx.x;
  const int64_t k = b >> 3;
  p = static_cast<int>(k & 1);
  const int64_t blk = (b & 7) * C + (k >> 1);
  if (blk >= B) return false;
  const int64_t t = blk * kBlock + threadIdx.x;
  if (t >= G.nsub) return false;
  const uint32_t tt = static_cast<uint32_t>(t);
  const uint32_t n0 = static_cast<uint32_t>(G.n0), n1 = static_cast<uint32_t>(G.n1);
  cx = static_cast<int>(tt % n0);
  const uint32_t r = tt / n0;
  cy = static_cast<int>(r % n1);
  cz = static_cast<int>(r / n1);
  s = slotOf(G, p, cx, cy, cz);
  return true;
}

inline unsigned siteGrid(const Geom& G) {
  const int64_t B = (G.nsub + kBlock - 1) / kBlock;
  const int64_t C = (B + 7) / 8;
  return static_cast<unsigned>(16 * C);
}

__device__ __forceinline__ bool nearFace(const Geom& G, int cx, int cy, int cz) {
  return cx < G.g || cx >= G.n0 - G.g || cy < G.g || cy >= G.n1 - G.g || cz < G.g || cz >= G.n2 - G.g;
}

// Calls fn(slot, sx, sy, sz) for every periodic image of sub-box cell
// (cx, cy, cz) that lies in the ghost layer; (sx, sy, sz) in {-1, 0, 1} is the
// image shift in box lengths (exchangeAtom: x_ghost = x + offset,
// lat_particle_packer.cpp:22-32,163-194).
template <typename Fn>
__device__ __forceinline__ void forEachImage(const Geom& G, int p, int cx, int cy, int cz, Fn fn) {
  if (!nearFace(G, cx, cy, cz)) return;
#pragma unroll
  for (int dz = -1; dz <= 1; dz++) {
    const int iz = cz + dz * G.n2;
    if (iz < -G.g || iz >= G.n2 + G.g) continue;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
      const int iy = cy + dy * G.n1;
      if (iy < -G.g || iy >= G.n1 + G.g) continue;
#pragma unroll
      for (int dx = -1; dx <= 1; dx++) {
        if (dx == 0 && dy == 0 && dz == 0) continue;
        const int ix = cx + dx * G.n0;
        if (ix < -G.g || ix >= G.n0 + G.g) continue;
        fn(slotOf(G, p, ix, iy, iz), dx, dy, dz);
      }
    }
  }
}

// atom::decide lattice test (atom.cpp:33-49) with every product and sum
// rounded separately, as the reference computes it.
__device__ __noinline__ bool leftSite(const Geom& G, int p, int cx, int cy, int cz, double x0, double x1,
                                      double x2) {
#pragma clang fp contract(off)
  const double i = static_cast<double>(2 * cx + p);
  const double xt = i * 0.5 * G.a;
  const double yt = (static_cast<double>(cy) + p * 0.5) * G.a;
  const double zt = (static_cast<double>(cz) + p * 0.5) * G.a;
  double dist = (x0 - xt) * (x0 - xt);
  dist += (x1 - yt) * (x1 - yt);
  dist += (x2 - zt) * (x2 - zt);
  return dist > G.thr;
}

// Reference linear index of sub-box cell (cx, cy, cz) of parity p.
__device__ __forceinline__ int64_t refIndex(const Geom& G, int p, int cx, int cy, int cz) {
  const int64_t sx = 2 * static_cast<int64_t>(G.NX);
  return (static_cast<int64_t>(cz + G.g) * G.NY + (cy + G.g)) * sx + (2 * (cx + G.g) + p);
}

struct StageArgs {
  int64_t cap;
  int64_t* key;
  int64_t* slot;
  uint64_t* id;
  int8_t* type;
  double* x[3];
  double* v[3];
};

__device__ __forceinline__ void stageInter(const StageArgs& st, Ctl* ctl, int64_t key, int64_t slot, uint64_t id,
                                           int8_t type, double x0, double x1, double x2, double v0, double v1,
                                           double v2) {
  const unsigned long long k = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->n_stage), 1ull);
  if (static_cast<int64_t>(k) >= st.cap) {
    atomicOr(reinterpret_cast<unsigned long long*>(&ctl->err_flags), 1ull);
    return;
  }
  st.key[k] = key;
  st.slot[k] = slot;
  st.id[k] = id;
  st.type[k] = type;
  st.x[0][k] = x0;
  st.x[1][k] = x1;
  st.x[2][k] = x2;
  st.v[0][k] = v0;
  st.v[1][k] = v1;
  st.v[2][k] = v2;
}

// Per-site sums: FP64 accumulators (ACC 0) or FP32 with Kahan compensation
// (ACC 1).
template <int ACC>
struct Sum1;
template <>
struct Sum1<0> {
  double s = 0;
  __device__ __forceinline__ void add(double v) { s += v; }
  __device__ __forceinline__ double value() const { return s; }
};
template <>
struct Sum1<1> {
  float s = 0, c = 0;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return static_cast<double>(s); }
};

template <int ACC>
struct Sum3;
template <>
struct Sum3<0> {
  double f0 = 0, f1 = 0, f2 = 0;
  template <typename T>
  __device__ __forceinline__ void add(double d0, double d1, double d2, T fp) {
    const double q = static_cast<double>(fp);
    f0 = fma(d0, q, f0);
    f1 = fma(d1, q, f1);
    f2 = fma(d2, q, f2);
  }
};
template <>
struct Sum3<1> {
  Sum1<1> a, b, c;
  double f0 = 0, f1 = 0, f2 = 0;
  __device__ __forceinline__ void add(double d0, double d1, double d2, float fp) {
    a.add(static_cast<float>(d0) * fp);
    b.add(static_cast<float>(d1) * fp);
    c.add(static_cast<float>(d2) * fp);
  }
};
template <int ACC>
__device__ __forceinline__ void sum3Value(const Sum3<ACC>& s, double& f0, double& f1, double& f2);
template <>
__device__ __forceinline__ void sum3Value<0>(const Sum3<0>& s, double& f0, double& f1, double& f2) {
  f0 = s.f0;
  f1 = s.f1;
  f2 = s.f2;
}
template <>
__device__ __forceinline__ void sum3Value<1>(const Sum3<1>& s, double& f0, double& f1, double& f2) {
  f0 = s.a.value();
  f1 = s.b.value();
  f2 = s.c.value();
}

// ---------------------------------------------------------------------------
// pass 1: rho and F'
// ---------------------------------------------------------------------------
template <typename T, int ACC>
__global__ __launch_bounds__(kBlock) void kRhoDf(Geom G, LatOffsets O, Tables<T> tb, CXPtr x,
                                                 const int8_t* __restrict__ type, T* __restrict__ df,
                                                 double* __restrict__ rho_out) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  T dfi = T(0);
  double rho = 0.0;
  if (!isVacant(ti)) {
    const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
    Sum1<ACC> acc;
    const int nk = O.n[p];
    for (int k = 0; k < nk; ++k) {
      const int64_t j = s + O.off[p][k];
      const int8_t tj = type[j];
      const double d0 = xi0 - x.c[0][j];
      const double d1 = xi1 - x.c[1][j];
      const double d2 = xi2 - x.c[2][j];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (!isVacant(tj) && r2 < G.rc2) acc.add(rhoPair(tb, species(tj), r2));
    }
    rho = acc.value();
    dfi = dEmbed(tb, species(ti), static_cast<T>(rho));
  }
  df[s] = dfi;
  if (rho_out) {
    rho_out[s] = rho;
    forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) {
      df[q] = dfi;
      rho_out[q] = rho;
    });
  } else {
    forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) { df[q] = dfi; });
  }
}

// ---------------------------------------------------------------------------
// pass 2: force and velocity Verlet
//   MODE 0: f stored, v += c f                        (secondstep)
//   MODE 1: v += c f; v += c f; x' = x + dt v; decide  (secondstep + next firststep)
//   MODE 2: f stored only                             (first force evaluation)
// ---------------------------------------------------------------------------
template <typename T, int ACC, int MODE>
__global__ __launch_bounds__(kBlock) void kForce(Geom G, LatOffsets O, Tables<T> tb, CXPtr x, XPtr xn,
                                                 const int8_t* __restrict__ type, const T* __restrict__ df,
                                                 XPtr v, XPtr f, const uint64_t* __restrict__ id,
                                                 VerletParams vp, StageArgs st, Ctl* ctl) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
  if (isVacant(ti)) {  // vacancy: keeps its stale position (atom.cpp:43-47)
    if (MODE == 1) {
      xn.c[0][s] = xi0;
      xn.c[1][s] = xi1;
      xn.c[2][s] = xi2;
      forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
        xn.c[0][q] = xi0 + sx * G.L0;
        xn.c[1][q] = xi1 + sy * G.L1;
        xn.c[2][q] = xi2 + sz * G.L2;
      });
    }
    return;
  }
  const int ei = species(ti);
  const T dfi = df[s];
  Sum3<ACC> acc;
  const int nk = O.n[p];
  for (int k = 0; k < nk; ++k) {
    const int64_t j = s + O.off[p][k];
    const int8_t tj = type[j];
    const double d0 = xi0 - x.c[0][j];
    const double d1 = xi1 - x.c[1][j];
    const double d2 = xi2 - x.c[2][j];
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (!isVacant(tj) && r2 < G.rc2) {
      const T fpair = pairForce(tb, ei, species(tj), r2, dfi, df[j]);
      acc.add(d0, d1, d2, fpair);
    }
  }
  double f0, f1, f2;
  sum3Value<ACC>(acc, f0, f1, f2);
  const double c = vp.c[ei];
  double v0 = v.c[0][s] + c * f0;
  double v1 = v.c[1][s] + c * f1;
  double v2 = v.c[2][s] + c * f2;
  if (MODE == 2) {
    f.c[0][s] = f0;
    f.c[1][s] = f1;
    f.c[2][s] = f2;
    return;
  }
  if (MODE == 0) {
    f.c[0][s] = f0;
    f.c[1][s] = f1;
    f.c[2][s] = f2;
    v.c[0][s] = v0;
    v.c[1][s] = v1;
    v.c[2][s] = v2;
    return;
  }
  v0 = v0 + c * f0;
  v1 = v1 + c * f1;
  v2 = v2 + c * f2;
  const double y0 = xi0 + vp.dt * v0;
  const double y1 = xi1 + vp.dt * v1;
  const double y2 = xi2 + vp.dt * v2;
  if (leftSite(G, p, cx, cy, cz, y0, y1, y2)) {
    stageInter(st, ctl, refIndex(G, p, cx, cy, cz), s, id[s], ti, y0, y1, y2, v0, v1, v2);
    v0 = v1 = v2 = 0.0;
  }
  v.c[0][s] = v0;
  v.c[1][s] = v1;
  v.c[2][s] = v2;
  xn.c[0][s] = y0;
  xn.c[1][s] = y1;
  xn.c[2][s] = y2;
  forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
    xn.c[0][q] = y0 + sx * G.L0;
    xn.c[1][q] = y1 + sy * G.L1;
    xn.c[2][q] = y2 + sz * G.L2;
  });
}

// firststep from stored forces + lattice decide + position halo.
__global__ __launch_bounds__(kBlock) void kFirst(Geom G, CXPtr x, XPtr xn, const int8_t* __restrict__ type,
                                                 XPtr v, CXPtr f, const uint64_t* __restrict__ id, VerletParams vp,
                                                 StageArgs st, Ctl* ctl) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  double y0 = x.c[0][s], y1 = x.c[1][s], y2 = x.c[2][s];
  if (!isVacant(ti)) {
    const double c = vp.c[species(ti)];
    double v0 = v.c[0][s] + c * f.c[0][s];
    double v1 = v.c[1][s] + c * f.c[1][s];
    double v2 = v.c[2][s] + c * f.c[2][s];
    y0 = y0 + vp.dt * v0;
    y1 = y1 + vp.dt * v1;
    y2 = y2 + vp.dt * v2;
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
                                                    const float* df32) {
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
  a.df = df64 ? df64[s] : static_cast<double>(df32[s]);
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

// Potential energy and virial with FP64 tables over the full list:
// per site F(rho_i) + 1/2 sum_j phi(r_ij) and 1/2 sum_j r_ij^2 fpair_ij.
__global__ __launch_bounds__(kBlock) void kEnergy(Geom G, LatOffsets O, Tables<double> tb, CXPtr x,
                                                  const int8_t* __restrict__ type, const double* __restrict__ rho,
                                                  double* partials) {
  int64_t s;
  int p, cx, cy, cz;
  double vals[3] = {0, 0, 0};
  if (siteOfThread(G, s, p, cx, cy, cz)) {
    const int8_t ti = type[s];
    if (!isVacant(ti)) {
      const int ei = species(ti);
      const double ri = rho[s];
      const double dfi = dEmbed(tb, ei, ri);
      vals[0] = embed(tb, ei, ri);
      const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
      for (int k = 0; k < O.n[p]; ++k) {
        const int64_t j = s + O.off[p][k];
        const int8_t tj = type[j];
        const double d0 = xi0 - x.c[0][j], d1 = xi1 - x.c[1][j], d2 = xi2 - x.c[2][j];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (!isVacant(tj) && r2 < G.rc2) {
          const int ej = species(tj);
          const double r = sqrt(r2);
          double q;
          const int m = knotOf(r, tb.rdr, tb.nr, false, q);
          vals[1] += 0.5 * splineVal(tb.z2 + pairIdx(ei, ej) * tb.sR + static_cast<int64_t>(m) * kKnot, q) / r;
          vals[2] += 0.5 * r2 * pairForce(tb, ei, ej, r2, dfi, dEmbed(tb, ej, rho[j]));
        }
      }
    }
  }
  blockSum<3>(vals, partials + 3 * blockIdx.x);
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

XPtr xp(double* const (&a)[3]) { return XPtr{{a[0], a[1], a[2]}}; }
CXPtr cxp(double* const (&a)[3]) { return CXPtr{{a[0], a[1], a[2]}}; }

StageArgs stageOf(const DevState& s) {
  StageArgs a;
  a.cap = s.in.cap;
  a.key = s.in.stage_key;
  a.slot = s.in.stage_slot;
  a.id = s.in.stage_id;
  a.type = s.in.stage_type;
  for (int k = 0; k < 3; k++) {
    a.x[k] = s.in.stage_x[k];
    a.v[k] = s.in.stage_v[k];
  }
  return a;
}

}  // namespace

void launchUpload(const DevState& s, const AtomElement* aos_dev, int64_t sx, int64_t sy, hipStream_t st) {
  kUpload<<<siteGrid(s.G), kBlock, 0, st>>>(s.G, aos_dev, sx, sy, xp(s.x[s.cur]), xp(s.v), xp(s.f), s.type, s.id);
}

void launchHaloAll(const DevState& s, int buf, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((2 * s.G.plane + kBlock - 1) / kBlock);
  kHaloAll<<<grid, kBlock, 0, st>>>(s.G, xp(s.x[buf]), s.type, s.df64, s.df32);
}

void launchDownload(const DevState& s, AtomElement* aos_dev, int64_t sx, int64_t sy, bool mixed, hipStream_t st) {
  const unsigned grid = static_cast<unsigned>((2 * s.G.plane + kBlock - 1) / kBlock);
  kDownload<<<grid, kBlock, 0, st>>>(s.G, aos_dev, sx, sy, cxp(s.x[s.cur]), cxp(s.v), cxp(s.f), s.type, s.id,
                                     s.rho, mixed ? nullptr : s.df64, s.df32);
}

void launchRhoDf(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                 bool mixed, int acc, bool store_rho, hipStream_t st) {
  double* rho = store_rho ? s.rho : nullptr;
  const unsigned grid = siteGrid(s.G);
  if (!mixed)
    kRhoDf<double, 0><<<grid, kBlock, 0, st>>>(s.G, o, t64, cxp(s.x[s.cur]), s.type, s.df64, rho);
  else if (acc == 0)
    kRhoDf<float, 0><<<grid, kBlock, 0, st>>>(s.G, o, t32, cxp(s.x[s.cur]), s.type, s.df32, rho);
  else
    kRhoDf<float, 1><<<grid, kBlock, 0, st>>>(s.G, o, t32, cxp(s.x[s.cur]), s.type, s.df32, rho);
}

void launchForce(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                 bool mixed, int acc, int mode, const VerletParams& vp, hipStream_t st) {
  const unsigned grid = siteGrid(s.G);
  const CXPtr x = cxp(s.x[s.cur]);
  const XPtr xn = xp(s.x[1 - s.cur]);
  const StageArgs sa = stageOf(s);
#define MISA_FORCE(T, A, M, TB, DF) \
  kForce<T, A, M><<<grid, kBlock, 0, st>>>(s.G, o, TB, x, xn, s.type, DF, xp(s.v), xp(s.f), s.id, vp, sa, s.ctl)
  if (!mixed) {
    if (mode == 0) MISA_FORCE(double, 0, 0, t64, s.df64);
    else if (mode == 1) MISA_FORCE(double, 0, 1, t64, s.df64);
    else MISA_FORCE(double, 0, 2, t64, s.df64);
  } else if (acc == 0) {
    if (mode == 0) MISA_FORCE(float, 0, 0, t32, s.df32);
    else if (mode == 1) MISA_FORCE(float, 0, 1, t32, s.df32);
    else MISA_FORCE(float, 0, 2, t32, s.df32);
  } else {
    if (mode == 0) MISA_FORCE(float, 1, 0, t32, s.df32);
    else if (mode == 1) MISA_FORCE(float, 1, 1, t32, s.df32);
    else MISA_FORCE(float, 1, 2, t32, s.df32);
  }
#undef MISA_FORCE
}

void launchFirst(const DevState& s, const VerletParams& vp, hipStream_t st) {
  kFirst<<<siteGrid(s.G), kBlock, 0, st>>>(s.G, cxp(s.x[s.cur]), xp(s.x[1 - s.cur]), s.type, xp(s.v), cxp(s.f),
                                           s.id, vp, stageOf(s), s.ctl);
}

int64_t launchThermo(const DevState& s, hipStream_t st) {
  const unsigned grid = siteGrid(s.G);
  kThermo<<<grid, kBlock, 0, st>>>(s.G, cxp(s.v), cxp(s.f), s.type, s.partials);
  return grid;
}

int64_t launchEnergy(const DevState& s, const LatOffsets& o, const Tables<double>& t64, hipStream_t st) {
  const unsigned grid = siteGrid(s.G);
  kEnergy<<<grid, kBlock, 0, st>>>(s.G, o, t64, cxp(s.x[s.cur]), s.type, s.rho, s.partials);
  return grid;
}

void launchScaleV(const DevState& s, double factor, hipStream_t st) {
  kScaleV<<<siteGrid(s.G), kBlock, 0, st>>>(s.G, xp(s.v), factor);
}

void launchSetv(const DevState& s, int64_t slot, double a, double b, double c, hipStream_t st) {
  kAddV<<<1, 1, 0, st>>>(xp(s.v), slot, a, b, c);
}

}  // namespace misa
