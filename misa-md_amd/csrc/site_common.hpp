// Device helpers shared by the resident-mode site kernels: site traversal,
// periodic images, the lattice ownership test, inter-atom staging and the
// per-site accumulators.
#pragma once
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "resident.hpp"

namespace misa {
namespace site {

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

// Hardware block -> (parity, site).  Block b runs on XCD c = b % 8 (the
// dispatcher deals workgroups to the XCDs round-robin); the k-th block of XCD
// c takes parity k & 1 and local block k >> 1 of that XCD's z-slab, whose
// sites are ordered (y-tile, z, y within tile, x).
// siteOfBlock: the same for thread tl of block b (also a virtual block of a
// persistent launch, whose XCD is b % 8 as well).
__device__ __forceinline__ bool siteOfBlock(const Geom& G, uint32_t b, uint32_t tl, int64_t& s, int& p, int& cx,
                                            int& cy, int& cz) {
  const int c = static_cast<int>(b & 7);
  const uint32_t k = b >> 3;
  p = static_cast<int>(k & 1);
  const int z0 = G.W.zlo[c], nzc = G.W.zlo[c + 1] - z0;
  const uint32_t t = (k >> 1) * kBlock + tl;
  const uint32_t n0 = static_cast<uint32_t>(G.n0), ty = static_cast<uint32_t>(G.W.ty);
  if (nzc <= 0 || t >= static_cast<uint32_t>(nzc) * static_cast<uint32_t>(G.n1) * n0) return false;
  cx = static_cast<int>(t % n0);
  uint32_t r = t / n0;
  const int yy = static_cast<int>(r % ty);
  r /= ty;
  cz = z0 + static_cast<int>(r % static_cast<uint32_t>(nzc));
  cy = static_cast<int>(r / static_cast<uint32_t>(nzc)) * G.W.ty + yy;
  s = slotOf(G, p, cx, cy, cz);
  return true;
}
// Site t of the plain slot order (t < 2 nsub): parity, then z, y, x with x
// fastest, so consecutive t are consecutive slots along a row.
__device__ __forceinline__ void siteInOrder(const Geom& G, int64_t t, int64_t& s, int& p, int& cx, int& cy,
                                            int& cz) {
  p = t >= G.nsub ? 1 : 0;
  const uint32_t r = static_cast<uint32_t>(t - p * G.nsub), n0 = static_cast<uint32_t>(G.n0);
  const uint32_t row = r / n0, n1 = static_cast<uint32_t>(G.n1);
  cx = static_cast<int>(r - row * n0);
  cz = static_cast<int>(row / n1);
  cy = static_cast<int>(row - static_cast<uint32_t>(cz) * n1);
  s = slotOf(G, p, cx, cy, cz);
}
__device__ __forceinline__ bool siteOfThread(const Geom& G, int64_t& s, int& p, int& cx, int& cy, int& cz) {
  return siteOfBlock(G, blockIdx.x, threadIdx.x, s, p, cx, cy, cz);
}

// Persistent launches of the passes that hold tables in LDS: blocks of
// kBlockP threads, a whole number of blocks per XCD (the grid is a multiple
// of 8), each walking virtual blocks of kBlock threads of its own XCD.  A
// virtual block covers the sites, and its wavefronts carry the ids, of the
// block of the same index in a launch of siteGrid blocks, so per-wavefront
// data (the wave masks) carries over between the two kinds of launch.
constexpr int kBlockP = 1024;
template <typename Body>
__device__ __forceinline__ void forVirtualBlocks(const Geom& G, Body body) {
  constexpr uint32_t V = kBlockP / kBlock;
  const uint32_t c = blockIdx.x & 7u, q = blockIdx.x >> 3, Q = gridDim.x >> 3;
  const uint32_t sub = __builtin_amdgcn_readfirstlane(threadIdx.x / kBlock);
  const uint32_t tl = threadIdx.x % kBlock;
  const uint32_t nvb = 2u * static_cast<uint32_t>(G.W.bpx);
  for (uint32_t k = q * V + sub; k < nvb; k += Q * V) body(c + 8u * k, tl);
}

inline unsigned siteGrid(const Geom& G) { return static_cast<unsigned>(16 * G.W.bpx); }
// Grid of a pass that strides over the siteGrid virtual blocks: up to 8
// blocks per compute unit, a whole number per XCD.
inline unsigned strideGrid(const Geom& G, int ncu) {
  const unsigned g = static_cast<unsigned>(ncu >= 8 ? ncu / 8 * 8 * 8 : 8);
  return siteGrid(G) < g ? siteGrid(G) : g;
}

// Work per z-plane of a pass (the data the XCD slabs are balanced by): the
// shader clock cycles a wavefront spent on its part of virtual block b, added
// to the plane of the wavefront's first site (empty wavefronts add nothing).
__device__ __forceinline__ void addPlaneWork(const Geom& G, unsigned long long* zw, uint32_t b, uint32_t tl,
                                             long long t0) {
  const long long t1 = clock64();
  int64_t s;
  int p, cx, cy, cz;
  const uint32_t t = __builtin_amdgcn_readfirstlane(tl & ~63u);
  if (siteOfBlock(G, b, t, s, p, cx, cy, cz) && (threadIdx.x & 63u) == 0)
    atomicAdd(zw + cz, static_cast<unsigned long long>(t1 - t0));
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
__device__ __forceinline__ bool leftSite(const Geom& G, int p, int cx, int cy, int cz, double x0, double x1,
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

// Site records of the mixed-precision passes: one 16-byte record per slot,
// {ux, uy, uz, w}.  u is the FP32 displacement of the atom from its lattice
// site (taken in FP64, then rounded), with the species in the two low
// mantissa bits of ux (a change of at most 3 ulp); w is F'(rho) of the site,
// written by the density pass.  A vacancy has a displacement of 10^4 A along
// each axis, finite so that a term multiplied by a zero pair factor stays
// zero, and so far away that every distance test against it fails.
// Periodic images carry the same record.
__device__ __forceinline__ void sitePos(const Geom& G, int p, int cx, int cy, int cz, double& X, double& Y,
                                        double& Z) {
#pragma clang fp contract(off)
  const double h = 0.5 * p;
  X = (cx + h) * G.a;
  Y = (cy + h) * G.a;
  Z = (cz + h) * G.a;
}
__device__ __forceinline__ float4 packRec(double u0, double u1, double u2, int sp, float w) {
  const uint32_t bx = (__float_as_uint(static_cast<float>(u0)) & ~3u) | static_cast<uint32_t>(sp);
  return make_float4(__uint_as_float(bx), static_cast<float>(u1), static_cast<float>(u2), w);
}
constexpr float kVacantDisp = 1.0e4f;
__device__ __forceinline__ float4 vacantRec() { return make_float4(kVacantDisp, kVacantDisp, kVacantDisp, 0.0f); }
__device__ __forceinline__ int recSpecies(float ux) { return static_cast<int>(__float_as_uint(ux) & 3u); }
// Record of the atom at x on site (p, cx, cy, cz) of species sp.
__device__ __forceinline__ float4 siteRec(const Geom& G, int p, int cx, int cy, int cz, double x0, double x1,
                                          double x2, int sp) {
#pragma clang fp contract(off)  // one rounding of the site position, one of the difference
  double X, Y, Z;
  sitePos(G, p, cx, cy, cz, X, Y, Z);
  return packRec(x0 - X, x1 - Y, x2 - Z, sp, 0.0f);
}

// Site records of the FP64 passes: per slot {x, y} and {z, w} in two arrays
// of 16-byte halves (a neighbour costs two 16-byte loads, at the offsets the
// mixed records use).  x, y, z are the atom's position as the position
// arrays hold it; w is F'(rho) of the site with the species in its two low
// mantissa bits (a change of at most 3 ulp), written by the density pass.
// A vacancy sits 10^6 A out (finite, beyond every cutoff); periodic images
// carry their own position and the site's w.
struct Rec64 {
  double2* xy;
  double2* zw;
};
constexpr double kVacantPos64 = 1.0e6;
__device__ __forceinline__ double packW64(double df, int sp) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(df));
  return __longlong_as_double(static_cast<long long>((b & ~3ull) | static_cast<unsigned long long>(sp)));
}
__device__ __forceinline__ int recSpecies64(double w) {
  return static_cast<int>(static_cast<unsigned long long>(__double_as_longlong(w)) & 3ull);
}
// The record of slot q for an atom of type t at (x0, x1, x2), F' not yet known.
__device__ __forceinline__ void putRec64(const Rec64& r, int64_t q, int8_t t, double x0, double x1, double x2) {
  if (isVacant(t)) {
    r.xy[q] = make_double2(kVacantPos64, kVacantPos64);
    r.zw[q] = make_double2(kVacantPos64, 0.0);
  } else {
    r.xy[q] = make_double2(x0, x1);
    r.zw[q] = make_double2(x2, packW64(0.0, species(t)));
  }
}

// |v|^2 with each product and sum rounded separately; non-negative doubles
// order like their bit patterns, so an integer maximum finds the largest.
__device__ __forceinline__ unsigned long long speed2Bits(double a, double b, double c) {
#pragma clang fp contract(off)
  const double v2 = (a * a + b * b) + c * c;
  return static_cast<unsigned long long>(__double_as_longlong(v2));
}

// Maximum of b over the 64 lanes of the wavefront into *dst; every lane of
// the wavefront must reach the call.  The word only grows, so a wavefront
// whose maximum does not exceed the value it reads skips the atomic.
__device__ __forceinline__ void waveMaxTo(unsigned long long b, unsigned long long* dst) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(b, o, 64);
    b = y > b ? y : b;
  }
  if ((threadIdx.x & 63) == 0 && b > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(dst, b);
}

// Step length and half-kick factor of species t for the current step: the
// fixed values, or those of the adaptive step in device memory, computed as
// NewtonMotion::preComputeDtInv2m does (newton_motion.cpp:19-28).
struct StepDt {
  double dt;
  double c;
};
__device__ __forceinline__ StepDt stepDt(const VerletParams& vp, int t) {
#pragma clang fp contract(off)
  StepDt r;
  if (!vp.dtp) {
    r.dt = vp.dt;
    r.c = t == 0 ? vp.c[0] : (t == 1 ? vp.c[1] : vp.c[2]);
    return r;
  }
  r.dt = *vp.dtp;
  const double h = 0.5 * r.dt * kFtm2vDev;
  r.c = h / speciesMass(t);
  return r;
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
  // one counter update per wavefront: the lanes that stage (the active ones
  // here) take consecutive entries; the decide orders the entries by key
  const uint64_t lanes = __ballot(1);
  const int leader = __builtin_ctzll(lanes);
  const int rank = __builtin_popcountll(lanes & ((1ull << (threadIdx.x & 63)) - 1ull));
  unsigned long long base = 0;
  if (static_cast<int>(threadIdx.x & 63) == leader)
    base = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->n_stage),
                     static_cast<unsigned long long>(__builtin_popcountll(lanes)));
  base = __shfl(base, leader, 64);
  const unsigned long long k = base + static_cast<unsigned long long>(rank);
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
#pragma clang fp contract(off)  // the compensation must see v rounded, in every instantiation
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return static_cast<double>(s); }
};

// plain FP32 sums (ACC 2)
template <>
struct Sum1<2> {
  float s = 0;
  __device__ __forceinline__ void add(float v) { s += v; }
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
  float g0 = 0, g1 = 0, g2 = 0;  // the open chunk of lattice terms (addGroup3)
  int chunk = 0;                 // its offsets: 8 chunk .. 8 chunk + 7
  __device__ __forceinline__ void add(double d0, double d1, double d2, float fp) {
#pragma clang fp contract(off)
    a.add(static_cast<float>(d0) * fp);
    b.add(static_cast<float>(d1) * fp);
    c.add(static_cast<float>(d2) * fp);
  }
  // a chunk's sum is exactly zero only as the lane's own terms make it
  __device__ __forceinline__ void flush() {
    if (g0 != 0.0f) a.add(g0);
    if (g1 != 0.0f) b.add(g1);
    if (g2 != 0.0f) c.add(g2);
    g0 = g1 = g2 = 0.0f;
  }
};
template <>
struct Sum3<2> {
  float f0 = 0, f1 = 0, f2 = 0;
  __device__ __forceinline__ void add(float d0, float d1, float d2, float fp) {
    f0 = fmaf(d0, fp, f0);
    f1 = fmaf(d1, fp, f1);
    f2 = fmaf(d2, fp, f2);
  }
};
// The force terms of a group of lattice neighbours (in[u]: neighbour u
// counts; ks[u]: its offset, ascending over a site's groups).  With
// compensated sums (ACC 1) the terms of the offsets 8t .. 8t+7 are summed in
// FP32 (fused multiply-adds) and that chunk's sum joins the compensated sum
// as one term, when the offsets move past the chunk and once more at the end
// (flushChunk3); a lane adds no empty chunk.  The chunks hang on the offsets
// alone, not on how they are grouped for the gathers nor on which other
// sites share the wavefront, so every launch shape gives the same bits.
// Otherwise term by term.
template <int ACC, int V>
__device__ __forceinline__ void addGroup3(Sum3<ACC>& acc, const int* ks, const bool (&in)[V], const float (&d0)[V],
                                          const float (&d1)[V], const float (&d2)[V], const float (&fp)[V]) {
#pragma clang fp contract(off)
  if constexpr (ACC == 1) {
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int t = ks[u] >> 3;  // the same in every lane
      if (t != acc.chunk) {
        acc.flush();
        acc.chunk = t;
      }
      // fp is 0 beyond the cutoff and the displacements are finite (a
      // vacancy's too), so those terms add an exact zero
      acc.g0 = fmaf(d0[u], fp[u], acc.g0);
      acc.g1 = fmaf(d1[u], fp[u], acc.g1);
      acc.g2 = fmaf(d2[u], fp[u], acc.g2);
    }
  } else {
#pragma unroll
    for (int u = 0; u < V; ++u)
      if (in[u]) acc.add(d0[u], d1[u], d2[u], fp[u]);
  }
}
template <int ACC>
__device__ __forceinline__ void flushChunk3(Sum3<ACC>& acc) {
  if constexpr (ACC == 1) acc.flush();
}

template <int ACC>
__device__ __forceinline__ void sum3Value(const Sum3<ACC>& s, double& f0, double& f1, double& f2);
template <>
__device__ __forceinline__ void sum3Value<0>(const Sum3<0>& s, double& f0, double& f1, double& f2) {
  f0 = s.f0;
  f1 = s.f1;
  f2 = s.f2;
}
template <>
__device__ __forceinline__ void sum3Value<2>(const Sum3<2>& s, double& f0, double& f1, double& f2) {
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

__host__ __device__ inline XPtr xp(double* const (&a)[3]) { return XPtr{{a[0], a[1], a[2]}}; }
__host__ __device__ inline CXPtr cxp(double* const (&a)[3]) { return CXPtr{{a[0], a[1], a[2]}}; }

// One box of a batch whose passes run as one launch for all boxes: what the
// single-box launches of the step take as arguments, with the position and
// record buffers of both parities of the step (cur is the buffer the first
// captured step reads; every step flips it).
struct BatchBox {
  Geom G;
  LatOffsets O;
  Tables<float> t32;
  InterView iv;
  InterDev in;
  double* x[2][3];
  double* v[3];
  double* f[3];
  int8_t* type;
  uint64_t* id;
  float4* rec[2];
  const float4* dl;
  uint64_t* wmask;
  uint64_t* smask;
  float rcs2;
  double skin;
  VerletParams vp;
  StageArgs st;
  Ctl* ctl;
  int cur;
};

// The box of block `blk` when box b owns blocks [start[b], start[b+1]).
__device__ __forceinline__ int boxOfBlock(const uint32_t* start, int nbox, uint32_t blk) {
  int lo = 0, hi = nbox - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

inline StageArgs stageOf(const DevState& s) {
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


}  // namespace site
}  // namespace misa
