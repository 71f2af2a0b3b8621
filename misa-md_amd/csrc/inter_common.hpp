// Device helpers for inter atoms (atoms that left their lattice site):
// Wigner-Seitz nearest site, out-of-box flags, slot decoding and the bucket
// index that maps a lattice slot to the inter atoms (and their periodic ghost
// images) whose nearest site it is.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"

namespace misa {

// The reference's full neighbour lists (NeighbourIndex even/odd,
// neighbour_index.inl:13-76) as parity-plane slot deltas; inter atoms and the
// lattice sites around them walk all of them.
// The site kernels read the offsets eight at a time (one scalar load of 32
// bytes): the arrays come first, 32-byte aligned, and the entries from n
// (nw) on are 0, the site itself.
static_assert(kMaxFullOffsets % 8 == 0, "offset rows are read eight at a time");
struct FullPlaneOffsets {
  alignas(32) int32_t off[2][kMaxFullOffsets];
  // the full-list offsets missing from the pruned lattice list (kWide sites)
  alignas(32) int32_t woff[2][kMaxFullOffsets];
  int n[2];
  int nw[2];
};

// What the site kernels need to see the inter atoms.
struct InterView {
  const int32_t* head;      // per plane slot: first sorted entry of its bucket, or -1
  const uint64_t* keys;     // sorted (slot << 24 | entry)
  const int32_t* e_src;     // entry -> local list index
  const double* e_x[3];     // entry positions (ghost images shifted)
  const int8_t* type;       // local list species
  const uint64_t* id;       // local list ids
  const double* df64;       // local list F'
  const float* df32;
  const Ctl* ctl;           // n_inter, n_ghost
  const FullPlaneOffsets* full;
};

constexpr int kEntryBits = 24;

// ws::VORONOY (ws_utils.cpp:160-187): global doubled-x lattice coordinates of
// the BCC site nearest to x.  Products with the +-1 normals are exact; sums
// are kept in the reference's order.
__device__ __forceinline__ void voronoyDev(double x0, double x1, double x2, double a, int64_t& lx, int64_t& ly,
                                           int64_t& lz) {
#pragma clang fp contract(off)
  const double q0 = x0 / a, q1 = x1 / a, q2 = x2 / a;
  lx = static_cast<int64_t>(lround(q0));
  ly = static_cast<int64_t>(lround(q1));
  lz = static_cast<int64_t>(lround(q2));
  const double dx = q0 - static_cast<double>(lx);
  const double dy = q1 - static_cast<double>(ly);
  const double dz = q2 - static_cast<double>(lz);
  const bool bx = dx > 0, by = dy > 0, bz = dz > 0;
  const double s = ((bx ? dx : -dx) + (by ? dy : -dy)) + (bz ? dz : -dz);
  lx = 2 * lx;
  if (s + (-3.0 / 4.0) >= 0.0) {
    lx += bx ? 1 : -1;
    ly += by ? 0 : -1;
    lz += bz ? 0 : -1;
  }
}

// ws::isOutBox (ws_utils.cpp:189-213) for the single-rank box at the origin.
__device__ __forceinline__ unsigned outBoxFlags(const Geom& G, double x0, double x1, double x2) {
  int64_t c0, c1, c2;
  voronoyDev(x0, x1, x2, G.a, c0, c1, c2);
  unsigned f = 0;
  if (c0 < 0) f |= 1u;
  else if (c0 >= 2 * static_cast<int64_t>(G.n0)) f |= 2u;
  if (c1 < 0) f |= 4u;
  else if (c1 >= G.n1) f |= 8u;
  if (c2 < 0) f |= 16u;
  else if (c2 >= G.n2) f |= 32u;
  return f;
}

// ghost-extended coordinates (x doubled) -> plane slot
__device__ __forceinline__ int64_t extSlot(const Geom& G, int64_t X, int64_t y, int64_t z) {
  return (X & 1) * G.plane + (z * G.NY + y) * G.NX + (X >> 1);
}

// ws::getNearLatCoord (ws_utils.cpp:254-261): nearest site in ghost-extended
// coordinates; returns its slot, or -1 outside the ghost-extended box.
__device__ __forceinline__ int64_t nearExtSlot(const Geom& G, double x0, double x1, double x2, int64_t& X, int64_t& y,
                                               int64_t& z) {
  voronoyDev(x0, x1, x2, G.a, X, y, z);
  X += 2 * G.g;
  y += G.g;
  z += G.g;
  if (X < 0 || y < 0 || z < 0 || X >= 2 * static_cast<int64_t>(G.NX) || y >= G.NY || z >= G.NZ) return -1;
  return extSlot(G, X, y, z);
}

// ws::findNearLatIndexInSubBox (ws_utils.cpp:231-252): nearest site if it is
// a sub-box site, else -1.
__device__ __forceinline__ int64_t nearSubBoxSlot(const Geom& G, double x0, double x1, double x2) {
  int64_t X, y, z;
  voronoyDev(x0, x1, x2, G.a, X, y, z);
  if (X < 0 || y < 0 || z < 0 || X >= 2 * static_cast<int64_t>(G.n0) || y >= G.n1 || z >= G.n2) return -1;
  return extSlot(G, X + 2 * G.g, y + G.g, z + G.g);
}

__device__ __forceinline__ void decodeSlot(const Geom& G, int64_t slot, int& p, int& cx, int& cy, int& cz) {
  p = slot >= G.plane ? 1 : 0;
  int64_t r = slot - p * G.plane;
  cx = static_cast<int>(r % G.NX) - G.g;
  r /= G.NX;
  cy = static_cast<int>(r % G.NY) - G.g;
  cz = static_cast<int>(r / G.NY) - G.g;
}

__device__ __forceinline__ bool isSubBoxSlot(const Geom& G, int64_t slot) {
  int p, cx, cy, cz;
  decodeSlot(G, slot, p, cx, cy, cz);
  return cx >= 0 && cx < G.n0 && cy >= 0 && cy < G.n1 && cz >= 0 && cz < G.n2;
}

template <typename T>
__device__ __forceinline__ T dfOf(const InterView& iv, int src);
template <>
__device__ __forceinline__ double dfOf<double>(const InterView& iv, int src) {
  return iv.df64[src];
}
template <>
__device__ __forceinline__ float dfOf<float>(const InterView& iv, int src) {
  return iv.df32[src];
}

// Visits the entries of the bucket of `slot` in insertion order.
template <typename Fn>
__device__ __forceinline__ void forBucket(const InterView& iv, int64_t slot, int ne, Fn fn) {
  const int h = iv.head[slot];
  if (h < 0) return;
  for (int q = h; q < ne; ++q) {
    const uint64_t k = iv.keys[q];
    if (static_cast<int64_t>(k >> kEntryBits) != slot) break;
    fn(static_cast<int>(k & ((1u << kEntryBits) - 1)));
  }
}

// Wave-uniform reads of data no kernel writes while it runs (the offset
// tables), through the scalar cache: a read through a plain pointer in a
// kernel that also stores to global memory goes through the vector memory
// path, and each such read then holds up the gather that needs it.
template <typename T>
__device__ __forceinline__ T scalarRead(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  using CT = const __attribute__((address_space(4))) T;
  return *(CT*)(p);
#else
  return *p;  // the host pass only parses device code
#endif
}
struct alignas(32) Offsets8 {
  int32_t o[8];
};
// Entries q0 .. q0 + 7 of an offset row (q0 a multiple of 8).
__device__ __forceinline__ Offsets8 offsets8(const int32_t* row, int q0) {
  return scalarRead(reinterpret_cast<const Offsets8*>(row + q0));
}

// Visits the entries of bucket `slot` in insertion order, from its head h.
template <typename Fn>
__device__ __forceinline__ void followBucket(const InterView& iv, int64_t slot, int h, int ne, Fn& fn) {
  for (int e = h; e < ne; ++e) {
    const uint64_t k = iv.keys[e];
    if (static_cast<int64_t>(k >> kEntryBits) != slot) break;
    fn(static_cast<int>(k & ((1u << kEntryBits) - 1)));
  }
}

// Visits, in full-list order (the site itself first), the entries of the
// buckets of site s (parity p, the same over the wavefront) and of its
// full-list neighbours.  The bucket heads are read eight at a time before
// any bucket is followed: nearly all are empty, and the reads then overlap
// instead of queueing one round trip behind the other.
template <typename Fn>
__device__ __forceinline__ void forNearBuckets(const InterView& iv, const FullPlaneOffsets& full, int p, int64_t s,
                                               int ne, Fn fn) {
  const int n = scalarRead(&full.n[p]);
  const int32_t* off = full.off[p];
  {
    const int h0 = iv.head[s];
    if (h0 >= 0) followBucket(iv, s, h0, ne, fn);
  }
  for (int q0 = 0; q0 < n; q0 += 8) {
    const Offsets8 o = offsets8(off, q0);
    int32_t h[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) h[u] = iv.head[s + o.o[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < n && h[u] >= 0) followBucket(iv, s + o.o[u], h[u], ne, fn);
  }
}

// Visits, in order, the full-list offsets of site s beyond the pruned list
// that hold a pair possibly within the cutoff (the site or the target a far
// lattice atom), calling fn(slot, type) for each; the target types are read
// eight at a time.
template <typename Fn>
__device__ __forceinline__ void forWideTargets(const FullPlaneOffsets& full, const int8_t* __restrict__ type, int p,
                                               int64_t s, int8_t ti, Fn fn) {
  const int n = scalarRead(&full.nw[p]);
  const int32_t* woff = full.woff[p];
  for (int q0 = 0; q0 < n; q0 += 8) {
    const Offsets8 o = offsets8(woff, q0);
    int8_t t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = type[s + o.o[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < n && wideTarget(ti, t[u])) fn(s + o.o[u], t[u]);
  }
}

}  // namespace misa
