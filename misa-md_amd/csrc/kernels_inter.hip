// Inter atoms on the device: ownership decisions, migration, periodic ghost
// images, the nearest-site bucket index, and their density, F', force and
// velocity-Verlet steps.
//
// Reference behaviour restated (single rank, periodic box):
//   atom::decide            atom.cpp:21-84
//   InterAtomList           inter_atom_list.cpp:16-75 (list order, makeIndex)
//   exchangeInter           inter_particle_packer.cpp:17-121
//   borderInter             inter_border_packer.cpp:12-106
//   atom::interRho          atom.cpp:194-284
//   atom::interForce        atom.cpp:360-473
//   NewtonMotion (inter)    newton_motion.cpp:46-54, 68-73
//
// The list is kept in the reference's order: new inter atoms are appended in
// the z->y->x order of the sites they left, recombined atoms are removed in
// place, migrating atoms are moved to the end (dimension by dimension, the
// atoms sent to the lower neighbour before those sent to the upper one), and
// ghost images follow the local atoms in the order borderInter appends them.
// Bucket order is insertion order, so every per-atom sum is taken in the
// reference's order.  Recombination claims a vacancy for the lowest list
// index among the atoms that target it, which is what the reference's
// sequential loop yields, since vacancies are only created before that loop.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>

#include "device_common.hpp"
#include "inter_common.hpp"
#include "resident.hpp"
#include "site_common.hpp"

namespace misa {

namespace {

using namespace site;

constexpr int kDT = 1024;  // threads of the single-workgroup decide kernel
constexpr int kIT = 256;   // threads per block of the per-atom kernels
constexpr int kIB = kInterBlocks;  // blocks of the per-atom kernels (grid-stride)

// Rounding allowance of the neighbour skin's displacement bound (A).
constexpr double kSkinMargin = 1e-3;

struct DecideArgs {
  Geom G;
  int mode;
  int cur;
  double* lx[3];  // lattice positions (current buffer)
  double* lv[3];
  int8_t* ltype;
  uint64_t* lid;
  float4* lrec;  // mixed precision: site records of the current buffer, else null
  Rec64 lrec64;  // FP64: site records of the current buffer, else null
  InterDev in;
  Ctl* ctl;
  double skin;  // neighbour skin of the mixed density pass (0: off)
};

// Exclusive block-wide prefix sum of v (1024 threads); `total` gets the sum.
__device__ int blockScan(int v, int& total) {
  __shared__ int wsum[kDT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (w == 0) {
    int t = lane < kDT / 64 ? wsum[lane] : 0;
#pragma unroll
    for (int o = 1; o < kDT / 64; o <<= 1) {
      const int y = __shfl_up(t, o, 64);
      if (lane >= o) t += y;
    }
    if (lane < kDT / 64) wsum[lane] = t;
  }
  __syncthreads();
  const int before = w > 0 ? wsum[w - 1] : 0;
  total = wsum[kDT / 64 - 1];
  __syncthreads();
  return before + x - v;
}

__device__ __forceinline__ void setFlag(unsigned long long* f, unsigned long long bit) { atomicOr(f, bit); }

// Copy local atom src (main arrays) to scratch slot dst, shifting x[d] by sh.
__device__ __forceinline__ void toScratch(InterDev& in, int src, int dst, int d, double sh) {
  in.id2[dst] = in.id[src];
  in.type2[dst] = in.type[src];
  for (int k = 0; k < 3; k++) {
    in.x2[k][dst] = in.x[k][src] + (k == d ? sh : 0.0);
    in.v2[k][dst] = in.v[k][src];
  }
}

__device__ __forceinline__ void fromScratch(InterDev& in, int n) {
  for (int i = threadIdx.x; i < n; i += kDT) {
    in.id[i] = in.id2[i];
    in.type[i] = in.type2[i];
    for (int k = 0; k < 3; k++) {
      in.x[k][i] = in.x2[k][i];
      in.v[k][i] = in.v2[k][i];
    }
  }
  __syncthreads();
}

// Forwarding region of borderInter in ghost-extended coordinates (x doubled):
// the ghost-wide layer of the sub-box next to face (dim, dir); dimensions
// already exchanged span the whole ghost-extended range.
__device__ __forceinline__ bool inForwardRegion(const Geom& G, int dim, int dir, int64_t X, int64_t y, int64_t z) {
  const int64_t c[3] = {X, y, z};
  const int64_t g[3] = {2 * G.g, G.g, G.g};
  const int64_t s[3] = {2 * static_cast<int64_t>(G.n0), G.n1, G.n2};
  const int64_t e[3] = {2 * static_cast<int64_t>(G.NX), G.NY, G.NZ};
  for (int k = 0; k < 3; k++) {
    int64_t lo, hi;
    if (k == dim) {
      lo = dir == 0 ? g[k] : s[k];
      hi = dir == 0 ? 2 * g[k] : s[k] + g[k];
    } else if (k < dim) {
      lo = 0;
      hi = e[k];
    } else {
      lo = g[k];
      hi = g[k] + s[k];
    }
    if (c[k] < lo || c[k] >= hi) return false;
  }
  return true;
}

// Type (and, for `pos`, position) of a lattice slot to its periodic images;
// in mixed precision also the site record: vacantRec for a vacancy, else the
// displacement of the (new) atom from its site.
__device__ __forceinline__ void writeLatticeImages(const DecideArgs& A, int64_t slot, bool pos, int8_t t) {
  int p, cx, cy, cz;
  decodeSlot(A.G, slot, p, cx, cy, cz);
  const double y0 = A.lx[0][slot], y1 = A.lx[1][slot], y2 = A.lx[2][slot];
  float4 r = vacantRec();
  if (A.lrec) {
    if (!isVacant(t)) r = siteRec(A.G, p, cx, cy, cz, y0, y1, y2, species(t));
    A.lrec[slot] = r;
  }
  if (A.lrec64.xy) putRec64(A.lrec64, slot, t, y0, y1, y2);
  forEachImage(A.G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
    A.ltype[q] = t;
    if (pos) {
      A.lx[0][q] = y0 + sx * A.G.L0;
      A.lx[1][q] = y1 + sy * A.G.L1;
      A.lx[2][q] = y2 + sz * A.G.L2;
    }
    if (A.lrec) A.lrec[q] = r;
    if (A.lrec64.xy) putRec64(A.lrec64, q, t, A.lx[0][q], A.lx[1][q], A.lx[2][q]);
  });
}

// fn(slot) for a sub-box slot and every periodic image of it in the ghost layer.
template <typename Fn>
__device__ __forceinline__ void forSlotAndImages(const Geom& G, int64_t slot, Fn fn) {
  int p, cx, cy, cz;
  decodeSlot(G, slot, p, cx, cy, cz);
  fn(slot);
  forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) { fn(q); });
}

// Ascending block sort of n 64-bit keys at a (LDS or global memory): a
// bitonic network over the next power of two, padded with ~0.
__device__ void blockSortU64(uint64_t* a, int n) {
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = n + static_cast<int>(threadIdx.x); i < P; i += kDT) a[i] = ~0ull;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += kDT) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const uint64_t x = a[i], y = a[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            a[i] = y;
            a[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
}

// Keys sorted in LDS when they fit (kSortLds of them), else in place in
// global memory.
constexpr int kSortLds = 8192;
__device__ void sortKeys(uint64_t* g, int n, uint64_t* lds) {
  int P = 1;
  while (P < n) P <<= 1;
  if (P > kSortLds) {
    blockSortU64(g, n);
    return;
  }
  for (int i = threadIdx.x; i < n; i += kDT) lds[i] = g[i];
  __syncthreads();
  blockSortU64(lds, n);
  for (int i = threadIdx.x; i < n; i += kDT) g[i] = lds[i];
  __syncthreads();
}

// The 8-bit type array in 32-bit words: set or clear bits of one byte with a
// word atomic, so that many threads may flag neighbouring sites at once.
__device__ __forceinline__ void typeBitsOr(int8_t* t, int64_t i, int8_t bits) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(t + i);
  const unsigned sh = static_cast<unsigned>(a & 3u) * 8u;
  atomicOr(reinterpret_cast<unsigned*>(a & ~uintptr_t(3)), static_cast<unsigned>(static_cast<uint8_t>(bits)) << sh);
}
__device__ __forceinline__ void typeBitsAnd(int8_t* t, int64_t i, int8_t keep) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(t + i);
  const unsigned sh = static_cast<unsigned>(a & 3u) * 8u;
  const unsigned m = ~(static_cast<unsigned>(static_cast<uint8_t>(~keep)) << sh);
  atomicAnd(reinterpret_cast<unsigned*>(a & ~uintptr_t(3)), m);
}

// Calls fn(t) for neighbour k (-1: the slot itself) of `slot` in the full
// list when that neighbour is a sub-box site.
template <typename Fn>
__device__ __forceinline__ void fullNeighbour(const Geom& G, const FullPlaneOffsets& full, int64_t slot, int k,
                                              Fn fn) {
  const int par = slot >= G.plane ? 1 : 0;
  if (k >= full.n[par]) return;
  const int64_t t = k < 0 ? slot : slot + full.off[par][k];
  if (t >= 0 && t < 2 * G.plane && isSubBoxSlot(G, t)) fn(t);
}

// Grid of the flag kernels: grid-stride over (entry, neighbour) items whose
// count is known only on the device.
constexpr int kFlagBlocks = 1024;

// Decide, step 1 (before the list work): the previous index's bucket heads
// and the flags it set, kNearInter around its entries and kWide around its
// far lattice atoms and their images (atom.cpp:21-84 rebuilds the index each
// step; inter_atom_list.cpp:27-45).
__device__ __forceinline__ void unflagBody(const Geom& G, const InterDev& in, int8_t* ltype, const Ctl* ctl,
                                           int64_t t0, int64_t stride) {
  const int nprev = static_cast<int>(ctl->n_prev);
  const int nfarPrev = static_cast<int>(ctl->n_far_prev);
  const FullPlaneOffsets& full = *in.full;
  const int nfull = 1 + max(full.n[0], full.n[1]);
  for (int64_t q = t0; q < nprev; q += stride) in.head[in.prev_slot[q]] = -1;
  for (int64_t q = t0; q < static_cast<int64_t>(nprev) * nfull; q += stride) {
    const int e = static_cast<int>(q / nfull), k = static_cast<int>(q % nfull) - 1;
    fullNeighbour(G, full, in.prev_slot[e], k, [&](int64_t t) {
      if (ltype[t] & kNearInter) typeBitsAnd(ltype, t, static_cast<int8_t>(~kNearInter));
    });
  }
  for (int64_t q = t0; q < static_cast<int64_t>(nfarPrev) * nfull; q += stride) {
    const int e = static_cast<int>(q / nfull), k = static_cast<int>(q % nfull) - 1;
    forSlotAndImages(G, in.far_prev[e], [&](int64_t slot) {
      fullNeighbour(G, full, slot, k, [&](int64_t t) {
        const int8_t v = ltype[t];
        if (!isVacant(v) && (v & kWide)) typeBitsAnd(ltype, t, static_cast<int8_t>(~kWide));
      });
    });
  }
  for (int64_t q = t0; q < nfarPrev; q += stride)
    forSlotAndImages(G, in.far_prev[q], [&](int64_t slot) {
      const int8_t v = ltype[slot];
      if (!isVacant(v) && (v & kFar)) typeBitsAnd(ltype, slot, static_cast<int8_t>(~kFar));
    });
}
__global__ __launch_bounds__(kBlock) void kUnflag(Geom G, InterDev in, int8_t* ltype, const Ctl* ctl) {
  unflagBody(G, in, ltype, ctl, static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x,
             static_cast<int64_t>(gridDim.x) * kBlock);
}

// Priority of a staged atom's claim on a vacated site (above every list
// index, ordered by the reference index of the site it left), and the stage
// key of an entry that kQuickRefill has put back.
constexpr int32_t kStagedPri = 1 << 30;
constexpr int64_t kQuickDone = (int64_t(1) << 42) - 1;
constexpr int kQuickBlocks = 128;

// Decide, step 2: the list work in one workgroup (the lists hold the
// interstitials and the atoms that moved more than 0.2a this step: hundreds
// to thousands in a developed cascade).
__device__ __forceinline__ void decideBody(const DecideArgs& A) {
  const Geom& G = A.G;
  InterDev in = A.in;
  const int tid = threadIdx.x;
  int nl = static_cast<int>(A.ctl->n_inter);
  __shared__ int sFar;   // far lattice atoms for this index (kFarScan's in mode 1)
  __shared__ int sNear;  // refilled atoms within 0.2a of their site
  __shared__ uint64_t sKeys[kSortLds];
  if (tid == 0) {
    sFar = static_cast<int>(min(A.ctl->n_far, in.cap));
    sNear = 0;
  }
  __syncthreads();
  unsigned long long* err = reinterpret_cast<unsigned long long*>(&A.ctl->err_flags);
  int refills = 0, staged = 0;

  if (A.mode == 0) {
    // P1: append the new inter atoms in the order of the sites they left
    // (atom.cpp:28-53 walks the sub-box z -> y -> x and appends): sort the
    // staged atoms by the reference linear index of their site.  Entries
    // that kQuickRefill already put back (key kQuickDone) are left out.
    const int nsAll = static_cast<int>(min(A.ctl->n_stage, in.cap));
    staged = nsAll;
    refills = static_cast<int>(A.ctl->n_quick);
    int ns = 0;  // the entries still staged, compacted before the sort
    for (int base = 0; base < nsAll; base += kDT) {
      const int i = base + tid;
      const int live = (i < nsAll && in.stage_key[i] != kQuickDone) ? 1 : 0;
      int tot;
      const int ex = blockScan(live, tot);
      if (live) in.keys[ns + ex] = (static_cast<uint64_t>(in.stage_key[i]) << 21) | static_cast<uint64_t>(i);
      ns += tot;
    }
    __syncthreads();
    sortKeys(in.keys, ns, sKeys);
    if (nl + ns > in.cap) {
      if (tid == 0) setFlag(err, 1ull);
      ns = static_cast<int>(in.cap) - nl;
    }
    for (int r = tid; r < ns; r += kDT) {
      const int i = static_cast<int>(in.keys[r] & ((1u << 21) - 1));
      const int dst = nl + r;
      in.id[dst] = in.stage_id[i];
      in.type[dst] = static_cast<int8_t>(species(in.stage_type[i]));
      for (int k = 0; k < 3; k++) {
        in.x[k][dst] = in.stage_x[k][i];
        in.v[k][dst] = in.stage_v[k][i];
        in.f[k][dst] = 0.0;
      }
      const int64_t slot = in.stage_slot[i];
      A.ltype[slot] = kVacant;  // type INVALID; its stale x stays (atom.cpp:43-47)
      writeLatticeImages(A, slot, false, kVacant);
    }
    __syncthreads();
    nl += ns;

    // P2: an inter atom whose Wigner-Seitz site is a vacant sub-box site whose
    // stale position is in the box refills it (atom.cpp:55-83)
    for (int i = tid; i < nl; i += kDT) {
      const int64_t t = nearSubBoxSlot(G, in.x[0][i], in.x[1][i], in.x[2][i]);
      int target = -1;
      if (t >= 0 && isVacant(A.ltype[t]) && outBoxFlags(G, A.lx[0][t], A.lx[1][t], A.lx[2][t]) == 0u) {
        atomicMin(&in.site_claim[t], i);
        target = static_cast<int>(t);
      }
      in.tmp[i] = target;
    }
    __syncthreads();
    for (int i = tid; i < nl; i += kDT) {
      const int t = in.tmp[i];
      if (t >= 0 && in.site_claim[t] == i) {
        A.lid[t] = in.id[i];
        A.ltype[t] = in.type[i];
        for (int k = 0; k < 3; k++) {
          A.lx[k][t] = in.x[k][i];
          A.lv[k][t] = in.v[k][i];
        }
        writeLatticeImages(A, t, true, in.type[i]);
        in.tmp[i] = -3 - t;  // removed; remembers its site for the claim reset
        int p, cx, cy, cz;
        decodeSlot(G, t, p, cx, cy, cz);
        if (leftSite(G, p, cx, cy, cz, in.x[0][i], in.x[1][i], in.x[2][i])) {
          const int q = atomicAdd(&sFar, 1);
          if (q < in.cap) in.far_new[q] = t;
          else setFlag(err, 1ull);
        } else {
          atomicAdd(&sNear, 1);
        }
      }
    }
    __syncthreads();
    // release every claim (winners' and losers') for the next step
    for (int i = tid; i < nl; i += kDT) {
      const int t = in.tmp[i];
      if (t >= 0) in.site_claim[t] = INT_MAX;
      else if (t <= -3) in.site_claim[-3 - t] = INT_MAX;
    }
    __syncthreads();

    // P3: remove the recombined atoms, keeping the list order
    int kept = 0;
    for (int base = 0; base < nl; base += kDT) {
      const int i = base + tid;
      const int keep = (i < nl && in.tmp[i] > -3) ? 1 : 0;
      int tot;
      const int ex = blockScan(keep, tot);
      if (keep) toScratch(in, i, kept + ex, -1, 0.0);
      kept += tot;
    }
    __syncthreads();
    refills += nl - kept;
    nl = kept;
    fromScratch(in, nl);
  }

  // P4: exchangeInter: per dimension, atoms out of the box below / above
  // move to the end of the list, shifted by +L / -L (inter_particle_packer.cpp)
  const double L[3] = {G.L0, G.L1, G.L2};
  for (int d = 0; d < 3; d++) {
    const unsigned lowBit = 1u << (2 * d), highBit = 2u << (2 * d);
    for (int i = tid; i < nl; i += kDT) {
      const unsigned fl = outBoxFlags(G, in.x[0][i], in.x[1][i], in.x[2][i]);
      in.tmp[i] = (fl & lowBit) ? 1 : ((fl & highBit) ? 2 : 0);
    }
    __syncthreads();
    int cnt[3] = {0, 0, 0};
    for (int grp = 0; grp < 3; grp++) {
      int run = 0;
      for (int base = 0; base < nl; base += kDT) {
        const int i = base + tid;
        const int is = (i < nl && in.tmp[i] == grp) ? 1 : 0;
        int tot;
        const int ex = blockScan(is, tot);
        if (is) {
          const int dst = (grp == 0 ? 0 : (grp == 1 ? cnt[0] : cnt[0] + cnt[1])) + run + ex;
          toScratch(in, i, dst, grp == 0 ? -1 : d, grp == 1 ? L[d] : (grp == 2 ? -L[d] : 0.0));
        }
        run += tot;
      }
      cnt[grp] = run;
    }
    __syncthreads();
    fromScratch(in, nl);
  }

  // P5: entries = local atoms, then borderInter's ghost images
  for (int i = tid; i < nl; i += kDT) {
    in.e_src[i] = i;
    for (int k = 0; k < 3; k++) in.e_x[k][i] = in.x[k][i];
  }
  __syncthreads();
  int ne = nl;
  for (int d = 0; d < 3; d++) {
    const int ne0 = ne;
    for (int dir = 0; dir < 2; dir++) {
      const double sh = dir == 0 ? L[d] : -L[d];
      for (int base = 0; base < ne0; base += kDT) {
        const int e = base + tid;
        int is = 0;
        if (e < ne0) {
          int64_t X, y, z;
          voronoyDev(in.e_x[0][e], in.e_x[1][e], in.e_x[2][e], G.a, X, y, z);
          is = inForwardRegion(G, d, dir, X + 2 * G.g, y + G.g, z + G.g) ? 1 : 0;
        }
        int tot;
        const int ex = blockScan(is, tot);
        if (is) {
          const int dst = ne + ex;
          if (dst < in.ecap) {
            in.e_src[dst] = in.e_src[e];
            for (int k = 0; k < 3; k++) in.e_x[k][dst] = in.e_x[k][e] + (k == d ? sh : 0.0);
          } else {
            setFlag(err, 2ull);
          }
        }
        ne += tot;
      }
      if (ne > in.ecap) ne = static_cast<int>(in.ecap);
      __syncthreads();
    }
  }

  // P6: the index: keys = (slot << 24 | entry) sorted, so buckets come in slot
  // order and each bucket in insertion order (inter_atom_list.cpp:27-45)
  const int nfar = min(sFar, static_cast<int>(in.cap));
  for (int e = tid; e < ne; e += kDT) {
    int64_t X, y, z;
    int64_t slot = nearExtSlot(G, in.e_x[0][e], in.e_x[1][e], in.e_x[2][e], X, y, z);
    if (slot < 0) {
      setFlag(err, 8ull);
      slot = 0;
    }
    in.e_slot[e] = slot;
    in.keys[e] = (static_cast<uint64_t>(slot) << kEntryBits) | static_cast<uint64_t>(e);
  }
  __syncthreads();
  sortKeys(in.keys, ne, sKeys);
  for (int q = tid; q < ne; q += kDT) {
    const int64_t slot = static_cast<int64_t>(in.keys[q] >> kEntryBits);
    if (q == 0 || static_cast<int64_t>(in.keys[q - 1] >> kEntryBits) != slot) in.head[slot] = q;
    in.prev_slot[q] = slot;
  }
  __syncthreads();
  if (tid == 0) {
    A.ctl->n_inter = nl;
    A.ctl->n_ghost = ne - nl;
    A.ctl->n_stage = 0;
    A.ctl->n_quick = 0;
    A.ctl->n_prev = ne;
    A.ctl->n_far_prev = nfar;
    A.ctl->n_far = 0;
    A.ctl->staged_total += staged;
    A.ctl->refill_total += refills;
    A.ctl->far_total += A.mode == 0 ? nfar : 0;
    // Neighbour skin of the density pass (both precisions).  Since the last full scan no
    // lattice atom has moved farther than disp_acc, so no pair has come closer
    // by more than 2 disp_acc: an offset beyond cutoff + skin for every lane
    // then is beyond the cutoff now while disp_acc < skin / 2.  A refill puts
    // a new atom on a site: when it lands farther than 0.2a from the site, its
    // whole neighbourhood carries kWide, and the wavefronts holding such
    // sites scan in full and record their masks afresh; any other refill, and
    // the first force pass after an upload (mode 1), need a full scan of all.
    Ctl* c = A.ctl;
    const double acc = c->disp_acc + sqrt(__longlong_as_double(static_cast<long long>(c->dispmax2)));
    const bool full = A.mode != 0 || !(A.skin > 0.0) || c->rescan != 0 || sNear > 0 ||
                      acc >= 0.5 * A.skin - kSkinMargin;
    c->skin_mode = full ? 0 : 1;
    c->disp_acc = full ? 0.0 : acc;
    c->dispmax2 = 0ull;
    c->rescan = 0;
  }
}
__global__ __launch_bounds__(kDT) void kDecide(DecideArgs A) { decideBody(A); }

// Decide, ahead of the list work: the staged atoms that the recombination
// loop would put straight back on the site they left (in a cascade, the
// thousands a step pushed beyond 0.2a that stay within their site's
// Wigner-Seitz cell).  atom.cpp:28-83 appends such an atom, finds its own
// vacated site nearest, and refills it with the first list entry that
// targets it.  The entries that can target a vacated site are the old inter
// atoms (list indices 0..n_inter-1) and the staged atoms, which follow them
// in the order of the reference index of the sites they left.  So the staged
// atom refills its own site exactly when no old inter atom and no staged
// atom of a lower site index targets that site as well.  Such an atom is put
// back here with everything decide writes for it (type, id, position,
// velocity, images, record, the far list); decide skips its stage entry,
// and its list work, list order and claims come out as before.
//   kQuickClaim:   a staged atom whose nearest site is its own, with its
//                  position in the box, opens a claim on that site with its
//                  priority (kStagedPri + site index);
//   kQuickCompete: every other entry whose nearest site carries a claim
//                  lowers it to its own priority (old inter atoms: their
//                  list index);
//   kQuickRefill:  a staged atom still holding its site's claim is put back;
//                  every opened claim is released.

__device__ __forceinline__ int quickStaged(const InterDev& in, const Ctl* ctl) {
  return static_cast<int>(min(ctl->n_stage, in.cap));
}
__device__ __forceinline__ bool ownSiteNearest(const Geom& G, const InterDev& in, int i) {
  const double y0 = in.stage_x[0][i], y1 = in.stage_x[1][i], y2 = in.stage_x[2][i];
  return nearSubBoxSlot(G, y0, y1, y2) == in.stage_slot[i] && outBoxFlags(G, y0, y1, y2) == 0u;
}

__global__ __launch_bounds__(kBlock) void kQuickClaim(Geom G, InterDev in, const Ctl* ctl) {
  const int ns = quickStaged(in, ctl);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < ns; i += gridDim.x * kBlock)
    if (ownSiteNearest(G, in, i))
      in.site_claim[in.stage_slot[i]] = kStagedPri + static_cast<int32_t>(in.stage_key[i]);
}

__global__ __launch_bounds__(kBlock) void kQuickCompete(Geom G, InterDev in, const Ctl* ctl) {
  const int nl = static_cast<int>(ctl->n_inter), ns = quickStaged(in, ctl);
  for (int q = blockIdx.x * kBlock + threadIdx.x; q < nl + ns; q += gridDim.x * kBlock) {
    double y0, y1, y2;
    int32_t pri;
    int64_t own = -1;
    if (q < nl) {
      y0 = in.x[0][q];
      y1 = in.x[1][q];
      y2 = in.x[2][q];
      pri = q;
    } else {
      const int i = q - nl;
      y0 = in.stage_x[0][i];
      y1 = in.stage_x[1][i];
      y2 = in.stage_x[2][i];
      pri = kStagedPri + static_cast<int32_t>(in.stage_key[i]);
      own = in.stage_slot[i];
    }
    const int64_t t = nearSubBoxSlot(G, y0, y1, y2);
    if (t >= 0 && t != own && in.site_claim[t] != INT_MAX) atomicMin(&in.site_claim[t], pri);
  }
}

__global__ __launch_bounds__(kBlock) void kQuickRefill(DecideArgs A) {
  InterDev& in = A.in;
  const int ns = quickStaged(in, A.ctl);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < ns; i += gridDim.x * kBlock) {
    if (!ownSiteNearest(A.G, in, i)) continue;
    const int64_t t = in.stage_slot[i];
    if (in.site_claim[t] == kStagedPri + static_cast<int32_t>(in.stage_key[i])) {
      const int8_t ty = static_cast<int8_t>(species(in.stage_type[i]));
      A.lid[t] = in.stage_id[i];
      A.ltype[t] = ty;
      for (int k = 0; k < 3; k++) {
        A.lx[k][t] = in.stage_x[k][i];
        A.lv[k][t] = in.stage_v[k][i];
      }
      writeLatticeImages(A, t, true, ty);
      // it left the site by more than 0.2a (that is why it was staged)
      const unsigned long long q = atomicAdd(reinterpret_cast<unsigned long long*>(&A.ctl->n_far), 1ull);
      if (q < static_cast<unsigned long long>(in.cap)) in.far_new[q] = t;
      else setFlag(reinterpret_cast<unsigned long long*>(&A.ctl->err_flags), 1ull);
      atomicAdd(reinterpret_cast<unsigned long long*>(&A.ctl->n_quick), 1ull);
      in.stage_key[i] = kQuickDone;
    }
    in.site_claim[t] = INT_MAX;
  }
}

// Decide, step 3: flag the sub-box sites within the full neighbour list of
// every entry (kNearInter: their sums visit the inter atoms), and those
// within the full neighbourhood of every lattice atom farther than 0.2a from
// its site and of its periodic images (kWide: their sums visit the whole full
// list); the far list becomes the one the next decide clears.
__device__ __forceinline__ void flagBody(const Geom& G, const InterDev& in, int8_t* ltype, const Ctl* ctl,
                                         int64_t t0, int64_t stride) {
  const int ne = static_cast<int>(ctl->n_prev);
  const int nfar = static_cast<int>(ctl->n_far_prev);
  const FullPlaneOffsets& full = *in.full;
  const int nfull = 1 + max(full.n[0], full.n[1]);
  for (int64_t q = t0; q < static_cast<int64_t>(ne) * nfull; q += stride) {
    const int e = static_cast<int>(q / nfull), k = static_cast<int>(q % nfull) - 1;
    fullNeighbour(G, full, in.e_slot[e], k, [&](int64_t t) {
      if (!(ltype[t] & kNearInter)) typeBitsOr(ltype, t, kNearInter);
    });
  }
  for (int64_t q = t0; q < static_cast<int64_t>(nfar) * nfull; q += stride) {
    const int e = static_cast<int>(q / nfull), k = static_cast<int>(q % nfull) - 1;
    forSlotAndImages(G, in.far_new[e], [&](int64_t slot) {
      fullNeighbour(G, full, slot, k, [&](int64_t t) {
        const int8_t v = ltype[t];
        if (!isVacant(v) && !(v & kWide)) typeBitsOr(ltype, t, kWide);
      });
    });
  }
  for (int64_t q = t0; q < nfar; q += stride) {
    forSlotAndImages(G, in.far_new[q], [&](int64_t slot) {
      const int8_t v = ltype[slot];
      if (!isVacant(v) && !(v & kFar)) typeBitsOr(ltype, slot, kFar);
    });
    in.far_prev[q] = in.far_new[q];
  }
}
__global__ __launch_bounds__(kBlock) void kFlag(Geom G, InterDev in, int8_t* ltype, const Ctl* ctl) {
  flagBody(G, in, ltype, ctl, static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x,
           static_cast<int64_t>(gridDim.x) * kBlock);
}

__global__ void kRescan(Ctl* ctl) { ctl->rescan = 1; }

// Lattice atoms farther than 0.2a from their site (the test of atom::decide,
// atom.cpp:33-49) -> far list, before an index built without a decide.
__global__ __launch_bounds__(kBlock) void kFarScan(Geom G, CXPtr x, const int8_t* __restrict__ type, InterDev in,
                                                   Ctl* ctl) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int8_t t = type[s];
  if (isVacant(t) || !leftSite(G, p, cx, cy, cz, x.c[0][s], x.c[1][s], x.c[2][s])) return;
  const unsigned long long q = atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->n_far), 1ull);
  if (static_cast<int64_t>(q) < in.cap) in.far_new[q] = s;
  else atomicOr(reinterpret_cast<unsigned long long*>(&ctl->err_flags), 1ull);
}

// ---------------------------------------------------------------------------
// per inter atom: rho/F', force + half kick, first half step
// ---------------------------------------------------------------------------

// Per inter atom one wavefront: lane l takes the sites of full-list items
// l - 1, l + 63, l + 127, ... (item -1 is the nearest site itself) and the
// inter atoms in their buckets; the lanes' partial sums then meet in a fixed
// shuffle tree, so the result does not depend on the launch.
constexpr int kWaveBlocks = 256;  // blocks of 4 wavefronts, grid-stride over the inter atoms

__device__ __forceinline__ double waveSum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// atom::interRho (atom.cpp:194-284): lattice (nearest site, full list), same
// bucket without itself, then the buckets of the full list; F' at once.
template <typename T>
__device__ __forceinline__ void interRhoBody(const Geom& G, const Tables<T>& tb, const InterView& iv, CXPtr lx,
                                             const int8_t* __restrict__ ltype, const InterDev& in, int w0,
                                             int wstride) {
  const int nl = static_cast<int>(iv.ctl->n_inter);
  const int ne = nl + static_cast<int>(iv.ctl->n_ghost);
  const FullPlaneOffsets& full = *iv.full;
  const int lane = threadIdx.x & 63;
  for (int i = w0; i < nl; i += wstride) {
    const int64_t n = in.e_slot[i];
    if (!isSubBoxSlot(G, n)) continue;
    const double xi0 = in.x[0][i], xi1 = in.x[1][i], xi2 = in.x[2][i];
    const uint64_t myid = in.id[i];
    const int par = n >= G.plane ? 1 : 0;
    double rho = 0.0;
    for (int k = lane - 1; k < full.n[par]; k += 64) {
      const int64_t j = k < 0 ? n : n + full.off[par][k];
      const int8_t tj = ltype[j];
      if (!isVacant(tj)) {
        const double d0 = xi0 - lx.c[0][j], d1 = xi1 - lx.c[1][j], d2 = xi2 - lx.c[2][j];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (r2 < G.rc2) rho += rhoPair(tb, species(tj), r2);
      }
      forBucket(iv, j, ne, [&](int e) {
        const int src = iv.e_src[e];
        if (k < 0 && e < nl && iv.id[src] == myid) return;  // itself (ghost images carry no id)
        const double d0 = xi0 - iv.e_x[0][e], d1 = xi1 - iv.e_x[1][e], d2 = xi2 - iv.e_x[2][e];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (r2 < G.rc2) rho += rhoPair(tb, species(iv.type[src]), r2);
      });
    }
    rho = waveSum(rho);
    if (lane == 0) {
      in.rho[i] = rho;
      checkRho(iv.ctl, rho, static_cast<double>(tb.rhomax));
      const T dfi = dEmbed(tb, species(in.type[i]), static_cast<T>(rho));
      if constexpr (sizeof(T) == 8)
        in.df64[i] = dfi;
      else
        in.df32[i] = dfi;
    }
  }
}
template <typename T>
__global__ __launch_bounds__(kIT) void kInterRho(Geom G, Tables<T> tb, InterView iv, CXPtr lx,
                                                 const int8_t* __restrict__ ltype, InterDev in) {
  interRhoBody<T>(G, tb, iv, lx, ltype, in, blockIdx.x * (kIT / 64) + static_cast<int>(threadIdx.x >> 6),
                  kWaveBlocks * (kIT / 64));
}

// atom::interForce (atom.cpp:360-473) + secondstep for the inter list.
template <typename T, int MODE>
__device__ __forceinline__ void interForceBody(const Geom& G, const Tables<T>& tb, const InterView& iv, CXPtr lx,
                                               const int8_t* __restrict__ ltype, const T* __restrict__ ldf,
                                               int ldfStride, const InterDev& in, const VerletParams& vp, int w0,
                                               int wstride) {
  const int nl = static_cast<int>(iv.ctl->n_inter);
  const int ne = nl + static_cast<int>(iv.ctl->n_ghost);
  const FullPlaneOffsets& full = *iv.full;
  const int lane = threadIdx.x & 63;
  for (int i = w0; i < nl; i += wstride) {
    const int64_t n = in.e_slot[i];
    double f0 = 0, f1 = 0, f2 = 0;
    const int ei = species(in.type[i]);
    if (isSubBoxSlot(G, n)) {
      const double xi0 = in.x[0][i], xi1 = in.x[1][i], xi2 = in.x[2][i];
      const uint64_t myid = in.id[i];
      const T dfi = dfOf<T>(iv, i);
      const int par = n >= G.plane ? 1 : 0;
      for (int k = lane - 1; k < full.n[par]; k += 64) {
        const int64_t j = k < 0 ? n : n + full.off[par][k];
        const int8_t tj = ltype[j];
        {
          const double d0 = xi0 - lx.c[0][j], d1 = xi1 - lx.c[1][j], d2 = xi2 - lx.c[2][j];
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          if (r2 < G.rc2 && !isVacant(tj)) {
            const double fp = static_cast<double>(pairForce(tb, ei, species(tj), r2, dfi, ldf[j * ldfStride]));
            f0 += d0 * fp;
            f1 += d1 * fp;
            f2 += d2 * fp;
          }
        }
        forBucket(iv, j, ne, [&](int e) {
          const int src = iv.e_src[e];
          if (k < 0 && e < nl && iv.id[src] == myid) return;
          const double d0 = xi0 - iv.e_x[0][e], d1 = xi1 - iv.e_x[1][e], d2 = xi2 - iv.e_x[2][e];
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          if (r2 < G.rc2) {
            const double fp =
                static_cast<double>(pairForce(tb, ei, species(iv.type[src]), r2, dfi, dfOf<T>(iv, src)));
            f0 += d0 * fp;
            f1 += d1 * fp;
            f2 += d2 * fp;
          }
        });
      }
      f0 = waveSum(f0);
      f1 = waveSum(f1);
      f2 = waveSum(f2);
    }
    if (lane == 0) {
      checkFinite(iv.ctl, f0, f1, f2);
      in.f[0][i] = f0;
      in.f[1][i] = f1;
      in.f[2][i] = f2;
      if (MODE == 0) {  // secondstep, each product and sum rounded separately
#pragma clang fp contract(off)
        const double c = stepDt(vp, ei).c;
        in.v[0][i] += c * f0;
        in.v[1][i] += c * f1;
        in.v[2][i] += c * f2;
      }
    }
  }
}
template <typename T, int MODE>
__global__ __launch_bounds__(kIT) void kInterForce(Geom G, Tables<T> tb, InterView iv, CXPtr lx,
                                                   const int8_t* __restrict__ ltype, const T* __restrict__ ldf,
                                                   int ldfStride, InterDev in, VerletParams vp) {
  interForceBody<T, MODE>(G, tb, iv, lx, ltype, ldf, ldfStride, in, vp,
                          blockIdx.x * (kIT / 64) + static_cast<int>(threadIdx.x >> 6), kWaveBlocks * (kIT / 64));
}

// NewtonMotion::firststep for the inter list (newton_motion.cpp:46-54).
__device__ __forceinline__ void interFirstBody(const InterDev& in, const Ctl* ctl, const VerletParams& vp, int i0,
                                               int istride) {
#pragma clang fp contract(off)  // the half kick and drift round as newton_motion.cpp does
  const int nl = static_cast<int>(ctl->n_inter);
  for (int i = i0; i < nl; i += istride) {
    const StepDt sd = stepDt(vp, species(in.type[i]));
    const double c = sd.c;
    for (int d = 0; d < 3; d++) {
      const double v = in.v[d][i] + c * in.f[d][i];
      in.v[d][i] = v;
      in.x[d][i] += sd.dt * v;
    }
  }
}
__global__ __launch_bounds__(kIT) void kInterFirst(InterDev in, const Ctl* ctl, VerletParams vp) {
  interFirstBody(in, ctl, vp, blockIdx.x * kIT + threadIdx.x, kIT * kIB);
}

__global__ __launch_bounds__(kIT) void kInterUpload(InterDev in, const AtomElement* __restrict__ a, int64_t n,
                                                    Ctl* ctl) {
  for (int64_t i = blockIdx.x * kIT + threadIdx.x; i < n; i += kIT * kIB) {
    in.id[i] = a[i].id;
    in.type[i] = static_cast<int8_t>(a[i].type);
    for (int k = 0; k < 3; k++) {
      in.x[k][i] = a[i].x[k];
      in.v[k][i] = a[i].v[k];
      in.f[k][i] = a[i].f[k];
    }
    in.rho[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->n_inter = n;
    ctl->n_ghost = 0;
    ctl->n_stage = 0;
  }
}

__global__ __launch_bounds__(kIT) void kInterDownload(InterDev in, AtomElement* __restrict__ a, const Ctl* ctl,
                                                      bool mixed) {
  const int64_t n = ctl->n_inter;
  for (int64_t i = blockIdx.x * kIT + threadIdx.x; i < n; i += kIT * kIB) {
    AtomElement e;
    e.id = in.id[i];
    e.type = in.type[i];
    for (int k = 0; k < 3; k++) {
      e.x[k] = in.x[k][i];
      e.v[k] = in.v[k][i];
      e.f[k] = in.f[k][i];
    }
    e.rho = in.rho[i];
    e.df = mixed ? static_cast<double>(in.df32[i]) : in.df64[i];
    a[i] = e;
  }
}

__global__ __launch_bounds__(kIT) void kInterScaleV(InterDev in, const Ctl* ctl, double factor) {
  const int nl = static_cast<int>(ctl->n_inter);
  for (int i = blockIdx.x * kIT + threadIdx.x; i < nl; i += kIT * kIB)
    for (int d = 0; d < 3; d++) in.v[d][i] *= factor;
}

template <int N>
__device__ __forceinline__ void blockSumI(double (&vals)[N], double* out) {
  __shared__ double red[N][kIT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < N; q++) {
    double v = vals[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[q][w] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < N; q++) {
      double v = 0;
      for (int k = 0; k < kIT / 64; k++) v += red[q][k];
      out[q] = v;
    }
}

__global__ __launch_bounds__(kIT) void kInterThermo(InterDev in, const Ctl* ctl, double* partials) {
  const int nl = static_cast<int>(ctl->n_inter);
  double vals[5] = {0, 0, 0, 0, 0};
  double vmax = 0;
  for (int i = blockIdx.x * kIT + threadIdx.x; i < nl; i += kIT * kIB) {
    const int t = species(in.type[i]);
    const double m = t == 0 ? 55.845 : (t == 1 ? 63.546 : 58.6934);
    const double a = in.v[0][i], b = in.v[1][i], c = in.v[2][i];
    const double vv = a * a + b * b + c * c;
    vals[0] += vv * m;
    vals[1] += in.f[0][i];
    vals[2] += in.f[1][i];
    vals[3] += in.f[2][i];
    vmax = fmax(vmax, vv);
  }
  __shared__ double mx[kIT];
  mx[threadIdx.x] = vmax;
  blockSumI<5>(vals, partials + 5 * blockIdx.x);
  __syncthreads();
  for (int o = kIT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[5 * blockIdx.x + 4] = mx[0];
}

// Energy of the inter atoms: F(rho_I) + phi to lattice neighbours (each such
// pair once, from the inter side) + 1/2 phi to inter neighbours; virial alike.
__global__ __launch_bounds__(kIT) void kInterEnergy(Geom G, Tables<double> tb, InterView iv, CXPtr lx,
                                                    const int8_t* __restrict__ ltype, const double* __restrict__ lrho,
                                                    InterDev in, double* partials) {
  const int nl = static_cast<int>(iv.ctl->n_inter);
  const int ne = nl + static_cast<int>(iv.ctl->n_ghost);
  const FullPlaneOffsets& full = *iv.full;
  double vals[3] = {0, 0, 0};
  for (int i = blockIdx.x * kIT + threadIdx.x; i < nl; i += kIT * kIB) {
    const int64_t n = in.e_slot[i];
    if (!isSubBoxSlot(G, n)) continue;
    const int ei = species(in.type[i]);
    const double ri = in.rho[i];
    const double dfi = dEmbed(tb, ei, ri);
    vals[0] += embed(tb, ei, ri);
    const double xi0 = in.x[0][i], xi1 = in.x[1][i], xi2 = in.x[2][i];
    const uint64_t myid = in.id[i];
    const int par = n >= G.plane ? 1 : 0;
    auto lat = [&](int64_t j) {
      const int8_t tj = ltype[j];
      const double d0 = xi0 - lx.c[0][j], d1 = xi1 - lx.c[1][j], d2 = xi2 - lx.c[2][j];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (r2 < G.rc2 && !isVacant(tj)) {
        const int ej = species(tj);
        const double r = sqrt(r2);
        double q;
        const int m = knotOf(r, tb.rdr, tb.nr, false, q);
        vals[1] += splineVal(tb.z2 + pairIdx(ei, ej) * tb.sR + static_cast<int64_t>(m) * kKnot, q) / r;
        vals[2] += r2 * pairForce(tb, ei, ej, r2, dfi, dEmbed(tb, ej, lrho[j]));
      }
    };
    auto ent = [&](int e, bool skip_self) {
      const int src = iv.e_src[e];
      const uint64_t eid = e < nl ? iv.id[src] : 0;
      if (skip_self && eid == myid) return;
      const double d0 = xi0 - iv.e_x[0][e], d1 = xi1 - iv.e_x[1][e], d2 = xi2 - iv.e_x[2][e];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (r2 < G.rc2) {
        const int ej = species(iv.type[src]);
        const double r = sqrt(r2);
        double q;
        const int m = knotOf(r, tb.rdr, tb.nr, false, q);
        vals[1] += 0.5 * splineVal(tb.z2 + pairIdx(ei, ej) * tb.sR + static_cast<int64_t>(m) * kKnot, q) / r;
        vals[2] += 0.5 * r2 * pairForce(tb, ei, ej, r2, dfi, dEmbed(tb, ej, in.rho[src]));
      }
    };
    lat(n);
    for (int k = 0; k < full.n[par]; k++) lat(n + full.off[par][k]);
    forBucket(iv, n, ne, [&](int e) { ent(e, true); });
    for (int k = 0; k < full.n[par]; k++) forBucket(iv, n + full.off[par][k], ne, [&](int e) { ent(e, false); });
  }
  blockSumI<3>(vals, partials + 3 * blockIdx.x);
}

// ---------------------------------------------------------------------------
// adaptive step length: dt = min(dt_max, max_disp / max|v|)
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void blockMaxToCtl(unsigned long long m, Ctl* ctl) {
  __shared__ unsigned long long red[N / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_down(m, o, 64);
    m = y > m ? y : m;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long b = 0;
    for (int k = 0; k < N / 64; k++) b = red[k] > b ? red[k] : b;
    atomicMax(&ctl->vmax2, b);
  }
}

// Grid-stride over the sub-box sites of both planes with a fixed number of
// blocks, so that only a few hundred atomics meet at the one counter.
constexpr int kVmaxBlocks = 512;
__global__ __launch_bounds__(kBlock) void kVmaxSites(Geom G, CXPtr v, const int8_t* __restrict__ type, Ctl* ctl) {
  unsigned long long m = 0;
  const int64_t n = 2 * G.nsub;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(kVmaxBlocks) * kBlock) {
    const int p = i >= G.nsub ? 1 : 0;
    int64_t r = i - p * G.nsub;
    const int cx = static_cast<int>(r % G.n0);
    r /= G.n0;
    const int cy = static_cast<int>(r % G.n1);
    const int cz = static_cast<int>(r / G.n1);
    const int64_t s = slotOf(G, p, cx, cy, cz);
    if (!isVacant(type[s])) {
      const unsigned long long b = speed2Bits(v.c[0][s], v.c[1][s], v.c[2][s]);
      m = b > m ? b : m;
    }
  }
  blockMaxToCtl<kBlock>(m, ctl);
}

__global__ __launch_bounds__(kIT) void kVmaxInter(InterDev in, Ctl* ctl) {
  const int nl = static_cast<int>(ctl->n_inter);
  unsigned long long m = 0;
  for (int i = blockIdx.x * kIT + threadIdx.x; i < nl; i += kIT * kIB) {
    const unsigned long long b = speed2Bits(in.v[0][i], in.v[1][i], in.v[2][i]);
    m = b > m ? b : m;
  }
  blockMaxToCtl<kIT>(m, ctl);
}

__global__ void kDtUpdate(Ctl* ctl, double dt_max, double max_disp) {
  const double vmax = sqrt(__longlong_as_double(static_cast<long long>(ctl->vmax2)));
  const double dt = vmax > 0.0 ? fmin(dt_max, max_disp / vmax) : dt_max;
  ctl->dt = dt;
  ctl->t_sim += dt;
  ctl->vmax2 = 0ull;
}

// ---------------------------------------------------------------------------
// a batch of boxes: each kernel above once for all boxes (kBatchBlocks
// blocks per box, or one workgroup per box for the list work)
// ---------------------------------------------------------------------------
constexpr int kBatchBlocks = 8;

__global__ __launch_bounds__(kBlock) void kUnflagB(const BatchBox* __restrict__ boxes) {
  const BatchBox& B = boxes[blockIdx.x / kBatchBlocks];
  unflagBody(B.G, B.in, B.type, B.ctl, static_cast<int64_t>(blockIdx.x % kBatchBlocks) * kBlock + threadIdx.x,
             static_cast<int64_t>(kBatchBlocks) * kBlock);
}
__global__ __launch_bounds__(kBlock) void kFlagB(const BatchBox* __restrict__ boxes) {
  const BatchBox& B = boxes[blockIdx.x / kBatchBlocks];
  flagBody(B.G, B.in, B.type, B.ctl, static_cast<int64_t>(blockIdx.x % kBatchBlocks) * kBlock + threadIdx.x,
           static_cast<int64_t>(kBatchBlocks) * kBlock);
}
// args[2 b + flip]: the decide arguments of box b for either buffer
__global__ __launch_bounds__(kDT) void kDecideB(const DecideArgs* __restrict__ args, int flip) {
  decideBody(args[2 * blockIdx.x + flip]);
}
__global__ __launch_bounds__(kIT) void kInterRhoB(const BatchBox* __restrict__ boxes, int flip) {
  const BatchBox& B = boxes[blockIdx.x / kBatchBlocks];
  const int c = B.cur ^ flip;
  interRhoBody<float>(B.G, B.t32, B.iv, cxp(B.x[c]), B.type, B.in,
                      static_cast<int>(blockIdx.x % kBatchBlocks) * (kIT / 64) + static_cast<int>(threadIdx.x >> 6),
                      kBatchBlocks * (kIT / 64));
}
__global__ __launch_bounds__(kIT) void kInterForceB(const BatchBox* __restrict__ boxes, int flip) {
  const BatchBox& B = boxes[blockIdx.x / kBatchBlocks];
  const int c = B.cur ^ flip;
  const float* ldf = reinterpret_cast<const float*>(B.rec[c]) + 3;  // lattice F' is the w of the records
  interForceBody<float, 0>(B.G, B.t32, B.iv, cxp(B.x[c]), B.type, ldf, 4, B.in, B.vp,
                           static_cast<int>(blockIdx.x % kBatchBlocks) * (kIT / 64) +
                               static_cast<int>(threadIdx.x >> 6),
                           kBatchBlocks * (kIT / 64));
}
__global__ __launch_bounds__(kIT) void kInterFirstB(const BatchBox* __restrict__ boxes) {
  const BatchBox& B = boxes[blockIdx.x / kBatchBlocks];
  interFirstBody(B.in, B.ctl, B.vp, static_cast<int>(blockIdx.x % kBatchBlocks) * kIT + threadIdx.x,
                 kBatchBlocks * kIT);
}

}  // namespace

size_t decideArgsBytes() { return sizeof(DecideArgs); }

void fillDecideArgs(const BatchBox& b, int c, void* out) {
  DecideArgs A;
  A.G = b.G;
  A.mode = 0;
  A.cur = c;
  for (int k = 0; k < 3; k++) {
    A.lx[k] = b.x[c][k];
    A.lv[k] = b.v[k];
  }
  A.ltype = b.type;
  A.lid = b.id;
  A.lrec = b.rec[c];
  A.lrec64 = Rec64{nullptr, nullptr};
  A.in = b.in;
  A.ctl = b.ctl;
  A.skin = b.skin;
  std::memcpy(out, &A, sizeof(A));
}

void launchInterBatch(int which, const BatchBox* boxes, int nbox, int flip, const void* decideArgs, hipStream_t st) {
  const unsigned g = static_cast<unsigned>(nbox * kBatchBlocks);
  switch (which) {
    case 0: kInterRhoB<<<g, kIT, 0, st>>>(boxes, flip); break;
    case 1: kInterForceB<<<g, kIT, 0, st>>>(boxes, flip); break;
    case 2: kInterFirstB<<<g, kIT, 0, st>>>(boxes); break;
    default:  // the decide of every box, on buffer cur ^ flip
      kUnflagB<<<g, kBlock, 0, st>>>(boxes);
      kDecideB<<<static_cast<unsigned>(nbox), kDT, 0, st>>>(static_cast<const DecideArgs*>(decideArgs), flip);
      kFlagB<<<g, kBlock, 0, st>>>(boxes);
      break;
  }
}

void launchAdaptiveDt(const DevState& s, double dt_max, double max_disp, bool sites, hipStream_t st) {
  if (sites) kVmaxSites<<<kVmaxBlocks, kBlock, 0, st>>>(s.G, cxp(s.v), s.type, s.ctl);
  kVmaxInter<<<kIB, kIT, 0, st>>>(s.in, s.ctl);
  kDtUpdate<<<1, 1, 0, st>>>(s.ctl, dt_max, max_disp);
}

InterView interView(const DevState& s, bool mixed) {
  InterView v;
  v.head = s.in.head;
  v.keys = s.in.keys;
  v.e_src = s.in.e_src;
  for (int k = 0; k < 3; k++) v.e_x[k] = s.in.e_x[k];
  v.type = s.in.type;
  v.id = s.in.id;
  v.df64 = s.in.df64;
  v.df32 = s.in.df32;
  v.ctl = s.ctl;
  v.full = s.in.full;
  (void)mixed;
  return v;
}

void launchFarScan(const DevState& s, hipStream_t st) {
  kFarScan<<<siteGrid(s.G), kBlock, 0, st>>>(s.G, cxp(s.x[s.cur]), s.type, s.in, s.ctl);
}

void launchDecide(const DevState& s, int mode, bool mixed, hipStream_t st) {
  DecideArgs A;
  A.G = s.G;
  A.mode = mode;
  A.cur = s.cur;
  for (int k = 0; k < 3; k++) {
    A.lx[k] = s.x[s.cur][k];
    A.lv[k] = s.v[k];
  }
  A.ltype = s.type;
  A.lid = s.id;
  A.lrec = mixed ? s.rec[s.cur] : nullptr;
  A.lrec64 = mixed ? Rec64{nullptr, nullptr} : Rec64{s.pxy[s.cur], s.pzw[s.cur]};
  A.in = s.in;
  A.ctl = s.ctl;
  A.skin = s.skin;
  kUnflag<<<kFlagBlocks, kBlock, 0, st>>>(s.G, s.in, s.type, s.ctl);
  if (mode == 0 && 2 * s.G.plane < kStagedPri) {  // site indices below the staged priority offset
    kQuickClaim<<<kQuickBlocks, kBlock, 0, st>>>(s.G, s.in, s.ctl);
    kQuickCompete<<<kQuickBlocks, kBlock, 0, st>>>(s.G, s.in, s.ctl);
    kQuickRefill<<<kQuickBlocks, kBlock, 0, st>>>(A);
  }
  kDecide<<<1, kDT, 0, st>>>(A);
  kFlag<<<kFlagBlocks, kBlock, 0, st>>>(s.G, s.in, s.type, s.ctl);
}

void launchRescan(const DevState& s, hipStream_t st) { kRescan<<<1, 1, 0, st>>>(s.ctl); }

void launchInterRho(const DevState& s, const Tables<double>& t64, const Tables<float>& t32, bool mixed,
                    hipStream_t st) {
  const InterView iv = interView(s, mixed);
  if (mixed)
    kInterRho<float><<<kWaveBlocks, kIT, 0, st>>>(s.G, t32, iv, cxp(s.x[s.cur]), s.type, s.in);
  else
    kInterRho<double><<<kWaveBlocks, kIT, 0, st>>>(s.G, t64, iv, cxp(s.x[s.cur]), s.type, s.in);
}

void launchInterForce(const DevState& s, const Tables<double>& t64, const Tables<float>& t32, bool mixed, int mode,
                      const VerletParams& vp, hipStream_t st) {
  const InterView iv = interView(s, mixed);
  const CXPtr lx = cxp(s.x[s.cur]);
  if (mixed) {  // lattice F' is the w component of the site records
    const float* ldf = reinterpret_cast<const float*>(s.rec[s.cur]) + 3;
    if (mode == 0)
      kInterForce<float, 0><<<kWaveBlocks, kIT, 0, st>>>(s.G, t32, iv, lx, s.type, ldf, 4, s.in, vp);
    else
      kInterForce<float, 2><<<kWaveBlocks, kIT, 0, st>>>(s.G, t32, iv, lx, s.type, ldf, 4, s.in, vp);
  } else {
    if (mode == 0)
      kInterForce<double, 0><<<kWaveBlocks, kIT, 0, st>>>(s.G, t64, iv, lx, s.type, s.df64, 1, s.in, vp);
    else
      kInterForce<double, 2><<<kWaveBlocks, kIT, 0, st>>>(s.G, t64, iv, lx, s.type, s.df64, 1, s.in, vp);
  }
}

void launchInterFirst(const DevState& s, const VerletParams& vp, hipStream_t st) {
  kInterFirst<<<kIB, kIT, 0, st>>>(s.in, s.ctl, vp);
}

void launchInterUpload(const DevState& s, const AtomElement* aos_dev, int64_t n, hipStream_t st) {
  kInterUpload<<<kIB, kIT, 0, st>>>(s.in, aos_dev, n, s.ctl);
}

void launchInterDownload(const DevState& s, AtomElement* aos_dev, bool mixed, hipStream_t st) {
  kInterDownload<<<kIB, kIT, 0, st>>>(s.in, aos_dev, s.ctl, mixed);
}

void launchInterScaleV(const DevState& s, double factor, hipStream_t st) {
  kInterScaleV<<<kIB, kIT, 0, st>>>(s.in, s.ctl, factor);
}

int64_t launchInterThermo(const DevState& s, hipStream_t st) {
  kInterThermo<<<kIB, kIT, 0, st>>>(s.in, s.ctl, s.partials);
  return kIB;
}

int64_t launchInterEnergy(const DevState& s, const Tables<double>& t64, hipStream_t st) {
  kInterEnergy<<<kIB, kIT, 0, st>>>(s.G, t64, interView(s, false), cxp(s.x[s.cur]), s.type, s.rho, s.in,
                                    s.partials);
  return kIB;
}

}  // namespace misa
