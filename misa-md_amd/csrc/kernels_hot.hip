// The two hot passes of the resident EAM step on MI355X (gfx950).
//
//   kRhoDf : rho_i = sum_j rho_{t_j}(r_ij) over the pruned full neighbour list,
//            F'(rho_i) -> df (+ ghost images)   (atom::latRho + latDf + the F'
//            forward halo, atom.cpp:102-149)
//   kForce : f_i = sum_j (x_i - x_j) fpair_ij, then the second half kick and,
//            in the fused mode, the next step's half kick, drift, the lattice
//            part of atom::decide and the position halo (atom::latForce,
//            NewtonMotion, atom::decide; simulation.cpp:164-194)
//
// One thread per site.  A full list (each lane owns its site's sums) replaces
// the reference's half list with ghost partials and reverse halos: no atomics,
// deterministic sums.  The list keeps the reference offsets whose ideal
// separation is below rc + 0.4a (+0.01a): a lattice atom normally stays within
// 0.2a of its site (decide() turns it into an inter atom otherwise), so every
// dropped offset is beyond the cutoff for such pairs.  The exception is an
// inter atom that has just refilled a vacancy anywhere in its Wigner-Seitz
// cell: the sites within its full neighbourhood carry kWide and also visit
// the dropped offsets, so the visited pair set is always the reference's.
//
// Memory access: every lane of a wavefront reads neighbour k at the same slot
// delta, so x, type and F' are fetched with buffer loads whose per-lane
// offset (site * size) is fixed for the whole loop and whose scalar offset is
// the (shifted, non-negative) delta: no per-neighbour address arithmetic.
// Neighbours are taken in groups of U: the group's position and type loads
// are all issued before the first distance test, and the spline records of
// the in-cutoff pairs are all issued before the first pair term, so a wave
// keeps U gathers in flight instead of waiting on each neighbour in turn.
// Spline records are packed per knot: the density value cubic (4 values) per
// species for the rho pass, and per ordered species pair the value cubics of
// r*phi and of the neighbour's density (8 values) for the force pass; the
// derivatives are taken from the value cubics, and the centre's density
// cubic is fetched separately only for unlike pairs.
#include <hip/hip_runtime.h>

#include <cmath>

#include "device_common.hpp"
#include "inter_common.hpp"
#include "resident.hpp"
#include "site_common.hpp"
#include "buffer_load.hpp"
#include "mixed_site.hpp"

namespace misa {

namespace {

using namespace site;

__device__ __forceinline__ double rsqT(double x) { return rsqrtRefined(x); }

// |x_i - x_j|^2 in FP64 with one fixed rounding, so that the density and
// force passes agree on which pairs are within the cutoff.
__device__ __forceinline__ double dist2D(double d0, double d1, double d2) {
  return __fma_rn(d2, d2, __fma_rn(d1, d1, __dmul_rn(d0, d0)));
}

// Index of the wavefront in the grid, in a scalar register (threadIdx.x >> 6
// is the same in every lane of a wavefront).
__device__ __forceinline__ uint32_t waveId() {
  return blockIdx.x * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}
// A wavefront-uniform 64-bit value moved to scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Calls visit(ks, nv) for the set bits of (b0, b1) in increasing order, U at
// a time, then one at a time (nv = U or 1; ks holds the offsets).
template <int U, typename Visit>
__device__ __forceinline__ void forMaskGroups(uint64_t b0, uint64_t b1, Visit visit) {
  int left = __builtin_popcountll(b0) + __builtin_popcountll(b1);
  for (; left > 0;) {
    const int V = left >= U ? U : 1;
    int ks[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (u < V) {
        if (b0) {
          ks[u] = __builtin_ctzll(b0);
          b0 &= b0 - 1;
        } else {
          ks[u] = 64 + __builtin_ctzll(b1);
          b1 &= b1 - 1;
        }
      } else {
        ks[u] = 0;
      }
    }
    visit(ks, V);
    left -= V;
  }
}

// Calls visit(ks, act) for the set bits of (b0, b1) in increasing order,
// always U at a time: in the last group the entries past the last bit
// (act[u] false) repeat the last offset, and the caller makes them add
// nothing.  One shape of group keeps the gathers of a group in flight
// together.
template <int U, typename Visit>
__device__ __forceinline__ void forMaskGroupsFull(uint64_t b0, uint64_t b1, Visit visit) {
  int k = 0;
  while (b0 | b1) {
    int ks[U];
    bool act[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      act[u] = (b0 | b1) != 0;
      if (b0) {
        k = __builtin_ctzll(b0);
        b0 &= b0 - 1;
      } else if (b1) {
        k = 64 + __builtin_ctzll(b1);
        b1 &= b1 - 1;
      }
      ks[u] = k;
    }
    visit(ks, act);
  }
}

// Set bit k of the 128-bit offset mask (m0, m1).
__device__ __forceinline__ void setBit(uint64_t& m0, uint64_t& m1, int k) {
  if (k < 64) m0 |= 1ull << k;
  else m1 |= 1ull << (k - 64);
}

// The leader (first gathering lane) writes the wavefront's masks.
__device__ __forceinline__ void writeMasks(bool gathers, uint32_t wv, uint64_t w0, uint64_t w1, uint64_t* wmask,
                                           bool full, uint64_t s0, uint64_t s1, uint64_t* smask) {
  const uint64_t gathered = __ballot(gathers);
  if (!gathered) return;
  const int leader = __builtin_ctzll(gathered);
  const uint64_t a0 = __shfl(w0, leader, 64), a1 = __shfl(w1, leader, 64);
  const uint64_t c0 = __shfl(s0, leader, 64), c1 = __shfl(s1, leader, 64);
  if (static_cast<int>(threadIdx.x & 63) == leader) {
    wmask[2 * wv] = a0;
    wmask[2 * wv + 1] = a1;
    if (full) {
      smask[2 * wv] = c0;
      smask[2 * wv + 1] = c1;
    }
  }
}

// ---------------------------------------------------------------------------
// pass 1: rho and F'
// FP64 and mixed density passes record, per wavefront, the pruned offsets
// within the cutoff for some lane (wmask) for the force pass, which visits
// only those.  A full scan also records the offsets within cutoff + skin
// (smask); while decide finds that no atom can have crossed the skin
// (Ctl::skin_mode) the density pass visits only those: the same pairs in the
// same order, so the sums do not change.
// ---------------------------------------------------------------------------
template <typename T, int ACC, int U>
__global__ __launch_bounds__(kBlock) void kRhoDf(Geom G, LatOffsets O, Tables<T> tb, CXPtr x,
                                                 const int8_t* __restrict__ type, T* __restrict__ df,
                                                 double* __restrict__ rho_out, uint64_t* __restrict__ wmask,
                                                 uint64_t* __restrict__ smask, double rcs2, InterView iv,
                                                 Rec64 r64) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  // a wavefront next to a far lattice atom (kWide) scans in full and records its masks afresh
  const bool skin =
      __builtin_amdgcn_readfirstlane(iv.ctl->skin_mode) != 0 && !__ballot(!isVacant(ti) && (ti & kWide));
  const uint32_t wv = waveId();
  T dfi = T(0);
  double rho = 0.0;
  uint64_t w0 = 0, w1 = 0, s0 = 0, s1 = 0;
  if (!isVacant(ti)) {
    const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
    // neighbours from the FP64 site records (a vacancy sits far out, so every
    // distance test against it fails as the type test did)
    const Rsrc rxy = rsrcBelow(r64.xy, O.lo), rzw = rsrcBelow(r64.zw, O.lo);
    const Rsrc rtab = rsrcAt(reinterpret_cast<uintptr_t>(tb.rhoV));
    const uint32_t o16 = static_cast<uint32_t>(s) * 16u;
    const uint32_t* sh = O.sh[p];
    const int nk = O.n[p];
    Sum1<ACC> acc;
    auto group = [&](auto UC, const int* ks) {
      constexpr int V = decltype(UC)::value;
      double2 xy[V], zw[V];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const uint32_t so = sh[ks[u]] * 16u;
        xy[u] = ldRec64(rxy, o16, so);
        zw[u] = ldRec64(rzw, o16, so);
      }
      bool in[V];
      T pp[V];
      T rec[V][kRhoRec];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        const double r2 = dist2D(xi0 - xy[u].x, xi1 - xy[u].y, xi2 - zw[u].x);
        in[u] = r2 < G.rc2;
        if (!skin && __ballot(r2 < rcs2)) setBit(s0, s1, ks[u]);
        if (__ballot(in[u])) setBit(w0, w1, ks[u]);
        if (in[u]) {
          const T q2 = static_cast<T>(r2);
          const int m = knotOf(q2 * rsqT(q2), tb.rdr, tb.nr, false, pp[u]);
          const uint32_t off = (static_cast<uint32_t>(recSpecies64(zw[u].y) * tb.nr + m)) *
                               static_cast<uint32_t>(kRhoRec * sizeof(T));
          ldRec<T, kRhoRec>(rtab, off, rec[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < V; ++u)
        if (in[u]) acc.add(((rec[u][0] * pp[u] + rec[u][1]) * pp[u] + rec[u][2]) * pp[u] + rec[u][3]);
    };
    if (!skin) {
      int ks[U];
      int k = 0;
      for (; k + U <= nk; k += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) ks[u] = k + u;
        group(std::integral_constant<int, U>{}, ks);
      }
      for (; k < nk; ++k) {
        ks[0] = k;
        group(std::integral_constant<int, 1>{}, ks);
      }
    } else {
      forMaskGroups<U>(uniform64(smask[2 * wv]), uniform64(smask[2 * wv + 1]), [&](const int* ks, int nv) {
        if (nv == U) group(std::integral_constant<int, U>{}, ks);
        else group(std::integral_constant<int, 1>{}, ks);
      });
    }
    if (ti & kNearInter) {  // inter atoms whose nearest site is within the full list (atom.cpp:224-240)
      const int ne = static_cast<int>(iv.ctl->n_inter + iv.ctl->n_ghost);
      const FullPlaneOffsets& full = *iv.full;
      forNearBuckets(iv, full, p, s, ne, [&](int e) {
        const double d0 = xi0 - iv.e_x[0][e], d1 = xi1 - iv.e_x[1][e], d2 = xi2 - iv.e_x[2][e];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (r2 < G.rc2) acc.add(rhoPair(tb, species(iv.type[iv.e_src[e]]), r2));
      });
    }
    if (ti & kWide) {  // full-list offsets beyond the pruned list, near a far lattice atom
      forWideTargets(*iv.full, type, p, s, ti, [&](int64_t j, int8_t tj) {
        const double d0 = xi0 - x.c[0][j], d1 = xi1 - x.c[1][j], d2 = xi2 - x.c[2][j];
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (r2 < G.rc2) acc.add(rhoPair(tb, species(tj), r2));
      });
    }
    rho = acc.value();
    checkRho(iv.ctl, rho, static_cast<double>(tb.rhomax));
    dfi = dEmbed(tb, species(ti), static_cast<T>(rho));
  }
  writeMasks(!isVacant(ti), wv, w0, w1, wmask, !skin, s0, s1, smask);
  df[s] = dfi;
  // F' into the w of the site's record (and of its images' records)
  double* w64 = reinterpret_cast<double*>(r64.zw) + 1;
  const double wi = packW64(static_cast<double>(dfi), species(ti));
  if (!isVacant(ti)) w64[2 * s] = wi;
  if (rho_out) {
    rho_out[s] = rho;
    forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) {
      df[q] = dfi;
      rho_out[q] = rho;
      if (!isVacant(ti)) w64[2 * q] = wi;
    });
  } else {
    forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) {
      df[q] = dfi;
      if (!isVacant(ti)) w64[2 * q] = wi;
    });
  }
}

// secondstep (and, fused, the next firststep and decide) of a lattice atom,
// newton_motion.cpp:46-77, with every product and sum rounded separately.
template <int MODE>
__device__ __forceinline__ void verletTail(const Geom& G, int64_t s, int p, int cx, int cy, int cz, int8_t ti,
                                           int ei, double xi0, double xi1, double xi2, double f0, double f1,
                                           double f2, XPtr xn, XPtr v, XPtr f, const uint64_t* __restrict__ id,
                                           const VerletParams& vp, const StageArgs& st, Ctl* ctl,
                                           const Rec64& rn64, unsigned long long& vb, unsigned long long& db) {
#pragma clang fp contract(off)
  const StepDt sd = stepDt(vp, ei);
  const double c = sd.c;
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
    vb = speed2Bits(v0, v1, v2);
    return;
  }
  v0 = v0 + c * f0;
  v1 = v1 + c * f1;
  v2 = v2 + c * f2;
  const double y0 = xi0 + sd.dt * v0;
  const double y1 = xi1 + sd.dt * v1;
  const double y2 = xi2 + sd.dt * v2;
  db = speed2Bits(y0 - xi0, y1 - xi1, y2 - xi2);
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
  putRec64(rn64, s, ti, y0, y1, y2);
  forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
    const double q0 = y0 + sx * G.L0, q1 = y1 + sy * G.L1, q2 = y2 + sz * G.L2;
    xn.c[0][q] = q0;
    xn.c[1][q] = q1;
    xn.c[2][q] = q2;
    putRec64(rn64, q, ti, q0, q1, q2);
  });
}

// ---------------------------------------------------------------------------
// pass 2: force and velocity Verlet
//   MODE 0: f stored, v += c f                        (secondstep)
//   MODE 1: v += c f; v += c f; x' = x + dt v; decide  (secondstep + next firststep)
//   MODE 2: f stored only                             (first force evaluation)
// ---------------------------------------------------------------------------
// The FP64 segments of a pair at knot m: rec = the r*phi value cubic c3..c6
// and the neighbour's density cubic c3..c6, reci = the centre's density
// cubic (rec's for a like pair).  With LDS a pair of Fe atoms at a knot from
// mlo on reads them from knD (7 doubles a knot, the density's c6 left out:
// no force or energy term uses it), every other pair from the tables in
// global memory; the values are the same.
template <bool LDS>
__device__ __forceinline__ void pairSegD(Rsrc rtab, Rsrc rtabR, const double* knD, int mlo, uint32_t pairBase,
                                         int nr, int ei, int ej, int m, double (&rec)[kPairRec],
                                         double (&reci)[kRhoRec]) {
  if (LDS && (ei | ej) == 0 && m >= mlo) {
    const double* k = knD + 7 * (m - mlo);
#pragma unroll
    for (int q = 0; q < 7; q++) rec[q] = k[q];
    rec[7] = 0.0;
#pragma unroll
    for (int q = 0; q < kRhoRec; q++) reci[q] = rec[4 + q];
    return;
  }
  ldRec<double, kPairRec>(rtab, (pairBase + static_cast<uint32_t>(ej * nr + m)) * (kPairRec * 8u), rec);
  if (ej != ei) {  // rho_ei differs from rho_ej only for unlike pairs
    ldRec<double, kRhoRec>(rtabR, static_cast<uint32_t>(ei * nr + m) * (kRhoRec * 8u), reci);
  } else {
#pragma unroll
    for (int q = 0; q < kRhoRec; q++) reci[q] = rec[4 + q];
  }
}

// One site of the FP64 force pass: thread tl of block b (also a virtual
// block of the persistent launch).  With LDS, knD holds the segments of the
// Fe-Fe pairs from knot mlo on, 7 doubles a knot: the r*phi value cubic
// c3..c6 and the Fe density cubic c3..c5 (its c6 enters no force); a pair
// of Fe atoms at such a knot reads them there, every other pair from the
// tables in global memory.  Both give the same values, so the same bits.
template <typename T, int ACC, int MODE, int U, bool LDS>
__device__ __forceinline__ void forceSiteD(const Geom& G, const LatOffsets& O, const Tables<T>& tb, CXPtr x, XPtr xn,
                                           const int8_t* __restrict__ type, const T* __restrict__ df,
                                           const uint64_t* __restrict__ wmask, XPtr v, XPtr f,
                                           const uint64_t* __restrict__ id, const VerletParams& vp,
                                           const StageArgs& st, Ctl* ctl, const InterView& iv, const Rec64& r64,
                                           const Rec64& rn64, const T* knD, int mlo, uint32_t b, uint32_t tl,
                                           unsigned long long& vb, unsigned long long& db) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfBlock(G, b, tl, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
  if (isVacant(ti)) {  // vacancy: keeps its stale position (atom.cpp:43-47)
    if (MODE != 1) {   // clearForce left 0 there and latForce skips it (atom.cpp:86-100, 311-358)
      f.c[0][s] = 0.0;
      f.c[1][s] = 0.0;
      f.c[2][s] = 0.0;
    } else {
      xn.c[0][s] = xi0;
      xn.c[1][s] = xi1;
      xn.c[2][s] = xi2;
      putRec64(rn64, s, ti, xi0, xi1, xi2);
      forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
        xn.c[0][q] = xi0 + sx * G.L0;
        xn.c[1][q] = xi1 + sy * G.L1;
        xn.c[2][q] = xi2 + sz * G.L2;
        putRec64(rn64, q, ti, xn.c[0][q], xn.c[1][q], xn.c[2][q]);
      });
    }
    return;
  }
  const int ei = species(ti);
  const T dfi = df[s];
  // neighbours from the FP64 site records: position and, in w, F' with the
  // species (a vacancy sits far out, so the distance test drops it)
  const Rsrc rxy = rsrcBelow(r64.xy, O.lo), rzw = rsrcBelow(r64.zw, O.lo);
  const Rsrc rtab = rsrcAt(reinterpret_cast<uintptr_t>(tb.pairF));
  const Rsrc rtabR = rsrcAt(reinterpret_cast<uintptr_t>(tb.rhoV));
  const uint32_t o16 = static_cast<uint32_t>(s) * 16u;
  const uint32_t pairBase = static_cast<uint32_t>(ei * 3) * static_cast<uint32_t>(tb.nr);
  const uint32_t* sh = O.sh[p];
  Sum3<ACC> acc;
  auto group = [&](auto UC, const int* ks) {
    constexpr int V = decltype(UC)::value;
    double d0[V], d1[V], d2[V], wj[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const uint32_t so = sh[ks[u]] * 16u;
      const double2 xy = ldRec64(rxy, o16, so), zw = ldRec64(rzw, o16, so);
      d0[u] = xy.x;
      d1[u] = xy.y;
      d2[u] = zw.x;
      wj[u] = zw.y;
    }
    bool in[V];
    T pp[V], ri[V], dfj[V];
    T rec[V][kPairRec], reci[V][kRhoRec];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      d0[u] = xi0 - d0[u];
      d1[u] = xi1 - d1[u];
      d2[u] = xi2 - d2[u];
      const double r2 = dist2D(d0[u], d1[u], d2[u]);
      in[u] = r2 < G.rc2;
      if (in[u]) {
        const T q2 = static_cast<T>(r2);
        ri[u] = rsqT(q2);
        const int m = knotOf(q2 * ri[u], tb.rdr, tb.nr, false, pp[u]);
        pairSegD<LDS>(rtab, rtabR, knD, mlo, pairBase, tb.nr, ei, recSpecies64(wj[u]), m, rec[u], reci[u]);
        dfj[u] = static_cast<T>(wj[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < V; ++u) {
      if (in[u]) {
        // value cubics c3..c6 in the knot-local coordinate q; derivatives
        // (3 c3 q + 2 c4) q + c5 per knot spacing, times 1/dr
        const T q = pp[u], rinv = ri[u];
        const T* z = rec[u];
        const T* rj = rec[u] + 4;
        const T* rI = reci[u];
        const T z2 = ((z[0] * q + z[1]) * q + z[2]) * q + z[3];
        const T z2p = ((T(3) * z[0] * q + T(2) * z[1]) * q + z[2]) * tb.rdr;
        const T rhojp = ((T(3) * rj[0] * q + T(2) * rj[1]) * q + rj[2]) * tb.rdr;
        const T rhoip = ((T(3) * rI[0] * q + T(2) * rI[1]) * q + rI[2]) * tb.rdr;
        const T phi = z2 * rinv;
        const T phip = z2p * rinv - phi * rinv;
        const T psip = dfi * rhojp + dfj[u] * rhoip + phip;
        acc.add(d0[u], d1[u], d2[u], -psip * rinv);
      }
    }
  };
  // the offsets the density pass found in range for some lane of this wavefront
  const uint32_t wv = b * (kBlock / 64) + __builtin_amdgcn_readfirstlane(tl >> 6);
  forMaskGroups<U>(uniform64(wmask[2 * wv]), uniform64(wmask[2 * wv + 1]), [&](const int* ks, int nv) {
    if (nv == U) group(std::integral_constant<int, U>{}, ks);
    else group(std::integral_constant<int, 1>{}, ks);
  });
  if (ti & kNearInter) {  // forces from inter atoms on lattice neighbours (atom.cpp:397-422)
    const int ne = static_cast<int>(iv.ctl->n_inter + iv.ctl->n_ghost);
    const FullPlaneOffsets& full = *iv.full;
    forNearBuckets(iv, full, p, s, ne, [&](int e) {
      const int src = iv.e_src[e];
      const double d0 = xi0 - iv.e_x[0][e], d1 = xi1 - iv.e_x[1][e], d2 = xi2 - iv.e_x[2][e];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (r2 < G.rc2) acc.add(d0, d1, d2, pairForce(tb, ei, species(iv.type[src]), r2, dfi, dfOf<T>(iv, src)));
    });
  }
  if (ti & kWide) {  // full-list offsets beyond the pruned list, near a far lattice atom
    forWideTargets(*iv.full, type, p, s, ti, [&](int64_t j, int8_t tj) {
      const double e0 = xi0 - x.c[0][j], e1 = xi1 - x.c[1][j], e2 = xi2 - x.c[2][j];
      const double r2 = e0 * e0 + e1 * e1 + e2 * e2;
      if (r2 < G.rc2) acc.add(e0, e1, e2, pairForce(tb, ei, species(tj), r2, dfi, df[j]));
    });
  }
  double f0, f1, f2;
  sum3Value<ACC>(acc, f0, f1, f2);
  checkFinite(ctl, f0, f1, f2);
  verletTail<MODE>(G, s, p, cx, cy, cz, ti, ei, xi0, xi1, xi2, f0, f1, f2, xn, v, f, id, vp, st, ctl, rn64, vb, db);
}

template <typename T, int ACC, int MODE, int U>
__global__ __launch_bounds__(kBlock) void kForce(Geom G, LatOffsets O, Tables<T> tb, CXPtr x, XPtr xn,
                                                 const int8_t* __restrict__ type, const T* __restrict__ df,
                                                 const uint64_t* __restrict__ wmask, XPtr v, XPtr f,
                                                 const uint64_t* __restrict__ id, VerletParams vp, StageArgs st,
                                                 Ctl* ctl, InterView iv, Rec64 r64, Rec64 rn64) {
  unsigned long long vb = 0;  // |v|^2 bits after the secondstep (MODE 0)
  unsigned long long db = 0;  // |dx|^2 bits of the drift (MODE 1)
  forceSiteD<T, ACC, MODE, U, false>(G, O, tb, x, xn, type, df, wmask, v, f, id, vp, st, ctl, iv, r64, rn64, nullptr,
                                     0, blockIdx.x, threadIdx.x, vb, db);
  if (MODE == 0 && vp.vmax) waveMaxTo(vb, vp.vmax);
  if (MODE == 1) waveMaxTo(db, &ctl->dispmax2);
}

// Potential energy and virial of site s (FP64 tables, FP64 site records):
// F(rho_i) into vals[0], 1/2 sum_j phi(r_ij) into vals[1] and
// 1/2 sum_j r_ij^2 fpair_ij into vals[2], fpair as the force pass computes it
// and F'(rho_j) from the records' w.  The offsets are those of the
// wavefront's skin mask (within the cutoff plus the skin of every pair at the
// last full density scan, so a superset of the pairs within the cutoff), or
// all pruned offsets when all is set (no skin).  With LDS the Fe-Fe segments
// come from knD as in forceSiteD.  Near a far lattice atom the rest of the
// full list follows from x and df.
template <bool LDS>
__device__ __forceinline__ void energySiteD(const Geom& G, const LatOffsets& O, const FullPlaneOffsets* full,
                                           const Tables<double>& tb, CXPtr x, const int8_t* __restrict__ type,
                                           const double* __restrict__ rho, const double* __restrict__ df,
                                           const uint64_t* __restrict__ smask, bool all, const Rec64& r64,
                                           const double* knD, int mlo, uint32_t b, uint32_t tl,
                                           double (&vals)[3]) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfBlock(G, b, tl, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  if (isVacant(ti)) return;
  const int ei = species(ti);
  const double ri = rho[s];
  const double dfi = dEmbed(tb, ei, ri);
  vals[0] += embed(tb, ei, ri);
  const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
  const Rsrc rxy = rsrcBelow(r64.xy, O.lo), rzw = rsrcBelow(r64.zw, O.lo);
  const Rsrc rtab = rsrcAt(reinterpret_cast<uintptr_t>(tb.pairF));
  const Rsrc rtabR = rsrcAt(reinterpret_cast<uintptr_t>(tb.rhoV));
  const uint32_t o16 = static_cast<uint32_t>(s) * 16u;
  const uint32_t pairBase = static_cast<uint32_t>(ei * 3) * static_cast<uint32_t>(tb.nr);
  const uint32_t* sh = O.sh[p];
  auto group = [&](auto UC, const int* ks) {
    constexpr int V = decltype(UC)::value;
    double d0[V], d1[V], d2[V], wj[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const uint32_t so = sh[ks[u]] * 16u;
      const double2 xy = ldRec64(rxy, o16, so), zw = ldRec64(rzw, o16, so);
      d0[u] = xi0 - xy.x;
      d1[u] = xi1 - xy.y;
      d2[u] = xi2 - zw.x;
      wj[u] = zw.y;
    }
    bool in[V];
    double r2[V], pp[V], rinv[V], rec[V][kPairRec], reci[V][kRhoRec];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      r2[u] = dist2D(d0[u], d1[u], d2[u]);
      in[u] = r2[u] < G.rc2;
      if (in[u]) {
        rinv[u] = rsqT(r2[u]);
        const int m = knotOf(r2[u] * rinv[u], tb.rdr, tb.nr, false, pp[u]);
        pairSegD<LDS>(rtab, rtabR, knD, mlo, pairBase, tb.nr, ei, recSpecies64(wj[u]), m, rec[u], reci[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < V; ++u) {
      if (!in[u]) continue;
      const double q = pp[u];
      const double* z = rec[u];
      const double* rj = rec[u] + 4;
      const double* rI = reci[u];
      const double z2 = ((z[0] * q + z[1]) * q + z[2]) * q + z[3];
      const double z2p = ((3.0 * z[0] * q + 2.0 * z[1]) * q + z[2]) * tb.rdr;
      const double rhojp = ((3.0 * rj[0] * q + 2.0 * rj[1]) * q + rj[2]) * tb.rdr;
      const double rhoip = ((3.0 * rI[0] * q + 2.0 * rI[1]) * q + rI[2]) * tb.rdr;
      const double phi = z2 * rinv[u];
      const double phip = z2p * rinv[u] - phi * rinv[u];
      const double psip = dfi * rhojp + wj[u] * rhoip + phip;
      vals[1] += 0.5 * phi;
      vals[2] += 0.5 * r2[u] * (-psip * rinv[u]);
    }
  };
  if (all) {
    const int nk = O.n[p];
    int ks[2];
    int k = 0;
    for (; k + 2 <= nk; k += 2) {
      ks[0] = k;
      ks[1] = k + 1;
      group(std::integral_constant<int, 2>{}, ks);
    }
    if (k < nk) {
      ks[0] = k;
      group(std::integral_constant<int, 1>{}, ks);
    }
  } else {
    const uint32_t wv = b * (kBlock / 64) + __builtin_amdgcn_readfirstlane(tl >> 6);
    forMaskGroups<2>(uniform64(smask[2 * wv]), uniform64(smask[2 * wv + 1]), [&](const int* ks, int nv) {
      if (nv == 2) group(std::integral_constant<int, 2>{}, ks);
      else group(std::integral_constant<int, 1>{}, ks);
    });
  }
  if (ti & kWide) {  // full-list offsets beyond the pruned list, near a far lattice atom
    forWideTargets(*full, type, p, s, ti, [&](int64_t j, int8_t tj) {
      const double e0 = xi0 - x.c[0][j], e1 = xi1 - x.c[1][j], e2 = xi2 - x.c[2][j];
      const double q2 = e0 * e0 + e1 * e1 + e2 * e2;
      if (q2 < G.rc2) {
        const int ej = species(tj);
        const double r = sqrt(q2);
        double q;
        const int m = knotOf(r, tb.rdr, tb.nr, false, q);
        vals[1] += 0.5 * splineVal(tb.z2 + pairIdx(ei, ej) * tb.sR + static_cast<int64_t>(m) * kKnot, q) / r;
        vals[2] += 0.5 * q2 * pairForce(tb, ei, ej, q2, dfi, df[j]);
      }
    });
  }
}

// Sums of a block's threads (fixed order), one value each into out.
template <int N, int THREADS>
__device__ __forceinline__ void blockSumD(double (&vals)[N], double* out) {
  __shared__ double red[N][THREADS / 64];
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
      for (int k = 0; k < THREADS / 64; k++) v += red[q][k];
      out[q] = v;
    }
  }
}

// Potential energy and virial, one block's partial sums each: energySiteD
// over the sites of a siteGrid launch, or in persistent blocks with the
// Fe-Fe segments in LDS.
__global__ __launch_bounds__(kBlock) void kEnergyD(Geom G, LatOffsets O, const FullPlaneOffsets* __restrict__ full,
                                                   Tables<double> tb, CXPtr x, const int8_t* __restrict__ type,
                                                   const double* __restrict__ rho, const double* __restrict__ df,
                                                   const uint64_t* __restrict__ smask, int all, Rec64 r64,
                                                   double* partials) {
  double vals[3] = {0, 0, 0};
  energySiteD<false>(G, O, full, tb, x, type, rho, df, smask, all != 0, r64, nullptr, 0, blockIdx.x, threadIdx.x,
                     vals);
  blockSumD<3, kBlock>(vals, partials + 3 * blockIdx.x);
}
__global__ __launch_bounds__(kBlockP) void kEnergyDL(Geom G, LatOffsets O, const FullPlaneOffsets* __restrict__ full,
                                                     Tables<double> tb, CXPtr x, const int8_t* __restrict__ type,
                                                     const double* __restrict__ rho, const double* __restrict__ df,
                                                     const uint64_t* __restrict__ smask, int all, Rec64 r64,
                                                     double* partials, int mlo, int nkn) {
  extern __shared__ double knD[];
  const double* z = tb.pairF + static_cast<size_t>(mlo) * kPairRec;  // row (Fe, Fe) is the first
  for (int i = threadIdx.x; i < 7 * nkn; i += kBlockP) {
    const int k = i / 7, q = i - 7 * k;
    knD[i] = z[static_cast<size_t>(k) * kPairRec + q];
  }
  __syncthreads();
  double vals[3] = {0, 0, 0};
  forVirtualBlocks(G, [&](uint32_t b, uint32_t tl) {
    energySiteD<true>(G, O, full, tb, x, type, rho, df, smask, all != 0, r64, knD, mlo, b, tl, vals);
  });
  blockSumD<3, kBlockP>(vals, partials + 3 * blockIdx.x);
}

// The FP64 force pass in persistent blocks, one per compute unit, with the
// Fe-Fe segments from knot mlo on in LDS (forceSiteD): per site the same
// pairs, in the same order and with the same values, as kForce.
template <int MODE, int U>
__global__ __launch_bounds__(kBlockP) void kForceDL(Geom G, LatOffsets O, Tables<double> tb, CXPtr x, XPtr xn,
                                                    const int8_t* __restrict__ type, const double* __restrict__ df,
                                                    const uint64_t* __restrict__ wmask, XPtr v, XPtr f,
                                                    const uint64_t* __restrict__ id, VerletParams vp, StageArgs st,
                                                    Ctl* ctl, InterView iv, Rec64 r64, Rec64 rn64, int mlo, int nkn) {
  extern __shared__ double knD[];
  const double* z = tb.pairF + static_cast<size_t>(mlo) * kPairRec;  // row (Fe, Fe) is the first
  for (int i = threadIdx.x; i < 7 * nkn; i += kBlockP) {
    const int k = i / 7, q = i - 7 * k;
    knD[i] = z[static_cast<size_t>(k) * kPairRec + q];
  }
  __syncthreads();
  unsigned long long vb = 0, db = 0;
  forVirtualBlocks(G, [&](uint32_t b, uint32_t tl) {
    unsigned long long vbs = 0, dbs = 0;
    forceSiteD<double, 0, MODE, U, true>(G, O, tb, x, xn, type, df, wmask, v, f, id, vp, st, ctl, iv, r64, rn64, knD,
                                         mlo, b, tl, vbs, dbs);
    vb = vbs > vb ? vbs : vb;
    db = dbs > db ? dbs : db;
  });
  if (MODE == 0 && vp.vmax) waveMaxTo(vb, vp.vmax);
  if (MODE == 1) waveMaxTo(db, &ctl->dispmax2);
}

// ---------------------------------------------------------------------------
// Mixed precision on site records: one 16-byte load per neighbour.
//   x_i - x_j = (u_i - u_j) - (site_j - site_i), the lattice vector of the
//   offset coming from a per-offset table in scalar registers.  The FP32
//   displacement differs from the FP64 one by a few 1e-7 relative, far inside
//   the mixed path's error bar, and the pair set differs from FP64's only for
//   pairs within that distance of the cutoff, where the tabulated functions
//   and their slopes vanish.
// ---------------------------------------------------------------------------
// The masks of pass 1 work as in FP64: both passes see the same records and
// round the distances the same way (dist2M), so a skipped offset has no pair.
// The density pass of block `blk` of one box (the kernel's own block, or a
// block of one box of a batch).
template <int ACC, int U>
__device__ __forceinline__ void rhoPassM(const Geom& G, const LatOffsets& O, const Tables<float>& tb, CXPtr x,
                                         const int8_t* __restrict__ type, float4* rec,
                                         const float4* __restrict__ dl, double* __restrict__ rho_out,
                                         uint64_t* __restrict__ wmask, uint64_t* __restrict__ smask, float rcs2,
                                         const InterView& iv, uint32_t blk) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfBlock(G, blk, threadIdx.x, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  // a wavefront next to a far lattice atom (kWide) scans in full and records its masks afresh
  const bool skin =
      __builtin_amdgcn_readfirstlane(iv.ctl->skin_mode) != 0 && !__ballot(!isVacant(ti) && (ti & kWide));
  const uint32_t wv = blk * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  Sum1<ACC> acc;
  uint64_t w0 = 0, w1 = 0;  // offsets in range for some lane (uniform among the lanes that gather)
  uint64_t s0 = 0, s1 = 0;  // offsets within cutoff + skin for some lane (full scan)
  if (!isVacant(ti)) {
    const float4 me = rec[s];
    const Rsrc rr = rsrcBelow(rec, O.lo);
    const Rsrc rtab = rsrcAt(reinterpret_cast<uintptr_t>(tb.rhoH));
    const uint32_t o16 = static_cast<uint32_t>(s) * 16u;
    // per pruned offset: the lattice vector and, in w, the record shift * 16
    const float4* D = dl + (2 + p) * kMaxLatOffsets;
    const float rc2 = static_cast<float>(G.rc2);
    const int nk = O.n[p];
    // U offsets at a time; the entries past the last repeat the group's first
    // offset, with the padding entry's lattice vector: they add nothing
    auto group = [&](const int (&ks)[U], const bool (&act)[U]) {
      float4 nb[U], d[U];
      bool in[U], sk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) d[u] = D[act[u] ? ks[u] : kPadOffset];
#pragma unroll
      for (int u = 0; u < U; ++u) nb[u] = ldRec4(rr, o16, __float_as_uint(d[u].w));
      rhoGroupM<ACC, U>(me, nb, d, rc2, rcs2, rtab, tb, acc, in, sk);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (__ballot(in[u])) setBit(w0, w1, ks[u]);
        if (!skin && __ballot(sk[u])) setBit(s0, s1, ks[u]);
      }
    };
    int ks[U];
    bool act[U];
    if (!skin) {
      for (int k = 0; k < nk; k += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          act[u] = k + u < nk;
          ks[u] = act[u] ? k + u : k;
        }
        group(ks, act);
      }
    } else {
      // the offsets of the last full scan's skin mask, in ascending order
      uint64_t b0 = uniform64(smask[2 * wv]), b1 = uniform64(smask[2 * wv + 1]);
      while (b0 | b1) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          act[u] = (b0 | b1) != 0;
          int k;
          if (b0) {
            k = __builtin_ctzll(b0);
            b0 &= b0 - 1;
          } else {
            k = b1 ? 64 + __builtin_ctzll(b1) : 0;
            b1 &= b1 - 1;
          }
          ks[u] = act[u] ? k : ks[0];
        }
        group(ks, act);
      }
    }
  }
  writeMasks(!isVacant(ti), wv, w0, w1, wmask, !skin, s0, s1, smask);
  rhoTailM<ACC>(G, s, p, cx, cy, cz, ti, acc, x, type, tb, iv, rec, rho_out);
}

template <int ACC, int U>
__global__ __launch_bounds__(kBlock) void kRhoDfM(Geom G, LatOffsets O, Tables<float> tb, CXPtr x,
                                                  const int8_t* __restrict__ type, float4* rec,
                                                  const float4* __restrict__ dl, double* __restrict__ rho_out,
                                                  uint64_t* __restrict__ wmask, uint64_t* __restrict__ smask,
                                                  float rcs2, InterView iv) {
  rhoPassM<ACC, U>(G, O, tb, x, type, rec, dl, rho_out, wmask, smask, rcs2, iv, blockIdx.x);
}

// The density pass of every box of a batch in one launch: the blocks of box
// b are [start[b], start[b+1]), in the box's own block order (the starts are
// multiples of 8, so a block keeps its XCD).
template <int ACC, int U>
__global__ __launch_bounds__(kBlock) void kRhoDfMB(const BatchBox* __restrict__ boxes, const uint32_t* start,
                                                   int nbox, int flip) {
  const int b = boxOfBlock(start, nbox, blockIdx.x);
  const BatchBox& B = boxes[b];
  const int c = B.cur ^ flip;
  rhoPassM<ACC, U>(B.G, B.O, B.t32, cxp(B.x[c]), B.type, B.rec[c], B.dl, nullptr, B.wmask, B.smask, B.rcs2, B.iv,
                   blockIdx.x - start[b]);
}

template <int ACC, int MODE, int U>
__global__ __launch_bounds__(kBlock) void kForceM(Geom G, LatOffsets O, Tables<float> tb, CXPtr x, XPtr xn,
                                                  const int8_t* __restrict__ type, const float4* __restrict__ rec,
                                                  float4* __restrict__ recn, const float4* __restrict__ dl,
                                                  const uint64_t* __restrict__ wmask, XPtr v, XPtr f,
                                                  const uint64_t* __restrict__ id, VerletParams vp, StageArgs st,
                                                  Ctl* ctl, InterView iv) {
  unsigned long long vb = 0;  // |v|^2 bits after the secondstep (MODE 0)
  unsigned long long db = 0;  // |dx|^2 bits of the drift (MODE 1)
  [&] {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfThread(G, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  if (isVacant(ti)) {  // vacancy: keeps its stale position (atom.cpp:43-47)
    vacancyForceM<MODE>(G, s, p, cx, cy, cz, x, xn, recn, f);
    return;
  }
  const int ei = species(ti);
  const float4 me = rec[s];
  const float dfi = me.w;
  const Rsrc rr = rsrcBelow(rec, O.lo);
  const Rsrc rtab = rsrcAt(reinterpret_cast<uintptr_t>(tb.pairZ));
  const Rsrc rtabD = rsrcAt(reinterpret_cast<uintptr_t>(tb.rhoD));
  const uint32_t o16 = static_cast<uint32_t>(s) * 16u;
  const uint32_t pairBase = static_cast<uint32_t>(ei * 3) * static_cast<uint32_t>(tb.nr);
  const uint32_t* sh = O.sh[p];
  const float4* D = dl + p * kMaxLatOffsets;
  const float rc2 = static_cast<float>(G.rc2);
  Sum3<ACC> acc;
  // the offsets the density pass found in range for some lane of this wavefront
  const uint32_t wv = waveId();
  // a group's entries past the mask's last offset get a lattice vector out
  // of range: beyond the cutoff, they add nothing
  forMaskGroupsFull<U>(uniform64(wmask[2 * wv]), uniform64(wmask[2 * wv + 1]), [&](const int* ks, const bool* act) {
    float4 nb[U], d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = act[u] ? ks[u] : kPadOffset;
      nb[u] = ldRec4(rr, o16, sh[k] * 16u);
      d[u] = D[k];
    }
    forceGroupM<ACC, U>(me, dfi, ei, pairBase, nb, d, rc2, rtab, rtabD, tb, ks, acc);
  });
  flushChunk3(acc);
  forceTailM<ACC, MODE>(G, s, p, cx, cy, cz, ti, ei, dfi, acc, x, xn, type, rec, recn, v, f, id, vp, st, ctl, tb,
                        iv, vb, db);
  }();
  if (MODE == 0 && vp.vmax) waveMaxTo(vb, vp.vmax);
  if (MODE == 1) waveMaxTo(db, &ctl->dispmax2);
}

// One virtual block of the LDS force pass (block b of a box): per site the
// same pairs, in the same order and with the same rounding, as kForceM.
// offT holds, per pruned offset of the centre's parity, the lattice vector
// and in w what shift(w) turns into the byte shift of the neighbour's record.
template <int ACC, int MODE, int U, typename Shift>
__device__ __forceinline__ void forceVBlockL(const Geom& G, const LatOffsets& O, const Tables<float>& tb, CXPtr x,
                                             XPtr xn, const int8_t* __restrict__ type,
                                             const float4* __restrict__ rec, float4* __restrict__ recn,
                                             const uint64_t* __restrict__ wmask, XPtr v, XPtr f,
                                             const uint64_t* __restrict__ id, const VerletParams& vp,
                                             const StageArgs& st, Ctl* ctl, const InterView& iv, const float4* offT,
                                             const float4* knT, int mlo, uint32_t b, uint32_t tl, Shift shift,
                                             unsigned long long& vb, unsigned long long& db) {
  int64_t s;
  int p, cx, cy, cz;
  if (!siteOfBlock(G, b, tl, s, p, cx, cy, cz)) return;
  const int8_t ti = type[s];
  if (isVacant(ti)) {  // vacancy: keeps its stale position (atom.cpp:43-47)
    vacancyForceM<MODE>(G, s, p, cx, cy, cz, x, xn, recn, f);
    return;
  }
  const Rsrc rr = rsrcBelow(rec, O.lo);
  const Rsrc rtab = rsrcAt(reinterpret_cast<uintptr_t>(tb.pairZ));
  const Rsrc rtabD = rsrcAt(reinterpret_cast<uintptr_t>(tb.rhoD));
  const float rc2 = static_cast<float>(G.rc2);
  const int ei = species(ti);
  const float4 me = rec[s];
  const float dfi = me.w;
  const uint32_t o16 = static_cast<uint32_t>(s) * 16u;
  const uint32_t pairBase = static_cast<uint32_t>(ei * 3) * static_cast<uint32_t>(tb.nr);
  const float4* T = offT + p * kMaxLatOffsets;
  Sum3<ACC> acc;
  const uint32_t wv = b * (kBlock / 64) + __builtin_amdgcn_readfirstlane(tl >> 6);
  // a group's entries past the mask's last offset get a lattice vector out
  // of range: beyond the cutoff, they add nothing
  forMaskGroupsFull<U>(uniform64(wmask[2 * wv]), uniform64(wmask[2 * wv + 1]), [&](const int* ks, const bool* act) {
    float4 nb[U], d[U];
#pragma unroll
    for (int u = 0; u < U; ++u) d[u] = T[act[u] ? ks[u] : kPadOffset];
#pragma unroll
    for (int u = 0; u < U; ++u) nb[u] = ldRec4(rr, o16 + shift(d[u].w), 0);
    forceGroupL<ACC, U>(me, dfi, ei, pairBase, nb, d, rc2, rtab, rtabD, tb, knT, mlo, ks, acc);
  });
  flushChunk3(acc);
  unsigned long long vbs = 0, dbs = 0;
  forceTailM<ACC, MODE>(G, s, p, cx, cy, cz, ti, ei, dfi, acc, x, xn, type, rec, recn, v, f, id, vp, st, ctl, tb,
                        iv, vbs, dbs);
  vb = vbs > vb ? vbs : vb;
  db = dbs > db ? dbs : db;
}

// The LDS knots of the force pass, per knot the Fe-Fe r*phi cubic and the Fe
// density slope cubic side by side, from knot mlo on (one block fills them).
__device__ __forceinline__ void fillKnots(const Tables<float>& tb, float4* knT, int mlo, int nkn) {
  const float4* z4 = reinterpret_cast<const float4*>(tb.pairZ);  // row (Fe, Fe) is the first
  const float4* s4 = reinterpret_cast<const float4*>(tb.rhoD);   // row Fe is the first
  for (int i = threadIdx.x; i < nkn; i += kBlockP) {
    knT[2 * i] = z4[mlo + i];
    knT[2 * i + 1] = s4[mlo + i];
  }
}

// The mixed force pass in persistent blocks, one per compute unit, that hold
// in LDS the lattice vector and record shift of every pruned offset (read as
// one broadcast per offset), the Fe-Fe r*phi cubics and the Fe density slope
// cubics from knot mlo on (one block fills them once).  Per site the same
// pairs, in the same order and with the same rounding, as kForceM.
template <int ACC, int MODE, int U>
__global__ __launch_bounds__(kBlockP) void kForceL(Geom G, LatOffsets O, Tables<float> tb, CXPtr x, XPtr xn,
                                                   const int8_t* __restrict__ type, const float4* __restrict__ rec,
                                                   float4* __restrict__ recn, const float4* __restrict__ dl,
                                                   const uint64_t* __restrict__ wmask, XPtr v, XPtr f,
                                                   const uint64_t* __restrict__ id, VerletParams vp, StageArgs st,
                                                   Ctl* ctl, InterView iv, int mlo, int nkn,
                                                   unsigned long long* zwork) {
  __shared__ float4 offT[2 * kMaxLatOffsets];  // {lattice vector, record shift * 16}
  extern __shared__ float4 knT[];              // per knot: Fe-Fe r*phi cubic, Fe density slope cubic
  for (int i = threadIdx.x; i < 2 * kMaxLatOffsets; i += kBlockP) {
    const int q = i / kMaxLatOffsets, k = i - q * kMaxLatOffsets;
    float4 d = dl[i];
    d.w = __uint_as_float(O.sh[q][k] * 16u);
    offT[i] = d;
  }
  fillKnots(tb, knT, mlo, nkn);
  __syncthreads();
  unsigned long long vb = 0;  // |v|^2 bits after the secondstep (MODE 0), largest over the lane's sites
  unsigned long long db = 0;  // |dx|^2 bits of the drift (MODE 1), likewise
  auto shift = [](float w) { return __float_as_uint(w); };
  forVirtualBlocks(G, [&](uint32_t b, uint32_t tl) {
    const long long t0 = zwork ? clock64() : 0;
    forceVBlockL<ACC, MODE, U>(G, O, tb, x, xn, type, rec, recn, wmask, v, f, id, vp, st, ctl, iv, offT, knT, mlo, b,
                               tl, shift, vb, db);
    if (zwork) addPlaneWork(G, zwork, b, tl, t0);
  });
  if (MODE == 0 && vp.vmax) waveMaxTo(vb, vp.vmax);
  if (MODE == 1) waveMaxTo(db, &ctl->dispmax2);
}

// The LDS force pass of every box of a batch in one persistent launch: on
// each XCD the blocks walk the virtual blocks of box 0, then box 1, ...
// (vstart[b] counts the virtual blocks per XCD before box b).  All boxes
// share the potential (the knots in LDS) and the pruned offsets (lattice
// vectors and packed deltas in dlp, 2 x kMaxLatOffsets); each turns a
// packed delta into its own record shift.
template <int ACC, int MODE, int U>
__global__ __launch_bounds__(kBlockP) void kForceLB(const BatchBox* __restrict__ boxes, const uint32_t* vstart,
                                                    int nbox, int flip, const float4* __restrict__ dlp, int mlo,
                                                    int nkn) {
  __shared__ float4 offT[2 * kMaxLatOffsets];
  extern __shared__ float4 knT[];
  for (int i = threadIdx.x; i < 2 * kMaxLatOffsets; i += kBlockP) offT[i] = dlp[i];
  fillKnots(boxes[0].t32, knT, mlo, nkn);
  __syncthreads();
  constexpr uint32_t V = kBlockP / kBlock;
  const uint32_t c = blockIdx.x & 7u, q = blockIdx.x >> 3, Q = gridDim.x >> 3;
  const uint32_t sub = __builtin_amdgcn_readfirstlane(threadIdx.x / kBlock);
  const uint32_t tl = threadIdx.x % kBlock;
  const uint32_t nvb = vstart[nbox];
  unsigned long long vb = 0, db = 0;
  int lastBox = -1;
  for (uint32_t k = q * V + sub; k < nvb; k += Q * V) {
    const int bi = boxOfBlock(vstart, nbox, k);
    const BatchBox& B = boxes[bi];
    if (bi != lastBox) {  // the maxima go to the box they came from
      if (lastBox >= 0) {
        if (MODE == 0 && boxes[lastBox].vp.vmax) waveMaxTo(vb, boxes[lastBox].vp.vmax);
        if (MODE == 1) waveMaxTo(db, &boxes[lastBox].ctl->dispmax2);
      }
      vb = db = 0;
      lastBox = bi;
    }
    const int cb = B.cur ^ flip;
    const int32_t plane = static_cast<int32_t>(B.G.plane), NX = B.G.NX, NY = B.G.NY, lo = B.O.lo;
    auto shift = [&](float w) {
      const uint32_t u = __builtin_amdgcn_readfirstlane(__float_as_uint(w));
      const int32_t dp = static_cast<int8_t>(u & 0xffu), dz = static_cast<int8_t>((u >> 8) & 0xffu);
      const int32_t dy = static_cast<int8_t>((u >> 16) & 0xffu), dcx = static_cast<int8_t>(u >> 24);
      return static_cast<uint32_t>(dp * plane + (dz * NY + dy) * NX + dcx - lo) * 16u;
    };
    forceVBlockL<ACC, MODE, U>(B.G, B.O, B.t32, cxp(B.x[cb]), xp(B.x[1 - cb]), B.type, B.rec[cb], B.rec[1 - cb],
                               B.wmask, xp(B.v), xp(B.f), B.id, B.vp, B.st, B.ctl, B.iv, offT, knT, mlo,
                               c + 8u * (k - vstart[bi]), tl, shift, vb, db);
  }
  if (lastBox >= 0) {
    if (MODE == 0 && boxes[lastBox].vp.vmax) waveMaxTo(vb, boxes[lastBox].vp.vmax);
    if (MODE == 1) waveMaxTo(db, &boxes[lastBox].ctl->dispmax2);
  }
}

// LDS knots of the force pass: the Fe-Fe r*phi and Fe slope cubics (32 bytes
// a knot) from knot mlo to the last, within one block per compute unit
// (160 KB of LDS) beside the kernel's static LDS (the offset table, 3.5 KB;
// 64 bytes are kept beyond it).
constexpr int kForceLdsBytes = 160 * 1024 - 2 * kMaxLatOffsets * 16 - 64;
struct LdsKnots {
  int mlo, nkn;
  size_t bytes;
};
LdsKnots ldsKnotsForce(int nr) {
  LdsKnots k;
  k.nkn = std::min(nr, kForceLdsBytes / 32);
  k.mlo = nr - k.nkn;
  k.bytes = static_cast<size_t>(32) * static_cast<size_t>(k.nkn);
  return k;
}
// The FP64 pass's LDS knots: 7 doubles (56 bytes) a knot, from knot mlo to
// the last, in one block per compute unit (the kernel has no static LDS;
// 64 bytes are kept).
constexpr int kForceLdsBytesD = 160 * 1024 - 64;
LdsKnots ldsKnotsForceD(int nr) {
  LdsKnots k;
  k.nkn = std::min(nr, kForceLdsBytesD / 56);
  k.mlo = nr - k.nkn;
  k.bytes = static_cast<size_t>(56) * static_cast<size_t>(k.nkn);
  return k;
}
// The energy pass's LDS knots: as the FP64 force pass's, beside its static
// LDS (the block sums; 512 bytes are kept).
constexpr int kEnergyLdsBytes = 160 * 1024 - 512;
LdsKnots ldsKnotsEnergy(int nr) {
  LdsKnots k;
  k.nkn = std::min(nr, kEnergyLdsBytes / 56);
  k.mlo = nr - k.nkn;
  k.bytes = static_cast<size_t>(56) * static_cast<size_t>(k.nkn);
  return k;
}
bool setLdsBytes(const void* fn, int bytes) {
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess;
}
// One persistent block per compute unit, a whole number per XCD.
unsigned persistentGrid(const DevState& s) { return static_cast<unsigned>(s.ncu / 8 * 8); }
// The LDS-table force pass: on request, or by default when every persistent
// block walks at least four virtual blocks per sub-block (a block's fill of
// its tables then costs little beside its share of the pass).
bool useLds(const DevState& s) {
  if (s.ncu < 8 || s.lds == kLdsOff) return false;
  if (s.lds == kLdsOn) return true;
  const int64_t perXcd = static_cast<int64_t>(persistentGrid(s) / 8) * (kBlockP / kBlock);
  return 2 * s.G.W.bpx >= 4 * perXcd;
}

constexpr int kUDouble = 2;

}  // namespace

void launchRhoDf(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                 bool mixed, int acc, bool store_rho, hipStream_t st) {
  double* rho = store_rho ? s.rho : nullptr;
  const unsigned grid = siteGrid(s.G);
  const InterView iv = interView(s, mixed);
  const CXPtr x = cxp(s.x[s.cur]);
  const double rcs = std::sqrt(s.G.rc2) + s.skin;
  if (!mixed)
    kRhoDf<double, 0, 4><<<grid, kBlock, 0, st>>>(s.G, o, t64, x, s.type, s.df64, rho, s.wmask, s.smask, rcs * rcs,
                                                   iv, Rec64{s.pxy[s.cur], s.pzw[s.cur]});
  else {
    const float rcs2 = static_cast<float>(rcs * rcs);
#define MISA_RHO_M(A, U)                                                                                   \
  kRhoDfM<A, U><<<grid, kBlock, 0, st>>>(s.G, o, t32, x, s.type, s.rec[s.cur], s.dl, rho, s.wmask, s.smask, \
                                        rcs2, iv)
    const bool u8 = s.uRho == 8;
    if (acc == 0) {
      if (u8) MISA_RHO_M(0, 8);
      else MISA_RHO_M(0, 4);
    } else if (acc == 1) {
      if (u8) MISA_RHO_M(1, 8);
      else MISA_RHO_M(1, 4);
    } else {
      if (u8) MISA_RHO_M(2, 8);
      else MISA_RHO_M(2, 4);
    }
#undef MISA_RHO_M
  }
}

int launchForce(const DevState& s, const LatOffsets& o, const Tables<double>& t64, const Tables<float>& t32,
                bool mixed, int acc, int mode, const VerletParams& vp, hipStream_t st) {
  int kind = mixed ? 1 : 0;
  const unsigned grid = siteGrid(s.G);
  const CXPtr x = cxp(s.x[s.cur]);
  const XPtr xn = xp(s.x[1 - s.cur]);
  const StageArgs sa = stageOf(s);
  const InterView iv = interView(s, mixed);
  const bool lds = useLds(s);
  const LdsKnots lk = ldsKnotsForce(t32.nr);
  const LdsKnots lkd = ldsKnotsForceD(t64.nr);
#define MISA_FORCE(M)                                                                                         \
  do {                                                                                                        \
    static const bool ldsReady =                                                                              \
        setLdsBytes(reinterpret_cast<const void*>(&kForceDL<M, kUDouble>), kForceLdsBytesD);                 \
    const Rec64 r64{s.pxy[s.cur], s.pzw[s.cur]}, rn64{s.pxy[1 - s.cur], s.pzw[1 - s.cur]};                   \
    if (lds && ldsReady) {                                                                                    \
      kind = 3;                                                                                               \
      kForceDL<M, kUDouble><<<persistentGrid(s), kBlockP, lkd.bytes, st>>>(                                   \
          s.G, o, t64, x, xn, s.type, s.df64, s.wmask, xp(s.v), xp(s.f), s.id, vp, sa, s.ctl, iv, r64, rn64,   \
          lkd.mlo, lkd.nkn);                                                                                  \
    } else {                                                                                                  \
      kForce<double, 0, M, kUDouble><<<grid, kBlock, 0, st>>>(s.G, o, t64, x, xn, s.type, s.df64, s.wmask,    \
                                                             xp(s.v), xp(s.f), s.id, vp, sa, s.ctl, iv, r64, \
                                                             rn64);                                          \
    }                                                                                                         \
  } while (0)
#define MISA_FORCE_MU(A, M, U)                                                                                \
  do {                                                                                                        \
    static const bool ldsReady =                                                                              \
        setLdsBytes(reinterpret_cast<const void*>(&kForceL<A, M, U>), kForceLdsBytes);                       \
    if (lds && ldsReady) {                                                                                    \
      kind = 2;                                                                                               \
      kForceL<A, M, U><<<persistentGrid(s), kBlockP, lk.bytes, st>>>(                                         \
          s.G, o, t32, x, xn, s.type, s.rec[s.cur], s.rec[1 - s.cur], s.dl, s.wmask, xp(s.v), xp(s.f), s.id, vp, \
          sa, s.ctl, iv, lk.mlo, lk.nkn, s.zmeasure ? s.zwork : nullptr);                                     \
    } else {                                                                                                  \
      kForceM<A, M, U><<<grid, kBlock, 0, st>>>(s.G, o, t32, x, xn, s.type, s.rec[s.cur], s.rec[1 - s.cur],    \
                                                s.dl, s.wmask, xp(s.v), xp(s.f), s.id, vp, sa, s.ctl, iv);    \
    }                                                                                                         \
  } while (0)
#define MISA_FORCE_M(A, M)            \
  do {                                \
    if (s.uForce == 2)                \
      MISA_FORCE_MU(A, M, 2);         \
    else                              \
      MISA_FORCE_MU(A, M, 4);         \
  } while (0)
  if (!mixed) {
    if (mode == 0) MISA_FORCE(0);
    else if (mode == 1) MISA_FORCE(1);
    else MISA_FORCE(2);
  } else if (acc == 0) {
    if (mode == 0) MISA_FORCE_M(0, 0);
    else if (mode == 1) MISA_FORCE_M(0, 1);
    else MISA_FORCE_M(0, 2);
  } else if (acc == 1) {
    if (mode == 0) MISA_FORCE_M(1, 0);
    else if (mode == 1) MISA_FORCE_M(1, 1);
    else MISA_FORCE_M(1, 2);
  } else {
    if (mode == 0) MISA_FORCE_M(2, 0);
    else if (mode == 1) MISA_FORCE_M(2, 1);
    else MISA_FORCE_M(2, 2);
  }
#undef MISA_FORCE
#undef MISA_FORCE_M
#undef MISA_FORCE_MU
  return kind;
}

int64_t launchEnergyPass(const DevState& s, const LatOffsets& o, const Tables<double>& t64, bool every,
                         hipStream_t st) {
  static const bool ldsReady = setLdsBytes(reinterpret_cast<const void*>(&kEnergyDL), kEnergyLdsBytes);
  const Rec64 r64{s.pxy[s.cur], s.pzw[s.cur]};
  const int all = every ? 1 : 0;
  if (useLds(s) && ldsReady && 3 * static_cast<int64_t>(persistentGrid(s)) <= s.npartials) {
    const LdsKnots lk = ldsKnotsEnergy(t64.nr);
    const unsigned grid = persistentGrid(s);
    kEnergyDL<<<grid, kBlockP, lk.bytes, st>>>(s.G, o, s.in.full, t64, cxp(s.x[s.cur]), s.type, s.rho, s.df64,
                                               s.smask, all, r64, s.partials, lk.mlo, lk.nkn);
    return grid;
  }
  const unsigned grid = siteGrid(s.G);  // 3 partials per block fit: the buffer holds 7 per site block
  kEnergyD<<<grid, kBlock, 0, st>>>(s.G, o, s.in.full, t64, cxp(s.x[s.cur]), s.type, s.rho, s.df64, s.smask, all,
                                    r64, s.partials);
  return grid;
}

void launchRhoDfBatch(const BatchBox* boxes, const uint32_t* rhoStart, int nbox, unsigned blocks, int acc,
                      int flip, hipStream_t st) {
  if (acc == 0) kRhoDfMB<0, 4><<<blocks, kBlock, 0, st>>>(boxes, rhoStart, nbox, flip);
  else if (acc == 1) kRhoDfMB<1, 4><<<blocks, kBlock, 0, st>>>(boxes, rhoStart, nbox, flip);
  else kRhoDfMB<2, 4><<<blocks, kBlock, 0, st>>>(boxes, rhoStart, nbox, flip);
}

bool launchForceBatch(const BatchBox* boxes, const uint32_t* vstart, int nbox, const float4* dl, int nr, int ncu,
                      int acc, int flip, hipStream_t st) {
  const LdsKnots lk = ldsKnotsForce(nr);
  const unsigned grid = static_cast<unsigned>(ncu / 8 * 8);
#define MISA_FORCE_B(A)                                                                                      \
  do {                                                                                                       \
    static const bool ready = setLdsBytes(reinterpret_cast<const void*>(&kForceLB<A, 1, 4>), kForceLdsBytes); \
    if (!ready) return false;                                                                                \
    kForceLB<A, 1, 4><<<grid, kBlockP, lk.bytes, st>>>(boxes, vstart, nbox, flip, dl, lk.mlo, lk.nkn);        \
  } while (0)
  if (acc == 0) MISA_FORCE_B(0);
  else if (acc == 1) MISA_FORCE_B(1);
  else MISA_FORCE_B(2);
#undef MISA_FORCE_B
  return true;
}

}  // namespace misa
