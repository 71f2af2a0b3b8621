// The per-site arithmetic of the mixed-precision passes, shared by the kernel
// that gathers neighbour records straight from memory and the kernel that
// stages them in LDS: the density and force terms of a group of neighbours,
// and what follows the lattice sum of a site (inter atoms, the full-list
// remainder near far lattice atoms, F', the half kicks, drift and staging,
// the stores and the periodic images).
//
// A neighbour is given by its 16-byte site record {u, w} and the lattice
// vector from the centre's site to its site:
//   x_i - x_j = (u_i - u_j) - (site_j - site_i).
#pragma once
#include <hip/hip_runtime.h>

#include "buffer_load.hpp"
#include "device_common.hpp"
#include "inter_common.hpp"
#include "resident.hpp"
#include "site_common.hpp"

namespace misa {
namespace site {

// x_i - x_j and |x_i - x_j|^2 of a neighbour, rounded the same way in every
// pass, so that the density and force passes see the same pair set.
__device__ __forceinline__ float dist2M(const float4& me, const float4& nb, const float4& d, float& d0, float& d1,
                                        float& d2) {
  d0 = __fsub_rn(__fsub_rn(me.x, nb.x), d.x);
  d1 = __fsub_rn(__fsub_rn(me.y, nb.y), d.y);
  d2 = __fsub_rn(__fsub_rn(me.z, nb.z), d.z);
  return __fmaf_rn(d2, d2, __fmaf_rn(d1, d1, __fmul_rn(d0, d0)));
}

// Value of the cubic Hermite segment between knots {y0, s0} and {y1, s1}
// (slopes per unit knot coordinate) at local coordinate q: the segment's
// coefficients c4 = 3 (y1 - y0) - 2 s0 - s1 and c3 = s0 + s1 - 2 (y1 - y0),
// as the eam/alloy interpolation defines them.  y1 - y0 is exact for
// neighbouring knots of like sign, so the value is as accurate as the knots.
__device__ __forceinline__ float hermiteValue(const float4& k, float q) {
#pragma clang fp contract(off)
  const float dy = k.z - k.x;
  const float c4 = fmaf(3.0f, dy, -fmaf(2.0f, k.y, k.w));
  const float c3 = fmaf(-2.0f, dy, k.y + k.w);
  return fmaf(fmaf(fmaf(c3, q, c4), q, k.y), q, k.x);
}

// rho_i += sum over the group of rho_{t_j}(r_ij) (pot eam::chargeDensity);
// in[u] tells whether neighbour u is within the cutoff, sk[u] whether it is
// within the cutoff plus the neighbour skin (r^2 < rcs2).
template <int ACC, int V>
__device__ __forceinline__ void rhoGroupM(const float4& me, const float4 (&nb)[V], const float4 (&d)[V], float rc2,
                                          float rcs2, Rsrc rtab, const Tables<float>& tb, Sum1<ACC>& acc,
                                          bool (&in)[V], bool (&sk)[V]) {
#pragma clang fp contract(off)  // explicit FMAs only: every instantiation rounds alike
  float pp[V];
  float4 cr[V];
#pragma unroll
  for (int u = 0; u < V; ++u) {
    float d0, d1, d2;
    const float r2 = dist2M(me, nb[u], d[u], d0, d1, d2);
    in[u] = r2 < rc2;
    sk[u] = r2 < rcs2;
    if (in[u]) {
      const int m = knotM(r2 * __builtin_amdgcn_rsqf(r2), tb.rdr, tb.nr, pp[u]);
      const uint32_t row = __umul24(static_cast<uint32_t>(recSpecies(nb[u].x)), static_cast<uint32_t>(tb.nr));
      cr[u] = ldRec4(rtab, (row + static_cast<uint32_t>(m)) * 8u, 0);  // {y0, s0, y1, s1}
    }
  }
#pragma unroll
  for (int u = 0; u < V; ++u)
    if (in[u]) acc.add(hermiteValue(cr[u], pp[u]));
}

// fpair of one pair (pot eam::toForce) from its segments: the value cubic
// of r*phi {c3, c4, c5, c6}, the slope cubic of the neighbour's density and,
// for an unlike pair, of the centre's density, each already divided by the
// knot spacing.  Shared by the force passes that gather the segments from
// global memory and from LDS, so that both round alike.
__device__ __forceinline__ float pairFpM(float q, float rinv, const float4& z, const f32x3& sj, const f32x3& si,
                                         bool like, float dfi, float dfj, float rdr) {
#pragma clang fp contract(off)  // explicit FMAs only: every instantiation rounds alike
  // c3 q^3 + c4 q^2 + c5 q + c6 and its slope 3 c3 q^2 + 2 c4 q + c5 from
  // one chain: a = c3 q + c4, b = a q + c5, slope = (c3 q + a) q + b
  const float a = fmaf(z.x, q, z.y);
  const float b = fmaf(a, q, z.z);
  const float z2 = fmaf(b, q, z.w);
  const float z2p = fmaf(fmaf(z.x, q, a), q, b) * rdr;
  const float rhojp = fmaf(fmaf(sj.x, q, sj.y), q, sj.z);
  const float rhoip = like ? rhojp : fmaf(fmaf(si.x, q, si.y), q, si.z);
  const float phi = z2 * rinv;
  const float phip = (z2p - phi) * rinv;
  const float psip = fmaf(dfi, rhojp, fmaf(dfj, rhoip, phip));
  return -psip * rinv;
}

// f_i += sum over the group of (x_i - x_j) fpair_ij, the segments gathered
// from the tables in global memory (16-byte knots).
template <int ACC, int V>
__device__ __forceinline__ void forceGroupM(const float4& me, float dfi, int ei, uint32_t pairBase,
                                            const float4 (&nb)[V], const float4 (&d)[V], float rc2, Rsrc rtab,
                                            Rsrc rtabD, const Tables<float>& tb, const int* ks, Sum3<ACC>& acc) {
#pragma clang fp contract(off)
  bool in[V], like[V];
  float d0[V], d1[V], d2[V], pp[V], ri[V];
  float4 zc[V];
  f32x3 sj[V], si[V];
#pragma unroll
  for (int u = 0; u < V; ++u) {
    const float r2 = dist2M(me, nb[u], d[u], d0[u], d1[u], d2[u]);
    in[u] = r2 < rc2;
    if (in[u]) {
      ri[u] = __builtin_amdgcn_rsqf(r2);
      const int m = knotM(r2 * ri[u], tb.rdr, tb.nr, pp[u]);
      const uint32_t ej = static_cast<uint32_t>(recSpecies(nb[u].x));
      const uint32_t rowj = __umul24(ej, static_cast<uint32_t>(tb.nr)) + static_cast<uint32_t>(m);
      zc[u] = ldRec4(rtab, (pairBase + rowj) * 16u, 0);
      sj[u] = ld3(rtabD, rowj * 16u);
      like[u] = ej == static_cast<uint32_t>(ei);
      si[u] = f32x3{0.0f, 0.0f, 0.0f};
      if (!like[u])  // rho_ei differs from rho_ej only for unlike pairs
        si[u] = ld3(rtabD, (__umul24(static_cast<uint32_t>(ei), static_cast<uint32_t>(tb.nr)) +
                            static_cast<uint32_t>(m)) * 16u);
    }
  }
  float fp[V];
#pragma unroll
  for (int u = 0; u < V; ++u) {
    fp[u] = 0.0f;  // the segments of a neighbour beyond the cutoff are not read
    if (in[u]) fp[u] = pairFpM(pp[u], ri[u], zc[u], sj[u], si[u], like[u], dfi, nb[u].w, tb.rdr);
  }
  addGroup3<ACC, V>(acc, ks, in, d0, d1, d2, fp);
}

// forceGroupM with the Fe-Fe r*phi cubics and the Fe density slope cubics
// from knot mlo on in LDS (knT, the two of a knot side by side): every lane
// reads its segments there,
// at a clamped knot; a lane whose pair involves another species, or lies
// below knot mlo, then re-reads what it needs from global memory, in a
// branch that only a wavefront with such a lane enters (after the LDS reads,
// so that no global load waits in front of them).
template <int ACC, int V>
__device__ __forceinline__ void forceGroupL(const float4& me, float dfi, int ei, uint32_t pairBase,
                                            const float4 (&nb)[V], const float4 (&d)[V], float rc2, Rsrc rtab,
                                            Rsrc rtabD, const Tables<float>& tb, const float4* knT, int mlo,
                                            const int* ks, Sum3<ACC>& acc) {
#pragma clang fp contract(off)
  bool in[V], like[V], fz[V];
  float d0[V], d1[V], d2[V], pp[V], ri[V];
  int mm[V];
  uint32_t ejv[V];
  float4 zc[V];
  f32x3 sj[V], si[V];
  bool any = false;
#pragma unroll
  for (int u = 0; u < V; ++u) {
    const float r2 = dist2M(me, nb[u], d[u], d0[u], d1[u], d2[u]);
    in[u] = r2 < rc2;
    ri[u] = __builtin_amdgcn_rsqf(r2);
    mm[u] = knotM(r2 * ri[u], tb.rdr, tb.nr, pp[u]);
    ejv[u] = static_cast<uint32_t>(recSpecies(nb[u].x));
    like[u] = ejv[u] == static_cast<uint32_t>(ei);
    const float4* k2 = knT + 2 * max(mm[u] - mlo, 0);  // {r*phi cubic, density slope cubic} of the knot
    zc[u] = k2[0];
    const float4 s4 = k2[1];
    sj[u] = f32x3{s4.x, s4.y, s4.z};
    si[u] = sj[u];
    fz[u] = in[u] && ((static_cast<uint32_t>(ei) | ejv[u]) != 0u || mm[u] < mlo);
    any |= fz[u];
  }
  if (__ballot(any)) {  // the pair needs r*phi from memory; the slopes as their species or knot say
#pragma unroll
    for (int u = 0; u < V; ++u) {
      if (!fz[u]) continue;
      const bool low = mm[u] < mlo;
      const uint32_t rowj = __umul24(ejv[u], static_cast<uint32_t>(tb.nr)) + static_cast<uint32_t>(mm[u]);
      zc[u] = ldRec4(rtab, (pairBase + rowj) * 16u, 0);
      if (ejv[u] != 0u || low) sj[u] = ld3(rtabD, rowj * 16u);
      if (!like[u] && (ei != 0 || low))
        si[u] = ld3(rtabD, (__umul24(static_cast<uint32_t>(ei), static_cast<uint32_t>(tb.nr)) +
                            static_cast<uint32_t>(mm[u])) * 16u);
    }
  }
  float fp[V];
#pragma unroll
  for (int u = 0; u < V; ++u) {
    fp[u] = 0.0f;  // the segments of a neighbour beyond the cutoff are not read
    if (in[u]) fp[u] = pairFpM(pp[u], ri[u], zc[u], sj[u], si[u], like[u], dfi, nb[u].w, tb.rdr);
  }
  addGroup3<ACC, V>(acc, ks, in, d0, d1, d2, fp);
}

// After the lattice sum of site s: the inter atoms and, near a far lattice
// atom, the rest of the full list; then F'(rho) into the w of the record and
// of its periodic images (and rho when asked for).  A vacancy gets F' = 0.
template <int ACC>
__device__ __forceinline__ void rhoTailM(const Geom& G, int64_t s, int p, int cx, int cy, int cz, int8_t ti,
                                         Sum1<ACC>& acc, CXPtr x, const int8_t* __restrict__ type,
                                         const Tables<float>& tb, const InterView& iv, float4* rec,
                                         double* __restrict__ rho_out) {
  float dfi = 0.0f;
  double rho = 0.0;
  if (!isVacant(ti)) {
    if (ti & (kNearInter | kWide)) {
      const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
      const FullPlaneOffsets& full = *iv.full;
      if (ti & kNearInter) {  // inter atoms whose nearest site is within the full list (atom.cpp:224-240)
        const int ne = static_cast<int>(iv.ctl->n_inter + iv.ctl->n_ghost);
        forNearBuckets(iv, full, p, s, ne, [&](int e) {
          const double d0 = xi0 - iv.e_x[0][e], d1 = xi1 - iv.e_x[1][e], d2 = xi2 - iv.e_x[2][e];
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          if (r2 < G.rc2) acc.add(rhoPair(tb, species(iv.type[iv.e_src[e]]), r2));
        });
      }
      if (ti & kWide) {  // full-list offsets beyond the pruned list, near a far lattice atom
        forWideTargets(full, type, p, s, ti, [&](int64_t j, int8_t tj) {
          const double d0 = xi0 - x.c[0][j], d1 = xi1 - x.c[1][j], d2 = xi2 - x.c[2][j];
          const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
          if (r2 < G.rc2) acc.add(rhoPair(tb, species(tj), r2));
        });
      }
    }
    rho = acc.value();
    checkRho(iv.ctl, rho, static_cast<double>(tb.rhomax));
    dfi = dEmbed(tb, species(ti), static_cast<float>(rho));
  }
  float* w = reinterpret_cast<float*>(rec) + 3;
  w[4 * s] = dfi;
  if (rho_out) rho_out[s] = rho;
  forEachImage(G, p, cx, cy, cz, [&](int64_t q, int, int, int) {
    w[4 * q] = dfi;
    if (rho_out) rho_out[q] = rho;
  });
}

// A vacancy in the force pass: f = 0 (clearForce, and latForce skips it), or
// in the fused pass its stale position and vacancy record into the next buffers.
template <int MODE>
__device__ __forceinline__ void vacancyForceM(const Geom& G, int64_t s, int p, int cx, int cy, int cz, CXPtr x,
                                              XPtr xn, float4* __restrict__ recn, XPtr f) {
  if (MODE != 1) {  // atom.cpp:86-100, 311-358
    f.c[0][s] = 0.0;
    f.c[1][s] = 0.0;
    f.c[2][s] = 0.0;
    return;
  }
  const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
  xn.c[0][s] = xi0;
  xn.c[1][s] = xi1;
  xn.c[2][s] = xi2;
  recn[s] = vacantRec();
  forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
    xn.c[0][q] = xi0 + sx * G.L0;
    xn.c[1][q] = xi1 + sy * G.L1;
    xn.c[2][q] = xi2 + sz * G.L2;
    recn[q] = vacantRec();
  });
}

// After the lattice sum of an occupied site: inter atoms, the full-list
// remainder near far lattice atoms, then
//   MODE 0: f stored, v += c f                        (secondstep)
//   MODE 1: v += c f; v += c f; x' = x + dt v; decide  (secondstep + next firststep)
//   MODE 2: f stored only                             (first force evaluation)
template <int ACC, int MODE>
__device__ __forceinline__ void forceTailM(const Geom& G, int64_t s, int p, int cx, int cy, int cz, int8_t ti,
                                           int ei, float dfi, Sum3<ACC>& acc, CXPtr x, XPtr xn,
                                           const int8_t* __restrict__ type, const float4* __restrict__ rec,
                                           float4* __restrict__ recn, XPtr v, XPtr f,
                                           const uint64_t* __restrict__ id, const VerletParams& vp,
                                           const StageArgs& st, Ctl* ctl, const Tables<float>& tb,
                                           const InterView& iv, unsigned long long& vb, unsigned long long& db) {
#pragma clang fp contract(off)  // the half kicks and drift round as newton_motion.cpp does
  const double xi0 = x.c[0][s], xi1 = x.c[1][s], xi2 = x.c[2][s];
  if (ti & (kNearInter | kWide)) {
    const FullPlaneOffsets& full = *iv.full;
    if (ti & kNearInter) {  // forces from inter atoms on lattice neighbours (atom.cpp:397-422)
      const int ne = static_cast<int>(iv.ctl->n_inter + iv.ctl->n_ghost);
      forNearBuckets(iv, full, p, s, ne, [&](int e) {
        const int src = iv.e_src[e];
        const double e0 = xi0 - iv.e_x[0][e], e1 = xi1 - iv.e_x[1][e], e2 = xi2 - iv.e_x[2][e];
        const double r2 = e0 * e0 + e1 * e1 + e2 * e2;
        if (r2 < G.rc2) acc.add(e0, e1, e2, pairForce(tb, ei, species(iv.type[src]), r2, dfi, iv.df32[src]));
      });
    }
    if (ti & kWide) {  // full-list offsets beyond the pruned list, near a far lattice atom
      forWideTargets(full, type, p, s, ti, [&](int64_t j, int8_t tj) {
        const double e0 = xi0 - x.c[0][j], e1 = xi1 - x.c[1][j], e2 = xi2 - x.c[2][j];
        const double r2 = e0 * e0 + e1 * e1 + e2 * e2;
        if (r2 < G.rc2) acc.add(e0, e1, e2, pairForce(tb, ei, species(tj), r2, dfi, rec[j].w));
      });
    }
  }
  double f0, f1, f2;
  sum3Value<ACC>(acc, f0, f1, f2);
  checkFinite(ctl, f0, f1, f2);
  const StepDt sd = stepDt(vp, ei);
  const double c = sd.c;
  double v0 = v.c[0][s] + c * f0;
  double v1 = v.c[1][s] + c * f1;
  double v2 = v.c[2][s] + c * f2;
  if (MODE != 1) {
    f.c[0][s] = f0;
    f.c[1][s] = f1;
    f.c[2][s] = f2;
    if (MODE == 0) {
      v.c[0][s] = v0;
      v.c[1][s] = v1;
      v.c[2][s] = v2;
      vb = speed2Bits(v0, v1, v2);
    }
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
  const float4 r = siteRec(G, p, cx, cy, cz, y0, y1, y2, ei);
  recn[s] = r;
  forEachImage(G, p, cx, cy, cz, [&](int64_t q, int sx, int sy, int sz) {
    xn.c[0][q] = y0 + sx * G.L0;
    xn.c[1][q] = y1 + sy * G.L1;
    xn.c[2][q] = y2 + sz * G.L2;
    recn[q] = r;
  });
}

}  // namespace site
}  // namespace misa
