// Hook-mode kernels: the reference's accelerator hooks hip_eam_rho_calc,
// hip_eam_df_calc and hip_eam_force_calc (src/arch/arch_imp.h:27-31) computed
// directly on a device copy of the host AtomList array (104-byte AtomElement
// slots, ghost-extended, doubled x index).
//
// One thread per sub-box slot, launched per parity so every wavefront walks
// one of the reference's full offset lists (NeighbourIndex even/odd,
// neighbour_index.inl:13-76) with one uniform linear offset per step.  The
// full list makes each lane's sum complete: rho and f of the sub-box slots
// come out final and the ghost slots keep the zeros atom::clearForce left,
// so the reference's reverse halos (rho_packer.cpp, force_packer.cpp) that
// follow the hooks add nothing.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "hook.hpp"

namespace misa {

namespace {

constexpr int kHookBlock = 256;

struct HookGeom {
  int n0, n1, n2, g;
  int64_t sx, sy;
  double rc2;
};

__device__ __forceinline__ bool hookSite(const HookGeom& G, int64_t& r) {
  const int p = blockIdx.y;
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kHookBlock + threadIdx.x;
  const int64_t nsub = static_cast<int64_t>(G.n0) * G.n1 * G.n2;
  if (t >= nsub) return false;
  const int64_t cx = t % G.n0, rest = t / G.n0;
  const int64_t cy = rest % G.n1, cz = rest / G.n1;
  r = ((cz + G.g) * G.sy + (cy + G.g)) * G.sx + (2 * (cx + G.g) + p);
  return true;
}

__global__ __launch_bounds__(kHookBlock) void kHookRho(HookGeom G, FullOffsets O, Tables<double> tb,
                                                       AtomElement* __restrict__ a) {
  int64_t r;
  if (!hookSite(G, r)) return;
  const AtomElement& c = a[r];
  if (c.type < 0) return;
  const int p = blockIdx.y;
  double rho = 0.0;
  for (int k = 0; k < O.n[p]; ++k) {
    const AtomElement& n = a[r + O.off[p][k]];
    if (n.type < 0) continue;
    const double d0 = c.x[0] - n.x[0], d1 = c.x[1] - n.x[1], d2 = c.x[2] - n.x[2];
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (r2 < G.rc2) rho += rhoPair(tb, n.type, r2);
  }
  a[r].rho += rho;
}

__global__ __launch_bounds__(kHookBlock) void kHookDf(HookGeom G, Tables<double> tb, AtomElement* __restrict__ a) {
  int64_t r;
  if (!hookSite(G, r)) return;
  if (a[r].type < 0) return;
  a[r].df = dEmbed(tb, a[r].type, a[r].rho);
}

__global__ __launch_bounds__(kHookBlock) void kHookForce(HookGeom G, FullOffsets O, Tables<double> tb,
                                                         AtomElement* __restrict__ a) {
  int64_t r;
  if (!hookSite(G, r)) return;
  const AtomElement& c = a[r];
  if (c.type < 0) return;
  const int p = blockIdx.y;
  double f0 = 0, f1 = 0, f2 = 0;
  for (int k = 0; k < O.n[p]; ++k) {
    const AtomElement& n = a[r + O.off[p][k]];
    const double d0 = c.x[0] - n.x[0], d1 = c.x[1] - n.x[1], d2 = c.x[2] - n.x[2];
    const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (r2 < G.rc2 && n.type >= 0) {
      const double fp = pairForce(tb, c.type, n.type, r2, c.df, n.df);
      f0 = fma(d0, fp, f0);
      f1 = fma(d1, fp, f1);
      f2 = fma(d2, fp, f2);
    }
  }
  a[r].f[0] += f0;
  a[r].f[1] += f1;
  a[r].f[2] += f2;
}

HookGeom hookGeom(const Box& b, double rc) {
  HookGeom G;
  G.n0 = static_cast<int>(b.n[0]);
  G.n1 = static_cast<int>(b.n[1]);
  G.n2 = static_cast<int>(b.n[2]);
  G.g = b.g;
  G.sx = b.sx;
  G.sy = b.sy;
  G.rc2 = rc * rc;
  return G;
}

dim3 hookGrid(const Box& b) {
  const int64_t nsub = b.n[0] * b.n[1] * b.n[2];
  return dim3(static_cast<unsigned>((nsub + kHookBlock - 1) / kHookBlock), 2, 1);
}

}  // namespace

void launchHook(int which, const Box& b, double rc, const FullOffsets& o, const Tables<double>& tb, AtomElement* a,
                hipStream_t st) {
  const HookGeom G = hookGeom(b, rc);
  const dim3 grid = hookGrid(b);
  if (which == 0)
    kHookRho<<<grid, kHookBlock, 0, st>>>(G, o, tb, a);
  else if (which == 1)
    kHookDf<<<grid, kHookBlock, 0, st>>>(G, tb, a);
  else
    kHookForce<<<grid, kHookBlock, 0, st>>>(G, o, tb, a);
}

}  // namespace misa
