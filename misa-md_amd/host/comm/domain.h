// The domain API of libcomm (v0.3.3, not part of the reference tree) that
// the reference's simulation, world builder, packers and WS helpers read,
// for one rank that owns the whole periodic box.  Field names and meanings
// follow their uses in the reference:
//   simulation.cpp:28-50 (Builder: phase space, cutoff factor, lattice
//     constant, ghost size ceil(f) + 1), atom.cpp:14-19,154-156,288-290,
//     316-318 (ghost-extended sizes; the sub-box starts at
//     dbx_lattice_size_ghost), output_base_interface.h:23-35 (begin/end of the
//     sub-box in the ghost-extended array), ws_utils.cpp:66-200 (measured
//     regions), atom_list.cpp:25-49 and inter_border_packer.cpp:24-36
//     (fwCommLocalRegion, Region::isIn).
// "dbx" fields count x in half lattice constants (the BCC basis doubles x).
// Lattice regions are [low, high) in global lattice coordinates; measured
// ("meas") regions are in Angstrom.
#pragma once
#include <cmath>
#include <cstdint>

namespace comm {

typedef int64_t _type_lattice_size;
typedef int64_t _type_lattice_coord;

enum { DIR_LOWER = 0, DIR_HIGHER = 1 };

template <typename T>
struct Region {
  T x_low = 0, y_low = 0, z_low = 0, x_high = 0, y_high = 0, z_high = 0;
  Region() = default;
  Region(T xl, T yl, T zl, T xh, T yh, T zh) : x_low(xl), y_low(yl), z_low(zl), x_high(xh), y_high(yh), z_high(zh) {}
  bool isIn(T x, T y, T z) const {
    return x >= x_low && x < x_high && y >= y_low && y < y_high && z >= z_low && z < z_high;
  }
  T volume() const { return (x_high - x_low) * (y_high - y_low) * (z_high - z_low); }
};

// the communicator of the domain: one rank here
struct mpi_process {
  int own_rank = 0;
  int all_ranks = 1;
};

struct Domain {
  int64_t phase_space[3] = {0, 0, 0};
  double lattice_const = 0;
  double cutoff_radius_factor = 0;
  _type_lattice_size cut_lattice = 0;
  // process grid: one rank, its own neighbour in every direction
  int grid_size[3] = {1, 1, 1};
  int grid_coord[3] = {0, 0, 0};
  int rank_id_neighbours[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  // measured box and sub-box
  double meas_global_length[3] = {0, 0, 0};
  Region<double> meas_global_region;
  Region<double> meas_sub_box_region;
  // lattice sizes and regions (x in lattice constants)
  _type_lattice_size sub_box_lattice_size[3] = {0, 0, 0};
  _type_lattice_size lattice_size_ghost[3] = {0, 0, 0};
  _type_lattice_size ghost_extended_lattice_size[3] = {0, 0, 0};
  Region<_type_lattice_coord> sub_box_lattice_region;
  Region<_type_lattice_coord> ghost_ext_lattice_region;
};

struct BccDomain : Domain {
  // the same with x in half lattice constants
  _type_lattice_size dbx_sub_box_lattice_size[3] = {0, 0, 0};
  _type_lattice_size dbx_lattice_size_ghost[3] = {0, 0, 0};
  _type_lattice_size dbx_ghost_extended_lattice_size[3] = {0, 0, 0};
  Region<_type_lattice_coord> dbx_sub_box_lattice_region;
  Region<_type_lattice_coord> dbx_ghost_ext_lattice_region;

  class Builder {
   public:
    Builder& setComm(const mpi_process& p, void* new_comm = nullptr) {
      proc_ = p;
      (void)new_comm;
      return *this;
    }
    Builder& setPhaseSpace(const int64_t ps[3]) {
      for (int d = 0; d < 3; d++) ps_[d] = ps[d];
      return *this;
    }
    Builder& setCutoffRadius(double f) {
      f_ = f;
      return *this;
    }
    Builder& setLatticeConst(double a) {
      a_ = a;
      return *this;
    }
    Builder& setGhostSize(int g) {
      g_ = g;
      return *this;
    }
    // null unless one rank, positive sizes, a > 0, f > 0
    BccDomain* build() const;

   private:
    mpi_process proc_;
    int64_t ps_[3] = {0, 0, 0};
    double f_ = 0, a_ = 0;
    int g_ = -1;  // default: cut_lattice
  };
};

inline BccDomain* BccDomain::Builder::build() const {
  if (proc_.all_ranks != 1 || proc_.own_rank != 0 || !(a_ > 0) || !(f_ > 0)) return nullptr;
  for (int d = 0; d < 3; d++)
    if (ps_[d] < 1) return nullptr;
  auto* b = new BccDomain;
  b->lattice_const = a_;
  b->cutoff_radius_factor = f_;
  b->cut_lattice = static_cast<_type_lattice_size>(std::ceil(f_));
  const _type_lattice_size g = g_ >= 0 ? g_ : b->cut_lattice;
  for (int d = 0; d < 3; d++) {
    const int64_t n = ps_[d];
    b->phase_space[d] = n;
    b->meas_global_length[d] = static_cast<double>(n) * a_;
    b->sub_box_lattice_size[d] = n;
    b->lattice_size_ghost[d] = g;
    b->ghost_extended_lattice_size[d] = n + 2 * g;
    const int64_t k = d == 0 ? 2 : 1;
    b->dbx_sub_box_lattice_size[d] = k * n;
    b->dbx_lattice_size_ghost[d] = k * g;
    b->dbx_ghost_extended_lattice_size[d] = k * (n + 2 * g);
  }
  const auto& L = b->meas_global_length;
  b->meas_global_region = Region<double>(0, 0, 0, L[0], L[1], L[2]);
  b->meas_sub_box_region = b->meas_global_region;
  const int64_t* n = b->phase_space;
  b->sub_box_lattice_region = Region<_type_lattice_coord>(0, 0, 0, n[0], n[1], n[2]);
  b->ghost_ext_lattice_region = Region<_type_lattice_coord>(-g, -g, -g, n[0] + g, n[1] + g, n[2] + g);
  b->dbx_sub_box_lattice_region = Region<_type_lattice_coord>(0, 0, 0, 2 * n[0], n[1], n[2]);
  b->dbx_ghost_ext_lattice_region =
      Region<_type_lattice_coord>(-2 * g, -g, -g, 2 * (n[0] + g), n[1] + g, n[2] + g);
  return b;
}

// The layer of the local ghost-extended array (indices from 0, x doubled)
// that is sent to the neighbour in (dimension, direction): the ghost-wide
// slab of the sub-box next to that face, spanning the whole ghost-extended
// range in the dimensions already exchanged and the sub-box in the later
// ones, so that corner and edge ghosts are forwarded through the faces.
inline Region<_type_lattice_coord> fwCommLocalRegion(const BccDomain* d, int dimension, int direction) {
  _type_lattice_coord lo[3], hi[3];
  for (int k = 0; k < 3; k++) {
    const _type_lattice_coord g = d->dbx_lattice_size_ghost[k], s = d->dbx_sub_box_lattice_size[k];
    if (k == dimension) {
      lo[k] = direction == DIR_LOWER ? g : s;
      hi[k] = direction == DIR_LOWER ? 2 * g : s + g;
    } else if (k < dimension) {
      lo[k] = 0;
      hi[k] = d->dbx_ghost_extended_lattice_size[k];
    } else {
      lo[k] = g;
      hi[k] = g + s;
    }
  }
  return Region<_type_lattice_coord>(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]);
}

// The ghost layer on the other side that the data sent in (dimension,
// direction) lands in at one rank: the same slab, moved across the box.
inline Region<_type_lattice_coord> fwCommRecvRegion(const BccDomain* d, int dimension, int direction) {
  Region<_type_lattice_coord> r = fwCommLocalRegion(d, dimension, direction);
  const _type_lattice_coord g = d->dbx_lattice_size_ghost[dimension], s = d->dbx_sub_box_lattice_size[dimension];
  _type_lattice_coord* lo = dimension == 0 ? &r.x_low : (dimension == 1 ? &r.y_low : &r.z_low);
  _type_lattice_coord* hi = dimension == 0 ? &r.x_high : (dimension == 1 ? &r.y_high : &r.z_high);
  *lo = direction == DIR_LOWER ? g + s : 0;
  *hi = direction == DIR_LOWER ? 2 * g + s : g;
  return r;
}

}  // namespace comm
