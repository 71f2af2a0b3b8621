// C-ABI over the one-rank mirror of libcomm (misa-md_amd/host/comm): the
// BccDomain fields of a geometry, its forwarding regions, and the lattice
// halo of exchangeAtomFirst (atom_list.cpp:25-49) with the LatPackerFirst
// packing (lat_particle_packer.cpp:22-160) on a host AtomElement array.
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/misa_eam.h"
#include "../host/comm/comm.h"
#include "../host/comm/domain.h"
#include "host_setup.hpp"

namespace {

std::unique_ptr<comm::BccDomain> domainOf(const misa_geom* g) {
  if (!g) return nullptr;
  // the ghost size simulation::createDomain sets (simulation.cpp:40-46)
  const int ghost = static_cast<int>(std::ceil(g->cutoff_radius_factor)) + 1;
  return std::unique_ptr<comm::BccDomain>(comm::BccDomain::Builder()
                                              .setComm(comm::mpi_process{})
                                              .setPhaseSpace(g->phase_space)
                                              .setCutoffRadius(g->cutoff_radius_factor)
                                              .setLatticeConst(g->lattice_const)
                                              .setGhostSize(ghost)
                                              .build());
}

template <typename T>
void put(const comm::Region<T>& r, T out[6]) {
  out[0] = r.x_low;
  out[1] = r.y_low;
  out[2] = r.z_low;
  out[3] = r.x_high;
  out[4] = r.y_high;
  out[5] = r.z_high;
}

struct LatParticleData {
  int32_t type;
  double r[3];
};

// LatPackerFirst: the send lists are the forwarding regions, the receive
// lists the ghost layers they land in, both walked z, y, x; a particle
// crossing the periodic boundary is shifted by the box length (setOffset).
class LatPacker : public comm::Packer<LatParticleData> {
 public:
  LatPacker(const comm::BccDomain& d, misa::AtomElement* atoms) : d_(d), atoms_(atoms), send_(6), recv_(6) {
    for (int dim = 0; dim < 3; dim++)
      for (int dir = comm::DIR_LOWER; dir <= comm::DIR_HIGHER; dir++) {
        walk(comm::fwCommLocalRegion(&d, dim, dir), send_[2 * dim + dir]);
        walk(comm::fwCommRecvRegion(&d, dim, dir), recv_[2 * dim + dir]);
      }
  }
  unsigned long sendLength(const int dimension, const int direction) override {
    return static_cast<unsigned long>(send_[2 * dimension + direction].size());
  }
  void onSend(LatParticleData buffer[], const unsigned long n, const int dimension, const int direction) override {
    double off[3] = {0.0, 0.0, 0.0};
    if (d_.grid_coord[dimension] == 0 && direction == comm::DIR_LOWER) off[dimension] = d_.meas_global_length[dimension];
    if (d_.grid_coord[dimension] == d_.grid_size[dimension] - 1 && direction == comm::DIR_HIGHER)
      off[dimension] = -d_.meas_global_length[dimension];
    const std::vector<int64_t>& sl = send_[2 * dimension + direction];
    for (unsigned long i = 0; i < n; i++) {
      const misa::AtomElement& a = atoms_[sl[i]];
      buffer[i].type = a.type;
      for (int k = 0; k < 3; k++) buffer[i].r[k] = a.x[k] + off[k];
    }
  }
  void onReceive(LatParticleData buffer[], const unsigned long n, const int dimension, const int direction) override {
    const std::vector<int64_t>& rl = recv_[2 * dimension + direction];
    for (unsigned long i = 0; i < n && i < rl.size(); i++) {
      misa::AtomElement& a = atoms_[rl[i]];
      a.type = buffer[i].type;
      for (int k = 0; k < 3; k++) a.x[k] = buffer[i].r[k];
    }
  }

 private:
  void walk(const comm::Region<comm::_type_lattice_coord>& r, std::vector<int64_t>& out) const {
    const int64_t sx = d_.dbx_ghost_extended_lattice_size[0], sy = d_.dbx_ghost_extended_lattice_size[1];
    for (int64_t z = r.z_low; z < r.z_high; z++)
      for (int64_t y = r.y_low; y < r.y_high; y++)
        for (int64_t x = r.x_low; x < r.x_high; x++) out.push_back((z * sy + y) * sx + x);
  }
  const comm::BccDomain& d_;
  misa::AtomElement* atoms_;
  std::vector<std::vector<int64_t>> send_, recv_;
};

}  // namespace

extern "C" {

int misa_host_domain(const misa_geom* g, misa_domain* out) {
  if (!out) return MISA_E_ARG;
  auto d = domainOf(g);
  if (!d) return MISA_E_ARG;
  std::memset(out, 0, sizeof(*out));
  for (int k = 0; k < 3; k++) {
    out->phase_space[k] = d->phase_space[k];
    out->grid_size[k] = d->grid_size[k];
    out->grid_coord[k] = d->grid_coord[k];
    out->rank_id_neighbours[2 * k] = d->rank_id_neighbours[k][0];
    out->rank_id_neighbours[2 * k + 1] = d->rank_id_neighbours[k][1];
    out->meas_global_length[k] = d->meas_global_length[k];
    out->sub_box_lattice_size[k] = d->sub_box_lattice_size[k];
    out->lattice_size_ghost[k] = d->lattice_size_ghost[k];
    out->ghost_extended_lattice_size[k] = d->ghost_extended_lattice_size[k];
    out->dbx_sub_box_lattice_size[k] = d->dbx_sub_box_lattice_size[k];
    out->dbx_lattice_size_ghost[k] = d->dbx_lattice_size_ghost[k];
    out->dbx_ghost_extended_lattice_size[k] = d->dbx_ghost_extended_lattice_size[k];
  }
  out->lattice_const = d->lattice_const;
  out->cutoff_radius_factor = d->cutoff_radius_factor;
  out->cut_lattice = d->cut_lattice;
  put(d->meas_global_region, out->meas_global_region);
  put(d->meas_sub_box_region, out->meas_sub_box_region);
  put(d->sub_box_lattice_region, out->sub_box_lattice_region);
  put(d->ghost_ext_lattice_region, out->ghost_ext_lattice_region);
  put(d->dbx_sub_box_lattice_region, out->dbx_sub_box_lattice_region);
  put(d->dbx_ghost_ext_lattice_region, out->dbx_ghost_ext_lattice_region);
  return MISA_OK;
}

int misa_host_fw_region(const misa_geom* g, int dimension, int direction, int64_t region[6]) {
  if (!region || dimension < 0 || dimension > 2 || direction < 0 || direction > 1) return MISA_E_ARG;
  auto d = domainOf(g);
  if (!d) return MISA_E_ARG;
  put(comm::fwCommLocalRegion(d.get(), dimension, direction), region);
  return MISA_OK;
}

int misa_host_exchange_atoms(const misa_geom* g, void* atoms) {
  if (!atoms) return MISA_E_ARG;
  auto d = domainOf(g);
  if (!d) return MISA_E_ARG;
  LatPacker packer(*d, static_cast<misa::AtomElement*>(atoms));
  comm::neiSendReceive<LatParticleData>(&packer, comm::mpi_process{}, 0, d->rank_id_neighbours);
  return MISA_OK;
}

}  // extern "C"
