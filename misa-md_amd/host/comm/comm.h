// libcomm's packer interface and neighbour exchange (comm::Packer,
// comm::neiSendReceive), for one rank that is its own neighbour in every
// direction.  The reference's packers implement Packer<T>
// (lat_particle_packer.h:18-75, rho_packer, force_packer, df_embed_packer,
// inter_particle_packer, inter_border_packer) and call neiSendReceive<T>
// (atom.cpp:114,131,145; atom_list.cpp:45,53; inter_atom_list.cpp:21,49);
// with REVERSE the dimensions run z, y, x (the reverse halos of rho and
// force, atom.cpp:114,145).
//
// Per dimension, both directions are packed before either is unpacked, as
// a send to both neighbours precedes the receives; at one rank the data
// sent in (dimension, direction) is what arrives for (dimension, direction).
#pragma once
#include <vector>

#include "domain.h"

namespace comm {

template <typename T>
class Packer {
 public:
  virtual ~Packer() = default;
  virtual unsigned long sendLength(const int dimension, const int direction) = 0;
  virtual void onSend(T buffer[], const unsigned long send_len, const int dimension, const int direction) = 0;
  virtual void onReceive(T buffer[], const unsigned long receive_len, const int dimension, const int direction) = 0;
};

template <typename T, bool REVERSE = false>
void neiSendReceive(Packer<T>* packer, const mpi_process& process, int data_type,
                    const int rank_id_neighbours[3][2]) {
  (void)process;
  (void)data_type;
  (void)rank_id_neighbours;
  std::vector<T> buf[2];
  for (int k = 0; k < 3; k++) {
    const int d = REVERSE ? 2 - k : k;
    for (int dir = DIR_LOWER; dir <= DIR_HIGHER; dir++) {
      const unsigned long n = packer->sendLength(d, dir);
      buf[dir].resize(n);
      packer->onSend(buf[dir].data(), n, d, dir);
    }
    for (int dir = DIR_LOWER; dir <= DIR_HIGHER; dir++)
      packer->onReceive(buf[dir].data(), static_cast<unsigned long>(buf[dir].size()), d, dir);
  }
}

}  // namespace comm
