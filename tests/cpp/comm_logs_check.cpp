// Exercises the one-rank libcomm and kiwi::logs mirrors from C++, as the
// reference's packers and simulation call them; prints one line per check
// for tests/test_host_comm.py.
#include <cstdio>
#include <string>
#include <vector>

#include "comm/comm.h"
#include "comm/domain.h"
#include "logs/logs.h"

struct Probe : comm::Packer<int> {
  std::vector<std::string> calls;
  unsigned long sendLength(const int dimension, const int direction) override {
    calls.push_back("len" + std::to_string(dimension) + std::to_string(direction));
    return 2 + dimension;
  }
  void onSend(int buffer[], const unsigned long n, const int dimension, const int direction) override {
    for (unsigned long i = 0; i < n; i++) buffer[i] = 100 * dimension + 10 * direction + static_cast<int>(i);
    calls.push_back("send" + std::to_string(dimension) + std::to_string(direction));
  }
  void onReceive(int buffer[], const unsigned long n, const int dimension, const int direction) override {
    bool ok = n == static_cast<unsigned long>(2 + dimension);
    for (unsigned long i = 0; i < n; i++) ok = ok && buffer[i] == 100 * dimension + 10 * direction + static_cast<int>(i);
    calls.push_back(std::string(ok ? "recv" : "BAD") + std::to_string(dimension) + std::to_string(direction));
  }
};

int main() {
  const int64_t ps[3] = {10, 12, 14};
  comm::BccDomain* d = comm::BccDomain::Builder()
                           .setComm(comm::mpi_process{})
                           .setPhaseSpace(ps)
                           .setCutoffRadius(5.6 / 2.855)
                           .setLatticeConst(2.855)
                           .setGhostSize(3)
                           .build();
  std::printf("domain %lld %lld %lld\n", static_cast<long long>(d->dbx_ghost_extended_lattice_size[0]),
              static_cast<long long>(d->ghost_extended_lattice_size[1]), static_cast<long long>(d->cut_lattice));
  comm::mpi_process two;
  two.all_ranks = 2;
  std::printf("multi-rank %s\n", comm::BccDomain::Builder().setComm(two).setPhaseSpace(ps).setCutoffRadius(2).setLatticeConst(2).build() ? "built" : "refused");
  for (int rev = 0; rev < 2; rev++) {
    Probe p;
    if (rev)
      comm::neiSendReceive<int, true>(&p, comm::mpi_process{}, 0, d->rank_id_neighbours);
    else
      comm::neiSendReceive<int>(&p, comm::mpi_process{}, 0, d->rank_id_neighbours);
    std::printf("%s", rev ? "reverse" : "forward");
    for (auto& c : p.calls) std::printf(" %s", c.c_str());
    std::printf("\n");
  }
  kiwi::logs::setSink(stdout);
  kiwi::logs::i("sim", "step {} of {}: {}\n", 3, 10, 2.5);
  kiwi::logs::e(0, "dump", "file {} failed\n", std::string("x.out"));
  kiwi::logs::w(1, "quiet", "rank one does not print here\n");
  delete d;
  return 0;
}
