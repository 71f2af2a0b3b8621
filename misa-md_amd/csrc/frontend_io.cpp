// Atom dumps of the reference frontend (frontend/io), for one rank, on the
// host AtomElement layout the C-ABI transfers.
//
//  - Copy mode (AtomDump, atom_dump.cpp:44-101; BufferedFileWriter,
//    buffered_io.cpp:17-49): 72-byte AtomInfoDump records
//    (atom_info_dump.h:14-22) after a 128-byte file header and one 128-byte
//    header per rank whose first 8 bytes hold the rank's record count
//    (writeDumpHeader), in blocks of 1024 records.  With one rank the blocks
//    follow each other, which is the layout tools/dump2text/out_convert.cpp
//    reads (HEADER_SIZE 128, _LOCAL_HEADER_SIZE 128, BLOCK_SIZE 1024 records).
//    The file header belongs to the kiwi storage library, which the reader
//    skips; it is written as zeros here.
//  - Debug mode (OutputDump::dump, output_dump.cpp:40-70): text, "print
//    atoms" then "id x y z" of every occupied sub-box site, "print inter" then
//    the inter atoms, numbers as std::ostream prints them by default (%g).
//
// A frame writes the inter atoms in list order, then the occupied sub-box
// sites in z, y, doubled-x order (atom_dump.cpp:66-84); every record has
// inter_type 0 and the frame's time step, as the reference writes them.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/misa_eam.h"
#include "host_setup.hpp"

namespace {

struct DumpRecord {
  uint64_t id;
  uint64_t step;
  int32_t type;
  int16_t inter_type;
  int16_t pad;
  double x[3];
  double v[3];
};
static_assert(sizeof(DumpRecord) == 72, "AtomInfoDump is 72 bytes");

constexpr long kFileHeader = 128;
constexpr long kRankHeader = 128;

bool boxOf(const misa_geom* g, misa::Box& b) {
  if (!g || g->phase_space[0] < 1 || g->phase_space[1] < 1 || g->phase_space[2] < 1 || !(g->lattice_const > 0) ||
      !(g->cutoff_radius_factor > 0))
    return false;
  b.init(g->phase_space, g->lattice_const, g->cutoff_radius_factor);
  return true;
}

// Inter atoms in list order, then the occupied sub-box sites (z, y, X).
template <typename Fn>
void forFrameAtoms(const misa::Box& b, const misa::AtomElement* atoms, const misa::AtomElement* inter,
                   int64_t n_inter, Fn fn) {
  for (int64_t i = 0; i < n_inter; i++) fn(inter[i]);
  for (int64_t z = b.g; z < b.g + b.n[2]; z++)
    for (int64_t y = b.g; y < b.g + b.n[1]; y++)
      for (int64_t X = 2 * b.g; X < 2 * (b.g + b.n[0]); X++) {
        const misa::AtomElement& a = atoms[b.refIndex(X, y, z)];
        if (a.type >= 0) fn(a);
      }
}

}  // namespace

struct misa_dump {
  FILE* f = nullptr;
  int64_t records = 0;
  std::vector<DumpRecord> buf;  // BufferedFileWriter's buffer of 1024 records
};

extern "C" {

int misa_dump_open(const char* path, misa_dump** out) {
  if (!path || !out) return MISA_E_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return MISA_E_IO;
  const char zeros[kFileHeader + kRankHeader] = {};
  if (std::fwrite(zeros, 1, sizeof(zeros), f) != sizeof(zeros)) {
    std::fclose(f);
    return MISA_E_IO;
  }
  misa_dump* d = new (std::nothrow) misa_dump;
  if (!d) {
    std::fclose(f);
    return MISA_E_ARG;
  }
  d->f = f;
  d->buf.reserve(1024);
  *out = d;
  return MISA_OK;
}

int misa_dump_frame(misa_dump* d, const misa_geom* g, const void* atoms, const void* inter, int64_t n_inter,
                    uint64_t time_step) {
  misa::Box b;
  if (!d || !d->f || !atoms || n_inter < 0 || (n_inter > 0 && !inter) || !boxOf(g, b)) return MISA_E_ARG;
  int rc = MISA_OK;
  auto flush = [&]() {
    if (d->buf.empty()) return;
    if (std::fwrite(d->buf.data(), sizeof(DumpRecord), d->buf.size(), d->f) != d->buf.size()) rc = MISA_E_IO;
    d->records += static_cast<int64_t>(d->buf.size());
    d->buf.clear();
  };
  forFrameAtoms(b, static_cast<const misa::AtomElement*>(atoms), static_cast<const misa::AtomElement*>(inter),
                n_inter, [&](const misa::AtomElement& a) {
                  DumpRecord r;
                  r.id = a.id;
                  r.step = time_step;
                  r.type = a.type;
                  r.inter_type = 0;
                  r.pad = 0;
                  for (int k = 0; k < 3; k++) {
                    r.x[k] = a.x[k];
                    r.v[k] = a.v[k];
                  }
                  d->buf.push_back(r);
                  if (d->buf.size() == 1024) flush();
                });
  flush();
  return rc;
}

int misa_dump_close(misa_dump* d, int64_t* records) {
  if (!d) return MISA_E_ARG;
  int rc = MISA_OK;
  if (d->f) {
    const uint64_t n = static_cast<uint64_t>(d->records);
    if (std::fseek(d->f, kFileHeader, SEEK_SET) != 0 || std::fwrite(&n, sizeof(n), 1, d->f) != 1) rc = MISA_E_IO;
    if (std::fclose(d->f) != 0) rc = MISA_E_IO;
  }
  if (records) *records = d->records;
  delete d;
  return rc;
}

int misa_dump_text(const char* path, const misa_geom* g, const void* atoms, const void* inter, int64_t n_inter) {
  misa::Box b;
  if (!path || !atoms || n_inter < 0 || (n_inter > 0 && !inter) || !boxOf(g, b)) return MISA_E_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return MISA_E_IO;
  const auto* at = static_cast<const misa::AtomElement*>(atoms);
  const auto* in = static_cast<const misa::AtomElement*>(inter);
  std::fprintf(f, "print atoms\n");
  forFrameAtoms(b, at, nullptr, 0, [&](const misa::AtomElement& a) {
    std::fprintf(f, "%llu %g %g %g\n", static_cast<unsigned long long>(a.id), a.x[0], a.x[1], a.x[2]);
  });
  std::fprintf(f, "print inter\n");
  for (int64_t i = 0; i < n_inter; i++)
    std::fprintf(f, "%llu %g %g %g\n", static_cast<unsigned long long>(in[i].id), in[i].x[0], in[i].x[1],
                 in[i].x[2]);
  return std::fclose(f) == 0 ? MISA_OK : MISA_E_IO;
}

}  // extern "C"
