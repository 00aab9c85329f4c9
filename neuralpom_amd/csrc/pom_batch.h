// The ensemble handle behind the C ABI's opaque pom_batch (shared by the
// translation units that implement entry points).
#pragma once
#include <vector>

#include "pom_kernels.h"

struct pom_batch {
  int32_t dtype;
  int32_t M;
  pom::Member* d_mem;
  int32_t* d_blk3;
  int32_t* d_blk2;
  int32_t nblk3, nblk2;
  int64_t n2, n3;
  int2* d_band = nullptr;           // edge-band columns (pom::Batch::band)
  int32_t nband[pom::BAND_W + 1] = {};
  std::vector<pom::Member> h_mem;   // host copy of the member table
  pom::Tuning tune;                 // launch shapes (pom_batch_tune)
};

namespace pom {
// the kernels' view of a batch
inline Batch batch_view(const pom_batch* b) {
  Batch B{b->d_mem, b->d_blk3, b->d_blk2, b->M};
  B.band = b->d_band;
  for (int w = 0; w <= BAND_W; ++w) B.nband[w] = b->nband[w];
  return B;
}
}  // namespace pom
