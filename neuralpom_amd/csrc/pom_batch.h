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
  std::vector<pom::Member> h_mem;   // host copy of the member table
  pom::Tuning tune;                 // launch shapes (pom_batch_tune)
};
