// fomo_nnlos_part.hip — instantiations of the fused NN + integration kernel k_nn_los for one
// group of lambda chunk sizes CH and one index kind. The build compiles this file once per group
// and kind (-DFOMO_NNLOS_GROUP=0..7, two CH values each, one of them at least 36 for the two-pass
// variant; -DFOMO_NNLOS_RECT=0 for candidate lookup grids, 1 for rectilinear datasets) so that the
// 352 instantiations compile in parallel.
#include <algorithm>

#include "fomo_common.h"

#ifndef FOMO_NNLOS_GROUP
#error "compile with -DFOMO_NNLOS_GROUP=<0..7>"
#endif
#ifndef FOMO_NNLOS_RECT
#error "compile with -DFOMO_NNLOS_RECT=<0|1>"
#endif

namespace fomo_dev {

namespace {

constexpr bool kRect = FOMO_NNLOS_RECT != 0;
constexpr int64_t kDirectBlocks = 512;  // blocks of the direct pass

template <int RS4, int CH, int G, int SL, int NP>
void launch_pair(unsigned blocks, hipStream_t s, const NnLosArgs& a) {
  const dim3 grid(blocks, (unsigned)a.nviews);
  k_nn_los<RS4, CH, G, false, SL, NP, kRect><<<grid, 256, 0, s>>>(a.pts, a.g, a.lslot, a.lcand, a.lk, a.mpts, a.nlo,
                                                                  a.nhi, a.bvh, a.geom, a.nx, a.ny, a.nz, a.rec, a.ri,
                                                                  a.ls, a.nl, a.flags, a.out_vs, a.flag_vs, a.dlist);
  // the direct pass walks the main pass's list of waves with flagged rays
  const unsigned dblocks = (unsigned)std::min<int64_t>((int64_t)blocks * a.nviews, kDirectBlocks);
  k_nn_los<RS4, CH, G, true, SL, NP, kRect><<<dblocks, 256, 0, s>>>(a.pts, a.g, a.lslot, a.lcand, a.lk, a.mpts,
                                                                    a.nlo, a.nhi, a.bvh, a.geom, a.nx, a.ny, a.nz,
                                                                    a.rec, a.ri, a.ls, a.nl, a.flags, a.out_vs,
                                                                    a.flag_vs, a.dlist);
}

}  // namespace

template <int CH, bool RECT>
void nn_los_launch(int variant, unsigned blocks, hipStream_t s, const NnLosArgs& a) {
  static_assert(RECT == kRect, "instantiated in the translation unit of its index kind");
  switch (variant) {
    case kNnLosSingle: launch_pair<2, CH, 4, 1, 1>(blocks, s, a); break;
    case kNnLosShared: launch_pair<2, CH, 4, 4, 1>(blocks, s, a); break;
    case kNnLosLines2: launch_pair<2, CH, 2, 0, 1>(blocks, s, a); break;
    case kNnLosLines4: launch_pair<2, CH, 4, 0, 1>(blocks, s, a); break;
    case kNnLosLines8: launch_pair<4, CH, 8, 0, 1>(blocks, s, a); break;
    case kNnLosSingle2:
      if constexpr (CH >= 36) launch_pair<2, CH, 4, 1, 2>(blocks, s, a);
      break;
    default: break;
  }
}

#define FOMO_NNLOS_INST(C) template void nn_los_launch<C, kRect>(int, unsigned, hipStream_t, const NnLosArgs&);
#if FOMO_NNLOS_GROUP == 0
FOMO_NNLOS_INST(4) FOMO_NNLOS_INST(64)
#elif FOMO_NNLOS_GROUP == 1
FOMO_NNLOS_INST(8) FOMO_NNLOS_INST(60)
#elif FOMO_NNLOS_GROUP == 2
FOMO_NNLOS_INST(12) FOMO_NNLOS_INST(56)
#elif FOMO_NNLOS_GROUP == 3
FOMO_NNLOS_INST(16) FOMO_NNLOS_INST(52)
#elif FOMO_NNLOS_GROUP == 4
FOMO_NNLOS_INST(20) FOMO_NNLOS_INST(48)
#elif FOMO_NNLOS_GROUP == 5
FOMO_NNLOS_INST(24) FOMO_NNLOS_INST(44)
#elif FOMO_NNLOS_GROUP == 6
FOMO_NNLOS_INST(28) FOMO_NNLOS_INST(40)
#elif FOMO_NNLOS_GROUP == 7
FOMO_NNLOS_INST(32) FOMO_NNLOS_INST(36)
#else
#error "FOMO_NNLOS_GROUP must be 0..7"
#endif
#undef FOMO_NNLOS_INST

}  // namespace fomo_dev
