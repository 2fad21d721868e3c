// kz_options.hip — the step-kernel variants of the reference's non-default
// options, in a code object of their own so that the default kernels of
// kz_step.hip are compiled and laid out as if the options did not exist:
//   b^2/rho floor in the ZAMO frame   invert.zig:649-684 (u2p pass, face flux)
//   REDUCEORDERATBH / AFTERFIXUP      sim.zig:1316-1337 (face flux)
//   synchrotron bridge                opacities.zig:208-297 (wavespeeds, implicit
//                                     ladder)
// kz_step.hip calls the launchers below when the parameter block asks for
// an option.
#include "kz_sim_kernels.h"

__global__ void __launch_bounds__(128) KZ_WPE(KZ_WPE_WS) k_wavespeeds_bridge(GridD g, Dev d, kz_physics ph) {
  wavespeeds_pass<true>(g, d, ph);
}

void opt_launch_wavespeeds(const GridD &g, const Dev &d, const kz_physics &ph, unsigned nb, hipStream_t st) {
  k_wavespeeds_bridge<<<nb, 128, 0, st>>>(g, d, ph);
}

__global__ void __launch_bounds__(128) KZ_WPE(KZ_WPE_U2P) k_u2p_zamo(GridD g, Dev d, kz_physics ph) {
  u2p_pass<true>(g, d, ph);
}

void opt_launch_u2p_zamo(const GridD &g, const Dev &d, const kz_physics &ph, unsigned nb, hipStream_t st) {
  k_u2p_zamo<<<nb, 128, 0, st>>>(g, d, ph);
}

// the face flux with the options read from the parameter block (order per
// cell, floor frame)
void opt_launch_face_flux(const GridD &g, const Dev &d, const kz_physics &ph, const FaceTiles &ft, unsigned nb,
                          hipStream_t st) {
  if (ph.radviscosity) k_face_flux<true, true><<<nb, 128, 0, st>>>(g, d, ph, ft);
  else k_face_flux<false, true><<<nb, 128, 0, st>>>(g, d, ph, ft);
}

template <class S>
void opt_launch_implicit_bridge(const S &cells, const ImpWork &w, const kz_physics &ph, long n, bool persist,
                                unsigned persist_blocks, hipStream_t st) {
  launch_implicit_br<S, true>(cells, w, ph, n, persist, persist_blocks, st);
}
template void opt_launch_implicit_bridge<GridCells>(const GridCells &, const ImpWork &, const kz_physics &, long, bool,
                                                    unsigned, hipStream_t);
template void opt_launch_implicit_bridge<MultiCells>(const MultiCells &, const ImpWork &, const kz_physics &, long,
                                                     bool, unsigned, hipStream_t);
template void opt_launch_implicit_bridge<BatchCells>(const BatchCells &, const ImpWork &, const kz_physics &, long,
                                                     bool, unsigned, hipStream_t);
