// Host-side launchers of the step, init, readout and force kernels (defined in lbm_step.hip).
#pragma once
#include "lbm_kernels.h"

namespace aero {
// Second stream for the per-cell kernel of a split step's "other" cells (boundary cells next to the body):
// it runs beside the bulk kernel (the two touch disjoint slots), forked from and joined back to the main
// stream with the two events.
struct SideLaunch {
  hipStream_t stream;
  hipEvent_t fork, join;
};
// Launches one step; *nblk receives the number of block partials it writes (the count to finalize).
// A split step records `mid`, when given, on the main stream right after its bulk kernel.
hipError_t launch_step(const StepArgs& a, bool fp16, int V, bool from_e, bool readout, hipStream_t s,
                       uint32_t* nblk, hipEvent_t mid = nullptr, const SideLaunch* side = nullptr);
// Partials written by a split (bulk + plain-fluid list + other list) step over N cells with groups of V.
uint32_t split_blocks(uint32_t N, int V, uint32_t nfluid, uint32_t nother);
// Compacted list of the cells outside bulk groups of V that a split step visits (hipMalloc'd; caller frees):
// first the *nfluid plain fluid cells, then the *nother others, each in memory order. *nbulk receives the
// number of cells the bulk kernel updates.
hipError_t build_special_list(const uint8_t* flags, const Geo& g, int V, int32_t** list, uint32_t* nfluid,
                              uint32_t* nother, uint32_t* nbulk, hipStream_t s);
hipError_t launch_finalize(double* partials, uint32_t n, StepState* st, hipStream_t s);
// Macro fields (0, 0, 0, 1) of every solid cell (mac: ux, uy, uz, rho of N cells each).
hipError_t launch_solid_macros(const uint8_t* flags, uint32_t N, float* mac, hipStream_t s);
uint32_t partials_len(uint32_t nblk);
hipError_t launch_init(const StepArgs& a, bool fp16, double u0, uint64_t seed, double amp, hipStream_t s);
hipError_t launch_load(const StepArgs& a, bool fp16, const float* p, hipStream_t s);
hipError_t launch_links(const uint8_t* flags, const int32_t* blist, uint32_t B, const Geo& g, uint32_t* bmask,
                        hipStream_t s);
// rowfluid[row] = 1 where every cell of row y + ny z is plain fluid, else 0 (ny nz entries).
hipError_t launch_row_fluid(const uint8_t* flags, const Geo& g, uint32_t* rowfluid, hipStream_t s);
hipError_t launch_forces(const StepArgs& a, bool fp16, const int32_t* blist, int aa_state, const float* halo_cur,
                         double* part, double* out, hipStream_t s);
uint32_t forces_blocks(uint32_t B);
uint32_t step_blocks(uint32_t N);
// Batched step in per-kind launches: the bulk kernel over `bulk` tiles (one block of sh.block threads per tile), then
// the plain-fluid list kernel over `fluid` tiles; the other list tiles run on the side stream beside them when
// `side` is given. launch_finalize_batch then sums each tunnel's partials.
hipError_t launch_step_batch(const StepArgs* targs, const BatchTile* bulk, uint32_t nbulk, const BatchTile* fluid,
                             uint32_t nfluid, const BatchTile* other, uint32_t nother, double* part, bool fp16,
                             bool from_e, const BatchShape& sh, hipStream_t s, const SideLaunch* side);
// The whole batched step in one persistent launch of sh.grid blocks (a multiple of 8), finalized by its last block.
hipError_t launch_step_batch_one(const BatchOneArgs& p, bool fp16, bool from_e, const BatchShape& sh, hipStream_t s);
hipError_t launch_finalize_batch(const double* part, const uint32_t* tile_begin, StepState* const* states,
                                 uint32_t ntun, hipStream_t s);
}  // namespace aero
