// adjoint.h — device buffers and launchers of the reverse-mode SPH step.
#pragma once

#include "fluid_internal.h"

namespace fluid {

// Gradient slots per particle in AdjBufs::acc (doubles, stride kAccStride):
// viscosity, surface tension, gravity x, y, z.
constexpr int kGradParams = 5;
constexpr int kAccStride = 8;

// Adjoint ("lambda") buffers, all indexed by the sorted slot of the step
// being reversed unless stated otherwise.
struct AdjBufs {
  const float4 *lin_x, *lin_v;  // adjoint of the step's output position / velocity
  float4 *lout_x, *lout_v;      // adjoint of its input state, previous step's slot order
  float4 *lx;                   // sorted position x
  float4 *lw;                   // post-XSPH velocity w
  float4 *la;                   // non-pressure acceleration a (force output)
  float4 *lu;                   // sorted (pre-XSPH) velocity u
  float4 *ln;                   // Akinci normal
  float2 *lrho;                 // (density, near density); (raw sum S, near) after
                                // the XSPH kernel folds in the wall factor
  float4 *lap[2];               // (pressure acceleration before the clamp, total
                                // predicted-density adjoint), ping-pong
  float *lP;                    // pressure of the iteration being reversed
  double *acc;                  // parameter gradients per slot (kAccStride)
  // lists built within one reversed step, in the wave blocks of
  // lane_block<L>: the solve skin list of the first reversed PCISPH
  // iteration (per lane, scnt per lane), and the in-range pairs of the force
  // adjoint (ballot-balanced like the density pass's list, pcnt per particle)
  uint32_t *slist, *scnt;       // kSolveListCapacity per particle
  uint32_t *plist, *pcnt;       // kListCapacity per particle
};

// What the forward step kept for its reverse (all in the step's slot order).
struct TapeStep {
  const float4 *x;     // sorted positions
  const float4 *u;     // sorted velocities (pre-XSPH)
  const float4 *w;     // post-XSPH velocities
  const float4 *rec;   // density record: (x, rho), (u, rho_near), (normal, 0)
  const float4 *vK;    // PCISPH's final predicted velocity
  const uint32_t *counts, *starts;  // the step's grid
  const uint32_t *gmap;             // slot -> previous step's slot
};

// One PCISPH iteration's saved inputs.
struct TapeIter {
  const float4 *xq;  // predicted position, pressure / max(rho*, 1)^2
  const float2 *pr;  // pressure, predicted density of the previous iteration
  const float4 *ap;  // previous iteration's unclamped pressure acceleration,
                     // null when that iteration is the first (it is zero)
};

// Canonical (original-index, 3 floats per particle) <-> slot-order float4.
hipError_t launch_adj_load(const uint32_t *sid, const float *dl_x3, const float *dl_v3,
                           float4 *lx, float4 *lv, uint32_t n, hipStream_t st);
hipError_t launch_adj_store(const uint32_t *sid, const float4 *lx, const float4 *lv,
                            float *out_x3, float *out_v3, uint32_t n, hipStream_t st);

// The carried adjoint state (lambda of position and velocity per slot, and
// the parameter accumulators) as mantissa x 2^E(scene): max_cur receives each
// scene's largest |lambda| (as float bits), e_step the exponent k with which
// the state is then scaled by 2^-k, e_total the running sum of the k; max_next
// (the other parity) is cleared for the next call.  scene_of null: one scene.
hipError_t launch_adj_rescale_step(const uint16_t *scene_of, float4 *lx, float4 *lv,
                                   double *acc, uint32_t n, uint32_t scenes, uint32_t *max_cur,
                                   uint32_t *max_next, int32_t *e_total, int32_t *e_step,
                                   hipStream_t st);
// lambda (and acc when given) *= 2^(mul * e[scene]).
hipError_t launch_adj_apply_scale(const uint16_t *scene_of, float4 *lx, float4 *lv, double *acc,
                                  uint32_t n, const int32_t *e, int mul, hipStream_t st);

// out[scene * kGradParams + q] = sum of acc[slot * kAccStride + q] over the
// scene's slots [slot_first[scene], slot_first[scene + 1]).
hipError_t launch_adj_reduce(const double *acc, const uint32_t *slot_first, uint32_t scenes,
                             double *out, hipStream_t st);

// The full reverse of one step.  it[k], k = 0..iters-1, are the inputs of
// PCISPH iteration k; last = the final iteration's output (its pr.y is the
// final predicted density, its ap the final unclamped acceleration).
hipError_t launch_adjoint_step(const Uni &u, AdjBufs a, const TapeStep &t, const TapeIter *it,
                               const TapeIter &last, uint32_t iters, uint32_t n, int lanes,
                               hipStream_t st);

}  // namespace fluid
