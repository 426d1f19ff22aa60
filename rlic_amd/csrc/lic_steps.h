// Streamline step loops shared by the single-field and the batched kernels.
#pragma once

#include "lic_device.h"

namespace rlic {

// Kernel weights are read through a copy padded with kKernelPad zeros on
// each side, so a group of 8 taps can be loaded without bounds checks.
constexpr int32_t kKernelPad = 8;

// Move-code layout: a direction's steps are cut into chunks of 32 steps
// (4 words of 8 four-bit codes, the first step of a word in its top nibble:
// move_code); chunk c of pixel p (row-major index
// i * nx + j) is the 16-byte vector codes4[(c0 + c) * ny * nx + p], so one
// lane moves 32 steps of codes with one load or store, and the tracing and
// replay kernels may tile the image differently.
constexpr int32_t kChunkSteps = 32;

__host__ __device__ inline int32_t chunks_of(int32_t nsteps) {
  return (nsteps + kChunkSteps - 1) / kChunkSteps;
}

// One streamline direction of one pixel (src/lib.rs:304-362), starting from
// the vector (pu, pv) of the start cell, already loaded.  The position is
// (row, j) with row = i * nx, both in the gather's units.  Returns the
// updated accumulator.  With REC, writes the move code of every step.
//
// The step is written without divergent branches: a lane whose line has
// ended (NaN vector) keeps executing with its position frozen and its
// accumulator untouched, and a lane with a zero vector keeps its position
// (src/lib.rs:242-244); steps run in groups of 8 (one code word, one batch
// of 8 weights) and a wave leaves the loop once all its lines have ended.
// POL and REC: 0 = off, 1 = on, kRuntime = decided by the wave-uniform
// arguments pol_on / rec_on (one kernel body for fields of mixed modes).
//
// The cells form a guarded array (guarded_cells: (ny + 2) x (nx + 2)
// with a ring of guard cells around the image) and `row` is the position.
// A guard cell holds the values of the cell that a step onto it ends on by
// the boundary rule (src/lib.rs:83-95: the edge cell itself for a closed
// side, the opposite edge cell for a periodic one) and, in its pad slot,
// the jump in cells to that cell (0 in image cells): a step is one add of
// +-1 cell or +-1 row, the gather, and one shifted add of the jump.
//
// LANE_DIR: the direction is the lane's own (forward where fwd_lane), one
// body for both (the paired trace, lic_trace_pair); FORWARD then names the
// weights the accumulator takes (the backward lanes' sum is not used).
// `sink(s, x, live)` sees each step's texture value.
constexpr int kRuntime = 2;

struct NoSink {
  template <typename T>
  __device__ __forceinline__ void operator()(int32_t, T, bool) const {}
};

template <typename T, bool FMA, bool BRANCHLESS, int POL, bool FORWARD, int REC, bool BUF,
          bool LANE_DIR = false, typename Sink = NoSink>
__device__ __forceinline__ T trace_direction(T acc, T pu, T pv, uint32_t row,
                                             const CellGather<T, BUF>& cell,
                                             const T* __restrict__ kernel, int32_t kmid,
                                             int32_t nsteps, const StepGeom& sg,
                                             uint4* __restrict__ codes, int64_t np, int64_t p,
                                             int32_t c0, bool pol_on = false,
                                             bool rec_on = false, bool fwd_lane = true,
                                             Sink sink = Sink{}) {
  const bool pol = POL == kRuntime ? pol_on : POL != 0;
  const bool rec = REC == kRuntime ? rec_on : REC != 0;
  const uint32_t lane_sign = fwd_lane ? 0u : 0x80000000u;  // LANE_DIR: negate on backward lanes
  T fx = T(0.5), fy = T(0.5);      // src/lib.rs:315-318
  T last_u = T(0), last_v = T(0);  // src/lib.rs:320-323
  bool alive = true;
  const uint32_t unit = cell.unit;
  const uint32_t row_up = lane_copy(sg.grow);  // row step, held in VGPRs across the loop
  const uint32_t row_down = lane_copy(0u - sg.grow);
  const int32_t nchunks = chunks_of(nsteps);
  for (int32_t ch = 0; ch < nchunks; ++ch) {
    uint32_t words[4] = {0u, 0u, 0u, 0u};
    bool any_alive = true;
#pragma unroll 1
    for (int g = 0; g < 4; ++g) {
      const int32_t s0 = ch * kChunkSteps + 8 * g;
      if (s0 >= nsteps || !any_alive) break;  // uniform
      T kw[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        kw[q] = kernel[FORWARD ? kmid + 1 + s0 + q : kmid - 1 - s0 - q];
      uint32_t word = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (s0 + q >= nsteps) break;  // uniform: last group only
        // a NaN component ends the line (src/lib.rs:336-338)
        const bool ordered = !__builtin_isunordered(pu, pv);
        const bool live = alive && ordered;
        if (POL != 0 && pol) {  // src/lib.rs:339-347, unfused dot product
          const T d1 = pu * last_u;
          const T d2 = pv * last_v;
          const bool flip = d1 + d2 < T(0);
          pu = flip ? -pu : pu;
          pv = flip ? -pv : pv;
          last_u = pu;
          last_v = pv;
        }
        // src/lib.rs:348-351
        const T mu = LANE_DIR ? xor_sign(pu, lane_sign) : (FORWARD ? pu : -pu);
        const T mv = LANE_DIR ? xor_sign(pv, lane_sign) : (FORWARD ? pv : -pv);
        // sign tests of mu, mv read pu, pv (negation flips the sign bit), so
        // that a backward step does not materialize -pu, -pv for them: the
        // divisions and products take the negation as a source modifier
        const bool neg_x = LANE_DIR  ? __builtin_signbit(mu)
                           : FORWARD ? __builtin_signbit(pu)
                                     : !__builtin_signbit(pu);
        const bool neg_y = LANE_DIR  ? __builtin_signbit(mv)
                           : FORWARD ? __builtin_signbit(pv)
                                     : !__builtin_signbit(pv);
        const T tx = time_to_next_pixel<T, FMA, BRANCHLESS>(mu, fx, neg_x);
        const T ty = time_to_next_pixel<T, FMA, BRANCHLESS>(mv, fy, neg_y);
        // x edge first when tx < ty; ties and NaN go to y (src/lib.rs:249-269)
        const bool xs = tx < ty;
        // no move when both components are +-0 (src/lib.rs:242-244)
        const bool moves = live && (mu != T(0) || mv != T(0));
        // update_state (src/lib.rs:208-233) for either axis, then keep one
        // pos_x only matters when the x edge comes first (tx < ty), which
        // needs a finite tx and so a nonzero, non-NaN mu: there it equals
        // mu >= 0 (src/lib.rs:217), and the sign test is the one the step
        // time already made.  pos_y is used for y steps, which also happen
        // with mv = -0 (tx and ty both infinite), so it stays mv >= 0.
        const bool pos_x = !neg_x;
        const bool pos_y = LANE_DIR ? mv >= T(0) : FORWARD ? pv >= T(0) : pv <= T(0);  // mv >= 0
        const T fy_after_x = macc<FMA>(tx, mv, fy);
        const T fx_after_y = macc<FMA>(ty, mu, fx);
        const T fx_new = xs ? (pos_x ? T(0) : T(1)) : fx_after_y;
        const T fy_new = xs ? fy_after_x : (pos_y ? T(0) : T(1));
        // The fractions need no "moves" gate: a lane that does not move
        // never moves again (an ended line stays on its NaN cell, a zero
        // vector stays on its cell, whose vector stays zero, polarization
        // flips included), so its fractions are never read again: they only
        // feed step times, whose results are masked out of its code bits and
        // its position update.
        fx = fx_new;
        fy = fy_new;
        const uint32_t step = xs ? (pos_x ? unit : 0u - unit) : (pos_y ? row_up : row_down);
        const uint32_t at = moves ? row + step : row;
        const CellVal<T> cv = cell(at);
        const uint32_t jump = pad_bits(cv.pad);
        const bool edge = jump != 0u;  // the step left the image
        row = at + jump * unit;
        const T next_acc = macc<FMA>(kw[q], cv.x, acc);  // src/lib.rs:353-360
        acc = live ? next_acc : acc;
        sink(s0 + q, cv.x, live);
        pu = cv.u;
        pv = cv.v;
        if (REC != 0 && rec) {
          // The code's four bits (moved, edge, y axis, negative direction;
          // the three low bits alone for the stop code) are combined from
          // the step's comparisons as wave lane masks on the scalar unit,
          // then shifted into the word one at a time with an add-with-carry
          // (word = 2 * word + bit), so the group's first step ends in the
          // top nibble (move_code).  A line that has ended stays on its NaN
          // cell, so it records the stop code again at every later step;
          // readers stop at the first one.
          const uint64_t m_stop = ~__builtin_amdgcn_ballot_w64(ordered);
          const uint64_t m_nz = __builtin_amdgcn_ballot_w64(mu != T(0)) |
                                __builtin_amdgcn_ballot_w64(mv != T(0));
          const uint64_t m_x = __builtin_amdgcn_ballot_w64(xs);
          const uint64_t m_edge = __builtin_amdgcn_ballot_w64(edge);
          const uint64_t m_neg = ~((m_x & __builtin_amdgcn_ballot_w64(pos_x)) |
                                   (~m_x & __builtin_amdgcn_ballot_w64(pos_y)));
          word = carry_in_bit(word, m_nz & ~m_stop);
          word = carry_in_bit(word, (m_nz & m_edge) | m_stop);
          word = carry_in_bit(word, (m_nz & ~m_x) | m_stop);
          word = carry_in_bit(word, (m_nz & m_neg) | m_stop);
        }
        alive = live;
      }
      // a last group of fewer than 8 steps: its first step to the top
      // nibble, "stay" codes after its last
      if (REC != 0 && rec && nsteps - s0 < 8) word <<= 4 * (8 - (nsteps - s0));
      words[g] = word;
      any_alive = __builtin_amdgcn_ballot_w64(alive) != 0;
    }
    if (REC != 0 && rec) codes[(c0 + ch) * np + p] = make_uint4(words[0], words[1], words[2], words[3]);
    if (!any_alive) break;  // every line of the wave ended
  }
  return acc;
}

// Replay of one direction from its move codes.  Chunks of 32 steps: the
// chunk after the current one is loaded before the current one is used, and
// within a chunk each group of 8 steps decodes its 8 positions through the
// delta table, issues the 8 texture gathers together and accumulates them
// in step order.  Codes are valid up to the step where the line ended (the
// tracing pass wrote the stop code there), so a lane stops reading codes
// once its line has ended and its later codes are taken as "stay".
template <typename T, bool FMA, bool FORWARD, bool BUF>
__device__ __forceinline__ T replay_direction(T acc, uint32_t pos,
                                              const ValueGather<T, BUF>& x,
                                              const T* __restrict__ kernel, int32_t kmid,
                                              int32_t nsteps, const uint32_t* delta,
                                              const uint4* __restrict__ codes, int64_t np,
                                              int64_t p, int32_t c0) {
  const int32_t nchunks = chunks_of(nsteps);
  bool alive = true;
  uint4 next = nchunks > 0 ? codes[c0 * np + p] : make_uint4(0u, 0u, 0u, 0u);
  for (int32_t ch = 0; ch < nchunks; ++ch) {
    const uint4 cur = next;
    if (ch + 1 < nchunks && alive) next = codes[(c0 + ch + 1) * np + p];
    const uint32_t words[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int32_t s0 = ch * kChunkSteps + 8 * g;
      if (s0 >= nsteps) break;  // uniform
      const uint32_t word = alive ? words[g] : 0u;  // 0 = stay
      T kw[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        kw[q] = kernel[FORWARD ? kmid + 1 + s0 + q : kmid - 1 - s0 - q];
      uint32_t at[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        pos += delta[move_code(word, q)];
        at[q] = pos;
      }
      T xv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) xv[q] = x(at[q]);
      const bool full = alive && !group_has_stop(word) && nsteps - s0 >= 8;
      if (__builtin_amdgcn_ballot_w64(!full) == 0) {
        // every lane accumulates all 8 steps (wave-uniform): no masks
#pragma unroll
        for (int q = 0; q < 8; ++q) acc = macc<FMA>(kw[q], xv[q], acc);
      } else {
        // no break for the last group: a break lets the compiler sink the
        // gathers past it and wait for each one in turn; steps past the end
        // read "stay" codes (position unchanged, weight from the padding)
        const int32_t valid = alive ? steps_before_stop(word) : 0;
        const int32_t lim = min(valid, nsteps - s0);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const T next_acc = macc<FMA>(kw[q], xv[q], acc);
          acc = q < lim ? next_acc : acc;
        }
        alive = valid == 8;
      }
    }
    if (__builtin_amdgcn_ballot_w64(alive) == 0) break;
  }
  return acc;
}

}  // namespace rlic
