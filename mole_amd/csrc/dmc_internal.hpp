// Internal structures of diffusion Monte Carlo on the device: the per-system
// state the step kernels share, and the launch arguments of the reductions,
// branching and bookkeeping around the Slater-Jastrow DMC steps
// (src/dmc/src/dmc.rs:75-202, src/dmc/src/branching.rs:7-91).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mole {

// Philox streams of DMC (key as for the walkers): the starting
// configuration is stream (walker 0xFFFFFFFF, kDmcInitEpoch); the branching
// of step s is stream (walker s, kDmcBranchEpoch).
constexpr uint32_t kDmcInitEpoch = 0x444D4300u;
constexpr uint32_t kDmcBranchEpoch = 0x444D4301u;

enum DmcBrancher : int { kDmcBranchSR = 0, kDmcBranchSimple = 1 };

// Device-resident state of one system's DMC run: written and read only by
// kernels, so a captured step replays without host input.
struct DmcState {
  double eref;                              // DmcRunner::reference_energy
  double sum_we, sum_w, sum_wnew, max_w;    // the last step's reductions
  double gw, norm;                          // SRBrancher: mean weight, W / max weight
  uint64_t steps;                           // steps since mole_dmc_init (branching stream)
  uint64_t total;                           // SRBrancher: total integer weight
  uint32_t it, j, block_nr, n_out;          // step of the call, step in block, block, blocks out
  uint32_t M, S;                            // SimpleBranching: list length, survivors
  uint32_t error;                           // 1: total integer weight 0 (the reference panics)
  uint32_t branch;                          // step whose branching is in progress (stream key)
};

struct DmcArgs {
  DmcState *st;
  double *dmc;  // [4][W] (see SJSysDev::dmc)
  int W;
  int brancher;
  uint64_t seed;
  uint32_t sys;
  uint32_t block_size, num_eq_blocks;
  uint32_t cap_steps;
  double *step_energy;  // [cap_steps]
  double *block_buf;    // [block_size]
  double *energies;     // [cap_blocks]
  double *vars;         // [cap_blocks]
  // scratch
  double *part;      // [4][n_parts]
  uint64_t *cnt_a;   // [3W] counts, then inclusive sums
  uint64_t *cnt_b;   // [W]
  uint64_t *bsum_a;  // scan block sums
  uint64_t *bsum_b;
  uint32_t *list;     // [3W]
  uint32_t *removed;  // [2W]
  uint32_t *src;      // [W]
  double *wnew;       // [W]
  // walker state (positions [pos_rows][W] f64, inverses [b_rows][W] of
  // b_elem bytes) and the copies gathered from it after branching, which the
  // next step reads (el_s the cached local energies, wnew the weights)
  double *pos;
  int pos_rows;
  void *B;
  int b_rows, b_elem;
  double *pos_s;
  void *B_s;
  double *el_s;
  // the sweep counters of the Slater-Jastrow launch state, which every DMC
  // step advances by one (set for one system per step, null for the rest)
  uint64_t *sweeps_done;
  uint32_t *trace_sweeps;
};

// scratch sizes for W walkers
size_t dmc_parts(int W);
size_t dmc_scan_blocks(size_t n);

// all walkers at one N(0,1) configuration from the init stream
hipError_t launch_dmc_init_configs(double *pos, int N, int W, uint64_t seed, uint32_t sys, hipStream_t stream);
// one system's work after a DMC step's sweep launch: reductions, the step
// energy, branching, the gather of walker state, the block bookkeeping
hipError_t launch_dmc_after_step(const DmcArgs &a, hipStream_t stream);
// The walkers between DMC calls are in the walker state (positions,
// inverses, weights and cached local energies in dmc rows 0-1); within a
// call, between steps, in the gathered copies that each step starts from.
// to_main copies the gathered copies into the walker state (a call's end),
// otherwise the walker state into the copies (its start).
hipError_t launch_dmc_sync_state(const DmcArgs &a, bool to_main, hipStream_t stream);
// start of a diffuse call: block counters to zero
hipError_t launch_dmc_call_start(DmcState *st, hipStream_t stream);
// *st = v by a kernel (capturable)
hipError_t launch_dmc_state_set(DmcState *st, const DmcState &v, hipStream_t stream);

}  // namespace mole
