// Launch-shape knobs shared by the launchers (set with awp_set_tuning).
#pragma once

struct AwpTuning {
  int vel_bj;       // j rows per velocity workgroup (64 k-lanes each)
  int str_bj;       // j rows per stress workgroup
  int vel_target_wgs;  // velocity: workgroups to aim for when splitting the i march
  int str_target_wgs;  // stress: same
  int vel_ichunk;   // >0 forces the velocity i-chunk length
  int str_ichunk;   // >0 forces the stress i-chunk length
  int xcd_remap;    // 1: XCD-aware tile order, 0: plain order
  int kalign;       // 1: start k tiles on a 32-float (128 B) line boundary
  int topo_vel_ichunk;  // planes per march of the curvilinear velocity kernel
  int topo_str_ichunk;  // same, curvilinear stress kernel
  int topo_q_ichunk;    // planes per march of the two-pass stress update
  int step_ichunk;      // >0 forces the i-chunk length of the fused step (awp_step)
  int step_target_wgs;  // fused step: workgroups to aim for when splitting the i march
  int vel_jfast;        // 1: j tiles fastest in the block order (velocity)
  int str_jfast;        // same, stress
  int step_jfast;       // same, fused step
  int topo_vel_split;   // 1: the curvilinear velocity update as one launch per component
  int adj_jgroup;       // adjoint kernels' block order: 0 plane by plane, else j tiles per group
  int topo_str_split;   // strain-rate pass: 0 one launch, 1 two (normal, shear), 2 three (normal, xy, xz+yz)
  int kw;               // k-run width of the Cartesian kernels: 0 by the idle-lane rule, 32, 64
  int topo_nseg;        // curvilinear stress kernel: waves per k chunk at most (1: one-wave segments)
  int topo_wsync;       // curvilinear kernels: 1 wave-scope LDS sync where a wave reads only its own
  int topo_rj;          // curvilinear stress kernel: > 0 forces the j rows per workgroup
};

AwpTuning& awp_tuning();

// Number of i-chunks and chunk length for a march over ni planes when the
// (k, j) tiling gives nkj workgroups per chunk.
static inline void awp_split_i(int ni, int nkj, int forced, int target, int* nic, int* ichunk) {
  int c;
  if (forced > 0) {
    c = forced;
  } else {
    int want = (target + nkj - 1) / nkj;
    if (want < 1) want = 1;
    if (want > ni) want = ni;
    c = (ni + want - 1) / want;
  }
  if (c < 1) c = 1;
  *ichunk = c;
  *nic = (ni + c - 1) / c;
}
