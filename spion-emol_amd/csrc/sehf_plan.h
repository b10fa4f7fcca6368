// sehf_plan.h -- sizes the storage planner shares with the scan engine.
#pragma once

#include <cstddef>

namespace sehf {

// What the scan's batch buffers depend on (besides batch width, checkpoint
// spacing and count, fit window and tangent).
struct BatchDims {
    size_t Np = 0, Nf = 0, Nt = 0, ns = 0, n_occ = 0, nstab = 0, n_start = 1, Nr = 0;
    bool streamed = false;    // tables rebuilt per radial step (adds the table scratch and segment B D^{-1})
    bool ckpt_host = false;   // checkpoints in pinned host memory
    size_t sinv_group = 1;    // radial points per Schur-inverse call (kSinvGroup with the SPD inverse)
};

// Radial points whose Schur matrices share one SPD-inverse call (one launch
// sequence over sinv_group x batch matrices): the order-961 inverse at l_cont
// = 30 costs 4.84 ms per 64 matrices alone and 3.16 ms per 64 in a call of
// 256 (its LDS leaves are one workgroup per matrix).  Beyond order 2048 the
// inverse is a few percent of the step and the group's memory is not spent.
inline int sinv_group_for(size_t n_psi, bool spd) { return (spd && n_psi <= 2048) ? 4 : 1; }
struct BatchBytes {
    size_t device = 0, host = 0;
};

// Bytes of one group's batch buffers, counted through the same buffer list
// the scan allocates from (sehf_scan.cpp, batch_arrays).
BatchBytes batch_bytes(const BatchDims& d, int B, int L, int n_ck, int n_fit, bool tan);
// Bytes of the Gauss-Jordan workspace of `batch` matrices of order n.
size_t gj_buffer_bytes(size_t n, size_t batch);

}  // namespace sehf
