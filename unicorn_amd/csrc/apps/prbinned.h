// Binned (propagation-blocked) layout of the pageRank graph: host-side planning and building,
// and the CPU twin of the GPU unit (same integer arithmetic, same bits). The device side and the
// layout's definition are in unicorn_hip.h (uh_pagerank_binned_args).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace prb {

// Regions of the graph address space, each starting on a 16-byte boundary:
//   deg (N u32) | unit_start (K+1) | bin_start (K+1) | run_ptr (K+1) | run_off (R+1) | run_pos (R)
//   | dst16 (slots u16) | src16 (E u16)
// A state address space is y (N floats, padded to a multiple of 4) | values (slots floats).
struct Layout {
    uint32_t pages = 0, per_subtask = 0, subtasks = 0, units = 0, runs = 0;
    uint64_t links = 0, slots = 0;        // slots: links plus the padding of each bin to 4
    size_t off_deg = 0, off_unit = 0, off_bin = 0, off_rptr = 0, off_roff = 0, off_rpos = 0, off_dst = 0,
           off_src = 0, graph_bytes = 0;
    size_t state_values_off = 0, state_bytes = 0;
    // host copies used to size subscriptions
    std::vector<uint32_t> unit_start;     // K+1
    std::vector<uint32_t> sub_unit;       // S+1: first unit of each subtask
    std::vector<uint32_t> bin_start;      // K+1
    std::vector<uint32_t> run_ptr;        // K+1
    std::vector<uint32_t> run_off;        // R+1
    // value slot of bin b where the links from subtask s's chunks begin: sub_start[b * (S+1) + s]
    // (entry S is the end of the bin's links, before its padding)
    std::vector<uint32_t> sub_start;
};

// Chooses the units (subtask-aligned, at most UH_PRB_UNIT_MAX pages; whole rounds of `cus` units
// while the runs keep about 128 links) and sizes the regions from the out-link CSR (0-based
// targets). Returns false when the graph cannot use the layout (2^32 or more link slots or runs).
bool plan(uint32_t pages, const uint32_t* out_row_ptr, const uint32_t* out_col, uint32_t per_subtask, uint32_t cus,
          Layout& L);

// Fills the graph address space (L.graph_bytes bytes) and the host tables from the out-link CSR
// (0-based targets). Returns 0, or -1 if the CSR does not match the plan.
int build(const uint32_t* out_row_ptr, const uint32_t* out_col, Layout& L, uint8_t* graph);

// floor(v * 2^s) of the accumulator (v >= 0 finite), as the device computes it.
uint64_t to_fixed(float v, int s);

// Exponent of the fixed-point scale for N pages of initial rank x0 (damping <= 1): every row sum
// stays below 2^62.
int scale_exp(uint32_t pages, float x0);

// Global-indexed pointers of one subtask (host memory), as the device launcher takes them.
struct Views {
    const uint32_t *deg, *unit_start, *bin_start, *run_ptr, *run_off, *run_pos;
    const uint16_t *dst16, *src16;
    const float* values_in;
    float* values_out;            // NULL: no next iteration
    float* y;
};

// The GPU unit on the host: kind 0 = first iteration (y = x0), 1 = iterate.
void unit_cpu(const Views& v, uint32_t u, int kind, float damping, float x0, int s, std::vector<uint64_t>& acc,
              std::vector<float>& w);

}  // namespace prb
