// Batches of frames of uneven sizes: one launch per stage over every frame of
// the batch, captured once as a hipGraph and replayed.
//
// Each frame of a batch has its own buffers (an Encoder bound to its slice
// of the batch's arena) and a FrameJob in device memory: its geometry,
// parameters, buffers, output, and for every stage the frame's share of
// that stage's batched grid (its first workgroup and its own grid shape).
// A batched kernel's workgroup finds its frame by a binary search over the
// jobs' first workgroups (uniform across the workgroup) and runs the stage's
// body with the frame's workgroup coordinates: the per-frame kernels and the
// batched ones share one body per stage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.h"

namespace cjx {

struct Constants;
struct CodeSet;
struct Layout;
struct DcBlobs;

enum BatchStage : int {
    kBRegion = 0, kBQuantize, kBDcMaps, kBDcHist, kBHuffDc, kBDcBlobs, kBDcScan, kBDcEmit,
    kBTokens, kBReduce, kBHuffAc, kBAcEmit, kBLayout, kBAssemble, kBStages
};

struct StageGrid {
    std::uint32_t first;  // the frame's first workgroup in the batched launch
    std::uint32_t nx, ny;  // the frame's own grid
    std::uint32_t pad;
};

struct FrameJob {
    StageGrid grid[kBStages];
    Geometry geo;
    FrameParams prm;
    const std::uint8_t* luma;
    std::size_t luma_pitch;
    const std::uint8_t* chroma;
    std::size_t chroma_pitch;
    std::uint16_t* coeffs;
    std::int8_t* acs;
    std::int32_t* qf;
    float* cfl_part;
    std::int8_t* ytox;
    std::int8_t* ytob;
    std::int16_t* ac;
    std::int32_t* dc;
    std::uint8_t* nz_grid;
    std::uint16_t* nz_count;
    std::uint32_t* hists;
    std::uint32_t* fb_count;
    CodeSet* codes;
    const Constants* k;
    DcBlobs* blobs;
    std::uint32_t* blob_words;
    std::uint32_t* tokens;
    std::uint16_t* item_tokens;
    std::uint32_t* item_list;
    std::uint32_t* item_count;
    std::uint32_t* hist_partial;
    std::uint32_t* ac_staging;
    std::uint32_t* dc_staging;
    std::uint64_t* ac_part_bits;
    std::uint32_t* ac_group_bytes;
    std::uint64_t* ac_group_bits;
    std::uint64_t* chunk_bits;
    std::uint32_t* dc_group_bytes;
    std::uint64_t* dc_group_bits;
    Layout* lay;
    std::uint64_t* ac_base;
    std::uint64_t* dc_base;
    std::uint32_t* out;       // the file (device address; may be mapped host memory)
    std::uint64_t cap;        // its capacity in bytes
    std::uint64_t* size_out;  // the file size (device address)
    Clears clr;
};

// Tunable launch shapes of the batched stages (the autotuner's search space).
struct BatchShape {
    std::uint32_t token_cap;         // tokenizer workgroups per frame, at most
    std::uint32_t assemble_words;    // output words per assembly thread (sets its grid)
    // Region stage: 0 for one workgroup per region, else a persistent grid of
    // this many workgroups (a multiple of 8) that claim the regions in turn.
    std::uint32_t region_wgs;
};

// Stage grids of one frame for a shape (nx, ny only; `first` is the caller's).
void batch_stage_grids(const Geometry& geo, const BatchShape& shape, StageGrid (&grid)[kBStages]);

// Queues the 14 batched launches of a batch (all frames in `mode`): the
// region kernel, the AC chain, the layout and the assembly on `s`, the DC
// chain on `s_dc`, forked after the region kernel (event `fork`) and joined
// before the layout (event `join`).  A persistent region stage
// (shape.region_wgs != 0) claims its regions through the eight device
// counters at `counters`, which a one-workgroup launch clears first.
void launch_batch(hipStream_t s, hipStream_t s_dc, hipEvent_t fork, hipEvent_t join, const FrameJob* jobs,
                  int count, const std::uint32_t (&total)[kBStages], std::uint32_t mode, const BatchShape& shape,
                  std::uint32_t* counters);

}  // namespace cjx
