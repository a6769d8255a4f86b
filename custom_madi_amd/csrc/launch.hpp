// Host-side launch declarations shared between the kernel translation units
// and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.hpp"
#include "phase.hpp"
#include "rng.hpp"

namespace madi {

struct WalkDev {
    int64_t n_walkers;
    int32_t n_steps;
    int32_t steps_per_h;
    int32_t n_grid;
    int32_t use_golden;
    double sigma;
    double ts;
    double pp;
    uint64_t seed;
    uint64_t first_walker;
    // Acceptance as integer thresholds on the 53-bit uniform mantissa k:
    // u = k * 2^-53 >= thr  <=>  k >= rej[m-1].  rej = 2^53 + 1 disables.
    uint64_t rej1;
    uint64_t rej2;
    // FP64 walk: 1 = graph labels with a certified ball
    // (MADI_CLASSIFY_FP64_CACHED), 0 = the exact classifier at every endpoint
    int32_t graph_labels;
};

// The walk's exit certificate (an inside walker's proposal past its cell,
// rejected for either crossing kind, needs no classification) pays where
// such exits come past a tier's ball, i.e. where the annulus is short next to
// a step: below 2.5 step deviations for the geometry's largest annulus.
__host__ __device__ inline bool exit_certificate_pays(float alpha_max, double sigma)
{
    return alpha_max < 2.5 * sigma;
}

// Longest Y grid (samples per walk) whose event table and occupancy counters
// the walk kernels keep in LDS.  The split walk handles only such entries;
// launches holding a longer one run the streamed walk (same results).
constexpr int kMaxLdsGrid = 256;

struct FusedOut {
    const PlanDev* plan;   // device copy
    Events ev;
    double* partial;       // tile slots: n_tiles x n_cols x 2
    int64_t tile_base;     // slot index of this launch's first tile
    long long* occupancy;  // n_grid
    long long* flags;      // MADI_N_FLAGS
};

// One work item of the persistent batch kernel: tiles [tile_lo, tile_lo +
// n_tiles) of one entry; work_begin is the global index of its first tile.
// Header of a batch launch in device memory: the tile counter the
// workgroups take tiles from (zero before the launch), the total number of
// tiles and of work items.  Kept in memory rather than in kernel arguments so
// a captured graph can run any batch.
struct BatchHdr {
    unsigned long long queue;
    int64_t total;
    int32_t n_work;
    int32_t pad;
    int64_t total_walkers;  // streamed walk: walkers of all items
};

// Streamed walk (lane refill): walkers [walker_lo, walker_lo + n_walk) of
// the entry, handed out from the item's counter (an array beside the table,
// zero before the launch, so the table itself stays read-only); walker_begin
// is the global index of the item's first walker.  Each walker leaves its
// moments (n_pairs x 3 floats) in `moments` and its active flag in `active`,
// both indexed by the entry's walker index.
struct WorkDev {
    GeomDev g;
    WalkDev w;
    int64_t tile_lo;
    int64_t n_tiles;
    int64_t work_begin;
    double* partial;
    long long* occupancy;
    long long* flags;
    int64_t walker_lo;
    int64_t n_walk;
    int64_t walker_begin;
    float* moments;
    uint8_t* active;
    double* y_out;  // NULL, or Y at every sample: walker x n_grid x 3 (streamed walks only)
};

// p_p of each of several FP64 walks run in one launch (MADI_MAX_MULTI_WALKS)
struct MultiPP {
    int32_t n;
    double pp[MADI_MAX_MULTI_WALKS];
};

void set_error(const char* fmt, ...);
int check_launch(const char* what);

int launch_classify_f64(const GeomDev& g, const double* pts, int64_t n, int32_t* cell, int8_t* inside, int8_t* ok,
                        unsigned long long* counters, hipStream_t s);
int launch_rng_init(uint64_t seed, uint64_t first, int64_t n, uint64_t* out, const Gf2Col* jump_pow, hipStream_t s);
int launch_walk_f64(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const double* gi, const double* gd,
                    const double* gu, double* Y, int8_t* itr, int8_t* esc, int8_t* err, unsigned long long* counters,
                    hipStream_t s);
int launch_walk_f64_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          unsigned long long* counters, hipStream_t s);
int launch_count_inside(const int8_t* itr, int64_t n_walks, int64_t n, int32_t G, long long* counts, hipStream_t s);
int launch_walk_f64_multi(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const MultiPP& mp, double* Y,
                          int8_t* itr, int8_t* esc, int8_t* err, unsigned long long* counters, hipStream_t s);
// Sample chunks of the column reduce for n walkers and n_cols columns (the
// partial buffer holds n_cols x chunks x 2 doubles).
int32_t reduce_chunks(int64_t n, int32_t n_cols);
// Yt: n * G * 3 doubles of scratch (Y in grid-major order).
int launch_reduce_f64(const double* Y, int64_t n, int32_t G, const int32_t* jd, const int32_t* jD, const int32_t* js,
                      const double* pc, int32_t n_cols, double* Yt, double* part, int32_t n_chunks, double* cos_out,
                      double* sin_out, hipStream_t s);

int launch_classify_graph(const GeomDev& g, const double* pts, int64_t n, int32_t* cell, int8_t* inside, int8_t* ok,
                          unsigned long long* counters, hipStream_t s);
int launch_walk_f32_fused(const GeomDev& g, const WalkDev& p, const Gf2Col* jump_pow, const FusedOut& fo,
                          float cap_width, int rounds, unsigned long long* counters, hipStream_t s);
int launch_batch_f32(const WorkDev* work, BatchHdr* hdr, const PlanDev* plan, const Events& ev, const Gf2Col* jump_pow,
                     float cap_width, int rounds, int32_t grid, unsigned long long* counters, hipStream_t s);
// exits: compile in the exit certificate (worth it when some entry of the
// batch passes exit_certificate_pays; results are the same either way)
int launch_stream_f32(const WorkDev* work, unsigned long long* item_next, const BatchHdr* hdr, const PlanDev* plan,
                      const Events& ev, const Gf2Col* jump_pow, float cap_width, int rounds, bool exits, int32_t grid,
                      unsigned long long* counters, hipStream_t s);
// shape 1: 3 walker waves + 1 classifier wave per workgroup; 2: 2 + 2
// verify (MADI_N_VERIFY counters, or NULL): the verification build, which
// checks every label against the exact FP64 KD classifier
int launch_split_f32(const WorkDev* work, unsigned long long* item_next, const BatchHdr* hdr, const PlanDev* plan,
                     const Events& ev, const Gf2Col* jump_pow, float cap_width, int shape, bool exits, int32_t grid,
                     unsigned long long* counters, unsigned long long* verify, hipStream_t s);
int launch_emit_tiles(const WorkDev* work, const BatchHdr* hdr, const PlanDev* plan, int n_pairs, int32_t grid,
                      hipStream_t s);
// Column sums of an entry's tile slots (see k_finalize_chunks for the order);
// scratch holds finalize_scratch_doubles(n_tiles, n_cols) doubles.
int64_t finalize_scratch_doubles(int64_t n_tiles, int32_t n_cols);
int launch_finalize_tiles(const double* partial, int64_t tile_lo, int64_t n_tiles, int32_t n_cols, double* cos_out,
                          double* sin_out, double* scratch, hipStream_t s);

}  // namespace madi
