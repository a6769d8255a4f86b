/*
 * madi_hip.h -- C ABI of the MI355X (gfx950) MADI random-walk library.
 *
 * Conventions
 *  - Every function returns an int status: MADI_OK (0) or a MADI_ERR_* code;
 *    madi_last_error() returns a thread-local message for the last failure.
 *  - Pointers named d_* are caller-owned device buffers (for example
 *    torch-ROCm tensors' data_ptr()); all other pointers are host memory.
 *  - void* stream is the caller's hipStream_t (NULL = the null stream).
 *    Functions enqueue work on it and return without synchronising unless
 *    the comment says otherwise.
 *  - No function aborts the process; device-side failures (fatal escape,
 *    classifier failure) come back as per-walker flags or counters.
 *
 * Each entry point names the reference interface it replaces; paths are
 * relative to the reference tree (jadenfigger/Custom_MADI).
 */
#ifndef MADI_HIP_H
#define MADI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MADI_ABI_VERSION 9

#define MADI_OK 0
#define MADI_ERR_ARG 1
#define MADI_ERR_HIP 2
#define MADI_ERR_OOM 3
#define MADI_ERR_CLASSIFIER 4
#define MADI_ERR_ESCAPE 5
#define MADI_ERR_BUSY 6

#define MADI_MAX_PAIRS 8
#define MADI_MAX_B 16
#define MADI_MAX_DIRS 64
/* Columns of one madi_reduce_f64 call (its grid holds 32 columns per block row). */
#define MADI_MAX_REDUCE_COLS 2000000

/* Counters returned by the walk kernels (int64 each). */
#define MADI_CNT_CLASSIFY 0      /* endpoint classifications run (graph or KD) */
#define MADI_CNT_SAFE_HIT 1      /* endpoints answered by the certified plane tier */
#define MADI_CNT_GRAPH_ITERS 2   /* Delaunay-graph anchor moves */
#define MADI_CNT_FP64_RECHECK 3  /* FP32 decisions re-decided in FP64 */
#define MADI_CNT_FACETS 4        /* facet (edge) evaluations */
#define MADI_CNT_KD_NODES 5      /* KD nodes visited (FP64 exact classifier) */
#define MADI_CNT_CROSSINGS 6     /* proposals with m > 0 */
#define MADI_CNT_REJECTS 7       /* rejected crossings */
#define MADI_CNT_WAVE_ITERS 8    /* FP32 walk: state-machine iterations, summed over waves */
#define MADI_CNT_LANE_ITERS 9    /* FP32 walk: iterations until each lane's walk ended, summed over lanes */
#define MADI_CNT_CLASS_ITERS 10  /* split FP32 walk: classifier-wave iterations, summed over waves */
#define MADI_CNT_CLASS_BUSY 11   /* split FP32 walk: busy classifier lanes, summed over those iterations */
#define MADI_CNT_EXITS 12        /* streamed FP32 walks: rejected exits certified by the tier (within SAFE_HIT) */
#define MADI_N_COUNTERS 13

/* Flags summed over walkers (int64 each). */
#define MADI_FLAG_ESCAPED 0
#define MADI_FLAG_CLASSIFIER_ERR 1
#define MADI_N_FLAGS 2

/* Counters of madi_walk_fused_verify: labels checked and labels that
 * disagree with the exact classifier, per source of the label. */
#define MADI_VERIFY_TIER_CHECKED 0        /* labels from the plane tier */
#define MADI_VERIFY_TIER_MISMATCH 1
#define MADI_VERIFY_EXIT_CHECKED 2        /* rejected exits certified without a classification */
#define MADI_VERIFY_EXIT_MISMATCH 3
#define MADI_VERIFY_CLASSIFIER_CHECKED 4  /* labels from a full FP32 classification (FP64 near a boundary) */
#define MADI_VERIFY_CLASSIFIER_MISMATCH 5
#define MADI_VERIFY_EXACT_FAILED 6        /* the exact classifier itself failed (coincident seeds) */
#define MADI_N_VERIFY 7

int madi_abi_version(void);
const char* madi_last_error(void);

typedef struct {
    int32_t device;
    int32_t cu_count;
    int32_t clock_khz;
    int32_t xcc_count;
    int64_t total_mem;
    int64_t free_mem;
    char name[64];
    char arch[32];
} madi_device_info;

/* Properties of the current device (no reference counterpart; used by the
 * benchmark's roofline and by admission control). */
int madi_get_device_info(madi_device_info* out);

/* ------------------------------------------------------------ geometry */

/* Host description of one ensemble geometry.  Replaces the per-walk device
 * upload of Ensemble arrays in madi/walker_gpu.py:871-877 (seeds, annulus
 * and the five KD arrays of madi/ensemble.py:525-563); the handle keeps one
 * device copy for every walk and k_io that uses the geometry.
 *
 * kd_*: the exact classifier's point set and its balanced KD tree in
 *   pre-order (node_seed/axis/left/right/parent as madi/ensemble.py builds
 *   them); n_kd = 0 gives a graph-only geometry for the FP32 walk.  kd_label maps each KD point to its cell (identity for the
 *   reference's non-periodic sets; periodic images map to their cell).
 * cell_*, nb_*: the Delaunay neighbour graph of the cells (CSR over cells,
 *   nb_vec = c_k + shift - c_a in FP64), used by the FP32 production walk.
 * hint: hint_n^3 ints over [0, L)^3, a cell near each box centre. */
typedef struct {
    int32_t n_kd;
    const double* kd_xyz;
    const double* kd_annulus;
    const int32_t* kd_label;
    const int32_t* node_seed;
    const int8_t* node_axis;
    const int32_t* node_left;
    const int32_t* node_right;
    const int32_t* node_parent;
    int32_t n_cells;
    const double* cell_xyz;
    const double* cell_annulus;
    int64_t n_edges;
    const int64_t* nb_offsets;
    const int32_t* nb_ids;
    const double* nb_vec;
    int32_t hint_n;
    const int32_t* hint;
    double L;
    double source_lo;
    double source_hi;
    int32_t periodic;
    int32_t free_water;
    /* optional (NULL: none), n_cells: 1 where a cell's neighbour list may
     * miss a facet (madi_voronoi_graph's `boxed`); the FP64 walk's graph
     * labels leave such cells to the KD classifier */
    const int8_t* cell_incomplete;
} madi_geom_desc;

typedef struct madi_geom madi_geom;

int madi_geom_create(const madi_geom_desc* desc, madi_geom** out);
/* Waits for all work on the device, then releases the handle. */
int madi_geom_destroy(madi_geom* g);
/* Releases the handle without waiting for the device: the memory is returned
 * to the stream-ordered allocator in the order of `stream` (a hipStream_t).
 * Every kernel that reads the geometry must have finished (an event recorded
 * after its last launch has completed): the freed memory may back a geometry
 * created afterwards before `stream` reaches this point.  The library checks
 * it: each entry point that launches kernels reading a geometry records an
 * event behind them (not while the stream is captured into a graph), and
 * while that event has not completed this call frees nothing and returns
 * MADI_ERR_BUSY.  No counterpart in
 * the reference (it re-uploads per walk); for schedulers that retire entries
 * while other work runs. */
int madi_geom_destroy_async(madi_geom* g, void* stream);
/* Device bytes held by the handle (admission control). */
int madi_geom_device_bytes(const madi_geom* g, int64_t* bytes);

/* ---------------------------------------------------------- host geometry */

/* Balanced KD tree of n points (pts: n x 3 FP64, host memory) in pre-order:
 * each node splits its subset on the axis of largest extent (the first on
 * ties) at the element of rank (size / 2) along that axis (equal coordinates
 * ordered by point index), its left subset before its right.  Outputs (host,
 * n each): node_seed, node_axis, left, right, parent (-1: none).  Replaces
 * build_kdtree (madi/ensemble.py:525-563); runs on the CPU, no device call. */
int madi_kd_build(const double* pts, int64_t n, int32_t* node_seed, int8_t* node_axis, int32_t* left, int32_t* right,
                  int32_t* parent);

/* Nearest of n points (pts: n x 3 FP64) to each of nq queries (q: nq x 3).
 * period > 0: the points and queries live in the periodic cube [0, period)^3
 * and distances take the minimum image; 0: plain distances.  exclude_self:
 * query i ignores point i (q == pts gives each point's nearest other point).
 * idx[nq] (nearest point, lowest index on ties) and dist[nq] (its distance,
 * sqrt((dx^2 + dy^2) + dz^2)) may each be NULL.  n_threads <= 0: up to 8
 * threads.  Replaces the cKDTree queries of the reference's geometry
 * construction: nearest-neighbour distances and the Eq. S10 certification
 * grid of _make_population / _nearest_distance_upper_bound
 * (madi/ensemble.py:566-606); runs on the CPU, no device call. */
int madi_nearest_points(const double* pts, int64_t n, const double* q, int64_t nq, double period,
                        int32_t exclude_self, int32_t n_threads, int32_t* idx, double* dist);

/* Voronoi neighbours (the Delaunay graph's edges) of the first n_owned of n
 * points (pts: n x 3 FP64).  Each cell is cut out of the points' bounding box
 * widened by pad (> 0) by the bisector planes of the other points in order of
 * distance, until every remaining point is farther than twice the cell's
 * largest vertex distance.  A plane within 1e-10 bin widths of a vertex is
 * kept as a neighbour (a superset of the exact edges is safe for the
 * classifier).  Output: offsets[n_owned + 1] and ids[offsets[i] ..
 * offsets[i+1]) ascending; n_edges = the total, also when it exceeds
 * capacity (then MADI_ERR_ARG and ids untouched); radius[n_owned] (or NULL)
 * each cell's largest vertex distance; boxed[n_owned] (or NULL) 1 where a cell
 * reaches the box.  Coincident points: MADI_ERR_CLASSIFIER (the reference's
 * classifier failure on such seeds, madi/walker_gpu.py:938-940).  No
 * counterpart in the reference, whose exact classifier scans a KD tree
 * (madi/walker_gpu.py:222-335); this graph feeds the FP32 walk's classifier.
 * Runs on the CPU, no device call. */
int madi_voronoi_graph(const double* pts, int64_t n, int64_t n_owned, double pad, int32_t n_threads,
                       int64_t* offsets, int32_t* ids, int64_t capacity, int64_t* n_edges, double* radius,
                       int8_t* boxed);

/* --------------------------------------------------------- classifiers */

/* Exact SI Eq. S2 label of n points (d_pts: n x 3 FP64).  FP64 stackless KD
 * traversal: nearest seed, then every seed within d1 + 2*alpha1.  Replaces
 * _classify_full_facet (madi/walker_gpu.py:272-335) and
 * Ensemble.classify_cpu (madi/ensemble.py:213-256).  d_ok[i] = 0 marks
 * coincident seeds (the reference's classifier failure). */
int madi_classify(const madi_geom* g, const double* d_pts, int64_t n, int32_t* d_cell, int8_t* d_inside,
                  int8_t* d_ok, void* stream);

/* The same labels from the FP32 Delaunay-graph classifier of the production
 * walk, certified in FP64 near every decision boundary.  d_counters may be
 * NULL (else MADI_N_COUNTERS int64, accumulated).  No reference counterpart
 * beyond madi_classify; it exists so tests can hold the production
 * classifier to the exact one. */
int madi_classify_graph(const madi_geom* g, const double* d_pts, int64_t n, int32_t* d_cell, int8_t* d_inside,
                        int8_t* d_ok, int64_t* d_counters, void* stream);

/* ----------------------------------------------------------------- RNG */

/* d_state[2i], d_state[2i+1] = xoroshiro128+ state of walker first+i for
 * seed (splitmix64 seeding, jump^(first+i)).  Replaces
 * create_xoroshiro128p_states (madi/walker_gpu.py:911, numba-src:224-291),
 * computed on the device. */
int madi_rng_init(uint64_t seed, uint64_t first, int64_t n, uint64_t* d_state, void* stream);

/* ---------------------------------------------------------------- walk */

typedef struct {
    int64_t n_walkers;
    int32_t n_steps;
    int32_t steps_per_h;
    int32_t n_grid;
    float tier_cap_um;      /* FP32 walk: plane-tier cap beyond max alpha (0 = default 5 um) */
    double sigma;
    double ts;
    double pp;
    uint64_t seed;          /* walk seed (device RNG) */
    uint64_t first_walker;  /* RNG subsequence index of walker 0 */
    int32_t scan_rounds;    /* FP32 walk: classification rounds per iteration, 1-2 (0 = the variant's default);
                               classify_mode 4: 1 = 3 walker + 1 classifier waves per workgroup, 2 = 2 + 2 */
    int32_t classify_mode;  /* FP32 walk: 0 default (4), 1 per-lane state machine per tile, 3 per-lane state
                               machine with lane refill (streamed), 4 streamed walker waves with classifier
                               waves (split); 2 (a removed lock-step variant) is rejected.  FP64 walk: 0
                               (or 1, 3, 4) the exact classifier at every endpoint, 5 graph labels with a
                               certified ball (MADI_CLASSIFY_FP64_CACHED; rejected by the FP32 walks) */
} madi_walk_params;

/* The FP64 walk's graph labels (the reference's "exact_cached" classifier
 * mode, madi/walker_gpu.py:337-518): each endpoint is labelled from the
 * geometry's Delaunay graph in FP64, starting at the walker's own cell (the
 * margins of its facets, greedy moves to a nearer seed), and the label also
 * certifies a ball around the endpoint (the distance to the nearest facet,
 * or to the contracted cell's boundary) within which later endpoints take it
 * without a scan.  A margin within 1e-9 um of a decision, a cell whose list
 * may be incomplete (madi_geom_desc.cell_incomplete) or seeds nearer than
 * 2e-3 um are left to the exact KD classifier, so every label is the exact
 * classifier's and the walk's results are its bit for bit.  Needs a geometry
 * with both the KD tree and the graph; tier_cap_um is not used. */
#define MADI_CLASSIFY_FP64_CACHED 5

/* Reference-semantics FP64 walk with the exact classifier at every endpoint.
 * Replaces _walk_kernel (madi/walker_gpu.py:520-656).  Random inputs come
 * from the device RNG (d_gold_* all NULL; states from seed/first_walker) or
 * from an injected stream (WalkRandomStream, madi/walker_gpu.py:86-127):
 * d_gold_init n x 3, d_gold_incr n_steps x n x 3, d_gold_u n_steps x n.
 * Outputs: d_Y n x n_grid x 3, d_inside_trace n x n_grid, d_escaped n,
 * d_err n.  d_counters may be NULL. */
int madi_walk_f64(const madi_geom* g, const madi_walk_params* p, const double* d_gold_init, const double* d_gold_incr,
                  const double* d_gold_u, double* d_Y, int8_t* d_inside_trace, int8_t* d_escaped, int8_t* d_err,
                  int64_t* d_counters, void* stream);

/* Several FP64 walks of one geometry that differ only in p_p, in one launch:
 * walk k is madi_walk_f64 with p->pp replaced by pp[k] (host array), bit for
 * bit, device RNG only.  The reference walks an ensemble once per k_io of its
 * sweep with one walk seed (run_simulation_multi_kio_reduced,
 * madi/walker_gpu.py:1170-1233, common random numbers); a walk of 50,000
 * walkers is 196 blocks, fewer than the GPU's 256 CUs, and the k_io walks
 * together fill it.  The walks of one walker run in adjacent lanes and share
 * each step's draws and their FP64 arithmetic (the same values as alone).
 * Outputs hold the walks one after another: d_Y n_walks x n x n_grid x 3,
 * d_inside_trace n_walks x n x n_grid, d_escaped and d_err n_walks x n,
 * d_counters (or NULL) n_walks x MADI_N_COUNTERS, zeroed by the caller.
 * 1 <= n_walks <= MADI_MAX_MULTI_WALKS. */
#define MADI_MAX_MULTI_WALKS 64
int madi_walk_f64_multi(const madi_geom* g, const madi_walk_params* p, int32_t n_walks, const double* pp,
                        double* d_Y, int8_t* d_inside_trace, int8_t* d_escaped, int8_t* d_err, int64_t* d_counters,
                        void* stream);

/* Occupancy of stacked walks, on the device: d_counts[k * n_grid + j] = the
 * sum over walkers w of d_inside_trace[(k * n_walkers + w) * n_grid + j].
 * The inside trace is as madi_walk_f64_multi writes it (n_walks x n_walkers
 * x n_grid int8 flags, 0 or 1), so each count is the number of walkers of
 * walk k inside at grid point j; the reference sums its inside trace on the
 * host (np.sum(inside, axis=0), madi/walker_gpu.py:1054).  d_counts: n_walks
 * x n_grid int64, overwritten. */
int madi_count_inside(const int8_t* d_inside_trace, int64_t n_walks, int64_t n_walkers, int32_t n_grid,
                      int64_t* d_counts, void* stream);

/* Workspace bytes of madi_walk_y_f32 for n_walkers. */
int madi_walk_y_f32_workspace(int64_t n_walkers, int64_t* bytes);

/* The FP32 production walk (the walk of madi_walk_fused at precision 32:
 * same random stream, labels and acceptance) writing Y instead of fused
 * moments: d_Y n x n_grid x 3 FP64, Y(t_j) being the walker's compensated
 * FP32 trapezoid sum widened to FP64 (sample 0 is zero).  With
 * madi_reduce_f64 it serves any column set, the universal grid included.
 * Replaces the Y output of _walk_kernel (madi/walker_gpu.py:520-656) on the
 * production path; d_occupancy (n_grid), d_flags (MADI_N_FLAGS) and
 * d_counters (may be NULL) as for madi_walk_fused.  classify_mode 0 or 4
 * (split walk) or 3 (streamed walk). */
int madi_walk_y_f32(const madi_geom* g, const madi_walk_params* p, void* d_workspace, int64_t workspace_bytes,
                    double* d_Y, int64_t* d_occupancy, int64_t* d_flags, int64_t* d_counters, void* stream);

/* Sigma cos / sigma sin over walkers and axes of pc*(Y[jd]+Y[jD]-Y[js]) per
 * column, FP64.  Replaces _reduce_kernel (madi/walker_gpu.py:658-695).  Every
 * jd/jD/js must lie in [0, n_grid); at most MADI_MAX_REDUCE_COLS columns.
 * Sums are formed in a fixed order (bitwise reproducible); consecutive
 * columns with equal grid indices (the b values of one pair) share dM. */
int madi_reduce_f64(const double* d_Y, int64_t n, int32_t n_grid, const int32_t* d_jd, const int32_t* d_jD,
                    const int32_t* d_js, const double* d_pc, int32_t n_cols, double* d_cos, double* d_sin,
                    void* stream);

/* Phase plan of the fused walk: n_pairs (delta, Delta) pairs given as Y-grid
 * indices, n_b phase coefficients per pair (gamma*G*1e-9), n_dirs unit
 * gradient directions.  Output column (pair, b, dir) has index
 * (pair * n_b + b) * n_dirs + dir. */
typedef struct {
    int32_t n_pairs;
    int32_t n_b;
    int32_t n_dirs;
    int32_t reserved;
    int32_t j_delta[MADI_MAX_PAIRS];
    int32_t j_Delta[MADI_MAX_PAIRS];
    int32_t j_sum[MADI_MAX_PAIRS];
    double pc[MADI_MAX_PAIRS * MADI_MAX_B];
    double dirs[MADI_MAX_DIRS * 3];
} madi_phase_plan;

/* Bytes of scratch the fused walk needs for n_walkers (fixed-slot per-tile
 * partial sums, the streamed variant's per-walker moments and its one-item
 * work table; any classify_mode). */
int madi_walk_fused_workspace(int64_t n_walkers, const madi_phase_plan* plan, int64_t* bytes);

/* Bytes of an entry's per-walker moment buffer (madi_work.d_moments):
 * n_walkers x n_pairs x 3 FP32 moments followed by n_walkers active flags.
 * No reference counterpart: the streamed variant hands walkers to lanes in
 * any order and forms the tile sums from this buffer afterwards. */
int madi_moments_bytes(int64_t n_walkers, const madi_phase_plan* plan, int64_t* bytes);

/* Production walk with the phase reduction fused in: FP32 seed-relative
 * state, Delaunay-graph classifier certified in FP64, compensated Y, FP64
 * cross-walker sums in a fixed order.  precision 32 selects this kernel;
 * precision 64 runs the same fused reduction on the FP64 exact-classifier
 * walk (the FP64 oracle run on the device).  Replaces the walk + reduce
 * pair of _walk_and_reduce_one_ensemble (madi/walker_gpu.py:989-1063) for
 * column sets that fit the plan.
 * Outputs: d_cos, d_sin (n_cols FP64 sums), d_occupancy (n_grid int64 sums of
 * the inside label at each Y-grid time), d_flags (MADI_N_FLAGS int64),
 * d_counters (MADI_N_COUNTERS int64, may be NULL).  All outputs are
 * overwritten. */
int madi_walk_fused(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan, int32_t precision,
                    void* d_workspace, int64_t workspace_bytes, double* d_cos, double* d_sin, int64_t* d_occupancy,
                    int64_t* d_flags, int64_t* d_counters, void* stream);

/* Verification build of the FP32 split walk (precision 32, classify_mode 0
 * or 4): the same walk as madi_walk_fused, bit for bit, in which every label
 * the walk takes -- from its plane tier, from an exit certificate, or from its
 * classifier wave -- is checked against the exact FP64 KD classifier of
 * madi_classify on the same represented point (the anchor's seed plus the
 * walker's FP32 offset, summed in FP64).  d_verify receives MADI_N_VERIFY
 * int64 counts (zeroed first).  The geometry needs both its KD tree and its
 * Delaunay graph.  Slower than madi_walk_fused (one exact classification per
 * step).  Replaces the reference's cached-vs-exact classifier equivalence
 * replay (scripts/validate_exact_classifier_cache.py:73-140) for this
 * build's tier, which has no bitwise twin: a tier answer keeps the walker's
 * anchor where a full classification moves it. */
int madi_walk_fused_verify(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan,
                           void* d_workspace, int64_t workspace_bytes, double* d_cos, double* d_sin,
                           int64_t* d_occupancy, int64_t* d_flags, int64_t* d_counters, int64_t* d_verify,
                           void* stream);

/* madi_walk_fused without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * scratch is a file-scope device array of madi_walk_fused_static_max_bytes
 * bytes (128 MiB; each GPU the library runs on holds it until the process
 * ends); an entry whose madi_walk_fused_workspace exceeds it is rejected with
 * MADI_ERR_ARG.  The call allocates nothing (no hipMalloc, hipFree or
 * stream-ordered allocation) and gives the same results as madi_walk_fused,
 * bit for bit.  Every call on a GPU shares that GPU's scratch and the scratch
 * keeps the last call's contents (each call clears what must start at zero):
 * calls that overlap on one GPU race, so serialise them (one stream, or
 * events). */
int madi_walk_fused_static_max_bytes(int64_t* bytes);
int madi_walk_fused_static(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan,
                           int32_t precision, double* d_cos, double* d_sin, int64_t* d_occupancy, int64_t* d_flags,
                           int64_t* d_counters, void* stream);

/* ------------------------------------------------------------- batching */

#define MADI_TILE_WALKERS 256

/* One work item of a batched launch: tiles [tile_lo, tile_lo + n_tiles) of
 * one library entry (tile t = walkers [256 t, 256 t + 256)).  d_partial
 * holds the entry's per-tile slots (ceil(n_walkers / 256) x n_cols x 2 FP64)
 * and d_occupancy (n_grid) / d_flags (MADI_N_FLAGS) accumulate over the
 * entry's launches; the caller zeroes them when it admits the entry. */
typedef struct {
    const madi_geom* geom;
    madi_walk_params walk;
    int64_t tile_lo;
    int64_t n_tiles;
    double* d_partial;
    int64_t* d_occupancy;
    int64_t* d_flags;
    void* d_moments;  /* the entry's madi_moments_bytes buffer; required by classify_mode 3 and 4 */
    double* d_Y;      /* NULL, or the entry's Y (n_walkers x n_grid x 3 FP64, zeroed by the caller): each
                         walker stores Y at samples 1.. (classify_mode 3 and 4; samples after an escape
                         are not written) */
} madi_work;

typedef struct {
    int32_t grid_per_cu;  /* persistent workgroups per CU (0 = 4) */
    float tier_cap_um;    /* plane-tier cap (0 = first item's walk.tier_cap_um) */
    int32_t scan_rounds;  /* 1-2 (split walk: its shape); 0 = first item's walk.scan_rounds */
    int32_t classify_mode;  /* as madi_walk_params.classify_mode; 0 = first item's */
} madi_tune;

/* Device workspace bytes of a batched launch of n_work items. */
int madi_batch_workspace(int32_t n_work, int64_t* bytes);

/* One persistent, work-stealing launch of the FP32 production walk over many
 * entries of uneven size: workgroups take tiles from an atomic counter until
 * every tile of every item is done (classify_mode 1), or waves take
 * walkers from per-item counters as lanes free up and a second kernel forms
 * the tile sums (classify_mode 3, 4).  Each tile writes its own slot, so an
 * entry's result is bitwise identical alone or in any mix, and equal across
 * classify modes.  Replaces the
 * one-entry-at-a-time loop of build_library_from_triplets
 * (madi/library.py:546-629) with one launch per batch or per serving
 * generation.  Asynchronous on stream. */
int madi_batch_launch(const madi_work* work, int32_t n_work, const madi_phase_plan* plan, const madi_tune* tune,
                      void* d_workspace, int64_t workspace_bytes, int64_t* d_counters, void* stream);

/* A batched launch captured once as a hipGraph (copy of a pinned staging
 * buffer holding the batch header and work table, then the persistent
 * kernel) and replayed for every batch with madi_batch_graph_launch: one
 * graph launch per serving generation instead of separate copies and
 * kernel launches.  The plan, the tuning (0 = defaults) and the counters
 * buffer are fixed at capture; a batch may hold up to `capacity` items.
 * Replaces the same loop as madi_batch_launch (madi/library.py:546-629). */
typedef struct madi_batch_graph madi_batch_graph;
int madi_batch_graph_create(const madi_phase_plan* plan, int32_t capacity, const madi_tune* tune, int64_t* d_counters,
                            madi_batch_graph** out);
/* Stage a batch and replay the graph on stream.  Waits for this graph's
 * previous replay to finish before overwriting its staging buffer (callers
 * alternate several graphs to keep batches in flight). */
int madi_batch_graph_launch(madi_batch_graph* graph, const madi_work* work, int32_t n_work, void* stream);
int madi_batch_graph_destroy(madi_batch_graph* graph);

/* Sum an entry's tile slots [0, n_tiles) into d_cos / d_sin (n_cols FP64
 * each) in a fixed order: chunks of 32 consecutive tiles in tile order, then
 * the chunk sums in chunk order (the order madi_walk_fused uses). */
int madi_tiles_finalize(const madi_phase_plan* plan, const double* d_partial, int64_t n_tiles, double* d_cos,
                        double* d_sin, void* stream);
/* madi_tiles_finalize with its scratch (madi_tiles_finalize_workspace bytes,
 * 8-byte aligned) passed in: no stream-ordered allocation, so the sum can be
 * queued behind work still running on `stream` while other streams allocate.
 * Same order and results.  No counterpart in the reference; for schedulers
 * that queue each entry's sum behind the launch that completes it. */
int madi_tiles_finalize_workspace(const madi_phase_plan* plan, int64_t n_tiles, int64_t* bytes);
int madi_tiles_finalize_ws(const madi_phase_plan* plan, const double* d_partial, int64_t n_tiles, double* d_cos,
                           double* d_sin, void* d_workspace, int64_t workspace_bytes, void* stream);

/* d[i] = value for i < n, by a fill kernel (usable where memset nodes are not). */
int madi_fill_i64(int64_t* d, int64_t n, int64_t value, void* stream);

/* MAP library matching, FP64.  d_measured is n_vox x n_feat and d_lib
 * n_lib x n_feat, both row-major; for each voxel row m the library row r of
 * smallest score, ties to the lowest row:
 *   fit_s0 = 0: score |r|^2 - 2 m.r, d_best_resid = |m|^2 + score
 *     (replaces fitters_gpu.map_match_gpu / _map_kernel_fixed,
 *      madi/fitters_gpu.py:96-118 and :414-438, the GPU path of
 *      library.match_voxels_batch, madi/library.py:1402-1450);
 *   fit_s0 = 1: s0 = m.r / rr, score |m|^2 - (m.r)^2 / rr with
 *     rr = max(|r|^2, 1e-30) and +inf where s0 <= 0; d_best_resid = score,
 *     d_best_s0 = s0 (replaces map_match_fits0_gpu / _map_kernel_free_s0,
 *     madi/fitters_gpu.py:120-154 and :441-463, the GPU path of
 *     library.match_voxels_batch_fits0, madi/library.py:1457-1522).
 * Dot products and squared norms accumulate over the features in order with
 * fused multiply-adds.  A row with no finite score returns row 0 (residual
 * +inf, or |m|^2 + inf, and s0 0), as the reference's scan does.  The
 * workspace holds madi_map_workspace bytes (8-byte aligned); d_best_s0 may be
 * NULL when fit_s0 = 0.  Asynchronous on stream. */
int madi_map_workspace(int64_t n_vox, int64_t n_lib, int64_t* bytes);
int madi_map_match(const double* d_measured, int64_t n_vox, const double* d_lib, int64_t n_lib, int32_t n_feat,
                   int32_t fit_s0, void* d_workspace, int64_t workspace_bytes, int64_t* d_best_idx,
                   double* d_best_resid, double* d_best_s0, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MADI_HIP_H */
