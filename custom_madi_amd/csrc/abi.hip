// C ABI of libmadi_hip.so: argument checks, geometry upload, jump matrices,
// and launches.  See include/madi_hip.h for the contract of each function.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"
#include "launch.hpp"
#include "phase.hpp"
#include "rng.hpp"

namespace madi {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int check_launch(const char* what)
{
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s launch failed: %s", what, hipGetErrorString(e));
        return MADI_ERR_HIP;
    }
    return MADI_OK;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
            return _e == hipErrorOutOfMemory ? MADI_ERR_OOM : MADI_ERR_HIP;            \
        }                                                                              \
    } while (0)

// ---------------------------------------------------- GF(2) jump matrices

struct HostState {
    uint64_t s0, s1;
};

static uint64_t host_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void host_next(HostState& s)
{
    uint64_t a = s.s0, b = s.s1;
    b ^= a;
    s.s0 = host_rotl(a, 55) ^ b ^ (b << 14);
    s.s1 = host_rotl(b, 36);
}

// Published xoroshiro128+ jump (2^64 steps).
static HostState host_jump(HostState s)
{
    static const uint64_t J[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
    uint64_t a = 0, b = 0;
    for (int w = 0; w < 2; ++w)
        for (int bit = 0; bit < 64; ++bit) {
            if (J[w] & (1ull << bit)) {
                a ^= s.s0;
                b ^= s.s1;
            }
            host_next(s);
        }
    return HostState{a, b};
}

static HostState mat_apply(const Gf2Col* M, HostState v)
{
    uint64_t lo = 0, hi = 0;
    for (int j = 0; j < 128; ++j) {
        const uint64_t bit = j < 64 ? (v.s0 >> j) & 1 : (v.s1 >> (j - 64)) & 1;
        if (bit) {
            lo ^= M[j].lo;
            hi ^= M[j].hi;
        }
    }
    return HostState{lo, hi};
}

// The jump matrices live in a file-scope device array (each GPU holds its
// own copy), filled once per device: no entry point allocates device memory
// for them.
static __device__ __align__(256) Gf2Col jump_table[kJumpPowers * 128];

struct DeviceTables {
    Gf2Col* jump_pow = nullptr;
};

static std::mutex g_tables_mu;
static DeviceTables g_tables[64];

static int get_jump_pow(const Gf2Col** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("device index %d out of range", dev);
        return MADI_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_tables_mu);
    if (!g_tables[dev].jump_pow) {
        std::vector<Gf2Col> P((size_t)kJumpPowers * 128);
        for (int j = 0; j < 128; ++j) {
            HostState e{j < 64 ? 1ull << j : 0ull, j < 64 ? 0ull : 1ull << (j - 64)};
            const HostState c = host_jump(e);
            P[j] = Gf2Col{c.s0, c.s1};
        }
        for (int b = 1; b < kJumpPowers; ++b) {
            const Gf2Col* A = &P[(size_t)(b - 1) * 128];
            for (int j = 0; j < 128; ++j) {
                const HostState col = mat_apply(A, HostState{A[j].lo, A[j].hi});
                P[(size_t)b * 128 + j] = Gf2Col{col.s0, col.s1};
            }
        }
        // self-check: matrix of one jump agrees with the jump routine
        HostState t{0x0123456789abcdefull, 0xfedcba9876543210ull};
        const HostState x = host_jump(t), y = mat_apply(&P[0], t);
        if (x.s0 != y.s0 || x.s1 != y.s1) {
            set_error("GF(2) jump matrix self-check failed");
            return MADI_ERR_HIP;
        }
        Gf2Col* d = nullptr;
        HIP_TRY(hipGetSymbolAddress((void**)&d, HIP_SYMBOL(jump_table)));
        HIP_TRY(hipMemcpy(d, P.data(), P.size() * sizeof(Gf2Col), hipMemcpyHostToDevice));
        g_tables[dev].jump_pow = d;
    }
    *out = g_tables[dev].jump_pow;
    return MADI_OK;
}

// ------------------------------------------------------------- geometry

}  // namespace madi

// An event recorded after a launch: complete once the launch (and all
// earlier work on its stream) has finished.  Shared by every geometry the
// launch reads; destroyed with its last reference.
struct UseEvent {
    hipEvent_t ev = nullptr;
    ~UseEvent()
    {
        if (ev) (void)hipEventDestroy(ev);
    }
};

struct madi_geom {
    madi::GeomDev dev;
    std::vector<void*> allocs;
    int64_t bytes = 0;
    int device = 0;
    // the last launch that reads this geometry (launches outside graph capture)
    mutable std::shared_ptr<UseEvent> used;
};

namespace madi {

// Geometry memory comes from the stream-ordered allocator on a per-device
// non-blocking stream (default priority): creating a geometry then waits
// for its own copies only, not for kernels running on the legacy default
// stream, and madi_geom_destroy_async can release it without waiting for the
// device.  A block freed on another stream may back the next geometry before
// that stream reaches the free, so a geometry is released only once every
// kernel that reads it has finished.
static int geom_stream(int dev, hipStream_t* out)
{
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    std::lock_guard<std::mutex> lock(mu);
    if ((int)streams.size() <= dev) streams.resize(dev + 1, nullptr);
    if (!streams[dev]) HIP_TRY(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    *out = streams[dev];
    return MADI_OK;
}

template <typename T>
static int upload(madi_geom* g, const T* host, size_t count, const T** out)
{
    *out = nullptr;
    if (count == 0) return MADI_OK;
    hipStream_t us = nullptr;
    int rc = geom_stream(g->device, &us);
    if (rc) return rc;
    void* d = nullptr;
    HIP_TRY(hipMallocAsync(&d, count * sizeof(T), us));
    g->allocs.push_back(d);
    g->bytes += (int64_t)(count * sizeof(T));
    HIP_TRY(hipMemcpyAsync(d, host, count * sizeof(T), hipMemcpyHostToDevice, us));
    HIP_TRY(hipStreamSynchronize(us));  // the host array may go away after this call
    *out = (const T*)d;
    return MADI_OK;
}

// Release after all work on the device (a plain destroy: whatever stream
// used the geometry has finished with it).
// Records on s an event behind everything queued there so far and makes it
// the last use of each geometry in gs.  Not while s is being captured into
// a graph: a captured record only runs at replay, so it says nothing about
// the kernels queued now (graph-captured uses are not tracked).
static int mark_used(const madi_geom* const* gs, int n, hipStream_t s)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return MADI_OK;
    auto u = std::make_shared<UseEvent>();
    HIP_TRY(hipEventCreateWithFlags(&u->ev, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(u->ev, s));
    for (int i = 0; i < n; ++i)
        if (gs[i]) gs[i]->used = u;
    return MADI_OK;
}

static int mark_used(const madi_geom* g, hipStream_t s) { return mark_used(&g, 1, s); }

static int mark_work_used(const madi_work* work, int32_t n_work, hipStream_t s)
{
    std::vector<const madi_geom*> gs((size_t)n_work);
    for (int32_t i = 0; i < n_work; ++i) gs[(size_t)i] = work[i].geom;
    return mark_used(gs.data(), n_work, s);
}

static void free_geom(madi_geom* g)
{
    if (g->allocs.empty()) return;
    hipStream_t us = nullptr;
    (void)hipDeviceSynchronize();
    if (geom_stream(g->device, &us) == MADI_OK) {
        for (void* p : g->allocs) (void)hipFreeAsync(p, us);
        (void)hipStreamSynchronize(us);
    }
    g->allocs.clear();
}

// Exact KD classifier arrays: the pre-order tree relabelled breadth-first so
// its first levels form a contiguous prefix (staged in LDS by the kernels).
static int build_kd(const madi_geom_desc* d, madi_geom* g)
{
    GeomDev& D = g->dev;
    const int n = d->n_kd;
    int rc;
    std::vector<int> order;
    order.reserve(n);
    std::vector<int> new_id(n, -1);
    order.push_back(0);
    new_id[0] = 0;
    for (size_t h = 0; h < order.size(); ++h) {
        const int o = order[h];
        for (int c : {d->node_left[o], d->node_right[o]}) {
            if (c < 0) continue;
            if (c >= n || new_id[c] >= 0) {
                set_error("geometry: KD tree is not a tree (node %d)", c);
                return MADI_ERR_ARG;
            }
            new_id[c] = (int)order.size();
            order.push_back(c);
        }
    }
    if ((int)order.size() != n) {
        set_error("geometry: KD tree reaches %zu of %d nodes", order.size(), n);
        return MADI_ERR_ARG;
    }
    std::vector<KdNodeF> nf(n);
    std::vector<KdNodeI> ni(n);
    for (int k = 0; k < n; ++k) {
        const int o = order[k];
        const int sid = d->node_seed[o];
        if (sid < 0 || sid >= n) {
            set_error("geometry: node_seed out of range");
            return MADI_ERR_ARG;
        }
        nf[k] = KdNodeF{d->kd_xyz[3 * sid], d->kd_xyz[3 * sid + 1], d->kd_xyz[3 * sid + 2], 0.0};
        const int L = d->node_left[o], R = d->node_right[o], P = d->node_parent[o];
        ni[k] = KdNodeI{L >= 0 ? new_id[L] : -1, R >= 0 ? new_id[R] : -1, P >= 0 ? new_id[P] : -1,
                        (sid << 2) | (d->node_axis[o] & 3)};
    }
    std::vector<int32_t> label(n);
    for (int k = 0; k < n; ++k) label[k] = d->kd_label ? d->kd_label[k] : k;
    if ((rc = upload(g, nf.data(), nf.size(), &D.kd_f))) return rc;
    if ((rc = upload(g, ni.data(), ni.size(), &D.kd_i))) return rc;
    if ((rc = upload(g, d->kd_annulus, (size_t)n, &D.kd_annulus))) return rc;
    if ((rc = upload(g, label.data(), label.size(), &D.kd_label))) return rc;
    if ((rc = upload(g, d->kd_xyz, (size_t)n * 3, &D.kd_seed_xyz))) return rc;
    D.n_kd = n;
    D.n_kd_lds = n < 511 ? n : 511;
    return MADI_OK;
}

static int build_geom(const madi_geom_desc* d, madi_geom* g)
{
    GeomDev& D = g->dev;
    memset(&D, 0, sizeof D);
    D.L = d->L;
    D.source_lo = d->source_lo;
    D.source_hi = d->source_hi;
    D.periodic = d->periodic;
    D.free_water = d->free_water;
    int rc;
    if (d->free_water) return MADI_OK;

    // exact KD classifier: relabel the pre-order tree breadth-first
    const int n = d->n_kd;
    if (n > 0 && (!d->kd_xyz || !d->kd_annulus || !d->node_seed || !d->node_axis || !d->node_left ||
                  !d->node_right || !d->node_parent)) {
        set_error("geometry: incomplete KD arrays (n_kd=%d)", n);
        return MADI_ERR_ARG;
    }
    if (n < 1 && !(d->n_cells > 0 && d->nb_offsets)) {
        set_error("geometry: a cellular geometry needs a KD tree or a Delaunay graph");
        return MADI_ERR_ARG;
    }
    if (n > 0 && (rc = build_kd(d, g))) return rc;

    // Delaunay graph (optional: walks that only use the exact classifier may omit it)
    const int nc = d->n_cells;
    if (nc > 0 && d->nb_offsets && d->n_edges > 0) {
        const int64_t E = d->n_edges;
        if (d->nb_offsets[0] != 0 || d->nb_offsets[nc] != E) {
            set_error("geometry: nb_offsets must run 0..n_edges");
            return MADI_ERR_ARG;
        }
        std::vector<int32_t> off(nc + 1);
        for (int c = 0; c < nc; ++c)
            if (d->nb_offsets[c + 1] < d->nb_offsets[c]) {
                set_error("geometry: nb_offsets must be non-decreasing");
                return MADI_ERR_ARG;
            }
        // padded offsets: each cell's list rounded up to a multiple of kEdgePad
        int64_t padded = 0;
        for (int c = 0; c < nc; ++c) {
            off[c] = (int32_t)padded;
            const int64_t deg = d->nb_offsets[c + 1] - d->nb_offsets[c];
            if (deg > kMaxCellEdges) {
                set_error("geometry: a cell has more than 2^24 - 8 edges");
                return MADI_ERR_ARG;
            }
            padded += (deg + kEdgePad - 1) / kEdgePad * kEdgePad;
            if (padded > INT32_MAX) {
                set_error("geometry: too many edges");
                return MADI_ERR_ARG;
            }
        }
        off[nc] = (int32_t)padded;
        // each cell's edges sorted by half length (the FP32 walk stops its edge
        // loop once h_j - |p| passes the bound it needs), then sentinels
        std::vector<float4> rec(padded, make_float4(0.f, 0.f, 0.f, INFINITY));
        std::vector<int32_t> ids(padded);
        std::vector<double> vec(3 * padded, 0.0);
        for (int c = 0; c < nc; ++c)
            for (int64_t k = off[c]; k < off[c + 1]; ++k) {
                ids[k] = c;
                vec[3 * k] = 1e30;
            }
        std::vector<float> alpha(nc), hmax(nc, 0.f);
        std::vector<double> alpha64(nc);
        std::vector<std::pair<double, int64_t>> order;
        float amin = INFINITY, amax = 0.f;
        for (int c = 0; c < nc; ++c) {
            alpha64[c] = d->cell_annulus[c];
            alpha[c] = (float)d->cell_annulus[c];
            amin = fminf(amin, alpha[c]);
            amax = fmaxf(amax, alpha[c]);
            order.clear();
            for (int64_t k = d->nb_offsets[c]; k < d->nb_offsets[c + 1]; ++k) {
                const int nb = d->nb_ids[k];
                if (nb < 0 || nb >= nc) {
                    set_error("geometry: neighbour id out of range");
                    return MADI_ERR_ARG;
                }
                const double ex = d->nb_vec[3 * k], ey = d->nb_vec[3 * k + 1], ez = d->nb_vec[3 * k + 2];
                const double len = std::sqrt(ex * ex + ey * ey + ez * ez);
                if (!(len > 0.0)) {
                    set_error("Poisson seed realization contains coincident seeds");
                    return MADI_ERR_CLASSIFIER;
                }
                order.emplace_back(len, k);
            }
            std::sort(order.begin(), order.end());
            int64_t dst = off[c];
            for (const auto& pr : order) {
                const int64_t k = pr.second;
                const double ex = d->nb_vec[3 * k], ey = d->nb_vec[3 * k + 1], ez = d->nb_vec[3 * k + 2];
                const double len = pr.first;
                rec[dst] = make_float4((float)(ex / len), (float)(ey / len), (float)(ez / len), (float)(0.5 * len));
                ids[dst] = d->nb_ids[k];
                vec[3 * dst] = ex;
                vec[3 * dst + 1] = ey;
                vec[3 * dst + 2] = ez;
                hmax[c] = fmaxf(hmax[c], (float)(0.5 * len));
                ++dst;
            }
        }
        if ((rc = upload(g, off.data(), off.size(), &D.nb_off))) return rc;
        if ((rc = upload(g, rec.data(), rec.size(), &D.nb_rec))) return rc;
        if ((rc = upload(g, ids.data(), ids.size(), &D.nb_id))) return rc;
        if ((rc = upload(g, vec.data(), vec.size(), &D.nb_e64))) return rc;
        if ((rc = upload(g, alpha.data(), alpha.size(), &D.cell_alpha))) return rc;
        if ((rc = upload(g, alpha64.data(), alpha64.size(), &D.cell_alpha64))) return rc;
        if ((rc = upload(g, hmax.data(), hmax.size(), &D.cell_hmax))) return rc;
        // one-load records of a cell and of each edge's neighbour, so a scan
        // of the next cell starts without looking up its list first
        auto bits = [](float v) {
            int32_t b;
            memcpy(&b, &v, sizeof b);
            return b;
        };
        std::vector<int4> crec(nc), krec(padded);
        std::vector<float> kalpha(padded);
        for (int c = 0; c < nc; ++c) crec[c] = make_int4(off[c], off[c + 1], bits(hmax[c]), bits(alpha[c]));
        for (int64_t j = 0; j < padded; ++j) {
            const int k = ids[j];
            krec[j] = make_int4(k, off[k], off[k + 1], bits(hmax[k]));
            kalpha[j] = alpha[k];
        }
        if ((rc = upload(g, crec.data(), crec.size(), &D.cell_rec))) return rc;
        if ((rc = upload(g, krec.data(), krec.size(), &D.nb_krec))) return rc;
        if ((rc = upload(g, kalpha.data(), kalpha.size(), &D.nb_kalpha))) return rc;
        if (n > 0) {
            // the FP64 walk's graph labels (classify_mode 5): FP64 unit edge
            // vectors and half lengths, and the cells whose lists may miss a
            // facet (cut by the construction's box) or hold a seed nearer than
            // 2e-3 um (where the KD classifier's nearest-seed rounding could
            // differ): those are left to the KD classifier
            std::vector<double4> u64(padded);
            for (int64_t j = 0; j < padded; ++j) {
                const double ex = vec[3 * j], ey = vec[3 * j + 1], ez = vec[3 * j + 2];
                if (ex >= 1e29) {
                    u64[j] = make_double4(0.0, 0.0, 0.0, INFINITY);
                    continue;
                }
                const double len = std::sqrt(ex * ex + ey * ey + ez * ez);
                u64[j] = make_double4(ex / len, ey / len, ez / len, 0.5 * len);
            }
            std::vector<int8_t> flags(nc, 0);
            for (int c = 0; c < nc; ++c) {
                if (d->cell_incomplete && d->cell_incomplete[c]) flags[c] = 1;
                for (int64_t j = off[c]; j < off[c + 1]; ++j)
                    if (u64[j].w < 1e-3) flags[c] = 1;
            }
            if ((rc = upload(g, u64.data(), u64.size(), &D.nb_u64))) return rc;
            if ((rc = upload(g, flags.data(), flags.size(), &D.cell_flags))) return rc;
        }
        if ((rc = upload(g, d->cell_xyz, (size_t)nc * 3, &D.cell_xyz))) return rc;
        if (d->hint_n < 1 || !d->hint) {
            set_error("geometry: a hint grid is required with the Delaunay graph");
            return MADI_ERR_ARG;
        }
        const size_t nh = (size_t)d->hint_n * d->hint_n * d->hint_n;
        for (size_t k = 0; k < nh; ++k)
            if (d->hint[k] < 0 || d->hint[k] >= nc) {
                set_error("geometry: hint cell out of range");
                return MADI_ERR_ARG;
            }
        if ((rc = upload(g, d->hint, nh, &D.hint))) return rc;
        D.n_cells = nc;
        D.hint_n = d->hint_n;
        D.hint_h = d->L / d->hint_n;
        D.alpha_min = amin;
        D.alpha_max = amax;
    }
    return MADI_OK;
}

static bool graph_ready(const GeomDev& g) { return g.free_water || g.nb_off != nullptr; }
static bool kd_ready(const GeomDev& g) { return g.free_water || g.kd_f != nullptr; }
// the FP64 walk's graph labels: FP64 edge records (uploaded with a KD tree)
static bool graph64_ready(const GeomDev& g) { return g.free_water || (g.nb_u64 != nullptr && g.cell_flags != nullptr); }

static float cap_width_of(float v) { return v > 0.f ? v : 5.0f; }
// FP32 walk classification mode -> kernel variant (0 per-lane state machine
// per tile, kVariantStream per-lane state machine with lane refill,
// kVariantSplit walker and classifier waves); 0 picks the default, the
// split one.  Mode 2 (a wave-cooperative lock-step variant) was removed.
constexpr int kVariantStream = 2, kVariantSplit = 3;
static int variant_of(int mode) { return mode == 1 ? 0 : (mode == 3 ? kVariantStream : kVariantSplit); }
// Variants that hand walkers to lanes one at a time (moments + tile emission).
static bool streamed(int variant) { return variant == kVariantStream || variant == kVariantSplit; }
// FP32 walk classification rounds per iteration (1-2); 0 picks the
// variant's default (1 streamed, where a round costs every lane of the wave
// but a walker that waits costs only its own lane; 2 per tile).  For the
// split variant the knob selects the workgroup's shape instead: 1 = three
// walker waves and one classifier wave, 2 = two and two.
static int rounds_of(int r, int variant)
{
    if (r <= 0) return streamed(variant) ? 1 : 2;
    return r > 2 ? 2 : r;
}

static int make_plan(const madi_phase_plan* plan, int32_t n_grid, PlanDev& P, Events& ev)
{
    memset(&P, 0, sizeof P);
    memset(&ev, 0, sizeof ev);
    if (!plan || plan->n_pairs < 1 || plan->n_pairs > MADI_MAX_PAIRS || plan->n_b < 1 || plan->n_b > MADI_MAX_B ||
        plan->n_dirs < 1 || plan->n_dirs > MADI_MAX_DIRS) {
        set_error("phase plan: need 1..%d pairs, 1..%d b values, 1..%d directions", MADI_MAX_PAIRS, MADI_MAX_B,
                  MADI_MAX_DIRS);
        return MADI_ERR_ARG;
    }
    P.n_pairs = plan->n_pairs;
    P.n_b = plan->n_b;
    P.n_dirs = plan->n_dirs;
    P.n_cols = plan->n_pairs * plan->n_b * plan->n_dirs;
    for (int p = 0; p < plan->n_pairs; ++p) {
        const int js[3] = {plan->j_delta[p], plan->j_Delta[p], plan->j_sum[p]};
        for (int t = 0; t < 3; ++t)
            if (js[t] < 0 || js[t] >= n_grid) {
                set_error("phase plan: grid index %d outside [0, %d)", js[t], n_grid);
                return MADI_ERR_ARG;
            }
        P.j_delta[p] = js[0];
        P.j_Delta[p] = js[1];
        P.j_sum[p] = js[2];
    }
    for (int k = 0; k < plan->n_pairs * plan->n_b; ++k) {
        P.pc[k] = plan->pc[k];
        P.pc32[k] = (float)plan->pc[k];
    }
    for (int k = 0; k < plan->n_dirs * 3; ++k) {
        P.dirs[k] = plan->dirs[k];
        P.dirs32[k] = (float)plan->dirs[k];
    }
    ev.n_pairs = plan->n_pairs;
    // events sorted by grid index (stable in pair order)
    for (int j = 0; j < n_grid; ++j)
        for (int p = 0; p < plan->n_pairs; ++p) {
            const int js[3] = {P.j_delta[p], P.j_Delta[p], P.j_sum[p]};
            const int sg[3] = {1, 1, -1};
            for (int t = 0; t < 3; ++t)
                if (js[t] == j) {
                    ev.j[ev.n] = j;
                    ev.pair[ev.n] = p;
                    ev.sign[ev.n] = sg[t];
                    ++ev.n;
                }
        }
    return MADI_OK;
}

static uint64_t reject_threshold(double thr)
{
    // smallest k with k * 2^-53 >= thr (k < 2^53); 2^53 never triggers
    if (!(thr > 0.0)) return 0;
    if (thr >= 1.0) return 1ull << 53;
    const double k = std::ceil(std::ldexp(thr, 53));
    return (uint64_t)k;
}

static int make_walk(const madi_walk_params* p, WalkDev& W)
{
    const bool shape_ok = p && p->n_walkers >= 0 && p->n_steps >= 0 && p->steps_per_h >= 1 && p->n_grid >= 1 &&
                          p->n_steps / p->steps_per_h < p->n_grid;
    if (!shape_ok) {
        set_error("walk params: need n_walkers >= 0, steps_per_h >= 1 and n_steps / steps_per_h < n_grid");
        return MADI_ERR_ARG;
    }
    if (!(p->pp >= 0.0 && p->pp <= 1.0)) {
        set_error("p_p must be finite and in [0, 1], got %g", p->pp);
        return MADI_ERR_ARG;
    }
    if (p->classify_mode < 0 || p->classify_mode > MADI_CLASSIFY_FP64_CACHED || p->classify_mode == 2 ||
        p->scan_rounds < 0 || p->scan_rounds > 2) {
        set_error("walk params: classify_mode must be 0, 1, 3, 4 or 5 and scan_rounds 0-2, got %d and %d",
                  p->classify_mode, p->scan_rounds);
        return MADI_ERR_ARG;
    }

    memset(&W, 0, sizeof W);
    W.n_walkers = p->n_walkers;
    W.n_steps = p->n_steps;
    W.steps_per_h = p->steps_per_h;
    W.n_grid = p->n_grid;
    W.sigma = p->sigma;
    W.ts = p->ts;
    W.pp = p->pp;
    W.seed = p->seed;
    W.first_walker = p->first_walker;
    W.graph_labels = p->classify_mode == MADI_CLASSIFY_FP64_CACHED ? 1 : 0;
    if (p->pp >= 1.0) {
        W.rej1 = W.rej2 = 1ull << 53;
    } else {
        W.rej1 = reject_threshold(p->pp);
        W.rej2 = reject_threshold(p->pp * p->pp);
    }
    return MADI_OK;
}

__global__ void k_fill_i64(long long* a, int64_t n, long long v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

static int fill_i64(long long* a, int64_t n, long long v, hipStream_t s)
{
    if (!a || n <= 0) return MADI_OK;
    hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, v);
    return check_launch("k_fill_i64");
}

static int zero_i64(long long* a, int n, hipStream_t s) { return fill_i64(a, n, 0, s); }

}  // namespace madi

using namespace madi;

// ------------------------------------------------------------- batching

namespace {

constexpr int64_t kAlign = 256;

int64_t align_up(int64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Batch workspace: [PlanDev | BatchHdr | WorkDev table | item walker
// counters], each part aligned.
static int64_t batch_off_hdr() { return align_up((int64_t)sizeof(PlanDev)); }
static int64_t batch_off_table() { return batch_off_hdr() + kAlign; }
static int64_t batch_off_next(int32_t n_work)
{
    return batch_off_table() + align_up((int64_t)n_work * (int64_t)sizeof(WorkDev));
}
int64_t batch_bytes(int32_t n_work)
{
    return batch_off_next(n_work) + align_up((int64_t)n_work * (int64_t)sizeof(unsigned long long));
}

// Grid-index span of a plan's events (a plan is valid for entries with at
// least this many Y-grid times).
static int32_t plan_grid_span(const madi_phase_plan* plan)
{
    int32_t m = 0;
    for (int p = 0; plan && p < plan->n_pairs && p < MADI_MAX_PAIRS; ++p)
        m = std::max(m, std::max(plan->j_sum[p], std::max(plan->j_Delta[p], plan->j_delta[p])));
    return m + 1;
}

// Launch knobs of a batch: tune's values, else the first item's, else defaults.
struct BatchKnobs {
    float cap_width;
    int rounds, variant, grid, emit_grid;
    bool exits;  // compile the exit certificate into the streamed walks
    unsigned long long* verify = nullptr;  // split walk: the verification build's counters
    int max_grid = 0;  // longest Y grid of the launch's entries (0: not known, at graph capture)
};

// Whether any item of a batch table can use the exit certificate.
static bool table_exits(const WorkDev* table, int32_t n_work)
{
    for (int i = 0; i < n_work; ++i)
        if (exit_certificate_pays(table[i].g.alpha_max, table[i].w.sigma)) return true;
    return false;
}

static int batch_knobs(const madi_tune* tune, const madi_walk_params* first, BatchKnobs& k)
{
    if (tune && (tune->classify_mode < 0 || tune->classify_mode > 4 || tune->classify_mode == 2 ||
                 tune->scan_rounds < 0 || tune->scan_rounds > 2)) {
        set_error("madi_tune: classify_mode must be 0, 1, 3 or 4 and scan_rounds 0-2, got %d and %d",
                  tune->classify_mode, tune->scan_rounds);
        return MADI_ERR_ARG;
    }
    k.cap_width = cap_width_of(tune && tune->tier_cap_um > 0.f ? tune->tier_cap_um : (first ? first->tier_cap_um : 0.f));
    k.variant = variant_of(tune && tune->classify_mode > 0 ? tune->classify_mode : (first ? first->classify_mode : 0));
    k.rounds = rounds_of(tune && tune->scan_rounds > 0 ? tune->scan_rounds : (first ? first->scan_rounds : 0),
                         k.variant);
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = tune && tune->grid_per_cu > 0 ? tune->grid_per_cu : 4;
    k.grid = cus * per_cu;
    k.emit_grid = cus * 8;
    k.exits = true;  // callers that know their entries narrow it
    return MADI_OK;
}

// Bytes of an entry's moment buffer: n_walkers x n_pairs x 3 FP32, then
// n_walkers active flags.
static int64_t moments_bytes(int64_t n_walkers, int n_pairs)
{
    return align_up(n_walkers * (int64_t)n_pairs * 3 * (int64_t)sizeof(float) + n_walkers);
}

// Workspace of a one-entry fused walk: [PlanDev | tile slots | BatchHdr |
// one WorkDev | its walker counter | moments | finalize chunk sums], each
// part aligned (the walker counter, table and moments for the streamed
// variant).
struct FusedLayout {
    int64_t partial, hdr, table, next, moments, final, total;
};

static FusedLayout fused_layout(int64_t n_walkers, const madi_phase_plan* plan)
{
    FusedLayout f;
    const int64_t tiles = (n_walkers + kBlock - 1) / kBlock;
    const int64_t cols = (int64_t)plan->n_pairs * plan->n_b * plan->n_dirs;
    f.partial = align_up((int64_t)sizeof(PlanDev));
    f.hdr = f.partial + align_up(tiles * cols * 2 * (int64_t)sizeof(double));
    f.table = f.hdr + kAlign;
    f.next = f.table + align_up((int64_t)sizeof(WorkDev));
    f.moments = f.next + kAlign;
    f.final = f.moments + moments_bytes(n_walkers, plan->n_pairs);
    f.total = f.final + align_up(finalize_scratch_doubles(tiles, (int32_t)cols) * (int64_t)sizeof(double));
    return f;
}

// Device table of a batch (validates every item); total = sum of tiles,
// walkers = sum of the items' walkers.  Items of a streamed launch (stream)
// need their entry's moment buffer.
static int batch_table(const madi_work* work, int32_t n_work, const madi_phase_plan* plan, WorkDev* table,
                       int64_t& total, const char* who, bool stream = false, int64_t* walkers = nullptr)
{
    int rc;
    total = 0;
    int64_t n_w = 0;
    for (int i = 0; i < n_work; ++i) {
        const madi_work& w = work[i];
        WorkDev& D = table[i];
        memset(&D, 0, sizeof D);
        if (!w.geom || !w.d_partial || !w.d_occupancy || !w.d_flags) {
            set_error("%s: work item %d has a NULL geometry or output", who, i);
            return MADI_ERR_ARG;
        }
        if ((rc = make_walk(&w.walk, D.w))) return rc;
        if (w.walk.classify_mode == MADI_CLASSIFY_FP64_CACHED) {
            set_error("%s: work item %d: classify_mode 5 is the FP64 walk's label cache; batches run the FP32 walk",
                      who, i);
            return MADI_ERR_ARG;
        }
        for (int p = 0; p < plan->n_pairs; ++p)
            if (plan->j_sum[p] >= w.walk.n_grid || plan->j_Delta[p] >= w.walk.n_grid ||
                plan->j_delta[p] >= w.walk.n_grid) {
                set_error("%s: work item %d has fewer Y-grid times than the plan needs", who, i);
                return MADI_ERR_ARG;
            }
        if (!graph_ready(w.geom->dev)) {
            set_error("%s: work item %d geometry has no Delaunay graph", who, i);
            return MADI_ERR_ARG;
        }
        const int64_t tiles = (w.walk.n_walkers + kBlock - 1) / kBlock;
        if (w.tile_lo < 0 || w.n_tiles < 0 || w.tile_lo + w.n_tiles > tiles) {
            set_error("%s: work item %d tiles [%lld, %lld) outside the entry's %lld tiles", who, i,
                      (long long)w.tile_lo, (long long)(w.tile_lo + w.n_tiles), (long long)tiles);
            return MADI_ERR_ARG;
        }
        D.g = w.geom->dev;
        D.tile_lo = w.tile_lo;
        D.n_tiles = w.n_tiles;
        D.work_begin = total;
        D.partial = w.d_partial;
        D.occupancy = (long long*)w.d_occupancy;
        D.flags = (long long*)w.d_flags;
        total += w.n_tiles;
        if (stream && !w.d_moments) {
            set_error("%s: work item %d has no moment buffer (classify_mode 3 and 4 need madi_work.d_moments)", who,
                      i);
            return MADI_ERR_ARG;
        }
        const int64_t hi = std::min<int64_t>(w.walk.n_walkers, (w.tile_lo + w.n_tiles) * kBlock);
        D.walker_lo = w.tile_lo * kBlock;
        D.n_walk = std::max<int64_t>(0, hi - D.walker_lo);
        D.walker_begin = n_w;
        D.moments = (float*)w.d_moments;
        D.active = w.d_moments ? (uint8_t*)w.d_moments + w.walk.n_walkers * (int64_t)plan->n_pairs * 3 * 4 : nullptr;
        if (w.d_Y && !stream) {
            set_error("%s: work item %d asks for Y, which only the streamed walks write (classify_mode 3 and 4)", who,
                      i);
            return MADI_ERR_ARG;
        }
        D.y_out = w.d_Y;
        n_w += D.n_walk;
    }
    if (walkers) *walkers = n_w;
    return MADI_OK;
}

// The kernels of one batch on s: the persistent tile kernel, or the
// streamed walk and the tile sums from its moments.
static int launch_batch(const WorkDev* table, unsigned long long* item_next, BatchHdr* hdr, const PlanDev* plan,
                        const Events& ev, const Gf2Col* jp, const BatchKnobs& k_in, int32_t grid,
                        unsigned long long* counters, hipStream_t s)
{
    // the split walk keeps the event table and occupancy counts of a walk in
    // LDS; a launch with a longer Y grid runs the streamed walk (same results)
    BatchKnobs k = k_in;
    if (k.variant == kVariantSplit && k.max_grid > kMaxLdsGrid) {
        k.variant = kVariantStream;
        k.rounds = rounds_of(0, kVariantStream);
    }
    if (!streamed(k.variant))
        return launch_batch_f32(table, hdr, plan, ev, jp, k.cap_width, k.rounds, grid, counters, s);
    int rc = k.variant == kVariantSplit
                 ? launch_split_f32(table, item_next, hdr, plan, ev, jp, k.cap_width, k.rounds, k.exits, grid, counters,
                                    k.verify, s)
                 : launch_stream_f32(table, item_next, hdr, plan, ev, jp, k.cap_width, k.rounds, k.exits, grid,
                                     counters, s);
    if (rc) return rc;
    return launch_emit_tiles(table, hdr, plan, ev.n_pairs, k.emit_grid, s);
}



// Scratch of madi_walk_fused_static: a file-scope device array, so that
// entry point allocates nothing (each GPU holds its own copy until the
// process ends).  Host code never passes the array's name to a kernel: its
// per-device address comes from hipGetSymbolAddress (static_scratch).
constexpr int64_t kStaticScratchBytes = 128ll << 20;
static __device__ __align__(256) char scratch[kStaticScratchBytes];

// Address of this device's scratch, looked up once per device with the
// device current.
static int static_scratch(char** out)
{
    static std::mutex mu;
    static char* ptrs[64] = {};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        set_error("device index %d out of range", dev);
        return MADI_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        char* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        ptrs[dev] = ptr;
    }
    *out = ptrs[dev];
    return MADI_OK;
}

// Largest staging image of a one-item fused walk (batch header, one work
// item, its walker counter), kept on the host stack.
constexpr int64_t kFusedImgBytes = 4096;

// The fused walk of one entry with its workspace at `base` (workspace_bytes
// long): argument checks, plan and table staging, launches.  Allocates no
// device or host memory.
static int fused_walk(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan, int32_t precision,
                      char* base, int64_t workspace_bytes, double* d_cos, double* d_sin, int64_t* d_occupancy,
                      int64_t* d_flags, int64_t* d_counters, hipStream_t s, const char* who,
                      int64_t* d_verify = nullptr, double* d_y = nullptr)
{
    WalkDev W;
    int rc = make_walk(p, W);
    if (rc) return rc;
    if (!g || !plan || !d_cos || !d_sin || !d_occupancy || !d_flags || !base) {
        set_error("%s: invalid arguments", who);
        return MADI_ERR_ARG;
    }
    if (precision != 32 && precision != 64) {
        set_error("%s: precision must be 32 or 64", who);
        return MADI_ERR_ARG;
    }
    if (precision == 32 && !graph_ready(g->dev)) {
        set_error("%s: the FP32 walk needs the geometry's Delaunay graph", who);
        return MADI_ERR_ARG;
    }
    if (precision == 32 && p->classify_mode == MADI_CLASSIFY_FP64_CACHED) {
        set_error("%s: classify_mode 5 is the FP64 walk's label cache (precision 64)", who);
        return MADI_ERR_ARG;
    }
    if (precision == 64 && !kd_ready(g->dev)) {
        set_error("%s: the FP64 walk needs the geometry's KD tree", who);
        return MADI_ERR_ARG;
    }
    if (precision == 64 && p->classify_mode == MADI_CLASSIFY_FP64_CACHED && !graph64_ready(g->dev)) {
        set_error("%s: the FP64 walk's graph labels (classify_mode 5) need the geometry's KD tree and Delaunay graph",
                  who);
        return MADI_ERR_ARG;
    }
    if (plan->n_pairs < 0 || plan->n_pairs > MADI_MAX_PAIRS) {
        set_error("%s: invalid phase plan", who);
        return MADI_ERR_ARG;
    }
    const FusedLayout lay = fused_layout(W.n_walkers, plan);
    if (workspace_bytes < lay.total) {
        set_error("%s: workspace of %lld bytes is smaller than the %lld required", who, (long long)workspace_bytes,
                  (long long)lay.total);
        return MADI_ERR_ARG;
    }
    if (lay.moments - lay.hdr > kFusedImgBytes) {
        set_error("%s: work table image exceeds %lld bytes", who, (long long)kFusedImgBytes);
        return MADI_ERR_ARG;
    }
    PlanDev P;
    Events ev;
    if ((rc = make_plan(plan, W.n_grid, P, ev))) return rc;
    PlanDev* dplan = (PlanDev*)base;
    double* partial = (double*)(base + lay.partial);
    HIP_TRY(hipMemcpyAsync(dplan, &P, sizeof P, hipMemcpyHostToDevice, s));
    if ((rc = zero_i64((long long*)d_occupancy, W.n_grid, s))) return rc;
    if ((rc = zero_i64((long long*)d_flags, MADI_N_FLAGS, s))) return rc;
    if (d_counters && (rc = zero_i64((long long*)d_counters, MADI_N_COUNTERS, s))) return rc;
    const Gf2Col* jp = nullptr;
    if ((rc = get_jump_pow(&jp))) return rc;
    const int64_t tiles = (W.n_walkers + kBlock - 1) / kBlock;
    if (tiles == 0) {
        // no walkers: every sum is +0.0 (all bits zero)
        if ((rc = zero_i64((long long*)d_cos, P.n_cols, s))) return rc;
        return zero_i64((long long*)d_sin, P.n_cols, s);
    }
    const int variant = variant_of(p->classify_mode);
    if (d_verify) {
        if (W.n_grid > kMaxLdsGrid) {
            set_error("%s: the verification build takes walks of at most %d samples", who, kMaxLdsGrid);
            return MADI_ERR_ARG;
        }
        if (precision != 32 || variant != kVariantSplit) {
            set_error("%s: the verification build is the FP32 split walk (precision 32, classify_mode 0 or 4)", who);
            return MADI_ERR_ARG;
        }
        if (!kd_ready(g->dev) || !graph_ready(g->dev)) {
            set_error("%s: verification needs the geometry's KD tree and Delaunay graph", who);
            return MADI_ERR_ARG;
        }
        if ((rc = zero_i64((long long*)d_verify, MADI_N_VERIFY, s))) return rc;
    }
    if (d_y && (precision != 32 || !streamed(variant))) {
        set_error("%s: Y output is written by the FP32 streamed and split walks (classify_mode 0, 3 or 4)", who);
        return MADI_ERR_ARG;
    }
    if (precision == 32 && streamed(variant)) {
        // one work item covering the entry
        madi_work item;
        memset(&item, 0, sizeof item);
        item.geom = g;
        item.walk = *p;
        item.tile_lo = 0;
        item.n_tiles = tiles;
        item.d_partial = partial;
        item.d_occupancy = d_occupancy;
        item.d_flags = d_flags;
        item.d_moments = base + lay.moments;
        item.d_Y = d_y;
        alignas(256) char img[kFusedImgBytes];
        const int64_t img_bytes = lay.moments - lay.hdr;
        memset(img, 0, (size_t)img_bytes);
        WorkDev* table = (WorkDev*)(img + (lay.table - lay.hdr));
        int64_t total = 0, walkers = 0;
        if ((rc = batch_table(&item, 1, plan, table, total, who, true, &walkers))) return rc;
        BatchHdr* h = (BatchHdr*)img;
        h->queue = 0;
        h->total = total;
        h->n_work = 1;
        h->total_walkers = walkers;
        // the copy zeroes the item's walker counter too (lay.next lies inside the image)
        HIP_TRY(hipMemcpyAsync(base + lay.hdr, img, (size_t)img_bytes, hipMemcpyHostToDevice, s));
        madi_tune tn;
        memset(&tn, 0, sizeof tn);
        tn.scan_rounds = rounds_of(p->scan_rounds, variant);
        tn.classify_mode = variant == kVariantSplit ? 4 : 3;  // the resolved variant
        tn.tier_cap_um = p->tier_cap_um;
        BatchKnobs k;
        if ((rc = batch_knobs(&tn, p, k))) return rc;
        k.exits = table_exits(table, 1);
        k.verify = (unsigned long long*)d_verify;
        k.max_grid = W.n_grid;
        const int grid = (int)std::min<int64_t>(k.grid, (walkers + 63) / 64);
        rc = launch_batch((WorkDev*)(base + lay.table), (unsigned long long*)(base + lay.next),
                          (BatchHdr*)(base + lay.hdr), dplan, ev, jp, k, grid, (unsigned long long*)d_counters, s);
        if (rc) return rc;
        return launch_finalize_tiles(partial, 0, tiles, P.n_cols, d_cos, d_sin, (double*)(base + lay.final), s);
    }
    FusedOut fo;
    fo.plan = dplan;
    fo.ev = ev;
    fo.partial = partial;
    fo.tile_base = 0;
    fo.occupancy = (long long*)d_occupancy;
    fo.flags = (long long*)d_flags;
    if (precision == 32)
        rc = launch_walk_f32_fused(g->dev, W, jp, fo, cap_width_of(p->tier_cap_um), rounds_of(p->scan_rounds, variant),
                                   (unsigned long long*)d_counters, s);
    else
        rc = launch_walk_f64_fused(g->dev, W, jp, fo, (unsigned long long*)d_counters, s);
    if (rc) return rc;
    return launch_finalize_tiles(partial, 0, tiles, P.n_cols, d_cos, d_sin, (double*)(base + lay.final), s);
}

}  // namespace

extern "C" {

int madi_abi_version(void) { return MADI_ABI_VERSION; }

const char* madi_last_error(void) { return g_err; }

int madi_get_device_info(madi_device_info* out)
{
    if (!out) {
        set_error("madi_get_device_info: out is NULL");
        return MADI_ERR_ARG;
    }
    memset(out, 0, sizeof *out);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    out->device = dev;
    out->cu_count = prop.multiProcessorCount;
    out->clock_khz = prop.clockRate;
    out->xcc_count = 0;
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    out->total_mem = (int64_t)tot;
    out->free_mem = (int64_t)fr;
    snprintf(out->name, sizeof out->name, "%s", prop.name);
    snprintf(out->arch, sizeof out->arch, "%s", prop.gcnArchName);
    return MADI_OK;
}

int madi_geom_create(const madi_geom_desc* desc, madi_geom** out)
{
    if (!desc || !out) {
        set_error("madi_geom_create: NULL argument");
        return MADI_ERR_ARG;
    }
    *out = nullptr;
    if (!(desc->L > 0.0)) {
        set_error("madi_geom_create: L must be positive");
        return MADI_ERR_ARG;
    }
    madi_geom* g = new madi_geom();
    (void)hipGetDevice(&g->device);
    const int rc = build_geom(desc, g);
    if (rc) {
        free_geom(g);
        delete g;
        return rc;
    }
    *out = g;
    return MADI_OK;
}

int madi_geom_destroy(madi_geom* g)
{
    if (!g) return MADI_OK;
    free_geom(g);
    delete g;
    return MADI_OK;
}

int madi_geom_destroy_async(madi_geom* g, void* stream)
{
    if (!g) return MADI_OK;
    if (g->used) {
        const hipError_t q = hipEventQuery(g->used->ev);
        if (q == hipErrorNotReady) {
            set_error("madi_geom_destroy_async: a kernel that reads the geometry has not finished (wait for its "
                      "stream, or use madi_geom_destroy); nothing was freed");
            return MADI_ERR_BUSY;
        }
        if (q != hipSuccess) {
            set_error("madi_geom_destroy_async: %s", hipGetErrorString(q));
            return MADI_ERR_HIP;
        }
    }
    hipError_t first = hipSuccess;
    for (void* p : g->allocs) {
        const hipError_t e = hipFreeAsync(p, (hipStream_t)stream);
        if (first == hipSuccess) first = e;
    }
    g->allocs.clear();
    delete g;
    if (first != hipSuccess) {
        set_error("madi_geom_destroy_async: %s", hipGetErrorString(first));
        return MADI_ERR_HIP;
    }
    return MADI_OK;
}

int madi_geom_device_bytes(const madi_geom* g, int64_t* bytes)
{
    if (!g || !bytes) {
        set_error("madi_geom_device_bytes: NULL argument");
        return MADI_ERR_ARG;
    }
    *bytes = g->bytes;
    return MADI_OK;
}

int madi_kd_build(const double* pts, int64_t n, int32_t* node_seed, int8_t* node_axis, int32_t* left, int32_t* right,
                  int32_t* parent)
{
    if (n < 0 || n > INT32_MAX || (n > 0 && (!pts || !node_seed || !node_axis || !left || !right || !parent))) {
        set_error("madi_kd_build: invalid arguments");
        return MADI_ERR_ARG;
    }
    std::vector<int32_t> idx((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        idx[(size_t)i] = (int32_t)i;
        left[i] = right[i] = parent[i] = -1;
    }
    // subsets [lo, hi) of idx with their parent node and side (0 left, 1
    // right, -1 root); right pushed before left, so ids come out in pre-order
    struct Item {
        int64_t lo, hi;
        int32_t par;
        int32_t side;
    };
    std::vector<Item> stack;
    stack.push_back({0, n, -1, -1});
    int32_t next = 0;
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        if (it.hi == it.lo) continue;
        double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t i = it.lo; i < it.hi; ++i) {
            const double* p = pts + 3 * (int64_t)idx[(size_t)i];
            for (int d = 0; d < 3; ++d) {
                mn[d] = std::min(mn[d], p[d]);
                mx[d] = std::max(mx[d], p[d]);
            }
        }
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
        const int64_t mid = (it.hi - it.lo) / 2;
        std::nth_element(idx.begin() + it.lo, idx.begin() + it.lo + mid, idx.begin() + it.hi,
                         [&](int32_t u, int32_t v) {
                             const double a = pts[3 * (int64_t)u + ax], b = pts[3 * (int64_t)v + ax];
                             return a < b || (a == b && u < v);
                         });
        const int32_t node = next++;
        node_seed[node] = idx[(size_t)(it.lo + mid)];
        node_axis[node] = (int8_t)ax;
        parent[node] = it.par;
        if (it.side == 0) left[it.par] = node;
        else if (it.side == 1) right[it.par] = node;
        stack.push_back({it.lo + mid + 1, it.hi, node, 1});
        stack.push_back({it.lo, it.lo + mid, node, 0});
    }
    return MADI_OK;
}

int madi_classify(const madi_geom* g, const double* d_pts, int64_t n, int32_t* d_cell, int8_t* d_inside, int8_t* d_ok,
                  void* stream)
{
    if (!g || n < 0 || (n > 0 && (!d_pts || !d_cell || !d_inside || !d_ok))) {
        set_error("madi_classify: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (!kd_ready(g->dev)) {
        set_error("madi_classify: geometry has no KD tree");
        return MADI_ERR_ARG;
    }
    const int rc = launch_classify_f64(g->dev, d_pts, n, d_cell, d_inside, d_ok, nullptr, (hipStream_t)stream);
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

int madi_classify_graph(const madi_geom* g, const double* d_pts, int64_t n, int32_t* d_cell, int8_t* d_inside,
                        int8_t* d_ok, int64_t* d_counters, void* stream)
{
    if (!g || n < 0 || (n > 0 && (!d_pts || !d_cell || !d_inside || !d_ok))) {
        set_error("madi_classify_graph: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (!graph_ready(g->dev)) {
        set_error("madi_classify_graph: geometry has no Delaunay graph");
        return MADI_ERR_ARG;
    }
    const int rc = launch_classify_graph(g->dev, d_pts, n, d_cell, d_inside, d_ok, (unsigned long long*)d_counters,
                                         (hipStream_t)stream);
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

int madi_rng_init(uint64_t seed, uint64_t first, int64_t n, uint64_t* d_state, void* stream)
{
    if (n < 0 || (n > 0 && !d_state) || first + (uint64_t)n >= (1ull << kJumpPowers)) {
        set_error("madi_rng_init: invalid arguments");
        return MADI_ERR_ARG;
    }
    const Gf2Col* jp = nullptr;
    int rc = get_jump_pow(&jp);
    if (rc) return rc;
    return launch_rng_init(seed, first, n, d_state, jp, (hipStream_t)stream);
}

int madi_walk_f64(const madi_geom* g, const madi_walk_params* p, const double* d_gold_init, const double* d_gold_incr,
                  const double* d_gold_u, double* d_Y, int8_t* d_inside_trace, int8_t* d_escaped, int8_t* d_err,
                  int64_t* d_counters, void* stream)
{
    WalkDev W;
    int rc = make_walk(p, W);
    if (rc) return rc;
    if (!g || (W.n_walkers > 0 && (!d_Y || !d_inside_trace || !d_escaped || !d_err))) {
        set_error("madi_walk_f64: invalid arguments");
        return MADI_ERR_ARG;
    }
    const int golden = d_gold_init != nullptr;
    if (golden && (!d_gold_incr || !d_gold_u)) {
        set_error("madi_walk_f64: an injected stream needs initial positions, increments and uniforms");
        return MADI_ERR_ARG;
    }
    W.use_golden = golden;
    if (!kd_ready(g->dev)) {
        set_error("madi_walk_f64: geometry has no KD tree");
        return MADI_ERR_ARG;
    }
    if (p->classify_mode == MADI_CLASSIFY_FP64_CACHED && !graph64_ready(g->dev)) {
        set_error("madi_walk_f64: graph labels (classify_mode 5) need the geometry's KD tree and Delaunay graph");
        return MADI_ERR_ARG;
    }
    const Gf2Col* jp = nullptr;
    if ((rc = get_jump_pow(&jp))) return rc;
    rc = launch_walk_f64(g->dev, W, jp, d_gold_init, d_gold_incr, d_gold_u, d_Y, d_inside_trace, d_escaped, d_err,
                         (unsigned long long*)d_counters, (hipStream_t)stream);
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

int madi_walk_f64_multi(const madi_geom* g, const madi_walk_params* p, int32_t n_walks, const double* pp,
                        double* d_Y, int8_t* d_inside_trace, int8_t* d_escaped, int8_t* d_err, int64_t* d_counters,
                        void* stream)
{
    WalkDev W;
    int rc = make_walk(p, W);
    if (rc) return rc;
    if (!g || n_walks < 1 || n_walks > MADI_MAX_MULTI_WALKS || !pp ||
        (W.n_walkers > 0 && (!d_Y || !d_inside_trace || !d_escaped || !d_err))) {
        set_error("madi_walk_f64_multi: invalid arguments (n_walks must lie in 1..%d)", MADI_MAX_MULTI_WALKS);
        return MADI_ERR_ARG;
    }
    MultiPP mp;
    memset(&mp, 0, sizeof mp);
    mp.n = n_walks;
    for (int k = 0; k < n_walks; ++k) {
        if (!(pp[k] >= 0.0 && pp[k] <= 1.0)) {
            set_error("p_p must be finite and in [0, 1], got %g (walk %d)", pp[k], k);
            return MADI_ERR_ARG;
        }
        mp.pp[k] = pp[k];
    }
    if (!kd_ready(g->dev)) {
        set_error("madi_walk_f64_multi: geometry has no KD tree");
        return MADI_ERR_ARG;
    }
    if (p->classify_mode == MADI_CLASSIFY_FP64_CACHED && !graph64_ready(g->dev)) {
        set_error("madi_walk_f64_multi: graph labels (classify_mode 5) need the geometry's KD tree and Delaunay graph");
        return MADI_ERR_ARG;
    }
    const Gf2Col* jp = nullptr;
    if ((rc = get_jump_pow(&jp))) return rc;
    rc = launch_walk_f64_multi(g->dev, W, jp, mp, d_Y, d_inside_trace, d_escaped, d_err,
                               (unsigned long long*)d_counters, (hipStream_t)stream);
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

int madi_count_inside(const int8_t* d_inside_trace, int64_t n_walks, int64_t n_walkers, int32_t n_grid,
                      int64_t* d_counts, void* stream)
{
    if (n_walks < 0 || n_walkers < 0 || n_grid < 0 || n_walks > 0x7fffffff ||
        (n_walks > 0 && n_grid > 0 && (!d_counts || (n_walkers > 0 && !d_inside_trace)))) {
        set_error("madi_count_inside: invalid arguments");
        return MADI_ERR_ARG;
    }
    return launch_count_inside(d_inside_trace, n_walks, n_walkers, n_grid, (long long*)d_counts,
                               (hipStream_t)stream);
}

int madi_reduce_f64(const double* d_Y, int64_t n, int32_t n_grid, const int32_t* d_jd, const int32_t* d_jD,
                    const int32_t* d_js, const double* d_pc, int32_t n_cols, double* d_cos, double* d_sin,
                    void* stream)
{
    if (n < 0 || n_grid < 1 || n_cols < 0 || (n_cols > 0 && (!d_jd || !d_jD || !d_js || !d_pc || !d_cos || !d_sin)) ||
        (n > 0 && !d_Y)) {
        set_error("madi_reduce_f64: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (n_cols == 0) return MADI_OK;
    if (n_cols > MADI_MAX_REDUCE_COLS) {
        set_error("madi_reduce_f64: more than MADI_MAX_REDUCE_COLS columns");
        return MADI_ERR_ARG;
    }
    const int32_t chunks = reduce_chunks(n, n_cols);
    double* part = nullptr;
    double* yt = nullptr;
    HIP_TRY(hipMallocAsync((void**)&part, (size_t)n_cols * chunks * 2 * sizeof(double), (hipStream_t)stream));
    if (n > 0) {
        const hipError_t e = hipMallocAsync((void**)&yt, (size_t)n * n_grid * 3 * sizeof(double), (hipStream_t)stream);
        if (e != hipSuccess) {
            (void)hipFreeAsync(part, (hipStream_t)stream);
            set_error("madi_reduce_f64: Y staging allocation failed: %s", hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? MADI_ERR_OOM : MADI_ERR_HIP;
        }
    }
    int rc = launch_reduce_f64(d_Y, n, n_grid, d_jd, d_jD, d_js, d_pc, n_cols, yt, part, chunks, d_cos, d_sin,
                               (hipStream_t)stream);
    if (yt) (void)hipFreeAsync(yt, (hipStream_t)stream);
    (void)hipFreeAsync(part, (hipStream_t)stream);
    return rc;
}

int madi_walk_fused_workspace(int64_t n_walkers, const madi_phase_plan* plan, int64_t* bytes)
{
    if (!plan || !bytes || n_walkers < 0 || plan->n_pairs < 0 || plan->n_pairs > MADI_MAX_PAIRS) {
        set_error("madi_walk_fused_workspace: invalid arguments");
        return MADI_ERR_ARG;
    }
    *bytes = fused_layout(n_walkers, plan).total;
    return MADI_OK;
}

int madi_moments_bytes(int64_t n_walkers, const madi_phase_plan* plan, int64_t* bytes)
{
    if (!plan || !bytes || n_walkers < 0 || plan->n_pairs < 0 || plan->n_pairs > MADI_MAX_PAIRS) {
        set_error("madi_moments_bytes: invalid arguments");
        return MADI_ERR_ARG;
    }
    *bytes = moments_bytes(n_walkers, plan->n_pairs);
    return MADI_OK;
}

int madi_walk_fused(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan, int32_t precision,
                    void* d_workspace, int64_t workspace_bytes, double* d_cos, double* d_sin, int64_t* d_occupancy,
                    int64_t* d_flags, int64_t* d_counters, void* stream)
{
    if (!d_workspace) {
        set_error("madi_walk_fused: invalid arguments");
        return MADI_ERR_ARG;
    }
    const int rc = fused_walk(g, p, plan, precision, (char*)d_workspace, workspace_bytes, d_cos, d_sin, d_occupancy,
                              d_flags, d_counters, (hipStream_t)stream, "madi_walk_fused");
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

// The phase plan of madi_walk_y_f32: one column (pair (0, 0, 0), pc 0), so
// the fused machinery runs unchanged and its single sum is ignored.
static madi_phase_plan y_walk_plan()
{
    madi_phase_plan plan;
    memset(&plan, 0, sizeof plan);
    plan.n_pairs = 1;
    plan.n_b = 1;
    plan.n_dirs = 1;
    plan.dirs[0] = 1.0;
    return plan;
}

int madi_walk_y_f32_workspace(int64_t n_walkers, int64_t* bytes)
{
    if (!bytes || n_walkers < 0) {
        set_error("madi_walk_y_f32_workspace: invalid arguments");
        return MADI_ERR_ARG;
    }
    const madi_phase_plan plan = y_walk_plan();
    *bytes = fused_layout(n_walkers, &plan).total + kAlign;  // + the dummy column's two sums
    return MADI_OK;
}

int madi_walk_y_f32(const madi_geom* g, const madi_walk_params* p, void* d_workspace, int64_t workspace_bytes,
                    double* d_Y, int64_t* d_occupancy, int64_t* d_flags, int64_t* d_counters, void* stream)
{
    if (!g || !p || !d_workspace || !d_Y || p->n_walkers < 0 || p->n_grid < 1) {
        set_error("madi_walk_y_f32: invalid arguments");
        return MADI_ERR_ARG;
    }
    const madi_phase_plan plan = y_walk_plan();
    const FusedLayout lay = fused_layout(p->n_walkers, &plan);
    if (workspace_bytes < lay.total + kAlign) {
        set_error("madi_walk_y_f32: workspace of %lld bytes is smaller than the %lld required",
                  (long long)workspace_bytes, (long long)(lay.total + kAlign));
        return MADI_ERR_ARG;
    }
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(hipMemsetAsync(d_Y, 0, (size_t)p->n_walkers * p->n_grid * 3 * sizeof(double), s));
    double* sums = (double*)((char*)d_workspace + lay.total);
    const int rc = fused_walk(g, p, &plan, 32, (char*)d_workspace, lay.total, sums, sums + 1, d_occupancy, d_flags,
                              d_counters, s, "madi_walk_y_f32", nullptr, d_Y);
    return rc ? rc : mark_used(g, s);
}

int madi_walk_fused_verify(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan,
                           void* d_workspace, int64_t workspace_bytes, double* d_cos, double* d_sin,
                           int64_t* d_occupancy, int64_t* d_flags, int64_t* d_counters, int64_t* d_verify,
                           void* stream)
{
    if (!d_workspace || !d_verify) {
        set_error("madi_walk_fused_verify: invalid arguments");
        return MADI_ERR_ARG;
    }
    const int rc = fused_walk(g, p, plan, 32, (char*)d_workspace, workspace_bytes, d_cos, d_sin, d_occupancy, d_flags,
                              d_counters, (hipStream_t)stream, "madi_walk_fused_verify", d_verify);
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

int madi_walk_fused_static_max_bytes(int64_t* bytes)
{
    if (!bytes) {
        set_error("madi_walk_fused_static_max_bytes: bytes is NULL");
        return MADI_ERR_ARG;
    }
    *bytes = kStaticScratchBytes;
    return MADI_OK;
}

int madi_walk_fused_static(const madi_geom* g, const madi_walk_params* p, const madi_phase_plan* plan,
                           int32_t precision, double* d_cos, double* d_sin, int64_t* d_occupancy, int64_t* d_flags,
                           int64_t* d_counters, void* stream)
{
    char* base = nullptr;
    int rc = static_scratch(&base);
    if (rc) return rc;
    rc = fused_walk(g, p, plan, precision, base, kStaticScratchBytes, d_cos, d_sin, d_occupancy, d_flags, d_counters,
                    (hipStream_t)stream, "madi_walk_fused_static");
    return rc ? rc : mark_used(g, (hipStream_t)stream);
}

}  // extern "C"


// A captured batch launch: [copy of the staged header and table to the
// device, persistent batch kernel (or the streamed walk and the tile-sum
// kernel)] as one hipGraph, replayed per batch.
struct madi_batch_graph {
    int device = -1;
    int32_t capacity = 0;
    madi_phase_plan plan;
    Events ev;
    BatchKnobs knobs;
    float tune_cap = 0.f;  // > 0 when the tune fixed the cap
    char* d_buf = nullptr;
    char* h_buf = nullptr;  // pinned staging, same layout as d_buf
    int64_t bytes = 0;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool launched = false;
    unsigned long long* counters = nullptr;
};

namespace {

static void graph_release(madi_batch_graph* gr)
{
    if (!gr) return;
    if (gr->launched && gr->done) (void)hipEventSynchronize(gr->done);
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    if (gr->cs) (void)hipStreamDestroy(gr->cs);
    if (gr->done) (void)hipEventDestroy(gr->done);
    if (gr->d_buf) (void)hipFree(gr->d_buf);
    if (gr->h_buf) (void)hipHostFree(gr->h_buf);
    delete gr;
}

}  // namespace

extern "C" {

int madi_batch_workspace(int32_t n_work, int64_t* bytes)
{
    if (!bytes || n_work < 0) {
        set_error("madi_batch_workspace: invalid arguments");
        return MADI_ERR_ARG;
    }
    *bytes = batch_bytes(n_work);
    return MADI_OK;
}

int madi_batch_launch(const madi_work* work, int32_t n_work, const madi_phase_plan* plan, const madi_tune* tune,
                      void* d_workspace, int64_t workspace_bytes, int64_t* d_counters, void* stream)
{
    if (n_work < 0 || (n_work > 0 && !work) || !d_workspace) {
        set_error("madi_batch_launch: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (workspace_bytes < batch_bytes(n_work)) {
        set_error("madi_batch_launch: workspace of %lld bytes is smaller than the %lld required",
                  (long long)workspace_bytes, (long long)batch_bytes(n_work));
        return MADI_ERR_ARG;
    }
    if (n_work == 0) return MADI_OK;
    int max_grid = 1;
    for (int i = 0; i < n_work; ++i) max_grid = std::max(max_grid, work[i].walk.n_grid);
    PlanDev P;
    Events ev;
    int rc = make_plan(plan, max_grid, P, ev);
    if (rc) return rc;
    // one host image of the workspace, one copy
    std::vector<char> img((size_t)batch_bytes(n_work), 0);
    memcpy(img.data(), &P, sizeof P);
    BatchKnobs k;
    if ((rc = batch_knobs(tune, &work[0].walk, k))) return rc;
    int64_t total = 0, walkers = 0;
    if ((rc = batch_table(work, n_work, plan, (WorkDev*)(img.data() + batch_off_table()), total,
                          "madi_batch_launch", streamed(k.variant), &walkers)))
        return rc;
    BatchHdr* h = (BatchHdr*)(img.data() + batch_off_hdr());
    h->queue = 0;
    h->total = total;
    h->n_work = n_work;
    h->total_walkers = walkers;
    if (total == 0) return MADI_OK;
    k.exits = table_exits((const WorkDev*)(img.data() + batch_off_table()), n_work);
    k.max_grid = max_grid;
    const Gf2Col* jp = nullptr;
    if ((rc = get_jump_pow(&jp))) return rc;
    hipStream_t s = (hipStream_t)stream;
    char* base = (char*)d_workspace;
    HIP_TRY(hipMemcpyAsync(base, img.data(), img.size(), hipMemcpyHostToDevice, s));
    const int grid = (int)std::min<int64_t>(k.grid, streamed(k.variant) ? (walkers + 63) / 64 : total);
    rc = launch_batch((WorkDev*)(base + batch_off_table()), (unsigned long long*)(base + batch_off_next(n_work)),
                      (BatchHdr*)(base + batch_off_hdr()), (PlanDev*)base, ev, jp, k, grid,
                      (unsigned long long*)d_counters, s);
    return rc ? rc : mark_work_used(work, n_work, s);
}

int madi_batch_graph_create(const madi_phase_plan* plan, int32_t capacity, const madi_tune* tune, int64_t* d_counters,
                            madi_batch_graph** out)
{
    if (!plan || !out || capacity < 1) {
        set_error("madi_batch_graph_create: invalid arguments");
        return MADI_ERR_ARG;
    }
    *out = nullptr;
    PlanDev P;
    madi_batch_graph* gr = new (std::nothrow) madi_batch_graph();
    if (!gr) {
        set_error("madi_batch_graph_create: out of host memory");
        return MADI_ERR_OOM;
    }
    int rc = make_plan(plan, plan_grid_span(plan), P, gr->ev);
    if (rc) {
        delete gr;
        return rc;
    }
    gr->plan = *plan;
    gr->capacity = capacity;
    gr->counters = (unsigned long long*)d_counters;
    gr->tune_cap = tune ? tune->tier_cap_um : 0.f;
    if ((rc = batch_knobs(tune, nullptr, gr->knobs))) {
        delete gr;
        return rc;
    }
    const Gf2Col* jp = nullptr;
    if ((rc = get_jump_pow(&jp))) {
        delete gr;
        return rc;
    }
    gr->bytes = batch_bytes(capacity);
    hipError_t e = hipGetDevice(&gr->device);
    if (e == hipSuccess) e = hipMalloc((void**)&gr->d_buf, (size_t)gr->bytes);
    if (e == hipSuccess) e = hipHostMalloc((void**)&gr->h_buf, (size_t)gr->bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&gr->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&gr->cs, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error("madi_batch_graph_create: %s", hipGetErrorString(e));
        graph_release(gr);
        return e == hipErrorOutOfMemory ? MADI_ERR_OOM : MADI_ERR_HIP;
    }
    memset(gr->h_buf, 0, (size_t)gr->bytes);
    memcpy(gr->h_buf, &P, sizeof P);
    e = hipStreamBeginCapture(gr->cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) e = hipMemcpyAsync(gr->d_buf, gr->h_buf, (size_t)gr->bytes, hipMemcpyHostToDevice, gr->cs);
    if (e == hipSuccess) {
        rc = launch_batch((WorkDev*)(gr->d_buf + batch_off_table()),
                          (unsigned long long*)(gr->d_buf + batch_off_next(capacity)),
                          (BatchHdr*)(gr->d_buf + batch_off_hdr()), (PlanDev*)gr->d_buf, gr->ev, jp, gr->knobs,
                          gr->knobs.grid, (unsigned long long*)d_counters, gr->cs);
        const hipError_t e2 = hipStreamEndCapture(gr->cs, &gr->graph);
        if (rc) {
            graph_release(gr);
            return rc;
        }
        e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        set_error("madi_batch_graph_create: capture failed: %s", hipGetErrorString(e));
        graph_release(gr);
        return MADI_ERR_HIP;
    }
    *out = gr;
    return MADI_OK;
}

int madi_batch_graph_launch(madi_batch_graph* gr, const madi_work* work, int32_t n_work, void* stream)
{
    if (!gr || n_work < 0 || (n_work > 0 && !work)) {
        set_error("madi_batch_graph_launch: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (n_work > gr->capacity) {
        set_error("madi_batch_graph_launch: %d work items exceed the graph's capacity of %d", n_work, gr->capacity);
        return MADI_ERR_ARG;
    }
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != gr->device) {
        set_error("madi_batch_graph_launch: graph belongs to device %d, current device is %d", gr->device, dev);
        return MADI_ERR_ARG;
    }
    if (n_work == 0) return MADI_OK;
    if (!(gr->tune_cap > 0.f) && cap_width_of(work[0].walk.tier_cap_um) != gr->knobs.cap_width) {
        set_error("madi_batch_graph_launch: the graph was captured for a tier cap of %g um", (double)gr->knobs.cap_width);
        return MADI_ERR_ARG;
    }
    // the previous replay has read the staging buffer once it completed
    if (gr->launched) HIP_TRY(hipEventSynchronize(gr->done));
    int64_t total = 0, walkers = 0;
    int rc = batch_table(work, n_work, &gr->plan, (WorkDev*)(gr->h_buf + batch_off_table()), total,
                         "madi_batch_graph_launch", streamed(gr->knobs.variant), &walkers);
    if (rc) return rc;
    BatchHdr* h = (BatchHdr*)(gr->h_buf + batch_off_hdr());
    h->queue = 0;
    h->total = total;
    h->n_work = n_work;
    h->total_walkers = walkers;
    if (total == 0) return MADI_OK;
    int max_grid = 0;
    for (int i = 0; i < n_work; ++i) max_grid = std::max(max_grid, work[i].walk.n_grid);
    if (gr->knobs.variant == kVariantSplit && max_grid > kMaxLdsGrid) {
        // the captured split walk keeps its grid in LDS: this batch runs
        // eagerly on the graph's buffers, and launch_batch picks the
        // streamed walk for it
        const Gf2Col* jp = nullptr;
        if ((rc = get_jump_pow(&jp))) return rc;
        hipStream_t s = (hipStream_t)stream;
        HIP_TRY(hipMemcpyAsync(gr->d_buf, gr->h_buf, (size_t)gr->bytes, hipMemcpyHostToDevice, s));
        BatchKnobs k = gr->knobs;
        k.max_grid = max_grid;
        rc = launch_batch((WorkDev*)(gr->d_buf + batch_off_table()),
                          (unsigned long long*)(gr->d_buf + batch_off_next(gr->capacity)),
                          (BatchHdr*)(gr->d_buf + batch_off_hdr()), (PlanDev*)gr->d_buf, gr->ev, jp, k, k.grid,
                          gr->counters, s);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(gr->done, s));
        gr->launched = true;
        return mark_work_used(work, n_work, s);
    }
    HIP_TRY(hipGraphLaunch(gr->exec, (hipStream_t)stream));
    HIP_TRY(hipEventRecord(gr->done, (hipStream_t)stream));
    gr->launched = true;
    return mark_work_used(work, n_work, (hipStream_t)stream);
}

int madi_batch_graph_destroy(madi_batch_graph* gr)
{
    graph_release(gr);
    return MADI_OK;
}

static bool finalize_plan_ok(const madi_phase_plan* plan)
{
    return plan && plan->n_pairs >= 0 && plan->n_pairs <= MADI_MAX_PAIRS && plan->n_b >= 0 && plan->n_b <= MADI_MAX_B &&
           plan->n_dirs >= 0 && plan->n_dirs <= MADI_MAX_DIRS;
}

int madi_tiles_finalize_workspace(const madi_phase_plan* plan, int64_t n_tiles, int64_t* bytes)
{
    if (!finalize_plan_ok(plan) || !bytes || n_tiles < 0) {
        set_error("madi_tiles_finalize_workspace: invalid arguments");
        return MADI_ERR_ARG;
    }
    const int32_t n_cols = plan->n_pairs * plan->n_b * plan->n_dirs;
    *bytes = std::max<int64_t>(1, finalize_scratch_doubles(n_tiles, n_cols)) * (int64_t)sizeof(double);
    return MADI_OK;
}

int madi_tiles_finalize_ws(const madi_phase_plan* plan, const double* d_partial, int64_t n_tiles, double* d_cos,
                           double* d_sin, void* d_workspace, int64_t workspace_bytes, void* stream)
{
    int64_t need = 0;
    if (!finalize_plan_ok(plan) || !d_partial || !d_cos || !d_sin || !d_workspace || n_tiles < 0 ||
        ((uintptr_t)d_workspace & 7) != 0) {
        set_error("madi_tiles_finalize_ws: invalid arguments");
        return MADI_ERR_ARG;
    }
    int rc = madi_tiles_finalize_workspace(plan, n_tiles, &need);
    if (rc) return rc;
    if (workspace_bytes < need) {
        set_error("madi_tiles_finalize_ws: workspace of %lld bytes is smaller than the %lld required",
                  (long long)workspace_bytes, (long long)need);
        return MADI_ERR_ARG;
    }
    const int32_t n_cols = plan->n_pairs * plan->n_b * plan->n_dirs;
    return launch_finalize_tiles(d_partial, 0, n_tiles, n_cols, d_cos, d_sin, (double*)d_workspace,
                                 (hipStream_t)stream);
}

int madi_tiles_finalize(const madi_phase_plan* plan, const double* d_partial, int64_t n_tiles, double* d_cos,
                        double* d_sin, void* stream)
{
    if (!finalize_plan_ok(plan) || !d_partial || !d_cos || !d_sin || n_tiles < 0) {
        set_error("madi_tiles_finalize: invalid arguments");
        return MADI_ERR_ARG;
    }
    int64_t need = 0;
    int rc = madi_tiles_finalize_workspace(plan, n_tiles, &need);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    double* scratch = nullptr;
    HIP_TRY(hipMallocAsync((void**)&scratch, (size_t)need, s));
    rc = madi_tiles_finalize_ws(plan, d_partial, n_tiles, d_cos, d_sin, scratch, need, stream);
    (void)hipFreeAsync(scratch, s);
    return rc;
}

int madi_fill_i64(int64_t* d, int64_t n, int64_t value, void* stream)
{
    if (n < 0 || (n > 0 && !d)) {
        set_error("madi_fill_i64: invalid arguments");
        return MADI_ERR_ARG;
    }
    if (n == 0) return MADI_OK;
    return fill_i64((long long*)d, n, (long long)value, (hipStream_t)stream);
}

}  // extern "C"
