// engine.h — device-resident state and step orchestration shared by a
// single World3 (one scene) and a Batch3 (many scenes).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/phieng_fluid3.h"
#include "sph_device.h"

namespace ph {

enum Kind : int {
    KIND_CELL_KEY = 0,
    KIND_SCAN_REDUCE,
    KIND_SCAN_PARTIALS,
    KIND_SCAN_APPLY,
    KIND_SCATTER,
    KIND_RANK_GATHER,
    KIND_DENSITY,
    KIND_FORCE,
    KIND_VMAX,
    KIND_EXPORT,
    KIND_MONITOR,
    KIND_APPEND,
    KIND_BBOX,
    KIND_FILL,
    KIND_SETVEL,
    KIND_PERSISTENT,
    KIND_SUR_STEP,
    KIND_SUR_AUX,
    KIND_SUR_COLLECT,
    KIND_GHOST,
    KIND_RESAMPLE,
    KIND_COUNT
};
const char *kind_name(int k);

struct HipError : std::runtime_error {
    HipError(const char *what, hipError_t e)
        : std::runtime_error(std::string(what) + ": " + hipGetErrorString(e)) {}
};
#define PH_CK(x)                                   \
    do {                                           \
        hipError_t e_ = (x);                       \
        if (e_ != hipSuccess) throw ph::HipError(#x, e_); \
    } while (0)

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void swap(DBuf &o) {
        T *q = p;
        p = o.p;
        o.p = q;
        const size_t k = n;
        n = o.n;
        o.n = k;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grow to at least m elements; contents are not preserved.
    void ensure(size_t m) {
        if (m <= n && p) return;
        release();
        size_t want = m ? m : 1;
        PH_CK(hipMalloc((void **)&p, want * sizeof(T)));
        n = want;
    }
};

// Host-side description of one scene (one fluid + its static bodies).
struct SceneHost {
    FluidConfig3Ffi cfg;
    float g[3];
    float dt;
    std::vector<BodyDev> bodies;
    int dims = 3;  // 2: a World2 fluid, particles and gravity in the z = 0 plane
};

// f32 constants of one scene, computed with the reference's operation
// order (fluid3.rs:220-251, 649, 1028; the 2D kernels fluid.rs:265-285) so
// device results match it.
void fill_scene_physics(SceneDev &s, const SceneHost &h);
// Boundary ghosts of the scene's static bodies in the reference's
// generation order (xyz, mass in w): fluid3.rs:716-841 for a 3D scene,
// fluid.rs:666-743 for a 2D one, with spacing = h_kernel / 2.
void make_ghosts(const SceneHost &h, std::vector<float4> &out);
// Pick a grid domain (cell box) for a scene from the particle cell box and
// its bodies.  The domain only affects speed: particles outside it land in
// the scene's outlier bucket and are still handled exactly.
void choose_domain(SceneDev &s, const int32_t *pbox, uint32_t nparticles, const std::vector<BodyDev> &bodies);

struct ListOut;  // sph_kernels.h

struct Engine {
    int device = -1;
    hipStream_t stream = nullptr;

    // particle arrays (two generations; cs/ct select the current ones)
    DBuf<float4> posm[2], vel[2];
    DBuf<uint32_t> tag[2];
    int cs = 0, ct = 0;
    size_t alloc_particles = 0;
    DBuf<uint32_t> keys, slots, tmp;
    DBuf<float2> rq;
    // Neighbour list of the last density pass: the candidates within
    // (1 + list_skin) h in reference order, four slots per 16-byte group
    // (group g of slot i at uint4 g * alloc_particles + i, SL_CAP / 4 groups
    // plus one spare row), the count word with its SLF_ flags, and (when
    // list_cells) each slot's cell.  The force pass reads it; a surrogate
    // substep reuses it with skin > 0.
    DBuf<uint32_t> nl_cnt, nl_idx;
    DBuf<int4> nl_cell;
    float list_skin = 0.0f;
    bool list_cells = false;
    // With list_coef > 0 the list radius is a Verlet skin from the current
    // speeds, at least list_skin (launch_list_radius)
    float list_coef = 0.0f;
    // where the list radius pass also writes the largest speed, and the
    // counter of particles whose list did not fit (both may be null)
    float *list_vmax_out = nullptr;
    uint32_t *list_slow = nullptr;
    DBuf<float> tile_vmax, list_f2;
    // bucket tables
    DBuf<uint32_t> counts, start, partial;
    size_t alloc_buckets = 0;
    // scene / tile / body tables
    DBuf<SceneDev> d_scenes;
    DBuf<TileDev> d_tiles;
    DBuf<BodyDev> d_bodies;
    DBuf<int32_t> d_bbox, d_monitor, d_box6;  // per-tile cell boxes, monitor output, set-up box
    DBuf<uint32_t> d_vmax;
    DBuf<float> staging_f, export_buf;
    DBuf<uint8_t> staging_u8;
    // recording of intermediates (parity readback)
    uint32_t rec_flags = 0;
    bool rec_valid = false;
    DBuf<int32_t> rec_cells;
    DBuf<uint32_t> rec_sorted, rec_count;
    DBuf<float> rec_rho, rec_p, rec_acc;

    std::vector<SceneDev> scenes;
    std::vector<TileDev> tiles;
    std::vector<BodyDev> bodies;
    uint32_t n_total = 0;
    uint32_t nbuckets = 0;
    bool scenes_dirty = true, tiles_dirty = true, bodies_dirty = true;

    // domain monitor (asynchronous readback of outlier counts + boxes)
    int32_t *h_monitor = nullptr;
    hipEvent_t mon_event = nullptr;
    bool mon_pending = false;
    uint64_t substeps_done = 0;
    uint32_t monitor_every = 16;

    // kernel timing
    bool timing = false;
    struct Ev {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Ev> pending_ev;
    std::vector<hipEvent_t> ev_pool;
    double total_ms[KIND_COUNT] = {};
    uint64_t launches[KIND_COUNT] = {};

    Engine() = default;
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;
    ~Engine();

    void init_device();
    // Row stride of the neighbour list (entries of group g of slot i sit at
    // g * stride + i): the particle count rounded up to 256, so a list
    // spans no more rows than the particles fill whatever the capacity.
    uint32_t list_stride() const {
        return (uint32_t)std::min<size_t>(alloc_particles, ((size_t)n_total + 255) & ~(size_t)255);
    }
    // Make room for `cap` particles, keeping the current state.
    void ensure_particles(uint32_t cap);
    void build_tiles();
    void assign_buckets();
    void upload_tables();
    void ensure_record();
    // One substep of every scene (dt per scene in SceneDev::dt):
    // sort_phase + plain_phase + finish_substep, or substep_verlet.
    void substep();
    // The plain substep over Verlet lists (verlet.hip): the decision, the
    // sort (its kernels return at once unless the device decided to
    // rebuild), the density pass walking (or rebuilding) the rows and
    // putting rows out of order back in order, the force pass over the
    // neighbour list.
    void substep_verlet();
    // Verlet lists: used by substep() when enabled (not for a world with
    // boundary ghosts, whose density pass reads this substep's cells).
    // vl_valid says that the rows and the displacement bound describe the
    // current state; every other way of moving or reordering particles
    // clears it, and the next Verlet substep then rebuilds.
    bool vl_enabled = true;
    // below this many particles the Verlet form's extra work (the rows'
    // rebuild sweeps and halo builds) costs more than its walk saves: the
    // list form runs
    uint32_t vl_min_particles = kVlAutoMin;
    static constexpr uint32_t kVlAutoMin = 131072;
    float vl_skin = 0.1f;
    bool vl_valid = false;
    uint64_t vl_epoch = 0;
    uint32_t vl_n = 0;
    DBuf<uint32_t> vl_cnt, vl_idx, vl_word[2], vl_halo, vl_halo_n;
    DBuf<uint32_t> vl_partial;  // the sort's block sums, added up by the cell-key pass; zero between sorts
    bool vl_partial_dirty = true;  // not known to be zero (new, or a sort cut short)
    DBuf<uint4> vl_row16;
    DBuf<VerletDev> vl_state;
    bool vl_state_ready = false;
    // Rows whose neighbours changed cell cost their wave a re-ordering each
    // substep; in a fast flow they outnumber what the sort and list passes
    // of the list form cost (the two forms give the same bits).  Above
    // vl_repair_limit() such rows per substep the list form runs for
    // vl_list_left substeps (doubling on every return to a flow still that
    // fast, reset after a long quiet stretch), then the rows are rebuilt.
    // Every fourth Verlet substep decides, from what the decision kernel of
    // the Verlet substep kVlLag before left in pinned host memory (slot
    // vl_k % 32 of a ring; every fourth substep records an event of a ring
    // of 8), so the decision is the same on every run and the host may run
    // kVlLag substeps ahead of the device.
    static constexpr uint32_t kVlLag = 2;
    uint32_t *vl_host = nullptr;  // [32][2]: rows re-ordered, rebuild flag
    hipEvent_t vl_ev[8] = {};
    uint64_t vl_k = 0;  // Verlet substeps since the list form last ran
    uint32_t vl_list_left = 0, vl_backoff = 16;
    uint64_t vl_list_substeps = 0;
    uint32_t vl_repair_limit() const { return std::max<uint32_t>(1024u, n_total / 512u); }
    void sort_phase(uint32_t *perm_out);
    // Density pass (writing the neighbour list), boundary ghosts, force pass
    // (with af, also the surrogate's history, records and judge sums).
    // density_done: the density pass of this sorted state already ran
    // (density_phase); only the ghosts and the force pass remain.
    void plain_phase(float4 *acc_sorted, const AuxFuse *af = nullptr, bool density_done = false);
    // The density pass alone (the first of an adaptive substep), with the
    // same outputs as plain_phase's.
    void density_phase();
    // Adaptive resolution (fluid3.rs:923-987) over the sorted state and
    // densities of this substep.  Writes the new particle set, in its new
    // index order, into the unsorted generation and returns its count;
    // `changed` is false when nothing merged or split (the sorted state is
    // then still the current one).  Synchronises the stream once.
    uint32_t resample_phase(float merge_hi, float split_lo, float min_mass, float h, uint32_t cap, bool &changed);
    void finish_substep();
    // After an exception inside a substep: drop a stale launch error and
    // make the generation selectors match what the device holds.  Counts
    // left non-zero by an interrupted sort are cleared by the next sort.
    void recover();
    bool counts_dirty = false;  // a sort started and did not clear counts
    // Sorts run, and changes of the bucket / tile / buffer layout: a skin
    // list is usable while both are as they were when it was built.
    uint64_t sorts_done = 0, layout_epoch = 0;
    bool gen_open = false;      // generation cs rewritten, ct not flipped yet
    // adaptive resolution work arrays (resample.h documents them)
    DBuf<unsigned long long> rs_ctr;
    unsigned long long *rs_host = nullptr;  // pinned copy of rs_ctr
    DBuf<uint32_t> rs_flag, rs_items, rs_minT, rs_pick, rs_state, rs_cK, rs_cE, rs_rK, rs_rE, rs_slot, rs_partial;
    // First size of the merge-candidate store (0: the particle capacity).
    // The store grows when a pass overflows it; the environment variable
    // PHIENG_RESAMPLE_STORE_INITIAL sets a small first size so tests can
    // force that grow-and-rerun path.
    size_t rs_items_initial = 0;
    DBuf<uint4> rs_takers;
    DBuf<float4> rs_mv;
    // surrogate history (previous rho, P = p/rho^2, its change) per slot
    DBuf<float4> aux[2];
    bool aux_on = false;
    DBuf<uint32_t> perm;
    DBuf<float4> acc_sorted;
    DBuf<float4> tile_err;
    DBuf<float> err_out;
    // Boundary ghosts of scene 0 (single-world engines; fluid3.rs:716-920):
    // cell-sorted positions + mass, per-substep state, cell table.
    DBuf<float4> g_pos, g_state;
    DBuf<int4> g_key;
    DBuf<uint2> g_range;
    uint32_t g_count = 0, g_mask = 0;
    std::vector<uint32_t> g_order;  // cell-sorted slot -> generation index
    // Replace the ghosts (generation order, xyz + mass) keyed by their cells
    // at the kernel smoothing length h_kernel (rebuild_boundary_grid).
    void set_ghosts(const std::vector<float4> &gen, float h_kernel);
    GhostDev ghost_dev();
    // Asynchronous domain monitor: returns true if some scene changed domain.
    bool poll_monitor(uint32_t nparticles_hint);
    // Cell box of the current particles of scene k (synchronous).
    void particle_cell_box(uint32_t k, int32_t out[6]);
    float max_speed(uint32_t begin, uint32_t end);

    template <class Fn>
    void run(int kind, Fn &&fn) {
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            a = get_event();
            b = get_event();
            PH_CK(hipEventRecord(a, stream));
        }
        hipError_t e = fn();
        if (e != hipSuccess) throw HipError(kind_name(kind), e);
        if (timing) {
            PH_CK(hipEventRecord(b, stream));
            pending_ev.push_back({kind, a, b});
            if (pending_ev.size() > 8192) harvest_timing();
        }
    }
    hipEvent_t get_event();
    void harvest_timing();
    void reset_timing();
};

}  // namespace ph
