// batch3.hip — many independent worlds advanced together.
//
// The worlds' particles are concatenated into one device state (scene k
// owns slots [part_begin, part_end)), every scene keeps its own grid
// domain, config and static bodies, and one substep of all scenes runs
// either as the eight launches of a single world or as one persistent
// launch with grid barriers, optionally replayed from a hipGraph.  The
// per-particle arithmetic is the same device code in every mode, so a
// scene's trajectory is bit-identical to the world run alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/phieng_fluid3.h"
#include "engine.h"
#include "sph_kernels.h"
#include "world3_internal.h"

namespace {
thread_local std::string g_batch_error;
void set_err(const char *where, const char *what) { g_batch_error = std::string(where) + ": " + what; }
}  // namespace

struct Batch3 {
    std::vector<World3 *> worlds;
    std::vector<uint32_t> offset;  // first global slot / original-index base of each world
    ph::Engine eng;
    int mode = PHIENG_BATCH_PERSISTENT_GRAPH;
    uint32_t blocks_per_cu = 0;  // persistent grid = blocks_per_cu x CUs (0 = occupancy limit)
    uint32_t grid = 0;
    int num_cus = 0;
    int max_blocks_per_cu = 0;
    uint32_t spin_limit = 1u << 22;
    ph::DBuf<uint32_t> bar;
    hipGraphExec_t exec[2] = {nullptr, nullptr};  // one per tag generation (ct = 0 / 1)
    uint32_t exec_nbuckets[2] = {0, 0};
    int exec_cs[2] = {0, 0};  // the state generation a graph was captured on (Verlet substeps alternate it)
    bool abort_seen = false;
    float dt_uploaded = -1.0f;

    ~Batch3() {
        for (auto &e : exec)
            if (e) (void)hipGraphExecDestroy(e);
    }

    void drop_graphs() {
        for (auto &e : exec)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    }

    void set_grid() {
        const int bpc = blocks_per_cu ? std::min<int>((int)blocks_per_cu, max_blocks_per_cu) : max_blocks_per_cu;
        uint32_t g = (uint32_t)std::max(1, bpc) * (uint32_t)num_cus;
        g = std::min<uint32_t>(g, std::max<uint32_t>((uint32_t)eng.tiles.size(), 8u));
        grid = std::max<uint32_t>(8u, g & ~7u);  // a multiple of 8 keeps items on their XCD
    }

    ph::PersistArgs args() const {
        const ph::Engine &e = eng;
        const int ns = 1 - e.cs, nt = 1 - e.ct;
        const bool rc = e.rec_flags & PHIENG_RECORD_CELLS, rd = e.rec_flags & PHIENG_RECORD_DENSITY,
                   ra = e.rec_flags & PHIENG_RECORD_ACCEL;
        ph::PersistArgs a{};
        a.tiles = e.d_tiles.p;
        a.ntiles = (uint32_t)e.tiles.size();
        a.scenes = e.d_scenes.p;
        a.bodies = e.d_bodies.p;
        a.posm_cur = e.posm[e.cs].p;
        a.vel_cur = e.vel[e.cs].p;
        a.posm_alt = e.posm[ns].p;
        a.vel_alt = e.vel[ns].p;
        a.tag_cur = e.tag[e.ct].p;
        a.tag_alt = e.tag[nt].p;
        a.keys = e.keys.p;
        a.slots = e.slots.p;
        a.tmp = e.tmp.p;
        a.counts = e.counts.p;
        a.start = e.start.p;
        a.partial = e.partial.p;
        a.rq = e.rq.p;
        a.lcnt = e.nl_cnt.p;
        a.lidx = e.nl_idx.p;
        a.mstride = e.list_stride();
        a.tile_bbox = e.d_bbox.p;
        a.n = e.n_total;
        a.nbuckets = e.nbuckets;
        a.bar = bar.p;
        a.spin_limit = spin_limit;
        a.rec_cells = rc ? e.rec_cells.p : nullptr;
        a.rec_sorted = rc ? e.rec_sorted.p : nullptr;
        a.rec_rho = rd ? e.rec_rho.p : nullptr;
        a.rec_p = rd ? e.rec_p.p : nullptr;
        a.rec_count = rd ? e.rec_count.p : nullptr;
        a.rec_acc = ra ? e.rec_acc.p : nullptr;
        return a;
    }

    void launch_persistent() {
        const ph::PersistArgs a = args();
        eng.run(ph::KIND_PERSISTENT, [&] { return ph::launch_step_persistent(eng.stream, grid, a); });
    }

    // One substep of all scenes in the current mode.
    void substep() {
        ph::Engine &e = eng;
        if (e.n_total == 0) return;
        // Verlet lists in the eager multi-kernel form only; the captured and
        // persistent forms reorder particles without keeping the rebuild
        // slots, so the lists are invalid after them.
        e.vl_enabled = mode == PHIENG_BATCH_KERNELS;
        if (!e.vl_enabled) e.vl_valid = false;
        if (mode == PHIENG_BATCH_KERNELS || (mode == PHIENG_BATCH_KERNELS_GRAPH && (e.timing || e.rec_flags))) {
            e.substep();
            return;
        }
        if (mode == PHIENG_BATCH_KERNELS_GRAPH) {
            // the sort and the two passes of one substep, captured once per
            // tag parity (the tags alternate between two buffers) and
            // replayed; a sort cut short earlier is cleared outside the graph
            if (e.counts_dirty) {
                e.run(ph::KIND_FILL, [&] { return ph::launch_fill_u32(e.stream, e.counts.p, (uint32_t)e.alloc_buckets, 0u); });
                e.counts_dirty = false;
            }
            const int k = e.ct;
            if (!exec[k] || exec_nbuckets[k] != e.nbuckets || exec_cs[k] != e.cs) {
                if (exec[k]) {
                    (void)hipGraphExecDestroy(exec[k]);
                    exec[k] = nullptr;
                }
                hipGraph_t g = nullptr;
                PH_CK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
                try {
                    e.sort_phase(nullptr);
                    e.plain_phase(nullptr);
                } catch (...) {
                    (void)hipStreamEndCapture(e.stream, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hipError_t ce = hipStreamEndCapture(e.stream, &g);
                if (ce != hipSuccess) throw ph::HipError("hipStreamEndCapture", ce);
                hipError_t ie = hipGraphInstantiate(&exec[k], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) throw ph::HipError("hipGraphInstantiate", ie);
                exec_nbuckets[k] = e.nbuckets;
                exec_cs[k] = e.cs;
            }
            PH_CK(hipGraphLaunch(exec[k], e.stream));
            e.counts_dirty = false;
            e.gen_open = true;
            e.finish_substep();
            return;
        }
        // The host cannot see whether a persistent launch finished its scan
        // (a timed-out barrier ends it early), so the multi-kernel path
        // clears the bucket counts before its next sort.
        e.counts_dirty = true;
        if (mode == PHIENG_BATCH_PERSISTENT || e.timing || e.rec_flags) {
            launch_persistent();
        } else {
            const int k = e.ct;
            if (!exec[k] || exec_nbuckets[k] != e.nbuckets || exec_cs[k] != e.cs) {
                if (exec[k]) {
                    (void)hipGraphExecDestroy(exec[k]);
                    exec[k] = nullptr;
                }
                hipGraph_t g = nullptr;
                PH_CK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
                const ph::PersistArgs a = args();
                hipError_t le = ph::launch_step_persistent(e.stream, grid, a);
                hipError_t ce = hipStreamEndCapture(e.stream, &g);
                if (le != hipSuccess) throw ph::HipError("capture launch", le);
                if (ce != hipSuccess) throw ph::HipError("hipStreamEndCapture", ce);
                hipError_t ie = hipGraphInstantiate(&exec[k], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) throw ph::HipError("hipGraphInstantiate", ie);
                exec_nbuckets[k] = e.nbuckets;
                exec_cs[k] = e.cs;
            }
            PH_CK(hipGraphLaunch(exec[k], e.stream));
        }
        e.gen_open = true;
        e.finish_substep();
    }

    void step(float dt) {
        ph::Engine &e = eng;
        if (e.n_total == 0) return;
        (void)hipGetLastError();  // a failed call before this step has already reported
        if (e.poll_monitor(e.n_total)) {
            e.upload_tables();
            drop_graphs();
        }
        if (dt != dt_uploaded) {
            for (auto &s : e.scenes) s.dt = dt;
            e.scenes_dirty = true;
            e.upload_tables();
            dt_uploaded = dt;
        }
        substep();
    }

    bool check_abort() {
        uint32_t h[3] = {0, 0, 0};
        PH_CK(hipMemcpyAsync(h, bar.p, sizeof(h), hipMemcpyDeviceToHost, eng.stream));
        PH_CK(hipStreamSynchronize(eng.stream));
        if (h[2]) abort_seen = true;
        return abort_seen;
    }
};

namespace {

// Copy every world's particles into the batch (tags shifted by the
// world's offset) and build one scene per world.
void build_batch(Batch3 &b) {
    ph::Engine &e = b.eng;
    e.init_device();
    uint32_t total = 0;
    for (World3 *w : b.worlds) {
        b.offset.push_back(total);
        total += world3_count(w);
    }
    if ((size_t)total > (size_t)PH_ID_MASK + 1) throw std::runtime_error("batch exceeds 2^24 particles");
    e.n_total = 0;
    e.ensure_particles(std::max<uint32_t>(total, 1));
    e.scenes.assign(b.worlds.size(), SceneDev{});
    e.bodies.clear();
    for (size_t k = 0; k < b.worlds.size(); ++k) {
        World3 *w = b.worlds[k];
        const uint32_t n = world3_count(w);
        SceneDev &s = e.scenes[k];
        ph::fill_scene_physics(s, world3_scene_host(w, 0.0f));
        s.part_begin = b.offset[k];
        s.part_end = b.offset[k] + n;
        s.body_begin = (uint32_t)e.bodies.size();
        const std::vector<BodyDev> &bd = world3_bodies(w);
        s.body_count = (uint32_t)bd.size();
        e.bodies.insert(e.bodies.end(), bd.begin(), bd.end());
        if (n) {
            const float4 *wp;
            const float4 *wv;
            const uint32_t *wt;
            hipStream_t ws;
            world3_device_state(w, &wp, &wv, &wt, &ws);
            PH_CK(hipStreamSynchronize(ws));
            PH_CK(hipMemcpyAsync(e.posm[e.cs].p + b.offset[k], wp, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                 e.stream));
            PH_CK(hipMemcpyAsync(e.vel[e.cs].p + b.offset[k], wv, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                 e.stream));
            e.run(ph::KIND_FILL, [&] {
                return ph::launch_tag_shift(e.stream, n, wt, e.tag[e.ct].p + b.offset[k], (int32_t)b.offset[k]);
            });
        }
    }
    e.n_total = total;
    e.bodies_dirty = true;
    e.build_tiles();
    e.upload_tables();
    for (size_t k = 0; k < b.worlds.size(); ++k) {
        int32_t box[6];
        e.particle_cell_box((uint32_t)k, box);
        ph::choose_domain(e.scenes[k], box, e.scenes[k].part_end - e.scenes[k].part_begin,
                          world3_bodies(b.worlds[k]));
    }
    e.assign_buckets();
    e.upload_tables();
    b.bar.ensure(4);
    PH_CK(hipMemsetAsync(b.bar.p, 0, 4 * sizeof(uint32_t), e.stream));
    PH_CK(hipDeviceGetAttribute(&b.num_cus, hipDeviceAttributeMultiprocessorCount, e.device));
    PH_CK(ph::persistent_max_blocks_per_cu(&b.max_blocks_per_cu));
    if (b.max_blocks_per_cu < 1) throw std::runtime_error("persistent step kernel cannot be resident");
    b.set_grid();
    PH_CK(hipStreamSynchronize(e.stream));
}

std::string tune_key(const Batch3 &b) {
    std::ostringstream k;
    k << "gfx950:n=" << b.eng.n_total << ":scenes=" << b.worlds.size() << ":cus=" << b.num_cus;
    return k.str();
}

}  // namespace

extern "C" {

const char *phieng_batch3_last_error(void) { return g_batch_error.c_str(); }

Batch3 *phieng_batch3_new(World3 *const *worlds, uint32_t count) {
    if (!worlds || count == 0) return nullptr;
    Batch3 *b = nullptr;
    try {
        b = new Batch3();
        for (uint32_t k = 0; k < count; ++k) {
            if (!worlds[k]) throw std::runtime_error("null world");
            const FluidConfig3Ffi &c = world3_config(worlds[k]);
            if (c.cfl > 0.0f && c.max_substeps > 1)
                throw std::runtime_error("CFL substepping differs per scene: not supported in a batch");
            if (!c.enabled) throw std::runtime_error("a disabled fluid cannot join a batch");
            if (c.boundary_particles)
                throw std::runtime_error("boundary ghosts are per world: a world with boundary_particles cannot join a batch");
            if (c.adaptive_resolution)
                throw std::runtime_error("adaptive resolution changes a world's particle count: it cannot join a batch");
            b->worlds.push_back(worlds[k]);
        }
        build_batch(*b);
        return b;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_new", ex.what());
        delete b;
        return nullptr;
    }
}

void phieng_batch3_free(Batch3 *batch) {
    try {
        delete batch;
    } catch (...) {
    }
}

int32_t phieng_batch3_set_mode(Batch3 *b, int32_t mode, uint32_t blocks_per_cu) {
    if (!b || mode < PHIENG_BATCH_KERNELS || mode > PHIENG_BATCH_KERNELS_GRAPH) return 0;
    try {
        b->mode = mode;
        b->blocks_per_cu = blocks_per_cu;
        b->set_grid();
        b->drop_graphs();
        b->eng.vl_valid = false;
        // A timed-out grid barrier leaves its arrival count and abort flag
        // set; a new mode starts from clear barrier words (a fill kernel on
        // the stream, outside any captured graph).
        // A launch that left at a barrier may also have counted keys into the
        // bucket counts without the scan that re-zeroes them.
        ph::Engine &e = b->eng;
        e.run(ph::KIND_FILL, [&] { return ph::launch_fill_u32(e.stream, b->bar.p, 4u, 0u); });
        if (e.counts.p)
            e.run(ph::KIND_FILL, [&] { return ph::launch_fill_u32(e.stream, e.counts.p, (uint32_t)e.alloc_buckets, 0u); });
        e.counts_dirty = false;
        PH_CK(hipStreamSynchronize(e.stream));
        b->abort_seen = false;
        return 1;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_set_mode", ex.what());
        return 0;
    }
}

int32_t phieng_batch3_set_spin_limit(Batch3 *b, uint32_t spins) {
    if (!b) return 0;
    b->spin_limit = spins;
    b->drop_graphs();  // captured launches carry the old limit
    return 1;
}

int32_t phieng_batch3_step(Batch3 *b, float dt) {
    if (!b || !std::isfinite(dt) || dt <= 0.0f) return 0;
    if (b->abort_seen) {
        set_err("phieng_batch3_step", "a grid barrier timed out earlier; call phieng_batch3_set_mode to recover");
        return 0;
    }
    try {
        b->step(std::min(dt, 1.0f / 15.0f));
        return 1;
    } catch (const std::exception &ex) {
        b->eng.recover();
        set_err("phieng_batch3_step", ex.what());
        return 0;
    }
}

int32_t phieng_batch3_sync(Batch3 *b) {
    if (!b) return 0;
    try {
        if (b->check_abort()) {
            set_err("phieng_batch3_sync", "a grid barrier timed out (blocks not co-resident); results invalid");
            return 0;
        }
        return 1;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_sync", ex.what());
        return 0;
    }
}

// Write the batch state back into its worlds (their fluid_get then reads it).
int32_t phieng_batch3_sync_worlds(Batch3 *b) {
    if (!b) return 0;
    try {
        if (b->check_abort()) {
            set_err("phieng_batch3_sync_worlds", "a grid barrier timed out; results invalid");
            return 0;
        }
        ph::Engine &e = b->eng;
        for (size_t k = 0; k < b->worlds.size(); ++k) {
            const SceneDev &s = e.scenes[k];
            const uint32_t n = s.part_end - s.part_begin;
            if (!n) continue;
            float4 *wp;
            float4 *wv;
            uint32_t *wt;
            hipStream_t ws;
            world3_device_state_mut(b->worlds[k], &wp, &wv, &wt, &ws);
            PH_CK(hipMemcpyAsync(wp, e.posm[e.cs].p + s.part_begin, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                 e.stream));
            PH_CK(hipMemcpyAsync(wv, e.vel[e.cs].p + s.part_begin, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                 e.stream));
            e.run(ph::KIND_FILL, [&] {
                return ph::launch_tag_shift(e.stream, n, e.tag[e.ct].p + s.part_begin, wt, -(int32_t)s.part_begin);
            });
        }
        PH_CK(hipStreamSynchronize(e.stream));
        return 1;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_sync_worlds", ex.what());
        return 0;
    }
}

int32_t phieng_batch3_set_timing(Batch3 *b, int32_t enable) {
    if (!b) return 0;
    try {
        b->eng.reset_timing();
        b->eng.timing = enable != 0;
        return 1;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_set_timing", ex.what());
        return 0;
    }
}

int32_t phieng_batch3_get_timing(Batch3 *b, double *total_ms, uint64_t *launches, int32_t max) {
    if (!b) return 0;
    try {
        b->eng.harvest_timing();
        for (int32_t k = 0; k < std::min<int32_t>(max, ph::KIND_COUNT); ++k) {
            if (total_ms) total_ms[k] = b->eng.total_ms[k];
            if (launches) launches[k] = b->eng.launches[k];
        }
        return ph::KIND_COUNT;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_get_timing", ex.what());
        return 0;
    }
}

// Autotuner: time each launch configuration for `trial_steps` substeps from
// the same saved state, keep the fastest, restore the state, and remember
// the choice in `cache_path` (one "key mode blocks_per_cu" line per shape).
// Every configuration runs the same device arithmetic, so the choice
// changes speed only.  Returns the chosen mode, or -1 on error.
int32_t phieng_batch3_autotune(Batch3 *b, float dt, uint32_t trial_steps, const char *cache_path,
                               uint32_t *out_blocks_per_cu, double *out_ms_per_step) {
    if (!b || !std::isfinite(dt) || dt <= 0.0f) return -1;
    try {
        const std::string key = tune_key(*b);
        if (cache_path) {
            std::ifstream in(cache_path);
            std::string k;
            int m;
            uint32_t bpc;
            double ms;
            while (in >> k >> m >> bpc >> ms) {
                if (k == key) {
                    phieng_batch3_set_mode(b, m, bpc);
                    if (out_blocks_per_cu) *out_blocks_per_cu = bpc;
                    if (out_ms_per_step) *out_ms_per_step = ms;
                    return m;
                }
            }
        }
        ph::Engine &e = b->eng;
        const uint32_t n = e.n_total;
        // save state
        ph::DBuf<float4> sp, sv;
        ph::DBuf<uint32_t> st;
        sp.ensure(n);
        sv.ensure(n);
        st.ensure(n);
        PH_CK(hipMemcpyAsync(sp.p, e.posm[e.cs].p, sizeof(float4) * n, hipMemcpyDeviceToDevice, e.stream));
        PH_CK(hipMemcpyAsync(sv.p, e.vel[e.cs].p, sizeof(float4) * n, hipMemcpyDeviceToDevice, e.stream));
        PH_CK(hipMemcpyAsync(st.p, e.tag[e.ct].p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, e.stream));
        const int ct0 = e.ct;
        std::vector<std::pair<int, uint32_t>> cands = {{PHIENG_BATCH_KERNELS, 0}, {PHIENG_BATCH_KERNELS_GRAPH, 0}};
        for (int bpc = 1; bpc <= b->max_blocks_per_cu; ++bpc) cands.push_back({PHIENG_BATCH_PERSISTENT_GRAPH, bpc});
        hipEvent_t t0, t1;
        PH_CK(hipEventCreate(&t0));
        PH_CK(hipEventCreate(&t1));
        double best = 1e30;
        std::pair<int, uint32_t> pick = cands.back();
        for (auto &c : cands) {
            phieng_batch3_set_mode(b, c.first, c.second);
            for (int w = 0; w < 2; ++w) b->step(dt);
            PH_CK(hipEventRecord(t0, e.stream));
            for (uint32_t s = 0; s < trial_steps; ++s) b->step(dt);
            PH_CK(hipEventRecord(t1, e.stream));
            PH_CK(hipEventSynchronize(t1));
            float ms = 0;
            PH_CK(hipEventElapsedTime(&ms, t0, t1));
            if (b->check_abort()) throw std::runtime_error("grid barrier timed out during tuning");
            const double per = (double)ms / std::max<uint32_t>(trial_steps, 1);
            if (per < best) {
                best = per;
                pick = c;
            }
            // restore
            PH_CK(hipMemcpyAsync(e.posm[e.cs].p, sp.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, e.stream));
            PH_CK(hipMemcpyAsync(e.vel[e.cs].p, sv.p, sizeof(float4) * n, hipMemcpyDeviceToDevice, e.stream));
            PH_CK(hipMemcpyAsync(e.tag[ct0].p, st.p, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, e.stream));
            e.ct = ct0;
            e.vl_valid = false;  // the Verlet rows described the trial's state
        }
        PH_CK(hipStreamSynchronize(e.stream));
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        phieng_batch3_set_mode(b, pick.first, pick.second);
        if (cache_path) {
            std::ofstream out(cache_path, std::ios::app);
            out << key << " " << pick.first << " " << pick.second << " " << best << "\n";
        }
        if (out_blocks_per_cu) *out_blocks_per_cu = pick.second;
        if (out_ms_per_step) *out_ms_per_step = best;
        return pick.first;
    } catch (const std::exception &ex) {
        set_err("phieng_batch3_autotune", ex.what());
        return -1;
    }
}

uint32_t phieng_batch3_count(const Batch3 *b) { return b ? b->eng.n_total : 0; }

}  // extern "C"
