// sph_kernels.h — host launchers of the SPH kernels (sph_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sph_device.h"

namespace ph {

// Arguments of the persistent one-launch substep (all scenes).
struct PersistArgs {
    const TileDev *tiles;
    uint32_t ntiles;
    const SceneDev *scenes;
    const BodyDev *bodies;
    float4 *posm_cur, *vel_cur, *posm_alt, *vel_alt;
    uint32_t *tag_cur, *tag_alt;
    uint32_t *keys, *slots, *tmp, *counts, *start, *partial;
    float2 *rq;
    uint32_t *lcnt, *lidx;  // the neighbour list between the two passes (ListOut)
    uint32_t mstride;
    int32_t *tile_bbox;
    uint32_t n, nbuckets;
    uint32_t *bar;  // [arrivals, generation, abort]
    uint32_t spin_limit;
    int32_t *rec_cells;
    uint32_t *rec_sorted, *rec_count;
    float *rec_rho, *rec_p, *rec_acc;
};

uint32_t scan_partials_needed(uint32_t nbuckets);

// vs (may be null): a Verlet substep, whose rebuild decision the kernel
// takes (vst) and records in vs->rebuild for the substep's later kernels;
// the neighbour words go to vword and the bucket keys are made only when
// it rebuilds.  pacc (may be null; zero on entry): the scan's block sums
// (scan_partials_needed entries) are added up there, in place of
// launch_scan_reduce.
hipError_t launch_cell_key(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                           const float4 *posm, const uint32_t *tag, uint32_t *keys, uint32_t *slots,
                           uint32_t *counts, int32_t *rec_cells, VerletDev *vs = nullptr,
                           const VlStep &vst = VlStep{}, uint32_t *vword = nullptr, uint32_t *pacc = nullptr);
// gate (may be null): the kernel runs only when *gate is nonzero.
hipError_t launch_scan_reduce(hipStream_t st, const uint32_t *counts, uint32_t nb, uint32_t *partial,
                              const uint32_t *gate = nullptr);
hipError_t launch_scan_partials(hipStream_t st, uint32_t *partial, uint32_t nb, uint32_t *start,
                                const uint32_t *gate = nullptr);
hipError_t launch_scan_apply(hipStream_t st, uint32_t *counts, uint32_t nb, const uint32_t *partial,
                             uint32_t *start, const uint32_t *gate = nullptr);
// scan_apply with the block offsets summed from the raw partials (no
// scan_partials launch), start[nb] = total.
hipError_t launch_scan_apply_sum(hipStream_t st, uint32_t *counts, uint32_t nb, const uint32_t *partial,
                                 uint32_t *start, const uint32_t *gate = nullptr);
// zero (may be null): nzero words set to zero when the kernel runs (the
// block sums the cell-key pass accumulated, once the scan has read them).
hipError_t launch_scatter(hipStream_t st, uint32_t n, const uint32_t *keys, const uint32_t *slots,
                          const uint32_t *start, uint32_t *tmp, const uint32_t *gate = nullptr,
                          uint32_t *zero = nullptr, uint32_t nzero = 0);
hipError_t launch_rank_gather(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                              const uint32_t *tmp, const uint32_t *keys, const uint32_t *start,
                              const float4 *posm_in, const float4 *vel_in, const uint32_t *tag_in,
                              float4 *posm_out, float4 *vel_out, uint32_t *tag_out, uint32_t *rec_sorted,
                              uint32_t *perm_out = nullptr, const VerletDev *vs = nullptr,
                              const uint32_t *vword_in = nullptr, uint32_t *vword_out = nullptr);
// The force pass of a Verlet substep (parity vl_par): the list-based pass
// over the neighbour list the Verlet density pass wrote, also taking the
// largest squared displacement (vs->dmax2[vl_par]), flagging positions
// beyond the reach of the mod-4 cell words (vs->slow[vl_par]), and writing
// each slot's next neighbour word (vword_next).
hipError_t launch_force_verlet(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                               const BodyDev *bodies, const uint32_t *start, const float4 *posm, const float4 *vel,
                               const float2 *rq, const uint32_t *tag, const uint32_t *lcnt, const uint32_t *lidx,
                               uint32_t mstride, float4 *posm_out, float4 *vel_out, float *rec_acc, int32_t *bbox,
                               const GhostDev &gh, VerletDev *vs, uint32_t *vword_next, const uint32_t *vhalo,
                               const uint32_t *vhalo_n, uint32_t vl_par);
// Verlet substep (verlet.hip).  Its first kernel (k_cell_key) decides
// whether this substep sorts and rebuilds every row: on request (forced),
// after a particle needed the per-cell path, or when the displacement
// bound (the substeps' largest displacements added up) passes thr.  The
// density pass then either sweeps the cells within (1 + skin) h of every
// particle, writing its row, or walks its row; a wave puts rows whose
// neighbours within h are out of visiting order back in order itself.
struct VerletArgs {
    const uint32_t *vword;   // neighbour word of each slot (this substep's order)
    uint32_t *vcnt;          // per row: entry count (| VLF_SLOW)
    uint32_t *vidx;          // rows: group g of slot i at uint4 g * vstride + i
    uint32_t vstride;
    float f2;                // (1 + skin)^2
    float skin;
    VerletDev *vs;           // the walk adds the rows it put in order to vs->qsum[par]
    uint32_t par;            // the substep's parity (VerletDev)
    uint32_t *tag_mirror;    // a rebuild copies the sorted tags here (the other tag generation)
    // halo form of the rows: per tile the sorted slots of its rows' union
    // (halo[tile * VL_HALO_CAP + k], count halo_n[tile] or VL_HALO_NONE),
    // and the rows as 16-bit halo indices, eight to a 16-byte group (group g
    // of slot i at rows16[g * vstride + i])
    uint32_t *halo;
    uint32_t *halo_n;
    uint4 *rows16;
    // the neighbours within h, for the force pass (ListOut layout: group g
    // of slot i at uint4 g * lstride + i; the count word in fcnt)
    uint32_t *lidx;
    uint32_t lstride;
    // a rebuild's density pass copies the sorted state (the pass's posm, and
    // sorted_vel) into the generation the walk and the force pass read
    float4 *keep_posm;
    const float4 *sorted_vel;
    float4 *keep_vel;
};
// posm_sorted: the sort's output (the rebuild reads it); posm: the state
// the walk reads (va.keep_posm, where the rebuild copies the sorted state).
hipError_t launch_density_verlet(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                                 const uint32_t *start, const float4 *posm_sorted, const float4 *posm,
                                 const uint32_t *tag, float2 *rq, float *rec_rho, float *rec_p,
                                 uint32_t *rec_count, uint32_t *fcnt, const VerletArgs &va);
// Skin list written by the density pass for surrogate substeps.
struct ListOut {
    int4 *cell = nullptr;
    uint32_t *cnt = nullptr;
    uint32_t *idx = nullptr;  // SL_CAP / 4 groups of 4 slots, group g of particle i at uint4 g * mstride + i
    float skin_factor2 = 1.0f;              // ((1 + skin))^2, or
    const float *skin_factor2_dev = nullptr;  // the same computed on the device (k_list_radius)
    uint32_t *slow = nullptr;  // non-null: counts the particles whose list did not fit (SLF_SLOW)
};
// List radius of a surrogate world from the current speeds (a Verlet skin):
// out = (1 + max(skin_min, coef * vmax))^2 with vmax the largest speed
// (per tile into tile_vmax, then over tiles) and coef = 4 x the substeps a
// list serves x dt / h: two particles approach at most at twice the
// largest speed, and speeds are allowed to double before the next list.
// vmax_out (may be null) receives vmax.
hipError_t launch_list_radius(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                              const float4 *vel, float *tile_vmax, float coef, float skin_min, float *out,
                              float *vmax_out = nullptr);
// Density of every slot from a kept skin list (ListOut cnt/idx of an
// earlier density pass over the same particle order): the entries within
// h in list order, then the density pass's clamp, equation of state and
// records.  A slot whose list did not fit (SLF_SLOW) sweeps the cells of
// the table in start.
hipError_t launch_density_skin(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                               const uint32_t *start, const float4 *posm, const uint32_t *tag, const uint32_t *lcnt,
                               const uint32_t *lidx, uint32_t mstride, float2 *rq, float *rec_rho, float *rec_p,
                               uint32_t *rec_count);
hipError_t launch_density_list(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                               const uint32_t *start, const float4 *posm, const uint32_t *tag, float2 *rq,
                               uint32_t mstride, float *rec_rho, float *rec_p, uint32_t *rec_count, const ListOut &lo);
// The force pass over the density pass's neighbour list (ListOut cnt/idx).
hipError_t launch_force_list(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                             const BodyDev *bodies, const uint32_t *start, const float4 *posm, const float4 *vel,
                             const float2 *rq, const uint32_t *tag, const uint32_t *lcnt, const uint32_t *lidx,
                             uint32_t mstride, float4 *posm_out, float4 *vel_out, float *rec_acc, int32_t *bbox,
                             float4 *acc_sorted, const GhostDev &gh, const AuxFuse &af = AuxFuse{});
// Boundary ghosts' density, pressure and q from the gathered fluid state.
hipError_t launch_ghost_density(hipStream_t st, const SceneDev *scenes, uint32_t scene, const uint32_t *start,
                                const float4 *posm, const GhostDev &gh, uint32_t count);
hipError_t launch_vmax(hipStream_t st, uint32_t n, const float4 *vel, uint32_t *out_bits);
hipError_t launch_export(hipStream_t st, uint32_t n, const float4 *posm, const float4 *vel, const uint32_t *tag,
                         uint32_t id_base, float *out6);
hipError_t launch_set_velocity(hipStream_t st, uint32_t n, const uint32_t *tag, const float *vel3,
                               float4 *vel);
hipError_t launch_append(hipStream_t st, uint32_t n0, uint32_t k, const float *xyz, const uint8_t *phase,
                         float4 masses, float4 *posm, float4 *vel, uint32_t *tag);
hipError_t launch_cell_bbox(hipStream_t st, uint32_t n, const float4 *posm, float h, int32_t *bb);
hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v);
hipError_t launch_monitor(hipStream_t st, const SceneDev *scenes, uint32_t nscenes, const uint32_t *start,
                          const int32_t *bbox, int32_t *out);

hipError_t launch_step_persistent(hipStream_t st, uint32_t blocks, const PersistArgs &a);
hipError_t persistent_max_blocks_per_cu(int *out);
hipError_t launch_tag_shift(hipStream_t st, uint32_t n, const uint32_t *src, uint32_t *dst, int32_t delta);

}  // namespace ph
