// sph_kernels.hip — HIP kernels of the SPH fluid step for gfx950 (CDNA4).
//
// One substep (crates/dynamics/src/fluid3.rs:990-1184) is
//   K1 cell_key      cell triple -> z-major bucket, slot = atomic count
//   K2 scan (3)      exclusive scan of bucket counts -> bucket start table
//   K3 scatter       slot table: sorted position (unranked) -> source slot
//   K3b rank_gather  rank inside each bucket by original index (reference
//                    bucket order, fluid3.rs:301-314), gather state
//   K4 density       rho, p, p/rho^2 per particle (fluid3.rs:639-655)
//   K5 force_integ.  pressure force + XSPH + integrate + boundary clamp
//                    (fluid3.rs:1029-1179, Jacobi: reads pre-step state)
// All per-particle kernels run over tiles of PH_TILE consecutive slots of
// one scene, so a batch of scenes uses the same kernels as one world.
#include "sph_kernels.h"

#include <hip/hip_runtime.h>

namespace {

constexpr int kScanItems = 8;
constexpr int kScanBlock = 256;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// Exclusive scan across a 256-thread block; returns the thread's prefix
// and the block total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t &total) {
    __shared__ uint32_t wsum[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; ++k) off += wsum[k];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + inc - v;
}

}  // namespace

// ---------------------------------------------------------------- K1
// Bucket key per particle and its slot inside the bucket.  Particles come
// in last substep's sorted order, so neighbouring lanes mostly share a
// bucket: one atomic per run of equal keys in a wave, not one per lane.
__device__ __forceinline__ void dev_cell_key(uint32_t bid, uint32_t nbl, const TileDev *__restrict__ tiles,
                                                      const SceneDev *__restrict__ scenes,
                                                      const float4 *__restrict__ posm,
                                                      const uint32_t *__restrict__ tag,
                                                      uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ slots,
                                                      uint32_t *__restrict__ counts,
                                                      int32_t *__restrict__ rec_cells,
                                                      uint32_t *__restrict__ pacc = nullptr) {
    const TileDev t = tiles[xcd_tile(bid, nbl)];
    const SceneDev s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    const bool active = i < s.part_end;
    const int lane = threadIdx.x & 63;
    uint32_t key = 0xFFFFFFFFu;
    int32_t cx = 0, cy = 0, cz = 0;
    if (active) {
        const float4 p = posm[i];
        cell_of(p.x, p.y, p.z, s.h_grid, cx, cy, cz);
        uint32_t lk;
        if (!finite3(p.x, p.y, p.z)) {
            lk = s.ncells + 1;  // left out of the grid (fluid3.rs:302-304)
        } else if (in_domain(s, cx, cy, cz)) {
            lk = (uint32_t)((((int64_t)cz - s.z0) * s.ny + ((int64_t)cy - s.y0)) * s.nx + ((int64_t)cx - s.x0));
        } else {
            lk = s.ncells;  // outlier bucket
        }
        key = s.bucket_base + lk;
    }
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || key != prev;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = ~0ull >> (63 - lane);
    const int run0 = 63 - __clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int run1 = after ? __ffsll((long long)after) - 1 : 64;
    uint32_t base = 0;
    if (active && head) base = atomicAdd(&counts[key], (uint32_t)(run1 - run0));
    base = __shfl(base, run0, 64);
    if (pacc) {
        // the scan's block sums (kScanBlock * kScanItems buckets each),
        // one atomic per run of lanes in one scan block (the wave's keys
        // are nearly sorted: one or two runs)
        const uint32_t pb = key / (uint32_t)(kScanBlock * kScanItems);
        const bool bhead = lane == 0 || pb != __shfl_up(pb, 1, 64);
        const unsigned long long bheads = __ballot(bhead);
        const int b0 = 63 - __clzll(bheads & upto);
        const unsigned long long bafter = bheads & ~upto;
        const int b1 = bafter ? __ffsll((long long)bafter) - 1 : 64;
        if (active && bhead) atomicAdd(&pacc[pb], (uint32_t)(b1 - b0));
    }
    if (!active) return;
    keys[i] = key;
    slots[i] = base + (uint32_t)(lane - run0);
    if (rec_cells) {
        const uint32_t id = tag[i] & PH_ID_MASK;
        rec_cells[3 * id] = cx;
        rec_cells[3 * id + 1] = cy;
        rec_cells[3 * id + 2] = cz;
    }
}

// With a Verlet list (vs non-null) this is the substep's first kernel: every
// block takes the substep's rebuild decision (vl_decide), one thread
// records it for the later kernels, and the kernel works only on a substep
// that sorts: every particle's neighbour word (original index, cell mod 4)
// into vword, then the bucket keys and counts.  (The cell record forces a
// sort on every substep.)
__global__ __launch_bounds__(PH_TILE) void k_cell_key(const TileDev *__restrict__ tiles,
                                                      const SceneDev *__restrict__ scenes,
                                                      const float4 *__restrict__ posm,
                                                      const uint32_t *__restrict__ tag,
                                                      uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ slots,
                                                      uint32_t *__restrict__ counts,
                                                      int32_t *__restrict__ rec_cells, VerletDev *vs, VlStep vst,
                                                      uint32_t *__restrict__ vword, uint32_t *__restrict__ pacc) {
    if (vs) {
        float bound;
        const bool rb = vl_decide(vs, vst, bound);
        if (blockIdx.x == 0 && threadIdx.x == 0) vl_commit(vs, vst, rb, bound);
        // between rebuilds the force pass has written the words already
        if (!rb) return;
        const TileDev t = tiles[xcd_tile(blockIdx.x, gridDim.x)];
        const SceneDev &s = scenes[t.scene];
        const uint32_t i = t.first + threadIdx.x;
        if (i < s.part_end) {
            const float4 p = posm[i];
            int32_t cx, cy, cz;
            cell_of(p.x, p.y, p.z, s.h_grid, cx, cy, cz);
            vword[i] = vl_word(tag[i], cx, cy, cz);
        }
    }
    dev_cell_key(blockIdx.x, gridDim.x, tiles, scenes, posm, tag, keys, slots, counts, rec_cells, pacc);
}

// ---------------------------------------------------------------- K2
__device__ __forceinline__ void dev_scan_reduce(uint32_t bid, uint32_t nbl, const uint32_t *__restrict__ counts,
                                                            uint32_t nb, uint32_t *__restrict__ partial) {
    const uint32_t base = bid * (kScanBlock * kScanItems) + threadIdx.x * kScanItems;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < nb) v += counts[base + k];
    uint32_t tot;
    block_excl_scan(v, tot);
    if (threadIdx.x == 0) partial[bid] = tot;
}

// gate (may be null): run only when *gate is nonzero (a Verlet substep that
// sorts); the launch sequence stays the same either way.
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t *__restrict__ counts,
                                                            uint32_t nb, uint32_t *__restrict__ partial,
                                                            const uint32_t *gate) {
    if (gate && !*gate) return;
    dev_scan_reduce(blockIdx.x, gridDim.x, counts, nb, partial);
}

// Single block: exclusive scan of the block partials, start[nb] = total.
__device__ __forceinline__ void dev_scan_partials(uint32_t bid, uint32_t nbl, uint32_t *__restrict__ partial, uint32_t np,
                                                              uint32_t *__restrict__ start, uint32_t nb) {
    uint32_t carry = 0;
    for (uint32_t c = 0; c < np; c += kScanBlock) {
        const uint32_t k = c + threadIdx.x;
        const uint32_t v = k < np ? partial[k] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan(v, tot);
        if (k < np) partial[k] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) start[nb] = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_partials(uint32_t *__restrict__ partial, uint32_t np,
                                                              uint32_t *__restrict__ start, uint32_t nb,
                                                              const uint32_t *gate) {
    if (gate && !*gate) return;
    dev_scan_partials(blockIdx.x, gridDim.x, partial, np, start, nb);
}

// Bucket start = block offset + in-block exclusive scan; counts are zeroed
// for the next substep's atomics.
__device__ __forceinline__ void dev_scan_apply(uint32_t bid, uint32_t nbl, uint32_t *__restrict__ counts, uint32_t nb,
                                                           const uint32_t *__restrict__ partial,
                                                           uint32_t *__restrict__ start) {
    const uint32_t base = bid * (kScanBlock * kScanItems) + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        c[k] = base + k < nb ? counts[base + k] : 0u;
        v += c[k];
    }
    uint32_t tot;
    uint32_t off = block_excl_scan(v, tot) + partial[bid];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < nb) {
            start[base + k] = off;
            counts[base + k] = 0u;
        }
        off += c[k];
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(uint32_t *__restrict__ counts, uint32_t nb,
                                                           const uint32_t *__restrict__ partial,
                                                           uint32_t *__restrict__ start, const uint32_t *gate) {
    if (gate && !*gate) return;
    dev_scan_apply(blockIdx.x, gridDim.x, counts, nb, partial, start);
}

// The same with the block's offset summed here from the raw block
// partials (a few hundred words: no separate scan of them), and the last
// block writing start[nb] = total.
__global__ __launch_bounds__(kScanBlock) void k_scan_apply_sum(uint32_t *__restrict__ counts, uint32_t nb,
                                                               const uint32_t *__restrict__ partial,
                                                               uint32_t *__restrict__ start, const uint32_t *gate) {
    if (gate && !*gate) return;
    const uint32_t bid = blockIdx.x;
    uint32_t s = 0;
    for (uint32_t k = threadIdx.x; k < bid; k += kScanBlock) s += partial[k];
    uint32_t off0;
    block_excl_scan(s, off0);
    const uint32_t base = bid * (kScanBlock * kScanItems) + threadIdx.x * kScanItems;
    uint32_t c[kScanItems];
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        c[k] = base + k < nb ? counts[base + k] : 0u;
        v += c[k];
    }
    uint32_t tot;
    uint32_t off = block_excl_scan(v, tot) + off0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < nb) {
            start[base + k] = off;
            counts[base + k] = 0u;
        }
        off += c[k];
    }
    if (bid + 1 == gridDim.x && threadIdx.x == 0) start[nb] = off0 + tot;
}

// ---------------------------------------------------------------- K3
__device__ __forceinline__ void dev_scatter(uint32_t bid, uint32_t nbl, uint32_t n, const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ slots,
                                                     const uint32_t *__restrict__ start,
                                                     uint32_t *__restrict__ tmp) {
    const uint32_t i = bid * PH_TILE + threadIdx.x;
    if (i >= n) return;
    // in range whenever the counts started at zero; the test keeps stale
    // counts (a sort cut short) from writing outside the table
    const uint32_t d = start[keys[i]] + slots[i];
    if (d < n) tmp[d] = i;
}

__global__ __launch_bounds__(PH_TILE) void k_scatter(uint32_t n, const uint32_t *__restrict__ keys,
                                                     const uint32_t *__restrict__ slots,
                                                     const uint32_t *__restrict__ start,
                                                     uint32_t *__restrict__ tmp, const uint32_t *gate,
                                                     uint32_t *__restrict__ zero, uint32_t nzero) {
    if (gate && !*gate) return;
    // block sums the cell-key pass accumulated, read by the scan before
    // this kernel: zero again for the next sort
    if (zero && blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < nzero; k += PH_TILE) zero[k] = 0u;
    dev_scatter(blockIdx.x, gridDim.x, n, keys, slots, start, tmp);
}

// Final slot = bucket start + rank of the particle inside its bucket.  In
// a grid cell the rank orders by original index, which is the order the
// reference's bucket Vec holds (fluid3.rs:301-314); in the outlier bucket
// it orders by (cell z, y, x) first so every outlier cell is contiguous.
__device__ __forceinline__ void dev_rank_gather(uint32_t bid, uint32_t nbl, const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes, const uint32_t *__restrict__ tmp,
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ start, const float4 *__restrict__ posm_in,
    const float4 *__restrict__ vel_in, const uint32_t *__restrict__ tag_in, float4 *__restrict__ posm_out,
    float4 *__restrict__ vel_out, uint32_t *__restrict__ tag_out, uint32_t *__restrict__ rec_sorted,
    uint32_t *__restrict__ perm_out, const VerletDev *vs = nullptr, const uint32_t *__restrict__ vword_in = nullptr,
    uint32_t *__restrict__ vword_out = nullptr) {
    const TileDev t = tiles[xcd_tile(bid, nbl)];
    const SceneDev s = scenes[t.scene];
    const uint32_t q = t.first + threadIdx.x;
    if (q >= s.part_end) return;
    if (vs && !vs->rebuild) {
        // A Verlet substep that keeps the order: the same slots, and the
        // passes read the state where it is (both tag generations already
        // hold the sorted tags, the rebuild's density pass mirrors them, and
        // the force pass has written the next words).
        if (perm_out) perm_out[q] = q;
        return;
    }
    const uint32_t src = tmp[q];
    if (src < s.part_begin || src >= s.part_end) return;  // only after a sort cut short (see dev_scatter)
    const uint32_t key = keys[src];
    const uint32_t b = min(start[key], s.part_end), e = min(start[key + 1], s.part_end);
    const uint32_t mytag = tag_in[src];
    const uint32_t myid = mytag & PH_ID_MASK;
    uint32_t rank = 0;
    if (key - s.bucket_base == s.ncells) {
        const float4 mp = posm_in[src];
        int32_t mx, my, mz;
        cell_of(mp.x, mp.y, mp.z, s.h_grid, mx, my, mz);
        for (uint32_t k = b; k < e; ++k) {
            const uint32_t o = tmp[k];
            const float4 op = posm_in[o];
            int32_t ox, oy, oz;
            cell_of(op.x, op.y, op.z, s.h_grid, ox, oy, oz);
            const uint32_t oid = tag_in[o] & PH_ID_MASK;
            const bool less = oz != mz ? oz < mz : (oy != my ? oy < my : (ox != mx ? ox < mx : oid < myid));
            rank += less ? 1u : 0u;
        }
    } else {
        for (uint32_t k = b; k < e; ++k) rank += (tag_in[tmp[k]] & PH_ID_MASK) < myid ? 1u : 0u;
    }
    const uint32_t d = b + rank;
    posm_out[d] = posm_in[src];
    vel_out[d] = vel_in[src];
    tag_out[d] = mytag;
    if (rec_sorted) rec_sorted[d] = myid;
    if (perm_out) perm_out[d] = src;
    if (vword_out) vword_out[d] = vword_in[src];
}

__global__ __launch_bounds__(PH_TILE) void k_rank_gather(const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes, const uint32_t *__restrict__ tmp,
    const uint32_t *__restrict__ keys, const uint32_t *__restrict__ start, const float4 *__restrict__ posm_in,
    const float4 *__restrict__ vel_in, const uint32_t *__restrict__ tag_in, float4 *__restrict__ posm_out,
    float4 *__restrict__ vel_out, uint32_t *__restrict__ tag_out, uint32_t *__restrict__ rec_sorted, uint32_t *__restrict__ perm_out,
    const VerletDev *vs, const uint32_t *__restrict__ vword_in, uint32_t *__restrict__ vword_out) {
    dev_rank_gather(blockIdx.x, gridDim.x, tiles, scenes, tmp, keys, start, posm_in, vel_in, tag_in, posm_out, vel_out, tag_out, rec_sorted, perm_out,
                    vs, vword_in, vword_out);
}


// ---------------------------------------------------------------- K4
// Density (fluid3.rs:639-655) in two phases per row: the cut-off test over
// the row's candidates builds a bit mask, then the set bits are summed in
// ascending slot order, i.e. the reference's neighbour order.  Rows with
// more than 32 candidates finish their tail inline.  rq[i] = (rho_i,
// p_i / max(rho_i, EPS)^2), the pressure quotient of the force pass
// (fluid3.rs:1046-1051).  The neighbours are written to the particle's
// list in the same order, four slots per 16-byte store (ListOut documents
// the layout); the count word carries SLF_SLOW / SLF_NONE where the force
// pass must sweep the cells or has nothing to sweep.
// SKIN: the list radius is sqrt(skin_factor2) h > h (the surrogate's
//   Verlet list, from skin_factor2_dev when set); else the list is exactly
//   the neighbours and reuses the cut-off test's masks.  The density
//   results are the same either way.
template <bool SKIN>
__device__ __forceinline__ void dev_density(uint32_t bid, uint32_t nbl, const TileDev *__restrict__ tiles,
                                                     const SceneDev *__restrict__ scenes,
                                                     const uint32_t *__restrict__ start,
                                                     const float4 *__restrict__ posm,
                                                     const uint32_t *__restrict__ tag, float2 *__restrict__ rq,
                                                     uint32_t mstride, float *__restrict__ rec_rho,
                                                     float *__restrict__ rec_p, uint32_t *__restrict__ rec_count,
                                                     const ph::ListOut &lo) {
    const TileDev t = tiles[xcd_tile(bid, nbl)];
    const SceneDev s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    if (i >= s.part_end) return;
    const float4 pi = posm[i];
    float rho = 0.0f;
    uint32_t cnt = 0;
    uint32_t lk = 0;                          // skin list length
    // skin list entries are written four at a time: group g of this particle
    // (entries 4g .. 4g + 3) is one 16-byte store at lgrp[g * mstride]
    uint4 *lgrp = reinterpret_cast<uint4 *>(lo.idx) + i;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    auto push = [&](uint32_t j) {
        const uint32_t p = lk & 3u;
        q0 = p == 0u ? j : q0;
        q1 = p == 1u ? j : q1;
        q2 = p == 2u ? j : q2;
        q3 = p == 3u ? j : q3;
        if (p == 3u && lk < SL_CAP) lgrp[(size_t)(lk >> 2) * mstride] = make_uint4(q0, q1, q2, q3);
        ++lk;
    };
    const float hs2 = SKIN ? s.h_grid2 * (lo.skin_factor2_dev ? *lo.skin_factor2_dev : lo.skin_factor2) : 0.0f;
    if (finite3(pi.x, pi.y, pi.z)) {
        int32_t cx, cy, cz;
        cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
        if (lo.cell) lo.cell[i] = make_int4(cx, cy, cz, 0);
        const float hg2 = s.h_grid2;
        if (!needs_cell_path(s, start, cx, cy, cz)) {
#pragma unroll 1
            for (int r = 0; r < 9; ++r) {
                uint32_t br, er;
                row_range(s, start, cx, cy, cz, r, br, er);
                const uint32_t lim = min(er, br + PH_ROW_BITS);
                uint32_t m = 0, ml = 0;
                uint32_t j = br;
                // Four independent loads in flight per lane.
                for (; j + 4 <= lim; j += 4) {
                    const float4 *q = posm + j;  // one address, immediate offsets
                    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
                    const uint32_t o = j - br;
                    const float d0 = dot3(q0.x - pi.x, q0.y - pi.y, q0.z - pi.z);
                    const float d1 = dot3(q1.x - pi.x, q1.y - pi.y, q1.z - pi.z);
                    const float d2 = dot3(q2.x - pi.x, q2.y - pi.y, q2.z - pi.z);
                    const float d3 = dot3(q3.x - pi.x, q3.y - pi.y, q3.z - pi.z);
                    m |= ((d0 < hg2 ? 1u : 0u) | (d1 < hg2 ? 2u : 0u) | (d2 < hg2 ? 4u : 0u) | (d3 < hg2 ? 8u : 0u)) << o;
                    if (SKIN)
                        ml |= ((d0 < hs2 ? 1u : 0u) | (d1 < hs2 ? 2u : 0u) | (d2 < hs2 ? 4u : 0u) |
                               (d3 < hs2 ? 8u : 0u)) << o;
                }
                for (; j < lim; ++j) {
                    const float4 pj = posm[j];
                    const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                    m |= (r2 < hg2 ? 1u : 0u) << (j - br);
                    if (SKIN) ml |= (r2 < hs2 ? 1u : 0u) << (j - br);
                }
                if (!SKIN) ml = m;
                while (ml) {
                    const uint32_t k = __ffs(ml) - 1;
                    ml &= ml - 1;
                    push(br + k);
                }
                while (m) {
                    const uint32_t k = __ffs(m) - 1;
                    m &= m - 1;
                    const float4 pj = posm[br + k];
                    const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                    ++cnt;
                    rho += pj.w * poly6(r2, s);
                }
                if (er > lim) {
                    for (uint32_t j = lim; j < er; ++j) {
                        const float4 pj = posm[j];
                        const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                        if (r2 < hg2) {
                            ++cnt;
                            rho += pj.w * poly6(r2, s);
                        }
                        if (SKIN ? r2 < hs2 : r2 < hg2) push(j);
                    }
                }
            }
            // the last, partly filled group
            if ((lk & 3u) != 0u && lk < SL_CAP) lgrp[(size_t)(lk >> 2) * mstride] = make_uint4(q0, q1, q2, q3);
            lo.cnt[i] = lk <= SL_CAP ? lk : SLF_SLOW;
            if (lk > SL_CAP && lo.slow) atomicAdd(lo.slow, 1u);
        } else {
            lo.cnt[i] = SLF_SLOW;
            if (lo.slow) atomicAdd(lo.slow, 1u);
            for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) {
                const float4 pj = posm[j];
                const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                if (r2 < hg2) {
                    ++cnt;
                    rho += pj.w * poly6(r2, s);
                }
            });
        }
    } else {
        lo.cnt[i] = SLF_NONE;  // no neighbours: the force pass skips the sweep
    }
    rho = rclampf(rho, 0.0f, s.rho_max);
    const uint32_t tg = tag[i];
    // Indexed through global memory: a lane-varying index into the scene
    // copy would put the whole copy in scratch.
    const float rho0 = scenes[t.scene].phase_rho0[min(tg >> 24, 3u)];
    const float p = eos(rho, rho0, s.stiffness, s.gamma);
    const float rc = rmaxf(rho, PH_EPS);
    rq[i] = make_float2(rho, p / (rc * rc));
    if (rec_rho) {
        const uint32_t id = tg & PH_ID_MASK;
        rec_rho[id] = rho;
        rec_p[id] = p;
        rec_count[id] = cnt;
    }
}

template <bool SKIN>
__global__ __launch_bounds__(PH_TILE) void k_density_list(const TileDev *__restrict__ tiles,
                                                          const SceneDev *__restrict__ scenes,
                                                          const uint32_t *__restrict__ start,
                                                          const float4 *__restrict__ posm,
                                                          const uint32_t *__restrict__ tag, float2 *__restrict__ rq,
                                                          uint32_t mstride, float *__restrict__ rec_rho,
                                                          float *__restrict__ rec_p, uint32_t *__restrict__ rec_count,
                                                          ph::ListOut lo) {
    dev_density<SKIN>(blockIdx.x, gridDim.x, tiles, scenes, start, posm, tag, rq, mstride, rec_rho, rec_p, rec_count,
                      lo);
}

// ---------------------------------------------------------------- K5
// Per-pair force (fluid3.rs:1042-1085) and XSPH (fluid3.rs:1114-1123) terms,
// accumulated in the reference's operation order.
struct PairAcc {
    float fx = 0.0f, fy = 0.0f, fz = 0.0f;  // force
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;  // XSPH smoothing
};

// ex, ey, ez: x_i - x_j; r2 = |x_i - x_j|^2 (the same bits as the cut-off
// test's |x_j - x_i|^2: the differences are exact negations).  The caller
// has tested r2 < h_grid^2 <= h^2.
__device__ __forceinline__ void pair_terms(const SceneDev &s, const float4 &pi, const float4 &vi, float rho_i,
                                           float q_i, const float4 &pj, const float4 &vj, const float2 &rqj,
                                           float ex, float ey, float ez, float r2, PairAcc &a) {
    const float h = s.h;
    const float rho_j = rmax_pos(rqj.x, PH_EPS);
    const float dist = sqrt_pair(r2);
    float gx = 0.0f, gy = 0.0f, gz = 0.0f;
    if (!(dist >= h || dist < PH_EPS)) {  // spiky_grad, fluid3.rs:232-239
        const float sc = s.spiky_c * (h - dist) * (h - dist);
        // dist in [1e-6, h) with h <= 1024, and |e| <= dist: the three
        // quotients share one reciprocal unless a component is 0 or tiny
        if (h <= 1024.0f && fminf(fminf(fabsf(ex), fabsf(ey)), fabsf(ez)) >= 0x1p-100f) {
            const float y = div_recip(dist);
            gx = div_by_recip(ex, dist, y) * sc;
            gy = div_by_recip(ey, dist, y) * sc;
            gz = div_by_recip(ez, dist, y) * sc;
        } else {
            gx = ex / dist * sc;
            gy = ey / dist * sc;
            gz = ez / dist * sc;
        }
    }
    float term = q_i + rqj.y;
    if (s.av_alpha > 0.0f && dist > PH_EPS) {  // artificial_pressure, fluid3.rs:254-269
        const float rho_bar = 0.5f * (rho_i + rho_j);
        const float vd = (vi.x - vj.x) * ex + (vi.y - vj.y) * ey + (vi.z - vj.z) * ez;
        float ap = 0.0f;
        if (!(vd >= 0.0f)) {
            const float mu = h * vd / (dist * dist + 0.01f * h * h);
            ap = -s.av_alpha * s.c_s * mu / rmaxf(rho_bar, PH_EPS);
        }
        term = term + ap;
    }
    const float sm = pj.w * term;
    a.fx = a.fx - gx * sm;
    a.fy = a.fy - gy * sm;
    a.fz = a.fz - gz * sm;
    if (s.kappa > 0.0f && dist > PH_EPS) {  // cohesion, fluid3.rs:272-278, 1062-1070
        const float coh = dist >= h ? 0.0f : 1.0f - dist / h;
        if (coh > 0.0f) {
            const float cs = s.kappa * pi.w * pj.w * coh;
            a.fx = a.fx - ex / dist * cs;
            a.fy = a.fy - ey / dist * cs;
            a.fz = a.fz - ez / dist * cs;
        }
    }
    if (s.viscosity > 0.0f) {
        const float rho_bar = 0.5f * (rho_i + rho_j);
        const float qm = pj.w / rho_bar;
        const float w = poly6_within(r2, s);
        a.sx = a.sx + (vj.x - vi.x) * qm * w;
        a.sy = a.sy + (vj.y - vi.y) * qm * w;
        a.sz = a.sz + (vj.z - vi.z) * qm * w;
    }
}

// Force + XSPH over the neighbours the density pass found (same state,
// same test), then integrate, speed cap, finite guards and the boundary
// clamp of this particle (fluid3.rs:1088-1179, Jacobi).  The neighbours come
// from the density pass's neighbour list: the candidates within (1 + skin) h
// in reference order, four per 16-byte group (group g of slot i at
// lgrp[g * mstride + i]), re-tested here against h.  The list takes one loop
// over a particle's own entries instead of nine row loops whose trip counts
// differ between lanes.
// VL: a Verlet substep; the pass also records the substep's largest
// displacement for the next rebuild decision.
template <bool VL = false>
__device__ __forceinline__ void dev_force_integrate(uint32_t bid, uint32_t nbl, const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes, const BodyDev *__restrict__ bodies,
    const uint32_t *__restrict__ start, const float4 *__restrict__ posm, const float4 *__restrict__ vel,
    const float2 *__restrict__ rq, const uint32_t *__restrict__ tag, const uint32_t *__restrict__ lcnt,
    const uint4 *__restrict__ lgrp, uint32_t mstride, float4 *__restrict__ posm_out,
    float4 *__restrict__ vel_out, float *__restrict__ rec_acc, int32_t *__restrict__ bbox,
    float4 *__restrict__ acc_sorted, const GhostDev &gh = GhostDev{}, const AuxFuse &af = AuxFuse{},
    VerletDev *vs = nullptr, uint32_t *__restrict__ vword_next = nullptr, const uint32_t *__restrict__ vhalo = nullptr,
    const uint32_t *__restrict__ vhalo_n = nullptr, uint32_t vl_par = 0u) {
    const uint32_t tile = xcd_tile(bid, nbl);
    const TileDev t = tiles[tile];
    const SceneDev s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    const bool active = i < s.part_end;
    // A Verlet walk substep's tile with a halo hands over 16-bit halo
    // indices (eight to a group); their slots are staged here.
    __shared__ uint32_t f_s[VL ? VL_HALO_CAP : 1];
    bool l16 = false;
    if (VL && vs->rebuild == 0u) {
        const uint32_t H = vhalo_n[tile];
        if (H != VL_HALO_NONE) {  // the same for the whole block
            l16 = true;
            const uint32_t *hl = vhalo + (size_t)tile * VL_HALO_CAP;
            constexpr uint32_t kR = (VL_HALO_CAP + PH_TILE - 1) / PH_TILE;
            uint32_t js[kR];  // every load issued before the stores wait
#pragma unroll
            for (uint32_t q = 0; q < kR; ++q) {
                const uint32_t k = threadIdx.x + q * PH_TILE;
                js[q] = k < H ? hl[k] : 0u;
            }
#pragma unroll
            for (uint32_t q = 0; q < kR; ++q) {
                const uint32_t k = threadIdx.x + q * PH_TILE;
                if (k < H) f_s[k] = js[q];
            }
            __syncthreads();
        }
    }
    int32_t bx0 = INT32_MAX, by0 = INT32_MAX, bz0 = INT32_MAX, bx1 = INT32_MIN, by1 = INT32_MIN, bz1 = INT32_MIN;
    float jerr = 0.0f, jcnt = 0.0f, jae = 0.0f, jam = 0.0f;  // judge sums (AuxFuse)
    float dmax2 = 0.0f;               // largest squared displacement (VL)
    if (active) {
        const float4 pi = posm[i];
        const float4 vi = vel[i];
        const float2 rqi = rq[i];
        const float rho_i = rmaxf(rqi.x, PH_EPS);
        const float q_i = rqi.y;
        PairAcc acc;
        const float hg2 = s.h_grid2;
        {
            const uint32_t cw = lcnt[i];
            // one list entry: the self entry and those outside h add nothing
            auto entry = [&](uint32_t j, const float4 &pj, const float4 &vj, const float2 &qj) {
                const float ex = pi.x - pj.x, ey = pi.y - pj.y, ez = pi.z - pj.z;
                const float r2 = dot3(ex, ey, ez);
                // the neighbour's velocity and density in registers before the
                // branch (no instruction): the compiler would otherwise issue
                // their loads inside it, a second memory wait per group
                asm volatile("" ::"v"(vj.x), "v"(vj.y), "v"(vj.z), "v"(qj.x), "v"(qj.y));
                if (j != i && r2 < hg2) pair_terms(s, pi, vi, rho_i, q_i, pj, vj, qj, ex, ey, ez, r2, acc);
            };
            if (cw & SLF_NONE) {
                // non-finite position: no neighbours
            } else if (cw & SLF_SLOW) {
                int32_t cx, cy, cz;
                cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
                for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) { entry(j, posm[j], vel[j], rq[j]); });
            } else {
                const uint32_t n = cw & SLF_COUNT;
                const uint4 *g = lgrp + i;
                // byte offsets fit 32 bits (slots < 2^27): one base register
                // per array and one shift per slot and element size
                const char *pb = reinterpret_cast<const char *>(posm), *vb = reinterpret_cast<const char *>(vel),
                           *qb = reinterpret_cast<const char *>(rq);
                auto p4 = [](const char *b, uint32_t j) { return *reinterpret_cast<const float4 *>(b + (j << 4)); };
                auto q2l = [](const char *b, uint32_t j) { return *reinterpret_cast<const float2 *>(b + (j << 3)); };
                // four neighbours (slots past the count are replaced by i,
                // which is skipped)
                auto four = [&](uint32_t j0, uint32_t j1, uint32_t j2, uint32_t j3) {
                    const float4 p0 = p4(pb, j0), p1 = p4(pb, j1), p2 = p4(pb, j2), p3 = p4(pb, j3);
                    const float4 v0 = p4(vb, j0), v1 = p4(vb, j1), v2 = p4(vb, j2), v3 = p4(vb, j3);
                    const float2 q0 = q2l(qb, j0), q1 = q2l(qb, j1), q2 = q2l(qb, j2), q3 = q2l(qb, j3);
                    entry(j0, p0, v0, q0);
                    entry(j1, p1, v1, q1);
                    entry(j2, p2, v2, q2);
                    entry(j3, p3, v3, q3);
                };
                // the next group is loaded while this one's neighbours are
                uint4 e = n ? *g : make_uint4(0u, 0u, 0u, 0u);
                if (VL && l16) {
#pragma unroll 1
                    for (uint32_t k = 0; k < n; k += 8) {
                        g += mstride;
                        const uint4 en = *(k + 8 < n ? g : lgrp + i);
                        const uint32_t ws[4] = {e.x, e.y, e.z, e.w};
#pragma unroll
                        for (uint32_t hq = 0; hq < 2; ++hq) {
                            uint32_t j[4];
#pragma unroll
                            for (uint32_t q = 0; q < 4; ++q) {
                                const uint32_t m = 4u * hq + q;
                                j[q] = k + m < n ? f_s[(ws[m >> 1] >> (16u * (m & 1u))) & 0xFFFFu] : i;
                            }
                            four(j[0], j[1], j[2], j[3]);
                        }
                        e = en;
                    }
                } else {
#pragma unroll 1
                    for (uint32_t k = 0; k < n; k += 4) {
                        g += mstride;
                        const uint4 en = *(k + 4 < n ? g : lgrp + i);
                        four(e.x, k + 1 < n ? e.y : i, k + 2 < n ? e.z : i, k + 3 < n ? e.w : i);
                        e = en;
                    }
                }
            }
        }
        if (s.ghost_count) ghost_force(s, gh, pi, q_i, acc.fx, acc.fy, acc.fz);
        // fluid3.rs:1095 acc = g + F / rho (the second division by rho).
        const float ax = s.gx + acc.fx / rho_i, ay = s.gy + acc.fy / rho_i, az = s.gz + acc.fz / rho_i;
        float ux = vi.x + ax * s.dt, uy = vi.y + ay * s.dt, uz = vi.z + az * s.dt;
        if (s.viscosity > 0.0f) {
            ux = ux + acc.sx * s.viscosity;
            uy = uy + acc.sy * s.viscosity;
            uz = uz + acc.sz * s.viscosity;
        }
        const float speed = sqrtf(dot3(ux, uy, uz));
        if (speed > s.max_velocity) {
            const float r = s.max_velocity / speed;
            ux = ux * r;
            uy = uy * r;
            uz = uz * r;
        }
        if (!finite3(ux, uy, uz)) ux = uy = uz = 0.0f;
        float x = pi.x + ux * s.dt, y = pi.y + uy * s.dt, z = pi.z + uz * s.dt;
        if (!finite3(x, y, z)) {
            x = pi.x;
            y = pi.y;
            z = pi.z;
        }
        if (af.defer) {
            // a particle touching a dynamic body leaves the clamp to the
            // ordered host pass, from this state
            const float x0 = x, y0 = y, z0 = z, u0 = ux, v0 = uy, w0 = uz;
            bool dyn = false;
            clamp_to_bodies(s, bodies, x, y, z, ux, uy, uz, &dyn);
            if (dyn) {
                x = x0;
                y = y0;
                z = z0;
                ux = u0;
                uy = v0;
                uz = w0;
                const uint32_t k = atomicAdd(af.defer, 1u);
                if (k < af.defer_cap) af.defer[1 + k] = i;
            }
        } else {
            clamp_to_bodies(s, bodies, x, y, z, ux, uy, uz);
        }
        if (VL) {
            const float d2 = dot3(x - pi.x, y - pi.y, z - pi.z);
            dmax2 = d2 == d2 ? d2 : 0.0f;  // a non-finite particle is nobody's neighbour
            // cells this far out saturate; the next substep sweeps cells
            if (finite3(x, y, z) && vl_far(x, y, z, s.h_grid)) atomicOr(&vs->slow[vl_par], 1u);
        }
        posm_out[i] = make_float4(x, y, z, pi.w);
        vel_out[i] = make_float4(ux, uy, uz, 0.0f);
        if (acc_sorted) acc_sorted[i] = make_float4(ax - s.gx, ay - s.gy, az - s.gz, 0.0f);
        if (rec_acc) {
            const uint32_t id = tag[i] & PH_ID_MASK;
            rec_acc[3 * id] = ax;
            rec_acc[3 * id + 1] = ay;
            rec_acc[3 * id + 2] = az;
        }
        if (af.aux_out) {
            const uint32_t o = af.perm ? af.perm[i] : i;
            const float4 ap = af.aux_in[o];
            const uint32_t tg = tag[i];
            const float dq = rqi.y - ap.y;
            af.aux_out[i] = make_float4(rqi.x, rqi.y, dq, 0.0f);
            if (af.rec_out.pos) {
                af.rec_out.pos[i] = make_float4(x, y, z, pack_pq(rqi.y, tg >> 24));
                af.rec_out.vd[i] = pack_vd(ux, uy, uz, dq);
            }
            const float r = rqi.x / scenes[t.scene].phase_rho0[min(tg >> 24, 3u)];
            if (r >= 0.9f) {
                jerr = fabsf(r - 1.0f);
                jcnt = 1.0f;
            }
            uint32_t ks;
            if (af.sur_acc && judge_sample(o, af.sur_stride, af.sur_off, ks)) {
                const float4 as = af.sur_acc[ks];
                const float px = ax - s.gx, py = ay - s.gy, pz = az - s.gz;
                jae = sqrtf(dot3(px - as.x, py - as.y, pz - as.z));
                jam = sqrtf(dot3(px, py, pz));
            }
        }
        if (finite3(x, y, z)) {
            cell_of(x, y, z, s.h_grid, bx0, by0, bz0);
            bx1 = bx0;
            by1 = by0;
            bz1 = bz0;
        }
        // the next substep's neighbour word of this slot (a non-finite
        // particle is nobody's neighbour; its word is never compared)
        if (VL) vword_next[i] = vl_word(tag[i], bx0, by0, bz0);
    }
    // Cell bounding box of the advanced particles for the domain monitor:
    // one plain store per tile (same-address atomics from every wave would
    // serialize across the XCDs).
    for (int d = 32; d >= 1; d >>= 1) {
        bx0 = min(bx0, __shfl_xor(bx0, d, 64));
        by0 = min(by0, __shfl_xor(by0, d, 64));
        bz0 = min(bz0, __shfl_xor(bz0, d, 64));
        bx1 = max(bx1, __shfl_xor(bx1, d, 64));
        by1 = max(by1, __shfl_xor(by1, d, 64));
        bz1 = max(bz1, __shfl_xor(bz1, d, 64));
    }
    if (VL)
        for (int d = 32; d >= 1; d >>= 1) dmax2 = fmaxf(dmax2, __shfl_xor(dmax2, d, 64));
    const bool judge = af.tile_err != nullptr;
    if (judge)
        for (int d = 32; d >= 1; d >>= 1) {
            jerr += __shfl_xor(jerr, d, 64);
            jcnt += __shfl_xor(jcnt, d, 64);
            jae += __shfl_xor(jae, d, 64);
            jam += __shfl_xor(jam, d, 64);
        }
    __shared__ int32_t wb[4][6];
    __shared__ float wj[4][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        wb[w][0] = bx0;
        wb[w][1] = by0;
        wb[w][2] = bz0;
        wb[w][3] = bx1;
        wb[w][4] = by1;
        wb[w][5] = bz1;
        wj[w][0] = jerr;
        wj[w][1] = jcnt;
        wj[w][2] = jae;
        wj[w][3] = jam;
        if (VL) wj[w][0] = dmax2;
    }
    __syncthreads();
    if (VL && threadIdx.x == 0)  // non-negative: integer order is float order
        atomicMax(&vs->dmax2[vl_par], __float_as_uint(fmaxf(fmaxf(wj[0][0], wj[1][0]), fmaxf(wj[2][0], wj[3][0]))));
    if (threadIdx.x < 6) {
        const int c = threadIdx.x;
        int32_t v = wb[0][c];
        for (int k = 1; k < PH_TILE / 64; ++k) v = c < 3 ? min(v, wb[k][c]) : max(v, wb[k][c]);
        bbox[6 * (size_t)tile + c] = v;
    }
    if (judge && threadIdx.x == 0)
        af.tile_err[tile] =
            make_float4(wj[0][0] + wj[1][0] + wj[2][0] + wj[3][0], wj[0][1] + wj[1][1] + wj[2][1] + wj[3][1],
                        wj[0][2] + wj[1][2] + wj[2][2] + wj[3][2], wj[0][3] + wj[1][3] + wj[2][3] + wj[3][3]);
}

// Largest speed of each tile's particles (the list radius's input).
__global__ __launch_bounds__(PH_TILE) void k_tile_vmax(const TileDev *__restrict__ tiles,
                                                       const SceneDev *__restrict__ scenes,
                                                       const float4 *__restrict__ vel, float *__restrict__ out) {
    const TileDev t = tiles[blockIdx.x];
    const uint32_t i = t.first + threadIdx.x;
    float v = 0.0f;
    if (i < scenes[t.scene].part_end) {
        const float4 u = vel[i];
        const float sp = sqrtf(dot3(u.x, u.y, u.z));
        v = sp == sp ? sp : 0.0f;
    }
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    __shared__ float wv[PH_TILE / 64];
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = wv[0];
        for (int k = 1; k < PH_TILE / 64; ++k) m = fmaxf(m, wv[k]);
        out[blockIdx.x] = m;
    }
}

// One block: (1 + max(skin_min, coef * vmax))^2 from the per-tile speeds.
__global__ __launch_bounds__(256) void k_list_radius(const float *__restrict__ tile_vmax, uint32_t ntiles, float coef,
                                                     float skin_min, float *__restrict__ out,
                                                     float *__restrict__ vmax_out) {
    float v = 0.0f;
    for (uint32_t k = threadIdx.x; k < ntiles; k += 256) v = fmaxf(v, tile_vmax[k]);
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    __shared__ float wv[4];
    if ((threadIdx.x & 63) == 0) wv[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float vm = fmaxf(fmaxf(wv[0], wv[1]), fmaxf(wv[2], wv[3]));
        const float f = 1.0f + fmaxf(skin_min, coef * vm);
        out[0] = f * f;
        if (vmax_out) vmax_out[0] = vm;
    }
}

__global__ __launch_bounds__(PH_TILE) void k_force_list(const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes,
    const BodyDev *__restrict__ bodies, const uint32_t *__restrict__ start, const float4 *__restrict__ posm,
    const float4 *__restrict__ vel, const float2 *__restrict__ rq, const uint32_t *__restrict__ tag,
    const uint32_t *__restrict__ lcnt, const uint4 *__restrict__ lgrp, uint32_t mstride, float4 *__restrict__ posm_out,
    float4 *__restrict__ vel_out, float *__restrict__ rec_acc, int32_t *__restrict__ bbox, float4 *__restrict__ acc_sorted,
    GhostDev gh, AuxFuse af) {
    dev_force_integrate(blockIdx.x, gridDim.x, tiles, scenes, bodies, start, posm, vel, rq, tag, lcnt, lgrp, mstride,
                        posm_out, vel_out, rec_acc, bbox, acc_sorted, gh, af);
}

__global__ __launch_bounds__(PH_TILE) void k_force_verlet(const TileDev *__restrict__ tiles, const SceneDev *__restrict__ scenes,
    const BodyDev *__restrict__ bodies, const uint32_t *__restrict__ start, const float4 *__restrict__ posm,
    const float4 *__restrict__ vel, const float2 *__restrict__ rq, const uint32_t *__restrict__ tag,
    const uint32_t *__restrict__ lcnt, const uint4 *__restrict__ lgrp, uint32_t mstride, float4 *__restrict__ posm_out,
    float4 *__restrict__ vel_out, float *__restrict__ rec_acc, int32_t *__restrict__ bbox, GhostDev gh,
    VerletDev *vs, uint32_t *__restrict__ vword_next, const uint32_t *__restrict__ vhalo,
    const uint32_t *__restrict__ vhalo_n, uint32_t vl_par) {
    dev_force_integrate<true>(blockIdx.x, gridDim.x, tiles, scenes, bodies, start, posm, vel, rq, tag, lcnt, lgrp,
                              mstride, posm_out, vel_out, rec_acc, bbox, nullptr, gh, AuxFuse{}, vs, vword_next,
                              vhalo, vhalo_n, vl_par);
}

// Boundary ghost density and pressure from the fluid (fluid3.rs:906-920):
// the fluid neighbours of each ghost (fluid_near, :883-903: the 27 cells at
// the raw smoothing length, reference order, d2 < h * h), rho = sum m W,
// stored clamped; the pressure takes the unclamped sum and the config's
// base rest density.  One thread per ghost (cell-sorted slot).
__global__ __launch_bounds__(256) void k_ghost_density(const SceneDev *__restrict__ scenes, uint32_t scene,
                                                       const uint32_t *__restrict__ start,
                                                       const float4 *__restrict__ posm, GhostDev gh,
                                                       uint32_t count) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= count) return;
    const SceneDev s = scenes[scene];
    const float4 gp = gh.pos[k];
    int32_t cx, cy, cz;
    cell_of(gp.x, gp.y, gp.z, s.h_grid, cx, cy, cz);
    float rho = 0.0f;
    for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) {
        const float4 pj = posm[j];
        const float d2 = dot3(pj.x - gp.x, pj.y - gp.y, pj.z - gp.z);
        if (d2 < s.h_grid2) rho = rho + pj.w * poly6(d2, s);
    });
    const float rc = rclampf(rho, 0.0f, s.rho_max);
    const float p = eos(rho, s.rest_density, s.stiffness, s.gamma);
    const float rb = rmaxf(rc, PH_EPS);
    gh.state[k] = make_float4(rc, p, p / (rb * rb), 0.0f);
}


// ---------------------------------------------------------------- persistent
// Grid-wide barrier for the persistent step.  bar[0] counts arrivals,
// bar[1] is the generation, bar[2] an abort flag.  Agent-scope release on
// arrival and acquire on departure make every block's writes of the phase
// visible to all XCDs.  The wait is bounded: a block that spins past
// spin_limit raises the abort flag and every block leaves at its next
// barrier, so the launch always ends (the host then reports the error).
__device__ __forceinline__ bool grid_barrier(uint32_t *bar, uint32_t nblocks, uint32_t spin_limit) {
    __shared__ uint32_t ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t good = 1;
        if (arrived == nblocks - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                    ++spins > spin_limit) {
                    __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

// One substep of every scene in a single launch: the eight phases of the
// multi-kernel path, separated by grid barriers.  Each block walks the work
// items of a phase with stride gridDim.x (a multiple of 8, so item k stays
// on XCD k mod 8 as in the multi-kernel launches).
__global__ __launch_bounds__(PH_TILE) void k_step_persistent(ph::PersistArgs a) {
    const uint32_t G = gridDim.x;
    const uint32_t np = (a.nbuckets + kScanBlock * kScanItems - 1) / (kScanBlock * kScanItems);
    const uint32_t nsb = (a.n + PH_TILE - 1) / PH_TILE;
    for (uint32_t k = blockIdx.x; k < a.ntiles; k += G) {
        dev_cell_key(k, a.ntiles, a.tiles, a.scenes, a.posm_cur, a.tag_cur, a.keys, a.slots, a.counts, a.rec_cells);
        __syncthreads();
    }
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    for (uint32_t k = blockIdx.x; k < np; k += G) {
        dev_scan_reduce(k, np, a.counts, a.nbuckets, a.partial);
        __syncthreads();
    }
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    if (blockIdx.x == 0) dev_scan_partials(0, 1, a.partial, np, a.start, a.nbuckets);
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    for (uint32_t k = blockIdx.x; k < np; k += G) {
        dev_scan_apply(k, np, a.counts, a.nbuckets, a.partial, a.start);
        __syncthreads();
    }
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    for (uint32_t k = blockIdx.x; k < nsb; k += G) dev_scatter(k, nsb, a.n, a.keys, a.slots, a.start, a.tmp);
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    for (uint32_t k = blockIdx.x; k < a.ntiles; k += G)
        dev_rank_gather(k, a.ntiles, a.tiles, a.scenes, a.tmp, a.keys, a.start, a.posm_cur, a.vel_cur, a.tag_cur,
                        a.posm_alt, a.vel_alt, a.tag_alt, a.rec_sorted, nullptr);
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    // the density and force passes of a single world (k_density_list<false>,
    // k_force_list): the neighbour list carries the neighbours between them
    ph::ListOut lo;
    lo.cnt = a.lcnt;
    lo.idx = a.lidx;
    lo.skin_factor2 = 1.0f;
    for (uint32_t k = blockIdx.x; k < a.ntiles; k += G)
        dev_density<false>(k, a.ntiles, a.tiles, a.scenes, a.start, a.posm_alt, a.tag_alt, a.rq, a.mstride,
                           a.rec_rho, a.rec_p, a.rec_count, lo);
    if (!grid_barrier(a.bar, G, a.spin_limit)) return;
    for (uint32_t k = blockIdx.x; k < a.ntiles; k += G) {
        dev_force_integrate(k, a.ntiles, a.tiles, a.scenes, a.bodies, a.start, a.posm_alt, a.vel_alt, a.rq,
                            a.tag_alt, a.lcnt, reinterpret_cast<const uint4 *>(a.lidx), a.mstride, a.posm_cur,
                            a.vel_cur, a.rec_acc, a.tile_bbox, nullptr);
        __syncthreads();
    }
}

// Tags of a world's particles moved into (delta > 0) or out of (delta < 0)
// a batch: original index shifted by the world's offset, phase kept.
__global__ __launch_bounds__(PH_TILE) void k_tag_shift(uint32_t n, const uint32_t *__restrict__ src,
                                                       uint32_t *__restrict__ dst, int32_t delta) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = src[i];
    dst[i] = (t & ~PH_ID_MASK) | ((uint32_t)((int32_t)(t & PH_ID_MASK) + delta) & PH_ID_MASK);
}

// ---------------------------------------------------------------- misc
// Largest |v| (fluid3.rs:627-629) as float bits (non-negative floats
// order like their bit patterns); NaN speeds are skipped as f32::max does.
__global__ __launch_bounds__(PH_TILE) void k_vmax(uint32_t n, const float4 *__restrict__ vel,
                                                  uint32_t *__restrict__ out_bits) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    uint32_t b = 0;
    if (i < n) {
        const float4 v = vel[i];
        const float sp = sqrtf(dot3(v.x, v.y, v.z));
        if (sp == sp) b = __float_as_uint(sp);
    }
    for (int d = 32; d >= 1; d >>= 1) b = max(b, (uint32_t)__shfl_xor((int)b, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, b);
}

// State in insertion order: out6[id] = (x, y, z, vx, vy, vz).
__global__ __launch_bounds__(PH_TILE) void k_export(uint32_t n, const float4 *__restrict__ posm,
                                                    const float4 *__restrict__ vel,
                                                    const uint32_t *__restrict__ tag, uint32_t id_base,
                                                    float *__restrict__ out6) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    if (i >= n) return;
    const float4 p = posm[i], v = vel[i];
    const uint32_t id = (tag[i] & PH_ID_MASK) - id_base;
    float *o = out6 + 6 * (size_t)id;
    o[0] = p.x;
    o[1] = p.y;
    o[2] = p.z;
    o[3] = v.x;
    o[4] = v.y;
    o[5] = v.z;
}

// Overwrite velocities from insertion-ordered vel3.
__global__ __launch_bounds__(PH_TILE) void k_set_velocity(uint32_t n, const uint32_t *__restrict__ tag,
                                                          const float *__restrict__ vel3,
                                                          float4 *__restrict__ vel) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    if (i >= n) return;
    const uint32_t id = tag[i] & PH_ID_MASK;
    vel[i] = make_float4(vel3[3 * id], vel3[3 * id + 1], vel3[3 * id + 2], 0.0f);
}

// Append particles [n0, n0 + k): zero velocity, mass and tag from the
// phase tables (fluid3.rs:196-212).
__global__ __launch_bounds__(PH_TILE) void k_append(uint32_t n0, uint32_t k, const float *__restrict__ xyz,
                                                    const uint8_t *__restrict__ phase, float4 m4lo,
                                                    float4 *__restrict__ posm, float4 *__restrict__ vel,
                                                    uint32_t *__restrict__ tag) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    if (i >= k) return;
    const uint32_t ph = phase ? phase[i] : 0u;
    const uint32_t pc = min(ph, 3u);
    const float m = pc == 0 ? m4lo.x : (pc == 1 ? m4lo.y : (pc == 2 ? m4lo.z : m4lo.w));
    posm[n0 + i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], m);
    vel[n0 + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    tag[n0 + i] = (ph << 24) | (n0 + i);
}

// Cell bounding box of finite particles (domain set-up).
__global__ __launch_bounds__(PH_TILE) void k_cell_bbox(uint32_t n, const float4 *__restrict__ posm, float h,
                                                       int32_t *__restrict__ bb) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    int32_t x0 = INT32_MAX, y0 = INT32_MAX, z0 = INT32_MAX, x1 = INT32_MIN, y1 = INT32_MIN, z1 = INT32_MIN;
    if (i < n) {
        const float4 p = posm[i];
        if (finite3(p.x, p.y, p.z)) {
            cell_of(p.x, p.y, p.z, h, x0, y0, z0);
            x1 = x0;
            y1 = y0;
            z1 = z0;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        x0 = min(x0, __shfl_xor(x0, d, 64));
        y0 = min(y0, __shfl_xor(y0, d, 64));
        z0 = min(z0, __shfl_xor(z0, d, 64));
        x1 = max(x1, __shfl_xor(x1, d, 64));
        y1 = max(y1, __shfl_xor(y1, d, 64));
        z1 = max(z1, __shfl_xor(z1, d, 64));
    }
    if ((threadIdx.x & 63) == 0 && x0 != INT32_MAX) {
        // the six words only move toward the result, so a word that already
        // holds a value at least as tight as this wave's needs no atomic
        // (millions of same-address atomics otherwise serialise)
        auto rd = [&](int k) { return __hip_atomic_load(bb + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
        if (x0 < rd(0)) atomicMin(bb + 0, x0);
        if (y0 < rd(1)) atomicMin(bb + 1, y0);
        if (z0 < rd(2)) atomicMin(bb + 2, z0);
        if (x1 > rd(3)) atomicMax(bb + 3, x1);
        if (y1 > rd(4)) atomicMax(bb + 4, y1);
        if (z1 > rd(5)) atomicMax(bb + 5, z1);
    }
}

__global__ __launch_bounds__(PH_TILE) void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per-scene outlier / non-finite counts and the cell bounding box reduced
// over the scene's tiles, into one small array the host copies back
// asynchronously (domain monitor).  One block per scene.
__global__ __launch_bounds__(256) void k_monitor(const SceneDev *__restrict__ scenes,
                                                 const uint32_t *__restrict__ start,
                                                 const int32_t *__restrict__ tile_bbox,
                                                 int32_t *__restrict__ out) {
    const uint32_t k = blockIdx.x;
    const SceneDev &s = scenes[k];
    int32_t v[6] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MIN, INT32_MIN, INT32_MIN};
    for (uint32_t t = s.tile_begin + threadIdx.x; t < s.tile_begin + s.tile_count; t += 256) {
        for (int c = 0; c < 3; ++c) v[c] = min(v[c], tile_bbox[6 * (size_t)t + c]);
        for (int c = 3; c < 6; ++c) v[c] = max(v[c], tile_bbox[6 * (size_t)t + c]);
    }
    __shared__ int32_t red[256][6];
    for (int c = 0; c < 6; ++c) red[threadIdx.x][c] = v[c];
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d)
            for (int c = 0; c < 6; ++c)
                red[threadIdx.x][c] = c < 3 ? min(red[threadIdx.x][c], red[threadIdx.x + d][c])
                                            : max(red[threadIdx.x][c], red[threadIdx.x + d][c]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[8 * k] = (int32_t)(start[s.bucket_base + s.ncells + 1] - start[s.bucket_base + s.ncells]);
        out[8 * k + 1] = (int32_t)(start[s.bucket_base + s.ncells + 2] - start[s.bucket_base + s.ncells + 1]);
        for (int c = 0; c < 6; ++c) out[8 * k + 2 + c] = red[0][c];
    }
}

// ---------------------------------------------------------------- launchers
namespace ph {

static inline uint32_t blocks_for(uint32_t n) { return (n + PH_TILE - 1) / PH_TILE; }

hipError_t launch_cell_key(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                           const float4 *posm, const uint32_t *tag, uint32_t *keys, uint32_t *slots,
                           uint32_t *counts, int32_t *rec_cells, VerletDev *vs, const VlStep &vst,
                           uint32_t *vword, uint32_t *pacc) {
    hipLaunchKernelGGL(k_cell_key, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, posm, tag, keys, slots,
                       counts, rec_cells, vs, vst, vword, pacc);
    return hipGetLastError();
}

uint32_t scan_partials_needed(uint32_t nbuckets) {
    return (nbuckets + kScanBlock * kScanItems - 1) / (kScanBlock * kScanItems);
}

hipError_t launch_scan_reduce(hipStream_t st, const uint32_t *counts, uint32_t nb, uint32_t *partial,
                              const uint32_t *gate) {
    hipLaunchKernelGGL(k_scan_reduce, dim3(scan_partials_needed(nb)), dim3(kScanBlock), 0, st, counts, nb,
                       partial, gate);
    return hipGetLastError();
}
hipError_t launch_scan_partials(hipStream_t st, uint32_t *partial, uint32_t nb, uint32_t *start,
                                const uint32_t *gate) {
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlock), 0, st, partial, scan_partials_needed(nb),
                       start, nb, gate);
    return hipGetLastError();
}
hipError_t launch_scan_apply(hipStream_t st, uint32_t *counts, uint32_t nb, const uint32_t *partial,
                             uint32_t *start, const uint32_t *gate) {
    hipLaunchKernelGGL(k_scan_apply, dim3(scan_partials_needed(nb)), dim3(kScanBlock), 0, st, counts, nb,
                       partial, start, gate);
    return hipGetLastError();
}
hipError_t launch_scan_apply_sum(hipStream_t st, uint32_t *counts, uint32_t nb, const uint32_t *partial,
                                 uint32_t *start, const uint32_t *gate) {
    hipLaunchKernelGGL(k_scan_apply_sum, dim3(scan_partials_needed(nb)), dim3(kScanBlock), 0, st, counts, nb,
                       partial, start, gate);
    return hipGetLastError();
}
hipError_t launch_scatter(hipStream_t st, uint32_t n, const uint32_t *keys, const uint32_t *slots,
                          const uint32_t *start, uint32_t *tmp, const uint32_t *gate, uint32_t *zero,
                          uint32_t nzero) {
    hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, n, keys, slots, start, tmp, gate, zero,
                       nzero);
    return hipGetLastError();
}
hipError_t launch_rank_gather(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                              const uint32_t *tmp, const uint32_t *keys, const uint32_t *start,
                              const float4 *posm_in, const float4 *vel_in, const uint32_t *tag_in,
                              float4 *posm_out, float4 *vel_out, uint32_t *tag_out, uint32_t *rec_sorted,
                              uint32_t *perm_out, const VerletDev *vs, const uint32_t *vword_in,
                              uint32_t *vword_out) {
    hipLaunchKernelGGL(k_rank_gather, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, tmp, keys, start,
                       posm_in, vel_in, tag_in, posm_out, vel_out, tag_out, rec_sorted, perm_out, vs, vword_in,
                       vword_out);
    return hipGetLastError();
}
hipError_t launch_force_verlet(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                               const BodyDev *bodies, const uint32_t *start, const float4 *posm, const float4 *vel,
                               const float2 *rq, const uint32_t *tag, const uint32_t *lcnt, const uint32_t *lidx,
                               uint32_t mstride, float4 *posm_out, float4 *vel_out, float *rec_acc, int32_t *bbox,
                               const GhostDev &gh, VerletDev *vs, uint32_t *vword_next, const uint32_t *vhalo,
                               const uint32_t *vhalo_n, uint32_t vl_par) {
    hipLaunchKernelGGL(k_force_verlet, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, bodies, start, posm, vel,
                       rq, tag, lcnt, reinterpret_cast<const uint4 *>(lidx), mstride, posm_out, vel_out, rec_acc,
                       bbox, gh, vs, vword_next, vhalo, vhalo_n, vl_par);
    return hipGetLastError();
}
hipError_t launch_density_list(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                               const uint32_t *start, const float4 *posm, const uint32_t *tag, float2 *rq,
                               uint32_t mstride, float *rec_rho, float *rec_p, uint32_t *rec_count,
                               const ListOut &lo) {
    if (lo.skin_factor2 != 1.0f || lo.skin_factor2_dev)
        hipLaunchKernelGGL(k_density_list<true>, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, start, posm, tag,
                           rq, mstride, rec_rho, rec_p, rec_count, lo);
    else
        hipLaunchKernelGGL(k_density_list<false>, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, start, posm, tag,
                           rq, mstride, rec_rho, rec_p, rec_count, lo);
    return hipGetLastError();
}
hipError_t launch_force_list(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                             const BodyDev *bodies, const uint32_t *start, const float4 *posm, const float4 *vel,
                             const float2 *rq, const uint32_t *tag, const uint32_t *lcnt, const uint32_t *lidx,
                             uint32_t mstride, float4 *posm_out, float4 *vel_out, float *rec_acc, int32_t *bbox,
                             float4 *acc_sorted, const GhostDev &gh, const AuxFuse &af) {
    hipLaunchKernelGGL(k_force_list, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, bodies, start, posm, vel, rq,
                       tag, lcnt, reinterpret_cast<const uint4 *>(lidx), mstride, posm_out, vel_out, rec_acc, bbox,
                       acc_sorted, gh, af);
    return hipGetLastError();
}
hipError_t launch_list_radius(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                              const float4 *vel, float *tile_vmax, float coef, float skin_min, float *out,
                              float *vmax_out) {
    hipLaunchKernelGGL(k_tile_vmax, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, vel, tile_vmax);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_list_radius, dim3(1), dim3(256), 0, st, tile_vmax, ntiles, coef, skin_min, out, vmax_out);
    return hipGetLastError();
}
hipError_t launch_ghost_density(hipStream_t st, const SceneDev *scenes, uint32_t scene, const uint32_t *start,
                                const float4 *posm, const GhostDev &gh, uint32_t count) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ghost_density, dim3((count + 255u) / 256u), dim3(256), 0, st, scenes, scene, start, posm, gh,
                       count);
    return hipGetLastError();
}
hipError_t launch_vmax(hipStream_t st, uint32_t n, const float4 *vel, uint32_t *out_bits) {
    hipLaunchKernelGGL(k_vmax, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, n, vel, out_bits);
    return hipGetLastError();
}
hipError_t launch_export(hipStream_t st, uint32_t n, const float4 *posm, const float4 *vel, const uint32_t *tag,
                         uint32_t id_base, float *out6) {
    hipLaunchKernelGGL(k_export, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, n, posm, vel, tag, id_base, out6);
    return hipGetLastError();
}
hipError_t launch_set_velocity(hipStream_t st, uint32_t n, const uint32_t *tag, const float *vel3,
                               float4 *vel) {
    hipLaunchKernelGGL(k_set_velocity, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, n, tag, vel3, vel);
    return hipGetLastError();
}
hipError_t launch_append(hipStream_t st, uint32_t n0, uint32_t k, const float *xyz, const uint8_t *phase,
                         float4 masses, float4 *posm, float4 *vel, uint32_t *tag) {
    hipLaunchKernelGGL(k_append, dim3(blocks_for(k)), dim3(PH_TILE), 0, st, n0, k, xyz, phase, masses, posm,
                       vel, tag);
    return hipGetLastError();
}
hipError_t launch_cell_bbox(hipStream_t st, uint32_t n, const float4 *posm, float h, int32_t *bb) {
    hipLaunchKernelGGL(k_cell_bbox, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, n, posm, h, bb);
    return hipGetLastError();
}
hipError_t launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v) {
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, p, n, v);
    return hipGetLastError();
}
hipError_t launch_monitor(hipStream_t st, const SceneDev *scenes, uint32_t nscenes, const uint32_t *start,
                          const int32_t *tile_bbox, int32_t *out) {
    hipLaunchKernelGGL(k_monitor, dim3(nscenes), dim3(256), 0, st, scenes, start, tile_bbox, out);
    return hipGetLastError();
}

hipError_t launch_step_persistent(hipStream_t st, uint32_t blocks, const PersistArgs &a) {
    hipLaunchKernelGGL(k_step_persistent, dim3(blocks), dim3(PH_TILE), 0, st, a);
    return hipGetLastError();
}
hipError_t persistent_max_blocks_per_cu(int *out) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(out, k_step_persistent, PH_TILE, 0);
}
hipError_t launch_tag_shift(hipStream_t st, uint32_t n, const uint32_t *src, uint32_t *dst, int32_t delta) {
    hipLaunchKernelGGL(k_tag_shift, dim3(blocks_for(n)), dim3(PH_TILE), 0, st, n, src, dst, delta);
    return hipGetLastError();
}

}  // namespace ph
