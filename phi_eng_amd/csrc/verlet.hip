// verlet.hip — neighbour search of a plain substep over Verlet lists.
//
// The reference finds a particle's neighbours by sweeping its 27 cells
// every substep (fluid3.rs:319-340): ~216 candidates for ~33 neighbours.
// Here each particle keeps a row of every particle that was within
// (1 + skin) h of it at the last rebuild.  While no particle has moved more
// than 0.45 skin h since then (the bound vl_decide keeps), every particle
// within h of it is in its row, and the row is walked instead of the cells.
//
// Between rebuilds the particles keep their slots: the sort's kernels are
// launched but return at once, and the gather copies the state unchanged,
// so a row holds slots and the host's launch sequence never depends on the
// device's decision.  Results stay those of the reference, bit for bit:
// - a neighbour is a particle j with d2 < h * h (the reference's f32 test)
//   whose cell is one of i's 27; with the cells mod 4 in the neighbour
//   words, vl_key decides that exactly;
// - the reference sums in cell-offset order (oz, oy, ox), then by original
//   index; a row is walked in its stored order and the keys of the
//   neighbours found must increase.  If not (a neighbour changed cell since
//   the row was last ordered) the wave sorts the row by key, stores it
//   back in that order and sums again;
// - a rebuild sorts the state (cell, then index) and sweeps the cells
//   within (1 + skin) h in ascending slot order, which is the reference
//   order, summing the density over the 27 cells in the same sweep.  A
//   particle the sweep cannot serve (its ball leaves the grid while
//   outliers exist, or more than VL_CAP candidates) takes the per-cell path
//   of the list-free pass, and the next substep sorts and rebuilds again.
// Every form writes the particle's neighbours within h, in reference order,
// to the list the force pass walks (as the list-free density pass does).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sph_kernels.h"

namespace {

// fluid3.rs:646-655 after the sum: clamp, equation of state, the force
// pass's pressure quotient, and the parity records (same operations as the
// list-free density pass).
__device__ __forceinline__ void density_out(const SceneDev &s, const SceneDev *__restrict__ scenes, uint32_t scene,
                                            uint32_t i, float rho, uint32_t cnt, const uint32_t *__restrict__ tag,
                                            float2 *__restrict__ rq, float *__restrict__ rec_rho,
                                            float *__restrict__ rec_p, uint32_t *__restrict__ rec_count) {
    rho = rclampf(rho, 0.0f, s.rho_max);
    const uint32_t tg = tag[i];
    const float rho0 = scenes[scene].phase_rho0[min(tg >> 24, 3u)];
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

// The neighbour list handed to the force pass: the slots within h in
// reference order, four to a 16-byte group (group g of slot i at
// g * stride + i), at most SL_CAP.
struct ListPush {
    uint4 *g;
    uint32_t stride;
    uint32_t k = 0;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    __device__ __forceinline__ ListPush(uint32_t *base, uint32_t stride_, uint32_t i)
        : g(reinterpret_cast<uint4 *>(base) + i), stride(stride_) {}
    __device__ __forceinline__ void push(uint32_t j) {
        const uint32_t p = k & 3u;
        q0 = p == 0u ? j : q0;
        q1 = p == 1u ? j : q1;
        q2 = p == 2u ? j : q2;
        q3 = p == 3u ? j : q3;
        if (p == 3u && k < SL_CAP) g[(size_t)(k >> 2) * stride] = make_uint4(q0, q1, q2, q3);
        ++k;
    }
    // the last, partly filled group; the count word for the force pass
    __device__ __forceinline__ uint32_t finish() {
        if ((k & 3u) != 0u && k < SL_CAP) g[(size_t)(k >> 2) * stride] = make_uint4(q0, q1, q2, q3);
        return k <= SL_CAP ? k : SLF_SLOW;
    }
};

// The same list from a tile that walks its halo: the neighbours' halo
// indices (16 bits), eight to a 16-byte group.  Each push shifts the group
// down by one entry and puts the new one on top, so after eight pushes
// entry m of the group sits at bits 16 m.
static_assert(VL_CAP <= SL_CAP, "a walked row's neighbours fit the force pass's list");
struct ListPush16 {
    uint4 *g;
    uint32_t stride;
    uint32_t k = 0;
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0;
    __device__ __forceinline__ ListPush16(uint32_t *base, uint32_t stride_, uint32_t i)
        : g(reinterpret_cast<uint4 *>(base) + i), stride(stride_) {}
    __device__ __forceinline__ void shift_in(uint32_t x) {
        w0 = __builtin_amdgcn_alignbit(w1, w0, 16u);
        w1 = __builtin_amdgcn_alignbit(w2, w1, 16u);
        w2 = __builtin_amdgcn_alignbit(w3, w2, 16u);
        w3 = __builtin_amdgcn_alignbit(x, w3, 16u);
    }
    // (a walked row has at most VL_CAP = SL_CAP entries, so every push fits)
    __device__ __forceinline__ void push(uint32_t x) {
        shift_in(x);
        if ((k & 7u) == 7u) g[(size_t)(k >> 3) * stride] = make_uint4(w0, w1, w2, w3);
        ++k;
    }
    // the last, partly filled group moved down to bit 0; the count word
    __device__ __forceinline__ uint32_t finish() {
        if ((k & 7u) != 0u && k < SL_CAP) {
            for (uint32_t r = k & 7u; r < 8u; ++r) shift_in(0u);
            g[(size_t)(k >> 3) * stride] = make_uint4(w0, w1, w2, w3);
        }
        return k <= SL_CAP ? k : SLF_SLOW;
    }
};

// The per-cell sum (reference order, outlier cells included).
__device__ __forceinline__ float density_cells(const SceneDev &s, const uint32_t *__restrict__ start,
                                               const float4 *__restrict__ posm, const float4 &pi, int32_t cx,
                                               int32_t cy, int32_t cz, uint32_t &cnt) {
    float rho = 0.0f;
    for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) {
        const float4 pj = posm[j];
        const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
        if (r2 < s.h_grid2) {
            ++cnt;
            rho += pj.w * poly6(r2, s);
        }
    });
    return rho;
}

// Cells per axis that the ball of radius (1 + skin) h around x can reach:
// the neighbour cells, and the next ones when x lies within the skin (plus
// a margin for the rounding of x / h) of a face.  Far from the origin the
// fraction is not resolved and both are taken.
__device__ __forceinline__ void reach(float x, float h, int32_t c, float skin, int64_t &lo, int64_t &hi) {
    const float q = x / h;
    const float fr = q - floorf(q);
    const float m = skin + 0.01f + fabsf(q) * 0x1p-22f;
    const bool wide = !(fabsf(q) < 0x1p22f);
    lo = (int64_t)c - 1 - ((wide || fr < m) ? 1 : 0);
    hi = (int64_t)c + 1 + ((wide || fr > 1.0f - m) ? 1 : 0);
}

// Exclusive prefix sum over a PH_TILE-thread block; total to every thread.
__device__ __forceinline__ uint32_t block_scan_u32(uint32_t v, uint32_t &total) {
    __shared__ uint32_t wsum[PH_TILE / 64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, 64);
        if ((int)lane >= d) inc += y;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
    for (uint32_t k = 0; k < PH_TILE / 64; ++k) {
        if (k < w) off += wsum[k];
        total += wsum[k];
    }
    __syncthreads();
    return off + inc - v;
}

}  // namespace

// The rebuild form of a particle: sweep the cells within (1 + skin) h,
// writing the row (every particle within (1 + skin) h, ascending slot) and
// summing the density over the 27 cells in the same sweep.
__device__ __forceinline__ void density_rebuild(const SceneDev &s, const SceneDev *__restrict__ scenes,
                                                uint32_t scene, const uint32_t *__restrict__ start,
                                                const float4 *__restrict__ posm, const uint32_t *__restrict__ tag,
                                                float2 *__restrict__ rq, float *__restrict__ rec_rho,
                                                float *__restrict__ rec_p, uint32_t *__restrict__ rec_count,
                                                uint32_t *__restrict__ fcnt, const ph::VerletArgs &va, uint32_t i,
                                                const float4 &pi) {
    uint32_t cnt = 0;
    float rho = 0.0f;
    ListPush lp(va.lidx, va.lstride, i);
    const float hg2 = s.h_grid2;
    int32_t cx, cy, cz;
    cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
    int64_t xlo, xhi, ylo, yhi, zlo, zhi;
    reach(pi.x, s.h_grid, cx, va.skin, xlo, xhi);
    reach(pi.y, s.h_grid, cy, va.skin, ylo, yhi);
    reach(pi.z, s.h_grid, cz, va.skin, zlo, zhi);
    const bool inside = in_domain(s, xlo, ylo, zlo) && in_domain(s, xhi, yhi, zhi);
    const uint32_t ob = start[s.bucket_base + s.ncells], oe = start[s.bucket_base + s.ncells + 1];
    if ((!inside && ob != oe) || vl_far(pi.x, pi.y, pi.z, s.h_grid)) {
        // the ball leaves the grid while outliers exist, or the cells
        // saturate: the per-cell path, and the next substep sorts again
        va.vcnt[i] = VLF_SLOW;
        fcnt[i] = SLF_SLOW;
        atomicOr(&va.vs->slow[va.par], 1u);
        rho = density_cells(s, start, posm, pi, cx, cy, cz, cnt);
        density_out(s, scenes, scene, i, rho, cnt, tag, rq, rec_rho, rec_p, rec_count);
        return;
    }
    const float R2 = hg2 * va.f2;
    uint32_t lk = 0;
    uint32_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    uint4 *lg = reinterpret_cast<uint4 *>(va.vidx) + i;
    auto push = [&](uint32_t j) {
        const uint32_t p = lk & 3u;
        q0 = p == 0u ? j : q0;
        q1 = p == 1u ? j : q1;
        q2 = p == 2u ? j : q2;
        q3 = p == 3u ? j : q3;
        if (p == 3u && lk < VL_CAP) lg[(size_t)(lk >> 2) * va.vstride] = make_uint4(q0, q1, q2, q3);
        ++lk;
    };
    for (int64_t z = max(zlo, (int64_t)s.z0); z <= min(zhi, (int64_t)s.z0 + s.nz - 1); ++z) {
        for (int64_t y = max(ylo, (int64_t)s.y0); y <= min(yhi, (int64_t)s.y0 + s.ny - 1); ++y) {
            const int64_t xl = max(xlo, (int64_t)s.x0), xh = min(xhi, (int64_t)s.x0 + s.nx - 1);
            if (xl > xh) continue;
            const uint32_t row = s.bucket_base + (uint32_t)(((z - s.z0) * s.ny + (y - s.y0)) * s.nx);
            const uint32_t b = start[row + (uint32_t)(xl - s.x0)], e = start[row + (uint32_t)(xh - s.x0) + 1];
            // the part of this row inside the 27 cells (empty off them)
            uint32_t b27 = 0, e27 = 0;
            if (llabs(z - cz) <= 1 && llabs(y - cy) <= 1) {
                const int64_t al = max((int64_t)cx - 1, (int64_t)s.x0),
                              ah = min((int64_t)cx + 1, (int64_t)s.x0 + s.nx - 1);
                if (al <= ah) {
                    b27 = start[row + (uint32_t)(al - s.x0)];
                    e27 = start[row + (uint32_t)(ah - s.x0) + 1];
                }
            }
            for (uint32_t j = b; j < e; ++j) {
                const float4 pj = posm[j];
                const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
                if (r2 < R2) push(j);
                if (j >= b27 && j < e27 && r2 < hg2) {
                    ++cnt;
                    rho += pj.w * poly6(r2, s);
                    lp.push(j);
                }
            }
        }
    }
    if ((lk & 3u) != 0u && lk < VL_CAP) lg[(size_t)(lk >> 2) * va.vstride] = make_uint4(q0, q1, q2, q3);
    const uint32_t fc = lp.finish();
    if (lk > VL_CAP || fc == SLF_SLOW) {
        va.vcnt[i] = VLF_SLOW;
        atomicOr(&va.vs->slow[va.par], 1u);
    } else {
        va.vcnt[i] = lk;
    }
    fcnt[i] = fc;
    density_out(s, scenes, scene, i, rho, cnt, tag, rq, rec_rho, rec_p, rec_count);
}

// Within one wave: make LDS stores of every lane visible to the loads of
// every lane that follow.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One block per tile: the sorted union of the tile's lists (its halo, at
// most cap slots) from a bitmap over the lists' slot span, and every entry
// rewritten as its rank in the halo (16 bits, eight to a 16-byte group:
// group g of slot i at out16[g * stride + i]).  A list is a count word per
// slot (bit 24: the slot takes the per-cell path; bit 25: no entries) and
// groups of four slots (group g of slot i at idx[g * stride + i]).  A tile
// whose span or halo is too large, or with a slot on the per-cell path,
// keeps the 32-bit lists (halo_n = VL_HALO_NONE).
constexpr uint32_t kHaloWords = VL_HALO_SPAN / 32u;
// LDS of the halo build: the span's bitmap, the halo rank of every fourth
// word's first bit (16-bit: ranks stay below VL_HALO_CAP), a reduction.
constexpr uint32_t kHaloLds = 4u * kHaloWords + 2u * (kHaloWords / 4u) + 4u * 3u * (PH_TILE / 64u);

__device__ __forceinline__ void halo_build_tile(uint32_t tile, const TileDev &t, const SceneDev &s,
                                                const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ idx,
                                                uint32_t stride, uint32_t cap, uint32_t *__restrict__ halo,
                                                uint32_t *__restrict__ halo_n, uint4 *__restrict__ out16,
                                                uint8_t *lds) {
    constexpr uint32_t kWords = kHaloWords, kPer = kWords / PH_TILE;
    static_assert(kPer % 4u == 0u, "a thread's words start a group of four");
    uint32_t *bm = reinterpret_cast<uint32_t *>(lds);                    // bit (slot - lo)
    uint16_t *pw = reinterpret_cast<uint16_t *>(lds + 4u * kWords);     // rank at word 4k
    uint32_t(*red)[PH_TILE / 64] = reinterpret_cast<uint32_t(*)[PH_TILE / 64]>(lds + 4u * kWords + 2u * (kWords / 4u));
    const uint32_t i = t.first + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint4 *grp = reinterpret_cast<const uint4 *>(idx) + i;
    const uint32_t raw = i < s.part_end ? cnt[i] : 0u;
    const bool slow = (raw & 0x01000000u) != 0u;
    const uint32_t n = (slow || (raw & 0x02000000u)) ? 0u : min(raw & 0x00FFFFFFu, VL_CAP);
    auto group = [&](uint32_t g) { return grp[(size_t)g * stride]; };
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (uint32_t m = 0; m < n; m += 4) {
        const uint4 e = group(m >> 2);
        const uint32_t e1 = m + 1 < n ? e.y : e.x, e2 = m + 2 < n ? e.z : e.x, e3 = m + 3 < n ? e.w : e.x;
        lo = min(lo, min(min(e.x, e1), min(e2, e3)));
        hi = max(hi, max(max(e.x, e1), max(e2, e3)));
    }
    uint32_t bad = slow ? 1u : 0u;
    for (int d = 32; d >= 1; d >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor(lo, d, 64));
        hi = max(hi, (uint32_t)__shfl_xor(hi, d, 64));
        bad |= (uint32_t)__shfl_xor(bad, d, 64);
    }
    if (lane == 0) {
        red[0][w] = lo;
        red[1][w] = hi;
        red[2][w] = bad;
    }
    __syncthreads();
    for (uint32_t k = 0; k < PH_TILE / 64; ++k) {
        lo = min(lo, red[0][k]);
        hi = max(hi, red[1][k]);
        bad |= red[2][k];
    }
    if (bad || (hi >= lo && hi - lo >= VL_HALO_SPAN)) {
        if (threadIdx.x == 0) halo_n[tile] = VL_HALO_NONE;
        return;
    }
    if (hi < lo) {  // no entries at all
        if (threadIdx.x == 0) halo_n[tile] = 0u;
        return;
    }
    const uint32_t nw = ((hi - lo) >> 5) + 1u;
    for (uint32_t k = threadIdx.x; k < nw; k += PH_TILE) bm[k] = 0u;
    __syncthreads();
    for (uint32_t m = 0; m < n; m += 4) {
        const uint4 e = group(m >> 2);
        const uint32_t es[4] = {e.x, e.y, e.z, e.w};
        for (uint32_t q = 0; q < 4 && m + q < n; ++q) {
            const uint32_t d = es[q] - lo;
            atomicOr(&bm[d >> 5], 1u << (d & 31u));
        }
    }
    __syncthreads();
    // ranks: thread k owns words [k kPer, (k + 1) kPer)
    uint32_t c = 0;
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t k = threadIdx.x * kPer + q;
        if (k < nw) c += __popc(bm[k]);
    }
    uint32_t total;
    const uint32_t base = block_scan_u32(c, total);
    if (total > cap) {
        if (threadIdx.x == 0) halo_n[tile] = VL_HALO_NONE;
        return;
    }
    uint32_t r = base;
    uint32_t *hl = halo + (size_t)tile * VL_HALO_CAP;
    for (uint32_t q = 0; q < kPer; ++q) {
        const uint32_t k = threadIdx.x * kPer + q;
        if (k >= nw) break;
        if ((k & 3u) == 0u) pw[k >> 2] = (uint16_t)r;
        for (uint32_t bits = bm[k]; bits; bits &= bits - 1u) hl[r++] = lo + k * 32u + (uint32_t)(__ffs(bits) - 1);
    }
    __syncthreads();
    uint4 *r16 = out16 + i;
    for (uint32_t m = 0; m < n; m += 8) {
        const uint4 e0 = group(m >> 2), e1 = m + 4 < n ? group((m >> 2) + 1) : e0;
        const uint32_t es[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
        uint32_t v[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (uint32_t q = 0; q < 8; ++q) {
            if (m + q < n) {
                // the words before d's in its group of four (never past nw
                // when used), then d's own bits below it
                const uint32_t d = es[q] - lo, wd = d >> 5, sub = wd & 3u;
                const uint4 b4 = *reinterpret_cast<const uint4 *>(bm + (wd & ~3u));
                const uint32_t own = sub == 0u ? b4.x : sub == 1u ? b4.y : sub == 2u ? b4.z : b4.w;
                const uint32_t rank = pw[wd >> 2] + (sub > 0u ? __popc(b4.x) : 0u) + (sub > 1u ? __popc(b4.y) : 0u) +
                                      (sub > 2u ? __popc(b4.z) : 0u) + __popc(own & ((1u << (d & 31u)) - 1u));
                v[q >> 1] |= rank << (16u * (q & 1u));
            }
        }
        r16[(size_t)(m >> 3) * stride] = make_uint4(v[0], v[1], v[2], v[3]);
    }
    if (threadIdx.x == 0) halo_n[tile] = total;
}

// The rebuild substep's density pass for one particle of a tile: the
// particle sweeps the cells for its row and density.
__device__ __forceinline__ void rebuild_particle(const SceneDev &s, const SceneDev *__restrict__ scenes,
                                                 uint32_t scene, const uint32_t *__restrict__ start,
                                                 const float4 *__restrict__ posm, const uint32_t *__restrict__ tag,
                                                 float2 *__restrict__ rq, float *__restrict__ rec_rho,
                                                 float *__restrict__ rec_p, uint32_t *__restrict__ rec_count,
                                                 uint32_t *__restrict__ fcnt, const ph::VerletArgs &va, uint32_t i) {
    // both tag generations hold the sorted tags until the next rebuild,
    // and the walk's generation the sorted state
    va.tag_mirror[i] = tag[i];
    const float4 pi = posm[i];
    va.keep_posm[i] = pi;
    va.keep_vel[i] = va.sorted_vel[i];
    if (!finite3(pi.x, pi.y, pi.z)) {
        // fluid3.rs:302-304: no neighbours, density 0; an empty row
        va.vcnt[i] = 0u;
        fcnt[i] = SLF_NONE;
        density_out(s, scenes, scene, i, 0.0f, 0u, tag, rq, rec_rho, rec_p, rec_count);
    } else {
        density_rebuild(s, scenes, scene, start, posm, tag, rq, rec_rho, rec_p, rec_count, fcnt, va, i, pi);
    }
}

// LDS of the walk: the halo's positions and words, and per wave one row
// being put in order (keys; entries with the within-h bit and
// contributions, overwritten in rank order).
constexpr uint32_t kWalkLds = (16u + 4u) * VL_HALO_CAP + 3u * 4u * (PH_TILE / 64u) * VL_CAP;
constexpr uint32_t kVerletLds = kWalkLds > kHaloLds ? kWalkLds : kHaloLds;

// The Verlet substep's density pass, one launch whichever the device
// decided:
// - a rebuild substep sweeps the cells for every particle of the tile
//   (rows, density, the force pass's list), then builds the tile's halo
//   from those rows, in the same block (a tile's halo reads only its own
//   rows);
// - a walk substep: a tile with a halo copies its halo's positions and
//   words to LDS, walks 16-bit rows and hands the force pass a 16-bit list
//   of halo indices; other tiles walk 32-bit rows of slots and hand over
//   slots.  A row group (eight entries) is loaded while the one before is
//   summed, and each group's LDS reads are issued together.
// The two forms share one LDS block.
__global__ __launch_bounds__(PH_TILE) void k_density_verlet(const TileDev *__restrict__ tiles,
                                                            const SceneDev *__restrict__ scenes,
                                                            const uint32_t *__restrict__ start,
                                                            const float4 *__restrict__ posm_sorted,
                                                            const float4 *__restrict__ posm,
                                                            const uint32_t *__restrict__ tag, float2 *__restrict__ rq,
                                                            float *__restrict__ rec_rho, float *__restrict__ rec_p,
                                                            uint32_t *__restrict__ rec_count,
                                                            uint32_t *__restrict__ fcnt, ph::VerletArgs va) {
    __shared__ __align__(16) uint8_t lds[kVerletLds];
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const TileDev t = tiles[tile];
    const SceneDev s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    const bool act = i < s.part_end;
    if (va.vs->rebuild != 0u) {
        if (act) rebuild_particle(s, scenes, t.scene, start, posm_sorted, tag, rq, rec_rho, rec_p, rec_count, fcnt,
                                  va, i);
        __syncthreads();  // the tile's rows, written above, are read below
        halo_build_tile(tile, t, s, va.vcnt, va.vidx, va.vstride, VL_HALO_CAP, va.halo, va.halo_n, va.rows16, lds);
        return;
    }
    float4 *h_p = reinterpret_cast<float4 *>(lds);                           // position and mass
    uint32_t *h_w = reinterpret_cast<uint32_t *>(lds + 16u * VL_HALO_CAP);  // neighbour word
    uint32_t(*s_key)[VL_CAP] = reinterpret_cast<uint32_t(*)[VL_CAP]>(lds + 20u * VL_HALO_CAP);
    uint32_t(*s_uj)[VL_CAP] = s_key + PH_TILE / 64u;
    float(*s_uc)[VL_CAP] = reinterpret_cast<float(*)[VL_CAP]>(s_uj + PH_TILE / 64u);
    const uint32_t H = va.halo_n[tile];
    const bool halo = H != VL_HALO_NONE;  // the same for the whole block
    if (halo) {
        // every load of a stage issued before the next stage waits: the
        // slots, then the positions and words, then the LDS stores
        const uint32_t *hl = va.halo + (size_t)tile * VL_HALO_CAP;
        constexpr uint32_t kR = (VL_HALO_CAP + PH_TILE - 1) / PH_TILE;
        uint32_t js[kR];
#pragma unroll
        for (uint32_t q = 0; q < kR; ++q) {
            const uint32_t k = threadIdx.x + q * PH_TILE;
            js[q] = k < H ? hl[k] : 0u;
        }
        float4 ps[kR];
        uint32_t ws[kR];
#pragma unroll
        for (uint32_t q = 0; q < kR; ++q) {  // past H: slot 0, read and dropped
            ps[q] = posm[js[q]];
            ws[q] = va.vword[js[q]];
        }
#pragma unroll
        for (uint32_t q = 0; q < kR; ++q) {
            const uint32_t k = threadIdx.x + q * PH_TILE;
            if (k < H) {
                h_p[k] = ps[q];
                h_w[k] = ws[q];
            }
        }
        __syncthreads();
    }
    const float4 pi = act ? posm[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool fin = finite3(pi.x, pi.y, pi.z);
    // Walk (every lane of the wave stays to the end: the repair below is
    // the whole wave's).
    const float hg2 = s.h_grid2;
    const bool mine = act && fin;
    if (act && !fin) {
        fcnt[i] = SLF_NONE;
        density_out(s, scenes, t.scene, i, 0.0f, 0u, tag, rq, rec_rho, rec_p, rec_count);
    }
    uint4 *rows = reinterpret_cast<uint4 *>(va.vidx);
    // never VLF_SLOW here (such a state sorts and rebuilds); bounded anyway
    const uint32_t n = mine ? min(va.vcnt[i], VL_CAP) : 0u;
    const uint32_t wi = mine ? va.vword[i] : 0u;
    const uint32_t ai = vl_addend(wi);
    bool bad = false;
    uint32_t cnt = 0, fc = 0;
    float rho = 0.0f;
    uint32_t next = 0;  // the next neighbour's key must be at least this
    // one row entry: true if it is a neighbour (within h, among the 27 cells)
    auto one = [&](const float4 &pj, uint32_t wj) {
        const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
        const uint32_t key = vl_key(ai, wj);
        const bool in = r2 < hg2 && key != VL_NOT27;
        if (in) {
            bad |= key < next;
            next = key + 1u;
            ++cnt;
            rho += pj.w * poly6_within(r2, s);
        }
        return in;
    };
    if (halo) {
        ListPush16 lp(va.lidx, va.lstride, i);
        const uint4 *row = va.rows16 + i;
        // entries past the count in a row's last group are 0 (a valid index)
        uint4 e = n ? *row : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll 1
        for (uint32_t k = 0; k < n; k += 8) {
            row += va.vstride;
            const uint4 en = *(k + 8 < n ? row : va.rows16 + i);
            const uint32_t ws[4] = {e.x, e.y, e.z, e.w};
            uint32_t x[8], wj[8];
            float4 pj[8];
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) x[q] = (ws[q >> 1] >> (16u * (q & 1u))) & 0xFFFFu;
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q) {
                pj[q] = h_p[x[q]];
                wj[q] = h_w[x[q]];
            }
#pragma unroll
            for (uint32_t q = 0; q < 8; ++q)
                if (k + q < n && one(pj[q], wj[q])) lp.push(x[q]);
            e = en;
        }
        fc = lp.finish();
    } else {
        ListPush lp(va.lidx, va.lstride, i);
        const uint4 *row = rows + i;
#pragma unroll 1
        for (uint32_t k = 0; k < n; k += 4) {
            const uint4 e = *row;
            row += va.vstride;
            const bool v1 = k + 1 < n, v2 = k + 2 < n, v3 = k + 3 < n;
            const uint32_t j0 = e.x, j1 = v1 ? e.y : e.x, j2 = v2 ? e.z : e.x, j3 = v3 ? e.w : e.x;
            const float4 p0 = posm[j0], p1 = posm[j1], p2 = posm[j2], p3 = posm[j3];
            const uint32_t w0 = va.vword[j0], w1 = va.vword[j1], w2 = va.vword[j2], w3 = va.vword[j3];
            if (one(p0, w0)) lp.push(j0);
            if (v1 && one(p1, w1)) lp.push(j1);
            if (v2 && one(p2, w2)) lp.push(j2);
            if (v3 && one(p3, w3)) lp.push(j3);
        }
        fc = lp.finish();
    }
    // Rows whose neighbours within h are out of visiting order (one of them
    // changed cell): the wave puts each in order in turn.  Every lane takes
    // entries lane and lane + 64, ranks them by key (ties, only between
    // entries off the 27 cells, by entry index), and the row is stored back
    // in that order; then the owner sums in that order.
    const unsigned long long badm = __ballot(mine && bad);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    if (badm != 0ull && lane == 0) atomicAdd(&va.vs->qsum[va.par], (uint32_t)__popcll(badm));  // statistics
    for (unsigned long long mm = badm; mm; mm &= mm - 1ull) {
        const int b = __ffsll((long long)mm) - 1;
        const uint32_t ri = __shfl(i, b, 64), rn = __shfl(n, b, 64), rw = __shfl(wi, b, 64);
        const float px = __shfl(pi.x, b, 64), py = __shfl(pi.y, b, 64), pz = __shfl(pi.z, b, 64);
        const uint32_t *ru = reinterpret_cast<const uint32_t *>(halo ? va.rows16 : rows);
        // entries lane and lane + 64 (rn <= VL_CAP < 128): the entry (slot,
        // or halo index) with bit 31 set for a neighbour within h, its
        // contribution, its key
#pragma unroll
        for (uint32_t h2 = 0; h2 < 2; ++h2) {
            const uint32_t e = lane + 64u * h2;
            if (e < rn) {
                uint32_t x;
                float4 pj;
                uint32_t wj;
                if (halo) {
                    x = (ru[((size_t)(e >> 3) * va.vstride + ri) * 4u + ((e & 7u) >> 1)] >> (16u * (e & 1u))) &
                        0xFFFFu;
                    pj = h_p[x];
                    wj = h_w[x];
                } else {
                    x = ru[((size_t)(e >> 2) * va.vstride + ri) * 4u + (e & 3u)];
                    pj = posm[x];
                    wj = va.vword[x];
                }
                const uint32_t key = vl_key(vl_addend(rw), wj);
                const float r2 = dot3(pj.x - px, pj.y - py, pj.z - pz);
                const bool in = r2 < hg2 && key != VL_NOT27;
                s_key[w][e] = key;
                s_uj[w][e] = x | (in ? 0x80000000u : 0u);
                s_uc[w][e] = in ? pj.w * poly6(r2, s) : 0.0f;
            }
        }
        wave_sync();
        // each lane's two entries: rank, entry and contribution read before
        // any lane writes them back in rank order
        uint32_t rk[2] = {0u, 0u}, uj[2] = {0u, 0u};
        float uc[2] = {0.0f, 0.0f};
#pragma unroll
        for (uint32_t h2 = 0; h2 < 2; ++h2) {
            const uint32_t e = lane + 64u * h2;
            if (e < rn) {
                const uint32_t ke = s_key[w][e];
                uint32_t r = 0;
                for (uint32_t m = 0; m < rn; ++m) {
                    const uint32_t km = s_key[w][m];
                    r += (km < ke || (km == ke && m < e)) ? 1u : 0u;
                }
                rk[h2] = r;
                uj[h2] = s_uj[w][e];
                uc[h2] = s_uc[w][e];
            }
        }
        wave_sync();
#pragma unroll
        for (uint32_t h2 = 0; h2 < 2; ++h2) {
            if (lane + 64u * h2 < rn) {
                s_uj[w][rk[h2]] = uj[h2];
                s_uc[w][rk[h2]] = uc[h2];
            }
        }
        wave_sync();
        // the row in visiting order, one 16-byte group per lane
        if (halo) {
            const uint32_t g0 = lane * 8u;
            if (g0 < rn) {
                uint32_t v[4] = {0u, 0u, 0u, 0u};
                for (uint32_t q = 0; q < 8 && g0 + q < rn; ++q)
                    v[q >> 1] |= (s_uj[w][g0 + q] & 0xFFFFu) << (16u * (q & 1u));
                va.rows16[(size_t)lane * va.vstride + ri] = make_uint4(v[0], v[1], v[2], v[3]);
            }
        } else {
            const uint32_t g0 = lane * 4u;
            if (g0 < rn) {
                const uint32_t e0 = s_uj[w][g0] & 0x7FFFFFFFu;
                const uint32_t e1 = g0 + 1 < rn ? s_uj[w][g0 + 1] & 0x7FFFFFFFu : 0u,
                               e2 = g0 + 2 < rn ? s_uj[w][g0 + 2] & 0x7FFFFFFFu : 0u,
                               e3 = g0 + 3 < rn ? s_uj[w][g0 + 3] & 0x7FFFFFFFu : 0u;
                rows[(size_t)lane * va.vstride + ri] = make_uint4(e0, e1, e2, e3);
            }
        }
        if (lane == (uint32_t)b) {
            rho = 0.0f;
            cnt = 0;
            if (halo) {
                ListPush16 lp(va.lidx, va.lstride, ri);
                for (uint32_t m = 0; m < rn; ++m) {
                    const uint32_t e = s_uj[w][m];
                    if (e & 0x80000000u) {
                        ++cnt;
                        rho += s_uc[w][m];
                        lp.push(e & 0xFFFFu);
                    }
                }
                fc = lp.finish();
            } else {
                ListPush lp(va.lidx, va.lstride, ri);
                for (uint32_t m = 0; m < rn; ++m) {
                    const uint32_t e = s_uj[w][m];
                    if (e & 0x80000000u) {
                        ++cnt;
                        rho += s_uc[w][m];
                        lp.push(e & 0x7FFFFFFFu);
                    }
                }
                fc = lp.finish();
            }
        }
        wave_sync();
    }
    if (mine) {
        fcnt[i] = fc;
        density_out(s, scenes, t.scene, i, rho, cnt, tag, rq, rec_rho, rec_p, rec_count);
    }
}

// Density from a kept skin list: a judging substep of a surrogate world
// that keeps the particle order and the list of an earlier sort (the list
// radius covers the displacement since).  Entries within h in list order,
// which sums the same terms as the cell order up to rounding.
__global__ __launch_bounds__(PH_TILE) void k_density_skin(const TileDev *__restrict__ tiles,
                                                          const SceneDev *__restrict__ scenes,
                                                          const uint32_t *__restrict__ start,
                                                          const float4 *__restrict__ posm,
                                                          const uint32_t *__restrict__ tag,
                                                          const uint32_t *__restrict__ lcnt,
                                                          const uint4 *__restrict__ lgrp, uint32_t mstride,
                                                          float2 *__restrict__ rq, float *__restrict__ rec_rho,
                                                          float *__restrict__ rec_p, uint32_t *__restrict__ rec_count) {
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const TileDev t = tiles[tile];
    const SceneDev s = scenes[t.scene];
    const uint32_t i = t.first + threadIdx.x;
    if (i >= s.part_end) return;
    const float4 pi = posm[i];
    const uint32_t cw = lcnt[i];
    const float hg2 = s.h_grid2;
    float rho = 0.0f;
    uint32_t cnt = 0;
    auto entry = [&](const float4 &pj) {
        const float r2 = dot3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
        if (r2 < hg2) {
            ++cnt;
            rho += pj.w * poly6(r2, s);
        }
    };
    if (cw & SLF_NONE) {
        // non-finite position: no neighbours
    } else if (cw & SLF_SLOW) {
        int32_t cx, cy, cz;
        cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
        for_each_candidate(s, start, posm, cx, cy, cz, [&](uint32_t j) { entry(posm[j]); });
    } else {
        const uint32_t n = cw & SLF_COUNT;
        const uint4 *g = lgrp + i;
#pragma unroll 1
        for (uint32_t k = 0; k < n; k += 4) {
            const uint4 e = *g;
            g += mstride;
            // slots past the count read the particle itself, then are masked
            const uint32_t j1 = k + 1 < n ? e.y : i, j2 = k + 2 < n ? e.z : i, j3 = k + 3 < n ? e.w : i;
            const float4 p0 = posm[e.x], p1 = posm[j1], p2 = posm[j2], p3 = posm[j3];
            entry(p0);
            if (k + 1 < n) entry(p1);
            if (k + 2 < n) entry(p2);
            if (k + 3 < n) entry(p3);
        }
    }
    density_out(s, scenes, t.scene, i, rho, cnt, tag, rq, rec_rho, rec_p, rec_count);
}

namespace ph {

hipError_t launch_density_skin(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                               const uint32_t *start, const float4 *posm, const uint32_t *tag, const uint32_t *lcnt,
                               const uint32_t *lidx, uint32_t mstride, float2 *rq, float *rec_rho, float *rec_p,
                               uint32_t *rec_count) {
    hipLaunchKernelGGL(k_density_skin, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, start, posm, tag, lcnt,
                       reinterpret_cast<const uint4 *>(lidx), mstride, rq, rec_rho, rec_p, rec_count);
    return hipGetLastError();
}

hipError_t launch_density_verlet(hipStream_t st, uint32_t ntiles, const TileDev *tiles, const SceneDev *scenes,
                                 const uint32_t *start, const float4 *posm_sorted, const float4 *posm,
                                 const uint32_t *tag, float2 *rq, float *rec_rho, float *rec_p,
                                 uint32_t *rec_count, uint32_t *fcnt, const VerletArgs &va) {
    hipLaunchKernelGGL(k_density_verlet, dim3(ntiles), dim3(PH_TILE), 0, st, tiles, scenes, start, posm_sorted,
                       posm, tag, rq, rec_rho, rec_p, rec_count, fcnt, va);
    return hipGetLastError();
}

}  // namespace ph
