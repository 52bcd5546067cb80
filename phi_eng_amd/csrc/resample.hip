// resample.hip — adaptive resolution on the device (FluidState3::resample,
// crates/dynamics/src/fluid3.rs:923-987; the 2D twin fluid.rs:826-891 is
// the same pass in the z = 0 plane).
//
// The reference runs it on the CPU in particle index order:
//   merge  for i ascending, a particle denser than merge_hi absorbs the
//          highest-index earlier dense particle within h/4 that is not yet
//          absorbed (mass summed, velocity mass-weighted, fluid3.rs:941-953);
//          the absorbed ones are removed in order (:957-965);
//   split  in the new order, a particle thinner than split_lo and heavier
//          than min_mass becomes two halves at +h/4 and -h/4 along x while
//          the output stays below the cap (:967-985).
// On the device, over the cell-sorted state of this substep:
//   k_rs_find     per particle: the merge candidates (dense, lower index,
//                 within h/4; all lie in the 27 cells around it since
//                 h/4 < the cell size), each taker's run written once;
//   k_rs_resolve  one workgroup resolves the index-order greedy in rounds:
//                 an undecided taker takes its highest available candidate
//                 as soon as no lower undecided taker lists that candidate.
//                 Every item a taker lists can only be taken by a lower
//                 taker or itself, so this is exactly the sequential result;
//                 the lowest undecided taker always decides, so the rounds
//                 end.  Then the merged mass / velocity along each chain
//                 (i absorbs j which absorbed k ...), lowest first;
//   k_rs_mark     per index: kept, split-eligible;
//   two scans     (the bucket scan kernels) give each index its kept and
//                 eligible prefix counts;
//   k_rs_count    the splits accepted: eligible with kept + eligible before
//                 it + 1 < cap (out.len() + 1 < cap, :972; both prefixes grow,
//                 so the accepted ones are the first of the eligible);
//   k_rs_write    each kept particle to its new index (kept before it +
//                 accepted splits before it), a split as two.
// The result is the reference's particle vector in its order; the caller
// sorts it by cell again.  Cross-thread state inside k_rs_resolve is read
// and written with agent-scope atomics, so no wave reads a stale line.
#include "resample.h"
#include "sph_kernels.h"

namespace ph {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kResolveThreads = 1024;

__device__ __forceinline__ uint32_t ald(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float4 ald4(const float4 *p) {
    const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
    return make_float4(__uint_as_float(ald(q)), __uint_as_float(ald(q + 1)), __uint_as_float(ald(q + 2)),
                       __uint_as_float(ald(q + 3)));
}
__device__ __forceinline__ void ast4(float4 *p, float4 v) {
    uint32_t *q = reinterpret_cast<uint32_t *>(p);
    ast(q, __float_as_uint(v.x));
    ast(q + 1, __float_as_uint(v.y));
    ast(q + 2, __float_as_uint(v.z));
    ast(q + 3, __float_as_uint(v.w));
}

// fluid3.rs:936-940: a dense particle is one the merge scan does not skip
// (`density <= merge_hi` skips, so a NaN density would count as dense).
__device__ __forceinline__ bool dense(float rho, float merge_hi) { return !(rho <= merge_hi); }

__global__ __launch_bounds__(PH_TILE) void k_rs_find(ResampleArgs a) {
    const uint32_t i = blockIdx.x * PH_TILE + threadIdx.x;
    if (i >= a.n) return;
    a.flag[i] = 0u;
    if (!dense(a.rq[i].x, a.merge_hi)) return;
    const float4 pi = a.posm[i];
    if (!finite3(pi.x, pi.y, pi.z)) return;  // d2 would be NaN for every candidate
    const SceneDev s = a.scenes[0];
    const uint32_t id = a.tag[i] & PH_ID_MASK;
    int32_t cx, cy, cz;
    cell_of(pi.x, pi.y, pi.z, s.h_grid, cx, cy, cz);
    // fluid3.rs:941: d2 = (p_i - p_j).length_squared() < (h/4)^2
    auto cand = [&](uint32_t j) {
        if (!dense(a.rq[j].x, a.merge_hi) || (a.tag[j] & PH_ID_MASK) >= id) return false;
        const float4 pj = a.posm[j];
        return dot3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z) < a.md2;
    };
    uint32_t cnt = 0;
    for_each_candidate(s, a.start, a.posm, cx, cy, cz, [&](uint32_t j) { cnt += cand(j) ? 1u : 0u; });
    if (!cnt) return;
    const uint32_t k = (uint32_t)atomicAdd(&a.ctr[0], 1ull);
    const unsigned long long off = atomicAdd(&a.ctr[1], (unsigned long long)cnt);
    if (off + cnt > a.pcap) return;  // the store is too small: the host grows it and runs the pass again
    a.takers[k] = make_uint4(i, id, (uint32_t)off, cnt);
    a.flag[i] = (k + 1u) << 2;
    uint32_t w = (uint32_t)off;
    for_each_candidate(s, a.start, a.posm, cx, cy, cz, [&](uint32_t j) {
        if (cand(j)) a.items[w++] = j;
    });
}

__global__ __launch_bounds__(kResolveThreads) void k_rs_resolve(ResampleArgs a) {
    if (a.ctr[1] > a.pcap) return;
    const uint32_t T = (uint32_t)a.ctr[0];
    const uint32_t tid = threadIdx.x;
    // Each taker is owned by one thread: its candidate run, sorted by index
    // descending (the scan order of fluid3.rs:936), and its undecided state.
    for (uint32_t k = tid; k < T; k += kResolveThreads) {
        const uint4 t = a.takers[k];
        uint32_t *it = a.items + t.z;
        for (uint32_t x = 1; x < t.w; ++x) {
            const uint32_t v = it[x], vid = a.tag[v] & PH_ID_MASK;
            uint32_t y = x;
            while (y > 0 && (a.tag[it[y - 1]] & PH_ID_MASK) < vid) {
                it[y] = it[y - 1];
                --y;
            }
            it[y] = v;
        }
        ast(&a.pick[k], kNone);
        ast(&a.state[k], 0u);
        const float4 v = a.vel[t.x];
        ast4(&a.mv[k], make_float4(v.x, v.y, v.z, a.posm[t.x].w));
    }
    __syncthreads();
    // Decisions.  At most T rounds: the lowest undecided taker decides in
    // every round.
    for (uint32_t round = 0; round <= T; ++round) {
        for (uint32_t k = tid; k < T; k += kResolveThreads) {
            if (ald(&a.state[k])) continue;
            const uint4 t = a.takers[k];
            for (uint32_t x = 0; x < t.w; ++x) {
                const uint32_t j = a.items[t.z + x];
                if (!(ald(&a.flag[j]) & 1u)) ast(&a.minT[j], kNone);
            }
        }
        __syncthreads();
        for (uint32_t k = tid; k < T; k += kResolveThreads) {
            if (ald(&a.state[k])) continue;
            const uint4 t = a.takers[k];
            for (uint32_t x = 0; x < t.w; ++x) {
                const uint32_t j = a.items[t.z + x];
                if (!(ald(&a.flag[j]) & 1u))
                    __hip_atomic_fetch_min(&a.minT[j], t.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __syncthreads();
        uint32_t left = 0;
        for (uint32_t k = tid; k < T; k += kResolveThreads) {
            if (ald(&a.state[k])) continue;
            const uint4 t = a.takers[k];
            uint32_t j = kNone;
            for (uint32_t x = 0; x < t.w && j == kNone; ++x) {
                const uint32_t c = a.items[t.z + x];
                if (!(ald(&a.flag[c]) & 1u)) j = c;
            }
            if (j == kNone) {
                ast(&a.state[k], 1u);  // every candidate already absorbed: no merge
            } else if (ald(&a.minT[j]) == t.y) {
                __hip_atomic_fetch_or(&a.flag[j], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ast(&a.pick[k], j);
                ast(&a.state[k], 1u);
            } else {
                ++left;
            }
        }
        if (__syncthreads_count(left != 0) == 0) break;
    }
    // Merged values: taker i's mass and velocity combine its own with those
    // of the particle it absorbed after that one's own merge (fluid3.rs:
    // 943-952 reads j's current values).  A chain resolves from its lowest
    // taker up, one link per round.
    for (uint32_t round = 0; round <= T; ++round) {
        uint32_t left = 0;
        for (uint32_t k = tid; k < T; k += kResolveThreads) {
            if (ald(&a.state[k]) != 1u) continue;
            const uint32_t j = ald(&a.pick[k]);
            if (j == kNone) {
                ast(&a.state[k], 2u);
                continue;
            }
            const uint32_t fj = ald(&a.flag[j]) >> 2;
            float4 w;
            if (fj) {
                const uint32_t kj = fj - 1u;
                if (ald(&a.state[kj]) != 2u) {
                    ++left;
                    continue;
                }
                w = ald4(&a.mv[kj]);
            } else {
                const float4 vj = a.vel[j];
                w = make_float4(vj.x, vj.y, vj.z, a.posm[j].w);
            }
            float4 u = ald4(&a.mv[k]);
            const float mi = u.w, mj = w.w, mt = mi + mj;
            if (mt > PH_EPS) {
                u.x = (u.x * mi + w.x * mj) / mt;
                u.y = (u.y * mi + w.y * mj) / mt;
                u.z = (u.z * mi + w.z * mj) / mt;
                u.w = mt;
            }
            ast4(&a.mv[k], u);
            ast(&a.state[k], 2u);
        }
        if (__syncthreads_count(left != 0) == 0) break;
    }
}

// The particle's mass and velocity after the merge pass.
__device__ __forceinline__ void merged(const ResampleArgs &a, uint32_t s, float4 &x, float4 &v) {
    x = a.posm[s];
    v = a.vel[s];
    const uint32_t fl = a.flag[s] >> 2;
    if (fl && a.pick[fl - 1u] != kNone) {
        const float4 u = a.mv[fl - 1u];
        v.x = u.x;
        v.y = u.y;
        v.z = u.z;
        x.w = u.w;
    }
}

// A pass whose candidate store overflowed (ctr[1] > pcap) writes nothing
// after k_rs_find: its takers' candidate runs are incomplete, so the host
// grows the store and runs the whole pass again.  k_rs_mark leaves cK / cE
// at zero (the scans re-zero them), k_rs_count and k_rs_write return, and
// the particle arrays stay as they were.
__global__ __launch_bounds__(PH_TILE) void k_rs_mark(ResampleArgs a) {
    if (a.ctr[1] > a.pcap) return;
    const uint32_t s = blockIdx.x * PH_TILE + threadIdx.x;
    if (s >= a.n) return;
    const uint32_t id = a.tag[s] & PH_ID_MASK;
    a.slot_of[id] = s;
    const bool kept = !(a.flag[s] & 1u);
    float4 x, v;
    merged(a, s, x, v);
    a.cK[id] = kept ? 1u : 0u;
    // fluid3.rs:972
    a.cE[id] = kept && a.rq[s].x < a.split_lo && x.w > a.min_mass ? 1u : 0u;
}

__global__ __launch_bounds__(PH_TILE) void k_rs_count(ResampleArgs a) {
    if (a.ctr[1] > a.pcap) return;
    const uint32_t p = blockIdx.x * PH_TILE + threadIdx.x;
    if (p == 0) a.ctr[3] = a.rK[a.n];
    if (p >= a.n) return;
    const bool e = a.rE[p + 1] != a.rE[p];
    if (e && (uint64_t)a.rK[p] + a.rE[p] + 1u < a.cap) atomicAdd(&a.ctr[2], 1ull);
}

__global__ __launch_bounds__(PH_TILE) void k_rs_write(ResampleArgs a) {
    if (a.ctr[1] > a.pcap) return;
    // nothing merged or split: the caller keeps the sorted state as it is
    if (a.ctr[3] == a.n && a.ctr[2] == 0) return;
    const uint32_t p = blockIdx.x * PH_TILE + threadIdx.x;
    if (p >= a.n || a.rK[p + 1] == a.rK[p]) return;  // absorbed
    const uint32_t s = a.slot_of[p];
    const uint32_t acc_total = (uint32_t)a.ctr[2];
    const bool split = a.rE[p + 1] != a.rE[p] && (uint64_t)a.rK[p] + a.rE[p] + 1u < a.cap;
    const uint32_t nid = a.rK[p] + min(a.rE[p], acc_total);
    const uint32_t ph = a.tag[s] & ~PH_ID_MASK;
    float4 x, v;
    merged(a, s, x, v);
    if (split) {
        // fluid3.rs:973-983: halves at p + (h/4, 0, 0) then p - (h/4, 0, 0)
        const float half = x.w * 0.5f;
        a.posm_out[nid] = make_float4(x.x + a.off, x.y + 0.0f, x.z + 0.0f, half);
        a.posm_out[nid + 1] = make_float4(x.x - a.off, x.y - 0.0f, x.z - 0.0f, half);
        a.vel_out[nid] = v;
        a.vel_out[nid + 1] = v;
        a.tag_out[nid] = nid | ph;
        a.tag_out[nid + 1] = (nid + 1u) | ph;
    } else {
        a.posm_out[nid] = x;
        a.vel_out[nid] = v;
        a.tag_out[nid] = nid | ph;
    }
}

__global__ __launch_bounds__(PH_TILE) void k_export_mass(uint32_t n, const float4 *__restrict__ posm,
                                                         const uint32_t *__restrict__ tag, float *__restrict__ mass,
                                                         uint8_t *__restrict__ phase) {
    const uint32_t s = blockIdx.x * PH_TILE + threadIdx.x;
    if (s >= n) return;
    const uint32_t t = tag[s], id = t & PH_ID_MASK;
    mass[id] = posm[s].w;
    phase[id] = (uint8_t)(t >> 24);
}

uint32_t grid(uint32_t n) { return (n + PH_TILE - 1) / PH_TILE; }

}  // namespace

hipError_t launch_export_mass(hipStream_t st, uint32_t n, const float4 *posm, const uint32_t *tag, float *mass,
                              uint8_t *phase) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_export_mass, dim3(grid(n)), dim3(PH_TILE), 0, st, n, posm, tag, mass, phase);
    return hipGetLastError();
}

hipError_t launch_resample(hipStream_t st, const ResampleArgs &a) {
    if (a.n == 0) return hipSuccess;
    const uint32_t g = grid(a.n);
    hipError_t e = launch_fill_u32(st, reinterpret_cast<uint32_t *>(a.ctr), 8u, 0u);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_find, dim3(g), dim3(PH_TILE), 0, st, a);
    hipLaunchKernelGGL(k_rs_resolve, dim3(1), dim3(kResolveThreads), 0, st, a);
    hipLaunchKernelGGL(k_rs_mark, dim3(g), dim3(PH_TILE), 0, st, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int k = 0; k < 2; ++k) {
        uint32_t *c = k ? a.cE : a.cK, *r = k ? a.rE : a.rK;
        if ((e = launch_scan_reduce(st, c, a.n, a.partial)) != hipSuccess) return e;
        if ((e = launch_scan_partials(st, a.partial, a.n, r)) != hipSuccess) return e;
        if ((e = launch_scan_apply(st, c, a.n, a.partial, r)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_rs_count, dim3(g), dim3(PH_TILE), 0, st, a);
    hipLaunchKernelGGL(k_rs_write, dim3(g), dim3(PH_TILE), 0, st, a);
    return hipGetLastError();
}

}  // namespace ph
