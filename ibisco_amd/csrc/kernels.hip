// gfx950 kernels of the per-step path: cell sort and Verlet-list rebuild,
// tabulated nonbonded forces, bonded forces, leap-frog integration and the
// per-system reductions.  Launch wrappers are at the end of the file.
#include <hip/hip_runtime.h>

#include "ibx_kernels.h"
#include "kernels_api.h"

namespace ibx {

// ------------------------------------------------------------------------------------------------
// helpers

// SHIFT.f90:37-43 in FP64 with no contraction, so the folded coordinates are
// bit-identical to the oracle's.
__device__ __forceinline__ void fold3(const double s[3], const double box2[3], double r[3]) {
#pragma clang fp contract(off)
    bool out = false;
    for (int k = 0; k < 3; ++k) out |= (s[k] > box2[k]) || (s[k] < -box2[k]);
    if (out) {
        for (int k = 0; k < 3; ++k) {
            double x = s[k] / box2[k];
            double sp = x - 2.0 * trunc(x / 2.0);
            r[k] = (sp - 2.0 * trunc(sp)) * box2[k];
        }
    } else {
        r[0] = s[0]; r[1] = s[1]; r[2] = s[2];
    }
}

// Pair criterion of NEW_NEIGHBOUR_WITHLIST.f90:64-69 in FP64: d = ri - rj; d -= ANINT(d/L)*L; r2 = dx*dx + dy*dy + dz*dz, each operation rounded on its own as in
// NEW_NEIGHBOUR_WITHLIST.f90:64-69 (FP64 operands; no contraction into FMA)
__device__ __forceinline__ double exact_r2(const double4 a, const double4 b, double l0, double l1, double l2,
                                           double li0, double li1, double li2) {
#pragma clang fp contract(off)
    double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
    d0 = d0 - round(d0 * li0) * l0;
    d1 = d1 - round(d1 * li1) * l1;
    d2 = d2 - round(d2 * li2) * l2;
    double r2 = d0 * d0;
    r2 = r2 + d1 * d1;
    r2 = r2 + d2 * d2;
    return r2;
}

__device__ __forceinline__ double exact_r2(const double4 a, const double4 b, const DevSys &S) {
    return exact_r2(a, b, S.box[0], S.box[1], S.box[2], S.boxinv[0], S.boxinv[1], S.boxinv[2]);
}

// Compiler-only barrier: loads above it stay above it.  Without it the
// compiler sinks a load below the next branch, into the block that uses it,
// and the kernel waits for it there instead of while it checks other things.
// The pointers a load reads through are passed in (they are uniform), so that
// restrict-qualified arguments do not exempt their loads.
__device__ __forceinline__ void issue_here() { asm volatile("" ::: "memory"); }
__device__ __forceinline__ void issue_here(const void *a, const void *b, const void *c) {
    asm volatile("" ::"s"(a), "s"(b), "s"(c) : "memory");
}

// Cross-lane moves inside each row of 16 lanes through DPP (VALU operand
// modifiers, no LDS traffic): the two quad permutations exchange lanes i and
// i^1, i and i^2; the mirrors take lane 7-i of each half row and 15-i of each
// row.  After the first two, each quad holds one value, so the mirrors pair
// every lane with the other quad (half row) and the other half row, as i^4 and
// i^8 would.  Every lane of the wave must be active.
constexpr int kDppXor1 = 0xb1;         // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4e;         // quad_perm [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;  // row_half_mirror
constexpr int kDppMirror = 0x140;      // row_mirror

template <int kCtrl>
__device__ __forceinline__ float dpp_f(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), kCtrl, 0xf, 0xf, false));
}

template <int kCtrl>
__device__ __forceinline__ double dpp_d(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, kCtrl, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), kCtrl, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ float lane_f(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

__device__ __forceinline__ double lane_d(double x, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes of a wave in a fixed order (pairs, quads, half rows,
// rows, then the four rows in order); every lane gets the same result.
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_d<kDppXor1>(x);
    x += dpp_d<kDppXor2>(x);
    x += dpp_d<kDppHalfMirror>(x);
    x += dpp_d<kDppMirror>(x);
    return (lane_d(x, 0) + lane_d(x, 16)) + (lane_d(x, 32) + lane_d(x, 48));
}

// Sum over aligned groups of kN lanes (kN = 1, 2, 4, 8 or 16), the same
// butterfly order as i^1, i^2, i^4, i^8; every lane of a group gets the sum.
template <int kN>
__device__ __forceinline__ float group_sum(float x) {
    static_assert(kN == 1 || kN == 2 || kN == 4 || kN == 8 || kN == 16, "group of 1 to 16 lanes");
    if (kN > 1) x += dpp_f<kDppXor1>(x);
    if (kN > 2) x += dpp_f<kDppXor2>(x);
    if (kN > 4) x += dpp_f<kDppHalfMirror>(x);
    if (kN > 8) x += dpp_f<kDppMirror>(x);
    return x;
}

// Deterministic block reduction of n doubles per thread; thread 0 gets the sums.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double *lds /* N*4 */) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double x = wave_sum(v[k]);
        if (lane == 0) lds[k * 4 + wid] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = ((lds[k * 4] + lds[k * 4 + 1]) + lds[k * 4 + 2]) + lds[k * 4 + 3];
    }
    __syncthreads();
}

constexpr int kMultiTabsLds(bool multi) { return multi ? kMaxNbTabs : 1; }
constexpr int kNbFields = 9;   // nonbonded partial fields: V_NB of three classes, virial xx yy zz xy xz yz

// Workgroups are dispatched to the XCDs in turn, and each XCD has its own L2.
// A kernel over n spatially ordered blocks is launched with xcd_grid(n)
// workgroups and works on block xcd_block(n), so each XCD runs one contiguous
// range of blocks (one region of the system) and finds its neighbours' data in
// its own L2.  Blocks at or past n return at once.
__device__ __forceinline__ int xcd_block(int n, int b) {
    const int per = (n + kXcd - 1) / kXcd;
    return (b % kXcd) * per + b / kXcd;
}
__device__ __forceinline__ int xcd_block(int n) { return xcd_block(n, (int)blockIdx.x); }
static inline int xcd_grid(int n) { return ((n + kXcd - 1) / kXcd) * kXcd; }

// Position word of slot j, addressed by a 32-bit byte offset from the array
// base so the gather keeps the base in scalar registers.
__device__ __forceinline__ int4 ld_pos(const int4 *__restrict__ p, int j) {
    return *reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(p) + (uint32_t)j * 16u);
}

__device__ __forceinline__ int pos_type(int w) { return w & 0xffff; }
__device__ __forceinline__ int pos_label(int w) { return (w >> 16) & 0xff; }

// ------------------------------------------------------------------------------------------------
// list rebuild (replaces LINKS.f90, MAPS.f90, NEW_NEIGHBOUR_WITHLIST.f90)

// Squared displacement, in working units^2, between two fixed-point positions.
__device__ __forceinline__ float fixed_d2(const int4 a, const int4 b, float rho1, float rho2) {
    const float dx = (float)(int)((uint32_t)a.x - (uint32_t)b.x);
    const float dy = (float)(int)((uint32_t)a.y - (uint32_t)b.y) * rho1;
    const float dz = (float)(int)((uint32_t)a.z - (uint32_t)b.z) * rho2;
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

__device__ __forceinline__ float fixed_d2(const int4 a, const int4 b, const DevSys &S) {
    return fixed_d2(a, b, S.rho[1], S.rho[2]);
}

// Largest value over a block of kBlock threads, into *dst (float bits, atomicMax on
// non-negative floats orders like the values).  Every thread must call it.
__device__ __forceinline__ void block_max_to(float x, int *dst, float *lds /* kBlock / kWave */) {
    x = fmaxf(x, dpp_f<kDppXor1>(x));
    x = fmaxf(x, dpp_f<kDppXor2>(x));
    x = fmaxf(x, dpp_f<kDppHalfMirror>(x));
    x = fmaxf(x, dpp_f<kDppMirror>(x));
    x = fmaxf(fmaxf(lane_f(x, 0), lane_f(x, 16)), fmaxf(lane_f(x, 32), lane_f(x, 48)));
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = lds[0];
        for (int w = 1; w < kBlock / kWave; ++w) m = fmaxf(m, lds[w]);
        if (m > 0.f) atomicMax(dst, __float_as_int(m));
    }
    __syncthreads();
}

__device__ void tune_skin(DevSys *sys, int nsys, const int *dmax, int nmove, int tune, int t);

// Cell of every member from its folded FP64 position (LINKS.f90:16-18, clamped).
// Workgroups past nwg run the list-skin tuner instead (one thread per system
// and group): it writes each group's inner radius and displacement bound, which
// the cell count does not read, so the two share a launch.
__global__ void __launch_bounds__(kBlock) k_cell_count(const DevSys *sys, const int *__restrict__ wg_sys, int nwg,
                                                       int g, const int *__restrict__ mem, const double *__restrict__ s,
                                                       int nitems_total, int *__restrict__ mem_cell,
                                                       int *__restrict__ mem_rank, int *__restrict__ cell_count,
                                                       DevSys *tune_sys, int nsys,
                                                       const int *__restrict__ dmax_all, int nmove, int tune,
                                                       int only) {
    if ((int)blockIdx.x >= nwg) {
        tune_skin(tune_sys, nsys, dmax_all, nmove, tune, ((int)blockIdx.x - nwg) * kBlock + threadIdx.x);
        return;
    }
    if (only >= 0 && wg_sys[blockIdx.x] != only) return;
    const DevSys &S = sys[wg_sys[blockIdx.x]];
    if (!alive_at(S, *S.step_ctr + 1)) return;   // a stopped system's cells stay empty
    const DevGroup &G = S.g[g];
    int m = blockIdx.x * kBlock + threadIdx.x;
    int it = mem[m];
    if (it < 0) return;
    double sp[3] = {s[it], s[nitems_total + it], s[2 * nitems_total + it]}, r[3];
    fold3(sp, S.box2, r);
    int c[3];
    for (int k = 0; k < 3; ++k) {
        int ck = (int)((r[k] * S.boxinv[k] + 0.5) * G.nc[k]);
        c[k] = ck < 0 ? 0 : (ck >= G.nc[k] ? G.nc[k] - 1 : ck);
    }
    int cell = (c[0] * G.nc[1] + c[1]) * G.nc[2] + c[2];
    mem_cell[m] = cell;
    mem_rank[m] = atomicAdd(&cell_count[G.cell_off + cell], 1);   // the member's place among its cell's arrivals
}

// Exclusive scan of the cell counts into cell starts (n + 1 of them: start[c + 1]
// is where cell c ends) in two passes over tiles of kBlock cells, so that every
// pass is spread over many workgroups (the counts of a large grid take one
// workgroup most of its time just to read): k_scan_tiles scans each tile,
// writes its total and clears the counts for the next rebuild, k_scan_mid turns
// the tile totals (one per kBlock cells) into their prefix in one workgroup, and
// k_scan_add adds each tile's prefix.  Integer sums, so the order is free.
__global__ void __launch_bounds__(kBlock) k_scan_tiles(int *__restrict__ count, int *__restrict__ start,
                                                      int *__restrict__ tile_sum, int n) {
    __shared__ int wtot[kBlock / kWave];
    const int i = blockIdx.x * kBlock + threadIdx.x, lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int c = i < n ? count[i] : 0;
    if (i < n) count[i] = 0;
    int x = c;   // inclusive scan over the wave
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wtot[wid] = x;
    __syncthreads();
    int pre = 0;
    for (int w = 0; w < wid; ++w) pre += wtot[w];
    if (i <= n) start[i] = pre + x - c;
    if (threadIdx.x == kBlock - 1) tile_sum[blockIdx.x] = pre + x;
}

// The tile sums replaced by their exclusive prefix, in one workgroup: each thread sums a contiguous run of
// tiles, the runs' totals are scanned across the workgroup, and each thread rewrites its run (linear in the
// number of tiles).
__global__ void __launch_bounds__(kBlock) k_scan_mid(int *__restrict__ tile_sum, int ntile) {
    __shared__ int wtot[kBlock / kWave];
    constexpr int U = 8;   // loads in flight per thread (indices clamped into the run, values past it dropped)
    const int per = (ntile + kBlock - 1) / kBlock;
    const int t0 = threadIdx.x * per, t1 = min(t0 + per, ntile);
    int c = 0;
    for (int t = t0; t < t1; t += U) {
        int v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = tile_sum[min(t + u, t1 - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) c += t + u < t1 ? v[u] : 0;
    }
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    int x = c;   // inclusive scan over the wave
    for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wtot[wid] = x;
    __syncthreads();
    int run = x - c;
    for (int w = 0; w < wid; ++w) run += wtot[w];
    for (int t = t0; t < t1; t += U) {
        int v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = tile_sum[min(t + u, t1 - 1)];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (t + u < t1) {
                tile_sum[t + u] = run;
                run += v[u];
            }
        }
    }
}

// Adds each tile's offset (the exclusive prefix of the tile sums) to its starts.
__global__ void __launch_bounds__(kBlock) k_scan_add(int *__restrict__ start, const int *__restrict__ tile_sum, int n) {
    const int b = blockIdx.x;
    if (b == 0) return;
    const int i = b * kBlock + threadIdx.x;
    if (i <= n) start[i] += tile_sum[b];
}

__global__ void __launch_bounds__(kBlock) k_place(const DevSys *__restrict__ sys, const int *__restrict__ wg_sys, int g,
                                                  const int *__restrict__ mem, const int *__restrict__ mem_cell,
                                                  const int *__restrict__ cell_start, const int *__restrict__ mem_rank,
                                                  int *__restrict__ tmp, int only) {
    if (only >= 0 && wg_sys[blockIdx.x] != only) return;
    const DevSys &S = sys[wg_sys[blockIdx.x]];
    if (!alive_at(S, *S.step_ctr + 1)) return;
    const DevGroup &G = S.g[g];
    int m = blockIdx.x * kBlock + threadIdx.x;
    if (mem[m] < 0) return;
    int gc = G.cell_off + mem_cell[m];
    int local = cell_start[gc] - cell_start[G.cell_off];
    const int pos = mem_rank[m];
    tmp[G.slot_off + local + pos] = m;
}

// Orders each cell by member index so the sorted layout (and every sum over
// it) is deterministic.  One workgroup per cell.
constexpr int kSortLanes = 4;    // lanes per cell in the cell sort (cells hold about one member on average)

__global__ void __launch_bounds__(kBlock) k_sort_cells(const int *__restrict__ cell_sys, const DevSys *__restrict__ sys, int g,
                                                       const int *__restrict__ cell_start,
                                                       const int *__restrict__ tmp, const int *__restrict__ mem,
                                                       int *__restrict__ perm, int *__restrict__ slot_cell,
                                                       int *__restrict__ flags, int ncell, int only) {
    // kSortLanes lanes per cell (cells hold a few members each), kBlock / kSortLanes cells per workgroup
    constexpr int kCap = 64;   // members of a cell ranked in LDS; a more crowded cell is ranked in global memory
    __shared__ int vals[kBlock / kSortLanes][kCap];
    const int w = threadIdx.x / kSortLanes, lane = threadIdx.x & (kSortLanes - 1);
    const int gc = blockIdx.x * (kBlock / kSortLanes) + w;
    int n = 0, base = 0, lc = 0;
    if (gc < ncell && (only < 0 || cell_sys[gc] == only)) {
        const DevSys &S = sys[cell_sys[gc]];
        const DevGroup &G = S.g[g];
        n = cell_start[gc + 1] - cell_start[gc];
        base = G.slot_off + cell_start[gc] - cell_start[G.cell_off];
        lc = gc - G.cell_off;
    }
    if (n > kCap) {   // too crowded for LDS: rank against the arrival copy in global memory (same order)
        for (int i = lane; i < n; i += kSortLanes) {
            const int v = tmp[base + i];
            int rank = 0;
            for (int k = 0; k < n; ++k) rank += tmp[base + k] < v;
            perm[base + rank] = mem[v];
            slot_cell[base + rank] = lc;
        }
        n = 0;
    }
    for (int i = lane; i < n; i += kSortLanes) vals[w][i] = tmp[base + i];
    __syncthreads();
    for (int i = lane; i < n; i += kSortLanes) {
        const int v = vals[w][i];
        int rank = 0;
        for (int k = 0; k < n; ++k) rank += vals[w][k] < v;
        perm[base + rank] = mem[v];
        slot_cell[base + rank] = lc;
    }
}

// Sorted copies of the positions: folded FP64 (list criterion) and fixed point
// (nonbonded kernel), plus the item -> slot map.
__global__ void __launch_bounds__(kBlock) k_pack(const DevSys *__restrict__ sys, const int *__restrict__ wg_sys, int g,
                                                 const int *__restrict__ perm, const double *__restrict__ s,
                                                 int nitems_total, const int *__restrict__ itype,
                                                 const int *__restrict__ label, double4 *__restrict__ r64,
                                                 int4 *__restrict__ pos4, int *__restrict__ inv_perm,
                                                 int4 *__restrict__ pos4_ref, int *__restrict__ dmax, int nsys,
                                                 int dmax_value, int only) {
    const DevSys &S = sys[wg_sys[blockIdx.x]];
    int slot = blockIdx.x * kBlock + threadIdx.x;
    // the group's displacement bound restarts (every system has a block, so slots 0..nsys-1 exist)
    if (slot < nsys && (only < 0 || slot == only)) dmax[slot] = dmax_value;
    if (only >= 0 && wg_sys[blockIdx.x] != only) return;
    if (!alive_at(S, *S.step_ctr + 1)) return;   // a stopped system keeps its sorted copies
    int it = perm[slot];
    if (it < 0) {
        pos4[slot] = make_int4(0, 0, 0, -1);
        if (pos4_ref) pos4_ref[slot] = make_int4(0, 0, 0, -1);
        r64[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    double sp[3] = {s[it], s[nitems_total + it], s[2 * nitems_total + it]}, r[3];
    fold3(sp, S.box2, r);
    r64[slot] = make_double4(r[0], r[1], r[2], 0.0);
    const int4 p = make_int4((int)to_fixed(sp[0], S.inv_box_frac[0]), (int)to_fixed(sp[1], S.inv_box_frac[1]),
                             (int)to_fixed(sp[2], S.inv_box_frac[2]), itype[it] | (label[it] << 16));
    pos4[slot] = p;
    if (pos4_ref) pos4_ref[slot] = p;   // positions of the rebuild, for the displacement bound
    inv_perm[it] = slot;
}

// RDF bin of a pair at squared distance r2u (working units^2), inv_dr_u bins per working unit.
__device__ __forceinline__ int rdf_bin(float r2u, float inv_dr_u) { return (int)(__builtin_amdgcn_sqrtf(r2u) * inv_dr_u); }

// Gap along one axis between a point at fraction f of its cell and the cell d
// cells away (cell edge a); 0 for its own cell.
__device__ __forceinline__ float cell_gap(int d, float f, float a) {
    return d > 0 ? ((float)d - f) * a : (d < 0 ? ((float)(-d - 1) + f) * a : 0.f);
}

// Full Verlet list of every group member (replaces NEW_NEIGHBOUR_WITHLIST.f90:5-127
// with LINKS/MAPS): every member j != i with r^2 < rlist^2, not excluded, not
// VS-VS.  The set of unordered pairs equals the reference's half list bit for
// bit: a fixed-point test settles pairs far from the boundary and the FP64
// criterion of the reference decides the rest.
//
// Cells have an edge of at least rlist/kCellDiv in x and y and rlist/kCellDivZ in
// z, and are stored x-major, so the cells of one (x, y) column are consecutive
// slots.  For member i the columns within kReach cells whose distance from i is
// below rlist give at most two slot ranges each (the z extent reachable from i,
// within kReachZ cells, split where it wraps), and
// kListLanes lanes walk their concatenation, 16 candidates per step.  Hits are
// compacted with a wave ballot, so each list is in candidate order and
// deterministic.  Pairs closer than rc + delta when the list is built fill the
// front run, the others the back run.  With fewer than min_cells(k) cells along
// axis k every member scans the whole group once.
template <bool kCubic, bool kRdf>
// Held to 64 VGPRs (eight waves per SIMD, measured faster than seven: the candidate loop waits on one position
// gather per round, and more waves hide more of it).
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
k_build_list(const DevSys *__restrict__ sys, const int *__restrict__ wg_sys,
                                                       int g, const int *__restrict__ perm,
                                                       const int *__restrict__ inv_perm,
                                                       const int4 *__restrict__ pos4, const double4 *__restrict__ r64,
                                                       const int *__restrict__ cell_start,
                                                       const int *__restrict__ excl_off, const int *__restrict__ excl,
                                                       int cap, int *__restrict__ list, int2 *__restrict__ nneb,
                                                       int *__restrict__ flags, int nblk, int only,
                                                       unsigned long long *__restrict__ rdf_hist,
                                                       int *__restrict__ rdf_samples, double *__restrict__ rdf_ivol,
                                                       int nbin, float inv_dr) {
    // range r of a member: {candidates up to its end, slot of candidate p in it minus p} (one LDS read)
    __shared__ int2 rng[kListAtoms][kListRanges];
    extern __shared__ unsigned int rdf_h[];   // nbin counts of this workgroup's listed pairs (RDF on)
    const int blk = xcd_block(nblk);
    if (blk >= nblk) return;
    const int slot0 = blk * kListAtoms;
    const int s_id = wg_sys[slot0 / kBlock];
    if (only >= 0 && s_id != only) return;
    const DevSys &S = sys[s_id];
    if (!alive_at(S, *S.step_ctr + 1)) return;
    const DevGroup &G = S.g[g];
    constexpr bool rdf = kRdf;   // the RDF of the listed pairs (kRdf: rdf_hist holds the histograms)
    if (rdf) {
        for (int b = threadIdx.x; b < nbin; b += kBlock) rdf_h[b] = 0;
        if (threadIdx.x == 0 && slot0 == G.slot_off) {   // one sample a build, and its 1/volume (NPT: the box varies)
            atomicAdd(&rdf_samples[s_id], 1);
            atomicAdd(&rdf_ivol[s_id], 1.0 / S.volume);
        }
    }
    const float inv_dr_u = (float)S.unit * inv_dr;   // bins in this system's working units
    const int sub = threadIdx.x & (kListLanes - 1), grp = threadIdx.x / kListLanes;
    const int i = slot0 + grp;
    const int I = perm[i];
    const bool valid = I >= 0;
    const int slot_off = G.slot_off, cell_off = G.cell_off;
    const int cs0 = cell_start[cell_off];
    const double box0 = S.box[0], box1 = S.box[1], box2 = S.box[2];
    const double bi0 = S.boxinv[0], bi1 = S.boxinv[1], bi2 = S.boxinv[2];
    const double rl2 = G.rl2;
    const bool hybrid = S.hybrid != 0;
    const float rho0 = S.rho[0], rho1 = S.rho[1], rho2 = S.rho[2];
    const float lo = G.rl2u_lo, hi = G.rl2u_hi, rin2u = G.rin2u;

    int4 pi = make_int4(0, 0, 0, 0);
    double4 ri = make_double4(0, 0, 0, 0);
    if (valid) {
        pi = pos4[i];
        ri = r64[i];
    }
    int ex0 = 0, ex1 = 0, li = 0;
    int ex[8] = {-1, -1, -1, -1, -1, -1, -1, -1};   // slots of the first exclusions
    if (valid) {
        li = pos_label(pi.w);
        ex0 = excl_off[I];
        ex1 = excl_off[I + 1];
        // unconditional loads, eight in flight at a time, while the candidate ranges are worked
        // out: entries past the member's read entry 0 of the array (which is allocated but holds
        // nothing when no item has exclusions), and what they read is replaced by item 0 before
        // it indexes inv_perm
        int xi[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) xi[e] = excl[excl_entry(ex0, e, ex1)];
#pragma unroll
        for (int e = 0; e < 8; ++e) xi[e] = inv_perm[excl_item(ex0, e, ex1, xi[e])];
#pragma unroll
        for (int e = 0; e < 8; ++e) ex[e] = ex0 + e < ex1 ? xi[e] : -1;
    }
    int T = 0;   // candidates of this member
    if (G.all_pairs) {
        if (sub == 0) {
            rng[grp][0] = make_int2(G.nmem, slot_off);
        }
        T = valid ? G.nmem : 0;
    } else {
        const int nc0 = G.nc[0], nc1 = G.nc[1], nc2 = G.nc[2];
        const double rv[3] = {ri.x, ri.y, ri.z}, bi[3] = {bi0, bi1, bi2};
        const int ncv[3] = {nc0, nc1, nc2};
        int c[3];
        float f[3], a[3];
        for (int k = 0; k < 3; ++k) {   // home cell as k_cell_count assigns it, and the position inside it
            const double u = (rv[k] * bi[k] + 0.5) * ncv[k];
            int ck = (int)u;
            ck = ck < 0 ? 0 : (ck >= ncv[k] ? ncv[k] - 1 : ck);
            c[k] = ck;
            f[k] = fminf(fmaxf((float)(u - ck), 0.f), 1.f);
            a[k] = (float)(S.box[k] / ncv[k]);
        }
        // reach: rlist with a margin, so a pruned cell is always beyond rlist
        const float reach2 = (float)(rl2 * (1.0 + 1e-4));
        // the kStencil x kStencil columns around the member's cell, column q in group q / kListLanes handled by
        // lane q % kListLanes; each column is one z range (two where it wraps)
        constexpr int kCols = kStencil * kStencil, kGroups = (kCols + kListLanes - 1) / kListLanes;
        int st[2 * kGroups], len[2 * kGroups], np = 0;   // pieces of this lane's columns, in column order
        int n_grp[kGroups];
#pragma unroll
        for (int h = 0; h < kGroups; ++h) {
            n_grp[h] = 0;
            const int q = sub + h * kListLanes;
            if (!valid || q >= kCols) continue;
            const int dx = q / kStencil - kReach, dy = q % kStencil - kReach;
            const float gx = cell_gap(dx, f[0], a[0]), gy = cell_gap(dy, f[1], a[1]);
            const float gxy = gx * gx + gy * gy;
            if (gxy >= reach2) continue;
            int dz0 = 0, dz1 = 0;
            for (int d = -kReachZ; d < 0; ++d) {
                const float gz = cell_gap(d, f[2], a[2]);
                if (gxy + gz * gz < reach2) { dz0 = d; break; }
            }
            for (int d = kReachZ; d > 0; --d) {
                const float gz = cell_gap(d, f[2], a[2]);
                if (gxy + gz * gz < reach2) { dz1 = d; break; }
            }
            int x = c[0] + dx, y = c[1] + dy;
            x += x < 0 ? nc0 : 0; x -= x >= nc0 ? nc0 : 0;
            y += y < 0 ? nc1 : 0; y -= y >= nc1 ? nc1 : 0;
            const int col = cell_off + (x * nc1 + y) * nc2;
            const int z0 = c[2] + dz0, z1 = c[2] + dz1;
            int za[2], zb[2], nz = 0;
            if (z0 < 0) { za[0] = z0 + nc2; zb[0] = nc2 - 1; za[1] = 0; zb[1] = z1; nz = 2; }
            else if (z1 >= nc2) { za[0] = z0; zb[0] = nc2 - 1; za[1] = 0; zb[1] = z1 - nc2; nz = 2; }
            else { za[0] = z0; zb[0] = z1; nz = 1; }
            for (int k = 0; k < nz; ++k) {
                const int b = cell_start[col + za[k]];
                const int e = cell_start[col + zb[k] + 1];
                st[np] = slot_off + b - cs0;
                len[np] = e - b;
                ++np;
                ++n_grp[h];
            }
        }
        // place the pieces: group-0 columns by lane, then group-1 columns by lane, ...
        int lsum[kGroups], pos[kGroups], cum[kGroups], tot_n = 0, tot_c = 0;
        {
            int k = 0;
#pragma unroll
            for (int h = 0; h < kGroups; ++h) {
                lsum[h] = 0;
                for (int u = 0; u < n_grp[h]; ++u, ++k) lsum[h] += len[k];
            }
        }
#pragma unroll
        for (int h = 0; h < kGroups; ++h) {
            int xn = n_grp[h], xc = lsum[h];
            for (int off = 1; off < kListLanes; off <<= 1) {
                const int yn = __shfl_up(xn, off, kListLanes), yc = __shfl_up(xc, off, kListLanes);
                if (sub >= off) { xn += yn; xc += yc; }
            }
            pos[h] = tot_n + xn - n_grp[h];
            cum[h] = tot_c + xc - lsum[h];
            tot_n += __shfl(xn, kListLanes - 1, kListLanes);
            tot_c += __shfl(xc, kListLanes - 1, kListLanes);
        }
        {
            int k = 0;
#pragma unroll
            for (int h = 0; h < kGroups; ++h) {
                for (int u = 0; u < n_grp[h]; ++u, ++k) {
                    const int r = pos[h]++;
                    if (r < kListRanges) {
                        const int rb = st[k] - cum[h];
                        cum[h] += len[k];
                        rng[grp][r] = make_int2(cum[h], rb);
                    }
                }
            }
        }
        T = tot_c;
        if (tot_n > kListRanges) T = 0;   // cannot happen: kListRanges holds every column split in two
    }
    __syncthreads();

    int tmax = T;   // the wave walks as far as its longest member
    for (int off = kListLanes; off < kWave; off <<= 1) tmax = max(tmax, __shfl_xor(tmax, off, kWave));
    // exclusion registers worth testing: the wave's largest count (entries past a member's own are -1)
    int nex = valid ? min(ex1 - ex0, 8) : 0;
    for (int off = 1; off < kWave; off <<= 1) nex = max(nex, __shfl_xor(nex, off, kWave));
    nex = __builtin_amdgcn_readfirstlane(nex);
    const int gshift = (threadIdx.x & (kWave - 1)) & ~(kListLanes - 1);   // this group's bits in a wave ballot
    const unsigned below = (1u << sub) - 1u;
    int *const lrow = list + (valid ? list_pos(i, 0, cap) : 0);   // entry k of the row: lrow[kofs(k)]
    int n_in = 0, n_out = 0, r = 0;
    // slot of candidate p (p ascending: the range walk only moves forward); the lane's current range stays in
    // registers, so LDS is read only when a candidate lies past its end
    int2 cur = rng[grp][0];
    auto cand_slot = [&](int p) {
        while (p >= cur.x) cur = rng[grp][++r];
        return cur.y + p;
    };
    for (int t0 = 0; t0 < tmax; t0 += kListLanes) {
        const int p = t0 + sub;
        int j = -1;
        int4 pj = make_int4(0, 0, 0, 0);
        bool hit = false, inner = false;
        float r2u = 0.f;
        if (p < T) {
            j = cand_slot(p);
            pj = ld_pos(pos4, j);
            float dx = (float)(int)((uint32_t)pi.x - (uint32_t)pj.x);
            float dy = (float)(int)((uint32_t)pi.y - (uint32_t)pj.y);
            float dz = (float)(int)((uint32_t)pi.z - (uint32_t)pj.z);
            if (!kCubic) { dx *= rho0; dy *= rho1; dz *= rho2; }
            r2u = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            bool cand = r2u < lo;
            // the member's own FP64 position is read again here (rare) rather than held in registers
            if (r2u >= lo && r2u <= hi) cand = exact_r2(r64[i], r64[j], box0, box1, box2, bi0, bi1, bi2) < rl2;
            // excluded: the member itself, and its exclusions -- the first eight held in registers (unused ones
            // -1) and compared without branches, four of them when no member of the wave has more
            bool skip = j == i;
            if (nex <= 4) {
#pragma unroll
                for (int e = 0; e < 4; ++e) skip |= ex[e] == j;
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) skip |= ex[e] == j;
            }
            if (hybrid && li == 3) skip |= pos_label(pj.w) == 3;   // VS-VS
            if (cand && !skip && ex1 - ex0 > 8)
                for (int e = ex0 + 8; e < ex1; ++e) skip |= (inv_perm[excl[e]] == j);
            hit = cand && !skip;
            inner = r2u < rin2u;
        }
        const unsigned gi = (unsigned)(__ballot(hit && inner) >> gshift) & 0xFFFFu;
        const unsigned go = (unsigned)(__ballot(hit && !inner) >> gshift) & 0xFFFFu;
        // entry k of the row: the front run from 0 up, the back run from cap - 1 down (both formed, one selected)
        const int k_in = n_in + __popc(gi & below);
        const int k_out = cap - 1 - n_out - __popc(go & below);
        const int k = inner ? k_in : k_out;
        if (hit && (unsigned)k < (unsigned)cap) lrow[((k >> 2) << 6) + (k & 3)] = j;   // list_pos(i, k, cap)
        if (rdf && hit) {
            const int b = rdf_bin(r2u, inv_dr_u);
            if (b < nbin) atomicAdd(&rdf_h[b], 1u);
        }
        n_in += __popc(gi);
        n_out += __popc(go);
    }
    if (valid && sub == 0) {
        nneb[i] = make_int2(n_in, n_out);
        if (n_in + n_out > cap) atomicMax(&flags[0], n_in + n_out);
    }
    if (rdf) {
        __syncthreads();
        for (int b = threadIdx.x; b < nbin; b += kBlock)
            if (rdf_h[b]) atomicAdd(&rdf_hist[(size_t)s_id * nbin + b], (unsigned long long)rdf_h[b]);
    }
}

// ------------------------------------------------------------------------------------------------
// nonbonded forces (replaces NONBONDED_FORCE.F90:11-109)
//
// One thread per sorted slot walks its full list; positions are int32 fixed
// point, so the minimum image is the wrap-around of the integer difference.
// Tables (F/r and V, interleaved with the next row) are staged in LDS.  Each
// pair is visited from both ends; energies and virial are halved.

// What a pair needs besides the two positions: the tables, the FP32 cutoff
// band, and what the exact cutoff test reads.
struct NbCtx {
    const float4 *rows;            // {F_k, V_k, F_k+1 - F_k, V_k+1 - V_k}
    const float *xoff;             // x_k/bin - k of non-uniform tables (global rows)
    const NbTab *tabs;
    const int *map;
    int ntype, xoff_base;
    NbTab t0;
    float rho[3];                  // L_k / L_x
    double rhod[3];                // the same in FP64 (non-cubic boxes)
    float band_lo, band_hi;        // FP32 r^2 below band_lo is inside rc, at or above band_hi outside
    // exact cutoff test
    const DevSys *S;
    const double *s;
    const int *perm;
    int n, I, g;
};

// r^2 < rc^2 decided as NONBONDED_FORCE.F90:58-63 does, from the folded FP64
// positions (SHIFT.f90:37-43) of items I and J.  Rare: only pairs within FP32
// rounding of the cutoff get here.
__device__ __noinline__ bool nb_inside_exact(const DevSys *__restrict__ S, int g, const double *__restrict__ s,
                                             int n, int I, int J) {
    const double si[3] = {s[I], s[n + I], s[2 * n + I]}, sj[3] = {s[J], s[n + J], s[2 * n + J]};
    double ri[3], rj[3];
    fold3(si, S->box2, ri);
    fold3(sj, S->box2, rj);
    const double4 a = make_double4(ri[0], ri[1], ri[2], 0.0), b = make_double4(rj[0], rj[1], rj[2], 0.0);
    return exact_r2(a, b, *S) < S->g[g].rc2;
}

// Contribution of one listed neighbour (NONBONDED_FORCE.F90:55-98).  The
// displacement is the wrap-around difference of the int32 fixed-point
// positions; the cutoff test and the force direction use it in FP32, and
// pairs within FP32 rounding of rc are settled by the FP64 test.  For pairs
// inside the cutoff the table coordinate r/bin is formed in FP64 from the
// exact integer r^2 (one Newton step from the FP32 root), so alpha = (r -
// x_k)/bin is exact to FP32: an FP32 r would leave a random error of ~1e-7
// of the summed energy, dominated by pairs on the steep repulsive wall.
// alpha uses the table's own abscissas (NONBONDED_FORCE.F90:66-68).
template <bool kMulti, int kGroup, bool kCubic>
__device__ __forceinline__ void nb_pair(const NbCtx &C, const int4 pi, int j, const int4 pj, int li, float &fx,
                                        float &fy, float &fz, float (&e)[3], float (&w)[6]) {
    const int ix = (int)((uint32_t)pi.x - (uint32_t)pj.x);
    const int iy = (int)((uint32_t)pi.y - (uint32_t)pj.y);
    const int iz = (int)((uint32_t)pi.z - (uint32_t)pj.z);
    float dx = (float)ix, dy = (float)iy, dz = (float)iz;
    if (!kCubic) { dx *= C.rho[0]; dy *= C.rho[1]; dz *= C.rho[2]; }
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    if (r2 >= C.band_hi) return;
    if (r2 >= C.band_lo && !nb_inside_exact(C.S, C.g, C.s, C.n, C.I, C.perm[j])) return;
    NbTab tb = C.t0;
    if (kMulti) tb = C.tabs[C.map[pos_type(pi.w) * C.ntype + pos_type(pj.w)]];
    double ddx = (double)ix, ddy = (double)iy, ddz = (double)iz;
    if (!kCubic) { ddx *= C.rhod[0]; ddy *= C.rhod[1]; ddz *= C.rhod[2]; }
    const double r2d = fma(ddx, ddx, fma(ddy, ddy, ddz * ddz));
    // r2 >= 1 for any two distinct positions; a coincident pair gets r = 0 (zero force, as the
    // reference's F/r table times a zero displacement).  Integer max of the bits of a non-negative float.
    const float rs = __builtin_amdgcn_rsqf(__int_as_float(max(__float_as_int(r2), 0x3f800000)));
    const double r0 = (double)(r2 * rs);
    const double rd = fma(fma(-r0, r0, r2d), (double)(0.5f * rs), r0);
    const double tt = rd * tb.c;
    const int kr = min((int)tt, tb.ndat);
    // tt - floor(tt), exact (kr = floor(tt) for every pair inside rc, so within the table)
    float a = (float)__builtin_amdgcn_fract(tt);
    if (!tb.uniform) a -= C.xoff[C.xoff_base + tb.row_off + kr];
    const float4 row = C.rows[tb.row_off + kr];
    const float F = fmaf(a, row.z, row.x);
    const float V = fmaf(a, row.w, row.y);
    fx = fmaf(F, dx, fx);
    fy = fmaf(F, dy, fy);
    fz = fmaf(F, dz, fz);
    if (kGroup == 0) {
        e[0] += V;
    } else {
        const int cls = (li == 2 && pos_label(pj.w) == 2) ? 1 : 2;
        e[cls] += V;
    }
    const float fdx = F * dx, fdy = F * dy;
    w[0] = fmaf(fdx, dx, w[0]);
    w[1] = fmaf(fdy, dy, w[1]);
    w[2] = fmaf(F * dz, dz, w[2]);
    w[3] = fmaf(fdy, dx, w[3]);
    w[4] = fmaf(fdx, dz, w[4]);
    w[5] = fmaf(fdy, dz, w[5]);
}


// One run of a slot's list: entries q, q + kLanes, ... at l[0], l[kStep], ...
// (kStep = +kWave for the front run, -kWave for the back run), kNbUnroll
// entries a round, with the next round's indices loaded ahead of their
// positions.  The pairs are added in entry order whatever the unroll.
constexpr int kNbUnroll = 4;

template <bool kMulti, int kGroup, bool kCubic, int kStep>
__device__ __forceinline__ void nb_run(const NbCtx &C, const int4 *__restrict__ pos4, const int *__restrict__ l,
                                       int nm, const int4 pi, int li, float &fx, float &fy, float &fz,
                                       float (&e)[3], float (&w)[6]) {
    constexpr int U = kNbUnroll;
    int m = 0;
    int jj[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jj[u] = 0;
    if (nm >= U) {
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = l[u * kStep];
    }
    for (; m + U <= nm; m += U) {
        int4 pp[U];
#pragma unroll
        for (int u = 0; u < U; ++u) pp[u] = ld_pos(pos4, jj[u]);
        l += U * kStep;
        int kk[U];
#pragma unroll
        for (int u = 0; u < U; ++u) kk[u] = jj[u];
        if (m + 2 * U <= nm) {
#pragma unroll
            for (int u = 0; u < U; ++u) kk[u] = l[u * kStep];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) nb_pair<kMulti, kGroup, kCubic>(C, pi, jj[u], pp[u], li, fx, fy, fz, e, w);
#pragma unroll
        for (int u = 0; u < U; ++u) jj[u] = kk[u];
    }
    for (; m < nm; ++m, l += kStep) {
        const int j = l[0];
        nb_pair<kMulti, kGroup, kCubic>(C, pi, j, ld_pos(pos4, j), li, fx, fy, fz, e, w);
    }
}

// Nonbonded forces (replaces NONBONDED_FORCE.F90:11-109).
//
// kLanes lanes per sorted slot walk its list: the run of pairs that were
// closer than rc + delta at the rebuild, and the rest only once some member has
// moved delta/2 since (before that they cannot be inside rc: the result is the
// reference's full list scan).  Positions are int32 fixed
// point, so the minimum image is the wrap-around of the integer difference.
// Tables are staged in LDS when they fit (kLds), else read through the cache.
// Each pair is visited from both ends; energies and virial are halved.  The
// loops are unrolled by four so four index and position loads are in flight.
#define IBX_NB_PARAMS                                                                                           \
    const DevSys *__restrict__ sys, const int *__restrict__ wg_sys, const int *__restrict__ perm,                  \
        const int4 *__restrict__ pos4, const int *__restrict__ list, const int2 *__restrict__ nneb,               \
        const int *__restrict__ dmax, int nsys, int cap, const float4 *__restrict__ nb_rows,                      \
        const float *__restrict__ nb_xoff, const NbTab *__restrict__ nb_tabs, const int *__restrict__ nb_map,     \
        int nitems_total, const double *__restrict__ s, float *__restrict__ fnb, double *__restrict__ part, int nblk
#define IBX_NB_ARGS sys, wg_sys, perm, pos4, list, nneb, dmax, nsys, cap, nb_rows, nb_xoff, nb_tabs, nb_map, \
                    nitems_total, s, fnb, part, nblk

// The work of workgroup `bidx` of a nonbonded launch (also one role of k_step_forces).
template <bool kMulti, bool kLds, int kGroup, bool kCubic>
__device__ __forceinline__ void nb_block(const int bidx, const DevSys *__restrict__ sys, const int *__restrict__ wg_sys,
                                                      const int *__restrict__ perm, const int4 *__restrict__ pos4,
                                                      const int *__restrict__ list, const int2 *__restrict__ nneb,
                                                      const int *__restrict__ dmax, int nsys, int cap,
                                                      const float4 *__restrict__ nb_rows,
                                                      const float *__restrict__ nb_xoff,
                                                      const NbTab *__restrict__ nb_tabs, const int *__restrict__ nb_map,
                                                      int nitems_total, const double *__restrict__ s,
                                                      float *__restrict__ fnb, double *__restrict__ part, int nblk,
                                                      int &lds_sys, float *acc_lds) {
    extern __shared__ float4 lds_rows[];
    __shared__ NbTab tabs[kMultiTabsLds(kMulti)];
    __shared__ int map[kMulti ? kMaxNbTypes * kMaxNbTypes : 1];
    __shared__ double red[kNbFields * 16];
    const int blk = xcd_block(nblk, bidx);
    if (blk >= nblk) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int a = lane / kLanes, q = lane % kLanes;
    const int i = blk * kNbAtomsPerBlock + (threadIdx.x / kWave) * kAtomsPerWave + a;
    // the slot's own loads are in flight while the tables are staged
    const int I = perm[i];
    const int4 pi = pos4[i];
    const int2 nn = nneb[i];
    issue_here(perm, pos4, nneb);
    const int sid = wg_sys[blk / (kBlock / kNbAtomsPerBlock)];
    const DevSys &S = sys[sid];
    if (!alive_at(S, *S.step_ctr + 1)) return;
    const DevGroup &G = S.g[kGroup];
    // lds_sys: the system whose tables the LDS holds (a workgroup that works on several blocks in turn
    // stages them once per system)
    const bool restage = sid != lds_sys;
    lds_sys = sid;
    if (kLds && restage) {
        // four rows per thread in flight at once (the last row is read again past the end)
        constexpr int kU = 4;
        const int nrow = S.nb_nrow;
        for (int r0 = threadIdx.x; r0 < nrow; r0 += kU * kBlock) {
            float4 v[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) v[u] = nb_rows[S.nb_row_off + min(r0 + u * kBlock, nrow - 1)];
#pragma unroll
            for (int u = 0; u < kU; ++u) lds_rows[min(r0 + u * kBlock, nrow - 1)] = v[u];   // same value twice at most
        }
    }
    for (int t = threadIdx.x; restage && t < S.nb_ntab && t < kMultiTabsLds(kMulti); t += kBlock) {
        NbTab tb = nb_tabs[S.nb_tab_off + t];
        tb.row_off -= S.nb_row_off;
        tabs[t] = tb;
    }
    if (kMulti && restage)
        for (int t = threadIdx.x; t < S.ntype * S.ntype; t += kBlock) {
            const int m = nb_map[S.nbmap_off + t];
            map[t] = m < 0 ? 0 : m - S.nb_tab_off;
        }
    __syncthreads();

    float fx = 0.f, fy = 0.f, fz = 0.f;
    float e[3] = {0.f, 0.f, 0.f};
    float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (I >= 0) {
        NbCtx C;
        C.rows = kLds ? lds_rows : nb_rows + S.nb_row_off;
        C.xoff = nb_xoff;
        C.xoff_base = S.nb_row_off;
        C.tabs = tabs;
        C.map = map;
        C.ntype = S.ntype;
        C.t0 = tabs[0];
        for (int k = 0; k < 3; ++k) {
            C.rho[k] = S.rho[k];
            C.rhod[k] = S.box[k] / S.box[0];
        }
        C.band_lo = G.rc2u_lo;
        C.band_hi = G.rc2u_hi;
        C.S = &S;
        C.s = s;
        C.perm = perm;
        C.n = nitems_total;
        C.I = I;
        C.g = kGroup;
        const int li = pos_label(pi.w);
        const int n_in = min(nn.x, cap);
        // the back run (rc + delta <= r < rlist at the rebuild) can only reach rc once some member of
        // the system has moved delta/2 since then
        const bool far_ok = __int_as_float(dmax[kGroup * nsys + sid]) < G.dmax2u;
        const int n_out = far_ok ? 0 : min(nn.y, cap - n_in);
        // front run: entry q + kLanes*m at lp[m*kWave]; back run: entry q + kLanes*m at lb[-m*kWave]
        const int *lp = list + list_pos(i, q, cap);
        const int *lb = list + list_pos(i, cap - 1 - q, cap);
        const int nm_in = n_in > q ? (n_in - q + kLanes - 1) / kLanes : 0;
        const int nm_out = n_out > q ? (n_out - q + kLanes - 1) / kLanes : 0;
        nb_run<kMulti, kGroup, kCubic, kWave>(C, pos4, lp, nm_in, pi, li, fx, fy, fz, e, w);
        if (nm_out > 0) nb_run<kMulti, kGroup, kCubic, -kWave>(C, pos4, lb, nm_out, pi, li, fx, fy, fz, e, w);
    }
    // the kLanes lanes of a slot hold partial forces: butterfly-sum them
    fx = group_sum<kLanes>(fx);
    fy = group_sum<kLanes>(fy);
    fz = group_sum<kLanes>(fz);
    if (I >= 0 && q == 0) {
        const float u = (float)S.unit;
        fnb[I] = fx * u;
        fnb[nitems_total + I] = fy * u;
        fnb[2 * nitems_total + I] = fz * u;
    }
    // Block sums of the nine energy and virial accumulators through LDS (this
    // kernel is VALU-bound, so no cross-lane arithmetic): the lane values go to
    // acc_lds (null: the table's staging area, no longer needed); sixteen
    // threads per field add every sixteenth value, then one thread per field
    // adds those sixteen sums, in FP64 and in a fixed order.
    const float acc[kNbFields] = {e[0], e[1], e[2], w[0], w[1], w[2], w[3], w[4], w[5]};
    if (!acc_lds) {
        acc_lds = reinterpret_cast<float *>(lds_rows);
        lds_sys = -1;   // the tables are overwritten
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kNbFields; ++k) acc_lds[k * kBlock + threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x < kNbFields * 16) {
        const float *pf = acc_lds + (threadIdx.x >> 4) * kBlock + (threadIdx.x & 15);
        double x = 0.0;
#pragma unroll
        for (int u = 0; u < kBlock / 16; ++u) x += (double)pf[16 * u];
        red[threadIdx.x] = x;
    }
    __syncthreads();
    if (threadIdx.x < kNbFields) {
        double x = 0.0;
        for (int u = 0; u < 16; ++u) x += red[threadIdx.x * 16 + u];
        // field-major partials: field k of block w at part[k * nblk + w]
        const double u2 = S.unit * S.unit;
        part[(size_t)threadIdx.x * nblk + blk] = threadIdx.x < 3 ? 0.5 * x : 0.5 * x * u2;
    }
}

template <bool kMulti, bool kLds, int kGroup, bool kCubic>
__global__ void __launch_bounds__(kBlock) k_nonbonded(IBX_NB_PARAMS) {
    int lds_sys = -1;
    nb_block<kMulti, kLds, kGroup, kCubic>((int)blockIdx.x, IBX_NB_ARGS, lds_sys, nullptr);
}

// ------------------------------------------------------------------------------------------------
// bonded forces (replaces BONDED_FORCE.F90 with BONDS/ANGLES/TORSIONS/OUTOFPLANES.inc)
//
// One thread per term, in FP64 from the unwrapped coordinates; each member's
// force goes to a contribution slot of its own, which the integrator gathers
// (no atomics, deterministic sums).

struct TermArgs {
    const DevSys *sys;    // per-system constants (retired systems are skipped)
    const int4 *mem;      // members (unused -1)
    const int *meta;      // kind | table << 2 | virtual-site members << kTermVsShift, -1 for padding
    const int4 *slot;     // contribution slot of each member
    const BdTab *tabs;    // all bonded table headers
    const float4 *rows;   // all bonded rows {F_k, V_k, F_k+1, V_k+1}
    const double *xs;     // abscissa of every bonded row
    const double *vs_com; // per virtual site: the mass-weighted centre of its atoms (SoA like s)
};

// Table row index from a non-negative abscissa ratio, clamped to [0, ndat]
// (NaN -> 0); ratios at or past ndat+1 set `bit` in *flag.
template <typename T>
__device__ __forceinline__ int row_index(T x, int ndat, int *flag, int bit) {
    if (!(x >= (T)0)) return 0;
    if (x >= (T)(ndat + 1)) {
        if (flag) atomicOr(flag, bit);
        return ndat;
    }
    return min((int)x, ndat);
}

// Interpolation in one table row {F_k, V_k, F_k+1, V_k+1}.  The callers load
// the row and its abscissa together, before either is used.
template <typename T>
__device__ __forceinline__ void lerp_row(const float4 r, T alpha, T &F, T &V) {
    F = (T)r.x * ((T)1 - alpha) + alpha * (T)r.z;
    V = (T)r.y * ((T)1 - alpha) + alpha * (T)r.w;
}

// Difference of two members' unwrapped positions (P holds the term's members).
template <typename T>
__device__ __forceinline__ void dvec(const double (&P)[4][3], int a, int b, T d[3]) {
    d[0] = (T)(P[a][0] - P[b][0]);
    d[1] = (T)(P[a][1] - P[b][1]);
    d[2] = (T)(P[a][2] - P[b][2]);
}

template <typename T>
__device__ __forceinline__ T nrm(const T a[3]) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

template <typename T>
__device__ __forceinline__ void cross(const T a[3], const T b[3], T c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = -a[0] * b[2] + a[2] * b[0];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// BONDS.inc:8-56.  Returns the force on member `role` and the energy.
template <typename T>
__device__ __forceinline__ void bond_term(const BdTab &t, const float4 *rows, const double *xs,
                                          const double (&P)[4][3], T fm[4][3], T &V, bool &over) {
    T d[3];
    dvec(P, 0, 1, d);
    const T rij = nrm(d);
    const T xr = (rij - (T)t.x0) * (T)t.inv_bin;
    over = xr >= (T)(t.ndat + 1);              // BONDS.inc:21-31: warn, clamp, continue
    const int ni = row_index(xr, t.ndat, (int *)nullptr, 0);
    const float4 r = rows[t.row_off + ni];
    const T alpha = (rij - (T)xs[t.row_off + ni]) * (T)t.inv_bin;   // RBOND(TIJ,NI), the table's own abscissa
    T F;
    lerp_row(r, alpha, F, V);
    for (int k = 0; k < 3; ++k) {
        fm[0][k] = F * d[k];
        fm[1][k] = -F * d[k];
    }
}

// ANGLES.inc:2-75
template <typename T>
__device__ __forceinline__ void angle_term(const BdTab &t, const float4 *rows, const double *xs,
                                           const double (&P)[4][3], T fm[4][3], T &V) {
    const T R2D = (T)180 / (T)3.14159265358;
    T ij[3], kj[3], M[3], N[3], Sv[3];
    dvec(P, 0, 1, ij);
    dvec(P, 2, 1, kj);
    const T rij = nrm(ij), rkj = nrm(kj);
    T ct = (ij[0] * kj[0] + ij[1] * kj[1] + ij[2] * kj[2]) / rij / rkj;
    ct = ct > (T)1 ? (T)1 : (ct < (T)-1 ? (T)-1 : ct);
    const T theta = acos(ct) * R2D;
    M[0] = -(ij[1] * kj[2] - ij[2] * kj[1]);
    M[1] = -(-ij[0] * kj[2] + ij[2] * kj[0]);
    M[2] = -(ij[0] * kj[1] - ij[1] * kj[0]);
    const T rm = nrm(M);
    for (int k = 0; k < 3; ++k) M[k] /= rm;
    N[0] = (M[1] * ij[2] - M[2] * ij[1]) / rij;
    N[1] = (-M[0] * ij[2] + M[2] * ij[0]) / rij;
    N[2] = (M[0] * ij[1] - M[1] * ij[0]) / rij;
    Sv[0] = -(M[1] * kj[2] - M[2] * kj[1]) / rkj;
    Sv[1] = -(-M[0] * kj[2] + M[2] * kj[0]) / rkj;
    Sv[2] = -(M[0] * kj[1] - M[1] * kj[0]) / rkj;
    const T xr = theta * (T)t.inv_bin;
    const int ni = row_index(xr, t.ndat, (int *)nullptr, 0);
    const float4 r = rows[t.row_off + ni];
    const T x = (T)xs[t.row_off + ni];
    T F;
    if (theta == (T)180) {
        F = (T)r.x;
        V = (T)r.y;
    } else {
        lerp_row(r, (theta - x) * (T)t.inv_bin, F, V);
    }
    for (int k = 0; k < 3; ++k) {
        const T fa = F * N[k] / rij, fk = F * Sv[k] / rkj;
        fm[0][k] = fa;
        fm[1][k] = -(fa + fk);
        fm[2][k] = fk;
    }
}

// TORSIONS.inc:1-152
template <typename T>
__device__ __forceinline__ void torsion_term(const BdTab &t, const float4 *rows, const double *xs,
                                             const double (&P)[4][3], T fm[4][3], T &V, int *flags) {
    const T R2D = (T)180 / (T)3.14159265358;
    T ij[3], jk[3], kl[3], M[3], N[3];
    dvec(P, 0, 1, ij);
    dvec(P, 1, 2, jk);
    dvec(P, 2, 3, kl);
    T rij = nrm(ij);
    const T rjk = nrm(jk);
    T rkl = nrm(kl);
    cross(ij, jk, M);
    T rm = nrm(M);
    if (rm == (T)0) {           // TORSIONS.inc:47-59: perturb x of atom i
        ij[0] = (T)(P[0][0] * 0.99 - P[1][0]);
        rij = nrm(ij);
        cross(ij, jk, M);
        rm = nrm(M);
    }
    for (int k = 0; k < 3; ++k) M[k] /= rm;
    const T cosm = (ij[0] * jk[0] + ij[1] * jk[1] + ij[2] * jk[2]) / rij / rjk;
    const T sinm = sqrt(fmax((T)0, (T)1 - cosm * cosm));
    const T cotm = cosm / sinm;
    cross(jk, kl, N);
    T rn = nrm(N);
    if (rn == (T)0) {           // TORSIONS.inc:75-87: perturb x of atom l
        kl[0] = (T)(P[2][0] - P[3][0] * 0.99);
        rkl = nrm(kl);
        cross(jk, kl, N);
        rn = nrm(N);
    }
    for (int k = 0; k < 3; ++k) N[k] /= rn;
    const T cosn = (jk[0] * kl[0] + jk[1] * kl[1] + jk[2] * kl[2]) / rjk / rkl;
    const T sinn = sqrt(fmax((T)0, (T)1 - cosn * cosn));
    const T cotn = cosn / sinn;
    T cost = M[0] * N[0] + M[1] * N[1] + M[2] * N[2];
    const T signt = -(M[0] * kl[0] + M[1] * kl[1] + M[2] * kl[2]);
    cost = cost > (T)1 ? (T)1 : (cost < (T)-1 ? (T)-1 : cost);
    T phi = acos(cost) * R2D;
    if (signt < (T)0) phi = (T)360 - phi;
    const int ni = row_index(phi * (T)t.inv_bin, t.ndat, flags, 1);
    const float4 r = rows[t.row_off + ni];
    const T x = (T)xs[t.row_off + ni];
    T F;
    if (phi == (T)360) {
        F = (T)r.x;
        V = (T)r.y;
    } else {
        lerp_row(r, (phi - x) * (T)t.inv_bin, F, V);
    }
    for (int k = 0; k < 3; ++k) {
        const T f1 = -F * M[k] / sinm / rij;
        const T f4 = F * N[k] / sinn / rkl;
        const T f12 = F * (cotm * M[k] + cotn * N[k]) / rjk;
        fm[0][k] = f1;
        fm[1][k] = -f1 + f12;
        fm[2][k] = -f4 - f12;
        fm[3][k] = f4;
    }
}

// OUTOFPLANES.inc:1-141
template <typename T>
__device__ __forceinline__ void oop_term(const BdTab &t, const float4 *rows, const double *xs,
                                         const double (&P)[4][3], T fm[4][3], T &V, int *flags) {
    const T R2D = (T)180 / (T)3.14159265358;
    T ij[3], jk[3], kl[3], M[3], N[3], mN[3];
    dvec(P, 0, 1, ij);
    dvec(P, 2, 1, jk);
    dvec(P, 2, 3, kl);
    const T rij = nrm(ij), rjk = nrm(jk), rkl = nrm(kl);
    cross(ij, jk, M);
    const T rm = nrm(M);
    for (int k = 0; k < 3; ++k) M[k] /= rm;
    const T cosm = (ij[0] * jk[0] + ij[1] * jk[1] + ij[2] * jk[2]) / rij / rjk;
    const T sinm = sqrt(fmax((T)0, (T)1 - cosm * cosm)), cotm = cosm / sinm;
    cross(jk, kl, N);
    const T rn = nrm(N);
    for (int k = 0; k < 3; ++k) N[k] /= rn;
    const T cosn = (jk[0] * kl[0] + jk[1] * kl[1] + jk[2] * kl[2]) / rjk / rkl;
    const T sinn = sqrt(fmax((T)0, (T)1 - cosn * cosn)), cotn = cosn / sinn;
    T cost = M[0] * N[0] + M[1] * N[1] + M[2] * N[2];
    cross(M, N, mN);
    const T signt = copysign((T)1, mN[0] * jk[0] + mN[1] * jk[1] + mN[2] * jk[2]);
    cost = cost > (T)1 ? (T)1 : (cost < (T)-1 ? (T)-1 : cost);
    const T phi = signt * acos(cost) * R2D;
    int ni = row_index(rint(((T)180 + phi) * (T)t.inv_bin), t.ndat, flags, 2);
    float4 r = rows[t.row_off + ni];
    T alpha = fabs(fabs(phi) - fabs((T)xs[t.row_off + ni])) * (T)t.inv_bin;
    if (alpha > (T)1) {
        ni = min(ni + 1, t.ndat);
        r = rows[t.row_off + ni];
        alpha = fabs(fabs(phi) - fabs((T)xs[t.row_off + ni])) * (T)t.inv_bin;
    }
    T F;
    lerp_row(r, alpha, F, V);
    for (int k = 0; k < 3; ++k) {
        const T f1 = F * M[k] / sinm / rij;
        const T f4 = -F * N[k] / sinn / rkl;
        const T f12 = F * (cotm * M[k] + cotn * N[k]) / rjk;
        fm[0][k] = f1;
        fm[1][k] = -f1 + f12;
        fm[2][k] = -f4 - f12;
        fm[3][k] = f4;
    }
}

// Bonded forces (replaces BONDED_FORCE.F90 with BONDS/ANGLES/TORSIONS/OUTOFPLANES.inc).
//
// One thread per term evaluates it in FP64 and stores the force on each member
// in its own contribution slot; the integration kernel gathers an item's slots
// in a fixed order, so there are no atomics and the sums are deterministic.
// Geometry differences are taken in FP64 from the unwrapped coordinates.
//
// Virial: NEW_FORCE.F90:73-80 sums (F - F_nb) . s over atoms with unwrapped
// coordinates.  Because the member forces of a term sum to zero, the same sum
// is accumulated per term as sum_m f_m (s_m - s_first) in FP64, which does not
// lose precision to the size of the unwrapped coordinates.  Component order as
// the reference: 11 22 33, then F_y s_x, F_z s_x, F_z s_y.
// Workgroup b of nwg (also one role of k_step_forces).
__device__ __forceinline__ void terms_block(const int b, const int nwg, const int *__restrict__ wg_sys, const TermArgs &A,
                                            const double *__restrict__ s, int nitems_total, int nterms,
                                            Contrib *__restrict__ contrib, int ncontrib, double *__restrict__ part,
                                            int *__restrict__ flags) {
    __shared__ double red[19 * 4];
    const int wsys = wg_sys[b];
    if (wsys & kWgRetired) return;   // a retired system: nothing loaded
    const DevSys &Sx = A.sys[wsys];
    const int live = alive_at(Sx, *Sx.step_ctr + 1);   // first, through the scalar cache
    const int t = b * kBlock + threadIdx.x;
    const int n = nitems_total;
    // the term and its members' positions are loaded before the system record
    // (the term arrays are padded to whole blocks), so those round trips overlap
    const int meta = t < nterms ? A.meta[t] : -1;
    const int4 m4 = A.mem[t];
    const int4 s4 = A.slot[t];
    const int mem[4] = {m4.x, m4.y, m4.z, m4.w};
    double P[4][3];   // unused members (and padding terms) read item 0, never used
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int mr = term_member(meta, mem[r]);
        for (int k = 0; k < 3; ++k) P[r][k] = s[k * n + mr];
    }
    const BdTab tb = A.tabs[term_table(meta)];
    issue_here(s, A.tabs, A.mem);
    if (!live) return;
    double en[19];   // 0..11 unused here: a term's energy is vt in class ek (see the sums below)
    for (int k = 0; k < 19; ++k) en[k] = 0.0;
    double vt = 0.0;
    int ek = -1;
    if (meta >= 0) {
        const int kind = meta & 3;
        double fm[4][3], V;
        int nm;
        if (kind == KBOND) {
            bool over;
            bond_term<double>(tb, A.rows, A.xs, P, fm, V, over);
            if (over) en[12] += 1.0;
            nm = 2;
        } else if (kind == KANGLE) {
            angle_term<double>(tb, A.rows, A.xs, P, fm, V);
            nm = 3;
        } else if (kind == KTORS) {
            torsion_term<double>(tb, A.rows, A.xs, P, fm, V, flags + 1);   // word 1: table bits
            nm = 4;
        } else {
            oop_term<double>(tb, A.rows, A.xs, P, fm, V, flags + 1);
            nm = 4;
        }
        ek = 3 * kind + tb.label - 1;
        vt = V;
        const int sl[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (r >= nm) break;
            contrib[sl[r]] = Contrib{(float)fm[r][0], (float)fm[r][1], (float)fm[r][2]};
        }
        // NEW_FORCE.F90:73-80 sums the bonded virial over atoms after DISTRIBUTE_VSFORCE.f90:27-37 has moved a
        // virtual site's force to its atoms (m_j/M each): a site's force counts at the centre of its atoms
        const int vsmask = meta >> kTermVsShift;
        if (vsmask)
            for (int r = 0; r < nm; ++r)
                if (vsmask & (1 << r))
                    for (int k = 0; k < 3; ++k) P[r][k] = A.vs_com[k * n + mem[r]];
#pragma unroll
        for (int r = 1; r < 4; ++r) {
            if (r >= nm) break;
            const double dx = P[r][0] - P[0][0], dy = P[r][1] - P[0][1], dz = P[r][2] - P[0][2];
            en[13] += fm[r][0] * dx;
            en[14] += fm[r][1] * dy;
            en[15] += fm[r][2] * dz;
            en[16] += fm[r][1] * dx;
            en[17] += fm[r][2] * dx;
            en[18] += fm[r][2] * dy;
        }
    }
    // Workgroup sums of the twelve energy classes (kind x label), the overextension
    // count and the virial.  The terms of a wave usually share one class (they are
    // grouped by kind and table), and such a wave sums its energies once instead
    // of once per class; the sums are the same either way.
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const unsigned long long valid = __ballot(ek >= 0);
    const int ek_ref = valid ? __builtin_amdgcn_readlane(ek, __ffsll((long long)valid) - 1) : 0;
    double w[19];
    if (__ballot(ek >= 0 && ek != ek_ref) == 0) {
        const double v = wave_sum(vt);
#pragma unroll
        for (int k = 0; k < 12; ++k) w[k] = k == ek_ref ? v : 0.0;
    } else {
#pragma unroll
        for (int k = 0; k < 12; ++k) w[k] = wave_sum(ek == k ? vt : 0.0);
    }
#pragma unroll
    for (int k = 12; k < 19; ++k) w[k] = wave_sum(en[k]);
    if (lane == 0)
        for (int k = 0; k < 19; ++k) red[k * 4 + wid] = w[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double *p = part + b;   // field-major
        for (int k = 0; k < 19; ++k)
            p[(size_t)k * nwg] = ((red[k * 4] + red[k * 4 + 1]) + red[k * 4 + 2]) + red[k * 4 + 3];
    }
}

__global__ void __launch_bounds__(kBlock) k_terms(const int *__restrict__ wg_sys, TermArgs A, const double *__restrict__ s,
                                                  int nitems_total, int nterms, Contrib *__restrict__ contrib,
                                                  int ncontrib, double *__restrict__ part, int *__restrict__ flags) {
    terms_block((int)blockIdx.x, (int)gridDim.x, wg_sys, A, s, nitems_total, nterms, contrib, ncontrib, part, flags);
}

// ------------------------------------------------------------------------------------------------
// integration (replaces MOVE.F90:9-118; MOMENTUM.f90 is applied here too)
//
// mode 0: leap-frog step; mode 1: forces only (combine, bonded virial, total force).
// The total force on an atom is its nonbonded plus bonded force plus the mass
// share of the forces on its virtual site (DISTRIBUTE_VSFORCE.f90:21-38); the
// bonded virial is sum over atoms of (F - F_nb) . s (NEW_FORCE.F90:73-80).

struct MoveArgs {
    double *s, *v;
    const float *fnb;
    const Contrib *contrib;  // bonded contribution slots; item i owns [inc_off[i], inc_off[i+1])
    const int *inc_off;
    int ncontrib;
    float *ftot;
    const int *itype, *label;
    const double *mass, *invmass;
    const int *vs_of;
    const double *vs_coeff;
    int4 *pos4[2];
    const int4 *pos4_ref[2];
    const int *inv_perm[2];
    int *dmax;                // [group][system] largest displacement^2 since the rebuild (float bits)
    int nsys;
    const double *sysstate;   // per system: temp_atom, temp_bead, pending momentum x,y,z
    double *part;
    int *step;
    int nitems_total;
    int mode;
    float nb_weight;          // multiple time step: k on the steps that evaluate the nonbonded force, else 0
};

// Adds the contribution slots [a, b) of one item in slot order, loading
// twelve slots at a time so their latencies overlap.  The loads are
// unconditional (slots past b read slot 0 and are dropped): a load under a
// per-lane condition would make the compiler wait for it before the next one.
__device__ __forceinline__ void gather_contrib(const Contrib *__restrict__ contrib, int nc, int a, int b, double f[3]) {
    (void)nc;
    for (int q = a; q < b; q += 12) {
        Contrib c[12];
#pragma unroll
        for (int u = 0; u < 12; ++u) c[u] = contrib[contrib_entry(q + u, b)];
#pragma unroll
        for (int u = 0; u < 12; ++u) {
            if (q + u < b) {
                f[0] += (double)c[u].x;
                f[1] += (double)c[u].y;
                f[2] += (double)c[u].z;
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_move(const DevSys *__restrict__ sys, const int *__restrict__ wg_sys, MoveArgs A) {
    // partials: kinetic tensor (xx yy zz xy xz yz) of the atom population, of the bead population, momentum
    constexpr int kNf = 15;
    __shared__ double red[kNf * 4];
    __shared__ float mx[kBlock / kWave];
    float disp[2] = {0.f, 0.f};   // squared displacement since the rebuild, per group
    const int it = blockIdx.x * kBlock + threadIdx.x;
    const int n = A.nitems_total;
    // The system's constants and state (temperatures of the two populations,
    // pending momentum) are read first, before any store (so through the scalar
    // cache, and not again after each store that might alias them).
    if (A.mode == 0 && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(A.step, 1);   // the step is done
    const int w = wg_sys[blockIdx.x];
    if (w & kWgRetired) return;   // a retired system: nothing loaded
    const int s_id = w;
    const DevSys &S = sys[s_id];
    const int ensemble = S.ensemble, hybrid = S.hybrid;
    const double fac = S.fac, temp_in = S.temp_in, dt = S.dt;
    const double ibf[3] = {S.inv_box_frac[0], S.inv_box_frac[1], S.inv_box_frac[2]};
    const float rho1 = S.rho[1], rho2 = S.rho[2];
    const double *stp = A.sysstate + (size_t)s_id * 8;
    const double st[5] = {stp[0], stp[1], stp[2], stp[3], stp[4]};
    // whether the system advances this step: written by the field sums of the step before (race-free, unlike
    // the step counter, which this launch advances)
    const int live = S.live && stp[5] != 0.0;
    // Every per-item load is issued before the system record is checked (the
    // arrays are padded to whole blocks): the item's own data, then what that
    // data points at (its contribution slots, its virtual site's, the reference
    // positions of its list slots), so the kernel waits for two round trips
    // rather than one per dependent load.
    const int ty = A.itype[it];
    const int lab = A.label[it];
    const int q0 = A.inc_off[it], q1 = A.inc_off[it + 1];
    const int vs = A.vs_of[it];
    const float fn[3] = {A.fnb[it], A.fnb[n + it], A.fnb[2 * n + it]};
    double sp[3] = {A.s[it], A.s[n + it], A.s[2 * n + it]};
    const double vin[3] = {A.v[it], A.v[n + it], A.v[2 * n + it]};
    const double m = A.mass[it], im = A.invmass[it];
    const int slot_g[2] = {A.inv_perm[0][it], A.inv_perm[1][it]};
    const bool own = ty >= 0 && lab != 3;   // an atom or bead (virtual sites are distributed, not moved)
    double fbv[3] = {0, 0, 0}, fvs[3] = {0, 0, 0}, fvs_nb[3] = {0, 0, 0};
    gather_contrib(A.contrib, A.ncontrib, q0, own ? q1 : q0, fbv);
    double vs_c = 0.0;
    if (own && vs >= 0) {
        // the virtual site's nonbonded and bonded forces, spread by mass fraction
        vs_c = A.vs_coeff[it];
        for (int k = 0; k < 3; ++k) fvs_nb[k] = (double)A.fnb[k * n + vs];
        gather_contrib(A.contrib, A.ncontrib, A.inc_off[vs], A.inc_off[vs + 1], fvs);
    }
    int4 ref[2];
    for (int g = 0; g < 2; ++g) ref[g] = A.pos4_ref[g][member_slot(slot_g[g])];   // used only where slot_g[g] >= 0
    issue_here();
    if (!live) return;
    double acc[kNf];
    for (int k = 0; k < kNf; ++k) acc[k] = 0.0;
    if (A.mode == 1 && ty >= 0 && lab == 3)
        for (int k = 0; k < 3; ++k) A.ftot[k * n + it] = 0.0f;   // distributed (DISTRIBUTE_VSFORCE.f90:40)
    if (own) {
        double F[3];
        const double wnb = (double)A.nb_weight;
        for (int k = 0; k < 3; ++k) F[k] = (double)fn[k] * wnb + fbv[k];
        if (vs >= 0)
            for (int k = 0; k < 3; ++k) F[k] += (fvs_nb[k] * wnb + fvs[k]) * vs_c;
        if (A.mode == 1) {
            for (int k = 0; k < 3; ++k) A.ftot[k * n + it] = (float)F[k];
        } else {
            const int pop = (lab == 2) ? 1 : 0;
            double lam = 1.0;
            if (ensemble == 1 || ensemble == 2)   // NVT or NPT (MOVE.F90:32)
                lam = sqrt(1.0 + fac * ((temp_in / (pop ? st[1] : st[0])) - 1.0));
            double vt[3];
            for (int k = 0; k < 3; ++k) {
                const double v0 = vin[k] - st[2 + k] * im;   // deferred MOMENTUM
                const double v1 = (v0 + dt * F[k] * im) * lam;
                A.v[k * n + it] = v1;
                sp[k] += dt * v1;
                A.s[k * n + it] = sp[k];
                vt[k] = 0.5 * (v1 + v0);
                acc[12 + k] = m * v1;
            }
            double *kk = acc + 6 * pop;
            kk[0] = m * vt[0] * vt[0];
            kk[1] = m * vt[1] * vt[1];
            kk[2] = m * vt[2] * vt[2];
            kk[3] = m * vt[0] * vt[1];
            kk[4] = m * vt[0] * vt[2];
            kk[5] = m * vt[1] * vt[2];
            // fixed-point copy for the next nonbonded evaluation
            const int g = (hybrid && lab == 2) ? 1 : 0;
            const int slot = slot_g[g];
            if (slot >= 0) {
                int4 p;
                p.x = (int)to_fixed(sp[0], ibf[0]);
                p.y = (int)to_fixed(sp[1], ibf[1]);
                p.z = (int)to_fixed(sp[2], ibf[2]);
                p.w = ty | (lab << 16);
                A.pos4[g][slot] = p;
                disp[g] = fixed_d2(p, ref[g], rho1, rho2);
            }
        }
    }
    if (A.mode == 0) {
        block_max_to(disp[0], A.dmax + s_id, mx);
        if (hybrid) block_max_to(disp[1], A.dmax + A.nsys + s_id, mx);
    }
    block_sum<kNf>(acc, red);
    if (threadIdx.x == 0) {
        double *p = A.part + blockIdx.x;   // field-major
        for (int k = 0; k < kNf; ++k) p[(size_t)k * gridDim.x] = acc[k];
    }
}

// VIRTUAL_DEF.f90:8-50 after the atoms moved (mass and vs_invmass are per item): centre-atom copy or mass-weighted
// centre (all three components accumulated), then the sorted fixed-point copy.
__global__ void __launch_bounds__(kBlock) k_vs_place(const DevSys *__restrict__ sys, const int *__restrict__ wg_sys,
                                                     double *__restrict__ s, int nitems_total, const int *__restrict__ label,
                                                     const int *__restrict__ itype, const int *__restrict__ vs_center,
                                                     const int *__restrict__ vs_moff, const int *__restrict__ vs_members,
                                                     const double *__restrict__ mass, const double *__restrict__ vs_invmass,
                                                     double *__restrict__ vs_com,
                                                     int4 *__restrict__ pos4, const int4 *__restrict__ pos4_ref,
                                                     const int *__restrict__ inv_perm, int *__restrict__ dmax) {
    __shared__ float mx[kBlock / kWave];
    const int w = wg_sys[blockIdx.x];
    if (w & kWgRetired) return;
    const int s_id = w;
    const DevSys &S = sys[s_id];
    if (!alive_at(S, *S.step_ctr)) return;   // after the move: the counter holds this step
    const int it = blockIdx.x * kBlock + threadIdx.x;
    const int n = nitems_total;
    float disp = 0.f;
    if (itype[it] >= 0 && label[it] == 3) {
        // the mass-weighted centre of the site's atoms: the site itself when it is defined as one, and where a
        // bonded force on it counts in the virial
        double p[3], sum[3] = {0, 0, 0};
        for (int q = vs_moff[it]; q < vs_moff[it + 1]; ++q) {
            const int a = vs_members[q];
            const double m = mass[a];
            for (int k = 0; k < 3; ++k) sum[k] += m * s[k * n + a];
        }
        for (int k = 0; k < 3; ++k) {
            sum[k] *= vs_invmass[it];
            vs_com[k * n + it] = sum[k];
        }
        const int c = vs_center[it];
        for (int k = 0; k < 3; ++k) p[k] = c >= 0 ? s[k * n + c] : sum[k];
        for (int k = 0; k < 3; ++k) s[k * n + it] = p[k];
        const int slot = inv_perm[it];
        if (slot >= 0) {
            const int4 q4 = make_int4((int)to_fixed(p[0], S.inv_box_frac[0]), (int)to_fixed(p[1], S.inv_box_frac[1]),
                                      (int)to_fixed(p[2], S.inv_box_frac[2]), itype[it] | (3 << 16));
            pos4[slot] = q4;
            disp = fixed_d2(q4, pos4_ref[slot], S);
        }
    }
    block_max_to(disp, dmax + s_id, mx);
}

// SCALEBP.f90:36-43 on the device, after the field sums of a step: where the system rescales its box this step
// (the field sums set kNptGo and the factor MIU), every item's position is multiplied by MIU (the reference
// scales the atoms; its virtual sites follow at the next VIRTUAL_DEF, which here ran already, so they are
// scaled with them, as are the centres of their atoms).  The fixed-point copies keep their values: a position's
// fraction of the box does not change.  One thread per item.
__global__ void __launch_bounds__(kBlock) k_npt_positions(const DevSys *__restrict__ sys, const int *__restrict__ wg_sys,
                                                          const double *__restrict__ npt, double *__restrict__ s,
                                                          double *__restrict__ vs_com, const int *__restrict__ itype,
                                                          const int *__restrict__ label, int nitems_total) {
    const int w = wg_sys[blockIdx.x];
    if (w & kWgRetired) return;
    const DevSys &S = sys[w];
    if (S.ensemble != 2 || !alive_at(S, *S.step_ctr)) return;   // after the move: the counter holds this step
    const double *q = npt + (size_t)w * kNptState;
    if (q[kNptGo] == 0.0) return;
    const double miu = q[kNptMiu];
    const int it = blockIdx.x * kBlock + threadIdx.x, n = nitems_total;
    if (itype[it] < 0) return;
    for (int k = 0; k < 3; ++k) s[k * n + it] *= miu;
    if (label[it] == 3)
        for (int k = 0; k < 3; ++k) vs_com[k * n + it] *= miu;
}

// The system's constants after its box is rescaled by MIU (SCALEBP.f90:38-40 and the cell grid of
// UPDATE_NEIGHBOURLIST.F90:14-48): box, its inverse and half, the fixed-point unit and every radius
// expressed in it, the nonbonded tables' abscissa scale, and the cell grid, which keeps cells of at least
// rlist / cell_div(k) (it shrinks with the box and grows back to the grid the cell range holds).  Pair distances
// change by MIU while the fixed-point displacements do not, so the back run of each list is scanned until the
// next rebuild.  One thread per system, after k_npt_positions (which reads kNptGo; cleared here).
__global__ void __launch_bounds__(kBlock) k_npt_box(DevSys *__restrict__ sys, int nsys, double *__restrict__ npt,
                                                    NbTab *__restrict__ nb_tabs, int *__restrict__ dmax) {
    const int s_id = blockIdx.x * kBlock + threadIdx.x;
    if (s_id >= nsys) return;
    DevSys &S = sys[s_id];
    if (S.ensemble != 2 || !alive_at(S, *S.step_ctr)) return;
    double *q = npt + (size_t)s_id * kNptState;
    if (q[kNptGo] == 0.0) return;
    q[kNptGo] = 0.0;
    const double miu = q[kNptMiu];
    for (int k = 0; k < 3; ++k) {
        S.box[k] *= miu;
        S.boxinv[k] = 1.0 / S.box[k];
        S.box2[k] = S.box[k] / 2.0;
        S.inv_box_frac[k] = 1.0 / S.box[k];
    }
    S.volume = S.box[0] * S.box[1] * S.box[2];
    const double u_old = S.unit;
    S.unit = S.box[0] / 4294967296.0;
    const double r = u_old / S.unit;   // a length in the new units is r times its value in the old ones
    const float r2 = (float)(r * r);
    for (int g = 0; g < S.ngroup; ++g) {
        DevGroup &G = S.g[g];
        const double rl = sqrt(G.rl2), rc = sqrt(G.rc2);
        for (int k = 0; k < 3; ++k) {
            int nc = (int)(cell_div(k) * S.box[k] / rl);
            nc = nc < 1 ? 1 : (nc > G.ncmax[k] ? G.ncmax[k] : nc);
            G.nc[k] = nc;
        }
        G.ncell = G.nc[0] * G.nc[1] * G.nc[2];
        G.all_pairs = (G.nc[0] < min_cells(0) || G.nc[1] < min_cells(1) || G.nc[2] < min_cells(2)) ? 1 : 0;
        const double u = S.unit;
        const double rl2u = (rl / u) * (rl / u);
        const double lband = 4e-6 + 16.0 * u / rl;
        G.rl2u_lo = (float)(rl2u * (1.0 - lband));
        G.rl2u_hi = (float)(rl2u * (1.0 + lband));
        const double band = 1e-6 + 16.0 * u / rc;
        G.rc2u = (float)((rc / u) * (rc / u));
        G.rc2u_lo = (float)((rc / u) * (rc / u) * (1.0 - band));
        G.rc2u_hi = (float)((rc / u) * (rc / u) * (1.0 + band));
        G.rin2u *= r2;     // the same distances rc + delta and 0.49 delta, in the new units
        G.dmax2u *= r2;
        dmax[g * nsys + s_id] = 0x7f800000;   // +inf: scan the back run until the next rebuild
    }
    for (int t = 0; t < S.nb_ntab; ++t) {
        NbTab &T = nb_tabs[S.nb_tab_off + t];
        T.c *= S.unit / u_old;
    }
}

// Sum of p[w] for w = a + t, a + t + stride, ... < b, eight loads in flight;
// the order of the additions is fixed.
__device__ __forceinline__ double strided_sum(const double *__restrict__ p, int a, int b, int t, int stride) {
    double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int w = a + t;
    for (; w + 7 * stride < b; w += 8 * stride) {
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] += p[w + u * stride];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
        if (w + u * stride < b) x[u] += p[w + u * stride];
    return ((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7]));
}

// Observables of one system from its field sums (AVERAGE.f90:10-47 window
// sums included), the temperatures the next MOVE uses, and the MOMENTUM
// correction when the step is a HALT_DRIFT step.
__device__ void finalize_system(const DevSys &S, int s_id, const double *red, double *__restrict__ obs,
                                double *__restrict__ avg, double *__restrict__ sysstate, const int *__restrict__ step,
                                int mode, double *__restrict__ npt) {
    // the field sums: 0..8 nonbonded, 9..21 bonded energies and overextension, 22..27 bonded
    // virial, 28..39 kinetic tensor of each population, 40..42 momentum
    double acc[kFields];
    for (int k = 0; k < kFields; ++k) acc[k] = red[k];
    const double mom[3] = {acc[40], acc[41], acc[42]};
    double *o = obs + (size_t)s_id * kObs;
    double *st = sysstate + (size_t)s_id * 8;
    for (int k = 0; k < 3; ++k) o[k] = acc[k];                 // V_NB
    for (int k = 0; k < 12; ++k) o[3 + k] = acc[9 + k];        // V_BOND, V_ANGLE, V_TORSION, V_OOP
    double pt[6];
    for (int k = 0; k < 6; ++k) pt[k] = acc[3 + k] + acc[22 + k];
    o[27] = acc[21];
    if (mode == 1) {
        for (int k = 0; k < 6; ++k) o[15 + k] = pt[k];
        return;
    }
    double ek0 = 0.5 * (acc[28] + acc[29] + acc[30]);
    double ek1 = 0.5 * (acc[34] + acc[35] + acc[36]);
    for (int k = 0; k < 6; ++k) o[15 + k] = (pt[k] + acc[28 + k] + acc[34 + k]) * (1.0 / S.volume);
    o[21] = ek0;
    o[22] = ek1;
    o[23] = (ek0 + ek1) * S.mktemp;
    o[24] = ek0 * S.mktemp_atom;
    o[25] = ek1 * S.mktemp_bead;
    o[26] = S.volume;
    st[0] = o[24];
    st[1] = o[25];
    const int cur = *step - S.step0;   // this system's own step
    if (S.ensemble == 2) {
        // SCALEBP.f90:18-34: this step's pressure into the ring; every LIMAVP-th step the ring's mean sets the
        // box scale MIU, applied after these sums (k_npt_positions, k_npt_box).  The volume the step reports
        // (and AVERAGE's window sums, taken after SCALEBP) is the rescaled one.
        double *q = npt + (size_t)s_id * kNptState;
        q[((cur - 1) % kLimAvp + kLimAvp) % kLimAvp] = (o[15] + o[16] + o[17]) / 3.0;
        double miu = 1.0;
        if (cur % kLimAvp == 0) {
            double avp = 0.0;
            for (int k = 0; k < kLimAvp; ++k) avp += q[k];
            miu = cbrt(1.0 + (avp / kLimAvp - S.pressure) * S.pfac);
        }
        q[kNptMiu] = miu;
        q[kNptGo] = cur % kLimAvp == 0 ? 1.0 : 0.0;
        o[26] = S.volume * miu * miu * miu;
    }
    if (cur % S.halt_drift == 0) {
        for (int k = 0; k < 3; ++k) st[2 + k] = mom[k] / (double)S.natoms;
    } else {
        st[2] = st[3] = st[4] = 0.0;
    }
    double *a = avg + (size_t)s_id * kObs;
    for (int k = 0; k < 26; ++k) a[k] += o[k];
    a[27] += o[27];
    a[28] += 1.0;
}

// Per-system sums of the workgroup partials: one workgroup per system, whose
// wave w sums fields w, w + 4, ... (0..8 nonbonded of both groups, 9..27
// bonded, 28..42 integrate), each in a fixed order (the lanes stride over the
// partials, then a fixed-order wave sum), so results are deterministic and a
// system's sums do not depend on the other systems of a batch.  The workgroup
// then forms the system's observables.  Workgroups past nsys (the grid is a
// whole number of XCD rounds) do nothing.
#define IBX_FIN_PARAMS                                                                                          \
    const DevSys *__restrict__ sys, const double *__restrict__ part_nb0, const double *__restrict__ part_nb1,      \
        const double *__restrict__ part_bd, const double *__restrict__ part_mv, int snb0, int snb1, int sbd,      \
        int smv, double *__restrict__ fsum, double *__restrict__ obs, double *__restrict__ avg,                   \
        double *__restrict__ sysstate, const int *__restrict__ step, int mode, int nsys, double *__restrict__ npt
#define IBX_FIN_ARGS sys, part_nb0, part_nb1, part_bd, part_mv, snb0, snb1, sbd, smv, fsum, obs, avg, sysstate, step, \
                     mode, nsys, npt

// Workgroup b of the field sums (also one role of k_step_forces).
__device__ __forceinline__ void field_sums_block(const int b, IBX_FIN_PARAMS) {
    __shared__ double red[kFields];
    const int s_id = b;
    if (s_id >= nsys) return;
    const DevSys &S = sys[s_id];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    constexpr int per = kBlock / kNbAtomsPerBlock;   // nonbonded workgroups per 256-slot block
    for (int f = wv; f < kFields; f += kBlock / kWave) {
        double x = 0.0;
        if (f < 9) {
            for (int g = 0; g < S.ngroup; ++g) {
                const double *pp = (g == 0 ? part_nb0 + (size_t)f * snb0 : part_nb1 + (size_t)f * snb1);
                x += strided_sum(pp, per * S.g[g].wg_off, per * (S.g[g].wg_off + S.g[g].nwg), lane, kWave);
            }
        } else if (f < 28) {
            x = strided_sum(part_bd + (size_t)(f - 9) * sbd, S.term_wg_off, S.term_wg_off + S.term_nwg, lane, kWave);
        } else {
            x = strided_sum(part_mv + (size_t)(f - 28) * smv, S.item_wg_off, S.item_wg_off + S.item_nwg, lane, kWave);
        }
        x = wave_sum(x);
        if (lane == 0) red[f] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < kFields; ++k) fsum[(size_t)s_id * kFields + k] = red[k];
        // the step summed is the counter's (the moves of its step are done, the next step's are not)
        const int t = *step;
        if (mode == 1 ? S.live : alive_at(S, t)) finalize_system(S, s_id, red, obs, avg, sysstate, step, mode, npt);
        if (mode == 0 && S.live) sysstate[(size_t)s_id * 8 + 5] = alive_at(S, t + 1) ? 1.0 : 0.0;   // next move
    }
}

__global__ void __launch_bounds__(kBlock) k_field_sums(IBX_FIN_PARAMS) { field_sums_block((int)blockIdx.x, IBX_FIN_ARGS); }

struct NbLaunchK {   // the arguments of k_nonbonded
    const DevSys *sys;
    const int *wg_sys, *perm;
    const int4 *pos4;
    const int *list;
    const int2 *nneb;
    const int *dmax;
    int nsys, cap;
    const float4 *rows;
    const float *xoff;
    const NbTab *tabs;
    const int *map;
    int nitems_total;
    const double *s;
    float *fnb;
    double *part;
    int nblk;
};

struct TermLaunchK {   // the arguments of k_terms
    const int *wg_sys;
    TermArgs A;
    const double *s;
    int nitems_total, nterms;
    Contrib *contrib;
    int ncontrib;
    double *part;
    int *flags;
    int nwg;
};

// One launch for the force phase of a step: the group-0 nonbonded workgroups
// (none on a step that skips the nonbonded force), then the bonded terms, then
// the field sums of the previous step (none on the first step of a run).  The
// roles share the device with one another the way kernels on two streams
// would, but with no dependency between streams to wait on; the field sums
// read the other of the two nonbonded and bonded partial buffers.  Block b of
// the launch's blocks: the previous step's field sums first (a whole number of
// XCD rounds, so the nonbonded blocks keep their XCD-aware numbering), so they
// run beside the nonbonded role rather than after it.
template <bool kMulti, bool kLds, bool kCubic>
__device__ __forceinline__ void step_forces_block(int b, const NbLaunchK &nb, int nb_grid, const TermLaunchK &tk,
                                                  const FinLaunch &fk, int fin_blocks, int &lds_sys, float *acc_lds) {
    if (b < fin_blocks) {
        field_sums_block(b, fk.sys, fk.part_nb0, fk.part_nb1, fk.part_bd, fk.part_mv, fk.snb0, fk.snb1, fk.sbd, fk.smv,
                         fk.fsum, fk.obs, fk.avg, fk.sysstate, fk.step, fk.mode, fk.nsys, fk.npt);
        return;
    }
    b -= fin_blocks;
    if (b < nb_grid) {
        nb_block<kMulti, kLds, 0, kCubic>(b, nb.sys, nb.wg_sys, nb.perm, nb.pos4, nb.list, nb.nneb, nb.dmax, nb.nsys,
                                          nb.cap, nb.rows, nb.xoff, nb.tabs, nb.map, nb.nitems_total, nb.s, nb.fnb,
                                          nb.part, nb.nblk, lds_sys, acc_lds);
        return;
    }
    b -= nb_grid;
    if (b < tk.nwg)
        terms_block(b, tk.nwg, tk.wg_sys, tk.A, tk.s, tk.nitems_total, tk.nterms, tk.contrib, tk.ncontrib, tk.part,
                    tk.flags);
}

// One workgroup per block.
template <bool kMulti, bool kLds, bool kCubic>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) k_step_forces(NbLaunchK nb, int nb_grid, TermLaunchK tk, FinLaunch fk,
                                                        int fin_blocks) {
    int lds_sys = -1;
    step_forces_block<kMulti, kLds, kCubic>((int)blockIdx.x, nb, nb_grid, tk, fk, fin_blocks, lds_sys, nullptr);
}

// Persistent: a grid sized to what the device holds at once works through the
// same blocks from per-XCD work queues.  Block b belongs to queue b % kXcd
// (the XCD the one-block-per-workgroup launch would run it on), and a
// workgroup takes `chunk` blocks at a time from its own XCD's queue in block
// order, then from the other XCDs' queues until every queue is drained; the
// nonbonded tables stay in LDS while consecutive blocks belong to one system.
// Every block computes what it computes in k_step_forces, so the results are
// the same bits.  queue[0 .. kXcd) count the blocks taken from each queue and
// queue[kXcd] the workgroups that have finished; the last to finish clears
// them all for the next launch (no memset, so graphs may hold the launch).
// Every workgroup ends once every queue is drained: no workgroup waits for
// another.
template <bool kMulti, bool kLds, bool kCubic>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
k_step_forces_q(NbLaunchK nb, int nb_grid, TermLaunchK tk, FinLaunch fk, int fin_blocks, int *__restrict__ queue,
                int chunk, int acc_off) {
    extern __shared__ float4 lds_dyn[];
    __shared__ int take, left[kXcd];
    float *const acc_lds = reinterpret_cast<float *>(lds_dyn + acc_off);
    const int total = fin_blocks + nb_grid + tk.nwg;
    int lds_sys = -1;
    int x = (int)blockIdx.x % kXcd;
    for (;;) {
        const int nx = total > x ? (total - x + kXcd - 1) / kXcd : 0;   // blocks x, x + kXcd, ... of queue x
        for (;;) {
            __syncthreads();   // the last block's LDS is no longer read, and take is free
            if (threadIdx.x == 0) take = atomicAdd(&queue[x], chunk);
            __syncthreads();
            const int k0 = take;
            if (k0 >= nx) break;
            const int k1 = min(k0 + chunk, nx);
            for (int k = k0; k < k1; ++k) {
                if (k > k0) __syncthreads();
                step_forces_block<kMulti, kLds, kCubic>(x + kXcd * k, nb, nb_grid, tk, fk, fin_blocks, lds_sys,
                                                        acc_lds);
            }
        }
        // queue x is drained: move to the next XCD's queue that still holds blocks (one read of each)
        if (threadIdx.x < kXcd) {
            const int y = threadIdx.x;
            const int ny = total > y ? (total - y + kXcd - 1) / kXcd : 0;
            left[y] = __hip_atomic_load(&queue[y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < ny;
        }
        __syncthreads();
        int nxt = -1;
        for (int h = 1; h < kXcd && nxt < 0; ++h)
            if (left[(x + h) % kXcd]) nxt = (x + h) % kXcd;
        if (nxt < 0) break;
        x = nxt;
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&queue[kXcd], 1) == (int)gridDim.x - 1) {
            for (int y = 0; y <= kXcd; ++y) atomicExch(&queue[y], 0);
        }
    }
}


// List-skin autotuner (no counterpart in the reference).  Before a scheduled
// rebuild, each system's inner skin delta (the front run holds the pairs closer
// than rc + delta at the build; the back run is scanned once a member has moved
// 0.49 delta) is chosen to minimise the list entries the nonbonded kernel is
// expected to read over the next interval: nmove front runs, plus a back run for
// every step whose displacement, extrapolated linearly from the largest one the
// system reached over the interval that ends (with a 25% margin), passes
// 0.49 delta.  In a uniform fluid the run lengths scale as (rc + delta)^3 and
// rlist^3 - (rc + delta)^3.  Which pairs are listed does not change, and the
// back-run rule keeps every pair within rc whatever delta is.  One thread per
// system and group: the choice depends only on the system's own trajectory, so
// a system tunes the same alone and in a batch.  tune == 0 restores the fixed
// delta = kInnerFrac (rlist - rc).  Runs only where a list is about to be built.
__device__ void tune_skin(DevSys *sys, int nsys, const int *dmax, int nmove, int tune, int t) {
    if (t >= 2 * nsys) return;
    const int g = t / nsys, s_id = t - g * nsys;
    DevSys &S = sys[s_id];
    if (g >= S.ngroup || !S.live) return;
    DevGroup &G = S.g[g];
    const double rc = sqrt(G.rc2), rl = sqrt(G.rl2), skin = fmax(0.0, rl - rc);
    const double u = S.unit;
    double best_delta = kInnerFrac * skin;
    const float d2u = __int_as_float(dmax[t]);
    if (tune && skin > 0.0 && d2u < 3.0e38f) {           // +inf: no interval observed since the last build
        const double d = 1.25 * sqrt((double)d2u) * u;   // nm, after nmove moves
        const double rl3 = rl * rl * rl;
        double best = 1e300;
        for (int i = 2; i <= 32; ++i) {
            const double delta = skin * i / 32.0;
            const double a = (rc + delta) * (rc + delta) * (rc + delta);
            // evaluation j = 0..nmove-1 of the interval sees j moves: scanned while j d / nmove >= 0.49 delta
            int nscan = 0;
            if (d > 0.0) {
                nscan = nmove - (int)ceil(0.49 * delta * nmove / d);
                nscan = nscan < 0 ? 0 : (nscan > nmove ? nmove : nscan);
            }
            const double cost = nmove * a + nscan * (rl3 - a);
            if (cost < best) {
                best = cost;
                best_delta = delta;
            }
        }
    }
    G.rin2u = (float)(((rc + best_delta) / u) * ((rc + best_delta) / u));
    G.dmax2u = (float)((0.49 * best_delta / u) * (0.49 * best_delta / u));
}

// MTS_LOOP.f90:71-224 on the device.  Explicit steps (MOD(STEP, 3+NMTS) = 1, 2, 3) keep the
// nonbonded force of every bead-group item (beads and virtual sites, label 2 or 3) in slot m
// (MTS_SAVEFORCE) and the group's per-workgroup energy and virial partials (MTS_SAVEAUXS, which
// never stores PT23: its slot keeps zero).  The other steps replace both by F3 + (I-3)/2 (F3 - F1)
// (MTS_APPROX, MTS_LOADAUXS).  One thread per item, and per partial.
__global__ void __launch_bounds__(kBlock) k_mts_save(const float *__restrict__ fnb, const int *__restrict__ label,
                                                     int n, const double *__restrict__ part, int npart,
                                                     float *__restrict__ mts_f, double *__restrict__ mts_part, int m) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const int lab = label[i];
        if (lab == 2 || lab == 3)
            for (int k = 0; k < 3; ++k) mts_f[((size_t)m * 3 + k) * n + i] = fnb[(size_t)k * n + i];
    }
    if (i < 9 * npart) {
        const int field = i / npart;
        mts_part[(size_t)m * 9 * npart + i] = field == 8 ? 0.0 : part[i];   // field 8 = PT23
    }
}

__global__ void __launch_bounds__(kBlock) k_mts_approx(float *__restrict__ fnb, const int *__restrict__ label, int n,
                                                       double *__restrict__ part, int npart,
                                                       const float *__restrict__ mts_f,
                                                       const double *__restrict__ mts_part, double I) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double w = (I - 3.0) * 0.5;
    if (i < n) {
        const int lab = label[i];
        if (lab == 2 || lab == 3)
            for (int k = 0; k < 3; ++k) {
                const double f1 = mts_f[(size_t)k * n + i], f3 = mts_f[((size_t)2 * 3 + k) * n + i];
                fnb[(size_t)k * n + i] = (float)(f3 + w * (f3 - f1));
            }
    }
    if (i < 9 * npart) {
        const double p1 = mts_part[i], p3 = mts_part[(size_t)2 * 9 * npart + i];
        part[i] = p3 + w * (p3 - p1);
    }
}

// ------------------------------------------------------------------------------------------------
// launch wrappers

void launch_mts(hipStream_t st, bool save, float *fnb, const int *label, int n, double *part, int npart,
                float *mts_f, double *mts_part, int m, double I) {
    const int work = n > 9 * npart ? n : 9 * npart;
    const dim3 grid((work + kBlock - 1) / kBlock);
    if (save)
        hipLaunchKernelGGL(k_mts_save, grid, dim3(kBlock), 0, st, fnb, label, n, part, npart, mts_f, mts_part, m);
    else
        hipLaunchKernelGGL(k_mts_approx, grid, dim3(kBlock), 0, st, fnb, label, n, part, npart, mts_f, mts_part, I);
}


__global__ void __launch_bounds__(kBlock) k_fill_int(int *__restrict__ p, int n, int v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_int(hipStream_t st, int *p, int n, int v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, p, n, v);
}

void launch_rebuild(hipStream_t st, const RebuildArgs &a) {
    const int ntune = a.tune >= 0 ? (2 * a.nsys + kBlock - 1) / kBlock : 0;
    hipLaunchKernelGGL(k_cell_count, dim3(a.nwg + ntune), dim3(kBlock), 0, st, a.sys, a.wg_sys, a.nwg, a.g, a.mem, a.s,
                       a.nitems_total, a.mem_cell, a.mem_rank, a.cell_count, a.tune_sys, a.nsys, a.dmax_all, a.nmove, a.tune,
                       a.only);
    const int ntile = (a.ncell_total + 1 + kBlock - 1) / kBlock;   // the starts have one more entry
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntile), dim3(kBlock), 0, st, a.cell_count, a.cell_start,
                       a.tile_sum, a.ncell_total);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(kBlock), 0, st, a.tile_sum, ntile);
    hipLaunchKernelGGL(k_scan_add, dim3(ntile), dim3(kBlock), 0, st, a.cell_start, a.tile_sum, a.ncell_total);
    hipLaunchKernelGGL(k_place, dim3(a.nwg), dim3(kBlock), 0, st, a.sys, a.wg_sys, a.g, a.mem, a.mem_cell, a.cell_start,
                       a.mem_rank, a.tmp, a.only);
    hipLaunchKernelGGL(k_sort_cells, dim3((a.ncell_total + kBlock / kSortLanes - 1) / (kBlock / kSortLanes)), dim3(kBlock), 0, st,
                       a.cell_sys, a.sys, a.g, a.cell_start,
                       a.tmp, a.mem, a.perm, a.slot_cell, a.flags, a.ncell_total, a.only);
    hipLaunchKernelGGL(k_pack, dim3(a.nwg), dim3(kBlock), 0, st, a.sys, a.wg_sys, a.g, a.perm, a.s, a.nitems_total,
                       a.itype, a.label, a.r64, a.pos4, a.inv_perm, a.pos4_ref, a.dmax, a.nsys, 0, a.only);
    const int nwg_list = a.nwg * (kBlock / kListAtoms);
    const size_t rdf_lds = a.rdf_hist ? (size_t)a.rdf_nbin * sizeof(unsigned int) : 0;
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(xcd_grid(nwg_list)), dim3(kBlock), rdf_lds, st, a.sys, a.wg_sys, a.g, a.perm,
                           a.inv_perm, a.pos4, a.r64, a.cell_start, a.excl_off, a.excl, a.cap, a.list, a.nneb, a.flags,
                           nwg_list, a.only, a.rdf_hist, a.rdf_samples, a.rdf_ivol, a.rdf_nbin, a.rdf_inv_dr);
    };
    if (a.rdf_hist) a.cubic ? go(k_build_list<true, true>) : go(k_build_list<false, true>);
    else a.cubic ? go(k_build_list<true, false>) : go(k_build_list<false, false>);
}

template <bool M, bool L, int G, bool K>
static void nb_go(hipStream_t st, const NbArgs &a, size_t lds) {
    const int nblk = a.nwg * (kBlock / kNbAtomsPerBlock);
    hipLaunchKernelGGL((k_nonbonded<M, L, G, K>), dim3(xcd_grid(nblk)), dim3(kBlock), lds, st, a.sys, a.wg_sys, a.perm,
                       a.pos4, a.list, a.nneb, a.dmax, a.nsys, a.cap, a.rows, a.xoff, a.tabs, a.map, a.nitems_total,
                       a.s, a.fnb, a.part, nblk);
}

template <bool M, bool L, int G>
static void nb_cubic(hipStream_t st, const NbArgs &a, size_t lds) {
    if (a.cubic) nb_go<M, L, G, true>(st, a, lds);
    else nb_go<M, L, G, false>(st, a, lds);
}

template <int G>
static void nb_dispatch(hipStream_t st, const NbArgs &a) {
    const bool lds = a.max_rows_lds > 0;
    // the staging area also holds the block's energy/virial accumulators at the end
    const size_t bytes = std::max(lds ? (size_t)a.max_rows_lds * sizeof(float4) : 0, (size_t)kNbFields * kBlock * sizeof(float));
    if (a.multi) {
        if (lds) nb_cubic<true, true, G>(st, a, bytes);
        else nb_cubic<true, false, G>(st, a, bytes);
    } else {
        if (lds) nb_cubic<false, true, G>(st, a, bytes);
        else nb_cubic<false, false, G>(st, a, bytes);
    }
}

void launch_nonbonded(hipStream_t st, const NbArgs &a) {
    if (a.g == 0) nb_dispatch<0>(st, a);
    else nb_dispatch<1>(st, a);
}

void launch_terms(hipStream_t st, int nwg, const int *wg_sys, const TermLaunch &b, const double *s, int nitems_total,
                  int nterms, Contrib *contrib, int ncontrib, double *part, int *flags) {
    TermArgs A{b.sys, b.mem, b.meta, b.slot, b.tabs, b.rows, b.xs, b.vs_com};
    hipLaunchKernelGGL(k_terms, dim3(nwg), dim3(kBlock), 0, st, wg_sys, A, s, nitems_total, nterms, contrib, ncontrib,
                       part, flags);
}

void launch_move(hipStream_t st, int nwg, const DevSys *sys, const int *wg_sys, const MoveLaunch &m) {
    MoveArgs A;
    A.s = m.s; A.v = m.v; A.fnb = m.fnb; A.contrib = m.contrib; A.inc_off = m.inc_off;
    A.ncontrib = m.ncontrib; A.ftot = m.ftot; A.itype = m.itype; A.label = m.label;
    A.mass = m.mass; A.invmass = m.invmass; A.vs_of = m.vs_of; A.vs_coeff = m.vs_coeff;
    A.pos4[0] = m.pos4[0]; A.pos4[1] = m.pos4[1]; A.inv_perm[0] = m.inv_perm[0]; A.inv_perm[1] = m.inv_perm[1];
    A.pos4_ref[0] = m.pos4_ref[0]; A.pos4_ref[1] = m.pos4_ref[1]; A.dmax = m.dmax; A.nsys = m.nsys;
    A.sysstate = m.sysstate; A.part = m.part; A.step = m.step; A.nitems_total = m.nitems_total; A.mode = m.mode;
    A.nb_weight = m.nb_weight;
    hipLaunchKernelGGL(k_move, dim3(nwg), dim3(kBlock), 0, st, sys, wg_sys, A);
}

void launch_vs_place(hipStream_t st, int nwg, const DevSys *sys, const int *wg_sys, double *s, int nitems_total,
                     const int *label, const int *itype, const int *vs_center, const int *vs_moff, const int *vs_members,
                     const double *mass, const double *vs_invmass, double *vs_com, int4 *pos4, const int4 *pos4_ref,
                     const int *inv_perm, int *dmax) {
    hipLaunchKernelGGL(k_vs_place, dim3(nwg), dim3(kBlock), 0, st, sys, wg_sys, s, nitems_total, label, itype, vs_center,
                       vs_moff, vs_members, mass, vs_invmass, vs_com, pos4, pos4_ref, inv_perm, dmax);
}

void launch_npt(hipStream_t st, int nwg, DevSys *sys, const int *wg_sys, int nsys, double *npt, double *s,
                double *vs_com, const int *itype, const int *label, int nitems_total, NbTab *nb_tabs, int *dmax) {
    hipLaunchKernelGGL(k_npt_positions, dim3(nwg), dim3(kBlock), 0, st, sys, wg_sys, npt, s, vs_com, itype, label,
                       nitems_total);
    hipLaunchKernelGGL(k_npt_box, dim3((nsys + kBlock - 1) / kBlock), dim3(kBlock), 0, st, sys, nsys, npt, nb_tabs,
                       dmax);
}

void launch_finalize(hipStream_t st, const FinLaunch &f) {
    hipLaunchKernelGGL(k_field_sums, dim3(f.nsys), dim3(kBlock), 0, st, f.sys, f.part_nb0, f.part_nb1,
                       f.part_bd, f.part_mv, f.snb0, f.snb1, f.sbd, f.smv, f.fsum, f.obs, f.avg, f.sysstate,
                       f.step, f.mode, f.nsys, f.npt);
}

template <bool M, bool L, bool K>
static void step_forces_go(hipStream_t st, dim3 grid, size_t lds, const NbLaunchK &nb, int nb_grid,
                           const TermLaunchK &tk, const FinLaunch &fk, int fin_blocks, const ForceTiles &ft,
                           int acc_off) {
    if (ft.persistent)
        hipLaunchKernelGGL((k_step_forces_q<M, L, K>), grid, dim3(kBlock), lds, st, nb, nb_grid, tk, fk, fin_blocks,
                           ft.queue, ft.chunk, acc_off);
    else
        hipLaunchKernelGGL((k_step_forces<M, L, K>), grid, dim3(kBlock), lds, st, nb, nb_grid, tk, fk, fin_blocks);
}

void launch_step_forces(hipStream_t st, const NbArgs *nb, int term_nwg, const int *wg_terms, const TermLaunch &b,
                        const double *s, int nitems_total, int nterms, Contrib *contrib, int ncontrib, double *part_bd,
                        int *flags, const FinLaunch *fin, const ForceTiles &ft) {
    NbLaunchK k{};
    int nb_grid = 0;
    bool multi = false, lds = false, cubic = true;
    size_t bytes = 0;
    if (nb && nb->g == 0) {   // the fused launch carries group 0 only
        k.sys = nb->sys; k.wg_sys = nb->wg_sys; k.perm = nb->perm; k.pos4 = nb->pos4; k.list = nb->list;
        k.nneb = nb->nneb; k.dmax = nb->dmax; k.nsys = nb->nsys; k.cap = nb->cap;
        k.rows = nb->rows; k.xoff = nb->xoff; k.tabs = nb->tabs; k.map = nb->map; k.nitems_total = nb->nitems_total;
        k.s = nb->s; k.fnb = nb->fnb; k.part = nb->part; k.nblk = nb->nwg * (kBlock / kNbAtomsPerBlock);
        nb_grid = xcd_grid(k.nblk);
        multi = nb->multi; lds = nb->max_rows_lds > 0; cubic = nb->cubic;
        bytes = std::max(lds ? (size_t)nb->max_rows_lds * sizeof(float4) : 0, (size_t)kNbFields * kBlock * sizeof(float));
    }
    TermLaunchK t{wg_terms, TermArgs{b.sys, b.mem, b.meta, b.slot, b.tabs, b.rows, b.xs, b.vs_com}, s, nitems_total, nterms,
                  contrib, ncontrib, part_bd, flags, term_nwg};
    FinLaunch f{};
    const int fin_blocks = fin ? xcd_grid(fin->nsys) : 0;
    if (fin) f = *fin;
    dim3 grid(nb_grid + term_nwg + fin_blocks);
    if (grid.x == 0) return;
    int acc_off = 0;   // persistent: the block sums are staged past the table rows (float4 units)
    if (ft.persistent) {
        grid.x = std::min<unsigned>(grid.x, (unsigned)xcd_grid(std::max(ft.grid, kXcd)));
        acc_off = lds ? nb->max_rows_lds : 0;
        bytes = (size_t)acc_off * sizeof(float4) + (size_t)kNbFields * kBlock * sizeof(float);
    }
    if (multi) {
        if (lds) cubic ? step_forces_go<true, true, true>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off)
                       : step_forces_go<true, true, false>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off);
        else cubic ? step_forces_go<true, false, true>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off)
                   : step_forces_go<true, false, false>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off);
    } else {
        if (lds) cubic ? step_forces_go<false, true, true>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off)
                       : step_forces_go<false, true, false>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off);
        else cubic ? step_forces_go<false, false, true>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off)
                   : step_forces_go<false, false, false>(st, grid, bytes, k, nb_grid, t, f, fin_blocks, ft, acc_off);
    }
}


}  // namespace ibx
