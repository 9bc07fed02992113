// nlist.hip -- cell-grid (x-y columns) rebuild and cluster -> particle pair lists.
//
// Replaces LIMA's compound-level O(nCompounds^2) list (NeighborLists::updateNlists,
// code/LIMA_ENGINE/src/Neighborlists.cuh:139-317) and the 1-nm block grid
// (Neighborlists.cuh:235-285, BoxGrid.cpp:5-39).  Every distance decision is made in
// integer Coord units (int64), so the list is bit-exact with the oracle's definition
// (oracle/lima_oracle.cpp, build_clusters) for identical positions:
//   columns  n_d = max(1, floor(L_d / cbrt(8 V / N))), column = floor(w_d n_d / Lq_d)
//   clusters 8 consecutive particles of a column ordered by (w_z, particle index)
//   entries  every particle b with shift s when dist^2(w_b + s Lq, bbox_I) <= R^2 (full list:
//            each pair appears once from each side), R = round((cutoff + skin) 1e8); 8-bit mask of
//            the i slots the pair is computed for (not intra, not the particle itself).
// Each i-cluster owns a fixed-capacity slice of the entry array (no count pass / scan).
// HBM bound: per rebuild it reads positions and writes the entry list (16 B / entry).
#include <algorithm>
#include <cmath>

#include "engine.h"

namespace lima {

namespace {

constexpr int MAXCOL = 2048;   // particles per x-y column held in LDS for the z sort

// Column of coordinate v in [0, Lq): floor(v n / Lq), from a double estimate corrected against
// the exact column starts xb[c] = ceil(c Lq / n) (no 64-bit integer division).
__device__ __forceinline__ int col_of(long long v, const long long* __restrict__ xb, int n, double rinv) {
    int c = (int)((double)v * rinv);
    c = min(max(c, 0), n - 1);
    if (xb[c] > v) c--;
    else if (c + 1 < n && xb[c + 1] <= v) c++;
    return c;
}

__device__ __forceinline__ int column_of(const uint4 p, const long long* xbx, const long long* xby, int ncx, int ncy, double rx,
                                         double ry) {
    return col_of((long long)p.y, xby, ncy, ry) * ncx + col_of((long long)p.x, xbx, ncx, rx);
}

constexpr int COL_ITEMS = 4;      // particles per thread in the column count / scatter
constexpr int COL_LDS_MAX = 8192; // columns counted in LDS; more columns use global atomics directly

// Column of every particle and the column populations.  Counts are gathered per block in LDS and
// added to the global counters once per (block, column): same-address global atomics from every
// particle serialise in a few L2 channels.
__global__ __launch_bounds__(256) void k_col_assign(int64_t N, const uint4* __restrict__ w, const long long* __restrict__ xbx,
                                                    const long long* __restrict__ xby, double rx, double ry, int ncx, int ncy,
                                                    int* __restrict__ col_of, int* __restrict__ col_count) {
    extern __shared__ int hist[];
    const int ncol = ncx * ncy;
    const bool lds = ncol <= COL_LDS_MAX;
    if (lds) for (int c = threadIdx.x; c < ncol; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const int64_t a0 = (int64_t)blockIdx.x * blockDim.x * COL_ITEMS;
#pragma unroll
    for (int k = 0; k < COL_ITEMS; k++) {
        const int64_t a = a0 + (int64_t)k * blockDim.x + threadIdx.x;
        if (a < N) {
            const int c = column_of(w[a], xbx, xby, ncx, ncy, rx, ry);
            col_of[a] = c;
            atomicAdd(lds ? &hist[c] : &col_count[c], 1);
        }
    }
    __syncthreads();
    if (lds)
        for (int c = threadIdx.x; c < ncol; c += blockDim.x)
            if (hist[c]) atomicAdd(&col_count[c], hist[c]);
}

// (w_z, particle) keys into their column's segment (order inside a column is settled by the sort).
// Each block reserves one range per column it touches and places its particles in it by their LDS rank.
__global__ __launch_bounds__(256) void k_col_scatter(int64_t N, int ncol, const uint4* __restrict__ w, const int* __restrict__ col_of,
                                                     const int* __restrict__ col_start, int* __restrict__ col_fill,
                                                     unsigned long long* __restrict__ keys) {
    extern __shared__ int hist[];
    const bool lds = ncol <= COL_LDS_MAX;
    if (lds) for (int c = threadIdx.x; c < ncol; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const int64_t a0 = (int64_t)blockIdx.x * blockDim.x * COL_ITEMS;
    int rank[COL_ITEMS], col[COL_ITEMS];
#pragma unroll
    for (int k = 0; k < COL_ITEMS; k++) {
        const int64_t a = a0 + (int64_t)k * blockDim.x + threadIdx.x;
        col[k] = a < N ? col_of[a] : -1;
        rank[k] = 0;
        if (col[k] >= 0) rank[k] = lds ? atomicAdd(&hist[col[k]], 1) : col_start[col[k]] + atomicAdd(&col_fill[col[k]], 1);
    }
    __syncthreads();
    if (lds)
        for (int c = threadIdx.x; c < ncol; c += blockDim.x)
            if (hist[c]) hist[c] = col_start[c] + atomicAdd(&col_fill[c], hist[c]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < COL_ITEMS; k++) {
        const int64_t a = a0 + (int64_t)k * blockDim.x + threadIdx.x;
        if (col[k] >= 0) {
            const int pos = lds ? hist[col[k]] + rank[k] : rank[k];
            keys[pos] = ((unsigned long long)w[a].z << 32) | (unsigned long long)(uint32_t)a;
        }
    }
}

// One block per column: bitonic sort of the column's (w_z, particle) keys in LDS.
__global__ __launch_bounds__(1024) void k_col_sort(const int* __restrict__ col_start, unsigned long long* __restrict__ keys,
                                                   int* __restrict__ col_ncl, int* __restrict__ flags) {
    __shared__ unsigned long long s[MAXCOL];
    const int c = blockIdx.x;
    const int b = col_start[c], n = col_start[c + 1] - b;
    if (threadIdx.x == 0) col_ncl[c] = (n + CLUSTER - 1) / CLUSTER;
    if (n > MAXCOL) {
        if (threadIdx.x == 0) atomicOr(&flags[DF_COLUMN], 1);
        return;
    }
    int P = 1;
    while (P < n) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x) s[i] = i < n ? keys[b + i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const bool asc = (i & k) == 0;
                    const unsigned long long x = s[i], y = s[ixj];
                    if ((x > y) == asc) { s[i] = y; s[ixj] = x; }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) keys[b + i] = s[i];
}

__device__ __forceinline__ int upper_bound_int(const int* a, int n, int v) {   // first index with a[idx] > v
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] <= v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Particle at sorted position p -> cluster slot.
__global__ void k_assign_slots(int64_t N, int ncol, const int* __restrict__ col_start, const int* __restrict__ col_cl_start,
                               const unsigned long long* __restrict__ keys, int* __restrict__ catoms, int* __restrict__ slot_of) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int c = upper_bound_int(col_start, ncol + 1, (int)p) - 1;
    const int k = (int)p - col_start[c];
    const int I = col_cl_start[c] + k / CLUSTER;
    const int a = (int)(uint32_t)(keys[p] & 0xffffffffull);
    catoms[I * CLUSTER + k % CLUSTER] = a;
    slot_of[a] = I * CLUSTER + k % CLUSTER;
}

// One thread per cluster slot; the 8 slots of a cluster are 8 consecutive lanes, which reduce the
// cluster's bounding box with xor shuffles.
__global__ __launch_bounds__(256) void k_cluster_props(const int* __restrict__ ncl_dev, int ncl_max, int ncol, const int* __restrict__ col_cl_start,
                                                       const int* __restrict__ catoms, const uint4* __restrict__ w,
                                                       const float4* __restrict__ prm, const int* __restrict__ comp,
                                                       int* __restrict__ ccol, uint4* __restrict__ cref, uint4* __restrict__ cmin,
                                                       uint4* __restrict__ cmax, float2* __restrict__ ljs, int* __restrict__ comps,
                                                       uint4* __restrict__ cw) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    const int I = slot / CLUSTER;
    const bool live = I < min(*ncl_dev, ncl_max);   // whole 8-lane groups share I
    const int a = live ? catoms[slot] : -1;
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    if (a >= 0) {
        const uint4 p = w[a];
        mn[0] = mx[0] = p.x; mn[1] = mx[1] = p.y; mn[2] = mx[2] = p.z;
        const float4 q = prm[a];
        ljs[slot] = make_float2(q.y, q.z * 4.898979485566356f);   // (sigma/2, sqrt(24 eps))
        comps[slot] = comp[a];
        cw[slot] = make_uint4(p.x, p.y, p.z, 0u);
    } else if (live) {
        ljs[slot] = make_float2(0.f, 0.f);
        comps[slot] = -1;
        cw[slot] = make_uint4(0u, 0u, 0u, 1u);
    }
#pragma unroll
    for (int o = 1; o < CLUSTER; o <<= 1)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            mn[d] = min(mn[d], (uint32_t)__shfl_xor((int)mn[d], o, 64));
            mx[d] = max(mx[d], (uint32_t)__shfl_xor((int)mx[d], o, 64));
        }
    if (live && (slot & (CLUSTER - 1)) == 0) {
        ccol[I] = upper_bound_int(col_cl_start, ncol + 1, I) - 1;
        cmin[I] = make_uint4(mn[0], mn[1], mn[2], 0);
        cmax[I] = make_uint4(mx[0], mx[1], mx[2], 0);
        cref[I] = make_uint4((uint32_t)(((uint64_t)mn[0] + mx[0]) >> 1), (uint32_t)(((uint64_t)mn[1] + mx[1]) >> 1),
                             (uint32_t)(((uint64_t)mn[2] + mx[2]) >> 1), 0);
    }
}

struct NlArgs {
    const int* ncl_dev;      // clusters of this list (device, written by the column scan)
    int ncl_max, ncx, ncy, nC, cap;
    int* flags;
    long long Lq[3];
    long long R, R2;
    const int* col_cl_start;
    const int* catoms;
    const int* comps;
    const uint8_t* bmask;
    const uint4* cw;
    const long long* xb[2];   // column starts along x, y
    double rinv[2];
    const uint4* cref;
    const uint4* cmin;
    const uint4* cmax;
    int* ecount;     // near entries of each cluster (padded)
    int* efar;       // far entries (padded), at the end of the slice
    long long F2;    // cutoff^2 in Coord^2: an entry whose image lies farther from the i box is far
    uint2* ent;
};

// range of column indices whose [lo, hi] interval (after shift) is non-empty
__device__ __forceinline__ bool col_range(long long lo, long long hi, long long Lq, int n, const long long* xb, double rinv, int* c0,
                                          int* c1) {
    if (hi < 0 || lo > Lq - 1) return false;
    lo = lo < 0 ? 0 : lo;
    hi = hi > Lq - 1 ? Lq - 1 : hi;
    *c0 = col_of(lo, xb, n, rinv);
    *c1 = col_of(hi, xb, n, rinv);
    return true;
}

// first index in [lo, hi) whose cluster max z >= z (clusters of a column are z ordered: a cluster
// holds 8 consecutive particles of the column's z order, so both its min z and max z are
// non-decreasing along the column)
__device__ __forceinline__ int first_cluster_reaching(const uint4* cmax, int lo, int hi, long long z) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)cmax[mid].z < z) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// first index in [lo, hi) whose cluster min z > z
__device__ __forceinline__ int first_cluster_beyond(const uint4* cmin, int lo, int hi, long long z) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)cmin[mid].z <= z) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ long long gap1(long long lo, long long hi, long long mn, long long mx) {   // [lo, hi] to [mn, mx]
    return lo > mx ? lo - mx : (hi < mn ? mn - hi : 0);
}

// The shifted column ranges along one axis: for shift s in -1..1, columns [c0[s], c1[s]] (c1 < c0:
// none) whose slab lies within R of [mn, mx].
struct AxisRanges {
    int c0[3], c1[3], n;
};
__device__ __forceinline__ AxisRanges axis_ranges(long long mn, long long mx, long long R, long long Lq, int ncol, const long long* xb,
                                                  double rinv) {
    AxisRanges r;
    r.n = 0;
    for (int s = 0; s < 3; s++) {
        int a0 = 0, a1 = -1;
        if (!col_range(mn - R - (s - 1) * Lq, mx + R - (s - 1) * Lq, Lq, ncol, xb, rinv, &a0, &a1)) { a0 = 0; a1 = -1; }
        r.c0[s] = a0; r.c1[s] = a1;
        r.n += a1 - a0 + 1;
    }
    return r;
}
// k-th column of the concatenated ranges -> (column, shift index 0..2)
__device__ __forceinline__ void axis_pick(const AxisRanges& r, int k, int* col, int* s) {
    int t = 0;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int len = r.c1[q] - r.c0[q] + 1;
        if (k >= 0 && k < len && t == 0) { *col = r.c0[q] + k; *s = q; t = 1; }
        k -= len;
    }
}

constexpr int NL_FLAT = 512;   // flat j-cluster table per wave (LDS), filled in chunks

// The i box moved by minus one image shift: a j box or particle shifted by s lies as far from the
// i box as the unshifted one from this box (exact integer arithmetic either way).  One per shift
// code, built once per wave, so the tests below need no shift decoding.
struct ShiftedBox {
    long long lo[3], hi[3];
};

// Is the shifted box of j-cluster J within R of the i box?  A particle is never nearer than its
// cluster's box, so dropping J here never drops an entry.
__device__ __forceinline__ bool jbox_near(const NlArgs& a, int J, const ShiftedBox& sb) {
    const uint4 lo = a.cmin[J], hi = a.cmax[J];
    const long long bx = min(gap1((long long)lo.x, (long long)hi.x, sb.lo[0], sb.hi[0]), a.R + 1);
    const long long by = min(gap1((long long)lo.y, (long long)hi.y, sb.lo[1], sb.hi[1]), a.R + 1);
    const long long bz = min(gap1((long long)lo.z, (long long)hi.z, sb.lo[2], sb.hi[2]), a.R + 1);
    return bx * bx + by * by + bz * bz <= a.R2;
}

// One wave per i-cluster.  Candidate j-clusters are z-ranges of x-y columns under the 27 image
// shifts.  Lanes resolve up to 64 (y column, x column, z shift) units at once: the unit is dropped
// when the column's x-y slab lies farther than R from the i box, else two binary searches on the
// column's z-ordered cluster bounds give its j-cluster range.  A wave prefix sum expands the ranges
// into an LDS table of (J, shift); lanes test 64 table entries at a time for a shifted box within R
// of the i box and pack the survivors into a second table.  The wave takes 8 survivors x 8 slots
// per pass, with the next pass's loads in flight, and tests each particle against the i box.  Entries go to the cluster's
// fixed-capacity slice [I cap, (I+1) cap): an entry whose image lies within the cutoff of the i box
// goes to the front, one farther away (within cutoff + skin) to the back, and each segment is
// padded to a multiple of 8 with mask-0 entries (counts in ecount[I], efar[I]).  The order does not
// change which entries exist (the integer bounding-box test with cutoff + skin).  A slice that
// would overflow sets the DF_PAIRLIST word and the host rebuilds with a larger capacity (at once, or
// by a rollback when the rebuild was not waited for).
template <bool NARROW>
__global__ __launch_bounds__(256) void k_nlist(NlArgs a) {
    __shared__ uint32_t flat_all[4][NL_FLAT], cand_all[4][NL_FLAT];
    __shared__ ShiftedBox sbox_all[4][27];   // per wave: the i box moved by minus each image shift
    __shared__ uint32_t yword_all[4][27];    // per wave: an entry's y word for each shift code
    const int lane = threadIdx.x & 63;
    // the wave's cluster, marked wave-uniform so that everything derived from it (bounds, compound
    // ids, column ranges) is loaded and kept in scalar registers
    const int I = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (I >= min(*a.ncl_dev, a.ncl_max)) return;   // uniform per wave
    uint32_t* flat = flat_all[threadIdx.x >> 6];   // candidates that passed the box test
    uint32_t* cand = cand_all[threadIdx.x >> 6];   // all candidates of the current chunk
    const uint4 bmn = a.cmin[I], bmx = a.cmax[I];
    const long long mn[3] = {(long long)bmn.x, (long long)bmn.y, (long long)bmn.z};
    const long long mx[3] = {(long long)bmx.x, (long long)bmx.y, (long long)bmx.z};
    ShiftedBox* sbox = sbox_all[threadIdx.x >> 6];
    uint32_t* yword = yword_all[threadIdx.x >> 6];
    if (lane < 27) {   // code = sx + 3 sy + 9 sz, each 0..2 for -1, 0, +1 box
        int cs[3];
        shift_code_axes(lane, &cs[0], &cs[1], &cs[2]);
        ShiftedBox b;
        uint32_t yw = (uint32_t)lane;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const long long sh = shift_of(cs[d], a.Lq[d]);
            b.lo[d] = mn[d] - sh;
            b.hi[d] = mx[d] - sh;
            yw |= (uint32_t)(cs[d] == 0) << (8 + 2 * d) | (uint32_t)(cs[d] == 2) << (9 + 2 * d);
        }
        sbox[lane] = b;
        yword[lane] = yw;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // the i slots' compounds and their bonded-compound bits, wave-uniform (scalar registers); an
    // empty slot gets a compound id no particle has, so it never matches
    int ci[CLUSTER];
    uint32_t bmi[CLUSTER];   // exclusion bits of slot s's compound, 0 when it has none
    uint32_t valid = 0;      // i slots holding a particle
#pragma unroll
    for (int s = 0; s < CLUSTER; s++) {
        const int c = a.comps[I * CLUSTER + s];
        bmi[s] = (c >= 0 && c < a.nC) ? (uint32_t)a.bmask[c] : 0u;
        if (c >= 0) valid |= 1u << s;
        ci[s] = c >= 0 ? c : -0x40000000;
    }
    const AxisRanges ry = axis_ranges(mn[1], mx[1], a.R, a.Lq[1], a.ncy, a.xb[1], a.rinv[1]);
    const AxisRanges rx = axis_ranges(mn[0], mx[0], a.R, a.Lq[0], a.ncx, a.xb[0], a.rinv[0]);
    const int nunits = ry.n * rx.n * 3;
    const int jsub = lane >> 3, sl = lane & 7;
    int cnear = 0, cfar = 0;   // entries of the near (front) and far (back) segments
    const int64_t base = (int64_t)I * a.cap;
    for (int ub = 0; ub < nunits; ub += 64) {
        // ---- this lane's unit: a z range of one shifted column ----
        int jf = 0, nj = 0, code = 13;
        {
            const int u = ub + lane;
            if (u < nunits) {
                const int szi = u % 3, t = u / 3;
                int cx = 0, sxi = 0, cy = 0, syi = 0;
                axis_pick(rx, t % rx.n, &cx, &sxi);
                axis_pick(ry, t / rx.n, &cy, &syi);
                const long long shx = shift_of(sxi, a.Lq[0]), shy = shift_of(syi, a.Lq[1]), shz = shift_of(szi, a.Lq[2]);
                // the column's slab holds x in [ceil(cx Lq / n), ceil((cx+1) Lq / n) - 1]
                const long long xlo = a.xb[0][cx], xhi = a.xb[0][cx + 1] - 1;
                const long long ylo = a.xb[1][cy], yhi = a.xb[1][cy + 1] - 1;
                const long long gx = min(gap1(xlo + shx, xhi + shx, mn[0], mx[0]), a.R + 1);
                const long long gy = min(gap1(ylo + shy, yhi + shy, mn[1], mx[1]), a.R + 1);
                const long long zlo = mn[2] - a.R - shz, zhi = mx[2] + a.R - shz;
                if (gx * gx + gy * gy <= a.R2 && zhi >= 0 && zlo <= a.Lq[2] - 1) {
                    const int col = cy * a.ncx + cx;
                    const int cs = a.col_cl_start[col], ce = a.col_cl_start[col + 1];
                    const int f0 = first_cluster_reaching(a.cmax, cs, ce, zlo);
                    const int f1 = first_cluster_beyond(a.cmin, f0, ce, zhi);
                    jf = f0;
                    nj = f1 - f0;
                    code = sxi + 3 * syi + 9 * szi;
                }
            }
        }
        // ---- candidates of all units, in chunks of NL_FLAT: expand, filter by box, pack ----
        int incl = nj;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        const int ncand = __shfl(incl, 63, 64);
        for (int k0 = 0; k0 < ncand; k0 += NL_FLAT) {
            const int k1 = min(ncand, k0 + NL_FLAT);
            for (int q = max(incl - nj, k0); q < min(incl, k1); q++)
                cand[q - k0] = (uint32_t)(jf + (q - (incl - nj))) | ((uint32_t)code << 24);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // one candidate per lane: keep it when its shifted box lies within R of the i box
            int c1 = 0;   // survivors packed into flat[0, c1)
            for (int r = k0; r < k1; r += 64) {
                const int k = r + lane;
                bool near = false;
                uint32_t t = 0;
                if (k < k1) {
                    t = cand[k - k0];
                    near = jbox_near(a, (int)(t & 0xffffffu), sbox[t >> 24]);
                }
                const unsigned long long bal = __ballot(near);
                if (near) flat[c1 + __popcll(bal & ((1ull << lane) - 1ull))] = t;
                c1 += __popcll(bal);
            }
            const int c0 = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // ---- 8 j-clusters x 8 slots per pass; next pass's loads in flight ----
            auto fetch = [&](int f, int& J, int& cd, uint4& pb, int& cb) {
                const bool ok = f < c1;
                const uint32_t t = flat[ok ? f - c0 : 0];
                J = ok ? (int)(t & 0xffffffu) : I;   // past the end: a valid cluster, never tested
                cd = ok ? (int)(t >> 24) : -1;
                pb = a.cw[J * CLUSTER + sl];
                cb = a.comps[J * CLUSTER + sl];
            };
            int J, cd, cb, nJ, ncd, ncb;
            uint4 pb, npb;
            fetch(c0 + jsub, J, cd, pb, cb);
            for (int f0 = c0; f0 < c1; f0 += CLUSTER) {
                fetch(f0 + CLUSTER + jsub, nJ, ncd, npb, ncb);
                uint32_t mask = 0;
                bool far = false;      // the j image is farther than the cutoff from the whole i box
                if (cd >= 0 && pb.w == 0u) {
                    const ShiftedBox& sb = sbox[cd];
                    bool in;
                    if (NARROW) {
                        // Lq + R + cluster extents < 2^31: every difference below is an exact int32
                        const int px = (int)pb.x, py = (int)pb.y, pz = (int)pb.z;
                        const uint32_t R1 = (uint32_t)a.R + 1u;
                        const uint32_t gx = min((uint32_t)max(0, max((int)sb.lo[0] - px, px - (int)sb.hi[0])), R1);
                        const uint32_t gy = min((uint32_t)max(0, max((int)sb.lo[1] - py, py - (int)sb.hi[1])), R1);
                        const uint32_t gz = min((uint32_t)max(0, max((int)sb.lo[2] - pz, pz - (int)sb.hi[2])), R1);
                        const unsigned long long g2 = (unsigned long long)gx * gx + (unsigned long long)gy * gy + (unsigned long long)gz * gz;
                        in = g2 <= (unsigned long long)a.R2;
                        far = g2 > (unsigned long long)a.F2;
                    } else {
                        const long long gx = min(gap1((long long)pb.x, (long long)pb.x, sb.lo[0], sb.hi[0]), a.R + 1);
                        const long long gy = min(gap1((long long)pb.y, (long long)pb.y, sb.lo[1], sb.hi[1]), a.R + 1);
                        const long long gz = min(gap1((long long)pb.z, (long long)pb.z, sb.lo[2], sb.hi[2]), a.R + 1);
                        const long long g2 = gx * gx + gy * gy + gz * gz;
                        in = g2 <= a.R2;
                        far = g2 > a.F2;
                    }
                    if (in) {
                        // a slot excludes cb only when its compound is within 2 of cb (LIMA's bonded
                        // compounds are neighbours in the compound order): test that first
                        // (branch-free: bitwise, not short-circuit, logic)
                        uint32_t close = 0u;
#pragma unroll
                        for (int s = 0; s < CLUSTER; s++) close |= (uint32_t)((uint32_t)(cb - ci[s] + 2) <= 4u) << s;
                        mask = valid;
                        if (close) {
                            const uint32_t cbin = (uint32_t)(cb >= 0) & (uint32_t)(cb < a.nC);
                            uint32_t excl = 0u;
#pragma unroll
                            for (int s = 0; s < CLUSTER; s++) {
                                const uint32_t dd2 = (uint32_t)(cb - ci[s] + 2);   // 0..4 when close
                                const uint32_t bonded = (bmi[s] >> (dd2 & 7u)) & cbin & ((close >> s) & 1u);
                                excl |= ((uint32_t)(dd2 == 2u) | bonded) << s;
                            }
                            mask = valid & ~excl;
                        }
                        if (J == I && cd == 13) mask &= ~(1u << sl);   // no self pair
                    }
                }
                const bool emit = mask != 0;
                const unsigned long long bn = __ballot(emit && !far), bf = __ballot(emit && far);
                if (emit) {
                    const unsigned long long lt = (1ull << lane) - 1ull;
                    const int pos = far ? cfar + __popcll(bf & lt) : cnear + __popcll(bn & lt);
                    if (pos < a.cap) {   // near from the front, far from the back; a collision is an overflow
                        const int64_t idx = base + (far ? a.cap - 1 - pos : pos);
                        a.ent[idx] = make_uint2((uint32_t)(J * CLUSTER + sl) | (mask << 24), yword[cd]);
                    }
                }
                cnear += __popcll(bn);
                cfar += __popcll(bf);
                J = nJ; cd = ncd; pb = npb; cb = ncb;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    const int pn = (cnear + CLUSTER - 1) & ~(CLUSTER - 1), pf = (cfar + CLUSTER - 1) & ~(CLUSTER - 1);
    if (pn + pf > a.cap) {
        if (lane == 0) {
            atomicOr(&a.flags[DF_PAIRLIST], 1);
            atomicMax(&a.flags[DF_LIST_NEED], pn + pf);
            a.ecount[I] = 0;
            a.efar[I] = 0;
        }
        return;
    }
    // padding entries: the cluster's own first slot, mask 0, no image shift
    const uint2 pad = make_uint2((uint32_t)(I * CLUSTER), 13u);
    for (int k = cnear + lane; k < pn; k += 64) a.ent[base + k] = pad;
    for (int k = cfar + lane; k < pf; k += 64) a.ent[base + a.cap - 1 - k] = pad;
    if (lane == 0) {
        a.ecount[I] = pn;
        a.efar[I] = pf;
    }
}

// Single-block exclusive scan of n ints; out[n] = total.
__global__ __launch_bounds__(1024) void k_exclusive_scan(const int* __restrict__ in, int* __restrict__ out, int n) {
    __shared__ int wsum[16];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < n ? in[i] : 0;
        int x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[wid] = x;
        __syncthreads();
        if (wid == 0) {
            int t = lane < 16 ? wsum[lane] : 0;
            for (int o = 1; o < 16; o <<= 1) {
                const int y = __shfl_up(t, o, 64);
                if (lane >= o) t += y;
            }
            if (lane < 16) wsum[lane] = t;
        }
        __syncthreads();
        const int wprefix = wid > 0 ? wsum[wid - 1] : 0;
        if (i < n) out[i] = carry + wprefix + x - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += wsum[15];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = carry;
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

void launch_exclusive_scan(const int* in, int* out, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_exclusive_scan, dim3(1), dim3(1024), 0, s, in, out, n);
}

void nlist_alloc(Engine& e) {
    NList& nl = e.nl;
    const double V = e.L[0] * e.L[1] * e.L[2];
    const double a = std::cbrt(8.0 * V / (double)std::max<int64_t>(1, e.N));
    nl.ncx = std::max(1, (int)std::floor(e.L[0] / a));
    nl.ncy = std::max(1, (int)std::floor(e.L[1] / a));
    nl.ncol = nl.ncx * nl.ncy;
    for (int d = 0; d < 2; d++) {   // column starts ceil(c Lq / n), exact integer arithmetic
        const int n = d == 0 ? nl.ncx : nl.ncy;
        const unsigned long long Lq = (unsigned long long)e.Lq[d];
        std::vector<long long> xb(n + 1);
        for (int c = 0; c <= n; c++) xb[c] = (long long)(((unsigned long long)c * Lq + (unsigned long long)n - 1) / (unsigned long long)n);
        HIP_CHECK(hipMalloc(&nl.xb[d], (n + 1) * sizeof(long long)));
        HIP_CHECK(hipMemcpy(nl.xb[d], xb.data(), (n + 1) * sizeof(long long), hipMemcpyHostToDevice));
        nl.rinv[d] = (double)n / (double)e.Lq[d];
    }
    nl.ncl_max = (int)(e.N / CLUSTER) + nl.ncol + 1;
    if (nl.ncl_max >= (1 << 24)) throw LimaFail(LIMA_ERR_UNSUPPORTED, "more than 2^24 particle clusters");
    const int64_t slots = (int64_t)nl.ncl_max * CLUSTER;
    HIP_CHECK(hipMalloc(&nl.col_of, e.N * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.col_count, (nl.ncol + 1) * sizeof(int)));
    HIP_CHECK(hipMemset(nl.col_count, 0, (nl.ncol + 1) * sizeof(int)));
    nl.ncl_dev = nl.col_count + nl.ncol;
    HIP_CHECK(hipMalloc(&nl.col_start, (nl.ncol + 1) * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.col_fill, (nl.ncol + 1) * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.keys, e.N * sizeof(unsigned long long)));
    HIP_CHECK(hipMalloc(&nl.col_ncl, (nl.ncol + 2) * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.catoms, slots * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.slot_of, e.N * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.ccol, nl.ncl_max * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.cref, nl.ncl_max * sizeof(uint4)));
    HIP_CHECK(hipMalloc(&nl.cmin, nl.ncl_max * sizeof(uint4)));
    HIP_CHECK(hipMalloc(&nl.cmax, nl.ncl_max * sizeof(uint4)));
    HIP_CHECK(hipMalloc(&nl.ljs, slots * sizeof(float2)));
    HIP_CHECK(hipMalloc(&nl.comps, slots * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.cw, slots * sizeof(uint4)));
    HIP_CHECK(hipMalloc(&nl.xq, slots * sizeof(float4)));
    HIP_CHECK(hipMalloc(&nl.ecount, (nl.ncl_max + 1) * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.efar, (nl.ncl_max + 1) * sizeof(int)));
    for (float4** f : {&e.fpair, &e.fintra, &e.fb, &e.fpme}) {
        HIP_CHECK(hipMalloc(f, slots * sizeof(float4)));
        HIP_CHECK(hipMemset(*f, 0, slots * sizeof(float4)));
    }
    HIP_CHECK(hipMalloc(&e.ephi, slots * sizeof(float)));
    HIP_CHECK(hipMemset(e.ephi, 0, slots * sizeof(float)));
    nl.nent_cap = 0;
    nl.cap = e.prm.nlist_capacity > 0 ? (e.prm.nlist_capacity + 7) & ~7 : 2048;
    nl.ent = nullptr;
}

void nlist_free(Engine& e) {
    NList& nl = e.nl;
    void* ptrs[] = {nl.col_of, nl.col_count, nl.col_start, nl.col_fill, nl.keys, nl.col_ncl, nl.catoms, nl.slot_of,
                    nl.ccol, nl.cref, nl.cmin, nl.cmax, nl.ljs, nl.comps, nl.cw, nl.xb[0], nl.xb[1], nl.xq, nl.ecount, nl.efar,
                    nl.ent, e.fpair, e.fintra, e.fb, e.fpme, e.ephi};
    for (void* p : ptrs) if (p) (void)hipFree(p);
}

namespace {

// The clears and the snapshot that open a rebuild, in one launch: column counters and fill cursors,
// and the PME tile bin counters when the integrator filled them for the old slots (bins !=
// nullptr).  When the list is built from the engine's own positions (snapshot != 0) the per-wave
// displacement maxima restart, and, if no error word is set, the MD state is saved (positions at
// the build, velocities, previous forces, thermostat words, list statistics, energy totals) and
// DF_SNAPSHOTS counts it.  Error words are sticky: a rebuild after an unresolved condition takes no
// snapshot, so the last one stays the state the host can restore.
struct OpenArgs {
    int64_t N;
    int snapshot;
    const uint4* w; const int4* img; const float4* v; const float4* fprev;
    const float* thermo; const int* lstats; const double* acc;
    uint4* wb; int4* ib; float4* vs; float4* fps; float* ths; int* lss; double* accs;
    float2* dpart;
    int* col_count; int* col_fill; int ncol1;
    int* flags;
    int* bins; int nbins;
    // the new list's slots: no particle (catoms -1) until assigned, per-kernel force records and
    // 1/2 q phi zero (a slot a kernel does not own keeps zero until the next rebuild)
    int* catoms; float4* fpair; float4* fintra; float4* fb; float4* fpme; float* ephi; int nslots;
};
__global__ void k_rebuild_open(OpenArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool clean = a.snapshot != 0;
#pragma unroll
    for (int k = 0; k < DF_ERRORS; k++) clean = clean && a.flags[k] == 0;
    if (a.snapshot && p < a.N && (p & 63) == 0) a.dpart[p >> 6] = make_float2(0.f, 0.f);
    if (clean && p < a.N) {
        a.wb[p] = a.w[p];
        a.ib[p] = a.img[p];
        a.vs[p] = a.v[p];
        a.fps[p] = a.fprev[p];
    }
    if (clean && p == 0) {
        a.ths[0] = a.thermo[0]; a.ths[1] = a.thermo[1];
        a.lss[0] = a.lstats[0]; a.lss[1] = a.lstats[1];
        for (int k = 0; k < ACC_N; k++) a.accs[k] = a.acc[k];
        a.flags[DF_SNAPSHOTS] += 1;   // only this thread touches the word; the others read words 0..3
    }
    if (p < a.ncol1) { a.col_count[p] = 0; a.col_fill[p] = 0; }
    if (a.bins && p < a.nbins) a.bins[p] = 0;
    if (p < a.nslots) {
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        a.catoms[p] = -1;
        a.fpair[p] = z; a.fintra[p] = z; a.fb[p] = z; a.fpme[p] = z;
        a.ephi[p] = 0.f;
    }
}

// Clears the error words and sets the snapshot count (after the host has dealt with a condition).
__global__ void k_reset_flags(int* flags, int snapshots) {
    const int k = threadIdx.x;
    if (k < DF_ERRORS || k == DF_LIST_NEED) flags[k] = 0;
    if (k == DF_SNAPSHOTS) flags[k] = snapshots;
}

// Completeness of the list being replaced, from the per-wave displacement maxima: a pair (i, j)
// inside the cutoff at some step of the interval is in the list when |dx_i| + |dx_j| <= skin at that
// step (the build listed every particle within cutoff + skin of the i-cluster's box).  Every
// step's largest plus second-largest displacement is at most G1 + G2, the two largest values of
// the multiset {running wave maximum, running wave second maximum} over all waves.  mode bit 0:
// statistics of a finished interval (*dmax2 = G1, stats[0] counts intervals with G1 + G2 > skin,
// stats[1] = sqrt(G1) as float bits); bit 1: raise DF_STALE when G1 + G2 > skin.
__device__ __forceinline__ void top2_insert(float& t1, float& t2, float v) {
    if (v > t1) { t2 = t1; t1 = v; }
    else if (v > t2) t2 = v;
}

// (t1 >= t2) and (u1 >= u2) -> the two largest of the four
__device__ __forceinline__ void top2_merge(float& t1, float& t2, float u1, float u2) {
    const float hi = fmaxf(t1, u1), lo = fminf(t1, u1);
    t1 = hi;
    t2 = fmaxf(lo, fmaxf(t2, u2));
}

__global__ __launch_bounds__(256) void k_list_stats(const float2* __restrict__ v, int n, float skin, int mode, float* dmax2,
                                                    int* stats, int* flags) {
    __shared__ float r1[4], r2[4];
    float t1 = 0.f, t2 = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float2 m = v[i];
        top2_insert(t1, t2, m.x);
        top2_insert(t1, t2, m.y);
    }
    for (int o = 1; o < 64; o <<= 1) top2_merge(t1, t2, __shfl_xor(t1, o, 64), __shfl_xor(t2, o, 64));
    if ((threadIdx.x & 63) == 0) { r1[threadIdx.x >> 6] = t1; r2[threadIdx.x >> 6] = t2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float g1 = r1[0], g2 = r2[0];
        for (int k = 1; k < 4; k++) top2_merge(g1, g2, r1[k], r2[k]);
        const bool stale = sqrtf(g1) + sqrtf(g2) > skin;
        if (mode & 1) {
            *dmax2 = g1;
            if (stale) stats[0] += 1;
            stats[1] = __float_as_int(sqrtf(g1));
        }
        if ((mode & 2) && stale) atomicOr(&flags[DF_STALE], 1);
    }
}

}  // namespace

int nlist_count(Engine& e) {
    int ncl = 0;
    HIP_CHECK(hipMemcpy(&ncl, e.nl.ncl_dev, sizeof(int), hipMemcpyDeviceToHost));
    e.nl.ncl = ncl;
    return ncl;
}

void nlist_grow(Engine& e, int need) {
    // the next build reallocates the entry buffer
    e.nl.cap = std::max((e.nl.cap * 3 / 2 + 7) & ~7, (int)(((int64_t)need * 3 / 2 + 7) & ~7));
}

void reset_flags(Engine& e, int64_t snapshots, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_flags, dim3(1), dim3(64), 0, s, e.dflags, (int)snapshots);
}

void list_interval_check(Engine& e, hipStream_t s) {
    hipLaunchKernelGGL(k_list_stats, dim3(1), dim3(256), 0, s, (const float2*)e.dpart, (int)((e.N + 63) / 64), e.skin, 2,
                       e.dmax2, e.lstats, e.dflags);
}

void nlist_build(Engine& e, const uint4* w, bool checked) {
    NList& nl = e.nl;
    hipStream_t s = e.s_main;
    const int64_t N = e.N;
    // Every kernel argument stays the same across rebuilds (grids and strides follow ncl_max, the
    // count is read from ncl_dev), so a captured step graph stays valid unless the entry buffer is
    // reallocated.  The integrator's packing of the old slots is stale.
    const bool clear_bins = e.packed && e.pme.tiled && e.elec && e.pme.ncharged > 0;   // tile bins filled for the old slots
    e.packed = false;
    const bool snapshot = w == e.p.w;
    for (int attempt = 0;; attempt++) {
        if (snapshot) {   // the host half of the snapshot this rebuild takes when no error word is set
            SnapRecord r;
            r.ordinal = ++e.snap_issued;
            r.step = e.step;
            r.log_frames = e.log_frames;
            r.n_rebuilds = e.n_rebuilds;
            r.temp_measured = e.temp_measured;
            r.energy_step = e.energy_step;
            r.list_interval = e.list_interval;
            r.builds_clean = e.builds_clean;
            e.snaps.push_back(r);
            if (e.snaps.size() > 16) e.snaps.erase(e.snaps.begin());
        }
        // completeness of the list being replaced, on the device (statistics read by lima_get_status);
        // an MD rebuild (deferred) raises DF_STALE before this rebuild's snapshot, so the snapshot of
        // the interval's start survives for the rollback that runs the interval again
        {
            const bool counted = e.n_rebuilds > 0 && !e.stats_skip;
            const int mode = counted ? (checked ? 1 : 3) : 0;
            hipLaunchKernelGGL(k_list_stats, dim3(1), dim3(256), 0, s, (const float2*)e.dpart, (int)((N + 63) / 64), e.skin,
                               mode, e.dmax2, e.lstats, e.dflags);
        }
        e.stats_skip = false;
        {
            OpenArgs o{};
            o.N = N; o.snapshot = snapshot ? 1 : 0;
            o.w = e.p.w; o.img = e.p.img; o.v = e.p.v; o.fprev = e.p.fprev; o.thermo = e.thermo; o.lstats = e.lstats; o.acc = e.acc;
            o.wb = e.w_build; o.ib = e.img_build; o.vs = e.v_snap; o.fps = e.fprev_snap; o.ths = e.thermo_snap; o.lss = e.lstats_snap;
            o.accs = e.acc_snap; o.dpart = e.dpart;
            o.col_count = nl.col_count; o.col_fill = nl.col_fill; o.ncol1 = nl.ncol + 1;
            o.flags = e.dflags;
            o.bins = clear_bins ? e.pme.bin_count : nullptr; o.nbins = e.pme.ntiles;
            o.catoms = nl.catoms; o.fpair = e.fpair; o.fintra = e.fintra; o.fb = e.fb; o.fpme = e.fpme; o.ephi = e.ephi;
            o.nslots = slots_max(e);
            const int64_t span = std::max<int64_t>(std::max<int64_t>(std::max<int64_t>(N, nl.ncol + 1), o.nslots),
                                                   std::max<int64_t>(1, clear_bins ? e.pme.ntiles : 0));
            hipLaunchKernelGGL(k_rebuild_open, dim3(nblk(span, 256)), dim3(256), 0, s, o);
        }
        const size_t col_lds = nl.ncol <= COL_LDS_MAX ? (size_t)nl.ncol * sizeof(int) : 0;
        hipLaunchKernelGGL(k_col_assign, dim3(nblk(N, 256 * COL_ITEMS)), dim3(256), col_lds, s, N, w, nl.xb[0], nl.xb[1], nl.rinv[0],
                           nl.rinv[1], nl.ncx, nl.ncy, nl.col_of, nl.col_count);
        launch_exclusive_scan(nl.col_count, nl.col_start, nl.ncol, s);
        hipLaunchKernelGGL(k_col_scatter, dim3(nblk(N, 256 * COL_ITEMS)), dim3(256), col_lds, s, N, nl.ncol, w, nl.col_of, nl.col_start,
                           nl.col_fill, nl.keys);
        hipLaunchKernelGGL(k_col_sort, dim3(nl.ncol), dim3(1024), 0, s, nl.col_start, nl.keys, nl.col_ncl, e.dflags);
        // col_count now holds the cluster start of every column and the total, ncl = col_count[ncol]
        // (at most N / 8 + ncol <= ncl_max: each column pads its last cluster)
        launch_exclusive_scan(nl.col_ncl, nl.col_count, nl.ncol, s);
        const int* col_cl_start = nl.col_count;
        const int ns = slots_max(e);
        hipLaunchKernelGGL(k_assign_slots, dim3(nblk(N, 256)), dim3(256), 0, s, N, nl.ncol, nl.col_start, col_cl_start, nl.keys,
                           nl.catoms, nl.slot_of);
        hipLaunchKernelGGL(k_cluster_props, dim3(nblk(ns, 256)), dim3(256), 0, s, nl.ncl_dev, nl.ncl_max, nl.ncol, col_cl_start,
                           nl.catoms, w, e.p.prm, e.p.comp, nl.ccol, nl.cref, nl.cmin, nl.cmax, nl.ljs, nl.comps, nl.cw);
        NlArgs a{};
        a.ncl_dev = nl.ncl_dev; a.ncl_max = nl.ncl_max; a.ncx = nl.ncx; a.ncy = nl.ncy; a.nC = e.nC;
        for (int d = 0; d < 3; d++) a.Lq[d] = e.Lq[d];
        a.R = llround(((double)e.rc + (double)e.skin) * 1e8);
        a.R2 = a.R * a.R;
        a.col_cl_start = col_cl_start; a.catoms = nl.catoms; a.comps = nl.comps; a.bmask = e.bmask; a.cw = nl.cw;
        a.xb[0] = nl.xb[0]; a.xb[1] = nl.xb[1]; a.rinv[0] = nl.rinv[0]; a.rinv[1] = nl.rinv[1];
        a.cref = nl.cref; a.cmin = nl.cmin; a.cmax = nl.cmax; a.ecount = nl.ecount; a.efar = nl.efar; a.flags = e.dflags;
        {
            const long long F = llround((double)e.rc * 1e8);
            a.F2 = F * F;
        }
        const int64_t need = (int64_t)nl.ncl_max * nl.cap;
        if (need > nl.nent_cap) {
            HIP_CHECK(hipStreamSynchronize(s));   // kernels of earlier steps may still read the old buffer
            HIP_CHECK(hipStreamSynchronize(e.s_pme));
            if (nl.ent) HIP_CHECK(hipFree(nl.ent));
            nl.ent = nullptr;
            HIP_CHECK(hipMalloc(&nl.ent, need * sizeof(uint2)));
            nl.nent_cap = need;
            e.step_graph_valid = false;   // the captured pair kernel names the old entry buffer
        }
        a.cap = nl.cap; a.ent = nl.ent;
        if (narrow_coords(e)) hipLaunchKernelGGL(k_nlist<true>, dim3(nblk(nl.ncl_max, 4)), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_nlist<false>, dim3(nblk(nl.ncl_max, 4)), dim3(256), 0, s, a);
        if (!checked) {
            // the status words go to pinned host memory behind an event; issue_steps reads them two
            // rebuilds later (bounded run-ahead), lima_synchronize at once
            const int par = e.rb_parity;
            HIP_CHECK(hipMemcpyAsync(e.hflags + par * DF_N, e.dflags, DF_N * sizeof(int), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipEventRecord(e.ev_rebuild[par], s));
            e.rb_pending[par] = true;
            e.rb_parity ^= 1;
            break;
        }
        // checked: the one host synchronisation of this rebuild
        int hf[DF_N];
        HIP_CHECK(hipMemcpyAsync(hf, e.dflags, sizeof(hf), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (hf[DF_COLUMN]) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "an x-y column holds more than 2048 particles");
        if (hf[DF_PMEBIN] || hf[DF_NONFINITE])
            throw LimaFail(LIMA_ERR_HIP, "internal: a checked rebuild found an unresolved PME or force condition");
        if (!hf[DF_PAIRLIST]) break;
        if (attempt >= 6) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "pair list capacity per cluster exceeded");
        nlist_grow(e, hf[DF_LIST_NEED]);
        // this attempt's snapshot (taken before the overflow) is superseded by the next attempt's
        if (snapshot) { e.snap_issued--; e.snaps.pop_back(); }
        reset_flags(e, e.snap_issued, s);
    }
    HIP_CHECK(hipGetLastError());
    e.steps_since_build = 0;
    e.n_rebuilds++;
}

}  // namespace lima
