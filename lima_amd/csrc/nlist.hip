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

__global__ void k_fill_int(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_col_assign(int64_t N, const uint4* __restrict__ w, long long Lqx, long long Lqy, int ncx, int ncy,
                             int* __restrict__ col_of, int* __restrict__ col_count) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const uint4 p = w[a];
    const int cx = (int)(((unsigned long long)p.x * (unsigned long long)ncx) / (unsigned long long)Lqx);
    const int cy = (int)(((unsigned long long)p.y * (unsigned long long)ncy) / (unsigned long long)Lqy);
    const int c = cy * ncx + cx;
    col_of[a] = c;
    atomicAdd(&col_count[c], 1);
}

__global__ void k_col_scatter(int64_t N, const uint4* __restrict__ w, const int* __restrict__ col_of,
                              const int* __restrict__ col_start, int* __restrict__ col_fill,
                              unsigned long long* __restrict__ keys) {
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int c = col_of[a];
    const int pos = col_start[c] + atomicAdd(&col_fill[c], 1);
    keys[pos] = ((unsigned long long)w[a].z << 32) | (unsigned long long)(uint32_t)a;
}

// One block per column: bitonic sort of the column's (w_z, particle) keys in LDS.
__global__ __launch_bounds__(1024) void k_col_sort(const int* __restrict__ col_start, unsigned long long* __restrict__ keys,
                                                   int* __restrict__ col_ncl, int* __restrict__ flags) {
    __shared__ unsigned long long s[MAXCOL];
    const int c = blockIdx.x;
    const int b = col_start[c], n = col_start[c + 1] - b;
    if (threadIdx.x == 0) col_ncl[c] = (n + CLUSTER - 1) / CLUSTER;
    if (n > MAXCOL) {
        if (threadIdx.x == 0) atomicOr(&flags[0], 1);
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

__global__ void k_cluster_props(int ncl, int ncol, const int* __restrict__ col_cl_start, const int* __restrict__ catoms,
                                const uint4* __restrict__ w, const float4* __restrict__ prm, const int* __restrict__ comp,
                                int* __restrict__ ccol, uint4* __restrict__ cref, uint4* __restrict__ cmin, uint4* __restrict__ cmax,
                                float2* __restrict__ ljs, int* __restrict__ comps) {
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= ncl) return;
    ccol[I] = upper_bound_int(col_cl_start, ncol + 1, I) - 1;
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0, 0, 0};
    for (int s = 0; s < CLUSTER; s++) {
        const int a = catoms[I * CLUSTER + s];
        if (a < 0) {
            ljs[I * CLUSTER + s] = make_float2(0.f, 0.f);
            comps[I * CLUSTER + s] = -1;
            continue;
        }
        const uint4 p = w[a];
        const uint32_t v[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; d++) { mn[d] = min(mn[d], v[d]); mx[d] = max(mx[d], v[d]); }
        const float4 q = prm[a];
        ljs[I * CLUSTER + s] = make_float2(q.y, q.z * 4.898979485566356f);   // (sigma/2, sqrt(24 eps))
        comps[I * CLUSTER + s] = comp[a];
    }
    cmin[I] = make_uint4(mn[0], mn[1], mn[2], 0);
    cmax[I] = make_uint4(mx[0], mx[1], mx[2], 0);
    cref[I] = make_uint4((uint32_t)(((uint64_t)mn[0] + mx[0]) >> 1), (uint32_t)(((uint64_t)mn[1] + mx[1]) >> 1),
                         (uint32_t)(((uint64_t)mn[2] + mx[2]) >> 1), 0);
}

struct NlArgs {
    int ncl, ncx, ncy, nC, cap;
    int* flags;
    long long Lq[3];
    long long R, R2;
    const int* col_cl_start;
    const int* catoms;
    const int* comps;
    const uint8_t* bmask;
    const uint4* w;
    const uint4* cref;
    const uint4* cmin;
    const uint4* cmax;
    int* ecount;
    const int* eoff;
    float4* ent;
    uint8_t* ecode;
};

__device__ __forceinline__ bool is_intra(int ca, int cb, int nC, const uint8_t* bmask) {
    if (ca == cb) return true;
    if (ca < 0 || cb < 0 || ca >= nC || cb >= nC) return false;
    const int d = cb - ca;
    if (d < -2 || d > 2) return false;
    return (bmask[ca] >> (d + 2)) & 1;
}

// range of column indices whose [lo, hi] interval (after shift) is non-empty
__device__ __forceinline__ bool col_range(long long lo, long long hi, long long Lq, int n, int* c0, int* c1) {
    if (hi < 0 || lo > Lq - 1) return false;
    lo = lo < 0 ? 0 : lo;
    hi = hi > Lq - 1 ? Lq - 1 : hi;
    *c0 = (int)(((unsigned long long)lo * (unsigned long long)n) / (unsigned long long)Lq);
    *c1 = (int)(((unsigned long long)hi * (unsigned long long)n) / (unsigned long long)Lq);
    return true;
}

// first index in [lo, hi) whose cluster max z >= z (clusters of a column are z ordered)
__device__ __forceinline__ int first_cluster_reaching(const uint4* cmax, int lo, int hi, long long z) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((long long)cmax[mid].z < z) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// One wave per i-cluster, one pass: entries go to the cluster's fixed-capacity slice
// [I cap, (I+1) cap); the count (padded to a multiple of 8 with mask-0 entries) goes to ecount[I].
// A slice that would overflow sets flags[1] and the host rebuilds with a larger capacity.
__global__ __launch_bounds__(256) void k_nlist(NlArgs a) {
    const int lane = threadIdx.x & 63;
    const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (I >= a.ncl) return;
    const uint4 bmn = a.cmin[I], bmx = a.cmax[I], refI = a.cref[I];
    const long long mn[3] = {(long long)bmn.x, (long long)bmn.y, (long long)bmn.z};
    const long long mx[3] = {(long long)bmx.x, (long long)bmx.y, (long long)bmx.z};
    int ci[CLUSTER];
    for (int s = 0; s < CLUSTER; s++) ci[s] = a.comps[I * CLUSTER + s];
    const int jsub = lane >> 3, sl = lane & 7;
    int count = 0;
    const int64_t base = (int64_t)I * a.cap;
    for (int sy = -1; sy <= 1; sy++) {
        int cy0, cy1;
        if (!col_range(mn[1] - a.R - sy * a.Lq[1], mx[1] + a.R - sy * a.Lq[1], a.Lq[1], a.ncy, &cy0, &cy1)) continue;
        for (int cy = cy0; cy <= cy1; cy++) {
            for (int sx = -1; sx <= 1; sx++) {
                int cx0, cx1;
                if (!col_range(mn[0] - a.R - sx * a.Lq[0], mx[0] + a.R - sx * a.Lq[0], a.Lq[0], a.ncx, &cx0, &cx1)) continue;
                for (int cx = cx0; cx <= cx1; cx++) {
                    const int col = cy * a.ncx + cx;
                    const int cs = a.col_cl_start[col], ce = a.col_cl_start[col + 1];
                    if (cs >= ce) continue;
                    for (int sz = -1; sz <= 1; sz++) {
                        const long long zlo = mn[2] - a.R - sz * a.Lq[2], zhi = mx[2] + a.R - sz * a.Lq[2];
                        if (zhi < 0 || zlo > a.Lq[2] - 1) continue;
                        const int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
                        const long long sh[3] = {sx * a.Lq[0], sy * a.Lq[1], sz * a.Lq[2]};
                        const int Jfirst = first_cluster_reaching(a.cmax, cs, ce, zlo);
                        for (int Jb = Jfirst; Jb < ce; Jb += CLUSTER) {
                            if ((long long)a.cmin[Jb].z > zhi) break;   // z ordered: nothing further reaches
                            const int J = Jb + jsub;
                            bool emit = false;
                            uint32_t mask = 0;
                            int bslot = 0;
                            if (J < ce && (long long)a.cmin[J].z <= zhi) {
                                bslot = J * CLUSTER + sl;
                                const int b = a.catoms[bslot];
                                if (b >= 0) {
                                    const uint4 pb = a.w[b];
                                    const long long pv[3] = {(long long)pb.x + sh[0], (long long)pb.y + sh[1], (long long)pb.z + sh[2]};
                                    long long d2 = 0;
                                    for (int d = 0; d < 3; d++) {
                                        long long g = 0;
                                        if (pv[d] < mn[d]) g = mn[d] - pv[d];
                                        else if (pv[d] > mx[d]) g = pv[d] - mx[d];
                                        d2 += g * g;
                                    }
                                    if (d2 <= a.R2) {
                                        const int cb = a.comps[bslot];
                                        for (int s = 0; s < CLUSTER; s++) {
                                            if (ci[s] < 0) continue;
                                            if (is_intra(ci[s], cb, a.nC, a.bmask)) continue;
                                            if (J == I && code == 13 && s == sl) continue;   // no self pair
                                            mask |= 1u << s;
                                        }
                                        emit = mask != 0;
                                    }
                                }
                            }
                            const unsigned long long bal = __ballot(emit);
                            if (emit) {
                                const int pos = count + __popcll(bal & ((1ull << lane) - 1ull));
                                if (pos < a.cap) {
                                    const uint4 rj = a.cref[J];
                                    const float sxf = (float)((long long)rj.x - (long long)refI.x + sh[0]) * LIMA_LIMA_TO_NANO_F;
                                    const float syf = (float)((long long)rj.y - (long long)refI.y + sh[1]) * LIMA_LIMA_TO_NANO_F;
                                    const float szf = (float)((long long)rj.z - (long long)refI.z + sh[2]) * LIMA_LIMA_TO_NANO_F;
                                    a.ent[base + pos] = make_float4(sxf, syf, szf, __uint_as_float((uint32_t)bslot | (mask << 24)));
                                    a.ecode[base + pos] = (uint8_t)code;
                                }
                            }
                            count += __popcll(bal);
                        }
                    }
                }
            }
        }
    }
    const int padded = (count + CLUSTER - 1) & ~(CLUSTER - 1);
    if (padded > a.cap) {
        if (lane == 0) { atomicOr(&a.flags[1], 1); a.ecount[I] = 0; }
        return;
    }
    for (int k = count + lane; k < padded; k += 64) {
        a.ent[base + k] = make_float4(0.f, 0.f, 0.f, __uint_as_float((uint32_t)(I * CLUSTER)));
        a.ecode[base + k] = 13;
    }
    if (lane == 0) {
        a.ecount[I] = padded;
        atomicAdd(&a.flags[2], padded);
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
    nl.ncl_max = (int)(e.N / CLUSTER) + nl.ncol + 1;
    const int64_t slots = (int64_t)nl.ncl_max * CLUSTER;
    HIP_CHECK(hipMalloc(&nl.col_of, e.N * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.col_count, (nl.ncol + 1) * sizeof(int)));
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
    HIP_CHECK(hipMalloc(&nl.xq, slots * sizeof(float4)));
    HIP_CHECK(hipMalloc(&nl.ecount, (nl.ncl_max + 1) * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.eoff, (nl.ncl_max + 2) * sizeof(int)));
    HIP_CHECK(hipMalloc(&nl.flags, 4 * sizeof(int)));
    HIP_CHECK(hipMalloc(&e.fs, 3 * slots * sizeof(float)));
    HIP_CHECK(hipMalloc(&e.es, slots * sizeof(float)));
    nl.nent_cap = 0;
    nl.cap = 2048;
    nl.ent = nullptr;
    nl.ecode = nullptr;
}

void nlist_free(Engine& e) {
    NList& nl = e.nl;
    void* ptrs[] = {nl.col_of, nl.col_count, nl.col_start, nl.col_fill, nl.keys, nl.col_ncl, nl.catoms, nl.slot_of,
                    nl.ccol, nl.cref, nl.cmin, nl.cmax, nl.ljs, nl.comps, nl.xq, nl.ecount, nl.eoff, nl.flags,
                    nl.ent, nl.ecode, e.fs, e.es};
    for (void* p : ptrs) if (p) (void)hipFree(p);
}

namespace {

__global__ void k_snapshot(int64_t N, const uint4* __restrict__ w, const int4* __restrict__ img, uint4* __restrict__ wb,
                           int4* __restrict__ ib, float* dpart) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    if ((p & 63) == 0) dpart[p >> 6] = 0.f;
    wb[p] = w[p];
    ib[p] = img[p];
}

__global__ __launch_bounds__(256) void k_max_reduce(const float* __restrict__ v, int n, float* out) {
    __shared__ float red[4];
    float m = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = fmaxf(m, v[i]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) *out = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

}  // namespace

void nlist_build(Engine& e, const uint4* w) {
    NList& nl = e.nl;
    hipStream_t s = e.s_main;
    const int64_t N = e.N;
    {   // a list is stale once two particles may have closed the skin: 2 max|dx| > skin
        float d2 = 0.f;
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, s, e.dpart, (int)((N + 63) / 64), e.dmax2);
        HIP_CHECK(hipMemcpyAsync(&d2, e.dmax2, sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        e.last_dmax = std::sqrt((double)d2);
        if (e.n_rebuilds > 0 && 2.0 * e.last_dmax > (double)e.skin) e.list_violations++;
    }
    if (w == e.p.w) hipLaunchKernelGGL(k_snapshot, dim3(nblk(N, 256)), dim3(256), 0, s, N, e.p.w, e.p.img, e.w_build, e.img_build, e.dpart);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nl.ncol + 1, 256)), dim3(256), 0, s, nl.col_count, nl.ncol + 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nl.ncol + 1, 256)), dim3(256), 0, s, nl.col_fill, nl.ncol + 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(4), 0, s, nl.flags, 4, 0);
    hipLaunchKernelGGL(k_col_assign, dim3(nblk(N, 256)), dim3(256), 0, s, N, w, (long long)e.Lq[0], (long long)e.Lq[1],
                       nl.ncx, nl.ncy, nl.col_of, nl.col_count);
    launch_exclusive_scan(nl.col_count, nl.col_start, nl.ncol, s);
    hipLaunchKernelGGL(k_col_scatter, dim3(nblk(N, 256)), dim3(256), 0, s, N, w, nl.col_of, nl.col_start, nl.col_fill, nl.keys);
    hipLaunchKernelGGL(k_col_sort, dim3(nl.ncol), dim3(1024), 0, s, nl.col_start, nl.keys, nl.col_ncl, nl.flags);
    launch_exclusive_scan(nl.col_ncl, nl.col_count, nl.ncol, s);   // col_count now holds the cluster start per column
    int hflags[4];
    int ncl = 0;
    HIP_CHECK(hipMemcpyAsync(hflags, nl.flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&ncl, nl.col_count + nl.ncol, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (hflags[0]) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "an x-y column holds more than 2048 particles");
    if (ncl > nl.ncl_max) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "cluster count exceeds capacity");
    nl.ncl = ncl;
    const int* col_cl_start = nl.col_count;
    hipLaunchKernelGGL(k_fill_int, dim3(nblk((int64_t)ncl * CLUSTER, 256)), dim3(256), 0, s, nl.catoms, ncl * CLUSTER, -1);
    hipLaunchKernelGGL(k_assign_slots, dim3(nblk(N, 256)), dim3(256), 0, s, N, nl.ncol, nl.col_start, col_cl_start, nl.keys,
                       nl.catoms, nl.slot_of);
    hipLaunchKernelGGL(k_cluster_props, dim3(nblk(ncl, 256)), dim3(256), 0, s, ncl, nl.ncol, col_cl_start, nl.catoms, w,
                       e.p.prm, e.p.comp, nl.ccol, nl.cref, nl.cmin, nl.cmax, nl.ljs, nl.comps);
    NlArgs a{};
    a.ncl = ncl; a.ncx = nl.ncx; a.ncy = nl.ncy; a.nC = e.nC;
    for (int d = 0; d < 3; d++) a.Lq[d] = e.Lq[d];
    a.R = llround(((double)e.rc + (double)e.skin) * 1e8);
    a.R2 = a.R * a.R;
    a.col_cl_start = col_cl_start; a.catoms = nl.catoms; a.comps = nl.comps; a.bmask = e.bmask; a.w = w;
    a.cref = nl.cref; a.cmin = nl.cmin; a.cmax = nl.cmax; a.ecount = nl.ecount; a.flags = nl.flags;
    for (int attempt = 0;; attempt++) {
        const int64_t need = (int64_t)nl.ncl_max * nl.cap;
        if (need > nl.nent_cap) {
            if (nl.ent) HIP_CHECK(hipFree(nl.ent));
            if (nl.ecode) HIP_CHECK(hipFree(nl.ecode));
            nl.ent = nullptr; nl.ecode = nullptr;
            HIP_CHECK(hipMalloc(&nl.ent, need * sizeof(float4)));
            HIP_CHECK(hipMalloc(&nl.ecode, need));
            nl.nent_cap = need;
        }
        a.cap = nl.cap; a.ent = nl.ent; a.ecode = nl.ecode;
        hipLaunchKernelGGL(k_nlist, dim3(nblk(ncl, 4)), dim3(256), 0, s, a);
        HIP_CHECK(hipMemcpyAsync(hflags, nl.flags, sizeof(hflags), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        if (!hflags[1]) {
            nl.nent = hflags[2];
            break;
        }
        if (attempt >= 6) throw LimaFail(LIMA_ERR_LIST_OVERFLOW, "pair list capacity per cluster exceeded");
        nl.cap = (nl.cap * 3 / 2 + 7) & ~7;
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(4), 0, s, nl.flags, 4, 0);
    }
    pme_bin(e, w, s);
    e.ereg = energy_regions(e);
    if (e.ereg.total > e.epart_cap) {
        if (e.epart) HIP_CHECK(hipFree(e.epart));
        e.epart_cap = (int64_t)(e.ereg.total * 1.25) + 64;
        HIP_CHECK(hipMalloc(&e.epart, e.epart_cap * ACC_N * sizeof(double)));
    }
    HIP_CHECK(hipGetLastError());
    e.steps_since_build = 0;
    e.n_rebuilds++;
}

}  // namespace lima
