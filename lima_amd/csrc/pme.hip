// pme.hip -- particle-mesh Ewald with LIMA's formulas (code/LIMA_ENGINE/src/PME.cuh).
//
//   spread   q * w_x w_y w_z at nodes ix-1..ix+2 of the centred cubic B-spline
//            (PME.cuh:263-300, DeviceAlgorithmsPrivate.cuh:7-13), accumulated as int64 fixed
//            point (2^32 per kC/mol) with global atomics, so the grid is order independent
//            (LIMA uses int32 fixed point for the same reason, PME.cuh:249-252).
//   solve    3D real FFT (own mixed-radix 2/3/4/5 Stockham in LDS), multiply by LIMA's
//            influence function G(k) (PME.cuh:434-495, spline factor multiplied), inverse FFT,
//            1/N (PME.cuh:585-602).  Five passes: x R2C, y, z (forward, G, inverse fused), y, x C2R.
//   gather   phi and E = -grad phi by central differences of the grid (PME.cuh:374-411),
//            rearranged into one 6x6x6 stencil: -sum_n w(n) (phi(n+1) - phi(n-1)) =
//            sum_m phi(m) (w(m+1) - w(m-1)).  Force q E, potE 1/2 (q phi + Eself) (PME.cuh:414-420).
// HBM / L2 bound: per step B = 48 N_q + 34 N_grid bytes of algorithmic traffic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.h"
#include "lima_terms.h"

namespace lima {

namespace {

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

struct FftPlan {
    int N, nrad;
    int radix[8];
    const float2* tw;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
__device__ __forceinline__ float2 rot(float2 a, bool inv) { return inv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x); }

template <int R>
__device__ __forceinline__ void dft(float2* u, bool inv) {
    if (R == 2) {
        const float2 a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    } else if (R == 3) {
        const float s = 0.86602540378443864676f;
        const float2 a = cadd(u[1], u[2]);
        const float2 t = csub(u[0], cscale(a, 0.5f));
        const float2 r = rot(cscale(csub(u[1], u[2]), s), inv);   // -i s (u1 - u2) forward
        u[0] = cadd(u[0], a);
        u[1] = cadd(t, r);
        u[2] = csub(t, r);
    } else if (R == 4) {
        const float2 a0 = cadd(u[0], u[2]), a1 = csub(u[0], u[2]);
        const float2 b0 = cadd(u[1], u[3]), b1 = rot(csub(u[1], u[3]), inv);
        u[0] = cadd(a0, b0);
        u[2] = csub(a0, b0);
        u[1] = cadd(a1, b1);
        u[3] = csub(a1, b1);
    } else {   // R == 5
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
        const float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
        const float2 t1 = cadd(u[0], cadd(cscale(a1, c1), cscale(a2, c2)));
        const float2 t2 = cadd(u[0], cadd(cscale(a1, c2), cscale(a2, c1)));
        const float2 r1 = rot(cadd(cscale(b1, s1), cscale(b2, s2)), inv);
        const float2 r2 = rot(csub(cscale(b1, s2), cscale(b2, s1)), inv);
        u[0] = cadd(u[0], cadd(a1, a2));
        u[1] = cadd(t1, r1);
        u[4] = csub(t1, r1);
        u[2] = cadd(t2, r2);
        u[3] = csub(t2, r2);
    }
}

// One radix-R Stockham stage over nl lines of length N (p = product of the earlier radices):
// butterfly k of a line reads a[k + t N/R], twiddles by exp(-+2 pi i t (k mod p) / (p R)) and
// writes b[(k - k mod p) R + k mod p + t p].  Index divisions use float reciprocals, exact for
// these sizes (nl * N <= 64 K).
template <int R>
__device__ __forceinline__ void fft_stage(const float2* __restrict__ a, float2* __restrict__ b, int nl, int N, int p,
                                          const float2* __restrict__ tw, bool inv) {
    const int nb = N / R;
    const int stride = N / (p * R);
    const float rnb = 1.0f / (float)nb, rp = 1.0f / (float)p;
    for (int idx = threadIdx.x; idx < nl * nb; idx += blockDim.x) {
        const int l = (int)(((float)idx + 0.5f) * rnb);
        const int k = idx - l * nb;
        const int q = (int)(((float)k + 0.5f) * rp);
        const int i = k - q * p;
        const float2* src = a + l * N + k;
        float2 u[R];
#pragma unroll
        for (int t = 0; t < R; t++) u[t] = src[t * nb];
#pragma unroll
        for (int t = 1; t < R; t++) {
            float2 w = tw[t * i * stride];
            if (inv) w.y = -w.y;
            u[t] = cmul(u[t], w);
        }
        dft<R>(u, inv);
        float2* dst = b + l * N + q * p * R + i;
#pragma unroll
        for (int t = 0; t < R; t++) dst[t * p] = u[t];
    }
}

// Stockham auto-sort FFT of nl lines of length N in LDS (line-major); returns the buffer
// holding the result (a or b).  All threads of the block participate.
__device__ float2* fft_lines(float2* a, float2* b, int nl, const FftPlan& P, bool inv) {
    const int N = P.N;
    int p = 1;
    for (int st = 0; st < P.nrad; st++) {
        const int R = P.radix[st];
        switch (R) {
            case 2: fft_stage<2>(a, b, nl, N, p, P.tw, inv); break;
            case 3: fft_stage<3>(a, b, nl, N, p, P.tw, inv); break;
            case 4: fft_stage<4>(a, b, nl, N, p, P.tw, inv); break;
            default: fft_stage<5>(a, b, nl, N, p, P.tw, inv); break;
        }
        __syncthreads();
        float2* tmp = a; a = b; b = tmp;
        p *= R;
    }
    return a;
}

// round(y) as int64 (nearest, ties to even) for |y| < 2^51: y + 1.5 2^52 in FP64 leaves the rounded
// integer in the low mantissa bits.  Four instructions instead of the generic float -> int64 sequence.
__device__ __forceinline__ long long fixed_round(float y) {
    const double m = 6755399441055744.0;   // 1.5 * 2^52
    return __double_as_longlong((double)y + m) - __double_as_longlong(m);
}

struct SpreadArgs {
    int64_t N;
    long long Lq[3];
    int K[3];
    const uint4* w;
    const float4* prm;
    long long* qgrid;
};

__global__ void k_spread(SpreadArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.N) return;
    const float q = a.prm[p].x;
    if (q == 0.f) return;
    const uint4 pw = a.w[p];
    int ix[3];
    float f[3], wx[4], wy[4], wz[4];
    grid_coord(pw.x, a.Lq[0], a.K[0], &ix[0], &f[0]);
    grid_coord(pw.y, a.Lq[1], a.K[1], &ix[1], &f[1]);
    grid_coord(pw.z, a.Lq[2], a.K[2], &ix[2], &f[2]);
    bspline(f[0], wx); bspline(f[1], wy); bspline(f[2], wz);
    for (int c = 0; c < 4; c++) {
        const int Z = wrapi(ix[2] - 1 + c, a.K[2]);
        for (int b = 0; b < 4; b++) {
            const int Y = wrapi(ix[1] - 1 + b, a.K[1]);
            const float qyz = q * wy[b] * wz[c];
            long long* row = a.qgrid + ((int64_t)Z * a.K[1] + Y) * a.K[0];
            for (int x = 0; x < 4; x++) {
                const int X = wrapi(ix[0] - 1 + x, a.K[0]);
                const long long v = fixed_round(qyz * wx[x] * 4294967296.0f);
                atomicAdd((unsigned long long*)&row[X], (unsigned long long)v);
            }
        }
    }
}

struct TileArgs {
    int64_t N;
    long long Lq[3];
    int K[3];
    int T[3], margin, E[3];
    int nt[3];
    int bincap;
    const uint4* bin_pos;   // the tile's particles: wrapped Coord, charge bits
    int* bin_count;
    float* tblock;
    int* flags;
};

// One block per tile: particles of the bin add q w_x w_y w_z (fixed point) into an E^3 LDS block
// whose origin is the tile corner minus one node (E = T + 3).  The block's sums are exact and
// independent of the order of the atomics; it is stored as float (rounded once), half the bytes
// of the int64 sums, and the x pass adds the at most 8 blocks covering a node in a fixed order.
__global__ __launch_bounds__(256) void k_spread_tiles(TileArgs a) {
    extern __shared__ unsigned long long blk[];
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int Ex = a.E[0], Ey = a.E[1], E3 = a.E[0] * a.E[1] * a.E[2];
    const int tx = t % a.nt[0], ty = (t / a.nt[0]) % a.nt[1], tz = t / (a.nt[0] * a.nt[1]);
    const int o[3] = {tx * a.T[0] - 1 - a.margin, ty * a.T[1] - 1 - a.margin, tz * a.T[2] - 1 - a.margin};
    for (int i = threadIdx.x; i < E3; i += blockDim.x) blk[i] = 0ull;
    __syncthreads();
    const int nb = min(a.bin_count[t], a.bincap);
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        const uint4 pw = a.bin_pos[(int64_t)t * a.bincap + k];
        const float q = __uint_as_float(pw.w);
        int ix[3];
        float f[3], wx[4], wy[4], wz[4];
        grid_coord(pw.x, a.Lq[0], a.K[0], &ix[0], &f[0]);
        grid_coord(pw.y, a.Lq[1], a.K[1], &ix[1], &f[1]);
        grid_coord(pw.z, a.Lq[2], a.K[2], &ix[2], &f[2]);
        bspline(f[0], wx); bspline(f[1], wy); bspline(f[2], wz);
        int l[3];
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            int v = ix[d] - 1 - o[d];   // local index of the first stencil node
            v = v < 0 ? v + a.K[d] : (v >= a.K[d] ? v - a.K[d] : v);
            l[d] = v;
            ok = ok && v >= 0 && v + 3 < a.E[d];
        }
        if (!ok) { atomicOr(a.flags, 2); continue; }   // cannot happen with current-position bins
        for (int c = 0; c < 4; c++)
            for (int b = 0; b < 4; b++) {
                const float qyz = q * wy[b] * wz[c];
                unsigned long long* row = blk + ((l[2] + c) * Ey + (l[1] + b)) * Ex + l[0];
                for (int x = 0; x < 4; x++)
                    atomicAdd(&row[x], (unsigned long long)fixed_round(qyz * wx[x] * 4294967296.0f));
            }
    }
    __syncthreads();
    float* out = a.tblock + (int64_t)t * E3;
    for (int i = threadIdx.x; i < E3; i += blockDim.x) out[i] = __ll2float_rn((long long)blk[i]) * (1.0f / 4294967296.0f);
}

// charge at grid point (x, y, z) summed over the tile blocks that cover it, in a fixed order
__device__ __forceinline__ float tiled_q(const float* __restrict__ tblock, const int4* __restrict__ cx,
                                            const int4* __restrict__ cy, const int4* __restrict__ cz, int x, int y, int z,
                                            int Ex, int Ey, int E3, int ntx, int nty) {
    const int4 X = cx[x], Y = cy[y], Z = cz[z];
    float acc = 0.f;
    const int zt[2] = {Z.x, Z.z}, zl[2] = {Z.y, Z.w};
    const int yt[2] = {Y.x, Y.z}, yl[2] = {Y.y, Y.w};
    const int xt[2] = {X.x, X.z}, xl[2] = {X.y, X.w};
    for (int c = 0; c < 2; c++) {
        if (zt[c] < 0) continue;
        for (int b = 0; b < 2; b++) {
            if (yt[b] < 0) continue;
            for (int k = 0; k < 2; k++) {
                if (xt[k] < 0) continue;
                const int t = (zt[c] * nty + yt[b]) * ntx + xt[k];
                acc += tblock[(int64_t)t * E3 + (zl[c] * Ey + yl[b]) * Ex + xl[k]];
            }
        }
    }
    return acc;
}

// x pass, forward, reading the tile blocks instead of a global grid.
__global__ __launch_bounds__(256) void k_fft_x_fwd_tiled(FftPlan P, int Ky, int H, int TY, float scale, const float* __restrict__ tblock,
                                                         const int4* __restrict__ cx, const int4* __restrict__ cy,
                                                         const int4* __restrict__ cz, int Ex, int Ey, int E3, int ntx, int nty,
                                                         float2* __restrict__ cg) {
    extern __shared__ float2 lds[];
    const int N = P.N;
    const int2 lb = xcd_block2();
    const int z = lb.y, y0 = lb.x * TY;
    const int nl = min(TY, Ky - y0);
    float2* a = lds;
    float2* b = lds + TY * N;
    for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
        const int l = i / N, x = i - l * N;
        a[i] = make_float2(tiled_q(tblock, cx, cy, cz, x, y0 + l, z, Ex, Ey, E3, ntx, nty) * scale, 0.f);
    }
    __syncthreads();
    const float2* r = fft_lines(a, b, nl, P, false);
    for (int i = threadIdx.x; i < nl * H; i += blockDim.x) {
        const int l = i / H, k = i - l * H;
        cg[((int64_t)z * Ky + y0 + l) * H + k] = r[l * N + k];
    }
}

// x pass, forward: TY rows of one z plane, fixed point -> density, complex FFT, keep H outputs.
__global__ __launch_bounds__(256) void k_fft_x_fwd(FftPlan P, int Ky, int H, int TY, float scale,
                                                   const long long* __restrict__ qgrid, float2* __restrict__ cg) {
    extern __shared__ float2 lds[];
    const int N = P.N;
    const int2 lb = xcd_block2();
    const int z = lb.y, y0 = lb.x * TY;
    const int nl = min(TY, Ky - y0);
    float2* a = lds;
    float2* b = lds + TY * N;
    for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
        const int l = i / N, x = i - l * N;
        const long long v = qgrid[((int64_t)z * Ky + y0 + l) * N + x];
        a[i] = make_float2(__ll2float_rn(v) * (1.0f / 4294967296.0f) * scale, 0.f);
    }
    __syncthreads();
    const float2* r = fft_lines(a, b, nl, P, false);
    for (int i = threadIdx.x; i < nl * H; i += blockDim.x) {
        const int l = i / H, k = i - l * H;
        cg[((int64_t)z * Ky + y0 + l) * H + k] = r[l * N + k];
    }
}

// x pass, inverse: Hermitian extension of H outputs to N, inverse FFT, real part * norm.
__global__ __launch_bounds__(256) void k_fft_x_inv(FftPlan P, int Ky, int H, int TY, float norm,
                                                   const float2* __restrict__ cg, float* __restrict__ phi) {
    extern __shared__ float2 lds[];
    const int N = P.N;
    const int2 lb = xcd_block2();
    const int z = lb.y, y0 = lb.x * TY;
    const int nl = min(TY, Ky - y0);
    float2* a = lds;
    float2* b = lds + TY * N;
    for (int i0 = 0; i0 < nl * N; i0 += 4 * (int)blockDim.x) {   // 4 loads in flight per thread
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {   // no branch around the loads: past the end, reload the last element
            const int i = min(i0 + u * (int)blockDim.x + (int)threadIdx.x, nl * N - 1);
            const int l = i / N, k = i - l * N;
            const float2* row = cg + ((int64_t)z * Ky + y0 + l) * H;
            v[u] = row[k < H ? k : N - k];
            if (k >= H) v[u].y = -v[u].y;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * (int)blockDim.x + (int)threadIdx.x;
            if (i < nl * N) a[i] = v[u];
        }
    }
    __syncthreads();
    const float2* r = fft_lines(a, b, nl, P, true);
    for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
        const int l = i / N, x = i - l * N;
        phi[((int64_t)z * Ky + y0 + l) * N + x] = r[i].x * norm;
    }
}

// Reciprocal-space virial, taken in the z pass of energy evaluations (rec != nullptr).  With
// E_rec = V_cell / (2 NG) sum_m G(m) |rho^(m)|^2 (the gather's 1/2 sum q phi, by Parseval) and the
// spline weights and sinc(pi m / K) fixed under a strain of the box, each wave vector adds
// W_ab = E_m (delta_ab - 2 (1 / k^2 + 1 / (4 kappa^2)) k_a k_b); the half-complex grid holds the
// x frequencies 1 .. (Kx - 1) / 2 for their conjugates too (weight 2).
struct PmeVir {
    float kf[3];     // 2 PI_L / L_d
    int K[3];
    float inv4k2;    // 1 / (4 kappa^2)
    float efac;      // V_cell / (2 NG)
    double* rec;     // one record per workgroup of the z pass, or nullptr
};

// Strided pass over lines (n * stride) for TK consecutive kx of one outer index.
// mode 0: forward, 1: inverse, 2: forward * green, inverse (z pass).
__global__ __launch_bounds__(256) void k_fft_strided(FftPlan P, int H, int TK, int64_t outer_stride, int64_t stride, int mode,
                                                     int green_y_count, float2* __restrict__ cg, const float* __restrict__ green,
                                                     PmeVir v) {
    extern __shared__ float2 lds[];
    const int N = P.N;
    const int2 lb = xcd_block2();
    const int o = lb.y, k0 = lb.x * TK;
    const int nl = min(TK, H - k0);
    float2* a = lds;
    float2* b = lds + TK * N;
    float2* base = cg + (int64_t)o * outer_stride + k0;
    for (int i0 = 0; i0 < nl * N; i0 += 4 * (int)blockDim.x) {   // 4 loads in flight per thread
        float2 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {   // no branch around the loads: past the end, reload the last element
            const int i = min(i0 + u * (int)blockDim.x + (int)threadIdx.x, nl * N - 1);
            const int n = i / nl, t = i - n * nl;
            v[u] = base[(int64_t)n * stride + t];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int i = i0 + u * (int)blockDim.x + (int)threadIdx.x;
            if (i < nl * N) {
                const int n = i / nl, t = i - n * nl;
                a[t * N + n] = v[u];
            }
        }
    }
    __syncthreads();
    float2* r = fft_lines(a, b, nl, P, mode == 1);
    if (mode == 2) {
        float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        const bool vir = v.rec != nullptr;   // kernel-uniform
        for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
            const int t = i / N, n = i - t * N;
            // z pass: line index n is z, outer o is y
            const float g = green[((int64_t)n * green_y_count + o) * H + k0 + t];
            if (vir) {
                const int kx = k0 + t;
                const int my = 2 * o <= v.K[1] ? o : o - v.K[1], mz = 2 * n <= v.K[2] ? n : n - v.K[2];
                const float ka = v.kf[0] * (float)kx, kb = v.kf[1] * (float)my, kc = v.kf[2] * (float)mz;
                const float k2 = ka * ka + kb * kb + kc * kc;
                if (k2 > 0.f) {
                    const float mult = (kx == 0 || 2 * kx == v.K[0]) ? 1.f : 2.f;
                    const float Em = v.efac * mult * g * (r[i].x * r[i].x + r[i].y * r[i].y);
                    const float c = 2.f * (1.f / k2 + v.inv4k2);
                    w[0] += Em * (1.f - c * ka * ka);
                    w[1] += Em * (1.f - c * kb * kb);
                    w[2] += Em * (1.f - c * kc * kc);
                    w[3] -= Em * c * ka * kb;
                    w[4] -= Em * c * ka * kc;
                    w[5] -= Em * c * kb * kc;
                }
            }
            r[i] = cscale(r[i], g);
        }
        if (vir) {   // FP64 workgroup sums in a fixed order, one record per workgroup
            __shared__ double red[6][4];
            const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const double sw = wave_sum_d((double)w[c]);
                if (lane == 0) red[c][wid] = sw;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                double vd[6];
                for (int c = 0; c < 6; c++) vd[c] = ((red[c][0] + red[c][1]) + red[c][2]) + red[c][3];
                write_record(v.rec + ((int64_t)o * gridDim.x + lb.x) * ACC_N, 0.0, 0.0, 0.0, 0.0, 0.0, vd);
            }
        }
        __syncthreads();
        float2* other = (r == a) ? b : a;
        r = fft_lines(r, other, nl, P, true);
    }
    for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
        const int n = i / nl, t = i - n * nl;
        base[(int64_t)n * stride + t] = r[t * N + n];
    }
}

struct GatherArgs {
    int64_t N;
    long long Lq[3];
    int K[3];
    float inv2h[3];
    float eself;
    const uint4* w;
    const float4* prm;
    const float* phi;
    const int* slot_of;
    float4* fpme;    // [nslots] (fx, fy, fz, potE share), the particle's slot written once
    float* ephi;     // [nslots] 1/2 q phi (energy evaluations)
};

template <bool ENERGY>
__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < a.N) {
        const float q = a.prm[p].x;
        if (q != 0.f) {
            const uint4 pw = a.w[p];
            int ix[3];
            float f[3], wx[4], wy[4], wz[4];
            grid_coord(pw.x, a.Lq[0], a.K[0], &ix[0], &f[0]);
            grid_coord(pw.y, a.Lq[1], a.K[1], &ix[1], &f[1]);
            grid_coord(pw.z, a.Lq[2], a.K[2], &ix[2], &f[2]);
            bspline(f[0], wx); bspline(f[1], wy); bspline(f[2], wz);
            // 6-point weights over nodes ix-2..ix+3: w6 (phi), d6 (difference stencil)
            float w6x[6], w6y[6], w6z[6], d6x[6], d6y[6], d6z[6];
            gather_weights(wx, w6x, d6x);
            gather_weights(wy, w6y, d6y);
            gather_weights(wz, w6z, d6z);
            float ph = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
            for (int mz = 0; mz < 6; mz++) {
                const int Z = wrapi(wrapi(ix[2] - 2 + mz, a.K[2]), a.K[2]);
                for (int my = 0; my < 6; my++) {
                    const int Y = wrapi(wrapi(ix[1] - 2 + my, a.K[1]), a.K[1]);
                    const float* row = a.phi + ((int64_t)Z * a.K[1] + Y) * a.K[0];
                    const float wyz = w6y[my] * w6z[mz];
                    const float dyz = d6y[my] * w6z[mz];
                    const float wdz = w6y[my] * d6z[mz];
                    float rw = 0.f, rd = 0.f;   // the same per-row order as the tiled gather
                    for (int mx = 0; mx < 6; mx++) {
                        const int X = wrapi(wrapi(ix[0] - 2 + mx, a.K[0]), a.K[0]);
                        const float v = row[X];
                        rw = fmaf(w6x[mx], v, rw);
                        rd = fmaf(d6x[mx], v, rd);
                    }
                    if (ENERGY) ph = fmaf(wyz, rw, ph);
                    ex = fmaf(wyz, rd, ex);
                    ey = fmaf(dyz, rw, ey);
                    ez = fmaf(wdz, rw, ez);
                }
            }
            const int s = a.slot_of[p];
            put_force<ENERGY>(a.fpme, s, q * ex * a.inv2h[0], q * ey * a.inv2h[1], q * ez * a.inv2h[2], 0.5f * (q * ph + a.eself));
            if (ENERGY) a.ephi[s] = 0.5f * q * ph;
        }
    }
}

struct GatherTileArgs {
    int K[3], T[3], nt[3];
    long long Lq[3];
    int bincap;
    float inv2h[3];
    float eself;
    const float* phi;
    const int* bin_slot;    // the tile's particles: slot, and wrapped Coord with charge bits
    const uint4* bin_pos;
    int* bin_count;
    float4* fpme;
    float* ephi;
};

__device__ __forceinline__ float pick6(const float* v, int i) {   // v[i] for a per-lane i in 0..5, as a masked sum (no stack array)
    float r = 0.f;
#pragma unroll
    for (int m = 0; m < 6; m++) r = fmaf(v[m], (float)(m == i), r);
    return r;
}

// One block per spreading tile: the potential on nodes corner-2 .. corner+T+2 of the tile (every node
// a 6-point gather stencil of a particle based in the tile reads) is staged in LDS.  Each particle
// of the tile's bin is gathered by 8 lanes: lane j < 6 sums the 36 nodes of plane mz = j, and xor
// shuffles add the partial sums.  The block clears the tile's bin counter for
// the next step's binning.  Energy: one FP64 record per wave.
template <bool ENERGY>
__global__ __launch_bounds__(256) void k_gather_tiles(GatherTileArgs a) {
    extern __shared__ float phib[];
    const int t = xcd_block(blockIdx.x, gridDim.x);
    const int G0 = a.T[0] + 5, G1 = a.T[1] + 5, G2 = a.T[2] + 5;
    const int tx = t % a.nt[0], ty = (t / a.nt[0]) % a.nt[1], tz = t / (a.nt[0] * a.nt[1]);
    const int o0 = tx * a.T[0] - 2, o1 = ty * a.T[1] - 2, o2 = tz * a.T[2] - 2;
    {   // the block's nodes, flat; each thread issues up to 8 global loads before it stores any, so
        // their latencies overlap (index divisions by float reciprocals, exact below 2^22)
        const int n = G0 * G1 * G2;
        const float rG0 = 1.f / (float)G0, rG1 = 1.f / (float)G1;
        for (int b0 = 0; b0 < n; b0 += 8 * (int)blockDim.x) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {   // no branch around the loads: past the end, reload the last node
                const int i = min(b0 + u * (int)blockDim.x + (int)threadIdx.x, n - 1);
                const int r = (int)(((float)i + 0.5f) * rG0), lx = i - r * G0;
                const int lz = (int)(((float)r + 0.5f) * rG1), ly = r - lz * G1;
                const int X = wrapi(o0 + lx, a.K[0]), Y = wrapi(o1 + ly, a.K[1]), Z = wrapi(o2 + lz, a.K[2]);
                v[u] = a.phi[((int64_t)Z * a.K[1] + Y) * a.K[0] + X];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = b0 + u * (int)blockDim.x + (int)threadIdx.x;
                if (i < n) phib[i] = v[u];
            }
        }
    }
    __syncthreads();
    // the bin's particles staged in LDS, GCHUNK at a time, read contiguously from the bin records:
    // one thread per particle computes its slot, charge, node-local base and fractional grid
    // coordinates once (the grid-coordinate division is FP64), and the 8 lanes that gather it
    // read them back
    constexpr int GCHUNK = 256;
    __shared__ int sp[GCHUNK];
    __shared__ float sq[GCHUNK];
    __shared__ float4 sf[GCHUNK];   // fractional coordinates, and the local base node l0 | l1 << 8 | l2 << 16
    const int nb = min(a.bin_count[t], a.bincap);
    const int j = threadIdx.x & 7;
    for (int c0 = 0; c0 < nb; c0 += GCHUNK) {   // uniform trip counts per block
        const int nc = min(GCHUNK, nb - c0);
        if ((int)threadIdx.x < nc) {
            const int64_t e = (int64_t)t * a.bincap + c0 + threadIdx.x;
            const uint4 pw = a.bin_pos[e];
            sp[threadIdx.x] = a.bin_slot[e];
            sq[threadIdx.x] = __uint_as_float(pw.w);
            int ix[3];
            float f[3];
            grid_coord(pw.x, a.Lq[0], a.K[0], &ix[0], &f[0]);
            grid_coord(pw.y, a.Lq[1], a.K[1], &ix[1], &f[1]);
            grid_coord(pw.z, a.Lq[2], a.K[2], &ix[2], &f[2]);
            // the particle's base node lies in this tile (binned from the same positions), so its
            // stencil ix-2 .. ix+3 is local nodes ix - corner .. ix - corner + 5
            const int lb = (ix[0] - tx * a.T[0]) | ((ix[1] - ty * a.T[1]) << 8) | ((ix[2] - tz * a.T[2]) << 16);
            sf[threadIdx.x] = make_float4(f[0], f[1], f[2], __int_as_float(lb));
        }
        __syncthreads();
    for (int k0 = 0; k0 < nc; k0 += blockDim.x >> 3) {
        const int k = k0 + (threadIdx.x >> 3);
        const bool live = k < nc;
        const int p = live ? sp[k] : 0;
        float q = 0.f, ph = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
        if (live) {
            q = sq[k];
            const float4 fk = sf[k];
            float wx[4], wy[4], wz[4];
            bspline(fk.x, wx); bspline(fk.y, wy); bspline(fk.z, wz);
            float w6x[6], w6y[6], w6z[6], d6x[6], d6y[6], d6z[6];
            gather_weights(wx, w6x, d6x);
            gather_weights(wy, w6y, d6y);
            gather_weights(wz, w6z, d6z);
            const int lb = __float_as_int(fk.w);
            const int l0 = lb & 0xff, l1 = (lb >> 8) & 0xff, l2 = lb >> 16;
            if (j < 6) {
                const float wz6 = pick6(w6z, j), dz6 = pick6(d6z, j);
                // per row: the x sums with the weights and their differences (two FMAs per node),
                // then the row's y and z factors
#pragma unroll
                for (int my = 0; my < 6; my++) {
                    const float wyz = w6y[my] * wz6, dyz = d6y[my] * wz6, wdz = w6y[my] * dz6;
                    const float* rowp = phib + ((l2 + j) * G1 + (l1 + my)) * G0 + l0;
                    float rw = 0.f, rd = 0.f;
#pragma unroll
                    for (int mx = 0; mx < 6; mx++) {
                        const float v = rowp[mx];
                        rw = fmaf(w6x[mx], v, rw);
                        rd = fmaf(d6x[mx], v, rd);
                    }
                    if (ENERGY) ph = fmaf(wyz, rw, ph);
                    ex = fmaf(wyz, rd, ex);
                    ey = fmaf(dyz, rw, ey);
                    ez = fmaf(wdz, rw, ez);
                }
            }
        }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            ph += __shfl_xor(ph, o, 64);
            ex += __shfl_xor(ex, o, 64);
            ey += __shfl_xor(ey, o, 64);
            ez += __shfl_xor(ez, o, 64);
        }
        if (live && j == 0) {
            const int s = p;   // the bin holds slots
            put_force<ENERGY>(a.fpme, s, q * ex * a.inv2h[0], q * ey * a.inv2h[1], q * ez * a.inv2h[2], 0.5f * (q * ph + a.eself));
            if (ENERGY) a.ephi[s] = 0.5f * q * ph;
        }
    }
        __syncthreads();   // the next chunk overwrites the staged particles
    }
    __syncthreads();
    if (threadIdx.x == 0) a.bin_count[t] = 0;
}

__global__ void k_zero_ll(long long* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

bool smooth235(int n) { for (int q : {2, 3, 5}) while (n % q == 0) n /= q; return n == 1; }

int derive_grid(double L, double spacing) {   // same rule as the oracle
    int n = (int)std::ceil(L / spacing - 1e-9);
    if (n < 4) n = 4;
    while (!smooth235(n)) n++;
    return n;
}

void plan_radices(int N, int* radix, int* nrad) {
    int n = N, k = 0;
    while (n % 4 == 0) { radix[k++] = 4; n /= 4; }
    while (n % 2 == 0) { radix[k++] = 2; n /= 2; }
    while (n % 3 == 0) { radix[k++] = 3; n /= 3; }
    while (n % 5 == 0) { radix[k++] = 5; n /= 5; }
    if (n != 1 || k > 8) throw LimaFail(LIMA_ERR_BAD_INPUT, "PME grid dimension must factor into 2, 3, 5 (at most 8 factors)");
    *nrad = k;
}

FftPlan make_plan(const Pme& m, int d) {
    FftPlan P{};
    P.N = m.K[d];
    P.nrad = m.nrad[d];
    for (int k = 0; k < 8; k++) P.radix[k] = m.radix[d][k];
    P.tw = m.tw[d];
    return P;
}

TileArgs tile_args(Engine& e, const uint4* w) {
    Pme& m = e.pme;
    TileArgs a{};
    a.N = e.N;
    for (int d = 0; d < 3; d++) { a.Lq[d] = e.Lq[d]; a.K[d] = m.K[d]; a.nt[d] = m.nt[d]; }
    for (int d = 0; d < 3; d++) { a.T[d] = m.T[d]; a.E[d] = m.E[d]; }
    a.margin = m.margin;
    a.bincap = m.bincap;
    a.bin_pos = m.bin_pos; a.bin_count = m.bin_count; a.tblock = m.tblock; a.flags = m.flags;
    return a;
}


}  // namespace

BinArgs bin_args(const Engine& e) {
    const Pme& m = e.pme;
    BinArgs b{};
    if (!m.tiled || !e.elec || m.ncharged == 0) return b;
    for (int d = 0; d < 3; d++) { b.Lq[d] = e.Lq[d]; b.K[d] = m.K[d]; b.T[d] = m.T[d]; b.nt[d] = m.nt[d]; }
    b.bincap = m.bincap;
    b.bin_count = m.bin_count;
    b.bin_slot = m.bin_slot;
    b.bin_pos = m.bin_pos;
    b.flags = m.flags;
    return b;
}


void pme_alloc(Engine& e, const std::vector<float>& q, const std::vector<uint4>& w) {
    Pme& m = e.pme;
    const double spacing = e.prm.pme_spacing_nm > 0 ? e.prm.pme_spacing_nm : 0.12;
    for (int d = 0; d < 3; d++) {
        m.K[d] = e.prm.pme_grid[d] > 0 ? e.prm.pme_grid[d] : derive_grid(e.L[d], spacing);
        if (m.K[d] < 4 || m.K[d] > 512) throw LimaFail(LIMA_ERR_BAD_INPUT, "PME grid dimension out of range [4, 512]");
        plan_radices(m.K[d], m.radix[d], &m.nrad[d]);
    }
    m.H = m.K[0] / 2 + 1;
    const int64_t NG = (int64_t)m.K[0] * m.K[1] * m.K[2];
    const int64_t NC = (int64_t)m.H * m.K[1] * m.K[2];
    HIP_CHECK(hipMalloc(&m.qgrid, NG * sizeof(long long)));
    HIP_CHECK(hipMalloc(&m.phi, NG * sizeof(float)));
    HIP_CHECK(hipMalloc(&m.cgrid, NC * sizeof(float2)));
    HIP_CHECK(hipMalloc(&m.green, NC * sizeof(float)));
    // twiddles exp(-2 pi i k / K) in double, stored as float
    for (int d = 0; d < 3; d++) {
        std::vector<float2> tw(m.K[d]);
        for (int k = 0; k < m.K[d]; k++) {
            const double ang = -2.0 * M_PI * (double)k / (double)m.K[d];
            tw[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
        }
        HIP_CHECK(hipMalloc(&m.tw[d], m.K[d] * sizeof(float2)));
        HIP_CHECK(hipMemcpy(m.tw[d], tw.data(), m.K[d] * sizeof(float2), hipMemcpyHostToDevice));
    }
    // influence function, PME.cuh:434-495, in double (setup only), LIMA's PI constant
    const double PI_L = (double)LIMA_PI_F;
    const double kappa = (double)e.pc.kappa;
    const double kc = (double)e.pc.kc;
    double h[3];
    for (int d = 0; d < 3; d++) h[d] = e.L[d] / m.K[d];
    std::vector<float> g(NC);
    for (int z = 0; z < m.K[2]; z++)
        for (int y = 0; y < m.K[1]; y++)
            for (int x = 0; x < m.H; x++) {
                const int mi[3] = {x, y, z};
                double k2 = 0, S = 1;
                for (int d = 0; d < 3; d++) {
                    const int half = m.K[d] / 2;
                    const int mm = mi[d] <= half ? mi[d] : mi[d] - m.K[d];
                    const double kd = 2.0 * PI_L * mm / e.L[d];
                    k2 += kd * kd;
                    const double kh = kd * h[d] * 0.5;
                    S *= (std::fabs(kh) < 1e-14) ? 1.0 : std::pow(std::sin(kh) / kh, 4);
                }
                double G = 0;
                if (k2 > 1e-14) G = 4.0 * PI_L / k2 * std::exp(-k2 / (4.0 * kappa * kappa)) * S * S * kc;
                g[((size_t)z * m.K[1] + y) * m.H + x] = (float)G;
            }
    HIP_CHECK(hipMemcpy(m.green, g.data(), NC * sizeof(float), hipMemcpyHostToDevice));
    // tiled spreading needs every tile block to fit the grid without wrapping onto itself
    {
        double hmin = 1e30;
        for (int d = 0; d < 3; d++) hmin = std::min(hmin, e.L[d] / m.K[d]);
        (void)hmin;
        m.margin = 0;   // bins are rebuilt from the current positions every step
        // per axis a tile edge T dividing K (5..12, closest to 8): with T >= 3 + 2 margin every grid
        // point is covered by at most two tile blocks per axis
        m.tiled = true;
        for (int d = 0; d < 3; d++) {
            int best = 0;
            for (int t = 5; t <= 12; t++)
                if (m.K[d] % t == 0 && t >= 3 + 2 * m.margin && (best == 0 || std::abs(t - 8) < std::abs(best - 8))) best = t;
            if (best == 0) { m.tiled = false; best = 8; }
            m.T[d] = best;
            m.E[d] = best + 3 + 2 * m.margin;
            m.nt[d] = m.K[d] / best;
            if (m.K[d] < m.E[d] || m.nt[d] < 2) m.tiled = false;
        }
        // large grids: grow the tiles (less halo per tile in the spread and gather blocks) while at
        // least 2048 tiles remain to spread the work over the CUs
        for (bool grew = m.tiled; grew;) {
            grew = false;
            for (int d = 0; d < 3; d++) {
                int t = m.T[d] + 1;
                while (t <= 12 && m.K[d] % t != 0) t++;
                if (t > 12) continue;
                const int nt = m.K[d] / t, E = t + 3 + 2 * m.margin;
                const int64_t tiles = (int64_t)m.nt[0] * m.nt[1] * m.nt[2] / m.nt[d] * nt;
                int64_t e3 = (int64_t)E;
                for (int k = 0; k < 3; k++) if (k != d) e3 *= m.E[k];
                if (tiles < 2048 || nt < 2 || m.K[d] < E || e3 * 8 > 56 * 1024) continue;
                m.T[d] = t; m.E[d] = E; m.nt[d] = nt;
                grew = true;
            }
        }
        if ((int64_t)m.E[0] * m.E[1] * m.E[2] * 8 > 56 * 1024) m.tiled = false;   // LDS block size
        m.ntiles = m.nt[0] * m.nt[1] * m.nt[2];
        if (m.tiled) {
            // bin capacity: twice the fullest bin of the starting positions (a multiple of 64, at
            // least 128); a bin that overflows later sets DF_PMEBIN and the capacity doubles
            {
                std::vector<int> cnt(m.ntiles, 0);
                for (size_t i = 0; i < q.size(); i++) {
                    if (q[i] == 0.f) continue;
                    const uint32_t wc[3] = {w[i].x, w[i].y, w[i].z};
                    int ix[3];
                    for (int d = 0; d < 3; d++) ix[d] = (int)(((uint64_t)wc[d] * (uint64_t)m.K[d]) / (uint64_t)e.Lq[d]);
                    cnt[((ix[2] / m.T[2]) * m.nt[1] + ix[1] / m.T[1]) * m.nt[0] + ix[0] / m.T[0]]++;
                }
                int mx = 0;
                for (int c : cnt) mx = std::max(mx, c);
                m.bincap = std::max(128, (2 * mx + 63) / 64 * 64);
            }
            HIP_CHECK(hipMalloc(&m.bin_count, (m.ntiles + 1) * sizeof(int)));
            HIP_CHECK(hipMemset(m.bin_count, 0, (m.ntiles + 1) * sizeof(int)));   // later cleared by the gather
            HIP_CHECK(hipMalloc(&m.bin_slot, (int64_t)m.ntiles * m.bincap * sizeof(int)));
            HIP_CHECK(hipMalloc(&m.bin_pos, (int64_t)m.ntiles * m.bincap * sizeof(uint4)));
            HIP_CHECK(hipMalloc(&m.tblock, (int64_t)m.ntiles * m.E[0] * m.E[1] * m.E[2] * sizeof(float)));
            m.flags = e.dflags + DF_PMEBIN;
            // per grid coordinate: the (at most two) tiles whose block covers it and the local index
            for (int d = 0; d < 3; d++) {
                std::vector<int4> cov(m.K[d]);
                for (int c = 0; c < m.K[d]; c++) {
                    int found[2][2] = {{-1, 0}, {-1, 0}}, nf = 0;
                    for (int t = 0; t < m.nt[d]; t++) {
                        int loc = c - (t * m.T[d] - 1 - m.margin);
                        loc = ((loc % m.K[d]) + m.K[d]) % m.K[d];
                        if (loc < m.E[d]) {
                            if (nf >= 2) throw LimaFail(LIMA_ERR_BAD_INPUT, "PME tile cover exceeds two tiles");
                            found[nf][0] = t; found[nf][1] = loc; nf++;
                        }
                    }
                    cov[c] = make_int4(found[0][0], found[0][1], found[1][0], found[1][1]);
                }
                HIP_CHECK(hipMalloc(&m.cover[d], m.K[d] * sizeof(int4)));
                HIP_CHECK(hipMemcpy(m.cover[d], cov.data(), m.K[d] * sizeof(int4), hipMemcpyHostToDevice));
            }
        }
    }
    double qsq = 0;
    m.ncharged = 0;
    for (float v : q) { qsq += (double)v * v; if (v != 0.f) m.ncharged++; }
    // PME.cuh:610-630: selfenergyCorrection = -kappa / sqrt(PI) * sum q^2 * kC (stored as float)
    m.eself_total = -kappa / std::sqrt(PI_L) * qsq * kc;
}

void pme_free(Engine& e) {
    Pme& m = e.pme;
    void* ptrs[] = {m.qgrid, m.phi, m.cgrid, m.green, m.tw[0], m.tw[1], m.tw[2], m.bin_count, m.bin_slot, m.bin_pos, m.tblock,
                    m.cover[0], m.cover[1], m.cover[2]};
    for (void* p : ptrs) if (p) (void)hipFree(p);
}

void pme_grow_bins(Engine& e) {
    Pme& m = e.pme;
    if (!m.tiled) return;
    HIP_CHECK(hipDeviceSynchronize());   // no kernel may still use the old bins
    if (m.bin_slot) HIP_CHECK(hipFree(m.bin_slot));
    if (m.bin_pos) HIP_CHECK(hipFree(m.bin_pos));
    m.bin_slot = nullptr;
    m.bin_pos = nullptr;
    m.bincap *= 2;
    HIP_CHECK(hipMalloc(&m.bin_slot, (int64_t)m.ntiles * m.bincap * sizeof(int)));
    HIP_CHECK(hipMalloc(&m.bin_pos, (int64_t)m.ntiles * m.bincap * sizeof(uint4)));
    e.step_graph_valid = false;   // bin arguments are part of the captured kernels
}

int pme_zpass_lines(const Pme& m) { return std::max(1, std::min(8, 3072 / m.K[2])); }

int pme_virial_records(const Engine& e) {
    if (!e.elec || e.pme.ncharged == 0) return 0;
    return (int)nblk(e.pme.H, pme_zpass_lines(e.pme)) * e.pme.K[1];
}

void pme_clear_bins(Engine& e, hipStream_t s) {
    if (e.pme.tiled && e.pme.bin_count) HIP_CHECK(hipMemsetAsync(e.pme.bin_count, 0, (e.pme.ntiles + 1) * sizeof(int), s));
}

void launch_pme(Engine& e, const uint4* w, bool energy, hipStream_t s) {
    if (!e.elec || e.pme.ncharged == 0) return;
    Pme& m = e.pme;
    const int Kx = m.K[0], Ky = m.K[1], Kz = m.K[2], H = m.H;
    const int64_t NG = (int64_t)Kx * Ky * Kz;
    const double vcell = (e.L[0] / Kx) * (e.L[1] / Ky) * (e.L[2] / Kz);
    const FftPlan Px = make_plan(m, 0), Py = make_plan(m, 1), Pz = make_plan(m, 2);
    // LDS tiles stay <= 48 KB: 2 buffers x lines x length x 8 B
    const int TY = std::max(1, std::min(8, 3072 / Kx));   // lines per workgroup: a few hundred workgroups per pass
    if (m.tiled) {
        TileArgs ta = tile_args(e, w);   // bins filled by k_pack (or the integrator) from the same positions
        const int E3 = m.E[0] * m.E[1] * m.E[2];
        hipLaunchKernelGGL(k_spread_tiles, dim3(m.ntiles), dim3(256), (size_t)E3 * sizeof(long long), s, ta);
        hipLaunchKernelGGL(k_fft_x_fwd_tiled, dim3(nblk(Ky, TY), Kz), dim3(256), 2 * TY * Kx * sizeof(float2), s, Px, Ky, H,
                           TY, (float)(1.0 / vcell), m.tblock, m.cover[0], m.cover[1], m.cover[2], m.E[0], m.E[1], E3, m.nt[0],
                           m.nt[1], m.cgrid);
    } else {
        hipLaunchKernelGGL(k_zero_ll, dim3(nblk(NG, 256)), dim3(256), 0, s, m.qgrid, NG);
        SpreadArgs sa{};
        sa.N = e.N;
        for (int d = 0; d < 3; d++) { sa.Lq[d] = e.Lq[d]; sa.K[d] = m.K[d]; }
        sa.w = w; sa.prm = e.p.prm; sa.qgrid = m.qgrid;
        hipLaunchKernelGGL(k_spread, dim3(nblk(e.N, 256)), dim3(256), 0, s, sa);
        hipLaunchKernelGGL(k_fft_x_fwd, dim3(nblk(Ky, TY), Kz), dim3(256), 2 * TY * Kx * sizeof(float2), s, Px, Ky, H, TY,
                           (float)(1.0 / vcell), m.qgrid, m.cgrid);
    }
    const int TKy = std::max(1, std::min(8, 3072 / Ky)), TKz = pme_zpass_lines(m);
    PmeVir none{};
    PmeVir vir{};
    if (energy) {
        const double PI_L = (double)LIMA_PI_F;
        for (int d = 0; d < 3; d++) { vir.kf[d] = (float)(2.0 * PI_L / e.L[d]); vir.K[d] = m.K[d]; }
        vir.inv4k2 = (float)(1.0 / (4.0 * (double)e.pc.kappa * (double)e.pc.kappa));
        vir.efac = (float)(vcell / (2.0 * (double)NG));
        vir.rec = e.epart + (int64_t)e.ereg.pme * ACC_N;
    }
    // y pass: outer z, lines along y (stride H)
    hipLaunchKernelGGL(k_fft_strided, dim3(nblk(H, TKy), Kz), dim3(256), 2 * TKy * Ky * sizeof(float2), s, Py, H, TKy,
                       (int64_t)Ky * H, (int64_t)H, 0, Ky, m.cgrid, m.green, none);
    // z pass with the influence function: outer y, lines along z (stride Ky*H)
    hipLaunchKernelGGL(k_fft_strided, dim3(nblk(H, TKz), Ky), dim3(256), 2 * TKz * Kz * sizeof(float2), s, Pz, H, TKz,
                       (int64_t)H, (int64_t)Ky * H, 2, Ky, m.cgrid, m.green, vir);
    hipLaunchKernelGGL(k_fft_strided, dim3(nblk(H, TKy), Kz), dim3(256), 2 * TKy * Ky * sizeof(float2), s, Py, H, TKy,
                       (int64_t)Ky * H, (int64_t)H, 1, Ky, m.cgrid, m.green, none);
    hipLaunchKernelGGL(k_fft_x_inv, dim3(nblk(Ky, TY), Kz), dim3(256), 2 * TY * Kx * sizeof(float2), s, Px, Ky, H, TY,
                       (float)(1.0 / (double)NG), m.cgrid, m.phi);
    if (m.tiled) {
        GatherTileArgs gt{};
        for (int d = 0; d < 3; d++) {
            gt.K[d] = m.K[d]; gt.T[d] = m.T[d]; gt.nt[d] = m.nt[d]; gt.Lq[d] = e.Lq[d];
            gt.inv2h[d] = (float)(m.K[d] / (2.0 * e.L[d]));
        }
        gt.bincap = m.bincap;
        gt.eself = (float)m.eself_total;
        gt.phi = m.phi; gt.bin_slot = m.bin_slot; gt.bin_pos = m.bin_pos;
        gt.bin_count = m.bin_count; gt.fpme = e.fpme; gt.ephi = e.ephi;
        const size_t lds = (size_t)(m.T[0] + 5) * (m.T[1] + 5) * (m.T[2] + 5) * sizeof(float);
        if (energy) hipLaunchKernelGGL(k_gather_tiles<true>, dim3(m.ntiles), dim3(256), lds, s, gt);
        else hipLaunchKernelGGL(k_gather_tiles<false>, dim3(m.ntiles), dim3(256), lds, s, gt);
        return;
    }
    GatherArgs ga{};
    ga.N = e.N;
    for (int d = 0; d < 3; d++) { ga.Lq[d] = e.Lq[d]; ga.K[d] = m.K[d]; ga.inv2h[d] = (float)(m.K[d] / (2.0 * e.L[d])); }
    ga.eself = (float)m.eself_total;
    ga.w = w; ga.prm = e.p.prm; ga.phi = m.phi; ga.slot_of = e.nl.slot_of; ga.fpme = e.fpme; ga.ephi = e.ephi;
    if (energy) hipLaunchKernelGGL(k_gather<true>, dim3(nblk(e.N, 256)), dim3(256), 0, s, ga);
    else hipLaunchKernelGGL(k_gather<false>, dim3(nblk(e.N, 256)), dim3(256), 0, s, ga);
}

}  // namespace lima
