/*
 * lima_oracle.cpp -- host FP64 oracle of LIMA's per-step force / minimization path.
 *
 * TEST INFRASTRUCTURE ONLY (see lima_oracle.h).  Never linked by the product.
 *
 * Reference lines restated here (paths relative to the LIMA tree):
 *   LJ pair                LIMA_ENGINE/src/LennardJonesInteractions.cuh:74-108
 *   erfc-Coulomb pair      LIMA_ENGINE/src/DeviceAlgorithmsPrivate.cuh:38-95
 *   Coulomb constant, kappa LIMA_BASE/include/PhysicsUtils.cuh:5,60-62; Constants.h:43-66
 *   cutoff test            LIMA_ENGINE/src/EngineUtils.cuh:175-180
 *   cutoff pair set        LennardJonesInteractions.cuh:164-203 (far kernel, EngineKernels.cuh:47-146)
 *   intra pair set         LennardJonesInteractions.cuh:121-161, EngineKernels.cuh:150-263
 *   exclusions             LIMA_TOOLS/src/CompoundBuilder.cpp:17-93, LimaTypes.cuh:415-490
 *   bonded                 LIMA_ENGINE/src/ForceComputations.cuh:17-264,335-376
 *   PME                    LIMA_ENGINE/src/PME.cuh:231-333 (spread), 335-430 (gather),
 *                          434-495 (Green's function), 574-608 (solve), 610-630 (self energy);
 *                          B-spline DeviceAlgorithmsPrivate.cuh:7-30
 *   integrator             LIMA_ENGINE/src/EngineUtils.cuh:31-41, EngineKernels.cuh:344-441;
 *                          Coord truncation LIMA_BASE/include/LimaTypes.cuh:297-334
 *   steepest descent       GROMACS integrator=steep (dev/gmxscripts/em.mdp:2-5): x' = x + h F/max|F|,
 *                          accept if E decreases then h*1.2, else h*0.2
 *
 * Generalisations of LIMA (shared with the product): orthorhombic non-integer boxes,
 * a free PME grid (LIMA fixes 10 points/nm, PME.cuh:16), per-particle sigma/epsilon,
 * all periodic images inside the cutoff for boxes shorter than 2*cutoff, and positions
 * kept as unwrapped int64 Coords so intra-molecular displacements never need an image.
 */
#include "lima_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_err;

/* ---- LIMA constants, Constants.h:43-66 (float-typed there; promoted here) ---- */
constexpr float PI_L = 3.14159f;                         // Constants.h:51
constexpr double NANO_D = 1e-9;
constexpr double AVOGADRO = 6.02214076e23;               // Constants.h:60
constexpr double COULOMB_SI = 8.9875517873681764e9;      // Constants.h:61
constexpr double BOLTZMANN = (double)1.38066e-23f;       // Constants.h:59 (float literal)
constexpr float KILO_F = 1000.f;
// PhysicsUtils.cuh:5, a constexpr float
const double KC = (double)(float)(COULOMB_SI / NANO_D / AVOGADRO * KILO_F * KILO_F);
constexpr double NANO_TO_LIMA = 1e8;                     // LimaTypes.cuh:299

struct Vec {
    double x = 0, y = 0, z = 0;
    Vec() = default;
    Vec(double a, double b, double c) : x(a), y(b), z(c) {}
    Vec operator+(const Vec& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vec operator-(const Vec& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vec operator*(double s) const { return {x * s, y * s, z * s}; }
    Vec operator-() const { return {-x, -y, -z}; }
    Vec& operator+=(const Vec& o) { x += o.x; y += o.y; z += o.z; return *this; }
    double dot(const Vec& o) const { return x * o.x + y * o.y + z * o.z; }
    Vec cross(const Vec& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    double len() const { return std::sqrt(dot(*this)); }
    Vec norm() const { double l = len(); return {x / l, y / l, z / l}; }
};

inline int64_t floordiv(int64_t a, int64_t b) { int64_t q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) --q; return q; }
inline int64_t floormod(int64_t a, int64_t b) { return a - floordiv(a, b) * b; }

}  // namespace

struct lima_oracle {
    int nC = 0, nT = 0;
    int64_t N = 0;
    double L[3]{};
    int64_t Lq[3]{};
    std::vector<int64_t> u;             // unwrapped Coord [N*3]
    std::vector<double> v, fprev;       // [N*3]
    std::vector<double> q, sig, eps, mass;
    std::vector<int> comp, slot;        // compound id (tinymols: nC + t), slot
    std::vector<int64_t> first;         // [nC+1] first global index of compound
    std::vector<int> key;
    std::vector<uint32_t> lut;          // [nC*5*32]
    std::vector<uint8_t> bmask;         // [nC] bit k: compound c+k-2 is same/bonded
    struct Term { int a[4]; double p[4]; };
    std::vector<Term> bonds, angles, dihedrals, impropers, pairs;
    // params
    double rc = 1.2, kappa = 2.5, dt_ns = 2e-6, skin = 0.1;
    bool elec = true;
    bool pme_divide_modulus = false;    // test switch: textbook smooth-PME deconvolution
    int K[3]{};
    int64_t step = 0;
    // thermostat (Thermostat.cuh:56-103, Engine.cu:162-174)
    bool thermo_apply = false;
    int64_t thermo_every = 200;       // steps_per_temperature_measurement
    int64_t log_interval = 5;         // data_logging_interval
    float thermo_scale = 1.f;         // DeviceConstants::thermostatScalar, initially 1 (Engine.cu:130-131)
    float temperature = NAN;          // RunStatus::current_temperature
    // last evaluation
    std::vector<double> F, pot;
    lima_energy E{};
    bool have_forces = false;
    // scratch for PME inspection
    std::vector<double> rho_last, phi_last;  // filled only while keep_grids is set
    bool keep_grids = false;
    std::vector<std::complex<double>> pme_grid;   // persistent work grid (reused across calls)
    // cluster list
    int ncol[2]{};

    int64_t gidx(int c, int s) const { return first[c] + s; }
    bool intra(int64_t i, int64_t j) const {
        const int ci = comp[i], cj = comp[j];
        if (ci == cj) return true;
        if (ci >= nC || cj >= nC) return false;
        const int d = cj - ci;
        if (d < -2 || d > 2) return false;
        return (bmask[ci] >> (d + 2)) & 1;
    }
    bool lut_excluded(int ci, int si, int cj, int sj) const {
        const int k = cj - ci + 2;
        return (lut[((size_t)ci * 5 + k) * 32 + si] >> sj) & 1u;
    }
    int64_t wrapped(int64_t i, int d) const { return floormod(u[i * 3 + d], Lq[d]); }
};

namespace {

int g_threads = 0;

/* ---------------------------------------------------------------------------------------
 * Pair interaction, LennardJonesInteractions.cuh:74-108 and DeviceAlgorithmsPrivate.cuh:38-95.
 * diff = x_other - x_self [nm].  Returns force on self [J/mol/nm]; adds the self half of
 * the pair energies (LJ 4 eps s(s-1)/2, Coulomb k qq erfc(kr)/r /2).
 * ------------------------------------------------------------------------------------- */
inline Vec pair_force(const Vec& diff, double sig, double eps, double qq, double kappa, double& elj, double& ec) {
    const double r2 = diff.dot(diff);
    const double inv_r2 = 1.0 / r2;
    double s = sig * sig * inv_r2;
    s = s * s * s;
    const double flj = eps * s * inv_r2 * (1.0 - 2.0 * s);           // [1/24 J/mol/nm^2]
    elj += 4.0 * eps * s * (s - 1.0) * 0.5;
    Vec f = diff * (24.0 * flj);
    if (qq != 0.0) {
        const double r = std::sqrt(r2);
        const double kr = kappa * r;
        const double sc = std::erfc(kr) + 2.0 * kappa / std::sqrt((double)PI_L) * r * std::exp(-kappa * kappa * r2);
        f += (-diff) * (KC * qq / (r2 * r) * sc);
        ec += KC * qq * std::erfc(kr) / r * 0.5;
    }
    return f;
}

/* Symmetric part of scale * r (x) f added to a virial tensor W (xx yy zz xy xz yz).  The virial
 * is W = -dU/d(strain) = sum_i x_i (x) F_i for a finite system; for pair terms evaluated from
 * both sides each ordered pair adds 1/2 (x_i - x_j) (x) F_i = -1/2 d (x) F_i with d = x_j - x_i. */
inline void add_virial(double* W, const Vec& r, const Vec& f, double scale) {
    W[0] += scale * r.x * f.x;
    W[1] += scale * r.y * f.y;
    W[2] += scale * r.z * f.z;
    W[3] += scale * 0.5 * (r.x * f.y + r.y * f.x);
    W[4] += scale * 0.5 * (r.x * f.z + r.z * f.x);
    W[5] += scale * 0.5 * (r.y * f.z + r.z * f.y);
}

/* ---- mixed-radix complex FFT in double (radix 2,3,4,5, generic odd prime by DFT) ---- */
using cd = std::complex<double>;

int nthreads() { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* Decimation in time: out[k + q m] = sum_t W_n^{t (k + q m)} FFT_m(in[t::r])[k], n = r m.
 * W holds W_N^j = exp(sign 2 pi i j / N) for the top-level length N; scr needs 2 n entries. */
void fft1d_rec(const cd* in, size_t stride, cd* out, int n, const cd* W, int N, cd* scr) {
    if (n == 1) { out[0] = in[0]; return; }
    int r = 0;
    for (int p : {4, 2, 3, 5}) if (n % p == 0) { r = p; break; }
    if (!r) { for (int p = 7; p <= n; p += 2) if (n % p == 0) { r = p; break; } }
    const int m = n / r;
    cd* sub = scr;
    for (int t = 0; t < r; t++) fft1d_rec(in + t * stride, stride * r, sub + t * m, m, W, N, scr + n);
    // W_n^{t (k + q m)} = W_n^{t k} W_r^{t q}: pre-twiddle, then an r-point DFT
    const long step = N / n;
    cd Wr[64];
    const bool small = r <= 8;
    if (small) for (int j = 0; j < r * r; j++) Wr[j] = W[(long)((j / r) * (j % r) % r) * (N / r)];
    cd x[8];
    for (int k = 0; k < m; k++) {
        if (small) {
            for (int t = 0; t < r; t++) x[t] = sub[t * m + k] * W[(long)t * k * step];
            for (int q = 0; q < r; q++) {
                cd acc = 0;
                for (int t = 0; t < r; t++) acc += x[t] * Wr[t * r + q];
                out[k + q * m] = acc;
            }
        } else {
            for (int q = 0; q < r; q++) {
                cd acc = 0;
                for (int t = 0; t < r; t++) {
                    const long e = ((long)t * (k + (long)q * m)) % n;
                    acc += sub[t * m + k] * W[e * step];
                }
                out[k + q * m] = acc;
            }
        }
    }
}

std::vector<cd> twiddles(int n, int sign) {
    std::vector<cd> W(n);
    for (int j = 0; j < n; j++) W[j] = std::polar(1.0, sign * 2.0 * M_PI * (double)j / (double)n);
    return W;
}

void fft3d(std::vector<cd>& a, int nx, int ny, int nz, int sign) {
    const int dims[3] = {nx, ny, nz};
    const size_t strides[3] = {1, (size_t)nx, (size_t)nx * ny};
    constexpr int B = 16;   // neighbouring x-lines transformed together (contiguous gathers)
    for (int ax = 0; ax < 3; ax++) {
        const int n = dims[ax];
        const std::vector<cd> W = twiddles(n, sign);
        if (ax == 0) {
#pragma omp parallel num_threads(nthreads())
            {
                std::vector<cd> res(n), scr(2 * (size_t)n);
#pragma omp for schedule(static)
                for (long l = 0; l < (long)ny * nz; l++) {
                    cd* p = &a[(size_t)l * nx];
                    fft1d_rec(p, 1, res.data(), n, W.data(), n, scr.data());
                    std::copy(res.begin(), res.end(), p);
                }
            }
            continue;
        }
        const int o = ax == 1 ? 2 : 1;                  // the other non-x axis
        const int nxb = (nx + B - 1) / B;
        const long nblk = (long)dims[o] * nxb;
#pragma omp parallel num_threads(nthreads())
        {
            std::vector<cd> tile((size_t)n * B), res(n), scr(2 * (size_t)n);
#pragma omp for schedule(static)
            for (long blk = 0; blk < nblk; blk++) {
                const int xb = (int)(blk % nxb) * B, io = (int)(blk / nxb);
                const int bw = std::min(B, nx - xb);
                cd* base = &a[io * strides[o] + xb];
                for (int i = 0; i < n; i++)
                    for (int b = 0; b < bw; b++) tile[(size_t)i * B + b] = base[i * strides[ax] + b];
                for (int b = 0; b < bw; b++) {
                    fft1d_rec(&tile[b], B, res.data(), n, W.data(), n, scr.data());
                    for (int i = 0; i < n; i++) tile[(size_t)i * B + b] = res[i];
                }
                for (int i = 0; i < n; i++)
                    for (int b = 0; b < bw; b++) base[i * strides[ax] + b] = tile[(size_t)i * B + b];
            }
        }
    }
}

/* Cubic B-spline weights at nodes ix-1..ix+2, DeviceAlgorithmsPrivate.cuh:7-13 */
inline void bspline4(double f, double* w) {
    w[0] = (1.0 - f) * (1.0 - f) * (1.0 - f) / 6.0;
    w[1] = (4.0 - 6.0 * f * f + 3.0 * f * f * f) / 6.0;
    w[2] = (1.0 + 3.0 * f + 3.0 * f * f - 3.0 * f * f * f) / 6.0;
    w[3] = (f * f * f) / 6.0;
}

bool smooth235(int n) { for (int p : {2, 3, 5}) while (n % p == 0) n /= p; return n == 1; }

int derive_grid(double L, double spacing) {
    int n = (int)std::ceil(L / spacing - 1e-9);
    if (n < 4) n = 4;
    while (!smooth235(n)) n++;
    return n;
}

// Displacements between the particles of one compound are unwrapped (LIMA keeps a compound's
// particles relative to one origo).  Between compounds LIMA moves the other compound's origo to
// its nearest image first (BoundaryCondition::applyHyperpos_Return, EngineKernels.cuh:235-236);
// here the compound's key particle stands for its origo.  Bonded terms join nearby particles, so
// the nearest image of each displacement is the hyperpos result.
inline int64_t min_image(int64_t d, int64_t Lq) {   // representative of d mod Lq in [-Lq/2, Lq/2]
    int64_t r = d % Lq;
    if (r < 0) r += Lq;
    if (2 * r > Lq) r -= Lq;
    return r;
}


Vec bdisp(const lima_oracle* o, int64_t a, int64_t b) {  // nearest image of x_b - x_a [nm]
    double d[3];
    for (int k = 0; k < 3; k++) d[k] = (double)min_image(o->u[b * 3 + k] - o->u[a * 3 + k], o->Lq[k]);
    return Vec(d[0], d[1], d[2]) * (1.0 / NANO_TO_LIMA);
}

/* ---------------------------------------------------------------------------------------
 * Nonbonded, cutoff pair set: every particle pair that is not intra (same or bonded
 * compound), summed over all periodic images with r^2 <= rc^2 (EngineUtils.cuh:175-180
 * includes r == rc).  LIMA evaluates each ordered pair from the self side
 * (EngineKernels.cuh:96-139); the oracle does the same: force on i, half energy on i.
 * ------------------------------------------------------------------------------------- */
void nb_cutoff(lima_oracle* o, std::vector<double>& F, std::vector<double>& pot, double& elj, double& ec, double* W) {
    const int64_t N = o->N;
    int nc[3];
    for (int d = 0; d < 3; d++) { nc[d] = std::max(1, (int)std::floor(o->L[d] / o->rc)); }
    const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    std::vector<int64_t> cell(N), head(ncell + 1, 0), order(N);
    for (int64_t i = 0; i < N; i++) {
        int64_t c[3];
        for (int d = 0; d < 3; d++) c[d] = std::min<int64_t>(nc[d] - 1, (int64_t)((__int128)o->wrapped(i, d) * nc[d] / o->Lq[d]));
        cell[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
        head[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) head[c + 1] += head[c];
    {
        std::vector<int64_t> fill(head.begin(), head.end() - 1);
        for (int64_t i = 0; i < N; i++) order[fill[cell[i]]++] = i;
    }
    const double rc2 = o->rc * o->rc;
    double elj_t = 0, ec_t = 0, w_t[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : elj_t, ec_t, w_t[:6]) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int64_t i = 0; i < N; i++) {
        int64_t ci[3] = {cell[i] % nc[0], (cell[i] / nc[0]) % nc[1], cell[i] / ((int64_t)nc[0] * nc[1])};
        // candidate (cell index, shift) per dim
        std::vector<std::pair<int, int>> cand[3];
        for (int d = 0; d < 3; d++) {
            if (nc[d] >= 3) {
                for (int off = -1; off <= 1; off++) {
                    int64_t c = ci[d] + off; int s = 0;
                    if (c < 0) { c += nc[d]; s = -1; } else if (c >= nc[d]) { c -= nc[d]; s = 1; }
                    cand[d].push_back({(int)c, s});
                }
            } else {
                for (int c = 0; c < nc[d]; c++) for (int s = -1; s <= 1; s++) cand[d].push_back({c, s});
            }
        }
        Vec fi;
        double pe = 0, el = 0, e2 = 0;
        const int64_t wi[3] = {o->wrapped(i, 0), o->wrapped(i, 1), o->wrapped(i, 2)};
        for (auto& cz : cand[2]) for (auto& cy : cand[1]) for (auto& cx : cand[0]) {
            const int64_t cc = ((int64_t)cz.first * nc[1] + cy.first) * nc[0] + cx.first;
            for (int64_t k = head[cc]; k < head[cc + 1]; k++) {
                const int64_t j = order[k];
                if (j == i || o->intra(i, j)) continue;
                const int64_t dq[3] = {o->wrapped(j, 0) + cx.second * o->Lq[0] - wi[0],
                                       o->wrapped(j, 1) + cy.second * o->Lq[1] - wi[1],
                                       o->wrapped(j, 2) + cz.second * o->Lq[2] - wi[2]};
                const Vec d((double)dq[0] / NANO_TO_LIMA, (double)dq[1] / NANO_TO_LIMA, (double)dq[2] / NANO_TO_LIMA);
                if (d.dot(d) > rc2) continue;
                const double sg = (o->sig[i] + o->sig[j]) * 0.5, ep = std::sqrt(o->eps[i] * o->eps[j]);
                double a = 0, b = 0;
                const Vec f = pair_force(d, sg, ep, o->q[i] * o->q[j], o->kappa, a, b);
                fi += f;
                add_virial(w_t, d, f, -0.5);
                el += a; e2 += b;
            }
        }
        pe = el + e2;
        F[i * 3 + 0] += fi.x; F[i * 3 + 1] += fi.y; F[i * 3 + 2] += fi.z;
        pot[i] += pe;
        elj_t += el; ec_t += e2;
    }
    elj += elj_t; ec += ec_t;
    for (int k = 0; k < 6; k++) W[k] += w_t[k];
}

/* Intra / bonded-compound pairs, no cutoff, LUT exclusions, compound-frame displacement.
 * EngineKernels.cuh:197-263 with LennardJonesInteractions.cuh:121-161. */
void nb_intra(lima_oracle* o, std::vector<double>& F, std::vector<double>& pot, double& elj, double& ec, double* W) {
    double elj_t = 0, ec_t = 0, w_t[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : elj_t, ec_t, w_t[:6]) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int A = 0; A < o->nC; A++) {
        const int nA = (int)(o->first[A + 1] - o->first[A]);
        for (int k = 0; k < 5; k++) {
            if (!((o->bmask[A] >> k) & 1)) continue;
            const int B = A + k - 2;
            const int nB = (int)(o->first[B + 1] - o->first[B]);
            int64_t sh[3] = {0, 0, 0};   // image shift of compound B: its key particle nearest to A's
            if (B != A) {
                const int64_t ka = o->gidx(A, o->key[A]), kb = o->gidx(B, o->key[B]);
                for (int d = 0; d < 3; d++) {
                    const int64_t dk = o->u[kb * 3 + d] - o->u[ka * 3 + d];
                    sh[d] = min_image(dk, o->Lq[d]) - dk;
                }
            }
            for (int si = 0; si < nA; si++) {
                const int64_t i = o->gidx(A, si);
                Vec fi; double el = 0, e2 = 0;
                for (int sj = 0; sj < nB; sj++) {
                    if (o->lut_excluded(A, si, B, sj)) continue;
                    const int64_t j = o->gidx(B, sj);
                    if (j == i) continue;
                    const Vec d((double)(o->u[j * 3 + 0] - o->u[i * 3 + 0] + sh[0]) / NANO_TO_LIMA,
                                (double)(o->u[j * 3 + 1] - o->u[i * 3 + 1] + sh[1]) / NANO_TO_LIMA,
                                (double)(o->u[j * 3 + 2] - o->u[i * 3 + 2] + sh[2]) / NANO_TO_LIMA);
                    const double sg = (o->sig[i] + o->sig[j]) * 0.5, ep = std::sqrt(o->eps[i] * o->eps[j]);
                    const Vec f = pair_force(d, sg, ep, o->q[i] * o->q[j], o->kappa, el, e2);
                    fi += f;
                    add_virial(w_t, d, f, -0.5);
                }
                F[i * 3 + 0] += fi.x; F[i * 3 + 1] += fi.y; F[i * 3 + 2] += fi.z;
                pot[i] += el + e2;
                elj_t += el; ec_t += e2;
            }
        }
    }
    elj += elj_t; ec += ec_t;
    for (int k = 0; k < 6; k++) W[k] += w_t[k];
}

inline void addF(std::vector<double>& F, int64_t i, const Vec& f) { F[i * 3 + 0] += f.x; F[i * 3 + 1] += f.y; F[i * 3 + 2] += f.z; }

/* Bonded terms, ForceComputations.cuh.  Energies are split over the term's particles
 * (ForceComputations.cuh:317-325). */
double bonded(lima_oracle* o, std::vector<double>& F, std::vector<double>& pot, double* W) {
    // each term adds sum_a (x_a - x_a0) (x) F_a, positions taken relative to its first particle
    double etot = 0;
    // bonds, calcSinglebondForces :17-51 (MD form; SD uses the exact gradient)
    for (auto& t : o->bonds) {
        const Vec dif = bdisp(o, t.a[1], t.a[0]);   // pos_a - pos_b
        const double r = dif.len(), err = r - t.p[0];
        const double e = 0.5 * t.p[1] * err * err;
        const Vec f = dif * (-t.p[1] * err / r);
        addF(F, t.a[0], f); addF(F, t.a[1], -f);
        add_virial(W, -dif, -f, 1.0);
        pot[t.a[0]] += 0.5 * e; pot[t.a[1]] += 0.5 * e; etot += e;
    }
    // angles + Urey-Bradley, calcAnglebondForces :53-115 (harmonic in theta; exact gradient)
    for (auto& t : o->angles) {
        const Vec r1 = bdisp(o, t.a[1], t.a[0]), r2 = bdisp(o, t.a[1], t.a[2]);
        const double l1 = r1.len(), l2 = r2.len();
        const Vec v1 = r1 * (1.0 / l1), v2 = r2 * (1.0 / l2);
        double c = v1.dot(v2); c = std::min(1.0, std::max(-1.0, c));
        const double th = std::acos(c), err = th - t.p[0];
        double e = 0.5 * t.p[1] * err * err;
        const double s = std::sqrt(std::max(1e-300, 1.0 - c * c));
        // dtheta/dx_left = -(v2 - c v1)/(l1 s)
        const Vec dl = (v2 - v1 * c) * (-1.0 / (l1 * s));
        const Vec dr = (v1 - v2 * c) * (-1.0 / (l2 * s));
        Vec fl = dl * (-t.p[1] * err), fr = dr * (-t.p[1] * err);
        if (t.p[3] != 0.0) {
            const Vec d13 = bdisp(o, t.a[2], t.a[0]);   // left - right
            const double r13 = d13.len(), eu = r13 - t.p[2];
            e += 0.5 * t.p[3] * eu * eu;
            const Vec fu = d13 * (-t.p[3] * eu / r13);
            fl += fu; fr += -fu;
        }
        addF(F, t.a[0], fl); addF(F, t.a[2], fr); addF(F, t.a[1], -(fl + fr));
        add_virial(W, bdisp(o, t.a[0], t.a[1]), -(fl + fr), 1.0);
        add_virial(W, bdisp(o, t.a[0], t.a[2]), fr, 1.0);
        for (int k = 0; k < 3; k++) pot[t.a[k]] += e / 3.0;
        etot += e;
    }
    // proper dihedrals, calcDihedralbondForces :119-210.  phi_L = -atan2(sin, cos) with
    // A = r12 x r23, B = r23 x r34 (LIMA's sign convention), V = k(1 + cos(n phi_L - phi0)).
    for (auto& t : o->dihedrals) {
        const Vec b1 = bdisp(o, t.a[0], t.a[1]), b2 = bdisp(o, t.a[1], t.a[2]), b3 = bdisp(o, t.a[2], t.a[3]);
        const Vec m = b1.cross(b2), n = b2.cross(b3);
        const double lb2 = b2.len();
        const double phi_bk = std::atan2(lb2 * b1.dot(n), m.dot(n));   // Blondel-Karplus angle
        const double phi = -phi_bk;                                    // LIMA: -atan2(sin, cos)
        const double e = t.p[1] * (1.0 + std::cos(t.p[2] * phi - t.p[0]));
        const double dVdphi = -t.p[1] * t.p[2] * std::sin(t.p[2] * phi - t.p[0]);
        // dphi_bk/dx (Blondel & Karplus 1996)
        const Vec g1 = m * (-lb2 / m.dot(m));
        const Vec g4 = n * (lb2 / n.dot(n));
        const double p = b1.dot(b2) / (lb2 * lb2), qd = b3.dot(b2) / (lb2 * lb2);
        const Vec g2 = g1 * (-(p + 1.0)) + g4 * qd;
        const Vec g3 = g1 * p - g4 * (qd + 1.0);
        // F = -dV/dphi_L * dphi_L/dx = +dV/dphi_L * dphi_bk/dx
        addF(F, t.a[0], g1 * dVdphi); addF(F, t.a[1], g2 * dVdphi);
        addF(F, t.a[2], g3 * dVdphi); addF(F, t.a[3], g4 * dVdphi);
        add_virial(W, b1, g2 * dVdphi, 1.0);                   // x_1 - x_0 = b1
        add_virial(W, b1 + b2, g3 * dVdphi, 1.0);
        add_virial(W, b1 + b2 + b3, g4 * dVdphi, 1.0);
        for (int k = 0; k < 4; k++) pot[t.a[k]] += e / 4.0;
        etot += e;
    }
    // improper dihedrals, calcImproperdihedralbondForces :214-264 (LIMA's force split, not a gradient)
    for (auto& t : o->impropers) {
        const Vec pi(0, 0, 0), pj = bdisp(o, t.a[0], t.a[1]), pk = bdisp(o, t.a[0], t.a[2]), pl = bdisp(o, t.a[0], t.a[3]);
        const Vec ij = (pj - pi).norm(), ik = (pk - pi).norm(), il = (pl - pi).norm();
        const Vec lj = (pj - pl).norm(), lk = (pk - pl).norm();
        const Vec n1 = ij.cross(ik).norm(), n2 = lj.cross(lk).norm();
        double c = n1.dot(n2); c = std::min(1.0, std::max(-1.0, c));
        double ang = std::acos(c);
        if (n1.dot(il) > 0.0) ang = -ang;
        const double err = ang - t.p[0];
        const double e = 0.5 * t.p[1] * err * err, torque = t.p[1] * err;
        auto dist_to_line = [](const Vec& p, const Vec& a, const Vec& b) { return ((b - a).cross(a - p)).len() / (b - a).len(); };
        const Vec fl = n2 * (torque / dist_to_line(pl, pj, pk));
        const Vec fi = -n1 * (torque / dist_to_line(pi, pj, pk));
        const Vec res = -(fi + fl);
        const double ratio = (pj - pi).len() / ((pj - pi).len() + (pk - pi).len());
        addF(F, t.a[0], fi); addF(F, t.a[1], res * (1.0 - ratio)); addF(F, t.a[2], res * ratio); addF(F, t.a[3], fl);
        add_virial(W, pj, res * (1.0 - ratio), 1.0);
        add_virial(W, pk, res * ratio, 1.0);
        add_virial(W, pl, fl, 1.0);
        for (int k = 0; k < 4; k++) pot[t.a[k]] += e / 4.0;
        etot += e;
    }
    // 1-4 pair LJ, computePairbondForces :335-376
    for (auto& t : o->pairs) {
        const Vec d = bdisp(o, t.a[0], t.a[1]);   // other - self (self = a0)
        double el = 0, ec = 0;
        const Vec f = pair_force(d, t.p[0], t.p[1], 0.0, o->kappa, el, ec);
        addF(F, t.a[0], f); addF(F, t.a[1], -f);
        add_virial(W, d, -f, 1.0);
        pot[t.a[0]] += el; pot[t.a[1]] += el; etot += 2 * el;
    }
    return etot;
}

/* ---------------------------------------------------------------------------------------
 * PME, LIMA structure (PME.cuh).  Spread q*w/V_cell with centred cubic B-splines at nodes
 * ix-1..ix+2 (PME.cuh:263-300), forward FFT, multiply by
 * G = kC 4 pi/k^2 exp(-k^2/4kappa^2) (prod_d sinc^4(k_d h_d/2))^2  (PME.cuh:434-495, spline
 * factor multiplied as in LIMA), inverse FFT / N (PME.cuh:585-602), gather phi and
 * E = -grad phi by central differences (PME.cuh:374-411), force q E, potE 1/2 (q phi + Eself)
 * with Eself = -kappa/sqrt(pi) kC sum_all q^2 on every charged particle (PME.cuh:414-420,610-630).
 * ------------------------------------------------------------------------------------- */
void pme(lima_oracle* o, std::vector<double>& F, std::vector<double>& pot, lima_energy& E) {
    const int Kx = o->K[0], Ky = o->K[1], Kz = o->K[2];
    const size_t NG = (size_t)Kx * Ky * Kz;
    double h[3];
    for (int d = 0; d < 3; d++) h[d] = o->L[d] / o->K[d];
    const double inv_vcell = 1.0 / (h[0] * h[1] * h[2]);
    std::vector<cd>& grid = o->pme_grid;
    grid.resize(NG);
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (long g = 0; g < (long)NG; g++) grid[g] = 0.0;
    struct Sten { int ix[3]; double w[3][4]; };
    std::vector<Sten> st(o->N);
    double qsq = 0;
    for (int64_t i = 0; i < o->N; i++) {
        qsq += o->q[i] * o->q[i];
        if (o->q[i] == 0.0) continue;
        Sten& s = st[i];
        for (int d = 0; d < 3; d++) {
            // grid coordinate of the wrapped position, exact from the integer Coord
            const int64_t w = o->wrapped(i, d);
            const __int128 num = (__int128)w * o->K[d];
            const int64_t ix = (int64_t)(num / o->Lq[d]);
            const double f = (double)(int64_t)(num - (__int128)ix * o->Lq[d]) / (double)o->Lq[d];
            s.ix[d] = (int)ix;
            bspline4(f, s.w[d]);
        }
        for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) for (int c = 0; c < 4; c++) {
            const int X = (int)floormod(s.ix[0] - 1 + a, Kx), Y = (int)floormod(s.ix[1] - 1 + b, Ky), Z = (int)floormod(s.ix[2] - 1 + c, Kz);
            grid[((size_t)Z * Ky + Y) * Kx + X] += o->q[i] * s.w[0][a] * s.w[1][b] * s.w[2][c] * inv_vcell;
        }
    }
    if (o->keep_grids) {
        o->rho_last.resize(NG);
        for (size_t g = 0; g < NG; g++) o->rho_last[g] = grid[g].real();
    }
    fft3d(grid, Kx, Ky, Kz, -1);
    const double twopi = 2.0 * (double)PI_L;
    const double fourpi = 4.0 * (double)PI_L;
    // per-axis wave numbers and spline factors; G is separable apart from exp(-k^2 / 4 kappa^2)
    std::vector<double> kk[3], kv[3], sf[3];
    for (int d = 0; d < 3; d++) {
        kk[d].resize(o->K[d]);
        kv[d].resize(o->K[d]);
        sf[d].resize(o->K[d]);
        const int half = o->K[d] / 2;
        for (int i = 0; i < o->K[d]; i++) {
            const int m = i <= half ? i : i - o->K[d];
            const double k = twopi * m / o->L[d];
            const double kh = k * h[d] * 0.5;
            kk[d][i] = k * k;
            kv[d][i] = k;
            sf[d][i] = (std::fabs(kh) < 1e-14) ? 1.0 : std::pow(std::sin(kh) / kh, 4);
        }
    }
    // Reciprocal-space virial.  E_rec = V_cell / (2 NG) sum_m G(m) |rho^(m)|^2 (the gather's
    // 1/2 sum q phi by Parseval), with |rho^|^2 V_cell^2 fixed under a strain of the box (the spline
    // weights depend on fractional coordinates only, and sinc(k h / 2) = sinc(pi m / K) does not
    // depend on L), so per wave vector W_ab = E_m (delta_ab - 2 (1/k^2 + 1/(4 kappa^2)) k_a k_b).
    const double vfac = 1.0 / (inv_vcell * 2.0 * (double)NG), inv4k2 = 1.0 / (4.0 * o->kappa * o->kappa);
    double wr[6] = {0, 0, 0, 0, 0, 0};
#pragma omp parallel for schedule(static) reduction(+ : wr[:6]) num_threads(nthreads())
    for (int z = 0; z < Kz; z++) for (int y = 0; y < Ky; y++) for (int x = 0; x < Kx; x++) {
        const double k2 = kk[0][x] + kk[1][y] + kk[2][z];
        const double S = sf[0][x] * sf[1][y] * sf[2][z];
        double G = 0.0;
        // LIMA multiplies the Green's function by the spline transform (sinc^4 per axis, squared:
        // spread and gather), where smooth PME divides by it.  The product keeps LIMA's form; the
        // divided form exists so tests can compare the rest of the pipeline with a direct Ewald sum.
        if (k2 > 1e-14) {
            G = fourpi / k2 * std::exp(-k2 / (4.0 * o->kappa * o->kappa)) * KC;
            G = o->pme_divide_modulus ? G / (S * S) : G * S * S;
            const double Em = vfac * G * std::norm(grid[((size_t)z * Ky + y) * Kx + x]);
            const double c = 2.0 * (1.0 / k2 + inv4k2);
            const double ka[3] = {kv[0][x], kv[1][y], kv[2][z]};
            wr[0] += Em * (1.0 - c * ka[0] * ka[0]);
            wr[1] += Em * (1.0 - c * ka[1] * ka[1]);
            wr[2] += Em * (1.0 - c * ka[2] * ka[2]);
            wr[3] -= Em * c * ka[0] * ka[1];
            wr[4] -= Em * c * ka[0] * ka[2];
            wr[5] -= Em * c * ka[1] * ka[2];
        }
        grid[((size_t)z * Ky + y) * Kx + x] *= G;
    }
    fft3d(grid, Kx, Ky, Kz, +1);
    // phi = Re(IFFT) / NG, kept in place in the real parts of the work grid
    const double inv_ng = 1.0 / (double)NG;
    if (o->keep_grids) {
        o->phi_last.resize(NG);
        for (size_t g = 0; g < NG; g++) o->phi_last[g] = grid[g].real() * inv_ng;
    }
    const double eself = -o->kappa / std::sqrt((double)PI_L) * qsq * KC;
    double erec = 0;
    int64_t ncharged = 0;
    for (int64_t i = 0; i < o->N; i++) {
        if (o->q[i] == 0.0) continue;
        ncharged++;
        const Sten& s = st[i];
        Vec Ef; double ph = 0;
        for (int a = 0; a < 4; a++) for (int b = 0; b < 4; b++) for (int c = 0; c < 4; c++) {
            const int X = (int)floormod(s.ix[0] - 1 + a, Kx), Y = (int)floormod(s.ix[1] - 1 + b, Ky), Z = (int)floormod(s.ix[2] - 1 + c, Kz);
            auto at = [&](int xx, int yy, int zz) {
                return grid[((size_t)floormod(zz, Kz) * Ky + floormod(yy, Ky)) * Kx + floormod(xx, Kx)].real() * inv_ng;
            };
            const double w = s.w[0][a] * s.w[1][b] * s.w[2][c];
            ph += w * at(X, Y, Z);
            Ef += Vec(-(at(X + 1, Y, Z) - at(X - 1, Y, Z)) / (2 * h[0]), -(at(X, Y + 1, Z) - at(X, Y - 1, Z)) / (2 * h[1]),
                      -(at(X, Y, Z + 1) - at(X, Y, Z - 1)) / (2 * h[2])) * w;
        }
        addF(F, i, Ef * o->q[i]);
        pot[i] += 0.5 * (o->q[i] * ph + eself);
        erec += 0.5 * o->q[i] * ph;
    }
    E.pme_recip = erec;
    for (int k = 0; k < 6; k++) E.virial[k] += wr[k];
    E.pme_self_lima = 0.5 * eself * (double)ncharged;
    E.pme_self_ewald = eself;
}

void evaluate(lima_oracle* o, int comps, std::vector<double>& F, std::vector<double>& pot, lima_energy& E) {
    F.assign(o->N * 3, 0.0);
    pot.assign(o->N, 0.0);
    E = lima_energy{};
    double elj = 0, ec = 0;
    if (comps & LIMA_ORC_NB_CUTOFF) nb_cutoff(o, F, pot, elj, ec, E.virial);
    if (comps & LIMA_ORC_NB_INTRA) nb_intra(o, F, pot, elj, ec, E.virial);
    E.lj = elj; E.coulomb_sr = ec;
    if (comps & LIMA_ORC_BONDED) E.bonded = bonded(o, F, pot, E.virial);
    if ((comps & LIMA_ORC_PME) && o->elec) pme(o, F, pot, E);
    E.potential = E.lj + E.coulomb_sr + E.bonded + E.pme_recip + E.pme_self_ewald;
    E.potential_lima = E.lj + E.coulomb_sr + E.bonded + E.pme_recip + E.pme_self_lima;
    double ke = 0;
    for (int64_t i = 0; i < o->N; i++) {
        const double v2 = o->v[i * 3] * o->v[i * 3] + o->v[i * 3 + 1] * o->v[i * 3 + 1] + o->v[i * 3 + 2] * o->v[i * 3 + 2];
        ke += 0.5 * o->mass[i] * v2;
    }
    E.kinetic = ke;
    E.total = E.potential + ke;
}

/* ---------------------------------------------------------------------------------------
 * Cluster-pair neighbour list.  This is the list the product builds; the oracle defines it:
 *  - x-y columns: n_d = max(1, floor(L_d / a)), a = cbrt(8 V / N); column of a particle
 *    floor(w_d n_d / Lq_d) from its wrapped integer Coord w.
 *  - within a column particles are ordered by (w_z, global index) and cut into clusters
 *    of 8 consecutive particles; clusters are numbered column by column (x fastest).
 *  - cluster bounding box: integer min/max of the members' wrapped Coords.
 *  - cluster I lists every particle b with image shift s in {-1,0,1}^3 whose integer squared
 *    distance from w_b + s Lq to I's box is <= round((rc+skin) 1e8)^2 (a full list: the pair
 *    kernel computes each pair from both sides, force on the i side only, as LIMA does).
 *  - the entry's 8-bit mask has bit i set for real slot i of I whose pair with b is not
 *    intra (same / bonded compound) and is not b itself (s == 0).  Entries with an empty mask
 *    are dropped.
 *  shift code = (sx+1) + 3 (sy+1) + 9 (sz+1); entry = global index of b | code << 24.
 * ------------------------------------------------------------------------------------- */
struct ClusterList {
    std::vector<int32_t> atoms;          // [nclust*8], -1 padding
    std::vector<int64_t> off;            // [nclust+1]
    std::vector<uint32_t> ent;
    std::vector<uint8_t> mask;
};

void column_dims(const lima_oracle* o, int* ncx, int* ncy) {
    const double V = o->L[0] * o->L[1] * o->L[2];
    const double a = std::cbrt(8.0 * V / (double)std::max<int64_t>(1, o->N));
    *ncx = std::max(1, (int)std::floor(o->L[0] / a));
    *ncy = std::max(1, (int)std::floor(o->L[1] / a));
}

ClusterList build_clusters(const lima_oracle* o) {
    ClusterList cl;
    int ncx, ncy;
    column_dims(o, &ncx, &ncy);
    const int64_t ncol = (int64_t)ncx * ncy;
    std::vector<std::vector<std::pair<int64_t, int64_t>>> cols(ncol);
    std::vector<int64_t> W(o->N * 3);
    for (int64_t i = 0; i < o->N; i++) for (int d = 0; d < 3; d++) W[i * 3 + d] = o->wrapped(i, d);
    for (int64_t i = 0; i < o->N; i++) {
        const int64_t cx = (int64_t)((__int128)W[i * 3] * ncx / o->Lq[0]);
        const int64_t cy = (int64_t)((__int128)W[i * 3 + 1] * ncy / o->Lq[1]);
        cols[cy * ncx + cx].push_back({W[i * 3 + 2], i});
    }
    std::vector<int64_t> clfirst(ncol + 1, 0);
    for (int64_t c = 0; c < ncol; c++) {
        std::sort(cols[c].begin(), cols[c].end());
        clfirst[c + 1] = clfirst[c] + ((int64_t)cols[c].size() + 7) / 8;
    }
    const int64_t ncl = clfirst[ncol];
    cl.atoms.assign(ncl * 8, -1);
    std::vector<int64_t> clof(o->N), slotof(o->N);
    for (int64_t c = 0; c < ncol; c++)
        for (size_t k = 0; k < cols[c].size(); k++) {
            const int64_t I = clfirst[c] + (int64_t)k / 8;
            cl.atoms[I * 8 + k % 8] = (int32_t)cols[c][k].second;
            clof[cols[c][k].second] = I;
            slotof[cols[c][k].second] = (int64_t)k % 8;
        }
    std::vector<int64_t> bmin(ncl * 3), bmax(ncl * 3);
    for (int64_t I = 0; I < ncl; I++) {
        for (int d = 0; d < 3; d++) { bmin[I * 3 + d] = INT64_MAX; bmax[I * 3 + d] = INT64_MIN; }
        for (int k = 0; k < 8; k++) {
            const int32_t a = cl.atoms[I * 8 + k];
            if (a < 0) continue;
            for (int d = 0; d < 3; d++) { bmin[I * 3 + d] = std::min(bmin[I * 3 + d], W[a * 3 + d]); bmax[I * 3 + d] = std::max(bmax[I * 3 + d], W[a * 3 + d]); }
        }
    }
    const int64_t R = llround((o->rc + o->skin) * NANO_TO_LIMA);
    const int64_t R2 = R * R;
    std::vector<std::vector<std::pair<uint32_t, uint8_t>>> lists(ncl);
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int64_t I = 0; I < ncl; I++) {
        auto& lst = lists[I];
        for (int64_t b = 0; b < o->N; b++) {
            const int64_t J = clof[b];
            for (int code = 0; code < 27; code++) {
                const int s[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
                // gaps beyond R end the test early: the squares of three gaps of up to two box
                // lengths would exceed int64 in boxes wider than ~9 nm
                int64_t d2 = 0;
                bool far = false;
                for (int d = 0; d < 3 && !far; d++) {
                    const int64_t p = W[b * 3 + d] + s[d] * o->Lq[d];
                    int64_t g = 0;
                    if (p < bmin[I * 3 + d]) g = bmin[I * 3 + d] - p;
                    else if (p > bmax[I * 3 + d]) g = p - bmax[I * 3 + d];
                    if (g > R) far = true;
                    else d2 += g * g;
                }
                if (far || d2 > R2) continue;
                uint8_t m = 0;
                for (int i = 0; i < 8; i++) {
                    const int32_t a = cl.atoms[I * 8 + i];
                    if (a < 0) continue;
                    if (o->intra(a, b)) continue;
                    if (J == I && code == 13 && i == slotof[b]) continue;
                    m |= (uint8_t)(1u << i);
                }
                if (m) lst.push_back({(uint32_t)b | ((uint32_t)code << 24), m});
            }
        }
    }
    cl.off.assign(ncl + 1, 0);
    for (int64_t I = 0; I < ncl; I++) cl.off[I + 1] = cl.off[I] + (int64_t)lists[I].size();
    cl.ent.resize(cl.off[ncl]);
    cl.mask.resize(cl.off[ncl]);
    for (int64_t I = 0; I < ncl; I++)
        for (size_t k = 0; k < lists[I].size(); k++) { cl.ent[cl.off[I] + k] = lists[I][k].first; cl.mask[cl.off[I] + k] = lists[I][k].second; }
    return cl;
}

}  // namespace

/* =========================================== C ABI =========================================== */
extern "C" {

const char* lima_oracle_last_error(void) { return g_err.c_str(); }
void lima_oracle_set_threads(int n) { g_threads = n; }

int lima_oracle_create(const lima_box_desc* b, const lima_simparams* p, lima_oracle** out) {
    try {
        if (!b || !p || !out) throw std::runtime_error("null argument");
        auto* o = new lima_oracle();
        o->nC = b->n_compounds; o->nT = b->n_tinymols;
        o->first.assign(o->nC + 1, 0);
        for (int c = 0; c < o->nC; c++) {
            const int n = b->compound_n_particles[c];
            if (n < 0 || n > 32) throw std::runtime_error("compound particle count out of range");
            o->first[c + 1] = o->first[c] + n;
        }
        o->N = o->first[o->nC] + o->nT;
        for (int d = 0; d < 3; d++) { o->L[d] = b->box_nm[d]; o->Lq[d] = llround(b->box_nm[d] * NANO_TO_LIMA); o->L[d] = (double)o->Lq[d] / NANO_TO_LIMA; }
        const int64_t N = o->N;
        o->u.resize(N * 3); o->v.assign(N * 3, 0.0); o->fprev.assign(N * 3, 0.0);
        o->q.resize(N); o->sig.resize(N); o->eps.resize(N); o->mass.resize(N); o->comp.resize(N); o->slot.resize(N);
        o->key.resize(o->nC);
        for (int c = 0; c < o->nC; c++) {
            o->key[c] = b->compound_key_particle ? b->compound_key_particle[c] : 0;
            for (int s = 0; s < b->compound_n_particles[c]; s++) {
                const int64_t i = o->gidx(c, s);
                const size_t ps = (size_t)c * 32 + s;
                for (int d = 0; d < 3; d++) {
                    o->u[i * 3 + d] = (int64_t)b->compound_origo[c * 3 + d] * (int64_t)LIMA_NANO_TO_LIMA + b->particle_coord[ps * 3 + d];
                    o->v[i * 3 + d] = b->particle_velocity ? b->particle_velocity[ps * 3 + d] : 0.0;
                }
                o->q[i] = p->enable_electrostatics ? b->particle_charge[ps] : 0.0;
                o->sig[i] = b->particle_sigma[ps]; o->eps[i] = b->particle_epsilon[ps]; o->mass[i] = b->particle_mass[ps];
                o->comp[i] = c; o->slot[i] = s;
            }
        }
        for (int t = 0; t < o->nT; t++) {
            const int64_t i = o->first[o->nC] + t;
            for (int d = 0; d < 3; d++) {
                o->u[i * 3 + d] = (int64_t)b->tinymol_block[t * 3 + d] * (int64_t)LIMA_NANO_TO_LIMA + b->tinymol_coord[t * 3 + d];
                o->v[i * 3 + d] = b->tinymol_velocity ? b->tinymol_velocity[t * 3 + d] : 0.0;
            }
            o->q[i] = 0.0; o->sig[i] = b->tinymol_sigma[t]; o->eps[i] = b->tinymol_epsilon[t]; o->mass[i] = b->tinymol_mass[t];
            o->comp[i] = o->nC + t; o->slot[i] = 0;
        }
        o->lut.assign((size_t)o->nC * 5 * 32, 0u);
        if (b->bonded_lut) std::memcpy(o->lut.data(), b->bonded_lut, o->lut.size() * 4);
        o->bmask.assign(o->nC, 0);
        for (int c = 0; c < o->nC; c++) {
            o->bmask[c] = 1u << 2;
            for (int k = 0; k < 5; k++) {
                const int B = c + k - 2;
                if (k == 2 || B < 0 || B >= o->nC) continue;
                bool any = false;
                for (int s = 0; s < 32; s++) any |= o->lut[((size_t)c * 5 + k) * 32 + s] != 0;
                if (any) o->bmask[c] |= (uint8_t)(1u << k);
            }
        }
        auto conv = [&](int32_t id) -> int {
            const int c = id / 32, s = id % 32;
            if (c < 0 || c >= o->nC || s >= b->compound_n_particles[c]) throw std::runtime_error("bonded particle id out of range");
            return (int)o->gidx(c, s);
        };
        auto load = [&](int n, const int32_t* ids, const float* prm, int na, int np, std::vector<lima_oracle::Term>& dst) {
            for (int k = 0; k < n; k++) {
                lima_oracle::Term t{};
                for (int a = 0; a < na; a++) t.a[a] = conv(ids[k * na + a]);
                for (int a = 0; a < np; a++) t.p[a] = prm[k * np + a];
                dst.push_back(t);
            }
        };
        load(b->n_bonds, b->bond_ids, b->bond_params, 2, 2, o->bonds);
        load(b->n_angles, b->angle_ids, b->angle_params, 3, 4, o->angles);
        load(b->n_dihedrals, b->dihedral_ids, b->dihedral_params, 4, 3, o->dihedrals);
        load(b->n_impropers, b->improper_ids, b->improper_params, 4, 2, o->impropers);
        load(b->n_pairs, b->pair_ids, b->pair_params, 2, 2, o->pairs);
        o->rc = (double)p->cutoff_nm;
        o->kappa = (double)(3.f / p->cutoff_nm);   // PhysicsUtils::CalcEwaldkappa, float
        o->dt_ns = (double)(p->dt_fs * 1e-6f);
        o->skin = p->nlist_skin_nm > 0 ? p->nlist_skin_nm : 0.1;
        o->elec = p->enable_electrostatics != 0;
        o->thermo_apply = p->apply_thermostat != 0;
        o->thermo_every = p->steps_per_temperature_measurement;
        o->log_interval = p->data_logging_interval;
        const double spacing = p->pme_spacing_nm > 0 ? p->pme_spacing_nm : 0.12;
        for (int d = 0; d < 3; d++) o->K[d] = p->pme_grid[d] > 0 ? p->pme_grid[d] : derive_grid(o->L[d], spacing);
        for (int d = 0; d < 3; d++) if (o->L[d] <= o->rc + o->skin) throw std::runtime_error("box edge must exceed cutoff + skin");
        *out = o;
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return LIMA_ERR_BAD_INPUT;
    }
}

void lima_oracle_destroy(lima_oracle* o) { delete o; }
int64_t lima_oracle_n_particles(lima_oracle* o) { return o->N; }

static void to_layout(const lima_oracle* o, const std::vector<double>& src, int ncomp, double* dst) {
    if (!dst) return;
    std::fill(dst, dst + ((size_t)o->nC * 32 + o->nT) * ncomp, 0.0);
    for (int64_t i = 0; i < o->N; i++) {
        const size_t at = o->comp[i] < o->nC ? (size_t)o->comp[i] * 32 + o->slot[i] : (size_t)o->nC * 32 + (o->comp[i] - o->nC);
        for (int k = 0; k < ncomp; k++) dst[at * ncomp + k] = src[i * ncomp + k];
    }
}

int lima_oracle_forces(lima_oracle* o, int comps, double* force, double* pot, lima_energy* e) {
    evaluate(o, comps, o->F, o->pot, o->E);
    o->have_forces = comps == LIMA_ORC_ALL;
    to_layout(o, o->F, 3, force);
    to_layout(o, o->pot, 1, pot);
    if (e) *e = o->E;
    return 0;
}

/* LIMA's thermostat.  hostMaster measures the temperature after a step when the logging buffer
 * is full (step % (5 * data_logging_interval) == 0, SimulationDevice.cuh:50-54) and the step is a
 * positive multiple of steps_per_temperature_measurement (Engine.cu:162-174).  Temperature from
 * the kinetic energy of the stored velocities with dof = 3N - 3 (Thermostat.cuh:87-103,
 * PhysicsUtils.cuh:32-34, BoxBuilder.cpp:108); the scale clamp(310 K / T, 1 -+ 0.001 / every)
 * (Thermostat.cuh:56-67) replaces the integrator's scale only when apply_thermostat is set. */
static void thermostat(lima_oracle* o) {
    const int64_t buf = 5 * o->log_interval;
    if (o->thermo_every <= 0 || buf <= 0 || o->step <= 0 || o->step % buf != 0 || o->step % o->thermo_every != 0) return;
    double ke = 0;
    for (int64_t i = 0; i < o->N * 3; i++) ke += 0.5 * o->mass[i / 3] * o->v[i] * o->v[i];
    const double dof = 3.0 * (double)o->N - 3.0;
    const float T = dof > 0 ? (float)(ke * (2.0 / dof) / ((double)1.38066e-23f * 6.02214076e23)) : 0.f;
    o->temperature = T;
    if (!o->thermo_apply) return;
    const float safe = T == 0.f ? 1.f : T;
    const float lim = 0.001f / (float)o->thermo_every;
    o->thermo_scale = std::clamp(310.f / safe, 1.f - lim, 1.f + lim);
}

int lima_oracle_thermostat(lima_oracle* o, float* out) {
    out[0] = o->thermo_scale;
    out[1] = o->temperature;
    return 0;
}

/* Velocity Verlet as LIMA integrates it (EngineUtils.cuh:31-41, EngineKernels.cuh:383-397):
 * v(t) = v(t-dt) + (F(t) + F(t-dt)) dt/2m ; x += Coord{v(t) dt + F(t) dt^2/2m} (truncation). */
int lima_oracle_step(lima_oracle* o, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps; s++) {
        evaluate(o, LIMA_ORC_ALL, o->F, o->pot, o->E);
        const double dt = o->dt_ns;
        double ke = 0;
        for (int64_t i = 0; i < o->N; i++) {
            const double m = o->mass[i];
            for (int d = 0; d < 3; d++) {
                const double f = o->F[i * 3 + d];
                const double vn = o->v[i * 3 + d] + (f + o->fprev[i * 3 + d]) * (dt * 0.5 / m);
                const double dx = vn * dt + f * (0.5 / m * dt * dt);
                o->u[i * 3 + d] += (int64_t)std::trunc(dx * NANO_TO_LIMA);
                o->v[i * 3 + d] = vn * (double)o->thermo_scale;   // EngineKernels.cuh:406-411
                o->fprev[i * 3 + d] = f;
                ke += 0.5 * m * vn * vn;
            }
        }
        o->E.kinetic = ke;
        o->E.total = o->E.potential + ke;
        o->step++;
        thermostat(o);
    }
    o->have_forces = false;
    return 0;
}

int lima_oracle_minimize(lima_oracle* o, double tol_kj, int64_t max_iter, double h0, lima_em_result* res,
                         double* trace_e, double* trace_f) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> F, pot, Ft, pt;
    lima_energy E{}, Et{};
    evaluate(o, LIMA_ORC_ALL, F, pot, E);
    auto fmax_of = [&](const std::vector<double>& f) {
        double m = 0;
        for (int64_t i = 0; i < o->N; i++) m = std::max(m, std::sqrt(f[i * 3] * f[i * 3] + f[i * 3 + 1] * f[i * 3 + 1] + f[i * 3 + 2] * f[i * 3 + 2]));
        return m;
    };
    double fmax = fmax_of(F), h = h0;
    int64_t it = 0, acc = 0;
    std::vector<int64_t> u0;
    bool conv = fmax / 1000.0 < tol_kj;
    while (!conv && it < max_iter) {
        u0 = o->u;
        for (int64_t i = 0; i < o->N * 3; i++) o->u[i] += (int64_t)std::trunc(h * F[i] / fmax * NANO_TO_LIMA);
        evaluate(o, LIMA_ORC_ALL, Ft, pt, Et);
        if (Et.potential < E.potential) {
            F.swap(Ft); pot.swap(pt); E = Et; fmax = fmax_of(F); h *= 1.2; acc++;
        } else {
            o->u = u0; h *= 0.2;
        }
        if (trace_e) trace_e[it] = E.potential;
        if (trace_f) trace_f[it] = fmax / 1000.0;
        it++;
        conv = fmax / 1000.0 < tol_kj;
    }
    o->F = F; o->pot = pot; o->E = E; o->have_forces = true;
    if (res) {
        res->iterations = it; res->accepted = acc; res->final_energy = E.potential; res->final_fmax = fmax / 1000.0;
        res->converged = conv ? 1 : 0;
        res->wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return 0;
}

int lima_oracle_get_positions(lima_oracle* o, int64_t* pos) {
    std::memcpy(pos, o->u.data(), o->u.size() * 8);
    return 0;
}

int lima_oracle_get_state(lima_oracle* o, lima_state_out* out) {
    for (int c = 0; c < o->nC; c++) {
        const int64_t kidx = o->gidx(c, o->key[c]);
        int64_t base[3];
        for (int d = 0; d < 3; d++) {
            const int64_t wk = floormod(o->u[kidx * 3 + d], o->Lq[d]);
            const int32_t org = (int32_t)floordiv(wk, (int64_t)LIMA_NANO_TO_LIMA);
            if (out->compound_origo) out->compound_origo[c * 3 + d] = org;
            base[d] = o->u[kidx * 3 + d] - wk + (int64_t)org * LIMA_NANO_TO_LIMA;
        }
        for (int s = 0; s < (int)(o->first[c + 1] - o->first[c]); s++) {
            const int64_t i = o->gidx(c, s);
            const size_t ps = (size_t)c * 32 + s;
            for (int d = 0; d < 3; d++) {
                if (out->particle_coord) out->particle_coord[ps * 3 + d] = (int32_t)(o->u[i * 3 + d] - base[d]);
                if (out->particle_velocity) out->particle_velocity[ps * 3 + d] = (float)o->v[i * 3 + d];
                if (out->particle_force) out->particle_force[ps * 3 + d] = o->F.empty() ? 0.f : (float)o->F[i * 3 + d];
            }
            if (out->particle_pot_energy) out->particle_pot_energy[ps] = o->pot.empty() ? 0.f : (float)o->pot[i];
        }
    }
    for (int t = 0; t < o->nT; t++) {
        const int64_t i = o->first[o->nC] + t;
        for (int d = 0; d < 3; d++) {
            const int64_t w = floormod(o->u[i * 3 + d], o->Lq[d]);
            const int32_t blk = (int32_t)floordiv(w, (int64_t)LIMA_NANO_TO_LIMA);
            if (out->tinymol_block) out->tinymol_block[t * 3 + d] = blk;
            if (out->tinymol_coord) out->tinymol_coord[t * 3 + d] = (int32_t)(w - (int64_t)blk * LIMA_NANO_TO_LIMA);
            if (out->tinymol_velocity) out->tinymol_velocity[t * 3 + d] = (float)o->v[i * 3 + d];
            if (out->tinymol_force) out->tinymol_force[t * 3 + d] = o->F.empty() ? 0.f : (float)o->F[i * 3 + d];
        }
        if (out->tinymol_pot_energy) out->tinymol_pot_energy[t] = o->pot.empty() ? 0.f : (float)o->pot[i];
    }
    return 0;
}

static thread_local ClusterList g_cl;

int lima_oracle_neighbor_counts(lima_oracle* o, int64_t* ncl, int64_t* nent) {
    g_cl = build_clusters(o);
    *ncl = (int64_t)g_cl.atoms.size() / 8;
    *nent = (int64_t)g_cl.ent.size();
    return 0;
}

int lima_oracle_neighbor_list(lima_oracle* o, int32_t* atoms, int64_t* off, uint32_t* ent, uint8_t* mask) {
    (void)o;
    std::memcpy(atoms, g_cl.atoms.data(), g_cl.atoms.size() * 4);
    std::memcpy(off, g_cl.off.data(), g_cl.off.size() * 8);
    std::memcpy(ent, g_cl.ent.data(), g_cl.ent.size() * 4);
    std::memcpy(mask, g_cl.mask.data(), g_cl.mask.size());
    return 0;
}

/* Brute-force exact pair set: unordered non-intra pairs (i<j) and image shift with
 * integer squared distance <= round(radius 1e8)^2; shift is that of j relative to i. */
static std::vector<int32_t> g_pi, g_pj, g_ps;
int64_t lima_oracle_pair_count(lima_oracle* o, double radius_nm) {
    g_pi.clear(); g_pj.clear(); g_ps.clear();
    const int64_t R = llround(radius_nm * NANO_TO_LIMA), R2 = R * R;
    for (int64_t i = 0; i < o->N; i++)
        for (int64_t j = i + 1; j < o->N; j++) {
            if (o->intra(i, j)) continue;
            for (int code = 0; code < 27; code++) {
                const int s[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
                int64_t d2 = 0;
                bool far = false;   // a component beyond R: its square could overflow int64 in wide boxes
                for (int d = 0; d < 3 && !far; d++) {
                    const int64_t dd = o->wrapped(j, d) + s[d] * o->Lq[d] - o->wrapped(i, d);
                    if (dd > R || dd < -R) far = true;
                    else d2 += dd * dd;
                }
                if (!far && d2 <= R2) { g_pi.push_back((int32_t)i); g_pj.push_back((int32_t)j); g_ps.push_back(code); }
            }
        }
    return (int64_t)g_pi.size();
}

int lima_oracle_pairs(lima_oracle* o, double radius_nm, int32_t* pi, int32_t* pj, int32_t* ps) {
    (void)o; (void)radius_nm;
    std::memcpy(pi, g_pi.data(), g_pi.size() * 4);
    std::memcpy(pj, g_pj.data(), g_pj.size() * 4);
    std::memcpy(ps, g_ps.data(), g_ps.size() * 4);
    return 0;
}

void lima_oracle_set_pme_divide_modulus(lima_oracle* o, int on) { o->pme_divide_modulus = on != 0; }

int lima_oracle_pme_grid_dims(lima_oracle* o, int32_t* dims) {
    for (int d = 0; d < 3; d++) dims[d] = o->K[d];
    return 0;
}

int lima_oracle_pme_grids(lima_oracle* o, double* rho, double* phi) {
    std::vector<double> F(o->N * 3), pot(o->N);
    lima_energy E{};
    o->keep_grids = true;
    pme(o, F, pot, E);
    o->keep_grids = false;
    if (rho) std::memcpy(rho, o->rho_last.data(), o->rho_last.size() * 8);
    if (phi) std::memcpy(phi, o->phi_last.data(), o->phi_last.size() * 8);
    return 0;
}

int lima_oracle_fft3d(int nx, int ny, int nz, const double* re, const double* im, double* ro, double* io, int inverse) {
    std::vector<cd> a((size_t)nx * ny * nz);
    for (size_t k = 0; k < a.size(); k++) a[k] = cd(re[k], im ? im[k] : 0.0);
    fft3d(a, nx, ny, nz, inverse ? +1 : -1);
    for (size_t k = 0; k < a.size(); k++) { ro[k] = a[k].real(); io[k] = a[k].imag(); }
    return 0;
}

}  // extern "C"
