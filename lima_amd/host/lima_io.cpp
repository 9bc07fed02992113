// lima_io.cpp -- host tools either side of the force path: energy analysis of a data log and
// TRR trajectory output.  Declarations and the reference behaviour they follow: include/lima_io.h.
#include "lima_io.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct IoFail : std::runtime_error {
    lima_status code;
    IoFail(lima_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class F>
lima_status guarded(F&& f) {
    try {
        f();
        return LIMA_OK;
    } catch (const IoFail& x) {
        g_err = x.what();
        return x.code;
    } catch (const std::exception& x) {
        g_err = x.what();
        return LIMA_ERR_BAD_INPUT;
    }
}

// ---- statistics in LIMA's precision (Statistics.cpp:106-125, Analyzer.cu:289-344) ----
float mean_f(const std::vector<float>& v) {   // double accumulation, float result
    double s = 0.0;
    for (float x : v) s += x;
    return (float)(s / (double)v.size());
}

float stddev_f(const std::vector<float>& v) {   // population deviation around the float mean
    if (v.empty()) return 0.f;
    const double m = mean_f(v);
    double var = 0.0;
    for (float x : v) {
        const double d = x - m;
        var += d * d;
    }
    return (float)std::sqrt(var / (double)v.size());
}

float variance_coefficient_f(const std::vector<float>& v) {
    if (v.empty()) return 0.f;
    const float sd = stddev_f(v), m = mean_f(v);
    if (sd == 0.f && m == 0.f) return 0.f;
    return sd / std::fabs(m);
}

// Least-squares slope of v against its index, with LIMA's float sums (the index enters as a
// size_t converted to float, i * i as an integer product), divided by mean.
float slope_coefficient_f(const std::vector<float>& v, float mean) {
    const size_t n = v.size();
    float sx = 0.f, sy = 0.f, sxy = 0.f, sxx = 0.f;
    for (size_t i = 0; i < n; i++) {
        sx += (float)i;
        sy += v[i];
        sxy += (float)i * v[i];
        sxx += (float)(i * i);
    }
    const float nf = (float)n;
    const float slope = (nf * sxy - sx * sy) / (nf * sxx - sx * sx);
    return slope / mean;
}

lima_energy_row row_f(const std::vector<float>& v) {
    lima_energy_row r{};
    if (v.empty()) return r;
    r.min = *std::min_element(v.begin(), v.end());
    r.max = *std::max_element(v.begin(), v.end());
    r.stddev = stddev_f(v);
    r.change = (v.back() - v.front()) / v.front();
    return r;
}

// ---- TRR (GROMACS trajectory) records ----
class TrrOut {
public:
    TrrOut(const std::string& path, bool host_order) : host_(host_order) {
        f_ = std::fopen(path.c_str(), "wb");
        if (!f_) throw IoFail(LIMA_ERR_BAD_INPUT, "cannot open '" + path + "' for writing");
    }
    ~TrrOut() {
        if (f_) std::fclose(f_);
    }
    void u32(uint32_t x) {
        if (!host_) x = __builtin_bswap32(x);   // XDR: big-endian
        put(&x, 4);
    }
    void i32(int32_t x) { u32((uint32_t)x); }
    void f32(float x) {
        uint32_t b;
        std::memcpy(&b, &x, 4);
        u32(b);
    }
    void bytes(const char* p, uint32_t n) {   // XDR opaque: length, bytes, zero padding to 4
        u32(n);
        put(p, n);
        static const char zero[4] = {0, 0, 0, 0};
        put(zero, (4 - n % 4) % 4);
    }
    void close() {
        if (f_ && std::fclose(f_) != 0) {
            f_ = nullptr;
            throw IoFail(LIMA_ERR_BAD_INPUT, "closing the trajectory failed");
        }
        f_ = nullptr;
    }

private:
    void put(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f_) != n) throw IoFail(LIMA_ERR_BAD_INPUT, "write failed");
    }
    std::FILE* f_ = nullptr;
    bool host_;
};

}  // namespace

extern "C" {

const char* lima_io_last_error(void) { return g_err.c_str(); }

lima_status lima_analyze_energy(const double* pot, const double* kin, int64_t n_logged, lima_energy_analysis* out) {
    return guarded([&] {
        if (!out || n_logged < 0 || (n_logged > 0 && (!pot || !kin))) throw IoFail(LIMA_ERR_BAD_INPUT, "bad arguments");
        std::memset(out, 0, sizeof(*out));
        if (n_logged < 3) return;
        // entries 1 .. n_logged - 2, each as LIMA's Float3(potE, kinE, potE + kinE) of float sums
        std::vector<float> vp, vk, vt;
        for (int64_t i = 1; i < n_logged - 1; i++) {
            vp.push_back((float)pot[i]);
            vk.push_back((float)kin[i]);
            vt.push_back((float)(pot[i] + kin[i]));
        }
        out->n_entries = (int64_t)vt.size();
        out->pot = row_f(vp);
        out->kin = row_f(vk);
        out->total = row_f(vt);
        out->mean_energy = mean_f(vt);
        out->energy_gradient = slope_coefficient_f(vt, out->mean_energy);
        out->variance_coefficient = variance_coefficient_f(vt);
    });
}

lima_status lima_write_trr(const char* path, int64_t n_frames, int64_t n_atoms, const double* box_nm,
                           const float* positions, const int64_t* steps, int32_t flags) {
    return guarded([&] {
        if (!path || !box_nm || n_frames < 0 || n_atoms < 0 || (n_frames > 0 && n_atoms > 0 && !positions))
            throw IoFail(LIMA_ERR_BAD_INPUT, "bad arguments");
        if (flags & ~LIMA_TRR_HOST_ORDER) throw IoFail(LIMA_ERR_BAD_INPUT, "unknown flags");
        const std::string p(path);
        if (p.size() < 4 || p.compare(p.size() - 4, 4, ".trr") != 0)
            throw IoFail(LIMA_ERR_BAD_INPUT, "expected extension .trr for '" + p + "'");
        if (n_atoms > (int64_t)(INT32_MAX / 12)) throw IoFail(LIMA_ERR_BAD_INPUT, "too many atoms for a TRR frame");
        TrrOut f(p, (flags & LIMA_TRR_HOST_ORDER) != 0);
        static const char version[] = "GMX_trn_file";
        const uint32_t vlen = (uint32_t)std::strlen(version);
        for (int64_t k = 0; k < n_frames; k++) {
            const int64_t step = steps ? steps[k] : k;
            if (step < INT32_MIN || step > INT32_MAX) throw IoFail(LIMA_ERR_BAD_INPUT, "step does not fit a TRR header");
            f.i32(1993);                       // magic
            f.u32(vlen + 1);                   // string length with its terminator
            f.bytes(version, vlen);
            // ir, e, box, vir, pres, top, sym, x, v, f sizes [bytes]; natoms, step, nre; time, lambda
            const int32_t sizes[10] = {0, 0, 9 * 4, 0, 0, 0, 0, (int32_t)(n_atoms * 12), 0, 0};
            for (int32_t s : sizes) f.i32(s);
            f.i32((int32_t)n_atoms);
            f.i32((int32_t)step);
            f.i32(0);
            f.f32(0.f);
            f.f32(0.f);
            const float box[9] = {(float)box_nm[0], 0.f, 0.f, 0.f, (float)box_nm[1], 0.f, 0.f, 0.f, (float)box_nm[2]};
            for (float b : box) f.f32(b);
            const float* x = positions + (size_t)k * (size_t)n_atoms * 3;
            for (int64_t i = 0; i < 3 * n_atoms; i++) f.f32(x[i]);
        }
        f.close();
    });
}

}  // extern "C"
