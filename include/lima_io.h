/*
 * lima_io.h -- host-side tools either side of the force path (host C++, liblima_host.so):
 * energy analysis of a data log and trajectory output.
 *
 * Errors: every function returns a lima_status; lima_io_last_error() returns a thread-local
 * message for the last failing call on the calling thread.
 */
#ifndef LIMA_IO_H
#define LIMA_IO_H

#include <stdint.h>
#include "lima_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SimAnalysis::analyzeEnergy + AnalyzedPackage (LIMA_MD/src/Analyzer.cu:216-366) ----
 * Input: the potential and kinetic energy of n_logged consecutive logged steps (LIMA sums the
 * per-particle potE and 1/2 m speed^2 of each logged step; lima_read_log's energy.potential_lima
 * and energy.kinetic are these sums in FP64).  As LIMA does, the first and the last entry are
 * dropped (Analyzer.cu:245,270) and the rest are analysed in FP32 with LIMA's formulas:
 *   row (printRow, Analyzer.cu:306-315): min, max, population standard deviation
 *       (Statistics::StdDev, LIMA_BASE/src/Statistics.cpp:112-125), change (last - first) / first;
 *   mean_energy: mean total energy (Statistics::Mean, Statistics.cpp:106-110);
 *   energy_gradient: least-squares slope of total energy against the entry index, divided by the
 *       mean (calculateSlopeLinearRegression, Analyzer.cu:327-344, FP32 sums as there);
 *   variance_coefficient: stddev / |mean| of total energy, 0 when both are 0
 *       (getVarianceCoefficient, Analyzer.cu:297-304).
 * With fewer than 3 logged entries nothing is analysed (n_entries 0, every field 0). */
typedef struct lima_energy_row {
    float min, max, stddev, change;
} lima_energy_row;
typedef struct lima_energy_analysis {
    int64_t n_entries;            /* entries analysed: n_logged - 2 */
    lima_energy_row pot, kin, total;
    float mean_energy;
    float energy_gradient;
    float variance_coefficient;
} lima_energy_analysis;
lima_status lima_analyze_energy(const double* pot, const double* kin, int64_t n_logged, lima_energy_analysis* out);

/* ---- MDFiles::TrrFile::Dump (LIMA_BASE/src/TrrFileInterface.cpp:123-196): positions-only TRR
 * frames.  Frame k holds positions[k][n_atoms][3] [nm], the box diag(box_nm), step steps[k]
 * (LIMA writes the frame index: pass NULL for that), time 0 and lambda 0, in single precision.
 * The path must end in ".trr" (as LIMA's TrrWriter requires).
 * flags 0: the GROMACS format, XDR big-endian.  LIMA_TRR_HOST_ORDER: the same fields in host
 * byte order, which is what LIMA's writer emits (its write_as_big_endian does not swap bytes). */
#define LIMA_TRR_HOST_ORDER 1
lima_status lima_write_trr(const char* path, int64_t n_frames, int64_t n_atoms, const double* box_nm,
                           const float* positions, const int64_t* steps, int32_t flags);

const char* lima_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
