/*
 * lima_oracle.h -- host FP64 oracle of LIMA's per-step force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the reference the HIP path is tested
 * against, and the CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * It restates LIMA's force-field formulas in double precision (OpenMP over
 * particles).  Each function in lima_oracle.cpp cites the reference lines it
 * follows.  Inputs are the same lima_box_desc / lima_simparams the product
 * takes; outputs use LIMA's [n_compounds][32] + [n_tinymols] layout.
 */
#ifndef LIMA_ORACLE_H
#define LIMA_ORACLE_H

#include <stdint.h>
#include "../include/lima_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lima_oracle lima_oracle;

int  lima_oracle_create(const lima_box_desc* box, const lima_simparams* params, lima_oracle** out);
void lima_oracle_destroy(lima_oracle* o);
const char* lima_oracle_last_error(void);

/* Number of OpenMP threads the oracle uses (0 = OpenMP default). */
void lima_oracle_set_threads(int n);

/* Component mask for lima_oracle_forces. */
enum {
    LIMA_ORC_NB_CUTOFF = 1,  /* cutoff pairs (compoundFarneighbor path) */
    LIMA_ORC_NB_INTRA  = 2,  /* intra / bonded-compound pairs (immediate path) */
    LIMA_ORC_BONDED    = 4,
    LIMA_ORC_PME       = 8,
    LIMA_ORC_ALL       = 15
};

/* Force/energy evaluation at the current state.  force [n_compounds*32*3 + n_tinymols*3],
 * pot [n_compounds*32 + n_tinymols] (LIMA per-particle potE), either may be NULL. */
int lima_oracle_forces(lima_oracle* o, int components, double* force, double* pot, lima_energy* e);

/* nsteps of LIMA's velocity Verlet on integer Coord positions, with LIMA's thermostat
 * (temperature measured, and the velocity scale replaced when apply_thermostat is set). */
int lima_oracle_step(lima_oracle* o, int64_t nsteps);

/* out[0] = velocity scale the integrator applies, out[1] = last measured temperature [K]
 * (NaN before the first measurement). */
int lima_oracle_thermostat(lima_oracle* o, float* out);

/* Steepest descent, FP64 acceptance.  trace_energy / trace_fmax (may be NULL)
 * receive the accepted energy and fmax after every iteration (max_iter entries). */
int lima_oracle_minimize(lima_oracle* o, double fmax_tol_kj, int64_t max_iter, double initial_step_nm,
                         lima_em_result* res, double* trace_energy, double* trace_fmax);

/* State in LIMA layout (same struct as the product, float fields filled from doubles). */
int lima_oracle_get_state(lima_oracle* o, lima_state_out* out);
/* Unwrapped integer positions [N][3] in Coord units, global particle order
 * (compound particles in compound order, then tinymols). */
int lima_oracle_get_positions(lima_oracle* o, int64_t* pos);
int64_t lima_oracle_n_particles(lima_oracle* o);

/* Cluster-pair neighbour list (same definition as the product; see lima_oracle.cpp). */
int lima_oracle_neighbor_counts(lima_oracle* o, int64_t* n_clusters, int64_t* n_entries);
int lima_oracle_neighbor_list(lima_oracle* o, int32_t* cluster_atoms, int64_t* entry_offsets,
                              uint32_t* entries, uint8_t* entry_masks);
/* Exact atom-pair set within (cutoff+skin) with images: pairs i<j (global order) and shift code. */
int64_t lima_oracle_pair_count(lima_oracle* o, double radius_nm);
int lima_oracle_pairs(lima_oracle* o, double radius_nm, int32_t* pi, int32_t* pj, int32_t* shift_code);

/* PME building blocks exposed for pinning tests. */
int lima_oracle_pme_grid_dims(lima_oracle* o, int32_t* dims);
/* Test switch: divide the Green's function by the B-spline transform (textbook smooth PME)
 * instead of multiplying (LIMA).  Off by default. */
void lima_oracle_set_pme_divide_modulus(lima_oracle* o, int on);
/* Charge-density grid after spreading [Kz][Ky][Kx] and potential grid after the solve. */
int lima_oracle_pme_grids(lima_oracle* o, double* rho, double* phi);
/* 1D/3D complex FFT used by the oracle (forward sign -1), for tests against numpy. */
int lima_oracle_fft3d(int nx, int ny, int nz, const double* re_in, const double* im_in, double* re_out, double* im_out, int inverse);

#ifdef __cplusplus
}
#endif
#endif
