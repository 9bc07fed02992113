/*
 * lima_gen.h -- deterministic synthetic box generators (host C++, liblima_host.so).
 *
 * They stand in for LIMA's topology -> compound builder (LIMA_TOOLS/src/CompoundBuilder.cpp,
 * BoxBuilder.cpp) for the synthetic water / chain workloads: they partition molecules into
 * compounds of <= 32 particles, build the BondedParticlesLUT exclusion tables
 * (CompoundBuilder.cpp:17-93) and convert parameters to LIMA's stored units
 * (Bodies.cpp:9-47, Forcefield.cpp:334-351).  All randomness comes from a project-written
 * SplitMix64 / xoshiro256** stream, so a seed gives the same box on every platform.
 */
#ifndef LIMA_GEN_H
#define LIMA_GEN_H

#include <stdint.h>
#include "lima_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lima_box lima_box;   /* owning container of every array a lima_box_desc points to */

/* n^3 rigid-geometry 3-site waters on a lattice in a cube of edge box_nm, random orientations,
 * Maxwell-Boltzmann velocities at temperature_K (0 = at rest). */
lima_status lima_gen_water_lattice(int32_t n_per_dim, double box_nm, uint64_t seed, double temperature_K, lima_box** out);

/* Chain slab + water: n_chains linear chains of chain_len atoms in two facing leaflets along z,
 * then exactly n_waters waters in the free volume: on a lattice at 1 g/cm^3 when they fit there,
 * otherwise on successively finer lattices with shorter clash distances (a denser start that
 * minimisation relaxes).  LIMA_ERR_BAD_INPUT when n_waters cannot be placed. */
lima_status lima_gen_chain_water(double bx, double by, double bz, int32_t n_chains, int32_t chain_len,
                                 int32_t n_waters, uint64_t seed, double temperature_K, lima_box** out);

/* One minimisation system: about target_atoms atoms of chains + water at 1 g/cm^3, positions
 * perturbed by Gaussian noise of noise_nm. */
lima_status lima_gen_em_system(int32_t target_atoms, uint64_t seed, double noise_nm, lima_box** out);

/* Large water box: n_waters waters plus n_extra neutral LJ tinymols (sigma, eps kJ/mol, mass g/mol)
 * in a cube at 1 g/cm^3. */
lima_status lima_gen_large_water(int32_t n_waters, int32_t n_extra, double extra_sigma, double extra_eps_kj,
                                 double extra_mass, uint64_t seed, double temperature_K, lima_box** out);

void     lima_box_describe(const lima_box* box, lima_box_desc* desc);
void     lima_box_free(lima_box* box);
uint64_t lima_box_hash(const lima_box* box);
int64_t  lima_box_n_particles(const lima_box* box);
const char* lima_gen_last_error(void);

/* Default LIMA SimParams for a config (cutoff 1.2 nm, PME spacing 0.12 nm, skin 0.1 nm). */
void lima_default_params(lima_simparams* p, float dt_fs, int64_t n_steps);

#ifdef __cplusplus
}
#endif
#endif
