/*
 * lima_topo.h -- GROMACS-format inputs to a LIMA box (host C++, liblima_host.so).
 *
 * Replaces LIMA's input side of the force path:
 *   - GroFile reading and writing (LIMA_BASE/include/MDFiles.h:21-53, src/MDFiles.cpp:13-187);
 *   - TopologyFile reading: sections, #include, #define / #ifdef (src/Topologyfile.cpp:8-400);
 *   - LIMAForcefield parameter lookup: atom types, bonded types with "X" wildcards, both atom
 *     orders, all equally specific matches, 1-4 pairs combined from atom types when no pair type
 *     exists (src/Forcefield.cpp:41-479), with LIMA's unit conversions (src/Bodies.cpp:9-47);
 *   - the molecule -> compound builder and its exclusion tables (LIMA_TOOLS/src/CompoundBuilder.cpp:17-692).
 *
 * Behaviour kept from LIMA:
 *   - molecules are the connected components of the bond graph; those of at most 3 atoms (other
 *     than residue "lxx") become single-site tinymols at their first atom, with the LJ parameters
 *     and mass of that atom's type (OW gains the mass of two HW), and their bonded terms dropped
 *     (CompoundBuilder.cpp:287-339, Forcefield.cpp:262-278);
 *   - compounds hold at most 32 particles of one molecule; particles sharing a bond, angle,
 *     dihedral or improper are excluded from the non-bonded pairs (CompoundBuilder.cpp:17-93);
 *   - the second [ dihedrals ] section of a file holds impropers, the second [ dihedraltypes ]
 *     section improper types (Topologyfile.cpp:8-47); parameters written on a term's line are
 *     used instead of the forcefield's;
 *   - FLEXIBLE is defined; includes whose name contains "posre" are skipped.
 * Differences: #ifdef / #ifndef / #else / #endif follow the C preprocessor; angle parameters are
 * read by function type (5: theta0 k ub0 kub; otherwise theta0 k, B-state values ignored), where
 * LIMA always reads four numbers; a molecule is split
 * into compounds by this library's own partition (the ends of long branches first); a term
 * whose particles lie in compounds more than 2 apart is an error (LIMA's exclusion tables reach
 * 2 compounds).  See lima_load_system for the flags.
 */
#ifndef LIMA_TOPO_H
#define LIMA_TOPO_H

#include <stdint.h>
#include "lima_gen.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lima_system lima_system;

/* Keep molecules of at most 3 atoms (e.g. 3-site water) as compounds with all their charges
 * and bonded terms, one compound per molecule, instead of LIMA's single-site tinymols. */
#define LIMA_TOPO_KEEP_SMALL_MOLECULES 1

/* Reads gro_path and top_path (Environment::CreateSimulation(GroFile, TopologyFile, ...) ->
 * LIMA_MOLECULEBUILD::buildMolecules, CompoundBuilder.cpp:597-692).  #include files are looked
 * up next to the including file, then in include_dirs[0 .. n_include_dirs-1] (LIMA looks in
 * resources/forcefields).  Atoms of the topology's [ molecules ] map to the .gro atoms in order. */
lima_status lima_load_system(const char* gro_path, const char* top_path, const char* const* include_dirs,
                             int32_t n_include_dirs, int32_t flags, lima_system** out);
void lima_system_free(lima_system* sys);

typedef struct lima_system_info {
    int64_t n_gro_atoms;
    int32_t n_molecules;          /* connected components with more than 3 atoms (or kept small ones) */
    int32_t n_compounds;
    int32_t n_tinymols;
    int32_t n_bonds, n_angles, n_dihedrals, n_impropers, n_pairs;   /* terms in the box */
} lima_system_info;
lima_status lima_system_get_info(const lima_system* sys, lima_system_info* out);

/* A new box holding the system (free with lima_box_free; describe with lima_box_describe).
 * Velocities are the .gro velocities converted to nm/ns, or zero. */
lima_status lima_system_make_box(const lima_system* sys, lima_box** out);

/* gro_index[n_compounds * 32 + n_tinymols]: the .gro atom (0-based) of each entry of LIMA's
 * buffer layout (a tinymol's first atom), -1 for unused compound slots. */
lima_status lima_system_layout(const lima_system* sys, int32_t* gro_index);

/* Environment::WriteBoxCoordinatesToFile + GroFile::printToFile (Environment.cpp:232-263,
 * MDFiles.cpp:57-80,162-187): writes the .gro with positions[n_compounds * 32 + n_tinymols][3]
 * [nm] (LIMA's buffer layout, e.g. a lima_read_log frame).  Compound particles take their entry;
 * the atoms of a tinymol move by the displacement of its first atom. */
lima_status lima_system_write_gro(const lima_system* sys, const float* positions, const char* path);

const char* lima_topo_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
