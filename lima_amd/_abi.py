"""ctypes mirrors of the C ABI structs in include/lima_hip.h and include/lima_gen.h."""
import ctypes as C

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u32p = C.POINTER(C.c_uint32)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class BoxDesc(C.Structure):
    _fields_ = [
        ("box_nm", C.c_double * 3),
        ("n_compounds", C.c_int32),
        ("n_tinymols", C.c_int32),
        ("compound_n_particles", i32p),
        ("compound_origo", i32p),
        ("compound_key_particle", i32p),
        ("particle_coord", i32p),
        ("particle_velocity", f32p),
        ("particle_charge", f32p),
        ("particle_sigma", f32p),
        ("particle_epsilon", f32p),
        ("particle_mass", f32p),
        ("bonded_lut", u32p),
        ("tinymol_block", i32p),
        ("tinymol_coord", i32p),
        ("tinymol_velocity", f32p),
        ("tinymol_sigma", f32p),
        ("tinymol_epsilon", f32p),
        ("tinymol_mass", f32p),
        ("n_bonds", C.c_int32), ("bond_ids", i32p), ("bond_params", f32p),
        ("n_angles", C.c_int32), ("angle_ids", i32p), ("angle_params", f32p),
        ("n_dihedrals", C.c_int32), ("dihedral_ids", i32p), ("dihedral_params", f32p),
        ("n_impropers", C.c_int32), ("improper_ids", i32p), ("improper_params", f32p),
        ("n_pairs", C.c_int32), ("pair_ids", i32p), ("pair_params", f32p),
    ]


class SimParams(C.Structure):
    _fields_ = [
        ("n_steps", C.c_int64),
        ("dt_fs", C.c_float),
        ("em_variant", C.c_int32),
        ("em_force_tolerance", C.c_float),
        ("bc_select", C.c_int32),
        ("enable_electrostatics", C.c_int32),
        ("cutoff_nm", C.c_float),
        ("data_logging_interval", C.c_int32),
        ("apply_thermostat", C.c_int32),
        ("steps_per_temperature_measurement", C.c_int64),
        ("pme_grid", C.c_int32 * 3),
        ("pme_spacing_nm", C.c_float),
        ("nlist_skin_nm", C.c_float),
        ("nlist_interval", C.c_int32),
        ("energy_interval", C.c_int32),
        ("launch_mode", C.c_int32),
        ("log_capacity", C.c_int32),
        ("nlist_capacity", C.c_int32),
    ]


class Energy(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "lj", "coulomb_sr", "bonded", "pme_recip", "pme_self_lima", "pme_self_ewald",
        "potential", "potential_lima", "kinetic", "total")] + [("virial", C.c_double * 6), ("step", C.c_int64)]

    VIRIAL_COMPONENTS = ("xx", "yy", "zz", "xy", "xz", "yz")

    def as_dict(self):
        return {n: (list(getattr(self, n)) if n == "virial" else getattr(self, n)) for n, _ in self._fields_}

    def virial_tensor(self):
        """W as a symmetric 3x3 numpy array [J/mol]."""
        import numpy as np
        v = list(self.virial)
        return np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])


class RunStatus(C.Structure):
    _fields_ = [
        ("current_step", C.c_int64),
        ("current_temperature", C.c_float),
        ("greatest_force", C.c_float),
        ("simulation_finished", C.c_int32),
        ("critical_error_occured", C.c_int32),
        ("n_list_rebuilds", C.c_int32),
        ("list_entries", C.c_int64),
        ("list_violations", C.c_int64),
        ("max_displacement_nm", C.c_float),
        ("capacity_rollbacks", C.c_int32),
        ("list_capacity", C.c_int32),
        ("list_max_entries", C.c_int32),
        ("list_interval", C.c_int32),
        ("list_stale_rollbacks", C.c_int32),
    ]


class StateOut(C.Structure):
    _fields_ = [
        ("compound_origo", i32p),
        ("particle_coord", i32p),
        ("particle_velocity", f32p),
        ("particle_force", f32p),
        ("particle_pot_energy", f32p),
        ("tinymol_block", i32p),
        ("tinymol_coord", i32p),
        ("tinymol_velocity", f32p),
        ("tinymol_force", f32p),
        ("tinymol_pot_energy", f32p),
    ]


class EmResult(C.Structure):
    _fields_ = [
        ("iterations", C.c_int64),
        ("accepted", C.c_int64),
        ("final_energy", C.c_double),
        ("final_fmax", C.c_double),
        ("wall_seconds", C.c_double),
        ("converged", C.c_int32),
    ]


class Timings(C.Structure):
    _fields_ = [
        ("pair_ms", C.c_double), ("intra_ms", C.c_double), ("bonded_ms", C.c_double),
        ("pme_ms", C.c_double), ("integrate_ms", C.c_double), ("nlist_ms", C.c_double),
        ("total_ms", C.c_double), ("steps_timed", C.c_int64),
        ("host_issue_ms", C.c_double), ("host_steps", C.c_int64),
    ]


class BatchTuning(C.Structure):
    _fields_ = [
        ("n_buckets", C.c_int32),
        ("bucket_max_atoms", C.c_int32 * 8),
        ("threads", C.c_int32 * 8),
        ("pair_unroll", C.c_int32 * 8),
        ("us_per_iter", C.c_double * 8),
    ]

    def as_list(self):
        return [{"max_atoms": self.bucket_max_atoms[k], "threads": self.threads[k], "pair_unroll": self.pair_unroll[k],
                 "us_per_iter": round(self.us_per_iter[k], 2)} for k in range(self.n_buckets)]


class LogInfo(C.Structure):
    _fields_ = [
        ("frames_logged", C.c_int64),
        ("first_held", C.c_int64),
        ("particles_per_frame", C.c_int64),
        ("interval", C.c_int32),
        ("capacity", C.c_int32),
    ]


class LogOut(C.Structure):
    _fields_ = [
        ("step", C.POINTER(C.c_int64)),
        ("energy", C.POINTER(Energy)),
        ("position", C.POINTER(C.c_float)),
        ("pot_energy", C.POINTER(C.c_float)),
        ("speed", C.POINTER(C.c_float)),
        ("force", C.POINTER(C.c_float)),
    ]


class EnergyRow(C.Structure):
    _fields_ = [("min", C.c_float), ("max", C.c_float), ("stddev", C.c_float), ("change", C.c_float)]


class EnergyAnalysis(C.Structure):
    _fields_ = [
        ("n_entries", C.c_int64),
        ("pot", EnergyRow), ("kin", EnergyRow), ("total", EnergyRow),
        ("mean_energy", C.c_float),
        ("energy_gradient", C.c_float),
        ("variance_coefficient", C.c_float),
    ]


class SystemInfo(C.Structure):
    _fields_ = [
        ("n_gro_atoms", C.c_int64),
        ("n_molecules", C.c_int32),
        ("n_compounds", C.c_int32),
        ("n_tinymols", C.c_int32),
        ("n_bonds", C.c_int32), ("n_angles", C.c_int32), ("n_dihedrals", C.c_int32),
        ("n_impropers", C.c_int32), ("n_pairs", C.c_int32),
    ]
