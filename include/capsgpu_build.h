/* capsgpu synthetic systems: the inputs of the parity cases and the benchmark, each with its flattened force field.
 *
 * capsgpu_build_kg replaces kremer_grest() + kremer_grest_forcefield() (core/include/caps/kremer_grest.hpp,
 * core/src/kremer_grest.cpp:15-124): the same chains from the same seed. capsgpu_build_melt has no counterpart in the
 * reference (its polymer builders are outside the force-and-step path): an all-atom polyethylene-like melt of
 * 128-atom chains for the 32,768- and 262,144-atom cases.
 */
#ifndef CAPSGPU_BUILD_H
#define CAPSGPU_BUILD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct capsgpu_system {
  int32_t n_atoms, n_types;
  double cell[12];              /* origin, a, b, c */
  double* x;                    /* 3N, Å (chains unwrapped) */
  int32_t* type;
  double* charge;               /* e */
  double* mass;                 /* g/mol */
  double* eps;                  /* nt·nt mixed ε */
  double* sigma;                /* nt·nt mixed σ */
  int32_t n_bonds;
  int32_t* bond_atoms;
  double* bond_param;           /* k, r0 */
  int32_t n_angles;
  int32_t* angle_atoms;
  double* angle_param;          /* k_θ, θ0 (radians) */
  int32_t n_fene;
  int32_t* fene_atoms;
  double* fene_param;           /* K, R0, ε, σ */
  int32_t* excl_offset;         /* CSR, sorted */
  int32_t* excl_index;
  double* v;                    /* 3N, Å/fs, or NULL: velocities, when the source holds them (a data file) */
  void* opaque;
} capsgpu_system;

const char* capsgpu_build_error(void);
/* chains of 128 atoms (CH3-(CH2)40-CH3) at density g/cm³ in a cubic box */
int32_t capsgpu_build_melt(int32_t chains, uint64_t seed, double density, capsgpu_system** out);
/* KgOptions{chains, beads, density, k_theta, bond, seed} (core/include/caps/kremer_grest.hpp:19-33) */
int32_t capsgpu_build_kg(int32_t chains, int32_t beads, double density, double k_theta, double bond, uint64_t seed,
                         capsgpu_system** out);
void capsgpu_system_free(capsgpu_system* s);

/* LAMMPS data files, atom_style full, units real: read_lammps_data / write_lammps_data (core/src/io_lammps.cpp:58-297,
 * :505-583), with the force field of this path carried in the coefficient sections — Pair Coeffs lj/cut (ε, σ per type,
 * geometric mixing) or PairIJ Coeffs, Bond Coeffs harmonic (K, r0; E = K (r - r0)^2) and fene (K, R0, ε, σ), Angle Coeffs
 * harmonic (K, θ0 in degrees). The reader unwraps positions by their image flags, keeps the Velocities section in v and
 * excludes 1-2 partners, and 1-3 ones when exclude13 is 1 (< 0: when the file has angles); dihedrals, impropers and
 * other bond styles are refused.
 * The writer writes 17 significant digits, rotates a general cell into LAMMPS's restricted triclinic frame, and writes
 * to path.tmp before renaming it to path, as the checkpoints of the command line are written (cli/main.cpp:7820-7846).
 * Both return CAPSGPU_OK or CAPSGPU_EINVAL with capsgpu_last_error() set. */
int32_t capsgpu_read_lammps(const char* path, int32_t exclude13, capsgpu_system** out);
int32_t capsgpu_write_lammps(const char* path, const capsgpu_system* s, const double* v, const char* title);

#ifdef __cplusplus
}
#endif
#endif
