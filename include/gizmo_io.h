/* gizmo_io.h -- GIZMO HDF5 initial conditions and snapshots for the hot path's gas fields.
 *
 * The file layout is the reference's (single file per snapshot, gas = group PartType0):
 *   Header attributes as written by write_header_attributes_in_hdf5 (src/file_io/io.cc:5324-5440)
 *   and read by read_header_attributes_in_hdf5 (src/file_io/read_ic.cc:2204-2260);
 *   dataset names as get_dataset_name (src/file_io/io.cc:4374-4800): Coordinates, Velocities,
 *   ParticleIDs, Masses, InternalEnergy, Density, KernelMaxRadius (read also under its legacy
 *   name SmoothingLength, get_dataset_name_legacy_alias io.cc:4848-4866), Pressure,
 *   HydroAcceleration, RateOfChangeOfInternalEnergy, VelocityDivergence, DensityGradient,
 *   VelocityGradient.
 * So a snapshot of this library can be read by the reference's readers and analysis scripts,
 * and the reference's ICs (and those its test generators write) can be loaded.
 *
 * Host code only (HDF5 C library); built as its own library, libgizmo_io.so, so the hot path
 * does not depend on HDF5.  Return codes: 0, or GZ_E_ARG / GZ_E_STATE (gizmo_hydro.h) with a
 * message in gz_io_last_error().
 */
#ifndef GIZMO_IO_H
#define GIZMO_IO_H

#include "gizmo_hydro.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    long long num_gas;     /* Header/NumPart_ThisFile[0] (= NumPart_Total[0]: one file) */
    double time;           /* Header/Time (All.Time) */
    double redshift;       /* Header/Redshift */
    double box_size;       /* Header/BoxSize (0 when absent) */
    double mass_table_gas; /* Header/MassTable[0]: > 0 when every gas particle has this mass */
    int double_precision;  /* Header/Flag_DoublePrecision: 1 FP64 datasets, 0 FP32 */
    int num_files;         /* Header/NumFilesPerSnapshot (only 1 is read) */
    int legacy_names;      /* writing: the kernel radius as SmoothingLength instead of KernelMaxRadius */
} gz_snap_header;

/* The header of an IC or snapshot file. */
int gz_ic_header(const char *path, gz_snap_header *hdr);

/* Read the n gas particles of `path` into the non-NULL arrays of f: pos, vel, mass, u, h, rho
   (vectors n x 3), and ids.  Coordinates, Velocities and InternalEnergy are required; mass
   comes from Masses or else from MassTable[0]; *has_h (if given) tells whether a kernel radius
   dataset was present (h is left untouched otherwise).  Datasets of any floating type. */
int gz_ic_read(const char *path, long n, gz_fields *f, long long *ids, int *has_h);

/* Write a snapshot of n gas particles: the header, then a PartType0 dataset for every
   non-NULL array of f that has one (Velocities and InternalEnergy take vel_kick and u_kick,
   the reference's P.Vel and CellP.InternalEnergy, when given, else vel and u), and ids
   (ParticleIDs; 1..n when NULL).  FP64 or FP32 datasets by hdr->double_precision.  With
   f->mass NULL, MassTable[0] = hdr->mass_table_gas (one mass for every gas particle). */
int gz_snapshot_write(const char *path, long n, const gz_fields *f, const long long *ids, const gz_snap_header *hdr);

const char *gz_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* GIZMO_IO_H */
