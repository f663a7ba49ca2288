/* State dump and restore in the reference's HDF5 wire format.
 *
 * Replaces the reference's dumper module (playground/src/utils/dumper.f90):
 * dump (:14-42), restore (:44-66) and clean (:68-72), with the same files and
 * datasets:
 *   - <dir>/dump_full.h5, written as <dir>/new_dump_full.h5 and renamed, plus
 *     a copy <dir>/dump_NNNNN.h5 numbered by statevars(ec_lastprint) (Fortran
 *     i5.5: five digits, "*****" when the number does not fit);
 *   - dataset "state": FP64, dims (49), max dims (unlimited), one chunk;
 *   - dataset "store": FP64, dims (N, 32) in C order (the Fortran (32, N) of
 *     store(es, i)), max dims (unlimited, unlimited), one chunk, with
 *     N = int(statevars(ec_realpn)) + int(statevars(ec_fixedpn)).
 * The fortdriver opens the same datasets with h5py and reads store[i, es - 1]
 * (playground/runner/fortdriver/core.py:322-324,453).
 *
 * Built as libsph_dump.so on the HDF5 C library; not part of the GPU path.
 * Return codes: 0 success, SPH_DUMP_ERR_* below.
 */
#ifndef SPH_DUMP_H
#define SPH_DUMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPH_DUMP_OK 0
#define SPH_DUMP_ERR_ARG (-1)   /* null pointer, ld < 32, N < 1, capacity too small */
#define SPH_DUMP_ERR_HDF5 (-2)  /* an HDF5 call failed (missing file, unreadable dataset, ...) */
#define SPH_DUMP_ERR_IO (-3)    /* rename or copy of the dump file failed */
#define SPH_DUMP_ERR_SHAPE (-4) /* a dataset of the file has an unexpected rank or size */

/* dump (dumper.f90:14-42): state with ec_time set to t, and the first N rows
 * of store (row i = particle i's 32 fields, rows ld >= 32 doubles apart). */
int sph_dump_write(const char *dir, const double state[49], const double *store, int64_t ld, double t);

/* The number of store rows in <dir>/dump_full.h5, for sizing a restore. */
int sph_dump_rows(const char *dir, int64_t *n);

/* restore (dumper.f90:44-66): the state and the store rows of
 * <dir>/dump_full.h5 into state[49] and store (cap_rows rows, ld apart). */
int sph_dump_read(const char *dir, double state[49], double *store, int64_t cap_rows, int64_t ld);

/* clean (dumper.f90:68-72): delete <dir>/new_dump_full.h5 if it exists. */
int sph_dump_clean(const char *dir);

#ifdef __cplusplus
}
#endif

#endif /* SPH_DUMP_H */
