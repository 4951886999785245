/* moloch_output.h -- the step's output files.
 *
 * The reference writes the surface pressure and the temperature with
 * grid_nc_create / grid_nc_write / grid_nc_destroy
 * (nompi/mod_mppparam.F90:121-480), called from the driver's output(istep)
 * before every iout-th step (nompi/regcm_dynamical_core.F90:202-215): one
 * netCDF classic file per variable, `<name>.nc`, created with nf90_clobber,
 * dimensions JX, IY[, KZ] and the unlimited KTAU, one nf90_double variable
 * named like the file over (JX, IY[, KZ], KTAU) in Fortran order, the
 * interior points 1:jx, 1:iy[, 1:kz], no attributes, synchronised after
 * every record.
 *
 * mlc_nc_* is that file format (netCDF-3 classic, big-endian; the 64-bit
 * offset variant when one record exceeds the classic format's 2 GiB
 * limits).  mlc_output_* writes ps.nc and t.nc from a handle without
 * stopping the step loop: the interior is packed on the handle's stream,
 * copied to pinned host memory on a second stream and written by a host
 * thread.
 */
#ifndef MOLOCH_OUTPUT_H
#define MOLOCH_OUTPUT_H

#include "moloch_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlc_ncfile mlc_ncfile;

/* Create (clobber) `path` holding one double variable `varname` over JX =
 * nx, IY = ny, KZ = nz (nz = 0: no KZ dimension) and the unlimited KTAU
 * (grid_nc_create_var2d / _var3d, nompi/mod_mppparam.F90:121-188,
 * :230-301). */
int mlc_nc_create(const char *path, const char *varname, int nx, int ny, int nz, mlc_ncfile **out);
/* Append one record: values[(k*ny + i)*nx + j] for j < nx, i < ny, k < max(nz,1)
 * (grid_nc_write_var2d / _var3d, :190-212, :303-330), then flush. */
int mlc_nc_append(mlc_ncfile *f, const double *values);
/* Records written so far. */
int mlc_nc_records(const mlc_ncfile *f);
/* Close the file (grid_nc_destroy, :214-226). */
int mlc_nc_close(mlc_ncfile *f);

typedef struct mlc_output mlc_output;

/* Create `dir`/ps.nc and `dir`/t.nc for the handle's grid. */
int mlc_output_open(mlc_handle *h, const char *dir, mlc_output **out);
/* Queue one record of ps and t as they are after the work queued so far on
 * the handle; returns before the data reach the files.  Steps queued
 * afterwards do not change the record. */
int mlc_output_write(mlc_output *o);
/* Wait until every queued record is in the files. */
int mlc_output_flush(mlc_output *o);
/* Flush, close both files and release the buffers. */
int mlc_output_close(mlc_output *o);

#ifdef __cplusplus
}
#endif
#endif
