/*
 * particles_amd_fits.h -- FITS primary-image I/O of the host driver.
 *
 * Replaces the reference's cfitsio calls (gkowal/particles
 * sources/fitsio.F90): get_dimensions (:173-196) and read_field (:208-269)
 * for the six field files, write_data (:281-540) for counter.fits (int64),
 * time.fits and state.fits (FP64, with the TUNIT / TEMP header keys).  Plain
 * C types; paths may carry cfitsio's leading '!' (overwrite), which is
 * accepted and ignored (files are always replaced); .fits.gz files are read
 * transparently.  Every function returns 0 or nonzero with a message for
 * ptc_fits_last_error().  Host-only: no GPU is touched.
 */
#ifndef PARTICLES_AMD_FITS_H
#define PARTICLES_AMD_FITS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* NAXIS, NAXIS1..8 and BITPIX of a file's primary image
 * (sources/fitsio.F90:173-196 reads NAXIS1..3 of the x-velocity file). */
int ptc_fits_dims(const char *path, int *naxis, int64_t dims[8], int *bitpix);

/* All n values of the primary image as FP64 (BSCALE/BZERO applied), any
 * BITPIX (8, 16, 32, 64, -32, -64); n must equal the image size
 * (sources/fitsio.F90:208-269, which reads FP64 files through the FP32 call;
 * this reads them as FP64). */
int ptc_fits_read(const char *path, int64_t n, double *out);

/* FP64 image (BITPIX -64) with optional extra real keys, e.g. TUNIT / TEMP
 * (sources/fitsio.F90:368-540). */
int ptc_fits_write_f64(const char *path, int naxis, const int64_t *dims, const double *data,
                       int nkeys, const char *const *keys, const double *values,
                       const char *const *comments);

/* int64 image (BITPIX 64), sources/fitsio.F90:281-356 (counter.fits). */
int ptc_fits_write_i64(const char *path, int naxis, const int64_t *dims, const int64_t *data);

const char *ptc_fits_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
