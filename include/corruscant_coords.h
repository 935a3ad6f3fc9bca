/*
 * corruscant_coords.h -- C ABI of libcoords.so (src/coordmath.c:38-80).
 *
 * Same names, argument order and output layout as the reference: parallel
 * input arrays in degrees, caller-allocated outputs, n points.  The
 * conversions run on the GPU in FP64 with correctly rounded sin and cos;
 * the reference calls the host libm, which (glibc 2.35) differs from the
 * correctly rounded value by one ulp for about 0.3% of arguments.
 */
#ifndef CORRUSCANT_COORDS_H
#define CORRUSCANT_COORDS_H

#ifdef __cplusplus
extern "C" {
#endif

/* src/coordmath.c:38-51: RA/DEC (degrees) -> unit-sphere xyz, (n,3) row-major */
void radec2cart64(double *ra, double *dec, double *xyz, int n);

/* src/coordmath.c:58-71: RA/DEC/distance -> xyz scaled by distance */
void radecdist2cart64(double *ra, double *dec, double *dist, double *xyz, int n);

/* src/coordmath.c:73-80: theta = (90 - dec) * pi/180, phi = ra * pi/180 */
void radec2sph64(double *ra, double *dec, double *theta, double *phi, int n);

/* HEALPix RING-scheme pixel index of each (theta, phi) in radians, the
 * pixelisation corruscant/fields.py:67 and :153 take from healpy's ang2pix
 * (no counterpart in the reference's C).  nside in [1, 2^29]; points with
 * theta outside [0, pi] or a non-finite phi get -1 and an error message. */
void ang2pix_ring64(const double *theta, const double *phi, long long *pix, int n, long long nside);

/* Message of the last failure on this thread ("" if none); the conversions
 * above return nothing, like the reference's, so errors are reported here. */
const char *coords_last_error(void);

/* Points of the last radec2cart64 / radecdist2cart64 call whose sin or cos
 * could not be certified correctly rounded (value within 2^-98 of a rounding
 * boundary, or |angle| >= 2^20 rad); 0 in practice. */
unsigned long long coords_last_uncertain(void);

#ifdef __cplusplus
}
#endif

#endif /* CORRUSCANT_COORDS_H */
