/*
 * spexcor -- the spex.cor checkpoint file of SPEX (C ABI, host side).
 *
 * Replaces writecor / readcor (correlation.f:1636-1952) with the file
 * layout of COR version 1 (version 0 files are read too) and the HDF5
 * wrapper's conventions (Hwrapper.f, including the division of datasets
 * above its chunk limit):
 * one group per k point named by the k vector in F12.10 format, the
 * matrices of each frequency as datasets "matrix_<i>" (packed upper) and
 * "matrixc_<i>" (full), complex arrays as double datasets with a leading
 * real/imaginary dimension.  Library libspexcor.so, kept apart from
 * libspexhip.so (it links HDF5; no device code).
 *
 * Arrays are Fortran-ordered host arrays, complex interleaved:
 *   coulomb  nbasm(nbasm+1)/2 complex (packed)   coul   nbas doubles
 *   ctrafo   nbasm x nbas complex                 cblock ncblock ints
 *   matrix   nbas(nbas+1)/2 x nfreq complex       matrixc nbas x nbas x nfreqc
 *   head     6 x nfreq,  wing 3 x nbas x nfreq    (Gamma point only)
 *   headc    3 x 3 x nfreqc, wingc 3 x nbas x 2 x nfreqc
 *   freq     nfreq doubles,  freqc nfreqc complex
 * Return 0 on success; otherwise spexcor_last_error() holds the message.
 */
#ifndef SPEXCOR_H
#define SPEXCOR_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

int spexcor_last_error(char *buf, size_t len);

/* writecor(typ) for typ = 1 (susceptibility), 2 (dielectric), 3 (inverse
 * dielectric); gamma != 0 writes plasma, head and wings (ikpt = 1).  Like
 * the reference it refuses to overwrite the matrices of a k point that is
 * already in the file, and checks nbasm and the frequency counts of an
 * existing file.                                                          */
int spexcor_write(const char *fname, const double kvec[3], int typ, const char *label, int nbasm,
                  const double *coulomb, int nbas, const double *coul, const double *ctrafo,
                  int ncblock, const int *cblock, int oselfc, const double contour[4],
                  int gamma, double plasma, int nfreq, const double *freq, const double *matrix,
                  const double *head, const double *wing, int nfreqc, const double *freqc,
                  const double *matrixc, const double *headc, const double *wingc);

/* Sizes for readcor: *typ = -1 if the k point is not in the file, 0 if it
 * holds only the Coulomb matrix, else the matrix type.                    */
int spexcor_query(const char *fname, const double kvec[3], int *typ, int *nbasm, int *nbas,
                  int *nfreq, int *nfreqc, int *ncblock);

/* readcor for a k point with matrices; any output pointer may be NULL.
 * label receives 6 characters (not terminated).                           */
int spexcor_read(const char *fname, const double kvec[3], char *label, double *coulomb,
                 double *coul, double *ctrafo, int *cblock, int *oselfc, double contour[4],
                 double *plasma, double *freq, double *matrix, double *head, double *wing,
                 double *freqc, double *matrixc, double *headc, double *wingc);

/* Layout of the arrays SPEX declares MCOMPLEX_dp (coulomb, ctrafo, matrix,
 * head, wing): complex (inv = 0, the default) or, as an inversion-symmetric
 * build (-DINV, cppmacro.h:50-60) stores them, real_dp (inv != 0); then
 * those arguments of spexcor_write/spexcor_read are real arrays of the
 * same element counts.  matrixc, headc, wingc and freqc stay complex.
 * spexcor_read refuses a file of the other layout.                        */
int spexcor_set_inv(int inv);

/* The layout of the k point's group in fname: *inv = 1 for the real
 * (INV build) layout, 0 for the complex one.                              */
int spexcor_file_inv(const char *fname, const double kvec[3], int *inv);

/* The chunk limit of the HDF5 wrapper (MaxChunk, cppmacro.h:143-144) in
 * bytes, 0 for the default 2^31-1: arrays of more than limit/(element
 * bytes) elements are stored as "<name>_0" and "<name>_1", divided again
 * until each part fits (Hwrapper.f:413-426).  Reads follow the same rule,
 * so a file must be read with the limit it was written with.             */
int spexcor_set_max_chunk(size_t bytes);

/* The name writecor writes to: prefix.cor, or prefix.cor.<color> for an
 * MPI color > 0 (correlation.f:1828).                                     */
int spexcor_file_name(const char *prefix, int color, char *fname, size_t len);

/* readcor's choice of file for k point kvec (correlation.f:1655-1672): the
 * file prefix.cor.<n> named by prefix.cor.map/<k point> when the map
 * directory holds "present", else prefix.cor.                             */
int spexcor_locate(const char *prefix, const double kvec[3], char *fname, size_t len);

/* writecor's map bookkeeping after a write (correlation.f:1911-1946): for
 * color > 0 the map directory, its "present" file and the entry holding
 * color are created; for color 0 an existing entry is removed.            */
int spexcor_map_record(const char *prefix, const double kvec[3], int color);

#ifdef __cplusplus
}
#endif
#endif
