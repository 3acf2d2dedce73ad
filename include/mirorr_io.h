/*
 * mirorr_io.h -- ITK-free image and transform file I/O of the command-line
 * drop-in (no GPU involved). Replaces what the reference's command line does
 * through ITK's readers and writers:
 *   - NIfTI-1 volumes (.nii, .nii.gz) read as float with ITK's LPS geometry
 *     (itkMirorrPyramidWrapper.txx:37-58 ReadImage, ITK NiftiImageIO);
 *   - the reorientation of every input to ARI (itkMirorrPyramidWrapper.txx:
 *     79-92, 208-211, ITK OrientImageFilter);
 *   - the centred initial transform (itkMirorrPyramidWrapper.txx:611-623,
 *     CenteredTransformInitializer with GeometryOn);
 *   - ITK transform files, "#Insight Transform File V1.0"
 *     (itkMirorrPyramidWrapper.txx:353-448, TransformFileReader/Writer).
 *
 * Geometry uses mirorr_geom of mirorr_gpu.h. Buffers returned by the readers
 * are allocated with malloc and released with mirorr_io_free. Every entry
 * point returns a mirorr_status; mirorr_io_last_error() describes the last
 * failure on the calling thread.
 */
#ifndef MIRORR_IO_H
#define MIRORR_IO_H

#include "mirorr_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed I/O call on this thread ("" if none). */
const char* mirorr_io_last_error(void);

void mirorr_io_free(void* p);

/* Reads a 3-D NIfTI-1 file (.nii or .nii.gz; uint8, int8, int16, uint16,
 * int32, uint32, float32 or float64 voxels, scaled by scl_slope/scl_inter
 * when the slope is non-zero) into a float volume. The geometry follows the
 * qform when qform_code > 0, else the sform when sform_code > 0, else the
 * voxel sizes alone; NIfTI's RAS coordinates become ITK's LPS (x and y
 * negated). *data receives dims[0]*dims[1]*dims[2] floats, x fastest. */
mirorr_status mirorr_nifti_read(const char* path, mirorr_geom* geom, float** data);

/* As mirorr_nifti_read for a mask, converted as ITK converts voxels into an
 * unsigned char image: fractions truncated, integers keep their low 8 bits. */
mirorr_status mirorr_nifti_read_mask(const char* path, mirorr_geom* geom, uint8_t** data);

/* Writes a float32 NIfTI-1 file (gzip-compressed when the name ends in
 * ".gz") with qform and sform both set (code 1) from the LPS geometry. */
mirorr_status mirorr_nifti_write(const char* path, const mirorr_geom* geom, const float* data);

/* Reorients a volume (and optionally its mask, same grid) in place to ITK's
 * ARI orientation: axis 0 runs anterior to posterior, axis 1 right to left,
 * axis 2 inferior to superior. Each axis goes to the LPS axis its direction
 * cosine is closest to; axes are permuted and flipped, and the origin and
 * direction change so that every voxel keeps its physical position. */
mirorr_status mirorr_reorient_ari(mirorr_geom* geom, float* data, uint8_t* mask);

/* Centred initial transform: rotation centre = the geometric centre of the
 * `centre_image` grid, translation = centre of `other` minus that centre
 * (CenteredTransformInitializer::GeometryOn with the reference's argument
 * order: the moving image as its fixed image). params gets 6 (rigid: zero
 * angles) or 12 (affine: identity matrix) values; center gets 3. */
mirorr_status mirorr_centered_init(const mirorr_geom* centre_image, const mirorr_geom* other,
                                   int affine, double* params, double center[3]);

/* ITK transform file, version 1.0. Writes one Euler3DTransform_double_3_3
 * (affine = 0: angles x, y, z and translation) or AffineTransform_double_3_3
 * (affine = 1: row-major matrix and translation), the centre as fixed
 * parameters. inverse = 1 writes the reference's inverse instead
 * (itkMirorrPyramidWrapper.txx:312-350 GetInverseTransform): the exact
 * inverse about the same centre with its matrix replaced by the nearest
 * rotation, for affine transforms too. Numbers get 17 significant digits. */
mirorr_status mirorr_tfm_write(const char* path, int affine, const double* params,
                               const double center[3], int inverse);

/* Reads the first transform of an ITK V1.0 transform file
 * (itkMirorrPyramidWrapper.txx:380-448). Euler3DTransform and
 * AffineTransform are accepted (double or float, 3-D); any other class fails
 * with MIRORR_ERR_ARG. An Euler matrix gets its third row rebuilt as the
 * cross product of the first two, as the reference does. The result is
 * returned as `want_affine` asks: 12 affine parameters, or 6 rigid ones (the
 * matrix must then be orthogonal within 1e-10, else MIRORR_ERR_ALGO, which
 * the reference reports as an unreadable file). inverse = 1 returns the
 * inverse as mirorr_tfm_write makes it. */
mirorr_status mirorr_tfm_read(const char* path, int want_affine, double* params,
                              double center[3], int inverse);

#ifdef __cplusplus
}
#endif

#endif /* MIRORR_IO_H */
