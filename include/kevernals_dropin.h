/* C entry points of libkevernals_dropin.so: the data formats on either side of
 * the projection path (geometry files, DICOM projections, TIFF dumps), for
 * callers that bind C (ctypes, cgo, JNI).  Host code only; no GPU needed.
 *
 * The C++ classes of the same library (TomoGeometry, DICOMReader,
 * vtkTIFFWriter, Projector, recons::) are the drop-in for the reference's
 * C++ interfaces; these functions expose the same readers and writer.
 * Return conventions: 0 = success for int results, NULL = failure for
 * pointers; a message goes into err (err_len bytes) when err is not NULL. */
#ifndef KEVERNALS_DROPIN_H
#define KEVERNALS_DROPIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel arguments of an XML geometry file, as TomoGeometry(xmlPath)
 * (TomoGeometry.cpp:13-437) marshalled by Projector.cu:876-940:
 * ints = {valid, nx, ny, nz, nu, nv, np}, floats = {sx, sy, sz, pu, pv,
 * fulcrum x, y, z}, det/src = 3 floats per view for the first min(np, cap)
 * views.  Returns np, or 0 when the file is not a valid geometry. */
int kvd_geometry_from_xml(const char *path, int32_t *ints, float *floats, float *det, float *src, int cap,
                          char *err, int err_len);

/* The detector region of an XML geometry file: roi = {x, y, width, height} in
 * detector pixels (TomoGeometry::GetProjectionROIsBLPixelPositionOnDetector
 * and projectionsRoisSize, TomoGeometry.cpp:415-434, 562-565), the region
 * DICOMReader crops to.  Returns 0 on success. */
int kvd_geometry_roi_from_xml(const char *path, int32_t *roi, char *err, int err_len);

/* TIFF output as vtkTIFFWriter writes the reference's dumps
 * (Reconstructors.h:49-52 and the other verbose dumps): data is
 * [nz][ny][nx][components] in VTK memory order; scalar_type 3 (unsigned
 * char), 5 (unsigned short) or 10 (float); spacing sets the resolution tags
 * (10 / spacing pixels per cm); compression 1 = PackBits (the writer's
 * default), 0 = none.  More than one z slice writes one page per slice (one
 * component only).  Returns 0 on success. */
int kvd_write_tiff(const char *path, const void *data, int scalar_type, int components, int nx, int ny, int nz,
                   double spacing_x, double spacing_y, int compression, char *err, int err_len);

/* DICOM projections (DICOMReader::Read / ReadDirectory, DICOMReader.cpp:
 * 26-86, 89-202): directory = 0 reads one file, 1 every .dcm/.DCM file of a
 * directory ordered by acquisition time with the indices remove[0..nremove)
 * of that order dropped; roi = {x, y, width, height} crops each image to the
 * detector region (NULL: whole image).  The result is a float image, rows
 * bottom-up, frames and files along z; free it with kvd_image_free. */
typedef struct kvd_image kvd_image;
kvd_image *kvd_dicom_read(const char *path, int directory, const int32_t *remove, int nremove, const int32_t *roi,
                          char *err, int err_len);
/* extent = {x0, x1, y0, y1, z0, z1} (inclusive, as vtkImageData), spacing =
 * 3 doubles.  Returns 0 on success. */
int kvd_image_info(const kvd_image *img, int32_t *extent, double *spacing);
/* [z][y][x] floats of the extent above. */
const float *kvd_image_data(const kvd_image *img);
void kvd_image_free(kvd_image *img);

#ifdef __cplusplus
}
#endif

#endif
