// TIFF output of volumes and projection stacks, as the reference writes its
// dumps with vtkTIFFWriter (KevLoq/Kevernals source/modules/reconstruction/
// Reconstructors.h:49-52,78-81,98-101,127-130 and the MLEM equivalents
// :166-169,195-198,218-221,246-249; mainReconstruct.cpp:127-148).
//
// The file layout is that of VTK 9.0.1's vtkTIFFWriter driving libtiff 4.0.x
// (neither is part of the reference):
//  * one z slice: rows from the highest y down with Orientation top-left;
//    more slices: one page per slice, rows in memory order, SubfileType page
//    and PageNumber (page, pages);
//  * PackBits compression (vtkTIFFWriter's default), one PackBits stream per
//    row; strips of max(1, 8192 / row bytes) rows;
//  * X/YResolution 10 / spacing in pixels per centimetre;
//  * each directory at the even-rounded end of the file, entries sorted by tag,
//    out-of-line values after it in libtiff's emission order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

#include "vtk_subset.h"

namespace kvtiff {

// VTK scalar type codes accepted by vtkTIFFWriter.
enum ScalarType { kUnsignedChar = 3, kUnsignedShort = 5, kFloat = 10 };
enum Compression { kNoCompression = 0, kPackBits = 1 };

// PackBits encoding of one row (libtiff's PackBitsEncode, tif_packbits.c):
// out must hold PackBitsBound(n) bytes; returns the encoded length.
size_t PackBitsRow(const uint8_t *row, size_t n, uint8_t *out);
inline size_t PackBitsBound(size_t n) { return 2 * n + 2; }

// Writes data[z][y][x][c] (VTK memory order) to path.  Returns false with a
// message in *err on unsupported input or an I/O failure.
bool WriteImage(const std::string &path, const void *data, int scalar_type, int components, int nx, int ny, int nz,
                double spacing_x, double spacing_y, int compression, std::string *err);

}  // namespace kvtiff

#ifdef KV_USE_REAL_VTK
#include <vtkTIFFWriter.h>
#else
// The part of vtkTIFFWriter the reference calls: New, SetFileName,
// SetInputData, Write (and the compression setters and error code).
class vtkTIFFWriter {
public:
    enum { NoCompression = 0, PackBits = 1 };
    static vtkTIFFWriter *New() { return new vtkTIFFWriter(); }
    void SetFileName(const char *name) { file_ = name ? name : ""; }
    const char *GetFileName() const { return file_.c_str(); }
    void SetInputData(vtkImageData *img) { input_ = img; }
    void SetCompression(int c) { compression_ = c; }
    void SetCompressionToNoCompression() { compression_ = NoCompression; }
    void SetCompressionToPackBits() { compression_ = PackBits; }
    // Single-component float images (the vtkImageData subset's scalars).
    void Write();
    // 0 after a successful Write, nonzero otherwise (message in GetErrorMessage).
    unsigned long GetErrorCode() const { return error_; }
    const char *GetErrorMessage() const { return message_.c_str(); }

private:
    std::string file_;
    vtkImageData *input_ = nullptr;
    int compression_ = PackBits;
    unsigned long error_ = 0;
    std::string message_;
};
#endif
