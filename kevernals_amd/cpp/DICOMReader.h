// Drop-in replacement of KevLoq/Kevernals
// source/modules/dataHandling/DICOMReader.h: projection images from DICOM
// files, cropped to the geometry's detector region.
//
// The reference reads files with vtk-dicom's vtkDICOMReader (VTK 9.0.1 module
// DICOM); this reader restates what that reader hands over for the files the
// projector consumes:
//  * DICOM Part 10 files (or bare datasets) in the implicit VR little endian,
//    explicit VR little endian, deflated explicit VR little endian and
//    explicit VR big endian transfer syntaxes; encapsulated (compressed)
//    pixel data is rejected;
//  * one sample per pixel, 8, 16 or 32 bits allocated, stored values (no
//    rescale applied) taken from the BitsStored bits ending at HighBit,
//    sign-extended when PixelRepresentation is 1;
//  * frames become z slices in file order, rows are stored bottom-up (the
//    first row of a frame lands at the highest y, vtkDICOMReader's default
//    memory row order), spacing (column, row, 1) from PixelSpacing, else
//    ImagerPixelSpacing, else 1.
// Output scalars are float: the values the reference's vtkImageCast to float
// produces from the reader's output (mainReconstruct.cpp:113-116).
#pragma once

#include <cstdint>
#include <string>
#include <system_error>
#include <vector>

#include "Reconstructors.h"
#include "TomoGeometry.h"

// DICOMReaderErrorCode.h:19-22 (the reference numbers it from 0)
enum class DICOMReaderErrorCode { DataFileReadingProblem };
namespace std {
template <>
struct is_error_code_enum<DICOMReaderErrorCode> : true_type {};
}  // namespace std
std::error_code make_error_code(DICOMReaderErrorCode e);

namespace kvdicom {

// One decoded file: frames x rows x columns stored values in file order.
struct Frames {
    int rows = 0, columns = 0, frames = 0;
    double spacing[2] = {1.0, 1.0};  // column (x), row (y)
    bool hasPatientID = false;
    std::string acquisitionTime;  // AcquisitionTime, trailing padding removed
    std::vector<float> values;    // [frame][row][column], first row of the file first
};

// Whether the file looks like DICOM (Part 10 preamble, or a bare dataset
// starting with a group 0002 or 0008 element).
bool CanReadFile(const std::string &path);
bool ReadFile(const std::string &path, Frames *out, std::string *err);

// An image of the drop-in's vtkImageData subset: x = column, y = bottom-up
// row, z = frame, cropped to x in [x0, x0 + w - 1], y in [y0, y0 + h - 1]
// (vtkExtractVOI, clipped to the image; extent kept).  Empty crop -> nullptr.
ImageDataPtr CropToImage(const Frames &f, int x0, int y0, int w, int h, std::string *err);

// One file cropped to x in [x0, x0 + w - 1], y in [y0, y0 + h - 1] (the
// steps of DICOMReader::Read), decoded straight into the image.
bool ReadImage(const std::string &path, int x0, int y0, int w, int h, ImageDataPtr *out, std::string *err);

// The directory steps of DICOMReader::ReadDirectory with an explicit region.
bool ReadDirectoryImages(const std::string &dir, const std::vector<int> &indicesToRemove, int x0, int y0, int w,
                         int h, bool verbose, ImageDataPtr *out, std::string *err);

}  // namespace kvdicom

class DICOMReader {
public:
    DICOMReader() = delete;
    explicit DICOMReader(TomoGeometry *p_tomoGeometry) : m_tomoGeometry(p_tomoGeometry) {}

    // DICOMReader.cpp:26-86: one file, cropped to the detector region.
    Result<ImageDataPtr> Read(std::string const &p_dicomFilePath) const;
    // DICOMReader.cpp:89-202: every .dcm / .DCM file of a directory, ordered
    // by AcquisitionTime (read only from files that carry a PatientID, the
    // reference's test), the given indices of that order removed, each
    // cropped and appended along z.  Files of equal time keep file-name
    // order (the reference leaves it to the directory and an unstable sort).
    Result<ImageDataPtr> ReadDirectory(std::string const &p_dicomFilesContainedDirPath,
                                       std::vector<int> const &p_indicesToRemove) const;

private:
    TomoGeometry *m_tomoGeometry;
};
