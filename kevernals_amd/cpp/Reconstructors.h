// Drop-in replacements of KevLoq/Kevernals
// source/modules/reconstruction/Projector.h and Reconstructors.h, running on
// libkevernals_amd (gfx950) through its C ABI.
//
// Signatures are those of the reference (Projector.h:10-28,
// Reconstructors.h:19-24,136-141,256-257,263-265); Result<T> carries a value or
// a std::error_code as the reference's outcome-based Result (Result.h:25-37).
// Differences from the reference, by design:
//  * no process exit on device errors: failures come back as error codes
//    (the reference called exit() from helper_cuda);
//  * ART/MLEM keep volume, data and work buffers on the device for all
//    iterations (the reference copied both ways every projection);
//  * with p_outputDirectoryPath, ART and MLEM run their steps apart on the
//    device to write the reference's TIFF dumps (same values as the fused
//    iterations).
#pragma once

//
// Inside the reference source tree, define KV_USE_REFERENCE_TREE (and
// KV_USE_REAL_VTK): the reference's own TomoGeometry, Result and error-code
// headers are used and only Projector / recons:: come from this file.
#include <optional>
#include <string>
#include <system_error>
#include <utility>
#include <vector>

#include "vtk_subset.h"

struct kv_context;

#ifdef KV_USE_REFERENCE_TREE
#include "commons/Result.h"
#include "modules/geometry/TomoGeometry.h"
#include "modules/reconstruction/ProjectorErrorCode.h"
#include "modules/reconstruction/ReconstructorsErrorCode.h"
#else
#include "TomoGeometry.h"

// ---- error codes (ReconstructorsErrorCode.h:19-25, ProjectorErrorCode.h:19-27)
enum class ReconstructorsErrorCode { ART = 1, BackProjectionFailed, ProjectionFailed, MLEM, ShiftAndAdd };
enum class ProjectorErrorCode {
    InvalidGeometryError = 1,
    GeometryInconsistency,
    VolumeProblemForProjection,
    ProjectionsDataFileReadingProblem,
    RadonMatrixComputation,
    BadAlgorithm,
};
namespace std {
template <>
struct is_error_code_enum<ReconstructorsErrorCode> : true_type {};
template <>
struct is_error_code_enum<ProjectorErrorCode> : true_type {};
}  // namespace std
std::error_code make_error_code(ReconstructorsErrorCode e);
std::error_code make_error_code(ProjectorErrorCode e);

// ---- Result<T>
template <class T>
class Result {
public:
    Result(T v) : value_(std::move(v)) {}
    Result(std::error_code e) : error_(e) {}
    template <class E, class = std::enable_if_t<std::is_error_code_enum<E>::value>>
    Result(E e) : error_(make_error_code(e)) {}
    bool has_value() const { return value_.has_value(); }
    bool has_error() const { return !value_.has_value(); }
    explicit operator bool() const { return has_value(); }
    T &value() { return *value_; }
    const T &value() const { return *value_; }
    std::error_code error() const { return error_; }

private:
    std::optional<T> value_;
    std::error_code error_;
};
#endif  // KV_USE_REFERENCE_TREE

// ---- Projector (Projector.h:10-28)
class Projector {
public:
    Projector(TomoGeometry *p_tomoGeometry);
    ~Projector();
    Projector(const Projector &) = delete;
    Projector &operator=(const Projector &) = delete;

    vtkSmartPointer<vtkImageData> PerformProjection(vtkSmartPointer<vtkImageData> p_volume) const;
    // Single-source projection into a caller container (unfinished in the
    // reference, Projector.cu:823-870): the volume is projected from
    // p_sourcePosition onto the container's views, whose dimensions give the
    // detector size and whose spacing gives the pixel pitch; view p's
    // detector origin is the geometry's region corner of view p at the
    // common detector height (Projector.cu:853-858).
    void PerformProjection(const vtkSmartPointer<vtkImageData> p_volume, const Position3D &p_sourcePosition,
                           vtkSmartPointer<vtkImageData> p_projectionsContainer) const;
    vtkSmartPointer<vtkImageData> PerformBackProjection(vtkSmartPointer<vtkImageData> p_projections) const;

    // Status of the last call (empty on success).
    std::error_code lastError() const { return lastError_; }
    std::string lastMessage() const { return lastMessage_; }
    // Device context (created on first use); for the reconstruction loops.
    kv_context *context() const;

private:
    bool ensureContext() const;
    TomoGeometry *m_tomoGeometry;
    mutable kv_context *ctx_ = nullptr;
    mutable std::error_code lastError_;
    mutable std::string lastMessage_;
};

// ---- recons:: (Reconstructors.h:17-381)
namespace recons {
Result<ImageDataPtr> ART(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages, int p_iterationNumber,
                         float p_relaxationCoefficient, std::optional<ImageDataPtr> p_initialVolume,
                         std::optional<std::string> p_outputDirectoryPath);
Result<ImageDataPtr> MLEM(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages, int p_iterationNumber,
                          float p_relaxationCoefficient, std::optional<ImageDataPtr> p_initialVolume,
                          std::optional<std::string> p_outputDirectoryPath);
Result<ImageDataPtr> BackProjection(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages);
Result<ImageDataPtr> ShiftAndAdd(TomoGeometry *p_tomoGeometry, ImageDataPtr p_projectionImages,
                                 std::optional<std::string> p_outputDirectoryPath);

// Accelerated MLEM, a sibling of MLEM with no counterpart in the reference:
// MLEM on line-integral data in its matched (mode-A) form,
// x <- x / A^T 1 * A^T(data / Ax), with the two-stretch trajectory-jump
// schedule (kv_mlem_accelerated): `first` iterations recorded from `start`
// on, a probe jump to iteration `mid` followed by `second` iterations, and a
// final jump from iteration `first` to an estimate of iteration `horizon`,
// kept only if it does not lower the Poisson log-likelihood.  With CNN
// parameters (KV_CNN_NPARAMS floats, input scale p_cnnScale) the CNN step on
// the jump's result is kept only if it does not lower the jump's
// log-likelihood.  As MLEM: the initial volume defaults to uniform 1
// (Reconstructors.h:152-153) and is otherwise updated in place and returned;
// failures return ReconstructorsErrorCode::MLEM.  p_stepsKept (optional)
// receives 0 (iteration `first` kept), 1 (the jump) or 2 (jump and CNN step).
struct AcceleratedSchedule {
    int start = 2, first = 10, mid = 40, second = 8, horizon = 100;
};
Result<ImageDataPtr> MLEMAccelerated(TomoGeometry *p_tomoGeometry, ImageDataPtr p_lineIntegrals,
                                     std::optional<ImageDataPtr> p_initialVolume,
                                     const std::vector<float> *p_cnnParameters = nullptr, float p_cnnScale = 20.f,
                                     AcceleratedSchedule p_schedule = {}, int *p_stepsKept = nullptr);
}  // namespace recons

#ifndef KV_USE_REFERENCE_TREE
// ---- PhantomMaker subset (PhantomMaker.h:39-70): the uniform volume used as
// the default initial volume of ART/MLEM (Reconstructors.h:35-36,152-153).
class PhantomMaker {
public:
    PhantomMaker(TomoVolume *p_volume, float p_backgroundDensity = 50.f)
        : m_volume(p_volume), m_backgroundDensity(p_backgroundDensity)
    {
    }
    ImageDataPtr GetPhantom() const;

private:
    TomoVolume *m_volume;
    float m_backgroundDensity;
};
#endif  // KV_USE_REFERENCE_TREE
