// The subset of VTK's image API that the projector and reconstruction
// interfaces use (KevLoq/Kevernals source/modules/reconstruction/Projector.h:5-6,
// Reconstructors.h:56-60, Projector.cu:965-969): vtkSmartPointer<vtkImageData>,
// New, Set/GetDimensions, Set/GetSpacing, Set/GetOrigin, Set/GetExtent,
// AllocateScalars(VTK_FLOAT, 1), GetScalarPointer.
//
// Building against a real VTK 9 installation: define KV_USE_REAL_VTK and the
// real headers are used instead; nothing else changes.
#pragma once

#ifdef KV_USE_REAL_VTK
#include <vtkImageData.h>
#include <vtkSmartPointer.h>
#else

#include <cstddef>
#include <memory>
#include <vector>

#ifndef VTK_FLOAT
#define VTK_FLOAT 10
#endif

class vtkImageData {
public:
    static vtkImageData *New() { return new vtkImageData(); }

    void SetDimensions(int x, int y, int z)
    {
        SetExtent(0, x - 1, 0, y - 1, 0, z - 1);
    }
    int *GetDimensions()
    {
        for (int i = 0; i < 3; ++i)
            dims_[i] = extent_[2 * i + 1] - extent_[2 * i] + 1;
        return dims_;
    }
    void GetDimensions(int d[3])
    {
        const int *s = GetDimensions();
        d[0] = s[0];
        d[1] = s[1];
        d[2] = s[2];
    }
    void SetExtent(int x0, int x1, int y0, int y1, int z0, int z1)
    {
        const int e[6] = {x0, x1, y0, y1, z0, z1};
        for (int i = 0; i < 6; ++i)
            extent_[i] = e[i];
    }
    int *GetExtent() { return extent_; }
    void SetSpacing(double x, double y, double z)
    {
        spacing_[0] = x;
        spacing_[1] = y;
        spacing_[2] = z;
    }
    double *GetSpacing() { return spacing_; }
    void GetSpacing(double s[3])
    {
        for (int i = 0; i < 3; ++i)
            s[i] = spacing_[i];
    }
    void SetOrigin(double x, double y, double z)
    {
        origin_[0] = x;
        origin_[1] = y;
        origin_[2] = z;
    }
    double *GetOrigin() { return origin_; }
    void GetOrigin(double o[3])
    {
        for (int i = 0; i < 3; ++i)
            o[i] = origin_[i];
    }
    // Only single-component float scalars are supported (VTK_FLOAT, 1).  As
    // in VTK, the values are left uninitialised.
    void AllocateScalars(int type, int components)
    {
        (void)type;
        (void)components;
        const int *d = GetDimensions();
        n_ = (size_t)d[0] * (size_t)d[1] * (size_t)d[2];
        data_.reset(n_ ? new float[n_] : nullptr);
    }
    void *GetScalarPointer() { return n_ ? data_.get() : nullptr; }
    void *GetScalarPointer(int x, int y, int z)
    {
        const int *d = GetDimensions();
        return data_.get() + ((size_t)(z - extent_[4]) * d[1] + (size_t)(y - extent_[2])) * d[0] +
               (size_t)(x - extent_[0]);
    }
    size_t GetNumberOfPoints()
    {
        const int *d = GetDimensions();
        return (size_t)d[0] * d[1] * d[2];
    }
    void Modified() {}

private:
    int extent_[6] = {0, -1, 0, -1, 0, -1};
    int dims_[3] = {0, 0, 0};
    double spacing_[3] = {1.0, 1.0, 1.0};
    double origin_[3] = {0.0, 0.0, 0.0};
    std::unique_ptr<float[]> data_;
    size_t n_ = 0;
};

template <class T>
class vtkSmartPointer {
public:
    vtkSmartPointer() = default;
    vtkSmartPointer(std::nullptr_t) {}
    static vtkSmartPointer New()
    {
        vtkSmartPointer p;
        p.p_.reset(T::New());
        return p;
    }
    T *operator->() const { return p_.get(); }
    T &operator*() const { return *p_; }
    T *Get() const { return p_.get(); }
    T *GetPointer() const { return p_.get(); }
    operator T *() const { return p_.get(); }
    bool operator==(std::nullptr_t) const { return !p_; }
    bool operator!=(std::nullptr_t) const { return (bool)p_; }

private:
    std::shared_ptr<T> p_;
};
#endif

using ImageDataPtr = vtkSmartPointer<vtkImageData>;  // commons/ImageDataPtr.h:5
