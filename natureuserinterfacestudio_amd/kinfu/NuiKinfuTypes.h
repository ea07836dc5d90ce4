// Value types used by the tracker / frame / scene interfaces.
//
// The reference's signatures use Eigen types (Matrix3frm = row-major float
// 3x3, src/Shape/NuiCameraPos.h:6; Vector3f; Affine3f) and the app's
// NuiCameraParams / NuiCameraPos (src/Shape/NuiCameraParams.h,
// src/Shape/NuiCameraPos.h).  Eigen is not part of this build, so the few
// operations the interfaces need are provided here with the same memory
// layout (Matrix3frm::data() is 9 row-major floats).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

typedef unsigned int UINT;
typedef uint16_t UINT16;
typedef int64_t INT64;

// src/stdafx.h:58-65
struct BGRQUAD {
    uint8_t rgbBlue, rgbGreen, rgbRed, rgbReserved;
};

// The Kinect SDK's ColorSpacePoint: where a depth pixel lands in the colour image (pixels).
struct ColorSpacePoint {
    float X, Y;
};

// src/Shape/NuiImageBuffer.h:7-57: an image of width x height pixels of T.
template <typename T>
class NuiImageBuffer {
public:
    T* AllocateBuffer(UINT width, UINT height) {
        m_nWidth = width;
        m_nHeight = height;
        m_buf.assign((size_t)width * height, T{});
        return m_buf.data();
    }
    void Clear() {
        m_buf.clear();
        m_nWidth = m_nHeight = 0;
    }
    T* GetBuffer() const { return m_buf.empty() ? nullptr : const_cast<T*>(m_buf.data()); }
    UINT GetWidth() const { return m_nWidth; }
    UINT GetHeight() const { return m_nHeight; }

private:
    std::vector<T> m_buf;
    UINT m_nWidth = 0, m_nHeight = 0;
};
typedef NuiImageBuffer<BGRQUAD> NuiColorImage;

class Vector3f {
public:
    Vector3f() { v_[0] = v_[1] = v_[2] = 0.f; }
    Vector3f(float x, float y, float z) { v_[0] = x; v_[1] = y; v_[2] = z; }
    static Vector3f Zero() { return Vector3f(); }
    static Vector3f Constant(float c) { return Vector3f(c, c, c); }
    float& operator()(int i) { return v_[i]; }
    float operator()(int i) const { return v_[i]; }
    float& operator[](int i) { return v_[i]; }
    float operator[](int i) const { return v_[i]; }
    float* data() { return v_; }
    const float* data() const { return v_; }
    Vector3f operator+(const Vector3f& o) const { return Vector3f(v_[0] + o.v_[0], v_[1] + o.v_[1], v_[2] + o.v_[2]); }
    Vector3f operator-(const Vector3f& o) const { return Vector3f(v_[0] - o.v_[0], v_[1] - o.v_[1], v_[2] - o.v_[2]); }
    float norm() const { return std::sqrt(v_[0] * v_[0] + v_[1] * v_[1] + v_[2] * v_[2]); }

private:
    float v_[3];
};

class Vector3i {
public:
    Vector3i() { v_[0] = v_[1] = v_[2] = 0; }
    Vector3i(int x, int y, int z) { v_[0] = x; v_[1] = y; v_[2] = z; }
    static Vector3i Constant(int c) { return Vector3i(c, c, c); }
    int& operator()(int i) { return v_[i]; }
    int operator()(int i) const { return v_[i]; }
    int& operator[](int i) { return v_[i]; }
    int operator[](int i) const { return v_[i]; }
    int* data() { return v_; }

private:
    int v_[3];
};

// Row-major float 3x3 (Eigen::Matrix<float,3,3,RowMajor>)
class Matrix3frm {
public:
    Matrix3frm() { setIdentity(); }
    static Matrix3frm Identity() { return Matrix3frm(); }
    void setIdentity() {
        std::memset(m_, 0, sizeof(m_));
        m_[0] = m_[4] = m_[8] = 1.f;
    }
    float& operator()(int r, int c) { return m_[3 * r + c]; }
    float operator()(int r, int c) const { return m_[3 * r + c]; }
    float* data() { return m_; }
    const float* data() const { return m_; }
    Matrix3frm operator*(const Matrix3frm& b) const {
        Matrix3frm r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                r.m_[3 * i + j] = m_[3 * i] * b.m_[j] + m_[3 * i + 1] * b.m_[3 + j] + m_[3 * i + 2] * b.m_[6 + j];
        return r;
    }
    Vector3f operator*(const Vector3f& v) const {
        return Vector3f(m_[0] * v[0] + m_[1] * v[1] + m_[2] * v[2], m_[3] * v[0] + m_[4] * v[1] + m_[5] * v[2],
                        m_[6] * v[0] + m_[7] * v[1] + m_[8] * v[2]);
    }
    // cofactor inverse, as Eigen's fixed-size 3x3 inverse
    Matrix3frm inverse() const {
        const float* m = m_;
        Matrix3frm r;
        const float c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
        const float inv = 1.0f / (m[0] * c00 + m[1] * c01 + m[2] * c02);
        r.m_[0] = c00 * inv;
        r.m_[3] = c01 * inv;
        r.m_[6] = c02 * inv;
        r.m_[1] = (m[2] * m[7] - m[1] * m[8]) * inv;
        r.m_[4] = (m[0] * m[8] - m[2] * m[6]) * inv;
        r.m_[7] = (m[1] * m[6] - m[0] * m[7]) * inv;
        r.m_[2] = (m[1] * m[5] - m[2] * m[4]) * inv;
        r.m_[5] = (m[2] * m[3] - m[0] * m[5]) * inv;
        r.m_[8] = (m[0] * m[4] - m[1] * m[3]) * inv;
        return r;
    }

private:
    float m_[9];
};

// Eigen::Affine3f subset: rotation (linear part) + translation
class Affine3f {
public:
    Matrix3frm& linear() { return R_; }
    const Matrix3frm& linear() const { return R_; }
    const Matrix3frm& rotation() const { return R_; }
    Vector3f& translation() { return t_; }
    const Vector3f& translation() const { return t_; }

private:
    Matrix3frm R_;
    Vector3f t_;
};

// src/Shape/NuiCameraParams.h:5-46
struct NuiCameraIntrinsics {
    float m_fx = 0.f, m_fy = 0.f, m_cx = 0.f, m_cy = 0.f;
    NuiCameraIntrinsics() = default;
    NuiCameraIntrinsics(float fx, float fy, float cx, float cy) : m_fx(fx), m_fy(fy), m_cx(cx), m_cy(cy) {}
};
struct NuiCameraParams {
    NuiCameraIntrinsics m_intrinsics;
    // src/Frame/NuiDepthFrame.h:15-16 defaults: 500 / 4500 mm
    float m_sensorDepthMin = 0.5f;
    float m_sensorDepthMax = 4.5f;
};

// src/Shape/NuiCameraPos.h:10-90
class NuiCameraPos {
public:
    void setRotation(const Matrix3frm& rot) { m_rotation = rot; }
    const Matrix3frm& getRotation() const { return m_rotation; }
    void setLocalTranslation(const Vector3f& t) { m_localTranslation = t; }
    const Vector3f& getLocalTranslation() const { return m_localTranslation; }
    void setOffsetTranslation(const Vector3f& t) { m_offsetTranslation = t; }
    Vector3f getGlobalTranslation() const { return m_localTranslation + m_offsetTranslation; }
    Affine3f getAffine() const {
        Affine3f a;
        a.linear() = m_rotation;
        a.translation() = getGlobalTranslation();
        return a;
    }
    void setIntrinsics(const NuiCameraIntrinsics& i) { m_params.m_intrinsics = i; }
    const NuiCameraIntrinsics& getIntrinsics() const { return m_params.m_intrinsics; }
    void setSensorDepthRange(float mn, float mx) {
        m_params.m_sensorDepthMin = mn;
        m_params.m_sensorDepthMax = mx;
    }
    float getSensorDepthMin() const { return m_params.m_sensorDepthMin; }
    float getSensorDepthMax() const { return m_params.m_sensorDepthMax; }

private:
    Matrix3frm m_rotation;
    Vector3f m_localTranslation;
    Vector3f m_offsetTranslation;
    NuiCameraParams m_params;
};
