// The reference's device-independent interfaces of the fusion path
// (src/SLAM/VisualOdometry/DeviceSpecific/NuiKinfu{Frame,FeedbackFrame,Tracker,
// Scene,CameraState}.h), kept with the same method names and meaning so that
// the engine and its callers are unchanged.  GL-mappable outputs
// (NuiCLMappableData) are replaced by plain host arrays.
#pragma once

#include <atomic>
#include <string>
#include <vector>

#include "NuiKinfuTypes.h"

class NuiKinfuCameraState;

// Host mesh returned by marching cubes (replaces the GL-shared NuiCLMappableData
// and NuiMeshShape outputs): positions (3 floats per vertex, 3 vertices per
// triangle) and BGRA colours per vertex.
struct NuiKinfuMesh {
    std::vector<float> positions;
    std::vector<float> colors;
    std::vector<uint32_t> indices;  // only for the deduplicated CPU-side mesh
    size_t triangleCount() const { return positions.size() / 9; }
};

typedef Vector3f SgVec3f;

// src/Shape/NuiMeshShape.h:7-61: shared points and triangle indices (Volume2Mesh's output for
// the SLAM controller, src/SLAM/Backend/NuiSLAMController.cpp:121-127); UVs and colours are
// not produced by the fusion path.
class NuiMeshShape {
public:
    void appendPoint(const SgVec3f& point) { m_points.push_back(point); }
    void appendTriangleIndex(int index) { m_triangle_indices.push_back(index); }
    size_t pointsNum() const { return m_points.size(); }
    const SgVec3f& getPoint(unsigned int id) const { return m_points.at(id); }
    size_t trianglesNum() const { return m_triangle_indices.size() / 3; }
    int triangleIndex(unsigned int id) const {
        return id >= m_triangle_indices.size() ? -1 : m_triangle_indices.at(id);
    }
    void clear() {
        m_points.clear();
        m_triangle_indices.clear();
    }

private:
    std::vector<SgVec3f> m_points;
    std::vector<int> m_triangle_indices;
};

// DS/NuiKinfuFrame.h:11-29
class NuiKinfuFrame {
public:
    virtual ~NuiKinfuFrame() {}
    void setDirty() { m_dirty = true; }
    void clearDirty() { m_dirty = false; }
    virtual void UpdateVertexBuffers(UINT16* pDepths, UINT nNum, NuiKinfuCameraState* pCameraState) = 0;
    virtual void UpdateColorBuffers(BGRQUAD* pColors, UINT nNum) = 0;
    // the colour image seen through the depth-to-colour map (and the depth distortion
    // lookup table, or NULL), DS/NuiKinfuFrame.h:22
    virtual void UpdateColorBuffers(ColorSpacePoint* pDepthToColor, UINT* pDepthDistortionLT, UINT nNum,
                                    const NuiColorImage& image) = 0;
    virtual UINT GetWidth() const = 0;
    virtual UINT GetHeight() const = 0;

protected:
    std::atomic<bool> m_dirty{false};
};

// DS/NuiKinfuFeedbackFrame.h:11-36
class NuiKinfuFeedbackFrame {
public:
    enum TrackerBufferType { eTracker_Vertices = 0, eTracker_Normals, eTracker_Ranges };
    virtual ~NuiKinfuFeedbackFrame() {}
    void setDirty() { m_dirty = true; }
    void clearDirty() { m_dirty = false; }
    virtual void UpdateBuffers(NuiKinfuFrame* frame, NuiKinfuCameraState* pCameraState) = 0;
    virtual UINT GetWidth() const = 0;
    virtual UINT GetHeight() const = 0;

protected:
    std::atomic<bool> m_dirty{false};
};

// DS/NuiKinfuTracker.h:9-23
class NuiKinfuTracker {
public:
    virtual ~NuiKinfuTracker() {}
    virtual bool EstimatePose(NuiKinfuFrame* pFrame, NuiKinfuFeedbackFrame* pFeedbackFrame,
                              NuiKinfuCameraState* pCameraState, Affine3f* hint) = 0;
    virtual bool hasColorData() const { return false; }
    virtual void log(const std::string& fileName) const {}
    virtual float getError() const { return 0.0f; }
    virtual int getCount() const { return 0; }
};

// DS/NuiKinfuScene.h:16-47
class NuiKinfuScene {
public:
    virtual ~NuiKinfuScene() {}
    void setDirty() { m_dirty = true; }
    void clearDirty() { m_dirty = false; }
    virtual bool log(const std::string& fileName) const = 0;
    virtual bool hasColorData() const { return true; }
    virtual bool needAcceleratedFeedback() const { return false; }
    virtual float getVoxelLeafSize() const { return 0.0f; }
    virtual void reset() = 0;
    virtual bool integrateVolume(NuiKinfuFrame* pFrame, NuiKinfuCameraState* pCameraState) = 0;
    virtual void raycastRender(NuiKinfuFeedbackFrame* pFeedbackFrame, NuiKinfuCameraState* pCameraState) = 0;
    // DS/NuiKinfuScene.h:39: the reference's hash scene defers its render to a GL thread that
    // calls this; the render here happens in raycastRender, so there is nothing left to do
    virtual void offlineRender() {}
    virtual bool Volume2CLVertices(NuiKinfuMesh* pData) { return false; }
    virtual bool Volume2CLMesh(NuiKinfuMesh* pData) = 0;
    virtual bool Volume2Mesh(NuiKinfuMesh* pMesh) = 0;
    // DS/NuiKinfuScene.h:43: the scene's mesh as a NuiMeshShape (points, then triangle
    // indices), appended to pMesh.  A scene whose mesh has shared vertices (the dense
    // volume) gives its indices; a triangle soup (the hash scene, whose reference
    // Volume2Mesh(NuiMeshShape*) appends nothing) gets indices 0, 1, 2, ... so that every
    // triangle of the marching cubes is in the shape.
    virtual bool Volume2Mesh(NuiMeshShape* pMesh) {
        if (!pMesh) return false;
        NuiKinfuMesh m;
        if (!Volume2Mesh(&m)) return false;
        const int base = (int)pMesh->pointsNum();
        for (size_t i = 0; i + 2 < m.positions.size(); i += 3)
            pMesh->appendPoint(SgVec3f(m.positions[i], m.positions[i + 1], m.positions[i + 2]));
        if (!m.indices.empty()) {
            for (uint32_t k : m.indices) pMesh->appendTriangleIndex(base + (int)k);
        } else {
            const int n = (int)(m.positions.size() / 9) * 3;
            for (int k = 0; k < n; ++k) pMesh->appendTriangleIndex(base + k);
        }
        return true;
    }

protected:
    std::atomic<bool> m_dirty{false};
};

// DS/NuiKinfuCameraState.h:5-31
class NuiKinfuCameraState {
public:
    virtual ~NuiKinfuCameraState() {}
    virtual void UpdateCameraTransform(const Matrix3frm& rot, const Vector3f& tran) {
        m_pos.setRotation(rot);
        m_pos.setLocalTranslation(tran);
    }
    virtual void UpdateCameraTransform(const Matrix3frm& rot, const Vector3f& tran, const Vector3f& offsetTran) {
        m_pos.setRotation(rot);
        m_pos.setLocalTranslation(tran);
        m_pos.setOffsetTranslation(offsetTran);
    }
    virtual void UpdateCameraParams(const NuiCameraParams& cameraParams, UINT nWidth, UINT nHeight) {
        m_pos.setIntrinsics(cameraParams.m_intrinsics);
        m_pos.setSensorDepthRange(cameraParams.m_sensorDepthMin, cameraParams.m_sensorDepthMax);
    }
    const NuiCameraPos& GetCameraPos() const { return m_pos; }

protected:
    NuiCameraPos m_pos;
};
