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

// DS/NuiKinfuFrame.h:11-29
class NuiKinfuFrame {
public:
    virtual ~NuiKinfuFrame() {}
    void setDirty() { m_dirty = true; }
    void clearDirty() { m_dirty = false; }
    virtual void UpdateVertexBuffers(UINT16* pDepths, UINT nNum, NuiKinfuCameraState* pCameraState) = 0;
    virtual void UpdateColorBuffers(BGRQUAD* pColors, UINT nNum) = 0;
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
    virtual bool Volume2CLVertices(NuiKinfuMesh* pData) { return false; }
    virtual bool Volume2CLMesh(NuiKinfuMesh* pData) = 0;
    virtual bool Volume2Mesh(NuiKinfuMesh* pMesh) = 0;

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
