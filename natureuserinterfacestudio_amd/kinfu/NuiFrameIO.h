// Frame ingest: the reference's recorded-frame file format
// (Foundation/NuiFileIOUtilities.cpp:13-160) and its frame loader / saver
// (DeviceManager/FileLoader/NuiFrameLoader.cpp:26-91,
// Frame/NuiFrameSaveManager.cpp:24-80), rebuilt as a prefetching pipeline that
// hands frames over already resident in device memory.
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/nuif_io.h"

namespace nuif_kinfu {

// A missing, truncated or undecodable file (status NUIF_ERR_IO at the C boundary).
struct IoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

namespace NuiFileIOUtilities {
// Foundation/NuiFileIOUtilities.cpp:13-52
void writeFrameImage(const std::string& fileName, bool bCompressed, int64_t timeStamp, uint32_t nWidth,
                     uint32_t nHeight, const void* pBuffer, uint32_t bufferSize);
// :54-75; false when the file cannot be opened
bool readFrameImageHeader(const std::string& fileName, int64_t* timeStamp, uint32_t* pWidth, uint32_t* pHeight);
// :77-110; bufferSize is the raw size the caller expects
void readFrameImageBuffer(const std::string& fileName, void* pBuffer, uint32_t bufferSize);
// :112-160
void writeCamera(const std::string& fileName, float fx, float fy, float cx, float cy);
bool readCamera(const std::string& fileName, float* fx, float* fy, float* cx, float* cy);
}  // namespace NuiFileIOUtilities

// Host (and, when uploading, device) memory of one part of one slot.
struct PartBuffer {
    void* host = nullptr;
    size_t host_bytes = 0;
    void* device = nullptr;
    size_t device_bytes = 0;
};

class NuiKinfuFrameLoader {
public:
    NuiKinfuFrameLoader(const std::string& dir, int slots, int threads, bool upload);
    ~NuiKinfuFrameLoader();
    NuiKinfuFrameLoader(const NuiKinfuFrameLoader&) = delete;
    NuiKinfuFrameLoader& operator=(const NuiKinfuFrameLoader&) = delete;

    void updateFileName(const std::string& dir);  // and reset()
    void reset();                                 // back to frame 0
    // 1: *out holds the next frame, 0: end of the recording; throws IoError on a bad file
    int next(nuif_loaded_frame* out);
    void release(int slot);

private:
    enum class State { Free, Loading, Ready, End, Failed, Taken };
    struct Slot {
        State state = State::Free;
        uint32_t id = 0;
        uint64_t generation = 0;
        nuif_loaded_frame frame{};
        PartBuffer part[4];
        std::string error;
    };
    void worker(int index, nuif_stream_t stream);
    // loads frame `id` into slot `s` (no lock held); false when no file of the frame exists
    bool load(Slot& s, uint32_t id, const std::string& dir, nuif_stream_t stream);
    void* grow_host(PartBuffer& b, size_t bytes);
    void* grow_device(PartBuffer& b, size_t bytes);

    std::string m_dir;
    const bool m_upload;
    std::vector<Slot> m_slots;
    std::vector<std::thread> m_threads;
    std::mutex m_mu;
    std::condition_variable m_work, m_ready;
    uint32_t m_claim = 0;               // next frame id a worker takes
    uint32_t m_out = 0;                 // next frame id handed to the consumer
    uint32_t m_end = UINT32_MAX;        // first frame id without any file
    uint64_t m_generation = 0;          // bumped by reset()
    bool m_stop = false;
};

// Frame/NuiFrameSaveManager.cpp:24-80 process() for one frame
void saveFrame(const std::string& dir, const nuif_loaded_frame& f, bool compressed);

}  // namespace nuif_kinfu
