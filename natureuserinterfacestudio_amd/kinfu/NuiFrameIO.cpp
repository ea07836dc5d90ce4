// Frame ingest: file format, prefetching loader and saver (see NuiFrameIO.h).
#include "NuiFrameIO.h"

#include <zlib.h>

#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "NuiKinfuHIP.h"

static_assert(sizeof(nuif_loaded_frame) == 144 && offsetof(nuif_loaded_frame, depth) == 80,
              "nuif_loaded_frame layout is part of the C ABI");

namespace nuif_kinfu {

namespace {

constexpr size_t kHeaderBytes = sizeof(int64_t) + 2 * sizeof(uint32_t);  // before the size field
// <id>.depth, .color, .cmap, .vertex and their bytes per pixel (UINT16, BGRQUAD,
// ColorSpacePoint, CameraSpacePoint: src/stdafx.h:122-140)
const char* const kExt[4] = {".depth", ".color", ".cmap", ".vertex"};
constexpr uint32_t kBpp[4] = {2, 4, 8, 12};

struct File {
    FILE* f = nullptr;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

std::string frame_path(const std::string& dir, uint32_t id, const char* ext) {
    std::string p = dir;
    if (!p.empty() && p.back() != '/') p.push_back('/');
    return p + std::to_string(id) + ext;
}

void read_exact(FILE* f, void* dst, size_t n, const std::string& name, const char* what) {
    if (n && std::fread(dst, 1, n, f) != n) throw IoError(name + ": truncated " + what);
}

void write_exact(FILE* f, const void* src, size_t n, const std::string& name) {
    if (n && std::fwrite(src, 1, n, f) != n) throw IoError(name + ": write failed");
}

// raw byte size of a width x height image of bpp-byte pixels; the format stores it in 32 bits
uint32_t raw_size(uint32_t w, uint32_t h, uint32_t bpp, const std::string& name) {
    const uint64_t n = (uint64_t)w * h * bpp;
    if (n > UINT32_MAX) throw IoError(name + ": image larger than the format's 32-bit size field");
    return (uint32_t)n;
}

// Payload of an open file positioned after the header: raw when the stored size equals the
// expected raw size, else a zlib stream that must inflate to exactly `raw` bytes
// (Foundation/NuiFileIOUtilities.cpp:88-103).
void read_payload(FILE* f, void* dst, uint32_t raw, const std::string& name) {
    uint32_t stored = 0;
    read_exact(f, &stored, sizeof(stored), name, "size field");
    if (stored == raw) {
        read_exact(f, dst, raw, name, "payload");
        return;
    }
    thread_local std::vector<unsigned char> z;
    z.resize(stored);
    read_exact(f, z.data(), stored, name, "compressed payload");
    uLongf out = raw;
    const int rc = uncompress(static_cast<Bytef*>(dst), &out, z.data(), stored);
    if (rc != Z_OK || out != raw)
        throw IoError(name + ": zlib payload does not inflate to " + std::to_string(raw) + " bytes");
}

}  // namespace

namespace NuiFileIOUtilities {

void writeFrameImage(const std::string& fileName, bool bCompressed, int64_t timeStamp, uint32_t nWidth,
                     uint32_t nHeight, const void* pBuffer, uint32_t bufferSize) {
    File f(std::fopen(fileName.c_str(), "wb"));
    if (!f.f) throw IoError(fileName + ": could not open file");
    write_exact(f.f, &timeStamp, sizeof(timeStamp), fileName);
    write_exact(f.f, &nWidth, sizeof(nWidth), fileName);
    write_exact(f.f, &nHeight, sizeof(nHeight), fileName);
    if (bCompressed) {
        std::vector<unsigned char> z(compressBound(bufferSize));
        uLongf zn = z.size();
        if (compress2(z.data(), &zn, static_cast<const Bytef*>(pBuffer), bufferSize, Z_BEST_SPEED) == Z_OK) {
            const uint32_t stored = (uint32_t)zn;
            write_exact(f.f, &stored, sizeof(stored), fileName);
            write_exact(f.f, z.data(), zn, fileName);
            return;
        }
    }
    write_exact(f.f, &bufferSize, sizeof(bufferSize), fileName);
    write_exact(f.f, pBuffer, bufferSize, fileName);
}

bool readFrameImageHeader(const std::string& fileName, int64_t* timeStamp, uint32_t* pWidth, uint32_t* pHeight) {
    File f(std::fopen(fileName.c_str(), "rb"));
    if (!f.f) return false;
    int64_t ts = 0;
    uint32_t wh[2] = {0, 0};
    read_exact(f.f, &ts, sizeof(ts), fileName, "header");
    read_exact(f.f, wh, sizeof(wh), fileName, "header");
    if (timeStamp) *timeStamp = ts;
    if (pWidth) *pWidth = wh[0];
    if (pHeight) *pHeight = wh[1];
    return true;
}

void readFrameImageBuffer(const std::string& fileName, void* pBuffer, uint32_t bufferSize) {
    File f(std::fopen(fileName.c_str(), "rb"));
    if (!f.f) throw IoError(fileName + ": could not open file");
    if (std::fseek(f.f, (long)kHeaderBytes, SEEK_SET) != 0) throw IoError(fileName + ": truncated header");
    read_payload(f.f, pBuffer, bufferSize, fileName);
}

void writeCamera(const std::string& fileName, float fx, float fy, float cx, float cy) {
    File f(std::fopen(fileName.c_str(), "wb"));
    if (!f.f) throw IoError(fileName + ": could not open file");
    const float v[4] = {fx, fy, cx, cy};
    write_exact(f.f, v, sizeof(v), fileName);
}

bool readCamera(const std::string& fileName, float* fx, float* fy, float* cx, float* cy) {
    File f(std::fopen(fileName.c_str(), "rb"));
    if (!f.f) return false;
    float v[4];
    read_exact(f.f, v, sizeof(v), fileName, "intrinsics");
    if (fx) *fx = v[0];
    if (fy) *fy = v[1];
    if (cx) *cx = v[2];
    if (cy) *cy = v[3];
    return true;
}

}  // namespace NuiFileIOUtilities

// ---- loader --------------------------------------------------------------------

NuiKinfuFrameLoader::NuiKinfuFrameLoader(const std::string& dir, int slots, int threads, bool upload)
    : m_dir(dir), m_upload(upload), m_slots((size_t)std::max(slots, 1)) {
    if (slots < 1 || threads < 1 || threads > 64) throw std::invalid_argument("frame loader: slots >= 1, threads 1..64");
    std::vector<nuif_stream_t> streams((size_t)threads, nullptr);
    if (upload)
        for (auto& s : streams) check(nuif_stream_create(&s), "frame loader stream");
    for (int i = 0; i < threads; ++i)
        m_threads.emplace_back([this, i, s = streams[(size_t)i]] {
            worker(i, s);
        });
}

NuiKinfuFrameLoader::~NuiKinfuFrameLoader() {
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_stop = true;
    }
    m_work.notify_all();
    m_ready.notify_all();
    for (auto& t : m_threads) t.join();
    for (auto& s : m_slots)
        for (auto& b : s.part) {
            if (b.host) m_upload ? (void)nuif_host_free(b.host) : std::free(b.host);
            if (b.device) (void)nuif_free(b.device);
        }
}

void* NuiKinfuFrameLoader::grow_host(PartBuffer& b, size_t bytes) {
    if (b.host_bytes >= bytes) return b.host;
    if (b.host) m_upload ? (void)nuif_host_free(b.host) : std::free(b.host);
    b.host = nullptr;
    b.host_bytes = 0;
    if (m_upload) {
        check(nuif_host_alloc(&b.host, bytes), "frame loader pinned buffer");
    } else {
        b.host = std::aligned_alloc(256, (bytes + 255) / 256 * 256);
        if (!b.host) throw std::bad_alloc();
    }
    b.host_bytes = bytes;
    return b.host;
}

void* NuiKinfuFrameLoader::grow_device(PartBuffer& b, size_t bytes) {
    if (b.device_bytes >= bytes) return b.device;
    if (b.device) (void)nuif_free(b.device);
    b.device = nullptr;
    b.device_bytes = 0;
    check(nuif_malloc(&b.device, bytes), "frame loader device buffer");
    b.device_bytes = bytes;
    return b.device;
}

bool NuiKinfuFrameLoader::load(Slot& s, uint32_t id, const std::string& dir, nuif_stream_t stream) {
    nuif_loaded_frame& fr = s.frame;
    fr = nuif_loaded_frame{};
    fr.frame_id = id;
    const void** hp[4] = {reinterpret_cast<const void**>(&fr.depth), reinterpret_cast<const void**>(&fr.color),
                          reinterpret_cast<const void**>(&fr.cmap), reinterpret_cast<const void**>(&fr.vertex)};
    for (int k = 0; k < 4; ++k) {
        const std::string name = frame_path(dir, id, kExt[k]);
        File f(std::fopen(name.c_str(), "rb"));
        if (!f.f) continue;
        int64_t ts = 0;
        uint32_t wh[2] = {0, 0};
        read_exact(f.f, &ts, sizeof(ts), name, "header");
        read_exact(f.f, wh, sizeof(wh), name, "header");
        if ((uint64_t)wh[0] * wh[1] == 0) continue;  // NuiImageFrame::loadFrame: nothing loaded
        const uint32_t raw = raw_size(wh[0], wh[1], kBpp[k], name);
        void* dst = grow_host(s.part[k], raw);
        read_payload(f.f, dst, raw, name);
        *hp[k] = dst;
        fr.timestamp[k] = ts;
        fr.width[k] = wh[0];
        fr.height[k] = wh[1];
        fr.parts |= 1u << k;
    }
    if (NuiFileIOUtilities::readCamera(frame_path(dir, id, ".camIntri"), &fr.fx, &fr.fy, &fr.cx, &fr.cy))
        fr.parts |= NUIF_FRAME_CAMERA;
    if (m_upload && (fr.parts & (NUIF_FRAME_DEPTH | NUIF_FRAME_COLOR))) {
        if (fr.parts & NUIF_FRAME_DEPTH) {
            const size_t n = (size_t)fr.width[0] * fr.height[0] * kBpp[0];
            void* d = grow_device(s.part[0], n);
            check(nuif_memcpy_h2d(d, fr.depth, n, stream), "frame loader depth upload");
            fr.device_depth = static_cast<const uint16_t*>(d);
        }
        if (fr.parts & NUIF_FRAME_COLOR) {
            const size_t n = (size_t)fr.width[1] * fr.height[1] * kBpp[1];
            void* d = grow_device(s.part[1], n);
            check(nuif_memcpy_h2d(d, fr.color, n, stream), "frame loader colour upload");
            fr.device_color = static_cast<const uint8_t*>(d);
        }
        check(nuif_stream_synchronize(stream), "frame loader upload");
    }
    return fr.parts != 0;
}

void NuiKinfuFrameLoader::worker(int, nuif_stream_t stream) {
    const size_t n = m_slots.size();
    std::unique_lock<std::mutex> lk(m_mu);
    for (;;) {
        m_work.wait(lk, [&] { return m_stop || (m_claim < m_end && m_slots[m_claim % n].state == State::Free); });
        if (m_stop) break;
        const uint32_t id = m_claim++;
        Slot& s = m_slots[id % n];
        const uint64_t gen = m_generation;
        const std::string dir = m_dir;
        s.state = State::Loading;
        s.id = id;
        s.generation = gen;
        lk.unlock();
        bool found = false;
        std::string err;
        try {
            found = load(s, id, dir, stream);
        } catch (const std::exception& e) {
            err = e.what();
        }
        lk.lock();
        if (gen != m_generation) {
            s.state = State::Free;  // loaded for a recording position that reset() abandoned
        } else if (!err.empty()) {
            s.state = State::Failed;
            s.error = "frame " + std::to_string(id) + ": " + err;
        } else if (!found) {
            s.state = State::End;
            if (id < m_end) m_end = id;
        } else {
            s.state = State::Ready;
        }
        m_ready.notify_all();
        m_work.notify_all();
    }
    lk.unlock();
    if (stream) (void)nuif_stream_destroy(stream);
}

int NuiKinfuFrameLoader::next(nuif_loaded_frame* out) {
    const size_t n = m_slots.size();
    std::unique_lock<std::mutex> lk(m_mu);
    const uint32_t want = m_out;
    Slot& s = m_slots[want % n];
    m_ready.wait(lk, [&] {
        return m_stop || (s.id == want && s.generation == m_generation &&
                          (s.state == State::Ready || s.state == State::End || s.state == State::Failed));
    });
    if (m_stop) throw std::runtime_error("frame loader stopped");
    if (s.state == State::End) return 0;
    if (s.state == State::Failed) throw IoError(s.error);
    s.state = State::Taken;
    *out = s.frame;
    out->slot = (int32_t)(want % n);
    ++m_out;
    return 1;
}

void NuiKinfuFrameLoader::release(int slot) {
    std::lock_guard<std::mutex> lk(m_mu);
    if (slot < 0 || (size_t)slot >= m_slots.size() || m_slots[(size_t)slot].state != State::Taken)
        throw std::invalid_argument("frame loader: slot " + std::to_string(slot) + " is not held");
    m_slots[(size_t)slot].state = State::Free;
    m_work.notify_all();
}

void NuiKinfuFrameLoader::reset() {
    {
        std::lock_guard<std::mutex> lk(m_mu);
        ++m_generation;
        m_claim = m_out = 0;
        m_end = UINT32_MAX;
        for (auto& s : m_slots)
            if (s.state != State::Loading) s.state = State::Free;
    }
    m_work.notify_all();
    m_ready.notify_all();
}

void NuiKinfuFrameLoader::updateFileName(const std::string& dir) {
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_dir = dir;
    }
    reset();
}

// ---- saver ---------------------------------------------------------------------

void saveFrame(const std::string& dir, const nuif_loaded_frame& f, bool compressed) {
    const void* hp[4] = {f.depth, f.color, f.cmap, f.vertex};
    for (int k = 0; k < 4; ++k) {
        if (!(f.parts & (1u << k))) continue;
        const std::string name = frame_path(dir, f.frame_id, kExt[k]);
        if (!hp[k]) throw std::invalid_argument(name + ": part listed without a buffer");
        const uint32_t raw = raw_size(f.width[k], f.height[k], kBpp[k], name);
        NuiFileIOUtilities::writeFrameImage(name, compressed, f.timestamp[k], f.width[k], f.height[k], hp[k], raw);
    }
    if (f.parts & NUIF_FRAME_CAMERA)
        NuiFileIOUtilities::writeCamera(frame_path(dir, f.frame_id, ".camIntri"), f.fx, f.fy, f.cx, f.cy);
}

}  // namespace nuif_kinfu
