// C++ layer of the controller side of the path, above the worker:
//  - SharedMemory    : ipc/SharedMemory.h:30-280, a System V segment with its
//                      size, a reader/writer control word and the user bytes,
//                      byte for byte the reference's layout and lock protocol
//                      (so the reference's frame readers attach to it);
//  - FrameOutputRW   : VsControllerFrameOutputRW (VsControllerFrameOutputRW.hpp:
//                      28-81), header + frame under the segment's locks;
//  - VsControllerHip : VsController + VsRenderer (VsController.cpp:42-196,
//                      VsRenderer.cpp:576-1192) for one GPU: input changes
//                      cancel the level-of-detail loop, wait for it, apply
//                      and restart it on a render thread that owns one
//                      rv_worker; frames go to the frame output segment.
// Mouse interpretation (MouseInterpreter) and the task-graph executor of the
// reference are not part of this layer: the camera is set as a matrix, and
// the worker boxes of a level are composed on the GPU by the worker.
#pragma once

#include "../revise_hip.h"
#include "vs_worker_hip.hpp"

#include <array>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace revise {

class SharedMemory {
public:
    // createNew (SharedMemory.h:50-52, 214-221): a new segment of byteCount
    // user bytes (IPC_PRIVATE), control word 0; removed when this object,
    // its creator, is destroyed.
    static std::shared_ptr<SharedMemory> create(std::size_t byteCount);
    // attachNew (SharedMemory.h:44-46, 204-211).
    static std::shared_ptr<SharedMemory> attach(int shmid);
    ~SharedMemory();
    SharedMemory(const SharedMemory&) = delete;
    SharedMemory& operator=(const SharedMemory&) = delete;

    int shmid() const noexcept { return m_shmid; }
    std::size_t size() const noexcept;

    // RAII read / write locks over the user bytes (ReadAccessor and
    // WriteAccessor, SharedMemory.h:101-187).
    class ReadAccessor {
    public:
        explicit ReadAccessor(const SharedMemory& m);
        ~ReadAccessor();
        ReadAccessor(const ReadAccessor&) = delete;
        ReadAccessor& operator=(const ReadAccessor&) = delete;
        const unsigned char* get() const noexcept { return m_data; }

    private:
        std::atomic<int>& m_ctl;
        const unsigned char* m_data;
    };
    class WriteAccessor {
    public:
        explicit WriteAccessor(SharedMemory& m);
        ~WriteAccessor();
        WriteAccessor(const WriteAccessor&) = delete;
        WriteAccessor& operator=(const WriteAccessor&) = delete;
        unsigned char* get() const noexcept { return m_data; }

    private:
        std::atomic<int>& m_ctl;
        unsigned char* m_data;
    };
    void read(void* dst, std::size_t start, std::size_t count) const;
    void write(const void* src, std::size_t start, std::size_t count);

    // The segment layout: user size, control word, user bytes.
    static constexpr std::size_t kControlOffset = sizeof(std::size_t);
    static constexpr std::size_t kUserOffset = kControlOffset + sizeof(int);
    // A writer holds the control word at -kWriterHold (MaxClientLimit).
    static constexpr int kWriterHold = 1000;

private:
    SharedMemory(int shmid, void* base, bool owner) : m_shmid(shmid), m_base(base), m_owner(owner) {}
    std::atomic<int>& control() const;
    unsigned char* user() const;
    int m_shmid = -1;
    void* m_base = nullptr;
    bool m_owner = false;
};

// VsControllerFrameOutputRW over a segment: the 24-byte header, then
// frameSize bytes of pixels.
class FrameOutputRW {
public:
    FrameOutputRW(std::shared_ptr<SharedMemory> shm, std::size_t frameSize);
    // readFrame: header always; pixels (frameSize bytes) when image != null.
    bool readFrame(FrameOutputHeader& hdr, void* image = nullptr) const;
    // writeFrame: header; pixels when image != null.
    bool writeFrame(const FrameOutputHeader& hdr, const void* image = nullptr);

private:
    std::shared_ptr<SharedMemory> m_shm;
    std::size_t m_frameSize;
};

// VsControllerFrameOutput (VsControllerFrameOutput.hpp:27-41).
struct FrameOutput {
    std::shared_ptr<SharedMemory> shmem;
    std::size_t frameSize = 0;
    int frameWidth = 0;
    int frameHeight = 0;
};

// One field of a dataset's field info file (BlockTreeFieldService::FieldInfo).
struct DatasetField {
    std::string name;
    std::array<float, 2> range{0.f, 1.f};
    bool animated = false;
};

class VsControllerHip {
public:
    // threadId: the worker's GPU is threadId % device count; splitCount:
    // worker boxes per level (ComputingCaps::workerThreadPerNodeCount);
    // measureTime: VsControllerOpts::measureRenderingTime.
    VsControllerHip(uint32_t threadId, unsigned splitCount, bool measureTime);
    ~VsControllerHip();
    VsControllerHip(const VsControllerHip&) = delete;
    VsControllerHip& operator=(const VsControllerHip&) = delete;

    // problemPath + timeFrame + primaryField + secondaryField.  Reads the
    // field info file at once (fieldNames works on return); the subtrees are
    // uploaded on the render thread.  Throws HipError(RV_ERR_INVALID) when
    // the dataset's files are not found.
    void setProblem(const std::string& path, unsigned timeFrame, const std::string& primary,
                    const std::string& secondary);
    std::vector<DatasetField> fields() const;
    unsigned timeFrameCount() const;
    // The render inputs of VsControllerInput (the state's field ranges are
    // replaced by the dataset's).
    void setInput(const RenderSharedState& in);
    void setCameraTransform(const std::array<float, 16>& m);
    std::array<float, 16> cameraTransform() const;
    // CameraController::resetCameraTransform (CameraController.cpp:88-101):
    // eye along (2, 1, 3) at L = 10 sqrt(3) 1.01 / (2 sin(fovY / 2)),
    // looking at the origin, y up; applied to the renderer at once.
    void resetCameraTransform();
    void setRenderLevel(int level);
    void start();
    FrameOutput frameOutput();
    // True if the render thread has nothing left to do within the timeout.
    bool waitIdle(std::chrono::milliseconds timeout, uint64_t* framesWritten = nullptr);
    // The failure of the render thread's last pass (RV_OK if none).
    int status(std::string* message = nullptr) const;
    void kill();

    static std::array<float, 16> defaultCameraTransform(float fovY);

private:
    struct Input {
        std::string problemPath, fieldPath, secondaryPath, metaPath;
        std::array<float, 2> range{0.f, 1.f}, secondaryRange{0.f, 1.f};
        bool hasProblem = false;
        bool hasInput = false;
        RenderSharedState state;
        int renderLevel = -1;
        uint64_t problemVersion = 0;
    };
    template <class F>
    void update(F&& f);
    void renderThread();
    void renderOnce(const Input& in);
    void setFailure(int code, const std::string& msg);
    FrameOutput frameOutputFor(uint32_t w, uint32_t h);

    const uint32_t m_threadId;
    const unsigned m_splitCount;
    const bool m_measureTime;

    mutable std::mutex m_mu;
    std::condition_variable m_cv;
    Input m_in;
    std::vector<DatasetField> m_fields;
    unsigned m_timeFrames = 0;
    bool m_started = false, m_quit = false, m_pending = false, m_busy = false;
    volatile int m_cancel = 0;
    uint64_t m_frames = 0;      // frames written (the next frame number)
    int m_status = RV_OK;
    std::string m_message;

    std::mutex m_outMu;
    FrameOutput m_out;

    // render thread only
    rv_worker* m_worker = nullptr;
    uint64_t m_loadedVersion = ~0ull;
    bool m_hasSecondary = false; // the secondary field's values are loaded
    unsigned m_maxFullLevel = 0;
    std::thread m_thread;
};

} // namespace revise
