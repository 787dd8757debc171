// Controller side of the path (see include/revise/vs_controller_hip.hpp):
// System V shared memory with the reference's segment layout and lock
// protocol, the frame output reader/writer over it, and a controller that
// runs the level-of-detail loop of one worker on a render thread, restarting
// it on every input change.  Host code only: every device operation goes
// through the worker's C ABI, on the render thread that created the worker.
#include "../../include/revise/vs_controller_hip.hpp"

#include <sys/ipc.h>
#include <sys/shm.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <new>
#include <stdexcept>

namespace rv {
void set_error(const std::string& msg);
}

namespace revise {

// ---------------------------------------------------------------- shared memory

static_assert(sizeof(std::size_t) == 8 && sizeof(std::atomic<int>) == 4 &&
                  std::atomic<int>::is_always_lock_free,
              "segment layout: 8-byte size, lock-free 4-byte control word");
static_assert(sizeof(FrameOutputHeader) == 24, "frame output header layout");

std::shared_ptr<SharedMemory> SharedMemory::create(std::size_t byteCount)
{
    const int id = shmget(IPC_PRIVATE, kUserOffset + byteCount, IPC_CREAT | 0666);
    if (id < 0)
        throw HipError(RV_ERR_INVALID, std::string("shmget failed: ") + std::strerror(errno));
    void* p = shmat(id, nullptr, 0);
    if (p == reinterpret_cast<void*>(-1)) {
        const int err = errno;
        shmctl(id, IPC_RMID, nullptr);
        throw HipError(RV_ERR_INVALID, std::string("shmat failed: ") + std::strerror(err));
    }
    // a new segment is zero-filled: size, then the control word at 0
    *static_cast<std::size_t*>(p) = byteCount;
    new (static_cast<unsigned char*>(p) + kControlOffset) std::atomic<int>(0);
    return std::shared_ptr<SharedMemory>(new SharedMemory(id, p, true));
}

std::shared_ptr<SharedMemory> SharedMemory::attach(int shmid)
{
    void* p = shmid >= 0 ? shmat(shmid, nullptr, 0) : reinterpret_cast<void*>(-1);
    if (p == reinterpret_cast<void*>(-1))
        throw HipError(RV_ERR_INVALID, "Wrong Shared Memory Id passed, please check");
    return std::shared_ptr<SharedMemory>(new SharedMemory(shmid, p, false));
}

SharedMemory::~SharedMemory()
{
    if (m_base)
        shmdt(m_base);
    if (m_owner)
        shmctl(m_shmid, IPC_RMID, nullptr);
}

std::size_t SharedMemory::size() const noexcept { return *static_cast<const std::size_t*>(m_base); }

std::atomic<int>& SharedMemory::control() const
{
    return *reinterpret_cast<std::atomic<int>*>(static_cast<unsigned char*>(m_base) + kControlOffset);
}

unsigned char* SharedMemory::user() const { return static_cast<unsigned char*>(m_base) + kUserOffset; }

// Readers count up while no writer holds the word (a writer makes it
// negative); a writer swaps 0 for -kWriterHold, so it waits for every reader
// to leave and keeps new readers out until it adds kWriterHold back.
SharedMemory::ReadAccessor::ReadAccessor(const SharedMemory& m) : m_ctl(m.control()), m_data(m.user())
{
    while (m_ctl.fetch_add(1) + 1 < 0) {
        m_ctl.fetch_sub(1);
        std::this_thread::yield();
    }
}

SharedMemory::ReadAccessor::~ReadAccessor() { m_ctl.fetch_sub(1); }

SharedMemory::WriteAccessor::WriteAccessor(SharedMemory& m) : m_ctl(m.control()), m_data(m.user())
{
    for (;;) {
        int expected = 0;
        if (m_ctl.compare_exchange_weak(expected, -kWriterHold))
            break;
        std::this_thread::yield();
    }
}

SharedMemory::WriteAccessor::~WriteAccessor() { m_ctl.fetch_add(kWriterHold); }

void SharedMemory::read(void* dst, std::size_t start, std::size_t count) const
{
    if (start > size() || count > size() - start)
        throw HipError(RV_ERR_INVALID, "SharedMemory::read: range outside the segment");
    ReadAccessor acc(*this);
    std::memcpy(dst, acc.get() + start, count);
}

void SharedMemory::write(const void* src, std::size_t start, std::size_t count)
{
    if (start > size() || count > size() - start)
        throw HipError(RV_ERR_INVALID, "SharedMemory::write: range outside the segment");
    WriteAccessor acc(*this);
    std::memcpy(acc.get() + start, src, count);
}

// ---------------------------------------------------------------- frame output

FrameOutputRW::FrameOutputRW(std::shared_ptr<SharedMemory> shm, std::size_t frameSize)
    : m_shm(std::move(shm)), m_frameSize(frameSize)
{
    if (!m_shm || sizeof(FrameOutputHeader) + m_frameSize > m_shm->size())
        throw HipError(RV_ERR_INVALID, "frame output: the segment cannot hold the header and a frame");
}

bool FrameOutputRW::readFrame(FrameOutputHeader& hdr, void* image) const
{
    SharedMemory::ReadAccessor acc(*m_shm);
    std::memcpy(&hdr, acc.get(), sizeof hdr);
    if (!image || !m_frameSize)
        return false;
    std::memcpy(image, acc.get() + sizeof hdr, m_frameSize);
    return true;
}

bool FrameOutputRW::writeFrame(const FrameOutputHeader& hdr, const void* image)
{
    SharedMemory::WriteAccessor acc(*m_shm);
    std::memcpy(acc.get(), &hdr, sizeof hdr);
    if (!image || !m_frameSize)
        return false;
    std::memcpy(acc.get() + sizeof hdr, image, m_frameSize);
    return true;
}

// ---------------------------------------------------------------- dataset names

namespace {

namespace fs = std::filesystem;

bool isFile(const std::string& p)
{
    std::error_code ec;
    return fs::is_regular_file(p, ec);
}

bool isDir(const std::string& p)
{
    std::error_code ec;
    return fs::is_directory(p, ec);
}

// The files of a problem path (firstOutputFrameFileName / frameOutputFileName,
// filename_util.hpp:36-94; BlockTreeFieldService::readFieldInfoFile and
// openMetadataFile, BlockTreeFieldService.hpp:177-191, 264-272): without time
// steps everything hangs off <base>; with time steps (<dir>/<stem> is a
// directory) frame k's files hang off <dir>/<stem>/<stem>_tNNNNNN<ext>, the
// info file is <dir>/<stem>/<stem><ext>.s3dmm-fields, and the metadata is
// frame 0's.
struct ProblemFiles {
    bool timeSteps = false;
    std::string info, meta;
    std::string frameBase(const std::string& base, unsigned frame) const
    {
        if (!timeSteps)
            return base;
        const fs::path p(base);
        char num[16];
        std::snprintf(num, sizeof num, "_t%06u", frame);
        return (p.parent_path() / p.stem() / (p.stem().string() + num + p.extension().string())).string();
    }
};

ProblemFiles resolveProblem(const std::string& base)
{
    ProblemFiles f;
    if (isFile(base) || isFile(base + ".s3dmm-meta")) {
        f.info = base + ".s3dmm-fields";
    } else {
        const fs::path p(base);
        const fs::path dir = p.parent_path() / p.stem();
        if (!isDir(dir.string()))
            throw HipError(RV_ERR_INVALID,
                           "Failed to find input file(s) matching the specified name '" + base + "'");
        f.timeSteps = true;
        f.info = (dir / (p.stem().string() + p.extension().string() + ".s3dmm-fields")).string();
    }
    f.meta = f.frameBase(base, 0) + ".s3dmm-meta";
    return f;
}

// resolveField (VsRenderer.cpp:57-72): the named field, else the first.
int resolveField(const std::vector<DatasetField>& fields, const std::string& name)
{
    for (size_t i = 0; i < fields.size(); ++i)
        if (fields[i].name == name)
            return (int)i;
    return fields.empty() ? -1 : 0;
}

std::array<float, 16> lookAt(const double eye[3])
{
    // row-major world->eye transform of a camera at eye looking at the
    // origin, y up, looking along its -z
    double f[3] = {-eye[0], -eye[1], -eye[2]};
    double n = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& v : f)
        v /= n;
    const double up[3] = {0.0, 1.0, 0.0};
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    n = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& v : s)
        v /= n;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    const double* R[3] = {s, u, f};
    std::array<float, 16> M{};
    for (int r = 0; r < 3; ++r) {
        const double sign = r == 2 ? -1.0 : 1.0;
        double t = 0.0;
        for (int c = 0; c < 3; ++c) {
            M[4 * r + c] = (float)(sign * R[r][c]);
            t -= sign * R[r][c] * eye[c];
        }
        M[4 * r + 3] = (float)t;
    }
    M[15] = 1.f;
    return M;
}

void check(int rc)
{
    if (rc != RV_OK)
        throw HipError(rc, rv_last_error());
}

} // namespace

// ---------------------------------------------------------------- controller

std::array<float, 16> VsControllerHip::defaultCameraTransform(float fovY)
{
    const double d[3] = {2.0, 1.0, 3.0};
    const double n = std::sqrt(14.0);
    const double L = 10.0 * std::sqrt(3.0) * 1.01 / (2.0 * std::sin(0.5 * fovY * M_PI / 180.0));
    const double eye[3] = {d[0] / n * L, d[1] / n * L, d[2] / n * L};
    return lookAt(eye);
}

VsControllerHip::VsControllerHip(uint32_t threadId, unsigned splitCount, bool measureTime)
    : m_threadId(threadId), m_splitCount(splitCount ? splitCount : 1u), m_measureTime(measureTime)
{
    m_in.state.cameraTransform = defaultCameraTransform(m_in.state.fovY);
    m_thread = std::thread([this] { renderThread(); });
}

VsControllerHip::~VsControllerHip() { kill(); }

void VsControllerHip::kill()
{
    {
        std::lock_guard<std::mutex> lk(m_mu);
        m_quit = true;
        m_cancel = 1;
    }
    m_cv.notify_all();
    if (m_thread.joinable())
        m_thread.join();
    std::lock_guard<std::mutex> lk(m_outMu);
    m_out = FrameOutput{};
}

// updateRenderSharedState (VsRenderer.cpp:1177-1192): cancel the loop, wait
// for the render thread to leave it, apply, and render again.
template <class F>
void VsControllerHip::update(F&& f)
{
    std::unique_lock<std::mutex> lk(m_mu);
    if (m_quit)
        throw HipError(RV_ERR_INVALID, "the controller was killed");
    m_cancel = 1;
    m_cv.wait(lk, [this] { return !m_busy; });
    f(m_in);
    m_cancel = 0;
    if (m_started)
        m_pending = true;
    lk.unlock();
    m_cv.notify_all();
}

void VsControllerHip::setProblem(const std::string& path, unsigned timeFrame,
                                 const std::string& primary, const std::string& secondary)
{
    const ProblemFiles pf = resolveProblem(path);
    uint32_t steps = 0, count = 0;
    int rc = rv_fields_info_read(pf.info.c_str(), &steps, nullptr, 0, &count);
    std::vector<rv_field_info> raw(std::max<uint32_t>(count, 1u));
    if (count)
        rc = rv_fields_info_read(pf.info.c_str(), &steps, raw.data(), count, &count);
    check(rc);
    std::vector<DatasetField> fields;
    for (uint32_t i = 0; i < count; ++i)
        fields.push_back({raw[i].name, {raw[i].min, raw[i].max}, raw[i].animated != 0});
    if (steps && timeFrame >= steps)
        throw HipError(RV_ERR_INVALID, "time frame out of range");
    const int ip = resolveField(fields, primary);
    const int is = resolveField(fields, secondary);
    const std::string base = pf.frameBase(path, timeFrame);
    update([&](Input& in) {
        in.problemPath = path;
        in.metaPath = pf.meta;
        in.hasProblem = ip >= 0;
        in.fieldPath = ip >= 0 ? base + ".s3dmm-field#" + fields[ip].name : std::string();
        in.range = ip >= 0 ? fields[ip].range : std::array<float, 2>{0.f, 1.f};
        in.secondaryPath = is >= 0 ? base + ".s3dmm-field#" + fields[is].name : std::string();
        in.secondaryRange = is >= 0 ? fields[is].range : std::array<float, 2>{0.f, 1.f};
        ++in.problemVersion;
        m_fields = fields;
        m_timeFrames = steps;
    });
}

std::vector<DatasetField> VsControllerHip::fields() const
{
    std::lock_guard<std::mutex> lk(m_mu);
    return m_fields;
}

unsigned VsControllerHip::timeFrameCount() const
{
    std::lock_guard<std::mutex> lk(m_mu);
    return m_timeFrames;
}

void VsControllerHip::setInput(const RenderSharedState& s)
{
    if (s.viewportSize[0] == 0 || s.viewportSize[1] == 0 || s.colorTable.empty() ||
        s.colorTable.size() % 4 != 0)
        throw HipError(RV_ERR_INVALID, "controller input: empty viewport or colour table");
    update([&](Input& in) {
        in.state = s;
        in.hasInput = true;
    });
}

void VsControllerHip::setCameraTransform(const std::array<float, 16>& m)
{
    update([&](Input& in) { in.state.cameraTransform = m; });
}

std::array<float, 16> VsControllerHip::cameraTransform() const
{
    std::lock_guard<std::mutex> lk(m_mu);
    return m_in.state.cameraTransform;
}

void VsControllerHip::resetCameraTransform()
{
    update([&](Input& in) { in.state.cameraTransform = defaultCameraTransform(in.state.fovY); });
}

void VsControllerHip::setRenderLevel(int level)
{
    update([&](Input& in) { in.renderLevel = level < 0 ? -1 : level; });
}

void VsControllerHip::start()
{
    {
        std::lock_guard<std::mutex> lk(m_mu);
        if (m_quit)
            throw HipError(RV_ERR_INVALID, "the controller was killed");
        if (!m_started) {
            m_started = true;
            m_pending = true;
        }
    }
    m_cv.notify_all();
}

// maybeInitFrameOutput (VsRenderer.cpp:81-103): a new segment whenever the
// frame size changes, first holding the "no frame" header.
FrameOutput VsControllerHip::frameOutputFor(uint32_t w, uint32_t h)
{
    std::lock_guard<std::mutex> lk(m_outMu);
    if (!m_out.shmem || m_out.frameWidth != (int)w || m_out.frameHeight != (int)h) {
        FrameOutput o;
        o.frameWidth = (int)w;
        o.frameHeight = (int)h;
        o.frameSize = (std::size_t)w * h * 4;
        o.shmem = SharedMemory::create(sizeof(FrameOutputHeader) + o.frameSize);
        FrameOutputHeader none;
        none.frameNumber = ~0ull;
        none.durationMs = 0;
        none.level = ~0u;
        FrameOutputRW(o.shmem, o.frameSize).writeFrame(none);
        m_out = o;
    }
    return m_out;
}

FrameOutput VsControllerHip::frameOutput()
{
    uint32_t w, h;
    {
        std::lock_guard<std::mutex> lk(m_mu);
        w = m_in.state.viewportSize[0];
        h = m_in.state.viewportSize[1];
    }
    return frameOutputFor(w, h);
}

bool VsControllerHip::waitIdle(std::chrono::milliseconds timeout, uint64_t* framesWritten)
{
    std::unique_lock<std::mutex> lk(m_mu);
    const bool idle = m_cv.wait_for(lk, timeout, [this] { return !m_pending && !m_busy; });
    if (framesWritten)
        *framesWritten = m_frames;
    return idle;
}

int VsControllerHip::status(std::string* message) const
{
    std::lock_guard<std::mutex> lk(m_mu);
    if (message)
        *message = m_message;
    return m_status;
}

void VsControllerHip::setFailure(int code, const std::string& msg)
{
    std::lock_guard<std::mutex> lk(m_mu);
    m_status = code;
    m_message = msg;
}

void VsControllerHip::renderThread()
{
    std::unique_lock<std::mutex> lk(m_mu);
    for (;;) {
        m_cv.wait(lk, [this] { return m_quit || m_pending; });
        if (m_quit)
            break;
        m_pending = false;
        m_busy = true;
        const Input in = m_in;
        m_status = RV_OK;
        m_message.clear();
        lk.unlock();
        try {
            renderOnce(in);
        } catch (const HipError& e) {
            setFailure(e.code, e.what());
        } catch (const std::bad_alloc&) {
            setFailure(RV_ERR_OOM, "host allocation failed");
        } catch (const std::exception& e) {
            setFailure(RV_ERR_INTERNAL, e.what());
        }
        lk.lock();
        m_busy = false;
        m_cv.notify_all();
    }
    lk.unlock();
    if (m_worker) {
        rv_worker_destroy(m_worker);
        m_worker = nullptr;
    }
}

// One pass of the render thread: (re)load the problem's subtrees when the
// problem changed, hand the inputs to the worker, then continueRendering
// (VsRenderer.cpp:1061-1166): each level from the first to maxFullLevel
// (or the render level alone) composed and written to the frame output,
// until cancelled.
void VsControllerHip::renderOnce(const Input& in)
{
    if (!in.hasProblem || !in.hasInput)
        return;
    const bool secondary = in.state.fieldMode == RV_FIELD_VALUE_ON_ISOSURFACE ||
                           in.state.fieldMode == RV_FIELD_VALUE_ON_ISOSURFACES;
    if (m_loadedVersion != in.problemVersion || (secondary && m_loadedVersion != ~0ull && !m_hasSecondary)) {
        if (m_worker) {
            rv_worker_destroy(m_worker);
            m_worker = nullptr;
        }
        m_loadedVersion = ~0ull;
        m_hasSecondary = false;
        rv_dataset* ds = nullptr;
        check(rv_dataset_open(in.metaPath.c_str(), &ds));
        std::unique_ptr<rv_dataset, void (*)(rv_dataset*)> hold(ds, rv_dataset_close);
        rv_dataset_info info;
        check(rv_dataset_get_info(ds, &info));
        check(rv_worker_create(m_threadId, &m_worker));
        uint64_t loaded = 0;
        check(rv_dataset_load_worker(ds, in.fieldPath.c_str(), m_worker, 0, info.max_full_level, &loaded));
        if (secondary) {
            check(rv_dataset_load_worker_secondary(ds, in.secondaryPath.c_str(), m_worker, 0,
                                                   info.max_full_level, &loaded));
            m_hasSecondary = true;
        }
        m_maxFullLevel = info.max_full_level;
        m_loadedVersion = in.problemVersion;
    }
    const RenderSharedState& s = in.state;
    rv_render_state rs{};
    rs.view16 = s.cameraTransform.data();
    rs.vp_w = s.viewportSize[0];
    rs.vp_h = s.viewportSize[1];
    rs.fov_y_deg = s.fovY;
    rs.quality = s.renderQuality;
    rs.threshold_rel = s.threshold;
    rs.alpha_cutoff = s.alphaCutoff;
    rs.fmin = in.range[0];
    rs.fmax = in.range[1];
    for (int i = 0; i < 3; ++i)
        rs.background[i] = s.backgroundColor[i];
    rs.tf = s.colorTable.data();
    rs.tf_n = (uint32_t)(s.colorTable.size() / 4);
    rs.field_mode = s.fieldMode;
    rs.iso_level_rel = s.isosurfaceLevel;
    rs.iso_opacity = s.isosurfaceOpacity;
    rs.iso_levels_rel = s.isosurfaceLevels.empty() ? nullptr : s.isosurfaceLevels.data();
    rs.iso_level_count = (uint32_t)s.isosurfaceLevels.size();
    rs.smin = in.secondaryRange[0];
    rs.smax = in.secondaryRange[1];
    rs.exact = s.exact ? 1u : 0u;
    check(rv_worker_set_state(m_worker, &rs));

    unsigned levelCount = m_maxFullLevel + 1, minLevel = 0;
    if (in.renderLevel >= 0) {
        minLevel = std::min((unsigned)in.renderLevel, levelCount - 1);
        levelCount = minLevel + 1;
    }
    std::vector<uint32_t> frame((size_t)rs.vp_w * rs.vp_h);
    for (unsigned level = minLevel; level < levelCount && !m_cancel; ++level) {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = rv_worker_compose_frame(m_worker, level, m_splitCount, frame.data(), &m_cancel);
        if (rc == RV_ERR_CANCELLED || m_cancel)
            break;
        check(rc);
        FrameOutputHeader h;
        {
            std::lock_guard<std::mutex> lk(m_mu);
            h.frameNumber = m_frames;
        }
        if (m_measureTime) {
            const std::chrono::duration<double, std::milli> ms = std::chrono::steady_clock::now() - t0;
            h.durationMs = (uint64_t)std::llround(ms.count());
        }
        h.level = level;
        const FrameOutput out = frameOutputFor(rs.vp_w, rs.vp_h);
        FrameOutputRW(out.shmem, out.frameSize).writeFrame(h, frame.data());
        std::lock_guard<std::mutex> lk(m_mu);
        ++m_frames;
    }
}

} // namespace revise

// ---------------------------------------------------------------- C ABI

struct rv_shm {
    std::shared_ptr<revise::SharedMemory> m;
};

struct rv_controller {
    std::unique_ptr<revise::VsControllerHip> c;
};

#define C_TRY try {
#define C_CATCH                                                                        \
    }                                                                                  \
    catch (const revise::HipError& e)                                                  \
    {                                                                                  \
        rv::set_error(e.what());                                                       \
        return e.code;                                                                 \
    }                                                                                  \
    catch (const std::bad_alloc&)                                                      \
    {                                                                                  \
        rv::set_error("host allocation failed");                                       \
        return RV_ERR_OOM;                                                             \
    }                                                                                  \
    catch (const std::exception& e)                                                    \
    {                                                                                  \
        rv::set_error(e.what());                                                       \
        return RV_ERR_INTERNAL;                                                        \
    }

static int bad(const char* msg)
{
    rv::set_error(msg);
    return RV_ERR_INVALID;
}

extern "C" int rv_shm_create(uint64_t bytes, rv_shm** out)
{
    C_TRY
    if (!out)
        return bad("rv_shm_create: null out");
    *out = nullptr;
    auto h = std::make_unique<rv_shm>();
    h->m = revise::SharedMemory::create(bytes);
    *out = h.release();
    return RV_OK;
    C_CATCH
}

extern "C" int rv_shm_attach(int shmid, rv_shm** out)
{
    C_TRY
    if (!out)
        return bad("rv_shm_attach: null out");
    *out = nullptr;
    auto h = std::make_unique<rv_shm>();
    h->m = revise::SharedMemory::attach(shmid);
    *out = h.release();
    return RV_OK;
    C_CATCH
}

extern "C" int rv_shm_id(const rv_shm* shm) { return shm && shm->m ? shm->m->shmid() : -1; }

extern "C" uint64_t rv_shm_size(const rv_shm* shm) { return shm && shm->m ? shm->m->size() : 0; }

extern "C" int rv_shm_read(const rv_shm* shm, void* dst, uint64_t start, uint64_t count)
{
    C_TRY
    if (!shm || (count && !dst))
        return bad("rv_shm_read: bad arguments");
    shm->m->read(dst, start, count);
    return RV_OK;
    C_CATCH
}

extern "C" int rv_shm_write(rv_shm* shm, const void* src, uint64_t start, uint64_t count)
{
    C_TRY
    if (!shm || (count && !src))
        return bad("rv_shm_write: bad arguments");
    shm->m->write(src, start, count);
    return RV_OK;
    C_CATCH
}

extern "C" void rv_shm_close(rv_shm* shm) { delete shm; }

extern "C" int rv_frame_output_write_shm(rv_shm* shm, uint64_t frame_size,
                                         const rv_frame_output_header* header, const void* pixels)
{
    C_TRY
    if (!shm || !header)
        return bad("rv_frame_output_write_shm: bad arguments");
    revise::FrameOutputHeader h;
    std::memcpy(&h, header, sizeof h);
    revise::FrameOutputRW(shm->m, frame_size).writeFrame(h, pixels);
    return RV_OK;
    C_CATCH
}

extern "C" int rv_frame_output_read_shm(const rv_shm* shm, uint64_t frame_size,
                                        rv_frame_output_header* header, void* pixels)
{
    C_TRY
    if (!shm || !header)
        return bad("rv_frame_output_read_shm: bad arguments");
    revise::FrameOutputHeader h;
    revise::FrameOutputRW(shm->m, frame_size).readFrame(h, pixels);
    std::memcpy(header, &h, sizeof h);
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_create(uint32_t thread_id, uint32_t split_count, int measure_time,
                                    rv_controller** out)
{
    C_TRY
    if (!out)
        return bad("rv_controller_create: null out");
    *out = nullptr;
    auto h = std::make_unique<rv_controller>();
    h->c = std::make_unique<revise::VsControllerHip>(thread_id, split_count, measure_time != 0);
    *out = h.release();
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_set_problem(rv_controller* c, const char* problem_path, uint32_t time_frame,
                                         const char* primary_field, const char* secondary_field)
{
    C_TRY
    if (!c || !problem_path)
        return bad("rv_controller_set_problem: bad arguments");
    c->c->setProblem(problem_path, time_frame, primary_field ? primary_field : "",
                     secondary_field ? secondary_field : "");
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_field_count(const rv_controller* c, uint32_t* count)
{
    if (!c || !count)
        return bad("rv_controller_field_count: bad arguments");
    *count = (uint32_t)c->c->fields().size();
    return RV_OK;
}

extern "C" int rv_controller_field_info(const rv_controller* c, uint32_t i, rv_field_info* info)
{
    if (!c || !info)
        return bad("rv_controller_field_info: bad arguments");
    const auto f = c->c->fields();
    if (i >= f.size())
        return bad("rv_controller_field_info: no such field");
    std::memset(info, 0, sizeof *info);
    std::strncpy(info->name, f[i].name.c_str(), sizeof info->name - 1);
    info->min = f[i].range[0];
    info->max = f[i].range[1];
    info->animated = f[i].animated ? 1 : 0;
    return RV_OK;
}

extern "C" int rv_controller_time_frame_count(const rv_controller* c, uint32_t* count)
{
    if (!c || !count)
        return bad("rv_controller_time_frame_count: bad arguments");
    *count = c->c->timeFrameCount();
    return RV_OK;
}

extern "C" int rv_controller_set_input(rv_controller* c, const rv_render_state* s)
{
    C_TRY
    if (!c || !s || !s->tf || s->tf_n == 0 || (s->iso_level_count && !s->iso_levels_rel))
        return bad("rv_controller_set_input: bad arguments");
    revise::RenderSharedState st;
    if (s->view16)
        std::copy(s->view16, s->view16 + 16, st.cameraTransform.begin());
    else
        st.cameraTransform = c->c->cameraTransform();
    st.viewportSize = {s->vp_w, s->vp_h};
    st.fovY = s->fov_y_deg;
    st.renderQuality = s->quality;
    st.threshold = s->threshold_rel;
    st.alphaCutoff = s->alpha_cutoff;
    st.backgroundColor = {s->background[0], s->background[1], s->background[2]};
    st.colorTable.assign(s->tf, s->tf + 4ull * s->tf_n);
    st.fieldMode = s->field_mode;
    st.isosurfaceLevel = s->iso_level_rel;
    st.isosurfaceOpacity = s->iso_opacity;
    if (s->iso_level_count)
        st.isosurfaceLevels.assign(s->iso_levels_rel, s->iso_levels_rel + s->iso_level_count);
    st.exact = s->exact != 0;
    c->c->setInput(st);
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_set_camera(rv_controller* c, const float* view16)
{
    C_TRY
    if (!c || !view16)
        return bad("rv_controller_set_camera: bad arguments");
    std::array<float, 16> m;
    std::copy(view16, view16 + 16, m.begin());
    c->c->setCameraTransform(m);
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_camera(const rv_controller* c, float* view16)
{
    if (!c || !view16)
        return bad("rv_controller_camera: bad arguments");
    const auto m = c->c->cameraTransform();
    std::copy(m.begin(), m.end(), view16);
    return RV_OK;
}

extern "C" int rv_controller_reset_camera(rv_controller* c)
{
    C_TRY
    if (!c)
        return bad("rv_controller_reset_camera: null controller");
    c->c->resetCameraTransform();
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_set_render_level(rv_controller* c, int32_t level)
{
    C_TRY
    if (!c)
        return bad("rv_controller_set_render_level: null controller");
    c->c->setRenderLevel(level);
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_start(rv_controller* c)
{
    C_TRY
    if (!c)
        return bad("rv_controller_start: null controller");
    c->c->start();
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_frame_output(rv_controller* c, int* shmid, uint64_t* frame_size,
                                          uint32_t* width, uint32_t* height)
{
    C_TRY
    if (!c || !shmid || !frame_size || !width || !height)
        return bad("rv_controller_frame_output: bad arguments");
    const revise::FrameOutput o = c->c->frameOutput();
    *shmid = o.shmem->shmid();
    *frame_size = o.frameSize;
    *width = (uint32_t)o.frameWidth;
    *height = (uint32_t)o.frameHeight;
    return RV_OK;
    C_CATCH
}

extern "C" int rv_controller_wait_idle(rv_controller* c, uint32_t timeout_ms, int* idle, uint64_t* frames)
{
    if (!c || !idle)
        return bad("rv_controller_wait_idle: bad arguments");
    uint64_t n = 0;
    *idle = c->c->waitIdle(std::chrono::milliseconds(timeout_ms), &n) ? 1 : 0;
    if (frames)
        *frames = n;
    return RV_OK;
}

extern "C" int rv_controller_status(const rv_controller* c, char* message, uint64_t cap)
{
    if (!c)
        return bad("rv_controller_status: null controller");
    std::string m;
    const int st = c->c->status(&m);
    if (message && cap) {
        const size_t n = std::min<size_t>(m.size(), cap - 1);
        std::memcpy(message, m.data(), n);
        message[n] = 0;
    }
    return st;
}

extern "C" void rv_controller_destroy(rv_controller* c) { delete c; }
