// Drop-in C++ worker layer (see include/revise/vs_worker_hip.hpp) and its
// C ABI wrappers.  Host code only: every device operation goes through the C
// ABI of this library on the worker's own stream.
#include "../../include/revise/vs_worker_hip.hpp"

#include <hip/hip_runtime_api.h>

#include "share_staging.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

namespace rv {
void set_error(const std::string& msg);
int bound_rect_launch(const uint32_t* d_pixels, uint32_t w, uint32_t h, int* d_buf, hipStream_t s);
int crop_launch(const uint32_t* d_pixels, uint32_t w, uint32_t h, const int* d_buf, uint32_t* d_part,
                uint64_t cap, int* rect_out, hipStream_t s);
int bound_rect_read(const int* d_buf, int32_t rect[4], int* nonempty, hipStream_t s);
}

namespace revise {

namespace {
// The presentation threshold of a field mode (DenseFieldRenderer_cu.cu:418-433):
// the isosurface level for Isosurface, else the threshold (both relative).
float presentationLevel(const RenderSharedState& st)
{
    return st.fieldMode == RV_FIELD_ISOSURFACE || st.fieldMode == RV_FIELD_VALUE_ON_ISOSURFACE
               ? st.isosurfaceLevel
               : st.threshold;
}

// The modes that colour by the secondary field (DenseFieldRenderer_cu.cu:290-292).
bool hasSecondaryField(const RenderSharedState& st)
{
    return st.fieldMode == RV_FIELD_VALUE_ON_ISOSURFACE || st.fieldMode == RV_FIELD_VALUE_ON_ISOSURFACES;
}
} // namespace

void throwIfFailed(int status)
{
    if (status != RV_OK)
        throw HipError(status, rv_last_error());
}

static void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw HipError(e == hipErrorOutOfMemory ? RV_ERR_OOM : RV_ERR_HIP,
                       std::string(what) + ": " + hipGetErrorString(e));
}

static hipStream_t hs(rv_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- field service

struct FieldServiceHip::Plan {
    unsigned depth = 0;
    std::vector<uint16_t> n2i; // the shape, to tell hash collisions apart
    void* plan = nullptr;
    ~Plan()
    {
        if (plan)
            (void)rv_interp_plan_free(plan);
    }
};

struct FieldServiceHip::Subtree {
    unsigned depth = 0;
    uint64_t n = 0;
    float* sparse = nullptr;
    float* sparse2 = nullptr; // secondary field values, or null
    std::shared_ptr<Plan> plan; // gather plan of the subtree's shape
    ~Subtree()
    {
        if (sparse)
            (void)hipFree(sparse);
        if (sparse2)
            (void)hipFree(sparse2);
    }
};

FieldServiceHip::FieldServiceHip() = default;
FieldServiceHip::~FieldServiceHip() = default;

// The plan of a subtree shape: a plan depends on (depth, n2i) only, so
// subtrees of one shape (every full subtree of a depth, say) share it, as
// the reference shares BlockTreeNodes per subtree (BlockTreeFieldService.hpp:
// 133-137).  FNV-1a over the shape finds candidates; the n2i is compared.
std::shared_ptr<FieldServiceHip::Plan> FieldServiceHip::planFor(unsigned depth, const uint16_t* n2i,
                                                                uint64_t n)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xffu;
            h *= 1099511628211ull;
        }
    };
    mix(depth);
    mix(n);
    for (uint64_t i = 0; i < 3 * n; i += 4) {
        uint64_t v = 0;
        for (uint64_t k = i; k < std::min<uint64_t>(i + 4, 3 * n); ++k)
            v |= (uint64_t)n2i[k] << (16 * (k - i));
        mix(v);
    }
    auto range = m_plans.equal_range(h);
    for (auto it = range.first; it != range.second;) {
        std::shared_ptr<Plan> pl = it->second.lock();
        if (!pl) {
            it = m_plans.erase(it);
            continue;
        }
        if (pl->depth == depth && pl->n2i.size() == 3 * n &&
            std::equal(pl->n2i.begin(), pl->n2i.end(), n2i))
            return pl;
        ++it;
    }
    auto pl = std::make_shared<Plan>();
    pl->depth = depth;
    pl->n2i.assign(n2i, n2i + 3 * n);
    throwIfFailed(rv_interp_plan_build(n2i, n, depth, &pl->plan, nullptr));
    m_plans.emplace(h, pl);
    return pl;
}

void FieldServiceHip::addSubtree(unsigned level, const std::array<uint32_t, 3>& index,
                                 unsigned depth, const uint16_t* n2i, uint64_t n, const float* sparse)
{
    if (depth > 8)
        throw HipError(RV_ERR_INVALID, "addSubtree: depth must be <= 8");
    auto st = std::make_unique<Subtree>();
    st->depth = depth;
    st->n = n;
    if (n) {
        hipCheck(hipMalloc((void**)&st->sparse, 4 * n), "hipMalloc");
        hipCheck(hipMemcpy(st->sparse, sparse, 4 * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    st->plan = planFor(depth, n2i, n);
    m_subtrees[{level, index[0], index[1], index[2]}] = std::move(st);
    ++m_version;
}

void FieldServiceHip::setFieldRange(float fmin, float fmax)
{
    m_fmin = fmin;
    m_fmax = fmax;
}

void FieldServiceHip::addSecondary(unsigned level, const std::array<uint32_t, 3>& i, const float* sparse)
{
    auto it = m_subtrees.find({level, i[0], i[1], i[2]});
    if (it == m_subtrees.end())
        throw HipError(RV_ERR_INVALID, "addSecondary: no such subtree (register it first)");
    Subtree& st = *it->second;
    if (st.n) {
        if (!st.sparse2)
            hipCheck(hipMalloc((void**)&st.sparse2, 4 * st.n), "hipMalloc");
        hipCheck(hipMemcpy(st.sparse2, sparse, 4 * st.n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    ++m_version;
}

void FieldServiceHip::setSecondaryRange(float fmin, float fmax)
{
    m_smin = fmin;
    m_smax = fmax;
}

bool FieldServiceHip::hasSubtree(unsigned level, const std::array<uint32_t, 3>& i) const
{
    return m_subtrees.count({level, i[0], i[1], i[2]}) != 0;
}

unsigned FieldServiceHip::subtreeDepth(unsigned level, const std::array<uint32_t, 3>& i) const
{
    auto it = m_subtrees.find({level, i[0], i[1], i[2]});
    if (it == m_subtrees.end())
        throw std::range_error("Block does not exist");
    return it->second->depth;
}

rv_brick_job FieldServiceHip::job(unsigned level, const std::array<uint32_t, 3>& i, float* d_dense,
                                  rv_stream_t stream, bool secondary)
{
    auto it = m_subtrees.find({level, i[0], i[1], i[2]});
    if (it == m_subtrees.end())
        throw std::range_error("Block does not exist");
    const Subtree& st = *it->second;
    (void)stream;
    if (secondary && st.n && !st.sparse2)
        throw HipError(RV_ERR_INVALID, "renderScenePart: a block has no secondary field values");
    rv_brick_job j{};
    j.sparse = secondary ? st.sparse2 : st.sparse;
    j.plan = st.plan->plan;
    j.dense = d_dense;
    j.depth = st.depth;
    return j;
}

void FieldServiceHip::interpolate(unsigned level, const std::array<uint32_t, 3>& i, float* d_dense,
                                  rv_stream_t stream)
{
    auto it = m_subtrees.find({level, i[0], i[1], i[2]});
    if (it == m_subtrees.end())
        throw std::range_error("Block does not exist");
    const Subtree& st = *it->second;
    throwIfFailed(rv_interp_brick_planned(st.sparse, st.plan->plan, st.depth, d_dense, stream));
}

void FieldServiceHip::interpolateSecondary(unsigned level, const std::array<uint32_t, 3>& i,
                                           float* d_dense, rv_stream_t stream)
{
    auto it = m_subtrees.find({level, i[0], i[1], i[2]});
    if (it == m_subtrees.end())
        throw std::range_error("Block does not exist");
    const Subtree& st = *it->second;
    if (st.n && !st.sparse2)
        throw HipError(RV_ERR_INVALID, "renderDenseField: the block has no secondary field values");
    throwIfFailed(rv_interp_brick_planned(st.sparse2, st.plan->plan, st.depth, d_dense, stream));
}

// ---------------------------------------------------------------- renderer

DenseFieldRendererHip::DenseFieldRendererHip()
{
    hipStream_t s;
    hipCheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    m_stream = reinterpret_cast<rv_stream_t>(s);
}

DenseFieldRendererHip::~DenseFieldRendererHip()
{
    if (m_tf)
        (void)hipFree(m_tf);
    if (m_brick)
        (void)hipFree(m_brick);
    if (m_brick2)
        (void)hipFree(m_brick2);
    if (m_stream)
        (void)hipStreamDestroy(hs(m_stream));
}

void DenseFieldRendererHip::setRenderSharedState(const RenderSharedState* s)
{
    if (s != m_state) {
        m_state = s;
        m_tfDirty = true;
        m_lastBlock = {~0u, 0, 0, 0};
        m_lastBlock2 = {~0u, 0, 0, 0};
    }
}

void DenseFieldRendererHip::clearViewport(uint32_t* d_viewport) const
{
    throwIfFailed(rv_clear_viewport_mode(d_viewport, m_state->viewportSize[0],
                                         m_state->viewportSize[1], m_state->fieldMode, m_stream));
}

void DenseFieldRendererHip::renderDenseField(uint32_t* vp, unsigned level,
                                             const std::array<uint32_t, 3>& index,
                                             const std::atomic<bool>& isCancelled)
{
    renderDenseField(vp, level, index, [&] { return isCancelled.load(); });
}

void DenseFieldRendererHip::renderDenseField(uint32_t* vp, unsigned level,
                                             const std::array<uint32_t, 3>& index,
                                             const std::function<bool()>& isCancelled)
{
    const RenderSharedState& st = *m_state;
    FieldServiceHip& fs = *st.fieldSvc;
    const unsigned depth = fs.subtreeDepth(level, index);
    // makeField (DenseFieldRenderer_cu.cu:504-545): one brick, re-interpolated
    // when the block changes.
    const uint64_t nodes = rv_brick_nodes(depth);
    if (nodes > m_brickNodes) {
        if (m_brick)
            (void)hipFree(m_brick);
        m_brick = nullptr;
        hipCheck(hipMalloc((void**)&m_brick, 4 * nodes), "hipMalloc");
        if (m_brick2)
            (void)hipFree(m_brick2);
        m_brick2 = nullptr;
        hipCheck(hipMalloc((void**)&m_brick2, 4 * nodes), "hipMalloc");
        m_brickNodes = nodes;
        m_lastBlock = {~0u, 0, 0, 0};
        m_lastBlock2 = {~0u, 0, 0, 0};
    }
    // the brick is also stale when the field service changed since it was
    // interpolated (new values of the block, or of its secondary field; the
    // reference marks its fields dirty likewise, DenseFieldRenderer_cu.cu:238-259)
    const std::array<uint32_t, 4> key{level, index[0], index[1], index[2]};
    if (key != m_lastBlock || m_lastVersion != fs.version()) {
        fs.interpolate(level, index, m_brick, m_stream);
        m_lastBlock = key;
        m_lastVersion = fs.version();
    }
    const bool second = hasSecondaryField(st);
    if (second && (key != m_lastBlock2 || m_lastVersion2 != fs.version())) {
        fs.interpolateSecondary(level, index, m_brick2, m_stream);
        m_lastBlock2 = key;
        m_lastVersion2 = fs.version();
    }
    ensureColorTable();
    if (isCancelled())
        return;
    const auto range = fs.fieldRange();
    rv_block_params p;
    int live = 0;
    throwIfFailed(rv_block_params_setup(st.cameraTransform.data(), st.viewportSize[0],
                                        st.viewportSize[1], st.fovY, level, index.data(), depth,
                                        range[0], range[1], presentationLevel(st), st.renderQuality,
                                        st.alphaCutoff, &p, &live));
    if (!live)
        return;
    p.flags = st.exact ? RV_RENDER_EXACT : 0u;
    const auto srange = fs.secondaryRange();
    throwIfFailed(rv_render_block_secondary(&p, st.fieldMode, st.isosurfaceOpacity,
                                            st.isosurfaceLevels.data(),
                                            (uint32_t)st.isosurfaceLevels.size(), m_brick,
                                            second ? m_brick2 : nullptr, srange[0], srange[1], m_tf,
                                            m_tfN, vp, m_stream));
}

// makeColorTransferTexture (DenseFieldRenderer_cu.cu:573-589): the table is
// uploaded when it changed.
void DenseFieldRendererHip::ensureColorTable()
{
    const RenderSharedState& st = *m_state;
    if (m_tfDirty) {
        const uint32_t n = (uint32_t)(st.colorTable.size() / 4);
        if (n == 0)
            throw HipError(RV_ERR_INVALID, "renderDenseField: empty colour table");
        if (n != m_tfN) {
            if (m_tf)
                (void)hipFree(m_tf);
            m_tf = nullptr;
            hipCheck(hipMalloc((void**)&m_tf, 16 * n), "hipMalloc");
            m_tfN = n;
        }
        hipCheck(hipMemcpyAsync(m_tf, st.colorTable.data(), 16 * n, hipMemcpyHostToDevice,
                                hs(m_stream)),
                 "hipMemcpyAsync");
        m_tfDirty = false;
    }
}

// ---------------------------------------------------------------- worker

// Device state of the fused share path.  Bricks are kept per block (level,
// index), each with its float count, so a block whose subtree grows deeper
// (a later addSubtree) gets a larger brick.  The batched interpolation
// handles are kept per (level, index box, secondary) -- composeFrame and the
// level-of-detail loop render several boxes per frame -- up to kMaxBatches,
// the least recently used dropped first; a field-service version change
// drops them all.
struct VsWorkerHip::ShareCache {
    static constexpr size_t kMaxBatches = 64;
    struct Brick {
        float* p = nullptr;
        uint64_t floats = 0;
    };
    struct Batch {
        rv_interp_batch* batch = nullptr;
        uint64_t used = 0;
    };
    uint64_t version = ~0ull;
    uint64_t clock = 0;
    std::map<std::array<uint32_t, 4>, Brick> bricks;
    std::map<std::array<uint32_t, 4>, Brick> bricks2; // secondary field (ValueOnIsosurface(s))
    std::map<std::array<uint32_t, 8>, Batch> batches;
    rv::ShareStaging table; // the block table and screen bins of the share launch
    void dropBatches()
    {
        for (auto& kv : batches)
            rv_interp_batch_destroy(kv.second.batch);
        batches.clear();
    }
    // The block's brick of at least `floats` floats; *grown = true when it
    // was (re)allocated.
    float* brick(std::map<std::array<uint32_t, 4>, Brick>& m, const std::array<uint32_t, 4>& key,
                 uint64_t floats, bool* grown)
    {
        Brick& b = m[key];
        if (b.floats < floats) {
            if (b.p)
                (void)hipFree(b.p);
            b.p = nullptr;
            b.floats = 0;
            hipCheck(hipMalloc((void**)&b.p, 4 * floats), "hipMalloc");
            b.floats = floats;
            *grown = true;
        }
        return b.p;
    }
    ~ShareCache()
    {
        dropBatches();
        for (auto& kv : bricks)
            (void)hipFree(kv.second.p);
        for (auto& kv : bricks2)
            (void)hipFree(kv.second.p);
        if (table.copied)
            (void)hipEventSynchronize(table.copied);
        table.release();
    }
};

VsWorkerHip::VsWorkerHip() = default;

VsWorkerHip::~VsWorkerHip()
{
    if (m_rectBuf)
        (void)hipFree(m_rectBuf);
    if (m_rectHost)
        (void)hipHostFree(m_rectHost);
    if (m_viewport)
        (void)hipFree(m_viewport);
    if (m_part)
        (void)hipFree(m_part);
}

void VsWorkerHip::initialize(const std::string&, bool, const std::string&, size_t threadId)
{
    int n = 0;
    hipCheck(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (n <= 0)
        throw HipError(RV_ERR_HIP, "no GPU");
    hipCheck(hipSetDevice((int)(threadId % (size_t)n)), "hipSetDevice"); // selectGpu.cu:6-9
}

void VsWorkerHip::setRenderSharedState(const RenderSharedState* s)
{
    m_state = s;
    m_renderer.setRenderSharedState(s);
}

void VsWorkerHip::updateState(unsigned flags)
{
    // VsWorker_cu::updateState (VsWorker_cu.cpp:122-152) does nothing for the
    // CUDA path; here a colour-table change must reach the device table.
    constexpr unsigned kUpdateColorTransfer = 1u << 7; // VsWorkerInterface.hpp:41-57
    if (flags & kUpdateColorTransfer)
        m_renderer.markColorTableDirty();
}

RgbaImagePart VsWorkerHip::renderScenePart(const SubtreeSetDescription& sub,
                                           const std::atomic<bool>& isCancelled)
{
    return renderScenePart(sub, [&] { return isCancelled.load(); });
}

RgbaImagePart VsWorkerHip::renderScenePart(const SubtreeSetDescription& sub,
                                           const std::function<bool()>& isCancelled,
                                           uint32_t* hostDst, uint64_t hostCap)
{
    RgbaImagePart part;
    m_partRect = {0, 0, 0, 0};
    m_cancelled = false;
    if (sub.empty())
        return part;
    const RenderSharedState& st = *m_state;
    const uint32_t W = st.viewportSize[0], H = st.viewportSize[1];
    const uint64_t px = (uint64_t)W * H;
    if (px > m_viewportPixels) {
        if (m_viewport)
            (void)hipFree(m_viewport);
        m_viewport = nullptr;
        hipCheck(hipMalloc((void**)&m_viewport, 4 * px), "hipMalloc");
        m_viewportPixels = px;
    }
    // the fused path's share launch writes the cleared viewport itself (it
    // is cleared here only when that launch does not happen)
    if (!m_fused)
        m_renderer.clearViewport(m_viewport);
    // the fused path interpolates the share's bricks first: the block order
    // and parameters are then computed on the host while the GPU works
    if (m_fused) {
        interpolateFused(sub, isCancelled);
        if (m_cancelled) {
            hipCheck(hipStreamSynchronize(hs(stream())), "hipStreamSynchronize");
            return part;
        }
    }
    // MacroBlock over the index box + BackToFrontOrder from the eye
    // (VsWorker_cu.cpp:254-262)
    float cmin[3], cmax[3], eye[3];
    throwIfFailed(rv_index_box_coords(sub.level, sub.boxMin.data(), sub.boxMax.data(), cmin, cmax));
    throwIfFailed(rv_eye_from_view(st.cameraTransform.data(), eye));
    uint64_t count = 1;
    for (int a = 0; a < 3; ++a)
        count *= sub.boxMax[a] - sub.boxMin[a];
    std::vector<uint32_t> order(3 * count);
    uint64_t got = 0;
    throwIfFailed(rv_back_to_front(sub.boxMin.data(), sub.boxMax.data(), cmin, cmax, eye,
                                   order.data(), count, &got));
    m_rectDone = false;
    if (m_fused) {
        renderFused(sub, order, got, isCancelled);
    } else {
        for (uint64_t i = 0; i < got; ++i) {
            if (isCancelled()) {
                m_cancelled = true;
                break;
            }
            const std::array<uint32_t, 3> idx{order[3 * i], order[3 * i + 1], order[3 * i + 2]};
            m_renderer.renderDenseField(m_viewport, sub.level, idx, isCancelled);
        }
    }
    if (m_cancelled || isCancelled()) {
        m_cancelled = true;
        markStagingCopied();
        hipCheck(hipStreamSynchronize(hs(stream())), "hipStreamSynchronize");
        return part;
    }
    if (m_fused && !m_rectDone)
        m_renderer.clearViewport(m_viewport); // no share launch (no live block)
    // Crop to the non-transparent rectangle (VsWorker_cu.cpp:279-306), on
    // the device: the rectangle and the packed part are computed in stream
    // order.  A pinned host destination receives the part and the rectangle
    // from the crop kernel itself (one synchronisation); otherwise the part
    // is packed in device memory, the rectangle read, and the part copied.
    int32_t rect[4];
    int nonempty = 0;
    if (!m_rectBuf)
        hipCheck(hipMalloc((void**)&m_rectBuf, 16), "hipMalloc");
    if (!m_rectDone)
        throwIfFailed(rv::bound_rect_launch(m_viewport, W, H, m_rectBuf, hs(stream())));
    void* mapped = hostDst ? mappedHost(hostDst) : nullptr;
    if (mapped) {
        if (!m_rectHost) {
            hipCheck(hipHostMalloc((void**)&m_rectHost, 16, hipHostMallocDefault), "hipHostMalloc");
            hipCheck(hipHostGetDevicePointer((void**)&m_rectHostDev, m_rectHost, 0), "hipHostGetDevicePointer");
        }
        throwIfFailed(rv::crop_launch(m_viewport, W, H, m_rectBuf, static_cast<uint32_t*>(mapped), hostCap,
                                      m_rectHostDev, hs(stream())));
        markStagingCopied();
        hipCheck(hipStreamSynchronize(hs(stream())), "hipStreamSynchronize");
        m_partRect = {0, 0, 0, 0}; // the part is not kept on the device
        if (m_rectHost[2] < 0)
            return part;
        rect[0] = m_rectHost[0];
        rect[1] = m_rectHost[1];
        rect[2] = m_rectHost[2] - m_rectHost[0] + 1;
        rect[3] = m_rectHost[3] - m_rectHost[1] + 1;
        if ((uint64_t)rect[2] * rect[3] > hostCap)
            throw HipError(RV_ERR_INVALID, "renderScenePart: output capacity too small");
        part.origin = {rect[0], rect[1]};
        part.size = {rect[2], rect[3]};
        return part;
    }
    if (px > m_partCap) {
        if (m_part)
            (void)hipFree(m_part);
        m_part = nullptr;
        m_partCap = 0;
        hipCheck(hipMalloc((void**)&m_part, 4 * px), "hipMalloc");
        m_partCap = px;
    }
    throwIfFailed(rv::crop_launch(m_viewport, W, H, m_rectBuf, m_part, m_partCap, nullptr, hs(stream())));
    markStagingCopied();
    throwIfFailed(rv::bound_rect_read(m_rectBuf, rect, &nonempty, hs(stream())));
    if (!nonempty)
        return part;
    const uint64_t n = (uint64_t)rect[2] * rect[3];
    part.origin = {rect[0], rect[1]};
    part.size = {rect[2], rect[3]};
    if (hostDst) {
        if (n > hostCap)
            throw HipError(RV_ERR_INVALID, "renderScenePart: output capacity too small");
        hipCheck(hipMemcpyAsync(hostDst, m_part, 4 * n, hipMemcpyDeviceToHost, hs(stream())),
                 "hipMemcpyAsync");
    } else {
        part.bits.resize(4 * n);
        hipCheck(hipMemcpyAsync(part.bits.data(), m_part, 4 * n, hipMemcpyDeviceToHost, hs(stream())),
                 "hipMemcpyAsync");
    }
    hipCheck(hipStreamSynchronize(hs(stream())), "hipStreamSynchronize");
    m_partRect = {rect[0], rect[1], rect[2], rect[3]};
    return part;
}

void* VsWorkerHip::mappedHost(void* p)
{
    // asked on every call: a buffer freed and another allocated at its
    // address may differ in kind
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) == hipSuccess && a.type == hipMemoryTypeHost && a.devicePointer)
        return a.devicePointer;
    (void)hipGetLastError(); // pageable memory: clear the query's error
    return nullptr;
}

void VsWorkerHip::interpolateFused(const SubtreeSetDescription& sub, const std::function<bool()>& isCancelled)
{
    const RenderSharedState& st = *m_state;
    FieldServiceHip& fs = *st.fieldSvc;
    rv_stream_t s = stream();
    if (!m_share)
        m_share = std::make_unique<ShareCache>();
    ShareCache& sc = *m_share;
    const bool second = hasSecondaryField(st);
    const std::array<uint32_t, 8> key{sub.level,     sub.boxMin[0], sub.boxMin[1], sub.boxMin[2],
                                      sub.boxMax[0], sub.boxMax[1], sub.boxMax[2], second ? 1u : 0u};
    if (sc.version != fs.version()) {
        sc.dropBatches();
        sc.version = fs.version();
    }
    auto hit = sc.batches.find(key);
    if (hit == sc.batches.end()) {
        // the box's blocks in index order (the interpolation needs no order)
        std::vector<std::array<uint32_t, 3>> blocks;
        for (uint32_t z = sub.boxMin[2]; z < sub.boxMax[2]; ++z)
            for (uint32_t y = sub.boxMin[1]; y < sub.boxMax[1]; ++y)
                for (uint32_t x = sub.boxMin[0]; x < sub.boxMax[0]; ++x)
                    blocks.push_back({x, y, z});
        std::vector<rv_brick_job> jobs;
        jobs.reserve(second ? 2 * blocks.size() : blocks.size());
        bool grown = false;
        for (const auto& idx : blocks) {
            const std::array<uint32_t, 4> bkey{sub.level, idx[0], idx[1], idx[2]};
            // the share's bricks in micro-bricks (fewer cache lines per
            // ray-cast corner load; the same values)
            const uint64_t floats = rv_brick_floats(fs.subtreeDepth(sub.level, idx), RV_BRICK_TILED4);
            jobs.push_back(fs.job(sub.level, idx, sc.brick(sc.bricks, bkey, floats, &grown), s));
            jobs.back().layout = RV_BRICK_TILED4;
            if (second) {
                jobs.push_back(fs.job(sub.level, idx, sc.brick(sc.bricks2, bkey, floats, &grown), s, true));
                jobs.back().layout = RV_BRICK_TILED4;
            }
        }
        // a reallocated brick invalidates every cached batch that wrote it
        if (grown)
            sc.dropBatches();
        if (sc.batches.size() >= ShareCache::kMaxBatches) {
            auto lru = sc.batches.begin();
            for (auto it = sc.batches.begin(); it != sc.batches.end(); ++it)
                if (it->second.used < lru->second.used)
                    lru = it;
            rv_interp_batch_destroy(lru->second.batch);
            sc.batches.erase(lru);
        }
        ShareCache::Batch b;
        throwIfFailed(rv_interp_batch_create(jobs.data(), (uint32_t)jobs.size(), &b.batch));
        hit = sc.batches.emplace(key, b).first;
    }
    hit->second.used = ++sc.clock;
    if (isCancelled()) {
        m_cancelled = true;
        return;
    }
    throwIfFailed(rv_interp_batch_run(hit->second.batch, s));
}

void VsWorkerHip::renderFused(const SubtreeSetDescription& sub, const std::vector<uint32_t>& order,
                              uint64_t count, const std::function<bool()>& isCancelled)
{
    const RenderSharedState& st = *m_state;
    FieldServiceHip& fs = *st.fieldSvc;
    rv_stream_t s = stream();
    ShareCache& sc = *m_share;
    const bool second = hasSecondaryField(st);
    // colour table (as the per-block renderer keeps it)
    m_renderer.ensureColorTable();
    const auto range = fs.fieldRange();
    std::vector<rv_block_params> params;
    std::vector<const float*> bricks, bricks2;
    params.reserve(count);
    bricks.reserve(count);
    for (uint64_t i = 0; i < count; ++i) {
        const std::array<uint32_t, 3> idx{order[3 * i], order[3 * i + 1], order[3 * i + 2]};
        rv_block_params p;
        int live = 0;
        throwIfFailed(rv_block_params_setup(st.cameraTransform.data(), st.viewportSize[0],
                                            st.viewportSize[1], st.fovY, sub.level, idx.data(),
                                            fs.subtreeDepth(sub.level, idx), range[0], range[1],
                                            presentationLevel(st), st.renderQuality, st.alphaCutoff,
                                            &p, &live));
        if (!live)
            continue;
        p.flags = st.exact ? RV_RENDER_EXACT : 0u;
        const std::array<uint32_t, 4> bkey{sub.level, idx[0], idx[1], idx[2]};
        params.push_back(p);
        bricks.push_back(sc.bricks[bkey].p);
        if (second)
            bricks2.push_back(sc.bricks2[bkey].p);
    }
    if (isCancelled()) {
        m_cancelled = true;
        return;
    }
    const auto srange = fs.secondaryRange();
    // the table and the screen bins go through pinned staging memory,
    // asynchronously (the next part waits for this copy before rewriting it)
    // the share launch also leaves the crop rectangle in m_rectBuf
    if (!m_rectBuf)
        hipCheck(hipMalloc((void**)&m_rectBuf, 16), "hipMalloc");
    throwIfFailed(rv::render_share_staged(
        params.data(), bricks.data(), second ? bricks2.data() : nullptr, (uint32_t)params.size(),
        RV_BRICK_TILED4, st.fieldMode, st.isosurfaceOpacity, st.isosurfaceLevels.data(),
        (uint32_t)st.isosurfaceLevels.size(), srange[0], srange[1], m_renderer.colorTable(),
        m_renderer.colorTableSize(), m_viewport, st.viewportSize[0], st.viewportSize[1], sc.table, s, m_rectBuf,
        &m_rectDone, st.fieldMode == RV_FIELD_MAX_INTENSITY_PROJECTION ? 0x7fc00000u : 0u, false));
}

// The staged table's copy event of the fused path, recorded after the
// part's last launch rather than between the table copy and the share
// launch (where it would delay the share kernel).
void VsWorkerHip::markStagingCopied()
{
    if (m_fused && m_share)
        hipCheck(rv::share_staging_copied(m_share->table, hs(stream())), "hipEventRecord");
}

const uint32_t* VsWorkerHip::devicePart(std::array<int, 4>& rect) const
{
    rect = m_partRect;
    return m_part;
}

// ---------------------------------------------------------------- compositing

namespace {

struct DevPart {
    int x = 0, y = 0, w = 0, h = 0;
    uint32_t* pix = nullptr;
};

struct DevBuffers {
    std::vector<uint32_t*> all;
    ~DevBuffers()
    {
        for (auto* p : all)
            (void)hipFree(p);
    }
    uint32_t* alloc(uint64_t n)
    {
        uint32_t* p = nullptr;
        hipCheck(hipMalloc((void**)&p, 4 * std::max<uint64_t>(n, 1)), "hipMalloc");
        all.push_back(p);
        return p;
    }
};

} // namespace

std::vector<uint32_t> composeFrame(VsWorkerHip& worker, unsigned level, unsigned splitCount,
                                   const std::function<bool()>& isCancelled)
{
    const RenderSharedState& st = *worker.renderSharedState();
    const uint32_t W = st.viewportSize[0], H = st.viewportSize[1];
    const bool mip = st.fieldMode == RV_FIELD_MAX_INTENSITY_PROJECTION;
    float eye[3];
    throwIfFailed(rv_eye_from_view(st.cameraTransform.data(), eye));
    std::vector<uint32_t> boxes(6ull * (splitCount + 8));
    uint64_t nb = 0;
    throwIfFailed(rv_index_cube_split_sorted(level, splitCount, eye, boxes.data(), splitCount + 8, &nb));
    DevBuffers bufs;
    hipStream_t s = hs(worker.stream());
    std::vector<DevPart> layer;
    for (uint64_t i = 0; i < nb; ++i) {
        SubtreeSetDescription sub;
        sub.level = level;
        for (int a = 0; a < 3; ++a) {
            sub.boxMin[a] = boxes[6 * i + a];
            sub.boxMax[a] = boxes[6 * i + 3 + a];
        }
        worker.renderScenePart(sub, isCancelled);
        if (worker.lastCancelled())
            return {};
        std::array<int, 4> r;
        const uint32_t* d = worker.devicePart(r);
        DevPart p;
        p.x = r[0];
        p.y = r[1];
        p.w = r[2];
        p.h = r[3];
        if (p.w * p.h > 0) {
            p.pix = bufs.alloc((uint64_t)p.w * p.h);
            hipCheck(hipMemcpyAsync(p.pix, d, 4ull * p.w * p.h, hipMemcpyDeviceToDevice, s),
                     "hipMemcpyAsync");
        }
        layer.push_back(p);
    }
    // AssembleFunc (VsRenderer.cpp:403-473) over the RenderGraphGenerator tree
    // (RenderGraphGenerator.cpp:64-84): parts paired in order, the earlier
    // (farther) part copied, the later blended over it; the bounding box
    // includes both parts' origin and far corner even for empty parts.
    while (layer.size() > 1) {
        std::vector<DevPart> next;
        for (size_t i = 1; i < layer.size(); i += 2) {
            const DevPart& a = layer[i - 1];
            const DevPart& b = layer[i];
            int x0 = std::min({a.x, a.x + a.w, b.x, b.x + b.w});
            int y0 = std::min({a.y, a.y + a.h, b.y, b.y + b.h});
            int x1 = std::max({a.x, a.x + a.w, b.x, b.x + b.w});
            int y1 = std::max({a.y, a.y + a.h, b.y, b.y + b.h});
            DevPart m;
            m.x = x0;
            m.y = y0;
            m.w = x1 - x0;
            m.h = y1 - y0;
            m.pix = bufs.alloc((uint64_t)m.w * m.h);
            // max intensity parts hold float bits, NaN = none (VsRenderer.cpp:440-445)
            throwIfFailed(rv_fill_u32(m.pix, (uint64_t)m.w * m.h, mip ? 0x7fc00000u : 0u, worker.stream()));
            auto blend = [&](const DevPart& q, int copy) {
                if (mip)
                    throwIfFailed(rv_blend_image_part_max_intensity(
                        reinterpret_cast<float*>(m.pix), m.w, m.h, m.x, m.y,
                        reinterpret_cast<const float*>(q.pix), q.w, q.h, q.x, q.y, copy, worker.stream()));
                else
                    throwIfFailed(rv_blend_image_part(m.pix, m.w, m.h, m.x, m.y, q.pix, q.w, q.h, q.x,
                                                      q.y, copy, worker.stream()));
            };
            if (a.pix)
                blend(a, 1);
            if (b.pix)
                blend(b, 0);
            next.push_back(m);
        }
        if (layer.size() % 2)
            next.push_back(layer.back());
        layer.swap(next);
    }
    // FinalComposeFunc (VsRenderer.cpp:475-519)
    uint32_t* frame = bufs.alloc((uint64_t)W * H);
    throwIfFailed(rv_fill_u32(frame, (uint64_t)W * H,
                              rv_background_argb(st.backgroundColor[0], st.backgroundColor[1],
                                                 st.backgroundColor[2]),
                              worker.stream()));
    if (!layer.empty() && layer[0].pix) {
        if (mip) // makeMaxIntensityProjection (VsRenderer.cpp:542-565) with the colour table
            throwIfFailed(rv_mip_colorize(frame, W, H, reinterpret_cast<const float*>(layer[0].pix),
                                          layer[0].w, layer[0].h, layer[0].x, layer[0].y,
                                          worker.colorTable(), worker.colorTableSize(), worker.stream()));
        else
            throwIfFailed(rv_blend_image_part(frame, W, H, 0, 0, layer[0].pix, layer[0].w, layer[0].h,
                                              layer[0].x, layer[0].y, 0, worker.stream()));
    }
    std::vector<uint32_t> out((size_t)W * H);
    hipCheck(hipMemcpyAsync(out.data(), frame, 4ull * W * H, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
    hipCheck(hipStreamSynchronize(s), "hipStreamSynchronize");
    return out;
}

unsigned renderProgressive(VsWorkerHip& worker, unsigned levelMin, unsigned levelMax,
                           unsigned splitCount, const std::function<bool()>& isCancelled,
                           uint64_t& frameNumber, bool measureTime,
                           const std::function<void(const FrameOutputHeader&, const uint32_t*)>& writeFrame)
{
    unsigned written = 0;
    for (unsigned level = levelMin; level <= levelMax && !isCancelled(); ++level) {
        const auto t0 = std::chrono::steady_clock::now();
        const std::vector<uint32_t> frame = composeFrame(worker, level, splitCount, isCancelled);
        if (frame.empty() || isCancelled())
            break;
        FrameOutputHeader h;
        h.frameNumber = frameNumber++;
        if (measureTime) {
            const std::chrono::duration<double, std::milli> ms = std::chrono::steady_clock::now() - t0;
            h.durationMs = (uint64_t)std::llround(ms.count());
        }
        h.level = level;
        writeFrame(h, frame.data());
        ++written;
    }
    return written;
}

} // namespace revise

// ---------------------------------------------------------------- C ABI

struct rv_worker {
    revise::FieldServiceHip field;
    revise::RenderSharedState state;
    revise::VsWorkerHip worker;
};

#define W_TRY try {
#define W_CATCH                                                                        \
    }                                                                                  \
    catch (const revise::HipError& e)                                                  \
    {                                                                                  \
        rv::set_error(e.what());                                                       \
        return e.code;                                                                 \
    }                                                                                  \
    catch (const std::range_error& e)                                                  \
    {                                                                                  \
        rv::set_error(e.what());                                                       \
        return RV_ERR_INVALID;                                                         \
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

extern "C" int rv_worker_create(uint32_t thread_id, rv_worker** out)
{
    W_TRY
    if (!out) {
        rv::set_error("rv_worker_create: null out");
        return RV_ERR_INVALID;
    }
    *out = nullptr;
    auto w = std::make_unique<rv_worker>();
    w->worker.initialize("", false, "", thread_id);
    w->state.fieldSvc = &w->field;
    w->worker.setRenderSharedState(&w->state);
    *out = w.release();
    return RV_OK;
    W_CATCH
}

extern "C" int rv_worker_add_subtree(rv_worker* w, uint32_t level, const uint32_t index[3],
                                     uint32_t depth, const uint16_t* n2i, uint64_t n,
                                     const float* sparse)
{
    W_TRY
    if (!w || !index || (n && (!n2i || !sparse))) {
        rv::set_error("rv_worker_add_subtree: bad arguments");
        return RV_ERR_INVALID;
    }
    w->field.addSubtree(level, {index[0], index[1], index[2]}, depth, n2i, n, sparse);
    return RV_OK;
    W_CATCH
}

extern "C" int rv_worker_add_subtree_secondary(rv_worker* w, uint32_t level, const uint32_t index[3],
                                               const float* sparse)
{
    W_TRY
    if (!w || !index || !sparse) {
        rv::set_error("rv_worker_add_subtree_secondary: bad arguments");
        return RV_ERR_INVALID;
    }
    w->field.addSecondary(level, {index[0], index[1], index[2]}, sparse);
    return RV_OK;
    W_CATCH
}

extern "C" int rv_worker_set_state(rv_worker* w, const rv_render_state* s)
{
    W_TRY
    if (!w || !s || !s->view16 || !s->tf || s->tf_n == 0 || !(s->fmax > s->fmin)) {
        rv::set_error("rv_worker_set_state: bad arguments");
        return RV_ERR_INVALID;
    }
    revise::RenderSharedState& st = w->state;
    std::copy(s->view16, s->view16 + 16, st.cameraTransform.begin());
    st.viewportSize = {s->vp_w, s->vp_h};
    st.fovY = s->fov_y_deg;
    st.renderQuality = s->quality;
    st.threshold = s->threshold_rel;
    st.alphaCutoff = s->alpha_cutoff;
    st.backgroundColor = {s->background[0], s->background[1], s->background[2]};
    // the device table is refreshed only when the table changed (a new
    // camera alone uploads nothing)
    const bool tf_changed = st.colorTable.size() != 4ull * s->tf_n ||
                            !std::equal(st.colorTable.begin(), st.colorTable.end(), s->tf);
    if (tf_changed)
        st.colorTable.assign(s->tf, s->tf + 4ull * s->tf_n);
    st.fieldMode = s->field_mode;
    st.isosurfaceLevel = s->iso_level_rel;
    st.isosurfaceOpacity = s->iso_opacity;
    if (s->iso_level_count && !s->iso_levels_rel) {
        rv::set_error("rv_worker_set_state: null isosurface levels");
        return RV_ERR_INVALID;
    }
    st.isosurfaceLevels.assign(s->iso_levels_rel, s->iso_levels_rel + s->iso_level_count);
    st.exact = s->exact != 0;
    if ((s->field_mode == RV_FIELD_VALUE_ON_ISOSURFACE || s->field_mode == RV_FIELD_VALUE_ON_ISOSURFACES) &&
        !(s->smax > s->smin)) {
        rv::set_error("rv_worker_set_state: the secondary field range must be non-empty");
        return RV_ERR_INVALID;
    }
    w->field.setFieldRange(s->fmin, s->fmax);
    w->field.setSecondaryRange(s->smin, s->smax);
    w->worker.setRenderSharedState(&st);
    if (tf_changed)
        w->worker.updateState(1u << 7);
    return RV_OK;
    W_CATCH
}

extern "C" int rv_worker_render_scene_part(rv_worker* w, uint32_t level, const uint32_t bmin[3],
                                           const uint32_t bmax[3], const volatile int* cancel,
                                           int32_t rect[4], uint32_t* bits, uint64_t cap)
{
    W_TRY
    if (!w || !bmin || !bmax || !rect) {
        rv::set_error("rv_worker_render_scene_part: bad arguments");
        return RV_ERR_INVALID;
    }
    revise::SubtreeSetDescription sub;
    sub.level = level;
    for (int a = 0; a < 3; ++a) {
        sub.boxMin[a] = bmin[a];
        sub.boxMax[a] = bmax[a];
    }
    // the part goes straight into the caller's buffer
    auto part = w->worker.renderScenePart(sub, [cancel] { return cancel && *cancel != 0; }, bits,
                                          bits ? cap : 0);
    rect[0] = part.origin[0];
    rect[1] = part.origin[1];
    rect[2] = part.size[0];
    rect[3] = part.size[1];
    if (w->worker.lastCancelled()) {
        rv::set_error("rv_worker_render_scene_part: cancelled");
        return RV_ERR_CANCELLED;
    }
    if (!bits && (uint64_t)part.size[0] * part.size[1] > 0) {
        rv::set_error("rv_worker_render_scene_part: output capacity too small");
        return RV_ERR_INVALID;
    }
    return RV_OK;
    W_CATCH
}

extern "C" int rv_worker_compose_frame(rv_worker* w, uint32_t level, uint32_t split_count,
                                       uint32_t* frame, const volatile int* cancel)
{
    W_TRY
    if (!w || !frame || split_count == 0) {
        rv::set_error("rv_worker_compose_frame: bad arguments");
        return RV_ERR_INVALID;
    }
    auto out = revise::composeFrame(w->worker, level, split_count,
                                    [cancel] { return cancel && *cancel != 0; });
    if (out.empty()) {
        rv::set_error("rv_worker_compose_frame: cancelled");
        return RV_ERR_CANCELLED;
    }
    std::memcpy(frame, out.data(), 4 * out.size());
    return RV_OK;
    W_CATCH
}

static_assert(sizeof(revise::FrameOutputHeader) == sizeof(rv_frame_output_header),
              "frame output header layout");

extern "C" int rv_worker_render_progressive(rv_worker* w, uint32_t level_min, uint32_t level_max,
                                            uint32_t split_count, const volatile int* cancel, void* out,
                                            uint64_t out_size, uint64_t* frame_number, int measure_time,
                                            uint32_t* frames_written)
{
    W_TRY
    if (!w || !out || !frame_number || split_count == 0 || level_min > level_max) {
        rv::set_error("rv_worker_render_progressive: bad arguments");
        return RV_ERR_INVALID;
    }
    const uint64_t px = (uint64_t)w->state.viewportSize[0] * w->state.viewportSize[1];
    if (out_size < sizeof(rv_frame_output_header) + 4 * px) {
        rv::set_error("rv_worker_render_progressive: the output block cannot hold a frame");
        return RV_ERR_INVALID;
    }
    unsigned char* dst = static_cast<unsigned char*>(out);
    const unsigned n = revise::renderProgressive(
        w->worker, level_min, level_max, split_count, [cancel] { return cancel && *cancel != 0; },
        *frame_number, measure_time != 0,
        [&](const revise::FrameOutputHeader& h, const uint32_t* pixels) {
            std::memcpy(dst, &h, sizeof h);
            std::memcpy(dst + sizeof h, pixels, 4 * px);
        });
    if (frames_written)
        *frames_written = n;
    return RV_OK;
    W_CATCH
}

extern "C" int rv_frame_output_read(const void* in, uint64_t in_size, uint32_t vp_w, uint32_t vp_h,
                                    rv_frame_output_header* header, uint32_t* pixels)
{
    const uint64_t px = (uint64_t)vp_w * vp_h;
    if (!in || !header || in_size < sizeof(rv_frame_output_header) + (pixels ? 4 * px : 0)) {
        rv::set_error("rv_frame_output_read: bad arguments");
        return RV_ERR_INVALID;
    }
    const unsigned char* src = static_cast<const unsigned char*>(in);
    std::memcpy(header, src, sizeof *header);
    if (pixels)
        std::memcpy(pixels, src + sizeof *header, 4 * px);
    return RV_OK;
}

extern "C" int rv_worker_set_fused(rv_worker* w, int fused)
{
    if (!w) {
        rv::set_error("rv_worker_set_fused: null worker");
        return RV_ERR_INVALID;
    }
    w->worker.setFused(fused != 0);
    return RV_OK;
}

extern "C" void rv_worker_destroy(rv_worker* w) { delete w; }
