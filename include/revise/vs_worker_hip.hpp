// C++ drop-in layer of the HIP path, mirroring the reference's plug-in
// interfaces for the Argb mode:
//  - FieldServiceHip      : the part of BlockTreeFieldService<3> the renderer
//                           uses (BlockTreeFieldService.hpp:30-274): subtree
//                           lookup by (level, index), n2i cache, field range,
//                           interpolateWith (:118-145);
//  - DenseFieldRendererHip: DenseFieldRenderer_cu (DenseFieldRenderer_cu.hpp:32-54);
//  - VsWorkerHip          : VsWorkerInterface (VsWorkerInterface.hpp:38-111) as
//                           implemented by VsWorker_cu (VsWorker_cu.cpp:96-313);
//  - composeFrame         : AssembleFunc / FinalComposeFunc of VsRenderer
//                           (VsRenderer.cpp:403-519) over the worker boxes of
//                           one level (IndexCubeSplitStorage::getBoxesSorted),
//                           blended on the device.
// The reference's own types (BlockTreeNodes, VsRenderSharedState, silver_bullets
// cancel checkers, Qt images) are absent here; the structs below carry the
// same fields under the same names.
#pragma once

#include "../revise_hip.h"

#include <array>
#include <atomic>
#include <functional>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace revise {

// C ABI status -> exception, the reference's error convention (cudaCheck.hpp:41).
struct HipError : std::runtime_error {
    int code;
    HipError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
void throwIfFailed(int status);

// SubtreeSetDescription (SubtreeSetDescription.hpp:27-58): a level and an index
// box (max exclusive).
struct SubtreeSetDescription {
    unsigned level = 0;
    std::array<uint32_t, 3> boxMin{0, 0, 0}, boxMax{0, 0, 0};
    bool empty() const
    {
        return boxMin[0] >= boxMax[0] || boxMin[1] >= boxMax[1] || boxMin[2] >= boxMax[2];
    }
};

// RgbaImagePart (types.hpp:47-57): origin, size and ARGB32 bytes (BB GG RR AA).
struct RgbaImagePart {
    std::array<int, 2> origin{0, 0};
    std::array<int, 2> size{0, 0};
    std::vector<unsigned char> bits;
};

// The field source: subtrees by (level, index) with their n2i and sparse
// values (BlockTreeFieldProvider layout, BlockTreeFieldProvider.hpp:156-168).
class FieldServiceHip {
public:
    FieldServiceHip();
    ~FieldServiceHip();
    FieldServiceHip(const FieldServiceHip&) = delete;
    FieldServiceHip& operator=(const FieldServiceHip&) = delete;

    // Registers a subtree root; values are copied to the device and the
    // subtree's gather plan is encoded from its n2i -- once per distinct
    // (depth, n2i): subtrees of the same shape share one plan.
    void addSubtree(unsigned level, const std::array<uint32_t, 3>& index, unsigned depth,
                    const uint16_t* n2i, uint64_t n, const float* sparse);
    void setFieldRange(float fmin, float fmax);
    std::array<float, 2> fieldRange() const { return {m_fmin, m_fmax}; }
    // The secondary field of the ValueOnIsosurface(s) modes
    // (RenderSharedState::secondaryFieldIndex, DenseFieldRenderer_cu.cu:288-307):
    // values of a registered subtree, in its n2i order (copied), and the
    // field's range.
    void addSecondary(unsigned level, const std::array<uint32_t, 3>& index, const float* sparse);
    void setSecondaryRange(float fmin, float fmax);
    std::array<float, 2> secondaryRange() const { return {m_smin, m_smax}; }
    bool hasSubtree(unsigned level, const std::array<uint32_t, 3>& index) const;
    unsigned subtreeDepth(unsigned level, const std::array<uint32_t, 3>& index) const;
    // interpolateWith<DenseFieldInterpolatorHip>: interpolate the subtree of
    // (level, index) into d_dense (rv_brick_nodes(depth) floats) on `stream`;
    // through the subtree's cached gather plan.  Throws
    // std::range_error for an absent block (CompressedBlockTree.hpp:77-78).
    void interpolate(unsigned level, const std::array<uint32_t, 3>& index, float* d_dense,
                     rv_stream_t stream);
    // The interpolation job of a subtree into d_dense, for batched
    // interpolation of many blocks in one launch.
    rv_brick_job job(unsigned level, const std::array<uint32_t, 3>& index, float* d_dense,
                     rv_stream_t stream, bool secondary = false);
    // The same for the secondary field (throws if the subtree has none).
    void interpolateSecondary(unsigned level, const std::array<uint32_t, 3>& index, float* d_dense,
                              rv_stream_t stream);
    // Bumped by addSubtree: cached batches of a previous field are stale.
    uint64_t version() const { return m_version; }

private:
    struct Subtree;
    struct Plan; // a gather plan shared by the subtrees of one shape
    std::shared_ptr<Plan> planFor(unsigned depth, const uint16_t* n2i, uint64_t n);
    std::map<std::array<uint32_t, 4>, std::unique_ptr<Subtree>> m_subtrees;
    std::multimap<uint64_t, std::weak_ptr<Plan>> m_plans; // by hash of (depth, n2i)
    float m_fmin = 0.f, m_fmax = 1.f;
    float m_smin = 0.f, m_smax = 1.f;
    uint64_t m_version = 0;
};

// The consumed fields of VsRenderSharedState / VsControllerInput
// (VsRenderSharedState.hpp:29-46, VsControllerInput.hpp:208-251), Argb mode.
struct RenderSharedState {
    FieldServiceHip* fieldSvc = nullptr;
    std::array<float, 16> cameraTransform{}; // row-major world->eye
    std::array<uint32_t, 2> viewportSize{800, 600};
    float fovY = 30.f;
    float renderQuality = 1.f;
    float threshold = 0.5f;
    float alphaCutoff = 0.996f;
    std::array<float, 3> backgroundColor{0.f, 0.f, 0.f};
    std::vector<float> colorTable; // N x 4 RGBA (makeColorTransferVector output)
    // VsControllerInput::fieldMode / fieldAllParam().isosurfaceLevel() /
    // isosurfaceOpacity() (VsControllerInput.hpp:208-251): a RV_FIELD_* mode;
    // the isosurface level is relative to the field range like threshold.
    uint32_t fieldMode = RV_FIELD_ARGB;
    float isosurfaceLevel = 0.5f;
    float isosurfaceOpacity = 1.f;
    std::vector<float> isosurfaceLevels; // fieldAllParam().isosurfaceLevels() (Isosurfaces)
    // The Argb sample sequence: false (default) the fast one, true the exact
    // one (RV_RENDER_EXACT, bit-identical to the CPU restatement).
    bool exact = false;
};

class DenseFieldRendererHip {
public:
    DenseFieldRendererHip();
    ~DenseFieldRendererHip();
    void setRenderSharedState(const RenderSharedState* s);
    const RenderSharedState* renderSharedState() const { return m_state; }
    // clearViewport (DenseFieldRenderer_cu.cu:268-277); d_viewport: W*H.
    void clearViewport(uint32_t* d_viewport) const;
    // renderDenseField (DenseFieldRenderer_cu.cu:279-486): interpolate the
    // block's brick and ray-cast it into d_viewport.
    void renderDenseField(uint32_t* d_viewport, unsigned level, const std::array<uint32_t, 3>& index,
                          const std::atomic<bool>& isCancelled);
    void renderDenseField(uint32_t* d_viewport, unsigned level, const std::array<uint32_t, 3>& index,
                          const std::function<bool()>& isCancelled);
    void markColorTableDirty() { m_tfDirty = true; }
    void ensureColorTable();
    const float* colorTable() const { return m_tf; }
    uint32_t colorTableSize() const { return m_tfN; }
    rv_stream_t stream() const { return m_stream; }

private:
    const RenderSharedState* m_state = nullptr;
    rv_stream_t m_stream = nullptr;
    float* m_tf = nullptr;
    uint32_t m_tfN = 0;
    bool m_tfDirty = true;
    float* m_brick = nullptr;
    uint64_t m_brickNodes = 0;
    std::array<uint32_t, 4> m_lastBlock{~0u, 0, 0, 0};
    uint64_t m_lastVersion = ~0ull; // field-service version the brick was interpolated at
    float* m_brick2 = nullptr; // secondary field brick (ValueOnIsosurface(s))
    std::array<uint32_t, 4> m_lastBlock2{~0u, 0, 0, 0};
    uint64_t m_lastVersion2 = ~0ull;
};

class VsWorkerHip {
public:
    VsWorkerHip();
    ~VsWorkerHip();
    // initialize (VsWorker_cu.cpp:96-104): selects GPU threadId % device count.
    void initialize(const std::string& shaderPath, bool logInfo, const std::string& logFileName,
                    size_t threadId);
    void setRenderSharedState(const RenderSharedState* s);
    const RenderSharedState* renderSharedState() const { return m_state; }
    void updateState(unsigned flags);
    // renderScenePart (VsWorker_cu.cpp:225-313).  The cropped part is also
    // kept on the device (devicePart) for on-device compositing.
    RgbaImagePart renderScenePart(const SubtreeSetDescription& subtrees,
                                  const std::atomic<bool>& isCancelled);
    // hostDst != nullptr: the part's pixels go straight to hostDst (hostCap
    // pixels available; pinned memory copies fastest) and bits stays empty.
    RgbaImagePart renderScenePart(const SubtreeSetDescription& subtrees,
                                  const std::function<bool()>& isCancelled,
                                  uint32_t* hostDst = nullptr, uint64_t hostCap = 0);
    // The last part on the device: pixels (w*h ARGB) and {x, y, w, h}.
    const uint32_t* devicePart(std::array<int, 4>& rect) const;
    // The fused path: the box's bricks interpolated in one batched launch
    // (interpolateFused), then one share launch over them in back-to-front
    // order (renderFused).
    void interpolateFused(const SubtreeSetDescription& sub, const std::function<bool()>& isCancelled);
    void renderFused(const SubtreeSetDescription& sub, const std::vector<uint32_t>& order,
                     uint64_t count, const std::function<bool()>& isCancelled);
    void markStagingCopied();
    bool lastCancelled() const { return m_cancelled; }
    rv_stream_t stream() const { return m_renderer.stream(); }
    // The device colour table of the current state (uploaded when it changed).
    const float* colorTable()
    {
        m_renderer.ensureColorTable();
        return m_renderer.colorTable();
    }
    uint32_t colorTableSize() const { return m_renderer.colorTableSize(); }
    // fused (default): the share's bricks are interpolated in one batched
    // launch and ray-cast in one launch (rv_render_share_argb), bit-identical
    // to the per-block loop; per-block: one renderDenseField call per block,
    // the reference's structure.  Cancellation is polled between blocks
    // (per-block) or before each launch (fused).
    void setFused(bool fused) { m_fused = fused; }

private:
    struct ShareCache;
    std::unique_ptr<ShareCache> m_share;
    bool m_fused = true;
    const RenderSharedState* m_state = nullptr;
    DenseFieldRendererHip m_renderer;
    uint32_t* m_viewport = nullptr;
    uint64_t m_viewportPixels = 0;
    uint32_t* m_part = nullptr;
    uint64_t m_partCap = 0;
    int* m_rectBuf = nullptr;     // 16 bytes: the crop rectangle's reduction
    bool m_rectDone = false;      // the share launch left the rectangle in m_rectBuf
    int* m_rectHost = nullptr;    // 16 pinned bytes: the rectangle for the host
    int* m_rectHostDev = nullptr; // the device's address of m_rectHost
    std::array<int, 4> m_partRect{0, 0, 0, 0};
    // The device address of a pinned host buffer, or null for pageable memory.
    void* mappedHost(void* p);
    bool m_cancelled = false;
};

// One frame of a level on one GPU: split the level's index cube into
// splitCount worker boxes sorted back to front, render each box's part,
// assemble the parts pairwise in that order and compose the result over the
// background (VsRenderer.cpp:1086-1156 with RenderGraphGenerator.cpp:37-129).
// Returns the W*H ARGB frame (rows as in the reference frame buffer).
std::vector<uint32_t> composeFrame(VsWorkerHip& worker, unsigned level, unsigned splitCount,
                                   const std::function<bool()>& isCancelled);

// VsControllerFrameOutputHeader (VsControllerFrameOutputHeader.hpp:30-50),
// with the same layout: 24 bytes, the level padded to 8.
struct FrameOutputHeader {
    uint64_t frameNumber = ~0ull; // ~0: no frame
    uint64_t durationMs = 0;
    uint32_t level = 0;
    uint32_t pad = 0;
};

// VsRenderer::continueRendering (VsRenderer.cpp:1061-1166): the progressive
// level-of-detail loop over levels levelMin..levelMax while not cancelled.
// Each level's composed frame goes to writeFrame(header, pixels) with the
// next frame number (then incremented) and, when measureTime, the level's
// wall time in ms (VsRenderer.cpp:1138-1152).  A level cancelled midway
// writes nothing.  Returns the number of frames written.
unsigned renderProgressive(VsWorkerHip& worker, unsigned levelMin, unsigned levelMax,
                           unsigned splitCount, const std::function<bool()>& isCancelled,
                           uint64_t& frameNumber, bool measureTime,
                           const std::function<void(const FrameOutputHeader&, const uint32_t*)>& writeFrame);

} // namespace revise
