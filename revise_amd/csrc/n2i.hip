// Node list (n2i) of a subtree built on the device: the same list as the
// host rv_subtree_n2i (BlockTreeNodes.hpp:66-80, 166-181 -- the corners of
// every block of a pre-order walk limited to `depth` levels, children in
// local order 0..7 with x = bit0, each corner appended the first time it is
// seen), computed without a walk or a hash set:
//
//   1. level by level, the blocks of the next level are the children of the
//      current one, compacted in (parent, child slot) order with a scan; each
//      block knows its location and the range of its children;
//   2. bottom-up, every block's subtree size (blocks visited below it,
//      itself included);
//   3. top-down, every block's pre-order position: a child's is its parent's
//      plus one plus the sizes of the siblings before it;
//   4. corner c of the block at position q is candidate 8 q + c; a dense
//      table over the (2^d+1)^3 lattice keeps each node's smallest candidate
//      (atomicMin), so a candidate is a node's first sighting iff the table
//      holds it;
//   5. the first sightings, in candidate order, are the list: a scan of
//      their flags gives each one its output slot.

#include "rv_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <vector>

namespace rv {

int fill_u32(uint32_t* d, uint64_t n, uint32_t value, hipStream_t s);

namespace {

struct Lvl {
    uint32_t* node;   // block's node index in the children array
    uint32_t* loc;    // 3 per block: x, y, z at this level
    uint32_t* first;  // per block: its first child's index in the next level (nblocks + 1)
    uint32_t* size;   // per block: subtree size
    uint32_t* pos;    // per block: pre-order position
    uint32_t n = 0;
};

__global__ void count_children_kernel(const uint32_t* __restrict__ children, uint32_t nnodes,
                                      const uint32_t* __restrict__ node, uint32_t n,
                                      uint32_t* __restrict__ cnt, uint32_t* __restrict__ err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t* ch = children + 8ull * node[i];
    uint32_t c = 0;
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = ch[k];
        if (v >= nnodes)
            atomicOr(err, 1u);
        c += (v != 0u && v < nnodes) ? 1u : 0u;
    }
    cnt[i] = c;
}

__global__ void emit_children_kernel(const uint32_t* __restrict__ children, uint32_t nnodes,
                                     const uint32_t* __restrict__ node,
                                     const uint32_t* __restrict__ loc, uint32_t n,
                                     const uint32_t* __restrict__ first,
                                     uint32_t* __restrict__ cnode, uint32_t* __restrict__ cloc)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t* ch = children + 8ull * node[i];
    uint32_t o = first[i];
    const uint32_t x = loc[3 * i], y = loc[3 * i + 1], z = loc[3 * i + 2];
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t v = ch[k];
        if (v == 0u || v >= nnodes)
            continue;
        cnode[o] = v;
        cloc[3 * o] = (x << 1) | (k & 1u);
        cloc[3 * o + 1] = (y << 1) | ((k >> 1) & 1u);
        cloc[3 * o + 2] = (z << 1) | ((k >> 2) & 1u);
        ++o;
    }
}

// size of a block = 1 + sizes of its children (next level, contiguous range)
__global__ void size_kernel(const uint32_t* __restrict__ first, uint32_t n,
                            const uint32_t* __restrict__ csize, uint32_t* __restrict__ size)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t s = 1;
    if (csize)
        for (uint32_t j = first[i]; j < first[i + 1]; ++j)
            s += csize[j];
    size[i] = s;
}

// children's positions: parent + 1 + sizes of the earlier siblings
__global__ void pos_kernel(const uint32_t* __restrict__ first, uint32_t n,
                           const uint32_t* __restrict__ pos, const uint32_t* __restrict__ csize,
                           uint32_t* __restrict__ cpos)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t p = pos[i] + 1u;
    for (uint32_t j = first[i]; j < first[i + 1]; ++j) {
        cpos[j] = p;
        p += csize[j];
    }
}

__device__ __forceinline__ uint32_t corner_key(const uint32_t* loc, uint32_t i, uint32_t k,
                                               uint32_t shift, uint32_t mask, uint32_t npe,
                                               uint32_t& cx, uint32_t& cy, uint32_t& cz)
{
    cx = ((loc[3 * i] << shift) & mask) + ((k & 1u) << shift);
    cy = ((loc[3 * i + 1] << shift) & mask) + (((k >> 1) & 1u) << shift);
    cz = ((loc[3 * i + 2] << shift) & mask) + (((k >> 2) & 1u) << shift);
    return (cz * npe + cy) * npe + cx;
}

__global__ void first_sight_kernel(const uint32_t* __restrict__ loc, const uint32_t* __restrict__ pos,
                                   uint32_t n, uint32_t shift, uint32_t mask, uint32_t npe,
                                   uint32_t* __restrict__ table)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8u * n)
        return;
    const uint32_t i = t >> 3, k = t & 7u;
    uint32_t cx, cy, cz;
    const uint32_t key = corner_key(loc, i, k, shift, mask, npe, cx, cy, cz);
    atomicMin(table + key, 8u * pos[i] + k);
}

__global__ void flag_kernel(const uint32_t* __restrict__ loc, const uint32_t* __restrict__ pos,
                            uint32_t n, uint32_t shift, uint32_t mask, uint32_t npe,
                            const uint32_t* __restrict__ table, uint32_t* __restrict__ flag)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8u * n)
        return;
    const uint32_t i = t >> 3, k = t & 7u;
    uint32_t cx, cy, cz;
    const uint32_t key = corner_key(loc, i, k, shift, mask, npe, cx, cy, cz);
    const uint32_t cand = 8u * pos[i] + k;
    flag[cand] = table[key] == cand ? 1u : 0u;
}

__global__ void write_kernel(const uint32_t* __restrict__ loc, const uint32_t* __restrict__ pos,
                             uint32_t n, uint32_t shift, uint32_t mask, uint32_t npe,
                             const uint32_t* __restrict__ table, const uint32_t* __restrict__ slot,
                             uint16_t* __restrict__ out, uint64_t cap)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 8u * n)
        return;
    const uint32_t i = t >> 3, k = t & 7u;
    uint32_t cx, cy, cz;
    const uint32_t key = corner_key(loc, i, k, shift, mask, npe, cx, cy, cz);
    const uint32_t cand = 8u * pos[i] + k;
    if (table[key] != cand)
        return;
    const uint32_t o = slot[cand];
    if (o < cap) {
        out[3ull * o] = (uint16_t)cx;
        out[3ull * o + 1] = (uint16_t)cy;
        out[3ull * o + 2] = (uint16_t)cz;
    }
}

__global__ void init_root_kernel(uint32_t* node, uint32_t* loc, uint32_t* pos)
{
    if (threadIdx.x == 0) {
        node[0] = 0;
        loc[0] = loc[1] = loc[2] = 0;
        pos[0] = 0;
    }
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

// Device buffers freed on every exit path.
struct Pool {
    std::vector<void*> all;
    ~Pool()
    {
        for (void* p : all)
            (void)hipFree(p);
    }
    template <class T>
    int get(T** p, uint64_t n)
    {
        void* v = nullptr;
        RV_HIP(hipMalloc(&v, sizeof(T) * (n ? n : 1)));
        all.push_back(v);
        *p = static_cast<T*>(v);
        return RV_OK;
    }
};

// Exclusive sum of in[0, n) into out[0, n] (out[n] = total).
int exclusive_sum(Pool& pool, const uint32_t* in, uint32_t* out, uint32_t n, hipStream_t s)
{
    size_t bytes = 0;
    RV_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n + 1, s));
    void* tmp = nullptr;
    if (int rc = pool.get(reinterpret_cast<uint8_t**>(&tmp), bytes))
        return rc;
    RV_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n + 1, s));
    return RV_OK;
}

} // namespace
} // namespace rv

using namespace rv;

extern "C" int rv_subtree_n2i_device(const uint32_t* d_children, uint32_t nnodes, uint32_t depth,
                                     uint16_t* d_n2i, uint64_t cap, uint64_t* count,
                                     rv_stream_t stream)
{
    RV_TRY
    RV_REQUIRE(d_children && nnodes && count, "rv_subtree_n2i_device: bad arguments");
    RV_REQUIRE(depth <= 8, "rv_subtree_n2i_device: depth must be <= 8");
    RV_REQUIRE(d_n2i || cap == 0, "rv_subtree_n2i_device: null output");
    hipStream_t s = S(stream);
    Pool pool;
    uint32_t* d_err = nullptr;
    int rc;
    if ((rc = pool.get(&d_err, 1)))
        return rc;
    RV_HIP(hipMemsetAsync(d_err, 0, 4, s));
    // 1. levels (the walk stops at `depth` levels below the root)
    std::vector<Lvl> L(depth + 1);
    L[0].n = 1;
    if ((rc = pool.get(&L[0].node, 1)) || (rc = pool.get(&L[0].loc, 3)) || (rc = pool.get(&L[0].pos, 1)))
        return rc;
    hipLaunchKernelGGL(init_root_kernel, dim3(1), dim3(64), 0, s, L[0].node, L[0].loc, L[0].pos);
    RV_LAUNCH_CHECK();
    uint64_t total = 1;
    for (uint32_t l = 0; l <= depth; ++l) {
        Lvl& c = L[l];
        if ((rc = pool.get(&c.first, (uint64_t)c.n + 1)) || (rc = pool.get(&c.size, c.n)))
            return rc;
        if (l == depth || c.n == 0) {
            RV_HIP(hipMemsetAsync(c.first, 0, 4ull * (c.n + 1), s));
            continue;
        }
        uint32_t* cnt = nullptr;
        if ((rc = pool.get(&cnt, (uint64_t)c.n + 1)))
            return rc;
        RV_HIP(hipMemsetAsync(cnt + c.n, 0, 4, s));
        hipLaunchKernelGGL(count_children_kernel, dim3(blocks_for(c.n)), dim3(256), 0, s, d_children,
                           nnodes, c.node, c.n, cnt, d_err);
        RV_LAUNCH_CHECK();
        if ((rc = exclusive_sum(pool, cnt, c.first, c.n, s)))
            return rc;
        uint32_t next = 0, err = 0;
        RV_HIP(hipMemcpyAsync(&next, c.first + c.n, 4, hipMemcpyDeviceToHost, s));
        RV_HIP(hipMemcpyAsync(&err, d_err, 4, hipMemcpyDeviceToHost, s));
        RV_HIP(hipStreamSynchronize(s));
        RV_REQUIRE(!err, "rv_subtree_n2i_device: child index out of range");
        Lvl& d = L[l + 1];
        d.n = next;
        total += next;
        RV_REQUIRE(total < (1ull << 29), "rv_subtree_n2i_device: too many blocks");
        if ((rc = pool.get(&d.node, next)) || (rc = pool.get(&d.loc, 3ull * next)) ||
            (rc = pool.get(&d.pos, next)))
            return rc;
        if (next)
            hipLaunchKernelGGL(emit_children_kernel, dim3(blocks_for(c.n)), dim3(256), 0, s, d_children,
                               nnodes, c.node, c.loc, c.n, c.first, d.node, d.loc);
        RV_LAUNCH_CHECK();
    }
    // 2. subtree sizes, bottom-up
    for (int l = (int)depth; l >= 0; --l) {
        Lvl& c = L[l];
        if (!c.n)
            continue;
        const uint32_t* csize = (l < (int)depth && L[l + 1].n) ? L[l + 1].size : nullptr;
        hipLaunchKernelGGL(size_kernel, dim3(blocks_for(c.n)), dim3(256), 0, s, c.first, c.n, csize,
                           c.size);
        RV_LAUNCH_CHECK();
    }
    // 3. pre-order positions, top-down
    for (uint32_t l = 0; l < depth; ++l) {
        if (!L[l].n || !L[l + 1].n)
            continue;
        hipLaunchKernelGGL(pos_kernel, dim3(blocks_for(L[l].n)), dim3(256), 0, s, L[l].first, L[l].n,
                           L[l].pos, L[l + 1].size, L[l + 1].pos);
        RV_LAUNCH_CHECK();
    }
    // 4. first sightings on the dense lattice
    const uint32_t npe = (1u << depth) + 1u, mask = (1u << depth) - 1u;
    const uint64_t V = (uint64_t)npe * npe * npe;
    uint32_t *table = nullptr, *flag = nullptr, *slot = nullptr;
    if ((rc = pool.get(&table, V)) || (rc = pool.get(&flag, 8 * total + 1)) ||
        (rc = pool.get(&slot, 8 * total + 1)))
        return rc;
    if ((rc = fill_u32(table, V, 0xFFFFFFFFu, s)))
        return rc;
    for (uint32_t l = 0; l <= depth; ++l)
        if (L[l].n)
            hipLaunchKernelGGL(first_sight_kernel, dim3(blocks_for(8ull * L[l].n)), dim3(256), 0, s,
                               L[l].loc, L[l].pos, L[l].n, depth - l, mask, npe, table);
    RV_LAUNCH_CHECK();
    // 5. flags in candidate order, their scan, and the list
    RV_HIP(hipMemsetAsync(flag + 8 * total, 0, 4, s));
    for (uint32_t l = 0; l <= depth; ++l)
        if (L[l].n)
            hipLaunchKernelGGL(flag_kernel, dim3(blocks_for(8ull * L[l].n)), dim3(256), 0, s, L[l].loc,
                               L[l].pos, L[l].n, depth - l, mask, npe, table, flag);
    RV_LAUNCH_CHECK();
    if ((rc = exclusive_sum(pool, flag, slot, (uint32_t)(8 * total), s)))
        return rc;
    uint32_t n = 0;
    RV_HIP(hipMemcpyAsync(&n, slot + 8 * total, 4, hipMemcpyDeviceToHost, s));
    RV_HIP(hipStreamSynchronize(s));
    *count = n;
    RV_REQUIRE(n <= cap, "rv_subtree_n2i_device: output capacity too small");
    for (uint32_t l = 0; l <= depth; ++l)
        if (L[l].n)
            hipLaunchKernelGGL(write_kernel, dim3(blocks_for(8ull * L[l].n)), dim3(256), 0, s, L[l].loc,
                               L[l].pos, L[l].n, depth - l, mask, npe, table, slot, d_n2i, cap);
    RV_LAUNCH_CHECK();
    RV_HIP(hipStreamSynchronize(s)); // the pool's buffers are freed on return
    return RV_OK;
    RV_CATCH
}
