// Reader of the reference's on-disk dataset: the block-tree metadata file
// (<base>.s3dmm-meta), the per-field sparse value files
// (<base>.s3dmm-field#<name>) and the field info text file
// (<base>.s3dmm-fields).  Host code only (no device calls), so that it runs
// wherever the library loads.
//
// Binary rules (binary_io.hpp:109-250): arithmetic values are written raw
// (little-endian on the hosts the reference targets), containers as a
// size_t count followed by the items, unordered_map items as (key, value)
// pairs, MultiIndex<3, T> as its 3 elements (MultiIndex_binary_io.hpp:27-45).
#include "../../include/revise_hip.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace rv {
void set_error(const std::string& msg);
void clear_error();
} // namespace rv

namespace {

struct FormatError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

int fail(int code, const std::string& msg)
{
    rv::set_error(msg);
    return code;
}

#define IO_REQUIRE(cond, msg)                                                          \
    do {                                                                               \
        if (!(cond))                                                                   \
            return fail(RV_ERR_INVALID, msg);                                          \
    } while (0)

#define IO_TRY try {
#define IO_CATCH                                                                       \
    }                                                                                  \
    catch (const FormatError& ex)                                                      \
    {                                                                                  \
        return fail(RV_ERR_INVALID, ex.what());                                        \
    }                                                                                  \
    catch (const std::bad_alloc&)                                                      \
    {                                                                                  \
        return fail(RV_ERR_OOM, "host allocation failed");                             \
    }                                                                                  \
    catch (const std::exception& ex)                                                   \
    {                                                                                  \
        return fail(RV_ERR_INTERNAL, ex.what());                                       \
    }                                                                                  \
    catch (...)                                                                        \
    {                                                                                  \
        return fail(RV_ERR_INTERNAL, "unknown exception");                             \
    }

// Sequential reader over a whole file held in memory; every read is bounds
// checked (a truncated file is an error, as the reference's failbit
// exceptions make it, binary_io.hpp:46-66).
struct Bytes {
    std::vector<char> buf;
    size_t pos = 0;
    std::string name;

    void need(size_t n) const
    {
        if (n > buf.size() || pos > buf.size() - n)
            throw FormatError(name + ": unexpected end of file at byte " + std::to_string(pos));
    }
    template <class T> T get()
    {
        need(sizeof(T));
        T v;
        std::memcpy(&v, buf.data() + pos, sizeof(T));
        pos += sizeof(T);
        return v;
    }
    void seek(uint64_t p)
    {
        if (p > buf.size())
            throw FormatError(name + ": position " + std::to_string(p) + " beyond the end of the file");
        pos = (size_t)p;
    }
};

bool read_file(const char* path, std::vector<char>& out)
{
    std::ifstream is(path, std::ios::binary);
    if (!is)
        return false;
    is.seekg(0, std::ios::end);
    const std::streamoff n = is.tellg();
    if (n < 0)
        return false;
    is.seekg(0);
    out.resize((size_t)n);
    if (n > 0 && !is.read(out.data(), n))
        return false;
    return true;
}

inline uint32_t popcount8(uint32_t v) { return (uint32_t)__builtin_popcount(v & 0xffu); }

inline uint64_t loc_key(uint32_t x, uint32_t y, uint32_t z)
{
    return ((uint64_t)z << 42) | ((uint64_t)y << 21) | (uint64_t)x;
}

} // namespace

struct rv_dataset {
    rv_dataset_info info{};
    // The block tree (CompressedBlockTree<3>::Data, CompressedBlockTree.hpp:306-311):
    // one byte per node, nodes in breadth-first order (node 0 = the root), bit
    // (7 - c) set when local child c (x = bit 0, y = bit 1, z = bit 2) exists
    // (BlockTreeCompressor.hpp:34-50).  first[i]: the index of node i's first
    // child, 1 + the children of all nodes before it (CompressedBlockTree.hpp:
    // 395-420).
    std::vector<uint8_t> mask;
    std::vector<uint64_t> first;
    std::vector<rv_subtree_info> subtrees;                       // ascending values_pos
    std::vector<std::unordered_map<uint64_t, uint64_t>> by_level; // location -> subtree

    // Tree node of the block (level, loc), descending from the root
    // (childBlockLocation: child location = 2 * parent + bits of c); -1 if absent.
    int64_t find_node(uint32_t level, const uint32_t loc[3]) const
    {
        uint64_t node = 0;
        for (uint32_t l = level; l-- > 0;) {
            const uint32_t c = ((loc[0] >> l) & 1u) | (((loc[1] >> l) & 1u) << 1) |
                               (((loc[2] >> l) & 1u) << 2);
            const uint32_t m = mask[node];
            if (!(m & (0x80u >> c)))
                return -1;
            node = first[node] + popcount8(m >> (8 - c)); // children before c: the top c bits
        }
        return (int64_t)node;
    }
};

namespace {

constexpr char kMetaId[] = "s3dmm-metadata"; // Metadata.hpp:495-497
constexpr uint32_t kMetaVersion = 1;         // Metadata.hpp:499

// Metadata<3>(std::istream&) (Metadata.hpp:182-214, 464-489).
void parse_meta(Bytes& b, rv_dataset& ds)
{
    rv_dataset_info& in = ds.info;
    const size_t idn = sizeof(kMetaId) - 1;
    if (b.buf.size() >= idn && std::memcmp(b.buf.data(), kMetaId, idn) == 0) {
        b.pos = idn;
        in.format_version = b.get<uint32_t>();
        if (in.format_version > kMetaVersion)
            throw FormatError(b.name + ": unsupported format version " + std::to_string(in.format_version));
        const uint64_t header_size = b.get<uint64_t>();
        const size_t hpos = b.pos;
        const uint32_t dim = b.get<uint32_t>();
        if (dim != 3)
            throw FormatError(b.name + ": wrong dimension " + std::to_string(dim));
        in.max_subtree_depth = b.get<uint32_t>();
        in.max_full_level = b.get<uint32_t>();
        if (b.pos - hpos != header_size)
            throw FormatError(b.name + ": invalid header");
    } else {
        // files without the identification string are version 0 and start
        // with the position prefix (Metadata.hpp:470-474)
        b.pos = 0;
        in.format_version = 0;
        in.max_full_level = 0;
    }
    b.seek(b.get<uint64_t>()); // end of the (absent) node lists: the tree follows

    // CompressedBlockTree: bounding cube (min, size), dataBitCount, byte vector
    // (BlockTree_io.hpp:64-80, BoundingCube_io.hpp:27-44)
    for (int a = 0; a < 3; ++a)
        in.bc_min[a] = b.get<float>();
    in.bc_size = b.get<float>();
    const uint64_t bits = b.get<uint64_t>();
    const uint64_t nbytes = b.get<uint64_t>();
    if (bits % 8 != 0 || bits / 8 != nbytes || nbytes == 0)
        throw FormatError(b.name + ": block tree of " + std::to_string(bits) + " bits in " +
                          std::to_string(nbytes) + " bytes");
    b.need(nbytes);
    ds.mask.assign(b.buf.begin() + (ptrdiff_t)b.pos, b.buf.begin() + (ptrdiff_t)(b.pos + nbytes));
    b.pos += nbytes;
    in.node_count = nbytes;
    ds.first.resize(nbytes);
    uint64_t next = 1;
    for (uint64_t i = 0; i < nbytes; ++i) {
        ds.first[i] = next;
        next += popcount8(ds.mask[i]);
    }
    if (next != nbytes) // every node but the root is some node's child (CompressedBlockTree.hpp:408)
        throw FormatError(b.name + ": block tree child counts do not add up to its node count");

    // Subtree maps: per level, (location, SubtreeData); version 0 entries
    // lack the depth, which is then maxSubtreeDepth (Metadata.hpp:195-211, 58-115)
    if (in.format_version == 0)
        in.max_subtree_depth = b.get<uint32_t>();
    const uint64_t nlevels = b.get<uint64_t>();
    if (nlevels > 32)
        throw FormatError(b.name + ": " + std::to_string(nlevels) + " subtree levels");
    in.level_count = (uint32_t)nlevels;
    ds.by_level.resize(nlevels);
    std::vector<std::unordered_map<uint64_t, rv_subtree_info>> maps(nlevels);
    for (uint32_t level = 0; level < nlevels; ++level) {
        const uint64_t count = b.get<uint64_t>();
        if (count > in.node_count)
            throw FormatError(b.name + ": more subtrees than blocks at level " + std::to_string(level));
        for (uint64_t k = 0; k < count; ++k) {
            rv_subtree_info s{};
            s.level = level;
            for (int a = 0; a < 3; ++a)
                s.location[a] = b.get<uint32_t>();
            s.node = b.get<uint32_t>();
            s.values_pos = b.get<uint64_t>();
            s.depth = in.format_version == 0 ? in.max_subtree_depth : (uint32_t)b.get<uint64_t>();
            // a repeated key keeps the last entry (binary_io.hpp:228-236)
            maps[level][loc_key(s.location[0], s.location[1], s.location[2])] = s;
        }
    }
    in.field_file_size = b.get<uint64_t>();

    for (auto& m : maps)
        for (auto& kv : m)
            ds.subtrees.push_back(kv.second);
    // field files hold the subtrees' values in walkSubtrees order, which is
    // the order of their positions (Metadata.hpp:408-433)
    std::sort(ds.subtrees.begin(), ds.subtrees.end(), [](const rv_subtree_info& a, const rv_subtree_info& c) {
        if (a.values_pos != c.values_pos)
            return a.values_pos < c.values_pos;
        if (a.level != c.level)
            return a.level < c.level;
        return loc_key(a.location[0], a.location[1], a.location[2]) <
               loc_key(c.location[0], c.location[1], c.location[2]);
    });
    in.subtree_count = ds.subtrees.size();
    for (uint64_t i = 0; i < ds.subtrees.size(); ++i) {
        const rv_subtree_info& s = ds.subtrees[i];
        const uint32_t lim = s.level >= 32 ? 0u : (1u << s.level);
        if (s.location[0] >= lim || s.location[1] >= lim || s.location[2] >= lim)
            throw FormatError(b.name + ": subtree location outside its level");
        if (s.depth > 15)
            throw FormatError(b.name + ": subtree depth " + std::to_string(s.depth) + " above 15");
        const int64_t node = ds.find_node(s.level, s.location);
        if (node < 0 || (uint64_t)node != s.node)
            throw FormatError(b.name + ": subtree root (level " + std::to_string(s.level) +
                              ") does not match the block tree");
        ds.by_level[s.level][loc_key(s.location[0], s.location[1], s.location[2])] = i;
    }
}

} // namespace

extern "C" int rv_dataset_open(const char* meta_path, rv_dataset** out)
{
    IO_TRY
    IO_REQUIRE(meta_path && out, "rv_dataset_open: null pointer");
    *out = nullptr;
    Bytes b;
    b.name = meta_path;
    IO_REQUIRE(read_file(meta_path, b.buf), std::string("rv_dataset_open: cannot read '") + meta_path + "'");
    auto ds = new rv_dataset();
    try {
        parse_meta(b, *ds);
    } catch (...) {
        delete ds;
        throw;
    }
    *out = ds;
    return RV_OK;
    IO_CATCH
}

extern "C" void rv_dataset_close(rv_dataset* ds) { delete ds; }

extern "C" int rv_dataset_get_info(const rv_dataset* ds, rv_dataset_info* info)
{
    IO_REQUIRE(ds && info, "rv_dataset_get_info: null pointer");
    *info = ds->info;
    return RV_OK;
}

extern "C" int rv_dataset_subtree(const rv_dataset* ds, uint64_t i, rv_subtree_info* out)
{
    IO_REQUIRE(ds && out, "rv_dataset_subtree: null pointer");
    IO_REQUIRE(i < ds->subtrees.size(), "rv_dataset_subtree: index out of range");
    *out = ds->subtrees[i];
    return RV_OK;
}

extern "C" int rv_dataset_find_subtree(const rv_dataset* ds, uint32_t level, const uint32_t loc[3],
                                       uint64_t* i)
{
    IO_REQUIRE(ds && loc && i, "rv_dataset_find_subtree: null pointer");
    if (level < ds->by_level.size()) {
        auto it = ds->by_level[level].find(loc_key(loc[0], loc[1], loc[2]));
        if (it != ds->by_level[level].end()) {
            *i = it->second;
            return RV_OK;
        }
    }
    return fail(RV_ERR_INVALID, "rv_dataset_find_subtree: no subtree root at this level and location");
}

extern "C" int rv_dataset_closest_subtree(const rv_dataset* ds, uint32_t level, const uint32_t block[3],
                                          uint64_t* i)
{
    IO_REQUIRE(ds && block && i, "rv_dataset_closest_subtree: null pointer");
    IO_REQUIRE(!ds->by_level.empty(), "rv_dataset_closest_subtree: the dataset has no subtrees");
    uint32_t b[3] = {block[0], block[1], block[2]};
    const uint32_t max_level = (uint32_t)ds->by_level.size() - 1;
    if (level > max_level) {
        for (int a = 0; a < 3; ++a)
            b[a] >>= level - max_level;
        level = max_level;
    }
    for (;; --level) {
        auto it = ds->by_level[level].find(loc_key(b[0], b[1], b[2]));
        if (it != ds->by_level[level].end()) {
            *i = it->second;
            return RV_OK;
        }
        if (level == 0)
            break;
        for (int a = 0; a < 3; ++a)
            b[a] >>= 1;
    }
    return fail(RV_ERR_INVALID, "rv_dataset_closest_subtree: no subtree root above this block");
}

namespace {

// Children table of subtree i (nnodes x 8, node 0 = the subtree root, 0 =
// no child), breadth first from the root down to the subtree depth; rows
// of the deepest level keep no children.
std::vector<uint32_t> subtree_rows(const rv_dataset& ds, uint64_t i)
{
    const rv_subtree_info& s = ds.subtrees[i];
    std::vector<uint64_t> tree_node{s.node};
    std::vector<uint32_t> rel_level{0};
    std::vector<uint32_t> rows(8, 0);
    for (uint64_t r = 0; r < tree_node.size(); ++r) {
        if (rel_level[r] >= s.depth)
            continue;
        const uint32_t m = ds.mask[tree_node[r]];
        uint64_t child = ds.first[tree_node[r]];
        for (uint32_t c = 0; c < 8; ++c) {
            if (!(m & (0x80u >> c)))
                continue;
            if (tree_node.size() >= 0xffffffffull)
                throw FormatError("subtree too large");
            rows[8 * r + c] = (uint32_t)tree_node.size();
            tree_node.push_back(child++);
            rel_level.push_back(rel_level[r] + 1);
            rows.resize(rows.size() + 8, 0);
        }
    }
    return rows;
}

} // namespace

extern "C" int rv_dataset_subtree_children(const rv_dataset* ds, uint64_t i, uint32_t* children,
                                           uint64_t cap_nodes, uint64_t* nnodes)
{
    IO_TRY
    IO_REQUIRE(ds && nnodes, "rv_dataset_subtree_children: null pointer");
    IO_REQUIRE(i < ds->subtrees.size(), "rv_dataset_subtree_children: index out of range");
    const std::vector<uint32_t> rows = subtree_rows(*ds, i);
    *nnodes = rows.size() / 8;
    IO_REQUIRE(children || cap_nodes == 0, "rv_dataset_subtree_children: null output");
    IO_REQUIRE(cap_nodes >= rows.size() / 8, "rv_dataset_subtree_children: output capacity too small");
    std::memcpy(children, rows.data(), rows.size() * sizeof(uint32_t));
    return RV_OK;
    IO_CATCH
}

extern "C" int rv_dataset_read_field(const rv_dataset* ds, const char* field_path, uint64_t i,
                                     float range[2], float* values, uint64_t n)
{
    IO_TRY
    IO_REQUIRE(ds && field_path && range && (values || n == 0), "rv_dataset_read_field: null pointer");
    IO_REQUIRE(i < ds->subtrees.size(), "rv_dataset_read_field: index out of range");
    const uint64_t pos = ds->subtrees[i].values_pos;
    std::ifstream is(field_path, std::ios::binary);
    IO_REQUIRE(is, std::string("rv_dataset_read_field: cannot open '") + field_path + "'");
    is.seekg(0, std::ios::end);
    const uint64_t size = (uint64_t)is.tellg();
    IO_REQUIRE(n <= (UINT64_MAX / 4) - 2 && pos <= size && (2 + n) * 4 <= size - pos,
               std::string("rv_dataset_read_field: '") + field_path + "' ends inside the subtree's values");
    // min, max, then the values in n2i order (BlockTreeFieldProvider.hpp:124-131, 156-168)
    is.seekg((std::streamoff)pos);
    IO_REQUIRE(is.read(reinterpret_cast<char*>(range), 2 * sizeof(float)), "rv_dataset_read_field: read failed");
    if (n)
        IO_REQUIRE(is.read(reinterpret_cast<char*>(values), (std::streamsize)(n * sizeof(float))),
                   "rv_dataset_read_field: read failed");
    return RV_OK;
    IO_CATCH
}

namespace {

int load_worker(const rv_dataset* ds, const char* field_path, rv_worker* w, uint32_t level_min,
                uint32_t level_max, uint64_t* loaded, bool secondary)
{
    IO_TRY
    IO_REQUIRE(ds && field_path && w, "rv_dataset_load_worker: null pointer");
    IO_REQUIRE(level_min <= level_max, "rv_dataset_load_worker: empty level range");
    std::vector<uint32_t> children;
    std::vector<uint16_t> n2i;
    std::vector<float> values;
    uint64_t count = 0;
    for (uint64_t i = 0; i < ds->subtrees.size(); ++i) {
        const rv_subtree_info& s = ds->subtrees[i];
        if (s.level < level_min || s.level > level_max)
            continue;
        children = subtree_rows(*ds, i);
        const uint64_t nnodes = children.size() / 8;
        // corners of the blocks to depth d: at most (2^d + 1)^3, and no more
        // than 8 per block
        const uint64_t npe = (1ull << s.depth) + 1;
        const uint64_t cap = std::min(npe * npe * npe, 8 * nnodes);
        n2i.resize(3 * cap);
        uint64_t n = 0;
        int rc = rv_subtree_n2i(children.data(), (uint32_t)nnodes, s.depth, n2i.data(), cap, &n);
        if (rc != RV_OK)
            return rc;
        values.resize(n);
        float range[2];
        rc = rv_dataset_read_field(ds, field_path, i, range, values.data(), n);
        if (rc != RV_OK)
            return rc;
        rc = secondary ? rv_worker_add_subtree_secondary(w, s.level, s.location, values.data())
                       : rv_worker_add_subtree(w, s.level, s.location, s.depth, n2i.data(), n,
                                               values.data());
        if (rc != RV_OK)
            return rc;
        ++count;
    }
    if (loaded)
        *loaded = count;
    return RV_OK;
    IO_CATCH
}

} // namespace

extern "C" int rv_dataset_load_worker(const rv_dataset* ds, const char* field_path, rv_worker* w,
                                      uint32_t level_min, uint32_t level_max, uint64_t* loaded)
{
    return load_worker(ds, field_path, w, level_min, level_max, loaded, false);
}

extern "C" int rv_dataset_load_worker_secondary(const rv_dataset* ds, const char* field_path,
                                                rv_worker* w, uint32_t level_min, uint32_t level_max,
                                                uint64_t* loaded)
{
    return load_worker(ds, field_path, w, level_min, level_max, loaded, true);
}

// BlockTreeFieldService::readFieldInfoFile (BlockTreeFieldService.hpp:177-228):
// "time_steps", the frame count, an empty line, the column header, then one
// line per field: name, min, max, animated (whitespace separated).
extern "C" int rv_fields_info_read(const char* path, uint32_t* time_steps, rv_field_info* fields,
                                   uint32_t cap, uint32_t* count)
{
    IO_TRY
    IO_REQUIRE(path && time_steps && count && (fields || cap == 0), "rv_fields_info_read: null pointer");
    std::ifstream is(path);
    IO_REQUIRE(is, std::string("rv_fields_info_read: cannot open '") + path + "'");
    std::string line;
    auto where = [&](unsigned ln) { return "line " + std::to_string(ln) + " of '" + std::string(path) + "'"; };
    auto predefined = [&](unsigned ln, const char* content) {
        if (!std::getline(is, line) && ln != 3)
            throw FormatError("unexpected end of file in " + where(ln));
        if (line != content)
            throw FormatError("unexpected content of " + where(ln) + ": expected '" + content +
                              "', found '" + line + "'");
    };
    predefined(1, "time_steps");
    if (!std::getline(is, line))
        throw FormatError("unexpected end of file in " + where(2));
    {
        std::istringstream ss(line);
        long long v = -1;
        if (!(ss >> v) || v < 0 || v > 0xffffffffll)
            throw FormatError("bad frame count in " + where(2));
        *time_steps = (uint32_t)v;
    }
    predefined(3, "");
    predefined(4, "field\tmin\tmax\tanimated");
    uint32_t n = 0;
    for (unsigned ln = 5; std::getline(is, line); ++ln) {
        std::istringstream ss(line);
        std::string name;
        float mn, mx;
        int animated;
        if (!(ss >> name >> mn >> mx >> animated))
            throw FormatError("bad field line in " + where(ln));
        if (n < cap) {
            rv_field_info& f = fields[n];
            std::memset(&f, 0, sizeof f);
            IO_REQUIRE(name.size() < sizeof f.name, "rv_fields_info_read: field name too long in " + where(ln));
            std::memcpy(f.name, name.data(), name.size());
            f.min = mn;
            f.max = mx;
            f.animated = animated != 0;
        }
        ++n;
    }
    *count = n;
    IO_REQUIRE(n <= cap, "rv_fields_info_read: output capacity too small");
    return RV_OK;
    IO_CATCH
}
