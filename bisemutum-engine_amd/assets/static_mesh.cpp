// static_mesh.cpp -- the engine's StaticMesh asset format read on the host (bi_assets.h).
// Layout per src/scene_basic/static_mesh.cpp:10-43, src/prelude/byte_stream.cpp:15-41
// and src/graphics/mesh.cpp:137-146; the zlib part is what miniz's mz_uncompress reads.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include <zlib.h>

#include "../../include/bi_assets.h"

namespace {

constexpr uint32_t kAssetMagic = 0x0B1A55E7u;
constexpr uint32_t kTriangleList = 3u;   // rhi::PrimitiveTopology::triangle_list

struct SubmeshDesc {
    uint32_t base_vertex, index_offset, num_indices;
    uint8_t topology, pad[3];
};
static_assert(sizeof(SubmeshDesc) == 16, "SubmeshDesc is 16 bytes");

struct AssetError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class Reader {
  public:
    Reader(const uint8_t* p, size_t n) : p_(p), n_(n) {}
    const uint8_t* take(uint64_t k) {
        if (k > n_ - off_) throw AssetError("truncated at byte " + std::to_string(off_));
        const uint8_t* q = p_ + off_;
        off_ += k;
        return q;
    }
    template <class T> T pod() {
        T v;
        std::memcpy(&v, take(sizeof(T)), sizeof(T));
        return v;
    }
    template <class T> std::vector<T> vec(size_t width = 1) {
        const uint64_t count = pod<uint64_t>();
        const uint64_t elem = sizeof(T) * width;
        if (count > (n_ - off_) / elem) throw AssetError("vector of " + std::to_string(count) + " elements exceeds the data");
        std::vector<T> v(count * width);
        if (count) std::memcpy(v.data(), take(count * elem), count * elem);
        return v;
    }
    std::string str() {
        const uint64_t len = pod<uint64_t>();
        const uint8_t* q = take(len);
        return std::string(reinterpret_cast<const char*>(q), len);
    }

  private:
    const uint8_t* p_;
    size_t n_;
    size_t off_ = 0;
};

void put_err(char* err, size_t len, const std::string& msg) {
    if (err && len) std::snprintf(err, len, "%s", msg.c_str());
}

}  // namespace

struct bi_static_mesh {
    std::vector<float> positions, normals, tangents, colors, texcoords, texcoords2;
    std::vector<uint32_t> indices;
    std::vector<SubmeshDesc> submeshes;
    std::vector<std::vector<uint32_t>> tri_indices;   // per submesh, built on request
};

namespace {

void read_mesh_data(Reader& r, bi_static_mesh& m) {
    m.positions = r.vec<float>(3);
    m.normals = r.vec<float>(3);
    m.tangents = r.vec<float>(4);
    m.colors = r.vec<float>(3);
    m.texcoords = r.vec<float>(2);
    m.texcoords2 = r.vec<float>(2);
    m.indices = r.vec<uint32_t>();
    m.submeshes = r.vec<SubmeshDesc>();
}

void parse(const uint8_t* data, size_t size, bi_static_mesh& m) {
    Reader r(data, size);
    const uint32_t magic = r.pod<uint32_t>();
    const std::string type = r.str();
    const uint32_t version = r.pod<uint32_t>();
    if (magic != kAssetMagic || type != "StaticMesh") throw AssetError("not a StaticMesh asset");
    if (version == 1) {
        read_mesh_data(r, m);
    } else if (version == 2) {
        const uint64_t raw_len = r.pod<uint64_t>();
        const uint64_t z_len = r.pod<uint64_t>();
        const uint8_t* z = r.take(z_len);
        // deflate expands at most 1032:1, so a larger stated length is a corrupt header
        // (and must not size an allocation)
        if (raw_len > z_len * 1032u + 1024u) throw AssetError("compressed part states an impossible length");
        std::vector<uint8_t> raw(raw_len);
        uLongf got = static_cast<uLongf>(raw_len);
        if (uncompress(raw.data(), &got, z, static_cast<uLong>(z_len)) != Z_OK || got != raw_len)
            throw AssetError("compressed part does not inflate to its stated length");
        Reader rr(raw.data(), raw.size());
        read_mesh_data(rr, m);
    } else {
        throw AssetError("unknown StaticMesh version " + std::to_string(version));
    }
    m.tri_indices.assign(m.submeshes.size(), {});
}

}  // namespace

extern "C" int bi_static_mesh_load(const void* data, size_t size, bi_static_mesh** out, char* err, size_t err_len) {
    if (!out || (!data && size)) {
        put_err(err, err_len, "null argument");
        return 1;
    }
    *out = nullptr;
    try {
        auto m = std::make_unique<bi_static_mesh>();
        parse(static_cast<const uint8_t*>(data), size, *m);
        *out = m.release();
        return 0;
    } catch (const std::bad_alloc&) {
        put_err(err, err_len, "out of host memory");
    } catch (const std::exception& e) {
        put_err(err, err_len, e.what());
    }
    return 1;
}

extern "C" void bi_static_mesh_destroy(bi_static_mesh* mesh) { delete mesh; }

extern "C" void bi_static_mesh_counts(const bi_static_mesh* m, uint64_t counts[8]) {
    counts[0] = m->positions.size() / 3;
    counts[1] = m->normals.size() / 3;
    counts[2] = m->tangents.size() / 4;
    counts[3] = m->colors.size() / 3;
    counts[4] = m->texcoords.size() / 2;
    counts[5] = m->texcoords2.size() / 2;
    counts[6] = m->indices.size();
    counts[7] = m->submeshes.size();
}

extern "C" int bi_static_mesh_submesh(const bi_static_mesh* mc, uint32_t s, bi_pt_mesh* out, char* err, size_t err_len) {
    auto* m = const_cast<bi_static_mesh*>(mc);
    if (!m || !out) return put_err(err, err_len, "null argument"), 1;
    if (s >= m->submeshes.size()) return put_err(err, err_len, "submesh index out of range"), 1;
    const SubmeshDesc& d = m->submeshes[s];
    if (d.topology != kTriangleList) return put_err(err, err_len, "submesh topology is not a triangle list"), 1;
    const uint64_t nv = m->positions.size() / 3;
    if (m->normals.size() / 3 != nv) return put_err(err, err_len, "the pass needs one normal per vertex"), 1;
    std::vector<uint32_t>& tri = m->tri_indices[s];
    tri.clear();
    uint64_t n = d.num_indices;
    if (!m->indices.empty()) {
        if (d.index_offset > m->indices.size()) return put_err(err, err_len, "index offset beyond the index data"), 1;
        n = std::min<uint64_t>(n, m->indices.size() - d.index_offset) / 3 * 3;
        tri.resize(n);
        for (uint64_t i = 0; i < n; ++i) {
            const uint64_t v = (uint64_t)m->indices[d.index_offset + i] + d.base_vertex;
            if (v >= nv) return put_err(err, err_len, "vertex index beyond the vertices"), 1;
            tri[i] = static_cast<uint32_t>(v);
        }
    } else {
        if (n == 0xFFFFFFFFull) n = nv > d.base_vertex ? nv - d.base_vertex : 0;
        n = n / 3 * 3;
        if (d.base_vertex + n > nv) return put_err(err, err_len, "vertex range beyond the vertices"), 1;
        tri.resize(n);
        for (uint64_t i = 0; i < n; ++i) tri[i] = static_cast<uint32_t>(d.base_vertex + i);
    }
    out->positions = m->positions.data();
    out->normals = m->normals.data();
    out->indices = tri.data();
    out->num_vertices = static_cast<uint32_t>(nv);
    out->num_indices = static_cast<uint32_t>(n);
    // texture coordinates and tangents when the asset has one per vertex
    out->texcoords = m->texcoords.size() == 2 * nv && nv ? m->texcoords.data() : nullptr;
    out->tangents = m->tangents.size() == 4 * nv && nv ? m->tangents.data() : nullptr;
    return 0;
}
