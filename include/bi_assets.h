/* bi_assets.h -- host-side loader of the engine's binary StaticMesh assets, feeding
 * bi_pt_set_scene (include/bi_pt.h).  No GPU calls.
 *
 * Replaces StaticMesh::load (src/scene_basic/static_mesh.cpp:10-43) with its byte
 * stream (src/prelude/byte_stream.cpp:15-41, MeshData::load_from_byte_stream
 * src/graphics/mesh.cpp:137-146) for callers outside the engine's runtime: the
 * file's bytes in, one bi_pt_mesh per submesh out.
 *
 * Format: u32 magic 0x0B1A55E7 (include/bisemutum/runtime/asset.hpp:15), u64-length
 * string "StaticMesh", u32 version; version 1: MeshData; version 2: u64 raw length,
 * u64 compressed length, zlib stream of MeshData.  MeshData: vectors (u64 count +
 * elements) of float3 positions, float3 normals, float4 tangents, float3 colors,
 * float2 texcoords, float2 texcoords2, u32 indices, and 16-byte SubmeshDesc records
 * {u32 base_vertex, u32 index_offset, u32 num_indices, u8 topology, 3 pad}.
 *
 * Errors: the reference logs and returns an empty asset; these return nonzero and
 * write a message into err (truncated to err_len, always NUL-terminated). */
#ifndef BI_ASSETS_H
#define BI_ASSETS_H

#include <stddef.h>
#include <stdint.h>

#include "bi_pt.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bi_static_mesh bi_static_mesh;

/* Parse a StaticMesh asset.  Returns 0 and *out on success. */
int bi_static_mesh_load(const void* data, size_t size, bi_static_mesh** out, char* err, size_t err_len);
void bi_static_mesh_destroy(bi_static_mesh* mesh);

/* Counts of the parsed MeshData vectors: positions, normals, tangents, colors,
 * texcoords, texcoords2, indices, submeshes (8 values). */
void bi_static_mesh_counts(const bi_static_mesh* mesh, uint64_t counts[8]);

/* Submesh s as the pass's mesh input.  The pointers are owned by `mesh` and stay valid
 * until bi_static_mesh_destroy or the next call for the same submesh; calls for one
 * mesh from several threads must not overlap.  Triangles
 * indices[index_offset .. +n] + base_vertex with n = num_indices (0xFFFFFFFF: every
 * index from index_offset on), clamped to the index data and cut to whole triangles
 * as the bottom-level build counts them (src/graphics/graphics_manager.cpp:634);
 * without index data, n vertices from base_vertex (include/bisemutum/graphics/mesh.hpp:19).
 * Fails for a topology other than triangle_list (3), an index beyond the vertices,
 * or normals missing. */
int bi_static_mesh_submesh(const bi_static_mesh* mesh, uint32_t s, bi_pt_mesh* out, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif
