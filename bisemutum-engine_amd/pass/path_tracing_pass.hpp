// path_tracing_pass.hpp -- C++20 stand-in for bisemutum's PathTracingPass on top of the
// bi_pt C ABI (include/bi_pt.h).
//
// Mirrors src/renderer/pass/path_tracing.hpp:12-35 (reference paths relative to bisemutum/):
//   PathTracingPass::update_params(LightsContext&, SkyboxContext&, PathTracingSettings const&)
//   PathTracingPass::render(Camera const&, InputData const&, PathTracingSettings const&) -> OutputData
// with the engine's types reduced to what the pass reads:
//   LightsContext        -> the DirLightData, PointLightData and RectLightData arrays, the rect
//                           lights' textures and the LTC tables (src/renderer/context/lights.hpp:
//                           12-22,65-131, lights.cpp:26-37,120-243)
//   SkyboxContext        -> the skybox cube (or an equirect), colour*strength, rotation
//                           (src/scene_basic/skybox.cpp:43-73)
//   PathTracingSettings  -> include/bisemutum/renderer/basic.hpp:76-81 (denoise is unused by the pass)
//   InputData            -> the scene that the TLAS is built from (the reference takes a
//                           ready AccelerationStructureHandle; here the pass builds its LBVH)
//   OutputData           -> device colour (RGBA32F instead of rgba16f) and reverse-Z depth
// Errors: the reference logs and continues (BI_ASSERT_MSG); this adapter throws
// bi::PathTracingError carrying the bi_pt status and message.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bi_pt.h"

namespace bi {

struct PathTracingError : std::runtime_error {
    bi_pt_status status;
    PathTracingError(bi_pt_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

// basic.hpp:76-81 plus the sample schedule of this pass.
struct PathTracingSettings {
    float ray_length = 100.0f;
    uint32_t max_bounces = 3;
    bool denoise = true;      // accepted and ignored, as in the reference pass
    bool accumulate = true;
    uint32_t spp = 1;
    uint32_t seed = 0;
    uint32_t flags = 0;
};

struct LightsContext {
    std::vector<bi_pt_dir_light> dir_lights;       // DirLightData records
    std::vector<bi_pt_point_light> point_lights;   // PointLightData records (with dir_lights, at most 4)
    std::vector<bi_pt_rect_light> rect_lights;     // RectLightData records (at most 16)
    std::vector<bi_pt_rect_texture> rect_light_textures;   // what texture_index >= 0 names
    // LTC tables (ltc_matrix_lut0..2 RGBA32F, ltc_norm_lut RG32F, 8 x 8 x 64 texels each);
    // empty: not loaded (then rect_lights must be empty)
    std::vector<float> ltc_matrix[3], ltc_norm;
};

struct SkyboxContext {
    std::vector<float> rgba32f;   // equirect texels, row 0 = +Y pole, or six cube faces (empty: black sky)
    bool cube = false;            // rgba32f holds the skybox cube: six width x width faces (height unused)
    uint32_t width = 0, height = 0;
    float color[3] = {0.0f, 0.0f, 0.0f};
    float rotation[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
};

struct Camera {
    float matrix_inv_view[16];
    float matrix_inv_proj[16];
    float matrix_proj_view[16];
    uint32_t width = 0, height = 0;
};

struct SceneGeometry {
    std::vector<bi_pt_mesh> meshes;
    std::vector<bi_pt_instance> instances;
    std::vector<bi_pt_material> materials;
    uint64_t version = 0;   // bump when geometry or materials change (like mesh/BLAS versions)
};

class PathTracingPass final {
public:
    struct InputData {
        const SceneGeometry* scene = nullptr;
    };
    struct OutputData {
        float* color = nullptr;   // device RGBA32F, width*height*4
        float* depth = nullptr;   // device float, width*height (reverse-Z, 0 = miss)
    };

    explicit PathTracingPass(int device = 0, uint64_t hbm_budget_bytes = 0);
    ~PathTracingPass();
    PathTracingPass(const PathTracingPass&) = delete;
    PathTracingPass& operator=(const PathTracingPass&) = delete;

    void update_params(LightsContext& lights_ctx, SkyboxContext& skybox_ctx, PathTracingSettings const& settings);
    OutputData render(Camera const& camera, InputData const& input, PathTracingSettings const& settings,
                      bi_pt_stream stream = nullptr);
    bi_pt_stats stats() const;

private:
    void check(bi_pt_status s) const;
    void ensure_outputs(uint32_t width, uint32_t height);

    bi_pt_ctx* ctx_ = nullptr;
    uint64_t scene_version_ = ~0ull;
    size_t num_point_lights_ = 0;   // point lights last given to the context
    const SceneGeometry* scene_ptr_ = nullptr;
    float* color_ = nullptr;
    float* depth_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
};

}  // namespace bi
