// path_tracing_pass.hpp -- C++20 stand-in for bisemutum's PathTracingPass on top of the
// bi_pt C ABI (include/bi_pt.h).
//
// Mirrors src/renderer/pass/path_tracing.hpp:12-35 (reference paths relative to bisemutum/):
//   PathTracingPass::update_params(LightsContext&, SkyboxContext&, PathTracingSettings const&)
//   PathTracingPass::render(Camera const&, InputData const&, PathTracingSettings const&) -> OutputData
// with the engine's types reduced to what the pass reads:
//   LightsContext        -> the DirLightData array (src/renderer/context/lights.hpp:12-22, 92-131)
//   SkyboxContext        -> equirect texels, colour*strength, rotation (src/scene_basic/skybox.cpp:43-73)
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
    std::vector<bi_pt_dir_light> dir_lights;   // DirLightData records
};

struct SkyboxContext {
    std::vector<float> rgba32f;   // equirect texels, row 0 = +Y pole (empty: black sky)
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
    const SceneGeometry* scene_ptr_ = nullptr;
    float* color_ = nullptr;
    float* depth_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
};

}  // namespace bi
