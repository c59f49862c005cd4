// path_tracing_pass.cpp -- PathTracingPass adapter over the bi_pt C ABI.
#include "path_tracing_pass.hpp"

#include <hip/hip_runtime.h>

namespace bi {

PathTracingPass::PathTracingPass(int device, uint64_t hbm_budget_bytes) {
    bi_pt_device_desc desc{device, 0u, hbm_budget_bytes};
    check(bi_pt_create(&desc, &ctx_));
}

PathTracingPass::~PathTracingPass() {
    if (color_) (void)hipFree(color_);
    if (depth_) (void)hipFree(depth_);
    if (ctx_) bi_pt_destroy(ctx_);
}

void PathTracingPass::check(bi_pt_status s) const {
    if (s != BI_PT_OK) throw PathTracingError(s, ctx_ ? bi_pt_last_error(ctx_) : "bi_pt_create failed");
}

// path_tracing.cpp:183-222: binds the light and sky contexts for the coming frame.
void PathTracingPass::update_params(LightsContext& lights_ctx, SkyboxContext& skybox_ctx,
                                    PathTracingSettings const&) {
    // directional and point lights share four slots: drop the old point lights first only
    // when the new directional ones would not fit beside them (unchanged lights keep the
    // running image)
    if (lights_ctx.dir_lights.size() + num_point_lights_ > 4) {
        check(bi_pt_set_point_lights(ctx_, nullptr, 0));
        num_point_lights_ = 0;
    }
    check(bi_pt_set_lights(ctx_, lights_ctx.dir_lights.data(), (uint32_t)lights_ctx.dir_lights.size()));
    check(bi_pt_set_point_lights(ctx_, lights_ctx.point_lights.empty() ? nullptr : lights_ctx.point_lights.data(),
                                 (uint32_t)lights_ctx.point_lights.size()));
    num_point_lights_ = lights_ctx.point_lights.size();
    if (!lights_ctx.ltc_norm.empty()) {
        for (const auto& t : lights_ctx.ltc_matrix)
            if (t.size() != 8u * 8u * 64u * 4u)
                throw PathTracingError(BI_PT_ERROR_INVALID_ARGUMENT, "LTC matrix table size");
        if (lights_ctx.ltc_norm.size() != 8u * 8u * 64u * 2u)
            throw PathTracingError(BI_PT_ERROR_INVALID_ARGUMENT, "LTC norm table size");
        check(bi_pt_set_ltc_luts(ctx_, lights_ctx.ltc_matrix[0].data(), lights_ctx.ltc_matrix[1].data(),
                                 lights_ctx.ltc_matrix[2].data(), lights_ctx.ltc_norm.data()));
    }
    check(bi_pt_set_rect_light_textures(ctx_, lights_ctx.rect_light_textures.data(),
                                        (uint32_t)lights_ctx.rect_light_textures.size()));
    check(bi_pt_set_rect_lights(ctx_, lights_ctx.rect_lights.empty() ? nullptr : lights_ctx.rect_lights.data(),
                                (uint32_t)lights_ctx.rect_lights.size()));
    const float* sky = skybox_ctx.rgba32f.empty() ? nullptr : skybox_ctx.rgba32f.data();
    if (skybox_ctx.cube)
        check(bi_pt_set_sky_cube(ctx_, sky, sky ? skybox_ctx.width : 0u, skybox_ctx.color, skybox_ctx.rotation));
    else
        check(bi_pt_set_sky(ctx_, sky, skybox_ctx.width, skybox_ctx.height, skybox_ctx.color, skybox_ctx.rotation));
}

void PathTracingPass::ensure_outputs(uint32_t width, uint32_t height) {
    if (width == width_ && height == height_ && color_) return;
    if (color_) (void)hipFree(color_);
    if (depth_) (void)hipFree(depth_);
    color_ = depth_ = nullptr;
    const size_t npix = (size_t)width * height;
    if (hipMalloc(&color_, npix * 4 * sizeof(float)) != hipSuccess || hipMalloc(&depth_, npix * sizeof(float)) != hipSuccess)
        throw PathTracingError(BI_PT_ERROR_OUT_OF_MEMORY, "output allocation failed");
    width_ = width;
    height_ = height;
}

// path_tracing.cpp:224-488: one frame; accumulation continues while the camera and
// settings are unchanged (has_history, :228-246), otherwise restarts.
PathTracingPass::OutputData PathTracingPass::render(Camera const& camera, InputData const& input,
                                                    PathTracingSettings const& settings, bi_pt_stream stream) {
    if (!input.scene) throw PathTracingError(BI_PT_ERROR_NO_SCENE, "InputData.scene is null");
    if (input.scene != scene_ptr_ || input.scene->version != scene_version_) {
        const SceneGeometry& g = *input.scene;
        check(bi_pt_set_scene(ctx_, g.meshes.data(), (uint32_t)g.meshes.size(), g.instances.data(),
                              (uint32_t)g.instances.size(), g.materials.data(), (uint32_t)g.materials.size()));
        scene_ptr_ = input.scene;
        scene_version_ = g.version;
    }
    ensure_outputs(camera.width, camera.height);
    bi_pt_camera cam{};
    for (int k = 0; k < 16; ++k) {
        cam.matrix_inv_view[k] = camera.matrix_inv_view[k];
        cam.matrix_inv_proj[k] = camera.matrix_inv_proj[k];
        cam.matrix_proj_view[k] = camera.matrix_proj_view[k];
    }
    cam.width = camera.width;
    cam.height = camera.height;
    bi_pt_settings st{};
    st.ray_length = settings.ray_length;
    st.max_bounces = settings.max_bounces;
    st.accumulate = settings.accumulate ? 1u : 0u;
    st.spp = settings.spp;
    st.seed = settings.seed;
    st.flags = settings.flags;
    check(bi_pt_render(ctx_, &cam, &st, color_, depth_, stream));
    return OutputData{color_, depth_};
}

bi_pt_stats PathTracingPass::stats() const {
    bi_pt_stats s{};
    check(bi_pt_get_stats(ctx_, &s));
    return s;
}

}  // namespace bi
