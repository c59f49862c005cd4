// pass_demo.cpp -- drives PathTracingPass the way BasicRenderer::render_camera does
// (src/renderer/basic.cpp:46-48 update_params, :157-161 render), frame after frame with
// accumulation, on a Lambertian plane under a constant sky and a directional light whose
// expected radiance is known in closed form: rho * Ls + rho / pi * cos(theta) * E.
// Exit status 0 when the accumulated image matches it within 0.5%.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "path_tracing_pass.hpp"

namespace {

void look_at_camera(bi::Camera& c, uint32_t w, uint32_t h) {
    // eye (0, 2, 0) looking at (0, 0, 0.3), up +Y, 40 degree vertical fov, reverse-Z (math.cpp:5-11)
    const double eye[3] = {0, 2, 0}, at[3] = {0, 0, 0.3}, up[3] = {0, 1, 0};
    double f[3] = {at[0] - eye[0], at[1] - eye[1], at[2] - eye[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    for (double& x : s) x /= sl;
    double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    // inverse view: columns s, u, -f and translation eye
    const double iv[16] = {s[0], u[0], -f[0], eye[0], s[1], u[1], -f[1], eye[1],
                           s[2], u[2], -f[2], eye[2], 0, 0, 0, 1};
    const double t = std::tan(40.0 * M_PI / 360.0), aspect = (double)w / h, n = 0.05, fr = 500.0;
    // inverse of the reverse-Z perspective P = [[1/(a t),0,0,0],[0,1/t,0,0],[0,0,n/(f-n),nf/(f-n)],[0,0,-1,0]]
    const double A = n / (fr - n), B = n * fr / (fr - n);
    const double ip[16] = {aspect * t, 0, 0, 0, 0, t, 0, 0, 0, 0, 0, -1, 0, 0, 1 / B, A / B};
    const double P[16] = {1 / (aspect * t), 0, 0, 0, 0, 1 / t, 0, 0, 0, 0, A, B, 0, 0, -1, 0};
    double V[16];   // view = inverse(iv) for a rigid transform
    const double R[3][3] = {{s[0], s[1], s[2]}, {u[0], u[1], u[2]}, {-f[0], -f[1], -f[2]}};
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) V[4 * r + k] = R[r][k];
        V[4 * r + 3] = -(R[r][0] * eye[0] + R[r][1] * eye[1] + R[r][2] * eye[2]);
    }
    V[12] = V[13] = V[14] = 0;
    V[15] = 1;
    for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 4; ++k) {
            double acc = 0;
            for (int m = 0; m < 4; ++m) acc += P[4 * r + m] * V[4 * m + k];
            c.matrix_proj_view[4 * r + k] = (float)acc;
            c.matrix_inv_view[4 * r + k] = (float)iv[4 * r + k];
            c.matrix_inv_proj[4 * r + k] = (float)ip[4 * r + k];
        }
    c.width = w;
    c.height = h;
}

}  // namespace

int main() {
    const float rho[3] = {0.5f, 0.4f, 0.3f}, ls[3] = {0.5f, 0.6f, 0.8f}, e[3] = {2.0f, 1.9f, 1.7f};
    const double sun_raw[3] = {0.3, 0.8, 0.2};
    const double sl = std::sqrt(0.09 + 0.64 + 0.04);
    const float sun[3] = {(float)(sun_raw[0] / sl), (float)(sun_raw[1] / sl), (float)(sun_raw[2] / sl)};

    // scene: one two-sided Lambertian quad of 200 x 200 at y = 0
    std::vector<float> pos = {-100, 0, -100, 100, 0, -100, 100, 0, 100, -100, 0, 100};
    std::vector<float> nrm = {0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0};
    std::vector<uint32_t> idx = {0, 1, 2, 0, 2, 3};
    bi::SceneGeometry scene;
    scene.meshes.push_back(bi_pt_mesh{pos.data(), nrm.data(), idx.data(), 4, 6});
    bi_pt_instance inst{};
    const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    for (int k = 0; k < 12; ++k) inst.transform[k] = id[k];
    scene.instances.push_back(inst);
    bi_pt_material mat{};
    for (int c = 0; c < 3; ++c) mat.base_color[c] = rho[c];
    mat.roughness = 1.0f;
    mat.ior = 1.5f;
    mat.two_sided = 1;   // f0 = f90 = 0: the lit BSDF reduces to Lambert
    scene.materials.push_back(mat);

    bi::LightsContext lights;
    bi_pt_dir_light L{};
    for (int c = 0; c < 3; ++c) { L.emission[c] = e[c]; L.direction[c] = sun[c]; }
    L.sm_index = 0;
    L.shadow_strength = 1.0f;
    lights.dir_lights.push_back(L);
    bi::SkyboxContext sky;
    sky.width = 64;
    sky.height = 32;
    sky.rgba32f.assign(64 * 32 * 4, 1.0f);
    for (int c = 0; c < 3; ++c) sky.color[c] = ls[c];

    bi::PathTracingSettings settings;
    settings.max_bounces = 3;
    settings.spp = 64;
    settings.seed = 11;
    settings.ray_length = 1000.0f;
    bi::Camera cam;
    look_at_camera(cam, 32, 24);

    try {
        bi::PathTracingPass pass(0);
        const int frames = 8;
        bi::PathTracingPass::OutputData out{};
        for (int f = 0; f < frames; ++f) {
            pass.update_params(lights, sky, settings);   // unchanged inputs keep the history
            out = pass.render(cam, bi::PathTracingPass::InputData{&scene}, settings);
        }
        std::vector<float> img(32 * 24 * 4);
        if (hipMemcpy(img.data(), out.color, img.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return 3;
        double mean[3] = {0, 0, 0};
        for (size_t p = 0; p < 32 * 24; ++p)
            for (int c = 0; c < 3; ++c) mean[c] += img[4 * p + c] / (32.0 * 24.0);
        bi_pt_stats st = pass.stats();
        int bad = 0;
        for (int c = 0; c < 3; ++c) {
            double want = rho[c] * ls[c] + rho[c] / M_PI * sun[1] * e[c];
            std::printf("channel %d: mean %.6f expected %.6f (rel %.2e)\n", c, mean[c], want, mean[c] / want - 1);
            if (std::fabs(mean[c] / want - 1) > 5e-3) bad = 1;
        }
        std::printf("frames %llu, samples accumulated %llu, rays %llu\n", (unsigned long long)st.frames,
                    (unsigned long long)st.samples_accumulated,
                    (unsigned long long)(st.rays_camera + st.rays_extension + st.rays_light_shadow + st.rays_sky_shadow));
        return bad;
    } catch (const bi::PathTracingError& ex) {
        std::fprintf(stderr, "PathTracingError %d: %s\n", (int)ex.status, ex.what());
        return 2;
    }
}
