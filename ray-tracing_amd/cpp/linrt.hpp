// C++ host layer over the C ABI of liblinrt_hip.so.
//
// Mirrors the roles of the reference's interfaces so a GPU integrator can
// stand in for the CPU one:
//   linrt::scene               <- LinRT::scene (scene.h:19-151) built from a
//                                 lrt_scene_desc and uploaded to one GPU
//   linrt::gpu_path_integrator <- LinRT::path_integrator (path_integrator.h:16)
//                                 driven by the frame loop of src/main.cpp:103-147:
//                                 render() fills pixels[y*W+x] with the mean
//                                 radiance of spp samples, row 0 at the bottom
// Errors from the library become linrt::error exceptions on this (host) side
// of the boundary; nothing throws across the C ABI itself.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/linrt.h"

namespace linrt {

struct error : std::runtime_error {
    int code;
    error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != LRT_OK) throw error(rc, lrt_last_error());
}

class scene {
public:
    scene(const lrt_scene_desc& desc, int device = 0) { check(lrt_scene_create(device, &desc, &h_)); }
    ~scene() { if (h_) lrt_scene_release(h_); }
    scene(const scene&) = delete;
    scene& operator=(const scene&) = delete;
    scene(scene&& o) noexcept : h_(std::exchange(o.h_, nullptr)) {}
    lrt_scene* handle() const { return h_; }
    lrt_scene_info info() const {
        lrt_scene_info i;
        check(lrt_scene_info_get(h_, &i));
        return i;
    }

private:
    lrt_scene* h_ = nullptr;
};

struct color {
    float r, g, b;
};

class gpu_path_integrator {
public:
    // path_integrator.h:173,19 hard-code 50 and 5; here they are parameters
    explicit gpu_path_integrator(uint32_t max_depth = 50, uint32_t rr_depth = 5, uint64_t seed = 1)
        : max_depth_(max_depth), rr_depth_(rr_depth), seed_(seed) {}

    // Renders width x height x spp samples; pixels[y*width + x] = mean radiance
    // (the reference's pixel_color / SAMPLES_PER_PIXEL, src/main.cpp:135-136).
    lrt_render_stats render(const scene& s, uint32_t width, uint32_t height, uint32_t spp, std::vector<color>& pixels) const {
        std::vector<double> sum((size_t)width * height * 3, 0.0);
        lrt_render_params p{};
        p.width = width;
        p.height = height;
        p.spp = spp;
        p.max_depth = max_depth_;
        p.rr_depth = rr_depth_;
        p.seed = seed_;
        lrt_render_stats st{};
        check(lrt_render(s.handle(), &p, sum.data(), nullptr, 0, &st));
        pixels.resize((size_t)width * height);
        for (size_t i = 0; i < pixels.size(); ++i)
            pixels[i] = color{(float)(sum[3 * i] / spp), (float)(sum[3 * i + 1] / spp), (float)(sum[3 * i + 2] / spp)};
        return st;
    }

private:
    uint32_t max_depth_, rr_depth_;
    uint64_t seed_;
};

// LinRT::photon_integrator (photon_integrator.h:13-123): preprocess() builds
// the global map and, when the scene has a delta BSDF, the caustic map with
// the reference's constants (250,000 photons each, lookup 120, radii 0.05
// and 0.0125 x the bounding-sphere radius, photon depth max_depth - 1, RR
// from 5); render() is the eye pass over those maps.
class gpu_photon_integrator {
public:
    explicit gpu_photon_integrator(uint32_t max_depth = 128, uint64_t seed = 1, uint32_t global_photons = 250000,
                                   uint32_t caustic_photons = 250000, uint32_t lookup = 120)
        : max_depth_(max_depth), seed_(seed), global_(global_photons), caustic_(caustic_photons), lookup_(lookup) {}

    void preprocess(const scene& s) const {
        const lrt_scene_info info = s.info();
        lrt_photon_params p{};
        p.kind = 0;
        p.requested = global_;
        p.max_depth = max_depth_ - 1;
        p.rr_depth = 5;
        p.seed = seed_;
        p.radius = 0.05f * info.bounding_radius;
        p.k = lookup_;
        check(lrt_photon_build(s.handle(), &p, nullptr));
        if (info.has_delta_bsdf && caustic_) {
            p.kind = 1;
            p.requested = caustic_;
            p.radius = 0.0125f * info.bounding_radius;
            check(lrt_photon_build(s.handle(), &p, nullptr));
        } else {
            check(lrt_photon_map_load(s.handle(), 1, nullptr, nullptr, 0, 1.0f, 0.0125f * info.bounding_radius, lookup_));
        }
    }

    // pixels[y*width + x] = mean radiance of spp eye samples, row 0 at the bottom
    lrt_render_stats render(const scene& s, uint32_t width, uint32_t height, uint32_t spp, std::vector<color>& pixels) const {
        std::vector<double> sum((size_t)width * height * 3, 0.0);
        lrt_render_params p{};
        p.width = width;
        p.height = height;
        p.spp = spp;
        p.max_depth = max_depth_;
        p.seed = seed_;
        lrt_render_stats st{};
        check(lrt_render_photon(s.handle(), &p, sum.data(), nullptr, 0, &st));
        pixels.resize((size_t)width * height);
        for (size_t i = 0; i < pixels.size(); ++i)
            pixels[i] = color{(float)(sum[3 * i] / spp), (float)(sum[3 * i + 1] / spp), (float)(sum[3 * i + 2] / spp)};
        return st;
    }

private:
    uint32_t max_depth_;
    uint64_t seed_;
    uint32_t global_, caustic_, lookup_;
};

// owner of a seeded generated scene (lrt_generate_scene)
class generated {
public:
    explicit generated(const lrt_gen_params& p, const std::vector<float>& tables = {}, const lrt_light_desc* sky = nullptr) {
        check(lrt_generate_scene(&p, tables.empty() ? nullptr : tables.data(), (uint32_t)(tables.size() / 101), sky, &g_));
    }
    ~generated() { if (g_) lrt_generated_free(g_); }
    generated(const generated&) = delete;
    generated& operator=(const generated&) = delete;
    const lrt_scene_desc& desc() const { return *lrt_generated_desc(g_); }

private:
    lrt_generated* g_ = nullptr;
};

}  // namespace linrt
