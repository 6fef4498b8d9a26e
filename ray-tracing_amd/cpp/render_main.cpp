// linrt_render: C++ driver in the shape of the reference's src/main.cpp
// (scene -> integrator -> timed frame loop -> image), rendering the seeded
// box-icosphere scene on the GPU through linrt::gpu_path_integrator (or,
// with --photon, linrt::gpu_photon_integrator: the reference's other
// integrator, selected there by #ifdef, src/main.cpp:92-101) and writing the
// image with lrt_image_write: OUT.png as the reference saves output.png
// (8-bit, rows flipped, src/main.cpp:151), or .pfm / .hdr / .ppm / .raw.
//
//   linrt_render [--photon] OUT.{png,pfm,hdr,ppm,raw} [width height spp max_depth rr_depth seed]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "linrt.hpp"

int main(int argc, char** argv) {
    const bool photon = argc > 1 && std::strcmp(argv[1], "--photon") == 0;
    if (photon) {
        --argc;
        ++argv;
    }
    if (argc < 2) {
        std::cerr << "usage: linrt_render [--photon] OUT.pfm [width height spp max_depth rr_depth seed]\n";
        return 2;
    }
    const uint32_t W = argc > 2 ? (uint32_t)std::atoi(argv[2]) : 256;
    const uint32_t H = argc > 3 ? (uint32_t)std::atoi(argv[3]) : 256;
    const uint32_t spp = argc > 4 ? (uint32_t)std::atoi(argv[4]) : 16;
    const uint32_t max_depth = argc > 5 ? (uint32_t)std::atoi(argv[5]) : 8;
    const uint32_t rr_depth = argc > 6 ? (uint32_t)std::atoi(argv[6]) : 3;
    const uint64_t seed = argc > 7 ? (uint64_t)std::atoll(argv[7]) : 1;
    try {
        lrt_gen_params gp;
        std::memset(&gp, 0, sizeof(gp));
        gp.kind = LRT_GEN_BOX_ICOSPHERE;
        gp.seed = 1;
        gp.subdivisions = 5;
        gp.aspect = (float)W / (float)H;
        linrt::generated gen(gp);
        linrt::scene scene(gen.desc());
        std::vector<linrt::color> pixels;
        const auto t0 = std::chrono::steady_clock::now();
        lrt_render_stats st;
        if (photon) {
            const linrt::gpu_photon_integrator integrator(argc > 5 ? max_depth : 128, seed);
            integrator.preprocess(scene);
            st = integrator.render(scene, W, H, spp, pixels);
        } else {
            const linrt::gpu_path_integrator integrator(max_depth, rr_depth, seed);
            st = integrator.render(scene, W, H, spp, pixels);
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("%u x %u x %u spp: %.3f s, %.2f Msamples/s, %llu extension + %llu shadow rays\n", W, H, spp, s,
                    (double)st.samples / s / 1e6, (unsigned long long)st.extension_rays,
                    (unsigned long long)st.shadow_rays);
        const std::string out = argv[1];
        const std::string ext = out.substr(out.find_last_of('.') + 1);
        int fmt = LRT_IMAGE_PFM;
        if (ext == "png") fmt = LRT_IMAGE_PNG;
        else if (ext == "hdr") fmt = LRT_IMAGE_HDR;
        else if (ext == "ppm") fmt = LRT_IMAGE_PPM;
        else if (ext == "raw") fmt = LRT_IMAGE_RAW;
        linrt::check(lrt_image_write(out.c_str(), &pixels[0].r, W, H, fmt, fmt == LRT_IMAGE_PNG ? 1 : 0));
    } catch (const linrt::error& e) {
        std::cerr << "linrt error " << e.code << ": " << e.what() << "\n";
        return 1;
    }
    return 0;
}
