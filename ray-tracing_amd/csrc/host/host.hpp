// Host-side structures of liblinrt_hip.so: kd-tree build, flattening, scene
// generators and the constructor-time precompute of the reference materials.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/linrt.h"

namespace lrt {

// Reference node numbering (sah_kdtree.h:124-132): the left child of an
// interior node is always the next node; retracted subtrees stay as orphans.
struct KdNode {
    uint8_t leaf = 0;
    uint8_t axis = 0;
    float split = 0.0f;
    int32_t left = -1, right = -1, first = 0, count = 0;
};

struct KdTree {
    std::vector<KdNode> nodes;
    std::vector<uint32_t> refs;
    int depth_limit = 1;
};

// SAH kd-tree over boxes (6 floats: min xyz, max xyz).  Same splits, same node
// numbering and same reference order as sah_kdtree::build (sah_kdtree.h:35-57,
// 175-347).  threads <= 0: hardware concurrency.
void kd_build(const float* boxes, uint32_t n, KdTree& out, int threads = 0);

// 8-byte device node (PBRT-style kd layout):
//   interior: w0 = split (float bits), w1 = axis | (right_child << 2); left child = index + 1
//   leaf:     w0 = first reference,    w1 = 3 | (count << 2)
struct FlatTree {
    std::vector<uint32_t> words;  // 2 per node
    std::vector<uint32_t> refs;
    int max_depth = 0;            // deepest root-to-leaf path (stack bound)
};
// Drops orphan nodes (pre-order renumbering keeps left = parent + 1).
void kd_flatten(const KdTree& t, FlatTree& out);

// fresnel.h:128-133 + gauss_lobatto.h:23-121 (FP32 adaptive Gauss-Lobatto)
float fresnel_diffuse_reflectance(float eta);
// rough_transmittance.h:17-24,67-185 from a ggx.dat image in memory; throws on
// a bad table or out-of-range eta / alpha like the reference.
void rough_transmittance_slice(const void* data, size_t bytes, float eta, float alpha, float out101[101]);

// camera.h:20-44 (aperture 0, focus distance 1)
struct CameraBasis {
    float origin[3], llc[3], horizontal[3], vertical[3];
};
CameraBasis camera_basis(const lrt_camera_desc& c);

// Median kd tree of photon positions (photon_map.h:83-114): leaves of at most
// 8 photons, split axis = largest extent of the range, std::nth_element at
// count/2, and the split value read from the median slot after both subtrees
// are built, as the reference does (the children's partitions may have moved
// another photon there; the search prunes with that value).  nodes: 4 words
// each; interior {split bits, axis, left, right}, leaf {0, 3, first, count}.
// order[i] = input index of the photon at tree position i.
struct PhotonKd {
    std::vector<uint32_t> nodes;
    std::vector<uint32_t> order;
    int depth = 0;  // deepest root-to-leaf path (nodes)
};
void photon_kd_build(const float* xyz, uint32_t n, PhotonKd& out, int threads = 0);

// Philox-4x32-10 on the host
void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

// load_obj (obj_loader.h:10-141): positions (3 per vertex, centred and scaled
// to a unit longest side), indices (3 per triangle), corner normals (9 per
// triangle); false with a message where the reference returns nullptr or
// would throw
bool load_obj(const char* path, std::vector<float>& positions, std::vector<uint32_t>& indices,
              std::vector<float>& corner_normals, std::string& error);

// image files (image_write.cpp:10-100): the reference's 8-bit mapping of a
// channel, a whole frame of it, and the writers of format LRT_IMAGE_*
uint8_t image_to_byte(float value);
void image_to_bytes(const float* rgb, uint32_t w, uint32_t h, bool flip, uint8_t* out);
bool write_image(const char* path, const float* rgb, uint32_t w, uint32_t h, int format, bool flip, std::string& err);
// FP64 accumulators of a progressive render with their JSON sidecar
bool write_accum(const char* path, const double* sum, const double* sumsq, const lrt_accum_info& info,
                 std::string& err);
bool read_accum(const char* path, double* sum, double* sumsq, lrt_accum_info& info, std::string& err);

// direction numbers of the reference sampler's 16-dimensional Gray-code Sobol
// sequence (sobol.h:43-106): out[d * 32 + b] for dimension d, bit b
void sobol_directions(uint32_t out[16 * 32]);

// Hosek-Wilkie RGB sky state from the model's RGB datasets (per channel 1080
// coefficients and 120 radiances): 9 configuration values and the radiance
// scale per channel (ArHosekSkyModel.cc:602-635)
constexpr size_t kHosekCoefficients = 2 * 10 * 6 * 9, kHosekRadiances = 2 * 10 * 6;
void sky_cook(const double* coefficients, const double* radiances, double turbidity, double albedo, double elevation,
              double* config27, double* radiance3);

void set_error(const std::string& msg);

}  // namespace lrt
